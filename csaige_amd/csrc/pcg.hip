// Batched preconditioned conjugate gradient on Sigma = tau0 * diag(1/W) + tau1 * GRM.
//
// Reference: PcgSolver::solve (saige-linalg/src/decomposition.rs:212-298) with the Sigma
// operator and diagonal preconditioner closures of saige-core/src/glmm/ai_reml.rs:142-164.
// k right-hand sides share one multi-column GRM product per iteration; every column keeps
// its own rz, alpha, beta, residual norm and stop flag, so each column follows exactly the
// iteration of a separate PcgSolver::solve:
//   x0 = 0 (the product A*x0 is exactly zero and is not formed), r = b, z = M^-1 r, p = z;
//   loop: stop if ||r|| < tol * max(||b||, 1); Ap; breakdown if |p'Ap| < 1e-30;
//         alpha = rz / p'Ap; x += alpha p; r -= alpha Ap; z = M^-1 r; beta = rz'/rz; p = z + beta p.
// Dot products and norms are fp64 with a deterministic two-stage reduction.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

namespace saige {

struct ColState {
  double rz, tol_abs, alpha, beta, rnorm;
  int active, iters, converged, pad;
};

// tau[0], tau[1] live in device memory (PcgWork scalars), so that a captured iteration reads
// the current solve's values.
__device__ __forceinline__ double precond_apply(double r, double w, double tau0, double tau1) {
  double diag = tau0 / fmax(w, 1e-30) + tau1;
  return fabs(diag) > 1e-30 ? r / diag : r;
}

__device__ __forceinline__ void block_sum3(double& a, double& b, double& c) {
  __shared__ double sa[kDotBlock], sb[kDotBlock], sc[kDotBlock];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int w = kDotBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sa[threadIdx.x] += sa[threadIdx.x + w];
      sb[threadIdx.x] += sb[threadIdx.x + w];
      sc[threadIdx.x] += sc[threadIdx.x + w];
    }
    __syncthreads();
  }
  a = sa[0];
  b = sb[0];
  c = sc[0];
}

// r = b, z = M^-1 r, p = z, x = 0; partial sums of r'z, b'b, r'r.
__global__ __launch_bounds__(kDotBlock) void pcg_init_kernel(
    const double* __restrict__ b, size_t ldb, double* __restrict__ x, size_t ldx,
    double* __restrict__ r, double* __restrict__ z, double* __restrict__ p, size_t ld,
    const double* __restrict__ w, const double* __restrict__ tau, size_t n, size_t nchunks,
    double* __restrict__ part) {
  const double tau0 = tau[0], tau1 = tau[1];
  const size_t c = blockIdx.y;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, n);
  double s_rz = 0.0, s_bb = 0.0, s_rr = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) {
    double bi = b[i + c * ldb];
    double zi = precond_apply(bi, w[i], tau0, tau1);
    r[i + c * ld] = bi;
    z[i + c * ld] = zi;
    p[i + c * ld] = zi;
    x[i + c * ldx] = 0.0;
    s_rz = fma(bi, zi, s_rz);
    s_bb = fma(bi, bi, s_bb);
  }
  s_rr = s_bb;
  block_sum3(s_rz, s_bb, s_rr);
  if (threadIdx.x == 0) {
    size_t k = gridDim.y;
    part[(0 * k + c) * nchunks + blockIdx.x] = s_rz;
    part[(1 * k + c) * nchunks + blockIdx.x] = s_bb;
    part[(2 * k + c) * nchunks + blockIdx.x] = s_rr;
  }
}

// Initial column states (one block of 1024 threads covers all columns) and the active list:
// n_active[0] is the number of active columns, n_active[1..] their indices in order.
__global__ __launch_bounds__(1024) void pcg_init_fin(const double* __restrict__ part,
                                                     size_t nchunks, size_t k, double tol,
                                                     ColState* __restrict__ st,
                                                     int* __restrict__ n_active,
                                                     int* __restrict__ iter) {
  __shared__ int flag[1024];
  if (threadIdx.x == 0) *iter = 0;
  const size_t c = threadIdx.x;
  int act = 0;
  if (c < k) {
    double rz = 0.0, bb = 0.0, rr = 0.0;
    for (size_t q = 0; q < nchunks; ++q) {
      rz += part[(0 * k + c) * nchunks + q];
      bb += part[(1 * k + c) * nchunks + q];
      rr += part[(2 * k + c) * nchunks + q];
    }
    ColState s;
    s.rz = rz;
    s.tol_abs = tol * fmax(sqrt(bb), 1.0);
    s.alpha = 0.0;
    s.beta = 0.0;
    s.rnorm = sqrt(rr);
    s.iters = 0;
    s.pad = 0;
    if (s.rnorm < s.tol_abs) {
      s.active = 0;
      s.converged = 1;
    } else {
      s.active = 1;
      s.converged = 0;
    }
    st[c] = s;
    act = s.active;
  }
  flag[threadIdx.x] = act;
  __syncthreads();
  if (threadIdx.x == 0) {
    int cnt = 0;
    for (size_t q = 0; q < k; ++q)
      if (flag[q]) n_active[1 + cnt++] = (int)q;
    n_active[0] = cnt;
  }
}

// Sigma p = tau0 p / max(w, 1e-30) + tau1 GRM p written over ap (ap holds GRM p on entry;
// the closure of ai_reml.rs:142-152) and the chunk's partial p'(Sigma p), for the active
// columns: one pass over p and ap.
__global__ __launch_bounds__(kDotBlock) void pcg_sigma_pap_kernel(
    const double* __restrict__ p, double* __restrict__ ap, size_t ld,
    const double* __restrict__ w, const double* __restrict__ tau, size_t n, size_t nchunks,
    const ColState* __restrict__ st, double* __restrict__ part) {
  const size_t c = blockIdx.y;
  if (!st[c].active) return;
  const double t0 = tau[0], t1 = tau[1];
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, n);
  double s = 0.0, d1 = 0.0, d2 = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) {
    const double pi = p[i + c * ld];
    const double v = t0 * pi / fmax(w[i], 1e-30) + t1 * ap[i + c * ld];
    ap[i + c * ld] = v;
    s = fma(pi, v, s);
  }
  block_sum3(s, d1, d2);
  if (threadIdx.x == 0) part[c * nchunks + blockIdx.x] = s;
}

// x += alpha p; r -= alpha Ap; z = M^-1 r; partial r'z and r'r.  Every block of column c
// first finishes p'Ap from the partials (same fixed order in every block) and takes
// alpha = r'z / p'Ap; on breakdown (|p'Ap| < 1e-30, decomposition.rs:257-264) the column
// stops, recorded by block 0.
__global__ __launch_bounds__(kDotBlock) void pcg_update_kernel(
    double* __restrict__ x, size_t ldx, double* __restrict__ r, double* __restrict__ z,
    const double* __restrict__ p, const double* __restrict__ ap, size_t ld,
    const double* __restrict__ w, const double* __restrict__ tau, size_t n, size_t nchunks,
    const double* __restrict__ pap_part, const int* __restrict__ iter_d,
    ColState* __restrict__ st, double* __restrict__ part) {
  const size_t c = blockIdx.y;
  if (!st[c].active) return;
  const double tau0 = tau[0], tau1 = tau[1];
  const int iter = *iter_d;
  __shared__ double alpha_s;
  __shared__ int stop_s;
  if (threadIdx.x == 0) {
    double pap = 0.0;
    for (size_t q = 0; q < nchunks; ++q) pap += pap_part[c * nchunks + q];
    stop_s = fabs(pap) < 1e-30;
    alpha_s = stop_s ? 0.0 : st[c].rz / pap;
  }
  __syncthreads();
  if (stop_s) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st[c].active = 0;
      st[c].converged = 0;
      st[c].iters = iter;
    }
    return;
  }
  const double alpha = alpha_s;
  const size_t i0 = (size_t)blockIdx.x * kDotChunk, i1 = min(i0 + (size_t)kDotChunk, n);
  double s_rz = 0.0, s_rr = 0.0, d = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += kDotBlock) {
    double pi = p[i + c * ld];
    x[i + c * ldx] += alpha * pi;
    double ri = r[i + c * ld] - alpha * ap[i + c * ld];
    r[i + c * ld] = ri;
    double zi = precond_apply(ri, w[i], tau0, tau1);
    z[i + c * ld] = zi;
    s_rz = fma(ri, zi, s_rz);
    s_rr = fma(ri, ri, s_rr);
  }
  block_sum3(s_rz, s_rr, d);
  if (threadIdx.x == 0) {
    size_t k = gridDim.y;
    part[(0 * k + c) * nchunks + blockIdx.x] = s_rz;
    part[(1 * k + c) * nchunks + blockIdx.x] = s_rr;
  }
}

// beta, new r'z and ||r|| per active column, the stop test of the next iteration, and the
// columns still active: n_active[0] is their number, n_active[1..] their indices in
// increasing order (one block of 1024 threads covers all columns).
__global__ __launch_bounds__(1024) void pcg_beta_fin(const double* __restrict__ part,
                                                     size_t nchunks, size_t k,
                                                     int* __restrict__ iter_d,
                                                     ColState* __restrict__ st,
                                                     int* __restrict__ n_active) {
  __shared__ int flag[1024];
  const size_t c = threadIdx.x;
  const int iter = *iter_d;
  if (c < k && st[c].active) {
    double rz = 0.0, rr = 0.0;
    for (size_t q = 0; q < nchunks; ++q) {
      rz += part[(0 * k + c) * nchunks + q];
      rr += part[(1 * k + c) * nchunks + q];
    }
    st[c].beta = rz / st[c].rz;
    st[c].rz = rz;
    st[c].rnorm = sqrt(rr);
    if (st[c].rnorm < st[c].tol_abs) {  // next top-of-loop test, decomposition.rs:242-250
      st[c].active = 0;
      st[c].converged = 1;
      st[c].iters = iter + 1;
    }
  }
  flag[threadIdx.x] = (c < k && st[c].active) ? 1 : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    int cnt = 0;
    for (size_t j = 0; j < k; ++j)
      if (flag[j]) n_active[1 + cnt++] = (int)j;
    n_active[0] = cnt;
    *iter_d = iter + 1;  // every thread has read it before the barrier above
  }
}

__global__ void pcg_set_tau_kernel(double* __restrict__ tau, double t0, double t1) {
  tau[0] = t0;
  tau[1] = t1;
}

// Active columns gathered into a contiguous block, and the product of the block scattered
// back: out[:, j] = in[:, idx[j]] (gather) or out[:, idx[j]] = in[:, j] (scatter).
__global__ void pcg_gather_kernel(const double* __restrict__ in, size_t ld, size_t n,
                                  const int* __restrict__ idx, int scatter,
                                  double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t j = blockIdx.y;
  if (i >= n) return;
  const size_t c = (size_t)idx[j];
  if (scatter) out[i + c * ld] = in[i + j * ld];
  else out[i + j * ld] = in[i + c * ld];
}

__global__ void pcg_p_kernel(double* __restrict__ p, const double* __restrict__ z, size_t ld,
                             size_t n, const ColState* __restrict__ st) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t c = blockIdx.y;
  if (i >= n || !st[c].active) return;
  p[i + c * ld] = z[i + c * ld] + st[c].beta * p[i + c * ld];
}

__global__ void pcg_final(size_t k, uint32_t max_iter, ColState* __restrict__ st) {
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k || !st[c].active) return;
  st[c].iters = (int)max_iter;
  st[c].converged = st[c].rnorm < st[c].tol_abs;
  st[c].active = 0;
}

static std::atomic<bool> g_pcg_graphs{true};
static std::atomic<uint64_t> g_graph_captures{0}, g_graph_replays{0}, g_graph_eager{0};
bool pcg_graphs_on() { return g_pcg_graphs.load(std::memory_order_relaxed); }

int pcg_sigma_device(saige_grm* h, ProductSink* sink, PcgWork& ws, const double* d_w,
                     const double tau[2], const double* d_b, size_t ldb, size_t k, double tol,
                     uint32_t max_iter, double* d_x, size_t ldx, saige_pcg_info* info,
                     int slices, const uint8_t* col_slices) {
  if (k == 0) return SAIGE_OK;
  SAIGE_ARG(k <= 1024, "k <= 1024 right-hand sides per solve");
  const size_t n = h->g.n;
  const size_t ld = round_up(n, 64);
  hipStream_t s = h->stream;
  const size_t nchunks = std::max<size_t>(1, ceil_div(n, kDotChunk));
  SAIGE_TRY(ws.vec.ensure(6 * ld * k * 8));
  SAIGE_TRY(ws.part.ensure(4 * k * nchunks * 8));
  const size_t list_bytes = round_up((k + 1) * sizeof(int), 256);
  SAIGE_TRY(ws.scal.ensure(list_bytes + k * sizeof(ColState) + 256));
  double* R = ws.vec.as<double>();
  double* Z = R + ld * k;
  double* P = Z + ld * k;
  double* AP = P + ld * k;
  double* PC = AP + ld * k;   // active columns of P, gathered
  double* APC = PC + ld * k;  // their product
  ColState* st = reinterpret_cast<ColState*>(ws.scal.as<char>() + list_bytes);
  int* d_nact = ws.scal.as<int>();  // [count, active column indices]
  // tau of this solve and the iteration counter, read by the iteration's kernels
  double* d_tau = reinterpret_cast<double*>(ws.scal.as<char>() + list_bytes + k * sizeof(ColState));
  int* d_iter = reinterpret_cast<int*>(d_tau + 2);
  double* part = ws.part.as<double>();
  double* pap_part = part + 3 * k * nchunks;  // p'Ap partials, apart from the update's
  const unsigned kb = (unsigned)ceil_div(k, 256);
  dim3 gchunk((unsigned)nchunks, (unsigned)k);

  pcg_set_tau_kernel<<<1, 1, 0, s>>>(d_tau, tau[0], tau[1]);
  pcg_init_kernel<<<gchunk, kDotBlock, 0, s>>>(d_b, ldb, d_x, ldx, R, Z, P, ld, d_w, d_tau, n,
                                               nchunks, part);
  pcg_init_fin<<<1, 1024, 0, s>>>(part, nchunks, k, tol, st, d_nact, d_iter);
  SAIGE_HIP(hipGetLastError());
  // [count, active list | column states] read back in one pinned copy per iteration
  const size_t state_bytes = list_bytes + k * sizeof(ColState);
  SAIGE_TRY(ws.host.ensure(state_bytes));
  int* list = ws.host.as<int>();
  ColState* hs = reinterpret_cast<ColState*>(ws.host.as<char>() + list_bytes);
  SAIGE_HIP(hipMemcpyAsync(list, d_nact, state_bytes, hipMemcpyDeviceToHost, s));
  SAIGE_HIP(hipStreamSynchronize(s));
  int n_active = list[0];

  // One iteration's launches at the current width (n_active): the GRM product of the active
  // columns (gathered when some have stopped), Sigma, p'Ap, the update and the next beta.
  // per-column slice requests of the active columns, in the gathered order (the host copy
  // of the active list is current whenever an iteration is enqueued)
  std::vector<uint8_t> act_sl;
  auto active_slices = [&](size_t width) -> const uint8_t* {
    if (!col_slices) return nullptr;
    act_sl.resize(width);
    for (size_t j = 0; j < width; ++j)
      act_sl[j] = col_slices[width == k ? j : (size_t)list[1 + j]];
    return act_sl.data();
  };
  auto enqueue = [&](size_t width) -> int {
    const uint8_t* sl = active_slices(width);
    if (width == k) {
      SAIGE_TRY(sink->product(P, ld, k, AP, ld, nullptr, nullptr, slices, sl));
    } else {
      // only the columns still iterating: their results do not depend on which other
      // columns share the product
      const dim3 gg((unsigned)ceil_div(n, 256), (unsigned)width);
      pcg_gather_kernel<<<gg, 256, 0, s>>>(P, ld, n, d_nact + 1, 0, PC);
      SAIGE_TRY(sink->product(PC, ld, width, APC, ld, nullptr, nullptr, slices, sl));
      pcg_gather_kernel<<<gg, 256, 0, s>>>(APC, ld, n, d_nact + 1, 1, AP);
    }
    const dim3 gn((unsigned)ceil_div(n, 256), (unsigned)k);
    pcg_sigma_pap_kernel<<<gchunk, kDotBlock, 0, s>>>(P, AP, ld, d_w, d_tau, n, nchunks, st,
                                                      pap_part);
    pcg_update_kernel<<<gchunk, kDotBlock, 0, s>>>(d_x, ldx, R, Z, P, AP, ld, d_w, d_tau, n,
                                                   nchunks, pap_part, d_iter, st, part);
    pcg_beta_fin<<<1, 1024, 0, s>>>(part, nchunks, k, d_iter, st, d_nact);
    pcg_p_kernel<<<gn, 256, 0, s>>>(P, Z, ld, n, st);
    SAIGE_HIP(hipGetLastError());
    return SAIGE_OK;
  };

  // Captured iterations: a graph per (block, width, buffers), captured the second time that
  // iteration shape is met with the DevBuf generation unchanged (the first run, eager, has
  // sized every workspace the product needs, so the capture allocates nothing) and replayed
  // while the generation stays.  Not with the kernel timer or the launch log on (their event
  // records are not captured), nor for sinks that join products across threads.
  const int tag = sink->graph_tag();
  const bool graphs = pcg_graphs_on() && tag >= 0 && !h->timer.on && !launch_log_on();
  for (uint32_t iter = 0; iter < max_iter && n_active > 0; ++iter) {
    const size_t width = (size_t)n_active;
    bool done = false;
    PcgGraph* g = nullptr;
    uint64_t key[12] = {};
    if (graphs) {
      const uint64_t gen = devbuf_generation();
      // the slice requests are part of the captured product: FNV-1a of the active columns'
      uint64_t slkey = (uint64_t)slices;
      if (const uint8_t* sl = active_slices(width)) {
        slkey = 1469598103934665603ull;
        for (size_t j = 0; j < width; ++j) slkey = (slkey ^ sl[j]) * 1099511628211ull;
      }
      const uint64_t kk[12] = {(uint64_t)(uintptr_t)h, (uint64_t)tag, slkey, k, width,
                               (uint64_t)(uintptr_t)d_x, ldx, (uint64_t)(uintptr_t)d_w,
                               (uint64_t)(uintptr_t)R, (uint64_t)(uintptr_t)part,
                               (uint64_t)(uintptr_t)st, n};
      std::memcpy(key, kk, sizeof(key));
      for (PcgGraph& c : ws.graphs)
        if (std::memcmp(c.key, key, sizeof(key)) == 0) g = &c;
      if (g && g->generation != gen) {  // a workspace moved since: forget the graph
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        g->exec = nullptr;
        g->generation = 0;
      }
      if (g && g->exec) {
        SAIGE_HIP(hipGraphLaunch(g->exec, s));
        sink->columns += width;
        h->products += width;
        g_graph_replays.fetch_add(1, std::memory_order_relaxed);
        done = true;
      } else if (g && g->generation == gen) {  // met before at this generation: capture
        hipGraph_t graph = nullptr;
        SAIGE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(width);
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc == SAIGE_OK && ec == hipSuccess && graph && devbuf_generation() == gen) {
          hipGraphExec_t exec = nullptr;
          if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
            g->exec = exec;
            SAIGE_HIP(hipGraphLaunch(exec, s));
            g_graph_captures.fetch_add(1, std::memory_order_relaxed);
            done = true;
          }
        }
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        if (!done) {  // run it eagerly below
          sink->columns -= rc == SAIGE_OK ? width : 0;  // the capture pass counted it
          h->products -= rc == SAIGE_OK ? width : 0;
          g->generation = 0;
        }
      }
    }
    if (!done) {
      SAIGE_TRY(enqueue(width));
      g_graph_eager.fetch_add(1, std::memory_order_relaxed);
      if (graphs) {  // remember the shape at the generation after its workspaces exist
        if (!g) {
          ws.graphs.emplace_back();
          g = &ws.graphs.back();
          std::memcpy(g->key, key, sizeof(key));
        }
        g->generation = devbuf_generation();
      }
    }
    SAIGE_HIP(hipMemcpyAsync(list, d_nact, state_bytes, hipMemcpyDeviceToHost, s));
    SAIGE_HIP(hipStreamSynchronize(s));
    n_active = list[0];
  }
  if (n_active > 0) {  // stopped at max_iter: the remaining columns' final states
    pcg_final<<<kb, 256, 0, s>>>(k, max_iter, st);
    SAIGE_HIP(hipGetLastError());
    if (info) {
      SAIGE_HIP(hipMemcpyAsync(hs, st, k * sizeof(ColState), hipMemcpyDeviceToHost, s));
      SAIGE_HIP(hipStreamSynchronize(s));
    }
  }
  if (info) {
    for (size_t c = 0; c < k; ++c) {
      info[c].iterations = (uint32_t)hs[c].iters;
      info[c].residual = hs[c].rnorm;
      info[c].converged = (uint8_t)hs[c].converged;
    }
  }
  return SAIGE_OK;
}

}  // namespace saige

using namespace saige;

extern "C" int saige_pcg_graphs(int enable) {
  const int prev = saige::g_pcg_graphs.load() ? 1 : 0;
  if (enable >= 0) saige::g_pcg_graphs.store(enable != 0);
  return prev;
}

extern "C" int saige_pcg_graph_counts(uint64_t* captured, uint64_t* replayed, uint64_t* eager) {
  if (captured) *captured = saige::g_graph_captures.load();
  if (replayed) *replayed = saige::g_graph_replays.load();
  if (eager) *eager = saige::g_graph_eager.load();
  return SAIGE_OK;
}

extern "C" int saige_pcg_solve_sigma(saige_grm* h, const double* w, const double tau[2],
                                     const double* b, size_t ldb, size_t k, double tol,
                                     uint32_t max_iter, double* x, size_t ldx,
                                     saige_pcg_info* info) {
  clear_error();
  SAIGE_ARG(h && w && tau && (k == 0 || (b && x)), "null pointer");
  SAIGE_ARG(ldb >= h->g.n && ldx >= h->g.n, "leading dimension < n");
  if (k == 0) return SAIGE_OK;
  std::lock_guard<std::recursive_mutex> lk(h->mu);
  SAIGE_HIP(hipSetDevice(h->device));
  const size_t n = h->g.n;
  SAIGE_TRY(h->pcg_w.ensure(n * 8));
  SAIGE_TRY(h->pcg_b.ensure(2 * n * k * 8));
  double* dw = h->pcg_w.as<double>();
  double* db = h->pcg_b.as<double>();
  double* dx = db + n * k;
  SAIGE_HIP(hipMemcpyAsync(dw, w, n * 8, hipMemcpyHostToDevice, h->stream));
  SAIGE_HIP(hipMemcpy2DAsync(db, n * 8, b, ldb * 8, n * 8, k, hipMemcpyHostToDevice, h->stream));
  DirectSink sink(h);
  SAIGE_TRY(pcg_sigma_device(h, &sink, h->pcg, dw, tau, db, n, k, tol, max_iter, dx, n, info));
  SAIGE_HIP(hipMemcpy2DAsync(x, ldx * 8, dx, n * 8, n * 8, k, hipMemcpyDeviceToHost, h->stream));
  SAIGE_HIP(hipStreamSynchronize(h->stream));
  return SAIGE_OK;
}
