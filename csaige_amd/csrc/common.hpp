// Internal helpers shared by the HIP translation units of libsaige_gpu.so.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/saige_gpu.h"

namespace saige {

void set_error(const char* fmt, ...);
void clear_error();

#define SAIGE_HIP(call)                                                              \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::saige::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                   \
                         hipGetErrorString(_e));                                     \
      return SAIGE_ERR_DEVICE;                                                       \
    }                                                                                \
  } while (0)

#define SAIGE_TRY(expr)                  \
  do {                                   \
    int _s = (expr);                     \
    if (_s != SAIGE_OK) return _s;       \
  } while (0)

#define SAIGE_ARG(cond, msg)                                        \
  do {                                                              \
    if (!(cond)) {                                                  \
      ::saige::set_error("invalid argument: %s (%s)", msg, #cond);  \
      return SAIGE_ERR_ARG;                                         \
    }                                                               \
  } while (0)

// Device memory cache.  A freed block is kept for its device and handed back to the next
// request of the same rounded size, so repeated fits, products and batches do not pay
// hipMalloc/hipFree (a 1.25 GB hipFree costs milliseconds).  Sizes round up to 64 KiB below
// 64 MiB and to 2 MiB above.  Handing out a cached block first waits for the device, so
// work enqueued before the block was freed cannot overlap its new use.  When hipMalloc
// fails, the device's cached blocks are released and the allocation is retried once.
// saige_release_cached_memory() empties the cache.
void* cache_alloc(size_t bytes, size_t* rounded, int* device);
void cache_free(void* p, size_t rounded, int device);
// Counts every allocation and release of a DevBuf in the process.  A captured HIP graph
// bakes in buffer addresses, so it is replayed only while this count is the one it was
// captured at.
uint64_t devbuf_generation();
void devbuf_generation_bump();

// Grow-only device buffer owned by a handle.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int device = 0;
  int ensure(size_t want) {
    if (want <= bytes) return SAIGE_OK;
    release();
    if (want == 0) return SAIGE_OK;
    size_t got = 0;
    p = cache_alloc(want, &got, &device);
    devbuf_generation_bump();
    if (!p) return SAIGE_ERR_NOMEM;
    bytes = got;
    return SAIGE_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  void release() {
    if (p) {
      cache_free(p, bytes, device);
      devbuf_generation_bump();
    }
    p = nullptr;
    bytes = 0;
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// Pinned host memory (hipHostMalloc) for the small device-to-host reads of the solver loops:
// a copy into it is one DMA transfer, with no staging through pageable memory.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return SAIGE_OK;
    release();
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return SAIGE_ERR_NOMEM;
    }
    bytes = want;
    return SAIGE_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  ~HostBuf() { release(); }
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
};

int use_device(int device);

// Optional per-handle timing of a handle's GEMM-class kernel launches.  Timed launches go
// through hipExtLaunchKernel with a start/stop event pair, so the events carry the
// dispatch's own begin and end timestamps; read() sums them (and resets).  At most
// kMaxPairs launches are recorded between reads.
struct KernelTimer {
  static constexpr size_t kMaxPairs = 4096;
  bool on = false;
  std::vector<hipEvent_t> ev;
  size_t used = 0;  // events handed out since the last read (always even)
  // Event pair for the next timed launch; null events when timing is off or full.
  int pair(hipEvent_t* a, hipEvent_t* b) {
    *a = *b = nullptr;
    if (!on || used / 2 >= kMaxPairs) return SAIGE_OK;
    while (ev.size() < used + 2) {
      hipEvent_t e;
      SAIGE_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    *a = ev[used];
    *b = ev[used + 1];
    used += 2;
    return SAIGE_OK;
  }
  int read(double* total_ms, uint64_t* launches) {
    double t = 0.0;
    if (used) SAIGE_HIP(hipEventSynchronize(ev[used - 1]));
    for (size_t i = 0; i + 1 < used; i += 2) {
      float ms = 0.f;
      SAIGE_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      t += ms;
    }
    if (total_ms) *total_ms = t;
    if (launches) *launches = used / 2;
    used = 0;
    return SAIGE_OK;
  }
  ~KernelTimer() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

// Process-wide launch log (diagnostic): while enabled, every GEMM-class launch of every
// handle (also the handles that the fit entry points create internally) gets a
// dispatch-stamped event pair and is filed under its operand width, the padded number of
// dense columns it multiplies.  Thread-safe; at most kMaxLaunches records between reads.
bool launch_log_on();
// The log's lock: held from handing out a launch's event pair until the launch is issued,
// so that a concurrent read or enable (which return every recorded pair to the pool) cannot
// recycle a pair whose launch has not been issued yet.
std::mutex& launch_log_mutex();
// What the launch log files a launch under: its operand width (the padded number of dense
// columns it multiplies) and the number of vectors those columns carry.
struct LaunchTag {
  uint32_t width, vectors;
};
// Event pair for a launch on the current device (null when the log is off or full); the
// caller holds launch_log_mutex().
int launch_log_pair_locked(LaunchTag tag, hipEvent_t* a, hipEvent_t* b);

// Launch a kernel on stream s, timed by t when its timing is on, else by the launch log
// when that is on.
template <class F, class... Args>
int timed_launch(KernelTimer& t, LaunchTag tag, F kernel, dim3 grid, dim3 block,
                 hipStream_t s, Args... args) {
  hipEvent_t a, b;
  SAIGE_TRY(t.pair(&a, &b));
  if (!a && launch_log_on()) {
    std::lock_guard<std::mutex> lk(launch_log_mutex());
    SAIGE_TRY(launch_log_pair_locked(tag, &a, &b));
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, a, b, 0, args...);
    SAIGE_HIP(hipGetLastError());
    return SAIGE_OK;
  }
  hipExtLaunchKernelGGL(kernel, grid, block, 0, s, a, b, 0, args...);
  SAIGE_HIP(hipGetLastError());
  return SAIGE_OK;
}

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }
inline size_t ceil_div(size_t x, size_t m) { return (x + m - 1) / m; }

// ----------------------------------------------------------------------------------
// Column dot products with a deterministic two-stage reduction: per-block partial sums
// (fixed tree inside the block) then a fixed-order sequential sum per column.
// ----------------------------------------------------------------------------------
constexpr int kDotBlock = 256;      // threads per block
constexpr int kDotChunk = 4096;     // samples per block

// Repacked genotype matrix for the GRM: SNP-major rows of row_bytes = n_pad / 4 bytes,
// samples in model order, padding samples and padding markers coded 01 (missing -> 0).
constexpr size_t kSamplePad = 512;
constexpr size_t kMarkerPad = 512;

struct GrmDevice {
  size_t n = 0, m = 0, n_pad = 0, m_pad = 0, row_bytes = 0;
  uint8_t* geno = nullptr;   // m_pad * row_bytes
  double* lut = nullptr;     // m_pad * 4 standardized values per code (fp64 path)
  float* scale = nullptr;    // m_pad: 1/sqrt(2p(1-p)) (0 if monomorphic)   (mixed path)
  float* shift = nullptr;    // m_pad: 2p * scale                           (mixed path)
  double* scale64 = nullptr; // m_pad
  double* shift64 = nullptr; // m_pad
  uint8_t* geno_t = nullptr; // mixed path: sample-major packed copy (n_pad rows of m_pad/4 bytes)
  int has_missing = 0;
};

}  // namespace saige

namespace saige {

// Where a fit sends its GRM products: straight to the handle, or to a batcher that joins
// the products of several concurrent fits (phenotypes sharing the GRM) into one
// multi-column product.  out = GRM v, or Sigma v = tau0 v / max(w, 1e-30) + tau1 GRM v when
// w is non-null.  `columns` counts the column-products this sink has served.
// `slices` asks the int8 forms for that many operand slices on these columns (3: the
// three-slice operand even on a SAIGE_PREC_INT8X2 handle; 0: the handle's own form); the
// other precisions ignore it.  A column's result depends only on its own slice count.
struct ProductSink {
  uint64_t columns = 0;
  // slices: int8 operand slices asked for every column (0: the handle's own form);
  // col_slices, when given, per column (k entries) instead
  virtual int product(const double* v, size_t ldv, size_t k, double* out, size_t ldo,
                      const double* w, const double* tau, int slices = 0,
                      const uint8_t* col_slices = nullptr) = 0;
  // Whether this sink's products may be captured in a HIP graph by the calling thread
  // (-1 no; otherwise a tag that distinguishes the product variants it issues).
  virtual int graph_tag() const { return -1; }
  // The next product's input equals that of any other product issued with the same nonzero
  // key (same columns, slices and left-out chromosome), so a sink that joins products may
  // compute it once.
  virtual void share_next(uint64_t key) { (void)key; }
  virtual ~ProductSink() = default;
};

// One captured PCG iteration: the launches of one iteration at one active width, for one
// block solve's buffers; replayed while the DevBuf generation is the one it was captured at.
struct PcgGraph {
  uint64_t key[12];
  uint64_t generation = 0;
  hipGraphExec_t exec = nullptr;  // null: seen once, not captured yet
};

// Workspace of one batched PCG solve (one per concurrent fit).
struct PcgWork {
  DevBuf vec, scal, part;
  HostBuf host;  // active-column list and column states read back each iteration
  std::vector<PcgGraph> graphs;
  void drop_graphs() {
    for (PcgGraph& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  ~PcgWork() { drop_graphs(); }
  PcgWork() = default;
  PcgWork(const PcgWork&) = delete;
  PcgWork& operator=(const PcgWork&) = delete;
};

// Process-wide switch of the PCG iteration graphs (on by default).
bool pcg_graphs_on();

}  // namespace saige

struct saige_grm {
  int device = 0;
  int precision = SAIGE_PREC_FP64;
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;
  saige::GrmDevice g;
  saige::DevBuf geno, lut, lut32, scale, shift, scale64, shift64, geno_t, miss_idx, miss_ptr;
  // reduced precision: per-marker 1/sd and 2p/sd (m_pad entries, 0 past m) on the host
  std::vector<double> host_scale, host_shift;
  // workspaces for products and solves
  saige::DevBuf ws_vt, ws_tpart, ws_t, ws_ypart, ws_in, ws_out;
  saige::DevBuf ws_vhi, ws_vlo, ws_tmp;
  // PCG workspaces of the handle's own solves
  saige::PcgWork pcg;
  saige::DevBuf pcg_b, pcg_w;
  uint64_t products = 0;   // column-products performed (diagnostic)
  saige::KernelTimer timer; // GEMM kernel launch timing (diagnostic)
  void* mixed = nullptr;   // mixed-precision state (grm_mixed.hip)
  // Leave-one-chromosome-out: chromosome code per marker (handle order) and M / M_c per
  // code (0 when a chromosome holds every marker); loco_req, set around one product call,
  // gives each column the chromosome its GRM leaves out (-1: none).
  saige::DevBuf loco_chrom, loco_scale;
  int loco_n_chrom = 0;
  const int* loco_req = nullptr;
  // Per column of the product call in progress: operand slices asked for (0: the handle's
  // own form); null for none.  Set around one product call, like loco_req.
  const uint8_t* slice_req = nullptr;
};

namespace saige {

// Columns of one product launch and the chromosome each leaves out of its GRM (-1: none).
constexpr int kLocoMaxCols = 256;
struct LocoCodes {
  int16_t code[kLocoMaxCols];
};
// Fills lc from h->loco_req[c0 .. c0 + kc); false when the product has no LOCO columns.
bool loco_codes(const saige_grm* h, size_t c0, size_t kc, LocoCodes& lc);
// The stage-1 result T (markers x columns, element (row, c) at t[row * rs + c * cs]) for
// LOCO: rows of the left-out chromosome become 0, the others are scaled by M / M_c, so the
// stage-2 product with the 1/M factor is the GRM of the kept markers.
void loco_mask(const saige_grm* h, double* t, size_t rs, size_t cs, size_t m, size_t kc,
               const LocoCodes& lc, hipStream_t s);

// GRM product on device: out[:, c] = scale_out * (Z Z^T v[:, c]) * (1/M) for c < k;
// v and out have leading dimensions ldv / ldo (>= n).  Rows >= n of out are not written.
// If w_tau is non-null the epilogue writes Sigma*v instead:
//   out = tau0 * v / max(w, 1e-30) + tau1 * GRM v.
int grm_product(saige_grm* h, const double* d_v, size_t ldv, size_t k, double* d_out, size_t ldo,
                const double* d_w = nullptr, const double* tau = nullptr);

// Column dots: out[c] = sum_i a[i + c*lda] * b[i + c*ldb] for c < k (host result).
int col_dots(hipStream_t s, const double* a, size_t lda, const double* b, size_t ldb, size_t n,
             size_t k, double* h_out, DevBuf& part, DevBuf& res);
// Gram block: out[a_i + na*b_j] = dot(A[:, a_i], B[:, b_j]) (host result, na x nb col-major).
int gram(hipStream_t s, const double* A, size_t lda, size_t na, const double* B, size_t ldb,
         size_t nb, size_t n, double* h_out, DevBuf& part, DevBuf& res);
// Column dots into device memory (d_out[c], same arithmetic as col_dots), no host wait; with
// read_doubles, several dot products share one device-to-host read.
int col_dots_dev(hipStream_t s, const double* a, size_t lda, const double* b, size_t ldb,
                 size_t n, size_t k, double* d_out, DevBuf& part);
int read_doubles(hipStream_t s, const double* d, size_t nq, double* h_out);

// The handle itself as a product sink; with loco >= 0 its products leave that chromosome
// out of the GRM.
struct DirectSink : ProductSink {
  saige_grm* h;
  int loco = -1;
  explicit DirectSink(saige_grm* hh) : h(hh) {}
  int graph_tag() const override { return loco < 0 ? 0 : 1 + loco; }
  int product(const double* v, size_t ldv, size_t k, double* out, size_t ldo, const double* w,
              const double* tau, int slices = 0, const uint8_t* col_slices = nullptr) override {
    columns += k;
    std::vector<int> codes(loco < 0 ? 0 : k, loco);
    std::vector<uint8_t> sl(slices ? k : 0, (uint8_t)slices);
    if (col_slices) sl.assign(col_slices, col_slices + k);
    h->loco_req = loco < 0 ? nullptr : codes.data();
    h->slice_req = sl.empty() ? nullptr : sl.data();
    const int rc = grm_product(h, v, ldv, k, out, ldo, w, tau);
    h->loco_req = nullptr;
    h->slice_req = nullptr;
    return rc;
  }
};

// Batched PCG on Sigma (device pointers; B, X: n x k with leading dimension ld), products
// through `sink`, vectors in `ws`; int8 operand slices for every column (slices) or per
// column of the block (col_slices, k entries).
int pcg_sigma_device(saige_grm* h, ProductSink* sink, PcgWork& ws, const double* d_w,
                     const double tau[2], const double* d_b, size_t ldb, size_t k, double tol,
                     uint32_t max_iter, double* d_x, size_t ldx, saige_pcg_info* info,
                     int slices = 0, const uint8_t* col_slices = nullptr);

}  // namespace saige
