#!/usr/bin/env python3
"""Benchmark: fit-null wall-seconds on one MI355X, mixed precision vs plain FP64 at equal
accuracy, plus the GRM product roofline, step-2 variants/s and a CPU baseline.

Workload (default): 10,000 nuclear families (2 HWE founders + 3 Mendelian offspring,
n = 50,000), 100,000 markers with MAF ~ U(0.01, 0.5), binary trait at 5% prevalence from a
liability with h2 = 0.4 plus two N(0,1) covariates; fit-null with 30 Hutchinson probes, so
every AI-REML iteration solves 1 + p + 30 = 34 right-hand sides as one block.  Synthetic,
seeded data; the .bed body is resident in HBM before the timed region.

A "step" is one whole fit-null (QC, VR/GRM split, GRM build, AI-REML, XVX^-1X'V and the
variance ratio) over that cohort.  `value` is the wall-seconds of one step of the mixed-
precision path with the int8 MFMA GRM product; the bf16 hi+lo form and the plain FP64 path
are timed beside it, and the null models are checked to give the same tau, variance ratio
and step-2 p-values.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fit-null wall-s + step-2 variants/s, 1 GPU: mixed vs FP64 at equal-accuracy p"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
I8_PEAK_TOPS = 5000.0
# HBM bytes per int8 GEMM launch at k = 34, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# (one counter per run; FETCH_SIZE scaled by 2.0 from a 1 GiB copy) for the default families
# workload: mean of the marker-side and sample-side launches.  Keyed by (families, markers,
# int8 form).
MEASURED_TRAFFIC = {(10000, 100000, "int8x2"): 1.45e9}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host the CPU baseline ran on: CPU model, logical CPUs, and the CPUs this process
    may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def time_product(sg, hip, d_bed, bpm, n, markers, af, precision, k, reps):
    """Average device time of one k-column GRM product, HIP events on the handle's stream."""
    L = sg.load()
    h = C.c_void_p()
    sg.check(L.saige_grm_create_device(0, C.c_void_p(d_bed.value), bpm, n, None, n,
                                       markers.ctypes.data_as(sg._lib.u32p), len(markers),
                                       np.ascontiguousarray(af).ctypes.data_as(sg._lib.dp),
                                       precision, C.byref(h)))
    d_v = hip.DeviceBuffer(n * k * 8)
    d_o = hip.DeviceBuffer(n * k * 8)
    d_v.upload(np.random.default_rng(0).standard_normal((k, n)))
    stream = L.saige_grm_stream(h)
    sg.check(L.saige_grm_mat_mat_device(h, d_v.value, n, k, d_o.value, n))
    sg.check(L.saige_grm_set_kernel_timing(h, 1))
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(reps):
        sg.check(L.saige_grm_mat_mat_device(h, d_v.value, n, k, d_o.value, n))
    e1.record(stream)
    ms = e0.elapsed_ms(e1) / reps
    kt, kl = C.c_double(), C.c_uint64()
    sg.check(L.saige_grm_kernel_timing(h, C.byref(kt), C.byref(kl)))
    L.saige_grm_destroy(h)
    return ms, kt.value / max(kl.value, 1), kl.value


def run_loco(args):
    """LOCO (glmm/loco.rs:54-114) on the families cohort: one AI-REML refit per chromosome
    (22 chromosomes of equal marker count), the refits' GRM products joined; the int8 MFMA
    path (both operand-slice forms) timed against the plain FP64 path and against one full-GRM
    fit; one JSON line."""
    import csaige_amd as sg
    from csaige_amd import synth
    t0 = time.time()
    c = synth.families(n_fam=args.n_fam, kids=3, m=args.markers, seed=2, h2=0.4, prevalence=0.05)
    n, m = c.n, c.bed.shape[0]
    n_chrom = 22
    chrom = (np.arange(m) * n_chrom // m).astype(np.int32)
    ac, nv, af, mac = sg.marker_stats(c.bed, c.bpm, n, m)
    log(f"loco: cohort n={n} m={m}, {n_chrom} chromosomes ({time.time() - t0:.1f}s)")
    forms = {"int8": sg.PREC_INT8, "int8x2": sg.PREC_INT8X2, "fp64": sg.PREC_FP64}
    head = args.precision
    other = "int8x2" if head == "int8" else "int8"
    out = {}
    for tag in [head, other] + ([] if args.no_fp64 else ["fp64"]):
        prec = forms[tag]
        g = sg.OnTheFlyGrm(c.bed, c.bpm, n, af, precision=prec)
        g.set_marker_chromosomes(chrom)
        sg.fit_ai_reml(g, c.y, c.x, tol=1e-5)  # warm-up: kernels, device memory cache
        t = time.perf_counter()
        full = sg.fit_ai_reml(g, c.y, c.x, tol=1e-5)
        dt_full = time.perf_counter() - t
        t = time.perf_counter()
        r = g.fit_ai_reml_loco(c.y, c.x, tol=1e-5)
        dt = time.perf_counter() - t
        log(f"loco {tag}: {n_chrom} refits in {dt:.2f}s ({dt / n_chrom:.3f}s per refit); one "
            f"full-GRM fit {dt_full:.3f}s; tau_g range {min(x['tau'][1] for x in r):.6g}.."
            f"{max(x['tau'][1] for x in r):.6g}, full {full['tau'][1]:.6g}")
        out[tag] = (dt, dt_full, r)
        g.close()
    dt8, full8, r8 = out[head]
    acc = None
    if "fp64" in out:
        rf = out["fp64"][2]
        acc = {}
        for tag in (head, other):
            ri = out[tag][2]
            acc[tag] = {"max_tau_rel_diff": float(max(abs(a["tau"][1] - b["tau"][1]) / b["tau"][1]
                                                      for a, b in zip(ri, rf))),
                        "iterations_equal": bool(all(a["iterations"] == b["iterations"]
                                                     for a, b in zip(ri, rf)))}
        log(f"loco accuracy vs fp64: {acc}")
    line = {
        "metric": "LOCO refits wall-s, 1 GPU: int8 mixed vs FP64",
        "value": dt8, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": dt8 * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"int8 MFMA ({'two' if head == 'int8x2' else 'three'} operand slices) / int32 "
                 "accumulate / fp64 solver",
        "data": "synthetic (seeded families cohort)",
        "config": {"workload": f"LOCO AI-REML refits, {args.n_fam} nuclear families x 5 = {n} "
                               f"samples, {m} markers in {n_chrom} chromosomes, binary 5%",
                   "precision": head, f"{other}_wall_s": out[other][0],
                   "per_refit_s": dt8 / n_chrom, "full_grm_fit_s": full8,
                   "fp64_wall_s": out["fp64"][0] if "fp64" in out else None,
                   "speedup_vs_fp64": out["fp64"][0] / dt8 if "fp64" in out else None,
                   "tau_g": [float(x["tau"][1]) for x in r8],
                   "iterations": [int(x["iterations"]) for x in r8],
                   "accuracy": acc},
    }
    print(json.dumps(line), flush=True)


def run_region(args):
    """Region tests (score_test/region.rs): BURDEN, SKAT and SKAT-O for 400 genes of 25
    variants at n = 50,000 (intercept-only null, 5% cases), f64 dosage rows resident in HBM
    (the same call from host rows, which copies them in, is timed beside it); the oracle's
    restatement on one core for a sample of the genes as the CPU baseline; one JSON line."""
    import csaige_amd as sg
    import oracle
    from csaige_amd import region as R, synth
    from oracle import region as OR
    n, n_genes, per_gene = 50000, 400, 25
    nf = synth.null_fit(n=n, p=1, eta_mean=float(np.log(0.05 / 0.95)), eta_sd=0.5, seed=7)
    a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    nv = n_genes * per_gene
    q = np.empty((nv, n), np.uint8)
    synth.dosage_u8(n, 0, nv, seed=8, maf=(5e-4, 0.05), out=q)
    G = q.astype(np.float64) / 127.0
    del q
    af = G.mean(axis=1) / 2
    w = R.beta_weights(af)
    from csaige_amd import hip
    e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.3), nf.x, a, 1.0, y=nf.y)
    d_g = hip.DeviceBuffer(G.nbytes)
    d_g.upload(G)  # resident in HBM before the timed call
    sizes = [per_gene] * n_genes
    e.test_regions(None, [per_gene], w[:per_gene], device_rows=(d_g.value, n, per_gene))  # warm
    t = time.perf_counter()
    res = e.test_regions(None, sizes, w, device_rows=(d_g.value, n, nv))
    dt = time.perf_counter() - t
    t = time.perf_counter()
    res_h = e.test_regions(G, sizes, w)
    dt_host = time.perf_counter() - t
    same = bool(np.array_equal(res["pvalue"], res_h["pvalue"]))
    log(f"region: {n_genes} genes x {per_gene} variants at n={n}: {dt:.3f}s with the f64 "
        f"dosages resident ({n_genes / dt:.0f} genes/s); {dt_host:.3f}s from host rows "
        f"({n_genes / dt_host:.0f} genes/s, upload inside the call; same p-values: {same}); "
        f"min SKAT-O p {res['pvalue'].min():.3g}")
    k_cpu = 8
    t = time.perf_counter()
    cpu = [OR.run_region_test(G[g * per_gene:(g + 1) * per_gene], w[g * per_gene:(g + 1) * per_gene],
                              nf.mu * (1 - nf.mu), nf.y - nf.mu, nf.x, a, 1.0,
                              R.default_rho_values()) for g in range(k_cpu)]
    dt_cpu = (time.perf_counter() - t) / k_cpu
    dlp = max(abs(np.log10(res["pvalue"][g]) - np.log10(cpu[g][0])) for g in range(k_cpu))
    log(f"region cpu: {dt_cpu:.3f}s per gene on 1 thread (oracle restatement); GPU vs oracle "
        f"max |d log10 p| {dlp:.2e} over {k_cpu} genes")
    line = {
        "metric": "region tests genes/s, 1 GPU (BURDEN + SKAT + SKAT-O)",
        "value": n_genes / dt, "unit": "genes/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded null fit, uint8-quantized "
                                                     "dosages as f64)",
        "config": {"workload": f"{n_genes} genes x {per_gene} variants, n={n}, "
                               f"Beta(MAF;1,25) weights, 11 rho values",
                   "dosages": "f64 rows resident in HBM (saige_region_tests_device)",
                   "host_rows_genes_per_s": n_genes / dt_host,
                   "host_rows_same_pvalues": same,
                   "max_abs_dlog10p_vs_oracle": float(dlp)},
        "cpu_baseline": {"value": 1.0 / dt_cpu, "unit": "genes/s", "cores": 1, "kind": "port",
                         "sample": f"oracle region restatement (numpy, explicit g~) on "
                                   f"{k_cpu} genes"},
    }
    print(json.dumps(line), flush=True)


def run_biobank(args):
    """Several phenotypes at biobank scale: 400,000 samples (15% in half-sibling pairs),
    150,000 markers resident in HBM, 4 binary traits at 0.5/1/5/20% prevalence fitted
    together (joined multi-column GRM products, per-trait W), int8 MFMA path (both
    operand-slice forms) timed against the plain FP64 path; one JSON line."""
    import csaige_amd as sg
    from csaige_amd import hip, synth
    t0 = time.time()
    prev = (0.005, 0.01, 0.05, 0.20)
    log(f"generating biobank cohort n={args.bb_n} m={args.bb_m}")
    c = synth.biobank(n=args.bb_n, m=args.bb_m, frac_related=0.15, prevalences=prev, seed=5)
    n, m = c.n, c.bed.shape[0]
    log(f"cohort generated in {time.time() - t0:.1f}s: {c.bed.nbytes / 1e9:.2f} GB packed, "
        f"{c.n_pairs} half-sib pairs, cases {c.y.sum(axis=1).astype(int).tolist()}")
    idx = np.arange(n, dtype=np.uint32)
    d_bed = hip.DeviceBuffer(c.bed.nbytes)
    d_bed.upload(c.bed)
    hip.device_synchronize()

    def fit(precision, which=None):
        Y = c.y if which is None else c.y[which:which + 1]
        hip.device_synchronize()
        t = time.perf_counter()
        r = sg.fit_null_multi(None, c.bpm, n, m, Y, c.x, sample_idx=idx, precision=precision,
                              d_bed=d_bed.value)
        hip.device_synchronize()
        return time.perf_counter() - t, r

    forms = {"int8": sg.PREC_INT8, "int8x2": sg.PREC_INT8X2}
    head = args.precision
    other = "int8x2" if head == "int8" else "int8"
    dt8, r8 = fit(forms[head])
    for t, r in enumerate(r8):
        log(f"{head} joined: trait {t} tau={r['tau'][1]:.8g} it={r['iterations']} "
            f"vr={r['vr']:.6g} pcg={r['pcg_iterations']} products={r['grm_products']}")
    log(f"{head} joined fit of {len(prev)} traits: {dt8:.2f}s")
    dto, ro = fit(forms[other])
    log(f"{other} joined fit of {len(prev)} traits: {dto:.2f}s")
    sep = None
    if args.bb_separate:
        sep = 0.0
        for t in range(len(prev)):
            dt, _ = fit(forms[head], t)
            sep += dt
        log(f"{head} one trait at a time: {sep:.2f}s")
    dtf, rf = (None, None)
    acc = None
    if not args.no_fp64:
        dtf, rf = fit(sg.PREC_FP64)
        log(f"fp64 joined fit: {dtf:.2f}s")
        acc = {head: [], other: []}
        ntest = 2048

        def pvals(r, t):
            e = sg.ScoreTestEngine(r["mu"], c.y[t] - r["mu"], r["tau"], c.x, r["xvx_inv_xv"],
                                   r["vr"], y=c.y[t])
            p = e.test_markers_plink(c.bed[:ntest], c.bpm, n, ntest)["pvalue"]
            e.close()
            return p

        for t in range(len(prev)):
            pf = pvals(rf[t], t)
            mask = pf >= 1e-10
            for tag, rr in ((head, r8), (other, ro)):
                pm = pvals(rr[t], t)
                acc[tag].append({
                    "tau_rel_diff": float(abs(rr[t]["tau"][1] - rf[t]["tau"][1]) / rf[t]["tau"][1]),
                    "vr_rel_diff": float(abs(rr[t]["vr"] - rf[t]["vr"]) / rf[t]["vr"]),
                    "max_abs_dlog10p": float(np.max(np.abs(np.log10(pf[mask]) -
                                                           np.log10(pm[mask]))))})
        log(f"accuracy vs fp64: {acc}")
    line = {
        "metric": "fit-null wall-s, several phenotypes, 1 GPU: int8 mixed vs FP64",
        "value": dt8, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": dt8 * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"int8 MFMA ({'two' if head == 'int8x2' else 'three'} operand slices) / int32 "
                 "accumulate / fp64 solver",
        "data": "synthetic (seeded biobank cohort; packed genotypes resident in HBM)",
        "config": {"workload": f"fit-null of {len(prev)} binary traits (prevalence "
                               f"{'/'.join(str(p) for p in prev)}), {n} samples (15% in half-sib "
                               f"pairs), {m} markers, 30 probes per trait, joined products",
                   "precision": head, f"{other}_wall_s": dto,
                   "fp64_wall_s": dtf, "speedup_vs_fp64": (dtf / dt8) if dtf else None,
                   f"{head}_one_trait_at_a_time_s": sep,
                   "tau_g": [float(r["tau"][1]) for r in r8],
                   "ai_reml_iterations": [int(r["iterations"]) for r in r8],
                   "pcg_iterations": [int(r["pcg_iterations"]) for r in r8],
                   "variance_ratio": [float(r["vr"]) for r in r8], "accuracy": acc},
    }
    print(json.dumps(line), flush=True)


def run_learned(args):
    """Learned SPA root proposal: train the MLP on ~2M saddle-point roots computed by the SPA
    kernel itself (n = 50,000, 1% and 10% prevalence, every variant SPA'd), then compare
    Newton iterations, p-values and SPA time with and without the proposal on held-out
    variants; one JSON line."""
    import csaige_amd as sg
    from csaige_amd import hip, synth
    n, B = args.ls_n, 4096
    nb = max(1, args.ls_variants // B)
    xs, ys = [], []
    t0 = time.time()
    q = np.empty((B, n), np.uint8)
    for prev, seed in ((0.01, 11), (0.10, 21)):
        eta = float(np.log(prev / (1 - prev)))
        nf = synth.null_fit(n=n, p=1, eta_mean=eta, eta_sd=0.5, seed=seed)
        a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
        e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.0, y=nf.y,
                               spa_pval_cutoff=1.0, spa_tol=1e-10)
        for b in range(nb):
            synth.dosage_u8(n, b * B, B, seed=seed + 1, maf=(5e-4, 0.5), out=q)
            x, y = sg.spa_training_set(e.test_markers_u8(q))
            xs.append(x)
            ys.append(y)
            if b % 32 == 31:
                log(f"prevalence {prev}: {b + 1}/{nb} batches, {sum(len(v) for v in ys)} roots")
        e.close()
    X, Y = np.concatenate(xs), np.concatenate(ys)
    t_gen = time.time() - t0
    log(f"{len(Y)} training roots in {t_gen:.1f}s")
    m = sg.SpaRootMlp(seed=5)
    t1 = time.time()
    hist = m.train(X, Y, epochs=args.ls_epochs, batch=4096, lr=2e-3, seed=7)
    t_train = time.time() - t1
    log(f"trained {args.ls_epochs} epochs in {t_train:.1f}s: loss {hist[0]:.4f} -> {hist[-1]:.4f}")

    def evaluate(tag, n_e, prev, seed, nbat, cutoff):
        eta = float(np.log(prev / (1 - prev)))
        nf = synth.null_fit(n=n_e, p=1, eta_mean=eta, eta_sd=0.5, seed=seed)
        a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
        qe = np.empty((nbat * B, n_e), np.uint8)
        synth.dosage_u8(n_e, 0, nbat * B, seed=seed + 1, maf=(1e-4, 0.5), out=qe)
        d_q = hip.DeviceBuffer(qe.nbytes)
        d_q.upload(qe)
        del qe
        rsize = C.sizeof(sg._lib.VariantResult)
        d_r = hip.DeviceBuffer(nbat * B * rsize)
        out = {}
        for mode in ("plain", "mlp"):
            e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                                   spa_pval_cutoff=cutoff, spa_tol=1e-10)
            if mode == "mlp":
                e.set_root_proposal(m)
            e.test_markers_u8_device(d_q.value, n_e, B, 1.0 / 127, d_r.value)  # warm
            hip.device_synchronize()
            s0, s1 = hip.Event(), hip.Event()
            s0.record(e.stream)
            for b in range(nbat):  # resident batches, timed on the engine's stream
                e.test_markers_u8_device(d_q.value + b * B * n_e, n_e, B, 1.0 / 127,
                                         d_r.value + b * B * rsize)
            s1.record(e.stream)
            ms = s0.elapsed_ms(s1)
            res = (sg._lib.VariantResult * (nbat * B))()
            host = np.empty(nbat * B * rsize, np.uint8)
            d_r.download(host)
            C.memmove(C.addressof(res), host.ctypes.data, host.nbytes)
            out[mode] = (sg.results_to_dict(res), ms * 1e-3)
            e.close()
        d_q.free()
        r0, r1 = out["plain"][0], out["mlp"][0]
        gated = r0["pvalue_noadj"] < cutoff
        ok = gated & (r0["spa_converged"] == 1)
        it0 = float((r0["spa_iters1"][ok] + r0["spa_iters2"][ok]).mean())
        it1 = float((r1["spa_iters1"][ok] + r1["spa_iters2"][ok]).mean())
        d = float(np.abs(np.log10(r0["pvalue"][ok]) - np.log10(r1["pvalue"][ok])).max())
        lost = gated & (r0["spa_converged"] == 0)
        # the plain method's outcome on every gated variant: the same convergence and SPA
        # flags, and the same (unadjusted) p where the plain search diverges
        diff = gated & ((r0["spa_converged"] != r1["spa_converged"]) |
                        (r0["is_spa"] != r1["is_spa"]))
        diff |= lost & (r0["pvalue"] != r1["pvalue"])
        res = {"workload": tag, "variants": int(nbat * B), "spa_variants": int(gated.sum()),
               "mean_newton_iterations_plain": it0, "mean_newton_iterations_mlp": it1,
               "max_abs_dlog10p_vs_plain": d,
               "plain_diverged": int(lost.sum()),
               "outcome_differences_vs_plain": int(diff.sum()),
               "step2_device_s_plain": out["plain"][1], "step2_device_s_mlp": out["mlp"][1]}
        log(f"{tag}: {res}")
        return res

    ev = [evaluate("held-out n=50,000 at 3% prevalence, SPA where p < 0.05", n, 0.03, 31, 16, 0.05),
          evaluate("step-2 shape n=400,000 at 1% prevalence, SPA where |z| > 2", 400000, 0.01, 3, 4,
                   0.04550026389635842)]
    line = {
        "metric": "learned SPA root proposal: mean Newton iterations and max |d -log10 p| vs plain",
        "value": ev[0]["mean_newton_iterations_mlp"], "unit": "Newton iterations per variant",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": None, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 weights / fp32 MLP; fp64 Newton",
        "data": "synthetic (roots computed by the SPA kernel on seeded null fits and dosages)",
        "config": {"workload": f"MLP 6-64-64-1 trained on {len(Y)} roots (n={n}, prevalence 1% "
                               f"and 10%, every variant SPA'd, tol 1e-10), {args.ls_epochs} epochs",
                   "roots_seconds": t_gen, "train_seconds": t_train,
                   "loss_first_last": [float(hist[0]), float(hist[-1])], "evaluation": ev},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-fam", type=int, default=10000)
    ap.add_argument("--markers", type=int, default=100000)
    ap.add_argument("--product-reps", type=int, default=5)
    ap.add_argument("--step2-n", type=int, default=400000)
    ap.add_argument("--step2-batches", type=int, default=4)
    ap.add_argument("--cpu-markers", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-step2-variants", type=int, default=1024,
                    help="variants of the step-2 batches the oracle tests on one core")
    ap.add_argument("--step2-reps", type=int, default=5,
                    help="passes over the resident step-2 batches inside the timed region")
    ap.add_argument("--cpu-threads", type=int,
                    default=max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", "16")))),
                    help="host threads of the CPU baseline (the box's CPU share)")
    ap.add_argument("--no-step2", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--precision", default="int8x2", choices=["int8", "int8x2", "bf16", "fp64"],
                    help="int8 form timed for the headline value: three operand slices (int8) "
                         "or two (int8x2); the other form is timed beside it")
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per GEMM launch measured with rocprofv3 counters")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device 3 times (counter calibration)")
    ap.add_argument("--library-int8", action="store_true",
                    help="only time the vendor int8 GEMM (torch._int_mm) at the product's shapes")
    ap.add_argument("--read-bandwidth", action="store_true",
                    help="only time plain streaming reads and copies of device memory")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "families", "biobank", "learned", "loco", "region",
                             "product"],
                    help="headline (default): 4 traits at 400,000 x 150,000 fitted together "
                         "plus step 2 over 1,000,000 variants; families: one trait at 50,000 x "
                         "100,000; biobank: the 4-trait fit alone; learned: the SPA root "
                         "proposal; loco: leave-one-chromosome-out refits; region: region tests")
    ap.add_argument("--step2-variants", type=int, default=1000000,
                    help="headline: step-2 variants streamed (generated on the device)")
    ap.add_argument("--step2-engines", type=int, default=2,
                    help="headline: score-test engines (streams and host threads) of step 2")
    ap.add_argument("--step2-resident", type=int, default=64,
                    help="headline: step-2 batches generated and resident per timed chunk")
    ap.add_argument("--cpu-fit-markers", type=int, default=500,
                    help="headline: markers of the CPU baseline's sampled f32 GRM product")
    ap.add_argument("--prod-k", type=int, default=34,
                    help="product: dense columns per product")
    ap.add_argument("--prod-biobank", action="store_true",
                    help="product: the 400,000 x 150,000 biobank cohort")
    ap.add_argument("--no-graphs", action="store_true",
                    help="run the PCG iterations without HIP graph capture")
    ap.add_argument("--no-families", action="store_true",
                    help="headline: skip the families extra keys")
    ap.add_argument("--ls-n", type=int, default=50000)
    ap.add_argument("--ls-variants", type=int, default=524288,
                    help="learned: variants per prevalence for training (2 roots each)")
    ap.add_argument("--ls-epochs", type=int, default=20)
    ap.add_argument("--bb-n", type=int, default=400000)
    ap.add_argument("--bb-m", type=int, default=150000)
    ap.add_argument("--bb-separate", action="store_true",
                    help="biobank: also time the traits one at a time")
    args = ap.parse_args()
    if args.copy_calibration:
        from csaige_amd import hip
        a, b = hip.DeviceBuffer(1 << 30), hip.DeviceBuffer(1 << 30)
        a.zero()
        for _ in range(3):
            b.copy_from(a)
        hip.device_synchronize()
        log("copied 3 x 1 GiB device to device")
        return
    if args.read_bandwidth:
        read_bandwidth_calibration()
        return
    if args.library_int8:
        library_int8_calibration()
        return
    if args.no_graphs:
        import csaige_amd as sg
        sg.pcg_graphs(False)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the multi-GPU "
              "sharding of the sample axis is a design, not built code", file=sys.stderr)
        sys.exit(2)
    if args.workload == "biobank":
        run_biobank(args)
        return
    if args.workload == "learned":
        run_learned(args)
        return
    if args.workload == "loco":
        run_loco(args)
        return
    if args.workload == "region":
        run_region(args)
        return
    if args.workload == "families":
        run_families(args)
        return
    if args.workload == "product":
        run_product(args)
        return
    run_headline(args)


def read_bandwidth_calibration():
    """What a plain streaming kernel reaches on this GPU: torch's sum over 15 GiB of int64
    (a read of the size of the 400,000 x 150,000 packed genotypes) and a 1 GiB device copy,
    as the practical ceiling beside the 8 TB/s peak."""
    import torch
    x = torch.ones(15 * 2**30 // 8, dtype=torch.int64, device="cuda")
    y = torch.ones(2**30 // 8, dtype=torch.int64, device="cuda")
    z = torch.empty_like(y)

    def timed(f, reps):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms = timed(lambda: x.sum(), 5)
    log(f"read 15 GiB (torch sum): {ms:.3f} ms, {15 * 2**30 / ms / 1e6:.0f} GB/s "
        f"({15 * 2**30 / ms / 1e6 / HBM_PEAK_GBS:.3f} of {HBM_PEAK_GBS:.0f})")
    ms = timed(lambda: z.copy_(y), 20)
    log(f"copy 1 GiB: {ms:.3f} ms, {2 * 2**30 / ms / 1e6:.0f} GB/s read + write")


def library_int8_calibration():
    """The vendor int8 GEMM (hipBLASLt through torch._int_mm) at the genotype product's
    shapes, as a point of comparison for i8_gemm_kernel: 148,992 rows (the biobank's GRM
    markers) by 65,536 reduction indices (a slice of the sample axis; the rate does not fall
    with longer reductions) by 80 to 512 columns, and two square shapes."""
    import torch
    shapes = [(8192, 8192, 8192), (16384, 16384, 16384)] + [
        (148992, 65536, n) for n in (80, 112, 224, 288, 512)]
    for m, k, n in shapes:
        a = torch.randint(-3, 3, (m, k), dtype=torch.int8, device="cuda")
        b = torch.randint(-100, 100, (n, k), dtype=torch.int8, device="cuda").t()
        torch._int_mm(a, b)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            torch._int_mm(a, b)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tops = 2.0 * m * k * n / ms / 1e9
        log(f"torch._int_mm m={m} k={k} n={n}: {ms:.3f} ms, {tops:.0f} TOP/s "
            f"({tops / I8_PEAK_TOPS:.3f} of int8 peak)")
        del a, b


def run_product(args):
    """GRM products only (for counter profiles): the families cohort (or, with --prod-biobank,
    the 400,000 x 150,000 biobank cohort), one handle of the chosen int8 form, --product-reps
    products of --prod-k dense random columns; one JSON line with the launch log per width."""
    import csaige_amd as sg
    from csaige_amd import hip, synth
    if args.prod_biobank:
        c = synth.biobank(n=args.bb_n, m=args.bb_m, frac_related=0.15, seed=5)
    else:
        c = synth.families(n_fam=args.n_fam, kids=3, m=args.markers, seed=2, h2=0.4,
                           prevalence=0.05)
    n, m = c.n, c.bed.shape[0]
    d_bed = hip.DeviceBuffer(c.bed.nbytes)
    d_bed.upload(c.bed)
    ac, nv, af, mac = sg.marker_stats(c.bed, c.bpm, n, m)
    markers = np.arange(m, dtype=np.uint32)
    prec = {"int8": sg.PREC_INT8, "int8x2": sg.PREC_INT8X2, "bf16": sg.PREC_MIXED,
            "fp64": sg.PREC_FP64}[args.precision]
    k = args.prod_k
    sg.launch_log_enable(True, 0)
    ms, kms, kl = time_product(sg, hip, d_bed, c.bpm, n, markers, af, prec, k, args.product_reps)
    summ = sg.launch_log_summary()
    sg.launch_log_enable(False)
    line = {"workload": f"GRM product, n={n} x M={m}, k={k}, {args.precision}",
            "product_ms": ms, "gemm_launch_ms": kms, "gemm_launches": kl,
            "per_width": {str(w): {"launches": v[1], "ms_per_launch": v[0] / max(v[1], 1)}
                          for w, v in summ.items()},
            "hbm_gbs_per_launch": float((n + 3) // 4) * m / (kms * 1e-3) / 1e9}
    print(json.dumps(line), flush=True)


def run_families(args):
    """Fit-null of the families cohort (50,000 samples x 100,000 markers, one binary trait):
    int8 MFMA forms, bf16 and FP64 timed, one JSON line."""
    import csaige_amd as sg
    from csaige_amd import hip, synth

    t0 = time.time()
    log(f"generating families n_fam={args.n_fam} markers={args.markers}")
    c = synth.families(n_fam=args.n_fam, kids=3, m=args.markers, seed=2, h2=0.4, prevalence=0.05)
    n, m = c.n, c.bed.shape[0]
    log(f"cohort n={n} m={m} cases={int(c.y.sum())} ({time.time() - t0:.1f}s)")
    idx = np.arange(n, dtype=np.uint32)
    d_bed = hip.DeviceBuffer(c.bed.nbytes)
    d_bed.upload(c.bed)
    hip.device_synchronize()

    def one_fit(precision):
        hip.device_synchronize()
        t = time.perf_counter()
        r = sg.fit_null(None, c.bpm, n, m, c.y, c.x, sample_idx=idx, precision=precision,
                        d_bed=d_bed.value)
        hip.device_synchronize()
        return time.perf_counter() - t, r

    def describe(tag, dt, r):
        log(f"{tag}: {dt:.3f}s tau={r['tau'][1]:.8g} iters={r['iterations']} vr={r['vr']:.8g} "
            f"products={r['grm_products']} pcg_iters={r['pcg_iterations']} "
            f"(qc {r['seconds_qc']:.3f} reml {r['seconds_reml']:.3f} vr {r['seconds_vr']:.3f})")

    int8_forms = {"int8": sg.PREC_INT8, "int8x2": sg.PREC_INT8X2}
    head = args.precision
    other = "int8x2" if head == "int8" else "int8"
    nsl = 3 if head == "int8" else 2  # int8 operand slices of the headline form
    # untimed warm-up fits load the kernels and fill the device memory cache, for each
    # precision that is timed below
    for w in range(args.warmup):
        log(f"warmup {w}")
        one_fit(sg.PREC_INT8)
        one_fit(sg.PREC_INT8X2)
        one_fit(sg.PREC_MIXED)
        if not args.no_fp64:
            one_fit(sg.PREC_FP64)
    # the dominant kernel inside the timed region: every genotype GEMM launch of the timed
    # fits is timed with HIP events on the stream it runs on (the library's launch log); the
    # 34-column block solves (y, 3 covariates, 30 probes) run at 16 * ceil(nsl * 34 / 16)
    # operand columns
    L = sg.load()
    fit_width = 16 * -(-nsl * 34 // 16)
    times, rm = [], None
    for s in range(args.steps):
        dt, rm = one_fit(int8_forms[head])
        times.append(dt)
        describe(f"step {s} mixed {head}", dt, rm)
    mixed_s = sum(times) / len(times)
    # the launch log in an untimed repetition (its event pairs stay out of the timed fits)
    sg.check(L.saige_launch_log_enable(1, fit_width))
    one_fit(int8_forms[head])
    lt, ln = C.c_double(), C.c_uint64()
    sg.check(L.saige_launch_log_read(fit_width, C.byref(lt), C.byref(ln)))
    sg.check(L.saige_launch_log_enable(0, 0))
    fit_launch_ms, fit_launches = lt.value / max(ln.value, 1), int(ln.value)
    log(f"PCG iterations captured / replayed / eager so far: {sg.pcg_graph_counts()}")
    log(f"{head} GEMM launches at {fit_width} operand columns in the timed fits: "
        f"{fit_launches}, {fit_launch_ms:.4f} ms each")
    other_s, ro = one_fit(int8_forms[other])
    describe(f"mixed {other}", other_s, ro)
    bf16_s, rb = one_fit(sg.PREC_MIXED)
    describe("mixed bf16", bf16_s, rb)
    fp64_s, rf = None, None
    if not args.no_fp64:
        fp64_s, rf = one_fit(sg.PREC_FP64)
        describe("fp64", fp64_s, rf)
    ref = rf if rf is not None else rm

    # ---- equal accuracy: step-2 p-values under the null models ----
    acc = {}
    if rf is not None:
        ntest = min(4096, m)

        def pvals(r):
            e = sg.ScoreTestEngine(r["mu"], c.y - r["mu"], r["tau"], c.x, r["xvx_inv_xv"], r["vr"],
                                   y=c.y)
            p = e.test_markers_plink(c.bed[:ntest], c.bpm, n, ntest)["pvalue"]
            e.close()
            return p

        pf = pvals(rf)
        mask = pf >= 1e-10
        for tag, r in ((head, rm), (other, ro), ("bf16", rb)):
            pm = pvals(r)
            acc[tag] = {
                "tau_rel_diff": float(abs(r["tau"][1] - rf["tau"][1]) / rf["tau"][1]),
                "vr_rel_diff": float(abs(r["vr"] - rf["vr"]) / rf["vr"]),
                "max_abs_dlog10p": float(np.max(np.abs(np.log10(pf[mask]) - np.log10(pm[mask])))),
            }
        acc["variants_compared"] = int(mask.sum())
        log(f"accuracy vs fp64: {acc}")

    # ---- GRM product roofline at the block-PCG shape ----
    grm_markers = np.where(ref["marker_class"] == 1)[0].astype(np.uint32)
    ac, nv, af, mac = sg.marker_stats(c.bed, c.bpm, n, m)
    mg = len(grm_markers)
    k = 34
    flops = 4.0 * n * mg * k
    prod, kern = {}, {}
    for tag, prec in (("int8", sg.PREC_INT8), ("int8x2", sg.PREC_INT8X2), ("bf16", sg.PREC_MIXED),
                      ("fp64", sg.PREC_FP64)):
        ms, kms, kl = time_product(sg, hip, d_bed, c.bpm, n, grm_markers, af[grm_markers], prec,
                                   k, args.product_reps)
        prod[tag] = ms
        kern[tag] = kms
        log(f"GRM product {tag} k={k}: {ms:.3f} ms, {flops / (ms * 1e-3) / 1e12:.2f} TFLOP/s "
            f"(useful), {2.0 * ((n + 3) // 4) * mg / (ms * 1e-3) / 1e9:.1f} GB/s packed G; "
            f"GEMM kernel {kms:.3f} ms/launch over {kl} launches")
    # the dominant kernel: one genotype GEMM launch = one of the two passes of a product,
    # 2 * n * M * k useful flop (exact genotypes times the k-column operand), timed over the
    # launches of the timed fits
    launch_flops = 2.0 * n * mg * k
    if fit_launches == 0:  # no block solve at that width (another covariate count)
        fit_launch_ms = kern[head]
        log(f"no {fit_width}-column launches in the timed fits: the roofline uses the "
            f"product loop's {kern[head]:.4f} ms/launch")
    mixed_tf = launch_flops / (fit_launch_ms * 1e-3) / 1e12
    # algorithmic bytes of a launch: one pass over the packed genotypes, ceil(n/4) * M; at
    # 2 * k = 68 useful ops per genotype (272 per byte, under the int8 ridge of 625) the
    # launch is bound by HBM on the algorithmic roofline
    launch_bytes = float((n + 3) // 4) * mg
    mixed_gbs = launch_bytes / (fit_launch_ms * 1e-3) / 1e9

    # ---- step 2: streamed score test + SPA (n = 400,000, 4,096-variant uint8 batches) ----
    step2 = None
    if not args.no_step2:
        n2 = args.step2_n
        nf = synth.null_fit(n=n2, p=1, eta_mean=-4.6, eta_sd=0.5, seed=3)
        a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
        cutoff = 0.04550026389635842  # 2 * (1 - Phi(2)): SPA where |score/sd| > 2
        eng = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                                 spa_pval_cutoff=cutoff)
        B = 4096
        nb = args.step2_batches
        log(f"step2: generating {nb} batches of {B} variants at n={n2}")
        q = np.empty((nb * B, n2), np.uint8)
        synth.dosage_u8(n2, 0, nb * B, seed=4, out=q)
        d_q = hip.DeviceBuffer(q.nbytes)
        d_q.upload(q)
        q_cpu = q[:args.cpu_step2_variants].copy() if not args.no_cpu else None
        del q
        rsize = C.sizeof(sg._lib.VariantResult)
        d_r = hip.DeviceBuffer(nb * B * rsize)
        es = eng.stream
        eng.test_markers_u8_device(d_q.value, n2, B, 1.0 / 127, d_r.value)  # warm
        hip.device_synchronize()
        reps = args.step2_reps
        s0, s1 = hip.Event(), hip.Event()
        s0.record(es)
        for _ in range(reps):
            for b in range(nb):
                eng.test_markers_u8_device(d_q.value + b * B * n2, n2, B, 1.0 / 127,
                                           d_r.value + b * B * rsize)
        s1.record(es)
        ms = s0.elapsed_ms(s1) / reps
        out = (sg._lib.VariantResult * (nb * B))()
        host = np.empty(nb * B * rsize, np.uint8)
        d_r.download(host)
        C.memmove(C.addressof(out), host.ctypes.data, host.nbytes)
        r = sg.results_to_dict(out)
        vps = nb * B / (ms * 1e-3)
        step2 = dict(n=n2, variants=nb * B, ms=ms, repetitions=reps, variants_per_s=vps,
                     spa_fraction=float(r["is_spa"].mean()),
                     dosage_gbs=nb * B * n2 / (ms * 1e-3) / 1e9)
        log(f"step2: {nb * B} variants in {ms:.2f} ms (mean of {reps}) -> {vps:.3g} variants/s, "
            f"SPA'd {r['is_spa'].mean() * 100:.2f}%")
        if q_cpu is not None:
            # the oracle's score test + fast SPA (saige-rs single_variant.rs / fast.rs
            # restated) on one core, as saige-rs runs step 2 (assoc_test.rs:165), on the
            # first variants of the same batches; its p-values also check the GPU's
            import oracle
            eo = oracle.ScoreEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                                    spa_cutoff=cutoff)
            t = time.perf_counter()
            want = eo.test_batch_u8(q_cpu, 1.0 / 127)
            dt = time.perf_counter() - t
            nv = len(q_cpu)
            got_p = r["pvalue"][:nv]
            finite = want["pvalue"] >= 1e-300
            dlp = float(np.max(np.abs(np.log10(got_p[finite]) -
                                      np.log10(want["pvalue"][finite]))))
            step2["cpu_variants_per_s"] = nv / dt
            step2["max_abs_dlog10p_vs_oracle"] = dlp
            step2["is_spa_mismatches_vs_oracle"] = int(np.sum(r["is_spa"][:nv] != want["is_spa"]))
            log(f"step2 cpu: {nv} variants in {dt:.2f} s on 1 thread -> {nv / dt:.1f} variants/s; "
                f"GPU vs oracle max |d log10 p| = {dlp:.2e}")

    # ---- CPU baseline: the oracle's f32 OnTheFlyGrm product (saige-rs restated) on the
    # host cores, split over the thread pool the way rayon splits it, and on one core ----
    cpu = None
    if not args.no_cpu:
        import oracle
        mc = min(args.cpu_markers, mg)
        sub = grm_markers[:mc]
        dos = sg.decode(c.bed[sub], c.bpm, n, len(sub))
        g32 = oracle.Grm(dos, af[sub], oracle.GRM_F32)
        v = np.random.default_rng(1).standard_normal(n)

        def per_product(threads, seconds):
            t = time.perf_counter()
            reps = 0
            while time.perf_counter() - t < seconds or reps < 1:
                g32.mat_vec_threads(v, threads)
                reps += 1
            return (time.perf_counter() - t) / reps * (mg / mc), reps

        pt, reps_t = per_product(args.cpu_threads, 8.0)
        p1, reps_1 = per_product(1, 4.0)
        cpu_fit = pt * ref["grm_products"]
        cpu = {"value": cpu_fit, "unit": "s (fit-null, extrapolated)", "cores": args.cpu_threads,
               "kind": "port", "host": host_cpu(),
               "sample": f"f32 OnTheFlyGrm products (saige-rs pcg.rs:128-180 restated, rayon-"
                         f"shaped chunks) at n={n} x {mc} markers: {reps_t} products on "
                         f"{args.cpu_threads} threads, {reps_1} on 1; scaled by {mg}/{mc} markers "
                         f"and by the {ref['grm_products']} column-products of the GPU fit-null; "
                         f"one full product = {pt:.3f} s on {args.cpu_threads} threads, "
                         f"{p1:.2f} s on 1 thread (fit-null {p1 * ref['grm_products']:.0f} s)"}
        log(f"cpu baseline: {pt:.3f} s/product on {args.cpu_threads} threads, {p1:.2f} s on 1 -> "
            f"{cpu_fit:.0f} s fit-null (extrapolated)")
        if step2 is not None and "cpu_variants_per_s" in step2:
            cpu["step2"] = {"value": step2["cpu_variants_per_s"], "unit": "variants/s",
                            "cores": 1, "kind": "port",
                            "sample": f"oracle score test + fast SPA of the first "
                                      f"{args.cpu_step2_variants} variants of the step-2 "
                                      f"batches (n={step2['n']}), single-threaded as saige-rs "
                                      f"runs step 2"}

    line = {
        "metric": METRIC,
        "value": mixed_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed_s * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"int8 MFMA (exact genotypes x {nsl} int8 operand slices) / int32 accumulate / "
                 "fp64 solver",
        "data": "synthetic (seeded families cohort; packed genotypes resident in HBM)",
        "config": {"workload": f"fit-null binary, {args.n_fam} nuclear families x 5 = {n} samples, "
                               f"{m} markers ({mg} GRM after the VR split), 30 probes",
                   "precision": head, f"{other}_wall_s": other_s, "bf16_wall_s": bf16_s,
                   "fp64_wall_s": fp64_s,
                   "speedup_vs_fp64": (fp64_s / mixed_s) if fp64_s else None,
                   "tau_g": float(rm["tau"][1]), "ai_reml_iterations": int(rm["iterations"]),
                   "variance_ratio": float(rm["vr"]),
                   "grm_column_products": int(rm["grm_products"]),
                   "pcg_iterations": int(rm["pcg_iterations"]), "accuracy": acc,
                   "product_ms": prod, "gemm_launch_ms": kern,
                   "fit_gemm_launch_ms": fit_launch_ms, "fit_gemm_launches": fit_launches},
        "roofline": {"bound": "hbm", "achieved": mixed_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": mixed_gbs / HBM_PEAK_GBS,
                     "traffic": (args.traffic_bytes if args.traffic_bytes is not None else
                                 MEASURED_TRAFFIC.get((args.n_fam, args.markers, head))),
                     "kernel": f"i8_gemm_kernel (int8 MFMA genotype GEMM, {head}), k={k}: "
                               f"{fit_launch_ms:.4f} ms/launch over the {fit_launches} launches "
                               f"at {fit_width} operand columns in one untimed fit (HIP events "
                               f"on the stream of the fit's GRM handle; "
                               f"{kern[head]:.3f} ms/launch on dense random operands; with no "
                               f"such launch, the random-operand figure), "
                               f"algorithmic bytes ceil(n/4)*M = {launch_bytes:.4g} B/launch "
                               f"and 2*n*M*k = {launch_flops:.3g} op/launch "
                               f"({mixed_tf:.0f} TOP/s useful = {mixed_tf / I8_PEAK_TOPS:.3f} of "
                               f"{I8_PEAK_TOPS:.0f}); the {nsl} operand slices execute {nsl}x "
                               f"the useful ops (padded to 16-column tiles), "
                               f"{mixed_tf * 16 * -(-nsl * k // 16) / k:.0f} TOP/s executed; "
                               f"whole product {prod[head]:.3f} ms; bf16 GEMM "
                               f"{kern['bf16']:.3f} ms/launch; fp64 GEMM {kern['fp64']:.3f} "
                               f"ms/launch = {launch_flops / (kern['fp64'] * 1e-3) / 1e12:.2f} "
                               f"TFLOP/s of {FP64_PEAK_TFLOPS}"},
        "cpu_baseline": cpu,
        "step2": step2,
    }
    print(json.dumps(line), flush=True)


# HBM bytes per launch of the dominant genotype GEMM at the 4-trait 400,000 x 150,000 fit,
# measured with rocprofv3 FETCH_SIZE and WRITE_SIZE (one counter per run; FETCH_SIZE scaled
# by the factor a 1 GiB device copy gives, 2.0): keyed by (n, markers, form, operand width).
MEASURED_TRAFFIC_BIOBANK = {(400000, 150000, "int8x2", 224): 1.78e10,
                            (400000, 150000, "int8x2", 288): 1.87e10}


def step2_stream(args, sg, hip, synth):
    """Step 2 over all of the workload's variants: n = 400,000, intercept-only null with
    mu = logistic(N(-4.6, 0.5)), VR 1.05, SPA where |score/sd| > 2, 4,096-variant uint8
    batches generated on the device (the seeded generator's device twin), in chunks of
    resident batches.  Each chunk is scored by a pool of engines on their own streams (the
    timed figure, host wall clock between device synchronisations) and by one engine alone
    (timed beside it; its results must equal the pool's bit for bit).  Returns the step2 dict
    and the CPU baseline of step 2."""
    n2, V, B = args.step2_n, args.step2_variants, 4096
    nf = synth.null_fit(n=n2, p=1, eta_mean=-4.6, eta_sd=0.5, seed=3)
    a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    cutoff = 0.04550026389635842  # 2 * (1 - Phi(2)): SPA where |score/sd| > 2
    pool = sg.ScoreTestPool(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                            spa_pval_cutoff=cutoff, engines=args.step2_engines)
    one = pool.engines[0]
    mafs = synth.dosage_mafs(0, V, seed=4)
    d_m = hip.DeviceBuffer(mafs.nbytes)
    d_m.upload(mafs)
    nb = -(-V // B)
    R = max(1, min(args.step2_resident, nb))
    d_q = [hip.DeviceBuffer(B * n2) for _ in range(R)]
    rsize = C.sizeof(sg._lib.VariantResult)
    d_r = [hip.DeviceBuffer(V * rsize), hip.DeviceBuffer(V * rsize)]
    d_w = hip.DeviceBuffer(B * rsize)
    es = one.stream
    sizes = sorted({min(B, V - b * B) for b in range(nb)})
    for nv in sizes:  # first use of each batch size tunes the sums kernel's shape: untimed
        synth.dosage_u8_device(n2, 0, nv, d_m.value, d_q[0].value, seed=4, stream=es)
        pool.test_markers_u8_device([(d_q[0].value, n2, nv, 1.0 / 127, d_w.value)])
    hip.device_synchronize()
    log(f"step2: {V} variants at n={n2} in {nb} batches of {B}, generated on the device in "
        f"chunks of {R} resident batches; {len(pool.engines)} engines")
    g0, g1 = hip.Event(), hip.Event()
    gen_ms = score_s = one_s = 0.0
    for c0 in range(0, nb, R):
        bl = [(b, b * B, min(B, V - b * B)) for b in range(c0, min(nb, c0 + R))]
        g0.record(es)
        for b, first, nv in bl:
            synth.dosage_u8_device(n2, first, nv, d_m.value + 8 * first, d_q[b - c0].value,
                                   seed=4, stream=es)
        g1.record(es)
        hip.device_synchronize()
        gen_ms += g0.elapsed_ms(g1)
        t = time.perf_counter()
        pool.test_markers_u8_device([(d_q[b - c0].value, n2, nv, 1.0 / 127,
                                      d_r[0].value + first * rsize) for b, first, nv in bl])
        hip.device_synchronize()
        score_s += time.perf_counter() - t
        t = time.perf_counter()
        for b, first, nv in bl:
            one.test_markers_u8_device(d_q[b - c0].value, n2, nv, 1.0 / 127,
                                       d_r[1].value + first * rsize)
        hip.device_synchronize()
        one_s += time.perf_counter() - t
    score_ms = score_s * 1e3
    host = [np.empty(V * rsize, np.uint8), np.empty(V * rsize, np.uint8)]
    d_r[0].download(host[0])
    d_r[1].download(host[1])
    same = bool(np.array_equal(host[0], host[1]))
    out = (sg._lib.VariantResult * V).from_buffer(host[0])
    r = sg.results_to_dict(out)
    vps = V / score_s
    step2 = dict(n=n2, variants=V, batches=nb, batch=B, resident_batches=R,
                 engines=len(pool.engines), ms=score_ms, variants_per_s=vps,
                 one_engine_ms=one_s * 1e3, one_engine_variants_per_s=V / one_s,
                 pool_equals_one_engine_bitwise=same,
                 generation_ms=gen_ms, spa_fraction=float(r["is_spa"].mean()),
                 dosage_gbs=V * n2 / score_s / 1e9,
                 timing="host wall clock between device synchronisations around each chunk's "
                        "score-test calls (dosages resident; generation timed apart)")
    log(f"step2: {V} variants scored in {score_ms:.1f} ms -> {vps:.3g} variants/s "
        f"({step2['dosage_gbs']:.0f} GB/s of dosages) with {len(pool.engines)} engines; one "
        f"engine {one_s * 1e3:.1f} ms ({V / one_s:.3g}/s); results equal bitwise: {same}; "
        f"generation {gen_ms:.1f} ms; SPA'd {r['is_spa'].mean() * 100:.2f}%")
    if not same:
        raise SystemExit("step 2: the engine pool's results differ from one engine's")
    cpu2 = None
    if not args.no_cpu:
        # the oracle's score test + fast SPA (saige-rs single_variant.rs / fast.rs restated) on
        # one core, as saige-rs runs step 2 (assoc_test.rs:165), on the first variants; its
        # p-values also check the GPU's
        import oracle
        nv = args.cpu_step2_variants
        q_cpu, _ = synth.dosage_u8(n2, 0, nv, seed=4)
        eo = oracle.ScoreEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.05, y=nf.y,
                                spa_cutoff=cutoff)
        t = time.perf_counter()
        want = eo.test_batch_u8(q_cpu, 1.0 / 127)
        dt = time.perf_counter() - t
        got_p = r["pvalue"][:nv]
        finite = want["pvalue"] >= 1e-300
        dlp = float(np.max(np.abs(np.log10(got_p[finite]) - np.log10(want["pvalue"][finite]))))
        step2["max_abs_dlog10p_vs_oracle"] = dlp
        step2["is_spa_mismatches_vs_oracle"] = int(np.sum(r["is_spa"][:nv] != want["is_spa"]))
        step2["cpu_variants_per_s"] = nv / dt
        cpu2 = {"value": nv / dt, "unit": "variants/s", "cores": 1, "kind": "port",
                "sample": f"oracle score test + fast SPA of the first {nv} of the {V} variants "
                          f"(n={n2}), single-threaded as saige-rs runs step 2 "
                          f"(assoc_test.rs:165); {dt:.1f} s"}
        log(f"step2 cpu: {nv} variants in {dt:.2f} s on 1 thread -> {nv / dt:.1f} variants/s; "
            f"GPU vs oracle max |d log10 p| = {dlp:.2e}")
    pool.close()
    return step2, cpu2


def run_headline(args):
    """The default line.  Fit-null of 4 binary traits (prevalence 0.5/1/5/20%) over 400,000
    samples (15% in half-sibling pairs) x 150,000 markers resident in HBM, the traits fitted
    together with joined GRM products: the int8 MFMA form with two operand slices (three for
    the narrow solves and near convergence) is `value`; the three-slice form and the plain
    FP64 path are timed beside it on the same cohort, and each form's tau, VR and step-2
    p-values are compared with FP64's.  Then step 2 over 1,000,000 variants, the roofline of
    the dominant GEMM launch and the CPU baselines."""
    import csaige_amd as sg
    from csaige_amd import hip, synth
    t0 = time.time()
    prev = (0.005, 0.01, 0.05, 0.20)
    log(f"generating biobank cohort n={args.bb_n} m={args.bb_m}")
    c = synth.biobank(n=args.bb_n, m=args.bb_m, frac_related=0.15, prevalences=prev, seed=5)
    n, m = c.n, c.bed.shape[0]
    log(f"cohort generated in {time.time() - t0:.1f}s: {c.bed.nbytes / 1e9:.2f} GB packed, "
        f"cases {c.y.sum(axis=1).astype(int).tolist()}")
    idx = np.arange(n, dtype=np.uint32)
    d_bed = hip.DeviceBuffer(c.bed.nbytes)
    d_bed.upload(c.bed)
    hip.device_synchronize()

    def fit(precision):
        hip.device_synchronize()
        t = time.perf_counter()
        r = sg.fit_null_multi(None, c.bpm, n, m, c.y, c.x, sample_idx=idx, precision=precision,
                              d_bed=d_bed.value)
        hip.device_synchronize()
        return time.perf_counter() - t, r

    def describe(tag, dt, rr):
        log(f"{tag}: {dt:.3f}s " + "; ".join(
            f"t{t} tau={r['tau'][1]:.7g} it={r['iterations']} vr={r['vr']:.6g} "
            f"pcg={r['pcg_iterations']}" for t, r in enumerate(rr)))

    forms = {"int8": sg.PREC_INT8, "int8x2": sg.PREC_INT8X2}
    head = args.precision
    other = "int8x2" if head == "int8" else "int8"
    for w in range(args.warmup):
        dt, rr = fit(forms[head])
        describe(f"warmup {w} {head}", dt, rr)
    times = []
    rm = None
    for s in range(args.steps):
        dt, rm = fit(forms[head])
        times.append(dt)
        describe(f"step {s} {head}", dt, rm)
    head_s = float(np.mean(times))
    # the dominant kernel, in an untimed repetition: every genotype GEMM launch of one fit is
    # timed with HIP events on the stream it runs on (the library's launch log)
    sg.launch_log_enable(True, 0)
    dt_log, _ = fit(forms[head])
    logsum = sg.launch_log_summary()
    sg.launch_log_enable(False)
    log(f"launch log ({dt_log:.3f}s fit): " + ", ".join(
        f"w{w}: {v[1]} x {v[0] / max(v[1], 1):.3f} ms" for w, v in logsum.items()))
    other_s, ro = fit(forms[other])
    describe(f"{other}", other_s, ro)
    log(f"PCG iterations captured / replayed / eager so far: {sg.pcg_graph_counts()}")
    fp64_s, rf = None, None
    if not args.no_fp64:
        fp64_s, rf = fit(sg.PREC_FP64)
        describe("fp64", fp64_s, rf)

    # ---- accuracy against FP64: tau, VR, step-2 -log10 p of 2,048 variants per trait ----
    acc, valid = None, {head: True, other: True}
    if rf is not None:
        acc = {head: [], other: []}
        ntest = 2048

        def pvals(r, t):
            e = sg.ScoreTestEngine(r["mu"], c.y[t] - r["mu"], r["tau"], c.x, r["xvx_inv_xv"],
                                   r["vr"], y=c.y[t])
            p = e.test_markers_plink(c.bed[:ntest], c.bpm, n, ntest)["pvalue"]
            e.close()
            return p

        for t in range(len(prev)):
            pf = pvals(rf[t], t)
            mask = pf >= 1e-10
            for tag, rr in ((head, rm), (other, ro)):
                pm = pvals(rr[t], t)
                a = {"tau_rel_diff": float(abs(rr[t]["tau"][1] - rf[t]["tau"][1]) / rf[t]["tau"][1]),
                     "vr_rel_diff": float(abs(rr[t]["vr"] - rf[t]["vr"]) / rf[t]["vr"]),
                     "max_abs_dlog10p": float(np.max(np.abs(np.log10(pf[mask]) -
                                                            np.log10(pm[mask])))),
                     "iterations": [int(rr[t]["iterations"]), int(rf[t]["iterations"])],
                     "converged": bool(rr[t]["converged"])}
                acc[tag].append(a)
                valid[tag] = valid[tag] and (a["converged"] and a["tau_rel_diff"] <= 1e-4 and
                                             a["max_abs_dlog10p"] <= 1e-3)
        log(f"accuracy vs fp64 (targets tau 1e-4, -log10 p 1e-3): {acc}")
    # the headline form must meet the mixed-precision targets; otherwise the three-slice form
    # (or, failing that, nothing) carries the line
    value, value_form, note = head_s, head, None
    if not valid[head]:
        if valid[other]:
            value, value_form = other_s, other
            note = f"{head} missed the accuracy targets in this run; {other} carries the value"
        else:
            note = "no int8 form met the accuracy targets in this run: value is not valid"
        log("WARNING: " + note)

    # ---- roofline of the dominant kernel: the genotype GEMM width with the most time ----
    dom_w = max(logsum, key=lambda w: logsum[w][0]) if logsum else None
    roof, kern = None, None
    if dom_w is not None:
        tot_ms, cnt, vecs = logsum[dom_w]
        launch_ms = tot_ms / cnt
        mg = int(rm[0]["n_grm_markers"])
        launch_bytes = float((n + 3) // 4) * mg          # one pass over the packed genotypes
        useful_ops = 2.0 * n * mg * vecs / cnt           # 2 n M k over the launches' real k
        exec_ops = 2.0 * n * mg * dom_w                  # the padded slice columns it executes
        gbs = launch_bytes / (launch_ms * 1e-3) / 1e9
        all_ms = sum(v[0] for v in logsum.values())
        kern = {"width": dom_w, "launches": cnt, "ms_per_launch": launch_ms,
                "mean_vectors": vecs / cnt, "share_of_gemm_time": tot_ms / all_ms,
                "useful_tops": useful_ops / (launch_ms * 1e-3) / 1e12,
                "executed_tops": exec_ops / (launch_ms * 1e-3) / 1e12,
                "per_width": {str(w): {"launches": v[1], "ms_per_launch": v[0] / max(v[1], 1),
                                       "mean_vectors": v[2] / max(v[1], 1)}
                              for w, v in logsum.items()}}
        # which roof bounds the launch on its algorithmic work: 2 n M k useful ops over
        # ceil(n/4) M genotype bytes is 8 k ops per byte against the int8 ridge of
        # 5.0 POP/s / 8 TB/s = 625
        useful_tops = useful_ops / (launch_ms * 1e-3) / 1e12
        intensity = useful_ops / launch_bytes
        mfma_bound = intensity >= I8_PEAK_TOPS / HBM_PEAK_GBS * 1e3
        traffic = MEASURED_TRAFFIC_BIOBANK.get((n, m, value_form, dom_w))
        text = (f"i8_gemm_kernel ({head}) at {dom_w} operand columns (mean {vecs / cnt:.1f} "
                f"vectors), the GEMM width with the most time in one {head} fit "
                f"({tot_ms / all_ms * 100:.0f}% of its GEMM time): {launch_ms:.3f} ms/launch over "
                f"{cnt} launches (HIP events on the fit's GRM stream, untimed repetition); "
                f"algorithmic work per launch 2*n*M*k = {useful_ops:.4g} int8 op over "
                f"ceil(n/4)*M = {launch_bytes:.4g} B ({intensity:.0f} op/B, int8 ridge 625): "
                f"{useful_tops:.0f} TOP/s useful, {exec_ops / (launch_ms * 1e-3) / 1e12:.0f} TOP/s "
                f"executed (operand slices and padding) of {I8_PEAK_TOPS:.0f}, {gbs:.0f} GB/s "
                f"of packed genotypes ({gbs / HBM_PEAK_GBS:.3f} of HBM)")
        if mfma_bound:
            roof = {"bound": "mfma", "achieved": useful_tops, "peak": I8_PEAK_TOPS,
                    "unit": "TFLOP/s", "frac": useful_tops / I8_PEAK_TOPS, "traffic": traffic,
                    "kernel": text + "; int8 ops counted as TFLOP/s"}
        else:
            roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "kernel": text}
        log(f"dominant GEMM: width {dom_w}, {launch_ms:.3f} ms/launch, {useful_tops:.0f} TOP/s "
            f"useful ({useful_tops / I8_PEAK_TOPS:.3f} of int8 peak), {gbs:.0f} GB/s "
            f"({gbs / HBM_PEAK_GBS:.3f} of HBM)")

    # ---- FP64 roofline floor: the FP64 fit's column-products at 78.6 TFLOP/s ----
    # grm_products counts each trait's columns; the joined fit computes the probes' GRM u
    # (30 columns, the same probes for every trait) once, so the other traits' copies are
    # not products
    fp64_floor = None
    if rf is not None:
        mg = int(rf[0]["n_grm_markers"])
        cols = sum(int(r["grm_products"]) for r in rf) - (len(rf) - 1) * 30
        fp64_floor = 4.0 * n * mg * cols / (FP64_PEAK_TFLOPS * 1e12)

    # ---- CPU baseline of the fit: sampled f32 OnTheFlyGrm products, extrapolated ----
    cpu = None
    if not args.no_cpu:
        import oracle
        ref = rf if rf is not None else rm
        mg = int(ref[0]["n_grm_markers"])
        grm_markers = np.flatnonzero(ref[0]["marker_class"] == 1)[:args.cpu_fit_markers]
        mc = len(grm_markers)
        ac, nv, af, mac = sg.marker_stats(c.bed[grm_markers], c.bpm, n, mc)
        dos = sg.decode(c.bed[grm_markers], c.bpm, n, mc)
        g32 = oracle.Grm(dos, af, oracle.GRM_F32)
        del dos
        v = np.random.default_rng(1).standard_normal(n)

        def per_product(threads, seconds):
            t = time.perf_counter()
            reps = 0
            while time.perf_counter() - t < seconds or reps < 1:
                g32.mat_vec_threads(v, threads)
                reps += 1
            return (time.perf_counter() - t) / reps * (mg / mc), reps

        pt, reps_t = per_product(args.cpu_threads, 8.0)
        p1, reps_1 = per_product(1, 4.0)
        cols = sum(int(r["grm_products"]) for r in ref)
        cpu = {"value": pt * cols, "unit": "s (fit-null of the 4 traits, extrapolated)",
               "cores": args.cpu_threads, "kind": "port", "host": host_cpu(),
               "sample": f"f32 OnTheFlyGrm products (saige-rs pcg.rs:128-180 restated, "
                         f"rayon-shaped chunks) at n={n} x {mc} markers: {reps_t} products on "
                         f"{args.cpu_threads} threads, {reps_1} on 1; scaled by {mg}/{mc} "
                         f"markers and by the {cols} column-products the GPU fits of the 4 "
                         f"traits performed (a lower bound: saige-rs also recomputes GRM u "
                         f"every iteration and needs ~720 GB for its f32 copy of Z at this "
                         f"size); one full product {pt:.2f} s on {args.cpu_threads} threads, "
                         f"{p1:.1f} s on 1"}
        log(f"cpu baseline: {pt:.3f} s/product on {args.cpu_threads} threads -> "
            f"{pt * cols:.0f} s fit-null (extrapolated)")

    # ---- step 2 over all variants ----
    step2, cpu2 = (None, None)
    if not args.no_step2:
        d_bed.free()
        step2, cpu2 = step2_stream(args, sg, hip, synth)
        if cpu is not None and cpu2 is not None:
            cpu["step2"] = cpu2

    # ---- the one-trait families workload, as extra keys ----
    fam = None
    if not args.no_families:
        fam = families_extra(args, sg, hip, synth, forms[head])

    line = {
        "metric": METRIC,
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"int8 MFMA (exact genotypes x {'2-3' if value_form == 'int8x2' else '3'} int8 "
                 "operand slices) / int32 accumulate / fp64 solver",
        "data": "synthetic (seeded biobank cohort, packed genotypes resident in HBM; step-2 "
                "dosages generated on the device)",
        "config": {"workload": f"fit-null of {len(prev)} binary traits (prevalence "
                               f"{'/'.join(str(p) for p in prev)}) fitted together, {n} samples "
                               f"(15% in half-sib pairs), {m} markers, 30 probes per trait, "
                               f"joined GRM products; step 2 over {args.step2_variants} "
                               f"variants at n={args.step2_n}",
                   "precision": value_form, "note": note,
                   f"{head}_wall_s": head_s, f"{other}_wall_s": other_s,
                   "fp64_wall_s": fp64_s,
                   "speedup_vs_fp64": (fp64_s / value) if fp64_s else None,
                   "fp64_roofline_floor_s": fp64_floor,
                   "speedup_vs_fp64_roofline_floor": (fp64_floor / value) if fp64_floor else None,
                   "accuracy_targets_met": valid if rf is not None else None,
                   "tau_g": [float(r["tau"][1]) for r in rm],
                   "ai_reml_iterations": [int(r["iterations"]) for r in rm],
                   "pcg_iterations": [int(r["pcg_iterations"]) for r in rm],
                   "grm_column_products": [int(r["grm_products"]) for r in rm],
                   "variance_ratio": [float(r["vr"]) for r in rm],
                   "accuracy_vs_fp64": acc, "step_times_s": times,
                   "dominant_gemm": kern, "families": fam},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step2": step2,
    }
    print(json.dumps(line), flush=True)


def families_extra(args, sg, hip, synth, prec):
    """The one-trait families workload (50,000 samples x 100,000 markers): the int8 form and
    FP64, for comparison with earlier lines."""
    c = synth.families(n_fam=args.n_fam, kids=3, m=args.markers, seed=2, h2=0.4, prevalence=0.05)
    n, m = c.n, c.bed.shape[0]
    idx = np.arange(n, dtype=np.uint32)
    d_bed = hip.DeviceBuffer(c.bed.nbytes)
    d_bed.upload(c.bed)

    def one(p):
        hip.device_synchronize()
        t = time.perf_counter()
        r = sg.fit_null(None, c.bpm, n, m, c.y, c.x, sample_idx=idx, precision=p,
                        d_bed=d_bed.value)
        hip.device_synchronize()
        return time.perf_counter() - t, r

    one(prec)
    ts = []
    for _ in range(3):
        dt, r = one(prec)
        ts.append(dt)
    out = {"workload": f"fit-null binary, {args.n_fam} nuclear families x 5 = {n} samples, "
                       f"{m} markers", "wall_s": float(np.mean(ts)),
           "tau_g": float(r["tau"][1]), "ai_reml_iterations": int(r["iterations"])}
    if not args.no_fp64:
        one(sg.PREC_FP64)
        dtf, rf = one(sg.PREC_FP64)
        out.update(fp64_wall_s=dtf, speedup_vs_fp64=dtf / out["wall_s"],
                   tau_rel_diff=float(abs(r["tau"][1] - rf["tau"][1]) / rf["tau"][1]),
                   fp64_ai_reml_iterations=int(rf["iterations"]))
    log(f"families: {out}")
    d_bed.free()
    return out


if __name__ == "__main__":
    main()
