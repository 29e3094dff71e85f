"""Writes the oracle's results on scaled-down versions of the two fit-null workloads, for the
GPU tests to compare with (the oracle needs minutes at these sizes; the GPU box runs the
tests in seconds).  The cohorts come from the seeded generator, which is bit-deterministic
(fixed-point polygenic scores), and the fixture keeps a hash of each so the test checks it
regenerated the same data.

  config_families.npz  400 nuclear families (2 founders + 3 offspring, n = 2,000), 4,000
                       markers, binary trait at 5% prevalence: oracle fit_null (FP64 GRM),
                       and the step-2 p-values of 256 uint8 test variants under its null.
  config_biobank.npz   8,000 samples (15% in half-sibling pairs), 4,000 markers, 4 binary
                       traits at 0.5/1/5/20% prevalence (40 cases at 0.5%): the oracle's
                       fit_null of each trait and the p-values of the same 256 test variants.

Run: python tests/golden/make_config_vectors.py  (CPU; about two minutes on 8 cores)
"""
import hashlib
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

OUT = os.path.dirname(os.path.abspath(__file__))
PREV = (0.005, 0.01, 0.05, 0.20)
NTEST = 256


def families():
    from csaige_amd import synth
    return synth.families(n_fam=400, kids=3, m=4000, seed=2, h2=0.4, prevalence=0.05)


def biobank():
    from csaige_amd import synth
    return synth.biobank(n=8000, m=4000, frac_related=0.15, prevalences=PREV, seed=5, h2=0.4,
                         n_causal=500)


def test_dosages(n):
    from csaige_amd import synth
    q, _ = synth.dosage_u8(n, 0, NTEST, seed=7, maf=(2e-3, 0.5))
    return q


def cohort_hash(c):
    h = hashlib.sha256()
    for a in (c.bed, c.y, c.x):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def fit_and_test(args):
    import oracle
    which, t = args
    c = families() if which == "families" else biobank()
    y = c.y if which == "families" else c.y[t]
    idx = np.arange(c.n, dtype=np.uint32)
    f = oracle.fit_null(c.bed, c.bpm, c.n, c.bed.shape[0], idx, y, c.x)
    eo = oracle.ScoreEngine(f["mu"], y - f["mu"], f["tau"], c.x, f["xvx_inv_xv"], f["vr"], y=y)
    r = eo.test_batch_u8(test_dosages(c.n), 1.0 / 127)
    return dict(tau=f["tau"], iterations=f["iterations"], vr=f["vr"],
                marker_class=f["marker_class"], pvalue=r["pvalue"], is_spa=r["is_spa"])


def main():
    jobs = [("families", 0)] + [("biobank", t) for t in range(len(PREV))]
    with ProcessPoolExecutor(max_workers=len(jobs)) as ex:
        res = list(ex.map(fit_and_test, jobs))
    fam = res[0]
    np.savez_compressed(os.path.join(OUT, "config_families.npz"), cohort_sha256=cohort_hash(families()),
                        **{k: np.asarray(v) for k, v in fam.items()})
    bb = res[1:]
    out = {"cohort_sha256": cohort_hash(biobank())}
    for k in bb[0]:
        out[k] = np.stack([np.asarray(r[k]) for r in bb])
    np.savez_compressed(os.path.join(OUT, "config_biobank.npz"), **out)
    for name, r in [("families", fam)] + [(f"biobank t{t}", r) for t, r in enumerate(bb)]:
        print(f"{name}: tau {r['tau']} it {r['iterations']} vr {r['vr']:.6g} "
              f"SPA'd {int(np.sum(r['is_spa']))}")


if __name__ == "__main__":
    main()
