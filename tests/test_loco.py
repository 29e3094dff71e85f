"""Leave-one-chromosome-out (saige-core/src/glmm/loco.rs:54-114).

CPU: the oracle's run_loco on the reference's own test case (loco.rs:120-171) and the
invariant that each per-chromosome GRM is the GRM of the kept markers.  GPU: the masked
product of each precision against the oracle GRM of the kept markers, and the joined
per-chromosome AI-REML refits against the oracle's one-at-a-time refits."""
import numpy as np
import pytest

import csaige_amd as sg
import oracle
from csaige_amd import synth
from tests.helpers import decode_np, rel


def _ref_case():
    # loco.rs:124-143: 20 samples, intercept only, 4 markers, 2 per chromosome
    n = 20
    y = np.array([1.0 + i * 0.1 for i in range(n)])
    x = np.ones((1, n))
    chroms = ["1", "1", "2", "2"]
    d = np.array([[float((i + m) % 3) for i in range(n)] for m in range(4)])
    afs = np.array([0.3, 0.4, 0.35, 0.45])
    return y, x, chroms, d, afs


def test_oracle_run_loco_reference_case():
    y, x, chroms, d, afs = _ref_case()
    res = oracle.run_loco(y, x, chroms, d, afs, trait=oracle.TRAIT_QUANTITATIVE, max_iter=5,
                          tol=1e-3, pcg_tol=1e-3, pcg_max_iter=50, n_random_vectors=3, seed=42)
    # the reference test's assertions (loco.rs:156-170)
    assert sorted(res) == ["1", "2"]
    for c, r in res.items():
        assert len(r["mu"]) == 20 and len(r["residuals"]) == 20
        assert r["tau"][0] > 0.0
    # each refit is fit_ai_reml on the GRM of the other chromosome's markers
    keep = np.array([c != "1" for c in chroms])
    direct = oracle.fit_ai_reml(oracle.Grm(d[keep], afs[keep]), y, x,
                                trait=oracle.TRAIT_QUANTITATIVE, max_iter=5, tol=1e-3,
                                pcg_tol=1e-3, pcg_max_iter=50, n_random_vectors=3, seed=42)
    np.testing.assert_array_equal(res["1"]["tau"], direct["tau"])


def _cohort(n=600, m=1800, seed=7):
    c = synth.unrelated(n=n, m_grm=m, m_test=0, seed=seed, h2=0.5, prevalence=0.2)
    d = decode_np(c.bed, c.n)
    af = np.array([oracle.compute_af(r)[1] for r in d])
    # three chromosomes of unequal size, not contiguous in marker order for one of them
    chrom = np.zeros(m, np.int32)
    chrom[m // 2:] = 1
    chrom[::7] = 2
    return c, d, af, chrom


@pytest.mark.gpu
@pytest.mark.parametrize("prec,bound", [(sg.PREC_FP64, 1e-12), (sg.PREC_INT8, 1e-6),
                                        (sg.PREC_INT8X2, 1e-4), (sg.PREC_MIXED, 2e-5)],
                         ids=["fp64", "int8", "int8x2", "bf16"])
def test_gpu_loco_products_vs_oracle(prec, bound):
    c, d, af, chrom = _cohort()
    g = sg.OnTheFlyGrm(c.bed, c.bpm, c.n, af, precision=prec)
    g.set_marker_chromosomes(chrom)
    rng = np.random.default_rng(1)
    V = rng.standard_normal((5, c.n))
    for cc in range(3):
        keep = chrom != cc
        ref = oracle.Grm(d[keep], af[keep], oracle.GRM_F64)
        got = g.mat_mat_loco(V, cc)
        for j in range(5):
            want = ref.mat_vec(V[j])
            err = np.linalg.norm(got[j] - want) / np.linalg.norm(want)
            assert err < bound, (cc, j, err)
    # the product without LOCO is untouched afterwards
    full = oracle.Grm(d, af, oracle.GRM_F64).mat_vec(V[0])
    got = g.mat_vec(V[0])
    assert np.linalg.norm(got - full) / np.linalg.norm(full) < bound


@pytest.mark.gpu
@pytest.mark.parametrize("trait", [sg.TRAIT_BINARY, sg.TRAIT_QUANTITATIVE])
def test_gpu_loco_refits_vs_oracle(trait):
    """The three refits run together (joined products, each column masked for its
    chromosome) and each follows the oracle's single refit: iterations, tau to 1e-6."""
    c, d, af, chrom = _cohort()
    y = c.y if trait == sg.TRAIT_BINARY else c.liability
    g = sg.OnTheFlyGrm(c.bed, c.bpm, c.n, af)
    g.set_marker_chromosomes(chrom)
    got = g.fit_ai_reml_loco(y, c.x, trait=trait, tol=1e-5)
    want = oracle.run_loco(y, c.x, [int(v) for v in chrom], d, af, trait=trait, tol=1e-5)
    assert len(got) == 3 and sorted(want) == [0, 1, 2]
    for cc in range(3):
        w = want[cc]
        print(f"chrom {cc}: tau gpu={got[cc]['tau']} oracle={w['tau']} "
              f"it {got[cc]['iterations']}/{w['iterations']}", flush=True)
        assert got[cc]["iterations"] == w["iterations"]
        assert rel(got[cc]["tau"], w["tau"]).max() < 1e-6
        np.testing.assert_allclose(got[cc]["mu"], w["mu"], rtol=1e-6)
        np.testing.assert_allclose(got[cc]["alpha"], w["alpha"], rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(got[cc]["residuals"], y - got[cc]["mu"], rtol=0, atol=0)
    # a refit is not the full-GRM fit
    full = sg.fit_ai_reml(g, y, c.x, trait=trait, tol=1e-5)
    assert rel(got[0]["tau"], full["tau"]).max() > 1e-6


@pytest.mark.gpu
def test_gpu_loco_rejects_chromosome_codes_without_markers():
    """Codes must be dense (0..C-1, each used): run_loco fits only the chromosomes present
    (loco.rs:62-95), so a gap in the codes (chromosomes coded 1..22, say) is refused instead of
    refitting a chromosome that does not exist."""
    c, d, af, chrom = _cohort(n=300, m=400)
    g = sg.OnTheFlyGrm(c.bed, c.bpm, c.n, af)
    with pytest.raises(sg.SaigeError):
        g.set_marker_chromosomes(chrom + 1)  # code 0 unused
    gap = chrom.copy()
    gap[gap == 1] = 3  # codes 0, 2, 3: code 1 unused
    with pytest.raises(sg.SaigeError):
        g.set_marker_chromosomes(gap)
    g.set_marker_chromosomes(chrom)  # dense codes are accepted
