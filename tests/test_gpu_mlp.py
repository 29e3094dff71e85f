"""GPU kernels of the learned saddle-point root proposal: gradients and forward pass
against the numpy checker, and what the trained proposal gains over starting at 0."""
import numpy as np
import pytest

import csaige_amd as sg
from csaige_amd import synth
from oracle import mlp as om
from tests.helpers import neglog10

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, np.float32).copy()
    u = a.view(np.uint32)
    u += 0x7FFF + ((u >> 16) & 1)   # round to nearest even
    u &= 0xFFFF0000
    return a


def test_loss_grad_vs_numpy():
    rng = np.random.default_rng(1)
    m = sg.SpaRootMlp(seed=3)
    p = (rng.standard_normal(om.NP) * 0.2).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(6), rng.uniform(0.5, 2.0, 6), [1.7]]).astype(np.float32)
    m.set_params(p, norm)
    x = rng.standard_normal((1000, 6)).astype(np.float32)
    y = rng.standard_normal(1000).astype(np.float32)
    loss, g = m.loss_grad(x, y)
    xn = om.standardize(x.astype(np.float64), norm)
    yn = om.target(y, norm[12])
    loss_ref, g_ref = om.loss_grad(p.astype(np.float64), xn, yn)
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    np.testing.assert_allclose(g, g_ref, rtol=2e-3, atol=2e-5 * np.abs(g_ref).max())


def test_predict_bf16_weights_vs_numpy():
    rng = np.random.default_rng(2)
    m = sg.SpaRootMlp(seed=4)
    p = (rng.standard_normal(om.NP) * 0.2).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(6), rng.uniform(0.5, 2.0, 6), [0.8]]).astype(np.float32)
    m.set_params(p, norm)
    x = rng.standard_normal((500, 6)).astype(np.float32)
    got = m.predict(x)
    want, _ = om.forward(_bf16(p).astype(np.float64), om.standardize(x, norm))
    np.testing.assert_allclose(got, np.sinh(want * norm[12]), rtol=2e-4, atol=2e-4)


def _spa_cohort(n, nvar, prevalence, seed):
    eta = np.log(prevalence / (1 - prevalence))
    nf = synth.null_fit(n=n, p=1, eta_mean=eta, eta_sd=0.5, seed=seed)
    a = sg.xvx_inv_xv(nf.x, nf.mu * (1 - nf.mu))
    q, _ = synth.dosage_u8(n, 0, nvar, seed=seed + 1, maf=(5e-4, 0.5))
    return nf, a, q


def test_trained_proposal_cuts_newton_iterations():
    """Train on roots from 1% and 10% prevalence runs, then compare on held-out variants:
    fewer Newton iterations and the plain method's outcome on every gated variant.  The
    proposal starts only searches whose root lies within 20 of 0; there the search from 0
    reaches the same unique root, and everywhere else the search runs from 0
    (find_root_k1_fast, fast.rs:53-132), so a root the plain search misses (divergence, then
    the unadjusted p of single_variant.rs) is missed the same way."""
    xs, ys = [], []
    for prev, seed in ((0.01, 11), (0.10, 21)):
        nf, a, q = _spa_cohort(20000, 6000, prev, seed)
        e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.0, y=nf.y,
                               spa_pval_cutoff=1.0, spa_tol=1e-10)
        r = e.test_markers_u8(q)
        x, y = sg.spa_training_set(r)
        xs.append(x)
        ys.append(y)
    x, y = np.concatenate(xs), np.concatenate(ys)
    m = sg.SpaRootMlp(seed=5)
    hist = m.train(x, y, epochs=40, batch=1024, lr=2e-3, seed=7)
    print(f"training roots {len(y)}; loss history {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.5 * hist[0]
    nf, a, q = _spa_cohort(20000, 3000, 0.03, 31)
    e = sg.ScoreTestEngine(nf.mu, nf.y - nf.mu, (1.0, 0.0), nf.x, a, 1.0, y=nf.y,
                           spa_pval_cutoff=0.05, spa_tol=1e-10)
    r0 = e.test_markers_u8(q)
    e.set_root_proposal(m)
    r1 = e.test_markers_u8(q)
    gated = r0["pvalue_noadj"] < 0.05
    plain_ok = gated & (r0["spa_converged"] == 1)
    lost = gated & (r0["spa_converged"] == 0)
    assert plain_ok.sum() > 50 and lost.sum() > 0
    it0 = (r0["spa_iters1"][plain_ok] + r0["spa_iters2"][plain_ok]).mean()
    it1 = (r1["spa_iters1"][plain_ok] + r1["spa_iters2"][plain_ok]).mean()
    d = np.abs(neglog10(r0["pvalue"][plain_ok]) - neglog10(r1["pvalue"][plain_ok])).max()
    print(f"SPA'd {plain_ok.sum()}: mean Newton iterations {it0:.2f} -> {it1:.2f}; "
          f"max |d -log10 p| {d:.2e}; plain search diverged on {lost.sum()}", flush=True)
    assert it1 < 0.85 * it0
    # the plain method's outcome on every gated variant
    np.testing.assert_array_equal(r1["spa_converged"][gated], r0["spa_converged"][gated])
    np.testing.assert_array_equal(r1["is_spa"][gated], r0["is_spa"][gated])
    assert d < 1e-6
    np.testing.assert_array_equal(r1["pvalue"][lost], r0["pvalue"][lost])
    np.testing.assert_array_equal(r1["pvalue"][lost], r1["pvalue_noadj"][lost])
