"""Region tests (BURDEN / SKAT / SKAT-O, saige-core/src/score_test/region.rs).

CPU: the reference's own region tests (region.rs:431-523) on the oracle and the host
helpers, the m = 1 closed form, and the library's chi-square tail against scipy.  GPU: the
device Gram + host p-values against the oracle's explicit g~ restatement."""
import math

import numpy as np
import pytest
from scipy.special import gammainc

import csaige_amd as sg
import oracle
from csaige_amd import region as R
from oracle import region as OR


def test_reference_region_kats():
    # region.rs:435-443
    w = R.beta_weights([0.001, 0.01, 0.05, 0.1, 0.3], 1.0, 25.0)
    assert w[0] > w[1] > w[2] > w[3]
    # region.rs:446-451
    rhos = R.default_rho_values()
    assert len(rhos) == 11 and rhos[0] == 0.0 and rhos[10] == 1.0
    # region.rs:454-467
    c = R.mac_category_counts([0.5, 1.0, 2.0, 5.0, 10.0, 15.0, 25.0, 50.0],
                              [1.5, 2.5, 5.5, 10.5, 20.5, 40.5])
    assert c == [2, 1, 1, 1, 1, 1, 1, 0]


def _ref_basic():
    # region.rs:470-506
    n = 100
    mu2 = np.full(n, 0.25)
    res = np.array([0.5 if i % 2 == 0 else -0.5 for i in range(n)])
    x = np.ones((1, n))
    a = np.full(n, 1.0 / n)
    G = np.array([[1.0 if (i + j) % 10 == 0 else 0.0 for i in range(n)] for j in range(5)])
    return G, np.ones(5), mu2, res, x, a


def test_reference_region_test_basic_on_oracle():
    G, w, mu2, res, x, a = _ref_basic()
    p_o, p_b, p_s, rho = OR.run_region_test(G, w, mu2, res, x, a, 1.0, R.default_rho_values())
    for p in (p_o, p_b, p_s):
        assert 0.0 <= p <= 1.0
    assert p_o <= max(p_b, p_s) + 1e-10  # region.rs:519-521


def test_oracle_single_variant_closed_form():
    """For one variant every test is the chi-square(1) test of S^2 / Phi, which with
    tau_e = 1 and weight 1 is the single-variant score test without variance ratio."""
    rng = np.random.default_rng(4)
    n = 400
    X = np.vstack([np.ones(n), rng.standard_normal(n)])
    mu = 1 / (1 + np.exp(-(-1.5 + 0.4 * X[1])))
    y = (rng.random(n) < mu).astype(float)
    a = oracle.xvx_inv_xv(X, mu * (1 - mu))
    g = rng.binomial(2, 0.2, n).astype(float)
    S, phi = OR.variant_scores(g[None], [1.0], mu * (1 - mu), y - mu, X, a, 1.0)
    want = 1.0 - gammainc(0.5, 0.5 * S[0] ** 2 / phi[0, 0])
    assert OR.burden(S, phi) == pytest.approx(want, rel=1e-13)
    assert OR.skat(S, phi) == pytest.approx(want, rel=1e-13)
    assert OR.skat_o(S, phi, R.default_rho_values())[0] == pytest.approx(want, rel=1e-13)
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.3), X, a, 1.0, y=y)
    assert eo.test_marker(g)["pvalue_noadj"] == pytest.approx(want, rel=1e-9)


@pytest.mark.parametrize("df", [0.3, 1.0, 2.0, 3.7, 10.0, 55.5, 300.0])
def test_library_chi2_tail_vs_scipy(df):
    lib = sg.load()
    for x in (1e-6, 0.01, 0.5, 1.0, df, 2 * df + 3, 5 * df + 30):
        want = 1.0 - gammainc(df / 2, x / 2)
        got = lib.saige_chi2_sf(df, x)
        assert abs(got - want) <= 1e-14 + 1e-12 * want, (df, x, got, want)
    assert lib.saige_chi2_sf(df, 0.0) == 1.0


def test_region_output_line(tmp_path):
    with open(tmp_path / "r.txt", "w") as f:
        R.write_region_header(f)
        R.write_region_line(f, "GENE1", 0.012, [2, 1, 0, 0, 0, 0, 0, 0], ["rs1", "rs2", "rs3"],
                            [0.001, 0.0025, 0.5], 0.5, 1e-7)
    lines = (tmp_path / "r.txt").read_text().splitlines()
    assert lines[0].split()[0] == "Gene" and len(lines[0].split()) == 17
    assert lines[1] == ("GENE1 0.012 2 1 0 0 0 0 0 0 rs1;rs2;rs3 0.001;0.0025;0.5 0.5 "
                        "0.0000001 0.012 0.5 0.0000001")


def _engine_case(n=3000, p=2, seed=3):
    rng = np.random.default_rng(seed)
    X = np.vstack([np.ones(n)] + [rng.standard_normal(n) for _ in range(p - 1)])
    mu = 1 / (1 + np.exp(-(-2.0 + 0.3 * X[-1] + 0.3 * rng.standard_normal(n))))
    y = (rng.random(n) < mu).astype(float)
    a = oracle.xvx_inv_xv(X, mu * (1 - mu))
    return X, mu, y, a, rng


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2])
def test_gpu_region_tests_vs_oracle(p):
    X, mu, y, a, rng = _engine_case(p=p)
    n = X.shape[1]
    sizes = [1, 3, 17, 64, 65, 130, 0, 9]
    nvar = sum(sizes)
    maf = rng.uniform(0.001, 0.05, nvar)
    G = rng.binomial(2, maf[:, None], (nvar, n)).astype(float)
    G[5, :40] = np.nan                      # missing calls read as 0
    G[20, y > 0.5] = 2.0                    # a signal
    w = R.beta_weights(maf)
    tau = (1.0, 0.4)
    e = sg.ScoreTestEngine(mu, y - mu, tau, X, a, 1.1, y=y)
    got = e.test_regions(G, sizes, w)
    s0 = 0
    for k, m in enumerate(sizes):
        want = OR.run_region_test(G[s0:s0 + m], w[s0:s0 + m], mu * (1 - mu), y - mu, X, a,
                                  tau[0], R.default_rho_values())
        s0 += m
        assert got["n_markers"][k] == m
        for f, wv in zip(("pvalue", "burden_pvalue", "skat_pvalue"), want[:3]):
            gv = got[f][k]
            # 1 - cdf carries an absolute error near 1e-16: tiny p-values agree absolutely
            assert abs(gv - wv) <= 1e-13 or abs(math.log10(gv) - math.log10(wv)) < 1e-6, \
                (k, f, gv, wv)
        assert got["rho"][k] == want[3] or abs(got["pvalue"][k] - want[0]) < 1e-12
    assert np.min(got["pvalue"]) < 1e-4  # the planted signal


@pytest.mark.gpu
def test_gpu_region_driver():
    X, mu, y, a, rng = _engine_case()
    n = X.shape[1]
    maf = rng.uniform(0.0005, 0.3, 60)
    G = rng.binomial(2, maf[:, None], (60, n)).astype(float)
    af = np.nanmean(G, axis=1) / 2
    mac = np.minimum(G.sum(1), 2 * n - G.sum(1))
    G[7] = 0.0
    mac[7], af[7] = 0.0, 0.0                # filtered out (MAC < 0.5)
    ids = [f"rs{i}" for i in range(60)]
    regions = {"A": list(range(0, 10)), "B": list(range(10, 35)), "C": [7], "D": [40, 41, 59]}
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.3), X, a, 1.0, y=y)
    out = R.run_regions(e, regions, G, af, mac, ids)
    assert [o["gene"] for o in out] == ["A", "B", "D"]   # C has no variant left
    assert out[0]["marker_ids"] == [f"rs{i}" for i in range(10) if i != 7]
    keep = [i for i in range(10) if i != 7]
    want = OR.run_region_test(G[keep], R.beta_weights(af[keep]), mu * (1 - mu), y - mu, X, a,
                              1.0, R.default_rho_values())
    assert abs(out[0]["pvalue"] - want[0]) <= 1e-13 or \
        abs(math.log10(out[0]["pvalue"]) - math.log10(want[0])) < 1e-6
    assert sum(out[1]["mac_category_counts"]) == 25


@pytest.mark.gpu
def test_gpu_region_tests_device_rows_match_host_rows():
    """saige_region_tests_device reads resident rows (with a leading dimension above n) and
    gives the host-row call's results bit for bit."""
    from csaige_amd import hip
    X, mu, y, a, rng = _engine_case(p=2, seed=9)
    n = X.shape[1]
    sizes = [4, 0, 33, 70, 1]
    nvar = sum(sizes)
    G = rng.binomial(2, rng.uniform(0.002, 0.05, nvar)[:, None], (nvar, n)).astype(float)
    G[3, :25] = np.nan
    w = R.beta_weights(np.nanmean(G, axis=1) / 2)
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.4), X, a, 1.0, y=y)
    want = e.test_regions(G, sizes, w)
    ld = n + 13
    padded = np.full((nvar, ld), 7.0)
    padded[:, :n] = G
    d = hip.DeviceBuffer(padded.nbytes)
    d.upload(np.ascontiguousarray(padded))
    got = e.test_regions(None, sizes, w, device_rows=(d.value, ld, nvar))
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
