"""CPU ORACLE -- test infrastructure only.

numpy restatement of the region tests of saige-core/src/score_test/region.rs, in the
reference's order of operations where it matters: compute_variant_scores (:75-135) with the
explicit g~ = gw - X (A gw), burden_test (:143-175), skat_test (:184-233), skat_o_test
(:246-339), run_region_test (:342-358).  The chi-square CDF (statrs, third-party) is
1 - scipy.special.gammainc(df / 2, x / 2) here: parity unpinned at the last bits.
Pinned by the reference's region tests (region.rs:431-523) and by the m = 1 closed form
(every test reduces to the chi-square(1) test of S^2 / Phi).
"""
import numpy as np
from scipy.special import gammainc


def _chi2_cdf(df, x):
    return 0.0 if x <= 0 else float(gammainc(df / 2.0, x / 2.0))


def variant_scores(G, weights, mu2, residuals, x_cm, a_flat, tau_e):
    """S_j = g~_j' res and Phi_jk = tau_e g~_j' diag(mu2) g~_k (region.rs:75-135)."""
    G = np.asarray(G, dtype=np.float64)
    x = np.asarray(x_cm, dtype=np.float64)            # p x n
    p, n = x.shape
    A = np.asarray(a_flat, dtype=np.float64).reshape(n, p).T  # col-major p x n view
    gts = []
    for g, w in zip(G, weights):
        gw = w * np.where(np.isnan(g), 0.0, g)
        gts.append(gw - x.T @ (A @ gw))
    gts = np.array(gts)
    S = gts @ np.asarray(residuals)
    m = len(gts)
    phi = np.zeros((m, m))
    for j in range(m):
        for k in range(j, m):
            d = float(np.sum(gts[j] * mu2 * gts[k])) * tau_e
            phi[j, k] = phi[k, j] = d
    return S, phi


def burden(S, phi):
    if len(S) == 0:
        return 1.0
    sum_phi = float(phi.sum())
    if sum_phi < 1e-30:
        return 1.0
    q = float(S.sum()) ** 2 / sum_phi
    return 1.0 - _chi2_cdf(1.0, q)


def skat(S, phi):
    if len(S) == 0:
        return 1.0
    q = float(np.sum(S * S))
    mean_q = float(np.trace(phi))
    if mean_q < 1e-30:
        return 1.0
    var_q = 2.0 * float(np.sum(phi * phi))
    if var_q < 1e-30:
        return 1.0
    kappa = var_q / (2.0 * mean_q)
    df = 2.0 * mean_q * mean_q / var_q
    return 1.0 - _chi2_cdf(df, q / kappa) if df > 0 and kappa > 0 else 1.0


def skat_o(S, phi, rho_values):
    m = len(S)
    if m == 0:
        return 1.0, 0.0
    q_skat = float(np.sum(S * S))
    q_burden = float(S.sum()) ** 2
    sum_phi = float(phi.sum())
    trace_phi = float(np.trace(phi))
    row = phi.sum(axis=1)
    best, best_rho = 1.0, 0.0
    for rho in rho_values:
        q_rho = (1.0 - rho) * q_skat + rho * q_burden
        mean_q = (1.0 - rho) * trace_phi + rho * sum_phi
        if mean_q < 1e-30:
            continue
        a = (1.0 - rho) * phi + rho * row[None, :]      # [j, k]: (1-rho) Phi_jk + rho row_k
        b = (1.0 - rho) * phi.T + rho * row[:, None]    # [j, k]: (1-rho) Phi_kj + rho row_j
        var_q = 2.0 * float(np.sum(a * b))
        if var_q < 1e-30:
            continue
        kappa = var_q / (2.0 * mean_q)
        df = 2.0 * mean_q * mean_q / var_q
        if df > 0 and kappa > 0:
            pv = 1.0 - _chi2_cdf(df, q_rho / kappa)
            if pv < best:
                best, best_rho = pv, rho
    return best, best_rho


def run_region_test(G, weights, mu2, residuals, x_cm, a_flat, tau_e, rho_values):
    """(p_skato, p_burden, p_skat, rho) of one region (region.rs:342-358)."""
    if len(G) == 0:
        return 1.0, 1.0, 1.0, 0.0
    S, phi = variant_scores(G, weights, mu2, residuals, x_cm, a_flat, tau_e)
    p_o, rho = skat_o(S, phi, rho_values)
    return p_o, burden(S, phi), skat(S, phi), rho
