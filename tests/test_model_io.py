"""The .saige.model file, SAIGE's result lines and the variance-ratio file (CPU).

The expected bytes of the model file are written out field by field from the bincode 1.3.3
encoding rules (little endian, u64 lengths, u32 enum index, arrays and tuples without a
length) and the field order of NullModel / VarianceRatioResult, for the model of the
reference's own round-trip test (saige-core/src/model/serialization.rs:70-105).
"""
import math

import numpy as np
import pytest

from csaige_amd import model as M


def _ref_test_model():
    # serialization.rs:73-93
    vr = M.VarianceRatioResult(0.94, [], 30, np.array([0.94]))
    return M.NullModel.new(M.TRAIT_BINARY, ["S1", "S2"], (1.0, 0.5), [0.1], [0.3, 0.7],
                           [0.0, 1.0], [1.0, 1.0], 1, [0.5, 0.5], vr)


def _h(s):
    return bytes.fromhex(s.replace(" ", ""))


def test_model_bytes_follow_bincode_layout():
    m = _ref_test_model()
    d = lambda *v: b"".join(np.array(v, dtype="<f8").tobytes() for v in v)  # noqa: E731
    want = b"".join([
        b"SGMD",                                    # magic [u8; 4], no length
        _h("01000000"),                             # version u32
        _h("00000000"),                             # trait_type: Binary = variant 0 (u32)
        _h("0200000000000000"),                     # sample_ids: Vec len 2
        _h("0200000000000000") + b"S1",             # String: u64 len + bytes
        _h("0200000000000000") + b"S2",
        _h("0200000000000000"),                     # n_samples usize
        _h("0100000000000000"),                     # n_covariates usize
        d(1.0, 0.5),                                # tau [f64; 2], no length
        _h("0100000000000000") + d(0.1),            # alpha
        _h("0200000000000000") + d(0.3, 0.7),       # mu
        _h("0200000000000000") + d(0.3 * 0.7, 0.7 * (1.0 - 0.7)),  # mu2 = mu (1 - mu)
        _h("0200000000000000") + d(0.0 - 0.3, 1.0 - 0.7),          # residuals = y - mu
        _h("0200000000000000") + d(0.0, 1.0),       # y
        _h("0200000000000000") + d(1.0, 1.0),       # x_flat
        _h("0100000000000000"),                     # x_ncols
        _h("0200000000000000") + d(0.5, 0.5),       # xvx_inv_xv_flat
        d(0.94),                                    # variance_ratio.variance_ratio
        _h("0000000000000000"),                     # categorical_vr: empty Vec
        _h("1e00000000000000"),                     # n_markers_used = 30
        _h("0100000000000000") + d(0.94),           # per_marker_vr
        _h("00"),                                   # use_loco: bool
        _h("0000000000000000") * 3,                 # loco_tau, loco_mu, loco_residuals
        _h("00"),                                   # use_sparse_grm
        _h("0000000000000000"),                     # n_markers_in_grm
    ])
    assert M.model_to_bytes(m) == want


def test_model_round_trip_and_reference_checks(tmp_path):
    m = _ref_test_model()
    path = tmp_path / "test.saige.model"
    M.save_model(m, path)
    got = M.load_model(path)
    # the reference test's assertions (serialization.rs:101-104)
    assert got.n_samples == 2
    assert got.tau == (1.0, 0.5)
    assert got.sample_ids == ["S1", "S2"]
    assert abs(got.variance_ratio.variance_ratio - 0.94) < 1e-10
    assert M.model_to_bytes(got) == M.model_to_bytes(m)


def test_model_round_trip_all_fields(tmp_path):
    rng = np.random.default_rng(3)
    n, p = 7, 2
    vr = M.VarianceRatioResult(1.02, [(1.5, 0.9), (10.5, 1.1), (math.inf, 1.0)], 33,
                               rng.random(33))
    m = M.NullModel.new(M.TRAIT_QUANTITATIVE, [f"id_{i}é" for i in range(n)], (0.7, 0.2),
                        rng.standard_normal(p), rng.random(n), rng.random(n),
                        rng.standard_normal(n * p), p, rng.standard_normal(n * p), vr)
    m.use_loco = True
    m.loco_tau = [("1", (0.7, 0.21)), ("22", (0.69, 0.19))]
    m.loco_mu = [("1", rng.random(n))]
    m.loco_residuals = [("1", rng.random(n)), ("2", rng.random(n))]
    m.use_sparse_grm = True
    m.n_markers_in_grm = 12345
    assert np.all(m.mu2 == 1.0)  # quantitative: mu2 = 1 (null_model.rs:87-90)
    M.save_model(m, tmp_path / "a.saige.model")
    g = M.load_model(tmp_path / "a.saige.model")
    assert M.model_to_bytes(g) == M.model_to_bytes(m)
    assert g.loco_tau == m.loco_tau and g.n_markers_in_grm == 12345 and g.use_sparse_grm
    assert g.variance_ratio.categorical_vr == vr.categorical_vr
    np.testing.assert_array_equal(g.loco_residuals[1][1], m.loco_residuals[1][1])
    # trailing bytes are ignored by bincode::deserialize
    assert M.model_to_bytes(M.model_from_bytes(M.model_to_bytes(m) + b"xyz")) == \
        M.model_to_bytes(m)


def test_model_rejects_bad_magic_and_truncation(tmp_path):
    b = bytearray(M.model_to_bytes(_ref_test_model()))
    b[0:4] = b"XXXX"
    (tmp_path / "bad.saige.model").write_bytes(bytes(b))
    with pytest.raises(ValueError, match=r"Invalid model file: expected magic bytes "
                                         r"\[83, 71, 77, 68\], got \[88, 88, 88, 88\]"):
        M.load_model(tmp_path / "bad.saige.model")
    (tmp_path / "short.saige.model").write_bytes(M.model_to_bytes(_ref_test_model())[:-3])
    with pytest.raises(ValueError, match="unexpected end of file"):
        M.load_model(tmp_path / "short.saige.model")
    b = bytearray(M.model_to_bytes(_ref_test_model()))
    b[8] = 7  # trait_type variant index out of range
    with pytest.raises(ValueError, match="variant index"):
        M.model_from_bytes(bytes(b))


@pytest.mark.parametrize("x,want", [
    (1.0, "1"), (100.0, "100"), (-2.5, "-2.5"), (0.1 + 0.2, "0.30000000000000004"),
    (1e-7, "0.0000001"), (1e21, "1000000000000000000000"), (0.94, "0.94"),
    (123456789012345680000.0, "123456789012345680000"), (0.0, "0"), (-0.0, "-0"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (-float("inf"), "-inf"),
    (5e-324, "0." + "0" * 323 + "5"), (1.7976931348623157e308, "17976931348623157" + "0" * 292),
    (0.000123, "0.000123"), (12.5e-3, "0.0125"), (2.0 ** 60, "1152921504606847000"),
])
def test_rust_f64_display(x, want):
    """Rust's `{}` for f64: shortest round-trip digits, never an exponent."""
    assert M.rust_f64(x) == want
    if math.isfinite(x):
        assert float(want) == x


def test_result_lines_and_header(tmp_path):
    r = dict(ac=3.0, af=0.0015, n=1000, beta=-0.25, se_beta=0.5, tstat=0.25, pvalue=0.617,
             pvalue_noadj=0.617, spa_converged=1, var_t=4.0, var_t_star=4.2,
             af_cases=float("nan"), af_controls=0.001, n_cases=0, n_controls=1000)
    res = {k: np.array([v, v]) for k, v in r.items()}
    M.write_results(tmp_path / "out.txt", res, ["1", "2"], [10, 20], ["rs1", "rs2"], ["A", "C"],
                    ["G", "T"])
    lines = (tmp_path / "out.txt").read_text().splitlines()
    # single_variant.rs:298-300
    assert lines[0] == ("CHR POS SNPID Allele1 Allele2 AC_Allele2 AF_Allele2 imputationInfo N "
                        "BETA SE Tstat p.value p.value.NA Is.SPA.converge varT varTstar "
                        "AF.Cases AF.Controls N.Cases N.Controls")
    assert lines[1] == "1 10 rs1 A G 3 0.0015 1 1000 -0.25 0.5 0.25 0.617 0.617 1 4 4.2 NA 0.001 0 1000"
    assert lines[2].startswith("2 20 rs2 C T ")
    assert len(lines) == 3


def test_variance_ratio_file_and_summaries(tmp_path):
    M.write_variance_ratio_file(M.VarianceRatioResult(0.94022084164312), tmp_path / "a.txt")
    assert (tmp_path / "a.txt").read_text() == "0.94022084164312 null 1\n"  # variance_ratio.rs:313-319
    M.write_variance_ratio_file(M.VarianceRatioResult(1.0, [(1.5, 0.8), (10.5, 0.9)]),
                                tmp_path / "b.txt")
    assert (tmp_path / "b.txt").read_text() == "0.8 null 1\n0.9 null 2\n"
    m = _ref_test_model()
    assert M.model_summary(m) == ("SAIGE Null Model v1\nTrait type: Binary\nSamples: 2\n"
                                  "Covariates: 1\nTau: [1.000000, 0.500000]\n"
                                  "Variance ratio: 0.9400\nN_eff: 2.0\nLOCO: no\nSparse GRM: no")
    # n_eff = 4 / (1/1 + 1/1) = 2 for one case and one control (null_model.rs:115-130)
    assert M.get_neff_report(m, "x.saige.model").splitlines()[3] == "N_eff: 2.00"
