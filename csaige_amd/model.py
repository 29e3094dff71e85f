"""The step-1 -> step-2 file boundary of saige-rs, so a null model fitted here can be tested by
``saige test`` and a model written by ``saige fit-null`` can be tested here.

- ``NullModel`` / ``VarianceRatioResult``: the fields of NullModel
  (saige-core/src/model/null_model.rs:17-60, constructor :70-113, n_eff :115-130) and of
  VarianceRatioResult (saige-core/src/glmm/variance_ratio.rs:61-73), in declaration order.
- ``save_model`` / ``load_model``: the .saige.model file (model/serialization.rs:13-34), the
  struct encoded as bincode 1.3.3 does with its default (legacy) options, which is what
  ``bincode::serialize`` / ``bincode::deserialize`` use: little endian, fixed-width integers,
  usize as u64, a u64 length before every Vec and String (UTF-8 bytes), fixed-size arrays
  and tuples without a length, a unit enum variant as its u32 index, bool as one byte 0/1.
  ``load_model`` then checks the magic bytes with the reference's message.
- ``model_summary`` and ``get_neff_report``: the texts of serialization.rs:43-64 and of
  ``saige get-neff`` (saige-cli/src/commands/get_neff.rs:17-30).
- ``write_results_header`` / ``write_result_line`` / ``write_results``: SAIGE's text output
  (saige-core/src/score_test/single_variant.rs:293-343), numbers printed as Rust's ``{}``
  prints f64 (``rust_f64``).
- ``write_variance_ratio_file``: the ``.varianceRatio.txt`` file (variance_ratio.rs:313-333).
- ``NullModel.from_fit_null`` builds the model the way fit_null.rs:382-394 does from a
  ``fit_null`` result, and ``engine_from_model`` the step-2 engine the way
  assoc_test.rs:116-140 does from a loaded model.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import BinaryIO, List, Sequence, Tuple

import numpy as np

from ._lib import TRAIT_BINARY, TRAIT_QUANTITATIVE, TRAIT_SURVIVAL

MAGIC = b"SGMD"  # null_model.rs:64
VERSION = 1      # null_model.rs:66
TRAIT_NAMES = {TRAIT_BINARY: "Binary", TRAIT_QUANTITATIVE: "Quantitative",
               TRAIT_SURVIVAL: "Survival"}  # link.rs:7-11, variant order = bincode index


@dataclass
class VarianceRatioResult:
    variance_ratio: float
    categorical_vr: List[Tuple[float, float]] = field(default_factory=list)
    n_markers_used: int = 0
    per_marker_vr: np.ndarray = field(default_factory=lambda: np.zeros(0))


@dataclass
class NullModel:
    trait_type: int
    sample_ids: List[str]
    n_samples: int
    n_covariates: int
    tau: Tuple[float, float]
    alpha: np.ndarray
    mu: np.ndarray
    mu2: np.ndarray
    residuals: np.ndarray
    y: np.ndarray
    x_flat: np.ndarray
    x_ncols: int
    xvx_inv_xv_flat: np.ndarray
    variance_ratio: VarianceRatioResult
    use_loco: bool = False
    loco_tau: List[Tuple[str, Tuple[float, float]]] = field(default_factory=list)
    loco_mu: List[Tuple[str, np.ndarray]] = field(default_factory=list)
    loco_residuals: List[Tuple[str, np.ndarray]] = field(default_factory=list)
    use_sparse_grm: bool = False
    n_markers_in_grm: int = 0
    magic: bytes = MAGIC
    version: int = VERSION

    @classmethod
    def new(cls, trait_type, sample_ids, tau, alpha, mu, y, x_flat, x_ncols, xvx_inv_xv_flat,
            variance_ratio):
        """NullModel::new (null_model.rs:70-113): mu2 = mu(1 - mu) for binary and survival
        traits, 1 for quantitative ones; residuals = y - mu; no LOCO, no sparse GRM."""
        mu = np.asarray(mu, dtype=np.float64).copy()
        y = np.asarray(y, dtype=np.float64).copy()
        n = len(sample_ids)
        if trait_type == TRAIT_QUANTITATIVE:
            mu2 = np.ones(n)
        else:
            mu2 = mu * (1.0 - mu)
        return cls(trait_type=int(trait_type), sample_ids=[str(s) for s in sample_ids],
                   n_samples=n, n_covariates=int(x_ncols), tau=(float(tau[0]), float(tau[1])),
                   alpha=np.asarray(alpha, dtype=np.float64).copy(), mu=mu, mu2=mu2,
                   residuals=y - mu, y=y,
                   x_flat=np.asarray(x_flat, dtype=np.float64).ravel().copy(),
                   x_ncols=int(x_ncols),
                   xvx_inv_xv_flat=np.asarray(xvx_inv_xv_flat, dtype=np.float64).ravel().copy(),
                   variance_ratio=variance_ratio)

    @classmethod
    def from_fit_null(cls, fit, y, x_cm, sample_ids, trait_type=TRAIT_BINARY):
        """The model fit_null.rs:382-394 saves, from a ``fit_null`` result: x_cm is the p x n
        covariate array (rows = covariates), i.e. X flat col-major n x p."""
        x = np.asarray(x_cm, dtype=np.float64)
        vr = VarianceRatioResult(float(fit["vr"]),
                                 [(float(u), float(r)) for u, r in fit.get("categorical_vr", [])],
                                 int(fit["vr_markers_used"]),
                                 np.asarray(fit["per_marker_vr"], dtype=np.float64).copy())
        return cls.new(trait_type, sample_ids, fit["tau"], fit["alpha"], fit["mu"], y,
                       x.ravel(), x.shape[0], fit["xvx_inv_xv"], vr)

    def n_eff(self) -> float:
        """null_model.rs:115-130: 4 / (1/cases + 1/controls) for binary traits."""
        if self.trait_type == TRAIT_BINARY:
            cases = float(np.sum(self.y > 0.5))
            controls = float(np.sum(self.y < 0.5))
            if cases > 0.0 and controls > 0.0:
                return 4.0 / (1.0 / cases + 1.0 / controls)
        return float(self.n_samples)

    def x_cm(self) -> np.ndarray:
        """X as a p x n array (rows = covariates)."""
        return self.x_flat.reshape(self.x_ncols, self.n_samples)


# ------------------------------------------------------------------------- bincode 1.3.3
class _Writer:
    def __init__(self):
        self.parts = []

    def u8(self, v):
        self.parts.append(struct.pack("<B", v))

    def boolean(self, v):
        self.parts.append(b"\x01" if v else b"\x00")

    def u32(self, v):
        self.parts.append(struct.pack("<I", v))

    def u64(self, v):
        self.parts.append(struct.pack("<Q", v))

    def f64(self, v):
        self.parts.append(struct.pack("<d", v))

    def string(self, s):
        b = s.encode("utf-8")
        self.u64(len(b))
        self.parts.append(b)

    def vec_f64(self, a):
        a = np.ascontiguousarray(a, dtype="<f8")
        self.u64(a.size)
        self.parts.append(a.tobytes())


class _Reader:
    def __init__(self, data: bytes):
        self.d = memoryview(data)
        self.o = 0

    def take(self, k):
        if self.o + k > len(self.d):
            raise ValueError("io error: unexpected end of file")
        b = self.d[self.o:self.o + k]
        self.o += k
        return b

    def u8(self):
        return self.take(1)[0]

    def boolean(self):
        v = self.u8()
        if v > 1:
            raise ValueError(f"invalid value: integer `{v}`, expected a boolean")
        return v == 1

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def f64(self):
        return struct.unpack("<d", self.take(8))[0]

    def string(self):
        k = self.u64()
        return bytes(self.take(k)).decode("utf-8")

    def vec_f64(self):
        k = self.u64()
        return np.frombuffer(bytes(self.take(8 * k)), dtype="<f8").astype(np.float64)


def model_to_bytes(m: NullModel) -> bytes:
    """bincode::serialize(&NullModel) (fields in null_model.rs:17-60 order)."""
    w = _Writer()
    if len(m.magic) != 4:
        raise ValueError("magic must be 4 bytes")
    for b in m.magic:
        w.u8(b)
    w.u32(m.version)
    w.u32(m.trait_type)
    w.u64(len(m.sample_ids))
    for s in m.sample_ids:
        w.string(s)
    w.u64(m.n_samples)
    w.u64(m.n_covariates)
    w.f64(m.tau[0])
    w.f64(m.tau[1])
    for a in (m.alpha, m.mu, m.mu2, m.residuals, m.y, m.x_flat):
        w.vec_f64(a)
    w.u64(m.x_ncols)
    w.vec_f64(m.xvx_inv_xv_flat)
    vr = m.variance_ratio
    w.f64(vr.variance_ratio)
    w.u64(len(vr.categorical_vr))
    for upper, ratio in vr.categorical_vr:
        w.f64(upper)
        w.f64(ratio)
    w.u64(vr.n_markers_used)
    w.vec_f64(vr.per_marker_vr)
    w.boolean(m.use_loco)
    w.u64(len(m.loco_tau))
    for chrom, t in m.loco_tau:
        w.string(chrom)
        w.f64(t[0])
        w.f64(t[1])
    for lst in (m.loco_mu, m.loco_residuals):
        w.u64(len(lst))
        for chrom, v in lst:
            w.string(chrom)
            w.vec_f64(v)
    w.boolean(m.use_sparse_grm)
    w.u64(m.n_markers_in_grm)
    return b"".join(w.parts)


def model_from_bytes(data: bytes) -> NullModel:
    """bincode::deserialize::<NullModel> (trailing bytes are ignored, as there)."""
    r = _Reader(data)
    magic = bytes(r.u8() for _ in range(4))
    version = r.u32()
    trait_type = r.u32()
    if trait_type > 2:
        raise ValueError(f"invalid value: integer `{trait_type}`, expected variant index 0 <= i < 3")
    ids = [r.string() for _ in range(r.u64())]
    n_samples = r.u64()
    n_cov = r.u64()
    tau = (r.f64(), r.f64())
    alpha, mu, mu2, residuals, y, x_flat = (r.vec_f64() for _ in range(6))
    x_ncols = r.u64()
    xvx = r.vec_f64()
    vr_value = r.f64()
    cat = [(r.f64(), r.f64()) for _ in range(r.u64())]
    used = r.u64()
    per = r.vec_f64()
    use_loco = r.boolean()
    loco_tau = []
    for _ in range(r.u64()):
        chrom = r.string()
        loco_tau.append((chrom, (r.f64(), r.f64())))
    lists = []
    for _ in range(2):
        lst = []
        for _ in range(r.u64()):
            chrom = r.string()
            lst.append((chrom, r.vec_f64()))
        lists.append(lst)
    use_sparse = r.boolean()
    n_in_grm = r.u64()
    return NullModel(trait_type=trait_type, sample_ids=ids, n_samples=n_samples,
                     n_covariates=n_cov, tau=tau, alpha=alpha, mu=mu, mu2=mu2,
                     residuals=residuals, y=y, x_flat=x_flat, x_ncols=x_ncols,
                     xvx_inv_xv_flat=xvx,
                     variance_ratio=VarianceRatioResult(vr_value, cat, used, per),
                     use_loco=use_loco, loco_tau=loco_tau, loco_mu=lists[0],
                     loco_residuals=lists[1], use_sparse_grm=use_sparse,
                     n_markers_in_grm=n_in_grm, magic=magic, version=version)


def save_model(model: NullModel, path) -> None:
    """save_model (serialization.rs:13-17)."""
    with open(path, "wb") as f:
        f.write(model_to_bytes(model))


def load_model(path) -> NullModel:
    """load_model (serialization.rs:20-34): decode, then check the magic bytes."""
    with open(path, "rb") as f:
        m = model_from_bytes(f.read())
    if m.magic != MAGIC:
        raise ValueError(f"Invalid model file: expected magic bytes {list(MAGIC)}, "
                         f"got {list(m.magic)}")
    return m


# ------------------------------------------------------------------------- text output
def rust_f64(x) -> str:
    """``format!("{}", x)`` for an f64: the shortest digits that read back to x, written
    without an exponent; "NaN", "inf", "-inf", "0" / "-0"."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0"
    r = repr(abs(x))  # the shortest round-trip digits (same as Rust's shortest strategy)
    if "e" in r:
        mant, exp = r.split("e")
        exp = int(exp)
    else:
        mant, exp = r, 0
    ip, _, fp = mant.partition(".")
    digits = ip + fp
    point = len(ip) + exp  # the decimal point sits after `point` digits
    stripped = digits.lstrip("0")
    point -= len(digits) - len(stripped)
    digits = stripped.rstrip("0")
    if point <= 0:
        return sign + "0." + "0" * (-point) + digits
    if point >= len(digits):
        return sign + digits + "0" * (point - len(digits))
    return sign + digits[:point] + "." + digits[point:]


def _fixed(x, d) -> str:
    """``format!("{:.d}", x)`` for an f64."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return f"{x:.{d}f}"


HEADER = ("CHR POS SNPID Allele1 Allele2 AC_Allele2 AF_Allele2 imputationInfo N BETA SE Tstat "
          "p.value p.value.NA Is.SPA.converge varT varTstar AF.Cases AF.Controls N.Cases "
          "N.Controls")


def write_results_header(f) -> None:
    """write_results_header (single_variant.rs:296-302)."""
    f.write(HEADER + "\n")


def write_result_line(f, chrom, pos, marker_id, ref_allele, alt_allele, r, info=1.0) -> None:
    """write_result_line (single_variant.rs:305-343) for one result ``r`` (a mapping with
    the SingleVariantResult fields of this package's results)."""
    def na(v):
        return "NA" if math.isnan(float(v)) else rust_f64(v)
    fields = [str(chrom), str(int(pos)), str(marker_id), str(ref_allele), str(alt_allele),
              rust_f64(r["ac"]), rust_f64(r["af"]), rust_f64(info), str(int(r["n"])),
              rust_f64(r["beta"]), rust_f64(r["se_beta"]), rust_f64(r["tstat"]),
              rust_f64(r["pvalue"]), rust_f64(r["pvalue_noadj"]),
              "1" if r["spa_converged"] else "0", rust_f64(r["var_t"]),
              rust_f64(r["var_t_star"]), na(r["af_cases"]), na(r["af_controls"]),
              str(int(r["n_cases"])), str(int(r["n_controls"]))]
    f.write(" ".join(fields) + "\n")


def write_results(path, results, chrom, pos, marker_id, ref_allele, alt_allele, info=None):
    """The single-variant output file of ``saige test`` (assoc_test.rs:154-214): the header,
    then one line per variant of ``results`` (the dict of arrays the engine returns) with
    the per-variant marker columns given as sequences."""
    nv = len(results["pvalue"])
    with open(path, "w") as f:
        write_results_header(f)
        for v in range(nv):
            r = {k: results[k][v] for k in ("ac", "af", "n", "beta", "se_beta", "tstat",
                                            "pvalue", "pvalue_noadj", "spa_converged", "var_t",
                                            "var_t_star", "af_cases", "af_controls", "n_cases",
                                            "n_controls")}
            write_result_line(f, chrom[v], pos[v], marker_id[v], ref_allele[v], alt_allele[v],
                              r, 1.0 if info is None else info[v])


def write_variance_ratio_file(vr: VarianceRatioResult, path) -> None:
    """write_variance_ratio_file (variance_ratio.rs:313-333)."""
    with open(path, "w") as f:
        if not vr.categorical_vr:
            f.write(f"{rust_f64(vr.variance_ratio)} null 1\n")
        else:
            for k, (_, ratio) in enumerate(vr.categorical_vr):
                f.write(f"{rust_f64(ratio)} null {k + 1}\n")


def model_summary(m: NullModel) -> str:
    """model_summary (serialization.rs:43-64)."""
    return (f"SAIGE Null Model v{m.version}\n"
            f"Trait type: {TRAIT_NAMES.get(m.trait_type, m.trait_type)}\n"
            f"Samples: {m.n_samples}\n"
            f"Covariates: {m.n_covariates}\n"
            f"Tau: [{_fixed(m.tau[0], 6)}, {_fixed(m.tau[1], 6)}]\n"
            f"Variance ratio: {_fixed(m.variance_ratio.variance_ratio, 4)}\n"
            f"N_eff: {_fixed(m.n_eff(), 1)}\n"
            f"LOCO: {'yes' if m.use_loco else 'no'}\n"
            f"Sparse GRM: {'yes' if m.use_sparse_grm else 'no'}")


def get_neff_report(m: NullModel, model_file: str) -> str:
    """The lines ``saige get-neff`` prints (get_neff.rs:17-30)."""
    return (f"Model: {model_file}\n"
            f"Trait type: {TRAIT_NAMES.get(m.trait_type, m.trait_type)}\n"
            f"N samples: {m.n_samples}\n"
            f"N_eff: {_fixed(m.n_eff(), 2)}\n"
            f"Tau: [{_fixed(m.tau[0], 6)}, {_fixed(m.tau[1], 6)}]\n"
            f"Variance ratio: {_fixed(m.variance_ratio.variance_ratio, 4)}")


def engine_from_model(m: NullModel, is_spa=True, is_fast_spa=True, spa_pval_cutoff=0.05,
                      device=0):
    """The step-2 engine assoc_test.rs:116-140 builds from a loaded model: SPA only for
    binary traits, case/control columns only for binary traits, spa_tol 1e-6."""
    from . import ScoreTestEngine
    binary = m.trait_type == TRAIT_BINARY
    return ScoreTestEngine(m.mu, m.residuals, m.tau, m.x_cm(), m.xvx_inv_xv_flat,
                           m.variance_ratio.variance_ratio, y=m.y if binary else None,
                           use_spa=bool(is_spa and binary), use_fast_spa=bool(is_fast_spa),
                           spa_tol=1e-6, spa_pval_cutoff=spa_pval_cutoff,
                           categorical_vr=list(m.variance_ratio.categorical_vr), mu2=m.mu2,
                           device=device)
