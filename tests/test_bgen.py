"""BGEN v1.2 input: the host reader and the oracle's layout-2 decode against the reference's
.bgen fixtures and the genotype calls of the same markers in their .vcf files (CPU), and the
device decode and score test against the oracle (GPU)."""
import math
import os
import struct

import numpy as np
import pytest

import csaige_amd as sg
import oracle
from csaige_amd import bgen

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bgen")


def _path(name):
    return os.path.join(GOLD, name)


def _expected(name):
    z = np.load(_path(name + ".expected.npz"))
    alt = z["alt"].astype(np.float64)
    alt[z["alt"] < 0] = np.nan
    return alt, list(z["samples"]), list(z["ids"])


def _same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)],
                                                                         b[~np.isnan(b)])


@pytest.mark.parametrize("name", ["genotype_100markers", "genotype_10markers.missingness",
                                  "genotype_100markers_2chr"])
def test_reader_and_oracle_decode_match_vcf_calls(name):
    r = bgen.BgenReader(_path(name + ".bgen"))
    alt, samples, ids = _expected(name)
    assert r.header.layout == 2 and r.header.compression == "zlib"
    assert r.n_markers() == len(ids) and [e.rsid for e in r.index] == ids
    assert r.n_samples() == len(samples)
    if name != "genotype_10markers.missingness":  # IDs from the .samples file or the file
        assert r.sample_ids == samples
    for v in range(r.n_markers()):
        d = oracle.bgen_layout2_dosages(r.layout2_block(v))
        assert _same(d, alt[v]), v
    assert r.marker_info(0) == dict(chrom="1", pos=int(r.index[0].pos), id=ids[0],
                                    ref_allele="A", alt_allele="C")


def test_reader_index_sources_and_subset():
    r = bgen.BgenReader(_path("genotype_100markers.bgen"))  # .bgi present
    seq = bgen.build_sequential_index(r.data, r.header)
    assert [(e.file_offset, e.size, e.chrom, e.pos, e.rsid, e.allele1, e.allele2)
            for e in r.index] == \
           [(e.file_offset, e.size, e.chrom, e.pos, e.rsid, e.allele1, e.allele2) for e in seq]
    r.set_sample_subset(["1a5", "nope", "1a2", "1a5"])  # bgen.rs:466-480
    assert r.sample_ids == ["1a5", "1a2", "1a5"]
    np.testing.assert_array_equal(r.sample_subset, [4, 1, 4])
    assert r.n_samples() == 3
    b = r.batch(0, 7)
    assert b.n_variants == 7 and b.n_file == 10000 and b.ld % 16 == 0
    assert np.all(b.info == (8 | 0x100))


def _layout2_block(probs, bits, missing=None, n_alleles=2, ploidy=(2, 2), phased=0):
    """A layout-2 block for biallelic diploid samples: probs is n x 2 integer (P(AA), P(AB))
    values below 2^bits, packed least significant bit first."""
    n = len(probs)
    missing = np.zeros(n, bool) if missing is None else missing
    head = struct.pack("<IHBB", n, n_alleles, ploidy[0], ploidy[1])
    pl = bytes(((0x80 if m else 0) | 2) for m in missing)
    acc, nbit, out = 0, 0, bytearray()
    for a, c in probs:
        for val in (int(a), int(c)):
            acc |= val << nbit
            nbit += bits
            while nbit >= 8:
                out.append(acc & 0xFF)
                acc >>= 8
                nbit -= 8
    if nbit:
        out.append(acc & 0xFF)
    return head + pl + bytes([phased, bits]) + bytes(out)


def _batch(blocks):
    n = struct.unpack_from("<I", blocks[0], 0)[0]
    info = []
    for b in blocks:
        k, lo, hi = struct.unpack_from("<HBB", b, 4)
        std = k == 2 and lo == 2 and hi == 2 and b[8 + n] == 0
        info.append(b[9 + n] | (0x100 if std else 0))
    ld = (max(len(b) for b in blocks) + 15) // 16 * 16
    arr = np.zeros((len(blocks), ld), np.uint8)
    for v, b in enumerate(blocks):
        arr[v, :len(b)] = np.frombuffer(b, np.uint8)
    return bgen.BgenBatch(arr, np.array(info, np.uint32), ld, n)


def _synthetic_blocks(n=333, seed=2):
    rng = np.random.default_rng(seed)
    out = []
    for bits in (8, 5, 16, 1, 32, 3):
        mx = (1 << bits) - 1
        a = rng.integers(0, mx + 1, n)
        c = np.minimum(rng.integers(0, mx + 1, n), mx - a)
        out.append(_layout2_block(np.stack([a, c], 1), bits, rng.random(n) < 0.05))
    p = np.stack([np.full(n, 200), np.full(n, 30)], 1)
    out.append(_layout2_block(p, 8, phased=1))          # phased: the reference's 0.0 fallback
    out.append(_layout2_block(p, 8, n_alleles=3))       # multi-allelic: same
    out.append(_layout2_block(p, 8, ploidy=(1, 2)))     # mixed ploidy: same
    return out


def test_oracle_decode_closed_forms():
    """dosage = P(AB) + 2 (1 - P(AA) - P(AB)) with P = value / (2^bits - 1), LSB-first
    packing, NaN for missing samples, 0 for the other layouts (bgen.rs:403-430)."""
    blocks = _synthetic_blocks()
    n = 333
    rng = np.random.default_rng(2)
    for bits in (8, 5, 16, 1, 32, 3):
        mx = (1 << bits) - 1
        a = rng.integers(0, mx + 1, n)
        c = np.minimum(rng.integers(0, mx + 1, n), mx - a)
        miss = rng.random(n) < 0.05
        d = oracle.bgen_layout2_dosages(blocks[[8, 5, 16, 1, 32, 3].index(bits)])
        want = c / mx + 2.0 * (1.0 - a / mx - c / mx)
        want[miss] = np.nan
        assert _same(d, want), bits
    for b in blocks[6:]:
        assert np.all(oracle.bgen_layout2_dosages(b) == 0.0)
    with pytest.raises(ValueError, match="0 bits per probability"):
        oracle.bgen_layout2_dosages(_layout2_block(np.zeros((4, 2), int), 0))


@pytest.mark.gpu
def test_gpu_bgen_decode_bit_exact():
    r = bgen.BgenReader(_path("genotype_10markers.missingness.bgen"))
    b = r.batch(0, r.n_markers())
    got = sg.bgen_dosages(b)
    for v in range(r.n_markers()):
        assert _same(got[v], oracle.bgen_layout2_dosages(r.layout2_block(v))), v
    idx = np.array([999, 4, 7, 0, 4], np.uint32)
    got = sg.bgen_dosages(b, sample_idx=idx)
    for v in range(r.n_markers()):
        assert _same(got[v], oracle.bgen_layout2_dosages(r.layout2_block(v))[idx]), v
    blocks = _synthetic_blocks()
    got = sg.bgen_dosages(_batch(blocks))
    for v, blk in enumerate(blocks):
        assert _same(got[v], oracle.bgen_layout2_dosages(blk)), v


@pytest.mark.gpu
@pytest.mark.parametrize("name,n_sub", [("genotype_100markers", 3000),
                                        ("genotype_10markers.missingness", None)])
def test_gpu_score_test_from_bgen_matches_dosage_path(name, n_sub):
    """test_markers_bgen (decode fused into the score test) against test_markers on the
    oracle's decoded dosages (identical sums) and against the oracle's score test: 100
    markers under a shuffled sample subset, and 10 markers with missing calls."""
    rng = np.random.default_rng(5)
    r = bgen.BgenReader(_path(name + ".bgen"))
    if n_sub:
        r.set_sample_subset([r.sample_ids[i] for i in rng.permutation(r.header.n_samples)[:n_sub]])
    n = r.n_samples()
    X = np.vstack([np.ones(n), rng.standard_normal(n)])
    mu = 1.0 / (1.0 + np.exp(-(-2.0 + 0.3 * X[1] + 0.3 * rng.standard_normal(n))))
    y = (rng.random(n) < mu).astype(float)
    a = oracle.xvx_inv_xv(X, mu * (1 - mu))
    e = sg.ScoreTestEngine(mu, y - mu, (1.0, 0.2), X, a, 1.05, y=y, spa_pval_cutoff=0.05)
    b = r.batch(0, r.n_markers())
    got = e.test_markers_bgen(b, sample_idx=r.sample_subset)
    sub = r.sample_subset if n_sub else slice(None)
    G = np.vstack([oracle.bgen_layout2_dosages(r.layout2_block(v))[sub]
                   for v in range(r.n_markers())])
    ref = e.test_markers(G)
    for k in ("ac", "n", "score", "var_t", "pvalue", "pvalue_noadj", "is_spa", "af_cases"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    if not n_sub:
        assert np.any(got["n"] < n)  # missing calls reached the engine
    eo = oracle.ScoreEngine(mu, y - mu, (1.0, 0.2), X, a, 1.05, y=y, spa_cutoff=0.05)
    for v, g in enumerate(G):
        w = eo.test_marker(g)
        assert got["n"][v] == w["n"]
        if np.isnan(w["score"]):  # a missing call makes A g and S NaN (single_variant.rs:137-144)
            assert np.isnan(got["score"][v]) and got["pvalue"][v] == w["pvalue"]
            continue
        assert got["ac"][v] == w["ac"]
        if w["pvalue"] > 1e-300:
            assert abs(math.log10(got["pvalue"][v]) - math.log10(w["pvalue"])) < 1e-6
