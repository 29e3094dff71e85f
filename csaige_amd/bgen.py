"""BGEN v1.2 input for step 2 (BgenReader, saige-geno/src/bgen.rs).

The host side does what is file work in the reference: the header (bgen.rs:114-171), the
sample IDs from the file (:173-193) or a .sample file (:195-209, candidates :84-88), the
variant index from the .bgi SQLite file (:211-234) or a sequential scan (:236-310), the
sample subset (:466-480) and the per-variant decompression (zlib; zstd when the
``zstandard`` module is importable) of :313-368.  The decompressed layout-2 blocks of a
batch of variants go to the device, where the probabilities become dosages inside the score
test (``ScoreTestEngine.test_markers_bgen``) or in ``csaige_amd.bgen_dosages``.
"""
from __future__ import annotations

import os
import sqlite3
import struct
import zlib
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

COMPRESSION = {0: "none", 1: "zlib", 2: "zstd"}


@dataclass
class BgenHeader:
    offset: int
    n_variants: int
    n_samples: int
    header_length: int
    compression: str
    layout: int
    has_sample_ids: bool


@dataclass
class BgenIndexEntry:
    file_offset: int
    size: int
    chrom: str
    pos: int
    rsid: str
    allele1: str
    allele2: str


@dataclass
class BgenBatch:
    """Decompressed layout-2 blocks of consecutive variants, zero-filled to a common stride."""
    blocks: np.ndarray   # uint8, n_variants x ld
    info: np.ndarray     # uint32: bits per probability | 0x100 for diploid biallelic unphased
    ld: int
    n_file: int

    @property
    def n_variants(self):
        return int(self.blocks.shape[0])


def _u16(d, o):
    return struct.unpack_from("<H", d, o)[0]


def _u32(d, o):
    return struct.unpack_from("<I", d, o)[0]


def _text(b: bytes) -> str:
    try:
        return b.decode("utf-8")
    except UnicodeDecodeError:
        return "?"  # bgen.rs: from_utf8(...).unwrap_or("?")


def parse_header(data: bytes) -> BgenHeader:
    """bgen.rs:114-171."""
    if len(data) < 20:
        raise ValueError("BGEN file too small")
    offset, header_length, n_variants, n_samples = struct.unpack_from("<IIII", data, 0)
    magic = data[16:20]
    if magic != b"bgen" and magic != b"\x00\x00\x00\x00":
        raise ValueError(f"Invalid BGEN magic number: {list(magic)}")
    if 20 + header_length - 20 > len(data):
        raise ValueError("BGEN header extends beyond file")
    flags = _u32(data, 4 + header_length - 4)
    comp = flags & 0x03
    if comp not in COMPRESSION:
        raise ValueError("Unknown BGEN compression type")
    layout = (flags >> 2) & 0x0F
    if layout != 2:
        raise ValueError(f"Only BGEN layout 2 (v1.2) is supported, got layout {layout}")
    return BgenHeader(offset, n_variants, n_samples, header_length, COMPRESSION[comp], layout,
                      bool((flags >> 31) & 1))


def parse_sample_ids(data: bytes, h: BgenHeader) -> List[str]:
    """bgen.rs:173-193: the sample identifier block after the header."""
    start = 4 + h.header_length
    n = _u32(data, start + 4)
    o = start + 8
    ids = []
    for _ in range(n):
        k = _u16(data, o)
        ids.append(_text(data[o + 2:o + 2 + k]))
        o += 2 + k
    return ids


def parse_sample_file(path) -> List[str]:
    """bgen.rs:195-209: first field of every line after the two header lines."""
    ids = []
    with open(path) as f:
        for i, line in enumerate(f.read().splitlines()):
            if i < 2:
                continue
            fields = line.split()
            if fields:
                ids.append(fields[0])
    return ids


def read_bgi_index(path) -> List[BgenIndexEntry]:
    """bgen.rs:211-234 (the Variant table of the .bgi SQLite index)."""
    con = sqlite3.connect(f"file:{os.fspath(path)}?mode=ro", uri=True)
    try:
        rows = con.execute("SELECT file_start_position, size_in_bytes, chromosome, position, "
                           "rsid, allele1, allele2 FROM Variant ORDER BY "
                           "file_start_position").fetchall()
    finally:
        con.close()
    return [BgenIndexEntry(int(r[0]), int(r[1]), str(r[2]), int(r[3]), str(r[4]), str(r[5]),
                           str(r[6])) for r in rows]


def build_sequential_index(data: bytes, h: BgenHeader) -> List[BgenIndexEntry]:
    """bgen.rs:236-310: walk the variant blocks from 4 + offset."""
    out = []
    pos = 4 + h.offset
    for _ in range(h.n_variants):
        if pos >= len(data):
            break
        o = pos
        k = _u16(data, o)
        o += 2 + k                                   # variant id
        k = _u16(data, o)
        rsid = _text(data[o + 2:o + 2 + k])
        o += 2 + k
        k = _u16(data, o)
        chrom = _text(data[o + 2:o + 2 + k])
        o += 2 + k
        vpos = _u32(data, o)
        n_alleles = _u16(data, o + 4)
        o += 6
        alleles = []
        for _ in range(n_alleles):
            k = _u32(data, o)
            alleles.append(_text(data[o + 4:o + 4 + k]))
            o += 4 + k
        c = _u32(data, o)
        total = (o + 4 - pos) + c
        out.append(BgenIndexEntry(pos, total, chrom, vpos, rsid,
                                  alleles[0] if alleles else "",
                                  alleles[1] if len(alleles) > 1 else ""))
        pos += total
    return out


class BgenReader:
    """BgenReader::new (bgen.rs:67-112) and the GenotypeReader methods used by step 2."""

    def __init__(self, path):
        path = os.fspath(path)
        with open(path, "rb") as f:
            self.data = f.read()
        self.header = parse_header(self.data)
        if self.header.has_sample_ids:
            self._ids = parse_sample_ids(self.data, self.header)
        else:
            stem = os.path.splitext(path)[0]
            cands = [stem + ".sample", path + ".sample", path + ".samples"]
            found = next((c for c in cands if os.path.exists(c)), None)
            self._ids = (parse_sample_file(found) if found else
                         [f"sample_{i}" for i in range(self.header.n_samples)])
        bgi = path + ".bgi"
        self.index = read_bgi_index(bgi) if os.path.exists(bgi) else \
            build_sequential_index(self.data, self.header)
        self.sample_ids = list(self._ids)
        self.sample_subset: Optional[np.ndarray] = None

    def n_markers(self):
        return len(self.index)

    def n_samples(self):
        return len(self.sample_subset) if self.sample_subset is not None else \
            self.header.n_samples

    def set_sample_subset(self, ids):
        """bgen.rs:466-480: keep the given IDs that the file has, in the given order."""
        pos = {}
        for k, s in enumerate(self._ids):
            pos.setdefault(s, k)  # Iterator::position: the first match
        idx, new_ids = [], []
        for s in ids:
            if s in pos:
                idx.append(pos[s])
                new_ids.append(s)
        self.sample_subset = np.asarray(idx, dtype=np.uint32)
        self.sample_ids = new_ids

    def marker_info(self, i):
        e = self.index[i]
        return dict(chrom=e.chrom, pos=e.pos, id=e.rsid, ref_allele=e.allele1,
                    alt_allele=e.allele2)

    def layout2_block(self, i) -> bytes:
        """The decompressed genotype data of variant i (bgen.rs:313-368)."""
        if i >= len(self.index):
            raise IndexError(f"Variant index {i} out of range ({len(self.index)})")
        d = self.data
        o = self.index[i].file_offset
        o += 2 + _u16(d, o)
        o += 2 + _u16(d, o)
        o += 2 + _u16(d, o)
        n_alleles = _u16(d, o + 4)
        o += 6
        for _ in range(n_alleles):
            o += 4 + _u32(d, o)
        c = _u32(d, o)
        o += 4
        comp = self.header.compression
        if comp == "none":
            return bytes(d[o:o + c])
        body = bytes(d[o + 4:o + c])  # the first 4 bytes of C are the decompressed length
        if comp == "zlib":
            return zlib.decompress(body)
        try:
            import zstandard
        except ImportError as e:  # not in this image; the reference links zstd
            raise RuntimeError("BGEN zstd blocks need the zstandard module") from e
        return zstandard.ZstdDecompressor().decompressobj().decompress(body)

    def batch(self, i0, i1) -> BgenBatch:
        """Blocks of variants [i0, i1) for the device, with their decode info."""
        raw = [self.layout2_block(i) for i in range(i0, i1)]
        n_file = _u32(raw[0], 0) if raw else self.header.n_samples
        info = np.zeros(len(raw), np.uint32)
        need = 10 + n_file
        for v, b in enumerate(raw):
            if len(b) < 10 + n_file or _u32(b, 0) != n_file:
                raise ValueError(f"variant {i0 + v}: layout-2 block does not hold {n_file} samples")
            n_alleles = _u16(b, 4)
            lo, hi = b[6], b[7]
            phased = b[8 + n_file] != 0
            bits = b[9 + n_file]
            if bits == 0:
                raise ValueError("BGEN: 0 bits per probability not supported")
            standard = n_alleles == 2 and lo == 2 and hi == 2 and not phased
            info[v] = bits | (0x100 if standard else 0)
            if standard:
                need = max(need, 10 + n_file + (2 * n_file * bits + 7) // 8)
        ld = max([need] + [len(b) for b in raw])
        ld = (ld + 15) // 16 * 16
        blocks = np.zeros((len(raw), ld), np.uint8)
        for v, b in enumerate(raw):
            blocks[v, :len(b)] = np.frombuffer(b, np.uint8)
        return BgenBatch(blocks, info, ld, n_file)
