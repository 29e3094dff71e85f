"""Writes the committed golden fixtures from data files the reference ships (its example
inputs and R SAIGE outputs under extdata/) and from OpenSSL.  Run in the build container
(it reads /root/reference); the fixtures it writes are plain data.

  frq_subset.npz  400 evenly spaced markers of extdata/input/plinkforGRM_1000samples_10kMarkers.bed
                  with PLINK 1.9 MAF and NCHROBS from the .frq (A1 = bim allele 1 for every
                  marker, so AC = MAF * NCHROBS): pins decode, AC, MAC.
  ac100.npz       the 32 markers of extdata/output/genotype_100markers_marker_plink.txt
                  (R SAIGE AC_Allele2, Allele2 = bim allele 1 = the saige-rs dosage allele)
                  with the .bed rows of extdata/input/genotype_100markers.bed and the model
                  sample order of extdata/input/pheno_1000samples.txt_withdosages_withBothTraitTypes.txt
  chacha20.npz    OpenSSL chacha20 keystream (20 rounds) for two keys, 4 blocks each.
"""
import os
import subprocess

import numpy as np

REF = "/root/reference/extdata"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_bim(path):
    return [line.split() for line in open(path) if line.strip()]


def frq_subset():
    base = f"{REF}/input/plinkforGRM_1000samples_10kMarkers"
    bim = read_bim(base + ".bim")
    fam = [line.split() for line in open(base + ".fam") if line.strip()]
    n = len(fam)
    bpm = (n + 3) // 4
    bed = np.fromfile(base + ".bed", dtype=np.uint8)
    assert bed[:3].tolist() == [0x6C, 0x1B, 0x01]
    body = bed[3:].reshape(len(bim), bpm)
    rows = [line.split() for line in open(base + ".frq")][1:]
    a1 = np.array([r[2] for r in rows])
    maf = np.array([float(r[4]) for r in rows])
    nchrobs = np.array([int(r[5]) for r in rows])
    allele1 = np.array([b[4] for b in bim])
    assert (a1 == allele1).all()  # .frq A1 is the bim allele 1 (the saige-rs dosage allele)
    pick = np.linspace(0, len(bim) - 1, 400).astype(int)
    np.savez_compressed(os.path.join(OUT, "frq_subset.npz"), bed=body[pick], marker=pick,
                        maf=maf[pick], nchrobs=nchrobs[pick], n=n)


def ac100():
    base = f"{REF}/input/genotype_100markers"
    bim = read_bim(base + ".bim")
    fam = [line.split() for line in open(base + ".fam") if line.strip()]
    iid_pos = {f[1]: i for i, f in enumerate(fam)}
    n_total = len(fam)
    bpm = (n_total + 3) // 4
    bed = np.fromfile(base + ".bed", dtype=np.uint8)
    assert bed[:3].tolist() == [0x6C, 0x1B, 0x01]
    body = bed[3:].reshape(len(bim), bpm)
    ph = [line.split() for line in open(f"{REF}/input/pheno_1000samples.txt_withdosages_withBothTraitTypes.txt")]
    head = ph[0]
    iid_col = head.index("IID")
    ids = [r[iid_col] for r in ph[1:]]
    sample_idx = np.array([iid_pos[i] for i in ids], dtype=np.uint32)
    res = [line.split("\t") for line in open(f"{REF}/output/genotype_100markers_marker_plink.txt")]
    h = res[0]
    mid = h.index("MarkerID")
    acc = h.index("AC_Allele2")
    a2 = h.index("Allele2")
    name_to_row = {b[1]: j for j, b in enumerate(bim)}
    markers, ac = [], []
    for r in res[1:]:
        j = name_to_row[r[mid]]
        assert bim[j][4] == r[a2], (bim[j], r[a2])
        markers.append(j)
        ac.append(float(r[acc]))
    markers = np.array(markers)
    np.savez_compressed(os.path.join(OUT, "ac100.npz"), bed=body[markers], marker=markers,
                        ac=np.array(ac), sample_idx=sample_idx, n_total=n_total)


def chacha():
    keys = [bytes(32), bytes(range(32))]
    blocks = []
    for k in keys:
        ks = subprocess.run(["openssl", "enc", "-chacha20", "-K", k.hex(), "-iv", "00" * 16,
                             "-nosalt"], input=b"\0" * 256, capture_output=True, check=True).stdout
        blocks.append(np.frombuffer(ks, dtype="<u4").reshape(4, 16))
    np.savez_compressed(os.path.join(OUT, "chacha20.npz"),
                        keys=np.stack([np.frombuffer(k, dtype="<u4") for k in keys]),
                        blocks=np.stack(blocks))


if __name__ == "__main__":
    frq_subset()
    ac100()
    chacha()
    print("written to", OUT)
