"""Writes the BGEN fixtures under tests/golden/bgen/ from data files the reference ships
(extdata/input).  Run in the build container (it reads /root/reference); what it writes is
plain data.

  genotype_100markers.bgen (+ .bgen.samples, .bgen.bgi)   10,000 samples x 100 markers, zlib,
                                                          8-bit probabilities, no missing
  genotype_10markers.missingness.bgen                     1,000 samples x 10 markers with
                                                          missing calls
  genotype_100markers_2chr.bgen                           sample IDs in the file
  <name>.expected.npz   from the .vcf of the same markers: per marker the ALT allele count of
                        every sample's GT (0/0 -> 0, 0/1 -> 1, 1/1 -> 2, ./. -> -1), the
                        sample names and the marker IDs.  ALT is the BGEN file's second
                        allele, whose dosage the reference decodes (bgen.rs:403-420).
"""
import os
import shutil

import numpy as np

REF = "/root/reference/extdata/input"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bgen")

FILES = [
    ("genotype_100markers", ["genotype_100markers.bgen", "genotype_100markers.bgen.samples",
                             "genotype_100markers.bgen.bgi"]),
    ("genotype_10markers.missingness", ["genotype_10markers.missingness.bgen"]),
    ("genotype_100markers_2chr", ["genotype_100markers_2chr.bgen"]),
]


def vcf_alt_counts(path):
    samples, ids, rows = None, [], []
    for line in open(path):
        if line.startswith("##"):
            continue
        f = line.rstrip("\n").split("\t")
        if line.startswith("#"):
            samples = f[9:]
            continue
        ids.append(f[2])
        row = []
        for g in f[9:]:
            gt = g.split(":")[0].replace("|", "/")
            a = gt.split("/")
            row.append(-1 if "." in a else sum(int(x) != 0 for x in a))
        rows.append(row)
    return np.array(rows, dtype=np.int8), np.array(samples), np.array(ids)


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, files in FILES:
        for fn in files:
            shutil.copyfile(os.path.join(REF, fn), os.path.join(OUT, fn))
        alt, samples, ids = vcf_alt_counts(os.path.join(REF, name + ".vcf"))
        np.savez_compressed(os.path.join(OUT, name + ".expected.npz"), alt=alt, samples=samples,
                            ids=ids)
        print(name, alt.shape, int((alt < 0).sum()), "missing")


if __name__ == "__main__":
    main()
