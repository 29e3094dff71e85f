"""Host side of the region tests (saige-core/src/score_test/region.rs and the region
driver of saige-cli/src/commands/assoc_test.rs:227-330): variant weights, MAC categories,
the group file and SAIGE's region output lines.  The tests themselves run in
``ScoreTestEngine.test_regions`` (device Gram, host Satterthwaite p-values)."""
from __future__ import annotations

import math
from typing import Dict, List, Sequence

import numpy as np

from .model import rust_f64

MAC_CATEGORIES = [1.5, 2.5, 5.5, 10.5, 20.5, 40.5, math.inf]  # region.rs:30-36


def default_rho_values() -> List[float]:
    """region.rs:368-370."""
    return [0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0]


def beta_weights(mafs, a1=1.0, a2=25.0) -> np.ndarray:
    """Beta(MAF; a1, a2) density (region.rs:374-386), MAF clamped to [1e-10, 1 - 1e-10];
    ln B(a1, a2) = ln G(a1) + ln G(a2) - ln G(a1 + a2) (statrs' ln_beta is third-party:
    equal to the last few bits)."""
    ln_b = math.lgamma(a1) + math.lgamma(a2) - math.lgamma(a1 + a2)
    out = []
    for maf in np.asarray(mafs, dtype=np.float64):
        m = min(max(float(maf), 1e-10), 1.0 - 1e-10)
        out.append(math.exp((a1 - 1.0) * math.log(m) + (a2 - 1.0) * math.log(1.0 - m) - ln_b))
    return np.array(out)


def mac_category_counts(mac_values, boundaries=MAC_CATEGORIES) -> List[int]:
    """region.rs:342-365: the first boundary b with MAC <= b; 8 categories."""
    n_cats = len(boundaries) + 1
    counts = [0] * min(n_cats, 8)
    for mac in mac_values:
        cat = next((i for i, b in enumerate(boundaries) if mac <= b), len(boundaries))
        if cat < len(counts):
            counts[cat] += 1
    return counts + [0] * (8 - len(counts))


REGION_HEADER = ("Gene Pvalue Nmarker_MACCate_1 Nmarker_MACCate_2 Nmarker_MACCate_3 "
                 "Nmarker_MACCate_4 Nmarker_MACCate_5 Nmarker_MACCate_6 Nmarker_MACCate_7 "
                 "Nmarker_MACCate_8 markerIDs markerAFs Pvalue_Burden Pvalue_SKAT Pvalue.NA "
                 "Pvalue_Burden.NA Pvalue_SKAT.NA")


def write_region_header(f) -> None:
    """write_region_header (region.rs:389-395)."""
    f.write(REGION_HEADER + "\n")


def write_region_line(f, name, pvalue, mac_counts, marker_ids, marker_afs, burden_pvalue,
                      skat_pvalue) -> None:
    """write_region_line (region.rs:398-428); the .NA columns repeat the p-values, as the
    driver sets them (assoc_test.rs:316-320)."""
    cats = list(mac_counts) + [0] * 8
    fields = [str(name), rust_f64(pvalue)] + [str(int(c)) for c in cats[:8]] + [
        ";".join(marker_ids), ";".join(rust_f64(a) for a in marker_afs),
        rust_f64(burden_pvalue), rust_f64(skat_pvalue), rust_f64(pvalue),
        rust_f64(burden_pvalue), rust_f64(skat_pvalue)]
    f.write(" ".join(fields) + "\n")


def run_regions(engine, regions: Dict[str, Sequence[int]], G, af, mac, marker_ids,
                min_mac=0.5, min_maf=0.0):
    """The per-gene loop of run_region_tests (assoc_test.rs:257-321) over dosage rows
    ``G`` with their AF and MAC: the MAC / MAF filter, Beta(MAF; 1, 25) weights, MAC
    categories, and one batched device call for all genes.  ``regions`` maps a gene name
    to row indices of G (genes with no row left are skipped).  Returns a list of dicts in
    gene order."""
    names, rows = [], []
    for gene, idx in regions.items():
        keep = [i for i in idx if not (mac[i] < min_mac) and not (af[i] < min_maf or
                                                                   af[i] > 1.0 - min_maf)]
        if keep:
            names.append(gene)
            rows.append(keep)
    if not names:
        return []
    order = [i for r in rows for i in r]
    Gr = np.asarray(G, dtype=np.float64)[order]
    afs = np.asarray(af, dtype=np.float64)[order]
    w = beta_weights(afs, 1.0, 25.0)
    res = engine.test_regions(Gr, [len(r) for r in rows], w)
    out = []
    for k, (gene, r) in enumerate(zip(names, rows)):
        out.append(dict(gene=gene, pvalue=float(res["pvalue"][k]),
                        burden_pvalue=float(res["burden_pvalue"][k]),
                        skat_pvalue=float(res["skat_pvalue"][k]), rho=float(res["rho"][k]),
                        mac_category_counts=mac_category_counts([mac[i] for i in r]),
                        marker_ids=[marker_ids[i] for i in r],
                        marker_afs=[float(af[i]) for i in r]))
    return out
