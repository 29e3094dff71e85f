"""The seeded cohorts are the same bit for bit whatever the thread count (CPU only)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_SCRIPT = r"""
import sys, hashlib
sys.path.insert(0, %r)
from csaige_amd import synth
u = synth.unrelated(n=700, m_grm=900, m_test=5, seed=1)
f = synth.families(n_fam=150, m=800, seed=2)
b = synth.biobank(n=1500, m=700, prevalences=(0.01, 0.2), seed=5, h2=0.5, n_causal=300)
h = hashlib.sha256()
for a in (u.liability, u.y, u.bed, f.y, f.bed, b.y, b.bed):
    h.update(a.tobytes())
print(h.hexdigest())
""" % ROOT


def _digest(threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", _SCRIPT], env=env, capture_output=True,
                         text=True, check=True)
    return out.stdout.strip()


def test_cohorts_independent_of_thread_count():
    """Polygenic scores add in fixed point, so the liabilities (not only the case labels) are
    identical on one thread and on several."""
    assert _digest(1) == _digest(3) == _digest(8)


@pytest.mark.gpu
def test_device_dosage_generator_matches_host():
    """The device twin of the step-2 dosage generator writes the host generator's bytes
    (hard calls, the noisy every-fourth variants, zero padding past n), for a batch that does
    not start at variant 0."""
    import numpy as np

    from csaige_amd import hip, synth
    n, ld, first, nvar = 1001, 1004, 5, 37
    want, mafs = synth.dosage_u8(n, first, nvar, seed=4, ld=ld)
    np.testing.assert_array_equal(mafs, synth.dosage_mafs(first, nvar, seed=4))
    d_m = hip.DeviceBuffer(mafs.nbytes)
    d_m.upload(mafs)
    d_q = hip.DeviceBuffer(nvar * ld)
    synth.dosage_u8_device(n, first, nvar, d_m.value, d_q.value, ld=ld, seed=4)
    hip.device_synchronize()
    got = np.empty((nvar, ld), np.uint8)
    d_q.download(got)
    assert (want[(3 - first) % 4::4] % 127 != 0).any()  # the noisy variants are there
    np.testing.assert_array_equal(got, want)


def test_config_fixture_cohorts_regenerate():
    """The committed oracle fixtures of the families and biobank configurations were made on
    the cohorts the generator makes today (same hash of genotypes, phenotypes, covariates)."""
    import numpy as np

    import tests.golden.make_config_vectors as mk
    gold = os.path.join(ROOT, "tests", "golden")
    fam = np.load(os.path.join(gold, "config_families.npz"))
    bb = np.load(os.path.join(gold, "config_biobank.npz"))
    assert mk.cohort_hash(mk.families()) == str(fam["cohort_sha256"])
    assert mk.cohort_hash(mk.biobank()) == str(bb["cohort_sha256"])
