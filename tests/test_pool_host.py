"""Host logic of ScoreTestPool without a GPU: stand-in engines record what they are given.
The first batch of each new size runs on engine 0 before the others start, the tuned shape
is then copied from engine 0 to every other engine, the remaining batches are dealt
round-robin (each exactly once), and an error on a worker thread reaches the caller."""
import threading

import pytest

import csaige_amd as sg


class _Engine:
    def __init__(self, j, log, fail_on=None):
        self.j, self.log, self.fail_on = j, log, fail_on
        self.copied_from = []

    def copy_tuning_from(self, other):
        self.log.append(("copy", self.j, other.j))
        self.copied_from.append(other.j)

    def test_markers_u8_device(self, d_q, ldq, nvar, scale, d_out):
        self.log.append(("run", self.j, d_q, nvar, threading.get_ident()))
        if self.fail_on == d_q:
            raise sg.SaigeError(1, "stand-in failure")


def _pool(k, fail_on=None):
    p = object.__new__(sg.ScoreTestPool)
    log = []
    p.engines = [_Engine(j, log, fail_on) for j in range(k)]
    p.n = 100
    p._sizes = set()
    return p, log


def test_first_batch_of_each_size_tunes_engine_zero_then_round_robin():
    p, log = _pool(3)
    batches = [(1000 + b, 100, 512 if b != 5 else 77, 1.0 / 127, 5000 + b) for b in range(9)]
    p.test_markers_u8_device(batches)
    runs = [e for e in log if e[0] == "run"]
    assert sorted(r[2] for r in runs) == [b[0] for b in batches]  # each batch once
    # the first 512-variant batch and the 77-variant batch run first, on engine 0
    assert [(r[1], r[2]) for r in runs[:2]] == [(0, 1000), (0, 1005)]
    first_copy = next(i for i, e in enumerate(log) if e[0] == "copy")
    assert first_copy == 2 and sorted(e[1] for e in log if e[0] == "copy") == [1, 2]
    assert all(e.copied_from == [0] for e in p.engines[1:])
    rest = [b[0] for b in batches if b[0] not in (1000, 1005)]
    for j in range(3):
        assert [r[2] for r in runs[2:] if r[1] == j] == rest[j::3]
    # a second call with known sizes copies nothing and tunes nothing first
    log.clear()
    p.test_markers_u8_device(batches[:4])
    assert not [e for e in log if e[0] == "copy"]
    assert sorted(e[2] for e in log) == [1000, 1001, 1002, 1003]


def test_engines_run_on_their_own_threads():
    p, log = _pool(2)
    p._sizes.add(512)
    p.test_markers_u8_device([(b, 100, 512, 1.0, b) for b in range(6)])
    tid = {}
    for e in log:
        tid.setdefault(e[1], set()).add(e[4])
    assert len(tid[0]) == 1 and len(tid[1]) == 1 and tid[0] != tid[1]


def test_worker_error_reaches_the_caller():
    p, _ = _pool(2, fail_on=3)
    p._sizes.add(512)
    with pytest.raises(sg.SaigeError, match="stand-in failure"):
        p.test_markers_u8_device([(b, 100, 512, 1.0, b) for b in range(6)])


def test_pool_needs_an_engine():
    with pytest.raises(ValueError):
        sg.ScoreTestPool(None, None, (1.0, 0.0), None, None, 1.0, engines=0)
