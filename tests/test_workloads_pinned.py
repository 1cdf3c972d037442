"""The generated benchmark workloads are pinned by the SHA-256 (first 16 hex
digits) of their arrays: a change in the generators, numpy's seeded
generator or scipy's gaussian_filter would silently change every figure
measured on them.  The same digests are recorded with the measurements."""
import hashlib

import numpy as np
import pytest

from sgldbench_amd import problems

DIGESTS = {
    "cfg0": {"voxels": "916b144867c34061", "E": "9d37a9102309aaf8", "F": "0dafbba5272e8360", "fixed": "5f9377b64667dc50"},
    "cfg1": {"voxels": "5d2bafc266e711ed", "E": "aa3b64ca0a57a0ac", "F": "a1a8de5c7086036b", "fixed": "888c5532daaee70c"},
    "cfg2": {"voxels": "becdc3835120ff88", "E": "a816bc9b30a932b9", "F": "5e66028b7b93365f", "fixed": "a15fd29a142e64b6"},
    "cfg3": {"voxels": "b70a752bfdf8d344", "E": "2e60a4a17c37223f", "F": "e25d0c23f43c28dd", "fixed": "da7e9b8e68bef71c"},
}
CFG4_ALL = "a2a5a8f7269bfebc"


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(np.asarray(a)).tobytes()).hexdigest()[:16]


@pytest.mark.parametrize("name", list(DIGESTS))
def test_workload_arrays_are_pinned(name):
    p = getattr(problems, name)()
    assert {k: digest(p[k]) for k in DIGESTS[name]} == DIGESTS[name]


def test_uneven_batch_is_pinned():
    h = hashlib.sha256()
    for d in problems.cfg4():
        for k in ("voxels", "x", "E", "F", "fixed"):
            h.update(np.ascontiguousarray(np.asarray(d[k])).tobytes())
    assert h.hexdigest()[:16] == CFG4_ALL
