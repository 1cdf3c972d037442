"""sgld_restrict_noalloc (device arrays, ghost array in the library's fixed
scratch) against sgld_restrict (host arrays, per-call allocation): bit-identical
outputs on two calls in a row, the second on a smaller mesh (fewer ghost
nodes: the scratch keeps the first call's values and must be cleared), plus
the limit and the id checks."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from sgldbench_amd import solver
from sgldbench_amd._lib import lib

pytestmark = pytest.mark.gpu


def voxels(shape, fill, seed):
    rng = np.random.default_rng(seed)
    v = (rng.random(shape) < fill).astype(np.uint8)
    v.ravel()[0] = 1
    return v


def mex_args(mesh, ii, seed):
    F, Cc = mesh.info[ii - 2], mesh.info[ii - 1]
    rng = np.random.default_rng(seed)
    rF = rng.standard_normal((F["numNodes"], 3))
    RIT = np.ascontiguousarray(mesh.get(ii, "RI").T)
    return (rF, mesh.get(ii, "transferMat"), RIT, mesh.get(ii, "eNodMat"), Cc["numNodes"],
            Cc["intermediateNumNodes"], mesh.get(ii, "solidNodeMapCoarser2Finer"),
            mesh.get(ii, "transferMatCoeffi"), mesh.colors(ii))


def noalloc(args, nghost_override=None, bad_id=False):
    import torch
    rF, tm, RIT, eN, nC, nG, mp, co, cols = args
    dev = torch.device("cuda:0")
    f64 = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))).to(dev)
    i32 = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel(order="F"))).to(dev)
    drF, dRIT, dco = f64(rF), f64(RIT), f64(co)
    tmf = np.asarray(tm, dtype=np.int32).ravel(order="F").copy()
    if bad_id:
        tmf[len(tmf) // 2] = int(nG) + 1
    dtm, deN, dmp = i32(tmf), i32(eN), i32(mp)
    off = np.zeros(len(cols) + 1, dtype=np.int64)
    for i, c in enumerate(cols):
        off[i + 1] = off[i] + len(c)
    dcol = i32(np.concatenate([np.asarray(c, dtype=np.int32) for c in cols]))
    dY = torch.full((3 * int(nC),), float("nan"), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = lib().sgld_restrict_noalloc(drF.data_ptr(), rF.shape[0], dtm.data_ptr(), tm.shape[0], dRIT.data_ptr(),
                                     8, tm.shape[0], deN.data_ptr(), eN.shape[0], int(nC),
                                     int(nG if nghost_override is None else nghost_override), dmp.data_ptr(),
                                     dco.data_ptr(), dcol.data_ptr(), off, len(cols), dY.data_ptr())
    return rc, dY.cpu().numpy().reshape(int(nC), 3, order="F")


def test_noalloc_restrict_equals_allocating_restrict_twice():
    big = orc.Mesh(voxels((16, 16, 16), 0.6, 21), 3)
    small = orc.Mesh(voxels((8, 8, 8), 0.5, 22), 2)
    for k, (mesh, ii) in enumerate([(big, 2), (small, 2), (big, 3)]):
        a = mex_args(mesh, ii, 40 + k)
        print("case: restrict_noalloc call", k, "nGhost", a[5], flush=True)
        ref = solver.Solving_Restriction_MatrixFree_mex(*a)
        rc, Y = noalloc(a)
        assert rc == 0, lib().sgld_last_error().decode()
        assert np.array_equal(Y, ref), k


def test_noalloc_restrict_rejects_oversize_and_bad_ids():
    mesh = orc.Mesh(voxels((8, 8, 8), 0.5, 23), 2)
    a = mex_args(mesh, 2, 50)
    print("case: restrict_noalloc limits", flush=True)
    rc, _ = noalloc(a, nghost_override=1_398_091)
    assert rc == 1 and "32 MiB" in lib().sgld_last_error().decode()
    rc, _ = noalloc(a, bad_id=True)
    assert rc == 1 and "transferMat" in lib().sgld_last_error().decode()
    rc, Y = noalloc(a)  # usable again after a rejected call
    assert rc == 0 and np.array_equal(Y, solver.Solving_Restriction_MatrixFree_mex(*a))
