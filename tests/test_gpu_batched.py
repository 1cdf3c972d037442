"""Batched multi-RHS path: the BF16 matrix-core K.u against the FP64 oracle,
and the solve of several load cases (interleaved 16-case iteration, and the
default sequential mixed-precision solves) against the oracle's FP64 solves
of each case."""
import numpy as np
import pytest

from oracle import oracle as orc
from sgldbench_amd import problems, solver

from tests.test_gpu_solver import holes

pytestmark = pytest.mark.gpu


def grf_E(vox, seed):
    x = problems.gaussian_random_density(vox.shape, seed, sigma=2.0).ravel(order="F")
    return problems.simp_modulus(x[vox.ravel(order="F") != 0])


@pytest.mark.parametrize("dims,nrhs", [((32, 16, 16), 16), ((38, 22, 18), 5), ((20, 12, 40), 16)])
def test_batched_kbyu_matches_oracle(dims, nrhs):
    import torch
    nx, ny, nz = dims
    print("case: batched kbyu", dims, nrhs, flush=True)
    vox = holes(nx, ny, nz, 3, nholes=3, r=3)
    _, fixed = problems.cantilever3(vox)
    E = grf_E(vox, 4)
    g = solver.FEASolver(vox, 2)
    g.set_fixed(fixed)
    g.set_modulus(E)
    m = orc.Mesh(vox, 2)
    m.set_fixed(fixed)
    m.set_modulus(E)
    rng = np.random.default_rng(7)
    U = rng.standard_normal((nrhs, g.numDOFs))
    dU = torch.from_numpy(U).cuda()
    dY = torch.zeros_like(dU)
    g.kbyu_batched_device(nrhs, dU.data_ptr(), dY.data_ptr())
    torch.cuda.synchronize()
    Y = dY.cpu().numpy()
    for j in range(nrhs):
        ref = m.kbyu(U[j])
        # BF16 hi/lo operands (16 significant bits) with FP32 accumulation
        assert np.abs(Y[j] - ref).max() <= 1e-4 * np.abs(ref).max(), j
        assert np.all(Y[j][fixed != 0] == 0)


@pytest.mark.parametrize("prec", [solver.PREC_INTERLEAVED, solver.PREC_MIXED])
@pytest.mark.parametrize("case", ["cuboid32x16x16", "holes32x16x16", "nondyadic32x16x16"])
def test_batched_solve_matches_oracle(case, prec):
    print("case: batched solve", case, "precision", prec, flush=True)
    if case == "holes32x16x16":
        vox, levels, nd = holes(32, 16, 16, 7), 3, 0
    else:
        vox, levels, nd = problems.cuboid(32, 16, 16), 4, int(case.startswith("nondyadic"))
    F, fixed, _ = problems.face_load_cases(vox, 6, seed=11)
    F[3] = 0.0  # a zero load case returns at once with U = 0
    E = grf_E(vox, 5)
    g = solver.FEASolver(vox, levels, nd)
    g.set_fixed(fixed)
    g.set_modulus(E)
    g.assemble(0.6)
    m = orc.Mesh(vox, levels, nd)
    m.set_fixed(fixed)
    m.set_modulus(E)
    m.assemble()
    U, its, rr = g.pcg_batched(F, tol=1e-8, maxIT=1000, precision=prec)
    assert g.converged
    assert its[3] == 0 and np.all(U[3] == 0)
    for j in range(F.shape[0]):
        if j == 3:
            continue
        ref, its_o, hist, rc = m.pcg(F[j], 0.6, 1e-8, 1000)
        assert rc == 0
        true = np.linalg.norm(F[j] - m.kbyu(U[j])) / np.linalg.norm(F[j])
        assert true <= 1.01e-8 and rr[j] <= 1e-8, (j, true, rr[j])
        c, c_ref = F[j] @ U[j], F[j] @ ref
        assert abs(c - c_ref) <= 1e-8 * abs(c_ref), j
        assert np.linalg.norm(U[j] - ref) <= 1e-5 * np.linalg.norm(ref), j
        assert its[j] <= 1.5 * its_o + 10, (j, its[j], its_o)


@pytest.mark.parametrize("prec", [solver.PREC_FP64, solver.PREC_MIXED])
def test_batched_sequential_modes_equal_single_solves(prec):
    vox = problems.cuboid(32, 16, 16)
    F, fixed, _ = problems.face_load_cases(vox, 3, seed=2)
    E = grf_E(vox, 6)
    g = solver.FEASolver(vox, 4)
    g.set_fixed(fixed)
    g.set_modulus(E)
    g.assemble(0.6)
    U, its, rr = g.pcg_batched(F, tol=1e-8, maxIT=1000, precision=prec)
    assert g.converged
    for j in range(3):
        u1, it1, r1 = g.pcg(F[j], tol=1e-8, maxIT=1000, precision=prec)
        # K.u sums LDS face forces in run-dependent order: equal up to rounding
        # (FP64), or to the FP32 inner solve's rounding (mixed)
        assert abs(int(its[j]) - it1) <= (1 if prec == solver.PREC_FP64 else 3) and rr[j] <= 1e-8
        assert np.linalg.norm(U[j] - u1) <= (1e-7 if prec == solver.PREC_FP64 else 1e-6) * np.linalg.norm(u1)


@pytest.mark.parametrize("prec", [solver.PREC_INTERLEAVED, solver.PREC_MIXED])
def test_batched_warm_start_and_maxit(prec):
    vox = problems.cuboid(32, 16, 16)
    F, fixed, _ = problems.face_load_cases(vox, 4, seed=3)
    g = solver.FEASolver(vox, 4)
    g.set_fixed(fixed)
    g.set_modulus(np.ones(int(vox.sum())))
    g.assemble(0.6)
    U, its, rr = g.pcg_batched(F, tol=1e-8, maxIT=1000, precision=prec)
    assert g.converged
    U2, its2, rr2 = g.pcg_batched(F, tol=1e-8, maxIT=1000, U0=U, precision=prec)
    assert np.all(its2 == 0) and np.all(rr2 <= 1e-8)
    U3, its3, rr3 = g.pcg_batched(F, tol=1e-12, maxIT=5, precision=prec)
    assert not g.converged and np.all(its3 <= 5)


def test_batched_rejects_unknown_precision():
    vox = problems.cuboid(16, 8, 8)
    F, fixed, _ = problems.face_load_cases(vox, 2, seed=4)
    g = solver.FEASolver(vox, 3)
    g.set_fixed(fixed)
    g.set_modulus(np.ones(int(vox.sum())))
    g.assemble(0.6)
    with pytest.raises(solver.SGLDError):
        g.pcg_batched(F, tol=1e-8, maxIT=100, precision=7)
    with pytest.raises(solver.SGLDError):  # the interleaved mode is a batched-solve mode only
        g.pcg(F[0], tol=1e-8, maxIT=100, precision=solver.PREC_INTERLEAVED)
