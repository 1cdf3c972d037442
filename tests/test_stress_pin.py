"""Pins the oracle's stress recovery (FEA_StressAnalysis.m) with the patch
test (a linear displacement field has the closed-form stress D eps at every
node) and an independent numpy restatement of the reference's operators."""
import numpy as np

from oracle import oracle as orc
from sgldbench_amd import problems

from tests.test_gpu_solver import holes


def D_matrix(nu):
    lam, mu = nu / ((1 + nu) * (1 - 2 * nu)), 1 / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[:3, :3] += 2 * mu * np.eye(3)
    D[3:, 3:] = mu * np.eye(3)
    return D


def test_patch_linear_field_gives_closed_form_stress():
    vox = problems.cuboid(4, 2, 2)
    m = orc.Mesh(vox, 2, nu=0.25)
    nodes = problems.solid_nodes(vox)
    xyz = problems.node_coords(vox, nodes).astype(float)
    A = np.array([[1e-3, 2e-3, 0.0], [0.0, -1e-3, 5e-4], [3e-4, 0.0, 2e-3]])
    E = np.full(int(vox.sum()), 2.0)
    m.set_modulus(E)
    st, vm = m.stress((xyz @ A.T).ravel())
    eps = np.array([A[0, 0], A[1, 1], A[2, 2], A[1, 2] + A[2, 1], A[2, 0] + A[0, 2], A[0, 1] + A[1, 0]])
    sig = 2.0 * D_matrix(0.25) @ eps
    np.testing.assert_allclose(st, np.tile(sig, (len(nodes), 1)), rtol=1e-12, atol=1e-15)
    vm_ref = np.sqrt(0.5 * ((sig[0] - sig[1]) ** 2 + (sig[1] - sig[2]) ** 2 + (sig[2] - sig[0]) ** 2)
                     + 3 * (sig[3] ** 2 + sig[4] ** 2 + sig[5] ** 2))
    np.testing.assert_allclose(vm, vm_ref, rtol=1e-12)


def test_random_field_matches_numpy_restatement():
    vox = holes(8, 6, 4, 2, nholes=2, r=2)
    ny, nx, nz = vox.shape
    m = orc.Mesh(vox, 2)
    ne = int(vox.sum())
    E = np.random.default_rng(4).uniform(0.1, 1, ne)
    m.set_modulus(E)
    nodes = problems.solid_nodes(vox)
    U = np.random.default_rng(5).standard_normal(3 * len(nodes))
    st, vm = m.stress(U)
    # independent restatement: corner signs, Gauss points, shape functions
    cs = np.array([-1, 1, 1, -1, -1, 1, 1, -1.0])
    ct = np.array([-1, -1, 1, 1, -1, -1, 1, 1.0])
    cp = np.array([-1, -1, -1, -1, 1, 1, 1, 1.0])
    g = 1 / np.sqrt(3)
    gp = np.stack([cs * g, ct * g, cp * g], axis=1)
    N = np.array([[0.125 * (1 + cs[j] * s) * (1 + ct[j] * t) * (1 + cp[j] * p) for j in range(8)] for s, t, p in gp])
    Ninv = np.linalg.inv(N)
    D = D_matrix(0.3)
    Bs = []
    for s, t, p in gp:
        dN = np.array([[0.125 * cs[j] * (1 + ct[j] * t) * (1 + cp[j] * p),
                        0.125 * ct[j] * (1 + cs[j] * s) * (1 + cp[j] * p),
                        0.125 * cp[j] * (1 + cs[j] * s) * (1 + ct[j] * t)] for j in range(8)]) * 2.0
        B = np.zeros((6, 24))
        for j in range(8):
            dx, dy, dz = dN[j]
            B[0, 3 * j], B[1, 3 * j + 1], B[2, 3 * j + 2] = dx, dy, dz
            B[3, 3 * j + 1], B[3, 3 * j + 2] = dz, dy
            B[4, 3 * j], B[4, 3 * j + 2] = dz, dx
            B[5, 3 * j], B[5, 3 * j + 1] = dy, dx
        Bs.append(D @ B)
    eN = m.get(1, "eNodMat") - 1
    acc = np.zeros((len(nodes), 6))
    cnt = np.zeros(len(nodes))
    for e in range(ne):
        ue = U.reshape(-1, 3)[eN[e]].ravel()
        sg = np.array([E[e] * (DB @ ue) for DB in Bs])     # 8 x 6
        sn = Ninv @ sg
        np.add.at(acc, eN[e], sn)
        np.add.at(cnt, eN[e], 1)
    want = acc / cnt[:, None]
    np.testing.assert_allclose(st, want, rtol=1e-11, atol=1e-13 * np.abs(want).max())
