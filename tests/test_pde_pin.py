"""Pins the oracle's PDE (Helmholtz) filter (TopOpti_BuildPDEfilter_matrixFree.m,
TopOpti_PDEFiltering_matrixFree.m): its element matrix against the exact
tensor-product integrals (2-point Gauss is exact for both terms), and the
filter against a scipy direct solve of the assembled system."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as orc
from sgldbench_amd import problems

from tests.test_gpu_solver import holes

CX = [0, 1, 1, 0, 0, 1, 1, 0]
CY = [0, 0, 1, 1, 0, 0, 1, 1]
CZ = [0, 0, 0, 0, 1, 1, 1, 1]


def exact_pde_element(rHatMin):
    K1 = np.array([[0.5, -0.5], [-0.5, 0.5]])        # int_{-1}^{1} phi_a' phi_b'
    M1 = np.array([[2 / 3, 1 / 3], [1 / 3, 2 / 3]])  # int_{-1}^{1} phi_a phi_b
    KEF0 = np.zeros((8, 8))
    KEF1 = np.zeros((8, 8))
    for a in range(8):
        for b in range(8):
            ax, ay, az, bx, by, bz = CX[a], CY[a], CZ[a], CX[b], CY[b], CZ[b]
            KEF0[a, b] = (K1[ax, bx] * M1[ay, by] * M1[az, bz] + M1[ax, bx] * K1[ay, by] * M1[az, bz]
                          + M1[ax, bx] * M1[ay, by] * K1[az, bz])
            KEF1[a, b] = M1[ax, bx] * M1[ay, by] * M1[az, bz] / 8.0   # detJ = 1/8
    r = rHatMin / 2 / np.sqrt(3)
    return r * r * KEF0 + KEF1


def test_pde_element_exact():
    for rh in (2.0, 8.0):
        np.testing.assert_allclose(orc.pde_element(rh), exact_pde_element(rh), rtol=1e-13, atol=1e-15)


def test_pde_filter_matches_direct_solve():
    vox = holes(16, 8, 8, 3, nholes=2, r=2)
    m = orc.Mesh(vox, 2)
    ne = int(vox.sum())
    eN = m.get(1, "eNodMat") - 1
    nn = eN.max() + 1
    Ke = orc.pde_element(4.0)
    # entry (j, k) of element e goes to (eN[e, j], eN[e, k])
    rows = eN[:, :, None].repeat(8, axis=2).ravel()
    cols = eN[:, None, :].repeat(8, axis=1).ravel()
    vals = np.tile(Ke.ravel(order="C"), ne)
    K = sp.csr_matrix((vals, (rows, cols)), shape=(nn, nn))
    src = np.random.default_rng(2).uniform(0, 1, ne)
    b = np.zeros(nn)
    np.add.at(b, eN.ravel(), np.repeat(src / 8, 8))
    xs = spla.spsolve(K.tocsc(), b)
    want = xs[eN].mean(axis=1)
    tar, its = m.pde_filter(src, 4.0, 1e-12, 1000)
    np.testing.assert_allclose(tar, want, rtol=1e-9)
    # constant in, constant out (the Helmholtz operator keeps constants with natural BCs)
    t1, _ = m.pde_filter(np.full(ne, 0.7), 4.0, 1e-12, 1000)
    np.testing.assert_allclose(t1, 0.7, rtol=1e-9)
