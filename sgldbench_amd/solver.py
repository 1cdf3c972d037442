"""Host-side mirror of SGLDBench's solve interface over libsgldfea.

The function names, argument order and meaning follow the reference's MEX
calls and MATLAB functions; each cites the file it replaces.  All numerics run
in the HIP library; this module only marshals arrays (1-based int32 ids,
column-major matrices, interleaved FP64 DOF vectors) across the C ABI.
"""
import ctypes as C

import numpy as np

from ._lib import SGLDError, check, lib, ERR_MAXIT

PREC_FP64 = 0
PREC_MIXED = 1
PREC_INTERLEAVED = 2  # pcg_batched only: the 16 cases interleaved per DOF, matrix-core K.u

# Kernel timer ids (SGLD_KT_* of include/sgldfea.h), in order.
KERNEL_TIMER_IDS = ("dir64", "dir32", "kbyu64", "kbyu32", "update_r64", "update_r32",
                    "restrict64", "restrict32")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _fortran_i32(a):
    return _i32(np.asarray(a).ravel(order="F"))


def _fortran_f64(a):
    return _f64(np.asarray(a).ravel(order="F"))


def _colors(colors):
    off = np.zeros(len(colors) + 1, dtype=np.int64)
    for i, c in enumerate(colors):
        off[i + 1] = off[i] + len(c)
    flat = np.concatenate([np.asarray(c, dtype=np.int32).ravel() for c in colors]) \
        if off[-1] else np.zeros(1, dtype=np.int32)
    return _i32(flat), off, len(colors)


def identifier(err):
    """MATLAB-style error identifier of an SGLDError message."""
    return str(err).split(": ")[0]


def Solving_KbyU_MatrixFree_mex(U, eNodMat, Ke, E, colors):
    """src/MEXfuncs/Solving_KbyU_MatrixFree_mex.cpp:6-103."""
    U = _f64(U).ravel()
    eNodMat = np.asarray(eNodMat)
    Ke = np.asarray(Ke, dtype=np.float64)
    E = _f64(E).ravel()
    flat, off, nc = _colors(colors)
    Y = np.zeros_like(U)
    kr, kc = (Ke.shape + (1,))[:2] if Ke.ndim == 2 else (Ke.size, 1)
    check(lib().sgld_kbyu(U, U.size, _fortran_i32(eNodMat), eNodMat.shape[0], _fortran_f64(Ke),
                          kr, kc, E, E.size, flat, off, nc, Y))
    return Y


def Solving_KbyU_MatrixFree8x8_mex(U, eNodMat, Ke, colors):
    """src/MEXfuncs/Solving_KbyU_MatrixFree8x8_mex.cpp:6-121 (one DOF per node)."""
    U = _f64(U).ravel()
    eNodMat = np.asarray(eNodMat)
    Ke = np.asarray(Ke, dtype=np.float64)
    flat, off, nc = _colors(colors)
    Y = np.zeros_like(U)
    kr, kc = (Ke.shape + (1,))[:2] if Ke.ndim == 2 else (Ke.size, 1)
    check(lib().sgld_kbyu8x8(U, U.size, _fortran_i32(eNodMat), eNodMat.shape[0], _fortran_f64(Ke),
                             kr, kc, flat, off, nc, Y))
    return Y


def Solving_Restriction_MatrixFree_mex(rF, transferMat, RIT, eNodMat, numNodesCoarse,
                                       numNodesFineGhost, solidNodeMapCoarser2Finer,
                                       transferMatCoeffi, colors):
    """src/MEXfuncs/Solving_Restriction_MatrixFree_mex.cpp:6-191 (rF: n_f x 3)."""
    rF = np.asarray(rF, dtype=np.float64)
    transferMat = np.asarray(transferMat)
    RIT = np.asarray(RIT, dtype=np.float64)
    eNodMat = np.asarray(eNodMat)
    flat, off, nc = _colors(colors)
    nC = int(numNodesCoarse)
    Y = np.zeros(3 * nC)
    check(lib().sgld_restrict(_fortran_f64(rF), rF.shape[0], _fortran_i32(transferMat),
                              transferMat.shape[0], _fortran_f64(RIT), RIT.shape[0], RIT.shape[1],
                              _fortran_i32(eNodMat), eNodMat.shape[0], nC, int(numNodesFineGhost),
                              _i32(solidNodeMapCoarser2Finer).ravel(), _f64(transferMatCoeffi).ravel(),
                              flat, off, nc, Y))
    return Y.reshape(nC, 3, order="F")


def Solving_Interpolation_MatrixFree_mex(xC, transferMat, RI, eNodMat, numNodesFine,
                                         numNodesFineGhost, solidNodeMapCoarser2Finer,
                                         transferMatCoeffi, colors):
    """src/MEXfuncs/Solving_Interpolation_MatrixFree_mex.cpp:6-190 (xC: n_c x 3)."""
    xC = np.asarray(xC, dtype=np.float64)
    transferMat = np.asarray(transferMat)
    RI = np.asarray(RI, dtype=np.float64)
    eNodMat = np.asarray(eNodMat)
    flat, off, nc = _colors(colors)
    nF = int(numNodesFine)
    Y = np.zeros(3 * nF)
    check(lib().sgld_interpolate(_fortran_f64(xC), xC.shape[0], _fortran_i32(transferMat),
                                 transferMat.shape[0], _fortran_f64(RI), RI.shape[0], RI.shape[1],
                                 _fortran_i32(eNodMat), eNodMat.shape[0], nF, int(numNodesFineGhost),
                                 _i32(solidNodeMapCoarser2Finer).ravel(),
                                 _f64(transferMatCoeffi).ravel(), flat, off, nc, Y))
    return Y.reshape(nF, 3, order="F")


def Solving_AssembleCmptStencilFromFinestLevel(Ke, E, elementUpwardMap, interpolatingKe=None,
                                               localMapping=None, numProjectNodes=27):
    """src/MEXfuncs/Solving_AssembleCmptStencilFromFinestLevel.cpp:12-169.
    interpolatingKe/localMapping are implied by the span ((s+1)^3 = numProjectNodes)."""
    Ke = np.asarray(Ke, dtype=np.float64)
    up = np.asarray(elementUpwardMap)
    E = _f64(E).ravel()
    nC = up.shape[0]
    Ks = np.zeros(576 * max(nC, 1))
    check(lib().sgld_assemble_finest(_fortran_f64(Ke), Ke.shape[0], Ke.shape[1], E, E.size,
                                     _fortran_i32(up), nC, up.shape[1], Ks))
    return Ks[:576 * nC].reshape(24, 24, nC, order="F")


def Solving_AssembleCmptStencilFromNonFinestLevel(KsPrevious, elementUpwardMap, interpolatingKe=None,
                                                  localMapping=None, numProjectNodes=27):
    """src/MEXfuncs/Solving_AssembleCmptStencilFromNonFinestLevel.cpp:12-143."""
    Kp = np.asarray(KsPrevious, dtype=np.float64)
    if Kp.ndim == 2:
        Kp = Kp[:, :, None]
    if Kp.shape[0] != 24 or Kp.shape[1] != 24:
        raise SGLDError(1, "MATLAB:mexFunction:KsPreviousDim: KsPrevious must be 24x24xN.")
    up = np.asarray(elementUpwardMap)
    nC = up.shape[0]
    Ks = np.zeros(576 * max(nC, 1))
    check(lib().sgld_assemble_nonfinest(_fortran_f64(Kp), Kp.shape[2], _fortran_i32(up), nC,
                                        up.shape[1], Ks))
    return Ks[:576 * nC].reshape(24, 24, nC, order="F")


def TopOpti_CmptUnitCompliance_mex(U, eNodMat, Ke):
    """src/MEXfuncs/TopOpti_CmptUnitCompliance_mex.cpp:6-59."""
    U = _f64(U).ravel()
    eNodMat = np.asarray(eNodMat)
    Ke = np.asarray(Ke, dtype=np.float64)
    ce = np.zeros(max(eNodMat.shape[0], 1))
    check(lib().sgld_unit_compliance(U, U.size, _fortran_i32(eNodMat), eNodMat.shape[0],
                                     _fortran_f64(Ke), Ke.shape[0], Ke.shape[1], ce))
    return ce[:eNodMat.shape[0]]


def FEA_VoxelBasedElementStiffnessMatrix(poissonRatio=0.3, cellSize=1.0):
    """src/FEA_VoxelBasedElementStiffnessMatrix.m:1-24."""
    Ke = np.zeros(576)
    check(lib().sgld_element_stiffness(poissonRatio, cellSize, Ke))
    return Ke.reshape(24, 24, order="F")


def TopOpti_SetupDensityFilter_mex(rMin, numElements, eleMapForward, resX, resY, resZ):
    """Hs (numElements) of the cone density filter; eleMapForward: voxels in
    (resY, resX, resZ) column-major order, 1-based element id or 0."""
    m = np.ascontiguousarray(np.asarray(eleMapForward).ravel(order="F"), dtype=np.int32)
    Hs = np.zeros(int(numElements))
    check(lib().sgld_density_filter_setup(float(rMin), int(numElements), m, int(resX), int(resY), int(resZ), Hs))
    return Hs


def TopOpti_PerformDensityFiltering_mex(x, rMin, numElements, eleMapForward, resX, resY, resZ):
    """Y = sum_e2 max(0, rMin - dist) x[e2] per solid element."""
    m = np.ascontiguousarray(np.asarray(eleMapForward).ravel(order="F"), dtype=np.int32)
    x = _f64(x).ravel()
    Y = np.zeros(int(numElements))
    check(lib().sgld_density_filter(x, float(rMin), int(numElements), m, int(resX), int(resY), int(resZ), Y))
    return Y


class FEASolver:
    """Device-resident meshHierarchy_ plus the solve functions that read it:
    FEA_SetupVoxelBased, Solving_AssembleFEAstencil, Solving_KbyU_MatrixFree,
    Solving_Vcycle, Solving_PreconditionedConjugateGradientSolver (as called by
    Solving_CG_GMGS) and TopOpti_ComputeUnitCompliance."""

    def __init__(self, voxels, numLevels, nonDyadic=0, poissonRatio=0.3, cellSize=1.0):
        vox = np.asarray(voxels)
        ny, nx, nz = vox.shape
        h = C.c_void_p()
        if vox.all():
            vp = None
        else:
            self._vox = np.ascontiguousarray(vox.ravel(order="F") != 0, dtype=np.uint8)
            vp = self._vox.ctypes.data_as(C.c_void_p)
        check(lib().sgld_ctx_create(C.byref(h), nx, ny, nz, vp, numLevels, nonDyadic,
                                    poissonRatio, cellSize))
        self._h = h
        self.numLevels = check_levels = lib().sgld_ctx_info(self._h, 0, np.zeros(8, np.int64))
        self.levels = []
        for lv in range(1, check_levels + 1):
            out = np.zeros(8, dtype=np.int64)
            check(lib().sgld_ctx_info(self._h, lv, out))
            self.levels.append(dict(zip(["resX", "resY", "resZ", "spanWidth", "numElements",
                                         "numNodes", "numDOFs", "latticeNodes"],
                                        [int(v) for v in out])))
        self.numDOFs = self.levels[0]["numDOFs"]
        self.numElements = self.levels[0]["numElements"]

    def close(self):
        if getattr(self, "_h", None):
            try:
                lib().sgld_ctx_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def __del__(self):
        self.close()

    def set_fixed(self, fixedDOFs):
        f = np.ascontiguousarray(np.asarray(fixedDOFs).ravel() != 0, dtype=np.uint8)
        assert f.size == self.numDOFs
        check(lib().sgld_ctx_set_fixed(self._h, f))

    def set_modulus(self, eleModulus):
        E = _f64(eleModulus).ravel()
        assert E.size == self.numElements
        check(lib().sgld_ctx_set_modulus(self._h, E))

    def assemble(self, weightFactorJacobi=0.6):
        """Solving_AssembleFEAstencil()"""
        check(lib().sgld_ctx_assemble(self._h, weightFactorJacobi))

    def kbyu(self, u):
        """productMV = Solving_KbyU_MatrixFree(uVec)"""
        u = _f64(u).ravel()
        y = np.zeros_like(u)
        check(lib().sgld_ctx_kbyu(self._h, u, y))
        return y

    def vcycle(self, r):
        """rTilde = Solving_Vcycle(r)"""
        r = _f64(r).ravel()
        z = np.zeros_like(r)
        check(lib().sgld_ctx_vcycle(self._h, r, z))
        return z

    def pcg(self, F, tol=1e-8, maxIT=500, U0=None, precision=PREC_FP64):
        """[U, its] = Solving_PreconditionedConjugateGradientSolver(@Solving_KbyU_MatrixFree,
        @Solving_Vcycle, F, tol, maxIT, printP, U0); also returns the FP64 true
        relative residual.  Reaching maxIT returns (as the reference warns)."""
        F = _f64(F).ravel()
        U = np.zeros_like(F) if U0 is None else _f64(U0).ravel().copy()
        its = C.c_int(0)
        rr = C.c_double(0)
        rc = check(lib().sgld_ctx_pcg(self._h, F, U, tol, maxIT, precision, C.byref(its),
                                      C.byref(rr)), allow=(ERR_MAXIT,))
        self.converged = rc == 0
        return U, its.value, rr.value

    def pcg_device(self, dF_ptr, dU_ptr, tol=1e-8, maxIT=500, precision=PREC_FP64):
        its = C.c_int(0)
        rr = C.c_double(0)
        rc = check(lib().sgld_ctx_pcg_device(self._h, C.c_void_p(dF_ptr), C.c_void_p(dU_ptr), tol,
                                             maxIT, precision, C.byref(its), C.byref(rr)),
                   allow=(ERR_MAXIT,))
        self.converged = rc == 0
        return its.value, rr.value

    def unit_compliance(self, U):
        """ceList = TopOpti_ComputeUnitCompliance()"""
        U = _f64(U).ravel()
        ce = np.zeros(self.numElements)
        check(lib().sgld_ctx_unit_compliance(self._h, U, ce))
        return ce

    def kbyu_device(self, du_ptr, dy_ptr, precision=PREC_FP64):
        """K.u on device vectors (compressed layout); FP32 arithmetic for
        PREC_MIXED (inputs/outputs stay FP64)."""
        check(lib().sgld_ctx_kbyu_device(self._h, C.c_void_p(du_ptr), C.c_void_p(dy_ptr), precision))

    def bench_kbyu(self, precision=PREC_FP64, reps=20):
        ms = C.c_double(0)
        check(lib().sgld_ctx_bench_kbyu(self._h, precision, reps, C.byref(ms)))
        return ms.value

    def pcg_batched(self, F, tol=1e-8, maxIT=500, U0=None, precision=PREC_MIXED):
        """Up to 16 load cases at once: F is (nrhs, numDOFs) (one case per row),
        each case solved as Solving_CG_GMGS would solve it alone.  Returns
        (U (nrhs, numDOFs), its[nrhs], relres[nrhs]); relres is the FP64 true
        relative residual of each case."""
        F = np.ascontiguousarray(np.atleast_2d(F), dtype=np.float64)
        k = F.shape[0]
        assert F.shape[1] == self.numDOFs and 1 <= k <= 16
        U = np.zeros_like(F) if U0 is None else np.ascontiguousarray(U0, dtype=np.float64).reshape(F.shape).copy()
        its = np.zeros(k, dtype=np.int32)
        rr = np.zeros(k)
        rc = check(lib().sgld_ctx_pcg_batched(self._h, k, F, U, tol, maxIT, precision, its, rr),
                   allow=(ERR_MAXIT,))
        self.converged = rc == 0
        return U, its, rr

    def pcg_batched_device(self, nrhs, dF_ptr, dU_ptr, tol=1e-8, maxIT=500, precision=PREC_MIXED):
        its = np.zeros(nrhs, dtype=np.int32)
        rr = np.zeros(nrhs)
        rc = check(lib().sgld_ctx_pcg_batched_device(self._h, nrhs, C.c_void_p(dF_ptr), C.c_void_p(dU_ptr), tol,
                                                     maxIT, precision, its, rr), allow=(ERR_MAXIT,))
        self.converged = rc == 0
        return its, rr

    def kbyu_batched_device(self, nrhs, dU_ptr, dY_ptr):
        """K U for nrhs <= 16 device vectors (case-major, compressed) through the
        batched matrix-core kernel."""
        check(lib().sgld_ctx_kbyu_batched_device(self._h, nrhs, C.c_void_p(dU_ptr), C.c_void_p(dY_ptr)))

    def bench_kbyu_batched(self, reps=20):
        ms = C.c_double(0)
        check(lib().sgld_ctx_bench_kbyu_batched(self._h, reps, C.byref(ms)))
        return ms.value

    def sensitivity(self, U, x, penal=3.0, E0=1.0, Emin=1e-9, scale=1.0):
        """dc = -TopOpti_DeMaterialInterpolation(x) .* ceList * scale and the
        compliance scale * E . ceList (TopOpti_GlobalVolumeConstraint.m:113-118)."""
        U = _f64(U).ravel()
        x = _f64(x).ravel()
        assert U.size == self.numDOFs and x.size == self.numElements
        dc = np.zeros(self.numElements)
        c = C.c_double(0)
        check(lib().sgld_ctx_sensitivity(self._h, U, x, penal, E0, Emin, scale, dc, C.byref(c)))
        return dc, c.value

    def sensitivity_device(self, dU_ptr, dx_ptr, ddc_ptr, penal=3.0, E0=1.0, Emin=1e-9, scale=1.0):
        c = C.c_double(0)
        check(lib().sgld_ctx_sensitivity_device(self._h, C.c_void_p(dU_ptr), C.c_void_p(dx_ptr), penal, E0,
                                                Emin, scale, C.c_void_p(ddc_ptr), C.byref(c)))
        return c.value

    def stress(self, U):
        """[cartesianStressField, vonMisesStressField] = FEA_StressAnalysis():
        (numNodes, 6) stresses (xx yy zz yz zx xy) and (numNodes,) von Mises."""
        U = _f64(U).ravel()
        nn = self.levels[0]["numNodes"]
        st = np.zeros(6 * nn)
        vm = np.zeros(nn)
        check(lib().sgld_ctx_stress(self._h, U, st, vm))
        return st.reshape(nn, 6, order="F"), vm

    def pde_filter(self, src, rHatMin=8.0, tol=1e-3, maxIT=500):
        """tar = TopOpti_PDEFiltering_matrixFree(src); returns (tar, its)."""
        src = _f64(src).ravel()
        tar = np.zeros(self.numElements)
        its = C.c_int(0)
        check(lib().sgld_ctx_pde_filter(self._h, src, float(rHatMin), float(tol), int(maxIT), tar, C.byref(its)))
        return tar, its.value

    def density_filter_device(self, dx_ptr, rMin, dy_ptr):
        """Density filter on this mesh with device vectors; dx_ptr 0 gives Hs."""
        check(lib().sgld_ctx_density_filter_device(self._h, C.c_void_p(dx_ptr) if dx_ptr else None, float(rMin),
                                                   C.c_void_p(dy_ptr)))

    def autotune(self, reps=10):
        """Picks the fastest K.u variant / z-chunk and span-2 restriction for
        this mesh (kept in the context)."""
        out = np.zeros(8)
        check(lib().sgld_ctx_autotune(self._h, reps, out))
        return dict(kbyu_variant="tiled" if out[0] == 1 else "dense", z_chunk=int(out[1]),
                    z_chunk_fp32=int(out[7]), kbyu_ms_fp64=out[2], kbyu_ms_fp32=out[3],
                    restriction="tiled" if out[4] == 1 else "node-centric", vcycle_ms=(out[5], out[6]))

    def set_schedule(self, kbyu_variant="tiled", z_chunk=16, restriction=None, blocks_per_cu=0):
        """Pins a K.u / restriction schedule (what autotune picks)."""
        rt = {None: -1, "node-centric": 0, "tiled": 1, "tiled-first": 2}[restriction]
        check(lib().sgld_ctx_set_schedule(self._h, 1 if kbyu_variant == "tiled" else 0, z_chunk, rt,
                                          blocks_per_cu))

    def stats(self):
        out = np.zeros(8)
        check(lib().sgld_ctx_stats(self._h, out))
        return dict(zip(["solve_s", "its", "true_relres", "outer", "setup_s", "filter_ms",
                         "kbyu_ms_fp64", "kbyu_ms_fp32"], out.tolist()))

    def kernel_timing(self, on=True):
        """HIP-event timing of the finest-level passes of the following solves
        (resets the counters)."""
        check(lib().sgld_ctx_kernel_timing(self._h, 1 if on else 0))

    def kernel_times(self):
        """{kernel: dict(launches, ms, work_launches, work_ms)} of the timed solves."""
        out = np.zeros(4 * len(KERNEL_TIMER_IDS))
        check(lib().sgld_ctx_kernel_times(self._h, out))
        return {name: dict(launches=int(out[4 * k]), ms=float(out[4 * k + 1]),
                           work_launches=int(out[4 * k + 2]), work_ms=float(out[4 * k + 3]),
                           kernel=lib().sgld_ctx_kernel_name(self._h, k).decode())
                for k, name in enumerate(KERNEL_TIMER_IDS)}


def multi_pcg_device(solvers, dF_ptrs, dU_ptrs, tol=1e-8, maxIT=500, precision=PREC_MIXED, batch=8):
    """Solves len(solvers) independent problems in one replayed graph
    (sgld_multi_pcg_device).  Returns (its, relres, latency_s, converged)."""
    n = len(solvers)
    ctxs = (C.c_void_p * n)(*[s._h.value for s in solvers])
    fp = (C.c_void_p * n)(*dF_ptrs)
    up = (C.c_void_p * n)(*dU_ptrs)
    its = np.zeros(n, dtype=np.int32)
    rr = np.zeros(n)
    lat = np.zeros(n)
    rc = check(lib().sgld_multi_pcg_device(ctxs, n, fp, up, tol, maxIT, precision, batch, its, rr, lat),
               allow=(ERR_MAXIT,))
    return its, rr, lat, rc == 0


def multi_sensitivity_device(solvers, dU_ptrs, dx_ptrs, ddc_ptrs, penal=3.0, E0=1.0, Emin=1e-9, scale=1.0):
    """FEASolver.sensitivity_device for every solver in one launch
    (sgld_multi_sensitivity_device).  Returns the compliances (numpy array)."""
    n = len(solvers)
    ctxs = (C.c_void_p * n)(*[s._h.value for s in solvers])
    up = (C.c_void_p * n)(*dU_ptrs)
    xp = (C.c_void_p * n)(*dx_ptrs)
    cp = (C.c_void_p * n)(*ddc_ptrs)
    comp = np.zeros(n)
    check(lib().sgld_multi_sensitivity_device(ctxs, n, up, xp, penal, E0, Emin, scale, cp, comp))
    return comp
