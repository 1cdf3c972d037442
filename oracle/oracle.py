"""ctypes binding of the CPU oracle (oracle/libsgld_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.  Every function follows the reference file:line cited in
oracle/sgld_oracle.c.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsgld_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i64 = C.c_int64


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_element_stiffness.argtypes = [C.c_double, C.c_double, _dp]
        L.or_kbyu_mex.argtypes = [_dp, _i64, _ip, _i64, _dp, _dp, _ip, _lp, _dp]
        L.or_kbyu8x8_mex.argtypes = [_dp, _i64, _ip, _i64, _dp, _ip, _lp, C.c_int, _dp]
        L.or_restrict_mex.argtypes = [_dp, _i64, _ip, C.c_int, _dp, _ip, _i64, _i64, _i64,
                                      _ip, _dp, _ip, _lp, _dp]
        L.or_interpolate_mex.argtypes = [_dp, _i64, _ip, C.c_int, _dp, _ip, _i64, _i64, _i64,
                                         _ip, _dp, _ip, _lp, _dp]
        L.or_assemble_finest.argtypes = [_dp, _dp, _ip, _i64, C.c_int, _dp]
        L.or_assemble_nonfinest.argtypes = [_dp, _ip, _i64, C.c_int, _dp]
        L.or_unit_compliance_mex.argtypes = [_dp, _ip, _i64, _dp, _dp]
        L.or_transfer_table.argtypes = [C.c_int, _dp]
        L.or_sub_ele_nod_mat.argtypes = [C.c_int, _ip]
        L.or_mesh_create.restype = C.c_void_p
        L.or_mesh_create.argtypes = [C.c_int, C.c_int, C.c_int, _bp, C.c_int, C.c_int,
                                     C.c_double, C.c_double]
        L.or_mesh_destroy.argtypes = [C.c_void_p]
        L.or_mesh_num_levels.argtypes = [C.c_void_p]
        L.or_mesh_level_info.argtypes = [C.c_void_p, C.c_int, _lp]
        L.or_mesh_get.restype = _i64
        L.or_mesh_get.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_void_p]
        L.or_mesh_set_fixed.argtypes = [C.c_void_p, _bp]
        L.or_mesh_set_modulus.argtypes = [C.c_void_p, _dp]
        L.or_mesh_assemble.argtypes = [C.c_void_p]
        L.or_mesh_set_lean.argtypes = [C.c_void_p, C.c_int]
        L.or_mesh_kbyu.argtypes = [C.c_void_p, _dp, _dp]
        L.or_mesh_restrict.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.or_mesh_interpolate.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.or_mesh_vcycle.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.or_mesh_pcg.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_double, C.c_int,
                                  _dp, C.POINTER(C.c_int)]
        L.or_mesh_unit_compliance.argtypes = [C.c_void_p, _dp, _dp]
        L.or_density_filter_setup.argtypes = [C.c_double, _i64, _ip, C.c_int, C.c_int, C.c_int, _dp]
        L.or_density_filter.argtypes = [_dp, C.c_double, _i64, _ip, C.c_int, C.c_int, C.c_int, _dp]
        L.or_mesh_stress.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_pde_element.argtypes = [C.c_double, _dp]
        L.or_mesh_pde_filter.argtypes = [C.c_void_p, C.c_double, _dp, C.c_double, C.c_int, _dp]
        L.or_set_threads.restype = C.c_int
        L.or_set_threads.argtypes = [C.c_int]
        L.or_mesh_time_pcg_iterations.restype = C.c_double
        L.or_mesh_time_pcg_iterations.argtypes = [C.c_void_p, C.c_double, _dp, C.c_int,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def element_stiffness(nu=0.3, cell_size=1.0):
    Ke = np.zeros(576)
    lib().or_element_stiffness(nu, cell_size, Ke)
    return Ke.reshape(24, 24, order="F")


def transfer_table(span):
    m = (span + 1) ** 3
    RI = np.zeros(m * 8)
    lib().or_transfer_table(span, RI)
    return RI.reshape(m, 8, order="F")


def sub_ele_nod_mat(span):
    out = np.zeros(span ** 3 * 8, dtype=np.int32)
    lib().or_sub_ele_nod_mat(span, out)
    return out.reshape(span ** 3, 8, order="F")


def _colors_flat(colors):
    off = np.zeros(9, dtype=np.int64)
    for i, c in enumerate(colors):
        off[i + 1] = off[i] + len(c)
    flat = i32(np.concatenate([np.asarray(c, dtype=np.int32) for c in colors]) if off[8] else np.zeros(1))
    return flat, off


def kbyu_mex(U, eNodMat, Ke, E, colors):
    """Solving_KbyU_MatrixFree_mex(U, eNodMat, Ke, E, colors)."""
    U = f64(U)
    eN = i32(np.asfortranarray(eNodMat).ravel(order="F"))
    flat, off = _colors_flat(colors)
    Y = np.zeros_like(U)
    lib().or_kbyu_mex(U, U.size, eN, eNodMat.shape[0], f64(np.asarray(Ke).ravel(order="F")),
                      f64(E), flat, off, Y)
    return Y


def kbyu8x8_mex(U, eNodMat, Ke, colors):
    """Solving_KbyU_MatrixFree8x8_mex(U, eNodMat, Ke, colors): any number of colours."""
    U = f64(U)
    eN = i32(np.asfortranarray(eNodMat).ravel(order="F"))
    off = np.zeros(len(colors) + 1, dtype=np.int64)
    for i, c in enumerate(colors):
        off[i + 1] = off[i] + len(c)
    flat = i32(np.concatenate([np.asarray(c, dtype=np.int32).ravel() for c in colors])
               if off[-1] else np.zeros(1))
    Y = np.zeros_like(U)
    lib().or_kbyu8x8_mex(U, U.size, eN, eNodMat.shape[0], f64(np.asarray(Ke).ravel(order="F")),
                         flat, off, len(colors), Y)
    return Y


def restrict_mex(rF, transferMat, RIT, eNodMatC, nC, nGhost, mapC2F, coeffi, colorsC):
    rF = np.asarray(rF, dtype=np.float64)
    nF = rF.shape[0]
    m = transferMat.shape[0]
    flat, off = _colors_flat(colorsC)
    Y = np.zeros(3 * nC)
    lib().or_restrict_mex(f64(rF.ravel(order="F")), nF, i32(transferMat.ravel(order="F")), m,
                          f64(np.asarray(RIT).ravel(order="F")), i32(eNodMatC.ravel(order="F")),
                          eNodMatC.shape[0], nC, nGhost, i32(mapC2F), f64(coeffi), flat, off, Y)
    return Y.reshape(nC, 3, order="F")


def interpolate_mex(xC, transferMat, RI, eNodMatC, nF, nGhost, mapC2F, coeffi, colorsC):
    xC = np.asarray(xC, dtype=np.float64)
    nC = xC.shape[0]
    m = transferMat.shape[0]
    flat, off = _colors_flat(colorsC)
    Y = np.zeros(3 * nF)
    lib().or_interpolate_mex(f64(xC.ravel(order="F")), nC, i32(transferMat.ravel(order="F")), m,
                             f64(np.asarray(RI).ravel(order="F")), i32(eNodMatC.ravel(order="F")),
                             eNodMatC.shape[0], nF, nGhost, i32(mapC2F), f64(coeffi), flat, off, Y)
    return Y.reshape(nF, 3, order="F")


def assemble_finest(Ke, E, upMap, span):
    nC = upMap.shape[0]
    Ks = np.zeros(576 * nC)
    lib().or_assemble_finest(f64(np.asarray(Ke).ravel(order="F")), f64(E),
                             i32(upMap.ravel(order="F")), nC, span, Ks)
    return Ks.reshape(24, 24, nC, order="F")


def assemble_nonfinest(KsPrev, upMap, span):
    nC = upMap.shape[0]
    Ks = np.zeros(576 * nC)
    lib().or_assemble_nonfinest(f64(np.asarray(KsPrev).ravel(order="F")),
                                i32(upMap.ravel(order="F")), nC, span, Ks)
    return Ks.reshape(24, 24, nC, order="F")


def unit_compliance_mex(U, eNodMat, Ke):
    ne = eNodMat.shape[0]
    ce = np.zeros(ne)
    lib().or_unit_compliance_mex(f64(U), i32(eNodMat.ravel(order="F")), ne,
                                 f64(np.asarray(Ke).ravel(order="F")), ce)
    return ce


def pde_element(rHatMin):
    Ke = np.zeros(64)
    lib().or_pde_element(rHatMin, Ke)
    return Ke.reshape(8, 8, order="F")


def density_filter_setup(rMin, eleMapForward, resX, resY, resZ):
    """Hs = TopOpti_SetupDensityFilter_mex(rMin, numElements, eleMapForward, resX, resY, resZ)."""
    efwd = i32(np.asarray(eleMapForward).ravel(order="F"))
    ne = int(efwd.max()) if efwd.size else 0
    Hs = np.zeros(ne)
    lib().or_density_filter_setup(rMin, ne, efwd, resX, resY, resZ, Hs)
    return Hs


def density_filter(x, rMin, eleMapForward, resX, resY, resZ):
    """Y = TopOpti_PerformDensityFiltering_mex(x, rMin, numElements, eleMapForward, resX, resY, resZ)."""
    efwd = i32(np.asarray(eleMapForward).ravel(order="F"))
    x = f64(x)
    Y = np.zeros(x.size)
    lib().or_density_filter(x, rMin, x.size, efwd, resX, resY, resZ, Y)
    return Y


class Mesh:
    """meshHierarchy_ restated (FEA_VoxelBasedDiscretization.m,
    Solving_BuildingMeshHierarchy.m, FEA_SetupVoxelBased.m,
    Solving_AssembleFEAstencil.m)."""

    def __init__(self, voxels, num_levels, non_dyadic=0, nu=0.3, cell_size=1.0):
        vox = np.asarray(voxels)
        ny, nx, nz = vox.shape
        flat = np.ascontiguousarray(vox.ravel(order="F") != 0, dtype=np.uint8)
        self._h = lib().or_mesh_create(nx, ny, nz, flat, num_levels, non_dyadic, nu, cell_size)
        if not self._h:
            raise ValueError("or_mesh_create failed (grid not divisible by the level spans?)")
        self.num_levels = lib().or_mesh_num_levels(self._h)
        self.info = []
        for lv in range(1, self.num_levels + 1):
            out = np.zeros(8, dtype=np.int64)
            lib().or_mesh_level_info(self._h, lv, out)
            self.info.append(dict(zip(
                ["resX", "resY", "resZ", "span", "numElements", "numNodes", "numDOFs",
                 "intermediateNumNodes"], [int(v) for v in out])))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_mesh_destroy(self._h)
            self._h = None

    def get(self, level, name):
        dt = {"colorOff": np.int64, "fixedDOFs": np.uint8, "transferMatCoeffi": np.float64,
              "RI": np.float64, "diagK": np.float64, "Ks": np.float64, "Ke": np.float64,
              "eleModulus": np.float64}.get(name, np.int32)
        n = lib().or_mesh_get(self._h, level, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n, dtype=dt)
        lib().or_mesh_get(self._h, level, name.encode(), out.ctypes.data_as(C.c_void_p))
        info = self.info[level - 1]
        s = info["span"]
        if name == "eNodMat":
            return out.reshape(info["numElements"], 8, order="F")
        if name == "elementUpwardMap":
            return out.reshape(info["numElements"], s ** 3, order="F")
        if name == "transferMat":
            return out.reshape((s + 1) ** 3, info["numElements"], order="F")
        if name == "RI":
            return out.reshape((s + 1) ** 3, 8, order="F")
        if name == "Ks":
            return out.reshape(24, 24, info["numElements"], order="F")
        if name == "Ke":
            return out.reshape(24, 24, order="F")
        return out

    def colors(self, level):
        flat = self.get(level, "colors")
        off = self.get(level, "colorOff")
        return [flat[off[i]:off[i + 1]] for i in range(8)]

    def set_fixed(self, fixed):
        lib().or_mesh_set_fixed(self._h, np.ascontiguousarray(fixed, dtype=np.uint8))

    def set_modulus(self, E):
        lib().or_mesh_set_modulus(self._h, f64(E))

    def assemble(self, check=True, lean=False):
        """lean: keep no level-2 Galerkin matrices (same results, less memory)."""
        lib().or_mesh_set_lean(self._h, 1 if lean else 0)
        rc = lib().or_mesh_assemble(self._h)
        if rc != 0 and check:
            raise RuntimeError("coarsest matrix not SPD")

    def kbyu(self, u):
        u = f64(u)
        y = np.zeros_like(u)
        lib().or_mesh_kbyu(self._h, u, y)
        return y

    def restrict(self, ii, rF):
        out = np.zeros(self.info[ii - 1]["numDOFs"])
        lib().or_mesh_restrict(self._h, ii, f64(rF), out)
        return out

    def interpolate(self, ii, xC):
        out = np.zeros(self.info[ii - 2]["numDOFs"])
        lib().or_mesh_interpolate(self._h, ii, f64(xC), out)
        return out

    def vcycle(self, r, omega):
        z = np.zeros(self.info[0]["numDOFs"])
        lib().or_mesh_vcycle(self._h, omega, f64(r), z)
        return z

    def pcg(self, b, omega, tol, max_it, x0=None):
        x = np.zeros(self.info[0]["numDOFs"]) if x0 is None else f64(x0).copy()
        hist = np.zeros(max_it)
        its = C.c_int(0)
        rc = lib().or_mesh_pcg(self._h, omega, f64(b), x, tol, max_it, hist, C.byref(its))
        return x, its.value, hist[:its.value], rc

    def unit_compliance(self, U):
        ce = np.zeros(self.info[0]["numElements"])
        lib().or_mesh_unit_compliance(self._h, f64(U), ce)
        return ce

    def stress(self, U):
        """[cartesianStressField, vonMisesStressField] = FEA_StressAnalysis()"""
        nn = self.info[0]["numNodes"]
        st = np.zeros(6 * nn)
        vm = np.zeros(nn)
        lib().or_mesh_stress(self._h, f64(U), st, vm)
        return st.reshape(nn, 6, order="F"), vm

    def pde_filter(self, src, rHatMin=8.0, tol=1e-3, max_it=500):
        """tar = TopOpti_PDEFiltering_matrixFree(src); returns (tar, its)."""
        tar = np.zeros(self.info[0]["numElements"])
        its = lib().or_mesh_pde_filter(self._h, rHatMin, f64(src), tol, max_it, tar)
        return tar, its

    def time_pcg_iterations(self, b, omega, its):
        """(seconds, K.u seconds, V-cycle seconds) of `its` FP64 PCG iterations."""
        tk = C.c_double(0)
        tv = C.c_double(0)
        t = lib().or_mesh_time_pcg_iterations(self._h, omega, f64(b), its, C.byref(tk), C.byref(tv))
        return t, tk.value, tv.value


def set_threads(n):
    """OpenMP team size of the oracle's following calls; returns the previous maximum."""
    return lib().or_set_threads(int(n))
