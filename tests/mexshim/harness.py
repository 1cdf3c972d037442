"""Builds this repository's MEX gateways (sgldbench_amd/mex/*.cpp) against the
test stand-in of the MEX API (tests/mexshim/mex.h) and calls them from
Python.  Test infrastructure only.

Two builds: "cpu" links the gateways to tests/mexshim/abi_oracle.c (the C ABI
over the CPU oracle), "gpu" to sgldbench_amd/libsgldfea.so.
"""
import ctypes as C
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GATEWAYS = sorted(glob.glob(os.path.join(ROOT, "sgldbench_amd", "mex", "*.cpp")))

mxCELL, mxLOGICAL, mxCHAR, mxDOUBLE, mxUINT8, mxINT32, mxUINT64 = 1, 3, 4, 6, 9, 12, 15
_NP = {np.dtype(np.float64): mxDOUBLE, np.dtype(np.int32): mxINT32, np.dtype(np.uint8): mxUINT8,
       np.dtype(np.bool_): mxLOGICAL, np.dtype(np.uint64): mxUINT64}
_CLS = {mxDOUBLE: np.float64, mxINT32: np.int32, mxUINT8: np.uint8, mxLOGICAL: np.bool_, mxUINT64: np.uint64}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(" ".join(cmd) + "\n" + r.stdout + r.stderr)


def build(variant):
    """Compiles the shim, the C-ABI provider of `variant` and every gateway
    into tests/mexshim/_build/<variant>/; returns that directory."""
    out = os.path.join(HERE, "_build", variant)
    os.makedirs(out, exist_ok=True)
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", os.path.join(out, "libmexshim.so"),
          os.path.join(HERE, "mexshim.cpp")])
    if variant == "cpu":
        odir = os.path.join(ROOT, "oracle")
        _run(["make", "-s", "-C", odir])
        _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", os.path.join(out, "libabi_oracle.so"),
              os.path.join(HERE, "abi_oracle.c"), "-L" + odir, "-lsgld_oracle", "-Wl,-rpath," + odir, "-lm"])
        abi = ["-L" + out, "-labi_oracle"]
    else:
        ldir = os.path.join(ROOT, "sgldbench_amd")
        abi = ["-L" + ldir, "-l:libsgldfea.so", "-Wl,-rpath," + ldir]
    for src in GATEWAYS:
        name = os.path.splitext(os.path.basename(src))[0]
        _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I" + HERE, "-I" + os.path.join(ROOT, "include"),
              "-o", os.path.join(out, name + ".so"), src, "-L" + out, "-lmexshim"] + abi +
             ["-Wl,-rpath," + out])
    return out


class MexError(Exception):
    def __init__(self, ident, msg):
        super().__init__("%s: %s" % (ident, msg))
        self.identifier = ident


class Sparse:
    """A sparse double argument (only its class is inspected by the gateways)."""

    def __init__(self, m, n):
        self.shape = (m, n)


class Gateways:
    def __init__(self, variant):
        d = os.path.join(HERE, "_build", variant)
        if not os.path.exists(os.path.join(d, "sgld_ctx_mex.so")):
            build(variant)
        self.shim = C.CDLL(os.path.join(d, "libmexshim.so"), mode=C.RTLD_GLOBAL)
        s = self.shim
        s.shim_array.restype = C.c_void_p
        s.shim_array.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_void_p]
        s.shim_sparse_flagged.restype = C.c_void_p
        s.shim_sparse_flagged.argtypes = [C.c_int64, C.c_int64]
        s.shim_cell.restype = C.c_void_p
        s.shim_cell.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        s.shim_string.restype = C.c_void_p
        s.shim_string.argtypes = [C.c_char_p]
        s.shim_free.argtypes = [C.c_void_p]
        s.shim_class.argtypes = [C.c_void_p]
        s.shim_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        s.shim_data.restype = C.c_void_p
        s.shim_data.argtypes = [C.c_void_p]
        s.shim_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p),
                                C.c_char_p, C.c_char_p, C.c_int]
        s.shim_last_warning.restype = C.c_char_p
        self._libs = {}
        self._dir = d

    def _fn(self, name):
        if name not in self._libs:
            self._libs[name] = C.CDLL(os.path.join(self._dir, name + ".so"))
        return C.cast(self._libs[name].mexFunction, C.c_void_p)

    def _to_mx(self, v):
        s = self.shim
        if isinstance(v, str):
            return s.shim_string(v.encode())
        if isinstance(v, Sparse):
            return s.shim_sparse_flagged(*v.shape)
        if isinstance(v, (list, tuple)):
            elems = (C.c_void_p * max(1, len(v)))(*[self._to_mx(e) for e in v])
            return s.shim_cell(len(v), elems)
        a = np.asarray(v)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        a = np.asfortranarray(a)
        dims = (C.c_int64 * a.ndim)(*a.shape)
        buf = np.ascontiguousarray(a.ravel(order="F"))
        return s.shim_array(_NP[a.dtype], a.ndim, dims, buf.ctypes.data_as(C.c_void_p))

    def _from_mx(self, p):
        s = self.shim
        dims = (C.c_int64 * 8)()
        nd = s.shim_dims(p, dims)
        shape = tuple(dims[k] for k in range(nd))
        dt = _CLS[s.shim_class(p)]
        n = int(np.prod(shape))
        buf = np.ctypeslib.as_array((C.c_uint8 * (n * np.dtype(dt).itemsize)).from_address(s.shim_data(p))) \
            if n else np.zeros(0, np.uint8)
        return buf.view(dt).reshape(shape, order="F").copy()

    def call(self, name, nlhs, *args):
        s = self.shim
        prhs = (C.c_void_p * max(1, len(args)))(*[self._to_mx(a) for a in args])
        plhs = (C.c_void_p * max(1, nlhs))()
        ident, msg = C.create_string_buffer(512), C.create_string_buffer(512)
        rc = s.shim_call(self._fn(name), nlhs, plhs, len(args), prhs, ident, msg, 512)
        for k in range(len(args)):
            s.shim_free(prhs[k])
        if rc:
            raise MexError(ident.value.decode(), msg.value.decode())
        out = [self._from_mx(plhs[k]) for k in range(nlhs) if plhs[k]]
        for k in range(nlhs):
            if plhs[k]:
                s.shim_free(plhs[k])
        return out

    def last_warning(self):
        return self.shim.shim_last_warning().decode()
