"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared with;
nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.environ.get("FDF_ORACLE_LIB") or os.path.join(ORACLE_DIR, "_build", "liboracle.so")
EXE_PATH = os.path.join(ORACLE_DIR, "_build", "fdfault_oracle")

_lib = None

IFACE_FIELDS = {"U": 0, "V": 1, "S": 2, "Sn": 3, "Ux": 4, "Vx": 5, "Sx": 6, "state": 7,
                "du": 8, "dux": 9, "dstate": 10}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_open.restype = C.c_void_p
        L.orc_open.argtypes = [C.c_char_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_close.argtypes = [C.c_void_p]
        L.orc_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_get_array.argtypes = [C.c_void_p, C.c_int, dp, C.c_long]
        L.orc_set_array.argtypes = [C.c_void_p, C.c_int, dp, C.c_long]
        L.orc_iface_get.restype = C.c_long
        L.orc_iface_get.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_long]
        L.orc_iface_shape.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.orc_front_get.restype = C.c_long
        L.orc_front_get.argtypes = [C.c_void_p, C.c_int, dp, C.c_long]
        L.orc_finish.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_char_p, C.c_int]
        L.orc_fd_coeffs.argtypes = [C.c_int, dp, dp, dp]
        L.orc_rk_coeffs.argtypes = [C.c_int, C.POINTER(C.c_int), dp, dp, dp]
        L.orc_stz_mu.restype = C.c_double
        L.orc_stz_mu.argtypes = [dp, C.POINTER(C.c_int)]
        L.orc_stz_func.restype = C.c_double
        L.orc_stz_func.argtypes = [C.c_double, dp]
        L.orc_plastic_flow.argtypes = [C.c_double, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, dp]
        L.orc_sw_mu.restype = C.c_double
        L.orc_sw_mu.argtypes = [C.c_double] * 7
        L.orc_stz_vpl.restype = C.c_double
        L.orc_stz_vpl.argtypes = [C.c_double, dp]
        L.orc_stz_dstatedt.restype = C.c_double
        L.orc_stz_dstatedt.argtypes = [C.c_double] * 8 + [C.POINTER(C.c_int)]
        L.orc_sw_solve.argtypes = [C.c_double] * 5 + [dp]
        L.orc_solve_locked.argtypes = [dp, C.c_double, C.c_double]
        L.orc_rotate.argtypes = [dp, dp, dp, dp, C.c_int, dp]
        L.orc_mlp_forward.argtypes = [dp, dp, C.c_long, dp]
        L.orc_mlp_grad.restype = C.c_double
        L.orc_mlp_grad.argtypes = [dp, dp, dp, C.c_long, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    pass


class Oracle:
    """One problem held by the CPU oracle."""

    def __init__(self, inpath: str):
        L = lib()
        self.h = L.orc_open(inpath.encode())
        if not self.h:
            raise OracleError(L.orc_last_error().decode())
        ints = (C.c_int * 16)()
        dbls = (C.c_double * 4)()
        L.orc_info(self.h, ints, dbls)
        self.nt, self.nstages, self.ndim, self.mode = ints[0], ints[1], ints[2], ints[3]
        self.nx = (ints[4], ints[5], ints[6])
        self.nf, self.nftot, self.nifaces = ints[7], ints[8], ints[9]
        self.dt, self.ttot, self.cfl = dbls[0], dbls[1], dbls[2]
        self.N = self.nx[0] * self.nx[1] * self.nx[2]

    def close(self):
        if self.h:
            lib().orc_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1, outputs=False):
        if lib().orc_step(self.h, n, 1 if outputs else 0) != 0:
            raise OracleError(lib().orc_last_error().decode())

    def stage(self, istep, stage):
        if lib().orc_stage(self.h, istep, stage) != 0:
            raise OracleError(lib().orc_last_error().decode())

    def array(self, which: str):
        w = {"f": 0, "df": 1, "x": 2, "metric": 3, "jac": 4}[which]
        n = {"f": self.nftot, "df": self.nf, "x": self.ndim, "metric": self.ndim * self.ndim, "jac": 1}[which]
        out = np.empty(n * self.N)
        if lib().orc_get_array(self.h, w, _dp(out), out.size) != 0:
            raise OracleError("bad array")
        return out.reshape((n,) + tuple(self.nx))

    def set_array(self, which: str, arr):
        w = {"f": 0, "df": 1, "x": 2, "metric": 3, "jac": 4}[which]
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if lib().orc_set_array(self.h, w, _dp(a), a.size) != 0:
            raise OracleError("bad array")

    def iface(self, i: int, which: str):
        w = IFACE_FIELDS[which]
        n = lib().orc_iface_get(self.h, i, w, None, 0)
        if n < 0:
            raise OracleError("bad interface field")
        out = np.empty(n)
        lib().orc_iface_get(self.h, i, w, _dp(out), n)
        return out

    def iface_shape(self, i: int):
        s = (C.c_int * 4)()
        lib().orc_iface_shape(self.h, i, s)
        return tuple(s)

    def front(self, i: int):
        n = lib().orc_front_get(self.h, i, None, 0)
        if n < 0:
            return None
        out = np.empty(n)
        lib().orc_front_get(self.h, i, _dp(out), n)
        return out

    def finish(self):
        if lib().orc_finish(self.h) != 0:
            raise OracleError(lib().orc_last_error().decode())


def run(inpath: str, quiet=True) -> int:
    return lib().orc_run(inpath.encode(), 1 if quiet else 0)


def mlp_forward(P, X):
    """Oracle forward of the proposal network (double precision)."""
    P = np.ascontiguousarray(P, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.empty(X.shape[0])
    lib().orc_mlp_forward(_dp(P), _dp(X), X.shape[0], _dp(Y))
    return Y


def mlp_grad(P, X, T):
    """Oracle gradient (4545) and mean loss of one batch (double precision)."""
    P = np.ascontiguousarray(P, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    G = np.empty(4545)
    loss = lib().orc_mlp_grad(_dp(P), _dp(X), _dp(T), X.shape[0], _dp(G))
    return G, loss
