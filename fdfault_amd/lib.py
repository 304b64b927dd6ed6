"""ctypes mirror of the C ABI in include/fdfault_hip.h.

Mirrors the reference's driver objects: :class:`Solver` plays
``problem`` (src/problem.cpp) -- construct from a ``.in`` file, ``step()``
runs ``problem::solve`` iterations, ``stage()`` one ``domain::do_rk_stage``;
arrays are read back like the reference's friend-access output code does.
Errors raise :class:`FdfaultError` with the reference's message text.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libfdfault_hip.so")
EXE_PATH = os.path.join(HERE, "_lib", "fdfault")

FP64, FP32 = 0, 1
IFACE_FIELDS = {"U": 0, "V": 1, "S": 2, "Sn": 3, "Ux": 4, "Vx": 5, "Sx": 6, "state": 7,
                "du": 8, "dux": 9, "dstate": 10}
ARRAYS = {"f": 0, "df": 1, "x": 2, "metric": 3, "jac": 4}


class FdfaultError(RuntimeError):
    pass


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("device", C.c_int), ("write_outputs", C.c_int),
                ("quiet", C.c_int), ("use_graph", C.c_int), ("exact_fp64", C.c_int)]


class Info(C.Structure):
    _fields_ = [("nt", C.c_int), ("nstages", C.c_int), ("ndim", C.c_int), ("mode", C.c_int),
                ("nx", C.c_int * 3), ("nfields", C.c_int), ("nfields_total", C.c_int),
                ("nifaces", C.c_int), ("sbporder", C.c_int), ("plastic", C.c_int), ("nblocks", C.c_int),
                ("step", C.c_int), ("precision", C.c_int),
                ("dt", C.c_double), ("ttot", C.c_double), ("cfl", C.c_double), ("t", C.c_double),
                ("tiled_blocks", C.c_int), ("row_blocks", C.c_int), ("curv_blocks", C.c_int),
                ("wide_blocks", C.c_int), ("fused_faults", C.c_int)]


EXPORTS = ["fdf_open", "fdf_step", "fdf_stage", "fdf_run", "fdf_read_field", "fdf_read_array",
           "fdf_write_array", "fdf_save_state", "fdf_load_state", "fdf_read_iface", "fdf_read_front", "fdf_finish", "fdf_get_info",
           "fdf_host_setup", "fdf_last_error", "fdf_close", "fdf_bulk_kernel_ms", "fdf_fault_kernel_ms",
           "fdf_open_batch", "fdf_batch_step", "fdf_batch_size", "fdf_batch_get_info", "fdf_batch_read_iface",
           "fdf_batch_read_front", "fdf_batch_read_array", "fdf_batch_tuning", "fdf_batch_persistent", "fdf_batch_finish",
           "fdf_batch_close", "fdf_mlp_param_count", "fdf_mlp_train", "fdf_mlp_grad", "fdf_mlp_grad_noalloc", "fdf_mlp_predict", "fdf_mlp_predict_mfma",
           "fdf_stz_set_model", "fdf_stz_log", "fdf_stz_log_read", "fdf_stz_stats"]

_lib = None


def load():
    """Load the HIP library; raises if it was not built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FdfaultError("fdfault_amd HIP library not built: %s (run fdfault_amd/build.py)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.fdf_open.argtypes = [C.c_char_p, C.POINTER(Opts), C.POINTER(C.c_void_p)]
    L.fdf_step.argtypes = [C.c_void_p, C.c_int]
    L.fdf_stage.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.fdf_run.argtypes = [C.c_char_p, C.POINTER(Opts)]
    L.fdf_read_field.argtypes = [C.c_void_p, C.c_int, ip, ip, ip, dp]
    L.fdf_read_array.argtypes = [C.c_void_p, C.c_int, dp, C.c_long]
    L.fdf_write_array.argtypes = [C.c_void_p, C.c_int, dp, C.c_long]
    L.fdf_save_state.argtypes = [C.c_void_p, C.c_char_p]
    L.fdf_load_state.argtypes = [C.c_void_p, C.c_char_p]
    L.fdf_read_iface.restype = C.c_long
    L.fdf_read_iface.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_long]
    L.fdf_read_front.restype = C.c_long
    L.fdf_read_front.argtypes = [C.c_void_p, C.c_int, dp, C.c_long]
    L.fdf_finish.argtypes = [C.c_void_p]
    L.fdf_get_info.argtypes = [C.c_void_p, C.POINTER(Info)]
    L.fdf_host_setup.argtypes = [C.c_char_p, C.POINTER(Info), dp, dp, dp]
    L.fdf_last_error.restype = C.c_char_p
    L.fdf_close.argtypes = [C.c_void_p]
    L.fdf_bulk_kernel_ms.restype = C.c_double
    L.fdf_bulk_kernel_ms.argtypes = [C.c_void_p, ip]
    L.fdf_fault_kernel_ms.restype = C.c_double
    L.fdf_fault_kernel_ms.argtypes = [C.c_void_p, ip]
    L.fdf_open_batch.argtypes = [C.POINTER(C.c_char_p), C.c_int, C.POINTER(Opts), C.POINTER(C.c_void_p)]
    L.fdf_batch_step.argtypes = [C.c_void_p, C.c_int]
    L.fdf_batch_size.argtypes = [C.c_void_p]
    L.fdf_batch_get_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(Info)]
    L.fdf_batch_read_iface.restype = C.c_long
    L.fdf_batch_read_iface.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp, C.c_long]
    L.fdf_batch_read_front.restype = C.c_long
    L.fdf_batch_read_front.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_long]
    L.fdf_batch_read_array.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, C.c_long]
    L.fdf_batch_tuning.argtypes = [C.c_void_p, ip, ip, dp, ip]
    L.fdf_batch_persistent.argtypes = [C.c_void_p]
    L.fdf_batch_finish.argtypes = [C.c_void_p]
    L.fdf_batch_close.argtypes = [C.c_void_p]
    fp = C.POINTER(C.c_float)
    L.fdf_mlp_train.argtypes = [dp, C.c_long, C.c_int, C.c_int, C.c_float, C.c_uint, fp, fp]
    L.fdf_mlp_grad.argtypes = [fp, fp, fp, C.c_long, fp, fp]
    vp = C.c_void_p
    L.fdf_mlp_grad_noalloc.argtypes = [C.c_int, vp, vp, vp, C.c_long, vp, vp, vp]
    L.fdf_mlp_predict.argtypes = [fp, fp, C.c_long, fp]
    L.fdf_mlp_predict_mfma.argtypes = [fp, fp, C.c_long, fp]
    L.fdf_stz_set_model.argtypes = [C.c_void_p, fp, C.c_int]
    L.fdf_stz_log.argtypes = [C.c_void_p, C.c_long]
    L.fdf_stz_log_read.restype = C.c_long
    L.fdf_stz_log_read.argtypes = [C.c_void_p, dp, C.c_long]
    L.fdf_stz_stats.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong)]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _err():
    return FdfaultError(load().fdf_last_error().decode())


def host_setup(inpath: str):
    """Grid, metric, Jacobian and dt built by the library's host problem
    compiler (no GPU needed)."""
    L = load()
    info = Info()
    if L.fdf_host_setup(inpath.encode(), C.byref(info), None, None, None) != 0:
        raise _err()
    N = info.nx[0] * info.nx[1] * info.nx[2]
    x = np.empty(info.ndim * N)
    m = np.empty(info.ndim * info.ndim * N)
    j = np.empty(N)
    if L.fdf_host_setup(inpath.encode(), None, _dp(x), _dp(m), _dp(j)) != 0:
        raise _err()
    shp = tuple(info.nx)
    return info, x.reshape((info.ndim,) + shp), m.reshape((info.ndim * info.ndim,) + shp), j.reshape(shp)


class Solver:
    """One problem on the GPU (mirror of the reference's ``problem`` class)."""

    def __init__(self, inpath: str, precision: int = FP64, device: int = 0, write_outputs: bool = False,
                 quiet: bool = True, exact: bool = False, graph: bool = False):
        L = load()
        self.opts = Opts(precision, device, 1 if write_outputs else 0, 1 if quiet else 0, 1 if graph else 0,
                         1 if exact else 0)
        h = C.c_void_p()
        if L.fdf_open(inpath.encode(), C.byref(self.opts), C.byref(h)) != 0:
            raise _err()
        self.h = h
        self.info = Info()
        L.fdf_get_info(self.h, C.byref(self.info))
        self.nx = tuple(self.info.nx)
        self.N = self.nx[0] * self.nx[1] * self.nx[2]
        self.ndim = self.info.ndim
        self.nt = self.info.nt
        self.dt = self.info.dt
        self.nstages = self.info.nstages

    def close(self):
        if getattr(self, "h", None):
            load().fdf_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n: int = 1):
        if load().fdf_step(self.h, n) != 0:
            raise _err()

    def stage(self, istep: int, stage: int):
        if load().fdf_stage(self.h, istep, stage) != 0:
            raise _err()

    def array(self, which: str):
        i = self.info
        n = {"f": i.nfields_total, "df": i.nfields, "x": i.ndim, "metric": i.ndim * i.ndim, "jac": 1}[which]
        out = np.empty(n * self.N)
        if load().fdf_read_array(self.h, ARRAYS[which], _dp(out), out.size) != 0:
            raise _err()
        return out.reshape((n,) + self.nx)

    def set_array(self, which: str, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if load().fdf_write_array(self.h, ARRAYS[which], _dp(a), a.size) != 0:
            raise _err()

    def save_state(self, path: str):
        """Snapshot of the evolving state after the last completed step
        (fdf_save_state)."""
        if load().fdf_save_state(self.h, path.encode()) != 0:
            raise _err()

    def load_state(self, path: str):
        """Resume from a snapshot of the same problem and precision
        (fdf_load_state)."""
        if load().fdf_load_state(self.h, path.encode()) != 0:
            raise _err()

    def field(self, q: int, lo, hi, st=(1, 1, 1)):
        lo = (C.c_int * 3)(*lo)
        hi_ = (C.c_int * 3)(*hi)
        st = (C.c_int * 3)(*st)
        n = [(hi_[d] - lo[d]) // st[d] + 1 for d in range(3)]
        out = np.empty(n[0] * n[1] * n[2])
        if load().fdf_read_field(self.h, q, lo, hi_, st, _dp(out)) != 0:
            raise _err()
        return out.reshape(n)

    def iface(self, i: int, which: str):
        L = load()
        n = L.fdf_read_iface(self.h, i, IFACE_FIELDS[which], None, 0)
        if n < 0:
            raise _err()
        out = np.empty(n)
        if L.fdf_read_iface(self.h, i, IFACE_FIELDS[which], _dp(out), n) < 0:
            raise _err()
        return out

    def front(self, i: int):
        L = load()
        n = L.fdf_read_front(self.h, i, None, 0)
        if n < 0:
            return None
        out = np.empty(n)
        L.fdf_read_front(self.h, i, _dp(out), n)
        return out

    def finish(self):
        if load().fdf_finish(self.h) != 0:
            raise _err()

    def stz_set_model(self, params):
        L = load()
        if params is None:
            rc = L.fdf_stz_set_model(self.h, None, 0)
        else:
            p = np.ascontiguousarray(params, np.float32)
            rc = L.fdf_stz_set_model(self.h, _fp(p), p.size)
        if rc != 0:
            raise _err()

    def stz_log(self, capacity):
        if load().fdf_stz_log(self.h, int(capacity)) != 0:
            raise _err()

    def stz_log_read(self):
        L = load()
        n = L.fdf_stz_log_read(self.h, None, 0)
        if n < 0:
            raise _err()
        out = np.empty((n, 5))
        if n and L.fdf_stz_log_read(self.h, _dp(out), n) < 0:
            raise _err()
        return out

    def stz_stats(self):
        """(solves, Newton iterations, proposals used, fallbacks) since the last call."""
        out = (C.c_ulonglong * 4)()
        if load().fdf_stz_stats(self.h, out) != 0:
            raise _err()
        return tuple(int(v) for v in out)

    def fault_kernel_ms(self):
        n = C.c_int()
        ms = load().fdf_fault_kernel_ms(self.h, C.byref(n))
        return ms, n.value

    def bulk_kernel_ms(self):
        n = C.c_int(0)
        ms = load().fdf_bulk_kernel_ms(self.h, C.byref(n))
        return ms, n.value


def run(inpath: str, precision: int = FP64, quiet: bool = True, exact: bool = False):
    o = Opts(precision, 0, 1, 1 if quiet else 0, 0, 1 if exact else 0)
    if load().fdf_run(inpath.encode(), C.byref(o)) != 0:
        raise _err()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def mlp_param_count():
    return load().fdf_mlp_param_count()


def mlp_train(rows, epochs=20, batch=1024, lr=3e-3, seed=0):
    """Train the STZ proposal network on the GPU from log rows
    (phi, sn, chi, eta, log V); returns (params, per-epoch mean loss)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    P = np.zeros(mlp_param_count(), np.float32)
    losses = np.zeros(epochs, np.float32)
    if load().fdf_mlp_train(_dp(rows), rows.shape[0], epochs, batch, lr, seed, _fp(P), _fp(losses)) != 0:
        raise _err()
    return P, losses


def mlp_grad(P, X, T):
    """GPU gradient and mean loss of one batch of normalised samples."""
    P = np.ascontiguousarray(P, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    G = np.zeros(4545, np.float32)
    loss = np.zeros(1, np.float32)
    if load().fdf_mlp_grad(_fp(P), _fp(X), _fp(T), X.shape[0], _fp(G), _fp(loss)) != 0:
        raise _err()
    return G, float(loss[0])


def mlp_grad_noalloc(device, params_ptr, X_ptr, T_ptr, n, grad_ptr, loss_ptr=None, stream=None):
    """fdf_mlp_grad_noalloc on device buffers given by address (params 4555,
    X n x 4, T n, grad 4545, loss 1 floats); queued on `stream` (a
    hipStream_t address, None for the null stream)."""
    if load().fdf_mlp_grad_noalloc(device, params_ptr, X_ptr, T_ptr, n, grad_ptr, loss_ptr, stream) != 0:
        raise _err()


def mlp_predict(P, X, mfma=False):
    """Normalised outputs for normalised inputs; mfma=True uses the fault
    kernel's evaluation (matrix cores, FP16 operands)."""
    P = np.ascontiguousarray(P, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.zeros(X.shape[0], np.float32)
    fn = load().fdf_mlp_predict_mfma if mfma else load().fdf_mlp_predict
    if fn(_fp(P), _fp(X), X.shape[0], _fp(Y)) != 0:
        raise _err()
    return Y


class Batch:
    """An ensemble of independent 2-D mode II problems stepped together
    (one fault launch and one persistent bulk launch per RK stage for all
    instances).  Per-instance readback mirrors :class:`Solver`."""

    def __init__(self, inpaths, precision: int = FP32, device: int = 0, write_outputs: bool = False,
                 quiet: bool = True, graph: bool = True):
        L = load()
        self.opts = Opts(precision, device, 1 if write_outputs else 0, 1 if quiet else 0, 1 if graph else 0, 0)
        arr = (C.c_char_p * len(inpaths))(*[p.encode() for p in inpaths])
        h = C.c_void_p()
        if L.fdf_open_batch(arr, len(inpaths), C.byref(self.opts), C.byref(h)) != 0:
            raise _err()
        self.h = h
        self.n = L.fdf_batch_size(self.h)
        self.infos = []
        for i in range(self.n):
            info = Info()
            L.fdf_batch_get_info(self.h, i, C.byref(info))
            self.infos.append(info)
        self.dt = self.infos[0].dt
        self.nstages = self.infos[0].nstages

    def close(self):
        if getattr(self, "h", None):
            load().fdf_batch_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        if load().fdf_batch_step(self.h, n) != 0:
            raise _err()

    def points(self, i: int):
        nx = self.infos[i].nx
        return nx[0] * nx[1] * nx[2]

    def tuning(self):
        tw, cx, fc = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        load().fdf_batch_tuning(self.h, C.byref(tw), C.byref(cx), C.byref(ms), C.byref(fc))
        return {"tile_width": tw.value, "chunk": cx.value, "stage_ms": ms.value, "from_cache": bool(fc.value),
                "persistent": bool(load().fdf_batch_persistent(self.h))}

    def iface(self, inst: int, i: int, which: str):
        L = load()
        n = L.fdf_batch_read_iface(self.h, inst, i, IFACE_FIELDS[which], None, 0)
        if n < 0:
            raise _err()
        out = np.empty(n)
        if L.fdf_batch_read_iface(self.h, inst, i, IFACE_FIELDS[which], _dp(out), n) < 0:
            raise _err()
        return out

    def front(self, inst: int, i: int):
        L = load()
        n = L.fdf_batch_read_front(self.h, inst, i, None, 0)
        if n < 0:
            return None
        out = np.empty(n)
        L.fdf_batch_read_front(self.h, inst, i, _dp(out), n)
        return out

    def array(self, inst: int, which: str):
        i = self.infos[inst]
        n = {"f": i.nfields_total, "df": i.nfields, "x": i.ndim, "metric": i.ndim * i.ndim, "jac": 1}[which]
        N = self.points(inst)
        out = np.empty(n * N)
        if load().fdf_batch_read_array(self.h, inst, ARRAYS[which], _dp(out), out.size) != 0:
            raise _err()
        return out.reshape((n,) + tuple(i.nx))

    def finish(self):
        if load().fdf_batch_finish(self.h) != 0:
            raise _err()
