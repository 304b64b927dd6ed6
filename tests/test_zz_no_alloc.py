"""The allocation-free gradient entry point (fdf_mlp_grad_noalloc) against
fdf_mlp_grad, through the C ABI, on device buffers the caller owns: two
calls in a row, the second with a smaller batch, must both reproduce the
allocating entry point's gradient (same kernels and fixed-order reductions:
bit-identical) and its loss (within the rounding of an atomic sum); batches
beyond the fixed scratch are rejected with an error.  The caller's buffers
come from the HIP runtime the library itself is linked to."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime already mapped into this process by the library."""
    from fdfault_amd import lib
    lib.load()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln and ln.split()[-1].startswith("/")}
    assert paths, "the HIP runtime is not loaded"
    hip = C.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, count):
        out = np.empty(count, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    P = np.zeros(4555, np.float32)
    P[:4545] = rng.normal(0., 0.2, 4545)
    X = rng.normal(size=(n, 4)).astype(np.float32)
    T = rng.normal(size=n).astype(np.float32)
    return P, X, T


def test_noalloc_gradient_matches_allocating_entry_point():
    from fdfault_amd import lib
    hip = _hip()
    nmax = 3000
    dP, dX, dT = DevBuf(hip, 4555 * 4), DevBuf(hip, nmax * 16), DevBuf(hip, nmax * 4)
    dG, dL = DevBuf(hip, 4545 * 4), DevBuf(hip, 4)
    try:
        for seed, n in ((11, 3000), (12, 700)):  # the second call is smaller
            P, X, T = _data(seed, n)
            Gref, lref = lib.mlp_grad(P, X, T)
            dP.put(P)
            dX.put(X)
            dT.put(T)
            dG.put(np.full(4545, np.nan, np.float32))
            dL.put(np.full(1, np.nan, np.float32))
            print("n", n, flush=True)
            lib.mlp_grad_noalloc(0, dP.p.value, dX.p.value, dT.p.value, n, dG.p.value, dL.p.value)
            G = dG.get(4545)  # (a blocking copy on the null stream, after the kernels)
            assert np.array_equal(G, Gref), n
            # the loss is a sum of per-block partials added with float
            # atomics, whose order varies from run to run in both entry points
            loss = float(dL.get(1)[0])
            assert abs(loss - lref) <= 1e-6 * abs(lref), (n, loss, lref)
        with pytest.raises(lib.FdfaultError, match="outside 1..4096"):
            lib.mlp_grad_noalloc(0, dP.p.value, dX.p.value, dT.p.value, 5000, dG.p.value, None)
    finally:
        for b in (dP, dX, dT, dG, dL):
            b.free()
