"""Order of accuracy of the GPU bulk kernels on the exact plane S wave of
tests/planewave.py (the oracle's own orders are pinned in
test_oracle_convergence.py): the 2-D tiled kernel and the 3-D row-tile
kernel in FP64 (interior orders 2, 4, 6 for sbporder 2, 3, 4), the
bit-faithful FP64 mode, and the FP32 bulk path where the errors stay well
above FP32 rounding."""
import numpy as np
import pytest

import planewave as PW
from fdfault_amd import lib

pytestmark = pytest.mark.gpu


def _open(precision, exact=False):
    return lambda p: lib.Solver(p, precision=precision, exact=exact)


@pytest.mark.parametrize("s", [2, 3, 4])
@pytest.mark.parametrize("mode", ["fp64", "exact"])
def test_gpu_interior_order_2d(tmp_path, s, mode):
    op = _open(lib.FP64, exact=(mode == "exact"))
    dxs = (0.05, 0.025) if s < 4 else (0.1, 0.05)
    e = [PW.error(op, str(tmp_path / ("g%d" % k)), dx, s, T=0.25, Lx=6.0, Ly=6.0) for k, dx in enumerate(dxs)]
    p = PW.order(*e)
    print("2-D", mode, "sbporder", s, "errors", e, "order %.2f" % p, flush=True)
    assert p >= 2 * (s - 1) - 0.3


@pytest.mark.parametrize("s", [2, 3])
def test_gpu_interior_order_2d_fp32(tmp_path, s):
    e = [PW.error(_open(lib.FP32), str(tmp_path / ("g%d" % k)), dx, s, T=0.25, Lx=6.0, Ly=6.0)
         for k, dx in enumerate((0.1, 0.05))]
    p = PW.order(*e)
    print("2-D fp32 sbporder", s, "errors", e, "order %.2f" % p, flush=True)
    assert p >= 2 * (s - 1) - 0.3


@pytest.mark.parametrize("s,prec", [(2, "fp64"), (3, "fp64"), (3, "fp32")])
def test_gpu_interior_order_3d(tmp_path, s, prec):
    op = _open(lib.FP64 if prec == "fp64" else lib.FP32)
    e = [PW.error(op, str(tmp_path / ("g%d" % k)), dx, s, T=0.25, Lx=6.0, Ly=6.0, ndim=3, Lz=9.0)
         for k, dx in enumerate((0.2, 0.1))]
    p = PW.order(*e)
    print("3-D", prec, "sbporder", s, "errors", e, "order %.2f" % p, flush=True)
    assert p >= 2 * (s - 1) - 0.3


def test_gpu_boundary_order_2d(tmp_path):
    """The pulse leaves through the absorbing x-plus face (sbporder 3): the
    boundary closure and SAT keep the SBP-SAT order s = 3."""
    e = [PW.error(_open(lib.FP64), str(tmp_path / ("g%d" % k)), dx, 3, T=0.95, Lx=6.0, Ly=10.0)
         for k, dx in enumerate((0.1, 0.05))]
    p = PW.order(*e)
    print("2-D boundary sbporder 3 errors", e, "order %.2f" % p, flush=True)
    assert p >= 2.5


@pytest.mark.parametrize("s,ndim", [(3, 2), (4, 2), (3, 3)])
def test_gpu_interior_order_varying_lambda(tmp_path, s, ndim):
    """Per-point material (lambda varying in x and y, uniform rho and G; the
    generic kernel's per-point branch) keeps the interior order of the plane
    S wave."""
    kw = dict(ndim=3, Lz=9.0) if ndim == 3 else {}
    dxs = (0.2, 0.1) if ndim == 3 else (0.1, 0.05)
    e = [PW.error(_open(lib.FP64), str(tmp_path / ("g%d" % k)), dx, s, T=0.25, Lx=6.0, Ly=6.0, lam_var=0.5, **kw)
         for k, dx in enumerate(dxs)]
    p = PW.order(*e)
    print("%d-D varying lambda sbporder" % ndim, s, "errors", e, "order %.2f" % p, flush=True)
    assert p >= 2 * (s - 1) - 0.3
