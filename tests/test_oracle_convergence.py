"""Order of accuracy of the oracle's bulk scheme on an exact plane S wave
(tests/planewave.py), the order-of-accuracy check the reference's
manufactured-solution hook (block.cpp:2052-2267) was written for.

* Interior: the pulse stays inside the blocks, so only the interior stencil
  of block.cpp:1048-1278 (fd.cpp:33-155, 2(s-1)-th order) and the RK4 update
  act on it; with dt proportional to dx the error falls as dx^(2(s-1)):
  2, 4, 6 for sbporder 2, 3, 4.
* Boundary: the pulse leaves through the absorbing x-plus face, exercising
  the boundary closures (order s - 1) and the SAT of boundary.cpp:265-503;
  an SBP-SAT scheme then converges at order s.
A wrong interior coefficient, a wrong boundary row or a wrong penalty breaks
the order.
"""
import numpy as np
import pytest

import oracle_lib as O
import planewave as PW


@pytest.mark.parametrize("s", [2, 3, 4])
def test_interior_order(tmp_path, s):
    e = [PW.error(O.Oracle, str(tmp_path / ("g%d" % k)), dx, s, T=0.25, Lx=6.0, Ly=6.0)
         for k, dx in enumerate((0.1, 0.05))]
    p = PW.order(*e)
    print("sbporder", s, "errors", e, "order %.2f" % p)
    assert p >= 2 * (s - 1) - 0.3


@pytest.mark.parametrize("s", [2, 3, 4])
def test_boundary_order(tmp_path, s):
    e = [PW.error(O.Oracle, str(tmp_path / ("g%d" % k)), dx, s, T=0.95, Lx=6.0, Ly=10.0)
         for k, dx in enumerate((0.1, 0.05))]
    p = PW.order(*e)
    print("sbporder", s, "errors", e, "order %.2f" % p)
    assert p >= s - 0.5


@pytest.mark.parametrize("s", [2, 3, 4])
def test_interior_order_varying_lambda(tmp_path, s):
    """Per-point material (lambda varying in x and y, uniform rho and G): the
    plane S wave is unchanged, and the per-point branch keeps the order."""
    e = [PW.error(O.Oracle, str(tmp_path / ("g%d" % k)), dx, s, T=0.25, Lx=6.0, Ly=6.0, lam_var=0.5)
         for k, dx in enumerate((0.1, 0.05))]
    p = PW.order(*e)
    print("sbporder", s, "varying lambda, errors", e, "order %.2f" % p)
    assert p >= 2 * (s - 1) - 0.3


def test_boundary_order_varying_lambda(tmp_path):
    e = [PW.error(O.Oracle, str(tmp_path / ("g%d" % k)), dx, 3, T=0.95, Lx=6.0, Ly=10.0, lam_var=0.5)
         for k, dx in enumerate((0.1, 0.05))]
    p = PW.order(*e)
    print("sbporder 3 varying lambda, boundary errors", e, "order %.2f" % p)
    assert p >= 2.5
