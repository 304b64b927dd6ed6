"""Host problem compiler of the library (grid, metric, Jacobian, dt) against
the oracle's restatement of block::set_grid / problem::set_time_step:
bit-identical, for Cartesian and curved block surfaces, 2-D and 3-D."""
import numpy as np
import pytest

import oracle_lib as O
from cases import curved_surfaces, write_case
from fdfault_amd import lib


def _compare(path):
    info, x, m, j = lib.host_setup(path)
    o = O.Oracle(path)
    assert info.dt == o.dt
    assert info.nt == o.nt
    np.testing.assert_array_equal(x, o.array("x"))
    np.testing.assert_array_equal(m, o.array("metric"))
    np.testing.assert_array_equal(j, o.array("jac")[0])


@pytest.mark.parametrize("sbp", [2, 3, 4])
def test_setup_2d(tmp_path, sbp):
    _compare(write_case(str(tmp_path), "s2d", nx=41, ny=21, sbporder=sbp))


def test_setup_3d(tmp_path):
    _compare(write_case(str(tmp_path), "s3d", ndim=3, nx=21, ny=13, nz=15, sbporder=4))


def test_setup_curved(tmp_path):
    surf = curved_surfaces(str(tmp_path), 41, 21, 0.1)
    _compare(write_case(str(tmp_path), "scurv", nx=41, ny=21, dx=0.1, surf=surf))


def test_cartesian_metric_closed_form(tmp_path):
    """Linear TFI grid: x exact multiples of dx, metric = 1/L, J = L^2 (2-D)."""
    info, x, m, j = lib.host_setup(write_case(str(tmp_path), "cf", nx=41, ny=21, dx=0.1))
    L = 4.0
    np.testing.assert_allclose(m[0], 1.0 / L, rtol=1e-13)
    np.testing.assert_allclose(m[3][:, :101], 1.0 / 2.0, rtol=1e-13)
    np.testing.assert_allclose(m[1], 0.0, atol=1e-13)
    assert (j > 0).all()


@pytest.mark.parametrize("name", ["test2d", "tpv4"])
def test_setup_reference_examples(tmp_path, name):
    """The reference's own example problems (tests/golden): grid, metric,
    Jacobian, dt and nt bit-identical to the oracle's."""
    from cases import reference_example
    _compare(reference_example(str(tmp_path), name, 10, str(tmp_path) + "/data/"))
