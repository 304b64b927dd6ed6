"""FP32-bulk / FP64-fault path against the FP64 oracle.

Accuracy bar: final slip within 1e-4 relative (max-norm over the fault)
and rupture-front times within one time step, with every result reporting
its error against the FP64 oracle."""
import numpy as np
import pytest

import oracle_lib as O
from cases import curved_surfaces, write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu


def _compare(path, nsteps):
    from fdfault_amd import lib
    o = O.Oracle(path)
    g = lib.Solver(path, precision=lib.FP32)
    o.step(nsteps)
    g.step(nsteps)
    U, Uo = g.iface(0, "U"), o.iface(0, "U")
    err_u = np.abs(U - Uo).max() / np.abs(Uo).max()
    fo, fg = o.front(0), g.front(0)
    both = (fo >= 0) & (fg >= 0)
    dt_err = np.abs(fo[both] - fg[both]).max() / o.dt if both.any() else 0.0
    mismatch = ((fo >= 0) != (fg >= 0)).sum()
    V, Vo = g.iface(0, "V"), o.iface(0, "V")
    err_v = np.abs(V - Vo).max() / np.abs(Vo).max()
    print("fp32 vs fp64: slip %.2e  slip-rate %.2e  front %.2f steps, %d unmatched" % (err_u, err_v, dt_err, mismatch))
    g.close()
    o.close()
    return err_u, dt_err, mismatch


def test_fp32_mode2_slip_and_front(tmp_path):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 8.0, 1.5, 0., 0., (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "f32", nx=161, ny=81, dx=0.1, nt=120, loads=loads, front=("V", 0.001))
    err_u, dt_err, mismatch = _compare(p, 120)
    assert err_u <= 1e-4
    assert dt_err <= 1.0
    assert mismatch <= 2


def test_fp32_3d_slip(tmp_path):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 2.0, 0.6, 1.5, 0.6, (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "f32d3", ndim=3, nx=41, ny=21, nz=31, dx=0.1, nt=30, loads=loads,
                   front=("V", 0.001))
    err_u, dt_err, mismatch = _compare(p, 30)
    assert err_u <= 1e-4
    assert dt_err <= 1.0


def test_fp32_curved_plastic_slip(tmp_path):
    """FP32 on the curvilinear tiled kernel: rough plastic fault, prestress
    from s0 (sxy 52 over syy -120), nucleation +30 MPa of shear."""
    loads = [Pert("boxcar", 0., 8.0, 1.5, 0., 0., (0., 30., 0.))]
    surf = curved_surfaces(str(tmp_path), 161, 81, 0.1, amp=0.2)
    p = write_case(str(tmp_path), "f32c", nx=161, ny=81, dx=0.1, nt=100, loads=loads, front=("V", 0.001),
                   surf=surf, material="plastic", plastic=dict(mu=0.7, c=0., beta=0.35, eta=0.05 * 32.04),
                   s0=(-100., 52., 0., -120., 0., -55.))
    err_u, dt_err, mismatch = _compare(p, 100)
    assert err_u <= 1e-4
    assert dt_err <= 1.0
    assert mismatch <= 2
