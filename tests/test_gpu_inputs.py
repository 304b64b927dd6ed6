"""GPU runs of the inputs and friction branches the benchmark configurations
depend on, against the oracle:

* a friction load file carrying a non-uniform prestress (the 3-D planar
  workload's Gaussian-random-field shear prestress comes this way,
  friction.cpp:305-311,505-567), on the 3-D row-tile kernel in FP64 and FP32;
* slip-weakening and STZ parameter files (slipweak.cpp:310-383,
  stz.cpp:358-434) and the STZ state file;
* the slip-weakening branches no other test reaches: time weakening,
  cohesion and the tensile branch (slipweak.cpp:206-308);
* an ensemble whose instances carry load files (the batched workload);
* 3-D STZ with FP32 bulk fields and the learned root proposal (the STZ
  workload): iterations, fallbacks, and slip against the plain solve and the
  FP64 oracle.

Each input file holds a non-uniform perturbation's values at every fault
point; the GPU run with the file must be bit-identical to the GPU run with
the perturbation itself (a transposed or misordered layout fails), and both
within 1e-10 of the oracle (FP64), or within 1e-4 in slip and one time step
in rupture time (FP32 bulk).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import STZ_PARAMS, write_case
from fdfault_amd.problems import Pert
from test_oracle_law_pins import file_cases

pytestmark = pytest.mark.gpu

TOL = 1e-10
FIELDS = ("U", "V", "S", "Sn", "Ux", "Vx", "Sx")


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _vs_oracle(g, o, fields, tol=TOL):
    sn = np.abs(o.iface(0, "Sn")).max()
    for w in fields:
        a, b = g.iface(0, w), o.iface(0, w)
        den = max(np.abs(b).max(), sn if w in ("S", "Sx") else 0., 1e-300)
        r = np.abs(a - b).max() / den
        assert r <= tol, (w, r)
    assert _rel(g.array("f"), o.array("f")) <= tol


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("which", ["load", "sw", "stz"])
def test_input_file_equals_perturbation_on_gpu(tmp_path, ndim, which):
    from fdfault_amd import lib
    pa, pb = file_cases(tmp_path, ndim, which)
    nt = 10 if ndim == 3 else 25
    print("case", which, ndim, flush=True)
    ga, gb = lib.Solver(pa), lib.Solver(pb)
    ga.step(nt)
    gb.step(nt)
    fields = FIELDS + (("state",) if which == "stz" else ())
    for w in fields:
        np.testing.assert_array_equal(ga.iface(0, w), gb.iface(0, w), err_msg=w)
    np.testing.assert_array_equal(ga.array("f"), gb.array("f"))
    o = O.Oracle(pb)
    o.step(nt)
    _vs_oracle(gb, o, fields)
    for h in (ga, gb, o):
        h.close()


def _grf_load_case(tmp, name, nx, ny, nz, nt, dx=0.1):
    """3-D two-block slip-weakening problem (sbporder 3, as the 3-D planar
    workload) whose prestress comes from a load file: sn = -120 MPa, shear 70 MPa + a seeded Gaussian random field
    (std 5 MPa, 2 km correlation) over the fault's x-z plane, and a square
    nucleation patch from a perturbation."""
    from fdfault_amd.generators import gaussian_random_field, write_load_file
    d = os.path.join(str(tmp), name)
    os.makedirs(d, exist_ok=True)
    tau = 70.0 + gaussian_random_field((nx, nz), dx, 2.0, 5.0, 77)
    assert tau.shape == (nx, nz) and np.ptp(tau) > 5.0
    lf = write_load_file(os.path.join(d, "load.bin"), np.full((nx, nz), -120.0), tau, np.zeros((nx, nz)))
    Lx, Lz = (nx - 1) * dx, (nz - 1) * dx
    loads = [Pert("boxcar", 0., 0.5 * Lx, 0.6, 0.5 * Lz, 0.6, (0., 15.0, 0.))]
    return write_case(str(tmp), name, ndim=3, nx=nx, ny=ny, nz=nz, dx=dx, nt=nt, loads=loads, loadfile=lf,
                      sbporder=3)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_3d_load_file_prestress_row_kernel(tmp_path, precision):
    from fdfault_amd import lib
    nt = 60
    p = _grf_load_case(tmp_path, "grf", 81, 31, 51, nt)
    prec = lib.FP32 if precision == "fp32" else lib.FP64
    g = lib.Solver(p, precision=prec)
    assert g.info.row_blocks == 2
    o = O.Oracle(p)
    g.step(nt)
    o.step(nt)
    V = o.iface(0, "V")
    print(precision, "max V", V.max(), "ruptured", int((o.front(0) >= 0).sum()), flush=True)
    assert V.max() > 0.1
    if prec == lib.FP64:
        _vs_oracle(g, o, FIELDS)
        assert np.abs(g.front(0) - o.front(0)).max() <= TOL * nt * o.dt
    else:
        assert _rel(g.iface(0, "U"), o.iface(0, "U")) <= 1e-4
        fg, fo = g.front(0), o.front(0)
        both = (fg >= 0) & (fo >= 0)
        assert np.array_equal(fg >= 0, fo >= 0) or ((fg >= 0) != (fo >= 0)).sum() <= 2
        assert np.abs(fg[both] - fo[both]).max() <= 1.0001 * o.dt
    g.close()
    o.close()


def _branch_case(tmp, name, ndim):
    """Slip weakening with every branch of slipweak.cpp:206-308 active:
    cohesion (c0 = 10 MPa outside the patch, as problems/tpv5.py sets for
    its barriers), forced time weakening (trup, tc) in a patch, and a patch
    under tensile normal stress that slips at the cohesion alone."""
    g = dict(nx=61, ny=21, dx=0.1) if ndim == 2 else dict(nx=41, ny=14, nz=25, dx=0.1)
    xc, zc = 0.5 * (g["nx"] - 1) * 0.1, 0.5 * (g.get("nz", 1) - 1) * 0.1
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., xc + 1.2, 0.5, zc, 0.5 if ndim == 3 else 0., (150., 0., 0.))]  # tensile patch
    sw = [Pert("constant", 0., 0., 0., 0., 0., (0.4, 0.677, 0.525, 10., 0., 0.)),
          Pert("boxcar", 0., xc, 0.6, zc, 0.6 if ndim == 3 else 0., (0., 0., 0., -10., 0.02, 0.04)),
          Pert("boxcar", 0., xc + 1.2, 0.5, zc, 0.5 if ndim == 3 else 0., (0., 0., 0., -7., 0., 0.))]
    return write_case(str(tmp), name, ndim=ndim, nt=40, loads=loads, sw=sw, **g)


@pytest.mark.parametrize("ndim", [2, 3])
def test_slipweak_cohesion_time_weakening_tensile(tmp_path, ndim):
    from fdfault_amd import lib
    p = _branch_case(tmp_path, "br%d" % ndim, ndim)
    nt = 40
    g, o = lib.Solver(p), O.Oracle(p)
    g.step(nt)
    o.step(nt)
    sn, V, S = o.iface(0, "Sn"), o.iface(0, "V"), o.iface(0, "S")
    tens = sn > 0
    print("tensile points", int(tens.sum()), "slipping there", int((V[tens] > 0).sum()),
          "max V", V.max(), flush=True)
    assert tens.any() and (V[tens] > 0).any()
    # tensile points that slip carry the patch's cohesion, 10 - 7 = 3 MPa
    np.testing.assert_allclose(S[tens & (V > 0)], 3.0, rtol=1e-12)
    assert V.max() > 0.5  # the time-weakened patch ruptured
    _vs_oracle(g, o, FIELDS)
    g.close()
    o.close()


def test_batch_instances_with_load_files(tmp_path, monkeypatch):
    """The ensemble workload's instances (random-field prestress from a load
    file each, uneven sizes): FP64 batch within 1e-10 of the oracle, FP32
    batch bit-identical to FP32 single runs."""
    from fdfault_amd import lib
    from fdfault_amd.ensemble import ensemble_workload
    monkeypatch.setenv("FDF_TUNE_CACHE", str(tmp_path / "tune.txt"))
    nt = 30
    paths, shapes = ensemble_workload(str(tmp_path / "ens"), n=3, seed=5, nt=nt,
                                      sizes=[(81, 50), (121, 40), (65, 64)])
    for p in paths:
        assert "load.bin" in open(p).read()
    b = lib.Batch(paths, precision=lib.FP64, graph=True)
    b.step(nt)
    for i, p in enumerate(paths):
        o = O.Oracle(p)
        o.step(nt)
        for w in ("U", "V", "S", "Sn"):
            assert _rel(b.iface(i, 0, w), o.iface(0, w)) <= TOL, (i, w)
        assert _rel(b.array(i, "f"), o.array("f")) <= TOL
        assert o.iface(0, "V").max() > 1e-3
        o.close()
    b.close()
    b = lib.Batch(paths, precision=lib.FP32, graph=True)
    b.step(nt)
    for i, p in enumerate(paths):
        s = lib.Solver(p, precision=lib.FP32)
        s.step(nt)
        assert np.array_equal(b.array(i, "f"), s.array("f")), i
        assert np.array_equal(b.iface(i, 0, "U"), s.iface(0, "U")), i
        s.close()
    b.close()


def _stz3d(tmp, name, seed, nx=49, ny=14, nz=33, nt=30):
    d = os.path.join(str(tmp), name)
    os.makedirs(d, exist_ok=True)
    chi = np.random.default_rng(seed).uniform(0.06, 0.08, size=(nx, nz))
    sp = os.path.join(d, "state.bin")
    chi.tofile(sp)
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
    return write_case(str(tmp), name, ndim=3, nx=nx, ny=ny, nz=nz, dx=0.1, nt=nt, law="stz",
                      stz=dict(STZ_PARAMS, chi0=0., statefile=sp), loads=loads)


def test_3d_stz_fp32_learned_proposal(tmp_path):
    """The STZ workload's configuration: 3-D, FP32 bulk, FP64 fault, the
    network trained on the GPU from plain solves of another seed."""
    from fdfault_amd import lib
    nt = 30
    a = lib.Solver(_stz3d(tmp_path, "sa", 1, nt=nt), precision=lib.FP32)
    a.stz_log(400000)
    a.step(nt)
    rows = a.stz_log_read()
    a.close()
    print("logged", rows.shape, flush=True)
    assert rows.shape[0] > 20000
    P, losses = lib.mlp_train(rows, epochs=30, batch=512, lr=3e-3, seed=0)
    print("loss", losses[0], losses[-1], flush=True)
    case = _stz3d(tmp_path, "sb", 2, nt=nt)
    plain = lib.Solver(case, precision=lib.FP32)
    plain.step(nt)
    sp = plain.stz_stats()
    prop = lib.Solver(case, precision=lib.FP32)
    prop.stz_set_model(P)
    prop.step(nt)
    sm = prop.stz_stats()
    mp, mm = sp[1] / sp[0], sm[1] / sm[0]
    print("plain %.2f iterations per solve; proposal %.2f, used %d of %d, fallbacks %d"
          % (mp, mm, sm[2], sm[0], sm[3]), flush=True)
    assert sp[0] == sm[0] and sp[2] == 0 and sp[3] == 0
    assert sm[2] >= 0.95 * sm[0] and sm[3] <= 0.05 * sm[0]
    assert mm < 0.6 * mp
    for w in ("U", "V", "S", "state"):
        r = _rel(prop.iface(0, w), plain.iface(0, w))
        print(w, "proposal vs plain (both FP32 bulk)", r, flush=True)
        assert r <= 1e-5, w
    o = O.Oracle(case)
    o.step(nt)
    for w in ("U", "V", "state"):
        r = _rel(prop.iface(0, w), o.iface(0, w))
        print(w, "FP32 proposal vs FP64 oracle", r, flush=True)
        if w != "V":  # final slip and state: the FP32 bar
            assert r <= 1e-4, w
    for h in (plain, prop, o):
        h.close()
