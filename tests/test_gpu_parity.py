"""GPU parity: the HIP path against the CPU oracle on the same .in files.

FP64 bar: every field, slip, slip rate, traction, state and rupture-front
time within 1e-10 relative (max-norm of the difference over the max-norm of
the oracle field; front times within 1e-10 of the run time).  In practice
the FP64 path is bit-identical: kernels keep the reference's association and
are built with -ffp-contract=off; transcendental calls (STZ exp/log) may
differ by an ulp, hence the tolerance rather than equality.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import STZ_PARAMS, curved_surfaces, write_case
from fdfault_amd.problems import Output, Pert

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def _check(path, nsteps, iface_fields=("U", "V", "S", "Sn", "Ux", "Vx", "Sx"), tol=TOL, state=False, exact=False):
    from fdfault_amd import lib
    print("case", os.path.basename(path), "steps", nsteps, "exact", exact, flush=True)
    o = O.Oracle(path)
    g = lib.Solver(path, exact=exact)
    o.step(nsteps)
    g.step(nsteps)
    fo, fg = o.array("f"), g.array("f")
    worst = 0.0
    for q in range(fo.shape[0]):
        r = _rel(fg[q], fo[q])
        worst = max(worst, r)
        assert r <= tol, ("field", q, r)
    r = _rel(g.array("df"), o.array("df"))
    assert r <= tol, ("df", r)
    for i in range(o.nifaces):
        shp = o.iface_shape(i)
        if shp[3] == 0:
            continue
        fields = list(iface_fields) + (["state"] if state else [])
        sn_scale = np.abs(o.iface(i, "Sn")).max()
        for w in fields:
            a, b = g.iface(i, w), o.iface(i, w)
            if w in ("S", "Sx"):
                # tractions relative to the normal-stress scale: on a
                # frictionless fault they are zero up to rounding noise
                r = np.abs(a - b).max() / max(np.abs(b).max(), sn_scale, 1e-300)
            else:
                r = _rel(a, b)
            assert r <= tol, (w, r)
        fo_, fg_ = o.front(i), g.front(i)
        if fo_ is not None:
            assert np.abs(fo_ - fg_).max() <= tol * max(1.0, nsteps * o.dt)
    g.close()
    o.close()
    return worst


def test_mode2_slipweak_sbp4_rk4(tmp_path):
    p = write_case(str(tmp_path), "m2", nx=81, ny=41, dx=0.1, nt=30)
    _check(p, 30)


@pytest.mark.parametrize("sbp,rk", [(2, 1), (2, 2), (3, 3), (3, 4), (4, 2)])
def test_mode2_orders(tmp_path, sbp, rk):
    p = write_case(str(tmp_path), "m2o", nx=61, ny=31, dx=0.1, nt=12, sbporder=sbp, rkorder=rk)
    _check(p, 12)


@pytest.mark.parametrize("bc", ["free", "rigid"])
def test_mode2_boundaries(tmp_path, bc):
    p = write_case(str(tmp_path), "m2b", nx=61, ny=31, dx=0.1, nt=15, bc=bc)
    _check(p, 15)


def test_mode2_dissipation(tmp_path):
    p = write_case(str(tmp_path), "m2d", nx=61, ny=31, dx=0.1, nt=15, cdiss=0.1)
    _check(p, 15)


def test_mode3_slipweak(tmp_path):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 0., 70.)),
             Pert("boxcar", 0., 3.0, 0.6, 0., 0., (0., 0., 11.6))]
    p = write_case(str(tmp_path), "m3", mode=3, nx=61, ny=31, dx=0.1, nt=20, loads=loads)
    _check(p, 20)


@pytest.mark.parametrize("law", ["locked", "frictionless"])
def test_mode2_other_laws(tmp_path, law):
    p = write_case(str(tmp_path), "m2l", nx=61, ny=31, dx=0.1, nt=12, law=law,
                   s0=(-50., 20., 0., -60., 0., 0.), front=None if law == "locked" else ("V", 0.001))
    fields = () if law == "locked" else ("U", "V", "S", "Sn", "Ux", "Vx", "Sx")
    _check(p, 12, iface_fields=fields)


def test_mode2_stz(tmp_path):
    p = write_case(str(tmp_path), "stz2", nx=61, ny=31, dx=0.1, nt=15, law="stz",
                   stz=dict(STZ_PARAMS, chi0=0.07))
    _check(p, 15, state=True)


NUC2 = [Pert("boxcar", 0., 3.0, 0.6, 0., 0., (0., 40., 0.))]
NUC3 = [Pert("boxcar", 0., 1.2, 0.6, 0.8, 0.6, (0., 40., 0.))]


def test_mode2_plastic(tmp_path):
    """Drucker-Prager close to yield: the rupture drives points over it."""
    plastic = dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04)
    p = write_case(str(tmp_path), "pl2", nx=61, ny=31, dx=0.1, nt=15, material="plastic", plastic=plastic,
                   s0=(-100., 52., 0., -120., 0., -55.), loads=NUC2)
    _check(p, 15)


def test_mode2_curved(tmp_path):
    surf = curved_surfaces(str(tmp_path), 61, 31, 0.1, amp=0.1)
    p = write_case(str(tmp_path), "curv", nx=61, ny=31, dx=0.1, nt=15, surf=surf)
    _check(p, 15)


def test_mode2_het_stress_material(tmp_path):
    nx, ny = 61, 31
    rng = np.random.default_rng(1)
    N = nx * 2 * ny
    s = rng.normal(0., 2., size=(3, nx, 2 * ny, 1))
    sp = str(tmp_path / "hs.bin")
    s.tofile(sp)
    mat = np.empty((3, nx, 2 * ny, 1))
    mat[0] = 2.67 * (1 + 0.05 * rng.random((nx, 2 * ny, 1)))
    mat[1] = 32.04 * (1 + 0.05 * rng.random((nx, 2 * ny, 1)))
    mat[2] = 32.04 * (1 + 0.05 * rng.random((nx, 2 * ny, 1)))
    mp = str(tmp_path / "mat.bin")
    mat.tofile(mp)
    p = write_case(str(tmp_path), "het", nx=nx, ny=ny, dx=0.1, nt=12, hetstressfile=sp, matfile=mp)
    _check(p, 12)


@pytest.mark.parametrize("exact", [True, False])
def test_3d_het_stress_material(tmp_path, exact):
    """3-D heterogeneous prestress (six components) and material (rho, lambda,
    G per point, fields.cpp:452-711): the generic kernels' per-point moduli
    and the fault solve's per-point impedances in 3-D."""
    nx, ny, nz = 25, 14, 17
    rng = np.random.default_rng(3)
    shape = (nx, 2 * ny, nz)
    s = rng.normal(0., 2., size=(6,) + shape)
    sp = str(tmp_path / "hs3.bin")
    s.tofile(sp)
    mat = np.empty((3,) + shape)
    mat[0] = 2.67 * (1 + 0.05 * rng.random(shape))
    mat[1] = 32.04 * (1 + 0.05 * rng.random(shape))
    mat[2] = 32.04 * (1 + 0.05 * rng.random(shape))
    mp = str(tmp_path / "mat3.bin")
    mat.tofile(mp)
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 1.2, 0.6, 0.8, 0.6, (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "het3", ndim=3, nx=nx, ny=ny, nz=nz, dx=0.1, nt=8, loads=loads, sbporder=3,
                   hetstressfile=sp, matfile=mp)
    _check(p, 8, exact=exact)


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_3d_rk_orders(tmp_path, rk):
    """RK orders 1-3 on the 3-D row kernel (1 to 3 stages: which stages read
    and write df differs from RK4's five)."""
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 1.5, 0.6, 0.8, 0.6, (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "rk3d%d" % rk, ndim=3, nx=31, ny=14, nz=17, dx=0.1, nt=8, loads=loads,
                   sbporder=3, rkorder=rk, cfl=0.2)
    from fdfault_amd import lib
    g = lib.Solver(p)
    assert g.info.row_blocks == 2
    g.close()
    _check(p, 8)


def test_3d_dissipation_rigid(tmp_path):
    """3-D artificial dissipation (the cdiss terms of block.cpp:1588-1598) with rigid
    outer faces."""
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 1.2, 0.6, 0.8, 0.6, (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "diss3", ndim=3, nx=25, ny=14, nz=17, dx=0.1, nt=8, loads=loads, sbporder=3,
                   cdiss=0.1, bc="rigid")
    _check(p, 8)


def test_mode3_het_stress(tmp_path):
    """Mode III with heterogeneous prestress (sxz, syz per point)."""
    nx, ny = 61, 31
    rng = np.random.default_rng(5)
    s = rng.normal(0., 2., size=(2, nx, 2 * ny, 1))
    sp = str(tmp_path / "hs_m3.bin")
    s.tofile(sp)
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 0., 70.)),
             Pert("boxcar", 0., 3.0, 0.6, 0., 0., (0., 0., 11.6))]
    p = write_case(str(tmp_path), "hm3", mode=3, nx=nx, ny=ny, dx=0.1, nt=15, loads=loads, hetstressfile=sp)
    _check(p, 15)


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("sbp", [2, 3, 4])
def test_3d_slipweak(tmp_path, sbp, exact):
    """exact=False exercises the rectilinear interior kernel (x-marching
    register window + LDS tile) on the interior and bulk_kernel on the shell."""
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 3.5, 0.6, 1.4, 0.6, (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "d3", ndim=3, nx=71, ny=24, nz=29, dx=0.1, nt=8, loads=loads, sbporder=sbp,
                   s0=(0., 0., 0., 0., 0., 0.))
    _check(p, 8, exact=exact)


def test_3d_stz(tmp_path):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
    p = write_case(str(tmp_path), "d3stz", ndim=3, nx=25, ny=14, nz=17, dx=0.1, nt=6, law="stz",
                   stz=dict(STZ_PARAMS, chi0=0.07), loads=loads)
    _check(p, 6, state=True)


def test_3d_plastic(tmp_path):
    plastic = dict(mu=0.675, c=0., beta=0.35, eta=0.05 * 32.04)
    p = write_case(str(tmp_path), "d3pl", ndim=3, nx=25, ny=14, nz=17, dx=0.1, nt=6, material="plastic",
                   plastic=plastic, s0=(-100., 52., 0., -120., 0., -55.), bc="free", loads=NUC3, plastic_tensor=1)
    _check(p, 6)


def test_kernel_level_random_state(tmp_path):
    """One RK stage from a random (f, df): every branch of the stencil,
    boundary and interface SAT is active at once."""
    from fdfault_amd import lib
    p = write_case(str(tmp_path), "rnd", nx=41, ny=21, dx=0.1, nt=2, cdiss=0.05)
    o = O.Oracle(p)
    g = lib.Solver(p)
    rng = np.random.default_rng(7)
    f = rng.normal(size=o.array("f").shape)
    df = rng.normal(size=o.array("df").shape)
    for h in (o, g):
        h.set_array("f", f)
        h.set_array("df", df)
    for st in range(o.nstages):
        o.stage(0, st)
        g.stage(0, st)
        assert _rel(g.array("df"), o.array("df")) <= TOL
        assert _rel(g.array("f"), o.array("f")) <= TOL


@pytest.mark.parametrize("sbp", [2, 3, 4])
def test_kernel_level_random_state_3d(tmp_path, sbp):
    """3-D stage from a random state: interior kernel + shell kernel together
    must reproduce the oracle's df and f (rounding-level differences only)."""
    from fdfault_amd import lib
    p = write_case(str(tmp_path), "rnd3", ndim=3, nx=45, ny=22, nz=71, dx=0.1, nt=2, sbporder=sbp)
    o = O.Oracle(p)
    g = lib.Solver(p)
    rng = np.random.default_rng(11)
    f = rng.normal(size=o.array("f").shape)
    df = rng.normal(size=o.array("df").shape)
    for h in (o, g):
        h.set_array("f", f)
        h.set_array("df", df)
    for st in range(o.nstages):
        o.stage(0, st)
        g.stage(0, st)
        assert _rel(g.array("df"), o.array("df")) <= 1e-12
        assert _rel(g.array("f"), o.array("f")) <= 1e-12


@pytest.mark.parametrize("plastic,sbp", [(False, 3), (True, 3), (True, 2), (False, 4), (True, 4)])
def test_kernel_level_random_state_curvilinear(tmp_path, plastic, sbp):
    """One RK stage from a random state on curvilinear 2-D blocks, elastic and
    plastic: the curvilinear tiled kernel (per-point metric, conservative
    velocity rates, per-point external SAT maps, szz rate) against the
    oracle's df and f."""
    from fdfault_amd import lib
    nx, ny = 61, 31
    surf = curved_surfaces(str(tmp_path), nx, ny, 0.1, amp=0.15)
    kw = dict(surf=surf)
    if plastic:
        kw.update(material="plastic", plastic=dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04),
                  s0=(-100., 52., 0., -120., 0., -55.))
    p = write_case(str(tmp_path), "rc", nx=nx, ny=ny, dx=0.1, nt=2, sbporder=sbp, **kw)
    o = O.Oracle(p)
    g = lib.Solver(p)
    assert g.info.curv_blocks == 2 and g.info.tiled_blocks == 0
    rng = np.random.default_rng(17)
    f = rng.normal(size=o.array("f").shape)
    df = rng.normal(size=o.array("df").shape)
    for h in (o, g):
        h.set_array("f", f)
        h.set_array("df", df)
    for st in range(o.nstages):
        o.stage(0, st)
        g.stage(0, st)
        edf, ef = _rel(g.array("df"), o.array("df")), _rel(g.array("f"), o.array("f"))
        print("plastic", plastic, "stage", st, "df %.2e f %.2e" % (edf, ef), flush=True)
        assert edf <= 1e-12 and ef <= 1e-12
    g.close()
    o.close()


@pytest.mark.parametrize("curved", [False, True])
def test_merged_launch_matches_per_block(tmp_path, monkeypatch, curved):
    """One bulk launch per stage over every block's tiles gives the same bits
    as one launch per block (rectilinear elastic and curvilinear plastic)."""
    from fdfault_amd import lib
    kw = dict(nx=61, ny=31, dx=0.1, nt=12)
    if curved:
        kw.update(surf=curved_surfaces(str(tmp_path), 61, 31, 0.1, amp=0.1), material="plastic",
                  plastic=dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04), s0=(-100., 52., 0., -120., 0., -55.),
                  loads=NUC2)
    p = write_case(str(tmp_path), "mg", **kw)
    res = []
    for m in ("1", "0"):
        monkeypatch.setenv("FDF_MERGE2D", m)
        g = lib.Solver(p, precision=lib.FP32)
        assert (g.info.curv_blocks if curved else g.info.tiled_blocks) == 2
        g.step(12)
        res.append((g.array("f"), g.iface(0, "U"), g.front(0)))
        g.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_mode2_curved_plastic(tmp_path):
    """Rough-fault plastic rupture on curvilinear blocks (the curvilinear
    tiled kernel with the plastic return map and friction) against the oracle."""
    surf = curved_surfaces(str(tmp_path), 61, 31, 0.1, amp=0.1)
    plastic = dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04)
    p = write_case(str(tmp_path), "cpl", nx=61, ny=31, dx=0.1, nt=15, surf=surf, material="plastic",
                   plastic=plastic, s0=(-100., 52., 0., -120., 0., -55.), loads=NUC2)
    _check(p, 15)


@pytest.mark.parametrize("precision,shape,nz,sbp", [("fp64", "auto", 71, 3), ("fp64", "0", 71, 3),
                                                    ("fp64", "auto", 181, 3), ("fp64", "auto", 71, 4),
                                                    ("fp32", "auto", 71, 3), ("fp32", "0", 71, 3),
                                                    ("fp32", "auto", 181, 3), ("fp32", "auto", 71, 4)])
@pytest.mark.parametrize("var", [1, 2])
def test_row_kernel_variants_random_state(tmp_path, monkeypatch, precision, shape, nz, sbp, var):
    """Every variant of the row-tile kernel (FDF_ROWS=1 single-buffered,
    2 double-buffered LDS tile; FP32 workgroups of 6 rows for rows of at
    most 160 points, 5 rows up to 204, or with FDF_ROWS_SHAPE=0 always 4;
    FP64 workgroups of 4 rows for rows of at most 128 points, else 2; FP32
    sbporder 4 in 6-row workgroups of 768 threads for rows of at most 128)
    reproduces the oracle's RK stages from a random state."""
    from fdfault_amd import lib
    monkeypatch.setenv("FDF_ROWS", str(var))
    if shape != "auto":
        monkeypatch.setenv("FDF_ROWS_SHAPE", shape)
    p = write_case(str(tmp_path), "rv", ndim=3, nx=37, ny=22, nz=nz, dx=0.1, nt=2, sbporder=sbp)
    o = O.Oracle(p)
    g = lib.Solver(p, precision=lib.FP32 if precision == "fp32" else lib.FP64)
    assert g.info.row_blocks == 2
    # FP32: rounding of ~30-term stencil sums accumulated over 5 stages of random
    # data stays near 1e-6; an index or coefficient error would be O(0.1)
    tol = 1e-12 if precision == "fp64" else 1e-5
    rng = np.random.default_rng(13)
    f = rng.normal(size=o.array("f").shape)
    df = rng.normal(size=o.array("df").shape)
    for h in (o, g):
        h.set_array("f", f)
        h.set_array("df", df)
    for st in range(o.nstages):
        o.stage(0, st)
        g.stage(0, st)
        edf, ef = _rel(g.array("df"), o.array("df")), _rel(g.array("f"), o.array("f"))
        print(precision, "shape", shape, "sbporder", sbp, "FDF_ROWS", var, "stage", st, "df %.2e f %.2e" % (edf, ef),
              flush=True)
        assert edf <= tol and ef <= tol
    g.close()
    o.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_row_tile_order_does_not_change_results(tmp_path, monkeypatch, precision):
    """The order in which the row-tile kernel takes its tiles (patches of gy y
    tiles by gx x-chunks, FDF_ROWS_GROUP) only schedules work: every tile is
    updated exactly once, so the fields after a stage from a random state are
    bit-identical for any patch shape, including ones that do not divide the
    tile grid (22 rows per block: 11 FP64 or 4 FP32 y tiles; 10 x-chunks)."""
    from fdfault_amd import lib
    p = write_case(str(tmp_path), "to", ndim=3, nx=37, ny=22, nz=71, dx=0.1, nt=2, sbporder=3)
    prec = lib.FP32 if precision == "fp32" else lib.FP64
    rng = np.random.default_rng(5)
    out = []
    for grp in ("16,2", "1000,1", "3,5", "7,2", "1,1", "2,1000"):
        monkeypatch.setenv("FDF_ROWS_GROUP", grp)
        g = lib.Solver(p, precision=prec)
        assert g.info.row_blocks == 2
        if not out:
            f = rng.normal(size=g.array("f").shape)
            df = rng.normal(size=g.array("df").shape)
        g.set_array("f", f)
        g.set_array("df", df)
        for st in range(g.nstages):
            g.stage(0, st)
        out.append((grp, g.array("f"), g.array("df")))
        g.close()
    for grp, fa, dfa in out[1:]:
        print(precision, "FDF_ROWS_GROUP", grp, flush=True)
        assert np.array_equal(fa, out[0][1]) and np.array_equal(dfa, out[0][2]), grp


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_paired_block_streams_do_not_change_results(tmp_path, monkeypatch, precision):
    """Launching the second 3-D block's bulk update on a second stream beside
    the first block's (FDF_PAIR3D=1) only overlaps independent work: the two
    blocks of a stage read the previous stage's fields and write disjoint
    arrays, so fields, slip and fronts after several steps, eager and from
    graph replay, are bit-identical to one stream (FDF_PAIR3D=0)."""
    from fdfault_amd import lib
    p = write_case(str(tmp_path), "pr", ndim=3, nx=37, ny=22, nz=71, dx=0.1, nt=8, sbporder=3)
    prec = lib.FP32 if precision == "fp32" else lib.FP64
    out = []
    for pair in ("0", "1"):
        monkeypatch.setenv("FDF_PAIR3D", pair)
        for graph in (False, True):
            g = lib.Solver(p, precision=prec, graph=graph)
            for k in range(3):
                g.step(2)
                out.append((pair, graph, k, g.array("f"), g.array("df")))
            g.close()
    n = len(out) // 4
    for i in range(1, 4):
        for k in range(n):
            pair, graph, kk, fa, dfa = out[i * n + k]
            print(precision, "FDF_PAIR3D", pair, "graph", graph, "call", kk, flush=True)
            assert np.array_equal(fa, out[k][3]) and np.array_equal(dfa, out[k][4]), (pair, graph, kk)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["plain2d", "plastic2d", "stz2d", "planar3d", "stz3d"])
def test_graph_replay_matches_eager(tmp_path, case, precision):
    """hipGraph replay of whole steps (device step counter) gives the same
    fields, slip and fronts as eager launches, over several replays with
    host reads in between, in both bulk precisions."""
    from fdfault_amd import lib
    kw = dict(nx=61, ny=31, dx=0.1, nt=12)
    if case == "plastic2d":
        kw.update(material="plastic", plastic=dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04),
                  s0=(-100., 52., 0., -120., 0., -55.), loads=NUC2)
    elif case.startswith("stz"):
        kw.update(law="stz", stz=dict(STZ_PARAMS, chi0=0.07))
    if case.endswith("3d"):
        kw.update(ndim=3, nx=31, ny=16, nz=25)
        if case == "planar3d":
            kw.update(loads=[Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))] + NUC3)
    p = write_case(str(tmp_path), "gr", **kw)
    a = lib.Solver(p, precision=precision)
    b = lib.Solver(p, precision=precision, graph=True)
    for k in range(4):
        a.step(3)
        b.step(3)
        assert np.array_equal(a.array("f"), b.array("f"))
        for w in ("U", "V", "S", "Sn") + (("state",) if case.startswith("stz") else ()):
            assert np.array_equal(a.iface(0, w), b.iface(0, w)), w
        assert np.array_equal(a.front(0), b.front(0))
    assert np.abs(a.iface(0, "V")).max() > 0
    a.close()
    b.close()


def test_3d_fast_path_drift_long_run(tmp_path):
    """The rectilinear interior kernel (block-constant metric) against the
    bit-faithful FP64 path over 150 steps of a rupture: within 1e-10."""
    from fdfault_amd import lib
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 3.0, 0.8, 2.0, 0.8, (0., 11.6, 0.))]
    p = write_case(str(tmp_path), "drift", ndim=3, nx=61, ny=31, nz=41, dx=0.1, nt=150, loads=loads, sbporder=3)
    a = lib.Solver(p, exact=True)
    b = lib.Solver(p, exact=False)
    a.step(150)
    b.step(150)
    for w in ("U", "V", "S"):
        r = _rel(b.iface(0, w), a.iface(0, w))
        print(w, "fast vs exact", r)
        assert r <= 1e-10, (w, r)
    r = _rel(b.array("f"), a.array("f"))
    print("f fast vs exact", r)
    assert r <= 1e-10
    fa, fb = a.front(0), b.front(0)
    assert np.abs(fa - fb).max() <= a.dt * 1e-6
