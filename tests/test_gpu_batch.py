"""2-D rectilinear tiled kernel and the ensemble (batch) path on the GPU.

* The 2-D tiled kernel (register window along x, LDS row along y, compact
  face SAT) against the bit-faithful FP64 path and the oracle: within 1e-10
  over a long rupture, 1e-12 for one stage from a random state.
* Tile shape changes only the work split: results are bit-identical for any
  tile width / x-chunk.
* A batch of instances of different sizes gives, per instance, bit-identical
  results to single runs (same kernels, same arithmetic per point), in FP32
  and FP64, eager and graph-replayed, also with instances of different
  friction laws in one batch; the autotuner's cache is reused.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _nuc(nx, dx):
    Lx = (nx - 1) * dx
    return [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
            Pert("boxcar", 0., 0.5 * Lx, 0.6, 0., 0., (0., 11.6, 0.))]


def _case(tmp, name, nx, ny, nt=40, dx=0.1, **kw):
    kw.setdefault("s0", (0., 0., 0., 0., 0., 0.))
    return write_case(str(tmp), name, nx=nx, ny=ny, dx=dx, nt=nt, loads=_nuc(nx, dx), front=("V", 0.001), **kw)


def test_mode2_fast_path_drift_long_run(tmp_path):
    """2-D tiled kernel (block-constant metric, fused multiply-adds) against
    the bit-faithful FP64 path over 300 steps of a rupture: within 1e-10."""
    from fdfault_amd import lib
    p = _case(tmp_path, "drift2", 121, 61, nt=300)
    a = lib.Solver(p, exact=True)
    b = lib.Solver(p, exact=False)
    assert b.info.tiled_blocks == 2 and a.info.tiled_blocks == 0
    a.step(300)
    b.step(300)
    for w in ("U", "V", "S"):
        r = _rel(b.iface(0, w), a.iface(0, w))
        print(w, "fast vs exact", r, flush=True)
        assert r <= 1e-10, (w, r)
    assert _rel(b.array("f"), a.array("f")) <= 1e-10
    assert np.abs(a.front(0) - b.front(0)).max() <= a.dt * 1e-6
    a.close()
    b.close()


@pytest.mark.parametrize("sbp", [2, 3, 4])
def test_kernel_level_random_state_2d_tiled(tmp_path, sbp):
    """Every branch of the 2-D tiled kernel at once (interior, one-sided rows
    in x and y, four external faces, the interface) from a random state."""
    from fdfault_amd import lib
    p = write_case(str(tmp_path), "rnd2", nx=77, ny=45, dx=0.1, nt=2, sbporder=sbp)
    o = O.Oracle(p)
    g = lib.Solver(p)
    assert g.info.tiled_blocks == 2
    rng = np.random.default_rng(5)
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
    g.close()
    o.close()


def test_tile_shape_does_not_change_results(tmp_path, monkeypatch):
    from fdfault_amd import lib
    p = _case(tmp_path, "shape", 97, 53, nt=20)
    res = []
    for tw, cx in ((64, 1), (128, 3), (256, 8)):
        monkeypatch.setenv("FDF_TILE2_TW", str(tw))
        monkeypatch.setenv("FDF_TILE2_CX", str(cx))
        g = lib.Solver(p, precision=lib.FP32)
        g.step(20)
        res.append((g.array("f"), g.iface(0, "U"), g.front(0)))
        g.close()
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0])
        assert np.array_equal(r[1], res[0][1])
        assert np.array_equal(r[2], res[0][2])


SIZES = [(61, 31), (97, 45), (45, 27), (130, 70)]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_runs(tmp_path, monkeypatch, precision, graph):
    from fdfault_amd import lib
    prec = lib.FP32 if precision == "fp32" else lib.FP64
    monkeypatch.setenv("FDF_TUNE_CACHE", str(tmp_path / "tune.txt"))
    paths = [_case(tmp_path, "b%d" % i, nx, ny, nt=30) for i, (nx, ny) in enumerate(SIZES)]
    print("batch", precision, "graph", graph, flush=True)
    b = lib.Batch(paths, precision=prec, graph=graph)
    t = b.tuning()
    print("tuning", t, flush=True)
    assert t["tile_width"] in (64, 128, 256) and t["chunk"] > 0
    for k in range(3):
        b.step(10)
    monkeypatch.setenv("FDF_TILE2_TW", str(t["tile_width"]))
    monkeypatch.setenv("FDF_TILE2_CX", str(t["chunk"]))
    for i, p in enumerate(paths):
        s = lib.Solver(p, precision=prec)
        s.step(30)
        assert np.array_equal(b.array(i, "f"), s.array("f")), i
        for w in ("U", "V", "S", "Sn"):
            assert np.array_equal(b.iface(i, 0, w), s.iface(0, w)), (i, w)
        assert np.array_equal(b.front(i, 0), s.front(0)), i
        assert b.infos[i].nx[0] == SIZES[i][0]
        s.close()
    b.close()


def test_batch_launch_modes_agree(tmp_path, monkeypatch):
    """The persistent grid and the one-workgroup-per-tile launch of the batched
    bulk kernel compute the same tiles: bit-identical results."""
    from fdfault_amd import lib
    paths = [_case(tmp_path, "m%d" % i, nx, ny, nt=20) for i, (nx, ny) in enumerate(SIZES)]
    monkeypatch.setenv("FDF_BATCH_TW", "64")
    monkeypatch.setenv("FDF_BATCH_CX", "2")
    res = []
    for mode in ("1", "0"):
        monkeypatch.setenv("FDF_BATCH_PERSIST", mode)
        b = lib.Batch(paths, precision=lib.FP32, graph=True)
        assert b.tuning()["persistent"] == (mode == "1")
        b.step(20)
        res.append([(b.array(i, "f"), b.iface(i, 0, "U")) for i in range(len(paths))])
        b.close()
    for (fa, ua), (fb, ub) in zip(res[0], res[1]):
        assert np.array_equal(fa, fb) and np.array_equal(ua, ub)


def test_batch_fp64_instance_matches_oracle(tmp_path, monkeypatch):
    from fdfault_amd import lib
    monkeypatch.setenv("FDF_TUNE_CACHE", str(tmp_path / "tune.txt"))
    paths = [_case(tmp_path, "o%d" % i, nx, ny, nt=25) for i, (nx, ny) in enumerate(SIZES[:2])]
    b = lib.Batch(paths, precision=lib.FP64, graph=True)
    b.step(25)
    for i, p in enumerate(paths):
        o = O.Oracle(p)
        o.step(25)
        for w in ("U", "V", "S"):
            assert _rel(b.iface(i, 0, w), o.iface(0, w)) <= 1e-10, (i, w)
        assert _rel(b.array(i, "f"), o.array("f")) <= 1e-10
        o.close()
    b.close()


def test_batch_autotune_cache_reused(tmp_path, monkeypatch):
    from fdfault_amd import lib
    monkeypatch.setenv("FDF_TUNE_CACHE", str(tmp_path / "tune.txt"))
    paths = [_case(tmp_path, "c%d" % i, nx, ny, nt=5) for i, (nx, ny) in enumerate(SIZES[:3])]
    a = lib.Batch(paths, precision=lib.FP32)
    ta = a.tuning()
    a.close()
    assert not ta["from_cache"]
    b = lib.Batch(paths, precision=lib.FP32)
    tb = b.tuning()
    b.close()
    assert tb["from_cache"] and (tb["tile_width"], tb["chunk"], tb["persistent"]) == (
        ta["tile_width"], ta["chunk"], ta["persistent"])
    assert os.path.getsize(str(tmp_path / "tune.txt")) > 0


def test_batch_rejects_unsupported_instance(tmp_path):
    from fdfault_amd import lib
    good = _case(tmp_path, "g", 61, 31, nt=5)
    diss = _case(tmp_path, "d", 61, 31, nt=5, cdiss=0.1)
    with pytest.raises(lib.FdfaultError, match="not rectilinear"):
        lib.Batch([good, diss], precision=lib.FP32)
    s2 = _case(tmp_path, "s2", 61, 31, nt=5, sbporder=2)
    with pytest.raises(lib.FdfaultError, match="share sbporder and rkorder"):
        lib.Batch([good, s2], precision=lib.FP32)
    d3 = write_case(str(tmp_path), "d3", ndim=3, nx=25, ny=14, nz=17, dx=0.1, nt=5)
    with pytest.raises(lib.FdfaultError, match="instance 1 is not a 2-D mode II elastic problem"):
        lib.Batch([good, d3], precision=lib.FP32)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_mixed_friction_laws(tmp_path, monkeypatch, precision):
    """Instances with different friction laws (slip weakening, STZ,
    frictionless) in one batch take the fault launch that dispatches on each
    interface's law; every instance is bit-identical to its single run,
    STZ state included."""
    from cases import STZ_PARAMS
    from fdfault_amd import lib
    prec = lib.FP32 if precision == "fp32" else lib.FP64
    monkeypatch.setenv("FDF_BATCH_TW", "64")
    monkeypatch.setenv("FDF_BATCH_CX", "2")
    monkeypatch.setenv("FDF_BATCH_PERSIST", "0")
    cases = [("sw", 61, 31, {}), ("stz", 61, 31, dict(law="stz", stz=dict(STZ_PARAMS, chi0=0.07))),
             ("fl", 45, 27, dict(law="frictionless")), ("sw2", 97, 45, {})]
    paths = [_case(tmp_path, n, nx, ny, nt=25, **kw) for n, nx, ny, kw in cases]
    b = lib.Batch(paths, precision=prec, graph=True)
    b.step(25)
    monkeypatch.setenv("FDF_TILE2_TW", "64")
    monkeypatch.setenv("FDF_TILE2_CX", "2")
    monkeypatch.setenv("FDF_FUSE2D", "0")  # single runs with the separate fault launch, as the batch
    for i, p in enumerate(paths):
        s = lib.Solver(p, precision=prec)
        s.step(25)
        assert np.array_equal(b.array(i, "f"), s.array("f")), i
        ws = ("U", "V", "S", "Sn") + (("state",) if cases[i][0] == "stz" else ())
        for w in ws:
            assert np.array_equal(b.iface(i, 0, w), s.iface(0, w)), (i, w)
        assert np.array_equal(b.front(i, 0), s.front(0)), i
        s.close()
    b.close()
