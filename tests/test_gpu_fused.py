"""The fault solve inside the 2-D stage launch (bulk_fault2d_batch).

A 2-D problem whose one slip-weakening interface runs along x between two
rectilinear blocks solves each x-chunk of fault points in the workgroup that
then updates the chunk's two interface tiles, so a stage is one launch
instead of a fault launch followed by a bulk launch.  Every fault point and
every bulk point is still computed once with the same arithmetic, so the
results are bit-identical to the separate launches (FDF_FUSE2D=0; the
fused launch is the default for single problems and forced for ensembles
by FDF_FUSE2D=1): fields,
df, slip, slip rate, tractions and rupture fronts, in FP32 and FP64, eager
and graph-replayed, for every SBP order, for x-chunks that do not divide the
fault (a ragged last chunk), and for the ensemble in both of its launch
modes.  Problems the fused launch cannot take keep the separate launches
(fused_faults = 0)."""
import numpy as np
import pytest

import oracle_lib as O
from cases import write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu


def _nuc(nx, dx):
    Lx = (nx - 1) * dx
    return [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
            Pert("boxcar", 0., 0.5 * Lx, 0.6, 0., 0., (0., 11.6, 0.))]


def _case(tmp, name, nx, ny, nt=30, dx=0.1, **kw):
    kw.setdefault("s0", (0., 0., 0., 0., 0., 0.))
    return write_case(str(tmp), name, nx=nx, ny=ny, dx=dx, nt=nt, loads=_nuc(nx, dx), front=("V", 0.001), **kw)


def _state(s):
    out = {"f": s.array("f"), "df": s.array("df"), "front": s.front(0)}
    for w in ("U", "V", "S", "Sn", "Ux", "Vx", "Sx"):
        out[w] = s.iface(0, w)
    return out


def _same(a, b, what):
    for k in a:
        assert np.array_equal(a[k], b[k]), (what, k, np.abs(a[k] - b[k]).max())


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("sbp", [2, 3, 4])
@pytest.mark.parametrize("graph", [False, True])
def test_fused_equals_separate_launches(tmp_path, monkeypatch, precision, sbp, graph):
    from fdfault_amd import lib
    p = _case(tmp_path, "fz", 97, 45, nt=30, sbporder=sbp)
    res = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("FDF_FUSE2D", fuse)
        print("precision", precision, "sbp", sbp, "graph", graph, "fuse", fuse, flush=True)
        s = lib.Solver(p, precision=precision, graph=graph)
        assert s.info.tiled_blocks == 2
        assert s.info.fused_faults == (1 if fuse == "1" else 0)
        s.step(30)
        res.append(_state(s))
        s.close()
    assert np.abs(res[0]["V"]).max() > 0.  # the rupture runs
    _same(res[0], res[1], "fused vs separate")


@pytest.mark.parametrize("tw,cx", [(64, 1), (64, 3), (128, 5), (256, 8)])
def test_fused_tile_shapes(tmp_path, monkeypatch, tw, cx):
    """x-chunks of 1 to 8 planes, including ones that leave a ragged last
    chunk (97 points), with the interface row in a partial y tile."""
    from fdfault_amd import lib
    p = _case(tmp_path, "sh", 97, 53, nt=20)
    monkeypatch.setenv("FDF_TILE2_TW", str(tw))
    monkeypatch.setenv("FDF_TILE2_CX", str(cx))
    res = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("FDF_FUSE2D", fuse)
        print("tw", tw, "cx", cx, "fuse", fuse, flush=True)
        s = lib.Solver(p, precision=lib.FP32)
        assert s.info.fused_faults == (1 if fuse == "1" else 0)
        s.step(20)
        res.append(_state(s))
        s.close()
    _same(res[0], res[1], "tile shape %d x %d" % (tw, cx))


def test_fused_stage_matches_oracle(tmp_path):
    """One stage at a time from the t = 0 state through a rupture: the fused
    launch against the oracle (FP64, 1e-10)."""
    from fdfault_amd import lib
    p = _case(tmp_path, "or", 81, 41, nt=25)
    g = lib.Solver(p)
    assert g.info.fused_faults == 1
    o = O.Oracle(p)
    g.step(25)
    o.step(25)
    for w in ("U", "V", "S"):
        a, b = g.iface(0, w), o.iface(0, w)
        assert np.abs(a - b).max() <= 1e-10 * max(np.abs(b).max(), 1e-300), w
    fa, fb = g.array("f"), o.array("f")
    assert np.abs(fa - fb).max() <= 1e-10 * np.abs(fb).max()
    g.close()
    o.close()


@pytest.mark.parametrize("persist", ["1", "0"])
@pytest.mark.parametrize("precision", [0, 1])
def test_fused_batch_equals_separate_launches(tmp_path, monkeypatch, persist, precision):
    from fdfault_amd import lib
    sizes = [(61, 31), (97, 45), (45, 27), (130, 70)]
    paths = [_case(tmp_path, "b%d" % i, nx, ny, nt=20) for i, (nx, ny) in enumerate(sizes)]
    monkeypatch.setenv("FDF_BATCH_TW", "64")
    monkeypatch.setenv("FDF_BATCH_CX", "3")
    monkeypatch.setenv("FDF_BATCH_PERSIST", persist)
    res = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("FDF_FUSE2D", fuse)
        print("batch persist", persist, "precision", precision, "fuse", fuse, flush=True)
        b = lib.Batch(paths, precision=precision, graph=True)
        assert all(i.fused_faults == (1 if fuse == "1" else 0) for i in b.infos)
        b.step(20)
        res.append([(b.array(i, "f"), b.iface(i, 0, "U"), b.iface(i, 0, "V"), b.front(i, 0))
                    for i in range(len(paths))])
        b.close()
    for i, (x, y) in enumerate(zip(res[0], res[1])):
        for a, c in zip(x, y):
            assert np.array_equal(a, c), i


def test_fused_default_per_precision(tmp_path, monkeypatch):
    from fdfault_amd import lib
    monkeypatch.delenv("FDF_FUSE2D", raising=False)
    p = _case(tmp_path, "df", 61, 31, nt=2)
    a = lib.Solver(p, precision=lib.FP64)
    b = lib.Solver(p, precision=lib.FP32)
    assert a.info.fused_faults == 1 and b.info.fused_faults == 1
    a.close()
    b.close()
    c = lib.Batch([p, p], precision=lib.FP64)
    assert all(i.fused_faults == 0 for i in c.infos)
    c.close()


def test_unfusable_problems_keep_separate_launches(tmp_path):
    """A frictionless interface, or a problem on the generic kernel, keeps
    the fault launch of its own."""
    from fdfault_amd import lib
    a = lib.Solver(_case(tmp_path, "fl", 61, 31, nt=2, law="frictionless"))
    assert a.info.fused_faults == 0 and a.info.tiled_blocks == 2
    a.close()
    b = lib.Solver(_case(tmp_path, "ex", 61, 31, nt=2), exact=True)
    assert b.info.fused_faults == 0 and b.info.tiled_blocks == 0
    b.close()
