"""State snapshot and resume (fdf_save_state / fdf_load_state; the reference
has no restart).  A run saved after n1 steps and continued for n2 more
matches, bit for bit, a fresh handle on the same problem that loads the
snapshot and steps n2: fields, df, plastic fields, every interface array and
the rupture fronts.  Snapshots of another problem or precision, truncated
files and other files are refused and leave the handle as it was."""
import os

import numpy as np
import pytest

from cases import STZ_PARAMS, write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu

IFACE = ("U", "V", "S", "Sn", "Ux", "Vx", "Sx", "du", "dux")


def _state(s, stz=False):
    out = {"f": s.array("f"), "df": s.array("df")}
    for i in range(s.info.nifaces):
        for w in IFACE + (("state", "dstate") if stz else ()):
            out["%d_%s" % (i, w)] = s.iface(i, w)
        fr = s.front(i)
        if fr is not None:
            out["%d_front" % i] = fr
    return out


CASES = {
    "planar2d": dict(nx=81, ny=41, dx=0.1),
    "stz2d": dict(nx=61, ny=31, dx=0.1, law="stz", stz=dict(STZ_PARAMS, chi0=0.07)),
    "plastic2d": dict(nx=61, ny=31, dx=0.1, plastic=dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04),
                      s0=(-100., 52., 0., -120., 0., -55.),
                      loads=[Pert("boxcar", 0., 3.0, 0.6, 0., 0., (0., 40., 0.))]),
    "planar3d": dict(ndim=3, nx=31, ny=21, nz=25, dx=0.1, sbporder=3),
}


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", sorted(CASES))
def test_resume_is_bit_identical(tmp_path, case, precision, graph):
    from fdfault_amd import lib
    kw = dict(CASES[case])
    if kw.get("plastic"):
        kw["material"] = "plastic"
    p = write_case(str(tmp_path), case, nt=40, **kw)
    snap = str(tmp_path / "snap.bin")
    print("case", case, "precision", precision, "graph", graph, flush=True)
    a = lib.Solver(p, precision=precision, graph=graph)
    a.step(12)
    a.save_state(snap)
    a.step(9)
    want = _state(a, stz=case == "stz2d")
    a.close()
    b = lib.Solver(p, precision=precision, graph=graph)
    b.load_state(snap)
    b.step(9)
    got = _state(b, stz=case == "stz2d")
    b.close()
    assert np.abs(want["0_V"]).max() > 0.  # the rupture runs
    for k in want:
        assert np.array_equal(want[k], got[k]), (k, np.abs(want[k] - got[k]).max())


def test_snapshot_refusals(tmp_path):
    from fdfault_amd import lib
    p = write_case(str(tmp_path), "a", nt=10, nx=61, ny=31, dx=0.1)
    q = write_case(str(tmp_path), "b", nt=10, nx=41, ny=31, dx=0.1)
    snap = str(tmp_path / "snap.bin")
    a = lib.Solver(p)
    a.step(3)
    a.save_state(snap)
    before = a.array("f")
    with pytest.raises(lib.FdfaultError, match="another problem"):
        lib.Solver(q).load_state(snap)
    with pytest.raises(lib.FdfaultError, match="another problem or precision"):
        lib.Solver(p, precision=lib.FP32).load_state(snap)
    cut = str(tmp_path / "cut.bin")
    with open(snap, "rb") as fi, open(cut, "wb") as fo:
        fo.write(fi.read()[: os.path.getsize(snap) // 2])
    with pytest.raises(lib.FdfaultError, match="truncated"):
        a.load_state(cut)
    other = str(tmp_path / "other.bin")
    np.zeros(16).tofile(other)
    with pytest.raises(lib.FdfaultError, match="not a state snapshot"):
        a.load_state(other)
    bad = str(tmp_path / "bad_step.bin")
    raw = bytearray(open(snap, "rb").read())
    raw[56:64] = np.int64(-1).tobytes()  # the step count, after the tag and six header words
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(lib.FdfaultError, match="not a state snapshot"):
        a.load_state(bad)
    assert np.array_equal(a.array("f"), before)  # refused loads left the state alone
    a.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_resume_multi_interface_layout(tmp_path, precision):
    """Four interfaces of three laws (slip weakening, locked, frictionless)
    on a 2 x 2 block grid: the snapshot carries every frictional
    interface's arrays, and the resumed run is bit-identical."""
    from fdfault_amd import lib
    from test_gpu_layouts import LAYOUTS, layout
    p = layout(tmp_path, "grid", nt=30, **LAYOUTS["grid_2x2"])
    snap = str(tmp_path / "snap.bin")

    def state(s):
        out = {"f": s.array("f"), "df": s.array("df")}
        for i in range(s.info.nifaces):
            for w in ("U", "V", "S", "Sn", "du"):
                try:
                    out["%d_%s" % (i, w)] = s.iface(i, w)
                except lib.FdfaultError:
                    pass  # the locked interface has no slip arrays
        return out

    a = lib.Solver(p, precision=precision)
    a.step(10)
    a.save_state(snap)
    a.step(8)
    want = state(a)
    a.close()
    b = lib.Solver(p, precision=precision)
    b.load_state(snap)
    b.step(8)
    got = state(b)
    b.close()
    assert len(want) >= 2 + 3 * 5
    for k in want:
        assert np.array_equal(want[k], got[k]), k
