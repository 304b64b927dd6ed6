"""Field arrays above 4 GiB.  The tiled kernels address the field arrays
through buffer descriptors with 32-bit offsets: one descriptor over all nf
fields while nf N elements fit in 4 GiB, else one descriptor per field (each
field within 4 GiB).

* Forced on small problems (FDF_WIDE=1), the per-field form must be
  bit-identical to the single-descriptor form on every tiled kernel: row
  tiles and 64-wide tiles in 3-D (FP64 and FP32), 2-D rectilinear blocks in
  one merged launch and one launch per block, curvilinear and plastic 2-D
  blocks.
* On problems just above the old limit (nf N w >= 4 GiB) the fast path must
  be taken, with one descriptor per field, and agree with the bit-faithful
  FP64 path (which the oracle pins bit-for-bit at small sizes) field by
  field, fault slip and slip rate included.
"""
import numpy as np
import pytest

from cases import curved_surfaces, write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu

GIB4 = 4294967296


def _nuc(x0, z0=0., w=0.6, ndim=2):
    return [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
            Pert("boxcar", 0., x0, w, z0, w if ndim == 3 else 0., (0., 30., 0.))]


def _small_cases(tmp):
    plastic = dict(mu=0.68, c=0., beta=0.35, eta=0.05 * 32.04)
    surf = curved_surfaces(str(tmp), 61, 31, 0.1, amp=0.1)
    return {
        # (path, precision, environment)
        "rows3d_fp64": (write_case(str(tmp), "r64", ndim=3, nx=33, ny=17, nz=21, dx=0.1, nt=6,
                                   loads=_nuc(1.6, 1.0, ndim=3)), 0, {"FDF_ROWS_SHAPE": "0"}),
        # (per-field descriptors take the 4-row FP32 / 2-row FP64 workgroups;
        # rounding in the one-sided y rows depends on the workgroup's row
        # count, so the single-descriptor runs take them as well)
        "rows3d_fp32": (write_case(str(tmp), "r32", ndim=3, nx=33, ny=17, nz=21, dx=0.1, nt=6, sbporder=3,
                                   loads=_nuc(1.6, 1.0, ndim=3)), 1, {"FDF_ROWS_SHAPE": "0"}),
        "tiled3d_fp64": (write_case(str(tmp), "t64", ndim=3, nx=33, ny=17, nz=21, dx=0.1, nt=6,
                                    loads=_nuc(1.6, 1.0, ndim=3)), 0, {"FDF_ROWS": "0"}),
        "tiled3d_fp32": (write_case(str(tmp), "t32", ndim=3, nx=33, ny=17, nz=21, dx=0.1, nt=6,
                                    loads=_nuc(1.6, 1.0, ndim=3)), 1, {"FDF_ROWS": "0"}),
        "merged2d": (write_case(str(tmp), "m2", nx=81, ny=41, dx=0.1, nt=6, loads=_nuc(4.0)), 0, {}),
        "perblock2d_fp32": (write_case(str(tmp), "p2", nx=81, ny=41, dx=0.1, nt=6, loads=_nuc(4.0)), 1,
                            {"FDF_MERGE2D": "0"}),
        "curv2d": (write_case(str(tmp), "c2", nx=61, ny=31, dx=0.1, nt=6, loads=_nuc(3.0), surf=surf), 0, {}),
        "curv2d_plastic_perblock": (write_case(str(tmp), "cp", nx=61, ny=31, dx=0.1, nt=6, loads=_nuc(3.0),
                                               surf=surf, material="plastic", plastic=plastic,
                                               s0=(-100., 52., 0., -120., 0., -55.)), 1, {"FDF_MERGE2D": "0"}),
    }


@pytest.mark.parametrize("case", ["rows3d_fp64", "rows3d_fp32", "tiled3d_fp64", "tiled3d_fp32", "merged2d",
                                  "perblock2d_fp32", "curv2d", "curv2d_plastic_perblock"])
def test_per_field_descriptors_bit_identical(tmp_path, monkeypatch, case):
    from fdfault_amd import lib
    path, prec, env = _small_cases(tmp_path)[case]
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    a = lib.Solver(path, precision=prec)
    monkeypatch.setenv("FDF_WIDE", "1")
    b = lib.Solver(path, precision=prec)
    monkeypatch.delenv("FDF_WIDE")
    nb = a.info.nblocks
    assert a.info.tiled_blocks + a.info.curv_blocks == nb and a.info.wide_blocks == 0
    assert b.info.tiled_blocks + b.info.curv_blocks == nb and b.info.wide_blocks == nb
    if case.startswith("rows"):
        assert a.info.row_blocks == nb and b.info.row_blocks == nb
    if case.startswith("tiled3d"):
        assert a.info.row_blocks == 0 and b.info.row_blocks == 0
    a.step(6)
    b.step(6)
    np.testing.assert_array_equal(a.array("f"), b.array("f"))
    np.testing.assert_array_equal(a.array("df"), b.array("df"))
    for w in ("U", "V", "S"):
        np.testing.assert_array_equal(a.iface(0, w), b.iface(0, w))
    a.close()
    b.close()


def _compare_large(path, prec, nsteps, nx, rtol):
    """The fast path on a problem above the old limit against the
    bit-faithful FP64 path, field by field (each read on its own: a whole f
    array of these problems is many GB)."""
    from fdfault_amd import lib
    g = lib.Solver(path, precision=prec)
    i = g.info
    w = 8 if prec == lib.FP64 else 4
    assert i.nfields * g.N * w >= GIB4 > g.N * w
    assert i.tiled_blocks == i.nblocks and i.wide_blocks == i.nblocks
    assert i.row_blocks == (i.nblocks if i.ndim == 3 else 0)
    e = lib.Solver(path, precision=lib.FP64, exact=True)
    assert e.info.tiled_blocks == 0 and e.info.wide_blocks == 0
    g.step(nsteps)
    e.step(nsteps)
    hi = [n - 1 for n in nx]
    # tolerance relative to the largest velocity or stress of the run
    peak = [np.abs(e.field(q, (0, 0, 0), hi)).max() for q in range(i.nfields)]
    for q in range(i.nfields):
        scale = max(peak[:i.ndim]) if q < i.ndim else max(peak[i.ndim:])
        assert scale > 0., q
        x = g.field(q, (0, 0, 0), hi)
        y = e.field(q, (0, 0, 0), hi)
        assert np.abs(x - y).max() <= rtol * scale, (q, np.abs(x - y).max(), scale)
        del x, y
    for wq in ("U", "V", "S"):
        x, y = g.iface(0, wq), e.iface(0, wq)
        assert np.abs(x - y).max() <= rtol * max(np.abs(y).max(), 1e-300), wq
    g.close()
    e.close()


def test_above_old_limit_3d_fp64_row_tiles(tmp_path):
    """2 x 1030 x 121 x 241 points (60.1 M): 9 N doubles are 4.33 GB."""
    from fdfault_amd import lib
    nx, ny, nz = 1030, 121, 241
    p = write_case(str(tmp_path), "l64", ndim=3, nx=nx, ny=ny, nz=nz, dx=0.1, nt=3,
                   loads=_nuc(51.4, 12.0, w=1.5, ndim=3))
    _compare_large(p, lib.FP64, 3, (nx, 2 * ny, nz), 1e-11)


def test_above_old_limit_3d_fp32_row_tiles(tmp_path):
    """2 x 2060 x 121 x 241 points (120.1 M): 9 N floats are 4.33 GB."""
    from fdfault_amd import lib
    nx, ny, nz = 2060, 121, 241
    p = write_case(str(tmp_path), "l32", ndim=3, nx=nx, ny=ny, nz=nz, dx=0.1, nt=3, sbporder=3,
                   loads=_nuc(102.9, 12.0, w=1.5, ndim=3))
    _compare_large(p, lib.FP32, 3, (nx, 2 * ny, nz), 2e-5)


def test_above_old_limit_2d_fp64_merged(tmp_path):
    """2 x 10400 x 5200 points (108.2 M): 5 N doubles are 4.33 GB."""
    from fdfault_amd import lib
    nx, ny = 10400, 5200
    p = write_case(str(tmp_path), "l2", nx=nx, ny=ny, dx=0.01, nt=3, loads=_nuc(52.0, w=0.3))
    _compare_large(p, lib.FP64, 3, (nx, 2 * ny, 1), 1e-11)
