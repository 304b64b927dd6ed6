"""FP32-bulk / FP64-fault path against the FP64 oracle on every kind of
block and fault the FP64 parity tests cover: mode III, sbporder 2 and 4 in
3-D, free and rigid outer faces, artificial dissipation, heterogeneous
stress and material, the STZ law in 2-D and 3-D, and plasticity in 3-D.
Each runs the bulk precision's own kernel (row tiles, 64-wide tiles, the
generic kernel) with the fault in FP64.

Bar: every field within 1e-4 of the oracle's (max-norm over its largest
value; the FP32 rounding of a few tens of steps is about 1e-6), final slip
within 1e-4 and rupture fronts within one time step, no point ruptured in
one run only."""
import numpy as np
import pytest

import oracle_lib as O
from cases import STZ_PARAMS, write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu

TOL = 1e-4
SHEAR = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
NUC2 = SHEAR + [Pert("boxcar", 0., 3.0, 0.6, 0., 0., (0., 11.6, 0.))]
NUC3 = SHEAR + [Pert("boxcar", 0., 1.5, 0.6, 1.2, 0.6, (0., 11.6, 0.))]


def _het_files(tmp, nx, ny):
    rng = np.random.default_rng(3)
    s = rng.normal(0., 2., size=(3, nx, 2 * ny, 1))
    sp = str(tmp / "hs.bin")
    s.tofile(sp)
    mat = np.empty((3, nx, 2 * ny, 1))
    for q, v in enumerate((2.67, 32.04, 32.04)):
        mat[q] = v * (1 + 0.05 * rng.random((nx, 2 * ny, 1)))
    mp = str(tmp / "mat.bin")
    mat.tofile(mp)
    return dict(hetstressfile=sp, matfile=mp)


def case_kw(tmp, name):
    d2 = dict(nx=61, ny=31, dx=0.1, nt=20)
    d3 = dict(ndim=3, nx=31, ny=16, nz=25, dx=0.1, nt=12)
    plastic = dict(mu=0.675, c=0., beta=0.35, eta=0.05 * 32.04)
    return {
        "mode3": dict(d2, mode=3, loads=[Pert("constant", 0., 0., 0., 0., 0., (-120., 0., 70.)),
                                         Pert("boxcar", 0., 3.0, 0.6, 0., 0., (0., 0., 11.6))]),
        "free_2d": dict(d2, bc="free", loads=NUC2),
        "rigid_2d": dict(d2, bc="rigid", loads=NUC2),
        "dissipation_2d": dict(d2, cdiss=0.1, loads=NUC2),
        "het_2d": dict(d2, loads=NUC2, **(_het_files(tmp, 61, 31) if name == "het_2d" else {})),
        "stz_2d": dict(d2, law="stz", stz=dict(STZ_PARAMS, chi0=0.07), loads=SHEAR),
        "sbp2_3d": dict(d3, sbporder=2, loads=NUC3),
        "sbp4_3d": dict(d3, sbporder=4, loads=NUC3),
        "dissipation_rigid_3d": dict(d3, cdiss=0.1, bc="rigid", loads=NUC3),
        "stz_3d": dict(d3, law="stz", stz=dict(STZ_PARAMS, chi0=0.07), loads=SHEAR),
        "plastic_3d": dict(d3, material="plastic", plastic=plastic, plastic_tensor=1, bc="free",
                           s0=(-100., 52., 0., -120., 0., -55.),
                           loads=[Pert("boxcar", 0., 1.5, 0.6, 1.2, 0.6, (0., 40., 0.))]),
    }[name]


CASES = ["mode3", "free_2d", "rigid_2d", "dissipation_2d", "het_2d", "stz_2d", "sbp2_3d", "sbp4_3d",
         "dissipation_rigid_3d", "stz_3d", "plastic_3d"]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", CASES)
def test_fp32_path_against_fp64_oracle(tmp_path, name):
    from fdfault_amd import lib
    kw = case_kw(tmp_path, name)
    nt = kw["nt"]
    p = write_case(str(tmp_path), name, front=("V", 0.001), **kw)
    print("case", name, flush=True)
    o = O.Oracle(p)
    g = lib.Solver(p, precision=lib.FP32)
    i = g.info
    print("blocks", i.nblocks, "tiled", i.tiled_blocks, "rows", i.row_blocks, "curv", i.curv_blocks, flush=True)
    o.step(nt)
    g.step(nt)
    fo, fg = o.array("f"), g.array("f")
    errs = [_rel(fg[q], fo[q]) for q in range(fo.shape[0]) if np.abs(fo[q]).max() > 0]
    U, Uo = g.iface(0, "U"), o.iface(0, "U")
    e_u = _rel(U, Uo)
    fr, fro = g.front(0), o.front(0)
    both = (fr >= 0) & (fro >= 0)
    e_front = np.abs(fr[both] - fro[both]).max() / o.dt if both.any() else 0.
    print("fields %.2e  slip %.2e  front %.2f steps  ruptured %d" % (max(errs), e_u, e_front, both.sum()), flush=True)
    assert np.abs(Uo).max() > 0  # the fault slips
    assert max(errs) <= TOL
    assert e_u <= TOL
    assert e_front <= 1.0 and ((fr >= 0) != (fro >= 0)).sum() == 0
    if name.startswith("stz"):
        assert _rel(g.iface(0, "state"), o.iface(0, "state")) <= TOL
    g.close()
    o.close()
