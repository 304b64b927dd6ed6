"""FP64 parity over whole runs, each problem run to its end by the CPU
oracle, the fast GPU path (tiled kernels, graph replay) and the bit-faithful
GPU path (eager launches):

* the 2-D planar-fault workload (two blocks of 801 x 401 points, 2000
  steps, the run the slip-weakening benchmark is quoted on);
* a 3-D planar fault with a Gaussian-random-field shear load file (two
  blocks of 101 x 51 x 51 points, 1500 steps);
* a 2-D rough fault with Drucker-Prager plasticity on a curvilinear grid
  (two blocks of 401 x 201 points, sixth-order SBP, 1000 steps);
* a 3-D STZ fault with a seeded effective-temperature state file (two
  blocks of 61 x 31 x 31 points, 600 steps).

Fields, slip, slip rate, traction (and the STZ state) and rupture-front
times must agree within 1e-10 relative at the end of the run
(problem.cpp:136-192 driving domain.cpp:228-305 every step).

The oracle needs minutes of CPU time for these, so the tests run only with
FDF_LONG_RUNS=1 in the environment; each prints one JSON line with the
differences."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import STZ_PARAMS
from fdfault_amd.generators import fault_surfaces, gaussian_random_field, self_similar_profile, write_load_file
from fdfault_amd.problems import Pert, two_block_fault

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("FDF_LONG_RUNS") != "1",
                                 reason="whole-run oracle comparison takes minutes of CPU; set FDF_LONG_RUNS=1")]

TOL = 1e-10


def planar2d(d, nx=801, ny=401, dx=0.04, nt=2000):
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.)),
             Pert("boxcar", 0., 16.0, 1.5, 0., 0., (0., 11.6, 0.))]
    p = two_block_fault(ndim=2, nx=nx, ny=ny, dx=dx, nt=nt, sbporder=3, loads=loads,
                        datadir=os.path.join(d, "data") + "/", name="planar2d", front=("V", 0.001))
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    return p.write(os.path.join(d, "planar2d.in"))


def planar3d(d, nx=101, ny=51, nz=51, dx=0.4, nt=1500, seed=1234):
    tau = 70.0 + gaussian_random_field((nx, nz), dx, 2.0, 5.0, seed)
    lf = write_load_file(os.path.join(d, "load3d.bin"), np.full((nx, nz), -120.0), tau, np.zeros((nx, nz)))
    Lx, Lz = (nx - 1) * dx, (nz - 1) * dx
    loads = [Pert("boxcar", 0., 0.5 * Lx, 1.5, 0.5 * Lz, 1.5, (0., 11.6, 0.))]
    p = two_block_fault(ndim=3, nx=nx, ny=ny, nz=nz, dx=dx, nt=nt, sbporder=3, loads=loads, loadfile=lf,
                        datadir=os.path.join(d, "data") + "/", name="planar3d", front=("V", 0.001))
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    return p.write(os.path.join(d, "planar3d.in"))


def rough2d(d, nx=401, ny=201, dx=0.1, nt=1000, seed=7):
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    h = self_similar_profile(nx, dx, 1.0, 1e-2, 0.2, 20.0, seed)
    surf = fault_surfaces(d, nx, ny, dx, h)
    loads = [Pert("boxcar", 0., 0.5 * (nx - 1) * dx, 1.5, 0., 0., (0., 30., 0.))]
    p = two_block_fault(ndim=2, nx=nx, ny=ny, dx=dx, nt=nt, sbporder=4, loads=loads, surf=surf,
                        material="plastic", plastic=dict(mu=0.7, c=0., beta=0.35, eta=0.05 * 32.04),
                        s0=(-100., 52., 0., -120., 0., -55.), datadir=os.path.join(d, "data") + "/",
                        name="rough2d", front=("V", 0.001))
    return p.write(os.path.join(d, "rough2d.in"))


def stz3d(d, nx=61, ny=31, nz=31, dx=0.1, nt=600, seed=3):
    os.makedirs(os.path.join(d, "data"), exist_ok=True)
    chi = np.random.default_rng(seed).uniform(0.06, 0.08, size=(nx, nz))
    sp = os.path.join(d, "state.bin")
    chi.tofile(sp)
    stz = dict(STZ_PARAMS, chi0=0., statefile=sp)
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 70., 0.))]
    p = two_block_fault(ndim=3, nx=nx, ny=ny, nz=nz, dx=dx, nt=nt, sbporder=3, law="stz", stz=stz,
                        loads=loads, datadir=os.path.join(d, "data") + "/", name="stz3d", front=("V", 0.001))
    return p.write(os.path.join(d, "stz3d.in"))


WORKLOADS = {"planar2d_2x801x401_2000_steps": planar2d, "planar3d_2x101x51x51_1500_steps": planar3d,
             "rough2d_plastic_2x401x201_1000_steps": rough2d, "stz3d_2x61x31x31_600_steps": stz3d}


def _run(h, nt, label):
    done = 0
    while done < nt:
        k = min(250, nt - done)
        h.step(k)
        done += k
        print(label, "step", done, flush=True)
    out = dict(f=h.array("f"), U=h.iface(0, "U"), V=h.iface(0, "V"), S=h.iface(0, "S"), front=h.front(0))
    if label.startswith("stz"):
        out["state"] = h.iface(0, "state")
    return out


def _diff(a, b, dt):
    rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))
    both = (a["front"] >= 0) & (b["front"] >= 0)
    return {"fields": max(rel(a["f"][q], b["f"][q]) for q in range(b["f"].shape[0])),
            "slip": rel(a["U"], b["U"]), "slip_rate": rel(a["V"], b["V"]), "traction": rel(a["S"], b["S"]),
            "front_steps": float(np.abs(a["front"][both] - b["front"][both]).max() / dt) if both.any() else 0.,
            "ruptured_in_one_only": int(((a["front"] >= 0) != (b["front"] >= 0)).sum()),
            "ruptured": int(both.sum()),
            **({"state": rel(a["state"], b["state"])} if "state" in b else {})}


@pytest.mark.parametrize("case", list(WORKLOADS))
def test_whole_run_against_oracle(tmp_path, case):
    from fdfault_amd import lib
    path = WORKLOADS[case](str(tmp_path))
    o = O.Oracle(path)
    dt = o.dt
    fast = lib.Solver(path, graph=True)
    nt = fast.nt
    i = fast.info
    assert i.tiled_blocks + i.curv_blocks == i.nblocks  # the fast kernels take every block
    res = {"fast": _run(fast, nt, case + " fast")}
    fast.close()
    exact = lib.Solver(path, exact=True, graph=False)
    res["exact"] = _run(exact, nt, case + " bit-faithful")
    exact.close()
    res["oracle"] = _run(o, nt, case + " oracle")
    o.close()
    out = {"case": case, "steps": nt,
           "fast_vs_oracle": _diff(res["fast"], res["oracle"], dt),
           "bit_faithful_vs_oracle": _diff(res["exact"], res["oracle"], dt)}
    print(json.dumps(out), flush=True)
    assert out["fast_vs_oracle"]["ruptured"] > 0
    for k in ("fast_vs_oracle", "bit_faithful_vs_oracle"):
        r = out[k]
        for q in ("fields", "slip", "slip_rate", "traction", "state"):
            if q in r:
                assert r[q] <= TOL, (k, q, r[q])
        assert r["front_steps"] * dt <= TOL * max(1.0, nt * dt), k
        assert r["ruptured_in_one_only"] == 0, k
