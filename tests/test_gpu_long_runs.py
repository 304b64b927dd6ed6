"""FP64 parity over whole runs: the 2-D planar-fault workload (two blocks of
801 x 401 points, 2000 steps, the run the slip-weakening benchmark is quoted
on) and a 3-D planar fault with a Gaussian-random-field shear load file
(two blocks of 101 x 51 x 51 points, 1500 steps), each run by the CPU oracle,
the fast GPU path (tiled kernels, graph replay) and the bit-faithful GPU path
(eager launches).  Fields, slip, slip rate, traction and rupture-front
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
from fdfault_amd.generators import gaussian_random_field, write_load_file
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


def _run(h, nt, label):
    done = 0
    while done < nt:
        k = min(250, nt - done)
        h.step(k)
        done += k
        print(label, "step", done, flush=True)
    return dict(f=h.array("f"), U=h.iface(0, "U"), V=h.iface(0, "V"), S=h.iface(0, "S"), front=h.front(0))


def _diff(a, b, dt):
    rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300))
    both = (a["front"] >= 0) & (b["front"] >= 0)
    return {"fields": max(rel(a["f"][q], b["f"][q]) for q in range(b["f"].shape[0])),
            "slip": rel(a["U"], b["U"]), "slip_rate": rel(a["V"], b["V"]), "traction": rel(a["S"], b["S"]),
            "front_steps": float(np.abs(a["front"][both] - b["front"][both]).max() / dt) if both.any() else 0.,
            "ruptured_in_one_only": int(((a["front"] >= 0) != (b["front"] >= 0)).sum()),
            "ruptured": int(both.sum())}


@pytest.mark.parametrize("case", ["planar2d_2x801x401_2000_steps", "planar3d_2x101x51x51_1500_steps"])
def test_whole_run_against_oracle(tmp_path, case):
    from fdfault_amd import lib
    path = planar2d(str(tmp_path)) if case.startswith("planar2d") else planar3d(str(tmp_path))
    o = O.Oracle(path)
    nt, dt = None, o.dt
    fast = lib.Solver(path, graph=True)
    nt = fast.nt
    assert fast.info.tiled_blocks == fast.info.nblocks
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
        for q in ("fields", "slip", "slip_rate", "traction"):
            assert r[q] <= TOL, (k, q, r[q])
        assert r["front_steps"] * dt <= TOL * max(1.0, nt * dt), k
        assert r["ruptured_in_one_only"] == 0, k
