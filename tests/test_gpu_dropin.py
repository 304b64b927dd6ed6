"""The drop-in executable `fdfault_amd/_lib/fdfault <file.in>` against the
oracle's whole-run equivalent: same set of output and front files, same
metadata text, and data byte-identical in the bit-faithful FP64 mode
(FDFAULT_EXACT=1) or within 1e-10 in the default mode, whose rectilinear
blocks run on the fast tiled kernel (2-D mode II, field and interface units,
strided boxes, fronts)."""
import filecmp
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from cases import write_case
from fdfault_amd.problems import Output

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("exact", [True, False])
def test_executable_files_match_oracle(tmp_path, exact):
    outs = [Output("vx", "vx", (0, 20, 5), (0, 60, 2), (0, 61, 3), (0, 0, 1)),
            Output("sxy", "sxy", (0, 20, 1), (30, 30, 1), (20, 40, 1), (0, 0, 1)),
            Output("V", "V", (0, 20, 2), (0, 60, 1), (31, 31, 1), (0, 0, 1)),
            Output("U", "U", (0, 20, 20), (0, 60, 3), (30, 30, 1), (0, 0, 1)),
            Output("Sn", "Sn", (0, 20, 4), (10, 50, 1), (31, 31, 1), (0, 0, 1)),
            Output("Vx", "Vx", (5, 15, 5), (0, 60, 1), (31, 31, 1), (0, 0, 1))]
    p = write_case(str(tmp_path), "run", nx=61, ny=31, dx=0.1, nt=20, outputs=outs, front=("V", 0.001))
    d = os.path.dirname(p)
    gdir = os.path.join(d, "data")
    odir = os.path.join(d, "odata")
    os.makedirs(odir)
    txt = open(p).read().replace(gdir + "/", odir + "/")
    po = os.path.join(d, "run_oracle.in")
    open(po, "w").write(txt)
    assert O.run(po) == 0
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    env = dict(os.environ, FDFAULT_EXACT="1" if exact else "0")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0
    assert "Problem finished" in r.stdout
    gfiles, ofiles = sorted(os.listdir(gdir)), sorted(os.listdir(odir))
    assert gfiles == ofiles and len(gfiles) >= 30
    for fn in gfiles:
        a, b = os.path.join(gdir, fn), os.path.join(odir, fn)
        if exact or not fn.endswith(".dat") or fn.endswith(("_x.dat", "_y.dat", "_z.dat", "_t.dat")):
            assert filecmp.cmp(a, b, shallow=False), fn
        else:
            x, y = np.fromfile(a), np.fromfile(b)
            assert x.shape == y.shape, fn
            scale = max(np.abs(y).max(), 1e-300)
            assert np.abs(x - y).max() <= 1e-10 * scale, fn


@pytest.mark.parametrize("nt", [0, 1, 3])
def test_short_runs_match_oracle(tmp_path, nt):
    """Runs of 0, 1 and 3 steps: units whose last time index lies past nt are
    clipped to nt and then to a whole number of strides (outputunit.cpp:301-308),
    a unit with one time index writes no _t.dat file (outputunit.cpp:683), and
    a run of no steps writes the t = 0 outputs and the front file only."""
    outs = [Output("vx", "vx", (0, 20, 2), (0, 60, 4), (0, 61, 6), (0, 0, 1)),
            Output("V", "V", (0, 0, 1), (0, 60, 1), (31, 31, 1), (0, 0, 1)),
            Output("U", "U", (1, 9, 3), (0, 60, 2), (30, 30, 1), (0, 0, 1))]
    p = write_case(str(tmp_path), "short", nx=61, ny=31, dx=0.1, nt=nt, outputs=outs, front=("V", 0.001))
    d = os.path.dirname(p)
    gdir, odir = os.path.join(d, "data"), os.path.join(d, "odata")
    os.makedirs(odir)
    po = os.path.join(d, "short_oracle.in")
    open(po, "w").write(open(p).read().replace(gdir + "/", odir + "/"))
    assert O.run(po) == 0
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    print("nt", nt, flush=True)
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=120, env=dict(os.environ, FDFAULT_EXACT="1"))
    assert r.returncode == 0, r.stderr
    gfiles, ofiles = sorted(os.listdir(gdir)), sorted(os.listdir(odir))
    assert gfiles == ofiles, (gfiles, ofiles)
    # vx: time indices 0, 2, ... up to nt, so more than one only from nt = 2
    assert "short_V_t.dat" not in gfiles and ("short_vx_t.dat" in gfiles) == (nt >= 2)
    for fn in gfiles:
        assert filecmp.cmp(os.path.join(gdir, fn), os.path.join(odir, fn), shallow=False), fn


def test_executable_graph_replay_and_eager_write_the_same_files(tmp_path):
    """FDFAULT_GRAPH=0 issues every launch of a step one by one instead of
    replaying the step's hipGraph: the output and front files are
    byte-identical (2-D with the fault solve in the stage launch, and 3-D on
    the row kernel)."""
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    outs = [Output("V", "V", (0, 20, 1), (0, 60, 1), (31, 31, 1), (0, 0, 1)),
            Output("vx", "vx", (0, 20, 5), (0, 60, 2), (0, 61, 3), (0, 0, 1))]
    outs3 = [Output("V", "V", (0, 10, 1), (0, 30, 1), (21, 21, 1), (0, 24, 1))]
    for name, kw in (("g2", dict(nx=61, ny=31, nt=20, outputs=outs)),
                     ("g3", dict(ndim=3, nx=31, ny=21, nz=25, nt=10, sbporder=3, outputs=outs3))):
        dirs = []
        for graph in ("1", "0"):
            p = write_case(str(tmp_path / graph), name, dx=0.1, front=("V", 0.001), **kw)
            r = subprocess.run([exe, p], capture_output=True, text=True, timeout=120,
                               env=dict(os.environ, FDFAULT_GRAPH=graph))
            assert r.returncode == 0, r.stderr[-400:]
            dirs.append(os.path.join(os.path.dirname(p), "data"))
        fa, fb = sorted(os.listdir(dirs[0])), sorted(os.listdir(dirs[1]))
        assert fa == fb and len(fa) >= 8
        for fn in fa:
            a, b = os.path.join(dirs[0], fn), os.path.join(dirs[1], fn)
            if fn.endswith(".dat"):
                assert filecmp.cmp(a, b, shallow=False), (name, fn)


@pytest.mark.parametrize("slots", ["1", "2"])
def test_output_files_with_small_staging_rings(tmp_path, monkeypatch, slots):
    """Output units written through staging rings of one or two slots (the
    host waits for each unit's oldest copy on almost every output step) give
    the same files, byte for byte, as the default rings, eager and graph
    replayed."""
    from fdfault_amd import lib
    outs = [Output("vx", "vx", (0, 30, 1), (0, 60, 2), (0, 61, 3), (0, 0, 1)),
            Output("V", "V", (0, 30, 1), (0, 60, 1), (31, 31, 1), (0, 0, 1)),
            Output("sxy", "sxy", (3, 27, 4), (10, 50, 1), (25, 36, 1), (0, 0, 1))]
    dirs = []
    for k, (env, graph) in enumerate(((None, False), (slots, False), (slots, True))):
        if env is None:
            monkeypatch.delenv("FDF_OUTPUT_SLOTS", raising=False)
        else:
            monkeypatch.setenv("FDF_OUTPUT_SLOTS", env)
        p = write_case(str(tmp_path), "r%d" % k, nx=61, ny=31, dx=0.1, nt=30, outputs=outs, front=("V", 0.001))
        s = lib.Solver(p, graph=graph, write_outputs=True)
        s.step(30)
        s.finish()
        s.close()
        dirs.append((os.path.join(os.path.dirname(p), "data"), "r%d" % k))
    base, bname = dirs[0]
    files = sorted(os.listdir(base))
    assert len(files) >= 15
    for d, name in dirs[1:]:
        for fn in files:
            other = os.path.join(d, name + fn[len(bname):])
            assert filecmp.cmp(os.path.join(base, fn), other, shallow=False), (name, fn)


def test_executable_error_message(tmp_path):
    bad = tmp_path / "bad.in"
    bad.write_text("[fdfault.problem]\nx\ndata/\n1\n0\n0\n0.3\n1\n7\n")
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    r = subprocess.run([exe, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "Error in initializing RK method" in r.stderr


def test_batch_output_files_match_single_runs(tmp_path):
    """An ensemble writing its instances' output units (gathered on the batch
    stream, copied and written asynchronously) gives the same files, byte for
    byte, as each instance run on its own, over more output steps than an
    output unit's staging ring holds."""
    from fdfault_amd import lib
    outs = [Output("U", "U", (0, 24, 1), (0, 80, 1), (31, 31, 1), (0, 0, 1)),
            Output("vy", "vy", (0, 24, 3), (0, 80, 4), (0, 61, 2), (0, 0, 1)),
            Output("sxy", "sxy", (2, 22, 2), (10, 70, 1), (28, 34, 1), (0, 0, 1))]
    single, batch = [], []
    for i, nx in enumerate((81, 97)):
        for tag, lst in (("s", single), ("b", batch)):
            lst.append(write_case(str(tmp_path), "%s%d" % (tag, i), nx=nx, ny=31, dx=0.1, nt=24,
                                  outputs=outs, front=("V", 0.001)))
    bt = lib.Batch(batch, precision=lib.FP32, graph=True, write_outputs=True)
    bt.step(24)
    bt.finish()
    bt.close()
    for ps, pb in zip(single, batch):
        s = lib.Solver(ps, precision=lib.FP32, graph=True, write_outputs=True)
        s.step(24)
        s.finish()
        s.close()
        ds, db = os.path.join(os.path.dirname(ps), "data"), os.path.join(os.path.dirname(pb), "data")
        fs = sorted(os.listdir(ds))
        assert len(fs) >= 15
        for fn in fs:
            fb = fn.replace(os.path.basename(os.path.dirname(ps)), os.path.basename(os.path.dirname(pb)), 1)
            assert filecmp.cmp(os.path.join(ds, fn), os.path.join(db, fb), shallow=False), fn
