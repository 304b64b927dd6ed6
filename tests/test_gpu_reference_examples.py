"""The reference's own example problems, problems/test2d.in (2-D mode II,
two 801 x 401 blocks, sbporder 4) and problems/tpv4.in (3-D, two
201 x 101 x 101 blocks, free surface, rupture front), kept verbatim under
tests/golden/, through the drop-in executable and through the oracle,
shortened to a few time steps: the same output and front files, the same
metadata text, data byte-identical in the bit-faithful FP64 mode and within
1e-10 in the default mode (tiled and row-tile kernels); the FP32-bulk mode
within its error bar."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from cases import reference_example

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name,nt", [("test2d", 40), ("tpv4", 10)])
@pytest.mark.parametrize("exact", [True, False])
def test_reference_example_files_match_oracle(tmp_path, name, nt, exact):
    gdir, odir = str(tmp_path / "gpu") + "/", str(tmp_path / "orc") + "/"
    os.makedirs(gdir)
    os.makedirs(odir)
    pg = reference_example(str(tmp_path / "g"), name, nt, gdir)
    po = reference_example(str(tmp_path / "o"), name, nt, odir)
    print("case", name, "nt", nt, "exact", exact, flush=True)
    assert O.run(po) == 0
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    env = dict(os.environ, FDFAULT_EXACT="1" if exact else "0")
    r = subprocess.run([exe, pg], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-400:], r.stderr[-400:])
    assert r.returncode == 0
    assert "Problem finished" in r.stdout
    gfiles, ofiles = sorted(os.listdir(gdir)), sorted(os.listdir(odir))
    assert gfiles == ofiles and len(gfiles) >= 6
    for fn in gfiles:
        a, b = os.path.join(gdir, fn), os.path.join(odir, fn)
        if not fn.endswith(".dat"):  # metadata text names the data directory
            ta, tb = open(a).read().replace(gdir, "<d>"), open(b).read().replace(odir, "<d>")
            assert ta == tb, fn
        elif exact or fn.endswith(("_x.dat", "_y.dat", "_z.dat", "_t.dat")):
            assert filecmp.cmp(a, b, shallow=False), fn
        else:
            x, y = np.fromfile(a), np.fromfile(b)
            assert x.shape == y.shape, fn
            scale = max(np.abs(y).max(), 1e-300)
            assert np.abs(x - y).max() <= 1e-10 * scale, fn


@pytest.mark.parametrize("name,nt", [("test2d", 60), ("tpv4", 30)])
def test_reference_example_fp32_dropin(tmp_path, name, nt):
    """The same examples through the drop-in executable with the FP32-bulk /
    FP64-fault path (FDFAULT_PRECISION=fp32): the same files and metadata, the
    coordinate and time files byte-identical, every output within 1e-4 of the
    FP64 oracle's relative to its largest value (at least 1e-6), and the
    rupture-front times within one time step, with the same points
    ruptured."""
    gdir, odir = str(tmp_path / "gpu") + "/", str(tmp_path / "orc") + "/"
    os.makedirs(gdir)
    os.makedirs(odir)
    pg = reference_example(str(tmp_path / "g"), name, nt, gdir)
    po = reference_example(str(tmp_path / "o"), name, nt, odir)
    assert O.run(po) == 0
    dt = O.Oracle(po).dt
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    r = subprocess.run([exe, pg], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, FDFAULT_PRECISION="fp32"))
    assert r.returncode == 0, r.stderr[-400:]
    gfiles, ofiles = sorted(os.listdir(gdir)), sorted(os.listdir(odir))
    assert gfiles == ofiles
    for fn in gfiles:
        a, b = os.path.join(gdir, fn), os.path.join(odir, fn)
        if not fn.endswith(".dat"):
            continue
        if "_front_" in fn and fn.endswith("_t.dat"):  # rupture times, -1 where none
            x, y = np.fromfile(a), np.fromfile(b)
            assert np.array_equal(x >= 0, y >= 0), fn
            both = (x >= 0) & (y >= 0)
            if both.any():
                assert np.abs(x[both] - y[both]).max() <= dt * (1 + 1e-9), fn
        elif fn.endswith(("_x.dat", "_y.dat", "_z.dat", "_t.dat")):
            assert filecmp.cmp(a, b, shallow=False), fn
        else:
            x, y = np.fromfile(a), np.fromfile(b)
            assert x.shape == y.shape, fn
            # a station the waves have not reached holds values far below
            # what FP32 fields resolve: an absolute floor of 1e-6 (m/s, MPa)
            assert np.abs(x - y).max() <= 1e-4 * max(np.abs(y).max(), 1e-6), fn
