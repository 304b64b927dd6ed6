"""The reference's Python-written examples (tests/golden/python: tpv5 in
3-D with cohesion barriers, 108 station outputs and a rupture front;
input_example in 2-D with a curved fault from fdfault.curve) through the
drop-in executable: the same output and front files as the oracle, byte-
identical in the bit-faithful mode and within 1e-10 in the default mode,
read by the tests' reader as the reference's own loaders read the oracle's
(tests/golden/python/analysis.json)."""
import filecmp
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from cases import GOLDEN_PY, python_example
from python_outputs import compare, read_outputs

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WANT = json.load(open(os.path.join(GOLDEN_PY, "analysis.json")))


@pytest.mark.parametrize("name", ["tpv5", "input_example"])
@pytest.mark.parametrize("exact", [True, False])
def test_python_example_dropin_matches_oracle(tmp_path, name, exact):
    nt = WANT[name]["nt"]
    gdir, odir = str(tmp_path / "gpu") + "/", str(tmp_path / "orc") + "/"
    os.makedirs(gdir)
    os.makedirs(odir)
    pg = python_example(str(tmp_path / "g"), name, nt, gdir)
    po = python_example(str(tmp_path / "o"), name, nt, odir)
    print("case", name, "nt", nt, "exact", exact, flush=True)
    assert O.run(po) == 0
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    env = dict(os.environ, FDFAULT_EXACT="1" if exact else "0")
    r = subprocess.run([exe, pg], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0 and "Problem finished" in r.stdout
    gfiles, ofiles = sorted(os.listdir(gdir)), sorted(os.listdir(odir))
    assert gfiles == ofiles
    # scale of each quantity (by its field name's letter: S traction, U
    # slip, V slip rate, v velocity, s stress) over every unit of the run:
    # station outputs that vanish by symmetry hold rounding noise only
    scale = {}
    for fn in ofiles:
        if fn.endswith(".dat") and not fn.endswith(("_x.dat", "_y.dat", "_z.dat", "_t.dat")) and "_front_" not in fn:
            q = fn[:-4].rsplit("_", 1)[1][:1]
            scale[q] = max(scale.get(q, 0.), float(np.abs(np.fromfile(os.path.join(odir, fn))).max()))
    for fn in gfiles:
        a, b = os.path.join(gdir, fn), os.path.join(odir, fn)
        if not fn.endswith(".dat"):
            assert open(a).read().replace(gdir, "<d>") == open(b).read().replace(odir, "<d>"), fn
        elif exact or fn.endswith(("_x.dat", "_y.dat", "_z.dat", "_t.dat")):
            assert filecmp.cmp(a, b, shallow=False), fn
        else:
            x, y = np.fromfile(a), np.fromfile(b)
            assert x.shape == y.shape, fn
            q = fn[:-4].rsplit("_", 1)[1][:1]
            assert np.abs(x - y).max() <= 1e-10 * max(scale.get(q, 0.), np.abs(y).max(), 1e-300), fn
    got, front = read_outputs(gdir, name)
    compare(got, front, WANT[name], rtol=1e-12 if exact else 1e-9, atol=0. if exact else 1e-12)
