"""Reader of output and front files by their .py metadata (the layouts of
outputunit.cpp:546-889 and front.cpp:127-359), and the statistics that
tests/golden/make_python_fixtures.py records from the reference's own
loaders (fdfault.analysis.output / front), for comparison."""
import glob
import os

import numpy as np


def meta(path):
    m = {}
    exec(open(path).read(), m)
    return m


def stats(a):
    flat = np.ascontiguousarray(a, dtype=np.float64).ravel()
    w = np.arange(1, flat.size + 1, dtype=np.float64)
    return {"shape": list(np.shape(a)), "sum": float(flat.sum()), "wsum": float((w * flat).sum())}


def read_outputs(data, name):
    """{unit: {"field", "data", "x", "y"[, "z"]} statistics} of every output
    unit of problem `name` in directory `data`, and the front statistics."""
    res = {}
    for mp in sorted(glob.glob(os.path.join(data, name + "_*.py"))):
        unit = os.path.basename(mp)[len(name) + 1:-3]
        if unit.startswith("front_"):
            continue
        m = meta(mp)
        shp = (m["nx"], m["ny"], m["nz"])
        base = os.path.join(data, name + "_" + unit + "_")
        rec = {"field": m["field"],
               "data": stats(np.squeeze(np.fromfile(base + m["field"] + ".dat").reshape((m["nt"],) + shp)))}
        for c in ("x", "y", "z"):
            if os.path.exists(base + c + ".dat"):
                rec[c] = stats(np.squeeze(np.fromfile(base + c + ".dat").reshape(shp)))
        res[unit] = rec
    front = None
    fp = os.path.join(data, name + "_front_0.py")
    if os.path.exists(fp):
        m = meta(fp)
        base = os.path.join(data, name + "_front_0_")
        front = {c: stats(np.squeeze(np.fromfile(base + c + ".dat").reshape(m["nx"], m["ny"]))) for c in ("t", "x")}
    return res, front


def compare(got, front, want, rtol=1e-12, atol=0.):
    assert sorted(got) == sorted(want["outputs"]), "output units differ"
    for unit, rec in want["outputs"].items():
        g = got[unit]
        assert g["field"] == rec["field"], unit
        for key in ("data", "x", "y", "z"):
            if key not in rec:
                continue
            assert g[key]["shape"] == rec[key]["shape"], (unit, key)
            for s in ("sum", "wsum"):
                a, b = g[key][s], rec[key][s]
                assert abs(a - b) <= rtol * abs(b) + atol, (unit, key, s, a, b)
    if want["front"] is None:
        assert front is None
    else:
        for c in ("t", "x"):
            assert front[c]["shape"] == want["front"][c]["shape"]
            for s in ("sum", "wsum"):
                a, b = front[c][s], want["front"][c][s]
                assert abs(a - b) <= rtol * abs(b) + atol, (c, s, a, b)
