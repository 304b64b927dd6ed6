"""Writes the fixtures of tests/golden/python/ with the reference's
unchanged Python setup and analysis module (python/fdfault of the reference
checkout).  Runs by hand, on a machine that has the reference checkout;
the tests read only the committed fixtures.

1. Runs the reference's example scripts problems/tpv5.py (3-D, cohesion
   barriers, on- and off-fault stations, a rupture front) and
   problems/input_example.py (2-D, a curved fault surface written by
   fdfault.curve, strided body outputs) unchanged, and keeps the .in files
   and binary side files they write.
2. Runs the CPU oracle on each, shortened to NT steps (cases.python_example),
   loads every output unit and the rupture fronts with
   fdfault.analysis.output / fdfault.analysis.front, and records each loaded
   array's shape, sum and index-weighted sum in analysis.json, so tests can
   check that the files this package and the oracle write are read by the
   reference's loaders exactly as the tests' own reader reads them.
"""
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "python")
NT = {"tpv5": 4, "input_example": 6}


def write_inputs():
    os.makedirs(OUT, exist_ok=True)
    sys.path.insert(0, os.path.join(REF, "python"))
    for name in NT:
        with tempfile.TemporaryDirectory() as d:
            cwd = os.getcwd()
            os.chdir(d)
            try:
                src = open(os.path.join(REF, "problems", name + ".py")).read()
                exec(compile(src, name + ".py", "exec"), {"__name__": "__main__"})
            finally:
                os.chdir(cwd)
            for fn in os.listdir(d):
                shutil.copy(os.path.join(d, fn), os.path.join(OUT, fn))
                print("wrote", fn)


def stats(a):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    w = np.arange(1, a.size + 1, dtype=np.float64)
    return {"shape": None, "sum": float(a.sum()), "wsum": float((w * a).sum())}


def record_outputs():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, ROOT)
    from cases import python_example
    from fdfault.analysis import front, output
    exe = os.path.join(ROOT, "oracle", "_build", "fdfault_oracle")
    res = {}
    for name, nt in NT.items():
        with tempfile.TemporaryDirectory() as d:
            data = os.path.join(d, "data") + "/"
            os.makedirs(data)
            p = python_example(d, name, nt, data)
            subprocess.run([exe, p], check=True, capture_output=True)
            entry = {"nt": nt, "outputs": {}, "front": None}
            for meta in sorted(glob.glob(os.path.join(data, name + "_*.py"))):
                unit = os.path.basename(meta)[len(name) + 1:-3]
                if unit.startswith("front_"):
                    continue
                o = output(name, unit, data)
                o.load()
                rec = {}
                for key, arr in (("data", o.fielddata), ("x", o.x), ("y", o.y)) + ((("z", o.z),) if hasattr(o, "z") else ()):
                    st = stats(arr)
                    st["shape"] = list(np.shape(arr))
                    rec[key] = st
                rec["field"] = o.field
                entry["outputs"][unit] = rec
            if os.path.exists(os.path.join(data, name + "_front_0.py")):
                fr = front(name, 0, data)
                fr.load()
                entry["front"] = {"t": dict(stats(fr.t), shape=list(np.shape(fr.t))),
                                  "x": dict(stats(fr.x), shape=list(np.shape(fr.x)))}
            res[name] = entry
            print(name, len(entry["outputs"]), "output units", "front" if entry["front"] else "")
    with open(os.path.join(OUT, "analysis.json"), "w") as fh:
        json.dump(res, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    write_inputs()
    record_outputs()
