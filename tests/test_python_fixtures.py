"""The reference's Python-written examples (tests/golden/python, written by
the unchanged fdfault setup module from problems/tpv5.py and
problems/input_example.py): the library's host problem compiler parses them
exactly as the oracle does (bit-identical grid, metric, Jacobian, dt), and
the output and front files the oracle writes for them are read by the tests'
reader exactly as the reference's own loaders (fdfault.analysis) read them
(shapes, sums and index-weighted sums recorded by
tests/golden/make_python_fixtures.py)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import GOLDEN_PY, python_example
from python_outputs import compare, read_outputs

WANT = json.load(open(os.path.join(GOLDEN_PY, "analysis.json")))


@pytest.mark.parametrize("name", ["tpv5", "input_example"])
def test_python_example_parses_like_oracle(tmp_path, name):
    from fdfault_amd import lib
    p = python_example(str(tmp_path), name, 2, str(tmp_path) + "/data/")
    info, x, m, j = lib.host_setup(p)
    o = O.Oracle(p)
    assert info.dt == o.dt and info.nt == o.nt == 2
    np.testing.assert_array_equal(x, o.array("x"))
    np.testing.assert_array_equal(m, o.array("metric"))
    np.testing.assert_array_equal(j, o.array("jac")[0])


@pytest.mark.parametrize("name", ["tpv5", "input_example"])
def test_oracle_outputs_read_like_reference_loaders(tmp_path, name):
    want = WANT[name]
    data = str(tmp_path / "data") + "/"
    os.makedirs(data)
    p = python_example(str(tmp_path), name, want["nt"], data)
    assert O.run(p) == 0
    got, front = read_outputs(data, name)
    compare(got, front, want)
