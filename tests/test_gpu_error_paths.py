"""Run-time error paths of the fault solve, forced from the inputs so that
every kernel still exits normally: the kernel sets a flag, the host reads it
after the step and stops with the reference's message (stz.cpp:341-344:
"Slip velocity exceeds melting value in stz.cpp" before MPI_Abort).  The
oracle stops at the same step with the same message; the drop-in executable
prints it and exits non-zero; a batch instance reports it too."""
import os
import subprocess

import pytest

import oracle_lib as O
from cases import STZ_PARAMS, write_case
from fdfault_amd.problems import Pert

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MELT = "Slip velocity exceeds melting value in stz.cpp"


def _melting_case(tmp, name, ramp=0.05):
    # the STZ law of the other tests with the melting velocity V1 at 1e-12 m/s:
    # the first slipping point exceeds it.  With the shear load ramped up over
    # `ramp` seconds (pert.cpp:45-57) nothing slips at t = 0, so the run opens
    # and stops in its first steps; with ramp = 0 the t = 0 interface solve
    # (domain.cpp:141-143) already stops it.
    v = (1.e-6, 0.3, 0.01, 0.1, 1.e4, 1.e-3, 0.1, 0.4, 1.e-12)
    stz = dict(STZ_PARAMS, chi0=0.07, perts=[Pert("constant", 0., 0., 0., 0., 0., v)])
    loads = [Pert("constant", 0., 0., 0., 0., 0., (-120., 0., 0.)),
             Pert("constant", ramp, 0., 0., 0., 0., (0., 70., 0.))]
    return write_case(str(tmp), name, nx=61, ny=31, dx=0.1, nt=10, law="stz", stz=stz, loads=loads)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("graph", [False, True])
def test_melting_velocity_stops_the_run(tmp_path, precision, graph):
    from fdfault_amd import lib
    p = _melting_case(tmp_path, "melt")
    o = O.Oracle(p)
    with pytest.raises(O.OracleError) as eo:
        o.step(10)
    assert str(eo.value) == MELT
    o.close()
    print("precision", precision, "graph", graph, flush=True)
    g = lib.Solver(p, precision=precision, graph=graph)
    with pytest.raises(lib.FdfaultError) as eg:
        g.step(10)
    assert str(eg.value) == MELT
    g.close()


def test_melting_velocity_at_open(tmp_path):
    from fdfault_amd import lib
    p = _melting_case(tmp_path, "melt0", ramp=0.)
    with pytest.raises(O.OracleError) as eo:
        O.Oracle(p)
    assert str(eo.value) == MELT
    with pytest.raises(lib.FdfaultError) as eg:
        lib.Solver(p)
    assert str(eg.value) == MELT


def test_melting_velocity_executable(tmp_path):
    p = _melting_case(tmp_path, "meltx")
    exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
    r = subprocess.run([exe, p], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert MELT in r.stderr


def test_melting_velocity_in_a_batch(tmp_path):
    from fdfault_amd import lib
    ok = write_case(str(tmp_path), "ok", nx=61, ny=31, dx=0.1, nt=10, law="stz", stz=dict(STZ_PARAMS, chi0=0.07))
    bad = _melting_case(tmp_path, "meltb")
    b = lib.Batch([ok, bad], precision=lib.FP32, graph=True)
    with pytest.raises(lib.FdfaultError) as e:
        b.step(10)
    assert MELT in str(e.value)
    b.close()


def test_folded_grid_on_the_device(tmp_path):
    """The grid built on the device stops on a folded block as the host
    compiler and the oracle do (tests/test_error_paths.py)."""
    from fdfault_amd import lib
    from test_error_paths import JAC, folded_case
    p = folded_case(tmp_path)
    with pytest.raises(lib.FdfaultError) as e:
        lib.Solver(p)
    assert str(e.value) == JAC


# Output-list errors, read only when outputs are written (outputlist.cpp:30-38,
# outputunit.cpp:13-343): the library opened with outputs on and the oracle's
# whole run stop with the reference's message.
OUTPUT_ERRORS = {
    "no_outputlist": (("[fdfault.outputlist]", "[fdfault.outputlistX]"), "Error reading outputlist from input file"),
    "bad_field": (("[fdfault.outputlist]\n\n", "[fdfault.outputlist]\nout\nvq\n0\n0\n1\n0\n40\n1\n0\n0\n1\n0\n0\n1\n\n"),
                  "Error in specifying output field in outputunit.cpp"),
    "no_state": (("[fdfault.outputlist]\n\n", "[fdfault.outputlist]\nout\nstate\n0\n0\n1\n0\n40\n1\n21\n21\n1\n0\n0\n1\n\n"),
                 "Interface for output does not have a state variable in outputunit.cpp"),
}


@pytest.mark.parametrize("case", sorted(OUTPUT_ERRORS))
def test_output_list_errors(tmp_path, case):
    from fdfault_amd import lib
    from test_error_paths import edited_case
    edit, msg = OUTPUT_ERRORS[case]
    p = edited_case(tmp_path, case, edit)
    assert O.run(p) != 0 and O.lib().orc_last_error().decode() == msg
    print("case", case, flush=True)
    with pytest.raises(lib.FdfaultError) as e:
        lib.Solver(p, write_outputs=True)
    assert str(e.value) == msg
    if case == "bad_field":
        exe = os.path.join(ROOT, "fdfault_amd", "_lib", "fdfault")
        r = subprocess.run([exe, p], capture_output=True, text=True, timeout=120)
        assert r.returncode != 0 and msg in r.stderr
