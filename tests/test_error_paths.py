"""Error paths of the input readers, on the CPU: a binary input file that
cannot be opened stops the set-up with the reference's message, in the
library's host problem compiler (fdf_host_setup) and in the oracle alike.
The reference prints these lines before MPI_Abort: fields.cpp:497 (stress
file), fields.cpp:627 (material file), friction.cpp:555 (load file),
friction.cpp:627 (state file), slipweak.cpp:358 and stz.cpp:406 (parameter
files), surface.cpp:44 (surface file).  A file that opens but is too short
is an error too (the reference would read past its end).  A folded block
grid (a non-positive Jacobian), which the reference leaves to an assert in
its metric code, stops the set-up with a message."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O
from cases import STZ_PARAMS, curved_surfaces, write_case
from fdfault_amd import lib

MISSING = "/nonexistent/fdfault_input.bin"


def _host_setup_error(path):
    L = lib.load()
    info = lib.Info()
    rc = L.fdf_host_setup(path.encode(), ctypes.byref(info), None, None, None)
    assert rc < 0
    return L.fdf_last_error().decode()


def _oracle_error(path):
    with pytest.raises(O.OracleError) as e:
        O.Oracle(path)
    return str(e.value)


CASES = {
    "stress": (dict(hetstressfile=MISSING), "Error opening stress file in fields.cpp"),
    "material": (dict(matfile=MISSING), "Error opening material file in fields.cpp"),
    "load": (dict(loadfile=MISSING), "Error opening load file in friction.cpp"),
    "sw_param": (dict(paramfile=MISSING), "Error opening file in friction.cpp"),
    "stz_param": (dict(law="stz", stz=dict(STZ_PARAMS, chi0=0.07, paramfile=MISSING)), "Error opening file in stz.cpp"),
    "stz_state": (dict(law="stz", stz=dict(STZ_PARAMS, chi0=0., statefile=MISSING)),
                  "Error opening state file in friction.cpp"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_missing_input_file_message(tmp_path, case):
    kw, msg = CASES[case]
    p = write_case(str(tmp_path), case, nx=41, ny=21, dx=0.1, nt=1, **kw)
    assert _host_setup_error(p) == msg
    assert _oracle_error(p) == msg


def test_missing_surface_file_message(tmp_path):
    surf = [[MISSING, "none", "none", "none"], ["none", "none", "none", "none"]]
    p = write_case(str(tmp_path), "surf", nx=41, ny=21, dx=0.1, nt=1, surf=surf)
    msg = "Error reading surface from file " + MISSING
    assert _host_setup_error(p) == msg
    assert _oracle_error(p) == msg


def test_short_input_file(tmp_path):
    short = str(tmp_path / "short.bin")
    np.zeros(5).tofile(short)  # a load file needs 3 x 41 values
    p = write_case(str(tmp_path), "short", nx=41, ny=21, dx=0.1, nt=1, loadfile=short)
    msg = "Error reading load from file " + short
    assert _host_setup_error(p) == msg
    assert _oracle_error(p) == msg
    assert os.path.getsize(short) == 40


JAC = "non-positive Jacobian in block grid (grid is not smooth enough)"


def folded_case(tmp):
    """A fault curve that dips below the lower block's bottom face
    (amplitude 1.5 block heights): that block's grid folds over."""
    nx, ny, dx = 41, 21, 0.1
    surf = curved_surfaces(str(tmp), nx, ny, dx, amp=1.5 * (ny - 1) * dx)
    return write_case(str(tmp), "fold", nx=nx, ny=ny, dx=dx, nt=1, surf=surf)


def test_folded_grid_message(tmp_path):
    p = folded_case(tmp_path)
    assert _host_setup_error(p) == JAC
    assert _oracle_error(p) == JAC


# Malformed input files: one edit of a valid two-block problem each, and the
# message the set-up stops with.  Missing sections print the reference's
# messages (problem.cpp:33, domain.cpp:43, fields.cpp:32, block.cpp:61,
# interface.cpp:38, friction.cpp:93, slipweak.cpp:59, frontlist.cpp:36), and
# so do bad FD and RK orders (fd.cpp:150, rk.cpp:31,93).  Where the reference
# stops on an assert (boundary type boundary.cpp:22, interface type
# domain.cpp:364, block size block.cpp:118) the message is this library's.
INPUT_ERRORS = {
    "no_problem": (("[fdfault.problem]", "[fdfault.problemX]"), "Error reading problem from input file"),
    "no_domain": (("[fdfault.domain]", "[fdfault.domainX]"), "Error reading domain from input file"),
    "no_fields": (("[fdfault.fields]", "[fdfault.fieldsX]"), "Error reading fields from input file"),
    "no_block": (("[fdfault.block010]", "[fdfault.block010X]"), "Error reading block 010 from input file"),
    "no_interface": (("[fdfault.interface0]", "[fdfault.interface0X]"), "Error reading interface 0 from input file"),
    "no_friction": (("[fdfault.friction]", "[fdfault.frictionX]"), "Error reading interface 0 from input file"),
    "no_slipweak": (("[fdfault.slipweak]", "[fdfault.slipweakX]"), "Error reading slipweak from input file"),
    "no_frontlist": (("[fdfault.frontlist]", "[fdfault.frontlistX]"), "Error reading frontlist from input file"),
    "boundary": (("absorbing\nabsorbing\nabsorbing\nnone", "sticky\nabsorbing\nabsorbing\nnone"),
                 "unknown boundary type sticky"),
    "iface_type": (("slipweak\n4\nelastic", "glue\n4\nelastic"), "unknown interface type glue"),
    "sbporder": (("slipweak\n4\nelastic", "slipweak\n5\nelastic"), "Error in initializing FD method"),
    "rkorder": (("1000\n4\n", "1000\n7\n"), "Error in initializing RK method"),
    "block_size": (("41 42 1\n1 2 1\n41\n21 21", "41 8 1\n1 2 1\n41\n4 4"),
                   "block too small for the finite difference method"),
    "material": (("slipweak\n4\nelastic", "slipweak\n4\nrubber"), "material must be elastic or plastic"),
}


def edited_case(tmp, name, edit):
    p = write_case(str(tmp), name, nx=41, ny=21, dx=0.1, nt=1)
    text = open(p).read()
    assert edit[0] in text
    open(p, "w").write(text.replace(edit[0], edit[1], 1))
    return p


@pytest.mark.parametrize("case", sorted(INPUT_ERRORS))
def test_malformed_input_message(tmp_path, case):
    edit, msg = INPUT_ERRORS[case]
    p = edited_case(tmp_path, case, edit)
    assert _host_setup_error(p) == msg
    assert _oracle_error(p) == msg
