"""Mutation check of the oracle's pins (test infrastructure; run by hand:
``python tests/oracle_mutants.py``).

Each mutant plants one plausible mistake in a copy of oracle/oracle.cpp --
a wrong sign, a swapped operand, a dropped term, a misread file layout --
builds it into oracle/_build/mutants/<name>/liboracle.so, and runs the pins
that should catch it with FDF_ORACLE_LIB pointing at the mutant.  A mutant
is killed when those pins fail.  The script prints one line per mutant and
exits non-zero if any survives, so every pin is shown to bite.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")
OUT = os.path.join(ROOT, "oracle", "_build", "mutants")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-w", "-shared"]

LAW = "tests/test_oracle_law_pins.py"
PINS = "tests/test_oracle_pins.py"

# name, original text, replacement, pins expected to fail
MUTANTS = [
    ("stz_f0_sign", "return v0 * exp(-f0 + mu / a - 1. / chi) * (1. - muy / mu);",
     "return v0 * exp(f0 + mu / a - 1. / chi) * (1. - muy / mu);", [LAW + "::test_stz_flow_law_known_answers"]),
    ("stz_yield_factor", "return v0 * exp(-f0 + mu / a - 1. / chi) * (1. - muy / mu);",
     "return v0 * exp(-f0 + mu / a - 1. / chi) * (1. + muy / mu);", [LAW + "::test_stz_flow_law_known_answers"]),
    ("stz_chi_sign", "return v0 * exp(-f0 + mu / a - 1. / chi) * (1. - muy / mu);",
     "return v0 * exp(-f0 + mu / a + 1. / chi) * (1. - muy / mu);", [LAW + "::test_stz_flow_law_known_answers"]),
    ("stz_relax_sign", "- Rt * exp(-betat / chi);", "- Rt * exp(betat / chi);",
     [LAW + "::test_stz_state_rate_known_answers"]),
    ("stz_chihat_form", "chiwt / log(v1t / V);", "chiwt * log(v1t / V);",
     [LAW + "::test_stz_state_rate_known_answers"]),
    ("sw_slip_ramp", "else f1 = u / dct;", "else f1 = 1. - u / dct;", [LAW + "::test_slipweak_mu_known_answers"]),
    ("sw_min_not_max", "return (f1 >= f2) ?", "return (f1 <= f2) ?", [LAW + "::test_slipweak_mu_known_answers"]),
    ("sw_time_ramp_origin", "f2 = (tct == 0.) ? 1. : (t - trupt) / tct;", "f2 = (tct == 0.) ? 1. : t / tct;",
     [LAW + "::test_slipweak_mu_known_answers"]),
    ("sw_no_cohesion", "else { S = c0t + mu * fabs(snc); V = (phi - S) / eta; }",
     "else { S = mu * fabs(snc); V = (phi - S) / eta; }", [LAW + "::test_slipweak_strength_branches"]),
    ("sw_tensile_branch", "if (phi > c0t) { S = c0t; V = (phi - S) / eta; }",
     "if (phi > 0.) { S = 0.; V = (phi - S) / eta; }", [LAW + "::test_slipweak_strength_branches"]),
    ("mode3_modulus", "double t1 = K.g * n0 * c * vz;", "double t1 = K.lam * n0 * c * vz;",
     [LAW + "::test_mode3_linear_fields_exact_rates"]),
    ("szz_ratio", "double nu = 0.5 * B.lam / (B.lam + B.G);", "double nu = B.lam / (B.lam + 2. * B.G);",
     [LAW + "::test_plastic_szz_rate_plane_strain"]),
    ("locked_impedance_swap", "o.s1 = (z2 * (c.s1 - z1 * c.v1) + z1 * (c.s2 + z2 * c.v2)) / (z1 + z2);",
     "o.s1 = (z1 * (c.s1 - z1 * c.v1) + z2 * (c.s2 + z2 * c.v2)) / (z1 + z2);",
     [LAW + "::test_interface_reflection_transmission"]),
    ("iface_normal_penalty_cs", "df[N + id1] -= I.cp1 * h1 * o1.v2;", "df[N + id1] -= I.cs1 * h1 * o1.v2;",
     [LAW + "::test_interface_penalty_one_stage"]),
    ("iface_shear_penalty_cp", "df[3 * N + id1] -= I.cs1 * h1 * o1.s12;", "df[3 * N + id1] -= I.cp1 * h1 * o1.s12;",
     [LAW + "::test_interface_penalty_one_stage"]),
    ("load_ramp_step", "double tfunc(double t) const { return (t >= t0) ? 1. : t / t0; }",
     "double tfunc(double t) const { return (t >= t0) ? 1. : 0.; }", [LAW + "::test_load_time_ramp_and_stage_time"]),
    ("stage_time_no_C", "const double ts = t + P.rk.C[stage] * dt;", "const double ts = t;",
     [LAW + "::test_load_time_ramp_and_stage_time"]),
    ("file_stack_interleaved",
     "for (int k = 0; k < nstack; k++) dst[k].assign(all.begin() + k * n, all.begin() + (k + 1) * n);",
     "for (int k = 0; k < nstack; k++) { dst[k].resize(n); for (size_t q = 0; q < n; q++) dst[k][q] = all[q * nstack + k]; }",
     [LAW + "::test_fault_input_file_equals_perturbation"]),
    ("stz_file_order", "prm[4] += I.prm[0][index];\n        prm[5] += I.prm[1][index];",
     "prm[4] += I.prm[1][index];\n        prm[5] += I.prm[0][index];", [LAW + "::test_fault_input_file_equals_perturbation"]),
    ("state_file_ignored",
     "if (sfile != \"none\") I.state = read_binary(sfile, n, \"Error opening state file in friction.cpp\", \"state\");", "",
     [LAW + "::test_stz_state_file_layout"]),
    ("sbp_interior_order", "row(fd.c[0], {-0.5, 0., 0.5});", "row(fd.c[0], {-0.5, 0.01, 0.49});",
     [PINS + "::test_sbp_rows_exact_on_polynomials"]),
]


def build(name, old, new):
    src = open(SRC).read()
    if src.count(old) != 1:
        return None, "pattern found %d times" % src.count(old)
    d = os.path.join(OUT, name)
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "oracle.cpp")
    with open(path, "w") as fh:
        fh.write(src.replace(old, new))
    # the mutant includes oracle.h from the original directory
    lib = os.path.join(d, "liboracle.so")
    r = subprocess.run(["g++"] + CXXFLAGS + ["-I", os.path.dirname(SRC), "-o", lib, path], capture_output=True,
                       text=True)
    if r.returncode:
        return None, r.stderr[-400:]
    return lib, ""


def check(m):
    name, old, new, tests = m
    lib, err = build(name, old, new)
    if lib is None:
        return name, "BUILD FAILED: " + err
    env = dict(os.environ, FDF_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + tests, cwd=ROOT,
                       env=env, capture_output=True, text=True)
    return name, ("killed" if r.returncode else "SURVIVED")


def main():
    bad = 0
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        for name, verdict in ex.map(check, MUTANTS):
            print("%-26s %s" % (name, verdict), flush=True)
            bad += verdict != "killed"
    print("%d mutants, %d not killed" % (len(MUTANTS), bad))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
