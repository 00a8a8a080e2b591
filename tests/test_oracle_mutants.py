"""Planted-mistake check of the oracle's LES pins (tests/test_oracle_les.py).

Each mutant is one plausible transcription slip in the oracle's Smagorinsky
branch or in the normal-stress moment rows it reads.  The oracle source is
copied with that one slip into a temporary directory, built there, and
tests/test_oracle_les.py is run against the build (TDLB_ORACLE_LIB): it must
fail on every mutant and pass on an unmutated build made the same way.  The
committed oracle is never modified.  CPU only.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "tdlb_oracle.cpp")
LES_TESTS = os.path.join(ROOT, "tests", "test_oracle_les.py")
FLAGS = ["-O1", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp", "-shared"]

# (name, text in the oracle, replacement); the text must occur exactly once
MUTANTS = [
    ("dudx reads M07", "0.5 * f[0] * ux - m[M05]", "0.5 * f[0] * ux - m[M07]"),
    ("dvdy reads M05", "0.5 * f[1] * uy - m[M07]", "0.5 * f[1] * uy - m[M05]"),
    ("dwdz reads M07", "0.5 * f[2] * uz - m[M10]", "0.5 * f[2] * uz - m[M07]"),
    ("dudypdvdx reads M08", "0.5 * f[1] * ux - m[M06]", "0.5 * f[1] * ux - m[M08]"),
    ("dudzpdwdx reads M09", "0.5 * f[2] * ux - m[M08]", "0.5 * f[2] * ux - m[M09]"),
    ("dvdzpdwdy reads M06", "0.5 * f[2] * uy - m[M09]", "0.5 * f[2] * uy - m[M06]"),
    ("dudypdvdx momentum M02", "((m[M03]) + 0.5 * f[1] * ux", "((m[M02]) + 0.5 * f[1] * ux"),
    ("dudzpdwdx momentum M03", "((m[M04]) + 0.5 * f[2] * ux", "((m[M03]) + 0.5 * f[2] * ux"),
    ("dvdzpdwdy momentum M03", "((m[M04]) + 0.5 * f[2] * uy", "((m[M03]) + 0.5 * f[2] * uy"),
    ("dudx drops momentum", "((m[M02] + 0.5 * f[0] * ux", "((0.5 * f[0] * ux"),
    ("dwdz drops force term", "m[M04] + 0.5 * f[2] * uz - m[M10]", "m[M04] - m[M10]"),
    ("dudypdvdx force term F_x", "0.5 * f[1] * ux - m[M06]", "0.5 * f[0] * ux - m[M06]"),
    ("dvdzpdwdy force term u_x", "0.5 * f[2] * uy - m[M09]", "0.5 * f[2] * ux - m[M09]"),
    ("fct without old Nu", "(1.0 + 6.0 * (nuCell + flow.nu))", "(1.0 + 6.0 * (flow.nu))"),
    ("divergence term dropped", " - (2.0 / 3.0) * divv * divv;", ";"),
    ("cross term not doubled", "T dudzpdwdx = 2 * fct", "T dudzpdwdx = fct"),
    ("relaxation with old Nu",
     "nuCell = flow.cs0 * flow.cs0 * std::sqrt(std::fabs(sh));\n            T nut = nuCell + flow.nu;",
     "T nut = nuCell + flow.nu;\n            nuCell = flow.cs0 * flow.cs0 * std::sqrt(std::fabs(sh));"),
    ("M05 row, z faces added", "m[M05] = -q[Q03] - q[Q04] - q[Q05] - q[Q06]",
     "m[M05] = -q[Q03] - q[Q04] + q[Q05] + q[Q06]"),
    ("M10 relaxes from M07", "* uz - m[M10] * c) * b;", "* uz - m[M07] * c) * b;"),
]


def build(tmp_path, name, text):
    d = tmp_path / name.replace(" ", "_")
    d.mkdir()
    src = d / "tdlb_oracle.cpp"
    src.write_text(text)
    lib = d / "liboracle.so"
    subprocess.run(["g++", *FLAGS, "-o", str(lib), str(src)], check=True)
    return str(lib)


def run_les_tests(lib):
    env = dict(os.environ, TDLB_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", LES_TESTS],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return r.returncode, r.stdout[-2000:]


def test_unmutated_build_passes(tmp_path):
    with open(SRC) as fh:
        rc, out = run_les_tests(build(tmp_path, "control", fh.read()))
    assert rc == 0, out


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_les_pins_catch_planted_mistake(tmp_path, name, old, new):
    with open(SRC) as fh:
        text = fh.read()
    assert text.count(old) == 1, f"mutant '{name}': text found {text.count(old)} times"
    rc, out = run_les_tests(build(tmp_path, name, text.replace(old, new)))
    assert rc != 0, f"mutant '{name}' passed tests/test_oracle_les.py:\n{out}"
