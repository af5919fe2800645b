"""Source compatibility: programs written for the reference build unchanged
against include/compat (its header names) + include/ and link with libquac.so.
The reference's own sources are compiled where they lie when the reference tree
is present (this check reads them, nothing is copied); skipped otherwise.

* examples/simple_jc_test.c and examples/simple_circuit.c compile and link;
* tests/time_step_tests.c, tests/time_dep_tests.c and tests/qvec_utilities_tests.c
  compile (their Unity runner, tests/unity.c, includes the reference's own
  PETSc-based operators.h and is not built)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
INC = ["-I", os.path.join(ROOT, "include", "compat"), "-I", os.path.join(ROOT, "include")]
pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")), reason="reference tree absent")


@pytest.mark.parametrize("name", ["simple_jc_test", "simple_circuit"])
def test_reference_example_builds_against_the_library(name, tmp_path):
    libdir = os.path.join(ROOT, "quac_amd")
    r = subprocess.run(["gcc", "-std=gnu11", "-w", *INC, os.path.join(REF, "examples", name + ".c"),
                        "-o", str(tmp_path / name), "-L", libdir, "-lquac", "-Wl,-rpath," + libdir, "-lm"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("name", ["time_step_tests", "time_dep_tests", "qvec_utilities_tests"])
def test_reference_test_sources_compile_against_the_headers(name):
    r = subprocess.run(["gcc", "-std=gnu11", "-fsyntax-only", "-Werror=implicit-function-declaration", *INC,
                        "-I", os.path.join(REF, "tests"), os.path.join(REF, "tests", name + ".c")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
