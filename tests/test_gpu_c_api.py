"""The C drop-in path: a reference-style C program (variadic API, monitor with
the Vec-aliasing idiom) compiled with gcc against include/quac.h and linked
with libquac.so; its populations must match the oracle's FP64 RK4 to 1e-10."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import recipes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_jc_c_program_matches_oracle(tmp_path):
    exe = tmp_path / "jc_drop_in"
    src = os.path.join(ROOT, "tests", "c", "jc_drop_in.c")
    libdir = os.path.join(ROOT, "quac_amd")
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", str(exe),
                           "-L", libdir, "-lquac", "-Wl,-rpath," + libdir, "-lm"])
    out = subprocess.run([str(exe), "-ts_rk_type", "4", "-ts_adapt_type", "none"], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    rows = [l.split() for l in out.stdout.splitlines() if l and l[0].isdigit()]
    assert [int(r[0]) for r in rows] == list(range(0, 1001, 100))
    r = recipes.jc_recipe()
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    N, loc = recipes.jc_init()
    x = np.zeros(N * N, np.complex128)
    x[loc] = 1
    for k, row in enumerate(rows):
        if k:
            x = m.rk4(x, 0.01, 100)
        pe = oracle.expect_dm(tot, x, [(0, 2, 1, -1)]).real
        pc = oracle.expect_dm(tot, x, [(0, 10, 2, -1)]).real
        assert abs(float(row[2]) - pe) < 1e-10 and abs(float(row[3]) - pc) < 1e-10, (row, pe, pc)
    tr = [l for l in out.stdout.splitlines() if l.startswith("trace")][0].split()[1]
    assert abs(float(tr) - 1.0) < 1e-12
