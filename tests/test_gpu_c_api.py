"""The C drop-in path: a reference-style C program (variadic API, monitor with
the Vec-aliasing idiom) compiled with gcc against include/quac.h and linked
with libquac.so; its populations must match the oracle's FP64 RK4 to 1e-10."""
import math
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


def _build(tmp_path, prog):
    exe = tmp_path / prog
    libdir = os.path.join(ROOT, "quac_amd")
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "c", prog + ".c"), "-o", str(exe),
                           "-L", libdir, "-lquac", "-Wl,-rpath," + libdir, "-lm"])
    return exe


def _run(exe, *opts, cwd=None):
    out = subprocess.run([str(exe), *opts], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120,
                         cwd=cwd)
    assert out.returncode == 0, out.stdout
    return out.stdout


def test_legacy_api_jc_program_matches_oracle(tmp_path):
    """the legacy global-operator API (create_op / add_to_ham / add_to_ham_mult2 /
    add_lin / set_initial_pop / create_full_dm / set_dm_from_initial_pop /
    set_ts_monitor / time_step / get_populations) on the JC config: populations
    every 100 RK4 steps within 1e-10 of the oracle's FP64 RK4"""
    text = _run(_build(tmp_path, "legacy_jc"), "-ts_rk_type", "4", "-ts_adapt_type", "none")
    rows = [l.split() for l in text.splitlines() if l and l[0].isdigit()]
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
    ex = float([l for l in text.splitlines() if l.startswith("expect")][0].split()[1])
    assert abs(ex - float(rows[-1][3])) < 1e-12
    tr = [l for l in text.splitlines() if l.startswith("trace")][0].split()
    assert abs(float(tr[1]) - 1.0) < 1e-12 and tr[2:] == ["2", "20"]


@pytest.mark.parametrize("opts,tol", [(("-ts_rk_type", "4", "-ts_adapt_type", "none"), 1e-9), ((), 2e-3)])
def test_legacy_api_vec_decay_closed_form(tmp_path, opts, tol):
    """a VEC three-level atom: |1> decays to |0> at rate 0.5 through the VEC-pair
    Lindblad operator |0><1| (add_lin_mult2), initial populations 3/4 in |1>, 1/4
    in |2> from set_initial_pop; at t = 2: p1 = 3/4 e^-1, p0 = 3/4 (1 - e^-1),
    p2 = 1/4, cavity <n> = 0.  Fixed-step RK4 to 1e-9; the legacy default
    (adaptive RK3BS at PETSc's default tolerances 1e-4) to 2e-3"""
    text = _run(_build(tmp_path, "legacy_vec"), *opts)
    f = [l for l in text.splitlines() if l.startswith("pops")][0].split()
    assert f[1] == "4"
    p0, p1, p2, nc = map(float, f[2:])
    e = math.exp(-1.0)
    assert abs(p1 - 0.75 * e) < tol and abs(p0 - 0.75 * (1 - e)) < tol and abs(p2 - 0.25) < tol and abs(nc) < tol, f
    # combine_ops_to_mat: |2><2| (x) I over 9 levels has trace 3; no operators: the identity
    c0 = [l.split() for l in text.splitlines() if l.startswith("comb0")][0]
    c1 = [l.split() for l in text.splitlines() if l.startswith("comb1")][0]
    assert c0[1] == "9" and abs(float(c0[2]) - 3.0) < 1e-15 and c1[1] == "9" and abs(float(c1[2]) - 9.0) < 1e-15


def test_legacy_api_steady_state_closed_form(tmp_path):
    """legacy steady_state(Vec) of a driven, detuned, damped qubit: rho_ee within
    1e-6 of (omega^2/4) / (delta^2 + gamma^2/4 + omega^2/2), unit trace"""
    text = _run(_build(tmp_path, "legacy_steady"))
    f = [l for l in text.splitlines() if l.startswith("steady")][0].split()
    omega, delta, gamma = 0.4, 0.3, 1.0
    ree = (omega ** 2 / 4) / (delta ** 2 + gamma ** 2 / 4 + omega ** 2 / 2)
    assert abs(float(f[1]) - ree) < 1e-6 and abs(float(f[2]) - 1.0) < 1e-9, (f, ree)


def test_legacy_api_circuit_flip_then_decay(tmp_path):
    """legacy circuits: sigma_x at t = 1 (add_gate_to_circuit, copied by
    add_circuit_to_circuit, started by start_circuit_at_time) on a qubit with
    decay 0.7, fixed-step RK4: <n>(2) = exp(-0.7) within 1e-8"""
    text = _run(_build(tmp_path, "legacy_circuit"), "-ts_rk_type", "4", "-ts_adapt_type", "none")
    n = float([l for l in text.splitlines() if l.startswith("n ")][0].split()[1])
    g = int([l for l in text.splitlines() if l.startswith("gates")][0].split()[1])
    assert g == 1 and abs(n - math.exp(-0.7)) < 1e-8, (n, math.exp(-0.7))


def test_legacy_dm_utilities_closed_forms(tmp_path):
    """legacy partial traces, concurrence and fidelity on two qubits:
    - rho_a (x) rho_b: partial_trace_over(b) = rho_a and partial_trace_keep(b) =
      rho_b, to 1e-14 (a swapped keep/over or subsystem order fails this, the
      factors differ); partial_trace_over_one over the middle factor of 1x2x2
      gives rho_b, of 2x2x1 rho_a;
    - Werner p |Phi+><Phi+| + (1 - p) I / 4: concurrence (3p - 1) / 2, left
      unclipped as in dm_utilities.c:1568-1582 (so -0.2 at p = 0.2), and
      fidelity to |Phi+> sqrt(p + (1 - p) / 4) (Tr sqrt(sqrt(rho) sigma
      sqrt(rho)), not squared: dm_utilities.c:1690-1699);
    - cos t |01> + e^{i f} sin t |10>: concurrence sin 2t;
    - measure_dm with a projector (dm_utilities.c:270-308);
    - the sparse printers' entries, on screen and in their files."""
    text = _run(_build(tmp_path, "legacy_dm"), cwd=str(tmp_path))
    body = text.split("sparse_begin\n")[1].split("sparse_end")[0]
    text = text.replace(body, "")
    lines = {l.split()[0] + (l.split()[1] if l.startswith("werner") else ""): l.split()
             for l in text.splitlines() if l.strip()}
    # sparse printers (dm_utilities.c:32-113): |Phi+><Phi+| has 1/2 at (0,0), (0,3),
    # (3,0), (3,3); a a^dag (x) I = |0><0| (x) I has 1 at (0,0), (1,1)
    want = [(i, j) for i in (0, 3) for j in (0, 3)]
    for src in (body, (tmp_path / "dm_sparse.txt").read_text()):
        rows = [l.split() for l in src.splitlines() if l.strip()]
        assert [(int(r[0]), int(r[1])) for r in rows] == want, src
        assert all(abs(float(r[2]) - 0.5) < 1e-6 and float(r[3]) == 0.0 for r in rows), src
    rows = [l.split() for l in (tmp_path / "mat_sparse.txt").read_text().splitlines() if l.strip()]
    assert [(int(r[0]), int(r[1]), float(r[2])) for r in rows] == [(0, 0, 1.0), (1, 1, 1.0)], rows
    ra = np.array([[0.7, 0.2 + 0.1j], [0.2 - 0.1j, 0.3]])
    rb = np.array([[0.1, -0.05j], [0.05j, 0.9]])
    for tag, want in (("over_b", ra), ("keep_b", rb), ("one_a", rb), ("one_b", ra)):
        v = [float(x) for x in lines[tag][1:]]
        got = np.array([complex(v[0], v[1]), complex(v[2], v[3]), complex(v[4], v[5]), complex(v[6], v[7])])
        assert np.abs(got - want.flatten(order="F")).max() < 1e-14, (tag, got, want)
    for p in (0.8, 0.2):
        f = lines["werner%.2f" % p]
        assert abs(float(f[2]) - (3 * p - 1) / 2) < 1e-12, (p, f)
        assert abs(float(f[3]) - math.sqrt(p + (1 - p) / 4)) < 1e-12, (p, f)
    assert abs(float(lines["pure"][1]) - math.sin(0.6)) < 1e-7, lines["pure"]
    # measure_dm with n of a (subsystem 0, the most significant digit): only
    # |10>, |11> survive, 2 ((1 - p)/4, (1 - p)/4 + p/2) = (0.1, 0.9) at p = 0.8,
    # the coherence to |00> removed
    v = [float(x) for x in lines["measure"][1:]]
    got = np.array([complex(v[2 * k], v[2 * k + 1]) for k in range(16)]).reshape(4, 4, order="F")
    want = np.zeros((4, 4))
    want[2, 2], want[3, 3] = 0.1, 0.9
    assert np.abs(got - want).max() < 1e-14, got


def test_c_sweep_through_time_step_sys_batch(tmp_path):
    """tests/c/jc_sweep_batch.c: eight Jaynes-Cummings systems of 4 .. 11 cavity
    levels advanced together by time_step_sys_batch (FP32 batch, unit trace every
    100 steps) against time_step_sys on each alone (FP64 RK4), and the first
    system's single run against the oracle"""
    exe = _build(tmp_path, "jc_sweep_batch")
    out = _run(exe, "-ts_rk_type", "4", "-ts_adapt_type", "none", "-quac_renorm", "100")
    rows = {}
    for l in out.splitlines():
        p = l.split()
        if len(p) == 5 and p[0] in ("batch", "single"):
            rows[(p[0], int(p[1]))] = [float(v) for v in p[2:]]
    assert len(rows) == 16
    for i in range(8):
        b, s = rows[("batch", i)], rows[("single", i)]
        assert abs(b[0] - s[0]) < 1e-6 and abs(b[1] - s[1]) < 1e-6, (i, b, s)
        assert abs(b[2] - 1.0) < 1e-6 and abs(s[2] - 1.0) < 1e-12
    r = recipes.jc_recipe(levels=4, wc=1.0 + 0.02 * (0 - 3.5), kappa=2e-2)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    x = np.zeros(tot * tot, np.complex128)
    x[4 * tot + 4] = 1.0
    x = m.rk4(x, 0.01, 1000)
    pe = oracle.expect_dm(tot, x, [(0, 2, 1, -1)]).real
    pc = oracle.expect_dm(tot, x, [(0, 4, 2, -1)]).real
    s = rows[("single", 0)]
    assert abs(s[0] - pe) < 1e-10 and abs(s[1] - pc) < 1e-10


def test_c_qvec_suite_mirrors_the_reference(tmp_path):
    """tests/c/qvec_suite.c: the state-container and observable cases of the
    reference's qvec_utilities_tests.c (creation and dimensions, expectation
    values, the random wavefunction fixture, bitstring probabilities of a
    wavefunction, a density matrix and an MCWF ensemble) through the C API"""
    exe = _build(tmp_path, "qvec_suite")
    out = _run(exe, os.path.join(ROOT, "tests", "golden", "pristine_matrices", "rand_wf.dat"))
    assert "FAIL" not in out, out
    last = [l for l in out.splitlines() if l.startswith("checks")][-1].split()
    assert int(last[1]) >= 90 and int(last[3]) == 0, out
