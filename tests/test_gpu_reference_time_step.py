"""The reference's time-stepping suite, tests/time_step_tests.c, mirrored through
the QuaC API with its systems, times, steps and tolerances, under the two option
sets the reference runs it with (Makefile:46-58: the default adaptive RK3BS, and
-ts_adapt_type none).  MCWF ensembles use a fixed seed here (the reference seeds
SPRNG), so their statistical tolerances are checked on one fixed sample.

Covered beside the cases tested elsewhere (T1 decay of one system,
time_step_tests.c:24-101, and the circuit decay of :337-396, in
test_gpu_stepping.py / test_gpu_mcwf.py):
* :145-191  MCWF T1 decay restarted ten times;
* :194-241  two vec-op subsystems, cascade 2 -> 1 -> 0 at rates 2 gamma, gamma;
* :243-285  a qubit and a 3-level mode, <n> = 1/e and 2/e;
* :287-335  the same under MCWF;
* :398-456  the circuit decay under MCWF;
* :459-564  circuits with gamma = 0: the ensemble equals the density matrix;
* :678-750  circuits applied to the states ten times between runs from t = 0:
            fidelity(dm, wf) against fidelity(wf, ensemble).
(The reference does not run :566-676, gamma = 1, whose two systems differ in
their dephasing terms.)  Also the one test tests/time_dep_tests.c runs (:142-213):
a time-dependent drive on a wavefunction checked at every monitored step."""
import ctypes
import math

import numpy as np
import pytest

import quac_amd as qa

pytestmark = pytest.mark.gpu
OPTS = {"adaptive": ("3bs", "basic"), "fixed": ("3bs", "none")}


@pytest.fixture(params=list(OPTS))
def opts(request):
    rk, ad = OPTS[request.param]
    qa.set_option("ts_rk_type", rk)
    qa.set_option("ts_adapt_type", ad)
    yield request.param
    qa.set_option("ts_rk_type", "3bs")
    qa.set_option("ts_adapt_type", "basic")


def _fidelity(a, b):
    f, v = ctypes.c_double(), ctypes.c_double()
    qa.lib().get_fidelity_qvec(a.h, b.h, ctypes.byref(f), ctypes.byref(v))
    qa._check()
    return f.value


def test_t1_decay_mcwf_restarted(opts):
    """time_step_tests.c:145-191"""
    n = 1000
    qs = qa.System()
    o = qs.create_op(2)
    qs.add_ham_term(1.0, o.n)
    qs.add_lin_term(1.0, o)
    qs.use_mcwf_solver(n, 11)
    x = qs.create_qvec()
    x.add(1.0, 1)
    qs.construct_matrix()
    for i in range(10):
        qs.time_step(x, i * 0.1, (i + 1) * 0.1, 0.01, 1000)
    assert abs(x.expectation(o.n).real - math.exp(-1)) < 2 / math.sqrt(n)
    qs.destroy()


def test_t1_decay_dm_vec_two_subsystems(opts):
    """time_step_tests.c:194-241"""
    qs = qa.System()
    v3 = qs.create_vec_op(3)
    v2 = qs.create_vec_op(2)
    qs.add_ham_term(1.0, v2[1], v2[1])
    qs.add_lin_term(1.0, v2[0], v2[1])
    qs.add_lin_term(1.0, v3[0], v3[1])
    qs.add_lin_term(2.0, v3[1], v3[2])
    x = qs.create_qvec()
    x.add(1.0, 5, 5)
    qs.construct_matrix()
    qs.time_step(x, 0.0, 1.0, 0.05, 200)
    assert abs(x.expectation(v2[1], v2[1]).real - math.exp(-1)) < 1e-4
    e = x.expectation(v3[1], v3[1]).real + 2 * x.expectation(v3[2], v3[2]).real
    assert abs(e - 2 * math.exp(-1)) < 2e-4
    qs.destroy()


def test_t1_decay_dm_two_subsystems(opts):
    """time_step_tests.c:243-285"""
    qs = qa.System()
    o3 = qs.create_op(3)
    o2 = qs.create_op(2)
    qs.add_ham_term(1.0, o2.n)
    qs.add_lin_term(1.0, o2)
    qs.add_lin_term(1.0, o3)
    x = qs.create_qvec()
    x.add(1.0, 5, 5)
    qs.construct_matrix()
    qs.time_step(x, 0.0, 1.0, 0.05, 200)
    assert abs(x.expectation(o2.n).real - math.exp(-1)) < 1e-4
    assert abs(x.expectation(o3.n).real - 2 * math.exp(-1)) < 2e-4
    qs.destroy()


def test_t1_decay_mcwf_two_subsystems(opts):
    """time_step_tests.c:287-335"""
    n = 1000
    qs = qa.System()
    o3 = qs.create_op(3)
    o2 = qs.create_op(2)
    qs.add_ham_term(1.0, o2.n)
    qs.add_lin_term(1.0, o2)
    qs.add_lin_term(1.0, o3)
    qs.use_mcwf_solver(n, 12)
    x = qs.create_qvec()
    x.add(1.0, 5)
    qs.construct_matrix()
    qs.time_step(x, 0.0, 1.0, 0.05, 2000)
    assert abs(x.expectation(o2.n).real - math.exp(-1)) < 2 / math.sqrt(n)
    assert abs(x.expectation(o3.n).real - 2 * math.exp(-1)) < 2 / math.sqrt(n)
    qs.destroy()


def test_circuit_decay_mcwf_two_subsystems(opts):
    """time_step_tests.c:398-456: SIGMAX on qubit 0 (the last-created subsystem)
    at t = 0.1 and on qubit 1 at t = 0.2, then decay to t = 1.1"""
    n = 1000
    qs = qa.System()
    o2 = qs.create_op(2)
    o3 = qs.create_op(2)
    qs.add_lin_term(1.0, o2)
    qs.add_lin_term(1.0, o3)
    qs.use_mcwf_solver(n, 13)
    x = qs.create_qvec()
    x.add(1.0, 0)
    qs.construct_matrix()
    c = qa.Circuit(5)
    c.add_gate(0.0, "SIGMAX", [0])
    c.add_gate(0.1, "SIGMAX", [1])
    qs.apply_circuit(c, 0.1)
    qs.time_step(x, 0.0, 1.1, 0.1, 250000)
    assert abs(x.expectation(o2.n).real - math.exp(-0.9)) < 2 / math.sqrt(n)
    assert abs(x.expectation(o3.n).real - math.exp(-1.0)) < 2 / math.sqrt(n)
    qs.destroy()


def _two_qubits(gamma, mcwf=None):
    qs = qa.System()
    a = qs.create_op(2)
    b = qs.create_op(2)
    qs.add_lin_term(gamma, a)
    qs.add_lin_term(gamma, b)
    if mcwf:
        qs.use_mcwf_solver(*mcwf)
    x = qs.create_qvec()
    if mcwf:
        x.add(1.0, 0)
    else:
        x.add(1.0, 0, 0)
    qs.construct_matrix()
    return qs, a, b, x


def test_circuit_ensemble_equals_density_matrix_without_decay(opts):
    """time_step_tests.c:459-564: gamma = 0, so no trajectory jumps and the
    ensemble average equals the density matrix (expectations and bitstring
    probabilities)"""
    qd, a0, a1, dm = _two_qubits(0.0)
    qw, b0, b1, wf = _two_qubits(0.0, (100, 14))
    c = qa.Circuit(10)
    c.add_gate(0.0, "SIGMAX", [0])
    c.add_gate(0.1, "SIGMAX", [1])
    c.add_gate(0.2, "CNOT", [0, 1])
    c.add_gate(0.3, "RX", [0], theta=0.1)
    c.add_gate(0.4, "RY", [1], theta=0.1)
    c.add_gate(0.5, "HADAMARD", [1])
    c.add_gate(1.0, "CNOT", [0, 1])
    qd.apply_circuit(c, 0.0001)
    qw.apply_circuit(c, 0.0001)
    qd.time_step(dm, 0.0, 1.0001, 0.05, 200)
    qw.time_step(wf, 0.0, 1.0001, 0.05, 200)
    assert abs(dm.expectation(a0.n).real - wf.expectation(b0.n).real) < 1e-6
    assert abs(dm.expectation(a1.n).real - wf.expectation(b1.n).real) < 1e-6
    assert np.abs(dm.bitstring_probs() - wf.bitstring_probs()).max() < 1e-6
    qd.destroy()
    qw.destroy()


def test_circuit_restarts_ensemble_against_density_matrix(opts):
    """time_step_tests.c:678-750: the circuit applied to each state at once, then a
    decay run from t = 0 to 1, ten times; the fidelity of the pure circuit state
    with the ensemble tracks its fidelity with the density matrix"""
    n = 1000
    qd, _, _, dm = _two_qubits(0.5)
    qw, _, _, wf_ens = _two_qubits(0.5, (n, 15))
    q3 = qa.System()
    c4 = q3.create_op(2)
    q3.create_op(2)
    q3.add_ham_term(0.0, c4)
    wf = q3.create_qvec()
    wf.add(1.0, 0)
    q3.construct_matrix()
    c = qa.Circuit(10)
    c.add_gate(0.5, "HADAMARD", [0])
    c.add_gate(0.5, "HADAMARD", [1])
    c.add_gate(0.3, "RX", [0], theta=0.1)
    c.add_gate(0.4, "RY", [1], theta=0.3)
    c.add_gate(1.0, "CNOT", [0, 1])
    c.add_gate(0.4, "RZ", [1], theta=0.9)
    c.add_gate(0.4, "RZ", [1], theta=0.4)
    for _ in range(10):
        qd.apply_circuit_to_qvec(c, dm)
        qw.apply_circuit_to_qvec(c, wf_ens)
        q3.apply_circuit_to_qvec(c, wf)
        qd.time_step(dm, 0.0, 1.0, 0.01, 200)
        qw.time_step(wf_ens, 0.0, 1.0, 0.01, 200)
        assert abs(_fidelity(dm, wf) - _fidelity(wf, wf_ens)) < 1 / math.sqrt(n)
    for s in (qd, qw, q3):
        s.destroy()


def test_time_dependent_wavefunction_every_step(opts):
    """time_dep_tests.c:142-213 (the test that file runs): H(t) = 0.5 omega
    e^{-alpha t} (a + a^dag) on a wavefunction from |0>, omega = 2 pi,
    alpha = 0.1, dt = 0.01 to t = 20; at every monitored step <sx> = 0,
    <sy> = -sin(theta), <sz> = cos(theta) with theta = omega (1 - e^{-alpha t}) /
    alpha, within the reference's 1e-4"""
    om, al = 2 * math.pi, 0.1
    qs = qa.System()
    o = qs.create_op(2)
    env = lambda t: math.exp(-al * t)
    qs.add_ham_term_time_dep(0.5 * om, env, o)
    qs.add_ham_term_time_dep(0.5 * om, env, o.dag)
    qs.construct_matrix()
    x = qs.create_wf()
    x.add_loc(1.0, 0)
    worst = [0.0, 0.0, 0.0]
    count = [0]

    def mon(step, t, v):
        p = v.get()
        c = p[0].conjugate() * p[1]
        th = om * (1 - math.exp(-al * t)) / al
        got = (2 * c.real, 2 * c.imag, abs(p[0]) ** 2 - abs(p[1]) ** 2)
        want = (0.0, -math.sin(th), math.cos(th))
        for k in range(3):
            worst[k] = max(worst[k], abs(got[k] - want[k]))
        count[0] += 1

    qs.set_monitor(mon)
    qs.time_step(x, 0.0, 20.0, 0.01, 10 ** 6)
    qs.set_monitor(None)
    assert count[0] > 20
    assert max(worst) < 1e-4, worst
    qs.destroy()
