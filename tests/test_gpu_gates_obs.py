"""Gates (both qubit conventions), observables, PETSc-binary state input, and the
steady-state GMRES on the device, against the oracle and closed forms (1e-12
max-abs for gates/observables in FP64)."""
import ctypes
import math
import os

import numpy as np
import pytest

import dense
import oracle
import petsc_io
import quac_amd as qa
import recipes
from qbuild import build_system

pytestmark = pytest.mark.gpu
FX = os.path.join(os.path.dirname(__file__), "golden", "pristine_matrices")
RNG = np.random.default_rng(7)


def _rand_rho(N):
    a = RNG.normal(size=(N, N)) + 1j * RNG.normal(size=(N, N))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


GATECASES = [("H", 1, (0,), 0.0), ("H", 1, (3,), 0.0), ("RZ", 8, (1,), 0.7), ("RX", 6, (2,), 1.3),
             ("SIGMAX", 2, (3,), 0.0), ("SIGMAY", 3, (0,), 0.0), ("CNOT", -1, (0, 1), 0.0),
             ("CNOT", -1, (3, 1), 0.0), ("CZ", -3, (2, 0), 0.0), ("CXZ", -2, (1, 3), 0.0),
             ("SIGMAZ", 4, (2,), 0.0), ("EYE", 5, (1,), 0.0), ("RY", 7, (3,), 2.1), ("CXZ", -2, (3, 0), 0.0)]


@pytest.mark.parametrize("name,code,qubits,theta", GATECASES)
def test_gate_superop_convention_vs_oracle(name, code, qubits, theta):
    nq = 4
    r = {"subs": [("op", 2)] * nq, "terms": [("lin", 0.0, [(0, "n")])], "mcwf": False}
    qs, _, _ = build_system(r, construct=False)
    x = qs.create_qvec()
    rho = _rand_rho(2 ** nq)
    x.set(dense.vec(rho))
    qs.apply_gate(x, name, list(qubits), theta=theta)
    ref = oracle.gate_superop([2] * nq, [1, 2, 4, 8], True, code, qubits, dense.vec(rho), theta=theta)
    assert np.abs(x.get() - ref).max() < 1e-12
    qs.destroy()


@pytest.mark.parametrize("gate", ["U1", "U2", "U3"])
def test_u_gates_vs_oracle(gate):
    """U1/U2/U3 take (theta, phi, lambda) into the OpenQASM U3 matrix
    (U3_gate_func_wf, quantum_circuits.c:2055-2110) on both paths"""
    nq, th, ph, la = 3, 1.1, -0.6, 2.3
    code = {"U1": 9, "U2": 10, "U3": 11}[gate]
    r = {"subs": [("op", 2)] * nq, "terms": [("lin", 0.0, [(0, "n")])], "mcwf": False}
    qs, _, _ = build_system(r, construct=False)
    rho = _rand_rho(2 ** nq)
    x = qs.create_qvec()
    x.set(dense.vec(rho))
    qs.apply_gate(x, gate, [1], theta=th, phi=ph, lam=la)
    ref = oracle.gate_superop([2] * nq, [1, 2, 4], True, code, (1,), dense.vec(rho), theta=th, phi=ph, lam=la)
    assert np.abs(x.get() - ref).max() < 1e-12
    x.set(dense.vec(rho))
    x.apply_gate2(gate, [2], theta=th, phi=ph, lam=la)
    ref = oracle.gate_inplace([2] * nq, True, code, (2,), dense.vec(rho), theta=th, phi=ph, lam=la)
    assert np.abs(x.get() - ref).max() < 1e-12
    qs.destroy()


@pytest.mark.parametrize("name,code,qubits,theta", GATECASES)
def test_gate_inplace_convention_vs_oracle(name, code, qubits, theta):
    nq = 4
    r = {"subs": [("op", 2)] * nq, "terms": [("lin", 0.0, [(0, "n")])], "mcwf": False}
    qs, _, _ = build_system(r, construct=False)
    x = qs.create_qvec()
    rho = _rand_rho(2 ** nq)
    x.set(dense.vec(rho))
    x.apply_gate2(name, list(qubits), theta=theta)
    ref = oracle.gate_inplace([2] * nq, True, code, qubits, dense.vec(rho), theta=theta)
    assert np.abs(x.get() - ref).max() < 1e-12
    qs.destroy()


def test_gate_on_wavefunction_and_mixed_levels():
    # qubit embedded between a 3-level system and a 2-level system (strides 2 and 1)
    qs = qa.System()
    a = qs.create_op(3)
    b = qs.create_op(2)
    c = qs.create_op(2)
    qs.add_ham_term(1.0, a.n)
    x = qs.create_wf()
    psi = RNG.normal(size=12) + 1j * RNG.normal(size=12)
    x.set(psi)
    qs.apply_gate(x, "H", [1])
    ref = oracle.gate_superop([3, 2, 2], [1, 3, 6], False, 1, (1,), psi)
    assert np.abs(x.get() - ref).max() < 1e-12
    qs.destroy()


def test_apply_circuit_to_qvec_and_layers():
    nq = 3
    r = {"subs": [("op", 2)] * nq, "terms": [("lin", 0.0, [(0, "n")])], "mcwf": False}
    qs, _, _ = build_system(r, construct=False)
    x = qs.create_qvec()
    rho = _rand_rho(8)
    x.set(dense.vec(rho))
    circ = qa.Circuit(10)
    gates = [("H", 1, (0,), 0), ("CNOT", -1, (0, 1), 0), ("RZ", 8, (2,), 0.4), ("RX", 6, (1,), 2.0),
             ("CNOT", -1, (2, 0), 0)]
    for g in gates:
        circ.add_gate(1.0, g[0], list(g[2]), theta=g[3])
    qs.apply_circuit_to_qvec(circ, x)
    ref = dense.vec(rho)
    for g in gates:
        ref = oracle.gate_superop([2] * nq, [1, 2, 4], True, g[1], g[2], ref, theta=g[3])
    assert np.abs(x.get() - ref).max() < 1e-12
    # greedy layering (quantum_circuits.c:316-404): H(0) | CNOT(0,1), RZ(2) | RX(1), CNOT(2,0)
    qs.schedule(circ)
    assert circ.num_layers == 3
    assert circ.layer_times() == [1.0, 2.0, 3.0]
    circ.destroy()
    qs.destroy()


def test_expectation_trace_populations_vs_oracle():
    r = {"subs": [("op", 2), ("op", 3), ("vec", 3)], "terms": [("lin", 0.0, [(0, "n")])], "mcwf": False}
    qs, subs, op = build_system(r, construct=False)
    tot = 18
    rho = _rand_rho(tot)
    x = qs.create_qvec()
    x.set(dense.vec(rho))
    for refs in ([(0, "lower"), (1, "raise")], [(1, "n")], [(0, "sx"), (0, "sy")],
                 [(2, ("vec", 1)), (2, ("vec", 2)), (1, "lower")]):
        got = x.expectation(*[op(q) for q in refs])
        _, terms = recipes.oracle_terms({"subs": r["subs"], "terms": [("ham", 1.0, refs)]})
        assert abs(got - oracle.expect_dm(tot, dense.vec(rho), terms[0][2])) < 1e-12
    assert abs(x.trace() - np.trace(rho)) < 1e-12
    assert abs(x.element(3, 5) - rho[3, 5]) < 1e-15
    assert np.abs(x.bitstring_probs() - np.real(np.diag(rho))).max() < 1e-15
    qs.destroy()


def test_wf_expectation_vs_oracle_and_known_answers():
    """<psi|O|psi> on the device (qvec_utilities.c:1353-1420 for a WF) against the
    oracle's qo_expect_wf (pinned by dense brute force) and the |1,1> known
    answers of qvec_utilities_tests.c:232-278"""
    r = {"subs": [("op", 2), ("op", 3), ("vec", 3)], "terms": [("ham", 0.0, [(0, "n")])], "mcwf": False}
    qs, subs, op = build_system(r, construct=False)
    tot = 18
    psi = RNG.normal(size=tot) + 1j * RNG.normal(size=tot)
    x = qs.create_wf()
    x.set(psi)
    for refs in ([(0, "lower"), (1, "raise")], [(1, "n")], [(0, "sy")], [(0, "sx"), (0, "sz")],
                 [(2, ("vec", 1)), (2, ("vec", 2)), (1, "lower")]):
        got = x.expectation(*[op(q) for q in refs])
        _, terms = recipes.oracle_terms({"subs": r["subs"], "terms": [("ham", 1.0, refs)]})
        want = oracle.expect_wf(tot, psi, terms[0][2])
        assert abs(got - want) < 1e-12 * max(1.0, abs(want)), refs
    qs.destroy()
    r2 = {"subs": [("op", 2), ("op", 2)], "terms": [("ham", 0.0, [(0, "n")])], "mcwf": False}
    qs, subs, op = build_system(r2, construct=False)
    x = qs.create_wf()
    x.add_loc(1.0, 3)  # |1,1>
    assert x.expectation(op((0, "n"))) == 1.0
    assert x.expectation(op((1, "n"))) == 1.0
    assert x.expectation(op((0, "raise")), op((1, "lower"))) == 0.0
    assert x.expectation(op((1, "raise")), op((0, "lower"))) == 0.0
    qs.destroy()


def test_read_dm_binary_fixture():
    """qvec_utilities_tests.c:788-870 reads rand_dm.dat; device copy must equal the file."""
    v = petsc_io.read_vec(os.path.join(FX, "rand_dm.dat"))
    x = qa.read_qvec_dm_binary(os.path.join(FX, "rand_dm.dat"))
    assert x.total_levels == 4 and x.is_dm
    assert np.array_equal(x.get(), v)
    assert abs(x.trace() - 1.0) < 1e-12
    x.destroy()


@pytest.mark.parametrize("omega,delta,gamma", [(1.0, 0.0, 0.5), (0.4, 0.3, 1.0)])
def test_steady_state_two_level_analytic(omega, delta, gamma):
    qs = qa.System()
    o = qs.create_op(2)
    qs.add_ham_term(omega / 2, o.sig_x)
    qs.add_ham_term(delta, o.n)
    qs.add_lin_term(gamma, o)
    qs.construct_matrix()
    x = qs.create_qvec()
    for jac in (False, True):
        x.set(np.zeros(4))
        it, res = qs.steady_state(x, tol=1e-12, jacobi=jac)
        rho = dense.unvec(x.get(), 2)
        ree = (omega ** 2 / 4) / (delta ** 2 + gamma ** 2 / 4 + omega ** 2 / 2)
        assert res < 1e-11 and abs(rho[1, 1].real - ree) < 1e-10 and abs(np.trace(rho) - 1) < 1e-10
    qs.destroy()


def test_steady_state_kerr_small_vs_oracle():
    p = recipes.kerr_params(4, 1)[0]
    r = recipes.kerr_recipe(p, levels=4)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    xo, ito, reso = m.gmres(tot, tol=1e-10, jacobi=True)
    qs, subs, op = build_system(r)
    x = qs.create_qvec()
    it, res = qs.steady_state(x, tol=1e-10, jacobi=True)
    assert res <= 1e-10
    got = x.get()
    assert np.abs(got - xo).max() < 1e-7
    # the solution is a fixed point: A rho ~ 0, unit trace, Hermitian
    assert np.abs(m.spmv(got)).max() < 1e-8
    rho = dense.unvec(got, tot)
    assert abs(np.trace(rho) - 1) < 1e-10 and np.abs(rho - rho.conj().T).max() < 1e-8
    qs.destroy()


def _apply_circuit2(circ, x):
    f = qa.lib().apply_circuit_to_qvec2
    f.restype = None
    f.argtypes = [qa.CircuitStruct, ctypes.c_void_p]
    f(circ.c, x.h)
    qa._check()


def test_apply_circuit_to_qvec2_inplace_convention_vs_oracle():
    """apply_circuit_to_qvec2 (quantum_circuits.c:123-140): every gate of the
    circuit with the in-place convention of apply_gate2, against the oracle"""
    nq = 3
    r = {"subs": [("op", 2)] * nq, "terms": [("lin", 0.0, [(0, "n")])], "mcwf": False}
    qs, _, _ = build_system(r, construct=False)
    x = qs.create_qvec()
    rho = _rand_rho(8)
    x.set(dense.vec(rho))
    circ = qa.Circuit(10)
    gates = [("H", 1, (0,), 0), ("CNOT", -1, (0, 1), 0), ("RZ", 8, (2,), 0.4), ("RX", 6, (1,), 2.0)]
    for g in gates:
        circ.add_gate(1.0, g[0], list(g[2]), theta=g[3])
    _apply_circuit2(circ, x)
    ref = dense.vec(rho)
    for g in gates:
        ref = oracle.gate_inplace([2] * nq, True, g[1], g[2], ref, theta=g[3])
    assert np.abs(x.get() - ref).max() < 1e-12
    circ.destroy()
    qs.destroy()


def test_circuits_act_on_every_trajectory_of_a_spawned_ensemble():
    """a spawned MCWF ensemble (quantum_circuits.c:127-132, 146-151): sigma_x on
    qubit 0 flips subsystem 0 of every trajectory with the superoperator
    convention (apply_circuit_to_qvec) and the last subsystem with the in-place
    one (apply_circuit_to_qvec2)"""
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qs = qa.System()
    q0, q1 = qs.create_op(2), qs.create_op(2)
    qs.add_ham_term(0.0, q0.n)
    qs.add_lin_term(0.2, q0)
    qs.use_mcwf_solver(40, 3)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 0)
    qs.time_step(x, 0.0, 0.0, 0.01, 10)        # spawns the trajectories, takes no step
    circ = qa.Circuit(2)
    circ.add_gate(1.0, "SIGMAX", [0])
    qs.apply_circuit_to_qvec(circ, x)
    assert abs(x.expectation(q0.n).real - 1.0) < 1e-12 and abs(x.expectation(q1.n).real) < 1e-12
    _apply_circuit2(circ, x)
    assert abs(x.expectation(q0.n).real - 1.0) < 1e-12 and abs(x.expectation(q1.n).real - 1.0) < 1e-12
    circ.destroy()
    qs.destroy()


def _ens_amp(x, i, loc):
    v = qa.Cx()
    qa.lib().get_wf_element_ens_i_qvec_local(x.h, i, loc, ctypes.byref(v))
    qa._check()
    return complex(v.re, v.im)


def test_qvec_mat_mult_and_per_trajectory_writes():
    """qvec_mat_mult (qvec_utilities.c:169-195) with A = 0.5 a0 a1^dag on two
    qubits, subsystem 0 the most significant digit: A|10> = 0.5|01>, so y[1] =
    0.5 x[2] and every other entry 0 (a transposed A would move entry 1 to 2).
    On a wavefunction, then on an ensemble whose trajectories were written one by
    one with add_to_wf_ens_loc (qvec_utilities.c:899-911): trajectory i holds
    (i + 1) at 2 and 0.3 at 0, the others start at zero, and after A each holds
    0.5 (i + 1) at 1 only.  A time step of zero length afterwards keeps them (the
    ensemble is not overwritten with the initial condition)."""
    qs = qa.System()
    q0, q1 = qs.create_op(2), qs.create_op(2)
    qs.add_ham_term(0.0, q0.n)
    qs.construct_matrix()
    A = qs.op_matrix_wf(0.5, q0, q1.dag)
    x = qs.create_qvec()
    rng = np.random.default_rng(5)
    v = rng.normal(size=4) + 1j * rng.normal(size=4)
    x.set(v)
    qa.lib().qvec_mat_mult(A.h, x.h)
    qa._check()
    want = np.zeros(4, complex)
    want[1] = 0.5 * v[2]
    assert np.abs(x.get() - want).max() < 1e-15, x.get()
    qs.destroy()

    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qs = qa.System()
    q0, q1 = qs.create_op(2), qs.create_op(2)
    qs.add_ham_term(0.0, q0.n)
    qs.add_lin_term(0.2, q0)
    qs.use_mcwf_solver(3, 3)
    qs.construct_matrix()
    A = qs.op_matrix_wf(0.5, q0, q1.dag)
    x = qs.create_qvec()
    for i in range(3):
        qa.lib().add_to_wf_ens_loc(x.h, i, qa.Cx.of(i + 1.0), 2)
        qa.lib().add_to_wf_ens_loc(x.h, i, qa.Cx.of(0.3), 0)
    qa._check()
    for i in range(3):
        got = [_ens_amp(x, i, k) for k in range(4)]
        assert np.abs(np.array(got) - [0.3, 0, i + 1.0, 0]).max() < 1e-15, (i, got)
    qa.lib().qvec_mat_mult(A.h, x.h)
    qa._check()
    qs.time_step(x, 0.0, 0.0, 0.01, 10)
    for i in range(3):
        got = [_ens_amp(x, i, k) for k in range(4)]
        assert np.abs(np.array(got) - [0, 0.5 * (i + 1), 0, 0]).max() < 1e-15, (i, got)
    qs.destroy()


def test_combine_circuit_to_mat_sys_matches_the_oracle():
    """combine_circuit_to_mat_sys (quantum_circuits.c:257-314): the 8 x 8 matrix of
    H(0), CNOT(0 -> 1), RZ(0.4) on 2, RX(2.0) on 1, CNOT(2 -> 0) on three qubits
    equals, column by column, the oracle's gates (qubit q = subsystem q) applied in
    order to the basis states -- H and CNOT do not commute, so a reversed product
    fails -- and is unitary.  The matrix arrives through qh_mat_from_csr, so its
    CSR (sorted columns, 0 < nnz) is read back as well."""
    nq = 3
    r = {"subs": [("op", 2)] * nq, "terms": [("ham", 0.0, [(0, "n")])], "mcwf": False}
    qs, _, _ = build_system(r, construct=False)
    circ = qa.Circuit(10)
    gates = [("H", 1, (0,), 0), ("CNOT", -1, (0, 1), 0), ("RZ", 8, (2,), 0.4), ("RX", 6, (1,), 2.0),
             ("CNOT", -1, (2, 0), 0)]
    for k, g in enumerate(gates):
        circ.add_gate(1.0 + k, g[0], list(g[2]), theta=g[3])
    h = ctypes.c_void_p()
    qa.lib().combine_circuit_to_mat_sys(qs.h, ctypes.byref(h), circ.c)
    qa._check()
    m = qa.Matrix(h, owned=True)
    rp, col, val = m.export()
    U = np.zeros((8, 8), complex)
    for row in range(8):
        assert np.all(np.diff(col[rp[row]:rp[row + 1]]) > 0)
        U[row, col[rp[row]:rp[row + 1]]] = val[rp[row]:rp[row + 1]]
    want = np.zeros((8, 8), complex)
    for j in range(8):
        e = np.zeros(8, complex)
        e[j] = 1.0
        for g in gates:
            e = oracle.gate_superop([2] * nq, [1, 2, 4], False, g[1], g[2], e, theta=g[3])
        want[:, j] = e
    assert np.abs(U - want).max() < 1e-14, np.abs(U - want).max()
    assert np.abs(U @ U.conj().T - np.eye(8)).max() < 1e-14
    del m
    circ.destroy()
    qs.destroy()


@pytest.mark.parametrize("kind", ["wf", "dm"])
def test_apply_op_to_qvec_mat_matches_krons(kind):
    """apply_op_to_qvec_mat (qvec_utilities.c:1553-1589; rows from kron.c:209-266):
    a dense complex 3 x 3 mat on the middle subsystem of qubit (x) qutrit (x)
    qubit, subsystem 0 the most significant digit: psi <- (I2 (x) mat (x) I2) psi,
    rho <- M rho M^dag, against numpy krons.  A dense mat on the middle subsystem
    catches a wrong stride, a transposed mat or a missing conjugate."""
    rng = np.random.default_rng(11)
    term = ("lin", 0.1, [(0, "lower")]) if kind == "dm" else ("ham", 0.1, [(0, "n")])
    qs, subs, _ = build_system({"subs": [("op", 2), ("op", 3), ("op", 2)], "terms": [term], "mcwf": False})
    x = qs.create_qvec()
    mat = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    rows = [(qa.Cx * 3)(*[qa.Cx.of(v) for v in mat[r]]) for r in range(3)]
    ptrs = (ctypes.POINTER(qa.Cx) * 3)(*[ctypes.cast(row, ctypes.POINTER(qa.Cx)) for row in rows])
    M = np.kron(np.kron(np.eye(2), mat), np.eye(2))
    if kind == "wf":
        v = rng.normal(size=12) + 1j * rng.normal(size=12)
        x.set(v)
        want = M @ v
    else:
        rho = _rand_rho(12)
        x.set(dense.vec(rho))
        want = dense.vec(M @ rho @ M.conj().T)
    qa.lib().apply_op_to_qvec_mat(x.h, ptrs, ctypes.cast(subs[1][1].ptr, ctypes.c_void_p))
    qa._check()
    assert np.abs(x.get() - want).max() < 1e-12 * max(1.0, np.abs(want).max())
    qs.destroy()
