"""Pins for the oracle functions that the pristine fixtures and the basic
closed forms leave open:

* every gate code of the oracle against the reference's hard-coded known-answer
  rows (tests/quantum_gates_test_sys.c) and, for the rotations and the U family,
  against exp(-i theta P / 2) and the OpenQASM U3 definition;
* qo_expect_wf against dense <psi|O|psi> and the |1,1> known answers of
  tests/qvec_utilities_tests.c:232-278;
* qo_rk3bs_trial against the closed forms of one Bogacki-Shampine step on a
  linear ODE (candidate and embedded error as polynomials in z = hA);
* qo_rk4_td against the analytic Rabi drive of tests/time_dep_tests.c:46-64,
  142-213 and against qo_rk4 on the summed matrix when every envelope is 1.
Each check is chosen so that a dropped term, a wrong sign or a transposed
operand in the oracle fails it.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import dense
import oracle
import recipes

RNG = np.random.default_rng(2024)
H_, SX, SY, SZ, EYE, RX, RY, RZ, U1, U2, U3 = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11
CNOT, CXZ, CZ = -1, -2, -3
S2 = 2 ** -0.5


def _wf_matrix(nq, gates):
    """the Hilbert-space matrix of a gate sequence (superoperator-path qubit
    numbering: qubit q = subsystem q, the most significant for q = 0), built from
    the oracle's WF action on the basis vectors"""
    N = 2 ** nq
    cols = []
    for j in range(N):
        e = np.zeros(N, np.complex128)
        e[j] = 1
        for (g, qs, th, ph, la) in gates:
            e = oracle.gate_superop([2] * nq, [2 ** s for s in range(nq)], False, g, qs, e, theta=th, phi=ph,
                                    lam=la)
        cols.append(e)
    return np.array(cols).T


def _rows(N, entries):
    """dense matrix from the known-answer rows {(row, col): value}"""
    M = np.zeros((N, N), np.complex128)
    for (r, c), v in entries.items():
        M[r, c] = v
    return M


# quantum_gates_test_sys.c: (line, qubits of the system, gate sequence, known rows)
KNOWN = [
    (1227, 1, [(SX, [0])], {(0, 1): 1, (1, 0): 1}),
    (577, 2, [(SX, [0])], {(0, 2): 1, (1, 3): 1, (2, 0): 1, (3, 1): 1}),
    (642, 2, [(SX, [1])], {(0, 1): 1, (1, 0): 1, (2, 3): 1, (3, 2): 1}),
    (687, 2, [(SX, [0]), (SX, [1])], {(0, 3): 1, (1, 2): 1, (2, 1): 1, (3, 0): 1}),
    (740, 2, [(SY, [0])], {(0, 2): -1j, (1, 3): -1j, (2, 0): 1j, (3, 1): 1j}),
    (805, 2, [(SY, [1])], {(0, 1): -1j, (1, 0): 1j, (2, 3): -1j, (3, 2): 1j}),
    (850, 2, [(SY, [0]), (SY, [1])], {(0, 3): -1, (1, 2): 1, (2, 1): 1, (3, 0): -1}),
    # test_sigmay (:1172) adds SIGMAZ, so its sigma-y rows never apply there; the
    # rows themselves are the 1-qubit sigma-y of :1201-1208
    (1201, 1, [(SY, [0])], {(0, 1): -1j, (1, 0): 1j}),
    (16, 1, [(SZ, [0])], {(0, 0): 1, (1, 1): -1}),
    (71, 2, [(SZ, [0])], {(0, 0): 1, (1, 1): 1, (2, 2): -1, (3, 3): -1}),
    (134, 2, [(SZ, [1])], {(0, 0): 1, (1, 1): -1, (2, 2): 1, (3, 3): -1}),
    (177, 2, [(SZ, [0]), (SZ, [1])], {(0, 0): 1, (1, 1): -1, (2, 2): -1, (3, 3): 1}),
    (1283, 1, [(H_, [0])], {(0, 0): S2, (0, 1): S2, (1, 0): S2, (1, 1): -S2}),
    (903, 2, [(H_, [0])], {(0, 0): S2, (0, 2): S2, (1, 1): S2, (1, 3): S2, (2, 0): S2, (2, 2): -S2,
                           (3, 1): S2, (3, 3): -S2}),
    (1009, 2, [(H_, [1])], {(0, 0): S2, (0, 1): S2, (1, 0): S2, (1, 1): -S2, (2, 2): S2, (2, 3): S2,
                            (3, 2): S2, (3, 3): -S2}),
    (1095, 2, [(H_, [0]), (H_, [1])],
     {(r, c): 0.5 * (-1) ** bin(r & c).count("1") for r in range(4) for c in range(4)}),
    (228, 2, [(CNOT, [0, 1])], {(0, 0): 1, (1, 1): 1, (2, 3): 1, (3, 2): 1}),
    (293, 2, [(CNOT, [1, 0])], {(0, 0): 1, (1, 3): 1, (2, 2): 1, (3, 1): 1}),
    (344, 2, [(CXZ, [0, 1])], {(0, 0): 1, (1, 1): 1, (2, 3): -1, (3, 2): 1}),
    (409, 2, [(CXZ, [1, 0])], {(0, 0): 1, (1, 3): -1, (2, 2): 1, (3, 1): 1}),
    (460, 2, [(CZ, [0, 1])], {(0, 0): 1, (1, 1): 1, (2, 2): 1, (3, 3): -1}),
    (525, 2, [(CZ, [1, 0])], {(0, 0): 1, (1, 1): 1, (2, 2): 1, (3, 3): -1}),
]


@pytest.mark.parametrize("line,nq,seq,rows", KNOWN, ids=[f"line{k[0]}" for k in KNOWN])
def test_gate_known_answer_rows(line, nq, seq, rows):
    """the reference's combine_circuit_to_mat_sys rows: values to 1e-10 and the
    number of entries per row (ncols) as the reference checks them"""
    got = _wf_matrix(nq, [(g, qs, 0.0, 0.0, 0.0) for (g, qs) in seq])
    want = _rows(2 ** nq, rows)
    assert np.abs(got - want).max() < 1e-10
    assert [int((np.abs(r) > 1e-14).sum()) for r in got] == [int((np.abs(r) > 0).sum()) for r in want]


PX = np.array([[0, 1], [1, 0]], np.complex128)
PY = np.array([[0, -1j], [1j, 0]], np.complex128)
PZ = np.array([[1, 0], [0, -1]], np.complex128)


def _u3_qasm(t, p, l):
    """OpenQASM 2 U3(theta, phi, lambda) (Cross, Bishop, Smolin, Gambetta 2017, eq. 2)"""
    return np.array([[math.cos(t / 2), -np.exp(1j * l) * math.sin(t / 2)],
                     [np.exp(1j * p) * math.sin(t / 2), np.exp(1j * (p + l)) * math.cos(t / 2)]])


@pytest.mark.parametrize("theta", [0.0, 0.3, 1.7, -2.9, 5.5])
def test_rotation_gates_are_pauli_exponentials(theta):
    one = lambda g, th, ph=0.0, la=0.0: _wf_matrix(1, [(g, [0], th, ph, la)])
    assert np.abs(one(RX, theta) - expm(-0.5j * theta * PX)).max() < 1e-14
    assert np.abs(one(RY, theta) - expm(-0.5j * theta * PY)).max() < 1e-14
    # QuaC's RZ is diag(1, e^{i theta}) (quantum_circuits.c:2489-2516): the
    # Pauli exponential up to the global phase e^{i theta/2}
    assert np.abs(one(RZ, theta) - np.exp(0.5j * theta) * expm(-0.5j * theta * PZ)).max() < 1e-14
    assert np.abs(one(EYE, theta) - np.eye(2)).max() == 0.0


@pytest.mark.parametrize("t,p,l", [(0.4, 1.1, -0.7), (2.2, -2.5, 0.3), (math.pi / 2, 0.0, math.pi)])
def test_u_family_is_qasm_u3(t, p, l):
    for g in (U1, U2, U3):  # the oracle takes theta, phi, lambda as given for all three
        got = _wf_matrix(1, [(g, [0], t, p, l)])
        assert np.abs(got - _u3_qasm(t, p, l)).max() < 1e-14
        assert np.abs(got.conj().T @ got - np.eye(2)).max() < 1e-14
    # identities tying U3 to the Pauli rotations: U3(t,-pi/2,pi/2) = RX(t), U3(t,0,0) = RY(t)
    assert np.abs(_wf_matrix(1, [(U3, [0], t, -math.pi / 2, math.pi / 2)]) - expm(-0.5j * t * PX)).max() < 1e-14
    assert np.abs(_wf_matrix(1, [(U3, [0], t, 0.0, 0.0)]) - expm(-0.5j * t * PY)).max() < 1e-14


def _embed(U, subs, nsub):
    N = 2 ** nsub
    M = np.zeros((N, N), np.complex128)
    k = len(subs)
    for i in range(N):
        bits = [(i >> (nsub - 1 - s)) & 1 for s in range(nsub)]
        si = 0
        for q in subs:
            si = 2 * si + bits[q]
        for sj in range(2 ** k):
            b2 = list(bits)
            for t, q in enumerate(subs):
                b2[q] = (sj >> (k - 1 - t)) & 1
            M[i, int("".join(map(str, b2)), 2)] += U[si, sj]
    return M


GATE_U = {SY: PY, SZ: PZ, EYE: np.eye(2), CZ: np.diag([1, 1, 1, -1]).astype(np.complex128),
          CXZ: _rows(4, {(0, 0): 1, (1, 1): 1, (2, 3): -1, (3, 2): 1})}


@pytest.mark.parametrize("gate,qubits,theta", [(SY, (1,), 0.0), (SZ, (2,), 0.0), (EYE, (0,), 0.0),
                                               (RY, (1,), 0.9), (U3, (2,), 1.3), (CZ, (2, 0), 0.0),
                                               (CXZ, (0, 2), 0.0), (CXZ, (2, 1), 0.0)])
def test_dm_gates_both_conventions(gate, qubits, theta):
    """U rho U^dagger on a 3-qubit density matrix for the gate codes the GPU
    tests use beyond H/X/RX/RZ/CNOT: superoperator path (qubit q = subsystem q)
    and in-place path (qubit q = subsystem n-1-q)"""
    nsub, N = 3, 8
    a = RNG.normal(size=(N, N)) + 1j * RNG.normal(size=(N, N))
    rho = a @ a.conj().T
    if gate == RY:
        U = expm(-0.5j * theta * PY)
    elif gate == U3:
        U = _u3_qasm(theta, 0.4, -1.1)
    else:
        U = GATE_U[gate]
    Us = _embed(U, list(qubits), nsub)
    got = oracle.gate_superop([2] * nsub, [1, 2, 4], True, gate, qubits, dense.vec(rho), theta=theta, phi=0.4,
                              lam=-1.1)
    assert np.abs(dense.unvec(got, N) - Us @ rho @ Us.conj().T).max() < 1e-12
    Ui = _embed(U, [nsub - 1 - q for q in qubits], nsub)
    got = oracle.gate_inplace([2] * nsub, True, gate, qubits, dense.vec(rho), theta=theta, phi=0.4, lam=-1.1)
    assert np.abs(dense.unvec(got, N) - Ui @ rho @ Ui.conj().T).max() < 1e-12


# ---------------------------------------------------------------- expect_wf
def test_expect_wf_known_answers_fock_11():
    """qvec_utilities_tests.c:232-278: |1,1> of two qubits: <n1> = <n2> = 1,
    <a1^dag a2> = <a2^dag a1> = 0"""
    r = {"subs": [("op", 2), ("op", 2)], "terms": []}
    psi = np.zeros(4, np.complex128)
    psi[3] = 1.0
    for ops, want in (([(0, "n")], 1.0), ([(1, "n")], 1.0), ([(0, "raise"), (1, "lower")], 0.0),
                      ([(1, "raise"), (0, "lower")], 0.0)):
        _, terms = recipes.oracle_terms({"subs": r["subs"], "terms": [("ham", 1.0, ops)]})
        assert oracle.expect_wf(4, psi, terms[0][2]) == want


def test_expect_wf_brute_force():
    r = {"subs": [("op", 2), ("op", 3), ("vec", 3)], "terms": []}
    tot, _ = recipes.layout(r)
    psi = RNG.normal(size=tot) + 1j * RNG.normal(size=tot)
    for ops in ([(0, "lower"), (1, "raise")], [(1, "n")], [(0, "sy")], [(0, "sx"), (0, "sz")],
                [(2, ("vec", 1)), (2, ("vec", 2)), (1, "lower")], [(1, "other"), (0, "raise")]):
        _, terms = recipes.oracle_terms({"subs": r["subs"], "terms": [("ham", 1.0, ops)]})
        got = oracle.expect_wf(tot, psi, terms[0][2])
        ref = psi.conj() @ dense.chain_matrix(r, ops) @ psi
        assert abs(got - ref) < 1e-12 * max(1.0, abs(ref))


# ------------------------------------------------------------- rk3bs_trial
def _dense_A(seed):
    rng = np.random.default_rng(seed)
    r = {"subs": [("op", 2), ("op", 3)],
         "terms": [("ham", float(rng.normal()), [(0, "sx")]), ("ham", complex(rng.normal(), rng.normal()),
                                                              [(0, "raise"), (1, "lower")]),
                   ("ham", 0.7, [(1, "n")]), ("lin", 0.3, [(1, "lower")]), ("lin", 0.2, [(0, "sz")])]}
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    rp, col, val = m.export()
    n = len(rp) - 1
    A = np.zeros((n, n), np.complex128)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            A[i, col[k]] += val[k]
    return m, A


@pytest.mark.parametrize("seed,h", [(0, 0.05), (1, 0.2), (2, 0.6)])
def test_rk3bs_trial_closed_forms(seed, h):
    """One BS3 step on x' = A x: xn = (1 + z + z^2/2 + z^3/6) x and the embedded
    error e = xn - x_hat = -(z^3 + z^4)/48 x (z = hA; b = (2/9, 1/3, 4/9, 0),
    b_hat = (7/24, 1/4, 1/3, 1/8)); err is its weighted RMS norm."""
    m, A = _dense_A(seed)
    n = A.shape[0]
    x = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    atol, rtol = 1e-6, 1e-4
    xn, err, ks = m.rk3bs_trial(x, h, atol=atol, rtol=rtol)
    Z = h * A
    Z2 = Z @ Z
    Z3 = Z2 @ Z
    Z4 = Z3 @ Z
    xn_cf = x + Z @ x + Z2 @ x / 2 + Z3 @ x / 6
    assert np.abs(xn - xn_cf).max() < 1e-12 * np.abs(xn_cf).max()
    e_cf = -(Z3 @ x + Z4 @ x) / 48
    sc = atol + rtol * np.maximum(np.abs(x), np.abs(xn_cf))
    err_cf = math.sqrt(np.mean(np.abs(e_cf) ** 2 / sc ** 2))
    assert abs(err - err_cf) < 1e-9 * err_cf
    assert np.abs(ks[0] - A @ x).max() < 1e-12 * np.abs(A @ x).max()
    assert np.abs(ks[3] - A @ xn_cf).max() < 1e-11 * np.abs(A @ xn_cf).max()
    # FSAL: k1 passed in gives the same step
    xn2, err2, _ = m.rk3bs_trial(x, h, atol=atol, rtol=rtol, k1=ks[0])
    assert np.array_equal(xn2, xn) and err2 == err


# ------------------------------------------------------------------ rk4_td
def _rabi(dm):
    om, al = 2 * math.pi, 0.1
    r = {"subs": [("op", 2)], "terms": [("td_ham", 0.5 * om, [(0, "lower")]), ("td_ham", 0.5 * om, [(0, "raise")])],
         "dm": dm}
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=dm)
    td = [oracle.construct(tot, terms, dm=dm, td_index=k) for k in range(2)]
    return om, al, m, td, [t[1] for t in terms]


@pytest.mark.parametrize("dm", [False, True])
def test_rk4_td_rabi_closed_form(dm):
    """time_dep_tests.c:142-213: H(t) = 0.5 omega e^{-alpha t} (a + a^dag), psi(0) = |0>:
    <sz> = cos(omega (1 - e^{-alpha t})/alpha), <sy> = -sin(...), <sx> = 0 (the
    reference allows 1e-4).  The error at t = 20 falls 16x when dt halves: the
    envelope is evaluated at the stage times t, t + h/2, t + h/2, t + h (an
    envelope taken at the wrong stage time drops the order below 4)."""
    om, al, m, td, a = _rabi(dm)
    f = lambda t, k: math.exp(-al * t)
    errs = []
    for dt in (0.02, 0.01):
        x = np.zeros(4 if dm else 2, np.complex128)
        x[0] = 1.0
        t, worst = 0.0, 0.0
        nblk = int(round(2.0 / dt))
        for _ in range(10):
            x = m.rk4_td(td, a, f, x, t, dt, nblk, wf_scale=not dm)
            t += nblk * dt
            rho = dense.unvec(x, 2) if dm else np.outer(x, x.conj())
            th = om * (1 - math.exp(-al * t)) / al
            for P, want in ((PZ, math.cos(th)), (PY, -math.sin(th)), (PX, 0.0)):
                worst = max(worst, abs(np.trace(P @ rho) - want))
        errs.append(worst)
    assert errs[1] < (2e-7 if not dm else 3e-6), errs
    assert 12.0 < errs[0] / errs[1] < 20.0, errs


@pytest.mark.parametrize("dm", [False, True])
def test_rk4_td_unit_envelopes_equal_rk4_on_summed_matrix(dm):
    """f_k = 1: A + sum_k a_k A_k (x -i for a WF) is the matrix of the same terms
    added as time-independent ones, so qo_rk4_td equals qo_rk4 on it"""
    r = {"subs": [("op", 2), ("op", 3)],
         "terms": [("ham", 0.4, [(1, "n")]), ("td_ham", 0.3 + 0.2j, [(0, "raise"), (1, "lower")]),
                   ("td_ham", 0.3 - 0.2j, [(0, "lower"), (1, "raise")]), ("td_ham", 0.9, [(0, "sz")])]
         + ([("lin", 0.25, [(1, "lower")])] if dm else []), "dm": dm}
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=dm)
    idx = [k for k, t in enumerate(terms) if t[0] < 0]
    td = [oracle.construct(tot, terms, dm=dm, td_index=k) for k in idx]
    a = [terms[k][1] for k in idx]
    const = [(abs(t[0]), t[1], t[2]) for t in terms]
    ms = oracle.construct(tot, const, dm=dm)
    n = ms.n
    x0 = RNG.normal(size=n) + 1j * RNG.normal(size=n)
    got = m.rk4_td(td, a, lambda t, k: 1.0, x0, 0.0, 0.01, 25, wf_scale=not dm)
    want = ms.rk4(x0, 0.01, 25)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
