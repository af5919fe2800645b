"""Pin the oracle's functions that the pristine fixtures do not cover, against
what the mathematics fixes: dense Lindbladians (VEC pairs, OTHER, mixed levels,
complex coefficients), the degree-4 Taylor polynomial of exp(hA) for RK4,
closed-form T1 decay (time_step_tests.c:24-101), Tr(O rho) brute force,
textbook U rho U^dagger gates in both qubit conventions, and the analytic
driven-damped two-level steady state for GMRES.
"""
import math

import numpy as np
import pytest

import dense
import oracle
import recipes

RNG = np.random.default_rng(12345)


def _random_recipe(seed, dm=True):
    rng = np.random.default_rng(seed)
    subs = [("op", 2), ("op", 3), ("vec", 3)]
    names2 = ["lower", "raise", "n", "eye", "sx", "sy", "sz", "other"]
    names3 = ["lower", "raise", "n", "eye", "sx", "other"]
    terms = []
    for t in range(6):
        ops = []
        for _ in range(rng.integers(1, 4)):
            s = int(rng.integers(0, 3))
            if subs[s][0] == "vec":
                ops += [(s, ("vec", int(rng.integers(0, 3)))), (s, ("vec", int(rng.integers(0, 3))))]
            else:
                pool = names2 if subs[s][1] == 2 else names3
                ops.append((s, pool[int(rng.integers(0, len(pool)))]))
        kind = "ham" if (t % 2 == 0 or not dm) else "lin"
        a = complex(rng.normal(), rng.normal()) if t % 3 else float(rng.normal())
        terms.append((kind, a, ops))
    return {"subs": subs, "terms": terms, "mcwf": False, "dm": dm}


def _dense_from_oracle(m):
    rp, col, val = m.export()
    n = len(rp) - 1
    A = np.zeros((n, n), np.complex128)
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            A[r, col[k]] += val[k]
    return A, rp, col


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("conj_a", (False, True))
def test_construct_matches_dense_lindbladian(seed, conj_a):
    r = _random_recipe(seed)
    tot, terms = recipes.oracle_terms(r)
    A, rp, col = _dense_from_oracle(oracle.construct(tot, terms, dm=True, conj_a=conj_a))
    ref = dense.superop(r, conj_a=conj_a)
    assert np.abs(A - ref).max() < 1e-12
    # every diagonal is stored, columns sorted and unique per row
    for i in range(len(rp) - 1):
        c = col[rp[i]:rp[i + 1]]
        assert i in c and np.all(np.diff(c) > 0)


@pytest.mark.parametrize("seed", range(4))
def test_construct_schrodinger_and_mcwf_dense(seed):
    r = _random_recipe(100 + seed, dm=False)
    tot, terms = recipes.oracle_terms(r)
    A, _, _ = _dense_from_oracle(oracle.construct(tot, terms, dm=False))
    assert np.abs(A - dense.superop(r)).max() < 1e-12
    r2 = _random_recipe(200 + seed, dm=True)
    r2["mcwf"] = True
    tot, terms = recipes.oracle_terms(r2)
    A, _, _ = _dense_from_oracle(oracle.construct(tot, terms, dm=True, mcwf=True))
    assert np.abs(A - dense.superop(r2)).max() < 1e-12


def test_spin_chain_structure_count():
    # n qubits, sx Hamiltonian + sz sz + sigma- and sz dissipators:
    # nnz = N^2 (1 + 2n) + n N^2/4 (diag, I x sx and sx^T x I per qubit, G*xG of sigma-)
    for n in (2, 3, 4):
        r = recipes.spin_chain_recipe(n, seed=1)
        tot, terms = recipes.oracle_terms(r)
        m = oracle.construct(tot, terms, dm=True)
        N2 = tot * tot
        assert m.nnz == N2 * (1 + 2 * n) + n * N2 // 4
    # JC 2x10 (400-row superoperator) has 1544 nnz
    r = recipes.jc_recipe()
    tot, terms = recipes.oracle_terms(r)
    assert oracle.construct(tot, terms, dm=True).nnz == 1544


def test_rk4_is_degree4_taylor():
    r = _random_recipe(7)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    A, _, _ = _dense_from_oracle(m)
    x0 = RNG.normal(size=tot * tot) + 1j * RNG.normal(size=tot * tot)
    h, steps = 0.01, 7
    P = np.eye(len(x0), dtype=np.complex128)
    T = np.eye(len(x0), dtype=np.complex128)
    for k in range(1, 5):
        T = T @ (h * A) / k
        P = P + T
    ref = x0.copy()
    for _ in range(steps):
        ref = P @ ref
    got = m.rk4(x0, h, steps)
    assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def _t1_recipe(vec=False):
    if vec:
        return {"subs": [("vec", 2)], "terms": [("lin", 1.0, [(0, ("vec", 0)), (0, ("vec", 1))])], "mcwf": False}
    return {"subs": [("op", 2)], "terms": [("lin", 1.0, [(0, "lower")])], "mcwf": False}


@pytest.mark.parametrize("vec", (False, True))
def test_t1_decay_closed_form(vec):
    # time_step_tests.c:24-101: <n>(1) = 1/e within 1e-4
    r = _t1_recipe(vec)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    x = np.zeros(4, np.complex128)
    x[1 * 2 + 1] = 1.0
    out = m.rk4(x, 0.01, 100)
    ops = [(2, 2, 1, 1), (2, 2, 1, 1)] if vec else [(0, 2, 1, -1)]
    assert abs(oracle.expect_dm(2, out, ops).real - math.exp(-1)) < 1e-4
    out2, steps, tend = m.rk3bs(x, 0.0, 1.0, 0.05, 200)
    assert abs(tend - 1.0) < 1e-12
    assert abs(oracle.expect_dm(2, out2, ops).real - math.exp(-1)) < 1e-4


def test_expect_dm_brute_force():
    r = {"subs": [("op", 2), ("op", 3), ("vec", 3)], "terms": [], "mcwf": False}
    tot, lay = recipes.layout(r)
    rho = RNG.normal(size=(tot, tot)) + 1j * RNG.normal(size=(tot, tot))
    x = dense.vec(rho)
    for ops in ([(0, "lower"), (1, "raise")], [(1, "n")], [(0, "sx"), (0, "sy")],
                [(2, ("vec", 1)), (2, ("vec", 2)), (1, "lower")]):
        _, terms = recipes.oracle_terms({"subs": r["subs"], "terms": [("ham", 1.0, ops)]})
        got = oracle.expect_dm(tot, x, terms[0][2])
        ref = np.trace(dense.chain_matrix(r, ops) @ rho)
        assert abs(got - ref) < 1e-12


def test_read_dm_binary_known_values():
    # qvec_utilities_tests.c:788-870 reads rand_dm.dat; the data themselves:
    # a 4x4 Hermitian unit-trace density matrix
    import os
    import petsc_io
    v = petsc_io.read_vec(os.path.join(os.path.dirname(__file__), "golden", "pristine_matrices", "rand_dm.dat"))
    rho = dense.unvec(v, 4)
    assert abs(np.trace(rho) - 1) < 1e-12
    assert np.abs(rho - rho.conj().T).max() < 1e-12
    ops = [(1, 2, 1, -1)]  # lower on the first qubit of a 2-qubit space
    got = oracle.expect_dm(4, v, ops)
    lay = {"subs": [("op", 2), ("op", 2)], "terms": []}
    assert abs(got - np.trace(dense.chain_matrix(lay, [(0, "lower")]) @ rho)) < 1e-14


def _u(gate, theta):
    s = 2 ** -0.5
    return {1: np.array([[s, s], [s, -s]]), 2: np.array([[0, 1], [1, 0]]),
            6: np.array([[math.cos(theta / 2), -1j * math.sin(theta / 2)],
                         [-1j * math.sin(theta / 2), math.cos(theta / 2)]]),
            8: np.diag([1, np.exp(1j * theta)]),
            -1: np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])}[gate].astype(np.complex128)


def _embed_gate(U, qs_sub, nsub):
    """U on subsystems qs_sub (MSB-first basis of 2-level subsystems)."""
    N = 2 ** nsub
    M = np.zeros((N, N), np.complex128)
    k = len(qs_sub)
    for i in range(N):
        bits = [(i >> (nsub - 1 - s)) & 1 for s in range(nsub)]
        si = 0
        for q in qs_sub:
            si = 2 * si + bits[q]
        for sj in range(2 ** k):
            b2 = list(bits)
            for t, q in enumerate(qs_sub):
                b2[q] = (sj >> (k - 1 - t)) & 1
            j = 0
            for b in b2:
                j = 2 * j + b
            M[i, j] += U[si, sj]
    return M


@pytest.mark.parametrize("gate,qubits,theta", [(1, (0,), 0.0), (1, (2,), 0.0), (8, (1,), 0.7), (6, (0,), 1.3),
                                               (2, (2,), 0.0), (-1, (0, 1), 0.0), (-1, (2, 0), 0.0)])
def test_gates_textbook(gate, qubits, theta):
    nsub = 3
    N = 8
    rho = RNG.normal(size=(N, N)) + 1j * RNG.normal(size=(N, N))
    rho = rho @ rho.conj().T
    U = _u(gate, theta)
    # superoperator path: qubit q -> subsystem q
    Us = _embed_gate(U, list(qubits), nsub)
    got = oracle.gate_superop([2] * nsub, [1, 2, 4], True, gate, qubits, dense.vec(rho), theta=theta)
    assert np.abs(dense.unvec(got, N) - Us @ rho @ Us.conj().T).max() < 1e-12
    # in-place path: qubit q -> subsystem nsub-1-q
    Ui = _embed_gate(U, [nsub - 1 - q for q in qubits], nsub)
    got = oracle.gate_inplace([2] * nsub, True, gate, qubits, dense.vec(rho), theta=theta)
    assert np.abs(dense.unvec(got, N) - Ui @ rho @ Ui.conj().T).max() < 1e-12


def test_gate_known_rows_cnot_hadamard():
    # quantum_gates_test_sys.c:228-340 (CNOT, control first) and the Hadamard
    # rows: single-qubit-space matrices U must equal the textbook ones.
    for gate, U in ((-1, _u(-1, 0)), (1, _u(1, 0))):
        nq = 2 if gate < 0 else 1
        N = 2 ** nq
        cols = []
        for j in range(N):
            e = np.zeros(N, np.complex128)
            e[j] = 1
            cols.append(oracle.gate_superop([2] * nq, [1, 2][:nq], False, gate, list(range(nq)), e))
        assert np.abs(np.array(cols).T - U).max() < 1e-15


@pytest.mark.parametrize("omega,delta,gamma", [(1.0, 0.0, 0.5), (0.4, 0.3, 1.0)])
def test_gmres_two_level_steady_state(omega, delta, gamma):
    r = {"subs": [("op", 2)],
         "terms": [("ham", omega / 2, [(0, "sx")]), ("ham", delta, [(0, "n")]), ("lin", gamma, [(0, "lower")])],
         "mcwf": False}
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    for jac in (False, True):
        x, it, res = m.gmres(tot, tol=1e-12, jacobi=jac)
        assert res < 1e-11
        rho = dense.unvec(x, 2)
        ree = (omega ** 2 / 4) / (delta ** 2 + gamma ** 2 / 4 + omega ** 2 / 2)
        assert abs(rho[1, 1].real - ree) < 1e-10
        assert abs(np.trace(rho) - 1) < 1e-10
