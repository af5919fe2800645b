"""State utilities of qvec_utilities.c on device states: partial trace,
|psi><psi|, fidelity, superfidelity, Hilbert-Schmidt distance, equality.

Pinned by the reference's own known-answer tests (tests/qvec_utilities_tests.c,
values restated here with their line numbers) and by the textbook definitions
computed independently with numpy / scipy on random complex states."""
import ctypes
import math

import numpy as np
import pytest
import scipy.linalg

import quac_amd as qa

pytestmark = pytest.mark.gpu
DM, WF = -1, 1


def _arb(dims, kind):
    h = ctypes.c_void_p()
    arr = (ctypes.c_int64 * len(dims))(*dims)
    qa.lib().create_arb_qvec_dims(ctypes.byref(h), len(dims), arr, kind)
    qa._check()
    return qa.QVec(h)


def _free(*qs):
    for q in qs:
        h = ctypes.c_void_p(q.h.value if isinstance(q.h, ctypes.c_void_p) else q.h)
        qa.lib().destroy_qvec(ctypes.byref(h))


def _ptrace(q, traced):
    out = ctypes.c_void_p()
    lst = (ctypes.c_int64 * len(traced))(*traced)
    qa.lib().ptrace_over_list_qvec(q.h, len(traced), lst, ctypes.byref(out))
    qa._check()
    return qa.QVec(out)


def _rho(q):
    """the density matrix (row, col) of a DM qvec (entry N col + row)"""
    n = q.total_levels
    return q.get().reshape(n, n).T


def _fid(a, b):
    f, v = ctypes.c_double(), ctypes.c_double()
    qa.lib().get_fidelity_qvec(a.h, b.h, ctypes.byref(f), ctypes.byref(v))
    qa._check()
    return f.value


def _rand_dm(rng, n, rank=None):
    a = rng.normal(size=(n, rank or n)) + 1j * rng.normal(size=(n, rank or n))
    r = a @ a.conj().T
    return r / np.trace(r)


def _set_dm(q, rho):
    q.set(rho.T.reshape(-1))


def test_ptrace_wavefunction_reference_vectors():
    """qvec_utilities_tests.c:380-437 (test_ptrace_wf_22_rand): a random
    two-qubit wavefunction, each qubit traced away in turn"""
    wf = _arb([2, 2], WF)
    for loc, v in enumerate([0.46076709 - 0.32805127j, -0.34063756 + 0.22835621j, 0.28068659 - 0.12410744j,
                             0.62797233 + 0.15283547j]):
        wf.add(v, loc)
    r1 = _rho(_ptrace(wf, [0]))
    want1 = np.array([[0.41411156544486266, -0.07457172896962921 - 0.11430734736991019j],
                      [-0.07457172896962921 + 0.11430734736991019j, 0.5858884340640674]])
    r0 = _rho(_ptrace(wf, [1]))
    want0 = np.array([[0.48810445290399873, -0.008965286971174186 + 0.1605679145006028j],
                      [-0.008965286971174186 - 0.1605679145006028j, 0.5118955466049314]])
    assert np.abs(r1 - want1).max() < 1e-7 and np.abs(r0 - want0).max() < 1e-7
    _free(wf)


@pytest.mark.parametrize("dims,bell,keep_q,keep_3", [
    ([3, 2, 2], (0, 3), [[0, 2], [0, 1]], [1, 2]),      # qvec_utilities_tests.c:590-649 (test_ptrace_dm_322)
    ([2, 2, 3], (0, 9), [[0, 2], [1, 2]], [0, 1]),      # :711-775 (test_ptrace_dm_223)
])
def test_ptrace_dm_reference_bell_cases(dims, bell, keep_q, keep_3):
    """a Bell pair of the two qubits next to a qutrit in |0>: tracing the
    qutrit and either qubit leaves I/2, tracing both qubits leaves |0><0|"""
    dm = _arb(dims + dims, DM)
    for i in bell:
        for j in bell:
            dm.add(0.5, i, j)
    for traced in keep_q:
        assert np.abs(_rho(_ptrace(dm, traced)) - 0.5 * np.eye(2)).max() < 1e-15
    r = _rho(_ptrace(dm, keep_3))
    want = np.zeros((3, 3))
    want[0, 0] = 1
    assert np.abs(r - want).max() < 1e-15
    _free(dm)


@pytest.mark.parametrize("kind", ["dm", "wf"])
def test_ptrace_random_states_vs_numpy(kind):
    """random complex states of (2, 3, 2) subsystems; every subset traced,
    including keeping two subsystems, against numpy's partial trace"""
    rng = np.random.default_rng(3)
    dims = [2, 3, 2]
    n = int(np.prod(dims))
    if kind == "dm":
        rho = _rand_dm(rng, n)
        q = _arb(dims + dims, DM)
        _set_dm(q, rho)
    else:
        psi = rng.normal(size=n) + 1j * rng.normal(size=n)
        psi /= np.linalg.norm(psi)
        rho = np.outer(psi, psi.conj())
        q = _arb(dims, WF)
        q.set(psi)
    t = rho.reshape(dims + dims)
    for traced in ([0], [1], [2], [0, 2], [1, 0], [2, 1]):
        keep = [k for k in range(3) if k not in traced]
        letters = "abc"
        rows = "".join(letters[k] if k in keep else "xyz"[k] for k in range(3))
        cols = "".join(letters[k].upper() if k in keep else "xyz"[k] for k in range(3))
        out = "".join(letters[k] for k in keep) + "".join(letters[k].upper() for k in keep)
        nk = int(np.prod([dims[k] for k in keep]))
        want = np.einsum(rows + cols + "->" + out, t).reshape(nk, nk)
        got_q = _ptrace(q, traced)
        assert got_q.total_levels == nk
        assert np.abs(_rho(got_q) - want).max() < 1e-14, traced
        _free(got_q)
    _free(q)


def test_fidelity_reference_known_answers():
    """qvec_utilities_tests.c:1023-1100: a Bell state against |00> is 0.5 in
    every pairing of wavefunction and density matrix"""
    wf1, wf2 = _arb([4], WF), _arb([4], WF)
    wf1.add(2 ** -0.5, 0)
    wf1.add(2 ** -0.5, 3)
    wf2.add(1.0, 0)
    dm0 = _arb([4, 4], DM)
    dm0.add_loc(1.0, 0)
    dmb = _arb([4, 4], DM)
    for i in (0, 3):
        for j in (0, 3):
            dmb.add(0.5, i, j)
    assert abs(_fid(wf1, wf2) - 0.5) < 1e-12
    assert abs(_fid(wf1, dm0) - 0.5) < 1e-12
    assert abs(_fid(dmb, wf2) - 0.5) < 1e-12
    assert abs(_fid(dmb, dm0) - 0.5) < 1e-7
    _free(wf1, wf2, dm0, dmb)


@pytest.mark.parametrize("n", [6, 80])
def test_fidelity_superfidelity_hs_random_vs_scipy(n):
    """DM-DM fidelity (Tr sqrt(sqrt(rho) sigma sqrt(rho)))^2 against scipy's
    sqrtm and numpy's Hermitian eigenvalues (sigma of rank 2), DM-WF
    <psi|rho|psi>, WF-WF |<a|b>|^2, superfidelity and the Hilbert-Schmidt
    distance against their definitions; F(rho, rho) = 1.  n = 80 takes the
    device eigensolver (embedding of size 160)"""
    rng = np.random.default_rng(9)
    rho, sig = _rand_dm(rng, n), _rand_dm(rng, n, rank=2)
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    phi = rng.normal(size=n) + 1j * rng.normal(size=n)
    phi /= np.linalg.norm(phi)
    qr, qs = _arb([n, n], DM), _arb([n, n], DM)
    _set_dm(qr, rho)
    _set_dm(qs, sig)
    qp, qf = _arb([n], WF), _arb([n], WF)
    qp.set(psi)
    qf.set(phi)
    sr = scipy.linalg.sqrtm(rho)
    m = sr @ sig @ sr
    want = np.sum(np.sqrt(np.clip(np.linalg.eigvalsh(0.5 * (m + m.conj().T)), 0, None))) ** 2
    # sigma has rank 2, so n - 2 eigenvalues of sqrt(rho) sigma sqrt(rho) are zero up
    # to rounding (~eps |M|) and their square roots (~1e-9 each) enter the trace:
    # up to ~2n x 1e-9 either way
    tol = 1e-9 if n < 10 else 2 * n * 2e-9
    assert abs(_fid(qr, qs) - want) < tol, (_fid(qr, qs), want)
    assert abs(_fid(qr, qr) - 1.0) < 1e-9
    assert abs(_fid(qr, qp) - (psi.conj() @ rho @ psi).real) < 1e-13
    assert abs(_fid(qp, qf) - abs(np.vdot(phi, psi)) ** 2) < 1e-13
    sf, hs = ctypes.c_double(), ctypes.c_double()
    qa.lib().get_superfidelity_qvec(qr.h, qs.h, ctypes.byref(sf))
    qa.lib().get_hilbert_schmidt_dist_qvec(qr.h, qs.h, ctypes.byref(hs))
    qa._check()
    t12, t11, t22 = np.trace(rho @ sig).real, np.trace(rho @ rho).real, np.trace(sig @ sig).real
    assert abs(sf.value - np.sqrt(t12 + np.sqrt(1 - t11) * np.sqrt(1 - t22))) < 1e-12
    assert abs(hs.value - np.sum(np.abs(rho - sig) ** 2)) < 1e-13
    _free(qr, qs, qp, qf)


def test_copy_wf_to_dm_and_equality():
    """copy_qvec of a wavefunction into a density matrix is |psi><psi|;
    check_qvec_equal is exact equality"""
    rng = np.random.default_rng(2)
    psi = rng.normal(size=5) + 1j * rng.normal(size=5)
    wf, d1, d2 = _arb([5], WF), _arb([5, 5], DM), _arb([5, 5], DM)
    wf.set(psi)
    qa.lib().copy_qvec(wf.h, d1.h)
    qa.lib().copy_qvec_wf_to_dm(wf.h, d2.h)
    qa._check()
    assert np.abs(_rho(d1) - np.outer(psi, psi.conj())).max() < 1e-15
    flag = ctypes.c_int()
    qa.lib().check_qvec_equal(d1.h, d2.h, ctypes.byref(flag))
    assert flag.value == 1
    d2.add_loc(1e-9, 7)
    qa.lib().check_qvec_equal(d1.h, d2.h, ctypes.byref(flag))
    assert flag.value == 0
    _free(wf, d1, d2)


def test_xeb_scores_of_states_and_index_helpers():
    """get_linear_xeb_fidelity / get_hog_score on a wavefunction (ideal) and a
    density matrix (experiment) use their bitstring probabilities; the local
    index digits (qvec_utilities.c:1427-1438) and change_qvec_dims"""
    rng = np.random.default_rng(6)
    n = 8
    psi = rng.normal(size=n) + 1j * rng.normal(size=n)
    psi /= np.linalg.norm(psi)
    rho = _rand_dm(rng, n)
    ref, exp = _arb([2, 2, 2], WF), _arb([2, 2, 2, 2, 2, 2], DM)
    ref.set(psi)
    _set_dm(exp, rho)
    L = qa.lib()
    P = ctypes.POINTER(ctypes.c_double)
    for name in ("get_linear_xeb_fidelity", "get_hog_score"):
        getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, P, P]
        getattr(L, name).restype = None
    a, b = ctypes.c_double(), ctypes.c_double()
    p_ref, p_exp = np.abs(psi) ** 2, np.diag(rho).real
    L.get_linear_xeb_fidelity(ref.h, exp.h, ctypes.byref(a), ctypes.byref(b))
    assert abs(a.value - (n * np.sum(p_exp * p_ref) - 1)) < 1e-12 and b.value == 0.0
    L.get_hog_score(ref.h, exp.h, ctypes.byref(a), ctypes.byref(b))
    assert abs(a.value - p_exp[p_ref >= np.log(2) / n].sum()) < 1e-12
    # digits of a global index: the last dimension the least significant
    L.get_qvec_local_idxs.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.get_qvec_local_idxs.restype = None
    d = (ctypes.c_int64 * 6)()
    L.get_qvec_local_idxs(exp.h, 45, d)
    assert list(d) == [int(c) for c in np.binary_repr(45, 6)]
    # regroup the three qubits as (4, 2) and trace the 4-level part
    L.change_qvec_dims.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.change_qvec_dims.restype = None
    L.change_qvec_dims(exp.h, 4, (ctypes.c_int64 * 4)(4, 2, 4, 2))
    qa._check()
    red = _rho(_ptrace(exp, [0]))
    assert np.abs(red - np.einsum("aiAi->aA", rho.reshape(4, 2, 4, 2).transpose(1, 0, 3, 2))).max() < 1e-14
    _free(ref, exp)


def test_readout_error_on_a_density_matrix_vector():
    """apply_single_qb_measurement_error_state (quantum_circuits.c:646-688): the
    readout mixing applied to the entry pairs of the density matrix's vector at
    the stride of the dimensions listed before qubit_num"""
    rng = np.random.default_rng(12)
    rho = _rand_dm(rng, 4)
    q = _arb([2, 2, 2, 2], DM)
    _set_dm(q, rho)
    x = rho.T.reshape(-1)
    L = qa.lib()
    L.apply_single_qb_measurement_error_state.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                          ctypes.c_int64]
    L.apply_single_qb_measurement_error_state.restype = None
    L.apply_single_qb_measurement_error_state(q.h, 0.05, 0.1, 1)
    qa._check()
    want = x.copy()
    for i in range(8):
        i0 = (i // 2) * 4 + i % 2
        i1 = i0 + 2
        want[i0] = 0.95 * x[i0] + 0.1 * x[i1]
        want[i1] = 0.05 * x[i0] + 0.9 * x[i1]
    assert np.abs(q.get() - want).max() < 1e-15
    _free(q)


@pytest.mark.parametrize("kind", ["wf", "dm"])
def test_projective_measurement_of_a_bell_pair(kind):
    """apply_projective_measurement_qubit_list (quantum_circuits.c:484-582) on
    (|00> + |11>)/sqrt(2): sigma_z of qubit 0 gives +-1 with probability 1/2 and
    collapses the pair, so sigma_z of qubit 1 then agrees (product +1) and the
    state is |00><00| or |11><11|; a following sigma_x of qubit 0 gives +-1 with
    probability 1/2 and leaves the trace / norm 1 (generator seeded by quac_seed
    on its first use in the process)"""
    qa.set_option("quac_seed", "17")
    qs = qa.System()
    q0, q1 = qs.create_op(2), qs.create_op(2)
    qs.add_ham_term(1.0, q0.n)
    if kind == "dm":
        qs.add_lin_term(0.1, q0)
    qs.construct_matrix()
    L = qa.lib()
    f = L.apply_projective_measurement_qubit_list
    f.restype = None
    f.argtypes = [ctypes.c_void_p, ctypes.POINTER(qa.Cx), ctypes.c_int64, ctypes.c_void_p]

    def measure(x, ops):
        mv = qa.Cx()
        arr = qa._oparray(ops)
        f(x.h, ctypes.byref(mv), len(ops), ctypes.cast(arr, ctypes.c_void_p))
        qa._check()
        return mv.value()

    outcomes = []
    for trial in range(200):
        x = qs.create_qvec()
        if kind == "dm":
            for i in (0, 3):
                for j in (0, 3):
                    x.add(0.5, i, j)
        else:
            x.add(2 ** -0.5, 0)
            x.add(2 ** -0.5, 3)
        assert abs(measure(x, [q0.sig_z, q1.sig_z]) - 1.0) < 1e-12
        v = x.get()
        probs = (np.abs(v) ** 2 if kind == "wf" else _rho(x).diagonal().real)
        assert abs(probs.sum() - 1) < 1e-12 and (abs(probs[0] - 1) < 1e-12 or abs(probs[3] - 1) < 1e-12)
        outcomes.append(probs[0] > 0.5)
        assert abs(abs(measure(x, [q0.sig_x])) - 1.0) < 1e-12
        tot = np.sum(np.abs(x.get()) ** 2) if kind == "wf" else _rho(x).trace().real
        assert abs(tot - 1) < 1e-12
    # binomial(200, 1/2): 4 standard deviations
    assert abs(sum(outcomes) - 100) < 4 * math.sqrt(50)
    qs.destroy()
