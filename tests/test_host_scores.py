"""Cross-entropy benchmark scores of bitstring distributions (host functions of
libquac.so, no device needed): linear XEB, unbiased linear XEB, log XEB and the
heavy-output score (qvec_utilities.c:1720-1991), against their definitions
computed with numpy, plus the values that fix them: an ideal sampler of a
Porter-Thomas distribution scores linear XEB ~ 1, a uniform sampler 0.  Also
the two-qubit concurrence routine against Wootters' formula in numpy."""
import ctypes
import math

import numpy as np
import pytest

import quac_amd as qa

D = 1024


def _call(name, p_ref, p_exp, v_exp):
    L = qa.lib()
    f = getattr(L, name)
    P = ctypes.POINTER(ctypes.c_double)
    f.restype = None
    f.argtypes = [P, P, ctypes.c_int64, P, P, ctypes.c_int64, ctypes.c_int64, P, P]
    a, b = ctypes.c_double(), ctypes.c_double()
    arrs = [np.ascontiguousarray(x, np.float64) for x in (p_ref, np.zeros_like(p_ref), p_exp, v_exp)]
    f(*(arrs[0].ctypes.data_as(P), arrs[1].ctypes.data_as(P)), len(p_ref), arrs[2].ctypes.data_as(P),
      arrs[3].ctypes.data_as(P), len(p_exp), D, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


@pytest.fixture(scope="module")
def dists():
    rng = np.random.default_rng(4)
    amp = rng.normal(size=D) + 1j * rng.normal(size=D)
    p_ref = np.abs(amp) ** 2
    p_ref /= p_ref.sum()                      # Porter-Thomas-like ideal distribution
    p_ref[7] = 0.0                            # one bitstring the ideal circuit never yields
    p_ref /= p_ref.sum()
    mix = 0.6 * p_ref + 0.4 / D               # depolarised "experiment"
    v = rng.uniform(0, 1e-6, size=D)
    return p_ref, mix, v


def test_linear_xeb(dists):
    p_ref, mix, v = dists
    f, var = _call("get_linear_xeb_fidelity_probs", p_ref, mix, v)
    assert abs(f - (D * np.sum(mix * p_ref) - 1)) < 1e-12 and abs(var - D * D * np.sum(v * p_ref ** 2)) < 1e-12
    ideal, _ = _call("get_linear_xeb_fidelity_probs", p_ref, p_ref, v)
    uniform, _ = _call("get_linear_xeb_fidelity_probs", p_ref, np.full(D, 1.0 / D), v)
    assert abs(ideal - 1.0) < 0.2 and abs(uniform) < 1e-12   # Porter-Thomas: D sum p^2 ~ 2
    assert abs(f - 0.6 * ideal) < 1e-12                      # linear in the depolarising weight


def test_unbiased_linear_xeb(dists):
    p_ref, mix, v = dists
    f, var = _call("get_linear_xeb_unbiased_fidelity_probs", p_ref, mix, v)
    den = D * np.sum(p_ref ** 2) - 1
    assert abs(f - (D * np.sum(mix * p_ref) - 1) / den) < 1e-12
    assert abs(var - D * D * np.sum(v * p_ref ** 2) / den ** 2) < 1e-12
    assert abs(f - 0.6) < 1e-12                              # exact for a depolarised sampler
    assert abs(_call("get_linear_xeb_unbiased_fidelity_probs", p_ref, p_ref, v)[0] - 1.0) < 1e-12


def test_log_xeb(dists):
    p_ref, mix, v = dists
    f, var = _call("get_log_xeb_fidelity_probs", p_ref, mix, v)
    nz = p_ref > 0
    want = math.log(D) + np.sum(mix[nz] * np.log(p_ref[nz])) + np.euler_gamma
    wvar = np.sum(v[nz] * np.log(p_ref[nz]) ** 2) + v[~nz].sum() * math.log(3e-33) ** 2
    assert abs(f - want) < 1e-12 and abs(var - wvar) < 1e-9 * wvar


def test_hog(dists):
    p_ref, mix, v = dists
    h, hv = _call("get_hog_score_probs", p_ref, mix, v)
    heavy = p_ref >= math.log(2) / D
    assert abs(h - mix[heavy].sum()) < 1e-12 and abs(hv - v[heavy].sum()) < 1e-15
    f, fv = _call("get_hog_score_fidelity_probs", p_ref, mix, v)
    assert abs(f - (2 * h - 1) / math.log(2)) < 1e-12 and abs(fv - (2 / math.log(2)) ** 2 * hv) < 1e-15


def test_readout_error_mixing_vs_numpy():
    """apply_single_qb_measurement_error_probs (quantum_circuits.c:620-644):
    qubit q at stride 2^q; [p0', p1'] = [[1 - p01, p10], [p01, 1 - p10]] [p0, p1]
    on that qubit, numpy on the reshaped distribution; total probability kept"""
    L = qa.lib()
    f = L.apply_single_qb_measurement_error_probs
    f.restype = None
    f.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int64]
    rng = np.random.default_rng(1)
    p = rng.uniform(size=16)
    p /= p.sum()
    m = np.array([[1 - 0.03, 0.08], [0.03, 1 - 0.08]])
    for q in range(4):
        got = p.copy()
        f(got.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 16, 0.03, 0.08, q)
        t = np.moveaxis(p.reshape(2, 2, 2, 2), 3 - q, 0)       # qubit q = bit q of the index
        want = np.moveaxis(np.tensordot(m, t, axes=1), 0, 3 - q).reshape(-1)
        assert np.abs(got - want).max() < 1e-15 and abs(got.sum() - 1) < 1e-14, q


def _wootters(rho):
    """l1 - l2 - l3 - l4 from the eigenvalues of rho (sy sy) rho* (sy sy), each
    computed by numpy's general eigensolver (no Hermitian structure used)"""
    sy = np.array([[0, -1j], [1j, 0]])
    yy = np.kron(sy, sy)
    ev = np.linalg.eigvals(rho @ yy @ rho.conj() @ yy).real
    ev = np.where(ev < 1e-14, 0.0, ev)
    l = np.sort(np.sqrt(ev))[::-1]
    return l[0] - l[1:].sum()


@pytest.mark.parametrize("seed", range(6))
def test_concurrence_matches_wootters_on_random_states(seed):
    """quac_concurrence_2q (host routine behind get_bipartite_concurrence,
    dm_utilities.c:1487-1602) against Wootters' formula evaluated with numpy on
    random two-qubit density matrices: ranks 1 to 4, so entangled (positive) and
    separable (negative, unclipped like the reference) cases both occur; and the
    subsystem-swap invariance C(SWAP rho SWAP) = C(rho)"""
    f = qa.lib().quac_concurrence_2q
    f.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    rng = np.random.default_rng(seed)
    rank = 1 + seed % 4
    g = rng.normal(size=(4, rank)) + 1j * rng.normal(size=(4, rank))
    rho = g @ g.conj().T
    rho /= np.trace(rho).real
    swap = np.eye(4)[[0, 2, 1, 3]]
    for m in (rho, swap @ rho @ swap):
        buf = np.ascontiguousarray(m.flatten(order="F"))
        out = ctypes.c_double(-9.0)
        f(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(out))
        assert abs(out.value - _wootters(rho)) < 1e-7, (seed, out.value, _wootters(rho))
