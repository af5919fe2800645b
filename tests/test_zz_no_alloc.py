"""qh_expect_dm_noalloc (Tr(O rho) with no device allocation, its op chain and
partial sums in a fixed device array) against qh_expect_dm through the C-ABI:
bit-identical results on two calls in a row, the second on a smaller system with
a shorter chain, on FP64 and FP32 (hi+lo) states; both checked against a dense
Tr(M rho).  Sizes beyond the fixed array are refused with QH_ERANGE."""
import ctypes
import math

import numpy as np
import pytest

import quac_amd as qa

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(11)


def _rand_rho(N):
    a = RNG.normal(size=(N, N)) + 1j * RNG.normal(size=(N, N))
    rho = a @ a.conj().T
    return rho / np.trace(rho)


def _chain(levels, spec):
    """qh_op / qh_level tables for a product of ladder operators; spec = list of
    (subsystem, kind) applied in order, kind in {"lower", "raise", "number"}.
    Returns the tables and the dense matrix M with Tr(M rho) equal to what the
    device sums (M[i, chain(i)] = value)."""
    tot = int(np.prod(levels))
    strides = [int(np.prod(levels[s + 1:])) for s in range(len(levels))]
    ops, tab = [], []
    for s, kind in spec:
        L = levels[s]
        start = len(tab)
        for l in range(L):
            if kind == "lower":
                nx, v = (l - 1, math.sqrt(l)) if l >= 1 else (-1, 0.0)
            elif kind == "raise":
                nx, v = (l + 1, math.sqrt(l + 1)) if l + 1 < L else (-1, 0.0)
            else:
                nx, v = (l, float(l)) if l > 0 else (-1, 0.0)
            tab.append((nx, v))
        ops.append((strides[s], L, start))
    prev = [-1] * len(tab)
    for (st, L, t0) in ops:
        for l in range(L):
            nx = tab[t0 + l][0]
            if nx >= 0:
                prev[t0 + nx] = l
    M = np.zeros((tot, tot), np.complex128)
    for i in range(tot):
        cur, val, ok = i, 1.0 + 0j, True
        for (st, L, t0) in ops:
            sub = (cur // st) % L
            nx, v = tab[t0 + sub]
            if nx < 0:
                ok = False
                break
            cur += (nx - sub) * st
            val *= v
        if ok:
            M[i, cur] = val
    c_ops = (qa.QhOp * len(ops))(*[qa.QhOp(st, L, t0) for (st, L, t0) in ops])
    c_lev = (qa.QhLevel * len(tab))(*[qa.QhLevel(nx, prev[k], v, 0.0) for k, (nx, v) in enumerate(tab)])
    return tot, c_ops, len(ops), c_lev, len(tab), M


def _expect(fn, vec, tot, c_ops, nops, c_lev, nlev):
    re, im = ctypes.c_double(), ctypes.c_double()
    rc = fn(vec.h, tot, nops, c_ops, nlev, c_lev, ctypes.byref(re), ctypes.byref(im))
    if rc != 0:
        raise qa.QuacError(qa.lib().qh_last_error().decode())
    return complex(re.value, im.value)


CASES = [([2, 3, 4, 5], [(1, "lower"), (3, "raise"), (0, "number"), (2, "raise")]),
         ([2, 3], [(1, "number")])]


@pytest.mark.parametrize("prec", [qa.FP64, qa.FP32])
def test_noalloc_matches_qh_expect_dm_twice(prec):
    L = qa.lib()
    for levels, spec in CASES:  # second call: smaller system, shorter chain
        tot, c_ops, nops, c_lev, nlev, M = _chain(levels, spec)
        rho = _rand_rho(tot)
        v = qa.DeviceVec(tot * tot, prec)
        v.set(rho.T.ravel())  # vec(rho) column-major: loc = N*col + row
        ref = _expect(L.qh_expect_dm, v, tot, c_ops, nops, c_lev, nlev)
        got = _expect(L.qh_expect_dm_noalloc, v, tot, c_ops, nops, c_lev, nlev)
        print(f"noalloc prec={prec} levels={levels} tot={tot} nops={nops}: {got} vs {ref}", flush=True)
        assert got.real == ref.real and got.imag == ref.imag, (got, ref)
        dense = np.trace(M @ rho)
        tol = 1e-12 if prec == qa.FP64 else 1e-6
        assert abs(got - dense) <= tol * max(1.0, abs(dense)), (got, dense)


def test_noalloc_refuses_sizes_beyond_its_array():
    L = qa.lib()
    tot, c_ops, nops, c_lev, nlev, _ = _chain([2, 2], [(0, "lower")])
    v = qa.DeviceVec(tot * tot, qa.FP64)
    many = (qa.QhOp * 65)(*([c_ops[0]] * 65))
    re, im = ctypes.c_double(), ctypes.c_double()
    rc = L.qh_expect_dm_noalloc(v.h, tot, 65, many, nlev, c_lev, ctypes.byref(re), ctypes.byref(im))
    assert rc == -4 and b"at most 64" in L.qh_last_error()
    big = (qa.QhLevel * 4097)()
    rc = L.qh_expect_dm_noalloc(v.h, tot, nops, c_ops, 4097, big, ctypes.byref(re), ctypes.byref(im))
    assert rc == -4 and b"at most 4096" in L.qh_last_error()
