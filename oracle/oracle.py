"""ctypes wrapper around the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package quac_amd/.

A term is ``(kind, a, ops)`` with kind in {HAM=1, LIN=2, TD_HAM=-1, TD_LIN=-2},
``a`` a Python complex and ``ops`` a list of ``(type, levels, n_before, position)``
tuples (the fields of ``struct operator``, src/operators.h:8-31).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

HAM, LIN, TD_HAM, TD_LIN = 1, 2, -1, -2
IG, GG, GI = -1, 0, 1

_lib = None

# qo_tdfunc: f(t, k, &re, &im), the envelope of time-dependent term k
TDFUNC = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double))


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        i64, dbl = ctypes.c_int64, ctypes.c_double
        L.qo_construct.argtypes = [i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   P(i64), P(i64), P(i64), P(dbl), P(i64), ctypes.c_int,
                                   P(ctypes.c_void_p)]
        L.qo_csr_rows.argtypes = [ctypes.c_void_p]
        L.qo_csr_rows.restype = i64
        L.qo_csr_nnz.argtypes = [ctypes.c_void_p]
        L.qo_csr_nnz.restype = i64
        L.qo_csr_export.argtypes = [ctypes.c_void_p, P(i64), P(i64), P(dbl)]
        L.qo_csr_free.argtypes = [ctypes.c_void_p]
        L.qo_spmv.argtypes = [ctypes.c_void_p, P(dbl), P(dbl)]
        L.qo_rk4.argtypes = [ctypes.c_void_p, P(dbl), dbl, i64, ctypes.c_int]
        L.qo_rk3bs.argtypes = [ctypes.c_void_p, P(dbl), dbl, dbl, dbl, i64, dbl, dbl, dbl, dbl,
                               P(i64), P(dbl)]
        L.qo_rk3bs_trial.argtypes = [ctypes.c_void_p, P(dbl), dbl, ctypes.c_int, P(dbl), P(dbl), P(dbl),
                                     P(dbl), P(dbl), dbl, dbl, P(dbl)]
        L.qo_rk4_td.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_void_p), P(dbl), ctypes.c_int,
                                TDFUNC, P(dbl), dbl, dbl, i64]
        L.qo_expect_dm.argtypes = [i64, P(dbl), ctypes.c_int, P(i64), P(dbl), P(dbl)]
        L.qo_expect_wf.argtypes = [i64, P(dbl), ctypes.c_int, P(i64), P(dbl), P(dbl)]
        L.qo_chain.argtypes = [i64, i64, ctypes.c_int, P(i64), ctypes.c_int, P(dbl), P(dbl), P(i64)]
        L.qo_gate_superop.argtypes = [ctypes.c_int, P(i64), P(i64), ctypes.c_int, ctypes.c_int,
                                      P(i64), dbl, dbl, dbl, P(dbl)]
        L.qo_gate_inplace.argtypes = [ctypes.c_int, P(i64), ctypes.c_int, ctypes.c_int, P(i64),
                                      dbl, dbl, dbl, ctypes.c_int, P(dbl)]
        L.qo_gmres.argtypes = [ctypes.c_void_p, i64, P(dbl), ctypes.c_int, dbl, i64, ctypes.c_int,
                               P(i64), P(dbl)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _pack_terms(terms):
    kind, nops, off, a, ops = [], [], [], [], []
    for (k, coef, oplist) in terms:
        kind.append(k)
        nops.append(len(oplist))
        off.append(len(ops))
        coef = complex(coef)
        a.extend([coef.real, coef.imag])
        ops.extend(oplist)
    return (np.array(kind, np.int64), np.array(nops, np.int64), np.array(off, np.int64),
            np.array(a, np.float64), np.array(ops, np.int64).reshape(-1, 4).copy())


class CSR:
    """Owned oracle CSR matrix (rows, nnz, rowptr, col, complex val)."""

    def __init__(self, handle):
        self.h = handle
        L = lib()
        self.n = L.qo_csr_rows(handle)
        self.nnz = L.qo_csr_nnz(handle)

    def export(self):
        L = lib()
        rp = np.empty(self.n + 1, np.int64)
        col = np.empty(self.nnz, np.int64)
        val = np.empty(2 * self.nnz, np.float64)
        L.qo_csr_export(self.h, _p(rp, ctypes.c_int64), _p(col, ctypes.c_int64),
                        _p(val, ctypes.c_double))
        return rp, col, val.view(np.complex128)

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty_like(x)
        lib().qo_spmv(self.h, _p(x.view(np.float64), ctypes.c_double), _p(y.view(np.float64), ctypes.c_double))
        return y

    def rk4(self, x, dt, nsteps, nthreads=1):
        x = np.array(x, np.complex128, copy=True)
        rc = lib().qo_rk4(self.h, _p(x.view(np.float64), ctypes.c_double), dt, nsteps, nthreads)
        assert rc == 0
        return x

    def rk3bs(self, x, t0, tmax, dt0, steps_max, atol=5e-7, rtol=5e-7, hmin=0.01, hmax=1.0):
        x = np.array(x, np.complex128, copy=True)
        st = ctypes.c_int64(0)
        te = ctypes.c_double(0)
        lib().qo_rk3bs(self.h, _p(x.view(np.float64), ctypes.c_double), t0, tmax, dt0, steps_max,
                       atol, rtol, hmin, hmax, ctypes.byref(st), ctypes.byref(te))
        return x, st.value, te.value

    def rk3bs_trial(self, x, h, atol=5e-7, rtol=5e-7, k1=None):
        """one Bogacki-Shampine trial step: (xn, err, [k1, k2, k3, k4]); k1 given = FSAL"""
        x = np.ascontiguousarray(x, np.complex128)
        ks = [np.array(k1, np.complex128, copy=True) if (i == 0 and k1 is not None) else np.empty_like(x)
              for i in range(4)]
        xn = np.empty_like(x)
        err = ctypes.c_double()
        d = ctypes.c_double
        rc = lib().qo_rk3bs_trial(self.h, _p(x.view(np.float64), d), h, int(k1 is not None),
                                  *[_p(k.view(np.float64), d) for k in ks], _p(xn.view(np.float64), d),
                                  atol, rtol, ctypes.byref(err))
        assert rc == 0
        return xn, err.value, ks

    def rk4_td(self, td, td_a, f, x, t0, dt, nsteps, wf_scale=False):
        """classic RK4 of dx/dt = A x + sum_k a_k f_k(t) A_k x (x -i when wf_scale);
        td: list of CSR (the A_k), td_a: complex a_k, f(t, k) -> complex"""
        x = np.array(x, np.complex128, copy=True)

        def cb(t, k, re, im):
            v = complex(f(t, k))
            re[0] = v.real
            im[0] = v.imag

        cfun = TDFUNC(cb)
        hs = (ctypes.c_void_p * max(1, len(td)))(*[m.h for m in td])
        a = np.array([[complex(c).real, complex(c).imag] for c in td_a] or [[0.0, 0.0]], np.float64).ravel()
        rc = lib().qo_rk4_td(self.h, len(td), hs, _p(a, ctypes.c_double), int(wf_scale), cfun,
                             _p(x.view(np.float64), ctypes.c_double), t0, dt, nsteps)
        assert rc == 0
        return x

    def gmres(self, tot, x0=None, restart=100, tol=1e-8, maxit=100000, jacobi=False):
        x = np.zeros(self.n, np.complex128) if x0 is None else np.array(x0, np.complex128, copy=True)
        it = ctypes.c_int64(0)
        res = ctypes.c_double(0)
        lib().qo_gmres(self.h, tot, _p(x.view(np.float64), ctypes.c_double), restart, tol, maxit,
                       int(jacobi), ctypes.byref(it), ctypes.byref(res))
        return x, it.value, res.value

    def __del__(self):
        try:
            if self.h:
                lib().qo_csr_free(self.h)
                self.h = None
        except Exception:
            pass


def construct(tot, terms, dm, mcwf=False, conj_a=False, td_index=-1):
    """construct_matrix (qsystem.c:593-676) restated; returns CSR."""
    kind, nops, off, a, ops = _pack_terms(terms)
    h = ctypes.c_void_p()
    rc = lib().qo_construct(tot, int(dm), int(mcwf), int(conj_a), len(terms),
                            _p(kind, ctypes.c_int64), _p(nops, ctypes.c_int64), _p(off, ctypes.c_int64),
                            _p(a, ctypes.c_double), _p(ops, ctypes.c_int64), td_index, ctypes.byref(h))
    if rc != 0:
        raise RuntimeError("qo_construct failed: %d" % rc)
    return CSR(h)


def chain(tot, i, ops, tc):
    o = np.array(ops, np.int64).reshape(-1, 4).copy()
    re, im, j = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
    rc = lib().qo_chain(tot, i, len(ops), _p(o, ctypes.c_int64), tc, ctypes.byref(re), ctypes.byref(im),
                        ctypes.byref(j))
    if rc:
        raise RuntimeError("qo_chain failed %d" % rc)
    return complex(re.value, im.value), j.value


def expect_dm(tot, x, ops):
    x = np.ascontiguousarray(x, np.complex128)
    o = np.array(ops, np.int64).reshape(-1, 4).copy()
    re, im = ctypes.c_double(), ctypes.c_double()
    rc = lib().qo_expect_dm(tot, _p(x.view(np.float64), ctypes.c_double), len(ops), _p(o, ctypes.c_int64),
                            ctypes.byref(re), ctypes.byref(im))
    assert rc == 0
    return complex(re.value, im.value)


def expect_wf(tot, x, ops):
    x = np.ascontiguousarray(x, np.complex128)
    o = np.array(ops, np.int64).reshape(-1, 4).copy()
    re, im = ctypes.c_double(), ctypes.c_double()
    rc = lib().qo_expect_wf(tot, _p(x.view(np.float64), ctypes.c_double), len(ops), _p(o, ctypes.c_int64),
                            ctypes.byref(re), ctypes.byref(im))
    assert rc == 0
    return complex(re.value, im.value)


def gate_superop(levels, nbefore, dm, gate, qubits, x, theta=0.0, phi=0.0, lam=0.0):
    x = np.array(x, np.complex128, copy=True)
    lv = np.array(levels, np.int64)
    nb = np.array(nbefore, np.int64)
    q = np.array(qubits, np.int64)
    rc = lib().qo_gate_superop(len(levels), _p(lv, ctypes.c_int64), _p(nb, ctypes.c_int64), int(dm), gate,
                               _p(q, ctypes.c_int64), theta, phi, lam, _p(x.view(np.float64), ctypes.c_double))
    assert rc == 0, rc
    return x


def gate_inplace(dims, dm, gate, qubits, x, theta=0.0, phi=0.0, lam=0.0, rz_conj_bug=False):
    x = np.array(x, np.complex128, copy=True)
    d = np.array(dims, np.int64)
    q = np.array(qubits, np.int64)
    rc = lib().qo_gate_inplace(len(dims), _p(d, ctypes.c_int64), int(dm), gate, _p(q, ctypes.c_int64),
                               theta, phi, lam, int(rz_conj_bug), _p(x.view(np.float64), ctypes.c_double))
    assert rc == 0, rc
    return x
