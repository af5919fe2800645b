"""Time-dependent terms (qsystem.c:1288-1330): y = A x + sum_k s_k(t) A_k x per
RK stage in each of its forms -- the merged A_k entries summed in A's sliced
stage kernel, as a pass of their own (sliced form or CSR lanes), or one fused
CSR kernel; dictionary or per-entry values; one-workgroup runs of steps for
small generators; the adaptive RK3BS trial step (qh_spmv_td, qh_rk4_td_step,
qh_rk4_td_steps, qh_rk3bs_td_step) -- against the oracle's qo_rk4_td (pinned by
the analytic Rabi drive and by qo_rk4 on the summed matrix) and the closed form
of time_dep_tests.c:142-213: FP64 to 1e-10, the FP32 state to 1e-6 on
populations."""
import cmath
import ctypes
import math

import numpy as np
import pytest

import oracle
import quac_amd as qa
import recipes
from qbuild import populations_np

pytestmark = pytest.mark.gpu

F = [lambda t: math.cos(1.3 * t), lambda t: cmath.exp(0.4j * t) * math.sin(0.7 * t + 0.2),
     lambda t: math.exp(-0.2 * t)]


def _recipe(dm, nc=3):
    terms = [("ham", 0.4, [(1, "n")]), ("td_ham", 0.3 + 0.2j, [(0, "raise"), (1, "lower")]),
             ("td_ham", 0.3 - 0.2j, [(0, "lower"), (1, "raise")])]
    if dm:
        terms += [("td_lin", 0.05, [(1, "lower")]), ("lin", 0.25, [(0, "lower")])]
    return {"subs": [("op", 2), ("op", nc)], "terms": terms, "dm": dm}


def _product(recipe, prec):
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qa.set_option("quac_precision", prec)
    qs = qa.System()
    q = qs.create_op(2)
    c = qs.create_op(recipe["subs"][1][1])
    op = {(0, "raise"): q.dag, (0, "lower"): q, (1, "raise"): c.dag, (1, "lower"): c, (1, "n"): c.n}
    k = 0
    for (kind, a, ops) in recipe["terms"]:
        plist = [op[o] for o in ops]
        if kind == "ham":
            qs.add_ham_term(a, *plist)
        elif kind == "lin":
            qs.add_lin_term(a, *plist)
        elif kind == "td_ham":
            qs.add_ham_term_time_dep(a, F[k], *plist)
            k += 1
        else:
            qs.add_lin_term_time_dep(a, F[k], *plist)
            k += 1
    qs.construct_matrix()
    return qs, q, c


def _oracle(recipe, x0, t0, dt, nsteps):
    dm = recipe["dm"]
    tot, terms = recipes.oracle_terms(recipe)
    m = oracle.construct(tot, terms, dm=dm)
    idx = [k for k, t in enumerate(terms) if t[0] < 0]
    td = [oracle.construct(tot, terms, dm=dm, td_index=k) for k in idx]
    a = [terms[k][1] for k in idx]
    return m.rk4_td(td, a, lambda t, k: F[k](t), x0, t0, dt, nsteps, wf_scale=not dm), m, td, a


@pytest.mark.parametrize("dm", [True, False])
def test_td_rk4_fp64_vs_oracle(dm):
    recipe = _recipe(dm)
    qs, q, c = _product(recipe, "fp64")
    try:
        N = 6
        x = qs.create_qvec() if dm else qs.create_wf()
        rng = np.random.default_rng(3)
        if dm:
            a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
            rho = a @ a.conj().T
            x0 = (rho / np.trace(rho)).flatten(order="F")
        else:
            x0 = rng.normal(size=N) + 1j * rng.normal(size=N)
            x0 /= np.linalg.norm(x0)
        x.set(x0)
        steps, tend = qs.time_step(x, 0.3, 2.3, 0.01, 10 ** 6)
        assert steps == 200 and abs(tend - 2.3) < 1e-9
        ref, _, _, _ = _oracle(recipe, x0, 0.3, 0.01, 200)
        assert np.abs(x.get() - ref).max() < 1e-10
    finally:
        qs.destroy()


def test_td_rk4_fp32_vs_oracle():
    recipe = _recipe(True)
    qa.set_option("quac_renorm", "20")
    qs, q, c = _product(recipe, "fp32")
    try:
        x = qs.create_qvec()
        x.add_loc(1.0, 3 * 6 + 3)  # |1,0><1,0|
        qs.time_step(x, 0.0, 3.0, 0.01, 10 ** 6)
        x0 = np.zeros(36, np.complex128)
        x0[3 * 6 + 3] = 1.0
        ref, _, _, _ = _oracle(recipe, x0, 0.0, 0.01, 300)
        assert np.abs(populations_np(x.get(), recipe) - populations_np(ref, recipe)).max() < 1e-6
        assert abs(x.trace() - 1.0) < 1e-6
    finally:
        qa.set_option("quac_precision", "fp64")
        qa.set_option("quac_renorm", "0")
        qs.destroy()


@pytest.mark.parametrize("form", ["fused_dict", "fused_full", "sliced_dict", "sliced_dict_pass", "sliced_full",
                                  "sliced_dict_l4"])
def test_spmv_td_matches_dense_sum(form):
    """qh_spmv_td against A + sum_k s_k A_k from the oracle's matrices, each form
    of the RHS (A forced onto sliced variant 134 for the sliced forms); a
    coefficient of magnitude <= 1e-10 is skipped (qsystem.c:34,1320)"""
    recipe = _recipe(True)
    qa.td_config(*FORMS[form])
    qs, q, c = _product(recipe, "fp64")
    try:
        if FORMS[form][1]:
            qs.matrix().set_lanes(134, 0)
        tot, terms = recipes.oracle_terms(recipe)
        m = oracle.construct(tot, terms, dm=True)
        idx = [k for k, t in enumerate(terms) if t[0] < 0]
        tds = [oracle.construct(tot, terms, dm=True, td_index=k) for k in idx]

        def dense(csr):
            rp, col, val = csr.export()
            D = np.zeros((csr.n, csr.n), np.complex128)
            for r in range(csr.n):
                D[r, col[rp[r]:rp[r + 1]]] += val[rp[r]:rp[r + 1]]
            return D

        A = dense(m)
        n = A.shape[0]
        rng = np.random.default_rng(8)
        xh = rng.normal(size=n) + 1j * rng.normal(size=n)
        s = np.array([0.7 - 0.3j, 1e-11, -1.2 + 0.5j])
        want = A @ xh + sum(s[k] * (dense(tds[k]) @ xh) for k in (0, 2))
        mats = (ctypes.c_void_p * 3)(*[qa.lib().quac_system_td_matrix(qs.h, k) for k in range(3)])
        sc = np.ascontiguousarray(np.column_stack([s.real, s.imag]).ravel())
        xv = qa.DeviceVec(n)
        xv.set(xh)
        yv = qa.DeviceVec(n)
        L = qa.lib()
        L.qh_spmv_td.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_void_p]
        L.qh_spmv_td.restype = ctypes.c_int
        qa._qh(L.qh_spmv_td(qs.matrix().h, 3, mats, sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), xv.h, yv.h))
        assert np.abs(yv.get() - want).max() < 1e-12 * np.abs(want).max()
    finally:
        qa.td_config()
        qs.destroy()


# the forms of the RHS: (dictionary limit, A's part on a sliced layout, lanes of the A_k pass)
# (lanes 0: the sliced 64-row form of the dictionary entries summed inside A's
# stage kernel for the one-row variants 128, 134, 136, else in a pass of its own;
# lanes 1: always that pass; a limit of 3 is below these systems' distinct
# (term, value) pairs, so the build falls back to per-entry values)
FORMS = {"sliced_dict": (4096, 1, 0), "sliced_dict_pass": (4096, 1, 1), "sliced_full": (0, 1, 0),
         "sliced_dict_l4": (4096, 1, 4),
         "sliced_full_l16": (0, 1, 16), "sliced_dict_l8": (4096, 1, 8), "sliced_over": (3, 1, 1),
         "fused_dict": (4096, 0, 0), "fused_full": (0, 0, 0)}
# forced sliced variants of A (FP64 / FP32): one-row layouts 128, 134, 136; row-pair 142, 143
VARIANTS = {"fp64": [134, 128, 136], "fp32": [142, 143, 134]}


def _forms_case(prec, form, variant, nc, steps):
    """a qubit and an nc-level cavity (density matrix), stepped under one form of
    the time-dependent RHS, against the oracle from the same random state"""
    qa.td_config(*FORMS[form])
    recipe = _recipe(True, nc)
    qs, q, c = _product(recipe, prec)
    try:
        if variant:
            qs.matrix().set_lanes(variant if prec == "fp64" else 0, variant if prec == "fp32" else 0)
        N = 2 * nc
        rng = np.random.default_rng(nc + steps)
        a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
        rho = a @ a.conj().T
        x0 = (rho / np.trace(rho)).flatten(order="F")
        x = qs.create_qvec()
        x.set(x0)
        print(f"td forms: {prec} {form} variant {variant} rows {N * N} steps {steps}", flush=True)
        qs.time_step(x, 0.1, 0.1 + 0.01 * steps, 0.01, 10 ** 6)
        ref, _, _, _ = _oracle(recipe, x0, 0.1, 0.01, steps)
        got = x.get()
        if prec == "fp64":
            assert np.abs(got - ref).max() < 1e-10
        else:
            assert np.abs(populations_np(got, recipe) - populations_np(ref, recipe)).max() < 1e-6
            assert np.abs(got - ref).max() < 2e-6
    finally:
        qa.td_config()
        qa.set_option("quac_precision", "fp64")
        qs.destroy()


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_td_rk4_forms_vs_oracle(prec, form):
    """every form of the RHS on a qubit and a 16-level cavity (1 024 rows, A on
    its first forced sliced variant where the form runs A sliced), 150 RK4 steps"""
    variant = VARIANTS[prec][0] if FORMS[form][1] else 0
    _forms_case(prec, form, variant, 16, 150)


@pytest.mark.parametrize("prec,variant", [(p, v) for p in VARIANTS for v in VARIANTS[p][1:]])
def test_td_rk4_sliced_variants_vs_oracle(prec, variant):
    """A's part on the other sliced variants: 12-level cavity (576 rows, sliced
    only because forced), 100 steps"""
    _forms_case(prec, "sliced_dict", variant, 12, 100)


def test_td_rk4_sliced_autotuned_vs_fused():
    """a 24-level cavity (2 304 rows): the autotuned variant of A (sliced or CSR,
    whichever the tuner picks) against the fused CSR kernel, FP64 and FP32, with
    the same coefficients; FP64 to 1e-12 relative"""
    recipe = _recipe(True, 24)
    out = {}
    for prec in ("fp64", "fp32"):
        for sliced in (1, 0):
            qa.td_config(4096, sliced, 0)
            qs, q, c = _product(recipe, prec)
            try:
                x = qs.create_qvec()
                x.add_loc(1.0, 24 * 48 + 24)  # |1,0><1,0|
                qs.time_step(x, 0.0, 0.5, 0.005, 10 ** 6)
                out[(prec, sliced)] = x.get()
            finally:
                qa.td_config()
                qa.set_option("quac_precision", "fp64")
                qs.destroy()
    a, b = out[("fp64", 1)], out[("fp64", 0)]
    assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
    a, b = out[("fp32", 1)], out[("fp32", 0)]
    assert np.abs(populations_np(a, recipe) - populations_np(b, recipe)).max() < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_td_rk4_one_workgroup_run_matches_stepwise_and_oracle(prec):
    """a qubit and a 12-level cavity (576 rows, under the 800-row limit): the
    fixed RK4 steps of time_step_sys run in one workgroup per chunk of steps
    (qh_rk4_td_steps, stage tables of every step computed ahead) against the
    same steps launched one at a time (per-entry values force that path) and
    against the oracle; 300 steps, more than one chunk of 256"""
    recipe = _recipe(True, 12)
    N = 24
    rng = np.random.default_rng(33)
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rho = a @ a.conj().T
    x0 = (rho / np.trace(rho)).flatten(order="F")
    out = {}
    for form, cfg in (("run", (4096, 1, 0)), ("stepwise", (0, 1, 0))):
        qa.td_config(*cfg)
        qs, q, c = _product(recipe, prec)
        try:
            x = qs.create_qvec()
            x.set(x0)
            print(f"td small run: {prec} {form}", flush=True)
            steps, tend = qs.time_step(x, 0.1, 3.1, 0.01, 10 ** 6)
            assert steps == 300 and abs(tend - 3.1) < 1e-9
            out[form] = x.get()
        finally:
            qa.td_config()
            qa.set_option("quac_precision", "fp64")
            qs.destroy()
    ref, _, _, _ = _oracle(recipe, x0, 0.1, 0.01, 300)
    if prec == "fp64":
        assert np.abs(out["run"] - out["stepwise"]).max() < 1e-13
        assert np.abs(out["run"] - ref).max() < 1e-10
    else:
        assert np.abs(populations_np(out["run"], recipe) - populations_np(ref, recipe)).max() < 1e-6
        assert np.abs(populations_np(out["stepwise"], recipe) - populations_np(ref, recipe)).max() < 1e-6


def test_td_rk3bs_forms_agree_and_match_fine_rk4():
    """the default integrator (adaptive RK3BS) with time-dependent terms on a
    qubit and a 16-level cavity (1 024 rows, A forced onto variant 134): the trial
    steps with the entries summed in A's four stage kernels, with the entries in
    a pass of their own after each sliced SpMV, and with the fused CSR SpMV take
    the same steps to the same state (1e-12 relative); the result matches the
    oracle's RK4 at dt = 0.002 within the integrator's tolerance"""
    recipe = _recipe(True, 16)
    N = 32
    rng = np.random.default_rng(21)
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    rho = a @ a.conj().T
    x0 = (rho / np.trace(rho)).flatten(order="F")
    res = {}
    for form in ("sliced_dict", "sliced_dict_pass", "fused_dict"):
        qa.td_config(*FORMS[form])
        qs, q, c = _product(recipe, "fp64")
        qa.set_option("ts_rk_type", "3bs")
        qa.set_option("ts_adapt_type", "basic")
        try:
            qs.matrix().set_lanes(134, 0)
            x = qs.create_qvec()
            x.set(x0)
            print(f"td rk3bs: {form}", flush=True)
            steps, tend = qs.time_step(x, 0.1, 1.6, 0.01, 10 ** 6)
            res[form] = (steps, x.get())
        finally:
            qa.td_config()
            qa.set_option("ts_rk_type", "3bs")
            qa.set_option("ts_adapt_type", "basic")
            qs.destroy()
    s0, x0f = res["sliced_dict"]
    for form in ("sliced_dict_pass", "fused_dict"):
        s1, x1 = res[form]
        assert s1 == s0
        assert np.abs(x1 - x0f).max() < 1e-12 * np.abs(x0f).max()
    ref, _, _, _ = _oracle(recipe, x0, 0.1, 0.002, 750)
    assert np.abs(populations_np(x0f, recipe) - populations_np(ref, recipe)).max() < 1e-5


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_td_rabi_wavefunction_closed_form(prec):
    """time_dep_tests.c:142-213: H(t) = 0.5 omega e^{-alpha t} (a + a^dag) on a
    wavefunction from |0>, omega = 2 pi, alpha = 0.1, RK4 dt = 0.01 to t = 20:
    <sz> = cos(omega (1 - e^{-alpha t})/alpha) (RK4's own error at dt = 0.01 is
    below 1e-7; the reference allows 1e-4)"""
    om, al = 2 * math.pi, 0.1
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qa.set_option("quac_precision", prec)
    qs = qa.System()
    try:
        o = qs.create_op(2)
        env = lambda t: math.exp(-al * t)
        qs.add_ham_term_time_dep(0.5 * om, env, o)
        qs.add_ham_term_time_dep(0.5 * om, env, o.dag)
        qs.construct_matrix()
        x = qs.create_wf()
        x.add_loc(1.0, 0)
        qs.time_step(x, 0.0, 20.0, 0.01, 10 ** 6)
        psi = x.get()
        th = om * (1 - math.exp(-al * 20.0)) / al
        sz = abs(psi[0]) ** 2 - abs(psi[1]) ** 2
        sy = 2 * (psi[0].conjugate() * psi[1]).imag
        tol = 2e-7 if prec == "fp64" else 1e-5
        assert abs(sz - math.cos(th)) < tol and abs(sy + math.sin(th)) < tol, (sz, sy)
    finally:
        qa.set_option("quac_precision", "fp64")
        qs.destroy()
