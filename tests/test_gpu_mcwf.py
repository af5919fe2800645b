"""Monte Carlo wavefunction trajectories (use_mcwf_solver) on the device.

Parity is statistical (the reference draws from SPRNG, which is not available):
ensemble averages must agree with the density-matrix solution within 4 standard
errors of the mean.  Jumps happen at the norm crossing located inside the step
(as the reference's event handler locates it), so there is no jump-timing bias
to allow for; the crossing itself is checked against a numpy root of the RK4
step's norm polynomial."""
import ctypes
import math

import numpy as np
import pytest

import oracle
import quac_amd as qa
import recipes
from qbuild import populations_np

pytestmark = pytest.mark.gpu


def _fixed_rk4():
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")


@pytest.mark.parametrize("levels", [2, 70])
def test_t1_decay_ensemble(levels):
    """time_step_tests.c:24-101 under MCWF: <n>(t) = exp(-gamma t) from one
    excitation (a qubit: several trajectories per wavefront; a 70-level mode:
    one workgroup per trajectory)"""
    _fixed_rk4()
    M, gamma, dt, T = 8000, 1.0, 0.01, 1.0
    qs = qa.System()
    o = qs.create_op(levels)
    qs.add_lin_term(gamma, o)
    qs.use_mcwf_solver(M, 7)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 1)
    qs.time_step(x, 0.0, T, dt, 100000)
    p = math.exp(-gamma * T)
    got = x.expectation(o.n).real
    sem = math.sqrt(p * (1 - p) / M)
    assert abs(got - p) < 4 * sem, (got, p, sem)
    qs.destroy()


@pytest.mark.parametrize("path", ["device_loop", "stepwise"])
def test_driven_jc_ensemble_vs_density_matrix(path, monkeypatch):
    """emitter (2) x cavity (4), exchange coupling and both losses; averages of
    M trajectories against the FP64 density-matrix RK4 of the same system, on the
    device-resident trajectory loop and on the step-by-step loop"""
    _fixed_rk4()
    if path == "stepwise":
        monkeypatch.setenv("QUAC_MCWF_STEPWISE", "1")
    M, dt, T = 4000, 0.01, 3.0
    g, wq, wc, kap, gam = 0.4, 1.0, 1.0, 0.3, 0.2
    qs = qa.System()
    q = qs.create_op(2)
    c = qs.create_op(4)
    qs.add_ham_term(wq, q.n)
    qs.add_ham_term(wc, c.n)
    qs.add_ham_term(g, q.dag, c)
    qs.add_ham_term(g, q, c.dag)
    qs.add_lin_term(gam, q)
    qs.add_lin_term(kap, c)
    qs.use_mcwf_solver(M, 11)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 4)          # emitter excited (index 1 * 4 + 0), cavity vacuum
    qs.time_step(x, 0.0, T, dt, 100000)
    nq = x.expectation(q.n).real
    nc = x.expectation(c.n).real
    recipe = {"subs": [("op", 2), ("op", 4)],
              "terms": [("ham", wq, [(0, "n")]), ("ham", wc, [(1, "n")]),
                        ("ham", g, [(0, "raise"), (1, "lower")]), ("ham", g, [(0, "lower"), (1, "raise")]),
                        ("lin", gam, [(0, "lower")]), ("lin", kap, [(1, "lower")])], "mcwf": False}
    tot, terms = recipes.oracle_terms(recipe)
    m = oracle.construct(tot, terms, dm=True)
    x0 = np.zeros(tot * tot, np.complex128)
    x0[4 * tot + 4] = 1.0
    ref = populations_np(m.rk4(x0, dt, int(round(T / dt)), nthreads=4), recipe)
    # per-trajectory values lie in [0, 1] (emitter) and [0, 3] (cavity): SEM <= range / (2 sqrt(M))
    tol_q = 4 * 0.5 / math.sqrt(M)
    tol_c = 4 * 1.5 / math.sqrt(M)
    assert abs(nq - ref[0]) < tol_q, (nq, ref[0], tol_q)
    assert abs(nc - ref[1]) < tol_c, (nc, ref[1], tol_c)
    qs.destroy()


def test_monitor_sees_the_ensemble_and_norms_decay():
    _fixed_rk4()
    M = 64
    qs = qa.System()
    o = qs.create_op(3)
    qs.add_ham_term(0.5, o.n)
    qs.add_lin_term(0.5, o)
    qs.use_mcwf_solver(M, 3)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 2)
    steps = []
    qs.set_monitor(lambda step, t, v: steps.append(step))
    qs.time_step(x, 0.0, 0.2, 0.01, 1000)
    qs.set_monitor(None)
    assert steps == list(range(21))
    e = x.expectation(o.n)
    assert 0.0 <= e.real <= 2.0 and abs(e.imag) < 1e-12
    qs.destroy()


def _jc_generator(gam=1.0, kap=0.8, levels=4):
    """-i Heff of an emitter (2) x cavity (levels) under MCWF (N = 2 levels)"""
    qs = qa.System()
    q = qs.create_op(2)
    c = qs.create_op(levels)
    qs.add_ham_term(1.0, q.n)
    qs.add_ham_term(0.9, c.n)
    qs.add_ham_term(0.4, q.dag, c)
    qs.add_ham_term(0.4, q, c.dag)
    qs.add_lin_term(gam, q)
    qs.add_lin_term(kap, c)
    qs.use_mcwf_solver(4, 1)
    qs.construct_matrix()
    return qs, qs.matrix()


def _qubit_ring_generator(nq):
    """-i Heff of nq qubits, sigma_x on each and sigma_z sigma_z neighbours, decay
    of each: every row holds the diagonal and nq sigma_x entries"""
    qs = qa.System()
    q = [qs.create_op(2) for _ in range(nq)]
    for k in range(nq):
        qs.add_ham_term(0.3 + 0.1 * k, q[k].sig_x)
        qs.add_ham_term(0.2, q[k].sig_z, q[(k + 1) % nq].sig_z)
        qs.add_lin_term(0.6, q[k])
    qs.use_mcwf_solver(4, 1)
    qs.construct_matrix()
    return qs, qs.matrix()


def _dense(A):
    rp, col, val = A.export()
    n = len(rp) - 1
    d = np.zeros((n, n), np.complex128)
    for r in range(n):
        d[r, col[rp[r]:rp[r + 1]]] += val[rp[r]:rp[r + 1]]
    return d


def _P(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _cross_np(Ad, psi0, h, thr):
    """the norm crossing inside an RK4 step of size h from psi0 (numpy): the step
    is P(tau) = sum_k tau^k A^k psi0 / k!, ||P(tau)||^2 a degree-8 polynomial;
    bisection for its first fall to thr on (0, h]"""
    v = [psi0]
    for k in range(1, 5):
        v.append(Ad @ v[-1] / k)
    c = np.zeros(9)
    for j in range(5):
        for k in range(5):
            c[j + k] += np.vdot(v[j], v[k]).real
    f = lambda t: np.polyval(c[::-1], t) - thr
    if f(h) > 0:
        return h, sum(h ** k * v[k] for k in range(5))
    a, b = 0.0, h
    for _ in range(200):
        m = 0.5 * (a + b)
        if m <= a or m >= b:
            break
        if f(m) > 0:
            a = m
        else:
            b = m
    return b, sum(b ** k * v[k] for k in range(5))


@pytest.mark.parametrize("system,row", [("jc2x4", 2), ("qubits2", 3), ("qubits3", 4), ("qubits4", 5),
                                        ("jc2x3", 2), ("jc2x5", 2), ("jc2x10", 2),
                                        ("jc2x40", 2), ("qubits7", 8), ("qubits10", 11)])
def test_mcwf_advance_matches_stepwise_ensemble_rk4(system, row):
    """qh_mcwf_advance (each trajectory runs from its step index to its first
    norm crossing, located inside the step; several trajectories per wavefront)
    against the ensemble RK4 kernel taken one step at a time with the crossing
    found in numpy: the step index and the time inside it at which every
    trajectory stops, and the amplitudes there, to 1e-12.  Covers thresholds
    that are never crossed, a trajectory that has already taken all steps, a late
    start and a shorter last step; generator rows of 2 to 5 entries (2-, 3-, 4-
    and 16-entry register layouts; 4 to 16 lanes per trajectory; segments of
    exactly 6, 10 and 20 lanes, 10, 6 and 3 trajectories per wavefront) and,
    one workgroup per trajectory, 80, 128 and 1024 amplitudes with rows of 2, 8
    and 11 entries."""
    if system.startswith("jc"):
        qs, A = _jc_generator(levels=int(system[4:]))
    else:
        qs, A = _qubit_ring_generator(int(system[6:]))
    L = qa.lib()
    N = A.info()[0]
    rp, _, _ = A.export()
    assert int(np.diff(rp).max()) == row
    Ad = _dense(A)
    fits = ctypes.c_int()
    qa._qh(L.qh_mcwf_fits(A.h, 2, ctypes.byref(fits)))
    assert fits.value == 1
    M, J, dt, hlast = 37, 60, 0.02, 0.013
    rng = np.random.default_rng(5)
    psi = rng.normal(size=(M, N)) + 1j * rng.normal(size=(M, N))
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    thr = rng.uniform(0.4, 0.99, M)
    thr[:3] = 0.0
    j0 = np.zeros(M, np.int32)
    j0[3] = J
    j0[4] = 17
    j0[5] = J - 1
    ens = qa.DeviceVec(M * N)
    ens.set(psi.ravel())
    jst = j0.copy()
    sub = np.zeros(M)
    jump = np.full(M, 7, np.int32)
    qa._qh(L.qh_mcwf_advance(A.h, ens.h, M, _P(thr, ctypes.c_double), _P(jst, ctypes.c_int32),
                             _P(sub, ctypes.c_double), J, dt, hlast, _P(jump, ctypes.c_int32)))
    got = ens.get().reshape(M, N)
    # reference: one ensemble RK4 step at a time; trajectories that have not
    # started are held at their initial state, stopped ones are frozen at the
    # crossing inside the step
    ref = qa.DeviceVec(M * N)
    cur = psi.copy()
    stop, tin, frozen = j0.copy(), np.zeros(M), {}
    done = j0 >= J
    for j in range(J):
        h = hlast if j == J - 1 else dt
        ref.set(cur.ravel())
        qa._qh(L.qh_ens_rk4(A.h, ref.h, M, h, 1))
        nxt = ref.get().reshape(M, N)
        for i in range(M):
            if done[i] or j < j0[i]:
                continue
            if np.vdot(nxt[i], nxt[i]).real < thr[i]:
                tau, st = _cross_np(Ad, cur[i], h, thr[i])
                inside = tau < h * (1 - 1e-14)
                done[i], frozen[i] = True, (st if inside else nxt[i])
                stop[i], tin[i] = (j, tau) if inside else (j + 1, 0.0)
            else:
                cur[i] = nxt[i]
        if done.all():
            break
    inside_count = 0
    for i in range(M):
        want = frozen.get(i, cur[i] if j0[i] < J else psi[i])
        exp_stop = stop[i] if i in frozen else J
        assert jst[i] == exp_stop, (i, jst[i], exp_stop)
        assert abs(sub[i] - tin[i]) <= 1e-12 * dt, (i, sub[i], tin[i])
        assert jump[i] == (1 if i in frozen else 0), (i, jump[i])
        assert np.abs(got[i] - want).max() < 1e-12, (i, np.abs(got[i] - want).max())
        if i in frozen:
            assert np.vdot(got[i], got[i]).real <= thr[i] * (1 + 1e-12)
            inside_count += tin[i] > 0
    assert len(frozen) >= 3 and M - len(frozen) >= 4   # crossings and runs to the end (4 by construction)
    assert inside_count >= 3
    # a second call resumes the trajectories that stopped (from the time inside
    # their step) and finishes the rest
    qa._qh(L.qh_mcwf_advance(A.h, ens.h, M, _P(np.zeros(M), ctypes.c_double), _P(jst, ctypes.c_int32),
                             _P(sub, ctypes.c_double), J, dt, hlast, _P(jump, ctypes.c_int32)))
    assert (jst == J).all() and (jump == 0).all() and (sub == 0).all()
    qs.destroy()


@pytest.mark.parametrize("nq", [3, 11])
def test_mcwf_cross_matches_numpy_root(nq):
    """qh_mcwf_cross (the crossing of one trajectory of any size, used by the
    step-by-step loop): the time inside the step and the re-taken state against
    the numpy bisection of the RK4 step's norm polynomial (8 and 2048 amplitudes)"""
    qs, A = _qubit_ring_generator(nq)
    L = qa.lib()
    L.qh_mcwf_cross.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    L.qh_mcwf_cross.restype = ctypes.c_int
    N = A.info()[0]
    Ad = _dense(A)
    rng = np.random.default_rng(12)
    psi = rng.normal(size=N) + 1j * rng.normal(size=N)
    psi /= np.linalg.norm(psi)
    h = 0.05
    v = [psi]
    for k in range(1, 5):
        v.append(Ad @ v[-1] / k)
    end = sum(h ** k * v[k] for k in range(5))
    thr = 0.5 * (1.0 + np.vdot(end, end).real)     # crossed inside the step
    tau_np, st_np = _cross_np(Ad, psi, h, thr)
    src, out = qa.DeviceVec(N), qa.DeviceVec(N)
    src.set(psi)
    tau = ctypes.c_double()
    qa._qh(L.qh_mcwf_cross(A.h, src.h, out.h, h, thr, ctypes.byref(tau)))
    assert 0 < tau.value < h and abs(tau.value - tau_np) < 1e-12 * h
    assert np.abs(out.get() - st_np).max() < 1e-12
    assert abs(np.vdot(st_np, st_np).real - thr) < 1e-12
    qs.destroy()


@pytest.mark.parametrize("levels", [4, 100])
def test_mcwf_jump_norms_and_apply_match_dense_products(levels):
    """qh_mcwf_jump_norms: ||C_k psi_i||^2 for every listed trajectory and
    operator; qh_mcwf_jump_apply: psi_i <- s_i C_{k_i} psi_i for the listed
    trajectories only, the others untouched (dense numpy products, 1e-12);
    8 and 200 amplitudes"""
    q1, A1 = _jc_generator(levels=levels)
    q2, A2 = _jc_generator(gam=0.3, kap=2.0, levels=levels)
    L = qa.lib()
    C = [_dense(A1), _dense(A2)]
    N = C[0].shape[0]
    M = 23
    rng = np.random.default_rng(8)
    psi = rng.normal(size=(M, N)) + 1j * rng.normal(size=(M, N))
    ens = qa.DeviceVec(M * N)
    ens.set(psi.ravel())
    ops = (ctypes.c_void_p * 2)(A1.h, A2.h)
    idx = np.array([0, 5, 5, 22, 9, 13], np.int64)
    p = np.empty(len(idx) * 2)
    qa._qh(L.qh_mcwf_jump_norms(2, ops, ens.h, N, _P(idx, ctypes.c_int64), len(idx), _P(p, ctypes.c_double)))
    want = np.array([[np.linalg.norm(C[k] @ psi[i]) ** 2 for k in range(2)] for i in idx]).ravel()
    assert np.allclose(p, want, rtol=1e-12, atol=0)
    idx = np.array([3, 17, 0, 22], np.int64)
    kk = np.array([1, 0, 1, 0], np.int32)
    sre, sim = rng.normal(size=4), rng.normal(size=4)
    qa._qh(L.qh_mcwf_jump_apply(2, ops, ens.h, N, _P(idx, ctypes.c_int64), 4, _P(kk, ctypes.c_int32),
                                _P(sre, ctypes.c_double), _P(sim, ctypes.c_double)))
    got = ens.get().reshape(M, N)
    want = psi.copy()
    for q, i in enumerate(idx):
        want[i] = (sre[q] + 1j * sim[q]) * (C[kk[q]] @ psi[i])
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    q1.destroy()
    q2.destroy()


def test_device_loop_is_deterministic_for_a_seed():
    """the same seed gives the same ensemble average on the device loop"""
    _fixed_rk4()
    vals = []
    for _ in range(2):
        qs = qa.System()
        o = qs.create_op(3)
        qs.add_ham_term(0.3, o.n)
        qs.add_lin_term(0.7, o)
        qs.use_mcwf_solver(500, 21)
        qs.construct_matrix()
        x = qs.create_qvec()
        x.add(1.0, 2)
        qs.time_step(x, 0.0, 2.0, 0.01, 100000)
        vals.append(x.expectation(o.n))
        qs.destroy()
    assert vals[0] == vals[1]


def test_device_loop_respects_steps_max_and_reports_stats():
    """time_step_sys under MCWF stops after steps_max steps (the step-by-step
    rule, also on the device loop) and reports the steps taken and the final
    time through quac_ts_stats; a second call continues from there"""
    _fixed_rk4()
    qs = qa.System()
    o = qs.create_op(3)
    qs.add_ham_term(0.4, o.n)
    qs.add_lin_term(0.5, o)
    qs.use_mcwf_solver(200, 4)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 2)
    steps, t = qs.time_step(x, 0.0, 1.0, 0.01, 37)      # returns quac_ts_stats
    assert steps == 37 and abs(t - 0.37) < 1e-12
    steps, t = qs.time_step(x, t, 1.0, 0.01, 1000)
    assert steps == 63 and abs(t - 1.0) < 1e-12
    e = x.expectation(o.n).real
    assert 0.0 <= e <= 2.0
    qs.destroy()
