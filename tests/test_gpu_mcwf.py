"""Monte Carlo wavefunction trajectories (use_mcwf_solver) on the device.

Parity is statistical (the reference draws from SPRNG, which is not available):
ensemble averages must agree with the density-matrix solution within 4 standard
errors of the mean plus the first-order jump-timing bias (gamma*dt/2 relative;
both bounds are written in each test)."""
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


def test_t1_decay_ensemble():
    """time_step_tests.c:24-101 under MCWF: <n>(t) = exp(-gamma t)"""
    _fixed_rk4()
    M, gamma, dt, T = 8000, 1.0, 0.01, 1.0
    qs = qa.System()
    o = qs.create_op(2)
    qs.add_lin_term(gamma, o)
    qs.use_mcwf_solver(M, 7)
    qs.construct_matrix()
    x = qs.create_qvec()
    x.add(1.0, 1)
    qs.time_step(x, 0.0, T, dt, 100000)
    p = math.exp(-gamma * T)
    got = x.expectation(o.n).real
    sem = math.sqrt(p * (1 - p) / M)
    assert abs(got - p) < 4 * sem + p * gamma * dt, (got, p, sem)
    qs.destroy()


def test_driven_jc_ensemble_vs_density_matrix():
    """emitter (2) x cavity (4), exchange coupling and both losses; averages of
    M trajectories against the FP64 density-matrix RK4 of the same system"""
    _fixed_rk4()
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
    tol_q = 4 * 0.5 / math.sqrt(M) + ref[0] * max(gam, kap) * dt
    tol_c = 4 * 1.5 / math.sqrt(M) + ref[1] * max(gam, kap) * dt
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
