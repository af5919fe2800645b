"""Batched parameter sweep: many Tavis-Cummings instances of uneven size advanced
together (block-diagonal CSR, fused FP32 stages, hipGraph stepping) must match
running each instance alone on the FP64 path: populations within 1e-6 (FP32
batch) and 1e-12 (FP64 batch)."""
import numpy as np
import pytest

import oracle
import quac_amd as qa
import recipes
from qbuild import build_system, subsystem_layout

pytestmark = pytest.mark.gpu


def _instances(k=12, seed=11):
    out = []
    for (nq, nc, g, d, ka, ga) in recipes.tc_instances(k, seed):
        nc = min(nc, 8)  # keep the oracle cross-check quick
        out.append((nq, nc, g, d, ka, ga))
    return out


@pytest.mark.parametrize("prec", [qa.FP32, qa.FP64])
def test_batch_matches_single_instance_fp64(prec):
    insts = _instances()
    systems, mats, tots, x0s, subs = [], [], [], [], []
    for p in insts:
        r = recipes.tc_recipe(*p)
        qs, _, _ = build_system(r)
        N, loc = recipes.tc_init(p[0], p[1])
        x0 = np.zeros(N * N, np.complex128)
        x0[loc] = 1
        systems.append((qs, r))
        mats.append(qs.matrix())
        tots.append(N)
        x0s.append(x0)
        subs.append(subsystem_layout(r))
    b = qa.Batch(mats, tots, prec=prec)
    for i, x0 in enumerate(x0s):
        b.set_state(i, x0)
    steps, dt = 150, 0.01
    b.rk4(dt, steps, renorm_every=50 if prec == qa.FP32 else 0)
    pops = b.populations(subs)
    worst = 0.0
    for i, (qs, r) in enumerate(systems):
        v = qa.DeviceVec(tots[i] ** 2, qa.FP64)
        v.set(x0s[i])
        qa.rk4_device(qs.matrix(), v, dt, steps)
        single = v.get()
        ref_pops = []
        for (L, st) in subs[i]:
            lev = (np.arange(tots[i]) // st) % L
            ref_pops.append(np.sum(lev * np.real(single[np.arange(tots[i]) * (tots[i] + 1)])))
        worst = max(worst, np.abs(np.array(ref_pops) - pops[i]).max())
        if i < 3:  # spot-check the single-instance FP64 path against the oracle
            tot, terms = recipes.oracle_terms(r)
            ref = oracle.construct(tot, terms, dm=True).rk4(x0s[i], dt, steps)
            assert np.abs(single - ref).max() < 1e-10
    assert worst < (1e-6 if prec == qa.FP32 else 1e-12), worst
    b.destroy()
    for qs, _ in systems:
        qs.destroy()


def _full_range_instances():
    """the first 20 instances of the 512-instance sweep (seed 2, cavity
    truncations 5..30) plus the largest shape the sweep allows (4 qubits, 30
    cavity levels: 230 400 superoperator rows)"""
    insts = list(recipes.tc_instances(20, 2))
    insts.append((4, 30, 0.2, 0.1, 1e-2, 1e-2))
    return insts


@pytest.mark.parametrize("prec", [qa.FP32, qa.FP64])
def test_batch_full_range_matches_single_instance_fp64(prec):
    """every cavity truncation of the sweep and the largest instance: the batch
    against each instance alone on the plain FP64 path (populations within 1e-6
    for the FP32 batch, 1e-12 for the FP64 batch), and the plain FP64 path
    against the oracle on the instances of at most 64 levels (1e-10)"""
    insts = _full_range_instances()
    assert max(p[1] for p in insts) == 30 and min(p[1] for p in insts) <= 8
    systems, mats, tots, x0s, subs = [], [], [], [], []
    for p in insts:
        r = recipes.tc_recipe(*p)
        qs, _, _ = build_system(r)
        N, loc = recipes.tc_init(p[0], p[1])
        x0 = np.zeros(N * N, np.complex128)
        x0[loc] = 1
        systems.append((qs, r))
        mats.append(qs.matrix())
        tots.append(N)
        x0s.append(x0)
        subs.append(subsystem_layout(r))
    print(f"batch of {len(insts)} instances, rows {min(t * t for t in tots)}..{max(t * t for t in tots)}, "
          f"precision {prec}", flush=True)
    b = qa.Batch(mats, tots, prec=prec)
    for i, x0 in enumerate(x0s):
        b.set_state(i, x0)
    steps, dt = 100, 0.01
    b.rk4(dt, steps, renorm_every=50 if prec == qa.FP32 else 0)
    pops = b.populations(subs)
    worst, checked = 0.0, 0
    for i, (qs, r) in enumerate(systems):
        N = tots[i]
        v = qa.DeviceVec(N * N, qa.FP64)
        v.set(x0s[i])
        qa.rk4_device(qs.matrix(), v, dt, steps)
        single = v.get()
        diag = np.real(single[np.arange(N) * (N + 1)])
        ref_pops = [np.sum(((np.arange(N) // st) % L) * diag) for (L, st) in subs[i]]
        worst = max(worst, np.abs(np.array(ref_pops) - pops[i]).max())
        if N <= 64 and prec == qa.FP64:
            tot, terms = recipes.oracle_terms(r)
            ref = oracle.construct(tot, terms, dm=True).rk4(x0s[i], dt, steps, nthreads=8)
            assert np.abs(single - ref).max() < 1e-10, i
            checked += 1
    assert worst < (1e-6 if prec == qa.FP32 else 1e-12), worst
    if prec == qa.FP64:
        assert checked >= 5
    b.destroy()
    for qs, _ in systems:
        qs.destroy()
