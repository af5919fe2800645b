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
