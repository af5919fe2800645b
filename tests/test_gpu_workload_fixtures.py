"""The device path against the committed workload fixtures
(tests/golden/workload_fixtures.json): the superoperators of the benchmark
systems, assembled on the device and exported, hash to the fixture's digests
(row pointers, column indices and complex128 values bit-exact); the
Jaynes-Cummings trajectory through time_step_sys (fixed-step RK4, FP64) and the
chain's first 20 RK4 steps (FP64 and FP32) match the fixture."""
import json
import os
import sys

import numpy as np
import pytest

import quac_amd as qa
import recipes
from qbuild import build_system, populations_np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)
import make_workload_fixtures as mk  # noqa: E402

pytestmark = pytest.mark.gpu
FX = json.load(open(os.path.join(GOLDEN, "workload_fixtures.json")))


@pytest.mark.parametrize("name", list(FX["structure"]))
def test_device_assembly_matches_structure_digest(name):
    print("workload", name, flush=True)
    r = dict(mk.workloads())[name]
    qs, _, _ = build_system(r)
    rp, col, val = qs.matrix().export()
    qs.destroy()
    assert mk.csr_digest(rp, col, val) == FX["structure"][name]


def test_jc_trajectory_through_time_step_sys():
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qa.set_option("quac_precision", "fp64")
    try:
        r = recipes.jc_recipe()
        N, loc = recipes.jc_init()
        qs, _, _ = build_system(r)
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        pops = np.array(FX["jc_trajectory"]["pops"])
        for k in range(1, 101):
            steps, tend = qs.time_step(x, k - 1.0, float(k), 0.01, 100000)
            assert steps == 100 and abs(tend - k) < 1e-9
            assert np.abs(populations_np(x.get(), r) - pops[k]).max() < 1e-10, k
        qs.destroy()
    finally:
        qa.set_option("ts_rk_type", "3bs")
        qa.set_option("ts_adapt_type", "basic")


@pytest.mark.parametrize("prec,tol", [(qa.FP64, 1e-10), (qa.FP32, 1e-6)])
def test_chain_first_steps(prec, tol):
    r = recipes.spin_chain_recipe(10, seed=1)
    N, loc = recipes.spin_chain_init(10)
    qs, _, _ = build_system(r)
    A = qs.matrix()
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1.0
    v = qa.DeviceVec(N * N, prec)
    v.set(x0)
    rows = np.array(FX["chain_populations"]["rows"])
    for k in range(20):
        qa.rk4_device(A, v, 0.005, 1, tot=N, renorm_every=0)
        x = v.get()
        tr = complex(np.sum(x[np.arange(N) * (N + 1)]))
        got = np.concatenate([populations_np(x, r), [tr.real, tr.imag]])
        assert np.abs(got - rows[k]).max() < tol, k
    qs.destroy()
