"""time_step_sys_batch, the QuaC-level entry point of the batched sweep: systems of
uneven size built through the QuaC API and advanced together must match
time_step_sys (fixed-step RK4, FP64) on each of them alone, over an interval
that ends with a shortened step and over a second call that reuses the kept
batch: states within 1e-12 (FP64 batch) and populations within 1e-6 (FP32
batch).  Systems the batch cannot take are reported as errors."""
import numpy as np
import pytest

import quac_amd as qa
import recipes
from qbuild import build_system, populations_np

pytestmark = pytest.mark.gpu


def _recipes():
    return [(recipes.jc_recipe(levels=4), recipes.jc_init(4)),
            (recipes.tc_recipe(2, 5, 0.12, 0.05, 4e-3, 6e-3), recipes.tc_init(2, 5)),
            (recipes.tc_recipe(1, 7, 0.08, -0.03, 2e-3, 9e-3), recipes.tc_init(1, 7)),
            (recipes.spin_chain_recipe(3, seed=5), recipes.spin_chain_init(3))]


def _states(systems):
    xs = []
    for (qs, r, (N, loc)) in systems:
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        xs.append(x)
    return xs


@pytest.fixture
def fixed_rk4():
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qa.set_option("quac_precision", "fp64")
    yield
    qa.set_option("ts_rk_type", "3bs")
    qa.set_option("ts_adapt_type", "basic")
    qa.set_option("quac_batch_precision", "fp32")
    qa.set_option("quac_renorm", "0")
    qa.lib().time_step_sys_batch_release()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batch_entry_point_matches_time_step_sys(prec, fixed_rk4):
    qa.set_option("quac_batch_precision", prec)
    qa.set_option("quac_renorm", "50" if prec == "fp32" else "0")
    systems = []
    for (r, init) in _recipes():
        qs, _, _ = build_system(r)
        systems.append((qs, r, init))
    xb = _states(systems)
    xs = _states(systems)
    for (t0, t1) in ((0.0, 1.234), (1.234, 2.0)):
        steps, tend = qa.time_step_batch([s[0] for s in systems], xb, t0, t1, 0.01, 100000)
        assert abs(tend - t1) < 1e-12
        assert steps == int(np.ceil((t1 - t0) / 0.01 - 1e-9))
        for (qs, r, _), x, y in zip(systems, xs, xb):
            qs.time_step(x, t0, t1, 0.01, 100000)
            a, b = x.get(), y.get()
            if prec == "fp64":
                assert np.abs(a - b).max() < 1e-12
            else:
                assert np.abs(populations_np(a, r) - populations_np(b, r)).max() < 1e-6
                n = int(round(np.sqrt(len(b))))
                assert abs(np.sum(b[np.arange(n) * (n + 1)]) - 1.0) < 1e-6
    for (qs, _, _) in systems:
        qs.destroy()


def test_batch_entry_point_rejects_what_it_cannot_step(fixed_rk4):
    (r, (N, loc)) = _recipes()[0]
    qs, _, _ = build_system(r)
    x = qs.create_qvec()
    x.add_loc(1.0, loc)
    # a state of another system's size
    r2, (N2, loc2) = _recipes()[1]
    qs2, _, _ = build_system(r2)
    y = qs2.create_qvec()
    with pytest.raises(qa.QuacError, match="does not match"):
        qa.time_step_batch([qs], [y], 0.0, 0.1, 0.01, 100)
    # time-dependent terms
    q3, subs, op = build_system(r, construct=False)
    q3.add_ham_term_time_dep(0.1, lambda t: np.cos(t), op((0, "lower")))
    q3.construct_matrix()
    z = q3.create_qvec()
    z.add_loc(1.0, loc)
    with pytest.raises(qa.QuacError, match="time-dependent"):
        qa.time_step_batch([q3], [z], 0.0, 0.1, 0.01, 100)
    with pytest.raises(qa.QuacError, match="dt"):
        qa.time_step_batch([qs], [x], 0.0, 0.1, -0.01, 100)
    for s in (qs, qs2, q3):
        s.destroy()
