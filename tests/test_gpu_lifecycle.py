"""Device memory across repeated use: systems built, stepped on every path
(fixed-step RK4 in FP64 and FP32, the adaptive integrator, the batch entry
point, an MCWF ensemble, the steady-state solve) and destroyed, twenty times
over, with their states destroyed as the C API requires (destroy_qvec), must
give their device memory back (what the library keeps on purpose, its
reduction scratch and the batch kept between calls, is allocated by the first
cycle or released explicitly)."""
import numpy as np
import pytest

import quac_amd as qa
import recipes
from qbuild import build_system

pytestmark = pytest.mark.gpu


def _cycle():
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    r = recipes.jc_recipe(levels=6)
    N, loc = recipes.jc_init(6)
    qs, _, _ = build_system(r)
    for prec in ("fp64", "fp32"):
        qa.set_option("quac_precision", prec)
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        qs.time_step(x, 0.0, 0.5, 0.01, 1000)
        x.destroy()
    qa.set_option("quac_precision", "fp64")
    qa.set_option("ts_rk_type", "3bs")
    qa.set_option("ts_adapt_type", "basic")
    x = qs.create_qvec()
    x.add_loc(1.0, loc)
    qs.time_step(x, 0.0, 0.5, 0.05, 1000)
    x.destroy()
    y = qs.create_qvec()
    qs.steady_state(y, tol=1e-8)
    y.destroy()
    # the batch entry point on three instances, then its kept batch released
    insts = [recipes.tc_recipe(*p) for p in recipes.tc_instances(3, 5)]
    systems = [build_system(rr)[0] for rr in insts]
    states = []
    for qb, p in zip(systems, recipes.tc_instances(3, 5)):
        v = qb.create_qvec()
        v.add_loc(1.0, recipes.tc_init(p[0], p[1])[1])
        states.append(v)
    qa.time_step_batch(systems, states, 0.0, 0.2, 0.01, 1000)
    qa.lib().time_step_sys_batch_release()
    for v in states:
        v.destroy()
    for qb in systems:
        qb.destroy()
    # an MCWF ensemble
    qm = qa.System()
    o = qm.create_op(4)
    qm.add_ham_term(1.0, o.n)
    qm.add_lin_term(0.5, o)
    qm.use_mcwf_solver(64, 3)
    w = qm.create_qvec()
    w.add(1.0, 2)
    qm.construct_matrix()
    qm.time_step(w, 0.0, 0.5, 0.01, 1000)
    w.destroy()
    qm.destroy()
    qs.destroy()


def test_repeated_use_returns_device_memory():
    _cycle()
    free0 = qa.mem_info()[0]
    for _ in range(20):
        _cycle()
    free1 = qa.mem_info()[0]
    # 20 cycles each allocate some tens of MB; a leak of one buffer per cycle
    # would show as several MB
    assert free0 - free1 < (8 << 20), (free0, free1)
