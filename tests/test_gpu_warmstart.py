"""The learned steady-state warm start end to end at a small size (3 Kerr
cavities x 4 levels): training data from the device solver, basis, network
trained on the device, then held-out solves with and without the warm start.

Two sweeps.  Over the wide parameter box of the full-size benchmark (every
parameter over its whole range) the network's prediction is far from the
steady state (the family is not low-dimensional there), so the test checks
the pipeline: every seeded or augmented solve reaches the plain solve's
residual, the warm starts are Hermitian with unit trace, training lowers the
loss, the prediction is closer to the steady state than the trivial guess (the
mean of the training solutions), and the solve whose search space is augmented
by the prediction needs no more iterations than the plain solve on any set.
Over a local sweep (every parameter within 5 % of a centre set) the network
predicts the steady state to about a percent, and the solve started from its
prediction needs fewer GMRES iterations than the solve from zero.  The
full-size comparisons are bench.py --workload kerr."""
import numpy as np
import pytest

import quac_amd as qa
import recipes
from qbuild import build_system
from quac_amd.warmstart import WarmStart, hermitian_unit_trace

pytestmark = pytest.mark.gpu

L = 4
N = L ** 3


def _solve(p, x0=None, tol=1e-8, aug=None):
    qs, _, _ = build_system(recipes.kerr_recipe(p, levels=L))
    x = qs.create_qvec()
    if x0 is not None:
        x.set(x0)
    u = None
    if aug is not None:
        u = qs.create_qvec()
        u.set(aug)
    it, res = qs.steady_state(x, restart=100, tol=tol, maxit=3000, precond="hbasis", aug=u)
    sol = x.get()
    qs.destroy()
    return sol, it, res


def test_warm_start_pipeline_small():
    train = recipes.kerr_sweep_params(4, 300)
    feats, sols = [], []
    for p in train:
        sol, it, res = _solve(p, tol=1e-10)
        assert res <= 1e-10
        feats.append(recipes.kerr_features(p))
        sols.append(sol)
    sols = np.array(sols)
    ws = WarmStart(N, rank=48, hidden=(128, 128, 128), seed=1)
    stats = ws.fit(np.array(feats), sols, epochs=400, batch=100, lr=3e-3)
    assert stats["losses"][-1][1] < stats["losses"][0][1] / 5
    trivial = hermitian_unit_trace(sols.mean(0), N)
    cold, warm, augm, e_net, e_triv = [], [], [], [], []
    for p in recipes.kerr_sweep_params(5, 24):
        sol, it_c, res_c = _solve(p)
        x0 = ws.predict([recipes.kerr_features(p)])[0]
        rho = x0.reshape(N, N, order="F")
        assert abs(np.trace(rho) - 1) < 1e-10 and np.abs(rho - rho.conj().T).max() < 1e-12
        _, it_w, res_w = _solve(p, x0=x0)
        sol_a, it_a, res_a = _solve(p, aug=x0)
        assert res_c <= 1e-8 and res_w <= 1e-8 and res_a <= 1e-8
        assert np.abs(sol_a - sol).max() < 1e-6
        assert it_a <= it_c, (it_a, it_c)     # one restart cycle: never more than the plain solve
        cold.append(it_c)
        warm.append(it_w)
        augm.append(it_a)
        n = np.linalg.norm(sol)
        e_net.append(np.linalg.norm(x0 - sol) / n)
        e_triv.append(np.linalg.norm(trivial - sol) / n)
    ws.destroy()
    print(f"Kerr 3x{L}: cold {np.mean(cold):.1f}, seeded {np.mean(warm):.1f}, augmented {np.mean(augm):.1f} "
          f"iterations; warm-start error {np.mean(e_net):.3f} (trivial guess {np.mean(e_triv):.3f})")
    assert np.mean(e_net) < 0.95 * np.mean(e_triv)


def test_warm_start_cuts_iterations_on_a_local_sweep():
    train = recipes.kerr_local_params(4, 300)
    feats, sols = [], []
    for p in train:
        sol, it, res = _solve(p, tol=1e-10)
        assert res <= 1e-10
        feats.append(recipes.kerr_features(p))
        sols.append(sol)
    ws = WarmStart(N, rank=32, hidden=(128, 128, 128), seed=1)
    ws.fit(np.array(feats), np.array(sols), epochs=600, batch=100, lr=3e-3)
    cold, warm, err = [], [], []
    for p in recipes.kerr_local_params(5, 16):
        sol, it_c, res_c = _solve(p)
        x0 = ws.predict([recipes.kerr_features(p)])[0]
        sol_w, it_w, res_w = _solve(p, x0=x0)
        assert res_c <= 1e-8 and res_w <= 1e-8
        assert np.abs(sol_w - sol).max() < 1e-6
        cold.append(it_c)
        warm.append(it_w)
        err.append(np.linalg.norm(x0 - sol) / np.linalg.norm(sol))
    ws.destroy()
    print(f"local sweep, Kerr 3x{L}: cold {np.mean(cold):.2f}, warm {np.mean(warm):.2f} iterations "
          f"({np.sum(cold) / np.sum(warm):.3f}x); warm-start error {np.mean(err):.4f}")
    assert np.mean(err) < 0.1
    assert np.mean(warm) < np.mean(cold)


def test_training_set_smaller_than_a_batch():
    """fit with fewer parameter sets than the batch size trains on all of them
    as one batch (and predicts Hermitian, unit-trace warm starts)"""
    train = recipes.kerr_sweep_params(4, 12)
    feats, sols = [], []
    for p in train:
        sol, it, res = _solve(p, tol=1e-10)
        feats.append(recipes.kerr_features(p))
        sols.append(sol)
    ws = WarmStart(N, rank=8, hidden=(32, 32), seed=2)
    stats = ws.fit(np.array(feats), np.array(sols), epochs=30, batch=250, lr=3e-3)
    assert stats["rank"] <= 8 and len(stats["losses"]) > 0
    x0 = ws.predict([recipes.kerr_features(recipes.kerr_sweep_params(5, 1)[0])])[0]
    rho = x0.reshape(N, N, order="F")
    assert abs(np.trace(rho) - 1) < 1e-10 and np.abs(rho - rho.conj().T).max() < 1e-12
    ws.destroy()
