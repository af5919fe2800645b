"""Steady state with the Hamiltonian-eigenbasis preconditioner (qh_precond.hip).

Pinned by mathematics, not by a restatement of the preconditioner:
  * when the decay operator commutes with H (no drive, equal loss on every
    cavity, excitation-conserving hopping), the eigenbasis of H also diagonalises
    sum a C^dag C, so M is exactly the no-jump Liouvillian L0 = L - J with
    J rho = sum a C rho C^dag built here from kron(conj(C), C): M^-1 (L0 v) = v;
  * the solves reach the requested residual, agree with the CPU oracle's GMRES
    solution and with the analytic two-level steady state, and need far fewer
    iterations than the Jacobi-preconditioned solve on driven Kerr cavities."""
import numpy as np
import pytest

import dense
import oracle
import petsc_io
import quac_amd as qa
import recipes
from qbuild import build_system

pytestmark = pytest.mark.gpu


def _annihilators(levels):
    out = []
    for k, L in enumerate(levels):
        a = np.diag(np.sqrt(np.arange(1, L)), 1)
        m = np.eye(1)
        for j, Lj in enumerate(levels):
            m = np.kron(m, a if j == k else np.eye(Lj))
        out.append(m)
    return out


def test_pc_inverts_no_jump_part_exactly_when_decay_commutes():
    levels = [3, 4]
    # equal frequencies 2.0 keep the excitation-number sectors of H apart, so the
    # eigenvectors of H are eigenvectors of the total number (and of the decay)
    p = {"terms": [("ham", 2.0, [(0, "n")]), ("ham", 2.0, [(1, "n")]),
                   ("ham", 0.3, [(0, "raise"), (1, "lower")]), ("ham", 0.3, [(1, "raise"), (0, "lower")]),
                   ("ham", 0.2, [(0, "raise"), (0, "raise"), (0, "lower"), (0, "lower")]),
                   ("lin", 0.15, [(0, "lower")]), ("lin", 0.15, [(1, "lower")])],
         "subs": [("op", L) for L in levels], "mcwf": False}
    tot, terms = recipes.oracle_terms(p)
    n = tot * tot
    Lm = petsc_io.to_dense(n, n, *oracle.construct(tot, terms, dm=True).export())
    J = sum(0.15 * np.kron(C.conj(), C) for C in _annihilators(levels))
    L0 = Lm - J
    rng = np.random.default_rng(3)
    v = rng.normal(size=tot * tot) + 1j * rng.normal(size=tot * tot)
    qs, _, _ = build_system(p)
    pc = qa.HBasisPC(qs)
    vi, vo = qa.DeviceVec(tot * tot), qa.DeviceVec(tot * tot)
    vi.set(L0 @ v)
    pc.apply(vi, vo)
    # the vacuum coherence |0><0| (vec index 0) is the null direction of L0 here
    # (no drive): L0 v carries none of it and M^-1 returns none of it
    got = vo.get()
    want = v.copy()
    want[0] = 0
    assert abs(got[0]) < 1e-4
    got[0] = 0
    assert np.abs(got - want).max() < 1e-10 * np.abs(v).max()
    pc.destroy()
    qs.destroy()


@pytest.mark.parametrize("omega,delta,gamma,drive", [(1.0, 0.0, 0.5, "x"), (0.4, 0.3, 1.0, "x"), (0.4, 0.3, 1.0, "y")])
def test_two_level_analytic_every_preconditioner(omega, delta, gamma, drive):
    """rho_ee of a driven, detuned, damped qubit in closed form, for every
    preconditioner; a sigma_y drive (imaginary matrix elements) gives the same
    populations and takes the complex eigensolver path of the eigenbasis
    preconditioner, a sigma_x drive the real one"""
    qs = qa.System()
    o = qs.create_op(2)
    qs.add_ham_term(omega / 2, o.sig_x if drive == "x" else o.sig_y)
    qs.add_ham_term(delta, o.n)
    qs.add_lin_term(gamma, o)
    qs.construct_matrix()
    x = qs.create_qvec()
    ree = (omega ** 2 / 4) / (delta ** 2 + gamma ** 2 / 4 + omega ** 2 / 2)
    for pc in ("none", "jacobi", "hbasis"):
        x.set(np.zeros(4))
        it, res = qs.steady_state(x, tol=1e-12, precond=pc)
        rho = dense.unvec(x.get(), 2)
        assert res < 1e-11 and abs(rho[1, 1].real - ree) < 1e-10 and abs(np.trace(rho) - 1) < 1e-10, pc
    qs.destroy()


def test_kerr_l4_hbasis_vs_oracle_and_jacobi():
    p = recipes.kerr_params(4, 1)[0]
    r = recipes.kerr_recipe(p, levels=4)
    tot, terms = recipes.oracle_terms(r)
    m = oracle.construct(tot, terms, dm=True)
    xo, ito, reso = m.gmres(tot, tol=1e-10, jacobi=True, maxit=200000)
    assert reso <= 1e-10
    qs, _, _ = build_system(r)
    x = qs.create_qvec()
    it_h, res_h = qs.steady_state(x, tol=1e-10, precond="hbasis")
    got = x.get()
    assert res_h <= 1e-10
    assert np.abs(got - xo).max() < 1e-7
    rho = dense.unvec(got, tot)
    assert abs(np.trace(rho) - 1) < 1e-10 and np.abs(rho - rho.conj().T).max() < 1e-8
    x.set(np.zeros(tot * tot))
    it_j, res_j = qs.steady_state(x, tol=1e-10, precond="jacobi", maxit=200000)
    print(f"Kerr 3x4 steady state: hbasis {it_h} iterations, Jacobi {it_j}, oracle Jacobi {ito}")
    assert it_h * 5 < it_j
    qs.destroy()


def test_warm_start_from_solution_returns_immediately():
    p = recipes.kerr_params(4, 1)[0]
    qs, _, _ = build_system(recipes.kerr_recipe(p, levels=3))
    x = qs.create_qvec()
    it, res = qs.steady_state(x, tol=1e-10)
    it2, res2 = qs.steady_state(x, tol=1e-10)
    assert it2 == 0 and res2 <= 1e-10
    qs.destroy()


def test_augmented_gmres_with_exact_and_poor_predictions():
    """qh_gmres_aug (steady_state_sys_aug): with the exact steady state as the
    augmentation vector the first iteration already meets the tolerance; with a
    poor prediction (a random Hermitian unit-trace state) the solve needs no more
    iterations than the plain one, and both reach the same solution"""
    r = recipes.kerr_recipe(recipes.kerr_sweep_params(9, 1)[0], levels=4)
    qs, _, _ = build_system(r)
    N = 64
    x = qs.create_qvec()
    it_c, res_c = qs.steady_state(x, restart=100, tol=1e-10, maxit=3000, precond="hbasis")
    exact = x.get()
    rng = np.random.default_rng(4)
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    poor = (a @ a.conj().T)
    poor = (poor / np.trace(poor)).reshape(-1, order="F")
    for pred, bound in ((exact, 1), (poor, it_c)):
        u = qs.create_qvec()
        u.set(pred)
        y = qs.create_qvec()
        it_a, res_a = qs.steady_state(y, restart=100, tol=1e-10, maxit=3000, precond="hbasis", aug=u)
        assert res_a <= 1e-10 and it_a <= bound, (it_a, bound)
        assert np.abs(y.get() - exact).max() < 1e-8
    qs.destroy()
