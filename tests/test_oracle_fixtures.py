"""Pin the CPU oracle against the reference's own golden data.

* the 48 PETSc-binary superoperators of tests/pristine_matrices (written by the
  reference's construct_matrix, t_helpers.c:7-13) must be reproduced bit-exactly:
  row lengths, sorted column indices, explicit zeros and every complex128 value;
* the ham_lin recipes must equal ham_ + lin_ fixtures within the reference's
  1e-9 Frobenius tolerance (t_helpers.c:26-59, t_helpers.h:6);
* the complex-a convention: fixtures use +i*a*conj(v_gi); the current kron.c:723
  (conj(a)) form differs on exactly the 12 complex-Hamiltonian DM fixtures.
"""
import os

import numpy as np
import pytest

import oracle
import petsc_io
import recipes

FX = os.path.join(os.path.dirname(__file__), "golden", "pristine_matrices")


def _build(fam, nop, fl, conj_a=False):
    r = recipes.fixture_recipe(fam, nop, fl)
    tot, terms = recipes.oracle_terms(r)
    return oracle.construct(tot, terms, dm=recipes.is_dm(r), mcwf=r["mcwf"], conj_a=conj_a).export()


@pytest.mark.parametrize("fam,nop,fl", recipes.FIXTURES)
def test_oracle_bit_exact_vs_pristine(fam, nop, fl):
    M, N, rp, col, val = petsc_io.read_mat(os.path.join(FX, f"{fam}_{nop}op_{fl}"))
    orp, ocol, oval = _build(fam, nop, fl)
    assert np.array_equal(orp, rp)
    assert np.array_equal(ocol, col)
    # bit-exact, including the sign of explicit zeros
    assert np.array_equal(oval.view(np.int64), val.view(np.int64))


@pytest.mark.parametrize("nop", (1, 2, 3))
@pytest.mark.parametrize("fl", ("br", "bc", "pr", "pc"))
def test_oracle_ham_lin_sum(nop, fl):
    _, _, rh, ch, vh = petsc_io.read_mat(os.path.join(FX, f"ham_{nop}op_{fl}"))
    _, _, rl, cl, vl = petsc_io.read_mat(os.path.join(FX, f"lin_{nop}op_{fl}"))
    n = len(rh) - 1
    ref = petsc_io.to_dense(n, n, rh, ch, vh) + petsc_io.to_dense(n, n, rl, cl, vl)
    orp, ocol, oval = _build("ham_lin", nop, fl)
    got = petsc_io.to_dense(n, n, orp, ocol, oval)
    assert np.linalg.norm(got - ref) <= 1e-9


def test_complex_a_convention():
    differing = []
    for (fam, nop, fl) in recipes.FIXTURES:
        a = _build(fam, nop, fl, conj_a=False)
        b = _build(fam, nop, fl, conj_a=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])  # structure identical
        if not np.array_equal(a[2], b[2]):
            differing.append((fam, nop, fl))
    assert sorted(differing) == sorted((("ham", n, f) for n in (1, 2, 3) for f in ("bc", "pc")))
