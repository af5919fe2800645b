"""Device superoperator assembly (qh_assemble via construct_matrix) against the
reference's pristine PETSc matrices and the CPU oracle: row pointers and column
indices bit-exact, and values bit-exact (the assembly kernel evaluates the same
C99 complex products in the same insertion order as PETSc's MatSetValue)."""
import os

import numpy as np
import pytest

import oracle
import petsc_io
import quac_amd as qa
import recipes
from qbuild import build_system

pytestmark = pytest.mark.gpu
FX = os.path.join(os.path.dirname(__file__), "golden", "pristine_matrices")


def _device_csr(recipe):
    qs, _, _ = build_system(recipe)
    rp, col, val = qs.matrix().export()
    qs.destroy()
    return rp, col, val


@pytest.mark.parametrize("fam,nop,fl", recipes.FIXTURES)
def test_device_assembly_bit_exact_vs_pristine(fam, nop, fl):
    print(f"fixture {fam}_{nop}op_{fl}", flush=True)
    M, N, rp, col, val = petsc_io.read_mat(os.path.join(FX, f"{fam}_{nop}op_{fl}"))
    drp, dcol, dval = _device_csr(recipes.fixture_recipe(fam, nop, fl))
    assert np.array_equal(drp, rp)
    assert np.array_equal(dcol, col)
    assert np.array_equal(dval.view(np.int64), val.view(np.int64))


def test_device_assembly_conj_a_option():
    qa.set_option("quac_kron_conj_a", 1)
    try:
        for (fam, nop, fl) in [("ham", 1, "bc"), ("ham", 3, "pc"), ("ham", 2, "br")]:
            r = recipes.fixture_recipe(fam, nop, fl)
            tot, terms = recipes.oracle_terms(r)
            orp, ocol, oval = oracle.construct(tot, terms, dm=True, conj_a=True).export()
            drp, dcol, dval = _device_csr(r)
            assert np.array_equal(drp, orp) and np.array_equal(dcol, ocol)
            assert np.array_equal(dval.view(np.int64), oval.view(np.int64))
    finally:
        qa.set_option("quac_kron_conj_a", 0)


def _random_recipe(seed):
    rng = np.random.default_rng(seed)
    subs = [("op", 2), ("op", 3), ("vec", 3), ("op", 2)]
    names2 = ["lower", "raise", "n", "eye", "sx", "sy", "sz", "other"]
    names3 = ["lower", "raise", "n", "eye", "sx", "other"]
    terms = []
    for t in range(10):
        ops = []
        for _ in range(rng.integers(1, 4)):
            s = int(rng.integers(0, 4))
            if subs[s][0] == "vec":
                ops += [(s, ("vec", int(rng.integers(0, 3)))), (s, ("vec", int(rng.integers(0, 3))))]
            else:
                pool = names2 if subs[s][1] == 2 else names3
                ops.append((s, pool[int(rng.integers(0, len(pool)))]))
        kind = "ham" if t % 2 == 0 else "lin"
        a = complex(rng.normal(), rng.normal()) if t % 3 else float(rng.normal())
        terms.append((kind, a, ops))
    return {"subs": subs, "terms": terms, "mcwf": False}


@pytest.mark.parametrize("seed", range(6))
def test_device_assembly_bit_exact_vs_oracle_random(seed):
    r = _random_recipe(seed)
    tot, terms = recipes.oracle_terms(r)
    orp, ocol, oval = oracle.construct(tot, terms, dm=True).export()
    drp, dcol, dval = _device_csr(r)
    assert np.array_equal(drp, orp) and np.array_equal(dcol, ocol)
    assert np.array_equal(dval.view(np.int64), oval.view(np.int64))


@pytest.mark.parametrize("name", ["jc", "chain8", "tc_big", "kerr"])
def test_device_assembly_workloads_vs_oracle(name):
    r = {"jc": lambda: recipes.jc_recipe(),
         "chain8": lambda: recipes.spin_chain_recipe(8, seed=1),
         "tc_big": lambda: recipes.tc_recipe(4, 12, 0.1, 0.05, 5e-3, 5e-3),
         "kerr": lambda: recipes.kerr_recipe(recipes.kerr_params(4, 1)[0])}[name]()
    tot, terms = recipes.oracle_terms(r)
    orp, ocol, oval = oracle.construct(tot, terms, dm=True).export()
    drp, dcol, dval = _device_csr(r)
    assert np.array_equal(drp, orp) and np.array_equal(dcol, ocol)
    assert np.array_equal(dval.view(np.int64), oval.view(np.int64))


def test_device_assembly_spin_chain_10_structure():
    """10-qubit chain: 1 048 576 rows; nnz = N^2 (1 + 2n) + n N^2 / 4 = 24 641 536,
    every diagonal stored, rows sorted, and the row-length histogram (21..31)."""
    r = recipes.spin_chain_recipe(10, seed=1)
    rp, col, val = _device_csr(r)
    n, N2 = 10, 4 ** 10
    assert len(rp) - 1 == N2
    assert rp[-1] == N2 * (1 + 2 * n) + n * N2 // 4 == 24641536
    lens = np.diff(rp)
    assert lens.min() == 21 and lens.max() == 31
    rows = np.repeat(np.arange(N2), lens)
    assert np.all(col[rows == col] == rows[rows == col]) and np.sum(rows == col) == N2
    # sortedness within rows
    d = np.diff(col)
    starts = rp[1:-1]
    mask = np.ones(len(d), bool)
    mask[starts - 1] = False
    assert np.all(d[mask] > 0)


def test_device_wf_and_mcwf_assembly():
    for fl in ("br", "bc", "pr", "pc"):
        for fam in ("ham_only", "mcwf"):
            M, N, rp, col, val = petsc_io.read_mat(os.path.join(FX, f"{fam}_2op_{fl}"))
            drp, dcol, dval = _device_csr(recipes.fixture_recipe(fam, 2, fl))
            assert np.array_equal(drp, rp) and np.array_equal(dcol, col)
            assert np.array_equal(dval.view(np.int64), val.view(np.int64))


def test_device_assembly_long_rows_general_path():
    """rows with more than 32 distinct columns (powers of the ladder operators of
    a 40-level mode: up to 81 entries per row) take the general assembly path
    (per-row insertion in memory); bit-exact against the oracle like the fast path"""
    terms = [("ham", 0.1 * k + 0.05j, [(0, "raise")] * k) for k in range(1, 21)]
    terms += [("ham", 0.1 * k - 0.05j, [(0, "lower")] * k) for k in range(1, 21)]
    terms += [("lin", 0.3, [(0, "lower")])]
    r = {"subs": [("op", 40)], "terms": terms, "mcwf": False}
    tot, oterms = recipes.oracle_terms(r)
    orp, ocol, oval = oracle.construct(tot, oterms, dm=True).export()
    assert np.diff(orp).max() > 32
    drp, dcol, dval = _device_csr(r)
    assert np.array_equal(drp, orp) and np.array_equal(dcol, ocol)
    assert np.array_equal(dval.view(np.int64), oval.view(np.int64))


def test_device_assembly_noisy_circuit_12_structure():
    """12 qubits with T1 decay and dephasing (superoperator 16 777 216 rows): the
    diagonal plus the sigma-minus G* x G entry of every qubit whose row and column
    digits are both 0, i.e. nnz = N^2 (1 + n/4) = 67 108 864, at most 13 per row"""
    r = recipes.noisy_circuit_recipe(12)
    qs, _, _ = build_system(r)
    rows, nnz = qs.matrix().info()
    assert rows == 4 ** 12 and nnz == 4 ** 12 * 4 == 67108864
    qs.destroy()
