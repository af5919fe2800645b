"""The committed workload fixtures (tests/golden/workload_fixtures.json, written by
tests/golden/make_workload_fixtures.py from the CPU oracle) against the oracle
itself and against what fixes them independently of it:

* the nonzero counts of the benchmark systems, computed by an independent sparse
  pattern build (Jaynes-Cummings 2 x 10: 1 544; 10-qubit chain: 24 641 536; Kerr
  3 x 8: 5 222 400; the largest Tavis-Cummings instance, 4 qubits and 30 cavity
  levels: 2 457 856);
* the Jaynes-Cummings trajectory: the Hamiltonian conserves the excitation number
  N = sigma^dag sigma + a^dag a, and with equal cavity and emitter loss rates
  (kappa = gamma = 0.01) d<N>/dt = -0.01 <N>, so <N>(t) = exp(-0.01 t) from one
  excitation (to the RK4 truncation error at dt = 0.01);
* the chain: every RK4 stage of a trace-preserving generator has zero trace, so
  the trace stays 1.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
import recipes
from qbuild import populations_np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
if GOLDEN not in sys.path:
    sys.path.insert(0, GOLDEN)
import make_workload_fixtures as mk  # noqa: E402

FX = json.load(open(os.path.join(GOLDEN, "workload_fixtures.json")))
NNZ = {"jc_2x10": 1544, "chain10": 24641536, "kerr_3x8_train0": 5222400}


def _digest(name):
    r = dict(mk.workloads())[name]
    tot, A = mk.oracle_csr(r)
    return mk.csr_digest(*A.export())


def test_fixture_counts_match_the_independent_pattern_build():
    s = FX["structure"]
    for k, v in NNZ.items():
        assert s[k]["nnz"] == v
    largest = [k for k in s if k.startswith("tc_largest")]
    assert len(largest) == 1 and s[largest[0]]["rows"] == 230400 and s[largest[0]]["nnz"] == 2457856
    assert s["chain10"]["rows"] == 1 << 20 and s["jc_2x10"]["rows"] == 400


@pytest.mark.parametrize("name", [k for k in FX["structure"] if k != "chain10"])
def test_oracle_reproduces_structure_digests(name):
    assert _digest(name) == FX["structure"][name]


def test_jc_trajectory_fixture():
    pops = np.array(FX["jc_trajectory"]["pops"])
    assert pops.shape == (101, 2)
    t = np.arange(101) * 1.0
    # excitation number decays at the common loss rate
    assert np.abs(pops.sum(1) - np.exp(-0.01 * t)).max() < 1e-9
    assert pops[0, 0] == 1.0 and pops[0, 1] == 0.0
    # the oracle reproduces the fixture
    r = recipes.jc_recipe()
    tot, terms = recipes.oracle_terms(r)
    A = oracle.construct(tot, terms, dm=True)
    N, loc = recipes.jc_init()
    x = np.zeros(N * N, np.complex128)
    x[loc] = 1.0
    for k in range(1, 101):
        x = A.rk4(x, 0.01, 100)
        assert np.abs(populations_np(x, r) - pops[k]).max() < 1e-12


def test_chain_structure_and_populations_fixture():
    rows = np.array(FX["chain_populations"]["rows"])
    assert rows.shape == (20, 12)
    assert np.abs(rows[:, 10] - 1.0).max() < 1e-12 and np.abs(rows[:, 11]).max() < 1e-12
    # |0101...>: the populations start at 0/1 and move by O(dt) per step
    assert np.abs(rows[0, :10] - np.array([i % 2 for i in range(10)])).max() < 0.01
    assert _digest("chain10") == FX["structure"]["chain10"]
    got = np.array(mk.chain_populations())
    assert np.abs(got - rows).max() < 1e-12
