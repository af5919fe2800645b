"""Write tests/golden/workload_fixtures.json from the CPU oracle (the C restatement
of kron.c / qsystem.c, never the reference itself):

* structure: SHA-256 of the superoperator CSR (row pointers and column indices as
  little-endian int64, values as little-endian complex128) for the Jaynes-Cummings
  2 x 10 system, the 10-qubit dissipative chain (seed 1), the first 3 x 8 Kerr
  training set (seed 4) and three Tavis-Cummings instances of the 512-instance
  sweep (seed 2: the first, the smallest and the largest);
* jc_trajectory: <sigma^dag sigma> and <a^dag a> of the Jaynes-Cummings system every
  100 FP64 RK4 steps (dt = 0.01) over 10 000 steps, from the excited emitter and
  the empty cavity;
* chain_populations: the 10 qubit populations and the trace of the chain after
  each of its first 20 FP64 RK4 steps (dt = 0.005) from |0101...>.

    python tests/golden/make_workload_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
import recipes  # noqa: E402
from qbuild import populations_np  # noqa: E402

OUT = os.path.join(HERE, "workload_fixtures.json")


def csr_digest(rp, col, val):
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes()).hexdigest()
    return {"rows": int(len(rp) - 1), "nnz": int(len(col)), "rowptr_sha256": h(np.asarray(rp, np.int64)),
            "col_sha256": h(np.asarray(col, np.int64)), "val_sha256": h(np.asarray(val, np.complex128))}


def oracle_csr(recipe):
    tot, terms = recipes.oracle_terms(recipe)
    return tot, oracle.construct(tot, terms, dm=True)


def tc_picks():
    insts = recipes.tc_instances(512, 2)
    rows = [((2 ** p[0]) * p[1]) ** 2 for p in insts]
    return {"first": 0, "smallest": int(np.argmin(rows)), "largest": int(np.argmax(rows))}


def workloads():
    """(name, recipe) of every system whose structure is recorded"""
    out = [("jc_2x10", recipes.jc_recipe()), ("chain10", recipes.spin_chain_recipe(10, seed=1)),
           ("kerr_3x8_train0", recipes.kerr_recipe(recipes.kerr_sweep_params(4, 1)[0], levels=8))]
    insts = recipes.tc_instances(512, 2)
    for tag, i in tc_picks().items():
        out.append((f"tc_{tag}_{i}", recipes.tc_recipe(*insts[i])))
    return out


def jc_trajectory():
    r = recipes.jc_recipe()
    tot, A = oracle_csr(r)
    N, loc = recipes.jc_init()
    x = np.zeros(N * N, np.complex128)
    x[loc] = 1.0
    pops = [populations_np(x, r).tolist()]
    for _ in range(100):
        x = A.rk4(x, 0.01, 100)
        pops.append(populations_np(x, r).tolist())
    return pops


def chain_populations(nthreads=8):
    r = recipes.spin_chain_recipe(10, seed=1)
    tot, A = oracle_csr(r)
    N, loc = recipes.spin_chain_init(10)
    x = np.zeros(N * N, np.complex128)
    x[loc] = 1.0
    rows = []
    for _ in range(20):
        x = A.rk4(x, 0.005, 1, nthreads)
        tr = complex(np.sum(x[np.arange(N) * (N + 1)]))
        rows.append(populations_np(x, r).tolist() + [tr.real, tr.imag])
    return rows


def main():
    fx = {"structure": {}, "tc_picks": tc_picks()}
    for name, r in workloads():
        tot, A = oracle_csr(r)
        fx["structure"][name] = csr_digest(*A.export())
        print(name, fx["structure"][name]["rows"], fx["structure"][name]["nnz"], flush=True)
    fx["jc_trajectory"] = {"dt": 0.01, "every": 100, "steps": 10000, "pops": jc_trajectory()}
    fx["chain_populations"] = {"dt": 0.005, "steps": 20, "rows": chain_populations()}
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
