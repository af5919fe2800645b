"""The noisy gate-circuit workload through the product's circuit API.

A seeded random circuit of H, CNOT on adjacent pairs, RZ and RX gates on a
register with per-qubit T1 decay (1e-3) and dephasing (2e-3) and no Hamiltonian.
The gates go through add_gate_to_circuit_sys (run time 0.1 each),
schedule_circuit_layers (greedy layers; layer k at time 1 + 0.1 k),
apply_circuit_to_sys and time_step_sys, whose gate events fire between fixed
RK4 steps of 0.01: ten steps between layers, as in the pattern of
time_step_tests.c:743-754 (quantum_circuits.c:17-120, 316-431;
qsystem.c:1007-1019).

* 6 and 8 qubits against the oracle, which replays the same layers with
  gate_superop and rk4: FP64 state to 1e-10, FP32 bulk populations and trace
  to 1e-6.
* 12 qubits, the full 300-gate workload (rho 4096 x 4096): trace 1, rho
  Hermitian, populations in [0, 1], FP32 within 1e-6 of FP64.
"""
import numpy as np
import pytest

import oracle
import quac_amd as qa
import recipes
from qbuild import build_system, populations_np

pytestmark = pytest.mark.gpu
CODE = {"H": 1, "CNOT": -1, "RZ": 8, "RX": 6}
DT, GAP = 0.01, 0.1


def _layers(gates, nq):
    """greedy ASAP layers (the rule of schedule_circuit_layers)"""
    cur = [0] * nq
    layers = []
    for g in gates:
        lv = max(cur[q] for q in g[1])
        while len(layers) <= lv:
            layers.append([])
        layers[lv].append(g)
        for q in g[1]:
            cur[q] = lv + 1
    return layers


def _run_device(nq, gates, prec):
    """time_step_sys from t = 0.9 over every layer event and 0.1 beyond the last"""
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qa.set_option("quac_precision", "fp32" if prec == qa.FP32 else "fp64")
    qa.set_option("quac_renorm", "10" if prec == qa.FP32 else "0")
    try:
        recipe = recipes.noisy_circuit_recipe(nq)
        qs, _, _ = build_system(recipe)
        circ = qa.Circuit(len(gates))
        for (name, qb, th) in gates:
            circ.add_gate(1.0, name, list(qb), theta=th, run_time=GAP)
        qs.schedule(circ)
        nl = circ.num_layers
        times = circ.layer_times()
        qs.apply_circuit(circ, 0.0)
        x = qs.create_qvec()
        x.add_loc(1.0, 0)
        t_end = 1.0 + GAP * nl
        steps, tend = qs.time_step(x, 0.9, t_end, DT, 10 ** 7)
        out = x.get()
        qs.destroy()
        circ.destroy()
        return out, nl, times, steps, tend
    finally:
        qa.set_option("quac_precision", "fp64")
        qa.set_option("quac_renorm", "0")


def _run_oracle(nq, layers):
    recipe = recipes.noisy_circuit_recipe(nq)
    tot, terms = recipes.oracle_terms(recipe)
    m = oracle.construct(tot, terms, dm=True)
    x = np.zeros(tot * tot, np.complex128)
    x[0] = 1.0
    nsteps = int(round(GAP / DT))
    x = m.rk4(x, DT, nsteps, nthreads=8)
    for layer in layers:
        for (name, qb, th) in layer:
            x = oracle.gate_superop([2] * nq, [2 ** s for s in range(nq)], True, CODE[name], qb, x, theta=th)
        x = m.rk4(x, DT, nsteps, nthreads=8)
    return x


@pytest.mark.parametrize("nq,ngates,seed", [(6, 60, 31), (8, 60, 32)])
def test_noisy_circuit_vs_oracle(nq, ngates, seed):
    gates = recipes.random_circuit(nq, ngates, seed)
    layers = _layers(gates, nq)
    print(f"noisy circuit nq={nq} gates={ngates} seed={seed} layers={len(layers)}", flush=True)
    ref = _run_oracle(nq, layers)
    recipe = recipes.noisy_circuit_recipe(nq)
    pops_ref = populations_np(ref, recipe)
    tot = 2 ** nq
    x64, nl, times, steps, tend = _run_device(nq, gates, qa.FP64)
    assert nl == len(layers)
    assert np.allclose(times, [1.0 + GAP * k for k in range(nl)], rtol=0, atol=1e-12)
    assert steps == 10 * (nl + 1) and abs(tend - (1.0 + GAP * nl)) < 1e-9
    assert np.abs(x64 - ref).max() < 1e-10
    x32, _, _, steps32, _ = _run_device(nq, gates, qa.FP32)
    assert steps32 == steps
    assert np.abs(populations_np(x32, recipe) - pops_ref).max() < 1e-6
    tr_ref = ref[np.arange(tot) * (tot + 1)].sum()
    assert abs(x32[np.arange(tot) * (tot + 1)].sum() - tr_ref) < 1e-6
    assert abs(tr_ref - 1.0) < 1e-12


def test_noisy_circuit_12_qubits_properties():
    nq = 12
    tot = 2 ** nq
    gates = recipes.random_circuit(nq, 300, 3)
    recipe = recipes.noisy_circuit_recipe(nq)
    diag = np.arange(tot) * (tot + 1)
    res = {}
    for prec in (qa.FP64, qa.FP32):
        print(f"noisy circuit 12 qubits, precision {prec}", flush=True)
        x, nl, _, steps, _ = _run_device(nq, gates, prec)
        rho = x.reshape(tot, tot).T  # vec(rho) is column-major
        herm = np.abs(rho - rho.conj().T).max()
        tr = x[diag].sum()
        pops = populations_np(x, recipe)
        tol = 1e-12 if prec == qa.FP64 else 1e-6
        assert herm < tol, herm
        assert abs(tr - 1.0) < tol, tr
        assert np.all(pops >= -tol) and np.all(pops <= 1 + tol), pops
        assert np.all(np.real(x[diag]) >= -tol)
        res[prec] = pops
        del rho, x
    assert np.abs(res[qa.FP32] - res[qa.FP64]).max() < 1e-6
