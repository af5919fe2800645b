"""Secondary workloads of bench.py (selected with --workload); each returns the
dict that bench.py prints as its JSON line.

tc512     : batched parameter sweep of Tavis-Cummings instances of uneven size
            (rows 100 .. 230 400), FP32 batch (one block-diagonal sliced matrix,
            fused stages, hipGraph) vs the plain FP64 path run one instance at a
            time; instance-steps/s and max population error.
circuit12 : 12-qubit noisy circuit on rho (4096 x 4096): 300 seeded gates in
            greedy layers, RK4 (dt = 0.01, 10 steps) between layers; FP32 bulk vs
            FP64, per-qubit populations and trace.
jc        : Jaynes-Cummings 2 x 10 (400-row superoperator), RK4 dt = 0.01 to
            t = 100 with populations every 100 steps (latency-bound).
kerr      : learned steady-state warm start on 3 driven Kerr cavities x 8
            levels (superoperator 262 144 rows): steady states of the training
            parameter sets from the device GMRES, low-rank basis, network
            trained on the device, then GMRES to 1e-8 on held-out sets with and
            without the network's warm start.
mcwf      : Monte Carlo wavefunction ensemble of the Jaynes-Cummings 2 x 10 system
            (10 000 trajectories, RK4 dt = 0.01 to t = 10) against the density-matrix
            RK4 of the same system: trajectory-steps/s and the ensemble error.
copy      : plain streaming copy of known size (complex128 -> complex64 hi/lo
            split: 16 B read + 16 B written per element), the calibration run
            for the HBM byte counters (FETCH_SIZE / WRITE_SIZE).
"""
import ctypes
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import quac_amd as qa  # noqa: E402
import recipes  # noqa: E402
from qbuild import build_system, populations_np, subsystem_layout  # noqa: E402


def _pops(vec, subs):
    """FP64 populations of a device DM vector (copied to FP64 first if FP32)"""
    lv = np.array([L for L, _ in subs], np.int64)
    st = np.array([s for _, s in subs], np.int64)
    out = np.zeros(len(subs))
    src = vec
    if vec.prec != qa.FP64:
        src = qa.DeviceVec(vec.n, qa.FP64)
        qa._qh(qa.lib().qh_vec_copy(vec.h, src.h))
    qa._qh(qa.lib().qh_populations(src.h, len(subs), lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                   st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def _trace(vec, N):
    src = vec
    if vec.prec != qa.FP64:
        src = qa.DeviceVec(vec.n, qa.FP64)
        qa._qh(qa.lib().qh_vec_copy(vec.h, src.h))
    re, im = ctypes.c_double(), ctypes.c_double()
    qa._qh(qa.lib().qh_trace(src.h, N, ctypes.byref(re), ctypes.byref(im)))
    return complex(re.value, im.value)


def tc512(args):
    n_inst = args.instances
    K = args.steps if args.steps != 2000 else 1000
    W = min(args.warmup, 10)
    dt = 0.01
    insts = recipes.tc_instances(n_inst, 2)
    t0 = time.perf_counter()
    systems, mats, tots, subs, x0s = [], [], [], [], []
    for p in insts:
        r = recipes.tc_recipe(*p)
        qs, _, _ = build_system(r)
        N, loc = recipes.tc_init(p[0], p[1])
        x0 = np.zeros(N * N, np.complex128)
        x0[loc] = 1.0
        systems.append(qs)
        mats.append(qs.matrix())
        tots.append(N)
        subs.append(subsystem_layout(r))
        x0s.append(x0)
    qa.synchronize()
    asm_s = time.perf_counter() - t0
    rows = sum(t * t for t in tots)
    nnz = sum(m.info()[1] for m in mats)

    # batched FP32: warmup W steps, then K timed steps
    b = qa.Batch(mats, tots, prec=qa.FP32)
    for i, x0 in enumerate(x0s):
        b.set_state(i, x0)
    b.rk4(dt, W, renorm_every=0)
    qa.synchronize()
    t = time.perf_counter()
    b.rk4(dt, K, renorm_every=100)
    qa.synchronize()
    el32 = time.perf_counter() - t
    pops32 = b.populations(subs)
    bl = b.matrix().lanes()

    # plain FP64 path, one instance at a time (same W + K steps; K timed)
    el64 = 0.0
    worst = 0.0
    for i, m in enumerate(mats):
        v = qa.DeviceVec(tots[i] ** 2, qa.FP64)
        v.set(x0s[i])
        qa.rk4_device(m, v, dt, W)
        qa.synchronize()
        t = time.perf_counter()
        qa.rk4_device(m, v, dt, K)
        qa.synchronize()
        el64 += time.perf_counter() - t
        worst = max(worst, float(np.abs(_pops(v, subs[i]) - pops32[i]).max()))
    b.destroy()
    for qs in systems:
        qs.destroy()
    v32 = n_inst * K / el32
    v64 = n_inst * K / el64
    return {"metric": "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain FP64; SpMV % HBM peak",
            "value": round(v32, 2), "unit": "instance RK4 steps/s", "n_gpus": 1, "steps": K, "warmup": W,
            "ms_per_step": round(1e3 * el32 / K, 5), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 batch vs fp64 one-at-a-time", "data": "synthetic",
            "config": {"workload": f"{n_inst} Tavis-Cummings instances (seed 2), rows {min(t*t for t in tots)}.."
                                   f"{max(t*t for t in tots)}, total {rows} rows {nnz} nnz, RK4 dt={dt}",
                       "batch_lanes": bl},
            "fp64_one_at_a_time_steps_per_s": round(v64, 2), "batched_over_plain": round(v32 / v64, 3),
            "max_pop_err_fp32_batch_vs_fp64": worst, "assembly_s": round(asm_s, 3)}


def sweep_eager(args):
    """the 512-instance sweep's FP32 stage kernels launched one by one (no step
    graph), for counter collection: the batch is built and the variant chosen
    as in the headline run (or --fp32-variant), W steps, then K timed steps with
    an event pair around every stage launch"""
    insts = recipes.tc_instances(args.instances, 2)
    mats, tots, x0s, systems = [], [], [], []
    for p in insts:
        qs, _, _ = build_system(recipes.tc_recipe(*p))
        N, loc = recipes.tc_init(p[0], p[1])
        x0 = np.zeros(N * N, np.complex128)
        x0[loc] = 1.0
        systems.append(qs)
        mats.append(qs.matrix())
        tots.append(N)
        x0s.append(x0)
    b = qa.Batch(mats, tots, prec=qa.FP32)
    for i, x0 in enumerate(x0s):
        b.set_state(i, x0)
    bA, bx = b.matrix(), b.state()
    if args.fp32_variant:
        bA.set_lanes(0, args.fp32_variant)
    qa.rk4_prepare(bA, bx, 0.01)
    qa.rk4_device(bA, bx, 0.01, max(args.warmup, 1), timed=True)
    ms = qa.rk4_device(bA, bx, 0.01, args.steps, timed=True)
    lanes = bA.lanes()
    b.destroy()
    for qs in systems:
        qs.destroy()
    return {"metric": "sweep stage kernel, eager launches", "value": round(ms, 5), "unit": "ms per stage launch",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False,
            "config": {"workload": f"{args.instances} Tavis-Cummings instances (seed 2), one FP32 batch"},
            "lanes": lanes}


def _td_ell_entries(qs, nterms, rows):
    """entries of the sliced 64-row form of the merged A_k (exact zeros left
    out, every row of a slice padded to the slice's longest)"""
    per_row = np.zeros(rows, np.int64)
    for k in range(nterms):
        rp, _, val = qs.td_matrix(k).export()
        owner = np.repeat(np.arange(rows), np.diff(rp))
        per_row += np.bincount(owner[val != 0], minlength=rows)
    pad = (-rows) % 64
    w = np.concatenate([per_row, np.zeros(pad, np.int64)]).reshape(-1, 64).max(axis=1)
    return int(per_row.sum()), int(64 * w.sum())


def timedep(args):
    """Time-dependent terms (qsystem.c:1288-1330) on the 10-qubit dissipative
    chain: its time-independent generator plus one driven sigma_x term per qubit,
    H_d(t) = sum_q 0.2 cos(w_q t) sigma_x^q (w_q = 1 + 0.1 q), fixed-step RK4
    through time_step_sys (dt = 0.005).  Every stage is A's stage kernel on its
    autotuned sliced layout summing, per row, also the merged A_k entries (a
    column and a 2-byte (term, value) dictionary index each, sliced 64-row
    entry-major), the envelopes evaluated on the host once per stage.  FP64 and
    the FP32 hi+lo state, populations of both after the same steps; for
    comparison, FP64 with the A_k entries in a pass of their own after A's
    kernel, and FP64 with every stage one fused kernel over A's CSR and the
    merged CSR carrying the term index and value per entry."""
    import math
    nq = 10
    recipe = recipes.spin_chain_recipe(nq, seed=1)
    N, loc = recipes.spin_chain_init(nq)
    K = args.steps if args.steps != 2000 else 200
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    t0 = time.perf_counter()
    qs, subs, op = build_system(recipe, construct=False)
    for q in range(nq):
        w = 1.0 + 0.1 * q
        qs.add_ham_term_time_dep(0.2, (lambda t, w=w: math.cos(w * t)), op((q, "sx")))
    qs.construct_matrix()
    qa.synchronize()
    asm_s = time.perf_counter() - t0
    rows, nnz = qs.matrix().info()
    td_nnz = sum(qs.td_matrix(k).info()[1] for k in range(nq))
    td_nonzero, td_ell = _td_ell_entries(qs, nq, rows)
    out = {}
    for prec, cfg in (("fp64", (4096, 1, 0)), ("fp32", (4096, 1, 0)), ("fp64_separate_pass", (4096, 1, 1)),
                      ("fp64_fused_csr", (0, 0, 0))):
        qa.td_config(*cfg)
        qa.set_option("quac_precision", prec[:4])
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        qs.time_step(x, 0.0, 0.01, 0.005, 10)        # set-up: merged structure, stage buffers
        qa.synchronize()
        t = time.perf_counter()
        steps, _ = qs.time_step(x, 0.01, 0.01 + K * 0.005, 0.005, 10 ** 7)
        qa.synchronize()
        el = time.perf_counter() - t
        out[prec] = (steps / el, populations_np(x.get(), recipe))
        x.destroy()
    qa.td_config()
    lanes = qs.matrix().lanes()
    qa.set_option("quac_precision", "fp64")
    # the default integrator (adaptive RK3BS, the reference's TS defaults) over
    # the same interval: every trial step four RHS evaluations through
    # qh_spmv_td, the combinations and the error norm on the device
    qa.set_option("ts_rk_type", "3bs")
    qa.set_option("ts_adapt_type", "basic")
    x = qs.create_qvec()
    x.add_loc(1.0, loc)
    qs.time_step(x, 0.0, 0.01, 0.005, 10)
    qa.synchronize()
    t = time.perf_counter()
    rk3_steps, _ = qs.time_step(x, 0.01, 0.01 + K * 0.005, 0.005, 10 ** 7)
    qa.synchronize()
    rk3_s = time.perf_counter() - t
    rk3_err = float(np.abs(populations_np(x.get(), recipe) - out["fp64"][1]).max())
    x.destroy()
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    qs.destroy()
    # bytes the A_k entries add to an FP64 stage summed in A's kernel: the sliced
    # form's column and index per entry and the slice offsets and widths (as a
    # pass of their own, K and the next stage input read and written again and
    # the stage input gathered once more: + 5 x 16 B per row)
    S = 16
    nsl = (rows + 63) // 64
    pass_bytes = td_ell * (4 + 2) + nsl * (8 + 4)
    return {"metric": "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain FP64; SpMV % HBM peak",
            "value": round(out["fp64"][0], 2), "unit": "RK4 steps/s (FP64, time-dependent terms)", "n_gpus": 1,
            "steps": K, "warmup": 2, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp64", "data": "synthetic",
            "config": {"workload": f"10-qubit dissipative chain + 10 driven sigma_x terms (cos envelopes), "
                                   f"{rows} rows, {nnz} nnz + {td_nnz} time-dependent entries, RK4 dt=0.005"},
            "fp32_steps_per_s": round(out["fp32"][0], 2),
            "fp64_steps_per_s_separate_pass": round(out["fp64_separate_pass"][0], 2),
            "fp64_steps_per_s_fused_csr_form": round(out["fp64_fused_csr"][0], 2),
            "max_pop_err_fp32_vs_fp64": float(np.abs(out["fp32"][1] - out["fp64"][1]).max()),
            "max_pop_err_fused_vs_sliced_fp64": float(np.abs(out["fp64_fused_csr"][1] - out["fp64"][1]).max()),
            "default_integrator_rk3bs": {"steps": int(rk3_steps), "seconds": round(rk3_s, 4),
                                         "steps_per_s": round(rk3_steps / rk3_s, 2),
                                         "max_pop_diff_vs_rk4": rk3_err},
            "a_variants_fp64_fp32": list(lanes), "td_nonzero_entries": td_nonzero, "td_sliced_form_entries": td_ell,
            "td_entry_bytes_per_stage": pass_bytes, "assembly_s": round(asm_s, 3)}


def circuit12(args):
    """the circuit through QuaC's circuit API: add_gate_to_circuit_sys (run time
    0.1 per gate), schedule_circuit_layers, apply_circuit_to_sys, and one
    time_step_sys call whose gate events fire between the fixed RK4 steps"""
    nq = 12
    N = 2 ** nq
    recipe = recipes.noisy_circuit_recipe(nq)
    gates = recipes.random_circuit(nq, 300, 3)
    subs = subsystem_layout(recipe)
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    res = {}
    t0 = time.perf_counter()
    qs, _, _ = build_system(recipe)
    qa.synchronize()
    asm_s = time.perf_counter() - t0
    nl = 0
    for prec in (qa.FP64, qa.FP32):
        qa.set_option("quac_precision", "fp32" if prec == qa.FP32 else "fp64")
        qa.set_option("quac_renorm", "10" if prec == qa.FP32 else "0")
        # warm-up: autotune and graph capture of this precision, outside the timed region
        w = qs.create_qvec()
        w.add_loc(1.0, 0)
        qs.time_step(w, 0.0, 0.1, 0.01, 100)
        w.destroy()
        circ = qa.Circuit(len(gates))
        for (name, qb, th) in gates:
            circ.add_gate(1.0, name, list(qb), theta=th, run_time=0.1)
        qs.schedule(circ)
        nl = circ.num_layers
        qs.apply_circuit(circ, 0.0)
        x = qs.create_qvec()
        x.add_loc(1.0, 0)
        qa.synchronize()
        t = time.perf_counter()
        steps, _ = qs.time_step(x, 0.9, 1.0 + 0.1 * nl, 0.01, 10 ** 7)
        qa.synchronize()
        el = time.perf_counter() - t
        v = x.data()
        res[prec] = (el, steps, _pops(v, subs), _trace(v, N))
        x.destroy()
        circ.destroy()
    qa.set_option("quac_precision", "fp64")
    qa.set_option("quac_renorm", "0")
    qs.destroy()
    el64, steps, p64, t64 = res[qa.FP64]
    el32, _, p32, t32 = res[qa.FP32]
    return {"metric": "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain FP64; SpMV % HBM peak",
            "value": round(steps / el32, 2), "unit": "RK4 steps/s (with gates)", "n_gpus": 1, "steps": steps,
            "warmup": 10, "ms_per_step": round(1e3 * el32 / steps, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32 bulk vs fp64", "data": "synthetic",
            "config": {"workload": f"12-qubit noisy circuit, rho 4096x4096, {len(gates)} gates (seed 3) in "
                                   f"{nl} layers through schedule_circuit_layers / apply_circuit_to_sys / "
                                   f"time_step_sys, RK4 10 x 0.01 between layers, T1 1e-3, dephasing 2e-3"},
            "fp64_steps_per_s": round(steps / el64, 2), "fp32_over_fp64": round(el64 / el32, 3),
            "wall_s_fp32": round(el32, 4), "wall_s_fp64": round(el64, 4), "assembly_s": round(asm_s, 4),
            "max_pop_err_fp32_vs_fp64": float(np.abs(p32 - p64).max()),
            "trace_fp32": abs(t32), "trace_fp64": abs(t64), "populations_fp64": [round(x, 10) for x in p64]}


def jc(args):
    recipe = recipes.jc_recipe()
    N, loc = recipes.jc_init()
    qs, _, _ = build_system(recipe)
    A = qs.matrix()
    subs = subsystem_layout(recipe)
    x0 = np.zeros(N * N, np.complex128)
    x0[loc] = 1.0
    out = {}
    for prec in (qa.FP64, qa.FP32):
        v = qa.DeviceVec(N * N, prec)
        v.set(x0)
        qa.rk4_device(A, v, 0.01, 2)
        v.set(x0)
        qa.synchronize()
        t = time.perf_counter()
        trace_pts = []
        for chunk in range(100):   # t = 100: 10 000 steps, populations every 100 steps
            qa.rk4_device(A, v, 0.01, 100)
            trace_pts.append(_pops(v, subs))
        qa.synchronize()
        out[prec] = (time.perf_counter() - t, np.array(trace_pts))
    # the default integrator of time_step_sys (adaptive RK3BS, FP64) to t = 100:
    # one workgroup with the controller on the device, and the host-driven loop
    # (-quac_small_path 0: one error read-back per trial step)
    adaptive = {}
    for mode in ("one_workgroup", "host_loop"):
        qa.set_option("ts_rk_type", "3bs")
        qa.set_option("ts_adapt_type", "basic")
        x = qs.create_qvec()
        x.add_loc(1.0, loc)
        qa.set_option("quac_small_path", "0" if mode == "host_loop" else "1")
        qa.synchronize()
        t = time.perf_counter()
        steps, tend = qs.time_step(x, 0.0, 100.0, 0.01, 10 ** 7)
        qa.synchronize()
        el = time.perf_counter() - t
        qa.set_option("quac_small_path", "1")
        adaptive[mode] = {"steps": steps, "seconds": round(el, 5), "steps_per_s": round(steps / el, 1),
                          "populations": [float(p) for p in populations_np(x.get(), recipe)]}
    qs.destroy()
    el64, p64 = out[qa.FP64]
    el32, p32 = out[qa.FP32]
    return {"metric": "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain FP64; SpMV % HBM peak",
            "value": round(10000 / el64, 2), "unit": "RK4 steps/s (FP64, incl. 100 population readbacks)",
            "n_gpus": 1, "steps": 10000, "warmup": 2, "ms_per_step": round(1e3 * el64 / 10000, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
            "data": "synthetic",
            "config": {"workload": "Jaynes-Cummings 2x10 (400-row superoperator), RK4 dt=0.01 to t=100"},
            "fp32_steps_per_s": round(10000 / el32, 2),
            "max_pop_err_fp32_vs_fp64": float(np.abs(p32 - p64).max()),
            "final_populations_fp64": [float(x) for x in p64[-1]],
            "default_integrator_rk3bs_to_t100": adaptive}


def copy(args):
    n = 1 << 24                        # 16 777 216 elements: 256 MiB read, 256 MiB written
    reps = 20
    src = qa.DeviceVec(n, qa.FP64)
    dst = qa.DeviceVec(n, qa.FP32)
    qa._qh(qa.lib().qh_vec_copy(src.h, dst.h))
    qa.synchronize()
    with qa.StreamTimer() as tm:
        for _ in range(reps):
            qa._qh(qa.lib().qh_vec_copy(src.h, dst.h))
    ms = tm.ms / reps
    rd = wr = 16 * n
    return {"metric": "HBM copy calibration", "value": round((rd + wr) / (ms * 1e-3) / 1e9, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(ms, 5), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 -> f32 pair", "data": "zeros",
            "config": {"workload": "k_split_d2 over 16 777 216 complex128 elements"},
            "read_bytes_per_launch": rd, "write_bytes_per_launch": wr,
            "stream_copy_read_gbs_1gib": [round(v, 1) for v in qa.bench_stream(1 << 30, 10)]}


def kerr(args):
    """Learned warm start of the steady-state GMRES solve: 3 driven Kerr cavities,
    network trained on the device solver's own steady states.  Two sweeps: the
    wide box (every parameter over its whole range; value = cold / warm
    iterations there) and a local sweep (every parameter within 5 % of a centre
    set).  Warm = GMRES started from the network's prediction, with plain GMRES
    from zero finishing any solve that does not reach the tolerance; augmented =
    the prediction added to each cycle's search space (qh_gmres_aug)."""
    from quac_amd.warmstart import WarmStart
    L = args.levels
    N = L ** 3
    tol_train, tol = 1e-10, 1e-8
    log = (lambda m: print(m, file=sys.stderr, flush=True))

    def solve(p, x0=None, t=tol, maxit=3000, aug=None):
        qs, _, _ = build_system(recipes.kerr_recipe(p, levels=L))
        x = qs.create_qvec()
        if x0 is not None:
            x.set(x0)
        u = None
        if aug is not None:
            u = qs.create_qvec()
            u.set(aug)
        t0 = time.perf_counter()
        it, res = qs.steady_state(x, restart=100, tol=t, maxit=maxit, precond="hbasis", aug=u)
        el = time.perf_counter() - t0
        sol = x.get()
        qs.destroy()
        return sol, it, res, el

    def leg(name, train, test, rank, epochs, diagnose):
        t0 = time.perf_counter()
        feats, sols, tr_its, skipped = [], [], [], 0
        for i, p in enumerate(train):
            sol, it, res, _ = solve(p, t=tol_train)
            if res > tol_train:
                skipped += 1
                continue
            feats.append(recipes.kerr_features(p))
            sols.append(sol)
            tr_its.append(it)
            if i % 200 == 0:
                log(f"{name}: train solve {i}/{len(train)}: {it} iterations")
        gen_s = time.perf_counter() - t0
        ws = WarmStart(N, rank=rank, seed=0)
        stats = ws.fit(np.array(feats), np.array(sols), epochs=epochs, log=log)
        del sols
        rows = []
        for i, p in enumerate(test):
            f = recipes.kerr_features(p)
            sol_c, it_c, res_c, t_c = solve(p)
            x0 = ws.predict([f])[0]
            sol_w, it_w, res_w, t_w = solve(p, x0=x0)
            fallback = 0
            if res_w > tol:                      # plain GMRES from zero finishes the solve
                _, it_f, res_w, t_f = solve(p)
                it_w += it_f
                t_w += t_f
                fallback = 1
            _, it_a, res_a, t_a = solve(p, aug=x0)
            nrm = np.linalg.norm(sol_c)
            r = {"cold": it_c, "warm": it_w, "aug": it_a, "res_cold": res_c, "res_warm": res_w, "res_aug": res_a,
                 "t_cold": t_c, "t_warm": t_w, "t_aug": t_a, "fallback": fallback,
                 "err_warm": float(np.linalg.norm(x0 - sol_c) / nrm)}
            if diagnose:
                xp = ws.project([sol_c])[0]
                r["basis_best"] = solve(p, x0=xp)[1]
                r["err_basis"] = float(np.linalg.norm(xp - sol_c) / nrm)
            rows.append(r)
            if i % 20 == 0:
                log(f"{name}: test {i}: cold {it_c} warm {it_w} augmented {it_a} err {r['err_warm']:.2e}")
        ws.destroy()
        cold = np.array([r["cold"] for r in rows], float)
        warm = np.array([r["warm"] for r in rows], float)
        augm = np.array([r["aug"] for r in rows], float)
        out = {"training_sets": len(train), "test_sets": len(test), "rank": stats["rank"], "epochs": epochs,
               "cold_mean": float(cold.mean()), "warm_mean": float(warm.mean()), "augmented_mean": float(augm.mean()),
               "warm_ratio": round(float(cold.sum() / warm.sum()), 4),
               "augmented_ratio": round(float(cold.sum() / augm.sum()), 4),
               "warm_fewer_sets": int(np.sum(warm < cold)), "warm_more_sets": int(np.sum(warm > cold)),
               "augmented_never_more": bool(np.all(augm <= cold)),
               "fallbacks": int(sum(r["fallback"] for r in rows)),
               "max_res": float(max(max(r["res_warm"], r["res_aug"], r["res_cold"]) for r in rows)),
               "warm_start_error_mean": float(np.mean([r["err_warm"] for r in rows])),
               "time_s": {"cold_mean": float(np.mean([r["t_cold"] for r in rows])),
                          "warm_mean": float(np.mean([r["t_warm"] for r in rows])),
                          "augmented_mean": float(np.mean([r["t_aug"] for r in rows])),
                          "training_data": round(gen_s, 2), "basis": round(stats["pca_s"], 2),
                          "network_training": round(stats["train_s"], 2)},
               "training": {"solved": len(feats), "skipped": skipped, "mean_iterations": float(np.mean(tr_its)),
                            "basis_err": stats["train_basis_err"], "network_err": stats["train_net_err"],
                            "loss_first_last": [stats["losses"][0][1], stats["losses"][-1][1]]}}
        if diagnose:
            out["basis_best_warm_mean"] = float(np.mean([r["basis_best"] for r in rows]))
            out["basis_best_error_mean"] = float(np.mean([r["err_basis"] for r in rows]))
        return out

    wide = leg("wide", recipes.kerr_sweep_params(4, args.train), recipes.kerr_sweep_params(5, args.test),
               args.rank, args.epochs, True)
    local = leg("local", recipes.kerr_local_params(4, args.local_train), recipes.kerr_local_params(5, args.local_test),
                args.local_rank, args.epochs, False)
    return {"metric": "GMRES iterations to a 1e-8 residual, cold / learned warm start (held-out sets)",
            "value": wide["warm_ratio"], "unit": "x fewer iterations", "n_gpus": 1, "steps": args.test,
            "warmup": 0, "ms_per_step": round(1e3 * wide["time_s"]["warm_mean"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp64 GMRES, fp32 network", "data": "synthetic",
            "config": {"workload": f"3 driven Kerr cavities x {L} levels (superoperator {N * N} rows), "
                                   f"{args.train} training / {args.test} held-out parameter sets (seeds 4 / 5) over "
                                   f"the full parameter box, network 10 -> 3x256 -> {2 * wide['rank']}, "
                                   f"{args.epochs} epochs, GMRES(100) with the Hamiltonian-eigenbasis preconditioner"},
            "warm_start_use": "value: GMRES started from the prediction (plain GMRES from zero finishes a solve "
                              "that misses the tolerance); augmented: the prediction added to each cycle's "
                              "search space",
            "wide": wide,
            "local": dict(local, sweep="every parameter within 5 % of a centre set (seeds 4 / 5)")}


def mcwf(args):
    M, dt, T = 10000, 0.01, 10.0
    steps = int(round(T / dt))
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    recipe = recipes.jc_recipe()
    N, loc = recipes.jc_init()
    # density-matrix reference on the device (FP64 drop-in path)
    qd, _, _ = build_system(recipe)
    xd = qd.create_qvec()
    xd.add_loc(1.0, loc)
    qd.time_step(xd, 0.0, T, dt, 10 ** 7)
    subs = subsystem_layout(recipe)
    ref = _pops(xd.data(), subs)
    qd.destroy()
    # trajectories
    def make(m, seed):
        qs = qa.System()
        ops = []
        for (kind, L) in recipe["subs"]:
            ops.append(qs.create_op(L))
        for (k, a, o) in recipe["terms"]:
            plist = [{"lower": ops[s_], "raise": ops[s_].dag, "n": ops[s_].n}[nm] for (s_, nm) in o]
            (qs.add_ham_term if k == "ham" else qs.add_lin_term)(a, *plist)
        qs.use_mcwf_solver(m, seed)
        qs.construct_matrix()
        x = qs.create_qvec()
        x.add(1.0, loc // N)      # |e,0> as a wavefunction: Hilbert index of the DM diagonal entry
        return qs, ops, x

    # warm-up on a separate ensemble of the same size to t = 2 (jumps occur):
    # the process's first launches of the stepping and jump kernels and its
    # first allocation of an ensemble of this size (about 6.5 ms in the HIP
    # runtime) stay out of the timed call
    qw, _, xw = make(M, 9)
    qw.time_step(xw, 0.0, 2.0, dt, 10 ** 7)
    qw.destroy()
    qs, ops, x = make(M, 5)
    qa.synchronize()
    # timed: the whole time_step call (jump operators, thresholds, trajectories)
    t0 = time.perf_counter()
    qs.time_step(x, 0.0, T, dt, 10 ** 7)
    qa.synchronize()
    el = time.perf_counter() - t0
    got = [x.expectation(ops[0].n).real, x.expectation(ops[1].n).real]
    qs.destroy()
    return {"metric": "MCWF trajectory RK4 steps/s (ensemble average vs density matrix)",
            "value": round(M * steps / el, 1), "unit": "trajectory RK4 steps/s", "n_gpus": 1, "steps": steps,
            "warmup": 200, "ms_per_step": round(1e3 * el / steps, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp64", "data": "synthetic",
            "config": {"workload": f"Jaynes-Cummings 2x10 wavefunctions, {M} trajectories, RK4 dt={dt} to t={T}"},
            "populations_mcwf": got, "populations_dm": [float(v) for v in ref],
            "max_abs_err": float(max(abs(a - b) for a, b in zip(got, ref))),
            "sem_bound": 4 * 4.5 / math.sqrt(M)}
