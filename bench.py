"""Benchmark of the metric "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain
FP64; SpMV % HBM peak".  One JSON line on stdout.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload ...]

Headline (default workload):
  * value: the batched parameter sweep of 512 Tavis-Cummings instances of uneven
    size (seed 2; 100 .. 230 400 superoperator rows), RK4 dt = 0.01, advanced as
    one FP32 batch (block-diagonal sliced matrix, fused stage kernels, hipGraph
    replay): instance RK4 steps per second over K timed steps, beside the plain
    FP64 path run one instance at a time over the same K steps, with the maximum
    population error of the batch against that FP64 path after the same W + K
    steps.
  * roofline: the 10-qubit dissipative spin chain (rho 1024 x 1024; superoperator
    1 048 576 rows, 24 641 536 nnz), RK4 dt = 0.005, FP32 and FP64 over K steps:
    the FP32 fused SpMV + RK4 stage kernel's launch duration against the bytes its
    format has to move (roofline.achieved / frac), the HBM traffic counted by
    rocprofv3 (roofline.traffic, from the committed counter summary), and the CSR
    model's bytes under roofline.csr_model.  GPU FP64 and FP32 populations are
    compared with the CPU oracle after the same W + K steps.
  * hbm_copy: the achievable bandwidth of this GPU (a one-pass streaming copy of
    1 GiB, and the same read alone), beside the 8 TB/s peak.
  * timedep: the same chain with ten driven sigma_x terms through time_step_sys
    (300 fixed RK4 steps of 0.005, FP64 and FP32; FP64 also with the A_k entries
    as a pass of their own and in the fused CSR form; the default adaptive
    integrator over the same interval), as --workload timedep reports it.
  * cpu_baseline: the C oracle's FP64 RK4 (same CSR, row-ordered MatMult, same RK4)
    on this host's cores, on a bounded sample of the sweep (every 16th instance),
    median of 3 runs at 1 thread and at all cores of the GPU's CPU share.
Before any timed region the autotuner runs and the step graph is captured and
instantiated (qh_rk4_prepare), whatever W is.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak (GB/s)
METRIC = "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain FP64; SpMV % HBM peak"
CPU_SHARE = 16         # host threads of one GPU's share of the box


def spmv_bytes(nnz, rows, S):
    """CSR-model bytes of one SpMV: nnz*(S+4) + 4*(rows+1) + 2*rows*S"""
    return nnz * (S + 4) + 4 * (rows + 1) + 2 * rows * S


def rk4_step_bytes(nnz, rows, S):
    """CSR-model bytes of one fused RK4 step: 4 matrix streams + 14 vector passes"""
    return 4 * (nnz * (S + 4) + 4 * (rows + 1)) + 14 * rows * S


def kernel_name(variant, real):
    """name of the stage kernel a lanes-per-row variant code selects"""
    if 128 <= variant <= 136:
        u, vv, pp, nt = {128: (2, 0, 0, 0), 129: (4, 0, 0, 0), 130: (8, 0, 0, 0), 131: (4, 1, 0, 0),
                         132: (8, 1, 0, 0), 133: (2, 0, 1, 0), 134: (4, 0, 1, 0), 135: (4, 0, 0, 1),
                         136: (4, 1, 0, 1)}[variant]
        tf = ("false", "true")
        return f"k_ks_stage<{real}, MODE, {u}, {tf[vv]}, {tf[pp]}, {tf[nt]}> (sliced layout)"
    if variant in (140, 141, 142, 143, 144):
        u, w, nt = {140: (2, "false", "false"), 141: (4, "false", "false"), 142: (2, "true", "false"),
                    143: (2, "false", "true"), 144: (2, "true", "true")}[variant]
        return f"k_ks2_stage<MODE, {u}, {w}, {nt}> (row-pair sliced layout)"
    return f"k_rk_stage<{real}, {variant}> (CSR, {variant} lanes per row)"


def format_step_bytes(A, rows):
    """bytes the FP32 sliced layout of A has to move per RK4 step: 4 x (stored
    values + slot metadata) + 20 vector passes of 8 B per row (stages 1-3: the
    gathered stage input, the state's high word for Y, the K and Y writes;
    stage 4: the gathered input, the state's high and low words, K1..K3, and the
    new high and low words written); and the layout statistics"""
    lanes = A.lanes()
    if lanes[1] in (140, 141, 142, 143, 144):  # row-pair layout: 40-byte slot metadata, values per lane pair
        layout = dict(A.pair_layout_info(), kind="row-pair (128 rows per wavefront)")
        meta = layout["meta_bytes"]
        vals = layout["stored_pairs"] * 16 + layout["stored_real_pairs"] * 8 + layout["stored_real_singles"] * 4
    else:
        layout = dict(A.layout_info(), kind="one row per lane (64 rows per wavefront)")
        meta = layout["slots"] * 32 + 12 * layout["slices"]
        vals = layout["stored_vals"] * 8 + layout["stored_reals"] * 4
    return 4 * (vals + meta) + 20 * rows * 8, layout


def read_traffic(path):
    """HBM bytes per stage launch from a committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by its copy-kernel factor); None when absent"""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return n, max(1, min(n, CPU_SHARE))


def roofline_obj(kname, workload, fmt_bytes, launch_ms, ev_ms, traffic, traffic_src, csr_bytes):
    ach = fmt_bytes / (launch_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname, "workload": workload,
           "bytes_per_launch": fmt_bytes, "bytes_basis": "bytes the sliced format moves per launch (stored values "
                                                         "+ slot metadata + 20/4 vector passes of 8 B per row)",
           "launch_ms": round(launch_ms, 5), "launch_ms_per_kernel_events": round(ev_ms, 5),
           "csr_model": {"bytes_per_launch": csr_bytes,
                         "gbs": round(csr_bytes / (launch_ms * 1e-3) / 1e9, 1),
                         "over_peak": round(csr_bytes / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    if traffic:
        tg = traffic / (launch_ms * 1e-3) / 1e9
        out.update(traffic_source=traffic_src, traffic_over_format_bytes=round(traffic / fmt_bytes, 3),
                   traffic_gbs=round(tg, 1), traffic_frac=round(tg / HBM_PEAK_GBS, 4))
    return out


# ---------------------------------------------------------------- sweep (value)
def run_sweep(qa, recipes, build_system, subsystem_layout, K, W, n_inst, renorm):
    from bench_workloads import _pops
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

    b = qa.Batch(mats, tots, prec=qa.FP32)
    for i, x0 in enumerate(x0s):
        b.set_state(i, x0)
    bA, bx = b.matrix(), b.state()
    qa.rk4_prepare(bA, bx, dt)                  # autotune + graph capture, untimed
    b.rk4(dt, W, renorm_every=0)
    qa.synchronize()
    t = time.perf_counter()
    with qa.StreamTimer() as tm32:
        b.rk4(dt, K, renorm_every=renorm)
    qa.synchronize()
    el32 = time.perf_counter() - t
    pops32 = b.populations(subs)
    lanes = bA.lanes()
    fmt_step, layout = format_step_bytes(bA, rows)
    nprof = min(20, max(K, 1))
    ev_ms = qa.rk4_device(bA, bx, dt, nprof, timed=True)   # per-launch events (state advanced past the check)
    launch_ms = tm32.ms / (4 * K) if K > 0 else ev_ms

    el64 = 0.0
    worst = 0.0
    for i, m in enumerate(mats):
        v = qa.DeviceVec(tots[i] ** 2, qa.FP64)
        v.set(x0s[i])
        qa.rk4_prepare(m, v, dt)
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
    traffic, src = read_traffic(os.path.join(ROOT, "profiles", "k_ks2_stage_tc512_traffic.txt"))
    rf = roofline_obj(kernel_name(lanes[1], "float") + " (mean of the 4 fused RK4 stage launches)",
                      f"{n_inst} Tavis-Cummings instances, one block-diagonal batch", fmt_step / 4, launch_ms,
                      ev_ms, traffic, src, rk4_step_bytes(nnz, rows, 8) / 4)
    return {"v32": v32, "v64": v64, "el32": el32, "el64": el64, "worst": worst, "asm_s": asm_s, "rows": rows,
            "nnz": nnz, "rows_min": min(t * t for t in tots), "rows_max": max(t * t for t in tots), "lanes": lanes,
            "roofline": rf, "layout": layout}


# --------------------------------------------------------- 10-qubit chain leg
def run_chain(qa, recipes, build_system, subsystem_layout, K, W, renorm, nq=10, variant=0):
    from bench_workloads import _pops, _trace
    dt = 0.005
    recipe = recipes.spin_chain_recipe(nq, seed=1)
    N, loc = recipes.spin_chain_init(nq)
    rows = N * N
    t0 = time.perf_counter()
    qs, _, _ = build_system(recipe)
    qa.synchronize()
    asm_s = time.perf_counter() - t0
    A = qs.matrix()
    if variant:
        A.set_lanes(0, variant)
    rows_, nnz = A.info()
    assert rows_ == rows
    subs = subsystem_layout(recipe)
    x0 = np.zeros(rows, np.complex128)
    x0[loc] = 1.0
    res = {}
    for prec in (qa.FP64, qa.FP32):
        rn = renorm if prec == qa.FP32 else 0
        v = qa.DeviceVec(rows, prec)
        v.set(x0)
        qa.rk4_prepare(A, v, dt, tot=N, renorm_every=rn)   # autotune + graph capture, untimed
        qa.rk4_device(A, v, dt, W, tot=N, renorm_every=rn)
        qa.synchronize()
        t = time.perf_counter()
        with qa.StreamTimer() as tm:
            qa.rk4_device(A, v, dt, K, tot=N, renorm_every=rn)
        qa.synchronize()
        res[prec] = (time.perf_counter() - t, tm.ms, _pops(v, subs), _trace(v, N).real)
    nprof = min(50, max(K, 1))
    ev = {}
    for prec in (qa.FP32, qa.FP64):
        v = qa.DeviceVec(rows, prec)
        v.set(x0)
        ev[prec] = qa.rk4_device(A, v, dt, nprof, tot=N, timed=True)
    spmv32 = qa.bench_spmv(A, qa.FP32, 30)
    spmv64 = qa.bench_spmv(A, qa.FP64, 30)
    lanes = A.lanes()
    fmt_step, layout = format_step_bytes(A, rows)
    qs.destroy()
    el64, dev64, p64, tr64 = res[qa.FP64]
    el32, dev32, p32, tr32 = res[qa.FP32]
    launch32 = dev32 / (4 * K) if K > 0 else ev[qa.FP32]
    launch64 = dev64 / (4 * K) if K > 0 else ev[qa.FP64]
    traffic, src = read_traffic(os.path.join(ROOT, "profiles", "k_ks2_stage_chain10_traffic.txt"))
    rf = roofline_obj(kernel_name(lanes[1], "float") + " (mean of the 4 fused RK4 stage launches)",
                      f"{nq}-qubit dissipative spin chain, superoperator {rows} rows {nnz} nnz, FP32",
                      fmt_step / 4, launch32, ev[qa.FP32], traffic, src, rk4_step_bytes(nnz, rows, 8) / 4)
    sp32 = spmv_bytes(nnz, rows, 8) / (spmv32 * 1e-3) / 1e9
    sp64 = spmv_bytes(nnz, rows, 16) / (spmv64 * 1e-3) / 1e9
    return {
        "recipe": recipe, "x0": x0, "N": N, "dt": dt, "pops64": p64, "pops32": p32, "roofline": rf,
        "out": {"workload": f"{nq}-qubit dissipative spin chain, rho {N}x{N}, superoperator {rows} rows {nnz} nnz, "
                            f"RK4 dt={dt}, h_i,J_i~U[0.5,1.5] seed 1, gamma1=1e-3, gamma2=5e-3",
                "fp32_steps_per_s": round(K / el32, 2), "fp64_steps_per_s": round(K / el64, 2),
                "fp32_over_fp64": round(el64 / el32, 3), "renorm_every_fp32": renorm,
                "max_pop_err_fp32_vs_fp64": float(np.abs(p32 - p64).max()), "trace_fp32": tr32, "trace_fp64": tr64,
                "fp64_stage": {"kernel": kernel_name(lanes[0], "double"), "launch_ms": round(launch64, 5),
                               "launch_ms_per_kernel_events": round(ev[qa.FP64], 5),
                               "csr_model_gbs": round(rk4_step_bytes(nnz, rows, 16) / 4 / (launch64 * 1e-3) / 1e9,
                                                      1)},
                "spmv": {"fp32_ms": round(spmv32, 5), "fp64_ms": round(spmv64, 5),
                         "fp32_csr_model_gbs": round(sp32, 1), "fp64_csr_model_gbs": round(sp64, 1),
                         "lanes_per_row": lanes},
                "assembly_s": round(asm_s, 4), "sliced_layout": layout}}


def measure_copy(qa):
    """achievable HBM bandwidth: one-pass streaming copy and read kernels over
    1 GiB buffers (qh_bench_stream), beside hipMemcpyAsync device to device of
    2^26 complex128 (1 GiB read + 1 GiB written)"""
    copy_gbs, read_gbs = qa.bench_stream(1 << 30, 10)
    n = 1 << 26
    src = qa.DeviceVec(n, qa.FP64)
    dst = qa.DeviceVec(n, qa.FP64)
    qa._qh(qa.lib().qh_vec_copy(src.h, dst.h))
    qa.synchronize()
    reps = 10
    with qa.StreamTimer() as tm:
        for _ in range(reps):
            qa._qh(qa.lib().qh_vec_copy(src.h, dst.h))
    ms = tm.ms / reps
    gbs = 2 * 16 * n / (ms * 1e-3) / 1e9
    return {"gbs": round(copy_gbs, 1), "frac_of_peak": round(copy_gbs / HBM_PEAK_GBS, 4),
            "read_gbs": round(read_gbs, 1), "bytes_per_copy": 2 << 30,
            "kind": "streaming copy kernel, 16 bytes per thread in one pass, 1 GiB read + 1 GiB written; "
                    "read_gbs: the same read alone",
            "hipmemcpy_d2d_gbs": round(gbs, 1)}


# ------------------------------------------------------------------- CPU legs
def cpu_sweep(recipes, n_inst, every, budget_s):
    """oracle FP64 RK4 over every `every`-th instance of the sweep, one instance at
    a time: instance-steps/s, median of 3 runs at 1 thread and at all threads"""
    import oracle
    nproc, allc = host_threads()
    insts = recipes.tc_instances(n_inst, 2)[::every]
    ms, xs = [], []
    t = time.perf_counter()
    for p in insts:
        tot, terms = recipes.oracle_terms(recipes.tc_recipe(*p))
        ms.append(oracle.construct(tot, terms, dm=True))
        N, loc = recipes.tc_init(p[0], p[1])
        x0 = np.zeros(N * N, np.complex128)
        x0[loc] = 1.0
        xs.append(x0)
    asm = time.perf_counter() - t

    def one(steps, th):
        t = time.perf_counter()
        for m, x in zip(ms, xs):
            m.rk4(x, 0.01, steps, nthreads=th)
        return time.perf_counter() - t

    out = {}
    for th in (1, allc):
        probe = one(1, th)
        steps = max(1, int(budget_s / 3 / max(probe, 1e-4)))
        runs = [len(insts) * steps / one(steps, th) for _ in range(3)]
        out[th] = (statistics.median(runs), steps, runs)
    return out, len(insts), asm, nproc, allc


def cpu_chain(recipes, chain, W, K):
    """the oracle on the 10-qubit chain: W + K steps at all threads (populations
    against the GPU's), and steps/s at 1 thread and all threads (median of 3)"""
    import oracle
    from qbuild import populations_np
    nproc, allc = host_threads()
    tot, terms = recipes.oracle_terms(chain["recipe"])
    t = time.perf_counter()
    m = oracle.construct(tot, terms, dm=True)
    asm = time.perf_counter() - t
    t = time.perf_counter()
    x = m.rk4(chain["x0"], chain["dt"], W + K, nthreads=allc)
    el = time.perf_counter() - t
    pops = populations_np(x, chain["recipe"])
    rate = {}
    for th, steps in ((1, 2), (allc, 10)):
        runs = []
        for _ in range(3):
            t = time.perf_counter()
            m.rk4(chain["x0"], chain["dt"], steps, nthreads=th)
            runs.append(steps / (time.perf_counter() - t))
        rate[th] = statistics.median(runs)
    return {"pop_err_fp64_vs_oracle": float(np.abs(chain["pops64"] - pops).max()),
            "pop_err_fp32_vs_oracle": float(np.abs(chain["pops32"] - pops).max()),
            "oracle_steps": W + K, "oracle_s": round(el, 3), "oracle_assembly_s": round(asm, 3),
            "cpu_steps_per_s_1thread": round(rate[1], 3), "cpu_steps_per_s_allthreads": round(rate[allc], 3),
            "threads": allc}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nq", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU time per thread count of the sweep sample")
    ap.add_argument("--renorm", type=int, default=10, help="FP32 trace renormalisation period (steps)")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "tc512", "circuit12", "jc", "copy", "kerr", "mcwf", "sweep_eager",
                             "timedep"])
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--levels", type=int, default=8)
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--rank", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--local-train", type=int, default=500, help="kerr: training sets of the local sweep")
    ap.add_argument("--local-test", type=int, default=50, help="kerr: held-out sets of the local sweep")
    ap.add_argument("--local-rank", type=int, default=64, help="kerr: basis rank of the local sweep")
    ap.add_argument("--skip-sweep", action="store_true", help="only the 10-qubit chain leg (profiling runs)")
    ap.add_argument("--fp32-variant", type=int, default=0,
                    help="force the chain's FP32 stage-kernel variant (qh_mat_set_lanes code; 0 = autotune). "
                         "Counter collection perturbs the autotuner's timings, so profiled runs name the variant "
                         "the plain run selects")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build targets a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import quac_amd as qa
    import recipes
    from qbuild import build_system, subsystem_layout

    if args.workload != "headline":
        import bench_workloads
        out = getattr(bench_workloads, args.workload)(args)
        print(json.dumps(out), flush=True)
        return

    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    K, W = args.steps, args.warmup
    log = (lambda m: print(m, file=sys.stderr, flush=True))

    sw = None
    if not args.skip_sweep:
        log(f"sweep: {args.instances} Tavis-Cummings instances, K={K} W={W}")
        sw = run_sweep(qa, recipes, build_system, subsystem_layout, K, W, args.instances, 50)
    log(f"10-qubit chain leg, K={K} W={W}")
    ch = run_chain(qa, recipes, build_system, subsystem_layout, K, W, args.renorm, args.nq, args.fp32_variant)
    log("copy bandwidth")
    cp = measure_copy(qa)
    log("time-dependent terms on the chain, 300 steps")
    import bench_workloads
    td = bench_workloads.timedep(argparse.Namespace(steps=300))
    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    td_keys = ("value", "fp32_steps_per_s", "fp64_steps_per_s_separate_pass", "fp64_steps_per_s_fused_csr_form",
               "max_pop_err_fp32_vs_fp64", "max_pop_err_fused_vs_sliced_fp64", "default_integrator_rk3bs",
               "a_variants_fp64_fp32", "td_entry_bytes_per_stage")
    timedep = {"fp64_steps_per_s" if k == "value" else k: td[k] for k in td_keys}
    timedep["steps"] = td["steps"]
    timedep["workload"] = td["config"]["workload"]

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        log("CPU oracle: chain populations and rates")
        ch["out"]["oracle_check"] = cpu_chain(recipes, ch, W, K)
        log("CPU oracle: sweep sample")
        every = 16
        rates, n_s, asm, nproc, allc = cpu_sweep(recipes, args.instances, every, args.cpu_seconds)
        cpu = {"value": round(rates[allc][0], 3), "unit": "instance RK4 steps/s (FP64)", "cores": allc,
               "kind": "port", "value_1thread": round(rates[1][0], 3), "nproc": nproc,
               "runs_allthreads": [round(r, 3) for r in rates[allc][2]],
               "runs_1thread": [round(r, 3) for r in rates[1][2]],
               "sample": f"every {every}th instance of the {args.instances}-instance sweep ({n_s} instances), "
                         f"{rates[allc][1]} RK4 steps each at {allc} threads and {rates[1][1]} at 1 thread, one "
                         f"instance at a time; median of 3 runs; {cpu_model()}; oracle assembly {asm:.2f} s",
               "chain10": {k: ch["out"]["oracle_check"][k] for k in ("cpu_steps_per_s_1thread",
                                                                      "cpu_steps_per_s_allthreads", "threads")}}

    if sw is not None:
        value, unit = round(sw["v32"], 2), "instance RK4 steps/s"
        ms_step = 1e3 * sw["el32"] / K
        config = {"workload": f"{args.instances} Tavis-Cummings instances (seed 2), rows {sw['rows_min']}.."
                              f"{sw['rows_max']}, total {sw['rows']} rows {sw['nnz']} nnz, RK4 dt=0.01, one FP32 "
                              f"batch vs the plain FP64 path one instance at a time",
                  "instances": args.instances, "parallelism": "single GPU", "renorm_every": 50,
                  "batch_lanes": sw["lanes"]}
    else:
        value, unit = ch["out"]["fp32_steps_per_s"], "RK4 steps/s"
        ms_step = 1e3 / value
        config = {"workload": ch["out"]["workload"], "parallelism": "single GPU"}
    out = {
        "metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64 matrix values and stage vectors, products and row sums in fp64, hi+lo complex64 "
                 "state; fp64 trace renormalisation and observables)",
        "data": "synthetic", "config": config,
        "roofline": ch["roofline"],
        "cpu_baseline": cpu,
    }
    if sw is not None:
        out.update({"plain_fp64_instance_steps_per_s": round(sw["v64"], 2),
                    "batched_over_plain": round(sw["v32"] / sw["v64"], 3),
                    "max_pop_err_fp32_batch_vs_fp64": sw["worst"], "sweep_assembly_s": round(sw["asm_s"], 3),
                    "roofline_sweep": sw["roofline"], "sweep_layout": sw["layout"]})
    out["chain10"] = ch["out"]
    out["hbm_copy"] = cp
    out["timedep"] = timedep
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
