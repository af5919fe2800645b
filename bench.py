"""Benchmark: RK4 steps/s of the fused FP32 bulk path (population error checked
against the plain FP64 path in the same run) and SpMV % of HBM peak, on the
10-qubit dissipative spin chain (rho 1024x1024; superoperator 1 048 576 rows,
24 641 536 nnz), RK4 dt = 0.005.  One JSON line on stdout.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 20] [--no-cpu]

Timed regions (each bracketed by a device synchronize on both sides):
  * FP64 drop-in path: K steps, hipGraph-captured fused stage kernels
  * FP32 bulk path:    K steps, same, with FP64 trace renormalisation every 100 steps
  * per-launch stage-kernel durations: the same FP32 steps launched eagerly with a
    HIP event pair around every stage kernel (qh_rk4_timed) -> roofline.achieved
The CPU baseline is the C oracle's FP64 RK4 on the host cores over a bounded
sample of steps of the same workload.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak (GB/s)


def spmv_bytes(nnz, rows, S):
    """algorithmic bytes of one SpMV: nnz*(S+4) + 4*(rows+1) + 2*rows*S"""
    return nnz * (S + 4) + 4 * (rows + 1) + 2 * rows * S


def rk4_step_bytes(nnz, rows, S):
    """algorithmic bytes of one fused RK4 step: 4 matrix streams + 14 vector passes"""
    return 4 * (nnz * (S + 4) + 4 * (rows + 1)) + 14 * rows * S


def kernel_name(variant, real):
    """name of the stage kernel a lanes-per-row variant code selects"""
    if 128 <= variant <= 134:
        u, vv, pp = {128: (2, 0, 0), 129: (4, 0, 0), 130: (8, 0, 0), 131: (4, 1, 0), 132: (8, 1, 0),
                     133: (2, 0, 1), 134: (4, 0, 1)}[variant]
        tf = ("false", "true")
        return f"k_ks_stage<{real}, MODE, {u}, {tf[vv]}, {tf[pp]}> (sliced layout)"
    if variant in (140, 141, 142):
        u, w = {140: (2, "false"), 141: (4, "false"), 142: (2, "true")}[variant]
        return f"k_ks2_stage<MODE, {u}, {w}> (row-pair sliced layout)"
    return f"k_rk_stage<{real}, {variant}> (CSR, {variant} lanes per row)"


def read_traffic(path):
    """HBM bytes per stage launch from the committed rocprofv3 --pmc summary
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nq", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--renorm", type=int, default=100)
    ap.add_argument("--workload", default="chain10", choices=["chain10", "tc512", "circuit12", "jc", "copy", "kerr", "mcwf"])
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--levels", type=int, default=8)
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--rank", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=2000)
    ap.add_argument("--fp32-variant", type=int, default=0,
                    help="force the FP32 stage-kernel variant (qh_mat_set_lanes code; 0 = autotune). Counter "
                         "collection perturbs the autotuner's timings, so profiled runs name the variant the "
                         "plain run selects")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: this build targets a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import quac_amd as qa
    import recipes
    from qbuild import build_system, subsystem_layout

    if args.workload != "chain10":
        import bench_workloads
        out = getattr(bench_workloads, args.workload)(args)
        print(json.dumps(out), flush=True)
        return

    qa.set_option("ts_rk_type", "4")
    qa.set_option("ts_adapt_type", "none")
    K, W, dt = args.steps, args.warmup, 0.005
    recipe = recipes.spin_chain_recipe(args.nq, seed=1)
    N, loc = recipes.spin_chain_init(args.nq)
    rows = N * N

    t0 = time.perf_counter()
    qs, _, _ = build_system(recipe)
    qa.synchronize()
    assemble_s = time.perf_counter() - t0
    A = qs.matrix()
    if args.fp32_variant:
        A.set_lanes(0, args.fp32_variant)
    rows_, nnz = A.info()
    assert rows_ == rows
    lay = subsystem_layout(recipe)
    levels, strides = [L for L, _ in lay], [s for _, s in lay]
    x0 = np.zeros(rows, np.complex128)
    x0[loc] = 1.0

    dev_ms = {}

    def run(prec, renorm):
        v = qa.DeviceVec(rows, prec)
        v.set(x0)
        qa.rk4_device(A, v, dt, W, tot=N, renorm_every=renorm)   # warmup (autotune + graph capture)
        qa.synchronize()
        t = time.perf_counter()
        with qa.StreamTimer() as tm:   # HIP events on the stream the stage kernels run on
            qa.rk4_device(A, v, dt, K, tot=N, renorm_every=renorm)
        qa.synchronize()
        el = time.perf_counter() - t
        dev_ms[prec] = tm.ms
        dv = qa.DeviceVec(rows, qa.FP64)
        qa._qh(qa.lib().qh_vec_copy(v.h, dv.h))
        pops = np.zeros(len(levels))
        import ctypes
        lv = np.array(levels, np.int64)
        st = np.array(strides, np.int64)
        qa._qh(qa.lib().qh_populations(dv.h, len(levels), lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1,
                                       pops.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        tr = ctypes.c_double()
        ti = ctypes.c_double()
        qa._qh(qa.lib().qh_trace(dv.h, N, ctypes.byref(tr), ctypes.byref(ti)))
        return el, pops, tr.value

    el64, pops64, tr64 = run(qa.FP64, 0)
    el32, pops32, tr32 = run(qa.FP32, args.renorm)
    pop_err = float(np.abs(pops32 - pops64).max())

    # per-launch kernel durations (events around every stage kernel), FP32 and FP64
    nprof = min(50, K)
    v = qa.DeviceVec(rows, qa.FP32)
    v.set(x0)
    stage_ms32 = qa.rk4_device(A, v, dt, nprof, tot=N, timed=True)
    v64 = qa.DeviceVec(rows, qa.FP64)
    v64.set(x0)
    stage_ms64 = qa.rk4_device(A, v64, dt, nprof, tot=N, timed=True)
    spmv_ms32 = qa.bench_spmv(A, qa.FP32, 30)
    spmv_ms64 = qa.bench_spmv(A, qa.FP64, 30)
    lanes = A.lanes()
    layout = A.layout_info()
    # bytes the sliced-layout step actually has to move: 4 x (stored values + slot
    # metadata) + 17 vector passes (4 gathers, 3 K and 3 Y writes, 3 x reads for
    # the Y's, K1..K3 + x reads and the x' write in the last epilogue)
    if lanes[1] in (140, 141, 142):  # row-pair layout: 40-byte slot metadata, values stored per lane pair
        layout = dict(A.pair_layout_info(), kind="row-pair (128 rows per wavefront)")
        meta32 = layout["meta_bytes"]
        vals32 = layout["stored_pairs"] * 16 + layout["stored_real_pairs"] * 8
    else:
        layout = dict(layout, kind="one row per lane (64 rows per wavefront)")
        meta32 = layout["slots"] * 32 + 12 * layout["slices"]
        vals32 = layout["stored_vals"] * 8 + layout["stored_reals"] * 4
    ks_step32 = 4 * (vals32 + meta32) + 17 * rows * 8

    # mean stage-kernel duration over the timed FP32 region: its device time
    # (event pair) over its 4K stage launches (includes the renormalisation
    # kernels every renorm steps and the gaps between graph nodes)
    launch_ms32 = dev_ms[qa.FP32] / (4 * K)
    launch_ms64 = dev_ms[qa.FP64] / (4 * K)
    b_step32 = rk4_step_bytes(nnz, rows, 8)
    b_step64 = rk4_step_bytes(nnz, rows, 16)
    ach_stage32 = (b_step32 / 4) / (launch_ms32 * 1e-3) / 1e9
    ach_stage64 = (b_step64 / 4) / (launch_ms64 * 1e-3) / 1e9
    spmv_gbs32 = spmv_bytes(nnz, rows, 8) / (spmv_ms32 * 1e-3) / 1e9
    spmv_gbs64 = spmv_bytes(nnz, rows, 16) / (spmv_ms64 * 1e-3) / 1e9

    kname = kernel_name(lanes[1], "float")
    traffic, traffic_src = read_traffic(os.path.join(ROOT, "profiles", "k_ks2_stage_chain10_traffic.txt"))

    steps_s32 = K / el32
    steps_s64 = K / el64

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline leg only
        cores = min(os.cpu_count() or 1, 16)
        tot, terms = recipes.oracle_terms(recipe)
        t = time.perf_counter()
        m = oracle.construct(tot, terms, dm=True)
        cpu_asm = time.perf_counter() - t
        xs = x0.copy()
        t = time.perf_counter()
        xs = m.rk4(xs, dt, 1, nthreads=cores)
        one = time.perf_counter() - t
        n_cpu = max(1, int(args.cpu_seconds / max(one, 1e-3)))
        t = time.perf_counter()
        xs = m.rk4(xs, dt, n_cpu, nthreads=cores)
        cpu_el = time.perf_counter() - t
        cpu = {"value": round(n_cpu / cpu_el, 4), "unit": "RK4 steps/s (FP64)", "cores": cores, "kind": "port",
               "sample": f"{n_cpu} RK4 steps of the same 10-qubit chain (1+{n_cpu} after a 1-step probe), "
                         f"OpenMP over {cores} threads, {cpu_model()}; oracle assembly {cpu_asm:.2f} s"}

    out = {
        "metric": "batched Lindblad RK4 steps/s at 1e-6 pop err vs plain FP64; SpMV % HBM peak",
        "value": round(steps_s32, 3),
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * el32 / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64 matrix values and stage vectors, products and row sums in fp64, "
                 "hi+lo complex64 state; fp64 trace renormalisation and observables)",
        "data": "synthetic",
        "config": {"workload": f"{args.nq}-qubit dissipative spin chain, rho {N}x{N}, superoperator {rows} rows "
                               f"{nnz} nnz, RK4 dt={dt}, h_i,J_i~U[0.5,1.5] seed 1, gamma1=1e-3, gamma2=5e-3",
                   "parallelism": "single GPU", "renorm_every": args.renorm},
        "roofline": {"bound": "hbm", "achieved": round(ach_stage32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_stage32 / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": f"{kname} (mean of the 4 fused RK4 stage launches)",
                     "bytes_per_launch": b_step32 / 4, "launch_ms": round(launch_ms32, 5),
                     "launch_ms_per_kernel_events": round(stage_ms32, 5),
                     "format_bytes_per_launch": ks_step32 / 4,
                     "format_gbs": round((ks_step32 / 4) / (launch_ms32 * 1e-3) / 1e9, 1),
                     "traffic_source": traffic_src,
                     "traffic_gbs": round(traffic / (launch_ms32 * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_frac": round(traffic / (launch_ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                     if traffic else None},
        "cpu_baseline": cpu,
        "fp64_steps_per_s": round(steps_s64, 3),
        "fp32_over_fp64": round(steps_s32 / steps_s64, 3),
        "max_pop_err_fp32_vs_fp64": pop_err,
        "trace_fp32": tr32, "trace_fp64": tr64,
        "fp64_stage": {"achieved_gbs": round(ach_stage64, 1), "frac": round(ach_stage64 / HBM_PEAK_GBS, 4),
                       "launch_ms": round(launch_ms64, 5), "launch_ms_per_kernel_events": round(stage_ms64, 5)},
        "spmv": {"fp32_gbs": round(spmv_gbs32, 1), "fp32_frac": round(spmv_gbs32 / HBM_PEAK_GBS, 4),
                 "fp64_gbs": round(spmv_gbs64, 1), "fp64_frac": round(spmv_gbs64 / HBM_PEAK_GBS, 4),
                 "fp32_ms": round(spmv_ms32, 5), "fp64_ms": round(spmv_ms64, 5), "lanes_per_row": lanes},
        "assembly_s": round(assemble_s, 4),
        "sliced_layout": dict(layout, step_bytes_fp32=ks_step32,
                              step_gbs_fp32=round(ks_step32 / (el32 / K) / 1e9, 1)),
    }
    print(json.dumps(out), flush=True)
    qs.destroy()


if __name__ == "__main__":
    main()
