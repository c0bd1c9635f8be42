"""Benchmark: full GPU presolve of the single large LP workload on one MI355X.

A step is one complete presolve (run_gpu_presolve_with_reduced_lp through the
C-ABI) of a seeded synthetic LP already resident in HBM: m = 1,000,000 rows,
n = 1,500,000 columns, about 7.5M nonzeros (row lengths uniform in 2..14,
Zipf(1.5) column popularity capped at 5,000 per column, dyadic values and the
planted reductions of the workload description).  `value` is LPs presolved
per second.

Extra objects on the JSON line:
  roofline      the row-activity kernel (gp_row_activity on a caller stream,
                timed with HIP events around back-to-back launches) against
                8 TB/s, with its algorithmic bytes 4(m+1) + 12 nnz + 16 n + 24 m;
  cpu_baseline  the serial CPU oracle presolving the same LP once on the
                host (single thread).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-presolver_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def activity_bytes(m, n, nnz):
    return 4 * (m + 1) + 12 * nnz + 16 * n + 24 * m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--activity-launches", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this library runs one LP presolve per GPU and "
              "the benchmark measures one GPU only (N=1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import gp_ffi
    import lp_synth

    if gp_ffi.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    t0 = time.time()
    lp = lp_synth.config1(seed=args.seed)
    m, n, nnz = lp["m"], lp["n"], len(lp["A_idx"])
    log(f"generated LP {m}x{n} nnz={nnz} in {time.time() - t0:.1f}s")
    dlp = gp_ffi.DeviceLp(lp)
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    H = gp_ffi.hip()

    def one_presolve():
        h = gp_ffi.Presolved(dlp, prm, keep_reduced=True)
        s = h.summary()
        return h, s

    for i in range(args.warmup):
        t = time.perf_counter()
        h, s = one_presolve()
        h.free()
        log(f"warmup {i}: {1e3 * (time.perf_counter() - t):.1f} ms -> {s['reduced_rows']}x{s['reduced_cols']}")
    gp_ffi.hip_check(H.hipDeviceSynchronize(), "sync")
    times = []
    last = None
    for i in range(args.steps):
        gp_ffi.hip_check(H.hipDeviceSynchronize(), "sync")
        t = time.perf_counter()
        h, s = one_presolve()
        gp_ffi.hip_check(H.hipDeviceSynchronize(), "sync")
        times.append(time.perf_counter() - t)
        if last is not None:
            last.free()
        last = h
        log(f"step {i}: {1e3 * times[-1]:.1f} ms")
    ms = 1e3 * sum(times) / len(times)
    summary = s

    # Postsolve of the last step's record (reported beside the presolve time).
    rec = last.record()
    red = last.reduced()
    x, y, z = lp_synth.reduced_solution(3, red["m"], red["n"], red["l"], red["u"])
    t = time.perf_counter()
    gp_ffi.postsolve(rec, x, y, z, lp)
    post_ms = 1e3 * (time.perf_counter() - t)
    last.free()
    log(f"postsolve (incl. host<->device copies of x/y/z and the record): {post_ms:.1f} ms")

    # Row-activity kernel roofline: back-to-back launches on one stream.
    stream = gp_ffi.Stream()
    mn, mx = gp_ffi.DeviceArray(size=m), gp_ffi.DeviceArray(size=m)
    ni, pi = gp_ffi.DeviceArray(size=m, dtype=np.int32), gp_ffi.DeviceArray(size=m, dtype=np.int32)
    L = gp_ffi.lib()
    import ctypes as C

    def launch():
        gp_ffi._check(L.gp_row_activity(C.byref(dlp.A.g), dlp.vec["l"].ptr, dlp.vec["u"].ptr, 1e-10, mn.ptr, mx.ptr,
                                        ni.ptr, pi.ptr, stream.ptr))

    for _ in range(5):
        launch()
    stream.sync()
    e0, e1 = gp_ffi.Event(), gp_ffi.Event()
    e0.record(stream)
    for _ in range(args.activity_launches):
        launch()
    e1.record(stream)
    act_ms = e0.elapsed_ms(e1) / args.activity_launches
    act_bytes = activity_bytes(m, n, nnz)
    achieved = act_bytes / (act_ms * 1e-3) / 1e9
    log(f"row activity: {1e3 * act_ms:.1f} us/launch, {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi

        t = time.perf_counter()
        s_cpu, _, _ = oracle_ffi.presolve(lp, oracle_ffi.params(enable_structure_specific_rules=0), False)
        cpu_s = time.perf_counter() - t
        log(f"cpu oracle presolve: {cpu_s:.2f} s")
        if (s_cpu["reduced_rows"], s_cpu["reduced_cols"], s_cpu["reduced_nnz"]) != (
                summary["reduced_rows"], summary["reduced_cols"], summary["reduced_nnz"]):
            raise SystemExit("bench.py: GPU and oracle reduced sizes differ")
        cpu = {"value": 1.0 / cpu_s, "unit": "LP/s", "cores": 1, "kind": "port",
               "sample": f"one full presolve of the same {m}x{n} LP (seed {args.seed}) by the serial C++ oracle "
                         f"({cpu_s:.2f} s)"}

    out = {
        "metric": "presolve ms + HBM GB/s %peak per LP; LPs/s and p99 latency in batched serving",
        "value": 1e3 / ms,
        "unit": "LP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded, dyadic values, planted reductions)",
        "config": {"workload": "single large sparse LP, plain presolve", "m": m, "n": n, "nnz": nnz,
                   "row_len": "U[2,14]", "col_popularity": "Zipf(1.5) cap 5000", "seed": args.seed,
                   "reduced": [summary["reduced_rows"], summary["reduced_cols"], summary["reduced_nnz"]],
                   "iterations": summary["iterations"], "postsolve_ms": post_ms},
        "roofline": {"bound": "hbm", "kernel": "row_activity", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "us_per_launch": 1e3 * act_ms, "algorithmic_bytes": act_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
