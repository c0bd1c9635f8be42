"""Benchmark: full GPU presolve of the single large LP workload on one MI355X.

A step is one complete presolve (run_gpu_presolve_with_reduced_lp through the
C-ABI) of a seeded synthetic LP already resident in HBM: m = 1,000,000 rows,
n = 1,500,000 columns, about 7.5M nonzeros (row lengths uniform in 2..14,
Zipf(1.5) column popularity capped at 5,000 per column, dyadic values and the
planted reductions of the workload description).  `value` is LPs presolved
per second.

Extra objects on the JSON line:
  roofline           the compaction (k_build_rank, k_compact_map, the tile-offset
                     scan and k_compact_write; the largest share of presolve
                     kernel time after the replay) on the full A with
                     15% of rows and columns dropped, timed with HIP events
                     around back-to-back launches on one stream, against 8 TB/s;
                     traffic comes from the committed rocprofv3 PMC summary;
  roofline_activity  one row-activity call (gp_row_activity on a caller
                     stream: k_pack_lu then k_row_activity_lu), algorithmic
                     bytes 4(m+1) + 12 nnz + 16 n + 24 m, with both kernels
                     also timed apart;
  roofline_transpose the transpose of the full A (three LSD sort passes over
                     the column indices, row pointers from the sorted keys,
                     one gather), algorithmic bytes
                     4(m+1) + 12 nnz + 4(n+1) + 12 nnz + 8 n;
  roofline_postsolve the prepared postsolve of the last record: tape records,
                     dependency levels of the two replay passes, time per level
                     and the bytes model (tape read per pass, outputs written);
  end_to_end         the solver hand-off on the same LP: presolve, the reduced
                     LP solved in HBM by gp_pdhg_solve (restarted PDHG, relative
                     KKT tolerance --solve-tol), one-shot postsolve of that
                     solution; the same as one gp_presolve_solve call (the
                     postsolve prepared on a second stream during the solve);
                     then the same solver on the original LP without presolve,
                     with the per-iteration bytes model 24 nnz + 76 n + 60 m
                     against 8 TB/s;
  cpu_baseline       the C++ oracle presolving the same LP on the host:
                     single thread and the OpenMP build on every usable core,
                     median of five runs after a warm-up, with the core count
                     and CPU model;
  long_rows          the same LP shape with row lengths U[2,64] (~30.5M nnz), on
                     which the batched singleton closure runs at its default
                     10M-nonzero threshold (presolve ms, reduced sizes);
  batched            4,096 LPs generated on the device, presolved together,
                     against one-at-a-time plain presolve (LP/s, speedup);
  serving            60 s of Poisson(200/s) presolve + postsolve requests whose
                     LPs are generated per request (p50/p99, LP/s), with the
                     capacity against serial plain processing.
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


def kernel_rooflines(gp_ffi, dlp, lp, launches):
    """Row activity (gp_row_activity on a caller stream) and compaction
    kernels, each timed with HIP events around back-to-back launches."""
    import ctypes as C

    import numpy as np

    m, n, nnz = lp["m"], lp["n"], len(lp["A_idx"])
    stream = gp_ffi.Stream()
    mn, mx = gp_ffi.DeviceArray(size=m), gp_ffi.DeviceArray(size=m)
    ni, pi = gp_ffi.DeviceArray(size=m, dtype=np.int32), gp_ffi.DeviceArray(size=m, dtype=np.int32)
    L = gp_ffi.lib()

    def launch():
        gp_ffi._check(L.gp_row_activity(C.byref(dlp.A.g), dlp.vec["l"].ptr, dlp.vec["u"].ptr, 1e-10, mn.ptr, mx.ptr,
                                        ni.ptr, pi.ptr, stream.ptr))

    for _ in range(5):
        launch()
    stream.sync()
    e0, e1 = gp_ffi.Event(), gp_ffi.Event()
    e0.record(stream)
    for _ in range(launches):
        launch()
    e1.record(stream)
    act_ms = e0.elapsed_ms(e1) / launches
    act_bytes = activity_bytes(m, n, nnz)
    act = {"us": 1e3 * act_ms, "bytes": act_bytes, "achieved": act_bytes / (act_ms * 1e-3) / 1e9,
           # the "effective" figure also counts the (l, u) gather of every entry
           "effective_bytes": act_bytes + 16 * nnz}
    act["effective_achieved"] = act["effective_bytes"] / (act_ms * 1e-3) / 1e9
    log(f"row activity (k_pack_lu + k_row_activity_lu): {act['us']:.1f} us/call, {act['achieved']:.0f} GB/s")
    act["kernels"] = measure_activity_kernels(gp_ffi, dlp, lp, stream, launches, mn.numpy(), mx.numpy())
    comp = measure_compaction(gp_ffi, dlp, lp, stream, launches)
    log(f"compaction: {comp['us']:.1f} us/call, {comp['achieved']:.0f} GB/s")
    comp["pair"] = measure_compaction_pair(gp_ffi, dlp, lp, stream, launches)
    tr = measure_transpose(gp_ffi, dlp, lp, stream, max(5, launches // 5))
    return act, comp, tr


def measure_activity_kernels(gp_ffi, dlp, lp, stream, launches, want_mn, want_mx):
    """The two kernels of one row-activity call timed apart: gpk_pack_lu
    (interleaves l and u) and gpk_row_activity_lu; the results must equal
    those of the gp_row_activity call."""
    import ctypes as C

    import numpy as np

    m, n, nnz = lp["m"], lp["n"], len(lp["A_idx"])
    L = gp_ffi.lib()
    vp = C.c_void_p
    L.gpk_pack_lu.argtypes = [vp, vp, C.c_int32, vp, vp]
    L.gpk_row_activity_lu.argtypes = [vp, vp, vp, C.c_int32, vp, C.c_double, vp, vp, vp, vp, vp]
    lu = gp_ffi.DeviceArray(size=2 * n)
    mn, mx = gp_ffi.DeviceArray(size=m), gp_ffi.DeviceArray(size=m)
    ni, pi = gp_ffi.DeviceArray(size=m, dtype=np.int32), gp_ffi.DeviceArray(size=m, dtype=np.int32)
    A = dlp.A.g

    def pack():
        gp_ffi.hip_check(L.gpk_pack_lu(dlp.vec["l"].ptr, dlp.vec["u"].ptr, n, lu.ptr, stream.ptr), "pack lu")

    def act():
        gp_ffi.hip_check(L.gpk_row_activity_lu(A.row_ptr, A.col_idx, A.val, m, lu.ptr, 1e-10, mn.ptr, mx.ptr,
                                               ni.ptr, pi.ptr, stream.ptr), "activity lu")

    pack()
    for _ in range(3):
        act()
    stream.sync()
    e0, e1, e2 = gp_ffi.Event(), gp_ffi.Event(), gp_ffi.Event()
    e0.record(stream)
    for _ in range(launches):
        act()
    e1.record(stream)
    for _ in range(launches):
        pack()
    e2.record(stream)
    us = 1e3 * e0.elapsed_ms(e1) / launches
    pack_us = 1e3 * e1.elapsed_ms(e2) / launches
    same = bool(np.array_equal(mn.numpy(), want_mn) and np.array_equal(mx.numpy(), want_mx))
    b = activity_bytes(m, n, nnz)
    if not same:
        raise SystemExit("bench.py: gpk_row_activity_lu differs from gp_row_activity")
    log(f"  k_row_activity_lu {us:.1f} us/launch ({b / (us * 1e-6) / 1e9:.0f} GB/s), k_pack_lu {pack_us:.1f} us")
    return {"k_row_activity_lu_us": us, "k_pack_lu_us": pack_us,
            "k_row_activity_lu_achieved": b / (us * 1e-6) / 1e9}


CALIBRATION_BYTES = 256 << 20


def calibration_copy(gp_ffi):
    """One device-to-device copy of 256 MiB: the reference point that scales
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    a = gp_ffi.DeviceArray(size=CALIBRATION_BYTES // 8)
    b = gp_ffi.DeviceArray(size=CALIBRATION_BYTES // 8)
    gp_ffi.hip_check(gp_ffi.hip().hipMemcpy(b.ptr, a.ptr, CALIBRATION_BYTES, gp_ffi.D2D), "calibration copy")
    gp_ffi.hip_check(gp_ffi.hip().hipDeviceSynchronize(), "sync")


def compaction_bytes(m, n, nnz, m_new, nnz_new):
    """Algorithmic bytes of one compaction in the count / scan / copy model of
    the reference's compaction (gpu_presolve.cu:1881-1978): a counting pass
    over the row pointers and column indices, the row and column maps, one
    read of every entry and one write of every kept entry, the new row
    pointers, and the row-pointer scan (read + write)."""
    return (4 * (m + 1) + 4 * nnz) + 4 * m + 4 * n + 12 * nnz + 12 * nnz_new + 4 * (m_new + 1) + 2 * 4 * m_new


def compaction_compulsory_bytes(m, n, nnz, m_new, nnz_new):
    """The floor under any compaction: read the row pointers, the maps and
    every entry once; write the kept entries and the new row pointers."""
    return 4 * (m + 1) + 4 * m + 4 * n + 12 * nnz + 12 * nnz_new + 4 * (m_new + 1)


def measure_compaction(gp_ffi, dlp, lp, stream, launches, frac=0.15, seed=7):
    """Times the compaction the presolve uses (gpk_build_rank, gpk_compact_map,
    the tile-offset scan and gpk_compact_write, back to back on one stream)
    on the full A with 15% of rows and columns dropped (a typical early
    apply), each kernel also timed apart, and checks the output against the
    host."""
    import ctypes as C

    import numpy as np

    class Lookback(C.Structure):
        _fields_ = [("status", C.c_void_p), ("ticket", C.c_void_p), ("error", C.c_void_p), ("epoch", C.c_uint32)]

    m, n = lp["m"], lp["n"]
    nnz = len(lp["A_idx"])
    rng = np.random.default_rng(seed)
    keep_r = rng.random(m) >= frac
    keep_c = rng.random(n) >= frac
    row_map = np.where(keep_r, np.cumsum(keep_r) - 1, -1).astype(np.int32)
    col_map = np.where(keep_c, np.cumsum(keep_c) - 1, -1).astype(np.int32)
    rows = np.repeat(np.arange(m), np.diff(lp["A_ptr"]))
    kept = keep_r[rows] & keep_c[lp["A_idx"]]
    m_new, nnz_new = int(keep_r.sum()), int(kept.sum())
    L = gp_ffi.lib()
    vp = C.c_void_p
    L.gpk_build_rank.argtypes = [vp, C.c_int32, vp, vp]
    L.gpk_compact_map.argtypes = [vp, vp, C.c_int32, vp, vp, vp, vp, vp, vp]
    L.gpk_compact_write.argtypes = [vp, vp, C.c_int32, vp, vp, vp, vp, C.c_int32, vp, vp, vp]
    L.gpk_compact_tiles.restype = C.c_int32
    L.gpk_compact_tiles.argtypes = [C.c_int32]
    L.gpk_scan_lb_i32.argtypes = [vp, vp, C.c_int64, Lookback, vp]
    L.gpk_lookback_tiles.restype = C.c_int64
    L.gpk_lookback_tiles.argtypes = [C.c_int64]
    d_rm, d_cm = gp_ffi.DeviceArray(row_map), gp_ffi.DeviceArray(col_map)
    d_rank = gp_ffi.DeviceArray(size=2 * ((n + 63) // 64), dtype=np.uint64)
    jn = gp_ffi.DeviceArray(size=nnz, dtype=np.int32)
    o_ptr = gp_ffi.DeviceArray(size=m_new + 1, dtype=np.int32)
    o_idx = gp_ffi.DeviceArray(size=nnz, dtype=np.int32)
    o_val = gp_ffi.DeviceArray(size=nnz)
    tiles = L.gpk_compact_tiles(m)
    t_off = gp_ffi.DeviceArray(size=tiles + 1, dtype=np.int32)
    status = gp_ffi.DeviceArray(np.zeros(max(1, L.gpk_lookback_tiles(tiles)), np.uint64))
    tickets = gp_ffi.DeviceArray(np.zeros(4 * launches + 16, np.int32))
    err = gp_ffi.DeviceArray(np.zeros(1, np.int32))
    A = dlp.A.g
    launch_no = [0]
    ck = gp_ffi.hip_check

    def rank():
        ck(L.gpk_build_rank(d_cm.ptr, n, d_rank.ptr, stream.ptr), "build rank")

    def cmap():
        ck(L.gpk_compact_map(A.row_ptr, A.col_idx, m, d_rm.ptr, d_rank.ptr, jn.ptr, o_ptr.ptr, t_off.ptr, stream.ptr),
           "map")

    def scan():
        k = launch_no[0]
        launch_no[0] += 1
        lb = Lookback(status.ptr, tickets.ptr + 4 * k, err.ptr, k + 1)
        ck(L.gpk_scan_lb_i32(t_off.ptr, t_off.ptr, tiles, lb, stream.ptr), "scan")

    def write():
        ck(L.gpk_compact_write(A.row_ptr, A.val, m, d_rm.ptr, jn.ptr, t_off.ptr, o_ptr.ptr, m_new, o_idx.ptr,
                               o_val.ptr, stream.ptr), "write")

    def whole():
        rank()
        cmap()
        scan()
        write()

    for _ in range(3):
        whole()
    stream.sync()
    ev = [gp_ffi.Event() for _ in range(6)]
    ev[0].record(stream)
    for _ in range(launches):
        whole()
    ev[1].record(stream)
    us = 1e3 * ev[0].elapsed_ms(ev[1]) / launches
    want_ptr = np.concatenate([[0], np.cumsum(np.bincount(rows[kept], minlength=m)[keep_r])]).astype(np.int32)
    if (err.numpy()[0] != 0 or not np.array_equal(o_ptr.numpy(), want_ptr)
            or not np.array_equal(o_idx.numpy()[:nnz_new], col_map[lp["A_idx"][kept]])
            or not np.array_equal(o_val.numpy()[:nnz_new], lp["A_val"][kept])):
        raise SystemExit("bench.py: the compaction differs from the host")
    # Each kernel apart (the map needs the rank, the write the scanned offsets
    # and the map's counts, so the write is timed after map + scan).
    parts = {}
    for name, f in (("k_build_rank", rank), ("k_compact_map", cmap), ("tile scan", None), ("k_compact_write", None)):
        if name == "tile scan":
            e0, e1 = gp_ffi.Event(), gp_ffi.Event()
            e0.record(stream)
            for _ in range(launches):
                cmap()
                scan()
            e1.record(stream)
            parts[name] = 1e3 * e0.elapsed_ms(e1) / launches - parts["k_compact_map"]
            continue
        if name == "k_compact_write":
            e0, e1 = gp_ffi.Event(), gp_ffi.Event()
            e0.record(stream)
            for _ in range(launches):
                cmap()
                scan()
                write()
            e1.record(stream)
            parts[name] = 1e3 * e0.elapsed_ms(e1) / launches - parts["k_compact_map"] - parts["tile scan"]
            continue
        e0, e1 = gp_ffi.Event(), gp_ffi.Event()
        e0.record(stream)
        for _ in range(launches):
            f()
        e1.record(stream)
        parts[name] = 1e3 * e0.elapsed_ms(e1) / launches
    nbytes = compaction_bytes(m, n, nnz, m_new, nnz_new)
    floor = compaction_compulsory_bytes(m, n, nnz, m_new, nnz_new)
    log("compaction kernels (us): " + ", ".join(f"{k} {v:.1f}" for k, v in parts.items()))
    return {"us": us, "bytes": nbytes, "achieved": nbytes / (us * 1e-6) / 1e9, "kernels_us": parts,
            "compulsory_bytes": floor, "compulsory_achieved": floor / (us * 1e-6) / 1e9,
            "workload": f"A of the LP ({m}x{n}, {nnz} nnz) with 15% of rows and columns dropped "
                        f"-> {m_new} rows, {nnz_new} nnz"}


def measure_compaction_pair(gp_ffi, dlp, lp, stream, launches, frac=0.15, seed=7):
    """Times the compaction a presolve plan application runs: A and its
    transpose by the same maps, with both rank-word arrays built, one map
    launch over both matrices' tiles (gpk_compact_map_pair), one scan of both
    sides' tile counts and one write launch (gpk_compact_write_pair), on the
    full LP with 15% of rows and columns dropped; checks both outputs against
    the host."""
    import ctypes as C

    import numpy as np

    class Lookback(C.Structure):
        _fields_ = [("status", C.c_void_p), ("ticket", C.c_void_p), ("error", C.c_void_p), ("epoch", C.c_uint32)]

    vp = C.c_void_p

    class Side(C.Structure):
        _fields_ = [("ptr", vp), ("idx", vp), ("val", vp), ("m_old", C.c_int32), ("m_new", C.c_int32),
                    ("row_map", vp), ("col_rank", vp), ("jn", vp), ("out_ptr", vp), ("out_idx", vp),
                    ("out_val", vp)]

    m, n = lp["m"], lp["n"]
    nnz = len(lp["A_idx"])
    rng = np.random.default_rng(seed)
    keep_r = rng.random(m) >= frac
    keep_c = rng.random(n) >= frac
    row_map = np.where(keep_r, np.cumsum(keep_r) - 1, -1).astype(np.int32)
    col_map = np.where(keep_c, np.cumsum(keep_c) - 1, -1).astype(np.int32)
    rows = np.repeat(np.arange(m), np.diff(lp["A_ptr"]))
    kept = keep_r[rows] & keep_c[lp["A_idx"]]
    m_new, n_new, nnz_new = int(keep_r.sum()), int(keep_c.sum()), int(kept.sum())
    L = gp_ffi.lib()
    L.gpk_build_rank.argtypes = [vp, C.c_int32, vp, vp]
    L.gpk_compact_map_pair.argtypes = [Side, Side, vp, vp, vp]
    L.gpk_compact_write_pair.argtypes = [Side, Side, vp, C.c_int32, vp]
    L.gpk_compact_tiles.restype = C.c_int32
    L.gpk_compact_tiles.argtypes = [C.c_int32]
    L.gpk_scan_lb_i32.argtypes = [vp, vp, C.c_int64, Lookback, vp]
    L.gpk_lookback_tiles.restype = C.c_int64
    L.gpk_lookback_tiles.argtypes = [C.c_int64]
    d_rm, d_cm = gp_ffi.DeviceArray(row_map), gp_ffi.DeviceArray(col_map)
    c_rank = gp_ffi.DeviceArray(size=2 * ((n + 63) // 64), dtype=np.uint64)
    r_rank = gp_ffi.DeviceArray(size=2 * ((m + 63) // 64), dtype=np.uint64)
    jn, t_jn = gp_ffi.DeviceArray(size=nnz, dtype=np.int32), gp_ffi.DeviceArray(size=nnz, dtype=np.int32)
    o_ptr = gp_ffi.DeviceArray(size=m + 1, dtype=np.int32)
    t_ptr = gp_ffi.DeviceArray(size=n + 1, dtype=np.int32)
    o_idx, t_idx = gp_ffi.DeviceArray(size=nnz, dtype=np.int32), gp_ffi.DeviceArray(size=nnz, dtype=np.int32)
    o_val, t_val = gp_ffi.DeviceArray(size=nnz), gp_ffi.DeviceArray(size=nnz)
    ta, tt = L.gpk_compact_tiles(m), L.gpk_compact_tiles(n)
    from_counts = ta <= 256 and tt <= 256
    t_off = gp_ffi.DeviceArray(size=ta + tt + 1, dtype=np.int32)
    status = gp_ffi.DeviceArray(np.zeros(max(1, L.gpk_lookback_tiles(ta + tt)), np.uint64))
    tickets = gp_ffi.DeviceArray(np.zeros(4 * launches + 16, np.int32))
    err = gp_ffi.DeviceArray(np.zeros(1, np.int32))
    A, AT = dlp.A.g, dlp.AT.g
    a = Side(A.row_ptr, A.col_idx, A.val, m, m_new, d_rm.ptr, c_rank.ptr, jn.ptr, o_ptr.ptr, o_idx.ptr, o_val.ptr)
    t = Side(AT.row_ptr, AT.col_idx, AT.val, n, n_new, d_cm.ptr, r_rank.ptr, t_jn.ptr, t_ptr.ptr, t_idx.ptr,
             t_val.ptr)
    launch_no = [0]
    ck = gp_ffi.hip_check

    def whole():
        ck(L.gpk_build_rank(d_cm.ptr, n, c_rank.ptr, stream.ptr), "build rank")
        ck(L.gpk_build_rank(d_rm.ptr, m, r_rank.ptr, stream.ptr), "build rank")
        ck(L.gpk_compact_map_pair(a, t, t_off.ptr, None, stream.ptr), "map pair")
        if not from_counts:
            k = launch_no[0]
            launch_no[0] += 1
            lb = Lookback(status.ptr, tickets.ptr + 4 * k, err.ptr, k + 1)
            ck(L.gpk_scan_lb_i32(t_off.ptr, t_off.ptr, ta + tt, lb, stream.ptr), "scan")
        ck(L.gpk_compact_write_pair(a, t, t_off.ptr, 1 if from_counts else 0, stream.ptr), "write pair")

    for _ in range(3):
        whole()
    stream.sync()
    e0, e1 = gp_ffi.Event(), gp_ffi.Event()
    e0.record(stream)
    for _ in range(launches):
        whole()
    e1.record(stream)
    us = 1e3 * e0.elapsed_ms(e1) / launches
    want_ptr = np.concatenate([[0], np.cumsum(np.bincount(rows[kept], minlength=m)[keep_r])]).astype(np.int32)
    at_rows = np.repeat(np.arange(n), np.diff(lp["AT_ptr"]))
    t_kept = keep_c[at_rows] & keep_r[lp["AT_idx"]]
    want_tptr = np.concatenate([[0], np.cumsum(np.bincount(at_rows[t_kept], minlength=n)[keep_c])]).astype(np.int32)
    if (err.numpy()[0] != 0 or not np.array_equal(o_ptr.numpy()[:m_new + 1], want_ptr)
            or not np.array_equal(o_idx.numpy()[:nnz_new], col_map[lp["A_idx"][kept]])
            or not np.array_equal(o_val.numpy()[:nnz_new], lp["A_val"][kept])
            or not np.array_equal(t_ptr.numpy()[:n_new + 1], want_tptr)
            or not np.array_equal(t_idx.numpy()[:nnz_new], row_map[lp["AT_idx"][t_kept]])
            or not np.array_equal(t_val.numpy()[:nnz_new], lp["AT_val"][t_kept])):
        raise SystemExit("bench.py: the paired compaction differs from the host")
    nbytes = compaction_bytes(m, n, nnz, m_new, nnz_new) + compaction_bytes(n, m, nnz, n_new, nnz_new)
    log(f"paired compaction (A and AT): {us:.1f} us/call, {nbytes / (us * 1e-6) / 1e9:.0f} GB/s")
    return {"us": us, "bytes": nbytes, "achieved": nbytes / (us * 1e-6) / 1e9,
            "workload": f"A and AT of the LP ({m}x{n}, {nnz} nnz) with 15% of rows and columns dropped "
                        f"-> {m_new}x{n_new}, {nnz_new} nnz"}


def transpose_bytes(m, n, nnz):
    """Algorithmic bytes of one transpose of an m x n CSR: read the row
    pointers and entries, write the transpose's row pointers and entries,
    and the column counts and their scan (8 n)."""
    return 4 * (m + 1) + 12 * nnz + 4 * (n + 1) + 12 * nnz + 8 * n


def measure_transpose(gp_ffi, dlp, lp, stream, launches):
    """Times the library's transpose (stable sort of the column indices
    carrying entry positions, row pointers from the sorted keys, rows of the
    entries, one gather of rows and values) on the full A, back to back on
    one stream, each kernel group also timed apart; the result must equal
    the LP's own AT."""
    import ctypes as C

    import numpy as np

    m, n, nnz = lp["m"], lp["n"], len(lp["A_idx"])
    L = gp_ffi.lib()
    vp = C.c_void_p
    L.gpk_radix_scratch_ints.restype = C.c_int64
    L.gpk_radix_scratch_ints.argtypes = [C.c_int64]
    L.gpk_transpose_sort.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, C.POINTER(C.c_int32), vp]
    L.gpk_col_bounds.argtypes = [vp, C.c_int32, C.c_int32, vp, vp]
    L.gpk_expand_rows.argtypes = [vp, vp, C.c_int32, vp, vp]
    L.gpk_transpose_gather.argtypes = [vp, vp, C.c_int32, vp, vp, vp]
    bits = max(1, int(n - 1).bit_length())
    ka, kb = gp_ffi.DeviceArray(size=nnz, dtype=np.uint32), gp_ffi.DeviceArray(size=nnz, dtype=np.uint32)
    va, vb = gp_ffi.DeviceArray(size=nnz, dtype=np.int32), gp_ffi.DeviceArray(size=nnz, dtype=np.int32)
    scratch = gp_ffi.DeviceArray(size=L.gpk_radix_scratch_ints(nnz), dtype=np.int32)
    ptr = gp_ffi.DeviceArray(size=n + 1, dtype=np.int32)
    idx, val = gp_ffi.DeviceArray(size=nnz, dtype=np.int32), gp_ffi.DeviceArray(size=nnz)
    rv = gp_ffi.DeviceArray(size=2 * nnz)  # 16-byte (row, value) pairs
    A = dlp.A.g
    in_b = C.c_int32(0)
    ck = gp_ffi.hip_check

    def sort():
        ck(L.gpk_transpose_sort(A.col_idx, nnz, bits, ka.ptr, va.ptr, kb.ptr, vb.ptr, scratch.ptr, C.byref(in_b),
                                stream.ptr), "transpose sort")

    def finish():
        keys, perm = (kb, vb) if in_b.value else (ka, va)
        ck(L.gpk_col_bounds(keys.ptr, nnz, n, ptr.ptr, stream.ptr), "col bounds")
        ck(L.gpk_expand_rows(A.row_ptr, A.val, m, rv.ptr, stream.ptr), "expand rows")
        ck(L.gpk_transpose_gather(perm.ptr, rv.ptr, nnz, idx.ptr, val.ptr, stream.ptr), "gather")

    for _ in range(2):
        sort()
        finish()
    stream.sync()
    e = [gp_ffi.Event() for _ in range(3)]
    e[0].record(stream)
    for _ in range(launches):
        sort()
        finish()
    e[1].record(stream)
    for _ in range(launches):
        sort()
    e[2].record(stream)
    us = 1e3 * e[0].elapsed_ms(e[1]) / launches
    sort_us = 1e3 * e[1].elapsed_ms(e[2]) / launches
    if not (np.array_equal(ptr.numpy(), lp["AT_ptr"]) and np.array_equal(idx.numpy(), lp["AT_idx"])
            and np.array_equal(val.numpy(), lp["AT_val"])):
        raise SystemExit("bench.py: the transpose differs from the LP's AT")
    b = transpose_bytes(m, n, nnz)
    log(f"transpose: {us:.1f} us/call ({b / (us * 1e-6) / 1e9:.0f} GB/s); sort {sort_us:.1f} us, "
        f"bounds + rows + gather {us - sort_us:.1f} us")
    return {"us": us, "bytes": b, "achieved": b / (us * 1e-6) / 1e9,
            "kernels_us": {"sort (3 LSD passes)": sort_us, "bounds + rows + gather": us - sort_us},
            "workload": f"A of the LP ({m}x{n}, {nnz} nnz) -> AT ({n}x{m})"}


def pdhg_iteration_bytes(m, n, nnz):
    """Algorithmic bytes of one PDHG iteration over an m x n LP: the primal
    step reads AT (row pointers 4 n, 12 B per entry), y once (8 m), c, l, u,
    x, x_sum (40 n) and writes x, xbar, x_sum (24 n); the dual step reads A
    (4 m, 12 B per entry), xbar once (8 n), AL, AU, y, y_sum (32 m) and
    writes y, y_sum (16 m)."""
    return 24 * nnz + 76 * n + 60 * m


def measure_pdhg_iterations(gp_ffi, dlp, lp, stream, launches, per_launch=64):
    """Times the solver's iteration kernels (k_rows<OpPrimal> over AT, then
    k_rows<OpDual> over A, per iteration) on the original LP, `per_launch`
    iterations per gpk_pdhg_iterations call, back to back on one stream with
    HIP events; the two sparse products alone (gpk_pdhg_spmv over A and AT)
    are timed apart.  The iterates are not those of a solve (unscaled LP,
    fixed small steps): the kernels' memory traffic does not depend on them."""
    import ctypes as C

    import numpy as np

    import gp_types as T

    class GpkRows(C.Structure):
        _fields_ = [("M", T.GpCsr), ("blk", C.c_void_p), ("blocks", C.c_int32)]

    m, n, nnz = lp["m"], lp["n"], len(lp["A_idx"])
    L = gp_ffi.lib()
    vp = C.c_void_p
    L.gpk_pdhg_row_blocks.restype = C.c_int32
    L.gpk_pdhg_row_blocks.argtypes = [vp, C.c_int32, vp]
    L.gpk_pdhg_iterations.argtypes = [GpkRows, GpkRows] + [vp] * 13 + [C.c_int32, C.c_int32, vp]
    L.gpk_pdhg_spmv.argtypes = [GpkRows, vp, vp, vp]
    keep = []

    def rows(g, host_ptr):
        ptr = np.ascontiguousarray(host_ptr, np.int32)
        blk = np.zeros(len(ptr), np.int32)
        nb = L.gpk_pdhg_row_blocks(ptr.ctypes.data, len(ptr) - 1, blk.ctypes.data)
        d = gp_ffi.DeviceArray(blk[:nb + 1])
        keep.append(d)
        return GpkRows(g, d.ptr, nb)

    A, AT = rows(dlp.A.g, lp["A_ptr"]), rows(dlp.AT.g, lp["AT_ptr"])
    v = dlp.vec
    x, xbar, xs = (gp_ffi.DeviceArray(np.zeros(n)) for _ in range(3))
    y, ys = (gp_ffi.DeviceArray(np.zeros(m)) for _ in range(2))
    step = gp_ffi.DeviceArray(np.array([1e-3, 1e-3, 0.0]))
    ck = gp_ffi.hip_check

    def iterate():
        ck(L.gpk_pdhg_iterations(A, AT, v["c"].ptr, v["l"].ptr, v["u"].ptr, v["AL"].ptr, v["AU"].ptr, x.ptr,
                                 xbar.ptr, xs.ptr, y.ptr, ys.ptr, x.ptr, y.ptr, step.ptr, 0, per_launch, stream.ptr),
           "iterations")

    def products():
        ck(L.gpk_pdhg_spmv(A, x.ptr, y.ptr, stream.ptr), "spmv A")
        ck(L.gpk_pdhg_spmv(AT, y.ptr, x.ptr, stream.ptr), "spmv AT")

    iterate()
    stream.sync()
    e = [gp_ffi.Event() for _ in range(3)]
    e[0].record(stream)
    for _ in range(launches):
        iterate()
    e[1].record(stream)
    for _ in range(launches * per_launch):
        products()
    e[2].record(stream)
    stream.sync()
    it_us = 1e3 * e[0].elapsed_ms(e[1]) / (launches * per_launch)
    prod_us = 1e3 * e[1].elapsed_ms(e[2]) / (launches * per_launch)
    b = pdhg_iteration_bytes(m, n, nnz)
    log(f"pdhg iteration kernels: {it_us:.1f} us/iteration ({b / (it_us * 1e-6) / 1e9:.0f} GB/s); "
        f"the two products alone {prod_us:.1f} us; blocks A {A.blocks}, AT {AT.blocks}")
    return {"bound": "hbm", "kernels": "k_rows<OpPrimal> over AT + k_rows<OpDual> over A",
            "workload": f"original {m}x{n} LP, {nnz} nnz, {launches} x {per_launch} iterations",
            "us_per_iteration": it_us, "bytes_per_iteration": b, "achieved": b / (it_us * 1e-6) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b / (it_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "us_two_products": prod_us, "row_blocks": [A.blocks, AT.blocks]}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_activity_compaction_transpose_1Mx1.5M.md")


def profiled_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by a device copy of
    known size), or None when the summary has no row for it."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == f"`{kernel}`":
            try:
                return float(cells[6]) * 1e6
            except ValueError:
                return None
    return None


def activity_traffic():
    """HBM bytes of one row-activity call (k_pack_lu + k_row_activity_lu) from
    the committed PMC summary, or None."""
    parts = [profiled_traffic(k) for k in ("k_pack_lu", "k_row_activity_lu")]
    return None if any(p is None for p in parts) else sum(parts)


def transpose_traffic():
    """HBM bytes of one transpose from the committed PMC summary, or None."""
    per_pass = [profiled_traffic(k) for k in ("k_radix_hist<unsigned int>", "k_scan_tiles<int>", "k_scan_tile_sums",
                                              "k_scan_down<int>", "k_radix_scatter<unsigned int, false>")]
    once = [profiled_traffic(k) for k in ("k_col_bounds", "k_expand_rows", "k_transpose_gather")]
    if any(p is None for p in per_pass + once):
        return None
    return 3 * sum(per_pass) + sum(once)


def compaction_traffic():
    """HBM bytes of one compaction call: the sum over its kernels in the
    committed PMC summary, or None when one of them is missing."""
    parts = [profiled_traffic(k) for k in ("k_build_rank", "k_compact_map", "k_scan_small<int, 0>",
                                           "k_compact_write<false>")]
    return None if any(p is None for p in parts) else sum(parts)


def _upload_pool(gp_ffi, lps):
    t = time.time()
    d = [gp_ffi.DeviceLp(lp) for lp in lps]
    log(f"uploaded {len(d)} LPs in {time.time() - t:.1f}s")
    return d


def _worker_candidates(args):
    return tuple(int(w) for w in args.workers.split(",") if w.strip())


def bench_batched(args):
    """4,096 independent LPs presolved together (every job arrives at t = 0)
    by the serving runtime, against the same LPs presolved one at a time
    through the plain entry point (gp_presolve, synchronous, hipMalloc'ed
    reduced LP).  The worker count comes from the autotuner on a 256-LP
    sample."""
    import numpy as np

    import gp_ffi
    import workloads

    t = time.time()
    lps = workloads.batched_pool(args.batch_lps)
    ms = np.array([lp["m"] for lp in lps])
    nnz = np.array([len(lp["A_idx"]) for lp in lps])
    log(f"generated {len(lps)} LPs (sum m {ms.sum()}, sum nnz {nnz.sum()}) in {time.time() - t:.1f}s")
    dlps = _upload_pool(gp_ffi, lps)
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    sample = np.random.default_rng(11).choice(len(dlps), size=min(256, len(dlps)), replace=False).astype(np.int32)

    def run(w, jobs=sample):
        return gp_ffi.serve_trace(dlps, jobs, prm, gp_ffi.serve_config(workers=w, postsolve=False))

    run(4, sample[:32])  # warm-up: contexts, pools, code objects
    best, table = workloads.autotune_workers(run, _worker_candidates(args))
    log(f"autotune (LP/s on a {len(sample)}-LP sample): {table} -> {best} workers")
    allj = np.arange(len(dlps), dtype=np.int32)
    res = run(best, allj)
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise SystemExit(f"bench.py: {len(bad)} batched jobs failed")
    st = workloads.latency_stats(res)
    log(f"batched: {st}")
    t = time.perf_counter()
    for d in dlps:
        h = gp_ffi.Presolved(d, prm, keep_reduced=True)
        h.free()
    plain_s = time.perf_counter() - t
    plain = len(dlps) / plain_s
    log(f"plain one at a time: {plain:.1f} LP/s ({plain_s:.1f} s)")
    out = {
        "metric": "batched presolve LPs/s vs one-at-a-time plain presolve",
        "value": st["lp_per_s"], "unit": "LP/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (seeded, dyadic values, planted reductions)",
        "config": {"workload": "batched presolve", "lps": len(dlps), "m": "logU[100,20000]", "n": "1.5 m",
                   "row_len": "Poisson(6), >= 1", "sum_m": int(ms.sum()), "sum_nnz": int(nnz.sum()),
                   "workers": best},
        "makespan_s": st["makespan_s"],
        "plain_one_at_a_time": {"value": plain, "unit": "LP/s", "seconds": plain_s},
        "speedup_vs_plain": st["lp_per_s"] / plain,
        "autotune_lp_per_s": {str(k): v for k, v in table.items()},
    }
    print(json.dumps(out), flush=True)


def bench_serving(args):
    """Poisson arrivals of presolve + postsolve requests (seeded reduced
    point in place of a solver) served with continuous admission under a
    200 GB budget; the requests reference a pool of pre-generated LPs drawn
    from the request size distribution, resident in HBM.  The same trace's
    first requests are also served one at a time (one worker) for the serial
    baseline."""
    import numpy as np

    import gp_ffi
    import workloads

    t = time.time()
    lps = workloads.serving_pool(args.serve_pool)
    ms = np.array([lp["m"] for lp in lps])
    nnz = np.array([len(lp["A_idx"]) for lp in lps])
    log(f"generated a pool of {len(lps)} LPs (mean m {ms.mean():.0f}, mean nnz {nnz.mean():.0f}) "
        f"in {time.time() - t:.1f}s")
    dlps = _upload_pool(gp_ffi, lps)
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    arrival, lp_of_job = workloads.poisson_trace(args.serve_rate, args.serve_seconds, len(dlps))
    budget = 200e9

    def run(w, jobs, arr=None):
        cfg = gp_ffi.serve_config(workers=w, postsolve=True, hbm_budget_bytes=budget, seed=99)
        return gp_ffi.serve_trace(dlps, jobs, prm, cfg, arr)

    sample = lp_of_job[:64]
    run(4, sample[:16])  # warm-up
    best, table = workloads.autotune_workers(lambda w: run(w, sample), _worker_candidates(args))
    log(f"autotune (LP/s, 64 jobs at once): {table} -> {best} workers")
    res = run(best, lp_of_job, arrival)
    bad = [r for r in res if r["status"] != 0]
    if bad:
        raise SystemExit(f"bench.py: {len(bad)} served jobs failed")
    st = workloads.latency_stats(res)
    log(f"served: {st}")
    k = int(np.searchsorted(arrival, min(10.0, args.serve_seconds)))
    ser = workloads.latency_stats(run(1, lp_of_job[:k], arrival[:k]))
    log(f"serial (first {k} requests, one worker): {ser}")
    out = {
        "metric": "served presolve+postsolve requests: LPs/s and p50/p99 latency",
        "value": st["lp_per_s"], "unit": "LP/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (seeded, dyadic values, planted reductions)",
        "config": {"workload": "serving", "rate_per_s": args.serve_rate, "seconds": args.serve_seconds,
                   "requests": st["jobs"], "pool_lps": len(dlps), "m": "logU[1000,500000]", "n": "1.5 m",
                   "row_len": "U[4,10]", "pool_mean_m": float(ms.mean()), "pool_mean_nnz": float(nnz.mean()),
                   "hbm_budget_bytes": budget, "workers": best},
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
        "serial_one_worker": dict(ser, requests=k, sample=f"first {k} requests of the same trace"),
        "capacity_lp_per_s_by_workers": {str(k2): v for k2, v in table.items()},
    }
    print(json.dumps(out), flush=True)


def bench_small(args):
    """Latency of one small LP from the batched size distribution (m rows,
    n = 1.5 m, mean row length 6) presolved one at a time through the plain
    entry point: ms per presolve, and its launch count under a kernel trace."""
    import numpy as np

    import gp_ffi
    import lp_synth

    m = args.small_rows
    lp = lp_synth.generate(m, int(round(1.5 * m)), seed=7 + m, mean_row_len=6.0)
    d = gp_ffi.DeviceLp(lp)
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    for _ in range(args.warmup):
        gp_ffi.Presolved(d, prm, keep_reduced=True).free()
    t = time.perf_counter()
    for _ in range(args.steps):
        gp_ffi.Presolved(d, prm, keep_reduced=True).free()
    ms = 1e3 * (time.perf_counter() - t) / args.steps
    print(json.dumps({"metric": "small-LP presolve latency", "value": ms, "unit": "ms", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "dtype": "f64",
                      "data": "synthetic (seeded, dyadic values, planted reductions)",
                      "config": {"workload": "small LP, plain presolve", "m": m, "n": int(round(1.5 * m)),
                                 "nnz": int(len(lp["A_idx"]))}}), flush=True)


def bench_requests(args):
    """Served requests processed one at a time: the first --requests requests
    of the serving workload's trace (m log-uniform in [1,000, 500,000],
    n = 1.5 m, 4-10 entries per general row), each generated on the device,
    presolved, given its seeded reduced point and postsolved through
    gp_plain_job_spec, as the serial leg of the serving object does.  LP/s of
    the serial processing; under a kernel trace, the kernels and launches one
    request costs."""
    import gp_ffi

    sizes = _sizes(4, args.requests, 1_000, 500_000)
    specs = [gp_ffi.lp_spec(0xC000 + k, int(m), row_len=(4, 10)) for k, m in enumerate(sizes)]
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    for k in range(min(args.warmup, len(specs))):
        gp_ffi.plain_job_spec(specs[k], prm, 99, k, True)
    busy = 0.0
    for k, spec in enumerate(specs):
        p = gp_ffi.plain_job_spec(spec, prm, 99, k, True)
        busy += p["t_end"] - p["t_start"]
    print(json.dumps({"metric": "served requests, one at a time", "value": len(specs) / busy, "unit": "LP/s",
                      "n_gpus": 1, "steps": len(specs), "warmup": args.warmup, "higher_is_better": True,
                      "dtype": "f64", "data": "synthetic (seeded, generated on the device)",
                      "config": {"workload": "serial served requests", "requests": len(specs),
                                 "mean_m": float(sizes.mean())}}), flush=True)


def _sizes(seed, count, lo, hi):
    import numpy as np

    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)


def long_rows(gp_ffi, args):
    """The single large LP with row lengths U[2,64] (1M x 1.5M, about 30.5M
    nonzeros): the working LP keeps more than 10M nonzeros into the fast
    phase, so the batched singleton closure runs at its default threshold.
    Presolve time (mean of 3 after one warm-up) and the reduced sizes, which
    must equal the oracle's (tests/golden/long_rows_presolve.json)."""
    import lp_synth

    t = time.time()
    lp = lp_synth.config1(seed=1, row_len=(2, 64))
    nnz = len(lp["A_idx"])
    log(f"long rows: generated LP {lp['m']}x{lp['n']} nnz={nnz} in {time.time() - t:.1f}s")
    d = gp_ffi.DeviceLp(lp)
    del lp
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    gp_ffi.Presolved(d, prm, keep_reduced=True).free()
    times, s = [], None
    for _ in range(3):
        gp_ffi.hip_check(gp_ffi.hip().hipDeviceSynchronize(), "sync")
        t = time.perf_counter()
        h = gp_ffi.Presolved(d, prm, keep_reduced=True)
        s = h.summary()
        gp_ffi.hip_check(gp_ffi.hip().hipDeviceSynchronize(), "sync")
        times.append(1e3 * (time.perf_counter() - t))
        h.free()
    with open(os.path.join(ROOT, "tests", "golden", "long_rows_presolve.json")) as f:
        want = json.load(f)["summary"]
    got = [s["reduced_rows"], s["reduced_cols"], s["reduced_nnz"]]
    if got != [want["reduced_rows"], want["reduced_cols"], want["reduced_nnz"]]:
        raise SystemExit(f"bench.py: long-row presolve gives {got}, the oracle {want}")
    ms = sum(times) / len(times)
    log(f"long rows: presolve {ms:.1f} ms (runs {[round(x, 1) for x in times]}) -> {got}")
    return {"ms_per_presolve": ms, "lp_per_s": 1e3 / ms, "m": 1_000_000, "n": 1_500_000, "nnz": nnz,
            "row_len": "U[2,64]", "reduced": got, "iterations": s["iterations"],
            "check": "reduced sizes equal the oracle's (record digests are compared in the GPU tests)"}


def end_to_end(gp_ffi, dlp, lp, prm, tol):
    """The solver hand-off: presolve, the reduced LP solved on the device by
    gp_pdhg_solve without leaving HBM, and the one-shot postsolve of that
    solution against the original LP; then the same solver on the original
    LP (no presolve), for the solve-time ratio.  Each stage is timed
    wall-clock with the device synchronised around it."""
    import ctypes as C

    import numpy as np

    import gp_types as T

    H = gp_ffi.hip()
    m, n, nnz = lp["m"], lp["n"], len(lp["A_idx"])
    sync = lambda: gp_ffi.hip_check(H.hipDeviceSynchronize(), "sync")
    sync()
    t = time.perf_counter()
    h = gp_ffi.Presolved(dlp, prm, keep_reduced=True)
    sync()
    pre_ms = 1e3 * (time.perf_counter() - t)
    try:
        view = h.reduced_view()
        m1, n1, nnz1 = view.A.rows, view.A.cols, view.A.nnz
        t = time.perf_counter()
        res, (dx, dy, dz) = gp_ffi.pdhg_solve_view(view, gp_ffi.pdhg_params(tol=tol))
        sync()
        solve_ms = 1e3 * (time.perf_counter() - t)
        rec = h.record()
    finally:
        h.free()
    rk = T.RecordKeepAlive(rec)
    out = (gp_ffi.DeviceArray(size=n), gp_ffi.DeviceArray(size=m), gp_ffi.DeviceArray(size=n))
    sync()
    t = time.perf_counter()
    gp_ffi._check(gp_ffi.lib().gp_postsolve(C.byref(rk.view), dx.ptr, dy.ptr, dz.ptr, C.byref(dlp.g),
                                            out[0].ptr, out[1].ptr, out[2].ptr))
    sync()
    post_ms = 1e3 * (time.perf_counter() - t)
    x = out[0].numpy()
    obj = float(np.dot(lp["c"], x)) + float(lp.get("obj", 0.0))
    log(f"end to end: presolve {pre_ms:.1f} ms, solve {solve_ms:.1f} ms ({res['iterations']} iterations, "
        f"reduced {m1}x{n1}), postsolve {post_ms:.1f} ms; objective {obj:.10g}")
    t = time.perf_counter()
    direct, _ = gp_ffi.pdhg_solve_view(dlp.g, gp_ffi.pdhg_params(tol=tol, max_time=120.0))
    sync()
    direct_ms = 1e3 * (time.perf_counter() - t)
    log(f"no presolve: solve {direct_ms:.0f} ms ({direct['iterations']} iterations, status {direct['status']}); "
        f"objective {direct['primal_obj']:.10g}")
    us_it = 1e6 * direct["elapsed_seconds"] / max(1, direct["iterations"])
    total = pre_ms + solve_ms + post_ms
    # The same three steps as one call (gp_presolve_solve): the postsolve is
    # prepared on a second stream while the solver runs; best of three.
    prm_s = gp_ffi.pdhg_params(tol=tol)
    pipe_out = (gp_ffi.DeviceArray(size=n), gp_ffi.DeviceArray(size=m), gp_ffi.DeviceArray(size=n))
    pipes = []
    for _ in range(3):
        sync()
        t = time.perf_counter()
        pr, _ = gp_ffi.presolve_solve(dlp, prm, prm_s, pipe_out)
        sync()
        pipes.append((1e3 * (time.perf_counter() - t), pr))
    pipe_ms, pr = min(pipes, key=lambda q: q[0])
    if not np.array_equal(pipe_out[0].numpy(), x):
        raise SystemExit("bench.py: gp_presolve_solve differs from the three steps")
    log(f"pipelined (gp_presolve_solve): {pipe_ms:.1f} ms (presolve {1e3 * pr['presolve_seconds']:.1f}, solve "
        f"{1e3 * pr['solve_seconds']:.1f} with the postsolve prepared beside it in "
        f"{1e3 * pr['prepare_seconds']:.1f}, postsolve {1e3 * pr['postsolve_seconds']:.1f})")
    return {"tol": tol, "presolve_ms": pre_ms, "solve_ms": solve_ms, "postsolve_ms": post_ms, "total_ms": total,
            "reduced": [m1, n1, nnz1], "solve_status": res["status"], "solve_iterations": res["iterations"],
            "solve_restarts": res["restarts"], "objective": obj,
            "reduced_kkt": [res["primal_res"], res["dual_res"], res["gap"]],
            "no_presolve_solve_ms": direct_ms, "no_presolve_status": direct["status"],
            "no_presolve_iterations": direct["iterations"], "no_presolve_objective": direct["primal_obj"],
            "no_presolve_kkt": [direct["primal_res"], direct["dual_res"], direct["gap"]],
            "speedup_total_vs_no_presolve": direct_ms / total,
            "pipelined_total_ms": pipe_ms,
            "pipelined_parts_ms": {k: 1e3 * pr[k + "_seconds"] for k in ("presolve", "solve", "prepare", "postsolve")},
            "speedup_pipelined_vs_no_presolve": direct_ms / pipe_ms,
            "no_presolve_us_per_iteration_with_checks": us_it,
            "iteration_roofline": measure_pdhg_iterations(gp_ffi, dlp, lp, gp_ffi.Stream(), 10)}


def many_batched(gp_ffi, args):
    """The batched workload on LPs generated on the device: 4,096 LPs,
    m log-uniform in [100, 20,000], n = 1.5 m, row lengths Poisson(6) with
    minimum 1, all submitted at once to the serving runtime (presolve only;
    worker count from a 256-LP sample), against the same LPs presolved one
    at a time through the plain synchronous entry points."""
    import numpy as np

    import workloads

    sizes = _sizes(3, args.batch_lps, 100, 20_000)
    t = time.time()
    specs = [gp_ffi.lp_spec(0xB000 + k, int(m), mean_len=6.0, min_len=1) for k, m in enumerate(sizes)]
    lps = [gp_ffi.GeneratedLp(s) for s in specs]
    nnz = sum(g.g.A.nnz for g in lps)
    log(f"batched: generated {len(lps)} LPs on the device (sum m {int(sizes.sum())}, sum nnz {nnz}) "
        f"in {time.time() - t:.1f}s")
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    sample = np.random.default_rng(11).choice(len(lps), size=min(256, len(lps)), replace=False).astype(np.int32)

    def run(w, jobs=sample):
        return gp_ffi.serve_trace(lps, jobs, prm, gp_ffi.serve_config(workers=w, postsolve=False))

    run(4, sample[:32])
    best, table = workloads.autotune_workers(run, _worker_candidates(args))
    res = run(best, np.arange(len(lps), dtype=np.int32))
    if any(r["status"] != 0 for r in res):
        raise SystemExit("bench.py: batched jobs failed")
    st = workloads.latency_stats(res)
    t = time.perf_counter()
    for k, g in enumerate(lps):
        p = gp_ffi.plain_job(g, prm, 0, k, postsolve=False)
        if (p["status"], p["reduced_rows"], p["reduced_nnz"]) != (0, res[k]["reduced_rows"], res[k]["reduced_nnz"]):
            raise SystemExit(f"bench.py: batched job {k} differs from its plain run")
    plain_s = time.perf_counter() - t
    for g in lps:
        g.free()
    plain = len(lps) / plain_s
    log(f"batched: {st['lp_per_s']:.0f} LP/s with {best} workers (autotune {table}); plain one at a time "
        f"{plain:.0f} LP/s -> {st['lp_per_s'] / plain:.2f}x")
    return {"value": st["lp_per_s"], "unit": "LP/s", "lps": len(lps), "sum_m": int(sizes.sum()), "sum_nnz": int(nnz),
            "m": "logU[100,20000]", "n": "1.5 m", "row_len": "Poisson(6), >= 1", "workers": best,
            "makespan_s": st["makespan_s"], "plain_one_at_a_time_lp_per_s": plain,
            "speedup_vs_plain": st["lp_per_s"] / plain, "autotune_lp_per_s": {str(k): v for k, v in table.items()},
            "check": "every job's reduced dimensions equal its plain run's"}


def many_serving(gp_ffi, args):
    """The serving workload: Poisson(rate) presolve + postsolve requests for
    --serve-seconds, each generating its own LP on a worker from the request
    size distribution (m log-uniform in [1,000, 500,000], n = 1.5 m, 4-10
    entries per general row), admitted under a 200 GB budget; p50/p99
    latency (arrival to result, generation included) and LP/s.  Beside it:
    the capacity with a sample of requests all at once, and serial plain
    processing (gp_plain_job_spec, one at a time) of the same requests,
    both as capacity and replayed against the trace's first arrivals."""
    import numpy as np

    import workloads

    rate, secs = args.serve_rate, args.serve_seconds
    arrival, _ = workloads.poisson_trace(rate, secs, 1, seed=5)
    sizes = _sizes(4, len(arrival), 1_000, 500_000)
    specs = [gp_ffi.lp_spec(0xC000 + k, int(m), row_len=(4, 10)) for k, m in enumerate(sizes)]
    prm = gp_ffi.params(enable_structure_specific_rules=0)
    budget = 200e9
    k_cap = min(len(specs), 400)

    def cfg_for(w):
        return gp_ffi.serve_config(workers=w, postsolve=True, hbm_budget_bytes=budget, seed=99)

    gp_ffi.serve_spec_trace(specs[:16], prm, cfg_for(4))  # warm-up
    if args.serve_workers > 0:
        workers, table = args.serve_workers, {}
    else:  # capacity of the first requests at each worker count
        cands = tuple(w for w in _worker_candidates(args) if w >= 8) or _worker_candidates(args)
        workers, table = workloads.autotune_workers(lambda w: gp_ffi.serve_spec_trace(specs[:k_cap], prm, cfg_for(w)),
                                                    cands)
    cfg = cfg_for(workers)
    cap = workloads.latency_stats(gp_ffi.serve_spec_trace(specs[:k_cap], prm, cfg))
    res = gp_ffi.serve_spec_trace(specs, prm, cfg, arrival)
    if any(r["status"] != 0 for r in res):
        raise SystemExit("bench.py: served requests failed")
    st = workloads.latency_stats(res)
    gen_ms = 1e3 * float(np.mean([r["t_ready"] - r["t_start"] for r in res]))
    # serial plain processing, one request at a time, of the first requests
    t_end, lat, busy = 0.0, [], 0.0
    k_ser = min(len(specs), k_cap)
    for k in range(k_ser):
        p = gp_ffi.plain_job_spec(specs[k], prm, 99, k, True)
        if p["checksum"] != res[k]["checksum"]:
            raise SystemExit(f"bench.py: served request {k} differs from its plain run")
        dt = p["t_end"] - p["t_start"]
        busy += dt
        t_end = max(t_end, arrival[k]) + dt
        lat.append(t_end - arrival[k])
    serial_cap = k_ser / busy
    lat = np.array(lat)
    log(f"serving: {st['lp_per_s']:.0f} LP/s at {rate:.0f} req/s offered, p50 {st['p50_ms']:.1f} ms, "
        f"p99 {st['p99_ms']:.1f} ms ({len(res)} requests, generation {gen_ms:.1f} ms mean); capacity "
        f"{cap['lp_per_s']:.0f} LP/s vs serial plain {serial_cap:.0f} LP/s -> {cap['lp_per_s'] / serial_cap:.2f}x")
    return {"value": st["lp_per_s"], "unit": "LP/s", "requests": len(res), "rate_per_s": rate, "seconds": secs,
            "m": "logU[1000,500000]", "n": "1.5 m", "row_len": "U[4,10]", "mean_m": float(sizes.mean()),
            "hbm_budget_bytes": budget, "workers": workers,
            "autotune_capacity_lp_per_s": {str(k): v for k, v in table.items()}, "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "max_ms": st["max_ms"], "generation_ms_mean": gen_ms,
            "capacity_lp_per_s": cap["lp_per_s"], "capacity_sample": k_cap,
            "serial_plain_capacity_lp_per_s": serial_cap, "capacity_vs_serial": cap["lp_per_s"] / serial_cap,
            "serial_plain_trace": {"requests": k_ser, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                                   "p99_ms": 1e3 * float(np.percentile(lat, 99))},
            "check": "every compared request's postsolve checksum equals its plain run's"}


def cpu_leg(args):
    """One CPU baseline leg in its own process: the oracle build named by
    ORACLE_LIB presolves the plain workload's LP once to warm up, then
    --cpu-runs more times; prints the times and the reduced sizes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lp_synth
    import oracle_ffi

    lp = lp_synth.config1(seed=args.seed)
    prm = oracle_ffi.params(enable_structure_specific_rules=0)
    times = []
    for k in range(args.cpu_runs + 1):
        t = time.perf_counter()
        s, _, _ = oracle_ffi.presolve(lp, prm, False)
        if k > 0:
            times.append(time.perf_counter() - t)
    print(json.dumps({"times": times, "reduced": [s["reduced_rows"], s["reduced_cols"], s["reduced_nnz"]]}),
          flush=True)


def _cpu_model():
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def cpu_baseline(args, summary, m, n):
    """The oracle (serial C++ restatement of the reference's passes) on the
    host cores: single thread, and the OpenMP build on all the threads the
    job may use (OMP_NUM_THREADS, else the usable cores); median of
    --cpu-runs runs after one warm-up, each leg in its own process."""
    import statistics
    import subprocess

    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    legs = {}
    for name, lib, nt in (("single_thread", "liboracle.so", 1), ("all_core", "liboracle_omp.so", threads)):
        env = dict(os.environ, ORACLE_LIB=os.path.join(ROOT, "oracle", "build", lib), OMP_NUM_THREADS=str(nt))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", "--seed", str(args.seed),
                            "--cpu-runs", str(args.cpu_runs)], capture_output=True, text=True, env=env, timeout=600)
        if r.returncode != 0:
            raise SystemExit(f"bench.py: cpu baseline leg {name} failed: {r.stderr[-2000:]}")
        res = json.loads(r.stdout.strip().splitlines()[-1])
        if res["reduced"] != [summary["reduced_rows"], summary["reduced_cols"], summary["reduced_nnz"]]:
            raise SystemExit("bench.py: GPU and oracle reduced sizes differ")
        med = statistics.median(res["times"])
        legs[name] = {"seconds_median": med, "seconds": res["times"], "threads": nt, "lp_per_s": 1.0 / med}
        log(f"cpu oracle presolve ({name}, {nt} threads): median {med:.2f} s of {res['times']}")
    best = legs["all_core"]
    return {"value": best["lp_per_s"], "unit": "LP/s", "cores": best["threads"], "kind": "port",
            "sample": f"full presolves of the same {m}x{n} LP (seed {args.seed}) by the C++ oracle, median of "
                      f"{args.cpu_runs} after one warm-up; all-core = OpenMP build (transpose, compaction and "
                      f"row activity threaded, identical results)",
            "single_thread": legs["single_thread"], "all_core": legs["all_core"],
            "host": {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "cpu_model": _cpu_model()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--activity-launches", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the row-activity and compaction kernel timings plus a 256 MiB device copy "
                         "(for hardware-counter runs)")
    ap.add_argument("--workload", choices=("plain", "batched", "serving", "small", "requests"), default="plain",
                    help="plain: the single large LP (default); batched: 4,096 small LPs at once against "
                         "one-at-a-time plain presolve; serving: Poisson presolve+postsolve requests; "
                         "small: latency of one small LP (--small-rows); requests: the serving workload's "
                         "first --requests requests one at a time")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--small-rows", type=int, default=1400)
    ap.add_argument("--batch-lps", type=int, default=4096)
    ap.add_argument("--workers", type=str, default="1,2,4,8,12,16,20,24",
                    help="worker counts the autotuner tries (batched and serving)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool", type=int, default=48)
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed oracle runs per CPU baseline leg")
    ap.add_argument("--serve-workers", type=int, default=0,
                    help="serving workers (0: the count with the highest capacity on the first requests)")
    ap.add_argument("--no-many", action="store_true",
                    help="plain workload only (no batched / serving objects on the line)")
    ap.add_argument("--no-solve", action="store_true", help="skip the presolve + solve + postsolve object")
    ap.add_argument("--solve-tol", type=float, default=1e-4, help="relative KKT tolerance of gp_pdhg_solve")
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        return cpu_leg(args)
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this library runs one LP presolve per GPU and "
              "the benchmark measures one GPU only (N=1).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "small":
        return bench_small(args)
    if args.workload == "requests":
        return bench_requests(args)

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

    if args.kernels_only:
        act, comp, tr = kernel_rooflines(gp_ffi, dlp, lp, args.activity_launches)
        pdhg = measure_pdhg_iterations(gp_ffi, dlp, lp, gp_ffi.Stream(), 10)
        calibration_copy(gp_ffi)
        print(json.dumps({"kernels_only": True, "row_activity": act, "compaction": comp, "transpose": tr,
                          "pdhg_iterations": pdhg}), flush=True)
        return

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
    ps = gp_ffi.PreparedPostsolve(rec, x, y, z, dlp)
    post = []
    for _ in range(3):
        t = time.perf_counter()
        ps.run()
        post.append(1e3 * (time.perf_counter() - t))
    post_ms = min(post)
    # The prepared form: tape, maps and schedules built once, each run replays
    # the captured graph of launches.
    t = time.perf_counter()
    plan = gp_ffi.PostsolvePlan(rec, original_dev=dlp, use_graph=True)
    prep_ms = 1e3 * (time.perf_counter() - t)
    runs = []
    for _ in range(4):
        t = time.perf_counter()
        ps.run_plan(plan)
        runs.append(1e3 * (time.perf_counter() - t))
    plan_ms = min(runs[1:])  # the first run captures the graph
    shape = plan.shape()
    plan.free()
    last.free()
    e2e = None if args.no_solve else end_to_end(gp_ffi, dlp, lp, prm, args.solve_tol)
    log(f"postsolve (inputs resident; record tape uploaded by the call): {post_ms:.1f} ms (runs {post}); "
        f"prepared: {prep_ms:.1f} ms once, then {plan_ms:.1f} ms per run (runs {runs})")

    # Postsolve model: each replay pass (primal, dual) reads the tape once, the
    # outputs x, z (n) and y (m) are cleared and written and the reduced point
    # scattered; the passes are chains of dependent levels, so the bound is the
    # level count times a level's latency, not bandwidth.
    levels = shape["primal_levels"] + shape["dual_levels"]
    post_bytes = 2 * shape["tape_bytes"] + 8 * 2 * (2 * n + m) + 8 * (2 * red["n"] + red["m"])
    post_roof = {"bound": "latency (dependent replay levels)", "records": shape["records"],
                 "primal_levels": shape["primal_levels"], "dual_levels": shape["dual_levels"],
                 "tape_bytes": shape["tape_bytes"], "ms_prepared_run": plan_ms,
                 "us_per_level": 1e3 * plan_ms / max(1, levels), "algorithmic_bytes": post_bytes,
                 "achieved": post_bytes / (plan_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": post_bytes / (plan_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"postsolve: {shape['records']} records, {shape['primal_levels']} + {shape['dual_levels']} levels, "
        f"{post_roof['us_per_level']:.1f} us per level")

    act, comp, tr = kernel_rooflines(gp_ffi, dlp, lp, args.activity_launches)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, summary, m, n)

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
                   "iterations": summary["iterations"], "postsolve_ms": post_ms,
                   "postsolve_prepare_ms": prep_ms, "postsolve_prepared_run_ms": plan_ms},
        "roofline": {"bound": "hbm",
                     "kernel": "compaction: k_build_rank + k_compact_map + tile-offset scan + k_compact_write",
                     "achieved": comp["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": comp["achieved"] / HBM_PEAK_GBS, "traffic": compaction_traffic(),
                     "us_per_call": comp["us"], "algorithmic_bytes": comp["bytes"],
                     "compulsory_bytes": comp["compulsory_bytes"],
                     "compulsory_frac": comp["compulsory_achieved"] / HBM_PEAK_GBS,
                     "workload": comp["workload"], "kernels_us": comp["kernels_us"]},
        "roofline_compaction_pair": {"bound": "hbm",
                                     "kernel": "paired compaction of A and AT: 2 x k_build_rank + "
                                               "k_compact_map_pair + tile-offset scan + k_compact_write_pair",
                                     "achieved": comp["pair"]["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": comp["pair"]["achieved"] / HBM_PEAK_GBS, "traffic": None,
                                     "us_per_call": comp["pair"]["us"],
                                     "algorithmic_bytes": comp["pair"]["bytes"],
                                     "workload": comp["pair"]["workload"]},
        "roofline_activity": {"bound": "hbm", "kernel": "k_pack_lu + k_row_activity_lu (one gp_row_activity call)",
                              "achieved": act["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": act["achieved"] / HBM_PEAK_GBS,
                              "traffic": activity_traffic(),
                              "us_per_call": act["us"], "algorithmic_bytes": act["bytes"],
                              "effective_bytes": act["effective_bytes"],
                              "effective_frac": act["effective_achieved"] / HBM_PEAK_GBS, **act["kernels"]},
        "roofline_transpose": {"bound": "hbm", "kernel": "transpose: 3 LSD sort passes + row pointers from the "
                                                          "sorted keys + entry rows + gather",
                               "achieved": tr["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": tr["achieved"] / HBM_PEAK_GBS, "traffic": transpose_traffic(),
                               "us_per_call": tr["us"], "algorithmic_bytes": tr["bytes"],
                               "workload": tr["workload"], "kernels_us": tr["kernels_us"]},
        "roofline_postsolve": post_roof,
        "end_to_end": e2e,
        "cpu_baseline": cpu,
    }
    if not args.no_many:
        del dlp
        out["long_rows"] = long_rows(gp_ffi, args)
        out["batched"] = many_batched(gp_ffi, args)
        out["serving"] = many_serving(gp_ffi, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
