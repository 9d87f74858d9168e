"""Benchmark: SpMV and power-method throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload all|powerlaw|poisson|stencil9|batch|serve|coo]

The default (`all`) prints ONE JSON line whose top level is the power-law
SpMV (N = 4 194 304, ~67M nnz, FP64 values, int32 indices) and whose
"configs" key carries the other workloads measured in the same run:

- config0_poisson: 2D 5-point Poisson 1024 x 1024 (84 MB per SpMV, small
  enough for the 256 MB Infinity Cache).  Its roofline figure comes from
  SpMVs rotating over several independent copies of the matrix and vectors
  whose arrays together exceed that cache, so every SpMV reads HBM; the rate
  of repeated SpMVs on one copy is reported beside it as `warm`.
- config2_stencil9: 9-point stencil 2896^2 + 2 random columns per row
  (~92M nnz), FP64 and FP32 storage, FP32 error against FP64.
- config3_batch: 512 uneven SPD power methods stepped together (solves/s
  against the plain FP64 per-matrix GPU loop, eigenvalue error against FP64).
- config4_serve: Poisson-arrival requests on the continuous-batching server
  (p50/p99 latency and solves/s against the FP64 per-request GPU path, the
  rate the server sustains, and each served result's error against FP64).

A step of an SpMV workload is one SpMV y = A x over the whole matrix with A,
x, y resident in HBM: W untimed steps, then K timed steps launched back to
back (no event or marker between them) between two device synchronisations;
`value` = algorithmic bytes per SpMV / wall time per step.  The algorithmic
bytes of an FP64/int32 CRS SpMV are 12 Z + 4 (N+1) + 16 N (values and column
indices once, row offsets once, x and y once).

All four GPU formats are measured (CRS, merge path, CSB, BlockCOH); the
fastest gives `value`, its dominant kernel's HIP-event time gives
`roofline.achieved`, and `roofline.traffic` comes from the committed rocprof
PMC summary for that workload and format.  The CPU baseline is the oracle's
restatements of the reference's CRSOMP::mv and CRS_Merge::mv (built at -O3
like the reference) timed on the host's cores.
"""
import argparse
import json
import math
import os
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20  # MI355X Infinity Cache
METRIC = "SpMV GB/s (% of HBM peak); batched power-method solves/s + p99 latency"
# host threads of the CPU baseline: the CPU share of one GPU on the GPU box
CPU_THREADS = 16


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(n, z, value_bytes=8, vector_bytes=8):
    return value_bytes * z + 4 * z + 4 * (n + 1) + 2 * vector_bytes * n


def poisson_crs(m, n):
    """2D 5-point Poisson matrix in CRS (4 on the diagonal, -1 to the four
    neighbours; row r = i*m + j, columns ascending), built with numpy."""
    N = m * n
    r = np.arange(N, dtype=np.int64)
    j = r % m
    offs = [(-m, r >= m), (-1, j != 0), (0, np.ones(N, bool)), (1, j != m - 1), (m, r < N - m)]
    cols, vals, keep = [], [], []
    for d, mask in offs:
        cols.append(r + d)
        vals.append(np.full(N, 4.0 if d == 0 else -1.0))
        keep.append(mask)
    C, V, K = np.stack(cols, 1), np.stack(vals, 1), np.stack(keep, 1)
    rs = np.concatenate([[0], np.cumsum(K.sum(1))]).astype(np.int32)
    return N, rs, C[K].astype(np.int32), V[K]


def make_workload(name, stencil_random=2):
    import thesis_spmv_amd as pw
    t0 = time.time()
    if name == "powerlaw":
        n = 4194304
        rs, ci, dv = pw.gen_powerlaw(n, alpha=1.8, xm=7.3333, cap=50000, struct_seed=1, value_seed=2)
        desc = "power-law CRS N=4194304, Pareto(alpha=1.8, mean 16, cap 50000) row lengths, N(0,1) values"
    elif name == "stencil9":
        side = 2896
        n = side * side
        rs, ci, dv = pw.gen_stencil9(side, stencil_random, 3)
        desc = f"9-point stencil 2896x2896 + {stencil_random} random off-band nnz per row"
    elif name == "poisson":
        n, rs, ci, dv = poisson_crs(1024, 1024)
        desc = "2D 5-point Poisson 1024x1024"
    else:
        raise SystemExit(f"unknown workload {name}")
    return n, rs, ci, dv, desc, time.time() - t0


def _host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# read before OpenMP starts: with OMP_PROC_BIND set, libgomp pins the calling
# thread to one place, and the affinity mask then shows a single core
HOST_CORES_AT_START = _host_cores()


def cpu_baseline(n, rs, ci, dv, x, bytes_per_mv, threads, budget_s=5.0, power_budget_s=6.0, y_gpu=None):
    """The oracle's restatements of CRSOMP::mv (Env_Implementations/CRSOMP.hpp:110-123,
    OpenMP dynamic,512) and CRS_Merge::mv (SOA_Implementations/CRS_Merge.hpp:95-141,
    P = threads), built at -O3 like the reference, on `threads` host cores
    with OMP_PROC_BIND=close.  Protocol of the reference's SpMV driver
    (driver_spmv_input.cpp:94-107): 50 warm-up calls, then up to 1000 timed
    calls, median, each leg bounded to about `budget_s` seconds (warm-ups
    included).  The power iteration (CRSOMP.hpp:135-157) is timed as calls of
    100 iterations: 3 warm-ups and 10 timed calls as in the reference's
    Poisson timings, as many as fit `power_budget_s`.  `value` is the faster
    SpMV.  As the checker, the sequential oracle's CRS::mv result gives the
    relative max-norm error of the GPU's y (`y_gpu`)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    A = oracle.CRSMatrix(n, n, rs, ci, dv)
    B = oracle.baseline_lib()
    xs = np.ascontiguousarray(x, np.float64)
    y = np.zeros(n)

    def timed(fn, warm, most, budget):
        t_start = time.time()
        t0 = time.perf_counter()
        fn()
        one = time.perf_counter() - t0
        # warm-ups and timed calls within the budget (at least one of each)
        warm = max(1, min(warm, int(0.2 * budget / max(one, 1e-9))))
        for _ in range(warm - 1):
            fn()
        times = []
        while len(times) < most and (not times or time.time() - t_start < budget):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), warm, len(times)

    t_omp, w_omp, k_omp = timed(lambda: B.orc_crsomp_mv(threads, n, A.row_start, A.col_ind, A.data, xs, y),
                                50, 1000, budget_s)
    t_mrg, w_mrg, k_mrg = timed(lambda: B.orc_merge_mv(threads, threads, n, A.nnz, A.row_start, A.col_ind, A.data,
                                                        xs, y), 50, 1000, budget_s)
    px, py = np.array(xs), np.zeros(n)
    t_pow, w_pow, k_pow = timed(lambda: B.orc_crsomp_power_iteration(threads, n, A.row_start, A.col_ind, A.data,
                                                                     px, py, 100), 3, 10, power_budget_s)
    best = min(t_omp, t_mrg)
    err = None
    if y_gpu is not None:
        y_ref = oracle.crs_mv(A, xs)
        err = float(np.max(np.abs(y_gpu - y_ref)) / max(float(np.max(np.abs(y_ref))), 1e-300))
    return {"value": round(bytes_per_mv / best / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
            "host_cores_visible": HOST_CORES_AT_START, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "gpu_y_rel_err_vs_oracle": err,
            "sample": f"median of {k_omp} CRSOMP::mv and {k_mrg} CRS_Merge::mv (P={threads}) full SpMVs of the same "
                      f"matrix after {w_omp} / {w_mrg} warm-ups: {t_omp * 1e3:.3f} ms and {t_mrg * 1e3:.3f} ms "
                      f"(oracle restatements, gcc -O3 -fopenmp, {threads} threads, OMP_PROC_BIND=close); "
                      f"power iteration {t_pow / 100 * 1e3:.3f} ms/iteration (median of {k_pow} calls of 100 "
                      f"after {w_pow} warm-ups)",
            "crsomp_ms": round(t_omp * 1e3, 4), "crs_merge_ms": round(t_mrg * 1e3, 4),
            "power_ms_per_iteration": round(t_pow / 100 * 1e3, 4)}


# ---------------------------------------------------------------- SpMV configs

def bench_spmv(args, workload, cpu=True):
    """One SpMV workload (configs 0-2): every format, the best one's roofline,
    the gather probe, the device power method and the CPU baseline."""
    import thesis_spmv_amd as pw
    values = args.values or ("f64,f32" if workload == "stencil9" else "f64")
    n, rs, ci, dv, desc, gen_s = make_workload(workload, args.stencil9_random)
    z = int(rs[-1])
    x_host = pw.gen_uniform(n, 0)
    bytes_mv = algorithmic_bytes(n, z)
    log(f"# workload: {desc}; nnz={z}; generated in {gen_s:.1f}s; {bytes_mv / 1e6:.1f} MB per SpMV")
    # a matrix that fits the Infinity Cache is also timed rotating over
    # enough copies (at least 4) that they overflow it twice over
    copies = max(4, math.ceil(2 * MALL_BYTES / bytes_mv)) if bytes_mv < MALL_BYTES else 0

    x = pw.DeviceVector.from_array(x_host)
    y = pw.DeviceVector(n)
    results = {}
    wanted = set(args.formats.split(","))
    csb_opts = {"block_bits": 14} if workload == "stencil9" else {}
    csb_opts.update({k: v for k, v in (("csb_groups", args.csb_groups), ("csb_unroll", args.csb_unroll),
                                       ("csb_layout", args.csb_layout), ("csb_window_bits", args.csb_window_bits),
                                       ("csb_sync_every", args.csb_sync_every),
                                       ("csb_sweep_unroll", args.csb_sweep_unroll))
                     if v})
    if args.csb_nt >= 0:
        csb_opts["csb_nt"] = args.csb_nt
    crs_opts = {"crs_kernel": args.crs_kernel} if args.crs_kernel else {}
    if args.crs_nt >= 0:
        crs_opts["crs_nt"] = args.crs_nt
    formats = (("crs", pw.CRSHip, crs_opts),
               ("merge_path", pw.CRSMergeHip, {}), ("csb", pw.CSBHip, csb_opts), ("blockcoh", pw.BlockCOHHip, {}))
    if workload == "stencil9":
        # the workload names beta = 2^14; the reference's own rule gives 2^15
        formats += (("csb_ref_beta", pw.CSBHip, {k: v for k, v in csb_opts.items() if k != "block_bits"}),)
    for name, cls, opts in formats:
        if name not in wanted:
            continue
        t_load = time.time()
        m = cls(-1, "f64", **opts)
        m.loadCRS(n, n, rs, ci, dv)
        t_load = time.time() - t_load
        for _ in range(args.warmup):
            m.time_mv(x, y, 1)
        # per-step events: the SpMV's device time and its dominant kernel
        ms_ev, kms = m.time_mv(x, y, args.steps)
        # wall: the same SpMVs back to back, no markers between the steps
        pw.synchronize()
        t0 = time.perf_counter()
        b2b_ms, _ = m.time_mv(x, y, args.steps, per_step=False)
        pw.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        ms_flush, kms_flush = (float("nan"), float("nan")) if args.no_flush else m.time_mv(x, y, 10, flush_bytes=512 << 20)
        rot = None
        if copies:
            # the same SpMV rotating over `copies` independent matrices and
            # vectors: each finds its arrays in HBM, not the Infinity Cache
            mats, xs, ys = [m], [x], [y]
            for _ in range(copies - 1):
                mk = cls(-1, "f64", **opts)
                mk.loadCRS(n, n, rs, ci, dv)
                mats.append(mk)
                xs.append(pw.DeviceVector.from_array(x_host))
                ys.append(pw.DeviceVector(n))
            pw.time_mv_rotating(mats, xs, ys, 2)  # warm-up
            r_ev, r_k = pw.time_mv_rotating(mats, xs, ys, args.steps)
            pw.synchronize()
            t0 = time.perf_counter()
            r_b2b, _ = pw.time_mv_rotating(mats, xs, ys, args.steps, per_step=False)
            pw.synchronize()
            r_wall = (time.perf_counter() - t0) * 1e3 / (args.steps * copies)
            rot = {"copies": copies, "wall_ms": r_wall, "event_ms": r_ev, "kernel_ms": r_k, "b2b_ms": r_b2b}
            del mats, xs, ys
        # the drop-in host API: host x in, host y out (PCIe included; never `value`)
        yh = np.empty(n)
        host_ms = []
        for _ in range(6):
            th = time.perf_counter()
            m.mv(x_host, yh)
            host_ms.append((time.perf_counter() - th) * 1e3)
        host_ms = float(np.median(host_ms[1:]))
        inf = m.info()
        mstats = m.stats()
        yv = y.download()
        results[name] = {"ms_per_step": wall_ms, "event_ms": ms_ev, "kernel_ms": kms, "b2b_ms": b2b_ms,
                         "kernel_ms_flushed": kms_flush, "event_ms_flushed": ms_flush, "rot": rot,
                         "GBs": bytes_mv / (wall_ms * 1e-3) / 1e9, "info": inf, "load_s": t_load,
                         "y_checksum": float(np.sum(np.abs(yv))), "y": yv, "host_ms": host_ms}
        if rot:
            results[name]["GBs_hbm"] = bytes_mv / (rot["wall_ms"] * 1e-3) / 1e9
        log(f"# {name}: wall {wall_ms * 1e3:.1f} us/step (events around them {b2b_ms * 1e3:.1f} us), "
            f"{inf['kernels']} kernel(s) per SpMV, events per step {ms_ev * 1e3:.1f} us, dominant kernel "
            f"{kms * 1e3:.1f} us ({bytes_mv / (kms * 1e-3) / 1e9:.0f} GB/s), flushed {kms_flush * 1e3:.1f} us, "
            f"host-array mv {host_ms * 1e3:.0f} us" +
            (f"; rotating over {copies} copies: wall {rot['wall_ms'] * 1e3:.1f} us (events around them "
             f"{rot['b2b_ms'] * 1e3:.1f} us), kernel "
             f"{rot['kernel_ms'] * 1e3:.1f} us ({bytes_mv / (rot['kernel_ms'] * 1e-3) / 1e9:.0f} GB/s)" if rot else ""))
        del m

    # the format that gives `value`: the highest HBM-resident wall rate (for
    # a matrix that fits the Infinity Cache, the rotating rate)
    rate_key = "GBs_hbm" if copies else "GBs"
    best = max(results, key=lambda k: results[k][rate_key])
    r = results[best]

    # FP32 values and vectors, FP32 row sums: rate and error against FP64
    f32 = {}
    if "f32" in values.split(","):
        x32 = pw.DeviceVector.from_array(x_host, np.float32)
        y32 = pw.DeviceVector(n, np.float32)
        bytes32 = algorithmic_bytes(n, z, 4, 4)
        for name, cls, opts in formats:
            if name not in results:
                continue
            m = cls(-1, "f32", **opts)
            m.loadCRS(n, n, rs, ci, dv)
            for _ in range(args.warmup):
                m.time_mv(x32, y32, 1)
            ms_ev, kms = m.time_mv(x32, y32, args.steps)
            pw.synchronize()
            t0 = time.perf_counter()
            m.time_mv(x32, y32, args.steps, per_step=False)
            pw.synchronize()
            wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
            y64 = results[name]["y"]
            err = float(np.max(np.abs(y32.download().astype(np.float64) - y64)) / np.max(np.abs(y64)))
            f32[name] = {"GBs_wall": round(bytes32 / (wall_ms * 1e-3) / 1e9, 1), "us_per_step": round(wall_ms * 1e3, 2),
                         "dominant_kernel_us": round(kms * 1e3, 2),
                         "kernel_GBs": round(bytes32 / (kms * 1e-3) / 1e9, 1), "y_rel_err_vs_f64": err}
            log(f"# {name} f32: wall {wall_ms * 1e3:.1f} us/step, kernel {kms * 1e3:.1f} us "
                f"({bytes32 / (kms * 1e-3) / 1e9:.0f} GB/s of {bytes32 / 1e6:.1f} MB), y rel.err vs FP64 {err:.2e}")
            del m
        del x32, y32

    # gather-rate probe on this workload's x: a 12-byte stream per entry plus
    # one random x gather each, over all of x, over an eighth of x per XCD (a
    # CSB column group's window) and over 128 KiB per XCD (L2-resident)
    gather = {}
    if n >= 8 and z >= 8192:
        m_eff = z // 8192 * 8192
        for label, window, per_xcd in (("stream_only", 0, False), ("all_x", n, False), ("eighth_per_xcd", n // 8, True),
                                       ("128KiB_per_xcd", min(16384, n // 8), True)):
            ms = pw.gather_rate(x, window, per_xcd, z)
            rate = m_eff / (ms * 1e-3)
            gather[label] = {"window_doubles": int(window), "Ggathers_per_s": round(rate / 1e9, 1),
                             "us_for_nnz": round(z / rate * 1e6, 1)}
            if window == 0:
                gather[label]["stream_GBps"] = round(12 * rate / 1e9, 1)
                log(f"# gather probe {label}: 12 B/entry stream at {12 * rate / 1e9:.0f} GB/s, "
                    f"{z / rate * 1e6:.1f} us for nnz entries")
                continue
            log(f"# gather probe {label} ({window} doubles{' per XCD' if per_xcd else ''}): "
                f"{rate / 1e9:.1f} G gathers/s with the 12 B/entry stream, {z / rate * 1e6:.1f} us for nnz entries")

    # power method on the device with the best format: ms per iteration
    cls, opts = [(c, o) for nm, c, o in formats if nm == best][0]
    m = cls(-1, "f64", **opts)
    m.loadCRS(n, n, rs, ci, dv)
    xp = pw.DeviceVector.from_array(x_host)
    m.power_method_device(xp, y, 3, -1.0)
    xp.upload(x_host)
    pw.synchronize()
    t0 = time.perf_counter()
    lam, _, _ = m.power_method_device(xp, y, 100, -1.0)
    pw.synchronize()
    power_ms = (time.perf_counter() - t0) * 1e3 / 100
    log(f"# power method ({best}, FP64, device vectors): {power_ms * 1e3:.1f} us/iteration")
    del m, xp
    # the dominant kernel's time, or for a CSB SpMV with column groups the
    # slab kernel and the group reduction together (HIP events around each
    # SpMV, the dominant kernel's end marked between them); for a matrix
    # that fits the Infinity Cache, the rotating (HBM-resident) times
    # A single-kernel SpMV (r["info"]["kernels"] == 1) is timed by the HIP
    # events around its back-to-back launches: the kernel's average launch
    # duration, gaps included, without the per-launch markers, which add
    # about 1-2 us to a 20 us kernel.
    two_kernels = best in ("csb", "csb_ref_beta") and r["info"]["group_size"] > 1 and r["info"]["kernels"] == 2
    single = r["info"]["kernels"] == 1
    src = r["rot"] if copies else {"event_ms": r["event_ms"], "kernel_ms": r["kernel_ms"], "b2b_ms": r["b2b_ms"]}
    t_roof = src["b2b_ms"] if single else src["event_ms"] if two_kernels else src["kernel_ms"]
    achieved = bytes_mv / (t_roof * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{workload}_{best}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    csb_layouts = {1: "slab pieces", 2: "slab streams", 3: "wide slab streams", 4: "swept slabs"}
    what = ("csb slab kernel + column-group reduction" if two_kernels else
            f"{best} SpMV kernel (its only launch; HIP events around {args.steps * max(copies, 1)} back-to-back launches)"
            if single else f"{best} dominant kernel")
    if best.startswith("csb"):
        what += f" ({csb_layouts.get(r['info']['layout'], 'layout ' + str(r['info']['layout']))})"
    if copies:
        what += f", rotating over {copies} copies of A, x, y ({copies * bytes_mv / 1e6:.0f} MB > 256 MB Infinity Cache)"
    wall_ms = r["rot"]["wall_ms"] if copies else r["ms_per_step"]
    out = {
        "metric": METRIC,
        "value": round(bytes_mv / (wall_ms * 1e-3) / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "format": best, "nnz": z, "rows": n, "index": "int32",
                   "algorithmic_bytes_per_spmv": bytes_mv,
                   "pct_of_hbm_peak": round(100 * bytes_mv / (wall_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 1)},
        "gflops": round(2.0 * z / (wall_ms * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "timed": f"{what}, {round(t_roof * 1e3, 2)} us"},
    }
    if copies:
        out["warm"] = {"note": "the same SpMV repeated on one copy, served from the Infinity Cache",
                       "GBs_wall": round(r["GBs"], 1), "us_per_step": round(r["ms_per_step"] * 1e3, 2),
                       "dominant_kernel_us": round(r["kernel_ms"] * 1e3, 2),
                       "back_to_back_event_us": round(r["b2b_ms"] * 1e3, 2),
                       "frac_of_hbm_peak": round(bytes_mv / ((r["b2b_ms"] if single else r["kernel_ms"]) * 1e-3) / 1e9
                                                 / HBM_PEAK_GBS, 4)}
    fm = {}
    for k, v in results.items():
        e = {"GBs_wall": round(v["GBs"], 1), "us_per_step": round(v["ms_per_step"] * 1e3, 2),
             "device_us": round(v["event_ms"] * 1e3, 2), "dominant_kernel_us": round(v["kernel_ms"] * 1e3, 2),
             "kernels_per_spmv": v["info"]["kernels"], "back_to_back_event_us": round(v["b2b_ms"] * 1e3, 2),
             "dominant_kernel_us_flushed": (None if v["kernel_ms_flushed"] != v["kernel_ms_flushed"]
                                            else round(v["kernel_ms_flushed"] * 1e3, 2)),
             "partitions": v["info"]["threads"], "group_size": v["info"]["group_size"],
             "long_rows": v["info"]["long_rows"], "block_bits": v["info"]["block_bits"],
             "unroll": v["info"]["unroll"], "non_temporal": v["info"]["non_temporal"], "layout": v["info"]["layout"],
             "high_occupancy": v["info"]["high_occupancy"],
             "load_s": round(v["load_s"], 2), "sum_abs_y": v["y_checksum"],
             "host_array_mv_us": round(v["host_ms"] * 1e3, 1)}
        if k == "crs":
            e["crs_kernel"] = "tiles" if v["info"]["group_size"] == 0 else "vector"
        if v["rot"]:
            e["hbm_rotating"] = {"copies": v["rot"]["copies"], "GBs_wall": round(v["GBs_hbm"], 1),
                                 "us_per_step": round(v["rot"]["wall_ms"] * 1e3, 2),
                                 "device_us": round(v["rot"]["event_ms"] * 1e3, 2),
                                 "dominant_kernel_us": round(v["rot"]["kernel_ms"] * 1e3, 2),
                                 "back_to_back_event_us": round(v["rot"]["b2b_ms"] * 1e3, 2)}
        fm[k] = e
    out["formats"] = fm
    out["power_method_us_per_iteration"] = round(power_ms * 1e3, 2)
    out["power_method_lambda_after_100"] = lam
    if gather:
        out["gather_probe"] = gather
    out["matrix_stats"] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in mstats.items()}
    if f32:
        out["f32"] = f32
    if cpu and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, rs, ci, dv, x_host, bytes_mv, CPU_THREADS, y_gpu=r["y"])
        log(f"# cpu baseline: {out['cpu_baseline']['sample']}")
    return out


# ---------------------------------------------------------------- batch

def bench_batch(args):
    """Batched mixed-precision power method: 512 SPD instances of uneven size
    (256 square 2D Poisson grids, side uniform in [32, 1024]; 256 symmetric
    power-law matrices, N log-uniform in [1e3, 2e5], diagonal = row |sum| + 1),
    FP32 values (with --batch-bf16-exact 1, BF16 for every instance whose
    values are exact in BF16, i.e. the Poisson grids) and FP32 vectors, FP64
    norms / Rayleigh quotients, stop at relative
    eigenvalue change <= 1e-7 or 2000 iterations; all unfinished instances
    advance together, one tile launch + one finish launch per step, chunks
    of steps replayed as a hipGraph.  Compared with the plain FP64 per-matrix
    GPU loop (one CRS handle per instance, fused SpMV+norm steps) from the
    same start vectors."""
    import thesis_spmv_amd as pw
    rng = np.random.default_rng(4)
    n_each = args.batch_instances // 2
    sides = rng.integers(32, 1025, n_each)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(2e5), n_each)).astype(np.int64)
    t0 = time.time()
    spd = [pw.gen_spd_powerlaw(int(n), seed=1000 + i) for i, n in enumerate(ns)]
    mixed = args.batch_values == "f32" and args.batch_bf16_exact
    opts = {"bf16_when_exact": 1} if mixed else {}
    if args.batch_autotune:
        opts["autotune"] = 1
    b = pw.Batch(args.batch_values, items_per_thread=args.batch_ipt, chunk=args.batch_chunk, graph=args.batch_graph,
                 **opts)
    ids = [b.add_poisson(int(s), int(s)) for s in sides] + [b.add_crs(*m) for m in spd]
    n_inst, rows, nnz = b.size()
    log(f"# batch: {n_inst} instances, {rows} rows, {nnz} nnz, built in {time.time() - t0:.1f}s")
    b.solve(5, -1.0, seed=0)  # warm-up (uploads the matrices, chooses the tile schedule)
    pw.synchronize()
    sched = b.schedule()
    log(f"# batch tile schedule: 256 x {sched['items_per_thread']} items, "
        f"{'persistent' if sched['persistent'] else 'plain'} grid"
        + (" (autotuned: " + ", ".join(f"{k[0]}{'p' if k[1] else ''} {v:.3f} ms"
                                        for k, v in sched["candidate_ms"].items()) + " per step)"
           if sched["tuned"] else ""))
    t1 = time.perf_counter()
    lam, its, ms = b.solve(args.max_it, args.tol, seed=0)
    pw.synchronize()
    wall = time.perf_counter() - t1
    solves = n_inst / wall
    label = "f32, bf16 where exact" if mixed else args.batch_values
    log(f"# batched {label}: {wall * 1e3:.1f} ms wall ({ms:.1f} ms device), {solves:.1f} solves/s, "
        f"iterations min/median/max {its.min()}/{int(np.median(its))}/{its.max()}")
    out = {"metric": METRIC, "value": round(solves, 2), "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+bf16" if mixed else args.batch_values, "data": "synthetic",
           "config": {"workload": "batched power method, 512 SPD instances (Poisson sides U[32,1024], power-law "
                                  "SPD N log-U[1e3,2e5])", "instances": n_inst, "rows": rows, "nnz": nnz,
                      "max_it": args.max_it, "tol": args.tol,
                      "storage": "FP32 values, BF16 for the instances whose values are exact in BF16 (the Poisson "
                                 "grids); FP32 vectors and row sums; FP64 norms and Rayleigh quotients" if mixed else
                                 f"{args.batch_values} values; FP64 norms and Rayleigh quotients"},
           "iterations": {"min": int(its.min()), "median": float(np.median(its)), "max": int(its.max())},
           "engine": b.stats(),
           "schedule": {"items_per_thread": sched["items_per_thread"], "persistent": sched["persistent"],
                        "tuned": sched["tuned"],
                        "candidate_ms_per_step": {f"{k[0]}{'p' if k[1] else ''}": round(v, 4)
                                                  for k, v in sched["candidate_ms"].items()}}}
    # algorithmic bytes of every step actually taken (values, indices, row
    # offsets, x and y once per step)
    vb0 = {"f64": 8, "f32": 4, "bf16": 2}[args.batch_values]
    xb = 8 if args.batch_values == "f64" else 4

    def exact16(v):  # the library's test: float32-exact with a zero low half
        f = np.asarray(v, np.float64).astype(np.float32)
        return bool(np.all(f.astype(np.float64) == v) and np.all((f.view(np.uint32) & 0xFFFF) == 0))

    sizes = [(int(s) ** 2, 5 * int(s) ** 2 - 4 * int(s), 2 if mixed else vb0) for s in sides] + \
            [(len(m[0]) - 1, int(m[0][-1]), 2 if mixed and exact16(m[2]) else vb0) for m in spd]
    traffic = sum(int(k) * ((vb + 4) * z + 4 * (n + 1) + 2 * xb * n) for k, (n, z, vb) in zip(its, sizes))
    out["achieved_GBps"] = round(traffic / wall / 1e9, 1)
    log(f"# algorithmic traffic {traffic / 1e12:.2f} TB over all steps -> {traffic / wall / 1e9:.0f} GB/s")
    # one step with every instance active: 128 steps without the stop test
    # (after a first such solve that captures their graphs)
    full = sum((vb + 4) * z + 4 * (n + 1) + 2 * xb * n for n, z, vb in sizes)
    b.solve(128, -1.0, seed=0)
    pw.synchronize()
    t2 = time.perf_counter()
    b.solve(128, -1.0, seed=0)
    pw.synchronize()
    step_ms = (time.perf_counter() - t2) * 1e3 / 128
    out["full_step"] = {"ms": round(step_ms, 3), "GB": round(full / 1e9, 2), "GBps": round(full / step_ms / 1e6, 1)}
    out["roofline"] = {"bound": "hbm", "achieved": round(full / step_ms / 1e6, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(full / step_ms / 1e6 / HBM_PEAK_GBS, 4), "traffic": None,
                       "timed": "one engine step with all 512 instances active (tile + finish launches), "
                                f"{step_ms * 1e3:.1f} us"}
    log(f"# one step with all {n_inst} instances active: {step_ms:.3f} ms for {full / 1e9:.2f} GB "
        f"({full / step_ms / 1e6:.0f} GB/s)")
    if not args.no_baseline:
        # plain FP64 per-matrix GPU loop, same start vectors
        lam64 = np.zeros(n_inst)
        its64 = np.zeros(n_inst, np.int64)
        t_loop = 0.0

        def fp64_handle(i):
            m = pw.CRSHip(-1, "f64")
            if i < n_each:
                m.generatePoissonMatrix(int(sides[i]), int(sides[i]), 1)
            else:
                rs, ci, dv = spd[i - n_each]
                m.loadCRS(len(rs) - 1, len(rs) - 1, rs, ci, dv)
            return m

        for i in range(n_inst):
            m = fp64_handle(i)
            x = pw.DeviceVector.from_array(b.start_vector(ids[i], seed=0))
            y = pw.DeviceVector(x.n)
            pw.synchronize()
            ts = time.perf_counter()
            lam64[i], _, its64[i] = m.power_method_device(x, y, args.max_it, args.tol)
            t_loop += time.perf_counter() - ts
            del m, x, y
        rel_loop = np.abs(lam - lam64) / np.abs(lam64)
        # precision error proper: FP64 eigenvalue after the SAME number of
        # steps as the batched instance (reruns only where the stop differs)
        lam_same = lam64.copy()
        redo = np.nonzero(its64 != its)[0]
        for i in redo:
            m = fp64_handle(i)
            x = pw.DeviceVector.from_array(b.start_vector(ids[i], seed=0))
            y = pw.DeviceVector(x.n)
            lam_same[i], _, _ = m.power_method_device(x, y, int(its[i]), -1.0)
            del m, x, y
        rel = np.abs(lam - lam_same) / np.abs(lam_same)
        out["fp64_loop"] = {"solves_per_s": round(n_inst / t_loop, 2), "seconds": round(t_loop, 3)}
        out["speedup_vs_fp64_loop"] = round((n_inst / wall) / (n_inst / t_loop), 2)
        out["eigenvalue_rel_err_vs_fp64_same_steps"] = {"max": float(rel.max()), "median": float(np.median(rel))}
        out["eigenvalue_rel_err_vs_fp64_loop"] = {"max": float(rel_loop.max()), "median": float(np.median(rel_loop)),
                                                  "instances_stopping_at_other_step": int(len(redo))}
        log(f"# FP64 per-matrix loop: {t_loop:.2f} s ({n_inst / t_loop:.1f} solves/s); batched speedup "
            f"{out['speedup_vs_fp64_loop']}x; eigenvalue rel.err at equal steps max {rel.max():.2e}, "
            f"vs the loop's own stop max {rel_loop.max():.2e} ({len(redo)} stop elsewhere)")
    return out


# ---------------------------------------------------------------- serve

def _serve_requests(n_req, rate, n_pool, seed=51):
    """Poisson arrivals at `rate` req/s; each request is an SpMV chain
    (length U{1..64}, 75%) or a power solve (tol 1e-7, <= 1000 steps, 25%);
    20% reuse a matrix that is resident, the rest bring a fresh copy of a
    pool matrix (upload and tile plan included in their latency).  The job
    mix and the unit-rate arrival gaps come from fixed seeds, so every rate
    serves the same requests and only their arrival times scale with 1/rate."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0, n_req)) / rate
    reqs = []
    for i in range(n_req):
        chain = rng.random() < 0.75
        reqs.append({"t": float(t[i]), "mat": int(rng.integers(n_pool)), "repeat": bool(rng.random() < 0.2),
                     "steps": int(rng.integers(1, 65)) if chain else 1000, "tol": -1.0 if chain else 1e-7,
                     "seed": 1000 + i})
    return reqs


def _percentiles(lat):
    lat = np.asarray(lat) * 1e3
    return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3)}


def bench_serve(args):
    """Served stream: requests arrive as a Poisson process and are served by
    the continuous-batching server (FP32 storage, FP64 reductions, 64 GB
    admission budget) and, for comparison, by the plain FP64 per-request GPU
    path (one CRS handle per matrix, requests one after another).  The
    capacity the server sustains is its completion rate at the highest
    arrival rates; below it latency stays bounded.  Each served result of one
    rate's requests is also compared with the FP64 path at the same number
    of steps (results do not depend on what else is in flight, so they come
    from an untimed replay)."""
    import thesis_spmv_amd as pw
    rng = np.random.default_rng(5)
    n_pool = args.serve_pool
    ns = np.exp(rng.uniform(np.log(1e4), np.log(args.serve_max_n), n_pool)).astype(np.int64)
    kinds = ["poisson", "powerlaw", "stencil9"]
    pool = []
    t0 = time.time()
    for i, n in enumerate(ns):
        kind = kinds[i % 3]
        if kind == "poisson":
            side = max(2, int(np.sqrt(n)))
            pool.append({"kind": kind, "side": side, "n": side * side, "nnz": 5 * side * side - 4 * side})
        elif kind == "powerlaw":
            rs, ci, dv = pw.gen_powerlaw(int(n), alpha=1.8, xm=7.3333, cap=50000, struct_seed=100 + i,
                                         value_seed=200 + i)
            pool.append({"kind": kind, "n": int(n), "nnz": len(ci), "crs": (rs, ci, dv)})
        else:
            side = max(3, int(np.sqrt(n)))
            rs, ci, dv = pw.gen_stencil9(side, 2, 300 + i)
            pool.append({"kind": kind, "n": side * side, "nnz": len(ci), "crs": (rs, ci, dv)})
    log(f"# serve pool: {n_pool} matrices, N {min(p['n'] for p in pool)}..{max(p['n'] for p in pool)}, "
        f"built in {time.time() - t0:.1f}s")

    def register(srv, p):
        return srv.add_poisson(p["side"], p["side"]) if p["kind"] == "poisson" else \
            srv.add_matrix(*p["crs"], borrow=True)

    def run_server(reqs):
        srv = pw.Server("f32", budget_bytes=64 << 30, max_jobs=1024, chunk=args.serve_chunk, graph=args.serve_graph,
                        bf16_when_exact=args.serve_bf16_exact, items_per_thread=args.serve_ipt)
        resident_ids = {}  # pool index -> list of server matrix ids
        lat, jobs, pending, k = [], {}, list(reqs), 0
        served_bytes = 0  # algorithmic bytes of every step served (FP32 values and vectors)
        reg_s = 0.0  # host time registering fresh matrix copies (validation included)
        start = time.perf_counter()
        while k < len(pending) or jobs:
            now = time.perf_counter() - start
            while k < len(pending) and pending[k]["t"] <= now:
                r = pending[k]
                p = pool[r["mat"]]
                ids = [m for m in resident_ids.get(r["mat"], []) if srv.resident(m)]
                if r["repeat"] and ids:
                    mid = ids[0]
                else:
                    t_reg = time.perf_counter()
                    mid = register(srv, p)
                    reg_s += time.perf_counter() - t_reg
                    resident_ids.setdefault(r["mat"], []).append(mid)
                jobs[srv.submit(mid, r["steps"], r["tol"], seed=r["seed"])] = r
                k += 1
            if not jobs:
                time.sleep(max(0.0, pending[k]["t"] - (time.perf_counter() - start)))
                continue
            srv.step()
            done_at = time.perf_counter() - start
            for j in srv.finished():
                r = jobs.pop(j)
                its = srv.result(j)[2]
                p = pool[r["mat"]]
                served_bytes += its * algorithmic_bytes(p["n"], p["nnz"], 4, 4)
                lat.append(done_at - r["t"])
        makespan = time.perf_counter() - start
        tm = srv.timing()
        log(f"#   server phases: registration {reg_s:.2f}s, admit {tm['admit_s']:.2f}s ({tm['uploads']} uploads), "
            f"run {tm['run_s']:.2f}s, retire {tm['retire_s']:.2f}s of {makespan:.2f}s; "
            f"{served_bytes / 1e9:.1f} GB of steps served")
        del srv
        return lat, makespan, served_bytes

    def fp64_handle(p):
        m = pw.CRSHip(-1, "f64")
        if p["kind"] == "poisson":
            m.generatePoissonMatrix(p["side"], p["side"], 1)
        else:
            rs, ci, dv = p["crs"]
            m.loadCRS(len(rs) - 1, len(rs) - 1, rs, ci, dv)
        return m

    def run_fp64(reqs):
        cache, lat = {}, []
        start = time.perf_counter()
        for r in reqs:
            now = time.perf_counter() - start
            if now < r["t"]:
                time.sleep(r["t"] - now)
            p = pool[r["mat"]]
            m = cache.get(r["mat"]) if r["repeat"] else None
            if m is None:
                m = fp64_handle(p)
                cache[r["mat"]] = m
            n = p["n"]
            x = pw.DeviceVector.from_array(pw.server_start_vector(n, r["seed"]))
            y = pw.DeviceVector(n)
            m.power_method_device(x, y, r["steps"], r["tol"])
            lat.append(time.perf_counter() - start - r["t"])
            del x, y
        return lat, time.perf_counter() - start

    def result_errors(reqs):
        """Every request's served result (untimed replay, all submitted at
        once) against the FP64 path at the same number of steps: eigenvalue
        error over max(|lambda_64|, |A x|_64) (the scale of the quotient's
        rounding errors; lambda itself can sit near 0 on the nonsymmetric
        power-law matrices), relative where lambda_64 is the dominant scale,
        and the max-norm difference of the normalised final vectors."""
        srv = pw.Server("f32", budget_bytes=64 << 30, max_jobs=1024, chunk=args.serve_chunk, graph=args.serve_graph,
                        bf16_when_exact=args.serve_bf16_exact, items_per_thread=args.serve_ipt)
        mids = {}
        jobs = {}
        for r in reqs:
            if r["mat"] not in mids:
                mids[r["mat"]] = register(srv, pool[r["mat"]])
            jobs[srv.submit(mids[r["mat"]], r["steps"], r["tol"], seed=r["seed"])] = r
        res = {}
        while len(res) < len(jobs):
            srv.step()
            for j in srv.finished():
                res[j] = srv.result(j, pool[jobs[j]["mat"]]["n"])
        del srv
        lam_err, vec_err, rel_err = [], [], []
        cache = {}
        for j, r in jobs.items():
            lam, yn, its, vec = res[j]
            p = pool[r["mat"]]
            m = cache.get(r["mat"]) or cache.setdefault(r["mat"], fp64_handle(p))
            x = pw.DeviceVector.from_array(pw.server_start_vector(p["n"], r["seed"]))
            y = pw.DeviceVector(p["n"])
            lam64, yn64, its64 = m.power_method_device(x, y, its, -1.0)
            v64 = (x if its64 % 2 == 0 else y).download()
            lam_err.append(abs(lam - lam64) / max(abs(lam64), yn64))
            rel_err.append(abs(lam - lam64) / abs(lam64))
            vec_err.append(float(np.max(np.abs(vec - v64))))
            del x, y
        lam_err, vec_err, rel_err = map(np.asarray, (lam_err, vec_err, rel_err))
        return {"results": len(lam_err),
                "lambda_err_scaled": {"max": float(lam_err.max()), "median": float(np.median(lam_err))},
                "lambda_rel_err": {"max": float(rel_err.max()), "median": float(np.median(rel_err)),
                                   "p90": float(np.percentile(rel_err, 90))},
                "vector_maxnorm_err": {"max": float(vec_err.max()), "median": float(np.median(vec_err))}}

    rows = []
    for rate in args.serve_rates:
        reqs = _serve_requests(args.serve_requests, rate, n_pool)
        lat, span, served = run_server(reqs)
        row = {"rate": rate, "requests": len(reqs), "server": {"solves_per_s": round(len(lat) / span, 2),
                                                               **_percentiles(lat),
                                                               "served_GB": round(served / 1e9, 2),
                                                               "effective_GBps": round(served / span / 1e9, 1)}}
        log(f"# rate {rate}/s: server {row['server']}")
        if not args.no_baseline:
            lat64, span64 = run_fp64(reqs)
            row["fp64_per_request"] = {"solves_per_s": round(len(lat64) / span64, 2), **_percentiles(lat64)}
            log(f"#   fp64 per-request: {row['fp64_per_request']}")
        rows.append(row)
    capacity = max(r["server"]["solves_per_s"] for r in rows)
    for row in rows:
        row["above_capacity"] = bool(row["rate"] > capacity)
    top = rows[-1]
    out = {"metric": METRIC, "value": top["server"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
           "steps": len(rows), "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"served stream: Poisson arrivals, 75% SpMV chains (1-64) / 25% power solves "
                                  f"(tol 1e-7, <=1000 steps), 20% resident repeats, 64 GB budget, pool of {n_pool} "
                                  f"matrices N log-U[1e4,{args.serve_max_n:.0e}]",
                      "rates": args.serve_rates, "requests_per_rate": args.serve_requests},
           "p50_ms": top["server"]["p50_ms"], "p99_ms": top["server"]["p99_ms"],
           "capacity_solves_per_s": capacity,
           "capacity_note": "the server's completion rate at saturation (the highest over the swept rates, "
                            "which all serve the same requests); rates below it keep p99 bounded by the longest "
                            "job",
           "roofline": {"bound": "hbm", "achieved": top["server"]["effective_GBps"], "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(top["server"]["effective_GBps"] / HBM_PEAK_GBS, 4),
                        "traffic": None,
                        "timed": "the top rate's whole served stream: algorithmic bytes of every step served "
                                 "(FP32 values and vectors, 8Z + 4(N+1) + 8N per step) over the makespan, uploads "
                                 "and host work included"},
           "rates": rows}
    if not args.no_baseline:
        out["speedup_vs_fp64_at_top_rate"] = round(top["server"]["solves_per_s"] /
                                                   top["fp64_per_request"]["solves_per_s"], 2)
        err_rate = args.serve_rates[0]
        reqs = _serve_requests(args.serve_requests, err_rate, n_pool)
        out["error_vs_fp64"] = {"rate": err_rate, **result_errors(reqs)}
        log(f"# served results vs FP64 at equal steps (rate {err_rate}/s requests): {out['error_vs_fp64']}")
    return out


# ---------------------------------------------------------------- gather probe

GATHER_KINDS = ((0, "plain global load"), (1, "buffer load"), (2, "buffer load sc0"), (3, "buffer load nt"),
                (4, "buffer load sc0 nt"), (17, "buffer load sc1"), (18, "buffer load sc0 sc1"),
                (19, "buffer load sc1 nt"), (20, "buffer load sc0 sc1 nt"))


def bench_gather(args):
    """The gather-rate probe (pwm_gather_rate) on the power-law workload's x
    (N = 4 194 304 doubles) with nnz = 67 081 149 entries, for every way of
    loading the gathered x values: plain global loads and buffer loads with
    each combination of the sc0 / sc1 / nt cache-policy bits; windows of all
    of x (a row-ordered CRS sweep) and of an eighth of x per XCD."""
    import ctypes as C
    import thesis_spmv_amd as pw
    n, z = 4194304, 67081149
    x = pw.DeviceVector.from_array(pw.gen_uniform(n, 0))
    m_eff = z // 8192 * 8192
    rows = []
    for kind, name in GATHER_KINDS:
        for per_xcd, window in ((0, n), (1, n // 8)):
            ms = C.c_double()
            pw._check(pw.lib().pwm_gather_rate(x.ptr, n, window, per_xcd | (kind << 8), z, C.byref(ms)))
            rate = m_eff / (ms.value * 1e-3)
            rows.append({"kind": kind, "loads": name, "window": "eighth of x per XCD" if per_xcd else "all of x",
                         "us_for_nnz": round(z / rate * 1e6, 1), "Ggathers_per_s": round(rate / 1e9, 1)})
            log(f"# {name:24s} {rows[-1]['window']:20s}: {rows[-1]['us_for_nnz']:8.1f} us for nnz, "
                f"{rows[-1]['Ggathers_per_s']:6.1f} G gathers/s")
    return {"metric": METRIC, "value": max(r["Ggathers_per_s"] for r in rows), "unit": "G gathers/s", "n_gpus": 1,
            "steps": 10, "warmup": 2, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "gather-rate probe on the power-law x, every gather load kind", "entries": z},
            "kinds": rows}


# ---------------------------------------------------------------- COO -> CRS

def bench_coo(args):
    """COO -> CRS (TripletToCRS, Util/TripletToCRS.hpp:135-172) of the
    power-law matrix (N=4194304, ~67M entries) given as shuffled triplets:
    loadFromTriplets of a CRS_Merge handle (host validation, H2D upload, the
    device radix sort, gathers, row starts and the merge plan), against the
    oracle's restatement of TripletToCRS on one host core.  The device arrays
    are checked bit-exact against the oracle's."""
    import thesis_spmv_amd as pw
    n, rs, ci, dv, desc, gen_s = make_workload("powerlaw")
    z = int(rs[-1])
    t0 = time.time()
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(rs))
    perm = np.random.default_rng(5).permutation(z)
    t = pw.Triplet(n, n, rows[perm], ci[perm], dv[perm])
    del rows, perm
    log(f"# coo: {desc}; {z} shuffled triplets built in {time.time() - t0:.1f}s")
    small = pw.Triplet(1000, 1000, np.arange(1000), np.arange(1000), np.ones(1000))
    pw.CRSMergeHip().loadFromTriplets(small)  # warm-up: loads the kernels
    times = []
    for _ in range(max(1, args.steps)):
        m = pw.CRSMergeHip()
        pw.synchronize()
        t1 = time.perf_counter()
        m.loadFromTriplets(t)
        pw.synchronize()
        times.append(time.perf_counter() - t1)
        last = m
    gpu_s = float(np.median(times))
    grs, gci, gdv = last.crs_export()
    del last, m
    # CPU baseline leg (the oracle's restatement) and the parity check
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ot = oracle.Triplets(n, n, t.row_coord, t.col_coord, t.data)
    t1 = time.perf_counter()
    ref = oracle.triplet_to_crs(ot)
    cpu_s = time.perf_counter() - t1
    exact = bool(np.array_equal(grs, ref.row_start) and np.array_equal(gci, ref.col_ind) and np.array_equal(gdv, ref.data))
    log(f"# coo: loadFromTriplets {gpu_s * 1e3:.1f} ms (median of {len(times)}), {z / gpu_s / 1e6:.0f} M entries/s; "
        f"oracle TripletToCRS {cpu_s * 1e3:.0f} ms; arrays bit-exact: {exact}")
    out = {"metric": METRIC, "value": round(z / gpu_s / 1e6, 1), "unit": "M entries/s", "n_gpus": 1,
           "steps": len(times), "warmup": 1, "ms_per_step": round(gpu_s * 1e3, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32 keys, f64 values", "data": "synthetic",
           "config": {"workload": "COO -> CRS of the power-law matrix (N=4194304), shuffled triplets, loadFromTriplets",
                      "entries": z, "bit_exact_vs_oracle": exact},
           "cpu_baseline": {"value": round(z / cpu_s / 1e6, 1), "unit": "M entries/s", "cores": 1, "kind": "port",
                            "sample": f"the oracle's TripletToCRS restatement on all {z} triplets, one call, "
                                      f"{cpu_s * 1e3:.0f} ms"}}
    if not exact:
        raise RuntimeError("COO -> CRS arrays differ from the oracle's")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="all",
                    help="all (power-law line + configs 0, 2, 3, 4) or one of powerlaw, poisson, stencil9, batch, "
                         "serve, coo, gather")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-flush", action="store_true", help="skip the Infinity-Cache-flushed timing")
    ap.add_argument("--formats", default="crs,merge_path,csb,blockcoh,csb_ref_beta")
    ap.add_argument("--crs-kernel", type=int, default=0,
                    help="CRS kernel: 0 = time the candidates at load, 1 = vector, 2 = row-aligned tiles")
    ap.add_argument("--csb-groups", type=int, default=0, help="CSB block-column groups (0 = time 1, 8, 16 at load)")
    ap.add_argument("--csb-unroll", type=int, default=0, help="CSB stream kernel unroll (0 = time 2 and 4 at load)")
    ap.add_argument("--csb-layout", type=int, default=0,
                    help="CSB layout: 0 = automatic, 1 slab pieces, 2 slab streams, 3 wide streams, 4 swept slabs")
    ap.add_argument("--csb-window-bits", type=int, default=0, help="swept slabs: log2 columns per window (0 = 1 MB)")
    ap.add_argument("--csb-sync-every", type=int, default=0, help="swept slabs: windows between workgroup barriers")
    ap.add_argument("--csb-sweep-unroll", type=int, default=0,
                    help="swept slabs: groups of 4 entries per lane per chunk (1, 2, 4; 0 = the library's 2)")
    ap.add_argument("--stencil9-random", type=int, default=2,
                    help="random off-band nonzeros per row of the stencil9 workload (2 in the configuration; 0 gives "
                         "the plain 9-point stencil)")
    ap.add_argument("--csb-nt", type=int, default=-1,
                    help="CSB stream kernel non-temporal entry loads: 1, 0, or -1 = time both at load")
    ap.add_argument("--crs-nt", type=int, default=-1,
                    help="CRS tiles kernel non-temporal entry loads: 1, 0, or -1 = time both at load")
    ap.add_argument("--values", default=None,
                    help="f64 and/or f32 SpMV passes (default f64; f64,f32 for stencil9)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="bytes: run the library's streaming copy kernel once first (profiler byte-counter reference)")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--batch-values", default="f32")
    ap.add_argument("--batch-ipt", type=int, default=16, help="merge items per thread of a batch tile (4, 8, 16)")
    ap.add_argument("--batch-bf16-exact", type=int, default=1,
                    help="FP32 batch: store instances whose values are exact in BF16 as BF16 (1) or not (0)")
    ap.add_argument("--batch-chunk", type=int, default=32, help="batch steps per graph replay")
    ap.add_argument("--batch-graph", type=int, default=1, help="0: the batch launches its steps one by one "
                    "instead of replaying them as hipGraphs")
    ap.add_argument("--batch-autotune", type=int, default=1,
                    help="1: the warm-up solve times tile sizes 256x8 / 256x16 with plain and persistent grids and "
                         "keeps the fastest (overrides --batch-ipt)")
    ap.add_argument("--max-it", type=int, default=2000)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[10.0, 20.0, 50.0, 200.0, 1000.0, 2000.0])
    ap.add_argument("--serve-requests", type=int, default=60, help="requests per arrival rate")
    ap.add_argument("--serve-pool", type=int, default=12, help="distinct matrices the requests draw from")
    ap.add_argument("--serve-max-n", type=float, default=1e7)
    ap.add_argument("--serve-chunk", type=int, default=8, help="server steps per chunk")
    ap.add_argument("--serve-ipt", type=int, default=16, choices=(4, 8, 16),
                    help="server tile size: 256 x this many merge items")
    ap.add_argument("--serve-bf16-exact", type=int, default=0,
                    help="1: the server stores matrices whose values are exact in BF16 (Poisson) as BF16")
    ap.add_argument("--serve-graph", type=int, default=1, help="0: the server launches its steps one by one "
                    "instead of replaying them as hipGraphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "sharding the rows over several GPUs is designed, not built", file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's OpenMP threads stay on neighbouring cores; libgomp
    # reads this when the oracle library is first loaded
    os.environ.setdefault("OMP_PROC_BIND", "close")

    import thesis_spmv_amd as pw
    if args.calibrate_copy > 0:
        nb = args.calibrate_copy // 16 * 16
        src, dst = pw.DeviceVector(nb // 8), pw.DeviceVector(nb // 8)
        src.fill(1.0)
        pw._check(pw.lib().pwm_device_copy(dst.ptr, src.ptr, nb))
        log(f"# calibration copy: {nb} bytes read and written once")
        del src, dst
    single = {"batch": bench_batch, "serve": bench_serve, "coo": bench_coo, "gather": bench_gather}
    if args.workload in single:
        out = single[args.workload](args)
    elif args.workload in ("powerlaw", "poisson", "stencil9"):
        out = bench_spmv(args, args.workload)
    elif args.workload == "all":
        t_all = time.time()
        out = bench_spmv(args, "powerlaw")
        extra = {}
        for key, fn in (("config0_poisson", lambda: bench_spmv(args, "poisson")),
                        ("config2_stencil9", lambda: bench_spmv(args, "stencil9")),
                        ("config3_batch", lambda: bench_batch(args)),
                        ("config4_serve", lambda: bench_serve(args))):
            t0 = time.time()
            log(f"## {key}")
            try:
                res = fn()
                for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data"):
                    res.pop(k, None)
                extra[key] = res
            except Exception as e:  # one failing extra config must not hide the others
                traceback.print_exc()
                extra[key] = {"error": repr(e)}
            extra[key]["bench_seconds"] = round(time.time() - t0, 1)
        out["configs"] = extra
        out["bench_seconds"] = round(time.time() - t_all, 1)
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
