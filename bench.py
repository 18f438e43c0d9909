"""Benchmark: sparse-grid stencil throughput on one MI355X.

Workload (one step = one pass of the hot path over the grid): 512^3 domain, spherical shell
200 <= r < 232 (18,795,780 points in 48,654 chunks of 8^3), two FP32 properties u, v
(u = 1, v = 0, 50 cubes of edge 8 with u = 0.5, v = 0.25, seed 1), one Gray-Scott
reaction-diffusion step (7-point Laplacian, Du = 2e-5, Dv = 1e-5, F = 0.04, k = 0.06,
dt = 1) ping-ponging (u, v) <-> (u', v').  Inputs are resident in HBM before timing.

Prints one JSON line: value = point-updates/s (10^9), roofline of the stencil kernel
(algorithmic bytes / kernel time, HIP events on the library's stream), and the CPU oracle
timed on this host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "sparse-grid stencil Gpt-updates/s + % HBM roofline; p50/p99 solve latency"


def log(msg):
    """Progress line on stderr (stdout carries only the JSON line)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def chunk_arrays(size, coords, values, edge=8):
    """Point coordinates -> sorted chunk ids, [n, 512] masks and per-property chunk data."""
    bs = [-(-s // edge) for s in size]
    c = coords.astype(np.int64)
    chunk = ((c[:, 2] // edge) * bs[1] + c[:, 1] // edge) * bs[0] + c[:, 0] // edge
    off = ((c[:, 2] % edge) * edge + c[:, 1] % edge) * edge + c[:, 0] % edge
    ids, inv = np.unique(chunk, return_inverse=True)
    cs = edge ** 3
    masks = np.zeros((ids.size, cs), np.uint8)
    masks[inv, off] = 1
    data = []
    for v in values:
        a = np.zeros((ids.size, cs), v.dtype)
        a[inv, off] = v
        data.append(a)
    return ids, masks, data


def traffic_from_profile(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, copy
    calibrated), committed as a one-line summary under profiles/; None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_gpu(args, ids, masks, data, dtype, edge=8):
    import torch
    from openfpm_data_amd import SparseGridGpu, autotune
    from openfpm_data_amd import workloads as W

    torch.cuda.set_device(0)
    g = SparseGridGpu(3, (512, 512, 512), [dtype] * 4, block_edge=edge)
    stream = torch.cuda.Stream()
    g.set_stream(stream)
    d_ids = torch.from_numpy(ids).cuda()
    d_m = torch.from_numpy(masks).cuda()
    d_u = torch.from_numpy(data[0].astype(dtype)).cuda()
    d_v = torch.from_numpy(data[1].astype(dtype)).cuda()
    z = torch.zeros_like(d_u)
    torch.cuda.synchronize()
    g.insert_block(d_ids, d_m, [d_u, d_v, z, z])
    g.flush(["sright"] * 4)
    g.find_neighbours()
    g.synchronize()
    del d_ids, d_m, d_u, d_v, z
    P = W.GS_PARAMS
    npts = g.count_existing_elements()
    nch = g.num_chunks()

    def step(i):
        if i % 2 == 0:
            g.conv2_gray_scott(0, 1, 2, 3, **P)
        else:
            g.conv2_gray_scott(2, 3, 0, 1, **P)

    tuned = None
    if edge == 8:
        # launch shape by measurement (cached per device and size class)
        tuned, times = autotune.tune_stencil(g, step, "gray_scott", dtype, stream, force=args.retune)
        for cfg, ms in sorted(times.items(), key=lambda kv: kv[1]):
            log("  candidate %s: %.4f ms" % (cfg, ms))
        g.synchronize()
        # restart from the initial state: tuning stepped the grid
        g.set_prop(0, data[0].astype(dtype))
        g.set_prop(1, data[1].astype(dtype))
        g.synchronize()
    for i in range(args.warmup):
        step(i)
    g.synchronize()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step(args.warmup + i)
    ev1.record(stream)
    ev1.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    res = {"points": npts, "chunks": nch, "wall_s": wall, "kernel_ms": kern_ms, "tuned": tuned,
           "checksum_u": float(np.float64(g.prop(0 if args.steps % 2 == 0 else 2)[masks == 1].astype(np.float64).sum()))}
    g.close()
    return res


def run_flush_secondary(cpu_threads, with_cpu):
    """2^24 point inserts into 1024^3 (64 Gaussian blobs, 30% repeats), int32 smax + uint32
    sadd flush, NNFull tagBoundaries, 27-point uint32 box sum; inserts resident in HBM."""
    import torch
    from openfpm_data_amd import SparseGridGpu, _capi as C
    from openfpm_data_amd import workloads as W

    w = W.blobs1024()
    n = w["coords"].shape[0]
    d_c = torch.from_numpy(w["coords"]).cuda()
    d_i = torch.from_numpy(w["int32"]).cuda()
    d_u = torch.from_numpy(w["uint32"].view(np.int32)).cuda()
    z = torch.zeros_like(d_u)
    stream = torch.cuda.Stream()
    times = []
    for rep in range(4):
        g = SparseGridGpu(3, w["size"], [np.int32, np.uint32, np.uint32], block_edge=8)
        g.set_stream(stream)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev[0].record(stream)
        g.insert_coords(d_c, [d_i, d_u, z])
        g.flush(["smax", "sadd", "none"])
        ev[1].record(stream)
        g.find_neighbours(C.NN_FULL)
        g.set_nn_type(C.NN_FULL)
        ev[2].record(stream)
        g.tag_boundaries(C.NN_FULL)
        ev[3].record(stream)
        g.conv_box(1, 2, np.ones(27, np.int64))
        ev[4].record(stream)
        ev[4].synchronize()
        times.append([ev[i].elapsed_time(ev[i + 1]) for i in range(4)])
        nch, npts = g.num_chunks(), g.count_existing_elements()
        g.close()
    t = np.median(np.array(times[1:]), axis=0)  # first repetition warms the allocator
    # algorithmic bytes: insert records in (8-byte key + 12 value bytes + mask byte), every
    # slot of the resulting chunks written once (3 x 4-byte values + mask); tagBoundaries
    # reads and writes each mask byte; the box reads one uint32 and the mask and writes one
    # uint32 per slot; both read 26 neighbour-table entries per chunk
    slots = nch * 512
    rl = lambda b, ms: {"bound": "hbm", "achieved": round(b / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,  # noqa: E731
                        "unit": "GB/s", "frac": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    out = {"workload": "blobs1024_flush", "inserts": n, "chunks": nch, "points": npts,
           "insert_flush_ms": round(float(t[0]), 3), "inserts_per_s": round(n / (t[0] * 1e-3) / 1e9, 4),
           "unit": "G inserts/s", "find_neighbours_full_ms": round(float(t[1]), 3),
           "tag_boundaries_full_ms": round(float(t[2]), 3), "box27_uint32_ms": round(float(t[3]), 3),
           "roofline": {"insert_flush": rl(n * 21 + slots * 13, t[0]),
                        "tag_boundaries_full": rl(slots * 2 + nch * 26 * 4, t[2]),
                        "box27_uint32": rl(slots * 9 + nch * 26 * 4, t[3])}}
    if with_cpu:
        from oracle import pysgcpu
        pysgcpu.set_threads(cpu_threads)
        g = pysgcpu.SparseGridCpu(3, w["size"], [np.int32, np.uint32, np.uint32], block_edge=8)
        t0 = time.perf_counter()
        g.insert_coords(w["coords"], [w["int32"], w["uint32"], np.zeros(n, np.uint32)])
        g.flush(["smax", "sadd", "none"])
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / el / 1e9, 5), "unit": "G inserts/s", "cores": cpu_threads,
                               "kind": "port", "sample": "oracle insert + flush of the same 2^24 records, %.1f s" % el}
        g.close()
    return out


def run_poisson_secondary():
    """256 independent Poisson grids (edge log-uniform in [32, 256], 5-40% of the chunks
    active, f uniform(-1, 1), seed 3) solved to relative residual 1e-10: all grids in one
    continuous batch in the mixed form (FP32 red-black SOR + FP64 defect correction, inner
    sweeps replayed from a hipGraph) against one plain FP64 solve per grid."""
    import torch
    from openfpm_data_amd import poisson as P, _capi as C
    from openfpm_data_amd import workloads as W

    specs = W.poisson_batch(256)
    fs = [torch.from_numpy(s["f"]).cuda() for s in specs]

    def grids():
        gs = [P.poisson_grid(s["size"], s["chunk_ids"], f) for s, f in zip(specs, fs)]
        torch.cuda.synchronize()
        return gs

    def timed(mode, batched, graph, tuning=None):
        gs = grids()
        t0 = time.perf_counter()
        if batched:
            res = P.solve(gs, P.make_params(mode=mode, tol=1e-10, use_graph=graph), tuning=tuning)
        else:
            res = [P.solve([g], P.make_params(mode=mode, tol=1e-10, use_graph=graph))[0] for g in gs]
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        for g in gs:
            g.close()
        return el, res

    from openfpm_data_amd import autotune
    gs = grids()
    tuned, _ = autotune.tune_batch(gs)
    for g in gs:
        g.close()
    tuning = {C.TUNE_BATCH_WAVES: tuned["waves"], C.TUNE_BATCH_MAX_WGS: tuned["max_wgs"]}
    timed(C.POISSON_MIXED, True, True, tuning)  # warm-up (allocator, code objects)
    t_mix, r_mix = timed(C.POISSON_MIXED, True, True, tuning)
    t_plain, r_plain = timed(C.POISSON_PLAIN_FP64, False, False)
    # the plain FP64 method batched the same way (shared launches, graph-replayed sweeps):
    # separates the precision gain from the batching gain
    t_pb, r_pb = timed(C.POISSON_PLAIN_FP64, True, True, tuning)
    sw = np.array([r["sweeps"] for r in r_mix])
    # sweep kernel rate: one outer iteration of K FP32 red-black sweeps over all 256 grids
    # (no graph: 2 K launches of rbsor_batch_kernel and two defect passes).  Algorithmic bytes
    # per slot: 13 per half sweep (x read and written, f read, mask), 33 per defect pass.
    K = 40
    slots = int(sum(s["chunk_ids"].size for s in specs)) * 512
    gs = grids()
    P.solve(gs, P.make_params(tol=1e-30, omega=1.5, first_sweeps=K, max_outer=1, use_graph=False), tuning=tuning)
    for g in gs:
        g.close()
    gs = grids()
    t0 = time.perf_counter()
    rk = P.solve(gs, P.make_params(tol=1e-30, omega=1.5, first_sweeps=K, max_outer=1, use_graph=False), tuning=tuning)
    torch.cuda.synchronize()
    t_k = time.perf_counter() - t0
    for g in gs:
        g.close()
    assert all(r["sweeps"] == K for r in rk)
    sweep_bytes = slots * (2 * K * 13 + 2 * 33)
    sweep_gbs = sweep_bytes / t_k / 1e9
    return {"workload": "poisson_batch256", "grids": len(specs),
            "chunks": int(sum(s["chunk_ids"].size for s in specs)), "tol": 1e-10,
            "solves_per_s": round(len(specs) / t_mix, 2), "unit": "solves/s",
            "batch_s": round(t_mix, 4), "converged": int(sum(r["converged"] for r in r_mix)),
            "max_rel_residual": max(r["rel_residual"] for r in r_mix),
            "sweeps_mean": round(float(sw.mean()), 1), "sweeps_max": int(sw.max()),
            "per_grid_plain_fp64_solves_per_s": round(len(specs) / t_plain, 2),
            "per_grid_plain_fp64_converged": int(sum(r["converged"] for r in r_plain)),
            "speedup_vs_per_grid_plain_fp64": round(t_plain / t_mix, 3),
            "batched_plain_fp64_solves_per_s": round(len(specs) / t_pb, 2),
            "batched_plain_fp64_converged": int(sum(r["converged"] for r in r_pb)),
            "speedup_vs_batched_plain_fp64": round(t_pb / t_mix, 3),
            "roofline": {"rbsor_batch_sweeps": {"bound": "hbm", "achieved": round(sweep_gbs, 1), "peak": HBM_PEAK_GBS,
                                                "unit": "GB/s", "frac": round(sweep_gbs / HBM_PEAK_GBS, 4),
                                                "sweeps": K, "seconds": round(t_k, 5)}},
            "tuned": tuned}


def run_serving_secondary(n_open=384, n_cap=64, budget_gib=32.0, batch_tuning=None):
    """Served Poisson requests (edge log-uniform in [32, 384], 5-40% of the chunks active,
    FP32 f, tolerance 1e-10 or 1e-6, seed 4), mixed method, HBM budget `budget_gib`:
    capacity measured closed-loop (first `n_cap` requests all present at t = 0), then
    `n_open` requests arriving as a Poisson process at 80% of that capacity (seed 5)."""
    from openfpm_data_amd import serving as S
    from openfpm_data_amd import workloads as W

    reqs = W.serving_requests(n_open)
    budget = int(budget_gib * (1 << 30))
    srv = S.PoissonServer(budget, tuning=batch_tuning)
    srv.run(reqs[:8], np.zeros(8))  # warm-up
    _, cap = srv.run(reqs[:n_cap], np.zeros(n_cap))
    rate = 0.8 * cap["solves_per_s"]
    _, s = srv.run(reqs, S.poisson_arrivals(n_open, rate, 5))
    srv.close()
    return {"workload": "poisson_serving", "requests": n_open, "budget_gib": budget_gib,
            "capacity_solves_per_s": round(cap["solves_per_s"], 2), "arrival_rate_per_s": round(rate, 2),
            "solves_per_s": round(s["solves_per_s"], 2), "unit": "solves/s",
            "p50_latency_ms": round(s["p50_latency_s"] * 1e3, 3), "p99_latency_ms": round(s["p99_latency_s"] * 1e3, 3),
            "point_updates_per_s": round(s["point_updates_per_s"] / 1e9, 3), "point_updates_unit": "G/s",
            "served": s["served"], "converged": s["converged"], "rejected": s["rejected"],
            "max_live": s["max_live"], "peak_reserved_gib": round(s["peak_reserved_bytes"] / (1 << 30), 3)}


def run_copy_bench(nbytes=1 << 30, reps=20):
    """Streaming device copies of `nbytes` (sg_stream_copy, 16-byte vectors, plain and
    non-temporal), HIP events on the copy's stream: GB/s counting the bytes read and written."""
    import ctypes
    import torch
    from openfpm_data_amd import _capi as C

    lib = C.load()
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    stream = torch.cuda.Stream()
    out = {"workload": "stream_copy", "bytes": nbytes, "reps": reps}
    for name, flags in (("plain", 0), ("nontemporal", C.COPY_NONTEMPORAL)):
        for _ in range(3):
            C.check(lib.sg_stream_copy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()), nbytes, flags,
                                       ctypes.c_void_p(stream.cuda_stream)))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            C.check(lib.sg_stream_copy(ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(a.data_ptr()), nbytes, flags,
                                       ctypes.c_void_p(stream.cuda_stream)))
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = {"ms": round(ms, 4), "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_of_peak": round(2 * nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    torch.cuda.synchronize()
    assert bool((b == 1).all()), "copy result"
    # a non-temporal copy reading what the previous kernel's non-temporal stores wrote
    # (distinct values, b and c cleared first): does the second copy see the new data?
    a.copy_(torch.arange(a.numel(), device="cuda", dtype=torch.float32))
    b.zero_()
    c = torch.zeros_like(a)
    torch.cuda.synchronize()
    for dst, src in ((b, a), (c, b)):
        C.check(lib.sg_stream_copy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                                   C.COPY_NONTEMPORAL, ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    out["nontemporal_roundtrip_equal"] = bool(torch.equal(c, a))
    return out


def host_cores():
    """CPU cores this process may run on (its affinity set) and the host's count."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return avail, os.cpu_count() or avail


def run_cpu_baseline(ids, masks, data, seconds, threads, edge=8):
    from oracle import pysgcpu
    from openfpm_data_amd import workloads as W

    pysgcpu.set_threads(threads)
    g = pysgcpu.SparseGridCpu(3, (512, 512, 512), [np.float32] * 4, block_edge=edge)
    z = np.zeros_like(data[0], dtype=np.float32)
    g.insert_block(ids, masks, [data[0].astype(np.float32), data[1].astype(np.float32), z, z])
    g.flush(["sright"] * 4)
    g.find_neighbours()
    P = W.GS_PARAMS
    npts = int(masks.sum())
    steps = 0
    t0 = time.perf_counter()
    while True:
        if steps % 2 == 0:
            g.conv2_gray_scott(0, 1, 2, 3, **P)
        else:
            g.conv2_gray_scott(2, 3, 0, 1, **P)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 1000:
            break
    g.close()
    return {"value": npts * steps / el / 1e9, "steps": steps, "seconds": el}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bounded CPU-baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core of the process's affinity set")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--retune", action="store_true", help="time every stencil launch shape instead of the cache")
    ap.add_argument("--no-secondary", action="store_true", help="skip the insertion/flush workload")
    ap.add_argument("--only", choices=["flush", "poisson", "serving"],
                    help="run one secondary workload alone and print its JSON (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 20 device copies of 1 GiB (reference for the FETCH/WRITE_SIZE counters)")
    ap.add_argument("--copy-bench", action="store_true",
                    help="only time the library's streaming copy of 1 GiB (plain and non-temporal): the "
                         "HBM ceiling the stencil is measured against")
    args = ap.parse_args()
    if args.copy_bench:
        print(json.dumps(run_copy_bench()))
        return
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "copies": 20}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the one-GPU "
              "figure is the kernel at its per-GPU shape); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.only:
        threads = args.cpu_threads or host_cores()[0]
        if args.only == "flush":
            print(json.dumps(run_flush_secondary(threads, not args.no_cpu)))
        elif args.only == "poisson":
            print(json.dumps(run_poisson_secondary()))
        else:
            print(json.dumps(run_serving_secondary()))
        return

    from openfpm_data_amd import workloads as W
    from openfpm_data_amd import autotune
    t_gen = time.perf_counter()
    w = W.shell512_gs()
    gen_s = time.perf_counter() - t_gen
    dtype = np.float32 if args.dtype == "f32" else np.float64
    P_GS = W.GS_PARAMS

    def gs_step(g, i):
        if i % 2 == 0:
            g.conv2_gray_scott(0, 1, 2, 3, **P_GS)
        else:
            g.conv2_gray_scott(2, 3, 0, 1, **P_GS)

    log("workload generated in %.1f s; tuning the chunk edge" % gen_s)
    # chunk edge by measurement (cached): one Gray-Scott step per point at edges 4 and 8
    edge, _ = autotune.tune_chunk_edge(w["size"], w["coords"], [w["u"], w["v"]], dtype, gs_step)
    ids, masks, data = chunk_arrays(w["size"], w["coords"], [w["u"].astype(np.float32), w["v"].astype(np.float32)],
                                    edge=edge)
    del w
    log("chunk edge %d; stencil run" % edge)
    r = run_gpu(args, ids, masks, data, dtype, edge=edge)
    log("stencil %.4f ms/step" % r["kernel_ms"])

    npts, nch = r["points"], r["chunks"]
    slots = nch * edge ** 3
    tsz = np.dtype(dtype).itemsize
    # algorithmic bytes of one step: per slot u, v read + u', v' written + mask byte;
    # per chunk the 6 face-neighbour entries of the neighbour table
    alg_bytes = slots * (2 * tsz * 2 + 1) + nch * 6 * 4
    kern_s = r["kernel_ms"] * 1e-3
    achieved = alg_bytes / kern_s / 1e9
    value = npts * args.steps / (r["kernel_ms"] * 1e-3 * args.steps) / 1e9
    traffic = traffic_from_profile("traffic_gray_scott_%s_shell512.txt" % args.dtype)

    cpu = None
    avail, nproc = host_cores()
    cpu_threads = args.cpu_threads or avail
    if not args.no_cpu:
        log("cpu baseline (%d threads, then 1)" % cpu_threads)
        c = run_cpu_baseline(ids, masks, data, args.cpu_seconds, cpu_threads, edge)
        c1 = run_cpu_baseline(ids, masks, data, args.cpu_seconds, 1, edge)
        cpu = {"value": round(c["value"], 6), "unit": "Gpt-updates/s", "cores": cpu_threads, "kind": "port",
               "nproc": nproc, "cores_available": avail,
               "sample": "oracle (oracle/sgcpu.cpp, OpenMP) on the same 512^3 shell FP32 Gray-Scott grid, "
                         "%d steps in %.1f s on %d threads" % (c["steps"], c["seconds"], cpu_threads),
               "one_core": {"value": round(c1["value"], 6), "cores": 1,
                            "sample": "%d steps in %.1f s on 1 thread" % (c1["steps"], c1["seconds"])}}

    secondary = None
    if not args.no_secondary:
        if dtype == np.float32:
            # the same workload computed in FP64 (its own kernel shape; 33 B per slot)
            log("FP64 stencil run")
            a64 = argparse.Namespace(**vars(args))
            a64.steps, a64.warmup = 200, 10
            r64 = run_gpu(a64, ids, masks, data, np.float64, edge=edge)
            b64 = slots * (2 * 8 * 2 + 1) + nch * 6 * 4
            ach64 = b64 / (r64["kernel_ms"] * 1e-3) / 1e9
            secondary_f64 = {"workload": "shell512_gray_scott_fp64", "steps": a64.steps,
                             "ms_per_step": round(r64["kernel_ms"], 5),
                             "value": round(r64["points"] / (r64["kernel_ms"] * 1e-3) / 1e9, 4), "unit": "Gpt-updates/s",
                             "tuned": r64["tuned"],
                             "roofline": {"bound": "hbm", "achieved": round(ach64, 2), "peak": HBM_PEAK_GBS,
                                          "unit": "GB/s", "frac": round(ach64 / HBM_PEAK_GBS, 4)}}
        log("flush workload")
        secondary = {"flush": run_flush_secondary(cpu_threads, not args.no_cpu),
                     "poisson": None}
        if dtype == np.float32:
            secondary["stencil_fp64"] = secondary_f64
        log("batched Poisson workload")
        secondary["poisson"] = run_poisson_secondary()
        tb = secondary["poisson"]["tuned"]
        from openfpm_data_amd import _capi as C
        log("served Poisson workload")
        secondary["serving"] = run_serving_secondary(
            batch_tuning={C.TUNE_BATCH_WAVES: tb["waves"], C.TUNE_BATCH_MAX_WGS: tb["max_wgs"]})

    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["kernel_ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {"workload": "shell512_gray_scott", "domain": [512, 512, 512], "shell_radius": [200, 232],
                   "points": npts, "chunks": nch, "chunk_edge": edge, "properties": 2, "tuned": r["tuned"],
                   "params": W.GS_PARAMS},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "secondary": secondary,
        "wall_s": round(r["wall_s"], 4),
        "setup_s": round(gen_s, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
