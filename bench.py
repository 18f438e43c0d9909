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


def run_gpu(args, ids, masks, data, dtype):
    import torch
    from openfpm_data_amd import SparseGridGpu
    from openfpm_data_amd import workloads as W

    torch.cuda.set_device(0)
    g = SparseGridGpu(3, (512, 512, 512), [dtype] * 4, block_edge=8)
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
    res = {"points": npts, "chunks": nch, "wall_s": wall, "kernel_ms": kern_ms,
           "checksum_u": float(np.float64(g.prop(0 if args.steps % 2 == 0 else 2)[masks == 1].astype(np.float64).sum()))}
    g.close()
    return res


def run_cpu_baseline(ids, masks, data, seconds, threads):
    from oracle import pysgcpu
    from openfpm_data_amd import workloads as W

    pysgcpu.set_threads(threads)
    g = pysgcpu.SparseGridCpu(3, (512, 512, 512), [np.float32] * 4, block_edge=8)
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
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, cpu count)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 20 device copies of 1 GiB (reference for the FETCH/WRITE_SIZE counters)")
    args = ap.parse_args()
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

    from openfpm_data_amd import workloads as W
    t_gen = time.perf_counter()
    w = W.shell512_gs()
    ids, masks, data = chunk_arrays(w["size"], w["coords"], [w["u"].astype(np.float32), w["v"].astype(np.float32)])
    del w
    gen_s = time.perf_counter() - t_gen
    dtype = np.float32 if args.dtype == "f32" else np.float64
    r = run_gpu(args, ids, masks, data, dtype)

    npts, nch = r["points"], r["chunks"]
    slots = nch * 512
    tsz = np.dtype(dtype).itemsize
    # algorithmic bytes of one step: per slot u, v read + u', v' written + mask byte;
    # per chunk the 6 face-neighbour entries of the neighbour table
    alg_bytes = slots * (2 * tsz * 2 + 1) + nch * 6 * 4
    kern_s = r["kernel_ms"] * 1e-3
    achieved = alg_bytes / kern_s / 1e9
    value = npts * args.steps / (r["kernel_ms"] * 1e-3 * args.steps) / 1e9
    traffic = traffic_from_profile("traffic_gray_scott_%s_shell512.txt" % args.dtype)

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        c = run_cpu_baseline(ids, masks, data, args.cpu_seconds, threads)
        cpu = {"value": round(c["value"], 6), "unit": "Gpt-updates/s", "cores": threads, "kind": "port",
               "sample": "oracle (oracle/sgcpu.cpp, OpenMP) on the same 512^3 shell FP32 Gray-Scott grid, "
                         "%d steps in %.1f s" % (c["steps"], c["seconds"])}

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
                   "points": npts, "chunks": nch, "chunk_edge": 8, "properties": 2,
                   "params": W.GS_PARAMS},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "wall_s": round(r["wall_s"], 4),
        "setup_s": round(gen_s, 2),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
