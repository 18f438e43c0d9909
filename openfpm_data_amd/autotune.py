"""Autotuner: launch shapes and chunk edge chosen by measurement on the device, cached in
JSON.

What is tuned
* 8^3 star stencils (Laplacian, heat, Gray-Scott): the kernel variant (one chunk per
  workgroup; persistent waves with a workgroups-per-XCD count; groups of 2 / 4 / 8 chunks
  per workgroup sharing x faces through LDS, whose LDS footprint grows with the group,
  with one LDS tile per property or one tile shared by the two Gray-Scott properties), the
  store mode (whole rows, plain or non-temporal) and the load mode (every row, or rows
  without points skipped by the occupancy table) -- ``tune_stencil``;
* batched red-black sweeps: waves (chunk tiles) per workgroup, i.e. the LDS tile footprint
  per workgroup, and a workgroup cap that makes waves loop over tiles -- ``tune_batch``;
* the chunk edge of a point set for a stencil workload: grids are built at each candidate
  edge and one stencil step is timed per point -- ``tune_chunk_edge``.

Each result is stored under a key naming the device, the workload kind, the data type and a
size class (log2 of the chunk count), in a JSON file: ``$OPENFPM_AMD_TUNE_CACHE`` or
``~/.cache/openfpm_data_amd/tune.json``.  ``openfpm_data_amd/tuned.json`` holds measured
entries shipped with the package; the user cache takes precedence.  Timing uses HIP events
recorded on the stream the kernels run on.
"""
import json
import math
import os
import time

import numpy as np

from . import _capi as C

_HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGED = os.path.join(_HERE, "tuned.json")

STENCIL_CANDIDATES = [
    {"variant": 0, "stores": 3},
    {"variant": 1, "wg_per_xcd": 160, "stores": 3},
    {"variant": 5, "stores": 3},
    {"variant": 3, "stores": 3},
    {"variant": 4, "stores": 3},
    {"variant": 3, "stores": 2},
    {"variant": 4, "stores": 2},
    {"variant": 4, "stores": 1},
    {"variant": 6, "stores": 3},
    {"variant": 7, "stores": 3},
    {"variant": 8, "stores": 3},
    {"variant": 3, "stores": 3, "lskip": 2},
    {"variant": 3, "stores": 3, "lskip": 3},
    {"variant": 4, "stores": 3, "lskip": 2},
    {"variant": 6, "stores": 3, "lskip": 2},
    {"variant": 7, "stores": 3, "lskip": 2},
    {"variant": 3, "stores": 3, "lskip": 4},
    {"variant": 6, "stores": 3, "lskip": 4},
    {"variant": 7, "stores": 3, "lskip": 4},
]

BATCH_CANDIDATES = [
    {"waves": 1, "max_wgs": 0},
    {"waves": 2, "max_wgs": 0},
    {"waves": 4, "max_wgs": 0},
    {"waves": 8, "max_wgs": 0},
    {"waves": 4, "max_wgs": 4096},
    {"waves": 8, "max_wgs": 2048},
]


def cache_path():
    p = os.environ.get("OPENFPM_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "openfpm_data_amd", "tune.json")


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load_cache():
    """Packaged entries overlaid by the user cache."""
    out = _read(PACKAGED)
    out.update(_read(cache_path()))
    return out


def store(key, value):
    path = cache_path()
    data = _read(path)
    data[key] = value
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the result is still returned, only not remembered


def device_name():
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return getattr(p, "gcnArchName", p.name).split(":")[0]


def size_class(nchunks):
    return int(round(math.log2(max(1, nchunks))))


def make_key(kind, dtype, nchunks):
    return "%s/%s/%s/c%d" % (device_name(), kind, np.dtype(dtype).name, size_class(nchunks))


def _event_ms(stream, fn, reps):
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(reps):
        fn(i)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def apply_stencil_tuning(grid, cfg):
    grid.set_tuning(C.TUNE_STENCIL_VARIANT, cfg["variant"])
    if "wg_per_xcd" in cfg:
        grid.set_tuning(C.TUNE_STENCIL_WG_PER_XCD, cfg["wg_per_xcd"])
    grid.set_tuning(C.TUNE_STENCIL_STORES, cfg["stores"])
    grid.set_tuning(C.TUNE_STENCIL_SKIP_EMPTY_ROWS, cfg.get("lskip", 0))


def tune_stencil(grid, step, kind, dtype, stream, reps=30, force=False, candidates=None):
    """Pick the 8^3 star-stencil launch shape for `grid`.  step(i) runs one stencil
    application (alternating buffers by i); the grid's tuning is left at the winner.
    Returns (config, {candidate: ms}) -- the timings are empty for a cached result."""
    key = make_key("stencil-" + kind, dtype, grid.num_chunks())
    cache = load_cache()
    if not force and key in cache:
        apply_stencil_tuning(grid, cache[key])
        return cache[key], {}
    times = {}
    best, best_ms = None, float("inf")
    for cfg in (candidates or STENCIL_CANDIDATES):
        apply_stencil_tuning(grid, cfg)
        _event_ms(stream, step, 4)  # warm-up
        ms = _event_ms(stream, step, reps)
        times[json.dumps(cfg, sort_keys=True)] = ms
        if ms < best_ms:
            best, best_ms = dict(cfg), ms
    best["ms"] = round(best_ms, 5)
    apply_stencil_tuning(grid, best)
    store(key, best)
    return best, times


def tune_batch(grids, mode=C.POISSON_MIXED, sweeps=16, force=False, candidates=None):
    """Pick the batched sweep kernel's launch shape for a set of Poisson grids (built by
    poisson.poisson_grid).  Each candidate runs one outer iteration (defect + `sweeps`
    red-black sweeps, fixed omega) of the whole set; the grids' values are scratch here.
    Returns (config, {candidate: ms})."""
    import torch
    from . import poisson as P
    total = sum(g.num_chunks() for g in grids)
    key = make_key("batch-sweep-%s" % ("mixed" if mode == C.POISSON_MIXED else "fp64"),
                   np.float32 if mode == C.POISSON_MIXED else np.float64, total)
    cache = load_cache()
    if not force and key in cache:
        return cache[key], {}
    stream = torch.cuda.Stream()
    params = P.make_params(mode=mode, tol=1e-30, omega=1.7, first_sweeps=sweeps, max_outer=1, use_graph=False)
    times = {}
    best, best_ms = None, float("inf")
    for rep in range(2):  # the first pass warms allocations and code objects
        for cfg in (candidates or BATCH_CANDIDATES):
            b = P.PoissonBatch(params, stream)
            b.set_tuning(C.TUNE_BATCH_WAVES, cfg["waves"])
            b.set_tuning(C.TUNE_BATCH_MAX_WGS, cfg["max_wgs"])
            for i, g in enumerate(grids):
                b.add(g, i)
            stream.synchronize()
            t0 = time.perf_counter()
            b.step()
            stream.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            b.close()
            if rep == 1:
                times[json.dumps(cfg, sort_keys=True)] = ms
                if ms < best_ms:
                    best, best_ms = dict(cfg), ms
    best["ms"] = round(best_ms, 4)
    store(key, best)
    return best, times


def tune_chunk_edge(size, coords, values, dtype, stencil, edges=(4, 8), reps=10, force=False):
    """Chunk edge for a point set under a stencil workload: a grid is built at each edge and
    one application of `stencil(grid, i)` is timed; returns (edge, {edge: ns per point}).
    `values` are the per-point property values (one array per property)."""
    import torch
    from .sparsegrid import SparseGridGpu
    key = "%s/chunk-edge/%s/n%d" % (device_name(), np.dtype(dtype).name, size_class(len(coords)))
    cache = load_cache()
    if not force and key in cache:
        return cache[key]["edge"], {}
    stream = torch.cuda.Stream()
    d_c = torch.from_numpy(np.ascontiguousarray(coords, np.int32)).cuda()
    d_v = [torch.from_numpy(np.ascontiguousarray(v.astype(dtype))).cuda() for v in values]
    res = {}
    for e in edges:
        g = SparseGridGpu(3, size, [dtype] * (2 * len(values)), block_edge=e)
        g.set_stream(stream)
        g.insert_coords(d_c, d_v + [torch.zeros_like(d_v[0])] * len(values))
        g.flush(["sright"] * (2 * len(values)))
        g.find_neighbours()
        g.synchronize()
        stencil(g, 0)
        ms = _event_ms(stream, lambda i: stencil(g, i), reps)
        res[e] = ms * 1e6 / len(coords)
        g.close()
    best = min(res, key=res.get)
    store(key, {"edge": int(best), "ns_per_point": {str(k): round(v, 6) for k, v in res.items()}})
    return best, res
