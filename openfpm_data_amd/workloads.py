"""Synthetic workloads of the sparse-grid benchmarks (deterministic, seeded numpy PCG64).

shell_points     points of a spherical shell lo <= |p - centre| < hi (centre = N/2, FP64 radius)
shell64_fp64     64^3 shell r in [20, 24), one FP64 property, values uniform(-1, 1), seed 0
shell512_gs      512^3 shell r in [200, 232), u = 1, v = 0, then 50 cubes of edge 8 (corners at
                 50 existing points drawn with seed 1) set u = 0.5, v = 0.25
blobs1024        2^24 point inserts in a 1024^3 domain from 64 Gaussian blobs (sigma 24 cells,
                 centres uniform in [4 sigma, 1024 - 4 sigma), seed 2), 30% of the inserts repeat
                 earlier coordinates; int32 and uint32 values uniform in [0, 2^20)
"""
import numpy as np


def shell_points(n, lo, hi, dim=3):
    """Integer coordinates [k, dim] (x fastest order) with lo <= r < hi about n/2."""
    c = n / 2.0
    ax = np.arange(n, dtype=np.float64) - c
    out = []
    if dim == 3:
        # slab by slab in z to bound memory at n = 512
        yy, xx = np.meshgrid(ax, ax, indexing="ij")
        r2xy = xx * xx + yy * yy
        for z in range(n):
            r = np.sqrt(r2xy + ax[z] * ax[z])
            ys, xs = np.nonzero((r >= lo) & (r < hi))
            if xs.size:
                out.append(np.stack([xs, ys, np.full_like(xs, z)], axis=1))
        if not out:
            return np.zeros((0, 3), dtype=np.int32)
        return np.concatenate(out).astype(np.int32)
    yy, xx = np.meshgrid(ax, ax, indexing="ij")
    r = np.sqrt(xx * xx + yy * yy)
    ys, xs = np.nonzero((r >= lo) & (r < hi))
    return np.stack([xs, ys], axis=1).astype(np.int32)


def shell64_fp64():
    pts = shell_points(64, 20.0, 24.0)
    vals = np.random.default_rng(0).uniform(-1.0, 1.0, pts.shape[0])
    return {"size": (64, 64, 64), "coords": pts, "values": vals}


def shell512_gs(n=512, lo=200.0, hi=232.0, ncubes=50, edge=8, seed=1):
    pts = shell_points(n, lo, hi)
    rng = np.random.default_rng(seed)
    corners = pts[rng.choice(pts.shape[0], size=ncubes, replace=False)].astype(np.int64)
    u = np.ones(pts.shape[0], dtype=np.float64)
    v = np.zeros(pts.shape[0], dtype=np.float64)
    mark = np.zeros((n, n, n), dtype=np.bool_)  # [z, y, x]
    for c in corners:
        x, y, z = (int(t) for t in c)
        mark[z:z + edge, y:y + edge, x:x + edge] = True
    inside = mark[pts[:, 2], pts[:, 1], pts[:, 0]]
    del mark
    u[inside] = 0.5
    v[inside] = 0.25
    return {"size": (n, n, n), "coords": pts, "u": u, "v": v, "corners": corners}


def blobs1024(n_insert=1 << 24, n=1024, nblobs=64, sigma=24.0, dup_frac=0.3, seed=2):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(4 * sigma, n - 4 * sigma, size=(nblobs, 3))
    n_dup = int(round(dup_frac * n_insert))
    n_base = n_insert - n_dup
    which = rng.integers(0, nblobs, n_base)
    base = np.rint(centres[which] + sigma * rng.standard_normal((n_base, 3)))
    base = np.clip(base, 0, n - 1).astype(np.int32)
    dup = base[rng.integers(0, n_base, n_dup)]
    coords = np.concatenate([base, dup])
    coords = coords[rng.permutation(n_insert)]
    vi = rng.integers(0, 1 << 20, n_insert).astype(np.int32)
    vu = rng.integers(0, 1 << 20, n_insert).astype(np.uint32)
    return {"size": (n, n, n), "coords": np.ascontiguousarray(coords), "int32": vi, "uint32": vu}


# Gray-Scott parameters of the 512^3 reaction-diffusion workload
GS_PARAMS = {"Du": 2e-5, "Dv": 1e-5, "F": 0.04, "k": 0.06, "dt": 1.0}


def random_chunk_grid(rng, edge_lo, edge_hi, occ_lo=0.05, occ_hi=0.40):
    """One sparse grid of the batched / served Poisson workloads: domain edge log-uniform in
    [edge_lo, edge_hi] rounded to a multiple of 8, a random subset of round(occ * nb^3)
    chunks (occ uniform in [occ_lo, occ_hi]) fully active."""
    e = float(np.exp(rng.uniform(np.log(edge_lo), np.log(edge_hi))))
    edge = int(min(edge_hi, max(edge_lo, 8 * round(e / 8))))
    nb = edge // 8
    occ = float(rng.uniform(occ_lo, occ_hi))
    k = max(1, int(round(occ * nb ** 3)))
    ids = np.sort(rng.choice(nb ** 3, size=k, replace=False)).astype(np.int64)
    return edge, occ, ids


def poisson_batch(n=256, seed=3):
    """256 independent grids (edge log-uniform in [32, 256], chunk occupancy in [5%, 40%]),
    f uniform(-1, 1) on the active points (FP64), u = 0 outside."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        edge, occ, ids = random_chunk_grid(rng, 32, 256)
        f = rng.uniform(-1.0, 1.0, (ids.size, 512))
        out.append({"size": (edge, edge, edge), "occupancy": occ, "chunk_ids": ids, "f": f})
    return out


def serving_requests(n, seed=4):
    """Request stream of the served Poisson workload: edge log-uniform in [32, 384],
    occupancy in [5%, 40%], FP32 RHS uniform(-1, 1), tolerance 1e-10 or 1e-6 (50/50)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        edge, occ, ids = random_chunk_grid(rng, 32, 384)
        f = rng.uniform(-1.0, 1.0, (ids.size, 512)).astype(np.float32)
        tol = 1e-10 if rng.random() < 0.5 else 1e-6
        out.append({"id": i, "size": (edge, edge, edge), "occupancy": occ, "chunk_ids": ids, "f": f, "tol": tol})
    return out
