"""Batched Poisson solves (-lap u = f, u = 0 outside the existing points) over
libsg_hip.so's continuous-batching engine (sg_batch_*).

Grid property layout: 0 u (f64), 1 u2 (f64), 2 f (f64), 3 e (f32), 4 r (f32); see
POISSON_PROPS.  ``PoissonBatch`` admits grids at any time; each ``step()`` runs one outer
iteration (FP64 defect correction + residual norms, then the FP32 red-black SOR inner
block) for every live grid in shared launches, and finished grids come back from ``poll()``.
"""
import ctypes

import numpy as np

from . import _capi as C
from .sparsegrid import SparseGridGpu

POISSON_PROPS = [np.float64, np.float64, np.float64, np.float32, np.float32]
U, U2, F, E, R = range(5)


def make_params(mode=C.POISSON_MIXED, tol=1e-10, omega=0.0, omega0=1.5, inner_target=0.0, first_sweeps=24,
                max_sweeps=20000, max_outer=100, use_graph=True):
    p = C.PoissonParams()
    p.mode = mode
    p.u_prop, p.u2_prop, p.f_prop, p.e_prop, p.r_prop = U, U2, F, E, R
    p.tol = tol
    p.omega = omega
    p.omega0 = omega0
    p.inner_target = inner_target
    p.first_sweeps = first_sweeps
    p.max_sweeps = max_sweeps
    p.max_outer = max_outer
    p.use_graph = 1 if use_graph else 0
    return p


def result_dict(r):
    return {"converged": bool(r.converged), "outer": r.outer, "sweeps": r.sweeps,
            "solution_prop": r.solution_prop, "rel_residual": r.rel_residual, "omega": r.omega}


def poisson_grid(size, chunk_ids, f_chunks, stream=None):
    """A grid holding the solver layout: chunks `chunk_ids` fully active, f given per chunk
    slot ([n, 512], host or device), u = u2 = e = r = 0."""
    g = SparseGridGpu(3, size, POISSON_PROPS, block_edge=8)
    if stream is not None:
        g.set_stream(stream)
    n = len(chunk_ids)
    dev = hasattr(f_chunks, "data_ptr")
    if dev:
        import torch
        ids = chunk_ids if hasattr(chunk_ids, "data_ptr") else torch.from_numpy(np.asarray(chunk_ids, np.int64)).cuda()
        masks = torch.ones((n, 512), dtype=torch.uint8, device="cuda")
        z64 = torch.zeros((n, 512), dtype=torch.float64, device="cuda")
        z32 = torch.zeros((n, 512), dtype=torch.float32, device="cuda")
        g.insert_block(ids, masks, [z64, z64, f_chunks.to(torch.float64), z32, z32])
    else:
        masks = np.ones((n, 512), np.uint8)
        z64 = np.zeros((n, 512))
        z32 = np.zeros((n, 512), np.float32)
        g.insert_block(np.asarray(chunk_ids, np.int64), masks, [z64, z64, np.asarray(f_chunks, np.float64), z32, z32])
    g.flush(["sright"] * 5)
    g.find_neighbours()
    return g


class PoissonBatch:
    """Continuous-batching Poisson engine (one per stream)."""

    def __init__(self, params, stream=None):
        self._lib = C.load()
        self.params = params
        raw = getattr(stream, "cuda_stream", stream) if stream is not None else None
        h = ctypes.c_void_p()
        C.check(self._lib.sg_batch_create(ctypes.byref(params), ctypes.c_void_p(raw), ctypes.byref(h)))
        self._h = h

    def add(self, grid, tag, tol=0.0):
        """Admit `grid` (tol > 0: its own relative residual target)."""
        C.check(self._lib.sg_batch_add(self._h, grid._h, int(tag), float(tol)))

    def live(self):
        return self._lib.sg_batch_live(self._h)

    def set_tuning(self, key, value):
        C.check(self._lib.sg_batch_set_tuning(self._h, int(key), int(value)))

    def step(self):
        C.check(self._lib.sg_batch_step(self._h))

    def poll(self, max_n=1024):
        tags = (ctypes.c_int64 * max_n)()
        res = (C.PoissonResult * max_n)()
        n = ctypes.c_int()
        C.check(self._lib.sg_batch_poll(self._h, tags, res, max_n, ctypes.byref(n)))
        return [(tags[i], result_dict(res[i])) for i in range(n.value)]

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(grids, params, tuning=None):
    """Solve every grid in one batch; returns one result dict per grid.  tuning: optional
    {key: value} of sg_batch_set_tuning (then the batch runs through PoissonBatch)."""
    if tuning:
        import torch
        b = PoissonBatch(params, torch.cuda.current_stream())
        for k, v in tuning.items():
            b.set_tuning(k, v)
        for i, g in enumerate(grids):
            b.add(g, i)
        out = [None] * len(grids)
        while b.live():
            b.step()
            for tag, r in b.poll():
                out[tag] = r
        for tag, r in b.poll():
            out[tag] = r
        b.close()
        return out
    lib = C.load()
    arr = (ctypes.c_void_p * len(grids))(*[g._h.value for g in grids])
    res = (C.PoissonResult * len(grids))()
    C.check(lib.sg_poisson_solve(arr, len(grids), ctypes.byref(params), res))
    return [result_dict(r) for r in res]
