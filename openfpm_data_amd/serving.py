"""Served Poisson solves: requests arrive over time, are admitted against an HBM budget and
are solved by continuous batching on one GPU.

Each request is one sparse grid (chunk ids + an FP32 right-hand side per chunk slot) and a
relative-residual tolerance.  The loop

1. moves arrived requests into a FIFO queue;
2. admits queued requests while the bytes reserved by resident grids plus the head's
   footprint fit the budget (``Admission``; the head blocks the queue, so no request is
   starved by smaller ones behind it);
3. builds each admitted grid on the device (insert of whole chunks, flush, neighbour
   table), frees its build temporaries (``sg_trim``) and books its measured resident bytes;
   the builds run in a builder thread on a stream of their own, so that they overlap the
   batched steps of the grids already live (the library's calls release the GIL);
4. runs one outer iteration of every live grid in shared launches (``sg_batch_step``):
   requests that arrived while others were on the GPU join between two outer iterations;
5. hands finished grids back: the solution is copied to the host, the grid is freed and
   its bytes return to the budget.

A request's latency is the time from its arrival to its solution being on the host.  A
grid's iterates do not depend on its batch mates (see ``sg_batch_add``), so the result of a
request is bitwise the same served alone or in any mix.
"""
import collections
import queue
import threading
import time

import numpy as np

from . import _capi as C
from . import poisson as P

# resident bytes per chunk slot of a Poisson grid: u, u2, f (8 B each), e, r (4 B each), mask
SLOT_BYTES = 3 * 8 + 2 * 4 + 1
# per chunk: chunk id (8 B) + NNStar table row (8 x 4 B)
CHUNK_BYTES = 8 + 8 * 4


def resident_bytes(nchunks):
    """Device bytes a built Poisson grid of `nchunks` chunks keeps (background chunk included)."""
    return (nchunks + 1) * (512 * SLOT_BYTES) + nchunks * CHUNK_BYTES


def build_bytes(nchunks):
    """Upper estimate of the extra bytes held while the grid is built (insert records of
    key + mask + values per slot, sort keys and permutation, inputs staged on the device),
    released by ``sg_trim`` once the grid is built."""
    slots = nchunks * 512
    return slots * (8 + 1 + SLOT_BYTES) * 2 + slots * 16 + slots * 8 + (1 << 20)


class Admission:
    """FIFO admission against a byte budget.  ``reserve`` is booked at admission (footprint
    while building), ``settle`` replaces it with the measured resident size, ``release``
    returns it."""

    def __init__(self, budget_bytes):
        self.budget = int(budget_bytes)
        self.used = 0
        self.peak = 0
        self.queue = collections.deque()
        self.booked = {}

    def arrive(self, req):
        self.queue.append(req)

    def footprint(self, req):
        n = len(req["chunk_ids"])
        return resident_bytes(n) + build_bytes(n)

    def admit(self, max_n=None):
        """Requests admitted now (FIFO, head-of-line blocking).  A request whose footprint
        alone exceeds the budget is rejected (returned in the second list)."""
        out, rejected = [], []
        while self.queue and (max_n is None or len(out) < max_n):
            req = self.queue[0]
            need = self.footprint(req)
            if need > self.budget:
                rejected.append(self.queue.popleft())
                continue
            if self.used + need > self.budget:
                break
            self.queue.popleft()
            self.used += need
            self.booked[req["id"]] = need
            self.peak = max(self.peak, self.used)
            out.append(req)
        return out, rejected

    def settle(self, req_id, actual):
        self.used += int(actual) - self.booked[req_id]
        self.booked[req_id] = int(actual)
        self.peak = max(self.peak, self.used)

    def release(self, req_id):
        self.used -= self.booked.pop(req_id)


def grid_device_bytes(grid):
    lib = C.load()
    import ctypes
    v = ctypes.c_int64()
    C.check(lib.sg_device_bytes(grid._h, ctypes.byref(v)))
    return v.value


class PoissonServer:
    """Continuous-batching Poisson server on one GPU.

    params: ``poisson.make_params(...)`` (method and inner-solve settings; each request
    carries its own tolerance).  budget_bytes: HBM budget for resident grids.  tuning:
    {key: value} of sg_batch_set_tuning (the sweep kernel's launch shape).  build_thread:
    build admitted grids in a builder thread (on a stream of its own) while the batch
    steps; False builds them inline, between two steps."""

    def __init__(self, budget_bytes, params=None, keep_solutions=False, tuning=None, build_thread=True):
        import torch
        self.torch = torch
        self.params = params if params is not None else P.make_params()
        self.stream = torch.cuda.Stream()
        self.build_stream = torch.cuda.Stream()
        self.batch = P.PoissonBatch(self.params, self.stream)
        for k, v in (tuning or {}).items():
            self.batch.set_tuning(k, v)
        self.adm = Admission(budget_bytes)
        self.keep = keep_solutions
        self.threaded = build_thread

    def close(self):
        self.batch.close()

    def _build(self, req, stream):
        """Builds the request's grid on `stream`; returns (grid, resident bytes) once the
        grid is complete (sg_trim waits for the stream)."""
        torch = self.torch
        with torch.cuda.stream(stream):
            f = torch.from_numpy(np.ascontiguousarray(req["f"])).to("cuda", non_blocking=False)
            ids = torch.from_numpy(np.asarray(req["chunk_ids"], np.int64)).to("cuda")
            g = P.poisson_grid(req["size"], ids, f, stream)
            del f, ids
        C.check(C.load().sg_trim(g._h))
        if stream is not self.stream:
            # later work on the grid runs on the batch's stream (the build stream is idle:
            # sg_trim waited for it, so the switch does not wait)
            g.set_stream(self.stream)
        return g, grid_device_bytes(g)

    def _builder(self, inq, outq):
        while True:
            req = inq.get()
            if req is None:
                return
            try:
                g, nbytes = self._build(req, self.build_stream)
                outq.put((req, g, nbytes, None))
            except BaseException as e:  # handed to the serving loop, which raises it
                outq.put((req, None, 0, e))

    def run(self, requests, arrivals):
        """Serve `requests` arriving at times `arrivals` (seconds from the start, sorted).
        Returns (per-request records, summary)."""
        order = np.argsort(np.asarray(arrivals), kind="stable")
        reqs = [requests[i] for i in order]
        arr = [float(arrivals[i]) for i in order]
        nxt = 0
        live = {}
        records = {}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        steps = 0
        max_live = 0
        t_build = t_step = t_fetch = 0.0
        inq, outq = queue.Queue(), queue.Queue()
        worker = None
        if self.threaded:
            worker = threading.Thread(target=self._builder, args=(inq, outq), daemon=True)
            worker.start()
        building = 0

        def join(item):
            # a built grid joins the batch
            r, g, nbytes, err = item
            if err is not None:
                raise err
            self.adm.settle(r["id"], nbytes)
            self.batch.add(g, r["id"], r["tol"])
            live[r["id"]] = (r, g, now())

        try:
            while len(records) < len(reqs):
                t = now()
                while nxt < len(reqs) and arr[nxt] <= t:
                    r = reqs[nxt]
                    r["_arrival"] = arr[nxt]
                    self.adm.arrive(r)
                    nxt += 1
                admitted, rejected = self.adm.admit()
                for r in rejected:
                    records[r["id"]] = {"id": r["id"], "rejected": True, "arrival": r["_arrival"]}
                for r in admitted:
                    if worker is not None:
                        inq.put(r)
                        building += 1
                    else:
                        tb = now()
                        g, nbytes = self._build(r, self.stream)
                        join((r, g, nbytes, None))
                        t_build += now() - tb
                while building:
                    try:
                        item = outq.get_nowait()
                    except queue.Empty:
                        break
                    building -= 1
                    join(item)
                max_live = max(max_live, len(live))
                if self.batch.live() == 0:
                    if building:
                        tb = now()
                        item = outq.get()  # nothing to step: wait for a build
                        building -= 1
                        join(item)
                        t_build += now() - tb
                    elif nxt < len(reqs) and not self.adm.queue:
                        time.sleep(max(0.0, arr[nxt] - now()))
                    continue
                ts = now()
                self.batch.step()
                steps += 1
                t_step += now() - ts
                for tag, res in self.batch.poll():
                    r, g, t_adm = live.pop(tag)
                    tf = now()
                    u = g.prop(res["solution_prop"])  # solution to the host
                    t_done = now()
                    t_fetch += t_done - tf
                    rec = {"id": tag, "arrival": r["_arrival"], "admitted": t_adm, "done": t_done,
                           "latency": t_done - r["_arrival"], "chunks": len(r["chunk_ids"]),
                           "points": len(r["chunk_ids"]) * 512, "tol": r["tol"], **res}
                    if self.keep:
                        rec["u"] = u
                    records[tag] = rec
                    g.close()
                    self.adm.release(tag)
        finally:
            if worker is not None:
                inq.put(None)
                worker.join()
        wall = now()
        recs = [records[r["id"]] for r in requests]  # in the caller's order
        done = [x for x in recs if not x.get("rejected")]
        lat = np.array([x["latency"] for x in done]) if done else np.zeros(1)
        summary = {
            "requests": len(reqs), "served": len(done), "rejected": len(recs) - len(done),
            "converged": int(sum(x["converged"] for x in done)),
            "wall_s": wall, "solves_per_s": len(done) / wall if wall > 0 else 0.0,
            "point_updates_per_s": sum(x["points"] * x["sweeps"] for x in done) / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "outer_steps": steps, "max_live": max_live,
            "budget_bytes": self.adm.budget, "peak_reserved_bytes": self.adm.peak,
            "build_s": t_build, "step_s": t_step, "fetch_s": t_fetch,
        }
        return recs, summary


def poisson_arrivals(n, rate, seed):
    """Arrival times of a Poisson process of `rate` requests/s (exponential gaps)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))
