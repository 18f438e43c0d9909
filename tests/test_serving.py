"""Serving loop: admission logic (CPU) and served results against solving alone (GPU)."""
import numpy as np
import pytest

from openfpm_data_amd import serving as S
from openfpm_data_amd import workloads as W


def _req(i, n):
    return {"id": i, "chunk_ids": np.arange(n), "size": (64, 64, 64), "tol": 1e-6}


def test_admission_fifo_budget_and_release():
    fp = S.resident_bytes(10) + S.build_bytes(10)
    a = S.Admission(2 * fp + fp // 2)
    for i in range(4):
        a.arrive(_req(i, 10))
    got, rej = a.admit()
    assert [r["id"] for r in got] == [0, 1] and not rej
    assert a.used == 2 * fp and a.peak == 2 * fp
    a.settle(0, S.resident_bytes(10))  # built: only the resident part stays booked
    got, _ = a.admit()
    assert [r["id"] for r in got] == [2]
    a.release(1)
    got, _ = a.admit()
    assert [r["id"] for r in got] == [3]
    for i in (0, 2, 3):
        a.release(i)
    assert a.used == 0 and not a.booked and a.peak <= a.budget


def test_admission_head_of_line_and_reject():
    small = S.resident_bytes(1) + S.build_bytes(1)
    a = S.Admission(3 * small)
    a.arrive(_req(0, 1))
    a.arrive(_req(1, 1000))  # never fits: rejected, does not block the queue
    a.arrive(_req(2, 1))
    got, rej = a.admit()
    assert [r["id"] for r in got] == [0, 2] and [r["id"] for r in rej] == [1]
    # head-of-line: a request that fits only later blocks the smaller ones behind it
    b = S.Admission(3 * small)
    b.arrive(_req(0, 1))
    b.arrive(_req(1, 1))
    b.arrive(_req(2, 1))
    got, _ = b.admit(max_n=1)
    assert [r["id"] for r in got] == [0]


def test_arrivals_are_poisson():
    t = S.poisson_arrivals(20000, 50.0, 7)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0) and abs(gaps.mean() - 1 / 50.0) < 0.05 / 50.0
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.05  # exponential: cv = 1


def test_request_generator_shapes():
    reqs = W.serving_requests(40)
    for r in reqs:
        e = r["size"][0]
        assert e % 8 == 0 and 32 <= e <= 384
        assert r["f"].dtype == np.float32 and r["f"].shape == (len(r["chunk_ids"]), 512)
        assert np.all(np.diff(r["chunk_ids"]) > 0) and r["chunk_ids"][-1] < (e // 8) ** 3
        assert r["tol"] in (1e-10, 1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("build_thread", [True, False])
def test_served_results_match_alone(build_thread):
    """Requests served with overlapping lifetimes under a budget that forces queueing give
    bitwise the same solutions as each solved alone (grids built by the builder thread or
    inline)."""
    import torch
    from openfpm_data_amd import poisson as P

    rng = np.random.default_rng(9)
    reqs = []
    for i in range(8):
        edge, occ, ids = W.random_chunk_grid(rng, 32, 96)
        reqs.append({"id": i, "size": (edge,) * 3, "chunk_ids": ids,
                     "f": rng.uniform(-1, 1, (ids.size, 512)).astype(np.float32),
                     "tol": 1e-10 if i % 2 == 0 else 1e-6})
    biggest = max(S.resident_bytes(len(r["chunk_ids"])) + S.build_bytes(len(r["chunk_ids"])) for r in reqs)
    srv = S.PoissonServer(budget_bytes=3 * biggest, keep_solutions=True, build_thread=build_thread)
    # four requests arrive together (their lifetimes overlap whatever the solve speed: the
    # budget admits at least two), the others one by one while those are on the GPU
    arrivals = np.concatenate([np.zeros(4), 0.002 * np.arange(1, 5)])
    recs, summ = srv.run(reqs, arrivals)
    srv.close()
    assert summ["served"] == 8 and summ["converged"] == 8
    assert summ["peak_reserved_bytes"] <= summ["budget_bytes"]
    assert summ["max_live"] >= 2  # lifetimes overlapped
    for r, rec in zip(reqs, recs):
        g = P.poisson_grid(r["size"], r["chunk_ids"], r["f"].astype(np.float64))
        res = P.solve([g], P.make_params(tol=r["tol"]))[0]
        assert res["sweeps"] == rec["sweeps"] and res["outer"] == rec["outer"]
        np.testing.assert_array_equal(g.prop(res["solution_prop"]), rec["u"])
        assert rec["rel_residual"] <= r["tol"]
        g.close()
    torch.cuda.synchronize()
