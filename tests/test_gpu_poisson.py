"""Batched Poisson solves on the GPU (sg_batch_* / sg_poisson_solve).

* kernel parity: one outer iteration of the mixed form equals the oracle's red-black SOR
  sweeps (FP32) and FP64 residual, bit-exact;
* convergence: mixed (FP32 sweeps + FP64 defect correction) and plain FP64 reach the
  relative residual 1e-10, re-verified with the oracle's FP64 residual on the returned u;
  the mixed solution is compared with the all-FP64 one (error bound stated below);
* continuous batching: a grid's solution is bitwise the same alone and in a mix.
"""
import numpy as np
import pytest

from openfpm_data_amd import _capi as C
from openfpm_data_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def P():
    import torch
    assert torch.cuda.is_available()
    from openfpm_data_amd import poisson
    return poisson


def _small_grids(n, seed, edge_hi=96):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        edge, occ, ids = W.random_chunk_grid(rng, 32, edge_hi)
        out.append({"size": (edge,) * 3, "chunk_ids": ids, "f": rng.uniform(-1, 1, (ids.size, 512))})
    return out


def _oracle_grid(oracle, spec, props):
    g = oracle.SparseGridCpu(3, spec["size"], props, block_edge=8)
    n = spec["chunk_ids"].size
    vals = [np.zeros((n, 512), p) for p in props]
    return g, vals


def _oracle_residual(oracle, spec, u_chunks):
    """FP64 ||f - A u|| / ||f|| with the oracle's residual stencil (u on the same chunks)."""
    g = oracle.SparseGridCpu(3, spec["size"], [np.float64] * 3, block_edge=8)
    n = spec["chunk_ids"].size
    g.insert_block(spec["chunk_ids"], np.ones((n, 512), np.uint8), [u_chunks, spec["f"], np.zeros((n, 512))])
    g.flush(["sright"] * 3)
    g.residual(0, 1, 2)
    r = g.prop(2)
    return float(np.sqrt((r * r).sum()) / np.sqrt((spec["f"] * spec["f"]).sum()))


def test_one_outer_iteration_matches_oracle(P, oracle):
    """max_outer = 1: defect (r = f), K FP32 red-black SOR sweeps of A e = r from e = 0,
    then u = 0 + e and its residual; u bit-exact against the oracle's sweeps."""
    specs = _small_grids(3, 11)
    K, w = 7, 1.37
    grids = [P.poisson_grid(s["size"], s["chunk_ids"], s["f"]) for s in specs]
    res = P.solve(grids, P.make_params(tol=1e-30, omega=w, first_sweeps=K, max_outer=1, use_graph=False))
    for s, g, r in zip(specs, grids, res):
        assert r["sweeps"] == K and r["outer"] == 1 and not r["converged"]
        u_gpu = g.prop(r["solution_prop"])
        n = s["chunk_ids"].size
        o = oracle.SparseGridCpu(3, s["size"], [np.float32, np.float32], block_edge=8)
        o.insert_block(s["chunk_ids"], np.ones((n, 512), np.uint8),
                       [np.zeros((n, 512), np.float32), s["f"].astype(np.float32)])
        o.flush(["sright", "sright"])
        for _ in range(K):
            o.rbsor_sweep(0, 1, w, 0)
            o.rbsor_sweep(0, 1, w, 1)
        e = o.prop(0).astype(np.float64)
        np.testing.assert_array_equal(u_gpu, 0.0 + e)
        rel = _oracle_residual(oracle, s, u_gpu)
        assert abs(rel - r["rel_residual"]) <= 1e-12 * rel


@pytest.mark.parametrize("mode", [C.POISSON_MIXED, C.POISSON_PLAIN_FP64])
@pytest.mark.parametrize("graph", [False, True])
def test_converges_to_1e_10(P, oracle, mode, graph):
    specs = _small_grids(6, 21 + mode)
    grids = [P.poisson_grid(s["size"], s["chunk_ids"], s["f"]) for s in specs]
    res = P.solve(grids, P.make_params(mode=mode, tol=1e-10, use_graph=graph))
    for s, g, r in zip(specs, grids, res):
        assert r["converged"], r
        assert r["rel_residual"] <= 1e-10
        rel = _oracle_residual(oracle, s, g.prop(r["solution_prop"]))
        assert rel <= 1.5e-10, (rel, r)


def test_mixed_against_all_fp64(P, oracle):
    """Error of the FP32+FP64 solution against the all-FP64 one.  Both have relative
    residual <= 1e-10; with kappa(A) = lambda_max / lambda_min <= 12 / 0.05 the relative
    2-norm difference of the solutions is bounded by 2 * kappa * 1e-10 = 4.8e-8."""
    specs = _small_grids(4, 31)
    gm = [P.poisson_grid(s["size"], s["chunk_ids"], s["f"]) for s in specs]
    gp = [P.poisson_grid(s["size"], s["chunk_ids"], s["f"]) for s in specs]
    rm = P.solve(gm, P.make_params(mode=C.POISSON_MIXED, tol=1e-10))
    rp = P.solve(gp, P.make_params(mode=C.POISSON_PLAIN_FP64, tol=1e-10))
    for a, b, ra, rb in zip(gm, gp, rm, rp):
        ua, ub = a.prop(ra["solution_prop"]), b.prop(rb["solution_prop"])
        err = np.linalg.norm(ua - ub) / np.linalg.norm(ub)
        assert err <= 4.8e-8, err


def test_alone_and_in_mix_identical(P):
    """Continuous batching: the iterates of a grid do not depend on its batch mates (each
    grid runs its own sweep counts inside the shared launches), so the solution is bitwise
    the same solved alone and admitted into a running batch of other grids."""
    import torch
    specs = _small_grids(7, 41)
    target = specs[3]
    g_alone = P.poisson_grid(target["size"], target["chunk_ids"], target["f"])
    r_alone = P.solve([g_alone], P.make_params(tol=1e-10))[0]
    u_alone = g_alone.prop(r_alone["solution_prop"])
    stream = torch.cuda.Stream()
    b = P.PoissonBatch(P.make_params(tol=1e-10), stream)
    grids = {}
    for i, s in enumerate(specs[:3]):
        grids[i] = P.poisson_grid(s["size"], s["chunk_ids"], s["f"], stream)
        b.add(grids[i], i)
    done = {}
    for step in range(200):
        if step == 2:  # admitted while the others are mid-solve
            for i, s in enumerate(specs[3:], start=3):
                grids[i] = P.poisson_grid(s["size"], s["chunk_ids"], s["f"], stream)
                b.add(grids[i], i)
        b.step()
        for tag, r in b.poll():
            done[tag] = r
        if len(done) == len(specs):
            break
    assert len(done) == len(specs) and all(r["converged"] for r in done.values())
    u_mix = grids[3].prop(done[3]["solution_prop"])
    assert done[3]["sweeps"] == r_alone["sweeps"]
    np.testing.assert_array_equal(u_mix, u_alone)
    b.close()


@pytest.mark.parametrize("waves,max_wgs", [(1, 0), (2, 7), (8, 0), (4, 33)])
def test_launch_shapes_identical(P, waves, max_wgs):
    """Waves per workgroup and looping waves (a workgroup cap) change only which wave
    sweeps which tile: solutions bitwise equal to the default launch shape."""
    specs = _small_grids(4, 51)
    ga = [P.poisson_grid(s["size"], s["chunk_ids"], s["f"]) for s in specs]
    gb = [P.poisson_grid(s["size"], s["chunk_ids"], s["f"]) for s in specs]
    ra = P.solve(ga, P.make_params(tol=1e-10))
    rb = P.solve(gb, P.make_params(tol=1e-10), tuning={C.TUNE_BATCH_WAVES: waves, C.TUNE_BATCH_MAX_WGS: max_wgs})
    for a, b, x, y in zip(ga, gb, ra, rb):
        assert x["sweeps"] == y["sweeps"] and y["converged"]
        np.testing.assert_array_equal(a.prop(x["solution_prop"]), b.prop(y["solution_prop"]))
