"""Autotuner: cache handling (CPU) and measured choices on the device (GPU)."""
import json

import numpy as np
import pytest

from openfpm_data_amd import autotune as A
from openfpm_data_amd import _capi as C


def test_cache_roundtrip_and_precedence(tmp_path, monkeypatch):
    user = tmp_path / "tune.json"
    monkeypatch.setenv("OPENFPM_AMD_TUNE_CACHE", str(user))
    packaged = tmp_path / "packaged.json"
    packaged.write_text(json.dumps({"k1": {"v": 1}, "k2": {"v": 2}}))
    monkeypatch.setattr(A, "PACKAGED", str(packaged))
    assert A.load_cache() == {"k1": {"v": 1}, "k2": {"v": 2}}
    A.store("k2", {"v": 20})
    A.store("k3", {"v": 3})
    assert json.loads(user.read_text()) == {"k2": {"v": 20}, "k3": {"v": 3}}
    assert A.load_cache() == {"k1": {"v": 1}, "k2": {"v": 20}, "k3": {"v": 3}}


def test_size_class():
    assert A.size_class(48654) == 16 and A.size_class(1) == 0 and A.size_class(0) == 0


def test_candidates_are_valid_settings():
    for c in A.STENCIL_CANDIDATES:
        assert 0 <= c["variant"] <= 8 and c["stores"] in (0, 1, 2, 3)
    for c in A.BATCH_CANDIDATES:
        assert c["waves"] in (1, 2, 4, 8) and c["max_wgs"] >= 0


@pytest.mark.gpu
def test_tuners_pick_a_candidate(tmp_path, monkeypatch):
    import torch
    from openfpm_data_amd import SparseGridGpu, poisson as P, workloads as W
    monkeypatch.setenv("OPENFPM_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    monkeypatch.setattr(A, "PACKAGED", str(tmp_path / "none.json"))
    # stencil on a 64^3 shell
    w = W.shell64_fp64()
    g = SparseGridGpu(3, w["size"], [np.float64] * 2, block_edge=8)
    s = torch.cuda.Stream()
    g.set_stream(s)
    g.insert_coords(w["coords"], [w["values"], np.zeros_like(w["values"])])
    g.flush(["sright", "sright"])
    g.find_neighbours()
    step = lambda i: g.conv_lap(i % 2, 1 - i % 2)  # noqa: E731
    best, times = A.tune_stencil(g, step, "lap", np.float64, s, reps=3)
    assert len(times) == len(A.STENCIL_CANDIDATES)
    assert {k: v for k, v in best.items() if k != "ms"} in A.STENCIL_CANDIDATES
    again, t2 = A.tune_stencil(g, step, "lap", np.float64, s)
    assert t2 == {} and again == best  # cached
    g.close()
    # batched sweeps
    rng = np.random.default_rng(1)
    grids = []
    for _ in range(3):
        edge, occ, ids = W.random_chunk_grid(rng, 32, 64)
        grids.append(P.poisson_grid((edge,) * 3, ids, rng.uniform(-1, 1, (ids.size, 512))))
    best, times = A.tune_batch(grids, sweeps=4)
    assert len(times) == len(A.BATCH_CANDIDATES)
    assert {"waves": best["waves"], "max_wgs": best["max_wgs"]} in A.BATCH_CANDIDATES
    # chunk edge
    e, res = A.tune_chunk_edge(w["size"], w["coords"], [w["values"]], np.float64,
                               lambda g, i: g.conv_lap(i % 2, 1 - i % 2), edges=(4, 8), reps=2)
    assert e in (4, 8) and set(res) == {4, 8}
