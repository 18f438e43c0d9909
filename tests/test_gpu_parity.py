"""GPU parity: libsg_hip.so (through its C ABI) against the oracle on the same seeded
inputs.  Integer/index/mask results bit-exact; FP results bit-exact as well (both sides
evaluate the same operation order without FMA contraction)."""
import ctypes

import numpy as np
import pytest

from openfpm_data_amd import _capi as C
from openfpm_data_amd import workloads as W
from tests import scenarios as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from openfpm_data_amd import SparseGridGpu
    return SparseGridGpu


@pytest.mark.parametrize("name", ["tag_boundaries_2d_1", "tag_boundaries_2d_2"])
def test_tag_boundaries_known_answer(G, golden, name):
    case = golden[name]
    g = S.known_tag_boundaries(G, case)
    got = S.padding_ids(g, int(np.prod(case["size"])))
    assert sorted(got.tolist()) == sorted(case["padding_linear_ids"])


@pytest.mark.parametrize("name", ["insert_2d", "insert_3d"])
def test_insert_known_answer(G, golden, name):
    case = golden[name]
    dim, size = case["dim"], case["size"]
    g = G(dim, size, [np.float32], block_edge=case["block_edge"])
    g.set_background_value(0, case["background"])
    grids = np.meshgrid(*[np.arange(s) for s in size], indexing="ij")
    coords = np.stack([x.reshape(-1) for x in grids], axis=1)
    g.insert_coords(coords, [coords[:, 0].astype(np.float32)])
    g.flush(["smax"])
    keys = np.arange(256)
    np.testing.assert_array_equal(g.get(0, keys), g.inv_lin_id(keys)[:, 0].astype(np.float32))


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("seed", range(12))
def test_flush_random_programs(G, oracle, seed, path):
    """Random insert / flush programs against the oracle and the dictionary model, with
    the point records grouped by chunk (path 0: records of 16 bytes, of 32 bytes, and
    beyond 32 bytes, where the point-key flush runs; one to four properties, whose loops
    the new-chunk fold unrolls, and six, which it loops over) and by point key (path 1)."""
    rng = np.random.default_rng(200 + seed)
    dim = 3 if seed % 2 else 2
    edge = [2, 4, 8][seed % 3]
    size = [int(rng.integers(edge, 6 * edge)) for _ in range(dim)]
    dtypes = [[np.float64], [np.int32, np.uint32], [np.float32, np.int64, np.uint8],
              [np.uint16, np.int8, np.float64, np.uint64], [np.float64, np.int64, np.float64, np.uint64],
              [np.uint8, np.int16, np.float32, np.int8, np.uint16, np.int32],
              ][seed % 6]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=3, nrec=150)
    bgs = [3, 1, 2, 5, 7, 9][: len(dtypes)]
    a = S.run_flush_program(G, dim, edge, size, dtypes, steps, background=bgs, flush_path=path)
    b = S.run_flush_program(oracle.SparseGridCpu, dim, edge, size, dtypes, steps, background=bgs)
    S.assert_same_state(a, b)
    S.compare_to_model(a, S.model_flush(dim, edge, size, dtypes, steps, background=bgs), dtypes)


@pytest.mark.parametrize("path", [0, 1])
@pytest.mark.parametrize("seed", range(4))
def test_flush_crowded_chunks(G, oracle, seed, path):
    """Thousands of point records in a few chunks, most of them on a handful of slots
    (many batches of 64 records per chunk and many records per slot within a batch), with
    order-sensitive folds (FP sadd of non-integer values, sright, sleft): bit-exact
    against the oracle on both flush paths."""
    rng = np.random.default_rng(900 + seed)
    dim, edge = [(3, 4), (3, 8), (2, 16), (2, 2)][seed]
    size = [4 * edge] * dim
    cs = edge ** dim
    dtypes = [np.float32, np.float64, np.int32]
    steps = []
    for f in range(2):
        n = 5000
        ch = rng.integers(0, 3, n)
        off = np.where(rng.random(n) < 0.8, rng.integers(0, 4, n), rng.integers(0, cs, n))
        keys = ch.astype(np.int64) * cs + off
        masks = rng.choice(np.array([1, 1, 1, 3, 0, 2], dtype=np.uint8), n)
        vals = [rng.uniform(-50, 50, n).astype(np.float32), rng.uniform(-50, 50, n),
                rng.integers(-1000, 1000, n).astype(np.int32)]
        steps.append(([("points", keys, masks, vals)], [["sadd", "sright", "sleft"], ["sadd", "sadd", "smax"]][f]))
    a = S.run_flush_program(G, dim, edge, size, dtypes, steps, background=[1, 2, 3], flush_path=path)
    b = S.run_flush_program(oracle.SparseGridCpu, dim, edge, size, dtypes, steps, background=[1, 2, 3])
    S.assert_same_state(a, b)
    S.compare_to_model(a, S.model_flush(dim, edge, size, dtypes, steps, background=[1, 2, 3]), dtypes)


@pytest.mark.parametrize("order", ["chunks_only", "chunks_first"])
@pytest.mark.parametrize("seed", range(6))
def test_flush_chunk_record_programs(G, oracle, seed, order):
    """Whole-chunk records flushed as chunks (chunks_only) and turned into point records
    when points follow them (chunks_first): same chunk ids, masks and existing values as
    the oracle (a slot that is PADDING without EXIST has no specified value in the
    reference, BlockMapGpu_kernels.cuh:361-367), and every mask-0 slot holds the
    background."""
    rng = np.random.default_rng(400 + seed)
    dim = 3 if seed % 2 else 2
    edge = [2, 4, 8][seed % 3]
    size = [int(rng.integers(edge, 6 * edge)) for _ in range(dim)]
    dtypes = [[np.float64], [np.int32, np.uint32], [np.float32, np.int64, np.uint8]][seed % 3]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=3, nrec=120, order=order)
    bgs = [3, 1, 2][: len(dtypes)]
    a = S.run_flush_program(G, dim, edge, size, dtypes, steps, background=bgs)
    b = S.run_flush_program(oracle.SparseGridCpu, dim, edge, size, dtypes, steps, background=bgs)
    S.assert_same_state(a, b)
    S.compare_to_model(a, S.model_flush(dim, edge, size, dtypes, steps, background=bgs), dtypes)
    empty = a.masks() == 0
    for p, dt in enumerate(dtypes):
        assert np.all(a.prop(p)[empty] == np.array(bgs[p], dt))


def _device_values(vals, dtypes):
    """Host value arrays -> device tensors (unsigned types as their signed views)."""
    import torch
    out = []
    for v, dt in zip(vals, dtypes):
        v = np.ascontiguousarray(np.asarray(v, dtype=dt).reshape(-1))
        if v.dtype.kind == "u" and v.dtype.itemsize > 1:
            v = v.view(np.dtype("i%d" % v.dtype.itemsize))
        out.append(torch.from_numpy(v).cuda())
    return out


@pytest.mark.parametrize("seed", range(8))
def test_staged_device_coordinate_inserts(G, oracle, seed):
    """Device-coordinate inserts are staged straight as the grouped flush's (chunk id,
    record) pairs while nothing else is pending.  Batches of them, mixed in any order with
    key inserts and chunk inserts (which move the staged records to the insert buffer
    first), the flush by point key chosen after staging, and batches holding coordinates
    outside the grid (the flush reports and drops those, the others stay pending and merge
    at the next flush), against the oracle fed the same records as keys."""
    import torch
    rng = np.random.default_rng(4000 + seed)
    dim, edge = (3, 8) if seed < 6 else (2, 16)
    # whole chunks: every point key of a chunk is then a coordinate inside the grid
    size = [edge * int(rng.integers(3, 9)) for _ in range(dim)]
    dtypes = [[np.float32, np.int32, np.uint32], [np.float64, np.int32, np.uint8], [np.int16],
              [np.uint8, np.float64], [np.float32], [np.int64, np.uint16, np.int8],
              [np.float32, np.float32], [np.uint32, np.uint32, np.uint32]][seed]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=3, nrec=3000,
                                   with_chunks=seed % 2 == 0, order=["points_first", "chunks_first"][seed % 4 // 2])
    a = G(dim, size, dtypes, block_edge=edge)
    b = oracle.SparseGridCpu(dim, size, dtypes, block_edge=edge)
    for g in (a, b):
        for p in range(len(dtypes)):
            g.set_background_value(p, p + 2)
    for fi, (recs, ops) in enumerate(steps):
        for kind, keys, m, v in recs:
            if kind != "points":
                a.insert_block(keys, m, v)
                b.insert_block(keys, m, v)
                continue
            if rng.random() < 0.75:
                # in two device batches, sometimes without masks (every record EXIST) or
                # without values (the background)
                nomask, noval = rng.random() < 0.2, rng.random() < 0.2
                if nomask:
                    m = np.ones_like(m)
                if noval:
                    v = [np.full(len(keys), p + 2, dt) for p, dt in enumerate(dtypes)]
                b.insert(keys, v, m)
                cut = int(rng.integers(0, len(keys) + 1))
                for lo, hi in ((0, cut), (cut, len(keys))):
                    if hi == lo:
                        continue
                    dc = torch.from_numpy(a.inv_lin_id(keys[lo:hi]).astype(np.int32)).cuda()
                    dm = None if nomask else torch.from_numpy(np.ascontiguousarray(m[lo:hi])).cuda()
                    dv = None if noval else _device_values([x[lo:hi] for x in v], dtypes)
                    a.insert_coords(dc, dv, dm)
            else:
                b.insert(keys, v, m)
                a.insert(keys, v, m)
        if fi == 1:
            # a staged batch with coordinates outside the grid: the flush drops them, the
            # rest stays pending in order
            n = 500
            coords = np.stack([rng.integers(0, s, n) for s in size], axis=1).astype(np.int32)
            bad = rng.choice(n, 7, replace=False)
            coords[bad[:4], 0] = size[0] + rng.integers(0, 5, 4)
            coords[bad[4:], dim - 1] = -1 - rng.integers(0, 3, 3)
            vals = [S._rand_vals(rng, dt, n) for dt in dtypes]
            masks = rng.choice(np.array([1, 1, 3, 0, 2], np.uint8), n)
            good = np.setdiff1d(np.arange(n), bad)
            pending = a.pending_inserts()
            a.insert_coords(torch.from_numpy(coords).cuda(), _device_values(vals, dtypes),
                            torch.from_numpy(masks).cuda())
            assert a.pending_inserts() == pending + n
            with pytest.raises(C.SgError):
                a.flush(ops)
            assert a.pending_inserts() == pending + len(good)
            b.insert(b.lin_id(coords[good]), [x[good] for x in vals], masks[good])
        if fi == 2 and seed % 3 == 1:
            a.set_tuning(C.TUNE_FLUSH_PATH, 1)  # the flush by point key takes the staged records
        a.flush(ops)
        b.flush(ops)
        assert a.pending_inserts() == 0
        # existing values equal (a PADDING slot without EXIST has no specified value), and
        # every mask-0 slot holds the background
        S.assert_same_state(a, b)
        empty = a.masks() == 0
        for p, dt in enumerate(dtypes):
            assert np.all(a.prop(p)[empty] == np.array(p + 2, dt))
    # staged records pending: they survive sg_trim, block a change of linearizer, and a
    # reset drops them
    e = G(dim, size, dtypes, block_edge=edge)
    e.insert_coords(torch.zeros((3, dim), dtype=torch.int32).cuda(), None)
    C.check(e._lib.sg_trim(e._h))
    assert e.pending_inserts() == 3
    assert e._lib.sg_set_linearizer(e._h, 1) != C.SG_OK
    e.remove_add_unpack_reset()
    assert e.pending_inserts() == 0
    e.flush(["none"] * len(dtypes))
    assert e.num_chunks() == 0


def test_flush_empty_and_repeat(G, oracle):
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (64, 64, 64), [np.int32], block_edge=8)
        g.flush(["smax"])  # nothing pending
        assert g.num_chunks() == 0
        assert g.get(0, [5])[0] == 0
        g.insert([5, 5, 5], [np.array([3, 9, 4], np.int32)])
        g.flush(["smax"])
        assert g.get(0, [5])[0] == 9
        g.insert([5], [np.array([2], np.int32)])
        g.flush(["sadd"])
        assert g.get(0, [5])[0] == 11


@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
@pytest.mark.parametrize("dim,edge", [(2, 8), (3, 4), (3, 8)])
def test_neighbour_tables(G, oracle, nn, dim, edge):
    rng = np.random.default_rng(dim * 10 + edge)
    size = [edge * 9] * dim
    nb = 9 ** dim
    chunks = rng.choice(nb, nb // 3, replace=False).astype(np.int64)
    out = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(dim, size, [np.float32], block_edge=edge)
        g.insert(chunks * edge ** dim, [np.ones(chunks.size, np.float32)])
        g.flush(["smax"])
        out.append(g.neighbours(nn))
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("lin", ["smb", "zmb"])
@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
def test_neighbour_tables_sparse_domain(G, oracle, nn, lin):
    """A few clustered chunks in a domain of 2^23 chunk positions: far more ids than 64 x
    the chunk count, where the neighbour search runs on the sorted ids instead of a table
    of positions by id (row-major and Morton ids), against the oracle."""
    rng = np.random.default_rng(77)
    size = [2048, 2048, 1024]
    centre = rng.integers(2, [254, 254, 126], size=(40, 3))
    coords = np.concatenate([centre + rng.integers(-1, 2, size=centre.shape) for _ in range(6)])
    coords = np.unique(coords, axis=0) * 8
    out = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, size, [np.float32], block_edge=8, linearizer=lin)
        g.insert_coords(coords, [np.ones(len(coords), np.float32)])
        g.flush(["smax"])
        out.append(g.neighbours(nn))
    assert len(out[0]) > 0
    np.testing.assert_array_equal(out[0], out[1])


def test_count_existing_through_planes_and_masks(G, oracle):
    """countExistingElements on 8^3 chunks reads the EXIST bit-planes while the flush
    keeps them current, the masks otherwise (after a removal, after the mask array was
    handed out); countBoundaryElements reads the masks: every count equals the oracle's."""
    rng = np.random.default_rng(91)
    size = (72, 64, 56)
    steps = S.random_flush_program(91, 3, 8, list(size), [np.int32], nflush=2, nrec=4000)
    grids = [S.run_flush_program(cls, 3, 8, list(size), [np.int32], steps) for cls in (G, oracle.SparseGridCpu)]

    def same():
        for f in ("count_existing_elements", "count_boundary_elements"):
            assert getattr(grids[0], f)() == getattr(grids[1], f)(), f

    same()
    for g in grids:
        g.tag_boundaries(C.NN_STAR)
    same()
    keys = rng.choice(grids[1].chunk_ids(), 20) * 512 + rng.integers(0, 512, 20)
    for g in grids:
        g.remove_points(keys)
    same()
    grids[0].device_ptr(-1)
    same()
    for g in grids:
        g.insert(keys, [np.arange(20, dtype=np.int32)])
        g.flush(["smax"])
    same()


def _shell(cls, n, lo, hi, props, edge=8, vals=None):
    pts = W.shell_points(n, lo, hi)
    g = cls(3, (n, n, n), props, block_edge=edge)
    g.insert_coords(pts, vals(pts) if vals else [np.ones(len(pts), p) for p in props])
    g.flush(["smax"] * len(props))
    return g


@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
@pytest.mark.parametrize("edge", [4, 8])
def test_tag_boundaries_shell(G, oracle, nn, edge):
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = _shell(cls, 64, 12.0, 19.0, [np.float32], edge=edge)
        g.set_nn_type(nn)
        g.tag_boundaries(nn)
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])
    assert gs[0].count_boundary_elements() == gs[1].count_boundary_elements() > 0
    assert gs[0].count_existing_elements() == gs[1].count_existing_elements()


def test_shell64_fp64_lap_x10_then_tag(G, oracle):
    """64^3 shell r in [20,24), FP64, seed 0 values, smax flush, 10 x 7-point Laplacian
    (-6, 1; absent neighbours 0) in place, then tagBoundaries: chunk map, masks bit-exact,
    field within 1e-12 relative (bit-exact here)."""
    w = W.shell64_fp64()
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, w["size"], [np.float64], block_edge=8)
        g.insert_coords(w["coords"], [w["values"]])
        g.flush(["smax"])
        for _ in range(10):
            g.conv_lap(0, 0)
        g.tag_boundaries()
        gs.append(g)
    assert gs[0].num_chunks() == 146 and gs[0].count_existing_elements() == 24376
    S.assert_same_state(gs[0], gs[1])
    ex = gs[0].masks() & 1 == 1
    a, b = gs[0].prop(0)[ex], gs[1].prop(0)[ex]
    assert np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("edge", [4, 8])
def test_gray_scott_steps(G, oracle, dtype, edge):
    rng = np.random.default_rng(21)
    pts = W.shell_points(48, 10.0, 17.0)
    u0 = rng.uniform(0.3, 1.0, len(pts)).astype(dtype)
    v0 = rng.uniform(0.0, 0.4, len(pts)).astype(dtype)
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (48, 48, 48), [dtype] * 4, block_edge=edge)
        g.insert_coords(pts, [u0, v0, np.zeros_like(u0), np.zeros_like(u0)])
        g.flush(["smax"] * 4)
        for _ in range(5):
            g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
            g.conv2_gray_scott(2, 3, 0, 1, **W.GS_PARAMS)
        g.conv2_gray_scott(0, 1, 0, 1, **W.GS_PARAMS)  # in place: Jacobi semantics
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])


@pytest.mark.parametrize("edge", [4, 8])
def test_heat_and_lap_with_box(G, oracle, edge):
    rng = np.random.default_rng(31)
    pts = W.shell_points(40, 8.0, 15.0)
    vals = [rng.uniform(-1, 1, len(pts)).astype(np.float32), np.zeros(len(pts), np.float32)]
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (40, 40, 40), [np.float32, np.float32], block_edge=edge)
        g.set_background_value(0, 0.25)
        g.insert_coords(pts, vals)
        g.flush(["smax", "smax"])
        g.tag_boundaries()
        g.conv_heat(0, 1, 0.1, box=((5, 9, 3), (30, 27, 33)))
        g.conv_lap(1, 0, -5.0, 0.5, box=((0, 0, 0), (39, 20, 39)))
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])


@pytest.mark.parametrize("edge", [4, 8])
def test_box_sum_uint32(G, oracle, edge):
    rng = np.random.default_rng(41)
    pts = W.shell_points(40, 6.0, 16.0)
    vals = [rng.integers(0, 2 ** 32, len(pts), dtype=np.uint64).astype(np.uint32), np.zeros(len(pts), np.uint32)]
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (40, 40, 40), [np.uint32, np.uint32], block_edge=edge)
        g.insert_coords(pts, vals)
        g.flush(["sadd", "sadd"])
        g.set_nn_type(C.NN_FULL)
        g.tag_boundaries(C.NN_FULL)
        g.conv_box(0, 1, np.ones(27, np.int64))
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])


def test_rbsor_and_residual(G, oracle):
    rng = np.random.default_rng(51)
    pts = W.shell_points(32, 5.0, 12.0)
    f = rng.uniform(-1, 1, len(pts))
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (32, 32, 32), [np.float64] * 3, block_edge=8)
        g.insert_coords(pts, [np.zeros(len(pts)), f, np.zeros(len(pts))])
        g.flush(["smax"] * 3)
        for _ in range(4):
            g.rbsor_sweep(0, 1, 1.5, 0)
            g.rbsor_sweep(0, 1, 1.5, 1)
        g.residual(0, 1, 2)
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])


def test_get_points(G, oracle):
    rng = np.random.default_rng(61)
    pts = W.shell_points(32, 5.0, 12.0)
    vals = [rng.uniform(-1, 1, len(pts))]
    keys = rng.integers(0, 32 ** 3, 5000)
    res = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (32, 32, 32), [np.float64], block_edge=8)
        g.set_background_value(0, -7.5)
        g.insert_coords(pts, vals)
        g.flush(["smax"])
        res.append(g.get(0, keys))
    np.testing.assert_array_equal(res[0], res[1])
    assert (res[0] == -7.5).sum() > 100


def test_blobs_flush_subsample(G, oracle):
    """2^20 of the 1024^3 Gaussian-blob inserts (30% repeats): int32 smax + uint32 sadd,
    then NNFull tagBoundaries and the 27-point uint32 box sum, bit-exact."""
    w = W.blobs1024(n_insert=1 << 20)
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, w["size"], [np.int32, np.uint32, np.uint32], block_edge=8)
        g.insert_coords(w["coords"], [w["int32"], w["uint32"], np.zeros(len(w["uint32"]), np.uint32)])
        g.flush(["smax", "sadd", "none"])
        g.set_nn_type(C.NN_FULL)
        g.tag_boundaries(C.NN_FULL)
        g.conv_box(1, 2, np.ones(27, np.int64))
        gs.append(g)
    S.assert_same_state(gs[0], gs[1], props=[0, 1, 2])


@pytest.mark.parametrize("variant,stores,lskip", [(0, 0, 0), (0, 1, 1), (1, 2, 0), (2, 1, 0), (3, 2, 1),
                                                  (4, 2, 0), (4, 0, 1), (5, 1, 0), (4, 3, 0), (3, 3, 1),
                                                  (0, 3, 0), (1, 3, 1), (6, 3, 0), (7, 3, 1), (8, 3, 0),
                                                  (6, 2, 1), (7, 0, 0), (8, 1, 1), (3, 3, 2), (3, 3, 3),
                                                  (4, 3, 2), (6, 3, 2), (7, 3, 3), (0, 3, 2), (1, 3, 2),
                                                  (2, 0, 3), (5, 1, 2), (8, 2, 3), (3, 3, 4), (6, 3, 4),
                                                  (4, 3, 4), (0, 3, 4), (1, 3, 4), (7, 2, 4), (5, 1, 4)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fast_stencil_variants(G, oracle, variant, stores, lskip, dtype):
    """Every selectable launch shape of the 8^3 star-stencil kernel gives the oracle's
    result (Gray-Scott ping-pong, heat with a box, Laplacian in place after tagBoundaries)."""
    rng = np.random.default_rng(71)
    pts = W.shell_points(56, 12.0, 22.0)
    u0 = rng.uniform(0.3, 1.0, len(pts)).astype(dtype)
    v0 = rng.uniform(0.0, 0.4, len(pts)).astype(dtype)
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (56, 56, 56), [dtype] * 4, block_edge=8)
        g.set_background_value(2, 0.5)
        if cls is G:
            g.set_tuning(C.TUNE_STENCIL_VARIANT, variant)
            g.set_tuning(C.TUNE_STENCIL_STORES, stores)
            g.set_tuning(C.TUNE_STENCIL_SKIP_EMPTY_ROWS, lskip)
            g.set_tuning(C.TUNE_STENCIL_WG_PER_XCD, 3)
        g.insert_coords(pts, [u0, v0, np.zeros_like(u0), np.zeros_like(u0)])
        g.flush(["smax"] * 4)
        for _ in range(3):
            g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
            g.conv2_gray_scott(2, 3, 0, 1, **W.GS_PARAMS)
        g.conv_heat(0, 2, 0.05, box=((3, 10, 0), (50, 44, 40)))
        g.tag_boundaries()
        g.conv_lap(2, 2)
        g.conv_lap(1, 3, -4.0, 0.5)
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])


def test_grouped_stencil_variants_agree_on_a_large_shell(G, oracle):
    """Groups of 2 / 4 / 8 chunks (two LDS tiles per wave or one) with every store mode,
    on a shell of several thousand chunks (many workgroups in flight, so that a missing
    wait between the waves of a group
    shows), give bitwise the same Gray-Scott steps as each other and as the oracle."""
    rng = np.random.default_rng(73)
    pts = W.shell_points(176, 60.0, 84.0)
    u0 = rng.uniform(0.3, 1.0, len(pts)).astype(np.float32)
    v0 = rng.uniform(0.0, 0.4, len(pts)).astype(np.float32)

    def run(cls, variant=None, stores=None, lskip=0):
        g = cls(3, (176, 176, 176), [np.float32] * 4, block_edge=8)
        if variant is not None:
            g.set_tuning(C.TUNE_STENCIL_VARIANT, variant)
            g.set_tuning(C.TUNE_STENCIL_STORES, stores)
            g.set_tuning(C.TUNE_STENCIL_SKIP_EMPTY_ROWS, lskip)
        g.insert_coords(pts, [u0, v0, np.zeros_like(u0), np.zeros_like(u0)])
        g.flush(["smax"] * 4)
        for _ in range(3):
            g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
            g.conv2_gray_scott(2, 3, 0, 1, **W.GS_PARAMS)
        return g

    ref = run(oracle.SparseGridCpu)
    for variant in (3, 4, 5, 6, 7, 8):
        for stores in (1, 2, 3):
            g = run(G, variant, stores)
            S.assert_same_state(g, ref)
            g.close()
    for variant in (3, 4, 6):
        for lskip in (2, 3, 4):
            g = run(G, variant, 3, lskip)
            S.assert_same_state(g, ref)
            g.close()
    ref.close()


@pytest.mark.parametrize("lskip", [2, 3, 4])
def test_occupancy_table_follows_mask_changes(G, oracle, lskip):
    """The row-occupancy table of the skipping loads is rebuilt after every change of the
    masks: points removed (whole rows emptied) and points added by a flush; each following
    Gray-Scott step equals the oracle's, with non-zero backgrounds in the skipped rows."""
    rng = np.random.default_rng(77)
    pts = W.shell_points(64, 14.0, 26.0)
    u0 = rng.uniform(0.3, 1.0, len(pts)).astype(np.float32)
    v0 = rng.uniform(0.0, 0.4, len(pts)).astype(np.float32)
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (64, 64, 64), [np.float32] * 4, block_edge=8)
        g.set_background_value(0, 0.25)
        g.set_background_value(1, -0.5)
        if cls is G:
            g.set_tuning(C.TUNE_STENCIL_VARIANT, 3)
            g.set_tuning(C.TUNE_STENCIL_STORES, 3)
            g.set_tuning(C.TUNE_STENCIL_SKIP_EMPTY_ROWS, lskip)
        g.insert_coords(pts, [u0, v0, np.zeros_like(u0), np.zeros_like(u0)])
        g.flush(["smax"] * 4)
        g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
        g.conv2_gray_scott(2, 3, 0, 1, **W.GS_PARAMS)
        # empty whole x-rows: every point with y % 3 == 0 and z % 2 == 0
        sel = pts[(pts[:, 1] % 3 == 0) & (pts[:, 2] % 2 == 0)]
        g.remove_points(g.lin_id(sel))
        g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
        g.conv2_gray_scott(2, 3, 0, 1, **W.GS_PARAMS)
        # refill some of them and add a new chunk
        extra = np.concatenate([sel[::2], np.array([[2, 2, 2], [3, 2, 2]], np.int32)])
        g.insert_coords(extra, [np.full(len(extra), 0.75, np.float32), np.full(len(extra), 0.125, np.float32),
                                np.zeros(len(extra), np.float32), np.zeros(len(extra), np.float32)])
        g.flush(["sright"] * 4)
        g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
        g.conv2_gray_scott(2, 3, 0, 1, **W.GS_PARAMS)
        gs.append(g)
    S.assert_same_state(gs[0], gs[1])


def test_blobs1024_full_flush_tag_box(G, oracle):
    """The full insertion workload: 2^24 inserts into 1024^3 from 64 Gaussian blobs (30%
    repeats), int32 smax + uint32 sadd, NNFull tagBoundaries, 27-point uint32 box sum:
    chunk map, masks and every property bit-exact against the oracle."""
    w = W.blobs1024()
    gs = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, w["size"], [np.int32, np.uint32, np.uint32], block_edge=8)
        g.insert_coords(w["coords"], [w["int32"], w["uint32"], np.zeros(len(w["uint32"]), np.uint32)])
        g.flush(["smax", "sadd", "none"])
        g.set_nn_type(C.NN_FULL)
        g.tag_boundaries(C.NN_FULL)
        g.conv_box(1, 2, np.ones(27, np.int64))
        gs.append(g)
    assert gs[0].num_chunks() == gs[1].num_chunks() == 317793
    S.assert_same_state(gs[0], gs[1])


def _no_shared_descs(rng, dim, size, kinds):
    out = []
    for kind in kinds:
        lo = [int(rng.integers(0, s // 2)) for s in size]
        hi = [int(rng.integers(l, s)) for l, s in zip(lo, size)]
        box = (lo, hi) if rng.random() < 0.5 else None
        apply = int(rng.choice([C.APPLY_INTERIOR, C.APPLY_EXISTING]))
        if kind == C.ST_LAP:
            out.append(C.make_desc(kind, [0], [1], apply, box, [-6.0, 1.0]))
        elif kind == C.ST_HEAT:
            out.append(C.make_desc(kind, [1], [1], apply, box, [0.125]))  # in place (Jacobi)
        elif kind == C.ST_GRAY_SCOTT:
            out.append(C.make_desc(kind, [0, 1], [2, 1], apply, box, [0.2, 0.1, 0.04, 0.06, 1.0]))
        elif kind == C.ST_RBSOR:
            out.append(C.make_desc(kind, [2, 0], [], apply, box, [1.3], [int(rng.integers(0, 2))]))
        elif kind == C.ST_RESIDUAL:
            out.append(C.make_desc(kind, [2, 0], [1], apply, box))
        else:
            out.append(C.make_desc(kind, [0], [1], apply, box, (), [int(v) for v in rng.integers(-3, 4, 27)]))
    return out


@pytest.mark.parametrize("seed", range(8))
def test_no_shared_mode(G, oracle, seed):
    """STENCIL_MODE_INPLACE_NO_SHARED (one thread per existing point, no LDS tile): the
    oracle's values in that mode and the LDS path's on every existing point, and no write
    to any other slot.  Random 2D/3D grids with padding and
    padding-only points, boxes, both apply modes, in-place outputs."""
    rng = np.random.default_rng(700 + seed)
    dim = 3 if seed % 2 else 2
    edge = [2, 4, 8][seed % 3]
    size = [int(rng.integers(2 * edge, 6 * edge)) for _ in range(dim)]
    integer = seed % 4 == 3
    dtypes = [np.int32] * 3 if integer else [[np.float64] * 3, [np.float32] * 3][seed % 2]
    kinds = [C.ST_BOX] if integer else [C.ST_LAP, C.ST_HEAT, C.ST_GRAY_SCOTT, C.ST_RBSOR, C.ST_RESIDUAL]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=2, nrec=250,
                                   ops=["smax"] * len(dtypes))
    descs = _no_shared_descs(rng, dim, size, kinds)
    grids = []
    for cls in (G, oracle.SparseGridCpu, G):
        g = S.run_flush_program(cls, dim, edge, size, dtypes, steps, background=[1, 2, 3])
        g.find_neighbours(C.NN_STAR)
        g.tag_boundaries(C.NN_STAR, C.CALCULATE_EXISTING_POINTS)
        if integer:
            # the 27-point box reads edge and corner ghosts: full ghost mapping, as the
            # reference's Conv3x3x3 tests set it (SparseGridGpu_tests.cu:887-952)
            g.find_neighbours(C.NN_FULL)
            g.set_nn_type(C.NN_FULL)
        grids.append(g)
    a, b, lds = grids
    # PADDING-only slots have no specified value after a flush (the two sides differ
    # there) and the stencils read them as neighbours: start from identical arrays
    for p in range(len(dtypes)):
        b.set_prop(p, a.prop(p))
    # the point list: EXIST points in chunk order, increasing offset
    n = ctypes.c_int64()
    C.check(a._lib.sg_existing_points(a._h, ctypes.byref(n)))
    ep = np.zeros(max(n.value, 1), np.int64)
    C.check(a._lib.sg_get_existing_points(a._h, ep.ctypes.data, C.SG_MEM_HOST))
    c, o = np.nonzero(a.masks() & 1)
    np.testing.assert_array_equal(ep[:n.value], c * a.chunk_size + o)
    for d in descs:
        # the LDS path starts from the same state (it also writes padding-only points
        # under APPLY_EXISTING, which later stencils read)
        before = [a.prop(p) for p in range(len(dtypes))]
        for p in range(len(dtypes)):
            lds.set_prop(p, before[p])
        d.mode = C.STENCIL_MODE_INPLACE_NO_SHARED
        a.apply_stencil(d)
        b.apply_stencil(d)
        d.mode = C.STENCIL_MODE_INPLACE
        lds.apply_stencil(d)
        S.assert_same_state(a, b)
        S.assert_same_state(a, lds)
        # the point list holds only EXIST points: nothing else is written (a PADDING-only
        # slot has no specified value after a flush, so this is checked on the GPU alone)
        other = (a.masks() & 1) == 0
        for p in range(len(dtypes)):
            np.testing.assert_array_equal(a.prop(p)[other], before[p][other])


@pytest.mark.parametrize("dim,edge", [(2, 4), (3, 2), (3, 8)])
def test_zmb_grid_parity(G, oracle, dim, edge):
    """SparseGridGpu_z (Morton chunk ids): flush, both neighbour tables, tagBoundaries,
    star stencils (the 8^3 fast path included), 27-point box, pack/unpack, against the
    oracle's Morton grid."""
    rng = np.random.default_rng(80 + dim * edge)
    size = [int(rng.integers(4 * edge, 9 * edge)) for _ in range(dim)]
    n = 3000
    coords = np.stack([rng.integers(0, s, n) for s in size], axis=1).astype(np.int32)
    vf = rng.integers(-20, 20, n).astype(np.float32)
    vi = rng.integers(0, 1000, n).astype(np.int32)
    grids = []
    for cls in (G, oracle.SparseGridCpu):
        g = cls(dim, size, [np.float32, np.float32, np.float32, np.int32, np.int32], block_edge=edge,
                linearizer="zmb")
        g.insert_coords(coords, [vf, vf, vf, vi, vi])
        g.flush(["sadd", "smax", "smin", "sadd", "smax"])
        grids.append(g)
    a, b = grids
    S.assert_same_state(a, b)
    np.testing.assert_array_equal(a.lin_id(coords), b.lin_id(coords))
    for nn in (C.NN_STAR, C.NN_FULL):
        for g in grids:
            g.find_neighbours(nn)
        np.testing.assert_array_equal(a.neighbours(nn), b.neighbours(nn))
    for g in grids:
        g.tag_boundaries(C.NN_STAR)
        g.conv_lap(0, 1)
        # a box restriction makes the 8^3 fast path decode chunk coordinates from ids
        g.conv_lap(1, 2, box=([s // 4 for s in size], [3 * s // 4 for s in size]))
        g.conv2_gray_scott(0, 1, 2, 0, Du=0.2, Dv=0.1, F=0.04, k=0.06, dt=1.0)
        g.set_nn_type(C.NN_FULL)  # the box reads edge and corner ghosts
        g.conv_box(3, 4, np.arange(27) % 5 - 2)
    S.assert_same_state(a, b)
    keys = a.lin_id(coords)
    for p in range(5):
        np.testing.assert_array_equal(a.get(p, keys), b.get(p, keys))
    boxes = [([0] * dim, [s // 2 for s in size]), ([s // 3 for s in size], [s - 1 for s in size])]
    bufs = [g.pack(boxes, [0, 3]) for g in grids]
    np.testing.assert_array_equal(bufs[0], bufs[1])
    for g, buf in zip(grids, bufs):
        g.remove_box(*boxes[0])
        g.unpack(buf, [1] * dim, [0, 3])
        g.remove_add_unpack_finalize([0, 3])
    S.assert_same_state(a, b)


@pytest.mark.parametrize("pairs", [1, 0])
@pytest.mark.parametrize("seed", range(6))
def test_tag_and_box_8cubed_random(G, oracle, seed, pairs):
    """The 8^3 tile kernels of tagBoundaries (bit-parallel, star and full) and of the
    27-point box (all scalar widths >= 4 bytes, box restriction, both apply modes; two
    chunks per wave sharing x ghosts, or one) on random 3D grids whose masks hold every
    bit pattern, against the oracle."""
    rng = np.random.default_rng(1000 + seed)
    size = [int(rng.integers(24, 60)) for _ in range(3)]
    dt = [np.float32, np.float64, np.int32, np.uint32, np.int64, np.uint64][seed]
    dtypes = [dt, dt]
    steps = S.random_flush_program(seed, 3, 8, size, dtypes, nflush=2, nrec=1500, ops=["smax", "smax"])
    nn = [C.NN_STAR, C.NN_FULL][seed % 2]
    lo = [int(rng.integers(0, s // 2)) for s in size]
    hi = [int(rng.integers(l, s)) for l, s in zip(lo, size)]
    integer = np.issubdtype(np.dtype(dt), np.integer)
    coef = rng.integers(-3, 4, 27)
    grids = []
    for cls in (G, oracle.SparseGridCpu):
        g = S.run_flush_program(cls, 3, 8, size, dtypes, steps, background=[1, 2])
        g.set_nn_type(C.NN_FULL)
        g.tag_boundaries(nn)
        grids.append(g)
    a, b = grids
    a.set_tuning(C.TUNE_BOX_PAIRS, pairs)
    S.assert_same_state(a, b)  # masks after tagging
    for p in range(2):
        b.set_prop(p, a.prop(p))  # PADDING-only slots have no specified value after a flush
    # random taps, and one coefficient for all taps (the integer kernel then sums the 27
    # values and multiplies once)
    for cf, (apply, box) in [(coef, (C.APPLY_EXISTING, None)), (coef, (C.APPLY_INTERIOR, (lo, hi))),
                             (coef, (C.APPLY_EXISTING, (lo, hi))), (np.full(27, -3), (C.APPLY_EXISTING, None))]:
        for g in grids:
            if integer:
                g.apply_stencil(C.make_desc(C.ST_BOX, [0], [1], apply, box, (), [int(v) for v in cf]))
            else:
                g.apply_stencil(C.make_desc(C.ST_BOX, [0], [1], apply, box, [float(v) for v in cf]))
        S.assert_same_state(a, b, exist_only=False)


@pytest.mark.parametrize("variant,cpw", [(0, 1), (1, 1), (1, 2), (1, 4)])
@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
def test_tag_boundaries_chunks_per_wave(G, oracle, variant, cpw, nn):
    """The 8^3 tagBoundaries kernels (neighbour EXIST bits from the bit-planes or from the
    mask rows, with 1, 2 or 4 chunks per wave; masks with every bit pattern) tag the same
    points as the oracle."""
    rng = np.random.default_rng(1100 + cpw)
    size = [int(rng.integers(60, 90)) for _ in range(3)]
    steps = S.random_flush_program(7 + cpw, 3, 8, size, [np.int32], nflush=2, nrec=4000, ops=["smax"])
    grids = []
    for cls in (G, oracle.SparseGridCpu):
        g = S.run_flush_program(cls, 3, 8, size, [np.int32], steps)
        if cls is G:
            g.set_tuning(C.TUNE_TAG_VARIANT, variant)
            g.set_tuning(C.TUNE_TAG_CHUNKS_PER_WAVE, cpw)
        g.set_nn_type(nn)
        g.tag_boundaries(nn)
        grids.append(g)
    S.assert_same_state(grids[0], grids[1])
    assert grids[0].count_boundary_elements() == grids[1].count_boundary_elements() > 0


@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
def test_tag_boundaries_after_mask_changes(G, oracle, nn):
    """The EXIST bit-planes follow every change of the masks: tagBoundaries after a flush
    that adds points to existing and new chunks, after removing points and a box, and
    after the masks were written through the device pointer, matches the oracle each
    time (tagging itself changes only PADDING bits and keeps the planes)."""
    import torch
    rng = np.random.default_rng(77)
    size = (72, 64, 80)
    steps = S.random_flush_program(31, 3, 8, size, [np.int32], nflush=1, nrec=3000, ops=["smax"])
    grids = [S.run_flush_program(cls, 3, 8, size, [np.int32], steps) for cls in (G, oracle.SparseGridCpu)]

    def tag_and_compare():
        for g in grids:
            g.set_nn_type(nn)
            g.tag_boundaries(nn)
        S.assert_same_state(grids[0], grids[1])

    tag_and_compare()
    tag_and_compare()  # again, on tagged masks
    pts = np.stack([rng.integers(0, s, 2000) for s in size], axis=1).astype(np.int32)
    vals = rng.integers(-100, 100, len(pts)).astype(np.int32)
    for g in grids:
        g.insert_coords(pts, [vals])
        g.flush(["smax"])
    tag_and_compare()
    # a flush into existing chunks only (no chunk added: planes rewritten in place), with
    # new points beside the existing ones
    ex = S.existing_points(grids[0])
    near = np.clip(ex[::7] + rng.integers(-1, 2, (len(ex[::7]), 3)), 0, np.array(size) - 1)
    same = (near // 8 == ex[::7] // 8).all(axis=1)
    near = near[same].astype(np.int32)
    nvals = rng.integers(-100, 100, len(near)).astype(np.int32)
    for g in grids:
        g.insert_coords(near, [nvals])
        g.flush(["smax"])
    tag_and_compare()
    for variant in (0, 1):
        grids[0].set_tuning(C.TUNE_TAG_VARIANT, variant)
        tag_and_compare()
    grids[0].set_tuning(C.TUNE_TAG_VARIANT, 0)
    keys = grids[0].lin_id(S.existing_points(grids[0])[::3])
    for g in grids:
        g.remove_points(keys)
        g.remove_box((8, 8, 8), (40, 30, 50))
        g.remove_add_unpack_finalize([])
    tag_and_compare()
    # masks written by the caller through the device pointer (EXIST cleared on every fifth
    # slot); the oracle gets the same masks as chunk records of a new grid
    m2 = grids[0].masks().copy()
    m2.reshape(-1)[::5] &= 0xFE
    ids = grids[0].chunk_ids()

    class _DeviceBytes:
        def __init__(self, ptr, n):
            self.__cuda_array_interface__ = {"shape": (n,), "typestr": "|u1", "data": (ptr, False),
                                             "version": 2}

    grids[0].synchronize()
    dev = torch.as_tensor(_DeviceBytes(grids[0].device_ptr(-1), m2.size), device="cuda")
    dev.copy_(torch.from_numpy(m2.reshape(-1)).cuda())
    torch.cuda.synchronize()
    o = oracle.SparseGridCpu(3, size, [np.int32], block_edge=8)
    o.insert_block(ids, m2, [np.zeros(m2.shape, np.int32)])
    o.flush(["smax"])
    for g in (grids[0], o):
        g.set_nn_type(nn)
        g.tag_boundaries(nn)
    np.testing.assert_array_equal(grids[0].chunk_ids(), o.chunk_ids())
    np.testing.assert_array_equal(grids[0].masks(), o.masks())
    for g in grids + [o]:
        g.close()


def test_new_grids_reuse_blocks_of_destroyed_ones(G, oracle):
    """Grids built one after another, each on a stream of its own and destroyed before
    the next (which then allocates from the blocks the library kept): flush, NNFull
    tagBoundaries and a Gray-Scott step match the oracle every time; sg_trim then frees
    the kept blocks."""
    import torch
    lib = C.load()
    for rep in range(3):
        rng = np.random.default_rng(300 + rep)
        size = [int(rng.integers(40, 70)) for _ in range(3)]
        dtypes = [np.float32, np.float32, np.float32, np.float32]
        steps = S.random_flush_program(30 + rep, 3, 8, size, dtypes, nflush=2, nrec=3000,
                                       ops=["smax", "smax", "smax", "smax"])
        grids = []
        for cls in (G, oracle.SparseGridCpu):
            g = cls(3, size, dtypes, block_edge=8)
            if cls is G:
                g.set_stream(torch.cuda.Stream())
            for recs, ops in steps:
                for kind, a, m, v in recs:
                    if kind == "points":
                        g.insert(a, v, m)
                    else:
                        g.insert_block(a, m, v)
                g.flush(ops)
            g.set_nn_type(C.NN_FULL)
            g.tag_boundaries(C.NN_FULL)
            grids.append(g)
        a, b = grids
        S.assert_same_state(a, b)
        for p in range(4):
            b.set_prop(p, a.prop(p))  # PADDING-only slots have no specified value after a flush
        for g in grids:
            g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
        S.assert_same_state(a, b)
        grids[0].close()
        grids[1].close()
    g = G(3, (16, 16, 16), [np.float32], block_edge=8)
    C.check(lib.sg_trim(g._h))
    g.close()


def test_to_kernel_view(G):
    """sg_to_kernel (toKernel): the device arrays of the grid, the full neighbour table it
    builds when stale, and the geometry."""
    pts = W.shell_points(40, 8.0, 14.0)
    g = G(3, (40, 40, 40), [np.float32, np.int32], block_edge=8)
    g.insert_coords(pts, [np.ones(len(pts), np.float32), np.ones(len(pts), np.int32)])
    g.flush(["smax", "smax"])
    v = C.KernelView()
    C.check(C.load().sg_to_kernel(g._h, C.NN_FULL, ctypes.byref(v)))
    assert v.nchunks == g.num_chunks() and v.nnn == 27 and v.dim == 3 and v.edge == 8
    assert v.chunk_size == 512 and v.nprops == 2 and v.zorder == 0
    assert list(v.size) == [40, 40, 40] and list(v.bsz) == [5, 5, 5]
    assert v.prop[0] == g.device_ptr(0) and v.prop[1] == g.device_ptr(1) and v.mask == g.device_ptr(-1)
    assert v.keys and v.nn
    g.close()


def test_flush_rejects_device_coords_outside_grid(G):
    """Coordinates handed over in device memory are checked by the flush: one outside
    the grid fails it (SgError), the grid is unchanged and the valid records stay pending;
    host coordinates fail at insertion."""
    import torch
    from openfpm_data_amd._capi import SgError
    g = G(3, (32, 32, 32), [np.float32], block_edge=8)
    g.insert_coords([[1, 2, 3]], [np.array([1.0], np.float32)])
    g.flush(["smax"])
    before = (g.chunk_ids().copy(), g.masks().copy(), g.prop(0).copy())
    coords = torch.tensor([[4, 5, 6], [3, 40, 2], [7, 7, 7]], dtype=torch.int32, device="cuda")
    g.insert_coords(coords, [torch.ones(3, dtype=torch.float32, device="cuda")])
    with pytest.raises(SgError):
        g.flush(["smax"])
    assert g.pending_inserts() == 2
    for a, b in zip(before, (g.chunk_ids(), g.masks(), g.prop(0))):
        np.testing.assert_array_equal(a, b)
    with pytest.raises(SgError):
        g.insert_coords([[3, 40, 2]], [np.array([1.0], np.float32)])
    g.flush(["smax"])
    assert g.pending_inserts() == 0 and g.count_existing_elements() == 3
    g.close()


def test_flush_rejects_device_keys_outside_grid(G):
    """Point keys handed over in device memory are checked by the flush: a key beyond the
    grid fails it (SgError) and is dropped, the grid is unchanged, the other records stay
    pending and the next flush merges them."""
    import torch
    from openfpm_data_amd._capi import SgError
    g = G(3, (32, 32, 32), [np.float32], block_edge=8)
    g.insert_coords([[1, 2, 3]], [np.array([1.0], np.float32)])
    g.flush(["smax"])
    before = (g.chunk_ids().copy(), g.masks().copy(), g.prop(0).copy())
    keys = torch.tensor([5, 32 * 32 * 32 + 7, 9], dtype=torch.int64, device="cuda")
    vals = torch.ones(3, dtype=torch.float32, device="cuda")
    g.insert(keys, [vals])
    with pytest.raises(SgError):
        g.flush(["smax"])
    assert g.pending_inserts() == 2
    for a, b in zip(before, (g.chunk_ids(), g.masks(), g.prop(0))):
        np.testing.assert_array_equal(a, b)
    g.flush(["smax"])
    assert g.count_existing_elements() == 3
    g.close()


def test_operations_on_an_empty_grid(G, oracle):
    """Every operation on a grid without chunks: nothing happens, nothing fails, and the
    results equal the oracle's (counts 0, empty packs, empty links, no copies)."""
    for cls in (G, oracle.SparseGridCpu):
        g = cls(3, (32, 32, 32), [np.float32, np.float32, np.float32, np.float32], block_edge=8)
        h = cls(3, (64, 64, 64), [np.float32, np.float32, np.float32, np.float32], block_edge=8)
        g.flush(["smax"] * 4)
        g.find_neighbours()
        g.tag_boundaries()
        g.conv2_gray_scott(0, 1, 2, 3, **W.GS_PARAMS)
        g.conv_lap(0, 1)
        assert g.num_chunks() == 0 and g.count_existing_elements() == 0
        assert g.pack_size([((0, 0, 0), (31, 31, 31))], [0]) == g.pack([((0, 0, 0), (31, 31, 31))], [0]).size
        g.remove_box((0, 0, 0), (31, 31, 31))
        g.remove_add_unpack_finalize([0])
        g.construct_link(h, 0, ((0, 0, 0), (31, 31, 31)), (0, 0, 0))
        pc, po, sc, lc, lo = g.links(0)
        assert len(pc) == 0 and len(lc) == 0 and list(sc) == [0]
        h.copy_to(g, ((0, 0, 0), (31, 31, 31)), ((0, 0, 0), (31, 31, 31)))
        for phase in range(3):
            g.remove_copy_to_finalize([0], phase)
            h.remove_copy_to_finalize([0], phase)
        assert h.num_chunks() == 0
        g.close()
        h.close()


def test_device_inputs_are_checked(G):
    """Device tensors handed to the inserts are checked for element type, contiguity and
    element count before they reach the C ABI (a wrong one raises, nothing is inserted)."""
    import torch
    g = G(3, (32, 32, 32), [np.float32, np.int32], block_edge=8)
    c = torch.tensor([[1, 2, 3], [4, 5, 6]], dtype=torch.int32, device="cuda")
    v = [torch.ones(2, device="cuda"), torch.ones(2, dtype=torch.int32, device="cuda")]
    bad = [
        (c.to(torch.int64), v, None),                      # int64 coordinates
        (c.t().contiguous().t(), v, None),                 # strided coordinates
        (c, [v[0].double(), v[1]], None),                  # float64 values for a float32 property
        (c, [v[0], v[1].float()], None),                   # float values for an int32 property
        (c, [torch.ones(3, device="cuda"), v[1]], None),   # wrong count
        (c, v, torch.ones(2, dtype=torch.int32, device="cuda")),  # int32 masks
    ]
    for coords, vals, masks in bad:
        with pytest.raises(ValueError):
            g.insert_coords(coords, vals, masks=masks)
    assert g.pending_inserts() == 0
    g.insert_coords(c, v)
    g.flush(["smax", "smax"])
    assert g.count_existing_elements() == 2
    g.close()


def test_device_coordinates_outside_the_grid(G, oracle):
    """Device coordinates outside the grid are found by the next flush: it fails, drops only
    those records and keeps the others pending; the next flush merges them as the oracle
    merges the valid records alone."""
    import torch
    from openfpm_data_amd._capi import SgError
    rng = np.random.default_rng(81)
    good = rng.integers(0, 40, size=(500, 3)).astype(np.int32)
    bad = np.array([[40, 0, 0], [-1, 3, 3], [5, 99, 5]], np.int32)
    coords = np.concatenate([good[:200], bad[:2], good[200:], bad[2:]])
    vals = rng.uniform(-1, 1, len(coords)).astype(np.float32)
    g = G(3, (40, 40, 40), [np.float32], block_edge=8)
    g.insert_coords(torch.from_numpy(coords).cuda(), [torch.from_numpy(vals).cuda()])
    with pytest.raises(SgError, match="3 inserted points outside the grid"):
        g.flush(["sright"])
    assert g.pending_inserts() == 500 and g.num_chunks() == 0
    g.flush(["sright"])
    keep = np.all((coords >= 0) & (coords < 40), axis=1)
    o = oracle.SparseGridCpu(3, (40, 40, 40), [np.float32], block_edge=8)
    o.insert_coords(coords[keep], [vals[keep]])
    o.flush(["sright"])
    S.assert_same_state(g, o)
    g.close()
    o.close()


def test_flush_known_answers_of_the_reference(G):
    """BlockMapGpu_tests.cu:126-264 and BlockMapGpu_kernels_tests.cu:16-492 (segmented fold
    with old data, merge maps, conflict resolution), end to end through insert + flush on
    the GPU, against the transcribed answers."""
    S.check_flush_known_answers(G)
