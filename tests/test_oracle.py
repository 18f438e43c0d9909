"""CPU tests that pin the oracle (oracle/sgcpu.cpp) before the GPU tests compare against it:
the reference's own known answers, brute-force models and closed forms."""
import numpy as np
import pytest

from openfpm_data_amd import _capi as C
from tests import scenarios as S


def test_linearizer_known_answers(oracle, golden):
    case = golden["grid_smb_linid"]
    g = oracle.SparseGridCpu(case["dim"], case["size"], [np.float32], block_edge=case["block_edge"])
    for crd, lin in case["cases"]:
        assert g.lin_id([crd])[0] == lin
        assert list(g.inv_lin_id([lin])[0]) == crd


def test_linearizer_closed_form_3d(oracle):
    g = oracle.SparseGridCpu(3, (40, 24, 17), [np.float32], block_edge=8)
    rng = np.random.default_rng(5)
    c = np.stack([rng.integers(0, s, 300) for s in (40, 24, 17)], axis=1)
    bs = (5, 3, 3)
    exp = (((c[:, 2] // 8) * bs[1] + c[:, 1] // 8) * bs[0] + c[:, 0] // 8) * 512 + \
        ((c[:, 2] % 8) * 8 + c[:, 1] % 8) * 8 + c[:, 0] % 8
    np.testing.assert_array_equal(g.lin_id(c), exp)
    np.testing.assert_array_equal(g.inv_lin_id(exp), c)


@pytest.mark.parametrize("name", ["insert_2d", "insert_3d"])
def test_insert_known_answer(oracle, golden, name):
    case = golden[name]
    dim, size = case["dim"], case["size"]
    g = oracle.SparseGridCpu(dim, size, [np.float32], block_edge=case["block_edge"])
    g.set_background_value(0, case["background"])
    grids = np.meshgrid(*[np.arange(s) for s in size], indexing="ij")
    coords = np.stack([x.reshape(-1) for x in grids], axis=1)
    g.insert_coords(coords, [coords[:, 0].astype(np.float32)])
    g.flush(["smax"])
    keys = np.arange(256)
    crd = g.inv_lin_id(keys)
    np.testing.assert_array_equal(g.get(0, keys), crd[:, 0].astype(np.float32))
    # a point of a missing chunk reads the background
    g2 = oracle.SparseGridCpu(dim, [2 * s for s in size], [np.float32], block_edge=case["block_edge"])
    g2.set_background_value(0, case["background"])
    g2.insert_coords(coords[:1], [np.array([5.0], np.float32)])
    g2.flush(["smax"])
    far = g2.lin_id([[s + 1 for s in size]])
    assert g2.get(0, far)[0] == case["background"]


@pytest.mark.parametrize("name", ["tag_boundaries_2d_1", "tag_boundaries_2d_2"])
def test_tag_boundaries_known_answer(oracle, golden, name):
    case = golden[name]
    g = S.known_tag_boundaries(oracle.SparseGridCpu, case)
    got = S.padding_ids(g, int(np.prod(case["size"])))
    assert sorted(got.tolist()) == sorted(case["padding_linear_ids"])


@pytest.mark.parametrize("seed", range(6))
def test_flush_against_dictionary_model(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    dim = 3 if seed % 2 else 2
    edge = [2, 4, 8][seed % 3]
    size = [int(rng.integers(edge, 5 * edge)) for _ in range(dim)]
    dtypes = [np.float64, np.int32, np.uint32, np.int16][: 1 + seed % 4]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=3, nrec=120)
    bgs = [7, 3, 9, 1][: len(dtypes)]
    g = S.run_flush_program(oracle.SparseGridCpu, dim, edge, size, dtypes, steps, background=bgs)
    model = S.model_flush(dim, edge, size, dtypes, steps, background=bgs)
    S.compare_to_model(g, model, dtypes)


@pytest.mark.parametrize("order", ["chunks_only", "chunks_first"])
@pytest.mark.parametrize("seed", range(4))
def test_flush_chunk_records_against_model(oracle, seed, order):
    rng = np.random.default_rng(300 + seed)
    dim = 3 if seed % 2 else 2
    edge = [2, 4, 8][seed % 3]
    size = [int(rng.integers(edge, 5 * edge)) for _ in range(dim)]
    dtypes = [np.float64, np.int32, np.uint32, np.int16][: 1 + seed % 4]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=3, nrec=80, order=order)
    bgs = [7, 3, 9, 1][: len(dtypes)]
    g = S.run_flush_program(oracle.SparseGridCpu, dim, edge, size, dtypes, steps, background=bgs)
    S.compare_to_model(g, S.model_flush(dim, edge, size, dtypes, steps, background=bgs), dtypes)


def _brute_neighbours(ids, bsz, dim, nn):
    """Geometric neighbour positions; NNFull reproduces the reference's unguarded -1 at the
    low faces through the plain linear-id formula (grid_smb::BlockLinId)."""
    pos = {int(c): i for i, c in enumerate(ids)}
    n = len(ids)
    out = []
    for c in ids:
        bc, b = [], int(c)
        for d in range(dim):
            bc.append(b % bsz[d])
            b //= bsz[d]
        row = []
        if nn == C.NN_STAR:
            for k in range(2 ** dim):
                if k >= 2 * dim:
                    row.append(-1)
                    continue
                q = list(bc)
                q[k // 2] += 1 if k % 2 else -1
                ok = all(0 <= q[d] < bsz[d] for d in range(dim))
                lin = 0
                for d in range(dim - 1, -1, -1):
                    lin = lin * bsz[d] + q[d]
                row.append(pos.get(lin, n) if ok else n)
        else:
            for k in range(3 ** dim):
                if k == (3 ** dim) // 2:
                    row.append(n)
                    continue
                q, kk = list(bc), k
                for d in range(dim):
                    q[d] += kk % 3 - 1
                    kk //= 3
                if any(q[d] >= bsz[d] for d in range(dim)):
                    row.append(n)
                    continue
                lin = 0
                for d in range(dim - 1, -1, -1):
                    lin = lin * bsz[d] + q[d]
                row.append(pos.get(lin, n) if lin >= 0 else n)
        out.append(row)
    return np.array(out, dtype=np.int32).reshape(n, -1)


@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
@pytest.mark.parametrize("dim", [2, 3])
def test_neighbour_tables_brute_force(oracle, nn, dim):
    rng = np.random.default_rng(7 + dim)
    size = [40] * dim
    g = oracle.SparseGridCpu(dim, size, [np.float32], block_edge=4)
    bsz = [10] * dim
    nb = int(np.prod(bsz))
    chunks = rng.choice(nb, nb // 3, replace=False)
    g.insert(chunks.astype(np.int64) * 4 ** dim, [np.ones(chunks.size, np.float32)])
    g.flush(["smax"])
    np.testing.assert_array_equal(g.neighbours(nn), _brute_neighbours(g.chunk_ids(), bsz, dim, nn))


def test_nnfull_low_face_aliasing(oracle):
    """At block x = 0 the NNFull -x neighbour is the chunk (bx_max, by - 1, bz): the
    reference's BlockLinId does not check -1 (grid_smb.hpp:258-276)."""
    g = oracle.SparseGridCpu(3, (32, 32, 32), [np.float32], block_edge=8)
    # chunk (0,1,1) and chunk (3,0,1) = its aliased -x neighbour
    ids = np.array([(1 * 4 + 1) * 4 + 0, (1 * 4 + 0) * 4 + 3], dtype=np.int64)
    g.insert(ids * 512, [np.ones(2, np.float32)])
    g.flush(["smax"])
    nnf = g.neighbours(C.NN_FULL)
    row = nnf[list(g.chunk_ids()).index(ids[0])]
    assert row[0 + 3 + 9] == list(g.chunk_ids()).index(ids[1])  # (-1, 0, 0) -> aliased chunk
    nns = g.neighbours(C.NN_STAR)
    assert nns[list(g.chunk_ids()).index(ids[0])][0] == g.num_chunks()  # NNStar guards it


def _shell_grid(G, n, lo, hi, props, edge=8, values=None):
    from openfpm_data_amd import workloads as W
    pts = W.shell_points(n, lo, hi)
    g = G(3, (n, n, n), props, block_edge=edge)
    vals = values(pts) if values else [np.ones(len(pts), dtype=p) for p in props]
    g.insert_coords(pts, vals)
    g.flush(["smax"] * len(props))
    return g, pts


@pytest.mark.parametrize("nn", [C.NN_STAR, C.NN_FULL])
def test_tag_boundaries_brute_force(oracle, nn):
    g, pts = _shell_grid(oracle.SparseGridCpu, 48, 10.0, 14.5, [np.float32])
    g.set_nn_type(nn)
    g.tag_boundaries(nn)
    ex = S.exist_dense(g) & 1
    pad = np.zeros_like(ex)
    P = np.pad(ex, 1)
    if nn == C.NN_STAR:
        offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    else:
        offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    for dz, dy, dx in offs:
        nbr = P[1 + dz:1 + dz + 48, 1 + dy:1 + dy + 48, 1 + dx:1 + dx + 48]
        pad |= ex & (nbr == 0)
    got = (S.exist_dense(g) & 2) >> 1
    np.testing.assert_array_equal(got, pad)


def _dense_lap(u, c0=-6.0):
    P = np.pad(u, 1)
    a = c0 * u
    n = u.shape[0]
    sl = lambda dz, dy, dx: P[1 + dz:1 + dz + n, 1 + dy:1 + dy + n, 1 + dx:1 + dx + n]
    a = a + (sl(0, 0, -1) + sl(0, 0, 1))
    a = a + (sl(0, -1, 0) + sl(0, 1, 0))
    a = a + (sl(-1, 0, 0) + sl(1, 0, 0))
    return a


def test_lap_against_dense_numpy(oracle):
    """The sparse Laplacian (background 0 for absent neighbours) equals a dense padded
    numpy Laplacian evaluated in the same operation order, bit for bit in FP64."""
    rng = np.random.default_rng(3)
    g, pts = _shell_grid(oracle.SparseGridCpu, 32, 6.0, 11.0, [np.float64, np.float64],
                         values=lambda p: [rng.uniform(-1, 1, len(p)), np.zeros(len(p))])
    u = S.dense_of(g, 0)
    g.conv_lap(0, 1)
    got = S.dense_of(g, 1)
    ex = (S.exist_dense(g) & 1) == 1
    np.testing.assert_array_equal(got[ex], _dense_lap(u)[ex])


def test_lap_closed_forms(oracle):
    """Interior points of a solid cube: a linear field has Laplacian 0 and x^2 + 2y^2 has 6."""
    n = 24
    g = oracle.SparseGridCpu(3, (n, n, n), [np.float64, np.float64], block_edge=8)
    grids = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    c = np.stack([x.reshape(-1) for x in grids], axis=1)
    g.insert_coords(c, [3.0 * c[:, 0] - 2.0 * c[:, 1] + 5.0 * c[:, 2] + 1.0, np.zeros(len(c))])
    g.flush(["smax", "smax"])
    g.tag_boundaries()
    g.conv_lap(0, 1)
    interior = ((S.exist_dense(g) & 3) == 1)
    assert interior.sum() == (n - 2) ** 3
    np.testing.assert_array_equal(S.dense_of(g, 1)[interior], 0.0)
    keys = g.lin_id(c)
    g.set_prop(0, np.zeros((g.num_chunks(), 512)))
    q = np.zeros(n ** 3)
    q[:] = (c[:, 0].astype(float) ** 2 + 2.0 * c[:, 1].astype(float) ** 2)
    arr = g.prop(0)
    ids = g.chunk_ids()
    arr[np.searchsorted(ids, keys // 512), keys % 512] = q
    g.set_prop(0, arr)
    g.conv_lap(0, 1)
    np.testing.assert_array_equal(S.dense_of(g, 1)[interior], 6.0)


def test_heat_steady_state_known_answer(oracle, golden):
    """SparseGridGpu_tests.cu:472-514: boundary fixed to 10*x/15, heat iterations converge to
    the linear profile within 1e-2."""
    case = golden["stencil_heat_2d"]
    g = oracle.SparseGridCpu(2, case["size"], [np.float32, np.float32], block_edge=8)
    grids = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    c = np.stack([x.reshape(-1) for x in grids], axis=1)
    g.insert_coords(c, [np.zeros(len(c), np.float32), np.zeros(len(c), np.float32)])
    g.flush(["smax", "smax"])
    g.tag_boundaries()
    m = g.masks()
    ids = g.chunk_ids()
    keys = (ids[:, None] * 64 + np.arange(64)[None, :]).reshape(-1)
    x = g.inv_lin_id(keys)[:, 0].reshape(m.shape)
    bnd = (m & 2) != 0
    u = np.where(bnd, 10.0 * x / 15.0, 0.0).astype(np.float32)
    g.set_prop(0, u)
    g.set_prop(1, u)
    for _ in range(case["iterations"]):
        g.conv_heat(0, 1, case["dt"][0])
        g.conv_heat(1, 0, case["dt"][1])
    got = g.prop(1)
    assert np.all(np.abs(got - 10.0 * x / 15.0) < case["tol"])


def test_box_all_ones_known_answer(oracle, golden):
    """SparseGridGpu_tests.cu:1012-1080: all-ones 3x3x3 conv is 27*value at non-padding points."""
    g, pts = _shell_grid(oracle.SparseGridCpu, 40, 8.0, 16.0, [np.uint32, np.uint32],
                         values=lambda p: [np.full(len(p), 5, np.uint32), np.zeros(len(p), np.uint32)])
    g.set_nn_type(C.NN_FULL)
    g.tag_boundaries(C.NN_FULL)
    g.conv_box(0, 1, np.ones(27, np.int64))
    m = g.masks()
    v = g.prop(1)
    assert (m == 1).sum() > 1000
    np.testing.assert_array_equal(v[m == 1], golden["conv3x3x3_all_ones"]["factor"] * 5)


def test_box_against_dense_numpy_uint32(oracle):
    rng = np.random.default_rng(11)
    g, pts = _shell_grid(oracle.SparseGridCpu, 32, 5.0, 12.0, [np.uint32, np.uint32],
                         values=lambda p: [rng.integers(0, 2 ** 32, len(p), dtype=np.uint64).astype(np.uint32),
                                           np.zeros(len(p), np.uint32)])
    g.set_nn_type(C.NN_FULL)
    g.tag_boundaries(C.NN_FULL)
    g.conv_box(0, 1, np.ones(27, np.int64))
    u = S.dense_of(g, 0)
    P = np.pad(u, 1).astype(np.uint64)
    tot = np.zeros(u.shape, np.uint64)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                tot += P[1 + dz:33 + dz, 1 + dy:33 + dy, 1 + dx:33 + dx]
    exp = (tot % (1 << 32)).astype(np.uint32)
    ex = (S.exist_dense(g) & 1) == 1
    np.testing.assert_array_equal(S.dense_of(g, 1)[ex], exp[ex])


def test_gray_scott_against_dense_numpy(oracle):
    rng = np.random.default_rng(12)
    g, pts = _shell_grid(oracle.SparseGridCpu, 32, 6.0, 12.0, [np.float32] * 4,
                         values=lambda p: [rng.uniform(0, 1, len(p)).astype(np.float32),
                                           rng.uniform(0, 0.5, len(p)).astype(np.float32),
                                           np.zeros(len(p), np.float32), np.zeros(len(p), np.float32)])
    P = dict(Du=2e-5, Dv=1e-5, F=0.04, k=0.06, dt=1.0)
    u = S.dense_of(g, 0)
    v = S.dense_of(g, 1)
    g.conv2_gray_scott(0, 1, 2, 3, **P)
    f = np.float32
    lu, lv = _dense_lap(u, f(-6.0)).astype(f), _dense_lap(v, f(-6.0)).astype(f)
    uvv = (u * v) * v
    nu = u + f(P["dt"]) * ((f(P["Du"]) * lu - uvv) + f(P["F"]) * (f(1) - u))
    nv = v + f(P["dt"]) * ((f(P["Dv"]) * lv + uvv) - (f(P["F"]) + f(P["k"])) * v)
    ex = (S.exist_dense(g) & 1) == 1
    np.testing.assert_array_equal(S.dense_of(g, 2)[ex], nu[ex])
    np.testing.assert_array_equal(S.dense_of(g, 3)[ex], nv[ex])


def test_inplace_is_jacobi(oracle):
    """In-place conv (src == dst) reads the pre-sweep state everywhere (ping-pong result)."""
    rng = np.random.default_rng(4)
    vals = lambda p: [rng.uniform(-1, 1, len(p)), np.zeros(len(p))]
    g1, _ = _shell_grid(oracle.SparseGridCpu, 32, 6.0, 11.0, [np.float64, np.float64], values=vals)
    rng = np.random.default_rng(4)
    g2, _ = _shell_grid(oracle.SparseGridCpu, 32, 6.0, 11.0, [np.float64, np.float64], values=vals)
    g1.conv_lap(0, 0)
    g2.conv_lap(0, 1)
    m = g1.masks() & 1 == 1
    np.testing.assert_array_equal(g1.prop(0)[m], g2.prop(1)[m])


def test_rbsor_colours_and_residual(oracle):
    """One red + one black sweep with w = 1 equals Gauss-Seidel written on a dense array;
    the residual of the exact discrete solution of a 1-point grid is 0."""
    rng = np.random.default_rng(9)
    g, pts = _shell_grid(oracle.SparseGridCpu, 24, 4.0, 9.0, [np.float64] * 3,
                         values=lambda p: [np.zeros(len(p)), rng.uniform(-1, 1, len(p)), np.zeros(len(p))])
    u = S.dense_of(g, 0)
    f = S.dense_of(g, 1)
    ex = (S.exist_dense(g) & 1) == 1
    n = 24
    zz, yy, xx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for color in (0, 1):
        g.rbsor_sweep(0, 1, 1.0, color)
        P = np.pad(u, 1)
        sl = lambda dz, dy, dx: P[1 + dz:1 + dz + n, 1 + dy:1 + dy + n, 1 + dx:1 + dx + n]
        nbs = 0.0 + (sl(0, 0, -1) + sl(0, 0, 1))
        nbs = nbs + (sl(0, -1, 0) + sl(0, 1, 0))
        nbs = nbs + (sl(-1, 0, 0) + sl(1, 0, 0))
        gs = (f + nbs) / 6.0
        new = u + 1.0 * (gs - u)
        sel = ex & (((xx + yy + zz) & 1) == color)
        u = np.where(sel, new, u)
    np.testing.assert_array_equal(S.dense_of(g, 0)[ex], u[ex])
    # residual of a single isolated point with u = f / 6 is zero
    h = oracle.SparseGridCpu(3, (16, 16, 16), [np.float64] * 3, block_edge=8)
    h.insert_coords([[5, 5, 5]], [np.array([0.5]), np.array([3.0]), np.array([7.0])])
    h.flush(["smax"] * 3)
    h.residual(0, 1, 2)
    assert h.get(2, h.lin_id([[5, 5, 5]]))[0] == 0.0


@pytest.mark.parametrize("seed", range(3))
def test_no_shared_mode_is_shared_on_existing_points(oracle, seed):
    """The point-list mode (applyStencilInPlaceNoShared) runs the same stencil over the EXIST
    points only: equal to the LDS mode there, and non-EXIST slots keep their values."""
    rng = np.random.default_rng(40 + seed)
    dim, edge = 3, [2, 4, 8][seed]
    size = [int(rng.integers(2 * edge, 5 * edge)) for _ in range(dim)]
    dtypes = [np.float64] * 3
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=2, nrec=200, ops=["smax"] * 3)
    for kind, src, dst, coef in [(C.ST_LAP, [0], [1], [-6.0, 1.0]), (C.ST_HEAT, [0], [1], [0.1]),
                                 (C.ST_GRAY_SCOTT, [0, 1], [2, 1], [0.2, 0.1, 0.04, 0.06, 1.0])]:
        for apply in (C.APPLY_INTERIOR, C.APPLY_EXISTING):
            gs = []
            for mode in (C.STENCIL_MODE_INPLACE, C.STENCIL_MODE_INPLACE_NO_SHARED):
                g = S.run_flush_program(oracle.SparseGridCpu, dim, edge, size, dtypes, steps)
                g.tag_boundaries(C.NN_STAR)
                before = [g.prop(p) for p in range(3)]
                g.apply_stencil(C.make_desc(kind, src, dst, apply, None, coef, mode=mode))
                gs.append((g, before))
            (a, _), (b, before) = gs
            ex = (b.masks() & 1) != 0
            for p in range(3):
                np.testing.assert_array_equal(a.prop(p)[ex], b.prop(p)[ex])
                np.testing.assert_array_equal(b.prop(p)[~ex], before[p][~ex])


def test_morton_known_answers(oracle):
    """lin_zid / invlin_zid known answers (util/test/zmorton_unit_tests.cpp:95-300)."""
    m = oracle.morton
    assert [m(2, k) for k in [(0, 0), (2, 2), (3, 2), (0, 3), (3, 0)]] == [0, 12, 13, 10, 5]
    assert m(2, (0xF, 0xF)) == 0xFF and m(2, (0xFF, 0xFF)) == 0xFFFF and m(2, (0xFFFF, 0xFFFF)) == 0xFFFFFFFF
    assert m(2, (0xFFFFFFFF, 0xFFFFFFFF)) == 0xFFFFFFFFFFFFFFFF
    assert [m(3, (v, v, v)) for v in (0, 2, 4, 8)] == [0, 56, 448, 3584]
    for k in range(4, 21):
        assert m(3, (1 << k,) * 3) == 7 << (3 * k)
    # 2D key (165, 347): bit i of x at bit 2i, of y at bit 2i+1
    r = m(2, (165, 347))
    assert all(((165 >> i) & 1) == ((r >> (2 * i)) & 1) and ((347 >> i) & 1) == ((r >> (2 * i + 1)) & 1)
               for i in range(16))


def test_morton_python_layer_matches_oracle(oracle):
    from openfpm_data_amd.sparsegrid import morton_decode, morton_encode
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        bc = rng.integers(0, 1 << (20 if dim == 3 else 30), (200, dim))
        ids = morton_encode(bc, dim)
        assert [int(i) for i in ids] == [oracle.morton(dim, b) for b in bc]
        np.testing.assert_array_equal(morton_decode(ids, dim), bc)


@pytest.mark.parametrize("dim,edge", [(2, 4), (3, 2), (3, 8)])
def test_zmb_grid_equals_smb_grid(oracle, dim, edge):
    """The same points in a Morton (SparseGridGpu_z) and a row-major grid: the same field
    after flush, NNStar tagBoundaries and Laplacian sweeps (chunk order differs only)."""
    rng = np.random.default_rng(60 + dim * edge)
    size = [int(rng.integers(3 * edge, 7 * edge)) for _ in range(dim)]
    n = 400
    coords = np.stack([rng.integers(0, s, n) for s in size], axis=1).astype(np.int32)
    vals = rng.integers(-20, 20, n).astype(np.float64)
    dense = []
    for lin in ("smb", "zmb"):
        g = oracle.SparseGridCpu(dim, size, [np.float64, np.float64], block_edge=edge, linearizer=lin)
        g.insert_coords(coords, [vals, vals])
        g.flush(["sadd", "smax"])
        g.find_neighbours(C.NN_STAR)
        g.tag_boundaries(C.NN_STAR)
        g.conv_lap(0, 1)
        g.conv_lap(1, 0, apply=C.APPLY_EXISTING)
        m = g.masks().reshape(-1)
        keys = (g.chunk_ids()[:, None] * g.chunk_size + np.arange(g.chunk_size)[None, :]).reshape(-1)[m != 0]
        md = np.zeros(size[::-1], np.uint8)
        crd = g.inv_lin_id(keys)
        md[tuple(crd[:, d] for d in range(dim - 1, -1, -1))] = m[m != 0]
        dense.append((S.dense_of(g, 0), S.dense_of(g, 1), md))
        keys = g.lin_id(coords)
        np.testing.assert_array_equal(g.inv_lin_id(keys), coords)
    for a, b in zip(dense[0], dense[1]):
        np.testing.assert_array_equal(a, b)


def test_flush_known_answers_of_the_reference(oracle):
    """BlockMapGpu_tests.cu:126-264 and BlockMapGpu_kernels_tests.cu:16-492 (segmented fold
    with old data, merge maps, conflict resolution), end to end through insert + flush."""
    S.check_flush_known_answers(oracle.SparseGridCpu)
