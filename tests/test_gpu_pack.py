"""Pack / unpack / remove on the GPU: the reference's known answers
(SparseGridGpu_tests.cu:1379-1934), packed bytes identical to the oracle's, and unpacked
grids in the same state as the oracle's."""
import numpy as np
import pytest

from tests import scenarios as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from openfpm_data_amd import SparseGridGpu
    return SparseGridGpu


def test_pack_known_answers_and_bytes(G, oracle):
    g, s = S.sphere_grid(G)
    o, _ = S.sphere_grid(oracle.SparseGridCpu)
    boxes = [((0, 0, 0), (321, 999, 999)), ((322, 0, 0), (999, 999, 999))]
    assert g.pack_size(boxes, [0]) == o.pack_size(boxes, [0])
    buf = g.pack(boxes, [0])
    assert S.pack_section_counts(buf, 4) == [(4685, 185807), (4475, 176787)]
    np.testing.assert_array_equal(buf, o.pack(boxes, [0]))
    # into a device buffer, and with 4-byte chunk ids
    import torch
    d = torch.zeros(len(buf) + 64, dtype=torch.uint8, device="cuda")
    n = g.pack(boxes, [0], out=d)
    np.testing.assert_array_equal(d[:n].cpu().numpy(), buf)
    np.testing.assert_array_equal(g.pack(boxes, [0], index_bytes=4), o.pack(boxes, [0], index_bytes=4))
    g.close()
    o.close()


def test_remove_sections(G, oracle):
    secs = [((310, 0, 0), (330, 999, 999)), ((0, 310, 0), (999, 330, 999)), ((0, 0, 310), (999, 999, 330))]
    grids = []
    for cls in (G, oracle.SparseGridCpu):
        g, s = S.sphere_grid(cls)
        for lo, hi in secs:
            g.remove_box(lo, hi)
        g.remove_add_unpack_finalize([])
        grids.append(g)
    S.assert_same_state(grids[0], grids[1], exist_only=False)
    p = S.existing_points(grids[0])
    for lo, hi in secs:
        assert not S.in_box(p, lo, hi).any()


def test_pack_unpack_known_counts(G, oracle):
    state = []
    for cls in (G, oracle.SparseGridCpu):
        src, s = S.sphere_grid(cls, vector=True)
        dst = cls(3, s["size"], [np.float32] * 4, block_edge=4)
        buf = S.pack_unpack_round(dst, src)
        assert S.pack_section_counts(buf, 16) == [(1107, 41003), (1420, 54276), (610, 20828), (739, 27283)]
        p = S.existing_points(dst)
        assert [int(S.in_box(p, lo, hi).sum()) for lo, hi in S.PACK_DST_BOXES] == [41003, 54276, 20828, 27283]
        sp = S.existing_points(src)
        src.remove_points(src.lin_id(sp[(sp.sum(axis=1) % 2) == 0]))
        buf2 = S.pack_unpack_round(dst, src)
        p = S.existing_points(dst)
        assert [int(S.in_box(p, lo, hi).sum()) for lo, hi in S.PACK_DST_BOXES] == [20520, 27152, 10423, 13649]
        for k, (lo, hi) in enumerate(S.PACK_DST_BOXES):
            assert S.shell_in_dst_box(p[S.in_box(p, lo, hi)], s["centre"], k).all()
        state.append((buf, buf2, dst, src))
    np.testing.assert_array_equal(state[0][0], state[1][0])
    np.testing.assert_array_equal(state[0][1], state[1][1])
    S.assert_same_state(state[0][2], state[1][2])
    S.assert_same_state(state[0][3], state[1][3], exist_only=False)


@pytest.mark.parametrize("seed", range(6))
def test_pack_unpack_random(G, oracle, seed):
    """Random grids (2D/3D, edges 2/4/8, mixed types), random boxes and property subsets:
    packed bytes equal, unpacked with a random shift into a grid that already holds points
    (one destination box removed first): same state as the oracle."""
    rng = np.random.default_rng(900 + seed)
    dim = 3 if seed % 2 else 2
    edge = [2, 4, 8][seed % 3]
    size = [int(rng.integers(3 * edge, 8 * edge)) for _ in range(dim)]
    dtypes = [[np.float64, np.int32], [np.float32, np.uint8, np.int64], [np.uint16, np.float64]][seed % 3]
    steps = S.random_flush_program(seed, dim, edge, size, dtypes, nflush=2, nrec=300)
    props = sorted(rng.choice(len(dtypes), size=int(rng.integers(1, len(dtypes) + 1)), replace=False).tolist())
    boxes = []
    for _ in range(3):
        lo = [int(rng.integers(0, s // 2)) for s in size]
        hi = [int(rng.integers(l, s)) for l, s in zip(lo, size)]
        boxes.append((lo, hi))
    res = []
    for cls in (G, oracle.SparseGridCpu):
        src = S.run_flush_program(cls, dim, edge, size, dtypes, steps)
        dst = S.run_flush_program(cls, dim, edge, size, dtypes, steps[:1])
        buf = src.pack(boxes, props, index_bytes=[4, 8][seed % 2])
        dst.remove_box(boxes[0][0], boxes[0][1])
        at = 0
        for lo, hi in boxes:
            ext = [h - l for l, h in zip(lo, hi)]
            dlo = np.minimum(np.array(lo) + 1, np.array(size) - 1 - np.array(ext)).tolist()
            at += dst.unpack(buf, dlo, props, index_bytes=[4, 8][seed % 2], offset=at)
        dst.remove_add_unpack_finalize(props)
        res.append((buf, dst))
    np.testing.assert_array_equal(res[0][0], res[1][0])
    S.assert_same_state(res[0][1], res[1][1])


@pytest.mark.parametrize("seed,props", [(0, (0,)), (1, (0, 1)), (2, (1,))])
def test_copy_to_ring(G, oracle, seed, props):
    """copy_to between three grids (one copying into itself, one removing a box), phases
    0, 1, 2 on every grid: the same grids as the oracle's."""
    res = [S.copy_ring(cls, seed, props)[0] for cls in (G, oracle.SparseGridCpu)]
    for a, b in zip(*res):
        S.assert_same_state(a, b, exist_only=False)
        a.close()
        b.close()


@pytest.mark.parametrize("dim,edge,lin", [(3, 4, "smb"), (3, 8, "zmb"), (2, 4, "smb")])
def test_links_match_oracle(G, oracle, dim, edge, lin):
    """construct_link_dw / construct_link_up: padding points, offsets and links equal the
    oracle's, both directions."""
    res = []
    for cls in (G, oracle.SparseGridCpu):
        c, f = S.link_levels(cls, 11, dim, edge, lin)
        box = ((0, 1, 0)[:dim], (19, 22, 15)[:dim])
        c.construct_link(f, 0, box, (1, 0, 1)[:dim])
        f.tag_boundaries()
        f.construct_link(c, 1, tuple(tuple(2 * v for v in b) for b in box), (0, 1, 1)[:dim])
        res.append((c.links(0), f.links(1)))
        c.close()
        f.close()
    for a, b in zip(res[0][0] + res[0][1], res[1][0] + res[1][1]):
        np.testing.assert_array_equal(a, b)
    assert len(res[0][0][3]) > 0 and len(res[0][1][3]) > 0


def test_unpack_outside_grid_fails(G):
    from openfpm_data_amd._capi import SgError
    g = G(3, (32, 32, 32), [np.float32], block_edge=8)
    g.insert_coords([[30, 30, 30]], [np.array([1.0], np.float32)])
    g.flush(["smax"])
    buf = g.pack([((28, 28, 28), (31, 31, 31))], [0])
    with pytest.raises(SgError):
        g.unpack(buf, (30, 30, 30), [0])  # the point would land at 32
    assert g.pending_inserts() == 0
    g.close()
