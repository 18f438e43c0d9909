"""Oracle pack / unpack / remove pinned against the reference's known answers
(SparseGridGpu_tests.cu:1379-1934): byte sizes of packed boxes, chunk and point counts of
each packed box, point counts after unpacking into shifted boxes and after removing the
even points, and the removal of box sections."""
import numpy as np
import pytest

from tests import scenarios as S


def al8(v):
    return (v + 7) // 8 * 8


def test_pack_size_two_boxes(oracle):
    """test_pack_request_with_iterator (SparseGridGpu_tests.cu:1379-1472)."""
    g, s = S.sphere_grid(oracle.SparseGridCpu)
    req = g.pack_size([((0, 0, 0), (321, 999, 999)), ((322, 0, 0), (999, 999, 999))], [0])
    tot = (2 * 8 + 2 * 2 * 3 * 4 + al8(4685 * 8) + al8(4475 * 8) + al8(4686 * 4) + al8(4476 * 4)
           + al8(185807 * 4) + al8(176787 * 4) + al8(185807 * 2) + al8(176787 * 2) + al8(185807) + al8(176787))
    assert req == tot
    # the same whole-grid box twice
    nch, cnt = g.num_chunks(), g.count_existing_elements()
    req = g.pack_size([((0, 0, 0), (999, 999, 999))] * 2, [0])
    assert req == 2 * 8 + 2 * 2 * 3 * 4 + 2 * (al8(nch * 8) + al8((nch + 1) * 4) + al8(cnt * 4) + al8(cnt * 2) + al8(cnt))
    buf = g.pack([((0, 0, 0), (321, 999, 999)), ((322, 0, 0), (999, 999, 999))], [0])
    assert S.pack_section_counts(buf, 4) == [(4685, 185807), (4475, 176787)]
    g.close()


def test_remove_sections(oracle):
    """sparsegridgpu_remove_test (SparseGridGpu_tests.cu:1474-1551)."""
    g, s = S.sphere_grid(oracle.SparseGridCpu)
    secs = [((310, 0, 0), (330, 999, 999)), ((0, 310, 0), (999, 330, 999)), ((0, 0, 310), (999, 999, 330))]
    before = S.existing_points(g)
    for lo, hi in secs:
        g.remove_box(lo, hi)
    g.remove_add_unpack_finalize([])
    p = S.existing_points(g)
    r = np.sqrt(((p - 320) ** 2).sum(axis=1))
    assert np.all((r >= 55.99) & (r <= 64.01))
    for lo, hi in secs:
        assert not S.in_box(p, lo, hi).any()
    # exactly the points outside the sections survive
    keep = ~np.any([S.in_box(before, lo, hi) for lo, hi in secs], axis=0)
    assert p.shape[0] == keep.sum() > 0
    # removed slots hold the background (0) with mask 0
    m = g.masks()
    assert np.all(g.prop(0)[m == 0] == 0)
    g.close()


def test_pack_unpack_known_counts(oracle):
    """sparsegridgpu_pack_unpack (SparseGridGpu_tests.cu:1684-1934)."""
    src, s = S.sphere_grid(oracle.SparseGridCpu, vector=True)
    dst = oracle.SparseGridCpu(3, s["size"], [np.float32] * 4, block_edge=4)
    buf = S.pack_unpack_round(dst, src)
    assert S.pack_section_counts(buf, 16) == [(1107, 41003), (1420, 54276), (610, 20828), (739, 27283)]
    p = S.existing_points(dst)
    counts = []
    for k, (lo, hi) in enumerate(S.PACK_DST_BOXES):
        ins = S.in_box(p, lo, hi)
        counts.append(int(ins.sum()))
        assert S.shell_in_dst_box(p[ins], s["centre"], k).all()
    assert counts == [41003, 54276, 20828, 27283]
    # unpacked values follow the points: x+y+z and x, y, z of the source position
    keys = dst.lin_id(p)
    shift = np.zeros((p.shape[0], 3), np.int64)
    for k, (lo, hi) in enumerate(S.PACK_DST_BOXES):
        shift[S.in_box(p, lo, hi), 0] = [0, -46, 44, 0][k]
    q = p - shift - s["centre"]
    np.testing.assert_array_equal(dst.get(1, keys), q[:, 0].astype(np.float32))
    np.testing.assert_array_equal(dst.get(0, keys), q.sum(axis=1).astype(np.float32))
    # remove the even points of the source, pack/unpack again
    sp = S.existing_points(src)
    even = sp[(sp.sum(axis=1) % 2) == 0]
    src.remove_points(src.lin_id(even))
    S.pack_unpack_round(dst, src)
    p = S.existing_points(dst)
    counts = [int(S.in_box(p, lo, hi).sum()) for lo, hi in S.PACK_DST_BOXES]
    assert counts == [20520, 27152, 10423, 13649]
    src.close()
    dst.close()


@pytest.mark.parametrize("seed,props", [(0, (0,)), (1, (0, 1)), (2, (1,))])
def test_copy_to_ring_matches_point_model(oracle, seed, props):
    """copy_to + removeCopyToFinalize (SparseGridGpu.hpp:1116-1290, :3314-3329, :3550-3559):
    removes before every copy, copies read their sources as the removes left them, copied
    properties take the source values, the others keep theirs, mask |= EXIST."""
    grids, before, sections, removes = S.copy_ring(oracle.SparseGridCpu, seed, props)
    model = S.model_copy_ring(before, sections, removes, props, (-1.5, 7))
    for g, m in zip(grids, model):
        assert S.point_state(g) == m
        g.close()


@pytest.mark.parametrize("dim,which,p", [(3, 0, (1, 0, 1)), (3, 0, (0, 0, 0)), (2, 0, (1, 1)), (3, 1, (1, 0, 0)),
                                         (2, 1, (0, 1))])
def test_links_match_coordinate_model(oracle, dim, which, p):
    """construct_link_dw / construct_link_up (SparseGridGpu.hpp:2090-2262) against the
    model on coordinates: padding points in the box, their children / parent that exist."""
    c, f = S.link_levels(oracle.SparseGridCpu, 5 + dim + which, dim)
    box = ((1, 2, 1)[:dim], (17, 20, 14)[:dim])
    if which == 0:
        c.construct_link(f, 0, box, p)
        got = c.links(0)
        want = S.model_links(c, f, 0, box, p)
    else:
        fbox = tuple(tuple(2 * v for v in b) for b in box)
        f.tag_boundaries()
        f.construct_link(c, 1, fbox, p)
        got = f.links(1)
        want = S.model_links(f, c, 1, fbox, p)
    assert len(want[0]) > 20 and len(want[3]) > 10
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, np.asarray(b, a.dtype))
    c.close()
    f.close()
