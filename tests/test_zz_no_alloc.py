"""sg_pack_noalloc (pack into device memory with the temporaries in a static per-GPU
scratch array) through the C ABI: its bytes equal sg_pack's into device memory on calls in
a row (the second on a smaller grid, so it reads a scratch holding the first call's rows),
it neither grows the grid's device memory nor changes the GPU's free memory, it leaves the
destination's bytes past the pack alone, and it refuses requests beyond its limits before
any device work."""
import ctypes

import numpy as np
import pytest

from openfpm_data_amd import _capi as C
from tests import scenarios as S

pytestmark = pytest.mark.gpu

FILL = 0xAB


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    from openfpm_data_amd import SparseGridGpu
    return SparseGridGpu


def _args(g, boxes, props):
    lo, hi = g._boxes(boxes)
    pr = np.ascontiguousarray(np.asarray(props, np.int32))
    return lo, hi, pr


def _noalloc(g, boxes, props, index_bytes, cap):
    """sg_pack_noalloc into a fresh device buffer of `cap` bytes filled with FILL; also the
    GPU's free memory right before and right after the call."""
    import torch
    lo, hi, pr = _args(g, boxes, props)
    d = torch.full((cap,), FILL, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    written = ctypes.c_int64(-1)
    free_before = torch.cuda.mem_get_info()[0]
    rc = g._lib.sg_pack_noalloc(g._h, len(boxes), lo.ctypes.data, hi.ctypes.data, pr.size, pr.ctypes.data,
                                int(index_bytes), ctypes.c_void_p(d.data_ptr()), cap, ctypes.byref(written))
    free_after = torch.cuda.mem_get_info()[0]
    return rc, written.value, d, (free_before, free_after)


def _device_bytes(g):
    v = ctypes.c_int64(0)
    C.check(g._lib.sg_device_bytes(g._h, ctypes.byref(v)))
    return v.value


def _check_against_pack(g, boxes, props, index_bytes):
    import torch
    n = g.pack_size(boxes, props, index_bytes)
    ref = torch.zeros(n + 64, dtype=torch.uint8, device="cuda")
    assert g.pack(boxes, props, index_bytes, out=ref) == n
    before_grid = _device_bytes(g)
    rc, written, d, (before_free, after_free) = _noalloc(g, boxes, props, index_bytes, n + 64)
    C.check(rc)
    assert written == n
    np.testing.assert_array_equal(d[:n].cpu().numpy(), ref[:n].cpu().numpy())
    assert bool((d[n:] == FILL).all())
    assert _device_bytes(g) == before_grid
    return n, before_free, after_free, d[:n].cpu().numpy()


def test_noalloc_pack_matches_pack_on_calls_in_a_row(G):
    print("sg_pack_noalloc: sphere shell grid, two boxes, property 0, 8-byte ids", flush=True)
    big, _ = S.sphere_grid(G)
    boxes = [((0, 0, 0), (321, 999, 999)), ((322, 0, 0), (999, 999, 999))]
    n1, _, _, buf = _check_against_pack(big, boxes, [0], 8)
    assert S.pack_section_counts(buf, 4) == [(4685, 185807), (4475, 176787)]

    print("sg_pack_noalloc: smaller 2D grid after it, three boxes, two properties, 4-byte ids", flush=True)
    rng = np.random.default_rng(7)
    small = G(2, (96, 80), [np.uint16, np.float64], block_edge=4)
    pts = np.unique(rng.integers(0, (96, 80), size=(1500, 2)).astype(np.int32), axis=0)
    small.insert_coords(pts, [rng.integers(0, 60000, len(pts)).astype(np.uint16), rng.standard_normal(len(pts))])
    small.flush(["smax", "sadd"])
    assert small.num_chunks() < big.num_chunks()
    sboxes = [((0, 0), (40, 79)), ((10, 5), (60, 30)), ((90, 70), (95, 79))]
    n2, free_before, free_after, _ = _check_against_pack(small, sboxes, [1, 0], 4)
    # no device allocation on a call after the first (lazy set-up done)
    assert free_after == free_before, (free_before, free_after)

    print("sg_pack_noalloc: sphere grid again, one small box; an empty box; an empty grid", flush=True)
    _check_against_pack(big, [((300, 300, 300), (340, 330, 335))], [0], 8)
    _check_against_pack(big, [((0, 0, 0), (1, 1, 1))], [0], 8)
    empty = G(3, (64, 64, 64), [np.float32], block_edge=8)
    _check_against_pack(empty, [((0, 0, 0), (63, 63, 63))], [0], 8)
    assert n1 > n2 > 0
    for g in (big, small, empty):
        g.close()


def test_noalloc_pack_refuses_beyond_its_limits(G):
    print("sg_pack_noalloc: limits (too many boxes, rows beyond the scratch, small buffer)", flush=True)
    rng = np.random.default_rng(3)
    # 2D, chunks of 2 x 2 points: 40,000 distinct chunks
    g = G(2, (1024, 1024), [np.float32], block_edge=2)
    cid = rng.choice(512 * 512, size=40_000, replace=False)
    pts = np.stack([(cid % 512) * 2, (cid // 512) * 2], axis=1).astype(np.int32)
    g.insert_coords(pts, [np.ones(len(pts), np.float32)])
    g.flush(["smax"])
    assert g.num_chunks() == 40_000
    box = ((0, 0), (1023, 1023))
    rc, written, _, _ = _noalloc(g, [box] * (C.PACK_NOALLOC_MAX_BOXES + 1), [0], 8, 1 << 20)
    assert rc == C.SG_E_RANGE and "boxes" in C.last_error()
    # 64 boxes x 40,001 rows x 16 bytes > 32 MiB
    rc, written, _, _ = _noalloc(g, [box] * 64, [0], 8, 1 << 20)
    assert rc == C.SG_E_RANGE and "scratch bytes" in C.last_error()
    # within the limits but a destination too small: refused, destination untouched
    n = g.pack_size([box], [0], 8)
    rc, written, d, _ = _noalloc(g, [box], [0], 8, n - 1)
    assert rc == C.SG_E_RANGE and "too small" in C.last_error()
    assert bool((d == FILL).all())
    # and the same grid within the limits
    _check_against_pack(g, [box, ((0, 0), (511, 255))], [0], 8)
    g.close()
