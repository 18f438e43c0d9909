"""Host-side SparseGridGpu over libsg_hip.so.

Mirrors the host API of openfpm_data's ``SparseGridGpu<dim, aggregate<...>, blockEdgeSize>``
(src/SparseGridGpu/SparseGridGpu.hpp) for Python callers: the same operations, argument
meaning and state machine (insert -> flush -> findNeighbours/tagBoundaries -> stencils),
with errors raised as :class:`SgError` instead of printed.  Buffers may be numpy arrays
(host) or torch tensors on the GPU (device, zero copy).
"""
import ctypes

import numpy as np

from . import _capi as C


LINEARIZERS = {"smb": 0, "zmb": 1}


def morton_encode(bc, dim):
    """grid_zmb chunk ids (lin_zid, util/zmorton.hpp): bit i of chunk coordinate d goes to
    bit dim*i + d.  bc: [n, dim] integer array."""
    bc = np.asarray(bc, dtype=np.int64).reshape(-1, dim).astype(np.uint64)
    out = np.zeros(bc.shape[0], dtype=np.uint64)
    for d in range(dim):
        for i in range(64 // dim):
            out |= ((bc[:, d] >> np.uint64(i)) & np.uint64(1)) << np.uint64(dim * i + d)
    return out.astype(np.int64)


def morton_decode(ids, dim):
    """invlin_zid: chunk ids -> [n, dim] chunk coordinates."""
    v = np.asarray(ids, dtype=np.int64).reshape(-1).astype(np.uint64)
    out = np.zeros((v.size, dim), dtype=np.int64)
    for d in range(dim):
        acc = np.zeros(v.size, dtype=np.uint64)
        for i in range(64 // dim):
            acc |= ((v >> np.uint64(dim * i + d)) & np.uint64(1)) << np.uint64(i)
        out[:, d] = acc.astype(np.int64)
    return out


def _is_device(x):
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _ptr(x):
    if x is None:
        return None
    if _is_device(x):
        return ctypes.c_void_p(x.data_ptr())
    return ctypes.c_void_p(x.ctypes.data)


def _check_device(t, dtype, numel, what):
    """A device tensor handed to the C ABI: contiguous, of the element size and kind (float or
    integer) of `dtype` (integers of either signedness share the bytes), `numel` elements."""
    dtype = np.dtype(dtype)
    if not t.is_contiguous():
        raise ValueError("%s: the device tensor must be contiguous" % what)
    if t.element_size() != dtype.itemsize or t.is_floating_point() != (dtype.kind == "f"):
        raise ValueError("%s: device tensor of %s where %s is expected" % (what, t.dtype, dtype.name))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s: %d elements where %d are expected" % (what, t.numel(), numel))
    return t


def _mem(*xs):
    dev = [_is_device(x) for x in xs if x is not None]
    if dev and all(dev):
        return C.SG_MEM_DEVICE
    if any(dev):
        raise ValueError("mix of host and device buffers in one call")
    return C.SG_MEM_HOST


class SparseGridGpu:
    """Block-sparse grid on one MI355X.

    ``props`` lists the scalar type of each property (the reference's ``aggregate<T...>``),
    ``block_edge`` the chunk edge (8 -> 8^3 chunks in 3D, the reference's default,
    SparseGridGpu.hpp:59-64).
    """

    def __init__(self, dim, size, props, block_edge=8, stencil_radius=1, linearizer="smb"):
        self._lib = C.load()
        self.dim = int(dim)
        self.size = tuple(int(s) for s in size)
        self.dtypes = [np.dtype(p) for p in props]
        self.block_edge = int(block_edge)
        self.chunk_size = self.block_edge ** self.dim
        self.block_sizes = tuple(-(-s // self.block_edge) for s in self.size)
        sz = (ctypes.c_int64 * 3)(*(list(self.size) + [1] * (3 - self.dim)))
        codes = (ctypes.c_int32 * max(1, len(self.dtypes)))(*[C.DTYPE_CODES[d] for d in self.dtypes])
        h = ctypes.c_void_p()
        C.check(self._lib.sg_create(self.dim, sz, self.block_edge, len(self.dtypes), codes,
                                    int(stencil_radius), ctypes.byref(h)))
        self._h = h
        # grid_smb (SparseGridGpu) or grid_zmb (SparseGridGpu_z) chunk ids
        self.linearizer = linearizer
        if linearizer != "smb":
            C.check(self._lib.sg_set_linearizer(h, LINEARIZERS[linearizer]))

    # -- lifetime ---------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.sg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """Run all later work on a HIP stream (``torch.cuda.Stream`` or a raw handle)."""
        raw = getattr(stream, "cuda_stream", stream)
        C.check(self._lib.sg_set_stream(self._h, ctypes.c_void_p(raw)))

    def set_tuning(self, key, value):
        C.check(self._lib.sg_set_tuning(self._h, int(key), int(value)))

    def synchronize(self):
        C.check(self._lib.sg_synchronize(self._h))

    # -- geometry (grid_smb) ----------------------------------------------------------
    def lin_id(self, coords):
        """grid_smb::LinId for an [n, dim] integer array (grid_smb.hpp:175-189)."""
        c = np.asarray(coords, dtype=np.int64).reshape(-1, self.dim)
        e = self.block_edge
        b = c[:, self.dim - 1] // e
        l = c[:, self.dim - 1] % e
        for d in range(self.dim - 2, -1, -1):
            b = b * self.block_sizes[d] + c[:, d] // e
            l = l * e + c[:, d] % e
        if self.linearizer == "zmb":
            b = morton_encode(c // e, self.dim)
        return b * self.chunk_size + l

    def inv_lin_id(self, keys):
        """grid_smb::InvLinId: point keys -> [n, dim] coordinates."""
        k = np.asarray(keys, dtype=np.int64).reshape(-1)
        b = k // self.chunk_size
        l = k % self.chunk_size
        out = np.empty((k.size, self.dim), dtype=np.int64)
        bc = morton_decode(b, self.dim) if self.linearizer == "zmb" else None
        for d in range(self.dim):
            cb = bc[:, d] if bc is not None else b % self.block_sizes[d]
            out[:, d] = cb * self.block_edge + l % self.block_edge
            b //= self.block_sizes[d]
            l //= self.block_edge
        return out

    # -- background / insert / flush ------------------------------------------------------
    def set_background_value(self, p, value):
        v = np.array([value], dtype=self.dtypes[p])
        C.check(self._lib.sg_set_background(self._h, p, v.ctypes.data))

    def _values(self, values, n, per=1):
        if values is None:
            return None, []
        if not isinstance(values, (list, tuple)):
            values = [values]
        if len(values) != len(self.dtypes):
            raise ValueError("one value array per property (None: the records carry the background)")
        keep = []
        arr = (ctypes.c_void_p * len(values))()
        for p, v in enumerate(values):
            if v is None:
                arr[p] = None
                continue
            if not _is_device(v):
                v = np.ascontiguousarray(np.asarray(v, dtype=self.dtypes[p]).reshape(-1))
                if v.size != n * per:
                    raise ValueError("property %d: expected %d values" % (p, n * per))
            else:
                _check_device(v, self.dtypes[p], n * per, "property %d" % p)
            keep.append(v)
            arr[p] = _ptr(v).value
        return arr, keep

    def _masks(self, masks, count):
        if masks is None:
            return None
        if _is_device(masks):
            return _check_device(masks, np.uint8, count, "masks")
        masks = np.ascontiguousarray(np.asarray(masks, dtype=np.uint8).reshape(-1))
        if masks.size != count:
            raise ValueError("masks: expected %d bytes" % count)
        return masks

    def insert(self, keys, values=None, masks=None):
        """Point insertion (``insert<p>(coord)`` once per point) by linear key."""
        if not _is_device(keys):
            keys = np.ascontiguousarray(np.asarray(keys, dtype=np.int64).reshape(-1))
        else:
            _check_device(keys, np.int64, None, "keys")
        n = keys.numel() if _is_device(keys) else keys.size
        masks = self._masks(masks, n)
        vals, keep = self._values(values, n)
        mem = _mem(keys, masks, *keep)
        C.check(self._lib.sg_insert_points(self._h, n, _ptr(keys), _ptr(masks), vals, mem))

    def insert_coords(self, coords, values=None, masks=None):
        """Point insertion by integer coordinates [n, dim]."""
        if not _is_device(coords):
            coords = np.ascontiguousarray(np.asarray(coords, dtype=np.int32).reshape(-1, self.dim))
            n = coords.shape[0]
        else:
            if coords.dim() != 2 or coords.shape[1] != self.dim:
                raise ValueError("coords: device tensor of shape (n, %d) expected" % self.dim)
            _check_device(coords, np.int32, None, "coords")
            n = coords.shape[0]
        masks = self._masks(masks, n)
        vals, keep = self._values(values, n)
        mem = _mem(coords, masks, *keep)
        C.check(self._lib.sg_insert_coords(self._h, n, _ptr(coords), _ptr(masks), vals, mem))

    def insert_block(self, chunk_ids, masks, values=None):
        """Chunk insertion (``insertBlock(blockId)``): masks/values are [n, chunk_size]."""
        if not _is_device(chunk_ids):
            chunk_ids = np.ascontiguousarray(np.asarray(chunk_ids, dtype=np.int64).reshape(-1))
            n = chunk_ids.size
        else:
            _check_device(chunk_ids, np.int64, None, "chunk_ids")
            n = chunk_ids.numel()
        masks = self._masks(masks, n * self.chunk_size)
        if masks is None:
            raise ValueError("insert_block needs the masks of the chunks")
        vals, keep = self._values(values, n, self.chunk_size)
        mem = _mem(chunk_ids, masks, *keep)
        C.check(self._lib.sg_insert_chunks(self._h, n, _ptr(chunk_ids), _ptr(masks), vals, mem))

    def pending_inserts(self):
        return self._lib.sg_pending_inserts(self._h)

    def flush(self, ops):
        """``flush<v_reduce...>(ctx, FLUSH_ON_DEVICE)``; ops: one name or code per property
        ('smax', 'smin', 'sadd', 'sright', 'sleft', 'sor', or 'none')."""
        codes = [C.OPS[o] if isinstance(o, str) else int(o) for o in ops]
        if len(codes) != len(self.dtypes):
            raise ValueError("one reduction per property")
        arr = (ctypes.c_int32 * max(1, len(codes)))(*codes)
        C.check(self._lib.sg_flush(self._h, arr))

    # -- neighbourhood / boundaries / stencils -------------------------------------------
    def find_neighbours(self, nn=C.NN_STAR):
        C.check(self._lib.sg_find_neighbours(self._h, nn))

    def set_nn_type(self, nn):
        C.check(self._lib.sg_set_nn_type(self._h, nn))

    def tag_boundaries(self, nn=C.NN_STAR, opt=C.NO_CALCULATE_EXISTING_POINTS):
        C.check(self._lib.sg_tag_boundaries(self._h, nn, opt))

    def apply_stencil(self, desc):
        C.check(self._lib.sg_apply_stencil(self._h, ctypes.byref(desc)))

    def conv_lap(self, src, dst, c0=-6.0, c1=1.0, box=None, apply=C.APPLY_INTERIOR):
        self.apply_stencil(C.make_desc(C.ST_LAP, [src], [dst], apply, box, [c0, c1]))

    def conv_heat(self, src, dst, dt, box=None):
        self.apply_stencil(C.make_desc(C.ST_HEAT, [src], [dst], C.APPLY_INTERIOR, box, [dt]))

    def conv2_gray_scott(self, u, v, u_out, v_out, Du, Dv, F, k, dt, box=None):
        self.apply_stencil(C.make_desc(C.ST_GRAY_SCOTT, [u, v], [u_out, v_out], C.APPLY_INTERIOR, box,
                                       [Du, Dv, F, k, dt]))

    def conv_box(self, src, dst, coef, box=None, apply=C.APPLY_EXISTING):
        coef = np.asarray(coef).reshape(-1)
        if np.issubdtype(self.dtypes[src], np.integer):
            d = C.make_desc(C.ST_BOX, [src], [dst], apply, box, (), [int(c) for c in coef])
        else:
            d = C.make_desc(C.ST_BOX, [src], [dst], apply, box, [float(c) for c in coef])
        self.apply_stencil(d)

    def rbsor_sweep(self, u, f, omega, color, box=None):
        self.apply_stencil(C.make_desc(C.ST_RBSOR, [u, f], [], C.APPLY_INTERIOR, box, [omega], [color]))

    def residual(self, u, f, r, box=None):
        self.apply_stencil(C.make_desc(C.ST_RESIDUAL, [u, f], [r], C.APPLY_INTERIOR, box))

    # -- queries ---------------------------------------------------------------------------
    def num_chunks(self):
        return self._lib.sg_num_chunks(self._h)

    def chunk_ids(self):
        n = self.num_chunks()
        out = np.empty(n, dtype=np.int64)
        C.check(self._lib.sg_get_chunk_ids(self._h, out.ctypes.data, C.SG_MEM_HOST))
        return out

    def masks(self):
        n = self.num_chunks()
        out = np.empty((n, self.chunk_size), dtype=np.uint8)
        C.check(self._lib.sg_get_masks(self._h, out.ctypes.data, C.SG_MEM_HOST))
        return out

    def prop(self, p):
        n = self.num_chunks()
        out = np.empty((n, self.chunk_size), dtype=self.dtypes[p])
        C.check(self._lib.sg_get_prop(self._h, p, out.ctypes.data, C.SG_MEM_HOST))
        return out

    def set_prop(self, p, data):
        if _is_device(data):
            C.check(self._lib.sg_set_prop(self._h, p, _ptr(data), C.SG_MEM_DEVICE))
            return
        a = np.ascontiguousarray(np.asarray(data, dtype=self.dtypes[p]))
        if a.size != self.num_chunks() * self.chunk_size:
            raise ValueError("expected [num_chunks, chunk_size] values")
        C.check(self._lib.sg_set_prop(self._h, p, a.ctypes.data, C.SG_MEM_HOST))

    def get(self, p, keys):
        """``get<p>(coord)`` for linear keys; background where the point does not exist."""
        k = np.ascontiguousarray(np.asarray(keys, dtype=np.int64).reshape(-1))
        out = np.empty(k.size, dtype=self.dtypes[p])
        C.check(self._lib.sg_get_points(self._h, p, k.size, k.ctypes.data, out.ctypes.data, C.SG_MEM_HOST))
        return out

    def neighbours(self, nn=C.NN_STAR):
        nnn = (2 ** self.dim) if nn == C.NN_STAR else 3 ** self.dim
        out = np.empty((self.num_chunks(), nnn), dtype=np.int32)
        C.check(self._lib.sg_get_neighbours(self._h, nn, out.ctypes.data, C.SG_MEM_HOST))
        return out

    def device_ptr(self, p):
        out = ctypes.c_void_p()
        C.check(self._lib.sg_device_ptr(self._h, p, ctypes.byref(out)))
        return out.value

    def count_existing_elements(self):
        v = ctypes.c_int64()
        C.check(self._lib.sg_count_existing(self._h, ctypes.byref(v)))
        return v.value

    def count_boundary_elements(self):
        v = ctypes.c_int64()
        C.check(self._lib.sg_count_boundary(self._h, ctypes.byref(v)))
        return v.value

    # -- pack / unpack / remove (SparseGridGpu.hpp:3040-3733) ---------------------------
    def _boxes(self, boxes):
        lo = np.ascontiguousarray(np.array([b[0] for b in boxes], np.int32).reshape(-1, self.dim))
        hi = np.ascontiguousarray(np.array([b[1] for b in boxes], np.int32).reshape(-1, self.dim))
        return lo, hi

    def pack_size(self, boxes, props, index_bytes=8):
        """packRequest + packCalculate: bytes of the packed inclusive boxes [(lo, hi), ...]."""
        lo, hi = self._boxes(boxes)
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        out = ctypes.c_int64()
        C.check(self._lib.sg_pack_size(self._h, len(boxes), lo.ctypes.data, hi.ctypes.data, pr.size,
                                       pr.ctypes.data, int(index_bytes), ctypes.byref(out)))
        return out.value

    def pack(self, boxes, props, index_bytes=8, out=None):
        """pack + packFinalize.  Returns a host uint8 array, or fills `out` (a torch uint8
        tensor on the device or a numpy array) and returns the byte count."""
        n = self.pack_size(boxes, props, index_bytes)
        lo, hi = self._boxes(boxes)
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        host = out is None
        buf = np.zeros(max(n, 1), np.uint8) if host else out
        cap = (buf.numel() if _is_device(buf) else buf.size)
        C.check(self._lib.sg_pack(self._h, len(boxes), lo.ctypes.data, hi.ctypes.data, pr.size, pr.ctypes.data,
                                  int(index_bytes), _ptr(buf), cap, _mem(buf)))
        return buf[:n] if host else n

    def unpack(self, buf, dst_lo, props, index_bytes=8, offset=0):
        """unpack one packed box starting at byte `offset` of `buf` (host or device) with its
        points moved to start at dst_lo; returns the bytes the box occupies."""
        lo = np.ascontiguousarray(np.asarray(dst_lo, np.int32))
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        if not _is_device(buf):
            buf = np.ascontiguousarray(np.asarray(buf, np.uint8))
        size = buf.numel() if _is_device(buf) else buf.size
        base = _ptr(buf).value + int(offset)
        used = ctypes.c_int64()
        C.check(self._lib.sg_unpack(self._h, ctypes.c_void_p(base), size - int(offset), _mem(buf), lo.ctypes.data,
                                    pr.size, pr.ctypes.data, int(index_bytes), ctypes.byref(used)))
        return used.value

    def remove_box(self, lo, hi):
        """remove(Box): queue an inclusive box for the next remove_add_unpack_finalize."""
        a = np.ascontiguousarray(np.asarray(lo, np.int32))
        b = np.ascontiguousarray(np.asarray(hi, np.int32))
        C.check(self._lib.sg_remove_box(self._h, a.ctypes.data, b.ctypes.data))

    def remove_points(self, keys):
        """Device remove(coord) for linear point keys (host array or device tensor)."""
        if not _is_device(keys):
            keys = np.ascontiguousarray(np.asarray(keys, np.int64).reshape(-1))
        n = keys.numel() if _is_device(keys) else keys.size
        C.check(self._lib.sg_remove_points(self._h, n, _ptr(keys), _mem(keys)))

    def remove_add_unpack_finalize(self, props):
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        C.check(self._lib.sg_remove_add_unpack_finalize(self._h, pr.size, pr.ctypes.data))

    def remove_add_unpack_reset(self):
        C.check(self._lib.sg_remove_add_unpack_reset(self._h))

    def construct_link(self, other, which, box, p):
        """construct_link_dw (which = 0, `other` finer) / construct_link_up (which = 1,
        `other` coarser) for the padding points in the inclusive box = (lo, hi)."""
        lo = np.ascontiguousarray(np.asarray(box[0], np.int32))
        hi = np.ascontiguousarray(np.asarray(box[1], np.int32))
        pp = np.ascontiguousarray(np.asarray(p, np.int32))
        C.check(self._lib.sg_construct_link(self._h, other._h, int(which), lo.ctypes.data, hi.ctypes.data, pp.ctypes.data))

    def links(self, which):
        """(pad_chunk, pad_off, offsets [n_pad + 1], link_chunk, link_off) of the last
        construct_link of that direction (getDownLinks / getDownLinksOffsets, ...)."""
        npad, nl = ctypes.c_int64(), ctypes.c_int64()
        C.check(self._lib.sg_link_sizes(self._h, int(which), ctypes.byref(npad), ctypes.byref(nl)))
        pc = np.zeros(npad.value, np.int32)
        po = np.zeros(npad.value, np.int16)
        sc = np.zeros(npad.value + 1, np.uint32)
        lc = np.zeros(nl.value, np.int32)
        lo = np.zeros(nl.value, np.int16)
        C.check(self._lib.sg_get_links(self._h, int(which), pc.ctypes.data, po.ctypes.data, sc.ctypes.data,
                                        lc.ctypes.data, lo.ctypes.data, C.SG_MEM_HOST))
        return pc, po, sc, lc, lo

    def copy_to(self, src, box_src, box_dst):
        """copy_to (SparseGridGpu.hpp:3550-3559): queue, on `src`, a copy of the existing
        points of the inclusive box_src = (lo, hi) of `src` into this grid, the box's low
        corner landing at box_dst's low corner.  Done by remove_copy_to_finalize."""
        lo = np.ascontiguousarray(np.asarray(box_src[0], np.int32))
        hi = np.ascontiguousarray(np.asarray(box_src[1], np.int32))
        dl = np.ascontiguousarray(np.asarray(box_dst[0], np.int32))
        C.check(self._lib.sg_copy_to(self._h, src._h, lo.ctypes.data, hi.ctypes.data, dl.ctypes.data))

    def remove_copy_to_finalize(self, props, phase):
        """removeCopyToFinalize (SparseGridGpu.hpp:3314-3329): phase 0 on every grid
        taking part, then phase 1 on every grid, then phase 2 on every grid."""
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        C.check(self._lib.sg_remove_copy_to_finalize(self._h, int(phase), pr.size, pr.ctypes.data))
