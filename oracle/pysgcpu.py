"""ctypes binding of the CPU oracle (oracle/libsgcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg.  Same method names as openfpm_data_amd.SparseGridGpu, host arrays only.
"""
import ctypes
import os
import subprocess

import numpy as np

from openfpm_data_amd import _capi as C

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgcpu.so")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int
_SIGS = {
    "sgcpu_create": (_INT, [_INT, ctypes.POINTER(_I64), _INT, _INT, ctypes.POINTER(ctypes.c_int32), _INT,
                            ctypes.POINTER(_P)]),
    "sgcpu_destroy": (_INT, [_P]),
    "sgcpu_last_error": (ctypes.c_char_p, []),
    "sgcpu_set_background": (_INT, [_P, _INT, _P]),
    "sgcpu_insert_points": (_INT, [_P, _I64, _P, _P, _P]),
    "sgcpu_insert_coords": (_INT, [_P, _I64, _P, _P, _P]),
    "sgcpu_insert_chunks": (_INT, [_P, _I64, _P, _P, _P]),
    "sgcpu_pending_inserts": (_I64, [_P]),
    "sgcpu_flush": (_INT, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "sgcpu_find_neighbours": (_INT, [_P, _INT]),
    "sgcpu_set_nn_type": (_INT, [_P, _INT]),
    "sgcpu_tag_boundaries": (_INT, [_P, _INT, _INT]),
    "sgcpu_apply_stencil": (_INT, [_P, ctypes.POINTER(C.StencilDesc)]),
    "sgcpu_num_chunks": (_I64, [_P]),
    "sgcpu_chunk_size": (_INT, [_P]),
    "sgcpu_get_chunk_ids": (_INT, [_P, _P]),
    "sgcpu_get_masks": (_INT, [_P, _P]),
    "sgcpu_get_prop": (_INT, [_P, _INT, _P]),
    "sgcpu_set_prop": (_INT, [_P, _INT, _P]),
    "sgcpu_get_points": (_INT, [_P, _INT, _I64, _P, _P]),
    "sgcpu_get_neighbours": (_INT, [_P, _INT, _P]),
    "sgcpu_count_existing": (_INT, [_P, ctypes.POINTER(_I64)]),
    "sgcpu_count_boundary": (_INT, [_P, ctypes.POINTER(_I64)]),
    "sgcpu_lin_id": (_I64, [_P, ctypes.POINTER(_I64)]),
    "sgcpu_inv_lin_id": (_INT, [_P, _I64, ctypes.POINTER(_I64)]),
    "sgcpu_set_threads": (_INT, [_INT]),
    "sgcpu_set_linearizer": (_INT, [_P, _INT]),
    "sgcpu_morton": (ctypes.c_uint64, [_INT, ctypes.POINTER(_I64)]),
    "sgcpu_pack_size": (_INT, [_P, _INT, _P, _P, _INT, _P, _INT, ctypes.POINTER(_I64)]),
    "sgcpu_pack": (_INT, [_P, _INT, _P, _P, _INT, _P, _INT, _P, _I64]),
    "sgcpu_unpack": (_INT, [_P, _P, _I64, _P, _INT, _P, _INT, ctypes.POINTER(_I64)]),
    "sgcpu_remove_box": (_INT, [_P, _P, _P]),
    "sgcpu_remove_points": (_INT, [_P, _I64, _P]),
    "sgcpu_remove_add_unpack_finalize": (_INT, [_P, _INT, _P]),
    "sgcpu_remove_add_unpack_reset": (_INT, [_P]),
    "sgcpu_copy_to": (_INT, [_P, _P, _P, _P, _P]),
    "sgcpu_construct_link": (_INT, [_P, _P, _INT, _P, _P, _P]),
    "sgcpu_link_sizes": (_INT, [_P, _INT, _P, _P]),
    "sgcpu_get_links": (_INT, [_P, _INT, _P, _P, _P, _P, _P]),
    "sgcpu_remove_copy_to_finalize": (_INT, [_P, _INT, _INT, _P]),
}
_lib = None


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def set_threads(n):
    load().sgcpu_set_threads(int(n))


def _check(rc):
    if rc != 0:
        raise RuntimeError("sgcpu error %d: %s" % (rc, load().sgcpu_last_error().decode()))


def _boxes(boxes, dim):
    lo = np.ascontiguousarray(np.array([b[0] for b in boxes], np.int32).reshape(-1, dim))
    hi = np.ascontiguousarray(np.array([b[1] for b in boxes], np.int32).reshape(-1, dim))
    return lo, hi


class SparseGridCpu:
    def __init__(self, dim, size, props, block_edge=8, stencil_radius=1, linearizer="smb"):
        self._lib = load()
        self.dim = int(dim)
        self.size = tuple(int(s) for s in size)
        self.dtypes = [np.dtype(p) for p in props]
        self.block_edge = int(block_edge)
        self.chunk_size = self.block_edge ** self.dim
        self.block_sizes = tuple(-(-s // self.block_edge) for s in self.size)
        sz = (ctypes.c_int64 * 3)(*(list(self.size) + [1] * (3 - self.dim)))
        codes = (ctypes.c_int32 * max(1, len(self.dtypes)))(*[C.DTYPE_CODES[d] for d in self.dtypes])
        h = ctypes.c_void_p()
        _check(self._lib.sgcpu_create(self.dim, sz, self.block_edge, len(self.dtypes), codes, int(stencil_radius),
                                      ctypes.byref(h)))
        self._h = h
        self.linearizer = linearizer
        if linearizer != "smb":
            _check(self._lib.sgcpu_set_linearizer(h, {"smb": 0, "zmb": 1}[linearizer]))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sgcpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lin_id(self, coords):
        """grid_smb::LinId through the oracle's C restatement."""
        c = np.asarray(coords, dtype=np.int64).reshape(-1, self.dim)
        out = np.empty(c.shape[0], dtype=np.int64)
        buf = (ctypes.c_int64 * 3)()
        for i in range(c.shape[0]):
            for d in range(self.dim):
                buf[d] = int(c[i, d])
            out[i] = self._lib.sgcpu_lin_id(self._h, buf)
        return out

    def inv_lin_id(self, keys):
        """grid_smb::InvLinId (numpy form of the oracle's restatement); grid_zmb through the
        oracle's C restatement."""
        k = np.asarray(keys, dtype=np.int64).reshape(-1)
        if self.linearizer != "smb":
            out = np.empty((k.size, self.dim), dtype=np.int64)
            buf = (ctypes.c_int64 * 3)()
            for i in range(k.size):
                _check(self._lib.sgcpu_inv_lin_id(self._h, int(k[i]), buf))
                out[i] = buf[:self.dim]
            return out
        b = k // self.chunk_size
        l = k % self.chunk_size
        out = np.empty((k.size, self.dim), dtype=np.int64)
        for d in range(self.dim):
            out[:, d] = (b % self.block_sizes[d]) * self.block_edge + l % self.block_edge
            b //= self.block_sizes[d]
            l //= self.block_edge
        return out

    def set_background_value(self, p, value):
        v = np.array([value], dtype=self.dtypes[p])
        _check(self._lib.sgcpu_set_background(self._h, p, v.ctypes.data))

    def _values(self, values, n, per=1):
        if values is None:
            return None, []
        if not isinstance(values, (list, tuple)):
            values = [values]
        keep = []
        arr = (ctypes.c_void_p * len(values))()
        for p, v in enumerate(values):
            if v is None:
                arr[p] = None
                continue
            v = np.ascontiguousarray(np.asarray(v, dtype=self.dtypes[p]).reshape(-1))
            assert v.size == n * per
            keep.append(v)
            arr[p] = v.ctypes.data
        return arr, keep

    def insert(self, keys, values=None, masks=None):
        k = np.ascontiguousarray(np.asarray(keys, dtype=np.int64).reshape(-1))
        m = None if masks is None else np.ascontiguousarray(np.asarray(masks, dtype=np.uint8).reshape(-1))
        vals, keep = self._values(values, k.size)
        _check(self._lib.sgcpu_insert_points(self._h, k.size, k.ctypes.data,
                                             None if m is None else m.ctypes.data, vals))

    def insert_coords(self, coords, values=None, masks=None):
        c = np.ascontiguousarray(np.asarray(coords, dtype=np.int32).reshape(-1, self.dim))
        m = None if masks is None else np.ascontiguousarray(np.asarray(masks, dtype=np.uint8).reshape(-1))
        vals, keep = self._values(values, c.shape[0])
        _check(self._lib.sgcpu_insert_coords(self._h, c.shape[0], c.ctypes.data,
                                             None if m is None else m.ctypes.data, vals))

    def insert_block(self, chunk_ids, masks, values=None):
        ids = np.ascontiguousarray(np.asarray(chunk_ids, dtype=np.int64).reshape(-1))
        m = np.ascontiguousarray(np.asarray(masks, dtype=np.uint8).reshape(-1))
        vals, keep = self._values(values, ids.size, self.chunk_size)
        _check(self._lib.sgcpu_insert_chunks(self._h, ids.size, ids.ctypes.data, m.ctypes.data, vals))

    def pending_inserts(self):
        return self._lib.sgcpu_pending_inserts(self._h)

    def flush(self, ops):
        codes = [C.OPS[o] if isinstance(o, str) else int(o) for o in ops]
        arr = (ctypes.c_int32 * max(1, len(codes)))(*codes)
        _check(self._lib.sgcpu_flush(self._h, arr))

    def find_neighbours(self, nn=C.NN_STAR):
        _check(self._lib.sgcpu_find_neighbours(self._h, nn))

    def set_nn_type(self, nn):
        _check(self._lib.sgcpu_set_nn_type(self._h, nn))

    def tag_boundaries(self, nn=C.NN_STAR, opt=C.NO_CALCULATE_EXISTING_POINTS):
        _check(self._lib.sgcpu_tag_boundaries(self._h, nn, opt))

    def apply_stencil(self, desc):
        _check(self._lib.sgcpu_apply_stencil(self._h, ctypes.byref(desc)))

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

    def num_chunks(self):
        return self._lib.sgcpu_num_chunks(self._h)

    def chunk_ids(self):
        out = np.empty(self.num_chunks(), dtype=np.int64)
        _check(self._lib.sgcpu_get_chunk_ids(self._h, out.ctypes.data))
        return out

    def masks(self):
        out = np.empty((self.num_chunks(), self.chunk_size), dtype=np.uint8)
        _check(self._lib.sgcpu_get_masks(self._h, out.ctypes.data))
        return out

    def prop(self, p):
        out = np.empty((self.num_chunks(), self.chunk_size), dtype=self.dtypes[p])
        _check(self._lib.sgcpu_get_prop(self._h, p, out.ctypes.data))
        return out

    def set_prop(self, p, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=self.dtypes[p]))
        _check(self._lib.sgcpu_set_prop(self._h, p, a.ctypes.data))

    def get(self, p, keys):
        k = np.ascontiguousarray(np.asarray(keys, dtype=np.int64).reshape(-1))
        out = np.empty(k.size, dtype=self.dtypes[p])
        _check(self._lib.sgcpu_get_points(self._h, p, k.size, k.ctypes.data, out.ctypes.data))
        return out

    def neighbours(self, nn=C.NN_STAR):
        nnn = (2 ** self.dim) if nn == C.NN_STAR else 3 ** self.dim
        out = np.empty((self.num_chunks(), nnn), dtype=np.int32)
        _check(self._lib.sgcpu_get_neighbours(self._h, nn, out.ctypes.data))
        return out

    def count_existing_elements(self):
        v = ctypes.c_int64()
        _check(self._lib.sgcpu_count_existing(self._h, ctypes.byref(v)))
        return v.value

    def count_boundary_elements(self):
        v = ctypes.c_int64()
        _check(self._lib.sgcpu_count_boundary(self._h, ctypes.byref(v)))
        return v.value

    # pack / unpack / remove (SparseGridGpu.hpp:3040-3733)
    def pack_size(self, boxes, props, index_bytes=8):
        lo, hi = _boxes(boxes, self.dim)
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        out = ctypes.c_int64()
        _check(self._lib.sgcpu_pack_size(self._h, len(boxes), lo.ctypes.data, hi.ctypes.data, pr.size, pr.ctypes.data,
                                         index_bytes, ctypes.byref(out)))
        return out.value

    def pack(self, boxes, props, index_bytes=8):
        n = self.pack_size(boxes, props, index_bytes)
        lo, hi = _boxes(boxes, self.dim)
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        buf = np.zeros(max(n, 1), np.uint8)
        _check(self._lib.sgcpu_pack(self._h, len(boxes), lo.ctypes.data, hi.ctypes.data, pr.size, pr.ctypes.data,
                                    index_bytes, buf.ctypes.data, n))
        return buf[:n]

    def unpack(self, buf, dst_lo, props, index_bytes=8, offset=0):
        b = np.ascontiguousarray(np.asarray(buf, np.uint8))
        lo = np.ascontiguousarray(np.asarray(dst_lo, np.int32))
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        used = ctypes.c_int64()
        _check(self._lib.sgcpu_unpack(self._h, b.ctypes.data + offset, b.size - offset, lo.ctypes.data, pr.size,
                                      pr.ctypes.data, index_bytes, ctypes.byref(used)))
        return used.value

    def remove_box(self, lo, hi):
        a = np.ascontiguousarray(np.asarray(lo, np.int32))
        b = np.ascontiguousarray(np.asarray(hi, np.int32))
        _check(self._lib.sgcpu_remove_box(self._h, a.ctypes.data, b.ctypes.data))

    def remove_points(self, keys):
        k = np.ascontiguousarray(np.asarray(keys, np.int64).reshape(-1))
        _check(self._lib.sgcpu_remove_points(self._h, k.size, k.ctypes.data))

    def remove_add_unpack_finalize(self, props):
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        _check(self._lib.sgcpu_remove_add_unpack_finalize(self._h, pr.size, pr.ctypes.data))

    def remove_add_unpack_reset(self):
        _check(self._lib.sgcpu_remove_add_unpack_reset(self._h))

    def construct_link(self, other, which, box, p):
        """construct_link_dw (which = 0, `other` finer) / construct_link_up (which = 1,
        `other` coarser) for the padding points in the inclusive box = (lo, hi)."""
        lo = np.ascontiguousarray(np.asarray(box[0], np.int32))
        hi = np.ascontiguousarray(np.asarray(box[1], np.int32))
        pp = np.ascontiguousarray(np.asarray(p, np.int32))
        _check(self._lib.sgcpu_construct_link(self._h, other._h, int(which), lo.ctypes.data, hi.ctypes.data, pp.ctypes.data))

    def links(self, which):
        """(pad_chunk, pad_off, offsets [n_pad + 1], link_chunk, link_off) of the last
        construct_link of that direction (getDownLinks / getDownLinksOffsets, ...)."""
        npad, nl = ctypes.c_int64(), ctypes.c_int64()
        _check(self._lib.sgcpu_link_sizes(self._h, int(which), ctypes.byref(npad), ctypes.byref(nl)))
        pc = np.zeros(npad.value, np.int32)
        po = np.zeros(npad.value, np.int16)
        sc = np.zeros(npad.value + 1, np.uint32)
        lc = np.zeros(nl.value, np.int32)
        lo = np.zeros(nl.value, np.int16)
        _check(self._lib.sgcpu_get_links(self._h, int(which), pc.ctypes.data, po.ctypes.data, sc.ctypes.data,
                                        lc.ctypes.data, lo.ctypes.data))
        return pc, po, sc, lc, lo

    def copy_to(self, src, box_src, box_dst):
        lo = np.ascontiguousarray(np.asarray(box_src[0], np.int32))
        hi = np.ascontiguousarray(np.asarray(box_src[1], np.int32))
        dl = np.ascontiguousarray(np.asarray(box_dst[0], np.int32))
        _check(self._lib.sgcpu_copy_to(self._h, src._h, lo.ctypes.data, hi.ctypes.data, dl.ctypes.data))

    def remove_copy_to_finalize(self, props, phase):
        pr = np.ascontiguousarray(np.asarray(props, np.int32))
        _check(self._lib.sgcpu_remove_copy_to_finalize(self._h, int(phase), pr.size, pr.ctypes.data))


def morton(dim, coord):
    """lin_zid of one chunk coordinate (the oracle's restatement of util/zmorton.hpp)."""
    c = (ctypes.c_int64 * 3)(*(list(int(v) for v in coord) + [0] * (3 - dim)))
    return int(load().sgcpu_morton(dim, c))
