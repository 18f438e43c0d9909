"""ctypes binding of libsg_hip.so (the C ABI declared in include/sparsegrid_gpu.h).

The library is built in-tree by ``openfpm_data_amd/csrc/Makefile`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsg_hip.so")

SG_OK = 0
SG_E_ARG, SG_E_HIP, SG_E_OOM, SG_E_STATE, SG_E_RANGE = -1, -2, -3, -4, -5
SG_MEM_HOST = 0
SG_MEM_DEVICE = 1

# dtypes (sg_dtype)
SG_F32, SG_F64, SG_I32, SG_U32, SG_I64, SG_U64, SG_I8, SG_U8, SG_I16, SG_U16 = range(10)
DTYPE_CODES = {
    np.dtype(np.float32): SG_F32,
    np.dtype(np.float64): SG_F64,
    np.dtype(np.int32): SG_I32,
    np.dtype(np.uint32): SG_U32,
    np.dtype(np.int64): SG_I64,
    np.dtype(np.uint64): SG_U64,
    np.dtype(np.int8): SG_I8,
    np.dtype(np.uint8): SG_U8,
    np.dtype(np.int16): SG_I16,
    np.dtype(np.uint16): SG_U16,
}
CODE_DTYPES = {v: k for k, v in DTYPE_CODES.items()}

# reductions (sg_op)
OP_NONE, OP_MAX, OP_MIN, OP_ADD, OP_RIGHT, OP_LEFT, OP_OR = range(7)
OPS = {"none": OP_NONE, "smax": OP_MAX, "smin": OP_MIN, "sadd": OP_ADD,
       "sright": OP_RIGHT, "sleft": OP_LEFT, "sor": OP_OR}

TUNE_STENCIL_WG_PER_XCD = 1
TUNE_STENCIL_VARIANT = 2
TUNE_STENCIL_STORES = 3
TUNE_STENCIL_SKIP_EMPTY_ROWS = 4
TUNE_TAG_CHUNKS_PER_WAVE = 7
TUNE_FLUSH_PATH = 8
TUNE_TAG_VARIANT = 9
TUNE_BOX_PAIRS = 10
TUNE_BATCH_WAVES = 5
TUNE_BATCH_MAX_WGS = 6
COPY_NONTEMPORAL = 1
# sg_pack_noalloc limits (static per-GPU scratch)
PACK_NOALLOC_SCRATCH_BYTES = 32 << 20
PACK_NOALLOC_MAX_BOXES = 64
LINK_DOWN = 0
LINK_UP = 1

EXIST = 1
PADDING = 2
NN_STAR = 0
NN_FULL = 1
NO_CALCULATE_EXISTING_POINTS = 0
CALCULATE_EXISTING_POINTS = 1

# stencil kinds (sg_stencil_kind) and apply modes
ST_LAP, ST_HEAT, ST_GRAY_SCOTT, ST_BOX, ST_RBSOR, ST_RESIDUAL = range(6)
STENCIL_MODE_INPLACE = 1
STENCIL_MODE_INPLACE_NO_SHARED = 3
APPLY_INTERIOR = 0
APPLY_EXISTING = 1


class StencilDesc(ctypes.Structure):
    """sg_stencil_desc (include/sparsegrid_gpu.h)."""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("apply", ctypes.c_int32),
        ("src", ctypes.c_int32 * 4),
        ("dst", ctypes.c_int32 * 4),
        ("use_box", ctypes.c_int32),
        ("box_lo", ctypes.c_int32 * 3),
        ("box_hi", ctypes.c_int32 * 3),
        ("mode", ctypes.c_int32),
        ("coef", ctypes.c_double * 27),
        ("icoef", ctypes.c_int64 * 27),
    ]


def make_desc(kind, src, dst=(), apply=APPLY_INTERIOR, box=None, coef=(), icoef=(), mode=STENCIL_MODE_INPLACE):
    d = StencilDesc()
    d.kind = kind
    d.apply = apply
    d.mode = mode
    for i, s in enumerate(src):
        d.src[i] = s
    for i, s in enumerate(dst):
        d.dst[i] = s
    if box is not None:
        d.use_box = 1
        lo, hi = box
        for i in range(len(lo)):
            d.box_lo[i] = int(lo[i])
            d.box_hi[i] = int(hi[i])
    for i, c in enumerate(coef):
        d.coef[i] = float(c)
    for i, c in enumerate(icoef):
        d.icoef[i] = int(c)
    return d


POISSON_MIXED = 0
POISSON_PLAIN_FP64 = 1


class PoissonParams(ctypes.Structure):
    """sg_poisson_params (include/sparsegrid_gpu.h)."""
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("u_prop", ctypes.c_int32), ("u2_prop", ctypes.c_int32), ("f_prop", ctypes.c_int32),
        ("e_prop", ctypes.c_int32), ("r_prop", ctypes.c_int32),
        ("tol", ctypes.c_double),
        ("omega", ctypes.c_double),
        ("omega0", ctypes.c_double),
        ("inner_target", ctypes.c_double),
        ("first_sweeps", ctypes.c_int32),
        ("max_sweeps", ctypes.c_int32),
        ("max_outer", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
    ]


class PoissonResult(ctypes.Structure):
    """sg_poisson_result (include/sparsegrid_gpu.h)."""
    _fields_ = [
        ("converged", ctypes.c_int32),
        ("outer", ctypes.c_int32),
        ("sweeps", ctypes.c_int32),
        ("solution_prop", ctypes.c_int32),
        ("rel_residual", ctypes.c_double),
        ("omega", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_INT = ctypes.c_int

_SIGS = {
    "sg_create": (_INT, [_INT, ctypes.POINTER(_I64), _INT, _INT, ctypes.POINTER(_I32), _INT, ctypes.POINTER(_P)]),
    "sg_destroy": (_INT, [_P]),
    "sg_set_linearizer": (_INT, [_P, _INT]),
    "sg_last_error": (ctypes.c_char_p, []),
    "sg_set_stream": (_INT, [_P, _P]),
    "sg_synchronize": (_INT, [_P]),
    "sg_set_background": (_INT, [_P, _INT, _P]),
    "sg_insert_points": (_INT, [_P, _I64, _P, _P, _P, _INT]),
    "sg_insert_coords": (_INT, [_P, _I64, _P, _P, _P, _INT]),
    "sg_insert_chunks": (_INT, [_P, _I64, _P, _P, _P, _INT]),
    "sg_pending_inserts": (_I64, [_P]),
    "sg_flush": (_INT, [_P, ctypes.POINTER(_I32)]),
    "sg_find_neighbours": (_INT, [_P, _INT]),
    "sg_set_nn_type": (_INT, [_P, _INT]),
    "sg_tag_boundaries": (_INT, [_P, _INT, _INT]),
    "sg_apply_stencil": (_INT, [_P, ctypes.POINTER(StencilDesc)]),
    "sg_existing_points": (_INT, [_P, ctypes.POINTER(_I64)]),
    "sg_get_existing_points": (_INT, [_P, _P, _INT]),
    "sg_set_tuning": (_INT, [_P, _INT, _I64]),
    "sg_num_chunks": (_I64, [_P]),
    "sg_chunk_size": (_INT, [_P]),
    "sg_get_chunk_ids": (_INT, [_P, _P, _INT]),
    "sg_get_masks": (_INT, [_P, _P, _INT]),
    "sg_get_prop": (_INT, [_P, _INT, _P, _INT]),
    "sg_set_prop": (_INT, [_P, _INT, _P, _INT]),
    "sg_get_points": (_INT, [_P, _INT, _I64, _P, _P, _INT]),
    "sg_set_chunks": (_INT, [_P, _I64, _P, _P, _INT]),
    "sg_restore_background": (_INT, [_P, _INT]),
    "sg_get_neighbours": (_INT, [_P, _INT, _P, _INT]),
    "sg_device_ptr": (_INT, [_P, _INT, ctypes.POINTER(_P)]),
    "sg_to_kernel": (_INT, [_P, _INT, _P]),
    "sg_insert_device_begin": (_INT, [_P, _I64, _P]),
    "sg_insert_device_end": (_INT, [_P]),
    "sg_count_existing": (_INT, [_P, ctypes.POINTER(_I64)]),
    "sg_count_boundary": (_INT, [_P, ctypes.POINTER(_I64)]),
    "sg_pack_size": (_INT, [_P, _INT, _P, _P, _INT, _P, _INT, ctypes.POINTER(_I64)]),
    "sg_pack": (_INT, [_P, _INT, _P, _P, _INT, _P, _INT, _P, _I64, _INT]),
    "sg_pack_noalloc": (_INT, [_P, _INT, _P, _P, _INT, _P, _INT, _P, _I64, ctypes.POINTER(_I64)]),
    "sg_unpack": (_INT, [_P, _P, _I64, _INT, _P, _INT, _P, _INT, ctypes.POINTER(_I64)]),
    "sg_remove_box": (_INT, [_P, _P, _P]),
    "sg_remove_points": (_INT, [_P, _I64, _P, _INT]),
    "sg_remove_add_unpack_finalize": (_INT, [_P, _INT, _P]),
    "sg_remove_add_unpack_reset": (_INT, [_P]),
    "sg_copy_to": (_INT, [_P, _P, _P, _P, _P]),
    "sg_construct_link": (_INT, [_P, _P, _INT, _P, _P, _P]),
    "sg_link_sizes": (_INT, [_P, _INT, _P, _P]),
    "sg_get_links": (_INT, [_P, _INT, _P, _P, _P, _P, _P, _INT]),
    "sg_remove_copy_to_finalize": (_INT, [_P, _INT, _INT, _P]),
    "sg_trim": (_INT, [_P]),
    "sg_stream_copy": (_INT, [_P, _P, _I64, _INT, _P]),
    "sg_device_bytes": (_INT, [_P, ctypes.POINTER(_I64)]),
    "sg_batch_create": (_INT, [ctypes.POINTER(PoissonParams), _P, ctypes.POINTER(_P)]),
    "sg_batch_destroy": (_INT, [_P]),
    "sg_batch_add": (_INT, [_P, _P, _I64, ctypes.c_double]),
    "sg_batch_live": (_INT, [_P]),
    "sg_batch_set_tuning": (_INT, [_P, _INT, _I64]),
    "sg_batch_step": (_INT, [_P]),
    "sg_batch_poll": (_INT, [_P, ctypes.POINTER(_I64), ctypes.POINTER(PoissonResult), _INT, ctypes.POINTER(_INT)]),
    "sg_poisson_solve": (_INT, [ctypes.POINTER(_P), _INT, ctypes.POINTER(PoissonParams), ctypes.POINTER(PoissonResult)]),
}

_lib = None


class KernelView(ctypes.Structure):
    """sg_kernel_view (the grid's device arrays, filled by sg_to_kernel)."""
    _fields_ = [("keys", _P), ("mask", _P), ("prop", _P * 16), ("nn", _P), ("stream", _P),
                ("nchunks", _I64), ("size", _I64 * 3), ("bsz", _I64 * 3),
                ("nnn", ctypes.c_int32), ("dim", ctypes.c_int32), ("edge", ctypes.c_int32),
                ("chunk_size", ctypes.c_int32), ("zorder", ctypes.c_int32), ("nprops", ctypes.c_int32)]


def exported_symbols():
    """Names of the entry points this binding expects the library to export."""
    return sorted(_SIGS)


def load():
    """Load libsg_hip.so once; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libsg_hip.so is not built (expected %s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SgError(RuntimeError):
    pass


def last_error():
    """The calling thread's last error message of the library."""
    msg = load().sg_last_error()
    return msg.decode() if msg else ""


def check(rc):
    if rc != SG_OK:
        raise SgError("libsg_hip error %d: %s" % (rc, last_error()))
    return rc
