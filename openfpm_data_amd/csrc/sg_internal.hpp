// sg_internal.hpp -- shared definitions of libsg_hip.so (host side + device helpers).
//
// Device layout of one grid (chunk-major SoA, the reference's memory_traits_inte
// DataBlock arrays, src/SparseGridGpu/BlockMapGpu.hpp:63-69):
//   keys  int64 [cap]                 sorted chunk ids (map_vector_sparse.hpp:790-791)
//   prop  T     [(cap+1) * csz]       one array per property, chunk c at c*csz
//   mask  uint8 [(cap+1) * csz]       EXIST bit 0, PADDING bit 1
//   chunk nChunks of every array is the background chunk (mask 0, values = background)
// Invariant kept by every kernel of this library: a slot whose mask byte is 0 holds
// the property's background value.  The reference substitutes the background for
// such slots whenever it reads them (loadGhostBlock_impl, SparseGridGpu_ker_util.hpp:551-552;
// BlockMapGpu_ker::get, BlockMapGpu_ker.cuh:119-135), so results are identical while
// the stencil kernels read neighbour data without reading neighbour masks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/sparsegrid_gpu.h"

namespace sg {

constexpr int kMaxProps = 16;

// ---- errors ------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define SG_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::sg::fail(SG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline size_t dtype_size(int dt)
{
    switch (dt) {
    case SG_F32: case SG_I32: case SG_U32: return 4;
    case SG_F64: case SG_I64: case SG_U64: return 8;
    case SG_I8: case SG_U8: return 1;
    case SG_I16: case SG_U16: return 2;
    default: return 0;
    }
}
inline bool dtype_is_float(int dt) { return dt == SG_F32 || dt == SG_F64; }

// ---- a per-handle caching device allocator -------------------------------------
// hipMalloc/hipFree synchronise the device; a grid reallocates its arrays on
// every flush that adds chunks, so freed blocks are kept and reused.  All work of a
// handle is ordered on the handle's stream, so reuse by the same handle is safe.
// When a handle is destroyed its blocks go to a process-wide cache of idle blocks (per
// device), from which later handles allocate before calling hipMalloc: creating and
// destroying grids over and over then costs no hipMalloc.  release() may only be called
// once the handle's stream is idle (its owners synchronise the stream first), since the
// blocks are then reused by other streams without further ordering.
class DevicePool {
public:
    ~DevicePool() { free_all(); }
    void *alloc(size_t bytes);
    void free(void *p);
    void release();   // every block of this handle to the process-wide idle cache
    void trim();      // hipFree this handle's cached blocks and the idle cache of the device
    size_t live_bytes() const { return live_; }
    size_t held_bytes() const;

private:
    void free_all();  // hipFree every block of this handle
    std::multimap<size_t, void *> cache_;
    std::map<void *, size_t> sizes_;
    std::map<void *, int> devs_;
    size_t live_ = 0;
};

// a device buffer owned by a pool, grown on demand
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};
int ensure(DevicePool &pool, DevBuf &b, size_t bytes);

}  // namespace sg

struct sg_grid {
    int dim = 3, edge = 8, csz = 512, r = 1;
    int64_t sz[3] = {1, 1, 1}, bsz[3] = {1, 1, 1};
    int np = 0;
    int dt[sg::kMaxProps];
    int ts[sg::kMaxProps];
    unsigned char bg[sg::kMaxProps][8];

    int64_t nchunks = 0, cap = 0;
    int64_t *d_keys = nullptr;
    void *d_prop[sg::kMaxProps] = {};
    uint8_t *d_mask = nullptr;

    // neighbour tables: NNStar [n][2^dim], NNFull [n][3^dim], int32 positions
    int32_t *d_nn[2] = {nullptr, nullptr};
    int64_t nn_cap[2] = {0, 0};
    bool nn_valid[2] = {false, false};
    int ghost_type = SG_NN_STAR;

    // insert buffer (point records)
    int64_t ins_n = 0, ins_cap = 0;
    int64_t *d_ins_key = nullptr;
    uint8_t *d_ins_mask = nullptr;
    void *d_ins_val[sg::kMaxProps] = {};
    // point records of device-coordinate inserts staged as the grouped flush's (chunk id,
    // record) pairs, while no other insert is pending: [n] chunk ids, [n] records of
    // 16 * pins_nv bytes, and a flag set when a coordinate was outside the grid
    int64_t pins_n = 0, pins_cap = 0;
    int pins_nv = 0;
    uint32_t *d_pins_ck = nullptr;
    void *d_pins_rec = nullptr;
    int *d_pins_bad = nullptr;
    // insert buffer of whole-chunk records (insertBlock), kept as chunks while no point
    // record is pending: [n] chunk ids, [n][csz] masks and values
    int64_t cins_n = 0, cins_cap = 0;
    int64_t *d_cins_id = nullptr;
    uint8_t *d_cins_mask = nullptr;
    void *d_cins_val[sg::kMaxProps] = {};

    // launch shape of the persistent 8^3 star-stencil kernel: workgroups (4 waves)
    // per XCD; set through sg_set_tuning (the autotuner picks it per workload)
    int64_t fast_wg_per_xcd = 160;
    // 0: one chunk per 64-lane workgroup; 1: persistent + pipelined; 2: persistent +
    // pipelined with registers capped for 4 waves per SIMD; 3 / 4 / 5: workgroups of
    // 4 / 8 / 2 consecutive chunks sharing x faces in LDS
    int fast_variant = 3;
    // stores of the 8^3 star-stencil kernel: 0 active points only, 1 whole rows where
    // every slot is active or empty, 2 as 1 with non-temporal stores, 3 as 2 through LDS
    // as lane-contiguous 16-byte vectors
    int fast_wmode = 3;
    // 8^3 tagBoundaries: chunks per wave (1, 2 or 4), all loaded before any is tagged
    int tag_cpw = 2;
    // flush of point records: 0 grouped by chunk where the records and chunk ids allow it,
    // 1 always by point key
    int flush_path = 0;
    // chunk records written by the caller's kernels (sg_insert_device_begin / _end):
    // [count (256 B)] [ids, cap + 1] [masks] [values per property], the last record a
    // sink for claims beyond the capacity
    sg::DevBuf dins;
    int64_t dins_cap = 0;
    bool dins_open = false;
    // 8^3 star stencils: 0 load every row; 1 read the mask first and skip loading 8-point
    // rows without points; 2 / 3 skip the own rows (2 only) and ghost rows without points
    // as the occupancy table says; 4 as 2 with each row's mask taken from the mask planes
    int fast_lskip = 0;
    // occupancy table of the 8^3 star stencils, [n][16] int32 per chunk: the 6 star
    // neighbour positions, the chunk's row occupancy (64 bits: row z*8+y has a slot with a
    // non-zero mask), the occupancy of the 32 face ghost rows it reads (32 bits), and the
    // row occupancy of its -x and +x neighbours; rebuilt when stale (chunk set or masks
    // changed)
    int32_t *d_occ_tab = nullptr;
    int64_t occ_tab_cap = 0;
    bool occ_tab_valid = false;
    // mask planes of the 8^3 star stencils' skip mode 4, [n + 1][16] uint64 (words 0-7: mask
    // byte non-zero, 8-15: EXIST without PADDING; word z bit x + 8 y), built with the
    // occupancy table and valid with it
    uint64_t *d_mplanes = nullptr;
    int64_t mplanes_cap = 0;
    // EXIST bit-planes of 8^3 chunks, [n + 1][8] uint64 (word z: bit x + 8 y is the EXIST
    // bit of point (x, y, z)), the background chunk's zero; rebuilt when stale (masks or
    // the chunk set may have changed; tagBoundaries changes only PADDING bits)
    uint64_t *d_ebits = nullptr;
    int64_t ebits_cap = 0;
    bool ebits_valid = false;
    // 8^3 tagBoundaries: 0 on the EXIST bit-planes (built first when stale), 1 on the
    // neighbours' mask rows
    int tag_variant = 0;
    // 8^3 27-point box: two consecutive chunks per wave, x ghosts between them shared
    int box_pairs = 1;

    // existing points (data positions chunk * csz + offset) for the point-list stencil
    // mode; rebuilt when stale (tagBoundaries with CALCULATE_EXISTING_POINTS builds it)
    int64_t *d_ep = nullptr;
    int64_t ep_n = 0, ep_cap = 0;
    bool ep_valid = false;

    // chunk linearizer: SG_LIN_SMB (row-major) or SG_LIN_ZMB (Morton)
    int lin = SG_LIN_SMB;

    // boxes queued by sg_remove_box, [k][dim] inclusive corners
    std::vector<int32_t> rem_lo, rem_hi;
    // links to the finer (SG_LINK_DOWN) and coarser (SG_LINK_UP) level built by
    // sg_construct_link: padding points of this grid, per padding point the first link
    // (n_pad + 1 entries), and the linked (chunk position, offset) of the other grid
    struct link_set {
        int32_t *pad_chunk = nullptr;
        int16_t *pad_off = nullptr;
        uint32_t *scan = nullptr;
        int32_t *link_chunk = nullptr;
        int16_t *link_off = nullptr;
        int64_t n_pad = 0, n_links = 0;
    };
    link_set links[2];

    // copies queued on this grid as the source by sg_copy_to: the points of [lo, hi]
    // go to `dst` with lo at dst_lo
    struct copy_section {
        sg_grid *dst;
        int32_t lo[3], hi[3], dst_lo[3];
    };
    std::vector<copy_section> copies;

    hipStream_t stream = nullptr;
    sg::DevicePool pool;
    sg::DevBuf scratch[14];  // flush / sort / scan temporaries
    sg::DevBuf hostlike;     // small device staging
    // pinned host words for device -> host copies of flags and totals: queued without
    // blocking the host, read after the next stream synchronisation
    uint32_t *h_words = nullptr;
    ~sg_grid()
    {
        if (h_words) (void)hipHostFree(h_words);
    }

    int nnn(int type) const { return type == SG_NN_STAR ? (1 << dim) : (dim == 2 ? 9 : 27); }
};

namespace sg {

// ---- chunk write-back --------------------------------------------------------------
// 16 bytes from LDS to global memory, non-temporal (raw bits)
template <typename T> __device__ inline void copy16_nt(T *dst, const T *lds)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u a = *(const v4u *)lds;
    __builtin_nontemporal_store(a, (v4u *)dst);
}

// Write back one property of an 8^3 chunk from a wave's LDS staging area `lds` (row r of
// the chunk at lds + 8 r, written there by the lane owning the row) as lane-contiguous
// 16-byte non-temporal stores, 1 KB per instruction.  fl is the calling lane's row flag:
// 0 nothing to store, 1 whole row, 2 active points only (bits 8-15: which); the lanes
// storing a row read its flag by a cross-lane shuffle.
template <typename T> __device__ inline void store_chunk_contig(T *dst, const T *lds, uint32_t fl, int lane)
{
    constexpr int VPR = (int)sizeof(T) / 2;   // 16-byte vectors per 8-point row
    constexpr int PPV = 16 / (int)sizeof(T);  // points per vector
#pragma unroll
    for (int k = 0; k < VPR; ++k) {
        const int vi = k * 64 + lane, row = vi / VPR, part = vi % VPR;
        const uint32_t f = (uint32_t)__shfl((int)fl, row, 64);
        const int off = row * 8 + part * PPV;
        if ((f & 3) == 1) {
            copy16_nt(dst + off, lds + off);
        } else if (f & 2) {
#pragma unroll
            for (int x = 0; x < PPV; ++x)
                if ((f >> (8 + part * PPV + x)) & 1u) dst[off + x] = lds[off + x];
        }
    }
}

// ---- chunk linearizers -----------------------------------------------------------
// grid_smb (src/Grid/Geometry/grid_smb.hpp:175-297): chunk id row-major, x fastest.
// grid_zmb (src/Grid/Geometry/grid_zmb.hpp:65-106, util/zmorton.hpp): chunk id = Morton
// code, bit i of the chunk coordinate d at bit dim*i + d.  Inside a chunk the local id is
// x fastest for both.
struct ChunkLin {
    int dim, zorder;
    int64_t bsz[3];
};

__host__ __device__ inline uint64_t zspread(uint64_t v, int dim)
{
    uint64_t r = 0;
    const int nbits = 64 / dim;
    for (int i = 0; i < nbits; ++i) r |= ((v >> i) & 1ull) << (dim * i);
    return r;
}

__host__ __device__ inline uint64_t zcompact(uint64_t v, int dim)
{
    uint64_t r = 0;
    const int nbits = 64 / dim;
    for (int i = 0; i < nbits; ++i) r |= ((v >> (dim * i)) & 1ull) << i;
    return r;
}

// chunk coordinates of chunk id `id`
// (loops run to the constant 3 and stop at dim, so that callers' coordinate arrays stay
// in registers)
__host__ __device__ inline void chunk_coords(const ChunkLin &L, int64_t id, int64_t *bc)
{
    if (L.zorder) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= L.dim) break;
            bc[d] = (int64_t)zcompact((uint64_t)id >> d, L.dim);
        }
        return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d >= L.dim) break;
        bc[d] = id % L.bsz[d];
        id /= L.bsz[d];
    }
}

// chunk id of chunk coordinates inside the domain
__host__ __device__ inline int64_t chunk_id(const ChunkLin &L, const int64_t *bc)
{
    if (L.zorder) {
        uint64_t r = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= L.dim) break;
            r |= zspread((uint64_t)bc[d], L.dim) << d;
        }
        return (int64_t)r;
    }
    int64_t v = 0;
#pragma unroll
    for (int d = 2; d >= 0; --d) {
        if (d >= L.dim) continue;
        v = v * L.bsz[d] + bc[d];
    }
    return v;
}

inline ChunkLin chunk_lin(const sg_grid *g);

// host launchers implemented in the .hip translation units
int flush_impl(sg_grid *g, const int32_t *ops);
// pending whole-chunk records -> point records (before point records are appended)
int convert_chunk_records(sg_grid *g);
// pending staged (chunk id, record) pairs -> the point insert buffer (before anything
// else is appended to it or reads it)
int convert_point_records(sg_grid *g);
// room for n more point records at *at of the insert buffer (chunk records converted first)
int reserve_point_records(sg_grid *g, int64_t n, int64_t *at);
int append_points(sg_grid *g, int64_t n, const int64_t *keys, const uint8_t *masks,
                  const void *const *values, int mem);
int append_coords(sg_grid *g, int64_t n, const int32_t *coords, const uint8_t *masks,
                  const void *const *values, int mem);
int append_chunks(sg_grid *g, int64_t n, const int64_t *chunk_ids, const uint8_t *masks,
                  const void *const *values, int mem);
int find_neighbours_impl(sg_grid *g, int type);
int insert_device_begin(sg_grid *g, int64_t nrec, sg_insert_view *out);
int insert_device_end(sg_grid *g);
int tag_boundaries_impl(sg_grid *g, int type);
// existing-point list of the point-list stencil mode (d_ep, ep_n)
int existing_points_impl(sg_grid *g);
int apply_stencil_impl(sg_grid *g, const sg_stencil_desc *d);
int get_points_impl(sg_grid *g, int prop, int64_t n, const int64_t *keys, void *out, int mem);
int fill_background_chunk(sg_grid *g, int prop);
int normalize_prop(sg_grid *g, int prop);
// occupancy table of the 8^3 star stencils (sg_grid::d_occ_tab), rebuilt when stale
int occ_table_impl(sg_grid *g);
// EXIST bit-planes of 8^3 chunks (sg_grid::d_ebits), rebuilt when stale
int ebits_impl(sg_grid *g);
int count_mask_impl(sg_grid *g, int which, int64_t *out);
// replace the chunk list and masks (hostToDevice of a host copy whose chunks changed)
int set_chunks_impl(sg_grid *g, int64_t n, const int64_t *ids, const uint8_t *masks, int mem);
// background into every mask-0 slot of property prop (every property for -1)
int restore_background_impl(sg_grid *g, int prop);

// device-wide primitives (sg_primitives.hip)
// exclusive scan of n uint32 values; *total (host) receives the sum when non-null
int exclusive_scan_u32(sg_grid *g, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total);
// the same on `stream` with a caller-provided workspace of ws_words >= scan_ws_words(n) words;
// d_total (device, may be null) receives the sum; allocates nothing and does not wait
size_t scan_ws_words(int64_t n);
int exclusive_scan_u32_ws(hipStream_t stream, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total,
                          uint32_t *ws, size_t ws_words);
// pinned host words of g (16 of them), allocated on first use; nullptr on failure
uint32_t *host_words(sg_grid *g);
// stable LSD radix sort of (key, value) pairs on the low `bits` bits of the key
int radix_sort_pairs_u64(sg_grid *g, uint64_t *keys, uint32_t *vals, uint64_t *keys_tmp,
                         uint32_t *vals_tmp, int64_t n, int bits, uint64_t **keys_out,
                         uint32_t **vals_out);
int radix_sort_pairs_u32(sg_grid *g, uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp,
                         uint32_t *vals_tmp, int64_t n, int bits, uint32_t **keys_out,
                         uint32_t **vals_out);
// the same for (key, record of nv 16-byte vectors, nv = 1 or 2): the records move with
// their keys.  With keys_tmp2 / recs_tmp2 the input is only read: the passes after the
// first alternate between the two temporaries
int radix_sort_records_u32(sg_grid *g, uint32_t *keys, uint4 *recs, uint32_t *keys_tmp, uint4 *recs_tmp, int nv,
                           int64_t n, int bits, uint32_t **keys_out, uint4 **recs_out,
                           uint32_t *keys_tmp2 = nullptr, uint4 *recs_tmp2 = nullptr);

inline unsigned grid_1d(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

inline ChunkLin chunk_lin(const sg_grid *g)
{
    ChunkLin L;
    L.dim = g->dim;
    L.zorder = g->lin == SG_LIN_ZMB;
    for (int d = 0; d < 3; ++d) L.bsz[d] = g->bsz[d];
    return L;
}

}  // namespace sg
