// sg_flush.hip -- insertion buffer and flush for gfx950.
//
// Reference pipeline (src/Vector/map_vector_sparse.hpp:1192-1226 ->
// src/SparseGridGpu/BlockMapGpu_kernels.cuh:700-807): every insert claims a whole
// chunk slot; slots are compacted, stably radix-sorted by chunk key, run-length
// segmented, merged with the old sorted chunk list, and each output chunk is the
// per-offset fold of its slots in sorted (= insertion) order followed by a fold
// with the old chunk: a = op(a,b) when both exist, a = b when only b exists, mask =
// OR of all masks.
//
// Here the records are points, not chunk slots (a 2^24-insert batch as chunk slots
// would need 2^24 * 4.6 KB): records are (point key, mask byte, values).  The fold a
// point receives is the same: the records of that point in insertion order, then
// the old value.  A chunk slot of the reference contributes to an offset only
// through its mask there (mask 0 leaves the fold unchanged), so splitting slots into
// points preserves the result at every existing point.
//
// Steps: stable LSD radix sort of (point key, record id) -> run heads + scan ->
// unique points -> chunk heads + scan -> unique chunks -> binary-search merge with
// the old chunk list (new chunks get positions by a merge-path count) -> one thread
// per unique point folds its records and the old value in place, every property and
// then the mask.
//
// Whole-chunk inserts (insertBlock) stay chunk records while no point record is
// pending: only the chunk ids are sorted, and one workgroup per touched chunk folds
// its records slot by slot with the same rules (a record takes part in the fold of the
// slots where its mask is non-zero).  Appending point records after chunk records
// first turns the chunk records into point records, so insertion order is kept.
#include "sg_internal.hpp"

#include <algorithm>

namespace sg {
namespace {

constexpr int kWG = 256;

template <typename T> __device__ inline T dev_add(T a, T b)
{
    if constexpr (std::is_integral<T>::value) {
        // wrapping arithmetic in an unsigned type of at least 32 bits (narrower types are
        // promoted to signed int, where a product can overflow), truncated to T
        typedef typename std::conditional<(sizeof(T) < 4), uint32_t, typename std::make_unsigned<T>::type>::type U;
        return (T)(U)((U)a + (U)b);
    } else {
        return a + b;
    }
}

// reduction operators of the reference (map_vector_sparse_cuda_kernels.cuh:42-450);
// max/min follow moderngpu's (a < b) ? b : a and (b < a) ? b : a
template <typename T> __device__ inline T dev_reduce(int op, T a, T b)
{
    switch (op) {
    case SG_OP_MAX: return (a < b) ? b : a;
    case SG_OP_MIN: return (b < a) ? b : a;
    case SG_OP_ADD: return dev_add(a, b);
    case SG_OP_RIGHT: return b;
    case SG_OP_LEFT: return a;
    case SG_OP_OR:
        if constexpr (std::is_integral<T>::value) return (T)(a | b);
        else return a;
    default: return b;
    }
}

// ---- insert buffer ------------------------------------------------------------
__global__ void fill_u8_kernel(uint8_t *p, uint8_t v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T> __global__ void fill_kernel(T *p, T v, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// coordinates -> grid_smb linear ids (grid_smb.hpp:175-189); a coordinate outside the grid
// gives key -1, which the next flush drops
__global__ void coords_to_keys_kernel(const int32_t *coords, int64_t *keys, int64_t n, int dim, int edge, int64_t sz0,
                                      int64_t sz1, int64_t sz2, ChunkLin L, int csz)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t c[3] = {0, 0, 0};
    int64_t sz[3] = {sz0, sz1, sz2};
    bool ok = true;
    for (int d = 0; d < dim; ++d) {
        c[d] = coords[i * dim + d];
        ok = ok && c[d] >= 0 && c[d] < sz[d];
    }
    if (!ok) {
        // device coordinates are not read back at insertion: the flush's key check finds it
        keys[i] = -1;
        return;
    }
    int64_t bc[3] = {0, 0, 0}, l = 0;
    for (int d = dim - 1; d >= 0; --d) {
        bc[d] = c[d] / edge;
        l = l * edge + c[d] % edge;
    }
    keys[i] = chunk_id(L, bc) * csz + l;
}

// the values of the properties of one batch of point inserts: source arrays (nullptr:
// the background), insert-buffer destinations
struct StageVals {
    int np;
    int ts[kMaxProps];
    const unsigned char *src[kMaxProps];
    unsigned char *dst[kMaxProps];
    uint64_t bg[kMaxProps];
};

constexpr int kStageItems = 4;  // records per thread of the streaming insert kernels

// Point inserts given by device coordinates, in one pass: the key (as
// coords_to_keys_kernel, in 32-bit arithmetic: coordinates are int32), the mask (masks
// or EXIST) and every property's value, each to its insert-buffer array.  A block takes
// kWG * kStageItems consecutive records, a thread every kWG-th of them, all loads of a
// step issued before its stores.
__global__ __launch_bounds__(kWG) void stage_coords_kernel(const int32_t *__restrict__ coords,
                                                           const uint8_t *__restrict__ masks, int64_t n, int dim,
                                                           int edge, int sz0, int sz1, int sz2, ChunkLin L, int csz,
                                                           int64_t *__restrict__ keys, uint8_t *__restrict__ omask,
                                                           StageVals V)
{
    const int64_t i0 = (int64_t)blockIdx.x * (kWG * kStageItems) + threadIdx.x;
    const int sz[3] = {sz0, sz1, sz2};
    int c[kStageItems][3];
#pragma unroll
    for (int t = 0; t < kStageItems; ++t) {
        const int64_t i = i0 + (int64_t)t * kWG;
#pragma unroll
        for (int d = 0; d < 3; ++d) c[t][d] = (d < dim && i < n) ? coords[i * dim + d] : 0;
    }
    const uint32_t e = (uint32_t)edge;
#pragma unroll
    for (int t = 0; t < kStageItems; ++t) {
        const int64_t i = i0 + (int64_t)t * kWG;
        if (i >= n) break;
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (d >= dim || (c[t][d] >= 0 && c[t][d] < sz[d]));
        int64_t key = -1;  // outside the grid: the next flush's key check finds it
        if (ok) {
            int64_t bc[3] = {0, 0, 0};
            uint32_t l = 0;
#pragma unroll
            for (int d = 2; d >= 0; --d) {
                if (d >= dim) continue;
                const uint32_t q = (uint32_t)c[t][d] / e;
                bc[d] = q;
                l = l * e + ((uint32_t)c[t][d] - q * e);
            }
            key = chunk_id(L, bc) * csz + l;
        }
        keys[i] = key;
    }
    {
        uint8_t m[kStageItems];
#pragma unroll
        for (int t = 0; t < kStageItems; ++t) {
            const int64_t i = i0 + (int64_t)t * kWG;
            m[t] = (masks && i < n) ? masks[i] : (uint8_t)SG_EXIST;
        }
#pragma unroll
        for (int t = 0; t < kStageItems; ++t) {
            const int64_t i = i0 + (int64_t)t * kWG;
            if (i < n) omask[i] = m[t];
        }
    }
    for (int p = 0; p < V.np; ++p) {
        const unsigned char *s = V.src[p];
        unsigned char *d = V.dst[p];
        const uint64_t bg = V.bg[p];
        uint64_t v[kStageItems];
#pragma unroll
        for (int t = 0; t < kStageItems; ++t) {
            const int64_t i = i0 + (int64_t)t * kWG;
            v[t] = bg;
            if (s && i < n) {
                switch (V.ts[p]) {
                case 8: v[t] = ((const uint64_t *)s)[i]; break;
                case 4: v[t] = ((const uint32_t *)s)[i]; break;
                case 2: v[t] = ((const uint16_t *)s)[i]; break;
                default: v[t] = s[i]; break;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < kStageItems; ++t) {
            const int64_t i = i0 + (int64_t)t * kWG;
            if (i >= n) break;
            switch (V.ts[p]) {
            case 8: ((uint64_t *)d)[i] = v[t]; break;
            case 4: ((uint32_t *)d)[i] = (uint32_t)v[t]; break;
            case 2: ((uint16_t *)d)[i] = (uint16_t)v[t]; break;
            default: d[i] = (unsigned char)v[t]; break;
            }
        }
    }
}

__global__ void check_keys_kernel(const int64_t *keys, int64_t n, int64_t limit, int *bad)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (keys[i] < 0 || keys[i] >= limit)) *bad = 1;
}

// chunk records -> count of point records (non-zero masks; an all-zero chunk keeps
// one record so that the chunk is still created, as the reference creates it)
__global__ __launch_bounds__(kWG) void chunk_count_kernel(const uint8_t *masks, int csz, uint32_t *cnt)
{
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const uint8_t *m = masks + (int64_t)blockIdx.x * csz;
    uint32_t c = 0;
    for (int o = threadIdx.x; o < csz; o += blockDim.x) c += (m[o] != 0);
    atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = s ? s : 1;
}

// expand chunk records into point records in offset order (one workgroup per chunk;
// a wave-ordered compaction keeps offsets ascending)
__global__ __launch_bounds__(kWG) void chunk_expand_kernel(const int64_t *chunk_ids, const uint8_t *masks,
                                                          const uint32_t *start, int csz, int64_t *okey,
                                                          uint8_t *omask, uint32_t *osrc, int64_t base)
{
    __shared__ uint32_t s_run;
    __shared__ uint32_t s_wave[kWG / 64];
    const int64_t c = blockIdx.x;
    const uint8_t *m = masks + c * csz;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    bool any = false;
    for (int o0 = 0; o0 < csz; o0 += kWG) {
        const int o = o0 + threadIdx.x;
        const bool has = (o < csz) && m[o] != 0;
        const uint64_t bal = __ballot(has);
        const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        if (lane == 0) s_wave[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t pre = s_run;
        uint32_t tot = 0;
        for (int q = 0; q < kWG / 64; ++q) {
            if (q < w) pre += s_wave[q];
            tot += s_wave[q];
        }
        if (has) {
            uint32_t pos = start[c] + pre + (uint32_t)__popcll(bal & lt);
            okey[pos] = chunk_ids[c] * csz + o;
            omask[pos] = m[o];
            osrc[pos] = (uint32_t)(c * csz + o);
        }
        any = any || tot != 0;
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
    if (!any && threadIdx.x == 0) {
        uint32_t pos = start[c];
        okey[pos] = chunk_ids[c] * csz;
        omask[pos] = 0;
        osrc[pos] = (uint32_t)(c * csz);
    }
    (void)base;
}

template <typename T> __global__ void gather_kernel(const T *src, const uint32_t *idx, T *dst, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// ---- flush kernels -------------------------------------------------------------
// sort keys and record ids; with `bad`, also the range check of the keys (flag set when
// one lies outside [0, limit))
template <typename K>
__global__ void keys_to_sort_kernel(const int64_t *keys, K *sk, uint32_t *sv, int64_t n, int64_t limit, int *bad)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int64_t k = keys[i];
        if (bad && (k < 0 || k >= limit)) *bad = 1;
        sk[i] = (K)k;
        sv[i] = (uint32_t)i;
    }
}

// run-start predicate (find_buffer_offsets_for_scan, src/util/cuda/kernels.cuh:57-68)
template <typename K> __global__ void run_head_kernel(const K *sk, uint32_t *head, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || sk[i - 1] != sk[i]) ? 1u : 0u;
}

// Runs of equal keys in a sorted uint32 array, in tiles of kRunTile items: thread t of a
// tile holds items t * kRunIPT .. + kRunIPT - 1 (consecutive, read as 16-byte vectors);
// item i starts a run when i == 0 or its key differs from item i - 1's.
constexpr int kRunIPT = 16;
constexpr int kRunTile = kWG * kRunIPT;

__device__ inline void run_items(const uint32_t *__restrict__ sk, int64_t n, int64_t i0, uint32_t (&k)[kRunIPT],
                                 uint32_t &prev)
{
    if (i0 + kRunIPT <= n) {
#pragma unroll
        for (int v = 0; v < kRunIPT / 4; ++v) {
            const uint4 q = ((const uint4 *)(sk + i0))[v];
            k[4 * v] = q.x;
            k[4 * v + 1] = q.y;
            k[4 * v + 2] = q.z;
            k[4 * v + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kRunIPT; ++j) k[j] = (i0 + j < n) ? sk[i0 + j] : 0u;
    }
    // the item before the thread's first (none for item 0: its key is taken as different)
    prev = (i0 > 0 && i0 <= n) ? sk[i0 - 1] : ~k[0];
}

// exclusive prefix of v over the kWG threads of the block (*total: the block's sum)
__device__ inline uint32_t run_block_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t s_w[kWG / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kWG / 64; ++q) {
        const uint32_t s = s_w[q];
        pre += (q < w) ? s : 0u;
        tot += s;
    }
    *total = tot;
    return pre + inc - v;
}

// the number of runs starting in each tile
__global__ __launch_bounds__(kWG) void run_count_kernel(const uint32_t *__restrict__ sk, int64_t n,
                                                       uint32_t *__restrict__ cnt)
{
    const int64_t i0 = (int64_t)blockIdx.x * kRunTile + (int64_t)threadIdx.x * kRunIPT;
    uint32_t k[kRunIPT], prev;
    run_items(sk, n, i0, k, prev);
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kRunIPT; ++j) {
        c += (i0 + j < n && k[j] != prev) ? 1u : 0u;
        prev = k[j];
    }
    uint32_t tot;
    run_block_scan(c, &tot);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// each run's key and first item at the run's rank (off: exclusive scan of the tiles'
// counts; start[nu] = n)
__global__ __launch_bounds__(kWG) void run_emit_kernel(const uint32_t *__restrict__ sk, int64_t n,
                                                      const uint32_t *__restrict__ off, int64_t *__restrict__ key,
                                                      uint32_t *__restrict__ start, uint32_t nu)
{
    const int64_t i0 = (int64_t)blockIdx.x * kRunTile + (int64_t)threadIdx.x * kRunIPT;
    uint32_t k[kRunIPT], prev;
    run_items(sk, n, i0, k, prev);
    uint32_t heads = 0, c = 0;
#pragma unroll
    for (int j = 0; j < kRunIPT; ++j) {
        const bool h = i0 + j < n && k[j] != prev;
        heads |= (h ? 1u : 0u) << j;
        c += h ? 1u : 0u;
        prev = k[j];
    }
    uint32_t tot;
    uint32_t u = off[blockIdx.x] + run_block_scan(c, &tot);
#pragma unroll
    for (int j = 0; j < kRunIPT; ++j) {
        if ((heads >> j) & 1u) {
            key[u] = (int64_t)k[j];
            start[u] = (uint32_t)(i0 + j);
            ++u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) start[nu] = (uint32_t)n;
}

// unique points (construct_index_unique, src/util/cuda/kernels.cuh:34-54)
template <typename K>
__global__ void unique_points_kernel(const K *sk, const uint32_t *head, const uint32_t *hscan, int64_t n,
                                     int64_t *ukey, uint32_t *rstart, uint32_t nu)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && head[i]) {
        ukey[hscan[i]] = (int64_t)sk[i];
        rstart[hscan[i]] = (uint32_t)i;
    }
    if (i == 0) rstart[nu] = (uint32_t)n;
}

__global__ void chunk_head_kernel(const int64_t *ukey, uint32_t *head, int64_t nu, int csz)
{
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nu) head[u] = (u == 0 || ukey[u - 1] / csz != ukey[u] / csz) ? 1u : 0u;
}

// unique chunk ids, and (ustart != nullptr) the first unique point of each, ustart[nn] = nu
__global__ void unique_chunks_kernel(const int64_t *ukey, const uint32_t *head, const uint32_t *hscan, int64_t nu,
                                     int csz, int64_t *nkey, uint32_t *ustart, uint32_t nn)
{
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nu && head[u]) {
        nkey[hscan[u]] = ukey[u] / csz;
        if (ustart) ustart[hscan[u]] = (uint32_t)u;
    }
    if (ustart && u == nu - 1) ustart[nn] = (uint32_t)nu;
}

__device__ inline int64_t lower_bound_dev(const int64_t *a, int64_t n, int64_t x)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// new chunk j: position in the old list or -1, and the brand-new flag
__global__ void old_lookup_kernel(const int64_t *nkey, int64_t nn, const int64_t *okeys, int64_t m, int64_t *oldpos,
                                  uint32_t *isnew)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    int64_t k = nkey[j];
    int64_t lb = lower_bound_dev(okeys, m, k);
    bool found = lb < m && okeys[lb] == k;
    oldpos[j] = found ? lb : -1;
    isnew[j] = found ? 0u : 1u;
}

__global__ void brand_new_kernel(const int64_t *nkey, const uint32_t *isnew, const uint32_t *nscan, int64_t nn,
                                 int64_t *bn)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nn && isnew[j]) bn[nscan[j]] = nkey[j];
}

// final chunk position of each new-touched chunk (merge-path rank in the union)
__global__ void chunk_pos_kernel(const int64_t *nkey, const int64_t *oldpos, const uint32_t *nscan, int64_t nn,
                                 const int64_t *okeys, int64_t m, const int64_t *bn, int64_t nb, int64_t *cpos)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    int64_t k = nkey[j];
    int64_t op = oldpos[j];
    if (op >= 0) cpos[j] = op + lower_bound_dev(bn, nb, k);
    else cpos[j] = (int64_t)nscan[j] + lower_bound_dev(okeys, m, k);
}

// scatter the old chunks into the enlarged arrays (copy_old_ker, BlockMapGpu_kernels.cuh:489-513)
// src_of[p] = the old chunk that lands at p, or -1 for a new chunk
__global__ void move_old_keys_kernel(const int64_t *okeys, int64_t m, const int64_t *bn, int64_t nb, int64_t *nkeys,
                                     int64_t *opos, int64_t *src_of)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t p = i + lower_bound_dev(bn, nb, okeys[i]);
    nkeys[p] = okeys[i];
    opos[i] = p;
    src_of[p] = i;
}

__global__ void place_new_keys_kernel(const int64_t *bn, int64_t nb, const int64_t *okeys, int64_t m, int64_t *nkeys,
                                      int64_t *bpos, int64_t *src_of)
{
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int64_t p = b + lower_bound_dev(okeys, m, bn[b]);
    nkeys[p] = bn[b];
    bpos[b] = p;
    src_of[p] = -1;
}

// The grown chunk arrays in one pass: for every array (mask, properties) and every output
// chunk position j < nout, chunk j is the old chunk src_of[j] or, for a new chunk and the
// background chunk at nout - 1, the array's fill pattern. 16-byte vectors in output order,
// so every store stream is sequential.
struct UnionArrays {
    int na;
    // mask, properties and the EXIST bit-planes (when kept)
    const uint4 *src[kMaxProps + 2];
    uint4 *dst[kMaxProps + 2];
    uint64_t pat[kMaxProps + 2];  // fill value replicated over 8 bytes
    int vshift[kMaxProps + 2];    // log2 of the 16-byte vectors per chunk
};

// skip (optional): output chunks written by someone else (the fused fold)
__global__ __launch_bounds__(kWG) void build_union_kernel(UnionArrays U, const int64_t *src_of, int64_t nout,
                                                         const uint8_t *skip)
{
    const int64_t stride = (int64_t)gridDim.x * kWG;
    // constant indices into the kernel argument (a runtime index would copy it to scratch)
#pragma unroll
    for (int a = 0; a < kMaxProps + 2; ++a) {
        if (a >= U.na) break;
        const int sh = U.vshift[a];
        const int64_t nv = nout << sh, lane_mask = ((int64_t)1 << sh) - 1;
        const uint4 *src = U.src[a];
        uint4 *dst = U.dst[a];
        const uint4 pat = make_uint4((uint32_t)U.pat[a], (uint32_t)(U.pat[a] >> 32), (uint32_t)U.pat[a],
                                     (uint32_t)(U.pat[a] >> 32));
        for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < nv; i += stride) {
            const int64_t j = i >> sh;
            if (skip && j < nout - 1 && skip[j]) continue;
            const int64_t c = (j < nout - 1) ? src_of[j] : -1;
            uint4 v = pat;
            if (c >= 0) v = src[(c << sh) + (i & lane_mask)];
            dst[i] = v;
        }
    }
}

// one workgroup per chunk: copy chunk `src` of an array to chunk `dst` of another
template <typename V>
__global__ __launch_bounds__(kWG) void move_chunks_kernel(const V *src, V *dst, const int64_t *dpos, int64_t m,
                                                         int vec_per_chunk)
{
    const int64_t c = blockIdx.x;
    if (c >= m) return;
    const int64_t d = dpos[c];
    for (int i = threadIdx.x; i < vec_per_chunk; i += kWG) dst[d * vec_per_chunk + i] = src[c * vec_per_chunk + i];
}

// fill chunks with one value (16-byte stores when the chunk size allows)
template <typename T>
__global__ __launch_bounds__(kWG) void init_chunks_kernel(T *dst, const int64_t *dpos, int64_t nb, int csz, T v)
{
    const int64_t b = blockIdx.x;
    if (b >= nb) return;
    const int64_t d = dpos ? dpos[b] : b;
    if (((size_t)csz * sizeof(T)) % 16 == 0) {
        union {
            T t[16 / sizeof(T)];
            uint4 q;
        } pat;
#pragma unroll
        for (int i = 0; i < (int)(16 / sizeof(T)); ++i) pat.t[i] = v;
        uint4 *p = (uint4 *)(dst + d * csz);
        const int nq = (int)((size_t)csz * sizeof(T) / 16);
        for (int i = threadIdx.x; i < nq; i += kWG) p[i] = pat.q;
        return;
    }
    for (int o = threadIdx.x; o < csz; o += kWG) dst[d * csz + o] = v;
}

// ---- fold of every property and the mask ------------------------------------------
struct FoldProps {
    int np;
    int dt[kMaxProps];
    int ts[kMaxProps];
    int op[kMaxProps];
    const unsigned char *rval[kMaxProps];  // record values (insert buffer, one array per property)
    unsigned char *out[kMaxProps];         // chunk arrays
    uint64_t bg[kMaxProps];
    // the records interleaved and in key order for the fold: the i-th record of the sort
    // at rec + i * stride, property q at byte off[q], mask byte at moff
    const unsigned char *rec;
    int stride, moff;
    int off[kMaxProps];
};

// insert buffer (one array per property + masks) -> interleaved records, in order
__global__ __launch_bounds__(kWG) void interleave_records_kernel(int64_t n, const uint8_t *rmask, FoldProps F,
                                                                unsigned char *rec)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    unsigned char *d = rec + r * F.stride;
    for (int q = 0; q < F.np; ++q) {
        const unsigned char *s = F.rval[q] + r * F.ts[q];
        unsigned char *t = d + F.off[q];
        switch (F.ts[q]) {
        case 8: *(uint64_t *)t = *(const uint64_t *)s; break;
        case 4: *(uint32_t *)t = *(const uint32_t *)s; break;
        case 2: *(uint16_t *)t = *(const uint16_t *)s; break;
        default: *t = *s; break;
        }
    }
    d[F.moff] = rmask[r];
}

// records in key order: out[i] = rec[perm[i]], in 16-byte vectors (stride = vecs * 16);
// one thread per vector, so the random gathers are all in flight together
__global__ __launch_bounds__(kWG) void permute_records_kernel(const uint4 *__restrict__ rec,
                                                             const uint32_t *__restrict__ perm, int64_t n, int vecs,
                                                             uint4 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * vecs) return;
    const int64_t r = i / vecs, k = i - r * vecs;
    out[i] = rec[(int64_t)perm[r] * vecs + k];
}

// The fold of one unique point, every property (segreduce_total_with_mask + applyOp,
// BlockMapGpu_kernels.cuh:326-477, BlockMapGpu_dimensionalityWrappers.cuh:214-226): the
// point's records in insertion order, a = op(a, b) when both exist, a = b when only b
// exists; then the old value as b; mask = OR of everything.
template <typename T> __device__ inline T fold_records(int op, uint32_t r0, uint32_t r1, const FoldProps &F, int q)
{
    const unsigned char *p = F.rec + (size_t)r0 * F.stride;
    T a = *(const T *)(p + F.off[q]);
    uint8_t am = p[F.moff];
    for (uint32_t i = r0 + 1; i < r1; ++i) {
        p = F.rec + (size_t)i * F.stride;
        const uint8_t bm = p[F.moff];
        const bool aE = am & 1, bE = bm & 1;
        const T b = *(const T *)(p + F.off[q]);
        if (op == SG_OP_NONE) {
            if (bE) a = b;
        } else if (aE && bE) {
            a = dev_reduce<T>(op, a, b);
        } else if (bE) {
            a = b;
        }
        am |= bm;
    }
    return a;
}

template <typename T> __device__ inline void fold_old(int op, T a, uint8_t am, T *slot, uint8_t om, T bgv)
{
    const bool aE = am & 1, oE = om & 1;
    if (op == SG_OP_NONE) {
        if (!aE && oE) a = *slot;
    } else if (aE && oE) {
        a = dev_reduce<T>(op, a, *slot);
    } else if (oE) {
        a = *slot;
    }
    *slot = ((am | om) == 0) ? bgv : a;
}

#define SG_FOLD_TYPES(X)                                                                                       \
    X(SG_F32, float) X(SG_F64, double) X(SG_I32, int32_t) X(SG_U32, uint32_t) X(SG_I64, int64_t)               \
    X(SG_U64, uint64_t) X(SG_I8, int8_t) X(SG_U8, uint8_t) X(SG_I16, int16_t) X(SG_U16, uint16_t)

// one thread per unique point: every property and then the mask, in place in the chunk
// arrays (the records' mask and order are read once per property, from cache after the
// first)
__global__ __launch_bounds__(kWG) void fold_points_kernel(const int64_t *ukey, const uint32_t *rstart,
                                                         const uint32_t *uchunk, const int64_t *cpos, int64_t nu,
                                                         int csz, FoldProps F, uint8_t *omask)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t r0 = rstart[u], r1 = rstart[u + 1];
    const int64_t slot = cpos[uchunk[u]] * csz + (ukey[u] % csz);
    uint8_t am = 0;
    for (uint32_t i = r0; i < r1; ++i) am |= F.rec[(size_t)i * F.stride + F.moff];
    const uint8_t om = omask[slot];
    for (int q = 0; q < F.np; ++q) {
        switch (F.dt[q]) {
#define SG_CASE(CODE, T)                                                                                        \
    case CODE: {                                                                                                \
        T bgv;                                                                                                  \
        __builtin_memcpy(&bgv, &F.bg[q], sizeof(T));                                                            \
        const T a = fold_records<T>(F.op[q], r0, r1, F, q);                                                   \
        fold_old<T>(F.op[q], a, am, (T *)F.out[q] + slot, om, bgv);                                            \
        break;                                                                                                  \
    }
            SG_FOLD_TYPES(SG_CASE)
#undef SG_CASE
        }
    }
    omask[slot] = (uint8_t)(om | am);
}

__global__ void mark_positions_kernel(const int64_t *pos, int64_t n, uint8_t *flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[pos[i]] = 1;
}

// LDS image of one chunk for the fused fold: property q at byte off[q], mask at moff,
// `bytes` per wave; src_* are the arrays touched old chunks are read from
struct FusedImg {
    int off[kMaxProps];
    int moff, bytes;
    uint64_t pat[kMaxProps];  // background replicated over 8 bytes
    const unsigned char *src[kMaxProps];
    const uint8_t *smask;
};

// 16-byte copies between global memory and a wave's LDS image (bytes % 16 == 0)
__device__ inline void image_in(unsigned char *img, const unsigned char *src, int bytes, uint64_t pat, int lane)
{
    uint4 *d = (uint4 *)img;
    const int nv = bytes >> 4;
    if (src) {
        const uint4 *s = (const uint4 *)src;
        for (int i = lane; i < nv; i += 64) d[i] = s[i];
    } else {
        const uint4 v = make_uint4((uint32_t)pat, (uint32_t)(pat >> 32), (uint32_t)pat, (uint32_t)(pat >> 32));
        for (int i = lane; i < nv; i += 64) d[i] = v;
    }
}

__device__ inline void image_out(unsigned char *dst, const unsigned char *img, int bytes, int lane)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const int nv = bytes >> 4;
    for (int i = lane; i < nv; i += 64) __builtin_nontemporal_store(((const v4u *)img)[i], (v4u *)dst + i);
}

// The fold of every unique point (same rules as fold_points_kernel), written to a compact
// array in key order instead of into the chunks: result u at res + u * F.stride, in the
// record layout (property q at byte F.off[q], mask at F.moff), so the stores are
// coalesced.  The old value comes from the chunk's old contents (I.src / I.smask at its
// old position) or is the background for a brand-new chunk.
__global__ __launch_bounds__(kWG) void fold_points_compact_kernel(const int64_t *__restrict__ ukey,
                                                                  const uint32_t *__restrict__ rstart,
                                                                  const uint32_t *__restrict__ uchunk,
                                                                  const int64_t *__restrict__ oldpos,
                                                                  const uint32_t *__restrict__ isnew, int64_t nu,
                                                                  int csz, FoldProps F, FusedImg I,
                                                                  unsigned char *__restrict__ res)
{
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint32_t r0 = rstart[u], r1 = rstart[u + 1];
    const uint32_t j = uchunk[u];
    const bool fresh = isnew[j] != 0;
    const int64_t slot = fresh ? 0 : oldpos[j] * csz + (ukey[u] % csz);
    uint8_t am = 0;
    for (uint32_t i = r0; i < r1; ++i) am |= F.rec[(size_t)i * F.stride + F.moff];
    const uint8_t om = fresh ? 0 : I.smask[slot];
    unsigned char *out = res + (size_t)u * F.stride;
    for (int q = 0; q < F.np; ++q) {
        switch (F.dt[q]) {
#define SG_CASE(CODE, T)                                                                                        \
    case CODE: {                                                                                                \
        T bgv;                                                                                                  \
        __builtin_memcpy(&bgv, &F.bg[q], sizeof(T));                                                            \
        const T a = fold_records<T>(F.op[q], r0, r1, F, q);                                                   \
        T v = fresh ? bgv : ((const T *)I.src[q])[slot];                                                        \
        fold_old<T>(F.op[q], a, am, &v, om, bgv);                                                               \
        *(T *)(out + F.off[q]) = v;                                                                             \
        break;                                                                                                  \
    }
            SG_FOLD_TYPES(SG_CASE)
#undef SG_CASE
        }
    }
    out[F.moff] = (uint8_t)(om | am);
}

// One wave per touched chunk j, one array at a time (each property, then the mask): the
// array's chunk is built in LDS (old contents, or the background / mask 0 when new), the
// folded results of the chunk's unique points are placed at their slots, and the array's
// chunk is written once, contiguously, at its final position.  A wave holds one array's
// chunk plus a 1 KB stage of its first 64 results in LDS, so more waves fit a CU than
// with the whole chunk image.
__global__ __launch_bounds__(256) void write_chunks_kernel(const int64_t *__restrict__ ukey,
                                                           const uint32_t *__restrict__ ustart,
                                                           const int64_t *__restrict__ cpos,
                                                           const int64_t *__restrict__ oldpos,
                                                           const uint32_t *__restrict__ isnew, int64_t nnew, int csz,
                                                           int waves, int img_bytes, FoldProps F, FusedImg I,
                                                           const unsigned char *__restrict__ res,
                                                           uint8_t *__restrict__ omask)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_img[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * waves + w;
    if (w >= waves || j >= nnew) return;  // waves are independent: no workgroup barrier
    unsigned char *img = smem_img + (size_t)w * (img_bytes + 1024);
    unsigned char *stage = img + img_bytes;  // 16-byte results of the lanes' first points
    const bool fresh = isnew[j] != 0;
    const int64_t src = fresh ? 0 : oldpos[j];
    const int64_t dst = cpos[j];
    const uint32_t u0 = ustart[j], u1 = ustart[j + 1];
    // the lane's first point (most chunks have fewer than 64): its key and, for 16-byte
    // results, the result are loaded once, before the first array's chunk is initialised
    const uint32_t ua = u0 + lane;
    const bool has = ua < u1, r16 = F.stride == 16;
    const int64_t keya = has ? ukey[ua] : 0;
    uint4 ra = make_uint4(0, 0, 0, 0);
    if (has && r16) ra = *(const uint4 *)(res + (size_t)ua * 16);
    if (r16) *(uint4 *)(stage + lane * 16) = ra;
    for (int a = 0; a <= F.np; ++a) {
        const bool m = a == F.np;  // the mask
        const int ts = m ? 1 : F.ts[a];
        const int bytes = csz * ts;
        const int off = m ? F.moff : F.off[a];
        image_in(img, fresh ? nullptr : (m ? I.smask + src * csz : I.src[a] + src * bytes), bytes, m ? 0 : I.pat[a],
                 lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t u = ua; u < u1; u += 64) {
            const int slot = (int)(((u == ua) ? keya : ukey[u]) % csz);
            const unsigned char *r = ((u == ua && r16) ? stage + lane * 16 : res + (size_t)u * F.stride) + off;
            unsigned char *d = img + slot * ts;
            switch (ts) {
            case 8: *(uint64_t *)d = *(const uint64_t *)r; break;
            case 4: *(uint32_t *)d = *(const uint32_t *)r; break;
            case 2: *(uint16_t *)d = *(const uint16_t *)r; break;
            default: *d = *r; break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        image_out(m ? omask + dst * csz : F.out[a] + dst * bytes, img, bytes, lane);
        // the image is initialised again for the next array
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---- point records grouped by chunk ------------------------------------------------
// A record of the chunk-grouped flush, 16 * nv bytes: uint16 slot in the chunk, uint8
// mask, then the value of every property at its natural alignment from byte 3 on.
struct RecLayout {
    int nv, bytes;
    int off[kMaxProps];
};

__device__ inline void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// bits `v` of `bytes` bytes at byte `off` of an NV-vector record held in registers
template <int NV> __device__ inline void rec_put(uint4 (&r)[NV], int off, int bytes, uint64_t v)
{
    const int w = off >> 2, sh = (off & 3) * 8;
    const uint32_t lo = bytes >= 4 ? (uint32_t)v : (((uint32_t)v & ((1u << (8 * bytes)) - 1u)) << sh);
    const uint32_t keep = bytes >= 4 ? 0u : ~(((1u << (8 * bytes)) - 1u) << sh);
    auto setq = [](uint4 &q, int k, uint32_t x, uint32_t kp) {
        switch (k & 3) {
        case 0: q.x = (q.x & kp) | x; break;
        case 1: q.y = (q.y & kp) | x; break;
        case 2: q.z = (q.z & kp) | x; break;
        default: q.w = (q.w & kp) | x; break;
        }
    };
    // the vector by a branch on the (uniform) word index, not by indexing the array
    auto set = [&](int k, uint32_t x, uint32_t kp) {
        if (NV == 1 || k < 4) setq(r[0], k, x, kp);
        else setq(r[NV - 1], k, x, kp);
    };
    set(w, lo, keep);
    if (bytes == 8) set(w + 1, (uint32_t)(v >> 32), 0u);
}

// insert buffer -> (chunk id, record) pairs in insertion order; a key outside [0, limit)
// sets *bad (the flush then drops those records before anything is merged).  Each thread
// assembles its record in registers (the layout is uniform) and stores it as 16-byte
// vectors, a wave's stores covering consecutive records.
template <int NV>
__global__ __launch_bounds__(kWG) void chunk_records_kernel(const int64_t *__restrict__ keys,
                                                           const uint8_t *__restrict__ mask, int64_t n, int csz,
                                                           int64_t limit, FoldProps F, RecLayout R,
                                                           uint32_t *__restrict__ ck, uint4 *__restrict__ rec,
                                                           int *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (i >= n) return;
    const int64_t k = keys[i];
    const uint8_t m = mask[i];
    const bool ok = k >= 0 && k < limit;
    if (!ok) *bad = 1;
    const int64_t kk = ok ? k : 0;
    ck[i] = (uint32_t)(kk / csz);
    uint4 r[NV];
    r[0] = make_uint4((uint32_t)(kk % csz) | ((uint32_t)m << 16), 0, 0, 0);
    if (NV == 2) r[NV - 1] = make_uint4(0, 0, 0, 0);
    for (int q = 0; q < F.np; ++q) {
        const unsigned char *s = F.rval[q];
        uint64_t v;
        switch (F.ts[q]) {
        case 8: v = ((const uint64_t *)s)[i]; break;
        case 4: v = ((const uint32_t *)s)[i]; break;
        case 2: v = ((const uint16_t *)s)[i]; break;
        default: v = s[i]; break;
        }
        rec_put<NV>(r, R.off[q], F.ts[q], v);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) rec[i * NV + v] = r[v];
}

// LDS fold state of one chunk, per wave: seen[csz] (0xff where a record reached the
// slot, else 0), am[csz] (OR of their masks), then per property csz values at voff[q]
struct FoldState {
    int bytes;  // per wave, a multiple of 16
    int roff;   // staging of a batch of 64 records
    int voff[kMaxProps];
};

// Device-coordinate point inserts staged straight as the grouped flush's (chunk id,
// record) pairs, in insertion order, so that the flush has no insert-buffer pass: the
// chunk id (as stage_coords_kernel computes the key), then the record: slot offset, mask
// (masks or EXIST) and each property's value (the source's, or the background) at the
// layout's offsets.  A coordinate outside the grid gives chunk id 0xffffffff and sets
// *bad.  A block takes kWG * kStageItems consecutive records, a thread every kWG-th; the
// records are assembled in registers (the layout is uniform) and stored as 16-byte
// vectors, a wave's stores covering consecutive records.

template <int NV>
__global__ __launch_bounds__(kWG) void stage_coord_records_kernel(const int32_t *__restrict__ coords,
                                                                  const uint8_t *__restrict__ masks, int64_t n,
                                                                  int dim, int edge, int sz0, int sz1, int sz2,
                                                                  ChunkLin L, StageVals V, RecLayout R,
                                                                  uint32_t *__restrict__ ck, uint4 *__restrict__ rec,
                                                                  int *__restrict__ bad)
{
    const int64_t b0 = (int64_t)blockIdx.x * (kWG * kStageItems);
    const int sz[3] = {sz0, sz1, sz2};
    const uint32_t e = (uint32_t)edge;
    // every load of a step is issued before its stores
    int c[kStageItems][3];
    uint8_t mk[kStageItems];
#pragma unroll
    for (int t = 0; t < kStageItems; ++t) {
        const int64_t i = b0 + t * kWG + threadIdx.x;
#pragma unroll
        for (int d = 0; d < 3; ++d) c[t][d] = (d < dim && i < n) ? coords[i * dim + d] : 0;
        mk[t] = (masks && i < n) ? masks[i] : (uint8_t)SG_EXIST;
    }
    uint4 r[kStageItems][NV];
#pragma unroll
    for (int t = 0; t < kStageItems; ++t) {
        const int64_t i = b0 + t * kWG + threadIdx.x;
        bool ok = i < n;
#pragma unroll
        for (int d = 0; d < 3; ++d) ok = ok && (d >= dim || (c[t][d] >= 0 && c[t][d] < sz[d]));
        uint32_t id = 0xffffffffu, l = 0;
        if (ok) {
            int64_t bc[3] = {0, 0, 0};
#pragma unroll
            for (int d = 2; d >= 0; --d) {
                if (d >= dim) continue;
                const uint32_t q = (uint32_t)c[t][d] / e;
                bc[d] = q;
                l = l * e + ((uint32_t)c[t][d] - q * e);
            }
            id = (uint32_t)chunk_id(L, bc);
        } else if (i < n) {
            *bad = 1;
        }
        if (i < n) ck[i] = id;
        r[t][0] = make_uint4(l | ((uint32_t)mk[t] << 16), 0, 0, 0);
        if (NV == 2) r[t][NV - 1] = make_uint4(0, 0, 0, 0);
    }
    for (int q = 0; q < V.np; ++q) {
        const unsigned char *src = V.src[q];
        const int ts = V.ts[q], off = R.off[q];
        uint64_t v[kStageItems];
#pragma unroll
        for (int t = 0; t < kStageItems; ++t) {
            const int64_t i = b0 + t * kWG + threadIdx.x;
            v[t] = V.bg[q];
            if (src && i < n) {
                switch (ts) {
                case 8: v[t] = ((const uint64_t *)src)[i]; break;
                case 4: v[t] = ((const uint32_t *)src)[i]; break;
                case 2: v[t] = ((const uint16_t *)src)[i]; break;
                default: v[t] = src[i]; break;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < kStageItems; ++t) rec_put<NV>(r[t], off, ts, v[t]);
    }
#pragma unroll
    for (int t = 0; t < kStageItems; ++t) {
        const int64_t i = b0 + t * kWG + threadIdx.x;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < NV; ++k) rec[i * NV + k] = r[t][k];
        }
    }
}

// staged (chunk id, record) pairs -> the point insert buffer (key -1 for chunk id
// 0xffffffff, the mask, each property's value into V.dst)
__global__ __launch_bounds__(kWG) void point_records_to_buffer_kernel(const uint32_t *__restrict__ ck,
                                                                      const unsigned char *__restrict__ rec, int64_t n,
                                                                      int csz, RecLayout R, StageVals V,
                                                                      int64_t *__restrict__ keys,
                                                                      uint8_t *__restrict__ mask)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned char *d = rec + i * R.bytes;
    const uint32_t id = ck[i];
    keys[i] = id == 0xffffffffu ? (int64_t)-1 : (int64_t)id * csz + *(const uint16_t *)d;
    mask[i] = d[2];
    for (int q = 0; q < V.np; ++q) {
        const unsigned char *f = d + R.off[q];
        unsigned char *o = V.dst[q];
        switch (V.ts[q]) {
        case 8: ((uint64_t *)o)[i] = *(const uint64_t *)f; break;
        case 4: ((uint32_t *)o)[i] = *(const uint32_t *)f; break;
        case 2: ((uint16_t *)o)[i] = *(const uint16_t *)f; break;
        default: o[i] = *f; break;
        }
    }
}

// the records of one batch for one property, rank by rank (a rank's records have distinct
// slots); first / ao come from the batch's mask pass.  The operator is a template
// parameter, so the rounds run without dispatch.
template <typename T, int OP>
__device__ inline void fold_batch_rounds(unsigned char *vals, const unsigned char *bp, int off, int rank, int maxr,
                                         bool first, uint8_t ao, uint8_t bm)
{
    T *a = (T *)vals + off;
    const T b = (rank >= 0) ? *(const T *)bp : T();
    const bool aE = ao & 1, bE = bm & 1;
    for (int r = 0; r <= maxr; ++r) {
        if (rank == r) {
            if (first) {
                *a = b;
            } else if (OP == SG_OP_NONE) {
                if (bE) *a = b;
            } else if (aE && bE) {
                *a = dev_reduce<T>(OP, *a, b);
            } else if (bE) {
                *a = b;
            }
        }
        if (maxr > 0) wave_sync_lds();
    }
}

template <typename T>
__device__ inline void fold_batch_prop(int op, unsigned char *vals, const unsigned char *bp, int off, int rank,
                                       int maxr, bool first, uint8_t ao, uint8_t bm)
{
    switch (op) {
    case SG_OP_MAX: fold_batch_rounds<T, SG_OP_MAX>(vals, bp, off, rank, maxr, first, ao, bm); break;
    case SG_OP_MIN: fold_batch_rounds<T, SG_OP_MIN>(vals, bp, off, rank, maxr, first, ao, bm); break;
    case SG_OP_ADD: fold_batch_rounds<T, SG_OP_ADD>(vals, bp, off, rank, maxr, first, ao, bm); break;
    case SG_OP_RIGHT: fold_batch_rounds<T, SG_OP_RIGHT>(vals, bp, off, rank, maxr, first, ao, bm); break;
    case SG_OP_LEFT: fold_batch_rounds<T, SG_OP_LEFT>(vals, bp, off, rank, maxr, first, ao, bm); break;
    case SG_OP_OR: fold_batch_rounds<T, SG_OP_OR>(vals, bp, off, rank, maxr, first, ao, bm); break;
    default: fold_batch_rounds<T, SG_OP_NONE>(vals, bp, off, rank, maxr, first, ao, bm); break;
    }
    wave_sync_lds();
}

// one property of a touched chunk: old contents (OLD; the background for a new chunk)
// combined with the folded records (fold_old) and written once at the final position
template <typename T, bool OLD>
__device__ inline void write_fold_prop(int op, T bgv, const unsigned char *__restrict__ vals,
                                       const uint8_t *__restrict__ seen, const uint8_t *__restrict__ am,
                                       const T *__restrict__ src, const uint8_t *__restrict__ smask, T *__restrict__ dst,
                                       int csz, int lane)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    constexpr int PPV = 16 / (int)sizeof(T);
    if ((csz % 16) == 0) {
        // gfx9 counts stores in vmcnt: a load between stores makes its wait cover the stores
        // issued before it.  So a new chunk's loop has no load at all, and an old chunk's
        // loop loads a group of vectors (values and mask bytes) before storing any of them.
        typedef typename std::conditional<PPV == 16, v4u,
                typename std::conditional<PPV == 8, uint64_t,
                typename std::conditional<PPV == 4, uint32_t, uint16_t>::type>::type>::type MaskW;
        const int nvec = csz / PPV;
        // groups of G vectors per lane: only an old chunk's loop loads, and for 1- and
        // 2-byte values a group would hold 16 or 8 unpacked values per vector (the kernel's
        // register peak, which sets its occupancy), so those load one vector at a time
        constexpr int G = OLD ? (sizeof(T) >= 4 ? 4 : 1) : 1;
        for (int v0 = lane; v0 < nvec; v0 += 64 * G) {
            union {
                v4u q;
                T t[PPV];
            } u[G];
            union {
                MaskW w;
                uint8_t b[PPV];
            } om[G];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int v = v0 + 64 * g;
                if (OLD && v < nvec) {
                    u[g].q = ((const v4u *)src)[v];
                    om[g].w = ((const MaskW *)smask)[v];
                } else {
#pragma unroll
                    for (int x = 0; x < PPV; ++x) {
                        u[g].t[x] = bgv;
                        om[g].b[x] = 0;
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int v = v0 + 64 * g;
                if (v >= nvec) break;
                // the vector's seen bytes in one LDS read (most vectors receive no record)
                union {
                    MaskW w;
                    uint8_t b[PPV];
                } sw;
                sw.w = ((const MaskW *)seen)[v];
                bool any = false;
#pragma unroll
                for (int x = 0; x < PPV; ++x) any = any || sw.b[x] != 0;
                if (any) {
#pragma unroll
                    for (int x = 0; x < PPV; ++x) {
                        if (sw.b[x]) {
                            const int s = v * PPV + x;
                            T o = u[g].t[x];
                            fold_old<T>(op, ((const T *)vals)[s], am[s], &o, om[g].b[x], bgv);
                            u[g].t[x] = o;
                        }
                    }
                }
                __builtin_nontemporal_store(u[g].q, (v4u *)dst + v);
            }
        }
        return;
    }
    for (int s = lane; s < csz; s += 64) {
        T o = OLD ? src[s] : bgv;
        if (seen[s]) fold_old<T>(op, ((const T *)vals)[s], am[s], &o, OLD ? smask[s] : (uint8_t)0, bgv);
        dst[s] = o;
    }
}

// EXIST bits (bit 0 of each byte) of an 8-byte mask row, byte i -> bit i: the bits sit
// at 8i, and multiplying by 0x0102040810204080 moves bit 8i to 56 + i with every partial
// product at its own bit position (no carries)
__device__ inline uint32_t exist_byte(uint64_t row)
{
    return (uint32_t)(((row & 0x0101010101010101ull) * 0x0102040810204080ull) >> 56);
}

// One wave per touched chunk j.  Its records are a contiguous run of the sorted records,
// in insertion order (the sort is stable).  64 at a time they are loaded with one
// coalesced load (the next 64 are loaded while the current ones are folded) and staged in
// LDS; each lane's record is applied to the LDS state of its slot, records of one batch
// that share a slot in lane order (one rank per round: ranks come from ballots over the
// slot bits).  Then every array of the chunk is combined with the chunk's old contents
// and written once, contiguously.
template <int NV>
__global__ __launch_bounds__(256) void fold_chunk_records_kernel(const uint4 *__restrict__ rec,
                                                                 const uint32_t *__restrict__ ustart,
                                                                 const int64_t *__restrict__ cpos,
                                                                 const int64_t *__restrict__ oldpos,
                                                                 const uint32_t *__restrict__ isnew, int64_t nnew,
                                                                 int csz, int obits, int waves, FoldState S,
                                                                 FoldProps F, RecLayout R, FusedImg I,
                                                                 uint8_t *__restrict__ omask, int skip_fresh,
                                                                 uint64_t *__restrict__ eb)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_fold[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= waves) return;  // waves are independent: no workgroup barrier
    const int64_t j = (int64_t)blockIdx.x * waves + w;
    if (j >= nnew) return;
    if (skip_fresh && isnew[j] != 0) return;  // folded by fold_fresh_chunks_kernel
    unsigned char *st = smem_fold + (size_t)w * S.bytes;
    uint8_t *seen = st, *am = st + csz;
    uint4 *stage = (uint4 *)(st + S.roff);  // the batch's records, lane l at stage[l * NV]
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t r0 = ustart[j], r1 = ustart[j + 1];
    uint4 nxt[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) nxt[v] = (r0 + lane < r1) ? rec[(size_t)(r0 + lane) * NV + v] : make_uint4(0, 0, 0, 0);
    {
        const bool fresh = isnew[j] != 0;
        const int64_t src = fresh ? 0 : oldpos[j], dst = cpos[j];
        if (csz % 16 == 0) {
            for (int v = lane; v < csz / 16; v += 64) ((uint4 *)seen)[v] = make_uint4(0, 0, 0, 0);
        } else {
            for (int s = lane; s < csz; s += 64) seen[s] = 0;
        }
        for (uint32_t b0 = r0; b0 < r1; b0 += 64) {
            const uint32_t k = b0 + (uint32_t)lane;
            const bool valid = k < r1;
#pragma unroll
            for (int v = 0; v < NV; ++v) stage[lane * NV + v] = nxt[v];
            // the next 64 records of the chunk, in flight while these are folded
            const uint32_t kn = k + 64;
#pragma unroll
            for (int v = 0; v < NV; ++v) nxt[v] = (kn < r1) ? rec[(size_t)kn * NV + v] : make_uint4(0, 0, 0, 0);
            wave_sync_lds();
            const unsigned char *rp = (const unsigned char *)(stage + lane * NV);
            const int off = valid ? (int)*(const uint16_t *)rp : 0;
            const uint8_t bm = rp[2];
            uint64_t peers = __ballot(valid);
            for (int b = 0; b < obits; ++b) {
                const bool bit = (off >> b) & 1;
                const uint64_t bal = __ballot(bit);
                peers &= bit ? bal : ~bal;
            }
            const int rank = valid ? (int)__popcll(peers & lt) : -1;
            int maxr = 0;
            while (__ballot(rank > maxr) != 0) ++maxr;
            // the mask pass, rank by rank: whether the record is its slot's first and the OR
            // of the masks of the slot's earlier records
            bool first = false;
            uint8_t ao = 0;
            for (int r = 0; r <= maxr; ++r) {
                if (rank == r) {
                    first = seen[off] == 0;
                    ao = first ? (uint8_t)0 : am[off];
                    am[off] = (uint8_t)(ao | bm);
                    seen[off] = 0xff;
                }
                wave_sync_lds();
            }
            // the values, property by property (type and operator dispatched once per batch)
            for (int q = 0; q < F.np; ++q) {
                unsigned char *vals = st + S.voff[q];
                const unsigned char *bp = rp + R.off[q];
                const int op = F.op[q];
                switch (F.dt[q]) {
#define SG_CASE(CODE, T)                                                                                        \
    case CODE: fold_batch_prop<T>(op, vals, bp, off, rank, maxr, first, ao, bm); break;
                    SG_FOLD_TYPES(SG_CASE)
#undef SG_CASE
                }
            }
        }
        const uint8_t *smask = fresh ? nullptr : I.smask + src * csz;
        for (int q = 0; q < F.np; ++q) {
            const size_t cb = (size_t)csz * F.ts[q];
            const unsigned char *sp = fresh ? nullptr : I.src[q] + src * cb;
            unsigned char *dp = F.out[q] + dst * cb;
            switch (F.dt[q]) {
#define SG_CASE(CODE, T)                                                                                        \
    case CODE: {                                                                                                \
        T bgv;                                                                                                  \
        __builtin_memcpy(&bgv, &F.bg[q], sizeof(T));                                                            \
        if (fresh)                                                                                              \
            write_fold_prop<T, false>(F.op[q], bgv, st + S.voff[q], seen, am, nullptr, nullptr, (T *)dp, csz, lane); \
        else                                                                                                    \
            write_fold_prop<T, true>(F.op[q], bgv, st + S.voff[q], seen, am, (const T *)sp, smask, (T *)dp, csz,  \
                                     lane);                                                                     \
        break;                                                                                                  \
    }
                SG_FOLD_TYPES(SG_CASE)
#undef SG_CASE
            }
        }
        // the mask: old | OR of the records
        uint8_t *dm = omask + dst * csz;
        if (csz % 16 == 0) {
            typedef unsigned int v4u __attribute__((ext_vector_type(4)));
            for (int v = lane; v < csz / 16; v += 64) {
                v4u q = smask ? ((const v4u *)smask)[v] : (v4u){0u, 0u, 0u, 0u};
                // seen bytes are 0 or 0xff: the OR of the records' masks where a record came
                q |= ((const v4u *)am)[v] & ((const v4u *)seen)[v];
                __builtin_nontemporal_store(q, (v4u *)dm + v);
                if (eb) {
                    // 8^3 chunks: the vector holds rows 2v and 2v + 1, whose EXIST bits are
                    // bytes 2v and 2v + 1 of the chunk's 64 plane bytes
                    const uint64_t lo = ((uint64_t)q.y << 32) | q.x, hi = ((uint64_t)q.w << 32) | q.z;
                    ((uint16_t *)eb)[dst * 32 + v] = (uint16_t)(exist_byte(lo) | (exist_byte(hi) << 8));
                }
            }
        } else {
            for (int s = lane; s < csz; s += 64) dm[s] = (uint8_t)((smask ? smask[s] : 0) | (seen[s] ? am[s] : 0));
        }
    }
}

// ---- fold of the records of brand-new chunks -------------------------------------
// A new chunk has no old contents, so its LDS state can be the chunk's image itself:
// every array starts as the background (masks 0), the records fold into it, and it is
// copied out as 16-byte vectors.  Record fields come straight from the registers the
// records were loaded into (no staging), the masks are ORed in by LDS atomics (order-
// free), and only the values of a slot reached by several records of one batch are
// folded rank by rank, in lane (= insertion) order.

struct FreshProp {
    int dt, op, ts, roff;  // type code, reduction, value bytes, byte offset in the record
    int voff, pad;         // LDS image offset of the property's chunk
    unsigned char *out;    // chunk array
    uint64_t pat;          // background replicated over 8 bytes
};

struct FreshArgs {
    int np;
    int moff, soff;  // LDS offsets of the mask image and the seen bits
    int bytes;       // LDS bytes per wave
    FreshProp p[kMaxProps];
};

// word i (uniform) of a record held in NV 16-byte registers
template <int NV> __device__ inline uint32_t rec_word(const uint4 *r, int i)
{
    const uint4 v = (NV == 1 || i < 4) ? r[0] : r[NV - 1];
    const int k = i & 3;
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// the value of a property at byte offset roff (uniform, naturally aligned)
template <typename T, int NV> __device__ inline T rec_value(const uint4 *r, int roff)
{
    T t;
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = (uint64_t)rec_word<NV>(r, roff >> 2) | ((uint64_t)rec_word<NV>(r, (roff >> 2) + 1) << 32);
        __builtin_memcpy(&t, &u, 8);
    } else if constexpr (sizeof(T) == 4) {
        const uint32_t u = rec_word<NV>(r, roff >> 2);
        __builtin_memcpy(&t, &u, 4);
    } else {
        const uint32_t u = rec_word<NV>(r, roff >> 2) >> (8 * (roff & 3));
        if constexpr (sizeof(T) == 2) {
            const uint16_t h = (uint16_t)u;
            __builtin_memcpy(&t, &h, 2);
        } else {
            const uint8_t h = (uint8_t)u;
            __builtin_memcpy(&t, &h, 1);
        }
    }
    return t;
}

// One property of one batch into the chunk image.  first: the record is its slot's first
// (its value is taken whatever the masks); aE: an earlier record of the slot carries
// EXIST; bE: this one does.  The same rule as the old-chunk fold's record rounds.
template <typename T, int OP, int NV>
__device__ inline void fresh_prop(unsigned char *img, const uint4 *r, int roff, int off, int rank, int maxr,
                                  bool first, bool aE, bool bE)
{
    T *a = (T *)img + off;
    const T b = rec_value<T, NV>(r, roff);
    for (int k = 0; k <= maxr; ++k) {
        if (rank == k) {
            const T cur = *a;
            T v;
            if (OP == SG_OP_NONE) v = bE ? b : cur;
            else v = (aE && bE) ? dev_reduce<T>(OP, cur, b) : (bE ? b : cur);
            *a = first ? b : v;
        }
        if (k < maxr) wave_sync_lds();
    }
}

template <typename T, int NV>
__device__ inline void fresh_prop_op(int op, unsigned char *img, const uint4 *r, int roff, int off, int rank, int maxr,
                                     bool first, bool aE, bool bE)
{
    switch (op) {
    case SG_OP_MAX: fresh_prop<T, SG_OP_MAX, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    case SG_OP_MIN: fresh_prop<T, SG_OP_MIN, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    case SG_OP_ADD: fresh_prop<T, SG_OP_ADD, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    case SG_OP_RIGHT: fresh_prop<T, SG_OP_RIGHT, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    case SG_OP_LEFT: fresh_prop<T, SG_OP_LEFT, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    case SG_OP_OR: fresh_prop<T, SG_OP_OR, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    default: fresh_prop<T, SG_OP_NONE, NV>(img, r, roff, off, rank, maxr, first, aE, bE); break;
    }
}

// the record's value bits of one property into the image, when every record of the batch
// is its slot's first (no fold: only the value size matters)
template <int NV> __device__ inline void fresh_store_bits(unsigned char *img, const uint4 *r, int roff, int ts, int off)
{
    switch (ts) {
    case 8: ((uint64_t *)img)[off] = rec_value<uint64_t, NV>(r, roff); break;
    case 4: ((uint32_t *)img)[off] = rec_value<uint32_t, NV>(r, roff); break;
    case 2: ((uint16_t *)img)[off] = rec_value<uint16_t, NV>(r, roff); break;
    default: img[off] = rec_value<uint8_t, NV>(r, roff); break;
    }
}

// f(q) for every property: unrolled when the count NP is known at compile time
template <int Q, int N, typename F> __device__ inline void props_unrolled(F &f)
{
    if constexpr (Q < N) {
        f(Q);
        props_unrolled<Q + 1, N>(f);
    }
}

template <int NP, typename F> __device__ inline void for_props(int np, F &&f)
{
    if constexpr (NP > 0) {
        props_unrolled<0, NP>(f);
    } else {
        for (int q = 0; q < np; ++q) f(q);
    }
}

// One wave per touched chunk; chunks that existed before (isnew[j] == 0) are left to
// fold_chunk_records_kernel, their waves return at once (isnew == nullptr: every
// touched chunk is new).  NP > 0: the property count, known at compile time (the
// per-property loops unroll and their parameters are loaded once); NP == 0: A.np.
// Needs csz % 16 == 0.
template <int NV, int NP>
__global__ __launch_bounds__(256) void fold_fresh_chunks_kernel(const uint4 *__restrict__ rec,
                                                                const uint32_t *__restrict__ ustart,
                                                                const int64_t *__restrict__ cpos,
                                                                const uint32_t *__restrict__ isnew, int64_t nnew,
                                                                int csz, int obits, int waves, FreshArgs A,
                                                                uint8_t *__restrict__ omask, uint64_t *__restrict__ eb)
{
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_fresh[];
    const int np = NP > 0 ? NP : A.np;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= waves) return;  // waves are independent: no workgroup barrier
    const int64_t j = (int64_t)blockIdx.x * waves + w;
    if (j >= nnew) return;
    if (isnew && isnew[j] == 0) return;
    unsigned char *img = smem_fresh + (size_t)w * A.bytes;
    uint8_t *mimg = img + A.moff;
    uint32_t *seen = (uint32_t *)(img + A.soff);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t r0 = ustart[j], r1 = ustart[j + 1];
    uint4 nxt[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) nxt[v] = (r0 + lane < r1) ? rec[(size_t)(r0 + lane) * NV + v] : make_uint4(0, 0, 0, 0);
    // the image: background values, masks 0, no slot seen
    for_props<NP>(np, [&](int q) {
        const FreshProp &P = A.p[q];
        const uint32_t lo = (uint32_t)P.pat, hi = (uint32_t)(P.pat >> 32);
        const v4u pat = {lo, hi, lo, hi};
        v4u *d = (v4u *)(img + P.voff);
        for (int v = lane; v < csz * P.ts / 16; v += 64) d[v] = pat;
    });
    for (int v = lane; v < csz / 16; v += 64) ((v4u *)mimg)[v] = (v4u){0u, 0u, 0u, 0u};
    for (int v = lane; v < (csz + 31) / 32; v += 64) seen[v] = 0u;
    wave_sync_lds();
    for (uint32_t b0 = r0; b0 < r1; b0 += 64) {
        uint4 cur[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) cur[v] = nxt[v];
        const uint32_t k = b0 + (uint32_t)lane;
        const bool valid = k < r1;
        const uint32_t kn = k + 64;
#pragma unroll
        for (int v = 0; v < NV; ++v) nxt[v] = (kn < r1) ? rec[(size_t)kn * NV + v] : make_uint4(0, 0, 0, 0);
        const int off = valid ? (int)(cur[0].x & 0xffffu) : 0;
        const uint8_t m = (uint8_t)(cur[0].x >> 16);
        // records of this batch that share the slot: ballots over the slot bits
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < obits; ++b) {
            const bool bit = (off >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = valid ? (int)__popcll(peers & lt) : -1;
        int maxr = 0;
        while (__ballot(rank > maxr) != 0) ++maxr;
        // the slot's state before this batch, read before any update of the batch
        const bool pseen = valid && ((seen[off >> 5] >> (off & 31)) & 1u);
        const bool pE = valid && (mimg[off] & 1);
        wave_sync_lds();
        if (valid) {
            atomicOr((uint32_t *)mimg + (off >> 2), (uint32_t)m << (8 * (off & 3)));
            if (rank == 0 && !pseen) atomicOr(seen + (off >> 5), 1u << (off & 31));
        }
        const bool first = rank == 0 && !pseen;
        if (maxr == 0 && __ballot(valid && !first) == 0) {
            // one record per slot, none reached before: the values are the records'
            for_props<NP>(np, [&](int q) {
                const FreshProp &P = A.p[q];
                if (valid) fresh_store_bits<NV>(img + P.voff, cur, P.roff, P.ts, off);
            });
        } else {
            const bool bE = valid && (m & 1);
            const uint64_t exist = __ballot(bE);
            const bool aE = pE || ((exist & peers & lt) != 0);
            for_props<NP>(np, [&](int q) {
                const FreshProp &P = A.p[q];
                unsigned char *vimg = img + P.voff;
                switch (P.dt) {
#define SG_CASE(CODE, T)                                                                                        \
    case CODE: fresh_prop_op<T, NV>(P.op, vimg, cur, P.roff, off, rank, maxr, first, aE, bE); break;
                    SG_FOLD_TYPES(SG_CASE)
#undef SG_CASE
                }
            });
        }
        wave_sync_lds();
    }
    // slots reached only by records without any mask bit hold the background (the old-
    // chunk fold gives the same with no old point); 8 slots per lane and step
    for (int s0 = lane * 8; s0 < csz; s0 += 512) {
        const uint32_t sb = (seen[s0 >> 5] >> (s0 & 31)) & 0xffu;
        const uint64_t mb = *(const uint64_t *)(mimg + s0);
        uint32_t fix = 0;
#pragma unroll
        for (int x = 0; x < 8; ++x) fix |= (((mb >> (8 * x)) & 0xffu) == 0 ? 1u : 0u) << x;
        fix &= sb;
        while (fix) {
            const int s = s0 + __builtin_ctz(fix);
            fix &= fix - 1;
            for (int q = 0; q < np; ++q) {
                const FreshProp &P = A.p[q];
                unsigned char *vp = img + P.voff + (size_t)s * P.ts;
                switch (P.ts) {
                case 8: *(uint64_t *)vp = P.pat; break;
                case 4: *(uint32_t *)vp = (uint32_t)P.pat; break;
                case 2: *(uint16_t *)vp = (uint16_t)P.pat; break;
                default: *vp = (unsigned char)P.pat; break;
                }
            }
        }
    }
    wave_sync_lds();
    // the image out, every array's chunk as whole 16-byte vectors
    const int64_t dst = cpos[j];
    if (eb) {
        // 8^3 chunks: EXIST bit-plane word z has row (y, z)'s bits as byte y, i.e. byte
        // l = 8z + y of the chunk's 64 plane bytes is lane l's row
        const uint64_t row = *(const uint64_t *)(mimg + lane * 8);
        ((uint8_t *)eb)[dst * 64 + lane] = exist_byte(row);
    }
    for_props<NP>(np, [&](int q) {
        const FreshProp &P = A.p[q];
        const int nvec = csz * P.ts / 16;
        const v4u *s = (const v4u *)(img + P.voff);
        v4u *d = (v4u *)(P.out + (size_t)dst * csz * P.ts);
        for (int v = lane; v < nvec; v += 64) __builtin_nontemporal_store(s[v], d + v);
    });
    {
        const v4u *s = (const v4u *)mimg;
        v4u *d = (v4u *)(omask + dst * csz);
        for (int v = lane; v < csz / 16; v += 64) __builtin_nontemporal_store(s[v], d + v);
    }
}

__global__ void uchunk_index_kernel(const uint32_t *hscan, const uint32_t *head, int64_t nu, uint32_t *uchunk)
{
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nu) uchunk[u] = hscan[u] + head[u] - 1;
}

// ---- whole-chunk records --------------------------------------------------------
// A chunk record takes part in the fold of the slots where its mask is non-zero (an
// all-zero record only creates its chunk), as the oracle's restatement of insertBlock.
// one workgroup per touched output chunk u: records perm[rstart[u] .. rstart[u+1]) in
// insertion order, folded per slot as fold_points_kernel folds a point's records
template <typename T>
__global__ __launch_bounds__(kWG) void fold_chunks_kernel(const uint32_t *rstart, const uint32_t *perm,
                                                         const int64_t *cpos, int csz, const uint8_t *rmask,
                                                         const T *rval, T *out,
                                                         const uint8_t *omask, int op, T bgv)
{
    const int64_t u = blockIdx.x;
    const uint32_t r0 = rstart[u], r1 = rstart[u + 1];
    const int64_t base = cpos[u] * csz;
    for (int o = threadIdx.x; o < csz; o += kWG) {
        bool have = false;
        T a = bgv;
        uint8_t am = 0;
        for (uint32_t i = r0; i < r1; ++i) {
            const uint32_t rec = perm[i];
            const uint8_t bm = rmask[(int64_t)rec * csz + o];
            if (bm == 0) continue;
            const T b = rval[(int64_t)rec * csz + o];
            if (!have) {
                a = b;
                am = bm;
                have = true;
                continue;
            }
            const bool aE = am & 1, bE = bm & 1;
            if (op == SG_OP_NONE) {
                if (bE) a = b;
            } else if (aE && bE) {
                a = dev_reduce<T>(op, a, b);
            } else if (bE) {
                a = b;
            }
            am |= bm;
        }
        if (!have) continue;
        const int64_t slot = base + o;
        const uint8_t om = omask[slot];
        const bool aE = am & 1, oE = om & 1;
        if (op == SG_OP_NONE) {
            if (!aE && oE) a = out[slot];
        } else if (aE && oE) {
            a = dev_reduce<T>(op, a, out[slot]);
        } else if (oE) {
            a = out[slot];
        }
        out[slot] = ((am | om) == 0) ? bgv : a;
    }
}

__global__ __launch_bounds__(kWG) void fold_chunk_masks_kernel(const uint32_t *rstart, const uint32_t *perm,
                                                              const int64_t *cpos, int csz, const uint8_t *rmask,
                                                              uint8_t *omask)
{
    const int64_t u = blockIdx.x;
    const uint32_t r0 = rstart[u], r1 = rstart[u + 1];
    const int64_t base = cpos[u] * csz;
    for (int o = threadIdx.x; o < csz; o += kWG) {
        uint8_t am = 0;
        for (uint32_t i = r0; i < r1; ++i) am |= rmask[(int64_t)perm[i] * csz + o];
        omask[base + o] |= am;
    }
}

template <typename F> void dispatch_type(int dt, F &&f)
{
    switch (dt) {
    case SG_F32: f(float()); break;
    case SG_F64: f(double()); break;
    case SG_I32: f(int32_t()); break;
    case SG_U32: f(uint32_t()); break;
    case SG_I64: f(int64_t()); break;
    case SG_U64: f(uint64_t()); break;
    case SG_I8: f(int8_t()); break;
    case SG_U8: f(uint8_t()); break;
    case SG_I16: f(int16_t()); break;
    case SG_U16: f(uint16_t()); break;
    }
}

// records whose key lies in [0, limit)
__global__ void valid_key_kernel(const int64_t *keys, int64_t n, int64_t limit, uint32_t *flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (keys[i] >= 0 && keys[i] < limit) ? 1u : 0u;
}

// stable compaction of ts-byte elements: src[i] -> dst[pos[i]] where flag[i]
__global__ void compact_bytes_kernel(const uint8_t *src, uint8_t *dst, const uint32_t *flag, const uint32_t *pos,
                                     int64_t n, int ts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    for (int b = 0; b < ts; ++b) dst[(int64_t)pos[i] * ts + b] = src[i * ts + b];
}

int grow_insert_buffer(sg_grid *g, int64_t need)
{
    if (need <= g->ins_cap) return SG_OK;
    int64_t ncap = std::max<int64_t>(need, std::max<int64_t>(1024, g->ins_cap * 2));
    int64_t *nk = (int64_t *)g->pool.alloc(ncap * sizeof(int64_t));
    uint8_t *nm = (uint8_t *)g->pool.alloc(ncap);
    if (!nk || !nm) return fail(SG_E_OOM, "insert buffer allocation failed");
    void *nv[kMaxProps] = {};
    for (int p = 0; p < g->np; ++p) {
        nv[p] = g->pool.alloc(ncap * g->ts[p]);
        if (!nv[p]) return fail(SG_E_OOM, "insert buffer allocation failed");
    }
    if (g->ins_n) {
        SG_HIP(hipMemcpyAsync(nk, g->d_ins_key, g->ins_n * sizeof(int64_t), hipMemcpyDeviceToDevice, g->stream));
        SG_HIP(hipMemcpyAsync(nm, g->d_ins_mask, g->ins_n, hipMemcpyDeviceToDevice, g->stream));
        for (int p = 0; p < g->np; ++p)
            SG_HIP(hipMemcpyAsync(nv[p], g->d_ins_val[p], g->ins_n * g->ts[p], hipMemcpyDeviceToDevice, g->stream));
    }
    g->pool.free(g->d_ins_key);
    g->pool.free(g->d_ins_mask);
    for (int p = 0; p < g->np; ++p) g->pool.free(g->d_ins_val[p]);
    g->d_ins_key = nk;
    g->d_ins_mask = nm;
    for (int p = 0; p < g->np; ++p) g->d_ins_val[p] = nv[p];
    g->ins_cap = ncap;
    return SG_OK;
}

// drop the point records whose key is outside [0, limit) from the insert buffer, keeping
// the others in order; *kept receives their count
int drop_invalid_records(sg_grid *g, int64_t limit, int64_t *kept)
{
    const int64_t n = g->ins_n;
    int rc;
    DevBuf &b = g->scratch[8];
    if ((rc = ensure(g->pool, b, (size_t)n * 8 + 64))) return rc;
    uint32_t *flag = (uint32_t *)b.p, *pos = flag + n;
    hipLaunchKernelGGL(valid_key_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_ins_key, n, limit, flag);
    SG_HIP(hipGetLastError());
    uint32_t nv = 0;
    if ((rc = exclusive_scan_u32(g, flag, pos, n, &nv))) return rc;
    int64_t *nk = (int64_t *)g->pool.alloc(g->ins_cap * sizeof(int64_t));
    uint8_t *nm = (uint8_t *)g->pool.alloc(g->ins_cap);
    if (!nk || !nm) return fail(SG_E_OOM, "insert buffer allocation failed");
    void *nvl[kMaxProps] = {};
    for (int p = 0; p < g->np; ++p)
        if (!(nvl[p] = g->pool.alloc(g->ins_cap * g->ts[p]))) return fail(SG_E_OOM, "insert buffer allocation failed");
    const dim3 grid(grid_1d(n, kWG)), blk(kWG);
    hipLaunchKernelGGL(compact_bytes_kernel, grid, blk, 0, g->stream, (const uint8_t *)g->d_ins_key, (uint8_t *)nk,
                       flag, pos, n, 8);
    hipLaunchKernelGGL(compact_bytes_kernel, grid, blk, 0, g->stream, (const uint8_t *)g->d_ins_mask, nm, flag, pos, n,
                       1);
    for (int p = 0; p < g->np; ++p)
        hipLaunchKernelGGL(compact_bytes_kernel, grid, blk, 0, g->stream, (const uint8_t *)g->d_ins_val[p],
                           (uint8_t *)nvl[p], flag, pos, n, g->ts[p]);
    SG_HIP(hipGetLastError());
    g->pool.free(g->d_ins_key);
    g->pool.free(g->d_ins_mask);
    for (int p = 0; p < g->np; ++p) g->pool.free(g->d_ins_val[p]);
    g->d_ins_key = nk;
    g->d_ins_mask = nm;
    for (int p = 0; p < g->np; ++p) g->d_ins_val[p] = nvl[p];
    g->ins_n = nv;
    *kept = nv;
    return SG_OK;
}

hipMemcpyKind kind_from(int mem) { return mem == SG_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice; }

// one past the largest point key of the domain (chunk ids grow with every chunk
// coordinate under both linearizers, so the last chunk has the largest id)
int64_t key_limit(const sg_grid *g)
{
    int64_t last[3] = {0, 0, 0};
    for (int d = 0; d < g->dim; ++d) last[d] = g->bsz[d] - 1;
    return (chunk_id(chunk_lin(g), last) + 1) * g->csz;
}

// The grouped flush's record layout for this grid's properties: slot offset (2 bytes),
// mask (1 byte), then each property at its natural alignment, in 16 or 32 bytes; S: the
// fold's LDS state per wave; cbits: significant bits of a chunk id.  False when the
// grouped flush does not take this grid.
bool point_record_layout(const sg_grid *g, RecLayout *R, FoldState *S, int *cbits)
{
    const int csz = g->csz;
    const int64_t nblocks = key_limit(g) / csz;
    int cb = 0;
    while (cb < 63 && (((int64_t)1 << cb) < nblocks)) ++cb;
    *cbits = cb;
    if (cb > 32 || csz > 65536) return false;
    int at = 3;
    for (int p = 0; p < g->np; ++p) {
        at = (at + g->ts[p] - 1) / g->ts[p] * g->ts[p];
        R->off[p] = at;
        at += g->ts[p];
    }
    if (at > 32) return false;
    R->nv = at <= 16 ? 1 : 2;
    R->bytes = 16 * R->nv;
    size_t b = (size_t)2 * csz;
    for (int p = 0; p < g->np; ++p) {
        b = (b + 15) / 16 * 16;
        S->voff[p] = (int)b;
        b += (size_t)csz * g->ts[p];
    }
    b = (b + 15) / 16 * 16;
    S->roff = (int)b;
    b += (size_t)64 * R->bytes;
    if (b > 65536) return false;
    S->bytes = (int)b;
    return true;
}

// room for `need` staged (chunk id, record) pairs of nv 16-byte vectors, the staged ones kept
int grow_point_records(sg_grid *g, int64_t need, int nv)
{
    if (need <= g->pins_cap && g->pins_nv == nv) return SG_OK;
    const int64_t ncap = std::max<int64_t>(need, std::max<int64_t>(1024, g->pins_cap * 2));
    uint32_t *nck = (uint32_t *)g->pool.alloc(ncap * sizeof(uint32_t));
    void *nrec = g->pool.alloc(ncap * 16 * nv);
    if (!nck || !nrec) return fail(SG_E_OOM, "insert buffer allocation failed");
    if (!g->d_pins_bad && !(g->d_pins_bad = (int *)g->pool.alloc(64)))
        return fail(SG_E_OOM, "insert buffer allocation failed");
    if (g->pins_n) {
        SG_HIP(hipMemcpyAsync(nck, g->d_pins_ck, g->pins_n * sizeof(uint32_t), hipMemcpyDeviceToDevice, g->stream));
        SG_HIP(hipMemcpyAsync(nrec, g->d_pins_rec, g->pins_n * 16 * nv, hipMemcpyDeviceToDevice, g->stream));
    }
    g->pool.free(g->d_pins_ck);
    g->pool.free(g->d_pins_rec);
    g->d_pins_ck = nck;
    g->d_pins_rec = nrec;
    g->pins_cap = ncap;
    g->pins_nv = nv;
    return SG_OK;
}

// append masks and values of n records at insert position `at`
int append_mask_values(sg_grid *g, int64_t at, int64_t n, const uint8_t *masks, const void *const *values, int mem)
{
    if (masks) SG_HIP(hipMemcpyAsync(g->d_ins_mask + at, masks, n, kind_from(mem), g->stream));
    else {
        hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_ins_mask + at,
                           (uint8_t)SG_EXIST, n);
        SG_HIP(hipGetLastError());
    }
    for (int p = 0; p < g->np; ++p) {
        unsigned char *dst = (unsigned char *)g->d_ins_val[p] + at * g->ts[p];
        if (values && values[p]) {
            SG_HIP(hipMemcpyAsync(dst, values[p], n * g->ts[p], kind_from(mem), g->stream));
        } else {
            dispatch_type(g->dt[p], [&](auto tag) {
                typedef decltype(tag) T;
                T v;
                memcpy(&v, g->bg[p], sizeof(T));
                hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, (T *)dst, v, n);
            });
            SG_HIP(hipGetLastError());
        }
    }
    return SG_OK;
}

}  // namespace

int convert_point_records(sg_grid *g)
{
    if (g->pins_n == 0) return SG_OK;
    const int64_t n = g->pins_n;
    RecLayout R;
    FoldState S;
    int cbits = 0;
    if (!point_record_layout(g, &R, &S, &cbits) || R.nv != g->pins_nv)
        return fail(SG_E_STATE, "staged point records of another layout");
    int rc = grow_insert_buffer(g, g->ins_n + n);
    if (rc) return rc;
    const int64_t at = g->ins_n;
    StageVals V;
    V.np = g->np;
    for (int p = 0; p < g->np; ++p) {
        V.ts[p] = g->ts[p];
        V.src[p] = nullptr;
        V.dst[p] = (unsigned char *)g->d_ins_val[p] + at * g->ts[p];
        V.bg[p] = 0;
    }
    hipLaunchKernelGGL(point_records_to_buffer_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_pins_ck,
                       (const unsigned char *)g->d_pins_rec, n, g->csz, R, V, g->d_ins_key + at, g->d_ins_mask + at);
    SG_HIP(hipGetLastError());
    g->ins_n += n;
    g->pins_n = 0;
    return SG_OK;
}

int append_points(sg_grid *g, int64_t n, const int64_t *keys, const uint8_t *masks, const void *const *values,
                  int mem)
{
    if (n < 0) return fail(SG_E_ARG, "negative record count");
    if (n == 0) return SG_OK;
    if (!keys) return fail(SG_E_ARG, "null keys");
    const int64_t limit = key_limit(g);
    if (mem == SG_MEM_HOST) {
        for (int64_t i = 0; i < n; ++i)
            if (keys[i] < 0 || keys[i] >= limit) return fail(SG_E_RANGE, "point key outside the grid");
    }
    int rc = convert_chunk_records(g);
    if (!rc) rc = convert_point_records(g);
    if (rc) return rc;
    rc = grow_insert_buffer(g, g->ins_n + n);
    if (rc) return rc;
    const int64_t at = g->ins_n;
    SG_HIP(hipMemcpyAsync(g->d_ins_key + at, keys, n * sizeof(int64_t), kind_from(mem), g->stream));
    rc = append_mask_values(g, at, n, masks, values, mem);
    if (rc) return rc;
    g->ins_n += n;
    return SG_OK;
}

int append_coords(sg_grid *g, int64_t n, const int32_t *coords, const uint8_t *masks, const void *const *values,
                  int mem)
{
    if (n < 0) return fail(SG_E_ARG, "negative record count");
    if (n == 0) return SG_OK;
    if (!coords) return fail(SG_E_ARG, "null coords");
    int rc = convert_chunk_records(g);
    if (rc) return rc;
    if (mem == SG_MEM_DEVICE && g->ins_n == 0 && g->flush_path == 0 && g->pins_n + n < ((int64_t)1 << 31)) {
        // nothing else pending: stage the records in the grouped flush's format
        RecLayout R;
        FoldState S;
        int cbits = 0;
        if (point_record_layout(g, &R, &S, &cbits) && cbits <= 31) {
            if ((rc = grow_point_records(g, g->pins_n + n, R.nv))) return rc;
            if (g->pins_n == 0) SG_HIP(hipMemsetAsync(g->d_pins_bad, 0, sizeof(int), g->stream));
            StageVals V;
            V.np = g->np;
            for (int p = 0; p < g->np; ++p) {
                V.ts[p] = g->ts[p];
                V.src[p] = values ? (const unsigned char *)values[p] : nullptr;
                V.dst[p] = nullptr;
                V.bg[p] = 0;
                memcpy(&V.bg[p], g->bg[p], g->ts[p]);
            }
            int szi[3];
            for (int d = 0; d < 3; ++d) szi[d] = (int)std::min<int64_t>(g->sz[d], INT32_MAX);
            uint4 *rd = (uint4 *)g->d_pins_rec + g->pins_n * R.nv;
            if (R.nv == 1)
                hipLaunchKernelGGL(stage_coord_records_kernel<1>, dim3(grid_1d(n, kWG * kStageItems)), dim3(kWG), 0,
                                   g->stream, coords, masks, n, g->dim, g->edge, szi[0], szi[1], szi[2], chunk_lin(g),
                                   V, R, g->d_pins_ck + g->pins_n, rd, g->d_pins_bad);
            else
                hipLaunchKernelGGL(stage_coord_records_kernel<2>, dim3(grid_1d(n, kWG * kStageItems)), dim3(kWG), 0,
                                   g->stream, coords, masks, n, g->dim, g->edge, szi[0], szi[1], szi[2], chunk_lin(g),
                                   V, R, g->d_pins_ck + g->pins_n, rd, g->d_pins_bad);
            SG_HIP(hipGetLastError());
            g->pins_n += n;
            return SG_OK;
        }
    }
    if ((rc = convert_point_records(g))) return rc;
    rc = grow_insert_buffer(g, g->ins_n + n);
    if (rc) return rc;
    const int64_t at = g->ins_n;
    const int32_t *dc = coords;
    DevBuf &cb = g->scratch[9];
    if (mem == SG_MEM_HOST) {
        for (int64_t i = 0; i < n * g->dim; ++i) {
            int d = (int)(i % g->dim);
            if (coords[i] < 0 || coords[i] >= g->sz[d]) return fail(SG_E_RANGE, "coordinate outside the grid");
        }
        rc = ensure(g->pool, cb, n * g->dim * sizeof(int32_t) + 64);
        if (rc) return rc;
        SG_HIP(hipMemcpyAsync(cb.p, coords, n * g->dim * sizeof(int32_t), hipMemcpyHostToDevice, g->stream));
        dc = (const int32_t *)cb.p;
    } else {
        // device arrays: keys, masks and values in one pass
        StageVals V;
        V.np = g->np;
        for (int p = 0; p < g->np; ++p) {
            V.ts[p] = g->ts[p];
            V.src[p] = values ? (const unsigned char *)values[p] : nullptr;
            V.dst[p] = (unsigned char *)g->d_ins_val[p] + at * g->ts[p];
            V.bg[p] = 0;
            memcpy(&V.bg[p], g->bg[p], g->ts[p]);
        }
        int szi[3];
        for (int d = 0; d < 3; ++d) szi[d] = (int)std::min<int64_t>(g->sz[d], INT32_MAX);
        hipLaunchKernelGGL(stage_coords_kernel, dim3(grid_1d(n, kWG * kStageItems)), dim3(kWG), 0, g->stream, dc,
                           masks, n, g->dim,
                           g->edge, szi[0], szi[1], szi[2], chunk_lin(g), g->csz, g->d_ins_key + at,
                           g->d_ins_mask + at, V);
        SG_HIP(hipGetLastError());
        g->ins_n += n;
        return SG_OK;
    }
    hipLaunchKernelGGL(coords_to_keys_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, dc,
                       g->d_ins_key + at, n, g->dim, g->edge, g->sz[0], g->sz[1], g->sz[2], chunk_lin(g), g->csz);
    SG_HIP(hipGetLastError());
    // device coordinates outside the grid became key -1: the next flush drops those records
    rc = append_mask_values(g, at, n, masks, values, mem);
    if (rc) return rc;
    g->ins_n += n;
    return SG_OK;
}

namespace {

// whole-chunk records -> point records appended to the point insert buffer
int expand_chunks(sg_grid *g, int64_t n, const int64_t *chunk_ids, const uint8_t *masks, const void *const *values,
                  int mem)
{
    const int csz = g->csz;
    // stage host chunk data on the device
    size_t bytes = 0;
    if (mem == SG_MEM_HOST) {
        bytes = n * sizeof(int64_t) + n * csz;
        for (int p = 0; p < g->np; ++p) bytes += (values && values[p]) ? n * csz * g->ts[p] + 16 : 0;
    }
    DevBuf &st = g->scratch[8];
    int rc = ensure(g->pool, st, bytes + 256 + (size_t)(n + 1) * 4 * 2);
    if (rc) return rc;
    char *q = (char *)st.p;
    const int64_t *dids = chunk_ids;
    const uint8_t *dm = masks;
    const void *dv[kMaxProps] = {};
    if (mem == SG_MEM_HOST) {
        SG_HIP(hipMemcpyAsync(q, chunk_ids, n * sizeof(int64_t), hipMemcpyHostToDevice, g->stream));
        dids = (const int64_t *)q;
        q += n * sizeof(int64_t);
        SG_HIP(hipMemcpyAsync(q, masks, n * csz, hipMemcpyHostToDevice, g->stream));
        dm = (const uint8_t *)q;
        q += n * csz;
        for (int p = 0; p < g->np; ++p) {
            if (values && values[p]) {
                q = (char *)(((uintptr_t)q + 15) & ~(uintptr_t)15);
                SG_HIP(hipMemcpyAsync(q, values[p], n * csz * g->ts[p], hipMemcpyHostToDevice, g->stream));
                dv[p] = q;
                q += n * csz * g->ts[p];
            }
        }
    } else {
        for (int p = 0; p < g->np; ++p) dv[p] = (values ? values[p] : nullptr);
    }
    q = (char *)(((uintptr_t)q + 15) & ~(uintptr_t)15);
    uint32_t *cnt = (uint32_t *)q;
    uint32_t *start = cnt + (n + 1);
    hipLaunchKernelGGL(chunk_count_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, dm, csz, cnt);
    SG_HIP(hipGetLastError());
    uint32_t total = 0;
    rc = exclusive_scan_u32(g, cnt, start, n, &total);
    if (rc) return rc;
    rc = grow_insert_buffer(g, g->ins_n + total);
    if (rc) return rc;
    const int64_t at = g->ins_n;
    DevBuf &sb = g->scratch[7];
    rc = ensure(g->pool, sb, (size_t)total * 4 + 64);
    if (rc) return rc;
    uint32_t *src = (uint32_t *)sb.p;
    hipLaunchKernelGGL(chunk_expand_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, dids, dm, start, csz,
                       g->d_ins_key + at, g->d_ins_mask + at, src, (int64_t)0);
    SG_HIP(hipGetLastError());
    for (int p = 0; p < g->np; ++p) {
        void *dst = (unsigned char *)g->d_ins_val[p] + at * g->ts[p];
        dispatch_type(g->dt[p], [&](auto tag) {
            typedef decltype(tag) T;
            if (dv[p]) {
                hipLaunchKernelGGL(gather_kernel<T>, dim3(grid_1d(total, kWG)), dim3(kWG), 0, g->stream,
                                   (const T *)dv[p], src, (T *)dst, (int64_t)total);
            } else {
                T v;
                memcpy(&v, g->bg[p], sizeof(T));
                hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_1d(total, kWG)), dim3(kWG), 0, g->stream, (T *)dst, v,
                                   (int64_t)total);
            }
        });
        SG_HIP(hipGetLastError());
    }
    g->ins_n += total;
    // the staging buffers are reused by the next call on this stream; keep them alive
    return SG_OK;
}

int grow_chunk_buffer(sg_grid *g, int64_t need)
{
    if (need <= g->cins_cap) return SG_OK;
    const int csz = g->csz;
    int64_t ncap = std::max<int64_t>(need, std::max<int64_t>(16, g->cins_cap * 2));
    int64_t *ni = (int64_t *)g->pool.alloc(ncap * sizeof(int64_t));
    uint8_t *nm = (uint8_t *)g->pool.alloc((size_t)ncap * csz);
    if (!ni || !nm) return fail(SG_E_OOM, "chunk insert buffer allocation failed");
    void *nv[kMaxProps] = {};
    for (int p = 0; p < g->np; ++p) {
        nv[p] = g->pool.alloc((size_t)ncap * csz * g->ts[p]);
        if (!nv[p]) return fail(SG_E_OOM, "chunk insert buffer allocation failed");
    }
    if (g->cins_n) {
        const int64_t n = g->cins_n;
        SG_HIP(hipMemcpyAsync(ni, g->d_cins_id, n * sizeof(int64_t), hipMemcpyDeviceToDevice, g->stream));
        SG_HIP(hipMemcpyAsync(nm, g->d_cins_mask, (size_t)n * csz, hipMemcpyDeviceToDevice, g->stream));
        for (int p = 0; p < g->np; ++p)
            SG_HIP(hipMemcpyAsync(nv[p], g->d_cins_val[p], (size_t)n * csz * g->ts[p], hipMemcpyDeviceToDevice,
                                  g->stream));
    }
    g->pool.free(g->d_cins_id);
    g->pool.free(g->d_cins_mask);
    for (int p = 0; p < g->np; ++p) g->pool.free(g->d_cins_val[p]);
    g->d_cins_id = ni;
    g->d_cins_mask = nm;
    for (int p = 0; p < g->np; ++p) g->d_cins_val[p] = nv[p];
    g->cins_cap = ncap;
    return SG_OK;
}

}  // namespace

int convert_chunk_records(sg_grid *g)
{
    if (g->cins_n == 0) return SG_OK;
    const int64_t n = g->cins_n;
    const void *vals[kMaxProps];
    for (int p = 0; p < g->np; ++p) vals[p] = g->d_cins_val[p];
    int rc = expand_chunks(g, n, g->d_cins_id, g->d_cins_mask, vals, SG_MEM_DEVICE);
    if (rc) return rc;
    g->cins_n = 0;
    return SG_OK;
}

int reserve_point_records(sg_grid *g, int64_t n, int64_t *at)
{
    int rc = convert_chunk_records(g);
    if (!rc) rc = convert_point_records(g);
    if (rc) return rc;
    rc = grow_insert_buffer(g, g->ins_n + n);
    if (rc) return rc;
    *at = g->ins_n;
    return SG_OK;
}

int append_chunks(sg_grid *g, int64_t n, const int64_t *chunk_ids, const uint8_t *masks, const void *const *values,
                  int mem)
{
    if (n < 0) return fail(SG_E_ARG, "negative record count");
    if (n == 0) return SG_OK;
    if (!chunk_ids || !masks) return fail(SG_E_ARG, "null chunk ids or masks");
    const int csz = g->csz;
    const int64_t nblocks = key_limit(g) / csz;
    if (mem == SG_MEM_HOST)
        for (int64_t i = 0; i < n; ++i)
            if (chunk_ids[i] < 0 || chunk_ids[i] >= nblocks) return fail(SG_E_RANGE, "chunk id outside the grid");
    // after point records, chunk records join them as points (insertion order is kept)
    if (g->pins_n > 0) {
        const int rc = convert_point_records(g);
        if (rc) return rc;
    }
    if (g->ins_n > 0) return expand_chunks(g, n, chunk_ids, masks, values, mem);
    int rc = grow_chunk_buffer(g, g->cins_n + n);
    if (rc) return rc;
    const int64_t at = g->cins_n;
    SG_HIP(hipMemcpyAsync(g->d_cins_id + at, chunk_ids, n * sizeof(int64_t), kind_from(mem), g->stream));
    SG_HIP(hipMemcpyAsync(g->d_cins_mask + at * csz, masks, (size_t)n * csz, kind_from(mem), g->stream));
    for (int p = 0; p < g->np; ++p) {
        unsigned char *dst = (unsigned char *)g->d_cins_val[p] + (size_t)at * csz * g->ts[p];
        if (values && values[p]) {
            SG_HIP(hipMemcpyAsync(dst, values[p], (size_t)n * csz * g->ts[p], kind_from(mem), g->stream));
        } else {
            dispatch_type(g->dt[p], [&](auto tag) {
                typedef decltype(tag) T;
                T v;
                memcpy(&v, g->bg[p], sizeof(T));
                hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_1d(n * csz, kWG)), dim3(kWG), 0, g->stream, (T *)dst, v,
                                   (int64_t)n * csz);
            });
            SG_HIP(hipGetLastError());
        }
    }
    g->cins_n += n;
    return SG_OK;
}

// Chunk records written by the caller's kernels: a zeroed staging area of nrec records
// (plus one sink record for claims beyond nrec) and a claim counter; insert_device_end
// reads the counter back and appends the claimed records as chunk records.
namespace {
inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

int insert_device_begin(sg_grid *g, int64_t nrec, sg_insert_view *out)
{
    if (nrec <= 0) return fail(SG_E_ARG, "record count must be positive");
    if (g->dins_open) return fail(SG_E_ARG, "a device insert buffer is open: end it (or flush) first");
    const size_t csz = (size_t)g->csz, nr = (size_t)nrec + 1;
    size_t total = 256 + round256(nr * 8) + round256(nr * csz);
    for (int p = 0; p < g->np; ++p) total += round256(nr * csz * g->ts[p]);
    int rc = ensure(g->pool, g->dins, total);
    if (rc) return rc;
    SG_HIP(hipMemsetAsync(g->dins.p, 0, total, g->stream));
    unsigned char *b = (unsigned char *)g->dins.p;
    memset(out, 0, sizeof(*out));
    out->count = (unsigned int *)b;
    b += 256;
    out->ids = (int64_t *)b;
    b += round256(nr * 8);
    out->mask = (uint8_t *)b;
    b += round256(nr * csz);
    for (int p = 0; p < g->np; ++p) {
        out->prop[p] = b;
        b += round256(nr * csz * g->ts[p]);
    }
    out->cap = nrec;
    out->chunk_size = g->csz;
    g->dins_cap = nrec;
    g->dins_open = true;
    return SG_OK;
}

int insert_device_end(sg_grid *g)
{
    if (!g->dins_open) return SG_OK;
    g->dins_open = false;
    unsigned int cnt = 0;
    SG_HIP(hipMemcpyAsync(&cnt, g->dins.p, sizeof(cnt), hipMemcpyDeviceToHost, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    if ((int64_t)cnt > g->dins_cap)
        return fail(SG_E_RANGE, "device insert buffer overflow: " + std::to_string(cnt) + " chunk records claimed, " +
                                    std::to_string(g->dins_cap) + " reserved");
    if (cnt == 0) return SG_OK;
    const size_t csz = (size_t)g->csz, nr = (size_t)g->dins_cap + 1;
    unsigned char *b = (unsigned char *)g->dins.p + 256;
    const int64_t *ids = (const int64_t *)b;
    b += round256(nr * 8);
    const uint8_t *mask = b;
    b += round256(nr * csz);
    const void *vals[kMaxProps] = {};
    for (int p = 0; p < g->np; ++p) {
        vals[p] = b;
        b += round256(nr * csz * g->ts[p]);
    }
    return append_chunks(g, cnt, ids, mask, vals, SG_MEM_DEVICE);
}

namespace {

// Deferred merge: the caller writes every touched chunk whole (the fused fold) from the
// arrays in src_* (the old arrays when the set grew); those are freed by finish_merge.
struct MergeDefer {
    const int64_t *oldpos = nullptr;  // [nnew] old position of each touched chunk
    const uint32_t *isnew = nullptr;  // [nnew] 1: brand-new chunk
    int64_t nfresh = 0;               // brand-new chunks among the nnew
    // in: keep the EXIST bit-planes (8^3 chunks) through the merge, the caller's fold
    // writing those of the touched chunks; out: false when they could not be kept
    bool planes = false;
    const uint8_t *src_mask = nullptr;
    const void *src_prop[kMaxProps] = {};
    void *free_keys = nullptr, *free_mask = nullptr, *free_prop[kMaxProps] = {};
};

void finish_merge(sg_grid *g, MergeDefer &df)
{
    // stream order: later allocations reuse these only after the fold has run
    if (df.free_keys) g->pool.free(df.free_keys);
    if (df.free_mask) g->pool.free(df.free_mask);
    for (int p = 0; p < g->np; ++p)
        if (df.free_prop[p]) g->pool.free(df.free_prop[p]);
    df = MergeDefer();
}


// brand-new chunks are placed by merge-path rank, the arrays grow (old chunks moved, new
// ones mask 0 / background), and cpos[j] is the final position of nkey[j].
int merge_chunk_set(sg_grid *g, const int64_t *nkey, int64_t nnew, int64_t **cpos_out, MergeDefer *df)
{
    const int csz = g->csz;
    const int64_t m = g->nchunks;
    int rc;
    DevBuf &b_c = g->scratch[5];
    if ((rc = ensure(g->pool, b_c, (size_t)nnew * (8 + 4 + 4 + 8 + 8) + 128))) return rc;
    int64_t *oldpos = (int64_t *)b_c.p;
    uint32_t *isnew = (uint32_t *)(oldpos + nnew);
    uint32_t *nscan = isnew + nnew;
    int64_t *cpos = (int64_t *)(((uintptr_t)(nscan + nnew) + 15) & ~(uintptr_t)15);
    int64_t *bn = cpos + nnew;
    hipLaunchKernelGGL(old_lookup_kernel, dim3(grid_1d(nnew, kWG)), dim3(kWG), 0, g->stream, nkey, nnew, g->d_keys,
                       m, oldpos, isnew);
    SG_HIP(hipGetLastError());
    uint32_t nb32 = 0;
    if ((rc = exclusive_scan_u32(g, isnew, nscan, nnew, &nb32))) return rc;
    const int64_t nb = nb32;
    hipLaunchKernelGGL(brand_new_kernel, dim3(grid_1d(nnew, kWG)), dim3(kWG), 0, g->stream, nkey, isnew, nscan, nnew,
                       bn);
    SG_HIP(hipGetLastError());
    hipLaunchKernelGGL(chunk_pos_kernel, dim3(grid_1d(nnew, kWG)), dim3(kWG), 0, g->stream, nkey, oldpos, nscan, nnew,
                       g->d_keys, m, bn, nb, cpos);
    SG_HIP(hipGetLastError());
    if (df) {
        df->oldpos = oldpos;
        df->isnew = isnew;
        df->nfresh = nb;
        df->src_mask = g->d_mask;
        for (int p = 0; p < g->np; ++p) df->src_prop[p] = g->d_prop[p];
    }

    if (nb > 0) {
        // the chunk set grows: build the union arrays (old chunks moved, new chunks
        // initialised to mask 0 / background), then fold into them
        const int64_t total = m + nb;
        int64_t ncap = std::max<int64_t>(total, g->cap);
        if (total > g->cap) ncap = std::max<int64_t>(total, (int64_t)(g->cap * 1.25) + 16);
        int64_t *nkeys = (int64_t *)g->pool.alloc(ncap * sizeof(int64_t));
        uint8_t *nmask = (uint8_t *)g->pool.alloc((size_t)(ncap + 1) * csz);
        void *nprop[kMaxProps] = {};
        if (!nkeys || !nmask) return fail(SG_E_OOM, "chunk array allocation failed");
        for (int p = 0; p < g->np; ++p) {
            nprop[p] = g->pool.alloc((size_t)(ncap + 1) * csz * g->ts[p]);
            if (!nprop[p]) return fail(SG_E_OOM, "chunk array allocation failed");
        }
        DevBuf &b_o = g->scratch[6];
        if ((rc = ensure(g->pool, b_o, (size_t)(m + nb) * (8 * 2 + 1) + 128))) return rc;
        int64_t *opos = (int64_t *)b_o.p;
        int64_t *bpos = opos + m;
        int64_t *src_of = bpos + nb;
        if (m > 0) {
            hipLaunchKernelGGL(move_old_keys_kernel, dim3(grid_1d(m, kWG)), dim3(kWG), 0, g->stream, g->d_keys, m, bn,
                               nb, nkeys, opos, src_of);
            SG_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(place_new_keys_kernel, dim3(grid_1d(nb, kWG)), dim3(kWG), 0, g->stream, bn, nb, g->d_keys,
                           m, nkeys, bpos, src_of);
        SG_HIP(hipGetLastError());
        // one streaming pass when every array's chunk is a power-of-two count of 16-byte vectors
        UnionArrays U;
        U.na = 1 + g->np;
        bool flat = true;
        uint64_t *neb = nullptr;
        for (int a = 0; a < U.na && flat; ++a) {
            const size_t cb = (size_t)csz * (a == 0 ? 1 : g->ts[a - 1]);
            flat = cb % 16 == 0 && ((cb / 16) & (cb / 16 - 1)) == 0;
            if (!flat) break;
            int sh = 0;
            while (((size_t)1 << sh) < cb / 16) ++sh;
            U.vshift[a] = sh;
            unsigned char pat[8];
            const int ts = (a == 0) ? 1 : g->ts[a - 1];
            for (int b = 0; b < 8; b += ts) {
                if (a == 0) pat[b] = 0;
                else memcpy(pat + b, g->bg[a - 1], ts);
            }
            memcpy(&U.pat[a], pat, 8);
            U.src[a] = (const uint4 *)(a == 0 ? (const void *)g->d_mask : g->d_prop[a - 1]);
            U.dst[a] = (uint4 *)(a == 0 ? (void *)nmask : nprop[a - 1]);
        }
        if (df && df->planes) {
            // the bit-planes (64 bytes per chunk, 0 for a new one) move with the chunks
            neb = flat ? (uint64_t *)g->pool.alloc((size_t)(ncap + 1) * 64) : nullptr;
            if (neb) {
                const int a = U.na++;
                U.vshift[a] = 2;
                U.pat[a] = 0;
                U.src[a] = (const uint4 *)g->d_ebits;
                U.dst[a] = (uint4 *)neb;
            } else {
                df->planes = false;
            }
        }
        if (flat && df && m == 0) {
            // no old chunk: the caller's fold writes every chunk, only the background
            // chunk (position total) is initialised here
            UnionArrays B = U;
            for (int a = 0; a < B.na; ++a) B.dst[a] += (size_t)total << B.vshift[a];
            hipLaunchKernelGGL(build_union_kernel, dim3(1), dim3(kWG), 0, g->stream, B, src_of, (int64_t)1,
                               (const uint8_t *)nullptr);
            SG_HIP(hipGetLastError());
        } else if (flat) {
            uint8_t *skip = nullptr;
            if (df) {
                // the touched chunks are written whole by the caller's fused fold
                skip = (uint8_t *)(src_of + (m + nb));
                hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_1d(total, kWG)), dim3(kWG), 0, g->stream, skip, (uint8_t)0,
                                   total);
                hipLaunchKernelGGL(mark_positions_kernel, dim3(grid_1d(nnew, kWG)), dim3(kWG), 0, g->stream, cpos,
                                   nnew, skip);
                SG_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(build_union_kernel, dim3(2048), dim3(kWG), 0, g->stream, U, src_of, total + 1, skip);
            SG_HIP(hipGetLastError());
        }
        // chunk data: copy as 16-byte vectors when the chunk size allows
        auto move = [&](const void *src, void *dst, size_t chunk_bytes) -> int {
            if (m == 0) return SG_OK;
            if (chunk_bytes % 16 == 0)
                hipLaunchKernelGGL(move_chunks_kernel<uint4>, dim3((unsigned)m), dim3(kWG), 0, g->stream,
                                   (const uint4 *)src, (uint4 *)dst, opos, m, (int)(chunk_bytes / 16));
            else
                hipLaunchKernelGGL(move_chunks_kernel<uint8_t>, dim3((unsigned)m), dim3(kWG), 0, g->stream,
                                   (const uint8_t *)src, (uint8_t *)dst, opos, m, (int)chunk_bytes);
            SG_HIP(hipGetLastError());
            return SG_OK;
        };
        if (!flat && (rc = move(g->d_mask, nmask, csz))) return rc;
        if (!flat) {
            hipLaunchKernelGGL(init_chunks_kernel<uint8_t>, dim3((unsigned)nb), dim3(kWG), 0, g->stream, nmask, bpos,
                               nb, csz, (uint8_t)0);
            SG_HIP(hipGetLastError());
            hipLaunchKernelGGL(init_chunks_kernel<uint8_t>, dim3(1), dim3(kWG), 0, g->stream, nmask + total * csz,
                               (const int64_t *)nullptr, (int64_t)1, csz, (uint8_t)0);
            SG_HIP(hipGetLastError());
        }
        for (int p = 0; p < g->np && !flat; ++p) {
            if ((rc = move(g->d_prop[p], nprop[p], (size_t)csz * g->ts[p]))) return rc;
            dispatch_type(g->dt[p], [&](auto tag) {
                typedef decltype(tag) T;
                T v;
                memcpy(&v, g->bg[p], sizeof(T));
                hipLaunchKernelGGL(init_chunks_kernel<T>, dim3((unsigned)nb), dim3(kWG), 0, g->stream, (T *)nprop[p],
                                   bpos, nb, csz, v);
                hipLaunchKernelGGL(init_chunks_kernel<T>, dim3(1), dim3(kWG), 0, g->stream,
                                   (T *)nprop[p] + total * csz, (const int64_t *)nullptr, (int64_t)1, csz, v);
            });
            SG_HIP(hipGetLastError());
        }
        if (df) {
            // the fused fold reads the touched old chunks from the old arrays: freed after it
            df->free_keys = g->d_keys;
            df->free_mask = g->d_mask;
            for (int p = 0; p < g->np; ++p) df->free_prop[p] = g->d_prop[p];
        } else {
            g->pool.free(g->d_keys);
            g->pool.free(g->d_mask);
            for (int p = 0; p < g->np; ++p) g->pool.free(g->d_prop[p]);
        }
        g->d_keys = nkeys;
        g->d_mask = nmask;
        for (int p = 0; p < g->np; ++p) g->d_prop[p] = nprop[p];
        g->nchunks = total;
        g->cap = ncap;
        if (neb) {
            // the old planes are read only by the union pass above (stream order)
            g->pool.free(g->d_ebits);
            g->d_ebits = neb;
            g->ebits_cap = ncap + 1;
        }
    }
    if (df && df->planes && (!g->d_ebits || g->ebits_cap < g->nchunks + 1)) df->planes = false;

    *cpos_out = cpos;
    return SG_OK;
}

// flush of whole-chunk records only: stable sort of (chunk id, record id), unique chunk
// ids + segment starts, merge into the chunk map, one workgroup per touched chunk folds
// its records slot by slot
int flush_chunks_impl(sg_grid *g, const int32_t *ops)
{
    const int64_t n = g->cins_n;
    if (n >= ((int64_t)1 << 31)) return fail(SG_E_RANGE, "more than 2^31 pending chunk records");
    const int csz = g->csz;
    const int64_t nblocks = key_limit(g) / csz;
    int rc;
    {
        if ((rc = ensure(g->pool, g->scratch[6], 64))) return rc;
        int *bad = (int *)g->scratch[6].p;
        SG_HIP(hipMemsetAsync(bad, 0, sizeof(int), g->stream));
        hipLaunchKernelGGL(check_keys_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_cins_id, n,
                           nblocks, bad);
        SG_HIP(hipGetLastError());
        int hbad = 0;
        SG_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
        SG_HIP(hipStreamSynchronize(g->stream));
        if (hbad) {
            g->cins_n = 0;
            return fail(SG_E_RANGE, "chunk id outside the grid in the insert buffer");
        }
    }
    int bits = 0;
    while (bits < 63 && (((int64_t)1 << bits) < nblocks)) ++bits;
    const bool k32 = bits <= 32;
    DevBuf &b_keys = g->scratch[0], &b_vals = g->scratch[1], &b_tk = g->scratch[2], &b_tv = g->scratch[3];
    DevBuf &b_u = g->scratch[4];
    const size_t kbytes = (k32 ? 4 : 8) * (size_t)n;
    if ((rc = ensure(g->pool, b_keys, kbytes)) || (rc = ensure(g->pool, b_tk, kbytes)) ||
        (rc = ensure(g->pool, b_vals, 4 * (size_t)n)) || (rc = ensure(g->pool, b_tv, 4 * (size_t)n)))
        return rc;
    void *sk = nullptr;
    uint32_t *perm = nullptr;
    if (k32) {
        hipLaunchKernelGGL(keys_to_sort_kernel<uint32_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           g->d_cins_id, (uint32_t *)b_keys.p, (uint32_t *)b_vals.p, n, (int64_t)0, (int *)nullptr);
        SG_HIP(hipGetLastError());
        uint32_t *ko;
        rc = radix_sort_pairs_u32(g, (uint32_t *)b_keys.p, (uint32_t *)b_vals.p, (uint32_t *)b_tk.p,
                                  (uint32_t *)b_tv.p, n, bits, &ko, &perm);
        sk = ko;
    } else {
        hipLaunchKernelGGL(keys_to_sort_kernel<uint64_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           g->d_cins_id, (uint64_t *)b_keys.p, (uint32_t *)b_vals.p, n, (int64_t)0, (int *)nullptr);
        SG_HIP(hipGetLastError());
        uint64_t *ko;
        rc = radix_sort_pairs_u64(g, (uint64_t *)b_keys.p, (uint32_t *)b_vals.p, (uint64_t *)b_tk.p,
                                  (uint32_t *)b_tv.p, n, bits, &ko, &perm);
        sk = ko;
    }
    if (rc) return rc;
    uint32_t *spare = (perm == (uint32_t *)b_vals.p) ? (uint32_t *)b_tv.p : (uint32_t *)b_vals.p;
    // b_u: hscan [n] u32 | ukey [n] i64 | rstart [n+1] u32
    if ((rc = ensure(g->pool, b_u, (size_t)n * (4 + 8 + 4) + 128))) return rc;
    uint32_t *head = spare;
    uint32_t *hscan = (uint32_t *)b_u.p;
    int64_t *ukey = (int64_t *)((char *)b_u.p + (((size_t)n * 4 + 15) & ~(size_t)15));
    uint32_t *rstart = (uint32_t *)(ukey + n);
    if (k32)
        hipLaunchKernelGGL(run_head_kernel<uint32_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint32_t *)sk, head, n);
    else
        hipLaunchKernelGGL(run_head_kernel<uint64_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint64_t *)sk, head, n);
    SG_HIP(hipGetLastError());
    uint32_t nu32 = 0;
    if ((rc = exclusive_scan_u32(g, head, hscan, n, &nu32))) return rc;
    const int64_t nu = nu32;
    if (k32)
        hipLaunchKernelGGL(unique_points_kernel<uint32_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint32_t *)sk, head, hscan, n, ukey, rstart, nu32);
    else
        hipLaunchKernelGGL(unique_points_kernel<uint64_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint64_t *)sk, head, hscan, n, ukey, rstart, nu32);
    SG_HIP(hipGetLastError());
    int64_t *cpos = nullptr;
    if ((rc = merge_chunk_set(g, ukey, nu, &cpos, nullptr))) return rc;
    for (int p = 0; p < g->np; ++p) {
        const int op = ops ? ops[p] : SG_OP_NONE;
        dispatch_type(g->dt[p], [&](auto tag) {
            typedef decltype(tag) T;
            T v;
            memcpy(&v, g->bg[p], sizeof(T));
            hipLaunchKernelGGL(fold_chunks_kernel<T>, dim3((unsigned)nu), dim3(kWG), 0, g->stream, rstart, perm, cpos,
                               csz, g->d_cins_mask, (const T *)g->d_cins_val[p], (T *)g->d_prop[p],
                               g->d_mask, op, v);
        });
        SG_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(fold_chunk_masks_kernel, dim3((unsigned)nu), dim3(kWG), 0, g->stream, rstart, perm, cpos, csz,
                       g->d_cins_mask, g->d_mask);
    SG_HIP(hipGetLastError());
    g->cins_n = 0;
    g->nn_valid[0] = g->nn_valid[1] = false;
    g->occ_tab_valid = false;
    g->ebits_valid = false;
    g->ep_valid = false;
    return SG_OK;
}

// Flush of point records grouped by chunk: the records become (chunk id, record of 16 or
// 32 bytes) pairs, stably sorted by chunk id with the records moving along (no gather
// through a permutation), and one wave per touched chunk folds the chunk's run of
// records in LDS and writes the chunk once.  The fold each point receives is the
// point-key flush's: its records in insertion order, then the old value.  Taken when the
// chunk ids fit 32 bits, a record fits 32 bytes and a chunk's fold state fits 64 KB of
// LDS (*taken = false otherwise, nothing done).
int flush_grouped_impl(sg_grid *g, const int32_t *ops, bool *taken)
{
    *taken = false;
    const int csz = g->csz;
    const int64_t limit = key_limit(g);
    RecLayout R;
    FoldState S;
    int cbits = 0;
    if (!point_record_layout(g, &R, &S, &cbits)) return SG_OK;
    if (g->ins_n + g->pins_n >= ((int64_t)1 << 31)) return SG_OK;
    *taken = true;
    int rc;
    uint32_t *hw = host_words(g);
    if (!hw) return fail(SG_E_OOM, "pinned host allocation failed");
    // records staged as (chunk id, record) pairs are sorted as they are (the sort only
    // reads them); when a staged coordinate was outside the grid (the flag comes back with
    // the first scan total below) they go through the insert buffer instead
    bool staged = g->pins_n > 0;
    if (staged && R.nv != g->pins_nv) {
        if ((rc = convert_point_records(g))) return rc;
        staged = false;
    }
    const int64_t n = staged ? g->pins_n : g->ins_n;
    if ((rc = ensure(g->pool, g->scratch[6], 64))) return rc;
    int *bad = (int *)g->scratch[6].p;
    SG_HIP(hipMemsetAsync(bad, 0, sizeof(int), g->stream));

    FoldProps F;
    F.np = g->np;
    for (int p = 0; p < g->np; ++p) {
        F.dt[p] = g->dt[p];
        F.ts[p] = g->ts[p];
        F.op[p] = ops ? ops[p] : SG_OP_NONE;
        F.rval[p] = (const unsigned char *)g->d_ins_val[p];
        F.bg[p] = 0;
        memcpy(&F.bg[p], g->bg[p], g->ts[p]);
    }
    // (chunk id, record) pairs and the sort's second buffers; staged pairs are the sort's
    // input, kept (a failure below leaves them pending), with b_ck / b_rec as its second
    // temporaries
    DevBuf &b_ck = g->scratch[0], &b_ckt = g->scratch[2], &b_rec = g->scratch[7], &b_rect = g->scratch[13];
    if ((rc = ensure(g->pool, b_ck, (size_t)n * 4 + 64)) || (rc = ensure(g->pool, b_ckt, (size_t)n * 4 + 64)) ||
        (rc = ensure(g->pool, b_rec, (size_t)n * R.bytes + 64)) ||
        (rc = ensure(g->pool, b_rect, (size_t)n * R.bytes + 64)))
        return rc;
    uint32_t *ck = nullptr;
    uint4 *rec = nullptr;
    if (staged) {
        rc = radix_sort_records_u32(g, g->d_pins_ck, (uint4 *)g->d_pins_rec, (uint32_t *)b_ckt.p, (uint4 *)b_rect.p,
                                    R.nv, n, cbits, &ck, &rec, (uint32_t *)b_ck.p, (uint4 *)b_rec.p);
    } else {
        if (R.nv == 1)
            hipLaunchKernelGGL(chunk_records_kernel<1>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_ins_key,
                               g->d_ins_mask, n, csz, limit, F, R, (uint32_t *)b_ck.p, (uint4 *)b_rec.p, bad);
        else
            hipLaunchKernelGGL(chunk_records_kernel<2>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_ins_key,
                               g->d_ins_mask, n, csz, limit, F, R, (uint32_t *)b_ck.p, (uint4 *)b_rec.p, bad);
        SG_HIP(hipGetLastError());
        rc = radix_sort_records_u32(g, (uint32_t *)b_ck.p, (uint4 *)b_rec.p, (uint32_t *)b_ckt.p, (uint4 *)b_rect.p,
                                    R.nv, n, cbits, &ck, &rec);
    }
    if (rc) return rc;
    // the range flags, read with the scan total below
    SG_HIP(hipMemcpyAsync(hw + 8, bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    if (staged) SG_HIP(hipMemcpyAsync(hw + 9, g->d_pins_bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    // touched chunks: runs of equal sorted chunk ids, counted per tile, tile offsets
    // scanned, then each run's id and first record written at its offset
    const int64_t rtiles = (n + kRunTile - 1) / kRunTile;
    DevBuf &b_h = g->scratch[4];
    if ((rc = ensure(g->pool, b_h, (size_t)rtiles * 8 + 64))) return rc;
    uint32_t *rcnt = (uint32_t *)b_h.p, *roff = rcnt + rtiles;
    hipLaunchKernelGGL(run_count_kernel, dim3((unsigned)rtiles), dim3(kWG), 0, g->stream, (const uint32_t *)ck, n,
                       rcnt);
    SG_HIP(hipGetLastError());
    uint32_t nn32 = 0;
    if ((rc = exclusive_scan_u32(g, rcnt, roff, rtiles, &nn32))) return rc;
    if (staged && hw[9]) {
        // the staged pairs, intact, through the insert buffer: the flush by chunk id over it
        // drops the records outside the grid and reports them
        if ((rc = convert_point_records(g))) return rc;
        return flush_grouped_impl(g, ops, taken);
    }
    if (hw[8]) {
        // the records outside the grid are dropped, the others stay pending (nothing is merged)
        int64_t kept = 0;
        if ((rc = drop_invalid_records(g, limit, &kept))) return rc;
        return fail(SG_E_RANGE, std::to_string(n - kept) + " inserted points outside the grid were dropped; the other " +
                                    std::to_string(kept) + " stay pending (flush again to merge them)");
    }
    const int64_t nnew = nn32;
    // unique chunk ids and the first record of each (ustart[nnew] = n)
    if ((rc = ensure(g->pool, g->scratch[12], (size_t)nnew * 8 + (size_t)(nnew + 1) * 4 + 64))) return rc;
    int64_t *nkey = (int64_t *)g->scratch[12].p;
    uint32_t *ustart = (uint32_t *)(nkey + nnew);
    hipLaunchKernelGGL(run_emit_kernel, dim3((unsigned)rtiles), dim3(kWG), 0, g->stream, (const uint32_t *)ck, n, roff,
                       nkey, ustart, nn32);
    SG_HIP(hipGetLastError());

    MergeDefer df;
    // 8^3 chunks: the EXIST bit-planes are kept up to date when they are valid (or the
    // grid has no chunk yet): the fold writes those of the touched chunks
    df.planes = g->dim == 3 && g->edge == 8 && (g->ebits_valid || g->nchunks == 0);
    int64_t *cpos = nullptr;
    if ((rc = merge_chunk_set(g, nkey, nnew, &cpos, &df))) return rc;
    const bool planes = df.planes;
    uint64_t *eb = planes ? g->d_ebits : nullptr;
    FusedImg I;
    for (int p = 0; p < g->np; ++p) {
        I.src[p] = (const unsigned char *)df.src_prop[p];
        F.out[p] = (unsigned char *)g->d_prop[p];
    }
    I.smask = df.src_mask;
    int obits = 0;
    while ((1 << obits) < csz) ++obits;
    // brand-new chunks: folded in an image of the chunk (csz % 16 == 0, image <= 64 KB)
    FreshArgs A;
    bool fresh_ok = (csz % 16) == 0 && df.nfresh > 0;
    {
        size_t at = 0;
        A.np = g->np;
        for (int p = 0; p < g->np; ++p) {
            FreshProp &P = A.p[p];
            P.dt = g->dt[p];
            P.op = F.op[p];
            P.ts = g->ts[p];
            P.roff = R.off[p];
            P.voff = (int)at;
            P.pad = 0;
            P.out = F.out[p];
            unsigned char pat[8];
            for (int b = 0; b < 8; b += g->ts[p]) memcpy(pat + b, g->bg[p], g->ts[p]);
            memcpy(&P.pat, pat, 8);
            at += (size_t)csz * g->ts[p];
        }
        A.moff = (int)at;
        at += csz;
        A.soff = (int)at;
        at += (size_t)(csz + 31) / 32 * 4;
        at = (at + 15) / 16 * 16;
        A.bytes = (int)std::min<size_t>(at, 1 << 30);
        fresh_ok = fresh_ok && at <= 65536;
    }
    if (fresh_ok) {
        const int fw = std::max(1, std::min(4, 65536 / A.bytes));
        const dim3 fgrid((unsigned)((nnew + fw - 1) / fw)), fblk(64 * fw);
        // every touched chunk new: no per-chunk check
        const uint32_t *chk = df.nfresh < nnew ? df.isnew : nullptr;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, fgrid, fblk, (size_t)fw * A.bytes, g->stream, rec, ustart, cpos, chk, nnew, csz,
                               obits, fw, A, g->d_mask, eb);
        };
        if (R.nv == 1) {
            switch (g->np) {
            case 1: launch(fold_fresh_chunks_kernel<1, 1>); break;
            case 2: launch(fold_fresh_chunks_kernel<1, 2>); break;
            case 3: launch(fold_fresh_chunks_kernel<1, 3>); break;
            case 4: launch(fold_fresh_chunks_kernel<1, 4>); break;
            default: launch(fold_fresh_chunks_kernel<1, 0>); break;
            }
        } else {
            switch (g->np) {
            case 1: launch(fold_fresh_chunks_kernel<2, 1>); break;
            case 2: launch(fold_fresh_chunks_kernel<2, 2>); break;
            case 3: launch(fold_fresh_chunks_kernel<2, 3>); break;
            case 4: launch(fold_fresh_chunks_kernel<2, 4>); break;
            default: launch(fold_fresh_chunks_kernel<2, 0>); break;
            }
        }
        SG_HIP(hipGetLastError());
    }
    if (!fresh_ok || df.nfresh < nnew) {
        // chunks that existed before (and new ones when the image path is not taken)
        const int waves = std::max(1, std::min(4, 65536 / S.bytes));
        // one wave per touched chunk (a wave's loop then runs once)
        const int64_t groups = (nnew + waves - 1) / waves;
        const dim3 fgrid((unsigned)groups), fblk(64 * waves);
        const int skip_fresh = fresh_ok ? 1 : 0;
        if (R.nv == 1)
            hipLaunchKernelGGL(fold_chunk_records_kernel<1>, fgrid, fblk, (size_t)waves * S.bytes, g->stream, rec,
                               ustart, cpos, df.oldpos, df.isnew, nnew, csz, obits, waves, S, F, R, I, g->d_mask,
                               skip_fresh, eb);
        else
            hipLaunchKernelGGL(fold_chunk_records_kernel<2>, fgrid, fblk, (size_t)waves * S.bytes, g->stream, rec,
                               ustart, cpos, df.oldpos, df.isnew, nnew, csz, obits, waves, S, F, R, I, g->d_mask,
                               skip_fresh, eb);
        SG_HIP(hipGetLastError());
    }
    finish_merge(g, df);
    g->ins_n = 0;
    g->pins_n = 0;
    g->nn_valid[0] = g->nn_valid[1] = false;
    g->occ_tab_valid = false;
    g->ebits_valid = planes;
    g->ep_valid = false;
    return SG_OK;
}

}  // namespace

int flush_impl(sg_grid *g, const int32_t *ops)
{
    // chunk records are pending only while no point record is (appending points converts them)
    if (g->cins_n > 0) return flush_chunks_impl(g, ops);
    if (g->ins_n + g->pins_n == 0) return SG_OK;  // nothing inserted (map_vector_sparse.hpp:1199-1213)
    if (g->ins_n + g->pins_n >= ((int64_t)1 << 31)) return fail(SG_E_RANGE, "more than 2^31 pending inserts");
    if (g->flush_path == 0) {
        bool taken = false;
        const int rc0 = flush_grouped_impl(g, ops, &taken);
        if (taken || rc0) return rc0;
    }
    // the flush by point key reads the insert buffer
    int rc = convert_point_records(g);
    if (rc) return rc;
    const int64_t n = g->ins_n;
    const int csz = g->csz;
    const int64_t limit = key_limit(g);

    // validate device-side keys: the flag comes back with the first scan total below (the
    // grid is not touched before it is read; out-of-range keys only reorder scratch data)
    uint32_t *hw = host_words(g);
    if (!hw) return fail(SG_E_OOM, "pinned host allocation failed");
    rc = ensure(g->pool, g->scratch[6], 64);
    if (rc) return rc;
    int *bad = (int *)g->scratch[6].p;
    SG_HIP(hipMemsetAsync(bad, 0, sizeof(int), g->stream));
    int bits = 0;
    while (bits < 63 && (((int64_t)1 << bits) < limit)) ++bits;
    const bool k32 = bits <= 32;

    // scratch layout
    DevBuf &b_keys = g->scratch[0], &b_vals = g->scratch[1], &b_tk = g->scratch[2], &b_tv = g->scratch[3];
    DevBuf &b_u = g->scratch[4];
    const size_t kbytes = (k32 ? 4 : 8) * (size_t)n;
    if ((rc = ensure(g->pool, b_keys, kbytes)) || (rc = ensure(g->pool, b_tk, kbytes)) ||
        (rc = ensure(g->pool, b_vals, 4 * (size_t)n)) || (rc = ensure(g->pool, b_tv, 4 * (size_t)n)))
        return rc;

    void *sk = nullptr;
    uint32_t *perm = nullptr;
    if (k32) {
        hipLaunchKernelGGL(keys_to_sort_kernel<uint32_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           g->d_ins_key, (uint32_t *)b_keys.p, (uint32_t *)b_vals.p, n, limit, bad);
        SG_HIP(hipGetLastError());
        uint32_t *ko;
        rc = radix_sort_pairs_u32(g, (uint32_t *)b_keys.p, (uint32_t *)b_vals.p, (uint32_t *)b_tk.p,
                                  (uint32_t *)b_tv.p, n, bits, &ko, &perm);
        sk = ko;
    } else {
        hipLaunchKernelGGL(keys_to_sort_kernel<uint64_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           g->d_ins_key, (uint64_t *)b_keys.p, (uint32_t *)b_vals.p, n, limit, bad);
        SG_HIP(hipGetLastError());
        uint64_t *ko;
        rc = radix_sort_pairs_u64(g, (uint64_t *)b_keys.p, (uint32_t *)b_vals.p, (uint64_t *)b_tk.p,
                                  (uint32_t *)b_tv.p, n, bits, &ko, &perm);
        sk = ko;
    }
    if (rc) return rc;
    // the range flag of the sort-key pass, read with the first scan total below
    SG_HIP(hipMemcpyAsync(hw + 8, bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    // the temporaries not holding the sorted result are free for reuse
    uint32_t *spare_a = (perm == (uint32_t *)b_vals.p) ? (uint32_t *)b_tv.p : (uint32_t *)b_vals.p;

    // unique points: run heads + scan (head -> spare_a, scan -> b_u)
    if ((rc = ensure(g->pool, b_u, (size_t)n * (4 + 4 + 8 + 4 + 4) + 64))) return rc;
    uint32_t *head = spare_a;
    uint32_t *hscan = (uint32_t *)b_u.p;
    if (k32)
        hipLaunchKernelGGL(run_head_kernel<uint32_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint32_t *)sk, head, n);
    else
        hipLaunchKernelGGL(run_head_kernel<uint64_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint64_t *)sk, head, n);
    SG_HIP(hipGetLastError());
    uint32_t nu32 = 0;
    if ((rc = exclusive_scan_u32(g, head, hscan, n, &nu32))) return rc;
    if (hw[8]) {  // the key check above, complete now
        // the records outside the grid are dropped, the others stay pending (nothing is merged)
        int64_t kept = 0;
        if ((rc = drop_invalid_records(g, limit, &kept))) return rc;
        return fail(SG_E_RANGE, std::to_string(n - kept) + " inserted points outside the grid were dropped; the other " +
                                    std::to_string(kept) + " stay pending (flush again to merge them)");
    }
    const int64_t nu = nu32;
    int64_t *ukey = (int64_t *)((char *)b_u.p + (((size_t)n * 4 + 15) & ~(size_t)15));
    uint32_t *rstart = (uint32_t *)(ukey + n);
    if (k32)
        hipLaunchKernelGGL(unique_points_kernel<uint32_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint32_t *)sk, head, hscan, n, ukey, rstart, nu32);
    else
        hipLaunchKernelGGL(unique_points_kernel<uint64_t>, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream,
                           (const uint64_t *)sk, head, hscan, n, ukey, rstart, nu32);
    SG_HIP(hipGetLastError());

    // unique chunks among the unique points (head -> hscan region reused)
    uint32_t *chead = head;  // head no longer needed
    uint32_t *cscan = hscan;
    hipLaunchKernelGGL(chunk_head_kernel, dim3(grid_1d(nu, kWG)), dim3(kWG), 0, g->stream, ukey, chead, nu, csz);
    SG_HIP(hipGetLastError());
    uint32_t nn32 = 0;
    if ((rc = exclusive_scan_u32(g, chead, cscan, nu, &nn32))) return rc;
    const int64_t nnew = nn32;

    // unique chunk ids, then their positions in the (grown) chunk map
    if ((rc = ensure(g->pool, g->scratch[12], (size_t)nnew * 8 + (size_t)(nnew + 1) * 4 + 64))) return rc;
    int64_t *nkey = (int64_t *)g->scratch[12].p;
    uint32_t *ustart = (uint32_t *)(nkey + nnew);
    hipLaunchKernelGGL(unique_chunks_kernel, dim3(grid_1d(nu, kWG)), dim3(kWG), 0, g->stream, ukey, chead, cscan, nu,
                       csz, nkey, ustart, nn32);
    SG_HIP(hipGetLastError());

    // fold every property and the mask of each unique point in place
    FoldProps F;
    F.np = g->np;
    for (int p = 0; p < g->np; ++p) {
        F.dt[p] = g->dt[p];
        F.ts[p] = g->ts[p];
        F.op[p] = ops ? ops[p] : SG_OP_NONE;
        F.rval[p] = (const unsigned char *)g->d_ins_val[p];
        F.bg[p] = 0;
        memcpy(&F.bg[p], g->bg[p], g->ts[p]);
    }
    // interleaved records (properties at their natural alignment, then the mask byte),
    // then put in key order
    {
        int at = 0;
        for (int p = 0; p < g->np; ++p) {
            at = (at + g->ts[p] - 1) / g->ts[p] * g->ts[p];
            F.off[p] = at;
            at += g->ts[p];
        }
        F.moff = at;
        F.stride = (at + 1 + 15) / 16 * 16;  // whole 16-byte vectors for the permutation
        if ((rc = ensure(g->pool, g->scratch[7], (size_t)n * F.stride + 64)) ||
            (rc = ensure(g->pool, g->scratch[13], (size_t)n * F.stride + 64)))
            return rc;
        hipLaunchKernelGGL(interleave_records_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, n,
                           g->d_ins_mask, F, (unsigned char *)g->scratch[7].p);
        SG_HIP(hipGetLastError());
        const int vecs = F.stride / 16;
        hipLaunchKernelGGL(permute_records_kernel, dim3(grid_1d(n * vecs, kWG)), dim3(kWG), 0, g->stream,
                           (const uint4 *)g->scratch[7].p, perm, n, vecs, (uint4 *)g->scratch[13].p);
        SG_HIP(hipGetLastError());
        F.rec = (const unsigned char *)g->scratch[13].p;
    }
    uint32_t *uchunk = rstart + (n + 1);
    hipLaunchKernelGGL(uchunk_index_kernel, dim3(grid_1d(nu, kWG)), dim3(kWG), 0, g->stream, cscan, chead, nu, uchunk);
    SG_HIP(hipGetLastError());
    // fold into a compact array, then one wave builds each touched chunk in LDS and writes
    // it once, when every array's chunk is whole 16-byte vectors and a chunk image fits
    // 64 KB of LDS
    FusedImg I;
    bool fused = csz % 16 == 0;
    {
        int at = 0;
        for (int p = 0; p < g->np; ++p) {
            fused = fused && ((size_t)csz * g->ts[p]) % 16 == 0;
            I.off[p] = at;
            at += csz * g->ts[p];
            unsigned char pat[8];
            for (int b = 0; b < 8; b += g->ts[p]) memcpy(pat + b, g->bg[p], g->ts[p]);
            memcpy(&I.pat[p], pat, 8);
        }
        I.moff = at;
        I.bytes = (at + csz + 15) / 16 * 16;
        fused = fused && I.bytes + 1024 <= 65536;
    }
    MergeDefer df;
    int64_t *cpos = nullptr;
    if ((rc = merge_chunk_set(g, nkey, nnew, &cpos, fused ? &df : nullptr))) return rc;
    for (int p = 0; p < g->np; ++p) F.out[p] = (unsigned char *)g->d_prop[p];
    if (fused) {
        for (int p = 0; p < g->np; ++p) I.src[p] = (const unsigned char *)df.src_prop[p];
        I.smask = df.src_mask;
        // the folded points, compact in key order (the permuted records are no longer needed)
        if ((rc = ensure(g->pool, g->scratch[7], (size_t)nu * F.stride + 64))) return rc;
        unsigned char *res = (unsigned char *)g->scratch[7].p;
        hipLaunchKernelGGL(fold_points_compact_kernel, dim3(grid_1d(nu, kWG)), dim3(kWG), 0, g->stream, ukey, rstart,
                           uchunk, df.oldpos, df.isnew, nu, csz, F, I, res);
        SG_HIP(hipGetLastError());
        int img_bytes = csz;  // the largest array's chunk
        for (int p = 0; p < g->np; ++p) img_bytes = std::max(img_bytes, csz * g->ts[p]);
        const int waves = std::max(1, std::min(4, 65536 / (img_bytes + 1024)));
        hipLaunchKernelGGL(write_chunks_kernel, dim3((unsigned)((nnew + waves - 1) / waves)), dim3(64 * waves),
                           (size_t)waves * (img_bytes + 1024), g->stream, ukey, ustart, cpos, df.oldpos, df.isnew,
                           nnew, csz, waves, img_bytes, F, I, (const unsigned char *)res, g->d_mask);
        SG_HIP(hipGetLastError());
        finish_merge(g, df);
    } else {
        hipLaunchKernelGGL(fold_points_kernel, dim3(grid_1d(nu, kWG)), dim3(kWG), 0, g->stream, ukey, rstart, uchunk,
                           cpos, nu, csz, F, g->d_mask);
        SG_HIP(hipGetLastError());
    }
    g->ins_n = 0;
    // flush marks the neighbour table stale (SparseGridGpu.hpp:778), and the point list
    g->nn_valid[0] = g->nn_valid[1] = false;
    g->occ_tab_valid = false;
    g->ebits_valid = false;
    g->ep_valid = false;
    return SG_OK;
}

// chunk c of a replaced chunk list: its values from the old chunk oldpos[c] (or the
// background when the chunk is new), the background wherever the new mask is 0; chunk n
// (the background chunk) holds the background
template <typename T>
__global__ __launch_bounds__(kWG) void remap_chunks_kernel(const T *__restrict__ src, T *__restrict__ dst,
                                                          const int64_t *__restrict__ oldpos,
                                                          const uint8_t *__restrict__ nmask, int64_t n, int csz, T bg)
{
    const int64_t c = blockIdx.x;
    if (c > n) return;
    const int64_t o = (c < n) ? oldpos[c] : -1;
    for (int i = threadIdx.x; i < csz; i += kWG) {
        const int64_t s = c * csz + i;
        dst[s] = (o >= 0 && nmask[s] != 0) ? src[o * csz + i] : bg;
    }
}

int set_chunks_impl(sg_grid *g, int64_t n, const int64_t *ids, const uint8_t *masks, int mem)
{
    if (n < 0 || (n > 0 && (!ids || !masks))) return fail(SG_E_ARG, "bad chunk list");
    if (g->dins_open) return fail(SG_E_STATE, "device insert buffer open: call sg_insert_device_end first");
    const int csz = g->csz;
    // the ids on the host, checked: strictly increasing chunk ids of the domain
    std::vector<int64_t> hid((size_t)n);
    if (n > 0) {
        if (mem == SG_MEM_DEVICE) {
            SG_HIP(hipMemcpyAsync(hid.data(), ids, (size_t)n * 8, hipMemcpyDeviceToHost, g->stream));
            SG_HIP(hipStreamSynchronize(g->stream));
        } else {
            memcpy(hid.data(), ids, (size_t)n * 8);
        }
    }
    const ChunkLin L = chunk_lin(g);
    for (int64_t c = 0; c < n; ++c) {
        if (c > 0 && hid[c] <= hid[c - 1]) return fail(SG_E_ARG, "chunk ids not strictly increasing");
        if (hid[c] < 0) return fail(SG_E_RANGE, "negative chunk id");
        int64_t bc[3] = {0, 0, 0};
        chunk_coords(L, hid[c], bc);
        for (int d = 0; d < g->dim; ++d)
            if (bc[d] < 0 || bc[d] >= g->bsz[d]) return fail(SG_E_RANGE, "chunk id outside the grid");
        if (chunk_id(L, bc) != hid[c]) return fail(SG_E_RANGE, "chunk id outside the grid");
    }
    const int64_t ncap = std::max<int64_t>(n, 1);
    int64_t *nkeys = (int64_t *)g->pool.alloc((size_t)ncap * 8);
    uint8_t *nmask = (uint8_t *)g->pool.alloc((size_t)(ncap + 1) * csz);
    void *nprop[kMaxProps] = {};
    bool ok = nkeys && nmask;
    for (int p = 0; p < g->np && ok; ++p) ok = (nprop[p] = g->pool.alloc((size_t)(ncap + 1) * csz * g->ts[p])) != nullptr;
    int rc = SG_OK;
    DevBuf &bp = g->scratch[5];
    if (ok && (rc = ensure(g->pool, bp, (size_t)ncap * 8 + (size_t)ncap * 4 + 64))) ok = false;
    if (!ok) {
        if (nkeys) g->pool.free(nkeys);
        if (nmask) g->pool.free(nmask);
        for (int p = 0; p < g->np; ++p)
            if (nprop[p]) g->pool.free(nprop[p]);
        return rc ? rc : fail(SG_E_OOM, "chunk array allocation failed");
    }
    int64_t *oldpos = (int64_t *)bp.p;
    uint32_t *isnew = (uint32_t *)(oldpos + ncap);
    const hipMemcpyKind kind = (mem == SG_MEM_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (n > 0) {
        SG_HIP(hipMemcpyAsync(nkeys, hid.data(), (size_t)n * 8, hipMemcpyHostToDevice, g->stream));
        SG_HIP(hipMemcpyAsync(nmask, masks, (size_t)n * csz, kind, g->stream));
        hipLaunchKernelGGL(old_lookup_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, nkeys, n, g->d_keys,
                           g->nchunks, oldpos, isnew);
        SG_HIP(hipGetLastError());
    }
    SG_HIP(hipMemsetAsync(nmask + (size_t)n * csz, 0, csz, g->stream));
    for (int p = 0; p < g->np; ++p) {
        dispatch_type(g->dt[p], [&](auto tag) {
            typedef decltype(tag) T;
            T v;
            memcpy(&v, g->bg[p], sizeof(T));
            hipLaunchKernelGGL(remap_chunks_kernel<T>, dim3((unsigned)(n + 1)), dim3(kWG), 0, g->stream,
                               (const T *)g->d_prop[p], (T *)nprop[p], oldpos, nmask, n, csz, v);
        });
        SG_HIP(hipGetLastError());
    }
    // the host ids and masks are read by the queued copies: wait before they may go away
    SG_HIP(hipStreamSynchronize(g->stream));
    g->pool.free(g->d_keys);
    g->pool.free(g->d_mask);
    for (int p = 0; p < g->np; ++p) g->pool.free(g->d_prop[p]);
    g->d_keys = nkeys;
    g->d_mask = nmask;
    for (int p = 0; p < g->np; ++p) g->d_prop[p] = nprop[p];
    g->nchunks = n;
    g->cap = ncap;
    g->nn_valid[0] = g->nn_valid[1] = false;
    g->occ_tab_valid = false;
    g->ebits_valid = false;
    g->ep_valid = false;
    return SG_OK;
}

int restore_background_impl(sg_grid *g, int prop)
{
    for (int p = 0; p < g->np; ++p) {
        if (prop >= 0 && p != prop) continue;
        int rc = normalize_prop(g, p);
        if (rc) return rc;
    }
    // the background chunk's mask stays 0
    SG_HIP(hipMemsetAsync(g->d_mask + (size_t)g->nchunks * g->csz, 0, g->csz, g->stream));
    return SG_OK;
}

}  // namespace sg
