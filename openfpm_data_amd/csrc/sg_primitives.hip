// sg_primitives.hip -- device-wide exclusive scan and stable LSD radix sort for gfx950.
//
// These replace the CUB/hipCUB calls of the reference flush
// (openfpm::scan -> DeviceScan::ExclusiveSum, src/util/cuda/scan_ofp.cuh:32-78;
//  openfpm::sort -> DeviceRadixSort::SortPairs, src/util/cuda/sort_ofp.cuh:262-402).
//
// Radix sort: 8-bit digits, one tile of 4096 keys per 256-thread workgroup.
//   pass = histogram (LDS atomics) -> device scan of the digit-major count table
//          -> scatter: a stable in-tile rank computed by wave64 ballot matching (for
//             each item, the lanes holding the same digit are found with 8 ballots;
//             rank = popcount of lower peer lanes + the wave's earlier items of the
//             digit + the digit's items in earlier waves) places the tile in LDS in digit
//             order, which is then written out so that each digit's run is contiguous in
//             global memory.
// Stability: items are ranked in index order inside a tile and tiles are ranked in
// tile order by the scanned table, so equal keys keep their input order (as
// CUB SortPairs does), which the flush relies on for its merge order.
#include "sg_internal.hpp"

#include <algorithm>

namespace sg {
namespace {

constexpr int kScanWG = 256;
constexpr int kScanIPT = 16;
constexpr int kScanTile = kScanWG * kScanIPT;
// a scan of up to kScanTile^(kScanMaxLevels-1) items (2^48 with 4096-item tiles)
constexpr int kScanMaxLevels = 5;

constexpr int kSortWG = 256;
constexpr int kSortIPT = 16;
constexpr int kSortTile = kSortWG * kSortIPT;

// inclusive scan across a wave64 with cross-lane shuffles
__device__ inline uint32_t wave_inclusive_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// exclusive scan of one value per thread over a 256-thread block; returns the
// block total in *total
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t s_wave[kScanWG / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_wave[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanWG / 64; ++i) {
        uint32_t s = s_wave[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Scan tiles are 4096 items read as four striped passes: pass j covers items
// [j * 1024, (j + 1) * 1024) of the tile, thread t holding items 4t .. 4t + 3 of it, so each
// load instruction of a wave reads 1 KB contiguous (one 16-byte vector per lane when the
// arrays are 16-byte aligned).
constexpr int kScanVec = 4;
constexpr int kScanPass = kScanIPT / kScanVec;  // 4 passes of 1024 items

template <bool VEC> __device__ inline void scan_load(const uint32_t *in, int64_t e, int64_t n, uint32_t (&v)[kScanVec])
{
    if (VEC && e + kScanVec <= n) {
        const uint4 q = *(const uint4 *)(in + e);
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
        v[3] = q.w;
        return;
    }
#pragma unroll
    for (int i = 0; i < kScanVec; ++i) v[i] = (e + i < n) ? in[e + i] : 0u;
}

template <bool VEC>
__global__ __launch_bounds__(kScanWG) void scan_reduce_kernel(const uint32_t *in, uint32_t *block_sums, int64_t n)
{
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanVec;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanPass; ++j) {
        uint32_t v[kScanVec];
        scan_load<VEC>(in, base + (int64_t)j * (kScanWG * kScanVec), n, v);
        s += v[0] + v[1] + v[2] + v[3];
    }
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// exclusive scan of the tile (in place allowed), offset by block_offsets[tile], or with
// block_sums (the tiles' totals, unscanned) by the sum of the earlier tiles' totals
template <bool VEC>
__global__ __launch_bounds__(kScanWG) void scan_down_kernel(const uint32_t *in, uint32_t *out,
                                                            const uint32_t *block_offsets, int64_t n,
                                                            uint32_t *total_out, const uint32_t *block_sums)
{
    __shared__ uint32_t s_wave[kScanPass][kScanWG / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t prior = 0;
    if (block_sums) {
        uint32_t p = 0;
        for (int64_t i = threadIdx.x; i < (int64_t)blockIdx.x; i += kScanWG) p += block_sums[i];
        block_exclusive_scan(p, &prior);
    }
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanVec;
    uint32_t v[kScanPass][kScanVec], inc[kScanPass];
#pragma unroll
    for (int j = 0; j < kScanPass; ++j) {
        scan_load<VEC>(in, base + (int64_t)j * (kScanWG * kScanVec), n, v[j]);
        inc[j] = wave_inclusive_scan(v[j][0] + v[j][1] + v[j][2] + v[j][3]);
        if (lane == 63) s_wave[j][w] = inc[j];
    }
    __syncthreads();
    uint32_t off = block_offsets ? block_offsets[blockIdx.x] : prior;
#pragma unroll
    for (int j = 0; j < kScanPass; ++j) {
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < kScanWG / 64; ++i) {
            const uint32_t sw = s_wave[j][i];
            if (i < w) pre += sw;
            tot += sw;
        }
        uint32_t run = off + pre + inc[j] - (v[j][0] + v[j][1] + v[j][2] + v[j][3]);
        const int64_t e = base + (int64_t)j * (kScanWG * kScanVec);
        uint32_t o[kScanVec];
#pragma unroll
        for (int i = 0; i < kScanVec; ++i) {
            o[i] = run;
            run += v[j][i];
            if (total_out && e + i == n - 1) *total_out = run;
        }
        if (VEC && e + kScanVec <= n) {
            *(uint4 *)(out + e) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int i = 0; i < kScanVec; ++i)
                if (e + i < n) out[e + i] = o[i];
        }
        off += tot;
    }
}

template <typename K>
__global__ __launch_bounds__(kSortWG) void radix_hist_kernel(const K *keys, uint32_t *counts, int64_t n, int shift,
                                                             int ntiles)
{
    __shared__ uint32_t s_hist[256];
    const int t = threadIdx.x;
    s_hist[t] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
    for (int j = 0; j < kSortIPT; ++j) {
        int64_t i = base + (int64_t)j * kSortWG + t;
        if (i < n) {
            uint32_t d = (uint32_t)(keys[i] >> shift) & 255u;
            atomicAdd(&s_hist[d], 1u);
        }
    }
    __syncthreads();
    counts[(int64_t)t * ntiles + blockIdx.x] = s_hist[t];
}

// Scatter of one pass: the tile's items are first placed in LDS in (digit, index) order
// (stable in-tile ranks from wave64 ballot matching), then written out in that order, so
// each digit's run of the tile goes to consecutive global addresses.  Wave w owns the
// tile's items [w * 64 * IPT, (w + 1) * 64 * IPT), item j of a lane at w * 64 * IPT +
// j * 64 + lane, so index order is (wave, j, lane): a wave ranks its items among equal
// digits by ballots and a running count per digit of its own (LDS, in wave order, no
// barrier); after one barrier the waves' counts are scanned per digit.
template <typename K>
__global__ __launch_bounds__(kSortWG) void radix_scatter_kernel(const K *keys, const uint32_t *vals, K *okeys,
                                                                uint32_t *ovals, const uint32_t *offsets,
                                                                const uint32_t *counts, int64_t n, int shift,
                                                                int ntiles)
{
    __shared__ K s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];
    __shared__ uint32_t s_gbase[256];  // global start of digit d for this tile
    __shared__ uint32_t s_lstart[256]; // start of digit d inside the tile
    __shared__ uint32_t s_wcnt[kSortWG / 64][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    s_gbase[t] = offsets[(int64_t)t * ntiles + blockIdx.x];
    uint32_t tot;
    s_lstart[t] = block_exclusive_scan(counts[(int64_t)t * ntiles + blockIdx.x], &tot);
#pragma unroll
    for (int q = 0; q < kSortWG / 64; ++q) s_wcnt[q][t] = 0;
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)w * 64 * kSortIPT;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    // every item of the thread is loaded before the first is ranked
    K kk[kSortIPT];
    uint32_t vv[kSortIPT];
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        const int64_t i = base + (int64_t)j * 64 + lane;
        kk[j] = (i < n) ? keys[i] : (K)0;
        vv[j] = (i < n) ? vals[i] : 0u;
    }
    __syncthreads();  // s_wcnt cleared
    uint32_t wr[kSortIPT];
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        const bool valid = base + (int64_t)j * 64 + lane < n;
        const uint32_t d = (uint32_t)(kk[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t lr = (uint32_t)__popcll(peers & lt);
        const uint32_t before = valid ? s_wcnt[w][d] : 0u;
        wr[j] = before + lr;
        if (valid && lr == 0) s_wcnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < kSortWG / 64; ++q) {
            const uint32_t c = s_wcnt[q][t];
            s_wcnt[q][t] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortIPT; ++j) {
        if (base + (int64_t)j * 64 + lane < n) {
            const uint32_t d = (uint32_t)(kk[j] >> shift) & 255u;
            const uint32_t pos = s_lstart[d] + s_wcnt[w][d] + wr[j];
            s_key[pos] = kk[j];
            s_val[pos] = vv[j];
        }
    }
    __syncthreads();
    const int64_t tbase = (int64_t)blockIdx.x * kSortTile;
    const int m = (int)((n - tbase < kSortTile) ? (n - tbase) : kSortTile);
    for (int k = t; k < m; k += kSortWG) {
        const K key = s_key[k];
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        const uint32_t pos = s_gbase[d] + (uint32_t)k - s_lstart[d];
        okeys[pos] = key;
        ovals[pos] = s_val[k];
    }
}

template <typename K>
int radix_sort_pairs(sg_grid *g, K *keys, uint32_t *vals, K *keys_tmp, uint32_t *vals_tmp, int64_t n, int bits,
                     K **keys_out, uint32_t **vals_out)
{
    *keys_out = keys;
    *vals_out = vals;
    if (n <= 1 || bits <= 0) return SG_OK;
    if (n >= (int64_t)1 << 31) return fail(SG_E_RANGE, "radix sort supports fewer than 2^31 items");
    const int ntiles = (int)((n + kSortTile - 1) / kSortTile);
    DevBuf &cbuf = g->scratch[10];
    int rc = ensure(g->pool, cbuf, (size_t)256 * ntiles * sizeof(uint32_t) * 2);
    if (rc) return rc;
    uint32_t *counts = (uint32_t *)cbuf.p;
    uint32_t *offs = counts + (size_t)256 * ntiles;
    K *src_k = keys, *dst_k = keys_tmp;
    uint32_t *src_v = vals, *dst_v = vals_tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(radix_hist_kernel<K>, dim3(ntiles), dim3(kSortWG), 0, g->stream, src_k, counts, n, shift,
                           ntiles);
        SG_HIP(hipGetLastError());
        rc = exclusive_scan_u32(g, counts, offs, (int64_t)256 * ntiles, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(radix_scatter_kernel<K>, dim3(ntiles), dim3(kSortWG), 0, g->stream, src_k, src_v, dst_k,
                           dst_v, offs, counts, n, shift, ntiles);
        SG_HIP(hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
    }
    *keys_out = src_k;
    *vals_out = src_v;
    return SG_OK;
}

// ---- stable LSD radix sort of (uint32 key, record of NV 16-byte vectors) ------------
// The same passes with the records travelling with their keys, so that a sorted run is
// read contiguously afterwards (no gather through a permutation).  Tiles of 2048 items
// keep the LDS staging (keys + records) small enough for several workgroups per CU.
constexpr int kRecIPT = 8;
constexpr int kRecTile = kSortWG * kRecIPT;

// the digit counts of kHistTiles consecutive tiles per workgroup (every key of the
// workgroup loaded before the first is counted)
constexpr int kHistTiles = 4;
__global__ __launch_bounds__(kSortWG) void radix_hist_rec_kernel(const uint32_t *keys, uint32_t *counts, int64_t n,
                                                                 int shift, uint32_t dmask, int ntiles)
{
    __shared__ uint32_t s_hist[kHistTiles][256];
    const int t = threadIdx.x;
#pragma unroll
    for (int h = 0; h < kHistTiles; ++h) s_hist[h][t] = 0;
    uint32_t d[kHistTiles][kRecIPT];
#pragma unroll
    for (int h = 0; h < kHistTiles; ++h) {
        const int64_t base = ((int64_t)blockIdx.x * kHistTiles + h) * kRecTile;
#pragma unroll
        for (int j = 0; j < kRecIPT; ++j) {
            const int64_t i = base + (int64_t)j * kSortWG + t;
            d[h][j] = (i < n) ? (keys[i] >> shift) & dmask : 256u;
        }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kHistTiles; ++h)
#pragma unroll
        for (int j = 0; j < kRecIPT; ++j)
            if (d[h][j] < 256u) atomicAdd(&s_hist[h][d[h][j]], 1u);
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kHistTiles; ++h) {
        const int64_t tile = (int64_t)blockIdx.x * kHistTiles + h;
        if (tile < ntiles && (uint32_t)t <= dmask) counts[(int64_t)t * ntiles + tile] = s_hist[h][t];
    }
}

template <int NV>
__global__ __launch_bounds__(kSortWG) void radix_scatter_rec_kernel(const uint32_t *__restrict__ keys,
                                                                    const uint4 *__restrict__ recs,
                                                                    uint32_t *__restrict__ okeys,
                                                                    uint4 *__restrict__ orecs,
                                                                    const uint32_t *__restrict__ offsets,
                                                                    const uint32_t *__restrict__ counts, int64_t n,
                                                                    int shift, int dbits, int ntiles)
{
    __shared__ uint32_t s_key[kRecTile];
    const uint32_t dmask = (1u << dbits) - 1u;
    __shared__ uint4 s_rec[kRecTile * NV];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_lstart[256];
    __shared__ uint32_t s_wcnt[kSortWG / 64][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // the count table holds the 2^dbits digits of this pass
    const bool dig = (uint32_t)t <= dmask;
    s_gbase[t] = dig ? offsets[(int64_t)t * ntiles + blockIdx.x] : 0u;
    uint32_t tot;
    s_lstart[t] = block_exclusive_scan(dig ? counts[(int64_t)t * ntiles + blockIdx.x] : 0u, &tot);
#pragma unroll
    for (int q = 0; q < kSortWG / 64; ++q) s_wcnt[q][t] = 0;
    // Wave w owns the tile's items [w * 64 * IPT, (w + 1) * 64 * IPT), item j of a lane at
    // w * 64 * IPT + j * 64 + lane, so index order is (wave, j, lane): a wave ranks its
    // items among equal digits by ballots and a running count per digit of its own (LDS,
    // in wave order, no barrier); after one barrier the waves' counts are scanned per digit.
    const int64_t base = (int64_t)blockIdx.x * kRecTile + (int64_t)w * 64 * kRecIPT;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    // every item of the thread is loaded before the first is ranked
    uint32_t kk[kRecIPT];
    uint4 rr[kRecIPT][NV];
#pragma unroll
    for (int j = 0; j < kRecIPT; ++j) {
        const int64_t i = base + (int64_t)j * 64 + lane;
        const bool valid = i < n;
        kk[j] = valid ? keys[i] : 0u;
#pragma unroll
        for (int v = 0; v < NV; ++v) rr[j][v] = valid ? recs[i * NV + v] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();  // s_wcnt cleared
    uint32_t wr[kRecIPT];
#pragma unroll
    for (int j = 0; j < kRecIPT; ++j) {
        const bool valid = base + (int64_t)j * 64 + lane < n;
        const uint32_t d = (kk[j] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < dbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t lr = (uint32_t)__popcll(peers & lt);
        const uint32_t before = valid ? s_wcnt[w][d] : 0u;
        wr[j] = before + lr;
        if (valid && lr == 0) s_wcnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        // per digit: the exclusive prefix of the waves' counts
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < kSortWG / 64; ++q) {
            const uint32_t c = s_wcnt[q][t];
            s_wcnt[q][t] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRecIPT; ++j) {
        if (base + (int64_t)j * 64 + lane < n) {
            const uint32_t d = (kk[j] >> shift) & dmask;
            const uint32_t pos = s_lstart[d] + s_wcnt[w][d] + wr[j];
            s_key[pos] = kk[j];
#pragma unroll
            for (int v = 0; v < NV; ++v) s_rec[pos * NV + v] = rr[j][v];
        }
    }
    __syncthreads();
    const int64_t tbase = (int64_t)blockIdx.x * kRecTile;
    const int m = (int)((n - tbase < kRecTile) ? (n - tbase) : kRecTile);
    for (int k = t; k < m; k += kSortWG) {
        const uint32_t key = s_key[k];
        const uint32_t d = (key >> shift) & dmask;
        const uint32_t pos = s_gbase[d] + (uint32_t)k - s_lstart[d];
        okeys[pos] = key;
#pragma unroll
        for (int v = 0; v < NV; ++v) orecs[(int64_t)pos * NV + v] = s_rec[k * NV + v];
    }
}

}  // namespace

int radix_sort_records_u32(sg_grid *g, uint32_t *keys, uint4 *recs, uint32_t *keys_tmp, uint4 *recs_tmp, int nv,
                           int64_t n, int bits, uint32_t **keys_out, uint4 **recs_out, uint32_t *keys_tmp2,
                           uint4 *recs_tmp2)
{
    *keys_out = keys;
    *recs_out = recs;
    if (n <= 1 || bits <= 0) return SG_OK;
    if (nv != 1 && nv != 2) return fail(SG_E_ARG, "records of 1 or 2 16-byte vectors");
    if (n >= (int64_t)1 << 31) return fail(SG_E_RANGE, "radix sort supports fewer than 2^31 items");
    const int ntiles = (int)((n + kRecTile - 1) / kRecTile);
    DevBuf &cbuf = g->scratch[10];
    int rc = ensure(g->pool, cbuf, (size_t)256 * ntiles * sizeof(uint32_t) * 2);
    if (rc) return rc;
    uint32_t *counts = (uint32_t *)cbuf.p;
    uint32_t *offs = counts + (size_t)256 * ntiles;
    uint32_t *src_k = keys, *dst_k = keys_tmp;
    uint4 *src_r = recs, *dst_r = recs_tmp;
    // as few passes as 8-bit digits need, the bits shared evenly among them (21 bits: 3
    // passes of 7, so that a tile's run of one digit is twice as long as with 256 digits)
    const int passes = (bits + 7) / 8, dbits = (bits + passes - 1) / passes;
    for (int shift = 0; shift < bits; shift += dbits) {
        hipLaunchKernelGGL(radix_hist_rec_kernel, dim3((ntiles + kHistTiles - 1) / kHistTiles), dim3(kSortWG), 0,
                           g->stream, src_k, counts, n, shift, (1u << dbits) - 1u, ntiles);
        SG_HIP(hipGetLastError());
        rc = exclusive_scan_u32(g, counts, offs, ((int64_t)1 << dbits) * ntiles, nullptr);
        if (rc) return rc;
        if (nv == 1)
            hipLaunchKernelGGL(radix_scatter_rec_kernel<1>, dim3(ntiles), dim3(kSortWG), 0, g->stream, src_k, src_r,
                               dst_k, dst_r, offs, counts, n, shift, dbits, ntiles);
        else
            hipLaunchKernelGGL(radix_scatter_rec_kernel<2>, dim3(ntiles), dim3(kSortWG), 0, g->stream, src_k, src_r,
                               dst_k, dst_r, offs, counts, n, shift, dbits, ntiles);
        SG_HIP(hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_r, dst_r);
        if (keys_tmp2 && dst_k == keys) {
            // keep the input: the next pass writes the second temporaries
            dst_k = keys_tmp2;
            dst_r = recs_tmp2;
        }
    }
    *keys_out = src_k;
    *recs_out = src_r;
    return SG_OK;
}

// level sizes of a scan of n items: n, ceil(n/T), ceil(ceil(n/T)/T), ... until one tile
static int scan_levels(int64_t n, int64_t (&sizes)[kScanMaxLevels])
{
    int nl = 0;
    sizes[nl++] = n;
    while (sizes[nl - 1] > kScanTile && nl < kScanMaxLevels) {
        sizes[nl] = (sizes[nl - 1] + kScanTile - 1) / kScanTile;
        ++nl;
    }
    return nl;
}

size_t scan_ws_words(int64_t n)
{
    if (n <= 0) return 0;
    int64_t sizes[kScanMaxLevels];
    const int nl = scan_levels(n, sizes);
    size_t need = 0;
    for (int l = 1; l < nl; ++l) need += 2 * (size_t)sizes[l];
    return need;
}

int exclusive_scan_u32_ws(hipStream_t stream, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total,
                          uint32_t *ws, size_t ws_words)
{
    if (n <= 0) return SG_OK;
    int64_t sizes[kScanMaxLevels];
    const int nl = scan_levels(n, sizes);
    if (sizes[nl - 1] > kScanTile) return fail(SG_E_RANGE, "scan too long");
    if (ws_words < scan_ws_words(n)) return fail(SG_E_RANGE, "scan workspace too small");
    // per level l >= 1: sums[l] (block sums of level l-1) and offs[l] (their scan)
    uint32_t *sums[kScanMaxLevels] = {}, *offs[kScanMaxLevels] = {};
    size_t at = 0;
    for (int l = 1; l < nl; ++l) {
        sums[l] = ws + at;
        at += sizes[l];
        offs[l] = ws + at;
        at += sizes[l];
    }
    auto aligned = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    const uint32_t *lin = in;
    for (int l = 0; l + 1 < nl; ++l) {
        if (aligned(lin))
            hipLaunchKernelGGL(scan_reduce_kernel<true>, dim3((unsigned)sizes[l + 1]), dim3(kScanWG), 0, stream,
                               lin, sums[l + 1], sizes[l]);
        else
            hipLaunchKernelGGL(scan_reduce_kernel<false>, dim3((unsigned)sizes[l + 1]), dim3(kScanWG), 0, stream,
                               lin, sums[l + 1], sizes[l]);
        SG_HIP(hipGetLastError());
        lin = sums[l + 1];
    }
    auto down = [&](const uint32_t *src, uint32_t *dst, const uint32_t *bo, int64_t len, uint32_t *tot,
                    const uint32_t *bs) {
        const dim3 grid((unsigned)((len + kScanTile - 1) / kScanTile));
        if (aligned(src) && aligned(dst))
            hipLaunchKernelGGL(scan_down_kernel<true>, grid, dim3(kScanWG), 0, stream, src, dst, bo, len, tot, bs);
        else
            hipLaunchKernelGGL(scan_down_kernel<false>, grid, dim3(kScanWG), 0, stream, src, dst, bo, len, tot, bs);
    };
    if (nl == 2) {
        // one level of tiles (up to kScanTile of them): each tile adds up the earlier
        // tiles' totals itself, no scan of the totals between the two launches
        down(in, out, nullptr, n, d_total, sums[1]);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    for (int l = nl - 1; l >= 1; --l) {
        // scan level l (its single top tile or using offsets from level l+1)
        const uint32_t *bo = (l + 1 < nl) ? offs[l + 1] : nullptr;
        down(sums[l], offs[l], bo, sizes[l], (uint32_t *)nullptr, (const uint32_t *)nullptr);
        SG_HIP(hipGetLastError());
    }
    // the inclusive total is written by the thread owning item n-1 (works in place)
    down(in, out, nl > 1 ? offs[1] : nullptr, n, d_total, (const uint32_t *)nullptr);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int exclusive_scan_u32(sg_grid *g, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total)
{
    if (n <= 0) {
        if (total) *total = 0;
        return SG_OK;
    }
    const size_t words = scan_ws_words(n) + 2;
    DevBuf &sb = g->scratch[11];
    int rc = ensure(g->pool, sb, words * sizeof(uint32_t));
    if (rc) return rc;
    uint32_t *d_total = (uint32_t *)sb.p + (words - 2);
    if ((rc = exclusive_scan_u32_ws(g->stream, in, out, n, total ? d_total : nullptr, (uint32_t *)sb.p, words - 2)))
        return rc;
    if (total) {
        // through a pinned word, so that copies queued before it need no wait of their own
        uint32_t *hw = host_words(g);
        if (!hw) return fail(SG_E_OOM, "pinned host allocation failed");
        SG_HIP(hipMemcpyAsync(hw, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
        SG_HIP(hipStreamSynchronize(g->stream));
        *total = hw[0];
    }
    return SG_OK;
}

uint32_t *host_words(sg_grid *g)
{
    if (!g->h_words && hipHostMalloc((void **)&g->h_words, 16 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        g->h_words = nullptr;
    return g->h_words;
}

int radix_sort_pairs_u64(sg_grid *g, uint64_t *keys, uint32_t *vals, uint64_t *keys_tmp, uint32_t *vals_tmp,
                         int64_t n, int bits, uint64_t **keys_out, uint32_t **vals_out)
{
    return radix_sort_pairs<uint64_t>(g, keys, vals, keys_tmp, vals_tmp, n, bits, keys_out, vals_out);
}

int radix_sort_pairs_u32(sg_grid *g, uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                         int64_t n, int bits, uint32_t **keys_out, uint32_t **vals_out)
{
    return radix_sort_pairs<uint32_t>(g, keys, vals, keys_tmp, vals_tmp, n, bits, keys_out, vals_out);
}

// streaming copy of 16-byte vectors (the HBM ceiling measured beside the stencil): every
// thread moves 4 vectors per loop trip, grid-stride, plain or non-temporal loads and stores
typedef unsigned int v4u_t __attribute__((ext_vector_type(4)));
template <bool NT> __global__ __launch_bounds__(256) void stream_copy_kernel(const v4u_t *__restrict__ src,
                                                                            v4u_t *__restrict__ dst, int64_t nvec)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += 4 * stride) {
        v4u_t v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t j = i + k * stride;
            if (j < nvec) v[k] = NT ? __builtin_nontemporal_load(src + j) : src[j];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t j = i + k * stride;
            if (j < nvec) {
                if (NT) __builtin_nontemporal_store(v[k], dst + j);
                else dst[j] = v[k];
            }
        }
    }
}

}  // namespace sg

extern "C" int sg_stream_copy(void *dst, const void *src, int64_t bytes, int flags, void *stream)
{
    using namespace sg;
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(SG_E_ARG, "bad copy arguments");
    if (bytes % 16 || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return fail(SG_E_ARG, "the copy needs 16-byte aligned buffers and a multiple of 16 bytes");
    const int64_t nvec = bytes / 16;
    if (nvec == 0) return SG_OK;
    int dev = 0, cus = 256;
    SG_HIP(hipGetDevice(&dev));
    SG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 8 workgroups of 256 threads per CU, each thread moving 4 vectors per trip
    const int64_t want = (nvec + 4 * 256 - 1) / (4 * 256);
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)cus * 8));
    if (flags & SG_COPY_NONTEMPORAL)
        hipLaunchKernelGGL(stream_copy_kernel<true>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (const v4u_t *)src, (v4u_t *)dst, nvec);
    else
        hipLaunchKernelGGL(stream_copy_kernel<false>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                           (const v4u_t *)src, (v4u_t *)dst, nvec);
    SG_HIP(hipGetLastError());
    return SG_OK;
}
