// sg_links.hip -- links between the levels of a multi-resolution set of grids on gfx950.
//
// construct_link_dw / construct_link_up (src/SparseGridGpu/SparseGridGpu.hpp:2090-2262) with
// the kernels count_paddings, collect_paddings, link_construct_{dw,up}_count and
// link_construct_insert_{dw,up} (src/SparseGridGpu/SparseGridGpu_kernels.cuh:775-1000):
//   * the padding points (mask bit PADDING) of this grid inside an inclusive box, listed
//     chunk by chunk in chunk-list order and by increasing offset inside a chunk (the
//     reference orders them inside a chunk through an atomic counter; any order is valid
//     there, this one is deterministic);
//   * down links: for padding point p, the points c = 2 p + p_dw + ((j >> k) & 1)_k,
//     j = 0 .. 2 dim - 1, of the finer grid that exist (EXIST bit); the reference loops
//     j < 2 * dim, so in 3D it visits 6 of the 8 children, and so does this;
//   * up links: the point (p - p_up) / 2 (C division) of the coarser grid, if it exists;
//   * offsets: per padding point the first of its links (exclusive scan), links: the
//     (chunk position, offset) of each linked point in the other grid.
// A point outside the other grid's domain does not exist (the reference's linearizer is
// undefined there).
#include "sg_internal.hpp"

namespace sg {
namespace {

constexpr int kWG = 256;

struct LinkGeom {
    ChunkLin L;
    int dim, edge, csz;
    int64_t sz[3];
};

LinkGeom link_geom(const sg_grid *g)
{
    LinkGeom G;
    G.L = chunk_lin(g);
    G.dim = g->dim;
    G.edge = g->edge;
    G.csz = g->csz;
    for (int d = 0; d < 3; ++d) G.sz[d] = g->sz[d];
    return G;
}

__device__ inline void point_coords(const LinkGeom &G, int64_t chunk_key, int o, int64_t *p)
{
    int64_t bc[3];
    chunk_coords(G.L, chunk_key, bc);
    for (int d = 0; d < G.dim; ++d) {
        p[d] = bc[d] * G.edge + o % G.edge;
        o /= G.edge;
    }
}

__device__ inline bool inside_box(const LinkGeom &G, const int64_t *p, const int32_t *lo, const int32_t *hi)
{
    for (int d = 0; d < G.dim; ++d)
        if (p[d] < lo[d] || p[d] > hi[d]) return false;
    return true;
}

// position (chunk, offset) of point p in a grid and whether it exists there
__device__ inline bool find_point(const LinkGeom &G, const int64_t *keys, int64_t n, const uint8_t *mask,
                                  const int64_t *p, int64_t *pos, int *off)
{
    int64_t bc[3] = {0, 0, 0};
    int o = 0, mul = 1;
    for (int d = 0; d < G.dim; ++d) {
        if (p[d] < 0 || p[d] >= G.sz[d]) return false;
        bc[d] = p[d] / G.edge;
        o += (int)(p[d] % G.edge) * mul;
        mul *= G.edge;
    }
    const int64_t id = chunk_id(G.L, bc);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= n || keys[lo] != id) return false;
    *pos = lo;
    *off = o;
    return (mask[lo * G.csz + o] & SG_EXIST) != 0;
}

struct LinkBox {
    int32_t lo[3], hi[3];
};

// count_paddings: padding points inside the box, per chunk (one workgroup per chunk)
__global__ __launch_bounds__(kWG) void count_pad_kernel(const int64_t *keys, const uint8_t *mask, LinkGeom G,
                                                       LinkBox B, uint32_t *cnt)
{
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const int64_t c = blockIdx.x;
    uint32_t k = 0;
    for (int o = threadIdx.x; o < G.csz; o += kWG) {
        if (!(mask[c * G.csz + o] & SG_PADDING)) continue;
        int64_t p[3];
        point_coords(G, keys[c], o, p);
        k += inside_box(G, p, B.lo, B.hi) ? 1u : 0u;
    }
    if (k) atomicAdd(&s, k);
    __syncthreads();
    if (threadIdx.x == 0) cnt[c] = s;
}

// collect_paddings: the padding points inside the box, by increasing offset in each chunk
__global__ __launch_bounds__(kWG) void collect_pad_kernel(const int64_t *keys, const uint8_t *mask, LinkGeom G,
                                                         LinkBox B, const uint32_t *start, int32_t *pc, int16_t *po)
{
    __shared__ uint32_t s_wave[kWG / 64];
    __shared__ uint32_t s_run;
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    for (int o0 = 0; o0 < G.csz; o0 += kWG) {
        const int o = o0 + threadIdx.x;
        bool has = o < G.csz && (mask[c * G.csz + o] & SG_PADDING);
        if (has) {
            int64_t p[3];
            point_coords(G, keys[c], o, p);
            has = inside_box(G, p, B.lo, B.hi);
        }
        const uint64_t bal = __ballot(has);
        if (lane == 0) s_wave[w] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t pre = s_run, tot = 0;
        for (int q = 0; q < kWG / 64; ++q) {
            if (q < w) pre += s_wave[q];
            tot += s_wave[q];
        }
        if (has) {
            const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
            const uint32_t at = start[c] + pre + (uint32_t)__popcll(bal & lt);
            pc[at] = (int32_t)c;
            po[at] = (int16_t)o;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
}

struct OtherGrid {
    LinkGeom G;
    const int64_t *keys;
    const uint8_t *mask;
    int64_t n;
};

// link_construct_{dw,up}_count (write == false) and link_construct_insert_{dw,up}
// (write == true): the linked points of padding point t
__global__ __launch_bounds__(kWG) void links_kernel(const int64_t *keys, LinkGeom G, const int32_t *pc,
                                                   const int16_t *po, int64_t npad, OtherGrid O, int down,
                                                   LinkBox P, uint32_t *cnt, const uint32_t *start, int32_t *lc,
                                                   int16_t *lo_out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npad) return;
    int64_t p[3];
    point_coords(G, keys[pc[t]], po[t], p);
    uint32_t k = 0;
    if (down) {
        for (int j = 0; j < 2 * G.dim; ++j) {
            int64_t q[3];
            for (int d = 0; d < G.dim; ++d) q[d] = p[d] * 2 + P.lo[d] + ((j >> d) & 1);
            int64_t pos;
            int off;
            if (find_point(O.G, O.keys, O.n, O.mask, q, &pos, &off)) {
                if (start) {
                    lc[start[t] + k] = (int32_t)pos;
                    lo_out[start[t] + k] = (int16_t)off;
                }
                ++k;
            }
        }
    } else {
        int64_t q[3];
        for (int d = 0; d < G.dim; ++d) q[d] = (p[d] - P.lo[d]) / 2;
        int64_t pos;
        int off;
        if (find_point(O.G, O.keys, O.n, O.mask, q, &pos, &off)) {
            if (start) {
                lc[start[t]] = (int32_t)pos;
                lo_out[start[t]] = (int16_t)off;
            }
            k = 1;
        }
    }
    if (cnt) cnt[t] = k;
}

__global__ void fill_u32_kernel(uint32_t *p, int64_t n, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int alloc_link(sg_grid *g, void **p, size_t bytes)
{
    g->pool.free(*p);
    *p = bytes ? g->pool.alloc(bytes) : nullptr;
    if (bytes && !*p) return fail(SG_E_OOM, "link allocation failed");
    return SG_OK;
}

}  // namespace
}  // namespace sg

using namespace sg;

extern "C" {

int sg_construct_link(sg_grid *g, sg_grid *other, int which, const int32_t *lo, const int32_t *hi, const int32_t *p)
{
    if (!g || !other || !lo || !hi || !p) return fail(SG_E_ARG, "null argument");
    if (which != SG_LINK_DOWN && which != SG_LINK_UP) return fail(SG_E_ARG, "link direction must be down or up");
    if (other->dim != g->dim) return fail(SG_E_ARG, "grids of different dimension");
    if (g->nchunks >= ((int64_t)1 << 31) || other->nchunks >= ((int64_t)1 << 31) || g->csz > 32768 ||
        other->csz > 32768)
        return fail(SG_E_RANGE, "link positions are int32 chunks and int16 offsets");
    // the other grid's queued work (its own stream) comes first
    if (other->stream != g->stream) SG_HIP(hipStreamSynchronize(other->stream));
    sg_grid::link_set &L = g->links[which];
    L.n_pad = L.n_links = 0;
    const int64_t n = g->nchunks;
    LinkBox B{}, P{};
    for (int d = 0; d < g->dim; ++d) {
        B.lo[d] = lo[d];
        B.hi[d] = hi[d];
        P.lo[d] = p[d];
    }
    const LinkGeom G = link_geom(g);
    int rc;
    if (n == 0) {
        if ((rc = alloc_link(g, (void **)&L.scan, 4))) return rc;
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(kWG), 0, g->stream, L.scan, (int64_t)1, 0u);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    // padding points of this grid in the box
    DevBuf &b = g->scratch[9];
    if ((rc = ensure(g->pool, b, (size_t)(n + 1) * 8 + 64))) return rc;
    uint32_t *cnt = (uint32_t *)b.p, *start = cnt + (n + 1);
    hipLaunchKernelGGL(count_pad_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_keys, g->d_mask, G, B, cnt);
    SG_HIP(hipGetLastError());
    uint32_t npad = 0;
    if ((rc = exclusive_scan_u32(g, cnt, start, n, &npad))) return rc;
    if ((rc = alloc_link(g, (void **)&L.pad_chunk, (size_t)npad * 4)) ||
        (rc = alloc_link(g, (void **)&L.pad_off, (size_t)npad * 2)) ||
        (rc = alloc_link(g, (void **)&L.scan, (size_t)(npad + 1) * 4)))
        return rc;
    L.n_pad = npad;
    if (npad == 0) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(kWG), 0, g->stream, L.scan, (int64_t)1, 0u);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    hipLaunchKernelGGL(collect_pad_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_keys, g->d_mask, G, B,
                       start, L.pad_chunk, L.pad_off);
    SG_HIP(hipGetLastError());
    // links per padding point, their scan, then the links
    OtherGrid O{link_geom(other), other->d_keys, other->d_mask, other->nchunks};
    if (other->nchunks == 0) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_1d(npad + 1, kWG)), dim3(kWG), 0, g->stream, L.scan,
                           (int64_t)npad + 1, 0u);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    DevBuf &c2 = g->scratch[8];
    if ((rc = ensure(g->pool, c2, (size_t)npad * 4 + 64))) return rc;
    uint32_t *lcnt = (uint32_t *)c2.p;
    const int down = which == SG_LINK_DOWN;
    hipLaunchKernelGGL(links_kernel, dim3(grid_1d(npad, kWG)), dim3(kWG), 0, g->stream, g->d_keys, G, L.pad_chunk,
                       L.pad_off, (int64_t)npad, O, down, P, lcnt, (const uint32_t *)nullptr, (int32_t *)nullptr,
                       (int16_t *)nullptr);
    SG_HIP(hipGetLastError());
    uint32_t nl = 0;
    if ((rc = exclusive_scan_u32(g, lcnt, L.scan, npad, &nl))) return rc;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(kWG), 0, g->stream, L.scan + npad, (int64_t)1, nl);
    SG_HIP(hipGetLastError());
    if ((rc = alloc_link(g, (void **)&L.link_chunk, (size_t)nl * 4)) ||
        (rc = alloc_link(g, (void **)&L.link_off, (size_t)nl * 2)))
        return rc;
    L.n_links = nl;
    if (nl > 0) {
        hipLaunchKernelGGL(links_kernel, dim3(grid_1d(npad, kWG)), dim3(kWG), 0, g->stream, g->d_keys, G,
                           L.pad_chunk, L.pad_off, (int64_t)npad, O, down, P, (uint32_t *)nullptr,
                           (const uint32_t *)L.scan, L.link_chunk, L.link_off);
        SG_HIP(hipGetLastError());
    }
    return SG_OK;
}

int sg_link_sizes(sg_grid *g, int which, int64_t *n_pad, int64_t *n_links)
{
    if (!g || !n_pad || !n_links) return fail(SG_E_ARG, "null argument");
    if (which != SG_LINK_DOWN && which != SG_LINK_UP) return fail(SG_E_ARG, "link direction must be down or up");
    *n_pad = g->links[which].n_pad;
    *n_links = g->links[which].n_links;
    return SG_OK;
}

int sg_get_links(sg_grid *g, int which, int32_t *pad_chunk, int16_t *pad_off, uint32_t *scan, int32_t *link_chunk,
                 int16_t *link_off, int mem)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    if (which != SG_LINK_DOWN && which != SG_LINK_UP) return fail(SG_E_ARG, "link direction must be down or up");
    const sg_grid::link_set &L = g->links[which];
    const hipMemcpyKind k = (mem == SG_MEM_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (pad_chunk && L.n_pad) SG_HIP(hipMemcpyAsync(pad_chunk, L.pad_chunk, (size_t)L.n_pad * 4, k, g->stream));
    if (pad_off && L.n_pad) SG_HIP(hipMemcpyAsync(pad_off, L.pad_off, (size_t)L.n_pad * 2, k, g->stream));
    if (scan && L.scan) SG_HIP(hipMemcpyAsync(scan, L.scan, (size_t)(L.n_pad + 1) * 4, k, g->stream));
    if (link_chunk && L.n_links) SG_HIP(hipMemcpyAsync(link_chunk, L.link_chunk, (size_t)L.n_links * 4, k, g->stream));
    if (link_off && L.n_links) SG_HIP(hipMemcpyAsync(link_off, L.link_off, (size_t)L.n_links * 2, k, g->stream));
    if (mem != SG_MEM_DEVICE) SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

}  // extern "C"
