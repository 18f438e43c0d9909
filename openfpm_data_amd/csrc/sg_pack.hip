// sg_pack.hip -- pack / unpack / remove of sparse-grid sections for gfx950.
//
// Reference: SparseGridGpu::packRequest / packCalculate / pack / packFinalize
// (src/SparseGridGpu/SparseGridGpu.hpp:3040-3340), unpack + removeAddUnpackFinalize
// (:3492-3499, :3665-3733), remove(Box) (:3528-3531) and the kernels
// calc_exist_points_with_boxes / pack_data / last_scan_point / remove_points
// (SparseGridGpu_kernels.cuh:1235-1697).
//
// Packed form of one box, sections in this order (A = sizeof(indexT), 4 or 8):
//   uint64 n_cnk | int32 box_start[dim] | int32 grid_size[dim] | indexT ids[n_cnk]
//   | uint32 scan[n_cnk + 1], padded to A | per packed property T[n_pnt], one after the
//   other, padded to A | int16 offsets[n_pnt], padded to A | uint8 masks[n_pnt], padded to A
// ids are the (sorted) chunks holding at least one existing point inside the inclusive
// box; scan[i] is the first point of chunk i and scan[n_cnk] = n_pnt.  Points of a chunk
// are written in increasing offset (the reference's order comes from an atomic counter,
// so any order is a valid instance of its format); padding bytes are zero.
//
// Kernels: a count pass (one workgroup per chunk, per box: existing points inside),
// per-box scans of the counts and of the non-empty flags, then a write pass (one
// workgroup per chunk, wave-ballot compaction keeps offsets ascending).  Unpack expands
// a section into point records (key = packed coordinate - box start + destination
// start) appended to the insert buffer; the finalize removes the queued boxes and
// flushes with the unpacked properties taking the new values.
#include "sg_internal.hpp"

#include <algorithm>
#include <mutex>

namespace sg {
namespace {

constexpr int kWG = 256;

struct Geom {
    int dim, edge, csz;
    ChunkLin L;
};

__device__ inline void chunk_origin(const Geom &G, int64_t id, int64_t *o)
{
    chunk_coords(G.L, id, o);
    for (int d = 0; d < G.dim; ++d) o[d] *= G.edge;
}

__device__ inline void local_coord(const Geom &G, int off, int64_t *l)
{
    for (int d = 0; d < G.dim; ++d) {
        l[d] = off % G.edge;
        off /= G.edge;
    }
}

struct Boxes {
    int n;
    const int32_t *lo, *hi;  // [n][dim]
};

// property set of a grid for kernels that write or read every property
struct PropSet {
    int np;
    void *p[kMaxProps];
    int ts[kMaxProps];
    uint64_t bg[kMaxProps];
};

__device__ inline void copy_scalar(unsigned char *dst, const unsigned char *src, int ts)
{
    // byte copies: packed property arrays follow each other without alignment
    for (int b = 0; b < ts; ++b) dst[b] = src[b];
}

__device__ inline void clear_slot(const PropSet &P, uint8_t *mask, int64_t slot)
{
    mask[slot] = 0;
    for (int q = 0; q < P.np; ++q)
        copy_scalar((unsigned char *)P.p[q] + slot * P.ts[q], (const unsigned char *)&P.bg[q], P.ts[q]);
}

__device__ inline bool chunk_meets_box(const Geom &G, const int64_t *org, const int32_t *lo, const int32_t *hi)
{
    for (int d = 0; d < G.dim; ++d)
        if (org[d] + G.edge - 1 < lo[d] || org[d] > hi[d]) return false;
    return true;
}

__device__ inline bool point_in_box(const Geom &G, const int64_t *c, const int32_t *lo, const int32_t *hi)
{
    for (int d = 0; d < G.dim; ++d)
        if (c[d] < lo[d] || c[d] > hi[d]) return false;
    return true;
}

// existing points of chunk c inside each box: cnt[k * stride + c], flag = cnt != 0
// (calc_exist_points_with_boxes, SparseGridGpu_kernels.cuh:1351-1420)
__global__ __launch_bounds__(kWG) void pack_count_kernel(const int64_t *__restrict__ keys, int64_t n,
                                                        const uint8_t *__restrict__ mask, Geom G, Boxes B,
                                                        uint32_t *__restrict__ cnt, uint32_t *__restrict__ flag,
                                                        int64_t stride)
{
    __shared__ uint32_t s_cnt;
    const int64_t c = blockIdx.x;
    if (c >= n) return;
    int64_t org[3] = {0, 0, 0};
    chunk_origin(G, keys[c], org);
    for (int k = 0; k < B.n; ++k) {
        const int32_t *lo = B.lo + k * G.dim, *hi = B.hi + k * G.dim;
        uint32_t mine = 0;
        if (chunk_meets_box(G, org, lo, hi)) {
            for (int o = threadIdx.x; o < G.csz; o += kWG) {
                if (!(mask[c * G.csz + o] & SG_EXIST)) continue;
                int64_t l[3] = {0, 0, 0}, p[3];
                local_coord(G, o, l);
                for (int d = 0; d < G.dim; ++d) p[d] = org[d] + l[d];
                mine += point_in_box(G, p, lo, hi) ? 1u : 0u;
            }
        }
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        if (mine) atomicAdd(&s_cnt, mine);
        __syncthreads();
        if (threadIdx.x == 0) {
            cnt[k * stride + c] = s_cnt;
            flag[k * stride + c] = s_cnt ? 1u : 0u;
        }
        __syncthreads();
    }
}

struct PackOut {
    unsigned char *ids;     // indexT [n_cnk]
    uint32_t *scan;         // [n_cnk + 1]
    unsigned char *data[kMaxProps];  // per packed property [n_pnt]
    int ts[kMaxProps];
    const unsigned char *src[kMaxProps];  // grid property arrays
    int nprop;
    int16_t *offs;          // [n_pnt]
    uint8_t *masks;         // [n_pnt]
    int A;                  // sizeof(indexT)
};

// write the points of box (lo, hi) of chunk c at their packed positions
// (get_exist_points_with_boxes + pack_data, SparseGridGpu_kernels.cuh:1437-1697)
__global__ __launch_bounds__(kWG) void pack_write_kernel(const int64_t *__restrict__ keys, int64_t n,
                                                        const uint8_t *__restrict__ mask, Geom G, const int32_t *lo,
                                                        const int32_t *hi, const uint32_t *__restrict__ pscan,
                                                        const uint32_t *__restrict__ cscan,
                                                        const uint32_t *__restrict__ cnt, PackOut O)
{
    __shared__ uint32_t s_wave[kWG / 64];
    __shared__ uint32_t s_run;
    const int64_t c = blockIdx.x;
    if (c >= n || cnt[c] == 0) return;
    const uint32_t slot = cscan[c], start = pscan[c];
    const int64_t key = keys[c];
    if (threadIdx.x == 0) {
        if (O.A == 8) copy_scalar(O.ids + (int64_t)slot * 8, (const unsigned char *)&key, 8);
        else {
            const int32_t k4 = (int32_t)key;
            copy_scalar(O.ids + (int64_t)slot * 4, (const unsigned char *)&k4, 4);
        }
        O.scan[slot] = start;
        s_run = 0;
    }
    int64_t org[3] = {0, 0, 0};
    chunk_origin(G, key, org);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    for (int o0 = 0; o0 < G.csz; o0 += kWG) {
        const int o = o0 + threadIdx.x;
        bool has = false;
        if (o < G.csz && (mask[c * G.csz + o] & SG_EXIST)) {
            int64_t l[3] = {0, 0, 0}, p[3];
            local_coord(G, o, l);
            for (int d = 0; d < G.dim; ++d) p[d] = org[d] + l[d];
            has = point_in_box(G, p, lo, hi);
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
            const int64_t j = (int64_t)start + pre + (uint32_t)__popcll(bal & lt);
            const int64_t slot_pt = c * G.csz + o;
            for (int q = 0; q < O.nprop; ++q)
                copy_scalar(O.data[q] + j * O.ts[q], O.src[q] + slot_pt * O.ts[q], O.ts[q]);
            O.offs[j] = (int16_t)o;
            O.masks[j] = mask[slot_pt];
        }
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
}

struct UnpackIn {
    const unsigned char *ids;
    const uint32_t *scan;
    const unsigned char *data[kMaxProps];  // per grid property: packed array or null
    const int16_t *offs;
    int A;
    int64_t pbsz[3];   // chunk counts of the packing grid
    int32_t shift[3];  // destination start - packed box start
};

// packed chunk i -> point records at base + scan[i] .. (copy_packed_data_to_chunks,
// SparseGridGpu_kernels.cuh:1591-1632; mask |= EXIST)
__global__ __launch_bounds__(kWG) void unpack_records_kernel(UnpackIn U, int64_t n_cnk, int64_t n_pnt, Geom G,
                                                            int64_t sz0, int64_t sz1, int64_t sz2, PropSet P,
                                                            int64_t *__restrict__ rkey, uint8_t *__restrict__ rmask,
                                                            int *__restrict__ bad)
{
    const int64_t i = blockIdx.x;
    if (i >= n_cnk) return;
    int64_t id;
    if (U.A == 8) copy_scalar((unsigned char *)&id, U.ids + i * 8, 8);
    else {
        int32_t k4;
        copy_scalar((unsigned char *)&k4, U.ids + i * 4, 4);
        id = k4;
    }
    // chunk coordinates in the packing grid (same linearizer, its own chunk counts)
    ChunkLin PL = G.L;
    for (int d = 0; d < 3; ++d) PL.bsz[d] = U.pbsz[d];
    int64_t bc[3] = {0, 0, 0};
    chunk_coords(PL, id, bc);
    const int64_t sz[3] = {sz0, sz1, sz2};
    // a malformed section (scan not starting at 0, not increasing or past n_pnt) is refused,
    // not followed: the records [0, n_pnt) are then exactly the chunks' ranges
    const int64_t j0 = U.scan[i], j1 = U.scan[i + 1];
    if (j0 > j1 || j1 > n_pnt || (i == 0 && j0 != 0)) {
        if (threadIdx.x == 0) *bad = 1;
        return;
    }
    for (int64_t j = j0 + threadIdx.x; j < j1; j += kWG) {
        const int off = (uint16_t)U.offs[j];
        int64_t l[3] = {0, 0, 0}, c[3] = {0, 0, 0};
        local_coord(G, off, l);
        bool ok = off < G.csz;
        for (int d = 0; d < G.dim; ++d) {
            c[d] = bc[d] * G.edge + l[d] + U.shift[d];
            ok = ok && c[d] >= 0 && c[d] < sz[d];
        }
        if (!ok) {
            *bad = 1;
            rkey[j] = 0;
            rmask[j] = 0;
            continue;
        }
        int64_t cb[3] = {0, 0, 0}, lo = 0;
        for (int d = G.dim - 1; d >= 0; --d) {
            cb[d] = c[d] / G.edge;
            lo = lo * G.edge + c[d] % G.edge;
        }
        rkey[j] = chunk_id(G.L, cb) * G.csz + lo;
        rmask[j] = SG_EXIST;
        for (int q = 0; q < P.np; ++q) {
            unsigned char *dst = (unsigned char *)P.p[q] + (int64_t)j * P.ts[q];
            if (U.data[q]) copy_scalar(dst, U.data[q] + (int64_t)j * P.ts[q], P.ts[q]);
            else copy_scalar(dst, (const unsigned char *)&P.bg[q], P.ts[q]);
        }
    }
}

// remove_points (SparseGridGpu_kernels.cuh:1301-1346): existing points inside a queued box
// get mask 0 (and the background value)
__global__ __launch_bounds__(kWG) void remove_boxes_kernel(const int64_t *__restrict__ keys, int64_t n,
                                                          uint8_t *__restrict__ mask, Geom G, Boxes B, PropSet P)
{
    const int64_t c = blockIdx.x;
    if (c >= n) return;
    int64_t org[3] = {0, 0, 0};
    chunk_origin(G, keys[c], org);
    bool any = false;
    for (int k = 0; k < B.n; ++k) any = any || chunk_meets_box(G, org, B.lo + k * G.dim, B.hi + k * G.dim);
    if (!any) return;
    for (int o = threadIdx.x; o < G.csz; o += kWG) {
        if (!(mask[c * G.csz + o] & SG_EXIST)) continue;
        int64_t l[3] = {0, 0, 0}, p[3];
        local_coord(G, o, l);
        for (int d = 0; d < G.dim; ++d) p[d] = org[d] + l[d];
        for (int k = 0; k < B.n; ++k)
            if (point_in_box(G, p, B.lo + k * G.dim, B.hi + k * G.dim)) {
                clear_slot(P, mask, c * G.csz + o);
                break;
            }
    }
}

__device__ inline int64_t lower_bound_keys(const int64_t *a, int64_t n, int64_t x)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// device remove(coord) (BlockMapGpu_ker.cuh:322-332) for a list of point keys
__global__ void remove_points_kernel(const int64_t *__restrict__ keys, int64_t n, uint8_t *__restrict__ mask,
                                     const int64_t *__restrict__ pk, int64_t m, int csz, PropSet P)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t k = pk[i];
    if (k < 0) return;
    const int64_t c = lower_bound_keys(keys, n, k / csz);
    if (c < n && keys[c] == k / csz) clear_slot(P, mask, c * csz + k % csz);
}

Geom geom_of(const sg_grid *g)
{
    Geom G;
    G.dim = g->dim;
    G.edge = g->edge;
    G.csz = g->csz;
    G.L = chunk_lin(g);
    return G;
}

PropSet grid_props(sg_grid *g)
{
    PropSet P;
    P.np = g->np;
    for (int q = 0; q < g->np; ++q) {
        P.p[q] = g->d_prop[q];
        P.ts[q] = g->ts[q];
        P.bg[q] = 0;
        memcpy(&P.bg[q], g->bg[q], g->ts[q]);
    }
    return P;
}

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

int check_pack_args(const sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop,
                    const int32_t *props, int A)
{
    if (nbox < 0 || nprop < 0) return fail(SG_E_ARG, "negative count");
    if (nbox > 0 && (!lo || !hi)) return fail(SG_E_ARG, "null boxes");
    if (nprop > 0 && !props) return fail(SG_E_ARG, "null property list");
    if (A != 4 && A != 8) return fail(SG_E_ARG, "index_bytes must be 4 or 8");
    for (int q = 0; q < nprop; ++q)
        if (props[q] < 0 || props[q] >= g->np) return fail(SG_E_ARG, "bad property");
    if (g->csz > 32768) return fail(SG_E_ARG, "chunk too large for 16-bit offsets");
    return SG_OK;
}

int64_t point_bytes(const sg_grid *g, int nprop, const int32_t *props)
{
    int64_t s = 0;
    for (int q = 0; q < nprop; ++q) s += g->ts[props[q]];
    return s;
}

int64_t section_bytes(const sg_grid *g, int64_t n_cnk, int64_t n_pnt, int64_t psz, int A)
{
    return 8 + 2 * g->dim * 4 + A * n_cnk + align_up((n_cnk + 1) * 4, A) + align_up(n_pnt * psz, A) +
           align_up(n_pnt * 2, A) + align_up(n_pnt, A);
}

// per box: point counts + their scan and the chunk flags + their scan, totals to the host
struct PackPlan {
    int64_t stride = 0;
    uint32_t *cnt = nullptr, *flag = nullptr, *pscan = nullptr, *cscan = nullptr;
    const int32_t *d_lo = nullptr, *d_hi = nullptr;  // device copy of the boxes, [nbox][dim]
    std::vector<int64_t> n_cnk, n_pnt;
};

int plan_pack(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, PackPlan &pl)
{
    const int64_t n = g->nchunks;
    pl.stride = n + 1;
    pl.n_cnk.assign(nbox, 0);
    pl.n_pnt.assign(nbox, 0);
    if (nbox == 0) return SG_OK;
    int rc;
    DevBuf &b_boxes = g->scratch[0], &b_cnt = g->scratch[1];
    const size_t bb = (size_t)nbox * g->dim * 4;
    if ((rc = ensure(g->pool, b_boxes, 2 * bb + 64))) return rc;
    const size_t per = (size_t)nbox * pl.stride;
    if ((rc = ensure(g->pool, b_cnt, per * 4 * sizeof(uint32_t) + 64))) return rc;
    int32_t *d_lo = (int32_t *)b_boxes.p, *d_hi = d_lo + (size_t)nbox * g->dim;
    SG_HIP(hipMemcpyAsync(d_lo, lo, bb, hipMemcpyHostToDevice, g->stream));
    SG_HIP(hipMemcpyAsync(d_hi, hi, bb, hipMemcpyHostToDevice, g->stream));
    pl.d_lo = d_lo;
    pl.d_hi = d_hi;
    pl.cnt = (uint32_t *)b_cnt.p;
    pl.flag = pl.cnt + per;
    pl.pscan = pl.flag + per;
    pl.cscan = pl.pscan + per;
    // the trailing entry of every box row is 0 (its scan is the row total)
    SG_HIP(hipMemsetAsync(pl.cnt, 0, per * 2 * sizeof(uint32_t), g->stream));
    if (n > 0) {
        Boxes B{nbox, d_lo, d_hi};
        hipLaunchKernelGGL(pack_count_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_keys, n, g->d_mask,
                           geom_of(g), B, pl.cnt, pl.flag, pl.stride);
        SG_HIP(hipGetLastError());
    }
    for (int k = 0; k < nbox; ++k) {
        uint32_t tp = 0, tc = 0;
        if ((rc = exclusive_scan_u32(g, pl.cnt + k * pl.stride, pl.pscan + k * pl.stride, pl.stride, &tp))) return rc;
        if ((rc = exclusive_scan_u32(g, pl.flag + k * pl.stride, pl.cscan + k * pl.stride, pl.stride, &tc))) return rc;
        pl.n_pnt[k] = tp;
        pl.n_cnk[k] = tc;
    }
    return SG_OK;
}

// header of one packed section: chunk count, box start and grid size, and the scan's last
// entry (n_pnt); passed by value to the kernel that writes them, so nothing is staged on
// the host
struct SectionHead {
    uint64_t nc;
    uint32_t np;
    int dim;
    int32_t lo[3], sz[3];
};

__global__ void pack_header_kernel(unsigned char *s0, uint32_t *scan_last, SectionHead h)
{
    if (threadIdx.x != 0) return;
    // sections start at multiples of 4 bytes only: byte copies
    copy_scalar(s0, (const unsigned char *)&h.nc, 8);
    for (int d = 0; d < h.dim; ++d) {
        copy_scalar(s0 + 8 + 4 * d, (const unsigned char *)&h.lo[d], 4);
        copy_scalar(s0 + 8 + 4 * (h.dim + d), (const unsigned char *)&h.sz[d], 4);
    }
    copy_scalar((unsigned char *)scan_last, (const unsigned char *)&h.np, 4);
}

// write the packed sections of the boxes to the device buffer buf (cap bytes) from a
// plan: per box k, rows k * stride of cnt / pscan / cscan and corners d_lo / d_hi + k * dim
// (device), lo + k * dim (host); queued on the grid's stream, allocates nothing
int pack_sections(sg_grid *g, int nbox, const int32_t *lo, int nprop, const int32_t *props, int A,
                  unsigned char *buf, int64_t cap, int64_t stride, const uint32_t *cnt, const uint32_t *pscan,
                  const uint32_t *cscan, const int32_t *d_lo, const int32_t *d_hi, const int64_t *n_cnk,
                  const int64_t *n_pnt)
{
    const int64_t psz = point_bytes(g, nprop, props);
    int64_t at = 0;
    const int64_t n = g->nchunks;
    const int hsz = 8 + 2 * g->dim * 4;
    for (int k = 0; k < nbox; ++k) {
        const int64_t nc = n_cnk[k], np = n_pnt[k];
        const int64_t sb = section_bytes(g, nc, np, psz, A);
        if (at + sb > cap) return fail(SG_E_RANGE, "pack buffer too small");
        unsigned char *s0 = buf + at;
        SG_HIP(hipMemsetAsync(s0, 0, (size_t)sb, g->stream));
        int64_t q = hsz;
        PackOut O;
        O.A = A;
        O.ids = s0 + q;
        q += A * nc;
        O.scan = (uint32_t *)(s0 + q);
        q += align_up((nc + 1) * 4, A);
        SectionHead h{};
        h.nc = (uint64_t)nc;
        h.np = (uint32_t)np;
        h.dim = g->dim;
        for (int d = 0; d < g->dim; ++d) {
            h.lo[d] = lo[k * g->dim + d];
            h.sz[d] = (int32_t)g->sz[d];
        }
        hipLaunchKernelGGL(pack_header_kernel, dim3(1), dim3(64), 0, g->stream, s0, O.scan + nc, h);
        SG_HIP(hipGetLastError());
        O.nprop = nprop;
        int64_t dq = q;
        for (int pi = 0; pi < nprop; ++pi) {
            O.data[pi] = s0 + dq;
            O.ts[pi] = g->ts[props[pi]];
            O.src[pi] = (const unsigned char *)g->d_prop[props[pi]];
            dq += np * g->ts[props[pi]];
        }
        q += align_up(np * psz, A);
        O.offs = (int16_t *)(s0 + q);
        q += align_up(np * 2, A);
        O.masks = s0 + q;
        if (n > 0 && np > 0) {
            hipLaunchKernelGGL(pack_write_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_keys, n, g->d_mask,
                               geom_of(g), d_lo + (size_t)k * g->dim, d_hi + (size_t)k * g->dim, pscan + k * stride,
                               cscan + k * stride, cnt + k * stride, O);
            SG_HIP(hipGetLastError());
        }
        at += sb;
    }
    return SG_OK;
}

// the packed sections from a plan, complete when the call returns
int pack_into_device(sg_grid *g, int nbox, const int32_t *lo, int nprop, const int32_t *props, int A,
                     unsigned char *buf, int64_t cap, const PackPlan &pl)
{
    int rc = pack_sections(g, nbox, lo, nprop, props, A, buf, cap, pl.stride, pl.cnt, pl.pscan, pl.cscan, pl.d_lo,
                           pl.d_hi, pl.n_cnk.data(), pl.n_pnt.data());
    if (rc) return rc;
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

// ---- pack without device allocation (sg_pack_noalloc) -----------------------------
// The count / scan temporaries live in one static device array per GPU.  Host code never
// uses the array's name as an address: its device address comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that address is passed on.
static __device__ __align__(256) unsigned char scratch[SG_PACK_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device
int noalloc_scratch(void **out)
{
    int dev = -1;
    SG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SG_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SG_OK;
}

constexpr size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// scratch bytes of a pack of nbox boxes over n chunks: the box corners, the four count /
// scan rows of every box, the per-box totals and the scan workspace
size_t noalloc_pack_bytes(int nbox, int dim, int64_t n)
{
    const size_t stride = (size_t)n + 1;
    return round256((size_t)2 * nbox * dim * 4) + round256((size_t)4 * nbox * stride * 4) +
           round256((size_t)2 * nbox * 4) + round256(scan_ws_words((int64_t)stride) * 4);
}

}  // namespace
}  // namespace sg

using namespace sg;

extern "C" {

int sg_pack_size(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop, const int32_t *props,
                 int index_bytes, int64_t *bytes)
{
    if (!g || !bytes) return fail(SG_E_ARG, "null argument");
    int rc = check_pack_args(g, nbox, lo, hi, nprop, props, index_bytes);
    if (rc) return rc;
    PackPlan pl;
    if ((rc = plan_pack(g, nbox, lo, hi, pl))) return rc;
    const int64_t psz = point_bytes(g, nprop, props);
    int64_t tot = 0;
    for (int k = 0; k < nbox; ++k) tot += section_bytes(g, pl.n_cnk[k], pl.n_pnt[k], psz, index_bytes);
    *bytes = tot;
    return SG_OK;
}

int sg_pack(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop, const int32_t *props,
            int index_bytes, void *buf, int64_t cap, int mem)
{
    if (!g || (!buf && cap > 0)) return fail(SG_E_ARG, "null argument");
    int rc = check_pack_args(g, nbox, lo, hi, nprop, props, index_bytes);
    if (rc) return rc;
    PackPlan pl;
    if ((rc = plan_pack(g, nbox, lo, hi, pl))) return rc;
    const int64_t psz = point_bytes(g, nprop, props);
    int64_t tot = 0;
    for (int k = 0; k < nbox; ++k) tot += section_bytes(g, pl.n_cnk[k], pl.n_pnt[k], psz, index_bytes);
    if (tot > cap) return fail(SG_E_RANGE, "pack buffer too small");
    if (mem == SG_MEM_DEVICE) {
        rc = pack_into_device(g, nbox, lo, nprop, props, index_bytes, (unsigned char *)buf, cap, pl);
        return rc;
    }
    DevBuf &st = g->scratch[2];
    if ((rc = ensure(g->pool, st, (size_t)tot + 64))) return rc;
    if ((rc = pack_into_device(g, nbox, lo, nprop, props, index_bytes, (unsigned char *)st.p, tot, pl)))
        return rc;
    SG_HIP(hipMemcpyAsync(buf, st.p, (size_t)tot, hipMemcpyDeviceToHost, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_pack_noalloc(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop, const int32_t *props,
                    int index_bytes, void *dev_buf, int64_t cap, int64_t *written)
{
    if (!g || !written || (!dev_buf && cap > 0)) return fail(SG_E_ARG, "null argument");
    int rc = check_pack_args(g, nbox, lo, hi, nprop, props, index_bytes);
    if (rc) return rc;
    if (nbox > SG_PACK_NOALLOC_MAX_BOXES)
        return fail(SG_E_RANGE, "sg_pack_noalloc: " + std::to_string(nbox) + " boxes, at most " +
                                    std::to_string(SG_PACK_NOALLOC_MAX_BOXES));
    const int64_t n = g->nchunks, stride = n + 1;
    const size_t need = noalloc_pack_bytes(nbox, g->dim, n);
    if (need > (size_t)SG_PACK_NOALLOC_SCRATCH_BYTES)
        return fail(SG_E_RANGE, "sg_pack_noalloc: " + std::to_string(nbox) + " boxes over " + std::to_string(n) +
                                    " chunks need " + std::to_string(need) + " scratch bytes, more than the " +
                                    std::to_string((size_t)SG_PACK_NOALLOC_SCRATCH_BYTES) + " reserved");
    *written = 0;
    if (nbox == 0) return SG_OK;
    void *base = nullptr;
    if ((rc = noalloc_scratch(&base))) return rc;
    // carve the scratch: corners | cnt, flag, pscan, cscan rows | totals | scan workspace
    unsigned char *p = (unsigned char *)base;
    const size_t bb = (size_t)nbox * g->dim * 4, per = (size_t)nbox * stride;
    int32_t *d_lo = (int32_t *)p, *d_hi = d_lo + (size_t)nbox * g->dim;
    p += round256(2 * bb);
    uint32_t *cnt = (uint32_t *)p, *flag = cnt + per, *pscan = flag + per, *cscan = pscan + per;
    p += round256(4 * per * 4);
    uint32_t *d_tot = (uint32_t *)p;
    p += round256((size_t)2 * nbox * 4);
    uint32_t *ws = (uint32_t *)p;
    const size_t ws_words = scan_ws_words(stride);

    SG_HIP(hipMemcpyAsync(d_lo, lo, bb, hipMemcpyHostToDevice, g->stream));
    SG_HIP(hipMemcpyAsync(d_hi, hi, bb, hipMemcpyHostToDevice, g->stream));
    // the scratch keeps the previous call's contents: the rows' trailing entries (which the
    // count pass does not write) must read 0
    SG_HIP(hipMemsetAsync(cnt, 0, per * 2 * sizeof(uint32_t), g->stream));
    if (n > 0) {
        Boxes B{nbox, d_lo, d_hi};
        hipLaunchKernelGGL(pack_count_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_keys, n, g->d_mask,
                           geom_of(g), B, cnt, flag, stride);
        SG_HIP(hipGetLastError());
    }
    for (int k = 0; k < nbox; ++k) {
        if ((rc = exclusive_scan_u32_ws(g->stream, cnt + k * stride, pscan + k * stride, stride, d_tot + 2 * k, ws,
                                        ws_words)))
            return rc;
        if ((rc = exclusive_scan_u32_ws(g->stream, flag + k * stride, cscan + k * stride, stride, d_tot + 2 * k + 1,
                                        ws, ws_words)))
            return rc;
    }
    uint32_t h_tot[2 * SG_PACK_NOALLOC_MAX_BOXES];
    SG_HIP(hipMemcpyAsync(h_tot, d_tot, (size_t)2 * nbox * 4, hipMemcpyDeviceToHost, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    int64_t n_pnt[SG_PACK_NOALLOC_MAX_BOXES], n_cnk[SG_PACK_NOALLOC_MAX_BOXES];
    const int64_t psz = point_bytes(g, nprop, props);
    int64_t tot = 0;
    for (int k = 0; k < nbox; ++k) {
        n_pnt[k] = h_tot[2 * k];
        n_cnk[k] = h_tot[2 * k + 1];
        tot += section_bytes(g, n_cnk[k], n_pnt[k], psz, index_bytes);
    }
    if (tot > cap) return fail(SG_E_RANGE, "pack buffer too small: " + std::to_string(tot) + " bytes needed");
    if ((rc = pack_sections(g, nbox, lo, nprop, props, index_bytes, (unsigned char *)dev_buf, cap, stride, cnt, pscan,
                            cscan, d_lo, d_hi, n_cnk, n_pnt)))
        return rc;
    SG_HIP(hipStreamSynchronize(g->stream));
    *written = tot;
    return SG_OK;
}

int sg_unpack(sg_grid *g, const void *buf, int64_t avail, int mem, const int32_t *dst_lo, int nprop,
              const int32_t *props, int index_bytes, int64_t *consumed)
{
    if (!g || !buf || !dst_lo || !consumed) return fail(SG_E_ARG, "null argument");
    int rc = check_pack_args(g, 0, nullptr, nullptr, nprop, props, index_bytes);
    if (rc) return rc;
    const int A = index_bytes;
    const int64_t hsz = 8 + 2 * g->dim * 4;
    if (avail < hsz) return fail(SG_E_RANGE, "packed section truncated");
    const hipMemcpyKind kind = (mem == SG_MEM_DEVICE) ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    std::vector<unsigned char> hdr(hsz);
    SG_HIP(hipMemcpyAsync(hdr.data(), buf, hsz, kind, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    uint64_t ncu;
    memcpy(&ncu, hdr.data(), 8);
    const int64_t nc = (int64_t)ncu;
    int32_t bstart[3] = {0, 0, 0}, psize[3] = {1, 1, 1};
    for (int d = 0; d < g->dim; ++d) {
        memcpy(&bstart[d], hdr.data() + 8 + 4 * d, 4);
        memcpy(&psize[d], hdr.data() + 8 + 4 * (g->dim + d), 4);
        if (psize[d] <= 0) return fail(SG_E_RANGE, "bad packed grid size");
    }
    if (nc < 0 || avail < hsz + A * nc + (nc + 1) * 4) return fail(SG_E_RANGE, "packed section truncated");
    uint32_t np32 = 0;
    SG_HIP(hipMemcpyAsync(&np32, (const unsigned char *)buf + hsz + A * nc + nc * 4, 4, kind, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    const int64_t np = np32;
    const int64_t psz = point_bytes(g, nprop, props);
    const int64_t sb = section_bytes(g, nc, np, psz, A);
    if (avail < sb) return fail(SG_E_RANGE, "packed section truncated");
    if (nc == 0 && np != 0) return fail(SG_E_RANGE, "malformed packed section: points without chunks");
    *consumed = sb;
    if (np == 0) return SG_OK;
    // the section on the device
    const unsigned char *s0 = (const unsigned char *)buf;
    if (mem != SG_MEM_DEVICE) {
        DevBuf &st = g->scratch[2];
        if ((rc = ensure(g->pool, st, (size_t)sb + 64))) return rc;
        SG_HIP(hipMemcpyAsync(st.p, buf, (size_t)sb, hipMemcpyHostToDevice, g->stream));
        s0 = (const unsigned char *)st.p;
    }
    int64_t at = 0;
    if ((rc = reserve_point_records(g, np, &at))) return rc;
    UnpackIn U;
    int64_t q = hsz;
    U.A = A;
    U.ids = s0 + q;
    q += A * nc;
    U.scan = (const uint32_t *)(s0 + q);
    q += align_up((nc + 1) * 4, A);
    for (int p = 0; p < kMaxProps; ++p) U.data[p] = nullptr;
    int64_t dq = q;
    for (int pi = 0; pi < nprop; ++pi) {
        U.data[props[pi]] = s0 + dq;
        dq += np * g->ts[props[pi]];
    }
    q += align_up(np * psz, A);
    U.offs = (const int16_t *)(s0 + q);
    for (int d = 0; d < 3; ++d) {
        U.pbsz[d] = (d < g->dim) ? (psize[d] + g->edge - 1) / g->edge : 1;
        U.shift[d] = (d < g->dim) ? dst_lo[d] - bstart[d] : 0;
    }
    // records go to the insert buffer at `at`; values of the properties not unpacked are
    // the background (the finalize keeps the old values there)
    PropSet R;
    R.np = g->np;
    for (int p = 0; p < g->np; ++p) {
        R.p[p] = (unsigned char *)g->d_ins_val[p] + at * g->ts[p];
        R.ts[p] = g->ts[p];
        R.bg[p] = 0;
        memcpy(&R.bg[p], g->bg[p], g->ts[p]);
    }
    if ((rc = ensure(g->pool, g->scratch[6], 64))) return rc;
    int *bad = (int *)g->scratch[6].p;
    SG_HIP(hipMemsetAsync(bad, 0, sizeof(int), g->stream));
    hipLaunchKernelGGL(unpack_records_kernel, dim3((unsigned)std::max<int64_t>(nc, 1)), dim3(kWG), 0, g->stream, U,
                       nc, np, geom_of(g), g->sz[0], g->sz[1], g->sz[2], R, g->d_ins_key + at, g->d_ins_mask + at, bad);
    SG_HIP(hipGetLastError());
    int hbad = 0;
    SG_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    if (hbad) return fail(SG_E_RANGE, "unpacked point outside the grid or malformed packed section");
    g->ins_n = at + np;
    return SG_OK;
}

int sg_remove_box(sg_grid *g, const int32_t *lo, const int32_t *hi)
{
    if (!g || !lo || !hi) return fail(SG_E_ARG, "null argument");
    for (int d = 0; d < g->dim; ++d) {
        g->rem_lo.push_back(lo[d]);
        g->rem_hi.push_back(hi[d]);
    }
    return SG_OK;
}

int sg_remove_points(sg_grid *g, int64_t n, const int64_t *keys, int mem)
{
    if (!g || (n > 0 && !keys)) return fail(SG_E_ARG, "null argument");
    if (n <= 0 || g->nchunks == 0) return SG_OK;
    const int64_t *dk = keys;
    if (mem != SG_MEM_DEVICE) {
        int rc = ensure(g->pool, g->scratch[9], (size_t)n * 8 + 64);
        if (rc) return rc;
        SG_HIP(hipMemcpyAsync(g->scratch[9].p, keys, (size_t)n * 8, hipMemcpyHostToDevice, g->stream));
        dk = (const int64_t *)g->scratch[9].p;
    }
    g->occ_tab_valid = false;
    g->ebits_valid = false;
    hipLaunchKernelGGL(remove_points_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_keys, g->nchunks,
                       g->d_mask, dk, n, g->csz, grid_props(g));
    SG_HIP(hipGetLastError());
    g->ep_valid = false;
    if (mem != SG_MEM_DEVICE) SG_HIP(hipStreamSynchronize(g->stream));  // the host keys are the caller's
    return SG_OK;
}

}  // extern "C"

namespace {

// clear the points of the boxes queued by sg_remove_box (mask 0, value background)
int remove_queued_boxes(sg_grid *g)
{
    const int nb = (int)(g->rem_lo.size() / std::max(1, g->dim));
    if (nb > 0 && g->nchunks > 0) {
        const size_t bb = (size_t)nb * g->dim * 4;
        int rc = ensure(g->pool, g->scratch[0], 2 * bb + 64);
        if (rc) return rc;
        int32_t *d_lo = (int32_t *)g->scratch[0].p, *d_hi = d_lo + (size_t)nb * g->dim;
        SG_HIP(hipMemcpyAsync(d_lo, g->rem_lo.data(), bb, hipMemcpyHostToDevice, g->stream));
        SG_HIP(hipMemcpyAsync(d_hi, g->rem_hi.data(), bb, hipMemcpyHostToDevice, g->stream));
        Boxes B{nb, d_lo, d_hi};
        g->occ_tab_valid = false;
        g->ebits_valid = false;
        hipLaunchKernelGGL(remove_boxes_kernel, dim3((unsigned)g->nchunks), dim3(kWG), 0, g->stream, g->d_keys,
                           g->nchunks, g->d_mask, geom_of(g), B, grid_props(g));
        SG_HIP(hipGetLastError());
        g->ep_valid = false;
        SG_HIP(hipStreamSynchronize(g->stream));  // the box vectors are cleared next
    }
    g->rem_lo.clear();
    g->rem_hi.clear();
    return SG_OK;
}

int check_props(sg_grid *g, int nprop, const int32_t *props)
{
    if (nprop > 0 && !props) return fail(SG_E_ARG, "null argument");
    for (int q = 0; q < nprop; ++q)
        if (props[q] < 0 || props[q] >= g->np) return fail(SG_E_ARG, "bad property");
    return SG_OK;
}

// unpacked properties: the unpacked value (left operand) wins over the old one; the
// others keep the old value (right operand), new points get the background
int flush_unpacked(sg_grid *g, int nprop, const int32_t *props)
{
    std::vector<int32_t> ops(g->np, SG_OP_RIGHT);
    for (int q = 0; q < nprop; ++q) ops[props[q]] = SG_OP_LEFT;
    return flush_impl(g, ops.data());
}

// phase 2 of removeCopyToFinalize: pack every queued section of g into one device buffer,
// then unpack each into its destination's insert buffer
int copy_sections_out(sg_grid *g, int nprop, const int32_t *props)
{
    if (g->copies.empty()) return SG_OK;
    const int n = (int)g->copies.size(), D = g->dim;
    for (const auto &c : g->copies) {
        sg_grid *d = c.dst;
        if (d->dim != D) return fail(SG_E_ARG, "copy_to: grids of different dimension");
        for (int q = 0; q < nprop; ++q)
            if (props[q] >= d->np || d->dt[props[q]] != g->dt[props[q]])
                return fail(SG_E_ARG, "copy_to: property types differ between the grids");
    }
    std::vector<int32_t> lo((size_t)n * D), hi((size_t)n * D);
    for (int k = 0; k < n; ++k)
        for (int a = 0; a < D; ++a) {
            lo[(size_t)k * D + a] = g->copies[k].lo[a];
            hi[(size_t)k * D + a] = g->copies[k].hi[a];
        }
    std::vector<int64_t> bytes(n);
    int64_t tot = 0;
    for (int k = 0; k < n; ++k) {
        int rc = sg_pack_size(g, 1, &lo[(size_t)k * D], &hi[(size_t)k * D], nprop, props, 8, &bytes[k]);
        if (rc) return rc;
        tot += bytes[k];
    }
    DevBuf &cb = g->scratch[3];
    int rc = ensure(g->pool, cb, (size_t)tot + 64);
    if (rc) return rc;
    rc = sg_pack(g, n, lo.data(), hi.data(), nprop, props, 8, cb.p, tot, SG_MEM_DEVICE);
    if (rc) return rc;
    SG_HIP(hipStreamSynchronize(g->stream));  // the destinations read it on their streams
    int64_t at = 0;
    for (int k = 0; k < n; ++k) {
        int64_t used = 0;
        rc = sg_unpack(g->copies[k].dst, (const unsigned char *)cb.p + at, tot - at, SG_MEM_DEVICE,
                       g->copies[k].dst_lo, nprop, props, 8, &used);
        if (rc) return rc;
        at += used;
    }
    g->copies.clear();
    return SG_OK;
}

}  // namespace

extern "C" {

int sg_remove_add_unpack_finalize(sg_grid *g, int nprop, const int32_t *props)
{
    if (!g) return fail(SG_E_ARG, "null argument");
    int rc = check_props(g, nprop, props);
    if (rc) return rc;
    if ((rc = remove_queued_boxes(g))) return rc;
    return flush_unpacked(g, nprop, props);
}

int sg_copy_to(sg_grid *dst, sg_grid *src, const int32_t *src_lo, const int32_t *src_hi, const int32_t *dst_lo)
{
    if (!dst || !src || !src_lo || !src_hi || !dst_lo) return fail(SG_E_ARG, "null argument");
    if (dst->dim != src->dim) return fail(SG_E_ARG, "copy_to: grids of different dimension");
    sg_grid::copy_section c{};
    c.dst = dst;
    for (int d = 0; d < 3; ++d) {
        c.lo[d] = d < src->dim ? src_lo[d] : 0;
        c.hi[d] = d < src->dim ? src_hi[d] : 0;
        c.dst_lo[d] = d < src->dim ? dst_lo[d] : 0;
    }
    src->copies.push_back(c);
    return SG_OK;
}

int sg_remove_copy_to_finalize(sg_grid *g, int phase, int nprop, const int32_t *props)
{
    if (!g) return fail(SG_E_ARG, "null argument");
    int rc = check_props(g, nprop, props);
    if (rc) return rc;
    switch (phase & 3) {
    case SG_COPY_PHASE1: return remove_queued_boxes(g);
    case SG_COPY_PHASE2: return copy_sections_out(g, nprop, props);
    case SG_COPY_PHASE3: return flush_unpacked(g, nprop, props);
    default: return fail(SG_E_ARG, "bad phase");
    }
}

int sg_remove_add_unpack_reset(sg_grid *g)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    g->rem_lo.clear();
    g->rem_hi.clear();
    g->copies.clear();
    g->ins_n = 0;
    g->cins_n = 0;
    g->pins_n = 0;
    return SG_OK;
}

}  // extern "C"
