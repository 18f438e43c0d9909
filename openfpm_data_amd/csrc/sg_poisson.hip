// sg_poisson.hip -- batched Poisson solves on 8^3-chunk sparse grids (gfx950).
//
// -lap(u) = f on the existing points of each grid, u = 0 outside (the background), with
// the 7-point operator A u = 6 u - sum of the 6 face neighbours.
//
// Precision form (SG_POISSON_MIXED): the state e and its right-hand side r are FP32 and
// swept with red-black SOR; FP64 is used only for the defect correction and the residual
// norms:   u' = u + e,   r = f - A u' (FP64),   ||r||, ||f|| (FP64),   e <- SOR sweeps of
// A e = (float) r from e = 0.  Plain form (SG_POISSON_PLAIN_FP64): red-black SOR sweeps of
// A u = f directly in FP64, residual checked between sweep blocks.
//
// Batching: every launch covers all live grids; workgroup waves take tiles of a (grid,
// chunk) work list (per-grid prefix sums of chunk counts, binary-searched per wave), so
// grids of uneven size share launches and small grids cost no launches of their own.
// The inner sweep sequence can be captured once per live set in a hipGraph and replayed.
// Grids join and leave the batch between outer iterations (continuous batching).
//
// SOR relaxation: omega is estimated per grid from the residual reduction rho of each
// outer iteration (the first runs with omega0): the red-black ordered 7-point matrix is
// consistently ordered, so the SOR eigenvalue lambda and the Jacobi spectral radius mu
// satisfy (lambda + w - 1)^2 = lambda w^2 mu^2 (Young); mu from (rho, w), then
// omega_opt = 2 / (1 + sqrt(1 - mu^2)).
#include "sg_internal.hpp"

#include <algorithm>
#include <cmath>

namespace sg {
namespace {

struct PGrid {
    const int32_t *nn;    // NNStar table [n][8]
    const uint8_t *mask;  // [n+1][512]
    void *x;              // swept unknown (e: f32 mixed, u: f64 plain)
    const void *rhs;      // r (f32) or f (f64)
    const double *u_in;   // defect: current u
    double *u_out;        // defect: u + e
    const double *f;      // defect / residual: f
    float *e;             // defect: correction (may be null: treated as 0)
    float *r32;           // defect: FP32 residual written for the inner solve
    double omega;
    int64_t n;            // chunks
    int32_t nsweeps;      // sweeps of this grid in the current inner block
};

template <typename T> struct R8 {
    T v[8];
};

template <typename T> __device__ inline void ld8(R8<T> &r, const T *p)
{
    if constexpr (sizeof(T) == 4) {
        const float4 a = ((const float4 *)p)[0], b = ((const float4 *)p)[1];
        r.v[0] = a.x; r.v[1] = a.y; r.v[2] = a.z; r.v[3] = a.w;
        r.v[4] = b.x; r.v[5] = b.y; r.v[6] = b.z; r.v[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double2 a = ((const double2 *)p)[i];
            r.v[2 * i] = a.x;
            r.v[2 * i + 1] = a.y;
        }
    }
}

template <typename T> __device__ inline void st8(T *p, const R8<T> &r)
{
    if constexpr (sizeof(T) == 4) {
        ((float4 *)p)[0] = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
        ((float4 *)p)[1] = make_float4(r.v[4], r.v[5], r.v[6], r.v[7]);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) ((double2 *)p)[i] = make_double2(r.v[2 * i], r.v[2 * i + 1]);
    }
}

__device__ inline void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ghost half row of the +-y / +-z faces for lane (see sg_stencil.hip): source row in the
// neighbour chunk and destination row of the 10x10-row tile
__device__ inline int g_src_row(int lane)
{
    const int face = lane >> 4, idx = (lane >> 1) & 7;
    const int ry = idx * 8 + ((face == 0) ? 7 : 0);
    const int rz = ((face == 2) ? 56 : 0) + idx;
    return (face < 2) ? ry : rz;
}
__device__ inline int g_tile_row(int lane)
{
    const int face = lane >> 4, idx = (lane >> 1) & 7;
    const int ty = (idx + 1) * 10 + ((face == 0) ? 0 : 9);
    const int tz = ((face == 2) ? 0 : 90) + idx + 1;
    return (face < 2) ? ty : tz;
}

// Stage a property's own rows + face ghost rows in the tile, return own row and x ghosts.
template <typename T>
__device__ inline void stage(const T *arr, int64_t c, const int64_t (&nbp)[6], T *tile, int lane, R8<T> &own, T &xm,
                             T &xp)
{
    ld8(own, arr + c * 512 + lane * 8);
    const int face = lane >> 4;
    const int64_t nbc = (face == 0) ? nbp[2] : (face == 1) ? nbp[3] : (face == 2) ? nbp[4] : nbp[5];
    const T *src = arr + nbc * 512 + g_src_row(lane) * 8 + (lane & 1) * 4;
    T *dst = tile + g_tile_row(lane) * 8 + (lane & 1) * 4;
    if constexpr (sizeof(T) == 4) {
        *(float4 *)dst = *(const float4 *)src;
    } else {
        ((double2 *)dst)[0] = ((const double2 *)src)[0];
        ((double2 *)dst)[1] = ((const double2 *)src)[1];
    }
    xm = arr[nbp[0] * 512 + lane * 8 + 7];
    xp = arr[nbp[1] * 512 + lane * 8 + 0];
    const int y = lane & 7, z = lane >> 3;
    st8(tile + ((z + 1) * 10 + (y + 1)) * 8, own);
}

template <typename T>
__device__ inline void neighbour_rows(const T *tile, int lane, R8<T> &ym, R8<T> &yp, R8<T> &zm, R8<T> &zp)
{
    const int y = lane & 7, z = lane >> 3;
    ld8(ym, tile + ((z + 1) * 10 + y) * 8);
    ld8(yp, tile + ((z + 1) * 10 + y + 2) * 8);
    ld8(zm, tile + (z * 10 + y + 1) * 8);
    ld8(zp, tile + ((z + 2) * 10 + y + 1) * 8);
}

// One colour of a red-black SOR sweep on one tile (chunk c of grid g): x += w ((rhs + sum
// nb) / 6 - x) on the points of colour `color` ((x + y + z) & 1; chunk origins are
// multiples of 8, so the local coordinates decide).  zero_x: x is taken as 0 everywhere
// (first half sweep of an inner solve): points of the other colour are written 0.  Whole
// rows are stored; values of the other colour are written back unchanged (no other wave
// changes them in this half sweep).
template <typename T>
__device__ inline void sweep_tile(const PGrid &g, int64_t c, T *tile, int lane, int color, int zero_x)
{
    T *x = (T *)g.x;
    const T *rhs = (const T *)g.rhs;
    const uint64_t mrow = *(const uint64_t *)(g.mask + c * 512 + lane * 8);
    R8<T> b;
    ld8(b, rhs + c * 512 + lane * 8);
    R8<T> own, ym, yp, zm, zp;
    T xm = 0, xp = 0;
    if (zero_x) {
#pragma unroll
        for (int i = 0; i < 8; ++i) own.v[i] = ym.v[i] = yp.v[i] = zm.v[i] = zp.v[i] = (T)0;
    } else {
        int64_t nbp[6];
        const int32_t *row = g.nn + c * 8;
#pragma unroll
        for (int k = 0; k < 6; ++k) nbp[k] = row[k];
        stage<T>(x, c, nbp, tile, lane, own, xm, xp);
        wave_sync_lds();
        neighbour_rows<T>(tile, lane, ym, yp, zm, zp);
    }
    const T wgt = (T)g.omega;
    const int par0 = ((lane & 7) + (lane >> 3)) & 1;  // parity of (y + z)
    R8<T> out;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint8_t cm = (uint8_t)(mrow >> (8 * i));
        const bool active = (cm & SG_EXIST) && !(cm & SG_PADDING);
        const bool mine = ((i + par0) & 1) == color;
        const T cur = own.v[i];
        const T l = (i == 0) ? xm : own.v[i - 1];
        const T r = (i == 7) ? xp : own.v[i + 1];
        T nbs = (T)0;
        nbs = nbs + (l + r);
        nbs = nbs + (ym.v[i] + yp.v[i]);
        nbs = nbs + (zm.v[i] + zp.v[i]);
        const T gs = (b.v[i] + nbs) / (T)6;
        const T nv = cur + wgt * (gs - cur);
        out.v[i] = (active && mine) ? nv : cur;
    }
    // rows with a point are written whole (both colours: the other colour keeps its
    // value): staged in the LDS tile, then written back as lane-contiguous 16-byte stores
    wave_sync_lds();  // every lane has read its neighbour rows
    st8(tile + lane * 8, out);
    wave_sync_lds();
    store_chunk_contig(x + c * 512, tile, mrow != 0 ? 1u : 0u, lane);
    // the next tile of this wave restages the LDS tile: finish reading it first
    wave_sync_lds();
}

// One colour of a red-black SOR sweep over the live grids.  Grids are ordered by
// decreasing sweep count, so the tiles of the grids still sweeping at sweep `sweep` are
// [0, act[sweep]); tile t belongs to grid tile_grid[t].  Wave v takes tiles v, v + V, ...
// (V = waves in the launch), so a launch of any size covers the active tiles.
template <typename T, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void rbsor_batch_kernel(const PGrid *__restrict__ grids,
                                                                const int64_t *__restrict__ prefix,
                                                                const int32_t *__restrict__ tile_grid,
                                                                const int64_t *__restrict__ act, int color,
                                                                int zero_x, int sweep)
{
    __shared__ __attribute__((aligned(16))) T tiles[WAVES][100 * 8];
    // uniform wave index: the tile's grid descriptor and neighbour-table row come by
    // scalar loads
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t nact = act[sweep];
    const int64_t nwaves = (int64_t)gridDim.x * WAVES;
    for (int64_t t = (int64_t)blockIdx.x * WAVES + w; t < nact; t += nwaves) {
        const int gi = tile_grid[t];
        sweep_tile<T>(grids[gi], t - prefix[gi], tiles[w], lane, color, zero_x);
    }
}

// tile -> grid map of a work list
__global__ __launch_bounds__(256) void tile_grid_kernel(const int64_t *__restrict__ prefix, int32_t *__restrict__ tg)
{
    const int gi = blockIdx.x;
    for (int64_t t = prefix[gi] + threadIdx.x; t < prefix[gi + 1]; t += 256) tg[t] = gi;
}

// Defect correction + residual, FP64: u_out = u_in + e, r = f - A u_out, r32 = (float) r;
// per-tile partial sums of r^2 and f^2 over the existing points (reduced per grid in a
// fixed order afterwards, so the norms are deterministic).  e == nullptr: e = 0,
// u_out == nullptr: residual only.
__global__ __launch_bounds__(256) void defect_batch_kernel(const PGrid *__restrict__ grids,
                                                           const int64_t *__restrict__ prefix,
                                                           const int32_t *__restrict__ tile_grid, int64_t ntiles,
                                                           double *__restrict__ partial)
{
    __shared__ __attribute__((aligned(16))) double tiles[4][100 * 8];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + w;
    if (t >= ntiles) return;
    const int gi = tile_grid[t];
    const PGrid g = grids[gi];
    const int64_t c = t - prefix[gi];
    int64_t nbp[6];
    const int32_t *row = g.nn + c * 8;
#pragma unroll
    for (int k = 0; k < 6; ++k) nbp[k] = row[k];
    const uint64_t mrow = *(const uint64_t *)(g.mask + c * 512 + lane * 8);
    R8<double> u, fr;
    double uxm, uxp;
    double *tile = tiles[w];
    // stage u (+ e) : u' = u + e for own rows and ghosts
    {
        ld8(u, g.u_in + c * 512 + lane * 8);
        const int face = lane >> 4;
        const int64_t nbc = (face == 0) ? nbp[2] : (face == 1) ? nbp[3] : (face == 2) ? nbp[4] : nbp[5];
        const int64_t goff = nbc * 512 + g_src_row(lane) * 8 + (lane & 1) * 4;
        double gh[4];
        const double2 a = ((const double2 *)(g.u_in + goff))[0], bb = ((const double2 *)(g.u_in + goff))[1];
        gh[0] = a.x; gh[1] = a.y; gh[2] = bb.x; gh[3] = bb.y;
        uxm = g.u_in[nbp[0] * 512 + lane * 8 + 7];
        uxp = g.u_in[nbp[1] * 512 + lane * 8 + 0];
        if (g.e) {
            R8<float> e;
            ld8(e, g.e + c * 512 + lane * 8);
#pragma unroll
            for (int i = 0; i < 8; ++i) u.v[i] = u.v[i] + (double)e.v[i];
            const float4 eg = *(const float4 *)(g.e + goff);
            gh[0] = gh[0] + (double)eg.x; gh[1] = gh[1] + (double)eg.y;
            gh[2] = gh[2] + (double)eg.z; gh[3] = gh[3] + (double)eg.w;
            uxm = uxm + (double)g.e[nbp[0] * 512 + lane * 8 + 7];
            uxp = uxp + (double)g.e[nbp[1] * 512 + lane * 8 + 0];
        }
        double *dst = tile + g_tile_row(lane) * 8 + (lane & 1) * 4;
        ((double2 *)dst)[0] = make_double2(gh[0], gh[1]);
        ((double2 *)dst)[1] = make_double2(gh[2], gh[3]);
        const int y = lane & 7, z = lane >> 3;
        st8(tile + ((z + 1) * 10 + (y + 1)) * 8, u);
    }
    ld8(fr, g.f + c * 512 + lane * 8);
    wave_sync_lds();
    R8<double> ym, yp, zm, zp;
    neighbour_rows<double>(tile, lane, ym, yp, zm, zp);
    double s_r = 0.0, s_f = 0.0;
    R8<float> r32;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint8_t cm = (uint8_t)(mrow >> (8 * i));
        const bool active = (cm & SG_EXIST) && !(cm & SG_PADDING);
        const double l = (i == 0) ? uxm : u.v[i - 1];
        const double rr = (i == 7) ? uxp : u.v[i + 1];
        double acc = 6.0 * u.v[i];
        acc = acc - (l + rr);
        acc = acc - (ym.v[i] + yp.v[i]);
        acc = acc - (zm.v[i] + zp.v[i]);
        const double res = fr.v[i] - acc;
        r32.v[i] = active ? (float)res : 0.0f;
        if (active) {
            s_r += res * res;
            s_f += fr.v[i] * fr.v[i];
        }
    }
    if (mrow != 0) {
        if (g.u_out) st8(g.u_out + c * 512 + lane * 8, u);
        if (g.r32) st8(g.r32 + c * 512 + lane * 8, r32);
    }
    // wave reduction (fixed shape)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s_r += __shfl_down(s_r, o, 64);
        s_f += __shfl_down(s_f, o, 64);
    }
    if (lane == 0) {
        partial[2 * t] = s_r;
        partial[2 * t + 1] = s_f;
    }
}

// per grid: sum of its tiles' partials in a fixed order (strided lanes + fixed tree)
__global__ __launch_bounds__(256) void norm_reduce_kernel(const double *__restrict__ partial,
                                                          const int64_t *__restrict__ prefix, int ng,
                                                          double *__restrict__ norms)
{
    __shared__ double sr[256], sf[256];
    const int gi = blockIdx.x;
    if (gi >= ng) return;
    double a = 0.0, b = 0.0;
    for (int64_t t = prefix[gi] + threadIdx.x; t < prefix[gi + 1]; t += 256) {
        a += partial[2 * t];
        b += partial[2 * t + 1];
    }
    sr[threadIdx.x] = a;
    sf[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] += sr[threadIdx.x + s];
            sf[threadIdx.x] += sf[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        norms[2 * gi] = sr[0];
        norms[2 * gi + 1] = sf[0];
    }
}

}  // namespace
}  // namespace sg

struct sg_batch {
    sg_poisson_params p;
    hipStream_t stream = nullptr;
    bool own_stream = false;  // created here when the caller passed none (graphs need one)
    struct Entry {
        sg_grid *g;
        int64_t tag;
        int cur_u;  // property holding the current u (mixed: u_prop or u2_prop)
        int outer = 0;
        int sweeps = 0;
        double omega = 0.0;
        double norm_f = -1.0;
        double last_rel = -1.0;
        double rel = 1.0;
        int inner = 0;  // inner sweeps for the next outer iteration
        double tol = 0.0;
    };
    std::vector<Entry> live;
    std::vector<std::pair<int64_t, sg_poisson_result>> done;
    // device work list (rebuilt when the live set changes)
    sg::DevicePool pool;
    sg::DevBuf d_grids, d_prefix, d_partial, d_norms, d_tg, d_act;
    bool dirty = true;
    std::vector<int64_t> h_prefix, h_act;
    // launch shape of the sweep kernel: waves per workgroup (1, 2, 4, 8) and a cap on the
    // workgroups of one launch (0: one tile per wave); set by sg_batch_set_tuning
    int waves = 4;
    int64_t max_wgs = 0;
    // graph of the inner sweep sequence for the current live set
    hipGraphExec_t graph = nullptr;
    int graph_sweeps = -1;
    uint64_t graph_version = 0, version = 0;
    ~sg_batch()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        pool.release();
    }
};

namespace {

using namespace sg;

// Device work list of the live grids in the order `order` (indices into live): grid
// descriptors, prefix sums of chunk counts, the tile -> grid map and, for the sweeps, the
// active tile count per sweep index (order must then be by decreasing inner count).
int upload_worklist(sg_batch *b, bool defect_view, const std::vector<int> &order)
{
    const int ng = (int)order.size();
    std::vector<PGrid> gd(ng);
    b->h_prefix.assign(ng + 1, 0);
    const bool mixed = b->p.mode == SG_POISSON_MIXED;
    int max_inner = 0;
    for (int i = 0; i < ng; ++i) {
        sg_batch::Entry &en = b->live[order[i]];
        sg_grid *g = en.g;
        PGrid &d = gd[i];
        d.nn = g->d_nn[SG_NN_STAR];
        d.mask = g->d_mask;
        d.n = g->nchunks;
        d.omega = en.omega;
        d.nsweeps = en.inner;
        d.f = (const double *)g->d_prop[b->p.f_prop];
        if (mixed) {
            d.x = g->d_prop[b->p.e_prop];
            d.rhs = g->d_prop[b->p.r_prop];
            d.u_in = (const double *)g->d_prop[en.cur_u];
            const int other = (en.cur_u == b->p.u_prop) ? b->p.u2_prop : b->p.u_prop;
            d.u_out = defect_view ? (double *)g->d_prop[other] : nullptr;
            // the first defect of a grid takes u as given (e may hold an earlier solve's state)
            d.e = (defect_view && en.outer == 0) ? nullptr : (float *)g->d_prop[b->p.e_prop];
            d.r32 = (float *)g->d_prop[b->p.r_prop];
        } else {
            d.x = g->d_prop[b->p.u_prop];
            d.rhs = g->d_prop[b->p.f_prop];
            d.u_in = (const double *)g->d_prop[b->p.u_prop];
            d.u_out = nullptr;
            d.e = nullptr;
            d.r32 = nullptr;
        }
        b->h_prefix[i + 1] = b->h_prefix[i] + g->nchunks;
        max_inner = std::max(max_inner, en.inner);
    }
    // act[s] = tiles of the grids with more than s sweeps (a prefix of the ordered list)
    b->h_act.assign(std::max(max_inner, 1), 0);
    if (!defect_view) {
        int k = ng;
        for (int sw = 0; sw < max_inner; ++sw) {
            while (k > 0 && gd[k - 1].nsweeps <= sw) --k;
            b->h_act[sw] = b->h_prefix[k];
        }
    }
    const int64_t nt = b->h_prefix[ng];
    int rc;
    if ((rc = ensure(b->pool, b->d_grids, sizeof(PGrid) * std::max(ng, 1))) ||
        (rc = ensure(b->pool, b->d_prefix, sizeof(int64_t) * (ng + 1))) ||
        (rc = ensure(b->pool, b->d_partial, sizeof(double) * 2 * std::max<int64_t>(nt, 1))) ||
        (rc = ensure(b->pool, b->d_norms, sizeof(double) * 2 * std::max(ng, 1))) ||
        (rc = ensure(b->pool, b->d_tg, sizeof(int32_t) * std::max<int64_t>(nt, 1))) ||
        (rc = ensure(b->pool, b->d_act, sizeof(int64_t) * b->h_act.size())))
        return rc;
    SG_HIP(hipMemcpyAsync(b->d_grids.p, gd.data(), sizeof(PGrid) * ng, hipMemcpyHostToDevice, b->stream));
    SG_HIP(hipMemcpyAsync(b->d_prefix.p, b->h_prefix.data(), sizeof(int64_t) * (ng + 1), hipMemcpyHostToDevice,
                          b->stream));
    SG_HIP(hipMemcpyAsync(b->d_act.p, b->h_act.data(), sizeof(int64_t) * b->h_act.size(), hipMemcpyHostToDevice,
                          b->stream));
    if (ng > 0) {
        hipLaunchKernelGGL(tile_grid_kernel, dim3((unsigned)ng), dim3(256), 0, b->stream,
                           (const int64_t *)b->d_prefix.p, (int32_t *)b->d_tg.p);
        SG_HIP(hipGetLastError());
    }
    // host vectors are reused on the next call: wait for the copies
    SG_HIP(hipStreamSynchronize(b->stream));
    return SG_OK;
}

template <typename T, int W> int launch_sweep_kernel(sg_batch *b, unsigned nblk, int color, int zx, int s)
{
    hipLaunchKernelGGL((rbsor_batch_kernel<T, W>), dim3(nblk), dim3(64 * W), 0, b->stream,
                       (const PGrid *)b->d_grids.p, (const int64_t *)b->d_prefix.p, (const int32_t *)b->d_tg.p,
                       (const int64_t *)b->d_act.p, color, zx, s);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

template <typename T> int launch_sweep(sg_batch *b, unsigned nblk, int color, int zx, int s)
{
    switch (b->waves) {
    case 1: return launch_sweep_kernel<T, 1>(b, nblk, color, zx, s);
    case 2: return launch_sweep_kernel<T, 2>(b, nblk, color, zx, s);
    case 8: return launch_sweep_kernel<T, 8>(b, nblk, color, zx, s);
    default: return launch_sweep_kernel<T, 4>(b, nblk, color, zx, s);
    }
}

int launch_sweeps(sg_batch *b, int sweeps, bool first_zero)
{
    const int64_t nt = b->h_act.empty() ? 0 : b->h_act[0];
    if (nt == 0) return SG_OK;
    int64_t nblk = (nt + b->waves - 1) / b->waves;
    if (b->max_wgs > 0) nblk = std::min(nblk, b->max_wgs);
    const bool mixed = b->p.mode == SG_POISSON_MIXED;
    for (int s = 0; s < sweeps; ++s) {
        for (int color = 0; color < 2; ++color) {
            const int zx = (first_zero && s == 0 && color == 0) ? 1 : 0;
            int rc = mixed ? launch_sweep<float>(b, (unsigned)nblk, color, zx, s)
                           : launch_sweep<double>(b, (unsigned)nblk, color, zx, s);
            if (rc) return rc;
        }
    }
    return SG_OK;
}

// run `sweeps` inner sweeps for the whole live set, through a cached hipGraph when asked
int run_sweeps(sg_batch *b, int sweeps, bool first_zero)
{
    if (!b->p.use_graph) return launch_sweeps(b, sweeps, first_zero);
    if (!b->graph || b->graph_sweeps != sweeps || b->graph_version != b->version) {
        if (b->graph) {
            SG_HIP(hipGraphExecDestroy(b->graph));
            b->graph = nullptr;
        }
        hipGraph_t gr;
        SG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_sweeps(b, sweeps, first_zero);
        hipError_t e2 = hipStreamEndCapture(b->stream, &gr);
        if (rc) return rc;
        if (e2 != hipSuccess) return fail(SG_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
        SG_HIP(hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0));
        SG_HIP(hipGraphDestroy(gr));
        b->graph_sweeps = sweeps;
        b->graph_version = b->version;
    }
    SG_HIP(hipGraphLaunch(b->graph, b->stream));
    return SG_OK;
}

// defect (mixed) or residual (plain) over the live set; per-grid norms to the host
int defect_and_norms(sg_batch *b, std::vector<double> &norms, bool defect_view)
{
    const int ng = (int)b->live.size();
    std::vector<int> order(ng);
    for (int i = 0; i < ng; ++i) order[i] = i;
    int rc = upload_worklist(b, defect_view, order);
    if (rc) return rc;
    const int64_t nt = b->h_prefix[ng];
    norms.assign(2 * ng, 0.0);
    if (nt > 0) {
        hipLaunchKernelGGL(defect_batch_kernel, dim3((unsigned)((nt + 3) / 4)), dim3(256), 0, b->stream,
                           (const PGrid *)b->d_grids.p, (const int64_t *)b->d_prefix.p, (const int32_t *)b->d_tg.p,
                           nt, (double *)b->d_partial.p);
        SG_HIP(hipGetLastError());
        hipLaunchKernelGGL(norm_reduce_kernel, dim3((unsigned)ng), dim3(256), 0, b->stream,
                           (const double *)b->d_partial.p, (const int64_t *)b->d_prefix.p, ng, (double *)b->d_norms.p);
        SG_HIP(hipGetLastError());
        SG_HIP(hipMemcpyAsync(norms.data(), b->d_norms.p, sizeof(double) * 2 * ng, hipMemcpyDeviceToHost, b->stream));
        SG_HIP(hipStreamSynchronize(b->stream));
    }
    return SG_OK;
}

// inner sweeps needed for a residual reduction `target` with SOR rate rho
int sweeps_for(double rho, double target, int lo, int hi)
{
    if (!(rho > 0.0) || rho >= 1.0) return hi;
    int k = (int)std::ceil(std::log(target) / std::log(rho));
    return std::max(lo, std::min(hi, k));
}

// Young's relation for consistently ordered matrices: given the SOR rate rho observed with
// relaxation w0, the Jacobi spectral radius mu and the optimal relaxation
double omega_from_rate(double rho, double w0)
{
    if (!(rho > 0.0) || rho >= 1.0) return w0;
    const double lam = rho;
    double mu = (lam + w0 - 1.0) / (w0 * std::sqrt(lam));
    mu = std::min(std::max(mu, 0.0), 0.9999);
    return 2.0 / (1.0 + std::sqrt(1.0 - mu * mu));
}

void retire(sg_batch *b, size_t i, int converged)
{
    sg_batch::Entry &en = b->live[i];
    sg_poisson_result r;
    r.converged = converged;
    r.outer = en.outer;
    r.sweeps = en.sweeps;
    r.rel_residual = en.rel;
    r.omega = en.omega;
    r.solution_prop = en.cur_u;
    b->done.emplace_back(en.tag, r);
}

}  // namespace

extern "C" {

int sg_batch_create(const sg_poisson_params *p, void *stream, sg_batch **out)
{
    if (!p || !out) return fail(SG_E_ARG, "null argument");
    if (p->mode != SG_POISSON_MIXED && p->mode != SG_POISSON_PLAIN_FP64) return fail(SG_E_ARG, "unknown mode");
    if (!(p->tol > 0.0)) return fail(SG_E_ARG, "tolerance must be positive");
    sg_batch *b = new sg_batch();
    b->p = *p;
    if (b->p.omega0 <= 0.0) b->p.omega0 = 1.5;
    if (b->p.first_sweeps <= 0) b->p.first_sweeps = 24;
    if (b->p.max_sweeps <= 0) b->p.max_sweeps = 20000;
    if (b->p.max_outer <= 0) b->p.max_outer = 100;
    if (b->p.inner_target <= 0.0) b->p.inner_target = (p->mode == SG_POISSON_MIXED) ? 1e-5 : 1e-3;
    b->stream = (hipStream_t)stream;
    if (!b->stream) {
        // stream capture is not possible on the legacy null stream; grids queued on it are
        // synchronized in sg_batch_add before the batch stream uses them
        hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete b;
            return fail(SG_E_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
        }
        b->own_stream = true;
    }
    *out = b;
    return SG_OK;
}

int sg_batch_destroy(sg_batch *b)
{
    if (!b) return SG_OK;
    (void)hipStreamSynchronize(b->stream);
    hipStream_t own = b->own_stream ? b->stream : nullptr;
    delete b;
    if (own) (void)hipStreamDestroy(own);
    return SG_OK;
}

int sg_batch_add(sg_batch *b, sg_grid *g, int64_t tag, double tol)
{
    if (!b || !g) return fail(SG_E_ARG, "null argument");
    if (g->dim != 3 || g->edge != 8) return fail(SG_E_ARG, "batched Poisson solves need 3D grids with 8^3 chunks");
    const sg_poisson_params &p = b->p;
    const bool mixed = p.mode == SG_POISSON_MIXED;
    auto need = [&](int prop, int dt) { return prop >= 0 && prop < g->np && g->dt[prop] == dt; };
    if (!need(p.u_prop, SG_F64) || !need(p.f_prop, SG_F64) ||
        (mixed && (!need(p.u2_prop, SG_F64) || !need(p.e_prop, SG_F32) || !need(p.r_prop, SG_F32))))
        return fail(SG_E_ARG, "grid properties do not match the solver layout (u, u2, f: f64; e, r: f32)");
    if (!g->nn_valid[SG_NN_STAR]) {
        int rc = find_neighbours_impl(g, SG_NN_STAR);
        if (rc) return rc;
    }
    // the grid's own queued work must be done before the batch stream uses it
    SG_HIP(hipStreamSynchronize(g->stream));
    sg_batch::Entry en;
    en.g = g;
    en.tag = tag;
    en.cur_u = p.u_prop;
    en.omega = (p.omega > 0.0) ? p.omega : p.omega0;
    en.inner = p.first_sweeps;
    en.tol = (tol > 0.0) ? tol : p.tol;
    b->live.push_back(en);
    b->dirty = true;
    ++b->version;
    return SG_OK;
}

int sg_batch_live(sg_batch *b) { return b ? (int)b->live.size() : 0; }

int sg_batch_set_tuning(sg_batch *b, int key, int64_t value)
{
    if (!b) return fail(SG_E_ARG, "null batch");
    switch (key) {
    case SG_TUNE_BATCH_WAVES:
        if (value != 1 && value != 2 && value != 4 && value != 8) return fail(SG_E_ARG, "waves must be 1, 2, 4 or 8");
        b->waves = (int)value;
        break;
    case SG_TUNE_BATCH_MAX_WGS:
        if (value < 0) return fail(SG_E_ARG, "workgroup cap must be >= 0");
        b->max_wgs = value;
        break;
    default: return fail(SG_E_ARG, "unknown tuning key");
    }
    ++b->version;  // the captured sweep graph has the old launch shape
    return SG_OK;
}

// One outer iteration for every live grid: residual (+ defect correction in the mixed
// form), convergence check and retirement, then the inner sweeps of the survivors.
int sg_batch_step(sg_batch *b)
{
    if (!b) return fail(SG_E_ARG, "null batch");
    if (b->live.empty()) return SG_OK;
    const bool mixed = b->p.mode == SG_POISSON_MIXED;
    std::vector<double> norms;
    int rc = defect_and_norms(b, norms, true);
    if (rc) return rc;
    // convergence / retirement
    std::vector<sg_batch::Entry> keep;
    for (size_t i = 0; i < b->live.size(); ++i) {
        sg_batch::Entry &en = b->live[i];
        // the mixed defect pass wrote u + e into the other u buffer
        if (mixed) en.cur_u = (en.cur_u == b->p.u_prop) ? b->p.u2_prop : b->p.u_prop;
        if (en.norm_f < 0.0) en.norm_f = std::sqrt(norms[2 * i + 1]);
        const double rn = std::sqrt(norms[2 * i]);
        en.last_rel = en.rel;
        en.rel = (en.norm_f > 0.0) ? rn / en.norm_f : 0.0;
        if (en.rel <= en.tol) {
            retire(b, i, 1);
            continue;
        }
        if (en.outer >= b->p.max_outer || en.sweeps >= b->p.max_sweeps) {
            retire(b, i, 0);
            continue;
        }
        // after each block: omega from the SOR rate observed with the block's omega.  The
        // first block starts from the rough residual f, whose fast modes decay early, so
        // its rate underestimates the asymptotic one and omega comes out low; later blocks
        // see the slow modes and refine it.  Below the optimum the dominant SOR eigenvalue
        // is real and Young's relation applies; at or above it the estimate returns the
        // block's own omega, so omega never decreases.
        if (en.outer >= 1 && b->p.omega <= 0.0 && en.last_rel > 0.0 && en.inner > 0) {
            const double rho = std::pow(std::max(en.rel / en.last_rel, 1e-300), 1.0 / en.inner);
            en.omega = std::max(en.omega, omega_from_rate(rho, en.omega));
        }
        // next block: sweeps for the wanted reduction at the asymptotic rate omega - 1 of
        // optimal SOR; never more than the remaining reduction to the tolerance asks
        if (en.outer == 0) {
            en.inner = b->p.first_sweeps;
        } else {
            const double target = std::max(b->p.inner_target, 0.5 * en.tol / en.rel);
            en.inner = sweeps_for(std::max(en.omega - 1.0, 1e-3), target, 4, 400);
        }
        keep.push_back(en);
    }
    const bool changed = keep.size() != b->live.size();
    b->live.swap(keep);
    if (changed || b->dirty) ++b->version;
    b->dirty = false;
    if (b->live.empty()) return SG_OK;
    // one launch sequence of max(inner) sweeps covers the live set; each grid stops after
    // its own count (its tiles leave the active prefix), so a grid's iterates do not
    // depend on its batch mates
    int sweeps = 0;
    for (auto &en : b->live) sweeps = std::max(sweeps, en.inner);
    // sweep work list: grids by decreasing inner count (ties in live order)
    std::vector<int> order(b->live.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->live[x].inner > b->live[y].inner; });
    rc = upload_worklist(b, false, order);
    if (rc) return rc;
    rc = run_sweeps(b, sweeps, mixed);
    if (rc) return rc;
    for (auto &en : b->live) {
        en.sweeps += en.inner;
        en.outer += 1;
    }
    return SG_OK;
}

int sg_batch_poll(sg_batch *b, int64_t *tags, sg_poisson_result *results, int max, int *n)
{
    if (!b || !n) return fail(SG_E_ARG, "null argument");
    int k = 0;
    while (k < max && !b->done.empty()) {
        tags[k] = b->done.front().first;
        results[k] = b->done.front().second;
        b->done.erase(b->done.begin());
        ++k;
    }
    *n = k;
    return SG_OK;
}

int sg_poisson_solve(sg_grid *const *grids, int n, const sg_poisson_params *p, sg_poisson_result *results)
{
    if (!grids || n < 0 || !p || !results) return fail(SG_E_ARG, "null argument");
    if (n == 0) return SG_OK;
    sg_batch *b = nullptr;
    int rc = sg_batch_create(p, grids[0]->stream, &b);
    if (rc) return rc;
    for (int i = 0; i < n && !rc; ++i) rc = sg_batch_add(b, grids[i], i, 0.0);
    while (!rc && !b->live.empty()) rc = sg_batch_step(b);
    if (!rc) {
        for (auto &d : b->done) results[d.first] = d.second;
    }
    sg_batch_destroy(b);
    return rc;
}

}  // extern "C"
