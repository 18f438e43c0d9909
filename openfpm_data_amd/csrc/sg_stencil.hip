// sg_stencil.hip -- neighbour tables, tagBoundaries, stencils and point gets for gfx950.
//
// Generic path (any dim 2/3, any chunk edge): one workgroup per chunk stages the
// chunk and its one-cell ghost layer in LDS (the reference's enlarged block,
// SparseGridGpu_kernels.cuh:258-276) and evaluates the stencil per point.
//
// Fast path (3D, 8^3 chunks, star stencils, f32/f64): one wave64 per chunk, lane l
// owns the x-row (y = l & 7, z = l >> 3) of 8 points in registers:
//   - own rows: two (f32) / four (f64) 16-byte loads per lane, the wave reads the
//     chunk's 2 KB / 4 KB contiguously (chunk-major SoA);
//   - +-y / +-z ghost faces: 32 rows of 8 values, one 16-byte load per lane per
//     half row, staged with the own rows in a 10x10-row LDS tile;
//   - +-x ghost faces: one scalar per lane from the x-neighbour chunks, kept in
//     registers (x-1 / x+1 of the row ends);
//   - neighbour chunk positions come from the findNeighbours table (scalar loads,
//     uniform per wave); a missing neighbour is the background chunk, whose values
//     are the background, so no mask of a neighbour is read (see the invariant in
//     sg_internal.hpp).
// The mask row (8 bytes) decides which points are written (EXIST && !PADDING, or
// mask != 0), as applyStencilInPlace does (SparseGridGpu.hpp:1005-1035).
#include "sg_internal.hpp"

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
template <typename T> __device__ inline T dev_mul(T a, T b)
{
    if constexpr (std::is_integral<T>::value) {
        typedef typename std::conditional<(sizeof(T) < 4), uint32_t, typename std::make_unsigned<T>::type>::type U;
        return (T)(U)((U)a * (U)b);
    } else {
        return a * b;
    }
}

struct GeomArgs {
    int dim, edge, csz;
    int64_t bsz0, bsz1, bsz2;
    int zorder;  // chunk ids are Morton codes (grid_zmb)
};

__device__ inline ChunkLin lin_of(const GeomArgs &ga)
{
    ChunkLin L;
    L.dim = ga.dim;
    L.zorder = ga.zorder;
    L.bsz[0] = ga.bsz0;
    L.bsz[1] = ga.bsz1;
    L.bsz[2] = ga.bsz2;
    return L;
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

// get_sparse: position of a chunk key or the background position n
// (map_vector_sparse_cuda_ker.cuh:75-92)
__device__ inline int64_t find_chunk(const int64_t *keys, int64_t n, int64_t key)
{
    int64_t lb = lower_bound_dev(keys, n, key);
    return (lb < n && keys[lb] == key) ? lb : n;
}

// findNeighbours (SparseGridGpu_kernels.cuh:1009-1035) with NNStar::getNNpos
// (SparseGridGpu_ker_util.hpp:91-112) or NNFull::getNNpos (SparseGridGpu.hpp:266-287).
// BlockLinId (grid_smb.hpp:258-276) returns -1 for a coordinate >= blockSz and does not
// check for -1: NNStar guards dPos < 0 itself, NNFull does not, so at the low x / y
// faces its linear id aliases another chunk; this table reproduces both.
// Position of chunk id `key` near chunk c (id kc), or n: in a sorted list of distinct ids
// the chunk with id kc + d is at most d positions after c (before c for d < 0), so the
// binary search runs in that window only (a few steps for the x and y neighbours).
__device__ inline int64_t find_chunk_near(const int64_t *keys, int64_t n, int64_t c, int64_t kc, int64_t key)
{
    int64_t lo, hi;
    if (key > kc) {
        const uint64_t d = (uint64_t)key - (uint64_t)kc;
        lo = c + 1;
        hi = (d >= (uint64_t)(n - c)) ? n : c + 1 + (int64_t)d;
    } else if (key < kc) {
        const uint64_t d = (uint64_t)kc - (uint64_t)key;
        hi = c;
        lo = (d >= (uint64_t)c) ? 0 : c - (int64_t)d;
    } else {
        return c;
    }
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && keys[lo] == key) ? lo : n;
}

// position of every chunk at its id (ids below the table size; the rest of the table -1)
__global__ void chunk_table_kernel(const int64_t *keys, int64_t n, int32_t *tab)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tab[keys[i]] = (int32_t)i;
}

// One thread per (chunk, table entry).  DIRECT: the neighbour's position comes from the
// chunk table (`tab`, ids [0, idmax)); otherwise from a binary search of the sorted ids
// confined to the positions the id distance allows.
template <bool DIRECT>
__global__ void find_neighbours_kernel(const int64_t *keys, int64_t n, GeomArgs ga, int type, int nnn,
                                       const int32_t *__restrict__ tab, int64_t idmax, int32_t *nn)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nnn) return;
    const int64_t c = t / nnn;
    const int k = (int)(t % nnn);
    const int dim = ga.dim;
    int64_t bs[3] = {ga.bsz0, ga.bsz1, ga.bsz2};
    int64_t bc[3] = {0, 0, 0};
    const ChunkLin L = lin_of(ga);
    chunk_coords(L, keys[c], bc);
    auto block_lin_id = [&](const int64_t *q) -> int64_t {
        // grid_zmb::BlockLinId is lin_zid without range checks: coordinates outside the
        // domain give ids of chunks that cannot exist (a -1 spreads to a huge code)
        if (L.zorder) return chunk_id(L, q);
        int64_t v = q[dim - 1];
        if (v >= bs[dim - 1]) return -1;
        for (int d = dim - 2; d >= 0; --d) {
            v *= bs[d];
            if (q[d] >= bs[d]) return -1;
            v += q[d];
        }
        return v;
    };
    // the position of chunk id `key`, n when there is no such chunk
    auto find = [&](int64_t key) -> int64_t {
        if constexpr (DIRECT) {
            if (key < 0 || key >= idmax) return n;
            const int32_t p = tab[key];
            return p < 0 ? n : (int64_t)p;
        } else {
            return find_chunk_near(keys, n, c, keys[c], key);
        }
    };
    int64_t pos;
    if (type == SG_NN_STAR) {
        pos = -1;
        if (k < 2 * dim) {
            const int d = k / 2;
            const int64_t dPos = bc[d] + (k % 2) * 2 - 1;
            bc[d] = dPos;
            int64_t bl = block_lin_id(bc);
            bl = (dPos < 0) ? -1 : bl;
            pos = find(bl);
        }
    } else {
        pos = n;
        if (k != nnn / 2) {
            int cnt = k;
            for (int d = 0; d < dim; ++d) {
                bc[d] += cnt % 3 - 1;
                cnt /= 3;
            }
            pos = find(block_lin_id(bc));
        }
    }
    nn[t] = (int32_t)pos;
}

// neighbour chunk of an enlarged-block cell; dir[d] in {-1,0,1}
// star tables index face neighbours as 2*d + (dir>0); full tables sum (dir+1)*3^d.
// Returns -1 when the cell is an edge/corner cell under a star table (never read by
// star stencils; the reference loads arbitrary data there).
__device__ inline int64_t ghost_source(const int32_t *nnrow, int type, int dim, const int *dir, int64_t self)
{
    int nz = 0, k = 0, acc = 1, star = -1;
    for (int d = 0; d < dim; ++d) {
        if (dir[d] != 0) {
            ++nz;
            star = 2 * d + (dir[d] > 0);
        }
        k += (dir[d] + 1) * acc;
        acc *= 3;
    }
    if (nz == 0) return self;
    if (type == SG_NN_STAR) return (nz == 1) ? (int64_t)nnrow[star] : -1;
    return (int64_t)nnrow[k];
}

// stage the enlarged block of one chunk (edge + 2 cells per dimension) in LDS
template <typename T>
__device__ void load_enlarged(T *eb, const T *arr, const int32_t *nnrow, int type, int64_t c, int64_t bgpos, int dim,
                              int edge, int csz, T bgv)
{
    const int e = edge + 2;
    const int ebs = (dim == 3) ? e * e * e : e * e;
    for (int i = threadIdx.x; i < ebs; i += blockDim.x) {
        int crd[3] = {0, 0, 0}, dir[3] = {0, 0, 0}, t = i, off = 0, acc = 1;
        for (int d = 0; d < dim; ++d) {
            crd[d] = t % e - 1;
            t /= e;
            dir[d] = (crd[d] < 0) ? -1 : (crd[d] >= edge ? 1 : 0);
            int w = crd[d] - dir[d] * edge;
            off += w * acc;
            acc *= edge;
        }
        int64_t src = ghost_source(nnrow, type, dim, dir, c);
        T v = bgv;
        if (src >= 0) v = arr[src * csz + off];
        eb[i] = v;
    }
    (void)bgpos;
}

struct StencilArgs {
    int kind, apply, use_box;
    int wmode;  // 8^3 fast path stores: 0 active points only, 1 whole rows, 2 whole rows non-temporal,
                // 3 as 2 through LDS as lane-contiguous 16-byte vectors
    int lo[3], hi[3];
    double bgd[2];  // destination backgrounds
    double bgs[2];  // source backgrounds
    int lskip;      // 8^3 fast path: read the mask first and skip loading rows without points
    int zorder;     // chunk ids are Morton codes (grid_zmb)
    // skip mode 4: per chunk 16 words, word z (0-7) bit x + 8 y = mask byte non-zero, word
    // 8 + z the same for EXIST without PADDING (the masks as the stencil needs them)
    const uint64_t *mplanes;
    double coef[27];
    int64_t icoef[27];
};

// generic stencil: one workgroup per chunk, enlarged blocks of up to two
// properties in dynamic LDS.  dst_alias[q] != 0 means dst q is also a source: it is
// then written for every slot (new value where active, the old value elsewhere)
// into a separate buffer that replaces the property afterwards.
template <typename T>
__global__ __launch_bounds__(kWG) void stencil_generic_kernel(const int64_t *keys, int64_t n, GeomArgs ga,
                                                             const int32_t *nn, int nnn, int type, const uint8_t *mask,
                                                             const T *src0, const T *src1, T *dst0, T *dst1,
                                                             int alias0, int alias1, T bg0, T bg1, StencilArgs sa)
{
    extern __shared__ unsigned char smem_raw[];
    T *eb0 = (T *)smem_raw;
    const int dim = ga.dim, edge = ga.edge, csz = ga.csz, e = edge + 2;
    const int ebs = (dim == 3) ? e * e * e : e * e;
    T *eb1 = eb0 + ebs;
    const int64_t c = blockIdx.x;
    if (c >= n) return;
    const int32_t *nnrow = nn + c * nnn;
    load_enlarged<T>(eb0, src0, nnrow, type, c, n, dim, edge, csz, bg0);
    if (src1) load_enlarged<T>(eb1, src1, nnrow, type, c, n, dim, edge, csz, bg1);
    __syncthreads();
    int64_t base[3] = {0, 0, 0};
    chunk_coords(lin_of(ga), keys[c], base);
#pragma unroll
    for (int d = 0; d < 3; ++d) base[d] *= edge;
    const int kind = sa.kind;
    for (int o = threadIdx.x; o < csz; o += blockDim.x) {
        int t = o, lin = 0, acc = 1;
        int64_t pc[3] = {0, 0, 0}, par = 0;
        // loops to the constant 3, stopping at dim: the small arrays stay in registers
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d >= dim) break;
            const int lc = t % edge;
            pc[d] = base[d] + lc;
            par += pc[d];
            lin += (lc + 1) * acc;
            t /= edge;
            acc *= e;
        }
        uint8_t cm = mask[c * csz + o];
        if (sa.use_box) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
                if (d < dim && (pc[d] < sa.lo[d] || pc[d] > sa.hi[d])) cm = 0;
        }
        bool active = (sa.apply == SG_APPLY_EXISTING) ? (cm != 0) : ((cm & SG_EXIST) && !(cm & SG_PADDING));
        const T cur = eb0[lin];
        T r0 = alias0 ? eb0[lin] : (T)0, r1 = alias1 ? eb1[lin] : (T)0;
        if (alias0 == 2) r0 = eb1[lin];
        if (alias1 == 1) r1 = eb0[lin];
        if (kind == SG_ST_RBSOR && (int)(par & 1) != (int)sa.icoef[0]) active = false;
        if (active) {
            if constexpr (!std::is_integral<T>::value) {
                if (kind == SG_ST_LAP || kind == SG_ST_HEAT || kind == SG_ST_RESIDUAL) {
                    T a;
                    if (kind == SG_ST_LAP) a = (T)sa.coef[0] * cur;
                    else if (kind == SG_ST_HEAT) a = (T)(-2.0 * dim) * cur;
                    else a = (T)(2.0 * dim) * cur;
                    int s = 1;
                    for (int d = 0; d < dim; ++d) {
                        T pair = eb0[lin - s] + eb0[lin + s];
                        if (kind == SG_ST_LAP) a = a + (T)sa.coef[1] * pair;
                        else if (kind == SG_ST_HEAT) a = a + pair;
                        else a = a - pair;
                        s *= e;
                    }
                    if (kind == SG_ST_LAP) r0 = a;
                    else if (kind == SG_ST_HEAT) r0 = cur + (T)sa.coef[0] * a;
                    else r0 = eb1[lin] - a;
                } else if (kind == SG_ST_GRAY_SCOTT) {
                    const T cu = cur, cv = eb1[lin];
                    T lu = (T)(-2.0 * dim) * cu, lv = (T)(-2.0 * dim) * cv;
                    int s = 1;
                    for (int d = 0; d < dim; ++d) {
                        lu = lu + (eb0[lin - s] + eb0[lin + s]);
                        lv = lv + (eb1[lin - s] + eb1[lin + s]);
                        s *= e;
                    }
                    const T Du = (T)sa.coef[0], Dv = (T)sa.coef[1], F = (T)sa.coef[2], k = (T)sa.coef[3],
                            dt = (T)sa.coef[4];
                    const T uvv = (cu * cv) * cv;
                    r0 = cu + dt * ((Du * lu - uvv) + F * ((T)1 - cu));
                    r1 = cv + dt * ((Dv * lv + uvv) - (F + k) * cv);
                } else if (kind == SG_ST_RBSOR) {
                    T nbs = 0;
                    int s = 1;
                    for (int d = 0; d < dim; ++d) {
                        nbs = nbs + (eb0[lin - s] + eb0[lin + s]);
                        s *= e;
                    }
                    const T w = (T)sa.coef[0];
                    const T gs = (eb1[lin] + nbs) / (T)(2 * dim);
                    r0 = cur + w * (gs - cur);
                }
            }
            if (kind == SG_ST_BOX) {
                // Conv3x3x3 order: x offset outermost, z innermost
                T tot = 0;
                const int nk = (dim == 3) ? 3 : 1;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        for (int k = 0; k < nk; ++k) {
                            const int off = (i - 1) + (j - 1) * e + ((dim == 3) ? (k - 1) * e * e : 0);
                            const int ci = (dim == 3) ? (i * 9 + j * 3 + k) : (i * 3 + j);
                            const T cf = std::is_integral<T>::value ? (T)sa.icoef[ci] : (T)sa.coef[ci];
                            tot = dev_add<T>(tot, dev_mul<T>(eb0[lin + off], cf));
                        }
                r0 = tot;
            }
        }
        const int64_t slot = c * csz + o;
        if (dst0 && (active || alias0)) dst0[slot] = r0;
        if (dst1 && (active || alias1)) dst1[slot] = r1;
    }
}

// Point-list stencil (applyStencilInPlaceNoShared, SparseGridGpu_kernels.cuh:1096-1149):
// one thread per existing point of the list `ep` (data positions chunk * csz + offset),
// neighbours read straight from the chunk arrays through the neighbour table (no LDS
// tile).  The arithmetic is stencil_generic_kernel's, term for term, so both modes give
// the same bits; a missing neighbour chunk is the background chunk.
template <typename T>
__device__ inline T point_fetch(const T *src, const int32_t *nnrow, int type, int64_t c, const int *l, int dx, int dy,
                                int dz, int dim, int edge, int csz)
{
    const int dd[3] = {dx, dy, dz};
    int disp[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    bool moved = false;
    for (int d = 0; d < dim; ++d) {
        int v = l[d] + dd[d];
        disp[d] = (v < 0) ? -1 : (v >= edge ? 1 : 0);
        w[d] = v - disp[d] * edge;
        moved = moved || disp[d] != 0;
    }
    int64_t nb = c;
    if (moved) {
        if (type == SG_NN_STAR) {
            int k = 0;
            for (int d = 0; d < dim; ++d)
                if (disp[d]) k = 2 * d + (disp[d] > 0 ? 1 : 0);
            nb = nnrow[k];
        } else {
            int k = 0, m = 1;
            for (int d = 0; d < dim; ++d) {
                k += (disp[d] + 1) * m;
                m *= 3;
            }
            nb = nnrow[k];
        }
    }
    int off = w[dim - 1];
    for (int d = dim - 2; d >= 0; --d) off = off * edge + w[d];
    return src[nb * csz + off];
}

template <typename T>
__global__ __launch_bounds__(kWG) void stencil_points_kernel(const int64_t *ep, int64_t ne, const int64_t *keys,
                                                            GeomArgs ga, const int32_t *nn, int nnn, int type,
                                                            const uint8_t *mask, const T *src0, const T *src1,
                                                            T *dst0, T *dst1, StencilArgs sa)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const int dim = ga.dim, edge = ga.edge, csz = ga.csz;
    const int64_t id = ep[i], c = id / csz;
    const int o = (int)(id % csz);
    int l[3] = {0, 0, 0};
    int64_t pc[3] = {0, 0, 0}, par = 0;
    chunk_coords(lin_of(ga), keys[c], pc);
    {
        int t = o;
        for (int d = 0; d < dim; ++d) {
            l[d] = t % edge;
            t /= edge;
            pc[d] = pc[d] * edge + l[d];
            par += pc[d];
        }
    }
    uint8_t cm = mask[id];
    if (sa.use_box)
        for (int d = 0; d < dim; ++d)
            if (pc[d] < sa.lo[d] || pc[d] > sa.hi[d]) cm = 0;
    bool active = (sa.apply == SG_APPLY_EXISTING) ? (cm != 0) : ((cm & SG_EXIST) && !(cm & SG_PADDING));
    const int kind = sa.kind;
    if (kind == SG_ST_RBSOR && (int)(par & 1) != (int)sa.icoef[0]) active = false;
    if (!active) return;
    const int32_t *nnrow = nn + c * nnn;
    auto at = [&](const T *src, int dx, int dy, int dz) {
        return point_fetch<T>(src, nnrow, type, c, l, dx, dy, dz, dim, edge, csz);
    };
    const T cur = src0[id];
    T r0 = (T)0, r1 = (T)0;
    if constexpr (!std::is_integral<T>::value) {
        if (kind == SG_ST_LAP || kind == SG_ST_HEAT || kind == SG_ST_RESIDUAL) {
            T a;
            if (kind == SG_ST_LAP) a = (T)sa.coef[0] * cur;
            else if (kind == SG_ST_HEAT) a = (T)(-2.0 * dim) * cur;
            else a = (T)(2.0 * dim) * cur;
            for (int d = 0; d < dim; ++d) {
                const T pair = at(src0, -(d == 0), -(d == 1), -(d == 2)) + at(src0, d == 0, d == 1, d == 2);
                if (kind == SG_ST_LAP) a = a + (T)sa.coef[1] * pair;
                else if (kind == SG_ST_HEAT) a = a + pair;
                else a = a - pair;
            }
            if (kind == SG_ST_LAP) r0 = a;
            else if (kind == SG_ST_HEAT) r0 = cur + (T)sa.coef[0] * a;
            else r0 = src1[id] - a;
        } else if (kind == SG_ST_GRAY_SCOTT) {
            const T cu = cur, cv = src1[id];
            T lu = (T)(-2.0 * dim) * cu, lv = (T)(-2.0 * dim) * cv;
            for (int d = 0; d < dim; ++d) {
                lu = lu + (at(src0, -(d == 0), -(d == 1), -(d == 2)) + at(src0, d == 0, d == 1, d == 2));
                lv = lv + (at(src1, -(d == 0), -(d == 1), -(d == 2)) + at(src1, d == 0, d == 1, d == 2));
            }
            const T Du = (T)sa.coef[0], Dv = (T)sa.coef[1], F = (T)sa.coef[2], k = (T)sa.coef[3], dt = (T)sa.coef[4];
            const T uvv = (cu * cv) * cv;
            r0 = cu + dt * ((Du * lu - uvv) + F * ((T)1 - cu));
            r1 = cv + dt * ((Dv * lv + uvv) - (F + k) * cv);
        } else if (kind == SG_ST_RBSOR) {
            T nbs = 0;
            for (int d = 0; d < dim; ++d)
                nbs = nbs + (at(src0, -(d == 0), -(d == 1), -(d == 2)) + at(src0, d == 0, d == 1, d == 2));
            const T w = (T)sa.coef[0];
            const T gs = (src1[id] + nbs) / (T)(2 * dim);
            r0 = cur + w * (gs - cur);
        }
    }
    if (kind == SG_ST_BOX) {
        T tot = 0;
        const int nk = (dim == 3) ? 3 : 1;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int k = 0; k < nk; ++k) {
                    const int ci = (dim == 3) ? (a * 9 + b * 3 + k) : (a * 3 + b);
                    const T cf = std::is_integral<T>::value ? (T)sa.icoef[ci] : (T)sa.coef[ci];
                    tot = dev_add<T>(tot, dev_mul<T>(at(src0, a - 1, b - 1, (dim == 3) ? k - 1 : 0), cf));
                }
        r0 = tot;
    }
    if (dst0) dst0[id] = r0;
    if (dst1) dst1[id] = r1;
}

// existing points per chunk, then their data positions in chunk order and increasing
// offset (calc_exist_points + fill_e_points, SparseGridGpu_kernels.cuh:1150-1233)
__global__ __launch_bounds__(kWG) void count_exist_kernel(const uint8_t *mask, int csz, uint32_t *cnt)
{
    __shared__ uint32_t s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    uint32_t c = 0;
    for (int o = threadIdx.x; o < csz; o += kWG) c += (mask[(int64_t)blockIdx.x * csz + o] & SG_EXIST) ? 1u : 0u;
    if (c) atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = s;
}

__global__ __launch_bounds__(kWG) void fill_exist_kernel(const uint8_t *mask, int csz, const uint32_t *start,
                                                        int64_t *out)
{
    __shared__ uint32_t s_wave[kWG / 64];
    __shared__ uint32_t s_run;
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_run = 0;
    __syncthreads();
    for (int o0 = 0; o0 < csz; o0 += kWG) {
        const int o = o0 + threadIdx.x;
        const bool has = o < csz && (mask[c * csz + o] & SG_EXIST);
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
            out[start[c] + pre + (uint32_t)__popcll(bal & lt)] = c * csz + o;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_run += tot;
        __syncthreads();
    }
}

// tagBoundaries (SparseGridGpu_kernels.cuh:667-737): padding = an existing point with
// a star (NNStar::isPadding) or full (NNfull_is_padding_impl) neighbour that does not
// exist; the enlarged mask block takes missing chunks from the background (mask 0).
__global__ __launch_bounds__(kWG) void tag_boundaries_kernel(int64_t n, GeomArgs ga, const int32_t *nn, int nnn,
                                                            int type, uint8_t *mask)
{
    extern __shared__ unsigned char smem_raw[];
    uint8_t *eb = smem_raw;
    const int dim = ga.dim, edge = ga.edge, csz = ga.csz, e = edge + 2;
    const int64_t c = blockIdx.x;
    if (c >= n) return;
    load_enlarged<uint8_t>(eb, mask, nn + c * nnn, type, c, n, dim, edge, csz, (uint8_t)0);
    __syncthreads();
    for (int o = threadIdx.x; o < csz; o += blockDim.x) {
        int t = o, lin = 0, acc = 1;
        for (int d = 0; d < dim; ++d) {
            lin += (t % edge + 1) * acc;
            t /= edge;
            acc *= e;
        }
        const uint8_t cur = eb[lin];
        if (!(cur & 1)) continue;
        bool pad = false;
        if (type == SG_NN_STAR) {
            int s = 1;
            for (int d = 0; d < dim; ++d) {
                pad = pad || !(eb[lin + s] & 1) || !(eb[lin - s] & 1);
                s *= e;
            }
        } else {
            const int nb = (dim == 3) ? 27 : 9;
            for (int k = 0; k < nb; ++k) {
                int kk = k, off = 0, s = 1;
                for (int d = 0; d < dim; ++d) {
                    off += (kk % 3 - 1) * s;
                    kk /= 3;
                    s *= e;
                }
                pad = pad || !(eb[lin + off] & 1);
            }
        }
        mask[c * csz + o] = pad ? (uint8_t)(cur | 2) : (uint8_t)(cur & ~2);
    }
}

// ---- 3D 8^3 tagBoundaries, bit-parallel ---------------------------------------------
// One wave per chunk, lane = x-row (y = l & 7, z = l >> 3).  Each wave builds a 10 x 10
// tile of 10-bit EXIST words, one per row (z, y) in [-1, 8]^2, bit j <-> x = j - 1, from
// the chunk, its face / edge / corner neighbours (through the neighbour table; a missing
// chunk is the background chunk, mask 0).  A point is padding when one of its star (6)
// or full (26) neighbours does not exist: AND of shifted words.  Same result as the
// generic kernel point by point; masks are read and written as 8-byte rows.
__device__ inline uint32_t exist_bits8(uint64_t row)
{
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) b |= (uint32_t)((row >> (8 * i)) & 1u) << i;
    return b;
}

// a mask row (8 bytes) with the PADDING bit of every existing point set to bit x of pad
// and cleared otherwise; points without EXIST are left as they are
__device__ inline uint64_t apply_padding(uint64_t row, uint32_t pad)
{
    // pad bit x -> bit 8x (four bits at a time: x * 0x00204081 places bit x at 8x with
    // no two partial products overlapping)
    const uint64_t p = (uint64_t)(((pad & 0xFu) * 0x00204081u) & 0x01010101u) |
                       ((uint64_t)((((pad >> 4) & 0xFu) * 0x00204081u) & 0x01010101u) << 32);
    const uint64_t e = row & 0x0101010101010101ull;
    return (row & ~(e << 1)) | ((p & e) << 1);
}

// ghost row i (0..35) of the 10 x 10 tile: (yy, zz) in [-1, 8]^2 outside [0, 7]^2
// (selects, not branches: branchy stores through the references end up in scratch)
__device__ inline void ghost_row_yz(int i, int &yy, int &zz)
{
    const bool zface = i < 20;  // rows 0-19: z = -1 / 8 faces; rows 20-35: y = -1 / 8 faces
    const int along = zface ? (i < 10 ? i - 1 : i - 11) : (i < 28 ? i - 20 : i - 28);
    const int face = zface ? (i < 10 ? -1 : 8) : (i < 28 ? -1 : 8);
    yy = zface ? along : face;
    zz = zface ? face : along;
}

// neighbour-table entry of chunk displacement (dx, dy, dz): star order -x +x -y +y -z +z,
// full index (dx+1) + 3 (dy+1) + 9 (dz+1); -1 when a star table has no such entry
__device__ inline int nn_entry(int type, int dx, int dy, int dz)
{
    if (type == SG_NN_FULL) return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
    const int nz = (dx != 0) + (dy != 0) + (dz != 0);
    if (nz == 0) return -2;  // the chunk itself
    if (nz > 1) return -1;
    if (dx) return dx < 0 ? 0 : 1;
    if (dy) return dy < 0 ? 2 : 3;
    return dz < 0 ? 4 : 5;
}

// One wave per CPW consecutive chunks (lane = x-row of each): every load of the wave's
// chunks (own mask rows, x-ghost EXIST bits, the 36 ghost rows of the y / z faces and, for
// the full neighbourhood, their x-ghost bits) is issued before any is used, so a wave
// keeps CPW chunks of loads in flight; then each chunk's 10 x 10 tile of EXIST words is
// staged in LDS and its points tagged.
template <int WAVES, int CPW>
__global__ __launch_bounds__(64 * WAVES) void tag_boundaries8_kernel(int64_t n, const int32_t *__restrict__ nn,
                                                                    int nnn, int type, uint8_t *__restrict__ mask,
                                                                    int64_t per_xcd)
{
    __shared__ uint16_t tile[WAVES][CPW][100];
    // uniform wave index: the neighbour-table rows are read by scalar loads
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // workgroup b runs in XCD b & 7, which walks its own contiguous range of workgroups'
    // chunks: neighbours' mask lines then mostly come from the same XCD's L2
    const int64_t wg = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const int64_t cb = (wg * WAVES + w) * CPW;
    if (cb >= n) return;  // waves are independent: no workgroup barrier below
    const int y = lane & 7, z = lane >> 3;
    int gyy, gzz;
    ghost_row_yz(lane < 36 ? lane : 0, gyy, gzz);
    const int gdy = (gyy < 0) ? -1 : (gyy > 7 ? 1 : 0), gdz = (gzz < 0) ? -1 : (gzz > 7 ? 1 : 0);
    const int gly = gyy - gdy * 8, glz = gzz - gdz * 8;
    uint64_t own[CPW], grow[CPW];
    uint32_t xm[CPW], xp[CPW], gxm[CPW], gxp[CPW];
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        const int64_t c = cb + k;
        own[k] = 0;
        grow[k] = 0;
        xm[k] = xp[k] = gxm[k] = gxp[k] = 0;
        if (c >= n) continue;
        const int32_t *row_nn = nn + c * nnn;
        auto chunk_of = [&](int dx, int dy, int dz) -> int64_t {
            const int e = nn_entry(type, dx, dy, dz);
            if (e == -2) return c;
            return (e < 0) ? -1 : (int64_t)row_nn[e];
        };
        // EXIST bit of x-row (ly, lz), point lx, of the chunk at displacement (dx, dy, dz)
        auto xbit = [&](int dx, int dy, int dz, int ly, int lz, int lx) -> uint32_t {
            const int64_t nb = chunk_of(dx, dy, dz);
            if (nb < 0) return 0u;
            return mask[nb * 512 + (lz * 8 + ly) * 8 + lx];
        };
        own[k] = *(const uint64_t *)(mask + c * 512 + lane * 8);
        xm[k] = xbit(-1, 0, 0, y, z, 7);
        xp[k] = xbit(1, 0, 0, y, z, 0);
        if (lane < 36) {
            const int64_t nb = chunk_of(0, gdy, gdz);
            if (nb >= 0) grow[k] = *(const uint64_t *)(mask + nb * 512 + (glz * 8 + gly) * 8);
            if (type == SG_NN_FULL) {
                gxm[k] = xbit(-1, gdy, gdz, gly, glz, 7);
                gxp[k] = xbit(1, gdy, gdz, gly, glz, 0);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        if (cb + k >= n) continue;
        tile[w][k][(z + 1) * 10 + (y + 1)] =
            (uint16_t)((xm[k] & 1u) | (exist_bits8(own[k]) << 1) | ((xp[k] & 1u) << 9));
        if (lane < 36)
            tile[w][k][(gzz + 1) * 10 + (gyy + 1)] =
                (uint16_t)((exist_bits8(grow[k]) << 1) | (gxm[k] & 1u) | ((gxp[k] & 1u) << 9));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        const int64_t c = cb + k;
        if (c >= n) continue;
        const uint16_t *t = tile[w][k];
        const uint32_t self = t[(z + 1) * 10 + (y + 1)];
        uint32_t all;  // bit j: every neighbour of x = j - 1 exists
        if (type == SG_NN_FULL) {
            all = 0x3FFu;
#pragma unroll
            for (int dz = 0; dz < 3; ++dz)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    const uint32_t r = t[(z + dz) * 10 + (y + dy)];
                    all &= r & (r << 1) & (r >> 1);
                }
        } else {
            all = (self << 1) & (self >> 1) & t[(z + 1) * 10 + y] & t[(z + 1) * 10 + y + 2] & t[z * 10 + y + 1] &
                  t[(z + 2) * 10 + y + 1];
        }
        const uint32_t pad = (self >> 1) & ~(all >> 1) & 0xFFu;  // existing and not surrounded
        const uint64_t o = own[k];
        const uint64_t out = apply_padding(o, pad);
        if (out != o) *(uint64_t *)(mask + c * 512 + lane * 8) = out;
    }
}

// ---- EXIST bit-planes of 8^3 chunks -------------------------------------------------
// 64 bytes per chunk: byte l = 8z + y holds the EXIST bits of row (y, z), bit x for point
// x (so word z has bit 8y + x).  A thread per row, kEbitsRows consecutive rows each (all
// loads first); the byte is gathered by one multiply: bit 0 of byte i sits at 8i, and
// times 0x0102040810204080 it lands on 56 + i, every partial product at its own bit.
// Chunk n (the background chunk) gets zero bytes.
constexpr int kEbitsRows = 8;
__global__ __launch_bounds__(256) void ebits_kernel(const uint8_t *__restrict__ mask, int64_t n,
                                                    uint64_t *__restrict__ eb)
{
    const int64_t rows = (n + 1) * 64;
    const int64_t r0 = (int64_t)blockIdx.x * (256 * kEbitsRows) + threadIdx.x;
    uint64_t row[kEbitsRows];
#pragma unroll
    for (int k = 0; k < kEbitsRows; ++k) {
        const int64_t r = r0 + (int64_t)k * 256;
        row[k] = (r < n * 64) ? *(const uint64_t *)(mask + r * 8) : 0ull;
    }
#pragma unroll
    for (int k = 0; k < kEbitsRows; ++k) {
        const int64_t r = r0 + (int64_t)k * 256;
        if (r < rows)
            ((uint8_t *)eb)[r] = (uint8_t)(((row[k] & 0x0101010101010101ull) * 0x0102040810204080ull) >> 56);
    }
}

// ---- 8^3 tagBoundaries on whole plane words -----------------------------------------
// Plane word z of a chunk holds its slice z: bit 8y + x = EXIST of point (x, y, z).  The
// neighbours of every point of a slice in x and y are shifts of the word, filled in at
// the slice's edges from the same word of the x / y neighbour chunks:
//   xlo(C, L): E(x - 1, y) (x = 0 from L's x = 7), xhi(C, R): E(x + 1, y),
//   ylo(C, M): E(x, y - 1) (y = 0 from M's y = 7), yhi(C, P): E(x, y + 1).
__device__ inline uint64_t pw_xlo(uint64_t c, uint64_t l)
{
    return ((c << 1) & 0xFEFEFEFEFEFEFEFEull) | ((l >> 7) & 0x0101010101010101ull);
}
__device__ inline uint64_t pw_xhi(uint64_t c, uint64_t r)
{
    return ((c >> 1) & 0x7F7F7F7F7F7F7F7Full) | ((r << 7) & 0x8080808080808080ull);
}
__device__ inline uint64_t pw_ylo(uint64_t c, uint64_t m) { return (c << 8) | (m >> 56); }
__device__ inline uint64_t pw_yhi(uint64_t c, uint64_t p) { return (c >> 8) | (p << 56); }

// A wave takes 6 consecutive chunks; lane l < 60 -> chunk k = l / 10 and slice zz =
// l % 10 - 1 in [-1, 8] (slices -1 and 8 are slice 7 of the chunk below and slice 0 of the
// chunk above).  Each lane loads the plane words of its slice and its x / y neighbour
// chunks (table entries and words in two dependent steps), and reduces them to S = AND
// over the 3 x 3 (x, y) neighbourhood (FULL) or to its star in the slice (STAR); slices
// z - 1 and z + 1 come from lanes l -/+ 1.  Not surrounded = E & ~(S(z-1) & S(z) &
// S(z+1)), or E & ~(star & E(z-1) & E(z+1)).  Then, lane = row, the chunks' mask rows
// get their PADDING bits (apply_padding).  Same result as tag_boundaries8_kernel.
template <int TYPE>
__global__ __launch_bounds__(256) void tag_planes_kernel(int64_t n, const int32_t *__restrict__ nn, int nnn,
                                                        const uint64_t *__restrict__ eb, uint8_t *__restrict__ mask,
                                                        int64_t per_xcd)
{
    constexpr int CPW = 6;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // workgroup b runs in XCD b & 7, which walks its own contiguous range of the chunks
    const int64_t wg = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    const int64_t cb = (wg * 4 + w) * CPW;
    if (cb >= n) return;  // waves are independent: no workgroup barrier
    // the chunks' mask rows (lane = row), used last, loaded first
    uint64_t own[CPW];
#pragma unroll
    for (int q = 0; q < CPW; ++q)
        own[q] = (cb + q < n) ? *(const uint64_t *)(mask + (cb + q) * 512 + lane * 8) : 0ull;
    const int k = lane / 10, s = lane - 10 * k, zz = s - 1;
    const int64_t c = cb + k;
    const bool act = lane < 10 * CPW && c < n;
    const int dz = zz < 0 ? -1 : (zz > 7 ? 1 : 0), wz = zz & 7;
    const int32_t *row = nn + c * nnn;
    uint64_t S, E;
    if (TYPE == SG_NN_FULL) {
        // chunks (dx, dy, dz): entry (dx + 1) + 3 (dy + 1) + 9 (dz + 1); the chunk itself at
        // 13.  The three entries of one dy are consecutive: one 12-byte load each.
        int64_t nb[3][3];
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            int32_t t3[3] = {-1, -1, -1};
            if (act) __builtin_memcpy(t3, row + 9 * (dz + 1) + 3 * (dy + 1), 12);
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
                nb[dy + 1][dx + 1] = (dz == 0 && dy == 0 && dx == 0) ? (act ? c : -1) : (int64_t)t3[dx + 1];
        }
        uint64_t W[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int x = 0; x < 3; ++x) W[r][x] = nb[r][x] >= 0 ? eb[nb[r][x] * 8 + wz] : 0ull;
        uint64_t X[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) X[r] = W[r][1] & pw_xlo(W[r][1], W[r][0]) & pw_xhi(W[r][1], W[r][2]);
        S = X[1] & pw_ylo(X[1], X[0]) & pw_yhi(X[1], X[2]);
        E = W[1][1];
    } else {
        // star table: -x +x -y +y -z +z; slices -1 / 8 only need the word of chunk (0, 0, dz)
        const int64_t self = !act ? -1 : (dz == 0 ? c : (int64_t)row[dz < 0 ? 4 : 5]);
        int64_t f[4] = {-1, -1, -1, -1};
#pragma unroll
        for (int q = 0; q < 4; ++q) f[q] = (act && dz == 0) ? (int64_t)row[q] : -1;
        E = self >= 0 ? eb[self * 8 + wz] : 0ull;
        uint64_t F[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) F[q] = f[q] >= 0 ? eb[f[q] * 8 + wz] : 0ull;
        S = E & pw_xlo(E, F[0]) & pw_xhi(E, F[1]) & pw_ylo(E, F[2]) & pw_yhi(E, F[3]);
    }
    // slices z - 1 and z + 1 from the neighbour lanes (valid for the slices 0..7 of a chunk)
    const uint32_t slo = (uint32_t)S, shi = (uint32_t)(S >> 32), elo = (uint32_t)E, ehi = (uint32_t)(E >> 32);
    uint64_t below, above;
    if (TYPE == SG_NN_FULL) {
        below = (uint64_t)(uint32_t)__shfl_up((int)slo, 1, 64) | ((uint64_t)(uint32_t)__shfl_up((int)shi, 1, 64) << 32);
        above = (uint64_t)(uint32_t)__shfl_down((int)slo, 1, 64) |
                ((uint64_t)(uint32_t)__shfl_down((int)shi, 1, 64) << 32);
    } else {
        below = (uint64_t)(uint32_t)__shfl_up((int)elo, 1, 64) | ((uint64_t)(uint32_t)__shfl_up((int)ehi, 1, 64) << 32);
        above = (uint64_t)(uint32_t)__shfl_down((int)elo, 1, 64) |
                ((uint64_t)(uint32_t)__shfl_down((int)ehi, 1, 64) << 32);
    }
    const uint64_t pad = E & ~(S & below & above);
    // lane = row (y, z) = (l & 7, l >> 3) of chunk q: its slice's pad word is lane
    // q * 10 + z + 1's
    const uint32_t plo = (uint32_t)pad, phi = (uint32_t)(pad >> 32);
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int src = q * 10 + (lane >> 3) + 1;
        const uint64_t pw = (uint64_t)(uint32_t)__shfl((int)plo, src, 64) |
                            ((uint64_t)(uint32_t)__shfl((int)phi, src, 64) << 32);
        if (cb + q >= n) continue;
        const uint64_t out = apply_padding(own[q], (uint32_t)(pw >> (8 * (lane & 7))) & 0xFFu);
        if (out != own[q]) *(uint64_t *)(mask + (cb + q) * 512 + lane * 8) = out;
    }
}

// ---- 3D 8^3 27-point box ----------------------------------------------------------
// One wave per chunk, lane = x-row.  The wave stages a 10 x 10 x 10 tile of the source
// (rows (z, y) in [-1, 8]^2, 10 values with the x ghosts) from the chunk and its 26
// neighbours (full neighbour table; a missing chunk is the background chunk), then each
// lane sums its 8 points over the 3 x 3 x 3 box in stencil_generic_kernel's order (x offset
// outermost, z innermost).  Rows whose points are all active or empty are stored whole
// (empty points take the destination background); other rows point by point.
// Per-lane loads of one chunk for the box: the own x-row with its two x-ghost values,
// one ghost row of the y / z faces (lanes 0-35) with its x ghosts, and the mask row.
template <typename T> struct BoxIn {
    T own[8], gr[8];
    T oxm, oxp, gxm, gxp;
    uint64_t mrow;
};

template <typename T>
__device__ inline void box_load(BoxIn<T> &in, int64_t c, const int32_t *__restrict__ nn,
                                const uint8_t *__restrict__ mask, const T *__restrict__ src, int lane,
                                bool skip_xm = false, bool skip_xp = false)
{
    const int32_t *row_nn = nn + c * 27;
    auto chunk_of = [&](int dx, int dy, int dz) -> int64_t {
        return (dx == 0 && dy == 0 && dz == 0) ? c : (int64_t)row_nn[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
    };
    in.mrow = *(const uint64_t *)(mask + c * 512 + lane * 8);
    const T *sr = src + c * 512 + lane * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) in.own[i] = sr[i];
    // skip_xm / skip_xp: the -x / +x ghosts come from a chunk the caller holds
    if (!skip_xm) in.oxm = src[chunk_of(-1, 0, 0) * 512 + lane * 8 + 7];
    if (!skip_xp) in.oxp = src[chunk_of(1, 0, 0) * 512 + lane * 8 + 0];
    if (lane < 36) {
        int yy, zz;
        ghost_row_yz(lane, yy, zz);
        const int dy = (yy < 0) ? -1 : (yy > 7 ? 1 : 0), dz = (zz < 0) ? -1 : (zz > 7 ? 1 : 0);
        const int lrow = (zz - dz * 8) * 8 + (yy - dy * 8);
        const T *gs = src + chunk_of(0, dy, dz) * 512 + lrow * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) in.gr[i] = gs[i];
        if (!skip_xm) in.gxm = src[chunk_of(-1, dy, dz) * 512 + lrow * 8 + 7];
        if (!skip_xp) in.gxp = src[chunk_of(1, dy, dz) * 512 + lrow * 8 + 0];
    }
}

// Stages chunk c's 10 x 10 x 10 tile from `in`, sums the 3 x 3 x 3 box of each of the
// lane's 8 points and writes the chunk back through the tile.
template <typename T, bool UNI>
__device__ inline void box_finish(const BoxIn<T> &in, int64_t c, T *tw, T *__restrict__ dst,
                                  const int64_t *__restrict__ keys, int64_t bsz0, int64_t bsz1, T bg_dst,
                                  const StencilArgs &sa, int lane)
{
    const int y = lane & 7, z = lane >> 3;
    {
        T *dstrow = tw + ((z + 1) * 10 + (y + 1)) * 10;
#pragma unroll
        for (int i = 0; i < 8; ++i) dstrow[i + 1] = in.own[i];
        dstrow[0] = in.oxm;
        dstrow[9] = in.oxp;
    }
    if (lane < 36) {
        int yy, zz;
        ghost_row_yz(lane, yy, zz);
        T *dstrow = tw + ((zz + 1) * 10 + (yy + 1)) * 10;
#pragma unroll
        for (int i = 0; i < 8; ++i) dstrow[i + 1] = in.gr[i];
        dstrow[0] = in.gxm;
        dstrow[9] = in.gxp;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint64_t mrow = in.mrow;
    // row flag: 0 no point (the destination row already holds the background), 1 whole
    // row, 2 active points only (bits 8-15: which)
    uint32_t fl = 0;
    T out[8];
    if (mrow != 0) {
        bool inbox[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) inbox[x] = true;
        if (sa.use_box) {
            ChunkLin L;
            L.dim = 3;
            L.zorder = sa.zorder;
            L.bsz[0] = bsz0;
            L.bsz[1] = bsz1;
            L.bsz[2] = 1 << 30;
            int64_t bc[3];
            chunk_coords(L, keys[c], bc);
            const int64_t py = bc[1] * 8 + y, pz = bc[2] * 8 + z;
            const bool yz = py >= sa.lo[1] && py <= sa.hi[1] && pz >= sa.lo[2] && pz <= sa.hi[2];
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const int64_t px = bc[0] * 8 + x;
                inbox[x] = yz && px >= sa.lo[0] && px <= sa.hi[0];
            }
        }
        T cf[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) cf[q] = std::is_integral<T>::value ? (T)sa.icoef[q] : (T)sa.coef[q];
        T isum[8];
        if constexpr (std::is_integral<T>::value) {
            // wrapping integer sums do not depend on the order of the terms, so each of the
            // 9 rows of the 3 x 3 (y, z) neighbourhood is read from LDS once (10 values)
            // instead of once per tap
#pragma unroll
            for (int x = 0; x < 8; ++x) isum[x] = 0;
            if constexpr (UNI) {
                // every tap has the same coefficient: c (a + b + ...) = c a + c b + ... in
                // wrapping arithmetic, so the 27 values are summed (row triples first) and
                // multiplied once, no multiply per tap
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const T *row = tw + ((z + k) * 10 + (y + j)) * 10;
                        T r[10];
#pragma unroll
                        for (int q = 0; q < 10; ++q) r[q] = row[q];
#pragma unroll
                        for (int x = 0; x < 8; ++x)
                            isum[x] = dev_add<T>(isum[x], dev_add<T>(dev_add<T>(r[x], r[x + 1]), r[x + 2]));
                    }
#pragma unroll
                for (int x = 0; x < 8; ++x) isum[x] = dev_mul<T>(isum[x], cf[0]);
            } else {
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const T *row = tw + ((z + k) * 10 + (y + j)) * 10;
                        T r[10];
#pragma unroll
                        for (int q = 0; q < 10; ++q) r[q] = row[q];
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int x = 0; x < 8; ++x)
                                isum[x] = dev_add<T>(isum[x], dev_mul<T>(r[x + i], cf[i * 9 + j * 3 + k]));
                    }
            }
        }
        bool act[8];
        bool whole = true;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            uint8_t cm = (uint8_t)(mrow >> (8 * x));
            if (!inbox[x]) cm = 0;
            act[x] = (sa.apply == SG_APPLY_EXISTING) ? (cm != 0) : ((cm & SG_EXIST) && !(cm & SG_PADDING));
            const uint8_t raw = (uint8_t)(mrow >> (8 * x));
            whole = whole && (act[x] || raw == 0);
            T tot = 0;
            if constexpr (std::is_integral<T>::value) {
                tot = isum[x];
            } else {
                // floating point: the reference order (x offset outermost, z innermost)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
#pragma unroll
                        for (int k = 0; k < 3; ++k)
                            tot = dev_add<T>(tot,
                                             dev_mul<T>(tw[((z + k) * 10 + (y + j)) * 10 + (x + i)], cf[i * 9 + j * 3 + k]));
            }
            out[x] = act[x] ? tot : bg_dst;
        }
        if (whole) {
            fl = 1;
        } else {
            fl = 2;
#pragma unroll
            for (int x = 0; x < 8; ++x) fl |= (uint32_t)act[x] << (8 + x);
        }
    }
    // the tile is free once every lane has summed its points: stage the rows there and
    // write the chunk back contiguously
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (fl != 0) {
#pragma unroll
        for (int x = 0; x < 8; ++x) tw[lane * 8 + x] = out[x];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    store_chunk_contig(dst + c * 512, tw, fl, lane);
}

// One wave per chunk (lane = x-row).  Every load of the chunk (source rows, ghosts, mask
// row) is issued into registers before any is used, then the 10 x 10 x 10 tile of the
// source (rows (z, y) in [-1, 8]^2, 10 values with the x ghosts, from the chunk and its 26
// neighbours; a missing chunk is the background chunk) is staged in LDS and each lane
// sums its 8 points over the 3 x 3 x 3 box in stencil_generic_kernel's order (x offset
// outermost, z innermost).  Rows whose points are all active or empty are stored whole
// (empty points take the destination background); other rows point by point.
template <typename T, int WAVES, bool UNI>
__global__ __launch_bounds__(64 * WAVES) void box8_kernel(int64_t n, const int32_t *__restrict__ nn,
                                                         const uint8_t *__restrict__ mask, const T *__restrict__ src,
                                                         T *__restrict__ dst, const int64_t *__restrict__ keys,
                                                         int64_t bsz0, int64_t bsz1, T bg_dst, StencilArgs sa)
{
    __shared__ T tile[WAVES][100 * 10];
    // uniform wave index: the neighbour-table row is read by scalar loads
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * WAVES + w;
    if (c >= n) return;  // waves are independent
    BoxIn<T> a;
    box_load<T>(a, c, nn, mask, src, lane);
    box_finish<T, UNI>(a, c, tile[w], dst, keys, bsz0, bsz1, bg_dst, sa, lane);
}

// Two consecutive chunks c0 = 2p, c1 = 2p + 1 per wave.  When c1 is c0's +x neighbour (the
// table says so; consecutive chunk ids along x sort next to each other), the x ghosts
// between them are not gathered: c0's +x ghosts of its own and ghost rows are x = 0 of
// c1's same rows, c1's -x ghosts x = 7 of c0's, all in the lane's own registers (a ghost
// row of c1 comes from the chunk at (1, dy, dz) from c0, whose row c0's +x ghost is).
// Those gathers of single values from 64 or 36 rows each are most of the kernel's
// texture-addresser work.  Otherwise each chunk gathers its own.
template <typename T, int WAVES, bool UNI>
__global__ __launch_bounds__(64 * WAVES) void box8_pair_kernel(int64_t n, const int32_t *__restrict__ nn,
                                                              const uint8_t *__restrict__ mask,
                                                              const T *__restrict__ src, T *__restrict__ dst,
                                                              const int64_t *__restrict__ keys, int64_t bsz0,
                                                              int64_t bsz1, T bg_dst, StencilArgs sa)
{
    __shared__ T tile[WAVES][100 * 10];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t c0 = ((int64_t)blockIdx.x * WAVES + w) * 2, c1 = c0 + 1;
    if (c0 >= n) return;  // waves are independent
    const bool two = c1 < n;
    // full table entry (dx + 1) + 3 (dy + 1) + 9 (dz + 1) of (+1, 0, 0): 14
    const bool adj = two && nn[c0 * 27 + 14] == (int32_t)c1;
    BoxIn<T> a, b;
    box_load<T>(a, c0, nn, mask, src, lane, false, adj);
    if (two) box_load<T>(b, c1, nn, mask, src, lane, adj, false);
    if (adj) {
        a.oxp = b.own[0];
        b.oxm = a.own[7];
        if (lane < 36) {
            a.gxp = b.gr[0];
            b.gxm = a.gr[7];
        }
    }
    box_finish<T, UNI>(a, c0, tile[w], dst, keys, bsz0, bsz1, bg_dst, sa, lane);
    if (two) {
        // the tile is staged again: every lane's reads of the first chunk come first
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        box_finish<T, UNI>(b, c1, tile[w], dst, keys, bsz0, bsz1, bg_dst, sa, lane);
    }
}

// ---- 3D 8^3 star fast path ------------------------------------------------------
// LDS tile per property: rows (z, y) with z, y in [-1, 8] -> index (z+1)*10 + (y+1),
// each row 8 values.
template <typename T> struct Row8 {
    T v[8];
};

template <typename T> __device__ inline void load_row(Row8<T> &r, const T *p)
{
    if constexpr (sizeof(T) == 4) {
        const float4 *q = (const float4 *)p;
        float4 a = q[0], b = q[1];
        r.v[0] = __builtin_bit_cast(T, a.x); r.v[1] = __builtin_bit_cast(T, a.y);
        r.v[2] = __builtin_bit_cast(T, a.z); r.v[3] = __builtin_bit_cast(T, a.w);
        r.v[4] = __builtin_bit_cast(T, b.x); r.v[5] = __builtin_bit_cast(T, b.y);
        r.v[6] = __builtin_bit_cast(T, b.z); r.v[7] = __builtin_bit_cast(T, b.w);
    } else {
        const double2 *q = (const double2 *)p;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double2 a = q[i];
            r.v[2 * i] = __builtin_bit_cast(T, a.x);
            r.v[2 * i + 1] = __builtin_bit_cast(T, a.y);
        }
    }
}

template <typename T> __device__ inline void store_row(T *p, const Row8<T> &r)
{
    if constexpr (sizeof(T) == 4) {
        float4 *q = (float4 *)p;
        q[0] = make_float4(__builtin_bit_cast(float, r.v[0]), __builtin_bit_cast(float, r.v[1]),
                           __builtin_bit_cast(float, r.v[2]), __builtin_bit_cast(float, r.v[3]));
        q[1] = make_float4(__builtin_bit_cast(float, r.v[4]), __builtin_bit_cast(float, r.v[5]),
                           __builtin_bit_cast(float, r.v[6]), __builtin_bit_cast(float, r.v[7]));
    } else {
        double2 *q = (double2 *)p;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            q[i] = make_double2(__builtin_bit_cast(double, r.v[2 * i]), __builtin_bit_cast(double, r.v[2 * i + 1]));
    }
}

template <typename T> __device__ inline void store_row_nt(T *p, const Row8<T> &r)
{
    if constexpr (sizeof(T) == 4) {
        typedef float v4f __attribute__((ext_vector_type(4)));
        v4f a = {__builtin_bit_cast(float, r.v[0]), __builtin_bit_cast(float, r.v[1]),
                 __builtin_bit_cast(float, r.v[2]), __builtin_bit_cast(float, r.v[3])};
        v4f b = {__builtin_bit_cast(float, r.v[4]), __builtin_bit_cast(float, r.v[5]),
                 __builtin_bit_cast(float, r.v[6]), __builtin_bit_cast(float, r.v[7])};
        __builtin_nontemporal_store(a, (v4f *)p);
        __builtin_nontemporal_store(b, (v4f *)p + 1);
    } else {
        typedef double v2d __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v2d a = {__builtin_bit_cast(double, r.v[2 * i]), __builtin_bit_cast(double, r.v[2 * i + 1])};
            __builtin_nontemporal_store(a, (v2d *)p + i);
        }
    }
}

__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T> struct Half4 {
    T v[4];
};

// Per-lane inputs of one chunk for one property: own x-row, one half ghost row of the
// +-y / +-z faces (lane -> face = lane >> 4, row = (lane >> 1) & 7, half = lane & 1),
// and the two x-ghost scalars of the own row.
template <typename T> struct PropIn {
    Row8<T> own;
    Half4<T> gh;
    T xm, xp;
};

// Ghost half rows of the +-y / +-z faces: lane -> face = lane >> 4 (0: -y, 1: +y, 2: -z,
// 3: +z), row index idx = (lane >> 1) & 7 (z for the y faces, y for the z faces), half
// = lane & 1.  Source row inside the neighbour chunk (row = z*8 + y) and destination
// row of the LDS tile ((z+1)*10 + (y+1)), chosen without branches.
__device__ inline int ghost_src_row(int lane)
{
    const int face = lane >> 4, idx = (lane >> 1) & 7;
    const int ry = idx * 8 + ((face == 0) ? 7 : 0);
    const int rz = ((face == 2) ? 56 : 0) + idx;
    return (face < 2) ? ry : rz;
}

__device__ inline int ghost_tile_row(int lane)
{
    const int face = lane >> 4, idx = (lane >> 1) & 7;
    const int ty = (idx + 1) * 10 + ((face == 0) ? 0 : 9);
    const int tz = ((face == 2) ? 0 : 90) + idx + 1;
    return (face < 2) ? ty : tz;
}

template <typename T> __device__ inline void load_half(Half4<T> &h, const T *src)
{
    if constexpr (sizeof(T) == 4) {
        const float4 a = *(const float4 *)src;
        h.v[0] = __builtin_bit_cast(T, a.x); h.v[1] = __builtin_bit_cast(T, a.y);
        h.v[2] = __builtin_bit_cast(T, a.z); h.v[3] = __builtin_bit_cast(T, a.w);
    } else {
        const double2 a = ((const double2 *)src)[0], b = ((const double2 *)src)[1];
        h.v[0] = __builtin_bit_cast(T, a.x); h.v[1] = __builtin_bit_cast(T, a.y);
        h.v[2] = __builtin_bit_cast(T, b.x); h.v[3] = __builtin_bit_cast(T, b.y);
    }
}

// neighbour-dependent loads of one property: the face half row and (optionally) the
// two x-ghost scalars of the own row
template <typename T>
__device__ inline void issue_ghosts(PropIn<T> &r, const T *__restrict__ arr, const int64_t (&nbp)[6], int lane,
                                    bool load_xm, bool load_xp)
{
    const int face = lane >> 4;
    const int64_t nbc = (face == 0) ? nbp[2] : (face == 1) ? nbp[3] : (face == 2) ? nbp[4] : nbp[5];
    load_half(r.gh, arr + nbc * 512 + ghost_src_row(lane) * 8 + (lane & 1) * 4);
    if (load_xm) r.xm = arr[nbp[0] * 512 + lane * 8 + 7];
    if (load_xp) r.xp = arr[nbp[1] * 512 + lane * 8 + 0];
}

// the occupancy table row of a chunk (see sg_grid::d_occ_tab): -x / +x neighbour
// positions, the +-y / +-z neighbour this lane's ghost half row comes from, own row
// occupancy, face ghost-row occupancy, -x / +x neighbour row occupancy
struct OccRow {
    int64_t nxm, nxp, face_nb;
    uint64_t own, xm, xp;
    uint32_t ghost;
};

// c is uniform across the wave: every entry goes through readfirstlane, and the lane's
// face neighbour is selected among those values (a select among the fields or array
// elements of a row held in memory is folded into an indexed load from a stack copy of
// the row, whose wait stalls every load issued before it)
__device__ inline void read_occ(OccRow &o, const int32_t *__restrict__ tab, int64_t c, int lane)
{
    const int32_t *row = tab + c * 16;
    const int32_t nxm = __builtin_amdgcn_readfirstlane(row[0]), nxp = __builtin_amdgcn_readfirstlane(row[1]);
    const int32_t nym = __builtin_amdgcn_readfirstlane(row[2]), nyp = __builtin_amdgcn_readfirstlane(row[3]);
    const int32_t nzm = __builtin_amdgcn_readfirstlane(row[4]), nzp = __builtin_amdgcn_readfirstlane(row[5]);
    const uint32_t own0 = (uint32_t)__builtin_amdgcn_readfirstlane(row[6]);
    const uint32_t own1 = (uint32_t)__builtin_amdgcn_readfirstlane(row[7]);
    const uint32_t gh = (uint32_t)__builtin_amdgcn_readfirstlane(row[8]);
    const uint32_t xm0 = (uint32_t)__builtin_amdgcn_readfirstlane(row[9]);
    const uint32_t xm1 = (uint32_t)__builtin_amdgcn_readfirstlane(row[10]);
    const uint32_t xp0 = (uint32_t)__builtin_amdgcn_readfirstlane(row[11]);
    const uint32_t xp1 = (uint32_t)__builtin_amdgcn_readfirstlane(row[12]);
    const int face = lane >> 4;  // 0: -y, 1: +y, 2: -z, 3: +z
    o.nxm = nxm;
    o.nxp = nxp;
    o.face_nb = (face & 2) ? ((face & 1) ? nzp : nzm) : ((face & 1) ? nyp : nym);
    o.own = (uint64_t)own0 | ((uint64_t)own1 << 32);
    o.ghost = gh;
    o.xm = (uint64_t)xm0 | ((uint64_t)xm1 << 32);
    o.xp = (uint64_t)xp0 | ((uint64_t)xp1 << 32);
}

// as issue_ghosts, reading only source rows that hold a point (the others hold the
// background: library invariant)
template <typename T>
__device__ inline void issue_ghosts_occ(PropIn<T> &r, const T *__restrict__ arr, const OccRow &o, int lane, bool load_xm,
                                        bool load_xp, T bg)
{
    if ((o.ghost >> (lane >> 1)) & 1u) {
        load_half(r.gh, arr + o.face_nb * 512 + ghost_src_row(lane) * 8 + (lane & 1) * 4);
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) r.gh.v[i] = bg;
    }
    if (load_xm) r.xm = ((o.xm >> lane) & 1u) ? arr[o.nxm * 512 + lane * 8 + 7] : bg;
    if (load_xp) r.xp = ((o.xp >> lane) & 1u) ? arr[o.nxp * 512 + lane * 8 + 0] : bg;
}

// LDS tile per property: rows (z, y) with z, y in [-1, 8] -> row index (z+1)*10 + (y+1),
// 8 values per row.  Own rows and the 32 face rows are written; corner rows are unused.
template <typename T> __device__ inline void stage_lds(const PropIn<T> &r, T *tile, int lane)
{
    const int y = lane & 7, z = lane >> 3;
    store_row(tile + ((z + 1) * 10 + (y + 1)) * 8, r.own);
    T *dst = tile + ghost_tile_row(lane) * 8 + (lane & 1) * 4;
    if constexpr (sizeof(T) == 4) {
        *(float4 *)dst = make_float4(__builtin_bit_cast(float, r.gh.v[0]), __builtin_bit_cast(float, r.gh.v[1]),
                                     __builtin_bit_cast(float, r.gh.v[2]), __builtin_bit_cast(float, r.gh.v[3]));
    } else {
        ((double2 *)dst)[0] = make_double2(__builtin_bit_cast(double, r.gh.v[0]), __builtin_bit_cast(double, r.gh.v[1]));
        ((double2 *)dst)[1] = make_double2(__builtin_bit_cast(double, r.gh.v[2]), __builtin_bit_cast(double, r.gh.v[3]));
    }
}

// Laplacian pair sums at x of a row, in the reference's accumulation order
// (x pair, then y pair, then z pair; HeatStencil, SparseGridGpu_util_test.cuh:240-255)
template <typename T>
__device__ inline void lap_terms(const PropIn<T> &r, const T *tile, int lane, T (&pxy)[8], T (&py)[8], T (&pz)[8])
{
    const int y = lane & 7, z = lane >> 3;
    Row8<T> ym, yp, zm, zp;
    load_row(ym, tile + ((z + 1) * 10 + y) * 8);
    load_row(yp, tile + ((z + 1) * 10 + y + 2) * 8);
    load_row(zm, tile + (z * 10 + y + 1) * 8);
    load_row(zp, tile + ((z + 2) * 10 + y + 1) * 8);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const T xm = (x == 0) ? r.xm : r.own.v[x - 1];
        const T xp = (x == 7) ? r.xp : r.own.v[x + 1];
        pxy[x] = xm + xp;
        py[x] = ym.v[x] + yp.v[x];
        pz[x] = zm.v[x] + zp.v[x];
    }
}

template <typename T, int KIND> struct ChunkIn {
    static constexpr bool kTwo = (KIND == SG_ST_GRAY_SCOTT);
    PropIn<T> u;
    PropIn<T> v;
    uint64_t mrow;
};

template <typename T, int KIND>
__device__ inline void issue_own(ChunkIn<T, KIND> &in, const uint8_t *__restrict__ mask, const T *__restrict__ src0,
                                 const T *__restrict__ src1, int64_t c, int lane, const StencilArgs &sa)
{
    if (sa.lskip) {
        // a row without points holds the background (library invariant): not loaded
        in.mrow = *(const uint64_t *)(mask + c * 512 + lane * 8);
        if (in.mrow != 0) {
            load_row(in.u.own, src0 + c * 512 + lane * 8);
            if constexpr (ChunkIn<T, KIND>::kTwo) load_row(in.v.own, src1 + c * 512 + lane * 8);
        } else {
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                in.u.own.v[x] = (T)sa.bgs[0];
                if constexpr (ChunkIn<T, KIND>::kTwo) in.v.own.v[x] = (T)sa.bgs[1];
            }
        }
        return;
    }
    load_row(in.u.own, src0 + c * 512 + lane * 8);
    if constexpr (ChunkIn<T, KIND>::kTwo) load_row(in.v.own, src1 + c * 512 + lane * 8);
    in.mrow = *(const uint64_t *)(mask + c * 512 + lane * 8);
}

// bit x of b -> bit 8 x
__device__ inline uint64_t spread8(uint32_t b)
{
    uint64_t r = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x) r |= (uint64_t)((b >> x) & 1u) << (8 * x);
    return r;
}

// own rows and mask of the chunk where the occupancy table says the row holds a point
// (the others hold the background: library invariant).  Skip mode 4 takes the row's mask
// from the two 64-bit planes (2 x 8 bytes per row, shared by the 8 rows of a z-plane,
// instead of the 8 mask bytes): byte x = 1 for an interior point, 2 for any other
// non-zero mask, 0 for none -- the three cases the fast kernels distinguish.
template <typename T, int KIND>
__device__ inline void issue_own_occ(ChunkIn<T, KIND> &in, const uint8_t *__restrict__ mask, const T *__restrict__ src0,
                                     const T *__restrict__ src1, int64_t c, int lane, const StencilArgs &sa,
                                     uint64_t own)
{
    if ((own >> lane) & 1ull) {
        load_row(in.u.own, src0 + c * 512 + lane * 8);
        if constexpr (ChunkIn<T, KIND>::kTwo) load_row(in.v.own, src1 + c * 512 + lane * 8);
        if (sa.lskip == 4) {
            const int z = lane >> 3, y8 = 8 * (lane & 7);
            const uint32_t nz = (uint32_t)(sa.mplanes[c * 16 + z] >> y8) & 0xFFu;
            const uint32_t it = (uint32_t)(sa.mplanes[c * 16 + 8 + z] >> y8) & 0xFFu;
            in.mrow = spread8(it) | (spread8(nz & ~it) << 1);
        } else {
            in.mrow = *(const uint64_t *)(mask + c * 512 + lane * 8);
        }
    } else {
        in.mrow = 0;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            in.u.own.v[x] = (T)sa.bgs[0];
            if constexpr (ChunkIn<T, KIND>::kTwo) in.v.own.v[x] = (T)sa.bgs[1];
        }
    }
}

template <typename T, int KIND>
__device__ inline void read_nn(int64_t (&nbp)[6], const int32_t *__restrict__ nn, int64_t c)
{
    const int32_t *row = nn + c * 8;
#pragma unroll
    for (int k = 0; k < 6; ++k) nbp[k] = row[k];
}

template <typename T, int KIND>
__device__ inline void issue_chunk(ChunkIn<T, KIND> &in, const int32_t *__restrict__ nn, const uint8_t *__restrict__ mask,
                                   const T *__restrict__ src0, const T *__restrict__ src1, int64_t c, int lane,
                                   const StencilArgs &sa)
{
    if (sa.lskip >= 2) {
        // nn is the occupancy table: rows without points are not loaded (lskip 3: ghost
        // rows only, the own rows are loaded before the table row arrives)
        if (sa.lskip == 3) issue_own<T, KIND>(in, mask, src0, src1, c, lane, sa);
        __builtin_amdgcn_sched_barrier(0);
        OccRow o;
        read_occ(o, nn, c, lane);
        if (sa.lskip == 2 || sa.lskip == 4) issue_own_occ<T, KIND>(in, mask, src0, src1, c, lane, sa, o.own);
        issue_ghosts_occ<T>(in.u, src0, o, lane, true, true, (T)sa.bgs[0]);
        if constexpr (ChunkIn<T, KIND>::kTwo) issue_ghosts_occ<T>(in.v, src1, o, lane, true, true, (T)sa.bgs[1]);
        return;
    }
    // own rows and mask first: they do not depend on the neighbour table
    issue_own<T, KIND>(in, mask, src0, src1, c, lane, sa);
    __builtin_amdgcn_sched_barrier(0);
    int64_t nbp[6];
    read_nn<T, KIND>(nbp, nn, c);
    issue_ghosts<T>(in.u, src0, nbp, lane, true, true);
    if constexpr (ChunkIn<T, KIND>::kTwo) issue_ghosts<T>(in.v, src1, nbp, lane, true, true);
}

// ONE: the two Gray-Scott properties pass through one LDS tile (tile0) one after the
// other (half the LDS per wave, so more workgroups fit a CU); the u Laplacian is reduced
// to one value per point before v is staged.
template <typename T, int KIND, bool ONE = false>
__device__ inline void process_chunk(const ChunkIn<T, KIND> &in, int64_t c, T *tile0, T *tile1, T *__restrict__ dst0,
                                     T *__restrict__ dst1, int alias, const StencilArgs &sa,
                                     const int64_t *__restrict__ keys, int64_t bsz0, int64_t bsz1, int lane)
{
    constexpr bool kTwo = ChunkIn<T, KIND>::kTwo;
    stage_lds<T>(in.u, tile0, lane);
    if constexpr (kTwo && !ONE) stage_lds<T>(in.v, tile1, lane);
    wave_lds_sync();

    bool act[8];
    {
        bool inbox[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) inbox[x] = true;
        if (sa.use_box) {
            ChunkLin L;
            L.dim = 3;
            L.zorder = sa.zorder;
            L.bsz[0] = bsz0;
            L.bsz[1] = bsz1;
            L.bsz[2] = 1 << 30;
            int64_t bcd[3];
            chunk_coords(L, keys[c], bcd);
            const int64_t bx = bcd[0], by = bcd[1], bz = bcd[2];
            const int64_t py = by * 8 + (lane & 7), pz = bz * 8 + (lane >> 3);
            const bool yz = py >= sa.lo[1] && py <= sa.hi[1] && pz >= sa.lo[2] && pz <= sa.hi[2];
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                const int64_t px = bx * 8 + x;
                inbox[x] = yz && px >= sa.lo[0] && px <= sa.hi[0];
            }
        }
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const uint8_t cm = (uint8_t)(in.mrow >> (8 * x));
            const bool a = (sa.apply == SG_APPLY_EXISTING) ? (cm != 0) : ((cm & SG_EXIST) && !(cm & SG_PADDING));
            act[x] = a && inbox[x];
        }
    }

    T px[8], py[8], pz[8];
    lap_terms<T>(in.u, tile0, lane, px, py, pz);
    Row8<T> r0, r1;
    if constexpr (KIND == SG_ST_LAP) {
        const T c0 = (T)sa.coef[0], c1 = (T)sa.coef[1];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            T a = c0 * in.u.own.v[x];
            a = a + c1 * px[x];
            a = a + c1 * py[x];
            a = a + c1 * pz[x];
            r0.v[x] = a;
        }
    } else if constexpr (KIND == SG_ST_HEAT) {
        const T dt = (T)sa.coef[0];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            T a = (T)(-6.0) * in.u.own.v[x];
            a = a + px[x];
            a = a + py[x];
            a = a + pz[x];
            r0.v[x] = in.u.own.v[x] + dt * a;
        }
    } else if constexpr (ONE) {
        const T Du = (T)sa.coef[0], Dv = (T)sa.coef[1], F = (T)sa.coef[2], k = (T)sa.coef[3], dt = (T)sa.coef[4];
        T lu[8], lv[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            T a = (T)(-6.0) * in.u.own.v[x];
            a = a + px[x];
            a = a + py[x];
            lu[x] = a + pz[x];
        }
        wave_lds_sync();  // the u rows are read: the tile takes v
        stage_lds<T>(in.v, tile0, lane);
        wave_lds_sync();
        lap_terms<T>(in.v, tile0, lane, px, py, pz);
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            T a = (T)(-6.0) * in.v.own.v[x];
            a = a + px[x];
            a = a + py[x];
            lv[x] = a + pz[x];
        }
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const T cu = in.u.own.v[x], cv = in.v.own.v[x];
            const T uvv = (cu * cv) * cv;
            r0.v[x] = cu + dt * ((Du * lu[x] - uvv) + F * ((T)1 - cu));
            r1.v[x] = cv + dt * ((Dv * lv[x] + uvv) - (F + k) * cv);
        }
    } else {
        T qx[8], qy[8], qz[8];
        lap_terms<T>(in.v, tile1, lane, qx, qy, qz);
        const T Du = (T)sa.coef[0], Dv = (T)sa.coef[1], F = (T)sa.coef[2], k = (T)sa.coef[3], dt = (T)sa.coef[4];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const T cu = in.u.own.v[x], cv = in.v.own.v[x];
            T lu = (T)(-6.0) * cu;
            lu = lu + px[x];
            lu = lu + py[x];
            lu = lu + pz[x];
            T lv = (T)(-6.0) * cv;
            lv = lv + qx[x];
            lv = lv + qy[x];
            lv = lv + qz[x];
            const T uvv = (cu * cv) * cv;
            r0.v[x] = cu + dt * ((Du * lu - uvv) + F * ((T)1 - cu));
            r1.v[x] = cv + dt * ((Dv * lv + uvv) - (F + k) * cv);
        }
    }
    // the LDS tiles are rewritten by the next chunk of this wave
    wave_lds_sync();

    bool all = true, whole = true;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        all = all && act[x];
        // a slot without a point holds the destination's background (library invariant),
        // so rewriting it with the background changes nothing
        whole = whole && (act[x] || (uint8_t)(in.mrow >> (8 * x)) == 0);
    }
    if (sa.wmode == 3 && !alias) {
        // The rows go through this wave's LDS tile so that every store instruction writes
        // 1 KB of the chunk contiguously (16 bytes per lane), non-temporal.  Row flags:
        // 0 no point (nothing stored: the row holds the background), 1 whole row (absent
        // points take the background), 2 active points only (bits 8-15: which).
        uint32_t fl = 0;
        if (in.mrow != 0) {
            if (whole) {
                fl = 1;
                const T b0 = (T)sa.bgd[0], b1 = (T)sa.bgd[1];
#pragma unroll
                for (int x = 0; x < 8; ++x) {
                    r0.v[x] = act[x] ? r0.v[x] : b0;
                    if constexpr (kTwo) r1.v[x] = act[x] ? r1.v[x] : b1;
                }
            } else {
                fl = 2;
#pragma unroll
                for (int x = 0; x < 8; ++x) fl |= (uint32_t)act[x] << (8 + x);
            }
        }
        if constexpr (kTwo && ONE) {
            store_row(tile0 + lane * 8, r0);
            wave_lds_sync();
            store_chunk_contig(dst0 + c * 512, tile0, fl, lane);
            wave_lds_sync();
            store_row(tile0 + lane * 8, r1);
            wave_lds_sync();
            store_chunk_contig(dst1 + c * 512, tile0, fl, lane);
            wave_lds_sync();
            return;
        }
        store_row(tile0 + lane * 8, r0);
        if constexpr (kTwo) store_row(tile1 + lane * 8, r1);
        wave_lds_sync();
        store_chunk_contig(dst0 + c * 512, tile0, fl, lane);
        if constexpr (kTwo) store_chunk_contig(dst1 + c * 512, tile1, fl, lane);
        wave_lds_sync();  // the tiles are rewritten by the next chunk of this wave
        return;
    }
    T *o0 = dst0 + c * 512 + lane * 8;
    T *o1 = kTwo ? dst1 + c * 512 + lane * 8 : nullptr;
    if (sa.wmode != 0 && !alias && in.mrow == 0) {
        // no point in the row: the destination row already holds the background
    } else if (sa.wmode != 0 && !alias && whole) {
        const T b0 = (T)sa.bgd[0], b1 = (T)sa.bgd[1];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            r0.v[x] = act[x] ? r0.v[x] : b0;
            if constexpr (kTwo) r1.v[x] = act[x] ? r1.v[x] : b1;
        }
        if (sa.wmode == 2) {
            store_row_nt(o0, r0);
            if constexpr (kTwo) store_row_nt(o1, r1);
        } else {
            store_row(o0, r0);
            if constexpr (kTwo) store_row(o1, r1);
        }
    } else if (alias) {
        // destination replaces a source: every slot is written (old value where inactive)
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            r0.v[x] = act[x] ? r0.v[x] : in.u.own.v[x];
            if constexpr (kTwo) r1.v[x] = act[x] ? r1.v[x] : in.v.own.v[x];
        }
        store_row(o0, r0);
        if constexpr (kTwo) store_row(o1, r1);
    } else if (all) {
        store_row(o0, r0);
        if constexpr (kTwo) store_row(o1, r1);
    } else {
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            if (act[x]) {
                o0[x] = r0.v[x];
                if constexpr (kTwo) o1[x] = r1.v[x];
            }
        }
    }
}

// Persistent wave-per-chunk star stencil over 8^3 chunks (KIND: LAP, HEAT, GRAY_SCOTT).
// Workgroups are dealt to the 8 XCDs round robin, so workgroup b works in XCD b & 7:
// XCD x owns the contiguous range [x*per_xcd, (x+1)*per_xcd) of the sorted chunk list,
// and each of its waves walks `per_wave` consecutive chunks.  Neighbour chunks in x and
// y are then mostly read from the same XCD's L2.  The loads of the next chunk are
// issued before the current chunk is computed (software pipelining across chunks).
// MINW: minimum waves per SIMD the register allocation must allow (1 = unconstrained).
template <typename T, int KIND, int MINW>
__global__ __launch_bounds__(256, MINW) void stencil_fast_kernel(const int32_t *__restrict__ nn, int64_t n,
                                                                 const uint8_t *__restrict__ mask,
                                                                 const T *__restrict__ src0, const T *__restrict__ src1,
                                                                 T *__restrict__ dst0, T *__restrict__ dst1, int alias,
                                                                 StencilArgs sa, const int64_t *__restrict__ keys,
                                                                 int64_t bsz0, int64_t bsz1, int64_t per_xcd,
                                                                 int64_t per_wave)
{
    constexpr bool kTwo = (KIND == SG_ST_GRAY_SCOTT);
    __shared__ __attribute__((aligned(16))) T tiles0[4][100 * 8];
    __shared__ __attribute__((aligned(16))) T tiles1[4][kTwo ? 100 * 8 : 8];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int64_t lo = xcd * per_xcd;
    const int64_t hi = (lo + per_xcd < n) ? lo + per_xcd : n;
    const int64_t begin = lo + (slot * 4 + w) * per_wave;
    const int64_t end = (begin + per_wave < hi) ? begin + per_wave : hi;
    if (begin >= end) return;
    T *tile0 = tiles0[w];
    T *tile1 = tiles1[w];
    ChunkIn<T, KIND> cur, nxt;
    issue_chunk<T, KIND>(cur, nn, mask, src0, src1, begin, lane, sa);
    for (int64_t c = begin; c < end; ++c) {
        if (c + 1 < end) issue_chunk<T, KIND>(nxt, nn, mask, src0, src1, c + 1, lane, sa);
        process_chunk<T, KIND>(cur, c, tile0, tile1, dst0, dst1, alias, sa, keys, bsz0, bsz1, lane);
        cur = nxt;
    }
}

// One chunk per 64-lane workgroup (no pipelining, low register use, many waves);
// workgroup b handles chunk (b & 7) * per_xcd + (b >> 3) so each XCD walks a
// contiguous range of the sorted chunk list.
template <typename T, int KIND>
__global__ __launch_bounds__(64) void stencil_wave_kernel(const int32_t *__restrict__ nn, int64_t n,
                                                          const uint8_t *__restrict__ mask, const T *__restrict__ src0,
                                                          const T *__restrict__ src1, T *__restrict__ dst0,
                                                          T *__restrict__ dst1, int alias, StencilArgs sa,
                                                          const int64_t *__restrict__ keys, int64_t bsz0, int64_t bsz1,
                                                          int64_t per_xcd)
{
    constexpr bool kTwo = (KIND == SG_ST_GRAY_SCOTT);
    __shared__ __attribute__((aligned(16))) T tile0[100 * 8];
    __shared__ __attribute__((aligned(16))) T tile1[kTwo ? 100 * 8 : 8];
    const int64_t b = blockIdx.x;
    const int64_t c = (b & 7) * per_xcd + (b >> 3);
    if (c >= n) return;
    ChunkIn<T, KIND> cur;
    issue_chunk<T, KIND>(cur, nn, mask, src0, src1, c, threadIdx.x, sa);
    process_chunk<T, KIND>(cur, c, tile0, tile1, dst0, dst1, alias, sa, keys, bsz0, bsz1, threadIdx.x);
}

// W waves per workgroup on W consecutive chunks of the sorted list.  Consecutive chunks
// are usually x-neighbours, so each wave publishes its x = 0 and x = 7 columns in LDS
// and takes its x-ghosts from the neighbouring wave when the neighbour table says the
// -x / +x chunk is the previous / next chunk of the group; only the other x-ghosts are
// loaded from memory.  One group: chunks grp * W .. grp * W + W - 1.
template <typename T, int KIND, int W, bool ONE = false> struct GroupLds {
    static constexpr bool kTwo = (KIND == SG_ST_GRAY_SCOTT);
    static constexpr int NP = kTwo ? 2 : 1;
    __attribute__((aligned(16))) T tiles0[W][100 * 8];
    __attribute__((aligned(16))) T tiles1[W][(kTwo && !ONE) ? 100 * 8 : 8];
    T cols[W][NP][2][64];  // [wave][property][x = 0 | x = 7][row]
};

template <typename T, int KIND, int W, bool ONE>
__device__ inline void group_step(GroupLds<T, KIND, W, ONE> &L, int64_t grp, int w, int lane,
                                  const int32_t *__restrict__ nn, int64_t n, const uint8_t *__restrict__ mask,
                                  const T *__restrict__ src0, const T *__restrict__ src1, T *__restrict__ dst0,
                                  T *__restrict__ dst1, int alias, const StencilArgs &sa,
                                  const int64_t *__restrict__ keys, int64_t bsz0, int64_t bsz1)
{
    constexpr bool kTwo = GroupLds<T, KIND, W, ONE>::kTwo;
    constexpr int NP = GroupLds<T, KIND, W, ONE>::NP;
    const int64_t c = grp * W + w;
    const bool valid = c < n;  // every wave reaches the barrier
    ChunkIn<T, KIND> in;
    bool xm_grp = false, xp_grp = false;
    if (valid && sa.lskip >= 2) {
        // nn is the occupancy table: rows without points are not loaded (lskip 3: ghost
        // rows only, the own rows are loaded before the table row arrives)
        if (sa.lskip == 3) issue_own<T, KIND>(in, mask, src0, src1, c, lane, sa);
        __builtin_amdgcn_sched_barrier(0);
        OccRow o;
        read_occ(o, nn, c, lane);
        if (sa.lskip == 2 || sa.lskip == 4) issue_own_occ<T, KIND>(in, mask, src0, src1, c, lane, sa, o.own);
        xm_grp = (w > 0) && o.nxm == c - 1;
        xp_grp = (w < W - 1) && (c + 1 < n) && o.nxp == c + 1;
        issue_ghosts_occ<T>(in.u, src0, o, lane, !xm_grp, !xp_grp, (T)sa.bgs[0]);
        if constexpr (kTwo) issue_ghosts_occ<T>(in.v, src1, o, lane, !xm_grp, !xp_grp, (T)sa.bgs[1]);
        L.cols[w][0][0][lane] = in.u.own.v[0];
        L.cols[w][0][1][lane] = in.u.own.v[7];
        if constexpr (kTwo) {
            L.cols[w][NP - 1][0][lane] = in.v.own.v[0];
            L.cols[w][NP - 1][1][lane] = in.v.own.v[7];
        }
    } else if (valid) {
        issue_own<T, KIND>(in, mask, src0, src1, c, lane, sa);
        __builtin_amdgcn_sched_barrier(0);
        int64_t nbp[6];
        read_nn<T, KIND>(nbp, nn, c);
        xm_grp = (w > 0) && nbp[0] == c - 1;
        xp_grp = (w < W - 1) && (c + 1 < n) && nbp[1] == c + 1;
        issue_ghosts<T>(in.u, src0, nbp, lane, !xm_grp, !xp_grp);
        if constexpr (kTwo) issue_ghosts<T>(in.v, src1, nbp, lane, !xm_grp, !xp_grp);
        L.cols[w][0][0][lane] = in.u.own.v[0];
        L.cols[w][0][1][lane] = in.u.own.v[7];
        if constexpr (kTwo) {
            L.cols[w][NP - 1][0][lane] = in.v.own.v[0];
            L.cols[w][NP - 1][1][lane] = in.v.own.v[7];
        }
    }
    __syncthreads();
    if (!valid) return;
    if (xm_grp) {
        in.u.xm = L.cols[w - 1][0][1][lane];
        if constexpr (kTwo) in.v.xm = L.cols[w - 1][NP - 1][1][lane];
    }
    if (xp_grp) {
        in.u.xp = L.cols[w + 1][0][0][lane];
        if constexpr (kTwo) in.v.xp = L.cols[w + 1][NP - 1][0][lane];
    }
    process_chunk<T, KIND, ONE>(in, c, L.tiles0[w], L.tiles1[w], dst0, dst1, alias, sa, keys, bsz0, bsz1, lane);
}

// one group per workgroup; group g runs in XCD g's slot (g & 7 of the launch order)
template <typename T, int KIND, int W, bool ONE>
__global__ __launch_bounds__(64 * W) void stencil_group_kernel(const int32_t *__restrict__ nn, int64_t n,
                                                               const uint8_t *__restrict__ mask,
                                                               const T *__restrict__ src0, const T *__restrict__ src1,
                                                               T *__restrict__ dst0, T *__restrict__ dst1, int alias,
                                                               StencilArgs sa, const int64_t *__restrict__ keys,
                                                               int64_t bsz0, int64_t bsz1, int64_t groups_per_xcd)
{
    __shared__ GroupLds<T, KIND, W, ONE> L;
    // the wave index is uniform: with it in a scalar register the neighbour-table row is
    // read by scalar loads, which the ghost loads wait for without waiting for the own
    // rows (vector loads complete in order)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t b = blockIdx.x;
    const int64_t grp = (b & 7) * groups_per_xcd + (b >> 3);
    group_step<T, KIND, W, ONE>(L, grp, w, lane, nn, n, mask, src0, src1, dst0, dst1, alias, sa, keys, bsz0, bsz1);
}

// row occupancy of 8^3 chunks: bit r of occ[c] set when row r (z*8 + y) of chunk c has a
// slot with a non-zero mask; one wave per chunk, lane r reads its row's 8 mask bytes;
// occ[n] (the background chunk) is 0.  Also the chunk's two mask planes (StencilArgs::
// mplanes): the row's 8 bits go to bits 8y .. 8y + 7 of word z, the 8 lanes of one z
// combine theirs by shuffles.
__global__ __launch_bounds__(256) void row_occ_kernel(const uint8_t *__restrict__ mask, int64_t n,
                                                      uint64_t *__restrict__ occ, uint64_t *__restrict__ planes)
{
    const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (c > n) return;
    const uint64_t m = (c < n) ? *(const uint64_t *)(mask + c * 512 + lane * 8) : 0ull;
    const uint64_t b = __ballot(m != 0);
    if (lane == 0) occ[c] = b;
    uint32_t nz = 0, it = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const uint32_t cm = (uint32_t)(m >> (8 * x)) & 0xFFu;
        nz |= (uint32_t)(cm != 0) << x;
        it |= (uint32_t)((cm & SG_EXIST) && !(cm & SG_PADDING)) << x;
    }
    const int sh = 8 * (lane & 7);
    uint32_t w[4] = {sh < 32 ? nz << sh : 0u, sh >= 32 ? nz << (sh - 32) : 0u, sh < 32 ? it << sh : 0u,
                     sh >= 32 ? it << (sh - 32) : 0u};
#pragma unroll
    for (int s = 1; s < 8; s <<= 1)
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] |= (uint32_t)__shfl_xor((int)w[k], s, 64);
    if ((lane & 7) == 0) {
        planes[c * 16 + (lane >> 3)] = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
        planes[c * 16 + 8 + (lane >> 3)] = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
    }
}

// the occupancy table row of every chunk from the star neighbour table and the row
// occupancy (one thread per chunk)
__global__ void occ_table_kernel(const int32_t *__restrict__ nn, const uint64_t *__restrict__ occ, int64_t n,
                                 int32_t *__restrict__ tab)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int64_t nbp[6];
    for (int k = 0; k < 6; ++k) nbp[k] = nn[c * 8 + k];
    uint32_t ghost = 0;
    for (int j = 0; j < 32; ++j) {
        const int face = j >> 3;
        const uint64_t o = occ[nbp[2 + face]];
        ghost |= (uint32_t)((o >> ghost_src_row(2 * j)) & 1ull) << j;
    }
    const uint64_t own = occ[c], xm = occ[nbp[0]], xp = occ[nbp[1]];
    int32_t *row = tab + c * 16;
    for (int k = 0; k < 6; ++k) row[k] = (int32_t)nbp[k];
    row[6] = (int32_t)(uint32_t)own;
    row[7] = (int32_t)(uint32_t)(own >> 32);
    row[8] = (int32_t)ghost;
    row[9] = (int32_t)(uint32_t)xm;
    row[10] = (int32_t)(uint32_t)(xm >> 32);
    row[11] = (int32_t)(uint32_t)xp;
    row[12] = (int32_t)(uint32_t)(xp >> 32);
    row[13] = row[14] = row[15] = 0;
}

template <typename T>
__global__ void get_points_kernel(const int64_t *keys, int64_t n, const uint8_t *mask, const T *arr, int csz,
                                  const int64_t *pk, int64_t m, T *out, T bgv)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t key = pk[i];
    const int64_t pos = find_chunk(keys, n, key / csz);
    const int64_t slot = pos * csz + key % csz;
    out[i] = (pos < n && (mask[slot] & 1)) ? arr[slot] : bgv;
}

template <typename T> __global__ void fill_range_kernel(T *p, int64_t n, T v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename T> __global__ void normalize_kernel(T *p, const uint8_t *mask, int64_t n, T v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && mask[i] == 0) p[i] = v;
}

// countExistingElements (which 0: EXIST bits) / countBoundaryElements (which 1: EXIST and
// PADDING), SparseGridGpu.hpp:2422-2470: 16 mask bytes per load, four bytes per popcount
__device__ inline uint32_t count_word(uint32_t w, int which)
{
    return (uint32_t)__popc((which == 0 ? w : (w & (w >> 1))) & 0x01010101u);
}

__global__ __launch_bounds__(kWG) void count_mask_kernel(const uint8_t *mask, int64_t n, int which,
                                                        unsigned long long *out)
{
    __shared__ unsigned long long s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    unsigned long long c = 0;
    const int64_t nv = n >> 4, stride = (int64_t)gridDim.x * blockDim.x;
    const uint4 *v = (const uint4 *)mask;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        const uint4 q = v[i];
        c += count_word(q.x, which) + count_word(q.y, which) + count_word(q.z, which) + count_word(q.w, which);
    }
    for (int64_t i = (nv << 4) + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint8_t m = mask[i];
        c += (which == 0) ? (m & 1) : ((m & 3) == 3);
    }
    // one shared atomic per wave
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0 && s) atomicAdd(out, s);
}

// points with EXIST from the 8^3 EXIST bit-planes (nw 64-bit words): 64 B per chunk
// instead of 512
__global__ __launch_bounds__(kWG) void count_planes_kernel(const uint64_t *eb, int64_t nw, unsigned long long *out)
{
    __shared__ unsigned long long s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    unsigned long long c = 0;
    // 8 consecutive words (one chunk) per thread and step, loaded before they are counted
    const int64_t stride = (int64_t)gridDim.x * blockDim.x * 8;
    for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8; i < nw; i += stride) {
        uint64_t w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] = (i + k < nw) ? eb[i + k] : 0ull;
#pragma unroll
        for (int k = 0; k < 8; ++k) c += __popcll(w[k]);
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s, c);
    __syncthreads();
    if (threadIdx.x == 0 && s) atomicAdd(out, s);
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

GeomArgs geom(const sg_grid *g)
{
    GeomArgs ga;
    ga.dim = g->dim;
    ga.edge = g->edge;
    ga.csz = g->csz;
    ga.bsz0 = g->bsz[0];
    ga.bsz1 = g->bsz[1];
    ga.bsz2 = g->bsz[2];
    ga.zorder = g->lin == SG_LIN_ZMB;
    return ga;
}

int ensure_nn(sg_grid *g, int type)
{
    if (g->nn_valid[type]) return SG_OK;
    return find_neighbours_impl(g, type);
}

}  // namespace

int find_neighbours_impl(sg_grid *g, int type)
{
    const int nnn = g->nnn(type);
    const int64_t n = g->nchunks;
    if (n * nnn > g->nn_cap[type]) {
        g->pool.free(g->d_nn[type]);
        int64_t cap = std::max<int64_t>(n * nnn, 64);
        g->d_nn[type] = (int32_t *)g->pool.alloc(cap * sizeof(int32_t));
        if (!g->d_nn[type]) return fail(SG_E_OOM, "neighbour table allocation failed");
        g->nn_cap[type] = cap;
    }
    if (n > 0) {
        // a table of chunk positions by id when the domain's id range is not much larger
        // than the chunk count (one read per lookup instead of a binary search)
        int64_t last[3] = {0, 0, 0};
        for (int d = 0; d < g->dim; ++d) last[d] = g->bsz[d] - 1;
        const int64_t idmax = chunk_id(chunk_lin(g), last) + 1;
        int32_t *tab = nullptr;
        if (idmax <= ((int64_t)1 << 26) && idmax <= 64 * std::max<int64_t>(n, 65536))
            tab = (int32_t *)g->pool.alloc((size_t)idmax * sizeof(int32_t));
        if (tab) {
            hipLaunchKernelGGL(fill_range_kernel<int32_t>, dim3(grid_1d(idmax, kWG)), dim3(kWG), 0, g->stream, tab,
                               idmax, (int32_t)-1);
            hipLaunchKernelGGL(chunk_table_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_keys, n, tab);
            hipLaunchKernelGGL(find_neighbours_kernel<true>, dim3(grid_1d(n * nnn, kWG)), dim3(kWG), 0, g->stream,
                               g->d_keys, n, geom(g), type, nnn, tab, idmax, g->d_nn[type]);
            SG_HIP(hipGetLastError());
            g->pool.free(tab);  // stream order: reused only after these kernels
        } else {
            hipLaunchKernelGGL(find_neighbours_kernel<false>, dim3(grid_1d(n * nnn, kWG)), dim3(kWG), 0, g->stream,
                               g->d_keys, n, geom(g), type, nnn, (const int32_t *)nullptr, (int64_t)0,
                               g->d_nn[type]);
            SG_HIP(hipGetLastError());
        }
    }
    g->nn_valid[type] = true;
    return SG_OK;
}

int tag_boundaries_impl(sg_grid *g, int type)
{
    const int64_t n = g->nchunks;
    if (n == 0) return SG_OK;
    // PADDING bits change: the stencils' mask planes (EXIST without PADDING) go stale with
    // the occupancy table they are built with; the EXIST bit-planes stay valid
    g->occ_tab_valid = false;
    int rc = ensure_nn(g, type);
    if (rc) return rc;
    if (g->dim == 3 && g->edge == 8 && g->tag_variant == 0) {
        // other waves read only the bit-planes, which tagging (PADDING bits) leaves valid
        if ((rc = ebits_impl(g))) return rc;
        const int64_t per_xcd = ((n + 23) / 24 + 7) / 8;
        const unsigned grid = (unsigned)(per_xcd * 8);
        if (type == SG_NN_FULL)
            hipLaunchKernelGGL((tag_planes_kernel<SG_NN_FULL>), dim3(grid), dim3(256), 0, g->stream, n, g->d_nn[type],
                               g->nnn(type), g->d_ebits, g->d_mask, per_xcd);
        else
            hipLaunchKernelGGL((tag_planes_kernel<SG_NN_STAR>), dim3(grid), dim3(256), 0, g->stream, n, g->d_nn[type],
                               g->nnn(type), g->d_ebits, g->d_mask, per_xcd);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    if (g->dim == 3 && g->edge == 8) {
        // the tile kernel writes masks in place: rows are read by other waves only through
        // ghost words (EXIST bits), which this kernel never changes
        auto launch = [&](auto cpw) {
            constexpr int CPW = decltype(cpw)::value;
            const int64_t per_wg = 4 * CPW, per_xcd = ((n + per_wg - 1) / per_wg + 7) / 8;
            hipLaunchKernelGGL((tag_boundaries8_kernel<4, CPW>), dim3((unsigned)(per_xcd * 8)), dim3(256), 0,
                               g->stream, n, g->d_nn[type], g->nnn(type), type, g->d_mask, per_xcd);
        };
        if (g->tag_cpw == 1) launch(std::integral_constant<int, 1>());
        else if (g->tag_cpw == 2) launch(std::integral_constant<int, 2>());
        else launch(std::integral_constant<int, 4>());
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    const int e = g->edge + 2;
    const size_t ebs = (g->dim == 3) ? (size_t)e * e * e : (size_t)e * e;
    const int wg = std::max(64, std::min(kWG, (g->csz + 63) / 64 * 64));  // one thread per slot up to kWG
    hipLaunchKernelGGL(tag_boundaries_kernel, dim3((unsigned)n), dim3(wg), ebs, g->stream, n, geom(g),
                       g->d_nn[type], g->nnn(type), type, g->d_mask);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int existing_points_impl(sg_grid *g)
{
    const int64_t n = g->nchunks;
    g->ep_n = 0;
    g->ep_valid = false;
    if (n == 0) {
        g->ep_valid = true;
        return SG_OK;
    }
    DevBuf &b = g->scratch[9];
    int rc = ensure(g->pool, b, (size_t)(n + 1) * 8 + 64);
    if (rc) return rc;
    uint32_t *cnt = (uint32_t *)b.p, *start = cnt + (n + 1);
    hipLaunchKernelGGL(count_exist_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_mask, g->csz, cnt);
    SG_HIP(hipGetLastError());
    uint32_t tot = 0;
    if ((rc = exclusive_scan_u32(g, cnt, start, n, &tot))) return rc;
    if ((int64_t)tot > g->ep_cap) {
        g->pool.free(g->d_ep);
        g->d_ep = (int64_t *)g->pool.alloc((size_t)tot * 8);
        if (!g->d_ep) {
            g->ep_cap = 0;
            return fail(SG_E_OOM, "existing-point list allocation failed");
        }
        g->ep_cap = tot;
    }
    hipLaunchKernelGGL(fill_exist_kernel, dim3((unsigned)n), dim3(kWG), 0, g->stream, g->d_mask, g->csz, start, g->d_ep);
    SG_HIP(hipGetLastError());
    g->ep_n = tot;
    g->ep_valid = true;
    return SG_OK;
}

int apply_stencil_impl(sg_grid *g, const sg_stencil_desc *d)
{
    const int kind = d->kind;
    if (kind < SG_ST_LAP || kind > SG_ST_RESIDUAL) return fail(SG_E_ARG, "unknown stencil kind");
    const int nsrc = (kind == SG_ST_GRAY_SCOTT || kind == SG_ST_RBSOR || kind == SG_ST_RESIDUAL) ? 2 : 1;
    const int ndst = (kind == SG_ST_GRAY_SCOTT) ? 2 : (kind == SG_ST_RBSOR ? 0 : 1);
    for (int q = 0; q < nsrc; ++q)
        if (d->src[q] < 0 || d->src[q] >= g->np) return fail(SG_E_ARG, "bad source property");
    for (int q = 0; q < ndst; ++q)
        if (d->dst[q] < 0 || d->dst[q] >= g->np) return fail(SG_E_ARG, "bad destination property");
    if (ndst == 2 && d->dst[0] == d->dst[1]) return fail(SG_E_ARG, "the two destinations must differ");
    const int dt = g->dt[d->src[0]];
    for (int q = 0; q < nsrc; ++q)
        if (g->dt[d->src[q]] != dt) return fail(SG_E_ARG, "stencil properties must share one type");
    for (int q = 0; q < ndst; ++q)
        if (g->dt[d->dst[q]] != dt) return fail(SG_E_ARG, "stencil properties must share one type");
    if (kind != SG_ST_BOX && !dtype_is_float(dt)) return fail(SG_E_ARG, "this stencil needs f32 or f64 properties");
    if (d->apply != SG_APPLY_INTERIOR && d->apply != SG_APPLY_EXISTING) return fail(SG_E_ARG, "bad apply mode");
    const bool no_shared = d->mode == SG_STENCIL_MODE_INPLACE_NO_SHARED;
    if (!no_shared && d->mode != 0 && d->mode != SG_STENCIL_MODE_INPLACE) return fail(SG_E_ARG, "bad stencil mode");
    const int64_t n = g->nchunks;
    if (n == 0) return SG_OK;
    const int type = (kind == SG_ST_BOX) ? SG_NN_FULL : SG_NN_STAR;
    int rc = ensure_nn(g, type);
    if (rc) return rc;

    StencilArgs sa;
    sa.kind = kind;
    sa.apply = d->apply;
    sa.use_box = d->use_box;
    sa.wmode = g->fast_wmode;
    sa.bgd[0] = sa.bgd[1] = 0.0;
    sa.bgs[0] = sa.bgs[1] = 0.0;
    sa.lskip = g->fast_lskip;
    sa.mplanes = nullptr;
    sa.zorder = g->lin == SG_LIN_ZMB;
    for (int q = 0; q < nsrc && q < 2; ++q) {
        if (g->dt[d->src[q]] == SG_F32) {
            float f;
            memcpy(&f, g->bg[d->src[q]], sizeof(f));
            sa.bgs[q] = f;
        } else if (g->dt[d->src[q]] == SG_F64) {
            memcpy(&sa.bgs[q], g->bg[d->src[q]], sizeof(double));
        }
    }
    for (int q = 0; q < ndst && q < 2; ++q) {
        if (g->dt[d->dst[q]] == SG_F32) {
            float f;
            memcpy(&f, g->bg[d->dst[q]], sizeof(f));
            sa.bgd[q] = f;
        } else if (g->dt[d->dst[q]] == SG_F64) {
            memcpy(&sa.bgd[q], g->bg[d->dst[q]], sizeof(double));
        }
    }
    for (int i = 0; i < 3; ++i) {
        sa.lo[i] = d->box_lo[i];
        sa.hi[i] = d->box_hi[i];
    }
    for (int i = 0; i < 27; ++i) {
        sa.coef[i] = d->coef[i];
        sa.icoef[i] = d->icoef[i];
    }
    // a destination that is also a source is computed into a fresh array that then
    // replaces the property (every chunk reads the pre-sweep state)
    int alias[2] = {0, 0};
    for (int q = 0; q < ndst; ++q)
        for (int s = 0; s < nsrc; ++s)
            if (d->dst[q] == d->src[s]) alias[q] = s + 1;
    void *outp[2] = {nullptr, nullptr};
    const size_t arr_bytes = (size_t)(g->cap + 1) * g->csz * g->ts[d->src[0]];
    for (int q = 0; q < ndst; ++q) {
        if (alias[q]) {
            outp[q] = g->pool.alloc(arr_bytes);
            if (!outp[q]) return fail(SG_E_OOM, "stencil output allocation failed");
        } else {
            outp[q] = g->d_prop[d->dst[q]];
        }
    }
    // the fast kernel writes old values of dst q from src q, so it takes dst q
    // aliasing src q (or no aliasing) only
    const bool alias_ok = (ndst == 1) ? (alias[0] == 0 || alias[0] == 1)
                                      : ((alias[0] == 0 && alias[1] == 0) || (alias[0] == 1 && alias[1] == 2));
    if (no_shared) {
        // the existing points (built by tagBoundaries with CALCULATE_EXISTING_POINTS, or
        // here when stale); aliased outputs start as copies of the old arrays, since only
        // the listed active points are written
        if (!g->ep_valid && (rc = existing_points_impl(g))) return rc;
        for (int q = 0; q < ndst; ++q)
            if (alias[q])
                SG_HIP(hipMemcpyAsync(outp[q], g->d_prop[d->dst[q]], arr_bytes, hipMemcpyDeviceToDevice, g->stream));
        if (g->ep_n > 0) {
            dispatch_type(dt, [&](auto tag) {
                typedef decltype(tag) T;
                const T *s0 = (const T *)g->d_prop[d->src[0]];
                const T *s1 = nsrc > 1 ? (const T *)g->d_prop[d->src[1]] : nullptr;
                T *o0 = (T *)(ndst > 0 ? outp[0] : (kind == SG_ST_RBSOR ? g->d_prop[d->src[0]] : nullptr));
                T *o1 = (T *)(ndst > 1 ? outp[1] : nullptr);
                hipLaunchKernelGGL(stencil_points_kernel<T>, dim3(grid_1d(g->ep_n, kWG)), dim3(kWG), 0, g->stream,
                                   g->d_ep, g->ep_n, g->d_keys, geom(g), g->d_nn[type], g->nnn(type), type, g->d_mask,
                                   s0, s1, o0, o1, sa);
            });
            SG_HIP(hipGetLastError());
        }
        for (int q = 0; q < ndst; ++q) {
            if (!alias[q]) continue;
            const int p = d->dst[q];
            g->pool.free(g->d_prop[p]);
            g->d_prop[p] = outp[q];
        }
        return SG_OK;
    }
    if (kind == SG_ST_BOX && g->dim == 3 && g->edge == 8 && alias[0] == 0) {
        bool done = false;
        dispatch_type(dt, [&](auto tag) {
            typedef decltype(tag) T;
            if constexpr (sizeof(T) >= 4) {
                T bgd;
                memcpy(&bgd, g->bg[d->dst[0]], sizeof(T));
                // integer boxes whose 27 coefficients are equal (as T) sum first, multiply once
                bool uni = std::is_integral<T>::value;
                for (int q = 1; q < 27 && uni; ++q) uni = (T)d->icoef[q] == (T)d->icoef[0];
                auto launch = [&](auto u) {
                    constexpr bool U = decltype(u)::value;
                    if (g->box_pairs)
                        hipLaunchKernelGGL((box8_pair_kernel<T, 4, U>), dim3((unsigned)((n + 7) / 8)), dim3(256), 0,
                                           g->stream, n, g->d_nn[type], g->d_mask, (const T *)g->d_prop[d->src[0]],
                                           (T *)g->d_prop[d->dst[0]], g->d_keys, g->bsz[0], g->bsz[1], bgd, sa);
                    else
                        hipLaunchKernelGGL((box8_kernel<T, 4, U>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0,
                                           g->stream, n, g->d_nn[type], g->d_mask, (const T *)g->d_prop[d->src[0]],
                                           (T *)g->d_prop[d->dst[0]], g->d_keys, g->bsz[0], g->bsz[1], bgd, sa);
                };
                if (uni) launch(std::true_type());
                else launch(std::false_type());
                done = true;
            }
        });
        if (done) {
            SG_HIP(hipGetLastError());
            return SG_OK;
        }
    }
    const bool fast = g->dim == 3 && g->edge == 8 && type == SG_NN_STAR &&
                      (kind == SG_ST_LAP || kind == SG_ST_HEAT || kind == SG_ST_GRAY_SCOTT) && alias_ok;
    // the fast kernels take the star table, or with lskip >= 2 the occupancy table
    const int32_t *fnn = g->d_nn[type];
    if (fast && g->fast_lskip >= 2) {
        if ((rc = occ_table_impl(g))) return rc;
        fnn = g->d_occ_tab;
        sa.mplanes = g->d_mplanes;
    }
    dispatch_type(dt, [&](auto tag) {
        typedef decltype(tag) T;
        const T *s0 = (const T *)g->d_prop[d->src[0]];
        const T *s1 = nsrc > 1 ? (const T *)g->d_prop[d->src[1]] : nullptr;
        T *o0 = (T *)(ndst > 0 ? outp[0] : (kind == SG_ST_RBSOR ? g->d_prop[d->src[0]] : nullptr));
        T *o1 = (T *)(ndst > 1 ? outp[1] : nullptr);
        if constexpr (std::is_same<T, float>::value || std::is_same<T, double>::value) {
            if (fast) {
                const int64_t per_xcd = (n + 7) / 8;
                const int al = alias[0] != 0;
                auto launch = [&](auto kind_tag) {
                    constexpr int K = decltype(kind_tag)::value;
                    if (g->fast_variant == 0) {
                        // one chunk per 64-lane workgroup
                        hipLaunchKernelGGL((stencil_wave_kernel<T, K>), dim3((unsigned)(per_xcd * 8)), dim3(64), 0,
                                           g->stream, fnn, n, g->d_mask, s0, s1, o0, o1, al, sa, g->d_keys,
                                           g->bsz[0], g->bsz[1], per_xcd);
                        return;
                    }
                    if (g->fast_variant >= 3) {
                        // groups of W consecutive chunks sharing x faces through LDS
                        // (variants 6 / 7 / 8: the same with one LDS tile per wave for both
                        // Gray-Scott properties)
                        auto grouped = [&](auto wtag, auto onetag) {
                            constexpr int GW = decltype(wtag)::value;
                            constexpr bool ONE = decltype(onetag)::value;
                            const int64_t ngroups = (n + GW - 1) / GW;
                            const int64_t gpx = (ngroups + 7) / 8;
                            hipLaunchKernelGGL((stencil_group_kernel<T, K, GW, ONE>), dim3((unsigned)(gpx * 8)),
                                               dim3(64 * GW), 0, g->stream, fnn, n, g->d_mask, s0, s1, o0,
                                               o1, al, sa, g->d_keys, g->bsz[0], g->bsz[1], gpx);
                        };
                        const std::false_type two_tiles;
                        const std::true_type one_tile;
                        switch (g->fast_variant) {
                        case 3: grouped(std::integral_constant<int, 4>(), two_tiles); break;
                        case 4: grouped(std::integral_constant<int, 8>(), two_tiles); break;
                        case 5: grouped(std::integral_constant<int, 2>(), two_tiles); break;
                        case 6: grouped(std::integral_constant<int, 4>(), one_tile); break;
                        case 7: grouped(std::integral_constant<int, 8>(), one_tile); break;
                        default: grouped(std::integral_constant<int, 2>(), one_tile); break;
                        }
                        return;
                    }
                    // persistent waves: wg_per_xcd workgroups of 4 waves per XCD, each wave
                    // walking per_wave consecutive chunks
                    const int64_t wg_per_xcd =
                        std::max<int64_t>(1, std::min<int64_t>(g->fast_wg_per_xcd, (per_xcd + 3) / 4));
                    const int64_t per_wave = (per_xcd + wg_per_xcd * 4 - 1) / (wg_per_xcd * 4);
                    const unsigned nblk = (unsigned)(wg_per_xcd * 8);
                    if (g->fast_variant == 1)
                        hipLaunchKernelGGL((stencil_fast_kernel<T, K, 1>), dim3(nblk), dim3(256), 0, g->stream,
                                           fnn, n, g->d_mask, s0, s1, o0, o1, al, sa, g->d_keys, g->bsz[0],
                                           g->bsz[1], per_xcd, per_wave);
                    else
                        hipLaunchKernelGGL((stencil_fast_kernel<T, K, 4>), dim3(nblk), dim3(256), 0, g->stream,
                                           fnn, n, g->d_mask, s0, s1, o0, o1, al, sa, g->d_keys, g->bsz[0],
                                           g->bsz[1], per_xcd, per_wave);
                };
                if (kind == SG_ST_LAP) launch(std::integral_constant<int, SG_ST_LAP>());
                else if (kind == SG_ST_HEAT) launch(std::integral_constant<int, SG_ST_HEAT>());
                else launch(std::integral_constant<int, SG_ST_GRAY_SCOTT>());
                return;
            }
        }
        T bg0, bg1;
        memcpy(&bg0, g->bg[d->src[0]], sizeof(T));
        memcpy(&bg1, g->bg[nsrc > 1 ? d->src[1] : d->src[0]], sizeof(T));
        const int e = g->edge + 2;
        const size_t ebs = (g->dim == 3) ? (size_t)e * e * e : (size_t)e * e;
        // one thread per slot up to kWG (small chunks leave no idle waves)
        const int wg = std::max(64, std::min(kWG, (g->csz + 63) / 64 * 64));
        hipLaunchKernelGGL(stencil_generic_kernel<T>, dim3((unsigned)n), dim3(wg), ebs * sizeof(T) * nsrc, g->stream,
                           g->d_keys, n, geom(g), g->d_nn[type], g->nnn(type), type, g->d_mask, s0, s1, o0, o1,
                           alias[0], alias[1], bg0, bg1, sa);
    });
    SG_HIP(hipGetLastError());
    for (int q = 0; q < ndst; ++q) {
        if (!alias[q]) continue;
        const int p = d->dst[q];
        g->pool.free(g->d_prop[p]);
        g->d_prop[p] = outp[q];
        rc = fill_background_chunk(g, p);
        if (rc) return rc;
    }
    return SG_OK;
}

// EXIST bit-planes of 8^3 chunks (rebuilt when stale)
int ebits_impl(sg_grid *g)
{
    if (g->ebits_valid) return SG_OK;
    const int64_t n = g->nchunks;
    if (n + 1 > g->ebits_cap || !g->d_ebits) {
        g->pool.free(g->d_ebits);
        const int64_t cap = n + 1;
        g->d_ebits = (uint64_t *)g->pool.alloc((size_t)cap * 8 * sizeof(uint64_t));
        if (!g->d_ebits) {
            g->ebits_cap = 0;
            return fail(SG_E_OOM, "EXIST bit-plane allocation failed");
        }
        g->ebits_cap = cap;
    }
    hipLaunchKernelGGL(ebits_kernel, dim3(grid_1d((n + 1) * 64, 256 * kEbitsRows)), dim3(256), 0, g->stream, g->d_mask,
                       n, g->d_ebits);
    SG_HIP(hipGetLastError());
    g->ebits_valid = true;
    return SG_OK;
}

// occupancy table of the 8^3 star stencils (rebuilt when stale)
int occ_table_impl(sg_grid *g)
{
    if (g->occ_tab_valid) return SG_OK;
    int rc;
    if (!g->nn_valid[SG_NN_STAR] && (rc = find_neighbours_impl(g, SG_NN_STAR))) return rc;
    const int64_t n = g->nchunks;
    if (n > g->occ_tab_cap || !g->d_occ_tab) {
        g->pool.free(g->d_occ_tab);
        const int64_t cap = std::max<int64_t>(n, 1);
        g->d_occ_tab = (int32_t *)g->pool.alloc((size_t)cap * 16 * sizeof(int32_t));
        if (!g->d_occ_tab) {
            g->occ_tab_cap = 0;
            return fail(SG_E_OOM, "occupancy table allocation failed");
        }
        g->occ_tab_cap = cap;
    }
    if (n + 1 > g->mplanes_cap || !g->d_mplanes) {
        g->pool.free(g->d_mplanes);
        g->d_mplanes = (uint64_t *)g->pool.alloc((size_t)(n + 1) * 16 * sizeof(uint64_t));
        if (!g->d_mplanes) {
            g->mplanes_cap = 0;
            return fail(SG_E_OOM, "mask plane allocation failed");
        }
        g->mplanes_cap = n + 1;
    }
    DevBuf &b = g->scratch[12];
    if ((rc = ensure(g->pool, b, (size_t)(n + 1) * sizeof(uint64_t)))) return rc;
    uint64_t *occ = (uint64_t *)b.p;
    hipLaunchKernelGGL(row_occ_kernel, dim3((unsigned)(n / 4 + 1)), dim3(256), 0, g->stream, g->d_mask, n, occ,
                       g->d_mplanes);
    SG_HIP(hipGetLastError());
    if (n > 0) {
        hipLaunchKernelGGL(occ_table_kernel, dim3(grid_1d(n, kWG)), dim3(kWG), 0, g->stream, g->d_nn[SG_NN_STAR], occ,
                           n, g->d_occ_tab);
        SG_HIP(hipGetLastError());
    }
    g->occ_tab_valid = true;
    return SG_OK;
}

int fill_background_chunk(sg_grid *g, int p)
{
    int rc = SG_OK;
    dispatch_type(g->dt[p], [&](auto tag) {
        typedef decltype(tag) T;
        T v;
        memcpy(&v, g->bg[p], sizeof(T));
        T *base = (T *)g->d_prop[p] + g->nchunks * g->csz;
        hipLaunchKernelGGL(fill_range_kernel<T>, dim3(grid_1d(g->csz, kWG)), dim3(kWG), 0, g->stream, base,
                           (int64_t)g->csz, v);
    });
    SG_HIP(hipGetLastError());
    return rc;
}

int normalize_prop(sg_grid *g, int p)
{
    const int64_t total = (g->nchunks + 1) * g->csz;
    dispatch_type(g->dt[p], [&](auto tag) {
        typedef decltype(tag) T;
        T v;
        memcpy(&v, g->bg[p], sizeof(T));
        hipLaunchKernelGGL(normalize_kernel<T>, dim3(grid_1d(total, kWG)), dim3(kWG), 0, g->stream, (T *)g->d_prop[p],
                           g->d_mask, total, v);
    });
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int get_points_impl(sg_grid *g, int prop, int64_t m, const int64_t *keys, void *out, int mem)
{
    if (m <= 0) return SG_OK;
    const size_t ts = g->ts[prop];
    DevBuf &b = g->scratch[9];
    int rc = ensure(g->pool, b, (size_t)m * (8 + ts) + 64);
    if (rc) return rc;
    const int64_t *dk = keys;
    void *dout = out;
    if (mem == SG_MEM_HOST) {
        SG_HIP(hipMemcpyAsync(b.p, keys, m * sizeof(int64_t), hipMemcpyHostToDevice, g->stream));
        dk = (const int64_t *)b.p;
        dout = (char *)b.p + m * sizeof(int64_t);
    }
    dispatch_type(g->dt[prop], [&](auto tag) {
        typedef decltype(tag) T;
        T v;
        memcpy(&v, g->bg[prop], sizeof(T));
        hipLaunchKernelGGL(get_points_kernel<T>, dim3(grid_1d(m, kWG)), dim3(kWG), 0, g->stream, g->d_keys,
                           g->nchunks, g->d_mask, (const T *)g->d_prop[prop], g->csz, dk, m, (T *)dout, v);
    });
    SG_HIP(hipGetLastError());
    if (mem == SG_MEM_HOST) {
        SG_HIP(hipMemcpyAsync(out, dout, m * ts, hipMemcpyDeviceToHost, g->stream));
        SG_HIP(hipStreamSynchronize(g->stream));
    }
    return SG_OK;
}

int count_mask_impl(sg_grid *g, int which, int64_t *out)
{
    DevBuf &b = g->scratch[9];
    int rc = ensure(g->pool, b, 64);
    if (rc) return rc;
    unsigned long long *d = (unsigned long long *)b.p;
    SG_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), g->stream));
    const int64_t total = g->nchunks * g->csz;
    if (total > 0 && which == 0 && g->dim == 3 && g->edge == 8 && g->ebits_valid) {
        // existing points: the EXIST bit-planes the flush keeps current
        const int64_t nw = g->nchunks * 8;
        // one workgroup per CU: each adds its count to the total by one global atomic
        unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_1d(nw, kWG * 8), 256));
        hipLaunchKernelGGL(count_planes_kernel, dim3(blocks), dim3(kWG), 0, g->stream, (const uint64_t *)g->d_ebits, nw,
                           d);
        SG_HIP(hipGetLastError());
    } else if (total > 0) {
        unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(grid_1d(total / 16, kWG), 256));
        hipLaunchKernelGGL(count_mask_kernel, dim3(blocks), dim3(kWG), 0, g->stream, g->d_mask, total, which, d);
        SG_HIP(hipGetLastError());
    }
    unsigned long long h = 0;
    SG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    *out = (int64_t)h;
    return SG_OK;
}

}  // namespace sg
