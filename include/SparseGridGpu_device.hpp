/* SparseGridGpu_device.hpp -- device side of the facade (hipcc only; included by
 * SparseGridGpu.hpp when __HIPCC__ is defined): the conv family for user __device__ lambdas
 * and the toKernel() view for device inserts (sparse_grid_gpu_ker, further below).
 *
 * Reference: SparseGridGpu::conv / conv2 / conv_cross / conv_cross_b / conv2_b / conv3_b
 * (src/SparseGridGpu/SparseGridGpu.hpp:2568-2683) with the functors stencil_cross_func_conv,
 * stencil_func_conv2, stencil_cross_func, stencil_cross_func_conv_block_read,
 * stencil_func_conv2_b, stencil_func_conv3_b (src/SparseGridGpu/SparseGridGpu_kernels.cuh:
 * 60-664), cp_block (src/SparseGrid/cp_block.hpp) and cross_stencil
 * (src/SparseGridGpu/SparseGridGpu_ker_util.hpp:31-36).
 *
 * One workgroup per chunk, one thread per chunk slot.  Each source property's chunk and
 * its ghost layer of `R` points ((B + 2R)^dim values, x fastest) is staged in LDS from the
 * chunk and its 3^dim - 1 neighbours (full neighbour table of sg_to_kernel); points
 * outside the domain and in missing chunks hold the background.  The lambda is called for
 * every point that exists, is not padding and lies in the inclusive box [start, stop];
 * the results are written to the destination properties there.  Differences from the
 * reference: the ghost layer includes the edge and corner neighbours (the reference
 * loads the 6 face neighbours only, so diagonal taps read unrelated values there), and a
 * destination may not be one of the sources (the reference's in-place update races with
 * the ghost loads of neighbouring chunks).  conv_cross writes every slot with a non-zero
 * mask (the result where the point is active, the source value elsewhere), as the
 * reference's block store does for those slots; empty slots keep the destination
 * background.
 */
#ifndef SPARSEGRIDGPU_DEVICE_HPP
#define SPARSEGRIDGPU_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <tuple>
#include <type_traits>

#include "sparsegrid_gpu.h"

namespace sg {

// cross_stencil<dim, T> (SparseGridGpu_ker_util.hpp:31-36): the 2 dim face neighbours
template <unsigned int dim, typename T> struct cross_stencil {
    T xm[dim];
    T xp[dim];
};

namespace conv_detail {

// the ghosted block of one property in LDS; (i, j[, k]) are chunk-local coordinates,
// valid in [-R, B + R)
template <typename T, unsigned int dim, unsigned int B, unsigned int R> struct cp_block {
    static constexpr int E = (int)(B + 2 * R);
    T *p;
    __device__ T &operator()(int i, int j) const { return p[(j + (int)R) * E + (i + (int)R)]; }
    __device__ T &operator()(int i, int j, int k) const
    {
        return p[((k + (int)R) * E + (j + (int)R)) * E + (i + (int)R)];
    }
    __device__ int LinId(int i, int j) const { return (j + (int)R) * E + (i + (int)R); }
    __device__ int LinId(int i, int j, int k) const { return ((k + (int)R) * E + (j + (int)R)) * E + (i + (int)R); }
};

// the chunk's own data (DataBlockLoad of the _b variants): get<p>()[offset]
template <typename Tuple> struct chunk_block {
    void *arr[16];
    int64_t base;  // chunk position * chunk size
    template <unsigned int p> __device__ typename std::tuple_element<p, Tuple>::type *get() const
    {
        return (typename std::tuple_element<p, Tuple>::type *)arr[p] + base;
    }
};

// conv_cross's lambda with the extra arguments of the call appended: f(cur, cs, args...)
template <typename F, typename... A> struct with_args {
    F f;
    std::tuple<A...> a;
    template <typename T, typename CS, size_t... I>
    __device__ auto call(T &cur, CS &cs, std::index_sequence<I...>) const
    {
        return f(cur, cs, std::get<I>(a)...);
    }
    template <typename T, typename CS> __device__ auto operator()(T &cur, CS &cs) const
    {
        return call(cur, cs, std::index_sequence_for<A...>());
    }
};

struct ConvArgs {
    const int64_t *keys;
    const uint8_t *mask;
    const int32_t *nn;
    int64_t n;
    int32_t nnn, zorder;
    int64_t bsz[3];
    int64_t lo[3], hi[3];  // inclusive box
    void *prop[16];
};

template <unsigned int dim> __device__ inline void chunk_coords(const ConvArgs &a, int64_t id, int64_t (&c)[3])
{
    c[0] = c[1] = c[2] = 0;
    if (a.zorder) {
        // grid_zmb: bit i of coordinate d at bit dim * i + d
        for (int i = 0; i < 64 / (int)dim; ++i)
            for (unsigned int d = 0; d < dim; ++d) c[d] |= ((id >> (dim * i + d)) & 1) << i;
        return;
    }
    for (unsigned int d = 0; d < dim; ++d) {
        c[d] = id % a.bsz[d];
        id /= a.bsz[d];
    }
}

// position of the chunk at displacement (dx, dy, dz) in {-1, 0, 1} from chunk c (at chunk
// coordinates cc); a neighbour outside the domain is the background chunk
template <unsigned int dim>
__device__ inline int64_t neighbour_pos(const ConvArgs &a, int64_t c, const int64_t (&cc)[3], int dx, int dy, int dz)
{
    const int dd[3] = {dx, dy, dz};
    int k = 0, pw = 1;
    for (unsigned int d = 0; d < dim; ++d) {
        const int64_t q = cc[d] + dd[d];
        if (q < 0 || q >= a.bsz[d]) return a.n;
        k += (dd[d] + 1) * pw;
        pw *= 3;
    }
    if (k == (pw - 1) / 2) return c;
    return a.nn[c * a.nnn + k];
}

// stage the ghosted block of `arr` for chunk c (all threads of the workgroup)
template <typename T, unsigned int dim, unsigned int B, unsigned int R>
__device__ inline void stage(T *lds, const T *arr, const ConvArgs &a, int64_t c, const int64_t (&cc)[3])
{
    constexpr int E = (int)(B + 2 * R);
    constexpr int NE = (dim == 3) ? E * E * E : E * E;
    constexpr int CS = (dim == 3) ? (int)(B * B * B) : (int)(B * B);
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        int q[3] = {e % E - (int)R, (e / E) % E - (int)R, dim == 3 ? e / (E * E) - (int)R : 0};
        int dv[3] = {0, 0, 0};
        for (unsigned int d = 0; d < dim; ++d) {
            dv[d] = (q[d] < 0) ? -1 : (q[d] >= (int)B ? 1 : 0);
            q[d] -= dv[d] * (int)B;
        }
        const int64_t pos = neighbour_pos<dim>(a, c, cc, dv[0], dv[1], dv[2]);
        const int local = (dim == 3) ? (q[2] * (int)B + q[1]) * (int)B + q[0] : q[1] * (int)B + q[0];
        lds[e] = arr[pos * CS + local];
    }
}

template <unsigned int dim, unsigned int B> struct PointCtx {
    int i, j, k;
    int64_t slot;  // position in the property arrays
    uint8_t m;
    bool active;
};

template <unsigned int dim, unsigned int B>
__device__ inline PointCtx<dim, B> point_ctx(const ConvArgs &a, int64_t c, const int64_t (&cc)[3])
{
    constexpr int CS = (dim == 3) ? (int)(B * B * B) : (int)(B * B);
    PointCtx<dim, B> p;
    const int t = threadIdx.x;
    p.i = t % (int)B;
    p.j = (t / (int)B) % (int)B;
    p.k = (dim == 3) ? t / (int)(B * B) : 0;
    p.slot = c * CS + t;
    p.m = a.mask[p.slot];
    const int64_t g[3] = {cc[0] * B + p.i, cc[1] * B + p.j, cc[2] * B + p.k};
    bool in = true;
    for (unsigned int d = 0; d < dim; ++d) in = in && g[d] >= a.lo[d] && g[d] <= a.hi[d];
    p.active = in && (p.m & SG_EXIST) && !(p.m & SG_PADDING);
    return p;
}

// MODE: 0 conv, 1 conv_cross, 2 conv_cross_b, 3 conv2, 4 conv2_b, 5 conv3_b
template <unsigned int dim, unsigned int B, unsigned int R, int MODE, typename Tuple, typename T1, typename T2,
          typename T3, unsigned int s1, unsigned int s2, unsigned int s3, unsigned int d1, unsigned int d2,
          unsigned int d3, typename F>
__global__ void conv_kernel(ConvArgs a, F f)
{
    constexpr int E = (int)(B + 2 * R);
    constexpr int NE = (dim == 3) ? E * E * E : E * E;
    constexpr int NS = (MODE == 5) ? 3 : (MODE >= 3 ? 2 : 1);
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    T1 *b1 = (T1 *)smem;
    T2 *b2 = (T2 *)(smem + ((NE * sizeof(T1) + 15) / 16) * 16);
    T3 *b3 = (T3 *)((unsigned char *)b2 + ((NE * sizeof(T2) + 15) / 16) * 16);
    const int64_t c = blockIdx.x;
    int64_t cc[3];
    chunk_coords<dim>(a, a.keys[c], cc);
    stage<T1, dim, B, R>(b1, (const T1 *)a.prop[s1], a, c, cc);
    if constexpr (NS >= 2) stage<T2, dim, B, R>(b2, (const T2 *)a.prop[s2], a, c, cc);
    if constexpr (NS >= 3) stage<T3, dim, B, R>(b3, (const T3 *)a.prop[s3], a, c, cc);
    __syncthreads();
    const PointCtx<dim, B> p = point_ctx<dim, B>(a, c, cc);
    cp_block<T1, dim, B, R> cpb1{b1};
    cp_block<T2, dim, B, R> cpb2{b2};
    cp_block<T3, dim, B, R> cpb3{b3};
    chunk_block<Tuple> blk;
    for (int q = 0; q < 16; ++q) blk.arr[q] = a.prop[q];
    blk.base = p.slot - threadIdx.x;
    const int offset = threadIdx.x;
    T1 *o1 = (T1 *)a.prop[d1];
    if constexpr (MODE == 1) {
        if (p.m == 0) return;
        T1 cur = (dim == 3) ? cpb1(p.i, p.j, p.k) : cpb1(p.i, p.j);
        T1 res = cur;
        if (p.active) {
            cross_stencil<dim, T1> cs;
            const int cd[3] = {p.i, p.j, p.k};
            for (unsigned int d = 0; d < dim; ++d) {
                int lo[3] = {cd[0], cd[1], cd[2]}, hi[3] = {cd[0], cd[1], cd[2]};
                lo[d] -= 1;
                hi[d] += 1;
                cs.xm[d] = (dim == 3) ? cpb1(lo[0], lo[1], lo[2]) : cpb1(lo[0], lo[1]);
                cs.xp[d] = (dim == 3) ? cpb1(hi[0], hi[1], hi[2]) : cpb1(hi[0], hi[1]);
            }
            res = f(cur, cs);
        }
        o1[p.slot] = res;
        return;
    } else {
        if (!p.active) return;
    }
    if constexpr (MODE == 0) {
        if constexpr (dim == 3) o1[p.slot] = f(cpb1, p.i, p.j, p.k);
        else o1[p.slot] = f(cpb1, p.i, p.j);
    } else if constexpr (MODE == 2) {
        if constexpr (dim == 3) o1[p.slot] = f(cpb1, blk, offset, p.i, p.j, p.k);
        else o1[p.slot] = f(cpb1, blk, offset, p.i, p.j);
    } else if constexpr (MODE == 3 || MODE == 4) {
        T1 r1 = 0;
        T2 r2 = 0;
        if constexpr (MODE == 3) {
            if constexpr (dim == 3) f(r1, r2, cpb1, cpb2, p.i, p.j, p.k);
            else f(r1, r2, cpb1, cpb2, p.i, p.j);
        } else {
            if constexpr (dim == 3) f(r1, r2, cpb1, cpb2, blk, offset, p.i, p.j, p.k);
            else f(r1, r2, cpb1, cpb2, blk, offset, p.i, p.j);
        }
        o1[p.slot] = r1;
        ((T2 *)a.prop[d2])[p.slot] = r2;
    } else if constexpr (MODE == 5) {
        T1 r1 = 0;
        T2 r2 = 0;
        T3 r3 = 0;
        if constexpr (dim == 3) f(r1, r2, r3, cpb1, cpb2, cpb3, blk, offset, p.i, p.j, p.k);
        else f(r1, r2, r3, cpb1, cpb2, cpb3, blk, offset, p.i, p.j);
        o1[p.slot] = r1;
        ((T2 *)a.prop[d2])[p.slot] = r2;
        ((T3 *)a.prop[d3])[p.slot] = r3;
    }
}

template <unsigned int dim, unsigned int B, unsigned int R, int MODE, typename Tuple, unsigned int s1,
          unsigned int s2, unsigned int s3, unsigned int d1, unsigned int d2, unsigned int d3, typename F>
inline int launch(sg_grid *h, const int64_t (&lo)[3], const int64_t (&hi)[3], F f)
{
    static_assert(R >= 1 && R <= B, "the stencil size must be in [1, chunk edge]");
    typedef typename std::tuple_element<s1, Tuple>::type T1;
    typedef typename std::tuple_element<s2, Tuple>::type T2;
    typedef typename std::tuple_element<s3, Tuple>::type T3;
    sg_kernel_view v;
    int rc = sg_to_kernel(h, SG_NN_FULL, &v);
    if (rc) return rc;
    if (v.nchunks == 0) return SG_OK;
    ConvArgs a;
    a.keys = v.keys;
    a.mask = v.mask;
    a.nn = v.nn;
    a.n = v.nchunks;
    a.nnn = v.nnn;
    a.zorder = v.zorder;
    for (int d = 0; d < 3; ++d) {
        a.bsz[d] = v.bsz[d];
        a.lo[d] = lo[d];
        a.hi[d] = hi[d];
    }
    for (int q = 0; q < 16; ++q) a.prop[q] = v.prop[q];
    constexpr int E = (int)(B + 2 * R);
    constexpr size_t NE = (dim == 3) ? (size_t)E * E * E : (size_t)E * E;
    const size_t lds = ((NE * sizeof(T1) + 15) / 16) * 16 + ((NE * sizeof(T2) + 15) / 16) * 16 +
                       ((NE * sizeof(T3) + 15) / 16) * 16;
    constexpr unsigned int CS = (dim == 3) ? B * B * B : B * B;
    static_assert(CS <= 1024, "one thread per chunk slot: the chunk must hold at most 1024 points");
    hipLaunchKernelGGL((conv_kernel<dim, B, R, MODE, Tuple, T1, T2, T3, s1, s2, s3, d1, d2, d3, F>),
                       dim3((unsigned)v.nchunks), dim3(CS), lds, (hipStream_t)v.stream, a, f);
    return hipGetLastError() == hipSuccess ? SG_OK : SG_E_HIP;
}

}  // namespace conv_detail
}  // namespace sg

namespace sg {

// a claimed chunk record of a device insert (insertBlock): get<p>()[offset] is property p
// of the record's slot `offset`, get<nprops>() the mask bytes
template <typename Tuple> struct insert_encap {
    static constexpr unsigned int np = std::tuple_size<Tuple>::value;
    void *prop[16];
    uint8_t *mask;
    template <unsigned int p>
    __device__ typename std::conditional<(p < np), typename std::tuple_element<(p < np ? p : 0), Tuple>::type,
                                         uint8_t>::type *
    get() const
    {
        typedef typename std::conditional<(p < np), typename std::tuple_element<(p < np ? p : 0), Tuple>::type,
                                          uint8_t>::type E;
        if constexpr (p < np) return (E *)prop[p];
        else return (E *)mask;
    }
};

// toKernel() (SparseGridGpu.hpp:1716-1751, SparseGridGpu_ker.cuh): the device view of a grid.
//  - reads (SparseGridGpu_ker.cuh:296-443, :670-693; BlockMapGpu_ker.cuh:105-135): get<p>(coord)
//    (the background where the point does not exist), get_sparse, exist, isPadding,
//    getNNPoint, get<p>(block_offset);
//  - the stencil-functor protocol of applyStencils (SparseGridGpu.hpp:954-1048, kernel
//    SparseGridGpu_kernels.cuh:1037-1094): loadGhostBlock / storeBlock and the enlarged-block
//    index helpers, for one workgroup of chunk-size threads per chunk;
//  - device inserts (BlockMapGpu_ker.cuh:171-227, SparseGridGpu_ker.cuh:445-490): init /
//    insertBlock / insert<p> / flush_block_insert, when setGPUInsertBuffer opened a buffer.
//    insertBlock claims one chunk record per distinct chunk id among the calling lanes of a
//    wave (the reference shares one per CUDA block); the records are merged by the next
//    flush in claim order.
// Ghost loads follow the neighbourhood type of the stencil (nn_star): NNStar reads the 2 dim
// face neighbours (background outside the domain; edge and corner cells hold the background,
// where the reference reads an unrelated neighbour), NNFull reads all 3^dim - 1 through the
// full neighbour table, exactly as the reference's table (so at the low x / y faces an
// out-of-domain neighbour aliases a chunk of the previous row, grid_smb.hpp:258-276).
template <unsigned int dim, unsigned int B, typename Tuple, typename indexT> struct sparse_grid_gpu_ker {
    static constexpr unsigned int dims = dim;
    static constexpr unsigned int d = dim;
    static constexpr unsigned int blockEdgeSize_ = B;
    static constexpr unsigned int pMask = std::tuple_size<Tuple>::value;
    static constexpr unsigned int chunk_size = (dim == 3) ? B * B * B : B * B;
    static constexpr unsigned int blockSize = chunk_size;
    static constexpr unsigned int stencilSupportRadius = 1;
    static constexpr int EB = (int)B + 2;  // edge of the enlarged block (radius 1)
    typedef typename block_aggregate_of<chunk_size, Tuple>::type AggregateBlockType;
    typedef AggregateBlockType AggregateType;
    typedef indexT indexT_;
    template <unsigned int p>
    using elem_t = typename std::conditional<(p < pMask), typename std::tuple_element<(p < pMask ? p : 0), Tuple>::type,
                                             unsigned char>::type;

    sg_insert_view iv;  // device insert buffer (count == nullptr: none open)
    sg_kernel_view gv;  // the grid's arrays and full neighbour table
    int64_t bsz[3];
    int zorder;
    int nn_star;        // ghost loads with NNStar semantics (else NNFull)

    __device__ void init() const {}
    __device__ void flush_block_insert() const {}
    __device__ static constexpr unsigned int getBlockEdgeSize() { return B; }
    __device__ static constexpr unsigned int getBlockSize() { return chunk_size; }
    __device__ static constexpr unsigned int getEnlargedBlockSize() { return dim == 3 ? EB * EB * EB : EB * EB; }

    // ---- geometry (grid_smb / grid_zmb) ----
    // chunk id of chunk coordinates (grid_smb row-major or grid_zmb Morton)
    __device__ indexT chunk_id(const int64_t (&c)[3]) const
    {
        if (zorder) {
            uint64_t r = 0;
            for (int i = 0; i < 64 / (int)dim; ++i)
                for (unsigned int q = 0; q < dim; ++q) r |= (uint64_t)((c[q] >> i) & 1) << (dim * i + q);
            return (indexT)r;
        }
        int64_t v = c[dim - 1];
        for (int q = (int)dim - 2; q >= 0; --q) v = v * bsz[q] + c[q];
        return (indexT)v;
    }
    __device__ void chunk_coords(int64_t id, int64_t (&c)[3]) const
    {
        c[0] = c[1] = c[2] = 0;
        if (zorder) {
            for (int i = 0; i < 64 / (int)dim; ++i)
                for (unsigned int q = 0; q < dim; ++q) c[q] |= ((id >> (dim * i + q)) & 1) << i;
            return;
        }
        for (unsigned int q = 0; q < dim; ++q) {
            c[q] = id % bsz[q];
            id /= bsz[q];
        }
    }
    // point key = chunk id * chunk size + x-fastest offset inside the chunk
    template <typename CoordT> __device__ indexT getLinId(const grid_key_dx<dim, CoordT> &k) const
    {
        int64_t c[3] = {0, 0, 0};
        int off = 0, mul = 1;
        for (unsigned int q = 0; q < dim; ++q) {
            c[q] = (int64_t)k.get(q) / B;
            off += (int)((int64_t)k.get(q) % B) * mul;
            mul *= (int)B;
        }
        return chunk_id(c) * (indexT)chunk_size + off;
    }
    __device__ grid_key_dx<dim, int> getCoord(indexT lin) const
    {
        int64_t c[3];
        chunk_coords((int64_t)(lin / (indexT)chunk_size), c);
        int off = (int)(lin % (indexT)chunk_size);
        grid_key_dx<dim, int> k;
        for (unsigned int q = 0; q < dim; ++q) {
            k.set_d(q, (int)(c[q] * B + off % (int)B));
            off /= (int)B;
        }
        return k;
    }
    template <typename CoordT> __device__ indexT getBlockLinId(const grid_key_dx<dim, CoordT> &blockCoord) const
    {
        int64_t c[3] = {0, 0, 0};
        for (unsigned int q = 0; q < dim; ++q) c[q] = (int64_t)blockCoord.get(q);
        return chunk_id(c);
    }
    __device__ grid_key_dx<dim, int> getBlockCoord(indexT blockId) const
    {
        int64_t c[3];
        chunk_coords((int64_t)blockId, c);
        grid_key_dx<dim, int> k;
        for (unsigned int q = 0; q < dim; ++q) k.set_d(q, (int)c[q]);
        return k;
    }
    // global coordinates of slot `offset` of the chunk at chunk coordinates blockCoord
    template <typename CoordT>
    __device__ grid_key_dx<dim, CoordT> getGlobalCoord(const grid_key_dx<dim, CoordT> &blockCoord, unsigned int offset) const
    {
        grid_key_dx<dim, CoordT> k;
        for (unsigned int q = 0; q < dim; ++q) {
            k.set_d(q, (CoordT)(blockCoord.get(q) * (CoordT)B + (CoordT)(offset % B)));
            offset /= B;
        }
        return k;
    }

    // ---- reads ----
    // position of chunk `blockId` in the chunk list, or nchunks (the background chunk)
    __device__ int64_t chunk_pos(int64_t blockId) const
    {
        int64_t lo = 0, hi = gv.nchunks;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (gv.keys[mid] < blockId) lo = mid + 1;
            else hi = mid;
        }
        return (lo < gv.nchunks && gv.keys[lo] == blockId) ? lo : gv.nchunks;
    }
    // get_sparse (SparseGridGpu_ker.cuh:409-414): chunk position (nchunks where the chunk does
    // not exist, or the point is outside the domain) and offset of a point
    template <typename CoordT>
    __device__ void get_sparse(const grid_key_dx<dim, CoordT> &coord, unsigned int &dataBlockPos, unsigned int &offset) const
    {
        bool in = true;
        for (unsigned int q = 0; q < dim; ++q) in = in && coord.get(q) >= 0 && (int64_t)coord.get(q) < gv.size[q];
        int off = 0, mul = 1;
        for (unsigned int q = 0; q < dim; ++q) {
            off += (int)(((int64_t)coord.get(q) % B + B) % B) * mul;
            mul *= (int)B;
        }
        offset = (unsigned int)off;
        dataBlockPos = (unsigned int)(in ? chunk_pos((int64_t)(getLinId(coord) / (indexT)chunk_size)) : gv.nchunks);
    }
    // get<p>(coord) (SparseGridGpu_ker.cuh:401-406, BlockMapGpu_ker.cuh:119-135)
    template <unsigned int p, typename CoordT> __device__ elem_t<p> get(const grid_key_dx<dim, CoordT> &coord) const
    {
        unsigned int pos, off;
        get_sparse(coord, pos, off);
        const int64_t s = (int64_t)pos * chunk_size + off;
        const elem_t<p> *a = (const elem_t<p> *)gv.prop[p];
        return (gv.mask[s] & SG_EXIST) ? a[s] : a[gv.nchunks * chunk_size + off];
    }
    // get<p>(block_offset) (SparseGridGpu_ker.cuh:424-443): the value at a chunk position
    template <unsigned int p, typename CoordT> __device__ elem_t<p> &get(const block_offset<CoordT> &bo) const
    {
        elem_t<p> *a = (p < pMask) ? (elem_t<p> *)gv.prop[p < pMask ? p : 0] : (elem_t<p> *)gv.mask;
        return a[(int64_t)bo.pos * chunk_size + bo.off];
    }
    template <typename CoordT> __device__ unsigned char getMask(const grid_key_dx<dim, CoordT> &coord) const
    {
        unsigned int pos, off;
        get_sparse(coord, pos, off);
        return gv.mask[(int64_t)pos * chunk_size + off];
    }
    template <typename CoordT> __device__ bool exist(const grid_key_dx<dim, CoordT> &coord) const
    {
        return getMask(coord) & SG_EXIST;
    }
    template <typename CoordT> __device__ bool isPadding(const grid_key_dx<dim, CoordT> &coord) const
    {
        return getMask(coord) & SG_PADDING;
    }
    // mask-byte helpers of BlockMapGpu_ker / SparseGridGpu_ker (bit 0 EXIST, bit 1 PADDING)
    template <typename BitMaskT> __device__ static bool getBit(const BitMaskT &m, unsigned char pos) { return (m >> pos) & 1u; }
    template <typename BitMaskT> __device__ static void setBit(BitMaskT &m, unsigned char pos) { m = m | (1u << pos); }
    template <typename BitMaskT> __device__ static void unsetBit(BitMaskT &m, unsigned char pos) { m = m & ~(1u << pos); }
    template <typename BitMaskT> __device__ static bool exist(const BitMaskT &m) { return getBit(m, 0); }
    template <typename BitMaskT> __device__ static void setExist(BitMaskT &m) { setBit(m, 0); }
    template <typename BitMaskT> __device__ static void unsetExist(BitMaskT &m) { unsetBit(m, 0); }
    template <typename BitMaskT> __device__ static bool isPadding(const BitMaskT &m) { return getBit(m, 1); }
    template <typename BitMaskT> __device__ static void setPadding(BitMaskT &m) { setBit(m, 1); }
    template <typename BitMaskT> __device__ static void unsetPadding(BitMaskT &m) { unsetBit(m, 1); }

    // position of the neighbour chunk of the chunk at `pos` in direction dd (entries -1/0/1)
    // under the stencil's neighbourhood (see above); nchunks = the background chunk
    __device__ int64_t neighbour_chunk(int64_t pos, const int (&dd)[3], bool star) const
    {
        int k = 0, pw = 1, nz = 0;
        for (unsigned int q = 0; q < dim; ++q) {
            k += (dd[q] + 1) * pw;
            pw *= 3;
            nz += dd[q] != 0;
        }
        if (nz == 0 || pos >= gv.nchunks) return pos;
        if (!star) return gv.nn[pos * gv.nnn + k];
        if (nz > 1) return gv.nchunks;
        int64_t c[3];
        chunk_coords(gv.keys[pos], c);
        for (unsigned int q = 0; q < dim; ++q) {
            const int64_t v = c[q] + dd[q];
            if (v < 0 || v >= bsz[q]) return gv.nchunks;
        }
        return gv.nn[pos * gv.nnn + k];
    }
    // getNNPoint<NN>(pos, offset, mov) (SparseGridGpu_ker.cuh:296-327): the chunk position and
    // offset of the point at displacement mov (at most one chunk edge) from slot `offset` of the
    // chunk at pos
    template <typename NN_type, typename indexT2>
    __device__ block_offset<indexT2> getNNPoint(openfpm::sparse_index<unsigned int> pos, unsigned int offset,
                                                const grid_key_dx<dim, indexT2> &mov) const
    {
        int dd[3] = {0, 0, 0};
        int off = 0, mul = 1;
        for (unsigned int q = 0; q < dim; ++q) {
            int l = (int)(offset % B) + (int)mov.get(q);
            offset /= B;
            dd[q] = (l < 0) ? -1 : (l >= (int)B ? 1 : 0);
            l -= dd[q] * (int)B;
            off += l * mul;
            mul *= (int)B;
        }
        block_offset<indexT2> bo;
        bo.pos = (indexT2)neighbour_chunk((int64_t)pos.id, dd, NN_type::type == SG_NN_STAR);
        bo.off = (indexT2)off;
        return bo;
    }

    // ---- enlarged block (chunk + ghost layer of radius 1, x fastest, in LDS) ----
    __device__ grid_key_dx<dim, int> getCoordInEnlargedBlock(unsigned int offset) const
    {
        grid_key_dx<dim, int> k;
        for (unsigned int q = 0; q < dim; ++q) {
            k.set_d(q, (int)(offset % B) + 1);
            offset /= B;
        }
        return k;
    }
    template <typename CoordT> __device__ static unsigned int enlarged_lin(const grid_key_dx<dim, CoordT> &c)
    {
        unsigned int lin = 0;
        for (int q = (int)dim - 1; q >= 0; --q) lin = lin * EB + (unsigned int)c.get(q);
        return lin;
    }
    __device__ unsigned int getLinIdInEnlargedBlock(unsigned int offset) const
    {
        return enlarged_lin(getCoordInEnlargedBlock(offset));
    }
    __device__ unsigned int posToEnlargedBlockPos(unsigned int pos) const { return getLinIdInEnlargedBlock(pos); }
    template <typename CoordT>
    __device__ unsigned int getNeighbourLinIdInEnlargedBlock(const grid_key_dx<dim, CoordT> &base,
                                                             const grid_key_dx<dim, CoordT> &offsets) const
    {
        grid_key_dx<dim, int> r;
        for (unsigned int q = 0; q < dim; ++q) r.set_d(q, (int)(base.get(q) + offsets.get(q)));
        return enlarged_lin(r);
    }
    template <typename CoordT>
    __device__ unsigned int getNeighbourLinIdInEnlargedBlock(const grid_key_dx<dim, CoordT> &base, unsigned int dimension,
                                                             char offset) const
    {
        grid_key_dx<dim, int> r;
        for (unsigned int q = 0; q < dim; ++q) r.set_d(q, (int)base.get(q) + (q == dimension ? (int)offset : 0));
        return enlarged_lin(r);
    }
    // loadGhostBlock<p> (SparseGridGpu_ker.cuh:555-570, loadGhostBlock_impl
    // SparseGridGpu_ker_util.hpp:487-945): the chunk at pos and its ghost layer into
    // sharedRegion (and the mask bytes into maskRegion), with the background wherever a
    // point does not exist.  Every thread of the workgroup takes part; the caller
    // synchronises before reading.
    template <unsigned int p, typename AggrWrapperT>
    __device__ void loadGhostBlock(const AggrWrapperT &, const openfpm::sparse_index<unsigned int> pos,
                                   elem_t<p> *sharedRegion, unsigned char *maskRegion = nullptr) const
    {
        const elem_t<p> *a = (p < pMask) ? (const elem_t<p> *)gv.prop[p < pMask ? p : 0] : (const elem_t<p> *)gv.mask;
        const int64_t bgbase = gv.nchunks * chunk_size;
        const unsigned int nthr = blockDim.x * blockDim.y * blockDim.z;
        const unsigned int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
        for (unsigned int e = tid; e < getEnlargedBlockSize(); e += nthr) {
            unsigned int r = e;
            int dd[3] = {0, 0, 0};
            int off = 0, mul = 1;
            for (unsigned int q = 0; q < dim; ++q) {
                int l = (int)(r % EB) - 1;
                r /= EB;
                dd[q] = (l < 0) ? -1 : (l >= (int)B ? 1 : 0);
                l -= dd[q] * (int)B;
                off += l * mul;
                mul *= (int)B;
            }
            const int64_t np = neighbour_chunk((int64_t)pos.id, dd, nn_star != 0);
            const int64_t s = np * chunk_size + off;
            const unsigned char m = gv.mask[s];
            sharedRegion[e] = (p == pMask || (m & SG_EXIST)) ? a[s] : a[bgbase + off];
            if (maskRegion) maskRegion[e] = m;
        }
    }
    // storeBlock<p> (SparseGridGpu_ker.cuh:872-891): the inner part of sharedRegion to the chunk
    template <unsigned int p, typename AggrWrapperT> __device__ void storeBlock(AggrWrapperT &block, elem_t<p> *sharedRegion) const
    {
        const unsigned int nthr = blockDim.x * blockDim.y * blockDim.z;
        const unsigned int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
        for (unsigned int o = tid; o < chunk_size; o += nthr) block.template get<p>()[o] = sharedRegion[getLinIdInEnlargedBlock(o)];
    }
    // the arrays of the chunk at position pos (the dataBlockLoad / dataBlockStore of the functors)
    __device__ insert_encap<Tuple> block_at(int64_t pos) const
    {
        insert_encap<Tuple> e;
        for (unsigned int q = 0; q < 16; ++q) e.prop[q] = nullptr;
        for (unsigned int q = 0; q < pMask; ++q)
            e.prop[q] = (unsigned char *)gv.prop[q] + (size_t)pos * chunk_size * sizeof_prop(q);
        e.mask = gv.mask + (size_t)pos * chunk_size;
        return e;
    }

    // ---- device inserts ----
    // insertBlock<nChunksPerBlocks>(blockId, stride) (BlockMapGpu_ker.cuh:199-227): the chunk
    // record of blockId for the calling lanes; the template parameter and the stride of the
    // reference's per-block sharing are accepted (records are shared per wave here)
    template <unsigned int nChunksPerBlocks = 1>
    __device__ insert_encap<Tuple> insertBlock(indexT blockId, unsigned int stride = 8192) const
    {
        (void)stride;
        const int lane = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)) & 63;
        uint64_t todo = __ballot(1);
        int64_t rec = 0;
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int64_t lid = __shfl((long long)blockId, leader);
            const uint64_t peers = __ballot((int64_t)blockId == lid) & todo;
            unsigned int idx = 0;
            if (lane == leader) {
                idx = atomicAdd(iv.count, 1u);
                if ((int64_t)idx < iv.cap) iv.ids[idx] = lid;
            }
            idx = (unsigned int)__shfl((int)idx, leader);
            if ((peers >> lane) & 1) rec = ((int64_t)idx < iv.cap) ? (int64_t)idx : iv.cap;  // the sink
            todo &= ~peers;
        }
        insert_encap<Tuple> e;
        for (unsigned int q = 0; q < 16; ++q) e.prop[q] = nullptr;
        for (unsigned int q = 0; q < pMask; ++q)
            e.prop[q] = (unsigned char *)iv.prop[q] + (size_t)rec * chunk_size * sizeof_prop(q);
        e.mask = iv.mask + (size_t)rec * chunk_size;
        return e;
    }
    // insert<p>(coord): the point exists after the next flush; returns its value of p
    template <unsigned int p, typename CoordT>
    __device__ typename std::tuple_element<p, Tuple>::type &insert(const grid_key_dx<dim, CoordT> &k) const
    {
        const indexT key = getLinId(k);
        insert_encap<Tuple> e = insertBlock(key / (indexT)chunk_size);
        const int off = (int)(key % (indexT)chunk_size);
        e.mask[off] |= (uint8_t)SG_EXIST;
        return e.template get<p>()[off];
    }

  private:
    template <size_t... I> __device__ static size_t sizeof_at(unsigned int q, std::index_sequence<I...>)
    {
        size_t s = 0;
        ((s = (q == I) ? sizeof(typename std::tuple_element<I, Tuple>::type) : s), ...);
        return s;
    }
    __device__ static size_t sizeof_prop(unsigned int q)
    {
        return sizeof_at(q, std::make_index_sequence<std::tuple_size<Tuple>::value>());
    }
};

// applyStencilInPlace (SparseGridGpu.hpp:954-1048, SparseGridGpu_kernels.cuh:1037-1094): one
// workgroup of chunk-size threads per chunk; the functor is called by every thread with the
// chunk's arrays as both dataBlockLoad and dataBlockStore (in place, as the reference) and
// curMask = the slot's mask, 0 outside the inclusive box [lo, hi].
template <typename Stencil, typename SGK, typename... Args>
__global__ void apply_stencil_functor_kernel(SGK sg, int lo0, int lo1, int lo2, int hi0, int hi1, int hi2, Args... args)
{
    constexpr unsigned int dim = SGK::dims;
    const int64_t pos = blockIdx.x;
    if (pos >= sg.gv.nchunks) return;  // uniform over the workgroup
    const unsigned int offset = threadIdx.x;
    auto blk = sg.block_at(pos);
    const int64_t id = sg.gv.keys[pos];
    grid_key_dx<dim, int> pointCoord = sg.getCoord((typename SGK::indexT_)(id * SGK::chunk_size + offset));
    unsigned char curMask = 0;
    if (offset < SGK::chunk_size) {
        curMask = sg.gv.mask[pos * SGK::chunk_size + offset];
        const int lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
        for (unsigned int q = 0; q < dim; ++q)
            if (pointCoord.get(q) < lo[q] || pointCoord.get(q) > hi[q]) curMask = 0;
    }
    openfpm::sparse_index<unsigned int> sp;
    sp.id = (unsigned int)pos;
    Stencil::stencil(sg, (unsigned int)id, sp, offset, pointCoord, blk, blk, curMask, args...);
}

// applyStencilInPlaceNoShared (SparseGridGpu.hpp:1051-1075, SparseGridGpu_kernels.cuh:1096-1192):
// one thread per existing point of the point list ep (data positions chunk * chunk size + offset)
template <typename Stencil, typename SGK, typename... Args>
__global__ void apply_stencil_functor_noshared_kernel(SGK sg, const int64_t *ep, int64_t nep, int lo0, int lo1, int lo2,
                                                      int hi0, int hi1, int hi2, Args... args)
{
    constexpr unsigned int dim = SGK::dims;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nep) return;
    const int64_t pos = ep[t] / SGK::chunk_size;
    const unsigned int offset = (unsigned int)(ep[t] % SGK::chunk_size);
    auto blk = sg.block_at(pos);
    const int64_t id = sg.gv.keys[pos];
    grid_key_dx<dim, int> pointCoord = sg.getCoord((typename SGK::indexT_)(id * SGK::chunk_size + offset));
    unsigned char curMask = sg.gv.mask[pos * SGK::chunk_size + offset];
    const int lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    for (unsigned int q = 0; q < dim; ++q)
        if (pointCoord.get(q) < lo[q] || pointCoord.get(q) > hi[q]) curMask = 0;
    openfpm::sparse_index<unsigned int> sp;
    sp.id = (unsigned int)pos;
    Stencil::stencil(sg, (unsigned int)id, sp, offset, pointCoord, blk, blk, curMask, args...);
}

}  // namespace sg

// BlockMapGpu_ker<>::setExist (BlockMapGpu_ker.cuh:75): sets the EXIST bit of a mask byte
template <typename = void> struct BlockMapGpu_ker {
    __device__ static void setExist(unsigned char &m) { m |= (unsigned char)SG_EXIST; }
    template <typename M> __device__ static void unsetExist(M &m) { m &= (M)~SG_EXIST; }
    template <typename M> __device__ static bool exist(const M &m) { return (m & SG_EXIST) != 0; }
    template <typename M> __device__ static bool getBit(const M &m, unsigned char pos) { return (m >> pos) & 1u; }
    template <typename M> __device__ static void setBit(M &m, unsigned char pos) { m = m | (1u << pos); }
    template <typename M> __device__ static void unsetBit(M &m, unsigned char pos) { m = m & ~(1u << pos); }
};

// the reference's names in user code: cross_stencil<dim, T> and
// GetCpBlockType<grid, prop, stencil_size>::type (the type of the lambdas' block argument)
using sg::cross_stencil;
template <typename SG, unsigned int prop, unsigned int stencil_size> struct GetCpBlockType {
    typedef sg::conv_detail::cp_block<typename SG::template prop_type<prop>, SG::dims, SG::blockEdgeSize_, stencil_size>
        type;
};

#endif  // SPARSEGRIDGPU_DEVICE_HPP
