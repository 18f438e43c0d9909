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

// toKernel() (SparseGridGpu.hpp:1716-1751) for device inserts (BlockMapGpu_ker.cuh:171-227,
// SparseGridGpu_ker.cuh:445-490): init / insertBlock / insert<p> / flush_block_insert.
// insertBlock claims one chunk record per distinct chunk id among the calling lanes of a
// wave (the reference shares one per CUDA block); the records are merged by the next
// flush in claim order.
template <unsigned int dim, unsigned int B, typename Tuple, typename indexT> struct sparse_grid_gpu_ker {
    static constexpr unsigned int dims = dim;
    static constexpr unsigned int d = dim;
    static constexpr unsigned int blockEdgeSize_ = B;
    static constexpr unsigned int pMask = std::tuple_size<Tuple>::value;
    static constexpr unsigned int chunk_size = (dim == 3) ? B * B * B : B * B;
    sg_insert_view iv;
    int64_t bsz[3];
    int zorder;

    __device__ void init() const {}
    __device__ void flush_block_insert() const {}
    __device__ static constexpr unsigned int getBlockEdgeSize() { return B; }

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

    __device__ insert_encap<Tuple> insertBlock(indexT blockId) const
    {
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

}  // namespace sg

// BlockMapGpu_ker<>::setExist (BlockMapGpu_ker.cuh:75): sets the EXIST bit of a mask byte
template <typename = void> struct BlockMapGpu_ker {
    __device__ static void setExist(unsigned char &m) { m |= (unsigned char)SG_EXIST; }
};

// the reference's names in user code: cross_stencil<dim, T> and
// GetCpBlockType<grid, prop, stencil_size>::type (the type of the lambdas' block argument)
using sg::cross_stencil;
template <typename SG, unsigned int prop, unsigned int stencil_size> struct GetCpBlockType {
    typedef sg::conv_detail::cp_block<typename SG::template prop_type<prop>, SG::dims, SG::blockEdgeSize_, stencil_size>
        type;
};

#endif  // SPARSEGRIDGPU_DEVICE_HPP
