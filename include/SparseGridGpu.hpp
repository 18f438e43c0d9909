/*
 * SparseGridGpu.hpp -- C++ drop-in facade of openfpm_data's SparseGridGpu over the C ABI of
 * libsg_hip.so (sparsegrid_gpu.h).  Host-only header: user code compiles with any C++17
 * compiler and links libsg_hip.so; the kernels behind it are hand-written gfx950 HIP.
 *
 * Kept from the reference (src/SparseGridGpu/SparseGridGpu.hpp):
 *   SparseGridGpu<dim, aggregate<T...>, blockEdgeSize, threadBlockSize, indexT> (:483-489),
 *   setBackgroundValue<p> (:2878), setGPUInsertBuffer (:2270), flush<v_reduce...>(ctx, opt)
 *   (:772), findNeighbours<NN> (:2396), setNNType<NN> (:1825), tagBoundaries<NN>(ctx, chk, opt)
 *   (:2290), applyStencils<Stencil>(box, mode, args...) (:2704), get<p>(coord) (:1897),
 *   insertFlush<p>(coord) (:2832), getLinId / getCoord (:1849-1857), getBox (:2690),
 *   countExistingElements / countBoundaryElements / measureBlockOccupancy (:2422-2552),
 *   deviceToHost / hostToDevice (BlockMapGpu.hpp:230-241), getIterator() / getIterator(start,
 *   stop) over the host copy (:3751-3764, Iterators/), packRequest / packCalculate / pack /
 *   packFinalize / unpack / remove / removeAddUnpackReset / removeAddUnpackFinalize (:3040-3733),
 *   copy_to / removeCopyToFinalize (:3550, :3314), construct_link_dw / construct_link_up and
 *   getDownLinks / getDownLinksOffsets / getUpLinks / getUpLinksOffsets (:2026-2262),
 *   reduction tags smax_ / smin_ / sadd_ / sRight_ / sLeft_ / sBitwiseOr_
 *   (src/Vector/cuda/map_vector_sparse_cuda_kernels.cuh), NNStar / NNFull, the test
 *   stencils HeatStencil and Conv3x3x3 (src/SparseGridGpu/tests/utils/SparseGridGpu_util_test.cuh).
 * Differences (a C ABI carries no device code):
 *   - insertion from user __global__ kernels (insert<p>() / insertBlock() on toKernel())
 *     needs a hipcc translation unit; host code inserts in batches (insertPoints /
 *     insertBlocks);
 *   - applyStencils takes the named stencils of this header (descriptors evaluated by
 *     library kernels), not arbitrary __device__ functors; user __device__ lambdas go
 *     through conv / conv2 / conv_cross / conv_cross_b / conv2_b / conv3_b, available in
 *     translation units compiled by hipcc (include/SparseGridGpu_device.hpp, over the
 *     toKernel view sg_to_kernel);
 *   - failures throw sg::error (the reference prints and continues).
 */
#ifndef SPARSEGRIDGPU_FACADE_HPP
#define SPARSEGRIDGPU_FACADE_HPP

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "sparsegrid_gpu.h"
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SPARSEGRIDGPU_HD __host__ __device__
#else
#define SPARSEGRIDGPU_HD
#endif

namespace sg {

struct error : std::runtime_error {
    explicit error(const std::string &m) : std::runtime_error(m) {}
};

inline void check(int rc)
{
    if (rc != SG_OK) throw error(std::string("libsg_hip: ") + sg_last_error());
}

template <typename T> struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = SG_F32; };
template <> struct dtype_of<double> { static constexpr int value = SG_F64; };
template <> struct dtype_of<int32_t> { static constexpr int value = SG_I32; };
template <> struct dtype_of<uint32_t> { static constexpr int value = SG_U32; };
template <> struct dtype_of<int64_t> { static constexpr int value = SG_I64; };
template <> struct dtype_of<uint64_t> { static constexpr int value = SG_U64; };
template <> struct dtype_of<int8_t> { static constexpr int value = SG_I8; };
template <> struct dtype_of<uint8_t> { static constexpr int value = SG_U8; };
template <> struct dtype_of<int16_t> { static constexpr int value = SG_I16; };
template <> struct dtype_of<uint16_t> { static constexpr int value = SG_U16; };

}  // namespace sg

// ---- reference vocabulary -------------------------------------------------------------

template <typename... T> struct aggregate {
    typedef std::tuple<T...> type;
    static constexpr unsigned int max_prop = sizeof...(T);
};
namespace sg {
template <typename Tuple> struct aggregate_of;
template <typename... T> struct aggregate_of<std::tuple<T...>> { typedef ::aggregate<T...> type; };
// scalar type of property p of an aggregate; p == number of properties is the mask byte
template <typename Agg, unsigned int p, bool prop = (p < Agg::max_prop)> struct scalar_of {
    typedef typename std::tuple_element<p, typename Agg::type>::type type;
};
template <typename Agg, unsigned int p> struct scalar_of<Agg, p, false> { typedef unsigned char type; };
template <int v> struct int_ { static constexpr int value = v; };
// the aggregate of a grid's chunks (AggregateBlockType / AggregateType of the device view):
// the property types and the chunk size N, so that BlockTypeOf<...>::size is N
template <unsigned int N, typename... T> struct block_aggregate {
    typedef std::tuple<T...> type;
    static constexpr unsigned int max_prop = sizeof...(T);
    static constexpr unsigned int chunk_size = N;
};
template <unsigned int N, typename Tuple> struct block_aggregate_of;
template <unsigned int N, typename... T> struct block_aggregate_of<N, std::tuple<T...>> {
    typedef block_aggregate<N, T...> type;
};
template <typename T, unsigned int N> struct data_block_type {
    typedef T scalarType;
    static constexpr unsigned int size = N;
};
}  // namespace sg
// ScalarTypeOf<AggregateT, p> of the stencil functors (the scalar of property p)
template <typename Agg, unsigned int p> using ScalarTypeOf = typename sg::scalar_of<Agg, p>::type;
// BlockTypeOf<AggregateType, p>: the chunk of property p (::size points of ScalarTypeOf)
template <typename Agg, unsigned int p> using BlockTypeOf = sg::data_block_type<ScalarTypeOf<Agg, p>, Agg::chunk_size>;
// IntPow<base, exp>::value (SparseGridGpu/TemplateUtils/mathUtils.hpp)
template <unsigned int base, unsigned int exp> struct IntPow {
    static constexpr unsigned int value = base * IntPow<base, exp - 1>::value;
};
template <unsigned int base> struct IntPow<base, 0> { static constexpr unsigned int value = 1; };
namespace openfpm {
// position of a chunk in the chunk list (openfpm::sparse_index, Vector/map_vector_sparse.hpp)
template <typename T> struct sparse_index { T id; };
}  // namespace openfpm
// a point as chunk position + offset in the chunk (block_offset, SparseGridGpu_ker.cuh)
template <typename T> struct block_offset {
    T pos;
    T off;
};

template <unsigned int dim, typename T = long int> class grid_key_dx
{
    T k[dim];

public:
    SPARSEGRIDGPU_HD grid_key_dx() : k{} {}
    SPARSEGRIDGPU_HD grid_key_dx(std::initializer_list<T> l)
    {
        unsigned int i = 0;
        for (T v : l)
            if (i < dim) k[i++] = v;
        for (; i < dim; ++i) k[i] = 0;
    }
    SPARSEGRIDGPU_HD T get(unsigned int i) const { return k[i]; }
    SPARSEGRIDGPU_HD void set_d(unsigned int i, T v) { k[i] = v; }
};

template <unsigned int dim, typename T> class Box
{
    T lo[dim], hi[dim];

public:
    Box() : lo{}, hi{} {}
    Box(std::initializer_list<T> l, std::initializer_list<T> h) : lo{}, hi{}
    {
        unsigned int i = 0;
        for (T v : l) if (i < dim) lo[i++] = v;
        i = 0;
        for (T v : h) if (i < dim) hi[i++] = v;
    }
    template <typename P> bool isInside(const P &p) const
    {
        for (unsigned int i = 0; i < dim; ++i)
            if ((T)p.get(i) < lo[i] || (T)p.get(i) > hi[i]) return false;
        return true;
    }
    void setLow(unsigned int i, T v) { lo[i] = v; }
    void setHigh(unsigned int i, T v) { hi[i] = v; }
    T getLow(unsigned int i) const { return lo[i]; }
    T getHigh(unsigned int i) const { return hi[i]; }
};

enum flush_type { FLUSH_ON_HOST = 0, FLUSH_ON_DEVICE = 1, FLUSH_NO_DATA = 2 };
enum tag_boundaries { NO_CALCULATE_EXISTING_POINTS = SG_NO_CALCULATE_EXISTING_POINTS,
                      CALCULATE_EXISTING_POINTS = SG_CALCULATE_EXISTING_POINTS };
enum StencilMode { STENCIL_MODE_INPLACE = 1, STENCIL_MODE_INPLACE_NO_SHARED = 3 };
enum mask_sparse { NOT_EXIST = 0, EXIST = 1, PADDING = 2, EXIST_AND_PADDING = 3 };
// removeCopyToFinalize options (Grid/grid_common.hpp:180-188)
enum rem_copy_opt { PHASE1 = SG_COPY_PHASE1, PHASE2 = SG_COPY_PHASE2, PHASE3 = SG_COPY_PHASE3,
                    KEEP_GEOMETRY = SG_COPY_KEEP_GEOMETRY };

struct No_check {};

namespace gpu {
// stream/temporary-storage context of the reference; the library keeps its own per grid
struct ofp_context_t {};
}  // namespace gpu

template <unsigned int p> struct smax_ { static constexpr unsigned int prop = p; static constexpr int op = SG_OP_MAX; };
template <unsigned int p> struct smin_ { static constexpr unsigned int prop = p; static constexpr int op = SG_OP_MIN; };
template <unsigned int p> struct sadd_ { static constexpr unsigned int prop = p; static constexpr int op = SG_OP_ADD; };
template <unsigned int p> struct sRight_ { static constexpr unsigned int prop = p; static constexpr int op = SG_OP_RIGHT; };
template <unsigned int p> struct sLeft_ { static constexpr unsigned int prop = p; static constexpr int op = SG_OP_LEFT; };
template <unsigned int p> struct sBitwiseOr_ { static constexpr unsigned int prop = p; static constexpr int op = SG_OP_OR; };

template <unsigned int dim> struct NNStar {
    static constexpr int type = SG_NN_STAR;
    static constexpr int nNN = 1 << dim;
};
template <unsigned int dim> struct NNFull {
    static constexpr int type = SG_NN_FULL;
    static constexpr int nNN = (dim == 2) ? 9 : 27;
};

#if defined(__HIPCC__)
#include "SparseGridGpu_device.hpp"
#endif

// ---- named stencils for applyStencils ----------------------------------------------------

namespace sg {
template <unsigned int dim> inline void box_to_desc(sg_stencil_desc &d, const Box<dim, int> &box)
{
    d.use_box = 1;
    for (unsigned int i = 0; i < dim; ++i) {
        d.box_lo[i] = box.getLow(i);
        d.box_hi[i] = box.getHigh(i);
    }
}
// a named stencil of this header: Stencil::desc(box, args...) exists
template <typename Stencil, typename BoxT, typename... Args> struct has_desc {
    template <typename S>
    static auto test(int) -> decltype(S::desc(std::declval<const BoxT &>(), std::declval<Args>()...), std::true_type());
    template <typename> static std::false_type test(...);
    static constexpr bool value = decltype(test<Stencil>(0))::value;
};
inline sg_stencil_desc blank_desc(int kind, int apply)
{
    sg_stencil_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.apply = apply;
    return d;
}
}  // namespace sg

// u_dst = u + dt * lap(u) (tests/utils/SparseGridGpu_util_test.cuh:201-260)
template <unsigned int dim, unsigned int p_src, unsigned int p_dst> struct HeatStencil {
    typedef NNStar<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    static constexpr unsigned int flops = 3 + 2 * dim;
    static sg_stencil_desc desc(const Box<dim, int> &box, double dt)
    {
        sg_stencil_desc d = sg::blank_desc(SG_ST_HEAT, SG_APPLY_INTERIOR);
        sg::box_to_desc<dim>(d, box);
        d.src[0] = p_src;
        d.dst[0] = p_dst;
        d.coef[0] = dt;
        return d;
    }
};

// u_dst = c0 * u + c1 * sum of the 2*dim face neighbours (conv with a 7-point Laplacian)
template <unsigned int dim, unsigned int p_src, unsigned int p_dst> struct LaplacianStencil {
    typedef NNStar<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    static sg_stencil_desc desc(const Box<dim, int> &box, double c0 = -2.0 * dim, double c1 = 1.0)
    {
        sg_stencil_desc d = sg::blank_desc(SG_ST_LAP, SG_APPLY_INTERIOR);
        sg::box_to_desc<dim>(d, box);
        d.src[0] = p_src;
        d.dst[0] = p_dst;
        d.coef[0] = c0;
        d.coef[1] = c1;
        return d;
    }
};

// two-property Gray-Scott step (conv2 form)
template <unsigned int dim, unsigned int pu, unsigned int pv, unsigned int pu_out, unsigned int pv_out>
struct GrayScottStencil {
    typedef NNStar<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    static sg_stencil_desc desc(const Box<dim, int> &box, double Du, double Dv, double F, double k, double dt)
    {
        sg_stencil_desc d = sg::blank_desc(SG_ST_GRAY_SCOTT, SG_APPLY_INTERIOR);
        sg::box_to_desc<dim>(d, box);
        d.src[0] = pu;
        d.src[1] = pv;
        d.dst[0] = pu_out;
        d.dst[1] = pv_out;
        d.coef[0] = Du;
        d.coef[1] = Dv;
        d.coef[2] = F;
        d.coef[3] = k;
        d.coef[4] = dt;
        return d;
    }
};

struct conv_coeff {
    float coeff[3][3][3];
};

// sum_{i,j,k} coeff[i][j][k] * u[x+i-1, y+j-1, z+k-1] on every existing point
// (Conv3x3x3, tests/utils/SparseGridGpu_util_test.cuh:357-420)
template <unsigned int dim, unsigned int p_src, unsigned int p_dst> struct Conv3x3x3 {
    typedef NNFull<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    static constexpr unsigned int flops = 2 * 27;
    static sg_stencil_desc desc(const Box<dim, int> &box, const conv_coeff &cc)
    {
        sg_stencil_desc d = sg::blank_desc(SG_ST_BOX, SG_APPLY_EXISTING);
        sg::box_to_desc<dim>(d, box);
        d.src[0] = p_src;
        d.dst[0] = p_dst;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    d.coef[i * 9 + j * 3 + k] = cc.coeff[i][j][k];
                    d.icoef[i * 9 + j * 3 + k] = (int64_t)std::llround(cc.coeff[i][j][k]);
                }
        return d;
    }
};

// ---- the grid ---------------------------------------------------------------------------

// One point of a grid's host copy: chunk position + offset in the chunk
// (sparse_grid_gpu_index, Iterators/SparseGridGpu_iterator.hpp:16-160).
template <typename SparseGridType> class sparse_grid_gpu_index
{
public:
    static constexpr unsigned int dims = SparseGridType::dims;
    sparse_grid_gpu_index(const SparseGridType &g, int cnk_pos_id, int data_id)
        : g_(&g), cnk_(cnk_pos_id), data_(data_id) {}
    int get_cnk_pos_id() const { return cnk_; }
    int get_data_id() const { return data_; }
    void set_cnk_pos_id(int c) { cnk_ = c; }
    void set_data_id(int d) { data_ = d; }
    // grid coordinates of the point
    grid_key_dx<dims, int> getKey() const
    {
        return g_->getCoord(g_->host_chunk_id(cnk_) * g_->getBlockSize() + data_);
    }
    int get(unsigned int i) const { return getKey().get(i); }

private:
    const SparseGridType *g_;
    int cnk_, data_;
};

// Iterator over every point of the host copy whose mask is not 0, chunk by chunk in the
// order of the chunk list (SparseGridGpu_iterator, Iterators/SparseGridGpu_iterator.hpp:162-306).
template <unsigned int dim, typename SparseGridType> class SparseGridGpu_iterator
{
public:
    explicit SparseGridGpu_iterator(const SparseGridType &g) : g_(&g) { settle(); }
    bool isNext() const { return chunk_ < g_->host_chunks(); }
    SparseGridGpu_iterator &operator++()
    {
        if (++pnt_ >= g_->getBlockSize()) {
            ++chunk_;
            pnt_ = 0;
        }
        settle();
        return *this;
    }
    sparse_grid_gpu_index<SparseGridType> get() const { return {*g_, (int)chunk_, (int)pnt_}; }

private:
    void settle()
    {
        const size_t cs = g_->getBlockSize(), nc = g_->host_chunks();
        const uint8_t *m = g_->host_mask();
        while (chunk_ < nc) {
            while (pnt_ < cs && m[chunk_ * cs + pnt_] == 0) ++pnt_;
            if (pnt_ < cs) return;
            ++chunk_;
            pnt_ = 0;
        }
    }
    const SparseGridType *g_;
    size_t chunk_ = 0, pnt_ = 0;
};

// Iterator over the EXIST points of the host copy inside the inclusive box [start, stop]:
// chunks in list order, points x-fastest within each chunk's part of the box
// (SparseGridGpu_iterator_sub, Iterators/SparseGridGpu_iterator_sub.hpp:13-260). pack,
// unpack and copy sections take their boxes from it; walking it needs the host copy
// (deviceToHost), which is looked at on the first isNext / get.
template <unsigned int dim, typename SparseGridType> class SparseGridGpu_iterator_sub
{
public:
    SparseGridGpu_iterator_sub(const SparseGridType &g, const grid_key_dx<dim, int> &start,
                               const grid_key_dx<dim, int> &stop)
        : g_(&g), start_(start), stop_(stop) {}
    const grid_key_dx<dim, int> &getStart() const { return start_; }
    const grid_key_dx<dim, int> &getStop() const { return stop_; }
    bool isNext() const
    {
        begin();
        return chunk_ < g_->host_chunks();
    }
    SparseGridGpu_iterator_sub &operator++()
    {
        begin();
        settle(true);
        return *this;
    }
    sparse_grid_gpu_index<SparseGridType> get() const
    {
        begin();
        return {*g_, (int)chunk_, local()};
    }

private:
    static constexpr int E = SparseGridType::blockEdgeSize_;
    void begin() const
    {
        if (started_) return;
        g_->require_host_copy();
        started_ = true;
        settle(false);
    }
    // chunk c's part of the box in local coordinates; false when it is empty
    bool enter(size_t c) const
    {
        const auto base = g_->getCoord(g_->host_chunk_id(c) * g_->getBlockSize());
        for (unsigned int d = 0; d < dim; ++d) {
            lo_[d] = std::max(start_.get(d), base.get(d)) - base.get(d);
            hi_[d] = std::min(stop_.get(d), base.get(d) + E - 1) - base.get(d);
            if (lo_[d] > hi_[d]) return false;
            cur_[d] = lo_[d];
        }
        return true;
    }
    bool step() const
    {
        for (unsigned int d = 0; d < dim; ++d) {
            if (++cur_[d] <= hi_[d]) return true;
            cur_[d] = lo_[d];
        }
        return false;
    }
    int local() const
    {
        int l = cur_[dim - 1];
        for (int d = (int)dim - 2; d >= 0; --d) l = l * E + cur_[d];
        return l;
    }
    void settle(bool advance) const
    {
        const size_t nc = g_->host_chunks(), cs = g_->getBlockSize();
        const uint8_t *m = g_->host_mask();
        while (chunk_ < nc) {
            if (!inside_) inside_ = enter(chunk_);
            else if (advance && !step()) inside_ = false;
            advance = false;
            if (inside_) {
                if (m[chunk_ * cs + local()] & 1) return;
                advance = true;
                continue;
            }
            ++chunk_;
        }
    }
    const SparseGridType *g_;
    grid_key_dx<dim, int> start_, stop_;
    mutable bool started_ = false, inside_ = false;
    mutable size_t chunk_ = 0;
    mutable int lo_[dim] = {}, hi_[dim] = {}, cur_[dim] = {};
};

// Host pack memory (the reference's ExtPreAlloc<CudaMemory> over a preallocated buffer):
// resize(req) after packCalculate, then pack/packFinalize write into it.
class PackMemory
{
public:
    void resize(size_t n) { buf_.assign(n, 0); }
    size_t size() const { return buf_.size(); }
    unsigned char *getPointer() { return buf_.data(); }
    const unsigned char *getPointer() const { return buf_.data(); }

private:
    std::vector<unsigned char> buf_;
};

// host memory of the whole-grid host pack (the reference's HeapMemory, openfpm_devices), and
// a view of it (ExtPreAlloc<HeapMemory>): pack / unpack read and write at a Pack_stat offset
class HeapMemory
{
public:
    bool allocate(size_t n)
    {
        buf_.assign(n, 0);
        return true;
    }
    bool resize(size_t n)
    {
        buf_.resize(n, 0);
        return true;
    }
    size_t size() const { return buf_.size(); }
    void *getPointer() { return buf_.data(); }
    const void *getPointer() const { return buf_.data(); }
    void incRef() {}
    void decRef() {}

private:
    std::vector<unsigned char> buf_;
};
template <typename Mem> class ExtPreAlloc
{
public:
    ExtPreAlloc(size_t n, Mem &m) : m_(&m)
    {
        if (m.size() < n) m.resize(n);
    }
    size_t size() const { return m_->size(); }
    unsigned char *getPointer() { return (unsigned char *)m_->getPointer(); }
    const unsigned char *getPointer() const { return (const unsigned char *)m_->getPointer(); }
    void incRef() {}
    void decRef() {}

private:
    Mem *m_;
};

// read position in a PackMemory (the reference's Pack_stat / Unpack_stat)
class Pack_stat
{
public:
    size_t getOffset() const { return off_; }
    void addOffset(size_t n) { off_ += n; }
    void reset() { off_ = 0; }

private:
    size_t off_ = 0;
};
typedef Pack_stat Unpack_stat;

// chunk edge and threadBlockSize defaults per dimension (SparseGridGpu.hpp:38-64)
template <unsigned int dim> struct default_edge { typedef sg::int_<2> type; };
template <> struct default_edge<1> { typedef sg::int_<256> type; typedef sg::int_<256> tb; };
template <> struct default_edge<2> { typedef sg::int_<16> type; typedef sg::int_<256> tb; };
template <> struct default_edge<3> { typedef sg::int_<8> type; typedef sg::int_<512> tb; };

// memory layouts of the reference's containers (memory_traits_inte: one array per property,
// the chunk-major layout this library implements; memory_traits_lin is declared so that
// code naming it compiles, and the grid rejects it)
template <typename T> struct memory_traits_inte { typedef T type; };
template <typename T> struct memory_traits_lin { typedef T type; };

// Chunk linearizers (src/Grid/Geometry/grid_smb.hpp, grid_zmb.hpp): the domain size in points
// and the chunk ids of grid_smb (row-major, x fastest) or grid_zmb (Morton: bit i of chunk
// coordinate d at bit dim * i + d).  Inside a chunk the offset is x fastest for both.
namespace sg {
template <unsigned int dim, unsigned int blockEdgeSize, typename indexT, bool morton> class chunk_linearizer
{
public:
    static constexpr bool is_morton = morton;
    static constexpr indexT blockSize = IntPow<blockEdgeSize, dim>::value;
    SPARSEGRIDGPU_HD chunk_linearizer() : blockSz{}, sz{} {}
    // from the domain size in points (grid_smb.hpp:58-83)
    template <typename I, typename = typename std::enable_if<std::is_integral<I>::value>::type>
    SPARSEGRIDGPU_HD chunk_linearizer(const I (&size)[dim])
    {
        for (unsigned int d = 0; d < dim; ++d) {
            sz[d] = (indexT)size[d];
            blockSz[d] = (indexT)(size[d] / blockEdgeSize + ((size[d] % blockEdgeSize) != 0));
        }
    }
    // a cube of domainBlockEdgeSize chunks per side (grid_smb.hpp:85-95)
    SPARSEGRIDGPU_HD explicit chunk_linearizer(size_t domainBlockEdgeSize)
    {
        for (unsigned int d = 0; d < dim; ++d) {
            blockSz[d] = (indexT)domainBlockEdgeSize;
            sz[d] = (indexT)domainBlockEdgeSize * blockEdgeSize;
        }
    }
#if defined(__HIPCC__)
    // chunk counts from a dim3 (grid_smb.hpp:112-148)
    SPARSEGRIDGPU_HD chunk_linearizer(const dim3 blockDimensions)
    {
        const unsigned int b[3] = {blockDimensions.x, blockDimensions.y, blockDimensions.z};
        for (unsigned int d = 0; d < dim; ++d) {
            blockSz[d] = (indexT)(d < 3 ? b[d] : 1u);
            sz[d] = blockSz[d] * (indexT)blockEdgeSize;
        }
    }
#endif
    SPARSEGRIDGPU_HD const indexT (&getSize() const)[dim] { return sz; }
    SPARSEGRIDGPU_HD const indexT &size(int i) const { return sz[i]; }
    SPARSEGRIDGPU_HD const indexT (&getBlockSize() const)[dim] { return blockSz; }
    // chunk id of chunk coordinates
    template <typename CoordT> SPARSEGRIDGPU_HD indexT BlockLinId(const grid_key_dx<dim, CoordT> &bc) const
    {
        if (morton) {
            uint64_t z = 0;
            for (unsigned int d = 0; d < dim; ++d)
                for (unsigned int i = 0; i < 64 / dim; ++i) z |= ((((uint64_t)bc.get(d)) >> i) & 1ull) << (dim * i + d);
            return (indexT)z;
        }
        indexT b = (indexT)bc.get(dim - 1);
        for (int d = (int)dim - 2; d >= 0; --d) b = b * blockSz[d] + (indexT)bc.get(d);
        return b;
    }
    // chunk coordinates of chunk id b
    SPARSEGRIDGPU_HD grid_key_dx<dim, int> BlockInvLinId(indexT b) const
    {
        grid_key_dx<dim, int> c;
        for (unsigned int d = 0; d < dim; ++d) {
            if (morton) {
                uint64_t v = 0;
                for (unsigned int i = 0; i < 64 / dim; ++i) v |= ((((uint64_t)b) >> (dim * i + d)) & 1ull) << i;
                c.set_d(d, (int)v);
            } else {
                c.set_d(d, (int)(b % blockSz[d]));
                b /= blockSz[d];
            }
        }
        return c;
    }
    // point id = chunk id * blockSize + x-fastest offset (grid_smb.hpp:175-189)
    template <typename CoordT> SPARSEGRIDGPU_HD indexT LinId(const grid_key_dx<dim, CoordT> &c) const
    {
        grid_key_dx<dim, CoordT> bc;
        indexT l = 0;
        for (int d = (int)dim - 1; d >= 0; --d) {
            bc.set_d(d, (CoordT)(c.get(d) / (CoordT)blockEdgeSize));
            l = l * blockEdgeSize + (indexT)(c.get(d) % (CoordT)blockEdgeSize);
        }
        return BlockLinId(bc) * blockSize + l;
    }
    SPARSEGRIDGPU_HD grid_key_dx<dim, int> InvLinId(indexT lin) const
    {
        grid_key_dx<dim, int> bc = BlockInvLinId(lin / blockSize), c;
        indexT l = lin % blockSize;
        for (unsigned int d = 0; d < dim; ++d) {
            c.set_d(d, (int)(bc.get(d) * (int)blockEdgeSize + (int)(l % blockEdgeSize)));
            l /= blockEdgeSize;
        }
        return c;
    }

protected:
    indexT blockSz[dim];
    indexT sz[dim];
};
}  // namespace sg

template <unsigned int dim, unsigned int blockEdgeSize, typename indexT = long int>
class grid_smb : public sg::chunk_linearizer<dim, blockEdgeSize, indexT, false>
{
public:
    using sg::chunk_linearizer<dim, blockEdgeSize, indexT, false>::chunk_linearizer;
    SPARSEGRIDGPU_HD grid_smb() {}
};
template <unsigned int dim, unsigned int blockEdgeSize, typename indexT = long int>
class grid_zmb : public sg::chunk_linearizer<dim, blockEdgeSize, indexT, true>
{
public:
    using sg::chunk_linearizer<dim, blockEdgeSize, indexT, true>::chunk_linearizer;
    SPARSEGRIDGPU_HD grid_zmb() {}
};

template <unsigned int dim, typename AggregateT, unsigned int blockEdgeSize = default_edge<dim>::type::value,
          unsigned int threadBlockSize = default_edge<dim>::tb::value, typename indexT = long int,
          template <typename> class layout_base = memory_traits_inte,
          typename linearizer = grid_smb<dim, blockEdgeSize, indexT>>
class SparseGridGpu
{
    static_assert(std::is_same<layout_base<int>, memory_traits_inte<int>>::value,
                  "the grid's arrays are one array per property (memory_traits_inte)");
    // chunk ids are Morton codes (grid_zmb, SparseGridGpu_z) or row-major (grid_smb)
    static constexpr bool zorder = linearizer::is_morton;

public:
    static constexpr unsigned int dims = dim;
    static constexpr unsigned int blockEdgeSize_ = blockEdgeSize;
    typedef typename AggregateT::type props_tuple;
    template <unsigned int p> using prop_type = typename std::tuple_element<p, props_tuple>::type;
    static constexpr unsigned int nprops = std::tuple_size<props_tuple>::value;

    static constexpr unsigned int chunk_size()
    {
        unsigned int s = 1;
        for (unsigned int i = 0; i < dim; ++i) s *= blockEdgeSize;
        return s;
    }

    // SparseGridGpu(const size_t (&res)[dim], r) (SparseGridGpu.hpp:1694-1698)
    explicit SparseGridGpu(const size_t (&res)[dim], unsigned int stencilSupportRadius = 1)
    {
        init_(res, stencilSupportRadius);
    }
    // SparseGridGpu(linearizer &, r) (SparseGridGpu.hpp:1704-1714): the domain of a linearizer
    explicit SparseGridGpu(const linearizer &gridGeometry, unsigned int stencilSupportRadius = 1)
    {
        size_t res[dim];
        for (unsigned int i = 0; i < dim; ++i) res[i] = (size_t)gridGeometry.getSize()[i];
        init_(res, stencilSupportRadius);
    }
    ~SparseGridGpu() { sg_destroy(h_); }
    SparseGridGpu(const SparseGridGpu &) = delete;
    SparseGridGpu &operator=(const SparseGridGpu &) = delete;
    SparseGridGpu(SparseGridGpu &&o) noexcept
        : pack_lo_(std::move(o.pack_lo_)), pack_hi_(std::move(o.pack_hi_)), host_copy_(o.host_copy_),
          host_ids_(std::move(o.host_ids_)), host_mask_(std::move(o.host_mask_)), h_(o.h_)
    {
        o.h_ = nullptr;
        o.host_copy_ = false;
        for (unsigned int i = 0; i < dim; ++i) {
            size_[i] = o.size_[i];
            bsz_[i] = o.bsz_[i];
        }
        for (unsigned int q = 0; q < (nprops > 0 ? nprops : 1); ++q) {
            host_prop_[q] = std::move(o.host_prop_[q]);
            host_has_[q] = o.host_has_[q];
            std::memcpy(bg_[q], o.bg_[q], sizeof(bg_[q]));
        }
        host_dirty_ = o.host_dirty_;
        nn_type_ = o.nn_type_;
    }

    sg_grid *handle() { return h_; }
    void setStream(void *hip_stream) { sg::check(sg_set_stream(h_, hip_stream)); }

    // -- geometry (grid_smb) --
    static constexpr unsigned int getBlockEdgeSize() { return blockEdgeSize; }
    constexpr unsigned int getBlockSize() const { return chunk_size(); }
    template <typename CoordT> indexT getLinId(const grid_key_dx<dim, CoordT> &c) const
    {
        indexT b = c.get(dim - 1) / blockEdgeSize, l = c.get(dim - 1) % blockEdgeSize;
        for (int d = (int)dim - 2; d >= 0; --d) {
            b = b * (indexT)bsz_[d] + c.get(d) / blockEdgeSize;
            l = l * blockEdgeSize + c.get(d) % blockEdgeSize;
        }
        if (zorder) {
            // lin_zid: bit i of chunk coordinate d at bit dim*i + d
            uint64_t z = 0;
            for (unsigned int d = 0; d < dim; ++d)
                for (unsigned int i = 0; i < 64 / dim; ++i)
                    z |= ((((uint64_t)(c.get(d) / blockEdgeSize)) >> i) & 1ull) << (dim * i + d);
            b = (indexT)z;
        }
        return b * (indexT)chunk_size() + l;
    }
    grid_key_dx<dim, int> getCoord(indexT lin) const
    {
        grid_key_dx<dim, int> c;
        indexT b = lin / chunk_size(), l = lin % chunk_size();
        for (unsigned int d = 0; d < dim; ++d) {
            indexT cb = b % (indexT)bsz_[d];
            if (zorder) {
                uint64_t v = 0;
                for (unsigned int i = 0; i < 64 / dim; ++i) v |= ((((uint64_t)(lin / chunk_size())) >> (dim * i + d)) & 1ull) << i;
                cb = (indexT)v;
            }
            c.set_d(d, (int)(cb * blockEdgeSize + l % blockEdgeSize));
            b /= (indexT)bsz_[d];
            l /= blockEdgeSize;
        }
        return c;
    }
    Box<dim, int> getBox() const
    {
        Box<dim, int> b;
        for (unsigned int i = 0; i < dim; ++i) {
            b.setLow(i, 0);
            b.setHigh(i, (int)size_[i]);
        }
        return b;
    }

    // -- background / insertion / flush --
    template <unsigned int p> void setBackgroundValue(prop_type<p> v)
    {
        sg::check(sg_set_background(h_, (int)p, &v));
        std::memcpy(bg_[p], &v, sizeof(v));
    }
    template <unsigned int p> prop_type<p> getBackgroundValue() const
    {
        prop_type<p> v;
        std::memcpy(&v, bg_[p], sizeof(v));
        return v;
    }
#if !defined(__HIPCC__)
    // host inserts only: the insert buffer grows on demand (the device insert buffer of
    // hipcc translation units is below)
    template <typename dim3T> void setGPUInsertBuffer(dim3T, dim3T) {}
#endif
    void preFlush() {}

    // point inserts (insert<p>(coord) once per point): keys are getLinId values, one
    // host value array per property (nullptr: the records carry the property's background,
    // which the flush merges like any value), masks default to EXIST
    void insertPoints(const indexT *keys, size_t n, const void *const *values, const uint8_t *masks = nullptr)
    {
        std::vector<int64_t> k(keys, keys + n);
        sg::check(sg_insert_points(h_, (int64_t)n, k.data(), masks, values, SG_MEM_HOST));
    }
    // chunk inserts (insertBlock(blockId)): masks and values are [n][blockSize]
    void insertBlocks(const indexT *chunk_ids, size_t n, const uint8_t *masks, const void *const *values)
    {
        std::vector<int64_t> k(chunk_ids, chunk_ids + n);
        sg::check(sg_insert_chunks(h_, (int64_t)n, k.data(), masks, values, SG_MEM_HOST));
    }

    // flush<v_reduce...>(ctx, opt) (SparseGridGpu.hpp:772-779).  FLUSH_ON_HOST: the
    // reference merges on the CPU and leaves the result in host memory
    // (map_vector_sparse.hpp:1375-1484); here the same merge runs on the device (same
    // result), with the host copy's edits (insertFlush) sent first and the host copy
    // refreshed afterwards, so that host reads see the merged grid either way.
    template <typename... v_reduce> void flush(gpu::ofp_context_t &, flush_type opt = FLUSH_ON_DEVICE)
    {
        int32_t ops[nprops > 0 ? nprops : 1];
        for (unsigned int i = 0; i < nprops; ++i) ops[i] = SG_OP_NONE;
        int dummy[] = {0, (ops[v_reduce::prop] = v_reduce::op, 0)...};
        (void)dummy;
        if (opt == FLUSH_ON_HOST && host_copy_ && host_dirty_) hostToDevice();
        sg::check(sg_flush(h_, ops));
        ins_view_ok_ = false;  // the flush took the records of the kernels' insert buffer
        if (opt == FLUSH_ON_HOST) deviceToHost();
    }

    // insertFlush<p>(coord) (SparseGridGpu.hpp:2780-2845, BlockMapGpu.hpp:186-222): insert the
    // point into the host copy (taken from the device, every property, when there is none): a
    // new chunk goes to its sorted place with mask 0 and the background, the point's mask
    // becomes EXIST (1, as the reference writes, clearing PADDING), its other properties keep
    // their values; returns the point's value of p in the host copy.  hostToDevice() sends the
    // edited copy (chunks, masks, properties) to the device; until then get<p>(coord) reads it.
    template <unsigned int p, typename CoordT> prop_type<p> &insertFlush(const grid_key_dx<dim, CoordT> &coord)
    {
        if (!host_copy_) deviceToHost();
        if (!host_has_[p]) {
            if (host_dirty_) throw sg::error("insertFlush: property not in the edited host copy");
            host_prop_[p].resize(host_mask_.size() * sizeof(prop_type<p>) + 1);
            if (!host_mask_.empty()) sg::check(sg_get_prop(h_, (int)p, host_prop_[p].data(), SG_MEM_HOST));
            host_has_[p] = true;
        }
        const indexT lin = getLinId(coord);
        const size_t s = host_chunk_(lin / (indexT)chunk_size(), true) * chunk_size() + (size_t)(lin % (indexT)chunk_size());
        host_mask_[s] = (uint8_t)SG_EXIST;
        host_dirty_ = true;
        return reinterpret_cast<prop_type<p> *>(host_prop_[p].data())[s];
    }
    // insertBlockFlush(coord, local_id) (SparseGridGpu.hpp:2799-2811): as insertFlush, returning
    // the chunk's host arrays (get<p>()[local_id]) and the point's offset in local_id
    struct host_block {
        SparseGridGpu *g;
        size_t base;
        template <unsigned int q> prop_type<q> *get() const
        {
            if (!g->host_has_[q]) throw sg::error("insertBlockFlush: property not in the host copy");
            return reinterpret_cast<prop_type<q> *>(g->host_prop_[q].data()) + base;
        }
        uint8_t *getMask() const { return g->host_mask_.data() + base; }
    };
    template <typename CoordT> host_block insertBlockFlush(const grid_key_dx<dim, CoordT> &coord, indexT &local_id)
    {
        if (!host_copy_) deviceToHost();
        const indexT lin = getLinId(coord);
        const size_t base = host_chunk_(lin / (indexT)chunk_size(), true) * chunk_size();
        local_id = lin % (indexT)chunk_size();
        host_mask_[base + (size_t)local_id] = (uint8_t)SG_EXIST;
        host_dirty_ = true;
        return host_block{this, base};
    }

    // -- neighbourhood, boundaries, stencils --
    template <typename NN = NNStar<dim>> void findNeighbours() { sg::check(sg_find_neighbours(h_, NN::type)); }
    template <typename NN> void setNNType()
    {
        sg::check(sg_set_nn_type(h_, NN::type));
        nn_type_ = NN::type;
    }
    template <typename NN = NNStar<dim>, typename checker_type = No_check>
    void tagBoundaries(gpu::ofp_context_t &, checker_type = checker_type(),
                       tag_boundaries opt = NO_CALCULATE_EXISTING_POINTS)
    {
        sg::check(sg_tag_boundaries(h_, NN::type, (int)opt));
    }
    // applyStencils<Stencil>(box, mode, args...) (SparseGridGpu.hpp:2704-2734).  A named
    // stencil of this header (one with a static desc(box, args...)) runs as a library kernel;
    // any other stencil is a user functor with the reference's protocol (typedef stencil_type;
    // supportRadius; __device__ static stencil(sparseGrid, chunkId, chunkPos, offset, coord,
    // dataBlockLoad, dataBlockStore, curMask, args...)), run by a kernel of
    // SparseGridGpu_device.hpp in translation units compiled by hipcc.  After a functor has
    // run, slots without a point are given the background again.
    template <typename Stencil, typename... Args> void applyStencils(const Box<dim, int> &box, StencilMode mode, Args... args)
    {
        if constexpr (sg::has_desc<Stencil, Box<dim, int>, Args...>::value) {
            sg_stencil_desc d = Stencil::desc(box, args...);
            d.mode = (int)mode;
            sg::check(sg_apply_stencil(h_, &d));
        } else {
#if defined(__HIPCC__)
            apply_functor_<Stencil>(box, mode, args...);
#else
            static_assert(sg::has_desc<Stencil, Box<dim, int>, Args...>::value,
                          "user stencil functors run in translation units compiled by hipcc");
#endif
        }
    }
    template <typename Stencil1, typename Stencil2, typename... Others, typename... Args>
    void applyStencils(const Box<dim, int> &box, StencilMode mode, Args... args)
    {
        applyStencils<Stencil1>(box, mode, args...);
        applyStencils<Stencil2, Others...>(box, mode, args...);
    }

#if defined(__HIPCC__)
    // -- device inserts (hipcc translation units) --
    // setGPUInsertBuffer (SparseGridGpu.hpp:2270-2278): room for nBlock x nSlot chunk
    // records (dim3 or counts) claimed by kernels that get toKernel(), until the next flush
    template <typename B3, typename S3> void setGPUInsertBuffer(B3 nBlock, S3 nSlot)
    {
        sg::check(sg_insert_device_end(h_));  // records of an earlier buffer join the inserts
        sg::check(sg_insert_device_begin(h_, count3(nBlock) * count3(nSlot), &ins_view_));
        ins_view_ok_ = true;
    }
    // toKernel() (SparseGridGpu.hpp:1716-1751): the device view: reads (get<p>, exist,
    // isPadding, get_sparse, getNNPoint), the stencil-functor helpers, and insertBlock /
    // insert<p> when setGPUInsertBuffer opened an insert buffer.  It stays valid until the
    // next call that changes the chunk set or a property array (flush, remove, ...).
    typedef sg::sparse_grid_gpu_ker<dim, blockEdgeSize, props_tuple, indexT> kernel_view;
    kernel_view toKernel() const
    {
        kernel_view k;
        std::memset((void *)&k, 0, sizeof(k));
        if (ins_view_ok_) k.iv = ins_view_;
        sg::check(sg_to_kernel(h_, SG_NN_FULL, &k.gv));
        for (unsigned int q = 0; q < 3; ++q) k.bsz[q] = q < dim ? (int64_t)bsz_[q] : 1;
        k.zorder = zorder ? 1 : 0;
        k.nn_star = nn_type_ == SG_NN_STAR;
        return k;
    }

    // -- conv family with user __device__ lambdas (SparseGridGpu.hpp:2568-2683), for
    //    translation units compiled by hipcc: include/SparseGridGpu_device.hpp.  The box
    //    [start, stop] is inclusive; stencil_size is the ghost width the lambda may reach.
    // res = func(cpb, i, j[, k])
    template <unsigned int prop_src, unsigned int prop_dst, unsigned int stencil_size, typename lambda_f,
              typename CoordT, typename... ArgsT>
    void conv(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, lambda_f func, ArgsT...)
    {
        static_assert(prop_dst != prop_src, "conv: the destination must not be the source");
        run_conv<stencil_size, 0, prop_src, prop_src, prop_src, prop_dst, prop_dst, prop_dst>(start, stop, func);
    }
    // res = func(cur, cross_stencil) (xm[d] / xp[d]: the face neighbours), extra arguments
    // appended
    template <unsigned int prop_src, unsigned int prop_dst, unsigned int stencil_size, typename lambda_f,
              typename CoordT, typename... ArgsT>
    void conv_cross(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, lambda_f func,
                    ArgsT... args)
    {
        static_assert(prop_dst != prop_src, "conv_cross: the destination must not be the source");
        run_conv<stencil_size, 1, prop_src, prop_src, prop_src, prop_dst, prop_dst, prop_dst>(
            start, stop, sg::conv_detail::with_args<lambda_f, ArgsT...>{func, {args...}});
    }
    // res = func(cpb, block, offset, i, j[, k]); block.template get<p>()[offset] reads the
    // chunk's own data of any property
    template <unsigned int prop_src, unsigned int prop_dst, unsigned int stencil_size, typename lambda_f,
              typename CoordT, typename... ArgsT>
    void conv_cross_b(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, lambda_f func,
                      ArgsT...)
    {
        static_assert(prop_dst != prop_src, "conv_cross_b: the destination must not be the source");
        run_conv<stencil_size, 2, prop_src, prop_src, prop_src, prop_dst, prop_dst, prop_dst>(start, stop, func);
    }
    // func(res1, res2, cpb1, cpb2, i, j[, k])
    template <unsigned int prop_src1, unsigned int prop_src2, unsigned int prop_dst1, unsigned int prop_dst2,
              unsigned int stencil_size, typename lambda_f, typename CoordT, typename... ArgsT>
    void conv2(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, lambda_f func, ArgsT...)
    {
        static_assert(prop_dst1 != prop_src1 && prop_dst1 != prop_src2 && prop_dst2 != prop_src1 &&
                          prop_dst2 != prop_src2 && prop_dst1 != prop_dst2,
                      "conv2: the destinations must differ from each other and from the sources");
        run_conv<stencil_size, 3, prop_src1, prop_src2, prop_src2, prop_dst1, prop_dst2, prop_dst2>(start, stop, func);
    }
    // func(res1, res2, cpb1, cpb2, block, offset, i, j[, k])
    template <unsigned int prop_src1, unsigned int prop_src2, unsigned int prop_dst1, unsigned int prop_dst2,
              unsigned int stencil_size, typename lambda_f, typename CoordT, typename... ArgsT>
    void conv2_b(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, lambda_f func,
                 ArgsT...)
    {
        static_assert(prop_dst1 != prop_src1 && prop_dst1 != prop_src2 && prop_dst2 != prop_src1 &&
                          prop_dst2 != prop_src2 && prop_dst1 != prop_dst2,
                      "conv2_b: the destinations must differ from each other and from the sources");
        run_conv<stencil_size, 4, prop_src1, prop_src2, prop_src2, prop_dst1, prop_dst2, prop_dst2>(start, stop, func);
    }
    // func(res1, res2, res3, cpb1, cpb2, cpb3, block, offset, i, j[, k])
    template <unsigned int prop_src1, unsigned int prop_src2, unsigned int prop_src3, unsigned int prop_dst1,
              unsigned int prop_dst2, unsigned int prop_dst3, unsigned int stencil_size, typename lambda_f,
              typename CoordT, typename... ArgsT>
    void conv3_b(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, lambda_f func,
                 ArgsT...)
    {
        static_assert(prop_dst1 != prop_src1 && prop_dst1 != prop_src2 && prop_dst1 != prop_src3 &&
                          prop_dst2 != prop_src1 && prop_dst2 != prop_src2 && prop_dst2 != prop_src3 &&
                          prop_dst3 != prop_src1 && prop_dst3 != prop_src2 && prop_dst3 != prop_src3 &&
                          prop_dst1 != prop_dst2 && prop_dst1 != prop_dst3 && prop_dst2 != prop_dst3,
                      "conv3_b: the destinations must differ from each other and from the sources");
        run_conv<stencil_size, 5, prop_src1, prop_src2, prop_src3, prop_dst1, prop_dst2, prop_dst3>(start, stop, func);
    }

    // the reference's forms with grid_key_dx<3> corners (SparseGridGpu.hpp:2568-2683), so that
    // braced corners such as {0, 0} compile in any dimension
    template <unsigned int prop_src, unsigned int prop_dst, unsigned int stencil_size, typename lambda_f, typename... ArgsT>
    void conv(grid_key_dx<3> start, grid_key_dx<3> stop, lambda_f func, ArgsT... args)
    {
        conv<prop_src, prop_dst, stencil_size>(key_of_(start), key_of_(stop), func, args...);
    }
    template <unsigned int prop_src, unsigned int prop_dst, unsigned int stencil_size, typename lambda_f, typename... ArgsT>
    void conv_cross(grid_key_dx<3> start, grid_key_dx<3> stop, lambda_f func, ArgsT... args)
    {
        conv_cross<prop_src, prop_dst, stencil_size>(key_of_(start), key_of_(stop), func, args...);
    }
    template <unsigned int prop_src, unsigned int prop_dst, unsigned int stencil_size, typename lambda_f, typename... ArgsT>
    void conv_cross_b(grid_key_dx<3> start, grid_key_dx<3> stop, lambda_f func, ArgsT... args)
    {
        conv_cross_b<prop_src, prop_dst, stencil_size>(key_of_(start), key_of_(stop), func, args...);
    }
    template <unsigned int prop_src1, unsigned int prop_src2, unsigned int prop_dst1, unsigned int prop_dst2,
              unsigned int stencil_size, typename lambda_f, typename... ArgsT>
    void conv2(grid_key_dx<3> start, grid_key_dx<3> stop, lambda_f func, ArgsT... args)
    {
        conv2<prop_src1, prop_src2, prop_dst1, prop_dst2, stencil_size>(key_of_(start), key_of_(stop), func, args...);
    }
    template <unsigned int prop_src1, unsigned int prop_src2, unsigned int prop_dst1, unsigned int prop_dst2,
              unsigned int stencil_size, typename lambda_f, typename... ArgsT>
    void conv2_b(grid_key_dx<3> start, grid_key_dx<3> stop, lambda_f func, ArgsT... args)
    {
        conv2_b<prop_src1, prop_src2, prop_dst1, prop_dst2, stencil_size>(key_of_(start), key_of_(stop), func, args...);
    }
    template <unsigned int prop_src1, unsigned int prop_src2, unsigned int prop_src3, unsigned int prop_dst1,
              unsigned int prop_dst2, unsigned int prop_dst3, unsigned int stencil_size, typename lambda_f,
              typename... ArgsT>
    void conv3_b(grid_key_dx<3> start, grid_key_dx<3> stop, lambda_f func, ArgsT... args)
    {
        conv3_b<prop_src1, prop_src2, prop_src3, prop_dst1, prop_dst2, prop_dst3, stencil_size>(key_of_(start),
                                                                                                key_of_(stop), func,
                                                                                                args...);
    }

private:
    // (int corners: the forms above then call the deduced-corner forms, never themselves)
    static grid_key_dx<dim, int> key_of_(const grid_key_dx<3> &k)
    {
        grid_key_dx<dim, int> r;
        for (unsigned int d = 0; d < dim; ++d) r.set_d(d, (int)k.get(d));
        return r;
    }
    static int64_t count3(const dim3 &d) { return (int64_t)d.x * d.y * d.z; }
    static int64_t count3(int64_t n) { return n; }
    sg_insert_view ins_view_ = {};
    template <unsigned int R, int MODE, unsigned int s1, unsigned int s2, unsigned int s3, unsigned int d1,
              unsigned int d2, unsigned int d3, typename CoordT, typename F>
    void run_conv(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop, F f)
    {
        static_assert(s1 < nprops && s2 < nprops && s3 < nprops && d1 < nprops && d2 < nprops && d3 < nprops,
                      "property index out of range");
        int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (unsigned int d = 0; d < dim; ++d) {
            lo[d] = (int64_t)start.get(d);
            hi[d] = (int64_t)stop.get(d);
        }
        sg::check(sg::conv_detail::launch<dim, blockEdgeSize, R, MODE, props_tuple, s1, s2, s3, d1, d2, d3>(h_, lo, hi,
                                                                                                          f));
    }

public:
#endif

    // -- access --
    // get<p>(coord) (SparseGridGpu.hpp:1897-1901): the point's value, the background where it
    // does not exist; read from the host copy while it holds edits not yet sent by
    // hostToDevice (insertFlush), else from the device
    template <unsigned int p, typename CoordT> prop_type<p> get(const grid_key_dx<dim, CoordT> &coord) const
    {
        int64_t k = (int64_t)getLinId(coord);
        prop_type<p> v;
        if (host_dirty_ && host_has_[p]) {
            const size_t c = const_cast<SparseGridGpu *>(this)->host_chunk_(k / (int64_t)chunk_size(), false);
            const size_t s = c * chunk_size() + (size_t)(k % (int64_t)chunk_size());
            if (c < host_ids_.size() && (host_mask_[s] & SG_EXIST))
                std::memcpy(&v, host_prop_[p].data() + s * sizeof(v), sizeof(v));
            else
                std::memcpy(&v, bg_[p], sizeof(v));
            return v;
        }
        sg::check(sg_get_points(h_, (int)p, 1, &k, &v, SG_MEM_HOST));
        return v;
    }
    size_t size() const { return (size_t)sg_num_chunks(h_) * chunk_size(); }
    std::vector<indexT> private_get_index_array() const
    {
        std::vector<int64_t> k((size_t)sg_num_chunks(h_));
        if (!k.empty()) sg::check(sg_get_chunk_ids(h_, k.data(), SG_MEM_HOST));
        return std::vector<indexT>(k.begin(), k.end());
    }
    template <unsigned int p> std::vector<prop_type<p>> private_get_data_array() const
    {
        std::vector<prop_type<p>> v((size_t)sg_num_chunks(h_) * chunk_size());
        if (!v.empty()) sg::check(sg_get_prop(h_, (int)p, v.data(), SG_MEM_HOST));
        return v;
    }
    std::vector<uint8_t> private_get_mask_array() const
    {
        std::vector<uint8_t> m((size_t)sg_num_chunks(h_) * chunk_size());
        if (!m.empty()) sg::check(sg_get_masks(h_, m.data(), SG_MEM_HOST));
        return m;
    }
    template <unsigned int p> void hostToDevice(const std::vector<prop_type<p>> &v)
    {
        sg::check(sg_set_prop(h_, (int)p, v.data(), SG_MEM_HOST));
    }

    // -- host copy (deviceToHost / hostToDevice, BlockMapGpu.hpp:230-241) --
    // deviceToHost<p...>() copies the chunk list, the masks and properties p... (every
    // property for an empty list) to host memory. The host iterators, get<p> and getFlag
    // of a sparse_grid_gpu_index read it, get<p> of an index writes it, and
    // hostToDevice<p...>() sends properties p... (every copied one for an empty list) back.
    template <unsigned int... p> void deviceToHost()
    {
        sg::check(sg_synchronize(h_));
        host_dirty_ = false;
        for (unsigned int q = 0; q < (nprops > 0 ? nprops : 1); ++q) host_has_[q] = false;
        host_ids_ = private_get_index_array();
        host_mask_ = private_get_mask_array();
        for (unsigned int q = 0; q < nprops; ++q)
            if (sizeof...(p) == 0 || listed<p...>(q)) {
                host_prop_[q].resize(host_mask_.size() * prop_bytes(q) + 1);
                if (!host_mask_.empty()) sg::check(sg_get_prop(h_, (int)q, host_prop_[q].data(), SG_MEM_HOST));
                host_has_[q] = true;
            }
        host_copy_ = true;
    }
    // hostToDevice<p...>() (BlockMapGpu.hpp:230-241): the chunk list and the masks of the host
    // copy (with the chunks insertFlush added), then properties p... (every property of the
    // copy for an empty list); device properties not sent keep their values per chunk
    template <unsigned int... p> void hostToDevice()
    {
        require_host_copy();
        for (unsigned int q = 0; q < nprops; ++q)
            if (sizeof...(p) != 0 && listed<p...>(q) && !host_has_[q])
                throw sg::error("hostToDevice: property not in the host copy");
        std::vector<int64_t> ids(host_ids_.begin(), host_ids_.end());
        sg::check(sg_set_chunks(h_, (int64_t)ids.size(), ids.data(), host_mask_.data(), SG_MEM_HOST));
        for (unsigned int q = 0; q < nprops; ++q)
            if ((sizeof...(p) == 0 && host_has_[q]) || (sizeof...(p) != 0 && listed<p...>(q)))
                if (!host_mask_.empty()) sg::check(sg_set_prop(h_, (int)q, host_prop_[q].data(), SG_MEM_HOST));
        host_dirty_ = false;
    }
    template <unsigned int p> prop_type<p> &get(const sparse_grid_gpu_index<SparseGridGpu> &k)
    {
        return host_value<p>(k);
    }
    template <unsigned int p> const prop_type<p> &get(const sparse_grid_gpu_index<SparseGridGpu> &k) const
    {
        return const_cast<SparseGridGpu *>(this)->host_value<p>(k);
    }
    unsigned char getFlag(const sparse_grid_gpu_index<SparseGridGpu> &k) const
    {
        require_host_copy();
        return host_mask_[(size_t)k.get_cnk_pos_id() * chunk_size() + k.get_data_id()];
    }
    typedef SparseGridGpu_iterator<dim, SparseGridGpu> iterator;
    iterator getIterator() const
    {
        require_host_copy();
        return iterator(*this);
    }
    // host-copy accessors of the iterators
    size_t host_chunks() const { return host_ids_.size(); }
    const uint8_t *host_mask() const { return host_mask_.data(); }
    indexT host_chunk_id(size_t c) const { return host_ids_[c]; }
    void require_host_copy() const
    {
        if (!host_copy_) throw sg::error("no host copy of the grid: call deviceToHost() first");
    }

    size_t countExistingElements() const
    {
        int64_t c = 0;
        sg::check(sg_count_existing(h_, &c));
        return (size_t)c;
    }
    size_t countBoundaryElements() const
    {
        int64_t c = 0;
        sg::check(sg_count_boundary(h_, &c));
        return (size_t)c;
    }
    // mean and standard deviation of the per-chunk share of EXIST && !PADDING points
    void measureBlockOccupancy(double &mean, double &deviation) const
    {
        std::vector<uint8_t> m = private_get_mask_array();
        const size_t nb = m.size() / chunk_size();
        std::vector<double> occ(nb, 0.0);
        for (size_t b = 0; b < nb; ++b) {
            size_t c = 0;
            for (size_t o = 0; o < chunk_size(); ++o) c += (m[b * chunk_size() + o] & 3) == 1;
            occ[b] = (double)c / chunk_size();
        }
        mean = 0.0;
        for (double o : occ) mean += o;
        mean = nb ? mean / nb : 0.0;
        double v = 0.0;
        for (double o : occ) v += (o - mean) * (o - mean);
        deviation = nb > 1 ? std::sqrt(v / (nb - 1)) : 0.0;
    }

    // -- whole-grid pack into host memory (packRequest(req) / pack(mem, sts) / unpack(mem, ps),
    //    SparseGridGpu.hpp:2895-2959): the reference serialises its index buffer, its data
    //    buffer and the linearizer.  Here, in host byte order, each part 8-byte aligned:
    //      uint64 n | indexT ids[n]                                 (the chunk list)
    //      uint64 n | uint64 k | int32 props[k] | T_q[n * chunk size] per listed property q
    //               | uint8 mask[n * chunk size]                    (the data)
    //      uint64 dim | uint64 size[dim]                            (the domain)
    //    with every property when prp... is empty.  The mask travels always (the reference's
    //    data buffer carries it only when every property is packed).  unpack replaces this
    //    grid's chunk list and masks with the packed ones (the domains must be equal), sets
    //    the packed properties and keeps the others per chunk (the background in new chunks).
    template <int... prp> void packRequest(size_t &req) const
    {
        const size_t n = (size_t)sg_num_chunks(h_), cs = chunk_size();
        const std::vector<int> pl = packed_props_<prp...>();
        size_t b = 8 + align8_(n * sizeof(indexT)) + 16 + align8_(4 * pl.size());
        for (int q : pl) b += align8_(n * cs * prop_bytes((unsigned)q));
        b += align8_(n * cs) + 8 + 8 * dim;
        req += b;
    }
    template <int... prp, typename Mem> void pack(ExtPreAlloc<Mem> &mem, Pack_stat &sts) const
    {
        size_t req = 0;
        packRequest<prp...>(req);
        if (sts.getOffset() + req > mem.size()) throw sg::error("pack: host memory too small");
        unsigned char *out = mem.getPointer() + sts.getOffset();
        std::memset(out, 0, req);
        const std::vector<indexT> ids = private_get_index_array();
        const uint64_t n = ids.size();
        const size_t cs = chunk_size();
        size_t at = 0;
        put_(out, at, &n, 8);
        put_(out, at, ids.data(), n * sizeof(indexT));
        const std::vector<int> pl = packed_props_<prp...>();
        const uint64_t k = pl.size();
        put_(out, at, &n, 8);
        put_(out, at, &k, 8);
        put_(out, at, pl.data(), 4 * pl.size());
        for (int q : pl) {
            if (n) sg::check(sg_get_prop(h_, q, out + at, SG_MEM_HOST));
            at += align8_(n * cs * prop_bytes((unsigned)q));
        }
        if (n) sg::check(sg_get_masks(h_, out + at, SG_MEM_HOST));
        at += align8_(n * cs);
        const uint64_t d = dim;
        put_(out, at, &d, 8);
        for (unsigned int i = 0; i < dim; ++i) {
            const uint64_t v = size_[i];
            put_(out, at, &v, 8);
        }
        sts.addOffset(at);
    }
    template <int... prp, typename Mem> void unpack(ExtPreAlloc<Mem> &mem, Unpack_stat &ps)
    {
        const unsigned char *in = mem.getPointer() + ps.getOffset();
        const size_t avail = mem.size() - ps.getOffset();
        size_t at = 0;
        uint64_t n = 0, n2 = 0, k = 0;
        get_(in, avail, at, &n, 8);
        std::vector<indexT> ids((size_t)n);
        get_(in, avail, at, ids.data(), n * sizeof(indexT));
        get_(in, avail, at, &n2, 8);
        get_(in, avail, at, &k, 8);
        if (n2 != n || k > nprops) throw sg::error("unpack: malformed packed grid");
        std::vector<int> pl((size_t)k);
        get_(in, avail, at, pl.data(), 4 * k);
        const size_t cs = chunk_size();
        std::vector<const unsigned char *> src((size_t)k);
        for (size_t i = 0; i < k; ++i) {
            if (pl[i] < 0 || pl[i] >= (int)nprops) throw sg::error("unpack: malformed packed grid");
            src[i] = in + at;
            skip_(avail, at, n * cs * prop_bytes((unsigned)pl[i]));
        }
        const unsigned char *masks = in + at;
        skip_(avail, at, n * cs);
        uint64_t d = 0;
        get_(in, avail, at, &d, 8);
        if (d != dim) throw sg::error("unpack: packed grid of another dimension");
        for (unsigned int i = 0; i < dim; ++i) {
            uint64_t v = 0;
            get_(in, avail, at, &v, 8);
            if (v != size_[i]) throw sg::error("unpack: packed grid of another domain size");
        }
        std::vector<int64_t> ids64(ids.begin(), ids.end());
        sg::check(sg_set_chunks(h_, (int64_t)n, ids64.data(), masks, SG_MEM_HOST));
        const std::vector<int> want = packed_props_<prp...>();
        for (size_t i = 0; i < k; ++i)
            if (n && std::find(want.begin(), want.end(), pl[i]) != want.end())
                sg::check(sg_set_prop(h_, pl[i], src[i], SG_MEM_HOST));
        host_copy_ = false;
        host_dirty_ = false;
        ps.addOffset(at);
    }

    // -- pack / unpack / remove (SparseGridGpu.hpp:3040-3733) --
    typedef SparseGridGpu_iterator_sub<dim, SparseGridGpu> sub_iterator;
    sub_iterator getIterator(const grid_key_dx<dim, int> &start, const grid_key_dx<dim, int> &stop) const
    {
        return sub_iterator(*this, start, stop);
    }
    template <typename CoordT>
    sub_iterator getIterator(const grid_key_dx<dim, CoordT> &start, const grid_key_dx<dim, CoordT> &stop) const
    {
        grid_key_dx<dim, int> a, b;
        for (unsigned int d = 0; d < dim; ++d) {
            a.set_d(d, (int)start.get(d));
            b.set_d(d, (int)stop.get(d));
        }
        return sub_iterator(*this, a, b);
    }
    void packReset()
    {
        pack_lo_.clear();
        pack_hi_.clear();
    }
    // queue a box to pack (req is left as is, as in the reference)
    template <int... prp> void packRequest(const sub_iterator &it, size_t &)
    {
        for (unsigned int d = 0; d < dim; ++d) {
            pack_lo_.push_back(it.getStart().get(d));
            pack_hi_.push_back(it.getStop().get(d));
        }
    }
    // add the bytes of every queued box to req
    template <int... prp> void packCalculate(size_t &req, gpu::ofp_context_t &)
    {
        const int32_t props[] = {prp..., 0};
        int64_t b = 0;
        sg::check(sg_pack_size(h_, npack(), pack_lo_.data(), pack_hi_.data(), (int)sizeof...(prp), props,
                               (int)sizeof(indexT), &b));
        req += (size_t)b;
    }
    // the reference reserves the box's space here; the boxes are written by packFinalize
    template <int... prp> void pack(PackMemory &, const sub_iterator &, Pack_stat &) {}
    template <int... prp> void packFinalize(PackMemory &mem, Pack_stat &sts, int = 0, bool = false)
    {
        const int32_t props[] = {prp..., 0};
        int64_t b = 0;
        sg::check(sg_pack_size(h_, npack(), pack_lo_.data(), pack_hi_.data(), (int)sizeof...(prp), props,
                               (int)sizeof(indexT), &b));
        if (sts.getOffset() + (size_t)b > mem.size()) throw sg::error("pack memory too small");
        sg::check(sg_pack(h_, npack(), pack_lo_.data(), pack_hi_.data(), (int)sizeof...(prp), props,
                          (int)sizeof(indexT), mem.getPointer() + sts.getOffset(), b, SG_MEM_HOST));
        sts.addOffset((size_t)b);
    }
    // queue a box whose existing points the next removeAddUnpackFinalize removes
    template <typename CoordT> void remove(const Box<dim, CoordT> &section)
    {
        int32_t lo[dim], hi[dim];
        for (unsigned int d = 0; d < dim; ++d) {
            lo[d] = (int32_t)section.getLow(d);
            hi[d] = (int32_t)section.getHigh(d);
        }
        sg::check(sg_remove_box(h_, lo, hi));
    }
    // the device remove(coord) of the reference for a batch of points
    void removePoints(const indexT *keys, size_t n)
    {
        std::vector<int64_t> k(keys, keys + n);
        sg::check(sg_remove_points(h_, (int64_t)n, k.data(), SG_MEM_HOST));
    }
    void removeAddUnpackReset() { sg::check(sg_remove_add_unpack_reset(h_)); }
    // unpack the next packed box of mem (at ps) so that it starts at sub's start
    template <unsigned int... prp>
    void unpack(PackMemory &mem, const sub_iterator &sub, Unpack_stat &ps, gpu::ofp_context_t &)
    {
        const int32_t props[] = {(int32_t)prp..., 0};
        int32_t lo[dim];
        for (unsigned int d = 0; d < dim; ++d) lo[d] = sub.getStart().get(d);
        int64_t used = 0;
        sg::check(sg_unpack(h_, mem.getPointer() + ps.getOffset(), (int64_t)(mem.size() - ps.getOffset()),
                            SG_MEM_HOST, lo, (int)sizeof...(prp), props, (int)sizeof(indexT), &used));
        ps.addOffset((size_t)used);
    }
    template <unsigned int... prp> void removeAddUnpackFinalize(gpu::ofp_context_t &, int)
    {
        const int32_t props[] = {(int32_t)prp..., 0};
        sg::check(sg_remove_add_unpack_finalize(h_, (int)sizeof...(prp), props));
    }
    // copy_to (SparseGridGpu.hpp:3550-3559): queue, on grid_src, a copy of its existing
    // points in box_src into this grid, box_src's low corner landing at box_dst's
    template <typename CoordT>
    void copy_to(SparseGridGpu &grid_src, const Box<dim, CoordT> &box_src, const Box<dim, CoordT> &box_dst)
    {
        int32_t lo[dim], hi[dim], dl[dim];
        for (unsigned int d = 0; d < dim; ++d) {
            lo[d] = (int32_t)box_src.getLow(d);
            hi[d] = (int32_t)box_src.getHigh(d);
            dl[d] = (int32_t)box_dst.getLow(d);
        }
        sg::check(sg_copy_to(h_, grid_src.h_, lo, hi, dl));
    }
    // -- links between levels (construct_link_dw / construct_link_up, SparseGridGpu.hpp:2090-2262) --
    // padding points of this grid inside db (chunk by chunk, by offset inside a chunk);
    // down: their existing children 2 x + p_dw + ((j >> k) & 1)_k, j < 2 dim, in grid_dw;
    // up: their existing parent (x - p_up) / 2 in grid_up
    struct link_entry {
        int chunk;        // chunk position in the other grid
        short int offset; // offset inside the chunk
    };
    template <typename PointT>
    void construct_link_dw(SparseGridGpu &grid_dw, const Box<dim, int> &db, const PointT &p_dw, gpu::ofp_context_t &)
    {
        construct_link_(grid_dw, SG_LINK_DOWN, db, p_dw);
    }
    template <typename PointT>
    void construct_link_up(SparseGridGpu &grid_up, const Box<dim, int> &db, const PointT &p_up, gpu::ofp_context_t &)
    {
        construct_link_(grid_up, SG_LINK_UP, db, p_up);
    }
    // per padding point the first of its links (getDownLinksOffsets / getUpLinksOffsets)
    std::vector<unsigned int> getDownLinksOffsets() const { return link_offsets_(SG_LINK_DOWN); }
    std::vector<unsigned int> getUpLinksOffsets() const { return link_offsets_(SG_LINK_UP); }
    std::vector<link_entry> getDownLinks() const { return links_(SG_LINK_DOWN); }
    std::vector<link_entry> getUpLinks() const { return links_(SG_LINK_UP); }
    // the padding points the links start from: (chunk position, offset) in this grid
    std::vector<link_entry> getLinkPaddingPoints(int which) const
    {
        int64_t np = 0, nl = 0;
        sg::check(sg_link_sizes(h_, which, &np, &nl));
        std::vector<int32_t> c((size_t)np);
        std::vector<int16_t> o((size_t)np);
        sg::check(sg_get_links(h_, which, c.data(), o.data(), nullptr, nullptr, nullptr, SG_MEM_HOST));
        std::vector<link_entry> out((size_t)np);
        for (size_t i = 0; i < out.size(); ++i) out[i] = {c[i], o[i]};
        return out;
    }

    // removeCopyToFinalize (SparseGridGpu.hpp:3314-3329): PHASE1 on every grid taking part,
    // then PHASE2 on every grid, then PHASE3 on every grid
    template <unsigned int... prp> void removeCopyToFinalize(gpu::ofp_context_t &, int opt)
    {
        const int32_t props[] = {(int32_t)prp..., 0};
        sg::check(sg_remove_copy_to_finalize(h_, opt, (int)sizeof...(prp), props));
    }

private:
    static size_t align8_(size_t b) { return (b + 7) & ~(size_t)7; }
    static void put_(unsigned char *out, size_t &at, const void *src, size_t bytes)
    {
        if (bytes) std::memcpy(out + at, src, bytes);
        at += align8_(bytes);
    }
    static void get_(const unsigned char *in, size_t avail, size_t &at, void *dst, size_t bytes)
    {
        if (at + bytes > avail) throw sg::error("unpack: packed grid truncated");
        if (bytes) std::memcpy(dst, in + at, bytes);
        at += align8_(bytes);
    }
    static void skip_(size_t avail, size_t &at, size_t bytes)
    {
        if (at + bytes > avail) throw sg::error("unpack: packed grid truncated");
        at += align8_(bytes);
    }
    template <int... prp> static std::vector<int> packed_props_()
    {
        std::vector<int> pl = {prp...};
        if (pl.empty())
            for (unsigned int q = 0; q < nprops; ++q) pl.push_back((int)q);
        return pl;
    }
    int npack() const { return (int)(pack_lo_.size() / dim); }
    std::vector<int32_t> pack_lo_, pack_hi_;

    template <typename PointT> void construct_link_(SparseGridGpu &other, int which, const Box<dim, int> &db, const PointT &p)
    {
        int32_t lo[dim], hi[dim], pp[dim];
        for (unsigned int d = 0; d < dim; ++d) {
            lo[d] = db.getLow(d);
            hi[d] = db.getHigh(d);
            pp[d] = (int32_t)p.get(d);
        }
        sg::check(sg_construct_link(h_, other.h_, which, lo, hi, pp));
    }
    std::vector<unsigned int> link_offsets_(int which) const
    {
        int64_t np = 0, nl = 0;
        sg::check(sg_link_sizes(h_, which, &np, &nl));
        std::vector<uint32_t> sc((size_t)np + 1);
        sg::check(sg_get_links(h_, which, nullptr, nullptr, sc.data(), nullptr, nullptr, SG_MEM_HOST));
        return std::vector<unsigned int>(sc.begin(), sc.end() - 1);  // the reference drops the total
    }
    std::vector<link_entry> links_(int which) const
    {
        int64_t np = 0, nl = 0;
        sg::check(sg_link_sizes(h_, which, &np, &nl));
        std::vector<int32_t> c((size_t)nl);
        std::vector<int16_t> o((size_t)nl);
        sg::check(sg_get_links(h_, which, nullptr, nullptr, nullptr, c.data(), o.data(), SG_MEM_HOST));
        std::vector<link_entry> out((size_t)nl);
        for (size_t i = 0; i < out.size(); ++i) out[i] = {c[i], o[i]};
        return out;
    }

    template <unsigned int... p> static bool listed(unsigned int q)
    {
        const unsigned int l[] = {p..., ~0u};
        for (unsigned int i = 0; i < sizeof...(p); ++i)
            if (l[i] == q) return true;
        return false;
    }
    template <size_t... I> static size_t prop_bytes_(unsigned int q, std::index_sequence<I...>)
    {
        const size_t b[] = {sizeof(typename std::tuple_element<I, props_tuple>::type)..., 0};
        return b[q];
    }
    static size_t prop_bytes(unsigned int q) { return prop_bytes_(q, std::make_index_sequence<nprops>()); }
    template <unsigned int p> prop_type<p> &host_value(const sparse_grid_gpu_index<SparseGridGpu> &k)
    {
        require_host_copy();
        if (!host_has_[p]) throw sg::error("property not in the host copy: deviceToHost<p>() first");
        const size_t i = (size_t)k.get_cnk_pos_id() * chunk_size() + k.get_data_id();
        return reinterpret_cast<prop_type<p> *>(host_prop_[p].data())[i];
    }
    // position of chunk id b in the host copy; when absent, the position it would take, where
    // insert == true puts it (mask 0, the background in every copied property)
    size_t host_chunk_(indexT b, bool insert)
    {
        const auto it = std::lower_bound(host_ids_.begin(), host_ids_.end(), b);
        const size_t c = (size_t)(it - host_ids_.begin());
        if (it != host_ids_.end() && *it == b) return c;
        if (!insert) return host_ids_.size();
        const size_t cs = chunk_size();
        host_ids_.insert(it, b);
        host_mask_.insert(host_mask_.begin() + c * cs, cs, (uint8_t)0);
        for (unsigned int q = 0; q < nprops; ++q) {
            if (!host_has_[q]) continue;
            const size_t ts = prop_bytes(q);
            std::vector<unsigned char> bgc(cs * ts);
            for (size_t i = 0; i < cs; ++i) std::memcpy(bgc.data() + i * ts, bg_[q], ts);
            host_prop_[q].insert(host_prop_[q].begin() + c * cs * ts, bgc.begin(), bgc.end());
        }
        return c;
    }
    void init_(const size_t (&res)[dim], unsigned int stencilSupportRadius)
    {
        int64_t sz[3] = {1, 1, 1};
        for (unsigned int i = 0; i < dim; ++i) {
            sz[i] = (int64_t)res[i];
            size_[i] = res[i];
            bsz_[i] = res[i] / blockEdgeSize + ((res[i] % blockEdgeSize) != 0);
        }
        int32_t dts[nprops > 0 ? nprops : 1];
        fill_dtypes(dts, std::make_index_sequence<nprops>());
        sg::check(sg_create((int)dim, sz, (int)blockEdgeSize, (int)nprops, dts, (int)stencilSupportRadius, &h_));
        if (zorder) sg::check(sg_set_linearizer(h_, SG_LIN_ZMB));
    }
#if defined(__HIPCC__)
    template <typename Stencil, typename... Args> void apply_functor_(const Box<dim, int> &box, StencilMode mode, Args... args)
    {
        static_assert(Stencil::supportRadius <= 1, "stencil functors of support radius 1");
        static_assert(chunk_size() <= 1024, "one thread per chunk slot: at most 1024 points per chunk");
        kernel_view k = toKernel();
        k.nn_star = Stencil::stencil_type::type == SG_NN_STAR;
        if (k.gv.nchunks == 0) return;
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (unsigned int d = 0; d < dim; ++d) {
            lo[d] = box.getLow(d);
            hi[d] = box.getHigh(d);
        }
        hipStream_t st = (hipStream_t)k.gv.stream;
        if (mode == STENCIL_MODE_INPLACE_NO_SHARED) {
            int64_t n = 0;
            sg::check(sg_existing_points(h_, &n));
            if (n > 0) {
                int64_t *ep = nullptr;
                if (hipMalloc((void **)&ep, (size_t)n * sizeof(int64_t)) != hipSuccess)
                    throw sg::error("applyStencils: point list allocation failed");
                const int rc = sg_get_existing_points(h_, ep, SG_MEM_DEVICE);
                if (rc == SG_OK) {
                    hipLaunchKernelGGL((sg::apply_stencil_functor_noshared_kernel<Stencil, kernel_view, Args...>),
                                       dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, k, (const int64_t *)ep, n,
                                       lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], args...);
                }
                const hipError_t le = hipGetLastError();
                (void)hipStreamSynchronize(st);
                (void)hipFree(ep);
                sg::check(rc);
                if (le != hipSuccess) throw sg::error(std::string("applyStencils: ") + hipGetErrorString(le));
            }
        } else {
            hipLaunchKernelGGL((sg::apply_stencil_functor_kernel<Stencil, kernel_view, Args...>),
                               dim3((unsigned)k.gv.nchunks), dim3(chunk_size()), 0, st, k, lo[0], lo[1], lo[2], hi[0],
                               hi[1], hi[2], args...);
            const hipError_t le = hipGetLastError();
            if (le != hipSuccess) throw sg::error(std::string("applyStencils: ") + hipGetErrorString(le));
        }
        sg::check(sg_restore_background(h_, -1));
    }
#endif
    bool host_copy_ = false;
    bool host_dirty_ = false;  // the host copy holds insertFlush edits not yet on the device
    int nn_type_ = SG_NN_STAR;
    unsigned char bg_[nprops > 0 ? nprops : 1][8] = {};
    std::vector<indexT> host_ids_;
    std::vector<uint8_t> host_mask_;
    std::vector<unsigned char> host_prop_[nprops > 0 ? nprops : 1];
    bool host_has_[nprops > 0 ? nprops : 1] = {};

    template <size_t... I> static void fill_dtypes(int32_t *dts, std::index_sequence<I...>)
    {
        int dummy[] = {0, (dts[I] = sg::dtype_of<typename std::tuple_element<I, props_tuple>::type>::value, 0)...};
        (void)dummy;
        (void)dts;
    }

    sg_grid *h_ = nullptr;
    size_t size_[dim];
    size_t bsz_[dim];
    bool ins_view_ok_ = false;
};

// SparseGridGpu_z / SparseGridGpu_zi (Morton chunk ids) and SparseGridGpu_i (int indices)
// (src/SparseGridGpu/SparseGridGpu.hpp:3947-3972)
template <unsigned int dim, typename AggregateT, unsigned int blockEdgeSize = default_edge<dim>::type::value,
          unsigned int threadBlockSize = default_edge<dim>::tb::value, typename indexT = long int,
          template <typename> class layout_base = memory_traits_inte,
          typename linearizer = grid_zmb<dim, blockEdgeSize, indexT>>
using SparseGridGpu_z = SparseGridGpu<dim, AggregateT, blockEdgeSize, threadBlockSize, indexT, layout_base, linearizer>;
template <unsigned int dim, typename AggregateT, unsigned int blockEdgeSize = default_edge<dim>::type::value,
          unsigned int threadBlockSize = default_edge<dim>::tb::value, typename indexT = int,
          template <typename> class layout_base = memory_traits_inte,
          typename linearizer = grid_zmb<dim, blockEdgeSize, indexT>>
using SparseGridGpu_zi = SparseGridGpu<dim, AggregateT, blockEdgeSize, threadBlockSize, indexT, layout_base, linearizer>;
template <unsigned int dim, typename AggregateT, unsigned int blockEdgeSize = default_edge<dim>::type::value,
          unsigned int threadBlockSize = default_edge<dim>::tb::value, typename indexT = int,
          template <typename> class layout_base = memory_traits_inte,
          typename linearizer = grid_smb<dim, blockEdgeSize, indexT>>
using SparseGridGpu_i = SparseGridGpu<dim, AggregateT, blockEdgeSize, threadBlockSize, indexT, layout_base, linearizer>;

#endif  // SPARSEGRIDGPU_FACADE_HPP
