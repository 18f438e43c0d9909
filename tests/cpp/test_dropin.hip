// Reference tests restated through the facade compiled by hipcc, with the reference's call
// forms: testFlushInsert (SparseGridGpu_tests.cu:802-880: host insertFlush, host get,
// hostToDevice, get<p> on toKernel() in a kernel), testStencil_lap_simplified (:521-576:
// grid_smb constructor, a device insert kernel, applyStencils with the boundary functor
// BoundaryStencilSetXRescaled, conv_cross Jacobi steps), and the stencil-functor protocol of
// applyStencils (HeatStencil and Conv3x3x3, tests/utils/SparseGridGpu_util_test.cuh:202-470)
// checked bitwise against the library's named stencils.
// Build: hipcc -std=c++17 --offload-arch=gfx950 -ffp-contract=off -I include
//        tests/cpp/test_dropin.hip -L openfpm_data_amd -lsg_hip
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "SparseGridGpu.hpp"

#define REQUIRE(c)                                                                      \
    do {                                                                                \
        if (!(c)) {                                                                     \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                  \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

// ---- functors with the reference's applyStencils protocol ------------------------------
namespace ref {

// value = maxValue * (x - minX) / (maxX - minX - 1) on boundary points, clamped outside
// [minX, maxX] (tests/utils/SparseGridGpu_testKernels.cuh:38-75)
template <unsigned int dim, unsigned int p_src, unsigned int p_dst, typename ScalarT = float>
struct BoundaryStencilSetXRescaled {
    typedef NNStar<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    template <typename SparseGridT, typename DataBlockWrapperT>
    static inline __device__ void stencil(SparseGridT &, const unsigned int, openfpm::sparse_index<unsigned int>,
                                          unsigned int offset, grid_key_dx<dim, int> &pointCoord,
                                          DataBlockWrapperT &, DataBlockWrapperT &dataBlockStore, unsigned char curMask,
                                          ScalarT minX, ScalarT maxX, ScalarT minValue, ScalarT maxValue)
    {
        if (curMask != mask_sparse::EXIST_AND_PADDING) return;
        const ScalarT x = pointCoord.get(0);
        ScalarT v = maxValue * (x - minX) / (maxX - minX - 1);
        if (x < minX) v = minValue;
        else if (x > maxX) v = maxValue;
        dataBlockStore.template get<p_dst>()[offset] = v;
    }
};

// u_dst = u + dt * (-2 dim u + sum over d of (u[-e_d] + u[+e_d])) on interior points, through
// the ghosted block in LDS (loadGhostBlock / storeBlock)
template <unsigned int dim, unsigned int p_src, unsigned int p_dst> struct HeatStencil {
    typedef NNStar<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    template <typename SparseGridT, typename DataBlockWrapperT>
    static inline __device__ void stencil(SparseGridT &sparseGrid, const unsigned int,
                                          const openfpm::sparse_index<unsigned int> dataBlockIdPos,
                                          const unsigned int offset, const grid_key_dx<dim, int> &,
                                          const DataBlockWrapperT &dataBlockLoad, DataBlockWrapperT &dataBlockStore,
                                          unsigned char curMask, float dt)
    {
        typedef typename SparseGridT::AggregateBlockType AggregateT;
        typedef ScalarTypeOf<AggregateT, p_src> ScalarT;
        constexpr unsigned int enlarged = IntPow<SparseGridT::getBlockEdgeSize() + 2 * supportRadius, dim>::value;
        __shared__ ScalarT eb[enlarged];
        sparseGrid.template loadGhostBlock<p_src>(dataBlockLoad, dataBlockIdPos, eb);
        __syncthreads();
        const bool active = (curMask & mask_sparse::EXIST) && !(curMask & mask_sparse::PADDING);
        unsigned int lin = 0;
        ScalarT res = 0;
        if (active) {
            const auto c = sparseGrid.getCoordInEnlargedBlock(offset);
            lin = sparseGrid.getLinIdInEnlargedBlock(offset);
            const ScalarT cur = eb[lin];
            ScalarT lap = -2.0 * dim * cur;
            for (unsigned int d = 0; d < dim; ++d)
                lap += eb[sparseGrid.getNeighbourLinIdInEnlargedBlock(c, d, -1)] +
                       eb[sparseGrid.getNeighbourLinIdInEnlargedBlock(c, d, 1)];
            res = cur + dt * lap;
        }
        __syncthreads();
        if (active) eb[lin] = res;
        __syncthreads();
        sparseGrid.template storeBlock<p_dst>(dataBlockStore, eb);
    }
};

// tot = sum over (i, j, k) of coeff[i][j][k] * u[x + i - 1, y + j - 1, z + k - 1] on every
// existing point, through the ghosted block (all 26 neighbours: NNFull)
template <unsigned int dim, unsigned int p_src, unsigned int p_dst> struct Conv3x3x3 {
    typedef NNFull<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    template <typename SparseGridT, typename DataBlockWrapperT>
    static inline __device__ void stencil(SparseGridT &sparseGrid, const unsigned int,
                                          openfpm::sparse_index<unsigned int> dataBlockIdPos, unsigned int offset,
                                          grid_key_dx<dim, int> &, DataBlockWrapperT &dataBlockLoad,
                                          DataBlockWrapperT &dataBlockStore, bool applyStencilHere, conv_coeff &cc)
    {
        typedef typename SparseGridT::AggregateBlockType AggregateT;
        typedef ScalarTypeOf<AggregateT, p_src> ScalarT;
        constexpr unsigned int enlarged = IntPow<SparseGridT::getBlockEdgeSize() + 2 * supportRadius, dim>::value;
        __shared__ ScalarT eb[enlarged];
        sparseGrid.template loadGhostBlock<p_src>(dataBlockLoad, dataBlockIdPos, eb);
        __syncthreads();
        if (!applyStencilHere) return;
        const auto c = sparseGrid.getCoordInEnlargedBlock(offset);
        ScalarT tot = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    grid_key_dx<dim, int> key;
                    key.set_d(0, i - 1);
                    key.set_d(1, j - 1);
                    key.set_d(2, k - 1);
                    tot += eb[sparseGrid.getNeighbourLinIdInEnlargedBlock(c, key)] * cc.coeff[i][j][k];
                }
        dataBlockStore.template get<p_dst>()[offset] = tot;
    }
};

// the same sum without LDS: each tap through getNNPoint + get<p>(block_offset), taps in the
// order z, y, x (the reference's Conv3x3x3_noshared, coefficient [i][j][k] at (k-1, j-1, i-1))
template <unsigned int dim, unsigned int p_src, unsigned int p_dst> struct Conv3x3x3_noshared {
    typedef NNFull<dim> stencil_type;
    static constexpr unsigned int supportRadius = 1;
    template <typename SparseGridT, typename DataBlockWrapperT>
    static inline __device__ void stencil(SparseGridT &sparseGrid, const unsigned int,
                                          openfpm::sparse_index<unsigned int> dataBlockIdPos, unsigned int offset,
                                          grid_key_dx<dim, int> &, DataBlockWrapperT &, DataBlockWrapperT &dataBlockStore,
                                          bool applyStencilHere, conv_coeff &cc)
    {
        typedef typename SparseGridT::AggregateBlockType AggregateT;
        typedef ScalarTypeOf<AggregateT, p_src> ScalarT;
        if (!applyStencilHere) return;
        ScalarT tot = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    grid_key_dx<dim, int> key;
                    key.set_d(0, k - 1);
                    key.set_d(1, j - 1);
                    key.set_d(2, i - 1);
                    auto pos = sparseGrid.template getNNPoint<stencil_type>(dataBlockIdPos, offset, key);
                    tot += sparseGrid.template get<p_src>(pos) * cc.coeff[i][j][k];
                }
        dataBlockStore.template get<p_dst>()[offset] = tot;
    }
};

}  // namespace ref

// ---- device kernels of the restated tests ------------------------------------------------

// every thread inserts its point (x, y[, z]) with `value` in property p (insertConstantValue,
// tests/utils/SparseGridGpu_util_test.cuh:115-146)
template <unsigned int p, typename SparseGridType, typename ScalarT>
__global__ void insertConstantValue(SparseGridType sparseGrid, ScalarT value)
{
    constexpr unsigned int pMask = SparseGridType::pMask;
    typedef BlockTypeOf<typename SparseGridType::AggregateType, p> BlockT;
    sparseGrid.init();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    grid_key_dx<SparseGridType::d, size_t> coord({(size_t)x, (size_t)y, (size_t)z});
    const auto pos = sparseGrid.getLinId(coord);
    const unsigned int dataBlockId = pos / BlockT::size;
    const unsigned int offset = pos % BlockT::size;
    auto encap = sparseGrid.template insertBlock<1>(dataBlockId, BlockT::size);
    encap.template get<p>()[offset] = value;
    BlockMapGpu_ker<>::setExist(encap.template get<pMask>()[offset]);
    __syncthreads();
    sparseGrid.flush_block_insert();
}

// get<0>, get<1>, exist and isPadding of one point on the device (sparse_grid_get_test,
// SparseGridGpu_tests.cu:795-800, widened)
template <typename SparseGridType> __global__ void readPoint(SparseGridType sparseGrid, grid_key_dx<3> key, float *out)
{
    out[0] = sparseGrid.template get<0>(key);
    out[1] = sparseGrid.template get<1>(key);
    out[2] = sparseGrid.exist(key) ? 1.0f : 0.0f;
    out[3] = sparseGrid.isPadding(key) ? 1.0f : 0.0f;
    unsigned int pos = 0, off = 0;
    sparseGrid.get_sparse(key, pos, off);
    out[4] = (float)off;
}

static void device_read(float *h, const SparseGridGpu<3, aggregate<float, float>, 4, 64>::kernel_view &k, int x, int y,
                        int z)
{
    float *d = nullptr;
    REQUIRE(hipMalloc((void **)&d, 5 * sizeof(float)) == hipSuccess);
    grid_key_dx<3> key({(long)x, (long)y, (long)z});
    hipLaunchKernelGGL((readPoint<SparseGridGpu<3, aggregate<float, float>, 4, 64>::kernel_view>), dim3(1), dim3(1), 0,
                       (hipStream_t)k.gv.stream, k, key, d);
    REQUIRE(hipMemcpy(h, d, 5 * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess);
    REQUIRE(hipFree(d) == hipSuccess);
}

// testFlushInsert (SparseGridGpu_tests.cu:802-880)
static void testFlushInsert()
{
    constexpr unsigned int dim = 3;
    constexpr unsigned int blockEdgeSize = 4;
    typedef aggregate<float, float> AggregateT;
    size_t sz[] = {137, 100, 57};
    SparseGridGpu<dim, AggregateT, blockEdgeSize, 64> sparseGrid(sz);
    sparseGrid.template setBackgroundValue<0>(0);
    const int pts[9][3] = {{3, 6, 7},    {13, 16, 17},  {13, 46, 27}, {36, 63, 11}, {37, 96, 47},
                           {130, 56, 37}, {131, 76, 17}, {36, 86, 27}, {34, 36, 7}};
    // the points, then a neighbour of each in the same chunk
    for (int i = 0; i < 9; ++i)
        sparseGrid.insertFlush<0>(grid_key_dx<3>({(long)pts[i][0], (long)pts[i][1], (long)pts[i][2]})) = 2.0f + i;
    const int dx[9] = {1, -1, -1, -1, 1, 1, 1, 1, 1};
    for (int i = 0; i < 9; ++i)
        sparseGrid.insertFlush<0>(grid_key_dx<3>({(long)pts[i][0] + dx[i], (long)pts[i][1], (long)pts[i][2]})) = 2.0f + i;
    for (int i = 0; i < 9; ++i) {
        REQUIRE(sparseGrid.get<0>(grid_key_dx<3>({(long)pts[i][0], (long)pts[i][1], (long)pts[i][2]})) == 2.0f + i);
        REQUIRE(sparseGrid.get<0>(grid_key_dx<3>({(long)pts[i][0] + dx[i], (long)pts[i][1], (long)pts[i][2]})) ==
                2.0f + i);
    }
    sparseGrid.template hostToDevice<0>();
    float h[5];
    device_read(h, sparseGrid.toKernel(), 3, 6, 7);
    REQUIRE(h[0] == 2.0f && h[2] == 1.0f && h[3] == 0.0f && h[4] == 3.0f + 4 * 2 + 16 * 3);
    device_read(h, sparseGrid.toKernel(), 131, 76, 17);
    REQUIRE(h[0] == 8.0f);
    // a point that was not inserted, in an inserted chunk and in a missing one: background
    device_read(h, sparseGrid.toKernel(), 2, 6, 7);
    REQUIRE(h[0] == 0.0f && h[2] == 0.0f);
    device_read(h, sparseGrid.toKernel(), 100, 1, 1);
    REQUIRE(h[0] == 0.0f && h[2] == 0.0f);
    REQUIRE(sparseGrid.countExistingElements() == 18);
    std::printf("testFlushInsert ok\n");
}

// insertFlush<p> writes property p and the mask only: the other properties of an existing
// point keep their values, and the mask becomes EXIST (the reference writes 1,
// SparseGridGpu.hpp:2842), also where it was EXIST | PADDING
static void testFlushInsertKeepsOtherProperties()
{
    typedef SparseGridGpu<3, aggregate<float, float>, 4, 64> SG;
    size_t sz[] = {32, 32, 32};
    SG sg(sz);
    gpu::ofp_context_t ctx;
    sg.template setBackgroundValue<1>(-1.0f);
    std::vector<long> keys;
    std::vector<float> a, b;
    for (int z = 8; z < 16; ++z)
        for (int y = 8; y < 16; ++y)
            for (int x = 8; x < 16; ++x) {
                keys.push_back(sg.getLinId(grid_key_dx<3, int>({x, y, z})));
                a.push_back(1.0f);
                b.push_back(7.0f + x);
            }
    const void *v[] = {a.data(), b.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.template flush<smax_<0>, smax_<1>>(ctx, FLUSH_ON_DEVICE);
    sg.template findNeighbours<NNStar<3>>();
    sg.template tagBoundaries<NNStar<3>>(ctx);
    sg.deviceToHost();
    // (8, 9, 9) is a boundary point: mask EXIST | PADDING before
    sg.insertFlush<0>(grid_key_dx<3>({8, 9, 9})) = 5.0f;
    // a new point in a new chunk: property 1 takes its background
    sg.insertFlush<0>(grid_key_dx<3>({30, 30, 30})) = 6.0f;
    REQUIRE(sg.get<1>(grid_key_dx<3>({8, 9, 9})) == 15.0f);
    REQUIRE(sg.get<1>(grid_key_dx<3>({30, 30, 30})) == -1.0f);
    sg.hostToDevice();
    float h[5];
    {
        typedef SparseGridGpu<3, aggregate<float, float>, 4, 64> SGT;
        SGT::kernel_view k = sg.toKernel();
        device_read(h, k, 8, 9, 9);
        REQUIRE(h[0] == 5.0f && h[1] == 15.0f && h[2] == 1.0f && h[3] == 0.0f);
        device_read(h, k, 30, 30, 30);
        REQUIRE(h[0] == 6.0f && h[1] == -1.0f && h[2] == 1.0f);
        device_read(h, k, 9, 9, 8);  // still a boundary point
        REQUIRE(h[0] == 1.0f && h[1] == 16.0f && h[3] == 1.0f);
    }
    std::vector<uint8_t> m = sg.private_get_mask_array();
    size_t n1 = 0;
    for (uint8_t x : m) n1 += (x & 1);
    REQUIRE(n1 == 513);
    std::printf("testFlushInsertKeepsOtherProperties ok\n");
}

// testStencil_lap_simplified (SparseGridGpu_tests.cu:521-576)
static void testStencil_lap_simplified()
{
    constexpr unsigned int dim = 2;
    constexpr unsigned int blockEdgeSize = 8;
    typedef aggregate<float, float> AggregateT;
    dim3 gridSize(2, 2);
    dim3 blockSizeInsert(blockEdgeSize, blockEdgeSize);
    grid_smb<dim, blockEdgeSize> blockGeometry(gridSize);
    SparseGridGpu<dim, AggregateT, blockEdgeSize, 64> sparseGrid(blockGeometry);
    gpu::ofp_context_t gpuContext;
    sparseGrid.template setBackgroundValue<0>(0);
    sparseGrid.setGPUInsertBuffer(gridSize, blockSizeInsert);
    hipLaunchKernelGGL((insertConstantValue<0>), gridSize, blockSizeInsert, 0, 0, sparseGrid.toKernel(), 0.0f);
    sparseGrid.template flush<smax_<0>>(gpuContext, flush_type::FLUSH_ON_DEVICE);
    sparseGrid.findNeighbours();
    sparseGrid.tagBoundaries(gpuContext);
    sparseGrid.template applyStencils<ref::BoundaryStencilSetXRescaled<dim, 0, 0>>(
        sparseGrid.getBox(), STENCIL_MODE_INPLACE, 0.0, gridSize.x * blockEdgeSize, 0.0, 10.0);
    for (unsigned int iter = 0; iter < 1000; ++iter) {
        sparseGrid.conv_cross<0, 1, 1>({0, 0}, {16, 16}, [] __device__(float &u, cross_stencil<2, float> &cs) {
            return u + (cs.xm[0] + cs.xp[0] + cs.xm[1] + cs.xp[1] - 4.0 * u) * 0.1;
        });
        sparseGrid.conv_cross<1, 0, 1>({0, 0}, {16, 16}, [] __device__(float &u, cross_stencil<2, float> &cs) {
            return u + (cs.xm[0] + cs.xp[0] + cs.xm[1] + cs.xp[1] - 4.0 * u) * 0.1;
        });
    }
    sparseGrid.deviceToHost<0, 1>();
    sparseGrid.template deviceToHost<0>();
    bool match = true;
    for (size_t i = 0; i < 64 * 4; i++) {
        auto coord = sparseGrid.getCoord(i);
        float expectedValue = 10.0 * coord.get(0) / (gridSize.x * blockEdgeSize - 1);
        match &= std::fabs(sparseGrid.template get<0>(coord) - expectedValue) < 1e-2;
    }
    REQUIRE(match);
    REQUIRE(sparseGrid.countExistingElements() == 256);
    std::printf("testStencil_lap_simplified ok\n");
}

// a 48^3 shell with non-integer values in properties 0 and 2, tagged with NNStar
template <typename SG> static void build_shell(SG &sg, bool full)
{
    gpu::ofp_context_t ctx;
    std::vector<long> keys;
    std::vector<float> a, z0, c;
    for (int z = 0; z < 48; ++z)
        for (int y = 0; y < 48; ++y)
            for (int x = 0; x < 48; ++x) {
                const double r = std::sqrt((x - 24.0) * (x - 24.0) + (y - 24.0) * (y - 24.0) + (z - 24.0) * (z - 24.0));
                if (r >= 9.0 && r < 17.0) {
                    keys.push_back(sg.getLinId(grid_key_dx<3, int>({x, y, z})));
                    a.push_back((float)std::sin(0.37 * x + 0.11 * y * y - 0.05 * z));
                    c.push_back((float)((x * 7 + y * 3 + z) % 17));
                    z0.push_back(0.0f);
                }
            }
    const void *v[] = {a.data(), z0.data(), c.data(), z0.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.template flush<smax_<0>, smax_<1>, smax_<2>, smax_<3>>(ctx, FLUSH_ON_DEVICE);
    if (full) {
        sg.template findNeighbours<NNFull<3>>();
        sg.template tagBoundaries<NNFull<3>>(ctx);
    } else {
        sg.template findNeighbours<NNStar<3>>();
        sg.template tagBoundaries<NNStar<3>>(ctx);
    }
}

typedef SparseGridGpu<3, aggregate<float, float, float, float>> SG4;

// HeatStencil as a user functor against the named HeatStencil: bitwise on interior points;
// on the others the functor's storeBlock writes the source value
template <typename SG> static void testHeatFunctor()
{
    size_t sz[] = {48, 48, 48};
    SG a(sz), b(sz);
    build_shell(a, false);
    build_shell(b, false);
    Box<3, int> box({0, 0, 0}, {47, 47, 47});
    a.template applyStencils<HeatStencil<3, 0, 1>>(box, STENCIL_MODE_INPLACE, 0.1);
    b.template applyStencils<ref::HeatStencil<3, 0, 1>>(box, STENCIL_MODE_INPLACE, 0.1f);
    const std::vector<float> ra = a.template private_get_data_array<1>(), rb = b.template private_get_data_array<1>();
    const std::vector<float> src = b.template private_get_data_array<0>();
    const std::vector<uint8_t> m = a.private_get_mask_array();
    REQUIRE(m == b.private_get_mask_array());
    size_t active = 0;
    for (size_t i = 0; i < m.size(); ++i) {
        if ((m[i] & 3) == 1) {
            REQUIRE(std::memcmp(&ra[i], &rb[i], 4) == 0);
            ++active;
        } else if (m[i] & 1) {
            REQUIRE(rb[i] == src[i]);
        } else {
            REQUIRE(rb[i] == 0.0f);  // slots without a point hold the background
        }
    }
    REQUIRE(active > 10000);
    // a box: points outside it keep the source value
    SG c(sz);
    build_shell(c, false);
    Box<3, int> half({0, 0, 0}, {23, 47, 47});
    c.template applyStencils<ref::HeatStencil<3, 0, 1>>(half, STENCIL_MODE_INPLACE, 0.1f);
    const std::vector<float> rc = c.template private_get_data_array<1>();
    const std::vector<long> ids = c.private_get_index_array();
    for (size_t i = 0; i < m.size(); ++i) {
        if ((m[i] & 3) != 1) continue;
        const int x = c.getCoord(ids[i / 512] * 512 + (long)(i % 512)).get(0);
        if (x <= 23) REQUIRE(std::memcmp(&rc[i], &ra[i], 4) == 0);
        else REQUIRE(rc[i] == src[i]);
    }
    std::printf("testHeatFunctor ok\n");
}

// Conv3x3x3 as a user functor (LDS, NNFull ghosts) and Conv3x3x3_noshared (point list,
// getNNPoint) against the named Conv3x3x3, on every existing point
template <typename SG> static void testConvFunctors()
{
    size_t sz[] = {48, 48, 48};
    conv_coeff cc;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) cc.coeff[i][j][k] = 0.25f + 0.125f * (i + 2 * j + 3 * k);
    SG a(sz), b(sz), c(sz);
    build_shell(a, true);
    build_shell(b, true);
    build_shell(c, true);
    Box<3, int> box({0, 0, 0}, {47, 47, 47});
    a.template applyStencils<Conv3x3x3<3, 0, 1>>(box, STENCIL_MODE_INPLACE, cc);
    b.template applyStencils<ref::Conv3x3x3<3, 0, 1>>(box, STENCIL_MODE_INPLACE, cc);
    const std::vector<float> ra = a.template private_get_data_array<1>(), rb = b.template private_get_data_array<1>();
    const std::vector<uint8_t> m = a.private_get_mask_array();
    size_t n = 0;
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] & 1) {
            REQUIRE(std::memcmp(&ra[i], &rb[i], 4) == 0);
            ++n;
        }
    REQUIRE(n > 10000);
    // the point-list form on integer-valued property 2 with symmetric integer coefficients:
    // every sum is exact, so its other tap order gives the same values
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) cc.coeff[i][j][k] = (float)(1 + (i == 1) + (j == 1) + (k == 1));
    a.template applyStencils<Conv3x3x3<3, 2, 3>>(box, STENCIL_MODE_INPLACE, cc);
    c.template applyStencils<ref::Conv3x3x3_noshared<3, 2, 3>>(box, STENCIL_MODE_INPLACE_NO_SHARED, cc);
    const std::vector<float> rd = a.template private_get_data_array<3>(), re = c.template private_get_data_array<3>();
    n = 0;
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] & 1) {
            REQUIRE(rd[i] == re[i]);
            n += rd[i] != 0.0f;
        }
    REQUIRE(n > 10000);
    std::printf("testConvFunctors ok\n");
}

int main()
{
    testFlushInsert();
    testFlushInsertKeepsOtherProperties();
    testStencil_lap_simplified();
    testHeatFunctor<SG4>();
    testHeatFunctor<SparseGridGpu_z<3, aggregate<float, float, float, float>>>();
    testConvFunctors<SG4>();
    testConvFunctors<SparseGridGpu_z<3, aggregate<float, float, float, float>>>();
    std::printf("dropin tests passed\n");
    return 0;
}
