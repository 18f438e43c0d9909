// The reference's own SparseGridGpu tests, restated through the C++ facade
// (include/SparseGridGpu.hpp) and checked against the same known answers.
//   testInsert                  src/SparseGridGpu/tests/SparseGridGpu_tests.cu:209-244
//   testTagBoundaries           src/SparseGridGpu/tests/SparseGridGpu_tests.cu:284-370
//   testStencilHeat             src/SparseGridGpu/tests/SparseGridGpu_tests.cu:472-514
//   test_convolution_3x3x3      src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1012-1080
#include <cmath>
#include <cstdio>
#include <set>
#include <vector>

#include "SparseGridGpu.hpp"

static int failures = 0;
#define REQUIRE(c)                                                     \
    do {                                                               \
        if (!(c)) {                                                    \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                                \
        }                                                              \
    } while (0)

static void testInsert()
{
    const size_t sz[2] = {16, 16};
    SparseGridGpu<2, aggregate<float>, 8> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(666);
    std::vector<long> keys;
    std::vector<float> vals;
    for (int y = 0; y < 16; ++y)
        for (int x = 0; x < 16; ++x) {
            keys.push_back(sg.getLinId(grid_key_dx<2, int>({x, y})));
            vals.push_back((float)x);
        }
    const void *v[] = {vals.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
    bool match = true;
    for (long i = 0; i < 4 * 64; ++i) {
        auto coord = sg.getCoord(i);
        match &= sg.get<0>(coord) == coord.get(0);
    }
    REQUIRE(match);
}

static void testTagBoundaries()
{
    const size_t sz[2] = {16, 16};
    SparseGridGpu<2, aggregate<float>, 8, 64> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(666);
    auto put = [&](std::vector<std::pair<int, int>> pts) {
        std::vector<long> keys;
        std::vector<float> vals;
        for (auto &p : pts) {
            keys.push_back(sg.getLinId(grid_key_dx<2, int>({p.first, p.second})));
            vals.push_back(1.0f);
        }
        const void *v[] = {vals.data()};
        sg.insertPoints(keys.data(), keys.size(), v);
        sg.flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
    };
    put({{0, 0}, {6, 6}, {7, 6}});
    put({{8, 6}});
    put({{6, 9}, {7, 9}, {6, 10}, {7, 10}, {6, 11}, {7, 11}});
    put({{8, 9}, {9, 9}, {8, 10}, {9, 10}, {8, 11}, {9, 11}});
    sg.findNeighbours();
    sg.tagBoundaries(ctx);
    std::vector<long> ids = sg.private_get_index_array();
    std::vector<uint8_t> m = sg.private_get_mask_array();
    std::set<long> got;
    for (size_t c = 0; c < ids.size(); ++c)
        for (int o = 0; o < 64; ++o)
            if (m[c * 64 + o] & PADDING) got.insert(ids[c] * 64 + o);
    std::set<long> expect = {0, 54, 55, 112, 142, 143, 200, 201, 150, 209, 158, 159, 216, 217};
    REQUIRE(got == expect);
}

static void testStencilHeat()
{
    const size_t sz[2] = {16, 16};
    SparseGridGpu<2, aggregate<float, float>, 8, 64> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(0);
    std::vector<long> keys;
    std::vector<float> zero(256, 0.0f);
    for (int y = 0; y < 16; ++y)
        for (int x = 0; x < 16; ++x) keys.push_back(sg.getLinId(grid_key_dx<2, int>({x, y})));
    const void *v[] = {zero.data(), zero.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.flush<smax_<0>, smax_<1>>(ctx, FLUSH_ON_DEVICE);
    sg.findNeighbours();
    sg.tagBoundaries(ctx);
    // boundary (padding) points fixed to 10 * x / 15 (BoundaryStencilSetXRescaled)
    std::vector<long> ids = sg.private_get_index_array();
    std::vector<uint8_t> m = sg.private_get_mask_array();
    std::vector<float> u = sg.private_get_data_array<0>();
    for (size_t i = 0; i < u.size(); ++i) {
        auto c = sg.getCoord(ids[i / 64] * 64 + (long)(i % 64));
        u[i] = (m[i] & PADDING) ? 10.0f * c.get(0) / 15.0f : 0.0f;
    }
    sg.hostToDevice<0>(u);
    sg.hostToDevice<1>(u);
    for (int it = 0; it < 1000; ++it) {
        sg.applyStencils<HeatStencil<2, 0, 1>>(sg.getBox(), STENCIL_MODE_INPLACE, 0.1);
        sg.applyStencils<HeatStencil<2, 1, 0>>(sg.getBox(), STENCIL_MODE_INPLACE, 0.01);
    }
    bool match = true;
    for (long i = 0; i < 64 * 4; ++i) {
        auto coord = sg.getCoord(i);
        float expected = 10.0f * coord.get(0) / 15.0f;
        match &= std::fabs(sg.get<1>(coord) - expected) < 1e-2;
    }
    REQUIRE(match);
}

static void testConv3x3x3()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu<3, aggregate<float, float>> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(0);
    std::vector<long> keys;
    std::vector<float> one, zero;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                double r = std::sqrt((x - 32.0) * (x - 32.0) + (y - 32.0) * (y - 32.0) + (z - 32.0) * (z - 32.0));
                if (r >= 10.0 && r < 20.0) {
                    keys.push_back(sg.getLinId(grid_key_dx<3, int>({x, y, z})));
                    one.push_back(1.0f);
                    zero.push_back(0.0f);
                }
            }
    const void *v[] = {one.data(), zero.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.flush<smax_<0>, smax_<1>>(ctx, FLUSH_ON_DEVICE);
    sg.findNeighbours<NNFull<3>>();
    sg.setNNType<NNFull<3>>();
    sg.tagBoundaries<NNFull<3>>(ctx);
    conv_coeff cc;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) cc.coeff[k][j][i] = 1.0;
    sg.applyStencils<Conv3x3x3<3, 0, 1>>(sg.getBox(), STENCIL_MODE_INPLACE, cc);
    std::vector<uint8_t> m = sg.private_get_mask_array();
    std::vector<float> out = sg.private_get_data_array<1>();
    bool match = true;
    size_t n_interior = 0;
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] == 1) {
            match &= out[i] == 27.0f;
            ++n_interior;
        }
    REQUIRE(match);
    REQUIRE(n_interior > 10000);
    double mean, dev;
    sg.measureBlockOccupancy(mean, dev);
    REQUIRE(mean > 0.0 && mean <= 1.0);
    REQUIRE(sg.countExistingElements() == keys.size());
}

int main()
{
    testInsert();
    testTagBoundaries();
    testStencilHeat();
    testConv3x3x3();
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("facade tests passed\n");
    return 0;
}
