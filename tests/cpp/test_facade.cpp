// The reference's own SparseGridGpu tests, restated through the C++ facade
// (include/SparseGridGpu.hpp) and checked against the same known answers.
//   testInsert                  src/SparseGridGpu/tests/SparseGridGpu_tests.cu:209-244
//   testTagBoundaries           src/SparseGridGpu/tests/SparseGridGpu_tests.cu:284-370
//   testStencilHeat             src/SparseGridGpu/tests/SparseGridGpu_tests.cu:472-514
//   test_convolution_3x3x3      src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1012-1080
//   test_pack_request_with_iterator  src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1379-1472
//   sparsegridgpu_remove_test   src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1474-1551
//   sparsegridgpu_pack_unpack   src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1553-1934
//                               (aggregate<float, float[3]> as four float properties)
//   test_sparse_grid_iterator_sub_host  src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1179-1229
//   test_sparse_grid_iterator_host      src/SparseGridGpu/tests/SparseGridGpu_tests.cu:1231-1282
//   testCopyTo: copy_to + removeCopyToFinalize (SparseGridGpu.hpp:3314-3329, :3550-3559) of
//   the ball of the sub-iterator test into an empty grid
//   testLinks: construct_link_dw / construct_link_up (SparseGridGpu.hpp:2090-2262) on a coarse
//   ball and a fine ball: every link is a child / the parent of its padding point
//   testHostPackUnpack: whole-grid pack into host memory (SparseGridGpu.hpp:2895-2959)
#include <cmath>
#include <cstdio>
#include <cstring>
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

// flush(ctx, FLUSH_ON_HOST): the same merge as on the device, and the host copy holds it
// (an insertFlush edit made before the flush is kept)
static void testFlushOnHost()
{
    const size_t sz[3] = {32, 32, 32};
    SparseGridGpu<3, aggregate<float, int>, 8> a(sz), b(sz);
    gpu::ofp_context_t ctx;
    std::vector<long> keys;
    std::vector<float> v0;
    std::vector<int> v1;
    for (int i = 0; i < 500; ++i) {
        keys.push_back((long)((i * 7919) % (32 * 32 * 32)));
        v0.push_back(0.5f * (float)i);
        v1.push_back(i % 13);
    }
    const void *v[] = {v0.data(), v1.data()};
    a.insertPoints(keys.data(), keys.size(), v);
    a.flush<smax_<0>, sadd_<1>>(ctx, FLUSH_ON_DEVICE);
    b.insertPoints(keys.data(), keys.size(), v);
    b.flush<smax_<0>, sadd_<1>>(ctx, FLUSH_ON_HOST);
    REQUIRE(a.private_get_index_array() == b.private_get_index_array());
    REQUIRE(a.private_get_mask_array() == b.private_get_mask_array());
    REQUIRE(a.private_get_data_array<0>() == b.private_get_data_array<0>());
    REQUIRE(a.private_get_data_array<1>() == b.private_get_data_array<1>());
    // a host edit, then more inserts flushed on the host
    b.insertFlush<1>(grid_key_dx<3>({1, 2, 3})) = 77;
    b.insertPoints(keys.data(), 100, v);
    b.flush<smax_<0>, sadd_<1>>(ctx, FLUSH_ON_HOST);
    REQUIRE(b.get<1>(grid_key_dx<3>({1, 2, 3})) == 77);
    a.insertPoints(keys.data(), 100, v);
    a.flush<smax_<0>, sadd_<1>>(ctx, FLUSH_ON_DEVICE);
    bool same = true;
    for (int i = 0; i < 500; ++i) {
        const auto c = a.getCoord(keys[i]);
        same &= a.get<0>(c) == b.get<0>(c) && a.get<1>(c) == b.get<1>(c);
    }
    REQUIRE(same);
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

// run on SparseGridGpu and on SparseGridGpu_z, as the reference does (:1170-1177)
template <typename SG> static void testConv3x3x3()
{
    const size_t sz[3] = {64, 64, 64};
    SG sg(sz);
    gpu::ofp_context_t ctx;
    sg.template setBackgroundValue<0>(0);
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
    sg.template flush<smax_<0>, smax_<1>>(ctx, FLUSH_ON_DEVICE);
    sg.template findNeighbours<NNFull<3>>();
    sg.template setNNType<NNFull<3>>();
    sg.template tagBoundaries<NNFull<3>>(ctx);
    conv_coeff cc;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) cc.coeff[k][j][i] = 1.0;
    sg.template applyStencils<Conv3x3x3<3, 0, 1>>(sg.getBox(), STENCIL_MODE_INPLACE, cc);
    std::vector<uint8_t> m = sg.private_get_mask_array();
    std::vector<float> out = sg.template private_get_data_array<1>();
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

// The chunks insertSphere3D_radius(V) writes (tests/utils/SparseGridGpu_testKernels.cuh:202-318):
// blocks 64..95 per axis of a 1000^3 grid with 4^3 chunks, active when 64 > r > 56 about 320.
template <typename SG> static void insertSphere(SG &sg, bool vector)
{
    std::vector<long> ids;
    std::vector<uint8_t> masks;
    std::vector<float> v[4];
    const int e = 4, bs = 250;
    for (int bz = 64; bz < 96; ++bz)
        for (int by = 64; by < 96; ++by)
            for (int bx = 64; bx < 96; ++bx) {
                uint8_t m[64];
                float val[4][64];
                bool any = false;
                for (int o = 0; o < 64; ++o) {
                    const long x = bx * e + o % e - 320, y = by * e + (o / e) % e - 320, z = bz * e + o / 16 - 320;
                    const float r = std::sqrt((float)(x * x + y * y + z * z));
                    const bool act = r < 64.0f && r > 56.0f;
                    m[o] = act ? 1 : 0;
                    val[0][o] = act ? (float)(x + y + z) : 0.0f;
                    val[1][o] = act ? (float)x : 0.0f;
                    val[2][o] = act ? (float)y : 0.0f;
                    val[3][o] = act ? (float)z : 0.0f;
                    any = any || act;
                }
                if (!any) continue;
                ids.push_back(((long)bz * bs + by) * bs + bx);
                masks.insert(masks.end(), m, m + 64);
                for (int q = 0; q < 4; ++q) v[q].insert(v[q].end(), val[q], val[q] + 64);
            }
    const void *vals[4] = {v[0].data(), v[1].data(), v[2].data(), v[3].data()};
    sg.insertBlocks(ids.data(), ids.size(), masks.data(), vals);
    gpu::ofp_context_t ctx;
    if (vector) sg.template flush<smax_<0>, smax_<1>, smax_<2>, smax_<3>>(ctx, FLUSH_ON_DEVICE);
    else sg.template flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
}

static size_t align8(size_t v) { return (v + 7) / 8 * 8; }

static void testPackRequestWithIterator()
{
    const size_t sz[3] = {1000, 1000, 1000};
    SparseGridGpu<3, aggregate<float>, 4, 64, long int> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(0);
    insertSphere(sg, false);
    size_t req = 0;
    sg.packReset();
    auto it1 = sg.getIterator(grid_key_dx<3>({0, 0, 0}), grid_key_dx<3>({321, 999, 999}));
    sg.packRequest<0>(it1, req);
    auto it2 = sg.getIterator(grid_key_dx<3>({322, 0, 0}), grid_key_dx<3>({999, 999, 999}));
    sg.packRequest<0>(it2, req);
    sg.packCalculate<0>(req, ctx);
    const size_t tot = 2 * 8 + 2 * 2 * 3 * 4 + align8(4685 * 8) + align8(4475 * 8) + align8(4686 * 4) +
                       align8(4476 * 4) + align8(185807 * 4) + align8(176787 * 4) + align8(185807 * 2) +
                       align8(176787 * 2) + align8(185807 * 1) + align8(176787 * 1);
    REQUIRE(req == tot);
}

template <typename SG> static std::vector<grid_key_dx<3, int>> existingPoints(const SG &sg)
{
    std::vector<grid_key_dx<3, int>> out;
    const std::vector<long> ids = sg.private_get_index_array();
    const std::vector<uint8_t> m = sg.private_get_mask_array();
    for (size_t c = 0; c < ids.size(); ++c)
        for (unsigned int o = 0; o < sg.getBlockSize(); ++o)
            if (m[c * sg.getBlockSize() + o] & 1) out.push_back(sg.getCoord(ids[c] * sg.getBlockSize() + o));
    return out;
}

static void testRemove()
{
    const size_t sz[3] = {1000, 1000, 1000};
    SparseGridGpu<3, aggregate<float>, 4, 64, long int> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(0);
    insertSphere(sg, false);
    Box<3, unsigned int> s1({310, 0, 0}, {330, 999, 999}), s2({0, 310, 0}, {999, 330, 999}),
        s3({0, 0, 310}, {999, 999, 330});
    sg.remove(s1);
    sg.remove(s2);
    sg.remove(s3);
    sg.removeAddUnpackFinalize<>(ctx, 0);
    bool match = true;
    for (const auto &p : existingPoints(sg)) {
        const float r = std::sqrt((float)((p.get(0) - 320) * (p.get(0) - 320) + (p.get(1) - 320) * (p.get(1) - 320) +
                                          (p.get(2) - 320) * (p.get(2) - 320)));
        match = match && !(r < 55.99f || r > 64.01f) && !s1.isInside(p) && !s2.isInside(p) && !s3.isInside(p);
    }
    REQUIRE(match);
}

typedef SparseGridGpu<3, aggregate<float, float, float, float>, 4, 64, long int> SGV;

static void packUnpack(SGV &dst, SGV &src, const Box<3, size_t> (&dbox)[4], bool test_pack)
{
    const Box<3, size_t> sbox[4] = {Box<3, size_t>({256, 256, 256}, {273, 390, 390}),
                                    Box<3, size_t>({320, 256, 256}, {337, 390, 390}),
                                    Box<3, size_t>({256, 256, 256}, {273, 320, 390}),
                                    Box<3, size_t>({320, 256, 256}, {337, 320, 390})};
    gpu::ofp_context_t ctx;
    src.packReset();
    size_t req = 0;
    for (const auto &b : sbox) {
        auto it = src.getIterator(grid_key_dx<3, size_t>({b.getLow(0), b.getLow(1), b.getLow(2)}),
                                  grid_key_dx<3, size_t>({b.getHigh(0), b.getHigh(1), b.getHigh(2)}));
        src.packRequest<0, 1, 2, 3>(it, req);
    }
    src.packCalculate<0, 1, 2, 3>(req, ctx);
    PackMemory mem;
    mem.resize(req);
    Pack_stat sts;
    src.packFinalize<0, 1, 2, 3>(mem, sts);
    if (test_pack) {
        const size_t expect[4][2] = {{1107, 41003}, {1420, 54276}, {610, 20828}, {739, 27283}};
        size_t at = 0;
        for (int k = 0; k < 4; ++k) {
            size_t ncnk;
            std::memcpy(&ncnk, mem.getPointer() + at, 8);
            REQUIRE(ncnk == expect[k][0]);
            size_t off = at + ncnk * 8 + 8 + 2 * 3 * 4;
            unsigned int npnt;
            std::memcpy(&npnt, mem.getPointer() + off + ncnk * 4, 4);
            REQUIRE(npnt == expect[k][1]);
            at = off + align8((ncnk + 1) * 4) + align8(npnt * 16) + align8(npnt * 2) + align8(npnt);
        }
    }
    Unpack_stat ps;
    dst.removeAddUnpackReset();
    for (const auto &b : dbox) {
        auto sub = dst.getIterator(grid_key_dx<3, size_t>({b.getLow(0), b.getLow(1), b.getLow(2)}),
                                   grid_key_dx<3, size_t>({b.getHigh(0), b.getHigh(1), b.getHigh(2)}));
        dst.remove(b);
        dst.unpack<0, 1, 2, 3>(mem, sub, ps, ctx);
    }
    dst.removeAddUnpackFinalize<0, 1, 2, 3>(ctx, 0);
}

static void countDst(const SGV &dst, const Box<3, size_t> (&dbox)[4], const size_t (&expect)[4])
{
    const long shift[4] = {0, -46, 44, 0};
    size_t cnt[4] = {0, 0, 0, 0};
    bool match = true;
    for (const auto &p : existingPoints(dst))
        for (int k = 0; k < 4; ++k)
            if (dbox[k].isInside(p)) {
                ++cnt[k];
                const long x = p.get(0) - (320 + shift[k]), y = p.get(1) - 320, z = p.get(2) - 320;
                const float r = std::sqrt((float)(x * x + y * y + z * z));
                match = match && r < 64.0f && r > 56.0f;
                break;
            }
    REQUIRE(match);
    for (int k = 0; k < 4; ++k) REQUIRE(cnt[k] == expect[k]);
}

static void testPackUnpack()
{
    const size_t sz[3] = {1000, 1000, 1000};
    const Box<3, size_t> dbox[4] = {Box<3, size_t>({256, 256, 256}, {273, 390, 390}),
                                    Box<3, size_t>({274, 256, 256}, {291, 390, 390}),
                                    Box<3, size_t>({300, 256, 256}, {317, 320, 390}),
                                    Box<3, size_t>({320, 256, 256}, {337, 320, 390})};
    SGV src(sz), dst(sz);
    src.setBackgroundValue<0>(0);
    dst.setBackgroundValue<0>(0);
    insertSphere(src, true);
    packUnpack(dst, src, dbox, true);
    countDst(dst, dbox, {41003, 54276, 20828, 27283});
    // remove the even points of the source (removeSphere3D_even_radiusV) and repeat
    std::vector<long> even;
    for (const auto &p : existingPoints(src))
        if ((p.get(0) + p.get(1) + p.get(2)) % 2 == 0) even.push_back(src.getLinId(p));
    src.removePoints(even.data(), even.size());
    packUnpack(dst, src, dbox, false);
    countDst(dst, dbox, {20520, 27152, 10423, 13649});
}

// The chunks insertSphere3D writes (tests/utils/SparseGridGpu_testKernels.cuh:143-200):
// 32^3 blocks from start / edge, points active when r1 > r >= r2 about start + 16 edges,
// value 1 on the active points of every block that has one.
template <typename SG> static void insertSphere3D(SG &sg, int start, float r1, float r2)
{
    const int e = SG::blockEdgeSize_, cs = e * e * e;
    const int c0 = start + 16 * e;
    std::vector<long> ids;
    std::vector<uint8_t> masks;
    std::vector<float> v;
    for (int bz = 0; bz < 32; ++bz)
        for (int by = 0; by < 32; ++by)
            for (int bx = 0; bx < 32; ++bx) {
                std::vector<uint8_t> m(cs);
                std::vector<float> val(cs, 0.0f);
                bool any = false;
                for (int o = 0; o < cs; ++o) {
                    const long x = start + bx * e + o % e - c0, y = start + by * e + (o / e) % e - c0,
                               z = start + bz * e + o / (e * e) - c0;
                    const float r = std::sqrt((float)(x * x + y * y + z * z));
                    const bool act = r < r1 && r >= r2;
                    m[o] = act ? 1 : 0;
                    val[o] = act ? 1.0f : 0.0f;
                    any = any || act;
                }
                if (!any) continue;
                grid_key_dx<3, int> b0({start + bx * e, start + by * e, start + bz * e});
                ids.push_back(sg.getLinId(b0) / cs);
                masks.insert(masks.end(), m.begin(), m.end());
                v.insert(v.end(), val.begin(), val.end());
            }
    const void *vals[2] = {v.data(), nullptr};
    sg.insertBlocks(ids.data(), ids.size(), masks.data(), vals);
    gpu::ofp_context_t ctx;
    sg.template flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
}

static void testIteratorSubHost()
{
    const size_t sz[3] = {768, 768, 768};
    SparseGridGpu<3, aggregate<float, float>, 4, 64, long int> sg(sz);
    sg.setBackgroundValue<0>(0);
    insertSphere3D(sg, 256, 32.0f, 0.0f);
    sg.deviceToHost<0>();
    bool match = true;
    int count = 0;
    auto it = sg.getIterator(grid_key_dx<3>({303, 303, 303}), grid_key_dx<3>({337, 337, 337}));
    std::set<long> seen;
    while (it.isNext()) {
        auto key = it.get();
        match = match && sg.get<0>(key) == 1.0f;
        const auto c = key.getKey();
        match = match && c.get(0) >= 303 && c.get(0) <= 337 && c.get(1) >= 303 && c.get(1) <= 337 &&
                c.get(2) >= 303 && c.get(2) <= 337;
        seen.insert(sg.getLinId(c));
        sg.get<0>(key) = 5.0f;
        ++count;
        ++it;
    }
    REQUIRE(match);
    REQUIRE(count == 42875);
    REQUIRE(seen.size() == 42875u);
    // the written host values travel back with hostToDevice
    sg.hostToDevice<0>();
    REQUIRE(sg.get<0>(grid_key_dx<3>({303, 337, 320})) == 5.0f);
    REQUIRE(sg.get<0>(grid_key_dx<3>({302, 320, 320})) == 1.0f);
}

static void testIteratorHost()
{
    const size_t sz[3] = {512, 512, 512};
    SparseGridGpu<3, aggregate<float, float>, 4, 64, long int> sg(sz);
    sg.setBackgroundValue<0>(0);
    insertSphere3D(sg, 256, 64.0f, 32.0f);
    bool threw = false;
    try {
        sg.getIterator();
    } catch (const sg::error &) {
        threw = true;
    }
    REQUIRE(threw);  // no host copy yet
    sg.deviceToHost<0>();
    bool match = true;
    size_t count = 0;
    auto it = sg.getIterator();
    while (it.isNext()) {
        auto key = it.get();
        match = match && sg.get<0>(key) == 1.0f && (sg.getFlag(key) & 1);
        ++count;
        ++it;
    }
    REQUIRE(sg.countExistingElements() != 0);
    REQUIRE(sg.countExistingElements() == count);
    REQUIRE(match);
}

static void testCopyTo()
{
    const size_t sz[3] = {768, 768, 768};
    typedef SparseGridGpu<3, aggregate<float, float>, 4, 64, long int> SG;
    SG a(sz), b(sz);
    gpu::ofp_context_t ctx;
    a.setBackgroundValue<0>(0);
    b.setBackgroundValue<0>(0);
    insertSphere3D(a, 256, 32.0f, 0.0f);
    const size_t na = a.countExistingElements();
    b.copy_to(a, Box<3, size_t>({303, 303, 303}, {337, 337, 337}), Box<3, size_t>({10, 20, 30}, {44, 54, 64}));
    a.remove(Box<3, int>({0, 0, 0}, {319, 767, 767}));
    for (int ph : {PHASE1, PHASE2, PHASE3}) {
        a.removeCopyToFinalize<0>(ctx, ph);
        b.removeCopyToFinalize<0>(ctx, ph);
    }
    // the removal comes first, so only the x >= 320 part of the box is copied
    REQUIRE(b.countExistingElements() == 18 * 35 * 35);
    REQUIRE(b.get<0>(grid_key_dx<3>({44, 54, 64})) == 1.0f);
    REQUIRE(b.get<0>(grid_key_dx<3>({27, 20, 30})) == 1.0f);
    REQUIRE(b.get<0>(grid_key_dx<3>({26, 20, 30})) == 0.0f);
    REQUIRE(a.countExistingElements() < na);
    size_t left = 0;
    for (const auto &p : existingPoints(a)) left += p.get(0) >= 320;
    REQUIRE(left == a.countExistingElements());
}

static void testLinks()
{
    typedef SparseGridGpu<3, aggregate<float>, 4, 64, long int> SG;
    const size_t szc[3] = {32, 32, 32}, szf[3] = {64, 64, 64};
    SG c(szc), f(szf), f2(szf);
    gpu::ofp_context_t ctx;
    auto ball = [](SG &g, int n, int r) {
        std::vector<long> keys;
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y)
                for (int x = 0; x < n; ++x)
                    if ((x - n / 2) * (x - n / 2) + (y - n / 2) * (y - n / 2) + (z - n / 2) * (z - n / 2) < r * r)
                        keys.push_back(g.getLinId(grid_key_dx<3>({x, y, z})));
        std::vector<float> v(keys.size(), 1.0f);
        const void *vals[1] = {v.data()};
        g.insertPoints(keys.data(), keys.size(), vals);
        gpu::ofp_context_t cx;
        g.flush<smax_<0>>(cx, FLUSH_ON_DEVICE);
    };
    // the children of the coarse boundary (radius ~12 -> ~24) lie in f (radius 27); the
    // parents of the boundary of f2 (radius ~20 -> ~10) lie in the coarse ball
    ball(c, 32, 12);
    ball(f, 64, 27);
    ball(f2, 64, 20);
    c.tagBoundaries(ctx);
    f2.tagBoundaries(ctx);
    c.construct_link_dw(f, Box<3, int>({0, 0, 0}, {31, 31, 31}), grid_key_dx<3>({0, 0, 0}), ctx);
    f2.construct_link_up(c, Box<3, int>({0, 0, 0}, {63, 63, 63}), grid_key_dx<3>({0, 0, 0}), ctx);
    const auto pad = c.getLinkPaddingPoints(SG_LINK_DOWN);
    const auto off = c.getDownLinksOffsets();
    const auto dw = c.getDownLinks();
    const std::vector<long> cid = c.private_get_index_array(), fid = f.private_get_index_array();
    REQUIRE(!pad.empty() && off.size() == pad.size() && !dw.empty());
    bool ok = true;
    for (size_t i = 0; i < pad.size(); ++i) {
        const auto x = c.getCoord(cid[pad[i].chunk] * 64 + pad[i].offset);
        const size_t e = (i + 1 < off.size()) ? off[i + 1] : dw.size();
        for (size_t k = off[i]; k < e; ++k) {
            const auto q = f.getCoord(fid[dw[k].chunk] * 64 + dw[k].offset);
            for (int d = 0; d < 3; ++d) ok = ok && (q.get(d) - 2 * x.get(d) == 0 || q.get(d) - 2 * x.get(d) == 1);
        }
    }
    REQUIRE(ok);
    const auto padu = f2.getLinkPaddingPoints(SG_LINK_UP);
    const auto offu = f2.getUpLinksOffsets();
    const auto up = f2.getUpLinks();
    const std::vector<long> f2id = f2.private_get_index_array();
    REQUIRE(!padu.empty() && offu.size() == padu.size() && up.size() == padu.size());
    for (size_t i = 0; i < padu.size(); ++i) {
        const auto x = f2.getCoord(f2id[padu[i].chunk] * 64 + padu[i].offset);
        const size_t e = (i + 1 < offu.size()) ? offu[i + 1] : up.size();
        ok = ok && e - offu[i] <= 1;
        for (size_t k = offu[i]; k < e; ++k) {
            const auto q = c.getCoord(cid[up[k].chunk] * 64 + up[k].offset);
            for (int d = 0; d < 3; ++d) ok = ok && q.get(d) == x.get(d) / 2;
        }
    }
    REQUIRE(ok);
}

// testInsert on a Morton grid: the same known answers through SparseGridGpu_z
static void testInsertZ()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu_z<3, aggregate<float>, 8> sg(sz);
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(666);
    std::vector<long> keys;
    std::vector<float> vals;
    for (int z = 0; z < 64; z += 7)
        for (int y = 0; y < 64; y += 5)
            for (int x = 0; x < 64; x += 3) {
                keys.push_back(sg.getLinId(grid_key_dx<3>({x, y, z})));
                vals.push_back((float)(x + 100 * y + 10000 * z));
            }
    const void *v[1] = {vals.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
    bool match = true;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                const grid_key_dx<3> k({x, y, z});
                const bool on = x % 3 == 0 && y % 5 == 0 && z % 7 == 0;
                const float e = on ? (float)(x + 100 * y + 10000 * z) : 666.0f;
                match = match && sg.get<0>(k) == e;
                const auto c = sg.getCoord(sg.getLinId(k));
                match = match && c.get(0) == x && c.get(1) == y && c.get(2) == z;
            }
    REQUIRE(match);
}

// whole-grid pack into host memory (packRequest(req) / pack(mem, sts) / unpack(mem, ps),
// SparseGridGpu.hpp:2895-2959): a grid packed and unpacked into an empty one gives the same
// chunks, masks and properties; a pack of one property keeps the receiver's others
static void testHostPackUnpack()
{
    typedef SparseGridGpu<3, aggregate<float, int>, 8> SG;
    const size_t sz[3] = {64, 64, 64};
    gpu::ofp_context_t ctx;
    SG a(sz);
    a.setBackgroundValue<1>(-3);
    std::vector<long> keys;
    std::vector<float> f;
    std::vector<int> iv;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                const double r = std::sqrt((x - 32.0) * (x - 32.0) + (y - 32.0) * (y - 32.0) + (z - 32.0) * (z - 32.0));
                if (r >= 12.0 && r < 17.0) {
                    keys.push_back(a.getLinId(grid_key_dx<3, int>({x, y, z})));
                    f.push_back((float)(0.5 * x - 0.25 * y + z));
                    iv.push_back(x * 100 + y);
                }
            }
    const void *v[] = {f.data(), iv.data()};
    a.insertPoints(keys.data(), keys.size(), v);
    a.flush<smax_<0>, smax_<1>>(ctx, FLUSH_ON_DEVICE);
    a.findNeighbours();
    a.tagBoundaries(ctx);
    size_t req = 0;
    a.packRequest(req);
    HeapMemory pmem;
    ExtPreAlloc<HeapMemory> &mem = *(new ExtPreAlloc<HeapMemory>(req, pmem));
    mem.incRef();
    Pack_stat sts;
    a.pack(mem, sts);
    REQUIRE(sts.getOffset() == req);
    SG b(sz);
    b.setBackgroundValue<1>(-3);
    Unpack_stat ps;
    b.unpack(mem, ps);
    REQUIRE(ps.getOffset() == req);
    REQUIRE(b.private_get_index_array() == a.private_get_index_array());
    REQUIRE(b.private_get_mask_array() == a.private_get_mask_array());
    REQUIRE(b.private_get_data_array<0>() == a.private_get_data_array<0>());
    REQUIRE(b.private_get_data_array<1>() == a.private_get_data_array<1>());
    REQUIRE(b.countBoundaryElements() == a.countBoundaryElements());
    // property 1 alone into a grid holding property 0 in a few of the chunks
    size_t req1 = 0;
    a.packRequest<1>(req1);
    REQUIRE(req1 < req);
    HeapMemory pmem1;
    ExtPreAlloc<HeapMemory> mem1(req1, pmem1);
    Pack_stat sts1;
    a.pack<1>(mem1, sts1);
    SG c(sz);
    c.setBackgroundValue<1>(-3);
    std::vector<long> k2(keys.begin(), keys.begin() + 100);
    std::vector<float> f2(100, 7.0f);
    std::vector<int> i2(100, 5);
    const void *v2[] = {f2.data(), i2.data()};
    c.insertPoints(k2.data(), k2.size(), v2);
    c.flush<smax_<0>, smax_<1>>(ctx, FLUSH_ON_DEVICE);
    Unpack_stat ps1;
    c.unpack<1>(mem1, ps1);
    REQUIRE(c.private_get_index_array() == a.private_get_index_array());
    REQUIRE(c.private_get_data_array<1>() == a.private_get_data_array<1>());
    for (long k : k2) REQUIRE(c.get<0>(c.getCoord(k)) == 7.0f);
    REQUIRE(c.get<0>(c.getCoord(keys[200])) == 0.0f);
    // another domain: refused
    const size_t sz2[3] = {32, 64, 64};
    SparseGridGpu<3, aggregate<float, int>, 8> d(sz2);
    Unpack_stat ps2;
    bool threw = false;
    try {
        d.unpack(mem, ps2);
    } catch (const sg::error &) {
        threw = true;
    }
    REQUIRE(threw);
    delete &mem;
}

int main()
{
    testInsert();
    testFlushOnHost();
    testTagBoundaries();
    testStencilHeat();
    testConv3x3x3<SparseGridGpu<3, aggregate<float, float>>>();
    testConv3x3x3<SparseGridGpu_z<3, aggregate<float, float>>>();
    testPackRequestWithIterator();
    testRemove();
    testPackUnpack();
    testInsertZ();
    testIteratorSubHost();
    testIteratorHost();
    testCopyTo();
    testLinks();
    testHostPackUnpack();
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("facade tests passed\n");
    return 0;
}
