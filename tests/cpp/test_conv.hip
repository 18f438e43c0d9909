// The conv family with user __device__ lambdas (SparseGridGpu.hpp:2568-2683 of the
// reference) and device inserts through toKernel(), through the facade compiled by hipcc.  Each lambda form is checked against a
// path pinned elsewhere: the library's named stencils (themselves checked against the
// oracle), or values computed on the host from the grid's host copy.
// Build: hipcc -std=c++17 --offload-arch=gfx950 -ffp-contract=off -I include
//        tests/cpp/test_conv.hip -L openfpm_data_amd -lsg_hip
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "SparseGridGpu.hpp"

#define REQUIRE(c)                                                                      \
    do {                                                                                \
        if (!(c)) {                                                                     \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                  \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

typedef aggregate<float, float, float, float, float, float> Props6;

// a 64^3 shell (10 <= r < 20) with small integer values in properties 0, 1 and 3 (sums of
// them are exact in FP32, whatever the order), then star tagBoundaries
template <typename SG> static void build_shell(SG &sg)
{
    gpu::ofp_context_t ctx;
    std::vector<long> keys;
    std::vector<float> a, b, zero, c;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                const double r = std::sqrt((x - 32.0) * (x - 32.0) + (y - 32.0) * (y - 32.0) + (z - 32.0) * (z - 32.0));
                if (r >= 10.0 && r < 20.0) {
                    keys.push_back(sg.getLinId(grid_key_dx<3, int>({x, y, z})));
                    a.push_back((float)((x * 7 + y * 3 + z) % 17));
                    b.push_back((float)((x + 2 * y + 5 * z) % 11));
                    c.push_back((float)((x * y + z) % 5));
                    zero.push_back(0.0f);
                }
            }
    const void *v[] = {a.data(), b.data(), zero.data(), c.data(), zero.data(), zero.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.template flush<smax_<0>, smax_<1>, smax_<2>, smax_<3>, smax_<4>, smax_<5>>(ctx, FLUSH_ON_DEVICE);
    sg.template findNeighbours<NNStar<3>>();
    sg.template tagBoundaries<NNStar<3>>(ctx);
}

static bool same(const std::vector<float> &x, const std::vector<float> &y)
{
    if (x.size() != y.size()) return false;
    for (size_t i = 0; i < x.size(); ++i)
        if (!(x[i] == y[i])) return false;
    return true;
}

// conv (cp_block taps) against the named Laplacian, over the whole grid and over a box
template <typename SG> static void testConvLaplacian()
{
    const size_t sz[3] = {64, 64, 64};
    for (int boxed = 0; boxed < 2; ++boxed) {
        SG sg(sz);
        build_shell(sg);
        grid_key_dx<3, int> lo({0, 0, 0}), hi({63, 63, 63});
        if (boxed) {
            lo = grid_key_dx<3, int>({14, 20, 9});
            hi = grid_key_dx<3, int>({40, 51, 33});
        }
        Box<3, int> box;
        for (int d = 0; d < 3; ++d) {
            box.setLow(d, lo.get(d));
            box.setHigh(d, hi.get(d));
        }
        sg.template conv<0, 2, 1>(lo, hi, [] __device__(auto &u, int i, int j, int k) -> float {
            float s = -6.0f * u(i, j, k);
            s = s + (u(i - 1, j, k) + u(i + 1, j, k));
            s = s + (u(i, j - 1, k) + u(i, j + 1, k));
            s = s + (u(i, j, k - 1) + u(i, j, k + 1));
            return s;
        });
        sg.template applyStencils<LaplacianStencil<3, 0, 4>>(box, STENCIL_MODE_INPLACE, -6.0, 1.0);
        REQUIRE(same(sg.template private_get_data_array<2>(), sg.template private_get_data_array<4>()));
    }
}

// conv_cross (face neighbours, an extra argument) against the named Laplacian on the
// interior points; padding points take the source value
static void testConvCross()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu<3, Props6> sg(sz);
    build_shell(sg);
    sg.template conv_cross<0, 2, 1>(grid_key_dx<3, int>({0, 0, 0}), grid_key_dx<3, int>({63, 63, 63}),
                                    [] __device__(float cur, const sg::cross_stencil<3, float> &cs, float c1) -> float {
                                        float s = -6.0f * cur;
                                        s = s + c1 * (cs.xm[0] + cs.xp[0]);
                                        s = s + c1 * (cs.xm[1] + cs.xp[1]);
                                        s = s + c1 * (cs.xm[2] + cs.xp[2]);
                                        return s;
                                    },
                                    1.0f);
    sg.template applyStencils<LaplacianStencil<3, 0, 4>>(sg.getBox(), STENCIL_MODE_INPLACE, -6.0, 1.0);
    const std::vector<uint8_t> m = sg.private_get_mask_array();
    const std::vector<float> u = sg.template private_get_data_array<0>(), a = sg.template private_get_data_array<2>(),
                             b = sg.template private_get_data_array<4>();
    size_t interior = 0, padding = 0;
    for (size_t i = 0; i < m.size(); ++i) {
        if (m[i] == 1) {
            REQUIRE(a[i] == b[i]);
            ++interior;
        } else if (m[i] == 3) {
            REQUIRE(a[i] == u[i]);
            ++padding;
        } else {
            REQUIRE(a[i] == 0.0f);
        }
    }
    REQUIRE(interior > 10000 && padding > 1000);
}

// conv with diagonal taps (all-ones 3x3x3) against the named Conv3x3x3 on the interior
static void testConvBox()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu<3, Props6> sg(sz);
    build_shell(sg);
    sg.template findNeighbours<NNFull<3>>();
    sg.template setNNType<NNFull<3>>();
    sg.template conv<0, 2, 1>(grid_key_dx<3, int>({0, 0, 0}), grid_key_dx<3, int>({63, 63, 63}),
                              [] __device__(auto &u, int i, int j, int k) -> float {
                                  float s = 0.0f;
                                  for (int a = -1; a <= 1; ++a)
                                      for (int b = -1; b <= 1; ++b)
                                          for (int c = -1; c <= 1; ++c) s = s + u(i + a, j + b, k + c);
                                  return s;
                              });
    conv_coeff cc;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) cc.coeff[k][j][i] = 1.0f;
    sg.template applyStencils<Conv3x3x3<3, 0, 4>>(sg.getBox(), STENCIL_MODE_INPLACE, cc);
    const std::vector<uint8_t> m = sg.private_get_mask_array();
    const std::vector<float> a = sg.template private_get_data_array<2>(), b = sg.template private_get_data_array<4>();
    size_t n = 0;
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] == 1) {
            REQUIRE(a[i] == b[i]);
            ++n;
        }
    REQUIRE(n > 10000);
}

// conv2 Gray-Scott in the library kernel's operation order against the named Gray-Scott
// (FP32, bit-exact: both evaluate the same operations without contraction)
static void testConv2GrayScott()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu<3, Props6> sg(sz);
    build_shell(sg);
    const float Du = (float)2e-5, Dv = (float)1e-5, F = (float)0.04, K = (float)0.06, dt = 1.0f;
    sg.template conv2<0, 1, 2, 3 + 1, 1>(
        grid_key_dx<3, int>({0, 0, 0}), grid_key_dx<3, int>({63, 63, 63}),
        [=] __device__(float &ru, float &rv, auto &u, auto &v, int i, int j, int k) {
            const float cu = u(i, j, k), cv = v(i, j, k);
            float lu = -6.0f * cu;
            lu = lu + (u(i - 1, j, k) + u(i + 1, j, k));
            lu = lu + (u(i, j - 1, k) + u(i, j + 1, k));
            lu = lu + (u(i, j, k - 1) + u(i, j, k + 1));
            float lv = -6.0f * cv;
            lv = lv + (v(i - 1, j, k) + v(i + 1, j, k));
            lv = lv + (v(i, j - 1, k) + v(i, j + 1, k));
            lv = lv + (v(i, j, k - 1) + v(i, j, k + 1));
            const float uvv = (cu * cv) * cv;
            ru = cu + dt * ((Du * lu - uvv) + F * (1.0f - cu));
            rv = cv + dt * ((Dv * lv + uvv) - (F + K) * cv);
        });
    const std::vector<float> ru = sg.template private_get_data_array<2>(), rv = sg.template private_get_data_array<4>();
    sg.template applyStencils<GrayScottStencil<3, 0, 1, 5, 3>>(sg.getBox(), STENCIL_MODE_INPLACE, 2e-5, 1e-5, 0.04,
                                                               0.06, 1.0);
    const std::vector<uint8_t> m = sg.private_get_mask_array();
    const std::vector<float> nu = sg.template private_get_data_array<5>(), nv = sg.template private_get_data_array<3>();
    size_t n = 0;
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] == 1) {
            REQUIRE(ru[i] == nu[i] && rv[i] == nv[i]);
            ++n;
        }
    REQUIRE(n > 10000);
}

// host reads of a 64^3 grid's arrays by coordinate (background 0 outside the chunks)
struct HostView {
    std::vector<long> ids;
    std::vector<uint8_t> m;
    std::vector<std::vector<float>> p;
    template <typename SG> explicit HostView(const SG &sg)
    {
        ids = sg.private_get_index_array();
        m = sg.private_get_mask_array();
        p.push_back(sg.template private_get_data_array<0>());
        p.push_back(sg.template private_get_data_array<1>());
        p.push_back(sg.template private_get_data_array<2>());
        p.push_back(sg.template private_get_data_array<3>());
        p.push_back(sg.template private_get_data_array<4>());
        p.push_back(sg.template private_get_data_array<5>());
    }
    long slot(int x, int y, int z) const
    {
        if (x < 0 || y < 0 || z < 0 || x >= 64 || y >= 64 || z >= 64) return -1;
        const long cid = ((long)(z / 8) * 8 + y / 8) * 8 + x / 8;
        const auto it = std::lower_bound(ids.begin(), ids.end(), cid);
        if (it == ids.end() || *it != cid) return -1;
        return (long)(it - ids.begin()) * 512 + ((z % 8) * 8 + y % 8) * 8 + x % 8;
    }
    float get(int q, int x, int y, int z) const
    {
        const long s = slot(x, y, z);
        return s < 0 ? 0.0f : p[q][s];
    }
    uint8_t flag(int x, int y, int z) const
    {
        const long s = slot(x, y, z);
        return s < 0 ? 0 : m[s];
    }
};

// the _b forms read the chunk's own data of another property through the block argument
static void testBlockForms()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu<3, Props6> sg(sz);
    build_shell(sg);
    const grid_key_dx<3, int> lo({0, 0, 0}), hi({63, 63, 63});
    sg.template conv_cross_b<0, 2, 1>(lo, hi, [] __device__(auto &u, auto &blk, int off, int i, int j, int k) -> float {
        return u(i + 1, j, k) + 2.0f * blk.template get<3>()[off];
    });
    sg.template conv2_b<0, 1, 4, 5, 1>(lo, hi,
                                       [] __device__(float &r1, float &r2, auto &u, auto &v, auto &blk, int off, int i,
                                                     int j, int k) {
                                           r1 = u(i, j - 1, k) - blk.template get<3>()[off];
                                           r2 = v(i, j, k + 1) + blk.template get<0>()[off];
                                       });
    const HostView h(sg);
    size_t n = 0;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                if (h.flag(x, y, z) != 1) continue;
                const float w = h.get(3, x, y, z), u = h.get(0, x, y, z);
                REQUIRE(h.get(2, x, y, z) == h.get(0, x + 1, y, z) + 2.0f * w);
                REQUIRE(h.get(4, x, y, z) == h.get(0, x, y - 1, z) - w);
                REQUIRE(h.get(5, x, y, z) == h.get(1, x, y, z + 1) + u);
                ++n;
            }
    REQUIRE(n > 10000);
    // conv3_b: three sources, three destinations
    SparseGridGpu<3, Props6> s3(sz);
    build_shell(s3);
    s3.template conv3_b<0, 1, 3, 2, 4, 5, 1>(lo, hi,
                                             [] __device__(float &r1, float &r2, float &r3, auto &a, auto &b, auto &c,
                                                           auto &blk, int off, int i, int j, int k) {
                                                 r1 = a(i - 1, j, k) + b(i, j, k);
                                                 r2 = c(i, j + 1, k) * 3.0f;
                                                 r3 = a(i, j, k - 1) + blk.template get<1>()[off];
                                             });
    const HostView h3(s3);
    n = 0;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                if (h3.flag(x, y, z) != 1) continue;
                REQUIRE(h3.get(2, x, y, z) == h3.get(0, x - 1, y, z) + h3.get(1, x, y, z));
                REQUIRE(h3.get(4, x, y, z) == h3.get(3, x, y + 1, z) * 3.0f);
                REQUIRE(h3.get(5, x, y, z) == h3.get(0, x, y, z - 1) + h3.get(1, x, y, z));
                ++n;
            }
    REQUIRE(n > 10000);
}

// a ghost layer of 2 points (stencil_size 2): taps two points away, across chunk faces,
// edges and the domain boundary (background 0 outside)
static void testConvRadius2()
{
    const size_t sz[3] = {64, 64, 64};
    SparseGridGpu<3, Props6> sg(sz);
    build_shell(sg);
    sg.template conv<0, 2, 2>(grid_key_dx<3, int>({0, 0, 0}), grid_key_dx<3, int>({63, 63, 63}),
                              [] __device__(auto &u, int i, int j, int k) -> float {
                                  return u(i + 2, j, k) - 2.0f * u(i, j - 2, k) + 3.0f * u(i - 1, j + 1, k + 2);
                              });
    const HostView h(sg);
    size_t n = 0;
    for (int z = 0; z < 64; ++z)
        for (int y = 0; y < 64; ++y)
            for (int x = 0; x < 64; ++x) {
                if (h.flag(x, y, z) != 1) continue;
                REQUIRE(h.get(2, x, y, z) ==
                        h.get(0, x + 2, y, z) - 2.0f * h.get(0, x, y - 2, z) + 3.0f * h.get(0, x - 1, y + 1, z + 2));
                ++n;
            }
    REQUIRE(n > 10000);
}

// 2D: conv of a 5-point Laplacian against the named Laplacian (FP64, integer values)
static void testConv2D()
{
    const size_t sz[2] = {200, 150};
    SparseGridGpu<2, aggregate<double, double, double>> sg(sz);
    gpu::ofp_context_t ctx;
    std::vector<long> keys;
    std::vector<double> a, zero;
    for (int y = 0; y < 150; ++y)
        for (int x = 0; x < 200; ++x) {
            const double r = std::sqrt((x - 100.0) * (x - 100.0) + (y - 75.0) * (y - 75.0));
            if (r >= 20.0 && r < 60.0) {
                keys.push_back(sg.getLinId(grid_key_dx<2, int>({x, y})));
                a.push_back((double)((x * 5 + y) % 13));
                zero.push_back(0.0);
            }
        }
    const void *v[] = {a.data(), zero.data(), zero.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.template flush<smax_<0>, smax_<1>, smax_<2>>(ctx, FLUSH_ON_DEVICE);
    sg.template findNeighbours<NNStar<2>>();
    sg.template tagBoundaries<NNStar<2>>(ctx);
    sg.template conv<0, 1, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({199, 149}),
                              [] __device__(auto &u, int i, int j) -> double {
                                  return -4.0 * u(i, j) + (u(i - 1, j) + u(i + 1, j)) + (u(i, j - 1) + u(i, j + 1));
                              });
    sg.template applyStencils<LaplacianStencil<2, 0, 2>>(sg.getBox(), STENCIL_MODE_INPLACE, -4.0, 1.0);
    const std::vector<uint8_t> m = sg.private_get_mask_array();
    const std::vector<double> x = sg.template private_get_data_array<1>(), y = sg.template private_get_data_array<2>();
    size_t n = 0;
    for (size_t i = 0; i < m.size(); ++i)
        if (m[i] == 1) {
            REQUIRE(x[i] == y[i]);
            ++n;
        }
    REQUIRE(n > 1000);
}

// ---- the reference's own conv tests (SparseGridGpu_tests.cu:519-792), restated ----

// 16 x 16 domain fully inserted (chunks of 8 x 8), star tagBoundaries; the boundary
// (padding) points of property q get vmax[q] * x / 15 (BoundaryStencilSetXRescaled), the
// others 0
typedef SparseGridGpu<2, aggregate<float, float, float, float>, 8, 64> Grid2;
static void build_square(Grid2 &sg, const float (&vmax)[4])
{
    gpu::ofp_context_t ctx;
    sg.setBackgroundValue<0>(0);
    std::vector<long> keys;
    std::vector<float> zero(256, 0.0f);
    for (int y = 0; y < 16; ++y)
        for (int x = 0; x < 16; ++x) keys.push_back(sg.getLinId(grid_key_dx<2, int>({x, y})));
    const void *v[] = {zero.data(), zero.data(), zero.data(), zero.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    sg.flush<smax_<0>, smax_<1>, smax_<2>, smax_<3>>(ctx, FLUSH_ON_DEVICE);
    sg.findNeighbours();
    sg.tagBoundaries(ctx);
    const std::vector<long> ids = sg.private_get_index_array();
    const std::vector<uint8_t> m = sg.private_get_mask_array();
    std::vector<float> u[4];
    for (int q = 0; q < 4; ++q) u[q].assign(m.size(), 0.0f);
    for (size_t i = 0; i < m.size(); ++i) {
        auto c = sg.getCoord(ids[i / 64] * 64 + (long)(i % 64));
        if (m[i] & PADDING)
            for (int q = 0; q < 4; ++q) u[q][i] = vmax[q] * c.get(0) / 15.0f;
    }
    sg.hostToDevice<0>(u[0]);
    sg.hostToDevice<1>(u[1]);
    sg.hostToDevice<2>(u[2]);
    sg.hostToDevice<3>(u[3]);
}

// steady state of 1000 Jacobi pairs: linear in x within 1e-2
static void require_linear(Grid2 &sg, int q, float vmax)
{
    const std::vector<long> ids = sg.private_get_index_array();
    std::vector<float> u;
    if (q == 0) u = sg.private_get_data_array<0>();
    else u = sg.private_get_data_array<1>();
    bool match = true;
    for (size_t i = 0; i < u.size(); ++i) {
        auto c = sg.getCoord(ids[i / 64] * 64 + (long)(i % 64));
        match &= std::fabs(u[i] - vmax * c.get(0) / 15.0f) < 1e-2f;
    }
    REQUIRE(match);
}

// testStencil_lap_simplified (:519-568): conv_cross Jacobi heat; only property 0 carries
// the boundary values, the block store of conv_cross copies them to property 1
static void testRefLapCross()
{
    const size_t sz[2] = {16, 16};
    Grid2 sg(sz);
    build_square(sg, {10.0f, 0.0f, 0.0f, 0.0f});
    for (int it = 0; it < 1000; ++it) {
        sg.conv_cross<0, 1, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({16, 16}),
                               [] __device__(float &u, cross_stencil<2, float> &cs) {
                                   return u + (cs.xm[0] + cs.xp[0] + cs.xm[1] + cs.xp[1] - 4.0 * u) * 0.1;
                               });
        sg.conv_cross<1, 0, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({16, 16}),
                               [] __device__(float &u, cross_stencil<2, float> &cs) {
                                   return u + (cs.xm[0] + cs.xp[0] + cs.xm[1] + cs.xp[1] - 4.0 * u) * 0.1;
                               });
    }
    require_linear(sg, 0, 10.0f);
}

// testStencil_lap_no_cross_simplified (:570-645): conv with cp_block taps
static void testRefLapConv()
{
    const size_t sz[2] = {16, 16};
    Grid2 sg(sz);
    build_square(sg, {10.0f, 10.0f, 0.0f, 0.0f});
    typedef GetCpBlockType<Grid2, 0, 1>::type CpBlockType;
    for (int it = 0; it < 1000; ++it) {
        sg.conv<0, 1, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({16, 16}),
                         [] __device__(CpBlockType & u, int i, int j) {
                             float c = u(i, j);
                             return c + (u(i - 1, j) + u(i + 1, j) + u(i, j - 1) + u(i, j + 1) - 4.0 * c) * 0.1;
                         });
        sg.conv<1, 0, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({16, 16}),
                         [] __device__(CpBlockType & u, int i, int j) {
                             float c = u(i, j);
                             return c + (u(i - 1, j) + u(i + 1, j) + u(i, j - 1) + u(i, j + 1) - 4.0 * c) * 0.1;
                         });
    }
    require_linear(sg, 0, 10.0f);
}

// testStencil_lap_no_cross_simplified2 (:647-736): conv2 on two fields at once
static void testRefLapConv2()
{
    const size_t sz[2] = {16, 16};
    Grid2 sg(sz);
    build_square(sg, {10.0f, 5.0f, 10.0f, 5.0f});
    typedef GetCpBlockType<Grid2, 0, 1>::type CpBlockType;
    auto step = [] __device__(float &u_out, float &v_out, CpBlockType &u, CpBlockType &v, int i, int j) {
        float cu = u(i, j);
        float cv = v(i, j);
        u_out = cu + (u(i - 1, j) + u(i + 1, j) + u(i, j - 1) + u(i, j + 1) - 4.0 * cu) * 0.1;
        v_out = cv + (v(i - 1, j) + v(i + 1, j) + v(i, j - 1) + v(i, j + 1) - 4.0 * cv) * 0.1;
    };
    for (int it = 0; it < 1000; ++it) {
        sg.conv2<0, 1, 2, 3, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({16, 16}), step);
        sg.conv2<2, 3, 0, 1, 1>(grid_key_dx<2, int>({0, 0}), grid_key_dx<2, int>({16, 16}), step);
    }
    require_linear(sg, 0, 10.0f);
    require_linear(sg, 1, 5.0f);
}

// testStencil_lap_no_cross_simplified_subset (:738-792): conv restricted to [3, 11]^2
static void testRefConvSubset()
{
    const size_t sz[2] = {16, 16};
    Grid2 sg(sz);
    build_square(sg, {0.0f, 0.0f, 0.0f, 0.0f});
    typedef GetCpBlockType<Grid2, 0, 1>::type CpBlockType;
    sg.conv<0, 1, 1>(grid_key_dx<2, int>({3, 3}), grid_key_dx<2, int>({11, 11}),
                     [] __device__(CpBlockType &, int, int) { return 5.0; });
    bool match = true;
    for (long i = 0; i < 64 * 4; ++i) {
        auto coord = sg.getCoord(i);
        const bool in = coord.get(0) >= 3 && coord.get(1) >= 3 && coord.get(0) <= 11 && coord.get(1) <= 11;
        match &= sg.get<1>(coord) == (in ? 5.0f : 0.0f);
    }
    REQUIRE(match);
}

// ---- device inserts through toKernel() (SparseGridGpu_tests.cu:15-75, :209-282) ----

template <unsigned int p, typename KerT> __global__ void insertValues(KerT sg)
{
    sg.init();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    grid_key_dx<KerT::d, int> coord({x, y, z});
    sg.template insert<p>(coord) = x;
    sg.flush_block_insert();
}

template <unsigned int p1, unsigned int p2, typename KerT> __global__ void insertConstantValue2(KerT sg, float value)
{
    constexpr unsigned int pMask = KerT::pMask;
    sg.init();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    grid_key_dx<KerT::d, int> coord({x, y, z});
    const long pos = sg.getLinId(coord);
    const long chunk = pos / KerT::chunk_size;
    const int offset = (int)(pos % KerT::chunk_size);
    auto encap = sg.insertBlock(chunk);
    encap.template get<p1>()[offset] = value;
    encap.template get<p2>()[offset] = value;
    BlockMapGpu_ker<>::setExist(encap.template get<pMask>()[offset]);
    __syncthreads();
    sg.flush_block_insert();
}

// testInsert (:209-244): 2D, 2 x 2 blocks of 8 x 8 threads, value = x, background 666
static void testDeviceInsert2D()
{
    const size_t sz[2] = {16, 16};
    SparseGridGpu<2, aggregate<float>, 8> sg(sz);
    sg.setBackgroundValue<0>(666);
    const dim3 gridSize(2, 2), blockSize(8, 8);
    sg.setGPUInsertBuffer(gridSize, blockSize);
    hipLaunchKernelGGL((insertValues<0>), gridSize, blockSize, 0, 0, sg.toKernel());
    gpu::ofp_context_t ctx;
    sg.flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
    bool match = true;
    for (long i = 0; i < 4 * 64; ++i) {
        auto coord = sg.getCoord(i);
        match &= sg.get<0>(coord) == (float)coord.get(0);
    }
    REQUIRE(match);
    REQUIRE(sg.countExistingElements() == 256);
}

// testInsert3D (:246-282): 3D chunks of 4^3, 2 x 2 x 2 blocks of 4^3 threads
static void testDeviceInsert3D()
{
    const size_t sz[3] = {8, 8, 8};
    SparseGridGpu<3, aggregate<float>, 4> sg(sz);
    sg.setBackgroundValue<0>(666);
    const dim3 gridSize(2, 2, 2), blockSize(4, 4, 4);
    sg.setGPUInsertBuffer(gridSize, blockSize);
    hipLaunchKernelGGL((insertValues<0>), gridSize, blockSize, 0, 0, sg.toKernel());
    gpu::ofp_context_t ctx;
    sg.flush<smax_<0>>(ctx, FLUSH_ON_DEVICE);
    bool match = true;
    for (long i = 0; i < 64 * 4; ++i) {
        auto coord = sg.getCoord(i);
        match &= sg.get<0>(coord) == (float)coord.get(0);
    }
    REQUIRE(match);
    REQUIRE(sg.countExistingElements() == 512);
}

// insertBlock + setExist of two properties (insertConstantValue2, :15-45) with 8 waves per
// block writing one chunk, merged with smax; then an overflowing buffer is reported
static void testDeviceInsertBlock()
{
    const size_t sz[3] = {16, 16, 16};
    SparseGridGpu<3, aggregate<float, float, float>> sg(sz);
    const dim3 gridSize(2, 2, 2), blockSize(8, 8, 8);
    sg.setGPUInsertBuffer(gridSize, blockSize);
    hipLaunchKernelGGL((insertConstantValue2<0, 2>), gridSize, blockSize, 0, 0, sg.toKernel(), 3.0f);
    gpu::ofp_context_t ctx;
    sg.flush<smax_<0>, smax_<1>, smax_<2>>(ctx, FLUSH_ON_DEVICE);
    REQUIRE(sg.countExistingElements() == 4096);
    const std::vector<float> a = sg.private_get_data_array<0>(), b = sg.private_get_data_array<1>(),
                             c = sg.private_get_data_array<2>();
    for (size_t i = 0; i < 4096; ++i) REQUIRE(a[i] == 3.0f && b[i] == 0.0f && c[i] == 3.0f);
    // a buffer of one record for eight chunks: claims beyond it go to the sink record, and
    // the flush reports the overflow
    SparseGridGpu<3, aggregate<float, float, float>> s2(sz);
    s2.setGPUInsertBuffer(1, 1);
    hipLaunchKernelGGL((insertConstantValue2<0, 1>), gridSize, blockSize, 0, 0, s2.toKernel(), 1.0f);
    bool thrown = false;
    try {
        s2.flush<smax_<0>, smax_<1>, smax_<2>>(ctx, FLUSH_ON_DEVICE);
    } catch (const sg::error &) {
        thrown = true;
    }
    REQUIRE(thrown);
}

// host point records first, then device chunk records into the same chunks: the flush
// folds them in that order (sadd of integer-valued floats: host value + device value)
static void testDeviceInsertAfterHostPoints()
{
    const size_t sz[3] = {8, 8, 8};
    SparseGridGpu<3, aggregate<float>, 4> sg(sz);
    std::vector<long> keys;
    std::vector<float> vals;
    for (int z = 0; z < 8; ++z)
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; x += 2) {
                keys.push_back(sg.getLinId(grid_key_dx<3, int>({x, y, z})));
                vals.push_back(100.0f);
            }
    const void *v[] = {vals.data()};
    sg.insertPoints(keys.data(), keys.size(), v);
    const dim3 gridSize(2, 2, 2), blockSize(4, 4, 4);
    sg.setGPUInsertBuffer(gridSize, blockSize);
    hipLaunchKernelGGL((insertValues<0>), gridSize, blockSize, 0, 0, sg.toKernel());
    gpu::ofp_context_t ctx;
    sg.flush<sadd_<0>>(ctx, FLUSH_ON_DEVICE);
    bool match = true;
    for (int z = 0; z < 8; ++z)
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x)
                match &= sg.get<0>(grid_key_dx<3, int>({x, y, z})) == (float)x + ((x % 2 == 0) ? 100.0f : 0.0f);
    REQUIRE(match);
    REQUIRE(sg.countExistingElements() == 512);
}

int main()
{
    testDeviceInsertAfterHostPoints();
    testDeviceInsert2D();
    testDeviceInsert3D();
    testDeviceInsertBlock();
    testRefLapCross();
    testRefLapConv();
    testRefLapConv2();
    testRefConvSubset();
    testConvLaplacian<SparseGridGpu<3, Props6>>();
    testConvLaplacian<SparseGridGpu_z<3, Props6>>();
    testConvCross();
    testConvBox();
    testConv2GrayScott();
    testBlockForms();
    testConvRadius2();
    testConv2D();
    std::printf("conv tests passed\n");
    return 0;
}
