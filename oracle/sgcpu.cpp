// sgcpu.cpp -- CPU restatement of openfpm_data's SparseGridGpu hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is the parity oracle for libsg_hip.so and the
// CPU baseline timed by bench.py; nothing in the product library links, loads or
// calls it.  Each function restates the reference algorithm step by step and cites
// the reference file:line it follows (paths relative to openfpm_data/src).
//
// Where this restatement is pinned (see tests/test_oracle_*.py):
//   linearizer      closed form + grid_smb_tests.cpp:11-36 known ids
//   flush           BlockMapGpu_tests.cu / SparseGridGpu_tests.cu known answers,
//                   brute-force dictionary model on random tiny inputs (all ops)
//   findNeighbours  brute-force geometric neighbour search (+ the NNFull low-face
//                   aliasing, checked against the closed form of grid_smb::BlockLinId)
//   tagBoundaries   SparseGridGpu_tests.cu:284-470 exact padding index lists,
//                   brute-force neighbour test on random 3D shells
//   stencils        closed forms (Laplacian of linear / quadratic fields), the
//                   3x3x3 all-ones conv = 27*value known answer
//                   (SparseGridGpu_tests.cu:1012-1177), HeatStencil steady state.
//
// Built by oracle/Makefile with g++ -O3 -fopenmp -ffp-contract=off (no FMA
// contraction, so FP results follow the written operation order exactly).

#include "../include/sparsegrid_gpu.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <omp.h>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

size_t dtype_size(int dt)
{
    switch (dt) {
    case SG_F32: case SG_I32: case SG_U32: return 4;
    case SG_F64: case SG_I64: case SG_U64: return 8;
    case SG_I8: case SG_U8: return 1;
    case SG_I16: case SG_U16: return 2;
    default: return 0;
    }
}

bool dtype_is_float(int dt) { return dt == SG_F32 || dt == SG_F64; }

// dispatch a generic lambda on the scalar type of a property
template <typename F>
void dispatch(int dt, F &&f)
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

// Wrapping add / mul for integers, done in an unsigned type of at least 32 bits (narrower
// types would be promoted to signed int, where a product can overflow) and truncated to T;
// plain + / * for FP.
template <typename T> T op_add(T a, T b)
{
    if constexpr (std::is_integral<T>::value) {
        typedef typename std::conditional<(sizeof(T) < 4), uint32_t, typename std::make_unsigned<T>::type>::type U;
        return (T)(U)((U)a + (U)b);
    } else {
        return a + b;
    }
}

template <typename T> T op_mul(T a, T b)
{
    if constexpr (std::is_integral<T>::value) {
        typedef typename std::conditional<(sizeof(T) < 4), uint32_t, typename std::make_unsigned<T>::type>::type U;
        return (T)(U)((U)a * (U)b);
    } else {
        return a * b;
    }
}

// Binary reduction operators.  smax_/smin_ use gpu::maximum_t / minimum_t from
// the (absent) openfpm_devices operators header, modelled on moderngpu's
// max(a,b) = (a < b) ? b : a and min(a,b) = (b < a) ? b : a;
// sadd_ = plus_t, sRight_ = b, sLeft_ = a (map_vector_sparse_cuda_kernels.cuh:42-97, :99-200).
template <typename T> T reduce_op(int op, T a, T b)
{
    switch (op) {
    case SG_OP_MAX: return (a < b) ? b : a;
    case SG_OP_MIN: return (b < a) ? b : a;
    case SG_OP_ADD: return op_add(a, b);
    case SG_OP_RIGHT: return b;
    case SG_OP_LEFT: return a;
    case SG_OP_OR:
        if constexpr (std::is_integral<T>::value) return (T)(a | b);
        else return a;
    default: return b;
    }
}

struct Grid {
    int dim = 3, edge = 8, csz = 512, r = 1;
    int64_t sz[3] = {1, 1, 1}, bsz[3] = {1, 1, 1};
    int np = 0;
    int dt[16];
    size_t ts[16];

    // chunk map (BlockMapGpu.hpp:63-69; map_vector_sparse.hpp:790-791)
    std::vector<int64_t> keys;                     // sorted chunk ids
    std::vector<std::vector<unsigned char>> data;  // per property, (n+1)*csz scalars
    std::vector<uint8_t> mask;                     // (n+1)*csz
    std::vector<std::vector<unsigned char>> bg;    // background scalar per property

    // insert slots in insertion order; slot s owns entries [s_beg[s], s_beg[s+1])
    std::vector<int64_t> s_key;
    std::vector<int64_t> s_beg{0};
    std::vector<int32_t> e_off;
    std::vector<uint8_t> e_mask;
    std::vector<std::vector<unsigned char>> e_val;  // per property, one scalar per entry

    // neighbour tables (findNeighbours) for NNStar (0) and NNFull (1)
    std::vector<int64_t> nn[2];
    bool nn_valid[2] = {false, false};
    int ghost_type = SG_NN_STAR;  // setNNType; initialize() uses NNStar (SparseGridGpu.hpp:922)

    // boxes queued by remove(Box) (rem_sects, SparseGridGpu.hpp:3528-3531), [k][dim]
    std::vector<int32_t> rem_lo, rem_hi;
    // links to the finer (0) and coarser (1) level (link_dw_scan / link_dw, link_up_scan /
    // link_up, SparseGridGpu.hpp:2090-2262), with the padding points they start from
    struct Links {
        std::vector<int32_t> pc, lc;
        std::vector<int16_t> po, lo;
        std::vector<uint32_t> scan{0};
    };
    Links links[2];
    // copies queued by copy_to on this grid as the source (copySect, SparseGridGpu.hpp:3550-3559)
    struct Section {
        void *dst;
        int32_t lo[3], hi[3], dst_lo[3];
    };
    std::vector<Section> copies;
    std::vector<int> ghost_lin, ghost_nb;

    int64_t n() const { return (int64_t)keys.size(); }
    int nnn(int type) const { return type == SG_NN_STAR ? (1 << dim) : (dim == 2 ? 9 : 27); }
    int eb_edge() const { return edge + 2 * r; }
    int eb_size() const { int s = 1; for (int d = 0; d < dim; ++d) s *= eb_edge(); return s; }

    unsigned char *val(int p, int64_t chunk, int off) { return &data[p][(chunk * csz + off) * ts[p]]; }

    // chunk linearizer: false = grid_smb (row-major), true = grid_zmb (Morton)
    bool zorder = false;

    // ---- grid_zmb (src/Grid/Geometry/grid_zmb.hpp:65-106; util/zmorton.hpp) -----
    // lin_zid: bit i of chunk coordinate d goes to bit dim*i + d (64/dim bits per axis)
    static uint64_t morton(const int64_t *bc, int dim)
    {
        uint64_t r = 0;
        for (int d = 0; d < dim; ++d)
            for (int i = 0; i < 64 / dim; ++i) r |= (((uint64_t)bc[d] >> i) & 1ull) << (dim * i + d);
        return r;
    }
    // invlin_zid
    static void inv_morton(uint64_t lin, int64_t *bc, int dim)
    {
        for (int d = 0; d < dim; ++d) {
            uint64_t v = 0;
            for (int i = 0; i < 64 / dim; ++i) v |= ((lin >> (dim * i + d)) & 1ull) << i;
            bc[d] = (int64_t)v;
        }
    }

    // ---- grid_smb (src/Grid/Geometry/grid_smb.hpp) ----------------------------
    // LinId, grid_smb.hpp:175-189 (grid_zmb.hpp:65-78: Morton chunk id, same local id)
    int64_t lin_id(const int64_t *c) const
    {
        if (zorder) {
            int64_t bc[3] = {0, 0, 0}, l = 0;
            for (int d = dim - 1; d >= 0; --d) {
                bc[d] = c[d] / edge;
                l = l * edge + c[d] % edge;
            }
            return (int64_t)morton(bc, dim) * csz + l;
        }
        int64_t b = c[dim - 1] / edge, l = c[dim - 1] % edge;
        for (int d = dim - 2; d >= 0; --d) {
            b = b * bsz[d] + c[d] / edge;
            l = l * edge + c[d] % edge;
        }
        return b * csz + l;
    }
    // InvLinId, grid_smb.hpp:240-255 (grid_zmb.hpp:80-97)
    void inv_lin(int64_t lin, int64_t *c) const
    {
        int64_t b = lin / csz, l = lin % csz;
        int64_t bc[3] = {0, 0, 0};
        block_inv(b, bc);
        for (int d = 0; d < dim; ++d) {
            c[d] = bc[d] * edge + l % edge;
            l /= edge;
        }
    }
    // BlockLinId, grid_smb.hpp:258-276: -1 for a coordinate >= blockSz, no check
    // for negative coordinates.  grid_zmb::BlockLinId (grid_zmb.hpp:99-103) is
    // lin_zid with no check at all.
    int64_t block_lin_id(const int64_t *bc) const
    {
        if (zorder) return (int64_t)morton(bc, dim);
        int64_t b = bc[dim - 1];
        if (b >= bsz[dim - 1]) return -1;
        for (int d = dim - 2; d >= 0; --d) {
            b *= bsz[d];
            if (bc[d] >= bsz[d]) return -1;
            b += bc[d];
        }
        return b;
    }
    // BlockInvLinId, grid_smb.hpp:313-323 (grid_zmb.hpp:105-111)
    void block_inv(int64_t b, int64_t *bc) const
    {
        if (zorder) {
            inv_morton((uint64_t)b, bc, dim);
            return;
        }
        for (int d = 0; d < dim; ++d) {
            bc[d] = b % bsz[d];
            b /= bsz[d];
        }
    }
    // vector_sparse get_sparse / _branchfree_search: position of key, or the
    // background position n() when absent (map_vector_sparse_cuda_ker.cuh:75-92)
    int64_t find(int64_t key) const
    {
        auto it = std::lower_bound(keys.begin(), keys.end(), key);
        if (it != keys.end() && *it == key) return it - keys.begin();
        return n();
    }

    void reset_background_chunk()
    {
        // resetBck: background chunk appended after the last chunk
        // (map_vector_sparse.hpp:1356-1364; BlockMapGpu.hpp:306-326 sets mask 0)
        int64_t nb = n();
        for (int p = 0; p < np; ++p) {
            data[p].resize((size_t)(nb + 1) * csz * ts[p]);
            for (int o = 0; o < csz; ++o) memcpy(val(p, nb, o), bg[p].data(), ts[p]);
        }
        mask.resize((size_t)(nb + 1) * csz);
        std::fill(mask.begin() + nb * csz, mask.end(), 0);
    }

    // ---- ghost layer mapping (SparseGridGpu.hpp:871-933) -----------------------
    // getNNskin of NNStar (SparseGridGpu_ker_util.hpp:114-139): face cells map to
    // their face neighbour, edge/corner cells to neighbour 0; NNFull
    // (SparseGridGpu.hpp:289-316): sum_i (c<0?0 : c>=edge?2 : 1) * 3^i.
    void compute_ghost_mapping(int type)
    {
        ghost_lin.clear();
        ghost_nb.clear();
        int e = eb_edge();
        int ebs = eb_size();
        for (int lin = 0; lin < ebs; ++lin) {
            int c[3] = {0, 0, 0}, t = lin;
            bool inner = true;
            for (int d = 0; d < dim; ++d) {
                c[d] = t % e;
                t /= e;
                if (c[d] < r || c[d] >= r + edge) inner = false;
            }
            if (inner) continue;
            int nb;
            if (type == SG_NN_STAR) {
                nb = -1;
                int ctr = 0;
                for (int j = 0; j < dim; ++j) {
                    int cc = c[j] - r;
                    if (cc < 0) { nb = 2 * j; ++ctr; }
                    else if (cc >= edge) { nb = 2 * j + 1; ++ctr; }
                }
                if (ctr > 1) nb = 0;
            } else {
                nb = 0;
                int acc = 1;
                for (int i = 0; i < dim; ++i) {
                    int cc = c[i] - r;
                    nb += (cc < 0 ? 0 : (cc >= edge ? 2 : 1)) * acc;
                    acc *= 3;
                }
            }
            ghost_lin.push_back(lin);
            ghost_nb.push_back(nb);
        }
        ghost_type = type;
    }

    // Offset inside the neighbour chunk of a ghost cell (loadGhostBlock_impl<1>,
    // SparseGridGpu_ker_util.hpp:515-528): every coordinate wrapped into [0, edge).
    int ghost_offset(int lin) const
    {
        int e = eb_edge(), ctr = lin, acc = 1, off = 0;
        for (int i = 0; i < dim; ++i) {
            int v = (ctr % e) - r;
            v = (v < 0) ? v + edge : v;
            v = (v >= edge) ? v - edge : v;
            off += v * acc;
            ctr /= e;
            acc *= edge;
        }
        return off;
    }

    // ---- findNeighbours (SparseGridGpu_kernels.cuh:1009-1035) ------------------
    void find_neighbours(int type)
    {
        int nNN = nnn(type);
        int64_t nc = n();
        nn[type].assign((size_t)nc * nNN, -1);
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nc; ++c) {
            int64_t bc0[3];
            block_inv(keys[c], bc0);
            for (int k = 0; k < nNN; ++k) {
                int64_t bc[3] = {bc0[0], bc0[1], bc0[2]};
                int64_t pos;
                if (type == SG_NN_STAR) {
                    // NNStar::getNNpos (SparseGridGpu_ker_util.hpp:91-112)
                    pos = -1;
                    if (k < 2 * dim) {
                        int d = k / 2;
                        int64_t dPos = bc[d] + (k % 2) * 2 - 1;
                        bc[d] = dPos;
                        int64_t bl = block_lin_id(bc);
                        bl = (dPos < 0) ? -1 : bl;
                        pos = find(bl);
                    }
                } else {
                    // NNFull::getNNpos (SparseGridGpu.hpp:266-287): the centre
                    // and out-of-range chunks map to the background; a -1 block
                    // coordinate is not guarded, so at the low x / y faces the
                    // linear id aliases another chunk (grid_smb.hpp:258-276).
                    pos = nc;
                    if (k != nNN / 2) {
                        int cnt = k;
                        for (int i = 0; i < dim; ++i) {
                            bc[i] += cnt % 3 - 1;
                            cnt /= 3;
                        }
                        pos = find(block_lin_id(bc));
                    }
                }
                nn[type][(size_t)c * nNN + k] = pos;
            }
        }
        nn_valid[type] = true;
    }

    // loadGhostBlock<p> into an enlarged block (SparseGridGpu_ker_util.hpp:487-564):
    // enlarged cells first hold the background (stencil_cross_func_conv,
    // SparseGridGpu_kernels.cuh:262-268), ghost cells then take the mapped
    // neighbour's value and inner cells the chunk's own value; any value whose own
    // mask byte is 0 is replaced by the background, except when loading the mask.
    // nn_type selects the neighbour table the kernel indexes (toKernelNN<nNN>).
    template <typename T>
    void load_enlarged(T *eb, int p, int64_t c, int nn_type, bool is_mask) const
    {
        int ebs = eb_size();
        T bgv;
        if (is_mask) bgv = 0;
        else memcpy(&bgv, bg[p].data(), sizeof(T));
        for (int i = 0; i < ebs; ++i) eb[i] = bgv;
        int nNN = nnn(nn_type);
        const int64_t *nnrow = &nn[nn_type][(size_t)c * nNN];
        for (size_t g = 0; g < ghost_lin.size(); ++g) {
            int nbn = ghost_nb[g];
            int64_t npos = (nbn >= 0 && nbn < nNN) ? nnrow[nbn] : n();
            if (npos < 0) npos = n();
            int off = ghost_offset(ghost_lin[g]);
            uint8_t gm = mask[npos * csz + off];
            T v;
            if (is_mask) v = (T)gm;
            else {
                memcpy(&v, &data[p][(npos * csz + off) * sizeof(T)], sizeof(T));
                if (gm == 0) v = bgv;
            }
            eb[ghost_lin[g]] = v;
        }
        int e = eb_edge();
        for (int o = 0; o < csz; ++o) {
            int t = o, lin = 0, acc = 1;
            for (int d = 0; d < dim; ++d) {
                lin += (t % edge + r) * acc;
                t /= edge;
                acc *= e;
            }
            uint8_t bm = mask[c * csz + o];
            T v;
            if (is_mask) v = (T)bm;
            else {
                memcpy(&v, &data[p][(c * csz + o) * sizeof(T)], sizeof(T));
                if (bm == 0) v = bgv;
            }
            eb[lin] = v;
        }
    }

    // ghost mapping used by a kernel whose neighbour table is `nn_type`:
    // the current setNNType mapping, except that a star table cannot be indexed by
    // the full mapping (out of bounds in the reference), where the star one is used.
    int ghost_for(int nn_type) const
    {
        if (ghost_type == SG_NN_FULL && nn_type == SG_NN_STAR) return SG_NN_STAR;
        return ghost_type;
    }
};

void clear_inserts(Grid &g)
{
    g.s_key.clear();
    g.s_beg.assign(1, 0);
    g.e_off.clear();
    g.e_mask.clear();
    for (auto &v : g.e_val) v.clear();
}

// ---- flush (map_vector_sparse.hpp:1192-1226 -> BlockMapGpu_kernels.cuh:700-807) ----
template <typename T>
void segreduce_prop(Grid &g, int p, int op, const std::vector<int64_t> &data_map,
                    const std::vector<int64_t> &seg, const std::vector<int64_t> &seg_old,
                    const std::vector<int64_t> &out_map, std::vector<unsigned char> &out,
                    std::vector<uint8_t> *out_mask)
{
    const int csz = g.csz;
    int64_t nseg = (int64_t)seg.size() - 1;
    T bgv;
    memcpy(&bgv, g.bg[p].data(), sizeof(T));
#pragma omp parallel
    {
        std::vector<T> A(csz);
        std::vector<uint8_t> AM(csz);
#pragma omp for schedule(dynamic, 16)
        for (int64_t j = 0; j < nseg; ++j) {
            // Phase 0: first slot of the segment is the first operand
            // (segreduce_total_with_mask, BlockMapGpu_kernels.cuh:361-367).  An
            // offset the slot does not carry holds an unspecified value with mask 0;
            // the background stands in for it here.
            for (int o = 0; o < csz; ++o) { A[o] = bgv; AM[o] = 0; }
            for (int64_t i = seg[j]; i < seg[j + 1]; ++i) {
                int64_t s = data_map[i];
                for (int64_t e = g.s_beg[s]; e < g.s_beg[s + 1]; ++e) {
                    int o = g.e_off[e];
                    uint8_t bm = g.e_mask[e];
                    T b;
                    memcpy(&b, &g.e_val[p][e * sizeof(T)], sizeof(T));
                    if (i == seg[j]) {
                        A[o] = b;
                        AM[o] = bm;
                        continue;
                    }
                    // applyOp (BlockMapGpu_dimensionalityWrappers.cuh:214-226):
                    // a = op(a,b) if both exist, a = b if only b exists.
                    bool aE = AM[o] & 1, bE = bm & 1;
                    if (op == SG_OP_NONE) {
                        if (bE) A[o] = b;
                    } else if (aE && bE) {
                        A[o] = reduce_op<T>(op, A[o], b);
                    } else if (bE) {
                        A[o] = b;
                    }
                    AM[o] = AM[o] | bm;
                }
            }
            // fold the old chunk when the key already existed (:445-455)
            int64_t so = seg_old[j];
            if (so != -1) {
                for (int o = 0; o < csz; ++o) {
                    uint8_t bm = g.mask[so * csz + o];
                    T b;
                    memcpy(&b, &g.data[p][(so * csz + o) * sizeof(T)], sizeof(T));
                    bool aE = AM[o] & 1, bE = bm & 1;
                    if (op == SG_OP_NONE) {
                        if (!aE && bE) A[o] = b;
                    } else if (aE && bE) {
                        A[o] = reduce_op<T>(op, A[o], b);
                    } else if (bE) {
                        A[o] = b;
                    }
                    AM[o] = AM[o] | bm;
                }
            }
            int64_t oid = out_map[j];
            memcpy(&out[(size_t)oid * csz * sizeof(T)], A.data(), csz * sizeof(T));
            if (out_mask) memcpy(&(*out_mask)[(size_t)oid * csz], AM.data(), csz);
        }
    }
}

int flush(Grid &g, const int32_t *ops)
{
    const int64_t ns = (int64_t)g.s_key.size();
    // nothing inserted: return without changes (map_vector_sparse.hpp:1199-1213)
    if (ns == 0) return SG_OK;
    const int64_t m = g.n();

    // make_continuos + reorder_indexes: stable radix sort of (key, slot)
    // (map_vector_sparse.hpp:880-995; sort_ofp.cuh:262-402 CUB SortPairs is stable)
    std::vector<int64_t> order(ns);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return g.s_key[a] < g.s_key[b]; });
    std::vector<int64_t> K(ns);
    for (int64_t i = 0; i < ns; ++i) K[i] = g.s_key[order[i]];

    // merge_indexes: run starts -> unique keys with segment starts
    // (find_buffer_offsets_for_scan + scan + construct_index_unique, kernels.cuh:34-68)
    std::vector<int64_t> U, seg;
    for (int64_t i = 0; i < ns; ++i) {
        if (i == 0 || K[i - 1] != K[i]) {
            U.push_back(K[i]);
            seg.push_back(i);
        }
    }
    const int64_t nU = (int64_t)U.size();
    seg.push_back(ns);  // extendSegments: trailing segment end

    // merge_by_key of old keys (vals 0..m-1) and unique new keys (vals m+j),
    // ties keep the old key first (merge_ofp.cuh:90-99)
    std::vector<int64_t> Mk, Mv;
    Mk.reserve(m + nU);
    Mv.reserve(m + nU);
    {
        int64_t a = 0, b = 0;
        while (a < m || b < nU) {
            if (b >= nU || (a < m && !(U[b] < g.keys[a]))) {
                Mk.push_back(g.keys[a]);
                Mv.push_back(a);
                ++a;
            } else {
                Mk.push_back(U[b]);
                Mv.push_back(m + b);
                ++b;
            }
        }
    }
    const int64_t len = (int64_t)Mk.size();

    // compute_predicate (BlockMapGpu_kernels.cuh:20-55) + 5 exclusive scans
    std::vector<int> pred[5], scan[5];
    for (int q = 0; q < 5; ++q) { pred[q].assign(len, 0); scan[q].assign(len, 0); }
    for (int64_t p = 0; p < len; ++p) {
        int64_t pp1 = (p + 1 == len) ? p : p + 1;
        int64_t pm1 = (p == 0) ? p : p - 1;
        int64_t id0 = Mv[pm1], id1 = Mv[p];
        int64_t k0 = Mk[pm1], k1 = Mk[p], k2 = Mk[pp1];
        pred[0][p] = ((k0 == k1) && (p != 0) && id0 < m) || (id1 < m && (k1 != k2));
        pred[1][p] = id1 >= m;
        pred[2][p] = (k1 != k2) | (p == len - 1);
        pred[3][p] = id1 < m && ((k1 != k2) | (p == len - 1));
        pred[4][p] = id1 < m;
    }
    int tot[5];
    for (int q = 0; q < 5; ++q) {
        int acc = 0;
        for (int64_t p = 0; p < len; ++p) { scan[q][p] = acc; acc += pred[q][p]; }
        tot[q] = acc;
    }
    // maps_create (BlockMapGpu_kernels.cuh:58-90)
    std::vector<int64_t> seg_old(tot[1]), out_map(tot[1]), cp_dst(tot[3]), cp_src(tot[3]);
    for (int64_t p = 0; p < len; ++p) {
        if (pred[1][p]) {
            seg_old[scan[1][p]] = pred[0][p] ? scan[0][p] : -1;
            out_map[scan[1][p]] = scan[2][p];
        }
        if (pred[3][p]) {
            cp_dst[scan[3][p]] = scan[2][p];
            cp_src[scan[3][p]] = scan[4][p];
        }
    }
    // output keys: copyKeyToDstIndexIfPredicate (BlockMapGpu_kernels.cuh:516-534)
    const int64_t ndata = tot[2];
    std::vector<int64_t> keysOut(ndata);
    for (int64_t p = 0; p < len; ++p) {
        int64_t pm1 = (p == 0) ? 0 : p - 1;
        if (scan[2][p] != scan[2][pm1] || p == 0) keysOut[scan[2][p]] = Mk[p];
    }

    // segreduce per property, mask written by the first reduction (:643-664)
    std::vector<std::vector<unsigned char>> out(g.np);
    std::vector<uint8_t> out_mask((size_t)(ndata + 1) * g.csz, 0);
    for (int p = 0; p < g.np; ++p) {
        out[p].assign((size_t)(ndata + 1) * g.csz * g.ts[p], 0);
        std::vector<uint8_t> *om = (p == 0) ? &out_mask : nullptr;
        dispatch(g.dt[p], [&](auto tag) {
            typedef decltype(tag) T;
            segreduce_prop<T>(g, p, ops[p], order, seg, seg_old, out_map, out[p], om);
        });
    }
    if (g.np == 0) {
        // masks still need the OR over the segment (kept for a mask-only grid)
        for (int64_t j = 0; j < nU; ++j) {
            uint8_t *am = &out_mask[out_map[j] * g.csz];
            for (int64_t i = seg[j]; i < seg[j + 1]; ++i) {
                int64_t s = order[i];
                for (int64_t e = g.s_beg[s]; e < g.s_beg[s + 1]; ++e) am[g.e_off[e]] |= g.e_mask[e];
            }
            if (seg_old[j] != -1)
                for (int o = 0; o < g.csz; ++o) am[o] |= g.mask[seg_old[j] * g.csz + o];
        }
    }
    // copy_old_ker (BlockMapGpu_kernels.cuh:489-513)
    for (size_t i = 0; i < cp_dst.size(); ++i) {
        for (int p = 0; p < g.np; ++p)
            memcpy(&out[p][(size_t)cp_dst[i] * g.csz * g.ts[p]],
                   &g.data[p][(size_t)cp_src[i] * g.csz * g.ts[p]], g.csz * g.ts[p]);
        memcpy(&out_mask[(size_t)cp_dst[i] * g.csz], &g.mask[(size_t)cp_src[i] * g.csz], g.csz);
    }
    g.keys.swap(keysOut);
    for (int p = 0; p < g.np; ++p) g.data[p].swap(out[p]);
    g.mask.swap(out_mask);
    g.reset_background_chunk();
    g.nn_valid[0] = g.nn_valid[1] = false;  // flush marks the neighbour table stale (SparseGridGpu.hpp:778)
    clear_inserts(g);
    return SG_OK;
}

// ---- tagBoundaries (SparseGridGpu_kernels.cuh:667-737) ----------------------------
int tag_boundaries(Grid &g, int nn_type)
{
    if (!g.nn_valid[nn_type]) g.find_neighbours(nn_type);
    const int64_t nc = g.n();
    const int ebs = g.eb_size(), e = g.eb_edge();
    std::vector<uint8_t> newmask(g.mask);
    const int gtype = g.ghost_for(nn_type);
    Grid *gp = &g;
    // the kernel reads neighbour masks while other chunks are rewritten; only the
    // PADDING bit changes and isPadding reads only EXIST, so a snapshot is exact.
    std::vector<int> save_lin = g.ghost_lin, save_nb = g.ghost_nb;
    int save_type = g.ghost_type;
    if (gtype != g.ghost_type) g.compute_ghost_mapping(gtype);
#pragma omp parallel
    {
        std::vector<uint8_t> eb(ebs);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < nc; ++c) {
            gp->load_enlarged<uint8_t>(eb.data(), -1, c, nn_type, true);
            for (int o = 0; o < g.csz; ++o) {
                int t = o, lin = 0, acc = 1, crd[3];
                for (int d = 0; d < g.dim; ++d) {
                    crd[d] = t % g.edge + g.r;
                    lin += crd[d] * acc;
                    t /= g.edge;
                    acc *= e;
                }
                uint8_t cur = eb[lin];
                if (!(cur & 1)) continue;
                bool pad = false;
                if (nn_type == SG_NN_STAR) {
                    // NNStar::isPadding (SparseGridGpu_ker_util.hpp:141-157)
                    int stride = 1;
                    for (int d = 0; d < g.dim; ++d) {
                        pad = pad || !(eb[lin + stride] & 1) || !(eb[lin - stride] & 1);
                        stride *= e;
                    }
                } else {
                    // NNfull_is_padding_impl (SparseGridGpu.hpp:200-259)
                    int nb = 1;
                    for (int d = 0; d < g.dim; ++d) nb *= 3;
                    for (int k = 0; k < nb; ++k) {
                        int kk = k, off = 0, s = 1;
                        for (int d = 0; d < g.dim; ++d) {
                            off += (kk % 3 - 1) * s;
                            kk /= 3;
                            s *= e;
                        }
                        pad = pad || !(eb[lin + off] & 1);
                    }
                }
                // setPadding / unsetPadding (SparseGridGpu_ker.cuh:684-693)
                newmask[c * g.csz + o] = pad ? (uint8_t)(cur | 2) : (uint8_t)(cur & ~2);
            }
        }
    }
    g.mask.swap(newmask);
    if (save_type != g.ghost_type) {
        g.ghost_lin = save_lin;
        g.ghost_nb = save_nb;
        g.ghost_type = save_type;
    }
    return SG_OK;
}

// ---- stencils: applyStencilInPlace + conv functors ---------------------------------
template <typename T>
int apply_stencil_t(Grid &g, const sg_stencil_desc *ds)
{
    const int kind = ds->kind;
    const int nn_type = (kind == SG_ST_BOX) ? SG_NN_FULL : SG_NN_STAR;
    if (!g.nn_valid[nn_type]) g.find_neighbours(nn_type);
    const int nsrc = (kind == SG_ST_GRAY_SCOTT || kind == SG_ST_RBSOR || kind == SG_ST_RESIDUAL) ? 2 : 1;
    const int ndst = (kind == SG_ST_GRAY_SCOTT) ? 2 : (kind == SG_ST_RBSOR ? 0 : 1);
    const int64_t nc = g.n();
    const int ebs = g.eb_size(), e = g.eb_edge(), csz = g.csz, dim = g.dim;
    const int gtype = g.ghost_for(nn_type);
    std::vector<int> save_lin = g.ghost_lin, save_nb = g.ghost_nb;
    int save_type = g.ghost_type;
    if (gtype != g.ghost_type) g.compute_ghost_mapping(gtype);

    // Outputs go to fresh copies of the destination arrays, so every chunk reads
    // the pre-sweep state (Jacobi semantics also when dst == src).  The red-black
    // SOR sweep writes src[0]: a colour reads only the other colour, so reading
    // the pre-sweep state is the same as updating in place.
    const int nout = (kind == SG_ST_RBSOR) ? 1 : ndst;
    int outp[2] = {ds->dst[0], ds->dst[1]};
    if (kind == SG_ST_RBSOR) outp[0] = ds->src[0];
    std::vector<std::vector<unsigned char>> outbuf(nout);
    for (int q = 0; q < nout; ++q) outbuf[q] = g.data[outp[q]];
    Grid *gp = &g;
#pragma omp parallel
    {
        std::vector<T> eb[2];
        for (int q = 0; q < nsrc; ++q) eb[q].resize(ebs);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < nc; ++c) {
            for (int q = 0; q < nsrc; ++q) gp->load_enlarged<T>(eb[q].data(), ds->src[q], c, nn_type, false);
            int64_t base[3];
            {
                int64_t crd[3];
                g.inv_lin(g.keys[c] * csz, crd);
                for (int d = 0; d < 3; ++d) base[d] = (d < dim) ? crd[d] : 0;
            }
            for (int o = 0; o < csz; ++o) {
                int t = o, lin = 0, acc = 1;
                int64_t pc[3] = {0, 0, 0};
                for (int d = 0; d < dim; ++d) {
                    int lc = t % g.edge;
                    pc[d] = base[d] + lc;
                    lin += (lc + g.r) * acc;
                    t /= g.edge;
                    acc *= e;
                }
                // curMask zeroed outside the inclusive box (SparseGridGpu.hpp:1019-1025)
                uint8_t cm = g.mask[c * csz + o];
                if (ds->use_box)
                    for (int d = 0; d < dim; ++d)
                        if (pc[d] < ds->box_lo[d] || pc[d] > ds->box_hi[d]) cm = 0;
                bool active = (ds->apply == SG_APPLY_EXISTING) ? (cm != 0)
                                                                : ((cm & SG_EXIST) && !(cm & SG_PADDING));
                // applyStencilInPlaceNoShared runs over e_points, the EXIST points only
                // (SparseGridGpu.hpp:1051-1075, fill_e_points SparseGridGpu_kernels.cuh:1150-1185)
                if (ds->mode == SG_STENCIL_MODE_INPLACE_NO_SHARED && !(g.mask[c * g.csz + o] & SG_EXIST))
                    active = false;
                if (!active) continue;
                const T *u = eb[0].data();
                T cur = u[lin];
                if (kind == SG_ST_LAP || kind == SG_ST_HEAT || kind == SG_ST_RESIDUAL) {
                    T acc_l;
                    if (kind == SG_ST_LAP) acc_l = (T)ds->coef[0] * cur;
                    else if (kind == SG_ST_HEAT) acc_l = (T)(-2.0 * dim) * cur;
                    else acc_l = (T)(2.0 * dim) * cur;
                    int s = 1;
                    for (int d = 0; d < dim; ++d) {
                        T pair = u[lin - s] + u[lin + s];
                        if (kind == SG_ST_LAP) acc_l = acc_l + (T)ds->coef[1] * pair;
                        else if (kind == SG_ST_HEAT) acc_l = acc_l + pair;
                        else acc_l = acc_l - pair;
                        s *= e;
                    }
                    T res;
                    if (kind == SG_ST_LAP) res = acc_l;
                    else if (kind == SG_ST_HEAT) res = cur + (T)ds->coef[0] * acc_l;
                    else res = eb[1][lin] - acc_l;
                    memcpy(&outbuf[0][(c * csz + o) * sizeof(T)], &res, sizeof(T));
                } else if (kind == SG_ST_GRAY_SCOTT) {
                    const T *v = eb[1].data();
                    T cu = cur, cv = v[lin];
                    T lu = (T)(-2.0 * dim) * cu, lv = (T)(-2.0 * dim) * cv;
                    int s = 1;
                    for (int d = 0; d < dim; ++d) {
                        lu = lu + (u[lin - s] + u[lin + s]);
                        lv = lv + (v[lin - s] + v[lin + s]);
                        s *= e;
                    }
                    const T Du = (T)ds->coef[0], Dv = (T)ds->coef[1], F = (T)ds->coef[2],
                            k = (T)ds->coef[3], dt = (T)ds->coef[4];
                    T uvv = (cu * cv) * cv;
                    T nu = cu + dt * ((Du * lu - uvv) + F * ((T)1 - cu));
                    T nv = cv + dt * ((Dv * lv + uvv) - (F + k) * cv);
                    memcpy(&outbuf[0][(c * csz + o) * sizeof(T)], &nu, sizeof(T));
                    memcpy(&outbuf[1][(c * csz + o) * sizeof(T)], &nv, sizeof(T));
                } else if (kind == SG_ST_BOX) {
                    // Conv3x3x3 loop order: x offset outermost, z innermost
                    // (tests/utils/SparseGridGpu_util_test.cuh:396-411)
                    T tot = 0;
                    int nk = (dim == 3) ? 3 : 1;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            for (int k = 0; k < nk; ++k) {
                                int off = (i - 1) + (j - 1) * e + ((dim == 3) ? (k - 1) * e * e : 0);
                                int ci = (dim == 3) ? (i * 9 + j * 3 + k) : (i * 3 + j);
                                T cf = std::is_integral<T>::value ? (T)ds->icoef[ci] : (T)ds->coef[ci];
                                tot = op_add<T>(tot, op_mul<T>(u[lin + off], cf));
                            }
                    memcpy(&outbuf[0][(c * csz + o) * sizeof(T)], &tot, sizeof(T));
                } else if (kind == SG_ST_RBSOR) {
                    int64_t par = 0;
                    for (int d = 0; d < dim; ++d) par += pc[d];
                    if ((int)(par & 1) != (int)ds->icoef[0]) continue;
                    T nbs = 0;
                    int s = 1;
                    for (int d = 0; d < dim; ++d) {
                        nbs = nbs + (u[lin - s] + u[lin + s]);
                        s *= e;
                    }
                    const T w = (T)ds->coef[0];
                    T gs = (eb[1][lin] + nbs) / (T)(2 * dim);
                    T res = cur + w * (gs - cur);
                    memcpy(&outbuf[0][(c * csz + o) * sizeof(T)], &res, sizeof(T));
                }
            }
        }
    }
    for (int q = 0; q < nout; ++q) g.data[outp[q]].swap(outbuf[q]);
    if (save_type != g.ghost_type) {
        g.ghost_lin = save_lin;
        g.ghost_nb = save_nb;
        g.ghost_type = save_type;
    }
    return SG_OK;
}

int append_records(Grid &g, int64_t n, const int64_t *chunk, const int32_t *off, const uint8_t *masks,
                   const void *const *values, int64_t per_record, bool skip_zero_mask)
{
    // per_record == 1: one point per record (chunk[i], off[i]);
    // per_record == csz: record i is a whole chunk slot
    for (int64_t i = 0; i < n; ++i) {
        g.s_key.push_back(chunk[i]);
        for (int64_t k = 0; k < per_record; ++k) {
            int64_t idx = i * per_record + k;
            uint8_t mk = masks ? masks[idx] : (uint8_t)SG_EXIST;
            if (skip_zero_mask && mk == 0) continue;
            g.e_off.push_back(per_record == 1 ? off[i] : (int32_t)k);
            g.e_mask.push_back(mk);
            for (int p = 0; p < g.np; ++p) {
                size_t t = g.ts[p];
                size_t at = g.e_val[p].size();
                g.e_val[p].resize(at + t);
                if (values && values[p]) memcpy(&g.e_val[p][at], (const unsigned char *)values[p] + idx * t, t);
                else memcpy(&g.e_val[p][at], g.bg[p].data(), t);
            }
        }
        g.s_beg.push_back((int64_t)g.e_off.size());
    }
    return SG_OK;
}

}  // namespace

extern "C" {

const char *sgcpu_last_error(void) { return g_err.c_str(); }

int sgcpu_create(int dim, const int64_t *size, int block_edge, int nprops, const int32_t *dtypes,
                 int stencil_radius, void **out)
{
    if (!out || !size) return fail(SG_E_ARG, "null argument");
    if (dim < 2 || dim > 3) return fail(SG_E_ARG, "dim must be 2 or 3");
    if (block_edge < 2) return fail(SG_E_ARG, "block_edge must be >= 2");
    if (stencil_radius != 1) return fail(SG_E_ARG, "stencil_radius must be 1");
    if (nprops < 0 || nprops > 16) return fail(SG_E_ARG, "nprops must be in [0,16]");
    Grid *g = new Grid();
    g->dim = dim;
    g->edge = block_edge;
    g->csz = 1;
    for (int d = 0; d < dim; ++d) g->csz *= block_edge;
    g->r = stencil_radius;
    for (int d = 0; d < dim; ++d) {
        if (size[d] <= 0) { delete g; return fail(SG_E_ARG, "size must be positive"); }
        g->sz[d] = size[d];
        g->bsz[d] = size[d] / block_edge + ((size[d] % block_edge) != 0);  // grid_smb.hpp:47-57
    }
    g->np = nprops;
    g->data.resize(nprops);
    g->bg.resize(nprops);
    g->e_val.resize(nprops);
    for (int p = 0; p < nprops; ++p) {
        g->dt[p] = dtypes[p];
        g->ts[p] = dtype_size(dtypes[p]);
        if (!g->ts[p]) { delete g; return fail(SG_E_ARG, "bad dtype"); }
        g->bg[p].assign(g->ts[p], 0);
    }
    g->compute_ghost_mapping(SG_NN_STAR);
    g->reset_background_chunk();
    *out = g;
    return SG_OK;
}

int sgcpu_destroy(void *h)
{
    delete (Grid *)h;
    return SG_OK;
}

int sgcpu_set_background(void *h, int p, const void *value)
{
    Grid &g = *(Grid *)h;
    if (p < 0 || p >= g.np) return fail(SG_E_ARG, "bad property");
    memcpy(g.bg[p].data(), value, g.ts[p]);
    g.reset_background_chunk();
    return SG_OK;
}

int sgcpu_insert_points(void *h, int64_t n, const int64_t *keys, const uint8_t *masks, const void *const *values)
{
    Grid &g = *(Grid *)h;
    std::vector<int64_t> ch(n);
    std::vector<int32_t> off(n);
    for (int64_t i = 0; i < n; ++i) {
        if (keys[i] < 0) return fail(SG_E_RANGE, "negative key");
        ch[i] = keys[i] / g.csz;
        off[i] = (int32_t)(keys[i] % g.csz);
    }
    return append_records(g, n, ch.data(), off.data(), masks, values, 1, false);
}

int sgcpu_insert_coords(void *h, int64_t n, const int32_t *coords, const uint8_t *masks, const void *const *values)
{
    Grid &g = *(Grid *)h;
    std::vector<int64_t> keys(n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        for (int d = 0; d < g.dim; ++d) {
            c[d] = coords[i * g.dim + d];
            if (c[d] < 0 || c[d] >= g.sz[d]) return fail(SG_E_RANGE, "coordinate outside the grid");
        }
        keys[i] = g.lin_id(c);
    }
    return sgcpu_insert_points(h, n, keys.data(), masks, values);
}

int sgcpu_insert_chunks(void *h, int64_t n, const int64_t *chunk_ids, const uint8_t *masks, const void *const *values)
{
    Grid &g = *(Grid *)h;
    return append_records(g, n, chunk_ids, nullptr, masks, values, g.csz, true);
}

int64_t sgcpu_pending_inserts(void *h) { return (int64_t)((Grid *)h)->s_key.size(); }

int sgcpu_flush(void *h, const int32_t *ops)
{
    Grid &g = *(Grid *)h;
    for (int p = 0; p < g.np; ++p)
        if (ops[p] == SG_OP_OR && dtype_is_float(g.dt[p])) return fail(SG_E_ARG, "bitwise or on a float property");
    return flush(g, ops);
}

int sgcpu_find_neighbours(void *h, int type)
{
    Grid &g = *(Grid *)h;
    if (type != SG_NN_STAR && type != SG_NN_FULL) return fail(SG_E_ARG, "bad nn type");
    g.find_neighbours(type);
    return SG_OK;
}

int sgcpu_set_nn_type(void *h, int type)
{
    Grid &g = *(Grid *)h;
    if (type != SG_NN_STAR && type != SG_NN_FULL) return fail(SG_E_ARG, "bad nn type");
    g.compute_ghost_mapping(type);
    return SG_OK;
}

int sgcpu_tag_boundaries(void *h, int type, int opt)
{
    (void)opt;
    Grid &g = *(Grid *)h;
    if (type != SG_NN_STAR && type != SG_NN_FULL) return fail(SG_E_ARG, "bad nn type");
    return tag_boundaries(g, type);
}

int sgcpu_apply_stencil(void *h, const sg_stencil_desc *ds)
{
    Grid &g = *(Grid *)h;
    int nsrc = (ds->kind == SG_ST_GRAY_SCOTT || ds->kind == SG_ST_RBSOR || ds->kind == SG_ST_RESIDUAL) ? 2 : 1;
    int ndst = (ds->kind == SG_ST_GRAY_SCOTT) ? 2 : (ds->kind == SG_ST_RBSOR ? 0 : 1);
    if (ds->kind < SG_ST_LAP || ds->kind > SG_ST_RESIDUAL) return fail(SG_E_ARG, "bad stencil kind");
    for (int q = 0; q < nsrc; ++q)
        if (ds->src[q] < 0 || ds->src[q] >= g.np) return fail(SG_E_ARG, "bad src property");
    for (int q = 0; q < ndst; ++q)
        if (ds->dst[q] < 0 || ds->dst[q] >= g.np) return fail(SG_E_ARG, "bad dst property");
    int dt = g.dt[ds->src[0]];
    for (int q = 0; q < nsrc; ++q) if (g.dt[ds->src[q]] != dt) return fail(SG_E_ARG, "mixed property types");
    for (int q = 0; q < ndst; ++q) if (g.dt[ds->dst[q]] != dt) return fail(SG_E_ARG, "mixed property types");
    if (ds->kind != SG_ST_BOX && !dtype_is_float(dt)) return fail(SG_E_ARG, "stencil needs a float property");
    int rc = SG_OK;
    dispatch(dt, [&](auto tag) {
        typedef decltype(tag) T;
        rc = apply_stencil_t<T>(g, ds);
    });
    return rc;
}

int64_t sgcpu_num_chunks(void *h) { return ((Grid *)h)->n(); }
int sgcpu_chunk_size(void *h) { return ((Grid *)h)->csz; }

int sgcpu_get_chunk_ids(void *h, int64_t *out)
{
    Grid &g = *(Grid *)h;
    if (g.n()) memcpy(out, g.keys.data(), g.n() * sizeof(int64_t));
    return SG_OK;
}

int sgcpu_get_masks(void *h, uint8_t *out)
{
    Grid &g = *(Grid *)h;
    if (g.n()) memcpy(out, g.mask.data(), g.n() * g.csz);
    return SG_OK;
}

int sgcpu_get_prop(void *h, int p, void *out)
{
    Grid &g = *(Grid *)h;
    if (p < 0 || p >= g.np) return fail(SG_E_ARG, "bad property");
    if (g.n()) memcpy(out, g.data[p].data(), g.n() * g.csz * g.ts[p]);
    return SG_OK;
}

int sgcpu_set_prop(void *h, int p, const void *in)
{
    Grid &g = *(Grid *)h;
    if (p < 0 || p >= g.np) return fail(SG_E_ARG, "bad property");
    if (g.n()) memcpy(g.data[p].data(), in, g.n() * g.csz * g.ts[p]);
    return SG_OK;
}

int sgcpu_get_points(void *h, int p, int64_t n, const int64_t *keys, void *out)
{
    // BlockMapGpu::get (BlockMapGpu.hpp:108-126): background if the exist bit is 0
    Grid &g = *(Grid *)h;
    if (p < 0 || p >= g.np) return fail(SG_E_ARG, "bad property");
    for (int64_t i = 0; i < n; ++i) {
        int64_t pos = g.find(keys[i] / g.csz);
        int off = (int)(keys[i] % g.csz);
        const unsigned char *src = (g.mask[pos * g.csz + off] & 1) ? g.val(p, pos, off) : g.bg[p].data();
        memcpy((unsigned char *)out + i * g.ts[p], src, g.ts[p]);
    }
    return SG_OK;
}

int sgcpu_get_neighbours(void *h, int type, int32_t *out)
{
    Grid &g = *(Grid *)h;
    if (!g.nn_valid[type]) g.find_neighbours(type);
    for (size_t i = 0; i < g.nn[type].size(); ++i) out[i] = (int32_t)g.nn[type][i];
    return SG_OK;
}

int sgcpu_count_existing(void *h, int64_t *out)
{
    Grid &g = *(Grid *)h;
    int64_t c = 0;
    for (int64_t i = 0; i < g.n() * g.csz; ++i) c += g.mask[i] & 1;
    *out = c;
    return SG_OK;
}

int sgcpu_count_boundary(void *h, int64_t *out)
{
    Grid &g = *(Grid *)h;
    int64_t c = 0;
    for (int64_t i = 0; i < g.n() * g.csz; ++i) c += (g.mask[i] & 3) == 3;
    *out = c;
    return SG_OK;
}

// ---- pack / unpack / remove ------------------------------------------------------
// Packed form of one box (packCalculate + pack + pack_data, SparseGridGpu.hpp:3072-3289,
// SparseGridGpu_kernels.cuh:1634-1697), sections in this order, A = index_bytes:
//   size_t n_cnk | int box_start[dim] | int grid_size[dim] | indexT ids[n_cnk]
//   | uint32 scan[n_cnk + 1], padded to A | per packed property T[n_pnt] (consecutive),
//   padded to A | int16 offsets[n_pnt], padded to A | uint8 masks[n_pnt], padded to A
// ids: chunks (sorted) with at least one existing point inside the box; scan: first
// point of each chunk, scan[n_cnk] = n_pnt; points of a chunk in increasing offset (the
// reference orders them by an atomic counter, i.e. in no fixed order).
static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

static bool inside(const Grid &g, const int64_t *c, const int32_t *lo, const int32_t *hi)
{
    for (int d = 0; d < g.dim; ++d)
        if (c[d] < lo[d] || c[d] > hi[d]) return false;
    return true;
}

// points of box k (existing, inside, inclusive box; Box::isInside) per chunk in order
static void box_points(const Grid &g, const int32_t *lo, const int32_t *hi, std::vector<int64_t> &chunks,
                       std::vector<uint32_t> &scan, std::vector<int32_t> &offs)
{
    chunks.clear();
    scan.clear();
    offs.clear();
    for (int64_t c = 0; c < g.n(); ++c) {
        const size_t before = offs.size();
        for (int o = 0; o < g.csz; ++o) {
            if (!(g.mask[c * g.csz + o] & SG_EXIST)) continue;
            int64_t pc[3];
            g.inv_lin(g.keys[c] * g.csz + o, pc);
            if (inside(g, pc, lo, hi)) offs.push_back(o);
        }
        if (offs.size() > before) {
            chunks.push_back(c);
            scan.push_back((uint32_t)before);
        }
    }
    scan.push_back((uint32_t)offs.size());
}

static int64_t section_bytes(const Grid &g, int64_t n_cnk, int64_t n_pnt, int nprop, const int32_t *props, int A)
{
    int64_t psz = 0;
    for (int q = 0; q < nprop; ++q) psz += (int64_t)g.ts[props[q]];
    return 8 + 2 * g.dim * 4 + A * n_cnk + align_up((n_cnk + 1) * 4, A) + align_up(n_pnt * psz, A) +
           align_up(n_pnt * 2, A) + align_up(n_pnt, A);
}

static int check_pack_args(const Grid &g, int nbox, int nprop, const int32_t *props, int A)
{
    if (nbox < 0 || nprop < 0) return fail(SG_E_ARG, "negative count");
    if (A != 4 && A != 8) return fail(SG_E_ARG, "index_bytes must be 4 or 8");
    for (int q = 0; q < nprop; ++q)
        if (props[q] < 0 || props[q] >= g.np) return fail(SG_E_ARG, "bad property");
    if (g.csz > 65536) return fail(SG_E_ARG, "chunk too large for 16-bit offsets");
    return SG_OK;
}

int sgcpu_pack_size(void *h, int nbox, const int32_t *lo, const int32_t *hi, int nprop, const int32_t *props,
                    int index_bytes, int64_t *bytes)
{
    Grid &g = *(Grid *)h;
    int rc = check_pack_args(g, nbox, nprop, props, index_bytes);
    if (rc) return rc;
    std::vector<int64_t> ch;
    std::vector<uint32_t> sc;
    std::vector<int32_t> of;
    int64_t tot = 0;
    for (int k = 0; k < nbox; ++k) {
        box_points(g, lo + k * g.dim, hi + k * g.dim, ch, sc, of);
        tot += section_bytes(g, (int64_t)ch.size(), (int64_t)of.size(), nprop, props, index_bytes);
    }
    *bytes = tot;
    return SG_OK;
}

int sgcpu_pack(void *h, int nbox, const int32_t *lo, const int32_t *hi, int nprop, const int32_t *props,
               int index_bytes, void *buf, int64_t cap)
{
    Grid &g = *(Grid *)h;
    int rc = check_pack_args(g, nbox, nprop, props, index_bytes);
    if (rc) return rc;
    const int A = index_bytes;
    unsigned char *out = (unsigned char *)buf;
    int64_t at = 0;
    std::vector<int64_t> ch;
    std::vector<uint32_t> sc;
    std::vector<int32_t> of;
    for (int k = 0; k < nbox; ++k) {
        const int32_t *bl = lo + k * g.dim, *bh = hi + k * g.dim;
        box_points(g, bl, bh, ch, sc, of);
        const int64_t nc = (int64_t)ch.size(), npt = (int64_t)of.size();
        const int64_t sb = section_bytes(g, nc, npt, nprop, props, A);
        if (at + sb > cap) return fail(SG_E_RANGE, "pack buffer too small");
        unsigned char *s0 = out + at;
        memset(s0, 0, (size_t)sb);
        uint64_t ncu = (uint64_t)nc;
        memcpy(s0, &ncu, 8);
        int64_t q = 8;
        for (int d = 0; d < g.dim; ++d, q += 4) memcpy(s0 + q, &bl[d], 4);
        for (int d = 0; d < g.dim; ++d, q += 4) {
            int32_t v = (int32_t)g.sz[d];
            memcpy(s0 + q, &v, 4);
        }
        for (int64_t i = 0; i < nc; ++i, q += A) {
            int64_t id = g.keys[ch[i]];
            if (A == 8) memcpy(s0 + q, &id, 8);
            else {
                int32_t id4 = (int32_t)id;
                memcpy(s0 + q, &id4, 4);
            }
        }
        memcpy(s0 + q, sc.data(), (size_t)(nc + 1) * 4);
        q += align_up((nc + 1) * 4, A);
        // point j of the box: chunk position pos_of[j], offset of[j]
        std::vector<int64_t> pos_of(npt);
        for (int64_t i = 0; i < nc; ++i)
            for (uint32_t j = sc[i]; j < sc[i + 1]; ++j) pos_of[j] = ch[i];
        int64_t dq = q;
        for (int pi = 0; pi < nprop; ++pi) {
            const int p = props[pi];
            for (int64_t j = 0; j < npt; ++j)
                memcpy(s0 + dq + j * g.ts[p], g.val(p, pos_of[j], of[j]), g.ts[p]);
            dq += npt * (int64_t)g.ts[p];
        }
        q = dq;
        int64_t psz = 0;
        for (int pi = 0; pi < nprop; ++pi) psz += (int64_t)g.ts[props[pi]];
        q = (q - npt * psz) + align_up(npt * psz, A);
        for (int64_t j = 0; j < npt; ++j) {
            int16_t o16 = (int16_t)of[j];
            memcpy(s0 + q + j * 2, &o16, 2);
        }
        q += align_up(npt * 2, A);
        for (int64_t j = 0; j < npt; ++j) s0[q + j] = g.mask[pos_of[j] * g.csz + of[j]];
        at += sb;
    }
    return SG_OK;
}

// unpack (SparseGridGpu.hpp:3665-3733 -> addAndConvertPackedChunkToTmp :1415-1560 ->
// copy_packed_data_to_chunks, SparseGridGpu_kernels.cuh:1591-1632): the point at packed
// coordinate p lands at p - box_start + dst_lo; it is queued as an insert that the
// finalize merges (unpacked properties take the packed values, mask |= EXIST).
int sgcpu_unpack(void *h, const void *buf, int64_t avail, const int32_t *dst_lo, int nprop, const int32_t *props,
                 int index_bytes, int64_t *consumed)
{
    Grid &g = *(Grid *)h;
    int rc = check_pack_args(g, 1, nprop, props, index_bytes);
    if (rc) return rc;
    const int A = index_bytes;
    const unsigned char *s0 = (const unsigned char *)buf;
    if (avail < 8 + 2 * g.dim * 4) return fail(SG_E_RANGE, "packed section truncated");
    uint64_t ncu;
    memcpy(&ncu, s0, 8);
    const int64_t nc = (int64_t)ncu;
    int32_t bstart[3], psize[3];
    int64_t q = 8;
    for (int d = 0; d < g.dim; ++d, q += 4) memcpy(&bstart[d], s0 + q, 4);
    for (int d = 0; d < g.dim; ++d, q += 4) memcpy(&psize[d], s0 + q, 4);
    if (avail < q + A * nc + (nc + 1) * 4) return fail(SG_E_RANGE, "packed section truncated");
    const int64_t ids_at = q;
    q += A * nc;
    std::vector<uint32_t> sc(nc + 1);
    memcpy(sc.data(), s0 + q, (size_t)(nc + 1) * 4);
    const int64_t npt = sc[nc];
    q += align_up((nc + 1) * 4, A);
    if (avail < section_bytes(g, nc, npt, nprop, props, A)) return fail(SG_E_RANGE, "packed section truncated");
    const int64_t data_at = q;
    int64_t psz = 0;
    for (int pi = 0; pi < nprop; ++pi) psz += (int64_t)g.ts[props[pi]];
    q += align_up(npt * psz, A);
    const int64_t off_at = q;
    // packed grid geometry: chunk counts of the packing grid (same chunk edge)
    int64_t pbsz[3];
    for (int d = 0; d < g.dim; ++d) pbsz[d] = (psize[d] + g.edge - 1) / g.edge;
    std::vector<int64_t> keys(npt);
    for (int64_t i = 0; i < nc; ++i) {
        int64_t id;
        if (A == 8) memcpy(&id, s0 + ids_at + i * 8, 8);
        else {
            int32_t id4;
            memcpy(&id4, s0 + ids_at + i * 4, 4);
            id = id4;
        }
        int64_t bc[3], b = id;
        if (g.zorder) Grid::inv_morton((uint64_t)id, bc, g.dim);
        else
            for (int d = 0; d < g.dim; ++d) {
                bc[d] = b % pbsz[d];
                b /= pbsz[d];
            }
        for (uint32_t j = sc[i]; j < sc[i + 1]; ++j) {
            int16_t o16;
            memcpy(&o16, s0 + off_at + j * 2, 2);
            int64_t l = (uint16_t)o16, c[3];
            for (int d = 0; d < g.dim; ++d) {
                c[d] = bc[d] * g.edge + l % g.edge - bstart[d] + dst_lo[d];
                l /= g.edge;
                if (c[d] < 0 || c[d] >= g.sz[d]) return fail(SG_E_RANGE, "unpacked point outside the grid");
            }
            keys[j] = g.lin_id(c);
        }
    }
    // values: packed properties from the data section, the others the background
    std::vector<std::vector<unsigned char>> vals(g.np);
    std::vector<const void *> vp(g.np, nullptr);
    int64_t dq = data_at;
    for (int pi = 0; pi < nprop; ++pi) {
        const int p = props[pi];
        vals[p].assign(s0 + dq, s0 + dq + npt * g.ts[p]);
        vp[p] = vals[p].data();
        dq += npt * (int64_t)g.ts[p];
    }
    std::vector<uint8_t> m(npt, SG_EXIST);
    *consumed = section_bytes(g, nc, npt, nprop, props, A);
    return sgcpu_insert_points(h, npt, keys.data(), m.data(), vp.data());
}

int sgcpu_remove_box(void *h, const int32_t *lo, const int32_t *hi)
{
    Grid &g = *(Grid *)h;
    for (int d = 0; d < g.dim; ++d) {
        g.rem_lo.push_back(lo[d]);
        g.rem_hi.push_back(hi[d]);
    }
    return SG_OK;
}

static void clear_point(Grid &g, int64_t c, int o)
{
    g.mask[c * g.csz + o] = 0;
    for (int p = 0; p < g.np; ++p) memcpy(g.val(p, c, o), g.bg[p].data(), g.ts[p]);
}

// device remove(coord) (BlockMapGpu_ker.cuh:322-332): mask of the point set to 0 (the
// value becomes the background, which the reference substitutes on every read)
int sgcpu_remove_points(void *h, int64_t n, const int64_t *keys)
{
    Grid &g = *(Grid *)h;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c = g.find(keys[i] / g.csz);
        if (c < g.n()) clear_point(g, c, (int)(keys[i] % g.csz));
    }
    return SG_OK;
}

// removeAddUnpackFinalize (SparseGridGpu.hpp:3492-3499): removePoints (remove_points,
// SparseGridGpu_kernels.cuh:1301-1346: existing points inside a queued box get mask 0),
// then the queued inserts are merged
static void remove_queued(Grid &g)
{
    const int nb = (int)(g.rem_lo.size() / g.dim);
    for (int64_t c = 0; c < g.n(); ++c)
        for (int o = 0; o < g.csz; ++o) {
            if (!(g.mask[c * g.csz + o] & SG_EXIST)) continue;
            int64_t pc[3];
            g.inv_lin(g.keys[c] * g.csz + o, pc);
            for (int k = 0; k < nb; ++k)
                if (inside(g, pc, &g.rem_lo[k * g.dim], &g.rem_hi[k * g.dim])) {
                    clear_point(g, c, o);
                    break;
                }
        }
    g.rem_lo.clear();
    g.rem_hi.clear();
}

static int flush_unpacked(Grid &g, int nprop, const int32_t *props)
{
    std::vector<int32_t> ops(g.np, SG_OP_RIGHT);
    for (int q = 0; q < nprop; ++q) ops[props[q]] = SG_OP_LEFT;
    return flush(g, ops.data());
}

int sgcpu_remove_add_unpack_finalize(void *h, int nprop, const int32_t *props)
{
    Grid &g = *(Grid *)h;
    remove_queued(g);
    return flush_unpacked(g, nprop, props);
}

// copy_to (SparseGridGpu.hpp:3550-3559): the section is queued on the source grid
int sgcpu_copy_to(void *dst, void *src, const int32_t *src_lo, const int32_t *src_hi, const int32_t *dst_lo)
{
    Grid &s = *(Grid *)src;
    Grid::Section c{};
    c.dst = dst;
    for (int d = 0; d < s.dim; ++d) {
        c.lo[d] = src_lo[d];
        c.hi[d] = src_hi[d];
        c.dst_lo[d] = dst_lo[d];
    }
    s.copies.push_back(c);
    return SG_OK;
}

// removeCopyToFinalize (SparseGridGpu.hpp:3314-3329).  PHASE1 = removeCopyToFinalize_phase1
// (:1116-1121, removePoints); PHASE2 = _phase2 (:1123-1174): pack every queued section of
// this grid (packRequest / packCalculate / pack / packFinalize), then preUnpack
// (:1095-1113) converts each packed box into its destination's insert buffer at the
// destination box's low corner; PHASE3 = _phase3 (:1176-1290): flush with sLeft_ on the
// copied properties, then copy_packed_data_to_chunks sets mask |= EXIST.
int sgcpu_remove_copy_to_finalize(void *h, int phase, int nprop, const int32_t *props)
{
    Grid &g = *(Grid *)h;
    switch (phase & 3) {
    case SG_COPY_PHASE1:
        remove_queued(g);
        return SG_OK;
    case SG_COPY_PHASE2: {
        const int n = (int)g.copies.size(), D = g.dim;
        if (n == 0) return SG_OK;
        std::vector<int32_t> lo((size_t)n * D), hi((size_t)n * D);
        for (int k = 0; k < n; ++k)
            for (int a = 0; a < D; ++a) {
                lo[(size_t)k * D + a] = g.copies[k].lo[a];
                hi[(size_t)k * D + a] = g.copies[k].hi[a];
            }
        int64_t tot = 0;
        int rc = sgcpu_pack_size(h, n, lo.data(), hi.data(), nprop, props, 8, &tot);
        if (rc) return rc;
        std::vector<unsigned char> buf((size_t)tot + 8);
        if ((rc = sgcpu_pack(h, n, lo.data(), hi.data(), nprop, props, 8, buf.data(), tot))) return rc;
        int64_t at = 0;
        for (int k = 0; k < n; ++k) {
            int64_t used = 0;
            rc = sgcpu_unpack(g.copies[k].dst, buf.data() + at, tot - at, g.copies[k].dst_lo, nprop, props, 8, &used);
            if (rc) return rc;
            at += used;
        }
        g.copies.clear();
        return SG_OK;
    }
    case SG_COPY_PHASE3: return flush_unpacked(g, nprop, props);
    default: return fail(SG_E_ARG, "bad phase");
    }
}

int sgcpu_remove_add_unpack_reset(void *h)
{
    Grid &g = *(Grid *)h;
    g.rem_lo.clear();
    g.rem_hi.clear();
    g.copies.clear();
    clear_inserts(g);
    return SG_OK;
}

// construct_link_dw / construct_link_up (SparseGridGpu.hpp:2090-2262).  count_paddings /
// collect_paddings (SparseGridGpu_kernels.cuh:775-838): the PADDING points inside the box,
// chunk by chunk (here by increasing offset inside a chunk).  link_construct_dw_count /
// link_construct_insert_dw (:846-966): for j < 2 * dim (the reference's bound), the point
// 2 x + p_dw + ((j >> k) & 1)_k of the finer grid, if its EXIST bit is set.
// link_construct_up_count / link_construct_insert_up (:886-1000): the point (x - p_up) / 2
// (C division) of the coarser grid, if it exists.  Points outside the other grid's domain
// do not exist.
static bool link_target(const Grid &o, const int64_t *q, int32_t *pos, int16_t *off)
{
    for (int d = 0; d < o.dim; ++d)
        if (q[d] < 0 || q[d] >= o.sz[d]) return false;
    const int64_t key = o.lin_id(q);
    const int64_t c = o.find(key / o.csz);
    if (c >= o.n()) return false;
    const int l = (int)(key % o.csz);
    if (!(o.mask[c * o.csz + l] & SG_EXIST)) return false;
    *pos = (int32_t)c;
    *off = (int16_t)l;
    return true;
}

int sgcpu_construct_link(void *h, void *other, int which, const int32_t *lo, const int32_t *hi, const int32_t *p)
{
    Grid &g = *(Grid *)h;
    const Grid &o = *(const Grid *)other;
    if (which != SG_LINK_DOWN && which != SG_LINK_UP) return fail(SG_E_ARG, "link direction must be down or up");
    Grid::Links L;
    for (int64_t c = 0; c < g.n(); ++c)
        for (int l = 0; l < g.csz; ++l) {
            if (!(g.mask[c * g.csz + l] & SG_PADDING)) continue;
            int64_t x[3] = {0, 0, 0};
            g.inv_lin(g.keys[c] * g.csz + l, x);
            bool in = true;
            for (int d = 0; d < g.dim; ++d) in = in && x[d] >= lo[d] && x[d] <= hi[d];
            if (!in) continue;
            L.pc.push_back((int32_t)c);
            L.po.push_back((int16_t)l);
            if (which == SG_LINK_DOWN) {
                for (int j = 0; j < 2 * g.dim; ++j) {
                    int64_t q[3] = {0, 0, 0};
                    for (int d = 0; d < g.dim; ++d) q[d] = x[d] * 2 + p[d] + ((j >> d) & 1);
                    int32_t pos;
                    int16_t off;
                    if (link_target(o, q, &pos, &off)) {
                        L.lc.push_back(pos);
                        L.lo.push_back(off);
                    }
                }
            } else {
                int64_t q[3] = {0, 0, 0};
                for (int d = 0; d < g.dim; ++d) q[d] = (x[d] - p[d]) / 2;
                int32_t pos;
                int16_t off;
                if (link_target(o, q, &pos, &off)) {
                    L.lc.push_back(pos);
                    L.lo.push_back(off);
                }
            }
            L.scan.push_back((uint32_t)L.lc.size());  // scan[i + 1]: end of padding point i's links
        }
    g.links[which] = L;
    return SG_OK;
}

int sgcpu_link_sizes(void *h, int which, int64_t *n_pad, int64_t *n_links)
{
    const Grid &g = *(const Grid *)h;
    *n_pad = (int64_t)g.links[which].pc.size();
    *n_links = (int64_t)g.links[which].lc.size();
    return SG_OK;
}

int sgcpu_get_links(void *h, int which, int32_t *pad_chunk, int16_t *pad_off, uint32_t *scan, int32_t *link_chunk,
                    int16_t *link_off)
{
    const Grid::Links &L = ((const Grid *)h)->links[which];
    if (pad_chunk) std::copy(L.pc.begin(), L.pc.end(), pad_chunk);
    if (pad_off) std::copy(L.po.begin(), L.po.end(), pad_off);
    if (scan) std::copy(L.scan.begin(), L.scan.end(), scan);
    if (link_chunk) std::copy(L.lc.begin(), L.lc.end(), link_chunk);
    if (link_off) std::copy(L.lo.begin(), L.lo.end(), link_off);
    return SG_OK;
}

int sgcpu_set_linearizer(void *h, int lin)
{
    Grid &g = *(Grid *)h;
    if (lin != SG_LIN_SMB && lin != SG_LIN_ZMB) return fail(SG_E_ARG, "unknown linearizer");
    if (g.n() != 0 || !g.s_key.empty()) return fail(SG_E_STATE, "the linearizer can only change on an empty grid");
    g.zorder = lin == SG_LIN_ZMB;
    return SG_OK;
}

// lin_zid of one key (known-answer tests of zmorton_unit_tests.cpp)
uint64_t sgcpu_morton(int dim, const int64_t *coord) { return Grid::morton(coord, dim); }

int64_t sgcpu_lin_id(void *h, const int64_t *coord) { return ((Grid *)h)->lin_id(coord); }

int sgcpu_inv_lin_id(void *h, int64_t lin, int64_t *coord)
{
    ((Grid *)h)->inv_lin(lin, coord);
    return SG_OK;
}

int sgcpu_set_threads(int n)
{
    omp_set_num_threads(n);
    return SG_OK;
}

}  // extern "C"
