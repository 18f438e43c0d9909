/*
 * sparsegrid_gpu.h -- C ABI of the MI355X (gfx950) block-sparse grid library
 * (libsg_hip.so).  Plain pointers and sizes only; no C++ or torch types.
 *
 * The library re-implements the SparseGridGpu hot path of openfpm_data:
 * chunk-key insertion + flush (sort, duplicate merge with smax/sadd/...,
 * merge into the sorted chunk map), findNeighbours, tagBoundaries,
 * applyStencils/conv over edge^dim chunks, and point get.  The C++ class
 * template in SparseGridGpu.hpp (same directory) is the drop-in facade over
 * these entry points; every entry point below names the reference interface
 * it replaces (paths relative to the openfpm_data source tree).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative SG_E* code; the
 *    message of the last failure on the calling thread is sg_last_error().
 *    (The reference prints to std::cout and continues,
 *    src/SparseGridGpu/SparseGridGpu.hpp:2350, :3158; an error code never
 *    throws across the ABI.)
 *  - Work is stream ordered on the handle's HIP stream (sg_set_stream);
 *    calls that return sizes to the host synchronise that stream.
 *  - `mem` arguments say where a caller buffer lives: SG_MEM_HOST or
 *    SG_MEM_DEVICE (a device pointer valid on the handle's device).
 *  - Point keys are grid_smb linear ids: chunkId * edge^dim + localId, with
 *    chunkId row-major x-fastest over ceil(size/edge) chunks and localId
 *    x-fastest inside the chunk (src/Grid/Geometry/grid_smb.hpp:175-189).
 *  - Layout in device memory (chunk-major SoA, as the reference's
 *    memory_traits_inte DataBlock arrays, src/SparseGridGpu/BlockMapGpu.hpp:63-69):
 *    one array per property of (nChunks+1)*edge^dim scalars and one uint8
 *    mask array of the same shape; chunk nChunks is the background chunk
 *    (src/Vector/map_vector_sparse.hpp:1356-1364).
 */
#ifndef SPARSEGRID_GPU_H
#define SPARSEGRID_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sg_grid sg_grid;

/* status codes */
#define SG_OK 0
#define SG_E_ARG (-1)     /* invalid argument */
#define SG_E_HIP (-2)     /* HIP runtime failure */
#define SG_E_OOM (-3)     /* device allocation failed */
#define SG_E_STATE (-4)   /* call not valid in this state */
#define SG_E_RANGE (-5)   /* size or coordinate out of the supported range */

/* where a caller buffer lives */
#define SG_MEM_HOST 0
#define SG_MEM_DEVICE 1

/* property scalar types */
typedef enum {
    SG_F32 = 0, SG_F64 = 1, SG_I32 = 2, SG_U32 = 3, SG_I64 = 4,
    SG_U64 = 5, SG_I8 = 6, SG_U8 = 7, SG_I16 = 8, SG_U16 = 9
} sg_dtype;

/* flush reductions: smax_, smin_, sadd_, sRight_, sLeft_, sBitwiseOr_
 * (src/Vector/cuda/map_vector_sparse_cuda_kernels.cuh:50-450).
 * SG_OP_NONE marks a property absent from the reference's v_reduce list;
 * the reference leaves such a property unwritten (undefined) for merged
 * chunks; this library defines it as "last inserted value wins, points
 * without an insert keep their old value". */
typedef enum {
    SG_OP_NONE = 0, SG_OP_MAX = 1, SG_OP_MIN = 2, SG_OP_ADD = 3,
    SG_OP_RIGHT = 4, SG_OP_LEFT = 5, SG_OP_OR = 6
} sg_op;

/* mask bits (src/SparseGridGpu/SparseGridGpu_kernels.cuh:21-27) */
#define SG_EXIST 1
#define SG_PADDING 2

/* neighbourhood types: NNStar (2*dim face chunks) and NNFull (3^dim chunks)
 * (src/SparseGridGpu/SparseGridGpu_ker_util.hpp:91-157,
 *  src/SparseGridGpu/SparseGridGpu.hpp:262-352) */
#define SG_NN_STAR 0
#define SG_NN_FULL 1

/* tagBoundaries option (src/SparseGridGpu/SparseGridGpu.hpp:31-35) */
#define SG_NO_CALCULATE_EXISTING_POINTS 0
#define SG_CALCULATE_EXISTING_POINTS 1

/* Stencil descriptors.  The reference takes an arbitrary __device__ functor
 * (src/SparseGridGpu/SparseGridGpu.hpp:2704-2734); a C ABI cannot carry one,
 * so the stencils used by the reference's tests and benchmarks are named
 * here and evaluated by hand-written kernels. */
typedef enum {
    /* dst = c0*u + sum_d c1*(u[-e_d] + u[+e_d]); conv with the 7-point
     * Laplacian of HeatStencil (tests/utils/SparseGridGpu_util_test.cuh:240-255) */
    SG_ST_LAP = 0,
    /* dst = u + coef[0] * (-2*dim*u + sum_d (u[-e_d] + u[+e_d]))
     * (HeatStencil, tests/utils/SparseGridGpu_util_test.cuh:201-260) */
    SG_ST_HEAT = 1,
    /* two-property Gray-Scott reaction-diffusion step (conv2 form,
     * src/SparseGridGpu/SparseGridGpu.hpp:2668-2683):
     * coef = {Du, Dv, F, k, dt};  src = {u, v}, dst = {u', v'} */
    SG_ST_GRAY_SCOTT = 2,
    /* 3^dim box convolution, tot = sum_{i,j,k} coef[i*9+j*3+k]*u[i-1,j-1,k-1]
     * (Conv3x3x3, tests/utils/SparseGridGpu_util_test.cuh:357-420);
     * integer properties use icoef[] and wrapping integer arithmetic */
    SG_ST_BOX = 3,
    /* red-black SOR sweep of -lap(u) = f, in place on src[0] with f = src[1]:
     * u = (1-w)*u + w*(f + sum of 2*dim neighbours)/(2*dim) on points with
     * (x+y+z)%2 == color; coef = {w}, color = icoef[0] */
    SG_ST_RBSOR = 4,
    /* residual r = f - A u with A = -lap (7 point): dst[0] = src[1] -
     * (2*dim*u - sum neighbours), u = src[0] */
    SG_ST_RESIDUAL = 5
} sg_stencil_kind;

/* how a stencil runs (StencilMode, src/SparseGridGpu/SparseGridGpu.hpp:150-154):
 * INPLACE stages chunk + ghost layer in LDS (applyStencilInPlace, :954-1048);
 * INPLACE_NO_SHARED runs one thread per existing point of the point list, reading
 * neighbours through the chunk table (applyStencilInPlaceNoShared, :1051-1075), so it
 * writes only existing (EXIST) points.  Both give the same values. */
#define SG_STENCIL_MODE_INPLACE 1
#define SG_STENCIL_MODE_INPLACE_NO_SHARED 3

/* which points a stencil writes */
#define SG_APPLY_INTERIOR 0  /* EXIST && !PADDING  (conv/conv2 functors, SparseGridGpu_kernels.cuh:281) */
#define SG_APPLY_EXISTING 1  /* mask != 0          (Conv3x3x3 called with bool(mask)) */

typedef struct {
    int32_t kind;          /* sg_stencil_kind */
    int32_t apply;         /* SG_APPLY_* */
    int32_t src[4];        /* property indices read */
    int32_t dst[4];        /* property indices written */
    int32_t use_box;       /* 0: whole grid; 1: restrict to [box_lo, box_hi] inclusive */
    int32_t box_lo[3];
    int32_t box_hi[3];
    int32_t mode;          /* SG_STENCIL_MODE_* (0 = SG_STENCIL_MODE_INPLACE) */
    double coef[27];       /* floating coefficients (see kinds) */
    int64_t icoef[27];     /* integer coefficients / integer parameters */
} sg_stencil_desc;

/* ---- lifetime ---------------------------------------------------------- */

/* SparseGridGpu(const size_t (&res)[dim], stencilSupportRadius)
 * (src/SparseGridGpu/SparseGridGpu.hpp:1694-1698) for
 * aggregate<dtypes...> and chunk edge block_edge (dim 2 or 3;
 * block_edge^dim <= 1024; stencil_radius must be 1). */
int sg_create(int dim, const int64_t *size, int block_edge, int nprops,
              const int32_t *dtypes, int stencil_radius, sg_grid **out);
/* waits for the handle's stream; its device blocks are kept, idle, for later handles on
 * the same device to allocate from (sg_trim on any handle of the device frees them) */
int sg_destroy(sg_grid *g);
/* chunk linearizer: SG_LIN_SMB = grid_smb, row-major chunk ids (SparseGridGpu, the
 * default); SG_LIN_ZMB = grid_zmb, Morton chunk ids (SparseGridGpu_z,
 * src/SparseGridGpu/SparseGridGpu.hpp:3947-3972, src/Grid/Geometry/grid_zmb.hpp).  Only on
 * an empty grid (no chunks, nothing pending). */
#define SG_LIN_SMB 0
#define SG_LIN_ZMB 1
int sg_set_linearizer(sg_grid *g, int lin);
const char *sg_last_error(void);
/* set the HIP stream (hipStream_t as void*) all later work goes to; NULL = default stream */
int sg_set_stream(sg_grid *g, void *stream);
int sg_synchronize(sg_grid *g);

/* setBackgroundValue<p> (src/SparseGridGpu/SparseGridGpu.hpp:2878-2884,
 * BlockMapGpu.hpp:306-326); value points to one host scalar of the property type */
int sg_set_background(sg_grid *g, int prop, const void *value);

/* ---- insertion + flush ---------------------------------------------------- */

/* Point insertion: sg.insert<p>(coord) per point, one insert slot per point
 * (src/SparseGridGpu/BlockMapGpu_ker.cuh:171-197,
 *  src/Vector/cuda/map_vector_sparse_cuda_ker.cuh:309-324).  Records are
 * appended to the handle's insert buffer in call order and merged by
 * sg_flush in that order.  masks may be NULL (every record EXIST);
 * values[p] may be NULL: the records then carry the property's background value, which
 * the flush merges with the property's reduction like any inserted value. */
int sg_insert_points(sg_grid *g, int64_t n, const int64_t *keys,
                     const uint8_t *masks, const void *const *values, int mem);
/* same, with integer coordinates [n][dim] instead of linear keys; host coordinates
 * outside the grid fail here (SG_E_RANGE), device ones fail the next sg_flush (which then
 * drops those records only: the other records stay pending, unmerged, for the next
 * sg_flush), so the call does not wait for the device.  Device coordinates inserted
 * while nothing else is pending are staged in the flush's own (chunk id, record) form;
 * any other insert moves them to the insert buffer first (same order, same result). */
int sg_insert_coords(sg_grid *g, int64_t n, const int32_t *coords,
                     const uint8_t *masks, const void *const *values, int mem);
/* Chunk insertion: insertBlock(blockId) per chunk slot
 * (src/SparseGridGpu/BlockMapGpu_ker.cuh:199-227); masks is [n][edge^dim],
 * values[p] is [n][edge^dim]. */
int sg_insert_chunks(sg_grid *g, int64_t n, const int64_t *chunk_ids,
                     const uint8_t *masks, const void *const *values, int mem);
/* number of point records waiting in the insert buffer (a whole-chunk record counts as
 * its slots with a non-zero mask, at least one); -1 on failure */
int64_t sg_pending_inserts(sg_grid *g);
/* flush<v_reduce...>(ctx, FLUSH_ON_DEVICE) (src/SparseGridGpu/SparseGridGpu.hpp:772-779,
 * src/Vector/map_vector_sparse.hpp:1192-1226); ops[p] for every property. */
int sg_flush(sg_grid *g, const int32_t *ops);

/* ---- neighbourhood, boundaries, stencils ---------------------------------- */

/* findNeighbours<NN>() (src/SparseGridGpu/SparseGridGpu.hpp:2396-2420) */
int sg_find_neighbours(sg_grid *g, int nn_type);
/* setNNType<NN>() (src/SparseGridGpu/SparseGridGpu.hpp:1825-1829) */
int sg_set_nn_type(sg_grid *g, int nn_type);
/* tagBoundaries<NN>(ctx, No_check(), opt) (src/SparseGridGpu/SparseGridGpu.hpp:2290-2394);
 * opt = SG_CALCULATE_EXISTING_POINTS also builds the existing-point list (e_points,
 * :2354-2391) that SG_STENCIL_MODE_INPLACE_NO_SHARED runs over */
int sg_tag_boundaries(sg_grid *g, int nn_type, int opt);
/* size and contents of the existing-point list (getPointBuffer, SparseGridGpu_ker.cuh:296-327):
 * data positions chunk * chunk_size + offset of the EXIST points, in chunk order and
 * increasing offset; rebuilt here when a flush or remove made it stale */
int sg_existing_points(sg_grid *g, int64_t *n);
int sg_get_existing_points(sg_grid *g, int64_t *out, int mem);
/* applyStencils / conv / conv2 (src/SparseGridGpu/SparseGridGpu.hpp:2568-2734) */
int sg_apply_stencil(sg_grid *g, const sg_stencil_desc *desc);

/* ---- batched Poisson solves (no counterpart in the reference: a solver built from the
 *      stencil sweeps of this path) ------------------------------------------------------
 * -lap(u) = f on the existing points, u = 0 outside; 3D grids with 8^3 chunks.
 * SG_POISSON_MIXED: FP32 red-black SOR on the correction e with right-hand side r, FP64
 * only for the defect correction u <- u + e, the residual r = f - A u and its norm.
 * SG_POISSON_PLAIN_FP64: red-black SOR on u in FP64, residual checked between blocks.
 * Property layout of each grid: u, u2, f double; e, r float (u2, e, r unused by the plain
 * form); u holds the initial guess when added (e, r, u2 need no initial values).
 * A grid's iterates do not depend on the other grids of its batch: it is solved to the
 * same bits alone or in any mix.  The stream may be null (the batch then makes its own;
 * graphs cannot be captured on the legacy null stream). */
#define SG_POISSON_MIXED 0
#define SG_POISSON_PLAIN_FP64 1

typedef struct {
    int32_t mode;          /* SG_POISSON_* */
    int32_t u_prop, u2_prop, f_prop, e_prop, r_prop;
    double tol;            /* relative residual ||f - A u|| / ||f|| to reach */
    double omega;          /* SOR relaxation; <= 0: estimated per grid */
    double omega0;         /* relaxation of the first block when estimating (default 1.5) */
    double inner_target;   /* residual reduction asked of one inner block (default 1e-5 mixed, 1e-3 plain) */
    int32_t first_sweeps;  /* sweeps of the first block (default 24) */
    int32_t max_sweeps;    /* per grid (default 20000) */
    int32_t max_outer;     /* per grid (default 100) */
    int32_t use_graph;     /* capture the inner sweep sequence in a hipGraph */
} sg_poisson_params;

typedef struct {
    int32_t converged;     /* 1 when rel_residual <= tol */
    int32_t outer;         /* outer iterations (residual evaluations - 1) */
    int32_t sweeps;        /* red-black sweeps */
    int32_t solution_prop; /* property holding the solution u */
    double rel_residual;   /* final ||f - A u|| / ||f|| (FP64) */
    double omega;          /* relaxation used */
} sg_poisson_result;

typedef struct sg_batch sg_batch;
/* continuous batching: grids join with sg_batch_add and leave when converged; every
 * sg_batch_step runs one outer iteration for all live grids in shared launches */
int sg_batch_create(const sg_poisson_params *p, void *stream, sg_batch **out);
int sg_batch_destroy(sg_batch *b);
/* tol > 0: this grid's relative residual target (else the batch's p->tol) */
int sg_batch_add(sg_batch *b, sg_grid *g, int64_t tag, double tol);
int sg_batch_live(sg_batch *b);
/* launch shape of the batched sweep kernel: waves (64 lanes, one chunk tile each) per
 * workgroup, 1 / 2 / 4 / 8 (LDS tile per wave: 3.2 KB FP32, 6.4 KB FP64), default 4 */
#define SG_TUNE_BATCH_WAVES 5
/* at most this many workgroups per launch, waves then loop over tiles; 0 (default): one
 * tile per wave */
#define SG_TUNE_BATCH_MAX_WGS 6
int sg_batch_set_tuning(sg_batch *b, int key, int64_t value);
int sg_batch_step(sg_batch *b);
int sg_batch_poll(sg_batch *b, int64_t *tags, sg_poisson_result *results, int max, int *n);
/* solve all grids to p->tol in one batch (results[i] for grids[i]) */
int sg_poisson_solve(sg_grid *const *grids, int n, const sg_poisson_params *p, sg_poisson_result *results);

/* ---- launch tuning (no counterpart in the reference: its launch shapes are
 *      compile-time template parameters, SparseGridGpu.hpp:483-489) -------------- */

/* workgroups (of 4 waves) per XCD of the persistent 8^3 star-stencil kernel */
#define SG_TUNE_STENCIL_WG_PER_XCD 1
/* 8^3 star-stencil kernel variant: 0 one chunk per 64-lane workgroup, 1 persistent
 * software-pipelined waves, 2 as 1 with registers capped for 4 waves per SIMD,
 * 3 / 4 / 5 workgroups of 4 / 8 / 2 consecutive chunks sharing x faces in LDS,
 * 6 / 7 / 8 as 3 / 4 / 5 with one LDS tile per wave for both Gray-Scott properties */
#define SG_TUNE_STENCIL_VARIANT 2
/* 8^3 star-stencil stores: 0 active points only, 1 whole 8-point rows where every point
 * is active or absent (absent points keep the background), 2 as 1, non-temporal, 3 as 2
 * with the rows passed through LDS so that each store instruction writes 1 KB of the
 * chunk contiguously */
#define SG_TUNE_STENCIL_STORES 3
/* 8^3 star-stencil loads: 0 every row; 1 read the mask first and skip rows without points;
 * 2 skip the own and ghost rows without points as the per-chunk occupancy table says
 * (rebuilt after any mask change); 3 as 2 for the ghost rows only; 4 as 2 with each row's
 * mask read from two 64-bit planes per chunk built with the table */
#define SG_TUNE_STENCIL_SKIP_EMPTY_ROWS 4
/* 8^3 tagBoundaries on the mask rows (SG_TUNE_TAG_VARIANT 1): chunks per wave, 1, 2
 * (default) or 4; a wave issues the loads of all its chunks before tagging any */
#define SG_TUNE_TAG_CHUNKS_PER_WAVE 7
/* flush of point records: 0 (default) sorts (chunk id, record) pairs and folds each
 * chunk's records in LDS where chunk ids fit 32 bits and a record 32 bytes; 1 always
 * sorts point keys.  Both give the same result. */
#define SG_TUNE_FLUSH_PATH 8
/* 8^3 tagBoundaries: 0 (default) works on per-chunk EXIST bit-planes (64 bytes per chunk,
 * one 64-bit word per z slice: kept current by the flush of point records grouped by
 * chunk, rebuilt after other mask changes); 1 reads the neighbours' mask rows.  Same
 * result. */
#define SG_TUNE_TAG_VARIANT 9
/* 8^3 27-point box: 1 (default) two consecutive chunks per wave, sharing the x ghosts
 * between them when they are x neighbours; 0 one chunk per wave.  Same result. */
#define SG_TUNE_BOX_PAIRS 10
int sg_set_tuning(sg_grid *g, int key, int64_t value);

/* ---- queries ---------------------------------------------------------------- */

int64_t sg_num_chunks(sg_grid *g);
int sg_chunk_size(sg_grid *g);
/* sorted chunk ids (private_get_index_array), [nChunks] int64 */
int sg_get_chunk_ids(sg_grid *g, int64_t *out, int mem);
/* masks [nChunks][edge^dim] (private_get_data_array<pMask>) */
int sg_get_masks(sg_grid *g, uint8_t *out, int mem);
/* property p [nChunks][edge^dim] (deviceToHost<p>) */
int sg_get_prop(sg_grid *g, int prop, void *out, int mem);
/* overwrite property p of the existing chunks (hostToDevice<p>) */
int sg_set_prop(sg_grid *g, int prop, const void *in, int mem);
/* hostToDevice of the chunk list and the masks (BlockMapGpu.hpp:230-241, the index buffer
 * and pMask of map_vector_sparse::hostToDevice): the grid's chunks become ids[n] (strictly
 * increasing chunk ids of the domain) with masks [n][edge^dim].  A chunk that was already
 * there keeps its property values, a new chunk takes the background; every slot whose new
 * mask is 0 takes the background.  Pending inserts stay pending; neighbour tables and the
 * existing-point list are rebuilt when next needed.  The host insertFlush of the facade
 * (SparseGridGpu.hpp:2780-2845) edits a host copy that goes back through this call. */
int sg_set_chunks(sg_grid *g, int64_t n, const int64_t *ids, const uint8_t *masks, int mem);
/* Restore the layout invariant after the caller's kernels wrote the arrays of sg_to_kernel
 * (a user applyStencils functor, a store through the device view): every slot whose mask is
 * 0 takes the background of property prop (every property for prop == -1).  The reference
 * substitutes the background for such slots on every read (loadGhostBlock_impl,
 * SparseGridGpu_ker_util.hpp:551-552), so reads see the same values. */
int sg_restore_background(sg_grid *g, int prop);
/* get<p>(coord) for n point keys; background where the point does not exist
 * (src/SparseGridGpu/SparseGridGpu.hpp:1897-1901, BlockMapGpu.hpp:108-126) */
int sg_get_points(sg_grid *g, int prop, int64_t n, const int64_t *keys,
                  void *out, int mem);
/* neighbour table [nChunks][nNN] int32 built by sg_find_neighbours
 * (private_get_neighborhood_array); nNN = 2^dim (star) or 3^dim (full) */
int sg_get_neighbours(sg_grid *g, int nn_type, int32_t *out, int mem);
/* device address of property p's array (prop == -1: the mask array) */
int sg_device_ptr(sg_grid *g, int prop, void **out);

/* SparseGridGpu::toKernel() (src/SparseGridGpu/SparseGridGpu.hpp:1716-1751): the grid's
 * device arrays, for kernels compiled by the caller (the facade's conv family,
 * include/SparseGridGpu_device.hpp).  Builds the nn_type neighbour table when it is stale.
 * The pointers stay valid until the next call that changes the chunk set or a property
 * array (flush, remove, unpack, a stencil writing into one of its sources, ...); work on
 * them belongs on `stream`, the handle's stream. */
typedef struct sg_kernel_view {
    const int64_t *keys;   /* sorted chunk ids [nchunks] */
    uint8_t *mask;         /* [(nchunks + 1) * chunk_size]; chunk nchunks is the background */
    void *prop[16];        /* property arrays, [(nchunks + 1) * chunk_size] each */
    const int32_t *nn;     /* neighbour table [nchunks][nnn] (positions; nchunks = missing) */
    void *stream;          /* hipStream_t of the handle */
    int64_t nchunks;
    int64_t size[3];       /* domain size in points */
    int64_t bsz[3];        /* domain size in chunks */
    int32_t nnn, dim, edge, chunk_size, zorder, nprops;
} sg_kernel_view;
int sg_to_kernel(sg_grid *g, int nn_type, sg_kernel_view *out);

/* setGPUInsertBuffer + device insertBlock / insert<p> / flush_block_insert
 * (src/SparseGridGpu/SparseGridGpu.hpp:2270-2278, BlockMapGpu.hpp:249-287,
 * BlockMapGpu_ker.cuh:171-227): a zeroed staging area of `nrec` chunk records for the
 * caller's kernels.  A kernel claims record r with atomicAdd(count, 1) (r >= cap: write
 * record cap, a sink, instead), writes ids[r], sets mask bits and values at
 * [r * chunk_size + offset]; sg_insert_device_end (or the next sg_flush) appends the
 * claimed records as chunk records (sg_insert_chunks) and fails with SG_E_RANGE when
 * more than `cap` were claimed.  Values of a property not written at an inserted point
 * are 0.  The kernels run on the handle's stream (or finish before sg_insert_device_end:
 * it reads the claim counter in the handle's stream order). */
typedef struct sg_insert_view {
    unsigned int *count;   /* records claimed so far */
    int64_t *ids;          /* [cap + 1] chunk id of each record */
    uint8_t *mask;         /* [(cap + 1) * chunk_size] */
    void *prop[16];        /* [(cap + 1) * chunk_size] per property */
    int64_t cap;
    int32_t chunk_size;
} sg_insert_view;
int sg_insert_device_begin(sg_grid *g, int64_t nrec, sg_insert_view *out);
int sg_insert_device_end(sg_grid *g);
/* countExistingElements / countBoundaryElements (src/SparseGridGpu/SparseGridGpu.hpp:2422-2490) */
int sg_count_existing(sg_grid *g, int64_t *out);
int sg_count_boundary(sg_grid *g, int64_t *out);

/* ---- pack / unpack / remove ---------------------------------------------------
 * Packed form of one box (packCalculate + pack + packFinalize,
 * src/SparseGridGpu/SparseGridGpu.hpp:3072-3340; pack_data, SparseGridGpu_kernels.cuh:1658-1697),
 * sections in this order, A = index_bytes = sizeof(indexT) (4 or 8):
 *   uint64 n_cnk | int32 box_start[dim] | int32 grid_size[dim] | indexT chunk_ids[n_cnk]
 *   | uint32 scan[n_cnk + 1] (first point of each chunk, scan[n_cnk] = n_pnt), padded to A
 *   | per packed property T[n_pnt], one after the other, padded to A
 *   | int16 offsets[n_pnt], padded to A | uint8 masks[n_pnt], padded to A
 * Boxes are inclusive [lo, hi] ([nbox][dim] corners, Box::isInside); only existing points
 * (EXIST) are packed; chunks without such a point are left out.  An array property of the
 * reference (T[N]) is N consecutive scalar properties here (same bytes). */
/* packRequest + packCalculate: total bytes of the packed boxes */
int sg_pack_size(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop,
                 const int32_t *props, int index_bytes, int64_t *bytes);
/* pack + packFinalize: write the packed boxes (in box order) to buf (cap bytes) */
int sg_pack(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop,
            const int32_t *props, int index_bytes, void *buf, int64_t cap, int mem);
/* sg_pack into device memory without allocating.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (loops replayed under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc).  The count and scan temporaries live in
 * one static device array of SG_PACK_NOALLOC_SCRATCH_BYTES per GPU (held until the process
 * ends) instead of memory the grid allocates; no hipMalloc / hipFree is called.  Limits:
 * nbox <= SG_PACK_NOALLOC_MAX_BOXES and 16 * nbox * (chunks + 1) bytes of rows plus the
 * corners and the scan workspace within the scratch (SG_E_RANGE otherwise, nothing done).
 * dev_buf is device memory of cap bytes; the bytes written equal sg_pack's with
 * SG_MEM_DEVICE, *written = their count; returns when the pack is complete.  Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU (several host threads,
 * or grids on several streams) race. */
#define SG_PACK_NOALLOC_SCRATCH_BYTES (32u << 20)
#define SG_PACK_NOALLOC_MAX_BOXES 64
int sg_pack_noalloc(sg_grid *g, int nbox, const int32_t *lo, const int32_t *hi, int nprop,
                    const int32_t *props, int index_bytes, void *dev_buf, int64_t cap, int64_t *written);
/* unpack (SparseGridGpu.hpp:3665-3733): read the packed box at buf; its points land at
 * (packed coordinate - box_start + dst_lo) and are queued for
 * sg_remove_add_unpack_finalize; *consumed = bytes of the section (the next box starts
 * there).  A point landing outside the grid fails the call (SG_E_RANGE). */
int sg_unpack(sg_grid *g, const void *buf, int64_t avail, int mem, const int32_t *dst_lo, int nprop,
              const int32_t *props, int index_bytes, int64_t *consumed);
/* remove(Box) (SparseGridGpu.hpp:3528-3531): queue an inclusive box whose existing points
 * the next sg_remove_add_unpack_finalize removes (chunks stay) */
int sg_remove_box(sg_grid *g, const int32_t *lo, const int32_t *hi);
/* device remove(coord) (SparseGridGpu_ker.cuh:665-668, BlockMapGpu_ker.cuh:322-332): the
 * given points get mask 0 now (value -> background); keys of absent chunks are ignored */
int sg_remove_points(sg_grid *g, int64_t n, const int64_t *keys, int mem);
/* removeAddUnpackFinalize (SparseGridGpu.hpp:3492-3499): remove the points of the queued
 * boxes, then merge the pending records: the listed (unpacked) properties take the new
 * values, the others keep theirs, mask |= EXIST */
int sg_remove_add_unpack_finalize(sg_grid *g, int nprop, const int32_t *props);
/* removeAddUnpackReset (SparseGridGpu.hpp:3362-3383): drop queued boxes, queued copies
 * and pending records */
int sg_remove_add_unpack_reset(sg_grid *g);

/* copy_to (SparseGridGpu.hpp:3550-3559, sparse_grid_section :467-480): queue, on the
 * source grid, a copy of the existing points of its inclusive box [src_lo, src_hi] into
 * dst, the box's low corner landing at dst_lo.  The grids must have the same dimension
 * and, for the copied properties, the same types; src may be dst. */
int sg_copy_to(sg_grid *dst, sg_grid *src, const int32_t *src_lo, const int32_t *src_hi, const int32_t *dst_lo);
/* removeCopyToFinalize (SparseGridGpu.hpp:3314-3329, phases :1116-1290; rem_copy_opt,
 * Grid/grid_common.hpp:180-188).  Call PHASE1 on every grid taking part, then PHASE2 on
 * every grid, then PHASE3 on every grid:
 *   PHASE1  remove the points of the boxes queued by sg_remove_box;
 *   PHASE2  pack this grid's queued copies (on the device) and queue their points in the
 *           destinations;
 *   PHASE3  merge the points queued in this grid: the listed properties take the copied
 *           values, the others keep theirs, mask |= EXIST.
 * The KEEP_GEOMETRY bit (4) of the reference is accepted and has no effect. */
#define SG_COPY_PHASE1 0
#define SG_COPY_PHASE2 1
#define SG_COPY_PHASE3 2
#define SG_COPY_KEEP_GEOMETRY 4
int sg_remove_copy_to_finalize(sg_grid *g, int phase, int nprop, const int32_t *props);

/* ---- links between levels (construct_link_dw / construct_link_up,
 *      SparseGridGpu.hpp:2090-2262; kernels SparseGridGpu_kernels.cuh:775-1000) ----
 * For the padding points (PADDING bit) of g inside the inclusive box [lo, hi] -- chunk by
 * chunk in chunk-list order, by increasing offset inside a chunk -- the linked points of
 * `other`: SG_LINK_DOWN, the existing points 2 x + p + ((j >> k) & 1)_k for j < 2 dim
 * (the reference's loop bound: 6 of the 8 children in 3D); SG_LINK_UP, the existing
 * point (x - p) / 2 (C division).  A point outside `other`'s domain does not exist.
 * The result is kept in g per direction until the next call of that direction. */
#define SG_LINK_DOWN 0
#define SG_LINK_UP 1
int sg_construct_link(sg_grid *g, sg_grid *other, int which, const int32_t *lo, const int32_t *hi, const int32_t *p);
/* number of padding points and of links of the last sg_construct_link of `which` */
int sg_link_sizes(sg_grid *g, int which, int64_t *n_pad, int64_t *n_links);
/* getDownLinks / getDownLinksOffsets / getUpLinks / getUpLinksOffsets: padding points
 * (chunk position, offset in chunk) [n_pad], offsets of each padding point's links
 * [n_pad + 1] (exclusive scan, last = n_links), links (chunk position, offset in chunk
 * of `other`) [n_links]; null outputs are skipped */
int sg_get_links(sg_grid *g, int which, int32_t *pad_chunk, int16_t *pad_off, uint32_t *scan, int32_t *link_chunk,
                 int16_t *link_off, int mem);

/* ---- device memory (no counterpart in the reference: its vectors keep their
 *      capacity; a server admitting grids against an HBM budget needs both) ---- */
/* wait for the handle's stream, then free the flush temporaries and cached blocks, and
 * the device blocks the library keeps from destroyed handles for reuse by new ones; the
 * chunk map, properties, masks and neighbour tables stay */
int sg_trim(sg_grid *g);
/* No counterpart in the reference: a streaming device copy of `bytes` (a multiple of 16,
 * 16-byte aligned buffers) on `stream` (hipStream_t as void*), plain or with non-temporal
 * loads and stores; the HBM ceiling that the stencil kernels are measured against. */
#define SG_COPY_NONTEMPORAL 1
int sg_stream_copy(void *dst, const void *src, int64_t bytes, int flags, void *stream);
/* bytes of device memory the handle holds (resident arrays + cached temporaries) */
int sg_device_bytes(sg_grid *g, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* SPARSEGRID_GPU_H */
