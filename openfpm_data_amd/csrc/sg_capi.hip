// sg_capi.hip -- the extern "C" entry points of libsg_hip.so (include/sparsegrid_gpu.h).
// Argument checking, handle lifetime, host<->device copies; the kernels live in
// sg_flush.hip, sg_stencil.hip and sg_primitives.hip.
#include "sg_internal.hpp"

#include <algorithm>
#include <mutex>

namespace sg {

static thread_local std::string t_err;

void set_error(const std::string &msg) { t_err = msg; }

int fail(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

namespace {

// idle blocks of destroyed handles: size -> (block, device)
std::mutex g_idle_mu;
std::multimap<size_t, std::pair<void *, int>> g_idle;

// an idle block of the device with at least `bytes` and at most twice that
std::pair<void *, size_t> take_idle(size_t bytes, int dev)
{
    std::lock_guard<std::mutex> lk(g_idle_mu);
    for (auto it = g_idle.lower_bound(bytes); it != g_idle.end() && it->first <= 2 * bytes; ++it) {
        if (it->second.second != dev) continue;
        const std::pair<void *, size_t> r(it->second.first, it->first);
        g_idle.erase(it);
        return r;
    }
    return {nullptr, 0};
}

void free_idle(int dev)
{
    std::lock_guard<std::mutex> lk(g_idle_mu);
    for (auto it = g_idle.begin(); it != g_idle.end();) {
        if (it->second.second == dev) {
            (void)hipFree(it->second.first);
            it = g_idle.erase(it);
        } else {
            ++it;
        }
    }
}

}  // namespace

void *DevicePool::alloc(size_t bytes)
{
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~(size_t)255;
    // reuse a cached block of at least `bytes` and at most twice that
    auto it = cache_.lower_bound(bytes);
    if (it != cache_.end() && it->first <= 2 * bytes) {
        void *p = it->second;
        cache_.erase(it);
        live_ += sizes_[p];
        return p;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    const std::pair<void *, size_t> idle = take_idle(bytes, dev);
    if (idle.first) {
        sizes_[idle.first] = idle.second;
        devs_[idle.first] = dev;
        live_ += idle.second;
        return idle.first;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        // drop the caches and retry once
        for (auto &kv : cache_) {
            (void)hipFree(kv.second);
            sizes_.erase(kv.second);
            devs_.erase(kv.second);
        }
        cache_.clear();
        free_idle(dev);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    sizes_[p] = bytes;
    devs_[p] = dev;
    live_ += bytes;
    return p;
}

void DevicePool::free(void *p)
{
    if (!p) return;
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    live_ -= it->second;
    cache_.emplace(it->second, p);
}

void DevicePool::release()
{
    {
        std::lock_guard<std::mutex> lk(g_idle_mu);
        for (auto &kv : sizes_) g_idle.emplace(kv.second, std::make_pair(kv.first, devs_[kv.first]));
    }
    sizes_.clear();
    devs_.clear();
    cache_.clear();
    live_ = 0;
}

void DevicePool::free_all()
{
    for (auto &kv : sizes_) (void)hipFree(kv.first);
    sizes_.clear();
    devs_.clear();
    cache_.clear();
    live_ = 0;
}

void DevicePool::trim()
{
    for (auto &kv : cache_) {
        (void)hipFree(kv.second);
        sizes_.erase(kv.second);
        devs_.erase(kv.second);
    }
    cache_.clear();
    int dev = 0;
    (void)hipGetDevice(&dev);
    free_idle(dev);
}

size_t DevicePool::held_bytes() const
{
    size_t n = 0;
    for (auto &kv : sizes_) n += kv.second;
    return n;
}

int ensure(DevicePool &pool, DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes && b.p) return SG_OK;
    pool.free(b.p);
    size_t want = std::max(bytes, b.bytes + b.bytes / 2);
    b.p = pool.alloc(want);
    if (!b.p) {
        b.bytes = 0;
        return fail(SG_E_OOM, "scratch allocation failed");
    }
    b.bytes = want;
    return SG_OK;
}

}  // namespace sg

using namespace sg;

namespace {

bool valid_prop(const sg_grid *g, int p) { return g && p >= 0 && p < g->np; }

int alloc_chunk_arrays(sg_grid *g)
{
    // empty grid: the background chunk only
    g->cap = 0;
    g->nchunks = 0;
    g->d_keys = (int64_t *)g->pool.alloc(sizeof(int64_t));
    g->d_mask = (uint8_t *)g->pool.alloc(g->csz);
    if (!g->d_keys || !g->d_mask) return fail(SG_E_OOM, "allocation failed");
    SG_HIP(hipMemsetAsync(g->d_mask, 0, g->csz, g->stream));
    for (int p = 0; p < g->np; ++p) {
        g->d_prop[p] = g->pool.alloc((size_t)g->csz * g->ts[p]);
        if (!g->d_prop[p]) return fail(SG_E_OOM, "allocation failed");
        int rc = fill_background_chunk(g, p);
        if (rc) return rc;
    }
    return SG_OK;
}

}  // namespace

extern "C" {

const char *sg_last_error(void) { return t_err.c_str(); }

int sg_create(int dim, const int64_t *size, int block_edge, int nprops, const int32_t *dtypes, int stencil_radius,
              sg_grid **out)
{
    if (!out || !size) return fail(SG_E_ARG, "null argument");
    *out = nullptr;
    if (dim < 2 || dim > 3) return fail(SG_E_ARG, "dim must be 2 or 3");
    if (block_edge < 2) return fail(SG_E_ARG, "block_edge must be >= 2");
    int64_t csz = 1;
    for (int d = 0; d < dim; ++d) csz *= block_edge;
    if (csz > 1024) return fail(SG_E_RANGE, "block_edge^dim must be <= 1024");
    if (stencil_radius != 1) return fail(SG_E_ARG, "stencil_radius must be 1 (the reference's kernels assume 1)");
    if (nprops < 0 || nprops > kMaxProps) return fail(SG_E_ARG, "nprops must be in [0, 16]");
    if (nprops > 0 && !dtypes) return fail(SG_E_ARG, "null dtypes");
    sg_grid *g = new sg_grid();
    g->dim = dim;
    g->edge = block_edge;
    g->csz = (int)csz;
    g->r = stencil_radius;
    int64_t nblocks = 1;
    for (int d = 0; d < dim; ++d) {
        if (size[d] <= 0) {
            delete g;
            return fail(SG_E_ARG, "size must be positive");
        }
        g->sz[d] = size[d];
        g->bsz[d] = size[d] / block_edge + ((size[d] % block_edge) != 0);  // grid_smb.hpp:47-57
        nblocks *= g->bsz[d];
    }
    if (nblocks > ((int64_t)1 << 31) - 2) {
        delete g;
        return fail(SG_E_RANGE, "more than 2^31 chunks in the domain");
    }
    g->np = nprops;
    for (int p = 0; p < nprops; ++p) {
        g->dt[p] = dtypes[p];
        g->ts[p] = (int)dtype_size(dtypes[p]);
        if (!g->ts[p]) {
            delete g;
            return fail(SG_E_ARG, "unknown dtype");
        }
        memset(g->bg[p], 0, sizeof(g->bg[p]));
    }
    int rc = alloc_chunk_arrays(g);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return SG_OK;
}

int sg_destroy(sg_grid *g)
{
    if (!g) return SG_OK;
    (void)hipStreamSynchronize(g->stream);
    g->pool.release();
    delete g;
    return SG_OK;
}

int sg_set_stream(sg_grid *g, void *stream)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    // work already queued on the old stream must finish before the new stream runs
    SG_HIP(hipStreamSynchronize(g->stream));
    g->stream = (hipStream_t)stream;
    return SG_OK;
}

int sg_set_linearizer(sg_grid *g, int lin)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    if (lin != SG_LIN_SMB && lin != SG_LIN_ZMB) return fail(SG_E_ARG, "unknown linearizer");
    if (g->nchunks != 0 || g->ins_n != 0 || g->cins_n != 0 || g->pins_n != 0)
        return fail(SG_E_STATE, "the linearizer can only change on an empty grid");
    if (lin == SG_LIN_ZMB) {
        // Morton ids of the domain's chunks: the largest must leave room for point keys
        const int64_t limit_bits = 62;
        int64_t maxb = 0;
        for (int d = 0; d < g->dim; ++d) maxb = std::max(maxb, g->bsz[d] - 1);
        int bits = 0;
        while (bits < 63 && (((int64_t)1 << bits) <= maxb)) ++bits;
        int cbits = 0;
        while (((int64_t)1 << cbits) < g->csz) ++cbits;
        if (bits * g->dim + cbits > limit_bits) return fail(SG_E_RANGE, "domain too large for Morton chunk ids");
    }
    g->lin = lin;
    return SG_OK;
}

int sg_synchronize(sg_grid *g)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_set_background(sg_grid *g, int prop, const void *value)
{
    if (!valid_prop(g, prop) || !value) return fail(SG_E_ARG, "bad property or value");
    memcpy(g->bg[prop], value, g->ts[prop]);
    // background chunk + every slot without a point takes the new value
    int rc = normalize_prop(g, prop);
    if (rc) return rc;
    return fill_background_chunk(g, prop);
}

int sg_insert_points(sg_grid *g, int64_t n, const int64_t *keys, const uint8_t *masks, const void *const *values,
                     int mem)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    int rc = append_points(g, n, keys, masks, values, mem);
    if (rc == SG_OK && mem == SG_MEM_HOST) SG_HIP(hipStreamSynchronize(g->stream));
    return rc;
}

int sg_insert_coords(sg_grid *g, int64_t n, const int32_t *coords, const uint8_t *masks, const void *const *values,
                     int mem)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    int rc = append_coords(g, n, coords, masks, values, mem);
    if (rc == SG_OK && mem == SG_MEM_HOST) SG_HIP(hipStreamSynchronize(g->stream));
    return rc;
}

int sg_insert_chunks(sg_grid *g, int64_t n, const int64_t *chunk_ids, const uint8_t *masks,
                     const void *const *values, int mem)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    int rc = append_chunks(g, n, chunk_ids, masks, values, mem);
    if (rc == SG_OK && mem == SG_MEM_HOST) SG_HIP(hipStreamSynchronize(g->stream));
    return rc;
}

int64_t sg_pending_inserts(sg_grid *g)
{
    if (!g) return 0;
    // chunk records count as the point records they stand for
    if (g->cins_n > 0 && convert_chunk_records(g) != SG_OK) return -1;
    return g->ins_n + g->pins_n;
}

int sg_flush(sg_grid *g, const int32_t *ops)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    for (int p = 0; p < g->np; ++p) {
        int op = ops ? ops[p] : SG_OP_NONE;
        if (op < SG_OP_NONE || op > SG_OP_OR) return fail(SG_E_ARG, "unknown reduction");
        if (op == SG_OP_OR && dtype_is_float(g->dt[p])) return fail(SG_E_ARG, "bitwise or on a float property");
    }
    if (g->dins_open) {
        // records the caller's kernels wrote join the pending inserts first
        const int rc = insert_device_end(g);
        if (rc) return rc;
    }
    return flush_impl(g, ops);
}

int sg_find_neighbours(sg_grid *g, int nn_type)
{
    if (!g || (nn_type != SG_NN_STAR && nn_type != SG_NN_FULL)) return fail(SG_E_ARG, "bad grid or nn type");
    return find_neighbours_impl(g, nn_type);
}

int sg_set_nn_type(sg_grid *g, int nn_type)
{
    if (!g || (nn_type != SG_NN_STAR && nn_type != SG_NN_FULL)) return fail(SG_E_ARG, "bad grid or nn type");
    // the kernels always stage the ghost cells their own neighbourhood needs
    g->ghost_type = nn_type;
    return SG_OK;
}

int sg_tag_boundaries(sg_grid *g, int nn_type, int opt)
{
    if (!g || (nn_type != SG_NN_STAR && nn_type != SG_NN_FULL)) return fail(SG_E_ARG, "bad grid or nn type");
    if (opt != SG_NO_CALCULATE_EXISTING_POINTS && opt != SG_CALCULATE_EXISTING_POINTS)
        return fail(SG_E_ARG, "bad option");
    int rc = tag_boundaries_impl(g, nn_type);
    if (rc == SG_OK && opt == SG_CALCULATE_EXISTING_POINTS) rc = existing_points_impl(g);
    return rc;
}

int sg_existing_points(sg_grid *g, int64_t *n)
{
    if (!g || !n) return fail(SG_E_ARG, "null argument");
    if (!g->ep_valid) {
        int rc = existing_points_impl(g);
        if (rc) return rc;
    }
    *n = g->ep_n;
    return SG_OK;
}

int sg_get_existing_points(sg_grid *g, int64_t *out, int mem)
{
    int64_t n = 0;
    int rc = sg_existing_points(g, &n);
    if (rc) return rc;
    if (n == 0) return SG_OK;
    if (!out) return fail(SG_E_ARG, "null output");
    SG_HIP(hipMemcpyAsync(out, g->d_ep, (size_t)n * 8, mem == SG_MEM_DEVICE ? hipMemcpyDeviceToDevice
                                                                            : hipMemcpyDeviceToHost, g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_apply_stencil(sg_grid *g, const sg_stencil_desc *desc)
{
    if (!g || !desc) return fail(SG_E_ARG, "null argument");
    return apply_stencil_impl(g, desc);
}

int sg_set_tuning(sg_grid *g, int key, int64_t value)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    if (key == SG_TUNE_STENCIL_WG_PER_XCD) {
        if (value < 1 || value > 4096) return fail(SG_E_RANGE, "workgroups per XCD must be in [1, 4096]");
        g->fast_wg_per_xcd = value;
        return SG_OK;
    }
    if (key == SG_TUNE_TAG_CHUNKS_PER_WAVE) {
        if (value != 1 && value != 2 && value != 4) return fail(SG_E_RANGE, "chunks per wave must be 1, 2 or 4");
        g->tag_cpw = (int)value;
        return SG_OK;
    }
    if (key == SG_TUNE_STENCIL_VARIANT) {
        if (value < 0 || value > 8) return fail(SG_E_RANGE, "stencil variant must be in [0, 8]");
        g->fast_variant = (int)value;
        return SG_OK;
    }
    if (key == SG_TUNE_STENCIL_STORES) {
        if (value < 0 || value > 3) return fail(SG_E_RANGE, "store mode must be 0, 1, 2 or 3");
        g->fast_wmode = (int)value;
        return SG_OK;
    }
    if (key == SG_TUNE_BOX_PAIRS) {
        if (value != 0 && value != 1) return fail(SG_E_RANGE, "box pairing must be 0 or 1");
        g->box_pairs = (int)value;
        return SG_OK;
    }
    if (key == SG_TUNE_TAG_VARIANT) {
        if (value != 0 && value != 1) return fail(SG_E_RANGE, "tagBoundaries variant must be 0 or 1");
        g->tag_variant = (int)value;
        return SG_OK;
    }
    if (key == SG_TUNE_FLUSH_PATH) {
        if (value != 0 && value != 1) return fail(SG_E_RANGE, "flush path must be 0 or 1");
        g->flush_path = (int)value;
        return SG_OK;
    }
    if (key == SG_TUNE_STENCIL_SKIP_EMPTY_ROWS) {
        if (value < 0 || value > 4) return fail(SG_E_RANGE, "skip-empty-rows must be in [0, 4]");
        g->fast_lskip = (int)value;
        return SG_OK;
    }
    return fail(SG_E_ARG, "unknown tuning key");
}

int64_t sg_num_chunks(sg_grid *g) { return g ? g->nchunks : 0; }

int sg_chunk_size(sg_grid *g) { return g ? g->csz : 0; }

static hipMemcpyKind out_kind(int mem) { return mem == SG_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost; }

int sg_get_chunk_ids(sg_grid *g, int64_t *out, int mem)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    if (g->nchunks)
        SG_HIP(hipMemcpyAsync(out, g->d_keys, g->nchunks * sizeof(int64_t), out_kind(mem), g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_get_masks(sg_grid *g, uint8_t *out, int mem)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    if (g->nchunks) SG_HIP(hipMemcpyAsync(out, g->d_mask, g->nchunks * g->csz, out_kind(mem), g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_get_prop(sg_grid *g, int prop, void *out, int mem)
{
    if (!valid_prop(g, prop) || !out) return fail(SG_E_ARG, "bad property or output");
    if (g->nchunks)
        SG_HIP(hipMemcpyAsync(out, g->d_prop[prop], (size_t)g->nchunks * g->csz * g->ts[prop], out_kind(mem),
                              g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_set_prop(sg_grid *g, int prop, const void *in, int mem)
{
    if (!valid_prop(g, prop) || !in) return fail(SG_E_ARG, "bad property or input");
    if (g->nchunks) {
        SG_HIP(hipMemcpyAsync(g->d_prop[prop], in, (size_t)g->nchunks * g->csz * g->ts[prop],
                              mem == SG_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, g->stream));
        // slots without a point hold the background
        int rc = normalize_prop(g, prop);
        if (rc) return rc;
    }
    if (mem == SG_MEM_HOST) SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_set_chunks(sg_grid *g, int64_t n, const int64_t *ids, const uint8_t *masks, int mem)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    return set_chunks_impl(g, n, ids, masks, mem);
}

int sg_restore_background(sg_grid *g, int prop)
{
    if (!g || (prop != -1 && !valid_prop(g, prop))) return fail(SG_E_ARG, "bad property");
    return restore_background_impl(g, prop);
}

int sg_get_points(sg_grid *g, int prop, int64_t n, const int64_t *keys, void *out, int mem)
{
    if (!valid_prop(g, prop) || (n > 0 && (!keys || !out))) return fail(SG_E_ARG, "bad argument");
    return get_points_impl(g, prop, n, keys, out, mem);
}

int sg_get_neighbours(sg_grid *g, int nn_type, int32_t *out, int mem)
{
    if (!g || !out || (nn_type != SG_NN_STAR && nn_type != SG_NN_FULL)) return fail(SG_E_ARG, "bad argument");
    if (!g->nn_valid[nn_type]) {
        int rc = find_neighbours_impl(g, nn_type);
        if (rc) return rc;
    }
    const size_t bytes = (size_t)g->nchunks * g->nnn(nn_type) * sizeof(int32_t);
    if (bytes) SG_HIP(hipMemcpyAsync(out, g->d_nn[nn_type], bytes, out_kind(mem), g->stream));
    SG_HIP(hipStreamSynchronize(g->stream));
    return SG_OK;
}

int sg_device_ptr(sg_grid *g, int prop, void **out)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    if (prop == -1) {
        *out = g->d_mask;
        g->occ_tab_valid = false;  // the caller may write masks
        g->ebits_valid = false;
    }
    else if (valid_prop(g, prop)) *out = g->d_prop[prop];
    else return fail(SG_E_ARG, "bad property");
    return SG_OK;
}

int sg_insert_device_begin(sg_grid *g, int64_t nrec, sg_insert_view *out)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    return insert_device_begin(g, nrec, out);
}

int sg_insert_device_end(sg_grid *g)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    return insert_device_end(g);
}

int sg_to_kernel(sg_grid *g, int nn_type, sg_kernel_view *out)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    if (nn_type != SG_NN_STAR && nn_type != SG_NN_FULL) return fail(SG_E_ARG, "bad nn type");
    if (!g->nn_valid[nn_type]) {
        const int rc = find_neighbours_impl(g, nn_type);
        if (rc) return rc;
    }
    memset(out, 0, sizeof(*out));
    g->occ_tab_valid = false;  // the caller's kernels may write masks
    g->ebits_valid = false;
    out->keys = g->d_keys;
    out->mask = g->d_mask;
    for (int p = 0; p < g->np; ++p) out->prop[p] = g->d_prop[p];
    out->nn = g->d_nn[nn_type];
    out->stream = (void *)g->stream;
    out->nchunks = g->nchunks;
    for (int d = 0; d < 3; ++d) {
        out->size[d] = d < g->dim ? g->sz[d] : 1;
        out->bsz[d] = d < g->dim ? g->bsz[d] : 1;
    }
    out->nnn = g->nnn(nn_type);
    out->dim = g->dim;
    out->edge = g->edge;
    out->chunk_size = g->csz;
    out->zorder = g->lin == SG_LIN_ZMB;
    out->nprops = g->np;
    return SG_OK;
}

int sg_count_existing(sg_grid *g, int64_t *out)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    return count_mask_impl(g, 0, out);
}

int sg_count_boundary(sg_grid *g, int64_t *out)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    return count_mask_impl(g, 1, out);
}

int sg_trim(sg_grid *g)
{
    if (!g) return fail(SG_E_ARG, "null grid");
    SG_HIP(hipStreamSynchronize(g->stream));
    if (g->ins_n == 0 && g->ins_cap > 0) {
        g->pool.free(g->d_ins_key);
        g->pool.free(g->d_ins_mask);
        g->d_ins_key = nullptr;
        g->d_ins_mask = nullptr;
        for (int p = 0; p < g->np; ++p) {
            g->pool.free(g->d_ins_val[p]);
            g->d_ins_val[p] = nullptr;
        }
        g->ins_cap = 0;
    }
    if (g->pins_n == 0 && g->pins_cap > 0) {
        g->pool.free(g->d_pins_ck);
        g->pool.free(g->d_pins_rec);
        g->d_pins_ck = nullptr;
        g->d_pins_rec = nullptr;
        g->pins_cap = 0;
    }
    if (g->cins_n == 0 && g->cins_cap > 0) {
        g->pool.free(g->d_cins_id);
        g->pool.free(g->d_cins_mask);
        g->d_cins_id = nullptr;
        g->d_cins_mask = nullptr;
        for (int p = 0; p < g->np; ++p) {
            g->pool.free(g->d_cins_val[p]);
            g->d_cins_val[p] = nullptr;
        }
        g->cins_cap = 0;
    }
    for (auto &b : g->scratch) {
        g->pool.free(b.p);
        b = DevBuf();
    }
    g->pool.free(g->hostlike.p);
    g->hostlike = DevBuf();
    g->pool.trim();
    return SG_OK;
}

int sg_device_bytes(sg_grid *g, int64_t *out)
{
    if (!g || !out) return fail(SG_E_ARG, "null argument");
    *out = (int64_t)g->pool.held_bytes();
    return SG_OK;
}

}  // extern "C"
