// lbm_batch_api.hip -- C ABI of the batched multi-domain solver.
//
// No reference counterpart: simuverse runs one lattice per simulator
// (simuverse/src/fluid/fluid_simulator.rs:26-133).  A batch holds many
// independent physics-mode domains of uneven sizes, each with its own tau,
// inlet velocity, body force and solids, in two shared population buffers;
// one persistent kernel launch steps all of them.  Every per-domain
// operation other than the step (reset, priming, host I/O, macro,
// reductions) runs the single-domain kernels on a view of that domain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lbm.h"
#include "lbm_internal.hpp"

using namespace lbm;

#define BATCH_TRY(expr)                                                                               \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return lbm_fail(LBM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

struct lbm_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    bool f16 = false;
    size_t elem = 4;
    float scale = 1.0f, inv_scale = 1.0f;
    std::vector<lbm_domain_desc> desc;
    std::vector<BatchDomainDev> dom;  // host copy of the device table
    std::vector<size_t> cell_off;     // first cell of each domain in mat
    size_t total_elems = 0;           // elements per population buffer
    size_t total_cells = 0;
    size_t total_words = 0;
    uint32_t max_cells = 0;
    void* buf[2] = {nullptr, nullptr};
    uint8_t* mat = nullptr;
    uint32_t* bits = nullptr;
    BatchDomainDev* ddom = nullptr;
    uint16_t* chunk_dom = nullptr;  // chunk -> domain for the one-step variant in use
    uint32_t nchunks = 0;
    uint16_t* tile_dom = nullptr;   // tile -> domain for a two- or three-step variant
    uint32_t ntiles = 0;
    float* tmp3 = nullptr;
    double* partial = nullptr;
    double* red = nullptr;
    int cur = 0;
    int variant = 0;
    int cus = 256;
    int graph_chunk = 100;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    PhysArgs args(int d) const { return dom[d].a; }
    void* view(int b, int d) const { return (char*)buf[b] + (size_t)dom[d].pop_off * elem; }
    // steps one launch of the chosen variant advances (1, 2 or 3)
    int steps() const { return batch_variant_steps(variant); }
    // the variant of single steps: the chosen one, or the fallback of a multi-step one
    int step1_variant() const { return steps() > 1 ? kBatchSingleStepFallback : variant; }
    uint32_t grid() const
    {
        const uint32_t g = (uint32_t)(kBatchVariants[step1_variant()].blocks_per_cu * cus);
        return nchunks < g ? (nchunks ? nchunks : 1) : g;
    }
};

namespace {

void batch_drop_graphs(lbm_batch* b)
{
    for (auto& g : b->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void batch_free(lbm_batch* b)
{
    batch_drop_graphs(b);
    for (void* p : {b->buf[0], b->buf[1], (void*)b->mat, (void*)b->bits, (void*)b->ddom, (void*)b->chunk_dom,
                    (void*)b->tile_dom, (void*)b->tmp3, (void*)b->partial, (void*)b->red})
        if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int batch_check(lbm_batch* b)
{
    if (!b) return lbm_fail(LBM_ERR_INVALID, "null batch");
    BATCH_TRY(hipSetDevice(b->dev));
    return LBM_OK;
}

int batch_domain_ok(lbm_batch* b, int d)
{
    if (d < 0 || d >= (int)b->desc.size()) return lbm_fail(LBM_ERR_INVALID, "no domain %d", d);
    return LBM_OK;
}

// Uploads a work-unit -> domain table.
int upload_table(lbm_batch* b, const std::vector<uint16_t>& table, uint16_t** dev)
{
    if (*dev) (void)hipFree(*dev);
    *dev = nullptr;
    BATCH_TRY(hipMalloc(dev, std::max<size_t>(table.size(), 1) * sizeof(uint16_t)));
    BATCH_TRY(hipMemcpyAsync(*dev, table.data(), table.size() * sizeof(uint16_t), hipMemcpyHostToDevice, b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

// Work tables of the current variant: chunks of block*vec cells for the
// one-step variant (the chosen one or the multi-step fallback), and tiles
// for a two- or three-step variant.
int build_chunks(lbm_batch* b)
{
    const BatchVariant& sv = kBatchVariants[b->step1_variant()];
    const uint32_t w = (uint32_t)(sv.block * sv.vec);
    std::vector<uint16_t> table, tiles;
    for (size_t d = 0; d < b->dom.size(); ++d) {
        BatchDomainDev& D = b->dom[d];
        D.chunk0 = (uint32_t)table.size();
        table.insert(table.end(), (D.ncells + w - 1) / w, (uint16_t)d);
        D.tile0 = (uint32_t)tiles.size();
        if (b->steps() > 1) {
            const BatchVariant& tv = kBatchVariants[b->variant];
            const uint32_t nt = (uint32_t)((D.a.nx + tv.vec - 1) / tv.vec) * (uint32_t)((D.a.ny + tv.tile_h - 1) / tv.tile_h);
            tiles.insert(tiles.end(), nt, (uint16_t)d);
        }
    }
    b->nchunks = (uint32_t)table.size();
    b->ntiles = (uint32_t)tiles.size();
    int rc = upload_table(b, table, &b->chunk_dom);
    if (rc == LBM_OK) rc = upload_table(b, tiles, &b->tile_dom);
    if (rc != LBM_OK) return rc;
    BATCH_TRY(hipMemcpyAsync(b->ddom, b->dom.data(), b->dom.size() * sizeof(BatchDomainDev), hipMemcpyHostToDevice,
                             b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

// Whether a variant can step every domain of the batch.
bool batch_variant_ok(const lbm_batch* b, int v)
{
    if (v < 0 || v >= kNumBatchVariants) return false;
    if (batch_variant_steps(v) == 1) return true;
    for (const lbm_domain_desc& q : b->desc)
        if (!two_step_fits(kBatchVariants[v].vec, q.nx, q.ny, q.x_edge != LBM_X_PERIODIC ? 1 : 0)) return false;
    return true;
}

// One step from buffer `swap` into buffer 1-swap.
hipError_t batch_enqueue_step(lbm_batch* b, int swap)
{
    return launch_batch_step(b->step1_variant(), b->f16, b->ddom, b->chunk_dom, b->nchunks, b->grid(),
                             b->buf[swap], b->buf[swap ^ 1], b->bits, b->stream);
}

// b->steps() steps from buffer `swap` into buffer 1-swap (multi-step variant).
hipError_t batch_enqueue_multi(lbm_batch* b, int swap)
{
    if (b->steps() == 3)
        return launch_batch_step3(b->variant, b->f16, b->ddom, b->tile_dom, b->ntiles, b->buf[swap],
                                  b->buf[swap ^ 1], b->bits, b->stream);
    return launch_batch_step2(b->variant, b->f16, b->ddom, b->tile_dom, b->ntiles, b->buf[swap], b->buf[swap ^ 1],
                              b->bits, b->stream);
}

// Buffer flips of one graph chunk: one per launch (its multi-step launches,
// then one-step launches for the steps left over).
int batch_chunk_flips(const lbm_batch* b)
{
    const int spl = b->steps();
    return (b->graph_chunk / spl + b->graph_chunk % spl) & 1;
}

int batch_ensure_graph(lbm_batch* b, int start)
{
    if (b->graph[start]) return LBM_OK;
    BATCH_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    // multi-step launches, then one-step launches for the steps left over
    const int spl = b->steps();
    const int multi = spl > 1 ? b->graph_chunk / spl : 0, singles = b->graph_chunk - multi * spl;
    int flips = 0;
    for (int s = 0; s < multi && le == hipSuccess; ++s, ++flips) le = batch_enqueue_multi(b, (start + flips) & 1);
    for (int s = 0; s < singles && le == hipSuccess; ++s, ++flips) le = batch_enqueue_step(b, (start + flips) & 1);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(b->stream, &g);
    if (le != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return lbm_fail(LBM_ERR_HIP, "batch graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ee));
    }
    hipError_t ie = hipGraphInstantiate(&b->graph[start], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        b->graph[start] = nullptr;
        return lbm_fail(LBM_ERR_HIP, "batch hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    return LBM_OK;
}

int batch_enqueue_run(lbm_batch* b, int nsteps)
{
    int left = nsteps;
    if (b->graph_chunk > 0) {
        while (left >= b->graph_chunk) {
            int rc = batch_ensure_graph(b, b->cur);
            if (rc != LBM_OK) return rc;
            BATCH_TRY(hipGraphLaunch(b->graph[b->cur], b->stream));
            left -= b->graph_chunk;
            b->cur ^= batch_chunk_flips(b);  // one flip per launch: even for one-step chunks
        }
    }
    const int spl = b->steps();
    if (spl > 1) {
        for (; left >= spl; left -= spl) {
            BATCH_TRY(batch_enqueue_multi(b, b->cur));
            b->cur ^= 1;
        }
    }
    for (int s = 0; s < left; ++s) {
        BATCH_TRY(batch_enqueue_step(b, b->cur));
        b->cur ^= 1;
    }
    return LBM_OK;
}

int batch_prime(lbm_batch* b, int buf, int d)
{
    BATCH_TRY(launch_phys_prime(b->f16, b->dom[d].x_edge, b->args(d), b->view(buf, d), b->mat + b->cell_off[d],
                                b->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create(const lbm_domain_desc* domains, int count, int storage, lbm_batch** out)
{
    if (!out) return lbm_fail(LBM_ERR_INVALID, "null out");
    *out = nullptr;
    if (!domains || count < 1 || count > 65535) return lbm_fail(LBM_ERR_INVALID, "domain count must be 1..65535");
    if (storage != LBM_STORAGE_F32 && storage != LBM_STORAGE_F16_DEV)
        return lbm_fail(LBM_ERR_INVALID, "unknown storage %d", storage);
    uint64_t elems = 0, cells = 0, words = 0;
    for (int d = 0; d < count; ++d) {
        const lbm_domain_desc& q = domains[d];
        if (q.nx < 8 || q.ny < 3 || q.nx % 4 != 0)
            return lbm_fail(LBM_ERR_INVALID, "domain %d: nx must be a multiple of 4 >= 8 and ny >= 3", d);
        if (!(q.tau > 0.0f) || !std::isfinite(q.tau)) return lbm_fail(LBM_ERR_INVALID, "domain %d: bad tau", d);
        if (q.x_edge != LBM_X_PERIODIC && q.x_edge != LBM_X_INLET_OUTLET)
            return lbm_fail(LBM_ERR_INVALID, "domain %d: unknown x_edge", d);
        if (!std::isfinite(q.inlet_ux) || !(q.inlet_ux < 1.0f) || !std::isfinite(q.force_x) ||
            !std::isfinite(q.force_y))
            return lbm_fail(LBM_ERR_INVALID, "domain %d: bad inlet velocity or force", d);
        for (int r : q.reserved)
            if (r) return lbm_fail(LBM_ERR_INVALID, "domain %d: reserved fields must be zero", d);
        const uint64_t n = (uint64_t)q.nx * (uint64_t)q.ny;
        const uint64_t pitch = ((uint64_t)q.nx + 15u) & ~15ull;
        elems += (9 * pitch * (uint64_t)q.ny + 63u) & ~63ull;
        cells += n;
        words += (uint64_t)solid_mask_words(q.nx) * (uint64_t)q.ny;
    }
    if (elems > 0xFFFFFFFFull) return lbm_fail(LBM_ERR_INVALID, "batch exceeds 32-bit population indexing");
    int dev = 0;
    BATCH_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    BATCH_TRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return lbm_fail(LBM_ERR_NO_DEVICE, "current HIP device %d is not a gfx950", dev);

    lbm_batch* b = new lbm_batch();
    b->dev = dev;
    b->cus = prop.multiProcessorCount;
    b->f16 = storage == LBM_STORAGE_F16_DEV;
    b->elem = b->f16 ? 2 : 4;
    b->scale = std::ldexp(1.0f, 10);
    b->inv_scale = std::ldexp(1.0f, -10);
    b->desc.assign(domains, domains + count);
    b->total_elems = elems;
    b->total_cells = cells;
    b->total_words = words;
    uint64_t pop = 0, cell = 0, word = 0;
    for (int d = 0; d < count; ++d) {
        const lbm_domain_desc& q = domains[d];
        BatchDomainDev D{};
        const uint32_t n = (uint32_t)q.nx * (uint32_t)q.ny;
        D.a.nx = q.nx;
        D.a.ny = q.ny;
        // row-interleaved populations, row pitch nx rounded up to 16
        const uint32_t pitch = ((uint32_t)q.nx + 15u) & ~15u;
        D.a.plane = pitch;
        D.a.rs = 9u * pitch;
        D.a.omega = 1.0f / q.tau;
        D.a.fcoef = 1.0f - 0.5f * D.a.omega;
        D.a.gx = q.force_x;
        D.a.gy = q.force_y;
        D.a.inlet_ux = q.inlet_ux;
        D.a.scale = b->scale;
        D.a.inv_scale = b->inv_scale;
        set_force_terms(D.a);
        D.pop_off = (uint32_t)pop;
        D.bits_off = (uint32_t)word;
        D.words = solid_mask_words(q.nx);
        D.ncells = n;
        D.x_edge = q.x_edge;
        D.force = (q.force_x != 0.0f || q.force_y != 0.0f) ? 1 : 0;
        b->dom.push_back(D);
        b->cell_off.push_back((size_t)cell);
        pop += ((uint64_t)D.a.rs * (uint64_t)q.ny + 63u) & ~63ull;
        cell += n;
        word += (uint64_t)D.words * (uint64_t)q.ny;
        b->max_cells = std::max(b->max_cells, n);
    }
    auto alloc = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return lbm_fail(LBM_ERR_OOM, "hipMalloc(%zu) failed", bytes);
        }
        return LBM_OK;
    };
    int rc;
    if ((rc = alloc(&b->buf[0], b->total_elems * b->elem)) || (rc = alloc(&b->buf[1], b->total_elems * b->elem)) ||
        (rc = alloc((void**)&b->mat, b->total_cells)) ||
        (rc = alloc((void**)&b->bits, b->total_words * sizeof(uint32_t))) ||
        (rc = alloc((void**)&b->ddom, b->dom.size() * sizeof(BatchDomainDev))) ||
        (rc = alloc((void**)&b->tmp3, (size_t)3 * b->max_cells * sizeof(float))) ||
        (rc = alloc((void**)&b->partial, (size_t)kReduceBlocks * 5 * sizeof(double))) ||
        (rc = alloc((void**)&b->red, 5 * sizeof(double)))) {
        batch_free(b);
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipMemsetAsync(b->mat, LBM_MAT_BULK, b->total_cells, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->buf[0], 0, b->total_elems * b->elem, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->buf[1], 0, b->total_elems * b->elem, b->stream);
    if (e != hipSuccess) {
        batch_free(b);
        return lbm_fail(LBM_ERR_HIP, "batch setup: %s", hipGetErrorString(e));
    }
    for (int d = 0; d < count && rc == LBM_OK; ++d) {
        e = launch_build_solid_bits(b->mat + b->cell_off[d], b->dom[d].a.nx, b->dom[d].a.ny,
                                    b->dom[d].x_edge == LBM_X_PERIODIC, b->dom[d].words, b->bits + b->dom[d].bits_off,
                                    b->stream);
        if (e != hipSuccess) rc = lbm_fail(LBM_ERR_HIP, "mask: %s", hipGetErrorString(e));
    }
    if (rc == LBM_OK) rc = build_chunks(b);
    if (rc == LBM_OK) rc = lbm_batch_reset(b);
    if (rc != LBM_OK) {
        std::string msg = lbm_last_error();
        batch_free(b);
        return lbm_fail(rc, "%s", msg.c_str());
    }
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* b)
{
    if (!b) return LBM_OK;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    return LBM_OK;
}

int lbm_batch_cells(lbm_batch* b, uint64_t* cells)
{
    if (!b || !cells) return lbm_fail(LBM_ERR_INVALID, "null argument");
    *cells = b->total_cells;
    return LBM_OK;
}

int lbm_batch_write_materials(lbm_batch* b, int d, const uint8_t* material)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if ((rc = batch_domain_ok(b, d))) return rc;
    if (!material) return lbm_fail(LBM_ERR_INVALID, "null material");
    const BatchDomainDev& D = b->dom[d];
    BATCH_TRY(hipMemcpyAsync(b->mat + b->cell_off[d], material, D.ncells, hipMemcpyHostToDevice, b->stream));
    BATCH_TRY(launch_build_solid_bits(b->mat + b->cell_off[d], D.a.nx, D.a.ny, D.x_edge == LBM_X_PERIODIC, D.words,
                                      b->bits + D.bits_off, b->stream));
    if ((rc = batch_prime(b, b->cur, d))) return rc;
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_reset(lbm_batch* b)
{
    int rc = batch_check(b);
    if (rc) return rc;
    for (int d = 0; d < (int)b->dom.size(); ++d)
        for (int k = 0; k < 2; ++k) {
            BATCH_TRY(launch_phys_fill(b->f16, b->args(d), b->view(k, d), b->mat + b->cell_off[d], b->stream));
            if ((rc = batch_prime(b, k, d))) return rc;
        }
    b->cur = 0;
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_write_populations(lbm_batch* b, int d, const float* f)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if ((rc = batch_domain_ok(b, d))) return rc;
    if (!f) return lbm_fail(LBM_ERR_INVALID, "null populations");
    const size_t nb = (size_t)b->dom[d].ncells * sizeof(float);
    float* stage = nullptr;
    BATCH_TRY(hipMalloc(&stage, 9 * nb));
    hipError_t e = hipMemcpyAsync(stage, f, 9 * nb, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = launch_host_to_dev(b->f16, b->dom[d].a, stage, b->view(b->cur, d), b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "population upload: %s", hipGetErrorString(e));
    if ((rc = batch_prime(b, b->cur, d))) return rc;
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_read_populations(lbm_batch* b, int d, float* f)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if ((rc = batch_domain_ok(b, d))) return rc;
    if (!f) return lbm_fail(LBM_ERR_INVALID, "null populations");
    const size_t nb = (size_t)b->dom[d].ncells * sizeof(float);
    float* stage = nullptr;
    BATCH_TRY(hipMalloc(&stage, 9 * nb));
    hipError_t e = launch_dev_to_host(b->f16, b->dom[d].a, b->view(b->cur, d), stage, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(f, stage, 9 * nb, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "population download: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_batch_run(lbm_batch* b, int nsteps)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if (nsteps < 0) return lbm_fail(LBM_ERR_INVALID, "nsteps < 0");
    return batch_enqueue_run(b, nsteps);
}

int lbm_batch_run_timed(lbm_batch* b, int nsteps, double* ms)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if (!ms || nsteps < 0) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    BATCH_TRY(hipEventRecord(b->ev0, b->stream));
    if ((rc = batch_enqueue_run(b, nsteps))) return rc;
    BATCH_TRY(hipEventRecord(b->ev1, b->stream));
    BATCH_TRY(hipEventSynchronize(b->ev1));
    float t = 0.0f;
    BATCH_TRY(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = (double)t;
    return LBM_OK;
}

int lbm_batch_synchronize(lbm_batch* b)
{
    int rc = batch_check(b);
    if (rc) return rc;
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_macro(lbm_batch* b, int d, float* rho, float* ux, float* uy)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if ((rc = batch_domain_ok(b, d))) return rc;
    const uint32_t n = b->dom[d].ncells;
    float* r = b->tmp3;
    float* u = r + n;
    float* v = u + n;
    BATCH_TRY(launch_phys_macro(b->f16, b->dom[d].x_edge, b->dom[d].force != 0, b->args(d), b->view(b->cur, d),
                                b->mat + b->cell_off[d], r, u, v, b->stream));
    if (rho) BATCH_TRY(hipMemcpyAsync(rho, r, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (ux) BATCH_TRY(hipMemcpyAsync(ux, u, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (uy) BATCH_TRY(hipMemcpyAsync(uy, v, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_reduce(lbm_batch* b, int d, double out[5])
{
    int rc = batch_check(b);
    if (rc) return rc;
    if ((rc = batch_domain_ok(b, d))) return rc;
    if (!out) return lbm_fail(LBM_ERR_INVALID, "null output");
    BATCH_TRY(launch_phys_reduce(b->f16, b->dom[d].x_edge, b->args(d), b->view(b->cur, d), b->mat + b->cell_off[d],
                                 b->partial, kReduceBlocks, b->red, b->stream));
    BATCH_TRY(hipMemcpyAsync(out, b->red, 5 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_variant_count(void) { return kNumBatchVariants; }

int lbm_batch_variant_info(int variant, int* block, int* vec, int* tile_h, int* flags)
{
    if (variant < 0 || variant >= kNumBatchVariants) return lbm_fail(LBM_ERR_INVALID, "no batch variant %d", variant);
    const BatchVariant& v = kBatchVariants[variant];
    if (block) *block = v.block;
    if (vec) *vec = v.vec;
    if (tile_h) *tile_h = v.tile_h;
    if (flags) *flags = v.flags;
    return LBM_OK;
}

int lbm_batch_set_variant(lbm_batch* b, int variant)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if (variant < 0 || variant >= kNumBatchVariants) return lbm_fail(LBM_ERR_INVALID, "no batch variant %d", variant);
    if (!batch_variant_ok(b, variant))
        return lbm_fail(LBM_ERR_INVALID, "batch variant %d cannot step every domain (an inlet/outlet domain leaves "
                        "a one-column tile)", variant);
    BATCH_TRY(hipStreamSynchronize(b->stream));
    if (variant == b->variant) return LBM_OK;
    batch_drop_graphs(b);
    b->variant = variant;
    return build_chunks(b);
}

int lbm_batch_get_variant(lbm_batch* b, int* variant)
{
    if (!b || !variant) return lbm_fail(LBM_ERR_INVALID, "null argument");
    *variant = b->variant;
    return LBM_OK;
}

int lbm_batch_set_graph_chunk(lbm_batch* b, int chunk)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if (chunk < 0 || (chunk & 1)) return lbm_fail(LBM_ERR_INVALID, "graph chunk must be even and >= 0");
    BATCH_TRY(hipStreamSynchronize(b->stream));
    batch_drop_graphs(b);
    b->graph_chunk = chunk;
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch* b, int steps, int* best_variant, double* ms_per_step)
{
    int rc = batch_check(b);
    if (rc) return rc;
    if (steps < 1) return lbm_fail(LBM_ERR_INVALID, "steps must be >= 1");
    BATCH_TRY(hipStreamSynchronize(b->stream));
    void* saved = nullptr;
    const size_t bytes = b->total_elems * b->elem;
    BATCH_TRY(hipMalloc(&saved, bytes));
    const int cur0 = b->cur, v0 = b->variant;
    hipError_t e = hipMemcpyAsync(saved, b->buf[cur0], bytes, hipMemcpyDeviceToDevice, b->stream);
    int best = v0;
    double best_ms = 1e300;
    for (int v = 0; v < kNumBatchVariants && e == hipSuccess; ++v) {
        if (!batch_variant_ok(b, v)) continue;
        b->variant = v;
        if (build_chunks(b) != LBM_OK) {
            e = hipErrorUnknown;
            break;
        }
        const int per = b->steps();  // steps per launch
        auto launch = [&]() {
            hipError_t r = per > 1 ? batch_enqueue_multi(b, b->cur) : batch_enqueue_step(b, b->cur);
            b->cur ^= 1;
            return r;
        };
        for (int s = 0; s < 2 && e == hipSuccess; ++s) e = launch();
        if (e == hipSuccess) e = hipEventRecord(b->ev0, b->stream);
        const int launches = (steps + per - 1) / per;
        for (int s = 0; s < launches && e == hipSuccess; ++s) e = launch();
        if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, b->ev0, b->ev1);
        if (e == hipSuccess && t / (launches * per) < best_ms * kTuneMargin) {
            best_ms = t / (launches * per);
            best = v;
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(b->buf[cur0], saved, bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(saved);
    b->cur = cur0;
    batch_drop_graphs(b);
    b->variant = best;
    rc = build_chunks(b);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "batch autotune: %s", hipGetErrorString(e));
    if (rc != LBM_OK) return rc;
    if (best_variant) *best_variant = best;
    if (ms_per_step) *ms_per_step = best_ms;
    return LBM_OK;
}

}  // extern "C"
