// lbm_api.hip -- the C ABI declared in include/lbm.h (single-domain handle).
//
// One handle is what simuverse's D2Q9Node owns (simuverse/src/fluid/
// d2q9_node.rs:13-27): two population buffers, the per-cell lattice info,
// the macro field, plus a HIP stream, the binary64 reduction scratch and the
// HIP graphs of 100-step chunks.
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

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(LBM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

bool is_gfx950(int dev)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return false;
    return std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

// Records the message of a failure for lbm_last_error(); used by every
// source file of the library.
int lbm_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

struct lbm_handle {
    lbm_desc d{};
    int dev = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0, plane = 0, rs = 0;  // population strides (PhysArgs)
    size_t pop_elems = 0;               // elements per population buffer
    bool f16 = false;
    size_t elem = 4;
    void* buf[2] = {nullptr, nullptr};
    uint8_t* mat = nullptr;
    int32_t* biter = nullptr;
    float2* fvel = nullptr;
    float* macro = nullptr;  // planes ux, uy, rho, w (stride n)
    uint32_t* bits = nullptr;  // packed solid mask (physics mode)
    uint32_t words = 0;        // mask words per row
    float* tmp3 = nullptr;   // lbm_macro output planes
    uint16_t* texels = nullptr;  // rgba16float staging of the macro / curl readbacks
    double* partial = nullptr;
    double* red = nullptr;
    int cur = 0;
    float omega = 1.0f;
    float scale = 1.0f, inv_scale = 1.0f;
    int variant = 6;
    int graph_chunk = 100;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const lbm_mlp* mlp = nullptr;   // learned collision correction, if set
    uint64_t graph_mlp_gen = 0;     // the network's generation the graphs were captured with
    int mlp_precision = LBM_MLP_BF16;
    uint32_t* health = nullptr;     // [max |u|^2 bits, non-finite flag]
    void* backup = nullptr;         // chunk start state of lbm_run_guarded

    bool physics() const { return d.mode == LBM_MODE_PHYSICS; }
    bool force() const { return d.force_x != 0.0f || d.force_y != 0.0f; }
    size_t pop_bytes() const { return pop_elems * elem; }
    PhysArgs phys() const
    {
        PhysArgs a;
        a.nx = d.nx;
        a.ny = d.ny;
        a.plane = plane;
        a.rs = rs;
        a.omega = omega;
        a.fcoef = 1.0f - 0.5f * omega;
        a.gx = d.force_x;
        a.gy = d.force_y;
        a.inlet_ux = d.inlet_ux;
        a.scale = scale;
        a.inv_scale = inv_scale;
        set_force_terms(a);
        return a;
    }
    WgslArgs wgsl() const
    {
        WgslArgs a;
        a.nx = d.nx;
        a.ny = d.ny;
        a.plane = plane;
        a.omega = omega;
        return a;
    }
};

namespace {

void drop_graphs(lbm_handle* h)
{
    // a graph launched on the stream may still run: let it finish first
    if (h->graph[0] || h->graph[1]) (void)hipStreamSynchronize(h->stream);
    for (auto& g : h->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

int check_handle(lbm_handle* h)
{
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    HIP_TRY(hipSetDevice(h->dev));
    return LBM_OK;
}

// One step from buffer `swap` into buffer 1-swap (no synchronisation).
hipError_t enqueue_step(lbm_handle* h, int swap, bool export_macro)
{
    const void* src = h->buf[swap];
    void* dst = h->buf[swap ^ 1];
    if (h->physics() && h->mlp) {
        float s_in, s_out;
        lbm_mlp_scales(h->mlp, &s_in, &s_out);
        float* macro = export_macro ? h->macro : nullptr;
        if (h->mlp_precision == LBM_MLP_BF16)
            return launch_phys_step_learned_mfma(h->f16, h->d.x_edge, h->force(), h->phys(), src, dst, h->bits,
                                                 h->words, lbm_mlp_device_packed(h->mlp), false,
                                                 lbm_mlp_device_params(h->mlp), s_in, s_out, macro, h->stream);
        if (h->mlp_precision == LBM_MLP_F16)
            return launch_phys_step_learned_mfma(h->f16, h->d.x_edge, h->force(), h->phys(), src, dst, h->bits,
                                                 h->words, lbm_mlp_device_packed16(h->mlp), true,
                                                 lbm_mlp_device_params(h->mlp), s_in, s_out, macro, h->stream);
        return launch_phys_step_learned(h->f16, h->d.x_edge, h->force(), h->phys(), src, dst, h->bits, h->words,
                                        lbm_mlp_device_params(h->mlp), s_in, s_out, macro, h->stream);
    }
    if (h->physics()) {
        const int v = step_variant_steps(h->variant) > 1 ? kSingleStepFallback : h->variant;
        return launch_phys_step(v, h->f16, h->d.x_edge, h->force(), export_macro, h->phys(), src, dst, h->bits,
                                h->words, h->macro, h->stream);
    }
    return launch_wgsl_step(export_macro, h->wgsl(), (const float*)src, (float*)dst, h->mat, h->biter, h->fvel,
                            h->macro, h->stream);
}

// Steps one launch advances: 2 or 3 for a multi-step variant on a plain
// physics handle, else 1.
int launch_steps(const lbm_handle* h) { return h->physics() && !h->mlp ? step_variant_steps(h->variant) : 1; }

// launch_steps(h) steps from buffer `swap` into buffer 1-swap in one launch.
hipError_t enqueue_multi(lbm_handle* h, int swap)
{
    if (launch_steps(h) == 3)
        return launch_phys_step3(h->variant, h->f16, h->d.x_edge, h->force(), h->phys(), h->buf[swap],
                                 h->buf[swap ^ 1], h->bits, h->words, h->stream);
    return launch_phys_step2(h->variant, h->f16, h->d.x_edge, h->force(), h->phys(), h->buf[swap],
                             h->buf[swap ^ 1], h->bits, h->words, h->stream);
}

// Buffer flips of one graph chunk: one per launch (its multi-step launches,
// then one-step launches for the steps left over).
int chunk_flips(const lbm_handle* h)
{
    const int spl = launch_steps(h);
    return (h->graph_chunk / spl + h->graph_chunk % spl) & 1;
}

int ensure_graph(lbm_handle* h, int b)
{
    // graphs of learned steps hold the network's scales as launch arguments
    if (h->mlp && lbm_mlp_generation(h->mlp) != h->graph_mlp_gen) {
        drop_graphs(h);
        h->graph_mlp_gen = lbm_mlp_generation(h->mlp);
    }
    if (h->graph[b]) return LBM_OK;
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t launch_err = hipSuccess;
    // multi-step launches, then one-step launches for the steps left over
    const int spl = launch_steps(h);
    const int multi = spl > 1 ? h->graph_chunk / spl : 0, singles = h->graph_chunk - multi * spl;
    int flips = 0;
    for (int s = 0; s < multi && launch_err == hipSuccess; ++s, ++flips)
        launch_err = enqueue_multi(h, (b + flips) & 1);
    for (int s = 0; s < singles && launch_err == hipSuccess; ++s, ++flips)
        launch_err = enqueue_step(h, (b + flips) & 1, false);
    hipGraph_t g = nullptr;
    hipError_t end_err = hipStreamEndCapture(h->stream, &g);
    if (launch_err != hipSuccess || end_err != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(LBM_ERR_HIP, "graph capture failed: %s",
                    hipGetErrorString(launch_err != hipSuccess ? launch_err : end_err));
    }
    hipError_t inst = hipGraphInstantiate(&h->graph[b], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (inst != hipSuccess) {
        h->graph[b] = nullptr;
        return fail(LBM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(inst));
    }
    return LBM_OK;
}

// Enqueue nsteps steps from the current buffer (no synchronisation).
int enqueue_run(lbm_handle* h, int nsteps, bool export_macro)
{
    int plain = export_macro ? nsteps - 1 : nsteps;
    if (h->graph_chunk > 0) {
        while (plain >= h->graph_chunk) {
            int rc = ensure_graph(h, h->cur);
            if (rc != LBM_OK) return rc;
            HIP_TRY(hipGraphLaunch(h->graph[h->cur], h->stream));
            plain -= h->graph_chunk;
            h->cur ^= chunk_flips(h);  // one flip per launch: even for one-step chunks
        }
    }
    const int spl = launch_steps(h);
    if (spl > 1) {
        for (; plain >= spl; plain -= spl) {
            HIP_TRY(enqueue_multi(h, h->cur));
            h->cur ^= 1;
        }
    }
    for (int s = 0; s < plain; ++s) {
        HIP_TRY(enqueue_step(h, h->cur, false));
        h->cur ^= 1;
    }
    if (export_macro) {
        HIP_TRY(enqueue_step(h, h->cur, true));
        h->cur ^= 1;
    }
    return LBM_OK;
}

int reprime(lbm_handle* h, int b)
{
    if (!h->physics()) return LBM_OK;
    HIP_TRY(launch_phys_prime(h->f16, h->d.x_edge, h->phys(), h->buf[b], h->mat, h->stream));
    return LBM_OK;
}

// Physics mode: the packed solid mask the step kernel reads, rebuilt from
// the material bytes after every material write.
int rebuild_mask(lbm_handle* h)
{
    if (!h->physics()) return LBM_OK;
    HIP_TRY(launch_build_solid_bits(h->mat, h->d.nx, h->d.ny, h->d.x_edge == LBM_X_PERIODIC ? 1 : 0, h->words,
                                    h->bits, h->stream));
    return LBM_OK;
}

int do_reset(lbm_handle* h)
{
    if (h->physics()) {
        for (int b = 0; b < 2; ++b) {
            HIP_TRY(launch_phys_fill(h->f16, h->phys(), h->buf[b], h->mat, h->stream));
            int rc = reprime(h, b);
            if (rc != LBM_OK) return rc;
        }
        HIP_TRY(hipMemsetAsync(h->macro, 0, (size_t)4 * h->n * sizeof(float), h->stream));
    } else {
        HIP_TRY(launch_wgsl_reset(h->wgsl(), h->d.fluid_ty, (float*)h->buf[0], (float*)h->buf[1], h->mat,
                                  h->biter, h->fvel, h->macro, h->stream));
    }
    h->cur = 0;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int validate_desc(const lbm_desc* d)
{
    if (!d) return fail(LBM_ERR_INVALID, "null desc");
    if (d->nx < 2 || d->ny < 2) return fail(LBM_ERR_INVALID, "lattice %dx%d: both sides must be >= 2", d->nx, d->ny);
    const uint64_t n = (uint64_t)d->nx * (uint64_t)d->ny;
    const uint64_t pitch = ((uint64_t)d->nx + 15u) & ~15ull;
    if (n * 9 + 9 * 64 > 0xFFFFFFFFull || pitch * (uint64_t)d->ny * 9 > 0xFFFFFFFFull)
        return fail(LBM_ERR_INVALID, "lattice %dx%d exceeds 32-bit population indexing", d->nx, d->ny);
    if (d->layout != LBM_LAYOUT_ROWS && d->layout != LBM_LAYOUT_PLANES)
        return fail(LBM_ERR_INVALID, "unknown layout %d", d->layout);
    if (!(d->tau > 0.0f) || !std::isfinite(d->tau)) return fail(LBM_ERR_INVALID, "tau must be finite and > 0");
    if (d->mode != LBM_MODE_PARITY_WGSL && d->mode != LBM_MODE_PHYSICS)
        return fail(LBM_ERR_INVALID, "unknown mode %d", d->mode);
    if (d->storage != LBM_STORAGE_F32 && d->storage != LBM_STORAGE_F16_DEV)
        return fail(LBM_ERR_INVALID, "unknown storage %d", d->storage);
    for (int r : d->reserved)
        if (r != 0) return fail(LBM_ERR_INVALID, "reserved fields must be zero");
    if (d->mode == LBM_MODE_PARITY_WGSL) {
        if (d->storage != LBM_STORAGE_F32)
            return fail(LBM_ERR_UNSUPPORTED, "parity mode stores populations in binary32 only");
        if (d->x_edge != LBM_X_PERIODIC || d->force_x != 0.0f || d->force_y != 0.0f || d->inlet_ux != 0.0f)
            return fail(LBM_ERR_INVALID,
                        "parity mode has no x_edge, body force or inlet_ux (the reference drives its inlet "
                        "through accelerate-cell records)");
        if (d->fluid_ty != 0 && d->fluid_ty != 1) return fail(LBM_ERR_INVALID, "fluid_ty must be 0 or 1");
    } else {
        if (d->x_edge != LBM_X_PERIODIC && d->x_edge != LBM_X_INLET_OUTLET)
            return fail(LBM_ERR_INVALID, "unknown x_edge %d", d->x_edge);
        if (d->x_edge == LBM_X_INLET_OUTLET && d->nx < 3)
            return fail(LBM_ERR_INVALID, "an inlet/outlet lattice needs nx >= 3");
        if (!std::isfinite(d->force_x) || !std::isfinite(d->force_y) || !std::isfinite(d->inlet_ux))
            return fail(LBM_ERR_INVALID, "force and inlet velocity must be finite");
        if (d->x_edge == LBM_X_INLET_OUTLET && !(d->inlet_ux < 1.0f))
            return fail(LBM_ERR_INVALID, "Zou-He inlet needs inlet_ux < 1");
    }
    if (d->dev_shift < 0 || d->dev_shift > 14) return fail(LBM_ERR_INVALID, "dev_shift must be in 0..14");
    return LBM_OK;
}

void free_handle(lbm_handle* h)
{
    drop_graphs(h);
    for (void* p : {h->buf[0], h->buf[1], (void*)h->mat, (void*)h->biter, (void*)h->fvel, (void*)h->macro,
                    (void*)h->tmp3, (void*)h->texels, (void*)h->partial, (void*)h->red, (void*)h->bits,
                    (void*)h->health, h->backup})
        if (p) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// Header of a saved state: what must match for the device image behind it
// to mean the same populations in another handle.
struct StateHeader {
    char magic[8];  // "LBMSTATE"
    int32_t abi;
    int32_t nx, ny;
    int32_t mode, storage, layout;
    int32_t dev_shift;  // effective binary16 scale exponent (0 for binary32)
    uint32_t plane, rs;
    int32_t reserved[5];
};
static_assert(sizeof(StateHeader) == 64, "64-byte state header");

StateHeader state_header(const lbm_handle* h)
{
    StateHeader hd{};
    std::memcpy(hd.magic, "LBMSTATE", 8);
    hd.abi = LBM_ABI_VERSION;
    hd.nx = h->d.nx;
    hd.ny = h->d.ny;
    hd.mode = h->d.mode;
    hd.storage = h->d.storage;
    hd.layout = h->physics() ? h->d.layout : LBM_LAYOUT_PLANES;
    hd.dev_shift = h->f16 ? (int32_t)std::lround(std::log2(h->scale)) : 0;
    hd.plane = h->plane;
    hd.rs = h->rs;
    return hd;
}

int upload_info(lbm_handle* h, const uint8_t* m, const int32_t* bi, const float2* v, size_t offset, size_t count)
{
    HIP_TRY(hipMemcpyAsync(h->mat + offset, m, count, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->biter + offset, bi, count * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipMemcpyAsync(h->fvel + offset, v, count * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    int rc = rebuild_mask(h);
    if (rc != LBM_OK) return rc;
    rc = reprime(h, h->cur);
    if (rc != LBM_OK) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

const char* lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_device_count(int* count)
{
    if (!count) return fail(LBM_ERR_INVALID, "null count");
    *count = 0;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) return LBM_OK;
    if (e != hipSuccess) return fail(LBM_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    for (int i = 0; i < n; ++i)
        if (is_gfx950(i)) ++*count;
    return LBM_OK;
}

int lbm_create(const lbm_desc* desc, lbm_handle** out)
{
    if (!out) return fail(LBM_ERR_INVALID, "null out");
    *out = nullptr;
    int rc = validate_desc(desc);
    if (rc != LBM_OK) return rc;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (!is_gfx950(dev)) return fail(LBM_ERR_NO_DEVICE, "current HIP device %d is not a gfx950", dev);

    lbm_handle* h = new lbm_handle();
    h->d = *desc;
    h->dev = dev;
    h->n = (uint32_t)((uint64_t)desc->nx * (uint64_t)desc->ny);
    if (desc->mode == LBM_MODE_PHYSICS && desc->layout == LBM_LAYOUT_ROWS) {
        const uint32_t pitch = ((uint32_t)desc->nx + 15u) & ~15u;
        h->plane = pitch;
        h->rs = 9u * pitch;
        h->pop_elems = (size_t)h->rs * (size_t)desc->ny;
    } else {
        h->plane = (h->n + 63u) & ~63u;
        h->rs = (uint32_t)desc->nx;
        h->pop_elems = (size_t)9 * h->plane;
    }
    h->f16 = desc->storage == LBM_STORAGE_F16_DEV;
    h->elem = h->f16 ? 2 : 4;
    h->words = solid_mask_words(desc->nx);
    h->omega = 1.0f / desc->tau;
    const int shift = desc->dev_shift == 0 ? 10 : desc->dev_shift;
    h->scale = std::ldexp(1.0f, shift);
    h->inv_scale = std::ldexp(1.0f, -shift);

    auto alloc = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            *p = nullptr;
            return fail(LBM_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        }
        return LBM_OK;
    };
    const size_t n = h->n;
    if ((rc = alloc(&h->buf[0], h->pop_bytes())) || (rc = alloc(&h->buf[1], h->pop_bytes())) ||
        (rc = alloc((void**)&h->mat, n)) || (rc = alloc((void**)&h->biter, n * sizeof(int32_t))) ||
        (rc = alloc((void**)&h->fvel, n * sizeof(float2))) || (rc = alloc((void**)&h->macro, 4 * n * sizeof(float))) ||
        (rc = alloc((void**)&h->partial, (size_t)kReduceBlocks * 5 * sizeof(double))) ||
        (rc = alloc((void**)&h->red, 5 * sizeof(double))) ||
        (rc = alloc((void**)&h->bits, (size_t)h->words * (size_t)desc->ny * sizeof(uint32_t)))) {
        free_handle(h);
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess) e = hipMemsetAsync(h->mat, LBM_MAT_BULK, n, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->biter, 0xFF, n * sizeof(int32_t), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->fvel, 0, n * sizeof(float2), h->stream);
    // Solid slots that no fluid neighbour ever writes stay zero, as after
    // the reference's init pass.
    if (e == hipSuccess) e = hipMemsetAsync(h->buf[0], 0, h->pop_bytes(), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->buf[1], 0, h->pop_bytes(), h->stream);
    if (e == hipSuccess && rebuild_mask(h) != LBM_OK) e = hipErrorLaunchFailure;
    if (e != hipSuccess) {
        free_handle(h);
        return fail(LBM_ERR_HIP, "stream/event setup: %s", hipGetErrorString(e));
    }
    rc = do_reset(h);
    if (rc != LBM_OK) {
        std::string msg = g_last_error;
        free_handle(h);
        return fail(rc, "%s", msg.c_str());
    }
    *out = h;
    return LBM_OK;
}

int lbm_destroy(lbm_handle* h)
{
    if (!h) return LBM_OK;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_handle(h);
    return LBM_OK;
}

int lbm_write_lattice_info(lbm_handle* h, const lbm_lattice_info* info, size_t offset, size_t count)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!info && count) return fail(LBM_ERR_INVALID, "null info");
    if (offset > h->n || count > h->n - offset)
        return fail(LBM_ERR_INVALID, "cells [%zu, %zu) outside the lattice of %u", offset, offset + count, h->n);
    if (!count) return LBM_OK;
    std::vector<uint8_t> m(count);
    std::vector<int32_t> bi(count);
    std::vector<float2> v(count);
    for (size_t k = 0; k < count; ++k) {
        const int32_t mm = info[k].material;
        if (mm < 0 || mm > 255) return fail(LBM_ERR_INVALID, "material %d of cell %zu is not a byte", mm, offset + k);
        m[k] = (uint8_t)mm;
        bi[k] = info[k].block_iter;
        v[k] = make_float2(info[k].vx, info[k].vy);
    }
    return upload_info(h, m.data(), bi.data(), v.data(), offset, count);
}

int lbm_read_lattice_info(lbm_handle* h, lbm_lattice_info* info, size_t offset, size_t count)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!info && count) return fail(LBM_ERR_INVALID, "null info");
    if (offset > h->n || count > h->n - offset) return fail(LBM_ERR_INVALID, "cells outside the lattice");
    if (!count) return LBM_OK;
    std::vector<uint8_t> m(count);
    std::vector<int32_t> bi(count);
    std::vector<float2> v(count);
    HIP_TRY(hipMemcpyAsync(m.data(), h->mat + offset, count, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(bi.data(), h->biter + offset, count * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(v.data(), h->fvel + offset, count * sizeof(float2), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < count; ++k) info[k] = lbm_lattice_info{(int32_t)m[k], bi[k], v[k].x, v[k].y};
    return LBM_OK;
}

int lbm_write_materials(lbm_handle* h, const uint8_t* material, size_t offset, size_t count)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!material && count) return fail(LBM_ERR_INVALID, "null material");
    if (offset > h->n || count > h->n - offset) return fail(LBM_ERR_INVALID, "cells outside the lattice");
    if (!count) return LBM_OK;
    std::vector<int32_t> bi(count, -1);
    std::vector<float2> v(count, make_float2(0.0f, 0.0f));
    return upload_info(h, material, bi.data(), v.data(), offset, count);
}

int lbm_reset(lbm_handle* h)
{
    int rc = check_handle(h);
    if (rc) return rc;
    return do_reset(h);
}

int lbm_write_populations(lbm_handle* h, int buf, const float* f)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (buf != 0 && buf != 1) return fail(LBM_ERR_INVALID, "buffer must be 0 or 1");
    if (!f) return fail(LBM_ERR_INVALID, "null populations");
    const size_t nb = (size_t)h->n * sizeof(float);
    if (!h->physics()) {
        // Parity mode keeps the reference's buffer contents: f itself.
        HIP_TRY(hipMemcpy2DAsync(h->buf[buf], (size_t)h->plane * 4, f, nb, nb, 9, hipMemcpyHostToDevice, h->stream));
    } else {
        float* stage = nullptr;
        HIP_TRY(hipMalloc(&stage, 9 * nb));
        hipError_t e = hipMemcpyAsync(stage, f, 9 * nb, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = launch_host_to_dev(h->f16, h->phys(), stage, h->buf[buf], h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(stage);
        if (e != hipSuccess) return fail(LBM_ERR_HIP, "population upload: %s", hipGetErrorString(e));
    }
    rc = reprime(h, buf);
    if (rc != LBM_OK) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_read_populations(lbm_handle* h, int buf, float* f)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (buf != 0 && buf != 1) return fail(LBM_ERR_INVALID, "buffer must be 0 or 1");
    if (!f) return fail(LBM_ERR_INVALID, "null populations");
    const size_t nb = (size_t)h->n * sizeof(float);
    if (!h->physics()) {
        HIP_TRY(hipMemcpy2DAsync(f, nb, h->buf[buf], (size_t)h->plane * 4, nb, 9, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    } else {
        float* stage = nullptr;
        HIP_TRY(hipMalloc(&stage, 9 * nb));
        hipError_t e = launch_dev_to_host(h->f16, h->phys(), h->buf[buf], stage, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(f, stage, 9 * nb, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(stage);
        if (e != hipSuccess) return fail(LBM_ERR_HIP, "population download: %s", hipGetErrorString(e));
    }
    return LBM_OK;
}

int lbm_set_tau(lbm_handle* h, float tau)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!(tau > 0.0f) || !std::isfinite(tau)) return fail(LBM_ERR_INVALID, "tau must be finite and > 0");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->d.tau = tau;
    h->omega = 1.0f / tau;
    drop_graphs(h);
    return LBM_OK;
}

int lbm_step(lbm_handle* h, int swap, int export_macro)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (swap != 0 && swap != 1) return fail(LBM_ERR_INVALID, "swap must be 0 or 1");
    HIP_TRY(enqueue_step(h, swap, export_macro != 0));
    h->cur = swap ^ 1;
    return LBM_OK;
}

int lbm_run(lbm_handle* h, int nsteps, int export_macro)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (nsteps < 0) return fail(LBM_ERR_INVALID, "nsteps < 0");
    if (nsteps == 0) return LBM_OK;
    return enqueue_run(h, nsteps, export_macro != 0);
}

int lbm_current_buffer(lbm_handle* h, int* buf)
{
    if (!h || !buf) return fail(LBM_ERR_INVALID, "null argument");
    *buf = h->cur;
    return LBM_OK;
}

int lbm_synchronize(lbm_handle* h)
{
    int rc = check_handle(h);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_run_timed(lbm_handle* h, int nsteps, double* ms)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!ms || nsteps < 0) return fail(LBM_ERR_INVALID, "bad arguments");
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    rc = enqueue_run(h, nsteps, false);
    if (rc != LBM_OK) return rc;
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = (double)t;
    return LBM_OK;
}

int lbm_read_macro(lbm_handle* h, float* rho, float* ux, float* uy)
{
    int rc = check_handle(h);
    if (rc) return rc;
    const size_t nb = (size_t)h->n * sizeof(float);
    if (ux) HIP_TRY(hipMemcpyAsync(ux, h->macro, nb, hipMemcpyDeviceToHost, h->stream));
    if (uy) HIP_TRY(hipMemcpyAsync(uy, h->macro + h->n, nb, hipMemcpyDeviceToHost, h->stream));
    if (rho) HIP_TRY(hipMemcpyAsync(rho, h->macro + 2 * (size_t)h->n, nb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

// The texel staging buffer, allocated on first use and kept with the handle
// (the readbacks run every frame).
static int ensure_texels(lbm_handle* h)
{
    if (!h->texels) HIP_TRY(hipMalloc(&h->texels, (size_t)h->n * 8));
    return LBM_OK;
}

int lbm_read_macro_rgba16f(lbm_handle* h, uint16_t* rgba)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!rgba) return fail(LBM_ERR_INVALID, "null output");
    if ((rc = ensure_texels(h))) return rc;
    HIP_TRY(launch_pack_rgba16f(h->macro, h->n, h->texels, h->stream));
    HIP_TRY(hipMemcpyAsync(rgba, h->texels, (size_t)h->n * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_read_curl_rgba16f(lbm_handle* h, uint16_t* rgba)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!rgba) return fail(LBM_ERR_INVALID, "null output");
    if ((rc = ensure_texels(h))) return rc;
    HIP_TRY(launch_curl_rgba16f(h->macro, h->d.nx, h->d.ny, h->texels, h->stream));
    HIP_TRY(hipMemcpyAsync(rgba, h->texels, (size_t)h->n * 8, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_macro(lbm_handle* h, float* rho, float* ux, float* uy)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->tmp3) HIP_TRY(hipMalloc(&h->tmp3, (size_t)3 * h->n * sizeof(float)));
    float* r = h->tmp3;
    float* u = r + h->n;
    float* v = u + h->n;
    if (h->physics())
        HIP_TRY(launch_phys_macro(h->f16, h->d.x_edge, h->force(), h->phys(), h->buf[h->cur], h->mat, r, u, v, h->stream));
    else
        HIP_TRY(launch_wgsl_macro(h->wgsl(), (const float*)h->buf[h->cur], h->mat, h->fvel, r, u, v, h->stream));
    const size_t nb = (size_t)h->n * sizeof(float);
    if (rho) HIP_TRY(hipMemcpyAsync(rho, r, nb, hipMemcpyDeviceToHost, h->stream));
    if (ux) HIP_TRY(hipMemcpyAsync(ux, u, nb, hipMemcpyDeviceToHost, h->stream));
    if (uy) HIP_TRY(hipMemcpyAsync(uy, v, nb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_reduce(lbm_handle* h, double out[5])
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    if (h->physics())
        HIP_TRY(launch_phys_reduce(h->f16, h->d.x_edge, h->phys(), h->buf[h->cur], h->mat, h->partial, kReduceBlocks,
                                   h->red, h->stream));
    else
        HIP_TRY(launch_wgsl_reduce(h->wgsl(), (const float*)h->buf[h->cur], h->mat, h->partial, kReduceBlocks, h->red,
                                   h->stream));
    HIP_TRY(hipMemcpyAsync(out, h->red, 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

// ---- tuning and diagnostics ----------------------------------------------

int lbm_variant_count(void) { return kNumStepVariants; }

int lbm_variant_info(int variant, int* block, int* cells_per_thread, int* vec, int* flags)
{
    if (variant < 0 || variant >= kNumStepVariants) return fail(LBM_ERR_INVALID, "no variant %d", variant);
    if (block) *block = kStepVariants[variant].block;
    if (cells_per_thread) *cells_per_thread = kStepVariants[variant].cells_per_thread;
    if (vec) *vec = kStepVariants[variant].vec;
    if (flags) *flags = kStepVariants[variant].flags;
    return LBM_OK;
}

int lbm_set_variant(lbm_handle* h, int variant)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (variant < 0 || variant >= kNumStepVariants) return fail(LBM_ERR_INVALID, "no variant %d", variant);
    if (!step_variant_ok(variant, h->d.nx, h->d.ny, h->d.x_edge))
        return fail(LBM_ERR_INVALID, "variant %d does not fit a %dx%d lattice with x edge %d (see lbm_variant_info)",
                    variant, h->d.nx, h->d.ny, h->d.x_edge);
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (variant != h->variant) drop_graphs(h);
    h->variant = variant;
    return LBM_OK;
}

int lbm_get_variant(lbm_handle* h, int* variant)
{
    if (!h || !variant) return fail(LBM_ERR_INVALID, "null argument");
    *variant = h->variant;
    return LBM_OK;
}

int lbm_set_graph_chunk(lbm_handle* h, int chunk)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (chunk < 0 || (chunk & 1)) return fail(LBM_ERR_INVALID, "graph chunk must be even and >= 0");
    HIP_TRY(hipStreamSynchronize(h->stream));
    drop_graphs(h);
    h->graph_chunk = chunk;
    return LBM_OK;
}

int lbm_autotune(lbm_handle* h, int steps, int* best_variant, double* ms_per_step)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (steps < 1) return fail(LBM_ERR_INVALID, "steps must be >= 1");
    if (!h->physics()) {
        if (best_variant) *best_variant = h->variant;
        if (ms_per_step) *ms_per_step = 0.0;
        return LBM_OK;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    // Tuning steps advance the state; keep the current buffer and restore it.
    void* saved = nullptr;
    HIP_TRY(hipMalloc(&saved, h->pop_bytes()));
    const int cur0 = h->cur;
    hipError_t e = hipMemcpyAsync(saved, h->buf[cur0], h->pop_bytes(), hipMemcpyDeviceToDevice, h->stream);
    int best = h->variant;
    double best_ms = 1e300;
    for (int v = 0; v < kNumStepVariants && e == hipSuccess; ++v) {
        if (!step_variant_ok(v, h->d.nx, h->d.ny, h->d.x_edge)) continue;
        const int per = step_variant_steps(v);  // steps per launch
        auto launch = [&]() {
            hipError_t r = per == 3   ? launch_phys_step3(v, h->f16, h->d.x_edge, h->force(), h->phys(),
                                                          h->buf[h->cur], h->buf[h->cur ^ 1], h->bits, h->words,
                                                          h->stream)
                           : per == 2 ? launch_phys_step2(v, h->f16, h->d.x_edge, h->force(), h->phys(),
                                                          h->buf[h->cur], h->buf[h->cur ^ 1], h->bits, h->words,
                                                          h->stream)
                                      : launch_phys_step(v, h->f16, h->d.x_edge, h->force(), false, h->phys(),
                                                         h->buf[h->cur], h->buf[h->cur ^ 1], h->bits, h->words,
                                                         h->macro, h->stream);
            h->cur ^= 1;
            return r;
        };
        for (int s = 0; s < 2 && e == hipSuccess; ++s) e = launch();
        if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
        const int launches = (steps + per - 1) / per;
        for (int s = 0; s < launches && e == hipSuccess; ++s) e = launch();
        if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, h->ev0, h->ev1);
        if (e == hipSuccess && t / (launches * per) < best_ms * kTuneMargin) {
            best_ms = t / (launches * per);
            best = v;
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h->buf[cur0], saved, h->pop_bytes(), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(saved);
    h->cur = cur0;
    if (e != hipSuccess) return fail(LBM_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    if (best != h->variant) drop_graphs(h);
    h->variant = best;
    if (best_variant) *best_variant = best;
    if (ms_per_step) *ms_per_step = best_ms;
    return LBM_OK;
}

int lbm_state_size(lbm_handle* h, size_t* bytes)
{
    if (!h || !bytes) return fail(LBM_ERR_INVALID, "null argument");
    *bytes = sizeof(StateHeader) + h->pop_bytes();
    return LBM_OK;
}

int lbm_save_state(lbm_handle* h, void* out, size_t bytes)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!out || bytes != sizeof(StateHeader) + h->pop_bytes())
        return fail(LBM_ERR_INVALID, "state buffer must be %zu bytes", sizeof(StateHeader) + h->pop_bytes());
    const StateHeader hd = state_header(h);
    std::memcpy(out, &hd, sizeof hd);
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(out) + sizeof hd, h->buf[h->cur], h->pop_bytes(),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_load_state(lbm_handle* h, const void* in, size_t bytes)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!in || bytes != sizeof(StateHeader) + h->pop_bytes())
        return fail(LBM_ERR_INVALID, "state buffer must be %zu bytes", sizeof(StateHeader) + h->pop_bytes());
    StateHeader got;
    std::memcpy(&got, in, sizeof got);
    const StateHeader want = state_header(h);
    if (std::memcmp(got.magic, want.magic, sizeof want.magic) != 0 || got.abi != want.abi)
        return fail(LBM_ERR_INVALID, "not a state saved by this library (ABI %d)", LBM_ABI_VERSION);
    if (got.nx != want.nx || got.ny != want.ny || got.mode != want.mode || got.storage != want.storage ||
        got.layout != want.layout || got.dev_shift != want.dev_shift || got.plane != want.plane ||
        got.rs != want.rs)
        return fail(LBM_ERR_INVALID,
                    "state of a %dx%d lattice (mode %d, storage %d, layout %d, dev_shift %d) does not fit this "
                    "%dx%d handle (mode %d, storage %d, layout %d, dev_shift %d)",
                    got.nx, got.ny, got.mode, got.storage, got.layout, got.dev_shift, want.nx, want.ny, want.mode,
                    want.storage, want.layout, want.dev_shift);
    HIP_TRY(hipMemcpyAsync(h->buf[h->cur], static_cast<const char*>(in) + sizeof got, h->pop_bytes(),
                           hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

// ---- learned collision ------------------------------------------------------

int lbm_set_learned_collision(lbm_handle* h, const lbm_mlp* m)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (m && !h->physics()) return lbm_fail(LBM_ERR_UNSUPPORTED, "the learned correction needs physics mode");
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->mlp = m;
    drop_graphs(h);
    return LBM_OK;
}

int lbm_set_learned_precision(lbm_handle* h, int precision)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (precision != LBM_MLP_BF16 && precision != LBM_MLP_F32 && precision != LBM_MLP_F16)
        return lbm_fail(LBM_ERR_INVALID, "unknown network precision %d", precision);
    h->mlp_precision = precision;
    drop_graphs(h);
    return LBM_OK;
}

int lbm_health(lbm_handle* h, float* umax, int* finite)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->physics()) return lbm_fail(LBM_ERR_UNSUPPORTED, "health check needs physics mode");
    if (!h->health) HIP_TRY(hipMalloc(&h->health, 2 * sizeof(uint32_t)));
    HIP_TRY(hipMemsetAsync(h->health, 0, 2 * sizeof(uint32_t), h->stream));
    HIP_TRY(launch_phys_health(h->f16, h->d.x_edge, h->force(), h->phys(), h->buf[h->cur], h->mat, h->health,
                               h->stream));
    uint32_t out[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(out, h->health, sizeof out, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    float u2;
    std::memcpy(&u2, &out[0], sizeof u2);
    if (umax) *umax = std::sqrt(u2);
    if (finite) *finite = out[1] ? 0 : 1;
    return LBM_OK;
}

int lbm_run_guarded(lbm_handle* h, int nsteps, int chunk, float umax_limit, int* fallbacks)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (nsteps < 0 || chunk < 2 || (chunk & 1) || !(umax_limit > 0.0f))
        return lbm_fail(LBM_ERR_INVALID, "bad arguments (chunk must be even and >= 2)");
    if (!h->physics()) return lbm_fail(LBM_ERR_UNSUPPORTED, "guarded runs need physics mode");
    if (!h->backup) HIP_TRY(hipMalloc(&h->backup, h->pop_bytes()));
    int falls = 0;
    for (int done = 0; done < nsteps;) {
        const int len = std::min(chunk, nsteps - done);
        const int cur0 = h->cur;
        HIP_TRY(hipMemcpyAsync(h->backup, h->buf[cur0], h->pop_bytes(), hipMemcpyDeviceToDevice, h->stream));
        if ((rc = enqueue_run(h, len, false))) return rc;
        float umax = 0.0f;
        int finite = 0;
        if ((rc = lbm_health(h, &umax, &finite))) return rc;
        if (h->mlp && (!finite || !(umax <= umax_limit))) {
            // The chunk with the learned correction left the healthy range:
            // restart it from its first state with plain BGK.
            HIP_TRY(hipMemcpyAsync(h->buf[cur0], h->backup, h->pop_bytes(), hipMemcpyDeviceToDevice, h->stream));
            h->cur = cur0;
            const lbm_mlp* keep = h->mlp;
            h->mlp = nullptr;
            drop_graphs(h);
            rc = enqueue_run(h, len, false);
            h->mlp = keep;
            drop_graphs(h);
            if (rc) return rc;
            ++falls;
        }
        done += len;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (fallbacks) *fallbacks = falls;
    return LBM_OK;
}

int lbm_restrict4(lbm_handle* h, float tau_coarse, float force_x_coarse, float force_y_coarse, float* coarse_f,
                  uint8_t* coarse_solid, float* coarse_ux, float* coarse_uy)
{
    int rc = check_handle(h);
    if (rc) return rc;
    if (!h->physics()) return lbm_fail(LBM_ERR_UNSUPPORTED, "restriction needs physics mode");
    if (h->d.nx % 4 || h->d.ny % 4) return lbm_fail(LBM_ERR_INVALID, "restriction needs nx and ny divisible by 4");
    if (!coarse_f || !coarse_solid || !(tau_coarse > 0.5f)) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    const size_t nc = (size_t)(h->d.nx / 4) * (size_t)(h->d.ny / 4);
    float* df = nullptr;
    uint8_t* ds = nullptr;
    float* du = nullptr;
    HIP_TRY(hipMalloc(&df, 11 * nc * sizeof(float)));
    du = df + 9 * nc;
    hipError_t e = hipMalloc(&ds, nc);
    const float nscale = 4.0f * tau_coarse / h->d.tau;
    if (e == hipSuccess)
        e = launch_restrict4(h->f16, h->d.x_edge, h->force(), h->phys(), h->buf[h->cur], h->mat, nscale,
                             force_x_coarse, force_y_coarse, df, ds, du, du + nc, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(coarse_f, df, 9 * nc * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(coarse_solid, ds, nc, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && coarse_ux)
        e = hipMemcpyAsync(coarse_ux, du, nc * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && coarse_uy)
        e = hipMemcpyAsync(coarse_uy, du + nc, nc * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(df);
    if (ds) (void)hipFree(ds);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "restriction: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_copy_bandwidth(size_t bytes, int reps, double* ms_per_copy)
{
    if (!ms_per_copy || reps < 1 || bytes < 16) return fail(LBM_ERR_INVALID, "bad arguments");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (!is_gfx950(dev)) return fail(LBM_ERR_NO_DEVICE, "current HIP device %d is not a gfx950", dev);
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, s);
    if (e == hipSuccess) e = launch_copy16(a, b, bytes, s);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy16(a, b, bytes, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, "copy bandwidth: %s", hipGetErrorString(e));
    *ms_per_copy = (double)t / reps;
    return LBM_OK;
}

}  // extern "C"
