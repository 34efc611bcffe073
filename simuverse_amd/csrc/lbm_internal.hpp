// lbm_internal.hpp -- host-side launchers shared by the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_device.hpp"

// Records the message of a failure for lbm_last_error() and returns `code`
// (defined in lbm_api.hip; every source file of the C ABI reports through it).
int lbm_fail(int code, const char* fmt, ...);

namespace lbm {

struct StepVariant {
    int block;             // threads per block
    int cells_per_thread;  // cells handled by one thread
    int vec;               // consecutive cells per lane (1: strided by block)
    int flags;             // kNtStore | kNtLoad | kXcdRemap | kShfl | kTwoStep | kThreeStep | kRowStream
};
// kTwoStep and kThreeStep variants: cells_per_thread = tile (or chunk)
// rows, vec = tile (or strip) width; one launch advances two steps
// (launch_phys_step2) or three (launch_phys_step3), and the steps left over
// use the one-step variant kSingleStepFallback.
constexpr int kSingleStepFallback = 6;
// The autotuners keep the fastest shape, but a later shape in the table
// replaces the best so far only when it is at least 1 % faster: shapes of
// nearly equal speed would otherwise swap places between runs on timing
// noise.  The tables list the preferred of such twins first (the XCD-
// remapped 64 x 24 tile, which reads one pass, before the unmapped one).
constexpr double kTuneMargin = 0.99;
extern const StepVariant kStepVariants[];
extern const int kNumStepVariants;
bool step_variant_ok(int variant, int nx, int ny, int xedge);
inline bool step_variant_two_step(int variant) { return (kStepVariants[variant].flags & kTwoStep) != 0; }
// Steps one launch of a variant advances: 1, 2 (kTwoStep) or 3 (kThreeStep).
inline int step_variant_steps(int variant)
{
    const int f = kStepVariants[variant].flags;
    return (f & kThreeStep) ? 3 : ((f & kTwoStep) ? 2 : 1);
}
// Whether two-step tiles of width bw fit an nx x ny lattice (lbm_step2.hip).
bool two_step_fits(int bw, int nx, int ny, int xedge);
// Blocks of bs threads for n work items.
inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// Words per row of the packed solid mask (bit x+1 of a row is cell x).
inline uint32_t solid_mask_words(int nx) { return (((uint32_t)nx + 2u) / 32u + 2u + 1u) & ~1u; }
hipError_t launch_build_solid_bits(const uint8_t* mat, int nx, int ny, int periodic_x, uint32_t words,
                                  uint32_t* bits, hipStream_t s);
hipError_t launch_phys_step(int variant, bool f16, int xedge, bool force, bool export_macro, const PhysArgs& a,
                            const void* src, void* dst, const uint32_t* bits, uint32_t words, float* macro,
                            hipStream_t s);
hipError_t launch_phys_step2(int variant, bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                             void* dst, const uint32_t* bits, uint32_t words, hipStream_t s);
hipError_t launch_phys_step3(int variant, bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                             void* dst, const uint32_t* bits, uint32_t words, hipStream_t s);
hipError_t launch_phys_macro(bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                             const uint8_t* mat, float* rho, float* ux, float* uy, hipStream_t s);
hipError_t launch_phys_prime(bool f16, int xedge, const PhysArgs& a, void* buf, const uint8_t* mat,
                             hipStream_t s);
hipError_t launch_phys_fill(bool f16, const PhysArgs& a, void* buf, const uint8_t* mat, hipStream_t s);
hipError_t launch_phys_reduce(bool f16, int xedge, const PhysArgs& a, const void* src, const uint8_t* mat,
                              double* partial, int nblocks, double* out, hipStream_t s);

hipError_t launch_wgsl_step(bool export_macro, const WgslArgs& a, const float* src, float* dst, uint8_t* mat,
                            int32_t* biter, const float2* fvel, float* macro, hipStream_t s);
hipError_t launch_wgsl_macro(const WgslArgs& a, const float* src, const uint8_t* mat, const float2* fvel,
                             float* rho, float* ux, float* uy, hipStream_t s);
hipError_t launch_wgsl_reset(const WgslArgs& a, int fluid_ty, float* A, float* B, uint8_t* mat, int32_t* biter,
                             float2* fvel, float* macro, hipStream_t s);
hipError_t launch_wgsl_reduce(const WgslArgs& a, const float* src, const uint8_t* mat, double* partial,
                              int nblocks, double* out, hipStream_t s);

// Host-format populations (dense binary32 f[i*nx*ny + y*nx + x]) <->
// physics-mode device storage of deviations (binary32, or binary16 scaled
// by 2^shift) in the layout a.plane / a.rs.
hipError_t launch_host_to_dev(bool f16, const PhysArgs& a, const float* in, void* out, hipStream_t s);
hipError_t launch_dev_to_host(bool f16, const PhysArgs& a, const void* in, float* out, hipStream_t s);
hipError_t launch_pack_rgba16f(const float* macro, uint32_t n, uint16_t* out, hipStream_t s);
hipError_t launch_curl_rgba16f(const float* macro, int nx, int ny, uint16_t* out, hipStream_t s);
hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s);

// Batched step shapes.  One-step variants: a chunk is block*vec consecutive
// cells of one domain and a persistent grid of blocks_per_cu blocks per CU
// walks the chunks.  Two-step variants (kTwoStep): one block per tile of
// tile_h rows x vec cells of one domain, two steps per launch; odd step
// counts take the one-step variant kBatchSingleStepFallback.
struct BatchVariant {
    int block;           // threads per block
    int vec;             // consecutive cells per lane; tile width when two- or three-step
    int blocks_per_cu;   // persistent grid = blocks_per_cu * CUs (one-step)
    int flags;           // kNtStore | kXcdRemap | kGridStride | kTwoStep | kThreeStep
    int tile_h;          // tile rows (two- and three-step)
};
constexpr int kBatchSingleStepFallback = 9;
extern const BatchVariant kBatchVariants[];
extern const int kNumBatchVariants;
inline bool batch_variant_two_step(int variant) { return (kBatchVariants[variant].flags & kTwoStep) != 0; }
// Steps one launch of a batch variant advances: 1, 2 (kTwoStep) or 3 (kThreeStep).
inline int batch_variant_steps(int variant)
{
    const int f = kBatchVariants[variant].flags;
    return (f & kThreeStep) ? 3 : ((f & kTwoStep) ? 2 : 1);
}
hipError_t launch_batch_step(int variant, bool f16, const BatchDomainDev* doms, const uint16_t* chunk_dom,
                             uint32_t nchunks, uint32_t grid, const void* src, void* dst, const uint32_t* bits,
                             hipStream_t s);
// Two steps of every domain: tile_dom[t] is the domain of tile t, domain d
// owns tiles [tile0, tile0 + ceil(nx/vec) * ceil(ny/tile_h)).
hipError_t launch_batch_step2(int variant, bool f16, const BatchDomainDev* doms, const uint16_t* tile_dom,
                              uint32_t ntiles, const void* src, void* dst, const uint32_t* bits, hipStream_t s);
// Three steps of every domain, with the tiles of a three-step variant.
hipError_t launch_batch_step3(int variant, bool f16, const BatchDomainDev* doms, const uint16_t* tile_dom,
                              uint32_t ntiles, const void* src, void* dst, const uint32_t* bits, hipStream_t s);

// Learned collision, restriction and health check.
hipError_t launch_phys_step_learned(bool f16, int xedge, bool force, const PhysArgs& a, const void* src, void* dst,
                                   const uint32_t* bits, uint32_t words, const float* params, float s_in,
                                   float s_out, float* macro, hipStream_t s);
// packed: the network's lane-ordered weights, binary16 when op_f16, else bf16
hipError_t launch_phys_step_learned_mfma(bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                                        void* dst, const uint32_t* bits, uint32_t words, const void* packed,
                                        bool op_f16, const float* params, float s_in, float s_out, float* macro,
                                        hipStream_t s);
hipError_t launch_restrict4(bool f16, int xedge, bool force, const PhysArgs& a, const void* src, const uint8_t* mat,
                            float nscale, float gcx, float gcy, float* fc, uint8_t* solid_c, float* uxc, float* uyc,
                            hipStream_t s);
hipError_t launch_phys_health(bool f16, int xedge, bool force, const PhysArgs& a, const void* src, const uint8_t* mat,
                              uint32_t* out, hipStream_t s);

}  // namespace lbm

struct lbm_mlp;
// Device pointer to a network's binary32 parameters, and its scales.
const float* lbm_mlp_device_params(const lbm_mlp* m);
// Device pointer to its lane-ordered bf16 weights (kPackElems).
const __bf16* lbm_mlp_device_packed(const lbm_mlp* m);
// The same weights in binary16.
const _Float16* lbm_mlp_device_packed16(const lbm_mlp* m);
void lbm_mlp_scales(const lbm_mlp* m, float* s_in, float* s_out);
// Changes whenever the network's parameters or scales change.
uint64_t lbm_mlp_generation(const lbm_mlp* m);

namespace lbm {

// Number of blocks of the binary64 reduction (partials = 5 doubles each).
constexpr int kReduceBlocks = 1024;

}  // namespace lbm
