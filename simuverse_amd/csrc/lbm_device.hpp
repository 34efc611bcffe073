// lbm_device.hpp -- device-side building blocks of the gfx950 D2Q9 kernels.
//
// Direction numbering and constants follow simuverse/src/fluid/mod.rs:22-52
// (the LbmUniform table the reference shaders read):
//     6 2 5
//     3 0 1
//     7 4 8
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

__host__ __device__ constexpr int ex_of(int i)
{
    return (i == 1 || i == 5 || i == 8) ? 1 : (i == 3 || i == 6 || i == 7) ? -1 : 0;
}
__host__ __device__ constexpr int ey_of(int i)
{
    return (i == 4 || i == 7 || i == 8) ? 1 : (i == 2 || i == 5 || i == 6) ? -1 : 0;
}
__host__ __device__ constexpr int inv_of(int i)
{
    return i == 0 ? 0 : i == 1 ? 3 : i == 2 ? 4 : i == 3 ? 1 : i == 4 ? 2 : i == 5 ? 7 : i == 6 ? 8 : i == 7 ? 5 : 6;
}

// Physics-mode weights 4/9, 1/9, 1/36 rounded once to binary32.
__host__ __device__ constexpr float w_phys(int i)
{
    return i == 0 ? 4.0f / 9.0f : (i <= 4 ? 1.0f / 9.0f : 1.0f / 36.0f);
}
// Parity-mode weights and post-collision caps: the reference's binary32
// literals (mod.rs:40-48).  Their sum is 0.9999988, not 1.
__host__ __device__ constexpr float w_wgsl(int i)
{
    return i == 0 ? 0.444444f : (i <= 4 ? 0.111111f : 0.0277777f);
}
__host__ __device__ constexpr float max_wgsl(int i)
{
    return i == 0 ? 0.6f : (i <= 4 ? 0.2222f : 0.1111f);
}

// Materials (fluid/lattice.rs:15-24, assets/wgsl/lbm/d2q9_fn.wgsl:19-24).
__host__ __device__ inline bool mat_solid(uint32_t m) { return m == 2u || m == 4u; }
__host__ __device__ inline bool mat_accel(uint32_t m) { return m == 3u || m == 6u; }

// Loads and stores with an optional non-temporal hint (streamed data that
// the next step reads back from HBM anyway).
template <bool NT, typename T>
__device__ inline T ld(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ inline void st(T* p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Step-kernel tuning flags.
enum : int {
    kNtStore = 1,  // non-temporal population stores
    kNtLoad = 2,   // non-temporal population loads
    kXcdRemap = 4  // give each XCD a contiguous range of rows
};

// Population storage of the physics mode.  Kernels work on deviations
// g_i = f_i - w_i from the rest state, and memory holds the deviation too, so
// the storage rounding is relative to g (~1e-4) rather than to f (~0.1):
//   float     stores f_i - w_i as binary32
//   _Float16  stores (f_i - w_i) * 2^shift as binary16
template <typename T>
struct Store;

template <>
struct Store<float> {
    static constexpr int bytes = 4;
    __device__ static inline float load_dev(const float* p, uint32_t k, float /*w*/, float /*inv_scale*/)
    {
        return p[k];
    }
    __device__ static inline void store_dev(float* p, uint32_t k, float g, float /*w*/, float /*scale*/)
    {
        p[k] = g;
    }
    __device__ static inline double load_f64(const float* p, uint32_t k, float w, float /*inv_scale*/)
    {
        return (double)w + (double)p[k];
    }
    __device__ static inline float to_storage(float g, float /*scale*/) { return g; }
    __device__ static inline float from_storage(float s, float /*inv_scale*/) { return s; }
    // Element already widened to binary32 <-> deviation.
    __device__ static inline float scale_in(float v, float /*inv_scale*/) { return v; }
    __device__ static inline float scale_out(float g, float /*scale*/) { return g; }
};

template <>
struct Store<_Float16> {
    static constexpr int bytes = 2;
    __device__ static inline float load_dev(const _Float16* p, uint32_t k, float /*w*/, float inv_scale)
    {
        return (float)p[k] * inv_scale;
    }
    __device__ static inline void store_dev(_Float16* p, uint32_t k, float g, float /*w*/, float scale)
    {
        p[k] = (_Float16)(g * scale);
    }
    __device__ static inline double load_f64(const _Float16* p, uint32_t k, float w, float inv_scale)
    {
        return (double)w + (double)((float)p[k] * inv_scale);
    }
    __device__ static inline _Float16 to_storage(float g, float scale) { return (_Float16)(g * scale); }
    __device__ static inline float from_storage(_Float16 s, float inv_scale) { return (float)s * inv_scale; }
    __device__ static inline float scale_in(float v, float inv_scale) { return v * inv_scale; }
    __device__ static inline float scale_out(float g, float scale) { return g * scale; }
};

// Arguments shared by the physics-mode kernels of one domain.
struct PhysArgs {
    int32_t nx, ny;
    uint32_t plane;   // elements between direction planes (>= nx*ny)
    float omega;      // 1/tau
    float fcoef;      // Guo prefactor 1 - omega/2
    float gx, gy;     // body force
    float inlet_ux;   // Zou-He inlet velocity (x_edge == 1)
    float scale;      // binary16 deviation scale 2^shift
    float inv_scale;  // 2^-shift
};

// One domain of a batch: its physics arguments and where its populations,
// solid mask and chunks start.
struct BatchDomainDev {
    PhysArgs a;
    uint32_t pop_off;   // element offset of the domain's plane 0 in a buffer
    uint32_t bits_off;  // word offset of the domain's solid mask
    uint32_t words;     // mask words per row
    uint32_t chunk0;    // index of the domain's first chunk
    uint32_t ncells;    // nx * ny
    int32_t x_edge;     // 0 periodic, 1 inlet/outlet
    int32_t force;      // non-zero when (gx, gy) != 0
    uint32_t pad;
};

// Arguments of the parity-mode (reference semantics) kernels.
struct WgslArgs {
    int32_t nx, ny;
    uint32_t plane;
    float omega;
};

// The learned collision correction's network 9 -> 64 -> 64 -> 9: offsets of
// W1[64][9], b1[64], W2[64][64], b2[64], W3[9][64], b3[9] in its flat
// parameter vector.
constexpr int kMlpIn = 9, kMlpH = 64, kMlpOut = 9;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpH * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH;
constexpr int kOffW3 = kOffB2 + kMlpH;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpH;
constexpr int kMlpParams = kOffB3 + kMlpOut;  // 5385

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.39894228040143268f;
}

// The network on one sample in binary32; P holds the parameter vector
// (in LDS in the kernels that call this).
__device__ inline void mlp_forward_one(const float* __restrict__ P, const float x[kMlpIn], float y[kMlpOut])
{
    float h1[kMlpH];
#pragma unroll 4
    for (int j = 0; j < kMlpH; ++j) {
        float s = P[kOffB1 + j];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) s += P[kOffW1 + j * kMlpIn + i] * x[i];
        h1[j] = gelu(s);
    }
#pragma unroll
    for (int m = 0; m < kMlpOut; ++m) y[m] = P[kOffB3 + m];
    for (int j = 0; j < kMlpH; ++j) {
        float s = P[kOffB2 + j];
#pragma unroll
        for (int k = 0; k < kMlpH; ++k) s += P[kOffW2 + j * kMlpH + k] * h1[k];
        const float h2 = gelu(s);
#pragma unroll
        for (int m = 0; m < kMlpOut; ++m) y[m] += P[kOffW3 + m * kMlpH + j] * h2;
    }
}

// Projection of a collision correction onto the non-conserved moments:
// d_i - w_i (a + 3 e_i.b) with a = sum d, b = sum e d (exact weights give
// sum w = 1, sum w e = 0, sum w e e = I/3), so mass and momentum of the
// result vanish.
__device__ inline void project_correction(float d[9])
{
    float a = 0.0f, bx = 0.0f, by = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        a += d[i];
        bx += (float)ex_of(i) * d[i];
        by += (float)ey_of(i) * d[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] -= w_phys(i) * (a + 3.0f * ((float)ex_of(i) * bx + (float)ey_of(i) * by));
}

// Linear block id -> logical tile id such that the blocks the hardware
// dispatches to one XCD (id mod 8) get a contiguous range of tiles.
__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t total)
{
    const uint32_t q = total >> 3, r = total & 7u;
    const uint32_t xcd = bid & 7u, local = bid >> 3;
    return (xcd < r) ? xcd * (q + 1u) + local : r * (q + 1u) + (xcd - r) * q + local;
}

}  // namespace lbm
