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

// Packed solid mask: row y holds bit x+1 for cell x; bit 0 and bit nx+1
// are the x-neighbours beyond the edges (the wrapped cells when x is
// periodic, empty otherwise).  `words` per row is even and covers bit
// nx+2, so the 64-bit window read below never leaves the row.
// Returns bits (x-1, x, x+1) of row y as bits 0, 1, 2.
__device__ inline uint32_t solid3(const uint32_t* __restrict__ bits, uint32_t words, int y, int x)
{
    const uint32_t* p = bits + (uint32_t)y * words + ((uint32_t)x >> 5);
    const uint64_t v = (uint64_t)p[0] | ((uint64_t)p[1] << 32);
    return (uint32_t)(v >> ((uint32_t)x & 31u)) & 7u;
}

// Sum over the 64 lanes of a wavefront (butterfly; every lane gets the sum).
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Step-kernel tuning flags.
enum : int {
    kNtStore = 1,  // non-temporal population stores
    kNtLoad = 2,   // non-temporal population loads
    kXcdRemap = 4, // give each XCD a contiguous range of rows
    kShfl = 8,     // aligned vector loads, x-neighbours from the adjacent lane
    kTwoStep = 16, // two steps per launch through an LDS tile (temporal blocking)
    kGridStride = 32, // batched: block b takes chunks b, b + grid, ... (else a contiguous range)
    kThreeStep = 64,  // three steps per launch through an LDS tile with a two-cell ring
    kRowStream = 128  // two-step: strips of block-size columns streamed down a chunk of rows
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

// Conversions of a kernel that works in the scaled binary16 domain (SC) or
// in plain deviations: SC keeps g * scale in registers, so loads and
// stores only convert the format.  The value is pinned in a register
// before the narrowing conversion (an empty asm barrier): otherwise the
// compiler may fold the collision's last fused multiply-add into the
// conversion (v_fma_mixlo_f16, one rounding), while the unscaled form
// rounds twice (to binary32, then to binary16) and the kernels would no
// longer agree bit for bit.
__device__ __forceinline__ float pin_f32(float v)
{
    asm volatile("" : "+v"(v));
    return v;
}

template <typename T, bool SC>
struct Dom {
    __device__ static inline float in(T v, float inv_scale)
    {
        if constexpr (SC) return (float)v;
        else return Store<T>::from_storage(v, inv_scale);
    }
    __device__ static inline T out(float g, float scale)
    {
        if constexpr (SC) return (T)pin_f32(g);
        else return Store<T>::to_storage(g, scale);
    }
    // element already widened to binary32
    __device__ static inline float in_w(float v, float inv_scale)
    {
        if constexpr (SC) return v;
        else return Store<T>::scale_in(v, inv_scale);
    }
    __device__ static inline float out_w(float g, float scale)
    {
        if constexpr (SC) return pin_f32(g);
        else return Store<T>::scale_out(g, scale);
    }
    __device__ static inline void push(T* p, uint32_t k, float g, float w, float scale)
    {
        if constexpr (SC) p[k] = (T)pin_f32(g);
        else Store<T>::store_dev(p, k, g, w, scale);
    }
};
// Whether a hot kernel keeps storage type T in the scaled domain.
template <typename T>
constexpr bool kScaledDom = Store<T>::bytes == 2;

// Arguments shared by the physics-mode kernels of one domain.  Population i
// of cell (x, y) is element y*rs + x + i*plane of a buffer:
//   row-interleaved  plane = pitch, rs = 9*pitch (the nine rows of one y are
//                    adjacent, so the 18 streams of a step walk 3 + 1
//                    contiguous regions of HBM instead of 18 far-apart ones)
//   planar           plane >= nx*ny, rs = nx
struct PhysArgs {
    int32_t nx, ny;
    uint32_t plane;   // elements between directions
    uint32_t rs;      // elements between rows
    float omega;      // 1/tau
    float fcoef;      // Guo prefactor 1 - omega/2
    float gx, gy;     // body force
    float inlet_ux;   // Zou-He inlet velocity (x_edge == 1)
    float scale;      // binary16 deviation scale 2^shift
    float inv_scale;  // 2^-shift
    // Guo forcing F_i = fcoef w_i (3 (e_i - u).g + 9 (e_i.u)(e_i.g))
    //                 = fa_i + fb_i (e_i.u) - fc_k (u.g)   (k: weight class)
    float fa[9];      // 3 fcoef w_i (e_i.g)
    float fb[9];      // 9 fcoef w_i (e_i.g)
    float fc[3];      // 3 fcoef w_k for the rest, axis and diagonal weights
    float rom;        // 1 - omega
    float ow[3];      // omega w_k
    // The same constants times `scale`, for kernels that keep binary16
    // deviations in their scaled form g * 2^shift through the step: every
    // product and sum of the collision is then the unscaled one times a
    // power of two, rounded identically, so both forms give the same bits.
    float sfa[9], sfb[9], sfc[3], sow[3];
};

// The Guo constants fa_i of the four direction pairs (1, 2, 5, 6; the
// scaled set when SC) held in vector registers.  The collision's
// D = ow_k d + fa_i has two wave-uniform operands, and a vector instruction
// reads at most one scalar register, so each cell would first copy fa_i
// into a vector register; pinned once per thread, the copies leave the
// per-cell code.  A tile kernel calls this on its own copy of the
// arguments.
template <bool SC>
__device__ inline void pin_pair_consts(PhysArgs& a)
{
    float* fa = SC ? a.sfa : a.fa;
    asm volatile("" : "+v"(fa[1]), "+v"(fa[2]), "+v"(fa[5]), "+v"(fa[6]));
}

// Fills fcoef and the Guo coefficient tables from omega, gx, gy.
__host__ __device__ inline void set_force_terms(PhysArgs& a)
{
    a.fcoef = 1.0f - 0.5f * a.omega;
    for (int i = 0; i < 9; ++i) {
        const float eg = (float)ex_of(i) * a.gx + (float)ey_of(i) * a.gy;
        a.fa[i] = 3.0f * a.fcoef * w_phys(i) * eg;
        a.fb[i] = 9.0f * a.fcoef * w_phys(i) * eg;
    }
    a.fc[0] = 3.0f * a.fcoef * w_phys(0);
    a.fc[1] = 3.0f * a.fcoef * w_phys(1);
    a.fc[2] = 3.0f * a.fcoef * w_phys(5);
    a.rom = 1.0f - a.omega;
    a.ow[0] = a.omega * w_phys(0);
    a.ow[1] = a.omega * w_phys(1);
    a.ow[2] = a.omega * w_phys(5);
    for (int i = 0; i < 9; ++i) {
        a.sfa[i] = a.fa[i] * a.scale;
        a.sfb[i] = a.fb[i] * a.scale;
    }
    for (int k = 0; k < 3; ++k) {
        a.sfc[k] = a.fc[k] * a.scale;
        a.sow[k] = a.ow[k] * a.scale;
    }
}

// Weight class of direction i: 0 rest, 1 axis, 2 diagonal.
__host__ __device__ constexpr int wclass(int i) { return i == 0 ? 0 : (i <= 4 ? 1 : 2); }

__host__ __device__ inline uint32_t pop_index(const PhysArgs& a, int x, int y, int i)
{
    return (uint32_t)y * a.rs + (uint32_t)x + (uint32_t)i * a.plane;
}

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
    uint32_t tile0;     // index of the domain's first two-step tile
};

// Moments of the post-streaming deviations (Guo velocity with FORCE).
// SC: g holds scaled deviations (g * scale; the sums are rescaled exactly).
template <bool FORCE, bool SC = false>
__device__ inline void phys_moments(const PhysArgs& a, const float g[9], float& rho, float& ux, float& uy)
{
    float dr = g[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) dr += g[i];
    const float jx = (g[1] - g[3]) + (g[5] - g[6]) + (g[8] - g[7]);
    const float jy = (g[4] - g[2]) + (g[7] - g[5]) + (g[8] - g[6]);
    if (SC) dr *= a.inv_scale;
    rho = 1.0f + dr;
    const float ir = __builtin_amdgcn_rcpf(rho);  // hardware reciprocal, 1 ulp
    // SC: the momentum sums carry the scale, which folds into the operation
    // that uses them (a product with a power of two is exact, so the
    // results are those of rescaling the sums first).
    if (FORCE) {
        ux = (SC ? fmaf(jx, a.inv_scale, 0.5f * a.gx) : jx + 0.5f * a.gx) * ir;
        uy = (SC ? fmaf(jy, a.inv_scale, 0.5f * a.gy) : jy + 0.5f * a.gy) * ir;
    } else {
        const float irs = SC ? ir * a.inv_scale : ir;
        ux = jx * irs;
        uy = jy * irs;
    }
}

// e_i.u for direction i (adds and negations only).
__device__ __forceinline__ float e_dot_u(int i, float ux, float uy)
{
    const int ex = ex_of(i), ey = ey_of(i);
    return ex == 0 ? (ey == 0 ? 0.0f : (ey > 0 ? uy : -uy))
                   : (ey == 0 ? (ex > 0 ? ux : -ux)
                              : (ex > 0 ? (ey > 0 ? ux + uy : ux - uy) : (ey > 0 ? uy - ux : -ux - uy)));
}

// BGK (+ Guo forcing) in deviation form:
//   g*_i = (1 - omega) g_i + omega geq_i + F_i
// with omega geq_i = omega w_k (rho q_i + rho - 1),
// q_i = e.u (3 + 4.5 e.u) - 1.5 u^2, and F_i = fa_i + fb_i (e_i.u) - fc_k (u.g)
// (PhysArgs).  Opposite directions i, inv(i) have opposite e.u, so with
//   s = 4.5 (e.u)^2 - 1.5 u^2,  c = rho s + rho - 1,  d = 3 rho e.u
// their rho q + rho - 1 are c + d and c - d, and their forcing terms
// A + fa_i and A - fa_i with A = fb_i (e_i.u) - fc_k (u.g) (fa and fb flip
// sign with e_i):
//   g*_i = (1 - omega) g_i + (S + D),  g*_inv(i) = (1 - omega) g_inv(i) + (S - D),
//   S = omega w_k c + A,  D = omega w_k d + fa_i,
// about eleven operations per pair of directions.
// SC: g holds scaled deviations and the result is scaled (the constants
// sow, sfa, sfb, sfc carry the scale; every sum and product is the unscaled
// one times a power of two).
template <bool FORCE, bool SC = false>
__device__ inline void phys_collide(const PhysArgs& a, float g[9], float rho, float ux, float uy)
{
    const float* const ow = SC ? a.sow : a.ow;
    const float* const fa = SC ? a.sfa : a.fa;
    const float* const fb = SC ? a.sfb : a.fb;
    const float* const fc = SC ? a.sfc : a.fc;
    const float dr = rho - 1.0f;
    const float usq = 1.5f * fmaf(ux, ux, uy * uy);
    const float r3 = 3.0f * rho;
    float fu[3] = {0.0f, 0.0f, 0.0f};
    if (FORCE) {
        const float ug = fmaf(ux, a.gx, uy * a.gy);
        fu[0] = -fc[0] * ug;
        fu[1] = -fc[1] * ug;
        fu[2] = -fc[2] * ug;
    }
    {
        // rest direction: e = 0, so rho q_0 + rho - 1 = dr - rho 1.5 u^2
        const float c0 = fmaf(-rho, usq, dr);
        const float S0 = FORCE ? fmaf(ow[0], c0, fu[0]) : ow[0] * c0;
        g[0] = fmaf(g[0], a.rom, S0);
    }
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        const int j = inv_of(i);
        if (j < i) continue;  // each pair once, from its lower index
        const int k = wclass(i);
        const float eu = e_dot_u(i, ux, uy);
        const float s = fmaf(4.5f * eu, eu, -usq);
        const float c = fmaf(rho, s, dr);
        const float d = r3 * eu;
        const float S = FORCE ? fmaf(ow[k], c, fmaf(fb[i], eu, fu[k])) : ow[k] * c;
        const float D = FORCE ? fmaf(ow[k], d, fa[i]) : ow[k] * d;
        g[i] = fmaf(g[i], a.rom, S + D);
        g[j] = fmaf(g[j], a.rom, S - D);
    }
}

// The learned correction's inputs and the moments it modulates, from the
// gathered deviations g and their moments rho, u:
//   q_i = g_i - geq_i = f_i - feq_i,
//   geq_i = w_i (rho - 1) + w_i rho (e_i.u (3 + 4.5 e_i.u) - 1.5 u^2),
//   x_i = q_i / (w_i s_in)                      (the network's input)
//   n_e  = sum m_e(i) q_i  = -4 q_0 - (q_1 + q_2 + q_3 + q_4) + 2 (q_5 + .. + q_8)
//   n_xx = sum m_xx(i) q_i = (q_1 + q_3) - (q_2 + q_4)
//   n_xy = sum m_xy(i) q_i = (q_6 + q_8) - (q_5 + q_7)
// (moment rows as in add_stress_correction).  inv_ws[k] = 1 / (w_k s_in) for
// the rest, axis and diagonal weight classes; two fused operations per
// direction for the equilibrium.
__host__ __device__ constexpr float inv_w_phys(int i) { return i == 0 ? 2.25f : (i <= 4 ? 9.0f : 36.0f); }
__device__ __forceinline__ void learned_inputs(const float g[9], float rho, float ux, float uy,
                                               const float inv_ws[3], float xin[9], float n[3])
{
    const float dr = rho - 1.0f;
    const float usq = 1.5f * fmaf(ux, ux, uy * uy);
    const float wr[3] = {w_phys(0) * rho, w_phys(1) * rho, w_phys(5) * rho};
    const float wd[3] = {w_phys(0) * dr, w_phys(1) * dr, w_phys(5) * dr};
    float q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int k = wclass(i);
        const float eu = e_dot_u(i, ux, uy);
        const float t = i == 0 ? -usq : fmaf(eu, fmaf(4.5f, eu, 3.0f), -usq);
        q[i] = g[i] - fmaf(wr[k], t, wd[k]);
        xin[i] = q[i] * inv_ws[k];
    }
    const float ax = q[1] + q[3], ay = q[2] + q[4];
    const float d57 = q[5] + q[7], d68 = q[6] + q[8];
    n[0] = fmaf(2.0f, d57 + d68, fmaf(-4.0f, q[0], -(ax + ay)));
    n[1] = ax - ay;
    n[2] = d68 - d57;
}
// inv_ws of learned_inputs for the scale s_in.
__host__ __device__ inline void learned_input_scales(float s_in, float inv_ws[3])
{
    const float inv_s = 1.0f / s_in;
    inv_ws[0] = inv_w_phys(0) * inv_s;
    inv_ws[1] = inv_w_phys(1) * inv_s;
    inv_ws[2] = inv_w_phys(5) * inv_s;
}

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

// ---------------------------------------------------------------------------
// The network on matrix cores: bf16 or binary16 operands, binary32
// accumulation.
//
// One wavefront evaluates 64 samples as two tiles of 32 with
// v_mfma_f32_32x32x16_bf16 (or _f16, the same layouts) on the transposed
// problem H = W X (M = units, N = samples, K = inputs).  Operand layouts of
// that instruction, lane l,
// h = l / 32, element j:
//   A (M x K)   row l % 32, k = 8 h + j            (j < 8)
//   B (K x N)   column l % 32, k = 8 h + j          (j < 8)
//   C (M x N)   column l % 32, row 8 (j / 4) + 4 h + (j % 4)   (j < 16)
// A layer's C tile therefore already sits in the lanes of the samples it
// belongs to; its 16 values per lane become the next layer's B operands for
// two K steps once the K axis of the next weight matrix is permuted to match
// (chain_unit).  No LDS round trip between layers.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Hidden unit that K index k (0..63) of a chained layer stands for: K step
// s = k / 16 reads C values j = 8 (s % 2) + k % 8 of M tile s / 2 from lane
// half h = (k / 8) % 2.
__host__ __device__ constexpr int chain_unit(int k)
{
    return 32 * ((k >> 4) >> 1) + 8 * ((8 * ((k >> 4) & 1) + (k & 7)) >> 2) + 4 * ((k >> 3) & 1) +
           ((8 * ((k >> 4) & 1) + (k & 7)) & 3);
}
// Row of C value j in lane half h of an M tile.
__host__ __device__ constexpr int c_row(int j, int h) { return 8 * (j >> 2) + 4 * h + (j & 3); }

// Lane-ordered weights in the operand type (8 per lane per operand, 16 B,
// read by one ds_read_b128):
//   W1p[mt][lane][8]      mt < 2, from W1 (K = inputs padded to 16)
//   W2p[mt][s][lane][8]   mt < 2, s < 4, from W2 with K permuted by chain_unit
//   W3p[s][lane][8]       s < 4, rows = outputs padded to 32
constexpr int kPackW1 = 0;
constexpr int kPackW2 = kPackW1 + 2 * 64 * 8;
constexpr int kPackW3 = kPackW2 + 2 * 4 * 64 * 8;
constexpr int kPackElems = kPackW3 + 4 * 64 * 8;  // 7168 elements = 14 KiB

// Packed weight element e (0 <= e < kPackElems) from the binary32 parameter
// vector P.
__host__ __device__ inline float pack_weight(const float* P, int e)
{
    if (e < kPackW2) {
        const int mt = e / 512, lane = (e / 8) % 64, j = e % 8;
        const int in = 8 * (lane >> 5) + j;
        return in < kMlpIn ? P[kOffW1 + (32 * mt + (lane & 31)) * kMlpIn + in] : 0.0f;
    }
    if (e < kPackW3) {
        const int r = e - kPackW2;
        const int mt = r / 2048, s = (r / 512) % 4, lane = (r / 8) % 64, j = r % 8;
        const int k = 16 * s + 8 * (lane >> 5) + j;
        return P[kOffW2 + (32 * mt + (lane & 31)) * kMlpH + chain_unit(k)];
    }
    const int r = e - kPackW3;
    const int s = r / 512, lane = (r / 8) % 64, j = r % 8;
    const int o = lane & 31, k = 16 * s + 8 * (lane >> 5) + j;
    return o < kMlpOut ? P[kOffW3 + o * kMlpH + chain_unit(k)] : 0.0f;
}

// GELU of two activations as x * Phi(x), Phi(x) = clamp(0.5 + x c(x^2), 0, 1),
// c a degree-4 polynomial in x^2 fitted to minimise the largest deviation
// from the erf form over [-10, 10].  Phi is evaluated in packed binary16
// (two activations per instruction: v_cvt_pkrtz_f16_f32, v_pk_mul_f16,
// five v_pk_fma_f16, the last with the [0, 1] clamp) and multiplies the
// binary32 x (v_fma_mix_f32): ten vector instructions per pair of
// activations, with the bf16 packing.  |GELU error| <= 5.2e-3 (at x = 3.35, where half a bf16 ulp
// of the activation is 7.8e-3), rms 9e-4 over |x| < 3.  The argument needs
// no clamp: checked over every finite binary16 x (binary16 coefficients,
// subnormal leading one included), Phi comes out exactly 1 for x >= 3.5
// and exactly 0 for x <= -3.5 -- x c(x^2) rises monotonically there, and
// overflow of x^2 or c gives +inf, which the clamp maps correctly.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f16x2 f16_splat(float v) { return (f16x2){(_Float16)v, (_Float16)v}; }
// Eight activations (four independent pairs, stage by stage, so dependent
// packed instructions are never adjacent) to bf16.
__device__ __forceinline__ bf16x8 gelu8_bf16(const float* v)
{
    f16x2 x[4], x2[4], c[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) x[p] = __builtin_bit_cast(f16x2, __builtin_amdgcn_cvt_pkrtz(v[2 * p], v[2 * p + 1]));
#pragma unroll
    for (int p = 0; p < 4; ++p) x2[p] = x[p] * x[p];
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(x2[p], f16_splat(1.02411032e-05f), f16_splat(-4.08110474e-04f));
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(c[p], x2[p], f16_splat(6.66348797e-03f));
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(c[p], x2[p], f16_splat(-5.97102903e-02f));
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(c[p], x2[p], f16_splat(3.94299256e-01f));
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        c[p] = __builtin_elementwise_fma(c[p], x[p], f16_splat(0.5f));
        c[p] = __builtin_elementwise_min(__builtin_elementwise_max(c[p], f16_splat(0.0f)), f16_splat(1.0f));
    }
    bf16x8 out;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        out[2 * p] = (__bf16)__builtin_fmaf(v[2 * p], (float)c[p].x, 0.0f);
        out[2 * p + 1] = (__bf16)__builtin_fmaf(v[2 * p + 1], (float)c[p].y, 0.0f);
    }
    return out;
}

// The same GELU for binary16 matrix operands: Phi as above, and the product
// x Phi(x) taken in packed binary16 too (v_pk_mul_f16), so a pair of
// activations costs eight vector instructions instead of ten (no binary32
// products, no bf16 packing).  x is the binary16 value rounded toward zero
// (11 significant bits against the 8 of a bf16 activation); beyond the
// binary16 range it saturates at 65504 instead of overflowing.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f16x8 gelu8_f16(const float* v)
{
    f16x2 x[4], x2[4], c[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) x[p] = __builtin_bit_cast(f16x2, __builtin_amdgcn_cvt_pkrtz(v[2 * p], v[2 * p + 1]));
#pragma unroll
    for (int p = 0; p < 4; ++p) x2[p] = x[p] * x[p];
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(x2[p], f16_splat(1.02411032e-05f), f16_splat(-4.08110474e-04f));
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(c[p], x2[p], f16_splat(6.66348797e-03f));
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(c[p], x2[p], f16_splat(-5.97102903e-02f));
#pragma unroll
    for (int p = 0; p < 4; ++p) c[p] = __builtin_elementwise_fma(c[p], x2[p], f16_splat(3.94299256e-01f));
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        c[p] = __builtin_elementwise_fma(c[p], x[p], f16_splat(0.5f));
        c[p] = __builtin_elementwise_min(__builtin_elementwise_max(c[p], f16_splat(0.0f)), f16_splat(1.0f));
    }
    f16x8 out;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f16x2 g = x[p] * c[p];
        out[2 * p] = g.x;
        out[2 * p + 1] = g.y;
    }
    return out;
}

// Operand element type E of the matrix-core network (bf16 or binary16):
// its 8-wide operand, matrix instruction and GELU.
template <typename E>
struct Mma;
template <>
struct Mma<__bf16> {
    typedef bf16x8 v8;
    __device__ static __forceinline__ f32x16 mfma(v8 a, v8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ v8 gelu8(const float* v) { return gelu8_bf16(v); }
};
template <>
struct Mma<_Float16> {
    typedef f16x8 v8;
    __device__ static __forceinline__ f32x16 mfma(v8 a, v8 b, f32x16 c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ v8 gelu8(const float* v) { return gelu8_f16(v); }
};

template <typename E>
__device__ __forceinline__ typename Mma<E>::v8 lds_b128(const E* p)
{
    return *reinterpret_cast<const typename Mma<E>::v8*>(p);
}

// LDS bias block of the matrix-core network: b1[64], b2[64], then b3 padded
// with zeros to the 32 rows of its M tile (16-byte aligned).
constexpr int kBiasElems = 2 * kMlpH + 32;

// The biases of one 32-row M tile as its initial accumulator: lane half h
// holds rows 8g + 4h .. 8g + 4h + 3 (c_row), four 16-byte LDS reads.  The
// row offset passes through an empty asm statement, so every tile reads its
// biases afresh straight into the accumulator registers instead of reusing
// one read and copying it (which cost a vector move per value).
__device__ __forceinline__ f32x16 bias_tile(const float* __restrict__ b, int h)
{
    int off = 4 * h;
    asm volatile("" : "+v"(off));
    f32x16 c;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const float4 v = *reinterpret_cast<const float4*>(b + off + 8 * g);
        c[4 * g] = v.x;
        c[4 * g + 1] = v.y;
        c[4 * g + 2] = v.z;
        c[4 * g + 3] = v.w;
    }
    return c;
}

// Network on the 64 samples of one wavefront.  W: packed weights (LDS);
// bias: b1[64], b2[64], b3[9] in binary32 (LDS); xs: 64 x 16 inputs (row per
// sample, inputs 9..15 zero); ys: 64 x 9 binary32 outputs; operands of
// element type E (bf16 or binary16), binary32 accumulation.  Every lane of
// the wavefront must call this (matrix instructions use all lanes).
template <typename E>
__device__ __forceinline__ void mlp_wave(const E* __restrict__ W, const float* __restrict__ bias,
                                         const E* __restrict__ xs, float* __restrict__ ys, int lane)
{
    typedef typename Mma<E>::v8 v8;
    const int h = lane >> 5, n = lane & 31;
    const float* b1 = bias;
    const float* b2 = bias + kMlpH;
    const float* b3 = bias + 2 * kMlpH;
    // The two sample tiles one after the other: fewer live accumulators,
    // and the other wavefronts of the SIMD fill the matrix-instruction
    // latency (2 % faster than interleaving them in one wavefront).
#pragma unroll 1
    for (int nt = 0; nt < 2; ++nt) {
        const v8 xb = lds_b128(xs + (32 * nt + n) * 16 + 8 * h);
        v8 hb[4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 c = bias_tile(b1 + 32 * mt, h);
            c = Mma<E>::mfma(lds_b128(W + kPackW1 + (mt * 64 + lane) * 8), xb, c);
{
                float v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) v[j] = c[j];
                hb[2 * mt] = Mma<E>::gelu8(v);
                hb[2 * mt + 1] = Mma<E>::gelu8(v + 8);
            }
        }
        v8 hb2[4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 c = bias_tile(b2 + 32 * mt, h);
#pragma unroll
            for (int s = 0; s < 4; ++s)
                c = Mma<E>::mfma(lds_b128(W + kPackW2 + ((mt * 4 + s) * 64 + lane) * 8), hb[s], c);
{
                float v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) v[j] = c[j];
                hb2[2 * mt] = Mma<E>::gelu8(v);
                hb2[2 * mt + 1] = Mma<E>::gelu8(v + 8);
            }
        }
        // b3 is stored padded with zeros to 32 rows
        f32x16 c = bias_tile(b3, h);
#pragma unroll
        for (int s = 0; s < 4; ++s)
            c = Mma<E>::mfma(lds_b128(W + kPackW3 + (s * 64 + lane) * 8), hb2[s], c);
        // Outputs 0..8 are rows 0-3 and 8 (lane half 0) and 4-7 (half 1):
        // both halves store c[0..3] at outputs 4h..4h+3 (the same register
        // indices in every lane, so no per-lane register selection).
        float* y = ys + (32 * nt + n) * kMlpOut + 4 * h;
        y[0] = c[0];
        y[1] = c[1];
        y[2] = c[2];
        y[3] = c[3];
        if (h == 0) y[8] = c[4];
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

// Projection of a collision correction onto the second-order non-conserved
// moments: the energy e and the deviatoric stresses p_xx, p_xy of the
// orthogonal D2Q9 moment basis (Lallemand & Luo, Phys. Rev. E 61 (2000)
// 6546), rows
//   e_i    = -4 (rest), -1 (axis), 2 (diagonal)       |e|^2    = 36
//   pxx_i  = ex^2 - ey^2                              |pxx|^2  = 4
//   pxy_i  = ex ey                                    |pxy|^2  = 4
// The rows are orthogonal to each other and to 1, ex, ey, so the result
// carries no mass and no momentum, and the higher (ghost) moments that do
// not enter the momentum flux are left to BGK.
__host__ __device__ constexpr float mom_e(int i) { return i == 0 ? -4.0f : (i <= 4 ? -1.0f : 2.0f); }
__device__ inline void project_stress(float d[9])
{
    float me = 0.0f, mxx = 0.0f, mxy = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float ex = (float)ex_of(i), ey = (float)ey_of(i);
        me += mom_e(i) * d[i];
        mxx += (ex * ex - ey * ey) * d[i];
        mxy += (ex * ey) * d[i];
    }
    me *= 1.0f / 36.0f;
    mxx *= 0.25f;
    mxy *= 0.25f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float ex = (float)ex_of(i), ey = (float)ey_of(i);
        d[i] = mom_e(i) * me + (ex * ex - ey * ey) * mxx + (ex * ey) * mxy;
    }
}

// Bounds of the learned modulation alpha of a moment's non-equilibrium part
// (learned_correction): BGK leaves (1 - omega) n_a after the collision and
// the correction adds alpha n_a, so |1 - omega + alpha| < 1 needs
// omega - 2 < alpha < omega.  The bound b = min((2 - omega)/2, omega/2)
// keeps half of the nearer margin on both sides: 1 - omega +- b lies in
// [-omega/2, 2 - 3 omega/2] for omega >= 1 and in [1 - 3 omega/2,
// 1 - omega/2] below, inside (-1, 1) for every omega in (0, 2).  The
// interval is symmetric, so clipped noise in the network's output does not
// shift the effective viscosity on average (near tau = 1/2 a small alpha
// changes it a lot).
__host__ __device__ inline float modulation_bound(float omega)
{
    return fminf(0.5f * (2.0f - omega), 0.5f * omega);
}
__host__ __device__ inline float modulation_lo(float omega) { return -modulation_bound(omega); }
__host__ __device__ inline float modulation_hi(float omega) { return modulation_bound(omega); }

// The learned collision correction of one cell.  The network's outputs
// 0, 1, 2 modulate the non-equilibrium parts n_a of the second-order
// non-conserved moments a = e, p_xx, p_xy of the cell (a learned,
// state-dependent change of how far BGK relaxes each of them):
//   dm_a = alpha_a n_a,   alpha_a = clamp(s_out y_a, modulation_lo, modulation_hi),
//   n_a = sum_i m_a(i) (f_i - feq_i),
// mapped back with the orthogonal basis (project_stress):
//   corr_i = sum_a m_a(i) dm_a / |m_a|^2,
// so the correction carries no mass, no momentum and no ghost moments.
// learned_inputs forms n_a before the network runs, so a kernel keeps three
// values, not nine, live across it.  With the rows written out:
//   corr_0 = -4 c_e,  corr_1,3 = c_xx - c_e,  corr_2,4 = -c_xx - c_e,
//   corr_5,7 = 2 c_e - c_xy,  corr_6,8 = 2 c_e + c_xy
// (c_e = alpha_e n_e / 36, c_xx = alpha_xx n_xx / 4, c_xy = alpha_xy n_xy / 4),
// added to g.  b is modulation_bound(omega).
__device__ __forceinline__ void add_stress_correction(const float n[3], const float y[3], float s_out, float b,
                                                      float g[9])
{
    const float ce = fminf(fmaxf(s_out * y[0], -b), b) * n[0] * (1.0f / 36.0f);
    const float cxx = fminf(fmaxf(s_out * y[1], -b), b) * n[1] * 0.25f;
    const float cxy = fminf(fmaxf(s_out * y[2], -b), b) * n[2] * 0.25f;
    const float ax = cxx - ce, ay = -cxx - ce;
    const float d57 = fmaf(2.0f, ce, -cxy), d68 = fmaf(2.0f, ce, cxy);
    g[0] = fmaf(-4.0f, ce, g[0]);
    g[1] += ax;
    g[3] += ax;
    g[2] += ay;
    g[4] += ay;
    g[5] += d57;
    g[7] += d57;
    g[6] += d68;
    g[8] += d68;
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
