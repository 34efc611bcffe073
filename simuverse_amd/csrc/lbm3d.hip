// lbm3d.hip -- gfx950 D3Q19 solver and the C ABI declared in include/lbm3d.h.
//
// The 2-D physics design carried to three dimensions: one fused kernel per
// step that pulls f_i from x - e_i (periodic in x, y and z), collides (BGK
// + Guo force, deviation form g = f - w) and stores the cell's own 19 slots;
// bounce-back is a push of f*_i into slot inv(i) of each solid neighbour
// x + e_i, so the next gather needs no branch.  Each solid slot has one
// writer, so the step is race-free.
//
// Layout in HBM: row r = z*ny + y of the lattice holds its 19 direction
// rows back to back, element r*rs + i*pitch + x with pitch = nx rounded up
// to 16 and rs = 19*pitch: the 38 streams of a step collapse into the 9
// (y, z) rows a cell touches.  Algorithmic traffic: 19 x 4 B read + 19 x 4 B
// written = 152 B per cell per step in binary32 (76 B in binary16).  The
// solid mask is one bit per cell with a wrapped halo bit at each end of a
// row, nine 64-bit windows per cell (L2-resident: 2 MB at 256^3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lbm.h"
#include "lbm3d.h"
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {
namespace d3 {

// Direction table: 0 rest; 1-6 faces; 7-18 edges; opposite pairs adjacent.
__host__ __device__ constexpr int ex3(int i)
{
    constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    return t[i];
}
__host__ __device__ constexpr int ey3(int i)
{
    constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    return t[i];
}
__host__ __device__ constexpr int ez3(int i)
{
    constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    return t[i];
}
__host__ __device__ constexpr int inv3(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
// weight class: 0 rest (1/3), 1 face (1/18), 2 edge (1/36)
__host__ __device__ constexpr int wclass3(int i) { return i == 0 ? 0 : (i <= 6 ? 1 : 2); }
__host__ __device__ constexpr float w3(int i)
{
    return i == 0 ? 1.0f / 3.0f : (i <= 6 ? 1.0f / 18.0f : 1.0f / 36.0f);
}

struct Args3 {
    int32_t nx, ny, nz;
    uint32_t pitch;   // elements between the direction rows of one lattice row
    uint32_t rs;      // 19 * pitch: elements between lattice rows
    float omega;      // 1/tau
    float fcoef;      // 1 - omega/2
    float gx, gy, gz; // body force
    float scale, inv_scale;
    // Guo forcing F_i = fa_i + fb_i (e_i.u) - fc_k (u.g)
    float fa[19];     // 3 fcoef w_i (e_i.g)
    float fb[19];     // 9 fcoef w_i (e_i.g)
    float fc[3];      // 3 fcoef w_k
};

void set_force_terms3(Args3& a)
{
    a.fcoef = 1.0f - 0.5f * a.omega;
    for (int i = 0; i < 19; ++i) {
        const float eg = (float)ex3(i) * a.gx + (float)ey3(i) * a.gy + (float)ez3(i) * a.gz;
        a.fa[i] = 3.0f * a.fcoef * w3(i) * eg;
        a.fb[i] = 9.0f * a.fcoef * w3(i) * eg;
    }
    for (int k = 0; k < 3; ++k) a.fc[k] = 3.0f * a.fcoef * w3(k == 0 ? 0 : (k == 1 ? 1 : 7));
}

// c * v for a lattice component c in {-1, 0, 1}, without multiplying.
__device__ __forceinline__ float sgn_mul(int c, float v) { return c == 0 ? 0.0f : (c > 0 ? v : -v); }

template <bool FORCE>
__device__ inline void moments3(const Args3& a, const float g[19], float& rho, float& ux, float& uy, float& uz)
{
    float dr = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        dr += g[i];
        if (ex3(i) != 0) jx += sgn_mul(ex3(i), g[i]);
        if (ey3(i) != 0) jy += sgn_mul(ey3(i), g[i]);
        if (ez3(i) != 0) jz += sgn_mul(ez3(i), g[i]);
    }
    rho = 1.0f + dr;
    const float ir = __builtin_amdgcn_rcpf(rho);
    if (FORCE) {
        ux = (jx + 0.5f * a.gx) * ir;
        uy = (jy + 0.5f * a.gy) * ir;
        uz = (jz + 0.5f * a.gz) * ir;
    } else {
        ux = jx * ir;
        uy = jy * ir;
        uz = jz * ir;
    }
}

// BGK + Guo in deviation form, as the 2-D phys_collide.
template <bool FORCE>
__device__ inline void collide3(const Args3& a, float g[19], float rho, float ux, float uy, float uz)
{
    const float dr = rho - 1.0f;
    const float usq = 1.5f * fmaf(ux, ux, fmaf(uy, uy, uz * uz));
    const float wr[3] = {w3(0) * rho, w3(1) * rho, w3(7) * rho};
    const float wd[3] = {w3(0) * dr, w3(1) * dr, w3(7) * dr};
    const float ug = FORCE ? fmaf(ux, a.gx, fmaf(uy, a.gy, uz * a.gz)) : 0.0f;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int k = wclass3(i);
        float eu = 0.0f;
        if (ex3(i) != 0) eu = sgn_mul(ex3(i), ux);
        if (ey3(i) != 0) eu = ex3(i) == 0 ? sgn_mul(ey3(i), uy) : eu + sgn_mul(ey3(i), uy);
        if (ez3(i) != 0) eu = (ex3(i) == 0 && ey3(i) == 0) ? sgn_mul(ez3(i), uz) : eu + sgn_mul(ez3(i), uz);
        const float q = fmaf(eu, fmaf(4.5f, eu, 3.0f), -usq);
        const float geq = fmaf(wr[k], q, wd[k]);
        float out = fmaf(-a.omega, g[i] - geq, g[i]);
        if (FORCE) out += fmaf(a.fb[i], eu, fmaf(-a.fc[k], ug, a.fa[i]));
        g[i] = out;
    }
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// The nine (y, z) rows around a cell and their solid windows: row[dz+1][dy+1]
// is lattice row (z+dz)*ny + (y+dy) (wrapped), s[dz+1][dy+1] its solid bits
// of cells x-1, x, x+1 as bits 0, 1, 2.
struct Nbr3 {
    uint32_t row[3][3];
    uint32_t s[3][3];
};

__device__ __forceinline__ Nbr3 nbr3(const Args3& a, const uint32_t* __restrict__ bits, uint32_t words, int x, int y,
                                     int z)
{
    Nbr3 nb;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
            const uint32_t r = (uint32_t)wrap(z + dz, a.nz) * (uint32_t)a.ny + (uint32_t)wrap(y + dy, a.ny);
            nb.row[dz + 1][dy + 1] = r;
            nb.s[dz + 1][dy + 1] = solid3(bits, words, (int)r, x);
        }
    return nb;
}

// Element (lattice row, direction i, column c) of a population buffer.  In
// binary16 it is a row start plus a 32-bit column: the row start depends
// only on (y, z), which a block shares, so it stays in scalar registers and
// the kernel needs 66 VGPRs instead of 112 (7 wavefronts per SIMD instead of
// 4; 9 % faster at 256^3).  Binary32 keeps one 32-bit element index: there
// the extra wavefronts make the HBM-bound kernel 4 % slower.
template <typename T>
__device__ __forceinline__ T* elem3(T* base, const Args3& a, uint32_t row, int i, uint32_t c)
{
    if constexpr (sizeof(T) == 2) return base + ((size_t)row * a.rs + (size_t)i * a.pitch) + c;
    else return base + (row * a.rs + (uint32_t)i * a.pitch + c);
}

template <typename T, bool FORCE, int FLAGS>
__device__ inline void cell3(const Args3& a, const T* __restrict__ src, T* __restrict__ dst,
                             const uint32_t* __restrict__ bits, uint32_t words, int x, int y, int z)
{
    const Nbr3 nb = nbr3(a, bits, words, x, y, z);
    if (nb.s[1][1] & 2u) return;  // solid: its slots are written by pushes
    const uint32_t xm = x == 0 ? (uint32_t)a.nx - 1u : (uint32_t)x - 1u;
    const uint32_t xp = x == a.nx - 1 ? 0u : (uint32_t)x + 1u;
    float g[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int ex = ex3(i), ey = ey3(i), ez = ez3(i);
        const uint32_t sx = ex == 1 ? xm : (ex == -1 ? xp : (uint32_t)x);
        g[i] = Store<T>::from_storage(*elem3(src, a, nb.row[1 - ez][1 - ey], i, sx), a.inv_scale);
    }
    float rho, ux, uy, uz;
    moments3<FORCE>(a, g, rho, ux, uy, uz);
    collide3<FORCE>(a, g, rho, ux, uy, uz);
#pragma unroll
    for (int i = 0; i < 19; ++i)
        st<(FLAGS & kNtStore) != 0>(elem3(dst, a, nb.row[1][1], i, (uint32_t)x), Store<T>::to_storage(g[i], a.scale));
    uint32_t any = 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) any |= nb.s[r][c];
    if (any == 0u) return;
#pragma unroll
    for (int i = 1; i < 19; ++i) {
        const int ex = ex3(i), ey = ey3(i), ez = ez3(i);
        if (!((nb.s[1 + ez][1 + ey] >> (1 + ex)) & 1u)) continue;
        const uint32_t tx = ex == 1 ? xp : (ex == -1 ? xm : (uint32_t)x);
        Store<T>::store_dev(elem3(dst, a, nb.row[1 + ez][1 + ey], inv3(i), tx), 0u, g[i], w3(i), a.scale);
    }
}

// Block = BS threads on a segment of BS*CPT cells of one lattice row; thread
// t takes cells t, t + BS, ... (every load and store instruction of a
// wavefront covers 64 consecutive cells of one direction row).
template <typename T, bool FORCE, int BS, int CPT, int FLAGS>
__global__ __launch_bounds__(BS) void k3_step(Args3 a, const T* __restrict__ src, T* __restrict__ dst,
                                              const uint32_t* __restrict__ bits, uint32_t words, uint32_t nseg)
{
    // with kXcdRemap the blocks one XCD receives take a contiguous range of
    // rows, so the neighbour rows a block gathers from sit in the same L2
    const uint32_t b = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t row = b / nseg;
    const int y = (int)(row % (uint32_t)a.ny), z = (int)(row / (uint32_t)a.ny);
    const int x0 = (int)(b % nseg) * (BS * CPT) + (int)threadIdx.x;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int x = x0 + k * BS;
        if (x < a.nx) cell3<T, FORCE, FLAGS>(a, src, dst, bits, words, x, y, z);
    }
}

// After a host write: slot k of each solid cell S gets the bounce-back value
// its fluid neighbour S + e_k will pull, that neighbour's own f_inv(k).
template <typename T>
__global__ __launch_bounds__(256) void k3_prime(Args3 a, T* __restrict__ buf, const uint8_t* __restrict__ mat)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny * (uint32_t)a.nz;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n || !mat_solid(mat[c])) return;
    const int x = (int)(c % (uint32_t)a.nx);
    const uint32_t r = c / (uint32_t)a.nx;
    const int y = (int)(r % (uint32_t)a.ny), z = (int)(r / (uint32_t)a.ny);
    for (int k = 1; k < 19; ++k) {
        const int tx = wrap(x + ex3(k), a.nx), ty = wrap(y + ey3(k), a.ny), tz = wrap(z + ez3(k), a.nz);
        const uint32_t tr = (uint32_t)tz * (uint32_t)a.ny + (uint32_t)ty;
        if (mat_solid(mat[tr * (uint32_t)a.nx + (uint32_t)tx])) continue;
        buf[r * a.rs + (uint32_t)k * a.pitch + (uint32_t)x] = buf[tr * a.rs + (uint32_t)inv3(k) * a.pitch + (uint32_t)tx];
    }
}

// Planar host populations f <-> device deviations.
template <typename T>
__global__ __launch_bounds__(256) void k3_host_to_dev(Args3 a, const float* __restrict__ f, T* __restrict__ buf)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny * (uint32_t)a.nz;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const uint32_t r = c / (uint32_t)a.nx, x = c % (uint32_t)a.nx;
#pragma unroll
    for (int i = 0; i < 19; ++i)
        buf[r * a.rs + (uint32_t)i * a.pitch + x] = Store<T>::to_storage(f[(size_t)i * n + c] - w3(i), a.scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k3_dev_to_host(Args3 a, const T* __restrict__ buf, float* __restrict__ f)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny * (uint32_t)a.nz;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const uint32_t r = c / (uint32_t)a.nx, x = c % (uint32_t)a.nx;
#pragma unroll
    for (int i = 0; i < 19; ++i)
        f[(size_t)i * n + c] = w3(i) + Store<T>::from_storage(buf[r * a.rs + (uint32_t)i * a.pitch + x], a.inv_scale);
}

// Gathered populations of fluid cell (x, y, z) in binary64 (f = w + g).
template <typename T>
__device__ inline void gather_f64(const Args3& a, const T* __restrict__ src, int x, int y, int z, double f[19])
{
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int sx = wrap(x - ex3(i), a.nx), sy = wrap(y - ey3(i), a.ny), sz = wrap(z - ez3(i), a.nz);
        const uint32_t r = (uint32_t)sz * (uint32_t)a.ny + (uint32_t)sy;
        f[i] = Store<T>::load_f64(src, r * a.rs + (uint32_t)i * a.pitch + (uint32_t)sx, w3(i), a.inv_scale);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k3_macro(Args3 a, const T* __restrict__ src, const uint8_t* __restrict__ mat,
                                                float* __restrict__ out)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny * (uint32_t)a.nz;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    if (mat_solid(mat[c])) {
        out[c] = out[c + n] = out[c + 2 * n] = out[c + 3 * n] = 0.0f;
        return;
    }
    const int x = (int)(c % (uint32_t)a.nx);
    const uint32_t r = c / (uint32_t)a.nx;
    double f[19];
    gather_f64<T>(a, src, x, (int)(r % (uint32_t)a.ny), (int)(r / (uint32_t)a.ny), f);
    double rho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        rho += f[i];
        jx += (double)ex3(i) * f[i];
        jy += (double)ey3(i) * f[i];
        jz += (double)ez3(i) * f[i];
    }
    out[c] = (float)rho;
    out[c + n] = (float)((jx + 0.5 * (double)a.gx) / rho);
    out[c + 2 * n] = (float)((jy + 0.5 * (double)a.gy) / rho);
    out[c + 3 * n] = (float)((jz + 0.5 * (double)a.gz) / rho);
}

constexpr int kSums = 7;

// 256 threads: 7 sums -> partial[blockIdx.x*7 + k], fixed order.
__device__ inline void block_sum7(double v[kSums], double* __restrict__ partial)
{
    __shared__ double sh[4][kSums];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kSums; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kSums) {
        const int k = threadIdx.x;
        partial[blockIdx.x * kSums + k] = (sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k3_reduce(Args3 a, const T* __restrict__ src, const uint8_t* __restrict__ mat,
                                                 double* __restrict__ partial)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny * (uint32_t)a.nz;
    double acc[kSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
        if (mat_solid(mat[c])) continue;
        const int x = (int)(c % (uint32_t)a.nx);
        const uint32_t r = c / (uint32_t)a.nx;
        const int y = (int)(r % (uint32_t)a.ny), z = (int)(r / (uint32_t)a.ny);
        double f[19];
        gather_f64<T>(a, src, x, y, z, f);
#pragma unroll
        for (int i = 0; i < 19; ++i) {
            acc[0] += f[i];
            acc[1] += (double)ex3(i) * f[i];
            acc[2] += (double)ey3(i) * f[i];
            acc[3] += (double)ez3(i) * f[i];
        }
        // momentum handed to solid neighbours: 2 e_i f*_i, f*_i being what
        // this cell pushed into slot inv(i) of the solid cell x + e_i
        for (int i = 1; i < 19; ++i) {
            const int tx = wrap(x + ex3(i), a.nx), ty = wrap(y + ey3(i), a.ny), tz = wrap(z + ez3(i), a.nz);
            const uint32_t tr = (uint32_t)tz * (uint32_t)a.ny + (uint32_t)ty;
            if (!mat_solid(mat[tr * (uint32_t)a.nx + (uint32_t)tx])) continue;
            const double v =
                Store<T>::load_f64(src, tr * a.rs + (uint32_t)inv3(i) * a.pitch + (uint32_t)tx, w3(i), a.inv_scale);
            acc[4] += 2.0 * (double)ex3(i) * v;
            acc[5] += 2.0 * (double)ey3(i) * v;
            acc[6] += 2.0 * (double)ez3(i) * v;
        }
    }
    block_sum7(acc, partial);
}

__global__ __launch_bounds__(256) void k3_reduce_final(const double* __restrict__ partial, int nblocks,
                                                       double* __restrict__ out)
{
    double acc[kSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
        for (int k = 0; k < kSums; ++k) acc[k] += partial[b * kSums + k];
    }
    __shared__ double sh[4][kSums];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kSums; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < kSums) {
        const int k = threadIdx.x;
        out[k] = (sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]);
    }
}

struct Variant3 {
    int block, cpt, flags;
};
const Variant3 kVariants3[] = {
    {256, 1, kNtStore}, {256, 1, 0}, {128, 1, kNtStore}, {512, 1, kNtStore},
    {1024, 1, kNtStore}, {256, 2, kNtStore}, {64, 4, kNtStore},
    {256, 1, kNtStore | kXcdRemap}, {512, 1, kNtStore | kXcdRemap}, {128, 1, kNtStore | kXcdRemap},
};
constexpr int kNumVariants3 = (int)(sizeof(kVariants3) / sizeof(kVariants3[0]));

template <typename T, bool FORCE, int BS, int CPT, int FLAGS>
hipError_t launch3(const Args3& a, const void* src, void* dst, const uint32_t* bits, uint32_t words, hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + BS * CPT - 1) / (BS * CPT));
    const uint32_t rows = (uint32_t)a.ny * (uint32_t)a.nz;
    hipLaunchKernelGGL((k3_step<T, FORCE, BS, CPT, FLAGS>), dim3(nseg * rows), dim3(BS), 0, s, a, (const T*)src,
                       (T*)dst, bits, words, nseg);
    return hipGetLastError();
}

template <typename T, bool FORCE>
hipError_t dispatch3(int v, const Args3& a, const void* src, void* dst, const uint32_t* bits, uint32_t words,
                     hipStream_t s)
{
    constexpr int S = kNtStore;
    switch (v) {
    case 0: return launch3<T, FORCE, 256, 1, S>(a, src, dst, bits, words, s);
    case 1: return launch3<T, FORCE, 256, 1, 0>(a, src, dst, bits, words, s);
    case 2: return launch3<T, FORCE, 128, 1, S>(a, src, dst, bits, words, s);
    case 3: return launch3<T, FORCE, 512, 1, S>(a, src, dst, bits, words, s);
    case 4: return launch3<T, FORCE, 1024, 1, S>(a, src, dst, bits, words, s);
    case 5: return launch3<T, FORCE, 256, 2, S>(a, src, dst, bits, words, s);
    case 6: return launch3<T, FORCE, 64, 4, S>(a, src, dst, bits, words, s);
    case 7: return launch3<T, FORCE, 256, 1, S | kXcdRemap>(a, src, dst, bits, words, s);
    case 8: return launch3<T, FORCE, 512, 1, S | kXcdRemap>(a, src, dst, bits, words, s);
    case 9: return launch3<T, FORCE, 128, 1, S | kXcdRemap>(a, src, dst, bits, words, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step3(int v, bool f16, bool force, const Args3& a, const void* src, void* dst, const uint32_t* bits,
                        uint32_t words, hipStream_t s)
{
    if (f16) return force ? dispatch3<_Float16, true>(v, a, src, dst, bits, words, s)
                          : dispatch3<_Float16, false>(v, a, src, dst, bits, words, s);
    return force ? dispatch3<float, true>(v, a, src, dst, bits, words, s)
                 : dispatch3<float, false>(v, a, src, dst, bits, words, s);
}

}  // namespace d3
}  // namespace lbm

using namespace lbm;
using namespace lbm::d3;

#define TRY3(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return lbm_fail(LBM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct lbm3d_handle {
    int dev = 0;
    hipStream_t stream = nullptr;
    lbm3d_desc d{};
    Args3 a{};
    bool f16 = false;
    bool force = false;
    size_t elem = 4;
    size_t cells = 0;     // nx*ny*nz
    size_t elems = 0;     // elements of one population buffer
    uint32_t words = 0;   // solid-mask words per lattice row
    void* buf[2] = {nullptr, nullptr};
    uint8_t* mat = nullptr;
    uint32_t* bits = nullptr;
    float* macro = nullptr;  // 4*cells, allocated on first use
    double* partial = nullptr;
    double* red = nullptr;
    int cur = 0;
    int variant = 0;
    int graph_chunk = 100;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

unsigned blocks256(size_t n) { return (unsigned)((n + 255) / 256); }

void drop_graphs3(lbm3d_handle* h)
{
    for (auto& g : h->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void free3(lbm3d_handle* h)
{
    drop_graphs3(h);
    for (void* p : {h->buf[0], h->buf[1], (void*)h->mat, (void*)h->bits, (void*)h->macro, (void*)h->partial,
                    (void*)h->red})
        if (p) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int check3(lbm3d_handle* h)
{
    if (!h) return lbm_fail(LBM_ERR_INVALID, "null handle");
    TRY3(hipSetDevice(h->dev));
    return LBM_OK;
}

hipError_t prime3(lbm3d_handle* h, int b)
{
    if (h->f16)
        hipLaunchKernelGGL((k3_prime<_Float16>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                           (_Float16*)h->buf[b], h->mat);
    else
        hipLaunchKernelGGL((k3_prime<float>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                           (float*)h->buf[b], h->mat);
    return hipGetLastError();
}

hipError_t enqueue_step3(lbm3d_handle* h, int swap)
{
    return launch_step3(h->variant, h->f16, h->force, h->a, h->buf[swap], h->buf[swap ^ 1], h->bits, h->words,
                        h->stream);
}

int ensure_graph3(lbm3d_handle* h, int b)
{
    if (h->graph[b]) return LBM_OK;
    TRY3(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int s = 0; s < h->graph_chunk && le == hipSuccess; ++s) le = enqueue_step3(h, (b + s) & 1);
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(h->stream, &g);
    if (le != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return lbm_fail(LBM_ERR_HIP, "3-D graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ee));
    }
    hipError_t ie = hipGraphInstantiate(&h->graph[b], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        h->graph[b] = nullptr;
        return lbm_fail(LBM_ERR_HIP, "3-D hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    return LBM_OK;
}

int enqueue_run3(lbm3d_handle* h, int nsteps)
{
    int left = nsteps;
    if (h->graph_chunk > 0) {
        while (left >= h->graph_chunk) {
            int rc = ensure_graph3(h, h->cur);
            if (rc != LBM_OK) return rc;
            TRY3(hipGraphLaunch(h->graph[h->cur], h->stream));
            left -= h->graph_chunk;  // an even chunk ends in the buffer it started from
        }
    }
    for (int s = 0; s < left; ++s) {
        TRY3(enqueue_step3(h, h->cur));
        h->cur ^= 1;
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm3d_create(const lbm3d_desc* desc, lbm3d_handle** out)
{
    if (!out) return lbm_fail(LBM_ERR_INVALID, "null out");
    *out = nullptr;
    if (!desc) return lbm_fail(LBM_ERR_INVALID, "null descriptor");
    const lbm3d_desc& q = *desc;
    if (q.nx < 3 || q.ny < 3 || q.nz < 3) return lbm_fail(LBM_ERR_INVALID, "nx, ny and nz must be >= 3");
    if (!std::isfinite(q.tau) || !(q.tau > 0.5f)) return lbm_fail(LBM_ERR_INVALID, "tau must be > 0.5");
    if (q.storage != LBM_STORAGE_F32 && q.storage != LBM_STORAGE_F16_DEV)
        return lbm_fail(LBM_ERR_INVALID, "unknown storage %d", q.storage);
    if (!std::isfinite(q.force_x) || !std::isfinite(q.force_y) || !std::isfinite(q.force_z))
        return lbm_fail(LBM_ERR_INVALID, "force must be finite");
    for (int r : q.reserved)
        if (r) return lbm_fail(LBM_ERR_INVALID, "reserved fields must be zero");
    const uint64_t pitch = ((uint64_t)q.nx + 15u) & ~15ull;
    const uint64_t rows = (uint64_t)q.ny * (uint64_t)q.nz;
    const uint64_t elems = 19ull * pitch * rows;
    const uint64_t cells = (uint64_t)q.nx * rows;
    if (elems + pitch >= 0xFFFFFFFFull || cells * 4 >= 0xFFFFFFFFull)
        return lbm_fail(LBM_ERR_INVALID, "lattice exceeds 32-bit indexing");
    int dev = 0;
    TRY3(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    TRY3(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return lbm_fail(LBM_ERR_NO_DEVICE, "current HIP device %d is not a gfx950", dev);

    lbm3d_handle* h = new lbm3d_handle();
    h->dev = dev;
    h->d = q;
    h->f16 = q.storage == LBM_STORAGE_F16_DEV;
    h->elem = h->f16 ? 2 : 4;
    h->force = q.force_x != 0.0f || q.force_y != 0.0f || q.force_z != 0.0f;
    h->cells = (size_t)cells;
    h->elems = (size_t)elems;
    h->words = solid_mask_words(q.nx);
    Args3& a = h->a;
    a.nx = q.nx;
    a.ny = q.ny;
    a.nz = q.nz;
    a.pitch = (uint32_t)pitch;
    a.rs = 19u * (uint32_t)pitch;
    a.omega = 1.0f / q.tau;
    a.gx = q.force_x;
    a.gy = q.force_y;
    a.gz = q.force_z;
    a.scale = std::ldexp(1.0f, 10);
    a.inv_scale = std::ldexp(1.0f, -10);
    set_force_terms3(a);
    auto alloc = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return lbm_fail(LBM_ERR_OOM, "hipMalloc(%zu) failed", bytes);
        }
        return LBM_OK;
    };
    int rc;
    if ((rc = alloc(&h->buf[0], h->elems * h->elem)) || (rc = alloc(&h->buf[1], h->elems * h->elem)) ||
        (rc = alloc((void**)&h->mat, h->cells)) ||
        (rc = alloc((void**)&h->bits, (size_t)h->words * rows * sizeof(uint32_t))) ||
        (rc = alloc((void**)&h->partial, (size_t)kReduceBlocks * kSums * sizeof(double))) ||
        (rc = alloc((void**)&h->red, kSums * sizeof(double)))) {
        free3(h);
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&h->ev0);
    if (e == hipSuccess) e = hipEventCreate(&h->ev1);
    if (e == hipSuccess) e = hipMemsetAsync(h->mat, LBM_MAT_BULK, h->cells, h->stream);
    if (e == hipSuccess)
        e = launch_build_solid_bits(h->mat, q.nx, (int)rows, 1, h->words, h->bits, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        free3(h);
        return lbm_fail(LBM_ERR_HIP, "3-D setup: %s", hipGetErrorString(e));
    }
    if ((rc = lbm3d_reset(h)) != LBM_OK) {
        std::string msg = lbm_last_error();
        free3(h);
        return lbm_fail(rc, "%s", msg.c_str());
    }
    *out = h;
    return LBM_OK;
}

int lbm3d_destroy(lbm3d_handle* h)
{
    if (!h) return LBM_OK;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free3(h);
    return LBM_OK;
}

int lbm3d_write_materials(lbm3d_handle* h, const uint8_t* material)
{
    int rc = check3(h);
    if (rc) return rc;
    if (!material) return lbm_fail(LBM_ERR_INVALID, "null material");
    TRY3(hipMemcpyAsync(h->mat, material, h->cells, hipMemcpyHostToDevice, h->stream));
    TRY3(launch_build_solid_bits(h->mat, h->d.nx, h->d.ny * h->d.nz, 1, h->words, h->bits, h->stream));
    TRY3(prime3(h, h->cur));
    TRY3(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm3d_reset(lbm3d_handle* h)
{
    int rc = check3(h);
    if (rc) return rc;
    // deviation 0 is the rest state f = w in both storages
    TRY3(hipMemsetAsync(h->buf[0], 0, h->elems * h->elem, h->stream));
    TRY3(hipMemsetAsync(h->buf[1], 0, h->elems * h->elem, h->stream));
    h->cur = 0;
    TRY3(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm3d_write_populations(lbm3d_handle* h, const float* f)
{
    int rc = check3(h);
    if (rc) return rc;
    if (!f) return lbm_fail(LBM_ERR_INVALID, "null populations");
    const size_t bytes = 19 * h->cells * sizeof(float);
    float* stage = nullptr;
    TRY3(hipMalloc(&stage, bytes));
    hipError_t e = hipMemcpyAsync(stage, f, bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        if (h->f16)
            hipLaunchKernelGGL((k3_host_to_dev<_Float16>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                               stage, (_Float16*)h->buf[h->cur]);
        else
            hipLaunchKernelGGL((k3_host_to_dev<float>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                               stage, (float*)h->buf[h->cur]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = prime3(h, h->cur);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "3-D population upload: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm3d_read_populations(lbm3d_handle* h, float* f)
{
    int rc = check3(h);
    if (rc) return rc;
    if (!f) return lbm_fail(LBM_ERR_INVALID, "null populations");
    const size_t bytes = 19 * h->cells * sizeof(float);
    float* stage = nullptr;
    TRY3(hipMalloc(&stage, bytes));
    if (h->f16)
        hipLaunchKernelGGL((k3_dev_to_host<_Float16>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                           (const _Float16*)h->buf[h->cur], stage);
    else
        hipLaunchKernelGGL((k3_dev_to_host<float>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                           (const float*)h->buf[h->cur], stage);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(f, stage, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(stage);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "3-D population download: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm3d_run(lbm3d_handle* h, int nsteps)
{
    int rc = check3(h);
    if (rc) return rc;
    if (nsteps < 0) return lbm_fail(LBM_ERR_INVALID, "nsteps < 0");
    return enqueue_run3(h, nsteps);
}

int lbm3d_run_timed(lbm3d_handle* h, int nsteps, double* ms)
{
    int rc = check3(h);
    if (rc) return rc;
    if (!ms || nsteps < 0) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    TRY3(hipEventRecord(h->ev0, h->stream));
    if ((rc = enqueue_run3(h, nsteps))) return rc;
    TRY3(hipEventRecord(h->ev1, h->stream));
    TRY3(hipEventSynchronize(h->ev1));
    float t = 0.0f;
    TRY3(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = (double)t;
    return LBM_OK;
}

int lbm3d_synchronize(lbm3d_handle* h)
{
    int rc = check3(h);
    if (rc) return rc;
    TRY3(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm3d_macro(lbm3d_handle* h, float* rho, float* ux, float* uy, float* uz)
{
    int rc = check3(h);
    if (rc) return rc;
    if (!h->macro) TRY3(hipMalloc(&h->macro, 4 * h->cells * sizeof(float)));
    if (h->f16)
        hipLaunchKernelGGL((k3_macro<_Float16>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                           (const _Float16*)h->buf[h->cur], h->mat, h->macro);
    else
        hipLaunchKernelGGL((k3_macro<float>), dim3(blocks256(h->cells)), dim3(256), 0, h->stream, h->a,
                           (const float*)h->buf[h->cur], h->mat, h->macro);
    TRY3(hipGetLastError());
    float* dsts[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
        if (dsts[k])
            TRY3(hipMemcpyAsync(dsts[k], h->macro + (size_t)k * h->cells, h->cells * sizeof(float),
                                hipMemcpyDeviceToHost, h->stream));
    TRY3(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm3d_reduce(lbm3d_handle* h, double out[7])
{
    int rc = check3(h);
    if (rc) return rc;
    if (!out) return lbm_fail(LBM_ERR_INVALID, "null output");
    if (h->f16)
        hipLaunchKernelGGL((k3_reduce<_Float16>), dim3(kReduceBlocks), dim3(256), 0, h->stream, h->a,
                           (const _Float16*)h->buf[h->cur], h->mat, h->partial);
    else
        hipLaunchKernelGGL((k3_reduce<float>), dim3(kReduceBlocks), dim3(256), 0, h->stream, h->a,
                           (const float*)h->buf[h->cur], h->mat, h->partial);
    TRY3(hipGetLastError());
    hipLaunchKernelGGL(k3_reduce_final, dim3(1), dim3(256), 0, h->stream, h->partial, kReduceBlocks, h->red);
    TRY3(hipGetLastError());
    TRY3(hipMemcpyAsync(out, h->red, kSums * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    TRY3(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm3d_variant_count(void) { return kNumVariants3; }

int lbm3d_variant_info(int variant, int* block, int* cells_per_thread, int* flags)
{
    if (variant < 0 || variant >= kNumVariants3) return lbm_fail(LBM_ERR_INVALID, "no 3-D variant %d", variant);
    if (block) *block = kVariants3[variant].block;
    if (cells_per_thread) *cells_per_thread = kVariants3[variant].cpt;
    if (flags) *flags = kVariants3[variant].flags;
    return LBM_OK;
}

int lbm3d_set_variant(lbm3d_handle* h, int variant)
{
    int rc = check3(h);
    if (rc) return rc;
    if (variant < 0 || variant >= kNumVariants3) return lbm_fail(LBM_ERR_INVALID, "no 3-D variant %d", variant);
    TRY3(hipStreamSynchronize(h->stream));
    if (variant != h->variant) drop_graphs3(h);
    h->variant = variant;
    return LBM_OK;
}

int lbm3d_get_variant(lbm3d_handle* h, int* variant)
{
    if (!h || !variant) return lbm_fail(LBM_ERR_INVALID, "null argument");
    *variant = h->variant;
    return LBM_OK;
}

int lbm3d_set_graph_chunk(lbm3d_handle* h, int chunk)
{
    int rc = check3(h);
    if (rc) return rc;
    if (chunk < 0 || (chunk & 1)) return lbm_fail(LBM_ERR_INVALID, "graph chunk must be even and >= 0");
    TRY3(hipStreamSynchronize(h->stream));
    drop_graphs3(h);
    h->graph_chunk = chunk;
    return LBM_OK;
}

int lbm3d_autotune(lbm3d_handle* h, int steps, int* best_variant, double* ms_per_step)
{
    int rc = check3(h);
    if (rc) return rc;
    if (steps < 1) return lbm_fail(LBM_ERR_INVALID, "steps must be >= 1");
    TRY3(hipStreamSynchronize(h->stream));
    void* saved = nullptr;
    const size_t bytes = h->elems * h->elem;
    TRY3(hipMalloc(&saved, bytes));
    const int cur0 = h->cur, v0 = h->variant;
    hipError_t e = hipMemcpyAsync(saved, h->buf[cur0], bytes, hipMemcpyDeviceToDevice, h->stream);
    int best = v0;
    double best_ms = 1e300;
    for (int v = 0; v < kNumVariants3 && e == hipSuccess; ++v) {
        h->variant = v;
        for (int s = 0; s < 2 && e == hipSuccess; ++s, h->cur ^= 1) e = enqueue_step3(h, h->cur);
        if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
        for (int s = 0; s < steps && e == hipSuccess; ++s, h->cur ^= 1) e = enqueue_step3(h, h->cur);
        if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, h->ev0, h->ev1);
        if (e == hipSuccess && t / steps < best_ms * kTuneMargin) {
            best_ms = t / steps;
            best = v;
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h->buf[cur0], saved, bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(saved);
    h->cur = cur0;
    if (best != v0) drop_graphs3(h);
    h->variant = best;
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "3-D autotune: %s", hipGetErrorString(e));
    if (best_variant) *best_variant = best;
    if (ms_per_step) *ms_per_step = best_ms;
    return LBM_OK;
}

}  // extern "C"
