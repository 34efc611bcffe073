// lbm_learned.hip -- the learned collision step (network on the vector ALUs
// or on the matrix cores), the 4x restriction that makes its training data,
// and the health check behind the guarded run, with their launchers.
#include "lbm_cell.hpp"
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {

// 4x restriction of a fine physics-mode state to a coarse lattice
// (nx/4 x ny/4), one coarse cell per thread.  A coarse cell is solid when at
// least 8 of its 16 fine cells are; otherwise, over its fluid fine cells,
//   rho_c = mean rho,  u_c = sum(rho u) / sum(rho),
//   f_c = feq(rho_c, u_c) + P(4 tau_c/tau_f * mean(f - feq)) - 3 w e.(g_c/2)
// with P the projection that removes mass and momentum: the non-equilibrium
// part is rescaled for a 4x coarser lattice under acoustic scaling
// (Dupuis & Chopard, Phys. Rev. E 67 (2003) 066707), and the last term makes
// the coarse Guo velocity (sum e f + g_c/2)/rho equal u_c.  Output: dense
// pre-collision f (binary32), the solid map and u_c.
template <typename T, int XEDGE, bool FORCE>
__global__ __launch_bounds__(256) void k_restrict4(PhysArgs a, const T* __restrict__ src,
                                                   const uint8_t* __restrict__ mat, float nscale, float gcx, float gcy,
                                                   float* __restrict__ fc, uint8_t* __restrict__ solid_c,
                                                   float* __restrict__ uxc, float* __restrict__ uyc)
{
    const int cx_n = a.nx / 4, cy_n = a.ny / 4;
    const uint32_t nc = (uint32_t)cx_n * (uint32_t)cy_n;
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= nc) return;
    const int X = (int)(k % (uint32_t)cx_n), Y = (int)(k / (uint32_t)cx_n);
    float sr = 0.0f, sjx = 0.0f, sjy = 0.0f, sneq[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) sneq[i] = 0.0f;
    int nfl = 0;
    for (int j = 0; j < 4; ++j)
        for (int i4 = 0; i4 < 4; ++i4) {
            const int x = 4 * X + i4, y = 4 * Y + j;
            if (mat_solid(mat[(uint32_t)y * (uint32_t)a.nx + (uint32_t)x])) continue;
            ++nfl;
            float g[9], rho, ux, uy;
            phys_gather<T, XEDGE>(a, src, mat, x, y, g);
            phys_moments<FORCE>(a, g, rho, ux, uy);
            sr += rho;
            sjx += rho * ux;
            sjy += rho * uy;
            const float dr = rho - 1.0f, usq = 1.5f * (ux * ux + uy * uy);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const float eu = (float)ex_of(i) * ux + (float)ey_of(i) * uy;
                sneq[i] += g[i] - w_phys(i) * (dr + rho * (3.0f * eu + 4.5f * (eu * eu) - usq));
            }
        }
    const uint32_t n = nc;
    if (nfl < 8) {
        solid_c[k] = 1;
#pragma unroll
        for (int i = 0; i < 9; ++i) fc[(size_t)i * n + k] = 0.0f;
        if (uxc) uxc[k] = 0.0f;
        if (uyc) uyc[k] = 0.0f;
        return;
    }
    solid_c[k] = 0;
    const float rho = sr / (float)nfl, ux = sjx / sr, uy = sjy / sr;
    float neq[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) neq[i] = nscale * sneq[i] / (float)nfl;
    project_correction(neq);
    const float usq = 1.5f * (ux * ux + uy * uy);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float exf = (float)ex_of(i), eyf = (float)ey_of(i);
        const float eu = exf * ux + eyf * uy;
        const float feq = w_phys(i) * rho * (1.0f + 3.0f * eu + 4.5f * (eu * eu) - usq);
        fc[(size_t)i * n + k] = feq + neq[i] - 1.5f * w_phys(i) * (exf * gcx + eyf * gcy);
    }
    if (uxc) uxc[k] = ux;
    if (uyc) uyc[k] = uy;
}

// Health of a state: largest |u|^2 over fluid cells (as the bits of a
// non-negative float, so unsigned max orders them) and whether any cell is
// not finite.
template <typename T, int XEDGE, bool FORCE>
__global__ __launch_bounds__(256) void k_phys_health(PhysArgs a, const T* __restrict__ src,
                                                     const uint8_t* __restrict__ mat, uint32_t* __restrict__ out)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    uint32_t u2bits = 0, bad = 0;
    if (c < n && !mat_solid(mat[c])) {
        const int y = (int)(c / (uint32_t)a.nx), x = (int)(c % (uint32_t)a.nx);
        float g[9], rho, ux, uy;
        phys_gather<T, XEDGE>(a, src, mat, x, y, g);
        phys_moments<FORCE>(a, g, rho, ux, uy);
        const float u2 = ux * ux + uy * uy;
        if (!(u2 <= 3.0e38f) || !(rho > 0.0f)) bad = 1;
        else u2bits = __float_as_uint(u2);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        u2bits = max(u2bits, (uint32_t)__shfl_xor((int)u2bits, off, 64));
        bad |= (uint32_t)__shfl_xor((int)bad, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(out, u2bits);
        if (bad) atomicOr(out + 1, 1u);
    }
}

// One step with the learned collision correction (binary32 network, its
// parameters staged in LDS once per block).  One cell per thread.  macro
// (or nullptr): the step also exports (ux, uy, rho, w) of the gathered
// populations, as the plain step does.
template <typename T, int XEDGE, bool FORCE>
__global__ __launch_bounds__(256) void k_phys_step_learned(PhysArgs a, const T* __restrict__ src,
                                                           T* __restrict__ dst, const uint32_t* __restrict__ bits,
                                                           uint32_t words, const float* __restrict__ params,
                                                           float s_in, float s_out, uint32_t nseg,
                                                           float* __restrict__ macro)
{
    __shared__ float P[kMlpParams];
    for (int k = threadIdx.x; k < kMlpParams; k += 256) P[k] = params[k];
    __syncthreads();
    const int y = (int)(blockIdx.x / nseg);
    const int x = (int)(blockIdx.x % nseg) * 256 + (int)threadIdx.x;
    if (x >= a.nx) return;
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    if (macro)
        phys_cell<T, XEDGE, FORCE, true, kNtStore, true>(a, src, dst, bits, words, macro, n, x, y, P, s_in, s_out);
    else
        phys_cell<T, XEDGE, FORCE, false, kNtStore, true>(a, src, dst, bits, words, nullptr, n, x, y, P, s_in,
                                                          s_out);
}

// One step with the learned correction evaluated on the matrix cores
// (operands of type E -- bf16 or binary16 -- and binary32 accumulation;
// mlp_wave).  Blocks of 4
// wavefronts loop over row segments of 256 cells, one cell per lane, so the
// 14 KiB of packed weights are staged in LDS once per block.  Every lane
// takes part in the matrix instructions; lanes beyond the row or on solid
// cells feed zeros and store nothing.
template <typename T, int XEDGE, bool FORCE, typename E>
__global__ __launch_bounds__(256) void k_phys_step_learned_mfma(PhysArgs a, const T* __restrict__ src,
                                                                T* __restrict__ dst, const uint32_t* __restrict__ bits,
                                                                uint32_t words, const E* __restrict__ packed,
                                                                const float* __restrict__ params, float s_in,
                                                                float s_out, uint32_t nseg,
                                                                float* __restrict__ macro)
{
    typedef typename Mma<E>::v8 v8;
    __shared__ __align__(16) E W[kPackElems];
    __shared__ __align__(16) float bias[kBiasElems];
    __shared__ __align__(16) E xs[4][64 * 16];
    __shared__ float ys[4][64 * kMlpOut];
    {
        const uint4* from = reinterpret_cast<const uint4*>(packed);
        uint4* to = reinterpret_cast<uint4*>(W);
        for (int k = threadIdx.x; k < kPackElems / 8; k += 256) to[k] = from[k];
        for (int k = threadIdx.x; k < kMlpH; k += 256) {
            bias[k] = params[kOffB1 + k];
            bias[kMlpH + k] = params[kOffB2 + k];
        }
        if (threadIdx.x < kBiasElems - 2 * kMlpH)
            bias[2 * kMlpH + threadIdx.x] = threadIdx.x < kMlpOut ? params[kOffB3 + threadIdx.x] : 0.0f;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t total = nseg * (uint32_t)a.ny;
    float inv_ws[3];
    learned_input_scales(s_in, inv_ws);
    const float bound = modulation_bound(a.omega);
    for (uint32_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        const int y = (int)(tile / nseg);
        const int x = (int)(tile % nseg) * 256 + (int)threadIdx.x;
        // Every lane runs the same straight-line arithmetic up to the
        // network: a lane past the row works on the row's last cell, a lane
        // on a solid cell on whatever its slots hold, and neither stores
        // anything nor feeds the network (the wavefront takes no branch and
        // sets no registers to zero for them).
        const bool in_row = x < a.nx;
        const int xc = in_row ? x : a.nx - 1;
        const CellNbr nb = cell_nbr(a, bits, words, xc, y);
        const bool fluid = in_row && !(nb.s0 & 2u);
        // trained on cells whose 3x3 neighbourhood is fluid: only those get
        // the correction
        const bool learned =
            in_row && (nb.sm | nb.s0 | nb.sp) == 0u && (XEDGE == 0 || (x > 0 && x < a.nx - 1));
        float g[9], rho, ux, uy;
        cell_gather<T, XEDGE, kNtStore>(a, src, nb, xc, y, g);
        phys_moments<FORCE>(a, g, rho, ux, uy);
        float xin[9], nneq[3];
        learned_inputs(g, rho, ux, uy, inv_ws, xin, nneq);
        if (macro && in_row) {
            // the macro field of the gathered populations (0 on solids)
            const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny, c = (uint32_t)y * (uint32_t)a.nx + (uint32_t)x;
            macro[c] = fluid ? ux : 0.0f;
            macro[c + n] = fluid ? uy : 0.0f;
            macro[c + 2 * n] = fluid ? rho : 0.0f;
            macro[c + 3 * n] = fluid ? 1.0f : 0.0f;
        }
        // BGK now, the correction after the network: only g, the three
        // moments and the neighbour bits stay live across it.
        phys_collide<FORCE>(a, g, rho, ux, uy);
        v8 lo, hi;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lo[i] = (E)xin[i];
            hi[i] = (E)(i == 0 ? xin[8] : 0.0f);
        }
        // cells that are not corrected feed zeros (their column of the
        // product is ignored anyway; zeros keep it finite)
        if (!learned) {
            lo = v8{};
            hi = v8{};
        }
        *reinterpret_cast<v8*>(&xs[wave][lane * 16]) = lo;
        *reinterpret_cast<v8*>(&xs[wave][lane * 16 + 8]) = hi;
        // xs[wave] and ys[wave] belong to this wavefront alone, whose LDS
        // accesses complete in issue order: a wavefront-scope fence keeps
        // the compiler from moving them across, and no block barrier is
        // needed (the wavefronts of a block run independently).
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        mlp_wave(W, bias, xs[wave], ys[wave], lane);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (learned) {
            float y3[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) y3[i] = ys[wave][lane * kMlpOut + i];
            add_stress_correction(nneq, y3, s_out, bound, g);
        }
        if (fluid) cell_store<T, XEDGE, kNtStore>(a, dst, nb, x, y, g);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_restrict4(bool f16, int xedge, bool force, const PhysArgs& a, const void* src, const uint8_t* mat,
                            float nscale, float gcx, float gcy, float* fc, uint8_t* solid_c, float* uxc, float* uyc,
                            hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)(a.nx / 4) * (size_t)(a.ny / 4), 256);
#define LBM_R(T, XE, FO)                                                                                       \
    hipLaunchKernelGGL((k_restrict4<T, XE, FO>), dim3(nb), dim3(256), 0, s, a, (const T*)src, mat, nscale, gcx, \
                       gcy, fc, solid_c, uxc, uyc)
    if (f16) {
        if (xedge) { if (force) LBM_R(_Float16, 1, true); else LBM_R(_Float16, 1, false); }
        else { if (force) LBM_R(_Float16, 0, true); else LBM_R(_Float16, 0, false); }
    } else {
        if (xedge) { if (force) LBM_R(float, 1, true); else LBM_R(float, 1, false); }
        else { if (force) LBM_R(float, 0, true); else LBM_R(float, 0, false); }
    }
#undef LBM_R
    return hipGetLastError();
}

hipError_t launch_phys_health(bool f16, int xedge, bool force, const PhysArgs& a, const void* src, const uint8_t* mat,
                              uint32_t* out, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
#define LBM_H(T, XE, FO) \
    hipLaunchKernelGGL((k_phys_health<T, XE, FO>), dim3(nb), dim3(256), 0, s, a, (const T*)src, mat, out)
    if (f16) {
        if (xedge) { if (force) LBM_H(_Float16, 1, true); else LBM_H(_Float16, 1, false); }
        else { if (force) LBM_H(_Float16, 0, true); else LBM_H(_Float16, 0, false); }
    } else {
        if (xedge) { if (force) LBM_H(float, 1, true); else LBM_H(float, 1, false); }
        else { if (force) LBM_H(float, 0, true); else LBM_H(float, 0, false); }
    }
#undef LBM_H
    return hipGetLastError();
}

hipError_t launch_phys_step_learned(bool f16, int xedge, bool force, const PhysArgs& a, const void* src, void* dst,
                                   const uint32_t* bits, uint32_t words, const float* params, float s_in,
                                   float s_out, float* macro, hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + 255) / 256);
    const dim3 grid(nseg * (uint32_t)a.ny), block(256);
#define LBM_L(T, XE, FO)                                                                                  \
    hipLaunchKernelGGL((k_phys_step_learned<T, XE, FO>), grid, block, 0, s, a, (const T*)src, (T*)dst, bits, \
                       words, params, s_in, s_out, nseg, macro)
    if (f16) {
        if (xedge) { if (force) LBM_L(_Float16, 1, true); else LBM_L(_Float16, 1, false); }
        else { if (force) LBM_L(_Float16, 0, true); else LBM_L(_Float16, 0, false); }
    } else {
        if (xedge) { if (force) LBM_L(float, 1, true); else LBM_L(float, 1, false); }
        else { if (force) LBM_L(float, 0, true); else LBM_L(float, 0, false); }
    }
#undef LBM_L
    return hipGetLastError();
}

hipError_t launch_phys_step_learned_mfma(bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                                        void* dst, const uint32_t* bits, uint32_t words, const void* packed,
                                        bool op_f16, const float* params, float s_in, float s_out, float* macro,
                                        hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + 255) / 256);
    const uint32_t tiles = nseg * (uint32_t)a.ny;
    const dim3 grid(tiles < 2048u ? tiles : 2048u), block(256);
#define LBM_L(T, XE, FO)                                                                                          \
    do {                                                                                                          \
        if (op_f16)                                                                                               \
            hipLaunchKernelGGL((k_phys_step_learned_mfma<T, XE, FO, _Float16>), grid, block, 0, s, a, (const T*)src, \
                               (T*)dst, bits, words, (const _Float16*)packed, params, s_in, s_out, nseg, macro);   \
        else                                                                                                      \
            hipLaunchKernelGGL((k_phys_step_learned_mfma<T, XE, FO, __bf16>), grid, block, 0, s, a, (const T*)src, \
                               (T*)dst, bits, words, (const __bf16*)packed, params, s_in, s_out, nseg, macro);     \
    } while (0)
    if (f16) {
        if (xedge) { if (force) LBM_L(_Float16, 1, true); else LBM_L(_Float16, 1, false); }
        else { if (force) LBM_L(_Float16, 0, true); else LBM_L(_Float16, 0, false); }
    } else {
        if (xedge) { if (force) LBM_L(float, 1, true); else LBM_L(float, 1, false); }
        else { if (force) LBM_L(float, 0, true); else LBM_L(float, 0, false); }
    }
#undef LBM_L
    return hipGetLastError();
}

}  // namespace lbm
