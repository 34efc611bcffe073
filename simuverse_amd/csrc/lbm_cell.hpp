// lbm_cell.hpp -- per-cell device functions of the physics mode shared by
// the one-step, two-step, batched and learned kernels: the neighbourhood
// of a cell in the packed solid mask, the pull gather (with the Zou-He
// inlet and the zero-gradient outlet), the store with the bounce-back push
// into solid neighbours, the whole fused cell step, and the four-cell
// vector form.  See lbm_kernels.hip for the streaming and bounce-back
// scheme.
#pragma once

#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {

// ---------------------------------------------------------------------------
// Physics mode
// ---------------------------------------------------------------------------

// Post-streaming deviations g_i = f_i - w_i of the fluid cell (x, y):
// pull with periodic y, periodic x or (x_edge 1) an own-slot bounce-back for
// populations that would come from outside [0, nx); zero-gradient outlet
// (column nx-1 gathers as column nx-2 when that cell is fluid) and Zou-He
// inlet on column 0.
template <typename T, int XEDGE>
__device__ inline void phys_gather(const PhysArgs& a, const T* __restrict__ src,
                                   const uint8_t* __restrict__ mat, int x, int y, float g[9])
{
    const int nx = a.nx, ny = a.ny;
    int cx = x;
    if (XEDGE == 1 && x == nx - 1 && !mat_solid(mat[(uint32_t)y * (uint32_t)nx + (uint32_t)(nx - 2)])) cx = nx - 2;
    const int ym = (y == 0) ? ny - 1 : y - 1;
    const int yp = (y == ny - 1) ? 0 : y + 1;
    int xm = cx - 1, xp = cx + 1;
    if (XEDGE == 0) {
        if (xm < 0) xm = nx - 1;
        if (xp >= nx) xp = 0;
    }
    const uint32_t rm = (uint32_t)ym * a.rs, r0 = (uint32_t)y * a.rs, rp = (uint32_t)yp * a.rs;
    const uint32_t cc = r0 + (uint32_t)cx;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const int sx = ex == 1 ? xm : (ex == -1 ? xp : cx);
        const uint32_t srow = ey == 1 ? rm : (ey == -1 ? rp : r0);
        bool oob = false;
        if (XEDGE == 1 && ex != 0) oob = (sx < 0) || (sx >= nx);
        if (oob)
            g[i] = Store<T>::load_dev(src, cc + (uint32_t)inv_of(i) * a.plane, w_phys(i), a.inv_scale);
        else
            g[i] = Store<T>::load_dev(src, srow + (uint32_t)sx + (uint32_t)i * a.plane, w_phys(i),
                                      a.inv_scale);
    }
    if (XEDGE == 1 && x == 0) {
        // Zou & He velocity inlet, in deviations: w0+w2+w4+2(w3+w6+w7) = 1.
        const float u0 = a.inlet_ux;
        const float G = g[0] + g[2] + g[4] + 2.0f * (g[3] + g[6] + g[7]);
        const float rho = (1.0f + G) / (1.0f - u0);
        const float ru = rho * u0;
        const float h = 0.5f * (g[4] - g[2]);
        g[1] = g[3] + (2.0f / 3.0f) * ru;
        g[5] = g[7] + h + (1.0f / 6.0f) * ru;
        g[8] = g[6] - h + (1.0f / 6.0f) * ru;
    }
}

// Neighbourhood of a cell: the wrapped rows above and below, and the solid
// flags (x-1, x, x+1) of rows ym, y, yp from three 64-bit windows of the
// packed mask.
struct CellNbr {
    int ym, yp;
    uint32_t sm, s0, sp;
};

__device__ __forceinline__ CellNbr cell_nbr(const PhysArgs& a, const uint32_t* __restrict__ bits, uint32_t words,
                                            int x, int y)
{
    CellNbr nb;
    nb.ym = (y == 0) ? a.ny - 1 : y - 1;
    nb.yp = (y == a.ny - 1) ? 0 : y + 1;
    nb.sm = solid3(bits, words, nb.ym, x);
    nb.s0 = solid3(bits, words, y, x);
    nb.sp = solid3(bits, words, nb.yp, x);
    return nb;
}

// The same for a cell whose rows y - 1 .. y + 1 lie inside the lattice
// (no wrap).  MUL24, for lattices with y, words < 2^24 (the tile kernels'
// fit rule): one 24-bit product (full rate; a 32-bit one issues at a
// quarter) and the neighbour rows' windows one `words` either side.
template <bool MUL24>
__device__ __forceinline__ CellNbr cell_nbr_inner(const uint32_t* __restrict__ bits, uint32_t words, int x, int y)
{
    CellNbr nb;
    nb.ym = y - 1;
    nb.yp = y + 1;
    if constexpr (MUL24) {
        const uint32_t* p = bits + (__umul24((uint32_t)y, words) + ((uint32_t)x >> 5));
        const uint32_t sh = (uint32_t)x & 31u;
        auto win = [sh](const uint32_t* q) {
            const uint64_t v = (uint64_t)q[0] | ((uint64_t)q[1] << 32);
            return (uint32_t)(v >> sh) & 7u;
        };
        nb.sm = win(p - words);
        nb.s0 = win(p);
        nb.sp = win(p + words);
    } else {
        nb.sm = solid3(bits, words, y - 1, x);
        nb.s0 = solid3(bits, words, y, x);
        nb.sp = solid3(bits, words, y + 1, x);
    }
    return nb;
}

// Zou & He velocity inlet on the gathered deviations of a column-0 cell
// (w0+w2+w4+2(w3+w6+w7) = 1); SC: g is scaled, and the unscaled sums and
// terms are formed exactly (power-of-two factors) so both forms agree.
template <bool SC>
__device__ __forceinline__ void zou_he_inlet(const PhysArgs& a, float g[9])
{
    const float u0 = a.inlet_ux;
    float G = g[0] + g[2] + g[4] + 2.0f * (g[3] + g[6] + g[7]);
    if (SC) G *= a.inv_scale;
    const float rho = (1.0f + G) / (1.0f - u0);
    const float ru = rho * u0;
    const float h = 0.5f * (g[4] - g[2]);
    float c23 = (2.0f / 3.0f) * ru, c16 = (1.0f / 6.0f) * ru;
    if (SC) {
        c23 *= a.scale;
        c16 *= a.scale;
    }
    g[1] = g[3] + c23;
    g[5] = g[7] + h + c16;
    g[8] = g[6] - h + c16;
}

// Post-streaming deviations of the fluid cell (x, y) (see phys_gather; the
// outlet test uses the west bit of the mask).
template <typename T, int XEDGE, int FLAGS, bool SC = false>
__device__ __forceinline__ void cell_gather(const PhysArgs& a, const T* __restrict__ src, const CellNbr& nb, int x,
                                            int y, float g[9])
{
    const int nx = a.nx;
    int cx = x;
    if (XEDGE == 1 && x == nx - 1 && !(nb.s0 & 1u)) cx = nx - 2;
    int xm = cx - 1, xp = cx + 1;
    if (XEDGE == 0) {
        if (xm < 0) xm = nx - 1;
        if (xp >= nx) xp = 0;
    }
    const uint32_t r0 = (uint32_t)y * a.rs, rm = (uint32_t)nb.ym * a.rs, rp = (uint32_t)nb.yp * a.rs;
    const uint32_t cc = r0 + (uint32_t)cx;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const int sx = ex == 1 ? xm : (ex == -1 ? xp : cx);
        const uint32_t srow = ey == 1 ? rm : (ey == -1 ? rp : r0);
        bool oob = false;
        if (XEDGE == 1 && ex != 0) oob = (sx < 0) || (sx >= nx);
        const uint32_t k = oob ? cc + (uint32_t)inv_of(i) * a.plane : srow + (uint32_t)sx + (uint32_t)i * a.plane;
        g[i] = Dom<T, SC>::in(ld<(FLAGS & kNtLoad) != 0>(src + k), a.inv_scale);
    }
    if (XEDGE == 1 && x == 0) zou_he_inlet<SC>(a, g);
}

// Stores the post-collision deviations of the fluid cell (x, y) in its own
// slots, and pushes f_i into slot inv(i) of each solid neighbour x + e_i
// (the mask's edge bits are empty for a non-periodic x, so nothing is pushed
// outside).
// The pushes of cell_store alone: f_i into slot inv(i) of each solid
// neighbour x + e_i of the fluid cell (x, y).
template <typename T, int XEDGE, bool SC = false>
__device__ __forceinline__ void cell_push(const PhysArgs& a, T* __restrict__ dst, const CellNbr& nb, int x, int y,
                                          const float g[9])
{
    if ((nb.sm | nb.sp | (nb.s0 & 5u)) == 0u) return;
    const int nx = a.nx;
    const uint32_t r0 = (uint32_t)y * a.rs;
    int xm = x - 1, xp = x + 1;
    if (XEDGE == 0) {
        if (xm < 0) xm = nx - 1;
        if (xp >= nx) xp = 0;
    }
    const uint32_t rm = (uint32_t)nb.ym * a.rs, rp = (uint32_t)nb.yp * a.rs;
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const uint32_t srow = ey == 1 ? nb.sp : (ey == -1 ? nb.sm : nb.s0);
        if (!((srow >> (1 + ex)) & 1u)) continue;
        const int tx = ex == 1 ? xp : (ex == -1 ? xm : x);
        const uint32_t trow = ey == 1 ? rp : (ey == -1 ? rm : r0);
        Dom<T, SC>::push(dst, trow + (uint32_t)tx + (uint32_t)inv_of(i) * a.plane, g[i], w_phys(i), a.scale);
    }
}

// cell_store with the cell's own element index y * rs + x given.
template <typename T, int XEDGE, int FLAGS, bool SC = false>
__device__ __forceinline__ void cell_store_at(const PhysArgs& a, T* __restrict__ dst, const CellNbr& nb, int x, int y,
                                              uint32_t own_index, const float g[9])
{
    // one per-lane pointer; the direction offsets are uniform across the wave
    T* const own = dst + own_index;
#pragma unroll
    for (int i = 0; i < 9; ++i)
        st<(FLAGS & kNtStore) != 0>(own + (size_t)((uint32_t)i * a.plane), Dom<T, SC>::out(g[i], a.scale));
    cell_push<T, XEDGE, SC>(a, dst, nb, x, y, g);
}

template <typename T, int XEDGE, int FLAGS, bool SC = false>
__device__ __forceinline__ void cell_store(const PhysArgs& a, T* __restrict__ dst, const CellNbr& nb, int x, int y,
                                           const float g[9])
{
    // one per-lane pointer; the direction offsets are uniform across the wave
    T* const own = dst + ((uint32_t)y * a.rs + (uint32_t)x);
#pragma unroll
    for (int i = 0; i < 9; ++i)
        st<(FLAGS & kNtStore) != 0>(own + (size_t)((uint32_t)i * a.plane), Dom<T, SC>::out(g[i], a.scale));
    cell_push<T, XEDGE, SC>(a, dst, nb, x, y, g);
}

// LEARNED adds the network's projected correction to the collision; mlp
// then points at its parameters (in LDS) and s_in, s_out are its scales.
template <typename T, int XEDGE, bool FORCE, bool EXPORT, int FLAGS = 0, bool LEARNED = false>
__device__ inline void phys_cell(const PhysArgs& a, const T* __restrict__ src, T* __restrict__ dst,
                                 const uint32_t* __restrict__ bits, uint32_t words,
                                 float* __restrict__ macro, uint32_t n, int x, int y,
                                 const float* __restrict__ mlp = nullptr, float s_in = 1.0f, float s_out = 1.0f)
{
    const uint32_t c = (uint32_t)y * (uint32_t)a.nx + (uint32_t)x;  // cell (macro planes)
    const CellNbr nb = cell_nbr(a, bits, words, x, y);
    if (nb.s0 & 2u) {
        if (EXPORT) {
            macro[c] = 0.0f;
            macro[c + n] = 0.0f;
            macro[c + 2 * n] = 0.0f;
            macro[c + 3 * n] = 0.0f;
        }
        return;
    }
    float g[9];
    // the learned correction works on unscaled deviations
    constexpr bool SC = kScaledDom<T> && !LEARNED;
    cell_gather<T, XEDGE, FLAGS, SC>(a, src, nb, x, y, g);
    float rho, ux, uy;
    phys_moments<FORCE, SC>(a, g, rho, ux, uy);
    if (EXPORT) {
        macro[c] = ux;
        macro[c + n] = uy;
        macro[c + 2 * n] = rho;
        macro[c + 3 * n] = 1.0f;
    }
    float nneq[3], y9[9];
    bool learned = false;
    if constexpr (LEARNED) {
        // The network is trained on cells whose 3x3 neighbourhood is fluid
        // (away from walls and open edges); only those get the correction.
        learned = (nb.sm | nb.s0 | nb.sp) == 0u && (XEDGE == 0 || (x > 0 && x < a.nx - 1));
        if (learned) {
            float inv_ws[3], xin[9];
            learned_input_scales(s_in, inv_ws);
            learned_inputs(g, rho, ux, uy, inv_ws, xin, nneq);
            mlp_forward_one(mlp, xin, y9);
        }
    }
    phys_collide<FORCE, SC>(a, g, rho, ux, uy);
    if (learned) add_stress_correction(nneq, y9, s_out, modulation_bound(a.omega), g);
    cell_store<T, XEDGE, FLAGS, SC>(a, dst, nb, x, y, g);
}

// Four consecutive cells of a row per lane.  Loads move 4 elements per
// direction (16 B in binary32; 4 B-aligned for the x-shifted rows), stores
// are aligned (the row stride and direction stride are multiples of 4).  Only lanes whose 4 cells and their x-neighbours lie inside
// the row (1 <= x0, x0 + 4 <= nx - 1) take this path; the per-cell
// arithmetic is phys_moments/phys_collide, as in phys_cell, so results are
// identical bit for bit.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    __device__ static inline void load(const float* p, float out[4])
    {
        float4 v;
        __builtin_memcpy(&v, p, sizeof v);
        out[0] = v.x;
        out[1] = v.y;
        out[2] = v.z;
        out[3] = v.w;
    }
    __device__ static inline void store(float* p, const float in[4])
    {
        *reinterpret_cast<float4*>(p) = make_float4(in[0], in[1], in[2], in[3]);
    }
};
template <>
struct Vec4<_Float16> {
    __device__ static inline void load(const _Float16* p, float out[4])
    {
        uint2 v;
        __builtin_memcpy(&v, p, sizeof v);
        out[0] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.x & 0xFFFFu));
        out[1] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.x >> 16));
        out[2] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.y & 0xFFFFu));
        out[3] = (float)__builtin_bit_cast(_Float16, (uint16_t)(v.y >> 16));
    }
    __device__ static inline void store(_Float16* p, const float in[4])
    {
        const uint32_t a = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)in[0]) |
                           ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)in[1]) << 16);
        const uint32_t b = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)in[2]) |
                           ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)in[3]) << 16);
        *reinterpret_cast<uint2*>(p) = make_uint2(a, b);
    }
};

// Bits (x0-1 .. x0+4) of row y as bits 0..5 (x0 is a multiple of 4).
__device__ inline uint32_t solid6(const uint32_t* __restrict__ bits, uint32_t words, int y, int x0)
{
    const uint32_t* p = bits + (uint32_t)y * words + ((uint32_t)x0 >> 5);
    const uint64_t v = (uint64_t)p[0] | ((uint64_t)p[1] << 32);
    return (uint32_t)(v >> ((uint32_t)x0 & 31u)) & 63u;
}

template <typename T, bool FORCE, bool EXPORT>
__device__ inline void phys_cells4(const PhysArgs& a, const T* __restrict__ src, T* __restrict__ dst,
                                   const uint32_t* __restrict__ bits, uint32_t words, float* __restrict__ macro,
                                   uint32_t n, int x0, int y)
{
    const int nx = a.nx, ny = a.ny;
    const int ym = (y == 0) ? ny - 1 : y - 1;
    const int yp = (y == ny - 1) ? 0 : y + 1;
    const uint32_t r0 = (uint32_t)y * a.rs, rm = (uint32_t)ym * a.rs, rp = (uint32_t)yp * a.rs;
    const uint32_t sm = solid6(bits, words, ym, x0);
    const uint32_t s0 = solid6(bits, words, y, x0);
    const uint32_t sp = solid6(bits, words, yp, x0);
    const uint32_t self = (s0 >> 1) & 15u;  // bit k: cell x0 + k is solid
    constexpr bool SC = kScaledDom<T>;
    const uint32_t c0 = (uint32_t)y * (uint32_t)nx + (uint32_t)x0;  // cell (macro planes)
    const uint32_t p0 = r0 + (uint32_t)x0;                         // population element
    if (self == 15u) {
        if (EXPORT)
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q) macro[c0 + k + q * n] = 0.0f;
        return;
    }
    float g[9][4];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const uint32_t srow = ey == 1 ? rm : (ey == -1 ? rp : r0);
        Vec4<T>::load(src + srow + (uint32_t)(x0 - ex) + (uint32_t)i * a.plane, g[i]);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float gc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) gc[i] = Dom<T, SC>::in_w(g[i][k], a.inv_scale);
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, gc, rho, ux, uy);
        if (EXPORT) {
            const bool s = (self >> k) & 1u;
            macro[c0 + k] = s ? 0.0f : ux;
            macro[c0 + k + n] = s ? 0.0f : uy;
            macro[c0 + k + 2 * n] = s ? 0.0f : rho;
            macro[c0 + k + 3 * n] = s ? 0.0f : 1.0f;
        }
        phys_collide<FORCE, SC>(a, gc, rho, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i][k] = gc[i];
    }
    if (self == 0u) {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            float o[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) o[k] = Dom<T, SC>::out_w(g[i][k], a.scale);
            Vec4<T>::store(dst + p0 + (uint32_t)i * a.plane, o);
        }
    } else {
        // A solid cell's slots receive pushes from its neighbours during
        // this step: write the fluid cells only.
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if ((self >> k) & 1u) continue;
#pragma unroll
            for (int i = 0; i < 9; ++i)
                Dom<T, SC>::push(dst, p0 + (uint32_t)k + (uint32_t)i * a.plane, g[i][k], w_phys(i), a.scale);
        }
    }
    if ((sm | sp | s0) == 0u) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if ((self >> k) & 1u) continue;
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            const int ex = ex_of(i), ey = ey_of(i);
            const uint32_t srow = ey == 1 ? sp : (ey == -1 ? sm : s0);
            if (!((srow >> (k + 1 + ex)) & 1u)) continue;
            const uint32_t trow = ey == 1 ? rp : (ey == -1 ? rm : r0);
            Dom<T, SC>::push(dst, trow + (uint32_t)(x0 + k + ex) + (uint32_t)inv_of(i) * a.plane, g[i][k],
                             w_phys(i), a.scale);
        }
    }
}

}  // namespace lbm
