// lbm_kernels.hip -- gfx950 kernels of the D2Q9 hot path (one domain).
//
// Layout in HBM: populations are structure-of-arrays, one plane per
// direction, plane stride `plane` elements (nx*ny rounded up to 64), row
// major inside a plane; two buffers ping-pong.  Materials are one byte per
// cell in the reference's LatticeType numbering.
//
// Streaming is the pull form (a cell gathers f_i from x - e_i and writes
// only its own slots), fused with the collision.  Bounce-back is folded in
// by a push into the solid cell: a fluid cell x whose neighbour x + e_i is
// solid also writes its post-collision f_i into slot inv(i) of that solid
// cell.  The next step's gather then reads the bounced value at x - e_inv(i)
// like any other population, so the gather has no branches and no extra
// loads.  Each solid slot has exactly one writer (x = S + e_k for slot k of
// solid S), so the fused step is race-free.  This is the reference's
// boundary pass (assets/wgsl/lbm/boundary.wgsl:16-33: dst[S, inv(i)] :=
// dst[S - e_i, i]) done by the producing cell instead of a second dispatch.
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
template <typename T, int XEDGE, int FLAGS, bool SC = false>
__device__ __forceinline__ void cell_store(const PhysArgs& a, T* __restrict__ dst, const CellNbr& nb, int x, int y,
                                           const float g[9])
{
    const int nx = a.nx;
    const uint32_t r0 = (uint32_t)y * a.rs;
    // one per-lane pointer; the direction offsets are uniform across the wave
    T* const own = dst + (r0 + (uint32_t)x);
#pragma unroll
    for (int i = 0; i < 9; ++i)
        st<(FLAGS & kNtStore) != 0>(own + (size_t)((uint32_t)i * a.plane), Dom<T, SC>::out(g[i], a.scale));
    if ((nb.sm | nb.sp | (nb.s0 & 5u)) == 0u) return;
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
    float corr[9];
    bool learned = false;
    if constexpr (LEARNED) {
        // The network is trained on cells whose 3x3 neighbourhood is fluid
        // (away from walls and open edges); only those get the correction.
        learned = (nb.sm | nb.s0 | nb.sp) == 0u && (XEDGE == 0 || (x > 0 && x < a.nx - 1));
        if (learned) {
            float xin[9], y9[9];
            cell_features(g, rho, ux, uy, s_in, xin);
            mlp_forward_one(mlp, xin, y9);
            learned_correction(xin, y9, s_in, s_out, a.omega, corr);
        }
    }
    phys_collide<FORCE, SC>(a, g, rho, ux, uy);
    if (learned) {
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] += corr[i];
    }
    cell_store<T, XEDGE, FLAGS, SC>(a, dst, nb, x, y, g);
}

// One step of the whole lattice.  Block = BS threads on a row segment of
// BS*CPT cells; thread t takes cells t, t+BS, ... so every load and store
// instruction of a wavefront covers 64 consecutive cells of one plane.
template <typename T, int XEDGE, bool FORCE, bool EXPORT, int BS, int CPT, int FLAGS>
__global__ __launch_bounds__(BS) void k_phys_step(PhysArgs a, const T* __restrict__ src,
                                                  T* __restrict__ dst, const uint32_t* __restrict__ bits,
                                                  uint32_t words, float* __restrict__ macro, uint32_t nseg)
{
    const uint32_t total = nseg * (uint32_t)a.ny;
    const uint32_t tile = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, total) : blockIdx.x;
    const int y = (int)(tile / nseg);
    const int x0 = (int)(tile % nseg) * (BS * CPT) + (int)threadIdx.x;
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int x = x0 + k * BS;
        if (x < a.nx) phys_cell<T, XEDGE, FORCE, EXPORT, FLAGS>(a, src, dst, bits, words, macro, n, x, y);
    }
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

template <typename T, int XEDGE, bool FORCE, bool EXPORT, int BS, int FLAGS>
__global__ __launch_bounds__(BS) void k_phys_step4(PhysArgs a, const T* __restrict__ src, T* __restrict__ dst,
                                                   const uint32_t* __restrict__ bits, uint32_t words,
                                                   float* __restrict__ macro, uint32_t nseg)
{
    const uint32_t total = nseg * (uint32_t)a.ny;
    const uint32_t tile = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, total) : blockIdx.x;
    const int y = (int)(tile / nseg);
    const int x0 = (int)(tile % nseg) * (BS * 4) + 4 * (int)threadIdx.x;
    if (x0 >= a.nx) return;
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    if (x0 >= 1 && x0 + 4 <= a.nx - 1) {
        phys_cells4<T, FORCE, EXPORT>(a, src, dst, bits, words, macro, n, x0, y);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            phys_cell<T, XEDGE, FORCE, EXPORT, FLAGS>(a, src, dst, bits, words, macro, n, x0 + k, y);
    }
}

// V consecutive cells per lane with aligned V-wide loads of every direction
// row: a wavefront covers 64*V consecutive cells of row y, and the x-shifted
// directions take their one missing element from the adjacent lane
// (__shfl_up / __shfl_down) or, at the wavefront's ends, from a load of the
// cell just outside it (the same address for every lane).  Periodic x only;
// nx must be a multiple of the block's BS*V cells.  The per-cell arithmetic
// is phys_moments/phys_collide, so results equal phys_cell's bit for bit.
template <typename T, bool FORCE, bool EXPORT, int V, int FLAGS>
__device__ inline void phys_cells_shfl(const PhysArgs& a, const T* __restrict__ src, T* __restrict__ dst,
                                       const uint32_t* __restrict__ bits, uint32_t words,
                                       float* __restrict__ macro, uint32_t n, int wbase, int lane, int y)
{
    typedef T vecT __attribute__((ext_vector_type(V)));
    constexpr bool SC = kScaledDom<T>;
    const int nx = a.nx, ny = a.ny;
    const int x0 = wbase + lane * V;
    const int ym = (y == 0) ? ny - 1 : y - 1;
    const int yp = (y == ny - 1) ? 0 : y + 1;
    const uint32_t r0 = (uint32_t)y * a.rs, rm = (uint32_t)ym * a.rs, rp = (uint32_t)yp * a.rs;
    const int xl = wbase == 0 ? nx - 1 : wbase - 1;
    const int xr = wbase + 64 * V == nx ? 0 : wbase + 64 * V;
    float g[9][V];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const T* row = src + (ey == 1 ? rm : (ey == -1 ? rp : r0)) + (uint32_t)i * a.plane;
        vecT v;
        if constexpr ((FLAGS & kNtLoad) != 0) v = __builtin_nontemporal_load(reinterpret_cast<const vecT*>(row + x0));
        else v = *reinterpret_cast<const vecT*>(row + x0);
        float f[V];
#pragma unroll
        for (int k = 0; k < V; ++k) f[k] = Dom<T, SC>::in(v[k], a.inv_scale);
        if (ex == 0) {
#pragma unroll
            for (int k = 0; k < V; ++k) g[i][k] = f[k];
        } else if (ex == 1) {  // pulls from x - 1
            float prev = __shfl_up(f[V - 1], 1, 64);
            const float edge = Dom<T, SC>::in(row[xl], a.inv_scale);
            g[i][0] = lane == 0 ? edge : prev;
#pragma unroll
            for (int k = 1; k < V; ++k) g[i][k] = f[k - 1];
        } else {  // pulls from x + 1
            float next = __shfl_down(f[0], 1, 64);
            const float edge = Dom<T, SC>::in(row[xr], a.inv_scale);
#pragma unroll
            for (int k = 0; k < V - 1; ++k) g[i][k] = f[k + 1];
            g[i][V - 1] = lane == 63 ? edge : next;
        }
    }
    // Solid bits of cells x0-1 .. x0+V of rows ym, y, yp (bit x+1 of a mask
    // row is cell x; x0 % 32 + V + 2 <= 64 keeps the window in two words).
    auto window = [&](int yy) {
        const uint32_t* p = bits + (uint32_t)yy * words + ((uint32_t)x0 >> 5);
        const uint64_t w64 = (uint64_t)p[0] | ((uint64_t)p[1] << 32);
        return (uint32_t)(w64 >> ((uint32_t)x0 & 31u)) & ((1u << (V + 2)) - 1u);
    };
    const uint32_t sm = window(ym), s0 = window(y), sp = window(yp);
    const uint32_t self = (s0 >> 1) & ((1u << V) - 1u);  // bit k: cell x0 + k is solid
    const uint32_t c0 = (uint32_t)y * (uint32_t)nx + (uint32_t)x0;
    const uint32_t p0 = r0 + (uint32_t)x0;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        float gc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) gc[i] = g[i][k];
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, gc, rho, ux, uy);
        if (EXPORT) {
            const bool sk = (self >> k) & 1u;
            macro[c0 + k] = sk ? 0.0f : ux;
            macro[c0 + k + n] = sk ? 0.0f : uy;
            macro[c0 + k + 2 * n] = sk ? 0.0f : rho;
            macro[c0 + k + 3 * n] = sk ? 0.0f : 1.0f;
        }
        phys_collide<FORCE, SC>(a, gc, rho, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i][k] = gc[i];
    }
    if (self == 0u) {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            vecT o;
#pragma unroll
            for (int k = 0; k < V; ++k) o[k] = Dom<T, SC>::out(g[i][k], a.scale);
            vecT* q = reinterpret_cast<vecT*>(dst + p0 + (uint32_t)i * a.plane);
            if constexpr ((FLAGS & kNtStore) != 0) __builtin_nontemporal_store(o, q);
            else *q = o;
        }
    } else {
        // A solid cell's slots receive pushes from its neighbours during
        // this step: write the fluid cells only.
#pragma unroll
        for (int k = 0; k < V; ++k) {
            if ((self >> k) & 1u) continue;
#pragma unroll
            for (int i = 0; i < 9; ++i)
                Dom<T, SC>::push(dst, p0 + (uint32_t)k + (uint32_t)i * a.plane, g[i][k], w_phys(i), a.scale);
        }
    }
    if ((sm | sp | s0) == 0u) return;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        if ((self >> k) & 1u) continue;
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            const int ex = ex_of(i), ey = ey_of(i);
            const uint32_t srow = ey == 1 ? sp : (ey == -1 ? sm : s0);
            if (!((srow >> (k + 1 + ex)) & 1u)) continue;
            int tx = x0 + k + ex;
            tx = tx < 0 ? nx - 1 : (tx >= nx ? 0 : tx);
            const uint32_t trow = ey == 1 ? rp : (ey == -1 ? rm : r0);
            Dom<T, SC>::push(dst, trow + (uint32_t)tx + (uint32_t)inv_of(i) * a.plane, g[i][k], w_phys(i),
                                a.scale);
        }
    }
}

template <typename T, bool FORCE, bool EXPORT, int BS, int V, int FLAGS>
__global__ __launch_bounds__(BS) void k_phys_step_shfl(PhysArgs a, const T* __restrict__ src, T* __restrict__ dst,
                                                       const uint32_t* __restrict__ bits, uint32_t words,
                                                       float* __restrict__ macro, uint32_t nseg)
{
    const uint32_t tile = blockIdx.x;
    const int y = (int)(tile / nseg);
    const int wbase = (int)(tile % nseg) * (BS * V) + (int)(threadIdx.x >> 6) * 64 * V;
    phys_cells_shfl<T, FORCE, EXPORT, V, FLAGS>(a, src, dst, bits, words, macro, (uint32_t)a.nx * (uint32_t)a.ny,
                                                wbase, (int)(threadIdx.x & 63u), y);
}

// Two steps per launch (temporal blocking).  A block owns a BW x BH tile of
// cells (x0, y0 its first cell; the last tiles of a row or column may be
// cut by the lattice edge).  Step 1 runs on the tile plus a one-cell ring,
// gathering from HBM exactly as the one-step kernel does, and leaves each
// cell's post-collision populations in LDS (rounded to the storage type, as
// the one-step path rounds them into its buffer).  Step 2 runs on the tile,
// gathering from LDS: a population that would come from a solid cell is the
// gathering cell's own post-collision population of the opposite direction
// (the half-way bounce-back the push rule produces), so step 1 needs no
// pushes.  With the inlet/outlet x edge (XEDGE 1) a population that would
// come from outside [0, nx) is likewise the cell's own opposite one, column
// 0 applies the Zou-He inlet, and the outlet column gathers as its west
// neighbour (whose 3x3 neighbourhood lies in the step-1 region when the
// tile holds at least two columns).  Step 2 stores to HBM with the usual
// pushes into solid neighbours, so the destination buffer ends up exactly
// as two one-step launches leave it.  HBM traffic per two steps: the ring's
// extra reads ((BW+2)(BH+2)/(BW BH) of 36 B) plus 36 B written, instead of
// 2 x 72 B.
template <int BW, int BH>
constexpr int kRegionCells = (BW + 2) * (BH + 2);

template <typename T, int XEDGE, bool FORCE, int BS, int BW, int BH, int FLAGS>
__device__ inline void two_step_tile(const PhysArgs& a, const T* __restrict__ src, T* __restrict__ dst,
                                     const uint32_t* __restrict__ bits, uint32_t words, int x0, int y0,
                                     T (*__restrict__ lds)[kRegionCells<BW, BH>])
{
    constexpr int RW = BW + 2, RC = kRegionCells<BW, BH>;
    constexpr int kIter1 = (RC + BS - 1) / BS, kIter2 = (BW * BH) / BS;
    static_assert((BW * BH) % BS == 0, "tile must be a multiple of the block");
    const int nx = a.nx, ny = a.ny;
    constexpr bool SC = kScaledDom<T>;
    // Step-2 neighbourhoods first when they cost few registers (binary32,
    // whose occupancy LDS bounds anyway): their mask loads do not depend on
    // step 1 and complete while the block computes it.
    constexpr bool kPrefetch = sizeof(T) == 4 && kIter2 <= 2;
    auto nbr2 = [&](int m) {
        const int k = (int)threadIdx.x + BS * m;
        const int x = x0 + k % BW, y = y0 + k / BW;
        CellNbr nb;
        if (x < nx && y < ny) nb = cell_nbr(a, bits, words, x, y);
        else nb.s0 = 2u;  // outside the lattice: skipped like a solid cell
        return nb;
    };
    CellNbr nb2[kPrefetch ? kIter2 : 1];
    if constexpr (kPrefetch) {
#pragma unroll
        for (int m = 0; m < kIter2; ++m) nb2[m] = nbr2(m);
    }
    // Step 1: region cells (q, r) = (x0 - 1 + q, y0 - 1 + r), y wrapped, x
    // wrapped when periodic.  With periodic x, solid cells are computed too
    // (from whatever their slots hold) instead of testing the mask: step 2
    // never reads their LDS entries, and the gather starts without waiting
    // for a mask load.
    // A tile whose region and its gather neighbourhood lie inside the lattice
    // (no wrap, no inlet or outlet column) takes the same pulls with one
    // pointer per cell and wave-uniform direction offsets (binary32; in
    // binary16 the extra path costs more occupancy than it saves).
    const bool inner = sizeof(T) == 4 && x0 >= 2 && x0 + BW + 1 <= nx - 1 && y0 >= 2 && y0 + BH + 1 <= ny - 1;
    if (inner) {
#pragma unroll
        for (int m = 0; m < kIter1; ++m) {
            const int k = (int)threadIdx.x + BS * m;
            if (k >= RC) break;
            const int r = k / RW, q = k % RW;
            const int x = x0 - 1 + q, y = y0 - 1 + r;
            if (XEDGE == 1 && (solid3(bits, words, y, x) & 2u)) continue;
            const T* const cell = src + ((uint32_t)y * a.rs + (uint32_t)x);
            float g[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                // x - e_i: row y - ey_i, column x - ex_i
                const ptrdiff_t off = (ptrdiff_t)((uint32_t)i * a.plane) - (ptrdiff_t)ey_of(i) * (ptrdiff_t)a.rs -
                                      (ptrdiff_t)ex_of(i);
                g[i] = Dom<T, SC>::in(ld<(FLAGS & kNtLoad) != 0>(cell + off), a.inv_scale);
            }
            float rho, ux, uy;
            phys_moments<FORCE, SC>(a, g, rho, ux, uy);
            phys_collide<FORCE, SC>(a, g, rho, ux, uy);
#pragma unroll
            for (int i = 0; i < 9; ++i) lds[i][k] = Dom<T, SC>::out(g[i], a.scale);
        }
    }
#pragma unroll
    for (int m = 0; m < kIter1 && !inner; ++m) {
        const int k = (int)threadIdx.x + BS * m;
        if (k >= RC) break;
        const int r = k / RW, q = k % RW;
        int x = x0 - 1 + q, y = y0 - 1 + r;
        if (x > nx || y > ny) continue;                  // beyond the ring of a cut tile
        if (XEDGE == 1 && (x < 0 || x >= nx)) continue;  // outside the channel: never read
        x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        CellNbr nb;
        nb.ym = (y == 0) ? ny - 1 : y - 1;
        nb.yp = (y == ny - 1) ? 0 : y + 1;
        nb.sm = nb.sp = 0u;
        nb.s0 = 0u;
        if (XEDGE == 1) {
            nb.s0 = solid3(bits, words, y, x);  // the outlet gathers through its west bit
            if (nb.s0 & 2u) continue;
        }
        float g[9];
        cell_gather<T, XEDGE, FLAGS, SC>(a, src, nb, x, y, g);
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g, rho, ux, uy);
        phys_collide<FORCE, SC>(a, g, rho, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) lds[i][k] = Dom<T, SC>::out(g[i], a.scale);
    }
    __syncthreads();
    // Step 2: tile cells, gathered from LDS.
#pragma unroll
    for (int m = 0; m < kIter2; ++m) {
        const int k = (int)threadIdx.x + BS * m;
        const int row = k / BW, col = k % BW;
        const int x = x0 + col, y = y0 + row;
        CellNbr nb;
        if constexpr (kPrefetch) nb = nb2[m];
        else nb = nbr2(m);
        if (nb.s0 & 2u) continue;
        float g[9];
        const bool outlet = XEDGE == 1 && x == nx - 1 && !(nb.s0 & 1u);
        if ((nb.sm | nb.sp | (nb.s0 & 5u)) == 0u && !outlet && (XEDGE == 0 || x > 0)) {
            // no solid neighbour, no edge: plain pulls from the region
            const int c = (row + 1) * RW + (col + 1);
#pragma unroll
            for (int i = 0; i < 9; ++i)
                g[i] = Dom<T, SC>::in(lds[i][c - ey_of(i) * RW - ex_of(i)], a.inv_scale);
        } else {
            // gathering cell: the outlet column gathers as its west neighbour
            int cx = x;
            CellNbr gb = nb;
            if (outlet) {
                cx = x - 1;
                gb = cell_nbr(a, bits, words, cx, y);
            }
            const int c = (row + 1) * RW + (cx - x0 + 1);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const int ex = ex_of(i), ey = ey_of(i);
                // source cell cx - e_i: bit (1 - ex) of row y - ey
                const uint32_t srow = ey == 1 ? gb.sm : (ey == -1 ? gb.sp : gb.s0);
                bool own = (srow >> (1 - ex)) & 1u;
                if (XEDGE == 1 && ex != 0) own = own || cx - ex < 0 || cx - ex >= nx;
                const T v = own ? lds[inv_of(i)][c] : lds[i][c - ey * RW - ex];
                g[i] = Dom<T, SC>::in(v, a.inv_scale);
            }
            if (XEDGE == 1 && x == 0) zou_he_inlet<SC>(a, g);
        }
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g, rho, ux, uy);
        phys_collide<FORCE, SC>(a, g, rho, ux, uy);
        cell_store<T, XEDGE, FLAGS, SC>(a, dst, nb, x, y, g);
    }
}

template <typename T, int XEDGE, bool FORCE, int BS, int BW, int BH, int FLAGS>
__global__ __launch_bounds__(BS) void k_phys_step2(PhysArgs a, const T* __restrict__ src, T* __restrict__ dst,
                                                   const uint32_t* __restrict__ bits, uint32_t words,
                                                   uint32_t tiles_x)
{
    __shared__ T lds[9][kRegionCells<BW, BH>];
    // with kXcdRemap the blocks of one XCD take a contiguous range of tiles,
    // so the ring rows and columns they share are read once into its L2
    const uint32_t t = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const int x0 = (int)(t % tiles_x) * BW;
    const int y0 = (int)(t / tiles_x) * BH;
    two_step_tile<T, XEDGE, FORCE, BS, BW, BH, FLAGS & ~kXcdRemap>(a, src, dst, bits, words, x0, y0, lds);
}

// ---------------------------------------------------------------------------
// Batched domains: one persistent launch steps every domain.
// ---------------------------------------------------------------------------

// The cells of chunk t (BS*V consecutive cells of one domain, row major).
template <typename T, int XEDGE, bool FORCE, int BS, int V, int FLAGS>
__device__ inline void batch_chunk(const BatchDomainDev& D, uint32_t t, const T* __restrict__ src,
                                   T* __restrict__ dst, const uint32_t* __restrict__ bits)
{
    const uint32_t c = (t - D.chunk0) * (uint32_t)(BS * V) + threadIdx.x * (uint32_t)V;
    if (c >= D.ncells) return;
    const int nx = D.a.nx;
    const int y = (int)(c / (uint32_t)nx), x0 = (int)(c % (uint32_t)nx);
    const T* s = src + D.pop_off;
    T* o = dst + D.pop_off;
    const uint32_t* m = bits + D.bits_off;
    if (V == 4) {
        if (x0 >= 1 && x0 + 4 <= nx - 1) {
            phys_cells4<T, FORCE, false>(D.a, s, o, m, D.words, nullptr, D.ncells, x0, y);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                phys_cell<T, XEDGE, FORCE, false, FLAGS>(D.a, s, o, m, D.words, nullptr, D.ncells, x0 + k, y);
        }
    } else {
        phys_cell<T, XEDGE, FORCE, false, FLAGS>(D.a, s, o, m, D.words, nullptr, D.ncells, x0, y);
    }
}

// Block b (after the XCD remap) steps chunks [b*per_block, (b+1)*per_block),
// or, with kGridStride, chunks b, b + grid, b + 2 grid, ... so that the
// blocks running at one time work on neighbouring chunks.
template <typename T, int BS, int V, int FLAGS>
__global__ __launch_bounds__(BS) void k_batch_step(const BatchDomainDev* __restrict__ doms,
                                                   const uint16_t* __restrict__ chunk_dom, uint32_t nchunks,
                                                   uint32_t per_block, const T* __restrict__ src, T* __restrict__ dst,
                                                   const uint32_t* __restrict__ bits)
{
    const uint32_t b = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const bool strided = (FLAGS & kGridStride) != 0;
    const uint32_t t0 = strided ? b : b * per_block;
    const uint32_t t1 = strided ? nchunks : min(t0 + per_block, nchunks);
    const uint32_t dt = strided ? gridDim.x : 1u;
    for (uint32_t t = t0; t < t1; t += dt) {
        const BatchDomainDev D = doms[chunk_dom[t]];
        if (D.x_edge) {
            if (D.force) batch_chunk<T, 1, true, BS, V, FLAGS>(D, t, src, dst, bits);
            else batch_chunk<T, 1, false, BS, V, FLAGS>(D, t, src, dst, bits);
        } else {
            if (D.force) batch_chunk<T, 0, true, BS, V, FLAGS>(D, t, src, dst, bits);
            else batch_chunk<T, 0, false, BS, V, FLAGS>(D, t, src, dst, bits);
        }
    }
}

// Two steps of every domain: block t takes tile t (after the XCD remap when
// kXcdRemap), of domain tile_dom[t], through the same tile function as the
// one-domain two-step kernel.
template <typename T, int BS, int BW, int BH, int FLAGS>
__global__ __launch_bounds__(BS) void k_batch_step2(const BatchDomainDev* __restrict__ doms,
                                                    const uint16_t* __restrict__ tile_dom, const T* __restrict__ src,
                                                    T* __restrict__ dst, const uint32_t* __restrict__ bits)
{
    __shared__ T lds[9][kRegionCells<BW, BH>];
    const uint32_t t = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const BatchDomainDev D = doms[tile_dom[t]];
    const uint32_t local = t - D.tile0;
    const uint32_t tiles_x = ((uint32_t)D.a.nx + (uint32_t)(BW - 1)) / (uint32_t)BW;
    const int x0 = (int)(local % tiles_x) * BW, y0 = (int)(local / tiles_x) * BH;
    const T* s = src + D.pop_off;
    T* o = dst + D.pop_off;
    const uint32_t* m = bits + D.bits_off;
    constexpr int F = FLAGS & ~kXcdRemap;
    if (D.x_edge) {
        if (D.force) two_step_tile<T, 1, true, BS, BW, BH, F>(D.a, s, o, m, D.words, x0, y0, lds);
        else two_step_tile<T, 1, false, BS, BW, BH, F>(D.a, s, o, m, D.words, x0, y0, lds);
    } else {
        if (D.force) two_step_tile<T, 0, true, BS, BW, BH, F>(D.a, s, o, m, D.words, x0, y0, lds);
        else two_step_tile<T, 0, false, BS, BW, BH, F>(D.a, s, o, m, D.words, x0, y0, lds);
    }
}

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
// parameters staged in LDS once per block).  One cell per thread.
template <typename T, int XEDGE, bool FORCE>
__global__ __launch_bounds__(256) void k_phys_step_learned(PhysArgs a, const T* __restrict__ src,
                                                           T* __restrict__ dst, const uint32_t* __restrict__ bits,
                                                           uint32_t words, const float* __restrict__ params,
                                                           float s_in, float s_out, uint32_t nseg)
{
    __shared__ float P[kMlpParams];
    for (int k = threadIdx.x; k < kMlpParams; k += 256) P[k] = params[k];
    __syncthreads();
    const int y = (int)(blockIdx.x / nseg);
    const int x = (int)(blockIdx.x % nseg) * 256 + (int)threadIdx.x;
    if (x >= a.nx) return;
    phys_cell<T, XEDGE, FORCE, false, kNtStore, true>(a, src, dst, bits, words, nullptr,
                                                      (uint32_t)a.nx * (uint32_t)a.ny, x, y, P, s_in, s_out);
}

// One step with the learned correction evaluated on the matrix cores
// (bf16 operands, binary32 accumulation; mlp_wave_bf16).  Blocks of 4
// wavefronts loop over row segments of 256 cells, one cell per lane, so the
// 14 KiB of packed weights are staged in LDS once per block.  Every lane
// takes part in the matrix instructions; lanes beyond the row or on solid
// cells feed zeros and store nothing.
template <typename T, int XEDGE, bool FORCE>
__global__ __launch_bounds__(256) void k_phys_step_learned_mfma(PhysArgs a, const T* __restrict__ src,
                                                                T* __restrict__ dst, const uint32_t* __restrict__ bits,
                                                                uint32_t words, const __bf16* __restrict__ packed,
                                                                const float* __restrict__ params, float s_in,
                                                                float s_out, uint32_t nseg)
{
    __shared__ __align__(16) __bf16 W[kPackElems];
    __shared__ float bias[2 * kMlpH + 16];
    __shared__ __align__(16) __bf16 xs[4][64 * 16];
    __shared__ float ys[4][64 * kMlpOut];
    {
        const uint4* from = reinterpret_cast<const uint4*>(packed);
        uint4* to = reinterpret_cast<uint4*>(W);
        for (int k = threadIdx.x; k < kPackElems / 8; k += 256) to[k] = from[k];
        for (int k = threadIdx.x; k < kMlpH; k += 256) {
            bias[k] = params[kOffB1 + k];
            bias[kMlpH + k] = params[kOffB2 + k];
        }
        if (threadIdx.x < kMlpOut) bias[2 * kMlpH + threadIdx.x] = params[kOffB3 + threadIdx.x];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t total = nseg * (uint32_t)a.ny;
    for (uint32_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
        const int y = (int)(tile / nseg);
        const int x = (int)(tile % nseg) * 256 + (int)threadIdx.x;
        bool fluid = false, learned = false;
        CellNbr nb{};
        float g[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        float rho = 1.0f, ux = 0.0f, uy = 0.0f;
        if (x < a.nx) {
            nb = cell_nbr(a, bits, words, x, y);
            fluid = !(nb.s0 & 2u);
            // trained on cells whose 3x3 neighbourhood is fluid: only those
            // get the correction
            learned = (nb.sm | nb.s0 | nb.sp) == 0u && (XEDGE == 0 || (x > 0 && x < a.nx - 1));
        }
        float xin[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (fluid) {
            cell_gather<T, XEDGE, kNtStore>(a, src, nb, x, y, g);
            phys_moments<FORCE>(a, g, rho, ux, uy);
            if (learned) cell_features(g, rho, ux, uy, s_in, xin);
        }
        bf16x8 lo, hi;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lo[i] = (__bf16)xin[i];
            hi[i] = (__bf16)(i == 0 ? xin[8] : 0.0f);
        }
        *reinterpret_cast<bf16x8*>(&xs[wave][lane * 16]) = lo;
        *reinterpret_cast<bf16x8*>(&xs[wave][lane * 16 + 8]) = hi;
        __syncthreads();
        mlp_wave_bf16(W, bias, xs[wave], ys[wave], lane);
        __syncthreads();
        if (fluid) {
            float corr[9], y9[9];
#pragma unroll
            for (int i = 0; i < 9; ++i) y9[i] = ys[wave][lane * kMlpOut + i];
            learned_correction(xin, y9, s_in, s_out, a.omega, corr);
            phys_collide<FORCE>(a, g, rho, ux, uy);
            if (learned) {
#pragma unroll
                for (int i = 0; i < 9; ++i) g[i] += corr[i];
            }
            cell_store<T, XEDGE, kNtStore>(a, dst, nb, x, y, g);
        }
    }
}

// Builds the packed solid mask from the material bytes (one 32-bit word per
// thread).
__global__ __launch_bounds__(256) void k_build_solid_bits(const uint8_t* __restrict__ mat, int nx, int ny,
                                                          int periodic_x, uint32_t words,
                                                          uint32_t* __restrict__ bits)
{
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= words * (uint32_t)ny) return;
    const uint32_t y = k / words, w = k % words;
    uint32_t v = 0;
    for (int b = 0; b < 32; ++b) {
        int x = (int)(32u * w + (uint32_t)b) - 1;
        if (x < -1 || x > nx) continue;
        if (x == -1 || x == nx) {
            if (!periodic_x) continue;
            x = (x == -1) ? nx - 1 : 0;
        }
        if (mat_solid(mat[y * (uint32_t)nx + (uint32_t)x])) v |= 1u << b;
    }
    bits[k] = v;
}

template <typename T, int XEDGE, bool FORCE>
__global__ __launch_bounds__(256) void k_phys_macro(PhysArgs a, const T* __restrict__ src,
                                                    const uint8_t* __restrict__ mat,
                                                    float* __restrict__ rho_out, float* __restrict__ ux_out,
                                                    float* __restrict__ uy_out)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const int y = (int)(c / (uint32_t)a.nx), x = (int)(c % (uint32_t)a.nx);
    if (mat_solid(mat[c])) {
        rho_out[c] = ux_out[c] = uy_out[c] = 0.0f;
        return;
    }
    float g[9], rho, ux, uy;
    phys_gather<T, XEDGE>(a, src, mat, x, y, g);
    phys_moments<FORCE>(a, g, rho, ux, uy);
    rho_out[c] = rho;
    ux_out[c] = ux;
    uy_out[c] = uy;
}

// Copies every fluid cell's f_i into slot inv(i) of its solid neighbour
// x + e_i, so that a state written from the host (or reset) carries its
// bounce-back values where the next gather expects them.
template <typename T, int XEDGE>
__global__ __launch_bounds__(256) void k_phys_prime(PhysArgs a, T* __restrict__ buf,
                                                    const uint8_t* __restrict__ mat)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n || mat_solid(mat[c])) return;
    const int nx = a.nx, ny = a.ny;
    const int y = (int)(c / (uint32_t)nx), x = (int)(c % (uint32_t)nx);
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        int tx = x + ex_of(i), ty = y + ey_of(i);
        if (ty < 0) ty += ny;
        else if (ty >= ny) ty -= ny;
        if (XEDGE == 0) {
            if (tx < 0) tx += nx;
            else if (tx >= nx) tx -= nx;
        } else if (tx < 0 || tx >= nx) {
            continue;
        }
        const uint32_t t = (uint32_t)ty * (uint32_t)nx + (uint32_t)tx;
        if (mat_solid(mat[t])) buf[pop_index(a, tx, ty, inv_of(i))] = buf[pop_index(a, x, y, i)];
    }
}

// Rest state: f = w on fluid cells (deviation 0), raw 0 on solid cells.
template <typename T>
__global__ __launch_bounds__(256) void k_phys_fill(PhysArgs a, T* __restrict__ buf,
                                                   const uint8_t* __restrict__ mat)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const bool s = mat_solid(mat[c]) != 0;
    const uint32_t p = pop_index(a, (int)(c % (uint32_t)a.nx), (int)(c / (uint32_t)a.nx), 0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (s) buf[p + (uint32_t)i * a.plane] = (T)0;
        else Store<T>::store_dev(buf, p + (uint32_t)i * a.plane, 0.0f, w_phys(i), a.scale);
    }
}

// ---------------------------------------------------------------------------
// Parity mode: the reference shaders' arithmetic, fused.
// ---------------------------------------------------------------------------

template <bool EXPORT>
__device__ inline void wgsl_cell(const WgslArgs& a, const float* __restrict__ src, float* __restrict__ dst,
                                 uint8_t* __restrict__ mat, int32_t* __restrict__ biter,
                                 const float2* __restrict__ fvel, float* __restrict__ macro, uint32_t n,
                                 int x, int y)
{
    const int nx = a.nx, ny = a.ny;
    const uint32_t c = (uint32_t)y * (uint32_t)nx + (uint32_t)x;
    const uint32_t m = mat[c];
    if (mat_solid(m)) {
        if (EXPORT) {
            macro[c] = 0.0f;
            macro[c + n] = 0.0f;
            macro[c + 2 * n] = 0.0f;
            macro[c + 3 * n] = 0.0f;
        }
        return;
    }
    // streaming_in (layout_and_fn.wgsl:38-51): single wrap on both axes.
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    float f[9];
    float vx = 0.0f, vy = 0.0f, rho = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const int sx = ex == 1 ? xm : (ex == -1 ? xp : x);
        const int sy = ey == 1 ? ym : (ey == -1 ? yp : y);
        f[i] = src[(uint32_t)sy * (uint32_t)nx + (uint32_t)sx + (uint32_t)i * a.plane];
        rho = rho + f[i];
        vx = vx + (float)ex * f[i];
        vy = vy + (float)ey * f[i];
    }
    rho = fminf(fmaxf(rho, 0.8f), 1.2f);
    vx = vx / rho;
    vy = vy / rho;
    float F[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = 0.0f;
    if (mat_accel(m)) {
        int32_t bi = biter[c];
        if (bi > 0) {
            bi = bi - 1;
            if (bi == 0) mat[c] = 1;
        }
        biter[c] = bi;
        const float2 fv = fvel[c];
        vx = fv.x * 0.5f / rho;
        vy = fv.y * 0.5f / rho;
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] = w_wgsl(i) * 3.0f * ((float)ex_of(i) * fv.x + (float)ey_of(i) * fv.y);
    }
    if (EXPORT) {
        macro[c] = vx;
        macro[c + n] = vy;
        macro[c + 2 * n] = rho;
        macro[c + 3 * n] = 1.0f;
    }
    const float usqr = 1.5f * (vx * vx + vy * vy);
    // boundary.wgsl:19 skips neighbours in the outer one-cell band.
    const bool band = x <= 0 || y <= 0 || x >= nx - 1 || y >= ny - 1;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float eu = (float)ex_of(i) * vx + (float)ey_of(i) * vy;
        const float feq = rho * w_wgsl(i) * (1.0f + 3.0f * eu + 4.5f * (eu * eu) - usqr);
        float t = f[i] - a.omega * (f[i] - feq) + F[i];
        if (t > max_wgsl(i)) t = max_wgsl(i);
        else if (t < 0.0f) t = 0.0f;
        bool pushed = false;
        if (i != 0 && !band) {
            const uint32_t tc = (uint32_t)(y + ey_of(i)) * (uint32_t)nx + (uint32_t)(x + ex_of(i));
            if (mat_solid(mat[tc])) {
                dst[tc + (uint32_t)inv_of(i) * a.plane] = t;
                pushed = true;
            }
        }
        dst[c + (uint32_t)i * a.plane] = pushed ? 0.0f : t;
    }
}

template <bool EXPORT>
__global__ __launch_bounds__(256) void k_wgsl_step(WgslArgs a, const float* __restrict__ src,
                                                   float* __restrict__ dst, uint8_t* __restrict__ mat,
                                                   int32_t* __restrict__ biter, const float2* __restrict__ fvel,
                                                   float* __restrict__ macro, uint32_t nseg)
{
    const uint32_t total = nseg * (uint32_t)a.ny;
    const uint32_t tile = xcd_remap(blockIdx.x, total);
    const int y = (int)(tile / nseg);
    const int x = (int)(tile % nseg) * 256 + (int)threadIdx.x;
    if (x >= a.nx) return;
    wgsl_cell<EXPORT>(a, src, dst, mat, biter, fvel, macro, (uint32_t)a.nx * (uint32_t)a.ny, x, y);
}

// Macro of the current state as the next collide_stream would export it,
// without the accelerate-cell countdown.
__global__ __launch_bounds__(256) void k_wgsl_macro(WgslArgs a, const float* __restrict__ src,
                                                    const uint8_t* __restrict__ mat,
                                                    const float2* __restrict__ fvel, float* __restrict__ rho_out,
                                                    float* __restrict__ ux_out, float* __restrict__ uy_out)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const int nx = a.nx, ny = a.ny;
    const int y = (int)(c / (uint32_t)nx), x = (int)(c % (uint32_t)nx);
    const uint32_t m = mat[c];
    if (mat_solid(m)) {
        rho_out[c] = ux_out[c] = uy_out[c] = 0.0f;
        return;
    }
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    float vx = 0.0f, vy = 0.0f, rho = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int ex = ex_of(i), ey = ey_of(i);
        const int sx = ex == 1 ? xm : (ex == -1 ? xp : x);
        const int sy = ey == 1 ? ym : (ey == -1 ? yp : y);
        const float f = src[(uint32_t)sy * (uint32_t)nx + (uint32_t)sx + (uint32_t)i * a.plane];
        rho = rho + f;
        vx = vx + (float)ex * f;
        vy = vy + (float)ey * f;
    }
    rho = fminf(fmaxf(rho, 0.8f), 1.2f);
    vx = vx / rho;
    vy = vy / rho;
    if (mat_accel(m)) {
        const float2 fv = fvel[c];
        vx = fv.x * 0.5f / rho;
        vy = fv.y * 0.5f / rho;
    }
    rho_out[c] = rho;
    ux_out[c] = vx;
    uy_out[c] = vy;
}

// init.wgsl:19-63 over both buffers.
__global__ __launch_bounds__(256) void k_wgsl_reset(WgslArgs a, int fluid_ty, float* __restrict__ A,
                                                    float* __restrict__ B, uint8_t* __restrict__ mat,
                                                    int32_t* __restrict__ biter, float2* __restrict__ fvel,
                                                    float* __restrict__ macro)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const uint32_t m = mat[c];
    if (mat_solid(m)) {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            A[c + (uint32_t)i * a.plane] = 0.0f;
            B[c + (uint32_t)i * a.plane] = 0.0f;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            A[c + (uint32_t)i * a.plane] = w_wgsl(i);
            B[c + (uint32_t)i * a.plane] = 0.0f;
        }
        if (fluid_ty == 0) {
            const float temp = w_wgsl(3) * 0.5f;
            A[c + 1u * a.plane] = w_wgsl(1) + temp;
            A[c + 3u * a.plane] = temp;
            B[c + 1u * a.plane] = w_wgsl(1) + temp;
            B[c + 3u * a.plane] = temp;
        }
    }
    if (mat_accel(m) && biter[c] > 0) {
        biter[c] = 0;
        mat[c] = 1;
        fvel[c] = make_float2(0.0f, 0.0f);
    }
    macro[c] = 0.0f;
    macro[c + n] = 0.0f;
    macro[c + 2 * n] = 0.0f;
    macro[c + 3 * n] = 1.0f;
}

// ---------------------------------------------------------------------------
// binary64 reductions: mass, momentum, momentum-exchange force.
// ---------------------------------------------------------------------------

// Block of 256 threads: 5 sums -> partial[blockIdx.x*5 + k], fixed order.
__device__ inline void block_sum5(double v[5], double* __restrict__ partial)
{
    __shared__ double sh[4][5];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int k = threadIdx.x;
        partial[blockIdx.x * 5 + k] = ((sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]));
    }
}

template <typename T, int XEDGE>
__global__ __launch_bounds__(256) void k_phys_reduce(PhysArgs a, const T* __restrict__ src,
                                                     const uint8_t* __restrict__ mat,
                                                     double* __restrict__ partial)
{
    const int nx = a.nx, ny = a.ny;
    const uint32_t n = (uint32_t)nx * (uint32_t)ny;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
        if (mat_solid(mat[c])) continue;
        const int y = (int)(c / (uint32_t)nx), x = (int)(c % (uint32_t)nx);
        int cx = x;
        if (XEDGE == 1 && x == nx - 1 && !mat_solid(mat[c - 1])) cx = nx - 2;
        const uint32_t cc = pop_index(a, cx, y, 0);
        double f[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            int sx = cx - ex_of(i), sy = y - ey_of(i);
            if (sy < 0) sy += ny;
            else if (sy >= ny) sy -= ny;
            bool oob = false;
            if (XEDGE == 0) {
                if (sx < 0) sx += nx;
                else if (sx >= nx) sx -= nx;
            } else {
                oob = sx < 0 || sx >= nx;
            }
            if (oob) f[i] = Store<T>::load_f64(src, cc + (uint32_t)inv_of(i) * a.plane, w_phys(i), a.inv_scale);
            else
                f[i] = Store<T>::load_f64(src, pop_index(a, sx, sy, i), w_phys(i), a.inv_scale);
        }
        if (XEDGE == 1 && x == 0) {
            const double u0 = (double)a.inlet_ux;
            const double rho = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - u0);
            f[1] = f[3] + (2.0 / 3.0) * rho * u0;
            f[5] = f[7] + 0.5 * (f[4] - f[2]) + (1.0 / 6.0) * rho * u0;
            f[8] = f[6] - 0.5 * (f[4] - f[2]) + (1.0 / 6.0) * rho * u0;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            acc[0] += f[i];
            acc[1] += (double)ex_of(i) * f[i];
            acc[2] += (double)ey_of(i) * f[i];
        }
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            int tx = x + ex_of(i), ty = y + ey_of(i);
            if (ty < 0) ty += ny;
            else if (ty >= ny) ty -= ny;
            if (XEDGE == 0) {
                if (tx < 0) tx += nx;
                else if (tx >= nx) tx -= nx;
            } else if (tx < 0 || tx >= nx) {
                continue;
            }
            const uint32_t t = (uint32_t)ty * (uint32_t)nx + (uint32_t)tx;
            if (!mat_solid(mat[t])) continue;
            const double v = Store<T>::load_f64(src, pop_index(a, tx, ty, inv_of(i)), w_phys(i), a.inv_scale);
            acc[3] += 2.0 * (double)ex_of(i) * v;
            acc[4] += 2.0 * (double)ey_of(i) * v;
        }
    }
    block_sum5(acc, partial);
}

__global__ __launch_bounds__(256) void k_wgsl_reduce(WgslArgs a, const float* __restrict__ src,
                                                     const uint8_t* __restrict__ mat,
                                                     double* __restrict__ partial)
{
    const int nx = a.nx, ny = a.ny;
    const uint32_t n = (uint32_t)nx * (uint32_t)ny;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
        if (mat_solid(mat[c])) continue;
        const int y = (int)(c / (uint32_t)nx), x = (int)(c % (uint32_t)nx);
        const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
        const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int ex = ex_of(i), ey = ey_of(i);
            const int sx = ex == 1 ? xm : (ex == -1 ? xp : x);
            const int sy = ey == 1 ? ym : (ey == -1 ? yp : y);
            const double f = (double)src[(uint32_t)sy * (uint32_t)nx + (uint32_t)sx + (uint32_t)i * a.plane];
            acc[0] += f;
            acc[1] += (double)ex * f;
            acc[2] += (double)ey * f;
        }
        const bool band = x <= 0 || y <= 0 || x >= nx - 1 || y >= ny - 1;
        if (band) continue;
#pragma unroll
        for (int i = 1; i < 9; ++i) {
            const uint32_t t = (uint32_t)(y + ey_of(i)) * (uint32_t)nx + (uint32_t)(x + ex_of(i));
            if (!mat_solid(mat[t])) continue;
            const double v = (double)src[t + (uint32_t)inv_of(i) * a.plane];
            acc[3] += 2.0 * (double)ex_of(i) * v;
            acc[4] += 2.0 * (double)ey_of(i) * v;
        }
    }
    block_sum5(acc, partial);
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ partial, int nblocks,
                                                      double* __restrict__ out)
{
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += 256) {
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] += partial[b * 5 + k];
    }
    __shared__ double sh[4][5];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const int k = threadIdx.x;
        out[k] = (sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]);
    }
}

// ---------------------------------------------------------------------------
// Format conversion and export.
// ---------------------------------------------------------------------------

// Dense host layout f[i*n + c] (binary32) <-> device storage of deviations.
template <typename T>
__global__ __launch_bounds__(256) void k_host_to_dev(PhysArgs a, const float* __restrict__ in, T* __restrict__ out)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const uint32_t p = pop_index(a, (int)(c % (uint32_t)a.nx), (int)(c / (uint32_t)a.nx), 0);
#pragma unroll
    for (int i = 0; i < 9; ++i)
        out[p + (uint32_t)i * a.plane] = Store<T>::to_storage(in[c + (uint32_t)i * n] - w_phys(i), a.scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_dev_to_host(PhysArgs a, const T* __restrict__ in, float* __restrict__ out)
{
    const uint32_t n = (uint32_t)a.nx * (uint32_t)a.ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const uint32_t p = pop_index(a, (int)(c % (uint32_t)a.nx), (int)(c / (uint32_t)a.nx), 0);
#pragma unroll
    for (int i = 0; i < 9; ++i)
        out[c + (uint32_t)i * n] = w_phys(i) + Store<T>::from_storage(in[p + (uint32_t)i * a.plane], a.inv_scale);
}

// (ux, uy, rho, w) planes -> rgba16float texels, round to nearest even.
__global__ __launch_bounds__(256) void k_pack_rgba16f(const float* __restrict__ macro, uint32_t n,
                                                      uint16_t* __restrict__ out)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const _Float16 h = (_Float16)macro[c + (uint32_t)k * n];
        out[4 * c + k] = __builtin_bit_cast(uint16_t, h);
    }
}

// assets/wgsl/lbm/curl_update.wgsl:12-34 on the macro texture: texels hold
// ux, uy as binary16; neighbours right (min(x+1, nx), y), left
// (max(x-1, 0), y), top (x, max(y-1, 0)), bottom (x, min(y+1, ny));
// curl = ((uy_right - uy_left) + ux_top) - ux_bottom, stored as
// rgba16float (curl * 3.5 + 0.5, 0, 0, 0).  The clamp to the lattice size
// (one past the last texel) makes the last column and row read outside the
// texture, whose value WGSL leaves to the implementation: here the zero
// texel.
__global__ __launch_bounds__(256) void k_curl_rgba16f(const float* __restrict__ macro, int nx, int ny,
                                                      uint16_t* __restrict__ out)
{
    const uint32_t n = (uint32_t)nx * (uint32_t)ny;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= n) return;
    const int x = (int)(c % (uint32_t)nx), y = (int)(c / (uint32_t)nx);
    auto texel = [&](int xx, int yy, int comp) -> float {
        if (xx >= nx || yy >= ny) return 0.0f;
        return (float)(_Float16)macro[(uint32_t)comp * n + (uint32_t)yy * (uint32_t)nx + (uint32_t)xx];
    };
    const float curl = ((texel(x + 1, y, 1) - texel(x > 0 ? x - 1 : 0, y, 1)) + texel(x, y > 0 ? y - 1 : 0, 0)) -
                       texel(x, y + 1, 0);
    const float r = curl * 3.5f + 0.5f;
    out[4 * c] = __builtin_bit_cast(uint16_t, (_Float16)r);
    out[4 * c + 1] = 0;
    out[4 * c + 2] = 0;
    out[4 * c + 3] = 0;
}

// Plain copy used to calibrate HBM byte counters and as the bandwidth
// ceiling a streaming kernel reaches: one 16-byte element per thread with
// non-temporal loads and stores (the fastest copy pattern measured on
// gfx950).
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const v4f* __restrict__ src, v4f* __restrict__ dst, size_t n4)
{
    const size_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + k), dst + k);
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------

static inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

template <typename T, int XEDGE, bool FORCE, bool EXPORT, int BS, int CPT, int FLAGS>
static hipError_t launch_phys_variant(const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                      uint32_t words, float* macro, hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + BS * CPT - 1) / (BS * CPT));
    const uint32_t total = nseg * (uint32_t)a.ny;
    hipLaunchKernelGGL((k_phys_step<T, XEDGE, FORCE, EXPORT, BS, CPT, FLAGS>), dim3(total), dim3(BS), 0, s, a,
                       (const T*)src, (T*)dst, bits, words, macro, nseg);
    return hipGetLastError();
}

template <typename T, int XEDGE, bool FORCE, bool EXPORT, int BS, int FLAGS>
static hipError_t launch_phys_variant4(const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                       uint32_t words, float* macro, hipStream_t s)
{
    if (a.nx % 4 != 0) return hipErrorInvalidValue;
    const uint32_t nseg = (uint32_t)((a.nx + BS * 4 - 1) / (BS * 4));
    const uint32_t total = nseg * (uint32_t)a.ny;
    hipLaunchKernelGGL((k_phys_step4<T, XEDGE, FORCE, EXPORT, BS, FLAGS>), dim3(total), dim3(BS), 0, s, a,
                       (const T*)src, (T*)dst, bits, words, macro, nseg);
    return hipGetLastError();
}

template <typename T, int XEDGE, bool FORCE, bool EXPORT, int BS, int V, int FLAGS>
static hipError_t launch_phys_variant_shfl(const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                           uint32_t words, float* macro, hipStream_t s)
{
    if constexpr (XEDGE != 0) {
        return hipErrorInvalidValue;
    } else {
        // whole blocks of BS*V cells per row: every lane's V-wide access and
        // both edge loads stay inside the row
        if (a.nx % (BS * V) != 0) return hipErrorInvalidValue;
        const uint32_t nseg = (uint32_t)(a.nx / (BS * V));
        hipLaunchKernelGGL((k_phys_step_shfl<T, FORCE, EXPORT, BS, V, FLAGS>), dim3(nseg * (uint32_t)a.ny), dim3(BS),
                           0, s, a, (const T*)src, (T*)dst, bits, words, macro, nseg);
        return hipGetLastError();
    }
}

// Whether the two-step tiles of width bw fit an nx x ny lattice: the ring
// wraps in y (and in x when periodic), so nx, ny >= 3, and with the
// inlet/outlet edge the last tile of a row holds at least two columns (the
// outlet gathers through its west neighbour's ring).
static bool two_step_fits(int bw, int nx, int ny, int xedge)
{
    return nx >= 3 && ny >= 3 && (xedge == 0 || nx % bw != 1);
}

template <typename T, int XEDGE, bool FORCE, int BS, int BW, int BH, int FLAGS>
static hipError_t launch_phys_variant2(const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                       uint32_t words, hipStream_t s)
{
    if (!two_step_fits(BW, a.nx, a.ny, XEDGE)) return hipErrorInvalidValue;
    const uint32_t tiles_x = (uint32_t)((a.nx + BW - 1) / BW);
    const uint32_t tiles = tiles_x * (uint32_t)((a.ny + BH - 1) / BH);
    hipLaunchKernelGGL((k_phys_step2<T, XEDGE, FORCE, BS, BW, BH, FLAGS>), dim3(tiles), dim3(BS), 0, s, a,
                       (const T*)src, (T*)dst, bits, words, tiles_x);
    return hipGetLastError();
}

// Tile shapes the step kernel is built for: threads per block, cells per
// thread, consecutive cells per lane (4 needs nx % 4 == 0) and the kNt* /
// kXcdRemap flags.  Every shape computes the same bits.
const StepVariant kStepVariants[] = {
    {256, 1, 1, 0},
    {256, 1, 1, kNtStore},
    {256, 1, 1, kNtStore | kNtLoad},
    {256, 1, 1, kXcdRemap},
    {256, 2, 1, kNtStore},
    {128, 1, 1, kNtStore},
    {512, 1, 1, kNtStore},
    {256, 4, 1, kNtStore},
    {256, 1, 1, kNtLoad},
    {64, 4, 4, kNtStore},
    {128, 4, 4, 0},
    {1024, 1, 1, kNtStore},
    {256, 1, 4, kNtStore | kShfl},
    {128, 1, 4, kNtStore | kShfl},
    {256, 1, 2, kNtStore | kShfl},
    {256, 1, 8, kNtStore | kShfl},
    {64, 1, 8, kNtStore | kShfl},
    {512, 1, 4, kNtStore | kShfl},
    // two steps per launch: (block, tile rows, tile width, flags)
    {256, 8, 128, kNtStore | kTwoStep},
    {256, 16, 64, kNtStore | kTwoStep},
    {256, 8, 256, kNtStore | kTwoStep},
    {256, 4, 256, kNtStore | kTwoStep},
    {512, 16, 64, kNtStore | kTwoStep},
    {1024, 16, 64, kNtStore | kTwoStep},
    {512, 8, 128, kNtStore | kTwoStep},
    {256, 8, 64, kNtStore | kTwoStep},
    {512, 32, 64, kNtStore | kTwoStep},
    {256, 16, 32, kNtStore | kTwoStep},
    {512, 8, 128, kNtStore | kTwoStep | kXcdRemap},
    {256, 16, 64, kNtStore | kTwoStep | kXcdRemap},
    {512, 16, 64, kNtStore | kTwoStep | kXcdRemap},
    {1024, 16, 64, kNtStore | kTwoStep | kXcdRemap},
    {256, 8, 128, kNtStore | kTwoStep | kXcdRemap},
};
const int kNumStepVariants = (int)(sizeof(kStepVariants) / sizeof(kStepVariants[0]));

bool step_variant_ok(int v, int nx, int ny, int xedge)
{
    if (v < 0 || v >= kNumStepVariants) return false;
    const StepVariant& sv = kStepVariants[v];
    if (sv.flags & kTwoStep) return two_step_fits(sv.vec, nx, ny, xedge);
    if (sv.flags & kShfl) return xedge == 0 && nx % (sv.block * sv.vec) == 0;
    return sv.vec == 1 || nx % sv.vec == 0;
}

template <typename T, int XEDGE, bool FORCE, bool EXPORT>
static hipError_t dispatch_variant(int v, const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                   uint32_t words, float* macro, hipStream_t s)
{
    constexpr int S = kNtStore, L = kNtLoad, R = kXcdRemap, H = kShfl;
#define LBM_V(BS, CPT, FL) return launch_phys_variant<T, XEDGE, FORCE, EXPORT, BS, CPT, FL>(a, src, dst, bits, words, macro, s)
#define LBM_V4(BS, FL) return launch_phys_variant4<T, XEDGE, FORCE, EXPORT, BS, FL>(a, src, dst, bits, words, macro, s)
#define LBM_VS(BS, V, FL) \
    return launch_phys_variant_shfl<T, XEDGE, FORCE, EXPORT, BS, V, FL>(a, src, dst, bits, words, macro, s)
    switch (v) {
    case 0: LBM_V(256, 1, 0);
    case 1: LBM_V(256, 1, S);
    case 2: LBM_V(256, 1, S | L);
    case 3: LBM_V(256, 1, R);
    case 4: LBM_V(256, 2, S);
    case 5: LBM_V(128, 1, S);
    case 6: LBM_V(512, 1, S);
    case 7: LBM_V(256, 4, S);
    case 8: LBM_V(256, 1, L);
    case 9: LBM_V4(64, S);
    case 10: LBM_V4(128, 0);
    case 11: LBM_V(1024, 1, S);
    case 12: LBM_VS(256, 4, S | H);
    case 13: LBM_VS(128, 4, S | H);
    case 14: LBM_VS(256, 2, S | H);
    case 15: LBM_VS(256, 8, S | H);
    case 16: LBM_VS(64, 8, S | H);
    case 17: LBM_VS(512, 4, S | H);
    default: return hipErrorInvalidValue;
    }
#undef LBM_V
#undef LBM_V4
#undef LBM_VS
}

template <typename T, int XEDGE, bool FORCE>
static hipError_t dispatch_variant2(int v, const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                    uint32_t words, hipStream_t s)
{
    constexpr int S = kNtStore, R = kXcdRemap;
    switch (v) {
    case 18: return launch_phys_variant2<T, XEDGE, FORCE, 256, 128, 8, S>(a, src, dst, bits, words, s);
    case 19: return launch_phys_variant2<T, XEDGE, FORCE, 256, 64, 16, S>(a, src, dst, bits, words, s);
    case 20: return launch_phys_variant2<T, XEDGE, FORCE, 256, 256, 8, S>(a, src, dst, bits, words, s);
    case 21: return launch_phys_variant2<T, XEDGE, FORCE, 256, 256, 4, S>(a, src, dst, bits, words, s);
    case 22: return launch_phys_variant2<T, XEDGE, FORCE, 512, 64, 16, S>(a, src, dst, bits, words, s);
    case 23: return launch_phys_variant2<T, XEDGE, FORCE, 1024, 64, 16, S>(a, src, dst, bits, words, s);
    case 24: return launch_phys_variant2<T, XEDGE, FORCE, 512, 128, 8, S>(a, src, dst, bits, words, s);
    case 25: return launch_phys_variant2<T, XEDGE, FORCE, 256, 64, 8, S>(a, src, dst, bits, words, s);
    case 26: return launch_phys_variant2<T, XEDGE, FORCE, 512, 64, 32, S>(a, src, dst, bits, words, s);
    case 27: return launch_phys_variant2<T, XEDGE, FORCE, 256, 32, 16, S>(a, src, dst, bits, words, s);
    case 28: return launch_phys_variant2<T, XEDGE, FORCE, 512, 128, 8, S | R>(a, src, dst, bits, words, s);
    case 29: return launch_phys_variant2<T, XEDGE, FORCE, 256, 64, 16, S | R>(a, src, dst, bits, words, s);
    case 30: return launch_phys_variant2<T, XEDGE, FORCE, 512, 64, 16, S | R>(a, src, dst, bits, words, s);
    case 31: return launch_phys_variant2<T, XEDGE, FORCE, 1024, 64, 16, S | R>(a, src, dst, bits, words, s);
    case 32: return launch_phys_variant2<T, XEDGE, FORCE, 256, 128, 8, S | R>(a, src, dst, bits, words, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_phys_step2(int variant, bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                             void* dst, const uint32_t* bits, uint32_t words, hipStream_t s)
{
#define LBM_D2(T, XE, FO) return dispatch_variant2<T, XE, FO>(variant, a, src, dst, bits, words, s)
    if (f16) {
        if (xedge) { if (force) LBM_D2(_Float16, 1, true); else LBM_D2(_Float16, 1, false); }
        else { if (force) LBM_D2(_Float16, 0, true); else LBM_D2(_Float16, 0, false); }
    }
    if (xedge) { if (force) LBM_D2(float, 1, true); else LBM_D2(float, 1, false); }
    if (force) LBM_D2(float, 0, true);
    LBM_D2(float, 0, false);
#undef LBM_D2
}

template <typename T, int XEDGE>
static hipError_t dispatch_force(int v, bool force, bool exp, const PhysArgs& a, const void* src, void* dst,
                                 const uint32_t* bits, uint32_t words, float* macro, hipStream_t s)
{
    if (force)
        return exp ? dispatch_variant<T, XEDGE, true, true>(v, a, src, dst, bits, words, macro, s)
                   : dispatch_variant<T, XEDGE, true, false>(v, a, src, dst, bits, words, macro, s);
    return exp ? dispatch_variant<T, XEDGE, false, true>(v, a, src, dst, bits, words, macro, s)
               : dispatch_variant<T, XEDGE, false, false>(v, a, src, dst, bits, words, macro, s);
}

hipError_t launch_phys_step(int variant, bool f16, int xedge, bool force, bool export_macro, const PhysArgs& a,
                            const void* src, void* dst, const uint32_t* bits, uint32_t words, float* macro,
                            hipStream_t s)
{
    if (f16)
        return xedge ? dispatch_force<_Float16, 1>(variant, force, export_macro, a, src, dst, bits, words, macro, s)
                     : dispatch_force<_Float16, 0>(variant, force, export_macro, a, src, dst, bits, words, macro, s);
    return xedge ? dispatch_force<float, 1>(variant, force, export_macro, a, src, dst, bits, words, macro, s)
                 : dispatch_force<float, 0>(variant, force, export_macro, a, src, dst, bits, words, macro, s);
}

// Batched step shapes: (threads per block, consecutive cells per lane,
// persistent blocks per CU).  A chunk is block*vec cells.
const BatchVariant kBatchVariants[] = {
    {256, 1, 8, kNtStore}, {256, 1, 8, kNtStore | kXcdRemap}, {512, 1, 4, kNtStore}, {256, 1, 8, 0},
    {64, 4, 16, kNtStore}, {128, 4, 8, 0},
    {256, 1, 8, kNtStore | kGridStride}, {512, 1, 4, kNtStore | kGridStride}, {1024, 1, 2, kNtStore | kGridStride},
    {256, 1, 16, kNtStore | kGridStride},
    // two steps per launch: (block, tile width, -, flags, tile rows)
    {256, 64, 0, kNtStore | kTwoStep, 16}, {512, 64, 0, kNtStore | kTwoStep, 16},
    {512, 128, 0, kNtStore | kTwoStep, 8}, {256, 32, 0, kNtStore | kTwoStep, 16},
    {256, 64, 0, kNtStore | kTwoStep, 8}, {256, 64, 0, kNtStore | kTwoStep | kXcdRemap, 16},
    {512, 64, 0, kNtStore | kTwoStep | kXcdRemap, 16}, {512, 128, 0, kNtStore | kTwoStep | kXcdRemap, 8},
    {256, 32, 0, kNtStore | kTwoStep | kXcdRemap, 16}, {1024, 64, 0, kNtStore | kTwoStep | kXcdRemap, 16},
};
const int kNumBatchVariants = (int)(sizeof(kBatchVariants) / sizeof(kBatchVariants[0]));

template <typename T>
static hipError_t launch_batch_typed(int v, const BatchDomainDev* doms, const uint16_t* chunk_dom, uint32_t nchunks,
                                     uint32_t grid, const void* src, void* dst, const uint32_t* bits, hipStream_t s)
{
    const uint32_t per_block = (nchunks + grid - 1) / grid;
#define LBM_BATCH(BS, V, FL)                                                                                   \
    hipLaunchKernelGGL((k_batch_step<T, BS, V, FL>), dim3(grid), dim3(BS), 0, s, doms, chunk_dom, nchunks,  \
                       per_block, (const T*)src, (T*)dst, bits)
    switch (v) {
    case 0: LBM_BATCH(256, 1, kNtStore); break;
    case 1: LBM_BATCH(256, 1, kNtStore | kXcdRemap); break;
    case 2: LBM_BATCH(512, 1, kNtStore); break;
    case 3: LBM_BATCH(256, 1, 0); break;
    case 4: LBM_BATCH(64, 4, kNtStore); break;
    case 5: LBM_BATCH(128, 4, 0); break;
    case 6: LBM_BATCH(256, 1, kNtStore | kGridStride); break;
    case 7: LBM_BATCH(512, 1, kNtStore | kGridStride); break;
    case 8: LBM_BATCH(1024, 1, kNtStore | kGridStride); break;
    case 9: LBM_BATCH(256, 1, kNtStore | kGridStride); break;
    default: return hipErrorInvalidValue;
    }
#undef LBM_BATCH
    return hipGetLastError();
}

hipError_t launch_batch_step(int variant, bool f16, const BatchDomainDev* doms, const uint16_t* chunk_dom,
                             uint32_t nchunks, uint32_t grid, const void* src, void* dst, const uint32_t* bits,
                             hipStream_t s)
{
    if (nchunks == 0) return hipSuccess;
    return f16 ? launch_batch_typed<_Float16>(variant, doms, chunk_dom, nchunks, grid, src, dst, bits, s)
               : launch_batch_typed<float>(variant, doms, chunk_dom, nchunks, grid, src, dst, bits, s);
}

template <typename T>
static hipError_t launch_batch2_typed(int v, const BatchDomainDev* doms, const uint16_t* tile_dom, uint32_t ntiles,
                                      const void* src, void* dst, const uint32_t* bits, hipStream_t s)
{
#define LBM_BATCH2(BS, BW, BH, FL)                                                                            \
    hipLaunchKernelGGL((k_batch_step2<T, BS, BW, BH, FL>), dim3(ntiles), dim3(BS), 0, s, doms, tile_dom,   \
                       (const T*)src, (T*)dst, bits)
    switch (v) {
    case 10: LBM_BATCH2(256, 64, 16, kNtStore); break;
    case 11: LBM_BATCH2(512, 64, 16, kNtStore); break;
    case 12: LBM_BATCH2(512, 128, 8, kNtStore); break;
    case 13: LBM_BATCH2(256, 32, 16, kNtStore); break;
    case 14: LBM_BATCH2(256, 64, 8, kNtStore); break;
    case 15: LBM_BATCH2(256, 64, 16, kNtStore | kXcdRemap); break;
    case 16: LBM_BATCH2(512, 64, 16, kNtStore | kXcdRemap); break;
    case 17: LBM_BATCH2(512, 128, 8, kNtStore | kXcdRemap); break;
    case 18: LBM_BATCH2(256, 32, 16, kNtStore | kXcdRemap); break;
    case 19: LBM_BATCH2(1024, 64, 16, kNtStore | kXcdRemap); break;
    default: return hipErrorInvalidValue;
    }
#undef LBM_BATCH2
    return hipGetLastError();
}

hipError_t launch_batch_step2(int variant, bool f16, const BatchDomainDev* doms, const uint16_t* tile_dom,
                              uint32_t ntiles, const void* src, void* dst, const uint32_t* bits, hipStream_t s)
{
    if (ntiles == 0) return hipSuccess;
    return f16 ? launch_batch2_typed<_Float16>(variant, doms, tile_dom, ntiles, src, dst, bits, s)
               : launch_batch2_typed<float>(variant, doms, tile_dom, ntiles, src, dst, bits, s);
}

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
                                   float s_out, hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + 255) / 256);
    const dim3 grid(nseg * (uint32_t)a.ny), block(256);
#define LBM_L(T, XE, FO)                                                                                  \
    hipLaunchKernelGGL((k_phys_step_learned<T, XE, FO>), grid, block, 0, s, a, (const T*)src, (T*)dst, bits, \
                       words, params, s_in, s_out, nseg)
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
                                        void* dst, const uint32_t* bits, uint32_t words, const __bf16* packed,
                                        const float* params, float s_in, float s_out, hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + 255) / 256);
    const uint32_t tiles = nseg * (uint32_t)a.ny;
    const dim3 grid(tiles < 2048u ? tiles : 2048u), block(256);
#define LBM_L(T, XE, FO)                                                                                       \
    hipLaunchKernelGGL((k_phys_step_learned_mfma<T, XE, FO>), grid, block, 0, s, a, (const T*)src, (T*)dst, bits, \
                       words, packed, params, s_in, s_out, nseg)
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

hipError_t launch_build_solid_bits(const uint8_t* mat, int nx, int ny, int periodic_x, uint32_t words,
                                  uint32_t* bits, hipStream_t s)
{
    hipLaunchKernelGGL(k_build_solid_bits, dim3(blocks_for((size_t)words * ny, 256)), dim3(256), 0, s, mat, nx, ny,
                       periodic_x, words, bits);
    return hipGetLastError();
}

hipError_t launch_phys_macro(bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                             const uint8_t* mat, float* rho, float* ux, float* uy, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
#define LBM_MACRO(T, XE, FO) \
    hipLaunchKernelGGL((k_phys_macro<T, XE, FO>), dim3(nb), dim3(256), 0, s, a, (const T*)src, mat, rho, ux, uy)
    if (f16) {
        if (xedge) { if (force) LBM_MACRO(_Float16, 1, true); else LBM_MACRO(_Float16, 1, false); }
        else { if (force) LBM_MACRO(_Float16, 0, true); else LBM_MACRO(_Float16, 0, false); }
    } else {
        if (xedge) { if (force) LBM_MACRO(float, 1, true); else LBM_MACRO(float, 1, false); }
        else { if (force) LBM_MACRO(float, 0, true); else LBM_MACRO(float, 0, false); }
    }
#undef LBM_MACRO
    return hipGetLastError();
}

hipError_t launch_phys_prime(bool f16, int xedge, const PhysArgs& a, void* buf, const uint8_t* mat,
                             hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
    if (f16) {
        if (xedge) hipLaunchKernelGGL((k_phys_prime<_Float16, 1>), dim3(nb), dim3(256), 0, s, a, (_Float16*)buf, mat);
        else hipLaunchKernelGGL((k_phys_prime<_Float16, 0>), dim3(nb), dim3(256), 0, s, a, (_Float16*)buf, mat);
    } else {
        if (xedge) hipLaunchKernelGGL((k_phys_prime<float, 1>), dim3(nb), dim3(256), 0, s, a, (float*)buf, mat);
        else hipLaunchKernelGGL((k_phys_prime<float, 0>), dim3(nb), dim3(256), 0, s, a, (float*)buf, mat);
    }
    return hipGetLastError();
}

hipError_t launch_phys_fill(bool f16, const PhysArgs& a, void* buf, const uint8_t* mat, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
    if (f16) hipLaunchKernelGGL((k_phys_fill<_Float16>), dim3(nb), dim3(256), 0, s, a, (_Float16*)buf, mat);
    else hipLaunchKernelGGL((k_phys_fill<float>), dim3(nb), dim3(256), 0, s, a, (float*)buf, mat);
    return hipGetLastError();
}

hipError_t launch_phys_reduce(bool f16, int xedge, const PhysArgs& a, const void* src, const uint8_t* mat,
                              double* partial, int nblocks, double* out, hipStream_t s)
{
    if (f16) {
        if (xedge) hipLaunchKernelGGL((k_phys_reduce<_Float16, 1>), dim3(nblocks), dim3(256), 0, s, a, (const _Float16*)src, mat, partial);
        else hipLaunchKernelGGL((k_phys_reduce<_Float16, 0>), dim3(nblocks), dim3(256), 0, s, a, (const _Float16*)src, mat, partial);
    } else {
        if (xedge) hipLaunchKernelGGL((k_phys_reduce<float, 1>), dim3(nblocks), dim3(256), 0, s, a, (const float*)src, mat, partial);
        else hipLaunchKernelGGL((k_phys_reduce<float, 0>), dim3(nblocks), dim3(256), 0, s, a, (const float*)src, mat, partial);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_wgsl_step(bool export_macro, const WgslArgs& a, const float* src, float* dst, uint8_t* mat,
                            int32_t* biter, const float2* fvel, float* macro, hipStream_t s)
{
    const uint32_t nseg = (uint32_t)((a.nx + 255) / 256);
    const uint32_t total = nseg * (uint32_t)a.ny;
    if (export_macro)
        hipLaunchKernelGGL((k_wgsl_step<true>), dim3(total), dim3(256), 0, s, a, src, dst, mat, biter, fvel, macro, nseg);
    else
        hipLaunchKernelGGL((k_wgsl_step<false>), dim3(total), dim3(256), 0, s, a, src, dst, mat, biter, fvel, macro, nseg);
    return hipGetLastError();
}

hipError_t launch_wgsl_macro(const WgslArgs& a, const float* src, const uint8_t* mat, const float2* fvel,
                             float* rho, float* ux, float* uy, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
    hipLaunchKernelGGL(k_wgsl_macro, dim3(nb), dim3(256), 0, s, a, src, mat, fvel, rho, ux, uy);
    return hipGetLastError();
}

hipError_t launch_wgsl_reset(const WgslArgs& a, int fluid_ty, float* A, float* B, uint8_t* mat, int32_t* biter,
                             float2* fvel, float* macro, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
    hipLaunchKernelGGL(k_wgsl_reset, dim3(nb), dim3(256), 0, s, a, fluid_ty, A, B, mat, biter, fvel, macro);
    return hipGetLastError();
}

hipError_t launch_wgsl_reduce(const WgslArgs& a, const float* src, const uint8_t* mat, double* partial,
                              int nblocks, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_wgsl_reduce, dim3(nblocks), dim3(256), 0, s, a, src, mat, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_host_to_dev(bool f16, const PhysArgs& a, const float* in, void* out, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
    if (f16)
        hipLaunchKernelGGL(k_host_to_dev<_Float16>, dim3(nb), dim3(256), 0, s, a, in, (_Float16*)out);
    else
        hipLaunchKernelGGL(k_host_to_dev<float>, dim3(nb), dim3(256), 0, s, a, in, (float*)out);
    return hipGetLastError();
}

hipError_t launch_dev_to_host(bool f16, const PhysArgs& a, const void* in, float* out, hipStream_t s)
{
    const unsigned nb = blocks_for((size_t)a.nx * a.ny, 256);
    if (f16)
        hipLaunchKernelGGL(k_dev_to_host<_Float16>, dim3(nb), dim3(256), 0, s, a, (const _Float16*)in, out);
    else
        hipLaunchKernelGGL(k_dev_to_host<float>, dim3(nb), dim3(256), 0, s, a, (const float*)in, out);
    return hipGetLastError();
}

hipError_t launch_pack_rgba16f(const float* macro, uint32_t n, uint16_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_rgba16f, dim3(blocks_for(n, 256)), dim3(256), 0, s, macro, n, out);
    return hipGetLastError();
}

hipError_t launch_curl_rgba16f(const float* macro, int nx, int ny, uint16_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_curl_rgba16f, dim3(blocks_for((size_t)nx * ny, 256)), dim3(256), 0, s, macro, nx, ny, out);
    return hipGetLastError();
}

hipError_t launch_copy16(const void* src, void* dst, size_t bytes, hipStream_t s)
{
    const size_t n4 = bytes / 16;
    const size_t blocks = (n4 + 255) / 256;
    if (blocks == 0 || blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, s, (const v4f*)src, (v4f*)dst, n4);
    return hipGetLastError();
}

}  // namespace lbm
