// lbm_kernels.hip -- gfx950 kernels of the D2Q9 hot path: the one-step
// physics kernels (one cell, four cells and lane-shuffle shapes), the
// parity-mode (reference semantics) kernels, materials, macro export,
// reductions, format conversion, and the launchers of all of them.  The
// per-cell functions shared with the two-step and learned kernels are in
// lbm_cell.hpp; the two-step and batched kernels in lbm_step2.hip; the
// learned, restriction and health kernels in lbm_learned.hip.
//
// Layout in HBM: physics mode keeps the nine direction rows of one lattice
// row adjacent (element y*rs + i*plane + x, plane = row pitch) or one plane
// per direction (plane >= nx*ny); parity mode keeps the reference's planes.
// Two buffers ping-pong.  Materials are one byte per cell in the
// reference's LatticeType numbering.
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
#include "lbm_cell.hpp"
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {

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
    {512, 32, 64, kNtStore | kTwoStep | kXcdRemap},
    // two steps per launch streamed down strips: (block, chunk rows, strip width = block, flags)
    {512, 128, 512, kNtStore | kTwoStep | kRowStream | kXcdRemap},
    {512, 64, 512, kNtStore | kTwoStep | kRowStream | kXcdRemap},
    {256, 64, 256, kNtStore | kTwoStep | kRowStream | kXcdRemap},
    {512, 128, 512, kNtStore | kTwoStep | kRowStream},
    {1024, 128, 1024, kNtStore | kTwoStep | kRowStream | kXcdRemap},
    {512, 256, 512, kNtStore | kTwoStep | kRowStream | kXcdRemap},
    {256, 128, 256, kNtStore | kTwoStep | kRowStream | kXcdRemap},
    // three steps per launch: (block, tile rows, tile width, flags)
    {512, 16, 64, kNtStore | kThreeStep},
    {512, 16, 64, kNtStore | kThreeStep | kXcdRemap},
    {512, 8, 128, kNtStore | kThreeStep},
    {256, 16, 32, kNtStore | kThreeStep},
    {256, 8, 64, kNtStore | kThreeStep},
    {768, 24, 64, kNtStore | kThreeStep | kXcdRemap},
    {768, 24, 64, kNtStore | kThreeStep},
};
const int kNumStepVariants = (int)(sizeof(kStepVariants) / sizeof(kStepVariants[0]));

bool step_variant_ok(int v, int nx, int ny, int xedge)
{
    if (v < 0 || v >= kNumStepVariants) return false;
    const StepVariant& sv = kStepVariants[v];
    if (sv.flags & (kTwoStep | kThreeStep)) return two_step_fits(sv.vec, nx, ny, xedge);
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
