// lbm_step3.hip -- three steps per launch for one domain (temporal blocking
// through an LDS tile with a two-cell ring), its launcher and shape table.
#include "lbm_cell.hpp"
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {

// Three steps per launch.  A block owns a BW x BH tile of cells (x0, y0 its
// first cell; the last tiles of a row or column may be cut by the lattice
// edge).  Step 1 runs on the tile plus a two-cell ring, gathering from HBM
// as the one-step kernel does, into LDS (region 1: (BW + 4) x (BH + 4)
// cells, rounded to the storage type as a one-step launch rounds them into
// its buffer).  Step 2 runs on the tile plus a one-cell ring (region 2),
// gathering from region 1 with the two-step kernel's rules: a population
// from a solid cell is the gathering cell's own opposite post-collision
// population (the half-way bounce-back the push rule produces), one from
// outside [0, nx) with the inlet/outlet edge likewise, column 0 applies the
// Zou-He inlet, and the outlet column gathers as its west neighbour.  Its
// results are held in registers across a barrier and then written over
// region 1 in region-2 layout, so one LDS buffer serves both.  Step 3 runs
// on the tile from region 2 and stores to HBM with the pushes into solid
// neighbours.  The destination ends up as three one-step launches leave it
// (at every slot a fluid cell can read), and in the same buffer: one launch
// is one buffer flip, as three are.  HBM traffic per three steps: the
// ring's extra reads ((BW + 4)(BH + 4) / (BW BH) of 36 B) plus 36 B written.
template <int BW, int BH, int R>
constexpr int kRegion = (BW + 2 * R) * (BH + 2 * R);

template <typename T, int XEDGE, bool FORCE, int BS, int BW, int BH, int FLAGS, bool MUL24 = true>
__device__ inline void three_step_tile(const PhysArgs& a_in, const T* __restrict__ src, T* __restrict__ dst,
                                       const uint32_t* __restrict__ bits, uint32_t words, int x0, int y0,
                                       T (*__restrict__ lds)[kRegion<BW, BH, 2>])
{
    PhysArgs a = a_in;
    if (FORCE) pin_pair_consts<kScaledDom<T>>(a);
    constexpr int RW1 = BW + 4, RC1 = kRegion<BW, BH, 2>;
    constexpr int RW2 = BW + 2, RC2 = kRegion<BW, BH, 1>;
    constexpr int kIter1 = (RC1 + BS - 1) / BS, kIter2 = (RC2 + BS - 1) / BS, kIter3 = (BW * BH) / BS;
    static_assert((BW * BH) % BS == 0, "tile must be a multiple of the block");
    constexpr bool SC = kScaledDom<T>;
    const int nx = a.nx, ny = a.ny;

    // Step 1: region-1 cells (q, r) = (x0 - 2 + q, y0 - 2 + r), y wrapped, x
    // wrapped when periodic.  With periodic x, solid cells are computed too
    // (step 2 never reads their entries).  A tile whose region and its
    // gather neighbourhood lie inside the lattice (no wrap, no inlet or
    // outlet column) pulls with one pointer per cell and wave-uniform
    // direction offsets.
    const bool inner = x0 >= 3 && x0 + BW + 3 <= nx && y0 >= 3 && y0 + BH + 3 <= ny;
#pragma unroll
    for (int m = 0; m < kIter1; ++m) {
        const int k = (int)threadIdx.x + BS * m;
        if (k >= RC1) break;
        const int r = k / RW1, q = k % RW1;
        int x = x0 - 2 + q, y = y0 - 2 + r;
        float g[9];
        if (inner) {
            if (XEDGE == 1 && (solid3(bits, words, y, x) & 2u)) continue;
            const uint32_t row = MUL24 ? __umul24((uint32_t)y, a.rs) : (uint32_t)y * a.rs;
            const T* const cell = src + (row + (uint32_t)x);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const ptrdiff_t off = (ptrdiff_t)((uint32_t)i * a.plane) - (ptrdiff_t)ey_of(i) * (ptrdiff_t)a.rs -
                                      (ptrdiff_t)ex_of(i);
                g[i] = Dom<T, SC>::in(ld<(FLAGS & kNtLoad) != 0>(cell + off), a.inv_scale);
            }
        } else {
            if (x > nx + 1 || y > ny + 1) continue;          // beyond the ring of a cut tile
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
            cell_gather<T, XEDGE, FLAGS, SC>(a, src, nb, x, y, g);
        }
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g, rho, ux, uy);
        phys_collide<FORCE, SC>(a, g, rho, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) lds[i][k] = Dom<T, SC>::out(g[i], a.scale);
    }
    __syncthreads();

    // Gather of the fluid cell (x, y) (wrapped; its LDS entry at index c of a
    // region whose rows are RW wide) from that region, with the edge and
    // bounce-back rules.
    auto gather_lds = [&](int RW, int c, int x, int y, const CellNbr& nb, float g[9]) {
        const bool outlet = XEDGE == 1 && x == nx - 1 && !(nb.s0 & 1u);
        if ((nb.sm | nb.sp | (nb.s0 & 5u)) == 0u && !outlet && (XEDGE == 0 || x > 0)) {
#pragma unroll
            for (int i = 0; i < 9; ++i) g[i] = Dom<T, SC>::in(lds[i][c - ey_of(i) * RW - ex_of(i)], a.inv_scale);
            return;
        }
        // gathering cell: the outlet column gathers as its west neighbour
        int cx = x, cc = c;
        CellNbr gb = nb;
        if (outlet) {
            cx = x - 1;
            cc = c - 1;
            gb = cell_nbr(a, bits, words, cx, y);
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int ex = ex_of(i), ey = ey_of(i);
            // source cell cx - e_i: bit (1 - ex) of row y - ey
            const uint32_t srow = ey == 1 ? gb.sm : (ey == -1 ? gb.sp : gb.s0);
            bool own = (srow >> (1 - ex)) & 1u;
            if (XEDGE == 1 && ex != 0) own = own || cx - ex < 0 || cx - ex >= nx;
            const T v = own ? lds[inv_of(i)][cc] : lds[i][cc - ey * RW - ex];
            g[i] = Dom<T, SC>::in(v, a.inv_scale);
        }
        if (XEDGE == 1 && x == 0) zou_he_inlet<SC>(a, g);
    };

    // Step 2: region-2 cells (q, r) = (x0 - 1 + q, y0 - 1 + r), computed
    // into registers, then written in region-2 layout after every read of
    // region 1 is done.
    float g2[kIter2][9];
    bool live[kIter2];
#pragma unroll
    for (int m = 0; m < kIter2; ++m) {
        const int k = (int)threadIdx.x + BS * m;
        live[m] = false;
        if (k >= RC2) continue;
        const int r = k / RW2, q = k % RW2;
        int x = x0 - 1 + q, y = y0 - 1 + r;
        CellNbr nb;
        if (inner) {
            // region 2 and its neighbours lie inside the lattice, away from
            // the inlet and outlet columns: no wrap, no cut
            nb = cell_nbr_inner<MUL24>(bits, words, x, y);
        } else {
            if (x > nx || y > ny) continue;                  // beyond the ring of a cut tile
            if (XEDGE == 1 && (x < 0 || x >= nx)) continue;  // outside the channel: never read
            x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
            y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
            nb = cell_nbr(a, bits, words, x, y);
        }
        if (nb.s0 & 2u) continue;  // solid: step 3 never reads its entries
        gather_lds(RW1, (r + 1) * RW1 + (q + 1), x, y, nb, g2[m]);
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g2[m], rho, ux, uy);
        phys_collide<FORCE, SC>(a, g2[m], rho, ux, uy);
        live[m] = true;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kIter2; ++m) {
        if (!live[m]) continue;
        const int k = (int)threadIdx.x + BS * m;
#pragma unroll
        for (int i = 0; i < 9; ++i) lds[i][k] = Dom<T, SC>::out(g2[m][i], a.scale);
    }
    __syncthreads();

    // Step 3: tile cells from region 2, stored to HBM with the pushes.
#pragma unroll
    for (int m = 0; m < kIter3; ++m) {
        const int k = (int)threadIdx.x + BS * m;
        const int row = k / BW, col = k % BW;
        const int x = x0 + col, y = y0 + row;
        CellNbr nb;
        if (inner) {
            nb = cell_nbr_inner<MUL24>(bits, words, x, y);
        } else {
            if (x >= nx || y >= ny) continue;
            nb = cell_nbr(a, bits, words, x, y);
        }
        if (nb.s0 & 2u) continue;
        float g[9];
        gather_lds(RW2, (row + 1) * RW2 + (col + 1), x, y, nb, g);
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g, rho, ux, uy);
        phys_collide<FORCE, SC>(a, g, rho, ux, uy);
        const uint32_t own = MUL24 ? __umul24((uint32_t)y, a.rs) : (uint32_t)y * a.rs;
        cell_store_at<T, XEDGE, FLAGS, SC>(a, dst, nb, x, y, own + (uint32_t)x, g);
    }
}

template <typename T, int XEDGE, bool FORCE, int BS, int BW, int BH, int FLAGS>
__global__ __launch_bounds__(BS) void k_phys_step3(PhysArgs a, const T* __restrict__ src, T* __restrict__ dst,
                                                   const uint32_t* __restrict__ bits, uint32_t words,
                                                   uint32_t tiles_x)
{
    __shared__ T lds[9][kRegion<BW, BH, 2>];
    // with kXcdRemap the blocks of one XCD take a contiguous range of tiles
    const uint32_t t = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const int x0 = (int)(t % tiles_x) * BW;
    const int y0 = (int)(t / tiles_x) * BH;
    three_step_tile<T, XEDGE, FORCE, BS, BW, BH, FLAGS & ~kXcdRemap>(a, src, dst, bits, words, x0, y0, lds);
}

template <typename T, int XEDGE, bool FORCE, int BS, int BW, int BH, int FLAGS>
static hipError_t launch_phys_variant3(const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                       uint32_t words, hipStream_t s)
{
    if (!two_step_fits(BW, a.nx, a.ny, XEDGE)) return hipErrorInvalidValue;
    const uint32_t tiles_x = (uint32_t)((a.nx + BW - 1) / BW);
    const uint32_t tiles = tiles_x * (uint32_t)((a.ny + BH - 1) / BH);
    hipLaunchKernelGGL((k_phys_step3<T, XEDGE, FORCE, BS, BW, BH, FLAGS>), dim3(tiles), dim3(BS), 0, s, a,
                       (const T*)src, (T*)dst, bits, words, tiles_x);
    return hipGetLastError();
}

template <typename T, int XEDGE, bool FORCE>
static hipError_t dispatch_variant3(int v, const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                    uint32_t words, hipStream_t s)
{
    constexpr int S = kNtStore, R = kXcdRemap;
    switch (v) {
    case 41: return launch_phys_variant3<T, XEDGE, FORCE, 512, 64, 16, S>(a, src, dst, bits, words, s);
    case 42: return launch_phys_variant3<T, XEDGE, FORCE, 512, 64, 16, S | R>(a, src, dst, bits, words, s);
    case 43: return launch_phys_variant3<T, XEDGE, FORCE, 512, 128, 8, S>(a, src, dst, bits, words, s);
    case 44: return launch_phys_variant3<T, XEDGE, FORCE, 256, 32, 16, S>(a, src, dst, bits, words, s);
    case 45: return launch_phys_variant3<T, XEDGE, FORCE, 256, 64, 8, S>(a, src, dst, bits, words, s);
    case 46: return launch_phys_variant3<T, XEDGE, FORCE, 768, 64, 24, S | R>(a, src, dst, bits, words, s);
    case 47: return launch_phys_variant3<T, XEDGE, FORCE, 768, 64, 24, S>(a, src, dst, bits, words, s);
    default: return hipErrorInvalidValue;
    }
}

// The region-1 LDS buffer in binary32: 64 x 16: 48 960 B (3 blocks of 512
// threads per CU), 128 x 8: 57 024 B, 64 x 24: 68 544 B (2 blocks of 768).
hipError_t launch_phys_step3(int variant, bool f16, int xedge, bool force, const PhysArgs& a, const void* src,
                             void* dst, const uint32_t* bits, uint32_t words, hipStream_t s)
{
#define LBM_D3(T, XE, FO) return dispatch_variant3<T, XE, FO>(variant, a, src, dst, bits, words, s)
    if (f16) {
        if (xedge) { if (force) LBM_D3(_Float16, 1, true); else LBM_D3(_Float16, 1, false); }
        else { if (force) LBM_D3(_Float16, 0, true); else LBM_D3(_Float16, 0, false); }
    }
    if (xedge) { if (force) LBM_D3(float, 1, true); else LBM_D3(float, 1, false); }
    if (force) LBM_D3(float, 0, true);
    LBM_D3(float, 0, false);
#undef LBM_D3
}

// ---------------------------------------------------------------------------
// Batched domains, three steps per launch: one block per tile of one domain
// (tile_dom[t]; domain d owns tiles [tile0, tile0 + ceil(nx/BW) ceil(ny/BH))),
// each domain with its own size, tau, inlet, force and x edge.
// ---------------------------------------------------------------------------
template <typename T, int BS, int BW, int BH, int FLAGS>
__global__ __launch_bounds__(BS) void k_batch_step3(const BatchDomainDev* __restrict__ doms,
                                                    const uint16_t* __restrict__ tile_dom, const T* __restrict__ src,
                                                    T* __restrict__ dst, const uint32_t* __restrict__ bits)
{
    __shared__ T lds[9][kRegion<BW, BH, 2>];
    const uint32_t t = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const BatchDomainDev D = doms[tile_dom[t]];
    const uint32_t local = t - D.tile0;
    const uint32_t tiles_x = ((uint32_t)D.a.nx + (uint32_t)(BW - 1)) / (uint32_t)BW;
    const int x0 = (int)(local % tiles_x) * BW, y0 = (int)(local / tiles_x) * BH;
    const T* s = src + D.pop_off;
    T* o = dst + D.pop_off;
    const uint32_t* m = bits + D.bits_off;
    constexpr int F = FLAGS & ~kXcdRemap;
    // 24-bit row products: measured 1.7 % faster batched in binary32 and
    // 34 % slower for the binary16 64 x 32 tile of 1 024 threads (same
    // registers, no spills; cause not found), so binary16 keeps the 32-bit
    // products here
    constexpr bool M24 = sizeof(T) == 4;
    if (D.x_edge) {
        if (D.force) three_step_tile<T, 1, true, BS, BW, BH, F, M24>(D.a, s, o, m, D.words, x0, y0, lds);
        else three_step_tile<T, 1, false, BS, BW, BH, F, M24>(D.a, s, o, m, D.words, x0, y0, lds);
    } else {
        if (D.force) three_step_tile<T, 0, true, BS, BW, BH, F, M24>(D.a, s, o, m, D.words, x0, y0, lds);
        else three_step_tile<T, 0, false, BS, BW, BH, F, M24>(D.a, s, o, m, D.words, x0, y0, lds);
    }
}

template <typename T>
static hipError_t launch_batch3_typed(int v, const BatchDomainDev* doms, const uint16_t* tile_dom, uint32_t ntiles,
                                      const void* src, void* dst, const uint32_t* bits, hipStream_t s)
{
    if (ntiles == 0) return hipSuccess;
#define LBM_BATCH3(BS, BW, BH, FL)                                                                             \
    hipLaunchKernelGGL((k_batch_step3<T, BS, BW, BH, FL>), dim3(ntiles), dim3(BS), 0, s, doms, tile_dom,    \
                       (const T*)src, (T*)dst, bits)
    switch (v) {
    case 23: LBM_BATCH3(512, 64, 16, kNtStore); break;
    case 24: LBM_BATCH3(512, 64, 16, kNtStore | kXcdRemap); break;
    case 25: LBM_BATCH3(256, 32, 16, kNtStore); break;
    case 26: LBM_BATCH3(256, 32, 16, kNtStore | kXcdRemap); break;
    case 27: LBM_BATCH3(256, 64, 8, kNtStore); break;
    case 28: LBM_BATCH3(768, 64, 24, kNtStore); break;
    case 29: LBM_BATCH3(768, 64, 24, kNtStore | kXcdRemap); break;
    case 30: LBM_BATCH3(1024, 64, 32, kNtStore); break;
    default: return hipErrorInvalidValue;
    }
#undef LBM_BATCH3
    return hipGetLastError();
}

hipError_t launch_batch_step3(int variant, bool f16, const BatchDomainDev* doms, const uint16_t* tile_dom,
                              uint32_t ntiles, const void* src, void* dst, const uint32_t* bits, hipStream_t s)
{
    return f16 ? launch_batch3_typed<_Float16>(variant, doms, tile_dom, ntiles, src, dst, bits, s)
               : launch_batch3_typed<float>(variant, doms, tile_dom, ntiles, src, dst, bits, s);
}

}  // namespace lbm
