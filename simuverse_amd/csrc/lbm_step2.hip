// lbm_step2.hip -- two steps per launch (temporal blocking through an LDS
// tile) for one domain, and the batched kernels that step many domains in
// one launch (one step per persistent chunk, or two per tile), with their
// launchers and shape tables.
#include "lbm_cell.hpp"
#include "lbm_device.hpp"
#include "lbm_internal.hpp"

namespace lbm {

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
__device__ inline void two_step_tile(const PhysArgs& a_in, const T* __restrict__ src, T* __restrict__ dst,
                                     const uint32_t* __restrict__ bits, uint32_t words, int x0, int y0,
                                     T (*__restrict__ lds)[kRegionCells<BW, BH>])
{
    PhysArgs a = a_in;
    if (FORCE) pin_pair_consts<kScaledDom<T>>(a);
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
    // pointer per cell and wave-uniform direction offsets.
    const bool inner = x0 >= 2 && x0 + BW + 1 <= nx - 1 && y0 >= 2 && y0 + BH + 1 <= ny - 1;
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

// a if pick, else b, as a bit merge under a mask the compiler cannot see
// through: both values stay separate LDS reads at fixed offsets from one
// address, instead of one read from a selected address per direction,
// which the compiler would then share with the plain pulls (nine address
// additions for every cell; with the merge, three, one per ring slot).
template <typename T>
__device__ __forceinline__ T lds_pick(bool pick, T a, T b)
{
    static_assert(sizeof(T) == 2 || sizeof(T) == 4, "16- or 32-bit element");
    using U = typename std::conditional<sizeof(T) == 2, uint16_t, uint32_t>::type;
    uint32_t m = pick ? 0xFFFFFFFFu : 0u;
    asm volatile("" : "+v"(m));
    const uint32_t ua = (uint32_t)__builtin_bit_cast(U, a), ub = (uint32_t)__builtin_bit_cast(U, b);
    return __builtin_bit_cast(T, (U)((ua & m) | (ub & ~m)));
}

// Two steps per launch, streamed down a strip of rows.  A block of BS
// threads owns the columns [x0, x0 + BS) of the rows [y0, y0 + H) (the last
// strip and chunk may be cut by the lattice edge) and walks down them one
// row at a time.  Step 1 of region row r -- lattice row y0 - 1 + r, the
// strip's columns plus one on each side -- gathers from HBM as the tile
// kernel's step 1 does and leaves its post-collision populations in slot
// r mod 4 of an LDS ring of four rows; after one barrier, step 2 of lattice
// row y0 + r - 2 gathers from the slots of region rows r - 2 .. r with the
// tile kernel's step-2 rules (own opposite population from a solid source
// or from outside the channel, the outlet gathering as its west neighbour,
// the Zou-He inlet) and stores to HBM with the pushes into solid
// neighbours, so the destination ends up exactly as two one-step launches
// leave it.  Four slots make the one barrier enough: step 1 of row r + 1,
// which a wave may start while another still runs step 2 of row r - 2,
// writes the one slot that step reads none of.
// Against the tile kernel: a step-1 gather or a step-2 store of one wave
// covers 64 consecutive cells of one direction row (128 B of binary16, a
// 1 KiB run per block row at 512 threads) instead of cells from two region
// rows, every direction row of the chunk is read once (the tile re-reads
// its ring), and the ring adds two cells per row of BS instead of
// 2 (BW + BH) + 4 per tile.  Step 1's two ring cells are computed by lanes 0
// and 1 of the first wave.
template <typename T, int XEDGE, bool FORCE, int BS, int H, int FLAGS>
__global__ __launch_bounds__(BS) void k_phys_step2_rows(PhysArgs a, const T* __restrict__ src, T* __restrict__ dst,
                                                        const uint32_t* __restrict__ bits, uint32_t words,
                                                        uint32_t strips)
{
    constexpr int RW = BS + 2;
    constexpr bool SC = kScaledDom<T>;
    __shared__ T lds[4][9][RW];
    const uint32_t t = (FLAGS & kXcdRemap) ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const int x0 = (int)(t % strips) * BS, y0 = (int)(t / strips) * H;
    const int nx = a.nx, ny = a.ny;
    const int rows = min(H, ny - y0);
    const int tid = (int)threadIdx.x;
    // The wave's step-1 cells x0 + 64w .. + 63 and their gathers lie inside
    // the lattice and away from the inlet and outlet columns (wave-uniform).
    const int xw = x0 + __builtin_amdgcn_readfirstlane(tid & ~63);
    const bool wave_inner = xw >= 1 && xw + 65 <= nx;

    // Step 1 of one cell (x, y), unwrapped, into column q of slot `slot`.
    auto step1_general = [&](int x, int y, int q, int slot) {
        if (x > nx) return;                              // beyond the ring of a cut strip
        if (XEDGE == 1 && (x < 0 || x >= nx)) return;    // outside the channel: never read
        x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        CellNbr nb;
        nb.ym = (y == 0) ? ny - 1 : y - 1;
        nb.yp = (y == ny - 1) ? 0 : y + 1;
        nb.sm = nb.sp = 0u;
        nb.s0 = 0u;
        if (XEDGE == 1) {
            nb.s0 = solid3(bits, words, y, x);  // the outlet gathers through its west bit
            if (nb.s0 & 2u) return;
        }
        float g[9];
        cell_gather<T, XEDGE, FLAGS, SC>(a, src, nb, x, y, g);
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g, rho, ux, uy);
        phys_collide<FORCE, SC>(a, g, rho, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; ++i) lds[slot][i][q] = Dom<T, SC>::out(g[i], a.scale);
    };
    auto step1_row = [&](int r) {
        const int y = y0 - 1 + r, slot = r & 3;
        const int x = x0 + tid;
        if (wave_inner && y >= 1 && y + 2 <= ny) {
            // no wrap, no edge column: one pointer per cell, wave-uniform
            // direction offsets (periodic x computes solid cells too: step 2
            // never reads their entries)
            if (!(XEDGE == 1 && (solid3(bits, words, y, x) & 2u))) {
                const T* const cell = src + ((uint32_t)y * a.rs + (uint32_t)x);
                float g[9];
#pragma unroll
                for (int i = 0; i < 9; ++i) {
                    const ptrdiff_t off = (ptrdiff_t)((uint32_t)i * a.plane) -
                                          (ptrdiff_t)ey_of(i) * (ptrdiff_t)a.rs - (ptrdiff_t)ex_of(i);
                    g[i] = Dom<T, SC>::in(ld<(FLAGS & kNtLoad) != 0>(cell + off), a.inv_scale);
                }
                float rho, ux, uy;
                phys_moments<FORCE, SC>(a, g, rho, ux, uy);
                phys_collide<FORCE, SC>(a, g, rho, ux, uy);
#pragma unroll
                for (int i = 0; i < 9; ++i) lds[slot][i][tid + 1] = Dom<T, SC>::out(g[i], a.scale);
            }
        } else {
            step1_general(x, y, tid + 1, slot);
        }
        if (tid < 2) step1_general(tid == 0 ? x0 - 1 : x0 + BS, y, tid == 0 ? 0 : BS + 1, slot);
    };
    // Step 2 of lattice row y0 + j from the slots of region rows j .. j + 2.
    auto step2_row = [&](int j) {
        const int x = x0 + tid, y = y0 + j;
        if (x >= nx) return;
        const CellNbr nb = cell_nbr(a, bits, words, x, y);
        if (nb.s0 & 2u) return;
        float g[9];
        const bool outlet = XEDGE == 1 && x == nx - 1 && !(nb.s0 & 1u);
        if ((nb.sm | nb.sp | (nb.s0 & 5u)) == 0u && !outlet && (XEDGE == 0 || x > 0)) {
            // no solid neighbour, no edge: plain pulls from the ring
#pragma unroll
            for (int i = 0; i < 9; ++i)
                g[i] = Dom<T, SC>::in(lds[(j + 1 - ey_of(i)) & 3][i][tid + 1 - ex_of(i)], a.inv_scale);
        } else {
            // gathering cell: the outlet column gathers as its west neighbour
            int cx = x;
            CellNbr gb = nb;
            if (outlet) {
                cx = x - 1;
                gb = cell_nbr(a, bits, words, cx, y);
            }
            const int c = cx - x0 + 1;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const int ex = ex_of(i), ey = ey_of(i);
                // source cell cx - e_i: bit (1 - ex) of row y - ey
                const uint32_t srow = ey == 1 ? gb.sm : (ey == -1 ? gb.sp : gb.s0);
                bool own = (srow >> (1 - ex)) & 1u;
                if (XEDGE == 1 && ex != 0) own = own || cx - ex < 0 || cx - ex >= nx;
                const T pulled = lds[(j + 1 - ey) & 3][i][c - ex];
                const T v = i == 0 ? pulled : lds_pick(own, lds[(j + 1) & 3][inv_of(i)][c], pulled);
                g[i] = Dom<T, SC>::in(v, a.inv_scale);
            }
            if (XEDGE == 1 && x == 0) zou_he_inlet<SC>(a, g);
        }
        float rho, ux, uy;
        phys_moments<FORCE, SC>(a, g, rho, ux, uy);
        phys_collide<FORCE, SC>(a, g, rho, ux, uy);
        cell_store<T, XEDGE, FLAGS, SC>(a, dst, nb, x, y, g);
    };

    step1_row(0);
    step1_row(1);
    for (int j = 0; j < rows; ++j) {
        step1_row(j + 2);
        __syncthreads();
        step2_row(j);
    }
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

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

// Whether the two- and three-step tiles of width bw fit an nx x ny
// lattice: the ring wraps in y (and in x when periodic), so nx, ny >= 3,
// and with the inlet/outlet edge the last tile of a row holds at least two
// columns (the outlet gathers through its west neighbour's ring).  The
// three-step tile forms row offsets with 24-bit products, so rows, the row
// stride (at most 9 (nx + 16) elements) and the mask words stay below 2^24.
bool two_step_fits(int bw, int nx, int ny, int xedge)
{
    return nx >= 3 && ny >= 3 && (xedge == 0 || nx % bw != 1) && 9ll * (nx + 16) < (1ll << 24) &&
           ny < (1 << 24);
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

template <typename T, int XEDGE, bool FORCE, int BS, int H, int FLAGS>
static hipError_t launch_phys_rows(const PhysArgs& a, const void* src, void* dst, const uint32_t* bits,
                                   uint32_t words, hipStream_t s)
{
    if (!two_step_fits(BS, a.nx, a.ny, XEDGE)) return hipErrorInvalidValue;
    const uint32_t strips = (uint32_t)((a.nx + BS - 1) / BS);
    const uint32_t blocks = strips * (uint32_t)((a.ny + H - 1) / H);
    hipLaunchKernelGGL((k_phys_step2_rows<T, XEDGE, FORCE, BS, H, FLAGS>), dim3(blocks), dim3(BS), 0, s, a,
                       (const T*)src, (T*)dst, bits, words, strips);
    return hipGetLastError();
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
    case 33: return launch_phys_variant2<T, XEDGE, FORCE, 512, 64, 32, S | R>(a, src, dst, bits, words, s);
    case 34: return launch_phys_rows<T, XEDGE, FORCE, 512, 128, S | R>(a, src, dst, bits, words, s);
    case 35: return launch_phys_rows<T, XEDGE, FORCE, 512, 64, S | R>(a, src, dst, bits, words, s);
    case 36: return launch_phys_rows<T, XEDGE, FORCE, 256, 64, S | R>(a, src, dst, bits, words, s);
    case 37: return launch_phys_rows<T, XEDGE, FORCE, 512, 128, S>(a, src, dst, bits, words, s);
    case 38: return launch_phys_rows<T, XEDGE, FORCE, 1024, 128, S | R>(a, src, dst, bits, words, s);
    case 39: return launch_phys_rows<T, XEDGE, FORCE, 512, 256, S | R>(a, src, dst, bits, words, s);
    case 40: return launch_phys_rows<T, XEDGE, FORCE, 256, 128, S | R>(a, src, dst, bits, words, s);
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
    {512, 64, 0, kNtStore | kTwoStep, 32}, {512, 64, 0, kNtStore | kTwoStep | kXcdRemap, 32},
    {1024, 64, 0, kNtStore | kTwoStep | kXcdRemap, 32},
    // three steps per launch (launch_batch_step3): (block, tile width, -, flags, tile rows)
    {512, 64, 0, kNtStore | kThreeStep, 16}, {512, 64, 0, kNtStore | kThreeStep | kXcdRemap, 16},
    {256, 32, 0, kNtStore | kThreeStep, 16}, {256, 32, 0, kNtStore | kThreeStep | kXcdRemap, 16},
    {256, 64, 0, kNtStore | kThreeStep, 8},
    {768, 64, 0, kNtStore | kThreeStep, 24}, {768, 64, 0, kNtStore | kThreeStep | kXcdRemap, 24},
    {1024, 64, 0, kNtStore | kThreeStep, 32},
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
    case 20: LBM_BATCH2(512, 64, 32, kNtStore); break;
    case 21: LBM_BATCH2(512, 64, 32, kNtStore | kXcdRemap); break;
    case 22: LBM_BATCH2(1024, 64, 32, kNtStore | kXcdRemap); break;
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

}  // namespace lbm
