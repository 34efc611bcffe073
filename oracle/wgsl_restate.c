/*
 * wgsl_restate.c -- CPU restatement, in binary32, of simuverse's three LBM
 * compute shaders.  TEST INFRASTRUCTURE ONLY: tests, the smoke check and the
 * CPU-baseline leg of the benchmark call it as the checker; the product
 * never links it.
 *
 * It follows the reference literally: the same constants, the same
 * evaluation order, the same two passes (collide_stream, then boundary) over
 * ping-pong buffers, and the same AoS LatticeInfo records.  Compile with
 * -ffp-contract=off: no operation may be fused that the WGSL writes as two.
 *
 * Reference files (jinleili/simuverse):
 *   constants           simuverse/src/fluid/mod.rs:32-55
 *   LatticeInfo         simuverse/src/fluid/lattice.rs:5-11
 *   predicates          assets/wgsl/lbm/d2q9_fn.wgsl:5-26
 *   pull gather         assets/wgsl/lbm/layout_and_fn.wgsl:38-51
 *   collide_stream      assets/wgsl/lbm/collide_stream.wgsl:18-22, 25-88
 *   boundary            assets/wgsl/lbm/boundary.wgsl:3-35
 *   init                assets/wgsl/lbm/init.wgsl:19-63
 *   ping-pong / swap    simuverse/src/fluid/d2q9_node.rs:128-135, 302-312
 *
 * Pinned (tests/test_oracle_wgsl.py) by: the literal constants, free
 * streaming of single pulses along the documented direction layout, the
 * half-way bounce-back delay, the band and wrap rules derived from
 * boundary.wgsl, the closed-form mass decay caused by the truncated weights,
 * both clamps, the accelerate-cell velocity override and countdown, and the
 * viscous decay rate of a shear wave.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

typedef struct {
    int32_t material;
    int32_t block_iter;
    float vx;
    float vy;
} ws_info;

/* LbmUniform::new (mod.rs:39-52): ex, ey, w, max as binary32 literals. */
static const float WS_EX[9] = {0.0f, 1.0f, 0.0f, -1.0f, 0.0f, 1.0f, -1.0f, -1.0f, 1.0f};
static const float WS_EY[9] = {0.0f, 0.0f, -1.0f, 0.0f, 1.0f, -1.0f, -1.0f, 1.0f, 1.0f};
static const float WS_W[9] = {0.444444f, 0.111111f, 0.111111f, 0.111111f, 0.111111f,
                              0.0277777f, 0.0277777f, 0.0277777f, 0.0277777f};
static const float WS_MAX[9] = {0.6f, 0.2222f, 0.2222f, 0.2222f, 0.2222f,
                                0.1111f, 0.1111f, 0.1111f, 0.1111f};
static const int WS_INV[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

/* d2q9_fn.wgsl:19-24 */
static int ws_is_solid(int32_t m) { return m == 2 || m == 4; }
static int ws_is_accel(int32_t m) { return m == 3 || m == 6; }

float ws_weight(int i) { return WS_W[i]; }
float ws_max_value(int i) { return WS_MAX[i]; }
float ws_omega(float tau) { return 1.0f / tau; } /* mod.rs:35 */

/* init.wgsl:19-63.  A is the buffer bound at binding 2 (buffer 0), B the one
 * at binding 3 (buffer 1); macro4 receives (ux, uy, rho, w) per cell. */
void ws_init(int nx, int ny, int fluid_ty, ws_info* info, float* A, float* B, float* macro4)
{
    const size_t n = (size_t)nx * (size_t)ny;
    for (size_t c = 0; c < n; ++c) {
        ws_info in = info[c];
        if (ws_is_solid(in.material)) {
            for (int i = 0; i < 9; ++i) {
                A[c + (size_t)i * n] = 0.0f;
                B[c + (size_t)i * n] = 0.0f;
            }
        } else if (fluid_ty == 0) {
            for (int i = 0; i < 9; ++i) {
                A[c + (size_t)i * n] = WS_W[i];
                B[c + (size_t)i * n] = 0.0f;
            }
            const float temp = WS_W[3] * 0.5f;
            A[c + 1 * n] = WS_W[1] + temp;
            A[c + 3 * n] = temp;
            B[c + 1 * n] = WS_W[1] + temp;
            B[c + 3 * n] = temp;
        } else {
            for (int i = 0; i < 9; ++i) {
                A[c + (size_t)i * n] = WS_W[i];
                B[c + (size_t)i * n] = 0.0f;
            }
        }
        if (ws_is_accel(in.material) && in.block_iter > 0) {
            in.block_iter = 0;
            in.material = 1;
            in.vx = 0.0f;
            in.vy = 0.0f;
            info[c] = in;
        }
        if (macro4) {
            macro4[4 * c + 0] = 0.0f;
            macro4[4 * c + 1] = 0.0f;
            macro4[4 * c + 2] = 0.0f;
            macro4[4 * c + 3] = 1.0f;
        }
    }
}

/* layout_and_fn.wgsl:38-51: source of direction i for the cell (x, y) is
 * (x, y) + e(inv(i)) = (x - ex_i, y - ey_i), wrapped once per axis. */
static size_t ws_stream_in(int nx, int ny, int x, int y, int i)
{
    int tx = x - (int)WS_EX[i];
    int ty = y - (int)WS_EY[i];
    if (tx < 0) tx = nx - 1;
    else if (tx >= nx) tx = 0;
    if (ty < 0) ty = ny - 1;
    else if (ty >= ny) ty = 0;
    return (size_t)tx + (size_t)ty * (size_t)nx + (size_t)i * (size_t)nx * (size_t)ny;
}

/* One cell of collide_stream.wgsl:25-88. */
static void ws_cell(int nx, int ny, float omega, ws_info* info, const float* src, float* dst,
                    float* macro4, int x, int y)
{
    const size_t n = (size_t)nx * (size_t)ny;
    const size_t c = (size_t)x + (size_t)y * (size_t)nx;
    ws_info in = info[c];
    if (ws_is_solid(in.material)) {
        if (macro4) {
            macro4[4 * c + 0] = 0.0f;
            macro4[4 * c + 1] = 0.0f;
            macro4[4 * c + 2] = 0.0f;
            macro4[4 * c + 3] = 0.0f;
        }
        return;
    }
    float f[9];
    float vx = 0.0f, vy = 0.0f, rho = 0.0f;
    for (int i = 0; i < 9; ++i) {
        f[i] = src[ws_stream_in(nx, ny, x, y, i)];
        rho = rho + f[i];
        vx = vx + WS_EX[i] * f[i];
        vy = vy + WS_EY[i] * f[i];
    }
    rho = fminf(fmaxf(rho, 0.8f), 1.2f); /* clamp(rho, 0.8, 1.2) */
    vx = vx / rho;
    vy = vy / rho;

    float F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (ws_is_accel(in.material)) {
        if (in.block_iter > 0) {
            in.block_iter = in.block_iter - 1;
            if (in.block_iter == 0) in.material = 1;
        }
        info[c] = in;
        const float fx = in.vx, fy = in.vy;
        vx = fx * 0.5f / rho;
        vy = fy * 0.5f / rho;
        for (int i = 0; i < 9; ++i) {
            const float d = WS_EX[i] * fx + WS_EY[i] * fy;
            F[i] = WS_W[i] * 3.0f * d;
        }
    }
    if (macro4) {
        macro4[4 * c + 0] = vx;
        macro4[4 * c + 1] = vy;
        macro4[4 * c + 2] = rho;
        macro4[4 * c + 3] = 1.0f;
    }
    const float usqr = 1.5f * (vx * vx + vy * vy);
    for (int i = 0; i < 9; ++i) {
        const float eu = WS_EX[i] * vx + WS_EY[i] * vy;
        const float feq = rho * WS_W[i] * (1.0f + 3.0f * eu + 4.5f * (eu * eu) - usqr);
        float t = f[i] - omega * (f[i] - feq) + F[i];
        if (t > WS_MAX[i]) t = WS_MAX[i];
        else if (t < 0.0f) t = 0.0f;
        dst[c + (size_t)i * n] = t;
    }
}

void ws_collide_stream(int nx, int ny, float omega, ws_info* info, const float* src, float* dst,
                       float* macro4)
{
    /* Every cell reads src and writes only its own dst slots and its own
     * info record, so the rows are independent. */
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) ws_cell(nx, ny, omega, info, src, dst, macro4, x, y);
}

/* boundary.wgsl:3-35 over dst.  The reference runs one GPU thread per
 * solid cell; threads of neighbouring solid cells race on solid-cell slots
 * that no fluid cell ever reads.  Raster order is one legal interleaving. */
void ws_boundary(int nx, int ny, const ws_info* info, float* dst)
{
    const size_t n = (size_t)nx * (size_t)ny;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t c = (size_t)x + (size_t)y * (size_t)nx;
            if (!ws_is_solid(info[c].material)) continue;
            for (int i = 0; i < 9; ++i) {
                const int nxp = x - (int)WS_EX[i];
                const int nyp = y - (int)WS_EY[i];
                if (nxp <= 0 || nyp <= 0 || nxp >= nx - 1 || nyp >= ny - 1) continue;
                const size_t nb = (size_t)nxp + (size_t)nyp * (size_t)nx + (size_t)i * n;
                const float v = dst[nb];
                dst[c + (size_t)WS_INV[i] * n] = v;
                dst[nb] = 0.0f;
            }
        }
}

/* D2Q9Node::compute_by_pass(swap): collide_stream then boundary on the
 * bind group `swap` (0: buf0 -> buf1, 1: buf1 -> buf0). */
void ws_compute_by_pass(int nx, int ny, float omega, ws_info* info, float* buf0, float* buf1,
                        float* macro4, int swap)
{
    float* src = swap ? buf1 : buf0;
    float* dst = swap ? buf0 : buf1;
    ws_collide_stream(nx, ny, omega, info, src, dst, macro4);
    ws_boundary(nx, ny, info, dst);
}

/* nsteps steps starting with bind group start_swap, alternating. */
void ws_run(int nx, int ny, float tau, ws_info* info, float* buf0, float* buf1, float* macro4,
            int nsteps, int start_swap)
{
    const float omega = ws_omega(tau);
    int swap = start_swap;
    for (int s = 0; s < nsteps; ++s) {
        ws_compute_by_pass(nx, ny, omega, info, buf0, buf1, macro4, swap);
        swap ^= 1;
    }
}
