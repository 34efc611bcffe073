/*
 * phys_oracle.c -- binary64 CPU oracle of the physics-mode D2Q9 solver.
 * TEST INFRASTRUCTURE ONLY: tests, the smoke check and the CPU-baseline leg
 * of the benchmark call it as the checker; the product never links it.
 *
 * The reference (jinleili/simuverse) has no physics-mode solver: its shaders
 * use truncated weights and clamps, and it has no Zou-He inlet, no
 * zero-gradient outlet and no body force (assets/wgsl/lbm/collide_stream.wgsl:
 * 49,55-71,79-83).  This oracle states the textbook scheme the physics mode
 * implements, written in the plain pull form:
 *
 *   state    post-collision populations f*_i(x) of the fluid cells
 *   stream   f_i(x, t+1) = f*_i(x - e_i, t), and when x - e_i is solid (or
 *            outside a non-periodic x range) the half-way bounce-back
 *            f_i(x, t+1) = f*_inv(i)(x, t)
 *   inlet    Zou & He, Phys. Fluids 9 (1997) 1591: velocity (u0, 0) on the
 *            fluid cells of column 0; unknowns f1, f5, f8
 *   outlet   zero gradient: column nx-1 takes the post-streaming
 *            populations of column nx-2 (when that cell is fluid)
 *   collide  BGK with Guo, Zheng & Shi, Phys. Rev. E 65 (2002) 046308
 *            forcing: u = (sum e f + g/2)/rho,
 *            F_i = (1 - omega/2) w_i [3(e_i - u) + 9 (e_i.u) e_i] . g
 *   drag     momentum exchange, Ladd, J. Fluid Mech. 271 (1994) 285:
 *            sum over links (x fluid, x + e_i solid) of 2 e_i f*_i(x)
 *
 * Pinned (tests/test_oracle_phys.py) by: the rest-state fixed point, exact
 * mass and momentum conservation in periodic boxes, the per-step momentum
 * identity  P(t+1) = P(t) + g N_fluid - drag(t), the shear-wave decay rate
 * nu = (tau - 1/2)/3, the exact body-force Poiseuille profile at
 * tau = 1/2 + sqrt(3/16), the Zou-He imposed velocity and the outlet's zero
 * gradient.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "phys_oracle.h"

static const int PO_EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int PO_EY[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
static const int PO_INV[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const double PO_W[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                               1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

double po_weight(int i) { return PO_W[i]; }

static int po_wrap(int v, int n)
{
    if (v < 0) return v + n;
    if (v >= n) return v - n;
    return v;
}

/* Post-streaming, post-boundary-condition populations of the fluid cell
 * (x, y) gathered from the post-collision state `post`. */
void po_gather(const po_params* p, const uint8_t* solid, const double* post, int x, int y,
               double f[9])
{
    const int nx = p->nx, ny = p->ny;
    const size_t n = (size_t)nx * (size_t)ny;
    int cx = x;
    if (p->x_edge == 1 && x == nx - 1 && !solid[(size_t)(nx - 2) + (size_t)y * nx]) cx = nx - 2;
    const size_t cc = (size_t)cx + (size_t)y * (size_t)nx;
    for (int i = 0; i < 9; ++i) {
        int sx = cx - PO_EX[i];
        const int sy = po_wrap(y - PO_EY[i], ny);
        int blocked = 0;
        if (p->x_edge == 0) sx = po_wrap(sx, nx);
        else if (sx < 0 || sx >= nx) blocked = 1;
        if (!blocked && solid[(size_t)sx + (size_t)sy * nx]) blocked = 1;
        if (blocked) f[i] = post[cc + (size_t)PO_INV[i] * n];
        else f[i] = post[(size_t)sx + (size_t)sy * (size_t)nx + (size_t)i * n];
    }
    if (p->x_edge == 1 && x == 0) {
        const double u0 = p->inlet_ux;
        const double rho = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - u0);
        f[1] = f[3] + (2.0 / 3.0) * rho * u0;
        f[5] = f[7] + 0.5 * (f[4] - f[2]) + (1.0 / 6.0) * rho * u0;
        f[8] = f[6] - 0.5 * (f[4] - f[2]) + (1.0 / 6.0) * rho * u0;
    }
}

/* rho and the Guo-corrected velocity of gathered populations. */
static void po_moments(const po_params* p, const double f[9], double* rho, double* ux, double* uy)
{
    double r = 0.0, jx = 0.0, jy = 0.0;
    for (int i = 0; i < 9; ++i) {
        r += f[i];
        jx += PO_EX[i] * f[i];
        jy += PO_EY[i] * f[i];
    }
    *rho = r;
    *ux = (jx + 0.5 * p->gx) / r;
    *uy = (jy + 0.5 * p->gy) / r;
}

void po_collide(const po_params* p, const double f[9], double out[9])
{
    const double omega = 1.0 / p->tau;
    double rho, ux, uy;
    po_moments(p, f, &rho, &ux, &uy);
    const double usq = ux * ux + uy * uy;
    for (int i = 0; i < 9; ++i) {
        const double eu = PO_EX[i] * ux + PO_EY[i] * uy;
        const double feq = PO_W[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
        const double fx = 3.0 * (PO_EX[i] - ux) + 9.0 * eu * PO_EX[i];
        const double fy = 3.0 * (PO_EY[i] - uy) + 9.0 * eu * PO_EY[i];
        const double Fi = (1.0 - 0.5 * omega) * PO_W[i] * (fx * p->gx + fy * p->gy);
        out[i] = f[i] - omega * (f[i] - feq) + Fi;
    }
}

/* One step: post_in (time t) -> post_out (time t+1), fluid cells only;
 * solid cells of post_out are set to 0. */
void po_step(const po_params* p, const uint8_t* solid, const double* post_in, double* post_out)
{
    const int nx = p->nx, ny = p->ny;
    const size_t n = (size_t)nx * (size_t)ny;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t c = (size_t)x + (size_t)y * (size_t)nx;
            if (solid[c]) {
                for (int i = 0; i < 9; ++i) post_out[c + (size_t)i * n] = 0.0;
                continue;
            }
            double f[9], o[9];
            po_gather(p, solid, post_in, x, y, f);
            po_collide(p, f, o);
            for (int i = 0; i < 9; ++i) post_out[c + (size_t)i * n] = o[i];
        }
}

void po_run(const po_params* p, const uint8_t* solid, double* a, double* b, int nsteps)
{
    const size_t n9 = (size_t)9 * (size_t)p->nx * (size_t)p->ny;
    for (int s = 0; s < nsteps; ++s) {
        po_step(p, solid, a, b);
        memcpy(a, b, n9 * sizeof(double));
    }
}

/* Moments of the state the next step gathers; 0 on solid cells. */
void po_macro(const po_params* p, const uint8_t* solid, const double* post, double* rho,
              double* ux, double* uy)
{
    const int nx = p->nx, ny = p->ny;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t c = (size_t)x + (size_t)y * (size_t)nx;
            if (solid[c]) {
                rho[c] = ux[c] = uy[c] = 0.0;
                continue;
            }
            double f[9];
            po_gather(p, solid, post, x, y, f);
            po_moments(p, f, &rho[c], &ux[c], &uy[c]);
        }
}

/* out[0] mass, out[1..2] momentum sum e_i f_i of the gathered populations,
 * out[3..4] momentum-exchange force on the solid cells from the last
 * collision (the post-collision values held in `post`). */
void po_reduce(const po_params* p, const uint8_t* solid, const double* post, double out[5])
{
    const int nx = p->nx, ny = p->ny;
    const size_t n = (size_t)nx * (size_t)ny;
    double m = 0.0, px = 0.0, py = 0.0, dx = 0.0, dy = 0.0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t c = (size_t)x + (size_t)y * (size_t)nx;
            if (solid[c]) continue;
            double f[9];
            po_gather(p, solid, post, x, y, f);
            for (int i = 0; i < 9; ++i) {
                m += f[i];
                px += PO_EX[i] * f[i];
                py += PO_EY[i] * f[i];
            }
            for (int i = 1; i < 9; ++i) {
                int tx = x + PO_EX[i];
                const int ty = po_wrap(y + PO_EY[i], ny);
                if (p->x_edge == 0) tx = po_wrap(tx, nx);
                else if (tx < 0 || tx >= nx) continue;
                if (!solid[(size_t)tx + (size_t)ty * nx]) continue;
                dx += 2.0 * PO_EX[i] * post[c + (size_t)i * n];
                dy += 2.0 * PO_EY[i] * post[c + (size_t)i * n];
            }
        }
    out[0] = m;
    out[1] = px;
    out[2] = py;
    out[3] = dx;
    out[4] = dy;
}

/* 4x4 block average (coarse cell (X, Y) = mean of fine cells 4X..4X+3,
 * 4Y..4Y+3) of a field; nx and ny are the fine sizes, multiples of 4. */
void po_restrict4(int nx, int ny, const double* fine, double* coarse)
{
    const int cx = nx / 4, cy = ny / 4;
    for (int Y = 0; Y < cy; ++Y)
        for (int X = 0; X < cx; ++X) {
            double s = 0.0;
            for (int j = 0; j < 4; ++j)
                for (int i = 0; i < 4; ++i) s += fine[(size_t)(4 * X + i) + (size_t)(4 * Y + j) * nx];
            coarse[(size_t)X + (size_t)Y * cx] = s / 16.0;
        }
}
