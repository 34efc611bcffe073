/*
 * learned_oracle.c -- binary64 CPU oracle of the learned collision step, of
 * the training targets the network is fitted to, and of the 4x restriction
 * that produces the coarse states those targets come from.
 * TEST INFRASTRUCTURE ONLY: tests call it as the checker; the product never
 * links it.
 *
 * No reference counterpart exists: jinleili/simuverse has no learned
 * collision.  The step modifies the BGK collision of the physics mode (the
 * reference's own BGK is assets/wgsl/lbm/collide_stream.wgsl:76-87), which
 * po_collide (phys_oracle.c) states.
 *
 * Moment rows (the orthogonal D2Q9 basis of Lallemand & Luo, Phys. Rev. E 61
 * (2000) 6546), orthogonal to each other and to 1, ex, ey:
 *     m_e  = (-4, -1, -1, -1, -1, 2, 2, 2, 2)    |m_e|^2  = 36
 *     m_xx = ex^2 - ey^2                         |m_xx|^2 = 4
 *     m_xy = ex ey                               |m_xy|^2 = 4
 *
 * lo_step -- one step.  Every fluid cell gathers f (po_gather: pull
 * streaming, half-way bounce-back, Zou-He inlet, zero-gradient outlet) and
 * forms rho = sum f, u = (sum e f + g/2) / rho and feq(rho, u).  A cell
 * whose 3x3 neighbourhood is fluid (y wrapped; x wrapped when periodic, else
 * columns 1 .. nx-2 only) is corrected:
 *     x_i     = (f_i - feq_i) / (w_i s_in)                network input
 *     y       = network(x)                                mlp_forward
 *     n_a     = sum_i m_a(i) (f_i - feq_i)                a = e, xx, xy
 *     alpha_a = clamp(s_out y_a, -b, b),  b = min((2 - omega)/2, omega/2)
 *     out_i   = BGK_i(f) + sum_a m_a(i) alpha_a n_a / |m_a|^2
 * and every other fluid cell takes BGK_i(f) (po_collide, Guo forcing).
 * The bound b keeps the relaxation factor 1 - omega + alpha of each moment
 * inside (-1, 1) for every omega in (0, 2), symmetric about 0.
 *
 * lo_transition -- training samples from two consecutive pre-collision
 * states s0, s1 (dense f[i*n + c]) of a lattice periodic in x and y.
 * Candidate k is cell c = k*stride + offset; a candidate whose periodic 3x3
 * neighbourhood holds a solid cell is skipped.  For the others, in candidate
 * order, with f = s0(c):
 *     X = x (as above)
 *     d_i = s1_i(c + e_i) - BGK_i(f)      the residual of one BGK step
 *     d_a = sum_i m_a(i) d_i
 *     eps2 = 1e-2 (n_xx^2 + n_xy^2 + n_e^2 / 9) + 1e-30
 *     T_a = clamp(d_a n_a / (n_a^2 + eps2), -b, b) / s_out   (a = 0, 1, 2)
 *     T_3..8 = 0
 * i.e. the modulation that reproduces the residual, regularised where n_a
 * is small against the cell's stress.
 *
 * lo_restrict4 -- 4x restriction of a state held as post-collision
 * populations (the state the next step gathers) to an (nx/4) x (ny/4)
 * coarse lattice.  A coarse cell is solid when fewer than 8 of its 16 fine
 * cells are fluid; otherwise, over its fluid fine cells (gathered with
 * po_gather, fine force in u):
 *     rho_c = mean rho,  u_c = sum(rho u) / sum(rho)
 *     q_i   = (4 tau_c / tau_f) mean(f_i - feq_i(rho, u))
 *     q_i  -= w_i (sum q + 3 e_i . sum e q)      no mass, no momentum
 *     fc_i  = feq_i(rho_c, u_c) + q_i - (3/2) w_i e_i . g_c
 * (acoustic rescaling of the non-equilibrium part for a 4x coarser lattice,
 * Dupuis & Chopard, Phys. Rev. E 67 (2003) 066707; the last term makes the
 * coarse Guo velocity (sum e fc + g_c/2) / rho_c equal u_c).
 *
 * Pinned (tests/test_oracle_learned.py) by: a state whose non-equilibrium
 * lies in one moment a comes out with (1 - omega + alpha_a) n_a in that
 * moment and the plain BGK value in every other; the correction's input
 * normalisation through a network that passes one input linearly to one
 * output; the clamp at both bounds over a sweep of tau; the mask (cells next
 * to a solid or an open edge equal plain BGK); targets of a transition made
 * by a known constant modulation equal alpha_a n_a^2 / (n_a^2 + eps2); the
 * restriction of a uniform shear state scaling its non-equilibrium part by
 * exactly 4 tau_c / tau_f; block mass and mass-weighted velocity of a
 * random state.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "phys_oracle.h"

static const int LO_EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int LO_EY[9] = {0, 0, -1, 0, 1, -1, -1, 1, 1};
static const double LO_W[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                               1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
static const double LO_ME[9] = {-4.0, -1.0, -1.0, -1.0, -1.0, 2.0, 2.0, 2.0, 2.0};

static double lo_mxx(int i) { return (double)(LO_EX[i] * LO_EX[i] - LO_EY[i] * LO_EY[i]); }
static double lo_mxy(int i) { return (double)(LO_EX[i] * LO_EY[i]); }

/* Bound b of the modulation for a relaxation rate omega. */
double lo_modulation_bound(double omega)
{
    const double a = 0.5 * (2.0 - omega), c = 0.5 * omega;
    return a < c ? a : c;
}

static double lo_clamp(double v, double b) { return v < -b ? -b : (v > b ? b : v); }

static int lo_wrap(int v, int n)
{
    if (v < 0) return v + n;
    if (v >= n) return v - n;
    return v;
}

/* rho, Guo velocity and f - feq of gathered populations (force gx, gy). */
static void lo_neq(const double f[9], double gx, double gy, double* rho, double* ux, double* uy, double neq[9])
{
    double r = 0.0, jx = 0.0, jy = 0.0;
    for (int i = 0; i < 9; ++i) {
        r += f[i];
        jx += LO_EX[i] * f[i];
        jy += LO_EY[i] * f[i];
    }
    const double u = (jx + 0.5 * gx) / r, v = (jy + 0.5 * gy) / r;
    const double usq = u * u + v * v;
    for (int i = 0; i < 9; ++i) {
        const double eu = LO_EX[i] * u + LO_EY[i] * v;
        neq[i] = f[i] - LO_W[i] * r * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
    }
    *rho = r;
    *ux = u;
    *uy = v;
}

/* The three non-equilibrium moments n_e, n_xx, n_xy. */
static void lo_moments(const double q[9], double n[3])
{
    n[0] = n[1] = n[2] = 0.0;
    for (int i = 0; i < 9; ++i) {
        n[0] += LO_ME[i] * q[i];
        n[1] += lo_mxx(i) * q[i];
        n[2] += lo_mxy(i) * q[i];
    }
}

/* Whether the 3x3 neighbourhood of (x, y) is fluid (y wrapped; x wrapped
 * when wrap_x, else cells beyond the x range count as not fluid). */
static int lo_nbhd_fluid(const uint8_t* solid, int nx, int ny, int x, int y, int wrap_x)
{
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int xx = x + dx;
            const int yy = lo_wrap(y + dy, ny);
            if (wrap_x) xx = lo_wrap(xx, nx);
            else if (xx < 0 || xx >= nx) return 0;
            if (solid[(size_t)xx + (size_t)yy * nx]) return 0;
        }
    return 1;
}

/* Correction of one learned cell from its non-equilibrium part q, its
 * network output y and the bound b. */
static void lo_correction(const double q[9], const double y[9], double s_out, double b, double corr[9])
{
    double n[3];
    lo_moments(q, n);
    const double ce = lo_clamp(s_out * y[0], b) * n[0] / 36.0;
    const double cxx = lo_clamp(s_out * y[1], b) * n[1] / 4.0;
    const double cxy = lo_clamp(s_out * y[2], b) * n[2] / 4.0;
    for (int i = 0; i < 9; ++i) corr[i] = LO_ME[i] * ce + lo_mxx(i) * cxx + lo_mxy(i) * cxy;
}

/* One learned step: post_in (time t) -> post_out (time t+1); solid cells
 * of post_out are 0.  learned (optional, n bytes) marks the corrected
 * cells.  Returns 0, or -1 when out of memory. */
int lo_step(const po_params* p, const uint8_t* solid, const double* post_in, double* post_out,
            const double* net, double s_in, double s_out, uint8_t* learned)
{
    const int nx = p->nx, ny = p->ny;
    const size_t n = (size_t)nx * (size_t)ny;
    const double omega = 1.0 / p->tau, b = lo_modulation_bound(omega);
    const int wrap_x = p->x_edge == 0;
    /* pass 1: gathered populations, BGK, and the inputs of learned cells */
    double* fg = (double*)malloc(sizeof(double) * 9 * n);
    double* xin = (double*)malloc(sizeof(double) * 9 * n);
    double* yout = (double*)malloc(sizeof(double) * 9 * n);
    int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * n);
    if (!fg || !xin || !yout || !slot) {
        free(fg);
        free(xin);
        free(yout);
        free(slot);
        return -1;
    }
    int nl = 0;
    for (size_t c = 0; c < n; ++c) {
        const int x = (int)(c % (size_t)nx), y = (int)(c / (size_t)nx);
        slot[c] = -1;
        if (solid[c]) continue;
        const int interior = wrap_x || (x > 0 && x < nx - 1);
        if (interior && lo_nbhd_fluid(solid, nx, ny, x, y, wrap_x)) slot[c] = nl++;
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)n; ++c) {
        if (solid[c]) continue;
        const int x = (int)(c % nx), y = (int)(c / nx);
        double* f = fg + 9 * (size_t)c;
        po_gather(p, solid, post_in, x, y, f);
        if (slot[c] >= 0) {
            double rho, ux, uy, q[9];
            lo_neq(f, p->gx, p->gy, &rho, &ux, &uy, q);
            for (int i = 0; i < 9; ++i) xin[9 * (size_t)slot[c] + i] = q[i] / (LO_W[i] * s_in);
        }
    }
    if (nl > 0) mlp_forward(net, nl, xin, yout);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)n; ++c) {
        if (learned) learned[c] = slot[c] >= 0;
        if (solid[c]) {
            for (int i = 0; i < 9; ++i) post_out[(size_t)c + (size_t)i * n] = 0.0;
            continue;
        }
        const double* f = fg + 9 * (size_t)c;
        double o[9];
        po_collide(p, f, o);
        if (slot[c] >= 0) {
            double rho, ux, uy, q[9], corr[9];
            lo_neq(f, p->gx, p->gy, &rho, &ux, &uy, q);
            lo_correction(q, yout + 9 * (size_t)slot[c], s_out, b, corr);
            for (int i = 0; i < 9; ++i) o[i] += corr[i];
        }
        for (int i = 0; i < 9; ++i) post_out[(size_t)c + (size_t)i * n] = o[i];
    }
    free(fg);
    free(xin);
    free(yout);
    free(slot);
    return 0;
}

/* nsteps learned steps in place (a: state, b: scratch of the same size). */
int lo_run(const po_params* p, const uint8_t* solid, double* a, double* b, const double* net, double s_in,
           double s_out, int nsteps)
{
    const size_t n9 = (size_t)9 * (size_t)p->nx * (size_t)p->ny;
    for (int s = 0; s < nsteps; ++s) {
        if (lo_step(p, solid, a, b, net, s_in, s_out, NULL)) return -1;
        memcpy(a, b, n9 * sizeof(double));
    }
    return 0;
}

/* Training samples of one transition (see the header); X and T hold room
 * for max_samples rows of 9.  Returns the number of samples written. */
int lo_transition(const po_params* p, const uint8_t* solid, const double* s0, const double* s1, double s_in,
                  double s_out, int stride, int offset, int max_samples, double* X, double* T)
{
    const int nx = p->nx, ny = p->ny;
    const size_t n = (size_t)nx * (size_t)ny;
    const double b = lo_modulation_bound(1.0 / p->tau);
    int count = 0;
    for (size_t k = 0;; ++k) {
        const size_t c = k * (size_t)stride + (size_t)offset;
        if (c >= n) break;
        const int x = (int)(c % (size_t)nx), y = (int)(c / (size_t)nx);
        if (!lo_nbhd_fluid(solid, nx, ny, x, y, 1)) continue;
        if (count >= max_samples) break;
        double f[9], q[9], bgk[9], d[9], rho, ux, uy;
        for (int i = 0; i < 9; ++i) f[i] = s0[c + (size_t)i * n];
        lo_neq(f, p->gx, p->gy, &rho, &ux, &uy, q);
        po_collide(p, f, bgk);
        for (int i = 0; i < 9; ++i) {
            const int tx = lo_wrap(x + LO_EX[i], nx), ty = lo_wrap(y + LO_EY[i], ny);
            d[i] = s1[(size_t)tx + (size_t)ty * nx + (size_t)i * n] - bgk[i];
        }
        double nm[3], dm[3];
        lo_moments(q, nm);
        lo_moments(d, dm);
        const double eps2 = 1e-2 * (nm[1] * nm[1] + nm[2] * nm[2] + nm[0] * nm[0] / 9.0) + 1e-30;
        double* xr = X + 9 * (size_t)count;
        double* tr = T + 9 * (size_t)count;
        for (int i = 0; i < 9; ++i) {
            xr[i] = q[i] / (LO_W[i] * s_in);
            tr[i] = 0.0;
        }
        for (int a = 0; a < 3; ++a) tr[a] = lo_clamp(dm[a] * nm[a] / (nm[a] * nm[a] + eps2), b) / s_out;
        ++count;
    }
    return count;
}

/* 4x restriction (see the header): fc dense (9 * nc), solid_c (nc bytes),
 * uxc, uyc (nc). */
void lo_restrict4(const po_params* p, const uint8_t* solid, const double* post, double tau_c, double gcx,
                  double gcy, double* fc, uint8_t* solid_c, double* uxc, double* uyc)
{
    const int nx = p->nx, cx = p->nx / 4, cy = p->ny / 4;
    const size_t nc = (size_t)cx * (size_t)cy;
    const double nscale = 4.0 * tau_c / p->tau;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)nc; ++k) {
        const int X = (int)(k % cx), Y = (int)(k / cx);
        double sr = 0.0, sjx = 0.0, sjy = 0.0, sq[9] = {0};
        int nfl = 0;
        for (int j = 0; j < 4; ++j)
            for (int i4 = 0; i4 < 4; ++i4) {
                const int x = 4 * X + i4, y = 4 * Y + j;
                if (solid[(size_t)x + (size_t)y * nx]) continue;
                ++nfl;
                double f[9], q[9], rho, ux, uy;
                po_gather(p, solid, post, x, y, f);
                lo_neq(f, p->gx, p->gy, &rho, &ux, &uy, q);
                sr += rho;
                sjx += rho * ux;
                sjy += rho * uy;
                for (int i = 0; i < 9; ++i) sq[i] += q[i];
            }
        if (nfl < 8) {
            solid_c[k] = 1;
            for (int i = 0; i < 9; ++i) fc[(size_t)k + (size_t)i * nc] = 0.0;
            uxc[k] = uyc[k] = 0.0;
            continue;
        }
        solid_c[k] = 0;
        const double rho = sr / nfl, ux = sjx / sr, uy = sjy / sr;
        double q[9], a = 0.0, bx = 0.0, by = 0.0;
        for (int i = 0; i < 9; ++i) {
            q[i] = nscale * sq[i] / nfl;
            a += q[i];
            bx += LO_EX[i] * q[i];
            by += LO_EY[i] * q[i];
        }
        const double usq = ux * ux + uy * uy;
        for (int i = 0; i < 9; ++i) {
            q[i] -= LO_W[i] * (a + 3.0 * (LO_EX[i] * bx + LO_EY[i] * by));
            const double eu = LO_EX[i] * ux + LO_EY[i] * uy;
            const double feq = LO_W[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
            fc[(size_t)k + (size_t)i * nc] = feq + q[i] - 1.5 * LO_W[i] * (LO_EX[i] * gcx + LO_EY[i] * gcy);
        }
        uxc[k] = ux;
        uyc[k] = uy;
    }
}
