/* san_check.c -- runs every function of the CPU oracle once on small
 * lattices, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (Makefile target lib/san_check; tests/test_oracle_sanitized.py).  Cases:
 * tiny periodic lattices (3x3, the wrap onto itself), channels with solids
 * on the inlet and outlet columns, the learned step with a saturating
 * network, training samples, the restriction.  TEST INFRASTRUCTURE ONLY. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "phys_oracle.h"

typedef struct {
    int32_t material, block_iter;
    float vx, vy;
} ws_info;

void ws_run(int nx, int ny, float tau, ws_info* info, float* buf0, float* buf1, float* macro4, int nsteps,
            int start_swap);
void ws_init(int nx, int ny, int fluid_ty, ws_info* info, float* A, float* B, float* macro4);
void po_run(const po_params* p, const uint8_t* solid, double* a, double* b, int nsteps);
void po_macro(const po_params* p, const uint8_t* solid, const double* post, double* rho, double* ux, double* uy);
void po_reduce(const po_params* p, const uint8_t* solid, const double* post, double out[5]);
void po_restrict4(int nx, int ny, const double* fine, double* coarse);
int mlp_param_count(void);
double mlp_loss_grad(const double* p, int batch, const double* x, const double* t, double* grad);
void mlp_adam(double* p, const double* grad, double* m, double* v, int step, double lr, double beta1, double beta2,
              double eps);
int lo_run(const po_params* p, const uint8_t* solid, double* a, double* b, const double* net, double s_in,
           double s_out, int nsteps);
int lo_transition(const po_params* p, const uint8_t* solid, const double* s0, const double* s1, double s_in,
                  double s_out, int stride, int offset, int max_samples, double* X, double* T);
void lo_restrict4(const po_params* p, const uint8_t* solid, const double* post, double tau_c, double gcx,
                  double gcy, double* fc, uint8_t* solid_c, double* uxc, double* uyc);

static double frand(unsigned* s)
{
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0 - 0.5;
}

static int finite_all(const double* a, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(a[i])) return 0;
    return 1;
}

static int phys_case(int nx, int ny, int x_edge, unsigned seed)
{
    const size_t n = (size_t)nx * (size_t)ny;
    po_params p = {nx, ny, 0.6, x_edge, x_edge ? 0.03 : 0.0, 1e-5, 0.0};
    uint8_t* solid = calloc(n, 1);
    double* a = calloc(9 * n, sizeof(double));
    double* b = calloc(9 * n, sizeof(double));
    double *rho = calloc(n, sizeof(double)), *ux = calloc(n, sizeof(double)), *uy = calloc(n, sizeof(double));
    for (size_t c = 0; c < n; ++c) solid[c] = (c * 7u + seed) % 11u == 0u;
    if (x_edge && nx >= 3) solid[0] = solid[(size_t)nx - 1] = 1; /* solids on the inlet and outlet columns */
    for (size_t k = 0; k < 9 * n; ++k) a[k] = 1e-3 * frand(&seed);
    po_run(&p, solid, a, b, 7);
    po_macro(&p, solid, a, rho, ux, uy);
    double red[5];
    po_reduce(&p, solid, a, red);
    int ok = finite_all(a, 9 * n) && finite_all(red, 5);
    free(solid), free(a), free(b), free(rho), free(ux), free(uy);
    return ok;
}

int main(void)
{
    unsigned seed = 12345u;
    int ok = 1;

    /* binary32 WGSL restatement, both init kinds, tiny and odd sizes */
    for (int t = 0; t < 3; ++t) {
        const int nx = t == 0 ? 3 : (t == 1 ? 17 : 32), ny = t == 0 ? 3 : (t == 1 ? 9 : 24);
        const size_t n = (size_t)nx * (size_t)ny;
        ws_info* info = calloc(n, sizeof(ws_info));
        float *A = calloc(9 * n, sizeof(float)), *B = calloc(9 * n, sizeof(float)), *m4 = calloc(4 * n, sizeof(float));
        for (size_t c = 0; c < n; ++c) {
            info[c].material = (c % 13u == 0u) ? 4 : ((c % 17u == 1u) ? 3 : 1);
            info[c].block_iter = (int32_t)(c % 5u) - 1;
            info[c].vx = 0.01f;
        }
        ws_init(nx, ny, t & 1, info, A, B, m4);
        ws_run(nx, ny, 0.6f, info, A, B, m4, 5, 0);
        free(info), free(A), free(B), free(m4);
    }

    /* binary64 physics: periodic, tiny, channel with edge solids */
    ok &= phys_case(3, 3, 0, 1u);
    ok &= phys_case(5, 4, 0, 2u);
    ok &= phys_case(24, 16, 0, 3u);
    ok &= phys_case(24, 16, 1, 4u);
    ok &= phys_case(3, 5, 1, 5u);

    /* network: loss gradient and an Adam step */
    const int np = mlp_param_count(), batch = 13;
    double* net = calloc((size_t)np, sizeof(double));
    double *g = calloc((size_t)np, sizeof(double)), *m = calloc((size_t)np, sizeof(double)),
           *v = calloc((size_t)np, sizeof(double));
    double *x = calloc(9 * (size_t)batch, sizeof(double)), *t = calloc(9 * (size_t)batch, sizeof(double));
    for (int k = 0; k < np; ++k) net[k] = 0.3 * frand(&seed);
    for (int k = 0; k < 9 * batch; ++k) x[k] = frand(&seed), t[k] = frand(&seed);
    const double loss = mlp_loss_grad(net, batch, x, t, g);
    mlp_adam(net, g, m, v, 1, 1e-3, 0.9, 0.999, 1e-8);
    ok &= isfinite(loss) && finite_all(g, (size_t)np);

    /* learned step (saturating network), transitions and the restriction */
    {
        const int nx = 16, ny = 12;
        const size_t n = (size_t)nx * (size_t)ny;
        po_params p = {nx, ny, 0.52, 0, 0.0, 1e-6, 0.0};
        uint8_t* solid = calloc(n, 1);
        solid[5 * nx + 7] = 1;
        double *a = calloc(9 * n, sizeof(double)), *b = calloc(9 * n, sizeof(double)),
               *s1 = calloc(9 * n, sizeof(double));
        for (size_t k = 0; k < 9 * n; ++k) a[k] = 1e-3 * frand(&seed);
        memcpy(s1, a, 9 * n * sizeof(double));
        for (int k = 0; k < np; ++k) net[k] = 5.0 * frand(&seed);
        ok &= lo_run(&p, solid, a, b, net, 1e-3, 0.05, 3) == 0 && finite_all(a, 9 * n);
        double *X = calloc(9 * n, sizeof(double)), *T = calloc(9 * n, sizeof(double));
        const int count = lo_transition(&p, solid, s1, a, 1e-3, 0.05, 1, 0, (int)n, X, T);
        ok &= count > 0 && count <= (int)n && finite_all(X, 9 * (size_t)count);
        const size_t nc = (size_t)(nx / 4) * (size_t)(ny / 4);
        double *fc = calloc(9 * nc, sizeof(double)), *uxc = calloc(nc, sizeof(double)),
               *uyc = calloc(nc, sizeof(double)), *coarse = calloc(9 * nc, sizeof(double));
        uint8_t* solid_c = calloc(nc, 1);
        lo_restrict4(&p, solid, a, 0.505, 4e-6, 0.0, fc, solid_c, uxc, uyc);
        po_restrict4(nx, ny, a, coarse);
        ok &= finite_all(fc, 9 * nc) && finite_all(coarse, 9 * nc);
        free(solid), free(a), free(b), free(s1), free(X), free(T), free(fc), free(uxc), free(uyc), free(coarse);
        free(solid_c);
    }
    free(net), free(g), free(m), free(v), free(x), free(t);
    printf(ok ? "san_check ok\n" : "san_check FAILED\n");
    return ok ? 0 : 1;
}
