/*
 * mlp_oracle.c -- binary64 CPU oracle of the learned collision correction's
 * network.  TEST INFRASTRUCTURE ONLY: tests and the smoke check call it as
 * the checker; the product never links it.
 *
 * No reference counterpart exists: jinleili/simuverse has no learned
 * collision.  The network is the one the learned-collision kernels
 * implement: 9 -> 64 -> 64 -> 9, GELU (erf form, 0.5 x (1 + erf(x/sqrt 2)))
 * after the two hidden layers, linear output.  Parameters are one flat
 * vector of 5385 values:
 *     W1[64][9]  b1[64]  W2[64][64]  b2[64]  W3[9][64]  b3[9]
 * (row = output unit).  The loss of a batch of B cells is
 *     L = 1/(9B) sum_b sum_k (y_bk - t_bk)^2 .
 * Adam follows Kingma & Ba (2015), Algorithm 1, with bias correction.
 *
 * Pinned (tests/test_oracle_mlp.py) by central finite differences of the
 * loss, the closed-form gradient of a network whose hidden layers are
 * switched off, the GELU values at 0 and at large |x| and a one-step Adam
 * update of known sign and size.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#define MLP_IN 9
#define MLP_H 64
#define MLP_OUT 9
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + MLP_H * MLP_IN)
#define OFF_W2 (OFF_B1 + MLP_H)
#define OFF_B2 (OFF_W2 + MLP_H * MLP_H)
#define OFF_W3 (OFF_B2 + MLP_H)
#define OFF_B3 (OFF_W3 + MLP_OUT * MLP_H)
#define MLP_NPARAM (OFF_B3 + MLP_OUT)

int mlp_param_count(void) { return MLP_NPARAM; }

double mlp_gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }

double mlp_gelu_grad(double x)
{
    return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / sqrt(2.0 * M_PI);
}

/* Forward pass of one cell; keeps the pre-activations for the backward. */
static void mlp_cell_forward(const double* p, const double* x, double* a1, double* h1, double* a2, double* h2,
                             double* y)
{
    for (int j = 0; j < MLP_H; ++j) {
        double s = p[OFF_B1 + j];
        for (int i = 0; i < MLP_IN; ++i) s += p[OFF_W1 + j * MLP_IN + i] * x[i];
        a1[j] = s;
        h1[j] = mlp_gelu(s);
    }
    for (int j = 0; j < MLP_H; ++j) {
        double s = p[OFF_B2 + j];
        for (int i = 0; i < MLP_H; ++i) s += p[OFF_W2 + j * MLP_H + i] * h1[i];
        a2[j] = s;
        h2[j] = mlp_gelu(s);
    }
    for (int k = 0; k < MLP_OUT; ++k) {
        double s = p[OFF_B3 + k];
        for (int i = 0; i < MLP_H; ++i) s += p[OFF_W3 + k * MLP_H + i] * h2[i];
        y[k] = s;
    }
}

void mlp_forward(const double* p, int batch, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int b = 0; b < batch; ++b) {
        double a1[MLP_H], h1[MLP_H], a2[MLP_H], h2[MLP_H];
        mlp_cell_forward(p, x + (size_t)b * MLP_IN, a1, h1, a2, h2, y + (size_t)b * MLP_OUT);
    }
}

/* Loss of the batch and its gradient with respect to every parameter. */
double mlp_loss_grad(const double* p, int batch, const double* x, const double* t, double* grad)
{
    memset(grad, 0, sizeof(double) * MLP_NPARAM);
    double loss = 0.0;
    const double scale = 1.0 / (9.0 * (double)batch);
    for (int b = 0; b < batch; ++b) {
        double a1[MLP_H], h1[MLP_H], a2[MLP_H], h2[MLP_H], y[MLP_OUT];
        const double* xb = x + (size_t)b * MLP_IN;
        mlp_cell_forward(p, xb, a1, h1, a2, h2, y);
        double dy[MLP_OUT];
        for (int k = 0; k < MLP_OUT; ++k) {
            const double r = y[k] - t[(size_t)b * MLP_OUT + k];
            loss += r * r;
            dy[k] = 2.0 * r * scale;
        }
        double dh2[MLP_H];
        for (int i = 0; i < MLP_H; ++i) dh2[i] = 0.0;
        for (int k = 0; k < MLP_OUT; ++k) {
            grad[OFF_B3 + k] += dy[k];
            for (int i = 0; i < MLP_H; ++i) {
                grad[OFF_W3 + k * MLP_H + i] += dy[k] * h2[i];
                dh2[i] += p[OFF_W3 + k * MLP_H + i] * dy[k];
            }
        }
        double da2[MLP_H], dh1[MLP_H];
        for (int j = 0; j < MLP_H; ++j) da2[j] = dh2[j] * mlp_gelu_grad(a2[j]);
        for (int i = 0; i < MLP_H; ++i) dh1[i] = 0.0;
        for (int j = 0; j < MLP_H; ++j) {
            grad[OFF_B2 + j] += da2[j];
            for (int i = 0; i < MLP_H; ++i) {
                grad[OFF_W2 + j * MLP_H + i] += da2[j] * h1[i];
                dh1[i] += p[OFF_W2 + j * MLP_H + i] * da2[j];
            }
        }
        for (int j = 0; j < MLP_H; ++j) {
            const double da1 = dh1[j] * mlp_gelu_grad(a1[j]);
            grad[OFF_B1 + j] += da1;
            for (int i = 0; i < MLP_IN; ++i) grad[OFF_W1 + j * MLP_IN + i] += da1 * xb[i];
        }
    }
    return loss * scale;
}

/* One Adam update; step is the 1-based update count. */
void mlp_adam(double* p, const double* grad, double* m, double* v, int step, double lr, double beta1, double beta2,
              double eps)
{
    const double c1 = 1.0 - pow(beta1, step);
    const double c2 = 1.0 - pow(beta2, step);
    for (int k = 0; k < MLP_NPARAM; ++k) {
        m[k] = beta1 * m[k] + (1.0 - beta1) * grad[k];
        v[k] = beta2 * v[k] + (1.0 - beta2) * grad[k] * grad[k];
        const double mh = m[k] / c1;
        const double vh = v[k] / c2;
        p[k] -= lr * mh / (sqrt(vh) + eps);
    }
}
