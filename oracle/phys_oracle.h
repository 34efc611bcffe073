/* phys_oracle.h -- parameters of the binary64 physics oracle (phys_oracle.c)
 * and the functions learned_oracle.c builds on.  TEST INFRASTRUCTURE ONLY. */
#ifndef LBM_PHYS_ORACLE_H
#define LBM_PHYS_ORACLE_H

#include <stdint.h>

typedef struct {
    int32_t nx, ny;
    double tau;
    int32_t x_edge; /* 0 periodic, 1 Zou-He inlet at x=0 + zero-gradient outlet */
    double inlet_ux;
    double gx, gy;
} po_params;

void po_gather(const po_params* p, const uint8_t* solid, const double* post, int x, int y, double f[9]);
void po_collide(const po_params* p, const double f[9], double out[9]);
void mlp_forward(const double* params, int batch, const double* x, double* y);

#endif
