/* lbm3d.h -- C ABI of the gfx950 D3Q19 lattice-Boltzmann solver.
 *
 * No reference counterpart to replace: simuverse declares a 3-D fluid
 * simulator type (simuverse/src/lib.rs:114, SimuType::D3Fluid) but
 * implements none.  The calls mirror the 2-D handle of lbm.h, which mirrors
 * D2Q9Node (simuverse/src/fluid/d2q9_node.rs): create ⟷ D2Q9Node::new,
 * write_materials ⟷ the lattice-info upload (d2q9_node.rs:250-254), run ⟷
 * compute_by_pass (d2q9_node.rs:302-312) per step, macro ⟷ the macro pass.
 *
 * Physics: D3Q19 BGK, periodic in x, y and z, half-way bounce-back on
 * solid cells, Guo body force; the same conventions as LBM_MODE_PHYSICS
 * (pull gather fused with the collision, bounce-back by a push into solid
 * neighbours, populations stored as deviations f_i - w_i).
 *
 * Direction order: 0 rest; 1-6 faces (+x, -x, +y, -y, +z, -z); 7-18 edges
 * (+x+y, -x-y, +x-y, -x+y, +x+z, -x-z, +x-z, -x+z, +y+z, -y-z, +y-z, -y+z);
 * opposite directions are adjacent.  Host arrays are planar:
 * f[i*N + (z*ny + y)*nx + x], N = nx*ny*nz; material and macro arrays are
 * (z*ny + y)*nx + x.
 *
 * Status codes, lbm_last_error() and the threading rules are those of
 * lbm.h (one HIP stream per handle; calls on a handle are stream-ordered).
 */
#ifndef SIMUVERSE_LBM3D_H
#define SIMUVERSE_LBM3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm3d_handle lbm3d_handle;

typedef struct lbm3d_desc {
    int32_t nx, ny, nz;   /* each >= 3                                        */
    float tau;            /* > 0.5                                            */
    int32_t storage;      /* LBM_STORAGE_F32 or LBM_STORAGE_F16_DEV (lbm.h)   */
    float force_x;        /* Guo body force                                   */
    float force_y;
    float force_z;
    int32_t reserved[8];  /* must be zero                                     */
} lbm3d_desc;

/* Allocates two population buffers (row-interleaved: the 19 direction rows
 * of one (y, z) row adjacent, row pitch nx rounded up to 16), the material
 * bytes and a packed solid mask on the current HIP device (a gfx950), and
 * sets every cell to fluid at rest (f = w). */
int lbm3d_create(const lbm3d_desc* desc, lbm3d_handle** out);
int lbm3d_destroy(lbm3d_handle* h);
/* Whole-lattice material map (N bytes, lbm.h's LBM_MAT_* numbering: 2 and 4
 * are solid, everything else fluid). */
int lbm3d_write_materials(lbm3d_handle* h, const uint8_t* material);
/* Every cell back to rest (f = w). */
int lbm3d_reset(lbm3d_handle* h);
/* Post-collision populations of every cell (19*N floats, planar); the values
 * at solid cells are ignored. */
int lbm3d_write_populations(lbm3d_handle* h, const float* f);
/* The current buffer (19*N floats, planar); solid cells hold the values
 * their fluid neighbours pushed into them. */
int lbm3d_read_populations(lbm3d_handle* h, float* f);
/* nsteps steps (HIP-graph chunks of lbm3d_set_graph_chunk steps). */
int lbm3d_run(lbm3d_handle* h, int nsteps);
/* lbm3d_run bracketed by HIP events on the handle's stream. */
int lbm3d_run_timed(lbm3d_handle* h, int nsteps, double* ms);
int lbm3d_synchronize(lbm3d_handle* h);
/* rho and u (Guo half-force velocity) of the populations the next step
 * gathers; zero at solid cells.  Any pointer may be NULL. */
int lbm3d_macro(lbm3d_handle* h, float* rho, float* ux, float* uy, float* uz);
/* Binary64 sums over fluid cells with a fixed reduction order: out[0] mass,
 * out[1..3] momentum (sum of e_i f_i of the gathered populations),
 * out[4..6] the force on solid cells from the last step (momentum exchange
 * 2 e_i f*_i over every fluid-to-solid link). */
int lbm3d_reduce(lbm3d_handle* h, double out[7]);
/* Step-kernel shapes: every shape computes bit-identical populations. */
int lbm3d_variant_count(void);
/* block: threads per block; cells_per_thread: cells of one row per thread
 * (strided by the block); flags: bit 0 non-temporal population stores,
 * bit 2 the blocks one XCD receives take a contiguous range of rows. */
int lbm3d_variant_info(int variant, int* block, int* cells_per_thread, int* flags);
int lbm3d_set_variant(lbm3d_handle* h, int variant);
int lbm3d_get_variant(lbm3d_handle* h, int* variant);
/* Times every shape over `steps` steps and keeps the fastest (a shape later
 * in the table must be at least 1 % faster than the best so far); the state
 * is restored afterwards. */
int lbm3d_autotune(lbm3d_handle* h, int steps, int* best_variant, double* ms_per_step);
/* Steps per HIP-graph chunk of lbm3d_run (even, default 100; 0: direct
 * launches). */
int lbm3d_set_graph_chunk(lbm3d_handle* h, int chunk);

#ifdef __cplusplus
}
#endif

#endif /* SIMUVERSE_LBM3D_H */
