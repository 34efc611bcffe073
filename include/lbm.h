/*
 * lbm.h -- C ABI of the MI355X (gfx950) D2Q9 lattice-Boltzmann hot path.
 *
 * This library replaces the wgpu compute pipelines, bind groups and buffers
 * that simuverse's fluid node builds for its three LBM shaders.  Each entry
 * point names the reference interface it stands in for (paths are inside the
 * jinleili/simuverse tree):
 *
 *   D2Q9Node::new              simuverse/src/fluid/d2q9_node.rs:31-209
 *   D2Q9Node::compute_by_pass  simuverse/src/fluid/d2q9_node.rs:302-312
 *   D2Q9Node::reset / init     simuverse/src/fluid/d2q9_node.rs:211-213,
 *                              assets/wgsl/lbm/init.wgsl:19-63
 *   info_buf writes            simuverse/src/fluid/d2q9_node.rs:243-244,250-254,297-298
 *   LbmUniform rewrite         simuverse/src/fluid/fluid_simulator.rs:175-193
 *   macro_tex (rgba16float)    simuverse/src/fluid/d2q9_node.rs:91-104,
 *                              assets/wgsl/lbm/collide_stream.wgsl:36,74
 *
 * Conventions
 *   - Every function returns an int status: LBM_OK (0) or an LBM_ERR_* code.
 *     Nothing aborts or throws across the ABI; lbm_last_error() returns the
 *     message of the last failure on the calling thread.
 *   - The caller owns host memory, the library owns device memory.
 *   - Each handle owns one HIP stream; calls on one handle are stream
 *     ordered.  A handle is not thread-safe; separate handles may be driven
 *     from separate threads.
 *   - Host-side population arrays are float32, structure-of-arrays, dense:
 *     f[i * nx * ny + y * nx + x], i in 0..8 in the reference's direction
 *     numbering (simuverse/src/fluid/mod.rs:22-52):
 *         6 2 5        i : 0  1  2  3  4  5  6  7  8
 *         3 0 1        ex: 0  1  0 -1  0  1 -1 -1  1
 *         7 4 8        ey: 0  0 -1  0  1 -1 -1  1  1
 *     (y grows downwards, as in the reference's lattice.)
 */
#ifndef SIMUVERSE_LBM_H
#define SIMUVERSE_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 1

typedef struct lbm_handle lbm_handle;

/* ---- status codes ---------------------------------------------------- */
enum {
    LBM_OK = 0,
    LBM_ERR_INVALID = 1,     /* bad argument or handle state            */
    LBM_ERR_HIP = 2,         /* a HIP runtime call failed               */
    LBM_ERR_OOM = 3,         /* device allocation failed                */
    LBM_ERR_UNSUPPORTED = 4, /* valid request the library does not do  */
    LBM_ERR_NO_DEVICE = 5    /* no gfx950 device visible                */
};

/* ---- solver semantics ------------------------------------------------- */
enum {
    /* The reference shaders' arithmetic: truncated f32 weights
     * (0.444444 / 0.111111 / 0.0277777), rho clamp [0.8, 1.2], post-collision
     * clamp [0, max_i], accelerate cells (materials 3 and 6) with their
     * block_iter countdown, periodic gather wrap in x and y, and the boundary
     * pass's bounce-back with its outer-band skip
     * (assets/wgsl/lbm/collide_stream.wgsl:25-88, boundary.wgsl:3-35). */
    LBM_MODE_PARITY_WGSL = 0,
    /* Textbook D2Q9 BGK: exact weights 4/9, 1/9, 1/36, no clamps, half-way
     * bounce-back on solid cells, Guo body force, optional Zou-He velocity
     * inlet at x = 0 with a zero-gradient outlet at x = nx-1. */
    LBM_MODE_PHYSICS = 1
};

/* ---- population storage on the device --------------------------------- */
enum {
    LBM_STORAGE_F32 = 0,    /* binary32, 72 B per cell per step: f_i in
                               parity mode (the reference's buffers), the
                               deviation f_i - w_i in physics mode         */
    LBM_STORAGE_F16_DEV = 1 /* (f_i - w_i) * 2^dev_shift as binary16 with
                               binary32 arithmetic: 36 B per cell per step.
                               Physics mode only. */
};
/* lbm_write_populations / lbm_read_populations always exchange f_i itself
 * in binary32; the library converts to and from its storage. */

/* ---- x-edge treatment (physics mode) ----------------------------------- */
enum {
    LBM_X_PERIODIC = 0,
    /* Zou-He velocity inlet (ux = inlet_ux, uy = 0) on the fluid cells of
     * column x = 0; zero-gradient outlet on column x = nx-1 (its
     * post-streaming populations are those of column nx-2). */
    LBM_X_INLET_OUTLET = 1
};
/* y is periodic in both modes; channel walls are solid cells. */

/* Device layout of the physics-mode populations.  Either way the host
 * exchanges the reference's planar f[i*nx*ny + y*nx + x].
 *   ROWS    the nine direction rows of each y are adjacent (row pitch
 *           nx rounded up to 16 elements): element (y*9 + i)*pitch + x
 *   PLANES  one plane per direction: element i*plane + y*nx + x          */
enum { LBM_LAYOUT_ROWS = 0, LBM_LAYOUT_PLANES = 1 };

/* ---- materials: the reference's LatticeType (fluid/lattice.rs:15-24) --- */
enum {
    LBM_MAT_BULK = 1,
    LBM_MAT_BOUNDARY = 2,       /* solid                                    */
    LBM_MAT_INLET = 3,          /* accelerate cell in parity mode           */
    LBM_MAT_OBSTACLE = 4,       /* solid                                    */
    LBM_MAT_OUTLET = 5,
    LBM_MAT_EXTERNAL_FORCE = 6, /* accelerate cell in parity mode           */
    LBM_MAT_GHOST = 7
};

/* The reference's 16-byte LatticeInfo record (fluid/lattice.rs:5-11,
 * assets/wgsl/lbm/struct/lattice_info.wgsl:2-8). */
typedef struct lbm_lattice_info {
    int32_t material;
    int32_t block_iter;
    float vx;
    float vy;
} lbm_lattice_info;

/* Everything D2Q9Node::new derives from its settings
 * (d2q9_node.rs:38-59: lattice size, tau, fluid_ty; LbmUniform). */
typedef struct lbm_desc {
    int32_t nx, ny;     /* lattice size; 2 <= nx, ny; nx*ny <= 2^31 / 9   */
    float tau;          /* relaxation time; omega = 1/tau in binary32     */
    int32_t mode;       /* LBM_MODE_*                                     */
    int32_t storage;    /* LBM_STORAGE_*                                  */
    int32_t fluid_ty;   /* parity-mode reset: 0 Poiseuille-biased start,
                           1 plain start (LbmUniform.fluid_ty)            */
    int32_t x_edge;     /* physics mode: LBM_X_*                          */
    float inlet_ux;     /* physics mode, LBM_X_INLET_OUTLET: inlet ux     */
    float force_x;      /* physics mode: uniform body force (Guo)         */
    float force_y;
    int32_t dev_shift;  /* LBM_STORAGE_F16_DEV: power-of-two scale of the
                           stored deviations (0..14; 0 means the default 10) */
    int32_t layout;     /* physics mode: LBM_LAYOUT_* of the device
                           populations (parity mode is always planar)     */
    int32_t reserved[4];/* must be zero                                   */
} lbm_desc;

/* ---- library ----------------------------------------------------------- */
int lbm_abi_version(void);
const char* lbm_last_error(void);
/* Number of visible HIP devices of architecture gfx950. */
int lbm_device_count(int* count);

/* ---- lifetime (D2Q9Node::new / drop) ----------------------------------- */
/* Allocates both population buffers (ping-pong, like the two lattice_buf
 * storage buffers at d2q9_node.rs:110-118), the material map (all
 * LBM_MAT_BULK) and the macro field.  The populations are reset as by
 * lbm_reset(). */
int lbm_create(const lbm_desc* desc, lbm_handle** out);
int lbm_destroy(lbm_handle* h);

/* ---- inputs ------------------------------------------------------------ */
/* queue.write_buffer(info_buf, 16 * offset, records) -- d2q9_node.rs:244,
 * 250-254, 298.  Cells offset .. offset+count-1 in row-major order. */
int lbm_write_lattice_info(lbm_handle* h, const lbm_lattice_info* info,
                           size_t offset, size_t count);
int lbm_read_lattice_info(lbm_handle* h, lbm_lattice_info* info,
                          size_t offset, size_t count);
/* Compact form: only the material of each cell (block_iter = -1, v = 0). */
int lbm_write_materials(lbm_handle* h, const uint8_t* material, size_t offset,
                        size_t count);

/* init.wgsl:19-63 on both buffers: parity mode -- solids zeroed, f = w (or
 * the Poiseuille-biased start when fluid_ty == 0) in buffer 0, 0 in buffer 1
 * (the biased f1, f3 in both), expiring accelerate cells; physics mode --
 * f = w_i (rest state at rho = 1) in buffer 0 and 1, solids zeroed.  The
 * current buffer becomes 0. */
int lbm_reset(lbm_handle* h);

/* Seeded initial states, readback and checkpoints.  buf is 0 or 1; the
 * array holds 9*nx*ny floats in the layout described at the top. */
int lbm_write_populations(lbm_handle* h, int buf, const float* f);
int lbm_read_populations(lbm_handle* h, int buf, float* f);

/* Checkpoint / resume (no reference counterpart: the reference only
 * re-runs init, fluid_simulator.rs:210-215).  The state is a 64-byte
 * header (ABI version, lattice size, mode, storage, layout, binary16 scale,
 * strides) followed by the device image of the current population buffer,
 * lbm_state_size() bytes in all; loading it into a handle with the same
 * descriptor and lattice info and stepping reproduces an uninterrupted run
 * bit for bit.  A state whose header does not match the handle is rejected
 * with LBM_ERR_INVALID. */
int lbm_state_size(lbm_handle* h, size_t* bytes);
int lbm_save_state(lbm_handle* h, void* out, size_t bytes);
int lbm_load_state(lbm_handle* h, const void* in, size_t bytes);

/* LbmUniform rewrite with a new tau (fluid_simulator.rs:175-193). */
int lbm_set_tau(lbm_handle* h, float tau);

/* ---- stepping ---------------------------------------------------------- */
/* One LBM step reading buffer `swap` and writing buffer 1-swap: the
 * reference's collide_stream + boundary dispatch pair
 * (d2q9_node.rs:302-312), fused into one kernel.  The current buffer
 * becomes 1-swap.  When export_macro is non-zero the step also writes the
 * macro field exactly as collide_stream.wgsl writes macro_tex: (ux, uy, rho)
 * of the gathered pre-collision populations, rho after the clamp in parity
 * mode, zero on solid cells. */
int lbm_step(lbm_handle* h, int swap, int export_macro);
/* nsteps steps starting from the current buffer, as 100-step HIP-graph
 * chunks plus a remainder.  FluidSimulator::compute (fluid_simulator.rs:
 * 217-232) is lbm_run(h, 2, 1).  export_macro applies to the last step. */
int lbm_run(lbm_handle* h, int nsteps, int export_macro);
int lbm_current_buffer(lbm_handle* h, int* buf);
int lbm_synchronize(lbm_handle* h);
/* lbm_run timed with HIP events recorded on the handle's stream; *ms is the
 * elapsed device time of the nsteps steps. */
int lbm_run_timed(lbm_handle* h, int nsteps, double* ms);

/* ---- outputs ----------------------------------------------------------- */
/* The macro field written by the last step that exported it (float32
 * planes, nx*ny each; any pointer may be NULL). */
int lbm_read_macro(lbm_handle* h, float* rho, float* ux, float* uy);
/* The same field packed like macro_tex: rgba16float (ux, uy, rho, w) per
 * texel, w = 1 on fluid and 0 on solid cells, round-to-nearest-even. */
int lbm_read_macro_rgba16f(lbm_handle* h, uint16_t* rgba);
/* The reference's curl pass (assets/wgsl/lbm/curl_update.wgsl:12-34, whose
 * dispatch fluid_simulator.rs:226,230 leaves commented out) on that macro
 * field: rgba16float texels (curl * 3.5 + 0.5, 0, 0, 0) with
 * curl = uy(x+1) - uy(x-1) + ux(y-1) - ux(y+1) of the binary16 texels,
 * neighbours clamped at 0; past the last column / row the texel is zero. */
int lbm_read_curl_rgba16f(lbm_handle* h, uint16_t* rgba);
/* Moments of the current state (the populations the next step gathers):
 * the values the next step would export. */
int lbm_macro(lbm_handle* h, float* rho, float* ux, float* uy);
/* Domain-wide sums over fluid cells, accumulated in binary64 with a fixed
 * reduction order: out[0] mass, out[1..2] momentum (sum of e_i f_i of the
 * gathered populations), out[3..4] the force exerted on solid cells by the
 * last step (momentum exchange over fluid-solid links). */
int lbm_reduce(lbm_handle* h, double out[5]);

/* ---- tuning and diagnostics (no reference counterpart: the reference
 * dispatches a fixed 64x4 workgroup, d2q9_node.rs:45) --------------------- */
/* Physics-mode step kernels are built for several tile shapes; every shape
 * computes bit-identical populations at every slot a fluid cell can read
 * (a two-step launch leaves older values in the solid-cell slots no fluid
 * cell pulls from). */
int lbm_variant_count(void);
/* block: threads per block; cells_per_thread: cells one thread updates;
 * vec: consecutive cells per lane (1, or 4 with nx % 4 == 0); flags: bit 0
 * non-temporal population stores, bit 1 non-temporal loads, bit 2 rows
 * grouped per XCD, bit 3 aligned vector loads with x-neighbours from the
 * adjacent lane (periodic x, nx a multiple of block*vec), bit 4 two steps
 * per launch through an LDS tile of cells_per_thread rows x vec cells
 * (nx, ny >= 3; with LBM_X_INLET_OUTLET, nx % vec != 1; odd step counts
 * end with a one-step launch), bit 6 three steps per launch through an LDS
 * tile with a two-cell ring (same conditions; the steps left over run one
 * per launch), bit 7 (with bit 4) the two steps streamed down strips of vec
 * columns in chunks of cells_per_thread rows.  Every variant gives the same
 * bits at every slot a fluid cell can read. */
int lbm_variant_info(int variant, int* block, int* cells_per_thread, int* vec, int* flags);
int lbm_set_variant(lbm_handle* h, int variant);
int lbm_get_variant(lbm_handle* h, int* variant);
/* Times every tile shape over `steps` steps on this handle's lattice and
 * keeps the fastest; a shape later in the table replaces the best so far
 * only when it is at least 1 % faster, so shapes of equal speed do not
 * swap places between runs.  The state is restored afterwards. */
int lbm_autotune(lbm_handle* h, int steps, int* best_variant, double* ms_per_step);
/* Steps per HIP-graph chunk used by lbm_run (even, default 100; 0 launches
 * every step directly). */
int lbm_set_graph_chunk(lbm_handle* h, int chunk);
/* Device-to-device copy of `bytes` with a plain streaming kernel on the
 * current device: the HBM ceiling a streaming kernel reaches, and the
 * calibration run for byte counters. */
int lbm_copy_bandwidth(size_t bytes, int reps, double* ms_per_copy);

/* ---- batched domains ----------------------------------------------------
 * No reference counterpart: simuverse runs one lattice per simulator
 * (fluid_simulator.rs:26-133).  A batch holds many independent physics-mode
 * domains of uneven sizes in two shared population buffers, each with its
 * own tau, x edges, inlet velocity, body force and solids; one persistent
 * kernel launch steps every domain (100-step HIP-graph chunks in
 * lbm_batch_run).  Each domain evolves exactly as a single-domain handle
 * with the same descriptor (bit for bit at every slot a fluid cell can
 * read). */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_domain_desc {
    int32_t nx, ny;      /* nx a multiple of 4, >= 8; ny >= 3              */
    float tau;
    int32_t x_edge;      /* LBM_X_*                                        */
    float inlet_ux;      /* LBM_X_INLET_OUTLET: Zou-He inlet velocity      */
    float force_x;       /* Guo body force                                 */
    float force_y;
    int32_t reserved[5]; /* must be zero                                   */
} lbm_domain_desc;

/* storage: LBM_STORAGE_F32 or LBM_STORAGE_F16_DEV (shift 10).  Domains
 * start at rest (f = w) with every cell LBM_MAT_BULK. */
int lbm_batch_create(const lbm_domain_desc* domains, int count, int storage, lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);
int lbm_batch_cells(lbm_batch* b, uint64_t* cells);
/* Whole-domain material map (nx*ny bytes, LBM_MAT_* numbering). */
int lbm_batch_write_materials(lbm_batch* b, int domain, const uint8_t* material);
int lbm_batch_reset(lbm_batch* b);
int lbm_batch_write_populations(lbm_batch* b, int domain, const float* f);
int lbm_batch_read_populations(lbm_batch* b, int domain, float* f);
int lbm_batch_run(lbm_batch* b, int nsteps);
int lbm_batch_run_timed(lbm_batch* b, int nsteps, double* ms);
int lbm_batch_synchronize(lbm_batch* b);
int lbm_batch_macro(lbm_batch* b, int domain, float* rho, float* ux, float* uy);
int lbm_batch_reduce(lbm_batch* b, int domain, double out[5]);
int lbm_batch_variant_count(void);
/* block: threads per block; vec: consecutive cells per lane, or the tile
 * width of a two- or three-step variant; tile_h: tile rows of such a
 * variant (0 otherwise); flags: bit 0 non-temporal stores, bit 2 work
 * grouped per XCD, bit 4 two steps per launch (one block per tile), bit 5
 * persistent blocks take every grid-th chunk, bit 6 three steps per launch
 * (one block per tile, a two-cell ring). */
int lbm_batch_variant_info(int variant, int* block, int* vec, int* tile_h, int* flags);
int lbm_batch_set_variant(lbm_batch* b, int variant);
int lbm_batch_get_variant(lbm_batch* b, int* variant);
int lbm_batch_set_graph_chunk(lbm_batch* b, int chunk);
/* Times every batched shape over `steps` steps on this batch and keeps the
 * fastest, by the same 1 % rule as lbm_autotune; the state is restored
 * afterwards. */
int lbm_batch_autotune(lbm_batch* b, int steps, int* best_variant, double* ms_per_step);

/* ---- learned collision correction --------------------------------------
 * No reference counterpart.  A per-cell network 9 -> 64 -> 64 -> 9 (GELU,
 * erf form) maps the non-equilibrium part of the gathered populations,
 * x_i = (f_i - feq_i) / (w_i * in_scale), to y; its outputs y_0..y_2
 * modulate the non-equilibrium parts n_a of the second-order non-conserved
 * moments a = e, p_xx, p_xy (orthogonal moment basis) by alpha_a * n_a,
 * alpha_a = clamp(out_scale * y_a, -b, b), b = min((2 - omega)/2, omega/2)
 * (a bound that keeps every relaxation factor 1 - omega + alpha_a inside
 * (-1, 1) for any tau > 1/2 and adds no viscosity on average),
 * mapped back to populations (no mass, no
 * momentum, no ghost moments) and added to the BGK post-collision
 * populations of cells whose 3x3 neighbourhood is fluid (the cells the
 * network is trained on; with LBM_X_INLET_OUTLET not the inlet and outlet
 * columns either).  Outputs 3..8 are unused (trained to zero).
 * Parameters: one vector of 5385 floats,
 * W1[64][9] b1[64] W2[64][64] b2[64] W3[9][64] b3[9] (row = output unit).
 * Training (forward, backward, Adam) runs on the device in binary32. */
typedef struct lbm_mlp lbm_mlp;

int lbm_mlp_param_count(void);
/* Glorot-uniform weights from SplitMix64(seed), zero biases; room for
 * `capacity` training samples on the device. */
int lbm_mlp_create(uint64_t seed, int capacity, lbm_mlp** out);
int lbm_mlp_destroy(lbm_mlp* m);
int lbm_mlp_set_params(lbm_mlp* m, const float* params);
int lbm_mlp_get_params(lbm_mlp* m, float* params);
int lbm_mlp_set_scales(lbm_mlp* m, float in_scale, float out_scale);
int lbm_mlp_get_scales(lbm_mlp* m, float* in_scale, float* out_scale);
/* The raw network on n samples (x: n*9 floats, y: n*9 floats, host). */
int lbm_mlp_forward(lbm_mlp* m, const float* x, int n, float* y);
/* The same on the matrix cores: inputs, weights and hidden activations
 * rounded to bf16, binary32 accumulation (LBM_MLP_BF16, the learned
 * step's default). */
int lbm_mlp_forward_bf16(lbm_mlp* m, const float* x, int n, float* y);
/* The same with binary16 operands and GELU (LBM_MLP_F16). */
int lbm_mlp_forward_f16(lbm_mlp* m, const float* x, int n, float* y);
/* Loss 1/(9n) sum (y - t)^2 of n host samples and its gradient (5385). */
int lbm_mlp_loss_grad(lbm_mlp* m, const float* x, const float* t, int n, float* grad, double* loss);
/* lbm_mlp_loss_grad without allocation.  No counterpart in the reference
 * (it has no learned collision); for callers that cannot allocate device
 * memory (code under HIP-graph capture, training or serving loops, harnesses
 * that forbid hipMalloc).  The same kernels and launch shape as
 * lbm_mlp_loss_grad, so the same bits; the staged samples, per-block
 * partials, gradient and loss live in a fixed 6.7 MB device array the
 * library holds on every device it runs on until the process ends.
 * n <= 16384 samples, else LBM_ERR_INVALID.  Every call on one device
 * shares that array: calls that overlap on one device (from several
 * threads, or several networks) race and must be serialised by the
 * caller. */
int lbm_mlp_loss_grad_noalloc(lbm_mlp* m, const float* x, const float* t, int n, float* grad, double* loss);
/* Training samples held on the device. */
int lbm_mlp_clear_samples(lbm_mlp* m);
int lbm_mlp_sample_count(lbm_mlp* m, int* n);
/* Copies stored samples offset .. offset+count-1 to the host (x, t: count*9
 * floats each; either may be NULL), in the order they were stored. */
int lbm_mlp_read_samples(lbm_mlp* m, int offset, int count, float* x, float* t);
int lbm_mlp_add_samples(lbm_mlp* m, const float* x, const float* t, int n);
/* Samples from two consecutive pre-collision states of a periodic lattice
 * (dense f[i*nx*ny + c], binary32): for every stride-th cell from offset
 * whose 3x3 neighbourhood is fluid, x = the network input at time t and
 * t_a = clamp(d_a n_a / (n_a^2 + eps^2)) / out_scale for a = e, p_xx, p_xy (0 for
 * outputs 3..8), d_a the moments of the residual f_i(c + e_i, t+1) -
 * BGK_i(f(c, t)), eps^2 = 1e-2 (n_xx^2 + n_xy^2 + n_e^2 / 9): the
 * modulation that reproduces the residual.  Uses the current scales.
 * Samples are appended in cell order (the same on every call with the same
 * arguments); those beyond the capacity are dropped, and *added counts the
 * stored ones. */
int lbm_mlp_add_transition(lbm_mlp* m, const float* state_t, const float* state_t1, const uint8_t* solid, int nx,
                           int ny, float tau, float force_x, float force_y, int stride, int offset, int* added);
/* Rescales the stored samples to unit root-mean-square (one factor for all
 * inputs, one for all targets) and folds the factors into the network's
 * scales, so a network trained afterwards sees order-one data; returns the
 * factors.  Call after adding samples and before training. */
int lbm_mlp_normalize_samples(lbm_mlp* m, float* in_rms, float* out_rms);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over `steps` minibatches of
 * `batch` samples drawn from the stored ones with SplitMix64(seed). */
int lbm_mlp_train(lbm_mlp* m, int steps, int batch, float lr, uint64_t seed, double* last_loss);

/* Physics-mode handles: steps add the network's projected correction to
 * the BGK collision; NULL switches back to plain BGK.  The network must
 * outlive its use.  Changing the network's parameters or scales later
 * (lbm_mlp_set_params, lbm_mlp_set_scales, lbm_mlp_normalize_samples,
 * lbm_mlp_train) is seen by every later step of every attached handle: the
 * parameter updates first wait for all work queued on the device, and
 * handles re-capture their HIP-graph chunks. */
int lbm_set_learned_collision(lbm_handle* h, const lbm_mlp* m);
/* Arithmetic of the learned step's network: LBM_MLP_BF16 (matrix cores,
 * bf16 operands, binary32 accumulation; the default), LBM_MLP_F16 (matrix
 * cores, binary16 operands and a binary16 GELU: faster and closer to
 * binary32) or LBM_MLP_F32 (binary32 on the vector ALUs). */
enum { LBM_MLP_BF16 = 0, LBM_MLP_F32 = 1, LBM_MLP_F16 = 2 };
int lbm_set_learned_precision(lbm_handle* h, int precision);
/* Largest |u| over fluid cells of the current state and whether every
 * fluid cell is finite with rho > 0. */
int lbm_health(lbm_handle* h, float* umax, int* finite);
/* nsteps steps in chunks of `chunk` (even): after each chunk the health is
 * checked; a chunk run with the learned correction that ends non-finite or
 * with |u| > umax_limit is rerun from its first state with plain BGK (the
 * fallback), and *fallbacks counts those chunks. */
int lbm_run_guarded(lbm_handle* h, int nsteps, int chunk, float umax_limit, int* fallbacks);
/* 4x restriction of the current physics-mode state to an (nx/4) x (ny/4)
 * lattice for tau_coarse and the coarse body force: dense pre-collision
 * coarse populations (9*nc floats), the coarse solid map (a coarse cell is
 * solid when >= 8 of its 16 fine cells are) and, optionally, the coarse
 * velocity (mass-weighted block mean of the fine velocity). */
int lbm_restrict4(lbm_handle* h, float tau_coarse, float force_x_coarse, float force_y_coarse, float* coarse_f,
                  uint8_t* coarse_solid, float* coarse_ux, float* coarse_uy);

#ifdef __cplusplus
}
#endif
#endif /* SIMUVERSE_LBM_H */
