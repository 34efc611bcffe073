// lbm_mlp.hip -- the learned collision correction's network: forward,
// backward and Adam kernels (binary32), the sample store, and its C ABI.
//
// No reference counterpart: jinleili/simuverse has no learned collision.
// The network is 9 -> 64 -> 64 -> 9 with GELU (erf form) after both hidden
// layers.  Parameters are one flat vector of 5385 floats:
//     W1[64][9] b1[64] W2[64][64] b2[64] W3[9][64] b3[9]   (row = output)
// The loss of a batch of B samples is L = 1/(9B) sum (y - t)^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lbm.h"
#include "lbm_internal.hpp"

using namespace lbm;

#define MLP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return lbm_fail(LBM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace lbm {

// SplitMix64 finaliser: sample indices of a minibatch.
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// Forward pass (binary32), one sample per thread, weights in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ params, const float* __restrict__ X,
                                                     float* __restrict__ Y, int n)
{
    __shared__ float P[kMlpParams];
    for (int k = threadIdx.x; k < kMlpParams; k += 256) P[k] = params[k];
    __syncthreads();
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float x[kMlpIn], y[kMlpOut];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) x[i] = X[(size_t)s * kMlpIn + i];
    mlp_forward_one(P, x, y);
#pragma unroll
    for (int m = 0; m < kMlpOut; ++m) Y[(size_t)s * kMlpOut + m] = y[m];
}

// ---------------------------------------------------------------------------
// Matrix-core forward (bf16 or binary16 operands, binary32 accumulation):
// lane-ordered weights in both operand types are rebuilt from the binary32
// parameters whenever those change; a block of 4 wavefronts takes 256
// samples.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ params, __bf16* __restrict__ packed,
                                                  _Float16* __restrict__ packed16)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= kPackElems) return;
    const float w = pack_weight(params, e);
    packed[e] = (__bf16)w;
    packed16[e] = (_Float16)w;
}

// Stages the packed weights and the biases b1, b2, b3 of a network in LDS.
template <typename E>
__device__ inline void stage_net(const E* __restrict__ packed, const float* __restrict__ params, E* W, float* bias)
{
    const uint4* src = reinterpret_cast<const uint4*>(packed);
    uint4* dst = reinterpret_cast<uint4*>(W);
    for (int k = threadIdx.x; k < kPackElems / 8; k += blockDim.x) dst[k] = src[k];
    for (int k = threadIdx.x; k < kMlpH; k += blockDim.x) {
        bias[k] = params[kOffB1 + k];
        bias[kMlpH + k] = params[kOffB2 + k];
    }
    for (int k = threadIdx.x; k < kBiasElems - 2 * kMlpH; k += blockDim.x)
        bias[2 * kMlpH + k] = k < kMlpOut ? params[kOffB3 + k] : 0.0f;
}

template <typename E>
__global__ __launch_bounds__(256) void k_mlp_forward_mma(const E* __restrict__ packed,
                                                         const float* __restrict__ params,
                                                         const float* __restrict__ X, float* __restrict__ Y, int n)
{
    __shared__ __align__(16) E W[kPackElems];
    __shared__ __align__(16) float bias[kBiasElems];
    __shared__ __align__(16) E xs[4][64 * 16];
    __shared__ float ys[4][64 * kMlpOut];
    stage_net(packed, params, W, bias);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s = blockIdx.x * 256 + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 16; ++i)
        xs[wave][lane * 16 + i] = (E)((s < n && i < kMlpIn) ? X[(size_t)s * kMlpIn + i] : 0.0f);
    __syncthreads();
    mlp_wave(W, bias, xs[wave], ys[wave], lane);
    __syncthreads();
    if (s >= n) return;
#pragma unroll
    for (int m = 0; m < kMlpOut; ++m) Y[(size_t)s * kMlpOut + m] = ys[wave][lane * kMlpOut + m];
}

// ---------------------------------------------------------------------------
// Loss gradient: a block of 1024 threads takes chunks of 64 samples (chunks
// b, b + grid, ...); for a chunk, thread (s, q) = (tid % 64, tid / 64)
// computes hidden units 4q..4q+3 of sample s, activations are staged in LDS
// (rows padded to 65 floats), and each thread adds the chunk's sample-sums
// of its parameters p = tid + 1024 r to registers.  One partial per block,
// summed over the (at most kGradBlocks) blocks in block order.  The LDS
// (about 130 KiB) admits one block per CU, so the block is as large as a
// block can be: four wavefronts per SIMD hide each other's LDS latency.
// Every sum runs in the same order for any block size (per parameter:
// samples in chunk order, t = 0..63 within a chunk).
// ---------------------------------------------------------------------------
constexpr int kTS = 64;       // samples per chunk
constexpr int kPadH = kMlpH + 1;
constexpr int kGradBlocks = 1024;
constexpr int kGradThreads = 1024;
constexpr int kUnitGroups = kGradThreads / kTS;  // 16
constexpr int kUnitsPerThread = kMlpH / kUnitGroups;  // 4
constexpr int kPerThread = (kMlpParams + kGradThreads - 1) / kGradThreads;
static_assert(kUnitGroups >= kMlpOut, "one output per unit group");

__global__ __launch_bounds__(kGradThreads) void k_mlp_grad_partial(const float* __restrict__ params,
                                                          const float* __restrict__ X, const float* __restrict__ T,
                                                          int n, int batch, uint64_t seed, int use_index,
                                                          float* __restrict__ partial, double* __restrict__ loss_part)
{
    __shared__ float P[kMlpParams];
    __shared__ float xs[kTS][kMlpIn], dy[kTS][kMlpOut];
    __shared__ float A1[kTS][kPadH], H1[kTS][kPadH], A2[kTS][kPadH], H2[kTS][kPadH], D1[kTS][kPadH],
        D2[kTS][kPadH];
    __shared__ double lsq[kTS][kMlpOut];
    const int tid = threadIdx.x;
    for (int k = tid; k < kMlpParams; k += kGradThreads) P[k] = params[k];
    const int s = tid & (kTS - 1), q = tid >> 6;
    float acc[kPerThread];
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) acc[r] = 0.0f;
    double lblock = 0.0;
    const int nchunks = (batch + kTS - 1) / kTS;
    for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    __syncthreads();  // the previous chunk's LDS reads are done
    const int g = chunk * kTS + s;  // position in the batch
    const bool live = g < batch;
    // sample index: the batch position itself, or a hashed draw from n
    const int idx = !live ? 0 : (use_index ? (int)(mix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(g + 1)) % (uint64_t)n) : g);
    if (q == 0) {
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) xs[s][i] = live ? X[(size_t)idx * kMlpIn + i] : 0.0f;
    }
    __syncthreads();
    // layer 1
    for (int jj = 0; jj < kUnitsPerThread; ++jj) {
        const int j = kUnitsPerThread * q + jj;
        float a = P[kOffB1 + j];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) a += P[kOffW1 + j * kMlpIn + i] * xs[s][i];
        A1[s][j] = a;
        H1[s][j] = gelu(a);
    }
    __syncthreads();
    // layer 2
    for (int jj = 0; jj < kUnitsPerThread; ++jj) {
        const int j = kUnitsPerThread * q + jj;
        float a = P[kOffB2 + j];
#pragma unroll 8
        for (int k = 0; k < kMlpH; ++k) a += P[kOffW2 + j * kMlpH + k] * H1[s][k];
        A2[s][j] = a;
        H2[s][j] = gelu(a);
    }
    __syncthreads();
    // output m = q, its squared residual and dL/dy
    if (q < kMlpOut) {
        const int m = q;
        const float inv = 2.0f / (9.0f * (float)batch);
        float y = P[kOffB3 + m];
#pragma unroll 8
        for (int k = 0; k < kMlpH; ++k) y += P[kOffW3 + m * kMlpH + k] * H2[s][k];
        const float r = live ? y - T[(size_t)idx * kMlpOut + m] : 0.0f;
        lsq[s][m] = (double)r * (double)r;
        dy[s][m] = r * inv;
    }
    __syncthreads();
    for (int jj = 0; jj < kUnitsPerThread; ++jj) {
        const int j = kUnitsPerThread * q + jj;
        float d = 0.0f;
#pragma unroll
        for (int m = 0; m < kMlpOut; ++m) d += P[kOffW3 + m * kMlpH + j] * dy[s][m];
        D2[s][j] = d * gelu_grad(A2[s][j]);
    }
    __syncthreads();
    for (int jj = 0; jj < kUnitsPerThread; ++jj) {
        const int j = kUnitsPerThread * q + jj;
        float d = 0.0f;
#pragma unroll 8
        for (int k = 0; k < kMlpH; ++k) d += P[kOffW2 + k * kMlpH + j] * D2[s][k];
        D1[s][j] = d * gelu_grad(A1[s][j]);
    }
    __syncthreads();
    // this chunk's sample-sums of the weight gradients, into registers
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) {
        const int p = tid + kGradThreads * r;
        if (p >= kMlpParams) break;
        float c = 0.0f;
        if (p < kOffB1) {
            const int j = p / kMlpIn, i = p % kMlpIn;
#pragma unroll 8
            for (int t = 0; t < kTS; ++t) c += D1[t][j] * xs[t][i];
        } else if (p < kOffW2) {
            const int j = p - kOffB1;
#pragma unroll 8
            for (int t = 0; t < kTS; ++t) c += D1[t][j];
        } else if (p < kOffB2) {
            const int j = (p - kOffW2) / kMlpH, k = (p - kOffW2) % kMlpH;
#pragma unroll 8
            for (int t = 0; t < kTS; ++t) c += D2[t][j] * H1[t][k];
        } else if (p < kOffW3) {
            const int j = p - kOffB2;
#pragma unroll 8
            for (int t = 0; t < kTS; ++t) c += D2[t][j];
        } else if (p < kOffB3) {
            const int m = (p - kOffW3) / kMlpH, k = (p - kOffW3) % kMlpH;
#pragma unroll 8
            for (int t = 0; t < kTS; ++t) c += dy[t][m] * H2[t][k];
        } else {
            const int m = p - kOffB3;
#pragma unroll 8
            for (int t = 0; t < kTS; ++t) c += dy[t][m];
        }
        acc[r] += c;
    }
    if (tid == 0)
        for (int t = 0; t < kTS; ++t) {
            double l = 0.0;
            for (int m = 0; m < kMlpOut; ++m) l += lsq[t][m];
            lblock += l;
        }
    }  // chunks
    float* out = partial + (size_t)blockIdx.x * kMlpParams;
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) {
        const int p = tid + kGradThreads * r;
        if (p < kMlpParams) out[p] = acc[r];
    }
    if (tid == 0) loss_part[blockIdx.x] = lblock;
}

// grad[p] = sum over blocks of partial[b][p] in block order.
__global__ __launch_bounds__(256) void k_mlp_grad_reduce(const float* __restrict__ partial, int nblocks,
                                                         float* __restrict__ grad, const double* __restrict__ loss_part,
                                                         double* __restrict__ loss, int batch)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < kMlpParams) {
        float acc = 0.0f;
        for (int b = 0; b < nblocks; ++b) acc += partial[(size_t)b * kMlpParams + p];
        grad[p] = acc;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double l = 0.0;
        for (int b = 0; b < nblocks; ++b) l += loss_part[b];
        *loss = l / (9.0 * (double)batch);
    }
}

// Sum of squares of a[0..count) in binary64: grid-stride, one partial per
// block (the host adds the partials in block order).
__global__ __launch_bounds__(256) void k_sumsq(const float* __restrict__ a, size_t count, double* __restrict__ partial)
{
    double acc = 0.0;
    for (size_t k = blockIdx.x * 256ull + threadIdx.x; k < count; k += (size_t)gridDim.x * 256ull)
        acc += (double)a[k] * (double)a[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ __launch_bounds__(256) void k_scale(float* __restrict__ a, size_t count, float s)
{
    for (size_t k = blockIdx.x * 256ull + threadIdx.x; k < count; k += (size_t)gridDim.x * 256ull) a[k] *= s;
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the binary32 master parameters.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= kMlpParams) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// ---------------------------------------------------------------------------
// Samples from a pair of coarse pre-collision states (state_t, state_t1,
// dense host layout f[i*n + c], binary32 f):
//   features  x_i = (f_i - feq_i) / (w_i s_in)          at cell c, time t
//   target    t_i = P(f_i(c + e_i, t+1) - BGK_i(f(c, t))) / (w_i s_out)
// where P removes the mass and momentum of a correction.  Cells that are
// solid or have a solid neighbour are skipped; y and x are periodic.
// Candidate k is cell k * stride + offset.  The samples are stored in
// candidate order (a count pass, an exclusive scan of the per-block counts,
// then each block writes at its base in thread order), so the sample set,
// and the training that draws from it, are the same on every run.
// ---------------------------------------------------------------------------
__device__ inline bool transition_cell(const uint8_t* __restrict__ solid, int nx, int ny, uint32_t k,
                                       uint32_t stride, uint32_t offset, uint32_t& c)
{
    const uint32_t n = (uint32_t)nx * (uint32_t)ny;
    const uint64_t c64 = (uint64_t)k * stride + offset;
    if (c64 >= n) return false;
    c = (uint32_t)c64;
    const int y = (int)(c / (uint32_t)nx), x = (int)(c % (uint32_t)nx);
    for (int dy2 = -1; dy2 <= 1; ++dy2)
        for (int dx2 = -1; dx2 <= 1; ++dx2) {
            const int xx = (x + dx2 + nx) % nx, yy = (y + dy2 + ny) % ny;
            if (solid[(uint32_t)yy * nx + xx]) return false;
        }
    return true;
}

__global__ __launch_bounds__(256) void k_mlp_transition_count(const uint8_t* __restrict__ solid, int nx, int ny,
                                                              uint32_t stride, uint32_t offset, int* __restrict__ counts)
{
    uint32_t c = 0;
    const bool take = transition_cell(solid, nx, ny, blockIdx.x * 256u + threadIdx.x, stride, offset, c);
    const int total = __syncthreads_count(take ? 1 : 0);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// block_base: first slot of each block's samples (the running sample count
// plus the exclusive scan of the counts, clamped to max_samples).
__global__ __launch_bounds__(256) void k_mlp_transition(const float* __restrict__ s0, const float* __restrict__ s1,
                                                        const uint8_t* __restrict__ solid, PhysArgs a, float s_in,
                                                        float s_out, uint32_t stride, uint32_t offset,
                                                        int max_samples, const int* __restrict__ block_base,
                                                        float* __restrict__ X, float* __restrict__ T)
{
    __shared__ int wave_total[4];
    const int nx = a.nx, ny = a.ny;
    const uint32_t n = (uint32_t)nx * (uint32_t)ny;
    uint32_t c = 0;
    const bool take = transition_cell(solid, nx, ny, blockIdx.x * 256u + threadIdx.x, stride, offset, c);
    // rank of this cell among the block's samples, in thread order
    const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
    const uint64_t mask = __ballot(take ? 1 : 0);
    const int below = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wave_total[wave] = __popcll(mask);
    __syncthreads();
    if (!take) return;
    int rank = below;
    for (int w = 0; w < wave; ++w) rank += wave_total[w];
    const int y = (int)(c / (uint32_t)nx), x = (int)(c % (uint32_t)nx);
    // The step's own arithmetic on deviations g = f - w: moments, features,
    // BGK (+ Guo).
    float g[9], gc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = s0[(size_t)i * n + c] - w_phys(i);
    float rho, ux, uy;
    phys_moments<true>(a, g, rho, ux, uy);
    float xin[9], d[9], nm[3], inv_ws[3];
    learned_input_scales(s_in, inv_ws);
    learned_inputs(g, rho, ux, uy, inv_ws, xin, nm);
#pragma unroll
    for (int i = 0; i < 9; ++i) gc[i] = g[i];
    phys_collide<true>(a, gc, rho, ux, uy);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int tx = (x + ex_of(i) + nx) % nx, ty = (y + ey_of(i) + ny) % ny;
        d[i] = (s1[(size_t)i * n + (uint32_t)ty * nx + tx] - w_phys(i)) - gc[i];
    }
    // Targets of the three modulations (add_stress_correction): residual
    // moment over non-equilibrium moment, d_a / n_a, regularised where n_a
    // is small against the cell's stress: d_a n_a / (n_a^2 + eps^2) with
    // eps^2 = 1e-2 (n_xx^2 + n_xy^2 + n_e^2 / 9), clamped to the bounds the
    // step applies.  nm holds n_e, n_xx, n_xy (learned_inputs).
    float de = 0.0f, dxx = 0.0f, dxy = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const float ex = (float)ex_of(i), ey = (float)ey_of(i);
        de += mom_e(i) * d[i];
        dxx += (ex * ex - ey * ey) * d[i];
        dxy += (ex * ey) * d[i];
    }
    const float ne = nm[0], nxx = nm[1], nxy = nm[2];
    const float eps2 = 1e-2f * (nxx * nxx + nxy * nxy + ne * ne * (1.0f / 9.0f)) + 1e-30f;
    const float lo = modulation_lo(a.omega), hi = modulation_hi(a.omega);
    auto ratio = [&](float dm, float n_a) { return fminf(fmaxf(dm * n_a / (n_a * n_a + eps2), lo), hi); };
    float t[9] = {ratio(de, ne), ratio(dxx, nxx), ratio(dxy, nxy), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const int slot = block_base[blockIdx.x] + rank;
    if (slot >= max_samples) return;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        X[(size_t)slot * 9 + i] = xin[i];
        T[(size_t)slot * 9 + i] = t[i] / s_out;
    }
}

}  // namespace lbm

struct lbm_mlp {
    int dev = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;  // binary32 master copy
    __bf16* packed = nullptr;      // lane-ordered bf16 weights of the matrix-core path
    _Float16* packed16 = nullptr;  // the same in binary16
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;
    float* partial = nullptr;  // per-block gradient partials
    double* loss_part = nullptr;
    double* loss = nullptr;
    int partial_blocks = 0;
    float* X = nullptr;  // samples
    float* T = nullptr;
    int capacity = 0;
    int samples = 0;
    int adam_step = 0;
    float s_in = 1.0f, s_out = 1.0f;
    // Bumped whenever the parameters or scales change: lattices with this
    // network attached rebuild the HIP graphs they captured (a graph holds
    // the scales its launches were recorded with).
    uint64_t generation = 0;
};

namespace {

int mlp_check(lbm_mlp* m)
{
    if (!m) return lbm_fail(LBM_ERR_INVALID, "null network");
    MLP_TRY(hipSetDevice(m->dev));
    return LBM_OK;
}

// Before the parameters change: lattices on other streams may still be
// running learned steps that read them, so every queue of the device drains
// first.  (Updates are rare next to steps; the wait costs one round trip.)
int mlp_quiesce_device(lbm_mlp* m)
{
    MLP_TRY(hipDeviceSynchronize());
    return LBM_OK;
}

void mlp_free(lbm_mlp* m)
{
    for (void* p : {(void*)m->params, (void*)m->packed, (void*)m->packed16, (void*)m->adam_m, (void*)m->adam_v, (void*)m->grad, (void*)m->partial,
                    (void*)m->loss_part, (void*)m->loss, (void*)m->X, (void*)m->T})
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

// Gradient of the loss over `batch` samples (the first `batch` stored
// samples, or hashed draws when use_index) into m->grad; loss into m->loss.
int mlp_grad(lbm_mlp* m, const float* X, const float* T, int n, int batch, uint64_t seed, int use_index)
{
    const int nb = std::min((batch + kTS - 1) / kTS, kGradBlocks);
    if (nb > m->partial_blocks) {
        if (m->partial) (void)hipFree(m->partial);
        if (m->loss_part) (void)hipFree(m->loss_part);
        m->partial = nullptr;
        m->loss_part = nullptr;
        MLP_TRY(hipMalloc(&m->partial, (size_t)nb * kMlpParams * sizeof(float)));
        MLP_TRY(hipMalloc(&m->loss_part, (size_t)nb * sizeof(double)));
        m->partial_blocks = nb;
    }
    hipLaunchKernelGGL(k_mlp_grad_partial, dim3(nb), dim3(kGradThreads), 0, m->stream, m->params, X, T, n, batch, seed,
                       use_index, m->partial, m->loss_part);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream, m->partial, nb,
                       m->grad, m->loss_part, m->loss, batch);
    MLP_TRY(hipGetLastError());
    return LBM_OK;
}

// Rebuilds the matrix-core weights after the binary32 parameters changed.
int mlp_repack(lbm_mlp* m)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3((kPackElems + 255) / 256), dim3(256), 0, m->stream, m->params, m->packed,
                       m->packed16);
    MLP_TRY(hipGetLastError());
    return LBM_OK;
}

}  // namespace

const __bf16* lbm_mlp_device_packed(const lbm_mlp* m) { return m->packed; }
const _Float16* lbm_mlp_device_packed16(const lbm_mlp* m) { return m->packed16; }

// Device pointer to the binary32 parameters and the two scales of a
// network (used by the learned collision step).
const float* lbm_mlp_device_params(const lbm_mlp* m) { return m->params; }
void lbm_mlp_scales(const lbm_mlp* m, float* s_in, float* s_out)
{
    *s_in = m->s_in;
    *s_out = m->s_out;
}
uint64_t lbm_mlp_generation(const lbm_mlp* m) { return m->generation; }

extern "C" {

int lbm_mlp_param_count(void) { return kMlpParams; }

int lbm_mlp_create(uint64_t seed, int capacity, lbm_mlp** out)
{
    if (!out) return lbm_fail(LBM_ERR_INVALID, "null out");
    *out = nullptr;
    if (capacity < 0) return lbm_fail(LBM_ERR_INVALID, "capacity must be >= 0");
    lbm_mlp* m = new lbm_mlp();
    hipError_t e = hipGetDevice(&m->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    const size_t pb = kMlpParams * sizeof(float);
    if (e == hipSuccess) e = hipMalloc(&m->params, pb);
    if (e == hipSuccess) e = hipMalloc(&m->packed, kPackElems * sizeof(__bf16));
    if (e == hipSuccess) e = hipMalloc(&m->packed16, kPackElems * sizeof(_Float16));
    if (e == hipSuccess) e = hipMalloc(&m->adam_m, pb);
    if (e == hipSuccess) e = hipMalloc(&m->adam_v, pb);
    if (e == hipSuccess) e = hipMalloc(&m->grad, pb);
    if (e == hipSuccess) e = hipMalloc(&m->loss, sizeof(double));
    if (e == hipSuccess && capacity > 0) e = hipMalloc(&m->X, (size_t)capacity * kMlpIn * sizeof(float));
    if (e == hipSuccess && capacity > 0) e = hipMalloc(&m->T, (size_t)capacity * kMlpOut * sizeof(float));
    if (e != hipSuccess) {
        mlp_free(m);
        return lbm_fail(LBM_ERR_OOM, "network allocation: %s", hipGetErrorString(e));
    }
    m->capacity = capacity;
    // Glorot-uniform weights from SplitMix64(seed), zero biases.
    std::vector<float> p(kMlpParams, 0.0f);
    uint64_t st = seed;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int k = 0; k < fan_out * fan_in; ++k) p[off + k] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(kOffW1, kMlpH, kMlpIn);
    fill(kOffW2, kMlpH, kMlpH);
    fill(kOffW3, kMlpOut, kMlpH);
    e = hipMemcpy(m->params, p.data(), pb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->adam_m, 0, pb);
    if (e == hipSuccess) e = hipMemset(m->adam_v, 0, pb);
    if (e != hipSuccess || mlp_repack(m) != LBM_OK || hipStreamSynchronize(m->stream) != hipSuccess) {
        mlp_free(m);
        return lbm_fail(LBM_ERR_HIP, "network init failed");
    }
    *out = m;
    return LBM_OK;
}

int lbm_mlp_destroy(lbm_mlp* m)
{
    if (!m) return LBM_OK;
    (void)hipSetDevice(m->dev);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    mlp_free(m);
    return LBM_OK;
}

int lbm_mlp_set_params(lbm_mlp* m, const float* params)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!params) return lbm_fail(LBM_ERR_INVALID, "null params");
    if ((rc = mlp_quiesce_device(m))) return rc;
    ++m->generation;
    MLP_TRY(hipMemcpyAsync(m->params, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemsetAsync(m->adam_m, 0, kMlpParams * sizeof(float), m->stream));
    MLP_TRY(hipMemsetAsync(m->adam_v, 0, kMlpParams * sizeof(float), m->stream));
    if ((rc = mlp_repack(m))) return rc;
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->adam_step = 0;
    return LBM_OK;
}

int lbm_mlp_get_params(lbm_mlp* m, float* params)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!params) return lbm_fail(LBM_ERR_INVALID, "null params");
    MLP_TRY(hipMemcpyAsync(params, m->params, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return LBM_OK;
}

int lbm_mlp_set_scales(lbm_mlp* m, float in_scale, float out_scale)
{
    if (!m) return lbm_fail(LBM_ERR_INVALID, "null network");
    if (!(in_scale > 0.0f) || !(out_scale > 0.0f) || !std::isfinite(in_scale) || !std::isfinite(out_scale))
        return lbm_fail(LBM_ERR_INVALID, "scales must be finite and > 0");
    if (in_scale != m->s_in || out_scale != m->s_out) ++m->generation;
    m->s_in = in_scale;
    m->s_out = out_scale;
    return LBM_OK;
}

int lbm_mlp_get_scales(lbm_mlp* m, float* in_scale, float* out_scale)
{
    if (!m || !in_scale || !out_scale) return lbm_fail(LBM_ERR_INVALID, "null argument");
    *in_scale = m->s_in;
    *out_scale = m->s_out;
    return LBM_OK;
}

static int mlp_forward_impl(lbm_mlp* m, const float* x, int n, float* y, int precision)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!x || !y || n < 0) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    if (n == 0) return LBM_OK;
    float *dx = nullptr, *dy = nullptr;
    MLP_TRY(hipMalloc(&dx, (size_t)n * kMlpIn * sizeof(float)));
    hipError_t e = hipMalloc(&dy, (size_t)n * kMlpOut * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        if (precision == LBM_MLP_BF16)
            hipLaunchKernelGGL(k_mlp_forward_mma<__bf16>, dim3((n + 255) / 256), dim3(256), 0, m->stream,
                               m->packed, m->params, dx, dy, n);
        else if (precision == LBM_MLP_F16)
            hipLaunchKernelGGL(k_mlp_forward_mma<_Float16>, dim3((n + 255) / 256), dim3(256), 0, m->stream,
                               m->packed16, m->params, dx, dy, n);
        else
            hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->params, dx, dy, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "network forward: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_mlp_forward(lbm_mlp* m, const float* x, int n, float* y)
{
    return mlp_forward_impl(m, x, n, y, LBM_MLP_F32);
}

int lbm_mlp_forward_bf16(lbm_mlp* m, const float* x, int n, float* y)
{
    return mlp_forward_impl(m, x, n, y, LBM_MLP_BF16);
}

int lbm_mlp_forward_f16(lbm_mlp* m, const float* x, int n, float* y)
{
    return mlp_forward_impl(m, x, n, y, LBM_MLP_F16);
}

int lbm_mlp_loss_grad(lbm_mlp* m, const float* x, const float* t, int n, float* grad, double* loss)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!x || !t || n < 1) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    float *dx = nullptr, *dt = nullptr;
    MLP_TRY(hipMalloc(&dx, (size_t)n * kMlpIn * sizeof(float)));
    hipError_t e = hipMalloc(&dt, (size_t)n * kMlpOut * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, t, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        rc = mlp_grad(m, dx, dt, n, n, 0, 0);
        if (rc != LBM_OK) e = hipErrorUnknown;
    }
    if (e == hipSuccess && grad)
        e = hipMemcpyAsync(grad, m->grad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (rc != LBM_OK) return rc;
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "network gradient: %s", hipGetErrorString(e));
    return LBM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Loss gradient without allocation (lbm_mlp_loss_grad_noalloc): the staged
// samples, the per-block partials, the gradient and the loss live in one
// file-scope device array, `scratch`, of fixed size (every device the
// library runs on holds its own copy until the process ends).  Its device
// address is looked up once per device with hipGetSymbolAddress; host code
// passes that address, never the array's name, to copies and kernels.
// Every call on one device shares the array: overlapping calls race.
// Nothing in it needs clearing: each block writes its whole partial, and
// the reduction writes the gradient and the loss.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocSamples = 16384;
constexpr int kNoAllocBlocks = (kNoAllocSamples + kTS - 1) / kTS;  // 256 gradient blocks
static_assert(kNoAllocBlocks <= kGradBlocks, "one partial per gradient block");
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kScrPartial = 0;
constexpr size_t kScrLossPart = kScrPartial + align256((size_t)kNoAllocBlocks * kMlpParams * sizeof(float));
constexpr size_t kScrGrad = kScrLossPart + align256((size_t)kNoAllocBlocks * sizeof(double));
constexpr size_t kScrLoss = kScrGrad + align256((size_t)kMlpParams * sizeof(float));
constexpr size_t kScrX = kScrLoss + align256(sizeof(double));
constexpr size_t kScrT = kScrX + align256((size_t)kNoAllocSamples * kMlpIn * sizeof(float));
constexpr size_t kScratchBytes = kScrT + align256((size_t)kNoAllocSamples * kMlpOut * sizeof(float));  // 6.7 MB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (which must be current).
int scratch_address(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= kMaxDevices) return lbm_fail(LBM_ERR_INVALID, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lbm_fail(LBM_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_mlp_loss_grad_noalloc(lbm_mlp* m, const float* x, const float* t, int n, float* grad, double* loss)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!x || !t || n < 1) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    if (n > kNoAllocSamples)
        return lbm_fail(LBM_ERR_INVALID, "%d samples exceed the %d that lbm_mlp_loss_grad_noalloc holds", n,
                        kNoAllocSamples);
    unsigned char* base = nullptr;
    if ((rc = scratch_address(m->dev, &base))) return rc;
    float* dx = reinterpret_cast<float*>(base + kScrX);
    float* dt = reinterpret_cast<float*>(base + kScrT);
    float* partial = reinterpret_cast<float*>(base + kScrPartial);
    double* loss_part = reinterpret_cast<double*>(base + kScrLossPart);
    float* dgrad = reinterpret_cast<float*>(base + kScrGrad);
    double* dloss = reinterpret_cast<double*>(base + kScrLoss);
    // the same launch shape as lbm_mlp_loss_grad, so the sums run in the
    // same order and the results are the same bits
    const int nb = std::min((n + kTS - 1) / kTS, kGradBlocks);
    MLP_TRY(hipMemcpyAsync(dx, x, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(dt, t, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_grad_partial, dim3(nb), dim3(kGradThreads), 0, m->stream, m->params, dx, dt, n, n,
                       (uint64_t)0, 0, partial, loss_part);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream, partial, nb,
                       dgrad, loss_part, dloss, n);
    MLP_TRY(hipGetLastError());
    if (grad) MLP_TRY(hipMemcpyAsync(grad, dgrad, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (loss) MLP_TRY(hipMemcpyAsync(loss, dloss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return LBM_OK;
}

int lbm_mlp_clear_samples(lbm_mlp* m)
{
    if (!m) return lbm_fail(LBM_ERR_INVALID, "null network");
    m->samples = 0;
    return LBM_OK;
}

int lbm_mlp_normalize_samples(lbm_mlp* m, float* in_rms, float* out_rms)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (m->samples < 1) return lbm_fail(LBM_ERR_INVALID, "no samples to normalize");
    constexpr int kBlocks = 256;
    double* part = nullptr;
    MLP_TRY(hipMalloc(&part, 2 * kBlocks * sizeof(double)));
    const size_t nx = (size_t)m->samples * kMlpIn, nt = (size_t)m->samples * kMlpOut;
    hipLaunchKernelGGL(k_sumsq, dim3(kBlocks), dim3(256), 0, m->stream, m->X, nx, part);
    hipLaunchKernelGGL(k_sumsq, dim3(kBlocks), dim3(256), 0, m->stream, m->T, nt, part + kBlocks);
    double host[2 * kBlocks];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, part, sizeof host, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(part);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "sample statistics: %s", hipGetErrorString(e));
    double sx = 0.0, st = 0.0;
    for (int b = 0; b < kBlocks; ++b) {
        sx += host[b];
        st += host[kBlocks + b];
    }
    const float rx = (float)std::sqrt(sx / (double)nx), rt = (float)std::sqrt(st / (double)nt);
    if (!(rx > 0.0f) || !(rt > 0.0f) || !std::isfinite(rx) || !std::isfinite(rt))
        return lbm_fail(LBM_ERR_INVALID, "samples have zero or non-finite spread");
    hipLaunchKernelGGL(k_scale, dim3(kBlocks), dim3(256), 0, m->stream, m->X, nx, 1.0f / rx);
    hipLaunchKernelGGL(k_scale, dim3(kBlocks), dim3(256), 0, m->stream, m->T, nt, 1.0f / rt);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->s_in *= rx;
    m->s_out *= rt;
    ++m->generation;
    if (in_rms) *in_rms = rx;
    if (out_rms) *out_rms = rt;
    return LBM_OK;
}

int lbm_mlp_sample_count(lbm_mlp* m, int* n)
{
    if (!m || !n) return lbm_fail(LBM_ERR_INVALID, "null argument");
    *n = m->samples;
    return LBM_OK;
}

int lbm_mlp_read_samples(lbm_mlp* m, int offset, int count, float* x, float* t)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (offset < 0 || count < 0 || offset > m->samples || count > m->samples - offset)
        return lbm_fail(LBM_ERR_INVALID, "samples [%d, %d) outside the %d stored", offset, offset + count, m->samples);
    if (count == 0) return LBM_OK;
    if (x)
        MLP_TRY(hipMemcpyAsync(x, m->X + (size_t)offset * kMlpIn, (size_t)count * kMlpIn * sizeof(float),
                               hipMemcpyDeviceToHost, m->stream));
    if (t)
        MLP_TRY(hipMemcpyAsync(t, m->T + (size_t)offset * kMlpOut, (size_t)count * kMlpOut * sizeof(float),
                               hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return LBM_OK;
}

int lbm_mlp_add_samples(lbm_mlp* m, const float* x, const float* t, int n)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!x || !t || n < 0 || m->samples + n > m->capacity)
        return lbm_fail(LBM_ERR_INVALID, "samples exceed the capacity of %d", m->capacity);
    MLP_TRY(hipMemcpyAsync(m->X + (size_t)m->samples * kMlpIn, x, (size_t)n * kMlpIn * sizeof(float),
                           hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->T + (size_t)m->samples * kMlpOut, t, (size_t)n * kMlpOut * sizeof(float),
                           hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->samples += n;
    return LBM_OK;
}

int lbm_mlp_add_transition(lbm_mlp* m, const float* state_t, const float* state_t1, const uint8_t* solid, int nx,
                           int ny, float tau, float force_x, float force_y, int stride, int offset, int* added)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (!state_t || !state_t1 || !solid || nx < 3 || ny < 3 || stride < 1 || offset < 0 || !(tau > 0.5f))
        return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    const size_t n = (size_t)nx * (size_t)ny;
    float *s0 = nullptr, *s1 = nullptr;
    uint8_t* sd = nullptr;
    MLP_TRY(hipMalloc(&s0, 9 * n * sizeof(float)));
    hipError_t e = hipMalloc(&s1, 9 * n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&sd, n);
    if (e == hipSuccess) e = hipMemcpyAsync(s0, state_t, 9 * n * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s1, state_t1, 9 * n * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sd, solid, n, hipMemcpyHostToDevice, m->stream);
    const size_t cand = (n - (size_t)std::min<int>(offset, (int)n) + stride - 1) / stride;
    const size_t nblk = (cand + 255) / 256;
    int* blk = nullptr;  // per-block counts, then per-block first slots
    std::vector<int> base(nblk);
    int64_t total = m->samples;
    if (e == hipSuccess && nblk > 0) e = hipMalloc(&blk, nblk * sizeof(int));
    if (e == hipSuccess && nblk > 0) {
        hipLaunchKernelGGL(k_mlp_transition_count, dim3((unsigned)nblk), dim3(256), 0, m->stream, sd, nx, ny,
                           (uint32_t)stride, (uint32_t)offset, blk);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(base.data(), blk, nblk * sizeof(int), hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e == hipSuccess) {
            // exclusive scan from the running count; bases past the capacity
            // are clamped (those blocks store nothing)
            for (size_t b = 0; b < nblk; ++b) {
                const int cnt = base[b];
                base[b] = (int)std::min<int64_t>(total, m->capacity);
                total += cnt;
            }
            e = hipMemcpyAsync(blk, base.data(), nblk * sizeof(int), hipMemcpyHostToDevice, m->stream);
        }
        if (e == hipSuccess) {
            PhysArgs pa{};
            pa.nx = nx;
            pa.ny = ny;
            pa.omega = 1.0f / tau;
            pa.gx = force_x;
            pa.gy = force_y;
            pa.scale = pa.inv_scale = 1.0f;
            set_force_terms(pa);
            hipLaunchKernelGGL(k_mlp_transition, dim3((unsigned)nblk), dim3(256), 0, m->stream, s0, s1, sd, pa,
                               m->s_in, m->s_out, (uint32_t)stride, (uint32_t)offset, m->capacity, blk, m->X, m->T);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(s0);
    if (s1) (void)hipFree(s1);
    if (sd) (void)hipFree(sd);
    if (blk) (void)hipFree(blk);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "transition samples: %s", hipGetErrorString(e));
    total = std::min<int64_t>(total, m->capacity);
    if (added) *added = (int)(total - m->samples);
    m->samples = (int)total;
    return LBM_OK;
}

int lbm_mlp_train(lbm_mlp* m, int steps, int batch, float lr, uint64_t seed, double* last_loss)
{
    int rc = mlp_check(m);
    if (rc) return rc;
    if (steps < 0 || batch < 1 || !(lr > 0.0f)) return lbm_fail(LBM_ERR_INVALID, "bad arguments");
    if (m->samples < 1) return lbm_fail(LBM_ERR_INVALID, "no samples to train on");
    if ((rc = mlp_quiesce_device(m))) return rc;
    ++m->generation;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < steps; ++s) {
        rc = mlp_grad(m, m->X, m->T, m->samples, batch, seed + 0xD1B54A32D192ED03ull * (uint64_t)(m->adam_step + 1), 1);
        if (rc != LBM_OK) return rc;
        ++m->adam_step;
        const float c1 = 1.0f - std::pow(b1, (float)m->adam_step);
        const float c2 = 1.0f - std::pow(b2, (float)m->adam_step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream, m->params, m->grad,
                           m->adam_m, m->adam_v, lr, b1, b2, eps, c1, c2);
        MLP_TRY(hipGetLastError());
    }
    if ((rc = mlp_repack(m))) return rc;
    if (last_loss) MLP_TRY(hipMemcpyAsync(last_loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return LBM_OK;
}

}  // extern "C"
