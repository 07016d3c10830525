// Learned planner: a 6 -> 64 -> 64 -> 2 FP32 MLP with ReLU, trained and run on
// the GPU by the kernels below.
//
// Inputs (per left tile, all log1p-scaled, then standardised):
//   left-tile nnz, distinct contraction keys, summed right-fiber lengths,
//   products bound, distinct-column sketch estimate, tile area TL*spanB.
// Outputs (log1p space, standardised), each relative to a tight upper bound
// known from the features:
//   r0 = log1p(tile output nnz)      - min(log1p(products), log1p(area))
//   r1 = log1p(distinct out columns) - log1p(summed fiber lengths)
// Both are <= 0; predicting the gap to the bound keeps predictions for tiles
// unlike the training set near the bound (safe) instead of near zero.
// The exact count pass produces the labels.
//
// The reference has no planner: its tile size is the caller's argument
// (contract.hpp:844-846); the features restate its per-tile statistics
// (num_nnz_in_tile index_tensors.hpp:562, num_active_columns :604,
// bound_output_nnz_* contract.hpp:180-211).
//
// Flat parameter layout (floats):
//   W1[6][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
//   in_mean[6] in_std[6] out_mean[2] out_std[2]
#include "fcc_mlp.h"

namespace fcc {

constexpr int NI = MLP_IN, NH = MLP_HID, NO = MLP_OUT;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO, OFF_IM = OFF_B3 + NO, OFF_IS = OFF_IM + NI,
              OFF_OM = OFF_IS + NI, OFF_OS = OFF_OM + NO, N_PARAMS = OFF_OS + NO;
static_assert(N_PARAMS == MLP_PARAMS, "parameter layout");
constexpr int N_TRAIN = OFF_IM;  // trainable prefix

// ---------------------------------------------------------------- features

__global__ void k_tile_features(const uint32_t *st_nnz, const uint32_t *st_keys, const uint64_t *st_sumfib,
                                const uint64_t *st_products, const uint32_t *st_sketch, uint32_t natiles,
                                float area, float *feat) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles) return;
    double h = ((double)st_sketch[2 * a] + (double)st_sketch[2 * a + 1] + 2.0) * 0.5;
    double dist = 4294967296.0 / h - 1.0;
    float *f = feat + (size_t)a * NI;
    f[0] = log1pf((float)st_nnz[a]);
    f[1] = log1pf((float)st_keys[a]);
    f[2] = log1pf((float)st_sumfib[a]);
    f[3] = log1pf((float)st_products[a]);
    f[4] = log1pf((float)(dist > 0 ? dist : 0));
    f[5] = log1pf(area);
}

// ------------------------------------------------------------- inference
// One thread per tile; weights staged in LDS and read as broadcasts. The
// second layer's 64 sums are taken 16 at a time: a rolled loop over the first
// layer's units j computes unit j and adds it to the 16 sums, so they are
// independent chains (each still adds in j order, then the output layer adds
// k in order: the plain row-by-row evaluation's result). The 16-sum chunks
// keep the kernel small in registers and code: it runs beside the dense-key
// GEMM on the other stream (a 400-register version waited for the GEMM to
// drain) and from the instruction cache.
constexpr int KC = 16;  // second-layer sums per chunk
__global__ __launch_bounds__(256) void k_mlp_infer(const float *P, const float *feat, uint32_t n, float *pred_nnz,
                                                   float *pred_cols) {
    __shared__ float w[N_PARAMS];
    for (int i = threadIdx.x; i < N_PARAMS; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const float *f = feat + (size_t)t * NI;
        float x[NI];
#pragma unroll
        for (int i = 0; i < NI; i++) x[i] = (f[i] - w[OFF_IM + i]) / w[OFF_IS + i];
        float y0 = w[OFF_B3], y1 = w[OFF_B3 + 1];
#pragma unroll 1
        for (int k0 = 0; k0 < NH; k0 += KC) {
            float h2[KC];
#pragma unroll
            for (int k = 0; k < KC; k++) h2[k] = w[OFF_B2 + k0 + k];
#pragma unroll 2
            for (int j = 0; j < NH; j++) {
                float h1 = w[OFF_B1 + j];
#pragma unroll
                for (int i = 0; i < NI; i++) h1 += x[i] * w[OFF_W1 + i * NH + j];
                h1 = h1 > 0.f ? h1 : 0.f;
#pragma unroll
                for (int k = 0; k < KC; k++) h2[k] += h1 * w[OFF_W2 + j * NH + k0 + k];
            }
#pragma unroll
            for (int k = 0; k < KC; k++) {
                const float r = h2[k] > 0.f ? h2[k] : 0.f;
                y0 += r * w[OFF_W3 + (k0 + k) * NO];
                y1 += r * w[OFF_W3 + (k0 + k) * NO + 1];
            }
        }
        float r0 = fminf(y0 * w[OFF_OS] + w[OFF_OM], 0.f);
        float r1 = fminf(y1 * w[OFF_OS + 1] + w[OFF_OM + 1], 0.f);
        float ub0 = fminf(f[3], f[5]), ub1 = f[2];
        pred_nnz[t] = fmaxf(expm1f(fminf(r0 + ub0, 80.f)), 0.f);
        if (pred_cols) pred_cols[t] = fmaxf(expm1f(fminf(r1 + ub1, 80.f)), 0.f);
    }
}

// ------------------------------------------------------ forward + backward
// Each block takes SB samples: forward per sample (one thread each), the
// per-sample deltas kept in LDS, then the weight gradients reduced over the
// block's samples with threads spread across the weight matrix, and added to
// the global gradient with one atomic per weight per block.
constexpr int SB = 64;

__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float *P, const float *feat, const float *label,
                                                     const uint32_t *perm, uint32_t n, float inv_count, float *grad,
                                                     float *loss, float *pred) {
    __shared__ float w[N_PARAMS];
    __shared__ float sx[SB][NI];
    __shared__ float sh1[SB][NH];
    __shared__ float sh2[SB][NH];
    __shared__ float sd2[SB][NH];  // dL/d(pre-activation of layer 2)
    __shared__ float sd1[SB][NH];  // dL/d(pre-activation of layer 1)
    __shared__ float sdy[SB][NO];
    __shared__ float sloss[SB];
    for (int i = threadIdx.x; i < N_PARAMS; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    const uint32_t s0 = blockIdx.x * SB;
    const int tid = threadIdx.x;
    // forward: threads 0..SB-1, one sample each
    if (tid < SB) {
        uint32_t si = s0 + tid;
        bool ok = si < n;
        uint32_t idx = ok ? (perm ? perm[si] : si) : 0;
        float x[NI];
        for (int i = 0; i < NI; i++) {
            x[i] = ok ? (feat[(size_t)idx * NI + i] - w[OFF_IM + i]) / w[OFF_IS + i] : 0.f;
            sx[tid][i] = x[i];
        }
        for (int j = 0; j < NH; j++) {
            float s = w[OFF_B1 + j];
            for (int i = 0; i < NI; i++) s += x[i] * w[OFF_W1 + i * NH + j];
            sh1[tid][j] = s > 0.f ? s : 0.f;
        }
        float y[NO];
        for (int o = 0; o < NO; o++) y[o] = w[OFF_B3 + o];
        for (int k = 0; k < NH; k++) {
            float s = w[OFF_B2 + k];
            for (int j = 0; j < NH; j++) s += sh1[tid][j] * w[OFF_W2 + j * NH + k];
            s = s > 0.f ? s : 0.f;
            sh2[tid][k] = s;
            for (int o = 0; o < NO; o++) y[o] += s * w[OFF_W3 + k * NO + o];
        }
        float l = 0.f;
        for (int o = 0; o < NO; o++) {
            float t = ok ? (label[(size_t)idx * NO + o] - w[OFF_OM + o]) / w[OFF_OS + o] : 0.f;
            float d = ok ? (y[o] - t) : 0.f;
            l += d * d;
            sdy[tid][o] = 2.f * d * inv_count;  // d(mean squared error)/dy
            if (pred && ok) pred[(size_t)si * NO + o] = y[o];
        }
        sloss[tid] = l * inv_count;
    }
    __syncthreads();
    // layer-2 deltas: thread (sample, k) pairs
    for (int e = tid; e < SB * NH; e += blockDim.x) {
        int sm = e / NH, k = e % NH;
        float d = 0.f;
        for (int o = 0; o < NO; o++) d += sdy[sm][o] * w[OFF_W3 + k * NO + o];
        sd2[sm][k] = sh2[sm][k] > 0.f ? d : 0.f;
    }
    __syncthreads();
    for (int e = tid; e < SB * NH; e += blockDim.x) {
        int sm = e / NH, j = e % NH;
        float d = 0.f;
        for (int k = 0; k < NH; k++) d += sd2[sm][k] * w[OFF_W2 + j * NH + k];
        sd1[sm][j] = sh1[sm][j] > 0.f ? d : 0.f;
    }
    __syncthreads();
    // weight gradients
    for (int e = tid; e < N_TRAIN; e += blockDim.x) {
        float g = 0.f;
        if (e < OFF_B1) {
            int i = e / NH, j = e % NH;
            for (int sm = 0; sm < SB; sm++) g += sx[sm][i] * sd1[sm][j];
        } else if (e < OFF_W2) {
            int j = e - OFF_B1;
            for (int sm = 0; sm < SB; sm++) g += sd1[sm][j];
        } else if (e < OFF_B2) {
            int q = e - OFF_W2, j = q / NH, k = q % NH;
            for (int sm = 0; sm < SB; sm++) g += sh1[sm][j] * sd2[sm][k];
        } else if (e < OFF_W3) {
            int k = e - OFF_B2;
            for (int sm = 0; sm < SB; sm++) g += sd2[sm][k];
        } else if (e < OFF_B3) {
            int q = e - OFF_W3, k = q / NO, o = q % NO;
            for (int sm = 0; sm < SB; sm++) g += sh2[sm][k] * sdy[sm][o];
        } else {
            int o = e - OFF_B3;
            for (int sm = 0; sm < SB; sm++) g += sdy[sm][o];
        }
        if (g != 0.f) atomicAdd(&grad[e], g);
    }
    if (tid == 0) {
        float l = 0.f;
        for (int sm = 0; sm < SB; sm++) l += sloss[sm];
        atomicAdd(loss, l);
    }
}

__global__ void k_adam(float *P, const float *grad, float *m, float *v, int n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = grad[i];
    float mi = b1 * m[i] + (1.f - b1) * g;
    float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_zero_f32(float *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ------------------------------------------------------------ host side

void mlp_tile_features(const Index &ix, uint64_t spanB, float *feat, hipStream_t s) {
    if (!ix.natiles) return;
    float area = (float)((double)ix.TL * (double)spanB);
    k_tile_features<<<(unsigned)div_up(ix.natiles, 256), 256, 0, s>>>(
        ix.st_nnz, ix.st_keys, ix.st_sumfib, ix.st_products, ix.st_sketch, ix.natiles, area, feat);
}

void mlp_infer(const float *dparams, const float *feat, uint32_t n, float *pred_nnz, float *pred_cols,
               hipStream_t s) {
    if (!n) return;
    // at most 2048 workgroups (each stages the weights once), looping over the rest
    k_mlp_infer<<<(unsigned)std::min<uint64_t>(div_up(n, 256), 2048), 256, 0, s>>>(dparams, feat, n, pred_nnz,
                                                                                     pred_cols);
}

void mlp_fwd_bwd(const float *dparams, const float *feat, const float *label, const uint32_t *perm, uint32_t n,
                 float *grad, float *loss, float *pred, hipStream_t s) {
    k_zero_f32<<<(unsigned)div_up(N_TRAIN, 256), 256, 0, s>>>(grad, N_TRAIN);
    k_zero_f32<<<1, 1, 0, s>>>(loss, 1);
    if (!n) return;
    float inv = 1.0f / (float)(n * NO);
    k_mlp_fwd_bwd<<<(unsigned)div_up(n, SB), 256, 0, s>>>(dparams, feat, label, perm, n, inv, grad, loss, pred);
}

void mlp_adam(float *dparams, const float *grad, float *m, float *v, float lr, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
    k_adam<<<(unsigned)div_up(N_TRAIN, 256), 256, 0, s>>>(dparams, grad, m, v, N_TRAIN, lr, b1, b2, eps, bc1, bc2);
}

int mlp_trainable_count() { return N_TRAIN; }

}  // namespace fcc
