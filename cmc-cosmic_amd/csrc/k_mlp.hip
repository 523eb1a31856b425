// Learned apsides brackets: a 6 -> 64 -> 64 -> 2 SiLU network (FP32) that predicts
// (log r_peri, log r_apo) of a star from (E, J, log r, phi(r), log(r^2 dphi/dr), m).
// The prediction only proposes the starting bracket of find_zero_Q
// (src/cmc/cmc_nr.c:316-383); the exact bisection decides the index and the full
// bracket is used whenever the proposal does not bracket the sign change.
//
// Training data are harvested from the method itself after each orbit phase
// (every sampled star's features and its exact apsides); training is minibatch
// Adam with the forward and backward passes below.
//
// Parameters: one flat float vector W1[64][6], b1[64], W2[64][64], b2[64],
// W3[2][64], b3[2]; inputs and targets are standardised with per-column mean and
// standard deviation fitted on the harvested rows (MlpNorm).
#include "kernels.hpp"

namespace {
constexpr int NIN = MLP_NIN, NH = MLP_NH, NOUT = MLP_NOUT;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NOUT * NH;
static_assert(OFF_B3 + NOUT == MLP_NPARAM, "parameter layout");
constexpr int TILE = 64;            // samples per tile
constexpr int FT = 256;             // threads per training block

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

typedef float float4_t __attribute__((ext_vector_type(4)));
constexpr int LDP = NH + 1;         // padded LDS row (conflict-free matrix-core operand reads)

// The wave's 16 x 64 block of A (64 x K) B (K x 64) on the matrix cores (16x16x4 f32
// MFMAs, f32 accumulation into acc): rows i0 .. i0 + 15 of A, all 64 columns of B in
// four 16-column blocks.  Lane l supplies A[i0 + l % 16][4 kc + l / 16] and
// B[4 kc + l / 16][16 cb + l % 16] and holds C[i0 + 4 (l / 16) + v][16 cb + l % 16] in
// acc[cb][v].
template <int K, class FA, class FB>
__device__ __forceinline__ void wave_gemm_16x64(int i0, FA a, FB b, float4_t (&acc)[4])
{
    const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll 4
    for (int kc = 0; kc < K / 4; kc++) {
        const float av = a(i0 + r, 4 * kc + q);
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b(4 * kc + q, 16 * cb + r), acc[cb], 0, 0, 0);
    }
}

// row r of the training set for minibatch position i: a multiplicative shuffle
__device__ __forceinline__ int64_t shuffled(int64_t i, int64_t count, uint64_t mult, uint64_t off)
{
    return (int64_t)(((unsigned __int128)(uint64_t)i * mult + off) % (uint64_t)count);
}

// One block: tiles blockIdx.x, blockIdx.x + gridDim.x, ... of the batch.  Writes the
// block's gradient partial (MLP_NPARAM floats) and its loss partial.  The three 64 x
// 64 x 64 products of a tile (layer 2 forward, layer 2 weight gradient, error at
// layer 1) run on the matrix cores, each wave taking 16 rows; the 6- and 2-wide
// layers stay on the vector units.
struct mlp_fwdbwd_kernel {
    static constexpr int LB = FT, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const float *__restrict__ theta, const MlpNorm *__restrict__ nrm,
                                                        const float *__restrict__ rows, int64_t count, int64_t first,
                                                        int64_t batch, uint64_t mult, uint64_t off, float inv_b,
                                                        float *__restrict__ gpart, double *__restrict__ lpart)
{
    __shared__ float W1T[NIN][NH], B1[NH], W2[NH][LDP], B2[NH], W3[NOUT][NH], B3[NOUT];
    __shared__ float X[TILE][NIN], T[TILE][NOUT], D3[TILE][NOUT];
    __shared__ float Z1[TILE][LDP], H1[TILE][LDP], Z2[TILE][LDP], H2[TILE][LDP];
    __shared__ float xm[NIN], xs[NIN], ym[NOUT], ys[NOUT];
    const int t = threadIdx.x, wv = t >> 6, lane = t & 63, lr = lane & 15, lq = lane >> 4;
    for (int e = t; e < NH * NIN; e += FT) W1T[e % NIN][e / NIN] = theta[OFF_W1 + e];
    for (int e = t; e < NH * NH; e += FT) W2[e / NH][e % NH] = theta[OFF_W2 + e];
    for (int e = t; e < NOUT * NH; e += FT) W3[e / NH][e % NH] = theta[OFF_W3 + e];
    if (t < NH) { B1[t] = theta[OFF_B1 + t]; B2[t] = theta[OFF_B2 + t]; }
    if (t < NOUT) { B3[t] = theta[OFF_B3 + t]; ym[t] = nrm->ym[t]; ys[t] = nrm->ys[t]; }
    if (t < NIN) { xm[t] = nrm->xm[t]; xs[t] = nrm->xs[t]; }
    // gradient accumulators owned by this thread (gW2: the wave's 16 rows of W2's
    // gradient in the matrix-core layout)
    float4_t gW2[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    float gW1[2] = {0.f, 0.f}, gB = 0.f, gW3 = 0.f, gB3 = 0.f;
    double loss = 0.0;
    __syncthreads();
    const int64_t ntiles = (batch + TILE - 1) / TILE;
    for (int64_t tile = kg_.bid; tile < ntiles; tile += kg_.nb) {
        // 1. inputs and targets (standardised); samples past the batch get zero weight
        if (t < TILE) {
            int64_t i = tile * TILE + t;
            bool ok = i < batch;
            const float *row = rows + 8 * shuffled(first + (ok ? i : 0), count, mult, off);
            for (int c = 0; c < NIN; c++) X[t][c] = ok ? (row[c] - xm[c]) / xs[c] : 0.f;
            for (int c = 0; c < NOUT; c++) T[t][c] = ok ? (row[NIN + c] - ym[c]) / ys[c] : 0.f;
            D3[t][0] = ok ? 1.f : 0.f;     // validity, replaced by the output error below
        }
        __syncthreads();
        // 2. layer 1
        for (int q = 0; q < 16; q++) {
            int e = t + FT * q, s = e / NH, u = e % NH;
            float z = B1[u];
            for (int c = 0; c < NIN; c++) z += W1T[c][u] * X[s][c];
            Z1[s][u] = z;
            H1[s][u] = silu(z);
        }
        __syncthreads();
        // 3. layer 2: Z2[s][u] = sum_v H1[s][v] W2[u][v] + B2[u]
        {
            float4_t acc[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
            wave_gemm_16x64<NH>(16 * wv, [&](int s, int v) { return H1[s][v]; },
                                [&](int v, int u) { return W2[u][v]; }, acc);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int s = 16 * wv + 4 * lq + v, u = 16 * cb + lr;
                    const float z = acc[cb][v] + B2[u];
                    Z2[s][u] = z;
                    H2[s][u] = silu(z);
                }
        }
        __syncthreads();
        // 4. output, loss, output error
        float err = 0.f;
        if (t < TILE * NOUT) {
            int s = t / NOUT, o = t % NOUT;
            float y = B3[o];
            for (int u = 0; u < NH; u++) y += W3[o][u] * H2[s][u];
            float r = (y - T[s][o]) * D3[s][0];
            loss += (double)r * (double)r;
            err = 2.f * r * inv_b;
        }
        __syncthreads();
        if (t < TILE * NOUT) D3[t / NOUT][t % NOUT] = err;
        __syncthreads();
        // 5. layer 3 gradients; error at layer 2 (in place of Z2)
        if (t >= 128) {
            int idx = t - 128, o = idx / NH, u = idx % NH;
            float g = 0.f;
            for (int s = 0; s < TILE; s++) g += D3[s][o] * H2[s][u];
            gW3 += g;
        }
        if (t < NOUT) {
            float g = 0.f;
            for (int s = 0; s < TILE; s++) g += D3[s][t];
            gB3 += g;
        }
        for (int q = 0; q < 16; q++) {
            int e = t + FT * q, s = e / NH, u = e % NH;
            float d = 0.f;
            for (int o = 0; o < NOUT; o++) d += D3[s][o] * W3[o][u];
            Z2[s][u] = d * dsilu(Z2[s][u]);
        }
        __syncthreads();
        // 6. layer 2 gradients (gW2[u][v] += sum_s Z2[s][u] H1[s][v]); error at layer 1
        // (in place of Z1: sum_u Z2[s][u] W2[u][v] times SiLU'(Z1[s][v]))
        wave_gemm_16x64<TILE>(16 * wv, [&](int u, int s) { return Z2[s][u]; },
                              [&](int s, int v) { return H1[s][v]; }, gW2);
        if (t >= NH && t < 2 * NH) {
            float g = 0.f;
            for (int s = 0; s < TILE; s++) g += Z2[s][t - NH];
            gB += g;
        }
        {
            float4_t acc[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
            wave_gemm_16x64<NH>(16 * wv, [&](int s, int u) { return Z2[s][u]; },
                                [&](int u, int v) { return W2[u][v]; }, acc);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int s = 16 * wv + 4 * lq + v, c = 16 * cb + lr;
                    Z1[s][c] = acc[cb][v] * dsilu(Z1[s][c]);
                }
        }
        __syncthreads();
        // 7. layer 1 gradients
        for (int k = 0; k < 2; k++) {
            int e = t + FT * k;
            if (e < NH * NIN) {
                int u = e / NIN, c = e % NIN;
                float g = 0.f;
                for (int s = 0; s < TILE; s++) g += Z1[s][u] * X[s][c];
                gW1[k] += g;
            }
        }
        if (t < NH) {
            float g = 0.f;
            for (int s = 0; s < TILE; s++) g += Z1[s][t];
            gB += g;
        }
        __syncthreads();
    }
    float *gp = gpart + (int64_t)kg_.bid * MLP_NPARAM;
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int v = 0; v < 4; v++) gp[OFF_W2 + (16 * wv + 4 * lq + v) * NH + 16 * cb + lr] = gW2[cb][v];
    for (int k = 0; k < 2; k++)
        if (t + FT * k < NH * NIN) gp[OFF_W1 + t + FT * k] = gW1[k];
    if (t < NH) gp[OFF_B1 + t] = gB;
    else if (t < 2 * NH) gp[OFF_B2 + t - NH] = gB;
    if (t >= 128) gp[OFF_W3 + t - 128] = gW3;
    if (t < NOUT) gp[OFF_B3 + t] = gB3;
    // block loss: fixed-order reduction
    __shared__ double lred[FT];
    lred[t] = loss;
    __syncthreads();
    if (t == 0) {
        double a = 0.0;
        for (int i = 0; i < FT; i++) a += lred[i];
        lpart[kg_.bid] = a;
    }
}
};

// gradient = sum of the block partials in block order; loss likewise (mean over batch)
struct mlp_reduce_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const float *__restrict__ gpart, const double *__restrict__ lpart, int nblocks,
                                  float *__restrict__ grad, double *__restrict__ loss, double inv_b)
{
    int p = kg_.bid * blockDim.x + threadIdx.x;
    if (p < MLP_NPARAM) {
        float a = 0.f;
        for (int b = 0; b < nblocks; b++) a += gpart[(int64_t)b * MLP_NPARAM + p];
        grad[p] = a;
    }
    if (p == 0) {
        double a = 0.0;
        for (int b = 0; b < nblocks; b++) a += lpart[b];
        *loss = a * inv_b;
    }
}
};

// Adam (Kingma and Ba 2015, Algorithm 1), step t counted from 1
struct mlp_adam_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, float *__restrict__ theta, const float *__restrict__ grad, float *__restrict__ m,
                                float *__restrict__ v, int64_t step, float lr)
{
    int p = kg_.bid * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float g = grad[p];
    float mm = b1 * m[p] + (1.f - b1) * g;
    float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float mh = mm / (1.f - powf(b1, (float)step));
    float vh = vv / (1.f - powf(b2, (float)step));
    theta[p] -= lr * mh / (sqrtf(vh) + eps);
}
};

// per-thread inference; weights are wave-uniform reads
struct mlp_predict_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const float *__restrict__ theta, const MlpNorm *__restrict__ nrm,
                                   const float *__restrict__ x, int64_t n, float *__restrict__ y)
{
    int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float in[NIN], out[NOUT];
    for (int c = 0; c < NIN; c++) in[c] = x[NIN * i + c];
    mlp_forward(theta, nrm, in, out);
    for (int o = 0; o < NOUT; o++) y[NOUT * i + o] = out[o];
}
};

// column sums of the harvested rows (mean and variance of every input and target)
struct mlp_colsum_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const float *__restrict__ rows, int64_t count,
                                                         double *__restrict__ part)
{
    __shared__ double lds[16][256];
    double acc[16];
    for (int c = 0; c < 16; c++) acc[c] = 0.0;
    for (int64_t i = (int64_t)kg_.bid * blockDim.x + threadIdx.x; i < count; i += (int64_t)kg_.nb * blockDim.x)
        for (int c = 0; c < 8; c++) {
            double v = rows[8 * i + c];
            acc[c] += v;
            acc[8 + c] += v * v;
        }
    for (int c = 0; c < 16; c++) lds[c][threadIdx.x] = acc[c];
    __syncthreads();
    if (threadIdx.x < 16) {
        double a = 0.0;
        for (int k = 0; k < 256; k++) a += lds[threadIdx.x][k];
        part[16 * kg_.bid + threadIdx.x] = a;
    }
}
};

struct mlp_norm_kernel {
    static constexpr int LB = 1024, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const double *__restrict__ part, int nblocks, int64_t count, MlpNorm *__restrict__ nrm)
{
    int c = threadIdx.x;
    if (c >= 8) return;
    double s = 0.0, s2 = 0.0;
    for (int b = 0; b < nblocks; b++) { s += part[16 * b + c]; s2 += part[16 * b + 8 + c]; }
    double mean = s / (double)count;
    double var = s2 / (double)count - mean * mean;
    float sd = var > 1e-24 ? (float)sqrt(var) : 1.f;
    if (c < NIN) { nrm->xm[c] = (float)mean; nrm->xs[c] = sd; }
    else { nrm->ym[c - NIN] = (float)mean; nrm->ys[c - NIN] = sd; }
}
};

// Training rows from the orbit phase just run: every sampled star (flag 256, set
// only by the orbit sampler before energy_conservation2) appends its features and
// exact apsides to the ring buffer.
struct mlp_harvest_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                          float *__restrict__ rows, int64_t cap,
                                                          unsigned long long *__restrict__ counter)
{
    int64_t j = 1 + (int64_t)kg_.bid * blockDim.x + threadIdx.x;
    int64_t n = s->n_max;
    bool take = j <= n && (f.flags[j] & 256) && !(f.flags[j] & (CMCG_SF_LINSCAN | CMCG_SF_BRENT));
    float x[NIN];
    if (take) {
        double rj = f.r[j], mj = f.m[j];
        double E = f.E[j] + phi_s_own(rj, mj, P.inv_nstar, rj);
        mlp_features(f.r, f.phi, n, j, rj, mj, E, f.J[j], x);
    }
    uint64_t bal = __ballot(take);
    if (!bal) return;
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == __ffsll((long long)bal) - 1) base = atomicAdd(counter, (unsigned long long)__popcll(bal));
    base = __shfl(base, __ffsll((long long)bal) - 1);
    if (!take) return;
    int64_t pos = (int64_t)((base + __popcll(bal & ((1ull << lane) - 1ull))) % (unsigned long long)cap);
    float *row = rows + 8 * pos;
    for (int c = 0; c < NIN; c++) row[c] = x[c];
    row[NIN] = logf((float)f.r_peri[j]);
    row[NIN + 1] = logf((float)f.r_apo[j]);
}
};

// ------------------------------------------------------------------ matrix-core inference
// The network on the matrix cores: a wave evaluates 16 stars at a time as transposed
// products H1^T = SiLU(W1 X^T + b1), H2^T = SiLU(W2 H1^T + b2), Y^T = W3 H2^T + b3 with
// 16x16x16 f16 MFMAs (f32 accumulation).  In that orientation the accumulator of one
// layer (lane l: rows 4(l/16)..+3 of a 16-neuron block, column = star l % 16) is
// exactly the B operand of the next layer's 16-wide K chunk, so activations never
// leave the registers; the weights stay in registers as A operands for all of the
// wave's tiles.  f16 keeps ~3 decimal digits: a bracket proposal only.
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));

struct MfmaNet {
    half4_t a1[4], a2[4][4], a3[4];
    float b1[4][4], b2[4][4], b3[2];
};

__device__ __forceinline__ void mfma_net_load(const float *__restrict__ theta, MfmaNet &w)
{
    const int lane = threadIdx.x & 63, row = lane & 15, quad = lane >> 4;
    const float *W1 = theta, *B1 = W1 + NH * NIN, *W2 = B1 + NH, *B2 = W2 + NH * NH, *W3 = B2 + NH,
                *B3 = W3 + NOUT * NH;
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int k = 4 * quad + v;
            w.a1[ob][v] = (_Float16)(k < NIN ? W1[(16 * ob + row) * NIN + k] : 0.f);
            w.a3[ob][v] = (_Float16)(row < NOUT ? W3[row * NH + 16 * ob + k] : 0.f);
            w.b1[ob][v] = B1[16 * ob + k];
            w.b2[ob][v] = B2[16 * ob + k];
#pragma unroll
            for (int kc = 0; kc < 4; kc++) w.a2[ob][kc][v] = (_Float16)W2[(16 * ob + row) * NH + 16 * kc + k];
        }
    }
    w.b3[0] = B3[0];
    w.b3[1] = B3[1];
}

// xb: this lane's B operand of the standardised inputs (k = 4(l/16) + v of star l % 16,
// zero beyond the 6 inputs).  Returns the standardised outputs in lanes 0..15.
__device__ __forceinline__ void mfma_net_forward(const MfmaNet &w, half4_t xb, float &y0, float &y1)
{
    half4_t h1[4], h2[4];
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a1[ob], xb, acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) h1[ob][v] = (_Float16)silu(acc[v] + w.b1[ob][v]);
    }
#pragma unroll
    for (int ob = 0; ob < 4; ob++) {
        float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < 4; kc++) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a2[ob][kc], h1[kc], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) h2[ob][v] = (_Float16)silu(acc[v] + w.b2[ob][v]);
    }
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < 4; kc++) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(w.a3[kc], h2[kc], acc, 0, 0, 0);
    y0 = acc[0] + w.b3[0];
    y1 = acc[1] + w.b3[1];
}

// The standardised inputs of one star as the lane's B operand (lanes with l / 16 = 0
// hold inputs 0..3, l / 16 = 1 inputs 4..5).
__device__ __forceinline__ half4_t mfma_inputs(const float (&x)[NIN], const MlpNorm *__restrict__ nrm, bool valid)
{
    const int quad = (threadIdx.x & 63) >> 4;
    half4_t xb = {(_Float16)0.f, (_Float16)0.f, (_Float16)0.f, (_Float16)0.f};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int k = 4 * quad + v;
        if (valid && k < NIN) xb[v] = (_Float16)((x[k] - nrm->xm[k]) / nrm->xs[k]);
    }
    return xb;
}

// Predictions for every star 1..n of the current state (the orbit phase's proposals):
// pred[j] = (log r_peri, log r_apo).  Persistent waves over 16-star tiles.
struct mlp_infer_stars_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, DevFields f, DevScalars *s, KParams P,
                                                              const float *__restrict__ theta,
                                                              const MlpNorm *__restrict__ nrm,
                                                              float2 *__restrict__ pred)
{
    MfmaNet w;
    mfma_net_load(theta, w);
    const int lane = threadIdx.x & 63, row = lane & 15, quad = lane >> 4;
    const int64_t n = s->n_max;
    const int64_t wave = ((int64_t)kg_.bid * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)kg_.nb * blockDim.x) >> 6;
    const float ym0 = nrm->ym[0], ys0 = nrm->ys[0], ym1 = nrm->ym[1], ys1 = nrm->ys[1];
    for (int64_t t = wave; 16 * t < n; t += nwaves) {
        const int64_t j = 1 + 16 * t + row;
        float x[NIN];
        const bool valid = j <= n && quad < 2;
        if (valid) {
            const double rj = f.r[j], mj = f.m[j];
            const double E = f.E[j] + phi_s_own(rj, mj, P.inv_nstar, rj);
            mlp_features(f.r, f.phi, n, j, rj, mj, E, f.J[j], x);
        }
        float y0, y1;
        mfma_net_forward(w, mfma_inputs(x, nrm, valid), y0, y1);
        if (quad == 0 && j <= n) pred[j] = make_float2(y0 * ys0 + ym0, y1 * ys1 + ym1);
    }
}
};

// The same network on given raw inputs x[6n] -> y[2n] (tests compare it with the
// per-thread FP32 path).
struct mlp_infer_rows_kernel {
    static constexpr int LB = 256, MB = 0;
    static __device__ __forceinline__ void run(KGrid kg_, const float *__restrict__ theta,
                                                             const MlpNorm *__restrict__ nrm,
                                                             const float *__restrict__ xin, int64_t n,
                                                             float *__restrict__ y)
{
    MfmaNet w;
    mfma_net_load(theta, w);
    const int lane = threadIdx.x & 63, row = lane & 15, quad = lane >> 4;
    const int64_t wave = ((int64_t)kg_.bid * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)kg_.nb * blockDim.x) >> 6;
    for (int64_t t = wave; 16 * t < n; t += nwaves) {
        const int64_t i = 16 * t + row;
        float x[NIN];
        const bool valid = i < n && quad < 2;
        if (valid)
            for (int c = 0; c < NIN; c++) x[c] = xin[NIN * i + c];
        float y0, y1;
        mfma_net_forward(w, mfma_inputs(x, nrm, valid), y0, y1);
        if (quad == 0 && i < n) {
            y[NOUT * i] = y0 * nrm->ys[0] + nrm->ym[0];
            y[NOUT * i + 1] = y1 * nrm->ys[1] + nrm->ym[1];
        }
    }
}
};
} // namespace

int mlp_grid_blocks(int64_t batch)
{
    int64_t nt = (batch + TILE - 1) / TILE;
    return (int)(nt < MLP_MAX_BLOCKS ? (nt > 0 ? nt : 1) : MLP_MAX_BLOCKS);
}

hipError_t launch_mlp_loss_grad(const MlpState &m, const float *rows, int64_t count, int64_t first, int64_t batch,
                                uint64_t mult, uint64_t off, hipStream_t st)
{
    int nb = mlp_grid_blocks(batch);
    kl<mlp_fwdbwd_kernel>(dim3(nb), dim3(FT), 0, st, (const float *)m.theta, (const MlpNorm *)m.norm,
                       rows, count, first, batch, mult, off, 1.0f / (float)batch, m.gpart, m.lpart);
    kl<mlp_reduce_kernel>(dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, (const float *)m.gpart,
                       (const double *)m.lpart, nb, m.grad, m.loss, 1.0 / (double)batch);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpState &m, int64_t step, float lr, hipStream_t st)
{
    kl<mlp_adam_kernel>(dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, m.theta,
                       (const float *)m.grad, m.m, m.v, step, lr);
    return hipGetLastError();
}

hipError_t launch_mlp_predict(const MlpState &m, const float *x, int64_t n, float *y, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    kl<mlp_predict_kernel>(dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       (const float *)m.theta, (const MlpNorm *)m.norm, x, n, y);
    return hipGetLastError();
}

static unsigned mfma_blocks(int64_t tiles)
{
    const int64_t b = (tiles + 3) / 4;                // four waves per block
    return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

hipError_t launch_mlp_infer_stars(const DevFields &f, DevScalars *s, const KParams &P, const MlpState &m,
                                  float2 *pred, int64_t n_ub, hipStream_t st)
{
    if (n_ub <= 0) return hipSuccess;
    kl<mlp_infer_stars_kernel>(dim3(mfma_blocks((n_ub + 15) / 16)), dim3(256), 0, st, f, s, P,
                       (const float *)m.theta, (const MlpNorm *)m.norm, pred);
    return hipGetLastError();
}

hipError_t launch_mlp_infer_rows(const MlpState &m, const float *x, int64_t n, float *y, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    kl<mlp_infer_rows_kernel>(dim3(mfma_blocks((n + 15) / 16)), dim3(256), 0, st,
                       (const float *)m.theta, (const MlpNorm *)m.norm, x, n, y);
    return hipGetLastError();
}

hipError_t launch_mlp_fit_norm(const MlpState &m, int64_t count, hipStream_t st)
{
    if (count <= 0) return hipErrorInvalidValue;
    int nb = 256;
    kl<mlp_colsum_kernel>(dim3(nb), dim3(256), 0, st, (const float *)m.rows, count, m.colsum);
    kl<mlp_norm_kernel>(dim3(1), dim3(64), 0, st, (const double *)m.colsum, nb, count, m.norm);
    return hipGetLastError();
}

hipError_t launch_mlp_harvest(const DevFields &f, DevScalars *s, const KParams &P, const MlpState &m, int64_t n_ub,
                              hipStream_t st)
{
    if (n_ub <= 0) return hipSuccess;
    kl<mlp_harvest_kernel>(dim3((unsigned)((n_ub + 255) / 256)), dim3(256), 0, st, f, s, P, m.rows,
                       m.cap, m.counter);
    return hipGetLastError();
}
