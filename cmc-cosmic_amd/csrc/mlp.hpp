// Learned apsides-bracket network: sizes, parameter state, input features and the
// per-thread forward pass (shared by the training, prediction and orbit kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

constexpr int MLP_NIN = 6, MLP_NH = 64, MLP_NOUT = 2;
constexpr int MLP_NPARAM = MLP_NH * MLP_NIN + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NOUT * MLP_NH + MLP_NOUT;
constexpr int MLP_MAX_BLOCKS = 256;

// per-column mean and standard deviation of the inputs and targets
struct MlpNorm { float xm[MLP_NIN], xs[MLP_NIN], ym[MLP_NOUT], ys[MLP_NOUT]; };

struct MlpState {
    float *theta, *grad, *m, *v;     // parameters, gradient, Adam moments
    MlpNorm *norm;
    float *gpart;                    // MLP_MAX_BLOCKS x MLP_NPARAM
    double *lpart, *loss;            // MLP_MAX_BLOCKS, 1
    float *rows;                     // harvested rows: 6 inputs + 2 targets
    int64_t cap;
    unsigned long long *counter;     // rows harvested so far (ring position)
    double *colsum;                  // 256 x 16
};

// Inputs of star j: E, J, log r, phi(r), log(r^2 dphi/dr) from the neighbouring
// nodes, m.
__device__ __forceinline__ void mlp_features(const double *__restrict__ r, const double *__restrict__ phi,
                                             int64_t n, int64_t j, double rj, double mj, double E, double J,
                                             float (&x)[MLP_NIN])
{
    double dr = r[j + 1] - r[j - 1];
    double slope = dr > 0.0 ? (phi[j + 1] - phi[j - 1]) / dr : 0.0;
    double g = slope * rj * rj;
    x[0] = (float)E;
    x[1] = (float)J;
    x[2] = (float)log(rj);
    x[3] = (float)phi[j];
    x[4] = (float)log(g > 1e-30 ? g : 1e-30);
    x[5] = (float)mj;
    (void)n;
}

// standardise, two SiLU layers, de-standardise: (log r_peri, log r_apo)
__device__ __forceinline__ void mlp_forward(const float *__restrict__ theta, const MlpNorm *__restrict__ nrm,
                                            const float (&in)[MLP_NIN], float (&out)[MLP_NOUT])
{
    const float *W1 = theta, *b1 = W1 + MLP_NH * MLP_NIN, *W2 = b1 + MLP_NH, *b2 = W2 + MLP_NH * MLP_NH,
                *W3 = b2 + MLP_NH, *b3 = W3 + MLP_NOUT * MLP_NH;
    float xs[MLP_NIN], h1[MLP_NH];
    for (int c = 0; c < MLP_NIN; c++) xs[c] = (in[c] - nrm->xm[c]) / nrm->xs[c];
    for (int u = 0; u < MLP_NH; u++) {
        float z = b1[u];
        for (int c = 0; c < MLP_NIN; c++) z += W1[MLP_NIN * u + c] * xs[c];
        h1[u] = z / (1.0f + __expf(-z));
    }
    float y0 = b3[0], y1 = b3[1];
    for (int u = 0; u < MLP_NH; u++) {
        float z = b2[u];
        for (int v = 0; v < MLP_NH; v++) z += W2[MLP_NH * u + v] * h1[v];
        float h = z / (1.0f + __expf(-z));
        y0 += W3[u] * h;
        y1 += W3[MLP_NH + u] * h;
    }
    out[0] = y0 * nrm->ys[0] + nrm->ym[0];
    out[1] = y1 * nrm->ys[1] + nrm->ym[1];
}
