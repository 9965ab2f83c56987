// rl2_mlp.h -- learned simplex proposal: sizes, parameter layout and launchers (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "rl2_internal.h"

namespace rl2 {

constexpr int kMlpRad = 16, kMlpAz = 16, kMlpBins = kMlpRad * kMlpAz;
constexpr int kMlpIn = 3 * kMlpBins + 2;  // 770
constexpr int kMlpHid = 256;
constexpr int kMlpOut = 9;
constexpr int kMlpLayers = 4;
// Parameter block (FP32 master copy): W1[H][In] b1[H] W2[H][H] b2[H] W3[H][H] b3[H]
// W4[Out][H] b4[Out]; W row-major [out][in].
constexpr long kMlpParams = (long)kMlpHid * kMlpIn + kMlpHid + 2L * (kMlpHid * kMlpHid + kMlpHid) +
                            (long)kMlpOut * kMlpHid + kMlpOut;

struct MlpGemm {
    int M, N, K;
    const float *A;
    long sam, sak;
    const float *B;
    long sbk, sbn;
    float *C;
    long scm, scn;
    const float *bias;  // optional, per column n
    float *pre;         // optional: pre-activation copy (same strides as C)
    int act;            // 1: GELU
};
void launch_mlp_gemm(hipStream_t s, const MlpGemm &g);
void launch_mlp_gelu_back(hipStream_t s, long n, const float *pre, const float *dout, float *dpre);
void launch_mlp_bias_grad(hipStream_t s, int B, int N, const float *dY, float *db);
void launch_mlp_nll(hipStream_t s, int B, const float *out, const float *y, float *dout,
                    float *loss);
void launch_mlp_adam(hipStream_t s, long n, float *p, const float *g, float *m, float *v,
                     float lr, int step);
void launch_mlp_features(hipStream_t s, const FacetDev *facets, int slot, double R_arcsec,
                         double flux, float *out);

// Layer l: weights at p + w_off(l), bias at p + b_off(l), sizes n_in(l) -> n_out(l).
inline int mlp_in(int l) { return l == 0 ? kMlpIn : kMlpHid; }
inline int mlp_out(int l) { return l == kMlpLayers - 1 ? kMlpOut : kMlpHid; }
inline long mlp_w_off(int l)
{
    long o = 0;
    for (int j = 0; j < l; j++) o += (long)mlp_out(j) * mlp_in(j) + mlp_out(j);
    return o;
}
inline long mlp_b_off(int l) { return mlp_w_off(l) + (long)mlp_out(l) * mlp_in(l); }

// Device-resident network, optimiser state and activations for batches up to `cap`.
struct Mlp {
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // kMlpParams each
    int step = 0;
    int cap = 0;
    float *x = nullptr;                       // cap * In
    float *act[kMlpLayers] = {};              // outputs of each layer, cap * n_out
    float *pre[kMlpLayers] = {};              // pre-activations
    float *dact = nullptr, *dpre = nullptr;   // cap * max width
    float *y = nullptr, *loss = nullptr;      // cap * 3, cap
    float *h_loss = nullptr;                  // pinned, cap
    ~Mlp();
    bool ensure(int batch);
    bool forward(hipStream_t s, int B);       // x -> act[L-1]
    bool backward(hipStream_t s, int B);      // needs y; fills g, loss
};

}  // namespace rl2

struct rl2_proposal {
    rl2::Mlp net;
};
