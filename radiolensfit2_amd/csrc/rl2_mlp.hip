// rl2_mlp.hip -- the learned simplex proposal: facet features, a 770-256-256-256-9 MLP
// with BF16 MFMA matrix products (FP32 accumulation), its backward pass and Adam.
//
// The network only chooses where the simplex starts and how large its first step is
// (galaxy_fitting.c:59-67); the exact likelihood still decides every move, and the fit
// falls back to the reference's start whenever the proposal's likelihood is lower.
#include <hip/hip_runtime.h>
#include <math.h>

#include "rl2_mlp.h"

namespace rl2 {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs).
__device__ inline short to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n]) with A(m, k) = A[m*sam + k*sak] and
// B(k, n) = B[k*sbk + n*sbn]; operands rounded to BF16, products accumulated in FP32 by
// v_mfma_f32_16x16x16_bf16.  One wave per 16x16 tile of C, four waves per work-group.
// MFMA 16x16x16 operand layout: lane l supplies A row l%16, k = 4(l/16)..4(l/16)+3 and B
// column l%16, the same k; it receives C rows 4(l/16)..+3 of column l%16.
__global__ __launch_bounds__(256) void mlp_gemm_kernel(MlpGemm g)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ntn = (g.N + 15) / 16, ntm = (g.M + 15) / 16;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= ntm * ntn) return;
    const int tm = tile / ntn, tn = tile % ntn;
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
    const int arow = tm * 16 + r16, bcol = tn * 16 + r16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int k = k0 + kq + i;
            const float av = (arow < g.M && k < g.K) ? g.A[(long)arow * g.sam + (long)k * g.sak] : 0.f;
            const float bv = (bcol < g.N && k < g.K) ? g.B[(long)k * g.sbk + (long)bcol * g.sbn] : 0.f;
            a[i] = to_bf16(av);
            b[i] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int r = tm * 16 + kq + i, c = bcol;
        if (r >= g.M || c >= g.N) continue;
        float x = acc[i] + (g.bias ? g.bias[c] : 0.f);
        if (g.pre) g.pre[(long)r * g.scm + (long)c * g.scn] = x;
        if (g.act) x = gelu(x);
        g.C[(long)r * g.scm + (long)c * g.scn] = x;
    }
}

void launch_mlp_gemm(hipStream_t s, const MlpGemm &g)
{
    const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    mlp_gemm_kernel<<<(tiles + 3) / 4, 256, 0, s>>>(g);
}

// d_pre = d_out * gelu'(pre)
__global__ void mlp_gelu_back_kernel(long n, const float *pre, const float *dout, float *dpre)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dpre[i] = dout[i] * gelu_grad(pre[i]);
}
void launch_mlp_gelu_back(hipStream_t s, long n, const float *pre, const float *dout, float *dpre)
{
    mlp_gelu_back_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, pre, dout, dpre);
}

// db[j] = sum_b dY[b][j]
__global__ void mlp_bias_grad_kernel(int B, int N, const float *dY, float *db)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += dY[(long)b * N + j];
    db[j] = s;
}
void launch_mlp_bias_grad(hipStream_t s, int B, int N, const float *dY, float *db)
{
    mlp_bias_grad_kernel<<<(N + 255) / 256, 256, 0, s>>>(B, N, dY, db);
}

// Gaussian negative log-likelihood of the target y = (e1, e2, log r) under the output
// o: mean o[0..2], precision L L^T with L lower triangular, L_ii = exp(o[3+i]),
// L_10 = o[6], L_20 = o[7], L_21 = o[8]:
//   nll = 1/2 |L^T (y - mu)|^2 - sum_i o[3+i];  dout = d nll / d o, divided by the batch.
__global__ void mlp_nll_kernel(int B, const float *out, const float *y, float *dout,
                               float *loss)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float *o = out + (long)b * kMlpOut;
    const float d0 = y[b * 3 + 0] - o[0], d1 = y[b * 3 + 1] - o[1], d2 = y[b * 3 + 2] - o[2];
    const float l00 = expf(o[3]), l11 = expf(o[4]), l22 = expf(o[5]);
    const float l10 = o[6], l20 = o[7], l21 = o[8];
    // z = L^T d
    const float z0 = l00 * d0 + l10 * d1 + l20 * d2;
    const float z1 = l11 * d1 + l21 * d2;
    const float z2 = l22 * d2;
    loss[b] = 0.5f * (z0 * z0 + z1 * z1 + z2 * z2) - (o[3] + o[4] + o[5]);
    const float inv = 1.f / B;
    float *g = dout + (long)b * kMlpOut;
    // d/dmu = -L z
    g[0] = -(l00 * z0) * inv;
    g[1] = -(l10 * z0 + l11 * z1) * inv;
    g[2] = -(l20 * z0 + l21 * z1 + l22 * z2) * inv;
    g[3] = (d0 * z0 * l00 - 1.f) * inv;
    g[4] = (d1 * z1 * l11 - 1.f) * inv;
    g[5] = (d2 * z2 * l22 - 1.f) * inv;
    g[6] = d1 * z0 * inv;
    g[7] = d2 * z0 * inv;
    g[8] = d2 * z1 * inv;
}
void launch_mlp_nll(hipStream_t s, int B, const float *out, const float *y, float *dout,
                    float *loss)
{
    mlp_nll_kernel<<<(B + 255) / 256, 256, 0, s>>>(B, out, y, dout, loss);
}

// Adam on the FP32 master parameters.
__global__ void mlp_adam_kernel(long n, float *p, const float *g, float *m, float *v, float lr,
                                float b1, float b2, float eps, float c1, float c2)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}
void launch_mlp_adam(hipStream_t s, long n, float *p, const float *g, float *m, float *v,
                     float lr, int step)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    mlp_adam_kernel<<<1024, 256, 0, s>>>(n, p, g, m, v, lr, b1, b2, 1e-8f, c1, c2);
}

// Facet features: cells binned by q = 2 pi |(u, v)| R (R the catalog scalelength in
// radians; 16 linear bins over q in [0, 4), the last open) and by position angle folded to
// [0, pi) (the conjugate point carries the same information, Im changes sign); per bin the
// asinh of the mean of D sigma^2 / flux (re, im) and log(1 + count / mean count); then
// log flux and log R in arcsec.  One work-group per facet slot.
__global__ __launch_bounds__(256) void mlp_features_kernel(const FacetDev *facets, int slot,
                                                          double q_scale, double flux,
                                                          double R_arcsec, float *out)
{
    __shared__ float sre[kMlpBins], sim[kMlpBins], sw[kMlpBins];
    for (int b = threadIdx.x; b < kMlpBins; b += blockDim.x) sre[b] = sim[b] = sw[b] = 0.f;
    __syncthreads();
    const FacetDev &F = facets[slot];
    const unsigned n = F.n;
    const double inv_flux = 1.0 / flux;
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
        const double u = F.uu[i], v = F.vv[i];
        const double q = q_scale * sqrt(u * u + v * v);
        double phi = atan2(v, u);
        double re = F.dre[i] * F.s2[i] * inv_flux, im = F.dim[i] * F.s2[i] * inv_flux;
        if (phi < 0) {
            phi += kPi;
            im = -im;
        }
        const int rb = min(kMlpRad - 1, (int)(q * (kMlpRad / 4.0)));
        const int ab = min(kMlpAz - 1, (int)(phi * (kMlpAz / kPi)));
        const int b = rb * kMlpAz + ab;
        atomicAdd(&sre[b], (float)re);
        atomicAdd(&sim[b], (float)im);
        atomicAdd(&sw[b], 1.f);
    }
    __syncthreads();
    // bin means are noise-dominated where few cells fall: asinh keeps them O(1..10)
    const float mean_count = n > 0 ? (float)n / kMlpBins : 1.f;
    for (int b = threadIdx.x; b < kMlpBins; b += blockDim.x) {
        const float w = sw[b];
        out[3 * b + 0] = w > 0 ? asinhf(sre[b] / w) : 0.f;
        out[3 * b + 1] = w > 0 ? asinhf(sim[b] / w) : 0.f;
        out[3 * b + 2] = log1pf(w / mean_count);
    }
    if (threadIdx.x == 0) {
        out[3 * kMlpBins] = (float)log(flux);
        out[3 * kMlpBins + 1] = (float)log(R_arcsec);
    }
}
void launch_mlp_features(hipStream_t s, const FacetDev *facets, int slot, double R_arcsec,
                         double flux, float *out)
{
    const double q_scale = kTwoPi * arcs2rad_scale(R_arcsec);
    mlp_features_kernel<<<1, 256, 0, s>>>(facets, slot, q_scale, flux, R_arcsec, out);
}

}  // namespace rl2
