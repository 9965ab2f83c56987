// rl2_proposal.cpp -- the learned proposal's network object and its rl2_proposal_* C ABI.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>

#include "rl2_mlp.h"
#include "rl2_obs.h"

namespace rl2 {
namespace {

template <class T>
bool dalloc(T **p, size_t count)
{
    *p = nullptr;
    return hip_ok(hipMalloc((void **)p, (count ? count : 1) * sizeof(T)), "hipMalloc");
}
template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// splitmix64 + Box-Muller for the initial weights
struct InitRng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u1 = uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
    }
};

}  // namespace

Mlp::~Mlp()
{
    (void)hipStreamSynchronize(stream());
    dfree(p);
    dfree(g);
    dfree(m);
    dfree(v);
    dfree(x);
    for (int l = 0; l < kMlpLayers; l++) {
        dfree(act[l]);
        dfree(pre[l]);
    }
    dfree(dact);
    dfree(dpre);
    dfree(y);
    dfree(loss);
    if (h_loss) (void)hipHostFree(h_loss);
}

bool Mlp::ensure(int batch)
{
    if (batch <= cap) return true;
    int c = 64;
    while (c < batch) c *= 2;
    dfree(x);
    for (int l = 0; l < kMlpLayers; l++) {
        dfree(act[l]);
        dfree(pre[l]);
    }
    dfree(dact);
    dfree(dpre);
    dfree(y);
    dfree(loss);
    if (h_loss) (void)hipHostFree(h_loss);
    h_loss = nullptr;
    bool ok = dalloc(&x, (size_t)c * kMlpIn) && dalloc(&dact, (size_t)c * kMlpHid) &&
              dalloc(&dpre, (size_t)c * kMlpHid) && dalloc(&y, (size_t)c * 3) &&
              dalloc(&loss, (size_t)c);
    for (int l = 0; l < kMlpLayers && ok; l++)
        ok = dalloc(&act[l], (size_t)c * mlp_out(l)) && dalloc(&pre[l], (size_t)c * mlp_out(l));
    ok = ok && hip_ok(hipHostMalloc((void **)&h_loss, c * sizeof(float)), "hipHostMalloc");
    if (ok) cap = c;
    return ok;
}

bool Mlp::forward(hipStream_t s, int B)
{
    for (int l = 0; l < kMlpLayers; l++) {
        const int in = mlp_in(l), out = mlp_out(l);
        MlpGemm gm;
        gm.M = B;
        gm.N = out;
        gm.K = in;
        gm.A = l == 0 ? x : act[l - 1];
        gm.sam = in;
        gm.sak = 1;
        gm.B = p + mlp_w_off(l);
        gm.sbk = 1;
        gm.sbn = in;
        gm.C = act[l];
        gm.scm = out;
        gm.scn = 1;
        gm.bias = p + mlp_b_off(l);
        gm.pre = pre[l];
        gm.act = l < kMlpLayers - 1;
        launch_mlp_gemm(s, gm);
    }
    return hip_ok(hipGetLastError(), "mlp forward");
}

bool Mlp::backward(hipStream_t s, int B)
{
    launch_mlp_nll(s, B, act[kMlpLayers - 1], y, dpre, loss);
    for (int l = kMlpLayers - 1; l >= 0; l--) {
        const int in = mlp_in(l), out = mlp_out(l);
        const float *xin = l == 0 ? x : act[l - 1];
        // dW = dpre^T x
        MlpGemm gw;
        gw.M = out;
        gw.N = in;
        gw.K = B;
        gw.A = dpre;
        gw.sam = 1;
        gw.sak = out;
        gw.B = xin;
        gw.sbk = in;
        gw.sbn = 1;
        gw.C = g + mlp_w_off(l);
        gw.scm = in;
        gw.scn = 1;
        gw.bias = nullptr;
        gw.pre = nullptr;
        gw.act = 0;
        launch_mlp_gemm(s, gw);
        launch_mlp_bias_grad(s, B, out, dpre, g + mlp_b_off(l));
        if (l == 0) break;
        // dact = dpre W ; dpre(l-1) = dact * gelu'(pre(l-1))
        MlpGemm gx;
        gx.M = B;
        gx.N = in;
        gx.K = out;
        gx.A = dpre;
        gx.sam = out;
        gx.sak = 1;
        gx.B = p + mlp_w_off(l);
        gx.sbk = in;
        gx.sbn = 1;
        gx.C = dact;
        gx.scm = in;
        gx.scn = 1;
        gx.bias = nullptr;
        gx.pre = nullptr;
        gx.act = 0;
        launch_mlp_gemm(s, gx);
        launch_mlp_gelu_back(s, (long)B * in, pre[l - 1], dact, dpre);
    }
    return hip_ok(hipGetLastError(), "mlp backward");
}

}  // namespace rl2

using namespace rl2;

extern "C" {

int rl2_proposal_num_params(void) { return (int)kMlpParams; }
int rl2_proposal_num_inputs(void) { return kMlpIn; }
int rl2_proposal_num_outputs(void) { return kMlpOut; }

int rl2_proposal_create(rl2_proposal **out, uint64_t seed)
{
    if (!out) return RL2_ERR_ARG;
    *out = nullptr;
    if (!ensure_device()) return RL2_ERR_NO_DEVICE;
    rl2_proposal *P = new rl2_proposal();
    Mlp &n = P->net;
    if (!(dalloc(&n.p, kMlpParams) && dalloc(&n.g, kMlpParams) && dalloc(&n.m, kMlpParams) &&
          dalloc(&n.v, kMlpParams))) {
        delete P;
        return RL2_ERR_HIP;
    }
    // W ~ N(0, 1/in) (GELU layers), output layer scaled by 0.1; biases 0
    std::vector<float> h(kMlpParams, 0.f);
    InitRng rng{seed ^ 0x5eed5eedull};
    for (int l = 0; l < kMlpLayers; l++) {
        const double sd = (l == kMlpLayers - 1 ? 0.1 : 1.0) / sqrt((double)mlp_in(l));
        float *w = h.data() + mlp_w_off(l);
        for (long i = 0; i < (long)mlp_out(l) * mlp_in(l); i++) w[i] = (float)(sd * rng.normal());
    }
    std::vector<float> zero(kMlpParams, 0.f);
    if (!(hip_ok(hipMemcpy(n.p, h.data(), kMlpParams * 4, hipMemcpyHostToDevice), "copy") &&
          hip_ok(hipMemcpy(n.m, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice), "copy") &&
          hip_ok(hipMemcpy(n.v, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice), "copy") &&
          hip_ok(hipMemcpy(n.g, zero.data(), kMlpParams * 4, hipMemcpyHostToDevice), "copy"))) {
        delete P;
        return RL2_ERR_HIP;
    }
    *out = P;
    return RL2_OK;
}

int rl2_proposal_destroy(rl2_proposal *p)
{
    delete p;
    return RL2_OK;
}

int rl2_proposal_get_params(rl2_proposal *p, float *host)
{
    if (!p || !host) return RL2_ERR_ARG;
    return hip_ok(hipMemcpy(host, p->net.p, kMlpParams * 4, hipMemcpyDeviceToHost), "copy")
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_proposal_set_params(rl2_proposal *p, const float *host)
{
    if (!p || !host) return RL2_ERR_ARG;
    return hip_ok(hipMemcpy(p->net.p, host, kMlpParams * 4, hipMemcpyHostToDevice), "copy")
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_proposal_get_grads(rl2_proposal *p, float *host)
{
    if (!p || !host) return RL2_ERR_ARG;
    return hip_ok(hipMemcpy(host, p->net.g, kMlpParams * 4, hipMemcpyDeviceToHost), "copy")
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_proposal_forward(rl2_proposal *p, int batch, const float *x, float *out)
{
    if (!p || batch <= 0 || !x || !out) return RL2_ERR_ARG;
    Mlp &n = p->net;
    if (!n.ensure(batch)) return RL2_ERR_HIP;
    hipStream_t s = stream();
    if (!hip_ok(hipMemcpyAsync(n.x, x, (size_t)batch * kMlpIn * 4, hipMemcpyHostToDevice, s),
                "copy") ||
        !n.forward(s, batch) ||
        !hip_ok(hipMemcpyAsync(out, n.act[kMlpLayers - 1], (size_t)batch * kMlpOut * 4,
                               hipMemcpyDeviceToHost, s),
                "copy") ||
        !hip_ok(hipStreamSynchronize(s), "sync"))
        return RL2_ERR_HIP;
    return RL2_OK;
}

// Forward + backward on one batch: gradients (mean over the batch) left in the network's
// gradient buffer; *mean_loss = mean negative log-likelihood.
static int grads(rl2_proposal *p, int batch, const float *x, const float *y, float *mean_loss)
{
    Mlp &n = p->net;
    if (!n.ensure(batch)) return RL2_ERR_HIP;
    hipStream_t s = stream();
    if (!hip_ok(hipMemcpyAsync(n.x, x, (size_t)batch * kMlpIn * 4, hipMemcpyHostToDevice, s),
                "copy") ||
        !hip_ok(hipMemcpyAsync(n.y, y, (size_t)batch * 3 * 4, hipMemcpyHostToDevice, s), "copy") ||
        !n.forward(s, batch) || !n.backward(s, batch) ||
        !hip_ok(hipMemcpyAsync(n.h_loss, n.loss, batch * 4, hipMemcpyDeviceToHost, s), "copy") ||
        !hip_ok(hipStreamSynchronize(s), "sync"))
        return RL2_ERR_HIP;
    if (mean_loss) {
        double t = 0.;
        for (int b = 0; b < batch; b++) t += n.h_loss[b];
        *mean_loss = (float)(t / batch);
    }
    return RL2_OK;
}

int rl2_proposal_gradients(rl2_proposal *p, int batch, const float *x, const float *y,
                           float *mean_loss)
{
    if (!p || batch <= 0 || !x || !y) return RL2_ERR_ARG;
    return grads(p, batch, x, y, mean_loss);
}

int rl2_proposal_train_step(rl2_proposal *p, int batch, const float *x, const float *y,
                            float lr, float *mean_loss)
{
    if (!p || batch <= 0 || !x || !y || !(lr > 0)) return RL2_ERR_ARG;
    int rc = grads(p, batch, x, y, mean_loss);
    if (rc != RL2_OK) return rc;
    Mlp &n = p->net;
    n.step++;
    launch_mlp_adam(stream(), kMlpParams, n.p, n.g, n.m, n.v, lr, n.step);
    return hip_ok(hipGetLastError(), "adam") && hip_ok(hipStreamSynchronize(stream()), "sync")
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_obs_features(rl2_obs *obs, int slot, double R_mu, double flux, float *out)
{
    if (!obs || slot < 0 || !out || !(flux > 0) || !(R_mu > 0)) return RL2_ERR_ARG;
    Obs &o = obs->o;
    if (slot >= (int)o.slots.size()) {
        set_error("rl2_obs_features: slot has no extracted facet");
        return RL2_ERR_STATE;
    }
    float *d = nullptr;
    if (!dalloc(&d, kMlpIn)) return RL2_ERR_HIP;
    launch_mlp_features(stream(), o.d_facets, slot, R_mu, flux, d);
    const bool ok = hip_ok(hipGetLastError(), "features") &&
                    hip_ok(hipMemcpyAsync(out, d, kMlpIn * 4, hipMemcpyDeviceToHost, stream()),
                           "copy") &&
                    hip_ok(hipStreamSynchronize(stream()), "sync");
    dfree(d);
    return ok ? RL2_OK : RL2_ERR_HIP;
}

}  // extern "C"
