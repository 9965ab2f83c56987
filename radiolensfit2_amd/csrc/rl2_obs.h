// rl2_obs.h -- the device-resident observation behind the rl2_obs handle (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <vector>

#include "rl2.h"
#include "rl2_internal.h"

namespace rl2 {

// One extraction target: dense accumulators and the compacted facet (the reference's
// sum_w / facet_visData / facet_sigma2 / facet_u / facet_v buffers,
// RadioLensfit2.cpp:425-501, sized for facet_size(RMAX)).
struct FacetSlot {
    double *acc = nullptr;   // 4 * fmax^2
    double *uu = nullptr, *vv = nullptr, *dre = nullptr, *dim = nullptr, *s2 = nullptr;
    double *sumw = nullptr;  // dense weights, allocated on first request
    float4 *pk = nullptr;    // FP32 {u, v, D_re, D_im} for the mixed-precision likelihood
    float *invs2 = nullptr;
    double *invs2_64 = nullptr;
    unsigned *scan = nullptr;
    uint32_t *cid = nullptr; // packed (row << 16 | column) per compacted cell
    double2 *dc = nullptr;   // packed {D_re, D_im} (fused Newton passes)
    uint32_t *loff = nullptr;// LDS record offsets per cell (fused Newton passes)
    double *ctr = nullptr;   // cell-centre coordinate per grid index (fmax)
    unsigned facet = 0;
    unsigned n = 0;          // host copy of the compacted size (valid after a sync)
    double rprior[kMaxR] = {};
};

// Buffers of one batched likelihood launch (items -> L_e).
struct LikBuffers {
    int cap = 0;
    LikItem *d_items = nullptr, *h_items = nullptr;
    double *d_Lr = nullptr, *h_Lr = nullptr;
    int *d_nit = nullptr, *h_nit = nullptr;
    // results block in pinned host memory, written by the kernels through its device view
    // d_out: status[2] u32 | FLOPs f64 | L_e[cap] | err[cap]
    char *d_out = nullptr, *h_out = nullptr;
    unsigned *d_status = nullptr, *h_status = nullptr;   // tasks still active, max steps
    double *d_flops = nullptr, *h_flops = nullptr;
    double *d_Le = nullptr, *h_Le = nullptr;
    int *d_err = nullptr, *h_err = nullptr;
    double *d_ws = nullptr;
    int *d_iws = nullptr;
    // tiled engine
    TileTask *d_tasks = nullptr;                         // cap * kMaxR
    double *d_partials = nullptr;
    size_t partials_cap = 0;                             // work-groups
    double *d_ppart = nullptr;                           // one-launch evaluations: chunk
    size_t ppart_cap = 0;                                // sums per pass (work-groups)
    unsigned *d_cnt = nullptr;                           // cap * kMaxR chunk counters
    double2 *d_tab = nullptr;                            // Newton phase tables
    size_t tab_cap = 0;                                  // entries
    int tile_passes = 6;     // Newton passes queued before the first status check
    bool ensure(int n);
    bool ensure_partials(size_t nwg);
    bool ensure_ppart(size_t nwg);
    bool ensure_tab(size_t entries);
    void release();
};

struct Obs {
    int prec = 0;
    unsigned nch = 0;
    unsigned long nrow = 0, nvis = 0;
    double len = 0.;
    unsigned fmax = 0;
    double *d_u = nullptr, *d_v = nullptr, *d_w = nullptr;
    uint32_t *d_perm = nullptr;  // sorted row -> original row
    void *d_res = nullptr;
    void *d_res_saved = nullptr;  // rl2_obs_checkpoint copy
    float *d_s2f = nullptr;
    double *d_wn = nullptr, *d_invlam = nullptr, *d_spec = nullptr;
    // Bounds that fix the scales of the gridding's integer accumulators: res_bound >= max
    // |re| + |im| of the residual (measured after every set-up change, raised by each
    // subtraction's largest possible change), s2_max = largest sigma^2, spec_max = largest
    // spectral factor (a model visibility is at most flux * spec_max).
    double res_bound = 0., res_bound_saved = 0., s2_max = 0., spec_max = 0.;
    bool measure_res_bound();
    std::vector<double> wn, spec, invlam;
    std::vector<FacetSlot> slots;
    FacetDev *d_facets = nullptr;
    std::vector<FacetDev> h_facets;
    LikBuffers lik;
    RGrid rg;
    int lik_block = 0;       // 0: tiled engine; 64..1024: one work-group per (item, r)
    int loglik_prec = 0;     // precision of rl2_obs_loglik requests (0 FP64, 1 mixed)
    float *d_feat = nullptr; // learned-proposal inputs of a batch (training data)
    size_t feat_cap = 0;
    // batched extraction: the batch descriptor (host staging and device copy), the work
    // queue, and the graphs of the batch's launches keyed by (targets, work-groups)
    BatchDesc *h_desc = nullptr;
    BatchDesc *d_desc = nullptr;
    unsigned *d_queue = nullptr;
    std::vector<std::pair<unsigned long long, hipGraphExec_t>> batch_graphs;
    ~Obs();
    bool ensure_slots(int n);
    bool facet_upload(int slot);  // push host descriptor (pointers, prior) to the device
};

// Batched likelihood evaluation for items over the observation's facet slots.
// Fills Le/err (and L_r, n_iter when non-null: nitems * kMaxR). Synchronous.
bool lik_eval(Obs &o, const std::vector<LikItem> &items, double *Le, int *err, double *Lr,
              int *nit);

// Stream and instrumentation shared by all translation units.
hipStream_t stream();
struct KTimer {
    int kernel;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(int k);
    ~KTimer();
};
void stats_add_work(int kernel, double flops, double bytes);

GalaxyGeom make_geom(double l0, double m0, double flux, double scale, double e1, double e2);
bool ensure_device();
// A post-fit update (galaxy position, flux, scalelength, fitted e) held back to be made in
// the same pass as the next extraction.
struct PendingSub {
    double l0, m0, flux, R_mu, e1, e2;
};
bool obs_extract(Obs &o, int slot, double l0, double m0, double flux, double R_mu,
                 unsigned facet, bool want_sumw, double e1 = 0., double e2 = 0.,
                 const PendingSub *sub = nullptr);
bool obs_subtract(Obs &o, double l0, double m0, double flux, double R_mu, double e1, double e2);
// One galaxy of a batched extraction into slot `slot`.
struct ExtractReq {
    int slot;
    double l0, m0, flux, R_mu;
    unsigned facet;
};
// Extract every request from one pass over the residual, after applying `subs` (in order)
// in the same pass (data_processing-mpi.c:65-216 batch semantics).
bool obs_extract_batch(Obs &o, const std::vector<ExtractReq> &reqs,
                       const std::vector<PendingSub> &subs);
// Apply the updates in order in one pass over the residual.
bool obs_subtract_batch(Obs &o, const std::vector<PendingSub> &subs);
bool process_list(Obs &o, const std::vector<unsigned> &todo, int pass, const double *l,
                  const double *m, const double *flux, const double *scale,
                  const rl2_fit_options &opt, rl2_gal_result *results,
                  std::vector<unsigned> *new_bad, double *t_ext, double *t_fit, double *t_sub,
                  const std::function<void(unsigned)> &on_done);

// Fit driver (rl2_fit.cpp).
struct FitRequest {
    int slot;
    rl2_gal_result *res;
    // likelihood_sampling only (likelihood.c:428-554): start from (e1, e2, maxL) in *res
    bool sampling_only = false;
    int np_max = kNpMax;
    // learned proposal: simplex start point and first step; the reference's start is
    // evaluated in the same launch and kept when its likelihood is higher
    bool has_prop = false;
    double prop_e1 = 0., prop_e2 = 0., prop_step = 0.1;
    int prop_chunk = 0;   // line-search steps per launch from the predicted width (0: default)
};
using BatchEval = std::function<bool(const std::vector<LikItem> &, std::vector<double> &,
                                     std::vector<int> &)>;
bool fit_batch(const std::vector<FitRequest> &reqs, const rl2_fit_options &opt,
               const BatchEval &eval);

}  // namespace rl2

struct rl2_obs {
    rl2::Obs o;
};
