// rl2_runtime.hip -- observation lifecycle, batched likelihood launches, the per-galaxy
// driver and the rl2_* C-ABI (include/rl2.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <math.h>
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

#include "rl2_mlp.h"
#include "rl2_obs.h"

namespace rl2 {

// ------------------------------------------------------------------ global state
namespace {
thread_local std::string g_err;
int g_device = -1;
hipStream_t g_stream = nullptr;
bool g_stats_on = false;
struct KStat {
    double ms = 0.;
    long long launches = 0;
    double flops = 0., bytes = 0.;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
KStat g_stats[RL2_K_COUNT];
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) e = nullptr;  // the timer then records nothing
    return e;
}
void drain_stats()
{
    for (auto &k : g_stats) {
        for (auto &p : k.pending) {
            float ms = 0.f;
            if (hipEventSynchronize(p.second) == hipSuccess &&
                hipEventElapsedTime(&ms, p.first, p.second) == hipSuccess)
                k.ms += ms;
            g_event_pool.push_back(p.first);
            g_event_pool.push_back(p.second);
        }
        k.pending.clear();
    }
}
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}
hipStream_t stream() { return g_stream; }
Tuning &tuning()
{
    static Tuning t;
    return t;
}

KTimer::KTimer(int k) : kernel(k)
{
    if (g_stats_on) {
        a = take_event();
        b = take_event();
        if (!a || !b || hipEventRecord(a, g_stream) != hipSuccess) {
            if (a) g_event_pool.push_back(a);
            if (b) g_event_pool.push_back(b);
            a = b = nullptr;
        }
    }
}
KTimer::~KTimer()
{
    if (g_stats_on && a && b) {
        if (hipEventRecord(b, g_stream) != hipSuccess) {
            g_event_pool.push_back(a);
            g_event_pool.push_back(b);
            return;
        }
        g_stats[kernel].pending.push_back({a, b});
        g_stats[kernel].launches++;
    }
}
void stats_add_work(int kernel, double flops, double bytes)
{
    if (!g_stats_on) return;
    g_stats[kernel].flops += flops;
    g_stats[kernel].bytes += bytes;
}

// Scratch of the allocation-free entry point rl2_marginalise_posterior_r_noalloc: the
// QUADPACK interval lists of one scalelength integration (kQagDoubles doubles, then
// kQagInts ints).  Each device holds its own copy until the process ends; host code uses
// only the address looked up once per device at set-up (g_scratch), never the array's name.
constexpr size_t kScratchBytes = kQagDoubles * sizeof(double) + kQagInts * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices] = {};

// With device `dev` current: its copy of the scratch array.
static bool init_scratch(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index beyond the scratch table");
        return false;
    }
    if (g_scratch[dev]) return true;
    void *p = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch"))
        return false;
    g_scratch[dev] = p;
    return true;
}

static bool ensure_init()
{
    if (g_stream) return true;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error("no HIP device available (librl2hip has no CPU path)");
        return false;
    }
    if (g_device < 0) g_device = 0;
    if (!hip_ok(hipSetDevice(g_device), "hipSetDevice")) return false;
    return init_scratch(g_device) && hip_ok(hipStreamCreate(&g_stream), "hipStreamCreate");
}

template <class T>
static bool dmalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    return hip_ok(hipMalloc((void **)p, count * sizeof(T)), "hipMalloc");
}
template <class T>
static void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// ------------------------------------------------------------------ likelihood buffers
bool LikBuffers::ensure(int n)
{
    if (n <= cap) return true;
    release();
    int c = 256;
    while (c < n) c *= 2;
    const size_t ob = 16 + 12 * (size_t)c;  // status[2] u32, FLOPs f64, L_e[c] f64, err[c] i32
    bool ok = dmalloc(&d_items, c) && dmalloc(&d_Lr, (size_t)c * kMaxR) &&
              dmalloc(&d_nit, (size_t)c * kMaxR) &&
              dmalloc(&d_ws, (size_t)c * kQagDoubles) && dmalloc(&d_iws, (size_t)c * kQagInts) &&
              dmalloc(&d_tasks, (size_t)c * kMaxR * 2) &&  // two generations of task state
              dmalloc(&d_cnt, (size_t)c * kMaxR);
    ok = ok && hip_ok(hipMemsetAsync(d_tasks, 0, (size_t)c * kMaxR * 2 * sizeof(TileTask),
                                     g_stream),
                      "memset");
    // the results block lives in pinned host memory that the kernels write directly: the
    // host reads it after the evaluation's one synchronisation, with no copy launch
    ok = ok && hip_ok(hipHostMalloc((void **)&h_out, ob), "hipHostMalloc") &&
         hip_ok(hipHostGetDevicePointer((void **)&d_out, h_out, 0), "hipHostGetDevicePointer") &&
         hip_ok(hipHostMalloc((void **)&h_items, c * sizeof(LikItem)), "hipHostMalloc") &&
         hip_ok(hipHostMalloc((void **)&h_Lr, (size_t)c * kMaxR * sizeof(double)), "hipHostMalloc") &&
         hip_ok(hipHostMalloc((void **)&h_nit, (size_t)c * kMaxR * sizeof(int)), "hipHostMalloc");
    if (!ok) return false;
    cap = c;
    d_status = (unsigned *)d_out;
    d_flops = (double *)(d_out + 8);
    d_Le = (double *)(d_out + 16);
    d_err = (int *)(d_out + 16 + 8 * (size_t)c);
    h_status = (unsigned *)h_out;
    h_flops = (double *)(h_out + 8);
    h_Le = (double *)(h_out + 16);
    h_err = (int *)(h_out + 16 + 8 * (size_t)c);
    return true;
}
void LikBuffers::release()
{
    dfree(d_items);
    dfree(d_Lr);
    dfree(d_nit);
    d_out = nullptr;  // device view of h_out
    dfree(d_ws);
    dfree(d_iws);
    dfree(d_tasks);
    dfree(d_cnt);
    if (h_out) (void)hipHostFree(h_out);
    if (h_items) (void)hipHostFree(h_items);
    if (h_Lr) (void)hipHostFree(h_Lr);
    if (h_nit) (void)hipHostFree(h_nit);
    h_out = nullptr;
    h_items = nullptr;
    h_Lr = nullptr;
    h_nit = nullptr;
    d_status = h_status = nullptr;
    d_flops = h_flops = nullptr;
    d_Le = h_Le = nullptr;
    d_err = h_err = nullptr;
    cap = 0;
}
bool LikBuffers::ensure_partials(size_t nwg)
{
    if (nwg <= partials_cap) return true;
    dfree(d_partials);
    size_t c = 4096;
    while (c < nwg) c *= 2;
    if (!dmalloc(&d_partials, c * 8)) {
        partials_cap = 0;
        return false;
    }
    partials_cap = c;
    return true;
}

bool LikBuffers::ensure_ppart(size_t nwg)
{
    if (nwg <= ppart_cap) return true;
    dfree(d_ppart);
    size_t c = 4096;
    while (c < nwg) c *= 2;
    if (!dmalloc(&d_ppart, c * 8 * kPersistPasses)) {
        ppart_cap = 0;
        return false;
    }
    ppart_cap = c;
    return true;
}

bool LikBuffers::ensure_tab(size_t entries)
{
    if (entries <= tab_cap) return true;
    dfree(d_tab);
    size_t c = 1 << 16;
    while (c < entries) c *= 2;
    if (!dmalloc(&d_tab, c)) {
        tab_cap = 0;
        return false;
    }
    tab_cap = c;
    return true;
}

// ------------------------------------------------------------------ observation
Obs::~Obs()
{
    (void)hipStreamSynchronize(g_stream);  // no kernel may still use what is freed below
    dfree(d_u);
    dfree(d_v);
    dfree(d_w);
    dfree(d_perm);
    dfree(d_res);
    dfree(d_res_saved);
    dfree(d_s2f);
    dfree(d_wn);
    dfree(d_invlam);
    dfree(d_spec);
    for (auto &s : slots) {
        dfree(s.acc);
        dfree(s.uu);
        dfree(s.vv);
        dfree(s.dre);
        dfree(s.dim);
        dfree(s.s2);
        dfree(s.sumw);
        dfree(s.pk);
        dfree(s.invs2);
        dfree(s.invs2_64);
        dfree(s.scan);
        dfree(s.cid);
        dfree(s.ctr);
        dfree(s.dc);
        dfree(s.loff);
    }
    dfree(d_facets);
    dfree(d_feat);
    for (auto &g : batch_graphs) (void)hipGraphExecDestroy(g.second);
    batch_graphs.clear();
    if (h_desc) (void)hipHostFree(h_desc);
    h_desc = nullptr;
    dfree(d_desc);
    dfree(d_queue);
    lik.release();
    dfree(lik.d_partials);
    lik.partials_cap = 0;
    dfree(lik.d_ppart);
    lik.ppart_cap = 0;
    dfree(lik.d_tab);
    lik.tab_cap = 0;
}

bool Obs::ensure_slots(int n)
{
    if ((int)slots.size() >= n) return true;
    size_t ncell = (size_t)fmax * fmax;
    int old = (int)slots.size();
    slots.resize(n);
    for (int i = old; i < n; i++) {
        FacetSlot &s = slots[i];
        // the cell arrays of the grid paths hold a zero pad cell after the last cell
        if (!(dmalloc(&s.acc, grid_acc_words(fmax)) && dmalloc(&s.uu, ncell) &&
              dmalloc(&s.vv, ncell) &&
              dmalloc(&s.dre, ncell + 1) && dmalloc(&s.dim, ncell + 1) &&
              dmalloc(&s.s2, ncell) && dmalloc(&s.pk, ncell) && dmalloc(&s.invs2, ncell) &&
              dmalloc(&s.invs2_64, ncell + 1) && dmalloc(&s.cid, ncell + 1) &&
              dmalloc(&s.dc, ncell + 1) && dmalloc(&s.loff, ncell + 1) &&
              dmalloc(&s.ctr, (size_t)fmax) &&
              dmalloc(&s.scan, compact_scratch_words(fmax) + 1)))
            return false;
        if (!hip_ok(hipMemsetAsync(s.acc, 0, grid_acc_words(fmax) * sizeof(double), g_stream),
                    "memset"))
            return false;
    }
    FacetDev *nd = nullptr;
    if (!dmalloc(&nd, n)) return false;
    if (d_facets) {
        const bool copied = hip_ok(hipMemcpyAsync(nd, d_facets, old * sizeof(FacetDev),
                                                  hipMemcpyDeviceToDevice, g_stream),
                                   "facet descriptors") &&
                            hip_ok(hipStreamSynchronize(g_stream), "facet descriptors");
        if (!copied) {
            dfree(nd);
            return false;
        }
        dfree(d_facets);
    }
    d_facets = nd;
    h_facets.resize(n);
    for (int i = old; i < n; i++) {
        FacetDev &f = h_facets[i];
        memset(&f, 0, sizeof f);
        f.uu = slots[i].uu;
        f.vv = slots[i].vv;
        f.dre = slots[i].dre;
        f.dim = slots[i].dim;
        f.s2 = slots[i].s2;
        f.mod = nullptr;
        f.pk = slots[i].pk;
        f.invs2 = slots[i].invs2;
        f.invs2_64 = slots[i].invs2_64;
        f.dc = slots[i].dc;
        f.loff = slots[i].loff;
        f.n = 0;
        // full descriptor once (n = 0 until a compaction or load writes it)
        if (!hip_ok(hipMemcpyAsync(d_facets + i, &f, sizeof(FacetDev), hipMemcpyHostToDevice,
                                   g_stream),
                    "facet descriptor"))
            return false;
    }
    return hip_ok(hipStreamSynchronize(g_stream), "ensure_slots");
}

// Pushes the slot descriptor except the device-written count n (updated by compaction).
bool Obs::facet_upload(int i)
{
    FacetDev &f = h_facets[i];
    for (int k = 0; k < kMaxR; k++) f.rprior[k] = slots[i].rprior[k];
    const size_t off = offsetof(FacetDev, rprior);
    return hip_ok(hipMemcpyAsync((char *)(d_facets + i) + off, (char *)&f + off,
                                 sizeof(FacetDev) - off, hipMemcpyHostToDevice, g_stream),
                  "facet descriptor") &&
           hip_ok(hipMemcpyAsync(d_facets + i, &f, offsetof(FacetDev, n), hipMemcpyHostToDevice,
                                 g_stream),
                  "facet descriptor");
}

bool Obs::measure_res_bound()
{
    unsigned long long *d = nullptr, h = 0;
    if (!dmalloc(&d, 1)) return false;
    bool ok = hip_ok(hipMemsetAsync(d, 0, sizeof *d, g_stream), "memset") ;
    if (ok) {
        launch_abs_bound(g_stream, prec, nvis, d_res, d);
        ok = hip_ok(hipGetLastError(), "abs_bound") &&
             hip_ok(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, g_stream), "copy") &&
             hip_ok(hipStreamSynchronize(g_stream), "sync");
    }
    dfree(d);
    double b;
    memcpy(&b, &h, sizeof b);
    res_bound = b;
    return ok;
}

GalaxyGeom make_geom(double l0, double m0, double flux, double scale, double e1, double e2)
{
    GalaxyGeom g;
    double detA = 1. - e1 * e1 - e2 * e2;
    double sc = arcs2rad_scale(scale);
    g.sf = (sc * sc) / (detA * detA);
    g.l0 = l0;
    g.m0 = m0;
    g.n0 = sqrt(1. - l0 * l0 - m0 * m0) - 1.;
    g.flux = flux;
    g.e1 = e1;
    g.e2 = e2;
    return g;
}

static double bytes_per_vis(const Obs &o, int kernel)
{
    const double res = o.prec == 0 ? 16. : 8.;
    if (kernel == RL2_K_EXTRACT) return res + 4.;  // residual read + sigma^2/flag
    return 2. * res;                                // read-modify-write
}

bool ensure_device() { return ensure_init(); }

// Second stream (default priority) and the events that order it against the library stream.
static hipStream_t g_stream2 = nullptr;
static hipEvent_t g_fork = nullptr, g_join = nullptr;
static bool ensure_stream2()
{
    if (g_stream2) return true;
    return hip_ok(hipStreamCreate(&g_stream2), "hipStreamCreate") &&
           hip_ok(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming), "hipEventCreate") &&
           hip_ok(hipEventCreateWithFlags(&g_join, hipEventDisableTiming), "hipEventCreate");
}

// Compaction of a slot's accumulators into its facet, the FP32 copy and the descriptor.
static bool finish_extract(Obs &o, int slot, unsigned facet, bool want_sumw, const AccScale &sc)
{
    FacetSlot &s = o.slots[slot];
    if (want_sumw && !s.sumw && !dmalloc(&s.sumw, (size_t)o.fmax * o.fmax)) return false;
    {
        KTimer t(RL2_K_COMPACT);
        launch_compact(g_stream, s.acc, facet, o.len, s.uu, s.vv, s.dre, s.dim, s.s2, s.cid,
                       s.ctr, &o.d_facets[slot].n, want_sumw ? s.sumw : nullptr, s.scan, &sc);
        launch_pack_facet(g_stream, &o.d_facets[slot], facet * facet, s.uu, s.vv, s.dre, s.dim,
                          s.s2, s.pk, s.invs2, s.invs2_64, s.cid, facet, s.dc, s.loff);
    }
    // the facet now has grid structure: cell ids and centres for the phase tables (packed
    // in 16-bit row / column fields)
    FacetDev &f = o.h_facets[slot];
    const unsigned gf = facet < 65536u ? facet : 0u;
    if (f.cid != s.cid || f.grid_facet != gf) {
        f.cid = gf ? s.cid : nullptr;
        f.ctr = gf ? s.ctr : nullptr;
        f.grid_facet = gf;
        if (!o.facet_upload(slot)) return false;
    }
    return true;
}

bool obs_extract(Obs &o, int slot, double l0, double m0, double flux, double R_mu,
                 unsigned facet, bool want_sumw, double e1, double e2, const PendingSub *sub)
{
    if (facet == 0 || facet > o.fmax) {
        set_error("facet size " + std::to_string(facet) + " outside the allocated facet (max " +
                  std::to_string(o.fmax) + ")");
        return false;
    }
    if (!o.ensure_slots(slot + 1)) return false;
    FacetSlot &s = o.slots[slot];
    s.facet = facet;
    GridLaunch a;
    a.mode = 0;
    a.prec = o.prec;
    a.nch = o.nch;
    a.nrow = o.nrow;
    a.u = o.d_u;
    a.v = o.d_v;
    a.w = o.d_w;
    a.wn = o.d_wn;
    a.inv_lambda = o.d_invlam;
    a.spec = o.d_spec;
    a.res = o.d_res;
    a.s2f = o.d_s2f;
    a.g = make_geom(l0, m0, flux, R_mu, e1, e2);
    a.len = o.len;
    a.inc = 2 * o.len / facet;
    a.facet = facet;
    a.acc = s.acc;
    // a visibility that is gridded is at most |round model| + |residual| per component
    if (sub) o.res_bound += 3. * fabs(sub->flux) * o.spec_max;
    const AccScale sc = fx_scales(fabs(flux) * o.spec_max + o.res_bound, o.s2_max, -1, o.nch);
    a.fx_scale = sc.fx_scale;
    a.fs_scale = sc.fs_scale;
    if (sub) {  // the previous galaxy's post-fit update in the same pass
        a.sub = 1;
        a.sfit = make_geom(sub->l0, sub->m0, sub->flux, sub->R_mu, sub->e1, sub->e2);
        a.srnd = make_geom(sub->l0, sub->m0, sub->flux, sub->R_mu, 0., 0.);
        a.res_w = o.d_res;
    }
    {
        KTimer t(RL2_K_EXTRACT);
        launch_grid(g_stream, a);
    }
    // read residual, sigma^2 and uvw; a fused pass also writes the residual back
    stats_add_work(RL2_K_EXTRACT, 0., o.nvis * (bytes_per_vis(o, RL2_K_EXTRACT) +
                                                (sub ? (o.prec == 0 ? 16. : 8.) : 0.)) +
                                          o.nrow * 24.);
    return finish_extract(o, slot, facet, want_sumw, sc) &&
           hip_ok(hipGetLastError(), "extract launch");
}

static unsigned g_num_cus = 0;
static unsigned num_cus()
{
    if (!g_num_cus) {
        hipDeviceProp_t p;
        g_num_cus = hipGetDeviceProperties(&p, g_device) == hipSuccess && p.multiProcessorCount > 0
                        ? (unsigned)p.multiProcessorCount
                        : 256u;
    }
    return g_num_cus;
}

static bool ensure_batch_buffers(Obs &o)
{
    if (o.d_desc) return true;
    return hip_ok(hipHostMalloc((void **)&o.h_desc, sizeof(BatchDesc)), "hipHostMalloc") &&
           dmalloc(&o.d_desc, 1) && dmalloc(&o.d_queue, 2) &&
           hip_ok(hipMemsetAsync(o.d_queue, 0, 2 * sizeof(unsigned), g_stream), "memset");
}

// The launches of one batched extraction: the residual pass, then the compactions and FP32
// copies of every target.  Their shapes depend only on the observation, the target capacity
// and the work-group count, and everything else comes from the device descriptor.
static void batch_launches(Obs &o, unsigned ncap, unsigned groups)
{
    GridBatchLaunch a = {};
    a.prec = o.prec;
    a.nch = o.nch;
    a.nrow = o.nrow;
    a.u = o.d_u;
    a.v = o.d_v;
    a.w = o.d_w;
    a.wn = o.d_wn;
    a.inv_lambda = o.d_invlam;
    a.spec = o.d_spec;
    a.res = o.d_res;
    a.s2f = o.d_s2f;
    a.len = o.len;
    a.fs_scale = fx_scales(1., o.s2_max, -1, o.nch).fs_scale;
    a.desc = o.d_desc;
    a.queue = o.d_queue;
    a.ntiles = (unsigned)((o.nrow + kBatchTileRows - 1) / kBatchTileRows);
    launch_grid_batch(g_stream, a, groups);
    launch_compact_batch(g_stream, o.d_desc, ncap, o.fmax, o.len);
}

bool obs_extract_batch(Obs &o, const std::vector<ExtractReq> &reqs,
                       const std::vector<PendingSub> &subs)
{
    const int nb = (int)reqs.size();
    if (nb == 0 || nb > kMaxBatch || (int)subs.size() > kMaxBatch) {
        set_error("batched extraction: 1 to " + std::to_string(kMaxBatch) +
                  " galaxies and pending updates per pass");
        return false;
    }
    int max_slot = 0;
    for (const ExtractReq &r : reqs) {
        if (r.facet == 0 || r.facet > o.fmax) {
            set_error("facet size " + std::to_string(r.facet) +
                      " outside the allocated facet (max " + std::to_string(o.fmax) + ")");
            return false;
        }
        max_slot = std::max(max_slot, r.slot);
    }
    if (!o.ensure_slots(max_slot + 1) || !ensure_batch_buffers(o)) return false;
    // the previous batch's descriptor may still be read by its launches
    if (!hip_ok(hipStreamSynchronize(g_stream), "batch sync")) return false;
    for (const PendingSub &p : subs) o.res_bound += 3. * fabs(p.flux) * o.spec_max;
    BatchDesc &D = *o.h_desc;
    D.ntarget = (unsigned)nb;
    D.nsub = (unsigned)subs.size();
    for (int k = 0; k < nb; k++) {
        const ExtractReq &r = reqs[k];
        FacetSlot &sl = o.slots[r.slot];
        sl.facet = r.facet;
        BatchTarget &T = D.targets[k];
        T.g = make_geom(r.l0, r.m0, r.flux, r.R_mu, 0., 0.);
        T.inc = 2 * o.len / r.facet;
        T.inv_inc = 1.0 / T.inc;
        const AccScale sc =
            fx_scales(fabs(r.flux) * o.spec_max + o.res_bound, o.s2_max, -1, o.nch);
        T.fx_scale = sc.fx_scale;
        T.acc = sl.acc;
        T.facet = r.facet;
        T.pad = 0;
        BatchCompact &Cp = D.compact[k];
        Cp.acc = sl.acc;
        Cp.fx_unscale = 1.0 / sc.fx_scale;
        Cp.fs_unscale = 1.0 / sc.fs_scale;
        Cp.uu = sl.uu;
        Cp.vv = sl.vv;
        Cp.dre = sl.dre;
        Cp.dim = sl.dim;
        Cp.s2 = sl.s2;
        Cp.cid = sl.cid;
        Cp.ctr = sl.ctr;
        Cp.dc = sl.dc;
        Cp.loff = sl.loff;
        Cp.pk = sl.pk;
        Cp.invs2 = sl.invs2;
        Cp.invs2_64 = sl.invs2_64;
        Cp.n_out = &o.d_facets[r.slot].n;
        Cp.scan = sl.scan;
        Cp.facet = r.facet;
        Cp.pad = 0;
    }
    for (size_t b = 0; b < subs.size(); b++) {
        const PendingSub &p = subs[b];
        D.subs[b].fit = make_geom(p.l0, p.m0, p.flux, p.R_mu, p.e1, p.e2);
        D.subs[b].rnd = make_geom(p.l0, p.m0, p.flux, p.R_mu, 0., 0.);
    }
    const size_t desc_bytes = offsetof(BatchDesc, compact) + nb * sizeof(BatchCompact);
    if (!hip_ok(hipMemcpyAsync(o.d_desc, o.h_desc, desc_bytes, hipMemcpyHostToDevice, g_stream),
                "batch descriptor"))
        return false;
    const unsigned groups = tuning().batch_groups ? tuning().batch_groups : 3 * num_cus();
    {
        KTimer t(RL2_K_EXTRACT);
        if (tuning().batch_graph) {
            // one graph per (target capacity, work-groups), captured on first use and replayed
            // for every later batch of that shape
            const unsigned ncap = (unsigned)nb;
            const unsigned long long key = ((unsigned long long)ncap << 32) | groups;
            hipGraphExec_t exec = nullptr;
            for (auto &g : o.batch_graphs)
                if (g.first == key) exec = g.second;
            if (!exec) {
                hipGraph_t graph = nullptr;
                if (!hip_ok(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal),
                            "hipStreamBeginCapture"))
                    return false;
                batch_launches(o, ncap, groups);
                if (!hip_ok(hipStreamEndCapture(g_stream, &graph), "hipStreamEndCapture"))
                    return false;
                const bool ok = hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0),
                                       "hipGraphInstantiate");
                (void)hipGraphDestroy(graph);
                if (!ok) return false;
                o.batch_graphs.push_back({key, exec});
            }
            if (!hip_ok(hipGraphLaunch(exec, g_stream), "hipGraphLaunch")) return false;
        } else {
            batch_launches(o, (unsigned)nb, groups);
        }
    }
    const double res_b = o.prec == 0 ? 16. : 8.;
    stats_add_work(RL2_K_EXTRACT, 0., o.nvis * (res_b + 4. + (subs.empty() ? 0. : res_b)) +
                                          o.nrow * 24.);
    // grid structure of each slot's facet for the phase tables (host-side descriptors)
    for (int k = 0; k < nb; k++) {
        const int slot = reqs[k].slot;
        FacetSlot &sl = o.slots[slot];
        FacetDev &f = o.h_facets[slot];
        const unsigned facet = reqs[k].facet;
        const unsigned gf = facet < 65536u ? facet : 0u;
        if (f.cid != sl.cid || f.grid_facet != gf) {
            f.cid = gf ? sl.cid : nullptr;
            f.ctr = gf ? sl.ctr : nullptr;
            f.grid_facet = gf;
            if (!o.facet_upload(slot)) return false;
        }
    }
    return hip_ok(hipGetLastError(), "batched extraction launch");
}

bool obs_subtract_batch(Obs &o, const std::vector<PendingSub> &subs)
{
    if (subs.empty()) return true;
    if ((int)subs.size() > kMaxBatch) {
        set_error("batched subtraction: at most " + std::to_string(kMaxBatch) + " updates");
        return false;
    }
    if (!ensure_batch_buffers(o) || !hip_ok(hipStreamSynchronize(g_stream), "batch sync"))
        return false;
    BatchDesc &D = *o.h_desc;
    for (size_t b = 0; b < subs.size(); b++) {
        const PendingSub &p = subs[b];
        o.res_bound += 3. * fabs(p.flux) * o.spec_max;
        D.subs[b].fit = make_geom(p.l0, p.m0, p.flux, p.R_mu, p.e1, p.e2);
        D.subs[b].rnd = make_geom(p.l0, p.m0, p.flux, p.R_mu, 0., 0.);
    }
    if (!hip_ok(hipMemcpyAsync(o.d_desc->subs, D.subs, subs.size() * sizeof(BatchSub),
                               hipMemcpyHostToDevice, g_stream),
                "batched subtraction upload"))
        return false;
    {
        KTimer t(RL2_K_SUBTRACT);
        launch_subtract_batch(g_stream, o.prec, o.nch, o.nrow, o.d_u, o.d_v, o.d_w, o.d_wn,
                              o.d_invlam, o.d_spec, o.d_desc->subs, (unsigned)subs.size(),
                              o.d_res);
    }
    stats_add_work(RL2_K_SUBTRACT, 0., o.nvis * bytes_per_vis(o, RL2_K_SUBTRACT) + o.nrow * 24.);
    return hip_ok(hipGetLastError(), "batched subtraction launch");
}

bool obs_subtract(Obs &o, double l0, double m0, double flux, double R_mu, double e1, double e2)
{
    o.res_bound += 3. * fabs(flux) * o.spec_max;  // |fitted model - round model| per component
    GalaxyGeom fit = make_geom(l0, m0, flux, R_mu, e1, e2);
    GalaxyGeom rnd = make_geom(l0, m0, flux, R_mu, 0., 0.);
    {
        KTimer t(RL2_K_SUBTRACT);
        launch_subtract(g_stream, o.prec, o.nch, o.nrow, o.d_u, o.d_v, o.d_w, o.d_wn, o.d_invlam,
                        o.d_spec, fit,
                        rnd, o.d_res);
    }
    stats_add_work(RL2_K_SUBTRACT, 0., o.nvis * bytes_per_vis(o, RL2_K_SUBTRACT) + o.nrow * 24.);
    return hip_ok(hipGetLastError(), "subtract launch");
}

bool lik_eval(Obs &o, const std::vector<LikItem> &items, double *Le, int *err, double *Lr,
              int *nit)
{
    const int n = (int)items.size();
    if (n == 0) return true;
    if (!o.lik.ensure(n)) return false;
    LikBuffers &b = o.lik;
    memcpy(b.h_items, items.data(), n * sizeof(LikItem));
    LikLaunch a;
    a.nitems = n;
    a.items = b.d_items;
    a.facets = o.d_facets;
    a.rg = o.rg;
    a.Lr = b.d_Lr;
    a.nit = b.d_nit;
    a.Le = b.d_Le;
    a.err = b.d_err;
    a.qag_ws = b.d_ws;
    a.qag_iws = b.d_iws;
    a.block = o.lik_block;
    a.prec = o.prec;
    std::vector<int> hp(n);
    for (int i = 0; i < n; i++) hp[i] = items[i].prec;
    a.host_prec = hp.data();
    bool tiled = o.lik_block == 0;
    for (int i = 0; i < n && tiled; i++)
        if (o.h_facets[items[i].slot].mod) tiled = false;  // given model: one group per (e, r)
    const bool want_detail = Lr || nit || (g_stats_on && !tiled);
    // Every copy, fill and stream dependency queued below ANDs its status into q_ok, which
    // is checked with the launches' before the results are read.
    bool q_ok = true;
    // one copy back per evaluation: status, FLOP count, L_e and the error flags
    auto copy_out = [&]() {  // the results block itself is written to host memory
        if (want_detail)
            q_ok = hip_ok(hipMemcpyAsync(b.h_Lr, b.d_Lr, (size_t)n * kMaxR * sizeof(double),
                                         hipMemcpyDeviceToHost, g_stream),
                          "L_r copy") &&
                   hip_ok(hipMemcpyAsync(b.h_nit, b.d_nit, (size_t)n * kMaxR * sizeof(int),
                                         hipMemcpyDeviceToHost, g_stream),
                          "step count copy") &&
                   q_ok;
    };
    if (tiled) {
        // chunks per (item, r) from the host copy of each slot's cell count.  An item's chunk
        // depends on its own facet only (the largest power of two from kTileMinCells that
        // still gives the item's tasks kTileItemGroups work-groups), so its chunk sums, and
        // with them its likelihood, do not depend on the other items of the launch.  A
        // launch typically carries 3 to 8 items (the simplex's candidates, a line search's
        // steps), which then fill the device with a few thousand work-groups.
        const int nr = o.rg.numr - 1;
        // FP32 items run the two-kernel passes, whose shorter work-groups favour the
        // smaller chunks of kTileItemGroups32
        auto item_chunk = [&](unsigned cells, int prec) {
            unsigned chunk = tuning().tile_cells;
            if (chunk == 0) {
                const unsigned long long target =
                    tuning().item_groups ? tuning().item_groups
                                         : (prec == 0 ? kTileItemGroups : kTileItemGroups32);
                chunk = kTileMinCells;
                while (chunk < kTileMaxCells &&
                       (unsigned long long)nr * ((cells + 2 * chunk - 1) / (2 * chunk)) >= target)
                    chunk *= 2;
            }
            return chunk;
        };
        unsigned chunks = 0;
        for (int i = 0; i < n; i++) {
            const unsigned cells = o.slots[items[i].slot].n;
            const unsigned chunk = item_chunk(cells, items[i].prec);
            b.h_items[i].chunk_first = chunks;
            b.h_items[i].chunk_cells = chunk;
            chunks += std::max(1u, (cells + chunk - 1) / chunk);
        }
        const unsigned total = chunks * (unsigned)nr;
        if (!b.ensure_partials(2 * (size_t)total)) return false;  // two generations
        // one chunk count for every item (the points of one galaxy): the kernels then find a
        // work-group's item by a division instead of a search
        unsigned item_chunks = n > 0 ? chunks / (unsigned)n : 0u;
        for (int i = 0; i < n && item_chunks; i++)
            if (b.h_items[i].chunk_first != (unsigned)i * item_chunks) item_chunks = 0;
        if (item_chunks * (unsigned)n != chunks) item_chunks = 0;
        // phase tables for FP64 items on grid facets: 2 * grid_facet entries per task
        size_t ntab = 0;
        unsigned tab_facet = 0;
        int ntab_items = 0;
        for (int i = 0; i < n; i++) {
            const FacetDev &f = o.h_facets[items[i].slot];
            b.h_items[i].tab_first = (unsigned)ntab;
            if (tuning().phase_tables && items[i].prec == 0 && f.cid && f.grid_facet) {
                ntab += (size_t)nr * 2 * f.grid_facet;
                tab_facet = std::max(tab_facet, f.grid_facet);
                ntab_items++;
            }
        }
        // offsets are 32-bit, and a work-group stages one task's column table (24 bytes per
        // column) in LDS: beyond either limit the passes evaluate sin/cos per cell
        if (ntab > 0xffffffffu || tab_facet * 24u > kTileLdsBytes) ntab = 0;
        if (ntab && !b.ensure_tab(ntab)) return false;
        if (n > kInlItems)  // few items travel with the first kernel's arguments instead
            q_ok = hip_ok(hipMemcpyAsync(b.d_items, b.h_items, n * sizeof(LikItem),
                                         hipMemcpyHostToDevice, g_stream),
                          "items copy") &&
                   q_ok;
        a.tasks = b.d_tasks;
        // the scalelength tasks in one or two groups: the second group's passes run on a
        // second stream, so one group's step kernels and launch gaps overlap the other's sums
        const int ng = (tuning().tile_streams == 2 && nr >= 2 && ensure_stream2()) ? 2 : 1;
        TileLaunch tg[2];
        for (int g = 0; g < ng; g++) {
            TileLaunch &t = tg[g];
            t.lk = a;
            t.tasks = b.d_tasks;
            t.r_lo = g == 0 ? 1 : 1 + (nr + 1) / 2;
            t.nr_g = ng == 1 ? nr : (g == 0 ? (nr + 1) / 2 : nr / 2);
            t.nwg = chunks * (unsigned)t.nr_g;
            t.partials = b.d_partials + (g == 0 ? 0 : (size_t)tg[0].nwg * 8);
            t.tab = ntab ? b.d_tab : nullptr;
            t.tab_lds_facet = ntab ? tab_facet : 0;
            t.tab_y_lds = ntab && tuning().phase_tables == 2 && tab_facet * 40u <= kTileLdsBytes;
            // the grid kernels when every item has tables (FP64 on an extracted facet); 2:
            // with the items in the kernel arguments
            t.grid = (ntab && ntab_items == n) ? (n <= kInlItems ? 2u : 1u) : 0u;
            t.items_w = b.d_items;
            t.item_chunks = item_chunks;
            t.n_inl = n <= kInlItems ? (unsigned)n : 0u;
            for (unsigned i = 0; i < t.n_inl; i++) t.inl[i] = b.h_items[i];
        }
        // Passes `first` .. `first + count - 1` (pass 0: zero-offset sums; every task's state
        // is (re)initialised by the pass-0 step kernel).  The second stream starts after
        // everything queued so far, and the r-marginalisation waits for both.
        auto passes_of = [&](int first, int count) {
            if (ng == 2)
                q_ok = hip_ok(hipEventRecord(g_fork, g_stream), "stream fork") &&
                       hip_ok(hipStreamWaitEvent(g_stream2, g_fork, 0), "stream fork") && q_ok;
            for (int p = first; p < first + count; p++) {
                const int pass = p == 0 ? 0 : 1;
                if (ng == 2) {
                    KTimer k(RL2_K_LIKELIHOOD);
                    launch_tile_pass(g_stream, tg[0], pass);
                    launch_tile_pass(g_stream2, tg[1], pass);
                    continue;
                }
                {
                    KTimer k(RL2_K_LIKELIHOOD);  // the chunk sums: the kernel the FLOPs count
                    launch_tile_sums(g_stream, tg[0], pass);
                }
                KTimer k(RL2_K_STEP);
                launch_tile_step(g_stream, tg[0], pass);
            }
            if (ng == 2)
                q_ok = hip_ok(hipEventRecord(g_join, g_stream2), "stream join") &&
                       hip_ok(hipStreamWaitEvent(g_stream, g_join, 0), "stream join") && q_ok;
        };
        // Fused passes (RL2_TUNE_FUSE_STEP): every pass after pass 0 takes the previous
        // pass's Newton step in its own work-groups, and one step kernel after the last
        // queued pass finishes the tasks that converged with it.  Task state and chunk sums
        // alternate between two generations (index si / pi: the latest).  They need every
        // item on a grid facet with its row table in LDS, and one stream.
        // (every item on a grid facet with its tables; facets up to kRecFacet wide use the
        // 40-byte records, wider ones, up to the column tables' LDS limit, column records)
        const bool fused = tuning().fuse_step && ng == 1 && tg[0].grid;
        // One launch per evaluation (RL2_TUNE_FUSE_STEP 2): every block resident, XCD x
        // owning an eighth of each task's chunks, each block an equal run of its XCD's
        // (task, chunk) list; runs that would span more tasks than a block holds fall
        // back to the fused passes.
        if (fused && tuning().fuse_step == 2 && tg[0].tab_y_lds) {
            TileLaunch &t0 = tg[0];
            const unsigned cap = persist_capacity(t0.tab_lds_facet * kRecBytes, (int)t0.grid);
            unsigned long long W[kXcds] = {}, minm[kXcds];
            unsigned long long wmax = 0;
            for (unsigned x = 0; x < kXcds; x++) {
                minm[x] = ~0ull;
                for (int i = 0; i < n; i++) {
                    const unsigned c1 = i + 1 < n ? b.h_items[i + 1].chunk_first : chunks;
                    const unsigned nc = c1 - b.h_items[i].chunk_first;
                    const unsigned m = (x + 1) * nc / kXcds - x * nc / kXcds;
                    W[x] += (unsigned long long)nr * m;
                    if (m) minm[x] = std::min<unsigned long long>(minm[x], m);
                }
                wmax = std::max(wmax, W[x]);
            }
            unsigned long long nbx = std::min<unsigned long long>(cap / kXcds, wmax);
            bool ok = nbx > 0;
            for (unsigned x = 0; x < kXcds && ok; x++) {
                if (!W[x]) continue;
                const unsigned long long L = (W[x] + nbx - 1) / nbx;
                ok = (L + minm[x] - 1) / minm[x] + 1 <= (unsigned long long)kPersistTasks;
            }
            if (ok) {
                if (!b.ensure_ppart(total)) return false;
                q_ok = hip_ok(hipMemsetAsync(b.d_cnt, 0, (size_t)n * kMaxR * sizeof(unsigned),
                                             g_stream),
                              "counter fill") &&
                       q_ok;
                PersistLaunch P;
                P.t = t0;
                P.t.tasks = b.d_tasks;
                P.part = b.d_ppart;
                P.cnt = b.d_cnt;
                P.nbx = (unsigned)nbx;
                P.spin_limit = 1u << 15;
                {
                    KTimer k(RL2_K_LIKELIHOOD);  // pass 0 into the sums' first slice
                    TileLaunch p0 = t0;
                    p0.partials = b.d_ppart;
                    launch_tile_sums(g_stream, p0, 0);
                }
                {
                    KTimer k(RL2_K_LIKELIHOOD);
                    if (!launch_persist(g_stream, P, (unsigned)nbx * kXcds)) {
                        set_error("likelihood: cooperative launch failed");
                        return false;
                    }
                }
                {
                    KTimer k(RL2_K_RMARG);
                    LikLaunch ra = a;
                    ra.tasks = b.d_tasks;
                    ra.status = b.d_status;
                    ra.flops = b.d_flops;
                    launch_rmarg(g_stream, ra);
                }
                copy_out();
                if (!hip_ok(hipGetLastError(), "likelihood launch") || !q_ok) return false;
                if (!hip_ok(hipStreamSynchronize(g_stream), "likelihood sync")) return false;
                if (b.h_status[0] != 0) {  // a task left unfinished: a wait gave up
                    set_error("likelihood: a device wait exceeded its bound");
                    return false;
                }
                if (g_stats_on) stats_add_work(RL2_K_LIKELIHOOD, *b.h_flops, 0.);
                memcpy(Le, b.h_Le, n * sizeof(double));
                memcpy(err, b.h_err, n * sizeof(int));
                if (Lr) memcpy(Lr, b.h_Lr, (size_t)n * kMaxR * sizeof(double));
                if (nit) memcpy(nit, b.h_nit, (size_t)n * kMaxR * sizeof(int));
                return true;
            }
        }
        if (fused) {
            TileTask *st[2] = {b.d_tasks, b.d_tasks + (size_t)b.cap * kMaxR};
            double *pt[2] = {b.d_partials, b.d_partials + (size_t)total * 8};
            int si = 0, pi = 0, mode = 0, passes = 0;
            int queued = tuning().tile_passes ? tuning().tile_passes : b.tile_passes;
            {
                TileLaunch p0 = tg[0];
                p0.partials = pt[0];
                KTimer k(RL2_K_LIKELIHOOD);
                launch_tile_sums(g_stream, p0, 0);
            }
            while (true) {
                for (int q = 0; q < queued; q++) {
                    TileLaunch f = tg[0];
                    f.mode = mode;
                    f.tasks = st[si];
                    f.tasks_out = st[si ^ 1];
                    f.part_in = pt[pi];
                    f.partials = pt[pi ^ 1];
                    {
                        KTimer k(RL2_K_LIKELIHOOD);
                        launch_tile_fused(g_stream, f);
                    }
                    if (mode != 2) si ^= 1;
                    pi ^= 1;
                    mode = 1;
                }
                passes += queued;
                {
                    TileLaunch sk = tg[0];
                    sk.tasks = st[si];
                    sk.tasks_out = st[si ^ 1];
                    sk.partials = pt[pi];
                    sk.tab = nullptr;  // no tables: the next pass, if any, makes its own
                    KTimer k(RL2_K_STEP);
                    launch_tile_step(g_stream, sk, 1);
                    si ^= 1;
                }
                {
                    KTimer k(RL2_K_RMARG);
                    LikLaunch ra = a;  // plus one work-group writing the status block
                    ra.tasks = st[si];
                    ra.status = b.d_status;
                    ra.flops = b.d_flops;
                    launch_rmarg(g_stream, ra);
                }
                copy_out();
                if (!hip_ok(hipGetLastError(), "likelihood launch") || !q_ok) return false;
                if (!hip_ok(hipStreamSynchronize(g_stream), "likelihood sync")) return false;
                if (b.h_status[0] == 0) break;
                if (passes > 50) {  // the step count is capped at 50 (likelihood.c:284)
                    set_error("likelihood: Newton passes did not terminate");
                    return false;
                }
                queued = 2;
                mode = 2;  // the step kernel took the last step: the next pass resumes from it
            }
            b.tile_passes = std::min(12, std::max(2, (int)b.h_status[1] + 1));
            if (g_stats_on) stats_add_work(RL2_K_LIKELIHOOD, *b.h_flops, 0.);
            memcpy(Le, b.h_Le, n * sizeof(double));
            memcpy(err, b.h_err, n * sizeof(int));
            if (Lr) memcpy(Lr, b.h_Lr, (size_t)n * kMaxR * sizeof(double));
            if (nit) memcpy(nit, b.h_nit, (size_t)n * kMaxR * sizeof(int));
            return true;
        }
        // Newton passes are queued ahead of the status check (the previous evaluation's
        // largest step count + 1); a pass leaves finished tasks at once.  The
        // r-marginalisation is queued with them and repeated if tasks were still iterating.
        int queued = tuning().tile_passes ? tuning().tile_passes : b.tile_passes;
        int passes = 0, first = 0;
        while (true) {
            passes_of(first, first == 0 ? queued + 1 : queued);
            first = 1;
            passes += queued;
            {
                KTimer k(RL2_K_RMARG);
                LikLaunch ra = a;  // plus one work-group writing the status block
                ra.status = b.d_status;
                ra.flops = b.d_flops;
                launch_rmarg(g_stream, ra);
            }
            copy_out();
            if (!hip_ok(hipGetLastError(), "likelihood launch") || !q_ok) return false;
            if (!hip_ok(hipStreamSynchronize(g_stream), "likelihood sync")) return false;
            if (b.h_status[0] == 0) break;
            if (passes > 50) {  // the step count is capped at 50 (likelihood.c:284)
                set_error("likelihood: Newton passes did not terminate");
                return false;
            }
            queued = 2;
        }
        b.tile_passes = std::min(12, std::max(2, (int)b.h_status[1] + 1));
        // FLOP-equivalents summed on the device: n * (19 + 18 + 25 * N_it) per (item, r)
        if (g_stats_on) stats_add_work(RL2_K_LIKELIHOOD, *b.h_flops, 0.);
    } else {
        q_ok = hip_ok(hipMemcpyAsync(b.d_items, b.h_items, n * sizeof(LikItem),
                                     hipMemcpyHostToDevice, g_stream),
                      "items copy");
        {
            KTimer t(RL2_K_LIKELIHOOD);
            launch_likelihood(g_stream, a);
        }
        {
            KTimer t(RL2_K_RMARG);
            launch_rmarg(g_stream, a);
        }
        if (!hip_ok(hipGetLastError(), "likelihood launch")) return false;
        copy_out();
        if (!q_ok || !hip_ok(hipStreamSynchronize(g_stream), "likelihood sync")) return false;
        if (g_stats_on) {
            double fl = 0.;
            for (int i = 0; i < n; i++) {
                const double nn = o.slots[items[i].slot].n;
                for (int r = 1; r < o.rg.numr; r++)
                    fl += nn * (37. + 25. * b.h_nit[(size_t)i * kMaxR + r]);
            }
            stats_add_work(RL2_K_LIKELIHOOD, fl, 0.);
        }
    }
    memcpy(Le, b.h_Le, n * sizeof(double));
    memcpy(err, b.h_err, n * sizeof(int));
    if (Lr) memcpy(Lr, b.h_Lr, (size_t)n * kMaxR * sizeof(double));
    if (nit) memcpy(nit, b.h_nit, (size_t)n * kMaxR * sizeof(int));
    return true;
}

static bool read_slot_n(Obs &o, int slot, unsigned *n)
{
    return hip_ok(hipMemcpyAsync(n, &o.d_facets[slot].n, sizeof *n, hipMemcpyDeviceToHost, g_stream),
                  "copy n") &&
           hip_ok(hipStreamSynchronize(g_stream), "sync");
}

static bool prior_for(Obs &o, int slot, double flux)
{
    const double mu = scale_mean(flux);
    FacetSlot &s = o.slots[slot];
    for (int k = 0; k < kMaxR; k++) s.rprior[k] = 0.;
    for (int nRo = 1; nRo < o.rg.numr; nRo++) s.rprior[nRo] = rfunc(mu, kRstd, o.rg.ro[nRo]);
    return o.facet_upload(slot);
}

// One pass of data_processing.c:49-142 over `todo` (batch > 1: data_processing-mpi.c:65-216).
// pass 0 collects failed fits in `new_bad`; pass 1 (re-fit) reports them with e = 0.
// Network inputs of the nb facets just extracted (slots 0..nb-1): copied out as training
// data and/or run through the learned proposal, whose peak and width become each fit's
// simplex start and first step.
static bool proposal_inputs(Obs &o, const std::vector<unsigned> &todo, size_t b0, size_t nb,
                            int pass, const double *flux, const double *scale,
                            const rl2_fit_options &opt, std::vector<FitRequest> &reqs)
{
    float *x = nullptr;
    if (opt.proposal) {
        if (!opt.proposal->net.ensure((int)nb)) return false;
        x = opt.proposal->net.x;
    } else {
        if (o.feat_cap < nb) {
            dfree(o.d_feat);
            if (!dmalloc(&o.d_feat, nb * kMlpIn)) return false;
            o.feat_cap = nb;
        }
        x = o.d_feat;
    }
    for (size_t k = 0; k < nb; k++) {
        const unsigned gal = todo[b0 + k];
        launch_mlp_features(g_stream, o.d_facets, (int)k, scale[gal], flux[gal], x + k * kMlpIn);
    }
    if (opt.features_out && pass == 0)
        for (size_t k = 0; k < nb; k++)
            if (!hip_ok(hipMemcpyAsync(opt.features_out + (size_t)todo[b0 + k] * kMlpIn,
                                       x + k * kMlpIn, kMlpIn * sizeof(float),
                                       hipMemcpyDeviceToHost, g_stream),
                        "features copy")) {
                (void)hipStreamSynchronize(g_stream);  // the copies queued before it land first
                return false;
            }
    if (opt.proposal) {
        Mlp &net = opt.proposal->net;
        if (!net.forward(g_stream, (int)nb)) return false;
        std::vector<float> out(nb * kMlpOut);
        const bool copied = hip_ok(hipMemcpyAsync(out.data(), net.act[kMlpLayers - 1],
                                                  out.size() * sizeof(float),
                                                  hipMemcpyDeviceToHost, g_stream),
                                   "proposal copy");
        if (!hip_ok(hipStreamSynchronize(g_stream), "proposal") || !copied) return false;
        for (size_t k = 0; k < nb; k++) {
            const float *q = out.data() + k * kMlpOut;
            // covariance of (e1, e2, log r) = (L L^T)^-1 for the precision factor L
            const double l00 = exp((double)q[3]), l11 = exp((double)q[4]), l22 = exp((double)q[5]);
            const double l10 = q[6], l20 = q[7], l21 = q[8];
            // inverse of the lower-triangular L
            const double i00 = 1. / l00, i11 = 1. / l11, i22 = 1. / l22;
            const double i10 = -l10 * i00 * i11;
            const double i21 = -l21 * i11 * i22;
            const double i20 = -(l20 * i00 + l21 * i10) * i22;
            // Sigma = L^-T L^-1: the e1, e2 variances
            const double v1 = i00 * i00 + i10 * i10 + i20 * i20;
            const double v2 = i11 * i11 + i21 * i21;
            const double sd = sqrt(fmax(v1, v2));
            if (!(isfinite(q[0]) && isfinite(q[1]) && isfinite(sd))) continue;
            reqs[k].has_prop = opt.proposal_mode == 0;
            reqs[k].prop_e1 = q[0];
            reqs[k].prop_e2 = q[1];
            reqs[k].prop_step = std::min(0.1, std::max(0.01, 2. * sd));
            // the sampling line searches stop near 2.45 sigma (L / Lmax = 0.05 for a
            // Gaussian): request that many 0.05 steps per launch (speculation only)
            reqs[k].prop_chunk = std::min(8, std::max(2, (int)ceil(2.45 * sd / 0.05) + 1));
        }
    }
    return hip_ok(hipGetLastError(), "proposal inputs") &&
           hip_ok(hipStreamSynchronize(g_stream), "proposal inputs");
}

// Training labels of nb fitted galaxies: (e1, e2, log r*) with r* maximising
// L_r + log prior at the fitted ellipticity; NaN for failed fits.
static bool fit_labels(Obs &o, const std::vector<unsigned> &todo, size_t b0, size_t nb,
                       const rl2_gal_result *results, double *labels)
{
    std::vector<LikItem> items(nb);
    for (size_t k = 0; k < nb; k++) {
        const rl2_gal_result &r = results[todo[b0 + k]];
        items[k] = {(int)k, 0, r.e1, r.e2};
    }
    std::vector<double> Le(nb), Lr(nb * kMaxR);
    std::vector<int> err(nb);
    if (!lik_eval(o, items, Le.data(), err.data(), Lr.data(), nullptr)) return false;
    for (size_t k = 0; k < nb; k++) {
        const unsigned gal = todo[b0 + k];
        const rl2_gal_result &r = results[gal];
        double *lab = labels + 3 * (size_t)gal;
        int best = -1;
        double bv = -1e300;
        for (int i = 1; i < o.rg.numr; i++) {
            const double lr = Lr[k * kMaxR + i];
            if (lr <= kBad) continue;
            const double val = lr + o.slots[k].rprior[i];
            if (val > bv) {
                bv = val;
                best = i;
            }
        }
        if (r.error || best < 0) {
            lab[0] = lab[1] = lab[2] = NAN;
        } else {
            lab[0] = r.e1;
            lab[1] = r.e2;
            lab[2] = log(o.rg.ro[best]);
        }
    }
    return true;
}

bool process_list(Obs &o, const std::vector<unsigned> &todo, int pass, const double *l,
                  const double *m, const double *flux, const double *scale,
                  const rl2_fit_options &opt, rl2_gal_result *results,
                  std::vector<unsigned> *new_bad, double *t_ext, double *t_fit, double *t_sub,
                  const std::function<void(unsigned)> &on_done)
{
    const int B = opt.batch > 0 ? opt.batch : 1;
    if (!o.ensure_slots(B)) return false;
    // The last subtraction of a batch can ride on the first extraction of the next one
    // (one pass over the residual instead of two): it is held here until then, and made on
    // its own at the end of the list.
    const bool fuse = tuning().fuse_subtract &&
                      (tuning().grid_variant == 3 || tuning().grid_variant >= 5);
    PendingSub pend;
    bool pending = false;
    // batches extracted in one residual pass each, carrying the previous batch's
    // subtractions (in order) in the same pass
    const bool batched = B > 1 && B <= kMaxBatch && tuning().batch_extract;
    std::vector<PendingSub> pend_list;
    o.lik_block = opt.lik_block > 0 ? opt.lik_block : 0;
    BatchEval eval = [&](const std::vector<LikItem> &items, std::vector<double> &Le,
                         std::vector<int> &err) {
        return lik_eval(o, items, Le.data(), err.data(), nullptr, nullptr);
    };
    for (size_t b0 = 0; b0 < todo.size(); b0 += B) {
        const size_t nb = std::min((size_t)B, todo.size() - b0);
        double t0 = now_s();
        if (batched) {
            std::vector<ExtractReq> reqs(nb);
            for (size_t k = 0; k < nb; k++) {
                const unsigned gal = todo[b0 + k];
                if (!prior_for(o, (int)k, flux[gal])) return false;
                const unsigned facet = facet_size(scale[gal], o.len);
                reqs[k] = {(int)k, l[gal], m[gal], flux[gal], scale[gal], facet};
                results[gal].facet = facet;
            }
            if (!obs_extract_batch(o, reqs, pend_list)) return false;
            pend_list.clear();
        }
        for (size_t k = 0; k < nb && !batched; k++) {
            const unsigned gal = todo[b0 + k];
            if (!prior_for(o, (int)k, flux[gal])) return false;
            const unsigned facet = facet_size(scale[gal], o.len);
            if (!obs_extract(o, (int)k, l[gal], m[gal], flux[gal], scale[gal], facet, false, 0.,
                             0., pending ? &pend : nullptr))
                return false;
            pending = false;
            results[gal].facet = facet;
        }
        for (size_t k = 0; k < nb; k++) {
            unsigned nn = 0;
            if (!read_slot_n(o, (int)k, &nn)) return false;
            o.slots[k].n = nn;
            rl2_gal_result &r = results[todo[b0 + k]];
            r.ncoords = nn;
            r.pass = pass;
            r.subtracted = 0;
        }
        std::vector<FitRequest> reqs;
        for (size_t k = 0; k < nb; k++) reqs.push_back({(int)k, &results[todo[b0 + k]]});
        if (opt.proposal || (opt.features_out && pass == 0)) {
            if (!proposal_inputs(o, todo, b0, nb, pass, flux, scale, opt, reqs)) return false;
        }
        double t1 = now_s();
        *t_ext += t1 - t0;
        if (!fit_batch(reqs, opt, eval)) return false;
        if (opt.labels_out && pass == 0 && !fit_labels(o, todo, b0, nb, results, opt.labels_out))
            return false;
        double t2 = now_s();
        *t_fit += t2 - t1;
        int last_good = -1;
        for (size_t k = 0; k < nb; k++)
            if (!results[todo[b0 + k]].error) last_good = (int)k;
        const bool more = b0 + nb < todo.size();
        for (size_t k = 0; k < nb; k++) {
            const unsigned gal = todo[b0 + k];
            rl2_gal_result &r = results[gal];
            if (r.error) {
                if (pass == 0) {
                    if (new_bad) new_bad->push_back(gal);
                } else {
                    r.e1 = 0.;
                    r.e2 = 0.;
                }
            } else {
                if (batched) {
                    pend_list.push_back({l[gal], m[gal], flux[gal], scale[gal], r.e1, r.e2});
                } else if (fuse && more && (int)k == last_good) {
                    pend = {l[gal], m[gal], flux[gal], scale[gal], r.e1, r.e2};
                    pending = true;
                } else if (!obs_subtract(o, l[gal], m[gal], flux[gal], scale[gal], r.e1, r.e2)) {
                    return false;
                }
                r.subtracted = 1;
            }
            if (on_done) on_done(gal);
        }
        if (!hip_ok(hipStreamSynchronize(g_stream), "subtract sync")) return false;
        *t_sub += now_s() - t2;
        if (opt.verbose && (b0 / B) % (size_t)opt.verbose == 0)
            fprintf(stderr, "[rl2] pass %d: %zu of %zu galaxies fitted\n", pass, b0 + nb,
                    todo.size());
    }
    if (pending && !obs_subtract(o, pend.l0, pend.m0, pend.flux, pend.R_mu, pend.e1, pend.e2))
        return false;
    if (!obs_subtract_batch(o, pend_list)) return false;
    return hip_ok(hipStreamSynchronize(g_stream), "subtract sync");
}

bool process(Obs &o, unsigned ngal, const double *l, const double *m, const double *flux,
             const double *scale, const rl2_fit_options &opt, rl2_gal_result *results,
             rl2_run_stats *stats)
{
    double t_ext = 0., t_fit = 0., t_sub = 0.;
    std::vector<unsigned> todo(ngal), bad;
    for (unsigned g = 0; g < ngal; g++) todo[g] = g;
    if (!process_list(o, todo, 0, l, m, flux, scale, opt, results, &bad, &t_ext, &t_fit, &t_sub,
                      nullptr))
        return false;
    const int n_bad_first = (int)bad.size();
    int n_bad_final = n_bad_first;
    if (opt.do_refit && !bad.empty()) {
        if (!process_list(o, bad, 1, l, m, flux, scale, opt, results, nullptr, &t_ext, &t_fit,
                          &t_sub, nullptr))
            return false;
        n_bad_final = 0;
        for (unsigned g : bad)
            if (results[g].error) n_bad_final++;
    }
    if (stats) {
        stats->extraction_s = t_ext;
        stats->fitting_s = t_fit;
        stats->subtraction_s = t_sub;
        stats->n_bad_first = n_bad_first;
        stats->n_bad_final = n_bad_final;
    }
    return true;
}

}  // namespace rl2

using namespace rl2;

// ====================================================================== C-ABI
extern "C" {

const char *rl2_last_error(void) { return g_err.c_str(); }

int rl2_init(int device)
{
    if (g_stream && device == g_device) return RL2_OK;
    if (g_stream) {
        set_error("rl2_init: already initialised on another device");
        return RL2_ERR_STATE;
    }
    g_device = device;
    return ensure_init() ? RL2_OK : RL2_ERR_NO_DEVICE;
}

int rl2_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return RL2_OK;
}

int rl2_r_grid(double *ro)
{
    // RadioLensfit2.cpp:370-392
    const double Rmin = kRmin, Rmax = kRmax;
    ro[0] = 0.;
    ro[1] = Rmin;
    int nRo = 2;
    while (nRo < kNumR && ro[nRo - 1] < Rmax) {
        double Rinterval = 0.08 + 0.4 * pow(((ro[nRo - 1] - Rmin) / (Rmax - Rmin)), 2);
        ro[nRo] = ro[nRo - 1] + Rinterval;
        nRo++;
    }
    return nRo;
}

int rl2_marginalise_posterior_r_noalloc(int num_marvals, const double *xmarvals,
                                        const double *ymarvals, double *out, void *stream)
{
    if (!g_stream) {  // set-up (stream, scratch address) is rl2_init's, not this call's
        set_error("rl2_marginalise_posterior_r_noalloc: call rl2_init first");
        return RL2_ERR_STATE;
    }
    if (num_marvals < 3 || num_marvals > kMaxR) {
        set_error("rl2_marginalise_posterior_r_noalloc: " + std::to_string(num_marvals) +
                  " points; 3 to " + std::to_string(kMaxR) + " fit its fixed scratch");
        return RL2_ERR_RANGE;
    }
    if (!xmarvals || !ymarvals || !out) {
        set_error("rl2_marginalise_posterior_r_noalloc: null device pointer");
        return RL2_ERR_ARG;
    }
    double *ws = (double *)g_scratch[g_device];
    int *iws = (int *)(ws + kQagDoubles);
    launch_marginalise_r(stream ? (hipStream_t)stream : g_stream, num_marvals, xmarvals,
                         ymarvals, out, ws, iws);
    return hip_ok(hipGetLastError(), "marginalise_posterior_r launch") ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_create(rl2_obs **out, unsigned int nch, unsigned long nrow, const double *u,
                   const double *v, const double *w, const double *wavenumbers,
                   const double *spec, const float *sigma2_vis, const unsigned char *flag,
                   double len, int precision)
{
    *out = nullptr;
    if (!ensure_init()) return RL2_ERR_NO_DEVICE;
    if (!u || !v || !w || !wavenumbers || !spec || !sigma2_vis || nch == 0 || nrow == 0 ||
        !(len > 0.) || (precision != RL2_PREC_FP64 && precision != RL2_PREC_MIXED)) {
        set_error("rl2_obs_create: invalid arguments");
        return RL2_ERR_ARG;
    }
    if (nrow >= (1ul << 32) || (unsigned long long)nch * nrow >= (1ull << 40)) {
        set_error("rl2_obs_create: observation too large");
        return RL2_ERR_RANGE;
    }
    rl2_obs *h = new rl2_obs;
    Obs &o = h->o;
    o.prec = precision;
    o.nch = nch;
    o.nrow = nrow;
    o.nvis = (unsigned long)nch * nrow;
    o.len = len;
    o.fmax = facet_size(kRmax, len);
    o.rg.numr = rl2_r_grid(o.rg.ro);
    for (int k = o.rg.numr; k < kMaxR; k++) o.rg.ro[k] = 0.;
    o.wn.assign(wavenumbers, wavenumbers + nch);
    o.spec.assign(spec, spec + nch);
    o.invlam.resize(nch);
    for (unsigned c = 0; c < nch; c++) {
        o.invlam[c] = wavenumbers[c] / (2. * kPi);
        o.spec_max = fmax(o.spec_max, fabs(spec[c]));
    }
    for (size_t i = 0; i < o.nvis; i++) o.s2_max = fmax(o.s2_max, (double)fabsf(sigma2_vis[i]));
    hipStream_t s = g_stream;
    double *tu = nullptr, *tv = nullptr, *tw = nullptr;
    uint32_t *keys = nullptr, *idx = nullptr, *keys2 = nullptr, *idx2 = nullptr;
    float *ts2 = nullptr;
    unsigned char *tfl = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = sort_rows_tmp_bytes(nrow);
    bool ok = dmalloc(&o.d_u, nrow) && dmalloc(&o.d_v, nrow) && dmalloc(&o.d_w, nrow) &&
              dmalloc(&o.d_perm, nrow) && dmalloc(&o.d_s2f, o.nvis) && dmalloc(&o.d_wn, nch) &&
              dmalloc(&o.d_invlam, nch) && dmalloc(&o.d_spec, nch) && dmalloc(&tu, nrow) &&
              dmalloc(&tv, nrow) && dmalloc(&tw, nrow) && dmalloc(&keys, nrow) &&
              dmalloc(&idx, nrow) && dmalloc(&keys2, nrow) && dmalloc(&idx2, nrow) &&
              dmalloc(&ts2, o.nvis) && (!flag || dmalloc(&tfl, o.nvis)) &&
              hip_ok(hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1), "hipMalloc");
    size_t res_bytes = o.nvis * (precision == RL2_PREC_FP64 ? 16 : 8);
    ok = ok && hip_ok(hipMalloc(&o.d_res, res_bytes), "hipMalloc residual");
    auto up = [&](void *d, const void *hsrc, size_t bytes) {
        return hip_ok(hipMemcpyAsync(d, hsrc, bytes, hipMemcpyHostToDevice, s), "obs upload");
    };
    ok = ok && up(tu, u, nrow * 8) && up(tv, v, nrow * 8) && up(tw, w, nrow * 8) &&
         up(ts2, sigma2_vis, o.nvis * 4) && (!flag || up(tfl, flag, o.nvis)) &&
         up(o.d_wn, o.wn.data(), nch * 8) && up(o.d_invlam, o.invlam.data(), nch * 8) &&
         up(o.d_spec, o.spec.data(), nch * 8);
    if (ok) {
        // spatial sort of the rows (Morton order of (u, v) over [-len_m, len_m])
        double umax = 0.;
        for (unsigned c = 0; c < nch; c++) umax = fmax(umax, len * 2. * kPi / wavenumbers[c]);
        umax *= 1.0000001;
        launch_morton_keys(s, nrow, tu, tv, umax, keys, idx);
        uint32_t *sorted = nullptr;
        ok = sort_rows(s, nrow, keys, idx, keys2, idx2, tmp, tmp_bytes, &sorted) &&
             hip_ok(hipMemcpyAsync(o.d_perm, sorted, nrow * 4, hipMemcpyDeviceToDevice, s),
                    "row order");
        if (ok) {
            launch_gather_rows(s, nrow, o.d_perm, tu, o.d_u);
            launch_gather_rows(s, nrow, o.d_perm, tv, o.d_v);
            launch_gather_rows(s, nrow, o.d_perm, tw, o.d_w);
            launch_pack_sigma(s, nch, nrow, o.d_perm, ts2, tfl, o.d_s2f);
            ok = hip_ok(hipMemsetAsync(o.d_res, 0, res_bytes, s), "residual fill");
        }
        ok = hip_ok(hipGetLastError(), "obs set-up") &&
             hip_ok(hipStreamSynchronize(s), "obs set-up sync") && ok;
    }
    if (!ok) (void)hipStreamSynchronize(s);  // nothing queued may still read what is freed
    dfree(tu);
    dfree(tv);
    dfree(tw);
    dfree(keys);
    dfree(idx);
    dfree(keys2);
    dfree(idx2);
    dfree(ts2);
    dfree(tfl);
    if (tmp) (void)hipFree(tmp);
    if (!ok) {
        delete h;
        return RL2_ERR_HIP;
    }
    *out = h;
    return RL2_OK;
}

int rl2_obs_destroy(rl2_obs *obs)
{
    delete obs;
    return RL2_OK;
}

int rl2_obs_set_vis(rl2_obs *obs, const double *visData, const double *visSkyMod)
{
    if (!obs || !visData) return RL2_ERR_ARG;
    Obs &o = obs->o;
    double *dD = nullptr, *dS = nullptr;
    bool ok = dmalloc(&dD, 2 * o.nvis) && (!visSkyMod || dmalloc(&dS, 2 * o.nvis));
    ok = ok &&
         hip_ok(hipMemcpyAsync(dD, visData, o.nvis * 16, hipMemcpyHostToDevice, g_stream),
                "visibility upload") &&
         (!dS || hip_ok(hipMemcpyAsync(dS, visSkyMod, o.nvis * 16, hipMemcpyHostToDevice,
                                       g_stream),
                        "sky model upload"));
    if (ok) {
        launch_residual_in(g_stream, o.prec, o.nch, o.nrow, o.d_perm, dD, dS, o.d_res);
        ok = hip_ok(hipGetLastError(), "set_vis") && hip_ok(hipStreamSynchronize(g_stream), "sync");
    }
    if (!ok) (void)hipStreamSynchronize(g_stream);  // nothing queued may still read dD, dS
    dfree(dD);
    dfree(dS);
    ok = ok && o.measure_res_bound();
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_get_residual(rl2_obs *obs, double *residual)
{
    if (!obs || !residual) return RL2_ERR_ARG;
    Obs &o = obs->o;
    double *d = nullptr;
    bool ok = dmalloc(&d, 2 * o.nvis);
    if (ok) {
        launch_residual_out(g_stream, o.prec, o.nch, o.nrow, o.d_perm, o.d_res, d);
        ok = hip_ok(hipGetLastError(), "get_residual") &&
             hip_ok(hipMemcpyAsync(residual, d, o.nvis * 16, hipMemcpyDeviceToHost, g_stream),
                    "residual download");
        ok = hip_ok(hipStreamSynchronize(g_stream), "sync") && ok;
    }
    dfree(d);
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_sync(rl2_obs *obs, double *visData, double *visSkyMod)
{
    int st = rl2_obs_get_residual(obs, visData);
    if (st == RL2_OK && visSkyMod) memset(visSkyMod, 0, obs->o.nvis * 16);
    return st;
}

int rl2_obs_add_sources(rl2_obs *obs, unsigned int ngal, const double *flux, const double *scale,
                        const double *l, const double *m, const double *e1, const double *e2,
                        double sign)
{
    if (!obs || !flux || !scale || !l || !m) return RL2_ERR_ARG;
    if (ngal == 0) return RL2_OK;
    Obs &o = obs->o;
    std::vector<GalaxyGeom> g(ngal);
    for (unsigned i = 0; i < ngal; i++)
        g[i] = make_geom(l[i], m[i], flux[i], scale[i], e1 ? e1[i] : 0., e2 ? e2[i] : 0.);
    GalaxyGeom *dg = nullptr;
    if (!dmalloc(&dg, ngal)) return RL2_ERR_HIP;
    bool ok = hip_ok(hipMemcpyAsync(dg, g.data(), ngal * sizeof(GalaxyGeom),
                                    hipMemcpyHostToDevice, g_stream),
                     "source upload");
    if (ok) {
        KTimer t(RL2_K_SOURCES);
        launch_add_sources(g_stream, o.prec, o.nch, o.nrow, o.d_u, o.d_v, o.d_w, o.d_wn, o.d_spec,
                           dg, ngal, sign, o.d_res);
    }
    ok = ok && hip_ok(hipGetLastError(), "add_sources");
    ok = hip_ok(hipStreamSynchronize(g_stream), "sync") && ok;
    dfree(dg);
    ok = ok && o.measure_res_bound();
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_add_noise(rl2_obs *obs, double sigma2, uint64_t seed)
{
    if (!obs || !(sigma2 >= 0.)) return RL2_ERR_ARG;
    Obs &o = obs->o;
    launch_add_noise(g_stream, o.prec, o.nvis, sqrt(sigma2), seed, o.d_res);
    return hip_ok(hipGetLastError(), "add_noise") &&
                   hip_ok(hipStreamSynchronize(g_stream), "sync") && o.measure_res_bound()
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_obs_checkpoint(rl2_obs *obs)
{
    if (!obs) return RL2_ERR_ARG;
    Obs &o = obs->o;
    const size_t bytes = o.nvis * (o.prec == RL2_PREC_FP64 ? 16 : 8);
    if (!o.d_res_saved && !hip_ok(hipMalloc(&o.d_res_saved, bytes), "hipMalloc checkpoint"))
        return RL2_ERR_HIP;
    o.res_bound_saved = o.res_bound;
    return hip_ok(hipMemcpyAsync(o.d_res_saved, o.d_res, bytes, hipMemcpyDeviceToDevice, g_stream),
                  "checkpoint") && hip_ok(hipStreamSynchronize(g_stream), "sync")
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_obs_restore(rl2_obs *obs)
{
    if (!obs) return RL2_ERR_ARG;
    Obs &o = obs->o;
    if (!o.d_res_saved) {
        set_error("rl2_obs_restore: no checkpoint");
        return RL2_ERR_STATE;
    }
    const size_t bytes = o.nvis * (o.prec == RL2_PREC_FP64 ? 16 : 8);
    o.res_bound = o.res_bound_saved;
    return hip_ok(hipMemcpyAsync(o.d_res, o.d_res_saved, bytes, hipMemcpyDeviceToDevice, g_stream),
                  "restore") && hip_ok(hipStreamSynchronize(g_stream), "sync")
               ? RL2_OK
               : RL2_ERR_HIP;
}

int rl2_obs_extract(rl2_obs *obs, int slot, double l0, double m0, double flux, double R_mu,
                    unsigned int facet, double *facet_vis, double *facet_sigma2, double *uu,
                    double *vv, double *sum_w, unsigned int *ncoords)
{
    if (!obs || slot < 0) return RL2_ERR_ARG;
    Obs &o = obs->o;
    if (!obs_extract(o, slot, l0, m0, flux, R_mu, facet, sum_w != nullptr)) return RL2_ERR_HIP;
    unsigned n = 0;
    if (!read_slot_n(o, slot, &n)) return RL2_ERR_HIP;
    FacetSlot &s = o.slots[slot];
    s.n = n;
    if (ncoords) *ncoords = n;
    std::vector<double> re, im;
    auto down = [](void *h, const void *d, size_t bytes) {
        return hip_ok(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_stream), "extract copy");
    };
    bool ok = true;
    if (facet_vis) {
        re.resize(n);
        im.resize(n);
        ok = down(re.data(), s.dre, n * 8) && down(im.data(), s.dim, n * 8);
    }
    if (facet_sigma2) ok = ok && down(facet_sigma2, s.s2, n * 8);
    if (uu) ok = ok && down(uu, s.uu, n * 8);
    if (vv) ok = ok && down(vv, s.vv, n * 8);
    if (sum_w) ok = ok && down(sum_w, s.sumw, (size_t)facet * facet * 8);
    ok = hip_ok(hipStreamSynchronize(g_stream), "extract copy") && ok;
    if (!ok) return RL2_ERR_HIP;
    if (facet_vis)
        for (unsigned i = 0; i < n; i++) {
            facet_vis[2 * i] = re[i];
            facet_vis[2 * i + 1] = im[i];
        }
    return RL2_OK;
}

int rl2_obs_load_facet(rl2_obs *obs, int slot, unsigned int n, const double *facet_vis,
                       const double *facet_sigma2, const double *uu, const double *vv)
{
    if (!obs || slot < 0 || !facet_vis || !facet_sigma2 || !uu || !vv) return RL2_ERR_ARG;
    Obs &o = obs->o;
    if (!o.ensure_slots(slot + 1)) return RL2_ERR_HIP;
    if ((size_t)n > (size_t)o.fmax * o.fmax) {
        set_error("facet larger than the allocated facet buffers");
        return RL2_ERR_RANGE;
    }
    FacetSlot &s = o.slots[slot];
    std::vector<double> re(n), im(n);
    for (unsigned i = 0; i < n; i++) {
        re[i] = facet_vis[2 * i];
        im[i] = facet_vis[2 * i + 1];
    }
    auto up = [](void *d, const void *h, size_t bytes) {
        return hip_ok(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_stream), "load_facet");
    };
    bool ok = up(s.dre, re.data(), n * 8) && up(s.dim, im.data(), n * 8) &&
              up(s.s2, facet_sigma2, n * 8) && up(s.uu, uu, n * 8) && up(s.vv, vv, n * 8) &&
              up(&o.d_facets[slot].n, &n, sizeof n);
    if (ok)
        launch_pack_facet(g_stream, &o.d_facets[slot], n, s.uu, s.vv, s.dre, s.dim, s.s2, s.pk,
                          s.invs2, s.invs2_64, nullptr, 0, s.dc, s.loff);
    FacetDev &f = o.h_facets[slot];  // caller-given cells: no grid structure
    if (ok && (f.cid || f.grid_facet)) {
        f.cid = nullptr;
        f.ctr = nullptr;
        f.grid_facet = 0;
        ok = o.facet_upload(slot);
    }
    ok = hip_ok(hipStreamSynchronize(g_stream), "load_facet") && ok;  // before n and re, im go
    if (!ok) return RL2_ERR_HIP;
    s.n = n;
    return RL2_OK;
}

int rl2_rmarg_eval(int nitems, int numr, const double *ro, const double *rprior,
                   const double *Lr, double *Le, int *err)
{
    if (nitems <= 0 || numr < 2 || numr > kMaxR || !ro || !rprior || !Lr || !Le || !err) {
        set_error("rl2_rmarg_eval: nitems > 0, numr in [2, 32] and every array required");
        return RL2_ERR_ARG;
    }
    if (!ensure_init()) return RL2_ERR_NO_DEVICE;
    FacetDev f;
    memset(&f, 0, sizeof f);
    for (int k = 0; k < kMaxR; k++) f.rprior[k] = k < numr ? rprior[k] : 0.;
    std::vector<LikItem> items(nitems);
    for (auto &it : items) {
        memset(&it, 0, sizeof it);
        it.slot = 0;
    }
    LikItem *d_items = nullptr;
    FacetDev *d_f = nullptr;
    double *d_Lr = nullptr, *d_Le = nullptr;
    int *d_err = nullptr;
    const size_t nlr = (size_t)nitems * kMaxR;
    bool ok = dmalloc(&d_items, nitems) && dmalloc(&d_f, 1) && dmalloc(&d_Lr, nlr) &&
              dmalloc(&d_Le, nitems) && dmalloc(&d_err, nitems) &&
              hip_ok(hipMemcpyAsync(d_items, items.data(), nitems * sizeof(LikItem),
                                    hipMemcpyHostToDevice, g_stream), "rmarg items") &&
              hip_ok(hipMemcpyAsync(d_f, &f, sizeof f, hipMemcpyHostToDevice, g_stream),
                     "rmarg prior") &&
              hip_ok(hipMemcpyAsync(d_Lr, Lr, nlr * sizeof(double), hipMemcpyHostToDevice,
                                    g_stream), "rmarg L_r");
    if (ok) {
        LikLaunch a{};
        a.nitems = nitems;
        a.items = d_items;
        a.facets = d_f;
        a.rg.numr = numr;
        for (int k = 0; k < kMaxR; k++) a.rg.ro[k] = k < numr ? ro[k] : 0.;
        a.Lr = d_Lr;
        a.Le = d_Le;
        a.err = d_err;
        launch_rmarg(g_stream, a);
        ok = hip_ok(hipGetLastError(), "rmarg launch") &&
             hip_ok(hipMemcpyAsync(Le, d_Le, nitems * sizeof(double), hipMemcpyDeviceToHost,
                                   g_stream), "rmarg L_e") &&
             hip_ok(hipMemcpyAsync(err, d_err, nitems * sizeof(int), hipMemcpyDeviceToHost,
                                   g_stream), "rmarg err") &&
             hip_ok(hipStreamSynchronize(g_stream), "rmarg sync");
    }
    dfree(d_items);
    dfree(d_f);
    dfree(d_Lr);
    dfree(d_Le);
    dfree(d_err);
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_marg_position_eval(int n, const double *x, double *out)
{
    if (n <= 0 || !x || !out) {
        set_error("rl2_marg_position_eval: n > 0 and both arrays required");
        return RL2_ERR_ARG;
    }
    if (!ensure_init()) return RL2_ERR_NO_DEVICE;
    double *d_x = nullptr, *d_out = nullptr;
    bool ok = dmalloc(&d_x, n) && dmalloc(&d_out, n) &&
              hip_ok(hipMemcpyAsync(d_x, x, n * sizeof(double), hipMemcpyHostToDevice,
                                    g_stream), "marg_position x");
    if (ok) {
        launch_marg_position(g_stream, n, d_x, d_out);
        ok = hip_ok(hipGetLastError(), "marg_position launch") &&
             hip_ok(hipMemcpyAsync(out, d_out, n * sizeof(double), hipMemcpyDeviceToHost,
                                   g_stream), "marg_position out") &&
             hip_ok(hipStreamSynchronize(g_stream), "marg_position sync");
    }
    if (!ok) (void)hipStreamSynchronize(g_stream);
    dfree(d_x);
    dfree(d_out);
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_set_prior(rl2_obs *obs, int slot, const double *rprior)
{
    if (!obs || slot < 0 || !rprior) return RL2_ERR_ARG;
    Obs &o = obs->o;
    if (!o.ensure_slots(slot + 1)) return RL2_ERR_HIP;
    for (int k = 0; k < kMaxR; k++) o.slots[slot].rprior[k] = (k < o.rg.numr) ? rprior[k] : 0.;
    if (!o.facet_upload(slot)) return RL2_ERR_HIP;
    return hip_ok(hipStreamSynchronize(g_stream), "set_prior") ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_loglik(rl2_obs *obs, int slot, int npts, const double *e1, const double *e2,
                   double *L_e, int *error, double *L_r, int *n_iter)
{
    if (!obs || slot < 0 || npts < 0 || (npts && (!e1 || !e2 || !L_e || !error)))
        return RL2_ERR_ARG;
    Obs &o = obs->o;
    if ((int)o.slots.size() <= slot) return RL2_ERR_STATE;
    std::vector<LikItem> items(npts);
    for (int i = 0; i < npts; i++) items[i] = {slot, o.loglik_prec, e1[i], e2[i]};
    std::vector<double> Lr;
    std::vector<int> nit;
    if (L_r) Lr.resize((size_t)npts * kMaxR);
    if (n_iter) nit.resize((size_t)npts * kMaxR);
    if (!lik_eval(o, items, L_e, error, L_r ? Lr.data() : nullptr, n_iter ? nit.data() : nullptr))
        return RL2_ERR_HIP;
    for (int i = 0; i < npts; i++) error[i] &= ~kErrFragile;
    const int nr = o.rg.numr;
    for (int i = 0; i < npts; i++)
        for (int r = 0; r < nr; r++) {
            if (L_r) L_r[(size_t)i * nr + r] = (r == 0) ? 0. : Lr[(size_t)i * kMaxR + r];
            if (n_iter) n_iter[(size_t)i * nr + r] = (r == 0) ? 0 : nit[(size_t)i * kMaxR + r];
        }
    return RL2_OK;
}

int rl2_obs_loglik_precision(rl2_obs *obs, int prec)
{
    if (!obs || (prec != 0 && prec != 1)) return RL2_ERR_ARG;
    obs->o.loglik_prec = prec;
    return RL2_OK;
}

int rl2_obs_fit(rl2_obs *obs, int slot, const rl2_fit_options *opt, rl2_gal_result *res)
{
    if (!obs || slot < 0 || !res) return RL2_ERR_ARG;
    Obs &o = obs->o;
    if ((int)o.slots.size() <= slot) return RL2_ERR_STATE;
    rl2_fit_options op = {};
    if (opt) op = *opt;
    if (!opt) {
        op.nm_speculate = 1;
        op.line_chunk = 4;
    }
    std::vector<FitRequest> reqs{{slot, res}};
    o.lik_block = op.lik_block > 0 ? op.lik_block : 0;
    bool ok = fit_batch(reqs, op, [&](const std::vector<LikItem> &items, std::vector<double> &Le,
                                      std::vector<int> &err) {
        return lik_eval(o, items, Le.data(), err.data(), nullptr, nullptr);
    });
    unsigned n = 0;
    if (ok && read_slot_n(o, slot, &n)) res->ncoords = n;
    res->facet = o.slots[slot].facet;
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_obs_subtract(rl2_obs *obs, double l0, double m0, double flux, double R_mu, double e1,
                     double e2)
{
    if (!obs) return RL2_ERR_ARG;
    if (!obs_subtract(obs->o, l0, m0, flux, R_mu, e1, e2)) return RL2_ERR_HIP;
    return hip_ok(hipStreamSynchronize(g_stream), "subtract") ? RL2_OK : RL2_ERR_HIP;
}

int rl2_process(rl2_obs *obs, unsigned int ngal, const double *l, const double *m,
                const double *flux, const double *scale, const rl2_fit_options *opt,
                rl2_gal_result *results, rl2_run_stats *stats)
{
    if (!obs || !l || !m || !flux || !scale || !results) return RL2_ERR_ARG;
    rl2_fit_options op = {};
    if (opt)
        op = *opt;
    else {
        op.batch = 1;
        op.do_refit = 1;
        op.nm_speculate = 1;
        op.line_chunk = 4;
    }
    for (unsigned g = 0; g < ngal; g++) memset(&results[g], 0, sizeof(rl2_gal_result));
    return process(obs->o, ngal, l, m, flux, scale, op, results, stats) ? RL2_OK : RL2_ERR_HIP;
}

// Coordinate search over the launch knobs that leave results unchanged up to rounding:
// every candidate value of one knob is timed on the same galaxies from the same residual
// (the first pass of the loop with the caller's options), the fastest is kept, then the
// next knob.  The residual is put back afterwards.
int rl2_autotune(rl2_obs *obs, unsigned int ngal, const double *l, const double *m,
                 const double *flux, const double *scale, const rl2_fit_options *opt,
                 rl2_tune_trial *trials, int max_trials, int *ntrials, int *lik_block)
{
    if (!obs || !l || !m || !flux || !scale || !opt || ngal == 0) return RL2_ERR_ARG;
    Obs &o = obs->o;
    const size_t bytes = o.nvis * (o.prec == RL2_PREC_FP64 ? 16 : 8);
    void *saved = nullptr;
    if (!hip_ok(hipMalloc(&saved, bytes), "hipMalloc autotune residual")) return RL2_ERR_HIP;
    const double bound0 = o.res_bound;
    bool ok = hip_ok(hipMemcpyAsync(saved, o.d_res, bytes, hipMemcpyDeviceToDevice, g_stream),
                     "autotune save");
    rl2_fit_options op = *opt;
    op.do_refit = 0;
    op.features_out = nullptr;
    op.labels_out = nullptr;
    const int B = op.batch > 0 ? op.batch : 1;
    const bool batched = B > 1 && B <= kMaxBatch && tuning().batch_extract;
    std::vector<rl2_gal_result> res(ngal);
    std::vector<unsigned> todo(ngal);
    for (unsigned g = 0; g < ngal; g++) todo[g] = g;
    int nt = 0;
    auto trial = [&](int knob, int value) -> double {
        if (!ok) return 0.;
        ok = hip_ok(hipMemcpyAsync(o.d_res, saved, bytes, hipMemcpyDeviceToDevice, g_stream),
                    "autotune restore") &&
             hip_ok(hipStreamSynchronize(g_stream), "sync");
        o.res_bound = bound0;
        double te = 0., tf = 0., ts = 0.;
        const double t0 = now_s();
        ok = ok && process_list(o, todo, 0, l, m, flux, scale, op, res.data(), nullptr, &te, &tf,
                                &ts, nullptr) &&
             hip_ok(hipStreamSynchronize(g_stream), "sync");
        const double dt = now_s() - t0;
        if (trials && nt < max_trials) trials[nt] = {knob, value, dt};
        nt++;
        return dt;
    };
    struct Knob {
        int id;
        std::vector<int> values;
    };
    std::vector<Knob> knobs;
    if (batched)
        knobs.push_back({RL2_TUNE_BATCH_GROUPS, {(int)num_cus(), 2 * (int)num_cus(),
                                                 3 * (int)num_cus(), 4 * (int)num_cus()}});
    else
        knobs.push_back({RL2_TUNE_GRID_VARIANT, {5, 3, 6, 4, 2, 1, 0}});
    knobs.push_back({RL2_TUNE_ITEM_GROUPS, {0, 256, 1024}});
    knobs.push_back({RL2_TUNE_TILE_CELLS, {0, 2048, 4096, 8192, 16384}});
    knobs.push_back({RL2_TUNE_TILE_STREAMS, {1, 2}});
    knobs.push_back({-1, {0, 256, 1024}});  // work-group threads per (e, r) task; 0: tiled
    for (const Knob &k : knobs) {
        int best = k.values[0];
        double tbest = 1e300;
        for (int v : k.values) {
            if (k.id >= 0 && rl2_tune_set(k.id, v) != RL2_OK) continue;
            if (k.id < 0) op.lik_block = v;
            const double dt = trial(k.id, v);
            if (ok && dt < tbest) {
                tbest = dt;
                best = v;
            }
        }
        if (k.id >= 0)
            rl2_tune_set(k.id, best);
        else
            op.lik_block = best;
    }
    ok = ok && hip_ok(hipMemcpyAsync(o.d_res, saved, bytes, hipMemcpyDeviceToDevice, g_stream),
                      "autotune restore") &&
         hip_ok(hipStreamSynchronize(g_stream), "sync");
    o.res_bound = bound0;
    (void)hipFree(saved);
    if (ntrials) *ntrials = nt;
    if (lik_block) *lik_block = op.lik_block;
    return ok ? RL2_OK : RL2_ERR_HIP;
}

int rl2_stats_enable(int on)
{
    if (!ensure_init()) return RL2_ERR_NO_DEVICE;
    if (!hip_ok(hipStreamSynchronize(g_stream), "stats sync")) return RL2_ERR_HIP;
    drain_stats();
    g_stats_on = on != 0;
    return RL2_OK;
}

int rl2_stats_reset(void)
{
    if (g_stream && !hip_ok(hipStreamSynchronize(g_stream), "stats sync")) return RL2_ERR_HIP;
    drain_stats();
    for (auto &k : g_stats) {
        k.ms = 0.;
        k.launches = 0;
        k.flops = 0.;
        k.bytes = 0.;
    }
    return RL2_OK;
}

int rl2_stats_get(int kernel, rl2_kernel_stat *out)
{
    if (kernel < 0 || kernel >= RL2_K_COUNT || !out) return RL2_ERR_ARG;
    if (g_stream && !hip_ok(hipStreamSynchronize(g_stream), "stats sync")) return RL2_ERR_HIP;
    drain_stats();
    out->ms = g_stats[kernel].ms;
    out->launches = g_stats[kernel].launches;
    out->flops = g_stats[kernel].flops;
    out->bytes = g_stats[kernel].bytes;
    return RL2_OK;
}

int rl2_synchronize(void)
{
    if (!g_stream) return RL2_OK;
    return hip_ok(hipStreamSynchronize(g_stream), "synchronize") ? RL2_OK : RL2_ERR_HIP;
}

int rl2_tune_set(int knob, int value)
{
    switch (knob) {
    case RL2_TUNE_GRID_VARIANT:
        if (value < 0 || value > 6) break;
        tuning().grid_variant = value;
        return RL2_OK;
    case RL2_TUNE_TILE_CELLS:
        if (value != 0 && (value < (int)kTileKnobMinCells || value > (int)kTileMaxCells ||
                           (value & (value - 1))))
            break;
        tuning().tile_cells = (unsigned)value;
        return RL2_OK;
    case RL2_TUNE_PHASE_TABLES:
        if (value < 0 || value > 2) break;
        tuning().phase_tables = value;
        return RL2_OK;
    case RL2_TUNE_FUSE_SUBTRACT:
        if (value != 0 && value != 1) break;
        tuning().fuse_subtract = value;
        return RL2_OK;
    case RL2_TUNE_TILE_BLOCKS:
        if (value < 0 || value % 8 || value > (1 << 20)) break;
        tuning().tile_blocks = (unsigned)value;
        return RL2_OK;
    case RL2_TUNE_BATCH_EXTRACT:
        if (value < 0 || value > 1) break;
        tuning().batch_extract = value;
        return RL2_OK;
    case RL2_TUNE_BATCH_GROUPS:
        if (value < 0 || value > 65536) break;
        tuning().batch_groups = (unsigned)value;
        return RL2_OK;
    case RL2_TUNE_BATCH_GRAPH:
        if (value < 0 || value > 1) break;
        tuning().batch_graph = value;
        return RL2_OK;
    case RL2_TUNE_ITEM_GROUPS:
        if (value < 0 || value > 65536) break;
        tuning().item_groups = (unsigned)value;
        return RL2_OK;
    case RL2_TUNE_FUSE_STEP:
        if (value < 0 || value > 2) break;
        tuning().fuse_step = value;
        return RL2_OK;
    case RL2_TUNE_TILE_PASSES:
        if (value < 0 || value > 50) break;
        tuning().tile_passes = value;
        return RL2_OK;
    case RL2_TUNE_P0_TASKS:
        if (value < 1 || value > 3) break;
        tuning().p0_tasks = value;
        return RL2_OK;
    case RL2_TUNE_TILE_STREAMS:
        if (value != 1 && value != 2) break;
        tuning().tile_streams = value;
        return RL2_OK;
    default:
        break;
    }
    set_error("rl2_tune_set: knob " + std::to_string(knob) + " does not take " +
              std::to_string(value));
    return RL2_ERR_ARG;
}

int rl2_tune_get(int knob, int *value)
{
    if (!value) return RL2_ERR_ARG;
    switch (knob) {
    case RL2_TUNE_GRID_VARIANT: *value = tuning().grid_variant; return RL2_OK;
    case RL2_TUNE_TILE_CELLS: *value = (int)tuning().tile_cells; return RL2_OK;
    case RL2_TUNE_PHASE_TABLES: *value = tuning().phase_tables; return RL2_OK;
    case RL2_TUNE_FUSE_SUBTRACT: *value = tuning().fuse_subtract; return RL2_OK;
    case RL2_TUNE_TILE_BLOCKS: *value = (int)tuning().tile_blocks; return RL2_OK;
    case RL2_TUNE_TILE_STREAMS: *value = tuning().tile_streams; return RL2_OK;
    case RL2_TUNE_BATCH_EXTRACT: *value = tuning().batch_extract; return RL2_OK;
    case RL2_TUNE_BATCH_GROUPS: *value = (int)tuning().batch_groups; return RL2_OK;
    case RL2_TUNE_BATCH_GRAPH: *value = tuning().batch_graph; return RL2_OK;
    case RL2_TUNE_ITEM_GROUPS: *value = (int)tuning().item_groups; return RL2_OK;
    case RL2_TUNE_FUSE_STEP: *value = tuning().fuse_step; return RL2_OK;
    case RL2_TUNE_TILE_PASSES: *value = tuning().tile_passes; return RL2_OK;
    case RL2_TUNE_P0_TASKS: *value = tuning().p0_tasks; return RL2_OK;
    default: set_error("rl2_tune_get: unknown knob"); return RL2_ERR_ARG;
    }
}

}  // extern "C"
