// rl2_dropin.cpp -- RadioLensfit2's per-galaxy functions with their reference signatures
// (include/rl2_reference.h), implemented on the GPU through the rl2 runtime.
//
// The observation arrays passed to source_extraction/data_processing are uploaded once per
// (uu_metres, num_coords, nchannels, len) and kept in HBM; the visibilities are re-read from
// the host arrays on every call so callers that update them on the host stay consistent.
// Likelihood-level functions upload the facet they are given unless it is the facet the
// last source_extraction produced, which is still resident.
#include <hip/hip_runtime.h>

#include <math.h>
#include <memory>
#include <string.h>
#include <vector>

#include "rl2_obs.h"
#include "rl2_reference.h"

using namespace rl2;

namespace {

// ---------------------------------------------------------------- small device buffers
// A device array of one call.  A failed allocation leaves p null and a failed transfer
// returns false, each with the library's error message set: callers check `allocated(...)`
// and the transfers before a launch takes the pointers.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    explicit DBuf(size_t count)
    {
        n = count ? count : 1;
        if (!hip_ok(hipMalloc((void **)&p, n * sizeof(T)), "hipMalloc")) p = nullptr;
    }
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    bool up(const T *h, size_t count)
    {
        return p && hip_ok(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice,
                                          stream()),
                           "upload");
    }
    bool down(T *h, size_t count) const
    {
        return p && hip_ok(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost,
                                          stream()),
                           "download");
    }
};

template <class... B>
bool allocated(const B &...b)
{
    return ((b.p != nullptr) && ...);
}

bool sync_ok() { return hip_ok(hipStreamSynchronize(stream()), "hipStreamSynchronize"); }

// Queues a device-to-host copy on the library's stream.
bool copy_down(void *h, const void *d, size_t bytes)
{
    return hip_ok(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, stream()), "download");
}

void report(const char *fn)
{
    fprintf(stderr, "librl2hip %s: %s\n", fn, rl2_last_error());
}

// ---------------------------------------------------------------- observation cache
struct ObsKey {
    const double *uu = nullptr;
    unsigned long nrow = 0;
    unsigned nch = 0;
    double len = 0.;
};
ObsKey g_obs_key;
std::unique_ptr<rl2_obs, int (*)(rl2_obs *)> g_obs(nullptr, rl2_obs_destroy);

Obs *observation(unsigned nch, unsigned long nrow, const double *u, const double *v,
                 const double *w, const double *wn, const double *spec, const float *sigma2,
                 const bool *flag, double len)
{
    if (g_obs && g_obs_key.uu == u && g_obs_key.nrow == nrow && g_obs_key.nch == nch &&
        g_obs_key.len == len)
        return &g_obs->o;
    g_obs.reset();
    rl2_obs *h = nullptr;
    if (rl2_obs_create(&h, nch, nrow, u, v, w, wn, spec, sigma2, (const unsigned char *)flag, len,
                       RL2_PREC_FP64) != RL2_OK)
        return nullptr;
    g_obs.reset(h);
    g_obs_key = {u, nrow, nch, len};
    return &g_obs->o;
}

// What the last source_extraction left resident in slot 0 of the cached observation.
struct LastFacet {
    const complexd *data = nullptr;
    const double *sigma2 = nullptr, *sum_w = nullptr, *uu = nullptr, *vv = nullptr;
    unsigned facet = 0, n = 0;
} g_last;

// Facet-only context for likelihood-level calls on host-supplied facets.
std::unique_ptr<rl2_obs, int (*)(rl2_obs *)> g_fctx(nullptr, rl2_obs_destroy);

Obs *facet_context(size_t n_needed)
{
    if (!ensure_device()) return nullptr;
    if (!g_fctx || (size_t)g_fctx->o.fmax * g_fctx->o.fmax < n_needed) {
        g_fctx.reset();
        rl2_obs *h = new rl2_obs;
        unsigned f = 64;
        while ((size_t)f * f < n_needed) f *= 2;
        h->o.fmax = f;
        h->o.len = 1.;
        h->o.rg.numr = rl2_r_grid(h->o.rg.ro);
        g_fctx.reset(h);
    }
    return &g_fctx->o;
}

// The scalelength grid of a call: ro[0..numr-1] (ro[0] unused, as in the reference).  The
// device buffers hold kMaxR points per request, so a larger grid is refused.
bool set_grid(Obs &o, int numr, const double *ro)
{
    if (numr < 2 || numr > kMaxR || !ro) {
        set_error("likelihood_params: numr must be in [2, " + std::to_string(kMaxR) +
                  "] with ro given (got " + std::to_string(numr) + ")");
        return false;
    }
    o.rg.numr = numr;
    for (int k = 0; k < kMaxR; k++) o.rg.ro[k] = (k < numr) ? ro[k] : 0.;
    return true;
}

// Facet slot holding par's facet: the resident extraction result when par points at it,
// otherwise an upload into the facet-only context.
bool facet_for(const likelihood_params *par, Obs **out, int *slot)
{
    const unsigned n = par->ncoords;
    if (g_obs && g_last.data == par->data && g_last.sigma2 == par->sigma2 &&
        g_last.uu == par->uu && g_last.vv == par->vv && g_last.n == n) {
        *out = &g_obs->o;
        *slot = 0;
    } else {
        Obs *o = facet_context(n);
        if (!o) return false;
        if (rl2_obs_load_facet(g_fctx.get(), 0, n, (const double *)par->data, par->sigma2, par->uu,
                               par->vv) != RL2_OK)
            return false;
        *out = o;
        *slot = 0;
    }
    Obs &o = **out;
    if (!set_grid(o, par->numr, par->ro) || !o.ensure_slots(1)) return false;
    for (int k = 0; k < kMaxR; k++) o.slots[0].rprior[k] = (k < par->numr) ? par->rprior[k] : 0.;
    o.slots[0].n = n;
    return o.facet_upload(0);
}

// Last arrays given to marginalise_posterior_r, for marf().
Spline g_marf_spline = {};

}  // namespace

extern "C" {

// ---------------------------------------------------------------- scalar helpers
double gaussian_shape(double k1, double k2, double scale_factor, double wave_factor)
{
    return exp(-0.5 * scale_factor * wave_factor * (k1 * k1 + k2 * k2));
}
double sersic_shape(double k1, double k2, double scale_factor, double wave_factor)
{
    return rl2::sersic_shape(k1, k2, scale_factor, wave_factor);
}
unsigned int facet_size(double theta_med, double len) { return rl2::facet_size(theta_med, len); }
double scale_mean(double flux) { return rl2::scale_mean(flux); }
double rfunc(const double mu, const double sigma, double r)
{
    if (r <= 0.) return NAN;
    return rl2::rfunc(mu, sigma, r);
}
double marginalise_over_position_shift(double x) { return rl2::marg_position(x); }

void set_posterior_values(int numR, double *L_r, double *rprior, double *Ro, double *xmarvals,
                          double *ymarvals, int *numvals)
{
    // marginalise_r.c:29-64
    int num = 0;
    int nRo = 1;
    while (nRo < numR && (L_r[nRo] <= kBad)) nRo++;
    if (nRo < numR)
        for (nRo = 1; nRo < numR; nRo++)
            if (L_r[nRo] > kBad) {
                xmarvals[num] = Ro[nRo];
                ymarvals[num] = L_r[nRo] + rprior[nRo];
                num++;
            }
    *numvals = num;
}

double marf(double x, void *)
{
    if (g_marf_spline.n < 2) return 0.;
    return spline_exp(g_marf_spline, x);
}

// ---------------------------------------------------------------- array functions (GPU)
double marginalise_posterior_r(int num_marvals, double *xmarvals, double *ymarvals)
{
    if (!ensure_device() || num_marvals < 3 || num_marvals > kMaxR) return kBad;
    DBuf<double> dx(num_marvals), dy(num_marvals), dout(num_marvals + 1);
    DBuf<double> ws(kQagDoubles);
    DBuf<int> iws(kQagInts);
    if (!allocated(dx, dy, dout, ws, iws) || !dx.up(xmarvals, num_marvals) ||
        !dy.up(ymarvals, num_marvals)) {
        report("marginalise_posterior_r");
        return kBad;
    }
    launch_marginalise_r(stream(), num_marvals, dx.p, dy.p, dout.p, ws.p, iws.p);
    std::vector<double> out(num_marvals + 1);
    if (!dout.down(out.data(), num_marvals + 1) || !sync_ok()) {
        report("marginalise_posterior_r");
        return kBad;
    }
    for (int i = 0; i < num_marvals; i++) ymarvals[i] = out[1 + i];
    spline_init(g_marf_spline, xmarvals, ymarvals, num_marvals, 0);
    return out[0];
}

void model_galaxy_visibilities_at_zero(double e1, double e2, double scale, unsigned int num_coords,
                                       double *grid_u, double *grid_v, const double *sigma2,
                                       double *Modvis)
{
    if (!ensure_device() || num_coords == 0) return;
    DBuf<double> du(num_coords), dv(num_coords), ds(num_coords), dm(num_coords),
        part((num_coords + 255) / 256);
    if (!allocated(du, dv, ds, dm, part) || !du.up(grid_u, num_coords) ||
        !dv.up(grid_v, num_coords) || !ds.up(sigma2, num_coords)) {
        report("model_galaxy_visibilities_at_zero");
        return;
    }
    launch_model_at_zero(stream(), e1, e2, scale, num_coords, du.p, dv.p, ds.p, dm.p, part.p);
    if (!dm.down(Modvis, num_coords) || !sync_ok()) report("model_galaxy_visibilities_at_zero");
}

void data_galaxy_visibilities(double spectra, double wavenumber, double, double, double e1,
                              double e2, double scale, double flux, double l, double m,
                              unsigned int num_coords, double *uu_metres, double *vv_metres,
                              double *ww_metres, complexd *vis)
{
    if (!ensure_device() || num_coords == 0) return;
    DBuf<double> du(num_coords), dv(num_coords), dw(num_coords), dvis(2 * (size_t)num_coords);
    if (!allocated(du, dv, dw, dvis) || !du.up(uu_metres, num_coords) ||
        !dv.up(vv_metres, num_coords) || !dw.up(ww_metres, num_coords)) {
        report("data_galaxy_visibilities");
        return;
    }
    GalaxyGeom g = make_geom(l, m, flux, scale, e1, e2);
    launch_galaxy_vis(stream(), spectra, wavenumber, g, num_coords, du.p, dv.p, dw.p, dvis.p);
    if (!dvis.down((double *)vis, 2 * (size_t)num_coords) || !sync_ok())
        report("data_galaxy_visibilities");
}

void data_visibilities_phase_shift(double wavenumber, double l, double m, unsigned int num_coords,
                                   double *uu_metres, double *vv_metres, double *ww_metres,
                                   complexd *vis)
{
    if (!ensure_device() || num_coords == 0) return;
    DBuf<double> du(num_coords), dv(num_coords), dw(num_coords), dvis(2 * (size_t)num_coords);
    if (!allocated(du, dv, dw, dvis) || !du.up(uu_metres, num_coords) ||
        !dv.up(vv_metres, num_coords) || !dw.up(ww_metres, num_coords) ||
        !dvis.up((const double *)vis, 2 * (size_t)num_coords)) {
        report("data_visibilities_phase_shift");
        return;
    }
    launch_phase_shift(stream(), wavenumber, l, m, num_coords, du.p, dv.p, dw.p, dvis.p);
    if (!dvis.down((double *)vis, 2 * (size_t)num_coords) || !sync_ok())
        report("data_visibilities_phase_shift");
}

// Shared by the three gridding entry points: grid (mode 1/2) natural-order rows.
static unsigned grid_natural(int mode, const double *wavenumbers, unsigned nch, unsigned ncoords,
                             const double *u, const double *v, const complexd *vis,
                             const float *sigma2, const bool *flag, double len, unsigned sizeg,
                             complexd *new_vis, double *new_sigma2, double *sum_w, double *gu,
                             double *gv)
{
    if (!ensure_device()) return 0;
    const size_t nvis = (size_t)nch * ncoords, ncell = (size_t)sizeg * sizeg;
    std::vector<uint32_t> iota(ncoords);
    for (unsigned i = 0; i < ncoords; i++) iota[i] = i;
    std::vector<double> invl(nch), spec(nch, 1.);
    for (unsigned c = 0; c < nch; c++) invl[c] = wavenumbers[c] / (2. * kPi);
    std::vector<float> s2(nvis);
    for (size_t i = 0; i < nvis; i++) s2[i] = sigma2 ? sigma2[i] : 1.f;
    DBuf<double> du(ncoords), dv(ncoords), dwn(nch), dinv(nch), dsp(nch), dres(2 * nvis);
    DBuf<uint32_t> dperm(ncoords);
    DBuf<float> ds2(nvis), ds2f(nvis);
    DBuf<unsigned char> dfl(flag ? nvis : 1);
    DBuf<double> acc(grid_acc_words(sizeg)), cu(ncell), cv(ncell), cre(ncell), cim(ncell),
        cs2(ncell), csw(ncell);
    DBuf<unsigned> scan(compact_scratch_words(sizeg) + 1), dn(1);
    if (!allocated(du, dv, dwn, dinv, dsp, dres, dperm, ds2, ds2f, dfl, acc, cu, cv, cre, cim,
                   cs2, csw, scan, dn) ||
        !du.up(u, ncoords) || !dv.up(v, ncoords) || !dwn.up(wavenumbers, nch) ||
        !dinv.up(invl.data(), nch) || !dsp.up(spec.data(), nch) ||
        !dperm.up(iota.data(), ncoords) || !ds2.up(s2.data(), nvis) ||
        (flag && !dfl.up((const unsigned char *)flag, nvis)) ||
        (mode == 1 && !dres.up((const double *)vis, 2 * nvis)) ||
        !hip_ok(hipMemsetAsync(acc.p, 0, grid_acc_words(sizeg) * sizeof(double), stream()),
                "hipMemsetAsync")) {
        report("gridding");
        return 0;
    }
    // scales of the integer accumulators from the largest gridded value and variance
    double vb = 0., sb = 0.;
    if (mode == 1)
        for (size_t i = 0; i < nvis; i++)
            vb = fmax(vb, fmax(fabs(vis[i].real), fabs(vis[i].imag)));
    for (size_t i = 0; i < nvis; i++) sb = fmax(sb, (double)fabsf(s2[i]));
    const AccScale sc = fx_scales(vb, sb, -1, nch);
    launch_pack_sigma(stream(), nch, ncoords, dperm.p, ds2.p, flag ? dfl.p : nullptr, ds2f.p);
    GridLaunch a = {};
    a.mode = mode;
    a.prec = 0;
    a.nch = nch;
    a.nrow = ncoords;
    a.u = du.p;
    a.v = dv.p;
    a.w = du.p;
    a.wn = dwn.p;
    a.inv_lambda = dinv.p;
    a.spec = dsp.p;
    a.res = dres.p;
    a.s2f = ds2f.p;
    a.len = len;
    a.inc = 2 * len / sizeg;
    a.facet = sizeg;
    a.acc = acc.p;
    a.fx_scale = sc.fx_scale;
    a.fs_scale = sc.fs_scale;
    launch_grid(stream(), a);
    launch_compact(stream(), acc.p, sizeg, len, cu.p, cv.p, cre.p, cim.p, cs2.p, nullptr, nullptr,
                   dn.p, csw.p, scan.p, &sc);
    unsigned n = 0;
    if (!dn.down(&n, 1) || !sync_ok()) {
        report("gridding");
        return 0;
    }
    std::vector<double> re(n), im(n);
    bool ok = true;
    if (new_vis) ok = cre.down(re.data(), n) && cim.down(im.data(), n);
    if (new_sigma2) ok = ok && cs2.down(new_sigma2, n);
    if (sum_w) ok = ok && csw.down(sum_w, ncell);
    if (gu) ok = ok && cu.down(gu, n);
    if (gv) ok = ok && cv.down(gv, n);
    if (!sync_ok() || !ok) {
        report("gridding copy");
        return 0;
    }
    if (new_vis)
        for (unsigned i = 0; i < n; i++) {
            new_vis[i].real = re[i];
            new_vis[i].imag = im[i];
        }
    return n;
}

unsigned int evaluate_uv_grid_size(int, int, double len, double *wavenumbers,
                                   unsigned int num_channels, unsigned int ncoords, double *u,
                                   double *v, unsigned int sizeg, bool *flag)
{
    return grid_natural(2, wavenumbers, num_channels, ncoords, u, v, nullptr, nullptr, flag, len,
                        sizeg, nullptr, nullptr, nullptr, nullptr, nullptr);
}

void gridding_visibilities(double *wavenumbers, unsigned int num_channels, unsigned int ncoords,
                           double *u, double *v, complexd *vis, float *sigma2, double len,
                           unsigned int sizeg, complexd *new_vis, double *new_sigma2, bool *flag,
                           double *sum_w)
{
    grid_natural(1, wavenumbers, num_channels, ncoords, u, v, vis, sigma2, flag, len, sizeg,
                 new_vis, new_sigma2, sum_w, nullptr, nullptr);
}

// Compaction of host-side accumulators (evaluate_facet_coords / average_facets).
static unsigned compact_host(unsigned sizeg, double len, const double *re, const double *im,
                             const double *s2, const double *w, double *gu, double *gv,
                             complexd *out_vis, double *out_s2)
{
    if (!ensure_device()) return 0;
    const size_t ncell = (size_t)sizeg * sizeg;
    std::vector<double> acc(4 * ncell, 0.);
    for (size_t p = 0; p < ncell; p++) {
        if (re) acc[p] = re[p];
        if (im) acc[ncell + p] = im[p];
        acc[2 * ncell + p] = s2 ? s2[p] : 1.;
        acc[3 * ncell + p] = w[p];
    }
    DBuf<double> dacc(4 * ncell), cu(ncell), cv(ncell), cre(ncell), cim(ncell), cs2(ncell);
    DBuf<unsigned> scan(compact_scratch_words(sizeg) + 1), dn(1);
    if (!allocated(dacc, cu, cv, cre, cim, cs2, scan, dn) || !dacc.up(acc.data(), 4 * ncell))
        return 0;
    launch_compact(stream(), dacc.p, sizeg, len, cu.p, cv.p, cre.p, cim.p, cs2.p, nullptr, nullptr,
                   dn.p, nullptr, scan.p, nullptr);
    unsigned n = 0;
    if (!dn.down(&n, 1) || !sync_ok()) return 0;
    std::vector<double> vr(n), vi(n);
    bool ok = true;
    if (gu) ok = cu.down(gu, n);
    if (gv) ok = ok && cv.down(gv, n);
    if (out_vis) ok = ok && cre.down(vr.data(), n) && cim.down(vi.data(), n);
    if (out_s2) ok = ok && cs2.down(out_s2, n);
    if (!sync_ok() || !ok) return 0;
    if (out_vis)
        for (unsigned i = 0; i < n; i++) {
            out_vis[i].real = vr[i];
            out_vis[i].imag = vi[i];
        }
    return n;
}

unsigned int evaluate_facet_coords(double *grid_u, double *grid_v, double len, unsigned int sizeg,
                                   double *count_w)
{
    if (g_obs && g_last.sum_w == count_w && g_last.facet == sizeg) {
        Obs &o = g_obs->o;
        FacetSlot &s = o.slots[0];
        if (!copy_down(grid_u, s.uu, g_last.n * 8) || !copy_down(grid_v, s.vv, g_last.n * 8) ||
            !sync_ok()) {
            report("evaluate_facet_coords");
            return 0;
        }
        g_last.uu = grid_u;
        g_last.vv = grid_v;
        return g_last.n;
    }
    return compact_host(sizeg, len, nullptr, nullptr, nullptr, count_w, grid_u, grid_v, nullptr,
                        nullptr);
}

void average_facets(unsigned long int size, complexd *grid_vis, double *grid_sigma2, double *sum_w)
{
    unsigned f = (unsigned)llround(sqrt((double)size));
    if ((unsigned long)f * f != size) {
        fprintf(stderr, "librl2hip average_facets: size %lu is not a square facet\n", size);
        return;
    }
    std::vector<double> re(size), im(size);
    for (unsigned long p = 0; p < size; p++) {
        re[p] = grid_vis[p].real;
        im[p] = grid_vis[p].imag;
    }
    compact_host(f, 1., re.data(), im.data(), grid_sigma2, sum_w, nullptr, nullptr, grid_vis,
                 grid_sigma2);
}

void source_extraction(int, unsigned int facet, likelihood_params *par, complexd *facet_vis,
                       double *facet_sigma2, double *sum_w, double l0, double m0, double flux,
                       double mu, double e1, double e2, complexd *visSkyMod, complexd *visData,
                       complexd *, float *sigma2_vis, bool *flag, unsigned long int num_coords,
                       double *uu_metres, double *vv_metres, double *ww_metres, double len)
{
    Obs *o = observation(par->nchannels, num_coords, uu_metres, vv_metres, ww_metres,
                         par->wavenumbers, par->spec, sigma2_vis, flag, len);
    if (!o) {
        report("source_extraction");
        return;
    }
    if (rl2_obs_set_vis(g_obs.get(), (const double *)visData, (const double *)visSkyMod) != RL2_OK ||
        !o->ensure_slots(1) || !obs_extract(*o, 0, l0, m0, flux, mu, facet, true, e1, e2)) {
        report("source_extraction");
        return;
    }
    unsigned n = 0;
    if (!copy_down(&n, &o->d_facets[0].n, sizeof n) || !sync_ok()) {
        report("source_extraction");
        return;
    }
    FacetSlot &s = o->slots[0];
    s.n = n;
    s.facet = facet;
    std::vector<double> re(n), im(n);
    if (!copy_down(re.data(), s.dre, n * 8) || !copy_down(im.data(), s.dim, n * 8) ||
        !copy_down(facet_sigma2, s.s2, n * 8) ||
        !copy_down(sum_w, s.sumw, (size_t)facet * facet * 8) || !sync_ok()) {
        report("source_extraction");
        return;
    }
    for (unsigned i = 0; i < n; i++) {
        facet_vis[i].real = re[i];
        facet_vis[i].imag = im[i];
    }
    g_last = {facet_vis, facet_sigma2, sum_w, nullptr, nullptr, facet, n};
}

double loglikelihood(void *params, double ee1, double ee2, int *error)
{
    likelihood_params *par = (likelihood_params *)params;
    Obs *o = nullptr;
    int slot = 0;
    *error = 1;
    if (!facet_for(par, &o, &slot)) {
        report("loglikelihood");
        return kBad;
    }
    std::vector<LikItem> items{{slot, 0, ee1, ee2}};
    double L = kBad;
    int err = 1;
    if (!lik_eval(*o, items, &L, &err, nullptr, nullptr)) {
        report("loglikelihood");
        return kBad;
    }
    *error = err;
    return L;
}

double f_likelihood(const gsl_vector *v, void *params)
{
    const double ee1 = v->data[0], ee2 = v->data[v->stride];
    int error = 0;
    double L_e;
    if (ee1 * ee1 + ee2 * ee2 < 0.65)
        L_e = loglikelihood(params, ee1, ee2, &error);
    else
        L_e = -1.e+10;
    return -L_e;
}

double loglikelihood_r(double ee1, double ee2, double scale, unsigned long int n_uv_coords,
                       const double *variance, double *grid_u, double *grid_v,
                       complexd *visData, double *visM)
{
    Obs *o = facet_context(n_uv_coords);
    if (!o || rl2_obs_load_facet(g_fctx.get(), 0, (unsigned)n_uv_coords, (const double *)visData,
                                 variance, grid_u, grid_v) != RL2_OK) {
        report("loglikelihood_r");
        return kBad;
    }
    const double ro[2] = {0., scale};
    set_grid(*o, 2, ro);  // numr 2: always accepted
    std::vector<LikItem> items{{0, 0, ee1, ee2}};
    double Le;
    int err;
    std::vector<double> Lr(kMaxR);
    std::vector<int> nit(kMaxR);
    if (!lik_eval(*o, items, &Le, &err, Lr.data(), nit.data())) {
        report("loglikelihood_r");
        return kBad;
    }
    if (visM) model_galaxy_visibilities_at_zero(ee1, ee2, scale, (unsigned)n_uv_coords, grid_u,
                                                grid_v, variance, visM);
    return Lr[1];
}

int cross_correlation(unsigned long int n_uv_coords, double *grid_u, double *grid_v,
                      complexd *visData, double *visMod, double *ho, double *det_sigma)
{
    std::vector<double> ones(n_uv_coords, 1.);
    Obs *o = facet_context(n_uv_coords);
    if (!o || rl2_obs_load_facet(g_fctx.get(), 0, (unsigned)n_uv_coords, (const double *)visData,
                                 ones.data(), grid_u, grid_v) != RL2_OK) {
        report("cross_correlation");
        return 1;
    }
    DBuf<double> dmod(n_uv_coords), dho(kMaxR), ddet(kMaxR);
    if (!allocated(dmod, dho, ddet) || !dmod.up(visMod, n_uv_coords)) {
        report("cross_correlation");
        return 1;
    }
    const double ro[2] = {0., 1.};
    set_grid(*o, 2, ro);  // numr 2: always accepted
    LikBuffers &b = o->lik;
    const LikItem it{0, 0, 0., 0.};
    o->h_facets[0].mod = dmod.p;
    bool ok = o->facet_upload(0) && b.ensure(1) &&
              hip_ok(hipMemcpyAsync(b.d_items, &it, sizeof it, hipMemcpyHostToDevice, stream()),
                     "upload");
    double h[kMaxR], d[kMaxR];
    if (ok) {
        LikLaunch a = {};
        a.nitems = 1;
        a.items = b.d_items;
        a.facets = o->d_facets;
        a.rg = o->rg;
        a.Lr = b.d_Lr;
        a.nit = b.d_nit;
        a.Le = b.d_Le;
        a.err = b.d_err;
        a.qag_ws = b.d_ws;
        a.qag_iws = b.d_iws;
        a.block = 256;
        a.ho = dho.p;
        a.det = ddet.p;
        launch_likelihood(stream(), a);
        ok = dho.down(h, kMaxR) && ddet.down(d, kMaxR);
    }
    ok = sync_ok() && ok;  // also before dmod is freed: nothing may still read it
    o->h_facets[0].mod = nullptr;
    ok = o->facet_upload(0) && ok;
    if (!ok) {
        report("cross_correlation");
        return 1;
    }
    *ho = h[1];
    *det_sigma = d[1];
    return d[1] <= 0 ? 1 : 0;
}

void likelihood_sampling(double *mes_e1, double *mes_e2, double maxL, void *params, int np_max,
                         double *var_e1, double *var_e2, double *oneDimvar)
{
    likelihood_params *par = (likelihood_params *)params;
    Obs *o = nullptr;
    int slot = 0;
    if (!facet_for(par, &o, &slot)) {
        report("likelihood_sampling");
        return;
    }
    rl2_gal_result r = {};
    r.e1 = *mes_e1;
    r.e2 = *mes_e2;
    r.maxL = maxL;
    FitRequest req{slot, &r, true, np_max};
    rl2_fit_options opt = {};
    opt.line_chunk = 4;
    bool ok = fit_batch({req}, opt, [&](const std::vector<LikItem> &items, std::vector<double> &Le,
                                        std::vector<int> &err) {
        return lik_eval(*o, items, Le.data(), err.data(), nullptr, nullptr);
    });
    if (!ok) {
        report("likelihood_sampling");
        return;
    }
    *mes_e1 = r.e1;
    *mes_e2 = r.e2;
    *var_e1 = r.var_e1;
    *var_e2 = r.var_e2;
    *oneDimvar = r.oneDimvar;
}

int source_fitting(int, likelihood_params *par, double *mes_e1, double *mes_e2, double *var_e1,
                   double *var_e2, double *oneDimvar, double *maxL)
{
    Obs *o = nullptr;
    int slot = 0;
    if (!facet_for(par, &o, &slot)) {
        report("source_fitting");
        return 1;
    }
    rl2_gal_result r = {};
    rl2_fit_options opt = {};
    opt.nm_speculate = 1;
    opt.line_chunk = 4;
    bool ok = fit_batch({{slot, &r}}, opt, [&](const std::vector<LikItem> &items,
                                               std::vector<double> &Le, std::vector<int> &err) {
        return lik_eval(*o, items, Le.data(), err.data(), nullptr, nullptr);
    });
    if (!ok) {
        report("source_fitting");
        return 1;
    }
    *mes_e1 = r.e1;
    *mes_e2 = r.e2;
    *var_e1 = r.var_e1;
    *var_e2 = r.var_e2;
    *oneDimvar = r.oneDimvar;
    *maxL = r.maxL;
    return r.error;
}

void sky_model(double *wavenumbers, double *spec, double, int, unsigned int num_channels,
               unsigned int n_gal, double *gflux, double *gscale, double *l, double *m,
               unsigned long int num_coords, double *uu_metres, double *vv_metres,
               double *ww_metres, complexd *, complexd *visMod)
{
    if (!ensure_device()) return;
    const size_t nvis = (size_t)num_channels * num_coords;
    DBuf<double> du(num_coords), dv(num_coords), dw(num_coords), dwn(num_channels),
        dsp(num_channels), dres(2 * nvis);
    DBuf<GalaxyGeom> dg(n_gal);
    std::vector<GalaxyGeom> g(n_gal);
    for (unsigned i = 0; i < n_gal; i++) g[i] = make_geom(l[i], m[i], gflux[i], gscale[i], 0., 0.);
    if (!allocated(du, dv, dw, dwn, dsp, dres, dg) || !du.up(uu_metres, num_coords) ||
        !dv.up(vv_metres, num_coords) || !dw.up(ww_metres, num_coords) ||
        !dwn.up(wavenumbers, num_channels) || !dsp.up(spec, num_channels) ||
        !dres.up((const double *)visMod, 2 * nvis) || !dg.up(g.data(), n_gal)) {
        report("sky_model");
        return;
    }
    launch_add_sources(stream(), 0, num_channels, num_coords, du.p, dv.p, dw.p, dwn.p, dsp.p, dg.p,
                       n_gal, 1.0, dres.p);
    if (!dres.down((double *)visMod, 2 * nvis) || !sync_ok()) report("sky_model");
}

void data_processing(bool re_fitting, unsigned int *bad_list, int nsources, double len,
                     unsigned int num_coords, FILE *pFile, likelihood_params *par, double *l,
                     double *m, double *gflux, double *gscale, double *ge1, double *ge2,
                     double *SNR_vis, double *, complexd *, complexd *visSkyMod, complexd *visData,
                     float *sigma2_vis, bool *flag, double *uu_metres, double *vv_metres,
                     double *ww_metres, double *fitting_time, int *bad)
{
    *bad = 0;
    Obs *o = observation(par->nchannels, num_coords, uu_metres, vv_metres, ww_metres,
                         par->wavenumbers, par->spec, sigma2_vis, flag, len);
    if (!o || rl2_obs_set_vis(g_obs.get(), (const double *)visData, (const double *)visSkyMod) !=
                  RL2_OK) {
        report("data_processing");
        return;
    }
    if (!set_grid(*o, par->numr, par->ro)) {
        report("data_processing");
        return;
    }
    std::vector<unsigned> todo(nsources);
    unsigned maxg = 0;
    for (int g = 0; g < nsources; g++) {
        todo[g] = re_fitting ? bad_list[g] : (unsigned)g;
        maxg = std::max(maxg, todo[g]);
    }
    std::vector<rl2_gal_result> res(nsources ? maxg + 1 : 0);
    std::vector<unsigned> new_bad;
    double t_ext = 0., t_fit = 0., t_sub = 0.;
    rl2_fit_options opt = {};
    opt.batch = 1;
    opt.nm_speculate = 1;
    opt.line_chunk = 4;
    auto write_line = [&](unsigned gal) {
        const rl2_gal_result &r = res[gal];
        if (r.error && !re_fitting) return;
        if (pFile)
            fprintf(pFile, "%f | %f | %f | %f | %f | %f | %f | %f | %f | %f | %f \n", gflux[gal],
                    ge1[gal], r.e1, sqrt(r.var_e1), ge2[gal], r.e2, sqrt(r.var_e2), r.oneDimvar,
                    SNR_vis[gal], l[gal] / (kPi / 648000.0), m[gal] / (kPi / 648000.0));
    };
    bool ok = process_list(*o, todo, re_fitting ? 1 : 0, l, m, gflux, gscale, opt, res.data(),
                           re_fitting ? nullptr : &new_bad, &t_ext, &t_fit, &t_sub, write_line);
    if (!ok) {
        report("data_processing");
        return;
    }
    int nbad = 0;
    if (!re_fitting) {
        for (size_t i = 0; i < new_bad.size(); i++) bad_list[i] = new_bad[i];
        nbad = (int)new_bad.size();
    } else {
        for (unsigned g : todo)
            if (res[g].error) nbad++;
    }
    *bad = nbad;
    *fitting_time += t_fit;
    if (rl2_obs_sync(g_obs.get(), (double *)visData, (double *)visSkyMod) != RL2_OK)
        report("data_processing");
}

// ---- test hooks: the shared scalar helpers the kernels use, evaluated on the host ----
void rl2_test_sincos_rev64(int n, const double *t, double *s, double *c)
{
    for (int i = 0; i < n; i++) rl2::sincos_rev64(t[i], s[i], c[i]);
}

// The simulated noise's generator as the kernel evaluates it (n counters, one key).
void rl2_test_philox4x32_10(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    for (int i = 0; i < n; i++) rl2::philox4x32_10(ctr + 4 * i, key, out + 4 * i);
}

// Cell of each (u, v) by the kernels' reciprocal-product form and by the reference's
// divisions; returns the number of disagreements.
int rl2_test_cell_index(int n, const double *u, const double *v, double inv_lambda, double len,
                        double inc, unsigned int facet, unsigned int *fast, unsigned int *exact)
{
    int bad = 0;
    const double inv_inc = 1.0 / inc;
    for (int i = 0; i < n; i++) {
        uint32_t a = 0, b = 0;
        const bool oka = rl2::cell_index_fast(u[i], v[i], inv_lambda, len, inc, inv_inc, facet, &a);
        const bool okb = rl2::cell_index(u[i], v[i], inv_lambda, len, inc, facet, &b);
        fast[i] = oka ? a : 0xffffffffu;
        exact[i] = okb ? b : 0xffffffffu;
        bad += fast[i] != exact[i];
    }
    return bad;
}

}  // extern "C"
