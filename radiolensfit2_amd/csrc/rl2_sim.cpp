// rl2_sim.cpp -- synthetic observations, catalogs and the RLMS container (host C++).
// Declarations and reference citations: include/rl2_sim.h.
#include "rl2_sim.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSidereal = 2.0 * kPi / 86164.0905;  // rad per second of time

// default_params.h:38-62
constexpr double kRmin = 0.3, kRmax = 3.5, kAdd = -0.93, kEsp = 0.33, kRstd = 0.3136;
constexpr double kMexp = -1.34;
constexpr double kEmax = 0.804, kE0 = 0.0732, kDisp = 0.2298, kNormE = 2.595;
constexpr int kNumOrient = 5, kNPoints = 1000;

struct MT19937 {
    static constexpr int N = 624, M = 397;
    uint32_t mt[N];
    int mti = N + 1;
    void seed(unsigned long s)
    {
        if (s == 0) s = 4357;  // gsl_rng_mt19937 default
        mt[0] = (uint32_t)(s & 0xffffffffUL);
        for (int i = 1; i < N; i++)
            mt[i] = (uint32_t)((1812433253UL * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xffffffffUL);
        mti = N;
    }
    uint32_t get()
    {
        const uint32_t UPPER = 0x80000000UL, LOWER = 0x7fffffffUL;
        if (mti >= N) {
            int kk;
            for (kk = 0; kk < N - M; kk++) {
                uint32_t y = (mt[kk] & UPPER) | (mt[kk + 1] & LOWER);
                mt[kk] = mt[kk + M] ^ (y >> 1) ^ ((y & 1) ? 0x9908b0dfUL : 0UL);
            }
            for (; kk < N - 1; kk++) {
                uint32_t y = (mt[kk] & UPPER) | (mt[kk + 1] & LOWER);
                mt[kk] = mt[kk + (M - N)] ^ (y >> 1) ^ ((y & 1) ? 0x9908b0dfUL : 0UL);
            }
            uint32_t y = (mt[N - 1] & UPPER) | (mt[0] & LOWER);
            mt[N - 1] = mt[M - 1] ^ (y >> 1) ^ ((y & 1) ? 0x9908b0dfUL : 0UL);
            mti = 0;
        }
        uint32_t k = mt[mti++];
        k ^= (k >> 11);
        k ^= (k << 7) & 0x9d2c5680UL;
        k ^= (k << 15) & 0xefc60000UL;
        k ^= (k >> 18);
        return k;
    }
    double uniform() { return get() / 4294967296.0; }
    double uniform_pos()
    {
        double x;
        do {
            x = uniform();
        } while (x == 0);
        return x;
    }
    // utils.c:77-101, polar Box-Muller
    double gaussian(double *another)
    {
        double x, y, r2, fac;
        do {
            x = 2.0 * uniform() - 1.0;
            y = 2.0 * uniform() - 1.0;
            r2 = x * x + y * y;
        } while (r2 >= 1.0 || r2 == 0.0);
        fac = sqrt(-2.0 * log(r2) / r2);
        x *= fac;
        if (another) *another = y * fac;
        return x;
    }
};

double round6(double x)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%f", x);
    return strtod(buf, nullptr);
}

// distributions.c:35-45
double e_pdf(double e)
{
    return kNormE * e * (1. - exp((e - kEmax) / kDisp)) / ((1. + e) * sqrt(e * e + kE0 * kE0));
}
double scale_mean(double flux) { return kAdd + kEsp * log(flux); }

// ---- sampling of the catalog priors (distributions.c:35-129 define the densities) ----
// Flux: density proportional to S^kMexp on [fmin, fmax]; the inverse of its cumulative
// distribution is closed-form, S = (fmin^a + u (fmax^a - fmin^a))^(1/a) with a = kMexp + 1.
double draw_flux(MT19937 &gen, double fmin, double fmax)
{
    const double a = kMexp + 1.;
    const double lo = pow(fmin, a), hi = pow(fmax, a);
    return pow(lo + gen.uniform() * (hi - lo), 1. / a);
}

// Scalelength: log-normal of mean log `mean` and width kRstd, truncated to [kRmin, kRmax].
// The target probability is placed between the two truncation points, and log r is found by
// bisection on the normal cumulative distribution (60 halvings: below one ulp of log r).
double draw_scale(MT19937 &gen, double mean)
{
    auto Phi = [&](double logr) { return 0.5 * (1. + erf((logr - mean) / (sqrt(2.) * kRstd))); };
    double lo = log(kRmin), hi = log(kRmax);
    const double target = Phi(lo) + gen.uniform() * (Phi(hi) - Phi(lo));
    for (int it = 0; it < 60; it++) {
        const double mid = 0.5 * (lo + hi);
        (Phi(mid) < target ? lo : hi) = mid;
    }
    return exp(0.5 * (lo + hi));
}

// Ellipticity modulus: the cumulative distribution of e_pdf on kNPoints intervals of
// [0, kEmax], each interval's integral by 8-point Gauss-Legendre quadrature (the density is
// smooth there), normalised to end at 1; a draw is the linear interpolation of the inverse
// on the interval that binary search finds.
const std::vector<double> &e_cdf_table()
{
    static const std::vector<double> F = [] {
        static const double xg[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267,
                                     0.9602898564975363};
        static const double wg[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745,
                                     0.1012285362903763};
        std::vector<double> t(kNPoints + 1, 0.);
        const double h = kEmax / kNPoints;
        for (int i = 0; i < kNPoints; i++) {
            const double c = (i + 0.5) * h, r = 0.5 * h;
            double q = 0.;
            for (int j = 0; j < 4; j++) q += wg[j] * (e_pdf(c - r * xg[j]) + e_pdf(c + r * xg[j]));
            t[i + 1] = t[i] + r * q;
        }
        const double total = t[kNPoints];
        for (double &v : t) v /= total;
        return t;
    }();
    return F;
}

double draw_modulus(MT19937 &gen)
{
    const std::vector<double> &F = e_cdf_table();
    const double u = gen.uniform();
    const size_t k = std::min<size_t>(
        kNPoints, std::max<size_t>(1, std::upper_bound(F.begin(), F.end(), u) - F.begin()));
    const double h = kEmax / kNPoints;
    const double span = F[k] - F[k - 1];
    return h * (k - 1) + (span > 0. ? h * (u - F[k - 1]) / span : 0.);
}

// Exponential-disk visibility of one galaxy in one channel (galaxy_visibilities.c:154-190).
inline void galaxy_vis(double spectra, double wavenumber, double e1, double e2, double scale,
                       double flux, double l, double m, unsigned long nrow, const double *u,
                       const double *v, const double *w, double *vis_acc, double *snr_acc)
{
    double detA = 1. - e1 * e1 - e2 * e2;
    scale *= kPi / 648000.0;
    double sf = (scale * scale) / (detA * detA);
    double wn2 = wavenumber * wavenumber;
    double n = sqrt(1. - l * l - m * m) - 1.;
    double snr = 0.;
    for (unsigned long i = 0; i < nrow; ++i) {
        double ui = u[i], vi = v[i], wi = w[i];
        double phase = ui * l + vi * m + wi * n;
        phase = wavenumber * phase;
        double k1 = (1. + e1) * ui + e2 * vi;
        double k2 = e2 * ui + (1. - e1) * vi;
        double den = 1. + sf * wn2 * (k1 * k1 + k2 * k2);
        double shape = 1. / (den * sqrt(den));
        shape *= spectra * flux;
        double re = shape * cos(phase), im = shape * sin(phase);
        snr += re * re + im * im;
        vis_acc[2 * i] += re;
        vis_acc[2 * i + 1] += im;
    }
    *snr_acc = snr;
}

}  // namespace

struct rl2sim_mt {
    MT19937 g;
};

extern "C" {

rl2sim_mt *rl2sim_mt_new(unsigned long seed)
{
    rl2sim_mt *p = new rl2sim_mt;
    p->g.seed(seed);
    return p;
}
void rl2sim_mt_free(rl2sim_mt *g) { delete g; }
uint32_t rl2sim_mt_get(rl2sim_mt *g) { return g->g.get(); }
double rl2sim_mt_uniform(rl2sim_mt *g) { return g->g.uniform(); }
double rl2sim_mt_uniform_pos(rl2sim_mt *g) { return g->g.uniform_pos(); }
double rl2sim_mt_gaussian(rl2sim_mt *g, double *another) { return g->g.gaussian(another); }

int rl2sim_layout(int kind, unsigned int n, unsigned long seed, double r_min, double r_max,
                  double core_frac, double r_core, double *east, double *north)
{
    MT19937 g;
    g.seed(seed);
    unsigned int ncore = (kind == 1) ? (unsigned int)(n * core_frac) : 0;
    for (unsigned int i = 0; i < n; i++) {
        double r;
        if (i < ncore)
            r = r_core * sqrt(g.uniform());
        else {
            double lo = (kind == 1) ? r_core : r_min;
            r = exp(log(lo) + g.uniform() * (log(r_max) - log(lo)));
        }
        double az = 2. * kPi * g.uniform();
        east[i] = r * sin(az);
        north[i] = r * cos(az);
    }
    return 0;
}

unsigned long rl2sim_uvw(unsigned int n, const double *east, const double *north, double lat_deg,
                         double dec_deg, unsigned int ntimes, double dt_s, double h0_hours,
                         double *u, double *v, double *w)
{
    const double phi = lat_deg * kPi / 180.0, dec = dec_deg * kPi / 180.0;
    const double sphi = sin(phi), cphi = cos(phi), sd = sin(dec), cd = cos(dec);
    unsigned long nbl = (unsigned long)n * (n - 1) / 2;
    std::vector<double> X(nbl), Y(nbl), Z(nbl);
    unsigned long b = 0;
    for (unsigned int i = 0; i < n; i++)
        for (unsigned int j = i + 1; j < n; j++) {
            double dE = east[j] - east[i], dN = north[j] - north[i];
            X[b] = -sphi * dN;  // up component is zero for a planar array
            Y[b] = dE;
            Z[b] = cphi * dN;
            b++;
        }
    for (unsigned int t = 0; t < ntimes; t++) {
        double H = h0_hours * (2. * kPi / 24.) + t * dt_s * kSidereal;
        double sH = sin(H), cH = cos(H);
        for (unsigned long k = 0; k < nbl; k++) {
            unsigned long r = (unsigned long)t * nbl + k;
            u[r] = sH * X[k] + cH * Y[k];
            v[r] = -sd * cH * X[k] + sd * sH * Y[k] + cd * Z[k];
            w[r] = cd * cH * X[k] - cd * sH * Y[k] + sd * Z[k];
        }
    }
    return nbl * ntimes;
}

unsigned int rl2sim_catalog(unsigned int nge, unsigned long seed, double fov_arcmin, double fmin,
                            double fmax, double *gflux, double *gscale, double *ge1, double *ge2,
                            double *l, double *m)
{
    MT19937 gen;
    gen.seed(seed);
    double fov_eff = fov_arcmin * 60. * kPi / 648000.0;
    const int NP = kNumOrient;
    unsigned int diffgal = nge / (2 * NP);
    unsigned int my_gal = diffgal * 2 * NP;
    // one flux, scalelength and ellipticity modulus per group of 2 NP galaxies: NP
    // orientations, each with its opposite, so that the group's intrinsic shapes average to
    // zero (the catalog layout of generate_catalog.c:97-117); groups in decreasing flux
    std::vector<double> f2(diffgal);
    for (unsigned int i = 0; i < diffgal; i++) f2[i] = draw_flux(gen, fmin, fmax);
    std::sort(f2.begin(), f2.end(), std::greater<double>());
    for (unsigned int grp = 0; grp < diffgal; grp++) {
        const double r = draw_scale(gen, scale_mean(f2[grp]));
        for (int k = 0; k < 2 * NP; k++) {
            gflux[grp * 2 * NP + k] = f2[grp];
            gscale[grp * 2 * NP + k] = r;
        }
    }
    for (unsigned int grp = 0; grp < diffgal; grp++) {
        const double modulus = draw_modulus(gen);
        const double phi0 = 2 * kPi * gen.uniform_pos();
        for (int k = 0; k < NP; k++) {
            const double phi = phi0 + k * (kPi / NP);
            const unsigned int g = grp * 2 * NP + 2 * k;
            ge1[g] = modulus * cos(phi);
            ge2[g] = modulus * sin(phi);
            ge1[g + 1] = -ge1[g];
            ge2[g + 1] = -ge2[g];
        }
    }
    for (unsigned int g = 0; g < my_gal; g++) {
        double radius = sqrt(gen.uniform()) * 0.5 * fov_eff;
        double orient = gen.uniform() * 2 * kPi;
        l[g] = radius * cos(orient);
        m[g] = radius * sin(orient);
    }
    for (unsigned int g = 0; g < my_gal; g++) {
        gflux[g] = round6(gflux[g]);
        gscale[g] = round6(gscale[g]);
        ge1[g] = round6(ge1[g]);
        ge2[g] = round6(ge2[g]);
        l[g] = round6(l[g]);
        m[g] = round6(m[g]);
    }
    return my_gal;
}

void rl2sim_shear(unsigned int ngal, double g1, double g2, double *ge1, double *ge2)
{
    for (unsigned int g = 0; g < ngal; g++) {
        double ee1 = ge1[g], ee2 = ge2[g];
        double z1r = ee1 + g1, z1i = ee2 + g2;
        double z2r = 1. + ee1 * g1 + ee2 * g2, z2i = ee2 * g1 - ee1 * g2;
        double den = z2r * z2r + z2i * z2i;
        ge1[g] = (z1r * z2r + z1i * z2i) / den;
        ge2[g] = (z1i * z2r - z1r * z2i) / den;
    }
}

int rl2sim_simulate(unsigned int nch, const double *wavenumbers, const double *spec,
                    unsigned long nrow, const double *u, const double *v, const double *w,
                    unsigned int ngal, const double *flux, const double *scale, const double *e1,
                    const double *e2, const double *l, const double *m, float sigma2,
                    int add_noise, unsigned long seed, int ms_roundtrip, double *vis, double *snr)
{
    std::vector<double> snr_ch(nch);
    for (unsigned int g = 0; g < ngal; g++) {
#pragma omp parallel for schedule(static)
        for (unsigned int ch = 0; ch < nch; ch++) {
            unsigned long ch_vis = (unsigned long)ch * nrow;
            galaxy_vis(spec[ch], wavenumbers[ch], e1[g], e2[g], scale[g], flux[g], l[g], m[g], nrow,
                       u, v, w, vis + 2 * ch_vis, &snr_ch[ch]);
        }
        if (snr) {
            double s = 0.;
            for (unsigned int ch = 0; ch < nch; ch++) s += snr_ch[ch];
            snr[g] = s;
        }
    }
    if (add_noise) {
        MT19937 gen;
        gen.seed(seed);
        double std = sqrt(sigma2);
        unsigned long nvis = (unsigned long)nch * nrow;
        for (unsigned long i = 0; i < nvis; i++) {
            double r2;
            double r1 = gen.gaussian(&r2);
            vis[2 * i] += r1 * std;
            vis[2 * i + 1] += r2 * std;
        }
    }
    if (snr)
        for (unsigned int g = 0; g < ngal; g++) snr[g] = sqrt(snr[g] / sigma2);
    if (ms_roundtrip) {
        unsigned long nvis = (unsigned long)nch * nrow;
#pragma omp parallel for schedule(static)
        for (unsigned long i = 0; i < 2 * nvis; i++) {
            float jy = (float)(vis[i] * 1e-6);
            double back = jy;
            back *= 1e+6;
            vis[i] = back;
        }
    }
    return 0;
}

// ---------------------------------------------------------------- RLMS container
namespace {
struct RlmsHeader {
    char magic[8];  // "RLMS0001"
    uint32_t nstations, nch;
    uint64_t nrow;
    double freq_start_hz, freq_inc_hz, time_inc_s;
};
}  // namespace

int rl2ms_write(const char *path, unsigned int nstations, unsigned int nch, unsigned long nrow,
                double freq_start_hz, double freq_inc_hz, double time_inc_s, const double *u,
                const double *v, const double *w, const unsigned char *flag_rc,
                const float *data_rc, const float *sigma_r)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    RlmsHeader h;
    memcpy(h.magic, "RLMS0001", 8);
    h.nstations = nstations;
    h.nch = nch;
    h.nrow = nrow;
    h.freq_start_hz = freq_start_hz;
    h.freq_inc_hz = freq_inc_hz;
    h.time_inc_s = time_inc_s;
    size_t ok = fwrite(&h, sizeof h, 1, f);
    std::vector<double> uvw(3 * nrow);
    for (unsigned long i = 0; i < nrow; i++) {
        uvw[3 * i] = u[i];
        uvw[3 * i + 1] = v[i];
        uvw[3 * i + 2] = w[i];
    }
    ok &= fwrite(uvw.data(), sizeof(double), 3 * nrow, f) == 3 * nrow;
    std::vector<unsigned char> fl((size_t)nrow * nch, 0);
    if (flag_rc) memcpy(fl.data(), flag_rc, fl.size());
    ok &= fwrite(fl.data(), 1, fl.size(), f) == fl.size();
    ok &= fwrite(data_rc, sizeof(float), 2 * (size_t)nrow * nch, f) == 2 * (size_t)nrow * nch;
    std::vector<float> sg(nrow, 0.f);
    if (sigma_r) memcpy(sg.data(), sigma_r, nrow * sizeof(float));
    ok &= fwrite(sg.data(), sizeof(float), nrow, f) == nrow;
    fclose(f);
    return ok ? 0 : -2;
}

int rl2ms_info(const char *path, unsigned int *nstations, unsigned int *nch, unsigned long *nrow,
               double *freq_start_hz, double *freq_inc_hz, double *time_inc_s)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    RlmsHeader h;
    size_t ok = fread(&h, sizeof h, 1, f);
    fclose(f);
    if (!ok || memcmp(h.magic, "RLMS0001", 8)) return -2;
    *nstations = h.nstations;
    *nch = h.nch;
    *nrow = h.nrow;
    *freq_start_hz = h.freq_start_hz;
    *freq_inc_hz = h.freq_inc_hz;
    *time_inc_s = h.time_inc_s;
    return 0;
}

double rl2ms_read(const char *path, double *u, double *v, double *w, unsigned char *flag,
                  double *vis, float *sigma2, int *status)
{
    *status = 0;
    FILE *f = fopen(path, "rb");
    if (!f) {
        *status = -1;
        return 0.;
    }
    RlmsHeader h;
    if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, "RLMS0001", 8)) {
        fclose(f);
        *status = -2;
        return 0.;
    }
    unsigned long nrow = h.nrow;
    unsigned int nch = h.nch;
    std::vector<double> uvw(3 * nrow);
    size_t ok = fread(uvw.data(), sizeof(double), 3 * nrow, f) == 3 * nrow;
    double umax = 0., vmax = 0.;
    for (unsigned long i = 0; i < nrow; i++) {
        u[i] = uvw[3 * i];
        v[i] = uvw[3 * i + 1];
        w[i] = uvw[3 * i + 2];
        if (fabs(u[i]) > umax) umax = fabs(u[i]);
        if (fabs(v[i]) > vmax) vmax = fabs(v[i]);
    }
    std::vector<unsigned char> fl((size_t)nrow * nch);
    ok &= fread(fl.data(), 1, fl.size(), f) == fl.size();
    std::vector<float> d(2 * (size_t)nrow * nch);
    ok &= fread(d.data(), sizeof(float), d.size(), f) == d.size();
    std::vector<float> sg(nrow);
    ok &= fread(sg.data(), sizeof(float), nrow, f) == nrow;
    fclose(f);
    if (!ok) {
        *status = -3;
        return 0.;
    }
    // measurement_set.cpp:286-341, :354-418: transpose row-major (row, channel) to
    // channel-major, Jy -> micro-Jy.
    for (unsigned int c = 0; c < nch; c++)
        for (unsigned long b = 0; b < nrow; b++) {
            size_t i = (size_t)b * nch + c, j = (size_t)c * nrow + b;
            if (flag) flag[j] = fl[i];
            double re = d[2 * i], im = d[2 * i + 1];
            re *= 1e+6;
            im *= 1e+6;
            vis[2 * j] = re;
            vis[2 * j + 1] = im;
        }
    // measurement_set.cpp:270-273: SIGMA^2 fills only the first nrow entries.
    if (sigma2) {
        for (unsigned long i = 0; i < (unsigned long)nch * nrow; i++) sigma2[i] = 0.f;
        for (unsigned long i = 0; i < nrow; i++) sigma2[i] = sg[i] * sg[i];
    }
    return ceil(fmax(umax, vmax));
}

}  // extern "C"
