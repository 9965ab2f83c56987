/*
 * rl2_oracle.cpp -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see rl2_oracle.h).
 *
 * Build: oracle/Makefile (g++ -O2 -fno-fast-math -ffp-contract=off -fopenmp).
 * Parallel decomposition follows the reference: OpenMP over channels for extraction,
 * subtraction and the sky model, OpenMP over scalelength samples for the likelihood.
 * Expressions keep the reference's evaluation order so the strict build reproduces its
 * FP64 rounding (e.g. the unparenthesised ARCS2RAD macro, datatype.h:40).
 */
#include "rl2_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <chrono>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "gk41_table.h"

/* datatype.h:28-42 and default_params.h:28-79 */
#define PI 3.14159265358979323846
#define C0 299792458.0
#define EFFICIENCY 0.9
#define SEFD_SKA 400e+6
#define REF_FREQ 1.4e+9
#define PSF_NAT 38
#define RMIN 0.3
#define RMAX 3.5
#define NUM_R 29
#define ADD -0.93
#define ESP 0.33
#define R_STD 0.3136
#define TOL 1.0e-3
#define NP_MAX 30
#define E_RES 3.0e-3
#define VAR 1e-4

static const double GSL_DBL_EPSILON_ = 2.2204460492503131e-16;
static const double GSL_DBL_MIN_ = 2.2250738585072014e-308;

/* Definitions take C linkage from the declarations in rl2_oracle.h. */

/* ======================================================================
 * Setup arithmetic
 * ====================================================================== */

/* RadioLensfit2.cpp:342-351: channel centre wavenumber and spectral factor. */
void orl2_channel_constants(unsigned int nch, double freq_start_hz, double bw_hz,
                            double *wavenumbers, double *spec)
{
    double ch_freq = freq_start_hz + 0.5 * bw_hz;
    for (unsigned int i = 0; i < nch; i++) {
        wavenumbers[i] = 2.0 * PI * ch_freq / C0;
        spec[i] = pow(ch_freq / REF_FREQ, -0.7);
        ch_freq += bw_hz;
    }
}

/* RadioLensfit2.cpp:255: theoretical visibility variance (stored as float). */
float orl2_default_sigma2(int time_acc, double bw_hz)
{
    const double efficiency = EFFICIENCY;
    const double SEFD = SEFD_SKA;
    float s = (SEFD * SEFD) / (2. * time_acc * bw_hz * efficiency * efficiency);
    return s;
}

/* measurement_set.cpp:223-240 (ceil of max |u|,|v| in metres) and RadioLensfit2.cpp:408. */
double orl2_len_lambda(const double *u, const double *v, unsigned long nrow, double wn_top)
{
    double umax = 0., vmax = 0.;
    for (unsigned long i = 0; i < nrow; ++i) {
        double ua = fabs(u[i]), va = fabs(v[i]);
        if (ua > umax) umax = ua;
        if (va > vmax) vmax = va;
    }
    double len = ceil(fmax(umax, vmax));
    len = len * wn_top / (2 * PI);
    return len;
}

/* RadioLensfit2.cpp:370-392: quadratic r grid; Ro[0]=0 is a placeholder. */
int orl2_r_grid(double *Ro)
{
    const double Rmin = RMIN;
    double Rmax = RMAX;
    int numR = NUM_R;
    Ro[0] = 0.;
    Ro[1] = Rmin;
    int nRo = 2;
    while (nRo < numR && Ro[nRo - 1] < Rmax) {
        double Rinterval = 0.08 + 0.4 * pow(((Ro[nRo - 1] - Rmin) / (Rmax - Rmin)), 2);
        Ro[nRo] = Ro[nRo - 1] + Rinterval;
        nRo++;
    }
    return nRo;
}

/* distributions.c:59-62 */
double orl2_scale_mean(double flux) { return ADD + ESP * log(flux); }

/* distributions.c:72-87 (exit on r<=0 replaced by NaN: the oracle never aborts a test) */
double orl2_rfunc(double mu, double sigma, double r)
{
    if (r <= 0.) return NAN;
    double x = (log(r) - mu) / sigma;
    double logprior = -(x * x) * 0.5 - log(r * sigma);
    return logprior;
}

/* ======================================================================
 * galaxy_visibilities.c
 * ====================================================================== */

/* galaxy_visibilities.c:48-53 */
double orl2_sersic_shape(double k1, double k2, double scale_factor, double wave_factor)
{
    double den = 1. + scale_factor * wave_factor * (k1 * k1 + k2 * k2);
    double vis = 1. / (den * sqrt(den));
    return vis;
}

/* galaxy_visibilities.c:40-44 (-DGAUSSIAN builds only) */
double orl2_gaussian_shape(double k1, double k2, double scale_factor, double wave_factor)
{
    return exp(-0.5 * scale_factor * wave_factor * (k1 * k1 + k2 * k2));
}

/* galaxy_visibilities.c:59-91 */
void orl2_model_at_zero(double e1, double e2, double scale, unsigned int num_coords,
                        const double *grid_u, const double *grid_v, const double *sigma2,
                        double *Modvis)
{
    double uu, vv, k1, k2;
    double detA = 1. - e1 * e1 - e2 * e2;
    scale *= PI / 648000.0;
    double scale_factor = (scale * scale) / (detA * detA);
    double sum = 0.;
    unsigned int nv = 0;
    double cc2 = 4 * PI * PI;
    for (unsigned int i = 0; i < num_coords; ++i) {
        uu = grid_u[i];
        vv = grid_v[i];
        k1 = (1. + e1) * uu + e2 * vv;
        k2 = e2 * uu + (1. - e1) * vv;
        Modvis[nv] = orl2_sersic_shape(k1, k2, scale_factor, cc2);
        sum += Modvis[nv] * Modvis[nv] / sigma2[nv];
        nv++;
    }
    sum = sqrt(sum);
    for (unsigned int k = 0; k < nv; k++) Modvis[k] /= sum;
}

/* galaxy_visibilities.c:154-190 */
void orl2_data_galaxy_vis(double spectra, double wavenumber, double e1, double e2, double scale,
                          double flux, double l, double m, unsigned long num_coords,
                          const double *uu_metres, const double *vv_metres,
                          const double *ww_metres, orl2_complexd *vis)
{
    double u, v, w, k1, k2, phase, shape;
    double detA = 1. - e1 * e1 - e2 * e2;
    scale *= PI / 648000.0;
    double scale_factor = (scale * scale) / (detA * detA);
    double wavenumber2 = wavenumber * wavenumber;
    double n = sqrt(1. - l * l - m * m) - 1.;
    for (unsigned long i = 0; i < num_coords; ++i) {
        u = uu_metres[i];
        v = vv_metres[i];
        w = ww_metres[i];
        phase = u * l + v * m + w * n;
        phase = wavenumber * phase;
        k1 = (1. + e1) * u + e2 * v;
        k2 = e2 * u + (1. - e1) * v;
        shape = orl2_sersic_shape(k1, k2, scale_factor, wavenumber2);
        shape *= spectra * flux;
        vis[i].real = shape * cos(phase);
        vis[i].imag = shape * sin(phase);
    }
}

/* galaxy_visibilities.c:244-270 */
void orl2_phase_shift(double wavenumber, double l, double m, unsigned long num_coords,
                      const double *uu_metres, const double *vv_metres, const double *ww_metres,
                      orl2_complexd *vis)
{
    double u, v, w, phase, sp, cp;
    orl2_complexd temp;
    double n = sqrt(1. - l * l - m * m) - 1.;
    for (unsigned long i = 0; i < num_coords; ++i) {
        u = uu_metres[i];
        v = vv_metres[i];
        w = ww_metres[i];
        phase = u * l + v * m + w * n;
        phase *= -wavenumber;
        sp = sin(phase);
        cp = cos(phase);
        temp.real = vis[i].real * cp - vis[i].imag * sp;
        temp.imag = vis[i].real * sp + vis[i].imag * cp;
        vis[i].real = temp.real;
        vis[i].imag = temp.imag;
    }
}

/* ======================================================================
 * evaluate_uv_grid.c
 * ====================================================================== */

/* evaluate_uv_grid.c:50-55. The macro expansion evaluates
 * len*PSF_NAT*theta*PI/648000.0 left to right (datatype.h:40 is unparenthesised). */
unsigned int orl2_facet_size(double theta_med, double len)
{
    unsigned int fs = (unsigned int)2 * ceil(len * PSF_NAT * theta_med * PI / 648000.0);
    return fs;
}

/* Cell index of one visibility: evaluate_uv_grid.c:75-77 / :166-169.
 * Returns false when the cell lies past the facet (pc >= sizeg^2): the reference writes
 * out of bounds there; the oracle drops such visibilities. */
static inline bool cell_index(double u, double v, double inv_lambda, double len, double inc,
                              unsigned int sizeg, unsigned long *pc_out)
{
    double uu = u * inv_lambda;
    double vv = v * inv_lambda;
    unsigned int pu = (unsigned int)((uu + len) / inc);
    unsigned int pv = (unsigned int)((vv + len) / inc);
    unsigned long pc = (unsigned long)pv * sizeg + pu;
    *pc_out = pc;
    return pc < (unsigned long)sizeg * sizeg;
}

/* evaluate_uv_grid.c:58-109 (non-MPI branch) */
unsigned int orl2_uv_grid_size(double len, const double *wavenumbers, unsigned int num_channels,
                               unsigned long ncoords, const double *u, const double *v,
                               unsigned int sizeg, const bool *flag)
{
    unsigned long size = (unsigned long)sizeg * sizeg;
    std::vector<unsigned long> count(size, 0);
    double inc = 2 * len / sizeg;
    unsigned long n = 0;
    for (unsigned int ch = 0; ch < num_channels; ch++) {
        double inv_lambda = wavenumbers[ch] / (2. * PI);
        for (unsigned long k = 0; k < ncoords; k++) {
            if (!flag || !flag[n]) {
                unsigned long pc;
                if (cell_index(u[k], v[k], inv_lambda, len, inc, sizeg, &pc)) count[pc]++;
            }
            n++;
        }
    }
    n = 0;
    for (unsigned long p = 0; p < size; p++)
        if (count[p]) n++;
    return (unsigned int)n;
}

/* evaluate_uv_grid.c:113-133 */
unsigned int orl2_facet_coords(double *grid_u, double *grid_v, double len, unsigned int sizeg,
                               const double *count_w)
{
    unsigned int i, j;
    unsigned long p, n;
    double inc = 2 * len / sizeg;
    unsigned long size = (unsigned long)sizeg * sizeg;
    n = 0;
    for (p = 0; p < size; p++) {
        if (count_w[p]) {
            j = p / sizeg;
            i = p % sizeg;
            grid_u[n] = (i + 0.5) * inc - len;
            grid_v[n] = (j + 0.5) * inc - len;
            n++;
        }
    }
    return (unsigned int)n;
}

/* evaluate_uv_grid.c:143-211 (non-MPI branch: accumulate then compact in cell order) */
void orl2_gridding(const double *wavenumbers, unsigned int num_channels, unsigned long ncoords,
                   const double *u, const double *v, const orl2_complexd *vis,
                   const float *sigma2, double len, unsigned int sizeg, orl2_complexd *new_vis,
                   double *new_sigma2, const bool *flag, double *sum_w)
{
    double inc = 2 * len / sizeg;
    unsigned long size = (unsigned long)sizeg * sizeg;
    for (unsigned long i = 0; i < size; i++) sum_w[i] = 0.;
    std::vector<orl2_complexd> temp_grid_vis(size, orl2_complexd{0., 0.});
    std::vector<double> temp_sigma2(size, 0.);
    unsigned long n = 0;
    const double weight = 1.;
    for (unsigned int ch = 0; ch < num_channels; ch++) {
        double inv_lambda = wavenumbers[ch] / (2. * PI);
        for (unsigned long k = 0; k < ncoords; k++) {
            if (!flag || !flag[n]) {
                unsigned long pc;
                if (cell_index(u[k], v[k], inv_lambda, len, inc, sizeg, &pc)) {
                    temp_grid_vis[pc].real += weight * vis[n].real;
                    temp_grid_vis[pc].imag += weight * vis[n].imag;
                    temp_sigma2[pc] += weight * weight * sigma2[n];
                    sum_w[pc] += weight;
                }
            }
            n++;
        }
    }
    n = 0;
    for (unsigned long p = 0; p < size; p++) {
        if (sum_w[p]) {
            new_vis[n].real = temp_grid_vis[p].real * sum_w[p] / temp_sigma2[p];
            new_vis[n].imag = temp_grid_vis[p].imag * sum_w[p] / temp_sigma2[p];
            new_sigma2[n] = temp_sigma2[p] / (sum_w[p] * sum_w[p]);
            n++;
        }
    }
}

/* evaluate_uv_grid.c:216-230 */
void orl2_average_facets(unsigned long size, orl2_complexd *facet_vis, double *facet_sigma2,
                         const double *sum_w)
{
    unsigned long n = 0;
    for (unsigned long p = 0; p < size; p++) {
        if (sum_w[p]) {
            facet_vis[n].real = facet_vis[p].real * sum_w[p] / facet_sigma2[p];
            facet_vis[n].imag = facet_vis[p].imag * sum_w[p] / facet_sigma2[p];
            facet_sigma2[n] = facet_sigma2[p] / (sum_w[p] * sum_w[p]);
            n++;
        }
    }
}

/* ======================================================================
 * source_extraction.c (FACET, :36-75)
 * ====================================================================== */
void orl2_source_extraction(unsigned int facet, unsigned int nch, const double *wavenumbers,
                            const double *spec, orl2_complexd *facet_vis, double *facet_sigma2,
                            double *sum_w, double l0, double m0, double flux, double mu,
                            double e1, double e2, const orl2_complexd *visSkyMod,
                            const orl2_complexd *visData, orl2_complexd *visGal,
                            const float *sigma2_vis, const bool *flag, unsigned long num_coords,
                            const double *uu_metres, const double *vv_metres,
                            const double *ww_metres, double len)
{
#pragma omp parallel for schedule(static)
    for (unsigned int ch = 0; ch < nch; ch++) {
        unsigned long ch_vis = (unsigned long)ch * num_coords;
        orl2_data_galaxy_vis(spec[ch], wavenumbers[ch], e1, e2, mu, flux, l0, m0, num_coords,
                             uu_metres, vv_metres, ww_metres, &visGal[ch_vis]);
        for (unsigned long i = ch_vis; i < ch_vis + num_coords; i++) {
            visGal[i].real += visData[i].real - visSkyMod[i].real;
            visGal[i].imag += visData[i].imag - visSkyMod[i].imag;
        }
        orl2_phase_shift(wavenumbers[ch], l0, m0, num_coords, uu_metres, vv_metres, ww_metres,
                         &visGal[ch_vis]);
    }
    orl2_gridding(wavenumbers, nch, num_coords, uu_metres, vv_metres, visGal, sigma2_vis, len,
                  facet, facet_vis, facet_sigma2, flag, sum_w);
}

/* ======================================================================
 * Exponential integral Ei(x), x > 0 (replaces gsl_sf_expint_Ei_e, likelihood.c:409-410).
 * Power series Ei(x) = gamma + ln x + sum x^k/(k k!) for x <= 40 (all terms positive),
 * asymptotic e^x/x sum k!/x^k beyond (smallest term < 1e-16 relative for x > 40).
 * ====================================================================== */
double orl2_expint_ei(double x)
{
    const double euler = 0.57721566490153286060651209008240243;
    if (x <= 0.) return NAN;
    if (x <= 40.) {
        double sum = 0., term = 1.;
        for (int k = 1; k < 1000; k++) {
            term *= x / k;
            double t = term / k;
            sum += t;
            if (t < DBL_EPSILON * 0.25 * sum) break;
        }
        return euler + log(x) + sum;
    }
    double sum = 1., term = 1.;
    for (int k = 1; k < 200; k++) {
        double prev = term;
        term *= k / x;
        if (term >= prev) break;
        sum += term;
        if (term < DBL_EPSILON * 0.25 * sum) break;
    }
    return exp(x) / x * sum;
}

/* likelihood.c:376-421 */
double orl2_marg_position(double x)
{
    double y, Delta, DeltaChiSquared;
    DeltaChiSquared = 5.991;
    Delta = 0.5 * DeltaChiSquared;
    if (x <= Delta) {
        y = -1.e+10;
        return y;
    }
    if (x > 100.) {
        y = x - log(Delta) - exp(-Delta);
        return y;
    }
    if ((x - Delta) < 0.0001) x = Delta + 0.0001;
    double yc = orl2_expint_ei(x);
    double ys = orl2_expint_ei(x - Delta);
    if (isfinite(yc) && isfinite(ys))
        y = log((yc - ys) / log(x / (x - Delta)));
    else
        y = -1.e+10;
    return y;
}

/* likelihood.c:195-358 (FACET branch) */
int orl2_cross_correlation(unsigned long n_uv_coords, const double *grid_u, const double *grid_v,
                           const orl2_complexd *visData, const double *visMod, double *ho,
                           double *det_sigma, int *n_iter)
{
    double a, real_part, imag_part, u, v;
    double det, d2h_dx2_xo, d2h_dy2_yo, dh_dx_xo, dh_dy_yo, d2h_dxdy_o, incx, incy;
    unsigned long i;
    std::vector<orl2_complexd> h_F(n_uv_coords);

    d2h_dx2_xo = 0.;
    d2h_dy2_yo = 0.;
    dh_dx_xo = 0.;
    dh_dy_yo = 0.;
    d2h_dxdy_o = 0;
    double value = 0;
    double cc = 2 * PI;
    for (i = 0; i < n_uv_coords; i++) {
        h_F[i].real = visData[i].real * visMod[i];
        h_F[i].imag = -visData[i].imag * visMod[i];
        value += h_F[i].real;
        u = cc * grid_u[i];
        v = cc * grid_v[i];
        dh_dx_xo -= h_F[i].imag * u;
        dh_dy_yo -= h_F[i].imag * v;
        d2h_dx2_xo -= h_F[i].real * u * u;
        d2h_dy2_yo -= h_F[i].real * v * v;
        d2h_dxdy_o -= h_F[i].real * u * v;
    }
    det = d2h_dx2_xo * d2h_dy2_yo - d2h_dxdy_o * d2h_dxdy_o;
    double xo = 0., yo = 0.;
    double tolerance = 1.e-7;
    double norm = 1.;
    int iter = 0;
    while (det > 0 && (norm > tolerance) && iter < 50) {
        iter++;
        incx = (dh_dx_xo * d2h_dy2_yo - dh_dy_yo * d2h_dxdy_o) / det;
        incy = (dh_dy_yo * d2h_dx2_xo - dh_dx_xo * d2h_dxdy_o) / det;
        xo -= incx;
        yo -= incy;
        d2h_dx2_xo = 0.;
        d2h_dy2_yo = 0.;
        dh_dx_xo = 0.;
        dh_dy_yo = 0.;
        d2h_dxdy_o = 0;
        value = 0;
        for (i = 0; i < n_uv_coords; ++i) {
            u = cc * grid_u[i];
            v = cc * grid_v[i];
            a = xo * u + yo * v;
            real_part = (h_F[i].real * cos(a) - h_F[i].imag * sin(a));
            imag_part = (h_F[i].real * sin(a) + h_F[i].imag * cos(a));
            value += real_part;
            dh_dx_xo -= imag_part * u;
            dh_dy_yo -= imag_part * v;
            d2h_dx2_xo -= real_part * u * u;
            d2h_dy2_yo -= real_part * v * v;
            d2h_dxdy_o -= real_part * u * v;
        }
        det = d2h_dx2_xo * d2h_dy2_yo - d2h_dxdy_o * d2h_dxdy_o;
        norm = sqrt(incx * incx + incy * incy);
    }
    *ho = value;
    *det_sigma = det;
    if (n_iter) *n_iter = iter;
    if (det <= 0) return 1;
    return 0;
}

/* likelihood.c:157-187 (FACET branch) */
double orl2_loglikelihood_r(double ee1, double ee2, double scale, unsigned long n_uv_coords,
                            const double *variance, const double *grid_u, const double *grid_v,
                            const orl2_complexd *visData, double *visM, int *n_iter)
{
    double L_er, ho, det_sigma;
    orl2_model_at_zero(ee1, ee2, scale, (unsigned int)n_uv_coords, grid_u, grid_v, variance, visM);
    int error = orl2_cross_correlation(n_uv_coords, grid_u, grid_v, visData, visM, &ho, &det_sigma,
                                       n_iter);
    if (error)
        L_er = -1.e10;
    else {
        double B = (ho * ho) * 0.5;
        L_er = orl2_marg_position(B);
        if (isnan(L_er))
            L_er = -1.e+10;
        else
            L_er -= 0.5 * log(det_sigma);
    }
    return L_er;
}

/* ======================================================================
 * marginalise_r.c: natural cubic spline (GSL cspline), linear interpolation and
 * QAG with the 41-point Gauss-Kronrod rule, restated from GSL 2.5 / QUADPACK.
 * ====================================================================== */
namespace {

struct Spline {
    int n = 0;
    bool linear = false;
    double x[32], y[32], c[32];
};

/* gsl_linalg_solve_symm_tridiag (LDL^T), as used by cspline_init */
static int solve_symm_tridiag(const double *diag, const double *offdiag, const double *b,
                              double *x, int N)
{
    double gamma[32], alpha[32], c[32], z[32];
    int status = 0;
    alpha[0] = diag[0];
    gamma[0] = (N > 1) ? offdiag[0] / alpha[0] : 0.;
    if (alpha[0] == 0) status = 1;
    for (int i = 1; i < N - 1; i++) {
        alpha[i] = diag[i] - offdiag[i - 1] * gamma[i - 1];
        gamma[i] = offdiag[i] / alpha[i];
        if (alpha[i] == 0) status = 1;
    }
    if (N > 1) alpha[N - 1] = diag[N - 1] - offdiag[N - 2] * gamma[N - 2];
    z[0] = b[0];
    for (int i = 1; i < N; i++) z[i] = b[i] - gamma[i - 1] * z[i - 1];
    for (int i = 0; i < N; i++) c[i] = z[i] / alpha[i];
    x[N - 1] = c[N - 1];
    if (N >= 2) {
        for (int i = N - 2; i >= 0; i--) x[i] = c[i] - gamma[i] * x[i + 1];
    }
    return status;
}

/* GSL cspline_init (natural boundary conditions) */
static int spline_init(Spline &s, const double *xa, const double *ya, int size, bool linear)
{
    s.n = size;
    s.linear = linear;
    for (int i = 0; i < size; i++) {
        s.x[i] = xa[i];
        s.y[i] = ya[i];
    }
    for (int i = 1; i < size; i++)
        if (!(xa[i - 1] < xa[i])) return 1;
    if (linear) return 0;
    int max_index = size - 1;
    int sys_size = max_index - 1;
    double diag[32], offdiag[32], g[32];
    s.c[0] = 0.0;
    s.c[max_index] = 0.0;
    for (int i = 0; i < sys_size; i++) {
        const double h_i = xa[i + 1] - xa[i];
        const double h_ip1 = xa[i + 2] - xa[i + 1];
        const double ydiff_i = ya[i + 1] - ya[i];
        const double ydiff_ip1 = ya[i + 2] - ya[i + 1];
        const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
        const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
        offdiag[i] = h_ip1;
        diag[i] = 2.0 * (h_ip1 + h_i);
        g[i] = 3.0 * (ydiff_ip1 * g_ip1 - ydiff_i * g_i);
    }
    if (sys_size == 1) {
        s.c[1] = g[0] / diag[0];
        return 0;
    }
    return solve_symm_tridiag(diag, offdiag, g, s.c + 1, sys_size);
}

/* gsl_interp_bsearch over the full range (the accelerator returns the same interval) */
static inline int bsearch_idx(const double *xa, double x, int lo, int hi)
{
    while (hi > lo + 1) {
        int i = (hi + lo) / 2;
        if (xa[i] > x)
            hi = i;
        else
            lo = i;
    }
    return lo;
}

/* gsl_spline_eval_e: 0 = success, 1 = domain error (y = NaN), 2 = dx <= 0 */
static inline int spline_eval(const Spline &s, double x, double *y)
{
    if (x < s.x[0] || x > s.x[s.n - 1]) {
        *y = NAN;
        return 1;
    }
    int index = bsearch_idx(s.x, x, 0, s.n - 1);
    double x_lo = s.x[index], x_hi = s.x[index + 1];
    double dx = x_hi - x_lo;
    if (!(dx > 0.0)) {
        *y = 0.0;
        return 2;
    }
    double y_lo = s.y[index], y_hi = s.y[index + 1];
    if (s.linear) {
        *y = y_lo + (x - x_lo) / dx * (y_hi - y_lo);
        return 0;
    }
    const double dy = y_hi - y_lo;
    double delx = x - x_lo;
    const double c_i = s.c[index];
    const double c_ip1 = s.c[index + 1];
    const double b_i = (dy / dx) - dx * (c_ip1 + 2.0 * c_i) / 3.0;
    const double d_i = (c_ip1 - c_i) / (3.0 * dx);
    *y = y_lo + delx * (b_i + delx * (c_i + delx * d_i));
    return 0;
}

/* marginalise_r.c:227-238 */
static double marf_impl(double x, void *params)
{
    const Spline *s = (const Spline *)params;
    double yval;
    if (spline_eval(*s, x, &yval))
        yval = 0.;
    else
        yval = exp(yval);
    return yval;
}

static double rescale_error(double err, const double result_abs, const double result_asc)
{
    err = fabs(err);
    if (result_asc != 0 && err != 0) {
        double scale = pow((200 * err / result_asc), 1.5);
        if (scale < 1)
            err = result_asc * scale;
        else
            err = result_asc;
    }
    if (result_abs > GSL_DBL_MIN_ / (50 * GSL_DBL_EPSILON_)) {
        double min_err = 50 * GSL_DBL_EPSILON_ * result_abs;
        if (min_err > err) err = min_err;
    }
    return err;
}

struct QagWork {
    unsigned long limit, size, nrmax, i, maximum_level;
    std::vector<double> alist, blist, rlist, elist;
    std::vector<unsigned long> order, level;
    explicit QagWork(unsigned long lim)
        : limit(lim), size(0), nrmax(0), i(0), maximum_level(0), alist(lim), blist(lim),
          rlist(lim), elist(lim), order(lim), level(lim) {}
};

static void qpsrt(QagWork &w)
{
    const unsigned long last = w.size - 1;
    const unsigned long limit = w.limit;
    double errmax, errmin;
    long i, k, top;
    unsigned long i_nrmax = w.nrmax;
    unsigned long i_maxerr = w.order[i_nrmax];
    if (last < 2) {
        w.order[0] = 0;
        w.order[1] = 1;
        w.i = i_maxerr;
        return;
    }
    errmax = w.elist[i_maxerr];
    while (i_nrmax > 0 && errmax > w.elist[w.order[i_nrmax - 1]]) {
        w.order[i_nrmax] = w.order[i_nrmax - 1];
        i_nrmax--;
    }
    if (last < (limit / 2 + 2))
        top = (long)last;
    else
        top = (long)(limit - last + 1);
    i = (long)i_nrmax + 1;
    while (i < top && errmax < w.elist[w.order[i]]) {
        w.order[i - 1] = w.order[i];
        i++;
    }
    w.order[i - 1] = i_maxerr;
    errmin = w.elist[last];
    k = top - 1;
    while (k > i - 2 && errmin >= w.elist[w.order[k]]) {
        w.order[k + 1] = w.order[k];
        k--;
    }
    w.order[k + 1] = last;
    i_maxerr = w.order[i_nrmax];
    w.i = i_maxerr;
    w.nrmax = i_nrmax;
}

static void qag_update(QagWork &w, double a1, double b1, double area1, double error1, double a2,
                       double b2, double area2, double error2)
{
    const unsigned long i_max = w.i;
    const unsigned long i_new = w.size;
    const unsigned long new_level = w.level[i_max] + 1;
    if (error2 > error1) {
        w.alist[i_max] = a2;
        w.rlist[i_max] = area2;
        w.elist[i_max] = error2;
        w.level[i_max] = new_level;
        w.alist[i_new] = a1;
        w.blist[i_new] = b1;
        w.rlist[i_new] = area1;
        w.elist[i_new] = error1;
        w.level[i_new] = new_level;
    } else {
        w.blist[i_max] = b1;
        w.rlist[i_max] = area1;
        w.elist[i_max] = error1;
        w.level[i_max] = new_level;
        w.alist[i_new] = a2;
        w.blist[i_new] = b2;
        w.rlist[i_new] = area2;
        w.elist[i_new] = error2;
        w.level[i_new] = new_level;
    }
    w.size++;
    if (new_level > w.maximum_level) w.maximum_level = new_level;
    qpsrt(w);
}

static inline bool subinterval_too_small(double a1, double a2, double b2)
{
    const double e = GSL_DBL_EPSILON_;
    const double u = GSL_DBL_MIN_;
    double tmp = (1 + 100 * e) * (fabs(a2) + 1000 * u);
    return fabs(a1) <= tmp && fabs(b2) <= tmp;
}

}  // namespace

/* gsl_integration_qk with the 41-point rule (n = 21) */
void orl2_qk41(double (*f)(double, void *), void *params, double a, double b, double *result,
               double *abserr, double *resabs, double *resasc)
{
    const int n = 21;
    double fv1[21], fv2[21];
    const double center = 0.5 * (a + b);
    const double half_length = 0.5 * (b - a);
    const double abs_half_length = fabs(half_length);
    const double f_center = f(center, params);
    double result_gauss = 0;
    double result_kronrod = f_center * ORL2_WGK41[n - 1];
    double result_abs = fabs(result_kronrod);
    double result_asc = 0;
    double mean = 0, err = 0;
    int j;
    for (j = 0; j < (n - 1) / 2; j++) {
        const int jtw = j * 2 + 1;
        const double abscissa = half_length * ORL2_XGK41[jtw];
        const double fval1 = f(center - abscissa, params);
        const double fval2 = f(center + abscissa, params);
        const double fsum = fval1 + fval2;
        fv1[jtw] = fval1;
        fv2[jtw] = fval2;
        result_gauss += ORL2_WG41[j] * fsum;
        result_kronrod += ORL2_WGK41[jtw] * fsum;
        result_abs += ORL2_WGK41[jtw] * (fabs(fval1) + fabs(fval2));
    }
    for (j = 0; j < n / 2; j++) {
        int jtwm1 = j * 2;
        const double abscissa = half_length * ORL2_XGK41[jtwm1];
        const double fval1 = f(center - abscissa, params);
        const double fval2 = f(center + abscissa, params);
        fv1[jtwm1] = fval1;
        fv2[jtwm1] = fval2;
        result_kronrod += ORL2_WGK41[jtwm1] * (fval1 + fval2);
        result_abs += ORL2_WGK41[jtwm1] * (fabs(fval1) + fabs(fval2));
    }
    mean = result_kronrod * 0.5;
    result_asc = ORL2_WGK41[n - 1] * fabs(f_center - mean);
    for (j = 0; j < n - 1; j++)
        result_asc += ORL2_WGK41[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    err = (result_kronrod - result_gauss) * half_length;
    result_kronrod *= half_length;
    result_abs *= abs_half_length;
    result_asc *= abs_half_length;
    *result = result_kronrod;
    *resabs = result_abs;
    *resasc = result_asc;
    *abserr = rescale_error(err, result_abs, result_asc);
}

/* gsl_integration_qag (qag.c) with GSL_INTEG_GAUSS41. Returns 0 on success, else the
 * GSL status: 14 (EROUND), 21 (ESING), 11 (EMAXITER), 5 (EFAILED), 4 (EINVAL). */
int orl2_qag41(double (*f)(double, void *), void *params, double a, double b, double epsabs,
               double epsrel, unsigned long limit, double *result, double *abserr,
               unsigned long *n_intervals)
{
    double area, errsum;
    double result0, abserr0, resabs0, resasc0;
    double tolerance;
    unsigned long iteration = 0;
    int roundoff_type1 = 0, roundoff_type2 = 0, error_type = 0;
    double round_off;
    QagWork w(limit);
    w.size = 0;
    w.nrmax = 0;
    w.i = 0;
    w.alist[0] = a;
    w.blist[0] = b;
    w.rlist[0] = 0.0;
    w.elist[0] = 0.0;
    w.order[0] = 0;
    w.level[0] = 0;
    w.maximum_level = 0;
    *result = 0;
    *abserr = 0;
    if (n_intervals) *n_intervals = 1;
    if (epsabs <= 0 && (epsrel < 50 * GSL_DBL_EPSILON_ || epsrel < 0.5e-28)) return 13;
    orl2_qk41(f, params, a, b, &result0, &abserr0, &resabs0, &resasc0);
    w.size = 1;
    w.rlist[0] = result0;
    w.elist[0] = abserr0;
    tolerance = fmax(epsabs, epsrel * fabs(result0));
    volatile double ro = 50 * GSL_DBL_EPSILON_ * resabs0;
    round_off = ro;
    if (abserr0 <= round_off && abserr0 > tolerance) {
        *result = result0;
        *abserr = abserr0;
        return 14;
    } else if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) {
        *result = result0;
        *abserr = abserr0;
        return 0;
    } else if (limit == 1) {
        *result = result0;
        *abserr = abserr0;
        return 11;
    }
    area = result0;
    errsum = abserr0;
    iteration = 1;
    do {
        double a1, b1, a2, b2;
        double a_i, b_i, r_i, e_i;
        double area1 = 0, area2 = 0, area12 = 0;
        double error1 = 0, error2 = 0, error12 = 0;
        double resasc1, resasc2;
        double resabs1, resabs2;
        a_i = w.alist[w.i];
        b_i = w.blist[w.i];
        r_i = w.rlist[w.i];
        e_i = w.elist[w.i];
        a1 = a_i;
        b1 = 0.5 * (a_i + b_i);
        a2 = b1;
        b2 = b_i;
        orl2_qk41(f, params, a1, b1, &area1, &error1, &resabs1, &resasc1);
        orl2_qk41(f, params, a2, b2, &area2, &error2, &resabs2, &resasc2);
        area12 = area1 + area2;
        error12 = error1 + error2;
        errsum += (error12 - e_i);
        area += area12 - r_i;
        if (resasc1 != error1 && resasc2 != error2) {
            double delta = r_i - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i) roundoff_type1++;
            if (iteration >= 10 && error12 > e_i) roundoff_type2++;
        }
        tolerance = fmax(epsabs, epsrel * fabs(area));
        if (errsum > tolerance) {
            if (roundoff_type1 >= 6 || roundoff_type2 >= 20) error_type = 2;
            if (subinterval_too_small(a1, a2, b2)) error_type = 3;
        }
        qag_update(w, a1, b1, area1, error1, a2, b2, area2, error2);
        iteration++;
    } while (iteration < limit && !error_type && errsum > tolerance);
    double sum = 0;
    for (unsigned long k = 0; k < w.size; k++) sum += w.rlist[k];
    *result = sum;
    *abserr = errsum;
    if (n_intervals) *n_intervals = w.size;
    if (errsum <= tolerance) return 0;
    if (error_type == 2) return 14;
    if (error_type == 3) return 21;
    if (iteration == limit) return 11;
    return 5;
}

/* marginalise_r.c:29-64 */
void orl2_set_posterior_values(int numR, const double *L_r, const double *rprior,
                               const double *Ro, double *xmarvals, double *ymarvals,
                               int *numvals)
{
    int num_marvals = 0;
    int nRo = 1;
    while (nRo < numR && (L_r[nRo] <= -1.e+10)) nRo++;
    if (nRo < numR) {
        double likelval;
        for (nRo = 1; nRo < numR; nRo++) {
            likelval = L_r[nRo];
            if (likelval > -1.e+10) {
                xmarvals[num_marvals] = Ro[nRo];
                ymarvals[num_marvals] = likelval + rprior[nRo];
                num_marvals++;
            }
        }
    }
    *numvals = num_marvals;
}

/* marginalise_r.c:81-223 (exit(EXIT_FAILURE) paths return NaN instead) */
double orl2_marginalise_posterior_r(int num_marvals, double *xmarvals, double *ymarvals)
{
    double sump, error;
    int gsl_status;
    double ymax = -10.0;
    int ymaxpos = 0;
    int ii;
    for (ii = 0; ii < num_marvals; ii++) {
        if (ymarvals[ii] > ymax) {
            ymax = ymarvals[ii];
            ymaxpos = ii;
        }
    }
    double xmin, xmax;
    xmin = xmax = xmarvals[ymaxpos];
    if (ymaxpos < num_marvals - 1) xmax = xmarvals[ymaxpos + 1];
    if (ymaxpos > 0) xmin = xmarvals[ymaxpos - 1];
    for (ii = 0; ii < num_marvals; ii++) {
        if (xmarvals[ii] < xmin) xmin = xmarvals[ii];
        if (xmarvals[ii] > xmax) xmax = xmarvals[ii];
        ymarvals[ii] -= ymax;
    }
    if (xmax <= xmin) return NAN;
    Spline s;
    if (spline_init(s, xmarvals, ymarvals, num_marvals, false)) return NAN;
    gsl_status = orl2_qag41(marf_impl, &s, xmin, xmax, 0., 1.e-3, 1000, &sump, &error, NULL);
    if (gsl_status) {
        if (spline_init(s, xmarvals, ymarvals, num_marvals, true)) return NAN;
        gsl_status = orl2_qag41(marf_impl, &s, xmin, xmax, 0., 1.e-3, 1000, &sump, &error, NULL);
    }
    if (gsl_status) return -1.e10;
    return log(sump) + ymax;
}

int orl2_cspline_eval_many(int num, const double *x, const double *y, int nq, const double *xq,
                           double *yq, int linear)
{
    Spline s;
    if (spline_init(s, x, y, num, linear != 0)) return 1;
    for (int i = 0; i < nq; i++) spline_eval(s, xq[i], &yq[i]);
    return 0;
}

/* ======================================================================
 * likelihood.c:106-149 loglikelihood (omp over r) for the facet in `par`.
 * ====================================================================== */
double orl2_loglikelihood_detail(orl2_params *par, double ee1, double ee2, int *error,
                                 double *L_r_out, int *n_iter_out)
{
    int numR = par->numr;
    double L_e;
    int numvals;
    double xmarvals[40], ymarvals[40], L_r[40];
    int n_iter[40];
    *error = 0;
    std::vector<double> mod((size_t)(numR - 1) * par->ncoords + 1);
#pragma omp parallel for schedule(dynamic, 1)
    for (int nRo = 1; nRo < numR; nRo++) {
        unsigned long index = (unsigned long)(nRo - 1) * (par->ncoords);
        L_r[nRo] = orl2_loglikelihood_r(ee1, ee2, (par->ro)[nRo], par->ncoords, par->sigma2,
                                        par->uu, par->vv, par->data, &mod[index], &n_iter[nRo]);
    }
    orl2_set_posterior_values(numR, L_r, par->rprior, par->ro, xmarvals, ymarvals, &numvals);
    if (numvals > 4)
        L_e = orl2_marginalise_posterior_r(numvals, xmarvals, ymarvals);
    else {
        *error = 1;
        L_e = -1.e+10;
    }
    if (L_e == -1.e+10) *error = 1;
    if (L_r_out)
        for (int k = 1; k < numR; k++) L_r_out[k] = L_r[k];
    if (n_iter_out)
        for (int k = 1; k < numR; k++) n_iter_out[k] = n_iter[k];
    return L_e;
}

double orl2_loglikelihood(orl2_params *par, double ee1, double ee2, int *error)
{
    return orl2_loglikelihood_detail(par, ee1, ee2, error, NULL, NULL);
}

/* ======================================================================
 * Nelder-Mead simplex, restated from GSL 2.5 multimin/simplex2.c (nmsimplex2), 2-D.
 * ====================================================================== */
namespace {

/* cblas dnrm2 with the scale/ssq recurrence */
static double dnrm2(const double *X, int N)
{
    double scale = 0.0, ssq = 1.0;
    if (N <= 0) return 0;
    if (N == 1) return fabs(X[0]);
    for (int i = 0; i < N; i++) {
        const double x = X[i];
        if (x != 0.0) {
            const double ax = fabs(x);
            if (scale < ax) {
                ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
                scale = ax;
            } else {
                ssq += (ax / scale) * (ax / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}

template <class F>
struct NMSimplex2 {
    static const int N = 2, P = 3;
    double x1[P][N];
    double y1[P];
    double center[N];
    double S2;
    F &f;
    explicit NMSimplex2(F &fn) : f(fn) {}

    void compute_center()
    {
        for (int j = 0; j < N; j++) center[j] = 0.;
        for (int i = 0; i < P; i++)
            for (int j = 0; j < N; j++) center[j] += 1.0 * x1[i][j];
        const double alpha = 1.0 / P;
        for (int j = 0; j < N; j++) center[j] *= alpha;
    }
    double compute_size()
    {
        double ss = 0.0;
        for (int i = 0; i < P; i++) {
            double s[N];
            for (int j = 0; j < N; j++) s[j] = x1[i][j];
            for (int j = 0; j < N; j++) s[j] += -1.0 * center[j];
            double t = dnrm2(s, N);
            ss += t * t;
        }
        S2 = (ss / P);
        return sqrt(ss / P);
    }
    double try_corner_move(double coeff, int corner, double *xc)
    {
        double alpha = (1 - coeff) * P / (P - 1.0);
        double beta = (P * coeff - 1.0) / (P - 1.0);
        for (int j = 0; j < N; j++) xc[j] = center[j];
        for (int j = 0; j < N; j++) xc[j] *= alpha;
        for (int j = 0; j < N; j++) xc[j] += beta * x1[corner][j];
        return f(xc[0], xc[1]);
    }
    void update_point(int i, const double *x, double val)
    {
        double delta[N], xmc[N];
        for (int j = 0; j < N; j++) delta[j] = x[j];
        for (int j = 0; j < N; j++) delta[j] += -1.0 * x1[i][j];
        for (int j = 0; j < N; j++) xmc[j] = x1[i][j];
        for (int j = 0; j < N; j++) xmc[j] += -1.0 * center[j];
        {
            double d = dnrm2(delta, N);
            double xmcd = 0.0;
            for (int j = 0; j < N; j++) xmcd += xmc[j] * delta[j];
            S2 += (2.0 / P) * xmcd + ((P - 1.0) / P) * (d * d / P);
        }
        {
            double alpha = 1.0 / P;
            for (int j = 0; j < N; j++) center[j] += -alpha * x1[i][j];
            for (int j = 0; j < N; j++) center[j] += alpha * x[j];
        }
        for (int j = 0; j < N; j++) x1[i][j] = x[j];
        y1[i] = val;
    }
    int contract_by_best(int best)
    {
        int status = 0;
        for (int i = 0; i < P; i++) {
            if (i != best) {
                for (int j = 0; j < N; j++) x1[i][j] = 0.5 * (x1[i][j] + x1[best][j]);
                double newval = f(x1[i][0], x1[i][1]);
                y1[i] = newval;
                if (!isfinite(newval)) status = 1;
            }
        }
        compute_center();
        compute_size();
        return status;
    }
    /* nmsimplex_set; returns size */
    int set(const double *x, const double *step, double *size)
    {
        double val = f(x[0], x[1]);
        if (!isfinite(val)) return 1;
        for (int j = 0; j < N; j++) x1[0][j] = x[j];
        y1[0] = val;
        for (int i = 0; i < N; i++) {
            double xtemp[N];
            for (int j = 0; j < N; j++) xtemp[j] = x[j];
            xtemp[i] = x[i] + step[i];
            val = f(xtemp[0], xtemp[1]);
            if (!isfinite(val)) return 1;
            for (int j = 0; j < N; j++) x1[i + 1][j] = xtemp[j];
            y1[i + 1] = val;
        }
        compute_center();
        *size = compute_size();
        return 0;
    }
    /* nmsimplex_iterate */
    int iterate(double *x, double *size, double *fval)
    {
        double xc[N], xc2[N];
        int hi, s_hi, lo;
        double dhi, ds_hi, dlo, val, val2;
        dhi = dlo = y1[0];
        hi = 0;
        lo = 0;
        ds_hi = y1[1];
        s_hi = 1;
        for (int i = 1; i < P; i++) {
            val = y1[i];
            if (val < dlo) {
                dlo = val;
                lo = i;
            } else if (val > dhi) {
                ds_hi = dhi;
                s_hi = hi;
                dhi = val;
                hi = i;
            } else if (val > ds_hi) {
                ds_hi = val;
                s_hi = i;
            }
        }
        int status = 0;
        val = try_corner_move(-1.0, hi, xc);
        if (isfinite(val) && val < y1[lo]) {
            val2 = try_corner_move(-2.0, hi, xc2);
            if (isfinite(val2) && val2 < y1[lo])
                update_point(hi, xc2, val2);
            else
                update_point(hi, xc, val);
        } else if (!isfinite(val) || val > y1[s_hi]) {
            if (isfinite(val) && val <= y1[hi]) update_point(hi, xc, val);
            val2 = try_corner_move(0.5, hi, xc2);
            if (isfinite(val2) && val2 <= y1[hi])
                update_point(hi, xc2, val2);
            else
                status = contract_by_best(lo, xc);
        } else {
            update_point(hi, xc, val);
        }
        if (status) return 5; /* GSL_EFAILED */
        /* gsl_vector_min_index */
        lo = 0;
        double mn = y1[0];
        for (int i = 0; i < P; i++) {
            if (y1[i] < mn) {
                mn = y1[i];
                lo = i;
            }
            if (isnan(y1[i])) {
                lo = i;
                break;
            }
        }
        for (int j = 0; j < N; j++) x[j] = x1[lo][j];
        *fval = y1[lo];
        if (S2 > 0)
            *size = sqrt(S2);
        else
            *size = compute_size();
        return 0;
    }
    int contract_by_best(int best, double *) { return contract_by_best(best); }
};

struct CbLik {
    orl2_loglik_cb cb;
    void *ctx;
    int n_evals;
    /* f_likelihood, likelihood.c:87-99 */
    double operator()(double ee1, double ee2)
    {
        double L_e;
        int error = 0;
        if (ee1 * ee1 + ee2 * ee2 < 0.65) {
            L_e = cb(ee1, ee2, &error, ctx);
            n_evals++;
        } else
            L_e = -1.e+10;
        return -L_e;
    }
};

struct CbF2 {
    orl2_f2_cb f;
    void *ctx;
    double operator()(double a, double b) { return f(a, b, ctx); }
};

/* likelihood.c:428-554, with loglikelihood supplied by the callback */
static void sampling_impl(orl2_loglik_cb cb, void *ctx, double *mes_e1, double *mes_e2,
                          double maxL, int np_max, double *var_e1, double *var_e2,
                          double *oneDimvar, int *n_evals)
{
    int error = 0;
    int ie1, ie2, np;
    double totL, L_e, average_1, average_2, cov, var1, var2, x_e1, x_e2;
    double xL1, xL2, mod2 = 0;
    int evals = 0;
    double L = 1.;
    double threshold = 0.05;
    double max_e1 = *mes_e1;
    double max_e2 = *mes_e2;
    average_1 = max_e1;
    average_2 = max_e2;
    var1 = max_e1 * max_e1;
    var2 = max_e2 * max_e2;
    cov = max_e1 * max_e2;
    totL = 1.;
    np = 1;
    float sampling = 0.05;
    if (max_e1 < 0) sampling = -0.05;
    int k1 = 0;
    x_e1 = max_e1;
    while (L > threshold && mod2 < 1. && !error) {
        x_e1 += sampling;
        mod2 = x_e1 * x_e1 + max_e2 * max_e2;
        if (mod2 < 1.) {
            L_e = cb(x_e1, max_e2, &error, ctx);
            evals++;
            if (!error) {
                L = exp(L_e - maxL);
                k1++;
            }
        }
    }
    if (L < threshold && k1 > 1) k1--;
    L = 1.;
    int k2 = 0;
    sampling = 0.05;
    if (max_e2 < 0) sampling = -0.05;
    x_e2 = max_e2;
    while (L > threshold && mod2 < 1. && !error) {
        x_e2 += sampling;
        mod2 = x_e2 * x_e2 + max_e1 * max_e1;
        if (mod2 < 1.) {
            L_e = cb(max_e1, x_e2, &error, ctx);
            evals++;
            if (!error) {
                L = exp(L_e - maxL);
                k2++;
            }
        }
    }
    if (L < threshold && k2 > 1) k2--;
    sampling = 0.05;
    float edim1 = k1 * sampling;
    float edim2 = k2 * sampling;
    float start1 = edim1;
    float start2 = edim2;
    while (np < np_max && sampling > E_RES) {
        for (ie1 = -start1 / sampling; ie1 <= start1 / sampling; ie1 += 2) {
            for (ie2 = -start2 / sampling; ie2 <= start2 / sampling; ie2 += 2) {
                if (ie1 != 0 || ie2 != 0) {
                    x_e1 = max_e1 + ie1 * sampling;
                    x_e2 = max_e2 + ie2 * sampling;
                    mod2 = x_e1 * x_e1 + x_e2 * x_e2;
                    if (mod2 < 1.) {
                        L_e = cb(x_e1, x_e2, &error, ctx);
                        evals++;
                        if (!error) {
                            L = exp(L_e - maxL);
                            if (L >= threshold && !isinf(L)) {
                                np++;
                                xL1 = x_e1 * L;
                                xL2 = x_e2 * L;
                                average_1 += xL1;
                                average_2 += xL2;
                                var1 += x_e1 * xL1;
                                var2 += x_e2 * xL2;
                                cov += x_e1 * xL2;
                                totL += L;
                            }
                        }
                    }
                }
            }
        }
        sampling *= 0.5;
        start1 = edim1 - sampling;
        start2 = edim2 - sampling;
    }
    average_1 /= totL;
    average_2 /= totL;
    var1 /= totL;
    var1 -= average_1 * average_1;
    var2 /= totL;
    var2 -= average_2 * average_2;
    cov /= totL;
    cov -= average_1 * average_2;
    *mes_e1 = average_1;
    *mes_e2 = average_2;
    *var_e1 = var1;
    *var_e2 = var2;
    *oneDimvar = sqrt((var1) * (var2)-cov * cov);
    if (n_evals) *n_evals = evals;
}

static double par_cb(double e1, double e2, int *error, void *ctx)
{
    return orl2_loglikelihood((orl2_params *)ctx, e1, e2, error);
}

}  // namespace

int orl2_nmsimplex2_minimize(orl2_f2_cb f, void *ctx, const double *x0, const double *step,
                             double tol, int max_iter, double *x_out, double *fval,
                             double *size_out, double *trace_best)
{
    CbF2 fn{f, ctx};
    NMSimplex2<CbF2> s(fn);
    double size;
    double x[2] = {x0[0], x0[1]};
    if (s.set(x0, step, &size)) return -1;
    int iter = 0, status;
    double fv = 0.;
    do {
        iter++;
        status = s.iterate(x, &size, &fv);
        if (trace_best) {
            trace_best[2 * (iter - 1)] = x[0];
            trace_best[2 * (iter - 1) + 1] = x[1];
        }
        if (status) break;
        status = (size < tol) ? 0 : -2;
    } while (status == -2 && iter < max_iter);
    x_out[0] = x[0];
    x_out[1] = x[1];
    *fval = fv;
    *size_out = size;
    return iter;
}

void orl2_likelihood_sampling_cb(orl2_loglik_cb cb, void *ctx, double *mes_e1, double *mes_e2,
                                 double maxL, int np_max, double *var_e1, double *var_e2,
                                 double *oneDimvar, int *n_evals)
{
    sampling_impl(cb, ctx, mes_e1, mes_e2, maxL, np_max, var_e1, var_e2, oneDimvar, n_evals);
}

void orl2_likelihood_sampling(double *mes_e1, double *mes_e2, double maxL, orl2_params *par,
                              int np_max, double *var_e1, double *var_e2, double *oneDimvar)
{
    sampling_impl(par_cb, par, mes_e1, mes_e2, maxL, np_max, var_e1, var_e2, oneDimvar, NULL);
}

/* galaxy_fitting.c:41-117 with the log-likelihood supplied by a callback */
int orl2_source_fitting_cb(orl2_loglik_cb cb, void *ctx, double start_e1, double start_e2,
                           double step, double *mes_e1, double *mes_e2, double *var_e1,
                           double *var_e2, double *oneDimvar, double *maxL,
                           orl2_fit_trace *trace)
{
    int np_max = NP_MAX;
    CbLik fn{cb, ctx, 0};
    NMSimplex2<CbLik> s(fn);
    double x[2] = {start_e1, start_e2};
    double sx[2] = {step, step};
    double size = 0., fval = 0.;
    s.set(x, sx, &size);
    int iter = 0;
    int status;
    do {
        iter++;
        status = s.iterate(x, &size, &fval);
        if (status) break;
        status = (size < TOL) ? 0 : -2; /* gsl_multimin_test_size */
    } while (status == -2 && iter < 50 && fval < 0.);
    *mes_e1 = x[0];
    *mes_e2 = x[1];
    *maxL = -fval;
    if (trace) {
        trace->nm_iter = iter;
        trace->nm_evals = fn.n_evals;
        trace->nm_e1 = x[0];
        trace->nm_e2 = x[1];
        trace->nm_size = size;
    }
    if ((*mes_e1) * (*mes_e1) + (*mes_e2) * (*mes_e2) > 0.65) *maxL = -1e+10;
    int error = 0;
    *var_e1 = 0.;
    *var_e2 = 0.;
    *oneDimvar = 0.;
    int samp_evals = 0;
    if (*maxL > -1.e+10) {
        sampling_impl(cb, ctx, mes_e1, mes_e2, *maxL, np_max, var_e1, var_e2, oneDimvar,
                      &samp_evals);
        if (*var_e1 < VAR || *var_e2 < VAR || *oneDimvar < VAR) error = 1;
    } else {
        error = 1;
    }
    if (trace) trace->n_evals = fn.n_evals + samp_evals;
    return error;
}

int orl2_source_fitting(orl2_params *par, double *mes_e1, double *mes_e2, double *var_e1,
                        double *var_e2, double *oneDimvar, double *maxL)
{
    return orl2_source_fitting_cb(par_cb, par, 0., 0., 0.1, mes_e1, mes_e2, var_e1, var_e2,
                                  oneDimvar, maxL, NULL);
}

/* ======================================================================
 * data_simulation.c:131-165 sky_model
 * ====================================================================== */
void orl2_sky_model(const double *wavenumbers, const double *spec, unsigned int num_channels,
                    unsigned int n_gal, const double *gflux, const double *gscale,
                    const double *l, const double *m, unsigned long num_coords,
                    const double *uu_metres, const double *vv_metres, const double *ww_metres,
                    orl2_complexd *visGal, orl2_complexd *visMod)
{
    for (unsigned int g = 0; g < n_gal; g++) {
        double R_mu = gscale[g];
        double l0 = l[g], m0 = m[g];
#pragma omp parallel for schedule(static)
        for (unsigned int ch = 0; ch < num_channels; ch++) {
            unsigned long ch_vis = (unsigned long)ch * num_coords;
            orl2_data_galaxy_vis(spec[ch], wavenumbers[ch], 0., 0., R_mu, gflux[g], l0, m0,
                                 num_coords, uu_metres, vv_metres, ww_metres, &visGal[ch_vis]);
            for (unsigned long i = ch_vis; i < ch_vis + num_coords; i++) {
                visMod[i].real += visGal[i].real;
                visMod[i].imag += visGal[i].imag;
            }
        }
    }
}

/* ======================================================================
 * data_processing.c:49-142 (batch == 1) and data_processing-mpi.c:65-216 (batch > 1):
 * extract `batch` consecutive galaxies against the same residual, fit them, then subtract
 * the good fits in order.  Followed by the bad-list re-fit pass (RadioLensfit2.cpp:597-606).
 * ====================================================================== */
namespace {
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
}  // namespace

int orl2_data_processing(unsigned int ngal, int batch, double len, unsigned int nch,
                         const double *wavenumbers, const double *spec, unsigned long nrow,
                         const double *u, const double *v, const double *w,
                         const float *sigma2_vis, const bool *flag, orl2_complexd *visData,
                         orl2_complexd *visSkyMod, const double *l, const double *m,
                         const double *gflux, const double *gscale, orl2_gal_result *results,
                         int do_refit, orl2_run_stats *stats)
{
    if (batch < 1) batch = 1;
    double Ro[40];
    int numR = orl2_r_grid(Ro);
    unsigned long nvis = (unsigned long)nch * nrow;
    std::vector<orl2_complexd> visGal(nvis);
    unsigned int fmax = orl2_facet_size(RMAX, len);
    unsigned long ncells = (unsigned long)fmax * fmax;
    double t_ext = 0., t_fit = 0., t_sub = 0.;

    struct Slot {
        std::vector<orl2_complexd> data;
        std::vector<double> sigma2, sum_w, uu, vv;
        double rprior[40];
        unsigned int facet = 0, ncoords = 0;
    };
    std::vector<Slot> slots(batch);
    for (auto &s : slots) {
        s.data.resize(ncells);
        s.sigma2.resize(ncells);
        s.sum_w.resize(ncells);
        s.uu.resize(ncells);
        s.vv.resize(ncells);
    }
    std::vector<unsigned int> bad_list;
    int n_bad_first = 0, n_bad_final = 0;

    for (int pass = 0; pass < (do_refit ? 2 : 1); pass++) {
        std::vector<unsigned int> todo;
        if (pass == 0)
            for (unsigned int g = 0; g < ngal; g++) todo.push_back(g);
        else
            todo = bad_list;
        std::vector<unsigned int> new_bad;
        for (size_t b0 = 0; b0 < todo.size(); b0 += batch) {
            size_t nb = std::min((size_t)batch, todo.size() - b0);
            double t0 = now_s();
            for (size_t k = 0; k < nb; k++) {
                unsigned int gal = todo[b0 + k];
                Slot &s = slots[k];
                double flux = gflux[gal];
                double mu = orl2_scale_mean(flux);
                double R_mu = gscale[gal];
                for (int nRo = 1; nRo < numR; nRo++) s.rprior[nRo] = orl2_rfunc(mu, R_STD, Ro[nRo]);
                s.facet = orl2_facet_size(R_mu, len);
                orl2_source_extraction(s.facet, nch, wavenumbers, spec, s.data.data(),
                                       s.sigma2.data(), s.sum_w.data(), l[gal], m[gal], flux,
                                       R_mu, 0., 0., visSkyMod, visData, visGal.data(),
                                       sigma2_vis, flag, nrow, u, v, w, len);
                s.ncoords = orl2_facet_coords(s.uu.data(), s.vv.data(), len, s.facet,
                                              s.sum_w.data());
            }
            double t1 = now_s();
            t_ext += t1 - t0;
            for (size_t k = 0; k < nb; k++) {
                unsigned int gal = todo[b0 + k];
                Slot &s = slots[k];
                orl2_params par;
                std::vector<double> mod(1);
                par.numr = numR;
                par.ro = Ro;
                par.rprior = s.rprior;
                par.nchannels = nch;
                par.band_factor = 0.;
                par.acc_time = 0.;
                par.spec = (double *)spec;
                par.wavenumbers = (double *)wavenumbers;
                par.uu = s.uu.data();
                par.vv = s.vv.data();
                par.l0 = l[gal];
                par.m0 = m[gal];
                par.data = s.data.data();
                par.sigma2 = s.sigma2.data();
                par.mod = mod.data();
                par.facet = s.facet;
                par.ncoords = s.ncoords;
                orl2_gal_result &r = results[gal];
                orl2_fit_trace tr;
                r.error = orl2_source_fitting_cb(par_cb, &par, 0., 0., 0.1, &r.e1, &r.e2,
                                                 &r.var_e1, &r.var_e2, &r.oneDimvar, &r.maxL,
                                                 &tr);
                r.pass = pass;
                r.facet = s.facet;
                r.ncoords = s.ncoords;
                r.n_evals = tr.n_evals;
                r.subtracted = 0;
            }
            double t2 = now_s();
            t_fit += t2 - t1;
            for (size_t k = 0; k < nb; k++) {
                unsigned int gal = todo[b0 + k];
                orl2_gal_result &r = results[gal];
                if (r.error) {
                    if (pass == 0)
                        new_bad.push_back(gal);
                    else {
                        r.e1 = 0.;
                        r.e2 = 0.;
                    }
                    continue;
                }
                double flux = gflux[gal], R_mu = gscale[gal], l0 = l[gal], m0 = m[gal];
#pragma omp parallel for schedule(static)
                for (unsigned int ch = 0; ch < nch; ch++) {
                    unsigned long ch_vis = (unsigned long)ch * nrow;
                    orl2_data_galaxy_vis(spec[ch], wavenumbers[ch], 0., 0., R_mu, flux, l0, m0,
                                         nrow, u, v, w, &visGal[ch_vis]);
                    for (unsigned long i = ch_vis; i < ch_vis + nrow; i++) {
                        visSkyMod[i].real -= visGal[i].real;
                        visSkyMod[i].imag -= visGal[i].imag;
                    }
                    orl2_data_galaxy_vis(spec[ch], wavenumbers[ch], r.e1, r.e2, R_mu, flux, l0,
                                         m0, nrow, u, v, w, &visGal[ch_vis]);
                    for (unsigned long i = ch_vis; i < ch_vis + nrow; i++) {
                        visData[i].real -= visGal[i].real;
                        visData[i].imag -= visGal[i].imag;
                    }
                }
                r.subtracted = 1;
            }
            t_sub += now_s() - t2;
        }
        if (pass == 0) {
            bad_list = new_bad;
            n_bad_first = (int)new_bad.size();
            n_bad_final = n_bad_first;
        } else {
            int nb = 0;
            for (unsigned int g : todo)
                if (results[g].error) nb++;
            n_bad_final = nb;
        }
    }
    if (stats) {
        stats->extraction_s = t_ext;
        stats->fitting_s = t_fit;
        stats->subtraction_s = t_sub;
        stats->n_bad_first = n_bad_first;
        stats->n_bad_final = n_bad_final;
    }
    return n_bad_first;
}

