// rl2_math.h -- scalar arithmetic of the shape-measurement path, shared by the gfx950
// kernels and the host driver.  Each helper keeps the reference's operation order so the
// FP64 path rounds like the reference (the library is built with -ffp-contract=off; FMAs
// appear only where written explicitly).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define RL2_HD __host__ __device__ inline
#if defined(__HIP_DEVICE_COMPILE__)
#define RL2_CONST __constant__ static const
#else
#define RL2_CONST static const
#endif

namespace rl2 {

// datatype.h:28-42, default_params.h:28-79
constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = 299792458.0;
constexpr double kTwoPi = 2 * kPi;  // the reference's (2*PI) and (2.*PI)
constexpr int kPsfNat = 38;
constexpr double kRmin = 0.3, kRmax = 3.5;
constexpr int kNumR = 29;         // buffer size of the r grid (23 models used)
constexpr int kMaxR = 32;
constexpr double kAdd = -0.93, kEsp = 0.33, kRstd = 0.3136;
constexpr double kTol = 1.0e-3;   // simplex size tolerance
constexpr int kNpMax = 30;
constexpr double kEres = 3.0e-3;
constexpr double kVar = 1e-4;
constexpr double kBad = -1.e+10;  // sentinel log-likelihood
constexpr double kDblEps = 2.2204460492503131e-16;
constexpr double kDblMin = 2.2250738585072014e-308;

// `x *= ARCS2RAD` expands to x *= PI/648000.0 (the right operand is evaluated first).
RL2_HD double arcs2rad_scale(double x) { return x * (kPi / 648000.0); }

// evaluate_uv_grid.c:50-55: 2*ceil(len*PSF_NAT*theta*PI/648000.0), evaluated left to right.
RL2_HD unsigned int facet_size(double theta_med, double len)
{
    return (unsigned int)(2.0 * ceil(len * kPsfNat * theta_med * kPi / 648000.0));
}

// distributions.c:59-62
RL2_HD double scale_mean(double flux) { return kAdd + kEsp * log(flux); }

// distributions.c:72-87 (r > 0 always on the grid)
RL2_HD double rfunc(double mu, double sigma, double r)
{
    double x = (log(r) - mu) / sigma;
    return -(x * x) * 0.5 - log(r * sigma);
}

// galaxy_visibilities.c:48-53: FT of the exponential disk, (1 + s*w*|k|^2)^-3/2.
RL2_HD double sersic_shape(double k1, double k2, double scale_factor, double wave_factor)
{
    double den = 1. + scale_factor * wave_factor * (k1 * k1 + k2 * k2);
    return 1. / (den * sqrt(den));
}

// evaluate_uv_grid.c:75-77, 166-169: pillbox cell of one visibility (exact IEEE
// mul/add/div, no contraction).  Returns false when pc >= facet^2.
RL2_HD bool cell_index(double u, double v, double inv_lambda, double len, double inc,
                       unsigned int facet, uint32_t *pc_out)
{
    double uu = u * inv_lambda;
    double vv = v * inv_lambda;
    unsigned int pu = (unsigned int)((uu + len) / inc);
    unsigned int pv = (unsigned int)((vv + len) / inc);
    unsigned long long pc = (unsigned long long)pv * facet + pu;
    *pc_out = (uint32_t)pc;
    return pc < (unsigned long long)facet * facet;
}

// sin(2 pi t), cos(2 pi t) in FP64.  t - rint(t) and the octant split y = r - q/4
// (q = rint(4 r)) are exact, |y| <= 1/8; theta = 2 pi y; Taylor series of sin and cos to
// degree 15 / 18 in Horner form with fused multiply-adds (truncation below 5e-17 and
// 3e-18 on |theta| <= pi/4).  The phase arrives in revolutions, so no multiple of 2 pi is
// ever subtracted.
RL2_HD void sincos_rev64(double t, double &sn, double &cs)
{
    const double r = t - rint(t);
    const double q = rint(4. * r);
    const double y = r - 0.25 * q;
    const double th = 6.283185307179586477 * y;
    const double z = th * th;
    double ps = -1. / 1307674368000.;
    ps = fma(ps, z, 1. / 6227020800.);
    ps = fma(ps, z, -1. / 39916800.);
    ps = fma(ps, z, 1. / 362880.);
    ps = fma(ps, z, -1. / 5040.);
    ps = fma(ps, z, 1. / 120.);
    ps = fma(ps, z, -1. / 6.);
    const double s = fma(th * z, ps, th);
    double pc = -1. / 6402373705728000.;
    pc = fma(pc, z, 1. / 20922789888000.);
    pc = fma(pc, z, -1. / 87178291200.);
    pc = fma(pc, z, 1. / 479001600.);
    pc = fma(pc, z, -1. / 3628800.);
    pc = fma(pc, z, 1. / 40320.);
    pc = fma(pc, z, -1. / 720.);
    pc = fma(pc, z, 1. / 24.);
    pc = fma(pc, z, -0.5);
    const double c = fma(pc, z, 1.);
    const int k = ((int)q) & 3;
    sn = (k == 0) ? s : (k == 1) ? c : (k == 2) ? -s : -c;
    cs = (k == 0) ? c : (k == 1) ? -s : (k == 2) ? -c : s;
}

// cell_index (evaluate_uv_grid.c:166-178) with the two divisions by the cell size replaced
// by products with its reciprocal.  The product is within a few ulp of the correctly
// rounded quotient, so their integer parts can only differ when the quotient lies within
// that distance of an integer; there (|q - rint(q)| < 1e-9, far above the few-ulp error
// for quotients below 2^21) the exact division is taken, so the cell is always the
// reference's.
RL2_HD bool cell_index_fast(double u, double v, double inv_lambda, double len, double inc,
                            double inv_inc, unsigned int facet, uint32_t *pc_out)
{
    const double a = u * inv_lambda + len;
    const double b = v * inv_lambda + len;
    double qa = a * inv_inc, qb = b * inv_inc;
    if (fabs(qa - rint(qa)) < 1e-9) qa = a / inc;
    if (fabs(qb - rint(qb)) < 1e-9) qb = b / inc;
    const unsigned int pu = (unsigned int)qa;
    const unsigned int pv = (unsigned int)qb;
    const unsigned long long pc = (unsigned long long)pv * facet + pu;
    *pc_out = (uint32_t)pc;
    return pc < (unsigned long long)facet * facet;
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw, SC'11): ten rounds of
// two 32x32 -> 64-bit products with the Weyl key schedule.  The simulated noise draws its
// uniforms from it, one counter per visibility.
RL2_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
RL2_HD void philox4x32_10(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3], k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        const uint32_t hi0 = mulhi32(M0, c0), lo0 = M0 * c0;
        const uint32_t hi1 = mulhi32(M1, c2), lo1 = M1 * c2;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Exponential integral Ei(x), x > 0: series gamma + ln x + sum x^k/(k k!) for x <= 40,
// asymptotic e^x/x sum k!/x^k above (replaces gsl_sf_expint_Ei_e, likelihood.c:409).
RL2_HD double expint_ei(double x)
{
    const double euler = 0.57721566490153286060651209008240243;
    if (x <= 40.) {
        double sum = 0., term = 1.;
        for (int k = 1; k < 1000; k++) {
            term *= x / k;
            double t = term / k;
            sum += t;
            if (t < 2.220446049250313e-16 * 0.25 * sum) break;
        }
        return euler + log(x) + sum;
    }
    double sum = 1., term = 1.;
    for (int k = 1; k < 200; k++) {
        double prev = term;
        term *= k / x;
        if (term >= prev) break;
        sum += term;
        if (term < 2.220446049250313e-16 * 0.25 * sum) break;
    }
    return exp(x) / x * sum;
}

// likelihood.c:376-421: position marginalisation of exp(B exp(-r^2)) over the 95% disk.
RL2_HD double marg_position(double x)
{
    const double Delta = 0.5 * 5.991;
    if (x <= Delta) return kBad;
    if (x > 100.) return x - log(Delta) - exp(-Delta);
    if ((x - Delta) < 0.0001) x = Delta + 0.0001;
    double yc = expint_ei(x), ys = expint_ei(x - Delta);
    return log((yc - ys) / log(x / (x - Delta)));
}

// ---------------------------------------------------------------------------------------
// Scalelength marginalisation (marginalise_r.c:81-223): natural cubic spline of the
// log-posterior, exp() integrated by adaptive Gauss-Kronrod-41 (QAG, epsrel 1e-3,
// limit 1000), linear interpolation as the fallback.
// ---------------------------------------------------------------------------------------
struct Spline {
    int n;
    int linear;
    double x[kMaxR], y[kMaxR], c[kMaxR];
};

// natural cubic spline: tridiagonal system solved by LDL^T (GSL cspline_init semantics);
// ws: 7 * kMaxR doubles of workspace (LDS on the device, so nothing spills to scratch)
RL2_HD int spline_init_ws(Spline &s, const double *xa, const double *ya, int size, int linear,
                          double *ws)
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
    const int max_index = size - 1, sys = max_index - 1;
    double *diag = ws, *off = ws + kMaxR, *g = ws + 2 * kMaxR;
    s.c[0] = 0.0;
    s.c[max_index] = 0.0;
    for (int i = 0; i < sys; i++) {
        const double h_i = xa[i + 1] - xa[i];
        const double h_ip1 = xa[i + 2] - xa[i + 1];
        const double yd_i = ya[i + 1] - ya[i];
        const double yd_ip1 = ya[i + 2] - ya[i + 1];
        const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
        const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
        off[i] = h_ip1;
        diag[i] = 2.0 * (h_ip1 + h_i);
        g[i] = 3.0 * (yd_ip1 * g_ip1 - yd_i * g_i);
    }
    if (sys == 1) {
        s.c[1] = g[0] / diag[0];
        return 0;
    }
    double *gam = ws + 3 * kMaxR, *alp = ws + 4 * kMaxR, *z = ws + 5 * kMaxR;
    int status = 0;
    alp[0] = diag[0];
    gam[0] = off[0] / alp[0];
    if (alp[0] == 0) status = 1;
    for (int i = 1; i < sys - 1; i++) {
        alp[i] = diag[i] - off[i - 1] * gam[i - 1];
        gam[i] = off[i] / alp[i];
        if (alp[i] == 0) status = 1;
    }
    alp[sys - 1] = diag[sys - 1] - off[sys - 2] * gam[sys - 2];
    z[0] = g[0];
    for (int i = 1; i < sys; i++) z[i] = g[i] - gam[i - 1] * z[i - 1];
    double *x = s.c + 1;
    double *cc = ws + 6 * kMaxR;
    for (int i = 0; i < sys; i++) cc[i] = z[i] / alp[i];
    x[sys - 1] = cc[sys - 1];
    for (int i = sys - 2; i >= 0; i--) x[i] = cc[i] - gam[i] * x[i + 1];
    return status;
}

RL2_HD int spline_init(Spline &s, const double *xa, const double *ya, int size, int linear)
{
    double ws[7 * kMaxR];
    return spline_init_ws(s, xa, ya, size, linear, ws);
}

// exp(spline(x)); 0 outside the knots (marf, marginalise_r.c:227-238)
RL2_HD double spline_exp(const Spline &s, double x)
{
    if (x < s.x[0] || x > s.x[s.n - 1]) return 0.;
    int lo = 0, hi = s.n - 1;
    while (hi > lo + 1) {
        int i = (hi + lo) / 2;
        if (s.x[i] > x)
            hi = i;
        else
            lo = i;
    }
    const double x_lo = s.x[lo], x_hi = s.x[lo + 1];
    const double dx = x_hi - x_lo;
    if (!(dx > 0.0)) return 0.;
    const double y_lo = s.y[lo], y_hi = s.y[lo + 1];
    double y;
    if (s.linear) {
        y = y_lo + (x - x_lo) / dx * (y_hi - y_lo);
    } else {
        const double dy = y_hi - y_lo;
        const double delx = x - x_lo;
        const double c_i = s.c[lo], c_ip1 = s.c[lo + 1];
        const double b_i = (dy / dx) - dx * (c_ip1 + 2.0 * c_i) / 3.0;
        const double d_i = (c_ip1 - c_i) / (3.0 * dx);
        y = y_lo + delx * (b_i + delx * (c_i + delx * d_i));
    }
    return exp(y);
}

}  // namespace rl2
