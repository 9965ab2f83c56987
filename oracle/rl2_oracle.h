/*
 * rl2_oracle.h -- FP64 CPU oracle for the per-galaxy shape-measurement path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the reference that the HIP product path
 * (radiolensfit2_amd/, librl2hip.so) is compared against.  Only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may load it.  Nothing in the product
 * links, calls or falls back to it.
 *
 * It restates, in strict IEEE double precision (built without -ffast-math and with
 * -ffp-contract=off), the FACET build of RadioLensfit2:
 *   galaxy_visibilities.c, evaluate_uv_grid.c, source_extraction.c, likelihood.c,
 *   marginalise_r.c, galaxy_fitting.c, distributions.c, data_processing.c,
 *   data_processing-mpi.c (batch semantics), data_simulation.c (sky_model).
 * GSL 2.5 is absent, so its routines used on the path are restated here:
 *   nmsimplex2 (multimin), expint_Ei, cspline + linear interpolation, QAG with GK41.
 * Each function cites the reference file:line it follows.
 *
 * All functions are extern "C" with an orl2_ prefix so they never collide with the
 * product's reference-named drop-in symbols.
 */
#ifndef RL2_ORACLE_H
#define RL2_ORACLE_H

#include <stddef.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* complexd / likelihood_params: same layout as datatype.h:44-74 (FACET fields). */
typedef struct { double real; double imag; } orl2_complexd;

typedef struct {
    int numr;
    double *ro;
    double *rprior;
    unsigned int nchannels;
    double band_factor;
    double acc_time;
    double *spec;
    double *wavenumbers;
    double *uu;
    double *vv;
    double l0;
    double m0;
    orl2_complexd *data;
    double *sigma2;
    double *mod;
    unsigned int facet;
    unsigned int ncoords;
} orl2_params;

/* ---- setup arithmetic (RadioLensfit2.cpp) ---- */
void orl2_channel_constants(unsigned int nch, double freq_start_hz, double bw_hz,
                            double *wavenumbers, double *spec);            /* :342-351 */
float orl2_default_sigma2(int time_acc, double bw_hz);                      /* :255 */
double orl2_len_lambda(const double *u, const double *v, unsigned long nrow,
                       double wavenumber_top);                              /* measurement_set.cpp:223-240, RadioLensfit2.cpp:408 */
int orl2_r_grid(double *ro /* >= 29 */);                                    /* :370-392, returns numR */
double orl2_scale_mean(double flux);                                        /* distributions.c:59-62 */
double orl2_rfunc(double mu, double sigma, double r);                       /* distributions.c:72-87 */

/* ---- galaxy_visibilities.c ---- */
double orl2_sersic_shape(double k1, double k2, double sf, double wf);
double orl2_gaussian_shape(double k1, double k2, double sf, double wf);
void orl2_model_at_zero(double e1, double e2, double scale, unsigned int n,
                        const double *gu, const double *gv, const double *sigma2, double *mod);
void orl2_data_galaxy_vis(double spectra, double wavenumber, double e1, double e2, double scale,
                          double flux, double l, double m, unsigned long n, const double *u,
                          const double *v, const double *w, orl2_complexd *vis);
void orl2_phase_shift(double wavenumber, double l, double m, unsigned long n, const double *u,
                      const double *v, const double *w, orl2_complexd *vis);

/* ---- evaluate_uv_grid.c ---- */
unsigned int orl2_facet_size(double theta_med, double len);
unsigned int orl2_uv_grid_size(double len, const double *wavenumbers, unsigned int nch,
                               unsigned long nrow, const double *u, const double *v,
                               unsigned int sizeg, const bool *flag);
unsigned int orl2_facet_coords(double *gu, double *gv, double len, unsigned int sizeg,
                               const double *sum_w);
void orl2_gridding(const double *wavenumbers, unsigned int nch, unsigned long nrow,
                   const double *u, const double *v, const orl2_complexd *vis, const float *sigma2,
                   double len, unsigned int sizeg, orl2_complexd *new_vis, double *new_sigma2,
                   const bool *flag, double *sum_w);
void orl2_average_facets(unsigned long size, orl2_complexd *facet_vis, double *facet_sigma2,
                         const double *sum_w);

/* ---- source_extraction.c (FACET) ---- */
void orl2_source_extraction(unsigned int facet, unsigned int nch, const double *wavenumbers,
                            const double *spec, orl2_complexd *facet_vis, double *facet_sigma2,
                            double *sum_w, double l0, double m0, double flux, double mu,
                            double e1, double e2, const orl2_complexd *visSkyMod,
                            const orl2_complexd *visData, orl2_complexd *visGal,
                            const float *sigma2_vis, const bool *flag, unsigned long nrow,
                            const double *u, const double *v, const double *w, double len);

/* ---- likelihood.c ---- */
double orl2_expint_ei(double x);
double orl2_marg_position(double x);                                        /* :376-421 */
int orl2_cross_correlation(unsigned long n, const double *gu, const double *gv,
                           const orl2_complexd *data, const double *mod, double *ho,
                           double *det_sigma, int *n_iter);                  /* :195-358 */
double orl2_loglikelihood_r(double e1, double e2, double scale, unsigned long n,
                            const double *variance, const double *gu, const double *gv,
                            const orl2_complexd *data, double *mod, int *n_iter); /* :157-187 */
double orl2_loglikelihood(orl2_params *par, double e1, double e2, int *error); /* :106-149 */
double orl2_loglikelihood_detail(orl2_params *par, double e1, double e2, int *error,
                                 double *L_r /* numr */, int *n_iter /* numr */);
void orl2_likelihood_sampling(double *mes_e1, double *mes_e2, double maxL, orl2_params *par,
                              int np_max, double *var_e1, double *var_e2, double *oneDimvar);

/* ---- marginalise_r.c ---- */
void orl2_set_posterior_values(int numR, const double *L_r, const double *rprior,
                               const double *Ro, double *x, double *y, int *numvals);
double orl2_marginalise_posterior_r(int num, double *x, double *y);
/* GSL-restated pieces, exposed for pinning */
int orl2_cspline_eval_many(int num, const double *x, const double *y, int nq, const double *xq,
                           double *yq, int linear);
void orl2_qk41(double (*f)(double, void *), void *params, double a, double b, double *result,
               double *abserr, double *resabs, double *resasc);
int orl2_qag41(double (*f)(double, void *), void *params, double a, double b, double epsabs,
               double epsrel, unsigned long limit, double *result, double *abserr,
               unsigned long *n_intervals);

/* ---- galaxy_fitting.c ---- */
int orl2_source_fitting(orl2_params *par, double *mes_e1, double *mes_e2, double *var_e1,
                        double *var_e2, double *oneDimvar, double *maxL);

/* Generic drivers with a caller-supplied log-likelihood (pins the control flow). */
typedef double (*orl2_loglik_cb)(double e1, double e2, int *error, void *ctx);
typedef struct {
    int nm_iter;          /* NM iterations executed */
    int n_evals;          /* loglikelihood calls (fitting + sampling) */
    int nm_evals;
    double nm_e1, nm_e2;  /* simplex end point */
    double nm_size;       /* last simplex size */
} orl2_fit_trace;
int orl2_source_fitting_cb(orl2_loglik_cb cb, void *ctx, double start_e1, double start_e2,
                           double step, double *mes_e1, double *mes_e2, double *var_e1,
                           double *var_e2, double *oneDimvar, double *maxL,
                           orl2_fit_trace *trace);
void orl2_likelihood_sampling_cb(orl2_loglik_cb cb, void *ctx, double *mes_e1, double *mes_e2,
                                 double maxL, int np_max, double *var_e1, double *var_e2,
                                 double *oneDimvar, int *n_evals);
/* Plain GSL-style nmsimplex2 minimiser of f(x) in 2-D: returns iterations; x_out, fval. */
typedef double (*orl2_f2_cb)(double x0, double x1, void *ctx);
int orl2_nmsimplex2_minimize(orl2_f2_cb f, void *ctx, const double *x0, const double *step,
                             double tol, int max_iter, double *x_out, double *fval,
                             double *size_out, double *trace_best /* 2*max_iter or NULL */);

/* ---- data_simulation.c ---- */
void orl2_sky_model(const double *wavenumbers, const double *spec, unsigned int nch,
                    unsigned int ngal, const double *gflux, const double *gscale,
                    const double *l, const double *m, unsigned long nrow, const double *u,
                    const double *v, const double *w, orl2_complexd *visGal,
                    orl2_complexd *visMod);

/* ---- data_processing.c (serial, batch == 1) / data_processing-mpi.c (batch > 1) ---- */
typedef struct {
    double e1, e2, var_e1, var_e2, oneDimvar, maxL;
    int error;        /* source_fitting status of the last fit of this galaxy */
    int pass;         /* 0 = first pass, 1 = re-fit pass */
    unsigned int facet, ncoords;
    int n_evals;      /* loglikelihood calls for this galaxy (last pass) */
    int subtracted;
} orl2_gal_result;

typedef struct {
    double extraction_s, fitting_s, subtraction_s;
    int n_bad_first, n_bad_final;
} orl2_run_stats;

int orl2_data_processing(unsigned int ngal, int batch, double len, unsigned int nch,
                         const double *wavenumbers, const double *spec, unsigned long nrow,
                         const double *u, const double *v, const double *w,
                         const float *sigma2_vis, const bool *flag,
                         orl2_complexd *visData, orl2_complexd *visSkyMod,
                         const double *l, const double *m, const double *gflux,
                         const double *gscale, orl2_gal_result *results, int do_refit,
                         orl2_run_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
