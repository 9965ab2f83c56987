/*
 * rl2_reference.h -- RadioLensfit2's per-galaxy function surface (FACET build) exported by
 * librl2hip.so.  Link against librl2hip.so instead of compiling galaxy_visibilities.c,
 * evaluate_uv_grid.c, source_extraction.c, likelihood.c, marginalise_r.c, galaxy_fitting.c
 * and data_processing.c: the callers (RadioLensfit2.cpp:597-606, data_processing.c:49-142)
 * keep their code, every array computation runs on the GPU.
 *
 * Same names, argument order and meaning, same in-place outputs and sentinel values
 * (-1e10 log-likelihood, int error flags) as the reference.  Differences a caller can see:
 *  - No GSL: f_likelihood takes a struct with the memory layout of gsl_vector (or GSL's own
 *    type when <gsl/gsl_vector.h> is included first).
 *  - Galaxies whose scalelength exceeds RMAX cannot overflow facet buffers: facet sizes
 *    above facet_size(RMAX, len) are rejected with an error.
 *  - data_processing() returns visData as the residual visData - visSkyMod and zeroes
 *    visSkyMod (the path only uses their difference).
 *  - Scalar helpers (sersic_shape, gaussian_shape, facet_size, scale_mean, rfunc,
 *    marginalise_over_position_shift, set_posterior_values, marf) evaluate the same
 *    host/device arithmetic the kernels use, on the calling thread.
 *  - Unlike the reference, fatal conditions (rfunc(r <= 0), failed spline set-up) return
 *    NaN / -1e10 instead of exit(EXIT_FAILURE).
 */
#ifndef RL2_REFERENCE_H
#define RL2_REFERENCE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* datatype.h:44-47 */
typedef struct {
    double real;
    double imag;
} complexd;

/* datatype.h:49-74 (FACET fields :64-67) */
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
    complexd *data;
    double *sigma2;
    double *mod;
    unsigned int facet;
    unsigned int ncoords;
} likelihood_params;

#ifndef __GSL_VECTOR_DOUBLE_H__
/* memory layout of GSL 2.5 gsl_vector (gsl_vector_double.h) */
typedef struct {
    size_t size;
    size_t stride;
    double *data;
    void *block;
    int owner;
} gsl_vector;
#endif

/* ---- galaxy_visibilities.h:31-55 ---- */
double gaussian_shape(double k1, double k2, double scale_factor, double wave_factor);
double sersic_shape(double k1, double k2, double scale_factor, double wave_factor);
void model_galaxy_visibilities_at_zero(double e1, double e2, double scale, unsigned int num_coords,
                                       double *grid_u, double *grid_v, const double *sigma2,
                                       double *Modvis);
void data_galaxy_visibilities(double spectra, double wavenumber, double band_factor,
                              double acc_time, double e1, double e2, double scale, double flux,
                              double l, double m, unsigned int num_coords, double *uu_metres,
                              double *vv_metres, double *ww_metres, complexd *vis);
void data_visibilities_phase_shift(double wavenumber, double l, double m, unsigned int num_coords,
                                   double *uu_metres, double *vv_metres, double *ww_metres,
                                   complexd *vis);

/* ---- evaluate_uv_grid.h:29-33 ---- */
unsigned int facet_size(double theta_med, double len);
unsigned int evaluate_uv_grid_size(int rank, int nprocs, double len, double *wavenumbers,
                                   unsigned int num_channels, unsigned int ncoords, double *u,
                                   double *v, unsigned int sizeg, bool *flag);
unsigned int evaluate_facet_coords(double *grid_u, double *grid_v, double len, unsigned int sizeg,
                                   double *count_w);
void gridding_visibilities(double *wavenumbers, unsigned int num_channels, unsigned int ncoords,
                           double *u, double *v, complexd *vis, float *sigma2, double len,
                           unsigned int sizeg, complexd *new_vis, double *new_sigma2, bool *flag,
                           double *sum_w);
void average_facets(unsigned long int size, complexd *grid_vis, double *grid_sigma2,
                    double *sum_w);

/* ---- source_extraction.h:31-34 (FACET) ---- */
void source_extraction(int rank, unsigned int facet, likelihood_params *par, complexd *facet_vis,
                       double *facet_sigma2, double *sum_w, double l0, double m0, double flux,
                       double mu, double e1, double e2, complexd *visSkyMod, complexd *visData,
                       complexd *visGal, float *sigma2_vis, bool *flag,
                       unsigned long int num_coords, double *uu_metres, double *vv_metres,
                       double *ww_metres, double len);

/* ---- likelihood.h:32-54 (FACET) ---- */
double f_likelihood(const gsl_vector *v, void *params);
double loglikelihood(void *params, double ee1, double ee2, int *error);
double loglikelihood_r(double ee1, double ee2, double scale, unsigned long int n_uv_coords,
                       const double *variance, double *grid_u, double *grid_v,
                       complexd *visData, double *visM);
int cross_correlation(unsigned long int n_uv_coords, double *grid_u, double *grid_v,
                      complexd *visData, double *visMod, double *ho, double *det_sigma);
double marginalise_over_position_shift(double x);
void likelihood_sampling(double *mes_e1, double *mes_e2, double maxL, void *params, int np_max,
                         double *var_e1, double *var_e2, double *oneDimvar);

/* ---- marginalise_r.h:24-27 ---- */
void set_posterior_values(int numR, double *L_r, double *rprior, double *Ro, double *xmarvals,
                          double *ymarvals, int *numvals);
double marginalise_posterior_r(int num_marvals, double *xmarvals, double *ymarvals);
double marf(double x, void *params);

/* ---- galaxy_fitting.h:30 ---- */
int source_fitting(int rank, likelihood_params *par, double *mes_e1, double *mes_e2,
                   double *var_e1, double *var_e2, double *oneDimvar, double *maxL);

/* ---- distributions.h (hot-path priors: scale_mean :59-62, rfunc :72-87) ---- */
double scale_mean(double flux);
double rfunc(const double mu, const double sigma, double r);

/* ---- data_simulation.h: sky_model (data_simulation.c:131-165) ---- */
void sky_model(double *wavenumbers, double *spec, double channel_bandwidth_hz, int time_acc,
               unsigned int num_channels, unsigned int n_gal, double *gflux, double *gscale,
               double *l, double *m, unsigned long int num_coords, double *uu_metres,
               double *vv_metres, double *ww_metres, complexd *visGal, complexd *visMod);

/* ---- data_processing.h:32-41 (serial signature) ---- */
void data_processing(bool re_fitting, unsigned int *bad_list, int nsources, double len,
                     unsigned int num_coords, FILE *pFile, likelihood_params *par, double *l,
                     double *m, double *gflux, double *gscale, double *ge1, double *ge2,
                     double *SNR_vis, double *sum_w, complexd *visGal, complexd *visSkyMod,
                     complexd *visData, float *sigma2_vis, bool *flag, double *uu_metres,
                     double *vv_metres, double *ww_metres, double *fitting_time, int *bad);

#ifdef __cplusplus
}
#endif
#endif
