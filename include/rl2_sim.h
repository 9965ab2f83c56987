/*
 * rl2_sim.h -- synthetic observations and catalogs (librl2sim.so, host C++).
 *
 * The reference reads a casacore Measurement Set written by its Simulate tool from a
 * catalog made by its Catalog tool (Simulate-from-catalog.cpp, generate_catalog.c).
 * casacore is absent here, so this library generates the same quantities in memory
 * (channel-major visibilities, per-row uvw in metres) with seeded generators:
 *   - rl2sim_mt_*:        Mersenne Twister with gsl_rng_mt19937 seeding and
 *                         gsl_rng_uniform / gsl_rng_uniform_pos semantics (utils.c:48-101
 *                         seeds it from /dev/random; here the seed is an argument).
 *   - rl2sim_layout_*:    antenna positions (east, north) in metres.
 *   - rl2sim_uvw:         earth-rotation synthesis of (u, v, w), rows time-major.
 *   - rl2sim_catalog:     generate_catalog.c:43-128 + generate_random_values.c:40-121.
 *   - rl2sim_simulate:    data_simulation.c:41-127 (shear, model sum, noise, SNR), followed
 *                         by the Measurement Set round trip of measurement_set.cpp:395-452
 *                         (complex64 in Jy, read back in micro-Jy).
 *   - rl2ms_*:            a raw on-disk container (RLMS) holding the fields the reference
 *                         reads from a Measurement Set (measurement_set.h:48-82).
 */
#ifndef RL2_SIM_H
#define RL2_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rl2sim_mt rl2sim_mt;
rl2sim_mt *rl2sim_mt_new(unsigned long seed);
void rl2sim_mt_free(rl2sim_mt *g);
uint32_t rl2sim_mt_get(rl2sim_mt *g);
double rl2sim_mt_uniform(rl2sim_mt *g);
double rl2sim_mt_uniform_pos(rl2sim_mt *g);
double rl2sim_mt_gaussian(rl2sim_mt *g, double *another); /* utils.c:77-101 */

/* Antenna layouts.  kind 0: log-uniform radius in [r_min, r_max], uniform azimuth.
 * kind 1: n*core_frac stations uniform in a disk of radius r_core, the rest log-uniform in
 * [r_core, r_max].  Draw order per station: radius, then azimuth. */
int rl2sim_layout(int kind, unsigned int n, unsigned long seed, double r_min, double r_max,
                  double core_frac, double r_core, double *east, double *north);

/* Earth-rotation synthesis: baselines (i<j) in order, rows = t * nbl + b, hour angle
 * H_t = h0_hours + t*dt_s (sidereal rate), site latitude lat_deg, pointing dec_deg. */
unsigned long rl2sim_uvw(unsigned int n, const double *east, const double *north, double lat_deg,
                         double dec_deg, unsigned int ntimes, double dt_s, double h0_hours,
                         double *u, double *v, double *w);

/* Catalog of ngal (multiple of 10) exponential galaxies ordered by decreasing flux;
 * values are rounded to 6 decimals (%f), as in the reference's catalog text files.
 * fov_arcmin: field diameter; fmin/fmax in micro-Jy. Returns the number generated. */
unsigned int rl2sim_catalog(unsigned int ngal, unsigned long seed, double fov_arcmin, double fmin,
                            double fmax, double *flux, double *scale, double *e1, double *e2,
                            double *l, double *m);

/* Shear e' = (e+g)/(1+g* e) (data_simulation.c:70-77), in place. */
void rl2sim_shear(unsigned int ngal, double g1, double g2, double *e1, double *e2);

/* data_simulation.c:41-127 on a channel-major complexd buffer (re,im interleaved, micro-Jy):
 * vis += sum_g model(e_g) [+ noise of variance sigma2 with the MT19937 stream `seed`].
 * snr (may be NULL) receives sqrt(sum |V_g|^2 / sigma2).  If ms_roundtrip != 0 the result is
 * stored as complex64 Jy and read back (x1e-6 -> float -> x1e6).  e1/e2 are used as given
 * (apply rl2sim_shear first). OpenMP over channels. */
int rl2sim_simulate(unsigned int nch, const double *wavenumbers, const double *spec,
                    unsigned long nrow, const double *u, const double *v, const double *w,
                    unsigned int ngal, const double *flux, const double *scale, const double *e1,
                    const double *e2, const double *l, const double *m, float sigma2,
                    int add_noise, unsigned long seed, int ms_roundtrip, double *vis,
                    double *snr);

/* RLMS container: write / read (fields as measurement_set.h:48-82; DATA complex64 Jy,
 * UVW float64 metres row-major, FLAG u8, SIGMA float32 per row). */
int rl2ms_write(const char *path, unsigned int nstations, unsigned int nch, unsigned long nrow,
                double freq_start_hz, double freq_inc_hz, double time_inc_s, const double *u,
                const double *v, const double *w, const unsigned char *flag_rc,
                const float *data_rc, const float *sigma_r);
int rl2ms_info(const char *path, unsigned int *nstations, unsigned int *nch, unsigned long *nrow,
               double *freq_start_hz, double *freq_inc_hz, double *time_inc_s);
/* Read into the reference's in-memory layout: uvw per row, flag/vis channel-major,
 * vis in micro-Jy (double), sigma2 = SIGMA^2 for the first nrow entries (the rest 0, see
 * measurement_set.cpp:249-274).  Returns ceil(max(|u|,|v|)) like ms_read_coords. */
double rl2ms_read(const char *path, double *u, double *v, double *w, unsigned char *flag,
                  double *vis, float *sigma2, int *status);

#ifdef __cplusplus
}
#endif
#endif
