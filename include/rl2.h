/*
 * rl2.h -- device-resident C-ABI of librl2hip.so (MI355X / gfx950).
 *
 * RadioLensfit2 keeps the whole observation in host memory and hands raw pointers to
 * every call on the per-galaxy path (data_processing.c:49-142).  On the GPU the
 * observation must stay resident in HBM across calls, so this layer adds an observation
 * handle (rl2_obs) that owns the device copy.  Every entry point takes plain pointers and
 * sizes, returns an int status (RL2_OK = 0, negative on error, message via
 * rl2_last_error()), and never falls back to the CPU: without a usable GPU the calls
 * return RL2_ERR_NO_DEVICE.
 *
 * Device representation of the reference's two visibility arrays: the path only ever uses
 * their difference visData - visSkyMod (source_extraction.c:56-61) and updates both by the
 * same galaxy (data_processing.c:109-135), so the device holds the residual
 * R = visData - visSkyMod.  rl2_obs_get_residual() returns R; rl2_obs_sync() writes
 * visData = R and visSkyMod = 0, which leaves every quantity the path computes unchanged.
 */
#ifndef RL2_H
#define RL2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RL2_OK = 0,
    RL2_ERR_NO_DEVICE = -1,
    RL2_ERR_ARG = -2,
    RL2_ERR_HIP = -3,
    RL2_ERR_RANGE = -4,
    RL2_ERR_STATE = -5,
};

/* Arithmetic of an observation's hot path.
 * RL2_PREC_FP64: residual complex128, FP64 model, likelihood and accumulators (the
 *   reference's arithmetic).
 * RL2_PREC_MIXED: residual complex64; FP64 phase-argument reduction and FP64 gridding
 *   accumulators; FP32 model and cross-correlation terms with FP64 chi-square sums;
 *   FP64 position marginalisation, log-sum-exp and scalelength integration. */
enum { RL2_PREC_FP64 = 0, RL2_PREC_MIXED = 1 };

typedef struct rl2_obs rl2_obs;

/* Per-galaxy result; the fields mirror the reference's output line
 * (data_processing.c:102,136) plus bookkeeping. */
typedef struct {
    double e1, e2, var_e1, var_e2, oneDimvar, maxL;
    int error;          /* source_fitting status of the last fit (galaxy_fitting.c:89-110) */
    int pass;           /* 0 first pass, 1 bad-list re-fit pass (RadioLensfit2.cpp:597-606) */
    unsigned int facet, ncoords;
    int n_evals;        /* loglikelihood evaluations requested by the fit */
    int subtracted;     /* 1 when the fitted model was removed from the residual */
    int start;          /* simplex start: 0 the reference's (0, 0), 1 the learned proposal */
    int rounds;         /* batched likelihood launches the fit waited for */
} rl2_gal_result;

typedef struct {
    double extraction_s, fitting_s, subtraction_s;
    int n_bad_first, n_bad_final;
} rl2_run_stats;

typedef struct rl2_proposal rl2_proposal;

/* Fit-driver options (0 = default for every field). */
typedef struct {
    int batch;             /* galaxies extracted/fitted per residual update (>=1); batch > 1
                              follows data_processing-mpi.c:65-216 */
    int do_refit;          /* 1: run the bad-list re-fit pass, as the reference always does */
    int nm_speculate;      /* 1: evaluate reflect/expand/contract candidates in one launch */
    int line_chunk;        /* line-search steps evaluated per launch (default 4) */
    int lik_block;         /* 0 (default): tiled likelihood, every (e, r) task split over
                              work-groups of 2048 facet cells; 64..1024: one work-group
                              of that many threads per (e, r) task */
    int verbose;           /* > 0: a progress line on stderr every `verbose` batches */
    int lik_precision;     /* 0: FP64 likelihood throughout; 1: mixed-precision likelihood
                              throughout; 2: mixed for the simplex, FP64 for the sampling;
                              3: FP64 simplex, mixed sampling; 4: as 3, with samples whose
                              threshold test lies within the FP32 error bound re-evaluated
                              in FP64 */
    rl2_proposal *proposal; /* learned proposal for the simplex start and first step, with
                              the reference's start (0, 0), step 0.1 as the fallback when the
                              proposal's likelihood is lower; NULL: the reference's start */
    float *features_out;   /* optional, rl2_process: ngal * 770 network inputs of each
                              galaxy's first-pass facet (training data) */
    double *labels_out;    /* optional, rl2_process: ngal * 3 = (e1, e2, log r) of each
                              first-pass fit, r maximising L_r + log prior at that e;
                              NaN for failed fits (training labels) */
    int proposal_mode;     /* 0: the proposal sets the simplex start, first step and the
                              sampling line-search depth; 1: only the line-search depth
                              (speculation: results identical to the reference start) */
} rl2_fit_options;

const char *rl2_last_error(void);
/* Select/initialise the device (hipSetDevice); idempotent. */
int rl2_init(int device);
int rl2_device_count(int *count);

/* ---- observation (replaces the arrays RadioLensfit2.cpp:157-535 allocates) ---- */
/* u,v,w: nrow metres; wavenumbers/spec: nch (RadioLensfit2.cpp:342-351);
 * sigma2_vis: nch*nrow variances (channel-major) after defaulting
 * (RadioLensfit2.cpp:255-270); flag: nch*nrow bytes or NULL; len: maximum uv distance in
 * wavelengths (RadioLensfit2.cpp:408). */
int rl2_obs_create(rl2_obs **out, unsigned int nch, unsigned long nrow, const double *u,
                   const double *v, const double *w, const double *wavenumbers,
                   const double *spec, const float *sigma2_vis, const unsigned char *flag,
                   double len, int precision);
int rl2_obs_destroy(rl2_obs *obs);
/* R = visData - visSkyMod (visSkyMod may be NULL); complexd arrays of nch*nrow. */
int rl2_obs_set_vis(rl2_obs *obs, const double *visData, const double *visSkyMod);
int rl2_obs_get_residual(rl2_obs *obs, double *residual /* 2*nch*nrow */);
int rl2_obs_sync(rl2_obs *obs, double *visData, double *visSkyMod);
/* R += sign * sum_g model_g (data_galaxy_visibilities, galaxy_visibilities.c:154-190).
 * sky_model (data_simulation.c:131-165) is sign = -1 with e = 0: visSkyMod += sum of round
 * models.  e1/e2 may be NULL (round models). */
int rl2_obs_add_sources(rl2_obs *obs, unsigned int ngal, const double *flux,
                        const double *scale, const double *l, const double *m,
                        const double *e1, const double *e2, double sign);
/* R += complex Gaussian noise of variance sigma2 per component (counter-based stream). */
int rl2_obs_add_noise(rl2_obs *obs, double sigma2, uint64_t seed);
/* Keep a device copy of the residual / put it back (repeatable runs over the same state). */
int rl2_obs_checkpoint(rl2_obs *obs);
int rl2_obs_restore(rl2_obs *obs);

/* ---- one galaxy at a time (the serial path of data_processing.c) ---- */
/* source_extraction + evaluate_facet_coords for one galaxy into facet slot `slot`.
 * Host outputs are optional (NULL): facet_vis (2*n), facet_sigma2 (n), uu/vv (n),
 * sum_w (facet^2 dense weights). *ncoords receives n. */
int rl2_obs_extract(rl2_obs *obs, int slot, double l0, double m0, double flux, double R_mu,
                    unsigned int facet, double *facet_vis, double *facet_sigma2, double *uu,
                    double *vv, double *sum_w, unsigned int *ncoords);
/* Upload a host facet into slot `slot` (used by the reference-signature drop-ins). */
int rl2_obs_load_facet(rl2_obs *obs, int slot, unsigned int n, const double *facet_vis,
                       const double *facet_sigma2, const double *uu, const double *vv);
/* log-prior of the scalelength grid for slot (rfunc, distributions.c:72-87) */
int rl2_obs_set_prior(rl2_obs *obs, int slot, const double *rprior /* numr */);
/* loglikelihood (likelihood.c:106-149) at npts ellipticities for the facet in `slot`.
 * L_r (npts*numr, may be NULL) receives the per-scalelength values of likelihood.c:157-187,
 * n_iter (npts*numr, may be NULL) the Newton iterations of cross_correlation. */
int rl2_obs_loglik(rl2_obs *obs, int slot, int npts, const double *e1, const double *e2,
                   double *L_e, int *error, double *L_r, int *n_iter);
/* Precision of rl2_obs_loglik evaluations: 0 FP64 (default), 1 mixed. */
int rl2_obs_loglik_precision(rl2_obs *obs, int prec);
/* source_fitting (galaxy_fitting.c:41-117) for the facet in `slot`. */
int rl2_obs_fit(rl2_obs *obs, int slot, const rl2_fit_options *opt, rl2_gal_result *res);
/* Remove a fitted galaxy: visSkyMod -= round model, visData -= fitted model
 * (data_processing.c:109-135). */
int rl2_obs_subtract(rl2_obs *obs, double l0, double m0, double flux, double R_mu, double e1,
                     double e2);

/* ---- the whole per-galaxy loop (data_processing.c:49-142, RadioLensfit2.cpp:597-606) ---- */
int rl2_process(rl2_obs *obs, unsigned int ngal, const double *l, const double *m,
                const double *flux, const double *scale, const rl2_fit_options *opt,
                rl2_gal_result *results, rl2_run_stats *stats);

/* ---- allocation-free entry point ----
 * No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  marginalise_posterior_r
 * (marginalise_r.c:81-223, as the function of the same name below) on device arrays:
 * xmarvals, ymarvals: num_marvals doubles each; out: 1 + num_marvals doubles receiving the
 * log of the integral, then ymarvals shifted by their maximum (the in-place update of the
 * reference).  Its QUADPACK workspace is a fixed 40 KB array per device held until the
 * process ends, so num_marvals is limited to 3..32 (RL2_ERR_RANGE beyond).  Queued on
 * `stream` (NULL: the library's stream) without waiting; calls that overlap on one device
 * share that workspace and race.  Calls no allocator; rl2_init must have run (it looks the
 * workspace up), else RL2_ERR_STATE. */
int rl2_marginalise_posterior_r_noalloc(int num_marvals, const double *xmarvals,
                                        const double *ymarvals, double *out, void *stream);

/* ---- verification entry point ----
 * No counterpart in the reference: the step of likelihood.c:126-146 that turns a request's
 * L_r into L_e (set_posterior_values, then marginalise_posterior_r when more than 4 points
 * remain), run by the kernel the likelihood engine launches (rmarg_kernel), for nitems rows
 * of L_r (kMaxR = 32 doubles each, L_r[0] unused) under one prior and scalelength grid
 * (numr in 2..32).  Le, err: nitems each (err 1 as loglikelihood's error flag).  For tests
 * and for callers that compute L_r themselves; synchronous. */
int rl2_rmarg_eval(int nitems, int numr, const double *ro, const double *rprior,
                   const double *Lr, double *Le, int *err);
/* No counterpart in the reference: marginalise_over_position_shift (likelihood.c:376-421,
 * declared in likelihood.h) evaluated by the device function the likelihood kernels apply,
 * for n host values x (out: n values, -1e10 for x <= Delta as the reference).  For tests;
 * synchronous. */
int rl2_marg_position_eval(int n, const double *x, double *out);

/* ---- scalelength grid and the fixed constants (RadioLensfit2.cpp:370-392) ---- */
int rl2_r_grid(double *ro /* >= 29 */);

/* ---- instrumentation: per-kernel device time measured with HIP events on the stream
 * the kernels run on, plus algorithmic work counters. ---- */
typedef struct {
    double ms;            /* summed kernel time */
    long long launches;
    double flops;         /* algorithmic FLOP-equivalents (likelihood) */
    double bytes;         /* algorithmic HBM bytes (streaming kernels) */
} rl2_kernel_stat;
/* RL2_K_LIKELIHOOD: the likelihood's chunk-sum kernels (pass 0 and the Newton passes, which
 * take the previous pass's step when steps are fused); RL2_K_STEP: the per-task step kernels
 * (after every pass with RL2_TUNE_FUSE_STEP 0, else once after an evaluation's last pass). */
enum { RL2_K_EXTRACT = 0, RL2_K_COMPACT = 1, RL2_K_LIKELIHOOD = 2, RL2_K_RMARG = 3,
       RL2_K_SUBTRACT = 4, RL2_K_SOURCES = 5, RL2_K_STEP = 6, RL2_K_COUNT = 7 };
int rl2_stats_enable(int on);
int rl2_stats_reset(void);
int rl2_stats_get(int kernel, rl2_kernel_stat *out);

/* Block until all work queued by the library has finished. */
int rl2_synchronize(void);

/* ---- launch-shape knobs (set by an autotuner; process-wide) ----
 * RL2_TUNE_GRID_VARIANT: extraction gridding, runs of consecutive rows merged per thread
 *   before one LDS table update: 0 / 1 / 2 / 3 / 4 runs of 2 / 2 / 8 / 4 / 4 rows with
 *   tables of 512 / 1024 / 2048 / 1024 / 2048 slots, flushed to the facet after every
 *   channel; 5 / 6 runs of 2 / 4 rows with 1024 / 2048 slots that keep their cells across
 *   channels and are flushed only when the next channel might not fit; default 5 (the
 *   fastest with the subtraction fused into the pass).  Every
 *   variant accumulates in integers, so the facet does not depend on the order of the
 *   work-groups (variants with the same rows per work-group give the same bits).
 * RL2_TUNE_TILE_CELLS: facet cells per work-group of the tiled likelihood, a power of two
 *   in [256, 65536], or 0 (default) to choose per launch (2048 to 65536).
 * RL2_TUNE_PHASE_TABLES: 2 (default) evaluates the FP64 Newton phases on extracted facets
 *   as products of per-pass column and row tables, both staged in LDS when they fit (40
 *   bytes per grid index within 40 KB); 1 stages only the column table; 0 evaluates
 *   sin/cos per cell.  All three give the same values up to rounding (1 and 2 identical).
 * RL2_TUNE_FUSE_SUBTRACT: 1 (default) makes the last subtraction of a batch in the same pass
 *   over the residual as the next extraction (gridding variant 3); 0 keeps them apart.
 * RL2_TUNE_TILE_BLOCKS: cap on the blocks of a likelihood pass, a multiple of 8 (each block
 *   then walks several chunks); 0 (default): one block per chunk.
 * RL2_TUNE_TILE_STREAMS: 2 splits the scalelength tasks of an evaluation in two groups
 *   whose Newton passes run on two streams; 1 (default) runs them all on one.
 * RL2_TUNE_BATCH_EXTRACT: 1 (default) extracts a batch of 2..64 galaxies
 *   (rl2_fit_options.batch) in one persistent pass over the residual, which also applies the
 *   previous batch's subtractions; 0 extracts and subtracts one galaxy per pass.  Both give
 *   identical results.
 * RL2_TUNE_BATCH_GROUPS: persistent work-groups of a batched extraction pass (0: three per
 *   compute unit).
 * RL2_TUNE_BATCH_GRAPH: 1 (default) replays a batch's extraction launches (the residual
 *   pass, compactions and FP32 copies) as a hipGraph captured once per batch shape; 0 launches
 *   them one by one.  Both give identical results.
 * RL2_TUNE_ITEM_GROUPS: work-groups the per-facet chunk choice aims at for one likelihood
 *   item (RL2_TUNE_TILE_CELLS 0): the chunk is the largest power of two from 2048 cells that
 *   still gives the item's 23 scalelength tasks that many work-groups; 0 (default): 256 for
 *   FP64 items, 512 for FP32 items.
 * RL2_TUNE_FUSE_STEP: 1 (default) folds each Newton step of the likelihood into the next
 *   pass's launch (its work-groups take the step and make their own phase tables), so a pass
 *   is one launch; 2 runs every Newton pass of an evaluation in one cooperative launch (the
 *   blocks wait on per-task counters between passes); 0 runs a step kernel after every pass.
 *   All three give identical results.
 * RL2_TUNE_TILE_PASSES: Newton passes queued before the first status check of an evaluation
 *   (1..50); 0 (default): the previous evaluation's largest step count + 1.  Passes beyond a
 *   task's last step leave it at once, and missing ones are queued after the check, so the
 *   results do not depend on it.
 * RL2_TUNE_P0_TASKS: scalelength tasks per work-group (1..3, default 3) in the zero-offset
 *   pass of FP64 items on extracted facets; the tasks of a work-group share each cell's
 *   loads and shape terms.  Identical results. */
enum { RL2_TUNE_GRID_VARIANT = 0, RL2_TUNE_TILE_CELLS = 1, RL2_TUNE_PHASE_TABLES = 2,
       RL2_TUNE_FUSE_SUBTRACT = 3, RL2_TUNE_TILE_BLOCKS = 4, RL2_TUNE_TILE_STREAMS = 5,
       RL2_TUNE_BATCH_EXTRACT = 6, RL2_TUNE_BATCH_GROUPS = 7, RL2_TUNE_BATCH_GRAPH = 8,
       RL2_TUNE_ITEM_GROUPS = 9, RL2_TUNE_FUSE_STEP = 10, RL2_TUNE_TILE_PASSES = 11,
       RL2_TUNE_P0_TASKS = 12 };
int rl2_tune_set(int knob, int value);
int rl2_tune_get(int knob, int *value);

/* Autotuner: times the candidate values of each launch knob in turn (coordinate search) on
 * the first pass of the loop over the given galaxies, from the observation's current
 * residual, and sets the fastest; the residual is unchanged on return.  Knobs: the batched
 * extraction's work-groups (batch > 1) or the gridding variant, the likelihood's facet cells
 * per work-group (RL2_TUNE_TILE_CELLS), the split of the scalelength tasks over two streams
 * and, as knob -1, the work-group threads per (e, r) task (0: the tiled engine), returned in
 * *lik_block for rl2_fit_options.lik_block.  Every knob leaves results unchanged up to
 * rounding.  trials (optional) receives up to max_trials timed trials. */
typedef struct {
    int knob, value;
    double seconds;
} rl2_tune_trial;
int rl2_autotune(rl2_obs *obs, unsigned int ngal, const double *l, const double *m,
                 const double *flux, const double *scale, const rl2_fit_options *opt,
                 rl2_tune_trial *trials, int max_trials, int *ntrials, int *lik_block);

/* ---- test hooks (host-side; no GPU needed) ----
 * The fit control flow driven by a caller-supplied log-likelihood; returns the number of
 * batched evaluation rounds. */
int rl2_fit_with_callback(double (*loglik)(double e1, double e2, int *error, void *ctx),
                          void *ctx, int nm_speculate, int line_chunk, rl2_gal_result *res);
/* Philox4x32-10 of n counters (4 words each) under one key (2 words), as the noise kernel
 * evaluates it. */
void rl2_test_philox4x32_10(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);
/* sin(2 pi t), cos(2 pi t) as the FP64 kernels evaluate them. */
void rl2_test_sincos_rev64(int n, const double *t, double *s, double *c);
/* Cell indices by the kernels' reciprocal-product form and by the reference's division;
 * returns the number of disagreements (0xffffffff marks a cell outside the facet). */
int rl2_test_cell_index(int n, const double *u, const double *v, double inv_lambda, double len,
                        double inc, unsigned int facet, unsigned int *fast, unsigned int *exact);

/* ---- learned simplex proposal ----
 * No counterpart in the reference: a 770-256-256-256-9 MLP (GELU, BF16 MFMA products with
 * FP32 accumulation, FP32 master weights, Adam) that reads a galaxy's binned facet and
 * predicts the likelihood peak (e1, e2, log r) and a 3x3 Cholesky factor of its precision.
 * It may only replace the simplex start point and first step (galaxy_fitting.c:59-67);
 * rl2_fit_options.proposal enables it in rl2_process / rl2_obs_fit. */
int rl2_proposal_num_params(void);
int rl2_proposal_num_inputs(void);   /* 770: 16x16 (re, im, weight) bins, log flux, log R */
int rl2_proposal_num_outputs(void);  /* 9: mu(3), log diag(3), off-diagonal (3) */
int rl2_proposal_create(rl2_proposal **out, uint64_t seed);
int rl2_proposal_destroy(rl2_proposal *p);
int rl2_proposal_get_params(rl2_proposal *p, float *host);
int rl2_proposal_set_params(rl2_proposal *p, const float *host);
int rl2_proposal_get_grads(rl2_proposal *p, float *host);
/* outputs[batch][9] for inputs[batch][770] */
int rl2_proposal_forward(rl2_proposal *p, int batch, const float *inputs, float *outputs);
/* mean Gaussian negative log-likelihood of targets[batch][3] = (e1, e2, log r) and its
 * gradient (left in the network, rl2_proposal_get_grads) */
int rl2_proposal_gradients(rl2_proposal *p, int batch, const float *inputs,
                           const float *targets, float *mean_loss);
/* one Adam step (lr) on the gradient of that loss */
int rl2_proposal_train_step(rl2_proposal *p, int batch, const float *inputs,
                            const float *targets, float lr, float *mean_loss);
/* network inputs of the facet extracted into `slot` (catalog scalelength R_mu, flux) */
int rl2_obs_features(rl2_obs *obs, int slot, double R_mu, double flux, float *out /* 770 */);

#ifdef __cplusplus
}
#endif
#endif
