// rl2_internal.h -- device descriptors and kernel launchers shared by the runtime,
// the fit driver and the reference-signature drop-ins.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rl2_math.h"

namespace rl2 {

// ---------------------------------------------------------------- descriptors (device)
// One extracted facet (the reference's par->uu/vv/data/sigma2, data_processing.c:82-84):
// compacted non-empty cells in row-major cell order, data already divided by sigma^2.
struct FacetDev {
    const double *uu, *vv;     // cell centres in wavelengths
    const double *dre, *dim;   // gridded data / variance
    const double *s2;          // gridded variance
    const double *mod;         // optional given model (drop-in cross_correlation); else null
    const float4 *pk;          // FP32 copy {u, v, D_re, D_im} for the mixed-precision kernel
    const float *invs2;        // FP32 1 / sigma^2
    const double *invs2_64;    // FP64 1 / sigma^2
    // Grid structure of a facet written by extraction (null / 0 for a caller-given facet):
    // cell k sits at row cid[k] >> 16, column cid[k] & 0xffff, and its centre is
    // (ctr[column], ctr[row]), bit-identical to uu[k], vv[k].
    const uint32_t *cid;
    const double *ctr;
    // Packed copies for the fused Newton passes (grid facets up to kRecFacet wide, else
    // unset): the data {D_re, D_im} in one 16-byte load, and each cell's LDS record offsets
    // (row * kRecBytes) << 16 | column * kRecBytes; both with the zero pad cell.
    const double2 *dc;
    const uint32_t *loff;
    unsigned int grid_facet;
    unsigned int pad0;
    unsigned int n;
    unsigned int pad;
    double rprior[kMaxR];      // log prior per scalelength index
};

// One likelihood evaluation request.
struct LikItem {
    int slot;
    int prec;                  // 0: FP64 kernel, 1: mixed-precision kernel
    double e1, e2;
    unsigned chunk_first;      // tiled engine: chunks of the items before this one
    unsigned tab_first;        // tiled engine: first phase-table entry of this item
    unsigned chunk_cells;      // tiled engine: facet cells per work-group of this item (a
                               // function of its facet alone, so that its sums are added in
                               // the same order whatever else the launch carries)
    unsigned pad2;
};

struct RGrid {
    int numr;                  // 24: ro[1..23] are the models (RadioLensfit2.cpp:370-392)
    int pad;
    double ro[kMaxR];
};

// Streaming-kernel arguments (one galaxy).
struct GalaxyGeom {
    double l0, m0, n0;         // n0 = sqrt(1 - l^2 - m^2) - 1
    double flux;
    double sf;                 // (scale * ARCS2RAD)^2 / detA^2
    double e1, e2;
};

// ---------------------------------------------------------------- launchers
// All launches go to `stream`; errors are reported through hipGetLastError by the caller.

// R[c][i] += sign * sum_g model_g ; gal: ngal GalaxyGeom on the device.
void launch_add_sources(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                        const double *u, const double *v, const double *w, const double *wn,
                        const double *spec, const GalaxyGeom *gal, unsigned ngal, double sign,
                        void *res);
void launch_add_noise(hipStream_t s, int prec, unsigned long nvis, double sigma, uint64_t seed,
                      void *res);

// Power-of-two scales of the gridding's integer accumulators: a value x is carried as
// rint(x * fx_scale) (visibility components) or rint(x * fs_scale) (sigma^2), with
// |x| * scale * (terms a table slot takes between flushes, in any gridding path) <= 2^62.
struct AccScale {
    double fx_scale = 1., fs_scale = 1.;
};
// The scales for visibility components bounded by vis_bound and variances by s2_bound.
// rows_log2 < 0: the bound common to every gridding path for nch channels (the most terms
// a slot takes between flushes in any of them), so that all paths give the same facets.
AccScale fx_scales(double vis_bound, double s2_bound, int rows_log2 = -1, unsigned nch = 1);
// Words (8 bytes) of a facet's gridding accumulators.
size_t grid_acc_words(unsigned facet);
// *out = max(*out, max over the residual of |re| + |im|), as the bits of a double (*out
// must start at 0 or at an earlier bound's bits).
void launch_abs_bound(hipStream_t s, int prec, unsigned long nvis, const void *res,
                      unsigned long long *out);

// Pillbox gridding into a facet's dense integer accumulators (grid_acc_words(facet) words,
// zero before the first launch; compaction clears them again).
// mode 0: extraction (round model of `g` added, phase-shifted to (l0,m0)), reads the
//         residual `res` (prec selects complex128/complex64).
// mode 1: plain gridding of complex128 `res` (drop-in gridding_visibilities).
// mode 2: weights only (evaluate_uv_grid_size).
struct GridLaunch {
    int mode, prec;
    unsigned nch;
    unsigned long nrow;
    const double *u, *v, *w;
    const double *wn, *inv_lambda, *spec;
    const void *res;
    const float *s2f;          // sigma^2 with the sign bit set for flagged visibilities
    GalaxyGeom g;
    double len, inc;
    unsigned facet;
    double *acc;               // grid_acc_words(facet) words
    double fx_scale = 1., fs_scale = 1.;  // AccScale of the launch
    // mode 0, gridding variant 3 only: apply this post-fit update to the residual first
    // (written back through res_w), then grid the updated visibilities
    int sub = 0;
    GalaxyGeom sfit, srnd;
    void *res_w = nullptr;
};
void launch_grid(hipStream_t s, const GridLaunch &a);

// Batched extraction (data_processing-mpi.c:65-216 semantics: every galaxy of a batch is
// extracted from the same residual).  One persistent launch streams the residual once for
// the whole batch: work-groups take tiles of 1024 rows from a device work queue; for each
// channel of a tile they first apply the pending post-fit updates of the previous batch, in
// order (written back: the same arithmetic as that many subtraction passes), then grid the
// tile into every target galaxy's facet in turn.
constexpr int kMaxBatch = 64;
struct BatchTarget {
    GalaxyGeom g;               // round model and phase centre of the galaxy
    double inc, inv_inc;        // cell width of its facet and its reciprocal
    double fx_scale;            // AccScale of its accumulators
    double *acc;                // grid_acc_words(facet) words
    unsigned facet;
    unsigned pad;
};
struct BatchSub {
    GalaxyGeom fit, rnd;        // R <- (R - fit) + rnd (data_processing.c:109-135)
};
// Compaction of one target's accumulators into its facet slot.
struct BatchCompact {
    double *acc;
    double fx_unscale, fs_unscale;
    double *uu, *vv, *dre, *dim, *s2;
    uint32_t *cid;
    double *ctr;
    double2 *dc;
    uint32_t *loff;
    float4 *pk;
    float *invs2;
    double *invs2_64;
    unsigned *n_out;            // the slot descriptor's cell count
    unsigned *scan;             // block counts, compact_scratch_words(facet) words
    unsigned facet;
    unsigned pad;
};
// Everything that changes from one batch to the next, in device memory, so that the same
// launches (and a graph of them) serve every batch.
struct BatchDesc {
    unsigned ntarget, nsub;     // galaxies of the batch, pending updates applied first
    unsigned pad[2];
    BatchTarget targets[kMaxBatch];
    BatchSub subs[kMaxBatch];
    BatchCompact compact[kMaxBatch];
};
struct GridBatchLaunch {
    int prec;
    unsigned nch;
    unsigned long nrow;
    const double *u, *v, *w, *wn, *inv_lambda, *spec;
    void *res;
    const float *s2f;
    double len, fs_scale;
    const BatchDesc *desc;      // device
    unsigned *queue;            // device: next tile, finished work-groups; zero before the
                                // launch, and the launch leaves them zero again
    unsigned ntiles;
};
// The batch's compactions and FP32 copies: four launches over ncap targets of facets up to
// fmax (targets beyond desc->ntarget and blocks beyond a facet exit at once).
void launch_compact_batch(hipStream_t s, const BatchDesc *desc, unsigned ncap, unsigned fmax,
                          double len);
constexpr unsigned kBatchTileRows = 1024;
// groups: persistent work-groups to launch (at most the tile count).
void launch_grid_batch(hipStream_t s, const GridBatchLaunch &a, unsigned groups);
// R <- (R - fit_b) + rnd_b for b = 0 .. nsub-1, in order, per visibility.
void launch_subtract_batch(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                           const double *u, const double *v, const double *w, const double *wn,
                           const double *inv_lambda, const double *spec, const BatchSub *subs,
                           unsigned nsub, void *res);

// Launch-shape knobs (rl2_tune_set): read by the launchers on the host.
struct Tuning {
    int grid_variant = 5;      // RL2_TUNE_GRID_VARIANT (runs of rows per thread, LDS table
                               // slots, flush per channel or deferred; see rl2.h)
    unsigned tile_cells = 0;   // RL2_TUNE_TILE_CELLS (0: chosen per launch)
    int phase_tables = 2;      // RL2_TUNE_PHASE_TABLES: separable Newton phases on grid facets
    int fuse_subtract = 1;     // RL2_TUNE_FUSE_SUBTRACT: a subtraction rides on the next extraction
    unsigned tile_blocks = 0;  // RL2_TUNE_TILE_BLOCKS: cap on the blocks of a likelihood pass
                               // (multiple of 8; 0: one block per chunk)
    int tile_streams = 1;      // RL2_TUNE_TILE_STREAMS: task groups (streams) of the passes
    int batch_extract = 1;     // RL2_TUNE_BATCH_EXTRACT: one residual pass per batch (1) or
                               // one per galaxy (0)
    unsigned batch_groups = 0; // RL2_TUNE_BATCH_GROUPS: persistent work-groups of a batched
                               // extraction (0: three per CU)
    unsigned item_groups = 0;  // RL2_TUNE_ITEM_GROUPS: work-groups an item's chunking aims at
                               // (0: kTileItemGroups)
    int batch_graph = 1;       // RL2_TUNE_BATCH_GRAPH: replay a batch's extraction launches
                               // as a hipGraph (1) or launch them one by one (0)
    int fuse_step = 1;         // RL2_TUNE_FUSE_STEP: Newton steps folded into the next pass
                               // (1) or a step kernel after every pass (0)
    int tile_passes = 0;       // RL2_TUNE_TILE_PASSES: Newton passes queued before the first
                               // status check (0: the previous evaluation's largest count + 1)
    int p0_tasks = 3;          // RL2_TUNE_P0_TASKS: scalelength tasks per work-group in pass
                               // 0 of the grid kernels (1..3)
};
Tuning &tuning();

// Compaction of dense accumulators into the facet arrays; clears the accumulators.
// fx: the integer accumulators of launch_grid with these scales; null: four double arrays
// acc[4][facet^2] (re, im, sigma^2, weight) given by a caller.
// n_out (device) receives the number of non-empty cells. sum_w_dense (optional, device)
// receives the dense weights before clearing.
// cid / ctr (optional): packed row and column of each compacted cell, and the centre
// coordinate of every grid index (facet entries).
void launch_compact(hipStream_t s, double *acc, unsigned facet, double len, double *uu,
                    double *vv, double *dre, double *dim, double *s2, uint32_t *cid, double *ctr,
                    unsigned *n_out,
                    double *sum_w_dense, unsigned *scratch_blocks, const AccScale *fx);
size_t compact_scratch_words(unsigned facet);
// FP32 packing of a compacted facet (count read from the device descriptor).
void launch_pack_facet(hipStream_t s, const FacetDev *dfacet, unsigned cap, const double *uu,
                       const double *vv, const double *dre, const double *dim, const double *s2,
                       float4 *pk, float *invs2, double *invs2_64, const uint32_t *cid,
                       unsigned facet, double2 *dc, uint32_t *loff);

struct TileTask {
    double xo, yo;             // position offset of the current Newton iterate
    double norm;               // |increment| of the last step (1 before the first)
    double S2;                 // sum M^2 / sigma^2 of the unnormalised model
    int iter;                  // Newton steps taken
    int active;                // 1 while another Newton pass is due
    int fragile;               // the finished L_r sits at a branch that FP32 sums can flip
    int pad;
};

// Likelihood: per (item, r) work-group L_r and Newton iterations; then r-marginalisation.
struct LikLaunch {
    int nitems;
    const LikItem *items;
    const FacetDev *facets;
    RGrid rg;
    double *Lr;                // nitems * kMaxR
    int *nit;                  // nitems * kMaxR
    double *Le;                // nitems
    int *err;                  // nitems
    double *qag_ws;            // nitems * kQagDoubles
    int *qag_iws;              // nitems * kQagInts
    int block;                 // threads per work-group (64..1024)
    int prec;
    double *ho = nullptr;      // optional nitems * kMaxR: cross-correlation peak value
    double *det = nullptr;     // optional nitems * kMaxR: Hessian determinant at the peak
    const int *host_prec = nullptr;  // host copy of items[].prec (selects the kernels)
    const TileTask *tasks = nullptr; // tiled engine: per-(item, r) state (fragile flags)
    // tiled engine: the r-marginalisation launch also writes the evaluation's status
    // (tasks still iterating, largest step count) and FLOP-equivalents when set
    unsigned *status = nullptr;
    double *flops = nullptr;
};
// err bit set for a mixed-precision item one of whose L_r is fragile
constexpr int kErrFragile = 0x100;
constexpr int kQagLimit = 1000;
constexpr int kQagDoubles = 4 * kQagLimit;
constexpr int kQagInts = 2 * kQagLimit;
void launch_likelihood(hipStream_t s, const LikLaunch &a);
void launch_rmarg(hipStream_t s, const LikLaunch &a);

// Tiled likelihood.  Every (item, scalelength) task is split into chunks of chunk_cells
// facet cells with one work-group per chunk.  A pass (the zero-offset sums, or one Newton
// iteration) is two launches: every work-group writes its chunk's sums, then one wave per
// task adds them in chunk order and takes the Newton step or finishes L_r.  The
// r-marginalisation follows as its own launch.
constexpr int kTileBS = 256;
constexpr unsigned kTileMinCells = 2048, kTileMaxCells = 65536;
constexpr unsigned kTileKnobMinCells = 256;   // smallest chunk RL2_TUNE_TILE_CELLS accepts
constexpr unsigned long long kTileItemGroups = 256;    // per FP64 item (fused Newton passes)
constexpr unsigned long long kTileItemGroups32 = 512;  // per FP32 item (two-kernel passes)
constexpr int kInlItems = 8;
// LDS record of one grid index in the fused Newton passes: {X (16 bytes), Y (16), ctr (8)}.
constexpr unsigned kRecBytes = 40;
constexpr unsigned kRecFacet = 1024;  // widest facet with records (40 KB of LDS)
constexpr unsigned kTileLdsBytes = 40960;  // column-table staging per work-group (<= 1706 columns)

struct TileLaunch {
    LikLaunch lk;              // items (with their first work-group), facets, r grid,
                               // outputs and QAG workspace
    TileTask *tasks;           // nitems * kMaxR
    double *partials;          // 8 doubles per work-group
    unsigned nwg;              // work-groups of a pass (nr_g * chunks summed over items)

    // The launch covers the scalelength tasks r_lo .. r_lo + nr_g - 1 of every item: the
    // tasks can be split in two groups whose passes run on two streams, so that one group's
    // step kernel and launch gaps overlap the other group's sums.
    int r_lo, nr_g;
    // Phase tables of the FP64 Newton passes on grid facets: per (item, r) task,
    // e^{2 pi i xo ctr[k]} for the columns then e^{2 pi i yo ctr[k]} for the rows
    // (2 * grid_facet entries from items[i].tab_first + (r - 1) * 2 * grid_facet); null: the
    // passes evaluate sin/cos per cell.
    double2 *tab;
    unsigned tab_lds_facet;    // widest grid_facet of the launch's table items (LDS staging)
    unsigned tab_y_lds;        // 1: the row tables are staged in LDS too
    unsigned grid;             // 1: every item is FP64 on an extracted facet with tables;
                               // 2: the same with the items in the kernel arguments
    // Up to kInlItems items travel in the kernel arguments instead of a host-to-device
    // copy: the tile kernels read them from here, and block 0 of the pass-0 tile kernel
    // stores them to items_w (= lk.items) for the step kernels and the r-marginalisation.
    unsigned n_inl;
    LikItem *items_w;
    LikItem inl[kInlItems];
    // Two generations of task state and chunk sums (fused passes, and the step kernel after
    // them): the state is read from `tasks` and written to `tasks_out` (null: `tasks`), the
    // previous pass's sums read from `part_in` while this pass's go to `partials`.
    TileTask *tasks_out = nullptr;
    const double *part_in = nullptr;
    int mode = 0;              // fused passes: 0 first step (from the pass-0 sums), 1 Newton
                               // step, 2 resume from the state in `tasks`
    unsigned item_chunks = 0;  // chunks of every item when all items have the same count
                               // (0: they differ; the kernels then search chunk_first)
};
void launch_tile_pass(hipStream_t s, const TileLaunch &t, int pass);
// One Newton pass with the previous pass's step folded in (grid items, row tables in LDS).
void launch_tile_fused(hipStream_t s, const TileLaunch &t);
// Work-groups are dispatched to the 8 XCDs round-robin (block b to XCD b % 8), and each XCD
// has its own L2.
constexpr unsigned kXcds = 8;
// All passes of one evaluation in one cooperative launch (lik_persist_kernel).
constexpr int kPersistTasks = 8;    // tasks one block holds
constexpr int kPersistPasses = 51;  // pass 0 and at most 50 Newton steps (likelihood.c:284)
struct PersistLaunch {
    TileLaunch t;              // items, facets, outputs; r_lo = 1, nr_g = all scalelengths
    double *part;              // [kPersistPasses][t.nwg][8] chunk sums
    unsigned *cnt;             // per task (item * kMaxR + r): chunks done, zero at launch
    unsigned nbx;              // blocks per XCD
    unsigned spin_limit;       // polls of one wait before the block gives up
};
unsigned persist_capacity(unsigned lds, int kind);
bool launch_persist(hipStream_t s, const PersistLaunch &p, unsigned blocks);
// The two launches of a pass separately: the chunk sums, then the per-task step.
void launch_tile_sums(hipStream_t s, const TileLaunch &t, int pass);
void launch_tile_step(hipStream_t s, const TileLaunch &t, int pass);

// Residual update for one fitted galaxy: R -= model(e) - model(0).
void launch_subtract(hipStream_t s, int prec, unsigned nch, unsigned long nrow, const double *u,
                     const double *v, const double *w, const double *wn,
                     const double *inv_lambda, const double *spec,
                     const GalaxyGeom &fit, const GalaxyGeom &round, void *res);

// Drop-in helpers on plain device arrays.
void launch_model_at_zero(hipStream_t s, double e1, double e2, double scale, unsigned n,
                          const double *uu, const double *vv, const double *s2, double *mod,
                          double *partials);
void launch_galaxy_vis(hipStream_t s, double spectra, double wavenumber, const GalaxyGeom &g,
                       unsigned long n, const double *u, const double *v, const double *w,
                       double *vis);
void launch_phase_shift(hipStream_t s, double wavenumber, double l, double m, unsigned long n,
                        const double *u, const double *v, const double *w, double *vis);
void launch_marginalise_r(hipStream_t s, int num, const double *x, const double *y,
                          double *out, double *qag_ws, int *qag_iws);
void launch_marg_position(hipStream_t s, int n, const double *x, double *out);

// Row permutation / packing at observation set-up.
void launch_gather_rows(hipStream_t s, unsigned long nrow, const uint32_t *perm,
                        const double *src, double *dst);
void launch_pack_sigma(hipStream_t s, unsigned nch, unsigned long nrow, const uint32_t *perm,
                       const float *sigma2, const unsigned char *flag, float *s2f);
void launch_residual_in(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                        const uint32_t *perm, const double *D, const double *S, void *res);
void launch_residual_out(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                         const uint32_t *perm, const void *res, double *out);
void launch_morton_keys(hipStream_t s, unsigned long nrow, const double *u, const double *v,
                        double umax, uint32_t *keys, uint32_t *idx);
bool sort_rows(hipStream_t s, unsigned long nrow, uint32_t *keys, uint32_t *idx,
               uint32_t *keys_alt, uint32_t *idx_alt, void *tmp, size_t tmp_bytes,
               uint32_t **sorted_idx);
size_t sort_rows_tmp_bytes(unsigned long nrow);

// Host-side error reporting shared by the ABI layer.
void set_error(const std::string &msg);
bool hip_ok(hipError_t e, const char *what);

}  // namespace rl2
