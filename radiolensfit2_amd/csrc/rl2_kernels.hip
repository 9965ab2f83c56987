// rl2_kernels.hip -- gfx950 kernels of the per-galaxy shape-measurement path.
//
// Residual layout: channel-major R[c * nrow + i] over rows i in the observation's
// spatially sorted order (Morton order of (u, v)), complex128 (RL2_PREC_FP64) or
// complex64 (RL2_PREC_MIXED).  sigma^2 is one float per visibility with the sign bit
// marking a flagged visibility (sigma^2 > 0 after defaulting, RadioLensfit2.cpp:266-269).
#include <hip/hip_runtime.h>
#include <map>

#include <algorithm>
#include <cmath>
#include <hipcub/hipcub.hpp>

#include "rl2_internal.h"
#include "rl2_gk41.h"

namespace rl2 {

constexpr uint32_t kEmpty = 0xffffffffu;

// ------------------------------------------------------------------ small helpers
__device__ inline double2 load_res(const void *res, int prec, size_t k)
{
    if (prec == 0) return reinterpret_cast<const double2 *>(res)[k];
    float2 f = reinterpret_cast<const float2 *>(res)[k];
    return make_double2((double)f.x, (double)f.y);
}
__device__ inline void store_res(void *res, int prec, size_t k, double2 v)
{
    if (prec == 0)
        reinterpret_cast<double2 *>(res)[k] = v;
    else
        reinterpret_cast<float2 *>(res)[k] = make_float2((float)v.x, (float)v.y);
}

// Round-model / fitted-model visibility of one galaxy in one channel at one row
// (galaxy_visibilities.c:154-190), returned with its phase rotor.
__device__ inline double model_shape(const GalaxyGeom &g, double u, double v, double wn2,
                                     double spectra)
{
    double k1 = (1. + g.e1) * u + g.e2 * v;
    double k2 = g.e2 * u + (1. - g.e1) * v;
    double shape = sersic_shape(k1, k2, g.sf, wn2);
    return shape * (spectra * g.flux);
}

// 1 / sqrt(a) for a >= 1: the hardware estimate refined once, r + r e (1/2 + 3/8 e) with
// e = 1 - a r^2 (bit-identical to rsqrt(), which adds a class test of the estimate for
// zero, infinite and denormal arguments that a >= 1 never has).
__device__ inline double rsqrt_ge1(double a)
{
    const double r = __builtin_amdgcn_rsq(a);
    const double e = fma(-a * r, r, 1.);
    return fma(r * e, fma(e, 0.375, 0.5), r);
}

// The same model with (1 + s w |k|^2)^(-3/2) as rsqrt(den)^3 (one reciprocal square
// root instead of a square root and a division; a few ulp from the reference's form).
__device__ inline double model_shape_rsq(const GalaxyGeom &g, double u, double v, double wn2,
                                         double spectra)
{
    const double k1 = fma(g.e2, v, (1. + g.e1) * u);
    const double k2 = fma(1. - g.e1, v, g.e2 * u);
    const double rr = rsqrt_ge1(fma(g.sf * wn2, fma(k1, k1, k2 * k2), 1.));
    return rr * rr * rr * (spectra * g.flux);
}

// One visibility of the post-fit update (data_processing.c:109-135): R = (R - fit) + round,
// both models with the phase rotor of the fitted galaxy's position (p = u l + v m + w n,
// the phase k p in revolutions is p / lambda).
__device__ inline double2 subtract_vis(const GalaxyGeom &fit, const GalaxyGeom &rnd, double p,
                                       double u, double v, double k2, double spectra,
                                       double inv_lambda, double2 r)
{
    double sn, cs;
    sincos_rev64(p * inv_lambda, sn, cs);
    const double sr = model_shape_rsq(rnd, u, v, k2, spectra);
    const double sfit = model_shape_rsq(fit, u, v, k2, spectra);
    r.x = (r.x - sfit * cs) + sr * cs;
    r.y = (r.y - sfit * sn) + sr * sn;
    return r;
}

// ------------------------------------------------------------------ sources / noise
// R += sign * sum_g flux_g spec_c shape_g exp(i k_c (u l + v m + w n)).
// One thread per row; galaxies staged through LDS in tiles of 128.
template <int PREC>
__global__ __launch_bounds__(256) void add_sources_kernel(unsigned nch, unsigned long nrow,
                                                          const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ wn,
                                                          const double *__restrict__ spec,
                                                          const GalaxyGeom *__restrict__ gal,
                                                          unsigned ngal, double sign, void *res)
{
    constexpr int TILE = 128;
    constexpr int CB = 8;  // channels per register block
    __shared__ GalaxyGeom sg[TILE];
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    bool ok = i < nrow;
    double ui = ok ? u[i] : 0., vi = ok ? v[i] : 0., wi = ok ? w[i] : 0.;
    for (unsigned c0 = 0; c0 < nch; c0 += CB) {
        double are[CB], aim[CB], wnc[CB], wn2[CB], spc[CB];
        int nc = min((int)(nch - c0), CB);
#pragma unroll
        for (int k = 0; k < CB; k++) {
            are[k] = 0.;
            aim[k] = 0.;
            wnc[k] = (k < nc) ? wn[c0 + k] : 0.;
            wn2[k] = wnc[k] * wnc[k];
            spc[k] = (k < nc) ? spec[c0 + k] : 0.;
        }
        for (unsigned g0 = 0; g0 < ngal; g0 += TILE) {
            __syncthreads();
            for (unsigned t = threadIdx.x; t < TILE && g0 + t < ngal; t += blockDim.x)
                sg[t] = gal[g0 + t];
            __syncthreads();
            unsigned gn = min((unsigned)TILE, ngal - g0);
            for (unsigned t = 0; t < gn; t++) {
                const GalaxyGeom g = sg[t];
                double p = ui * g.l0 + vi * g.m0 + wi * g.n0;
#pragma unroll
                for (int k = 0; k < CB; k++) {
                    if (k < nc) {
                        double phase = wnc[k] * p;
                        double shape = model_shape(g, ui, vi, wn2[k], spc[k]);
                        double sn, cs;
                        sincos(phase, &sn, &cs);
                        are[k] += shape * cs;
                        aim[k] += shape * sn;
                    }
                }
            }
        }
        if (ok) {
            for (int k = 0; k < nc; k++) {
                size_t idx = (size_t)(c0 + k) * nrow + i;
                double2 r = load_res(res, PREC, idx);
                r.x += sign * are[k];
                r.y += sign * aim[k];
                store_res(res, PREC, idx, r);
            }
        }
    }
}

void launch_add_sources(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                        const double *u, const double *v, const double *w, const double *wn,
                        const double *spec, const GalaxyGeom *gal, unsigned ngal, double sign,
                        void *res)
{
    dim3 grid((unsigned)((nrow + 255) / 256));
    if (prec == 0)
        add_sources_kernel<0><<<grid, 256, 0, s>>>(nch, nrow, u, v, w, wn, spec, gal, ngal,
                                                    sign, res);
    else
        add_sources_kernel<1><<<grid, 256, 0, s>>>(nch, nrow, u, v, w, wn, spec, gal, ngal,
                                                    sign, res);
}

template <int PREC>
__global__ void add_noise_kernel(unsigned long nvis, double sigma, uint64_t seed, void *res)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nvis) return;
    // Philox4x32-10 (rl2_math.h) of counter (i, 0x5eed, 0) under the seed -> Box-Muller
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0x5eedu, 0u};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t rr[4];
    philox4x32_10(ctr, key, rr);
    const uint4 r = make_uint4(rr[0], rr[1], rr[2], rr[3]);
    // 53-bit uniforms in (0, 1]
    double u1 = ((double)(((uint64_t)r.x << 21) ^ (r.y >> 11)) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(((uint64_t)r.z << 21) ^ (r.w >> 11)) + 0.5) * (1.0 / 9007199254740992.0);
    double rad = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    double2 x = load_res(res, PREC, i);
    x.x += sigma * rad * cs;
    x.y += sigma * rad * sn;
    store_res(res, PREC, i, x);
}

void launch_add_noise(hipStream_t s, int prec, unsigned long nvis, double sigma, uint64_t seed,
                      void *res)
{
    dim3 grid((unsigned)((nvis + 255) / 256));
    if (prec == 0)
        add_noise_kernel<0><<<grid, 256, 0, s>>>(nvis, sigma, seed, res);
    else
        add_noise_kernel<1><<<grid, 256, 0, s>>>(nvis, sigma, seed, res);
}

// ------------------------------------------------------------------ extraction + gridding
// Deterministic accumulation.  The gridded sums of a cell are kept as integers, so the
// order in which work-groups and threads add them does not change a single bit:
//  - a value x (a visibility component, or sigma^2) is carried as rint(x * 2^S), with S
//    chosen per launch from a bound B >= |x| (the residual's bound plus the galaxy's flux;
//    the largest sigma^2) so that B * 2^S * (rows of a work-group) <= 2^62;
//  - a work-group adds its runs in LDS as int64 (the slot sums of a channel cover at most
//    the work-group's rows, so they stay below 2^62);
//  - the flush splits each slot sum into its high word (arithmetic shift by 32) and its low
//    32 bits, added by integer atomics into two 64-bit accumulators per cell and component
//    (no carries are lost: the low sums stay below 2^64 for fewer than 2^32 flushes);
//  - compaction recombines hi * 2^32 + lo and scales by 2^-S.
// Each visibility is rounded on its own (runs of a cell are then added as integers), and
// every gridding path takes S from one bound (the most terms any of them puts in a slot
// between flushes), so the facet is the same bit for bit whatever the variant, the batch
// path, the order of work-groups or the run.  A term's rounding is at most 2^-(S+1), below
// B * 2^-48 for 16 channels, and the sum itself is exact, so the facet agrees with the
// reference's sequential FP64 sums to their own rounding.
// Accumulator layout: 7 arrays of facet^2 words: re hi, re lo, im hi, im lo, sigma^2 hi,
// sigma^2 lo (uint64, two's complement), weight (double, integer-valued).
enum { kAccReHi = 0, kAccReLo, kAccImHi, kAccImLo, kAccS2Hi, kAccS2Lo, kAccW, kAccWords };

__device__ inline long long fx_round(double x, double scale)
{
    return (long long)rint(x * scale);
}

__device__ inline void fx_flush(unsigned long long *hi, unsigned long long *lo, long long v)
{
    atomicAdd(hi, (unsigned long long)(v >> 32));
    atomicAdd(lo, (unsigned long long)(v & 0xffffffffll));
}

// Variants 2-4.  Each thread owns RPT consecutive sorted rows and, per channel, adds runs
// of consecutive visibilities that fall in the same cell in registers before one LDS
// update per run: in the dense uv core, where many rows share a cell, this removes most of
// the same-address LDS atomics that serialise across a wave.  A channel inserts at most
// RPT*256 runs into the HS-slot table (HS >= RPT*256, so probing always finds a slot);
// flush as variant 0.
// SUB: first apply the previous galaxy's post-fit update (a.sfit, a.srnd) to each of the
// thread's visibilities and write it back, then grid the updated value: the subtraction and
// the next extraction share one pass over the residual, and every visibility gets the same
// arithmetic, in the same order, as subtract_kernel followed by this kernel.
// LAZY: the table keeps its cells from one channel to the next and is flushed only when
// another channel's runs might not fit (more than HS - RPT*256 slots in use) and after the
// last channel.  Neighbouring channels put a dense-core visibility in the same cell, so the
// global atomics of a cell are made once for several channels; the integer sums make the
// facet the same bit for bit whenever the flushes happen.
template <int MODE, int PREC, int RPT, int HS, bool SUB = false, bool LAZY = false>
__global__ __launch_bounds__(256) void grid_run_kernel(GridLaunch a)
{
    constexpr int BS = 256;
    static_assert(HS >= RPT * BS, "a free slot must exist for every run of a channel");
    constexpr int HBITS = (HS == 2048) ? 11 : (HS == 1024) ? 10 : (HS == 4096) ? 12 : 9;
    static_assert((1 << HBITS) == HS, "hash size must be a power of two");
    static_assert(RPT * BS <= 4096, "slot sums of a channel must stay below 2^62");
    __shared__ uint32_t hkey[HS];
    __shared__ unsigned long long hre[HS], him[HS], hs2[HS];  // rint(x * 2^S), int64
    __shared__ uint32_t hcnt[HS];
    __shared__ uint32_t hlist[HS];
    __shared__ uint32_t hn;
    const int tid = threadIdx.x;
    for (int s = tid; s < HS; s += BS) {
        hkey[s] = kEmpty;
        hre[s] = 0ull;
        him[s] = 0ull;
        hs2[s] = 0ull;
        hcnt[s] = 0;
    }
    if (tid == 0) hn = 0;
    const unsigned long row0 = (unsigned long)blockIdx.x * (RPT * BS) + (unsigned long)tid * RPT;
    const unsigned long ncell = (unsigned long)a.facet * a.facet;
    const double inv_inc = 1.0 / a.inc;
    unsigned long long *acc = (unsigned long long *)a.acc;
    double *acc_w = a.acc + kAccW * ncell;
    // a run's sums are already integers (each visibility rounded on its own, so the run
    // length, and with it the variant, does not change a bit)
    auto emit = [&](uint32_t key, long long xr, long long xi, long long xs, uint32_t n) {
        uint32_t h = (key * 2654435761u) >> (32 - HBITS);
        while (true) {
            const uint32_t prev = atomicCAS(&hkey[h], kEmpty, key);
            if (prev == kEmpty) hlist[atomicAdd(&hn, 1u)] = h;
            if (prev == kEmpty || prev == key) break;
            h = (h + 1) & (HS - 1);
        }
        if (MODE != 2) {
            atomicAdd(&hre[h], (unsigned long long)xr);
            atomicAdd(&him[h], (unsigned long long)xi);
        }
        atomicAdd(&hs2[h], (unsigned long long)xs);
        atomicAdd(&hcnt[h], n);
    };
    // the rows' u, v and phase numerator p = ul + vm + wn stay in registers for all
    // channels (the channel only scales them)
    double ru[RPT], rv[RPT], rp[RPT], rps[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const unsigned long i = row0 + k;
        const bool in = i < a.nrow;
        ru[k] = in ? a.u[i] : 0.;
        rv[k] = in ? a.v[i] : 0.;
        const double wi = (in && MODE == 0) ? a.w[i] : 0.;
        rp[k] = (in && MODE == 0) ? ru[k] * a.g.l0 + rv[k] * a.g.m0 + wi * a.g.n0 : 0.;
        rps[k] = (in && SUB) ? ru[k] * a.sfit.l0 + rv[k] * a.sfit.m0 + wi * a.sfit.n0 : 0.;
    }
    __syncthreads();
    for (unsigned c = 0; c < a.nch; c++) {
        const double wn = a.wn[c], inv_lambda = a.inv_lambda[c];
        const double wn2 = wn * wn, spectra = a.spec[c];
        uint32_t rkey = kEmpty, rn = 0;
        long long rre = 0, rim = 0, rs2 = 0;
        // the thread's RPT consecutive visibilities are loaded together, so each load
        // instruction's cache lines are still resident for the next one (issued one row
        // at a time between the gridding work, they were fetched from HBM ~3 times over)
        float vs2[RPT];
        double2 vres[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const unsigned long i = row0 + k;
            const size_t idx = (size_t)c * a.nrow + i;
            vs2[k] = i < a.nrow ? a.s2f[idx] : -1.f;
            vres[k] = (MODE != 2 && i < a.nrow) ? load_res(a.res, PREC, idx) : make_double2(0., 0.);
        }
        if (SUB) {
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                const unsigned long i = row0 + k;
                if (i >= a.nrow) continue;
                const size_t idx = (size_t)c * a.nrow + i;
                const double2 r = subtract_vis(a.sfit, a.srnd, rps[k], ru[k], rv[k], wn2,
                                               spectra, inv_lambda, vres[k]);
                store_res(a.res_w, PREC, idx, r);
                // the value a separate pass would read back
                vres[k] = PREC == 0 ? r : make_double2((double)(float)r.x, (double)(float)r.y);
            }
        }
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const unsigned long i = row0 + k;
            if (i >= a.nrow) continue;
            const float s2 = vs2[k];
            if (signbit(s2)) continue;  // flagged
            const double u = ru[k], v = rv[k];
            uint32_t key;
            if (!cell_index_fast(u, v, inv_lambda, a.len, a.inc, inv_inc, a.facet, &key)) continue;
            double xr = 0., xi = 0.;
            if (MODE == 0) {
                // visGal = round model + (visData - visSkyMod), then shift by -k(ul+vm+wn)
                // (source_extraction.c:51-65); the phase k p in revolutions is p / lambda.
                const double p = rp[k];
                const double shape = model_shape_rsq(a.g, u, v, wn2, spectra);
                double sn, cs;
                sincos_rev64(p * inv_lambda, sn, cs);
                const double2 r = vres[k];
                const double gre = shape * cs + r.x;
                const double gim = shape * sn + r.y;
                xr = gre * cs - gim * (-sn);
                xi = gre * (-sn) + gim * cs;
            } else if (MODE == 1) {
                const double2 r = vres[k];
                xr = r.x;
                xi = r.y;
            }
            const long long ir = MODE != 2 ? fx_round(xr, a.fx_scale) : 0;
            const long long ii = MODE != 2 ? fx_round(xi, a.fx_scale) : 0;
            const long long is = fx_round((double)s2, a.fs_scale);
            if (key == rkey) {
                rre += ir;
                rim += ii;
                rs2 += is;
                rn++;
            } else {
                if (rn) emit(rkey, rre, rim, rs2, rn);
                rkey = key;
                rre = ir;
                rim = ii;
                rs2 = is;
                rn = 1;
            }
        }
        if (rn) emit(rkey, rre, rim, rs2, rn);
        __syncthreads();
        const uint32_t nocc = hn;
        if (LAZY && c + 1 < a.nch && nocc <= (uint32_t)(HS - RPT * BS)) continue;
        for (uint32_t j = tid; j < nocc; j += BS) {
            const uint32_t s = hlist[j];
            const uint32_t key = hkey[s];
            if (MODE != 2) {
                fx_flush(acc + kAccReHi * ncell + key, acc + kAccReLo * ncell + key,
                         (long long)hre[s]);
                fx_flush(acc + kAccImHi * ncell + key, acc + kAccImLo * ncell + key,
                         (long long)him[s]);
            }
            fx_flush(acc + kAccS2Hi * ncell + key, acc + kAccS2Lo * ncell + key,
                     (long long)hs2[s]);
            unsafeAtomicAdd(&acc_w[key], (double)hcnt[s]);  // integer-valued: exact
            hkey[s] = kEmpty;
            hre[s] = 0ull;
            him[s] = 0ull;
            hs2[s] = 0ull;
            hcnt[s] = 0;
        }
        __syncthreads();
        if (tid == 0) hn = 0;
        __syncthreads();
    }
}

template <int MODE, int PREC>
static void launch_grid_variant(hipStream_t s, const GridLaunch &a)
{
    if (tuning().grid_variant == 5) {
        // runs of 2 rows, 1024 slots, flushed when another channel might not fit
        constexpr int RPT = 2;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        if (MODE == 0 && a.sub)
            grid_run_kernel<MODE, PREC, RPT, 1024, true, true><<<grid, 256, 0, s>>>(a);
        else
            grid_run_kernel<MODE, PREC, RPT, 1024, false, true><<<grid, 256, 0, s>>>(a);
    } else if (tuning().grid_variant == 6) {
        // runs of 4 rows, 2048 slots, flushed when another channel might not fit
        constexpr int RPT = 4;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        if (MODE == 0 && a.sub)
            grid_run_kernel<MODE, PREC, RPT, 2048, true, true><<<grid, 256, 0, s>>>(a);
        else
            grid_run_kernel<MODE, PREC, RPT, 2048, false, true><<<grid, 256, 0, s>>>(a);
    } else if (tuning().grid_variant == 2) {
        constexpr int RPT = 8;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        grid_run_kernel<MODE, PREC, RPT, 2048><<<grid, 256, 0, s>>>(a);
    } else if (tuning().grid_variant == 3) {
        constexpr int RPT = 4;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        if (MODE == 0 && a.sub)
            grid_run_kernel<MODE, PREC, RPT, 1024, true><<<grid, 256, 0, s>>>(a);
        else
            grid_run_kernel<MODE, PREC, RPT, 1024><<<grid, 256, 0, s>>>(a);
    } else if (tuning().grid_variant == 4) {
        constexpr int RPT = 4;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        grid_run_kernel<MODE, PREC, RPT, 2048><<<grid, 256, 0, s>>>(a);
    } else if (tuning().grid_variant == 1) {
        constexpr int RPT = 2;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        grid_run_kernel<MODE, PREC, RPT, 1024><<<grid, 256, 0, s>>>(a);
    } else {
        // runs of 2 rows: 512 rows per work-group and a 512-slot table (18 KB of LDS)
        constexpr int RPT = 2;
        dim3 grid((unsigned)((a.nrow + RPT * 256 - 1) / (RPT * 256)));
        grid_run_kernel<MODE, PREC, RPT, 512><<<grid, 256, 0, s>>>(a);
    }
}

void launch_grid(hipStream_t s, const GridLaunch &a)
{
    if (a.mode == 0) {
        if (a.prec == 0)
            launch_grid_variant<0, 0>(s, a);
        else
            launch_grid_variant<0, 1>(s, a);
    } else if (a.mode == 1) {
        launch_grid_variant<1, 0>(s, a);
    } else {
        launch_grid_variant<2, 0>(s, a);
    }
}

// ------------------------------------------------------------------ compaction
// Cells with weight != 0 in row-major order (evaluate_uv_grid.c:196-210 and :113-133):
//   vis = sum_vis * w / sum_sigma2, sigma2 = sum_sigma2 / w^2, (u,v) = cell centre.
constexpr int kCompactBS = 256, kCompactPT = 16, kCompactTile = kCompactBS * kCompactPT;

size_t compact_scratch_words(unsigned facet)
{
    size_t ncell = (size_t)facet * facet;
    return (ncell + kCompactTile - 1) / kCompactTile + 1;
}

__global__ __launch_bounds__(kCompactBS) void compact_count_kernel(const double *acc_w,
                                                                   unsigned long ncell,
                                                                   unsigned *bcount)
{
    __shared__ unsigned part[kCompactBS / 64];
    unsigned long base = (unsigned long)blockIdx.x * kCompactTile + threadIdx.x * kCompactPT;
    unsigned cnt = 0;
    for (int k = 0; k < kCompactPT; k++) {
        unsigned long p = base + k;
        if (p < ncell && acc_w[p] != 0.) cnt++;
    }
    for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int i = 0; i < kCompactBS / 64; i++) t += part[i];
        bcount[blockIdx.x] = t;
    }
}

// exclusive scan of nb block counts in place (single work-group), total -> bcount[nb]
__global__ __launch_bounds__(1024) void compact_scan_kernel(unsigned *bcount, unsigned nb,
                                                            unsigned *n_out)
{
    __shared__ unsigned buf[1024];
    __shared__ unsigned carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (unsigned base = 0; base < nb; base += 1024) {
        unsigned i = base + threadIdx.x;
        unsigned x = (i < nb) ? bcount[i] : 0;
        buf[threadIdx.x] = x;
        __syncthreads();
        for (unsigned off = 1; off < 1024; off <<= 1) {
            unsigned t = (threadIdx.x >= off) ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bcount[i] = carry + buf[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bcount[nb] = carry;
        *n_out = carry;
    }
}

// Accumulators of cell p: FX (device gridding) the integer layout above, scaled back by
// 2^-S; otherwise four double arrays (re, im, sigma^2, weight) given by a caller.
template <bool FX>
__device__ inline void acc_read(double *acc, unsigned long ncell, unsigned long p,
                                double fx_unscale, double fs_unscale, double &re, double &im,
                                double &sg)
{
    if (FX) {
        const unsigned long long *q = (const unsigned long long *)acc;
        auto get = [&](int hi, int lo, double unscale) {
            const unsigned long long l = q[lo * ncell + p];
            const long long h = (long long)q[hi * ncell + p] + (long long)(l >> 32);
            return ((double)h * 4294967296.0 + (double)(l & 0xffffffffull)) * unscale;
        };
        re = get(kAccReHi, kAccReLo, fx_unscale);
        im = get(kAccImHi, kAccImLo, fx_unscale);
        sg = get(kAccS2Hi, kAccS2Lo, fs_unscale);
    } else {
        re = acc[p];
        im = acc[ncell + p];
        sg = acc[2 * ncell + p];
    }
}

template <bool FX>
__device__ inline void acc_clear(double *acc, unsigned long ncell, unsigned long p)
{
    if (FX) {
        unsigned long long *q = (unsigned long long *)acc;
        for (int k = 0; k < kAccW; k++) q[k * ncell + p] = 0ull;
        acc[kAccW * ncell + p] = 0.;
    } else {
        for (int k = 0; k < 4; k++) acc[k * ncell + p] = 0.;
    }
}

// One work-group's tile of the compaction (block `blk` of the nb blocks of one facet).
template <bool FX>
__device__ inline void compact_write_tile(
    unsigned blk, unsigned nb, double *acc, unsigned facet, unsigned long ncell, double len,
    const unsigned *boff, double *uu, double *vv, double *dre, double *dim, double *s2o,
    uint32_t *cid, double *ctr, double *sum_w_dense, double fx_unscale, double fs_unscale,
    unsigned *tscan)
{
    const double *acc_w = acc + (FX ? kAccW : 3) * ncell;
    unsigned long base = (unsigned long)blk * kCompactTile + threadIdx.x * kCompactPT;
    unsigned cnt = 0;
    for (int k = 0; k < kCompactPT; k++) {
        unsigned long p = base + k;
        if (p < ncell && acc_w[p] != 0.) cnt++;
    }
    tscan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kCompactBS; off <<= 1) {
        unsigned t = (threadIdx.x >= (unsigned)off) ? tscan[threadIdx.x - off] : 0;
        __syncthreads();
        tscan[threadIdx.x] += t;
        __syncthreads();
    }
    unsigned n = boff[blk] + tscan[threadIdx.x] - cnt;
    const double inc = 2 * len / facet;
    if (cid && blk == 0 && threadIdx.x == 0) {
        // zero pad cell after the last one (the arrays hold facet^2 + 1 entries): lanes past
        // the end of a chunk read it instead of masking their data
        const unsigned nt = boff[nb];
        dre[nt] = 0.;
        dim[nt] = 0.;
        cid[nt] = 0;
    }
    if (ctr) {  // centre of grid index k, the same expression as uu / vv below
        const unsigned k = blk * kCompactBS + threadIdx.x;
        if (k < facet) ctr[k] = (k + 0.5) * inc - len;
    }
    for (int k = 0; k < kCompactPT; k++) {
        unsigned long p = base + k;
        if (p >= ncell) break;
        double wgt = acc_w[p];
        if (sum_w_dense) sum_w_dense[p] = wgt;
        if (wgt != 0.) {
            double re, im, sg;
            acc_read<FX>(acc, ncell, p, fx_unscale, fs_unscale, re, im, sg);
            dre[n] = re * wgt / sg;
            dim[n] = im * wgt / sg;
            s2o[n] = sg / (wgt * wgt);
            unsigned j = (unsigned)(p / facet), i = (unsigned)(p % facet);
            uu[n] = (i + 0.5) * inc - len;
            vv[n] = (j + 0.5) * inc - len;
            if (cid) cid[n] = (j << 16) | i;
            n++;
            acc_clear<FX>(acc, ncell, p);
        }
    }
}

template <bool FX>
__global__ __launch_bounds__(kCompactBS) void compact_write_kernel(
    double *acc, unsigned facet, unsigned long ncell, double len, const unsigned *boff,
    double *uu, double *vv, double *dre, double *dim, double *s2o, uint32_t *cid, double *ctr,
    double *sum_w_dense, double fx_unscale, double fs_unscale)
{
    __shared__ unsigned tscan[kCompactBS];
    compact_write_tile<FX>(blockIdx.x, gridDim.x, acc, facet, ncell, len, boff, uu, vv, dre, dim,
                           s2o, cid, ctr, sum_w_dense, fx_unscale, fs_unscale, tscan);
}

// The packed data and record offsets of the fused Newton passes (cell i and the zero pad
// cell n); offsets only for grid facets with records (facet <= kRecFacet).
__device__ inline void pack_cell_records(unsigned i, unsigned n, const double *dre,
                                         const double *dim, const uint32_t *cid, unsigned facet,
                                         double2 *dc, uint32_t *loff)
{
    const bool in = i < n;
    dc[i] = in ? make_double2(dre[i], dim[i]) : make_double2(0., 0.);
    if (cid && facet <= kRecFacet) {
        const uint32_t c = in ? cid[i] : 0u;
        loff[i] = ((c >> 16) * kRecBytes) << 16 | (c & 0xffffu) * kRecBytes;
    }
}

// ---- batched compaction: grid.y = target of a BatchDesc; targets beyond desc->ntarget and
// blocks beyond a target's facet exit at once, so one launch shape serves every batch.
__global__ __launch_bounds__(kCompactBS) void compact_count_batch_kernel(const BatchDesc *desc)
{
    const unsigned t = blockIdx.y;
    if (t >= desc->ntarget) return;
    const BatchCompact &C = desc->compact[t];
    const unsigned long ncell = (unsigned long)C.facet * C.facet;
    const unsigned nb = (unsigned)((ncell + kCompactTile - 1) / kCompactTile);
    if (blockIdx.x >= nb) return;
    __shared__ unsigned part[kCompactBS / 64];
    const double *acc_w = C.acc + kAccW * ncell;
    unsigned long base = (unsigned long)blockIdx.x * kCompactTile + threadIdx.x * kCompactPT;
    unsigned cnt = 0;
    for (int k = 0; k < kCompactPT; k++) {
        unsigned long p = base + k;
        if (p < ncell && acc_w[p] != 0.) cnt++;
    }
    for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned tt = 0;
        for (int i = 0; i < kCompactBS / 64; i++) tt += part[i];
        C.scan[blockIdx.x] = tt;
    }
}

__global__ __launch_bounds__(1024) void compact_scan_batch_kernel(const BatchDesc *desc)
{
    const unsigned t = blockIdx.x;
    if (t >= desc->ntarget) return;
    const BatchCompact &C = desc->compact[t];
    const unsigned long ncell = (unsigned long)C.facet * C.facet;
    const unsigned nb = (unsigned)((ncell + kCompactTile - 1) / kCompactTile);
    unsigned *bcount = C.scan;
    __shared__ unsigned buf[1024];
    __shared__ unsigned carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (unsigned base = 0; base < nb; base += 1024) {
        unsigned i = base + threadIdx.x;
        unsigned x = (i < nb) ? bcount[i] : 0;
        buf[threadIdx.x] = x;
        __syncthreads();
        for (unsigned off = 1; off < 1024; off <<= 1) {
            unsigned tt = (threadIdx.x >= off) ? buf[threadIdx.x - off] : 0;
            __syncthreads();
            buf[threadIdx.x] += tt;
            __syncthreads();
        }
        if (i < nb) bcount[i] = carry + buf[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bcount[nb] = carry;
        *C.n_out = carry;
    }
}

__global__ __launch_bounds__(kCompactBS) void compact_write_batch_kernel(const BatchDesc *desc,
                                                                          double len)
{
    const unsigned t = blockIdx.y;
    if (t >= desc->ntarget) return;
    const BatchCompact &C = desc->compact[t];
    const unsigned long ncell = (unsigned long)C.facet * C.facet;
    const unsigned nb = (unsigned)((ncell + kCompactTile - 1) / kCompactTile);
    if (blockIdx.x >= nb) return;
    __shared__ unsigned tscan[kCompactBS];
    compact_write_tile<true>(blockIdx.x, nb, C.acc, C.facet, ncell, len, C.scan, C.uu, C.vv,
                             C.dre, C.dim, C.s2, C.cid, C.ctr, nullptr, C.fx_unscale,
                             C.fs_unscale, tscan);
}

__global__ void pack_facet_batch_kernel(const BatchDesc *desc)
{
    const unsigned t = blockIdx.y;
    if (t >= desc->ntarget) return;
    const BatchCompact &C = desc->compact[t];
    const unsigned n = *C.n_out;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        C.pk[i] = make_float4((float)C.uu[i], (float)C.vv[i], (float)C.dre[i], (float)C.dim[i]);
        const double w = 1.0 / C.s2[i];
        C.invs2[i] = (float)w;
        C.invs2_64[i] = w;
        pack_cell_records(i, n, C.dre, C.dim, C.cid, C.facet, C.dc, C.loff);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the zero pad cell
        C.invs2_64[n] = 0.;
        pack_cell_records(n, n, C.dre, C.dim, C.cid, C.facet, C.dc, C.loff);
    }
}

void launch_compact_batch(hipStream_t s, const BatchDesc *desc, unsigned ncap, unsigned fmax,
                          double len)
{
    const unsigned long ncell = (unsigned long)fmax * fmax;
    const unsigned nb = (unsigned)((ncell + kCompactTile - 1) / kCompactTile);
    compact_count_batch_kernel<<<dim3(nb, ncap), kCompactBS, 0, s>>>(desc);
    compact_scan_batch_kernel<<<ncap, 1024, 0, s>>>(desc);
    compact_write_batch_kernel<<<dim3(nb, ncap), kCompactBS, 0, s>>>(desc, len);
    const unsigned pb = std::min(256u, (unsigned)((ncell + 255) / 256) + 1);
    pack_facet_batch_kernel<<<dim3(pb, ncap), 256, 0, s>>>(desc);
}

size_t grid_acc_words(unsigned facet) { return (size_t)kAccWords * facet * facet; }

// ------------------------------------------------------------------ batched extraction
// One persistent pass over the residual for a whole batch (GridBatchLaunch).  A work-group
// takes the next tile of kBatchTileRows sorted rows from the queue; each thread owns 4
// consecutive rows.  Per channel: the thread's visibilities are loaded once, the pending
// updates are applied in order and written back, then every target galaxy grids them with
// the arithmetic of grid_run_kernel (runs of equal cells merged in registers, an LDS table
// of integer sums, flushed to the target's accumulators before the next target).  No
// work-group waits for another, so the launch may hold any number of them; the last one
// to finish leaves the queue counters at zero for the next launch.
template <int PREC>
__global__ __launch_bounds__(256) void grid_batch_kernel(GridBatchLaunch a)
{
    constexpr int BS = 256, RPT = 4, HS = 1024, HBITS = 10;
    static_assert(RPT * BS == kBatchTileRows, "tile rows");
    __shared__ uint32_t hkey[HS];
    __shared__ unsigned long long hre[HS], him[HS], hs2[HS];
    __shared__ uint32_t hcnt[HS];
    __shared__ uint32_t hlist[HS];
    __shared__ uint32_t hn;
    __shared__ unsigned s_tile;
    const int tid = threadIdx.x;
    for (int s = tid; s < HS; s += BS) {
        hkey[s] = kEmpty;
        hre[s] = 0ull;
        him[s] = 0ull;
        hs2[s] = 0ull;
        hcnt[s] = 0;
    }
    if (tid == 0) hn = 0;
    const unsigned ntarget = a.desc->ntarget, nsub = a.desc->nsub;
    while (true) {
        if (tid == 0) s_tile = atomicAdd(&a.queue[0], 1u);
        __syncthreads();
        const unsigned tile = s_tile;
        if (tile >= a.ntiles) break;
        const unsigned long row0 = (unsigned long)tile * kBatchTileRows + (unsigned long)tid * RPT;
        double ru[RPT], rv[RPT], rw[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const unsigned long i = row0 + k;
            const bool in = i < a.nrow;
            ru[k] = in ? a.u[i] : 0.;
            rv[k] = in ? a.v[i] : 0.;
            rw[k] = in ? a.w[i] : 0.;
        }
        for (unsigned c = 0; c < a.nch; c++) {
            const double wn = a.wn[c], inv_lambda = a.inv_lambda[c];
            const double wn2 = wn * wn, spectra = a.spec[c];
            float vs2[RPT];
            double2 vres[RPT];
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                const unsigned long i = row0 + k;
                const size_t idx = (size_t)c * a.nrow + i;
                vs2[k] = i < a.nrow ? a.s2f[idx] : -1.f;
                vres[k] = i < a.nrow ? load_res(a.res, PREC, idx) : make_double2(0., 0.);
            }
            if (nsub) {
                for (unsigned b = 0; b < nsub; b++) {
                    const GalaxyGeom fit = a.desc->subs[b].fit, rnd = a.desc->subs[b].rnd;
#pragma unroll
                    for (int k = 0; k < RPT; k++) {
                        const double p = ru[k] * fit.l0 + rv[k] * fit.m0 + rw[k] * fit.n0;
                        vres[k] = subtract_vis(fit, rnd, p, ru[k], rv[k], wn2, spectra,
                                               inv_lambda, vres[k]);
                        // each update a pass of its own: a complex64 residual is stored
                        // between them
                        if (PREC != 0)
                            vres[k] = make_double2((double)(float)vres[k].x,
                                                   (double)(float)vres[k].y);
                    }
                }
#pragma unroll
                for (int k = 0; k < RPT; k++) {
                    const unsigned long i = row0 + k;
                    if (i >= a.nrow) continue;
                    store_res(a.res, PREC, (size_t)c * a.nrow + i, vres[k]);
                }
            }
            for (unsigned t = 0; t < ntarget; t++) {
                const BatchTarget T = a.desc->targets[t];
                uint32_t rkey = kEmpty, rn = 0;
                long long rre = 0, rim = 0, rs2 = 0;  // integer run sums, as grid_run_kernel
                auto emit = [&](uint32_t key, long long xr, long long xi, long long xs,
                                uint32_t n) {
                    uint32_t h = (key * 2654435761u) >> (32 - HBITS);
                    while (true) {
                        const uint32_t prev = atomicCAS(&hkey[h], kEmpty, key);
                        if (prev == kEmpty) hlist[atomicAdd(&hn, 1u)] = h;
                        if (prev == kEmpty || prev == key) break;
                        h = (h + 1) & (HS - 1);
                    }
                    atomicAdd(&hre[h], (unsigned long long)xr);
                    atomicAdd(&him[h], (unsigned long long)xi);
                    atomicAdd(&hs2[h], (unsigned long long)xs);
                    atomicAdd(&hcnt[h], n);
                };
#pragma unroll
                for (int k = 0; k < RPT; k++) {
                    const unsigned long i = row0 + k;
                    if (i >= a.nrow) continue;
                    const float s2 = vs2[k];
                    if (signbit(s2)) continue;  // flagged
                    const double u = ru[k], v = rv[k];
                    uint32_t key;
                    if (!cell_index_fast(u, v, inv_lambda, a.len, T.inc, T.inv_inc, T.facet, &key))
                        continue;
                    // visGal = round model + residual, shifted by -k(ul + vm + wn)
                    // (source_extraction.c:51-65)
                    const double p = u * T.g.l0 + v * T.g.m0 + rw[k] * T.g.n0;
                    const double shape = model_shape_rsq(T.g, u, v, wn2, spectra);
                    double sn, cs;
                    sincos_rev64(p * inv_lambda, sn, cs);
                    const double gre = shape * cs + vres[k].x;
                    const double gim = shape * sn + vres[k].y;
                    const double xr = gre * cs - gim * (-sn);
                    const double xi = gre * (-sn) + gim * cs;
                    const long long ir = fx_round(xr, T.fx_scale);
                    const long long ii = fx_round(xi, T.fx_scale);
                    const long long is = fx_round((double)s2, a.fs_scale);
                    if (key == rkey) {
                        rre += ir;
                        rim += ii;
                        rs2 += is;
                        rn++;
                    } else {
                        if (rn) emit(rkey, rre, rim, rs2, rn);
                        rkey = key;
                        rre = ir;
                        rim = ii;
                        rs2 = is;
                        rn = 1;
                    }
                }
                if (rn) emit(rkey, rre, rim, rs2, rn);
                __syncthreads();
                const unsigned long ncell = (unsigned long)T.facet * T.facet;
                unsigned long long *acc = (unsigned long long *)T.acc;
                double *acc_w = T.acc + kAccW * ncell;
                const uint32_t nocc = hn;
                for (uint32_t j = tid; j < nocc; j += BS) {
                    const uint32_t s = hlist[j];
                    const uint32_t key = hkey[s];
                    fx_flush(acc + kAccReHi * ncell + key, acc + kAccReLo * ncell + key,
                             (long long)hre[s]);
                    fx_flush(acc + kAccImHi * ncell + key, acc + kAccImLo * ncell + key,
                             (long long)him[s]);
                    fx_flush(acc + kAccS2Hi * ncell + key, acc + kAccS2Lo * ncell + key,
                             (long long)hs2[s]);
                    unsafeAtomicAdd(&acc_w[key], (double)hcnt[s]);
                    hkey[s] = kEmpty;
                    hre[s] = 0ull;
                    him[s] = 0ull;
                    hs2[s] = 0ull;
                    hcnt[s] = 0;
                }
                __syncthreads();
                if (tid == 0) hn = 0;
                __syncthreads();
            }
        }
        __syncthreads();  // every thread has read s_tile before thread 0 takes the next tile
    }
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(&a.queue[1], 1u) == gridDim.x - 1) {
            atomicExch(&a.queue[0], 0u);
            atomicExch(&a.queue[1], 0u);
        }
    }
}

void launch_grid_batch(hipStream_t s, const GridBatchLaunch &a, unsigned groups)
{
    groups = std::max(1u, std::min(groups, a.ntiles));
    if (a.prec == 0)
        grid_batch_kernel<0><<<groups, 256, 0, s>>>(a);
    else
        grid_batch_kernel<1><<<groups, 256, 0, s>>>(a);
}

template <int PREC>
__global__ __launch_bounds__(256) void subtract_batch_kernel(
    unsigned nch, unsigned long nrow, const double *__restrict__ u, const double *__restrict__ v,
    const double *__restrict__ w, const double *__restrict__ wn,
    const double *__restrict__ inv_lambda, const double *__restrict__ spec, const BatchSub *subs,
    unsigned nsub, void *res)
{
    const unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    const double ui = u[i], vi = v[i], wi = w[i];
    for (unsigned c = 0; c < nch; c++) {
        const double k = wn[c], k2 = k * k, spectra = spec[c];
        const size_t idx = (size_t)c * nrow + i;
        double2 r = load_res(res, PREC, idx);
        for (unsigned b = 0; b < nsub; b++) {
            const GalaxyGeom fit = subs[b].fit, rnd = subs[b].rnd;
            const double p = ui * fit.l0 + vi * fit.m0 + wi * fit.n0;
            r = subtract_vis(fit, rnd, p, ui, vi, k2, spectra, inv_lambda[c], r);
            if (PREC != 0 && b + 1 < nsub)  // each update a pass of its own: stored as complex64
                r = make_double2((double)(float)r.x, (double)(float)r.y);
        }
        store_res(res, PREC, idx, r);
    }
}

void launch_subtract_batch(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                           const double *u, const double *v, const double *w, const double *wn,
                           const double *inv_lambda, const double *spec, const BatchSub *subs,
                           unsigned nsub, void *res)
{
    if (nsub == 0) return;
    dim3 grid((unsigned)((nrow + 255) / 256));
    if (prec == 0)
        subtract_batch_kernel<0><<<grid, 256, 0, s>>>(nch, nrow, u, v, w, wn, inv_lambda, spec,
                                                      subs, nsub, res);
    else
        subtract_batch_kernel<1><<<grid, 256, 0, s>>>(nch, nrow, u, v, w, wn, inv_lambda, spec,
                                                      subs, nsub, res);
}

// Rows of a gridding work-group (256 threads x rows per thread) for the current variant.
// log2 of the visibilities one LDS slot can receive between flushes: the rows of a
// work-group, times the channels for the variants that keep cells across channels.
// The most terms any gridding path puts in one table slot between flushes, as a power of
// two: 2048 rows (runs of 8), or 1024 rows times every channel (runs of 4 kept across
// channels).  Every path (each variant, the batched kernel) takes its scale from this one
// bound, so their integer sums, and the facets, are the same bit for bit.
static int grid_rows_log2(unsigned nch)
{
    int ch = 0;
    while ((1u << ch) < nch) ch++;
    return std::max(11, 10 + ch);
}

AccScale fx_scales(double vis_bound, double s2_bound, int rows_log2, unsigned nch)
{
    // 2^S with bound * 2^S * (terms a slot takes between flushes) <= 2^62
    if (rows_log2 < 0) rows_log2 = grid_rows_log2(nch);
    auto scale = [&](double b) {
        if (!(b > 0.) || !std::isfinite(b)) b = 1.;
        const int S = std::max(-1000, std::min(1000, 62 - rows_log2 - (ilogb(b) + 1)));
        return ldexp(1.0, S);
    };
    AccScale a;
    a.fx_scale = scale(vis_bound);
    a.fs_scale = scale(s2_bound);
    return a;
}

// Bound max |re| + |im| over the residual (as the bit pattern of a non-negative double, so
// that integer atomicMax orders it).
template <int PREC>
__global__ __launch_bounds__(256) void abs_bound_kernel(const void *res, unsigned long n,
                                                        unsigned long long *out)
{
    double m = 0.;
    for (unsigned long i = blockIdx.x * 256ul + threadIdx.x; i < n; i += (unsigned long)gridDim.x * 256ul) {
        const double2 x = load_res(res, PREC, i);
        m = fmax(m, fabs(x.x) + fabs(x.y));
    }
    for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

void launch_abs_bound(hipStream_t s, int prec, unsigned long nvis, const void *res,
                      unsigned long long *out)
{
    const unsigned grid = (unsigned)std::min<unsigned long>(4096ul, (nvis + 255) / 256);
    if (prec == 0)
        abs_bound_kernel<0><<<grid, 256, 0, s>>>(res, nvis, out);
    else
        abs_bound_kernel<1><<<grid, 256, 0, s>>>(res, nvis, out);
}

void launch_compact(hipStream_t s, double *acc, unsigned facet, double len, double *uu,
                    double *vv, double *dre, double *dim, double *s2, uint32_t *cid, double *ctr,
                    unsigned *n_out, double *sum_w_dense, unsigned *scratch, const AccScale *fx)
{
    unsigned long ncell = (unsigned long)facet * facet;
    unsigned nb = (unsigned)((ncell + kCompactTile - 1) / kCompactTile);
    // ctr needs a thread per grid index: nb * kCompactBS = ceil(facet^2 / 4096) * 256 >= facet
    // for every facet size
    compact_count_kernel<<<nb, kCompactBS, 0, s>>>(acc + (fx ? kAccW : 3) * ncell, ncell, scratch);
    compact_scan_kernel<<<1, 1024, 0, s>>>(scratch, nb, n_out);
    if (fx)
        compact_write_kernel<true><<<nb, kCompactBS, 0, s>>>(
            acc, facet, ncell, len, scratch, uu, vv, dre, dim, s2, cid, ctr, sum_w_dense,
            1.0 / fx->fx_scale, 1.0 / fx->fs_scale);
    else
        compact_write_kernel<false><<<nb, kCompactBS, 0, s>>>(
            acc, facet, ncell, len, scratch, uu, vv, dre, dim, s2, cid, ctr, sum_w_dense, 1.0,
            1.0);
}

// ------------------------------------------------------------------ likelihood (K3)
// One work-group per (item, scalelength): the exponential-disk model at zero
// (galaxy_visibilities.c:59-91), the cross-correlation sums and the Newton search of its
// peak (likelihood.c:195-358), and the position marginalisation (likelihood.c:157-187,
// 376-421).  The model normalisation 1/sqrt(sum M^2/sigma^2) is applied at the end: the
// Newton increments are ratios of sums of equal degree in M, so ho = h'/sqrt(S2) and
// det = det'/S2 with h', det' computed from the unnormalised model.
template <int BS, int N>
__device__ inline void block_allreduce(double (&v)[N], double *lds)
{
#pragma unroll
    for (int j = 0; j < N; j++)
        for (int off = 32; off >= 1; off >>= 1) v[j] += __shfl_xor(v[j], off);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int NW = BS / 64;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; j++) lds[wave * N + j] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; j++) {
        double t = 0.;
        for (int q = 0; q < NW; q++) t += lds[q * N + j];
        v[j] = t;
    }
    __syncthreads();
}

// likelihood.c:174-183: L_r = log of the position-marginalised likelihood minus
// 1/2 log det of the cross-correlation Hessian; -1e10 when the peak search failed.
__device__ inline void finish_lr(const LikLaunch &a, int item, int r, double ho,
                                 double det_sigma, int iter)
{
    double L;
    if (det_sigma <= 0)
        L = kBad;
    else {
        const double B = (ho * ho) * 0.5;
        L = marg_position(B);
        if (isnan(L))
            L = kBad;
        else
            L -= 0.5 * log(det_sigma);
    }
    a.Lr[(size_t)item * kMaxR + r] = L;
    a.nit[(size_t)item * kMaxR + r] = iter;
    if (a.ho) a.ho[(size_t)item * kMaxR + r] = ho;
    if (a.det) a.det[(size_t)item * kMaxR + r] = det_sigma;
}

template <int BS, int PPT>
__global__ __launch_bounds__(BS) void likelihood_kernel(LikLaunch a)
{
    __shared__ double red[(BS / 64) * 8];
    const int nr = a.rg.numr - 1;
    const int item = blockIdx.x / nr;
    const int r = blockIdx.x % nr + 1;
    const LikItem it = a.items[item];
    if (it.prec != 0) return;  // evaluated by the mixed-precision kernel
    const FacetDev &F = a.facets[it.slot];
    const unsigned n = F.n;
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(a.rg.ro[r]);
    const double sf = (scale * scale) / (detA * detA);
    const double cc2 = 4 * kPi * kPi;
    const double sw = sf * cc2;
    const double a1 = 1. + e1, a2 = 1. - e1;
    const double *__restrict__ gu = F.uu;
    const double *__restrict__ gv = F.vv;
    const double *__restrict__ dre = F.dre;
    const double *__restrict__ dim = F.dim;
    const double *__restrict__ gs2 = F.s2;
    const double *__restrict__ gmod = F.mod;
    const int tid = threadIdx.x;

    // Point terms of the first PPT*BS cells stay in registers (static indices so the
    // cache is not demoted to scratch memory); cells beyond are recomputed per pass.
    double hrc[PPT > 0 ? PPT : 1], hic[PPT > 0 ? PPT : 1];
    double s[7] = {0., 0., 0., 0., 0., 0., 0.};  // S2, value, dx, dy, dxx, dyy, dxy
    // (1 + s w |k|^2)^(-3/2) as rsqrt(den)^3: one reciprocal square root instead of a
    // square root and a division (a few ulp from the reference's 1/(den*sqrt(den))).
    auto model_at = [&](unsigned i, double uu, double vv) -> double {
        if (gmod) return gmod[i];
        const double k1 = a1 * uu + e2 * vv;
        const double k2 = e2 * uu + a2 * vv;
        const double rr = rsqrt(1. + sw * (k1 * k1 + k2 * k2));
        return rr * rr * rr;
    };
    auto pass0 = [&](unsigned i, double &hr, double &hi) {
        const double uu = gu[i], vv = gv[i];
        const double M = model_at(i, uu, vv);
        s[0] += M * M / gs2[i];
        hr = dre[i] * M;
        hi = -dim[i] * M;
        const double U = kTwoPi * uu, V = kTwoPi * vv;
        s[1] += hr;
        s[2] -= hi * U;
        s[3] -= hi * V;
        s[4] -= hr * U * U;
        s[5] -= hr * V * V;
        s[6] -= hr * U * V;
    };
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const unsigned i = tid + k * BS;
        hrc[k] = 0.;
        hic[k] = 0.;
        if (i < n) pass0(i, hrc[k], hic[k]);
    }
    for (unsigned i = tid + PPT * BS; i < n; i += BS) {
        double hr, hi;
        pass0(i, hr, hi);
    }
    block_allreduce<BS, 7>(s, red);
    const double S2 = gmod ? 1.0 : s[0];
    double value = s[1], dx = s[2], dy = s[3], dxx = s[4], dyy = s[5], dxy = s[6];
    double det = dxx * dyy - dxy * dxy;
    double xo = 0., yo = 0., norm = 1.;
    int iter = 0;
    while (det > 0 && norm > 1.e-7 && iter < 50) {
        iter++;
        const double incx = (dx * dyy - dy * dxy) / det;
        const double incy = (dy * dxx - dx * dxy) / det;
        xo -= incx;
        yo -= incy;
        double t[6] = {0., 0., 0., 0., 0., 0.};
        auto newton_term = [&](unsigned i, double hr, double hi) {
            const double U = kTwoPi * gu[i], V = kTwoPi * gv[i];
            const double ph = xo * U + yo * V;
            double sn, cs;
            sincos(ph, &sn, &cs);
            const double re = hr * cs - hi * sn;
            const double im = hr * sn + hi * cs;
            t[0] += re;
            t[1] -= im * U;
            t[2] -= im * V;
            t[3] -= re * U * U;
            t[4] -= re * V * V;
            t[5] -= re * U * V;
        };
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const unsigned i = tid + k * BS;
            if (i < n) newton_term(i, hrc[k], hic[k]);
        }
        for (unsigned i = tid + PPT * BS; i < n; i += BS) {
            const double M = model_at(i, gu[i], gv[i]);
            newton_term(i, dre[i] * M, -dim[i] * M);
        }
        block_allreduce<BS, 6>(t, red);
        value = t[0];
        dx = t[1];
        dy = t[2];
        dxx = t[3];
        dyy = t[4];
        dxy = t[5];
        det = dxx * dyy - dxy * dxy;
        norm = sqrt(incx * incx + incy * incy);
    }
    if (tid == 0)
        finish_lr(a, item, r, gmod ? value : value / sqrt(S2), gmod ? det : det / S2, iter);
}

// ------------------------------------------------------------------ likelihood, mixed
// FP32 model and cross-correlation terms from the packed facet {u, v, D_re, D_im}; the
// Newton phase x.u + y.v is formed and reduced in FP64 (revolutions), sin/cos evaluated
// in FP32 on the reduced octant; per-thread partial sums are flushed to FP64 every
// kFlush cells; the Newton update, position and scalelength marginalisation run in FP64.
constexpr int kFlush = 8;

// sin(2 pi t), cos(2 pi t): t reduced to [-1/2, 1/2] in FP64, then to an octant
// |y| <= 1/8 in FP32; Taylor series of sin / cos of 2 pi y to degree 9 / 10
// (truncation < 3e-9 on the octant).
__device__ inline void sincos_rev(double t, float &sn, float &cs)
{
    const float x = (float)(t - rint(t));
    const float q = rintf(4.f * x);
    const float y = x - 0.25f * q;
    const float th = 6.28318530717958648f * y;
    const float z = th * th;
    const float ps = fmaf(fmaf(fmaf(1.f / 362880, z, -1.f / 5040), z, 1.f / 120), z, -1.f / 6);
    const float s = fmaf(th * z, ps, th);
    const float pc = fmaf(fmaf(fmaf(fmaf(-1.f / 3628800, z, 1.f / 40320), z, -1.f / 720), z,
                               1.f / 24), z, -0.5f);
    const float c = fmaf(pc, z, 1.f);
    const int k = ((int)q) & 3;
    sn = (k == 0) ? s : (k == 1) ? c : (k == 2) ? -s : -c;
    cs = (k == 0) ? c : (k == 1) ? -s : (k == 2) ? -c : s;
}

template <int BS, int PPT>
__global__ __launch_bounds__(BS) void likelihood32_kernel(LikLaunch a)
{
    __shared__ double red[(BS / 64) * 8];
    const int nr = a.rg.numr - 1;
    const int item = blockIdx.x / nr;
    const int r = blockIdx.x % nr + 1;
    const LikItem it = a.items[item];
    if (it.prec != 1) return;
    const FacetDev &F = a.facets[it.slot];
    const unsigned n = F.n;
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(a.rg.ro[r]);
    const double sf = (scale * scale) / (detA * detA);
    const float sw = (float)(sf * (4 * kPi * kPi));
    const float a1 = (float)(1. + e1), a2 = (float)(1. - e1), fe2 = (float)e2;
    const float tp = 6.28318530717958648f;
    const float4 *__restrict__ pk = F.pk;
    const float *__restrict__ is2 = F.invs2;
    const int tid = threadIdx.x;

    float hrc[PPT > 0 ? PPT : 1], hic[PPT > 0 ? PPT : 1];
    double s[7] = {0., 0., 0., 0., 0., 0., 0.};
    float f[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int cnt = 0;
    auto model = [&](const float4 p) -> float {
        const float k1 = a1 * p.x + fe2 * p.y;
        const float k2 = fe2 * p.x + a2 * p.y;
        const float rr = rsqrtf(1.f + sw * (k1 * k1 + k2 * k2));
        return rr * rr * rr;
    };
    auto flush7 = [&]() {
#pragma unroll
        for (int j = 0; j < 7; j++) {
            s[j] += (double)f[j];
            f[j] = 0.f;
        }
    };
    auto pass0 = [&](unsigned i, float &hr, float &hi) {
        const float4 p = pk[i];
        const float M = model(p);
        hr = p.z * M;
        hi = -p.w * M;
        const float U = tp * p.x, V = tp * p.y;
        f[0] += M * M * is2[i];
        f[1] += hr;
        f[2] -= hi * U;
        f[3] -= hi * V;
        f[4] -= hr * U * U;
        f[5] -= hr * V * V;
        f[6] -= hr * U * V;
        if (++cnt == kFlush) {
            flush7();
            cnt = 0;
        }
    };
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const unsigned i = tid + k * BS;
        hrc[k] = 0.f;
        hic[k] = 0.f;
        if (i < n) pass0(i, hrc[k], hic[k]);
    }
    for (unsigned i = tid + PPT * BS; i < n; i += BS) {
        float hr, hi;
        pass0(i, hr, hi);
    }
    flush7();
    block_allreduce<BS, 7>(s, red);
    const double S2 = s[0];
    double value = s[1], dx = s[2], dy = s[3], dxx = s[4], dyy = s[5], dxy = s[6];
    double det = dxx * dyy - dxy * dxy;
    double xo = 0., yo = 0., norm = 1.;
    int iter = 0;
    while (det > 0 && norm > 1.e-7 && iter < 50) {
        iter++;
        const double incx = (dx * dyy - dy * dxy) / det;
        const double incy = (dy * dxx - dx * dxy) / det;
        xo -= incx;
        yo -= incy;
        double t[6] = {0., 0., 0., 0., 0., 0.};
        float g[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        cnt = 0;
        auto newton_term = [&](const float4 p, float hr, float hi) {
            float sn, cs;
            sincos_rev(xo * (double)p.x + yo * (double)p.y, sn, cs);
            const float U = tp * p.x, V = tp * p.y;
            const float re = hr * cs - hi * sn;
            const float im = hr * sn + hi * cs;
            g[0] += re;
            g[1] -= im * U;
            g[2] -= im * V;
            g[3] -= re * U * U;
            g[4] -= re * V * V;
            g[5] -= re * U * V;
            if (++cnt == kFlush) {
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    t[j] += (double)g[j];
                    g[j] = 0.f;
                }
                cnt = 0;
            }
        };
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const unsigned i = tid + k * BS;
            if (i < n) newton_term(pk[i], hrc[k], hic[k]);
        }
        for (unsigned i = tid + PPT * BS; i < n; i += BS) {
            const float4 p = pk[i];
            const float M = model(p);
            newton_term(p, p.z * M, -p.w * M);
        }
#pragma unroll
        for (int j = 0; j < 6; j++) t[j] += (double)g[j];
        block_allreduce<BS, 6>(t, red);
        value = t[0];
        dx = t[1];
        dy = t[2];
        dxx = t[3];
        dyy = t[4];
        dxy = t[5];
        det = dxx * dyy - dxy * dxy;
        norm = sqrt(incx * incx + incy * incy);
    }
    if (tid == 0) finish_lr(a, item, r, value / sqrt(S2), det / S2, iter);
}

void launch_likelihood(hipStream_t s, const LikLaunch &a)
{
    if (a.nitems <= 0) return;
    dim3 grid((unsigned)(a.nitems * (a.rg.numr - 1)));
    bool any64 = false, any32 = false;
    for (int i = 0; i < a.nitems; i++) {  // host copy of the request precisions
        if (a.host_prec && a.host_prec[i]) any32 = true; else any64 = true;
    }
    if (any64) {
        switch (a.block) {
        // PPT: register-cached cells per thread; larger caches spill FP64 sincos state.
        case 64: likelihood_kernel<64, 4><<<grid, 64, 0, s>>>(a); break;
        case 128: likelihood_kernel<128, 4><<<grid, 128, 0, s>>>(a); break;
        case 512: likelihood_kernel<512, 0><<<grid, 512, 0, s>>>(a); break;
        case 1024: likelihood_kernel<1024, 0><<<grid, 1024, 0, s>>>(a); break;
        default: likelihood_kernel<256, 4><<<grid, 256, 0, s>>>(a); break;
        }
    }
    if (any32) {
        switch (a.block) {
        case 128: likelihood32_kernel<128, 8><<<grid, 128, 0, s>>>(a); break;
        case 512: likelihood32_kernel<512, 4><<<grid, 512, 0, s>>>(a); break;
        case 1024: likelihood32_kernel<1024, 2><<<grid, 1024, 0, s>>>(a); break;
        default: likelihood32_kernel<256, 8><<<grid, 256, 0, s>>>(a); break;
        }
    }
}

__global__ void pack_facet_kernel(const FacetDev *dfacet, const double *uu, const double *vv,
                                  const double *dre, const double *dim, const double *s2,
                                  float4 *pk, float *invs2, double *invs2_64,
                                  const uint32_t *cid, unsigned facet, double2 *dc,
                                  uint32_t *loff)
{
    const unsigned n = dfacet->n;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        pk[i] = make_float4((float)uu[i], (float)vv[i], (float)dre[i], (float)dim[i]);
        const double w = 1.0 / s2[i];
        invs2[i] = (float)w;
        invs2_64[i] = w;
        pack_cell_records(i, n, dre, dim, cid, facet, dc, loff);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the zero pad cell
        invs2_64[n] = 0.;
        pack_cell_records(n, n, dre, dim, cid, facet, dc, loff);
    }
}

void launch_pack_facet(hipStream_t s, const FacetDev *dfacet, unsigned cap, const double *uu,
                       const double *vv, const double *dre, const double *dim, const double *s2,
                       float4 *pk, float *invs2, double *invs2_64, const uint32_t *cid,
                       unsigned facet, double2 *dc, uint32_t *loff)
{
    unsigned blocks = std::min(2048u, (cap + 255) / 256 + 1);
    pack_facet_kernel<<<blocks, 256, 0, s>>>(dfacet, uu, vv, dre, dim, s2, pk, invs2, invs2_64,
                                             cid, facet, dc, loff);
}

// ------------------------------------------------------------------ r-marginalisation (K4)
// One thread per item: set_posterior_values (marginalise_r.c:29-64), then the spline +
// QAG-GK41 integral of marginalise_posterior_r (marginalise_r.c:81-223).
struct QagWs {
    double *alist, *blist, *rlist, *elist;
    int *order, *level;
};

// QUADPACK's error rescaling (GSL 2.5 err.c).  (200 err / asc)^1.5 is formed as x sqrt(x):
// within 1.5 ulp of the correctly rounded power (sqrt is correctly rounded), an order of
// magnitude fewer instructions than pow on the r-marginalisation's serial chain, and the
// value only feeds comparisons with tolerances.
__device__ inline double rescale_error(double err, double result_abs, double result_asc)
{
    err = fabs(err);
    if (result_asc != 0 && err != 0) {
        const double x = 200 * err / result_asc;
        double scale = x * sqrt(x);
        err = (scale < 1) ? result_asc * scale : result_asc;
    }
    if (result_abs > kDblMin / (50 * kDblEps)) {
        double min_err = 50 * kDblEps * result_abs;
        if (min_err > err) err = min_err;
    }
    return err;
}

__device__ inline void qk41_spline(const Spline &sp, double a, double b, double *result,
                                   double *abserr, double *resabs, double *resasc)
{
    const int n = 21;
    double fv1[21], fv2[21];
    const double center = 0.5 * (a + b);
    const double half_length = 0.5 * (b - a);
    const double abs_half_length = fabs(half_length);
    const double f_center = spline_exp(sp, center);
    double result_gauss = 0;
    double result_kronrod = f_center * kWgk41[n - 1];
    double result_abs = fabs(result_kronrod);
    for (int j = 0; j < (n - 1) / 2; j++) {
        const int jtw = j * 2 + 1;
        const double abscissa = half_length * kXgk41[jtw];
        const double f1 = spline_exp(sp, center - abscissa);
        const double f2 = spline_exp(sp, center + abscissa);
        const double fsum = f1 + f2;
        fv1[jtw] = f1;
        fv2[jtw] = f2;
        result_gauss += kWg41[j] * fsum;
        result_kronrod += kWgk41[jtw] * fsum;
        result_abs += kWgk41[jtw] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < n / 2; j++) {
        const int jtwm1 = j * 2;
        const double abscissa = half_length * kXgk41[jtwm1];
        const double f1 = spline_exp(sp, center - abscissa);
        const double f2 = spline_exp(sp, center + abscissa);
        fv1[jtwm1] = f1;
        fv2[jtwm1] = f2;
        result_kronrod += kWgk41[jtwm1] * (f1 + f2);
        result_abs += kWgk41[jtwm1] * (fabs(f1) + fabs(f2));
    }
    const double mean = result_kronrod * 0.5;
    double result_asc = kWgk41[n - 1] * fabs(f_center - mean);
    for (int j = 0; j < n - 1; j++)
        result_asc += kWgk41[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    const double err = (result_kronrod - result_gauss) * half_length;
    result_kronrod *= half_length;
    result_abs *= abs_half_length;
    result_asc *= abs_half_length;
    *result = result_kronrod;
    *resabs = result_abs;
    *resasc = result_asc;
    *abserr = rescale_error(err, result_abs, result_asc);
}

// QUADPACK qpsrt: keep the error list descending
__device__ inline void qpsrt(QagWs &w, int size, int limit, int &nrmax, int &imax)
{
    const int last = size - 1;
    int i_nrmax = nrmax;
    int i_maxerr = w.order[i_nrmax];
    if (last < 2) {
        w.order[0] = 0;
        w.order[1] = 1;
        imax = i_maxerr;
        return;
    }
    const double errmax = w.elist[i_maxerr];
    while (i_nrmax > 0 && errmax > w.elist[w.order[i_nrmax - 1]]) {
        w.order[i_nrmax] = w.order[i_nrmax - 1];
        i_nrmax--;
    }
    int top = (last < (limit / 2 + 2)) ? last : limit - last + 1;
    int i = i_nrmax + 1;
    while (i < top && errmax < w.elist[w.order[i]]) {
        w.order[i - 1] = w.order[i];
        i++;
    }
    w.order[i - 1] = i_maxerr;
    const double errmin = w.elist[last];
    int k = top - 1;
    while (k > i - 2 && errmin >= w.elist[w.order[k]]) {
        w.order[k + 1] = w.order[k];
        k--;
    }
    w.order[k + 1] = last;
    imax = w.order[i_nrmax];
    nrmax = i_nrmax;
}

// gsl_integration_qag semantics (GSL 2.5 qag.c) with the 41-point rule; 0 = success.
__device__ int qag41_spline(const Spline &sp, double a, double b, double epsrel, int limit,
                            QagWs &w, double *result)
{
    double result0, abserr0, resabs0, resasc0;
    int roundoff_type1 = 0, roundoff_type2 = 0, error_type = 0;
    w.alist[0] = a;
    w.blist[0] = b;
    w.rlist[0] = 0.0;
    w.elist[0] = 0.0;
    w.order[0] = 0;
    w.level[0] = 0;
    int size = 0, nrmax = 0, imax = 0;
    *result = 0;
    qk41_spline(sp, a, b, &result0, &abserr0, &resabs0, &resasc0);
    size = 1;
    w.rlist[0] = result0;
    w.elist[0] = abserr0;
    double tolerance = fmax(0., epsrel * fabs(result0));
    const double round_off = 50 * kDblEps * resabs0;
    if (abserr0 <= round_off && abserr0 > tolerance) {
        *result = result0;
        return 14;
    } else if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) {
        *result = result0;
        return 0;
    } else if (limit == 1) {
        *result = result0;
        return 11;
    }
    double area = result0, errsum = abserr0;
    int iteration = 1;
    do {
        const double a_i = w.alist[imax], b_i = w.blist[imax];
        const double r_i = w.rlist[imax], e_i = w.elist[imax];
        const double a1 = a_i, b1 = 0.5 * (a_i + b_i), a2 = b1, b2 = b_i;
        double area1, area2, error1, error2, resabs1, resabs2, resasc1, resasc2;
        qk41_spline(sp, a1, b1, &area1, &error1, &resabs1, &resasc1);
        qk41_spline(sp, a2, b2, &area2, &error2, &resabs2, &resasc2);
        const double area12 = area1 + area2;
        const double error12 = error1 + error2;
        errsum += (error12 - e_i);
        area += area12 - r_i;
        if (resasc1 != error1 && resasc2 != error2) {
            const double delta = r_i - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i) roundoff_type1++;
            if (iteration >= 10 && error12 > e_i) roundoff_type2++;
        }
        tolerance = fmax(0., epsrel * fabs(area));
        if (errsum > tolerance) {
            if (roundoff_type1 >= 6 || roundoff_type2 >= 20) error_type = 2;
            const double tmp = (1 + 100 * kDblEps) * (fabs(a2) + 1000 * kDblMin);
            if (fabs(a1) <= tmp && fabs(b2) <= tmp) error_type = 3;
        }
        // update(): bisected halves, larger error first in place
        const int i_max = imax, i_new = size;
        const int new_level = w.level[i_max] + 1;
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
        size++;
        qpsrt(w, size, limit, nrmax, imax);
        iteration++;
    } while (iteration < limit && !error_type && errsum > tolerance);
    double sum = 0;
    for (int k = 0; k < size; k++) sum += w.rlist[k];
    *result = sum;
    if (errsum <= tolerance) return 0;
    if (error_type == 2) return 14;
    if (error_type == 3) return 21;
    if (iteration == limit) return 11;
    return 5;
}

// marginalise_r.c:81-223 on (num, x, y); y is modified (y -= ymax).
__device__ double marginalise_posterior(int num, const double *x, double *y, QagWs &w)
{
    double ymax = -10.0;
    int ymaxpos = 0;
    for (int ii = 0; ii < num; ii++)
        if (y[ii] > ymax) {
            ymax = y[ii];
            ymaxpos = ii;
        }
    double xmin = x[ymaxpos], xmax = x[ymaxpos];
    if (ymaxpos < num - 1) xmax = x[ymaxpos + 1];
    if (ymaxpos > 0) xmin = x[ymaxpos - 1];
    for (int ii = 0; ii < num; ii++) {
        if (x[ii] < xmin) xmin = x[ii];
        if (x[ii] > xmax) xmax = x[ii];
        y[ii] -= ymax;
    }
    if (xmax <= xmin) return NAN;
    Spline sp;
    double sump = 0.;
    if (spline_init(sp, x, y, num, 0)) return NAN;
    int st = qag41_spline(sp, xmin, xmax, 1.e-3, kQagLimit, w, &sump);
    if (st) {
        spline_init(sp, x, y, num, 1);
        st = qag41_spline(sp, xmin, xmax, 1.e-3, kQagLimit, w, &sump);
    }
    if (st) return kBad;
    return log(sump) + ymax;
}

// Wave-cooperative variants: one work-group of two waves per item.  Lane j of a wave
// evaluates the integrand at Gauss-Kronrod node j (0..19 left of the centre, 20..39 right,
// 40 the centre) and contributes its weighted terms to the Kronrod, Gauss, |f| and
// |f - mean| sums, which wave_sum completes; every lane ends with the same sums (lane
// 63's, read into scalar registers), so all lanes follow the same control flow.  The sums
// differ from QUADPACK's sequential order only by rounding.  Each bisection integrates its
// two halves on the two waves at once.

// Sum of v over the 64 lanes of a wave, in lane 63, by data-parallel-primitive moves (VALU
// lane shifts, no LDS round trips): shifts by 1, 2, 4 and 8 within each row of 16 lanes
// leave each row's total in its lane 15; row_bcast:15 adds row 0's total into row 1 and
// row 2's into row 3, row_bcast:31 adds lane 31's into rows 2 and 3.  Lanes whose source
// lies outside the row or is masked off add 0.  A fixed order, so the sum is deterministic.
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ inline double dpp_f64(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, BANK_MASK, false);
    const int rhi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, BANK_MASK, false);
    return __hiloint2double(rhi, rlo);
}
__device__ inline double wave_sum_last(double v)
{
    v += dpp_f64<0x111, 0xf, 0xf>(v);  // row_shr:1
    v += dpp_f64<0x112, 0xf, 0xf>(v);  // row_shr:2
    v += dpp_f64<0x114, 0xf, 0xf>(v);  // row_shr:4
    v += dpp_f64<0x118, 0xf, 0xf>(v);  // row_shr:8
    v += dpp_f64<0x142, 0xa, 0xf>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc, 0xf>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// The wave's sum in every lane: wave_sum_last, then lane 63's value read into scalar
// registers (every lane gets the same value, so the wave takes the same branches after it).
__device__ inline double wave_sum(double v)
{
    v = wave_sum_last(v);
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}

// Lane values read into scalar registers (l uniform over the wave).
__device__ inline double read_lane(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// The wave's maximum (MAX) or minimum in every lane, by the lane shifts of wave_sum_last; a
// lane whose source lies outside its row, or in a row masked off, keeps its own value, which
// leaves the extreme unchanged.  fmax and fmin are exact (and skip NaN), so the result does
// not depend on the order.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_keep_f64(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
    const int rhi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
    return __hiloint2double(rhi, rlo);
}
template <bool MAX>
__device__ inline double wave_extreme(double v)
{
    auto op = [](double p, double q) { return MAX ? fmax(p, q) : fmin(p, q); };
    v = op(v, dpp_keep_f64<0x111, 0xf>(v));  // row_shr:1
    v = op(v, dpp_keep_f64<0x112, 0xf>(v));  // row_shr:2
    v = op(v, dpp_keep_f64<0x114, 0xf>(v));  // row_shr:4
    v = op(v, dpp_keep_f64<0x118, 0xf>(v));  // row_shr:8
    v = op(v, dpp_keep_f64<0x142, 0xa>(v));  // row_bcast:15 into rows 1 and 3
    v = op(v, dpp_keep_f64<0x143, 0xc>(v));  // row_bcast:31 into rows 2 and 3
    return read_lane(v, 63);
}

// The spline of the r-marginalisation in LDS with the cubic coefficients b_i and d_i of each
// interval (cspline_eval's), formed once per spline: spline_exp_lds then takes no division
// per evaluation and, as b_i and d_i are the same expressions of the same knot data, gives
// spline_exp's values bit for bit.
struct SplineLds {
    Spline s;
    double b[kMaxR], d[kMaxR];
};

// The interval is found by counting the inner knots at or below x: with strictly increasing
// knots (spline_build_wave checks them) that is the index cspline_eval's bisection finds, and
// the reads are independent of each other rather than a chain of dependent LDS round trips.
__device__ inline double spline_exp_lds(const SplineLds &t, double x)
{
    const Spline &s = t.s;
    const int n = s.n;
    if (x < s.x[0] || x > s.x[n - 1]) return 0.;
    int lo = 0;
#pragma unroll 8
    for (int j = 1; j < n - 1; j++) lo += s.x[j] <= x ? 1 : 0;
    const double x_lo = s.x[lo], x_hi = s.x[lo + 1];
    const double dx = x_hi - x_lo;
    if (!(dx > 0.0)) return 0.;
    const double y_lo = s.y[lo];
    double y;
    if (s.linear) {
        y = y_lo + (x - x_lo) / dx * (s.y[lo + 1] - y_lo);
    } else {
        const double delx = x - x_lo;
        y = y_lo + delx * (t.b[lo] + delx * (s.c[lo] + delx * t.d[lo]));
    }
    return exp(y);
}

// spline_init_ws (GSL cspline_init semantics) by each wave: lane p holds knot (xp, yp) for
// p < n, the knots already in t.s.x and t.s.y.  The rows of the tridiagonal system are formed
// by their lanes at once; the LDL^T recurrences run over lane values read into scalar
// registers, in spline_init_ws's order and with its expressions, so c, and the b_i and d_i
// formed from it, are the same bit for bit.  Wave 0 writes n, linear = 0, c, b and d.
// Returns spline_init_ws's status, the same in every lane.
__device__ int spline_build_wave(SplineLds &t, int n, double xp, double yp)
{
    const int lane = threadIdx.x & 63;
    const bool w0 = threadIdx.x < 64;
    const double x1 = __shfl_down(xp, 1), x2 = __shfl_down(xp, 2);
    const double y1 = __shfl_down(yp, 1), y2 = __shfl_down(yp, 2);
    if (__any(lane + 1 < n && !(xp < x1))) return 1;  // knots strictly increasing
    const int max_index = n - 1, sys = max_index - 1;
    double off = 0., diag = 0., g = 0.;
    if (lane < sys) {
        const double h_i = x1 - xp;
        const double h_ip1 = x2 - x1;
        const double yd_i = y1 - yp;
        const double yd_ip1 = y2 - y1;
        const double g_i = (h_i != 0.0) ? 1.0 / h_i : 0.0;
        const double g_ip1 = (h_ip1 != 0.0) ? 1.0 / h_ip1 : 0.0;
        off = h_ip1;
        diag = 2.0 * (h_ip1 + h_i);
        g = 3.0 * (yd_ip1 * g_ip1 - yd_i * g_i);
    }
    double c = 0.;  // lane p: c_p, with c_0 = c_{n-1} = 0
    int status = 0;
    if (sys == 1) {
        const double c1 = read_lane(g, 0) / read_lane(diag, 0);
        if (lane == 1) c = c1;
    } else if (sys > 1) {
        // step i: alp_i from gam_{i-1}, the forward substitution's z_i, then gam_i =
        // off_i / alp_i, whose division is the chain's latency; the next step's lane values
        // are read before it
        double alp_l = 0., gam_l = 0., z_l = 0.;  // lane i: alp_i, gam_i, z_i
        double al = read_lane(diag, 0);
        double of = read_lane(off, 0);
        double z = read_lane(g, 0);
        double gm = of / al;
        if (al == 0) status = 1;
        if (lane == 0) {
            alp_l = al;
            gam_l = gm;
            z_l = z;
        }
        double dg = read_lane(diag, 1), gg = read_lane(g, 1), of_next = read_lane(off, 1);
        for (int i = 1; i < sys - 1; i++) {
            al = dg - of * gm;
            z = gg - gm * z;
            of = of_next;
            dg = read_lane(diag, i + 1);
            gg = read_lane(g, i + 1);
            of_next = read_lane(off, i + 1);
            gm = of / al;
            if (al == 0) status = 1;
            if (lane == i) {
                alp_l = al;
                gam_l = gm;
                z_l = z;
            }
        }
        al = dg - of * gm;  // alp_{sys-1}: no gam, and no status check, in spline_init_ws
        z = gg - gm * z;
        if (lane == sys - 1) {
            alp_l = al;
            z_l = z;
        }
        const double cc = lane < sys ? z_l / alp_l : 0.;
        double xs = read_lane(cc, sys - 1);
        if (lane == sys) c = xs;
        for (int i = sys - 2; i >= 0; i--) {
            xs = read_lane(cc, i) - read_lane(gam_l, i) * xs;
            if (lane == i + 1) c = xs;
        }
    }
    const double c1 = __shfl_down(c, 1);
    if (w0) {
        if (lane == 0) {
            t.s.n = n;
            t.s.linear = 0;
        }
        if (lane < n) t.s.c[lane] = c;
        if (lane + 1 < n) {
            const double dx = x1 - xp;
            const double dy = y1 - yp;
            t.b[lane] = (dy / dx) - dx * (c1 + 2.0 * c) / 3.0;
            t.d[lane] = (c1 - c) / (3.0 * dx);
        }
    }
    return status;
}

// Lane j's Gauss-Kronrod-41 node (0..19 left of the centre, 20..39 right, 40 the centre) and
// its Kronrod and Gauss weights, read once per work-group.
struct Gk41Node {
    double x, wk, wg;
    int side;  // -1 left, 1 right, 0 the centre (and the unused lanes 41..63)
};

__device__ inline Gk41Node gk41_node()
{
    const int lane = threadIdx.x & 63;
    const int k = lane < 20 ? lane : lane - 20;
    Gk41Node g{0., 0., 0., 0};
    if (lane < 41) {
        g.side = lane < 20 ? -1 : (lane < 40 ? 1 : 0);
        g.x = lane < 40 ? kXgk41[k] : 0.;
        g.wk = lane < 40 ? kWgk41[k] : kWgk41[20];
        g.wg = (lane < 40 && (k & 1)) ? kWg41[k >> 1] : 0.;
    }
    return g;
}

__device__ inline void qk41_wave(const SplineLds &sp, const Gk41Node &g, double a, double b,
                                 double *result, double *abserr, double *resabs,
                                 double *resasc)
{
    const int lane = threadIdx.x & 63;
    const double center = 0.5 * (a + b);
    const double half_length = 0.5 * (b - a);
    const double abs_half_length = fabs(half_length);
    double f = 0.;
    if (lane < 41) {
        double x;
        if (g.side < 0)
            x = center - half_length * g.x;
        else if (g.side > 0)
            x = center + half_length * g.x;
        else
            x = center;
        f = spline_exp_lds(sp, x);
    }
    double result_kronrod = wave_sum(g.wk * f);
    const double result_gauss = wave_sum(g.wg * f);
    double result_abs = wave_sum(g.wk * fabs(f));
    const double mean = result_kronrod * 0.5;
    double result_asc = wave_sum(lane < 41 ? g.wk * fabs(f - mean) : 0.);
    const double err = (result_kronrod - result_gauss) * half_length;
    result_kronrod *= half_length;
    result_abs *= abs_half_length;
    result_asc *= abs_half_length;
    *result = result_kronrod;
    *resabs = result_abs;
    *resasc = result_asc;
    *abserr = rescale_error(err, result_abs, result_asc);
}

// qag41_spline's QUADPACK sequence by the work-group.  Each bisection takes two barriers: one
// after the halves' results are exchanged, one after lane 0 has updated the lists (qpsrt
// included) and published the next interval; every thread reads the exchange before the
// second and the lists after it, so no write can overtake a read of the previous bisection.
__device__ int qag41_wave(const SplineLds &sp, const Gk41Node &g, double a, double b,
                          double epsrel, int limit, QagWs &w, double *result)
{
    double result0, abserr0, resabs0, resasc0;
    int roundoff_type1 = 0, roundoff_type2 = 0, error_type = 0;
    const bool writer = threadIdx.x == 0;
    if (writer) {
        w.alist[0] = a;
        w.blist[0] = b;
        w.rlist[0] = 0.0;
        w.elist[0] = 0.0;
        w.order[0] = 0;
        w.level[0] = 0;
    }
    int size = 0, nrmax = 0, imax = 0;
    *result = 0;
    qk41_wave(sp, g, a, b, &result0, &abserr0, &resabs0, &resasc0);
    size = 1;
    if (writer) {
        w.rlist[0] = result0;
        w.elist[0] = abserr0;
    }
    double tolerance = fmax(0., epsrel * fabs(result0));
    const double round_off = 50 * kDblEps * resabs0;
    if (abserr0 <= round_off && abserr0 > tolerance) {
        *result = result0;
        return 14;
    } else if ((abserr0 <= tolerance && abserr0 != resasc0) || abserr0 == 0.0) {
        *result = result0;
        return 0;
    } else if (limit == 1) {
        *result = result0;
        return 11;
    }
    double area = result0, errsum = abserr0;
    int iteration = 1;
    __shared__ double xch[8];
    __shared__ int sh_imax, sh_nrmax;
    __syncthreads();
    do {
        const double a_i = w.alist[imax], b_i = w.blist[imax];
        const double r_i = w.rlist[imax], e_i = w.elist[imax];
        const int lev = w.level[imax];
        const double a1 = a_i, b1 = 0.5 * (a_i + b_i), a2 = b1, b2 = b_i;
        double area1, area2, error1, error2, resasc1, resasc2;  // |f| integrals unused by qag
        // the two halves at once: wave 0 integrates [a1, b1], wave 1 [a2, b2]
        {
            const int wv = threadIdx.x >> 6;
            double r, e, ra, rc;
            qk41_wave(sp, g, wv ? a2 : a1, wv ? b2 : b1, &r, &e, &ra, &rc);
            if ((threadIdx.x & 63) == 0) {
                xch[4 * wv + 0] = r;
                xch[4 * wv + 1] = e;
                xch[4 * wv + 2] = ra;
                xch[4 * wv + 3] = rc;
            }
        }
        __syncthreads();
        area1 = xch[0];
        error1 = xch[1];
        resasc1 = xch[3];
        area2 = xch[4];
        error2 = xch[5];
        resasc2 = xch[7];
        const double area12 = area1 + area2;
        const double error12 = error1 + error2;
        errsum += (error12 - e_i);
        area += area12 - r_i;
        if (resasc1 != error1 && resasc2 != error2) {
            const double delta = r_i - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && error12 >= 0.99 * e_i) roundoff_type1++;
            if (iteration >= 10 && error12 > e_i) roundoff_type2++;
        }
        tolerance = fmax(0., epsrel * fabs(area));
        if (errsum > tolerance) {
            if (roundoff_type1 >= 6 || roundoff_type2 >= 20) error_type = 2;
            const double tmp = (1 + 100 * kDblEps) * (fabs(a2) + 1000 * kDblMin);
            if (fabs(a1) <= tmp && fabs(b2) <= tmp) error_type = 3;
        }
        size++;
        if (writer) {
            const int i_max = imax, i_new = size - 1;
            const int new_level = lev + 1;
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
            qpsrt(w, size, limit, nrmax, imax);
            sh_imax = imax;
            sh_nrmax = nrmax;
        }
        __syncthreads();
        imax = sh_imax;
        nrmax = sh_nrmax;
        iteration++;
    } while (iteration < limit && !error_type && errsum > tolerance);
    double sum = 0;
    for (int k = 0; k < size; k++) sum += w.rlist[k];
    *result = sum;
    if (errsum <= tolerance) return 0;
    if (error_type == 2) return 14;
    if (error_type == 3) return 21;
    if (iteration == limit) return 11;
    return 5;
}

// likelihood.c:126-146 with marginalise_posterior_r (marginalise_r.c:81-223) for one item, by
// the two waves of the work-group.  Lane r (1 <= r < numR) of each wave reads L_r[r]; the
// points above kBad go, in r order, to knots 0..num-1 in LDS; ymax, xmin and xmax are wave
// reductions (xmin and xmax are the extreme abscissae: the reference's start from the
// neighbours of the maximum does not change them); the QUADPACK lists live in LDS.
__device__ void rmarg_item(const LikLaunch &a, int item, SplineLds &sp, double *qd, int *qi)
{
    const int lane = threadIdx.x & 63;
    const int numR = a.rg.numr;
    const bool mine = lane >= 1 && lane < numR;
    // every global read of the item is issued here, the node tables first
    const Gk41Node g = gk41_node();
    const LikItem it = a.items[item];
    const size_t base = (size_t)item * kMaxR;
    const double lv = mine ? a.Lr[base + lane] : kBad;
    bool act = false, frag = false;
    if (a.tasks && mine) {
        act = a.tasks[base + lane].active;
        frag = a.tasks[base + lane].fragile;
    }
    if (__any(act)) return;  // tiled engine: an item with a task still iterating waits
    // the node tables have arrived with the loads above: keep them from being re-read later
    asm volatile("" ::"v"(g.x), "v"(g.wk), "v"(g.wg));
    const bool fragile = it.prec == 1 && __any(frag);
    const bool ok = mine && lv > kBad;
    const double xr = mine ? a.rg.ro[lane] : 0.;
    const double yr = ok ? lv + a.facets[it.slot].rprior[lane] : 0.;
    const uint64_t m = __ballot(ok);
    const int num = __popcll(m);
    int error = 0;
    double L_e;
    if (num > 4) {
        const double ymax = wave_extreme<true>(ok && yr > -10.0 ? yr : -10.0);
        const double xmin = wave_extreme<false>(ok ? xr : INFINITY);
        const double xmax = wave_extreme<true>(ok ? xr : -INFINITY);
        const int pos = __popcll(m & ((1ull << lane) - 1ull));
        if (threadIdx.x < 64 && ok) {
            sp.s.x[pos] = xr;
            sp.s.y[pos] = yr - ymax;
        }
        __syncthreads();
        const double xp = lane < num ? sp.s.x[lane] : 0.;
        const double yp = lane < num ? sp.s.y[lane] : 0.;
        if (xmax <= xmin || spline_build_wave(sp, num, xp, yp)) {
            L_e = NAN;
        } else {
            QagWs w;
            w.alist = qd;
            w.blist = qd + kQagLimit;
            w.rlist = qd + 2 * kQagLimit;
            w.elist = qd + 3 * kQagLimit;
            w.order = qi;
            w.level = qi + kQagLimit;
            __syncthreads();
            double sump = 0.;
            int st = qag41_wave(sp, g, xmin, xmax, 1.e-3, kQagLimit, w, &sump);
            if (st) {  // linear interpolation (marginalise_r.c: the retry after a failure)
                __syncthreads();
                if (threadIdx.x == 0) sp.s.linear = 1;
                __syncthreads();
                st = qag41_wave(sp, g, xmin, xmax, 1.e-3, kQagLimit, w, &sump);
            }
            L_e = st ? kBad : log(sump) + ymax;
        }
    } else {
        error = 1;
        L_e = kBad;
    }
    if (L_e == kBad) error = 1;
    if (fragile) error |= kErrFragile;
    if (threadIdx.x == 0) {
        a.Le[item] = L_e;
        a.err[item] = error;
    }
}

__device__ void tile_status(const LikLaunch &a);

// The spline and the QUADPACK interval lists live in LDS (41 KB): lane 0 updates the lists
// and every lane reads them back after each bisection, a round trip through L2 when they
// were in global memory.
__global__ __launch_bounds__(128) void rmarg_kernel(LikLaunch a)
{
    if (blockIdx.x == (unsigned)a.nitems) {  // the status group (tiled engine)
        tile_status(a);
        return;
    }
    __shared__ SplineLds sp;
    __shared__ double qd[kQagDoubles];
    __shared__ int qi[kQagInts];
    rmarg_item(a, blockIdx.x, sp, qd, qi);
}

// ------------------------------------------------------------------ likelihood, tiled
// Largest item whose first work-group is <= wg (first work-groups are non-decreasing and
// every item has at least one work-group).
__device__ inline int tile_item(const LikItem *items, int nitems, unsigned wg, int nr_g)
{
    int lo = 0, hi = nitems;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (items[mid].chunk_first * (unsigned)nr_g <= wg)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Work-groups [first, first + nr_g * chunks) of `item`, chunk-major: the nr_g scalelength
// tasks of one chunk are consecutive, so they read the same cells close together in time.
__device__ inline void tile_range(const TileLaunch &t, const LikItem *items, int item,
                                  unsigned *first, unsigned *nch)
{
    if (t.item_chunks) {  // every item of the launch has item_chunks chunks
        *first = (unsigned)item * t.item_chunks * (unsigned)t.nr_g;
        *nch = t.item_chunks;
        return;
    }
    const unsigned c0 = items[item].chunk_first;
    const unsigned c1 = (item + 1 < t.lk.nitems) ? items[item + 1].chunk_first
                                                  : t.nwg / (unsigned)t.nr_g;
    *first = c0 * (unsigned)t.nr_g;
    *nch = c1 - c0;
}

// The item, first work-group and chunk count of logical work-group wg: a division when every
// item of the launch has the same chunk count (one galaxy's points: the serial loop), else
// the search over the items' first chunks.
__device__ inline void tile_locate(const TileLaunch &t, const LikItem *items, unsigned wg,
                                   int *item, unsigned *first, unsigned *nch)
{
    *item = t.item_chunks ? (int)(wg / (t.item_chunks * (unsigned)t.nr_g))
                          : tile_item(items, t.lk.nitems, wg, t.nr_g);
    tile_range(t, items, *item, first, nch);
}

// Work-groups are dispatched to the 8 XCDs round-robin (block b to XCD b % 8) and each XCD
// has its own L2.  The blocks of XCD x walk the logical work-groups x * per ... (x + 1) *
// per - 1 (per = ceil(nwg / 8)), so the tasks of a chunk run on one XCD and share its L2.
// Blocks of a pass: one per logical work-group, or at most `cap` (a multiple of 8).
inline unsigned tile_grid(unsigned nwg, unsigned cap)
{
    const unsigned g = (nwg + kXcds - 1) / kXcds * kXcds;
    return (cap && g > cap) ? cap : g;
}

// Loads through the global address space: the facet arrays are reached through pointers
// read from memory, which the compiler would otherwise address as flat (generic) memory,
// and a flat load also holds the LDS counter every wait has to drain.
#define RL2_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ inline T gload(const T *p, size_t i)
{
    return ((const RL2_GLOBAL T *)p)[i];
}
// HIP's vector structs through the matching builtin vector types (one 16-byte load)
typedef double rl2_d2 __attribute__((ext_vector_type(2)));
typedef float rl2_f4 __attribute__((ext_vector_type(4)));
__device__ inline double2 gload(const double2 *p, size_t i)
{
    const rl2_d2 v = ((const RL2_GLOBAL rl2_d2 *)p)[i];
    return make_double2(v.x, v.y);
}
__device__ inline float4 gload(const float4 *p, size_t i)
{
    const rl2_f4 v = ((const RL2_GLOBAL rl2_f4 *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

// Cells per thread whose loads are issued together before any of them is used: a thread
// walks its cells beg + tid, beg + tid + 256, ... in batches of kTileU, so each memory
// round trip serves kTileU cells.  Lanes past the chunk end load cell `beg` with zero data:
// their terms are exactly zero, and every thread adds its cells in increasing order as
// before.
constexpr int kTileU = 4;
// Cells per thread and batch of the paths with two batches in flight (the Newton passes'
// record path, pass 0 on extracted facets; 3 and 4 measured slower).
constexpr int kRecU = 2;

// Chunk sums in FP64.  PASS 0: {S2, value, dx, dy, dxx, dyy, dxy} of the unnormalised
// model at zero offset (likelihood.c:214-236); PASS 1: {value, dx, dy, dxx, dyy, dxy} at
// the offset (xo, yo) (likelihood.c:296-327).
template <int PASS>
__device__ inline void tile_sums64(const FacetDev &F, const LikItem &it, double ro, double xo,
                                   double yo, unsigned beg, unsigned end, double (&s)[7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(ro);
    const double sf = (scale * scale) / (detA * detA);
    const double sw = sf * (4 * kPi * kPi);
    const double a1 = 1. + e1, a2 = 1. - e1;
#pragma unroll
    for (int j = 0; j < 7; j++) s[j] = 0.;
    for (unsigned base = beg + threadIdx.x; base < end; base += kTileU * kTileBS) {
        double uu[kTileU], vv[kTileU], dr[kTileU], di[kTileU], s2[kTileU];
#pragma unroll
        for (int q = 0; q < kTileU; q++) {
            const unsigned i = base + q * kTileBS;
            const bool in = i < end;
            const unsigned k = in ? i : beg;
            uu[q] = gload(F.uu, k);
            vv[q] = gload(F.vv, k);
            const double r_ = gload(F.dre, k), m_ = gload(F.dim, k);
            dr[q] = in ? r_ : 0.;
            di[q] = in ? m_ : 0.;
            s2[q] = PASS == 0 ? gload(F.s2, k) : 1.;
        }
#pragma unroll
        for (int q = 0; q < kTileU; q++) {
            const double k1 = fma(e2, vv[q], a1 * uu[q]);
            const double k2 = fma(a2, vv[q], e2 * uu[q]);
            const double rr = rsqrt_ge1(fma(sw, fma(k1, k1, k2 * k2), 1.));
            const double M = rr * rr * rr;
            const double hr = dr[q] * M;
            const double hi = -di[q] * M;
            const double U = kTwoPi * uu[q], V = kTwoPi * vv[q];
            if (PASS == 0) {
                if (base + q * kTileBS < end) s[0] += M * M / s2[q];
                s[1] += hr;
                s[2] = fma(-hi, U, s[2]);
                s[3] = fma(-hi, V, s[3]);
                const double hU = hr * U;
                s[4] = fma(-hU, U, s[4]);
                s[5] = fma(-hr * V, V, s[5]);
                s[6] = fma(-hU, V, s[6]);
            } else {
                // phase xo U + yo V = 2 pi (xo u + yo v): evaluated in revolutions
                double sn, cs;
                sincos_rev64(fma(xo, uu[q], yo * vv[q]), sn, cs);
                const double re = fma(hr, cs, -hi * sn);
                const double im = fma(hr, sn, hi * cs);
                s[0] += re;
                s[1] = fma(-im, U, s[1]);
                s[2] = fma(-im, V, s[2]);
                const double rU = re * U;
                s[3] = fma(-rU, U, s[3]);
                s[4] = fma(-re * V, V, s[4]);
                s[5] = fma(-rU, V, s[5]);
            }
        }
    }
}

// PASS 0 sums of tile_sums64 on a facet with grid structure: the cell centres from the
// LDS copy of ctr (sC) by cell id, so a cell costs 28 bytes of vector memory instead of 40,
// and M^2 / sigma^2 as M^2 times the packed 1 / sigma^2 (one rounding apart from the
// quotient; the division sequence was a fifth of the pass's instructions).
template <int U>
__device__ inline void tile_sums64_p0_grid(const FacetDev &F, const LikItem &it, double ro,
                                           const double *sC, unsigned beg, unsigned end,
                                           double (&s)[7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(ro);
    const double sf = (scale * scale) / (detA * detA);
    const double sw = sf * (4 * kPi * kPi);
    const double a1 = 1. + e1, a2 = 1. - e1;
#pragma unroll
    for (int j = 0; j < 7; j++) s[j] = 0.;
    for (unsigned base = beg + threadIdx.x; base < end; base += U * kTileBS) {
        uint32_t c[U];
        double dr[U], di[U], s2[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const unsigned i = base + q * kTileBS;
            const unsigned k = i < end ? i : F.n;  // past the end: the zero pad cell
            c[q] = gload(F.cid, k);
            dr[q] = gload(F.dre, k);
            di[q] = gload(F.dim, k);
            s2[q] = gload(F.invs2_64, k);  // 1 / sigma^2
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            const double uu = sC[c[q] & 0xffffu], vv = sC[c[q] >> 16];
            const double k1 = fma(e2, vv, a1 * uu);
            const double k2 = fma(a2, vv, e2 * uu);
            const double rr = rsqrt_ge1(fma(sw, fma(k1, k1, k2 * k2), 1.));
            const double M = rr * rr * rr;
            const double hr = dr[q] * M;
            const double hi = -di[q] * M;
            // sums over u, v: the 2 pi and 4 pi^2 factors are applied to the totals below
            s[0] = fma(M * M, s2[q], s[0]);
            s[1] += hr;
            s[2] = fma(-hi, uu, s[2]);
            s[3] = fma(-hi, vv, s[3]);
            const double hU = hr * uu;
            s[4] = fma(-hU, uu, s[4]);
            s[5] = fma(-hr * vv, vv, s[5]);
            s[6] = fma(-hU, vv, s[6]);
        }
    }
    s[2] *= kTwoPi;
    s[3] *= kTwoPi;
    constexpr double kTwoPi2 = kTwoPi * kTwoPi;
    s[4] *= kTwoPi2;
    s[5] *= kTwoPi2;
    s[6] *= kTwoPi2;
}

// tile_sums64_p0_grid's sums with two batches of kRecU cells per thread in flight: the next
// batch's loads are issued before the current batch is summed, the visibility pair comes from
// the packed copy dc, and `stage` (the centres into sC, ending with a barrier) runs after the
// first batch's loads are issued.  A thread's cells, and so its sums, are those of
// tile_sums64_p0_grid in the same order (past the end: the zero pad cell, which adds zeros).
// NT scalelength tasks of the same item and chunk at once (ro[t], s[t]): the cell's loads,
// centres and the shape terms k1, k2 and k1^2 + k2^2 are shared, the rest is each task's own,
// with the expressions of the one-task form, so every task's sums are the same bit for bit.
template <int NT, class Stage>
__device__ inline void tile_sums64_p0_db(const FacetDev &F, const LikItem &it,
                                         const double (&ro)[NT], const double *sC,
                                         const Stage &stage, unsigned beg, unsigned end,
                                         double (&s)[NT][7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    double sw[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const double scale = arcs2rad_scale(ro[t]);
        const double sf = (scale * scale) / (detA * detA);
        sw[t] = sf * (4 * kPi * kPi);
    }
    const double a1 = 1. + e1, a2 = 1. - e1;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int j = 0; j < 7; j++) s[t][j] = 0.;
    const uint32_t *cid = F.cid;
    const double2 *dc = F.dc;
    const double *w2 = F.invs2_64;
    const unsigned pad = F.n;
    struct Cells {
        uint32_t c[kRecU];
        double dr[kRecU], di[kRecU], s2[kRecU];
    };
    auto fetch = [&](Cells &b, unsigned base) {
#pragma unroll
        for (int q = 0; q < kRecU; q++) {
            const unsigned i = base + q * kTileBS;
            const unsigned k = i < end ? i : pad;
            b.c[q] = gload(cid, k);
            const double2 d = gload(dc, k);
            b.dr[q] = d.x;
            b.di[q] = d.y;
            b.s2[q] = gload(w2, k);
        }
    };
    auto eval = [&](const Cells &b) {
#pragma unroll
        for (int q = 0; q < kRecU; q++) {
            const double uu = sC[b.c[q] & 0xffffu], vv = sC[b.c[q] >> 16];
            const double k1 = fma(e2, vv, a1 * uu);
            const double k2 = fma(a2, vv, e2 * uu);
            const double kk = fma(k1, k1, k2 * k2);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const double rr = rsqrt_ge1(fma(sw[t], kk, 1.));
                const double M = rr * rr * rr;
                const double hr = b.dr[q] * M;
                const double hi = -b.di[q] * M;
                s[t][0] = fma(M * M, b.s2[q], s[t][0]);
                s[t][1] += hr;
                s[t][2] = fma(-hi, uu, s[t][2]);
                s[t][3] = fma(-hi, vv, s[t][3]);
                const double hU = hr * uu;
                s[t][4] = fma(-hU, uu, s[t][4]);
                s[t][5] = fma(-hr * vv, vv, s[t][5]);
                s[t][6] = fma(-hU, vv, s[t][6]);
            }
        }
    };
    constexpr unsigned S = kRecU * kTileBS;
    Cells A, B;
    fetch(A, beg + threadIdx.x);
    stage();
#pragma unroll
    for (int q = 0; q < kRecU; q++)
        asm volatile("" : "+v"(A.c[q]), "+v"(A.dr[q]), "+v"(A.di[q]), "+v"(A.s2[q]));
    for (unsigned base = beg + threadIdx.x; base < end; base += 2 * S) {
        fetch(B, base + S);
        __builtin_amdgcn_sched_barrier(0);
        eval(A);
        if (base + S >= end) break;
        fetch(A, base + 2 * S);
        __builtin_amdgcn_sched_barrier(0);
        eval(B);
    }
    constexpr double kTwoPi2 = kTwoPi * kTwoPi;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        s[t][2] *= kTwoPi;
        s[t][3] *= kTwoPi;
        s[t][4] *= kTwoPi2;
        s[t][5] *= kTwoPi2;
        s[t][6] *= kTwoPi2;
    }
}

// PASS 1 sums of tile_sums64 on a facet with grid structure.  The Newton phase is
// separable on the grid, 2 pi (xo u + yo v) with u = ctr[column], v = ctr[row], so
// e^{i phase} is the product of the task's column entry X[column] = e^{2 pi i xo u} and row
// entry Y[row] = e^{2 pi i yo v}, computed once per pass (lik_step_kernel) instead of a
// sin/cos per cell.  The cell centres come from ctr, bit-identical to uu / vv.  The work-group
// copies X, ctr and (YL) Y to LDS first (sX, sC, sY): per cell only the cell id and data then
// go through the vector memory pipeline, which otherwise carried 68 bytes per cell and task.
// Without YL the row entry is a global load, nearly uniform across a wave.
// One batch of cells of the table path: cell ids and data.
constexpr int kTabU = 2;
struct TabCells {
    uint32_t c[kTabU];
    double dr[kTabU], di[kTabU];
};

// `stage` fills sX, sC and (YL) sY and ends with a barrier; it runs after the first cells'
// loads are issued, so their round trip overlaps it.  It returns false (the same in every
// thread) when the work-group has nothing to sum; the function then returns false too.
template <bool YL, class Stage>
__device__ inline bool tile_sums64_tab(const FacetDev &F, const LikItem &it, double ro,
                                       const Stage &stage, const double2 *Yg, double2 *sX,
                                       double2 *sY, double *sC, unsigned beg, unsigned end,
                                       double (&s)[7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(ro);
    const double sf = (scale * scale) / (detA * detA);
    const double sw = sf * (4 * kPi * kPi);
    const double a1 = 1. + e1, a2 = 1. - e1;
#pragma unroll
    for (int j = 0; j < 7; j++) s[j] = 0.;
    // The facet fields in registers (the scheduling barriers below would otherwise make the
    // compiler reload them from the descriptor every iteration).  Lanes past the end of the
    // chunk read the zero pad cell after the facet's last one, so the loads are
    // unconditional and no lane masks its data.
    const uint32_t *cid = F.cid;
    const double *dre = F.dre, *dim = F.dim;
    const unsigned pad = F.n;
    auto fetch = [&](TabCells &b, unsigned base) {
#pragma unroll
        for (int q = 0; q < kTabU; q++) {
            const unsigned i = base + q * kTileBS;
            const unsigned k = i < end ? i : pad;
            b.c[q] = gload(cid, k);
            b.dr[q] = gload(dre, k);
            b.di[q] = gload(dim, k);
        }
    };
    auto eval = [&](const TabCells &b) {
        double uu[kTabU], vv[kTabU];
        double2 x[kTabU], y[kTabU];
#pragma unroll
        for (int q = 0; q < kTabU; q++) {
            const unsigned ci = b.c[q] & 0xffffu, cj = b.c[q] >> 16;
            uu[q] = sC[ci];
            vv[q] = sC[cj];
            x[q] = sX[ci];
            y[q] = YL ? sY[cj] : gload(Yg, cj);
        }
#pragma unroll
        for (int q = 0; q < kTabU; q++) {
            const double k1 = fma(e2, vv[q], a1 * uu[q]);
            const double k2 = fma(a2, vv[q], e2 * uu[q]);
            const double rr = rsqrt_ge1(fma(sw, fma(k1, k1, k2 * k2), 1.));
            const double M = rr * rr * rr;
            const double hr = b.dr[q] * M;
            const double hi = -b.di[q] * M;
            const double cs = fma(x[q].x, y[q].x, -(x[q].y * y[q].y));
            const double sn = fma(x[q].y, y[q].x, x[q].x * y[q].y);
            const double re = fma(hr, cs, -hi * sn);
            const double im = fma(hr, sn, hi * cs);
            // the gradient and Hessian sums over u, v; their 2 pi and 4 pi^2 factors are
            // applied once to the thread's totals instead of to every cell's U = 2 pi u
            s[0] += re;
            s[1] = fma(-im, uu[q], s[1]);
            s[2] = fma(-im, vv[q], s[2]);
            const double rU = re * uu[q];
            s[3] = fma(-rU, uu[q], s[3]);
            s[4] = fma(-re * vv[q], vv[q], s[4]);
            s[5] = fma(-rU, vv[q], s[5]);
        }
    };
    // Software-pipelined over two register sets: the loads of the next batch are issued
    // before the current batch is evaluated, so their latency overlaps its arithmetic.  A
    // set is loaded only after its previous batch has been used, so the loop needs no
    // register copies (a copy of a just-issued load waits for it at once), and the
    // scheduling barriers keep the loads above the arithmetic.  The first batch passes
    // through an empty asm so that the compiler does not fold the loop-carried set back
    // into loads at the point of use.  Cells are still added in increasing order.
    constexpr unsigned S = kTabU * kTileBS;
    TabCells A, B;
    // the first cells are requested before the tables are staged in LDS, so the two
    // memory round trips at the start of the work-group overlap
    fetch(A, beg + threadIdx.x);
    if (!stage()) return false;
#pragma unroll
    for (int q = 0; q < kTabU; q++) asm volatile("" : "+v"(A.c[q]), "+v"(A.dr[q]), "+v"(A.di[q]));
    for (unsigned base = beg + threadIdx.x; base < end; base += 2 * S) {
        fetch(B, base + S);
        __builtin_amdgcn_sched_barrier(0);
        eval(A);
        if (base + S >= end) break;
        fetch(A, base + 2 * S);
        __builtin_amdgcn_sched_barrier(0);
        eval(B);
    }
    s[1] *= kTwoPi;
    s[2] *= kTwoPi;
    constexpr double kTwoPi2 = kTwoPi * kTwoPi;
    s[3] *= kTwoPi2;
    s[4] *= kTwoPi2;
    s[5] *= kTwoPi2;
    return true;
}

// The Newton-pass sums of tile_sums64_tab from LDS records: grid index k's record at byte
// k * kRecBytes holds {X[k], Y[k], ctr[k]} as five doubles (8-byte aligned: read in pairs),
// and a cell's packed loff gives both its records' offsets, so per cell the vector memory
// carries loff and the packed {D_re, D_im}, and the LDS addresses need no arithmetic
// beyond unpacking.  The arithmetic and its order are tile_sums64_tab's.
struct RecEntry {
    double xr, xi, yr, yi, c;
};
// The cell centres as the records hold them (RecEntry::c), indexed like the centre array.
struct RecCtr {
    const RecEntry *rec;
    __device__ double operator[](unsigned k) const { return rec[k].c; }
};
static_assert(sizeof(RecEntry) == kRecBytes, "record layout");
struct RecCells {
    uint32_t o[kRecU];
    double dr[kRecU], di[kRecU];
};
template <class Stage>
__device__ inline bool tile_sums64_rec(unsigned pad, const uint32_t *loff, const double2 *dc,
                                       const LikItem &it, double ro,
                                       const Stage &stage, const char *rec, unsigned beg,
                                       unsigned end, double (&s)[7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(ro);
    const double sf = (scale * scale) / (detA * detA);
    const double sw = sf * (4 * kPi * kPi);
    const double a1 = 1. + e1, a2 = 1. - e1;
#pragma unroll
    for (int j = 0; j < 7; j++) s[j] = 0.;
    auto fetch = [&](RecCells &b, unsigned base) {
#pragma unroll
        for (int q = 0; q < kRecU; q++) {
            const unsigned i = base + q * kTileBS;
            const unsigned k = i < end ? i : pad;
            b.o[q] = gload(loff, k);
            const double2 d = gload(dc, k);
            b.dr[q] = d.x;
            b.di[q] = d.y;
        }
    };
    auto eval = [&](const RecCells &b) {
        double uu[kRecU], vv[kRecU];
        double2 x[kRecU], y[kRecU];
#pragma unroll
        for (int q = 0; q < kRecU; q++) {
            const RecEntry &col = *reinterpret_cast<const RecEntry *>(rec + (b.o[q] & 0xffffu));
            const RecEntry &row = *reinterpret_cast<const RecEntry *>(rec + (b.o[q] >> 16));
            uu[q] = col.c;
            vv[q] = row.c;
            x[q] = make_double2(col.xr, col.xi);
            y[q] = make_double2(row.yr, row.yi);
        }
#pragma unroll
        for (int q = 0; q < kRecU; q++) {
            const double k1 = fma(e2, vv[q], a1 * uu[q]);
            const double k2 = fma(a2, vv[q], e2 * uu[q]);
            const double rr = rsqrt_ge1(fma(sw, fma(k1, k1, k2 * k2), 1.));
            const double M = rr * rr * rr;
            const double hr = b.dr[q] * M;
            const double hi = -b.di[q] * M;
            const double cs = fma(x[q].x, y[q].x, -(x[q].y * y[q].y));
            const double sn = fma(x[q].y, y[q].x, x[q].x * y[q].y);
            const double re = fma(hr, cs, -hi * sn);
            const double im = fma(hr, sn, hi * cs);
            s[0] += re;
            s[1] = fma(-im, uu[q], s[1]);
            s[2] = fma(-im, vv[q], s[2]);
            const double rU = re * uu[q];
            s[3] = fma(-rU, uu[q], s[3]);
            s[4] = fma(-re * vv[q], vv[q], s[4]);
            s[5] = fma(-rU, vv[q], s[5]);
        }
    };
    constexpr unsigned S = kRecU * kTileBS;
    RecCells A, B;
    fetch(A, beg + threadIdx.x);
    if (!stage()) return false;
#pragma unroll
    for (int q = 0; q < kRecU; q++) asm volatile("" : "+v"(A.o[q]), "+v"(A.dr[q]), "+v"(A.di[q]));
    for (unsigned base = beg + threadIdx.x; base < end; base += 2 * S) {
        fetch(B, base + S);
        __builtin_amdgcn_sched_barrier(0);
        eval(A);
        if (base + S >= end) break;
        fetch(A, base + 2 * S);
        __builtin_amdgcn_sched_barrier(0);
        eval(B);
    }
    s[1] *= kTwoPi;
    s[2] *= kTwoPi;
    constexpr double kTwoPi2 = kTwoPi * kTwoPi;
    s[3] *= kTwoPi2;
    s[4] *= kTwoPi2;
    s[5] *= kTwoPi2;
    return true;
}

// The same sums from the FP32 facet copy: model and terms in FP32, the per-thread FP32
// partials added into FP64 every kFlush cells; the Newton phase formed and reduced in
// FP64 revolutions.
template <int PASS>
__device__ inline void tile_sums32(const FacetDev &F, const LikItem &it, double ro, double xo,
                                   double yo, unsigned beg, unsigned end, double (&s)[7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(ro);
    const double sf = (scale * scale) / (detA * detA);
    const float sw = (float)(sf * (4 * kPi * kPi));
    const float a1 = (float)(1. + e1), a2 = (float)(1. - e1), fe2 = (float)e2;
    const float tp = 6.28318530717958648f;
#pragma unroll
    for (int j = 0; j < 7; j++) s[j] = 0.;
    for (unsigned base = beg + threadIdx.x; base < end; base += kFlush * kTileBS) {
        float f[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const unsigned stop = min(end, base + kFlush * kTileBS);
#pragma unroll 2
        for (unsigned i = base; i < stop; i += kTileBS) {
            const float4 p = gload(F.pk, i);
            const float k1 = fmaf(fe2, p.y, a1 * p.x);
            const float k2 = fmaf(a2, p.y, fe2 * p.x);
            const float rr = rsqrtf(fmaf(sw, fmaf(k1, k1, k2 * k2), 1.f));
            const float M = rr * rr * rr;
            const float hr = p.z * M;
            const float hi = -p.w * M;
            const float U = tp * p.x, V = tp * p.y;
            if (PASS == 0) {
                f[0] = fmaf(M * M, gload(F.invs2, i), f[0]);
                f[1] += hr;
                f[2] = fmaf(-hi, U, f[2]);
                f[3] = fmaf(-hi, V, f[3]);
                const float hU = hr * U;
                f[4] = fmaf(-hU, U, f[4]);
                f[5] = fmaf(-hr * V, V, f[5]);
                f[6] = fmaf(-hU, V, f[6]);
            } else {
                float sn, cs;
                sincos_rev(fma(xo, (double)p.x, yo * (double)p.y), sn, cs);
                const float re = fmaf(hr, cs, -hi * sn);
                const float im = fmaf(hr, sn, hi * cs);
                f[0] += re;
                f[1] = fmaf(-im, U, f[1]);
                f[2] = fmaf(-im, V, f[2]);
                const float rU = re * U;
                f[3] = fmaf(-rU, U, f[3]);
                f[4] = fmaf(-re * V, V, f[4]);
                f[5] = fmaf(-rU, V, f[5]);
            }
        }
#pragma unroll
        for (int j = 0; j < 7; j++) s[j] += (double)f[j];
    }
}

// The Newton loop of likelihood.c:284-350 one step at a time: the task's state T advances
// by the step from the sums v (PASS 0: the zero-offset sums, which also start the state);
// returns 1 when the task has finished, with ho and the normalised determinant for L_r.
template <int PASS>
__device__ int newton_update(TileTask &T, const double (&v)[7], double *ho_out,
                             double *det_out)
{
    double value, dx, dy, dxx, dyy, dxy;
    if (PASS == 0) {
        T.S2 = v[0];
        value = v[1];
        dx = v[2];
        dy = v[3];
        dxx = v[4];
        dyy = v[5];
        dxy = v[6];
        T.xo = 0.;
        T.yo = 0.;
        T.norm = 1.;
        T.iter = 0;
        T.fragile = 0;
    } else {
        value = v[0];
        dx = v[1];
        dy = v[2];
        dxx = v[3];
        dyy = v[4];
        dxy = v[5];
    }
    const double det = dxx * dyy - dxy * dxy;
    if (det > 0 && T.norm > 1.e-7 && T.iter < 50) {
        const double incx = (dx * dyy - dy * dxy) / det;
        const double incy = (dy * dxx - dx * dxy) / det;
        T.xo -= incx;
        T.yo -= incy;
        T.iter += 1;
        T.norm = sqrt(incx * incx + incy * incy);
        T.active = 1;
        return 0;
    }
    T.active = 0;
    const double S2 = T.S2;
    // branches an FP32 sum could decide differently: the determinant sign, the Newton
    // step cap, the position-marginalisation cut x <= Delta (likelihood.c:395)
    const double ho = value / sqrt(S2);
    const double scale = fabs(dxx * dyy) + dxy * dxy;
    const double B = ho * ho * 0.5, Delta = 0.5 * 5.991;
    T.fragile = (fabs(det) <= 1e-5 * scale) || T.iter >= 50 ||
                fabs(B - Delta) <= 1e-4 * fabs(B) || !isfinite(det);
    *ho_out = ho;
    *det_out = det / S2;
    return 1;
}

// newton_update on the task's state in memory, writing L_r when the task finishes (the
// step kernel of the two-kernel passes).
// The state is read from T and written to To (the same task in another buffer when the
// passes keep two generations of state; T itself otherwise).
template <int PASS>
__device__ int tile_step(const LikLaunch &a, const TileTask *T, TileTask *To, int item, int r,
                         const double (&v)[7], double *xo, double *yo)
{
    TileTask S = *T;
    double ho, det;
    const int done = newton_update<PASS>(S, v, &ho, &det);
    *To = S;
    if (done) finish_lr(a, item, r, ho, det, S.iter);
    *xo = S.xo;
    *yo = S.yo;
    return done;
}

// A task's chunk sums of one pass added in chunk order by one wave: lanes over chunks, then
// across lanes into lane 63.  P0 is the task's first chunk; chunk c follows at c * nr * 8.
__device__ inline void task_sums(const double *P0, unsigned nch, int nr, int lane,
                                 double (&v)[7])
{
#pragma unroll
    for (int j = 0; j < 7; j++) v[j] = 0.;
    for (unsigned c = lane; c < nch; c += 64)
#pragma unroll
        for (int j = 0; j < 7; j++) v[j] += P0[(size_t)c * nr * 8 + j];
#pragma unroll
    for (int j = 0; j < 7; j++) v[j] = wave_sum_last(v[j]);  // totals in lane 63
}

// Newton phase tables e^{2 pi i o ctr[k]} (o = xo for the columns, yo for the rows) by runs
// of kTabRun grid indices.  The cell centres are evenly spaced, ctr[k] = (k + 0.5) inc - len
// (evaluate_uv_grid.c:127-128), so ctr[run * kTabRun + j] - ctr[run * kTabRun] equals
// ctr[j] - ctr[0] to within a few ulp of len: entry k is the product of the run's first
// entry E[k / kTabRun] (sincos_rev64) and W[k % kTabRun] = e^{2 pi i o (ctr[j] - ctr[0])}, one
// sin/cos per run and per run offset instead of one per entry.  The step kernel and the
// fused passes build their tables this same way, so both give the same entries.
constexpr unsigned kTabRun = 16;
template <unsigned NRUN>
struct TabRot {
    double2 ex[NRUN], ey[NRUN], wx[kTabRun], wy[kTabRun];
};
__device__ inline double2 cis_rev(double t)
{
    double sn, cs;
    sincos_rev64(t, sn, cs);
    return make_double2(cs, sn);
}
__device__ inline double2 tab_entry(const double2 &e, const double2 &w)
{
    return make_double2(fma(e.x, w.x, -(e.y * w.y)), fma(e.y, w.x, e.x * w.y));
}
// The runs' first entries (columns: every run of the gf columns; rows: runs yb0..yb1 and,
// with y0, run 0) and both W tables, spread over the block's threads; the caller
// synchronises before reading R.  ctr(k): the cell centre k, from global memory or from a
// copy in LDS.
template <unsigned NRUN, class Ctr>
__device__ inline void tab_rot_stage(TabRot<NRUN> &R, const Ctr &ctr, unsigned gf, double xo,
                                     double yo, unsigned yb0, unsigned yb1, bool y0)
{
    const unsigned nb = (gf + kTabRun - 1) / kTabRun;
    const unsigned ny = yb1 - yb0 + 1 + (y0 ? 1u : 0u);
    const unsigned nw = min(gf, kTabRun);
    const double c0 = ctr[0];
    static_assert(sizeof(TabRot<NRUN>) == (2 * NRUN + 2 * kTabRun) * sizeof(double2),
                  "ex, ey, wx, wy contiguous");
    double2 *E = reinterpret_cast<double2 *>(&R);  // ex, ey, wx, wy in that order
    for (unsigned q = threadIdx.x; q < nb + ny + 2 * nw; q += blockDim.x) {
        // the entry's centre index, factor and slot by selects, then one centre read and one
        // sin/cos whatever the kind of entry: with them in branches, a wave holding several
        // kinds ran each branch's read and sin/cos one after the other
        const bool isx = q < nb, isy = !isx && q < nb + ny;
        const unsigned b = q - nb <= yb1 - yb0 ? yb0 + (q - nb) : 0u;  // a row run (isy)
        const unsigned w = q - nb - ny;                                  // < 2 nw (W entries)
        const unsigned j = w < nw ? w : w - nw;
        const unsigned ci = isx ? q * kTabRun : (isy ? b * kTabRun : j);
        const double o = isx ? xo : (isy ? yo : (w < nw ? xo : yo));
        const unsigned slot =
            isx ? q : (isy ? NRUN + b : 2 * NRUN + (w < nw ? 0u : kTabRun) + j);
        const double c = ctr[ci];
        const double run = o * c, off = o * (c - c0);  // a run's first entry; a W entry
        E[slot] = cis_rev(isx || isy ? run : off);
    }
}
template <unsigned NRUN>
__device__ inline double2 tab_x(const TabRot<NRUN> &R, unsigned k)
{
    return tab_entry(R.ex[k / kTabRun], R.wx[k % kTabRun]);
}
template <unsigned NRUN>
__device__ inline double2 tab_y(const TabRot<NRUN> &R, unsigned k)
{
    return tab_entry(R.ey[k / kTabRun], R.wy[k % kTabRun]);
}

// The chunk sums of a tile work-group into P[0..N-1]: each wave's total from lane 63, then
// thread j adds the waves' value j in wave order.
template <int N>
__device__ inline void tile_reduce(double (&v)[N], double *lds, double *P)
{
    constexpr int NW = kTileBS / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = wave_sum_last(v[j]);
    if (lane == 63)
#pragma unroll
        for (int j = 0; j < N; j++) lds[wave * N + j] = v[j];
    __syncthreads();
    if (threadIdx.x < N) {
        double t = 0.;
#pragma unroll
        for (int q = 0; q < NW; q++) t += lds[q * N + threadIdx.x];
        P[threadIdx.x] = t;
    }
    __syncthreads();  // lds is reused by the block's next logical work-group
}

template <int PASS, int KIND>
__device__ inline void tile_chunk(const TileLaunch &t, unsigned wg, double *red);

// Waves per SIMD the tile kernels are compiled for.  The generic kernel carries every path
// (the FP32 sums and the per-cell sin/cos among them) in 128 VGPRs; the grid kernel only
// the FP64 paths of extracted facets (pass 0 with the cell centres in LDS, Newton passes
// from the phase tables), which fit a smaller register budget, so more waves stay resident
// to cover the memory latency.
constexpr int kTileWaves = 4;
#ifndef RL2_TILE_GRID_WAVES
#define RL2_TILE_GRID_WAVES 4
#endif
constexpr int kTileGridWaves = RL2_TILE_GRID_WAVES;

// A block takes the logical work-groups q, q + nbx, q + 2 nbx, ... of its XCD's eighth of
// the logical order (q = block / 8, nbx = blocks per XCD): with one block per logical
// work-group that is exactly one, and a launch capped at the resident block count walks
// the whole pass without further dispatches.
template <int PASS, int KIND>
__global__ __launch_bounds__(kTileBS, KIND ? kTileGridWaves : kTileWaves) void lik_tile_kernel(
    TileLaunch t)
{
    __shared__ double red[(kTileBS / 64) * 8];
    if (PASS == 0 && blockIdx.x == 0 && threadIdx.x < t.n_inl)
        t.items_w[threadIdx.x] = t.inl[threadIdx.x];
    const unsigned per = (t.nwg + kXcds - 1) / kXcds;
    const unsigned nbx = gridDim.x / kXcds;
    for (unsigned j = blockIdx.x / kXcds; j < per; j += nbx) {
        const unsigned wg = (blockIdx.x % kXcds) * per + j;
        if (wg >= t.nwg) break;
        tile_chunk<PASS, KIND>(t, wg, red);
    }
}

// KIND 0: any items (the generic kernel).  KIND 1 and 2 (the grid kernels): every item of
// the launch is an FP64 item on an extracted facet with phase tables, which the host checks
// when it picks the kernel; KIND 2 when the items travel in the kernel arguments.
template <int PASS, int KIND>
__device__ inline void tile_chunk(const TileLaunch &t, unsigned wg, double *red)
{
    constexpr bool GRID = KIND != 0;
    const LikLaunch &a = t.lk;
    // KIND 2: the items from the kernel arguments (scalar loads that stay in the scalar
    // cache instead of a chain of dependent vector loads for the search).  Pass 0 reads them
    // there whenever they travel there: its block 0 is what stores them to a.items.
    const LikItem *items =
        KIND == 2 ? t.inl : ((PASS == 0 && t.n_inl) ? t.inl : a.items);
    int item;
    unsigned first, nch;
    tile_locate(t, items, wg, &item, &first, &nch);
    const unsigned local = wg - first;
    const unsigned chunk = local / (unsigned)t.nr_g;
    const int r = (int)(local - chunk * (unsigned)t.nr_g) + t.r_lo;
    const TileTask *T = t.tasks + (size_t)item * kMaxR + r;
    if (PASS == 1 && !T->active) return;  // finished task: the whole work-group moves on
    const LikItem it = items[item];
    const unsigned beg = chunk * it.chunk_cells;
    double s[7];
    if constexpr (GRID) {
        auto grid_sums = [&](const FacetDev &F) {
            const unsigned end = min(F.n, beg + it.chunk_cells);
            extern __shared__ double2 tab_lds[];
            if (PASS == 1) {
                // t.tab_lds_facet * (16 + 8) bytes, + 16 with the rows: X, ctr, Y
                const unsigned gf = F.grid_facet;
                const double2 *X = t.tab + it.tab_first + (size_t)(r - 1) * 2 * gf;
                const double2 *Y = X + gf;
                double2 *sX = tab_lds;
                double *sC = reinterpret_cast<double *>(tab_lds + t.tab_lds_facet);
                double2 *sY = reinterpret_cast<double2 *>(sC + t.tab_lds_facet);
                const bool yl = t.tab_y_lds;
                auto copy = [&]() {  // the step kernel's tables and the cell centres
                    const double *ctr = F.ctr;
                    for (unsigned k = threadIdx.x; k < gf; k += kTileBS) {
                        sX[k] = gload(X, k);
                        sC[k] = gload(ctr, k);
                        if (yl) sY[k] = gload(Y, k);
                    }
                    __syncthreads();
                    return true;
                };
                if (yl)
                    tile_sums64_tab<true>(F, it, a.rg.ro[r], copy, Y, sX, sY, sC, beg, end, s);
                else
                    tile_sums64_tab<false>(F, it, a.rg.ro[r], copy, Y, sX, sY, sC, beg, end, s);
            } else {
                double *sC = reinterpret_cast<double *>(tab_lds);  // tab_lds_facet * 8 bytes
                auto stage = [&]() {
                    for (unsigned k = threadIdx.x; k < F.grid_facet; k += kTileBS)
                        sC[k] = gload(F.ctr, k);
                    __syncthreads();
                };
                const double ro1[1] = {a.rg.ro[r]};
                tile_sums64_p0_db<1>(F, it, ro1, sC, stage, beg, end,
                                     reinterpret_cast<double(&)[1][7]>(s));
            }
        };
        grid_sums(a.facets[it.slot]);
    } else {
        const FacetDev &F = a.facets[it.slot];
        const unsigned end = min(F.n, beg + it.chunk_cells);
        if (PASS == 1 && it.prec == 0 && t.tab && F.cid) {
            extern __shared__ double2 tab_lds[];
            const unsigned gf = F.grid_facet;
            const double2 *X = t.tab + it.tab_first + (size_t)(r - 1) * 2 * gf;
            const double2 *Y = X + gf;
            double2 *sX = tab_lds;
            double *sC = reinterpret_cast<double *>(tab_lds + t.tab_lds_facet);
            double2 *sY = reinterpret_cast<double2 *>(sC + t.tab_lds_facet);
            const bool yl = t.tab_y_lds;
            auto copy = [&]() {
                const double *ctr = F.ctr;
                for (unsigned k = threadIdx.x; k < gf; k += kTileBS) {
                    sX[k] = gload(X, k);
                    sC[k] = gload(ctr, k);
                    if (yl) sY[k] = gload(Y, k);
                }
                __syncthreads();
                return true;
            };
            if (yl)
                tile_sums64_tab<true>(F, it, a.rg.ro[r], copy, Y, sX, sY, sC, beg, end, s);
            else
                tile_sums64_tab<false>(F, it, a.rg.ro[r], copy, Y, sX, sY, sC, beg, end, s);
        } else if (PASS == 0 && it.prec == 0 && F.cid && F.grid_facet <= t.tab_lds_facet) {
            extern __shared__ double2 tab_lds[];  // t.tab_lds_facet * 8 bytes: ctr
            double *sC = reinterpret_cast<double *>(tab_lds);
            for (unsigned k = threadIdx.x; k < F.grid_facet; k += kTileBS)
                sC[k] = gload(F.ctr, k);
            __syncthreads();
            tile_sums64_p0_grid<kTileU>(F, it, a.rg.ro[r], sC, beg, end, s);
        } else if (it.prec == 0)
            tile_sums64<PASS>(F, it, a.rg.ro[r], PASS == 1 ? T->xo : 0.,
                              PASS == 1 ? T->yo : 0., beg, end, s);
        else
            tile_sums32<PASS>(F, it, a.rg.ro[r], PASS == 1 ? T->xo : 0.,
                              PASS == 1 ? T->yo : 0., beg, end, s);
    }
    tile_reduce<7>(s, red, t.partials + (size_t)wg * 8);
}

// One work-group per (item, r) task.  Wave 0 adds the task's chunk sums in a fixed order
// (lanes over chunks, then across lanes into lane 63), and lane 63 takes the Newton step or
// ends the task; when
// the task goes on over a grid facet in FP64, the whole group (16 waves, so each lane
// writes at most 3 entries of the 2 * facet) then writes its phase tables for the next pass.
constexpr int kStepBS = 1024;
template <int PASS>
__global__ __launch_bounds__(kStepBS) void lik_step_kernel(TileLaunch t)
{
    __shared__ int s_go;
    __shared__ double s_off[2];
    const LikLaunch &a = t.lk;
    const int nr = t.nr_g;
    const unsigned k = blockIdx.x;
    const int item = (int)(k / nr), r = (int)(k % nr) + t.r_lo;
    TileTask *T = t.tasks + (size_t)item * kMaxR + r;
    TileTask *To = (t.tasks_out ? t.tasks_out : t.tasks) + (size_t)item * kMaxR + r;
    // the item and its facet fields are requested by every wave at once, so their latency
    // overlaps the reduction instead of following the barrier
    const LikItem it = a.items[item];
    const FacetDev &F = a.facets[it.slot];
    const uint32_t *fcid = F.cid;
    const double *fctr = F.ctr;
    const unsigned gf = F.grid_facet;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int go = 0;
        if (PASS == 0 || T->active) {  // read by the whole wave before lane 63 updates it
            unsigned first, nch;
            tile_range(t, a.items, item, &first, &nch);
            double v[7];
            task_sums(t.partials + (size_t)(first + (unsigned)(r - t.r_lo)) * 8, nch, nr, lane,
                      v);
            if (lane == 63) go = !tile_step<PASS>(a, T, To, item, r, v, &s_off[0], &s_off[1]);
        } else if (lane == 63 && To != T) {
            *To = *T;  // a finished task keeps its state in the other generation too
        }
        if (lane == 63) s_go = go;
    }
    __syncthreads();
    if (!s_go || !t.tab) return;
    if (it.prec != 0 || !fcid) return;
    double2 *X = t.tab + it.tab_first + (size_t)(r - 1) * 2 * gf;
    // every column and every row (gf <= 1706: the host's LDS limit for the column tables)
    __shared__ TabRot<128> R;
    const unsigned nb = (gf + kTabRun - 1) / kTabRun;
    tab_rot_stage(R, fctr, gf, s_off[0], s_off[1], 0u, nb - 1, false);
    __syncthreads();
    for (unsigned q = threadIdx.x; q < 2 * gf; q += blockDim.x)
        X[q] = q < gf ? tab_x(R, q) : tab_y(R, q - gf);
}


// status[0] = tasks still iterating, status[1] = largest Newton step count,
// *flops = sum over tasks of n (37 + 25 N_it) (model + normalisation 19, zero-offset sums
// 18, 25 per Newton pass; transcendentals, sqrt and division counted as one).  Run by the
// first wave of one extra work-group of the r-marginalisation launch.
__device__ void tile_status(const LikLaunch &a)
{
    if (threadIdx.x >= 64) return;
    const int nr = a.rg.numr - 1;
    const unsigned ntask = (unsigned)a.nitems * nr;
    unsigned c = 0, m = 0;
    double f = 0.;
    for (unsigned k = threadIdx.x; k < ntask; k += 64) {
        const int item = (int)(k / nr);
        const TileTask &T = a.tasks[(size_t)item * kMaxR + k % nr + 1];
        c += T.active ? 1u : 0u;
        m = max(m, (unsigned)T.iter);
        f += (double)a.facets[a.items[item].slot].n * (37. + 25. * T.iter);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        c += __shfl_xor(c, off);
        m = max(m, (unsigned)__shfl_xor((int)m, off));
        f += __shfl_xor(f, off);
    }
    if (threadIdx.x == 0) {
        a.status[0] = c;
        a.status[1] = m;
        *a.flops = f;
    }
}

// ------------------------------------------------------------------ Newton passes, fused
// What a fused pass's work-group reads of its task before anything else, so that these
// loads are in flight with the facet's and the cells': the task's state S (the same address
// in every thread: scalar loads) and wave 0's share of the previous pass's chunk sums of the
// task (lane c: chunk c, when nch <= 64; task_sums reads them itself otherwise).
struct FusedPre {
    TileTask S;
    double p[7];
};

__device__ __attribute__((always_inline)) inline FusedPre fused_pre(const TileLaunch &t,
                                                                    unsigned first,
                                                                    unsigned nch, int item,
                                                                    int r)
{
    FusedPre f;
    f.S = t.tasks[(size_t)item * kMaxR + r];
#pragma unroll
    for (int j = 0; j < 7; j++) f.p[j] = 0.;
    const unsigned lane = threadIdx.x;
    if (t.mode != 2 && nch <= 64 && lane < nch) {
        const double *P = t.part_in + (size_t)(first + (unsigned)(r - t.r_lo)) * 8 +
                          (size_t)lane * t.nr_g * 8;
#pragma unroll
        for (int j = 0; j < 7; j++) f.p[j] = P[j];
    }
    return f;
}

// The step a fused pass's work-group takes before its sums (wave 0; the caller
// synchronises), for a task that is still iterating (fused_idle has dealt with the others):
// the task's chunk sums of the previous pass added in chunk order, the Newton step, the
// state stored by the task's chunk-0 work-group (to the other generation) with L_r when the
// task ends; *go and off[2] receive whether the task goes on and its new offset.
__device__ __attribute__((always_inline)) inline void fused_step(const TileLaunch &t, unsigned first,
                                                                 unsigned nch,
                                  unsigned chunk, int item, int r, const FusedPre &pre,
                                  int *go_out, double *off)
{
    if (threadIdx.x >= 64) return;
    const LikLaunch &a = t.lk;
    const size_t ti = (size_t)item * kMaxR + r;
    const int lane = threadIdx.x;
    int go = 0;
    double xo = 0., yo = 0.;
    if (t.mode == 2) {
        if (lane == 63) {
            go = pre.S.active;
            xo = pre.S.xo;
            yo = pre.S.yo;
        }
    } else {
        double v[7];
        if (nch <= 64) {  // task_sums' arithmetic on the chunk sums read at the start
#pragma unroll
            for (int j = 0; j < 7; j++) {
                double pj = pre.p[j];
                asm volatile("" : "+v"(pj));  // consumed here, not where the load was issued
                v[j] = 0.;
                v[j] += pj;
            }
#pragma unroll
            for (int j = 0; j < 7; j++) v[j] = wave_sum_last(v[j]);
        } else {
            task_sums(t.part_in + (size_t)(first + (unsigned)(r - t.r_lo)) * 8, nch, t.nr_g,
                      lane, v);
        }
        if (lane == 63) {
            TileTask S = t.mode == 0 ? TileTask{} : pre.S;
            double ho, det;
            const int done = t.mode == 0 ? newton_update<0>(S, v, &ho, &det)
                                         : newton_update<1>(S, v, &ho, &det);
            if (chunk == 0) {
                t.tasks_out[ti] = S;
                if (done) finish_lr(a, item, r, ho, det, S.iter);
            }
            go = !done;
            xo = S.xo;
            yo = S.yo;
        }
    }
    if (lane == 63) {
        *go_out = go;
        off[0] = xo;
        off[1] = yo;
    }
}

// Whether a fused pass's work-group has a task that ended in an earlier pass (mode 1), in
// which case the task's chunk-0 work-group keeps its state in both generations.  Every wave
// holds the flag (read with scalar loads, whose wait does not include the cells' vector
// loads in flight), so an idle work-group leaves without a barrier, which would wait for
// those loads.
__device__ inline bool fused_idle(const TileLaunch &t, unsigned chunk, int item, int r,
                                  const FusedPre &pre)
{
    if (t.mode != 1 || pre.S.active) return false;
    if (chunk == 0 && threadIdx.x == 63) t.tasks_out[(size_t)item * kMaxR + r] = pre.S;
    return true;
}

// One Newton pass with the previous pass's step folded in, so that a pass is one launch
// instead of two.  Each work-group first adds its task's chunk sums of the previous pass
// (wave 0, in the step kernel's order) and takes the Newton step itself, so every
// work-group of a task holds the same state bit for bit.  The task's chunk-0 work-group
// stores that state to the other generation of the task buffer (the rest of the pass still
// reads this one) and writes L_r when the task ends.  The work-group then evaluates its own
// phase tables in LDS with the step kernel's arithmetic: the column table over the facet's
// width, the row table over the rows its chunk spans (cells are in row-major order) and row
// 0, which the zero pad cell reads.  t.mode: 0 the first step, from the pass-0 sums; 1 a
// Newton step from the previous pass's sums; 2 resume from the state the step kernel left.
template <int KIND>
__global__ __launch_bounds__(kTileBS, kTileGridWaves) void lik_fused_kernel(TileLaunch t)
{
    __shared__ double red[(kTileBS / 64) * 8];
    __shared__ int s_go;
    __shared__ double s_off[2];
    // one block per logical work-group, in the XCD-aware order of lik_tile_kernel
    const unsigned per = (t.nwg + kXcds - 1) / kXcds;
    const unsigned wg = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (wg >= t.nwg) return;
    const LikLaunch &a = t.lk;
    const LikItem *items = KIND == 2 ? t.inl : a.items;
    int item;
    unsigned first, nch;
    tile_locate(t, items, wg, &item, &first, &nch);
    const unsigned local = wg - first;
    const unsigned chunk = local / (unsigned)t.nr_g;
    const int r = (int)(local - chunk * (unsigned)t.nr_g) + t.r_lo;
    const FusedPre pre = fused_pre(t, first, nch, item, r);
    const LikItem it = items[item];
    const FacetDev &F = a.facets[it.slot];
    const unsigned fn = F.n;
    const unsigned gf = F.grid_facet;  // <= 1024: the host checks the LDS tables fit
    const uint32_t *cid = F.cid, *loff = F.loff;
    const double *ctr = F.ctr;
    const double2 *dc = F.dc;
    // the facet's fields arrive with the task's state: one wait for both, before the
    // cells' loads are issued
    asm volatile("" ::"s"(fn), "s"(gf), "s"(cid), "s"(loff), "s"(ctr), "s"(dc));
    const unsigned beg = chunk * it.chunk_cells;
    const unsigned end = min(fn, beg + it.chunk_cells);
    extern __shared__ double2 tab_lds[];  // t.tab_lds_facet records of kRecBytes
    RecEntry *rec = reinterpret_cast<RecEntry *>(tab_lds);
    // Runs after the first cells' loads are issued (inside tile_sums64_rec): the cell
    // centres this thread's records need are requested too, then wave 0 adds the previous
    // pass's chunk sums and takes the step while those loads are in flight.
    __shared__ TabRot<1024 / kTabRun> R;  // gf <= 1024
    auto make = [&]() -> bool {
        if (fused_idle(t, chunk, item, r, pre)) return false;
        constexpr int NC = 1024 / kTileBS;
        double c[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) {  // loads without branches: one count of them to wait on
            const unsigned k = threadIdx.x + j * kTileBS;
            const double v = gload(ctr, k < gf ? k : 0u);
            c[j] = k < gf ? v : 0.;
        }
        uint32_t id_lo = end > beg ? gload(cid, beg) : 0u;
        uint32_t id_hi = end > beg ? gload(cid, end - 1) : 0u;
        fused_step(t, first, nch, chunk, item, r, pre, &s_go, s_off);
        // the records' centres do not depend on the step: stored while wave 0 takes it, and
        // the tables read their runs' centres from them
#pragma unroll
        for (int j = 0; j < NC; j++) {
            const unsigned k = threadIdx.x + j * kTileBS;
            if (k < gf) rec[k].c = c[j];
        }
        __syncthreads();
        if (!s_go) return false;
        // the chunk's first and last cell ids are waited for only here: a wait for them
        // before the step would also wait for every cell and centre load issued before them
        asm volatile("" : "+v"(id_lo), "+v"(id_hi));
        const unsigned row_lo = __builtin_amdgcn_readfirstlane(id_lo >> 16);
        const unsigned row_hi = __builtin_amdgcn_readfirstlane(id_hi >> 16);
        // the column runs, the runs of the rows the chunk spans (cells are in row-major
        // order) and run 0 (row 0: the zero pad cell's), then the entries
        const unsigned yb0 = row_lo / kTabRun, yb1 = row_hi / kTabRun;
        tab_rot_stage(R, RecCtr{rec}, gf, s_off[0], s_off[1], yb0, yb1, yb0 > 0);
        __syncthreads();
        // the column entries: every run and W value read before any is used
        double2 ex[NC], wx[NC];
#pragma unroll
        for (int j = 0; j < NC; j++) {
            const unsigned k = threadIdx.x + j * kTileBS;
            const unsigned kk = k < gf ? k : 0u;
            ex[j] = R.ex[kk / kTabRun];
            wx[j] = R.wx[kk % kTabRun];
        }
#pragma unroll
        for (int j = 0; j < NC; j++) {
            const unsigned k = threadIdx.x + j * kTileBS;
            if (k < gf) {
                const double2 x = tab_entry(ex[j], wx[j]);
                rec[k].xr = x.x;
                rec[k].xi = x.y;
            }
        }
        const unsigned span = row_hi - row_lo + 1;
        for (unsigned q = threadIdx.x; q < span + (row_lo > 0 ? 1u : 0u); q += kTileBS) {
            const unsigned k = q < span ? row_lo + q : 0u;
            const double2 y = tab_y(R, k);
            rec[k].yr = y.x;
            rec[k].yi = y.y;
        }
        __syncthreads();
        return true;
    };
    double s[7];
    if (tile_sums64_rec(fn, loff, dc, it, a.rg.ro[r], make, reinterpret_cast<const char *>(rec),
                        beg, end, s))
        tile_reduce<7>(s, red, t.partials + (size_t)wg * 8);
}

// ------------------------------------------------------------------ one launch per evaluation
// tile_reduce, with the chunk's sums published to the other work-groups of the launch: the
// stores are made at agent scope, wave 0 (which issued them) waits until they have
// completed, and only then counts the chunk as done on the task's counter.  A reader that
// sees the count loads the sums at agent scope too.  Every sum lives at its own address for
// the whole launch (one slice per pass), so no reader can hold an older copy of it.
template <int N>
__device__ inline void tile_reduce_publish(double (&v)[N], double *lds, double *P,
                                           unsigned *cnt)
{
    constexpr int NW = kTileBS / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = wave_sum_last(v[j]);
    if (lane == 63)
#pragma unroll
        for (int j = 0; j < N; j++) lds[wave * N + j] = v[j];
    __syncthreads();
    if (threadIdx.x < 64) {
        if (threadIdx.x < N) {
            double t = 0.;
#pragma unroll
            for (int q = 0; q < NW; q++) t += lds[q * N + threadIdx.x];
            __hip_atomic_store(P + threadIdx.x, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (threadIdx.x == 0)
            __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();  // lds is reused by the next chunk
}

// A task's chunk sums of one pass as task_sums, loaded at agent scope (published by
// tile_reduce_publish in other work-groups of the launch).
__device__ inline void task_sums_published(const double *P0, unsigned nch, int nr, int lane,
                                           double (&v)[7])
{
#pragma unroll
    for (int j = 0; j < 7; j++) v[j] = 0.;
    for (unsigned c = lane; c < nch; c += 64)
#pragma unroll
        for (int j = 0; j < 7; j++)
            v[j] += __hip_atomic_load(P0 + (size_t)c * nr * 8 + j, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int j = 0; j < 7; j++) v[j] = wave_sum_last(v[j]);
}

// One block's share of an evaluation: a task, and the run of its chunks the block sums.
struct PersistTask {
    int item, r;
    unsigned c0, c1;       // chunks [c0, c1) of the task
    unsigned nch;          // the task's chunk count
    unsigned first;        // its item's first work-group index (chunk-major layout)
    int active;
    int pad;
    TileTask S;            // the block's copy of the task's Newton state
};

// Every Newton pass of one likelihood evaluation, to convergence, in one cooperative launch
// after the pass-0 launch, so that no Newton pass pays a launch, a ramp and a tail.  Work layout:
// XCD x (blocks b with b % 8 == x) owns chunks [x nc / 8, (x + 1) nc / 8) of every task
// (nc: the task's chunk count), so the tasks that read the same cells share that XCD's L2;
// within the XCD the (task, chunk) pairs are listed task by task and each block takes an
// equal contiguous run of the list, so a block holds one or a few tasks and takes their
// Newton steps and builds their phase tables once per pass.  The chunks, their sums and
// the steps are those of the fused passes, bit for bit.
// Between passes a task's chunk sums of pass q >= 1 are complete when its counter reaches
// q nc (pass 0's are complete at launch).  A block waits on that count only for its own task's previous pass, and it
// has finished all its work of every earlier pass, so a waiting block always depends on
// blocks that are running or done: with every block resident (cooperative launch) the
// waits end.  Each wait is bounded all the same (p.spin_limit polls): a block that gives
// up leaves its tasks marked active, which the host reports as an error.

template <int KIND>
__global__ __launch_bounds__(kTileBS, kTileGridWaves) void lik_persist_kernel(PersistLaunch p)
{
    __shared__ double red[(kTileBS / 64) * 8];
    __shared__ PersistTask s_task[kPersistTasks];
    __shared__ int s_ntask, s_go, s_fail;
    __shared__ double s_off[2];
    __shared__ TabRot<1024 / kTabRun> R;
    extern __shared__ double2 tab_lds[];  // t.tab_lds_facet records
    RecEntry *rec = reinterpret_cast<RecEntry *>(tab_lds);
    const TileLaunch &t = p.t;
    const LikLaunch &a = t.lk;
    const LikItem *items = KIND == 2 ? t.inl : a.items;
    const int nr = t.nr_g;
    const unsigned x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    // items that travel in the arguments are stored for the r-marginalisation
    if (KIND == 2 && blockIdx.x == 0 && threadIdx.x < t.n_inl)
        t.items_w[threadIdx.x] = t.inl[threadIdx.x];
    if (threadIdx.x == 0) {
        // the XCD's list: item by item, task by task, the task's chunks in the XCD's eighth
        auto chunks = [&](int i) {
            const unsigned c1 = i + 1 < a.nitems ? items[i + 1].chunk_first : t.nwg / nr;
            return c1 - items[i].chunk_first;
        };
        unsigned long long W = 0;
        for (int i = 0; i < a.nitems; i++) {
            const unsigned nc = chunks(i);
            W += (unsigned long long)nr * ((x + 1) * nc / kXcds - x * nc / kXcds);
        }
        const unsigned long long q0 = j * W / p.nbx, q1 = (j + 1) * W / p.nbx;
        int nt = 0;
        unsigned long long P = 0;
        for (int i = 0; i < a.nitems && P < q1 && nt >= 0; i++) {
            const unsigned nc = chunks(i);
            const unsigned lo = x * nc / kXcds, m = (x + 1) * nc / kXcds - lo;
            for (int r = 1; r <= nr && m; r++, P += m) {
                if (P + m <= q0 || P >= q1) continue;
                if (nt == kPersistTasks) {
                    nt = -1;
                    break;
                }
                PersistTask &k = s_task[nt++];
                k.item = i;
                k.r = r;
                k.c0 = lo + (unsigned)(q0 > P ? q0 - P : 0);
                k.c1 = lo + (unsigned)(q1 < P + m ? q1 - P : m);
                k.nch = nc;
                k.first = items[i].chunk_first * (unsigned)nr;
                k.active = 1;
            }
        }
        s_ntask = nt;
        s_fail = 0;
    }
    __syncthreads();
    const int ntask = s_ntask;  // < 0: more tasks than a block holds (the host prevents it)
    // the owner of a task's chunk 0 publishes its state: active until it finishes
    for (int k = 0; k < ntask; k++)
        if (threadIdx.x == 0 && s_task[k].c0 == 0)
            p.t.tasks[(size_t)s_task[k].item * kMaxR + s_task[k].r].active = 1;
    // pass 0's sums are in slice 0 already (the pass-0 launch before this one, stream-ordered);
    // passes 1 .. kPersistPasses - 1 sum; the last iteration only takes the final step
    for (int pass = 1; pass <= kPersistPasses && ntask > 0; pass++) {
        bool any = false;
        for (int k = 0; k < ntask; k++) {
            PersistTask &K = s_task[k];
            if (!K.active) continue;  // uniform: read after a barrier
            const LikItem it = items[K.item];
            const FacetDev &F = a.facets[it.slot];
            const unsigned gf = F.grid_facet;
            const size_t ti = (size_t)K.item * kMaxR + K.r;
            const double ro = a.rg.ro[K.r];
            {
                // the previous pass's sums of this task, then its Newton step
                if (threadIdx.x == 0 && pass > 1) {
                    const unsigned want = (unsigned)(pass - 1) * K.nch;
                    unsigned spins = 0;
                    while (__hip_atomic_load(p.cnt + ti, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) < want) {
                        if (++spins > p.spin_limit) {
                            s_fail = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                __syncthreads();
                if (s_fail) return;  // the task stays marked active: the host reports it
                if (threadIdx.x < 64) {
                    const int lane = threadIdx.x;
                    double v[7];
                    task_sums_published(p.part + ((size_t)(pass - 1) * t.nwg + K.first +
                                                  (unsigned)(K.r - 1)) * 8,
                                        K.nch, nr, lane, v);
                    if (lane == 63) {
                        TileTask S = pass == 1 ? TileTask{} : K.S;
                        double ho, det;
                        const int done = pass == 1 ? newton_update<0>(S, v, &ho, &det)
                                                   : newton_update<1>(S, v, &ho, &det);
                        K.S = S;
                        K.active = !done;
                        if (done && K.c0 == 0) {
                            p.t.tasks[ti] = S;
                            finish_lr(a, K.item, K.r, ho, det, S.iter);
                        }
                        s_off[0] = S.xo;
                        s_off[1] = S.yo;
                    }
                }
                __syncthreads();
                if (!K.active) continue;
                // phase tables: every column, the rows this block's cells span, row 0
                const unsigned beg = K.c0 * it.chunk_cells;
                const unsigned end = min(F.n, K.c1 * it.chunk_cells);
                const uint32_t *cid = F.cid;
                const unsigned row_lo = end > beg ? gload(cid, beg) >> 16 : 0u;
                const unsigned row_hi = end > beg ? gload(cid, end - 1) >> 16 : 0u;
                const unsigned yb0 = row_lo / kTabRun, yb1 = row_hi / kTabRun;
                tab_rot_stage(R, F.ctr, gf, s_off[0], s_off[1], yb0, yb1, yb0 > 0);
                __syncthreads();
                for (unsigned q = threadIdx.x; q < gf; q += kTileBS) {
                    const double2 xx = tab_x(R, q);
                    rec[q].xr = xx.x;
                    rec[q].xi = xx.y;
                    rec[q].c = gload(F.ctr, q);
                }
                const unsigned span = row_hi - row_lo + 1;
                for (unsigned q = threadIdx.x; q < span + (row_lo > 0 ? 1u : 0u); q += kTileBS) {
                    const unsigned kk = q < span ? row_lo + q : 0u;
                    const double2 y = tab_y(R, kk);
                    rec[kk].yr = y.x;
                    rec[kk].yi = y.y;
                }
                __syncthreads();
            }
            any = true;
            for (unsigned c = K.c0; c < K.c1; c++) {
                const unsigned beg = c * it.chunk_cells;
                const unsigned end = min(F.n, beg + it.chunk_cells);
                double s[7];
                auto none = []() { return true; };  // the records are already staged
                tile_sums64_rec(F.n, F.loff, F.dc, it, ro, none,
                                reinterpret_cast<const char *>(rec), beg, end, s);
                tile_reduce_publish<7>(
                    s, red, p.part + ((size_t)pass * t.nwg + K.first + c * nr + (K.r - 1)) * 8,
                    p.cnt + ti);
            }
            __syncthreads();  // the records are restaged for the next task
        }
        if (!any) break;
    }
}

// Largest block count of a cooperative persistent launch with `lds` bytes of dynamic LDS
// (0: the kernel cannot run).
unsigned persist_capacity(unsigned lds, int kind)
{
    static std::map<std::pair<unsigned, int>, unsigned> cache;
    const auto key = std::make_pair(lds, kind);
    auto f = cache.find(key);
    if (f != cache.end()) return f->second;
    int per = 0, dev = 0, cus = 0;
    const bool queried =
        hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        (kind == 2 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, lik_persist_kernel<2>,
                                                                  kTileBS, lds)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, lik_persist_kernel<1>,
                                                                  kTileBS, lds)) == hipSuccess;
    if (!queried) return 0;  // not cached: the next call asks again
    const unsigned cap = (unsigned)std::max(0, per) * (unsigned)std::max(0, cus);
    cache[key] = cap;
    return cap;
}

bool launch_persist(hipStream_t s, const PersistLaunch &p, unsigned blocks)
{
    const unsigned lds = p.t.tab_lds_facet * kRecBytes;
    void *args[] = {const_cast<PersistLaunch *>(&p)};
    const void *fn = p.t.grid == 2 ? (const void *)lik_persist_kernel<2>
                                   : (const void *)lik_persist_kernel<1>;
    return hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(kTileBS), args, lds, s) ==
           hipSuccess;
}

// Fused Newton passes on facets wider than kRecFacet (up to kWideFacet columns): a 24-byte
// record {X, ctr} per column in LDS, and the row-table entries of the rows the work-group's
// chunk spans in a small cache (entry `span` holds row 0, which the zero pad cell reads;
// a span beyond the cache evaluates the row entries per cell from the run tables).  Cell
// ids from cid; the arithmetic and its order are tile_sums64_tab's.
constexpr unsigned kWideRows = 256;
struct ColEntry {
    double xr, xi, c;
};
// The cell centres as the column records hold them (ColEntry::c).
struct ColCtr {
    const ColEntry *col;
    __device__ double operator[](unsigned k) const { return col[k].c; }
};
template <bool CACHED, class Stage>
__device__ inline bool tile_sums64_col(const FacetDev &F, const LikItem &it, double ro,
                                       const Stage &stage, const ColEntry *col,
                                       const double2 *sY, const TabRot<128> &R,
                                       unsigned row_lo, unsigned span, unsigned beg,
                                       unsigned end, double (&s)[7])
{
    const double e1 = it.e1, e2 = it.e2;
    const double detA = 1. - e1 * e1 - e2 * e2;
    const double scale = arcs2rad_scale(ro);
    const double sf = (scale * scale) / (detA * detA);
    const double sw = sf * (4 * kPi * kPi);
    const double a1 = 1. + e1, a2 = 1. - e1;
#pragma unroll
    for (int j = 0; j < 7; j++) s[j] = 0.;
    const uint32_t *cid = F.cid;
    const double2 *dc = F.dc;
    const unsigned pad = F.n;
    auto fetch = [&](RecCells &b, unsigned base) {
#pragma unroll
        for (int q = 0; q < kTabU; q++) {
            const unsigned i = base + q * kTileBS;
            const unsigned k = i < end ? i : pad;
            b.o[q] = gload(cid, k);
            const double2 d = gload(dc, k);
            b.dr[q] = d.x;
            b.di[q] = d.y;
        }
    };
    auto eval = [&](const RecCells &b) {
        double uu[kTabU], vv[kTabU];
        double2 x[kTabU], y[kTabU];
#pragma unroll
        for (int q = 0; q < kTabU; q++) {
            const unsigned ci = b.o[q] & 0xffffu, cj = b.o[q] >> 16;
            const ColEntry &cc = col[ci];
            uu[q] = cc.c;
            vv[q] = col[cj].c;
            x[q] = make_double2(cc.xr, cc.xi);
            // rows below row_lo (only row 0, the pad cell's) wrap past span: entry span
            y[q] = CACHED ? sY[min(cj - row_lo, span)] : tab_y(R, cj);
        }
#pragma unroll
        for (int q = 0; q < kTabU; q++) {
            const double k1 = fma(e2, vv[q], a1 * uu[q]);
            const double k2 = fma(a2, vv[q], e2 * uu[q]);
            const double rr = rsqrt_ge1(fma(sw, fma(k1, k1, k2 * k2), 1.));
            const double M = rr * rr * rr;
            const double hr = b.dr[q] * M;
            const double hi = -b.di[q] * M;
            const double cs = fma(x[q].x, y[q].x, -(x[q].y * y[q].y));
            const double sn = fma(x[q].y, y[q].x, x[q].x * y[q].y);
            const double re = fma(hr, cs, -hi * sn);
            const double im = fma(hr, sn, hi * cs);
            s[0] += re;
            s[1] = fma(-im, uu[q], s[1]);
            s[2] = fma(-im, vv[q], s[2]);
            const double rU = re * uu[q];
            s[3] = fma(-rU, uu[q], s[3]);
            s[4] = fma(-re * vv[q], vv[q], s[4]);
            s[5] = fma(-rU, vv[q], s[5]);
        }
    };
    constexpr unsigned S = kTabU * kTileBS;
    RecCells A, B;
    fetch(A, beg + threadIdx.x);
    if (!stage()) return false;
#pragma unroll
    for (int q = 0; q < kTabU; q++) asm volatile("" : "+v"(A.o[q]), "+v"(A.dr[q]), "+v"(A.di[q]));
    for (unsigned base = beg + threadIdx.x; base < end; base += 2 * S) {
        fetch(B, base + S);
        __builtin_amdgcn_sched_barrier(0);
        eval(A);
        if (base + S >= end) break;
        fetch(A, base + 2 * S);
        __builtin_amdgcn_sched_barrier(0);
        eval(B);
    }
    s[1] *= kTwoPi;
    s[2] *= kTwoPi;
    constexpr double kTwoPi2 = kTwoPi * kTwoPi;
    s[3] *= kTwoPi2;
    s[4] *= kTwoPi2;
    s[5] *= kTwoPi2;
    return true;
}

template <int KIND>
__global__ __launch_bounds__(kTileBS, kTileGridWaves) void lik_fused_wide_kernel(TileLaunch t)
{
    __shared__ double red[(kTileBS / 64) * 8];
    __shared__ int s_go;
    __shared__ double s_off[2];
    __shared__ TabRot<128> R;  // gf <= kWideFacet < 128 runs of kTabRun
    const unsigned per = (t.nwg + kXcds - 1) / kXcds;
    const unsigned wg = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
    if (wg >= t.nwg) return;
    const LikLaunch &a = t.lk;
    const LikItem *items = KIND == 2 ? t.inl : a.items;
    int item;
    unsigned first, nch;
    tile_locate(t, items, wg, &item, &first, &nch);
    const unsigned local = wg - first;
    const unsigned chunk = local / (unsigned)t.nr_g;
    const int r = (int)(local - chunk * (unsigned)t.nr_g) + t.r_lo;
    const FusedPre pre = fused_pre(t, first, nch, item, r);
    const LikItem it = items[item];
    const FacetDev &F = a.facets[it.slot];
    const unsigned beg = chunk * it.chunk_cells;
    const unsigned end = min(F.n, beg + it.chunk_cells);
    const unsigned gf = F.grid_facet;
    // t.tab_lds_facet (rounded up to even) column records, then kWideRows row entries
    extern __shared__ double2 tab_lds[];
    ColEntry *col = reinterpret_cast<ColEntry *>(tab_lds);
    double2 *sY = tab_lds + (size_t)((t.tab_lds_facet + 1) & ~1u) * 3 / 2;
    const uint32_t *cid = F.cid;
    const double *ctr = F.ctr;
    const unsigned row_lo = end > beg ? gload(cid, beg) >> 16 : 0u;
    const unsigned row_hi = end > beg ? gload(cid, end - 1) >> 16 : 0u;
    const unsigned span = row_hi - row_lo + 1;
    const bool cached = span < kWideRows;  // entry span holds row 0
    auto make = [&]() -> bool {
        if (fused_idle(t, chunk, item, r, pre)) return false;
        fused_step(t, first, nch, chunk, item, r, pre, &s_go, s_off);
        // the records' centres do not depend on the step: stored by waves 1.. while wave 0
        // takes it, and the tables read their runs' centres from them
        if (threadIdx.x >= 64)
            for (unsigned k = threadIdx.x - 64; k < gf; k += kTileBS - 64)
                col[k].c = gload(ctr, k);
        __syncthreads();
        if (!s_go) return false;
        const unsigned yb0 = row_lo / kTabRun, yb1 = row_hi / kTabRun;
        tab_rot_stage(R, ColCtr{col}, gf, s_off[0], s_off[1], yb0, yb1, yb0 > 0);
        __syncthreads();
        for (unsigned k = threadIdx.x; k < gf; k += kTileBS) {
            const double2 x = tab_x(R, k);
            col[k].xr = x.x;
            col[k].xi = x.y;
        }
        if (cached)
            for (unsigned q = threadIdx.x; q <= span; q += kTileBS)
                sY[q] = tab_y(R, q < span ? row_lo + q : 0u);
        __syncthreads();
        return true;
    };
    double s[7];
    const bool ran =
        cached ? tile_sums64_col<true>(F, it, a.rg.ro[r], make, col, sY, R, row_lo, span, beg,
                                       end, s)
               : tile_sums64_col<false>(F, it, a.rg.ro[r], make, col, sY, R, row_lo, span, beg,
                                        end, s);
    if (ran) tile_reduce<7>(s, red, t.partials + (size_t)wg * 8);
}

void launch_tile_fused(hipStream_t s, const TileLaunch &t)
{
    const unsigned g = tile_grid(t.nwg, 0);
    if (t.tab_lds_facet > kRecFacet) {
        const unsigned lds =
            ((t.tab_lds_facet + 1) & ~1u) * (unsigned)sizeof(ColEntry) + kWideRows * 16;
        if (t.grid == 2)
            lik_fused_wide_kernel<2><<<g, kTileBS, lds, s>>>(t);
        else
            lik_fused_wide_kernel<1><<<g, kTileBS, lds, s>>>(t);
        return;
    }
    const unsigned lds = t.tab_lds_facet * kRecBytes;
    if (t.grid == 2)
        lik_fused_kernel<2><<<g, kTileBS, lds, s>>>(t);
    else
        lik_fused_kernel<1><<<g, kTileBS, lds, s>>>(t);
}

// Pass 0 on extracted facets with TPW scalelength tasks per work-group (tasks r_lo + TPW g ..
// of one item and chunk; the last group holds what is left): they share each cell's loads,
// centres and shape terms (tile_sums64_p0_db<NT>), and each task's chunk sums go where its
// own work-group of lik_tile_kernel<0, KIND> puts them, with the same values bit for bit.
// Logical work-group wg2 = chunk * groups + g (chunk counted over the launch's items), in the
// XCD-aware order of lik_tile_kernel.
template <int KIND, int NTT>
__device__ inline void p0_group_chunk(const TileLaunch &t, const LikItem *items, int item,
                                      unsigned cg, unsigned chunk, int ra, double *red)
{
    const LikLaunch &a = t.lk;
    const LikItem it = items[item];
    const FacetDev &F = a.facets[it.slot];
    const unsigned beg = chunk * it.chunk_cells;
    const unsigned end = min(F.n, beg + it.chunk_cells);
    extern __shared__ double2 tab_lds[];
    double *sC = reinterpret_cast<double *>(tab_lds);  // tab_lds_facet * 8 bytes
    auto stage = [&]() {
        for (unsigned k = threadIdx.x; k < F.grid_facet; k += kTileBS) sC[k] = gload(F.ctr, k);
        __syncthreads();
    };
    constexpr int NT = NTT;
    double ro[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) ro[q] = a.rg.ro[ra + q];
    double s[NT][7];
    tile_sums64_p0_db<NT>(F, it, ro, sC, stage, beg, end, s);
    const size_t w0 = (size_t)cg * (unsigned)t.nr_g + (unsigned)(ra - t.r_lo);
#pragma unroll
    for (int q = 0; q < NT; q++) tile_reduce<7>(s[q], red, t.partials + (w0 + q) * 8);
}

template <int KIND, int TPW>
__global__ __launch_bounds__(kTileBS, kTileGridWaves) void lik_p0_group_kernel(TileLaunch t)
{
    __shared__ double red[(kTileBS / 64) * 8];
    if (KIND == 2 && blockIdx.x == 0 && threadIdx.x < t.n_inl)
        t.items_w[threadIdx.x] = t.inl[threadIdx.x];
    const LikItem *items = KIND == 2 ? t.inl : t.lk.items;
    const unsigned nr = (unsigned)t.nr_g, groups = (nr + TPW - 1) / TPW;
    const unsigned nwg2 = t.nwg / nr * groups;
    const unsigned per = (nwg2 + kXcds - 1) / kXcds;
    const unsigned nbx = gridDim.x / kXcds;
    for (unsigned j = blockIdx.x / kXcds; j < per; j += nbx) {
        const unsigned wg2 = (blockIdx.x % kXcds) * per + j;
        if (wg2 >= nwg2) break;
        const unsigned cg = wg2 / groups, g = wg2 - cg * groups;
        const int item = t.item_chunks ? (int)(cg / t.item_chunks)
                                       : tile_item(items, t.lk.nitems, cg * nr, (int)nr);
        const unsigned c0 = t.item_chunks ? (unsigned)item * t.item_chunks
                                          : items[item].chunk_first;
        const int ra = t.r_lo + TPW * (int)g;
        const unsigned left = nr - TPW * g;
        if (left >= (unsigned)TPW)
            p0_group_chunk<KIND, TPW>(t, items, item, cg, cg - c0, ra, red);
        else if (TPW == 3 && left == 2)
            p0_group_chunk<KIND, 2>(t, items, item, cg, cg - c0, ra, red);
        else
            p0_group_chunk<KIND, 1>(t, items, item, cg, cg - c0, ra, red);
    }
}

void launch_tile_sums(hipStream_t s, const TileLaunch &t, int pass)
{
    const unsigned g = tile_grid(t.nwg, tuning().tile_blocks);
    if (pass == 0) {
        const unsigned lds = t.tab_lds_facet * 8;
        // the grid kernels' pass 0 with RL2_TUNE_P0_TASKS tasks per work-group (default 3)
        const unsigned tpw = (unsigned)tuning().p0_tasks;
        const unsigned g2 = tile_grid(
            t.nwg / (unsigned)t.nr_g * (((unsigned)t.nr_g + tpw - 1) / tpw), tuning().tile_blocks);
        if (t.grid && tpw == 3) {
            if (t.grid == 2)
                lik_p0_group_kernel<2, 3><<<g2, kTileBS, lds, s>>>(t);
            else
                lik_p0_group_kernel<1, 3><<<g2, kTileBS, lds, s>>>(t);
        } else if (t.grid && tpw == 2) {
            if (t.grid == 2)
                lik_p0_group_kernel<2, 2><<<g2, kTileBS, lds, s>>>(t);
            else
                lik_p0_group_kernel<1, 2><<<g2, kTileBS, lds, s>>>(t);
        } else if (t.grid == 2)
            lik_tile_kernel<0, 2><<<g, kTileBS, lds, s>>>(t);
        else if (t.grid == 1)
            lik_tile_kernel<0, 1><<<g, kTileBS, lds, s>>>(t);
        else
            lik_tile_kernel<0, 0><<<g, kTileBS, lds, s>>>(t);
    } else {
        const unsigned lds = t.tab ? t.tab_lds_facet * (t.tab_y_lds ? 40 : 24) : 0;
        if (t.grid == 2)
            lik_tile_kernel<1, 2><<<g, kTileBS, lds, s>>>(t);
        else if (t.grid == 1)
            lik_tile_kernel<1, 1><<<g, kTileBS, lds, s>>>(t);
        else
            lik_tile_kernel<1, 0><<<g, kTileBS, lds, s>>>(t);
    }
}

void launch_tile_step(hipStream_t s, const TileLaunch &t, int pass)
{
    const unsigned ntask = (unsigned)t.lk.nitems * t.nr_g;
    // the step is wave 0's; the other waves only write the next pass's tables, so a step
    // without tables (the one after an evaluation's last fused pass) runs one wave
    const unsigned bs = t.tab ? kStepBS : 64u;
    if (pass == 0)
        lik_step_kernel<0><<<ntask, bs, 0, s>>>(t);
    else
        lik_step_kernel<1><<<ntask, bs, 0, s>>>(t);
}

void launch_tile_pass(hipStream_t s, const TileLaunch &t, int pass)
{
    launch_tile_sums(s, t, pass);
    launch_tile_step(s, t, pass);
}

void launch_rmarg(hipStream_t s, const LikLaunch &a)
{
    if (a.nitems <= 0) return;
    rmarg_kernel<<<a.nitems + (a.status ? 1 : 0), 128, 0, s>>>(a);
}

__global__ void marginalise_r_kernel(int num, const double *x, const double *y, double *out,
                                     double *ws, int *iws)
{
    double xx[kMaxR], yy[kMaxR];
    for (int i = 0; i < num; i++) {
        xx[i] = x[i];
        yy[i] = y[i];
    }
    QagWs w{ws, ws + kQagLimit, ws + 2 * kQagLimit, ws + 3 * kQagLimit, iws, iws + kQagLimit};
    out[0] = marginalise_posterior(num, xx, yy, w);
    for (int i = 0; i < num; i++) out[1 + i] = yy[i];
}

void launch_marginalise_r(hipStream_t s, int num, const double *x, const double *y,
                          double *out, double *qag_ws, int *qag_iws)
{
    marginalise_r_kernel<<<1, 1, 0, s>>>(num, x, y, out, qag_ws, qag_iws);
}

// marg_position (likelihood.c:376-421) on n given arguments, one per thread: the device
// function finish_lr applies to B = ho^2 / 2.
__global__ void marg_position_kernel(int n, const double *x, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = marg_position(x[i]);
}

void launch_marg_position(hipStream_t s, int n, const double *x, double *out)
{
    if (n <= 0) return;
    marg_position_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, x, out);
}

// ------------------------------------------------------------------ subtraction (K2)
// data_processing.c:109-135: visSkyMod -= round model, visData -= fitted model, i.e.
// R = (R - fit) + round with one phase rotor per visibility.
template <int PREC>
__global__ __launch_bounds__(256) void subtract_kernel(unsigned nch, unsigned long nrow,
                                                       const double *__restrict__ u,
                                                       const double *__restrict__ v,
                                                       const double *__restrict__ w,
                                                       const double *__restrict__ wn,
                                                       const double *__restrict__ inv_lambda,
                                                       const double *__restrict__ spec,
                                                       GalaxyGeom fit, GalaxyGeom rnd, void *res)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    const double ui = u[i], vi = v[i], wi = w[i];
    const double p = ui * fit.l0 + vi * fit.m0 + wi * fit.n0;
    for (unsigned c = 0; c < nch; c++) {
        const double k = wn[c], k2 = k * k, spectra = spec[c];
        const size_t idx = (size_t)c * nrow + i;
        const double2 r = subtract_vis(fit, rnd, p, ui, vi, k2, spectra, inv_lambda[c],
                                       load_res(res, PREC, idx));
        store_res(res, PREC, idx, r);
    }
}

void launch_subtract(hipStream_t s, int prec, unsigned nch, unsigned long nrow, const double *u,
                     const double *v, const double *w, const double *wn,
                     const double *inv_lambda, const double *spec,
                     const GalaxyGeom &fit, const GalaxyGeom &round, void *res)
{
    dim3 grid((unsigned)((nrow + 255) / 256));
    if (prec == 0)
        subtract_kernel<0><<<grid, 256, 0, s>>>(nch, nrow, u, v, w, wn, inv_lambda, spec, fit,
                                                 round, res);
    else
        subtract_kernel<1><<<grid, 256, 0, s>>>(nch, nrow, u, v, w, wn, inv_lambda, spec, fit,
                                                 round, res);
}

// ------------------------------------------------------------------ drop-in helpers
// galaxy_visibilities.c:59-91 in two steps: model + per-block partial sums of M^2/sigma^2,
// then normalisation by sqrt of the fixed-order total.
__global__ __launch_bounds__(256) void model_zero_kernel(double e1, double e2, double sf, unsigned n,
                                                         const double *uu, const double *vv,
                                                         const double *s2, double *mod,
                                                         double *partials)
{
    __shared__ double red[4];
    unsigned i = blockIdx.x * 256 + threadIdx.x;
    double t = 0.;
    if (i < n) {
        double k1 = (1. + e1) * uu[i] + e2 * vv[i];
        double k2 = e2 * uu[i] + (1. - e1) * vv[i];
        double M = sersic_shape(k1, k2, sf, 4 * kPi * kPi);
        mod[i] = M;
        t = M * M / s2[i];
    }
    double v[1] = {t};
    block_allreduce<256, 1>(v, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}
__global__ __launch_bounds__(256) void model_norm_kernel(unsigned n, unsigned nb,
                                                         const double *partials, double *mod)
{
    __shared__ double tot;
    if (threadIdx.x == 0) {
        double s = 0.;
        for (unsigned b = 0; b < nb; b++) s += partials[b];
        tot = sqrt(s);
    }
    __syncthreads();
    for (unsigned i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) mod[i] /= tot;
}
void launch_model_at_zero(hipStream_t s, double e1, double e2, double scale, unsigned n,
                          const double *uu, const double *vv, const double *s2, double *mod,
                          double *partials)
{
    double detA = 1. - e1 * e1 - e2 * e2;
    double sc = arcs2rad_scale(scale);
    double sf = (sc * sc) / (detA * detA);
    unsigned nb = (n + 255) / 256;
    if (nb == 0) return;
    model_zero_kernel<<<nb, 256, 0, s>>>(e1, e2, sf, n, uu, vv, s2, mod, partials);
    model_norm_kernel<<<min(nb, 1024u), 256, 0, s>>>(n, nb, partials, mod);
}

__global__ void galaxy_vis_kernel(double spectra, double wn, GalaxyGeom g, unsigned long n,
                                  const double *u, const double *v, const double *w, double *vis)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ui = u[i], vi = v[i], wi = w[i];
    double phase = ui * g.l0 + vi * g.m0 + wi * g.n0;
    phase = wn * phase;
    double shape = model_shape(g, ui, vi, wn * wn, spectra);
    double sn, cs;
    sincos(phase, &sn, &cs);
    vis[2 * i] = shape * cs;
    vis[2 * i + 1] = shape * sn;
}
void launch_galaxy_vis(hipStream_t s, double spectra, double wavenumber, const GalaxyGeom &g,
                       unsigned long n, const double *u, const double *v, const double *w,
                       double *vis)
{
    if (!n) return;
    galaxy_vis_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(spectra, wavenumber, g, n, u, v,
                                                                  w, vis);
}

__global__ void phase_shift_kernel(double wn, double l, double m, double nn, unsigned long n,
                                   const double *u, const double *v, const double *w, double *vis)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double phase = u[i] * l + v[i] * m + w[i] * nn;
    phase *= -wn;
    double sp, cp;
    sincos(phase, &sp, &cp);
    double re = vis[2 * i], im = vis[2 * i + 1];
    vis[2 * i] = re * cp - im * sp;
    vis[2 * i + 1] = re * sp + im * cp;
}
void launch_phase_shift(hipStream_t s, double wavenumber, double l, double m, unsigned long n,
                        const double *u, const double *v, const double *w, double *vis)
{
    if (!n) return;
    double nn = sqrt(1. - l * l - m * m) - 1.;
    phase_shift_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(wavenumber, l, m, nn, n, u, v,
                                                                   w, vis);
}

// ------------------------------------------------------------------ observation set-up
__global__ void gather_rows_kernel(unsigned long nrow, const uint32_t *perm, const double *src,
                                   double *dst)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i < nrow) dst[i] = src[perm[i]];
}
void launch_gather_rows(hipStream_t s, unsigned long nrow, const uint32_t *perm,
                        const double *src, double *dst)
{
    gather_rows_kernel<<<(unsigned)((nrow + 255) / 256), 256, 0, s>>>(nrow, perm, src, dst);
}

__global__ void pack_sigma_kernel(unsigned nch, unsigned long nrow, const uint32_t *perm,
                                  const float *sigma2, const unsigned char *flag, float *s2f)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    unsigned long src = perm[i];
    for (unsigned c = 0; c < nch; c++) {
        float sg = sigma2[(size_t)c * nrow + src];
        bool fl = flag && flag[(size_t)c * nrow + src];
        s2f[(size_t)c * nrow + i] = fl ? -fabsf(sg) - 0.0f : sg;
    }
}
void launch_pack_sigma(hipStream_t s, unsigned nch, unsigned long nrow, const uint32_t *perm,
                       const float *sigma2, const unsigned char *flag, float *s2f)
{
    pack_sigma_kernel<<<(unsigned)((nrow + 255) / 256), 256, 0, s>>>(nch, nrow, perm, sigma2, flag,
                                                                     s2f);
}

template <int PREC>
__global__ void residual_in_kernel(unsigned nch, unsigned long nrow, const uint32_t *perm,
                                   const double *D, const double *S, void *res)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    unsigned long src = perm[i];
    for (unsigned c = 0; c < nch; c++) {
        size_t k = (size_t)c * nrow + src;
        double re = D[2 * k], im = D[2 * k + 1];
        if (S) {
            re = re - S[2 * k];
            im = im - S[2 * k + 1];
        }
        store_res(res, PREC, (size_t)c * nrow + i, make_double2(re, im));
    }
}
void launch_residual_in(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                        const uint32_t *perm, const double *D, const double *S, void *res)
{
    dim3 grid((unsigned)((nrow + 255) / 256));
    if (prec == 0)
        residual_in_kernel<0><<<grid, 256, 0, s>>>(nch, nrow, perm, D, S, res);
    else
        residual_in_kernel<1><<<grid, 256, 0, s>>>(nch, nrow, perm, D, S, res);
}

template <int PREC>
__global__ void residual_out_kernel(unsigned nch, unsigned long nrow, const uint32_t *perm,
                                    const void *res, double *out)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    unsigned long dst = perm[i];
    for (unsigned c = 0; c < nch; c++) {
        double2 r = load_res(res, PREC, (size_t)c * nrow + i);
        size_t k = (size_t)c * nrow + dst;
        out[2 * k] = r.x;
        out[2 * k + 1] = r.y;
    }
}
void launch_residual_out(hipStream_t s, int prec, unsigned nch, unsigned long nrow,
                         const uint32_t *perm, const void *res, double *out)
{
    dim3 grid((unsigned)((nrow + 255) / 256));
    if (prec == 0)
        residual_out_kernel<0><<<grid, 256, 0, s>>>(nch, nrow, perm, res, out);
    else
        residual_out_kernel<1><<<grid, 256, 0, s>>>(nch, nrow, perm, res, out);
}

// Morton (Z-order) key of (u, v) quantised to 16 bits per axis over [-umax, umax].
__device__ inline uint32_t spread16(uint32_t x)
{
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__global__ void morton_kernel(unsigned long nrow, const double *u, const double *v, double umax,
                              uint32_t *keys, uint32_t *idx)
{
    unsigned long i = blockIdx.x * (unsigned long)blockDim.x + threadIdx.x;
    if (i >= nrow) return;
    double s = 65535.0 / (2.0 * umax);
    double qu = fmin(fmax((u[i] + umax) * s, 0.0), 65535.0);
    double qv = fmin(fmax((v[i] + umax) * s, 0.0), 65535.0);
    keys[i] = spread16((uint32_t)qu) | (spread16((uint32_t)qv) << 1);
    idx[i] = (uint32_t)i;
}
void launch_morton_keys(hipStream_t s, unsigned long nrow, const double *u, const double *v,
                        double umax, uint32_t *keys, uint32_t *idx)
{
    morton_kernel<<<(unsigned)((nrow + 255) / 256), 256, 0, s>>>(nrow, u, v, umax, keys, idx);
}

size_t sort_rows_tmp_bytes(unsigned long nrow)
{
    size_t bytes = 0;
    hipcub::DoubleBuffer<uint32_t> k(nullptr, nullptr), v(nullptr, nullptr);
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, (int)nrow) != hipSuccess)
        return 0;  // the sort itself then reports the failure
    return bytes;
}

bool sort_rows(hipStream_t s, unsigned long nrow, uint32_t *keys, uint32_t *idx,
               uint32_t *keys_alt, uint32_t *idx_alt, void *tmp, size_t tmp_bytes,
               uint32_t **sorted_idx)
{
    hipcub::DoubleBuffer<uint32_t> k(keys, keys_alt), v(idx, idx_alt);
    const bool ok = hip_ok(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k, v, (int)nrow,
                                                              0, 32, s),
                           "row sort");
    *sorted_idx = v.Current();
    return ok;
}

}  // namespace rl2
