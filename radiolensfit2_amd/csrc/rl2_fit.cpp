// rl2_fit.cpp -- per-galaxy fit control flow (galaxy_fitting.c:41-117, likelihood.c:87-99
// and :428-554, GSL 2.5 multimin/simplex2.c) as C++20 coroutines.
//
// Each galaxy's fit is written as the reference's straight-line algorithm; wherever the
// reference calls loglikelihood() the coroutine suspends with the ellipticities it needs.
// The scheduler gathers the requests of every suspended galaxy into one batched GPU launch
// (one work-group per (request, scalelength)), then resumes them.  Two kinds of speculation
// evaluate points the sequential algorithm may or may not ask for next; the replay uses the
// values exactly where the reference would have computed them, so the trajectory is the
// sequential one:
//   - simplex: the reflected, expanded and both contracted candidates of an iteration;
//   - line searches: `line_chunk` steps ahead along the search direction.
#include <coroutine>
#include <exception>
#include <math.h>
#include <stdio.h>

#include "rl2_obs.h"

namespace rl2 {
namespace {

struct Pt {
    double e1, e2;
    double L;
    int err;
    int frag = 0;  // mixed-precision value sits at a branch FP32 sums can flip
};

struct Task {
    struct promise_type {
        Task get_return_object()
        {
            return Task{std::coroutine_handle<promise_type>::from_promise(*this)};
        }
        std::suspend_always initial_suspend() noexcept { return {}; }
        std::suspend_always final_suspend() noexcept { return {}; }
        void return_void() {}
        void unhandled_exception() { std::terminate(); }
    };
    std::coroutine_handle<promise_type> h;
    explicit Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Task(Task &&o) noexcept : h(o.h) { o.h = nullptr; }
    Task(const Task &) = delete;
    ~Task()
    {
        if (h) h.destroy();
    }
};

struct Ctx {
    int slot = 0;
    rl2_gal_result *res = nullptr;
    std::vector<Pt> *pending = nullptr;
    bool speculate = true;
    int line_chunk = 4;
    double start_e1 = 0., start_e2 = 0., step = 0.1;
    bool sampling_only = false;
    int np_max = kNpMax;
    int n_evals = 0;
    int n_verified = 0;   // FP64 re-evaluations requested by the verification
    int prec_nm = 0, prec_samp = 0;  // likelihood precision of the simplex / sampling phase
    bool verify = false;  // re-evaluate in FP64 the mixed values near a sampling threshold
    bool has_prop = false;  // learned proposal for the simplex start
    double prop_e1 = 0., prop_e2 = 0., prop_step = 0.1;
    int start = 0;          // 1 when the simplex started from the proposal
    int rounds = 0;         // batched launches this fit waited for
    int cur_prec = 0;                // precision of the pending requests
};

struct Await {
    Ctx *c;
    std::vector<Pt> *v;
    bool await_ready() const noexcept { return v->empty(); }
    void await_suspend(std::coroutine_handle<>) noexcept { c->pending = v; }
    void await_resume() const noexcept {}
};

// cblas dnrm2 (scale / sum-of-squares recurrence), as used by simplex2.c
double dnrm2(const double *x, int n)
{
    double scale = 0.0, ssq = 1.0;
    if (n == 1) return fabs(x[0]);
    for (int i = 0; i < n; i++) {
        if (x[i] != 0.0) {
            const double ax = fabs(x[i]);
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

// nmsimplex2 state for n = 2 (three corners), with GSL's incremental centre / size update.
struct Simplex {
    static constexpr int N = 2, P = 3;
    double x1[P][N];
    double y1[P];
    double center[N];
    double S2;

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
            for (int j = 0; j < N; j++) s[j] = x1[i][j] + -1.0 * center[j];
            const double t = dnrm2(s, N);
            ss += t * t;
        }
        S2 = ss / P;
        return sqrt(ss / P);
    }
    // xc = alpha * centre + beta * x_corner (try_corner_move)
    void corner(double coeff, int c, double *xc) const
    {
        const double alpha = (1 - coeff) * P / (P - 1.0);
        const double beta = (P * coeff - 1.0) / (P - 1.0);
        for (int j = 0; j < N; j++) xc[j] = center[j];
        for (int j = 0; j < N; j++) xc[j] *= alpha;
        for (int j = 0; j < N; j++) xc[j] += beta * x1[c][j];
    }
    void update_point(int i, const double *x, double val)
    {
        double delta[N], xmc[N];
        for (int j = 0; j < N; j++) delta[j] = x[j] + -1.0 * x1[i][j];
        for (int j = 0; j < N; j++) xmc[j] = x1[i][j] + -1.0 * center[j];
        const double d = dnrm2(delta, N);
        double xmcd = 0.0;
        for (int j = 0; j < N; j++) xmcd += xmc[j] * delta[j];
        S2 += (2.0 / P) * xmcd + ((P - 1.0) / P) * (d * d / P);
        const double alpha = 1.0 / P;
        for (int j = 0; j < N; j++) center[j] += -alpha * x1[i][j];
        for (int j = 0; j < N; j++) center[j] += alpha * x[j];
        for (int j = 0; j < N; j++) x1[i][j] = x[j];
        y1[i] = val;
    }
};

Task fit_task(Ctx *c)
{
    rl2_gal_result &R = *c->res;
    std::vector<Pt> req;
    req.reserve(256);
    constexpr int kNone = -2;  // candidate not requested
    // f_likelihood (likelihood.c:87-99): points with |e|^2 >= 0.65 are 1e10 without a call
    auto add_f = [&](const double *x) -> int {
        if (x[0] * x[0] + x[1] * x[1] < 0.65) {
            req.push_back({x[0], x[1], 0., 0});
            return (int)req.size() - 1;
        }
        return -1;
    };
    auto fv = [&](int k) -> double { return k < 0 ? 1.e+10 : -req[k].L; };
    int nref = 0;  // loglikelihood calls of the sequential algorithm

    double mes_e1, mes_e2, maxL;
    if (c->sampling_only) {
        mes_e1 = R.e1;
        mes_e2 = R.e2;
        maxL = R.maxL;
        goto sampling;
    }
    c->cur_prec = c->prec_nm;
    {
    // ---- gsl_multimin_fminimizer_set (nmsimplex_set) ----
    Simplex s;
    double x0[2] = {c->start_e1, c->start_e2};
    double step = c->step;
    double xs[2][2] = {{x0[0] + step, x0[1]}, {x0[0], x0[1] + step}};
    req.clear();
    int k0 = add_f(x0), ka = add_f(xs[0]), kb = add_f(xs[1]);
    int kp0 = -1, kpa = -1, kpb = -1;
    double xp[2] = {c->prop_e1, c->prop_e2};
    double xps[2][2] = {{xp[0] + c->prop_step, xp[1]}, {xp[0], xp[1] + c->prop_step}};
    if (c->has_prop) {
        // the proposal's start triangle in the same launch; it replaces the reference's
        // start (galaxy_fitting.c:59-67) only when its likelihood is higher
        kp0 = add_f(xp);
        kpa = add_f(xps[0]);
        kpb = add_f(xps[1]);
    }
    co_await Await{c, &req};
    nref += (k0 >= 0) + (ka >= 0) + (kb >= 0) + (kp0 >= 0) + (kpa >= 0) + (kpb >= 0);
    if (c->has_prop && kp0 >= 0 && isfinite(fv(kp0)) && fv(kp0) < fv(k0)) {
        c->start = 1;
        x0[0] = xp[0];
        x0[1] = xp[1];
        step = c->prop_step;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) xs[i][j] = xps[i][j];
        k0 = kp0;
        ka = kpa;
        kb = kpb;
    }
    for (int j = 0; j < 2; j++) {
        s.x1[0][j] = x0[j];
        s.x1[1][j] = xs[0][j];
        s.x1[2][j] = xs[1][j];
    }
    s.y1[0] = fv(k0);
    s.y1[1] = fv(ka);
    s.y1[2] = fv(kb);
    s.compute_center();
    double size = s.compute_size();

    // ---- iterate loop (galaxy_fitting.c:72-82) ----
    int iter = 0, status;
    double x[2] = {x0[0], x0[1]}, fval = 0.;
    do {
        iter++;
        int hi = 0, s_hi = 1, lo = 0;
        double dhi = s.y1[0], dlo = s.y1[0], ds_hi = s.y1[1];
        for (int i = 1; i < 3; i++) {
            const double val = s.y1[i];
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
        double xr[2], xe[2], xci[2], xcu[2];
        s.corner(-1.0, hi, xr);
        s.corner(-2.0, hi, xe);
        s.corner(0.5, hi, xci);  // contraction when the reflected point is rejected
        {
            Simplex su = s;  // contraction after the reflected point replaced the worst
            su.update_point(hi, xr, 0.);
            su.corner(0.5, hi, xcu);
        }
        req.clear();
        int ir = add_f(xr), ie = kNone, icu = kNone, ici = kNone;
        if (c->speculate) {
            ie = add_f(xe);
            icu = add_f(xcu);
            ici = add_f(xci);
        }
        co_await Await{c, &req};
        nref += (ir >= 0);
        int st = 0;
        const double val = fv(ir);
        if (isfinite(val) && val < s.y1[lo]) {
            if (ie == kNone) {
                req.clear();
                ie = add_f(xe);
                co_await Await{c, &req};
            }
            nref += (ie >= 0);
            const double val2 = fv(ie);
            if (isfinite(val2) && val2 < s.y1[lo])
                s.update_point(hi, xe, val2);
            else
                s.update_point(hi, xr, val);
        } else if (!isfinite(val) || val > s.y1[s_hi]) {
            const double *xc = xci;
            int ic = ici;
            if (isfinite(val) && val <= s.y1[hi]) {
                s.update_point(hi, xr, val);
                xc = xcu;
                ic = icu;
            }
            if (ic == kNone) {
                req.clear();
                ic = add_f(xc);
                co_await Await{c, &req};
            }
            nref += (ic >= 0);
            const double val2 = fv(ic);
            if (isfinite(val2) && val2 <= s.y1[hi]) {
                s.update_point(hi, xc, val2);
            } else {
                // contract_by_best: every corner but the best moves half way to it
                int idx[3] = {-3, -3, -3};
                req.clear();
                for (int i = 0; i < 3; i++)
                    if (i != lo) {
                        for (int j = 0; j < 2; j++) s.x1[i][j] = 0.5 * (s.x1[i][j] + s.x1[lo][j]);
                        idx[i] = add_f(s.x1[i]);
                    }
                co_await Await{c, &req};
                for (int i = 0; i < 3; i++)
                    if (i != lo) {
                        nref += (idx[i] >= 0);
                        s.y1[i] = fv(idx[i]);
                        if (!isfinite(s.y1[i])) st = 1;
                    }
                s.compute_center();
                s.compute_size();
            }
        } else {
            s.update_point(hi, xr, val);
        }
        if (st) break;  // GSL_EFAILED from the contraction
        lo = 0;
        double mn = s.y1[0];
        for (int i = 0; i < 3; i++) {
            if (s.y1[i] < mn) {
                mn = s.y1[i];
                lo = i;
            }
            if (isnan(s.y1[i])) {
                lo = i;
                break;
            }
        }
        x[0] = s.x1[lo][0];
        x[1] = s.x1[lo][1];
        fval = s.y1[lo];
        size = (s.S2 > 0) ? sqrt(s.S2) : s.compute_size();
        status = (size < kTol) ? 0 : -2;  // gsl_multimin_test_size
    } while (status == -2 && iter < 50 && fval < 0.);

    mes_e1 = x[0];
    mes_e2 = x[1];
    maxL = -fval;
    if (mes_e1 * mes_e1 + mes_e2 * mes_e2 > 0.65) maxL = -1e+10;
    }
sampling:
    const double nm_e1 = mes_e1, nm_e2 = mes_e2;
    double var_e1 = 0., var_e2 = 0., oneDimvar = 0.;
    int error = 0;
    // With mixed-precision sampling and verification, a galaxy whose variance test
    // (galaxy_fitting.c:100) lands within 2% of its threshold is sampled again in FP64, so
    // the good/bad decision (and with it the residual update) is the FP64 one.
    for (int attempt = 0; attempt < 2; attempt++) {
    mes_e1 = nm_e1;
    mes_e2 = nm_e2;
    c->cur_prec = attempt == 0 ? c->prec_samp : 0;
    const int samp_prec = c->cur_prec;
    const bool check_var = attempt == 0 && c->verify && samp_prec != 0;
    var_e1 = var_e2 = oneDimvar = 0.;
    error = 0;
    if (maxL > -1.e+10 || c->sampling_only) {
        // ---- likelihood_sampling (likelihood.c:428-554) ----
        int lerr = 0, np;
        double totL, average_1, average_2, cov, var1, var2;
        double mod2 = 0;
        double L = 1.;
        const double threshold = 0.05;
        // FP64 re-evaluation of the mixed-precision values whose acceptance test
        // L / Lmax >= threshold is closer to its boundary than the FP32 error bound
        // (measured |L32 - L64| <= 1.3e-7 |L| on the full-size workload; bound 5e-7 |L|),
        // and of points whose error flag decides a line search.
        std::vector<Pt> ver;
        std::vector<size_t> vidx;
        const double log_thr = log(threshold);
        const double bound = 5e-7 * fabs(maxL) + 1e-5;
        const double max_e1 = mes_e1, max_e2 = mes_e2;
        average_1 = max_e1;
        average_2 = max_e2;
        var1 = max_e1 * max_e1;
        var2 = max_e2 * max_e2;
        cov = max_e1 * max_e2;
        totL = 1.;
        np = 1;
        const int K = c->line_chunk > 0 ? c->line_chunk : 1;
        float sampling = 0.05;
        if (max_e1 < 0) sampling = -0.05;
        int k1 = 0;
        double x_e1 = max_e1;
        while (L > threshold && mod2 < 1. && !lerr) {
            req.clear();
            double xs1 = x_e1;
            for (int j = 0; j < K; j++) {
                xs1 += sampling;
                if (xs1 * xs1 + max_e2 * max_e2 < 1.)
                    req.push_back({xs1, max_e2, 0., 0});
                else
                    break;
            }
            co_await Await{c, &req};
            if (c->verify) {
                ver.clear();
                vidx.clear();
                for (size_t j = 0; j < req.size(); j++)
                    if (req[j].err || req[j].frag || fabs(req[j].L - maxL - log_thr) < bound) {
                        ver.push_back(req[j]);
                        vidx.push_back(j);
                    }
                if (!ver.empty()) {
                    c->cur_prec = 0;
                    co_await Await{c, &ver};
                    c->cur_prec = samp_prec;
                    c->n_verified += (int)ver.size();
                    for (size_t j = 0; j < ver.size(); j++) {
                        req[vidx[j]].L = ver[j].L;
                        req[vidx[j]].err = ver[j].err;
                        req[vidx[j]].frag = 0;
                    }
                }
            }
            size_t j = 0;
            do {
                x_e1 += sampling;
                mod2 = x_e1 * x_e1 + max_e2 * max_e2;
                if (mod2 < 1.) {
                    lerr = req[j].err;
                    const double L_e = req[j].L;
                    j++;
                    nref++;
                    if (!lerr) {
                        L = exp(L_e - maxL);
                        k1++;
                    }
                }
            } while (L > threshold && mod2 < 1. && !lerr && j < req.size());
        }
        if (L < threshold && k1 > 1) k1--;
        L = 1.;
        int k2 = 0;
        sampling = 0.05;
        if (max_e2 < 0) sampling = -0.05;
        double x_e2 = max_e2;
        while (L > threshold && mod2 < 1. && !lerr) {
            req.clear();
            double xs2 = x_e2;
            for (int j = 0; j < K; j++) {
                xs2 += sampling;
                if (xs2 * xs2 + max_e1 * max_e1 < 1.)
                    req.push_back({max_e1, xs2, 0., 0});
                else
                    break;
            }
            co_await Await{c, &req};
            if (c->verify) {
                ver.clear();
                vidx.clear();
                for (size_t j = 0; j < req.size(); j++)
                    if (req[j].err || req[j].frag || fabs(req[j].L - maxL - log_thr) < bound) {
                        ver.push_back(req[j]);
                        vidx.push_back(j);
                    }
                if (!ver.empty()) {
                    c->cur_prec = 0;
                    co_await Await{c, &ver};
                    c->cur_prec = samp_prec;
                    c->n_verified += (int)ver.size();
                    for (size_t j = 0; j < ver.size(); j++) {
                        req[vidx[j]].L = ver[j].L;
                        req[vidx[j]].err = ver[j].err;
                        req[vidx[j]].frag = 0;
                    }
                }
            }
            size_t j = 0;
            do {
                x_e2 += sampling;
                mod2 = x_e2 * x_e2 + max_e1 * max_e1;
                if (mod2 < 1.) {
                    lerr = req[j].err;
                    const double L_e = req[j].L;
                    j++;
                    nref++;
                    if (!lerr) {
                        L = exp(L_e - maxL);
                        k2++;
                    }
                }
            } while (L > threshold && mod2 < 1. && !lerr && j < req.size());
        }
        if (L < threshold && k2 > 1) k2--;
        sampling = 0.05;
        const float edim1 = k1 * sampling;
        const float edim2 = k2 * sampling;
        float start1 = edim1, start2 = edim2;
        while (np < c->np_max && sampling > kEres) {
            req.clear();
            for (int ie1 = (int)(-start1 / sampling); (float)ie1 <= start1 / sampling; ie1 += 2)
                for (int ie2 = (int)(-start2 / sampling); (float)ie2 <= start2 / sampling;
                     ie2 += 2)
                    if (ie1 != 0 || ie2 != 0) {
                        const double xe1 = max_e1 + (float)(ie1 * sampling);
                        const double xe2 = max_e2 + (float)(ie2 * sampling);
                        if (xe1 * xe1 + xe2 * xe2 < 1.) req.push_back({xe1, xe2, 0., 0});
                    }
            co_await Await{c, &req};
            if (c->verify) {
                ver.clear();
                vidx.clear();
                for (size_t j = 0; j < req.size(); j++)
                    if (req[j].err || req[j].frag || fabs(req[j].L - maxL - log_thr) < bound) {
                        ver.push_back(req[j]);
                        vidx.push_back(j);
                    }
                if (!ver.empty()) {
                    c->cur_prec = 0;
                    co_await Await{c, &ver};
                    c->cur_prec = samp_prec;
                    c->n_verified += (int)ver.size();
                    for (size_t j = 0; j < ver.size(); j++) {
                        req[vidx[j]].L = ver[j].L;
                        req[vidx[j]].err = ver[j].err;
                        req[vidx[j]].frag = 0;
                    }
                }
            }
            for (const Pt &p : req) {
                nref++;
                lerr = p.err;
                if (!lerr) {
                    L = exp(p.L - maxL);
                    if (L >= threshold && !isinf(L)) {
                        np++;
                        const double xL1 = p.e1 * L, xL2 = p.e2 * L;
                        average_1 += xL1;
                        average_2 += xL2;
                        var1 += p.e1 * xL1;
                        var2 += p.e2 * xL2;
                        cov += p.e1 * xL2;
                        totL += L;
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
        mes_e1 = average_1;
        mes_e2 = average_2;
        var_e1 = var1;
        var_e2 = var2;
        oneDimvar = sqrt((var1) * (var2)-cov * cov);
        if (!c->sampling_only && (var_e1 < kVar || var_e2 < kVar || oneDimvar < kVar)) error = 1;
        if (check_var) {
            auto near_var = [](double x) { return fabs(x / kVar - 1.) < 0.02; };
            if (near_var(var_e1) || near_var(var_e2) || near_var(oneDimvar)) continue;
        }
    } else {
        error = 1;
    }
    break;
    }
    R.e1 = mes_e1;
    R.e2 = mes_e2;
    R.var_e1 = var_e1;
    R.var_e2 = var_e2;
    R.oneDimvar = oneDimvar;
    R.maxL = maxL;
    R.error = error;
    R.n_evals = nref;
    R.start = c->start;
    R.rounds = c->rounds;
    co_return;
}

}  // namespace

bool fit_batch(const std::vector<FitRequest> &reqs, const rl2_fit_options &opt,
               const BatchEval &eval)
{
    const size_t nb = reqs.size();
    std::vector<Ctx> ctx(nb);
    std::vector<Task> tasks;
    tasks.reserve(nb);
    for (size_t i = 0; i < nb; i++) {
        ctx[i].slot = reqs[i].slot;
        ctx[i].res = reqs[i].res;
        ctx[i].speculate = opt.nm_speculate != 0;
        ctx[i].line_chunk = opt.line_chunk > 0 ? opt.line_chunk : 4;
        ctx[i].sampling_only = reqs[i].sampling_only;
        ctx[i].np_max = reqs[i].np_max;
        // lik_precision: 0 FP64; 1 mixed; 2 mixed simplex + FP64 sampling; 3 FP64 simplex
        // + mixed sampling; 4 as 3 with FP64 verification of near-threshold samples
        const int lp = opt.lik_precision;
        ctx[i].prec_nm = (lp == 1 || lp == 2) ? 1 : 0;
        ctx[i].prec_samp = (lp == 1 || lp == 3 || lp == 4) ? 1 : 0;
        ctx[i].verify = lp == 4;
        ctx[i].has_prop = reqs[i].has_prop;
        ctx[i].prop_e1 = reqs[i].prop_e1;
        ctx[i].prop_e2 = reqs[i].prop_e2;
        ctx[i].prop_step = reqs[i].prop_step;
        if (reqs[i].prop_chunk > 0) ctx[i].line_chunk = reqs[i].prop_chunk;
        tasks.push_back(fit_task(&ctx[i]));
    }
    for (auto &t : tasks) t.h.resume();
    std::vector<LikItem> items;
    std::vector<double> Le;
    std::vector<int> err;
    while (true) {
        items.clear();
        for (size_t i = 0; i < nb; i++)
            if (!tasks[i].h.done() && ctx[i].pending)
                for (const Pt &p : *ctx[i].pending)
                    items.push_back({ctx[i].slot, ctx[i].cur_prec, p.e1, p.e2});
        if (items.empty()) break;
        Le.assign(items.size(), 0.);
        err.assign(items.size(), 0);
        if (!eval(items, Le, err)) return false;
        size_t k = 0;
        for (size_t i = 0; i < nb; i++) {
            if (tasks[i].h.done() || !ctx[i].pending) continue;
            for (Pt &p : *ctx[i].pending) {
                p.L = Le[k];
                p.err = err[k] & ~kErrFragile;
                p.frag = (err[k] & kErrFragile) != 0;
                k++;
            }
            ctx[i].pending = nullptr;
            ctx[i].rounds++;
            tasks[i].h.resume();
        }
    }
    return true;
}

}  // namespace rl2

// ---- test hook: the same control flow driven by a caller-supplied log-likelihood ----
// Lets the host logic be pinned against the oracle on analytic likelihoods without a GPU.
extern "C" int rl2_fit_with_callback(double (*loglik)(double, double, int *, void *), void *ctx,
                                     int nm_speculate, int line_chunk, rl2_gal_result *res)
{
    rl2_fit_options opt = {};
    opt.nm_speculate = nm_speculate;
    opt.line_chunk = line_chunk;
    std::vector<rl2::FitRequest> reqs{{0, res}};
    int launches = 0;
    bool ok = rl2::fit_batch(reqs, opt,
                             [&](const std::vector<rl2::LikItem> &items, std::vector<double> &Le,
                                 std::vector<int> &err) {
                                 launches++;
                                 for (size_t i = 0; i < items.size(); i++) {
                                     int e = 0;
                                     Le[i] = loglik(items[i].e1, items[i].e2, &e, ctx);
                                     err[i] = e;
                                 }
                                 return true;
                             });
    return ok ? launches : -1;
}
