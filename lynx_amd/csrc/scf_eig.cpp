// Eigen-stage driver: the contract between LYNX's SCF loop and the CheFSI
// eigensolver, run on the library's C ABI.  Follows
// SCF::estimate_spectral_bounds (SCF.cpp:883-935) for the first call and
// SCF::solve_eigenproblem (SCF.cpp:310-467) for every SCF iteration:
// rho_trigger CheFSI passes in iteration 0 and nchefsi afterwards; each pass
// solves every (spin, k-point) block, then sets eig_min[s] = eig(s, k0)[0]
// (iterations > 0 only), lambda_cutoff = max_s eig(s, k0)[N-1] + 0.1, and
// recomputes the occupations and Fermi level (Occupation.cpp:53-176).
// eig_max stays at its Lanczos value.  The filter runs in FP32 while the SCF
// residual is above a threshold and in FP64 below it.
#include "lynx_chefsi.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

struct lynx_scf_eig {
    lynx_ctx *ctx = nullptr;
    lynx_scf_eig_params p{};
    std::vector<double> kpts, kw;
    long nd = 0;
    bool have_bounds = false;
    std::vector<double> eig_min, eig_max;
    double lambda_c = 0.0;
    double fermi = 0.0;
    int last_precision = LYNX_FILTER_FP64;
    std::string err;
    std::vector<lynx_batch *> batch;   // per spin: its k-points' passes on device blocks
};

namespace {

struct Fail {
    lynx_status st;
    std::string msg;
};

void ok(lynx_status st, const lynx_scf_eig *d, const char *what)
{
    if (st != LYNX_OK) throw Fail{st, std::string(what) + ": " + lynx_last_error(d->ctx)};
}

double smear(double x, double beta, int type)
{
    if (type == 1) return 0.5 * std::erfc(beta * x);
    const double a = x * beta;
    if (a > 50.0) return 0.0;
    if (a < -50.0) return 1.0;
    return 1.0 / (1.0 + std::exp(a));
}

// Brent's root finder (Brent 1973, in the form of NumericalMethods.cpp:14-78):
// inverse quadratic interpolation or secant steps kept inside the bracket,
// bisection otherwise.  Which point of a flat stretch it returns (the
// Fermi level of a gapped spectrum) depends on these steps, so they are
// kept as the reference takes them.
template <class F>
double brent(F &&f, double a, double b, double tol, int max_iter)
{
    double fa = f(a), fb = f(b);
    double c = a, fc = fa, d = b - a, e = d;
    for (int it = 0; it < max_iter; ++it) {
        if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) {
            c = a; fc = fa;
            d = e = b - a;
        }
        if (std::fabs(fc) < std::fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tol1 = 2.0 * 2.220446049250313e-16 * std::fabs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if (std::fabs(xm) <= tol1 || fb == 0.0) return b;
        if (std::fabs(e) >= tol1 && std::fabs(fa) > std::fabs(fb)) {
            const double s = fb / fa;
            double p, q;
            if (a == c) {
                p = 2.0 * xm * s;
                q = 1.0 - s;
            } else {
                const double qq = fa / fc, r = fb / fc;
                p = s * (2.0 * xm * qq * (qq - r) - (b - a) * (r - 1.0));
                q = (qq - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0) q = -q;
            p = std::fabs(p);
            if (2.0 * p < std::min(3.0 * xm * q - std::fabs(tol1 * q), std::fabs(e * q))) {
                e = d;
                d = p / q;
            } else {
                d = xm;
                e = d;
            }
        } else {
            d = xm;
            e = d;
        }
        a = b;
        fa = fb;
        b += std::fabs(d) > tol1 ? d : (xm >= 0 ? tol1 : -tol1);
        fb = f(b);
    }
    return b;
}

// Fermi level: the root of sum_i w_i f(e_i - Ef) = Ne on the reference's
// bracket (10/beta beyond the extreme levels, else 100/beta), by Brent to
// 1e-14 in at most 200 steps (Occupation.cpp:53-84).
double fermi_level(const std::vector<double> &e, const std::vector<double> &w, double ne, double beta, int type)
{
    auto occ = [&](double ef) {
        double s = 0.0;
        for (size_t i = 0; i < e.size(); ++i) s += w[i] * smear(e[i] - ef, beta, type);
        return s - ne;
    };
    const double emin = *std::min_element(e.begin(), e.end());
    const double emax = *std::max_element(e.begin(), e.end());
    double a = emin - 10.0 / beta, b = emax + 10.0 / beta;
    if (occ(a) * occ(b) > 0) {
        a = emin - 100.0 / beta;
        b = emax + 100.0 / beta;
        if (occ(a) * occ(b) > 0) throw Fail{LYNX_E_ARG, "cannot bracket the Fermi level"};
    }
    return brent(occ, a, b, 1e-14, 200);
}

}  // namespace

extern "C" {

lynx_status lynx_scf_eig_create(lynx_scf_eig **out, lynx_ctx *ctx, const lynx_scf_eig_params *p,
                                const double *kpts, const double *kpt_weights)
{
    if (!out || !ctx || !p) return LYNX_E_ARG;
    *out = nullptr;
    if (p->nspin < 1 || p->nspin > 2 || p->nkpt < 1 || p->nband < 1 || p->cheb_degree < 1 || p->rho_trigger < 1 ||
        p->nchefsi < 1 || p->beta <= 0 || (p->smearing != 0 && p->smearing != 1))
        return LYNX_E_ARG;
    int dims[3] = {0, 0, 0};
    const lynx_status st = lynx_get_grid_dims(ctx, dims);
    if (st != LYNX_OK) return st;
    auto *d = new lynx_scf_eig();
    d->ctx = ctx;
    d->p = *p;
    d->nd = (long)dims[0] * dims[1] * dims[2];
    d->kpts.assign(3 * p->nkpt, 0.0);
    if (kpts) d->kpts.assign(kpts, kpts + 3 * p->nkpt);
    d->kw.assign(p->nkpt, 1.0 / p->nkpt);
    if (kpt_weights) d->kw.assign(kpt_weights, kpt_weights + p->nkpt);
    d->eig_min.assign(p->nspin, 0.0);
    d->eig_max.assign(p->nspin, 0.0);
    *out = d;
    return LYNX_OK;
}

lynx_status lynx_scf_eig_destroy(lynx_scf_eig *d)
{
    if (d)
        for (lynx_batch *b : d->batch)
            if (b) (void)lynx_batch_destroy(b);
    delete d;
    return LYNX_OK;
}

const char *lynx_scf_eig_last_error(const lynx_scf_eig *d) { return d ? d->err.c_str() : "null driver"; }

lynx_status lynx_scf_eig_step(lynx_scf_eig *d, int scf_iter, double scf_residual, const double *veff,
                              void *const *psi, int ld, int mem, double *eig, double *occ, double *fermi,
                              int *precision_used)
{
    return lynx_scf_eig_step_veff(d, scf_iter, scf_residual, veff, LYNX_MEM_HOST, psi, ld, mem, eig, occ, fermi,
                                  precision_used);
}

lynx_status lynx_scf_eig_step_veff(lynx_scf_eig *d, int scf_iter, double scf_residual, const double *veff,
                                   int veff_mem, void *const *psi, int ld, int mem, double *eig, double *occ,
                                   double *fermi, int *precision_used)
{
    if (!d || !veff || !psi || !eig) return LYNX_E_ARG;
    // the context's filter and Gram precisions are the caller's settings:
    // the step sets its own and restores the caller's on every return
    int caller_precision = LYNX_FILTER_FP64, caller_gram = LYNX_GRAM_FP32;
    if (lynx_get_filter_precision(d->ctx, &caller_precision) != LYNX_OK) return LYNX_E_ARG;
    if (lynx_get_gram_precision(d->ctx, &caller_gram) != LYNX_OK) return LYNX_E_ARG;
    struct Restore {
        lynx_ctx *c;
        int p, g;
        ~Restore()
        {
            (void)lynx_set_filter_precision(c, p);
            (void)lynx_set_gram_precision(c, g);
        }
    } restore{d->ctx, caller_precision, caller_gram};
    try {
        const lynx_scf_eig_params &p = d->p;
        const int dtype = p.complex_kpt ? LYNX_C128 : LYNX_F64;
        const int N = p.nband;
        // spectral bounds per spin at the first k-point, once per SCF run:
        // at its first iteration (SCF.cpp:912-929)
        if (!d->have_bounds || scf_iter == 0) {
            for (int s = 0; s < p.nspin; ++s) {
                ok(lynx_set_veff(d->ctx, veff + s * d->nd, veff_mem), d, "lynx_set_veff");
                if (p.complex_kpt) ok(lynx_set_kpoint(d->ctx, &d->kpts[0]), d, "lynx_set_kpoint");
                ok(lynx_lanczos_bounds(d->ctx, dtype, &d->eig_min[s], &d->eig_max[s], 1e-2, 1000, 1, nullptr), d,
                   "lynx_lanczos_bounds");
            }
            d->lambda_c = 0.5 * (d->eig_min[0] + d->eig_max[0]);
            d->have_bounds = true;
        }
        const int passes = scf_iter == 0 ? p.rho_trigger : p.nchefsi;
        const bool single = p.fp64_switch_residual > 0 && (scf_iter == 0 || scf_residual > p.fp64_switch_residual);
        d->last_precision = !single ? LYNX_FILTER_FP64
                                    : (p.low_precision == LYNX_FILTER_BF16 ? LYNX_FILTER_BF16 : LYNX_FILTER_FP32);
        ok(lynx_set_filter_precision(d->ctx, d->last_precision), d, "lynx_set_filter_precision");
        ok(lynx_set_gram_precision(d->ctx, p.gram_precision), d, "lynx_set_gram_precision");
        const double spin_fac = p.nspin == 1 ? 2.0 : 1.0;
        // several k-points of device blocks: the k-points of a spin solved
        // together (lynx_batch_solve: the same pass per k-point, with
        // batched launches, GEMMs and rocSOLVER calls).  Reduced-precision
        // passes run its filter in complex64 and the subspace steps in
        // complex128.
        // (one batch per spin, so each keeps its plan and graph for its own
        // blocks)
        const bool batched = p.complex_kpt && p.nkpt > 1 && mem == LYNX_MEM_DEVICE;
        if (batched) {
            d->batch.resize(p.nspin, nullptr);
            for (int s = 0; s < p.nspin; ++s) {
                if (!d->batch[s]) {
                    std::vector<lynx_ctx *> ctxs(p.nkpt, d->ctx);
                    lynx_batch *b = nullptr;
                    const lynx_status bs = lynx_batch_create(&b, p.nkpt, ctxs.data(), d->kpts.data());
                    if (bs != LYNX_OK) throw Fail{bs, "lynx_batch_create failed"};
                    d->batch[s] = b;
                }
                const lynx_status bs =
                    lynx_batch_set_filter_precision(d->batch[s], single ? LYNX_FILTER_FP32 : LYNX_FILTER_FP64);
                if (bs != LYNX_OK) throw Fail{bs, "lynx_batch_set_filter_precision failed"};
            }
        }
        for (int pass = 0; pass < passes; ++pass) {
            for (int s = 0; s < p.nspin; ++s) {
                ok(lynx_set_veff(d->ctx, veff + s * d->nd, veff_mem), d, "lynx_set_veff");
                if (batched) {
                    const int nk = p.nkpt;
                    std::vector<void *> ps(psi + (long)s * nk, psi + (long)(s + 1) * nk);
                    std::vector<long> lds(nk, ld);
                    std::vector<int> nc(nk, N);
                    std::vector<double> lc(nk, d->lambda_c), lo(nk, d->eig_min[s]), hi(nk, d->eig_max[s]);
                    const lynx_status bs = lynx_batch_solve(d->batch[s], dtype, ps.data(), lds.data(), nc.data(),
                                                            eig + (long)s * nk * N, lc.data(), lo.data(), hi.data(),
                                                            p.cheb_degree, nullptr);
                    if (bs != LYNX_OK)
                        throw Fail{bs, std::string("lynx_batch_solve: ") + lynx_batch_last_error(d->batch[s])};
                    continue;
                }
                for (int k = 0; k < p.nkpt; ++k) {
                    if (p.complex_kpt) ok(lynx_set_kpoint(d->ctx, &d->kpts[3 * k]), d, "lynx_set_kpoint");
                    ok(lynx_chefsi_solve(d->ctx, dtype, psi[s * p.nkpt + k], ld, N, eig + (long)(s * p.nkpt + k) * N,
                                         d->lambda_c, d->eig_min[s], d->eig_max[s], p.cheb_degree, nullptr, mem),
                       d, "lynx_chefsi_solve");
                }
            }
            // bound update after every pass (SCF.cpp:436-449)
            double last_max = -1e30;
            for (int s = 0; s < p.nspin; ++s) {
                const double *e0 = eig + (long)(s * p.nkpt) * N;
                if (scf_iter > 0) d->eig_min[s] = e0[0];
                last_max = std::max(last_max, e0[N - 1]);
            }
            d->lambda_c = last_max + 0.1;
            // occupations over every spin and k-point
            std::vector<double> e, w;
            for (int s = 0; s < p.nspin; ++s)
                for (int k = 0; k < p.nkpt; ++k)
                    for (int n = 0; n < N; ++n) {
                        e.push_back(eig[(long)(s * p.nkpt + k) * N + n]);
                        w.push_back(d->kw[k] * spin_fac);
                    }
            d->fermi = fermi_level(e, w, p.nelectron, p.beta, p.smearing);
            if (occ)
                for (size_t i = 0; i < e.size(); ++i) occ[i] = smear(e[i] - d->fermi, p.beta, p.smearing);
        }
        if (fermi) *fermi = d->fermi;
        if (precision_used) *precision_used = d->last_precision;
        return LYNX_OK;
    } catch (const Fail &f) {
        d->err = f.msg;
        return f.st;
    }
}

lynx_status lynx_fermi_level(int nspin, int nkpt, int nband, const double *eig, const double *kpt_weights,
                             double nelectron, double beta, int smearing, double *occ, double *fermi)
{
    if (nspin < 1 || nspin > 2 || nkpt < 1 || nband < 1 || !eig || !kpt_weights || !fermi || !(beta > 0) ||
        (smearing != 0 && smearing != 1))
        return LYNX_E_ARG;
    try {
        const double spin_fac = nspin == 1 ? 2.0 : 1.0;
        std::vector<double> e, w;
        for (int s = 0; s < nspin; ++s)
            for (int k = 0; k < nkpt; ++k)
                for (int n = 0; n < nband; ++n) {
                    e.push_back(eig[((long)s * nkpt + k) * nband + n]);
                    w.push_back(kpt_weights[k] * spin_fac);
                }
        *fermi = fermi_level(e, w, nelectron, beta, smearing);
        if (occ)
            for (size_t i = 0; i < e.size(); ++i) occ[i] = smear(e[i] - *fermi, beta, smearing);
        return LYNX_OK;
    } catch (const Fail &f) {
        return f.st;
    }
}

lynx_status lynx_scf_eig_state(const lynx_scf_eig *d, double *lambda_cutoff, double *eig_min, double *eig_max)
{
    if (!d) return LYNX_E_ARG;
    if (lambda_cutoff) *lambda_cutoff = d->lambda_c;
    for (int s = 0; s < d->p.nspin; ++s) {
        if (eig_min) eig_min[s] = d->eig_min[s];
        if (eig_max) eig_max[s] = d->eig_max[s];
    }
    return LYNX_OK;
}

}  // extern "C"
