// SCF loop around the CheFSI eigen stage, FP64 on the GPU: density,
// exchange-correlation, Poisson, Pulay/Kerker mixing and the total energy.
// Follows SCF::run and its steps (SCF.cpp:110-291, 469-791, 937-1028),
// EffectivePotential::compute (EffectivePotential.cpp:44-176),
// PoissonSolver::solve (PoissonSolver.cpp:43-166), LinearSolver::aar
// (LinearSolver.cpp:23-130), KerkerPreconditioner::apply
// (KerkerPreconditioner.cpp:37-95), Mixer::mix (Mixer.cpp:116-266) and
// Energy::compute_all (Energy.cpp:140-200).  Every Nd-sized step runs in the
// kernels of kernels_scf.hip, the fused stencil or rocFFT; the host solves
// the Anderson / Pulay normal equations (at most 7 x 7) and sums partials.
//
// The Poisson and Kerker equations have constant coefficients on a periodic
// grid, so besides the reference's AAR iteration they are solved exactly in
// the Fourier basis: the FD operator's symbol at wavevector theta is
//   L(theta) = -sum_a (D2_a[0] + 2 sum_p D2_a[p] cos(p theta_a))
//              + 4 sum_{(a,b)} (sum_p D2m_ab[p] sin(p theta_a)) (sum_q D1_b[q] sin(q theta_b))
// (mixed pairs xy, xz, yz of a non-orthogonal cell), the same operator the
// stencil kernels apply, evaluated from the same coefficients.
#include "ctx.hpp"
#include "kernels_scf.hpp"

#include <rocfft/rocfft.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace lynx;
using namespace lynx::detail;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kAarM = 7, kAarP = 6;          // Anderson history and period
constexpr double kAarOmega = 0.6, kAarBeta = 0.6;
constexpr double kKerkerKTF = 1.0, kKerkerIdiemac = 0.1;

// NumericalMethods.cpp:84-120: Gaussian elimination with partial pivoting,
// rows with a vanishing pivot skipped.
void gauss_solve(std::vector<double> A, std::vector<double> &b, int n)
{
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > std::fabs(A[piv * n + k])) piv = i;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[piv * n + j]);
            std::swap(b[k], b[piv]);
        }
        const double d = A[k * n + k];
        if (std::fabs(d) < 1e-14) continue;
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / d;
            for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
            b[i] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        if (std::fabs(A[k * n + k]) < 1e-14) continue;
        for (int j = k + 1; j < n; ++j) b[k] -= A[k * n + j] * b[j];
        b[k] /= A[k * n + k];
    }
}

// symbol of -lap on the half-spectrum (nx/2+1, ny, nz) layout of a
// real-to-complex transform, x fastest
std::vector<double> fd_symbol(const Stencil &st)
{
    const int nx = st.nx, ny = st.ny, nz = st.nz, nxh = nx / 2 + 1, R = st.fdn;
    auto cosine = [&](const std::vector<double> &c, int n, int m) {
        double v = c[0];
        for (int p = 1; p <= R; ++p) v += 2.0 * c[p] * std::cos(2.0 * kPi * p * m / n);
        return v;
    };
    auto sine = [&](const std::vector<double> &c, int n, int m) {
        double v = 0.0;
        for (int p = 1; p <= R; ++p) v += c[p] * std::sin(2.0 * kPi * p * m / n);
        return v;
    };
    std::vector<double> cx(nxh), sxy(nxh), sxz(nxh), cy(ny), dy(ny), syz(ny), cz(nz), dz(nz);
    for (int i = 0; i < nxh; ++i) {
        cx[i] = cosine(st.D2[0], nx, i);
        sxy[i] = sine(st.D2m[0], nx, i);
        sxz[i] = sine(st.D2m[1], nx, i);
    }
    for (int j = 0; j < ny; ++j) {
        cy[j] = cosine(st.D2[1], ny, j);
        dy[j] = sine(st.D1[1], ny, j);
        syz[j] = sine(st.D2m[2], ny, j);
    }
    for (int k = 0; k < nz; ++k) {
        cz[k] = cosine(st.D2[2], nz, k);
        dz[k] = sine(st.D1[2], nz, k);
    }
    std::vector<double> L((size_t)nxh * ny * nz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nxh; ++i)
                L[i + (size_t)nxh * (j + (size_t)ny * k)] =
                    -(cx[i] + cy[j] + cz[k]) + 4.0 * (sxy[i] * dy[j] + sxz[i] * dz[k] + syz[j] * dz[k]);
    return L;
}

void fft_ok(rocfft_status s, const char *what)
{
    if (s != rocfft_status_success) fail(LYNX_E_HIP, std::string(what) + " failed (rocFFT status " +
                                                         std::to_string((int)s) + ")");
}

// forward and inverse real 3-D transforms of one grid, their work buffer,
// the half spectrum and the symbol
struct Spectral {
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    DevBuf work, spec, sym, in;
    long ncplx = 0;

    void init(const Stencil &st, hipStream_t hs)
    {
        static std::once_flag once;
        rocfft_status setup = rocfft_status_success;
        std::call_once(once, [&] { setup = rocfft_setup(); });
        fft_ok(setup, "rocfft_setup");
        const size_t len[3] = {(size_t)st.nx, (size_t)st.ny, (size_t)st.nz};
        fft_ok(rocfft_plan_create(&fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                  rocfft_precision_double, 3, len, 1, nullptr),
               "rocfft_plan_create (forward)");
        fft_ok(rocfft_plan_create(&inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                  rocfft_precision_double, 3, len, 1, nullptr),
               "rocfft_plan_create (inverse)");
        size_t w1 = 0, w2 = 0;
        fft_ok(rocfft_plan_get_work_buffer_size(fwd, &w1), "rocfft_plan_get_work_buffer_size");
        fft_ok(rocfft_plan_get_work_buffer_size(inv, &w2), "rocfft_plan_get_work_buffer_size");
        fft_ok(rocfft_execution_info_create(&info), "rocfft_execution_info_create");
        const size_t wb = std::max(w1, w2);
        if (wb > 0) {
            work.get(wb);
            fft_ok(rocfft_execution_info_set_work_buffer(info, work.p, wb), "rocfft_execution_info_set_work_buffer");
        }
        fft_ok(rocfft_execution_info_set_stream(info, hs), "rocfft_execution_info_set_stream");
        const std::vector<double> L = fd_symbol(st);
        ncplx = (long)L.size();
        spec.get((size_t)ncplx * 2 * sizeof(double));
        in.get((size_t)st.nx * st.ny * st.nz * sizeof(double));
        upload(sym, L, hs);
    }

    // x = F^-1 g(L) F b; b is left unchanged (a transform may overwrite its
    // input, so it reads a copy)
    void solve(const double *b, double *x, long nd, int kind, double amix, double c1, double c2, hipStream_t hs)
    {
        HIP_OK(hipMemcpyAsync(in.p, b, nd * sizeof(double), hipMemcpyDeviceToDevice, hs));
        void *src[1] = {in.p}, *dst[1] = {spec.p};
        fft_ok(rocfft_execute(fwd, src, dst, info), "rocfft_execute (forward)");
        gpu::scf_spectral_scale(spec.as<double>(), sym.as<double>(), ncplx, kind, amix, c1, c2, 1.0 / (double)nd,
                                hs);
        void *src2[1] = {spec.p}, *dst2[1] = {x};
        fft_ok(rocfft_execute(inv, src2, dst2, info), "rocfft_execute (inverse)");
    }

    void release()
    {
        if (fwd) rocfft_plan_destroy(fwd);
        if (inv) rocfft_plan_destroy(inv);
        if (info) rocfft_execution_info_destroy(info);
        fwd = inv = nullptr;
        info = nullptr;
        for (DevBuf *b : {&work, &spec, &sym, &in}) b->release();
    }
};

}  // namespace

struct lynx_scf {
    lynx_ctx *ctx = nullptr;
    lynx_scf_params p{};
    std::vector<double> kpts, kw;
    std::string err;
    lynx_scf_eig *stage = nullptr;
    long nd = 0, ldpsi = 0;
    int nblk = 0;
    bool cplx = false;
    double dV = 0, poisson_tol = 0, precond_tol = 0;
    int nelectron = 0, natom = 0;
    double eself = 0, ec = 0;
    bool nlcc = false;

    DevBuf psi, w;                               // orbital blocks; per-column density weights
    DevBuf b, core, zeros, ones;                 // pseudocharge, core density, a zero Veff, ones
    DevBuf rho;                                  // input density [total | up | down] (one spin: total)
    DevBuf rho_out;                              // new density, same layout
    DevBuf veff, vxc, exc, phi;                  // potentials (nspin x Nd), energy density, phi
    DevBuf xcin;                                 // XC input densities (core added, floored)
    DevBuf grad, aux;                            // gradients / GGA work
    DevBuf rhs, ar, af, afo, axo, aX, aF;        // AAR work
    DevBuf mfk, mxkm1, mR, mF, mxw, mfw, mpf, mx, mg;   // mixer state
    DevBuf partials;
    Spectral fft;                                // LYNX_SOLVE_SPECTRAL
    int mix_iter = 0;
    std::vector<double> hpart;

    std::vector<double> eig, occ;
    double fermi = 0;
    bool ran = false;                            // a lynx_scf_run completed
    std::vector<double> rho_init;                // restart density, layout of rho (empty: atomic start)
    lynx_scf_result res{};
    std::vector<double> h_e, h_err, h_t;
    std::vector<int> h_prec;

    ~lynx_scf()
    {
        if (stage) lynx_scf_eig_destroy(stage);
        DevBuf *bufs[] = {&psi, &w, &b, &core, &zeros, &ones, &rho, &rho_out, &veff, &vxc, &exc, &phi, &xcin, &grad,
                          &aux, &rhs, &ar, &af, &afo, &axo, &aX, &aF, &mfk, &mxkm1, &mR, &mF, &mxw, &mfw,
                          &mpf, &mx, &mg, &partials};
        if (ctx && ctx->device >= 0) (void)hipSetDevice(ctx->device);
        fft.release();
        for (DevBuf *x : bufs) x->release();
    }
};

namespace {

double *dptr(DevBuf &b) { return b.as<double>(); }

// fixed-order sum of the nb partials (block-major, k per block) of pair k
std::vector<double> sum_partials(lynx_scf *S, int nb, int npairs)
{
    S->hpart.resize((size_t)nb * npairs);
    HIP_OK(hipMemcpyAsync(S->hpart.data(), S->partials.p, S->hpart.size() * sizeof(double), hipMemcpyDeviceToHost,
                          S->ctx->stream));
    HIP_OK(hipStreamSynchronize(S->ctx->stream));
    std::vector<double> out(npairs, 0.0);
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < npairs; ++k) out[k] += S->hpart[(size_t)b * npairs + k];
    return out;
}

std::vector<double> dots(lynx_scf *S, const std::vector<const double *> &a, const std::vector<const double *> &b,
                         long n)
{
    std::vector<double> out;
    for (size_t k0 = 0; k0 < a.size(); k0 += 40) {
        const int np = (int)std::min<size_t>(40, a.size() - k0);
        const int nb = gpu::scf_multidot(a.data() + k0, b.data() + k0, np, n, dptr(S->partials), S->ctx->stream);
        auto part = sum_partials(S, nb, np);
        out.insert(out.end(), part.begin(), part.end());
    }
    return out;
}

double dot1(lynx_scf *S, const double *a, const double *b, long n) { return dots(S, {a}, {b}, n)[0]; }

void dcopy(lynx_scf *S, double *dst, const double *src, long n)
{
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, S->ctx->stream));
}

// r = b - (-lap x + k2 x) through the fused stencil at k = 0:
// a (-1/2 lap x + s x) + old with a = -2, s = k2 / 2, old = b
void residual(lynx_scf *S, const double *x, const double *bvec, double *r, double k2)
{
    lynx_ctx *c = S->ctx;
    const double k0[3] = {0, 0, 0};
    const gpu::StepCfg &cfg = step_config(c, gpu::F64, x, S->nd, r, S->nd, 1);
    gpu::stencil_step(gpu::F64, geom_at(c, k0), x, S->nd, bvec, S->nd, r, S->nd, S->zeros.p, 1, -2.0,
                      bvec ? 1.0 : 0.0, 0.5 * k2, cfg, c->stream);
}

// LinearSolver::aar / PoissonSolver::solve for (-lap + k2) x = b with a
// Jacobi preconditioner m_inv; x is the initial guess and the result.
int aar_solve(lynx_scf *S, const double *bvec, double *x, double tol, double k2, double m_inv, int max_iter)
{
    const long n = S->nd;
    double *r = dptr(S->ar), *f = dptr(S->af), *fo = dptr(S->afo), *xo = dptr(S->axo);
    double *X = dptr(S->aX), *F = dptr(S->aF);
    dcopy(S, xo, x, n);
    const double abs_tol = tol * std::sqrt(dot1(S, bvec, bvec, n));
    residual(S, x, bvec, r, k2);
    double rnorm = abs_tol + 1.0;   // the first norm is not checked (as the reference)
    int it = 0;
    while (rnorm > abs_tol && it < max_iter) {
        const bool anderson = (it + 1) % kAarP == 0 && it > 0;
        gpu::scf_aar_pre(x, xo, f, fo, r, X, F, it > 0 ? (it - 1) % kAarM : -1, m_inv, kAarOmega, !anderson, n,
                         S->ctx->stream);
        if (anderson) {
            const int cols = std::min(it, kAarM);
            std::vector<const double *> a, b;
            for (int i = 0; i < cols; ++i) {
                a.push_back(F + (long)i * n);
                b.push_back(f);
                for (int j = 0; j <= i; ++j) {
                    a.push_back(F + (long)i * n);
                    b.push_back(F + (long)j * n);
                }
            }
            const auto d = dots(S, a, b, n);
            std::vector<double> FtF((size_t)cols * cols), g(cols);
            size_t k = 0;
            for (int i = 0; i < cols; ++i) {
                g[i] = d[k++];
                for (int j = 0; j <= i; ++j) FtF[i * cols + j] = FtF[j * cols + i] = d[k++];
            }
            gauss_solve(FtF, g, cols);
            gpu::scf_aar_anderson(x, xo, f, X, F, g.data(), cols, kAarBeta, n, S->ctx->stream);
            residual(S, x, bvec, r, k2);
            rnorm = std::sqrt(dot1(S, r, r, n));
        } else {
            residual(S, x, bvec, r, k2);
        }
        ++it;
    }
    return it;
}

double stencil_diag(const lynx_ctx *c) { return c->st.D2[0][0] + c->st.D2[1][0] + c->st.D2[2][0]; }

// EffectivePotential::solve_poisson: -lap phi = 4 pi (rho + b), the right
// side shifted to zero mean, phi shifted to zero mean (periodic gauge)
void solve_poisson(lynx_scf *S, const double *rho_total)
{
    const long n = S->nd;
    double *rhs = dptr(S->rhs);
    dcopy(S, rhs, rho_total, n);
    gpu::scf_axpby(rhs, dptr(S->b), 1.0, 1.0, n, S->ctx->stream);
    gpu::scf_scale(rhs, 4.0 * kPi, n, S->ctx->stream);
    if (S->p.linear_solver == LYNX_SOLVE_SPECTRAL) {
        // zero-mean right side and solution: the q = 0 coefficient is dropped
        S->fft.solve(rhs, dptr(S->phi), n, 0, 0.0, 0.0, 0.0, S->ctx->stream);
        return;
    }
    const double ntot = (double)n;
    const double mean = dot1(S, rhs, dptr(S->ones), n) / ntot;
    gpu::scf_axpby(rhs, dptr(S->ones), -mean, 1.0, n, S->ctx->stream);
    const double d = stencil_diag(S->ctx);
    const double m_inv = std::fabs(d) < 1e-14 ? 1.0 : -1.0 / d;
    aar_solve(S, rhs, dptr(S->phi), S->poisson_tol, 0.0, m_inv, 3000);
    const double pm = dot1(S, dptr(S->phi), dptr(S->ones), n) / ntot;
    gpu::scf_axpby(dptr(S->phi), dptr(S->ones), -pm, 1.0, n, S->ctx->stream);
}

// gradient of a field along the three lattice directions into g[0..2]
void gradient3(lynx_scf *S, const double *x, double *const g[3])
{
    const lynx_ctx *c = S->ctx;
    for (int a = 0; a < 3; ++a)
        gpu::scf_grad(x, g[a], a, c->st.nx, c->st.ny, c->st.nz, c->st.D1[a].data(), c->st.fdn, 1.0, false,
                      S->ctx->stream);
}

// out -= sum_a D_a f_a
void sub_divergence(lynx_scf *S, double *const f[3], double *out)
{
    const lynx_ctx *c = S->ctx;
    for (int a = 0; a < 3; ++a)
        gpu::scf_grad(f[a], out, a, c->st.nx, c->st.ny, c->st.nz, c->st.D1[a].data(), c->st.fdn, -1.0, true,
                      S->ctx->stream);
}

void metric(const lynx_ctx *c, double M[9])
{
    for (int i = 0; i < 9; ++i) M[i] = c->lat.orth ? (i % 4 == 0 ? 1.0 : 0.0) : c->lat.lapcT[i];
}

// EffectivePotential::compute: XC of the input density (core density
// added, floored at 1e-14 when present), Poisson, Veff_s = Vxc_s + phi
void build_potential(lynx_scf *S)
{
    const long n = S->nd;
    hipStream_t st = S->ctx->stream;
    const int kind = S->p.xc;
    const bool gga = kind == LYNX_XC_GGA_PBE;
    double M[9];
    metric(S->ctx, M);
    double *rho = dptr(S->rho), *xin = dptr(S->xcin), *g = dptr(S->grad), *aux = dptr(S->aux);
    double *vxc = dptr(S->vxc), *exc = dptr(S->exc);
    if (S->p.nspin == 1) {
        const double *r = rho;
        if (S->nlcc) {
            dcopy(S, xin, rho, n);
            gpu::scf_axpby(xin, dptr(S->core), 1.0, 1.0, n, st);
            gpu::scf_clamp_floor(xin, 1e-14, n, st);
            r = xin;
        }
        if (!gga) {
            gpu::scf_xc_unpol(kind, r, nullptr, exc, vxc, nullptr, n, st);
        } else {
            double *gr[3] = {g, g + n, g + 2 * n};
            gradient3(S, r, gr);
            double *sigma = aux, *v2 = aux + n;
            const double *cg[3] = {gr[0], gr[1], gr[2]};
            gpu::scf_sigma(M, cg, cg, sigma, n, st);
            gpu::scf_xc_unpol(kind, r, sigma, exc, vxc, v2, n, st);
            double *f[3] = {g + 3 * n, g + 4 * n, g + 5 * n};
            gpu::scf_gga_flux(M, v2, cg, nullptr, cg, f, n, st);
            sub_divergence(S, f, vxc);
        }
    } else {
        // [total | up | down] with core halves added, floored (EffectivePotential.cpp:103-125)
        double *xt = xin, *xu = xin + n, *xd = xin + 2 * n;
        dcopy(S, xu, rho + n, n);
        dcopy(S, xd, rho + 2 * n, n);
        if (S->nlcc) {
            gpu::scf_axpby(xu, dptr(S->core), 0.5, 1.0, n, st);
            gpu::scf_axpby(xd, dptr(S->core), 0.5, 1.0, n, st);
        }
        gpu::scf_clamp_floor(xu, 0.5e-14, n, st);
        gpu::scf_clamp_floor(xd, 0.5e-14, n, st);
        dcopy(S, xt, xu, n);
        gpu::scf_axpby(xt, xd, 1.0, 1.0, n, st);
        if (!gga) {
            gpu::scf_xc_pol(kind, xu, xd, nullptr, nullptr, nullptr, exc, vxc, vxc + n, nullptr, nullptr, nullptr, n,
                            st);
        } else {
            double *gu[3] = {g, g + n, g + 2 * n}, *gd[3] = {g + 3 * n, g + 4 * n, g + 5 * n};
            gradient3(S, xu, gu);
            gradient3(S, xd, gd);
            const double *cu[3] = {gu[0], gu[1], gu[2]}, *cd[3] = {gd[0], gd[1], gd[2]};
            double *suu = aux, *sud = aux + n, *sdd = aux + 2 * n, *vsuu = aux + 3 * n, *vsud = aux + 4 * n,
                   *vsdd = aux + 5 * n;
            gpu::scf_sigma(M, cu, cu, suu, n, st);
            gpu::scf_sigma(M, cu, cd, sud, n, st);
            gpu::scf_sigma(M, cd, cd, sdd, n, st);
            gpu::scf_xc_pol(kind, xu, xd, suu, sud, sdd, exc, vxc, vxc + n, vsuu, vsud, vsdd, n, st);
            // f_up = M (2 vs_uu grad up + vs_ud grad dn), f_dn = M (vs_ud grad up + 2 vs_dd grad dn)
            double *two = aux + 6 * n;
            double *f[3] = {g + 6 * n, g + 7 * n, g + 8 * n};
            dcopy(S, two, vsuu, n);
            gpu::scf_scale(two, 2.0, n, st);
            gpu::scf_gga_flux(M, two, cu, vsud, cd, f, n, st);
            sub_divergence(S, f, vxc);
            dcopy(S, two, vsdd, n);
            gpu::scf_scale(two, 2.0, n, st);
            gpu::scf_gga_flux(M, two, cd, vsud, cu, f, n, st);
            sub_divergence(S, f, vxc + n);
        }
    }
    solve_poisson(S, rho);
    for (int s = 0; s < S->p.nspin; ++s) {
        double *v = dptr(S->veff) + (long)s * n;
        dcopy(S, v, vxc + (long)s * n, n);
        gpu::scf_axpby(v, dptr(S->phi), 1.0, 1.0, n, st);
    }
}

// ElectronDensity::compute: rho_s = sum_k w_k spin_fac sum_n f_n |psi|^2
void compute_density(lynx_scf *S, double *out)
{
    const long n = S->nd;
    hipStream_t st = S->ctx->stream;
    const int N = S->p.nband, nk = S->p.nkpt, ns = S->p.nspin;
    const double spin_fac = ns == 1 ? 2.0 : 1.0;
    const size_t es = S->cplx ? 16 : 8;
    HIP_OK(hipMemsetAsync(out, 0, (size_t)(ns == 1 ? 1 : 3) * n * sizeof(double), st));
    std::vector<double> wh((size_t)S->nblk * N);
    for (int s = 0; s < ns; ++s)
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < N; ++j) {
                const double f = S->occ[((size_t)s * nk + k) * N + j];
                wh[((size_t)s * nk + k) * N + j] = f < 1e-16 ? 0.0 : spin_fac * S->kw[k] * f;
            }
    HIP_OK(hipMemcpyAsync(S->w.p, wh.data(), wh.size() * sizeof(double), hipMemcpyHostToDevice, st));
    for (int s = 0; s < ns; ++s) {
        double *rs = ns == 1 ? out : out + (long)(1 + s) * n;
        for (int k = 0; k < nk; ++k) {
            const int blk = s * nk + k;
            const char *P = static_cast<const char *>(S->psi.p) + (size_t)blk * S->ldpsi * N * es;
            gpu::scf_density(S->cplx, P, S->ldpsi, N, dptr(S->w) + (size_t)blk * N, rs, n, st);
        }
    }
    if (ns == 2) {
        dcopy(S, out, out + n, n);
        gpu::scf_axpby(out, out + 2 * n, 1.0, 1.0, n, st);
    }
    HIP_OK(hipStreamSynchronize(st));
}

// KerkerPreconditioner::apply on one column: Lf = (lap - idiemac kTF^2) r,
// (-lap + kTF^2) z = Lf by AAR from z = 0, z *= -amix
void kerker(lynx_scf *S, const double *r, double *z, double amix)
{
    const long n = S->nd;
    const double k2 = kKerkerKTF * kKerkerKTF;
    if (S->p.linear_solver == LYNX_SOLVE_SPECTRAL) {
        // z = -amix (-L - idiemac k2) / (L + k2) r
        S->fft.solve(r, z, n, 1, amix, kKerkerIdiemac * k2, k2, S->ctx->stream);
        return;
    }
    double *Lf = dptr(S->rhs);
    lynx_ctx *c = S->ctx;
    const double k0[3] = {0, 0, 0};
    const gpu::StepCfg &cfg = step_config(c, gpu::F64, r, n, Lf, n, 1);
    gpu::stencil_step(gpu::F64, geom_at(c, k0), r, n, nullptr, 0, Lf, n, S->zeros.p, 1, -2.0, 0.0,
                      0.5 * kKerkerIdiemac * k2, cfg, c->stream);
    const double md = stencil_diag(c) - k2;
    const double m_inv = std::fabs(md) < 1e-14 ? 1.0 : -1.0 / md;
    HIP_OK(hipMemsetAsync(z, 0, n * sizeof(double), c->stream));
    aar_solve(S, Lf, z, S->precond_tol, k2, m_inv, 1000);
    gpu::scf_scale(z, -amix, n, c->stream);
}

// Mixer::mix with the density constraint: x (in/out) and g, ncol columns of Nd
void pulay_mix(lynx_scf *S, double *x, const double *g, int ncol)
{
    const long nd = S->nd, N = nd * ncol;
    const int m = S->p.mixing_history;
    hipStream_t st = S->ctx->stream;
    double *fk = dptr(S->mfk), *xkm1 = dptr(S->mxkm1), *R = dptr(S->mR), *F = dptr(S->mF);
    double *xw = dptr(S->mxw), *fw = dptr(S->mfw), *Pf = dptr(S->mpf);
    const int it = S->mix_iter;
    gpu::scf_mix_residual(x, g, xkm1, fk, R, F, it > 0 ? (it - 1) % m : -1, N, st);
    if (it > 0) {
        const int cols = std::min(it, m);
        std::vector<const double *> a, b;
        for (int i = 0; i < cols; ++i) {
            a.push_back(F + (long)i * N);
            b.push_back(fk);
            for (int j = 0; j <= i; ++j) {
                a.push_back(F + (long)i * N);
                b.push_back(F + (long)j * N);
            }
        }
        const auto d = dots(S, a, b, N);
        std::vector<double> FtF((size_t)cols * cols), gm(cols);
        size_t k = 0;
        for (int i = 0; i < cols; ++i) {
            gm[i] = d[k++];
            for (int j = 0; j <= i; ++j) FtF[i * cols + j] = FtF[j * cols + i] = d[k++];
        }
        gauss_solve(FtF, gm, cols);
        gpu::scf_mix_wavg(x, fk, R, F, gm.data(), cols, xw, fw, N, st);
    } else {
        dcopy(S, xw, x, N);
        dcopy(S, fw, fk, N);
    }
    const double amix = S->p.mixing_param;
    if (S->p.kerker) {
        kerker(S, fw, Pf, amix);
        for (int c = 1; c < ncol; ++c) {
            dcopy(S, Pf + c * nd, fw + c * nd, nd);
            gpu::scf_scale(Pf + c * nd, amix, nd, st);
        }
    } else {
        dcopy(S, Pf, fw, N);
        gpu::scf_scale(Pf, amix, N, st);
    }
    dcopy(S, xkm1, x, N);
    dcopy(S, x, xw, N);
    gpu::scf_axpby(x, Pf, 1.0, 1.0, N, st);
    // Mixer.cpp:70-114: clamp to non-negative densities, rescale to Ne
    const int nb = gpu::scf_clamp_density(x, ncol, nd, dptr(S->partials), st);
    const double ne = sum_partials(S, nb, 1)[0] * S->dV;
    if (ne > 1e-10) gpu::scf_scale(x, (double)S->nelectron / ne, N, st);
    ++S->mix_iter;
}

}  // namespace

extern "C" {

lynx_status lynx_scf_create(lynx_scf **out, lynx_ctx *ctx, const lynx_scf_params *p, const double *kpts,
                            const double *kpt_weights)
{
    if (!out || !ctx || !p) return LYNX_E_ARG;
    *out = nullptr;
    auto *S = new lynx_scf();
    S->ctx = ctx;
    S->p = *p;
    const lynx_status st = guarded(ctx, [&] {
        need_gpu_stencil(ctx);
        if (p->nspin < 1 || p->nspin > 2 || p->nkpt < 1 || p->nband < 1 || p->cheb_degree < 1 ||
            p->rho_trigger < 1 || p->nchefsi < 1 || !(p->beta > 0) || p->xc < 0 || p->xc > 2 || !(p->tol > 0) ||
            p->max_iter < 1 || p->mixing_history < 1 || p->mixing_history > 8 || !(p->mixing_param > 0) ||
            (p->linear_solver != LYNX_SOLVE_SPECTRAL && p->linear_solver != LYNX_SOLVE_AAR))
            fail(LYNX_E_ARG, "bad SCF parameters");
        if (!p->complex_kpt && p->nkpt != 1) fail(LYNX_E_ARG, "real orbitals are for one (Gamma) k-point");
        const ElecHost &E = electrostatics(ctx);
        S->nd = ctx->nd;
        S->ldpsi = padded_ld(ctx->nd);
        S->cplx = p->complex_kpt != 0;
        S->nblk = p->nspin * p->nkpt;
        S->dV = ctx->st.dV;
        S->nelectron = E.nelectron;
        S->natom = E.natom;
        S->eself = E.eself;
        S->ec = E.ec;
        S->nlcc = E.has_nlcc;
        S->kpts.assign(3 * p->nkpt, 0.0);
        if (kpts) S->kpts.assign(kpts, kpts + 3 * p->nkpt);
        S->kw.assign(p->nkpt, 1.0 / p->nkpt);
        if (kpt_weights) S->kw.assign(kpt_weights, kpt_weights + p->nkpt);
        // ParameterDefaults.cpp:10-47
        const double *h = ctx->st.h;
        const double heff = (std::fabs(h[0] - h[1]) < 1e-12 && std::fabs(h[1] - h[2]) < 1e-12)
                                ? h[0]
                                : std::sqrt(3.0 / (1 / (h[0] * h[0]) + 1 / (h[1] * h[1]) + 1 / (h[2] * h[2])));
        S->poisson_tol = p->poisson_tol > 0 ? p->poisson_tol : p->tol * 0.01;
        S->precond_tol = p->precond_tol > 0 ? p->precond_tol : heff * heff * 1e-3;
        const long n = S->nd;
        const int ns = p->nspin;
        const int ncol_mix = ns;
        const long N = n * ncol_mix;
        hipStream_t hs = ctx->stream;
        S->psi.get((size_t)S->nblk * S->ldpsi * p->nband * (S->cplx ? 16 : 8));
        S->w.get((size_t)S->nblk * p->nband * sizeof(double));
        upload(S->b, E.b, hs);
        upload(S->core, E.rho_core, hs);
        S->zeros.get(n * sizeof(double));
        HIP_OK(hipMemsetAsync(S->zeros.p, 0, n * sizeof(double), hs));
        S->ones.get(n * sizeof(double));
        gpu::scf_fill(dptr(S->ones), 1.0, n, hs);
        // the stencil autotuner (first Poisson solve) reads the context's
        // Veff: start it at zero; the eigen stage sets each spin's own
        {
            const lynx_status vs = lynx_set_veff(ctx, dptr(S->zeros), LYNX_MEM_DEVICE);
            if (vs != LYNX_OK) fail(vs, std::string("lynx_set_veff: ") + lynx_last_error(ctx));
        }
        const long nrho = ns == 1 ? n : 3 * n;
        S->rho.get(nrho * sizeof(double));
        S->rho_out.get(nrho * sizeof(double));
        S->veff.get(ns * n * sizeof(double));
        S->vxc.get(ns * n * sizeof(double));
        S->exc.get(n * sizeof(double));
        S->phi.get(n * sizeof(double));
        HIP_OK(hipMemsetAsync(S->phi.p, 0, n * sizeof(double), hs));
        S->xcin.get(3 * n * sizeof(double));
        S->grad.get(9 * n * sizeof(double));
        S->aux.get(7 * n * sizeof(double));
        for (DevBuf *x : {&S->rhs, &S->ar, &S->af, &S->afo, &S->axo}) x->get(n * sizeof(double));
        S->aX.get(kAarM * n * sizeof(double));
        S->aF.get(kAarM * n * sizeof(double));
        for (DevBuf *x : {&S->mfk, &S->mxkm1, &S->mxw, &S->mfw, &S->mpf, &S->mx, &S->mg})
            x->get(N * sizeof(double));
        S->mR.get((size_t)p->mixing_history * N * sizeof(double));
        S->mF.get((size_t)p->mixing_history * N * sizeof(double));
        S->partials.get((size_t)gpu::kScfPartialBlocks * 40 * sizeof(double));
        if (p->linear_solver == LYNX_SOLVE_SPECTRAL) S->fft.init(ctx->st, hs);
        lynx_scf_eig_params ep{};
        ep.nspin = ns;
        ep.nkpt = p->nkpt;
        ep.complex_kpt = p->complex_kpt;
        ep.nband = p->nband;
        ep.cheb_degree = p->cheb_degree;
        ep.rho_trigger = p->rho_trigger;
        ep.nchefsi = p->nchefsi;
        ep.nelectron = E.nelectron;
        ep.beta = p->beta;
        ep.smearing = p->smearing;
        ep.fp64_switch_residual = p->fp64_switch_residual;
        ep.low_precision = p->low_precision;
        ep.gram_precision = p->gram_precision;
        const lynx_status es = lynx_scf_eig_create(&S->stage, ctx, &ep, S->kpts.data(), S->kw.data());
        if (es != LYNX_OK) fail(es, "lynx_scf_eig_create failed");
        HIP_OK(hipStreamSynchronize(hs));
    });
    if (st != LYNX_OK) {
        delete S;
        return st;
    }
    *out = S;
    return LYNX_OK;
}

lynx_status lynx_scf_destroy(lynx_scf *S)
{
    delete S;
    return LYNX_OK;
}

const char *lynx_scf_last_error(const lynx_scf *S) { return S ? S->err.c_str() : "null SCF"; }

lynx_status lynx_scf_build_potential(lynx_scf *S, const double *rho)
{
    if (!S || !rho) return LYNX_E_ARG;
    const lynx_status st = guarded(S->ctx, [&] {
        const long n = S->nd;
        if (S->p.nspin == 1) {
            HIP_OK(hipMemcpyAsync(S->rho.p, rho, n * sizeof(double), hipMemcpyHostToDevice, S->ctx->stream));
        } else {
            HIP_OK(hipMemcpyAsync(dptr(S->rho) + n, rho, 2 * n * sizeof(double), hipMemcpyHostToDevice,
                                  S->ctx->stream));
            dcopy(S, dptr(S->rho), dptr(S->rho) + n, n);
            gpu::scf_axpby(dptr(S->rho), dptr(S->rho) + 2 * n, 1.0, 1.0, n, S->ctx->stream);
        }
        build_potential(S);
        HIP_OK(hipStreamSynchronize(S->ctx->stream));
    });
    if (st != LYNX_OK) S->err = S->ctx->err;
    return st;
}

lynx_status lynx_scf_set_initial_density(lynx_scf *S, const double *rho_total, int nd, const double *mag)
{
    if (!S) return LYNX_E_ARG;
    if (!rho_total) {
        S->rho_init.clear();
        return LYNX_OK;
    }
    if (nd != S->nd) {
        S->err = "lynx_scf_set_initial_density: nd " + std::to_string(nd) + " is not the grid's " +
                 std::to_string(S->nd);
        return LYNX_E_ARG;
    }
    // SCF::set_initial_density (SCF.cpp:83-107): the total as given; two
    // spins split as (total +- mag) / 2 floored at zero, or evenly without mag
    const long n = S->nd;
    if (S->p.nspin == 1) {
        S->rho_init.assign(rho_total, rho_total + n);
        return LYNX_OK;
    }
    S->rho_init.assign(3 * n, 0.0);
    double *t = S->rho_init.data(), *up = t + n, *dn = t + 2 * n;
    for (long i = 0; i < n; ++i) {
        t[i] = rho_total[i];
        if (mag) {
            up[i] = std::max(0.5 * (rho_total[i] + mag[i]), 0.0);
            dn[i] = std::max(0.5 * (rho_total[i] - mag[i]), 0.0);
        } else {
            up[i] = 0.5 * rho_total[i];
            dn[i] = 0.5 * rho_total[i];
        }
    }
    return LYNX_OK;
}

lynx_status lynx_scf_run(lynx_scf *S, lynx_scf_result *out)
{
    if (!S) return LYNX_E_ARG;
    const auto t0 = std::chrono::steady_clock::now();
    const lynx_status st = guarded(S->ctx, [&] {
        lynx_ctx *c = S->ctx;
        hipStream_t hs = c->stream;
        const long n = S->nd;
        const int ns = S->p.nspin, nk = S->p.nkpt, N = S->p.nband;
        const ElecHost &E = electrostatics(c);
        // initial density: a restart density given by
        // lynx_scf_set_initial_density, else the isolated-atom superposition
        // (SCF.cpp:705-725) split evenly between spins
        if (!S->rho_init.empty()) {
            HIP_OK(hipMemcpyAsync(S->rho.p, S->rho_init.data(), S->rho_init.size() * sizeof(double),
                                  hipMemcpyHostToDevice, hs));
        } else {
            HIP_OK(hipMemcpyAsync(S->rho.p, E.rho_at.data(), n * sizeof(double), hipMemcpyHostToDevice, hs));
        }
        if (ns == 2 && S->rho_init.empty()) {
            dcopy(S, dptr(S->rho) + n, dptr(S->rho), n);
            gpu::scf_scale(dptr(S->rho) + n, 0.5, n, hs);
            dcopy(S, dptr(S->rho) + 2 * n, dptr(S->rho) + n, n);
        }
        HIP_OK(hipMemsetAsync(S->phi.p, 0, n * sizeof(double), hs));
        {
            const lynx_status vs = lynx_set_veff(c, dptr(S->zeros), LYNX_MEM_DEVICE);
            if (vs != LYNX_OK) fail(vs, std::string("lynx_set_veff: ") + lynx_last_error(c));
        }
        S->mix_iter = 0;
        build_potential(S);
        // orbitals: Wavefunction::randomize per (spin, k-point), seed
        // spin * 1000 + 1 (SCF.cpp:863-881)
        const int dt = S->cplx ? LYNX_C128 : LYNX_F64;
        const size_t es = S->cplx ? 16 : 8;
        std::vector<void *> blocks(S->nblk);
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k < nk; ++k) {
                const int blk = s * nk + k;
                blocks[blk] = static_cast<char *>(S->psi.p) + (size_t)blk * S->ldpsi * N * es;
                const lynx_status rs = lynx_randomize(c, dt, blocks[blk], (int)n, (int)S->ldpsi, N,
                                                      (unsigned)(s * 1000 + 1), LYNX_MEM_DEVICE);
                if (rs != LYNX_OK) fail(rs, std::string("lynx_randomize: ") + lynx_last_error(c));
            }
        S->eig.assign((size_t)S->nblk * N, 0.0);
        S->occ.assign((size_t)S->nblk * N, 0.0);
        S->h_e.clear(); S->h_err.clear(); S->h_t.clear(); S->h_prec.clear();
        double err = 1.0;
        bool converged = false;
        int iters = 0;
        lynx_scf_result R{};
        const double spin_fac = ns == 1 ? 2.0 : 1.0;
        for (int it = 0; it < S->p.max_iter; ++it) {
            const auto ti = std::chrono::steady_clock::now();
            int prec = 0;
            const lynx_status es = lynx_scf_eig_step_veff(S->stage, it, err, dptr(S->veff), LYNX_MEM_DEVICE,
                                                          blocks.data(), (int)S->ldpsi, LYNX_MEM_DEVICE,
                                                          S->eig.data(), S->occ.data(), &S->fermi, &prec);
            if (es != LYNX_OK) fail(es, std::string("eigen stage: ") + lynx_scf_eig_last_error(S->stage));
            compute_density(S, dptr(S->rho_out));
            // energy of the input density and the new eigenvalues (Energy.cpp:140-200)
            double eband = 0, entropy = 0;
            for (int s = 0; s < ns; ++s)
                for (int k = 0; k < nk; ++k)
                    for (int j = 0; j < N; ++j) {
                        const size_t i = ((size_t)s * nk + k) * N + j;
                        const double wk = S->kw[k] * spin_fac, f = S->occ[i];
                        eband += wk * f * S->eig[i];
                        if (S->p.smearing == 0) {
                            if (f > 1e-14 && f < 1.0 - 1e-14)
                                entropy += wk * (-(f * std::log(f) + (1.0 - f) * std::log(1.0 - f)));
                        } else {
                            const double x = S->p.beta * (S->eig[i] - S->fermi);
                            entropy += wk * (0.5 / std::sqrt(kPi)) * std::exp(-x * x);
                        }
                    }
            entropy = -entropy / S->p.beta;
            const double *rt = dptr(S->rho);
            double *tmp = dptr(S->aux);
            dcopy(S, tmp, rt, n);
            if (S->nlcc) gpu::scf_axpby(tmp, dptr(S->core), 1.0, 1.0, n, hs);
            dcopy(S, tmp + n, rt, n);
            gpu::scf_axpby(tmp + n, dptr(S->b), 1.0, 1.0, n, hs);
            std::vector<const double *> a = {tmp, rt, tmp + n}, b = {dptr(S->exc), dptr(S->phi), dptr(S->phi)};
            if (ns == 1) {
                a.push_back(rt);
                b.push_back(dptr(S->vxc));
            } else {
                a.push_back(rt + n);
                b.push_back(dptr(S->vxc));
                a.push_back(rt + 2 * n);
                b.push_back(dptr(S->vxc) + n);
            }
            // the residual ||rho_out - rho_in|| / ||rho_out|| of the totals
            double *diff = tmp + 2 * n;
            dcopy(S, diff, dptr(S->rho_out), n);
            gpu::scf_axpby(diff, rt, -1.0, 1.0, n, hs);
            a.push_back(diff);
            b.push_back(diff);
            a.push_back(dptr(S->rho_out));
            b.push_back(dptr(S->rho_out));
            const auto d = dots(S, a, b, n);
            const double exc = d[0] * S->dV, e3 = d[1] * S->dV, ehart = 0.5 * d[2] * S->dV;
            const double e2 = (ns == 1 ? d[3] : d[3] + d[4]) * S->dV;
            const double sdiff = d[ns == 1 ? 4 : 5], sout = d[ns == 1 ? 5 : 6];
            err = sout > 0 ? std::sqrt(sdiff / sout) : 0.0;
            R.eband = eband;
            R.exc = exc;
            R.ehart = ehart;
            R.eself = S->eself;
            R.ec = S->ec;
            R.entropy = entropy;
            R.etotal = eband - e2 + exc - e3 + ehart + S->eself + S->ec + entropy;
            R.fermi = S->fermi;
            R.scf_error = err;
            iters = it + 1;
            S->h_e.push_back(R.etotal);
            S->h_err.push_back(err);
            S->h_prec.push_back(prec);
            const bool done = it >= S->p.min_iter && err < S->p.tol;
            if (!done) {
                // mix [total] or [total | magnetization], then the potential
                // of the mixed density (SCF.cpp:636-791)
                if (ns == 1) {
                    pulay_mix(S, dptr(S->rho), dptr(S->rho_out), 1);
                } else {
                    double *x = dptr(S->mx), *g = dptr(S->mg);
                    const double *ri = dptr(S->rho), *ro = dptr(S->rho_out);
                    dcopy(S, x, ri, n);
                    dcopy(S, x + n, ri + n, n);
                    gpu::scf_axpby(x + n, ri + 2 * n, -1.0, 1.0, n, hs);
                    dcopy(S, g, ro, n);
                    dcopy(S, g + n, ro + n, n);
                    gpu::scf_axpby(g + n, ro + 2 * n, -1.0, 1.0, n, hs);
                    pulay_mix(S, x, g, 2);
                    double *rr = dptr(S->rho);
                    dcopy(S, rr, x, n);
                    // up = (total + mag) / 2, down = (total - mag) / 2
                    dcopy(S, rr + n, x, n);
                    gpu::scf_axpby(rr + n, x + n, 0.5, 0.5, n, hs);
                    dcopy(S, rr + 2 * n, x, n);
                    gpu::scf_axpby(rr + 2 * n, x + n, -0.5, 0.5, n, hs);
                }
                build_potential(S);
            }
            HIP_OK(hipStreamSynchronize(hs));
            S->h_t.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - ti).count());
            if (done) {
                converged = true;
                break;
            }
        }
        R.iterations = iters;
        R.converged = converged ? 1 : 0;
        R.natom = S->natom;
        S->res = R;
    });
    S->res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st != LYNX_OK) {
        S->err = S->ctx->err;
        return st;
    }
    S->ran = true;
    if (out) *out = S->res;
    return LYNX_OK;
}

lynx_status lynx_scf_forces(lynx_scf *S, double *forces, double *parts)
{
    if (!S || !forces) return LYNX_E_ARG;
    if (!S->ran) return LYNX_E_STATE;
    const lynx_status st = guarded(S->ctx, [&] {
        lynx_ctx *c = S->ctx;
        hipStream_t hs = c->stream;
        const long n = S->nd;
        const int ns = S->p.nspin, nk = S->p.nkpt, N = S->p.nband;
        const ElecHost &E = electrostatics(c);
        const Stencil &stn = c->st;
        const int natom = S->natom;
        // local and core-correction terms: gradients of phi and Vc on the
        // device, the per-atom box sums on the host (Forces.cpp:186-420, 1108-1275)
        double *g = dptr(S->grad);   // 9 Nd
        double *vc = dptr(S->aux);
        HIP_OK(hipMemcpyAsync(vc, E.vc.data(), n * sizeof(double), hipMemcpyHostToDevice, hs));
        for (int a = 0; a < 3; ++a) {
            gpu::scf_grad(dptr(S->phi), g + a * n, a, stn.nx, stn.ny, stn.nz, stn.D1[a].data(), stn.fdn, 1.0, false,
                          hs);
            gpu::scf_grad(vc, g + (3 + a) * n, a, stn.nx, stn.ny, stn.nz, stn.D1[a].data(), stn.fdn, 1.0, false, hs);
        }
        std::vector<double> hg(6 * n), vxc;
        HIP_OK(hipMemcpyAsync(hg.data(), g, 6 * n * sizeof(double), hipMemcpyDeviceToHost, hs));
        if (E.has_nlcc) {
            // the first spin's Vxc, as the reference passes SCF::Vxc()
            vxc.resize(n);
            HIP_OK(hipMemcpyAsync(vxc.data(), S->vxc.p, n * sizeof(double), hipMemcpyDeviceToHost, hs));
        }
        HIP_OK(hipStreamSynchronize(hs));
        const double *dphi[3] = {hg.data(), hg.data() + n, hg.data() + 2 * n};
        const double *dvc[3] = {hg.data() + 3 * n, hg.data() + 4 * n, hg.data() + 5 * n};
        std::vector<double> f_loc, f_xc;
        local_forces(c->lat, stn, c->psps, c->atom_types, c->atom_pos, E, dphi, dvc,
                     E.has_nlcc ? vxc.data() : nullptr, f_loc, f_xc);
        // nonlocal term (Forces.cpp:509-800): alpha = dV Gamma <chi|psi> and
        // beta_d = dV <chi|D_d psi> per (atom, projector, band), both with
        // the image Bloch factors, through the apply path's gather
        std::vector<double> f_nl(3 * (size_t)natom, 0.0);
        if (c->has_nl && c->nl64.n_atom > 0) {
            const gpu::NlDev &nlg = c->nl64;
            const int npm = nlg.npm;
            gpu::NlDev nlb = nlg;
            DevBuf dvg;
            upload(dvg, std::vector<double>((size_t)nlg.n_atom * npm, S->dV), hs);
            nlb.gamma = dvg.p;
            const int e = S->cplx ? gpu::C128 : gpu::F64;
            const size_t es = S->cplx ? 16 : 8;
            const size_t per_col = (size_t)nlg.n_atom * npm;
            DevBuf dpsi, ga, gb;
            dpsi.get((size_t)S->ldpsi * N * es);
            ga.get(per_col * N * es);
            gb.get(per_col * N * es);
            const int vals = S->cplx ? 2 : 1;
            std::vector<double> ha(per_col * N * vals), hb(per_col * N * vals);
            const double spn_fac = (ns == 1 ? 2.0 : 1.0) * 2.0;
            const double k0[3] = {0, 0, 0};
            for (int s = 0; s < ns; ++s)
                for (int k = 0; k < nk; ++k) {
                    const int blk = s * nk + k;
                    const double *kp = S->cplx ? &S->kpts[3 * k] : k0;
                    if (S->cplx) {
                        const lynx_status ks = lynx_set_kpoint(c, kp);
                        if (ks != LYNX_OK) fail(ks, std::string("lynx_set_kpoint: ") + lynx_last_error(c));
                    }
                    const void *psi = static_cast<const char *>(S->psi.p) + (size_t)blk * S->ldpsi * N * es;
                    gpu::nl_gather(e, nlg, psi, S->ldpsi, ga.p, N, hs);
                    HIP_OK(hipMemcpyAsync(ha.data(), ga.p, ha.size() * sizeof(double), hipMemcpyDeviceToHost, hs));
                    const gpu::StencilGeom geo = geom_at(c, kp);
                    for (int a = 0; a < 3; ++a) {
                        gpu::scf_grad_block(S->cplx, static_cast<const double *>(psi), S->ldpsi, dpsi.as<double>(),
                                            S->ldpsi, N, stn.nx, stn.ny, stn.nz, a, stn.D1[a].data(), stn.fdn,
                                            geo.phase[a][0], geo.phase[a][1], hs);
                        gpu::nl_gather(e, nlb, dpsi.p, S->ldpsi, gb.p, N, hs);
                        HIP_OK(hipMemcpyAsync(hb.data(), gb.p, hb.size() * sizeof(double), hipMemcpyDeviceToHost,
                                              hs));
                        HIP_OK(hipStreamSynchronize(hs));
                        for (int j = 0; j < N; ++j) {
                            const double gn = S->occ[(size_t)blk * N + j];
                            if (std::fabs(gn) < 1e-15) continue;
                            const double w = spn_fac * S->kw[k] * gn;
                            for (int at = 0; at < natom; ++at) {
                                double fj = 0.0;
                                for (int q = 0; q < npm; ++q) {
                                    const size_t i = ((size_t)j * per_col + (size_t)at * npm + q) * vals;
                                    fj += vals == 2 ? ha[i] * hb[i] + ha[i + 1] * hb[i + 1] : ha[i] * hb[i];
                                }
                                f_nl[3 * (size_t)at + a] -= w * fj;
                            }
                        }
                    }
                }
        }
        // lattice-aligned components to Cartesian, F = U^-1 F_nc
        // (Forces.cpp:132-170), then the mean removed (Forces.cpp:1277-1288)
        auto to_cart = [&](std::vector<double> &f) {
            if (c->lat.orth) return;
            const double *U = c->lat.uvec_inv;
            for (int at = 0; at < natom; ++at) {
                const double x = f[3 * at], y = f[3 * at + 1], z = f[3 * at + 2];
                for (int i = 0; i < 3; ++i) f[3 * at + i] = U[3 * i] * x + U[3 * i + 1] * y + U[3 * i + 2] * z;
            }
        };
        std::vector<double> tot(3 * (size_t)natom);
        for (size_t i = 0; i < tot.size(); ++i) tot[i] = f_loc[i] + f_nl[i] + f_xc[i];
        to_cart(tot);
        to_cart(f_loc);
        to_cart(f_nl);
        to_cart(f_xc);
        for (int a = 0; a < 3; ++a) {
            double avg = 0.0;
            for (int at = 0; at < natom; ++at) avg += tot[3 * at + a];
            avg /= natom;
            for (int at = 0; at < natom; ++at) tot[3 * at + a] -= avg;
        }
        std::copy(tot.begin(), tot.end(), forces);
        if (parts) {
            std::copy(f_loc.begin(), f_loc.end(), parts);
            std::copy(f_nl.begin(), f_nl.end(), parts + 3 * natom);
            std::copy(f_xc.begin(), f_xc.end(), parts + 6 * natom);
        }
    });
    if (st != LYNX_OK) S->err = S->ctx->err;
    return st;
}

lynx_status lynx_scf_history(const lynx_scf *S, int cap, double *etotal, double *error, double *seconds,
                             int *precision, int *n)
{
    if (!S) return LYNX_E_ARG;
    const int m = (int)S->h_e.size();
    for (int i = 0; i < m && i < cap; ++i) {
        if (etotal) etotal[i] = S->h_e[i];
        if (error) error[i] = S->h_err[i];
        if (seconds) seconds[i] = i < (int)S->h_t.size() ? S->h_t[i] : 0.0;
        if (precision) precision[i] = S->h_prec[i];
    }
    if (n) *n = m;
    return LYNX_OK;
}

lynx_status lynx_scf_get(const lynx_scf *S, double *eig, double *occ, double *rho, double *veff)
{
    if (!S) return LYNX_E_ARG;
    lynx_scf *M = const_cast<lynx_scf *>(S);
    return guarded(M->ctx, [&] {
        if (eig) std::memcpy(eig, S->eig.data(), S->eig.size() * sizeof(double));
        if (occ) std::memcpy(occ, S->occ.data(), S->occ.size() * sizeof(double));
        const long n = S->nd;
        if (rho) {
            const double *src = S->p.nspin == 1 ? M->rho.as<double>() : M->rho.as<double>() + n;
            HIP_OK(hipMemcpyAsync(rho, src, (size_t)S->p.nspin * n * sizeof(double), hipMemcpyDeviceToHost,
                                  M->ctx->stream));
        }
        if (veff)
            HIP_OK(hipMemcpyAsync(veff, M->veff.p, (size_t)S->p.nspin * n * sizeof(double), hipMemcpyDeviceToHost,
                                  M->ctx->stream));
        HIP_OK(hipStreamSynchronize(M->ctx->stream));
    });
}

lynx_status lynx_scf_get_fields(const lynx_scf *S, double *phi, double *exc, double *vxc)
{
    if (!S) return LYNX_E_ARG;
    lynx_scf *M = const_cast<lynx_scf *>(S);
    return guarded(M->ctx, [&] {
        const size_t b = (size_t)S->nd * sizeof(double);
        if (phi) HIP_OK(hipMemcpyAsync(phi, M->phi.p, b, hipMemcpyDeviceToHost, M->ctx->stream));
        if (exc) HIP_OK(hipMemcpyAsync(exc, M->exc.p, b, hipMemcpyDeviceToHost, M->ctx->stream));
        if (vxc) HIP_OK(hipMemcpyAsync(vxc, M->vxc.p, b * S->p.nspin, hipMemcpyDeviceToHost, M->ctx->stream));
        HIP_OK(hipStreamSynchronize(M->ctx->stream));
    });
}

}  // extern "C"
