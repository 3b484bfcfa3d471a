// HIP kernels of the SCF tail around the CheFSI path, all FP64: electron
// density from the orbital blocks, exchange-correlation (LDA: Slater
// exchange with Perdew-Zunger 1981 or Perdew-Wang 1992 correlation; GGA:
// PBE), gradients along lattice directions, and the vector steps of the
// Anderson-Richardson (AAR) linear solver and the Pulay mixer.  Grid-sized
// and memory-bound except XC (a few hundred flops per point).  The XC
// derivatives come from forward-mode dual numbers on the energy density, so
// the potential is the exact derivative of the evaluated energy.
#include "kernels_scf.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

namespace lynx {
namespace gpu {

namespace {

constexpr int kNT = 256;

inline int blocks_for(long n, int per_block = kNT) { return (int)std::min<long>((n + per_block - 1) / per_block, 65535L * 8); }

void check_launch(const char *what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) fail(2, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------- dual numbers
template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> var(double v, int k)
{
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = i == k ? 1.0 : 0.0;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> cst(double v)
{
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = 0.0;
    return r;
}
// f(a) with f(a.v) = fv and f'(a.v) = fd
template <int N>
__device__ __forceinline__ Dual<N> chain(const Dual<N> &a, double fv, double fd)
{
    Dual<N> r;
    r.v = fv;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = fd * a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    const double ib = 1.0 / b.v;
    r.v = a.v * ib;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double b) { Dual<N> r = a; r.v += b; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator+(double b, const Dual<N> &a) { return a + b; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double b) { Dual<N> r = a; r.v -= b; return r; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(double b, const Dual<N> &a)
{
    Dual<N> r;
    r.v = b - a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) { return 0.0 - a; }
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double b, const Dual<N> &a) { return a * b; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double b) { return a * (1.0 / b); }
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N> &b) { return cst<N>(a) / b; }
template <int N>
__device__ __forceinline__ Dual<N> dlog(const Dual<N> &a) { return chain(a, log(a.v), 1.0 / a.v); }
template <int N>
__device__ __forceinline__ Dual<N> dexp(const Dual<N> &a) { const double e = exp(a.v); return chain(a, e, e); }
template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N> &a) { const double s = sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int N>
__device__ __forceinline__ Dual<N> dcbrt(const Dual<N> &a) { const double c = cbrt(a.v); return chain(a, c, 1.0 / (3.0 * c * c)); }

// ----------------------------------------------------- XC energy densities
// Each returns rho * eps_xc (energy per volume) of its inputs.
constexpr double kPi = 3.14159265358979323846;

// Slater exchange of an unpolarized density: rho eps_x = -(3/4)(3/pi)^(1/3) rho^(4/3)
template <int N>
__device__ Dual<N> fx_slater(const Dual<N> &rho)
{
    const double cx = -0.75 * cbrt(3.0 / kPi);
    return cx * rho * dcbrt(rho);
}

// Perdew-Zunger 1981 correlation per particle, unpolarized (pol = 0) or
// fully polarized (pol = 1).
template <int N>
__device__ Dual<N> ec_pz(const Dual<N> &rs, int pol)
{
    const double g = pol ? -0.0843 : -0.1423, b1 = pol ? 1.3981 : 1.0529, b2 = pol ? 0.2611 : 0.3334;
    const double A = pol ? 0.01555 : 0.0311, B = pol ? -0.0269 : -0.048, C = pol ? 0.0007 : 0.0020,
                 D = pol ? -0.0048 : -0.0116;
    if (rs.v >= 1.0) return g / (1.0 + b1 * dsqrt(rs) + b2 * rs);
    const Dual<N> lr = dlog(rs);
    return A * lr + B + C * rs * lr + D * rs;
}

// Perdew-Wang 1992: G(rs; A, alpha1, beta1..beta4), p = 1.
struct PwSet {
    double A, a1, b1, b2, b3, b4;
};
template <int N>
__device__ Dual<N> pw_g(const Dual<N> &rs, const PwSet &p)
{
    const Dual<N> srs = dsqrt(rs);
    const Dual<N> den = 2.0 * p.A * (p.b1 * srs + p.b2 * rs + p.b3 * rs * srs + p.b4 * rs * rs);
    return -2.0 * p.A * (1.0 + p.a1 * rs) * dlog(1.0 + 1.0 / den);
}
// the three fits (unpolarized, polarized, -spin stiffness) and f''(0); mod
// selects the more precise A values and f''(0) PBE uses
__device__ __forceinline__ PwSet pw_set(int which, bool mod)
{
    if (which == 0) return {mod ? 0.0310907 : 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294};
    if (which == 1) return {mod ? 0.01554535 : 0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517};
    return {mod ? 0.0168869 : 0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671};
}
__device__ __forceinline__ double pw_fz20(bool mod) { return mod ? 1.709920934161365617563962776245 : 1.709921; }

// f(zeta) = ((1+z)^(4/3) + (1-z)^(4/3) - 2) / (2^(4/3) - 2)
template <int N>
__device__ Dual<N> fzeta(const Dual<N> &z)
{
    const Dual<N> a = 1.0 + z, b = 1.0 - z;
    return (a * dcbrt(a) + b * dcbrt(b) - 2.0) / (2.0 * cbrt(2.0) - 2.0);
}

template <int N>
__device__ Dual<N> rs_of(const Dual<N> &rho) { return dcbrt(3.0 / (4.0 * kPi * rho)); }

// correlation per particle at (rs, zeta): PZ81 (kind 0) or PW92 (kind 1 / mod)
template <int N>
__device__ Dual<N> ec_unif(const Dual<N> &rs, const Dual<N> *zeta, int kind, bool mod)
{
    if (kind == 0) {
        const Dual<N> eu = ec_pz(rs, 0);
        if (!zeta) return eu;
        return eu + fzeta(*zeta) * (ec_pz(rs, 1) - eu);
    }
    const Dual<N> e0 = pw_g(rs, pw_set(0, mod));
    if (!zeta) return e0;
    const Dual<N> e1 = pw_g(rs, pw_set(1, mod));
    const Dual<N> ac = -pw_g(rs, pw_set(2, mod));
    const Dual<N> f = fzeta(*zeta);
    const Dual<N> z2 = (*zeta) * (*zeta);
    const Dual<N> z4 = z2 * z2;
    return e0 + ac * f * (1.0 - z4) / pw_fz20(mod) + (e1 - e0) * f * z4;
}

// PBE (Perdew, Burke & Ernzerhof 1996) constants
constexpr double kKappa = 0.8040;
constexpr double kBeta = 0.06672455060314922;
constexpr double kMu = kBeta * kPi * kPi / 3.0;

// PBE exchange of an unpolarized density with sigma = |grad rho|^2
template <int N>
__device__ Dual<N> fx_pbe(const Dual<N> &rho, const Dual<N> &sigma)
{
    const double c = 4.0 * cbrt(9.0 * kPi * kPi * kPi * kPi);   // 4 (3 pi^2)^(2/3)
    const Dual<N> r13 = dcbrt(rho);
    const Dual<N> s2 = sigma / (c * rho * rho * r13 * r13);
    const Dual<N> fx = 1.0 + kKappa - kKappa / (1.0 + kMu * s2 / kKappa);
    return fx_slater(rho) * fx;
}

// PBE correlation: rho (eps_c^PW(rs, zeta) + H(rs, zeta, t)), sigma the
// squared gradient of the total density; zeta NULL for unpolarized
template <int N>
__device__ Dual<N> fc_pbe(const Dual<N> &rho, const Dual<N> &sigma, const Dual<N> *zeta)
{
    const double gam = (1.0 - log(2.0)) / (kPi * kPi);
    const Dual<N> rs = rs_of(rho);
    const Dual<N> ec = ec_unif(rs, zeta, 1, true);
    Dual<N> phi = cst<N>(1.0);
    if (zeta) {
        const Dual<N> a = dcbrt(1.0 + *zeta), b = dcbrt(1.0 - *zeta);
        phi = 0.5 * (a * a + b * b);
    }
    const Dual<N> phi3 = phi * phi * phi;
    const Dual<N> kf = dcbrt(3.0 * kPi * kPi * rho);
    const Dual<N> ks2 = 4.0 * kf / kPi;
    const Dual<N> t2 = sigma / (4.0 * phi * phi * ks2 * rho * rho);
    const Dual<N> A = (kBeta / gam) / (dexp(-ec / (gam * phi3)) - 1.0);
    const Dual<N> At2 = A * t2;
    const Dual<N> H = gam * phi3 * dlog(1.0 + (kBeta / gam) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
    return rho * (ec + H);
}

constexpr double kRhoMin = 1e-15;   // below this density the XC outputs are zero

// unpolarized LDA (kind 0 PZ, 1 PW) or PBE (kind 2)
__global__ void __launch_bounds__(kNT) k_xc_unpol(int kind, const double *__restrict__ rho,
                                                   const double *__restrict__ sigma, double *__restrict__ exc,
                                                   double *__restrict__ vrho, double *__restrict__ v2, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double r = rho[i];
        if (!(r > kRhoMin)) {
            exc[i] = 0.0;
            vrho[i] = 0.0;
            if (v2) v2[i] = 0.0;
            continue;
        }
        if (kind == 2) {
            const double s = fmax(sigma[i], 1e-40);
            const Dual<2> R = var<2>(r, 0), S = var<2>(s, 1);
            const Dual<2> f = fx_pbe(R, S) + fc_pbe<2>(R, S, nullptr);
            exc[i] = f.v / r;
            vrho[i] = f.d[0];
            v2[i] = 2.0 * f.d[1];
        } else {
            const Dual<1> R = var<1>(r, 0);
            const Dual<1> f = fx_slater(R) + R * ec_unif<1>(rs_of(R), nullptr, kind, false);
            exc[i] = f.v / r;
            vrho[i] = f.d[0];
            if (v2) v2[i] = 0.0;
        }
    }
}

// spin-polarized LDA (kind 0, 1) or PBE (kind 2).  Inputs: up and down
// densities, and for PBE sigma_uu, sigma_ud, sigma_dd.  Outputs: exc per
// particle of the total density, vrho up/down, and for PBE the sigma
// derivatives (vs_uu, vs_ud, vs_dd).
__global__ void __launch_bounds__(kNT) k_xc_pol(int kind, const double *__restrict__ ru, const double *__restrict__ rd,
                                                 const double *__restrict__ suu, const double *__restrict__ sud,
                                                 const double *__restrict__ sdd, double *__restrict__ exc,
                                                 double *__restrict__ vu, double *__restrict__ vd,
                                                 double *__restrict__ vsuu, double *__restrict__ vsud,
                                                 double *__restrict__ vsdd, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double a = ru[i], b = rd[i];
        if (!(a + b > kRhoMin)) {
            exc[i] = vu[i] = vd[i] = 0.0;
            if (vsuu) vsuu[i] = vsud[i] = vsdd[i] = 0.0;
            continue;
        }
        if (kind == 2) {
            const Dual<5> U = var<5>(a, 0), D = var<5>(b, 1);
            const Dual<5> Suu = var<5>(fmax(suu[i], 1e-40), 2), Sud = var<5>(sud[i], 3), Sdd = var<5>(fmax(sdd[i], 1e-40), 4);
            const Dual<5> R = U + D;
            const Dual<5> z = (U - D) / R;
            const Dual<5> fx = 0.5 * fx_pbe(2.0 * U, 4.0 * Suu) + 0.5 * fx_pbe(2.0 * D, 4.0 * Sdd);
            const Dual<5> f = fx + fc_pbe(R, Suu + 2.0 * Sud + Sdd, &z);
            exc[i] = f.v / (a + b);
            vu[i] = f.d[0];
            vd[i] = f.d[1];
            vsuu[i] = f.d[2];
            vsud[i] = f.d[3];
            vsdd[i] = f.d[4];
        } else {
            const Dual<2> U = var<2>(a, 0), D = var<2>(b, 1);
            const Dual<2> R = U + D;
            const Dual<2> z = (U - D) / R;
            const Dual<2> f = 0.5 * fx_slater(2.0 * U) + 0.5 * fx_slater(2.0 * D) + R * ec_unif(rs_of(R), &z, kind, false);
            exc[i] = f.v / (a + b);
            vu[i] = f.d[0];
            vd[i] = f.d[1];
            if (vsuu) vsuu[i] = vsud[i] = vsdd[i] = 0.0;
        }
    }
}

// ------------------------------------------------------------------ density
// rho[i] += sum_n w[n] |psi[i + n ld]|^2
template <bool CPLX>
__global__ void __launch_bounds__(kNT) k_density(const double *__restrict__ psi, long ld, int ncol,
                                                  const double *__restrict__ w, double *__restrict__ rho, long nd)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nd; i += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int n = 0; n < ncol; ++n) {
            const double wn = w[n];
            if (wn == 0.0) continue;
            if (CPLX) {
                const double2 v = reinterpret_cast<const double2 *>(psi)[i + n * ld];
                s += wn * (v.x * v.x + v.y * v.y);
            } else {
                const double v = psi[i + n * ld];
                s += wn * v * v;
            }
        }
        rho[i] += s;
    }
}

// ------------------------------------------------------------------ gradient
// y = sum_p d1[p] (x(+p) - x(-p)) along lattice direction dir, periodic
struct GradArgs {
    int nx, ny, nz, R, dir;
    double d1[kMaxFdn + 1];
};
__global__ void __launch_bounds__(kNT) k_grad(GradArgs g, const double *__restrict__ x, double *__restrict__ y,
                                               double scale, int accumulate)
{
    const long nd = (long)g.nx * g.ny * g.nz;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nd; i += (long)gridDim.x * blockDim.x) {
        const int ix = (int)(i % g.nx);
        const int iy = (int)((i / g.nx) % g.ny);
        const int iz = (int)(i / ((long)g.nx * g.ny));
        const int n = g.dir == 0 ? g.nx : (g.dir == 1 ? g.ny : g.nz);
        const int c = g.dir == 0 ? ix : (g.dir == 1 ? iy : iz);
        const long stride = g.dir == 0 ? 1 : (g.dir == 1 ? g.nx : (long)g.nx * g.ny);
        const long base = i - (long)c * stride;
        double s = 0.0;
        for (int p = 1; p <= g.R; ++p) {
            int cp = c + p, cm = c - p;
            cp -= (cp >= n) ? n * ((cp) / n) : 0;
            while (cm < 0) cm += n;
            s += g.d1[p] * (x[base + cp * stride] - x[base + cm * stride]);
        }
        y[i] = accumulate ? y[i] + scale * s : scale * s;
    }
}

// ---------------------------------------------------- small vector kernels
__global__ void k_axpby(double *__restrict__ y, const double *__restrict__ x, double a, double b, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}
struct GradCoef {
    double d1[kMaxFdn + 1];
};

template <bool CPLX>
__global__ void __launch_bounds__(kNT) k_grad_block(const double *__restrict__ x, long ldx, double *__restrict__ y,
                                                     long ldy, int ncol, int nx, int ny, int nz, int dir,
                                                     GradCoef c, int R, double lo_re, double lo_im, double hi_re,
                                                     double hi_im)
{
    const long nd = (long)nx * ny * nz;
    const long total = nd * ncol;
    const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
    const long stride = dir == 0 ? 1 : (dir == 1 ? (long)nx : (long)nx * ny);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long col = t / nd, g = t - col * nd;
        const int i = (int)(dir == 0 ? g % nx : (dir == 1 ? (g / nx) % ny : g / ((long)nx * ny)));
        const long base = g - (long)i * stride;
        if (CPLX) {
            const double2 *xc = reinterpret_cast<const double2 *>(x) + col * ldx;
            double re = 0.0, im = 0.0;
            for (int p = 1; p <= R; ++p) {
                const int ip = i + p, im_ = i - p;
                double2 u = xc[base + (long)(ip >= n ? ip - n : ip) * stride];
                double2 v = xc[base + (long)(im_ < 0 ? im_ + n : im_) * stride];
                if (ip >= n) u = make_double2(u.x * hi_re - u.y * hi_im, u.x * hi_im + u.y * hi_re);
                if (im_ < 0) v = make_double2(v.x * lo_re - v.y * lo_im, v.x * lo_im + v.y * lo_re);
                re += c.d1[p] * (u.x - v.x);
                im += c.d1[p] * (u.y - v.y);
            }
            reinterpret_cast<double2 *>(y)[col * ldy + g] = make_double2(re, im);
        } else {
            const double *xr = x + col * ldx;
            double acc = 0.0;
            for (int p = 1; p <= R; ++p) {
                const int ip = i + p, im_ = i - p;
                acc += c.d1[p] * (xr[base + (long)(ip >= n ? ip - n : ip) * stride] -
                                  xr[base + (long)(im_ < 0 ? im_ + n : im_) * stride]);
            }
            y[col * ldy + g] = acc;
        }
    }
}

__global__ void k_spectral_scale(double2 *__restrict__ spec, const double *__restrict__ sym, long n, int kind,
                                 double amix, double c1, double c2, double norm)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double L = sym[i];
        const double g = kind == 0 ? (i == 0 ? 0.0 : norm / L) : norm * amix * (L + c1) / (L + c2);
        double2 v = spec[i];
        v.x *= g;
        v.y *= g;
        spec[i] = v;
    }
}

__global__ void k_fill(double *__restrict__ y, double v, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = v;
}

// AAR step, first half: f = m r; history X(:,col) = x - x_old,
// F(:,col) = f - f_old (col >= 0); x_old = x; f_old = f; and for a
// Richardson step x += omega f
__global__ void __launch_bounds__(kNT) k_aar_pre(double *__restrict__ x, double *__restrict__ xo,
                                                  double *__restrict__ f, double *__restrict__ fo,
                                                  const double *__restrict__ r, double *__restrict__ X,
                                                  double *__restrict__ F, long col_off, double m, double omega,
                                                  int richardson, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double fi = m * r[i];
        const double xi = x[i];
        if (col_off >= 0) {
            X[col_off + i] = xi - xo[i];
            F[col_off + i] = fi - fo[i];
        }
        xo[i] = xi;
        fo[i] = fi;
        f[i] = fi;
        if (richardson) x[i] = xi + omega * fi;
    }
}

struct Coef8 {
    double g[8];
};

// Anderson extrapolation: x = x_old + beta f - sum_j g_j (X_j + beta F_j)
__global__ void __launch_bounds__(kNT) k_aar_anderson(double *__restrict__ x, const double *__restrict__ xo,
                                                       const double *__restrict__ f, const double *__restrict__ X,
                                                       const double *__restrict__ F, Coef8 c, int cols, double beta,
                                                       long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double v = xo[i] + beta * f[i];
        for (int j = 0; j < cols; ++j) v -= c.g[j] * (X[j * n + i] + beta * F[j * n + i]);
        x[i] = v;
    }
}

// Pulay residual and history: f_new = g - x; R(:,col) = x - x_km1,
// F(:,col) = f_new - f_k (col >= 0); f_k = f_new
__global__ void __launch_bounds__(kNT) k_mix_residual(const double *__restrict__ x, const double *__restrict__ g,
                                                       const double *__restrict__ xkm1, double *__restrict__ fk,
                                                       double *__restrict__ R, double *__restrict__ F, long col_off,
                                                       long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double fn = g[i] - x[i];
        if (col_off >= 0) {
            R[col_off + i] = x[i] - xkm1[i];
            F[col_off + i] = fn - fk[i];
        }
        fk[i] = fn;
    }
}

// x_wavg = x - sum g_j R_j, f_wavg = f - sum g_j F_j
__global__ void __launch_bounds__(kNT) k_mix_wavg(const double *__restrict__ x, const double *__restrict__ f,
                                                   const double *__restrict__ R, const double *__restrict__ F,
                                                   Coef8 c, int cols, double *__restrict__ xw,
                                                   double *__restrict__ fw, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double a = x[i], b = f[i];
        for (int j = 0; j < cols; ++j) {
            a -= c.g[j] * R[j * n + i];
            b -= c.g[j] * F[j * n + i];
        }
        xw[i] = a;
        fw[i] = b;
    }
}

// Many dot products in one pass: out[blk * nd + k] = partial of
// a_k . b_k over this block's slice, for k < nd pairs (pointers in a table)
struct DotTable {
    const double *a[40];
    const double *b[40];
    int n;
};
__global__ void __launch_bounds__(kNT) k_multidot(DotTable t, long n, double *__restrict__ part)
{
    __shared__ double red[kNT / 64][40];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = 0; k < t.n; ++k) {
        double s = 0.0;
        for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
            s += t.a[k][i] * t.b[k][i];
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < t.n) {
        double s = 0.0;
        for (int w = 0; w < kNT / 64; ++w) s += red[w][threadIdx.x];
        part[(long)blockIdx.x * t.n + threadIdx.x] = s;
    }
}

// density constraint of the mixer: one column (rho >= 0) or [total | mag]
// (up, down >= 0); writes per-block partial sums of the total density
__global__ void __launch_bounds__(kNT) k_clamp_density(double *__restrict__ x, int ncol, long nd,
                                                        double *__restrict__ part)
{
    __shared__ double red[kNT / 64];
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nd; i += (long)gridDim.x * blockDim.x) {
        if (ncol == 2) {
            double up = 0.5 * (x[i] + x[nd + i]), dn = 0.5 * (x[i] - x[nd + i]);
            if (up < 0.0) up = 0.0;
            if (dn < 0.0) dn = 0.0;
            x[i] = up + dn;
            x[nd + i] = up - dn;
        } else if (x[i] < 0.0) {
            x[i] = 0.0;
        }
        s += x[i];
    }
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kNT / 64; ++w) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ void k_floor(double *__restrict__ x, double lo, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (x[i] < lo) x[i] = lo;
}

__global__ void k_scale(double *__restrict__ x, double s, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] *= s;
}

// GGA flux f_a = v2 (M grad rho)_a with M = lapc_T (identity for orthogonal
// cells), for one spin channel: f = 2 vs_self grad rho_self + vs_cross grad rho_other
struct FluxArgs {
    double M[9];
};
__global__ void __launch_bounds__(kNT) k_gga_flux(FluxArgs m, const double *__restrict__ c1, const double *__restrict__ gx,
                                                   const double *__restrict__ gy, const double *__restrict__ gz,
                                                   const double *__restrict__ c2, const double *__restrict__ hx,
                                                   const double *__restrict__ hy, const double *__restrict__ hz,
                                                   double *__restrict__ fx, double *__restrict__ fy,
                                                   double *__restrict__ fz, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double ax = c1[i] * gx[i], ay = c1[i] * gy[i], az = c1[i] * gz[i];
        if (c2) {
            ax += c2[i] * hx[i];
            ay += c2[i] * hy[i];
            az += c2[i] * hz[i];
        }
        fx[i] = m.M[0] * ax + m.M[1] * ay + m.M[2] * az;
        fy[i] = m.M[3] * ax + m.M[4] * ay + m.M[5] * az;
        fz[i] = m.M[6] * ax + m.M[7] * ay + m.M[8] * az;
    }
}

// sigma = (M ga) . gb
__global__ void __launch_bounds__(kNT) k_sigma(FluxArgs m, const double *__restrict__ ax, const double *__restrict__ ay,
                                                const double *__restrict__ az, const double *__restrict__ bx,
                                                const double *__restrict__ by, const double *__restrict__ bz,
                                                double *__restrict__ s, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double px = m.M[0] * ax[i] + m.M[1] * ay[i] + m.M[2] * az[i];
        const double py = m.M[3] * ax[i] + m.M[4] * ay[i] + m.M[5] * az[i];
        const double pz = m.M[6] * ax[i] + m.M[7] * ay[i] + m.M[8] * az[i];
        s[i] = px * bx[i] + py * by[i] + pz * bz[i];
    }
}

// element-wise out = a * x * y (+ out)
__global__ void k_mul(double *__restrict__ out, const double *__restrict__ x, const double *__restrict__ y, double a,
                      int accumulate, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = accumulate ? out[i] + a * x[i] * y[i] : a * x[i] * y[i];
}

}  // namespace

// ---------------------------------------------------------------- launchers
void scf_xc_unpol(int kind, const double *rho, const double *sigma, double *exc, double *vrho, double *v2, long n,
                  hipStream_t st)
{
    k_xc_unpol<<<blocks_for(n), kNT, 0, st>>>(kind, rho, sigma, exc, vrho, v2, n);
    check_launch("k_xc_unpol");
}

void scf_xc_pol(int kind, const double *ru, const double *rd, const double *suu, const double *sud, const double *sdd,
                double *exc, double *vu, double *vd, double *vsuu, double *vsud, double *vsdd, long n, hipStream_t st)
{
    k_xc_pol<<<blocks_for(n), kNT, 0, st>>>(kind, ru, rd, suu, sud, sdd, exc, vu, vd, vsuu, vsud, vsdd, n);
    check_launch("k_xc_pol");
}

void scf_density(bool cplx, const void *psi, long ld, int ncol, const double *w, double *rho, long nd, hipStream_t st)
{
    if (cplx)
        k_density<true><<<blocks_for(nd), kNT, 0, st>>>(static_cast<const double *>(psi), ld, ncol, w, rho, nd);
    else
        k_density<false><<<blocks_for(nd), kNT, 0, st>>>(static_cast<const double *>(psi), ld, ncol, w, rho, nd);
    check_launch("k_density");
}

void scf_grad(const double *x, double *y, int dir, int nx, int ny, int nz, const double *d1, int R, double scale,
              bool accumulate, hipStream_t st)
{
    if (R > kMaxFdn) fail(8, "gradient stencil radius above the kernel's maximum");
    GradArgs g{};
    g.nx = nx; g.ny = ny; g.nz = nz; g.R = R; g.dir = dir;
    for (int p = 0; p <= R; ++p) g.d1[p] = d1[p];
    const long nd = (long)nx * ny * nz;
    k_grad<<<blocks_for(nd), kNT, 0, st>>>(g, x, y, scale, accumulate ? 1 : 0);
    check_launch("k_grad");
}

void scf_axpby(double *y, const double *x, double a, double b, long n, hipStream_t st)
{
    k_axpby<<<blocks_for(n), kNT, 0, st>>>(y, x, a, b, n);
    check_launch("k_axpby");
}

void scf_grad_block(bool cplx, const double *x, long ldx, double *y, long ldy, int ncol, int nx, int ny, int nz,
                    int dir, const double *d1, int R, const double ph_lo[2], const double ph_hi[2], hipStream_t st)
{
    const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
    if (R < 1 || R > kMaxFdn || R > n || dir < 0 || dir > 2)
        fail(8, "scf_grad_block: stencil radius outside 1..min(12, grid length)");   // LYNX_E_UNSUPPORTED
    GradCoef c{};
    for (int p = 0; p <= R; ++p) c.d1[p] = d1[p];
    const long total = (long)nx * ny * nz * ncol;
    if (cplx)
        k_grad_block<true><<<blocks_for(total), kNT, 0, st>>>(x, ldx, y, ldy, ncol, nx, ny, nz, dir, c, R, ph_lo[0],
                                                              ph_lo[1], ph_hi[0], ph_hi[1]);
    else
        k_grad_block<false><<<blocks_for(total), kNT, 0, st>>>(x, ldx, y, ldy, ncol, nx, ny, nz, dir, c, R, 0, 0, 0,
                                                               0);
    check_launch("k_grad_block");
}

void scf_spectral_scale(double *spec, const double *sym, long n, int kind, double amix, double c1, double c2,
                        double norm, hipStream_t st)
{
    k_spectral_scale<<<blocks_for(n), kNT, 0, st>>>(reinterpret_cast<double2 *>(spec), sym, n, kind, amix, c1, c2,
                                                    norm);
    check_launch("k_spectral_scale");
}

void scf_fill(double *y, double v, long n, hipStream_t st)
{
    k_fill<<<blocks_for(n), kNT, 0, st>>>(y, v, n);
    check_launch("k_fill");
}

void scf_scale(double *x, double s, long n, hipStream_t st)
{
    k_scale<<<blocks_for(n), kNT, 0, st>>>(x, s, n);
    check_launch("k_scale");
}

void scf_clamp_floor(double *x, double lo, long n, hipStream_t st)
{
    k_floor<<<blocks_for(n), kNT, 0, st>>>(x, lo, n);
    check_launch("k_floor");
}

void scf_mul(double *out, const double *x, const double *y, double a, bool accumulate, long n, hipStream_t st)
{
    k_mul<<<blocks_for(n), kNT, 0, st>>>(out, x, y, a, accumulate ? 1 : 0, n);
    check_launch("k_mul");
}

void scf_aar_pre(double *x, double *xo, double *f, double *fo, const double *r, double *X, double *F, int col,
                 double m, double omega, bool richardson, long n, hipStream_t st)
{
    k_aar_pre<<<blocks_for(n), kNT, 0, st>>>(x, xo, f, fo, r, X, F, col >= 0 ? (long)col * n : -1L, m, omega,
                                             richardson ? 1 : 0, n);
    check_launch("k_aar_pre");
}

void scf_aar_anderson(double *x, const double *xo, const double *f, const double *X, const double *F,
                      const double *g, int cols, double beta, long n, hipStream_t st)
{
    Coef8 c{};
    for (int j = 0; j < cols && j < 8; ++j) c.g[j] = g[j];
    k_aar_anderson<<<blocks_for(n), kNT, 0, st>>>(x, xo, f, X, F, c, cols, beta, n);
    check_launch("k_aar_anderson");
}

void scf_mix_residual(const double *x, const double *g, const double *xkm1, double *fk, double *R, double *F, int col,
                      long n, hipStream_t st)
{
    k_mix_residual<<<blocks_for(n), kNT, 0, st>>>(x, g, xkm1, fk, R, F, col >= 0 ? (long)col * n : -1L, n);
    check_launch("k_mix_residual");
}

void scf_mix_wavg(const double *x, const double *f, const double *R, const double *F, const double *g, int cols,
                  double *xw, double *fw, long n, hipStream_t st)
{
    Coef8 c{};
    for (int j = 0; j < cols && j < 8; ++j) c.g[j] = g[j];
    k_mix_wavg<<<blocks_for(n), kNT, 0, st>>>(x, f, R, F, c, cols, xw, fw, n);
    check_launch("k_mix_wavg");
}

int scf_multidot(const double *const *a, const double *const *b, int npairs, long n, double *part, hipStream_t st)
{
    if (npairs > 40) fail(1, "too many dot products in one pass");
    DotTable t{};
    for (int k = 0; k < npairs; ++k) {
        t.a[k] = a[k];
        t.b[k] = b[k];
    }
    t.n = npairs;
    const int nb = (int)std::min<long>(kScfPartialBlocks, (n + kNT - 1) / kNT);
    k_multidot<<<std::max(nb, 1), kNT, 0, st>>>(t, n, part);
    check_launch("k_multidot");
    return std::max(nb, 1);
}

int scf_clamp_density(double *x, int ncol, long nd, double *part, hipStream_t st)
{
    const int nb = (int)std::min<long>(kScfPartialBlocks, (nd + kNT - 1) / kNT);
    k_clamp_density<<<std::max(nb, 1), kNT, 0, st>>>(x, ncol, nd, part);
    check_launch("k_clamp_density");
    return std::max(nb, 1);
}

void scf_gga_flux(const double M[9], const double *c1, const double *const g[3], const double *c2,
                  const double *const h[3], double *const f[3], long n, hipStream_t st)
{
    FluxArgs m{};
    for (int i = 0; i < 9; ++i) m.M[i] = M[i];
    k_gga_flux<<<blocks_for(n), kNT, 0, st>>>(m, c1, g[0], g[1], g[2], c2, c2 ? h[0] : nullptr, c2 ? h[1] : nullptr,
                                              c2 ? h[2] : nullptr, f[0], f[1], f[2], n);
    check_launch("k_gga_flux");
}

void scf_sigma(const double M[9], const double *const a[3], const double *const b[3], double *s, long n, hipStream_t st)
{
    FluxArgs m{};
    for (int i = 0; i < 9; ++i) m.M[i] = M[i];
    k_sigma<<<blocks_for(n), kNT, 0, st>>>(m, a[0], a[1], a[2], b[0], b[1], b[2], s, n);
    check_launch("k_sigma");
}

}  // namespace gpu
}  // namespace lynx
