// Launchers of the SCF-tail kernels (kernels_scf.hip).  Internal.
#pragma once

#include "common.hpp"

#include <hip/hip_runtime.h>

namespace lynx {
namespace gpu {

constexpr int kMaxFdn = 12;               // largest FD radius of the gradient kernel
constexpr int kScfPartialBlocks = 512;    // partial sums per reduction

// exchange-correlation kind: 0 LDA (PZ81), 1 LDA (PW92), 2 GGA (PBE)
void scf_xc_unpol(int kind, const double *rho, const double *sigma, double *exc, double *vrho, double *v2, long n,
                  hipStream_t st);
void scf_xc_pol(int kind, const double *ru, const double *rd, const double *suu, const double *sud, const double *sdd,
                double *exc, double *vu, double *vd, double *vsuu, double *vsud, double *vsdd, long n, hipStream_t st);
// rho += sum_n w[n] |psi_n|^2 (w on the device)
void scf_density(bool cplx, const void *psi, long ld, int ncol, const double *w, double *rho, long nd, hipStream_t st);
// y (+)= scale * D_dir x, periodic central differences with coefficients d1[0..R]
void scf_grad(const double *x, double *y, int dir, int nx, int ny, int nz, const double *d1, int R, double scale,
              bool accumulate, hipStream_t st);
void scf_axpby(double *y, const double *x, double a, double b, long n, hipStream_t st);
void scf_fill(double *y, double v, long n, hipStream_t st);
void scf_scale(double *x, double s, long n, hipStream_t st);
void scf_clamp_floor(double *x, double lo, long n, hipStream_t st);   // x = max(x, lo)
void scf_mul(double *out, const double *x, const double *y, double a, bool accumulate, long n, hipStream_t st);
void scf_aar_pre(double *x, double *xo, double *f, double *fo, const double *r, double *X, double *F, int col,
                 double m, double omega, bool richardson, long n, hipStream_t st);
void scf_aar_anderson(double *x, const double *xo, const double *f, const double *X, const double *F,
                      const double *g, int cols, double beta, long n, hipStream_t st);
void scf_mix_residual(const double *x, const double *g, const double *xkm1, double *fk, double *R, double *F, int col,
                      long n, hipStream_t st);
void scf_mix_wavg(const double *x, const double *f, const double *R, const double *F, const double *g, int cols,
                  double *xw, double *fw, long n, hipStream_t st);
// npairs dot products a[k] . b[k]; returns the number of blocks whose
// partials (block-major, npairs per block) are in part
int scf_multidot(const double *const *a, const double *const *b, int npairs, long n, double *part, hipStream_t st);
int scf_clamp_density(double *x, int ncol, long nd, double *part, hipStream_t st);
void scf_gga_flux(const double M[9], const double *c1, const double *const g[3], const double *c2,
                  const double *const h[3], double *const f[3], long n, hipStream_t st);
void scf_sigma(const double M[9], const double *const a[3], const double *const b[3], double *s, long n, hipStream_t st);
// multiply the n interleaved complex Fourier coefficients spec by a factor
// of the symbol L of -lap at each: kind 0 (Poisson), norm / L with the
// q = 0 coefficient set to zero; kind 1 (Kerker), norm amix (L + c1) / (L + c2)
// y[:, n] = D_dir x[:, n] for ncol columns of a real (cplx false) or
// interleaved complex block: periodic FD first derivative along lattice
// axis dir with coefficients d1[1..R] (R <= n_dir); complex values that wrap
// across the upper (lower) face are multiplied by ph_hi (ph_lo)
void scf_grad_block(bool cplx, const double *x, long ldx, double *y, long ldy, int ncol, int nx, int ny, int nz,
                    int dir, const double *d1, int R, const double ph_lo[2], const double ph_hi[2], hipStream_t st);
void scf_spectral_scale(double *spec, const double *sym, long n, int kind, double amix, double c1, double c2,
                        double norm, hipStream_t st);

}  // namespace gpu
}  // namespace lynx
