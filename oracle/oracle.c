/*
 * oracle.c -- CPU restatement of LYNX's Chebyshev-filtered subspace iteration
 * hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the checker, not the product.  Only tests/, the smoke() entry
 * and bench.py's CPU-baseline leg may load it.  Every routine follows the
 * reference algorithm step for step, in the reference's summation order, so
 * that the HIP path can be compared against it on identical inputs.
 *
 * Reference files followed (paths relative to the LYNX source tree):
 *   halo / Bloch wrap ........ src/parallel/HaloExchange.cpp:34-82, 104-189
 *   stencil + diag ........... src/operators/Hamiltonian.cpp:126-174 (orth),
 *                              176-267 (non-orth)
 *   nonlocal projector ....... src/operators/NonlocalProjector.cpp:121-227
 *                              (Gamma), 229-347 (k-point)
 *   H apply .................. src/operators/Hamiltonian.cpp:64-106
 *   Chebyshev filter ......... src/solvers/EigenSolver.cpp:558-619
 *   Lanczos bounds ........... src/solvers/Lanczos.hpp:104-198
 *
 * Complex arrays are interleaved (re, im) doubles, i.e. C99 double complex.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zd;

/* Operator description shared by every routine below.  All arrays are owned
 * by the caller (the Python side keeps them alive). */
typedef struct {
    int nx, ny, nz, fdn, orth;
    const double *cx, *cy, *cz;        /* D2 coefficients per axis, fdn+1   */
    const double *cxy, *cxz, *cyz;     /* mixed D2 coefficients (non-orth)  */
    const double *d1y, *d1z;           /* D1 coefficients (non-orth)        */
    /* nonlocal projector, one entry per influence image, reference order */
    int n_img;
    const int *img_atom;               /* physical (type-major) atom index  */
    const int *img_ndc;                /* points in the image's rc sphere   */
    const int *img_gpos_off;           /* offset into gpos                  */
    const int *gpos;                   /* flat grid indices                 */
    const int *img_chi_off;            /* offset into chi (ndc x nproj)     */
    const int *img_nproj;
    const double *chi;                 /* column-major per image            */
    const double *img_shift;           /* 3 per image: periodic shift       */
    int n_atom;
    const int *ip_displ;               /* n_atom+1 projector offsets        */
    const double *gamma;               /* per global projector              */
    double dV;
    double kpt[3];                     /* k in the lattice-aligned basis    */
    double cell[3];                    /* cell lengths                      */
} orc_op;

/* ------------------------------------------------------------------------ */
/* Halo: HaloExchange.cpp:34-82 (real), 104-189 (complex with Bloch phase). */
/* Interior copy, then z faces, then y faces (full extended rows), then x. */

static void halo_real(const orc_op *op, const double *x, double *xe, int ncol)
{
    const int nx = op->nx, ny = op->ny, nz = op->nz, R = op->fdn;
    const int ex = nx + 2 * R, ey = ny + 2 * R, ez = nz + 2 * R;
    const long pl = (long)ex * ey, vol = pl * ez, nd = (long)nx * ny * nz;
    memset(xe, 0, sizeof(double) * vol * ncol);
    for (int n = 0; n < ncol; ++n) {
        const double *s = x + n * nd;
        double *d = xe + n * vol;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                memcpy(d + (k + R) * pl + (long)(j + R) * ex + R,
                       s + (long)k * nx * ny + (long)j * nx, sizeof(double) * nx);
        for (int k = 0; k < R; ++k) {
            memcpy(d + k * pl, d + (nz + k) * pl, sizeof(double) * pl);
            memcpy(d + (nz + R + k) * pl, d + (R + k) * pl, sizeof(double) * pl);
        }
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < R; ++j) {
                memcpy(d + k * pl + (long)j * ex, d + k * pl + (long)(ny + j) * ex,
                       sizeof(double) * ex);
                memcpy(d + k * pl + (long)(ny + R + j) * ex, d + k * pl + (long)(R + j) * ex,
                       sizeof(double) * ex);
            }
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < ey; ++j) {
                double *row = d + k * pl + (long)j * ex;
                for (int i = 0; i < R; ++i) {
                    row[i] = row[nx + i];
                    row[nx + R + i] = row[R + i];
                }
            }
    }
}

static void halo_cplx(const orc_op *op, const zd *x, zd *xe, int ncol)
{
    const int nx = op->nx, ny = op->ny, nz = op->nz, R = op->fdn;
    const int ex = nx + 2 * R, ey = ny + 2 * R, ez = nz + 2 * R;
    const long pl = (long)ex * ey, vol = pl * ez, nd = (long)nx * ny * nz;
    double th[3];
    for (int a = 0; a < 3; ++a) th[a] = op->kpt[a] * op->cell[a];
    const zd lx = cos(th[0]) - I * sin(th[0]), rx = cos(th[0]) + I * sin(th[0]);
    const zd ly = cos(th[1]) - I * sin(th[1]), ry = cos(th[1]) + I * sin(th[1]);
    const zd lz = cos(th[2]) - I * sin(th[2]), rz = cos(th[2]) + I * sin(th[2]);
    memset(xe, 0, sizeof(zd) * vol * ncol);
    for (int n = 0; n < ncol; ++n) {
        const zd *s = x + n * nd;
        zd *d = xe + n * vol;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                memcpy(d + (k + R) * pl + (long)(j + R) * ex + R,
                       s + (long)k * nx * ny + (long)j * nx, sizeof(zd) * nx);
        for (int k = 0; k < R; ++k)
            for (long q = 0; q < pl; ++q) {
                d[k * pl + q] = d[(nz + k) * pl + q] * lz;
                d[(nz + R + k) * pl + q] = d[(R + k) * pl + q] * rz;
            }
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < R; ++j)
                for (int i = 0; i < ex; ++i) {
                    d[k * pl + (long)j * ex + i] = d[k * pl + (long)(ny + j) * ex + i] * ly;
                    d[k * pl + (long)(ny + R + j) * ex + i] = d[k * pl + (long)(R + j) * ex + i] * ry;
                }
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < ey; ++j) {
                zd *row = d + k * pl + (long)j * ex;
                for (int i = 0; i < R; ++i) {
                    row[i] = row[nx + i] * lx;
                    row[nx + R + i] = row[R + i] * rx;
                }
            }
    }
}

/* ------------------------------------------------------------------------ */
/* Stencil + diagonal: Hamiltonian.cpp:126-174 (orth) and 176-267.         */
/* y = -0.5*lap + (Veff + c)*x, evaluated on the halo-extended array.       */

#define STENCIL_BODY(T)                                                        \
    const int nx = op->nx, ny = op->ny, nz = op->nz, R = op->fdn;              \
    const int ex = nx + 2 * R, ey = ny + 2 * R;                                \
    const long pl = (long)ex * ey, vol = pl * (nz + 2 * R);                    \
    const long nd = (long)nx * ny * nz;                                        \
    const double *cx = op->cx, *cy = op->cy, *cz = op->cz;                     \
    const double diag = cx[0] + cy[0] + cz[0];                                 \
    const int has_xy = !op->orth && fabs(op->cxy[1]) > 1e-30;                  \
    const int has_xz = !op->orth && fabs(op->cxz[1]) > 1e-30;                  \
    const int has_yz = !op->orth && fabs(op->cyz[1]) > 1e-30;                  \
    for (int n = 0; n < ncol; ++n) {                                           \
        const T *xn = xe + n * vol;                                            \
        T *yn = y + n * nd;                                                    \
        for (int k = 0; k < nz; ++k)                                           \
            for (int j = 0; j < ny; ++j)                                       \
                for (int i = 0; i < nx; ++i) {                                 \
                    const long c0 = (k + R) * pl + (long)(j + R) * ex + R + i; \
                    const long loc = (long)k * nx * ny + (long)j * nx + i;     \
                    T lap = diag * xn[c0];                                     \
                    for (int p = 1; p <= R; ++p) {                             \
                        lap += cx[p] * (xn[c0 + p] + xn[c0 - p]);              \
                        lap += cy[p] * (xn[c0 + p * ex] + xn[c0 - p * ex]);    \
                        lap += cz[p] * (xn[c0 + p * pl] + xn[c0 - p * pl]);    \
                    }                                                          \
                    if (has_xy)                                                \
                        for (int p = 1; p <= R; ++p)                           \
                            for (int q = 1; q <= R; ++q)                       \
                                lap += op->cxy[p] * op->d1y[q] *               \
                                       (xn[c0 + p + q * ex] - xn[c0 + p - q * ex] - \
                                        xn[c0 - p + q * ex] + xn[c0 - p - q * ex]); \
                    if (has_xz)                                                \
                        for (int p = 1; p <= R; ++p)                           \
                            for (int q = 1; q <= R; ++q)                       \
                                lap += op->cxz[p] * op->d1z[q] *               \
                                       (xn[c0 + p + q * pl] - xn[c0 + p - q * pl] - \
                                        xn[c0 - p + q * pl] + xn[c0 - p - q * pl]); \
                    if (has_yz)                                                \
                        for (int p = 1; p <= R; ++p)                           \
                            for (int q = 1; q <= R; ++q)                       \
                                lap += op->cyz[p] * op->d1z[q] *               \
                                       (xn[c0 + p * ex + q * pl] - xn[c0 + p * ex - q * pl] - \
                                        xn[c0 - p * ex + q * pl] + xn[c0 - p * ex - q * pl]); \
                    const double v = veff ? veff[loc] : 0.0;                   \
                    yn[loc] = -0.5 * lap + (v + c) * xn[c0];                   \
                }                                                              \
    }

static void stencil_real(const orc_op *op, const double *xe, const double *veff,
                         double *y, int ncol, double c)
{
    STENCIL_BODY(double)
}

static void stencil_cplx(const orc_op *op, const zd *xe, const double *veff,
                         zd *y, int ncol, double c)
{
    STENCIL_BODY(zd)
}

/* ------------------------------------------------------------------------ */
/* Nonlocal projector: NonlocalProjector.cpp:121-227 / 229-347.             */
/* alpha layout [atom][column][projector]; images accumulate into their    */
/* physical atom; Gamma applied per projector; scatter back to all images. */

static void nonlocal_real(const orc_op *op, const double *psi, double *hpsi, int ncol)
{
    if (op->n_img == 0) return;
    const long nd = (long)op->nx * op->ny * op->nz;
    const int tot = op->ip_displ[op->n_atom];
    double *alpha = calloc((size_t)tot * ncol, sizeof(double));
    for (int a = 0; a < op->n_img; ++a) {
        const int ndc = op->img_ndc[a], np = op->img_nproj[a];
        if (ndc == 0 || np == 0) continue;
        const int *g = op->gpos + op->img_gpos_off[a];
        const double *chi = op->chi + op->img_chi_off[a];
        const long off = (long)op->ip_displ[op->img_atom[a]] * ncol;
        for (int n = 0; n < ncol; ++n) {
            const double *pn = psi + n * nd;
            for (int j = 0; j < np; ++j) {
                double dot = 0.0;
                for (int ig = 0; ig < ndc; ++ig) dot += chi[ig + (long)j * ndc] * pn[g[ig]];
                alpha[off + (long)n * np + j] += dot * op->dV;
            }
        }
    }
    for (int at = 0; at < op->n_atom; ++at) {
        const int np = op->ip_displ[at + 1] - op->ip_displ[at];
        const long off = (long)op->ip_displ[at] * ncol;
        for (int n = 0; n < ncol; ++n)
            for (int j = 0; j < np; ++j)
                alpha[off + (long)n * np + j] *= op->gamma[op->ip_displ[at] + j];
    }
    for (int a = 0; a < op->n_img; ++a) {
        const int ndc = op->img_ndc[a], np = op->img_nproj[a];
        if (ndc == 0 || np == 0) continue;
        const int *g = op->gpos + op->img_gpos_off[a];
        const double *chi = op->chi + op->img_chi_off[a];
        const long off = (long)op->ip_displ[op->img_atom[a]] * ncol;
        for (int n = 0; n < ncol; ++n) {
            double *hn = hpsi + n * nd;
            for (int ig = 0; ig < ndc; ++ig) {
                double v = 0.0;
                for (int j = 0; j < np; ++j) v += chi[ig + (long)j * ndc] * alpha[off + (long)n * np + j];
                hn[g[ig]] += v;
            }
        }
    }
    free(alpha);
}

static void nonlocal_cplx(const orc_op *op, const zd *psi, zd *hpsi, int ncol)
{
    if (op->n_img == 0) return;
    const long nd = (long)op->nx * op->ny * op->nz;
    const int tot = op->ip_displ[op->n_atom];
    zd *alpha = calloc((size_t)tot * ncol, sizeof(zd));
    for (int a = 0; a < op->n_img; ++a) {
        const int ndc = op->img_ndc[a], np = op->img_nproj[a];
        if (ndc == 0 || np == 0) continue;
        const int *g = op->gpos + op->img_gpos_off[a];
        const double *chi = op->chi + op->img_chi_off[a];
        const double *sh = op->img_shift + 3 * a;
        const double th = -(op->kpt[0] * sh[0] + op->kpt[1] * sh[1] + op->kpt[2] * sh[2]);
        const zd scale = (cos(th) + I * sin(th)) * op->dV;
        const long off = (long)op->ip_displ[op->img_atom[a]] * ncol;
        for (int n = 0; n < ncol; ++n) {
            const zd *pn = psi + n * nd;
            for (int j = 0; j < np; ++j) {
                zd dot = 0.0;
                for (int ig = 0; ig < ndc; ++ig) dot += chi[ig + (long)j * ndc] * pn[g[ig]];
                alpha[off + (long)n * np + j] += scale * dot;
            }
        }
    }
    for (int at = 0; at < op->n_atom; ++at) {
        const int np = op->ip_displ[at + 1] - op->ip_displ[at];
        const long off = (long)op->ip_displ[at] * ncol;
        for (int n = 0; n < ncol; ++n)
            for (int j = 0; j < np; ++j)
                alpha[off + (long)n * np + j] *= op->gamma[op->ip_displ[at] + j];
    }
    for (int a = 0; a < op->n_img; ++a) {
        const int ndc = op->img_ndc[a], np = op->img_nproj[a];
        if (ndc == 0 || np == 0) continue;
        const int *g = op->gpos + op->img_gpos_off[a];
        const double *chi = op->chi + op->img_chi_off[a];
        const double *sh = op->img_shift + 3 * a;
        const double th = -(op->kpt[0] * sh[0] + op->kpt[1] * sh[1] + op->kpt[2] * sh[2]);
        const zd conjfac = cos(th) - I * sin(th);
        const long off = (long)op->ip_displ[op->img_atom[a]] * ncol;
        for (int n = 0; n < ncol; ++n) {
            zd *hn = hpsi + n * nd;
            for (int ig = 0; ig < ndc; ++ig) {
                zd v = 0.0;
                for (int j = 0; j < np; ++j) v += chi[ig + (long)j * ndc] * alpha[off + (long)n * np + j];
                hn[g[ig]] += conjfac * v;
            }
        }
    }
    free(alpha);
}

/* ------------------------------------------------------------------------ */
/* Public entry points.                                                     */

static long ext_volume(const orc_op *op)
{
    const int R = op->fdn;
    return (long)(op->nx + 2 * R) * (op->ny + 2 * R) * (op->nz + 2 * R);
}

/* Halo on its own (for direct comparison with the reference's test data). */
void orc_halo_real(const orc_op *op, const double *x, double *xe, int ncol) { halo_real(op, x, xe, ncol); }
void orc_halo_cplx(const orc_op *op, const zd *x, zd *xe, int ncol) { halo_cplx(op, x, xe, ncol); }

/* Hamiltonian.cpp:64-75: local part, then nonlocal accumulated into y. */
void orc_h_apply_real(const orc_op *op, const double *psi, const double *veff,
                      double *y, int ncol, double c, int with_nonlocal)
{
    double *xe = malloc(sizeof(double) * ext_volume(op) * ncol);
    halo_real(op, psi, xe, ncol);
    stencil_real(op, xe, veff, y, ncol, c);
    free(xe);
    if (with_nonlocal) nonlocal_real(op, psi, y, ncol);
}

/* Hamiltonian.cpp:94-106. */
void orc_h_apply_cplx(const orc_op *op, const zd *psi, const double *veff,
                      zd *y, int ncol, double c, int with_nonlocal)
{
    zd *xe = malloc(sizeof(zd) * ext_volume(op) * ncol);
    halo_cplx(op, psi, xe, ncol);
    stencil_cplx(op, xe, veff, y, ncol, c);
    free(xe);
    if (with_nonlocal) nonlocal_cplx(op, psi, y, ncol);
}

/* EigenSolver.cpp:558-619: scaled three-term Chebyshev recurrence. */
#define FILTER_BODY(T, APPLY)                                                  \
    const long tot = (long)op->nx * op->ny * op->nz * ncol;                     \
    const double e = (eig_max - lambda_c) / 2.0;                               \
    const double cc = (eig_max + lambda_c) / 2.0;                              \
    const double sigma1 = e / (eig_min - cc);                                  \
    double sigma = sigma1;                                                     \
    T *xold = malloc(sizeof(T) * tot), *xnew = malloc(sizeof(T) * tot);        \
    T *hx = malloc(sizeof(T) * tot);                                           \
    APPLY(op, x, veff, hx, ncol, 0.0, with_nonlocal);                          \
    const double s0 = sigma / e;                                               \
    for (long i = 0; i < tot; ++i) y[i] = s0 * (hx[i] - cc * x[i]);            \
    memcpy(xold, x, sizeof(T) * tot);                                          \
    T *py = y, *pold = xold, *pnew = xnew;                                     \
    for (int k = 2; k <= degree; ++k) {                                        \
        const double sn = 1.0 / (2.0 / sigma1 - sigma);                        \
        const double g = 2.0 * sn / e;                                         \
        APPLY(op, py, veff, hx, ncol, 0.0, with_nonlocal);                     \
        const double ss = sigma * sn;                                          \
        for (long i = 0; i < tot; ++i) pnew[i] = g * (hx[i] - cc * py[i]) - ss * pold[i]; \
        T *t = pold; pold = py; py = pnew; pnew = t;                           \
        sigma = sn;                                                            \
    }                                                                          \
    if (py != y) memcpy(y, py, sizeof(T) * tot);                               \
    free(xold); free(xnew); free(hx);

void orc_cheb_filter_real(const orc_op *op, const double *x, double *y, const double *veff,
                          int ncol, double lambda_c, double eig_min, double eig_max,
                          int degree, int with_nonlocal)
{
    FILTER_BODY(double, orc_h_apply_real)
}

void orc_cheb_filter_cplx(const orc_op *op, const zd *x, zd *y, const double *veff,
                          int ncol, double lambda_c, double eig_min, double eig_max,
                          int degree, int with_nonlocal)
{
    FILTER_BODY(zd, orc_h_apply_cplx)
}

/* The same filter with the columns shared among nthreads OpenMP threads,
 * one column per call of the routine above: columns are independent in
 * every step (halo, stencil, nonlocal, recurrence), so the result is the
 * same, bit for bit, as one call over all columns.  The reference runs its
 * CPU filter with OpenMP over columns and planes (Hamiltonian.cpp:145-147). */
void orc_cheb_filter_real_par(const orc_op *op, const double *x, double *y, const double *veff,
                              int ncol, double lambda_c, double eig_min, double eig_max,
                              int degree, int with_nonlocal, int nthreads)
{
    const long nd = (long)op->nx * op->ny * op->nz;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int n = 0; n < ncol; ++n)
        orc_cheb_filter_real(op, x + n * nd, y + n * nd, veff, 1, lambda_c, eig_min, eig_max, degree,
                             with_nonlocal);
}

/* ------------------------------------------------------------------------ */
/* Lanczos bounds: Lanczos.hpp:104-198.                                     */
/* The tridiagonal eigenvalue problem (the reference calls LAPACK dsterf)  */
/* is solved here for its two extremal eigenvalues by Sturm-sequence       */
/* bisection, which converges to the same values to rounding.              */

static int sturm_count(const double *d, const double *e, int n, double x)
{
    /* number of eigenvalues of T strictly below x */
    int cnt = 0;
    double q = d[0] - x;
    if (q < 0) ++cnt;
    for (int i = 1; i < n; ++i) {
        if (q == 0.0) q = 1e-300;
        q = d[i] - x - e[i - 1] * e[i - 1] / q;
        if (q < 0) ++cnt;
    }
    return cnt;
}

static double tridiag_eig(const double *d, const double *e, int n, int which)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        double r = 0.0;
        if (i > 0) r += fabs(e[i - 1]);
        if (i < n - 1) r += fabs(e[i]);
        if (d[i] - r < lo) lo = d[i] - r;
        if (d[i] + r > hi) hi = d[i] + r;
    }
    const double span = hi - lo + 1e-300;
    lo -= 1e-12 * span; hi += 1e-12 * span;
    for (int it = 0; it < 200 && hi - lo > 4e-16 * (fabs(lo) + fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        const int c = sturm_count(d, e, n, mid);
        if (which == 0) { if (c >= 1) hi = mid; else lo = mid; }
        else            { if (c >= n) hi = mid; else lo = mid; }
    }
    return 0.5 * (lo + hi);
}

void orc_tridiag_extremes(const double *d, const double *e, int n, double *mn, double *mx)
{
    *mn = tridiag_eig(d, e, n, 0);
    *mx = tridiag_eig(d, e, n, 1);
}

#define LANCZOS_BODY(T, APPLY, DOT, NRM, RAND0, RANDSYM)                        \
    const long N = (long)op->nx * op->ny * op->nz;                             \
    T *vjm1 = malloc(sizeof(T) * N), *vj = malloc(sizeof(T) * N);             \
    T *vjp1 = malloc(sizeof(T) * N);                                           \
    double *a = calloc(max_iter + 1, sizeof(double));                          \
    double *b = calloc(max_iter + 1, sizeof(double));                          \
    srand(seed);                                                               \
    for (long i = 0; i < N; ++i) vjm1[i] = RAND0;                              \
    double vs = 1.0 / NRM(vjm1, N);                                            \
    for (long i = 0; i < N; ++i) vjm1[i] *= vs;                                \
    APPLY(op, vjm1, veff, vj, 1, 0.0, with_nonlocal);                          \
    a[0] = DOT(vjm1, vj, N);                                                   \
    for (long i = 0; i < N; ++i) vj[i] -= a[0] * vjm1[i];                      \
    b[0] = NRM(vj, N);                                                         \
    if (b[0] == 0.0) {                                                         \
        for (long i = 0; i < N; ++i) vj[i] = RANDSYM;                          \
        T pr = 0.0;                                                            \
        for (long i = 0; i < N; ++i) pr += PROJ_TERM(vj[i], vjm1[i]);          \
        for (long i = 0; i < N; ++i) vj[i] -= pr * vjm1[i];                    \
        b[0] = NRM(vj, N);                                                     \
    }                                                                          \
    vs = (b[0] == 0.0) ? 1.0 : 1.0 / b[0];                                     \
    for (long i = 0; i < N; ++i) vj[i] *= vs;                                  \
    double emin = 0.0, emax = 0.0, pmin = 0.0, pmax = 0.0;                     \
    double errmin = tol + 1.0, errmax = tol + 1.0;                             \
    int j = 0;                                                                 \
    double *dd = malloc(sizeof(double) * (max_iter + 2));                      \
    while ((errmin > tol || errmax > tol) && j < max_iter) {                   \
        APPLY(op, vj, veff, vjp1, 1, 0.0, with_nonlocal);                      \
        a[j + 1] = DOT(vj, vjp1, N);                                           \
        for (long i = 0; i < N; ++i) {                                         \
            vjp1[i] -= a[j + 1] * vj[i] + b[j] * vjm1[i];                      \
            vjm1[i] = vj[i];                                                   \
        }                                                                      \
        b[j + 1] = NRM(vjp1, N);                                               \
        if (b[j + 1] == 0.0) break;                                            \
        vs = 1.0 / b[j + 1];                                                   \
        for (long i = 0; i < N; ++i) vj[i] = vjp1[i] * vs;                     \
        const int n = j + 2;                                                   \
        orc_tridiag_extremes(a, b, n, &emin, &emax);                           \
        errmin = fabs(emin - pmin); errmax = fabs(emax - pmax);                \
        pmin = emin; pmax = emax;                                              \
        ++j;                                                                   \
    }                                                                          \
    *eig_min = emin - 0.1;                                                     \
    *eig_max = emax * 1.01;                                                    \
    *iters = j;                                                                \
    free(vjm1); free(vj); free(vjp1); free(a); free(b); free(dd);

static double dot_real(const double *x, const double *y, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}
static double nrm_real(const double *x, long n) { return sqrt(dot_real(x, x, n)); }
static double dot_cplx(const zd *x, const zd *y, long n)
{
    zd s = 0.0;
    for (long i = 0; i < n; ++i) s += conj(x[i]) * y[i];
    return creal(s);
}
static double nrm_cplx(const zd *x, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    return sqrt(s);
}
static double urand(void) { return (double)rand() / RAND_MAX; }

void orc_lanczos_real(const orc_op *op, const double *veff, double tol, int max_iter,
                      unsigned seed, int with_nonlocal, double *eig_min, double *eig_max,
                      int *iters)
{
#define PROJ_TERM(u, v) ((u) * (v))
    LANCZOS_BODY(double, orc_h_apply_real, dot_real, nrm_real, urand(), -1.0 + 2.0 * urand())
#undef PROJ_TERM
}

/* Complex start entry of Lanczos.hpp random_init: Complex(rand()/RAND_MAX,
 * rand()/RAND_MAX).  The order of the two calls is unspecified in C++; g++
 * (the reference's compiler here) evaluates the imaginary argument first,
 * which the reference subset confirms, so the first draw is the imaginary
 * part. */
static zd urand_cplx(void)
{
    const double im = urand();
    const double re = urand();
    return re + I * im;
}

/* The start vector before normalisation, for checking against the
 * reference's own random_init; out holds n doubles or n (re, im) pairs. */
void orc_lanczos_start(int complex_, long n, unsigned seed, double *out)
{
    srand(seed);
    if (complex_) {
        for (long i = 0; i < n; ++i) {
            const zd z = urand_cplx();
            out[2 * i] = creal(z);
            out[2 * i + 1] = cimag(z);
        }
    } else {
        for (long i = 0; i < n; ++i) out[i] = urand();
    }
}

void orc_lanczos_cplx(const orc_op *op, const double *veff, double tol, int max_iter,
                      unsigned seed, int with_nonlocal, double *eig_min, double *eig_max,
                      int *iters)
{
#define PROJ_TERM(u, v) (conj(u) * (v))
    LANCZOS_BODY(zd, orc_h_apply_cplx, dot_cplx, nrm_cplx,
                 urand_cplx(), ((-1.0 + 2.0 * urand()) + I * (-1.0 + 2.0 * urand())))
#undef PROJ_TERM
}

/* glibc rand() stream, U[lo, hi] per entry, column by column: the start
 * vectors used by Wavefunction::randomize (Wavefunction.cpp:53-76). */
void orc_rand_uniform(unsigned seed, double lo, double hi, double *out, long n)
{
    srand(seed);
    for (long i = 0; i < n; ++i) out[i] = lo + (hi - lo) * urand();
}
