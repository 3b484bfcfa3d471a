// Thin C ABI over a subset of the reference sources, compiled from
// /root/reference by oracle/ref_build.py into oracle/_ref/libref.so.
// TEST INFRASTRUCTURE ONLY: used by tests/test_ref_subset.py to validate the
// oracle's restatement (and the library's host set-up) against the
// reference's own code for the leaf steps that build without MPI, LAPACK or
// the rest of the program: FD stencil, lattice/grid, halo wrap, psp8 reader
// and splines, k-points, the Lanczos bound estimate (its template, with the
// caller's matvec and the LAPACK dsterf_ of the MKL runtime), the random
// initial wavefunctions, the occupations and Fermi level (Occupation with
// null communicators, linked against MPICH without MPI_Init), and the
// local-pseudopotential set-up (Crystal's atom boxes and Electrostatics).
#include "core/Domain.hpp"
#include "core/FDGrid.hpp"
#include "core/KPoints.hpp"
#include "core/Lattice.hpp"
#include "core/NumericalMethods.hpp"
#include "atoms/Pseudopotential.hpp"
#include "operators/FDStencil.hpp"
#include "parallel/HaloExchange.hpp"
#include "solvers/Lanczos.hpp"
#include "electronic/Occupation.hpp"
#include "electronic/Wavefunction.hpp"
#include "atoms/AtomType.hpp"
#include "atoms/Crystal.hpp"
#include "physics/Electrostatics.hpp"

#include <complex>
#include <cstring>
#include <exception>
#include <vector>

using namespace lynx;

namespace {
Lattice make_lattice(const double *lv, int orth)
{
    Mat3 m;
    for (int i = 0; i < 9; ++i) m.data[i] = lv[i];
    return Lattice(m, orth ? CellType::Orthogonal : CellType::NonOrthogonal);
}
}  // namespace

extern "C" {

// FDStencil(order, grid, lattice) for a periodic grid.
int ref_stencil(int order, int nx, int ny, int nz, const double *lv, int orth, double *w1, double *w2,
                double *D1, double *D2, double *D2m, double *max_eig, double *h, double *dV)
{
    try {
        Lattice lat = make_lattice(lv, orth);
        FDGrid g(nx, ny, nz, lat, BCType::Periodic, BCType::Periodic, BCType::Periodic);
        FDStencil st(order, g, lat);
        const int m = order / 2 + 1;
        std::memcpy(w1, st.weights_D1(), m * sizeof(double));
        std::memcpy(w2, st.weights_D2(), m * sizeof(double));
        const double *d1[3] = {st.D1_coeff_x(), st.D1_coeff_y(), st.D1_coeff_z()};
        const double *d2[3] = {st.D2_coeff_x(), st.D2_coeff_y(), st.D2_coeff_z()};
        const double *dm[3] = {st.D2_coeff_xy(), st.D2_coeff_xz(), st.D2_coeff_yz()};
        for (int a = 0; a < 3; ++a) {
            std::memcpy(D1 + a * m, d1[a], m * sizeof(double));
            std::memcpy(D2 + a * m, d2[a], m * sizeof(double));
            std::memcpy(D2m + a * m, dm[a], m * sizeof(double));
        }
        *max_eig = st.max_eigval_half_lap();
        h[0] = g.dx(); h[1] = g.dy(); h[2] = g.dz();
        *dV = g.dV();
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

// HaloExchange::execute / execute_kpt on ncol columns.
int ref_halo(int nx, int ny, int nz, int fdn, int ncol, const double *x, double *xe, int cplx, const double *kpt,
             const double *L)
{
    try {
        Mat3 m;
        m(0, 0) = L[0]; m(1, 1) = L[1]; m(2, 2) = L[2];
        Lattice lat(m, CellType::Orthogonal);
        FDGrid g(nx, ny, nz, lat, BCType::Periodic, BCType::Periodic, BCType::Periodic);
        DomainVertices v{0, nx - 1, 0, ny - 1, 0, nz - 1};
        Domain d(g, v);
        HaloExchange h(d, fdn);
        if (cplx) {
            h.execute_kpt(reinterpret_cast<const std::complex<double> *>(x), reinterpret_cast<std::complex<double> *>(xe),
                          ncol, Vec3(kpt[0], kpt[1], kpt[2]), Vec3(L[0], L[1], L[2]));
        } else {
            h.execute(x, xe, ncol);
        }
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

// Pseudopotential::load_psp8: header values and the (l, p) radial tables.
// sizes[0..4] = lmax, lloc, mmax, nproj_per_atom, n_tables; tables holds
// for each nonlocal (l, p) in order: udv[mmax], udv_d[mmax]; gamma and rc as
// flat arrays (gamma per table, rc per l).
int ref_psp8(const char *path, int *sizes, double *zion, double *r, double *tables, double *gamma, double *rc,
             int cap_tables)
{
    try {
        Pseudopotential p;
        p.load_psp8(path);
        const int mmax = p.grid_size();
        sizes[0] = p.lmax(); sizes[1] = p.lloc(); sizes[2] = mmax; sizes[3] = p.nproj_per_atom();
        *zion = p.Zval();
        std::memcpy(r, p.radial_grid().data(), mmax * sizeof(double));
        int t = 0;
        for (int l = 0; l <= p.lmax(); ++l) {
            rc[l] = p.rc()[l];
            if (l == p.lloc()) continue;
            for (int q = 0; q < p.ppl()[l]; ++q) {
                if (t >= cap_tables) return 2;
                std::memcpy(tables + (2 * t) * mmax, p.UdV()[l][q].data(), mmax * sizeof(double));
                std::memcpy(tables + (2 * t + 1) * mmax, p.UdV_spline_d()[l][q].data(), mmax * sizeof(double));
                gamma[t] = p.Gamma()[l][q];
                ++t;
            }
        }
        sizes[4] = t;
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

// Cubic Hermite spline of NumericalMethods: derivatives and evaluation.
void ref_spline(const double *x, const double *y, int n, const double *xi, int m, double *yd, double *yi)
{
    std::vector<double> vx(x, x + n), vy(y, y + n);
    SplineData s = spline_setup(vx, vy);
    std::memcpy(yd, s.dydx.data(), n * sizeof(double));
    spline_eval_array(s, xi, yi, m);
}

// KPoints::generate; returns the number of reduced k-points.
int ref_kpoints(int kx, int ky, int kz, const double *shift, const double *lv, int orth, double *kcart,
                double *weights, int cap)
{
    Lattice lat = make_lattice(lv, orth);
    KPoints k;
    k.generate(kx, ky, kz, Vec3(shift[0], shift[1], shift[2]), lat);
    const int n = k.Nkpts();
    for (int i = 0; i < n && i < cap; ++i) {
        kcart[3 * i] = k.kpts_cart()[i].x;
        kcart[3 * i + 1] = k.kpts_cart()[i].y;
        kcart[3 * i + 2] = k.kpts_cart()[i].z;
        weights[i] = k.weights()[i];
    }
    return n;
}

// Lanczos start vector before normalisation (Lanczos.hpp random_init after
// srand(seed)), as this compiler orders the two rand() calls of the complex
// case; out holds n doubles (real) or n interleaved (re, im) pairs.
void ref_lanczos_start(int complex_, int n, unsigned seed, double *out)
{
    std::srand(seed);
    if (complex_) {
        std::vector<std::complex<double>> v(n);
        lanczos_detail::random_init(v.data(), n);
        std::memcpy(out, v.data(), sizeof(double) * 2 * n);
    } else {
        lanczos_detail::random_init(out, n);
    }
}

// lynx::lanczos_bounds<T> (Lanczos.hpp:104-198) on the caller's operator:
// matvec(x, y) with n doubles (real) or n interleaved (re, im) pairs.
// *nmatvec receives the number of operator applications.
typedef void (*ref_matvec_fn)(const double *x, double *y);
void ref_lanczos(int complex_, int n, ref_matvec_fn mv, double tol, int max_iter, int seed, double *emin,
                 double *emax, int *nmatvec)
{
    int calls = 0;
    if (complex_) {
        std::function<void(const std::complex<double> *, std::complex<double> *)> f =
            [&](const std::complex<double> *x, std::complex<double> *y) {
                ++calls;
                mv(reinterpret_cast<const double *>(x), reinterpret_cast<double *>(y));
            };
        lanczos_bounds<std::complex<double>>(f, n, *emin, *emax, tol, max_iter, seed);
    } else {
        std::function<void(const double *, double *)> f = [&](const double *x, double *y) {
            ++calls;
            mv(x, y);
        };
        lanczos_bounds<double>(f, n, *emin, *emax, tol, max_iter, seed);
    }
    *nmatvec = calls;
}

// Wavefunction::randomize / randomize_kpt (Wavefunction.cpp:53-76) of one
// (spin, k-point) block of nd x nband: out holds nd*nband doubles (real) or
// interleaved (re, im) pairs, column-major.
void ref_randomize(int complex_, int nd, int nband, unsigned seed, double *out)
{
    Wavefunction w;
    w.allocate(nd, nband, 1, 1, complex_ != 0);
    if (complex_) {
        w.randomize_kpt(0, 0, seed);
        for (int j = 0; j < nband; ++j)
            std::memcpy(out + 2L * nd * j, w.psi_kpt(0, 0).col(j), sizeof(double) * 2 * nd);
    } else {
        w.randomize(0, 0, seed);
        for (int j = 0; j < nband; ++j) std::memcpy(out + (long)nd * j, w.psi(0, 0).col(j), sizeof(double) * nd);
    }
}

// Occupation::compute (Occupation.cpp:86-176) with null communicators:
// eigs holds nspin x nkpt x nband values; smearing 0 Fermi-Dirac, 1
// Gaussian (this library's numbering); returns the Fermi level and fills
// occ in the same layout.
double ref_occupation(int nspin, int nkpt, int nband, const double *eigs, const double *kpt_weights,
                      double nelectron, double beta, int smearing, double *occ)
{
    Wavefunction w;
    w.allocate(1, nband, nspin, nkpt, false);
    for (int s = 0; s < nspin; ++s)
        for (int k = 0; k < nkpt; ++k)
            for (int n = 0; n < nband; ++n) w.eigenvalues(s, k)(n) = eigs[((long)s * nkpt + k) * nband + n];
    std::vector<double> kw(kpt_weights, kpt_weights + nkpt);
    MPIComm none;
    const SmearingType st = smearing == 1 ? SmearingType::GaussianSmearing : SmearingType::FermiDirac;
    const double ef = Occupation::compute(w, nelectron, beta, st, kw, none, none, 0);
    for (int s = 0; s < nspin; ++s)
        for (int k = 0; k < nkpt; ++k)
            for (int n = 0; n < nband; ++n) occ[((long)s * nkpt + k) * nband + n] = w.occupations(s, k)(n);
    return ef;
}

// Crystal::compute_atom_influence + Electrostatics (pseudocharge, Vc, Ec,
// atomic and core densities) for natom atoms at fractional coordinates of
// ntype psp8 files (type-major), with the rc_max of Crystal::setup
// (Crystal.cpp:327-336).  out_* hold nd doubles; scalars[5] = Eself, Ec,
// int_b, number of boxes, has-NLCC.
int ref_electrostatics(int nx, int ny, int nz, const double *lv, int orth, int order, int ntype,
                       const char *const *paths, int natom, const int *types, const double *frac, double *out_b,
                       double *out_bref, double *out_vc, double *out_rho, double *out_core, double *scalars)
{
    try {
        Lattice lat = make_lattice(lv, orth);
        FDGrid g(nx, ny, nz, lat, BCType::Periodic, BCType::Periodic, BCType::Periodic);
        DomainVertices v{0, nx - 1, 0, ny - 1, 0, nz - 1};
        Domain d(g, v);
        FDStencil st(order, g, lat);
        std::vector<AtomType> at;
        std::vector<Vec3> pos;
        std::vector<int> ti;
        int nelec = 0;
        for (int t = 0; t < ntype; ++t) {
            int n = 0;
            for (int i = 0; i < natom; ++i) n += types[i] == t;
            Pseudopotential tmp;
            tmp.load_psp8(paths[t]);
            AtomType a("X", 1.0, tmp.Zval(), n);
            a.psd().load_psp8(paths[t]);
            for (int i = 0; i < natom; ++i) {
                if (types[i] != t) continue;
                Vec3 f(frac[3 * i], frac[3 * i + 1], frac[3 * i + 2]);
                Vec3 p;
                if (orth) {
                    p = lat.frac_to_cart(f);
                } else {
                    Vec3 L = lat.lengths();
                    p = Vec3(f.x * L.x, f.y * L.y, f.z * L.z);
                }
                pos.push_back(p);
                ti.push_back(t);
            }
            nelec += (int)tmp.Zval() * n;
            at.push_back(std::move(a));
        }
        Crystal cr(std::move(at), pos, ti, lat);
        double rc_max = 0.0;
        for (int t = 0; t < cr.n_types(); ++t) {
            const auto &ps = cr.types()[t].psd();
            for (auto rc : ps.rc()) rc_max = std::max(rc_max, rc);
            if (!ps.radial_grid().empty()) rc_max = std::max(rc_max, ps.radial_grid().back());
        }
        rc_max += 8.0 * std::max({g.dx(), g.dy(), g.dz()});
        std::vector<AtomInfluence> inf;
        cr.compute_atom_influence(d, rc_max, inf);
        Electrostatics el;
        el.compute_pseudocharge(cr, inf, d, g, st);
        const int nd = nx * ny * nz;
        std::vector<double> vloc(nd, 0.0);
        el.compute_Vloc(cr, inf, d, g, vloc.data());
        el.compute_Ec(vloc.data(), nd, g.dV());
        el.compute_atomic_density(cr, inf, d, g, out_rho, nelec);
        const bool nlcc = el.compute_core_density(cr, inf, d, g, out_core);
        std::memcpy(out_b, el.pseudocharge().data(), nd * sizeof(double));
        std::memcpy(out_bref, el.pseudocharge_ref().data(), nd * sizeof(double));
        std::memcpy(out_vc, vloc.data(), nd * sizeof(double));
        int nbox = 0;
        for (auto &x : inf) nbox += x.n_atom;
        scalars[0] = el.Eself();
        scalars[1] = el.Ec();
        scalars[2] = el.int_b();
        scalars[3] = nbox;
        scalars[4] = nlcc ? 1.0 : 0.0;
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

}  // extern "C"
