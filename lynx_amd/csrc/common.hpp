// Internal declarations shared by the host set-up code, the HIP kernels and
// the C ABI.  Nothing here is part of the public interface.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lynx {

// ---------------------------------------------------------------- errors --
// Internal failures travel as this exception and are converted to a
// lynx_status at the C-ABI boundary.
struct Error {
    int status;
    std::string msg;
};
[[noreturn]] void fail(int status, const std::string &msg);

// ---------------------------------------------------- lattice and stencil --
struct Lattice {
    double latvec[9];      // rows are lattice vectors
    double L[3];           // lengths
    double uvec[9];        // unit lattice vectors (rows)
    double uvec_inv[9];
    double lapcT[9];       // (U^-1)^T U^-1
    double ext[3];         // ||column_i(U^-1)||: non-Cartesian sphere extent
    double jacobian;
    bool orth;
};
void lattice_init(Lattice &lat, const double latvec[9], bool orth);
// lattice-aligned (non-Cartesian) coordinates -> Cartesian
void noncart_to_cart(const Lattice &lat, const double nc[3], double cart[3]);

struct Stencil {
    int order = 0, fdn = 0;
    int nx = 0, ny = 0, nz = 0;
    double h[3] = {0, 0, 0};
    double dV = 0;
    std::vector<double> w1, w2;        // fdn+1 raw weights
    std::vector<double> D1[3], D2[3];  // scaled per axis
    std::vector<double> D2m[3];        // mixed xy, xz, yz
    double max_eig = 0;
};
void fd_weights(int order, std::vector<double> &w1, std::vector<double> &w2);
void stencil_init(Stencil &st, int order, int nx, int ny, int nz, const Lattice &lat);

// ------------------------------------------------------------ psp8 data --
struct Psp8 {
    double zion = 0;
    int pspxc = 0, lmax = 0, lloc = 0, mmax = 0;
    std::vector<int> ppl;                            // projectors per l
    std::vector<double> r;                           // radial grid
    std::vector<std::vector<std::vector<double>>> udv, udv_d;   // [l][p][i]
    std::vector<std::vector<double>> gamma;          // [l][p]
    std::vector<double> rc;                          // per l
    // local part and densities (Pseudopotential.cpp:133-305)
    double vloc0 = 0;                                // Vloc(r = 0) from the file
    std::vector<double> rvloc, rvloc_d;              // r Vloc(r) and its spline derivative
    double fchrg = 0;                                // model core charge switch
    std::vector<double> rho_c, rho_c_d;              // core density / (4 pi), when fchrg > 0
    std::vector<double> rho_iso, rho_iso_d;          // isolated-atom valence density / (4 pi)
    int nproj_per_atom() const;
    double rc_max() const;
};
// The reference's single-point Hermite evaluation (NumericalMethods.cpp:194-225):
// the interval from the uniform-grid guess, checked, else bisection.
double spline_eval1(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &yd,
                    double xi);
void psp8_load(const std::string &path, Psp8 &out);
void spline_deriv(const double *x, const double *y, int n, double *yd);
void spline_eval(const double *x, const double *y, const double *yd, int n,
                 const double *xi, double *yi, int m);
double ylm_real(int l, int m, double x, double y, double z, double r);

// Nonlocal projector data in the host layout of the C ABI (one entry per
// influence image, accumulation order).
struct NonlocalHost {
    int n_img = 0, n_atom = 0;
    std::vector<int> img_atom, img_ndc, img_nproj;
    std::vector<int> gpos;
    std::vector<double> chi;           // per image (ndc x nproj) column-major
    std::vector<double> img_shift;     // 3 per image
    std::vector<int> ip_displ;         // n_atom+1
    std::vector<double> gamma;         // per projector
};
// Monkhorst-Pack k-points reduced by time reversal (KPoints.cpp:26-90):
// k_cart = 2 pi k_red / L per lattice direction; a point whose negative is
// already listed (modulo a reciprocal vector) doubles that point's weight.
struct KPointSet {
    std::vector<double> cart, red, weight;   // 3 per point (cart, red), 1 per point
};
void kpoints_generate(int k1, int k2, int k3, const double shift[3], const Lattice &lat, KPointSet &out);

void build_nonlocal(const Lattice &lat, const Stencil &st, const std::vector<Psp8> &psps,
                    const std::vector<int> &atom_type, const std::vector<double> &pos_nc,
                    NonlocalHost &out);

}  // namespace lynx
