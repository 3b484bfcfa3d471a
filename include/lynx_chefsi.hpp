// lynx_chefsi.hpp -- C++ classes with the reference's EigenSolver and
// Hamiltonian signatures, on top of the C ABI of liblynx_chefsi.so.
//
// A caller of LYNX's src/solvers/EigenSolver.hpp and
// src/operators/Hamiltonian.hpp switches to the MI355X library by building
// these objects instead.  The work calls keep the reference's argument
// lists: EigenSolver::solve / solve_kpt / lanczos_bounds[_kpt] and
// Hamiltonian::apply / apply_kpt, with the same meaning (column-major psi with
// leading dimension ld, Nd_d grid points, eigvals ascending, Veff of Nd_d
// doubles) and the same side effect on lambda_cutoff().  The set-up calls do
// not: the reference's EigenSolver::setup(const LynxContext&, const
// Hamiltonian&) (EigenSolver.hpp:32) and Hamiltonian::setup(stencil, domain,
// grid, halo, vnl) (Hamiltonian.hpp:49-53) take LYNX's own grid objects,
// which this library does not have; here Hamiltonian::setup takes the grid
// as point counts, lattice rows and FD order, set_atoms the psp8 files and
// positions, and EigenSolver::setup the Hamiltonian alone.
// Failures throw std::runtime_error, as the reference's solve does on a
// Cholesky or eigensolver failure (EigenSolver.cpp:633-635, 667-669).
#pragma once

#include "lynx_chefsi.h"

#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace lynx_mi355x {

using Complex = std::complex<double>;

struct Vec3 {
    double x = 0, y = 0, z = 0;
};

inline void check(lynx_status s, const lynx_ctx *c)
{
    if (s != LYNX_OK) {
        std::string m = lynx_status_string(s);
        if (c) m += std::string(": ") + lynx_last_error(c);
        throw std::runtime_error(m);
    }
}

// One operator H = -1/2 lap + Veff + V_nl on one grid.  setup() plays the
// role of Hamiltonian::setup(stencil, domain, grid, halo, vnl)
// (Hamiltonian.hpp:49-53) with the grid described by its lattice rows, the
// point counts and the FD order.
class Hamiltonian {
public:
    explicit Hamiltonian(int device_id = 0) { check(lynx_ctx_create(&ctx_, device_id), nullptr); }
    ~Hamiltonian() { lynx_ctx_destroy(ctx_); }
    Hamiltonian(const Hamiltonian &) = delete;
    Hamiltonian &operator=(const Hamiltonian &) = delete;

    void setup(int nx, int ny, int nz, const double latvec[9], bool orthogonal, int fd_order = 12)
    {
        check(lynx_set_grid(ctx_, nx, ny, nz, latvec, orthogonal ? LYNX_CELL_ORTHOGONAL : LYNX_CELL_NONORTHOGONAL,
                            fd_order), ctx_);
        nd_ = nx * ny * nz;
        last_veff_.clear();
    }
    // Nonlocal projectors from psp8 files (Crystal::setup + NonlocalProjector::setup).
    void set_atoms(const std::vector<std::string> &psp8, const std::vector<int> &type,
                   const std::vector<double> &coords, bool fractional)
    {
        std::vector<const char *> p;
        for (auto &s : psp8) p.push_back(s.c_str());
        check(lynx_set_atoms(ctx_, (int)p.size(), p.data(), (int)type.size(), type.data(), coords.data(),
                             fractional ? 1 : 0), ctx_);
    }

    // Hamiltonian::apply (Hamiltonian.hpp:64-65): psi and y packed (stride Nd).
    void apply(const double *psi, const double *Veff, double *y, int ncol, double c = 0.0) const
    {
        set_veff(Veff);
        check(lynx_h_apply(ctx_, LYNX_F64, psi, nd_, y, nd_, ncol, c, LYNX_MEM_HOST), ctx_);
    }
    // Hamiltonian::apply_kpt (Hamiltonian.hpp:78-80).
    void apply_kpt(const Complex *psi, const double *Veff, Complex *y, int ncol, const Vec3 &kpt_cart,
                   const Vec3 & /*cell_lengths*/, double c = 0.0) const
    {
        set_veff(Veff);
        const double k[3] = {kpt_cart.x, kpt_cart.y, kpt_cart.z};
        check(lynx_set_kpoint(ctx_, k), ctx_);
        check(lynx_h_apply(ctx_, LYNX_C128, psi, nd_, y, nd_, ncol, c, LYNX_MEM_HOST), ctx_);
    }

    lynx_ctx *ctx() const { return ctx_; }
    int Nd() const { return nd_; }

    // Uploads Veff unless its Nd_d values are bit-identical to the last ones
    // this object uploaded (the caller may change an array in place between
    // calls) and no other caller has set a potential on the context since
    // (ScfEigenStage sets each spin's Veff itself).
    void set_veff(const double *Veff) const
    {
        const size_t bytes = sizeof(double) * static_cast<size_t>(nd_);
        long gen = -1;
        check(lynx_get_veff_generation(ctx_, &gen), ctx_);
        if (!last_veff_.empty() && gen == last_gen_ && std::memcmp(last_veff_.data(), Veff, bytes) == 0)
            return;
        // a refused potential leaves the context without one: forget the
        // previous values first so that they are uploaded again next time
        last_veff_.clear();
        check(lynx_set_veff(ctx_, Veff, LYNX_MEM_HOST), ctx_);
        last_veff_.assign(Veff, Veff + nd_);
        check(lynx_get_veff_generation(ctx_, &last_gen_), ctx_);
    }
    // Forget the cached potential so that the next call uploads it.
    void invalidate_veff() const { last_veff_.clear(); }

private:
    lynx_ctx *ctx_ = nullptr;
    int nd_ = 0;
    mutable std::vector<double> last_veff_;
    mutable long last_gen_ = -1;
};

// EigenSolver with the reference's solve/solve_kpt/lanczos_bounds signatures
// (EigenSolver.hpp:32-59).
class EigenSolver {
public:
    void setup(const Hamiltonian &H) { H_ = &H; }

    void solve(double *psi, double *eigvals, const double *Veff, int Nd_d, int Nband, double lambda_cutoff,
               double eigval_min, double eigval_max, int cheb_degree = 20, int ld = 0)
    {
        need(Nd_d);
        H_->set_veff(Veff);
        check(lynx_chefsi_solve(H_->ctx(), LYNX_F64, psi, ld ? ld : Nd_d, Nband, eigvals, lambda_cutoff, eigval_min,
                                eigval_max, cheb_degree, &lambda_cutoff_, LYNX_MEM_HOST), H_->ctx());
    }

    void solve_kpt(Complex *psi, double *eigvals, const double *Veff, int Nd_d, int Nband, double lambda_cutoff,
                   double eigval_min, double eigval_max, const Vec3 &kpt_cart, const Vec3 & /*cell_lengths*/,
                   int cheb_degree = 20, int ld = 0)
    {
        need(Nd_d);
        H_->set_veff(Veff);
        const double k[3] = {kpt_cart.x, kpt_cart.y, kpt_cart.z};
        check(lynx_set_kpoint(H_->ctx(), k), H_->ctx());
        check(lynx_chefsi_solve(H_->ctx(), LYNX_C128, psi, ld ? ld : Nd_d, Nband, eigvals, lambda_cutoff,
                                eigval_min, eigval_max, cheb_degree, &lambda_cutoff_, LYNX_MEM_HOST), H_->ctx());
    }

    void lanczos_bounds(const double *Veff, int Nd_d, double &eigmin, double &eigmax, double tol = 1e-2,
                        int max_iter = 1000)
    {
        need(Nd_d);
        H_->set_veff(Veff);
        check(lynx_lanczos_bounds(H_->ctx(), LYNX_F64, &eigmin, &eigmax, tol, max_iter, 1, nullptr), H_->ctx());
    }

    void lanczos_bounds_kpt(const double *Veff, int Nd_d, const Vec3 &kpt_cart, const Vec3 & /*cell_lengths*/,
                            double &eigmin, double &eigmax, double tol = 1e-2, int max_iter = 1000)
    {
        need(Nd_d);
        H_->set_veff(Veff);
        const double k[3] = {kpt_cart.x, kpt_cart.y, kpt_cart.z};
        check(lynx_set_kpoint(H_->ctx(), k), H_->ctx());
        check(lynx_lanczos_bounds(H_->ctx(), LYNX_C128, &eigmin, &eigmax, tol, max_iter, 1, nullptr), H_->ctx());
    }

    double lambda_cutoff() const { return lambda_cutoff_; }
    void set_lambda_cutoff(double v) { lambda_cutoff_ = v; }

    // Precision of solve/solve_kpt (no reference counterpart): FP32 filter,
    // Gram products and rotations with FP64 reductions, Cholesky and
    // eigensolve (see lynx_set_filter_precision).
    void set_filter_fp32(bool single)
    {
        if (!H_) throw std::runtime_error("EigenSolver::setup has not been called");
        check(lynx_set_filter_precision(H_->ctx(), single ? LYNX_FILTER_FP32 : LYNX_FILTER_FP64), H_->ctx());
    }
    // LYNX_FILTER_FP64, LYNX_FILTER_FP32 or LYNX_FILTER_BF16
    void set_precision(int precision)
    {
        if (!H_) throw std::runtime_error("EigenSolver::setup has not been called");
        check(lynx_set_filter_precision(H_->ctx(), precision), H_->ctx());
    }

private:
    void need(int Nd_d) const
    {
        if (!H_) throw std::runtime_error("EigenSolver::setup has not been called");
        if (Nd_d != H_->Nd()) throw std::runtime_error("Nd_d differs from the Hamiltonian's grid");
    }
    const Hamiltonian *H_ = nullptr;
    double lambda_cutoff_ = 0.0;
};

// The eigen stage of one SCF iteration, in place of SCF::solve_eigenproblem
// (SCF.cpp:310-467) with estimate_spectral_bounds (SCF.cpp:883-935): the
// caller's loop passes the iteration index, its density residual and Veff.
class ScfEigenStage {
public:
    ScfEigenStage(const Hamiltonian &H, const lynx_scf_eig_params &p, const std::vector<Vec3> &kpts = {},
                  const std::vector<double> &kpt_weights = {})
        : ctx_(H.ctx())
    {
        std::vector<double> k;
        for (const Vec3 &v : kpts) k.insert(k.end(), {v.x, v.y, v.z});
        check(lynx_scf_eig_create(&d_, ctx_, &p, k.empty() ? nullptr : k.data(),
                                  kpt_weights.empty() ? nullptr : kpt_weights.data()), ctx_);
    }
    ~ScfEigenStage() { lynx_scf_eig_destroy(d_); }
    ScfEigenStage(const ScfEigenStage &) = delete;
    ScfEigenStage &operator=(const ScfEigenStage &) = delete;

    // psi[s * nkpt + k]: blocks of nband columns (leading dimension ld);
    // returns the Fermi level; eig and occ hold nspin x nkpt x nband values.
    double step(int scf_iter, double scf_residual, const double *Veff, const std::vector<void *> &psi, int ld,
                double *eig, double *occ, bool device_buffers = false)
    {
        double ef = 0.0;
        const lynx_status st = lynx_scf_eig_step(d_, scf_iter, scf_residual, Veff, psi.data(), ld,
                                                 device_buffers ? LYNX_MEM_DEVICE : LYNX_MEM_HOST, eig, occ, &ef,
                                                 &last_precision_);
        if (st != LYNX_OK)
            throw std::runtime_error(std::string(lynx_status_string(st)) + ": " + lynx_scf_eig_last_error(d_));
        return ef;
    }
    // LYNX_FILTER_FP32 while the residual was above the switch, else FP64
    int last_filter_precision() const { return last_precision_; }
    double lambda_cutoff() const
    {
        double lc = 0.0;
        lynx_scf_eig_state(d_, &lc, nullptr, nullptr);
        return lc;
    }

private:
    lynx_ctx *ctx_ = nullptr;
    lynx_scf_eig *d_ = nullptr;
    int last_precision_ = LYNX_FILTER_FP64;
};

// Chebyshev filter of many (Hamiltonian, k-point) units in one pass per
// step (batched form of the per-k-point loop, SCF.cpp:391-411).
class BatchFilter {
public:
    BatchFilter(const std::vector<const Hamiltonian *> &H, const std::vector<Vec3> &kpts = {})
    {
        std::vector<lynx_ctx *> c;
        std::vector<double> k;
        for (size_t u = 0; u < H.size(); ++u) {
            c.push_back(H[u]->ctx());
            const Vec3 v = u < kpts.size() ? kpts[u] : Vec3{};
            k.insert(k.end(), {v.x, v.y, v.z});
        }
        const lynx_status st = lynx_batch_create(&b_, (int)c.size(), c.data(), k.data());
        if (st != LYNX_OK) throw std::runtime_error(std::string("lynx_batch_create: ") + lynx_status_string(st));
    }
    ~BatchFilter() { lynx_batch_destroy(b_); }
    BatchFilter(const BatchFilter &) = delete;
    BatchFilter &operator=(const BatchFilter &) = delete;

    // Device buffers: x[u] -> y[u] (leading dimensions ld[u], ncol[u]
    // columns), dtype LYNX_F64 / LYNX_C128 / LYNX_F32 / LYNX_C64.
    void filter(int dtype, const std::vector<void *> &x, const std::vector<void *> &y, const std::vector<long> &ld,
                const std::vector<int> &ncol, int degree, const std::vector<double> &lambda_cutoff,
                const std::vector<double> &eig_min, const std::vector<double> &eig_max)
    {
        const lynx_status st = lynx_batch_filter(b_, dtype, x.data(), ld.data(), y.data(), ld.data(), ncol.data(),
                                                 degree, lambda_cutoff.data(), eig_min.data(), eig_max.data(),
                                                 LYNX_MEM_DEVICE);
        if (st != LYNX_OK)
            throw std::runtime_error(std::string(lynx_status_string(st)) + ": " + lynx_batch_last_error(b_));
    }

private:
    lynx_batch *b_ = nullptr;
};

}  // namespace lynx_mi355x
