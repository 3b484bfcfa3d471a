// The C++ drop-in classes on the GPU: Hamiltonian + EigenSolver (the
// reference's solve signature) and ScfEigenStage, on Si8 at 16^3 with the
// Veff and start block the calling test writes.  Prints eigenvalues so the
// test can compare them with the same calls made through the Python API.
#include "lynx_chefsi.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <stdexcept>
#include <vector>

static std::vector<double> read_doubles(const char *path, size_t n)
{
    std::vector<double> v(n);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    const char *psp = argv[1];
    const int n = 16, nb = 20;
    const double a = 10.26;
    const double lv[9] = {a, 0, 0, 0, a, 0, 0, 0, a};
    lynx_mi355x::Hamiltonian H(0);
    H.setup(n, n, n, lv, true, 12);
    std::vector<double> frac = {0, 0, 0, .25, .25, .25, .5, .5, 0, .75, .75, .25,
                                .5, 0, .5, .75, .25, .75, 0, .5, .5, .25, .75, .75};
    H.set_atoms({psp}, std::vector<int>(8, 0), frac, true);
    const int nd = H.Nd();
    std::vector<double> veff = read_doubles(argv[2], nd);
    std::vector<double> psi = read_doubles(argv[3], (size_t)nd * nb);
    std::vector<double> psi2 = psi;

    lynx_mi355x::EigenSolver es;
    es.setup(H);
    double lo = 0, hi = 0;
    es.lanczos_bounds(veff.data(), nd, lo, hi);
    std::vector<double> eig(nb);
    es.solve(psi.data(), eig.data(), veff.data(), nd, nb, 0.5 * (lo + hi), lo, hi, 10);
    std::printf("bounds %.15e %.15e\n", lo, hi);
    std::printf("solve");
    for (double e : eig) std::printf(" %.15e", e);
    std::printf("\n");

    lynx_scf_eig_params p{};
    p.nspin = 1; p.nkpt = 1; p.complex_kpt = 0; p.nband = nb; p.cheb_degree = 20;
    p.rho_trigger = 4; p.nchefsi = 1; p.nelectron = 32.0; p.beta = 315.77; p.smearing = 0;
    p.fp64_switch_residual = 0.0;
    lynx_mi355x::ScfEigenStage stage(H, p);
    std::vector<double> e2(nb), occ(nb);
    const double ef = stage.step(0, 1.0, veff.data(), {psi2.data()}, nd, e2.data(), occ.data());
    std::printf("stage %.15e", ef);
    for (double e : e2) std::printf(" %.15e", e);
    std::printf("\n");

    // a refused potential throws and leaves the context without one; the
    // array set before it is then uploaded again by the next call
    std::vector<double> bad = veff;
    bad[3] = std::numeric_limits<double>::quiet_NaN();
    int threw = 0;
    try {
        H.set_veff(bad.data());
    } catch (const std::runtime_error &) {
        threw = 1;
    }
    double lo2 = 0, hi2 = 0;
    es.lanczos_bounds(veff.data(), nd, lo2, hi2);
    std::printf("refused %d %.15e %.15e\n", threw, lo2, hi2);

    // a change made in place to one entry of the same array reaches the
    // device: H(v + d e_j) x - H(v) x = d x_j e_j
    std::vector<double> x(psi2.begin(), psi2.begin() + nd), y0(nd), y1(nd);
    int j = 0;
    for (int i = 1; i < nd; ++i)
        if (std::fabs(x[i]) > std::fabs(x[j])) j = i;
    const double d = 0.75;
    H.apply(x.data(), veff.data(), y0.data(), 1);
    veff[j] += d;
    H.apply(x.data(), veff.data(), y1.data(), 1);
    double err = 0, ymax = 0;
    for (int i = 0; i < nd; ++i) {
        const double want = (i == j) ? d * x[i] : 0.0;
        err = std::fmax(err, std::fabs(y1[i] - y0[i] - want));
        ymax = std::fmax(ymax, std::fabs(y0[i]));
    }
    std::printf("inplace %.6e %.6e %.6e\n", err, ymax, std::fabs(d * x[j]));
    return 0;
}
