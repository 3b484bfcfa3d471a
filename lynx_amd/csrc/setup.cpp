// Host-side set-up of the operator: lattice metrics, FD stencil
// coefficients, psp8 pseudopotentials, Hermite splines, real spherical
// harmonics and the Kleinman-Bylander projectors of every atom image whose
// rc sphere touches the grid.  Runs once per geometry; no GPU needed.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lynx {

void fail(int status, const std::string &msg) { throw Error{status, msg}; }

// ------------------------------------------------------------------ lattice
static void inv3(const double a[9], double out[9])
{
    const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                       a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (std::fabs(det) < 1e-300) fail(1, "singular lattice matrix");
    const double id = 1.0 / det;
    out[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    out[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    out[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    out[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    out[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    out[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    out[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    out[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    out[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// Lattice.cpp:13-54: unit vectors, |det|, lapc_T = (U^-1)^T U^-1 and the
// per-axis extent of a Cartesian sphere in lattice-aligned coordinates.
void lattice_init(Lattice &lat, const double latvec[9], bool orth)
{
    for (int i = 0; i < 9; ++i) lat.latvec[i] = latvec[i];
    lat.orth = orth;
    for (int i = 0; i < 3; ++i) {
        const double *v = latvec + 3 * i;
        lat.L[i] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (lat.L[i] < 1e-14) fail(1, "lattice vector of zero length");
        for (int j = 0; j < 3; ++j) lat.uvec[3 * i + j] = v[j] / lat.L[i];
    }
    const double *a = latvec;
    lat.jacobian = std::fabs(a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                             a[2] * (a[3] * a[7] - a[4] * a[6]));
    if (lat.jacobian < 1e-14) fail(1, "coplanar lattice vectors");
    inv3(lat.uvec, lat.uvec_inv);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += lat.uvec_inv[3 * k + i] * lat.uvec_inv[3 * k + j];
            lat.lapcT[3 * i + j] = s;
        }
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) s += lat.uvec_inv[3 * j + i] * lat.uvec_inv[3 * j + i];
        lat.ext[i] = std::sqrt(s);
    }
}

void noncart_to_cart(const Lattice &lat, const double nc[3], double cart[3])
{
    for (int j = 0; j < 3; ++j)
        cart[j] = nc[0] * lat.uvec[0 + j] + nc[1] * lat.uvec[3 + j] + nc[2] * lat.uvec[6 + j];
}

// ------------------------------------------------------------------ stencil
// Central-difference weights of order 2n: with f(n,p) = n!^2 / ((n-p)!(n+p)!),
// w1[p] = (-1)^(p+1) f / p and w2[p] = (-1)^(p+1) 2 f / p^2, w2[0] = -sum 2/p^2
// (FDStencil.cpp:7-44).  Numerator n!/(n-p)! and denominator (n+p)!/n! are
// integer products, exact in double for n <= 12, so f is one correctly
// rounded quotient.
void fd_weights(int order, std::vector<double> &w1, std::vector<double> &w2)
{
    if (order < 2 || order % 2 || order > 24) fail(1, "FD order must be even, 2..24");
    const int n = order / 2;
    w1.assign(n + 1, 0.0);
    w2.assign(n + 1, 0.0);
    for (int p = 1; p <= n; ++p) {
        double num = 1.0, den = 1.0;
        for (int i = n - p + 1; i <= n; ++i) num *= i;
        for (int i = n + 1; i <= n + p; ++i) den *= i;
        const double f = num / den;
        const double sgn = (p & 1) ? 1.0 : -1.0;
        w1[p] = sgn * f / p;
        w2[p] = sgn * 2.0 * f / (double(p) * p);
        w2[0] -= 2.0 / (double(p) * p);
    }
}

// FDStencil.cpp:46-130: per-axis scaling by 1/h and 1/h^2 (times lapc_T for
// non-orthogonal cells, plus mixed terms 2 lapc_T(a,b) w1 / h_a), and the
// bound of the -1/2 Laplacian spectrum.
void stencil_init(Stencil &st, int order, int nx, int ny, int nz, const Lattice &lat)
{
    st.order = order;
    st.fdn = order / 2;
    st.nx = nx; st.ny = ny; st.nz = nz;
    fd_weights(order, st.w1, st.w2);
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) st.h[a] = lat.L[a] / n[a];
    st.dV = lat.jacobian / (double(nx) * ny * nz);
    const int m = st.fdn + 1;
    for (int a = 0; a < 3; ++a) {
        st.D1[a].resize(m);
        st.D2[a].resize(m);
        st.D2m[a].assign(m, 0.0);
        const double t = lat.orth ? 1.0 : lat.lapcT[4 * a];
        for (int p = 0; p < m; ++p) {
            st.D1[a][p] = st.w1[p] / st.h[a];
            st.D2[a][p] = t * st.w2[p] / (st.h[a] * st.h[a]);
        }
    }
    if (!lat.orth) {
        for (int p = 0; p < m; ++p) {
            st.D2m[0][p] = 2.0 * lat.lapcT[1] * st.w1[p] / st.h[0];
            st.D2m[1][p] = 2.0 * lat.lapcT[2] * st.w1[p] / st.h[0];
            st.D2m[2][p] = 2.0 * lat.lapcT[5] * st.w1[p] / st.h[1];
        }
    }
    if (lat.orth) {
        // symbol of the stencil at the largest grid wavevector (pi for even n)
        double s = st.D2[0][0] + st.D2[1][0] + st.D2[2][0];
        for (int p = 1; p < m; ++p)
            for (int a = 0; a < 3; ++a) {
                const double scal = double(n[a] - n[a] % 2) / n[a];
                s += 2.0 * st.D2[a][p] * std::cos(M_PI * p * scal);
            }
        st.max_eig = -0.5 * s;
    } else {
        double s = 0;
        for (int a = 0; a < 3; ++a) s += std::fabs(st.D2[a][0]);
        for (int p = 1; p < m; ++p)
            for (int a = 0; a < 3; ++a)
                s += 2.0 * (std::fabs(st.D2[a][p]) + std::fabs(st.D2m[a][p]));
        st.max_eig = 0.5 * s;
    }
}

// --------------------------------------------------------------------- psp8
int Psp8::nproj_per_atom() const
{
    int t = 0;
    for (int l = 0; l <= lmax; ++l)
        if (l != lloc) t += ppl[l] * (2 * l + 1);
    return t;
}

double Psp8::rc_max() const
{
    double m = 0;
    for (double v : rc) m = std::max(m, v);
    return m;
}

// Reads the leading numeric fields of a line; Fortran 'D' exponents allowed.
static std::vector<double> numbers(std::string s, int maxn = 1 << 20)
{
    for (char &c : s)
        if (c == 'D' || c == 'd') c = 'E';
    std::vector<double> v;
    std::istringstream is(s);
    std::string tok;
    while ((int)v.size() < maxn && is >> tok) {
        char *end = nullptr;
        const double x = std::strtod(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0') break;
        v.push_back(x);
    }
    return v;
}

// psp8 layout (Pseudopotential.cpp:28-333): 6 header lines (plus an nprojso
// line when extension_switch >= 2), then for each l a header "l ekb..."
// followed by mmax rows "i r rp_0(r) rp_1(r) ..." for nonlocal channels or
// "i r Vloc(r)" for l = lloc.  Projectors are stored as p(r) = (r p(r)) / r,
// with p(0) copied from p(r_1).
void psp8_load(const std::string &path, Psp8 &ps)
{
    std::ifstream f(path);
    if (!f) fail(6, "cannot open pseudopotential file: " + path);
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) lines.push_back(line);
    size_t pos = 0;
    auto next = [&](int maxn = 1 << 20) {
        if (pos >= lines.size()) fail(6, "truncated psp8 file: " + path);
        return numbers(lines[pos++], maxn);
    };
    ++pos;                                  // title
    auto h2 = next(2);
    if (h2.size() < 2) fail(6, "bad psp8 line 2: " + path);
    ps.zion = h2[1];
    auto h3 = next(5);
    if (h3.size() < 5) fail(6, "bad psp8 line 3: " + path);
    if ((int)h3[0] != 8) fail(6, "not a psp8 file (pspcod != 8): " + path);
    ps.pspxc = (int)h3[1];
    ps.lmax = (int)h3[2];
    ps.lloc = (int)h3[3];
    ps.mmax = (int)h3[4];
    if (ps.lmax < 0 || ps.lmax > 3 || ps.mmax < 3) fail(8, "unsupported psp8 (lmax > 3 or tiny mesh): " + path);
    auto h4 = next(3);                      // rchrg fchrg qchrg
    ps.fchrg = h4.size() >= 2 ? h4[1] : 0.0;
    auto np = next(ps.lmax + 1);
    if ((int)np.size() < ps.lmax + 1) fail(6, "bad nproj line: " + path);
    ps.ppl.resize(ps.lmax + 1);
    for (int l = 0; l <= ps.lmax; ++l) ps.ppl[l] = (int)np[l];
    auto ext = next(1);
    const int extension = ext.empty() ? 0 : (int)ext[0];
    std::vector<int> nprojso(ps.lmax + 1, 0);
    if (extension >= 2) {                   // spin-orbit projector counts for l = 1..lmax
        auto so = next(ps.lmax + 1);
        for (int l = 1; l <= ps.lmax && l - 1 < (int)so.size(); ++l) nprojso[l] = (int)so[l - 1];
    }
    const int mm = ps.mmax;
    ps.r.assign(mm, 0.0);
    ps.udv.assign(ps.lmax + 1, {});
    ps.udv_d.assign(ps.lmax + 1, {});
    ps.gamma.assign(ps.lmax + 1, {});
    bool have_r = false;
    auto read_rows = [&](int ncols, std::vector<std::vector<double>> &cols) {
        cols.assign(ncols, std::vector<double>(mm));
        std::vector<double> rr(mm);
        for (int i = 0; i < mm; ++i) {
            auto v = next(2 + ncols);
            if ((int)v.size() < 2 + ncols) fail(6, "short radial row: " + path);
            rr[i] = v[1];
            for (int c = 0; c < ncols; ++c) cols[c][i] = v[2 + c];
        }
        if (!have_r) ps.r = rr;
        have_r = true;
    };
    for (int l = 0; l <= ps.lmax; ++l) {
        auto hdr = next(1 + ps.ppl[l]);
        const int nprj = ps.ppl[l];
        if (l != ps.lloc && nprj > 0) {
            if ((int)hdr.size() < 1 + nprj) fail(6, "bad projector header: " + path);
            ps.gamma[l].assign(hdr.begin() + 1, hdr.begin() + 1 + nprj);
            std::vector<std::vector<double>> cols;
            read_rows(nprj, cols);
            for (int p = 0; p < nprj; ++p) {
                for (int i = 0; i < mm; ++i) cols[p][i] = ps.r[i] > 1e-10 ? cols[p][i] / ps.r[i] : 0.0;
                cols[p][0] = cols[p][1];
            }
            ps.udv[l] = cols;
        } else if (l == ps.lloc) {
            std::vector<std::vector<double>> cols;
            read_rows(1, cols);
            ps.vloc0 = cols[0][0];
            ps.rvloc.resize(mm);
            for (int i = 0; i < mm; ++i) ps.rvloc[i] = ps.r[i] * cols[0][i];
        }
    }
    if (ps.lloc > ps.lmax) {                // the local channel follows the nonlocal ones
        next();
        std::vector<std::vector<double>> cols;
        read_rows(1, cols);
        ps.vloc0 = cols[0][0];
        ps.rvloc.resize(mm);
        for (int i = 0; i < mm; ++i) ps.rvloc[i] = ps.r[i] * cols[0][i];
    }
    if (ps.rvloc.empty()) fail(6, "psp8 file without a local potential: " + path);
    if (ps.vloc0 == 0.0 && mm > 1) ps.vloc0 = ps.r[0] < 1e-14 ? ps.rvloc[1] / ps.r[1] : ps.rvloc[0] / ps.r[0];
    if (extension >= 2 && ps.lmax >= 1)     // spin-orbit projectors: skipped
        for (int l = 1; l <= ps.lmax; ++l)
            if (nprojso[l] > 0) {
                next();
                for (int i = 0; i < mm; ++i) next();
            }
    // model core charge, then the isolated-atom valence density, each
    // "i r value ..." rows stored as value / (4 pi)
    auto density_rows = [&](std::vector<double> &out, bool required) {
        out.assign(mm, 0.0);
        for (int i = 0; i < mm; ++i) {
            if (pos >= lines.size()) {
                if (required) fail(6, "truncated core density: " + path);
                if (i == 0) out.clear();
                return;
            }
            auto v = numbers(lines[pos], 3);
            if (v.size() < 3) {
                if (required) fail(6, "short core density row: " + path);
                if (i == 0) out.clear();
                return;
            }
            ++pos;
            out[i] = v[2] / (4.0 * M_PI);
        }
    };
    if (ps.fchrg > 1e-10) density_rows(ps.rho_c, true);
    density_rows(ps.rho_iso, false);
    ps.rc.assign(ps.lmax + 1, 0.0);
    for (int l = 0; l <= ps.lmax; ++l) {
        if (ps.ppl[l] > 0 && l != ps.lloc) {
            ps.rc[l] = ps.r.back();
            for (int i = mm - 1; i >= 0; --i)
                if (std::fabs(ps.udv[l][0][i]) > 1e-12) { ps.rc[l] = ps.r[i]; break; }
        }
    }
    for (int l = 0; l <= ps.lmax; ++l) {
        ps.udv_d[l].resize(ps.udv[l].size());
        for (size_t p = 0; p < ps.udv[l].size(); ++p) {
            ps.udv_d[l][p].resize(mm);
            spline_deriv(ps.r.data(), ps.udv[l][p].data(), mm, ps.udv_d[l][p].data());
        }
    }
    auto deriv = [&](const std::vector<double> &y, std::vector<double> &d) {
        d.assign(y.size(), 0.0);
        if (!y.empty()) spline_deriv(ps.r.data(), y.data(), mm, d.data());
    };
    deriv(ps.rvloc, ps.rvloc_d);
    deriv(ps.rho_c, ps.rho_c_d);
    deriv(ps.rho_iso, ps.rho_iso_d);
}

double spline_eval1(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &yd,
                    double xi)
{
    const int n = (int)x.size();
    if (xi <= x[0]) return y[0];
    if (xi >= x[n - 1]) return y[n - 1];
    int j = (int)((xi - x[0]) / (x[1] - x[0]));
    if (j >= n - 1) j = n - 2;
    if (xi < x[j] || xi > x[j + 1]) {
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (x[mid] > xi) hi = mid; else lo = mid;
        }
        j = lo;
    }
    const double w = 1.0 / (x[j + 1] - x[j]);
    const double s = (y[j + 1] - y[j]) * w;
    const double a2 = w * (3.0 * s - 2.0 * yd[j] - yd[j + 1]);
    const double a3 = w * w * (-2.0 * s + yd[j] + yd[j + 1]);
    const double t = xi - x[j];
    return ((a3 * t + a2) * t + yd[j]) * t + y[j];
}

// ------------------------------------------------------------------ splines
// Hermite cubic with first derivatives from the tridiagonal C1/C2 system of
// NumericalMethods.cpp:128-192 (non-standard end rows), solved by a forward
// elimination and back substitution.
void spline_deriv(const double *x, const double *y, int n, double *yd)
{
    if (n < 2) { if (n == 1) yd[0] = 0; return; }
    if (n == 2) { yd[0] = yd[1] = (y[1] - y[0]) / (x[1] - x[0]); return; }
    std::vector<double> lo(n, 0.0), dg(n, 0.0), up(n, 0.0), f(n, 0.0);
    {
        const double a = x[1] - x[0], b = x[2] - x[1];
        const double s0 = (y[1] - y[0]) / a, s1 = (y[2] - y[1]) / b;
        dg[0] = b * (a + b);
        up[0] = (a + b) * (a + b);
        yd[0] = s0 * (3 * a * b + 2 * b * b) + s1 * a * a;
    }
    for (int i = 1; i < n - 1; ++i) {
        const double a = x[i] - x[i - 1], b = x[i + 1] - x[i];
        const double s0 = (y[i] - y[i - 1]) / a, s1 = (y[i + 1] - y[i]) / b;
        lo[i] = b;
        dg[i] = 2 * (a + b);
        up[i] = a;
        yd[i] = 3 * (s0 * b + s1 * a);
    }
    {
        const double a = x[n - 2] - x[n - 3], b = x[n - 1] - x[n - 2];
        const double s0 = (y[n - 2] - y[n - 3]) / a, s1 = (y[n - 1] - y[n - 2]) / b;
        lo[n - 1] = (a + b) * (a + b);
        dg[n - 1] = a * (a + b);
        yd[n - 1] = s0 * b * b + s1 * (3 * a * b + 2 * a * a);
    }
    double piv = dg[0];
    yd[0] /= piv;
    for (int j = 1; j < n; ++j) {
        f[j] = up[j - 1] / piv;
        piv = dg[j] - lo[j] * f[j];
        if (std::fabs(piv) < 1e-30) piv = 1e-30;
        yd[j] = (yd[j] - lo[j] * yd[j - 1]) / piv;
    }
    for (int j = n - 2; j >= 0; --j) yd[j] -= f[j + 1] * yd[j + 1];
}

// Evaluation clamped to the table ends; uniform tables use direct interval
// lookup, others bisection (NumericalMethods.cpp:194-290).
void spline_eval(const double *x, const double *y, const double *yd, int n, const double *xi,
                 double *yi, int m)
{
    const double dx0 = x[1] - x[0];
    bool uniform = true;
    for (int i = 2; i < n && uniform; ++i)
        if (std::fabs((x[i] - x[i - 1]) - dx0) > 1e-10 * std::max(1.0, dx0)) uniform = false;
    for (int k = 0; k < m; ++k) {
        const double r = xi[k];
        if (r <= x[0]) { yi[k] = y[0]; continue; }
        if (r >= x[n - 1]) { yi[k] = y[n - 1]; continue; }
        int j;
        if (uniform) {
            j = (int)((r - x[0]) / dx0);
            if (j >= n - 1) j = n - 2;
        } else {
            int a = 0, b = n - 1;
            while (b - a > 1) {
                const int c = (a + b) / 2;
                if (x[c] > r) b = c; else a = c;
            }
            j = a;
        }
        const double w = 1.0 / (x[j + 1] - x[j]);
        const double s = (y[j + 1] - y[j]) * w;
        const double c2 = w * (3 * s - 2 * yd[j] - yd[j + 1]);
        const double c3 = w * w * (-2 * s + yd[j] + yd[j + 1]);
        const double t = r - x[j];
        yi[k] = ((c3 * t + c2) * t + yd[j]) * t + y[j];
    }
}

// Real spherical harmonics, l <= 3, from their Cartesian closed forms
// (m < 0: sin-type, m > 0: cos-type), the convention of math_utils.hpp:174-.
double ylm_real(int l, int m, double x, double y, double z, double r)
{
    if (l == 0) return 0.5 * std::sqrt(1.0 / M_PI);
    if (r < 1e-14) return 0.0;
    const double u = x / r, v = y / r, w = z / r;
    if (l == 1) {
        const double c = std::sqrt(3.0 / (4.0 * M_PI));
        return m < 0 ? c * v : (m == 0 ? c * w : c * u);
    }
    if (l == 2) {
        const double a = 0.5 * std::sqrt(15.0 / M_PI);
        switch (m) {
            case -2: return a * u * v;
            case -1: return a * v * w;
            case 0: return 0.25 * std::sqrt(5.0 / M_PI) * (3 * w * w - 1.0);
            case 1: return a * u * w;
            default: return 0.25 * std::sqrt(15.0 / M_PI) * (u * u - v * v);
        }
    }
    if (l == 3) {
        const double a3 = 0.25 * std::sqrt(35.0 / (2.0 * M_PI));
        const double a1 = 0.25 * std::sqrt(21.0 / (2.0 * M_PI));
        switch (m) {
            case -3: return a3 * v * (3 * u * u - v * v);
            case -2: return 0.5 * std::sqrt(105.0 / M_PI) * u * v * w;
            case -1: return a1 * v * (5 * w * w - 1.0);
            case 0: return 0.25 * std::sqrt(7.0 / M_PI) * w * (5 * w * w - 3.0);
            case 1: return a1 * u * (5 * w * w - 1.0);
            case 2: return 0.25 * std::sqrt(105.0 / M_PI) * w * (u * u - v * v);
            default: return a3 * u * (u * u - 3 * v * v);
        }
    }
    fail(8, "spherical harmonic with l > 3");
}

// ----------------------------------------------------------- nonlocal set-up
// Crystal.cpp:136-255 and NonlocalProjector.cpp:13-119: for every atom (type
// order), every periodic image within reach (iz, iy, ix loops), collect the
// grid points with |r - R_img| <= rc_max and evaluate
// Chi[g, (l,p,m)] = Y_lm(r - R) * UdV_lp(|r - R|), l != lloc.
void build_nonlocal(const Lattice &lat, const Stencil &st, const std::vector<Psp8> &psps,
                    const std::vector<int> &atom_type, const std::vector<double> &pos_nc,
                    NonlocalHost &out)
{
    out = NonlocalHost{};
    const int natom = (int)atom_type.size();
    out.n_atom = natom;
    out.ip_displ.assign(natom + 1, 0);
    for (int ia = 0; ia < natom; ++ia) {
        const Psp8 &ps = psps[atom_type[ia]];
        out.ip_displ[ia + 1] = out.ip_displ[ia] + ps.nproj_per_atom();
        for (int l = 0; l <= ps.lmax; ++l) {
            if (l == ps.lloc) continue;
            for (int p = 0; p < ps.ppl[l]; ++p)
                for (int m = -l; m <= l; ++m) out.gamma.push_back(ps.gamma[l][p]);
        }
    }
    const int n[3] = {st.nx, st.ny, st.nz};
    for (int it = 0; it < (int)psps.size(); ++it) {
        const Psp8 &ps = psps[it];
        const double rc = ps.rc_max();
        if (rc < 1e-14) continue;
        const int nproj = ps.nproj_per_atom();
        double ext[3] = {1, 1, 1};
        if (!lat.orth) for (int a = 0; a < 3; ++a) ext[a] = lat.ext[a];
        for (int ia = 0; ia < natom; ++ia) {
            if (atom_type[ia] != it) continue;
            const double *pos = &pos_nc[3 * ia];
            int nimg[3];
            for (int a = 0; a < 3; ++a) nimg[a] = (int)std::ceil(rc * ext[a] / lat.L[a]);
            for (int iz = -nimg[2]; iz <= nimg[2]; ++iz)
                for (int iy = -nimg[1]; iy <= nimg[1]; ++iy)
                    for (int ix = -nimg[0]; ix <= nimg[0]; ++ix) {
                        const int im[3] = {ix, iy, iz};
                        double c[3];
                        int s[3], e[3];
                        for (int a = 0; a < 3; ++a) {
                            c[a] = pos[a] + im[a] * lat.L[a];
                            if (lat.orth) {
                                s[a] = (int)std::ceil((c[a] - rc) / st.h[a]);
                                e[a] = (int)std::floor((c[a] + rc) / st.h[a]);
                            } else {
                                const double rr = rc * ext[a] / st.h[a];
                                s[a] = (int)std::ceil(c[a] / st.h[a] - rr);
                                e[a] = (int)std::floor(c[a] / st.h[a] + rr);
                            }
                            s[a] = std::max(s[a], 0);
                            e[a] = std::min(e[a], n[a] - 1);
                        }
                        if (s[0] > e[0] || s[1] > e[1] || s[2] > e[2]) continue;
                        std::vector<int> g;
                        std::vector<double> dx, dy, dz, dr;
                        for (int k = s[2]; k <= e[2]; ++k)
                            for (int j = s[1]; j <= e[1]; ++j)
                                for (int i = s[0]; i <= e[0]; ++i) {
                                    double d[3] = {i * st.h[0] - c[0], j * st.h[1] - c[1], k * st.h[2] - c[2]};
                                    if (!lat.orth) {
                                        double cc[3];
                                        noncart_to_cart(lat, d, cc);
                                        d[0] = cc[0]; d[1] = cc[1]; d[2] = cc[2];
                                    }
                                    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                                    if (r2 <= rc * rc) {
                                        g.push_back(i + j * st.nx + k * st.nx * st.ny);
                                        dx.push_back(d[0]); dy.push_back(d[1]); dz.push_back(d[2]);
                                        dr.push_back(std::sqrt(r2));
                                    }
                                }
                        if (g.empty()) continue;
                        const int ndc = (int)g.size();
                        out.img_atom.push_back(ia);
                        out.img_ndc.push_back(ndc);
                        out.img_nproj.push_back(nproj);
                        out.gpos.insert(out.gpos.end(), g.begin(), g.end());
                        for (int a = 0; a < 3; ++a) out.img_shift.push_back(im[a] * lat.L[a]);
                        const size_t base = out.chi.size();
                        out.chi.resize(base + (size_t)ndc * nproj);
                        std::vector<double> u(ndc);
                        int col = 0;
                        for (int l = 0; l <= ps.lmax; ++l) {
                            if (l == ps.lloc) continue;
                            for (int p = 0; p < ps.ppl[l]; ++p) {
                                spline_eval(ps.r.data(), ps.udv[l][p].data(), ps.udv_d[l][p].data(),
                                            ps.mmax, dr.data(), u.data(), ndc);
                                for (int m = -l; m <= l; ++m, ++col)
                                    for (int q = 0; q < ndc; ++q)
                                        out.chi[base + (size_t)col * ndc + q] =
                                            ylm_real(l, m, dx[q], dy[q], dz[q], dr[q]) * u[q];
                            }
                        }
                        out.n_img++;
                    }
        }
    }
}

}  // namespace lynx

namespace lynx {

// ---------------------------------------------------------------- k-points --
void kpoints_generate(int k1, int k2, int k3, const double shift[3], const Lattice &lat, KPointSet &out)
{
    if (k1 < 1 || k2 < 1 || k3 < 1) fail(1, "k-point grid must be at least 1 x 1 x 1");
    const int K[3] = {k1, k2, k3};
    const double tol = 1e-12;
    int start[3];
    double sums[3];
    for (int a = 0; a < 3; ++a) {
        start[a] = -(int)std::floor((K[a] - 1) / 2.0);
        sums[a] = 2.0 * M_PI / lat.L[a];
    }
    out.cart.clear();
    out.red.clear();
    out.weight.clear();
    for (int n1 = start[0]; n1 < start[0] + K[0]; ++n1)
        for (int n2 = start[1]; n2 < start[1] + K[1]; ++n2)
            for (int n3 = start[2]; n3 < start[2] + K[2]; ++n3) {
                const int nn[3] = {n1, n2, n3};
                double kr[3], kc[3];
                for (int a = 0; a < 3; ++a) {
                    // fold into (-1/2, 1/2] with the reference's tolerance shift
                    const double v = (double)nn[a] / K[a];
                    kr[a] = std::fmod(v + shift[a] / K[a] + 0.5 - tol, 1.0) - 0.5 + tol;
                    kc[a] = kr[a] * 2.0 * M_PI / lat.L[a];
                }
                long hit = -1;
                const long np = (long)out.weight.size();
                for (long q = 0; q < np && hit < 0; ++q) {
                    bool same = true;
                    for (int a = 0; a < 3; ++a) {
                        const double s = kc[a] + out.cart[3 * q + a];
                        same = same && (std::fabs(s) < tol || std::fabs(s - sums[a]) < tol);
                    }
                    if (same) hit = q;
                }
                if (hit >= 0) {
                    out.weight[hit] = 2.0;
                } else {
                    for (int a = 0; a < 3; ++a) {
                        out.cart.push_back(kc[a]);
                        out.red.push_back(kr[a]);
                    }
                    out.weight.push_back(1.0);
                }
            }
    const double tot = (double)K[0] * K[1] * K[2];
    for (double &w : out.weight) w /= tot;
}

}  // namespace lynx
