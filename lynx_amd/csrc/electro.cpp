// Local-pseudopotential set-up of the SCF: the pseudocharge b of every atom
// (b_J = -1/(4 pi) lap V_J on a box around the atom, with the same FD
// stencil as the Hamiltonian), the reference charge b_ref of the smooth
// -Z/r potential, the correction potential Vc = V_ref - V_J, the self and
// correction energies, the superposed isolated-atom valence density and the
// model core density.  Follows Electrostatics.cpp:16-520 and
// Crystal::compute_atom_influence (Crystal.cpp:45-134).  Host code, run once
// per geometry; the boxes are processed one after another, each spread over
// threads by planes, so every grid point is accumulated in the same order
// whatever the thread count.
#include "common.hpp"
#include "electro.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <thread>

namespace lynx {

namespace {

// Runs fn(i) for i in [0, n) on up to 16 threads, contiguous ranges.
void parallel_for(int n, const std::function<void(int)> &fn)
{
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::min({hw, 16, std::max(1, n)});
    if (nt <= 1 || n < 4) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> th;
    const int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([a, b, &fn] {
            for (int i = a; i < b; ++i) fn(i);
        });
    }
    for (auto &x : th) x.join();
}

// Smooth reference potential of charge Z (Electrostatics.cpp:16-33): -Z/r
// outside rc, a polynomial inside.
double v_ref(double r, double rc, double Z)
{
    if (r > rc) return -Z / r;
    const double rc2 = rc * rc, rc3 = rc2 * rc, rc6 = rc3 * rc3, rc7 = rc6 * rc, rc8 = rc7 * rc;
    const double r2 = r * r, r3 = r2 * r;
    const double a0 = 2.4 / rc, a3 = -2.8 / rc3, a5 = 5.6 / rc6, a6 = -6.0 / rc7, a7 = 1.8 / rc8;
    return -Z * (r2 * (r3 * (a5 + r * (a6 + r * a7)) + a3) + a0);
}

// Local pseudopotential V_J(r) of one species: Vloc(0) at the nucleus, the
// spline of r Vloc(r) divided by r on the table, -Z/r beyond it.
double v_local(const Psp8 &ps, double r)
{
    if (r < 1e-10 || std::isnan(r)) return ps.vloc0;
    if (r < ps.r.back()) return spline_eval1(ps.r, ps.rvloc, ps.rvloc_d, r) / r;
    return -ps.zion / r;
}

// Distance of a displacement given in the grid's lattice-aligned coordinates.
double distance(const Lattice &lat, double x, double y, double z)
{
    if (lat.orth) return std::sqrt(x * x + y * y + z * z);
    const double d[3] = {x, y, z};
    double c[3];
    noncart_to_cart(lat, d, c);
    return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}

struct Box {
    int type, atom;
    double c[3];      // image position, lattice-aligned coordinates
    int s[3], e[3];   // grid index range (inclusive), within the grid
};

// Crystal.cpp:45-134: every periodic image of every atom (type order, then
// atom, then iz, iy, ix) whose rc_max box meets the grid.
std::vector<Box> atom_boxes(const Lattice &lat, const Stencil &st, const std::vector<int> &types,
                            const std::vector<double> &pos, double rc_max)
{
    std::vector<Box> out;
    const int n[3] = {st.nx, st.ny, st.nz};
    int ntypes = 0;
    for (int t : types) ntypes = std::max(ntypes, t + 1);
    for (int it = 0; it < ntypes; ++it)
        for (int ia = 0; ia < (int)types.size(); ++ia) {
            if (types[ia] != it) continue;
            double ext[3] = {1, 1, 1};
            if (!lat.orth)
                for (int a = 0; a < 3; ++a) ext[a] = lat.ext[a];
            int nimg[3];
            for (int a = 0; a < 3; ++a) nimg[a] = (int)std::ceil(rc_max * ext[a] / lat.L[a]);
            for (int iz = -nimg[2]; iz <= nimg[2]; ++iz)
                for (int iy = -nimg[1]; iy <= nimg[1]; ++iy)
                    for (int ix = -nimg[0]; ix <= nimg[0]; ++ix) {
                        const int im[3] = {ix, iy, iz};
                        Box b;
                        b.type = it;
                        b.atom = ia;
                        bool empty = false;
                        for (int a = 0; a < 3; ++a) {
                            b.c[a] = pos[3 * ia + a] + im[a] * lat.L[a];
                            int s, e;
                            if (lat.orth) {
                                s = (int)std::ceil((b.c[a] - rc_max) / st.h[a]);
                                e = (int)std::floor((b.c[a] + rc_max) / st.h[a]);
                            } else {
                                const double rr = rc_max * ext[a] / st.h[a];
                                s = (int)std::ceil(b.c[a] / st.h[a] - rr);
                                e = (int)std::floor(b.c[a] / st.h[a] + rr);
                            }
                            b.s[a] = std::max(s, 0);
                            b.e[a] = std::min(e, n[a] - 1);
                            if (b.s[a] > b.e[a]) empty = true;
                        }
                        if (!empty) out.push_back(b);
                    }
        }
    return out;
}

// Crystal.cpp:327-336: rc_max over every species' projector radii and
// radial table, plus 8 grid spacings
double influence_radius(const std::vector<Psp8> &psps, const Stencil &st)
{
    double rc_max = 0.0;
    for (const Psp8 &ps : psps) {
        for (double r : ps.rc) rc_max = std::max(rc_max, r);
        if (!ps.r.empty()) rc_max = std::max(rc_max, ps.r.back());
    }
    return rc_max + 8.0 * std::max({st.h[0], st.h[1], st.h[2]});
}

// -1/(4 pi) times the FD Laplacian (mixed terms unfactored, as
// Electrostatics.cpp:48-127) of a box-extended field V at index idx
struct BoxLaplacian {
    const Lattice &lat;
    const Stencil &st;
    long sy, sz;
    double lap(const std::vector<double> &V, long idx) const
    {
        const int R = st.fdn;
        const double coef = -0.25 / M_PI;
        const double *cx = st.D2[0].data(), *cy = st.D2[1].data(), *cz = st.D2[2].data();
        const double *cxy = st.D2m[0].data(), *cxz = st.D2m[1].data(), *cyz = st.D2m[2].data();
        const double *d1y = st.D1[1].data(), *d1z = st.D1[2].data();
        double val = (cx[0] + cy[0] + cz[0]) * coef * V[idx];
        for (int p = 1; p <= R; ++p) {
            val += coef * cx[p] * (V[idx + p] + V[idx - p]);
            val += coef * cy[p] * (V[idx + p * sy] + V[idx - p * sy]);
            val += coef * cz[p] * (V[idx + p * sz] + V[idx - p * sz]);
        }
        if (!lat.orth) {
            if (std::fabs(cxy[1]) > 1e-30)
                for (int p = 1; p <= R; ++p)
                    for (int q = 1; q <= R; ++q)
                        val += coef * cxy[p] * d1y[q] *
                               (V[idx + p + q * sy] - V[idx + p - q * sy] - V[idx - p + q * sy] + V[idx - p - q * sy]);
            if (std::fabs(cxz[1]) > 1e-30)
                for (int p = 1; p <= R; ++p)
                    for (int q = 1; q <= R; ++q)
                        val += coef * cxz[p] * d1z[q] *
                               (V[idx + p + q * sz] - V[idx + p - q * sz] - V[idx - p + q * sz] + V[idx - p - q * sz]);
            if (std::fabs(cyz[1]) > 1e-30)
                for (int p = 1; p <= R; ++p)
                    for (int q = 1; q <= R; ++q)
                        val += coef * cyz[p] * d1z[q] *
                               (V[idx + p * sy + q * sz] - V[idx + p * sy - q * sz] - V[idx - p * sy + q * sz] +
                                V[idx - p * sy - q * sz]);
        }
        return val;
    }
    // FD first derivative along lattice axis a
    double d1(const std::vector<double> &V, long idx, int a) const
    {
        const long s = a == 0 ? 1 : (a == 1 ? sy : sz);
        double v = 0.0;
        for (int p = 1; p <= st.fdn; ++p) v += (V[idx + p * s] - V[idx - p * s]) * st.D1[a][p];
        return v;
    }
};

}  // namespace

void build_electrostatics(const Lattice &lat, const Stencil &st, const std::vector<Psp8> &psps,
                          const std::vector<int> &types, const std::vector<double> &pos, ElecHost &out)
{
    out = ElecHost{};
    const long nd = (long)st.nx * st.ny * st.nz;
    const int R = st.fdn;
    const double dV = st.dV;
    const double rc_ref = 0.5;   // REFERENCE_CUTOFF of the reference potential
    out.b.assign(nd, 0.0);
    out.b_ref.assign(nd, 0.0);
    out.vc.assign(nd, 0.0);
    out.rho_at.assign(nd, 0.0);
    out.rho_core.assign(nd, 0.0);
    const double rc_max = influence_radius(psps, st);
    out.nelectron = 0;
    for (int t : types) out.nelectron += (int)psps[t].zion;
    out.natom = (int)types.size();
    for (const Psp8 &ps : psps)
        if (ps.fchrg > 1e-10) out.has_nlcc = true;

    const std::vector<Box> boxes = atom_boxes(lat, st, types, pos, rc_max);
    out.n_boxes = (int)boxes.size();
    const int nxg = st.nx, nyg = st.ny;
    // Laplacian coefficients scaled by -1/(4 pi) (Electrostatics.cpp:48-127)
    const double coef = -0.25 / M_PI;
    const double *cx = st.D2[0].data(), *cy = st.D2[1].data(), *cz = st.D2[2].data();
    const double *cxy = st.D2m[0].data(), *cxz = st.D2m[1].data(), *cyz = st.D2m[2].data();
    const double *d1y = st.D1[1].data(), *d1z = st.D1[2].data();
    const bool has_xy = !lat.orth && std::fabs(cxy[1]) > 1e-30;
    const bool has_xz = !lat.orth && std::fabs(cxz[1]) > 1e-30;
    const bool has_yz = !lat.orth && std::fabs(cyz[1]) > 1e-30;
    const double wdiag = (cx[0] + cy[0] + cz[0]) * coef;

    for (const Box &B : boxes) {
        const Psp8 &ps = psps[B.type];
        const double Z = ps.zion;
        const int nl[3] = {B.e[0] - B.s[0] + 1, B.e[1] - B.s[1] + 1, B.e[2] - B.s[2] + 1};
        const int np[3] = {nl[0] + 2 * R, nl[1] + 2 * R, nl[2] + 2 * R};
        const long npl = (long)np[0] * np[1];
        std::vector<double> VJ((size_t)npl * np[2]), VR((size_t)npl * np[2]);
        // V_J and V_ref on the box extended by the stencil radius
        parallel_for(np[2], [&](int kk) {
            const double rz = (B.s[2] - R + kk) * st.h[2] - B.c[2];
            for (int jj = 0; jj < np[1]; ++jj) {
                const double ry = (B.s[1] - R + jj) * st.h[1] - B.c[1];
                for (int ii = 0; ii < np[0]; ++ii) {
                    const double rx = (B.s[0] - R + ii) * st.h[0] - B.c[0];
                    const double r = distance(lat, rx, ry, rz);
                    const long idx = ii + jj * (long)np[0] + kk * npl;
                    double vj = v_local(ps, r);
                    double vr = v_ref(r, rc_ref, Z);
                    if (std::isnan(vj)) vj = ps.vloc0;
                    if (std::isnan(vr)) vr = v_ref(0.0, rc_ref, Z);
                    VJ[idx] = vj;
                    VR[idx] = vr;
                }
            }
        });
        // b_J = -1/(4 pi) lap V_J on the box; accumulate, and the box's
        // share of Eself = -1/2 dV sum b_ref V_ref (per plane, then summed
        // over planes in order)
        std::vector<double> es_plane(nl[2], 0.0);
        auto lap = [&](const std::vector<double> &V, long idx) {
            double val = wdiag * V[idx];
            for (int p = 1; p <= R; ++p) {
                val += coef * cx[p] * (V[idx + p] + V[idx - p]);
                val += coef * cy[p] * (V[idx + p * np[0]] + V[idx - p * np[0]]);
                val += coef * cz[p] * (V[idx + p * npl] + V[idx - p * npl]);
            }
            if (has_xy)
                for (int p = 1; p <= R; ++p)
                    for (int q = 1; q <= R; ++q)
                        val += coef * cxy[p] * d1y[q] *
                               (V[idx + p + q * np[0]] - V[idx + p - q * np[0]] - V[idx - p + q * np[0]] +
                                V[idx - p - q * np[0]]);
            if (has_xz)
                for (int p = 1; p <= R; ++p)
                    for (int q = 1; q <= R; ++q)
                        val += coef * cxz[p] * d1z[q] *
                               (V[idx + p + q * npl] - V[idx + p - q * npl] - V[idx - p + q * npl] +
                                V[idx - p - q * npl]);
            if (has_yz)
                for (int p = 1; p <= R; ++p)
                    for (int q = 1; q <= R; ++q)
                        val += coef * cyz[p] * d1z[q] *
                               (V[idx + p * np[0] + q * npl] - V[idx + p * np[0] - q * npl] -
                                V[idx - p * np[0] + q * npl] + V[idx - p * np[0] - q * npl]);
            return val;
        };
        parallel_for(nl[2], [&](int kk) {
            double es = 0.0;
            for (int jj = 0; jj < nl[1]; ++jj)
                for (int ii = 0; ii < nl[0]; ++ii) {
                    const long idx = (ii + R) + (jj + R) * (long)np[0] + (kk + R) * npl;
                    const double bj = lap(VJ, idx), br = lap(VR, idx);
                    const long g = (B.s[0] + ii) + (B.s[1] + jj) * (long)nxg + (B.s[2] + kk) * (long)nxg * nyg;
                    out.b[g] += bj;
                    out.b_ref[g] += br;
                    es -= br * VR[idx];
                    // Vc = V_ref - V_J; the valence and core densities
                    const double rx = (B.s[0] + ii) * st.h[0] - B.c[0];
                    const double ry = (B.s[1] + jj) * st.h[1] - B.c[1];
                    const double rz = (B.s[2] + kk) * st.h[2] - B.c[2];
                    const double r = distance(lat, rx, ry, rz);
                    out.vc[g] += v_ref(r, rc_ref, Z) - v_local(ps, r);
                    if (!ps.rho_iso.empty() && r < ps.r.back())
                        out.rho_at[g] += spline_eval1(ps.r, ps.rho_iso, ps.rho_iso_d, r);
                    if (ps.fchrg > 1e-10 && !ps.rho_c.empty() && r < ps.r.back())
                        out.rho_core[g] += spline_eval1(ps.r, ps.rho_c, ps.rho_c_d, r);
                }
            es_plane[kk] = es;
        });
        for (double v : es_plane) out.eself += v;
    }
    out.eself *= 0.5 * dV;
    double sb = 0.0, ec = 0.0, sr = 0.0;
    for (long i = 0; i < nd; ++i) {
        sb += out.b[i];
        ec += (out.b[i] + out.b_ref[i]) * out.vc[i];
        sr += out.rho_at[i];
    }
    out.int_b = sb * dV;
    out.ec = 0.5 * dV * ec;
    // the valence density rescaled to hold Nelectron (Electrostatics.cpp:455-463)
    const double int_rho = sr * dV;
    if (int_rho > 1e-10) {
        const double scale = out.nelectron / int_rho;
        for (double &v : out.rho_at) v *= scale;
    }
}

void local_forces(const Lattice &lat, const Stencil &st, const std::vector<Psp8> &psps,
                  const std::vector<int> &types, const std::vector<double> &pos, const ElecHost &E,
                  const double *const dphi[3], const double *const dvc[3], const double *vxc,
                  std::vector<double> &f_loc, std::vector<double> &f_xc)
{
    const int natom = (int)types.size();
    f_loc.assign(3 * (size_t)natom, 0.0);
    f_xc.assign(3 * (size_t)natom, 0.0);
    const int R = st.fdn;
    const double dV = st.dV;
    const double rc_ref = 0.5;
    const std::vector<Box> boxes = atom_boxes(lat, st, types, pos, influence_radius(psps, st));
    const long nxg = st.nx, nyg = st.ny;
    for (const Box &B : boxes) {
        const Psp8 &ps = psps[B.type];
        const double Z = ps.zion;
        const bool core = vxc && ps.fchrg > 1e-10 && !ps.rho_c.empty();
        const int nl[3] = {B.e[0] - B.s[0] + 1, B.e[1] - B.s[1] + 1, B.e[2] - B.s[2] + 1};
        const int np[3] = {nl[0] + 2 * R, nl[1] + 2 * R, nl[2] + 2 * R};
        const long npl = (long)np[0] * np[1];
        const size_t nbox = (size_t)npl * np[2];
        std::vector<double> VJ(nbox), VR(nbox), VC(nbox), RC(core ? nbox : 0);
        parallel_for(np[2], [&](int kk) {
            const double rz = (B.s[2] - R + kk) * st.h[2] - B.c[2];
            for (int jj = 0; jj < np[1]; ++jj) {
                const double ry = (B.s[1] - R + jj) * st.h[1] - B.c[1];
                for (int ii = 0; ii < np[0]; ++ii) {
                    const double rx = (B.s[0] - R + ii) * st.h[0] - B.c[0];
                    const double r = distance(lat, rx, ry, rz);
                    const long idx = ii + jj * (long)np[0] + kk * npl;
                    double vj = v_local(ps, r), vr = v_ref(r, rc_ref, Z);
                    if (std::isnan(vj)) vj = ps.vloc0;
                    if (std::isnan(vr)) vr = v_ref(0.0, rc_ref, Z);
                    VJ[idx] = vj;
                    VR[idx] = vr;
                    VC[idx] = vr - vj;
                    if (core) RC[idx] = r <= ps.r.back() ? spline_eval1(ps.r, ps.rho_c, ps.rho_c_d, r) : 0.0;
                }
            }
        });
        const BoxLaplacian L{lat, st, (long)np[0], npl};
        // per-plane sums, added over planes in order
        std::vector<double> part((size_t)nl[2] * 6, 0.0);
        parallel_for(nl[2], [&](int kk) {
            double f[3] = {0, 0, 0}, corr[3] = {0, 0, 0}, fx[3] = {0, 0, 0};
            for (int jj = 0; jj < nl[1]; ++jj)
                for (int ii = 0; ii < nl[0]; ++ii) {
                    const long idx = (ii + R) + (jj + R) * (long)np[0] + (kk + R) * npl;
                    const long g = (B.s[0] + ii) + (B.s[1] + jj) * nxg + (B.s[2] + kk) * nxg * nyg;
                    const double bj = L.lap(VJ, idx), br = L.lap(VR, idx);
                    const double bb = E.b[g] + E.b_ref[g];
                    for (int a = 0; a < 3; ++a) {
                        f[a] -= bj * dphi[a][g];
                        corr[a] += L.d1(VC, idx, a) * bb - dvc[a][g] * (bj + br);
                        if (core) fx[a] += vxc[g] * L.d1(RC, idx, a);
                    }
                }
            for (int a = 0; a < 3; ++a) {
                part[(size_t)kk * 6 + a] = f[a] + 0.5 * corr[a];
                part[(size_t)kk * 6 + 3 + a] = fx[a];
            }
        });
        double fl[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
        for (int kk = 0; kk < nl[2]; ++kk)
            for (int a = 0; a < 3; ++a) {
                fl[a] += part[(size_t)kk * 6 + a];
                fc[a] += part[(size_t)kk * 6 + 3 + a];
            }
        for (int a = 0; a < 3; ++a) {
            f_loc[3 * (size_t)B.atom + a] += fl[a] * dV;
            f_xc[3 * (size_t)B.atom + a] += fc[a] * dV;
        }
    }
}

}  // namespace lynx
