// C ABI of liblynx_chefsi.so: context, operator set-up and the CheFSI
// orchestration (filter, Cholesky-QR, Rayleigh-Ritz, rotation, Lanczos) on
// one HIP stream.  Dense N x N factorisations stay in FP64 on the GPU
// through rocSOLVER; everything Nd-sized runs in the project's own kernels.
#include "lynx_chefsi.h"

#include "ctx.hpp"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

using namespace lynx;

namespace {

// Scratch of lynx_h_apply_noalloc: the nonlocal projector coefficients
// beta (columns x atoms x padded projectors, compute type) of one call.  A
// file-scope device array: every GPU the library runs on holds its own copy
// until the process ends.  Host code never uses the array's name as an
// address; it takes the per-device address from hipGetSymbolAddress
// (scratch_address) and passes that.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

// Address of this device's copy of the scratch array, looked up once per
// device with that device current (the caller has made it current).
void *scratch_address(int device)
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void *p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(LYNX_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    addr[device] = p;
    return p;
}

}  // namespace

namespace lynx {
namespace detail {


void need_grid(const lynx_ctx *c)
{
    if (!c->has_grid) fail(LYNX_E_STATE, "lynx_set_grid has not been called");
}

void need_device(const lynx_ctx *c)
{
    if (c->device < 0) fail(LYNX_E_STATE, "host-only context (device -1) cannot run GPU work");
}

void need_gpu_stencil(const lynx_ctx *c)
{
    need_device(c);
    need_grid(c);
    if (c->st.fdn != gpu::kMaxR) fail(LYNX_E_UNSUPPORTED, "GPU stencil kernels are built for FD order 12");
    if (c->st.nx < gpu::kMaxR || c->st.ny < gpu::kMaxR || c->st.nz < gpu::kMaxR)
        fail(LYNX_E_ARG, "grid dimensions must be at least the stencil radius");
}

void need_veff(const lynx_ctx *c)
{
    if (!c->has_veff) fail(LYNX_E_STATE, "lynx_set_veff has not been called");
}

void check_dtype(int dtype)
{
    if (dtype < 0 || dtype > 4) fail(LYNX_E_ARG, "unknown dtype");
}

gpu::StencilGeom geom_at(const lynx_ctx *c, const double *kpt);

gpu::StencilGeom geom(const lynx_ctx *c) { return geom_at(c, c->kpt); }

// Stencil description of the context's grid with the Bloch phases of kpt.
gpu::StencilGeom geom_at(const lynx_ctx *c, const double *kpt)
{
    gpu::StencilGeom g{};
    g.nx = c->st.nx; g.ny = c->st.ny; g.nz = c->st.nz; g.R = c->st.fdn;
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k <= gpu::kMaxR; ++k) g.c[a][k] = k <= c->st.fdn ? c->st.D2[a][k] : 0.0;
    // factored form for the non-orthogonal kernel (FDStencil.cpp:59-101)
    g.orth = c->lat.orth;
    for (int k = 0; k <= gpu::kMaxR; ++k) {
        g.w1[k] = k <= c->st.fdn ? c->st.w1[k] : 0.0;
        g.w2[k] = k <= c->st.fdn ? c->st.w2[k] : 0.0;
    }
    const double *T = c->lat.lapcT;
    for (int a = 0; a < 3; ++a) {
        g.h[a] = c->st.h[a];
        g.beta[a] = (c->lat.orth ? 1.0 : T[4 * a]) / (g.h[a] * g.h[a]);
    }
    g.alpha[0] = c->lat.orth ? 0.0 : 2.0 * T[1] / g.h[0];
    g.alpha[1] = c->lat.orth ? 0.0 : 2.0 * T[2] / g.h[0];
    g.alpha[2] = c->lat.orth ? 0.0 : 2.0 * T[5] / g.h[1];
    for (int a = 0; a < 3; ++a) {
        const double th = kpt[a] * c->lat.L[a];
        g.phase[a][0][0] = std::cos(th);
        g.phase[a][0][1] = -std::sin(th);
        g.phase[a][1][0] = std::cos(th);
        g.phase[a][1][1] = std::sin(th);
    }
    return g;
}

const void *veff_for(const lynx_ctx *c, int e)
{
    return gpu::elem_single(e) ? c->veff32.p : c->veff64.p;
}

const gpu::NlDev &nl_for(const lynx_ctx *c, int e) { return gpu::elem_single(e) ? c->nl32 : c->nl64; }

// ---------------------------------------------------------------- timing --
int ev_index(lynx_ctx *c, size_t i)
{
    while (c->ev_pool.size() <= i) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
    }
    return (int)i;
}

struct Marker {
    lynx_ctx *c;
    int cls;
    size_t idx = 0;
    Marker(lynx_ctx *ctx, int k) : c(ctx), cls(k)
    {
        if (!c->timing) return;
        idx = c->ev_marks.size() * 2;
        ev_index(c, idx + 1);
        HIP_OK(hipEventRecord(c->ev_pool[idx], c->stream));
    }
    ~Marker()
    {
        if (!c->timing) return;
        if (hipEventRecord(c->ev_pool[idx + 1], c->stream) == hipSuccess)
            c->ev_marks.push_back({cls, (int)idx});
    }
};

void timing_begin(lynx_ctx *c) { c->ev_marks.clear(); }

void timing_end(lynx_ctx *c)
{
    if (!c->timing) return;
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) { c->t_last[k] = 0; c->n_last[k] = 0; }
    for (auto &m : c->ev_marks) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, c->ev_pool[m.second], c->ev_pool[m.second + 1]));
        c->t_last[m.first] += ms;
        c->n_last[m.first] += 1;
    }
    if (!c->ev_marks.empty()) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, c->ev_pool[c->ev_marks.front().second],
                                   c->ev_pool[c->ev_marks.back().second + 1]));
        c->t_last[0] = ms;
        c->n_last[0] = (int)c->ev_marks.size();
    }
}

// ------------------------------------------------------- nonlocal upload --
void upload_nonlocal(lynx_ctx *c)
{
    const NonlocalHost &h = c->nlh;
    const double dV = c->st.dV;
    int maxp = 0;
    for (int i = 0; i < h.n_img; ++i) maxp = std::max(maxp, h.img_nproj[i]);
    for (int a = 0; a < h.n_atom; ++a) maxp = std::max(maxp, h.ip_displ[a + 1] - h.ip_displ[a]);
    if (maxp > 32) fail(LYNX_E_UNSUPPORTED, "more than 32 projectors per atom");
    const int npm = maxp <= 8 ? 8 : (maxp <= 16 ? 16 : 32);

    std::vector<long> img_goff(h.n_img + 1, 0), img_coff(h.n_img + 1, 0);
    for (int i = 0; i < h.n_img; ++i) {
        img_goff[i + 1] = img_goff[i] + h.img_ndc[i];
        img_coff[i + 1] = img_coff[i] + (long)h.img_ndc[i] * h.img_nproj[i];
    }
    std::vector<int> atom_pt_off(h.n_atom + 1, 0), pt_gpos, pt_img, pt_atom;
    std::vector<double> chi;
    for (int a = 0; a < h.n_atom; ++a) {
        for (int i = 0; i < h.n_img; ++i) {
            if (h.img_atom[i] != a) continue;
            const int ndc = h.img_ndc[i], np = h.img_nproj[i];
            for (int g = 0; g < ndc; ++g) {
                pt_gpos.push_back(h.gpos[img_goff[i] + g]);
                pt_img.push_back(i);
                pt_atom.push_back(a);
                for (int j = 0; j < npm; ++j)
                    chi.push_back(j < np ? h.chi[img_coff[i] + (long)j * ndc + g] : 0.0);
            }
        }
        atom_pt_off[a + 1] = (int)pt_gpos.size();
    }
    const int npts = (int)pt_gpos.size();
    std::vector<double> gam((size_t)h.n_atom * npm, 0.0);
    for (int a = 0; a < h.n_atom; ++a)
        for (int j = 0; j < h.ip_displ[a + 1] - h.ip_displ[a]; ++j)
            gam[(size_t)a * npm + j] = dV * h.gamma[h.ip_displ[a] + j];
    // coverage lists: entries of each covered grid point, in gather order
    std::vector<int> order(npts);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pt_gpos[x] < pt_gpos[y]; });
    std::vector<int> cov_off{0}, cov_gpos, ent_pt;
    for (int k = 0; k < npts; ++k) {
        const int pt = order[k];
        if (cov_gpos.empty() || cov_gpos.back() != pt_gpos[pt]) {
            if (!cov_gpos.empty()) cov_off.push_back((int)ent_pt.size());
            cov_gpos.push_back(pt_gpos[pt]);
        }
        ent_pt.push_back(pt);
    }
    if (!cov_gpos.empty()) cov_off.push_back((int)ent_pt.size());
    for (int g : pt_gpos)
        if (g < 0 || g >= c->nd) fail(LYNX_E_ARG, "nonlocal grid index out of range");

    hipStream_t s = c->stream;
    upload(c->d_atom_pt_off, atom_pt_off, s);
    upload(c->d_pt_gpos, pt_gpos, s);
    upload(c->d_pt_img, pt_img, s);
    upload(c->d_pt_atom, pt_atom, s);
    upload(c->d_cov_off, cov_off, s);
    upload(c->d_cov_gpos, cov_gpos, s);
    upload(c->d_ent_pt, ent_pt, s);
    upload(c->d_chi64, chi, s);
    upload(c->d_gam64, gam, s);
    std::vector<float> chi32(chi.begin(), chi.end()), gam32(gam.begin(), gam.end());
    upload(c->d_chi32, chi32, s);
    upload(c->d_gam32, gam32, s);
    HIP_OK(hipStreamSynchronize(s));

    for (gpu::NlDev *v : {&c->nl64, &c->nl32}) {
        v->n_atom = h.n_atom;
        v->npm = npm;
        v->n_pts = npts;
        v->n_cov = (int)cov_gpos.size();
        v->n_img = h.n_img;
        v->max_atom_pts = 0;
        for (int a = 0; a < h.n_atom; ++a)
            v->max_atom_pts = std::max(v->max_atom_pts, atom_pt_off[a + 1] - atom_pt_off[a]);
        v->atom_pt_off = c->d_atom_pt_off.as<int>();
        v->pt_gpos = c->d_pt_gpos.as<int>();
        v->pt_img = c->d_pt_img.as<int>();
        v->pt_atom = c->d_pt_atom.as<int>();
        v->cov_off = c->d_cov_off.as<int>();
        v->cov_gpos = c->d_cov_gpos.as<int>();
        v->ent_pt = c->d_ent_pt.as<int>();
    }
    c->nl64.chi = c->d_chi64.p; c->nl64.gamma = c->d_gam64.p;
    c->nl32.chi = c->d_chi32.p; c->nl32.gamma = c->d_gam32.p;
    c->has_nl = h.n_img > 0;
}

// Bloch factors e^{-i k . R_img} of the images for the current k-point
// (NonlocalProjector.cpp:268-272).
void upload_phases(lynx_ctx *c)
{
    const NonlocalHost &h = c->nlh;
    std::vector<double> ph(2 * std::max(1, h.n_img));
    for (int i = 0; i < h.n_img; ++i) {
        const double *sh = &h.img_shift[3 * i];
        const double th = -(c->kpt[0] * sh[0] + c->kpt[1] * sh[1] + c->kpt[2] * sh[2]);
        ph[2 * i] = std::cos(th);
        ph[2 * i + 1] = std::sin(th);
    }
    upload(c->d_ph64, ph, c->stream);
    std::vector<float> ph32(ph.begin(), ph.end());
    upload(c->d_ph32, ph32, c->stream);
    HIP_OK(hipStreamSynchronize(c->stream));
    c->nl64.phase = c->d_ph64.p;
    c->nl32.phase = c->d_ph32.p;
}

// ------------------------------------------------------------ autotuner --
// Candidate launch shapes of the fused stencil step: points per thread
// (register ring depth x points) and number of z chunks per column.  Each is
// timed on the caller's own data (plain H application into the destination
// block, which the real step overwrites afterwards); the fastest is cached
// for this grid shape, element type and column count.
std::vector<gpu::StepCfg> step_candidates(const lynx_ctx *c, int e)
{
    std::vector<gpu::StepCfg> v;
    for (int vec : {0, 1, 2, 3})
      for (int align : {0, 1})
        for (int nt : {256, 512})
            for (int div : {1, 2})
                for (int nz : {1, 2, 3, 4}) {
                    if (align && !vec) continue;
                    gpu::StepCfg k;
                    k.rp_div = div;
                    k.nzch = nz;
                    k.nt = nt;
                    k.vec = vec;
                    k.align = align;
                    if (nz > 1 && c->st.nz / nz < 2 * gpu::kMaxR) continue;
                    if (vec && !c->lat.orth) continue;   // the non-orthogonal kernel has one form
                    if (!gpu::stencil_shape_fits(e, c->st.nx, k)) continue;
                    v.push_back(k);
                }
    return v;
}

const gpu::StepCfg &step_config(lynx_ctx *c, int e, const void *x, long ldx, void *y, long ldy, int ncol)
{
    if (c->has_forced) return c->forced;
    std::vector<long> key = {c->st.nx, c->st.ny, c->st.nz, e, ncol};
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) return it->second;
    gpu::StepCfg best{};
    if (c->autotune) {
        const auto cands = step_candidates(c, e);
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        double tbest = 1e30;
        // the timed step reads the destination as its Y_{k-2} stream (in
        // place, as the recurrence does), so all three streams are measured.
        // The whole block is timed (the tail of the last wave of workgroups
        // depends on the column count).  Two timing passes: every candidate once
        // after an untimed launch, then the six fastest three more times;
        // each candidate's time is its minimum.
        auto time_one = [&](const gpu::StepCfg &k) {
            float ms = 0;
            HIP_OK(hipEventRecord(e0, c->stream));
            gpu::stencil_step(e, geom(c), x, ldx, y, ldy, y, ldy, veff_for(c, e), ncol, 1.0, 1e-300, 0.0, k,
                              c->stream);
            HIP_OK(hipEventRecord(e1, c->stream));
            HIP_OK(hipEventSynchronize(e1));
            HIP_OK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms;
        };
        std::vector<double> times;
        for (const auto &k : cands) {
            (void)time_one(k);
            times.push_back(time_one(k));
        }
        std::vector<size_t> order(cands.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::sort(order.begin(), order.end(), [&](size_t i, size_t j) { return times[i] < times[j]; });
        for (size_t r = 0; r < std::min<size_t>(6, order.size()); ++r)
            for (int rep = 0; rep < 3; ++rep) times[order[r]] = std::min(times[order[r]], time_one(cands[order[r]]));
        for (size_t i = 0; i < cands.size(); ++i)
            if (times[i] < tbest) { tbest = times[i]; best = cands[i]; }
        // among shapes within 1.5% of the fastest, take the one with the
        // fewest z chunks (each chunk re-reads 2R planes): timing noise
        // should not pick the shape with more HBM traffic
        for (size_t i = 0; i < cands.size(); ++i)
            if (times[i] <= 1.015 * tbest && cands[i].vec == best.vec && cands[i].align == best.align &&
                cands[i].nt == best.nt && cands[i].rp_div == best.rp_div && cands[i].nzch < best.nzch)
                best = cands[i];
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c->tune_log[key] = times;
    }
    if (!gpu::stencil_shape_fits(e, c->st.nx, best)) {
        // untuned: the first shape that fits the rows (long rows need 512 threads)
        const auto cands = step_candidates(c, e);
        if (!cands.empty()) best = cands.front();
    }
    return c->tuned[key] = best;
}

// The bfloat16 filter runs on the run-vector stencil kernels only: an
// orthogonal cell whose rows hold an even number of points that fits one of
// their shapes.
bool bf16_filter_fits(const lynx_ctx *c)
{
    if (!c->lat.orth) return false;
    for (const auto &k : step_candidates(c, gpu::B16))
        if (k.vec) return true;
    return false;
}

// ------------------------------------------------------------- operator --
// y = a * (H - shift) x + b * old, H = -1/2 lap + Veff + V_nl.
void apply_op(lynx_ctx *c, int e, const void *x, long ldx, const void *old, long ldo, void *y, long ldy,
              int ncol, double a, double b, double shift, const void *coef = nullptr)
{
    const gpu::StepCfg cfg = step_config(c, e, x, ldx, y, ldy, ncol);
    const bool nl = c->has_nl;
    const gpu::NlDev &nd = nl_for(c, e);
    if (nl) {
        c->beta.get((size_t)ncol * nd.n_atom * nd.npm * gpu::elem_size(gpu::elem_compute(e)));
        Marker m(c, 2);
        gpu::nl_gather(e, nd, x, ldx, c->beta.p, ncol, c->stream);
    }
    {
        Marker m(c, 1);
        gpu::stencil_step(e, geom(c), x, ldx, old, ldo, y, ldy, veff_for(c, e), ncol, a, b, shift, cfg,
                          c->stream, coef);
    }
    if (nl) {
        Marker m(c, 3);
        gpu::nl_scatter(e, nd, c->beta.p, y, ldy, ncol, a, c->stream, coef);
    }
    HIP_OK(hipGetLastError());
}

// filters of higher degree run without a graph
constexpr int kMaxGraphDegree = 256;

// EigenSolver.cpp:558-619 as a sequence of fused steps.  x is read only; the
// final block lands in y; w is a work block of the same shape.
//
// The first call for a given key (buffers, column count, element type,
// launch shape and set-up epoch) runs the launches directly (and tunes the
// launch shape); the second captures them into a HIP graph whose kernels
// read each step's (a, b, shift) from a device table; later calls upload the
// table and replay the graph.  With timing on, the graph also holds the
// event records around each kernel, which every replay records again.
void filter(lynx_ctx *c, int e, const void *x, long ldx, void *y, long ldy, void *w, long ldw, int ncol,
            int degree, double lambda_c, double eig_min, double eig_max)
{
    if (degree < 1) fail(LYNX_E_ARG, "degree must be >= 1");
    const double ehalf = (eig_max - lambda_c) / 2.0;
    const double cen = (eig_max + lambda_c) / 2.0;
    if (!(ehalf > 0)) fail(LYNX_E_ARG, "eig_max must exceed lambda_cutoff");
    // (a, b, shift) of every step
    std::vector<double> sc(3 * degree);
    const double sigma1 = ehalf / (eig_min - cen);
    double sigma = sigma1;
    sc[0] = sigma / ehalf; sc[1] = 0.0; sc[2] = -cen;
    for (int k = 2; k <= degree; ++k) {
        const double sn = 1.0 / (2.0 / sigma1 - sigma);
        sc[3 * (k - 1)] = 2.0 * sn / ehalf;
        sc[3 * (k - 1) + 1] = -sigma * sn;
        sc[3 * (k - 1) + 2] = -cen;
        sigma = sn;
    }
    // odd steps write D1, even steps D2; the last step must land in y
    void *D1 = (degree % 2 == 1) ? y : w;
    void *D2 = (degree % 2 == 1) ? w : y;
    const long ld1 = (degree % 2 == 1) ? ldy : ldw;
    const long ld2 = (degree % 2 == 1) ? ldw : ldy;
    // table element: the real scalar type of the kernels' (a, b, shift)
    // (float for single precision and bfloat16 storage, complex included)
    const size_t ce = gpu::elem_single(e) ? sizeof(float) : sizeof(double);
    auto run = [&](bool table) {
        auto cf = [&](int k) -> const void * {
            return table ? static_cast<const char *>(c->fcoef.p) + (size_t)3 * (k - 1) * ce : nullptr;
        };
        apply_op(c, e, x, ldx, nullptr, 0, D1, ld1, ncol, sc[0], sc[1], sc[2], cf(1));
        const void *prev = D1;   // Y_{k-1}
        long ldp = ld1;
        const void *pprev = x;   // Y_{k-2}
        long ldpp = ldx;
        for (int k = 2; k <= degree; ++k) {
            void *dst = (k % 2 == 1) ? D1 : D2;
            const long ldd = (k % 2 == 1) ? ld1 : ld2;
            apply_op(c, e, prev, ldp, pprev, ldpp, dst, ldd, ncol, sc[3 * (k - 1)], sc[3 * (k - 1) + 1],
                     sc[3 * (k - 1) + 2], cf(k));
            pprev = prev; ldpp = ldp;
            prev = dst; ldp = ldd;
        }
    };
    if (!c->use_graphs || degree > kMaxGraphDegree) {
        run(false);
        return;
    }
    const gpu::StepCfg cfg = step_config(c, e, x, ldx, D1, ld1, ncol);
    if (c->has_nl) {
        // beta sized before the key: the graph's launches hold its address
        const gpu::NlDev &nd = nl_for(c, e);
        c->beta.get((size_t)ncol * nd.n_atom * nd.npm * gpu::elem_size(gpu::elem_compute(e)));
    }
    // everything a captured launch holds by value: buffers, shapes, the
    // set-up epoch (grid, atoms, projectors, forced launch shapes), Veff and
    // beta addresses, the k-point (Bloch phases of the stencil) and the shape
    long kb[3];
    std::memcpy(kb, c->kpt, sizeof(kb));
    std::vector<long> key = {e, (long)(uintptr_t)x, ldx, (long)(uintptr_t)y, ldy, (long)(uintptr_t)w, ldw, ncol,
                             degree, c->epoch, (long)(uintptr_t)veff_for(c, e), (long)(uintptr_t)c->beta.p,
                             kb[0], kb[1], kb[2], cfg.rp_div, cfg.nzch, cfg.nt, cfg.vec, cfg.align,
                             c->timing ? 1L : 0L};
    lynx_ctx::FilterGraph *g = nullptr;
    for (auto &fg : c->graphs)
        if (fg.key == key) g = &fg;
    if (!g && std::find(c->graph_seen.begin(), c->graph_seen.end(), key) == c->graph_seen.end()) {
        // first call with this key: direct launches
        if (c->graph_seen.size() >= 16) c->graph_seen.erase(c->graph_seen.begin());
        c->graph_seen.push_back(key);
        run(false);
        return;
    }
    if (!c->fcoef.p) {
        c->fcoef.get((size_t)3 * kMaxGraphDegree * sizeof(double));
        HIP_OK(hipHostMalloc(&c->fcoef_host, (size_t)3 * kMaxGraphDegree * sizeof(double)));
        HIP_OK(hipEventCreateWithFlags(&c->fcoef_copied, hipEventDisableTiming));
    }
    if (!g) {
        // second call: capture (graphs of older set-up epochs are dropped)
        for (auto &old : c->graphs)
            if (old.key[9] != c->epoch) {
                (void)hipGraphExecDestroy(old.exec);
                old.exec = nullptr;
            }
        c->graphs.erase(std::remove_if(c->graphs.begin(), c->graphs.end(),
                                       [](const lynx_ctx::FilterGraph &fg) { return !fg.exec; }),
                        c->graphs.end());
        if (c->graphs.size() >= 8) {
            (void)hipGraphExecDestroy(c->graphs.front().exec);
            c->graphs.erase(c->graphs.begin());
        }
        // timing events are created before the capture (three kernels and
        // six events per step)
        if (c->timing) ev_index(c, c->ev_marks.size() * 2 + (size_t)6 * degree + 2);
        hipGraph_t graph = nullptr;
        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            run(true);
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIP_OK(hipStreamEndCapture(c->stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t r = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_OK(r);
        c->graphs.push_back({key, exec, c->ev_marks});
        g = &c->graphs.back();
    }
    if (c->timing) c->ev_marks = g->marks;
    // the table in the arithmetic type, through the pinned staging buffer
    // (free again once the previous call's upload has run)
    HIP_OK(hipEventSynchronize(c->fcoef_copied));
    if (ce == sizeof(double)) std::memcpy(c->fcoef_host, sc.data(), sc.size() * ce);
    else
        for (size_t i = 0; i < sc.size(); ++i) static_cast<float *>(c->fcoef_host)[i] = (float)sc[i];
    HIP_OK(hipMemcpyAsync(c->fcoef.p, c->fcoef_host, sc.size() * ce, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipEventRecord(c->fcoef_copied, c->stream));
    HIP_OK(hipGraphLaunch(g->exec, c->stream));
}

long padded_ld(long nd) { return (nd + 63) / 64 * 64; }

// ------------------------------------------------------ subspace GEMMs --
// Plain GEMMs on rocBLAS (MFMA).  Gram: C = alpha A^H B (n x n) with the
// long K = Nd dimension split into chunks, one strided-batched GEMM over the
// chunks plus the tail, then a fixed-order sum of the partial products, so
// that a few hundred workgroups cover the small n x n output.
int wide_of(int e) { return (e == gpu::F32) ? gpu::F64 : (e == gpu::C64 ? gpu::C128 : e); }

// Mixed pass, by the spread of the Cholesky factor's diagonal (a lower bound
// of the block's condition number kappa): one Cholesky-QR leaves the block
// orthonormal to about kappa^2 times the FP32 rounding of the overlap
// (~1e-6 with partial sums of a few thousand terms); a second pass restores
// it while kappa^2 * 1e-6 << 1.  Beyond kCholQrFp64Ratio the pass is redone
// in FP64.
constexpr double kCholQr2Ratio = 10.0;
constexpr double kCholQrFp64Ratio = 100.0;

// Single-precision operands (the mixed-precision pass) use up to 64 chunks,
// so that each partial product sums at most a few thousand terms in FP32
// before the FP64 reduction.  C is FP64 (wide_of(e)).
// hermitian: the caller takes the result's Hermitian part (the projection
// X^H (H X)), so for FP64 only its upper tiles need forming.
void gram(lynx_ctx *c, int e, const void *A, long lda, const void *B, long ldb, int n, double alpha, void *C,
          bool hermitian = false)
{
    const long nd = c->nd;
    if (e == gpu::F32 && c->gram_precision != LYNX_GRAM_FP32 && lda % 4 == 0 && ldb % 4 == 0 &&
        reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0) {
        const int ns = gpu::gram_bf16x3_splits(nd, n);
        void *W = c->gwork.get((size_t)ns * n * n * sizeof(float));
        const int used = gpu::gram_bf16x3(static_cast<const float *>(A), lda, static_cast<const float *>(B), ldb, nd,
                                          n, static_cast<float *>(W), ns, c->stream);
        gpu::gram_reduce(e, W, used, n, alpha, C, c->stream);
        return;
    }
    const bool sym = hermitian || (A == B && lda == ldb);
    if (e == gpu::F64 && sym &&
        gpu::gram_f64_fits(static_cast<const double *>(A), lda, static_cast<const double *>(B), ldb)) {
        // a symmetric result (the overlap A^T A, or the projection whose
        // Hermitian part is taken): the hand-written FP64 MFMA kernel forms
        // only the upper tiles (5.8 against 7.5 ms for rocBLAS's full
        // product at 729,000 x 480); general products stay rocBLAS GEMMs,
        // which run them faster than this kernel
        const int ns = gpu::gram_f64_splits(nd, n);
        void *W = c->gwork.get((size_t)ns * n * n * sizeof(double));
        const int used = gpu::gram_f64(static_cast<const double *>(A), lda, static_cast<const double *>(B), ldb, nd,
                                       n, static_cast<double *>(W), ns, true, c->stream);
        gpu::gram_reduce(e, W, used, n, alpha, C, c->stream);
        return;
    }
    const bool single = e == gpu::F32 || e == gpu::C64;
    const int ks = (int)std::max<long>(1, std::min<long>(single ? 64 : 16, nd / 8192));
    const long kch = nd / ks;
    const long tail = nd - (long)ks * kch;
    const int nsplit = ks + (tail > 0 ? 1 : 0);
    const size_t es = gpu::elem_size(e);
    void *W = c->gwork.get((size_t)nsplit * n * n * es);
    const rocblas_stride sc = (rocblas_stride)n * n;
    if (e == gpu::F64) {
        const double one = 1.0, zero = 0.0;
        BLAS_OK(rocblas_dgemm_strided_batched(c->blas, rocblas_operation_transpose, rocblas_operation_none, n, n,
                                              (int)kch, &one, (const double *)A, (int)lda, kch, (const double *)B,
                                              (int)ldb, kch, &zero, (double *)W, n, sc, ks));
        if (tail > 0)
            BLAS_OK(rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, n, n, (int)tail,
                                  &one, (const double *)A + ks * kch, (int)lda, (const double *)B + ks * kch,
                                  (int)ldb, &zero, (double *)W + (size_t)ks * n * n, n));
    } else if (e == gpu::C128) {
        const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
        using Z = rocblas_double_complex;
        BLAS_OK(rocblas_zgemm_strided_batched(c->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none,
                                              n, n, (int)kch, &one, (const Z *)A, (int)lda, kch, (const Z *)B,
                                              (int)ldb, kch, &zero, (Z *)W, n, sc, ks));
        if (tail > 0)
            BLAS_OK(rocblas_zgemm(c->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, n,
                                  (int)tail, &one, (const Z *)A + ks * kch, (int)lda, (const Z *)B + ks * kch,
                                  (int)ldb, &zero, (Z *)W + (size_t)ks * n * n, n));
    } else if (e == gpu::F32) {
        const float one = 1.0f, zero = 0.0f;
        BLAS_OK(rocblas_sgemm_strided_batched(c->blas, rocblas_operation_transpose, rocblas_operation_none, n, n,
                                              (int)kch, &one, (const float *)A, (int)lda, kch, (const float *)B,
                                              (int)ldb, kch, &zero, (float *)W, n, sc, ks));
        if (tail > 0)
            BLAS_OK(rocblas_sgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, n, n, (int)tail,
                                  &one, (const float *)A + ks * kch, (int)lda, (const float *)B + ks * kch,
                                  (int)ldb, &zero, (float *)W + (size_t)ks * n * n, n));
    } else {
        const rocblas_float_complex one(1.0f, 0.0f), zero(0.0f, 0.0f);
        using Z = rocblas_float_complex;
        BLAS_OK(rocblas_cgemm_strided_batched(c->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none,
                                              n, n, (int)kch, &one, (const Z *)A, (int)lda, kch, (const Z *)B,
                                              (int)ldb, kch, &zero, (Z *)W, n, sc, ks));
        if (tail > 0)
            BLAS_OK(rocblas_cgemm(c->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, n,
                                  (int)tail, &one, (const Z *)A + ks * kch, (int)lda, (const Z *)B + ks * kch,
                                  (int)ldb, &zero, (Z *)W + (size_t)ks * n * n, n));
    }
    gpu::gram_reduce(e, W, nsplit, n, alpha, C, c->stream);
}

// Rotation: C (Nd x n) = A (Nd x n) Q (n x n).
void rotate(lynx_ctx *c, int e, const void *A, long lda, const void *Q, int n, void *C, long ldc)
{
    const int nd = c->nd;
    if (e == gpu::F64) {
        const double one = 1.0, zero = 0.0;
        BLAS_OK(rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nd, n, n, &one,
                              (const double *)A, (int)lda, (const double *)Q, n, &zero, (double *)C, (int)ldc));
    } else if (e == gpu::C128) {
        const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
        using Z = rocblas_double_complex;
        BLAS_OK(rocblas_zgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nd, n, n, &one, (const Z *)A,
                              (int)lda, (const Z *)Q, n, &zero, (Z *)C, (int)ldc));
    } else if (e == gpu::F32) {
        const float one = 1.0f, zero = 0.0f;
        BLAS_OK(rocblas_sgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nd, n, n, &one,
                              (const float *)A, (int)lda, (const float *)Q, n, &zero, (float *)C, (int)ldc));
    } else {
        const rocblas_float_complex one(1.0f, 0.0f), zero(0.0f, 0.0f);
        using Z = rocblas_float_complex;
        BLAS_OK(rocblas_cgemm(c->blas, rocblas_operation_none, rocblas_operation_none, nd, n, n, &one, (const Z *)A,
                              (int)lda, (const Z *)Q, n, &zero, (Z *)C, (int)ldc));
    }
}

void *block(DevBuf &b, int e, long ld, int ncol) { return b.get((size_t)ld * ncol * gpu::elem_size(e)); }

// ------------------------------------------------------- dense FP64 steps --
void check_info(lynx_ctx *c, int status, const char *what)
{
    int info = 0;
    HIP_OK(hipMemcpyAsync(&info, c->info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (info != 0) fail(status, std::string(what) + " failed (info=" + std::to_string(info) + ")");
}

// S = dV X^H X; R^H R = S; returns R^-1 (upper, zero below) in c->s1, in
// FP64 whatever the precision e of X.
void cholesky_inverse(lynx_ctx *c, int e, const void *X, long ldx, int n)
{
    const int ew = wide_of(e);
    void *S = c->s1.get((size_t)n * n * gpu::elem_size(ew));
    gram(c, e, X, ldx, X, ldx, n, c->st.dV, S);
    c->info.get(sizeof(int));
    e = ew;
    if (e == gpu::F64) {
        BLAS_OK(rocsolver_dpotrf(c->blas, rocblas_fill_upper, n, (double *)S, n, c->info.as<rocblas_int>()));
        check_info(c, LYNX_E_CHOLESKY, "Cholesky factorisation of the overlap matrix");
        BLAS_OK(rocsolver_dtrtri(c->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, n, (double *)S, n,
                                 c->info.as<rocblas_int>()));
    } else {
        BLAS_OK(rocsolver_zpotrf(c->blas, rocblas_fill_upper, n, (rocblas_double_complex *)S, n,
                                 c->info.as<rocblas_int>()));
        check_info(c, LYNX_E_CHOLESKY, "Cholesky factorisation of the overlap matrix");
        BLAS_OK(rocsolver_ztrtri(c->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, n,
                                 (rocblas_double_complex *)S, n, c->info.as<rocblas_int>()));
    }
    check_info(c, LYNX_E_CHOLESKY, "triangular inverse");
    gpu::zero_lower(e, S, n, c->stream);
}

// Spread max|r_ii| / min|r_ii| of the Cholesky factor whose inverse
// cholesky_inverse left in c->s1 (FP64 or complex128, n x n): a lower bound
// of the condition number of the factorised block.
double triangular_diag_ratio(lynx_ctx *c, int ew, int n)
{
    const size_t es = gpu::elem_size(ew);
    std::vector<double> d(ew == gpu::C128 ? 2 * n : n);
    // gather the diagonal on the device (column j's element j: n columns of
    // one element at leading dimension n + 1), then one contiguous copy; a
    // strided 2D copy of n single elements runs as n separate transfers
    void *diag = c->evals.get((size_t)n * es);
    gpu::copy_cols(ew, c->s1.p, n + 1, diag, 1, 1, n, c->stream);
    HIP_OK(hipMemcpyAsync(d.data(), diag, (size_t)n * es, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    double lo = 1e300, hi = 0;
    for (int i = 0; i < n; ++i) {
        const double v = ew == gpu::C128 ? std::hypot(d[2 * i], d[2 * i + 1]) : std::fabs(d[i]);
        if (!std::isfinite(v)) return 1e300;   // a non-finite factor: the pass is redone in FP64
        lo = std::min(lo, v);
        hi = std::max(hi, v);
    }
    return lo > 0 ? hi / lo : 1e300;
}

// Hermitian eigensolve of c->s2 (n x n) in place; eigenvalues to host.
void subspace_eig(lynx_ctx *c, int e, int n, double *eig_host)
{
    double *D = (double *)c->dvals.get(n * sizeof(double));
    double *E = (double *)c->evals.get(n * sizeof(double));
    c->info.get(sizeof(int));
    if (e == gpu::F64)
        BLAS_OK(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, n, (double *)c->s2.p, n, D,
                                 E, c->info.as<rocblas_int>()));
    else
        BLAS_OK(rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper, n,
                                 (rocblas_double_complex *)c->s2.p, n, D, E, c->info.as<rocblas_int>()));
    check_info(c, LYNX_E_EIG, "subspace eigensolve");
    HIP_OK(hipMemcpyAsync(eig_host, D, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(eig_host[i])) fail(LYNX_E_NONFINITE, "non-finite Ritz value");
}

// Sturm-sequence bisection for the extremal eigenvalues of the symmetric
// tridiagonal matrix (d, e); the reference obtains all of them with dsterf.
double tridiag_extreme(const double *d, const double *e, int n, bool largest)
{
    double lo = d[0], hi = d[0];
    for (int i = 0; i < n; ++i) {
        double r = 0;
        if (i > 0) r += std::fabs(e[i - 1]);
        if (i < n - 1) r += std::fabs(e[i]);
        lo = std::min(lo, d[i] - r);
        hi = std::max(hi, d[i] + r);
    }
    const double span = hi - lo + 1e-300;
    lo -= 1e-12 * span;
    hi += 1e-12 * span;
    auto below = [&](double x) {
        int cnt = 0;
        double q = d[0] - x;
        if (q < 0) ++cnt;
        for (int i = 1; i < n; ++i) {
            if (q == 0.0) q = 1e-300;
            q = d[i] - x - e[i - 1] * e[i - 1] / q;
            if (q < 0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200 && hi - lo > 4e-16 * (std::fabs(lo) + std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        const int k = below(mid);
        if (!largest) { if (k >= 1) hi = mid; else lo = mid; }
        else { if (k >= n) hi = mid; else lo = mid; }
    }
    return 0.5 * (lo + hi);
}

double host_sum(lynx_ctx *c, int nb)
{
    c->hpart.resize(nb);
    HIP_OK(hipMemcpyAsync(c->hpart.data(), c->partials.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    double s = 0;
    for (int i = 0; i < nb; ++i) s += c->hpart[i];
    return s;
}

const ElecHost &electrostatics(lynx_ctx *c)
{
    need_grid(c);
    if (!c->has_atoms) fail(LYNX_E_STATE, "lynx_set_atoms has not been called");
    if (!c->has_elec) {
        build_electrostatics(c->lat, c->st, c->psps, c->atom_types, c->atom_pos, c->elec);
        c->has_elec = true;
    }
    return c->elec;
}

}  // namespace detail
}  // namespace lynx

using namespace lynx::detail;

// ======================================================================= API

extern "C" {

const char *lynx_status_string(lynx_status s)
{
    switch (s) {
        case LYNX_OK: return "ok";
        case LYNX_E_ARG: return "invalid argument";
        case LYNX_E_HIP: return "HIP runtime or library failure";
        case LYNX_E_CHOLESKY: return "overlap matrix not positive definite";
        case LYNX_E_EIG: return "eigensolver failure";
        case LYNX_E_NONFINITE: return "non-finite value";
        case LYNX_E_IO: return "pseudopotential file error";
        case LYNX_E_STATE: return "required set-up call missing";
        case LYNX_E_UNSUPPORTED: return "unsupported input";
        default: return "unknown status";
    }
}

lynx_status lynx_ctx_create(lynx_ctx **out, int device_id)
{
    if (!out) return LYNX_E_ARG;
    *out = nullptr;
    lynx_ctx *c = new lynx_ctx();
    c->device = device_id;
    if (device_id == -1) {   // host-only context: set-up and inspection, no GPU
        *out = c;
        return LYNX_OK;
    }
    lynx_status s = guarded(c, [&] {
        int n = 0;
        HIP_OK(hipGetDeviceCount(&n));
        if (device_id < 0 || device_id >= n) fail(LYNX_E_ARG, "no such HIP device");
        HIP_OK(hipSetDevice(device_id));
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        BLAS_OK(rocblas_create_handle(&c->blas));
        BLAS_OK(rocblas_set_stream(c->blas, c->stream));
        c->scratch = scratch_address(device_id);
    });
    if (s != LYNX_OK) {
        delete c;
        return s;
    }
    *out = c;
    return LYNX_OK;
}

void lynx_ctx_destroy(lynx_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
    delete ctx;
}

const char *lynx_last_error(const lynx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

lynx_status lynx_ctx_stream(const lynx_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return LYNX_E_ARG;
    *stream = ctx->stream;
    return LYNX_OK;
}

lynx_status lynx_ctx_synchronize(lynx_ctx *ctx)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] { HIP_OK(hipStreamSynchronize(ctx->stream)); });
}

lynx_status lynx_set_grid(lynx_ctx *ctx, int nx, int ny, int nz, const double latvec[9], int cell_type,
                          int fd_order)
{
    if (!ctx || !latvec) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        ++ctx->epoch;   // captured filter graphs hold the old grid / projector data
        if (nx <= 0 || ny <= 0 || nz <= 0) fail(LYNX_E_ARG, "grid dimensions must be positive");
        if ((long)nx * ny * nz > (1L << 31) - 1) fail(LYNX_E_ARG, "grid too large");
        if (cell_type != LYNX_CELL_ORTHOGONAL && cell_type != LYNX_CELL_NONORTHOGONAL)
            fail(LYNX_E_ARG, "unknown cell type");
        Lattice lat;
        lattice_init(lat, latvec, cell_type == LYNX_CELL_ORTHOGONAL);
        Stencil st;
        stencil_init(st, fd_order, nx, ny, nz, lat);
        ctx->lat = lat;
        ctx->st = st;
        ctx->nd = nx * ny * nz;
        ctx->has_grid = true;
        ctx->has_veff = false;
        ctx->has_nl = false;
        ctx->nlh = NonlocalHost{};
        ctx->has_atoms = ctx->has_elec = false;
    });
}

lynx_status lynx_get_grid_info(const lynx_ctx *ctx, double h[3], double *dV, double *max_eig)
{
    if (!ctx) return LYNX_E_ARG;
    if (!ctx->has_grid) return LYNX_E_STATE;
    if (h) for (int a = 0; a < 3; ++a) h[a] = ctx->st.h[a];
    if (dV) *dV = ctx->st.dV;
    if (max_eig) *max_eig = ctx->st.max_eig;
    return LYNX_OK;
}

lynx_status lynx_get_stencil(const lynx_ctx *ctx, double *D1, double *D2, double *D2m)
{
    if (!ctx) return LYNX_E_ARG;
    if (!ctx->has_grid) return LYNX_E_STATE;
    const int m = ctx->st.fdn + 1;
    for (int a = 0; a < 3; ++a)
        for (int p = 0; p < m; ++p) {
            if (D1) D1[a * m + p] = ctx->st.D1[a][p];
            if (D2) D2[a * m + p] = ctx->st.D2[a][p];
            if (D2m) D2m[a * m + p] = ctx->st.D2m[a][p];
        }
    return LYNX_OK;
}

lynx_status lynx_set_atoms(lynx_ctx *ctx, int ntypes, const char *const *paths, int natom, const int *atom_type,
                           const double *coords, int frac)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        ++ctx->epoch;   // captured filter graphs hold the old grid / projector data
        need_grid(ctx);
        if (ntypes <= 0 || natom < 0 || !paths || (natom > 0 && (!atom_type || !coords)))
            fail(LYNX_E_ARG, "bad atom arguments");
        std::vector<Psp8> psps(ntypes);
        for (int t = 0; t < ntypes; ++t) psp8_load(paths[t], psps[t]);
        std::vector<int> types(atom_type, atom_type + natom);
        for (int i = 0; i < natom; ++i) {
            if (types[i] < 0 || types[i] >= ntypes) fail(LYNX_E_ARG, "atom type out of range");
            if (i > 0 && types[i] < types[i - 1]) fail(LYNX_E_ARG, "atoms must be listed type-major");
        }
        // Crystal.cpp:284-295: lattice-aligned coordinates
        std::vector<double> pos(3 * natom);
        for (int i = 0; i < natom; ++i) {
            const double *v = coords + 3 * i;
            double *o = &pos[3 * i];
            if (frac) {
                if (ctx->lat.orth) {
                    for (int j = 0; j < 3; ++j)
                        o[j] = v[0] * ctx->lat.latvec[j] + v[1] * ctx->lat.latvec[3 + j] + v[2] * ctx->lat.latvec[6 + j];
                } else {
                    for (int j = 0; j < 3; ++j) o[j] = v[j] * ctx->lat.L[j];
                }
            } else if (ctx->lat.orth) {
                for (int j = 0; j < 3; ++j) o[j] = v[j];
            } else {
                // cart -> fractional (rows of latvec are vectors) -> scaled
                double inv[9], f[3];
                const double *a = ctx->lat.latvec;
                const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                                   a[2] * (a[3] * a[7] - a[4] * a[6]);
                inv[0] = (a[4] * a[8] - a[5] * a[7]) / det; inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
                inv[2] = (a[1] * a[5] - a[2] * a[4]) / det; inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
                inv[4] = (a[0] * a[8] - a[2] * a[6]) / det; inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
                inv[6] = (a[3] * a[7] - a[4] * a[6]) / det; inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
                inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
                for (int j = 0; j < 3; ++j) f[j] = v[0] * inv[j] + v[1] * inv[3 + j] + v[2] * inv[6 + j];
                for (int j = 0; j < 3; ++j) o[j] = f[j] * ctx->lat.L[j];
            }
        }
        build_nonlocal(ctx->lat, ctx->st, psps, types, pos, ctx->nlh);
        ctx->psps = std::move(psps);
        ctx->atom_types = std::move(types);
        ctx->atom_pos = std::move(pos);
        ctx->has_atoms = true;
        ctx->has_elec = false;
        if (ctx->device >= 0) {
            upload_nonlocal(ctx);
            upload_phases(ctx);
        }
    });
}

lynx_status lynx_set_nonlocal(lynx_ctx *ctx, int n_img, const int *img_atom, const int *img_ndc,
                              const int *img_nproj, const int *gpos, const double *chi, const double *img_shift,
                              int n_atom, const int *ip_displ, const double *gamma)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        ++ctx->epoch;   // captured filter graphs hold the old grid / projector data
        need_grid(ctx);
        if (n_img < 0 || n_atom < 0) fail(LYNX_E_ARG, "negative sizes");
        NonlocalHost h;
        h.n_img = n_img;
        h.n_atom = n_atom;
        h.img_atom.assign(img_atom, img_atom + n_img);
        h.img_ndc.assign(img_ndc, img_ndc + n_img);
        h.img_nproj.assign(img_nproj, img_nproj + n_img);
        long ng = 0, nc = 0;
        for (int i = 0; i < n_img; ++i) {
            if (h.img_atom[i] < 0 || h.img_atom[i] >= n_atom) fail(LYNX_E_ARG, "image atom out of range");
            if (h.img_ndc[i] < 0 || h.img_nproj[i] < 0) fail(LYNX_E_ARG, "negative image size");
            ng += h.img_ndc[i];
            nc += (long)h.img_ndc[i] * h.img_nproj[i];
        }
        h.gpos.assign(gpos, gpos + ng);
        h.chi.assign(chi, chi + nc);
        h.img_shift.assign(img_shift, img_shift + 3L * n_img);
        h.ip_displ.assign(ip_displ, ip_displ + n_atom + 1);
        for (int a = 0; a < n_atom; ++a)
            if (h.ip_displ[a + 1] < h.ip_displ[a]) fail(LYNX_E_ARG, "ip_displ must be non-decreasing");
        for (int i = 0; i < n_img; ++i) {
            const int a = h.img_atom[i];
            if (h.img_nproj[i] != h.ip_displ[a + 1] - h.ip_displ[a])
                fail(LYNX_E_ARG, "image projector count differs from its atom's");
        }
        h.gamma.assign(gamma, gamma + h.ip_displ[n_atom]);
        for (int g : h.gpos)
            if (g < 0 || g >= ctx->nd) fail(LYNX_E_ARG, "nonlocal grid index out of range");
        ctx->nlh = std::move(h);
        ctx->has_atoms = ctx->has_elec = false;
        if (ctx->device >= 0) {
            upload_nonlocal(ctx);
            upload_phases(ctx);
        }
    });
}

lynx_status lynx_get_electrostatics(lynx_ctx *ctx, double *b, double *b_ref, double *vc, double *rho_atom,
                                    double *rho_core, double *scalars)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        need_grid(ctx);
        if (!ctx->has_atoms) fail(LYNX_E_STATE, "lynx_set_atoms has not been called");
        if (!ctx->has_elec) {
            build_electrostatics(ctx->lat, ctx->st, ctx->psps, ctx->atom_types, ctx->atom_pos, ctx->elec);
            ctx->has_elec = true;
        }
        const ElecHost &E = ctx->elec;
        const size_t bytes = E.b.size() * sizeof(double);
        if (b) std::memcpy(b, E.b.data(), bytes);
        if (b_ref) std::memcpy(b_ref, E.b_ref.data(), bytes);
        if (vc) std::memcpy(vc, E.vc.data(), bytes);
        if (rho_atom) std::memcpy(rho_atom, E.rho_at.data(), bytes);
        if (rho_core) std::memcpy(rho_core, E.rho_core.data(), bytes);
        if (scalars) {
            scalars[0] = E.eself;
            scalars[1] = E.ec;
            scalars[2] = E.int_b;
            scalars[3] = E.nelectron;
            scalars[4] = E.has_nlcc ? 1.0 : 0.0;
            scalars[5] = E.n_boxes;
        }
    });
}

lynx_status lynx_get_nonlocal_sizes(const lynx_ctx *ctx, int *n_img, long *n_gpos, long *n_chi, int *n_atom,
                                    int *n_proj)
{
    if (!ctx) return LYNX_E_ARG;
    const NonlocalHost &h = ctx->nlh;
    if (n_img) *n_img = h.n_img;
    if (n_gpos) *n_gpos = (long)h.gpos.size();
    if (n_chi) *n_chi = (long)h.chi.size();
    if (n_atom) *n_atom = h.n_atom;
    if (n_proj) *n_proj = h.ip_displ.empty() ? 0 : h.ip_displ.back();
    return LYNX_OK;
}

lynx_status lynx_get_nonlocal(const lynx_ctx *ctx, int *img_atom, int *img_ndc, int *img_nproj, int *gpos,
                              double *chi, double *img_shift, int *ip_displ, double *gamma)
{
    if (!ctx) return LYNX_E_ARG;
    const NonlocalHost &h = ctx->nlh;
    auto cp = [](auto *dst, const auto &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(img_atom, h.img_atom); cp(img_ndc, h.img_ndc); cp(img_nproj, h.img_nproj); cp(gpos, h.gpos);
    cp(chi, h.chi); cp(img_shift, h.img_shift); cp(ip_displ, h.ip_displ); cp(gamma, h.gamma);
    return LYNX_OK;
}

lynx_status lynx_set_veff(lynx_ctx *ctx, const double *veff, int mem)
{
    if (!ctx || !veff) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        need_device(ctx);
        need_grid(ctx);
        const size_t n = ctx->nd;
        // the device copy is overwritten before the check: a refused Veff
        // leaves the context without one until the next successful call
        ++ctx->veff_gen;
        ctx->has_veff = false;
        ctx->veff64.get(n * sizeof(double));
        HIP_OK(hipMemcpyAsync(ctx->veff64.p, veff, n * sizeof(double),
                              mem == LYNX_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              ctx->stream));
        // finiteness on the device: 0*v is 0 for a finite v and NaN for an
        // infinite or NaN one, so sum_i v_i (0 v_i) is finite exactly when
        // every entry is (and cannot overflow)
        void *zv = ctx->lz[0].get(n * sizeof(double));
        ctx->partials.get(gpu::kPartialBlocks * sizeof(double));
        gpu::scale_copy(gpu::F64, ctx->veff64.p, zv, 0.0, (long)n, ctx->stream);
        const double chk = host_sum(
            ctx, gpu::dot_partials(gpu::F64, ctx->veff64.p, zv, (long)n, ctx->partials.as<double>(), ctx->stream));
        if (!std::isfinite(chk)) fail(LYNX_E_NONFINITE, "non-finite effective potential");
        // the FP32 copy for the single-precision kernels (round to nearest)
        ctx->veff32.get(n * sizeof(float));
        gpu::convert_cols(gpu::F64, ctx->veff64.p, (long)n, gpu::F32, ctx->veff32.p, (long)n, (long)n, 1, ctx->stream);
        HIP_OK(hipStreamSynchronize(ctx->stream));
        ctx->has_veff = true;
    });
}

lynx_status lynx_set_kpoint(lynx_ctx *ctx, const double kpt[3])
{
    if (!ctx || !kpt) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        for (int a = 0; a < 3; ++a) ctx->kpt[a] = kpt[a];
        if (ctx->nlh.n_img > 0 && ctx->device >= 0) upload_phases(ctx);
    });
}

lynx_status lynx_h_apply(lynx_ctx *ctx, int dtype, const void *psi, int ld_psi, void *y, int ld_y, int ncol,
                         double c, int mem)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_gpu_stencil(ctx);
        need_veff(ctx);
        if (ncol <= 0 || !psi || !y) fail(LYNX_E_ARG, "bad block arguments");
        const long nd = ctx->nd;
        const long ldp = ld_psi ? ld_psi : nd, ldy = ld_y ? ld_y : nd;
        if (ldp < nd || ldy < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        const size_t es = gpu::elem_size(dtype);
        timing_begin(ctx);
        if (mem == LYNX_MEM_DEVICE) {
            apply_op(ctx, dtype, psi, ldp, nullptr, 0, y, ldy, ncol, 1.0, 0.0, c);
        } else {
            void *dx = block(ctx->wA, dtype, nd, ncol);
            void *dy = block(ctx->wB, dtype, nd, ncol);
            HIP_OK(hipMemcpy2DAsync(dx, nd * es, psi, ldp * es, nd * es, ncol, hipMemcpyHostToDevice, ctx->stream));
            apply_op(ctx, dtype, dx, nd, nullptr, 0, dy, nd, ncol, 1.0, 0.0, c);
            HIP_OK(hipMemcpy2DAsync(y, ldy * es, dy, nd * es, nd * es, ncol, hipMemcpyDeviceToHost, ctx->stream));
        }
        HIP_OK(hipStreamSynchronize(ctx->stream));
        timing_end(ctx);
    });
}

lynx_status lynx_h_apply_noalloc(lynx_ctx *ctx, int dtype, const void *psi, int ld_psi, void *y, int ld_y,
                                 int ncol, double c)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_gpu_stencil(ctx);
        need_veff(ctx);
        if (ncol <= 0 || !psi || !y) fail(LYNX_E_ARG, "bad block arguments");
        if (!ctx->scratch) fail(LYNX_E_STATE, "no scratch array on this context's device");
        const long nd = ctx->nd;
        const long ldp = ld_psi ? ld_psi : nd, ldy = ld_y ? ld_y : nd;
        if (ldp < nd || ldy < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        const int e = dtype;
        const bool nl = ctx->has_nl;
        const gpu::NlDev &nld = nl_for(ctx, e);
        if (nl) {
            const size_t per_col = (size_t)nld.n_atom * nld.npm * gpu::elem_size(gpu::elem_compute(e));
            if ((size_t)ncol * per_col > kScratchBytes)
                fail(LYNX_E_ARG, "the projector coefficients of " + std::to_string(ncol) +
                                     " columns exceed the 32 MiB scratch array: at most " +
                                     std::to_string(kScratchBytes / per_col) + " columns per call");
        }
        // launch shape: the forced one, or the one already tuned for this
        // grid, dtype and column count, else the default shape (tuning
        // creates timing events, which this entry point does not)
        gpu::StepCfg cfg{};
        if (ctx->has_forced) {
            cfg = ctx->forced;
        } else {
            const std::vector<long> key = {ctx->st.nx, ctx->st.ny, ctx->st.nz, e, ncol};
            auto it = ctx->tuned.find(key);
            if (it != ctx->tuned.end()) cfg = it->second;
        }
        if (nl) gpu::nl_gather(e, nld, psi, ldp, ctx->scratch, ncol, ctx->stream);
        gpu::stencil_step(e, geom(ctx), psi, ldp, nullptr, 0, y, ldy, veff_for(ctx, e), ncol, 1.0, 0.0, c, cfg,
                          ctx->stream);
        if (nl) gpu::nl_scatter(e, nld, ctx->scratch, y, ldy, ncol, 1.0, ctx->stream);
        HIP_OK(hipGetLastError());
    });
}

lynx_status lynx_cheb_filter(lynx_ctx *ctx, int dtype, const void *x, int ld_x, void *y, int ld_y, int ncol,
                             int degree, double lambda_c, double eig_min, double eig_max, int mem)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_gpu_stencil(ctx);
        need_veff(ctx);
        if (ncol <= 0 || !x || !y) fail(LYNX_E_ARG, "bad block arguments");
        const long nd = ctx->nd;
        const long ldx = ld_x ? ld_x : nd, ldy = ld_y ? ld_y : nd;
        if (ldx < nd || ldy < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        const size_t es = gpu::elem_size(dtype);
        const long ldw = padded_ld(nd);
        timing_begin(ctx);
        if (mem == LYNX_MEM_DEVICE) {
            void *w = block(ctx->wC, dtype, ldw, ncol);
            filter(ctx, dtype, x, ldx, y, ldy, w, ldw, ncol, degree, lambda_c, eig_min, eig_max);
        } else {
            void *dx = block(ctx->wA, dtype, nd, ncol);
            void *dy = block(ctx->wB, dtype, nd, ncol);
            void *w = block(ctx->wC, dtype, ldw, ncol);
            HIP_OK(hipMemcpy2DAsync(dx, nd * es, x, ldx * es, nd * es, ncol, hipMemcpyHostToDevice, ctx->stream));
            filter(ctx, dtype, dx, nd, dy, nd, w, ldw, ncol, degree, lambda_c, eig_min, eig_max);
            HIP_OK(hipMemcpy2DAsync(y, ldy * es, dy, nd * es, nd * es, ncol, hipMemcpyDeviceToHost, ctx->stream));
        }
        HIP_OK(hipStreamSynchronize(ctx->stream));
        timing_end(ctx);
    });
}

lynx_status lynx_orthogonalize(lynx_ctx *ctx, int dtype, void *x, int ld, int ncol, int mem)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_device(ctx);
        need_grid(ctx);
        if (dtype != LYNX_F64 && dtype != LYNX_C128) fail(LYNX_E_UNSUPPORTED, "orthogonalisation runs in FP64");
        if (ncol <= 0 || !x) fail(LYNX_E_ARG, "bad block arguments");
        const long nd = ctx->nd, ldx = ld ? ld : nd;
        if (ldx < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        const size_t es = gpu::elem_size(dtype);
        void *dx = x;
        long ldd = ldx;
        if (mem != LYNX_MEM_DEVICE) {
            dx = block(ctx->wA, dtype, nd, ncol);
            ldd = nd;
            HIP_OK(hipMemcpy2DAsync(dx, nd * es, x, ldx * es, nd * es, ncol, hipMemcpyHostToDevice, ctx->stream));
        }
        cholesky_inverse(ctx, dtype, dx, ldd, ncol);
        void *tmp = block(ctx->wB, dtype, padded_ld(nd), ncol);
        rotate(ctx, dtype, dx, ldd, ctx->s1.p, ncol, tmp, padded_ld(nd));
        gpu::copy_cols(dtype, tmp, padded_ld(nd), dx, ldd, nd, ncol, ctx->stream);
        if (mem != LYNX_MEM_DEVICE)
            HIP_OK(hipMemcpy2DAsync(x, ldx * es, dx, nd * es, nd * es, ncol, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

lynx_status lynx_project_hamiltonian(lynx_ctx *ctx, int dtype, const void *x, int ld, int ncol, void *hs, int mem)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_gpu_stencil(ctx);
        need_veff(ctx);
        if (dtype != LYNX_F64 && dtype != LYNX_C128) fail(LYNX_E_UNSUPPORTED, "projection runs in FP64");
        if (ncol <= 0 || !x || !hs) fail(LYNX_E_ARG, "bad block arguments");
        const long nd = ctx->nd, ldx = ld ? ld : nd;
        if (ldx < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        const size_t es = gpu::elem_size(dtype);
        const void *dx = x;
        long ldd = ldx;
        if (mem != LYNX_MEM_DEVICE) {
            void *t = block(ctx->wA, dtype, nd, ncol);
            HIP_OK(hipMemcpy2DAsync(t, nd * es, x, ldx * es, nd * es, ncol, hipMemcpyHostToDevice, ctx->stream));
            dx = t;
            ldd = nd;
        }
        const long ldw = padded_ld(nd);
        void *hx = block(ctx->wB, dtype, ldw, ncol);
        apply_op(ctx, dtype, dx, ldd, nullptr, 0, hx, ldw, ncol, 1.0, 0.0, 0.0);
        void *H = ctx->s2.get((size_t)ncol * ncol * es);
        gram(ctx, dtype, dx, ldd, hx, ldw, ncol, ctx->st.dV, H, true);
        gpu::hermitian_part(dtype, H, ncol, ctx->stream);
        HIP_OK(hipMemcpyAsync(hs, H, (size_t)ncol * ncol * es, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

lynx_status lynx_chefsi_solve(lynx_ctx *ctx, int dtype, void *psi, int ld, int nband, double *eigvals,
                              double lambda_c, double eig_min, double eig_max, int degree, double *lambda_c_out,
                              int mem)
{
    if (!ctx) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_gpu_stencil(ctx);
        need_veff(ctx);
        if (dtype != LYNX_F64 && dtype != LYNX_C128)
            fail(LYNX_E_UNSUPPORTED, "the solve takes FP64 wavefunctions (real or complex)");
        if (nband <= 0 || !psi || !eigvals) fail(LYNX_E_ARG, "bad block arguments");
        const long nd = ctx->nd, ldp = ld ? ld : nd;
        if (ldp < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        if (nband > nd) fail(LYNX_E_ARG, "more bands than grid points");
        const size_t es = gpu::elem_size(dtype);
        const long ldw = padded_ld(nd);
        void *P = psi;
        long ldP = ldp;
        if (mem != LYNX_MEM_DEVICE) {
            P = block(ctx->wA, dtype, nd, nband);
            ldP = nd;
            HIP_OK(hipMemcpy2DAsync(P, nd * es, psi, ldp * es, nd * es, nband, hipMemcpyHostToDevice, ctx->stream));
        }
        void *B = block(ctx->wB, dtype, ldw, nband);
        void *C = block(ctx->wC, dtype, ldw, nband);
        if (ctx->filter_precision != LYNX_FILTER_FP64) {
            // Mixed precision: the filter, the Nd-sized products and the
            // rotations in single precision (inside the FP64 work blocks:
            // two single-precision blocks per FP64 block); the overlap and
            // projected matrices summed and factorised in FP64 (Cholesky,
            // triangular inverse, eigensolve).  With LYNX_FILTER_BF16 (real
            // blocks whose rows allow the run-vector stencil) the filter
            // stores its blocks in bfloat16 (arithmetic in float) and the
            // filtered block is widened to FP32 for the rest of the pass.
            const int es1 = dtype == LYNX_F64 ? LYNX_F32 : LYNX_C64;
            const size_t half = (size_t)ldw * nband * gpu::elem_size(es1);
            const bool bf = ctx->filter_precision == LYNX_FILTER_BF16 && dtype == LYNX_F64 && bf16_filter_fits(ctx);
            void *Xs = B, *Ys = C;
            void *Ws = static_cast<char *>(C) + half;
            if (bf) {
                // bf16 X, Y and work block in the first quarters of B and C
                const size_t quarter = (size_t)ldw * nband * 2;
                void *Xb = B, *Yb = C, *Wb = static_cast<char *>(C) + quarter;
                gpu::convert_cols(dtype, P, ldP, LYNX_BF16, Xb, ldw, nd, nband, ctx->stream);
                timing_begin(ctx);
                filter(ctx, LYNX_BF16, Xb, ldw, Yb, ldw, Wb, ldw, nband, degree, lambda_c, eig_min, eig_max);
                timing_end(ctx);
                Ys = B;
                Xs = static_cast<char *>(B) + half;
                Ws = C;
                gpu::convert_cols(LYNX_BF16, Yb, ldw, es1, Ys, ldw, nd, nband, ctx->stream);
            } else {
                gpu::convert_cols(dtype, P, ldP, es1, Xs, ldw, nd, nband, ctx->stream);
                timing_begin(ctx);
                filter(ctx, es1, Xs, ldw, Ys, ldw, Ws, ldw, nband, degree, lambda_c, eig_min, eig_max);
                timing_end(ctx);
            }
            // a9: Cs = Ys R^-1 with dV Ys^H Ys = R^H R (R in FP64).  The
            // orthonormality of Cs degrades with the square of the block's
            // condition number times the FP32 rounding of the overlap, so a
            // block whose R spreads its diagonal by more than kCholQr2Ratio
            // gets a second Cholesky-QR pass on Cs (CholeskyQR2).  A block
            // too ill-conditioned for that (spread above kCholQrFp64Ratio, an
            // overlap that is not positive definite, or Ritz values outside
            // the spectral bounds; the first passes from a random start,
            // where the filter aligns the columns with the lowest
            // eigenvectors) is redone in FP64 from psi, still unchanged.
            auto fp32_pass = [&]() -> bool {
                auto chol = [&](const void *Y) {
                    try {
                        cholesky_inverse(ctx, es1, Y, ldw, nband);
                        return true;
                    } catch (const Error &err) {
                        if (err.status != LYNX_E_CHOLESKY) throw;
                        return false;
                    }
                };
                // with LYNX_GRAM_BF16X3_ROT the rotations of real blocks run as
                // split-BF16x3 matrix-core products from the FP64 factors
                // directly; otherwise as FP32 GEMMs of FP32 copies
                const bool rot3 = es1 == gpu::F32 && ctx->gram_precision == LYNX_GRAM_BF16X3_ROT;
                void *Rs = rot3 ? ctx->s3.get(gpu::rotate_bf16x3_work(nband, nband))
                                : ctx->s3.get((size_t)nband * nband * gpu::elem_size(es1));
                auto rot = [&](const void *A, const void *Q64, void *C, long ldc, bool c_double) {
                    if (rot3) {
                        gpu::rotate_bf16x3(static_cast<const float *>(A), ldw, nd, static_cast<const double *>(Q64),
                                           nband, nband, nband, C, ldc, c_double, Rs, ctx->stream);
                        return;
                    }
                    gpu::convert_cols(dtype, Q64, nband, es1, Rs, nband, nband, nband, ctx->stream);
                    if (!c_double) {
                        rotate(ctx, es1, A, ldw, Rs, nband, C, ldc);
                        return;
                    }
                    void *tmp = A == Xs ? Ys : Xs;    // the other single-precision block is free
                    rotate(ctx, es1, A, ldw, Rs, nband, tmp, ldw);
                    gpu::convert_cols(es1, tmp, ldw, dtype, C, ldc, nd, nband, ctx->stream);
                };
                void *Cs = Xs, *Free = Ys;
                if (!chol(Ys)) return false;
                rot(Ys, ctx->s1.p, Cs, ldw, false);
                const double spread = triangular_diag_ratio(ctx, dtype, nband);
                if (spread > kCholQrFp64Ratio) return false;
                if (spread > kCholQr2Ratio) {
                    if (!chol(Cs)) return false;
                    rot(Cs, ctx->s1.p, Free, ldw, false);
                    std::swap(Cs, Free);
                    ++ctx->cholqr2_passes;
                }
                // a10: Hs = dV Cs^H H Cs
                apply_op(ctx, es1, Cs, ldw, nullptr, 0, Ws, ldw, nband, 1.0, 0.0, 0.0);
                void *H = ctx->s2.get((size_t)nband * nband * es);
                gram(ctx, es1, Cs, ldw, Ws, ldw, nband, ctx->st.dV, H);
                gpu::hermitian_part(dtype, H, nband, ctx->stream);
                // a11 in FP64 (a failed or non-finite eigensolve of the
                // single-precision block also sends the pass to FP64)
                try {
                    subspace_eig(ctx, dtype, nband, eigvals);
                } catch (const Error &err) {
                    if (err.status != LYNX_E_EIG && err.status != LYNX_E_NONFINITE) throw;
                    return false;
                }
                if (!(eigvals[0] >= eig_min - 0.05 * (eig_max - eig_min)) || !(eigvals[nband - 1] <= eig_max))
                    return false;
                // a12: psi = Cs Q, written to the FP64 psi
                rot(Cs, H, P, ldP, true);
                return true;
            };
            if (fp32_pass()) {
                if (mem != LYNX_MEM_DEVICE)
                    HIP_OK(hipMemcpy2DAsync(psi, ldp * es, P, nd * es, nd * es, nband, hipMemcpyDeviceToHost,
                                            ctx->stream));
                HIP_OK(hipStreamSynchronize(ctx->stream));
                if (lambda_c_out) *lambda_c_out = eigvals[nband - 1] + 0.1;
                ++ctx->solves;
                return;
            }
            ++ctx->fp64_overlap_fallbacks;
        }
        timing_begin(ctx);
        // a8: Y = p_m(H) psi -> B
        filter(ctx, dtype, P, ldP, B, ldw, C, ldw, nband, degree, lambda_c, eig_min, eig_max);
        timing_end(ctx);
        // a9: C = B R^-1 with dV B^H B = R^H R
        cholesky_inverse(ctx, dtype, B, ldw, nband);
        rotate(ctx, dtype, B, ldw, ctx->s1.p, nband, C, ldw);
        // a10: Hs = dV C^H H C (B <- H C)
        apply_op(ctx, dtype, C, ldw, nullptr, 0, B, ldw, nband, 1.0, 0.0, 0.0);
        void *H = ctx->s2.get((size_t)nband * nband * es);
        gram(ctx, dtype, C, ldw, B, ldw, nband, ctx->st.dV, H, true);
        gpu::hermitian_part(dtype, H, nband, ctx->stream);
        // a11: Hs = Q diag(eig) Q^H
        subspace_eig(ctx, dtype, nband, eigvals);
        // a12: psi = C Q
        rotate(ctx, dtype, C, ldw, H, nband, P, ldP);
        if (mem != LYNX_MEM_DEVICE)
            HIP_OK(hipMemcpy2DAsync(psi, ldp * es, P, nd * es, nd * es, nband, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        if (lambda_c_out) *lambda_c_out = eigvals[nband - 1] + 0.1;
        ++ctx->solves;
    });
}

lynx_status lynx_lanczos_bounds(lynx_ctx *ctx, int dtype, double *eig_min, double *eig_max, double tol,
                                int max_iter, unsigned seed, int *iters)
{
    if (!ctx || !eig_min || !eig_max) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        check_dtype(dtype);
        need_gpu_stencil(ctx);
        need_veff(ctx);
        if (dtype != LYNX_F64 && dtype != LYNX_C128) fail(LYNX_E_UNSUPPORTED, "Lanczos runs in FP64");
        if (max_iter < 1) fail(LYNX_E_ARG, "max_iter must be positive");
        const long n = ctx->nd;
        const bool cplx = dtype == LYNX_C128;
        const int e = dtype;
        const size_t es = gpu::elem_size(e);
        void *vjm1 = ctx->lz[0].get(n * es);
        void *vj = ctx->lz[1].get(n * es);
        void *vjp1 = ctx->lz[2].get(n * es);
        double *part = (double *)ctx->partials.get(gpu::kPartialBlocks * sizeof(double));
        (void)part;
        // Lanczos.hpp:104-110: glibc rand() start vector, normalised on the
        // host.  The complex entries are Complex(rand(), rand()) there, whose
        // two calls g++ evaluates imaginary part first (Lanczos.hpp:64-66).
        const long nr = cplx ? 2 * n : n;
        std::vector<double> v(nr);
        std::srand(seed);
        if (cplx) {
            for (long i = 0; i < n; ++i) {
                v[2 * i + 1] = (double)std::rand() / RAND_MAX;
                v[2 * i] = (double)std::rand() / RAND_MAX;
            }
        } else {
            for (long i = 0; i < n; ++i) v[i] = (double)std::rand() / RAND_MAX;
        }
        double s = 0;
        if (cplx) for (long i = 0; i < n; ++i) s += v[2 * i] * v[2 * i] + v[2 * i + 1] * v[2 * i + 1];
        else for (long i = 0; i < n; ++i) s += v[i] * v[i];
        const double vs = 1.0 / std::sqrt(s);
        for (double &x : v) x *= vs;
        HIP_OK(hipMemcpyAsync(vjm1, v.data(), nr * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        std::vector<double> a(max_iter + 2, 0.0), b(max_iter + 2, 0.0);
        apply_op(ctx, e, vjm1, n, nullptr, 0, vj, n, 1, 1.0, 0.0, 0.0);
        a[0] = host_sum(ctx, gpu::dot_partials(e, vjm1, vj, n, ctx->partials.as<double>(), ctx->stream));
        int nb = gpu::lanczos_update(e, vj, vjm1, vjp1, a[0], 0.0, n, ctx->partials.as<double>(), ctx->stream);
        b[0] = std::sqrt(host_sum(ctx, nb));
        if (b[0] == 0.0) fail(LYNX_E_UNSUPPORTED, "Lanczos start vector spans an invariant subspace");
        gpu::scale_copy(e, vj, vj, 1.0 / b[0], n, ctx->stream);
        double emin = 0, emax = 0, pmin = 0, pmax = 0, errmin = tol + 1, errmax = tol + 1;
        int j = 0;
        while ((errmin > tol || errmax > tol) && j < max_iter) {
            apply_op(ctx, e, vj, n, nullptr, 0, vjp1, n, 1, 1.0, 0.0, 0.0);
            a[j + 1] = host_sum(ctx, gpu::dot_partials(e, vj, vjp1, n, ctx->partials.as<double>(), ctx->stream));
            nb = gpu::lanczos_update(e, vjp1, vj, vjm1, a[j + 1], b[j], n, ctx->partials.as<double>(), ctx->stream);
            b[j + 1] = std::sqrt(host_sum(ctx, nb));
            if (b[j + 1] == 0.0) break;
            gpu::scale_copy(e, vjp1, vj, 1.0 / b[j + 1], n, ctx->stream);
            const int m = j + 2;
            emin = tridiag_extreme(a.data(), b.data(), m, false);
            emax = tridiag_extreme(a.data(), b.data(), m, true);
            errmin = std::fabs(emin - pmin);
            errmax = std::fabs(emax - pmax);
            pmin = emin;
            pmax = emax;
            ++j;
        }
        HIP_OK(hipStreamSynchronize(ctx->stream));
        *eig_min = emin - 0.1;
        *eig_max = emax * 1.01;
        if (iters) *iters = j;
    });
}

lynx_status lynx_randomize(lynx_ctx *ctx, int dtype, void *psi, int nd_, int ld, int nband, unsigned seed, int mem)
{
    if (!ctx && mem == LYNX_MEM_DEVICE) return LYNX_E_ARG;
    lynx_ctx host_only;
    host_only.device = -1;
    lynx_ctx *c = ctx ? ctx : &host_only;
    return guarded(c, [&] {
        check_dtype(dtype);
        if (dtype == LYNX_BF16) fail(LYNX_E_UNSUPPORTED, "initial wavefunctions are FP64 or FP32 blocks");
        if (nband <= 0 || nd_ <= 0 || !psi) fail(LYNX_E_ARG, "bad block arguments");
        if (mem == LYNX_MEM_DEVICE) need_device(c);
        const long nd = nd_, ldp = ld ? ld : nd;
        if (ldp < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
        const bool cplx = gpu::elem_complex(dtype), single = gpu::elem_single(dtype);
        // glibc's rand() stream after srand(seed) (random() with its default
        // 128-byte state), drawn from a private state so that the caller's
        // own rand() sequence is left alone
        char state[128];
        random_data rd;
        std::memset(&rd, 0, sizeof(rd));
        if (initstate_r(seed, state, sizeof(state), &rd) != 0) fail(LYNX_E_ARG, "initstate_r failed");
        auto draw = [&] {
            int32_t r = 0;
            random_r(&rd, &r);
            return -0.5 + 1.0 * ((double)r / RAND_MAX);
        };
        // Wavefunction.cpp:53-76: column by column; complex entries draw the
        // real part, then the imaginary part
        const long vals = cplx ? 2 * nd : nd;
        const size_t es = gpu::elem_size(dtype);
        // device blocks: staged 32 columns at a time, one 2D copy each
        const int chunk = mem == LYNX_MEM_DEVICE ? 32 : 1;
        std::vector<double> buf((size_t)vals * chunk);
        std::vector<float> buff(single ? (size_t)vals * chunk : 0);
        for (int j0 = 0; j0 < nband; j0 += chunk) {
            const int nc = std::min(chunk, nband - j0);
            for (long i = 0; i < vals * nc; ++i) buf[i] = draw();
            const void *src = buf.data();
            if (single) {
                for (long i = 0; i < vals * nc; ++i) buff[i] = (float)buf[i];
                src = buff.data();
            }
            char *dst = static_cast<char *>(psi) + (size_t)j0 * ldp * es;
            if (mem == LYNX_MEM_DEVICE) {
                HIP_OK(hipMemcpy2DAsync(dst, ldp * es, src, nd * es, nd * es, nc, hipMemcpyHostToDevice, c->stream));
                HIP_OK(hipStreamSynchronize(c->stream));
            } else {
                std::memcpy(dst, src, nd * es);
            }
        }
    });
}

lynx_status lynx_fd_weights(int fd_order, double *w1, double *w2)
{
    try {
        std::vector<double> a, b;
        fd_weights(fd_order, a, b);
        for (size_t i = 0; i < a.size(); ++i) {
            if (w1) w1[i] = a[i];
            if (w2) w2[i] = b[i];
        }
        return LYNX_OK;
    } catch (const Error &e) {
        return e.status;
    }
}

lynx_status lynx_kpoints(int k1, int k2, int k3, const double shift[3], const double latvec[9], int cell,
                         double *kpt_cart, double *kpt_red, double *weights, int capacity, int *nkpt)
{
    try {
        if (!latvec || !nkpt) fail(LYNX_E_ARG, "null argument");
        if (cell != LYNX_CELL_ORTHOGONAL && cell != LYNX_CELL_NONORTHOGONAL) fail(LYNX_E_ARG, "unknown cell type");
        Lattice lat;
        lattice_init(lat, latvec, cell == LYNX_CELL_ORTHOGONAL);
        const double zero[3] = {0, 0, 0};
        KPointSet ks;
        kpoints_generate(k1, k2, k3, shift ? shift : zero, lat, ks);
        const int n = (int)ks.weight.size();
        *nkpt = n;
        if (capacity < n) return (kpt_cart || kpt_red || weights) ? LYNX_E_ARG : LYNX_OK;
        if (kpt_cart) std::copy(ks.cart.begin(), ks.cart.end(), kpt_cart);
        if (kpt_red) std::copy(ks.red.begin(), ks.red.end(), kpt_red);
        if (weights) std::copy(ks.weight.begin(), ks.weight.end(), weights);
        return LYNX_OK;
    } catch (const Error &e) {
        return e.status;
    }
}

lynx_status lynx_psp8_info(const char *path, double *zion, int *lmax, int *lloc, int *nproj, double *rc_max)
{
    if (!path) return LYNX_E_ARG;
    try {
        Psp8 p;
        psp8_load(path, p);
        if (zion) *zion = p.zion;
        if (lmax) *lmax = p.lmax;
        if (lloc) *lloc = p.lloc;
        if (nproj) *nproj = p.nproj_per_atom();
        if (rc_max) *rc_max = p.rc_max();
        return LYNX_OK;
    } catch (const Error &e) {
        return e.status;
    }
}

lynx_status lynx_last_filter_timing(const lynx_ctx *ctx, double times[4], int launches[4])
{
    if (!ctx) return LYNX_E_ARG;
    for (int k = 0; k < 4; ++k) {
        if (times) times[k] = ctx->t_last[k];
        if (launches) launches[k] = ctx->n_last[k];
    }
    return LYNX_OK;
}

lynx_status lynx_set_stencil_config(lynx_ctx *ctx, int threads, int points_div, int z_chunks, int vector_runs)
{
    if (!ctx) return LYNX_E_ARG;
    if (threads == 0 && points_div == 0 && z_chunks == 0) {
        ctx->has_forced = false;
        return LYNX_OK;
    }
    if ((threads != 256 && threads != 512) || (points_div != 1 && points_div != 2) || z_chunks < 1 ||
        vector_runs < 0 || vector_runs > 7 || vector_runs == 4)
        return LYNX_E_ARG;
    ctx->forced.nt = threads;
    ctx->forced.rp_div = points_div;
    ctx->forced.nzch = z_chunks;
    ctx->forced.vec = vector_runs & 3;
    ctx->forced.align = (vector_runs >> 2) & 1;
    ctx->has_forced = true;
    return LYNX_OK;
}

lynx_status lynx_get_stencil_config(const lynx_ctx *ctx, int dtype, int ncol, int *threads, int *points_div,
                                    int *z_chunks, int *vector_runs, double *tune_ms, int *n_tune)
{
    if (!ctx) return LYNX_E_ARG;
    const gpu::StepCfg *k = nullptr;
    std::vector<long> key = {ctx->st.nx, ctx->st.ny, ctx->st.nz, dtype, ncol};
    if (ctx->has_forced) k = &ctx->forced;
    else {
        auto it = ctx->tuned.find(key);
        if (it == ctx->tuned.end()) return LYNX_E_STATE;
        k = &it->second;
    }
    if (threads) *threads = k->nt;
    if (points_div) *points_div = k->rp_div;
    if (z_chunks) *z_chunks = k->nzch;
    if (vector_runs) *vector_runs = k->vec | (k->align << 2);
    if (n_tune) {
        auto it = ctx->tune_log.find(key);
        const int n = it == ctx->tune_log.end() ? 0 : (int)it->second.size();
        if (tune_ms)
            for (int i = 0; i < std::min(n, *n_tune); ++i) tune_ms[i] = it->second[i];
        *n_tune = n;
    }
    return LYNX_OK;
}

lynx_status lynx_get_grid_dims(const lynx_ctx *ctx, int dims[3])
{
    if (!ctx || !dims) return LYNX_E_ARG;
    if (!ctx->has_grid) return LYNX_E_STATE;
    dims[0] = ctx->st.nx;
    dims[1] = ctx->st.ny;
    dims[2] = ctx->st.nz;
    return LYNX_OK;
}

lynx_status lynx_set_filter_precision(lynx_ctx *ctx, int precision)
{
    if (!ctx) return LYNX_E_ARG;
    if (precision != LYNX_FILTER_FP64 && precision != LYNX_FILTER_FP32 && precision != LYNX_FILTER_BF16) {
        ctx->err = "filter precision must be LYNX_FILTER_FP64, LYNX_FILTER_FP32 or LYNX_FILTER_BF16";
        return LYNX_E_ARG;
    }
    ctx->filter_precision = precision;
    return LYNX_OK;
}

lynx_status lynx_set_gram_precision(lynx_ctx *ctx, int mode)
{
    if (!ctx) return LYNX_E_ARG;
    if (mode != LYNX_GRAM_FP32 && mode != LYNX_GRAM_BF16X3 && mode != LYNX_GRAM_BF16X3_ROT) {
        ctx->err = "Gram precision must be LYNX_GRAM_FP32, LYNX_GRAM_BF16X3 or LYNX_GRAM_BF16X3_ROT";
        return LYNX_E_ARG;
    }
    ctx->gram_precision = mode;
    return LYNX_OK;
}

lynx_status lynx_get_gram_precision(const lynx_ctx *ctx, int *mode)
{
    if (!ctx || !mode) return LYNX_E_ARG;
    *mode = ctx->gram_precision;
    return LYNX_OK;
}

lynx_status lynx_gram(lynx_ctx *ctx, int dtype, const void *a, int lda, const void *b, int ldb, int n, void *c,
                      int mem)
{
    if (!ctx || !a || !b || !c || n < 1) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        need_grid(ctx);
        if (dtype != LYNX_F64 && dtype != LYNX_C128 && dtype != LYNX_F32 && dtype != LYNX_C64)
            fail(LYNX_E_ARG, "lynx_gram: dtype must be f64, c128, f32 or c64");
        if (lda < ctx->nd || ldb < ctx->nd) fail(LYNX_E_ARG, "lynx_gram: leading dimension below Nd");
        const int e = dtype;
        const size_t es = gpu::elem_size(e);
        const void *da = a, *db = b;
        if (mem == LYNX_MEM_HOST) {
            // padded device copies (16-byte aligned columns)
            const long ld = padded_ld(ctx->nd);
            void *A = ctx->wA.get((size_t)ld * n * es);
            HIP_OK(hipMemcpy2DAsync(A, ld * es, a, (size_t)lda * es, ctx->nd * es, n, hipMemcpyHostToDevice,
                                    ctx->stream));
            void *B = A;
            if (b != a) {
                B = ctx->wB.get((size_t)ld * n * es);
                HIP_OK(hipMemcpy2DAsync(B, ld * es, b, (size_t)ldb * es, ctx->nd * es, n, hipMemcpyHostToDevice,
                                        ctx->stream));
            }
            da = A;
            db = B;
            lda = ldb = (int)ld;
        } else if (mem != LYNX_MEM_DEVICE) {
            fail(LYNX_E_ARG, "lynx_gram: mem must be LYNX_MEM_HOST or LYNX_MEM_DEVICE");
        }
        const int ew = wide_of(e);
        const size_t ews = gpu::elem_size(ew);
        void *C = ctx->s2.get((size_t)n * n * ews);
        gram(ctx, e, da, lda, db, ldb, n, ctx->st.dV, C);
        HIP_OK(hipMemcpyAsync(c, C, (size_t)n * n * ews, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

lynx_status lynx_rotate(lynx_ctx *ctx, int dtype, const void *a, int lda, const double *q, int n, void *c, int ldc,
                        int mem)
{
    if (!ctx || !a || !q || !c || n < 1) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        need_grid(ctx);
        if (dtype != LYNX_F64 && dtype != LYNX_C128 && dtype != LYNX_F32 && dtype != LYNX_C64)
            fail(LYNX_E_ARG, "lynx_rotate: dtype must be f64, c128, f32 or c64");
        if (lda < ctx->nd || ldc < ctx->nd) fail(LYNX_E_ARG, "lynx_rotate: leading dimension below Nd");
        if (mem != LYNX_MEM_HOST && mem != LYNX_MEM_DEVICE)
            fail(LYNX_E_ARG, "lynx_rotate: mem must be LYNX_MEM_HOST or LYNX_MEM_DEVICE");
        const int e = dtype, ew = wide_of(e);
        const size_t es = gpu::elem_size(e), ews = gpu::elem_size(ew);
        const long nd = ctx->nd;
        const void *A = a;
        void *C = c;
        long la = lda, lc = ldc;
        if (mem == LYNX_MEM_HOST) {
            const long ld = padded_ld(nd);
            void *dA = ctx->wA.get((size_t)ld * n * es);
            HIP_OK(hipMemcpy2DAsync(dA, ld * es, a, (size_t)lda * es, nd * es, n, hipMemcpyHostToDevice,
                                    ctx->stream));
            A = dA;
            C = ctx->wB.get((size_t)ld * n * es);
            la = lc = ld;
        }
        void *Q = ctx->s2.get((size_t)n * n * ews);
        HIP_OK(hipMemcpyAsync(Q, q, (size_t)n * n * ews, hipMemcpyHostToDevice, ctx->stream));
        if (e == gpu::F32 && ctx->gram_precision == LYNX_GRAM_BF16X3_ROT) {
            void *W = ctx->s3.get(gpu::rotate_bf16x3_work(n, n));
            gpu::rotate_bf16x3(static_cast<const float *>(A), la, nd, static_cast<const double *>(Q), n, n, n, C, lc,
                               false, W, ctx->stream);
        } else if (e != ew) {
            void *Qs = ctx->s3.get((size_t)n * n * es);
            gpu::convert_cols(ew, Q, n, e, Qs, n, n, n, ctx->stream);
            rotate(ctx, e, A, la, Qs, n, C, lc);
        } else {
            rotate(ctx, e, A, la, Q, n, C, lc);
        }
        if (mem == LYNX_MEM_HOST)
            HIP_OK(hipMemcpy2DAsync(c, (size_t)ldc * es, C, lc * es, nd * es, n, hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
    });
}

lynx_status lynx_get_filter_precision(const lynx_ctx *ctx, int *precision)
{
    if (!ctx || !precision) return LYNX_E_ARG;
    *precision = ctx->filter_precision;
    return LYNX_OK;
}

lynx_status lynx_solve_count(const lynx_ctx *ctx, long *count)
{
    if (!ctx || !count) return LYNX_E_ARG;
    *count = ctx->solves;
    return LYNX_OK;
}

lynx_status lynx_get_veff_generation(const lynx_ctx *ctx, long *generation)
{
    if (!ctx || !generation) return LYNX_E_ARG;
    *generation = ctx->veff_gen;
    return LYNX_OK;
}

lynx_status lynx_set_timing(lynx_ctx *ctx, int enable)
{
    if (!ctx) return LYNX_E_ARG;
    ctx->timing = enable != 0;
    return LYNX_OK;
}

lynx_status lynx_set_filter_graphs(lynx_ctx *ctx, int enable)
{
    if (!ctx) return LYNX_E_ARG;
    ctx->use_graphs = enable != 0;
    return LYNX_OK;
}

lynx_status lynx_mixed_fallbacks(const lynx_ctx *ctx, long *count, long *cholqr2)
{
    if (!ctx || !count) return LYNX_E_ARG;
    *count = ctx->fp64_overlap_fallbacks;
    if (cholqr2) *cholqr2 = ctx->cholqr2_passes;
    return LYNX_OK;
}

lynx_status lynx_stream_ceiling(lynx_ctx *ctx, long nd, int ncol, double *gbps)
{
    if (!ctx || !gbps) return LYNX_E_ARG;
    return guarded(ctx, [&] {
        if (nd <= 0 || ncol <= 0 || nd % 2) fail(LYNX_E_ARG, "nd must be positive and even, ncol positive");
        const long n16 = nd * ncol / 2, nv16 = nd / 2;
        struct Bufs {
            DevBuf b[4];
            ~Bufs() { for (DevBuf &d : b) d.release(); }
        } bufs;
        DevBuf &x = bufs.b[0], &o = bufs.b[1], &v = bufs.b[2], &y = bufs.b[3];
        x.get((size_t)n16 * 16);
        o.get((size_t)n16 * 16);
        y.get((size_t)n16 * 16);
        v.get((size_t)nv16 * 16);
        HIP_OK(hipMemsetAsync(x.p, 0, (size_t)n16 * 16, ctx->stream));
        HIP_OK(hipMemsetAsync(o.p, 0, (size_t)n16 * 16, ctx->stream));
        HIP_OK(hipMemsetAsync(v.p, 0, (size_t)nv16 * 16, ctx->stream));
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        float best = 1e30f;
        for (int blocks : {1024, 4096, 16384})
            for (int rep = 0; rep < 6; ++rep) {
                HIP_OK(hipEventRecord(e0, ctx->stream));
                gpu::stream_probe(x.p, o.p, v.p, y.p, n16, nv16, blocks, ctx->stream);
                HIP_OK(hipEventRecord(e1, ctx->stream));
                HIP_OK(hipEventSynchronize(e1));
                float ms = 0;
                HIP_OK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) best = std::min(best, ms);
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIP_OK(hipStreamSynchronize(ctx->stream));
        *gbps = (3.0 * 8.0 * nd * ncol + 8.0 * nd) / (best * 1e-3) / 1e9;
    });
}

}  // extern "C"

// ----------------------------------------------------------------- batches --
// Many independent filter units (k-points of one structure, several
// structures with their own grids) filtered together: every Chebyshev step
// is one launch per kernel for all units (gather, stencil, scatter per
// group of like units), and the whole filter is captured once into a HIP
// graph and replayed.
struct lynx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    struct Unit {
        lynx_ctx *ctx;
        double kpt[3];
        DevBuf ph64, ph32;     // image phases e^{-i k.R} of this unit's k-point
        DevBuf work, dx, dy;   // recurrence work block; staging for host buffers
        DevBuf sy;             // lynx_batch_solve: the filtered block, then H psi, then the rotated block
        DevBuf sx32, sy32;     // lynx_batch_solve with the FP32 filter: single-precision copies
        long epoch = -1;       // the context's set-up epoch the phases were built for
    };
    std::vector<Unit> units;
    // plan and graph of the last call; rebuilt when buffers, shapes or dtype change
    std::vector<long> key;
    gpu::BatchPlan *plan = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_degree = 0;
    bool use_graph = true;
    int filter_precision = LYNX_FILTER_FP64;   // of lynx_batch_solve's filter
    // lynx_batch_solve: the H psi step's plan (its own coefficient row, so
    // the filter graph's table is never resized), a BLAS handle on the
    // batch's stream and per-group dense work
    std::vector<long> hkey;
    gpu::BatchPlan *hplan = nullptr;
    rocblas_handle blas = nullptr;
    DevBuf dmat, dhs, dev, dwork, dinfo, dptrs;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0;
    int last_launches = 0;
    ~lynx_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        gpu::batch_plan_destroy(plan);
        gpu::batch_plan_destroy(hplan);
        if (blas) (void)rocblas_destroy_handle(blas);
        for (auto &u : units) {
            u.ph64.release(); u.ph32.release(); u.work.release(); u.dx.release(); u.dy.release(); u.sy.release();
            u.sx32.release(); u.sy32.release();
        }
        for (DevBuf *x : {&dmat, &dhs, &dev, &dwork, &dinfo, &dptrs}) x->release();
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
// Bloch factors e^{-i k.R_img} of the unit's images at its own k-point,
// from the context's current projector set-up (rebuilt when a set-up call
// on the context has changed its images).
void unit_phases(lynx_batch::Unit &U, hipStream_t st)
{
    const NonlocalHost &h = U.ctx->nlh;
    std::vector<double> ph(2 * std::max(1, h.n_img));
    for (int i = 0; i < h.n_img; ++i) {
        const double *sh = &h.img_shift[3 * i];
        const double th = -(U.kpt[0] * sh[0] + U.kpt[1] * sh[1] + U.kpt[2] * sh[2]);
        ph[2 * i] = std::cos(th);
        ph[2 * i + 1] = std::sin(th);
    }
    upload(U.ph64, ph, st);
    upload(U.ph32, std::vector<float>(ph.begin(), ph.end()), st);
    U.epoch = U.ctx->epoch;
}

template <class F>
lynx_status bguarded(lynx_batch *b, F &&f)
{
    try {
        HIP_OK(hipSetDevice(b->device));
        f();
        return LYNX_OK;
    } catch (const Error &e) {
        b->err = e.msg;
        return e.status;
    } catch (const std::exception &e) {
        b->err = e.what();
        return LYNX_E_HIP;
    }
}
}  // namespace

lynx_status lynx_batch_create(lynx_batch **out, int nunit, lynx_ctx *const *ctxs, const double *kpts)
{
    if (!out || nunit <= 0 || !ctxs) return LYNX_E_ARG;
    *out = nullptr;
    auto *b = new lynx_batch();
    b->device = ctxs[0] ? ctxs[0]->device : -1;
    const lynx_status st = bguarded(b, [&] {
        for (int u = 0; u < nunit; ++u) {
            lynx_ctx *c = ctxs[u];
            if (!c) fail(LYNX_E_ARG, "null context in batch");
            need_gpu_stencil(c);
            need_veff(c);
            if (c->device != b->device) fail(LYNX_E_ARG, "batch contexts must share one device");
        }
        HIP_OK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&b->e0));
        HIP_OK(hipEventCreate(&b->e1));
        b->units.resize(nunit);
        for (int u = 0; u < nunit; ++u) {
            lynx_batch::Unit &U = b->units[u];
            U.ctx = ctxs[u];
            for (int a = 0; a < 3; ++a) U.kpt[a] = kpts ? kpts[3 * u + a] : 0.0;
            unit_phases(U, b->stream);
        }
        HIP_OK(hipStreamSynchronize(b->stream));
    });
    if (st != LYNX_OK) {
        delete b;
        return st;
    }
    *out = b;
    return LYNX_OK;
}

lynx_status lynx_batch_destroy(lynx_batch *b)
{
    if (!b) return LYNX_E_ARG;
    if (b->device >= 0) (void)hipSetDevice(b->device);
    delete b;
    return LYNX_OK;
}

const char *lynx_batch_last_error(const lynx_batch *b) { return b ? b->err.c_str() : "null batch"; }

lynx_status lynx_batch_set_filter_precision(lynx_batch *b, int precision)
{
    if (!b) return LYNX_E_ARG;
    if (precision != LYNX_FILTER_FP64 && precision != LYNX_FILTER_FP32) {
        b->err = "lynx_batch_set_filter_precision: LYNX_FILTER_FP64 or LYNX_FILTER_FP32";
        return LYNX_E_ARG;
    }
    b->filter_precision = precision;
    return LYNX_OK;
}

lynx_status lynx_batch_set_graph(lynx_batch *b, int enable)
{
    if (!b) return LYNX_E_ARG;
    b->use_graph = enable != 0;
    return LYNX_OK;
}

lynx_status lynx_batch_last_timing(const lynx_batch *b, double *ms, int *launches)
{
    if (!b) return LYNX_E_ARG;
    if (ms) *ms = b->last_ms;
    if (launches) *launches = b->last_launches;
    return LYNX_OK;
}

lynx_status lynx_batch_filter(lynx_batch *b, int dtype, void *const *x, const long *ldx, void *const *y,
                              const long *ldy, const int *ncol, int degree, const double *lambda_c,
                              const double *eig_min, const double *eig_max, int mem)
{
    if (!b) return LYNX_E_ARG;
    return bguarded(b, [&] {
        check_dtype(dtype);
        if (dtype == LYNX_BF16) fail(LYNX_E_UNSUPPORTED, "batches filter in FP64 or FP32");
        const int n = (int)b->units.size();
        if (!x || !y || !ncol || !lambda_c || !eig_min || !eig_max) fail(LYNX_E_ARG, "null batch argument");
        if (degree < 1) fail(LYNX_E_ARG, "degree must be >= 1");
        const size_t es = gpu::elem_size(dtype);
        const bool single = gpu::elem_single(dtype);
        const bool cplx = gpu::elem_complex(dtype);
        // the contexts' own set-up work (Veff, projectors) must be complete;
        // a context whose grid or projectors were set up again since the
        // last call gets new phase tables, and its new set-up epoch and Veff
        // buffers enter the key, so the plan and graph are rebuilt
        for (auto &U : b->units) {
            HIP_OK(hipStreamSynchronize(U.ctx->stream));
            need_gpu_stencil(U.ctx);
            need_veff(U.ctx);
            if (U.ctx->epoch != U.epoch) unit_phases(U, b->stream);
        }
        HIP_OK(hipStreamSynchronize(b->stream));
        std::vector<gpu::BatchUnit> bu(n);
        std::vector<long> key = {dtype, degree % 2, mem};
        for (int u = 0; u < n; ++u) {
            lynx_batch::Unit &U = b->units[u];
            lynx_ctx *c = U.ctx;
            const long nd = c->nd;
            if (ncol[u] <= 0 || !x[u] || !y[u]) fail(LYNX_E_ARG, "bad block of batch unit " + std::to_string(u));
            const long lx = (ldx && ldx[u]) ? ldx[u] : nd, ly = (ldy && ldy[u]) ? ldy[u] : nd;
            if (lx < nd || ly < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
            const long ldw = padded_ld(nd);
            void *w = U.work.get((size_t)ldw * ncol[u] * es);
            const void *X = x[u];
            void *Y = y[u];
            long lX = lx, lY = ly;
            if (mem != LYNX_MEM_DEVICE) {
                void *dx = U.dx.get((size_t)nd * ncol[u] * es);
                Y = U.dy.get((size_t)nd * ncol[u] * es);
                HIP_OK(hipMemcpy2DAsync(dx, nd * es, x[u], lx * es, nd * es, ncol[u], hipMemcpyHostToDevice,
                                        b->stream));
                X = dx;
                lX = lY = nd;
            }
            gpu::BatchUnit &B = bu[u];
            const double k0[3] = {0, 0, 0};
            B.g = geom_at(c, cplx ? U.kpt : k0);
            B.veff = single ? c->veff32.p : c->veff64.p;
            if (c->has_nl) {
                B.nl = single ? c->nl32 : c->nl64;
                B.nl.phase = single ? U.ph32.p : U.ph64.p;
            }
            // odd steps write D1, even steps D2; the last step lands in Y
            void *D1 = (degree % 2 == 1) ? Y : w;
            void *D2 = (degree % 2 == 1) ? w : Y;
            B.buf[0] = X; B.buf[1] = D1; B.buf[2] = D2;
            B.ld[0] = lX;
            B.ld[1] = (degree % 2 == 1) ? lY : ldw;
            B.ld[2] = (degree % 2 == 1) ? ldw : lY;
            B.ncol = ncol[u];
            key.push_back(reinterpret_cast<long>(X));
            key.push_back(reinterpret_cast<long>(Y));
            key.push_back(lX);
            key.push_back(lY);
            key.push_back(ncol[u]);
            key.push_back(c->epoch);
            key.push_back(reinterpret_cast<long>(c->veff64.p));
            key.push_back(reinterpret_cast<long>(c->veff32.p));
            key.push_back(reinterpret_cast<long>(U.ph64.p));
            key.push_back(reinterpret_cast<long>(U.ph32.p));
        }
        if (key != b->key || !b->plan) {
            if (b->graph) (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
            gpu::batch_plan_destroy(b->plan);
            b->plan = nullptr;
            b->plan = gpu::batch_plan_create(dtype, bu.data(), n, b->stream);
            b->key = key;
        }
        // filter scalars of every step and unit (EigenSolver.cpp:558-619)
        std::vector<double> coef((size_t)degree * n * 3);
        for (int u = 0; u < n; ++u) {
            const double ehalf = (eig_max[u] - lambda_c[u]) / 2.0;
            const double cen = (eig_max[u] + lambda_c[u]) / 2.0;
            if (!(ehalf > 0)) fail(LYNX_E_ARG, "eig_max must exceed lambda_cutoff (unit " + std::to_string(u) + ")");
            const double sigma1 = ehalf / (eig_min[u] - cen);
            double sigma = sigma1;
            coef[u * 3 + 0] = sigma / ehalf;
            coef[u * 3 + 1] = 0.0;
            coef[u * 3 + 2] = -cen;
            for (int k = 2; k <= degree; ++k) {
                const double sn = 1.0 / (2.0 / sigma1 - sigma);
                double *cf = &coef[((size_t)(k - 1) * n + u) * 3];
                cf[0] = 2.0 * sn / ehalf;
                cf[1] = -sigma * sn;
                cf[2] = -cen;
                sigma = sn;
            }
        }
        gpu::batch_set_coef(b->plan, coef.data(), degree, b->stream);
        auto record = [&] {
            gpu::batch_step(b->plan, 0, 0, 3, 1, b->stream);
            int prev = 1;
            for (int k = 2; k <= degree; ++k) {
                const int dst = (k % 2 == 1) ? 1 : 2;
                const int old = (k == 2) ? 0 : dst;
                gpu::batch_step(b->plan, k - 1, prev, old, dst, b->stream);
                prev = dst;
            }
        };
        HIP_OK(hipEventRecord(b->e0, b->stream));
        if (b->use_graph) {
            if (!b->graph || b->graph_degree != degree) {
                if (b->graph) (void)hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
                hipGraph_t g = nullptr;
                HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                record();
                HIP_OK(hipStreamEndCapture(b->stream, &g));
                const hipError_t r = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIP_OK(r);
                b->graph_degree = degree;
            }
            HIP_OK(hipGraphLaunch(b->graph, b->stream));
        } else {
            record();
        }
        HIP_OK(hipEventRecord(b->e1, b->stream));
        if (mem != LYNX_MEM_DEVICE) {
            for (int u = 0; u < n; ++u) {
                lynx_batch::Unit &U = b->units[u];
                const long nd = U.ctx->nd;
                const long ly = (ldy && ldy[u]) ? ldy[u] : nd;
                HIP_OK(hipMemcpy2DAsync(y[u], ly * es, U.dy.p, nd * es, nd * es, ncol[u], hipMemcpyDeviceToHost,
                                        b->stream));
            }
        }
        HIP_OK(hipStreamSynchronize(b->stream));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, b->e0, b->e1));
        b->last_ms = ms;
        b->last_launches = degree * gpu::batch_launches_per_step(b->plan);
    });
}

namespace {

// The operator part of batch unit u (grid, Bloch phases, Veff, projectors)
// with three block buffers.
gpu::BatchUnit batch_unit(lynx_batch::Unit &U, int dtype, const void *b0, long l0, const void *b1, long l1,
                          const void *b2, long l2, int ncol)
{
    lynx_ctx *c = U.ctx;
    const bool single = gpu::elem_single(dtype), cplx = gpu::elem_complex(dtype);
    gpu::BatchUnit B;
    const double k0[3] = {0, 0, 0};
    B.g = geom_at(c, cplx ? U.kpt : k0);
    B.veff = single ? c->veff32.p : c->veff64.p;
    if (c->has_nl) {
        B.nl = single ? c->nl32 : c->nl64;
        B.nl.phase = single ? U.ph32.p : U.ph64.p;
    }
    B.buf[0] = b0; B.buf[1] = b1; B.buf[2] = b2;
    B.ld[0] = l0; B.ld[1] = l1; B.ld[2] = l2;
    B.ncol = ncol;
    return B;
}

void check_infos(lynx_batch *b, const rocblas_int *dinfo, int m, const std::vector<int> &units, int status,
                 const char *what)
{
    std::vector<rocblas_int> info(m);
    HIP_OK(hipMemcpyAsync(info.data(), dinfo, m * sizeof(rocblas_int), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    for (int i = 0; i < m; ++i)
        if (info[i] != 0)
            fail(status, std::string(what) + " failed for batch unit " + std::to_string(units[i]) + " (info=" +
                             std::to_string(info[i]) + ")");
}

}  // namespace

lynx_status lynx_batch_solve(lynx_batch *b, int dtype, void *const *psi, const long *ldpsi, const int *ncol,
                             double *eigvals, const double *lambda_c, const double *eig_min, const double *eig_max,
                             int degree, double *lambda_c_out)
{
    if (!b) return LYNX_E_ARG;
    return bguarded(b, [&] {
        if (dtype != LYNX_F64 && dtype != LYNX_C128) fail(LYNX_E_ARG, "lynx_batch_solve runs in FP64 or complex128");
        if (!psi || !ncol || !eigvals || !lambda_c || !eig_min || !eig_max) fail(LYNX_E_ARG, "null batch argument");
        const int n = (int)b->units.size();
        const bool cplx = dtype == LYNX_C128;
        const size_t es = gpu::elem_size(dtype);
        std::vector<void *> ys(n);
        std::vector<long> lys(n), lps(n), eoff(n + 1, 0);
        for (int u = 0; u < n; ++u) {
            lynx_batch::Unit &U = b->units[u];
            const long nd = U.ctx->nd;
            if (ncol[u] <= 0 || !psi[u]) fail(LYNX_E_ARG, "bad block of batch unit " + std::to_string(u));
            lps[u] = (ldpsi && ldpsi[u]) ? ldpsi[u] : nd;
            if (lps[u] < nd) fail(LYNX_E_ARG, "leading dimension smaller than Nd");
            if (ncol[u] > nd) fail(LYNX_E_ARG, "more orbitals than grid points in batch unit " + std::to_string(u));
            lys[u] = padded_ld(nd);
            ys[u] = U.sy.get((size_t)lys[u] * ncol[u] * es);
            eoff[u + 1] = eoff[u] + ncol[u];
        }
        // 1. Y = p(H) psi for every unit (EigenSolver.cpp:558-619); with the
        // FP32 filter on single-precision copies, widened back into Y
        if (b->filter_precision == LYNX_FILTER_FP32) {
            const int e1 = cplx ? LYNX_C64 : LYNX_F32;
            const size_t es1 = gpu::elem_size(e1);
            std::vector<void *> xs(n), y1(n);
            for (int u = 0; u < n; ++u) {
                lynx_batch::Unit &U = b->units[u];
                HIP_OK(hipStreamSynchronize(U.ctx->stream));   // psi as the context's stream left it
                xs[u] = U.sx32.get((size_t)lys[u] * ncol[u] * es1);
                y1[u] = U.sy32.get((size_t)lys[u] * ncol[u] * es1);
                gpu::convert_cols(dtype, psi[u], lps[u], e1, xs[u], lys[u], U.ctx->nd, ncol[u], b->stream);
            }
            const lynx_status fs = lynx_batch_filter(b, e1, xs.data(), lys.data(), y1.data(), lys.data(), ncol, degree,
                                                     lambda_c, eig_min, eig_max, LYNX_MEM_DEVICE);
            if (fs != LYNX_OK) fail(fs, "batch filter: " + b->err);
            for (int u = 0; u < n; ++u)
                gpu::convert_cols(e1, y1[u], lys[u], dtype, ys[u], lys[u], b->units[u].ctx->nd, ncol[u], b->stream);
        } else {
            const lynx_status fs = lynx_batch_filter(b, dtype, psi, lps.data(), ys.data(), lys.data(), ncol, degree,
                                                     lambda_c, eig_min, eig_max, LYNX_MEM_DEVICE);
            if (fs != LYNX_OK) fail(fs, "batch filter: " + b->err);
        }
        if (!b->blas) {
            BLAS_OK(rocblas_create_handle(&b->blas));
            BLAS_OK(rocblas_set_stream(b->blas, b->stream));
        }
        hipStream_t st = b->stream;
        // units of equal (Nd, ncol, ldpsi, dV) form one batched group
        std::map<std::vector<long>, std::vector<int>> groups;
        for (int u = 0; u < n; ++u) {
            const lynx_ctx *c = b->units[u].ctx;
            long dvbits;
            std::memcpy(&dvbits, &c->st.dV, sizeof(dvbits));
            groups[{(long)c->nd, ncol[u], lps[u], dvbits}].push_back(u);
        }
        size_t mat = 0, ev = 0;
        for (auto &g : groups) {
            const size_t N = (size_t)g.first[1];
            mat += g.second.size() * N * N;
            ev += g.second.size() * N;
        }
        char *M = static_cast<char *>(b->dmat.get(mat * es));      // S, then R^-1
        char *Hm = static_cast<char *>(b->dhs.get(mat * es));      // Hs, then Q
        b->dev.get(ev * sizeof(double));
        b->dwork.get(ev * sizeof(double));
        b->dinfo.get((size_t)n * sizeof(rocblas_int));
        double *D = b->dev.as<double>(), *E = b->dwork.as<double>();
        rocblas_int *info = b->dinfo.as<rocblas_int>();
        // device pointer tables, four per unit: Y, psi, the unit's N x N matrix
        // in M, the one in Hm (group by group)
        std::vector<void *> hp(4 * (size_t)n);
        struct Grp {
            std::vector<int> units;
            long nd, ld, ldy;
            int N;
            double dV;
            size_t moff, eoff, poff;
        };
        std::vector<Grp> G;
        {
            size_t moff = 0, eo = 0, po = 0;
            for (auto &g : groups) {
                Grp q;
                q.units = g.second;
                q.nd = g.first[0];
                q.N = (int)g.first[1];
                q.ld = g.first[2];
                q.ldy = padded_ld(q.nd);
                q.dV = b->units[q.units[0]].ctx->st.dV;
                q.moff = moff; q.eoff = eo; q.poff = po;
                const int m = (int)q.units.size();
                for (int i = 0; i < m; ++i) {
                    const int u = q.units[i];
                    hp[po + i] = ys[u];
                    hp[po + m + i] = psi[u];
                    hp[po + 2 * m + i] = M + (moff + (size_t)i * q.N * q.N) * es;
                    hp[po + 3 * m + i] = Hm + (moff + (size_t)i * q.N * q.N) * es;
                }
                moff += (size_t)m * q.N * q.N;
                eo += (size_t)m * q.N;
                po += 4 * (size_t)m;
                G.push_back(q);
            }
        }
        void **dp = static_cast<void **>(b->dptrs.get(hp.size() * sizeof(void *)));
        HIP_OK(hipMemcpyAsync(dp, hp.data(), hp.size() * sizeof(void *), hipMemcpyHostToDevice, st));
        using Z = rocblas_double_complex;
        auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int mm, int nn, int kk, double alpha, void **A,
                        long lda, void **B, long ldb, void **C, long ldc, int cnt) {
            if (cplx) {
                const Z a(alpha, 0.0), z(0.0, 0.0);
                BLAS_OK(rocblas_zgemm_batched(b->blas, ta, tb, mm, nn, kk, &a, (const Z *const *)A, (int)lda,
                                              (const Z *const *)B, (int)ldb, &z, (Z *const *)C, (int)ldc, cnt));
            } else {
                const double z = 0.0;
                BLAS_OK(rocblas_dgemm_batched(b->blas, ta == rocblas_operation_conjugate_transpose
                                                           ? rocblas_operation_transpose
                                                           : ta,
                                              tb, mm, nn, kk, &alpha, (const double *const *)A, (int)lda,
                                              (const double *const *)B, (int)ldb, &z, (double *const *)C, (int)ldc,
                                              cnt));
            }
        };
        const rocblas_operation OC = rocblas_operation_conjugate_transpose, ON = rocblas_operation_none;
        // 2. Cholesky-QR per group: S = dV Y^H Y, R^H R = S, psi = Y R^-1 (EigenSolver.cpp:621-641)
        for (const Grp &q : G) {
            const int m = (int)q.units.size(), N = q.N;
            void **Yp = dp + q.poff, **Pp = dp + q.poff + m, **Mp = dp + q.poff + 2 * m;
            gemm(OC, ON, N, N, (int)q.nd, q.dV, Yp, q.ldy, Yp, q.ldy, Mp, N, m);
            void *Mg = M + q.moff * es;
            const rocblas_stride sN = (rocblas_stride)N * N;
            if (cplx)
                BLAS_OK(rocsolver_zpotrf_strided_batched(b->blas, rocblas_fill_upper, N, (Z *)Mg, N, sN, info, m));
            else
                BLAS_OK(rocsolver_dpotrf_strided_batched(b->blas, rocblas_fill_upper, N, (double *)Mg, N, sN, info,
                                                         m));
            check_infos(b, info, m, q.units, LYNX_E_CHOLESKY, "Cholesky factorisation of the overlap matrix");
            if (cplx)
                BLAS_OK(rocsolver_ztrtri_strided_batched(b->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, N,
                                                         (Z *)Mg, N, sN, info, m));
            else
                BLAS_OK(rocsolver_dtrtri_strided_batched(b->blas, rocblas_fill_upper, rocblas_diagonal_non_unit, N,
                                                         (double *)Mg, N, sN, info, m));
            check_infos(b, info, m, q.units, LYNX_E_CHOLESKY, "triangular inverse");
            for (int i = 0; i < m; ++i) gpu::zero_lower(dtype, M + (q.moff + (size_t)i * N * N) * es, N, st);
            gemm(ON, ON, (int)q.nd, N, N, 1.0, Yp, q.ldy, Mp, N, Pp, q.ld, m);
        }
        // 3. H psi into Y for every unit: one batched step (a = 1, no old
        // term, no shift) from its own plan
        {
            std::vector<gpu::BatchUnit> bu(n);
            std::vector<long> key = {dtype};
            for (int u = 0; u < n; ++u) {
                lynx_batch::Unit &U = b->units[u];
                bu[u] = batch_unit(U, dtype, psi[u], lps[u], ys[u], lys[u], ys[u], lys[u], ncol[u]);
                for (long v : {(long)(uintptr_t)psi[u], lps[u], (long)(uintptr_t)ys[u], lys[u], (long)ncol[u],
                               U.ctx->epoch, (long)(uintptr_t)U.ctx->veff64.p, (long)(uintptr_t)U.ph64.p})
                    key.push_back(v);
            }
            if (key != b->hkey || !b->hplan) {
                gpu::batch_plan_destroy(b->hplan);
                b->hplan = nullptr;
                b->hplan = gpu::batch_plan_create(dtype, bu.data(), n, st);
                b->hkey = key;
            }
            std::vector<double> coef((size_t)n * 3, 0.0);
            for (int u = 0; u < n; ++u) coef[3 * u] = 1.0;
            gpu::batch_set_coef(b->hplan, coef.data(), 1, st);
            gpu::batch_step(b->hplan, 0, 0, 3, 1, st);
        }
        // 4. Rayleigh-Ritz per group: Hs = dV psi^H (H psi), Hermitian part,
        // Hs = Q Lambda Q^H, psi = psi Q (EigenSolver.cpp:643-681)
        std::vector<double> hev(ev);
        for (const Grp &q : G) {
            const int m = (int)q.units.size(), N = q.N;
            void **Yp = dp + q.poff, **Pp = dp + q.poff + m, **Hp = dp + q.poff + 3 * m;
            gemm(OC, ON, N, N, (int)q.nd, q.dV, Pp, q.ld, Yp, q.ldy, Hp, N, m);
            for (int i = 0; i < m; ++i) gpu::hermitian_part(dtype, Hm + (q.moff + (size_t)i * N * N) * es, N, st);
            void *Hg = Hm + q.moff * es;
            const rocblas_stride sN = (rocblas_stride)N * N;
            if (cplx)
                BLAS_OK(rocsolver_zheevd_strided_batched(b->blas, rocblas_evect_original, rocblas_fill_upper, N,
                                                         (Z *)Hg, N, sN, D + q.eoff, N, E + q.eoff, N, info, m));
            else
                BLAS_OK(rocsolver_dsyevd_strided_batched(b->blas, rocblas_evect_original, rocblas_fill_upper, N,
                                                         (double *)Hg, N, sN, D + q.eoff, N, E + q.eoff, N, info, m));
            check_infos(b, info, m, q.units, LYNX_E_EIG, "subspace eigensolve");
            gemm(ON, ON, (int)q.nd, N, N, 1.0, Pp, q.ld, Hp, N, Yp, q.ldy, m);
            for (int i = 0; i < m; ++i) {
                const int u = q.units[i];
                gpu::copy_cols(dtype, ys[u], lys[u], psi[u], lps[u], q.nd, N, st);
            }
        }
        HIP_OK(hipMemcpyAsync(hev.data(), D, ev * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (const Grp &q : G)
            for (size_t i = 0; i < q.units.size(); ++i) {
                const int u = q.units[i];
                const double *src = hev.data() + q.eoff + i * q.N;
                for (int j = 0; j < q.N; ++j) {
                    if (!std::isfinite(src[j])) fail(LYNX_E_NONFINITE, "non-finite Ritz value in batch unit " +
                                                                           std::to_string(u));
                    eigvals[eoff[u] + j] = src[j];
                }
                if (lambda_c_out) lambda_c_out[u] = src[q.N - 1] + 0.1;
            }
        for (int u = 0; u < n; ++u) ++b->units[u].ctx->solves;
    });
}
