// Internal state of a lynx_ctx and helpers shared by the translation units
// of liblynx_chefsi.so (the C ABI in capi.cpp, the SCF driver in scf.cpp).
// Not part of the public interface.
#pragma once

#include "lynx_chefsi.h"

#include "common.hpp"
#include "electro.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

namespace lynx {
namespace detail {


#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            fail(LYNX_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

#define BLAS_OK(expr)                                                                        \
    do {                                                                                     \
        rocblas_status s_ = (expr);                                                          \
        if (s_ != rocblas_status_success)                                                    \
            fail(LYNX_E_HIP, std::string(#expr) + ": " + rocblas_status_to_string(s_));      \
    } while (0)

// Grow-only device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    void *get(size_t bytes)
    {
        if (bytes > n) {
            if (p) HIP_OK(hipFree(p));
            p = nullptr;
            n = 0;
            HIP_OK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
            n = bytes;
        }
        return p;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

template <class T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t st)
{
    b.get(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) HIP_OK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

}  // namespace detail
}  // namespace lynx

using lynx::detail::DevBuf;
using lynx::ElecHost;
using lynx::Lattice;
using lynx::NonlocalHost;
using lynx::Psp8;
using lynx::Stencil;
namespace gpu = lynx::gpu;

struct lynx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    std::string err;

    bool has_grid = false;
    Lattice lat{};
    Stencil st{};
    int nd = 0;

    bool has_veff = false;
    long veff_gen = 0;                             // bumped by every lynx_set_veff call
    DevBuf veff64, veff32;
    double kpt[3] = {0, 0, 0};

    bool has_nl = false;
    NonlocalHost nlh;
    // atoms of lynx_set_atoms (psp8 data, types, lattice-aligned positions)
    // and their local-pseudopotential set-up, built on first use
    bool has_atoms = false;
    std::vector<Psp8> psps;
    std::vector<int> atom_types;
    std::vector<double> atom_pos;
    bool has_elec = false;
    ElecHost elec;
    gpu::NlDev nl64, nl32;                         // views for double / float kernels
    DevBuf d_atom_pt_off, d_pt_gpos, d_pt_img, d_pt_atom, d_cov_off, d_cov_gpos, d_ent_pt;
    DevBuf d_chi64, d_chi32, d_gam64, d_gam32, d_ph64, d_ph32;
    DevBuf beta;

    DevBuf wA, wB, wC, lz[3];                      // Nd-sized work blocks
    DevBuf s1, s2, s3, gwork, dvals, evals, info, partials;
    std::vector<double> hpart;

    // stencil launch shape: forced by lynx_set_stencil_config, else tuned
    // per (grid, dtype, ncol) on first use and cached
    gpu::StepCfg forced{};
    bool has_forced = false;
    bool autotune = true;
    std::map<std::vector<long>, gpu::StepCfg> tuned;
    std::map<std::vector<long>, std::vector<double>> tune_log;

    // precision of the filter inside lynx_chefsi_solve (LYNX_FILTER_*)
    int filter_precision = 0;
    int gram_precision = 0;            // LYNX_GRAM_* for single-precision real Gram products
    long solves = 0;                   // completed lynx_chefsi_solve calls
    long fp64_overlap_fallbacks = 0;   // FP32 passes redone in FP64
    long cholqr2_passes = 0;           // FP32 passes with a second Cholesky-QR

    // filter graphs: the step scalars come from a device table uploaded per
    // call, so one captured graph serves every call with the same buffers,
    // column count and set-up (key); epoch is bumped by each set-up call
    // whose values the launches capture
    struct FilterGraph {
        std::vector<long> key;
        hipGraphExec_t exec = nullptr;
        std::vector<std::pair<int, int>> marks;   // timing events recorded by the graph's nodes
    };
    std::vector<FilterGraph> graphs;
    std::vector<std::vector<long>> graph_seen;
    DevBuf fcoef;                      // kMaxGraphDegree steps of (a, b, shift)
    void *fcoef_host = nullptr;        // pinned staging of the table
    hipEvent_t fcoef_copied = nullptr; // the last upload from the staging buffer
    long epoch = 0;
    bool use_graphs = true;

    void *scratch = nullptr;           // this device's scratch array (lynx_h_apply_noalloc)

    bool timing = false;
    double t_last[4] = {0, 0, 0, 0};
    int n_last[4] = {0, 0, 0, 0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_marks;     // (class, start index)

    ~lynx_ctx()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        DevBuf *bufs[] = {&veff64, &veff32, &d_atom_pt_off, &d_pt_gpos, &d_pt_img, &d_pt_atom, &d_cov_off,
                          &d_cov_gpos, &d_ent_pt, &d_chi64, &d_chi32, &d_gam64, &d_gam32, &d_ph64, &d_ph32,
                          &beta, &wA, &wB, &wC, &lz[0], &lz[1], &lz[2], &s1, &s2, &s3, &gwork, &dvals, &evals,
                          &info, &partials};
        for (auto &g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        fcoef.release();
        if (fcoef_copied) (void)hipEventDestroy(fcoef_copied);
        if (fcoef_host) (void)hipHostFree(fcoef_host);
        for (DevBuf *b : bufs) b->release();
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
        if (blas) rocblas_destroy_handle(blas);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace lynx {
namespace detail {

template <class F>
lynx_status guarded(lynx_ctx *ctx, F &&f)
{
    try {
        if (ctx && ctx->device >= 0) HIP_OK(hipSetDevice(ctx->device));
        f();
        return LYNX_OK;
    } catch (const Error &e) {
        if (ctx) ctx->err = e.msg;
        return e.status;
    } catch (const std::exception &e) {
        if (ctx) ctx->err = e.what();
        return LYNX_E_HIP;
    }
}

void need_grid(const lynx_ctx *c);
void need_device(const lynx_ctx *c);
void need_gpu_stencil(const lynx_ctx *c);
void need_veff(const lynx_ctx *c);
// stencil description of the context's grid with the Bloch phases of kpt
gpu::StencilGeom geom_at(const lynx_ctx *c, const double *kpt);
// launch shape of the fused stencil for (grid, dtype, ncol), tuned on first use
const gpu::StepCfg &step_config(lynx_ctx *c, int e, const void *x, long ldx, void *y, long ldy, int ncol);
// sum of nb partials in c->partials, on the host (waits for the stream)
double host_sum(lynx_ctx *c, int nb);
// builds the local-pseudopotential set-up on first use (lynx_set_atoms contexts)
const ElecHost &electrostatics(lynx_ctx *c);
long padded_ld(long nd);

}  // namespace detail
}  // namespace lynx
