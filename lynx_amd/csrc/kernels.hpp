// Launch interface of the HIP kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace lynx {
namespace gpu {

// Element types of wavefunction blocks; values match enum lynx_dtype.
// B16: real bfloat16 storage with float arithmetic (filter steps only;
// its Veff, projector data and beta are float)
enum Elem { F64 = 0, C128 = 1, F32 = 2, C64 = 3, B16 = 4 };
inline size_t elem_size(int e) { return e == F64 ? 8 : e == C128 ? 16 : e == F32 ? 4 : e == B16 ? 2 : 8; }
inline bool elem_complex(int e) { return e == C128 || e == C64; }
inline bool elem_single(int e) { return e == F32 || e == C64 || e == B16; }
// element type of beta and the other arithmetic-side buffers of e
inline int elem_compute(int e) { return e == B16 ? F32 : e; }

constexpr int kMaxR = 6;   // stencil radius handled by the fused kernels (FD order 12)

// Grid, stencil and Bloch wrap for the fused stencil kernel.
struct StencilGeom {
    int nx, ny, nz, R;
    double c[3][kMaxR + 1];          // D2 coefficients per axis
    double phase[3][2][2];           // [axis][0: e^{-i th}, 1: e^{+i th}][re, im]
    // non-orthogonal cells: D2_a[k] = beta_a w2[k], mixed coefficient
    // c_ab[k] = alpha_ab w1[k] (ab = xy, xz, yz), D1_a[k] = w1[k] / h_a
    bool orth = true;
    double w1[kMaxR + 1], w2[kMaxR + 1];
    double beta[3], alpha[3], h[3];
};

// Launch shape of the fused stencil step; chosen per grid shape by the
// tuner in the C ABI layer.
struct StepCfg {
    int rp_div = 1;    // points per thread = base / rp_div
    int nzch = 1;      // number of z chunks per column
    int nt = 256;      // threads per workgroup (256 or 512)
    int vec = 1;       // real types: threads own 16-byte x runs (falls back when unaligned);
                       // 2: the same with z-terms accumulated forward, 3: 2 with loads two planes ahead
    int align = 0;     // 16-byte-run kernels: band rows start at lane multiples of 16
};

// out = a * ( -1/2 lap x + (veff + s) x ) + b * old    (old may alias out).
// With coef (device memory, three elements of the arithmetic type), a, b, s
// are read from coef[0..2] when the kernel runs instead.
void stencil_step(int e, const StencilGeom &g, const void *x, long ldx, const void *old, long ldo,
                  void *out, long ldout, const void *veff, int ncol, double a, double b, double s,
                  const StepCfg &cfg, hipStream_t st, const void *coef = nullptr);
int stencil_points_per_thread(int e, const StepCfg &cfg);
// True when the shape can run the grid's x rows (halo slots and LDS pitch);
// for vec shapes, also when nx is a multiple of the 16-byte run.
bool stencil_shape_fits(int e, int nx, const StepCfg &cfg);

// Nonlocal projector data in device layout.
//   points: all (image, sphere point) pairs grouped by physical atom;
//   chi_pm: point-major projector rows padded to npm entries;
//   coverage: unique grid points covered by any sphere, each listing its
//   entries (indices into points) for the owner-computes scatter.
struct NlDev {
    int n_atom = 0, npm = 0, n_pts = 0, n_cov = 0, n_img = 0;
    int max_atom_pts = 0;               // most sphere points of one atom (all images)
    const int *atom_pt_off = nullptr;   // n_atom+1
    const int *pt_gpos = nullptr;       // n_pts
    const int *pt_img = nullptr;        // n_pts
    const int *pt_atom = nullptr;       // n_pts
    const void *chi = nullptr;          // n_pts*npm (double for F64/C128, float otherwise)
    const void *gamma = nullptr;        // n_atom*npm, dV*Gamma (same precision as chi)
    const int *cov_off = nullptr;       // n_cov+1
    const int *cov_gpos = nullptr;      // n_cov
    const int *ent_pt = nullptr;        // entries
    const void *phase = nullptr;        // n_img complex e^{-i k.R_img} (complex types only)
};

// beta[n][atom*npm + j] = dV*Gamma_j * sum_points chi_j * x (with Bloch phase)
void nl_gather(int e, const NlDev &nl, const void *x, long ldx, void *beta, int ncol,
               hipStream_t st);
// y[g] += a * sum_entries chi . beta (conjugate Bloch phase); with acoef
// (device memory, arithmetic type) a is read from acoef[0] instead
void nl_scatter(int e, const NlDev &nl, const void *beta, void *y, long ldy, int ncol, double a,
                hipStream_t st, const void *acoef = nullptr);

// ---- batched filter: many independent units, one launch per kernel ----
// One unit: an operator (grid, coefficients, Bloch phases, Veff, nonlocal
// data with this unit's image phases) and its three column blocks (input X
// and the two work blocks D1, D2 of the recurrence).
struct BatchUnit {
    StencilGeom g;
    const void *veff = nullptr;        // device, in the batch's precision
    NlDev nl;                          // n_atom == 0: no nonlocal part
    const void *buf[3] = {nullptr, nullptr, nullptr};
    long ld[3] = {0, 0, 0};
    int ncol = 0;
};

// Device tables of a batch for one element type (opaque to callers).
struct BatchPlan;
BatchPlan *batch_plan_create(int e, const BatchUnit *units, int n, hipStream_t st);
void batch_plan_destroy(BatchPlan *p);
// Filter scalars of every step: coef[(step * n + u) * 3 + {0, 1, 2}] = (a, b, s).
void batch_set_coef(BatchPlan *p, const double *coef, int steps, hipStream_t st);
// One recurrence step for every unit: buf[iout] = a (H - shift) buf[ix] + b buf[iold]
// (iold = 3: no old term).  Launches gather, stencil and scatter per group.
void batch_step(BatchPlan *p, int step, int ix, int iold, int iout, hipStream_t st);
int batch_launches_per_step(const BatchPlan *p);

// C (n x n) = alpha * sum of nsplit consecutive n x n partial products in work
// (fixed order).  C is FP64 (F64 or C128) whatever the precision e of the
// partials: single-precision partials are summed in double precision.
void gram_reduce(int e, const void *work, int nsplit, int n, double alpha, void *C, hipStream_t st);
// Split-BF16x3 Gram partials on the matrix cores (kernels_mfma.hip): for
// real single-precision column-major A, B (K x N, leading dimensions and
// base addresses 16-byte aligned), part[s] (N x N, FP32) = A_s^T B_s over
// consecutive K chunks, each operand split into bfloat16 hi + lo and the
// product taken as hi.hi + hi.lo + lo.hi with FP32 accumulation.  At most
// nsplit chunks; returns the number written (sum them with gram_reduce).
int gram_bf16x3(const float *A, long lda, const float *B, long ldb, long K, int N, float *part, int nsplit,
                hipStream_t st);
int gram_bf16x3_splits(long K, int N);
// FP64 Gram partials on the matrix cores (kernels_mfma.hip,
// v_mfma_f64_16x16x4f64): part[s] (N x N) = A_s^T B_s over consecutive K
// chunks; with sym only the upper tiles, mirrored into the lower ones
// (exact for A == B; for a product whose Hermitian part is taken, that
// part).  Columns must be 16-byte aligned (gram_f64_fits).  At most nsplit
// chunks; returns the number written (sum them with gram_reduce).
int gram_f64(const double *A, long lda, const double *B, long ldb, long K, int N, double *part, int nsplit,
             bool sym, hipStream_t st);
int gram_f64_splits(long K, int N);
bool gram_f64_fits(const double *A, long lda, const double *B, long ldb);
// Split-BF16x3 rotation on the matrix cores (kernels_mfma.hip): C (M x N) =
// A (M x K) Q (K x N) for a real FP32 block A (column-major, ld lda) and an
// FP64 Q (ld ldq), both split into bfloat16 hi + lo, hi.hi + hi.lo + lo.hi
// accumulated in FP32; C is written as float, or as double when c_double.
// work: rotate_bf16x3_work(K, N) bytes of device memory (Q's split halves).
size_t rotate_bf16x3_work(int K, int N);
void rotate_bf16x3(const float *A, long lda, long M, const double *Q, int ldq, int K, int N, void *C, long ldc,
                   bool c_double, void *work, hipStream_t st);
// Small dense helpers on n x n column-major matrices.
void hermitian_part(int e, void *H, int n, hipStream_t st);
void zero_lower(int e, void *H, int n, hipStream_t st);

// Vector helpers (Lanczos).  Partial sums per block are written to
// partials (one double per block); returns the number of blocks used.
int dot_partials(int e, const void *x, const void *y, long n, double *partials, hipStream_t st);
// y -= a*x + b*z; z <- x; partial sums of |y|^2
int lanczos_update(int e, void *y, void *x, void *z, double a, double b, long n, double *partials,
                   hipStream_t st);
void scale_copy(int e, const void *src, void *dst, double s, long n, hipStream_t st);

// Copies ncol columns between leading dimensions (device to device).
void copy_cols(int e, const void *src, long lds, void *dst, long ldd, long nd, int ncol,
               hipStream_t st);
// Same, converting between precisions of the same kind (FP64 <-> FP32,
// complex128 <-> complex64).
// y = a x + b o + v over n16 16-byte vectors (v: nv16 vectors, re-read
// cyclically): the stencil step's global traffic without the stencil.
void stream_probe(const void *x, const void *o, const void *v, void *y, long n16, long nv16, int blocks,
                  hipStream_t st);
void convert_cols(int e_src, const void *src, long lds, int e_dst, void *dst, long ldd, long nd, int ncol,
                  hipStream_t st);
constexpr int kPartialBlocks = 512;

}  // namespace gpu
}  // namespace lynx
