// HIP kernels for gfx950 (MI355X): fused FD stencil + local potential +
// Chebyshev recurrence, nonlocal projector gather/scatter, subspace GEMMs and
// small helpers.  Launch wrappers at the bottom.
#include "kernels.hpp"
#include "common.hpp"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <vector>
#include <string>

namespace lynx {
namespace gpu {
namespace {

// ------------------------------------------------------------ element ops --
template <class S> struct Tr;
template <> struct Tr<double> { using A = double; static constexpr bool C = false; };
template <> struct Tr<float> { using A = float; static constexpr bool C = false; };
template <> struct Tr<double2> { using A = double; static constexpr bool C = true; };
template <> struct Tr<float2> { using A = float; static constexpr bool C = true; };

// the FP64 type of each element type (partial sums of single-precision
// products are added in double precision)
template <class S> struct Wide { using T = S; };
template <> struct Wide<float> { using T = double; };
template <> struct Wide<float2> { using T = double2; };
__device__ __forceinline__ double widen(double x) { return x; }
__device__ __forceinline__ double widen(float x) { return (double)x; }
__device__ __forceinline__ double2 widen(double2 x) { return x; }
__device__ __forceinline__ double2 widen(float2 x) { return make_double2(x.x, x.y); }

// bfloat16 storage (the filter's BF16 mode): 8-bit mantissa in memory,
// arithmetic in float.  Conversion to float is exact; from float it rounds
// to nearest even (finite values).
struct bf16 {
    unsigned short b;
};
__device__ __forceinline__ float to_compute(bf16 v) { return __uint_as_float((unsigned)v.b << 16); }
__device__ __forceinline__ float to_compute(float v) { return v; }
__device__ __forceinline__ double to_compute(double v) { return v; }
__device__ __forceinline__ double2 to_compute(double2 v) { return v; }
__device__ __forceinline__ float2 to_compute(float2 v) { return v; }
template <class T> struct StoreAs;
template <> struct StoreAs<bf16> {
    static __device__ __forceinline__ bf16 from(float f)
    {
        unsigned u = __float_as_uint(f);
        u += 0x7fffu + ((u >> 16) & 1u);
        return bf16{(unsigned short)(u >> 16)};
    }
};
template <> struct StoreAs<float> { static __device__ __forceinline__ float from(float f) { return f; } };
template <> struct StoreAs<double> { static __device__ __forceinline__ double from(double f) { return f; } };
template <> struct StoreAs<float2> { static __device__ __forceinline__ float2 from(float2 f) { return f; } };
template <> struct StoreAs<double2> { static __device__ __forceinline__ double2 from(double2 f) { return f; } };

template <class S> __device__ __forceinline__ S zero();
template <> __device__ __forceinline__ double zero<double>() { return 0.0; }
template <> __device__ __forceinline__ float zero<float>() { return 0.0f; }
template <> __device__ __forceinline__ double2 zero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zero<float2>() { return make_float2(0.0f, 0.0f); }

__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ float add(float a, float b) { return a + b; }
__device__ __forceinline__ double2 add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

// scalar * element
__device__ __forceinline__ double scl(double a, double x) { return a * x; }
__device__ __forceinline__ float scl(float a, float x) { return a * x; }
__device__ __forceinline__ double2 scl(double a, double2 x) { return make_double2(a * x.x, a * x.y); }
__device__ __forceinline__ float2 scl(float a, float2 x) { return make_float2(a * x.x, a * x.y); }

// a * x + y with scalar a
__device__ __forceinline__ double fm(double a, double x, double y) { return fma(a, x, y); }
__device__ __forceinline__ float fm(float a, float x, float y) { return fmaf(a, x, y); }
__device__ __forceinline__ double2 fm(double a, double2 x, double2 y) { return make_double2(fma(a, x.x, y.x), fma(a, x.y, y.y)); }
__device__ __forceinline__ float2 fm(float a, float2 x, float2 y) { return make_float2(fmaf(a, x.x, y.x), fmaf(a, x.y, y.y)); }

// x * (pr + i pi)
__device__ __forceinline__ double2 cmul(double2 x, double pr, double pi)
{
    return make_double2(x.x * pr - x.y * pi, x.x * pi + x.y * pr);
}
__device__ __forceinline__ float2 cmul(float2 x, float pr, float pi)
{
    return make_float2(x.x * pr - x.y * pi, x.x * pi + x.y * pr);
}
template <class S, class A>
__device__ __forceinline__ S phased(S v, A pr, A pi)
{
    if constexpr (Tr<S>::C) return cmul(v, pr, pi);
    else return v;
}

// conj(a) * b and its accumulation
__device__ __forceinline__ double cfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float cfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c)
{
    return make_double2(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 c)
{
    return make_float2(fmaf(a.x, b.x, fmaf(a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(-a.y, b.x, c.y)));
}
// a * b + c (complex)
__device__ __forceinline__ double mfma_(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float mfma_(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double2 mfma_(double2 a, double2 b, double2 c)
{
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ float2 mfma_(float2 a, float2 b, float2 c)
{
    return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}

__device__ __forceinline__ double re2(double a) { return a * a; }
__device__ __forceinline__ double re2(float a) { return double(a) * a; }
__device__ __forceinline__ double re2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double re2(float2 a) { return double(a.x) * a.x + double(a.y) * a.y; }
__device__ __forceinline__ double redot(double a, double b) { return a * b; }
__device__ __forceinline__ double redot(float a, float b) { return double(a) * b; }
__device__ __forceinline__ double redot(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ double redot(float2 a, float2 b) { return double(a.x) * b.x + double(a.y) * b.y; }

// ------------------------------------------------------- fused stencil step --
// One workgroup owns a band of full x-rows of one column and marches along z
// over a chunk of planes.  Each thread owns RP points of the band and keeps
// their 2R+1 z-neighbours in a register ring; the current xy-plane (band plus
// R halo rows above and below, plus R ghost columns on each side) is staged in
// LDS, double-buffered, so x/y neighbours come from LDS and every value of x
// is read from HBM about once per step.  Periodic and Bloch wrap are applied
// while loading: ghost values are wrapped copies times e^{-i k L} (low side)
// or e^{+i k L} (high side), as the reference halo does
// (HaloExchange.cpp:104-189).  The work item order puts all bands of one
// (column, z-chunk) on the same XCD, which dispatches workgroups round-robin.
template <class A>
struct StepArgs {
    const void *x;
    const void *old;
    void *out;
    const A *veff;
    long ldx, ldo, ldout;
    int nx, ny, nz, ncol;
    int tyr, nbands, zch, nzch, nwork;
    int lpr;            // 16-byte-run kernels: lanes per band row (>= runs per row)
    A c[3][kMaxR + 1];
    A a, b, s;
    A ph[3][2][2];
    // when set, (a, b, s) are read from this device triple instead (the
    // filter's steps in a replayed HIP graph take their scalars from a
    // table uploaded per call)
    const A *coef = nullptr;
};

// Per-launch operands of a step (buffers and recurrence scalars), kept
// apart from the operator record so that a batch can read the record from
// device memory while these arrive as kernel arguments.
template <class A>
struct StepIO {
    const void *x;
    const void *old;
    void *out;
    long ldx, ldo, ldout;
    A a, b, s;
};

template <class A>
__device__ inline StepIO<A> step_io(const StepArgs<A> &p)
{
    if (p.coef) return StepIO<A>{p.x, p.old, p.out, p.ldx, p.ldo, p.ldout, p.coef[0], p.coef[1], p.coef[2]};
    return StepIO<A>{p.x, p.old, p.out, p.ldx, p.ldo, p.ldout, p.a, p.b, p.s};
}

// Loads of one z-step, issued one step ahead of their use.
template <class S, int RP, int HP>
struct StepLoads {
    S front[RP];                 // plane z+R of the owned points (ring input)
    S old[RP];                   // Y_{k-2} at plane z
    typename Tr<S>::A v[RP];     // Veff at plane z
    S halo[HP];                  // halo-row values of plane z
};

// element i of a uniform base pointer, addressed by a 32-bit byte offset so
// the load can use a scalar base plus a vector offset
template <class T>
__device__ __forceinline__ T ld_at(const T *base, unsigned i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st_at(T *base, unsigned i, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (unsigned)sizeof(T)) = v;
}

// XCD-aware work order: the hardware hands workgroups to the 8 XCDs round
// robin, so workgroup b runs on XCD b % 8; mapping it to work item
// (b % 8) * per + b / 8 gives each XCD a contiguous range of work items.
__device__ __forceinline__ int xcd_work_item()
{
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

template <class S, int R, int RP, int NT, int HP, int W>
__device__ __forceinline__ void stencil_step_body(const StepArgs<typename Tr<S>::A> &p, const StepIO<typename Tr<S>::A> io, const int w)
{
    using A = typename Tr<S>::A;
    constexpr bool C = Tr<S>::C;
    constexpr int NR = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    S *lds = reinterpret_cast<S *>(smem_raw);

    // W: LDS row pitch, a compile-time constant >= nx + 2R so that every
    // neighbour read is one base address plus an immediate offset
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    // each LDS plane buffer ends with one dummy slot per lane that absorbs
    // the writes of idle lanes and of elements without a ghost copy (no
    // branches, and no two lanes of a wave writing the same address)
    const int dummy = W * (p.tyr + 2 * R) + (int)threadIdx.x;
    const int plane = W * (p.tyr + 2 * R) + NT;

    const int band = w % p.nbands;
    const int rest = w / p.nbands;
    const int zc = rest % p.nzch;
    const int col = rest / p.nzch;
    const int y0 = band * p.tyr;
    const int rows = min(p.tyr, ny - y0);
    const int z0 = zc * p.zch;
    const int z1 = min(z0 + p.zch, nz);
    const int nxny = nx * ny;

    const S *__restrict__ X = reinterpret_cast<const S *>(io.x) + (long)col * io.ldx;
    const S *__restrict__ O = reinterpret_cast<const S *>(io.old) + (long)col * io.ldo;
    S *__restrict__ Y = reinterpret_cast<S *>(io.out) + (long)col * io.ldout;
    const bool has_old = io.old != nullptr && io.b != A(0);   // no Y_(k-2) stream without a buffer

    const int tid = threadIdx.x;
    const int npts = rows * nx;
    // owned points: plane offset (0 for idle lanes), LDS index and the LDS
    // indices of its wrapped x-ghost copies (dummy when none)
    unsigned qoff[RP];   // element offsets within a plane
    int qli[RP], qgh[RP], qgl[RP];
    bool qv[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int q = tid + r * NT;
        qv[r] = q < npts;
        const int qq = qv[r] ? q : 0;
        const int qy = qq / nx;
        const int qx = qq - qy * nx;
        qoff[r] = qv[r] ? (y0 + qy) * nx + qx : 0;
        const int li = (qy + R) * W + qx + R;
        qli[r] = qv[r] ? li : dummy;
        qgh[r] = (qv[r] && qx < R) ? li + nx : dummy;
        qgl[r] = (qv[r] && qx >= nx - R) ? li - nx : dummy;
    }
    // halo slots: R rows below and R above the band, wrapped in y
    const int nh = 2 * R * nx;
    unsigned hsrc[HP];
    int hli[HP], hgh[HP], hgl[HP], hys[HP];
#pragma unroll
    for (int h = 0; h < HP; ++h) {
        const int t = tid + h * NT;
        const bool v = t < nh;
        const int tt = v ? t : 0;
        const int i = tt / nx;
        const int xx = tt - i * nx;
        const int yb = (i < R) ? (i - R) : (rows + i - R);
        int yg = y0 + yb, ys = 0;
        if (yg < 0) { yg += ny; ys = 1; }
        else if (yg >= ny) { yg -= ny; ys = 2; }
        hsrc[h] = v ? yg * nx + xx : 0;
        const int li = (yb + R) * W + xx + R;
        hli[h] = v ? li : dummy;
        hgh[h] = (v && xx < R) ? li + nx : dummy;
        hgl[h] = (v && xx >= nx - R) ? li - nx : dummy;
        hys[h] = ys;
    }

    auto load_plane = [&](int zz, S *dst) {
        int side = -1;
        if (zz < 0) { zz += nz; side = 0; }
        else if (zz >= nz) { zz -= nz; side = 1; }
        const S *src = X + zz * nxny;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            S v = ld_at(src, qoff[r]);
            if constexpr (C) {
                if (side >= 0) v = phased(v, p.ph[2][side][0], p.ph[2][side][1]);
            }
            dst[r] = v;
        }
    };
    // every load of plane zs (clamped into the chunk, so the look-ahead past
    // the last plane stays in bounds)
    auto issue = [&](int zs, StepLoads<S, RP, HP> &L) {
        zs = min(zs, z1 - 1);
        load_plane(zs + R, L.front);
        const int zo = zs * nxny;
        const S *Oz = O + zo;
        const A *Vz = p.veff + zo;
        const S *Xz = X + zo;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            L.old[r] = has_old ? ld_at(Oz, qoff[r]) : zero<S>();
            L.v[r] = ld_at(Vz, qoff[r]);
        }
#pragma unroll
        for (int h = 0; h < HP; ++h) L.halo[h] = ld_at(Xz, hsrc[h]);
    };
    // an element and its wrapped x-ghost copies into the LDS plane
    auto put = [&](S *L, int li, int gh, int gl, S v) {
        L[li] = v;
        L[gh] = phased(v, p.ph[0][1][0], p.ph[0][1][1]);
        L[gl] = phased(v, p.ph[0][0][0], p.ph[0][0][1]);
    };

    S ring[RP][NR];
#pragma unroll
    for (int i = 0; i < 2 * R; ++i) {
        S tmp[RP];
        load_plane(z0 - R + i, tmp);
#pragma unroll
        for (int r = 0; r < RP; ++r) ring[r][i] = tmp[r];
    }

    const A diag = p.c[0][0] + p.c[1][0] + p.c[2][0];
    int buf = 0;
    // one z-step: prefetch the next step first, then consume this one
    auto step = [&](int z, StepLoads<S, RP, HP> &cur, StepLoads<S, RP, HP> &nxt) {
        issue(z + 1, nxt);
        S *L = lds + buf * plane;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            ring[r][2 * R] = cur.front[r];
            put(L, qli[r], qgh[r], qgl[r], ring[r][R]);
        }
#pragma unroll
        for (int h = 0; h < HP; ++h) {
            S v = cur.halo[h];
            if constexpr (C) {
                const A pr = hys[h] == 1 ? p.ph[1][0][0] : (hys[h] == 2 ? p.ph[1][1][0] : A(1));
                const A pi = hys[h] == 1 ? p.ph[1][0][1] : (hys[h] == 2 ? p.ph[1][1][1] : A(0));
                v = cmul(v, pr, pi);
            }
            put(L, hli[h], hgh[h], hgl[h], v);
        }
        __syncthreads();
        S *Yz = Y + z * nxny;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            const S *c = L + qli[r];
            S lap = scl(diag, ring[r][R]);
#pragma unroll
            for (int k = 1; k <= R; ++k) {
                lap = fm(p.c[0][k], add(c[k], c[-k]), lap);
                lap = fm(p.c[1][k], add(c[k * W], c[-k * W]), lap);
                lap = fm(p.c[2][k], add(ring[r][R + k], ring[r][R - k]), lap);
            }
            const S hval = fm(cur.v[r] + io.s, ring[r][R], scl(A(-0.5), lap));
            S res = scl(io.a, hval);
            if (has_old) res = fm(io.b, cur.old[r], res);
            if (qv[r]) st_at(Yz, qoff[r], res);
        }
#pragma unroll
        for (int r = 0; r < RP; ++r)
#pragma unroll
            for (int i = 0; i < 2 * R; ++i) ring[r][i] = ring[r][i + 1];
        buf ^= 1;
    };

    StepLoads<S, RP, HP> la, lb;
    issue(z0, la);
    int z = z0;
    for (; z + 1 < z1; z += 2) {
        step(z, la, lb);
        step(z + 1, lb, la);
    }
    if (z < z1) step(z, la, lb);
}

template <class S, int R, int RP, int NT, int HP, int W>
__global__ void __launch_bounds__(NT) k_stencil_step(StepArgs<typename Tr<S>::A> p)
{
    const int w = xcd_work_item();
    if (w >= p.nwork) return;
    stencil_step_body<S, R, RP, NT, HP, W>(p, step_io(p), w);
}

// --------------------------------------- fused stencil, non-orthogonal cell --
// The non-orthogonal Laplacian adds three mixed terms to the orthogonal one
// (Hamiltonian.cpp:176-267):
//   sum_pq cxy[p] d1y[q] (x(+p,+q) - x(+p,-q) - x(-p,+q) + x(-p,-q))  (x,y)
// and the same for (x,z) with cxz, d1z and (y,z) with cyz, d1z.  Each double
// sum factors into two one-dimensional antisymmetric sums:
//   xy: sum_p cxy[p] (g(x+p) - g(x-p)),  g = sum_q d1y[q] (x(y+q) - x(y-q))
//   xz + yz: sum_q d1z[q] (h(z+q) - h(z-q)),
//            h = sum_p cxz[p] (x(x+p) - x(x-p)) + cyz[p] (x(y+p) - x(y-p))
// so a point costs 36 LDS reads and about 60 FMAs instead of 108 four-point
// groups.  g and h are in-plane quantities: each plane passes through LDS
// once; g goes back to LDS (with x ghosts) for its x-differences, and h
// enters the z-difference by forward accumulation: plane f adds its
// contributions to the 2R+1 pending outputs f-R..f+R held in registers, and
// output f-R is complete after plane f (the potential term joins the
// accumulator when the output's own plane passes, so no plane is kept).  A chunk of planes [z0, z1) thus
// streams planes z0-R .. z1-1+R.  One barrier per plane: the (plane,
// g-plane) pair is double-buffered and the xy term of plane f is taken one
// step later, once its g plane is complete.
// Coefficients are passed factored, D2_a[k] = beta_a w2[k], mixed
// c_ab[k] = alpha_ab w1[k] and D1_a[k] = w1[k] / h_a, so that the kernel
// holds 24 scalars instead of 56 (they live in SGPRs).
template <class A>
struct NonorthArgs {
    StepArgs<A> b;
    A w1[kMaxR + 1];
    A w2[kMaxR + 1];
    A beta[3];
    A axy_hy;            // alpha_xy / h_y
    A axz, ayz;          // alpha_xz, alpha_yz
    A ihz;               // 1 / h_z
};

// PD: planes of global loads in flight ahead of the plane being consumed
// (1: the next plane; 2: the two next planes, for short-row grids whose
// steps are latency-bound)
template <class S, int R, int RP, int NT, int HP, int W, int PD = 1>
__device__ __forceinline__ void stencil_nonorth_body(const NonorthArgs<typename Tr<S>::A> &q, const StepIO<typename Tr<S>::A> io, const int w)
{
    using A = typename Tr<S>::A;
    constexpr bool C = Tr<S>::C;
    constexpr int NA = 2 * R + 1;
    const StepArgs<A> &p = q.b;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int dummy = W * (p.tyr + 2 * R) + (int)threadIdx.x;
    const int plane = W * (p.tyr + 2 * R) + NT;
    // LDS: two (plane, g-plane) pairs used alternately, so one barrier per
    // plane suffices: the g of plane f is read one step later
    S *const lds0 = reinterpret_cast<S *>(smem_raw);

    const int band = w % p.nbands;
    const int rest = w / p.nbands;
    const int zc = rest % p.nzch;
    const int col = rest / p.nzch;
    const int y0 = band * p.tyr;
    const int rows = min(p.tyr, ny - y0);
    const int z0 = zc * p.zch;
    const int z1 = min(z0 + p.zch, nz);
    const int nxny = nx * ny;

    const S *__restrict__ X = reinterpret_cast<const S *>(io.x) + (long)col * io.ldx;
    const S *__restrict__ O = reinterpret_cast<const S *>(io.old) + (long)col * io.ldo;
    S *__restrict__ Y = reinterpret_cast<S *>(io.out) + (long)col * io.ldout;
    const bool has_old = io.old != nullptr && io.b != A(0);   // no Y_(k-2) stream without a buffer

    const int tid = threadIdx.x;
    const int npts = rows * nx;
    unsigned qoff[RP];
    int qli[RP], qgh[RP], qgl[RP];
    bool qv[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        const int t = tid + r * NT;
        qv[r] = t < npts;
        const int tt = qv[r] ? t : 0;
        const int qy = tt / nx;
        const int qx = tt - qy * nx;
        qoff[r] = qv[r] ? (y0 + qy) * nx + qx : 0;
        const int li = (qy + R) * W + qx + R;
        qli[r] = qv[r] ? li : dummy;
        qgh[r] = (qv[r] && qx < R) ? li + nx : dummy;
        qgl[r] = (qv[r] && qx >= nx - R) ? li - nx : dummy;
    }
    const int nh = 2 * R * nx;
    unsigned hsrc[HP];
    int hli[HP], hgh[HP], hgl[HP], hys[HP];
#pragma unroll
    for (int h = 0; h < HP; ++h) {
        const int t = tid + h * NT;
        const bool v = t < nh;
        const int tt = v ? t : 0;
        const int i = tt / nx;
        const int xx = tt - i * nx;
        const int yb = (i < R) ? (i - R) : (rows + i - R);
        int yg = y0 + yb, ys = 0;
        if (yg < 0) { yg += ny; ys = 1; }
        else if (yg >= ny) { yg -= ny; ys = 2; }
        hsrc[h] = v ? yg * nx + xx : 0;
        const int li = (yb + R) * W + xx + R;
        hli[h] = v ? li : dummy;
        hgh[h] = (v && xx < R) ? li + nx : dummy;
        hgl[h] = (v && xx >= nx - R) ? li - nx : dummy;
        hys[h] = ys;
    }

    // A workgroup that owns the whole periodic z range marches the nz planes
    // once (no 2R wrapped planes): the R outputs at each end then lack the
    // contributions of the planes across the z boundary.  The first R steps
    // leave the bottom outputs' wrapped part (planes 0..R-1 seen from
    // nz-R..nz-1, Bloch factor e^{+ikL}) as partial results in Y, the next R
    // steps store the top outputs 0..R-1 without their wrapped part, and
    // after the march the ring holds the rest of both, which is added to
    // those lines of Y (the planes nz-R..nz-1 seen from 0..R-1 with e^{-ikL}).
    const bool wrapz = p.nzch == 1 && nz >= 2 * R;
    const int f_end = wrapz ? z1 - 1 : z1 - 1 + R;
    // every load of plane f (owned points and halo rows, z-wrapped) and of
    // the output plane f - R (Y_{k-2}, Veff), clamped into range
    auto issue = [&](int f, StepLoads<S, RP, HP> &L) {
        f = min(f, f_end);
        int zz = f;
        if (zz < 0) zz += nz;
        else if (zz >= nz) zz -= nz;
        const S *Xz = X + zz * nxny;
#pragma unroll
        for (int r = 0; r < RP; ++r) L.front[r] = ld_at(Xz, qoff[r]);
#pragma unroll
        for (int h = 0; h < HP; ++h) L.halo[h] = ld_at(Xz, hsrc[h]);
        const int zo = max(f - R, z0) * nxny;
        const int zv = min(max(f, z0), z1 - 1) * nxny;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            L.old[r] = has_old ? ld_at(O + zo, qoff[r]) : zero<S>();
            L.v[r] = ld_at(p.veff + zv, qoff[r]);
        }
    };
    auto put = [&](S *L, int li, int gh, int gl, S v) {
        L[li] = v;
        L[gh] = phased(v, p.ph[0][1][0], p.ph[0][1][1]);
        L[gl] = phased(v, p.ph[0][0][0], p.ph[0][0][1]);
    };

    // acc[j] accumulates -2 (H - s) x at output plane f - R + j; the output's
    // own plane adds -2 (Veff + s) x there, so H x = -acc / 2 at the end
    S acc[RP][NA];
#pragma unroll
    for (int r = 0; r < RP; ++r)
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[r][j] = zero<S>();
    const A diag = (q.beta[0] + q.beta[1] + q.beta[2]) * q.w2[0];

    int buf = 0;
    auto step = [&](int f, StepLoads<S, RP, HP> &cur, StepLoads<S, RP, HP> &nxt) {
        issue(f + PD, nxt);
        S *LX = lds0 + buf * 2 * plane;
        S *LG = LX + plane;
        const S *LGp = lds0 + (buf ^ 1) * 2 * plane + plane;   // g of plane f - 1
        // z-wrap phase of this plane (uniform)
        A zr = A(1), zi = A(0);
        if constexpr (C) {
            if (f < 0) { zr = p.ph[2][0][0]; zi = p.ph[2][0][1]; }
            else if (f >= nz) { zr = p.ph[2][1][0]; zi = p.ph[2][1][1]; }
        }
        S xf[RP];
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            xf[r] = phased(cur.front[r], zr, zi);
            put(LX, qli[r], qgh[r], qgl[r], xf[r]);
        }
#pragma unroll
        for (int h = 0; h < HP; ++h) {
            S v = cur.halo[h];
            if constexpr (C) {
                const A pr = hys[h] == 1 ? p.ph[1][0][0] : (hys[h] == 2 ? p.ph[1][1][0] : A(1));
                const A pi = hys[h] == 1 ? p.ph[1][0][1] : (hys[h] == 2 ? p.ph[1][1][1] : A(0));
                v = cmul(cmul(v, pr, pi), zr, zi);
            }
            put(LX, hli[h], hgh[h], hgl[h], v);
        }
        __syncthreads();
        // xy term of plane f - 1 (its g plane is complete), into output f - 1
        if (f - 1 >= z0 && f - 1 < z1) {
#pragma unroll
            for (int r = 0; r < RP; ++r) {
                const S *g = LGp + qli[r];
                S mxy = zero<S>();
#pragma unroll
                for (int k = 1; k <= R; ++k) mxy = fm(q.w1[k], add(g[k], scl(A(-1), g[-k])), mxy);
                acc[r][R - 1] = fm(q.axy_hy, mxy, acc[r][R - 1]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        S hz[RP], xzs[RP];
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            const S *c = LX + qli[r];
            S sx2 = zero<S>(), sy2 = zero<S>(), sx1 = zero<S>(), sy1 = zero<S>();
            // x then y neighbours; the scheduling barrier keeps the compiler
            // from hoisting all 24 LDS reads at once (register pressure)
#pragma unroll
            for (int k = 1; k <= R; ++k) {
                const S xp = c[k], xm = c[-k];
                sx2 = fm(q.w2[k], add(xp, xm), sx2);
                sx1 = fm(q.w1[k], add(xp, scl(A(-1), xm)), sx1);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 1; k <= R; ++k) {
                const S yp = c[k * W], ym = c[-k * W];
                sy2 = fm(q.w2[k], add(yp, ym), sy2);
                sy1 = fm(q.w1[k], add(yp, scl(A(-1), ym)), sy1);
            }
            // in-plane second differences, the potential term (-2 (V + s) x)
            S a2 = fm(diag - A(2) * (cur.v[r] + io.s), xf[r], scl(q.beta[0], sx2));
            acc[r][R] = add(acc[r][R], fm(q.beta[1], sy2, a2));
            hz[r] = scl(q.ihz, fm(q.axz, sx1, scl(q.ayz, sy1)));
            xzs[r] = scl(q.beta[2], xf[r]);
            put(LG, qli[r], qgh[r], qgl[r], sy1);
        }
        const bool out = f - R >= z0;
        S *Yz = Y + (f - R) * nxny;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            // z-terms of plane f for outputs f-R+j, j != R (distance d = R-j)
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                const int d = R - j;
                if (d != 0) {
                    const int ad = d > 0 ? d : -d;
                    acc[r][j] = fm(q.w2[ad], xzs[r], acc[r][j]);
                    acc[r][j] = fm(d > 0 ? q.w1[ad] : -q.w1[ad], hz[r], acc[r][j]);
                }
            }
            if (out) {
                S res = scl(A(-0.5) * io.a, acc[r][0]);
                if (has_old) res = fm(io.b, cur.old[r], res);
                if (qv[r]) st_at(Yz, qoff[r], res);
            } else if (wrapz && qv[r]) {
                // bottom output nz + f - R: its contributions from planes
                // 0..f, across the boundary
                const int zo = (nz + f - R) * nxny;
                S res = scl(A(-0.5) * io.a, phased(acc[r][0], p.ph[2][1][0], p.ph[2][1][1]));
                if (has_old) res = fm(io.b, ld_at(O + zo, qoff[r]), res);
                st_at(Y + zo, qoff[r], res);
            }
#pragma unroll
            for (int j = 0; j + 1 < NA; ++j) acc[r][j] = acc[r][j + 1];
            acc[r][NA - 1] = zero<S>();
        }
        buf ^= 1;
    };

    // a ring of PD + 1 load sets: the step on plane f consumes its set and
    // refills the one the previous step consumed with plane f + PD
    StepLoads<S, RP, HP> lring[PD + 1];
    int f = wrapz ? z0 : z0 - R;
#pragma unroll
    for (int d = 0; d < PD; ++d) issue(f + d, lring[d]);
    for (; f <= f_end; f += PD + 1) {
#pragma unroll
        for (int d = 0; d <= PD; ++d)
            if (f + d <= f_end) step(f + d, lring[d], lring[(d + PD) % (PD + 1)]);
    }
    if (!wrapz) return;
    // the xy term of the last plane (its g plane is complete after a barrier)
    __syncthreads();
    {
        const S *LGp = lds0 + (buf ^ 1) * 2 * plane + plane;
#pragma unroll
        for (int r = 0; r < RP; ++r) {
            const S *g = LGp + qli[r];
            S mxy = zero<S>();
#pragma unroll
            for (int k = 1; k <= R; ++k) mxy = fm(q.w1[k], add(g[k], scl(A(-1), g[-k])), mxy);
            acc[r][R - 1] = fm(q.axy_hy, mxy, acc[r][R - 1]);
        }
    }
    // ring slot j holds output nz - R + j: the rest of the bottom outputs
    // (j < R), and for j >= R the planes nz-R..nz-1 seen from top output
    // j - R across the boundary
#pragma unroll
    for (int r = 0; r < RP; ++r) {
        if (!qv[r]) continue;
#pragma unroll
        for (int j = 0; j < 2 * R; ++j) {
            const int o = j < R ? nz - R + j : j - R;
            const S c = j < R ? acc[r][j] : phased(acc[r][j], p.ph[2][0][0], p.ph[2][0][1]);
            S *Yo = Y + o * nxny;
            st_at(Yo, qoff[r], fm(A(-0.5) * io.a, c, ld_at(Yo, qoff[r])));
        }
    }
}

template <class S, int R, int RP, int NT, int HP, int W>
__global__ void __launch_bounds__(NT) k_stencil_nonorth(NonorthArgs<typename Tr<S>::A> q)
{
    const int w = xcd_work_item();
    if (w >= q.b.nwork) return;
    stencil_nonorth_body<S, R, RP, NT, HP, W>(q, step_io(q.b), w);
}

// ------------------------------------------------ fused stencil, 16-B runs --
// Same step as k_stencil_step for real element types, with each thread owning
// runs of VW consecutive x points: 16-byte runs (two doubles, four floats),
// or 8-byte runs of two floats when nx is not a multiple of four.  A run is
// loaded, stored and kept in LDS as one vector; its x neighbours come from
// 1 + 2*XO/VW aligned LDS reads shared by the run, its y neighbours from one
// read per offset.  This cuts the LDS instructions per point by about 2.5x
// and makes every global access a dwordx4 (dwordx2).  Requires nx % VW == 0
// and columns aligned to the run size.
template <class S, int VW>
struct alignas(sizeof(S) * VW) VRun {
    S v[VW];
};

template <class S, int VW>
__device__ __forceinline__ VRun<S, VW> vzero()
{
    VRun<S, VW> r;
#pragma unroll
    for (int e = 0; e < VW; ++e) r.v[e] = S(0);
    return r;
}

template <int VW, class S>
__device__ __forceinline__ VRun<S, VW> vld(const S *base, unsigned i)
{
    return *reinterpret_cast<const VRun<S, VW> *>(reinterpret_cast<const char *>(base) + i * (unsigned)sizeof(S));
}

// a run of storage type T loaded into compute type S, and stored back
template <int VW, class S, class T>
__device__ __forceinline__ VRun<S, VW> vldc(const T *base, unsigned i)
{
    if constexpr (std::is_same<S, T>::value) {
        return vld<VW>(base, i);
    } else {
        const VRun<T, VW> t = vld<VW>(base, i);
        VRun<S, VW> r;
#pragma unroll
        for (int e = 0; e < VW; ++e) r.v[e] = to_compute(t.v[e]);
        return r;
    }
}
template <int VW, class S, class T>
__device__ __forceinline__ void vstc(T *base, unsigned i, const VRun<S, VW> &v)
{
    VRun<T, VW> t;
#pragma unroll
    for (int e = 0; e < VW; ++e) t.v[e] = StoreAs<T>::from(v.v[e]);
    *reinterpret_cast<VRun<T, VW> *>(reinterpret_cast<char *>(base) + i * (unsigned)sizeof(T)) = t;
}

template <class S, int VW, int RV, int HP>
struct VecLoads {
    VRun<S, VW> front[RV];
    VRun<S, VW> old[RV];
    VRun<S, VW> v[RV];
    VRun<S, VW> halo[HP];
};

template <class S, int R, int RV, int NT, int HP, int W, int VW, class T = S>
__device__ __forceinline__ void stencil_vec_body(const StepArgs<S> &p, const StepIO<S> io, const int w)
{
    constexpr int XO = (R + VW - 1) / VW * VW;   // LDS column of x = 0, vector aligned
    constexpr int NXV = 1 + 2 * XO / VW;         // 16-byte reads covering x-R .. x+VW-1+R
    constexpr int NR = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    S *lds = reinterpret_cast<S *>(smem_raw);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int base = W * (p.tyr + 2 * R);
    const int plane = base + NT * VW;            // + one dummy run per lane

    const int band = w % p.nbands;
    const int rest = w / p.nbands;
    const int zc = rest % p.nzch;
    const int col = rest / p.nzch;
    const int y0 = band * p.tyr;
    const int rows = min(p.tyr, ny - y0);
    const int z0 = zc * p.zch;
    const int z1 = min(z0 + p.zch, nz);
    const int nxny = nx * ny;

    const T *__restrict__ X = reinterpret_cast<const T *>(io.x) + (long)col * io.ldx;
    const T *__restrict__ O = reinterpret_cast<const T *>(io.old) + (long)col * io.ldo;
    T *__restrict__ Y = reinterpret_cast<T *>(io.out) + (long)col * io.ldout;
    const bool has_old = io.old != nullptr && io.b != S(0);   // no Y_(k-2) stream without a buffer

    const int tid = threadIdx.x;
    const int dummy = base + tid * VW;
    // owned runs: plane offset (element index), LDS index and x of the first
    // element (kNoX for lanes without a run, whose stores go to the dummy).
    // Lane q takes run q % lpr of band row q / lpr; with lpr a multiple of 16
    // no group of 16 lanes (one LDS cycle of 16-byte accesses) straddles two
    // rows, so row starts never collide in the same banks.
    constexpr int kNoX = -(1 << 20);
    const int runs = nx / VW;
    const int lpr = p.lpr;
    unsigned qoff[RV];
    int qli[RV], qx0[RV];
    bool qv[RV];
#pragma unroll
    for (int r = 0; r < RV; ++r) {
        const int q = tid + r * NT;
        const int qy = q / lpr;
        const int qr = q - qy * lpr;
        qv[r] = qy < rows && qr < runs;
        const int qx = qv[r] ? qr * VW : 0;
        const int qyy = qv[r] ? qy : 0;
        qoff[r] = qv[r] ? (y0 + qyy) * nx + qx : 0;
        qli[r] = qv[r] ? (qyy + R) * W + XO + qx : dummy;
        qx0[r] = qv[r] ? qx : kNoX;
    }
    unsigned hsrc[HP];
    int hli[HP], hx0[HP];
#pragma unroll
    for (int h = 0; h < HP; ++h) {
        const int t = tid + h * NT;
        const int ti = t / lpr;
        const int tr = t - ti * lpr;
        const bool v = ti < 2 * R && tr < runs;
        const int i = v ? ti : 0;
        const int xx = v ? tr * VW : 0;
        const int yb = (i < R) ? (i - R) : (rows + i - R);
        int yg = y0 + yb;
        if (yg < 0) yg += ny;
        else if (yg >= ny) yg -= ny;
        hsrc[h] = v ? yg * nx + xx : 0;
        hli[h] = v ? (yb + R) * W + XO + xx : dummy;
        hx0[h] = v ? xx : kNoX;
    }

    auto load_plane = [&](int zz, VRun<S, VW> *dst) {
        if (zz < 0) zz += nz;
        else if (zz >= nz) zz -= nz;
        const T *src = X + zz * nxny;
#pragma unroll
        for (int r = 0; r < RV; ++r) dst[r] = vldc<VW, S>(src, qoff[r]);
    };
    auto issue = [&](int zs, VecLoads<S, VW, RV, HP> &L) {
        zs = min(zs, z1 - 1);
        load_plane(zs + R, L.front);
        const int zo = zs * nxny;
#pragma unroll
        for (int r = 0; r < RV; ++r) {
            L.old[r] = has_old ? vldc<VW, S>(O + zo, qoff[r]) : vzero<S, VW>();
            L.v[r] = vld<VW>(p.veff + zo, qoff[r]);
        }
#pragma unroll
        for (int h = 0; h < HP; ++h) L.halo[h] = vldc<VW, S>(X + zo, hsrc[h]);
    };
    // a run, and one wrapped copy of it when it lies within XO of an x face
    // (x + nx to the right of the row, x - nx to the left; nx >= 2 XO, so a
    // run is near at most one face): one 16-byte write per run, to the
    // lane's dummy slot when no copy is needed
    auto put = [&](S *L, int li, int x0, const VRun<S, VW> &v) {
        *reinterpret_cast<VRun<S, VW> *>(L + li) = v;
        const bool lo = unsigned(x0) < unsigned(XO);
        const bool hi = unsigned(x0 - (nx - XO)) < unsigned(XO);
        *reinterpret_cast<VRun<S, VW> *>(L + (lo ? li + nx : (hi ? li - nx : dummy))) = v;
    };

    VRun<S, VW> ring[RV][NR];
#pragma unroll
    for (int i = 0; i < 2 * R; ++i) {
        VRun<S, VW> tmp[RV];
        load_plane(z0 - R + i, tmp);
#pragma unroll
        for (int r = 0; r < RV; ++r) ring[r][i] = tmp[r];
    }

    const S diag = p.c[0][0] + p.c[1][0] + p.c[2][0];
    int buf = 0;
    auto step = [&](int z, VecLoads<S, VW, RV, HP> &cur, VecLoads<S, VW, RV, HP> &nxt) {
        issue(z + 1, nxt);
        S *L = lds + buf * plane;
#pragma unroll
        for (int r = 0; r < RV; ++r) {
            ring[r][2 * R] = cur.front[r];
            put(L, qli[r], qx0[r], ring[r][R]);
        }
#pragma unroll
        for (int h = 0; h < HP; ++h) put(L, hli[h], hx0[h], cur.halo[h]);
        __syncthreads();
        T *Yz = Y + z * nxny;
#pragma unroll
        for (int r = 0; r < RV; ++r) {
            S xr[NXV * VW];
#pragma unroll
            for (int j = 0; j < NXV; ++j) {
                const VRun<S, VW> t = *reinterpret_cast<const VRun<S, VW> *>(L + qli[r] - XO + j * VW);
#pragma unroll
                for (int e = 0; e < VW; ++e) xr[j * VW + e] = t.v[e];
            }
            S lap[VW];
#pragma unroll
            for (int e = 0; e < VW; ++e) lap[e] = diag * ring[r][R].v[e];
#pragma unroll
            for (int k = 1; k <= R; ++k) {
                const VRun<S, VW> up = *reinterpret_cast<const VRun<S, VW> *>(L + qli[r] + k * W);
                const VRun<S, VW> dn = *reinterpret_cast<const VRun<S, VW> *>(L + qli[r] - k * W);
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    lap[e] = fm(p.c[0][k], add(xr[XO + e + k], xr[XO + e - k]), lap[e]);
                    lap[e] = fm(p.c[1][k], add(up.v[e], dn.v[e]), lap[e]);
                    lap[e] = fm(p.c[2][k], add(ring[r][R + k].v[e], ring[r][R - k].v[e]), lap[e]);
                }
            }
            VRun<S, VW> res;
#pragma unroll
            for (int e = 0; e < VW; ++e) {
                const S hval = fm(cur.v[r].v[e] + io.s, ring[r][R].v[e], S(-0.5) * lap[e]);
                res.v[e] = io.a * hval;
                if (has_old) res.v[e] = fm(io.b, cur.old[r].v[e], res.v[e]);
            }
            if (qv[r]) vstc<VW>(Yz, qoff[r], res);
        }
#pragma unroll
        for (int r = 0; r < RV; ++r)
#pragma unroll
            for (int i = 0; i < 2 * R; ++i) ring[r][i] = ring[r][i + 1];
        buf ^= 1;
    };

    VecLoads<S, VW, RV, HP> la, lb;
    issue(z0, la);
    int z = z0;
    for (; z + 1 < z1; z += 2) {
        step(z, la, lb);
        step(z + 1, lb, la);
    }
    if (z < z1) step(z, la, lb);
}

template <class S, int R, int RV, int NT, int HP, int W, int VW, class T>
__global__ void __launch_bounds__(NT) k_stencil_vec(StepArgs<S> p)
{
    const int w = xcd_work_item();
    if (w >= p.nwork) return;
    stencil_vec_body<S, R, RV, NT, HP, W, VW, T>(p, step_io(p), w);
}

// The same 16-byte-run step with the z-direction accumulated forward
// instead of kept as a ring of planes: the LDS plane is the newest plane f,
// whose in-plane terms go to output f and whose z-neighbour terms go to the
// 2R pending outputs f-R..f+R (2R+1 accumulators per run, as many registers
// as the ring).  Output f-R is complete after plane f.  The halo rows of
// plane f are then read in the same step in which the neighbouring bands
// read those rows as their own newest plane, so they come from L2 instead of
// being re-fetched R steps later.  Planes outside the chunk only feed the
// z-terms and skip LDS.
template <class S, int R, int RV, int NT, int HP, int W, int PD, int VW, class T = S>
__device__ __forceinline__ void stencil_vacc_body(const StepArgs<S> &p, const StepIO<S> io, const int w)
{
    constexpr int XO = (R + VW - 1) / VW * VW;
    constexpr int NXV = 1 + 2 * XO / VW;
    constexpr int NA = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    S *lds = reinterpret_cast<S *>(smem_raw);

    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const int base = W * (p.tyr + 2 * R);
    const int plane = base + NT * VW;

    const int band = w % p.nbands;
    const int rest = w / p.nbands;
    const int zc = rest % p.nzch;
    const int col = rest / p.nzch;
    const int y0 = band * p.tyr;
    const int rows = min(p.tyr, ny - y0);
    const int z0 = zc * p.zch;
    const int z1 = min(z0 + p.zch, nz);
    const int nxny = nx * ny;

    const T *__restrict__ X = reinterpret_cast<const T *>(io.x) + (long)col * io.ldx;
    const T *__restrict__ O = reinterpret_cast<const T *>(io.old) + (long)col * io.ldo;
    T *__restrict__ Y = reinterpret_cast<T *>(io.out) + (long)col * io.ldout;
    const bool has_old = io.old != nullptr && io.b != S(0);   // no Y_(k-2) stream without a buffer

    const int tid = threadIdx.x;
    const int dummy = base + tid * VW;
    // lane-to-run mapping as in stencil_vec_body
    constexpr int kNoX = -(1 << 20);
    const int runs = nx / VW;
    const int lpr = p.lpr;
    unsigned qoff[RV];
    int qli[RV], qx0[RV];
    bool qv[RV];
#pragma unroll
    for (int r = 0; r < RV; ++r) {
        const int q = tid + r * NT;
        const int qy = q / lpr;
        const int qr = q - qy * lpr;
        qv[r] = qy < rows && qr < runs;
        const int qx = qv[r] ? qr * VW : 0;
        const int qyy = qv[r] ? qy : 0;
        qoff[r] = qv[r] ? (y0 + qyy) * nx + qx : 0;
        qli[r] = qv[r] ? (qyy + R) * W + XO + qx : dummy;
        qx0[r] = qv[r] ? qx : kNoX;
    }
    unsigned hsrc[HP];
    int hli[HP], hx0[HP];
#pragma unroll
    for (int h = 0; h < HP; ++h) {
        const int t = tid + h * NT;
        const int ti = t / lpr;
        const int tr = t - ti * lpr;
        const bool v = ti < 2 * R && tr < runs;
        const int i = v ? ti : 0;
        const int xx = v ? tr * VW : 0;
        const int yb = (i < R) ? (i - R) : (rows + i - R);
        int yg = y0 + yb;
        if (yg < 0) yg += ny;
        else if (yg >= ny) yg -= ny;
        hsrc[h] = v ? yg * nx + xx : 0;
        hli[h] = v ? (yb + R) * W + XO + xx : dummy;
        hx0[h] = v ? xx : kNoX;
    }

    const int f_end = z1 - 1 + R;
    // Odd z-chunks run downwards: plane f of the loop is physical plane
    // z0 + z1 - 1 - f (the z-stencil is symmetric, so the accumulation is
    // the same with the roles of +d and -d swapped).  Neighbouring chunks of
    // a column then reach their shared boundary planes (the R planes each
    // reads beyond its own range) at the same time, at both ends, and the
    // second read of those planes hits L2 instead of HBM.
    const bool down = (zc & 1) != 0;
    auto phys = [&](int f) { return down ? z0 + z1 - 1 - f : f; };
    // loads for plane f: its owned runs (z-wrapped), and, for planes of the
    // chunk, its halo runs and Veff; Y_{k-2} of output plane f - R
    auto issue = [&](int f, VecLoads<S, VW, RV, HP> &L) {
        f = min(f, f_end);
        int zz = phys(f);
        if (zz < 0) zz += nz;
        else if (zz >= nz) zz -= nz;
        const T *Xz = X + zz * nxny;
#pragma unroll
        for (int r = 0; r < RV; ++r) L.front[r] = vldc<VW, S>(Xz, qoff[r]);
        if (f >= z0 && f < z1) {
#pragma unroll
            for (int h = 0; h < HP; ++h) L.halo[h] = vldc<VW, S>(Xz, hsrc[h]);
#pragma unroll
            for (int r = 0; r < RV; ++r) L.v[r] = vld<VW>(p.veff + zz * nxny, qoff[r]);
        }
        if (has_old && f - R >= z0) {
            const int zo = phys(f - R) * nxny;
#pragma unroll
            for (int r = 0; r < RV; ++r) L.old[r] = vldc<VW, S>(O + zo, qoff[r]);
        }
    };
    auto put = [&](S *L, int li, int x0, const VRun<S, VW> &v) {   // as in stencil_vec_body
        *reinterpret_cast<VRun<S, VW> *>(L + li) = v;
        const bool lo = unsigned(x0) < unsigned(XO);
        const bool hi = unsigned(x0 - (nx - XO)) < unsigned(XO);
        *reinterpret_cast<VRun<S, VW> *>(L + (lo ? li + nx : (hi ? li - nx : dummy))) = v;
    };

    // acc[j]: -2 (H - s) x at output plane f - R + j, so H x = -acc / 2
    VRun<S, VW> acc[RV][NA];
#pragma unroll
    for (int r = 0; r < RV; ++r)
#pragma unroll
        for (int j = 0; j < NA; ++j) acc[r][j] = vzero<S, VW>();

    const S diag = p.c[0][0] + p.c[1][0] + p.c[2][0];
    int buf = 0;
    // PD: planes of loads in flight ahead of the one consumed (1 or 2)
    auto step = [&](int f, VecLoads<S, VW, RV, HP> &cur, VecLoads<S, VW, RV, HP> &nxt) {
        issue(f + PD, nxt);
        if (f >= z0 && f < z1) {
            S *L = lds + buf * plane;
#pragma unroll
            for (int r = 0; r < RV; ++r) put(L, qli[r], qx0[r], cur.front[r]);
#pragma unroll
            for (int h = 0; h < HP; ++h) put(L, hli[h], hx0[h], cur.halo[h]);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < RV; ++r) {
                S xr[NXV * VW];
#pragma unroll
                for (int j = 0; j < NXV; ++j) {
                    const VRun<S, VW> t = *reinterpret_cast<const VRun<S, VW> *>(L + qli[r] - XO + j * VW);
#pragma unroll
                    for (int e = 0; e < VW; ++e) xr[j * VW + e] = t.v[e];
                }
                S lap[VW];
#pragma unroll
                for (int e = 0; e < VW; ++e)
                    lap[e] = fm(diag - S(2) * (cur.v[r].v[e] + io.s), cur.front[r].v[e], acc[r][R].v[e]);
#pragma unroll
                for (int k = 1; k <= R; ++k) {
                    const VRun<S, VW> up = *reinterpret_cast<const VRun<S, VW> *>(L + qli[r] + k * W);
                    const VRun<S, VW> dn = *reinterpret_cast<const VRun<S, VW> *>(L + qli[r] - k * W);
#pragma unroll
                    for (int e = 0; e < VW; ++e) {
                        lap[e] = fm(p.c[0][k], add(xr[XO + e + k], xr[XO + e - k]), lap[e]);
                        lap[e] = fm(p.c[1][k], add(up.v[e], dn.v[e]), lap[e]);
                    }
                }
#pragma unroll
                for (int e = 0; e < VW; ++e) acc[r][R].v[e] = lap[e];
            }
            buf ^= 1;
        }
        // z-neighbour terms of plane f for the outputs within R of it
#pragma unroll
        for (int r = 0; r < RV; ++r)
#pragma unroll
            for (int j = 0; j < NA; ++j) {
                if (j == R) continue;
                const int d = j < R ? R - j : j - R;
#pragma unroll
                for (int e = 0; e < VW; ++e) acc[r][j].v[e] = fm(p.c[2][d], cur.front[r].v[e], acc[r][j].v[e]);
            }
        if (f - R >= z0) {
            const int zp = phys(f - R);
            T *Yz = Y + zp * nxny;
#pragma unroll
            for (int r = 0; r < RV; ++r) {
                VRun<S, VW> res;
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    res.v[e] = (S(-0.5) * io.a) * acc[r][0].v[e];
                    if (has_old) res.v[e] = fm(io.b, cur.old[r].v[e], res.v[e]);
                }
                if (qv[r]) vstc<VW>(Yz, qoff[r], res);
            }
        }
#pragma unroll
        for (int r = 0; r < RV; ++r) {
#pragma unroll
            for (int j = 0; j + 1 < NA; ++j) acc[r][j] = acc[r][j + 1];
            acc[r][NA - 1] = vzero<S, VW>();
        }
    };

    int f = z0 - R;
    if constexpr (PD == 1) {
        VecLoads<S, VW, RV, HP> la, lb;
        issue(f, la);
        for (; f + 1 <= f_end; f += 2) {
            step(f, la, lb);
            step(f + 1, lb, la);
        }
        if (f <= f_end) step(f, la, lb);
    } else {
        // three load sets in rotation: the set refilled at step f is the
        // one consumed at step f - 1
        VecLoads<S, VW, RV, HP> la, lb, lc;
        issue(f, la);
        issue(f + 1, lb);
        for (; f + 2 <= f_end; f += 3) {
            step(f, la, lc);
            step(f + 1, lb, la);
            step(f + 2, lc, lb);
        }
        if (f <= f_end) step(f, la, lc);
        if (f + 1 <= f_end) step(f + 1, lb, la);
    }
}

template <class S, int R, int RV, int NT, int HP, int W, int PD, int VW, class T>
__global__ void __launch_bounds__(NT) k_stencil_vacc(StepArgs<S> p)
{
    const int w = xcd_work_item();
    if (w >= p.nwork) return;
    stencil_vacc_body<S, R, RV, NT, HP, W, PD, VW, T>(p, step_io(p), w);
}

// --------------------------------------------------------- nonlocal gather --
// One workgroup per (physical atom, column).  The four waves take
// interleaved runs of the atom's sphere points (all images); each lane
// accumulates npm projections, then a wave reduction, and the four wave sums
// are added in a fixed order through LDS (deterministic, no atomics).
template <class A>
struct NlArgs {
    const int *atom_pt_off, *pt_gpos, *pt_img, *pt_atom, *cov_off, *cov_gpos, *ent_pt;
    const A *chi, *gamma, *phase;
    int n_atom, n_cov, npm;
};

// SPLIT = false: one wave per column (four columns per workgroup, index c
// is the column group); SPLIT = true: the four waves share one column (c)
// and split the points, for atoms with many sphere points.
template <class S, int NPM, bool SPLIT, class T = S>
__device__ __forceinline__ void nl_gather_body(const NlArgs<typename Tr<S>::A> &p, const T *__restrict__ X,
                                               long ldx, S *__restrict__ beta, int ncol, int at, int c)
{
    using A = typename Tr<S>::A;
    __shared__ S part[4][NPM];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int n = SPLIT ? c : c * 4 + wave;
    if (n >= ncol) return;       // uniform over the workgroup when SPLIT
    const int pb = p.atom_pt_off[at], pe = p.atom_pt_off[at + 1];
    const T *Xn = X + (long)n * ldx;
    S acc[NPM];
#pragma unroll
    for (int j = 0; j < NPM; ++j) acc[j] = zero<S>();
    constexpr int kStride = SPLIT ? 256 : 64;
    for (int e = pb + (SPLIT ? wave * 64 : 0) + lane; e < pe; e += kStride) {
        S v = to_compute(Xn[p.pt_gpos[e]]);
        if constexpr (Tr<S>::C) {
            const int im = p.pt_img[e];
            v = cmul(v, p.phase[2 * im], p.phase[2 * im + 1]);
        }
        const A *ch = p.chi + (long)e * NPM;
#pragma unroll
        for (int j = 0; j < NPM; ++j) acc[j] = fm(ch[j], v, acc[j]);
    }
#pragma unroll
    for (int j = 0; j < NPM; ++j) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            if constexpr (Tr<S>::C) {
                acc[j].x += __shfl_xor(acc[j].x, off);
                acc[j].y += __shfl_xor(acc[j].y, off);
            } else {
                acc[j] += __shfl_xor(acc[j], off);
            }
        }
    }
    if constexpr (!SPLIT) {
        S *bn = beta + (long)n * p.n_atom * NPM + (long)at * NPM;
#pragma unroll
        for (int j = 0; j < NPM; ++j)
            if (lane == j) bn[j] = scl(p.gamma[at * NPM + j], acc[j]);
        return;
    }
#pragma unroll
    for (int j = 0; j < NPM; ++j)
        if (lane == j) part[wave][j] = acc[j];
    __syncthreads();
    if (wave == 0 && lane < NPM) {
        S t = zero<S>();
        // fixed order over the waves
#pragma unroll
        for (int w = 0; w < 4; ++w) t = add(t, part[w][lane]);
        S *bn = beta + (long)n * p.n_atom * NPM + (long)at * NPM;
        bn[lane] = scl(p.gamma[at * NPM + lane], t);
    }
}

template <class S, int NPM, bool SPLIT, class T>
__global__ void __launch_bounds__(256) k_nl_gather(NlArgs<typename Tr<S>::A> p, const T *__restrict__ X,
                                                   long ldx, S *__restrict__ beta, int ncol)
{
    nl_gather_body<S, NPM, SPLIT, T>(p, X, ldx, beta, ncol, blockIdx.x, blockIdx.y);
}

// atoms with more sphere points than this are gathered by four waves (one
// operator) or by parts (batches, k_nl_gather_part_batch)
constexpr int kGatherSplitPts = 1024;

// -------------------------------------------------------- nonlocal scatter --
// Owner-computes: one lane per covered grid point, one wave per column; sums
// chi . beta over every (image, point) entry at that grid point, so no two
// lanes write the same address and the result is deterministic.
// Many columns per wave: each entry's metadata and projector row are loaded
// once and reused for the wave's cpw columns (at most CPWMAX), so per column
// only the atom's beta row (shared by the wave's neighbouring points) and y
// itself move.  Per column the entries are summed in list order, then
// y += a * sum.
template <class S, int NPM, int CPWMAX, class T>
__device__ __forceinline__ void nl_scatter_mc_body(const NlArgs<typename Tr<S>::A> &p, const S *__restrict__ beta,
                                                   T *__restrict__ Y, long ldy, int ncol, typename Tr<S>::A a,
                                                   int cblock, int cgroup, int cpw)
{
    using A = typename Tr<S>::A;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int u = cblock * 64 + lane;
    const int c0 = (cgroup * 4 + wave) * cpw;
    if (u >= p.n_cov || c0 >= ncol) return;
    const int nc = min(cpw, ncol - c0);
    const int e0 = p.cov_off[u], e1 = p.cov_off[u + 1];
    const long g = p.cov_gpos[u];
    const long nb = (long)p.n_atom * NPM;
    if (e1 - e0 == 1) {
        // one entry (spheres do not overlap here): each column's y is
        // updated as soon as its product is formed
        const int pt = p.ent_pt[e0];
        A ch[NPM];
#pragma unroll
        for (int j = 0; j < NPM; ++j) ch[j] = p.chi[(long)pt * NPM + j];
        const long boff = (long)p.pt_atom[pt] * NPM;
        A pr = A(1), pi = A(0);
        if constexpr (Tr<S>::C) {
            const int im = p.pt_img[pt];
            pr = p.phase[2 * im];
            pi = -p.phase[2 * im + 1];
        }
        for (int c = c0; c < c0 + nc; ++c) {
            const S *bt = beta + c * nb + boff;
            S s = zero<S>();
#pragma unroll
            for (int j = 0; j < NPM; ++j) s = fm(ch[j], bt[j], s);
            if constexpr (Tr<S>::C) s = cmul(s, pr, pi);
            const long gi = (long)c * ldy + g;
            Y[gi] = StoreAs<T>::from(fm(a, s, to_compute(Y[gi])));
        }
        return;
    }
    S acc[CPWMAX];
#pragma unroll
    for (int k = 0; k < CPWMAX; ++k) acc[k] = zero<S>();
    for (int e = e0; e < e1; ++e) {
        const int pt = p.ent_pt[e];
        A ch[NPM];
#pragma unroll
        for (int j = 0; j < NPM; ++j) ch[j] = p.chi[(long)pt * NPM + j];
        const long boff = (long)p.pt_atom[pt] * NPM;
        A pr = A(1), pi = A(0);
        if constexpr (Tr<S>::C) {
            const int im = p.pt_img[pt];
            pr = p.phase[2 * im];
            pi = -p.phase[2 * im + 1];
        }
#pragma unroll
        for (int k = 0; k < CPWMAX; ++k) {
            if (k < nc) {
                const S *bt = beta + (c0 + k) * nb + boff;
                S s = zero<S>();
#pragma unroll
                for (int j = 0; j < NPM; ++j) s = fm(ch[j], bt[j], s);
                if constexpr (Tr<S>::C) s = cmul(s, pr, pi);
                acc[k] = add(acc[k], s);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < CPWMAX; ++k) {
        if (k < nc) {
            const long gi = (long)(c0 + k) * ldy + g;
            Y[gi] = StoreAs<T>::from(fm(a, acc[k], to_compute(Y[gi])));
        }
    }
}

template <class S, int NPM, int CPW, class T>
__global__ void __launch_bounds__(256) k_nl_scatter_mc(NlArgs<typename Tr<S>::A> p, const S *__restrict__ beta,
                                                       T *__restrict__ Y, long ldy, int ncol,
                                                       typename Tr<S>::A a, const typename Tr<S>::A *acoef)
{
    if (acoef) a = acoef[0];
    nl_scatter_mc_body<S, NPM, CPW, T>(p, beta, Y, ldy, ncol, a, blockIdx.x, blockIdx.y, CPW);
}

// ------------------------------------------------------- batched launches --
// A batch holds many independent units (k-points of one structure, or
// different structures with their own grids): each unit has its own
// operator record, buffers and filter coefficients, and one launch of each
// kernel covers every unit.  Work items of a unit are contiguous, so the
// XCD-aware order keeps a unit's bands on one XCD; a table maps each work
// item to its unit.
template <class A>
struct BatchRec {
    NonorthArgs<A> q;           // operator of the unit; pointers and a, b, s set per step
    const void *buf[4];         // X, D1, D2, none
    long ld[4];
    int item0;                  // first stencil work item
    int g_item0, s_item0;       // first gather / scatter block
    int ncol;
    int s_cpw;                  // scatter columns per wave
    long beta_off;              // first beta element of the unit
    NlArgs<A> nl;
    // gather by parts (units with large spheres): sphere point ranges
    // [pb, pe) of at most kGatherPartPts points of one atom, the first part
    // of each atom (n_atom + 1), and the unit's first partial sum
    const int2 *parts;
    const int *atom_part;
    int nparts;
    int r_item0;                // first reduction block
    long part_off;
};

// KIND: 0 orthogonal (scalar body), 1 non-orthogonal, 2 orthogonal with
// 16-byte runs and row-aligned lanes (real types, RP runs per thread; the
// forward-accumulation body with loads two planes ahead), 3 the same with
// 8-byte runs of two floats
template <class S, int R, int RP, int NT, int HP, int W, int KIND>
__global__ void __launch_bounds__(NT) k_stencil_batch(const BatchRec<typename Tr<S>::A> *__restrict__ recs,
                                                      const int *__restrict__ item_unit, int nitems,
                                                      const typename Tr<S>::A *__restrict__ coef, int ix,
                                                      int iold, int iout)
{
    using A = typename Tr<S>::A;
    const int w = xcd_work_item();
    if (w >= nitems) return;
    const int u = item_unit[w];
    const BatchRec<A> &r = recs[u];
    const StepIO<A> io{r.buf[ix], r.buf[iold], const_cast<void *>(r.buf[iout]), r.ld[ix], r.ld[iold], r.ld[iout],
                       coef[3 * u], coef[3 * u + 1], coef[3 * u + 2]};
    if constexpr (KIND == 1) {
        // batched units are small grids: two planes of loads in flight
        stencil_nonorth_body<S, R, RP, NT, HP, W, 2>(r.q, io, w - r.item0);
    } else if constexpr (KIND == 2 || KIND == 3) {
        if constexpr (!Tr<S>::C)
            stencil_vacc_body<S, R, RP, NT, HP, W, 2, KIND == 3 ? 2 : int(16 / sizeof(S))>(r.q.b, io, w - r.item0);
    } else {
        stencil_step_body<S, R, RP, NT, HP, W>(r.q.b, io, w - r.item0);
    }
}

// Batched gather of units whose atoms have small spheres: one workgroup per
// (atom, group of four columns), one wave per column.
template <class S, int NPM>
__global__ void __launch_bounds__(256) k_nl_gather_batch(const BatchRec<typename Tr<S>::A> *__restrict__ recs,
                                                         const int *__restrict__ item_unit, int nitems,
                                                         S *__restrict__ beta, int ix)
{
    using A = typename Tr<S>::A;
    const int w = blockIdx.x;
    if (w >= nitems) return;
    const int u = item_unit[w];
    const BatchRec<A> &r = recs[u];
    const NlArgs<A> nl = r.nl;
    const int lb = w - r.g_item0;
    nl_gather_body<S, NPM, false>(nl, reinterpret_cast<const S *>(r.buf[ix]), r.ld[ix], beta + r.beta_off,
                                  r.ncol, lb % nl.n_atom, lb / nl.n_atom);
}

// Gather by parts, for units whose atoms have large spheres: one workgroup
// per (part of at most kGatherPartPts sphere points of one atom, group of
// kGatherPartCols columns).  The part's x values (Bloch phase applied) are
// staged in LDS with point-contiguous loads; then the lanes are laid out
// projector-major (lane j of each NPM-lane subgroup holds projector j) and
// the subgroups take interleaved points, so one coalesced load fetches a
// point's whole projector row, used for every column of the group.  The
// subgroup sums are added in a fixed order through LDS and stored as the
// part's partial sums; k_nl_gather_reduce_batch adds the parts of each atom.
constexpr int kGatherPartPts = 256;
constexpr int kGatherPartCols = 8;

template <class S, int NPM>
__global__ void __launch_bounds__(256) k_nl_gather_part_batch(const BatchRec<typename Tr<S>::A> *__restrict__ recs,
                                                              const int *__restrict__ item_unit, int nitems,
                                                              S *__restrict__ partial, int ix)
{
    using A = typename Tr<S>::A;
    constexpr int CG = kGatherPartCols, PP = kGatherPartPts, NSUB = 256 / NPM;
    static_assert(CG * PP >= NSUB * CG * NPM, "the staging buffer holds the subgroup sums");
    __shared__ S xs[CG * PP];
    const int w = blockIdx.x;
    if (w >= nitems) return;
    const int u = item_unit[w];
    const BatchRec<A> &r = recs[u];
    const NlArgs<A> nl = r.nl;
    const int lb = w - r.g_item0;
    const int q = lb % r.nparts;
    const int c0 = (lb / r.nparts) * CG;
    const int ncol = r.ncol;
    const int2 pr = r.parts[q];
    const int np = pr.y - pr.x;
    const S *__restrict__ X = reinterpret_cast<const S *>(r.buf[ix]);
    const long ldx = r.ld[ix];
    const int tid = threadIdx.x;
    for (int i = tid; i < CG * PP; i += 256) {
        const int k = i / PP, e = i % PP;
        S v = zero<S>();
        if (e < np && c0 + k < ncol) {
            v = X[(long)(c0 + k) * ldx + nl.pt_gpos[pr.x + e]];
            if constexpr (Tr<S>::C) {
                const int im = nl.pt_img[pr.x + e];
                v = cmul(v, nl.phase[2 * im], nl.phase[2 * im + 1]);
            }
        }
        xs[i] = v;
    }
    __syncthreads();
    const int j = tid % NPM, sub = tid / NPM;
    S acc[CG];
#pragma unroll
    for (int k = 0; k < CG; ++k) acc[k] = zero<S>();
    for (int e = sub; e < np; e += NSUB) {
        const A ch = nl.chi[(long)(pr.x + e) * NPM + j];
#pragma unroll
        for (int k = 0; k < CG; ++k) acc[k] = fm(ch, xs[k * PP + e], acc[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < CG; ++k) xs[(sub * CG + k) * NPM + j] = acc[k];
    __syncthreads();
    for (int i = tid; i < CG * NPM; i += 256) {
        const int k = i / NPM, jj = i % NPM;
        if (c0 + k >= ncol) continue;
        S t = zero<S>();
        for (int sg = 0; sg < NSUB; ++sg) t = add(t, xs[(sg * CG + k) * NPM + jj]);
        partial[r.part_off + ((long)q * ncol + c0 + k) * NPM + jj] = t;
    }
}

// beta[n][atom][j] = gamma[atom][j] * (sum of the atom's part sums, in part
// order); one thread per (column, atom, projector) of a unit.
template <class S, int NPM>
__global__ void __launch_bounds__(256) k_nl_gather_reduce_batch(const BatchRec<typename Tr<S>::A> *__restrict__ recs,
                                                                const int *__restrict__ item_unit, int nitems,
                                                                const S *__restrict__ partial, S *__restrict__ beta)
{
    using A = typename Tr<S>::A;
    const int w = blockIdx.x;
    if (w >= nitems) return;
    const int u = item_unit[w];
    const BatchRec<A> &r = recs[u];
    const NlArgs<A> nl = r.nl;
    const long i = (long)(w - r.r_item0) * 256 + threadIdx.x;
    if (i >= (long)r.ncol * nl.n_atom * NPM) return;
    const int jj = (int)(i % NPM);
    const int at = (int)((i / NPM) % nl.n_atom);
    const int c = (int)(i / ((long)NPM * nl.n_atom));
    S t = zero<S>();
    const int q0 = r.atom_part[at], q1 = r.atom_part[at + 1];
    // loads issued four parts ahead; the sum stays in part order
#pragma unroll 4
    for (int q = q0; q < q1; ++q) t = add(t, partial[r.part_off + ((long)q * r.ncol + c) * NPM + jj]);
    beta[r.beta_off + i] = scl(nl.gamma[at * NPM + jj], t);
}

// batched scatter: columns per wave at most (one workgroup covers 4 cpw)
constexpr int kBatchScatterCpw = 4;

template <class S, int NPM>
__global__ void __launch_bounds__(256) k_nl_scatter_batch(const BatchRec<typename Tr<S>::A> *__restrict__ recs,
                                                          const int *__restrict__ item_unit, int nitems,
                                                          const S *__restrict__ beta,
                                                          const typename Tr<S>::A *__restrict__ coef, int iout)
{
    using A = typename Tr<S>::A;
    const int w = blockIdx.x;
    if (w >= nitems) return;
    const int u = item_unit[w];
    const BatchRec<A> &r = recs[u];
    const NlArgs<A> nl = r.nl;
    const int lb = w - r.s_item0;
    const int ncb = (nl.n_cov + 63) / 64;
    nl_scatter_mc_body<S, NPM, kBatchScatterCpw, S>(nl, beta + r.beta_off,
                                                    reinterpret_cast<S *>(const_cast<void *>(r.buf[iout])),
                                                    r.ld[iout], r.ncol, coef[3 * u], lb % ncb, lb / ncb, r.s_cpw);
}

// ------------------------------------------------------ Gram reduction --
// Sum of split-K partial Gram products in a fixed order, times alpha.
template <class S>
__global__ void k_gram_reduce(const S *__restrict__ work, int nsplit, int n, double alpha,
                              typename Wide<S>::T *C)
{
    using W = typename Wide<S>::T;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nn = (long)n * n;
    if (idx >= nn) return;
    W s = zero<W>();
    for (int z = 0; z < nsplit; ++z) s = add(s, widen(work[(long)z * nn + idx]));
    C[idx] = scl(alpha, s);
}


// ------------------------------------------------------------ small helpers --
template <class S>
__global__ void k_hermitian_part(S *H, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n || j >= n || i > j) return;
    if constexpr (Tr<S>::C) {
        if (i == j) {
            H[(long)i * n + i].y = 0;
            return;
        }
        const S a = H[(long)j * n + i], b = H[(long)i * n + j];
        S m;
        m.x = 0.5 * (a.x + b.x);
        m.y = 0.5 * (a.y - b.y);
        H[(long)j * n + i] = m;
        m.y = -m.y;
        H[(long)i * n + j] = m;
    } else {
        if (i == j) return;
        const S m = 0.5 * (H[(long)j * n + i] + H[(long)i * n + j]);
        H[(long)j * n + i] = m;
        H[(long)i * n + j] = m;
    }
}

template <class S>
__global__ void k_zero_lower(S *H, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n && i > j) H[(long)j * n + i] = zero<S>();
}

template <class S>
__global__ void __launch_bounds__(256) k_dot(const S *x, const S *y, long n, double *part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        s += redot(x[i], y[i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <class S>
__global__ void __launch_bounds__(256) k_lanczos_update(S *y, S *x, S *z, double a, double b,
                                                        long n, double *part)
{
    using A = typename Tr<S>::A;
    __shared__ double red[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        // y -= a*x + b*z  (evaluated as in Lanczos.hpp:162-165)
        S t = add(scl((A)a, x[i]), scl((A)b, z[i]));
        S v = add(y[i], scl(A(-1), t));
        y[i] = v;
        z[i] = x[i];
        s += re2(v);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <class S>
__global__ void k_scale_copy(const S *src, S *dst, double s, long n)
{
    using A = typename Tr<S>::A;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = scl((A)s, src[i]);
}

template <class S>
__global__ void k_copy_cols(const S *src, long lds, S *dst, long ldd, long nd)
{
    const S *s = src + (long)blockIdx.y * lds;
    S *d = dst + (long)blockIdx.y * ldd;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += (long)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// Precision conversion of a column block (double <-> float, element-wise;
// complex blocks are converted as interleaved pairs).
template <class D, class T>
__global__ void k_convert_cols(const T *src, long lds, D *dst, long ldd, long n)
{
    const T *s = src + (long)blockIdx.y * lds;
    D *d = dst + (long)blockIdx.y * ldd;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d[i] = static_cast<D>(s[i]);
}

// Streaming probe with the stencil step's global traffic and no stencil:
// y = a x + b o + v, x/o/y column blocks of n 16-byte vectors, v (one
// column, nv vectors) re-read per column from L2.  Its rate is the
// practical HBM ceiling for the step's access pattern.
__global__ void k_stream3(const double2 *__restrict__ x, const double2 *__restrict__ o,
                          const double2 *__restrict__ v, double2 *__restrict__ y, long n, long nv)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double2 a = x[i], b = o[i], c = v[i % nv];
        y[i] = make_double2(fma(0.5, a.x, fma(0.25, b.x, c.x)), fma(0.5, a.y, fma(0.25, b.y, c.y)));
    }
}

// Rounding to and widening from bfloat16 storage (real blocks; FP64 input is
// rounded through float).
template <class T>
__global__ void k_to_bf16_cols(const T *src, long lds, bf16 *dst, long ldd, long n)
{
    const T *s = src + (long)blockIdx.y * lds;
    bf16 *d = dst + (long)blockIdx.y * ldd;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d[i] = StoreAs<bf16>::from((float)s[i]);
}
template <class D>
__global__ void k_from_bf16_cols(const bf16 *src, long lds, D *dst, long ldd, long n)
{
    const bf16 *s = src + (long)blockIdx.y * lds;
    D *d = dst + (long)blockIdx.y * ldd;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d[i] = (D)to_compute(s[i]);
}

// ---------------------------------------------------------- dispatch helpers --
template <class F>
void dispatch(int e, F &&f)
{
    switch (e) {
        case F64: f(double{}); break;
        case C128: f(double2{}); break;
        case F32: f(float{}); break;
        case C64: f(float2{}); break;
        default: fail(1, "element type not supported by this operation");
    }
}

template <class S> constexpr int base_rp() { return int(32 / sizeof(S)); }

// LDS of one workgroup (gfx950: 160 KiB per CU)
constexpr long kLdsBytes = 160 * 1024;

// rows per band of the non-orthogonal body within the LDS: four planes of
// pitch * (rows + 2R) + NT elements (two (plane, g-plane) pairs)
template <class S>
int nonorth_max_rows(int pitch, int nt)
{
    return (int)((kLdsBytes / (4 * (long)sizeof(S)) - nt) / pitch) - 2 * kMaxR;
}

// LDS pitch of the scalar and non-orthogonal stencil tiles for rows of nx
// points: 64, 128, or 256 for long rows (nx + 2R > 128; 512-thread shapes,
// with kLongHalo halo slots per thread for up to 256-point rows)
inline int tile_pitch(int nx) { return nx + 2 * kMaxR <= 64 ? 64 : (nx + 2 * kMaxR <= 128 ? 128 : 256); }
constexpr int kLongHalo = 6;

// rows per band of the scalar body within the LDS: two planes of
// pitch * (rows + 2R) + NT elements
template <class S>
int step_max_rows(int pitch, int nt)
{
    return (int)((kLdsBytes / (2 * (long)sizeof(S)) - nt) / pitch) - 2 * kMaxR;
}

// Whether rows of nx points fit the scalar / non-orthogonal tiles of NT
// threads and RP points per thread (row pitch, halo slots, LDS).
template <class S>
bool step_rows_fit(int nx, int nt, int rp, bool nonorth)
{
    const int pitch = tile_pitch(nx);
    if (nx + 2 * kMaxR > 256 || nx > (long)nt * rp) return false;
    const int hb = pitch == 256 ? (nt >= 512 ? kLongHalo : 0) : (nt >= 512 ? 3 : 6);
    if (2 * kMaxR * nx > nt * hb) return false;
    return (nonorth ? nonorth_max_rows<S>(pitch, nt) : step_max_rows<S>(pitch, nt)) >= 1;
}

template <class S, int RP, int NT>
void launch_step_rp(const StencilGeom &g, const void *x, long ldx, const void *old, long ldo,
                    void *out, long ldout, const void *veff, int ncol, double a, double b, double s,
                    int nzch, hipStream_t st, const void *coef = nullptr)
{
    using A = typename Tr<S>::A;
    constexpr int HS = NT >= 512 ? 2 : 3;      // halo slots: rows up to 64 (256) or 85 (512) points
    constexpr int HB = NT >= 512 ? 3 : 6;      // rows up to 128 points
    StepArgs<A> p;
    p.x = x; p.old = old; p.out = out; p.veff = static_cast<const A *>(veff);
    p.ldx = ldx; p.ldo = ldo; p.ldout = ldout;
    p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.ncol = ncol;
    const long pts = (long)NT * RP;
    if (!step_rows_fit<S>(g.nx, NT, RP, false))
        fail(8, "grid rows longer than the fused stencil kernel's tile are not supported");
    const int pitch = tile_pitch(g.nx);
    const int max_rows = step_max_rows<S>(pitch, NT);
    int nb = (int)(((long)g.nx * g.ny + pts - 1) / pts);
    nb = std::max(1, std::min(nb, g.ny));
    p.tyr = (g.ny + nb - 1) / nb;
    while ((long)p.tyr * g.nx > pts || p.tyr > max_rows) { ++nb; p.tyr = (g.ny + nb - 1) / nb; }
    p.nbands = (g.ny + p.tyr - 1) / p.tyr;
    nzch = std::max(1, std::min(nzch, g.nz));
    p.zch = (g.nz + nzch - 1) / nzch;
    p.nzch = (g.nz + p.zch - 1) / p.zch;
    p.nwork = p.nbands * p.nzch * ncol;
    for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k <= kMaxR; ++k) p.c[ax][k] = (A)g.c[ax][k];
    p.a = (A)a; p.b = (A)b; p.s = (A)s;
    p.coef = static_cast<const A *>(coef);
    for (int ax = 0; ax < 3; ++ax)
        for (int sd = 0; sd < 2; ++sd)
            for (int q = 0; q < 2; ++q) p.ph[ax][sd][q] = (A)g.phase[ax][sd][q];
    const int blocks = ((p.nwork + 7) / 8) * 8;
    const size_t shm = 2 * sizeof(S) * ((size_t)pitch * (p.tyr + 2 * kMaxR) + NT);
    const bool small_halo = 2 * kMaxR * g.nx <= NT * HS;
    if (pitch == 64) {
        if (small_halo) hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HS, 64>), dim3(blocks), dim3(NT), shm, st, p);
        else hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HB, 64>), dim3(blocks), dim3(NT), shm, st, p);
    } else if (pitch == 128) {
        if (small_halo) hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HS, 128>), dim3(blocks), dim3(NT), shm, st, p);
        else hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HB, 128>), dim3(blocks), dim3(NT), shm, st, p);
    } else if constexpr (NT >= 512) {
        hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, kLongHalo, 256>), dim3(blocks), dim3(NT), shm, st, p);
    }
}

template <class S, int RP, int NT>
void launch_step_nonorth(const StencilGeom &g, const void *x, long ldx, const void *old, long ldo, void *out,
                         long ldout, const void *veff, int ncol, double a, double b, double s, int nzch,
                         hipStream_t st, const void *coef = nullptr)
{
    using A = typename Tr<S>::A;
    constexpr int HS = NT >= 512 ? 2 : 3;
    constexpr int HB = NT >= 512 ? 3 : 6;
    const long pts = (long)NT * RP;
    if (!step_rows_fit<S>(g.nx, NT, RP, true))
        fail(8, "grid rows longer than the fused stencil kernel's tile are not supported");
    NonorthArgs<A> q;
    StepArgs<A> &p = q.b;
    p.x = x; p.old = old; p.out = out; p.veff = static_cast<const A *>(veff);
    p.ldx = ldx; p.ldo = ldo; p.ldout = ldout;
    p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.ncol = ncol;
    // short rows: a 40-point pitch, so more workgroups fit a CU
    const int pitch = g.nx + 2 * kMaxR <= 40 ? 40 : tile_pitch(g.nx);
    const int max_rows = nonorth_max_rows<S>(pitch, NT);
    int nb = (int)(((long)g.nx * g.ny + pts - 1) / pts);
    nb = std::max(1, std::min(nb, g.ny));
    p.tyr = (g.ny + nb - 1) / nb;
    while ((long)p.tyr * g.nx > pts || p.tyr > max_rows) { ++nb; p.tyr = (g.ny + nb - 1) / nb; }
    p.nbands = (g.ny + p.tyr - 1) / p.tyr;
    nzch = std::max(1, std::min(nzch, g.nz));
    p.zch = (g.nz + nzch - 1) / nzch;
    p.nzch = (g.nz + p.zch - 1) / p.zch;
    p.nwork = p.nbands * p.nzch * ncol;
    for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k <= kMaxR; ++k) p.c[ax][k] = (A)g.c[ax][k];
    for (int k = 0; k <= kMaxR; ++k) {
        q.w1[k] = (A)g.w1[k];
        q.w2[k] = (A)g.w2[k];
    }
    for (int ax = 0; ax < 3; ++ax) q.beta[ax] = (A)g.beta[ax];
    q.axy_hy = (A)(g.alpha[0] / g.h[1]);
    q.axz = (A)g.alpha[1];
    q.ayz = (A)g.alpha[2];
    q.ihz = (A)(1.0 / g.h[2]);
    p.a = (A)a; p.b = (A)b; p.s = (A)s;
    p.coef = static_cast<const A *>(coef);
    for (int ax = 0; ax < 3; ++ax)
        for (int sd = 0; sd < 2; ++sd)
            for (int k = 0; k < 2; ++k) p.ph[ax][sd][k] = (A)g.phase[ax][sd][k];
    const int blocks = ((p.nwork + 7) / 8) * 8;
    const size_t shm = 4 * sizeof(S) * ((size_t)pitch * (p.tyr + 2 * kMaxR) + NT);
    const bool small_halo = 2 * kMaxR * g.nx <= NT * HS;
    if (pitch == 40) {
        if (small_halo) hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, HS, 40>), dim3(blocks), dim3(NT), shm, st, q);
        else hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, HB, 40>), dim3(blocks), dim3(NT), shm, st, q);
    } else if (pitch == 64) {
        if (small_halo) hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, HS, 64>), dim3(blocks), dim3(NT), shm, st, q);
        else hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, HB, 64>), dim3(blocks), dim3(NT), shm, st, q);
    } else if (pitch == 128) {
        if (small_halo) hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, HS, 128>), dim3(blocks), dim3(NT), shm, st, q);
        else hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, HB, 128>), dim3(blocks), dim3(NT), shm, st, q);
    } else if constexpr (NT >= 512 && sizeof(S) <= 8) {
        // long rows; complex128 planes of pitch 256 do not fit the LDS (step_rows_fit)
        hipLaunchKernelGGL((k_stencil_nonorth<S, kMaxR, RP, NT, kLongHalo, 256>), dim3(blocks), dim3(NT), shm, st, q);
    }
}

// run width of the vector stencil for rows of nx points: 16-byte runs, else
// (floats) 8-byte runs, else 0 (no vector form)
template <class S>
int vec_width(int nx)
{
    constexpr int V16 = int(16 / sizeof(S));
    if (nx % V16 == 0) return V16;
    if (V16 == 4 && nx % 2 == 0) return 2;
    return 0;
}

template <class S, int VW, int RV, int NT, int MODE, class T = S>
bool launch_step_vec(const StencilGeom &g, const void *x, long ldx, const void *old, long ldo,
                     void *out, long ldout, const void *veff, int ncol, double a, double b, double s,
                     int nzch, bool align, hipStream_t st, const void *coef = nullptr)
{
    constexpr int XO = (kMaxR + VW - 1) / VW * VW;
    constexpr int HS = NT >= 512 ? 1 : 2;      // halo runs per thread
    constexpr int HB = NT >= 512 ? 2 : 3;
    constexpr uintptr_t AM = sizeof(T) * VW - 1, AV = sizeof(S) * VW - 1;
    auto aligned = [](const void *ptr) { return (reinterpret_cast<uintptr_t>(ptr) & AM) == 0; };
    if (g.nx % VW || (ldx % VW) || (old && ldo % VW) || (ldout % VW) || !aligned(x) || (old && !aligned(old)) ||
        !aligned(out) || (reinterpret_cast<uintptr_t>(veff) & AV))
        return false;
    const int runs = g.nx / VW;
    const int lpr = align ? (runs + 15) / 16 * 16 : runs;   // lanes per band row
    const int rpb = NT * RV / lpr;                          // band rows per workgroup
    if (rpb < 1 || 2 * kMaxR * lpr > NT * HB || g.nx + 2 * XO > 128 || g.nx < 2 * XO) return false;
    StepArgs<S> p;
    p.x = x; p.old = old; p.out = out; p.veff = static_cast<const S *>(veff);
    p.ldx = ldx; p.ldo = ldo; p.ldout = ldout;
    p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.ncol = ncol;
    p.lpr = lpr;
    const int nb = (g.ny + rpb - 1) / rpb;
    p.tyr = (g.ny + nb - 1) / nb;
    p.nbands = (g.ny + p.tyr - 1) / p.tyr;
    nzch = std::max(1, std::min(nzch, g.nz));
    p.zch = (g.nz + nzch - 1) / nzch;
    p.nzch = (g.nz + p.zch - 1) / p.zch;
    p.nwork = p.nbands * p.nzch * ncol;
    for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k <= kMaxR; ++k) p.c[ax][k] = (S)g.c[ax][k];
    p.a = (S)a; p.b = (S)b; p.s = (S)s;
    p.coef = static_cast<const S *>(coef);
    for (int ax = 0; ax < 3; ++ax)
        for (int sd = 0; sd < 2; ++sd)
            for (int q = 0; q < 2; ++q) p.ph[ax][sd][q] = (S)g.phase[ax][sd][q];
    const int blocks = ((p.nwork + 7) / 8) * 8;
    const int pitch = g.nx + 2 * XO <= 64 ? 64 : 128;
    const size_t shm = 2 * sizeof(S) * ((size_t)pitch * (p.tyr + 2 * kMaxR) + NT * VW);
    const bool small_halo = 2 * kMaxR * lpr <= NT * HS;
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(blocks), dim3(NT), shm, st, p); };
    if constexpr (MODE > 0) {
        if (pitch == 64) {
            if (small_halo) go(k_stencil_vacc<S, kMaxR, RV, NT, HS, 64, MODE, VW, T>);
            else go(k_stencil_vacc<S, kMaxR, RV, NT, HB, 64, MODE, VW, T>);
        } else {
            if (small_halo) go(k_stencil_vacc<S, kMaxR, RV, NT, HS, 128, MODE, VW, T>);
            else go(k_stencil_vacc<S, kMaxR, RV, NT, HB, 128, MODE, VW, T>);
        }
    } else {
        if (pitch == 64) {
            if (small_halo) go(k_stencil_vec<S, kMaxR, RV, NT, HS, 64, VW, T>);
            else go(k_stencil_vec<S, kMaxR, RV, NT, HB, 64, VW, T>);
        } else {
            if (small_halo) go(k_stencil_vec<S, kMaxR, RV, NT, HS, 128, VW, T>);
            else go(k_stencil_vec<S, kMaxR, RV, NT, HB, 128, VW, T>);
        }
    }
    return true;
}

}  // namespace

// ------------------------------------------------------------ host wrappers --
int stencil_points_per_thread(int e, const StepCfg &cfg)
{
    int rp = 0;
    dispatch(e, [&](auto t) { rp = base_rp<decltype(t)>() / std::max(1, cfg.rp_div); });
    return rp;
}

bool stencil_shape_fits(int e, int nx, const StepCfg &cfg)
{
    if (e == B16) {
        // the run-vector kernels with float geometry
        if (!cfg.vec) return false;
        e = F32;
    }
    bool ok = false;
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        const int nt = cfg.nt >= 512 ? 512 : 256;
        if (cfg.vec) {
            if constexpr (Tr<S>::C) {
                ok = false;
            } else {
                const int VW = vec_width<S>(nx);
                const int XO = VW ? (kMaxR + VW - 1) / VW * VW : 0;
                const int hb = nt >= 512 ? 2 : 3;
                const int runs = VW ? nx / VW : 0;
                const int lpr = cfg.align ? (runs + 15) / 16 * 16 : runs;
                const int lanes = nt * (cfg.rp_div >= 2 ? 1 : 2);
                ok = VW > 0 && lpr <= lanes && 2 * kMaxR * lpr <= nt * hb && nx + 2 * XO <= 128 && nx >= 2 * XO;
            }
        } else {
            ok = step_rows_fit<S>(nx, nt, base_rp<S>() / std::max(1, cfg.rp_div), false);
        }
    });
    return ok;
}

// The run-vector kernels of a real compute type S with storage T: the form
// cfg.vec (1 ring, 2 forward accumulation, 3 the same two planes ahead), 16-
// or 8-byte float runs by the row length; false when the shape or the
// buffers' alignment does not allow it.
template <class S, class T>
bool launch_vec_modes(const StencilGeom &g, const void *x, long ldx, const void *old, long ldo, void *out,
                      long ldout, const void *veff, int ncol, double a, double b, double s, const StepCfg &cfg,
                      hipStream_t st, const void *coef)
{
    const int vw = vec_width<S>(g.nx);
    auto run = [&](auto rv, auto nt, auto acc) {
        constexpr int V16 = int(16 / sizeof(S));
        if constexpr (V16 == 4) {
            if (vw == 2)
                return launch_step_vec<S, 2, decltype(rv)::value, decltype(nt)::value, decltype(acc)::value, T>(
                    g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, cfg.align != 0, st, coef);
        }
        return launch_step_vec<S, V16, decltype(rv)::value, decltype(nt)::value, decltype(acc)::value, T>(
            g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, cfg.align != 0, st, coef);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using N256 = std::integral_constant<int, 256>;
    using N512 = std::integral_constant<int, 512>;
    auto by_mode = [&](auto mode) {
        using M = decltype(mode);
        if (cfg.nt >= 512) return cfg.rp_div >= 2 ? run(I1{}, N512{}, M{}) : run(I2{}, N512{}, M{});
        return cfg.rp_div >= 2 ? run(I1{}, N256{}, M{}) : run(I2{}, N256{}, M{});
    };
    // base RV = 2 runs per thread, RV = 1 for rp_div 2
    if (cfg.vec == 3) return by_mode(std::integral_constant<int, 2>{});
    if (cfg.vec == 2) return by_mode(std::integral_constant<int, 1>{});
    return by_mode(std::integral_constant<int, 0>{});
}

void stencil_step(int e, const StencilGeom &g, const void *x, long ldx, const void *old, long ldo,
                  void *out, long ldout, const void *veff, int ncol, double a, double b, double s,
                  const StepCfg &cfg, hipStream_t st, const void *coef)
{
    if (e == B16) {
        // bfloat16 storage, float arithmetic: run-vector kernels only
        if (!g.orth || !cfg.vec ||
            !launch_vec_modes<float, bf16>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg, st, coef))
            fail(8, "the bfloat16 filter needs an orthogonal cell, even row length and 4-byte aligned columns");
        return;
    }
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        constexpr int B = base_rp<S>();
        if (!g.orth) {
            // non-orthogonal cells: 2R+1 accumulators per point, so a quarter
            // of the orthogonal kernel's points per thread (8 B per thread for
            // points_div 1, half that for 2, at least one point)
            constexpr int B2 = B / 4 > 0 ? B / 4 : 1;
            constexpr int B4 = B / 8 > 0 ? B / 8 : 1;
            if (cfg.nt >= 512) {
                if (cfg.rp_div >= 2) launch_step_nonorth<S, B4, 512>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
                else launch_step_nonorth<S, B2, 512>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
            } else {
                if (cfg.rp_div >= 2) launch_step_nonorth<S, B4, 256>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
                else launch_step_nonorth<S, B2, 256>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
            }
            return;
        }
        if constexpr (!Tr<S>::C) {
            if (cfg.vec && launch_vec_modes<S, S>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg, st, coef))
                return;
        }
        if (cfg.nt >= 512) {
            if (cfg.rp_div >= 2)
                launch_step_rp<S, B / 2, 512>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
            else
                launch_step_rp<S, B, 512>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
        } else {
            if (cfg.rp_div >= 2)
                launch_step_rp<S, B / 2, 256>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
            else
                launch_step_rp<S, B, 256>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st, coef);
        }
    });
}

template <class S, int NPM, class T = S>
static void gather_npm(const NlDev &nl, const void *x, long ldx, void *beta, int ncol, hipStream_t st)
{
    using A = typename Tr<S>::A;
    NlArgs<A> p{nl.atom_pt_off, nl.pt_gpos, nl.pt_img, nl.pt_atom, nl.cov_off, nl.cov_gpos, nl.ent_pt,
                (const A *)nl.chi, (const A *)nl.gamma, (const A *)nl.phase, nl.n_atom, nl.n_cov, nl.npm};
    if (nl.max_atom_pts > kGatherSplitPts) {
        dim3 grid(nl.n_atom, ncol);
        hipLaunchKernelGGL((k_nl_gather<S, NPM, true, T>), grid, dim3(256), 0, st, p, (const T *)x, ldx,
                           (S *)beta, ncol);
    } else {
        dim3 grid(nl.n_atom, (ncol + 3) / 4);
        hipLaunchKernelGGL((k_nl_gather<S, NPM, false, T>), grid, dim3(256), 0, st, p, (const T *)x, ldx,
                           (S *)beta, ncol);
    }
}

template <class S, int NPM, class T = S>
static void scatter_npm(const NlDev &nl, const void *beta, void *y, long ldy, int ncol, double a,
                        const void *acoef, hipStream_t st)
{
    using A = typename Tr<S>::A;
    NlArgs<A> p{nl.atom_pt_off, nl.pt_gpos, nl.pt_img, nl.pt_atom, nl.cov_off, nl.cov_gpos, nl.ent_pt,
                (const A *)nl.chi, (const A *)nl.gamma, (const A *)nl.phase, nl.n_atom, nl.n_cov, nl.npm};
    constexpr int CPW = 4;
    dim3 grid((nl.n_cov + 63) / 64, (ncol + 4 * CPW - 1) / (4 * CPW));
    hipLaunchKernelGGL((k_nl_scatter_mc<S, NPM, CPW, T>), grid, dim3(256), 0, st, p, (const S *)beta, (T *)y, ldy,
                       ncol, (A)a, (const A *)acoef);
}

void nl_gather(int e, const NlDev &nl, const void *x, long ldx, void *beta, int ncol, hipStream_t st)
{
    if (nl.n_atom == 0 || nl.n_pts == 0) return;
    if (e == B16) {
        // bfloat16 storage: x read as bf16, beta in float
        if (nl.npm == 8) gather_npm<float, 8, bf16>(nl, x, ldx, beta, ncol, st);
        else if (nl.npm == 16) gather_npm<float, 16, bf16>(nl, x, ldx, beta, ncol, st);
        else gather_npm<float, 32, bf16>(nl, x, ldx, beta, ncol, st);
        return;
    }
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        if (nl.npm == 8) gather_npm<S, 8>(nl, x, ldx, beta, ncol, st);
        else if (nl.npm == 16) gather_npm<S, 16>(nl, x, ldx, beta, ncol, st);
        else gather_npm<S, 32>(nl, x, ldx, beta, ncol, st);
    });
}

void nl_scatter(int e, const NlDev &nl, const void *beta, void *y, long ldy, int ncol, double a,
                hipStream_t st, const void *acoef)
{
    if (nl.n_cov == 0) return;
    if (e == B16) {
        if (nl.npm == 8) scatter_npm<float, 8, bf16>(nl, beta, y, ldy, ncol, a, acoef, st);
        else if (nl.npm == 16) scatter_npm<float, 16, bf16>(nl, beta, y, ldy, ncol, a, acoef, st);
        else scatter_npm<float, 32, bf16>(nl, beta, y, ldy, ncol, a, acoef, st);
        return;
    }
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        if (nl.npm == 8) scatter_npm<S, 8>(nl, beta, y, ldy, ncol, a, acoef, st);
        else if (nl.npm == 16) scatter_npm<S, 16>(nl, beta, y, ldy, ncol, a, acoef, st);
        else scatter_npm<S, 32>(nl, beta, y, ldy, ncol, a, acoef, st);
    });
}

// ------------------------------------------------------------ batch plan --
struct BatchPlan {
    int e = 0, n = 0;
    void *recs = nullptr;
    struct SGroup {
        int kind = 0;             // 0 orthogonal, 1 non-orthogonal, 2 / 3 orthogonal 16- / 8-byte runs
        int W = 64;
        int *items = nullptr;
        int nitems = 0;
        int max_tyr = 0;
        bool small_halo = true;   // fewer halo slots per thread suffice for every unit of the group
        int nt = 256;             // threads per workgroup (192 for short-row non-orthogonal groups, 512 for long rows)
    };
    struct NGroup {
        int npm = 8;
        bool split = false;     // gather by parts (k_nl_gather_part_batch), then a reduction
        int *gitems = nullptr, *sitems = nullptr, *ritems = nullptr;
        int ngitems = 0, nsitems = 0, nritems = 0;
    };
    std::vector<SGroup> sg;
    std::vector<NGroup> ng;
    void *beta = nullptr;
    void *partial = nullptr;    // part sums of the gather by parts
    int2 *parts = nullptr;      // part point ranges, all units
    int *atom_part = nullptr;   // first part of each atom, all units
    void *coef = nullptr;
    int coef_steps = 0;
};

namespace {
void hip_check(hipError_t r, const char *what)
{
    if (r != hipSuccess) fail(2, std::string(what) + ": " + hipGetErrorString(r));
}

template <class T>
T *dev_upload(const std::vector<T> &v, hipStream_t st)
{
    T *d = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void **>(&d), std::max<size_t>(1, v.size() * sizeof(T))), "hipMalloc");
    if (!v.empty()) hip_check(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st), "copy");
    return d;
}

// points per thread of the batched stencil shapes (256 threads per group)
template <class S> constexpr int batch_rp_orth() { return base_rp<S>() / 2 > 0 ? base_rp<S>() / 2 : 1; }
template <class S> constexpr int batch_rp_nonorth() { return base_rp<S>() / 4 > 0 ? base_rp<S>() / 4 : 1; }
constexpr int kBatchNT = 256;
// runs per thread of the batched run-vector kinds (two, for taller bands
// and less halo, ran slower: more registers, fewer waves)
constexpr int kBatchRuns = 1;
// LDS row pitch of batched non-orthogonal units with rows of at most
// 40 - 2R points
constexpr int kShortNonorthPitch = 40;

template <class S>
void build_batch(BatchPlan *P, const BatchUnit *units, int n, hipStream_t st)
{
    using A = typename Tr<S>::A;
    std::vector<BatchRec<A>> recs(n);
    // stencil groups by (orth, pitch, threads), nonlocal groups by npm
    std::vector<int> sgi(n), ngi(n, -1);
    std::vector<int> kind(n), ntu(n, kBatchNT);
    // lanes of a non-orthogonal tile with nt threads that hold a point
    auto nonorth_fill = [&](const StencilGeom &g, int nt, int W) {
        const long pts = (long)nt * batch_rp_nonorth<S>();
        const int max_rows = nonorth_max_rows<S>(W, nt);
        if (max_rows < 1 || g.nx > pts) return 0.0;
        int nb = (int)std::max<long>(1, std::min<long>(((long)g.nx * g.ny + pts - 1) / pts, g.ny));
        int tyr = (g.ny + nb - 1) / nb;
        while ((long)tyr * g.nx > pts || tyr > max_rows) { ++nb; tyr = (g.ny + nb - 1) / nb; }
        nb = (g.ny + tyr - 1) / tyr;
        return (double)g.nx * g.ny / ((double)nb * pts);
    };
    for (int u = 0; u < n; ++u) {
        const StencilGeom &g = units[u].g;
        if (g.nx + 2 * kMaxR > 256) fail(8, "grid rows longer than the fused stencil kernel's tile are not supported");
        kind[u] = g.orth ? 0 : 1;
        int W = tile_pitch(g.nx);
        if (!g.orth && g.nx + 2 * kMaxR <= kShortNonorthPitch) W = kShortNonorthPitch;
        bool sh = 2 * kMaxR * g.nx <= kBatchNT * 3;
        if (W == 256) {
            // long rows: 512 threads, a 256-point pitch and kLongHalo halo
            // slots per thread, scalar bodies (no vector runs)
            if (!g.orth && sizeof(S) > 8)
                fail(8, "complex128 non-orthogonal grids with rows over 116 points are not supported");
            ntu[u] = 512;
            sh = false;
        } else if (!g.orth && 2 * kMaxR * g.nx <= 192 * 2 && nonorth_fill(g, 192, W) > nonorth_fill(g, kBatchNT, W)) {
            // short non-orthogonal rows: 192 threads (two halo slots each)
            // when they leave fewer lanes idle than 256, e.g. 8 rows of 24
            // points
            ntu[u] = 192;
            sh = true;
        }
        if constexpr (!Tr<S>::C) {
            // kind 2: 16-byte runs; kind 3: 8-byte runs of two floats for
            // rows whose length is not a multiple of four
            const BatchUnit &U = units[u];
            auto try_vec = [&](int VW, int kd) {
                const int XO = (kMaxR + VW - 1) / VW * VW;
                const uintptr_t am = (uintptr_t)(VW * sizeof(S)) - 1;
                auto al = [&](const void *ptr, long ld) {
                    return (reinterpret_cast<uintptr_t>(ptr) & am) == 0 && ld % VW == 0;
                };
                const int lpr = (g.nx / VW + 15) / 16 * 16;
                if (g.orth && g.nx % VW == 0 && g.nx + 2 * XO <= 128 && g.nx >= 2 * XO && lpr <= kBatchNT &&
                    al(U.buf[0], U.ld[0]) && al(U.buf[1], U.ld[1]) && al(U.buf[2], U.ld[2]) && al(U.veff, 0)) {
                    kind[u] = kd;
                    W = g.nx + 2 * XO <= 64 ? 64 : 128;
                    sh = 2 * kMaxR * lpr <= kBatchNT * 2;
                    return true;
                }
                return false;
            };
            if (!try_vec(int(16 / sizeof(S)), 2) && sizeof(S) == 4) try_vec(2, 3);
        }
        int k = -1;
        for (size_t i = 0; i < P->sg.size(); ++i)
            if (P->sg[i].kind == kind[u] && P->sg[i].W == W && P->sg[i].small_halo == sh && P->sg[i].nt == ntu[u])
                k = (int)i;
        if (k < 0) {
            P->sg.push_back({kind[u], W, nullptr, 0, 0, sh, ntu[u]});
            k = (int)P->sg.size() - 1;
        }
        sgi[u] = k;
        if (units[u].nl.n_atom > 0 && units[u].nl.n_pts > 0) {
            const bool split = units[u].nl.max_atom_pts > kGatherSplitPts;
            int j = -1;
            for (size_t i = 0; i < P->ng.size(); ++i)
                if (P->ng[i].npm == units[u].nl.npm && P->ng[i].split == split) j = (int)i;
            if (j < 0) {
                BatchPlan::NGroup G;
                G.npm = units[u].nl.npm;
                G.split = split;
                P->ng.push_back(G);
                j = (int)P->ng.size() - 1;
            }
            ngi[u] = j;
        }
    }
    std::vector<std::vector<int>> sitems(P->sg.size()), gitems(P->ng.size()), citems(P->ng.size()),
        ritems(P->ng.size());
    long beta_off = 0;
    std::vector<int2> parts_h;
    std::vector<int> atom_part_h;
    std::vector<long> part_first(n, -1), range_first(n, -1);
    long part_elems = 0;
    for (int u = 0; u < n; ++u) {
        const BatchUnit &U = units[u];
        const StencilGeom &g = U.g;
        BatchRec<A> &r = recs[u];
        NonorthArgs<A> &q = r.q;
        StepArgs<A> &p = q.b;
        p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.ncol = U.ncol;
        if (kind[u] == 2 || kind[u] == 3) {
            // kBatchRuns runs per thread, band rows at lane multiples of 16
            const int VW = kind[u] == 3 ? 2 : int(16 / sizeof(S));
            p.lpr = (g.nx / VW + 15) / 16 * 16;
            const int rpb = kBatchRuns * kBatchNT / p.lpr;
            const int nb = (g.ny + rpb - 1) / rpb;
            p.tyr = (g.ny + nb - 1) / nb;
        } else {
            const int nt = ntu[u];
            const long pts = (long)nt * (g.orth ? batch_rp_orth<S>() : batch_rp_nonorth<S>());
            const int hslots = nt == 512 ? kLongHalo : (nt == kBatchNT ? 6 : 2);
            if (g.nx > pts || 2 * kMaxR * g.nx > nt * hslots)
                fail(8, "grid rows longer than the batched stencil kernel's tile are not supported");
            p.lpr = 0;
            const int pitch = P->sg[sgi[u]].W;
            const int max_rows = g.orth ? step_max_rows<S>(pitch, nt) : nonorth_max_rows<S>(pitch, nt);
            if (max_rows < 1) fail(8, "grid rows longer than the batched stencil kernel's tile are not supported");
            int nb = (int)(((long)g.nx * g.ny + pts - 1) / pts);
            nb = std::max(1, std::min(nb, g.ny));
            p.tyr = (g.ny + nb - 1) / nb;
            while ((long)p.tyr * g.nx > pts || p.tyr > max_rows) { ++nb; p.tyr = (g.ny + nb - 1) / nb; }
        }
        p.nbands = (g.ny + p.tyr - 1) / p.tyr;
        p.zch = g.nz;
        p.nzch = 1;
        p.nwork = p.nbands * U.ncol;
        p.veff = static_cast<const A *>(U.veff);
        for (int ax = 0; ax < 3; ++ax)
            for (int k = 0; k <= kMaxR; ++k) p.c[ax][k] = (A)g.c[ax][k];
        for (int ax = 0; ax < 3; ++ax)
            for (int sd = 0; sd < 2; ++sd)
                for (int k = 0; k < 2; ++k) p.ph[ax][sd][k] = (A)g.phase[ax][sd][k];
        for (int k = 0; k <= kMaxR; ++k) {
            q.w1[k] = (A)g.w1[k];
            q.w2[k] = (A)g.w2[k];
        }
        for (int ax = 0; ax < 3; ++ax) q.beta[ax] = (A)g.beta[ax];
        q.axy_hy = (A)(g.alpha[0] / g.h[1]);
        q.axz = (A)g.alpha[1];
        q.ayz = (A)g.alpha[2];
        q.ihz = (A)(1.0 / g.h[2]);
        for (int i = 0; i < 3; ++i) {
            r.buf[i] = U.buf[i];
            r.ld[i] = U.ld[i];
        }
        r.buf[3] = nullptr;
        r.ld[3] = 0;
        r.ncol = U.ncol;
        P->sg[sgi[u]].max_tyr = std::max(P->sg[sgi[u]].max_tyr, p.tyr);
        auto &si = sitems[sgi[u]];
        r.item0 = (int)si.size();
        for (int w = 0; w < p.nwork; ++w) si.push_back(u);
        const NlDev &nl = U.nl;
        r.nl = NlArgs<A>{nl.atom_pt_off, nl.pt_gpos, nl.pt_img, nl.pt_atom, nl.cov_off, nl.cov_gpos, nl.ent_pt,
                         (const A *)nl.chi, (const A *)nl.gamma, (const A *)nl.phase, nl.n_atom, nl.n_cov, nl.npm};
        r.beta_off = beta_off;
        r.g_item0 = r.s_item0 = r.r_item0 = 0;
        r.s_cpw = 1;
        r.parts = nullptr;
        r.atom_part = nullptr;
        r.nparts = 0;
        r.part_off = 0;
        if (ngi[u] >= 0) {
            auto &gi = gitems[ngi[u]];
            r.g_item0 = (int)gi.size();
            if (P->ng[ngi[u]].split) {
                // parts of at most kGatherPartPts points of one atom
                std::vector<int> off(nl.n_atom + 1);
                hip_check(hipMemcpy(off.data(), nl.atom_pt_off, off.size() * sizeof(int), hipMemcpyDeviceToHost),
                          "copy");
                part_first[u] = (long)atom_part_h.size();
                range_first[u] = (long)parts_h.size();
                for (int at = 0; at < nl.n_atom; ++at) {
                    atom_part_h.push_back(r.nparts);
                    for (int pb = off[at]; pb < off[at + 1]; pb += kGatherPartPts) {
                        parts_h.push_back(make_int2(pb, std::min(pb + kGatherPartPts, off[at + 1])));
                        ++r.nparts;
                    }
                }
                atom_part_h.push_back(r.nparts);
                r.part_off = part_elems;
                part_elems += (long)r.nparts * U.ncol * nl.npm;
                const int groups = (U.ncol + kGatherPartCols - 1) / kGatherPartCols;
                for (int w = 0; w < r.nparts * groups; ++w) gi.push_back(u);
                auto &ri = ritems[ngi[u]];
                r.r_item0 = (int)ri.size();
                const long nred = (long)U.ncol * nl.n_atom * nl.npm;
                for (long w = 0; w < (nred + 255) / 256; ++w) ri.push_back(u);
            } else {
                // (atom, group of four columns)
                const int cg = (U.ncol + 3) / 4;
                for (int w = 0; w < nl.n_atom * cg; ++w) gi.push_back(u);
            }
            // scatter: column groups of 4 cpw columns, cpw chosen so the
            // groups divide the columns evenly (24 columns: two groups of 12)
            const int sgroups = (U.ncol + 4 * kBatchScatterCpw - 1) / (4 * kBatchScatterCpw);
            r.s_cpw = (U.ncol + 4 * sgroups - 1) / (4 * sgroups);
            auto &ci = citems[ngi[u]];
            r.s_item0 = (int)ci.size();
            for (int w = 0; w < ((nl.n_cov + 63) / 64) * sgroups; ++w) ci.push_back(u);
            beta_off += (long)U.ncol * nl.n_atom * nl.npm;
        }
    }
    if (!parts_h.empty()) {
        P->parts = dev_upload(parts_h, st);
        P->atom_part = dev_upload(atom_part_h, st);
        hip_check(hipMalloc(&P->partial, std::max<size_t>(1, part_elems * sizeof(S))), "hipMalloc");
        for (int u = 0; u < n; ++u)
            if (part_first[u] >= 0) {
                recs[u].parts = P->parts + range_first[u];
                recs[u].atom_part = P->atom_part + part_first[u];
            }
    }
    P->recs = dev_upload(recs, st);
    for (size_t i = 0; i < P->sg.size(); ++i) {
        P->sg[i].nitems = (int)sitems[i].size();
        P->sg[i].items = dev_upload(sitems[i], st);
    }
    for (size_t i = 0; i < P->ng.size(); ++i) {
        P->ng[i].ngitems = (int)gitems[i].size();
        P->ng[i].gitems = dev_upload(gitems[i], st);
        P->ng[i].nsitems = (int)citems[i].size();
        P->ng[i].sitems = dev_upload(citems[i], st);
        P->ng[i].nritems = (int)ritems[i].size();
        P->ng[i].ritems = dev_upload(ritems[i], st);
    }
    hip_check(hipMalloc(&P->beta, std::max<size_t>(1, beta_off * sizeof(S))), "hipMalloc");
    hip_check(hipStreamSynchronize(st), "batch upload");
}

template <class S, int W, int HP>
void launch_stencil_batch(const BatchPlan *P, const BatchPlan::SGroup &G, const void *coef, int ix, int iold,
                          int iout, hipStream_t st)
{
    using A = typename Tr<S>::A;
    const int blocks = ((G.nitems + 7) / 8) * 8;
    const auto *recs = static_cast<const BatchRec<A> *>(P->recs);
    const A *cf = static_cast<const A *>(coef);
    auto go = [&](auto kern, int nt, size_t shm) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(nt), shm, st, recs, G.items, G.nitems, cf, ix, iold, iout);
    };
    if constexpr (W == 256) {
        // long rows: 512 threads, scalar bodies
        const size_t shm = (G.kind == 1 ? 4 : 2) * sizeof(S) * ((size_t)W * (G.max_tyr + 2 * kMaxR) + 512);
        if (G.kind == 0) go(k_stencil_batch<S, kMaxR, batch_rp_orth<S>(), 512, kLongHalo, W, 0>, 512, shm);
        else if constexpr (sizeof(S) <= 8)
            go(k_stencil_batch<S, kMaxR, batch_rp_nonorth<S>(), 512, kLongHalo, W, 1>, 512, shm);
    } else if constexpr (W == kShortNonorthPitch) {
        // short non-orthogonal rows: the tight pitch leaves room for more
        // workgroups per CU (the body holds four planes of this pitch)
        const size_t shm = 4 * sizeof(S) * ((size_t)W * (G.max_tyr + 2 * kMaxR) + G.nt);
        if (G.nt == 192) go(k_stencil_batch<S, kMaxR, batch_rp_nonorth<S>(), 192, 2, W, 1>, 192, shm);
        else go(k_stencil_batch<S, kMaxR, batch_rp_nonorth<S>(), kBatchNT, HP, W, 1>, kBatchNT, shm);
    } else if (G.kind == 2 || G.kind == 3) {
        if constexpr (!Tr<S>::C) {
            // vector kinds: halo slots counted in runs (2 or 3 per thread)
            constexpr int HV = HP == 3 ? 2 : 3;
            const int VW = G.kind == 3 ? 2 : int(16 / sizeof(S));
            const size_t shm = 2 * sizeof(S) * ((size_t)W * (G.max_tyr + 2 * kMaxR) + kBatchNT * VW);
            if (G.kind == 2) go(k_stencil_batch<S, kMaxR, kBatchRuns, kBatchNT, HV, W, 2>, kBatchNT, shm);
            else if constexpr (sizeof(S) == 4)
                go(k_stencil_batch<S, kMaxR, kBatchRuns, kBatchNT, HV, W, 3>, kBatchNT, shm);
        }
    } else if (G.kind == 1 && G.nt == 192) {
        // short non-orthogonal rows, two halo slots per thread
        if constexpr (W == 64) {
            const size_t shm = 4 * sizeof(S) * ((size_t)W * (G.max_tyr + 2 * kMaxR) + 192);
            go(k_stencil_batch<S, kMaxR, batch_rp_nonorth<S>(), 192, 2, W, 1>, 192, shm);
        }
    } else {
        // the non-orthogonal body uses two (plane, g-plane) pairs
        const size_t shm = (G.kind == 1 ? 4 : 2) * sizeof(S) * ((size_t)W * (G.max_tyr + 2 * kMaxR) + kBatchNT);
        if (G.kind == 0) go(k_stencil_batch<S, kMaxR, batch_rp_orth<S>(), kBatchNT, HP, W, 0>, kBatchNT, shm);
        else go(k_stencil_batch<S, kMaxR, batch_rp_nonorth<S>(), kBatchNT, HP, W, 1>, kBatchNT, shm);
    }
}

template <class S, int NPM>
void launch_nl_batch(const BatchPlan *P, const BatchPlan::NGroup &G, const void *coef, int ix, int iout,
                     bool gather, hipStream_t st)
{
    using A = typename Tr<S>::A;
    const auto *recs = static_cast<const BatchRec<A> *>(P->recs);
    if (gather) {
        if (G.ngitems > 0) {
            if (G.split) {
                hipLaunchKernelGGL((k_nl_gather_part_batch<S, NPM>), dim3(G.ngitems), dim3(256), 0, st, recs,
                                   G.gitems, G.ngitems, static_cast<S *>(P->partial), ix);
                hipLaunchKernelGGL((k_nl_gather_reduce_batch<S, NPM>), dim3(G.nritems), dim3(256), 0, st, recs,
                                   G.ritems, G.nritems, static_cast<const S *>(P->partial),
                                   static_cast<S *>(P->beta));
            } else
                hipLaunchKernelGGL((k_nl_gather_batch<S, NPM>), dim3(G.ngitems), dim3(256), 0, st, recs,
                                   G.gitems, G.ngitems, static_cast<S *>(P->beta), ix);
        }
    } else if (G.nsitems > 0) {
        hipLaunchKernelGGL((k_nl_scatter_batch<S, NPM>), dim3(G.nsitems), dim3(256), 0, st, recs, G.sitems,
                           G.nsitems, static_cast<const S *>(P->beta), static_cast<const A *>(coef), iout);
    }
}
}  // namespace

BatchPlan *batch_plan_create(int e, const BatchUnit *units, int n, hipStream_t st)
{
    auto *P = new BatchPlan();
    P->e = e;
    P->n = n;
    try {
        dispatch(e, [&](auto t) { build_batch<decltype(t)>(P, units, n, st); });
    } catch (...) {
        batch_plan_destroy(P);
        throw;
    }
    return P;
}

void batch_plan_destroy(BatchPlan *P)
{
    if (!P) return;
    (void)hipFree(P->recs);
    for (auto &g : P->sg) (void)hipFree(g.items);
    for (auto &g : P->ng) {
        (void)hipFree(g.gitems);
        (void)hipFree(g.sitems);
        (void)hipFree(g.ritems);
    }
    (void)hipFree(P->beta);
    (void)hipFree(P->partial);
    (void)hipFree(P->parts);
    (void)hipFree(P->atom_part);
    (void)hipFree(P->coef);
    delete P;
}

void batch_set_coef(BatchPlan *P, const double *coef, int steps, hipStream_t st)
{
    const size_t cnt = (size_t)steps * P->n * 3;
    const bool single = elem_single(P->e);
    const size_t bytes = cnt * (single ? sizeof(float) : sizeof(double));
    if (steps > P->coef_steps) {
        (void)hipFree(P->coef);
        P->coef = nullptr;
        hip_check(hipMalloc(&P->coef, bytes), "hipMalloc");
        P->coef_steps = steps;
    }
    if (single) {
        std::vector<float> f(coef, coef + cnt);
        hip_check(hipMemcpyAsync(P->coef, f.data(), bytes, hipMemcpyHostToDevice, st), "copy");
        hip_check(hipStreamSynchronize(st), "copy");
    } else {
        hip_check(hipMemcpyAsync(P->coef, coef, bytes, hipMemcpyHostToDevice, st), "copy");
    }
}

void batch_step(BatchPlan *P, int step, int ix, int iold, int iout, hipStream_t st)
{
    if (step >= P->coef_steps) fail(1, "batch step beyond the uploaded filter coefficients");
    dispatch(P->e, [&](auto t) {
        using S = decltype(t);
        using A = typename Tr<S>::A;
        const void *cf = static_cast<const A *>(P->coef) + (size_t)step * P->n * 3;
        for (const auto &G : P->ng) {
            if (G.npm == 8) launch_nl_batch<S, 8>(P, G, cf, ix, iout, true, st);
            else if (G.npm == 16) launch_nl_batch<S, 16>(P, G, cf, ix, iout, true, st);
            else launch_nl_batch<S, 32>(P, G, cf, ix, iout, true, st);
        }
        for (const auto &G : P->sg) {
            if (G.W == kShortNonorthPitch) {
                if (G.small_halo) launch_stencil_batch<S, kShortNonorthPitch, 3>(P, G, cf, ix, iold, iout, st);
                else launch_stencil_batch<S, kShortNonorthPitch, 6>(P, G, cf, ix, iold, iout, st);
            } else if (G.W == 64) {
                if (G.small_halo) launch_stencil_batch<S, 64, 3>(P, G, cf, ix, iold, iout, st);
                else launch_stencil_batch<S, 64, 6>(P, G, cf, ix, iold, iout, st);
            } else if (G.W == 256) {
                launch_stencil_batch<S, 256, kLongHalo>(P, G, cf, ix, iold, iout, st);
            } else {
                if (G.small_halo) launch_stencil_batch<S, 128, 3>(P, G, cf, ix, iold, iout, st);
                else launch_stencil_batch<S, 128, 6>(P, G, cf, ix, iold, iout, st);
            }
        }
        for (const auto &G : P->ng) {
            if (G.npm == 8) launch_nl_batch<S, 8>(P, G, cf, ix, iout, false, st);
            else if (G.npm == 16) launch_nl_batch<S, 16>(P, G, cf, ix, iout, false, st);
            else launch_nl_batch<S, 32>(P, G, cf, ix, iout, false, st);
        }
    });
}

int batch_launches_per_step(const BatchPlan *P)
{
    int n = (int)(P->sg.size() + 2 * P->ng.size());
    for (const auto &G : P->ng) n += G.split && G.ngitems > 0;   // the part sums' reduction
    return n;
}

void gram_reduce(int e, const void *work, int nsplit, int n, double alpha, void *C, hipStream_t st)
{
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        const long nn = (long)n * n;
        hipLaunchKernelGGL((k_gram_reduce<S>), dim3((nn + 255) / 256), dim3(256), 0, st, (const S *)work, nsplit, n,
                           alpha, (typename Wide<S>::T *)C);
    });
}


void hermitian_part(int e, void *H, int n, hipStream_t st)
{
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_hermitian_part<S>), dim3((n + 63) / 64, n), dim3(64), 0, st, (S *)H, n);
    });
}

void zero_lower(int e, void *H, int n, hipStream_t st)
{
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_zero_lower<S>), dim3((n + 63) / 64, n), dim3(64), 0, st, (S *)H, n);
    });
}

int dot_partials(int e, const void *x, const void *y, long n, double *partials, hipStream_t st)
{
    const int nb = (int)std::min<long>(kPartialBlocks, std::max<long>(1, (n + 255) / 256));
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_dot<S>), dim3(nb), dim3(256), 0, st, (const S *)x, (const S *)y, n, partials);
    });
    return nb;
}

int lanczos_update(int e, void *y, void *x, void *z, double a, double b, long n, double *partials,
                   hipStream_t st)
{
    const int nb = (int)std::min<long>(kPartialBlocks, std::max<long>(1, (n + 255) / 256));
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_lanczos_update<S>), dim3(nb), dim3(256), 0, st, (S *)y, (S *)x, (S *)z, a, b, n,
                           partials);
    });
    return nb;
}

void scale_copy(int e, const void *src, void *dst, double s, long n, hipStream_t st)
{
    const int nb = (int)std::min<long>(2048, std::max<long>(1, (n + 255) / 256));
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_scale_copy<S>), dim3(nb), dim3(256), 0, st, (const S *)src, (S *)dst, s, n);
    });
}

void copy_cols(int e, const void *src, long lds, void *dst, long ldd, long nd, int ncol, hipStream_t st)
{
    if (ncol <= 0 || nd <= 0) return;
    const int nb = (int)std::min<long>(1024, std::max<long>(1, (nd + 255) / 256));
    if (e == B16) {
        hipLaunchKernelGGL((k_copy_cols<unsigned short>), dim3(nb, ncol), dim3(256), 0, st,
                           (const unsigned short *)src, lds, (unsigned short *)dst, ldd, nd);
        return;
    }
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_copy_cols<S>), dim3(nb, ncol), dim3(256), 0, st, (const S *)src, lds, (S *)dst,
                           ldd, nd);
    });
}

void stream_probe(const void *x, const void *o, const void *v, void *y, long n16, long nv16, int blocks,
                  hipStream_t st)
{
    hipLaunchKernelGGL(k_stream3, dim3(blocks), dim3(256), 0, st, (const double2 *)x, (const double2 *)o,
                       (const double2 *)v, (double2 *)y, n16, nv16);
}

void convert_cols(int e_src, const void *src, long lds, int e_dst, void *dst, long ldd, long nd, int ncol,
                  hipStream_t st)
{
    if (ncol <= 0 || nd <= 0) return;
    if (elem_complex(e_src) != elem_complex(e_dst)) fail(1, "conversion between real and complex blocks");
    const int m = elem_complex(e_src) ? 2 : 1;      // scalars per element
    const long n = nd * m;
    const int nb = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
    if (e_src == B16 || e_dst == B16) {
        if (e_src == B16 && e_dst == B16)
            hipLaunchKernelGGL((k_copy_cols<unsigned short>), dim3(nb, ncol), dim3(256), 0, st,
                               (const unsigned short *)src, lds, (unsigned short *)dst, ldd, nd);
        else if (e_dst == B16 && e_src == F64)
            hipLaunchKernelGGL((k_to_bf16_cols<double>), dim3(nb, ncol), dim3(256), 0, st, (const double *)src, lds,
                               (bf16 *)dst, ldd, nd);
        else if (e_dst == B16)
            hipLaunchKernelGGL((k_to_bf16_cols<float>), dim3(nb, ncol), dim3(256), 0, st, (const float *)src, lds,
                               (bf16 *)dst, ldd, nd);
        else if (e_dst == F64)
            hipLaunchKernelGGL((k_from_bf16_cols<double>), dim3(nb, ncol), dim3(256), 0, st, (const bf16 *)src, lds,
                               (double *)dst, ldd, nd);
        else
            hipLaunchKernelGGL((k_from_bf16_cols<float>), dim3(nb, ncol), dim3(256), 0, st, (const bf16 *)src, lds,
                               (float *)dst, ldd, nd);
        return;
    }
    const bool sd = !elem_single(e_src), dd = !elem_single(e_dst);
    if (sd && !dd)
        hipLaunchKernelGGL((k_convert_cols<float, double>), dim3(nb, ncol), dim3(256), 0, st, (const double *)src,
                           lds * m, (float *)dst, ldd * m, n);
    else if (!sd && dd)
        hipLaunchKernelGGL((k_convert_cols<double, float>), dim3(nb, ncol), dim3(256), 0, st, (const float *)src,
                           lds * m, (double *)dst, ldd * m, n);
    else
        copy_cols(e_src, src, lds, dst, ldd, nd, ncol, st);
}

}  // namespace gpu
}  // namespace lynx
