// HIP kernels for gfx950 (MI355X): fused FD stencil + local potential +
// Chebyshev recurrence, nonlocal projector gather/scatter, subspace GEMMs and
// small helpers.  Launch wrappers at the bottom.
#include "kernels.hpp"
#include "common.hpp"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

namespace lynx {
namespace gpu {
namespace {

// ------------------------------------------------------------ element ops --
template <class S> struct Tr;
template <> struct Tr<double> { using A = double; static constexpr bool C = false; };
template <> struct Tr<float> { using A = float; static constexpr bool C = false; };
template <> struct Tr<double2> { using A = double; static constexpr bool C = true; };
template <> struct Tr<float2> { using A = float; static constexpr bool C = true; };

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
    A c[3][kMaxR + 1];
    A a, b, s;
    A ph[3][2][2];
};

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

template <class S, int R, int RP, int NT, int HP, int W>
__global__ void __launch_bounds__(NT) k_stencil_step(StepArgs<typename Tr<S>::A> p)
{
    using A = typename Tr<S>::A;
    constexpr bool C = Tr<S>::C;
    constexpr int NR = 2 * R + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    S *lds = reinterpret_cast<S *>(smem_raw);

    // W: LDS row pitch, a compile-time constant >= nx + 2R so that every
    // neighbour read is one base address plus an immediate offset
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    // each LDS plane buffer ends with one dummy slot that absorbs the writes
    // of idle lanes and of elements without a ghost copy (no branches)
    const int dummy = W * (p.tyr + 2 * R);
    const int plane = dummy + 1;

    const int per = gridDim.x >> 3;
    const int w = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (w >= p.nwork) return;
    const int band = w % p.nbands;
    const int rest = w / p.nbands;
    const int zc = rest % p.nzch;
    const int col = rest / p.nzch;
    const int y0 = band * p.tyr;
    const int rows = min(p.tyr, ny - y0);
    const int z0 = zc * p.zch;
    const int z1 = min(z0 + p.zch, nz);
    const int nxny = nx * ny;

    const S *__restrict__ X = reinterpret_cast<const S *>(p.x) + (long)col * p.ldx;
    const S *__restrict__ O = reinterpret_cast<const S *>(p.old) + (long)col * p.ldo;
    S *__restrict__ Y = reinterpret_cast<S *>(p.out) + (long)col * p.ldout;
    const bool has_old = p.b != A(0);

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
            const S hval = fm(cur.v[r] + p.s, ring[r][R], scl(A(-0.5), lap));
            S res = scl(p.a, hval);
            if (has_old) res = fm(p.b, cur.old[r], res);
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

// --------------------------------------------------------- nonlocal gather --
// One workgroup per physical atom and 4 columns (one wave per column).  Lanes
// stride over the atom's sphere points (all images); each lane accumulates
// npm projections, then a wave reduction and one store per projector.
template <class A>
struct NlArgs {
    const int *atom_pt_off, *pt_gpos, *pt_img, *pt_atom, *cov_off, *cov_gpos, *ent_pt;
    const A *chi, *gamma, *phase;
    int n_atom, n_cov, npm;
};

template <class S, int NPM>
__global__ void __launch_bounds__(256) k_nl_gather(NlArgs<typename Tr<S>::A> p, const S *__restrict__ X,
                                                   long ldx, S *__restrict__ beta, int ncol)
{
    using A = typename Tr<S>::A;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int at = blockIdx.x;
    const int n = blockIdx.y * 4 + wave;
    if (n >= ncol) return;
    const int pb = p.atom_pt_off[at], pe = p.atom_pt_off[at + 1];
    const S *Xn = X + (long)n * ldx;
    S acc[NPM];
#pragma unroll
    for (int j = 0; j < NPM; ++j) acc[j] = zero<S>();
    for (int e = pb + lane; e < pe; e += 64) {
        S v = Xn[p.pt_gpos[e]];
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
    S *bn = beta + (long)n * p.n_atom * NPM + (long)at * NPM;
#pragma unroll
    for (int j = 0; j < NPM; ++j)
        if (lane == j) bn[j] = scl(p.gamma[at * NPM + j], acc[j]);
}

// -------------------------------------------------------- nonlocal scatter --
// Owner-computes: one lane per covered grid point, one wave per column; sums
// chi . beta over every (image, point) entry at that grid point, so no two
// lanes write the same address and the result is deterministic.
template <class S, int NPM>
__global__ void __launch_bounds__(256) k_nl_scatter(NlArgs<typename Tr<S>::A> p,
                                                    const S *__restrict__ beta, S *__restrict__ Y,
                                                    long ldy, int ncol, typename Tr<S>::A a)
{
    using A = typename Tr<S>::A;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int u = blockIdx.x * 64 + lane;
    const int n = blockIdx.y * 4 + wave;
    if (u >= p.n_cov || n >= ncol) return;
    const S *bn = beta + (long)n * p.n_atom * NPM;
    S acc = zero<S>();
    for (int e = p.cov_off[u]; e < p.cov_off[u + 1]; ++e) {
        const int pt = p.ent_pt[e];
        const A *ch = p.chi + (long)pt * NPM;
        const S *bt = bn + (long)p.pt_atom[pt] * NPM;
        S s = zero<S>();
#pragma unroll
        for (int j = 0; j < NPM; ++j) s = fm(ch[j], bt[j], s);
        if constexpr (Tr<S>::C) {
            const int im = p.pt_img[pt];
            s = cmul(s, p.phase[2 * im], -p.phase[2 * im + 1]);
        }
        acc = add(acc, s);
    }
    const long gi = (long)n * ldy + p.cov_gpos[u];
    Y[gi] = fm(a, acc, Y[gi]);
}

// ---------------------------------------------------------------- GEMMs --
// Gram: C = A^H B over a split K range; tiles of 64x64 outputs per
// workgroup, 16 K-rows staged in LDS per step; partial tiles go to work and a
// second kernel sums them in a fixed order (deterministic).
constexpr int GT = 64, GK = 16;

template <class S>
__global__ void __launch_bounds__(256) k_gram_partial(const S *__restrict__ A, long lda,
                                                      const S *__restrict__ B, long ldb, int nd,
                                                      int n, int kchunk, S *__restrict__ work)
{
    __shared__ S As[GK][GT + 1];
    __shared__ S Bs[GK][GT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.x * GT, j0 = blockIdx.y * GT;
    const long kb = (long)blockIdx.z * kchunk;
    const long ke = min((long)nd, kb + kchunk);
    S acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zero<S>();
    for (long k0 = kb; k0 < ke; k0 += GK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int t = threadIdx.x + 256 * r;
            const int kk = t & (GK - 1), ii = t >> 4;
            const long k = k0 + kk;
            const int ia = i0 + ii, jb = j0 + ii;
            As[kk][ii] = (k < ke && ia < n) ? A[(long)ia * lda + k] : zero<S>();
            Bs[kk][ii] = (k < ke && jb < n) ? B[(long)jb * ldb + k] : zero<S>();
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            S av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[kk][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = cfma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
    S *wz = work + (long)blockIdx.z * n * n;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            if (i < n && j < n) wz[(long)j * n + i] = acc[a][b];
        }
}

template <class S>
__global__ void k_gram_reduce(const S *__restrict__ work, int nsplit, int n, double alpha, S *C)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nn = (long)n * n;
    if (idx >= nn) return;
    S s = zero<S>();
    for (int z = 0; z < nsplit; ++z) s = add(s, work[(long)z * nn + idx]);
    C[idx] = scl((typename Tr<S>::A)alpha, s);
}

// Rotation: C = A Q for tall A; 64 rows x 64 columns per workgroup.
template <class S>
__global__ void __launch_bounds__(256) k_rotate(const S *__restrict__ A, long lda,
                                                const S *__restrict__ Q, int n, S *__restrict__ C,
                                                long ldc, int nd)
{
    __shared__ S As[GK][GT + 1];
    __shared__ S Qs[GK][GT + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long k0 = (long)blockIdx.x * GT;
    const int j0 = blockIdx.y * GT;
    S acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = zero<S>();
    for (int i0 = 0; i0 < n; i0 += GK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int t = threadIdx.x + 256 * r;
            // A tile: 64 rows (k) x 16 (i), consecutive threads -> consecutive rows
            const int kk = t & 63, ii = t >> 6;
            const long k = k0 + kk;
            const int i = i0 + ii;
            As[ii][kk] = (k < nd && i < n) ? A[(long)i * lda + k] : zero<S>();
            // Q tile: 16 (i) x 64 (j)
            const int qi = t & 15, qj = t >> 4;
            Qs[qi][qj] = (i0 + qi < n && j0 + qj < n) ? Q[(long)(j0 + qj) * n + i0 + qi] : zero<S>();
        }
        __syncthreads();
#pragma unroll
        for (int ii = 0; ii < GK; ++ii) {
            S av[4], qv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[ii][tx + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) qv[b] = Qs[ii][ty + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = mfma_(av[a], qv[b], acc[a][b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const long k = k0 + tx + 16 * a;
            const int j = j0 + ty + 16 * b;
            if (k < nd && j < n) C[(long)j * ldc + k] = acc[a][b];
        }
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

// ---------------------------------------------------------- dispatch helpers --
template <class F>
void dispatch(int e, F &&f)
{
    switch (e) {
        case F64: f(double{}); break;
        case C128: f(double2{}); break;
        case F32: f(float{}); break;
        default: f(float2{}); break;
    }
}

template <class S> constexpr int base_rp() { return int(32 / sizeof(S)); }

template <class S, int RP, int NT>
void launch_step_rp(const StencilGeom &g, const void *x, long ldx, const void *old, long ldo,
                    void *out, long ldout, const void *veff, int ncol, double a, double b, double s,
                    int nzch, hipStream_t st)
{
    using A = typename Tr<S>::A;
    constexpr int HS = NT >= 512 ? 2 : 3;      // halo slots: rows up to 64 (256) or 85 (512) points
    constexpr int HB = NT >= 512 ? 3 : 6;      // rows up to 128 points
    StepArgs<A> p;
    p.x = x; p.old = old; p.out = out; p.veff = static_cast<const A *>(veff);
    p.ldx = ldx; p.ldo = ldo; p.ldout = ldout;
    p.nx = g.nx; p.ny = g.ny; p.nz = g.nz; p.ncol = ncol;
    const long pts = (long)NT * RP;
    if (g.nx > pts || 2 * kMaxR * g.nx > NT * HB || g.nx + 2 * kMaxR > 128)
        fail(8, "grid rows longer than the fused stencil kernel's tile are not supported");
    int nb = (int)(((long)g.nx * g.ny + pts - 1) / pts);
    nb = std::max(1, std::min(nb, g.ny));
    p.tyr = (g.ny + nb - 1) / nb;
    while ((long)p.tyr * g.nx > pts) { ++nb; p.tyr = (g.ny + nb - 1) / nb; }
    p.nbands = (g.ny + p.tyr - 1) / p.tyr;
    nzch = std::max(1, std::min(nzch, g.nz));
    p.zch = (g.nz + nzch - 1) / nzch;
    p.nzch = (g.nz + p.zch - 1) / p.zch;
    p.nwork = p.nbands * p.nzch * ncol;
    for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k <= kMaxR; ++k) p.c[ax][k] = (A)g.c[ax][k];
    p.a = (A)a; p.b = (A)b; p.s = (A)s;
    for (int ax = 0; ax < 3; ++ax)
        for (int sd = 0; sd < 2; ++sd)
            for (int q = 0; q < 2; ++q) p.ph[ax][sd][q] = (A)g.phase[ax][sd][q];
    const int blocks = ((p.nwork + 7) / 8) * 8;
    const int pitch = g.nx + 2 * kMaxR <= 64 ? 64 : 128;
    const size_t shm = 2 * sizeof(S) * ((size_t)pitch * (p.tyr + 2 * kMaxR) + 1);
    const bool small_halo = 2 * kMaxR * g.nx <= NT * HS;
    if (pitch == 64) {
        if (small_halo) hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HS, 64>), dim3(blocks), dim3(NT), shm, st, p);
        else hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HB, 64>), dim3(blocks), dim3(NT), shm, st, p);
    } else {
        if (small_halo) hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HS, 128>), dim3(blocks), dim3(NT), shm, st, p);
        else hipLaunchKernelGGL((k_stencil_step<S, kMaxR, RP, NT, HB, 128>), dim3(blocks), dim3(NT), shm, st, p);
    }
}

}  // namespace

// ------------------------------------------------------------ host wrappers --
int stencil_points_per_thread(int e, const StepCfg &cfg)
{
    int rp = 0;
    dispatch(e, [&](auto t) { rp = base_rp<decltype(t)>() / std::max(1, cfg.rp_div); });
    return rp;
}

void stencil_step(int e, const StencilGeom &g, const void *x, long ldx, const void *old, long ldo,
                  void *out, long ldout, const void *veff, int ncol, double a, double b, double s,
                  const StepCfg &cfg, hipStream_t st)
{
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        constexpr int B = base_rp<S>();
        if (cfg.nt >= 512) {
            if (cfg.rp_div >= 2)
                launch_step_rp<S, B / 2, 512>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st);
            else
                launch_step_rp<S, B, 512>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st);
        } else {
            if (cfg.rp_div >= 2)
                launch_step_rp<S, B / 2, 256>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st);
            else
                launch_step_rp<S, B, 256>(g, x, ldx, old, ldo, out, ldout, veff, ncol, a, b, s, cfg.nzch, st);
        }
    });
}

template <class S, int NPM>
static void gather_npm(const NlDev &nl, const void *x, long ldx, void *beta, int ncol, hipStream_t st)
{
    using A = typename Tr<S>::A;
    NlArgs<A> p{nl.atom_pt_off, nl.pt_gpos, nl.pt_img, nl.pt_atom, nl.cov_off, nl.cov_gpos, nl.ent_pt,
                (const A *)nl.chi, (const A *)nl.gamma, (const A *)nl.phase, nl.n_atom, nl.n_cov, nl.npm};
    dim3 grid(nl.n_atom, (ncol + 3) / 4);
    hipLaunchKernelGGL((k_nl_gather<S, NPM>), grid, dim3(256), 0, st, p, (const S *)x, ldx, (S *)beta, ncol);
}

template <class S, int NPM>
static void scatter_npm(const NlDev &nl, const void *beta, void *y, long ldy, int ncol, double a,
                        hipStream_t st)
{
    using A = typename Tr<S>::A;
    NlArgs<A> p{nl.atom_pt_off, nl.pt_gpos, nl.pt_img, nl.pt_atom, nl.cov_off, nl.cov_gpos, nl.ent_pt,
                (const A *)nl.chi, (const A *)nl.gamma, (const A *)nl.phase, nl.n_atom, nl.n_cov, nl.npm};
    dim3 grid((nl.n_cov + 63) / 64, (ncol + 3) / 4);
    hipLaunchKernelGGL((k_nl_scatter<S, NPM>), grid, dim3(256), 0, st, p, (const S *)beta, (S *)y, ldy,
                       ncol, (A)a);
}

void nl_gather(int e, const NlDev &nl, const void *x, long ldx, void *beta, int ncol, hipStream_t st)
{
    if (nl.n_atom == 0 || nl.n_pts == 0) return;
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        if (nl.npm == 8) gather_npm<S, 8>(nl, x, ldx, beta, ncol, st);
        else if (nl.npm == 16) gather_npm<S, 16>(nl, x, ldx, beta, ncol, st);
        else gather_npm<S, 32>(nl, x, ldx, beta, ncol, st);
    });
}

void nl_scatter(int e, const NlDev &nl, const void *beta, void *y, long ldy, int ncol, double a,
                hipStream_t st)
{
    if (nl.n_cov == 0) return;
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        if (nl.npm == 8) scatter_npm<S, 8>(nl, beta, y, ldy, ncol, a, st);
        else if (nl.npm == 16) scatter_npm<S, 16>(nl, beta, y, ldy, ncol, a, st);
        else scatter_npm<S, 32>(nl, beta, y, ldy, ncol, a, st);
    });
}

static int gram_splits(int nd, int n)
{
    const int tiles = ((n + GT - 1) / GT) * ((n + GT - 1) / GT);
    int ks = std::max(1, 1024 / tiles);
    const int maxks = std::max(1, nd / 512);
    return std::min(ks, maxks);
}

size_t gram_work_bytes(int e, int nd, int n)
{
    return (size_t)gram_splits(nd, n) * n * n * elem_size(e);
}

void gram(int e, const void *A, long lda, const void *B, long ldb, int nd, int n, double alpha, void *C,
          void *work, hipStream_t st)
{
    const int ks = gram_splits(nd, n);
    int kchunk = (nd + ks - 1) / ks;
    kchunk = (kchunk + GK - 1) / GK * GK;
    const int nsplit = (nd + kchunk - 1) / kchunk;
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        dim3 grid((n + GT - 1) / GT, (n + GT - 1) / GT, nsplit);
        hipLaunchKernelGGL((k_gram_partial<S>), grid, dim3(256), 0, st, (const S *)A, lda, (const S *)B, ldb,
                           nd, n, kchunk, (S *)work);
        const long nn = (long)n * n;
        hipLaunchKernelGGL((k_gram_reduce<S>), dim3((nn + 255) / 256), dim3(256), 0, st, (const S *)work,
                           nsplit, n, alpha, (S *)C);
    });
}

void rotate(int e, const void *A, long lda, const void *Q, int n, void *C, long ldc, int nd, hipStream_t st)
{
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        dim3 grid((nd + GT - 1) / GT, (n + GT - 1) / GT);
        hipLaunchKernelGGL((k_rotate<S>), grid, dim3(256), 0, st, (const S *)A, lda, (const S *)Q, n, (S *)C,
                           ldc, nd);
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
    dispatch(e, [&](auto t) {
        using S = decltype(t);
        hipLaunchKernelGGL((k_copy_cols<S>), dim3(nb, ncol), dim3(256), 0, st, (const S *)src, lds, (S *)dst,
                           ldd, nd);
    });
}

}  // namespace gpu
}  // namespace lynx
