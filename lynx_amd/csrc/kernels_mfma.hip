// Split-BF16x3 Gram products on the CDNA4 matrix cores.
//
// C = A^T B for real single-precision blocks A, B (K x N, column-major,
// K = Nd grid points, N orbitals).  Each operand is split into a bfloat16
// head and a bfloat16 tail, x = hi + lo (hi = bf16(x), lo = bf16(x - hi)),
// and the product is hi.hi + hi.lo + lo.hi, three v_mfma_f32_32x32x16_bf16
// per 16-deep step with FP32 accumulation (the lo.lo term, 2^-16 below, is
// dropped).  K is split into chunks; each chunk's FP32 partial product is
// written out and the chunks are summed in FP64 in a fixed order by
// gram_reduce, so no FP32 sum runs over more than one chunk.
//
// Work decomposition: a workgroup of five waves owns a 160 x 160 tile of C
// and one K chunk (for A^T A only the tiles on and above the diagonal,
// written to both halves).  Per 32-deep block it stages the tile's 160
// columns of A and of B as (hi, lo) bfloat16 pairs in LDS, column-major with
// k contiguous, the split done once per element by the loading thread, and
// the next block's loads are in flight while the current block's products
// run.  Wave w computes rows 32w..32w+31 of the tile against all five
// 32-column blocks: its A fragment is read once per step and reused by five
// products.
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (checked on gfx950): lane l
// supplies A[l % 32][8 (l / 32) + i] and B[8 (l / 32) + i][l % 32],
// i = 0..7, and holds D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32], j = 0..15.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace lynx {
namespace gpu {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kTile = 160;            // C tile edge
constexpr int kWaves = kTile / 32;    // one 32-row strip per wave
constexpr int kThreads = 64 * kWaves;
constexpr int kKB = 32;               // K rows per LDS block
constexpr int kPitch = kKB + 8;       // bf16 per LDS column (padded against bank conflicts)

// round to nearest even
__device__ inline unsigned short bf16_bits(float x)
{
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_value(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// This thread's share of a 32-deep block of a 160-column tile: 16
// consecutive K rows (from k0 + koff, masked by kend and the column count)
// of column c = t / 2.
struct Slab {
    float v[16];
};

__device__ inline void load_slab(const float *__restrict__ X, long ldx, long k0, long kend, int col0, int N, int t,
                                 Slab &s)
{
    const int c = t >> 1;
    const int koff = (t & 1) * 16;
    const int gc = col0 + c;
    const long r0 = k0 + koff;
    if (gc < N && r0 + 16 <= kend) {
        const float4 *p = reinterpret_cast<const float4 *>(X + (long)gc * ldx + r0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 f = p[q];
            s.v[4 * q] = f.x; s.v[4 * q + 1] = f.y; s.v[4 * q + 2] = f.z; s.v[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) s.v[i] = (gc < N && r0 + i < kend) ? X[(long)gc * ldx + r0 + i] : 0.0f;
    }
}

// Splits the slab into bfloat16 hi and lo and stores both into LDS.
__device__ inline void store_slab(const Slab &s, unsigned short *hi, unsigned short *lo, int t)
{
    const int c = t >> 1;
    const int koff = (t & 1) * 16;
    unsigned short h[16], l[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        h[i] = bf16_bits(s.v[i]);
        l[i] = bf16_bits(s.v[i] - bf16_value(h[i]));
    }
    uint4 *dh = reinterpret_cast<uint4 *>(hi + c * kPitch + koff);
    uint4 *dl = reinterpret_cast<uint4 *>(lo + c * kPitch + koff);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint4 a, b;
        a.x = h[8 * q] | ((unsigned)h[8 * q + 1] << 16);
        a.y = h[8 * q + 2] | ((unsigned)h[8 * q + 3] << 16);
        a.z = h[8 * q + 4] | ((unsigned)h[8 * q + 5] << 16);
        a.w = h[8 * q + 6] | ((unsigned)h[8 * q + 7] << 16);
        b.x = l[8 * q] | ((unsigned)l[8 * q + 1] << 16);
        b.y = l[8 * q + 2] | ((unsigned)l[8 * q + 3] << 16);
        b.z = l[8 * q + 4] | ((unsigned)l[8 * q + 5] << 16);
        b.w = l[8 * q + 6] | ((unsigned)l[8 * q + 7] << 16);
        dh[q] = a;
        dl[q] = b;
    }
}

__device__ inline bf16x8 frag(const unsigned short *s, int col, int k)
{
    return *reinterpret_cast<const bf16x8 *>(s + col * kPitch + k);
}

// eight bf16 of row r (pitch 16 + 8) from k on, for the rotation's LDS rows
__device__ inline bf16x8 frag_p(const unsigned short *s, int r, int k)
{
    return *reinterpret_cast<const bf16x8 *>(s + r * (16 + 8) + k);
}

// Tile pair of workgroup b: all tiles_n^2 pairs, or for A^T A only the
// pairs tm <= tn (row-major over the upper triangle); the lower tiles of
// A^T A are written as the transposes of the upper ones.
__device__ inline void tile_of(int b, int tiles_n, int same, int &tm, int &tn)
{
    if (!same) {
        tm = b / tiles_n;
        tn = b % tiles_n;
        return;
    }
    tm = 0;
    int rem = b;
    while (rem >= tiles_n - tm) {
        rem -= tiles_n - tm;
        ++tm;
    }
    tn = tm + rem;
}

// One-dimensional grid of 8 * per_xcd workgroups over items = pairs x
// chunks (item = chunk * pairs + pair; chunk s covers rows
// [s * kch, min(K, (s + 1) * kch))).  Workgroup b runs on XCD b mod 8 and
// takes item (b mod 8) * per_xcd + b / 8, so consecutive items, the tiles
// of one chunk, share an XCD's L2; workgroups past the last item return.
// The next block's global loads are issued before the current block's
// products, so they overlap.
__global__ void __launch_bounds__(kThreads) k_gram_bf16x3(const float *__restrict__ A, long lda,
                                                         const float *__restrict__ B, long ldb, long K, int N,
                                                         long kch, int tiles_n, int same, int pairs, int items,
                                                         int per_xcd, float *__restrict__ part)
{
    const int item = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (item >= items) return;
    const int chunk = item / pairs, pair = item % pairs;
    __shared__ __attribute__((aligned(16))) unsigned short sAh[kTile * kPitch], sAl[kTile * kPitch];
    __shared__ __attribute__((aligned(16))) unsigned short sBh[kTile * kPitch], sBl[kTile * kPitch];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    int tm, tn;
    tile_of(pair, tiles_n, same, tm, tn);
    const int m0 = tm * kTile, n0 = tn * kTile;
    const long kbeg = (long)chunk * kch, kend = std::min(K, kbeg + kch);
    const bool one = same && tm == tn;          // B's tile is A's
    const unsigned short *bh = one ? sAh : sBh;
    const unsigned short *bl = one ? sAl : sBl;
    const float *Bsrc = same ? A : B;
    const long ldbs = same ? lda : ldb;
    floatx16 acc[kWaves];
#pragma unroll
    for (int j = 0; j < kWaves; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[j][r] = 0.0f;
    const int row = w * 32 + (l & 31), kl = 8 * (l >> 5);
    Slab sa, sb;
    if (kbeg < kend) {
        load_slab(A, lda, kbeg, kend, m0, N, t, sa);
        if (!one) load_slab(Bsrc, ldbs, kbeg, kend, n0, N, t, sb);
    }
    for (long k0 = kbeg; k0 < kend; k0 += kKB) {
        store_slab(sa, sAh, sAl, t);
        if (!one) store_slab(sb, sBh, sBl, t);
        __syncthreads();
        if (k0 + kKB < kend) {
            load_slab(A, lda, k0 + kKB, kend, m0, N, t, sa);
            if (!one) load_slab(Bsrc, ldbs, k0 + kKB, kend, n0, N, t, sb);
        }
#pragma unroll
        for (int ks = 0; ks < kKB; ks += 16) {
            const bf16x8 ah = frag(sAh, row, ks + kl), al = frag(sAl, row, ks + kl);
#pragma unroll
            for (int j = 0; j < kWaves; ++j) {
                const bf16x8 b_h = frag(bh, j * 32 + (l & 31), ks + kl);
                const bf16x8 b_l = frag(bl, j * 32 + (l & 31), ks + kl);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b_h, acc[j], 0, 0, 0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, b_l, acc[j], 0, 0, 0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b_h, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float *P = part + (size_t)chunk * N * N;
    const bool mirror = same && tm != tn;
#pragma unroll
    for (int j = 0; j < kWaves; ++j) {
        const int n = n0 + j * 32 + (l & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int m = m0 + w * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
            if (m < N && n < N) {
                P[m + (size_t)n * N] = acc[j][r];
                if (mirror) P[n + (size_t)m * N] = acc[j][r];
            }
        }
    }
}

// ------------------------------------------------------------- FP64 Gram --
// C = A^T B for FP64 blocks on v_mfma_f64_16x16x4f64 (lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[l / 16 + 4 v][l % 16],
// v = 0..3; checked on gfx950).  Same decomposition as k_gram_bf16x3: one
// workgroup per (tile pair, K chunk); for A^T A, or a product whose
// Hermitian part is wanted (upper), only the tiles on and above the
// diagonal, their upper entries written to both halves; one FP64 partial
// product per chunk summed in a fixed order by gram_reduce.  A workgroup of four waves owns a
// 128 x 128 tile; wave w takes the 64 x 64 quarter (w / 2, w % 2), 4 x 4
// products per 4-deep step, each fragment read once from LDS and used four
// times.  Per 32-deep block the tile's 128 columns of A and of B are staged
// in LDS with k contiguous (33 doubles per column, so the 16 lanes reading
// one k of 16 columns hit distinct banks); the next block's loads are in
// flight while the current block's products run.
constexpr int kDT = 128;          // C tile edge
constexpr int kDKB = 32;          // K rows per LDS block
constexpr int kDP = kDKB + 1;     // doubles per LDS column
typedef double double4v __attribute__((ext_vector_type(4)));

// This thread's share of a 32-deep block of a 128-column tile: 16
// consecutive K rows (from k0 + koff, masked by kend and the column count)
// of column c = t / 2.
struct DSlab {
    double v[kDKB / 2];
};

__device__ inline void load_dslab(const double *__restrict__ X, long ldx, long k0, long kend, int col0, int N, int t,
                                  DSlab &s)
{
    const int c = t >> 1;
    const int koff = (t & 1) * (kDKB / 2);
    const int gc = col0 + c;
    const long r0 = k0 + koff;
    if (gc < N && r0 + kDKB / 2 <= kend) {
        const double2 *p = reinterpret_cast<const double2 *>(X + (long)gc * ldx + r0);
#pragma unroll
        for (int q = 0; q < kDKB / 4; ++q) {
            const double2 f = p[q];
            s.v[2 * q] = f.x;
            s.v[2 * q + 1] = f.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kDKB / 2; ++i) s.v[i] = (gc < N && r0 + i < kend) ? X[(long)gc * ldx + r0 + i] : 0.0;
    }
}

__device__ inline void store_dslab(const DSlab &s, double *lds, int t)
{
    double *d = lds + (t >> 1) * kDP + (t & 1) * (kDKB / 2);
#pragma unroll
    for (int i = 0; i < kDKB / 2; ++i) d[i] = s.v[i];
}

__global__ void __launch_bounds__(256) k_gram_f64(const double *__restrict__ A, long lda,
                                                  const double *__restrict__ B, long ldb, long K, int N, long kch,
                                                  int tiles_n, int same, int upper, int pairs, int items,
                                                  int per_xcd, double *__restrict__ part)
{
    const int item = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (item >= items) return;
    const int chunk = item / pairs, pair = item % pairs;
    __shared__ double sA[kDT * kDP], sB[kDT * kDP];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    int tm, tn;
    tile_of(pair, tiles_n, upper, tm, tn);
    const int m0 = tm * kDT, n0 = tn * kDT;
    const long kbeg = (long)chunk * kch, kend = std::min(K, kbeg + kch);
    const bool one = same && tm == tn;          // B's tile is A's
    const double *sb = one ? sA : sB;
    const double *Bsrc = same ? A : B;
    const long ldbs = same ? lda : ldb;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
    const int lr = l & 15, lk = l >> 4;
    double4v acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = double4v{0.0, 0.0, 0.0, 0.0};
    DSlab xa, xb;
    if (kbeg < kend) {
        load_dslab(A, lda, kbeg, kend, m0, N, t, xa);
        if (!one) load_dslab(Bsrc, ldbs, kbeg, kend, n0, N, t, xb);
    }
    for (long k0 = kbeg; k0 < kend; k0 += kDKB) {
        store_dslab(xa, sA, t);
        if (!one) store_dslab(xb, sB, t);
        __syncthreads();
        if (k0 + kDKB < kend) {
            load_dslab(A, lda, k0 + kDKB, kend, m0, N, t, xa);
            if (!one) load_dslab(Bsrc, ldbs, k0 + kDKB, kend, n0, N, t, xb);
        }
#pragma unroll
        for (int ks = 0; ks < kDKB; ks += 4) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[(wm + 16 * i + lr) * kDP + ks + lk];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sb[(wn + 16 * j + lr) * kDP + ks + lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    double *P = part + (size_t)chunk * N * N;
    // upper: the tile's entries above the diagonal are mirrored below it;
    // a diagonal tile of a non-symmetric product keeps its upper half only
    const bool mirror = upper && tm != tn;
    const bool diag_upper = upper && !same && tm == tn;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + wn + 16 * j + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * i + lk + 4 * v;
                if (m < N && n < N) {
                    if (diag_upper) {
                        if (m <= n) {
                            P[m + (size_t)n * N] = acc[i][j][v];
                            P[n + (size_t)m * N] = acc[i][j][v];
                        }
                    } else {
                        P[m + (size_t)n * N] = acc[i][j][v];
                        if (mirror) P[n + (size_t)m * N] = acc[i][j][v];
                    }
                }
            }
        }
}

// ---------------------------------------------------------------- rotation --
// C (M x N) = A (M x K) Q (K x N) for a real FP32 block A (M grid points,
// column-major, ld lda) and an FP64 Q, as split-BF16x3 products: computed as
// C^T = Q^T A^T, so that the matrix cores' output rows run along the grid
// points and each 32-lane half-wave stores 32 consecutive points of one
// column (coalesced).  Q is split once (k_split_q) into bfloat16 hi and lo,
// n-major with k contiguous and zero padding to Kp x Npad; A is split while
// it is staged.  A workgroup of 4 * NS waves owns 128 grid points and up to
// 160 * NS columns; wave w takes the 32-point block w % 4 and the five
// 32-column blocks of strip w / 4.  Per 16-deep K block, the first four
// waves stage A (lane = grid point, eight K values, each load coalesced
// across the lanes) and all waves stage Q, both as (hi, lo) rows with k
// contiguous; the next block's loads are in flight during the products.
constexpr int kRB = 128;             // grid points per workgroup
constexpr int kRK = 16;              // K per LDS block (one MFMA step)
constexpr int kRP = kRK + 8;         // bf16 per LDS row, padded
constexpr int kRNB = 5;              // 32-column blocks per wave

__global__ void k_split_q(const double *__restrict__ Q, int ldq, int K, int N, int Kp, int Npad,
                          unsigned short *__restrict__ qh, unsigned short *__restrict__ ql)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Npad * Kp) return;
    const int n = (int)(i / Kp), k = (int)(i % Kp);
    const double q = (n < N && k < K) ? Q[k + (long)n * ldq] : 0.0;
    const unsigned short h = bf16_bits((float)q);
    qh[i] = h;
    ql[i] = bf16_bits((float)(q - (double)bf16_value(h)));
}

template <int NS, bool DOUT>
__global__ void __launch_bounds__(256 * NS) k_rotate_bf16x3(const float *__restrict__ A, long lda, long M, int K,
                                                           const unsigned short *__restrict__ qh,
                                                           const unsigned short *__restrict__ ql, int Kp, int N,
                                                           void *__restrict__ Cv, long ldc, int mtiles, int per_xcd)
{
    constexpr int NT = 256 * NS;
    constexpr int NC = 160 * NS;                    // columns per workgroup
    constexpr int QCH = NC * 2 * (kRK / 8);         // 16-byte (hi or lo) chunks of Q per K block
    constexpr int QPT = (QCH + NT - 1) / NT;
    const int item = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (item >= mtiles) return;
    __shared__ __attribute__((aligned(16))) unsigned short sAh[kRB * kRP], sAl[kRB * kRP];
    __shared__ __attribute__((aligned(16))) unsigned short sQh[NC * kRP], sQl[NC * kRP];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const long m0 = (long)item * kRB;
    const int n0 = blockIdx.y * NC;                 // first column of this workgroup
    const int Npad = (N + 31) / 32 * 32;
    const int mb = (w & 3) * 32, ns = w >> 2;       // this wave's 32 points and column strip
    // columns of the strip that exist (blocks of 32 beyond N are skipped)
    int nblk = 0;
#pragma unroll
    for (int j = 0; j < kRNB; ++j)
        if (n0 + (ns * kRNB + j) * 32 < Npad) nblk = j + 1;

    // A staging: thread t < 256 holds grid point m0 + t % 128, K values
    // 8 (t / 128) .. + 7 of the block
    const long am = m0 + (t & 127);
    const int akg = (t >> 7) & 1;
    float av[8];
    uint4 qv[QPT];
    auto load = [&](int k0) {
        if (t < 256) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = k0 + 8 * akg + i;
                av[i] = (am < M && k < K) ? A[am + (long)k * lda] : 0.0f;
            }
        }
#pragma unroll
        for (int q = 0; q < QPT; ++q) {
            const int c = t + q * NT;
            if (c < QCH) {
                // chunk c: column n = c / 4 of this workgroup, half (hi, lo) and k offset
                const int n = c >> 2, part = (c >> 1) & 1, ko = (c & 1) * 8;
                const int gn = n0 + n;
                const unsigned short *src = part ? ql : qh;
                qv[q] = gn < Npad ? *reinterpret_cast<const uint4 *>(src + (long)gn * Kp + k0 + ko)
                                  : make_uint4(0, 0, 0, 0);
            }
        }
    };
    auto store = [&]() {
        if (t < 256) {
            unsigned short h[8], lo[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                h[i] = bf16_bits(av[i]);
                lo[i] = bf16_bits(av[i] - bf16_value(h[i]));
            }
            uint4 a, b;
            a.x = h[0] | ((unsigned)h[1] << 16); a.y = h[2] | ((unsigned)h[3] << 16);
            a.z = h[4] | ((unsigned)h[5] << 16); a.w = h[6] | ((unsigned)h[7] << 16);
            b.x = lo[0] | ((unsigned)lo[1] << 16); b.y = lo[2] | ((unsigned)lo[3] << 16);
            b.z = lo[4] | ((unsigned)lo[5] << 16); b.w = lo[6] | ((unsigned)lo[7] << 16);
            *reinterpret_cast<uint4 *>(sAh + (t & 127) * kRP + 8 * akg) = a;
            *reinterpret_cast<uint4 *>(sAl + (t & 127) * kRP + 8 * akg) = b;
        }
#pragma unroll
        for (int q = 0; q < QPT; ++q) {
            const int c = t + q * NT;
            if (c < QCH) {
                const int n = c >> 2, part = (c >> 1) & 1, ko = (c & 1) * 8;
                *reinterpret_cast<uint4 *>((part ? sQl : sQh) + n * kRP + ko) = qv[q];
            }
        }
    };

    floatx16 acc[kRNB];
#pragma unroll
    for (int j = 0; j < kRNB; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[j][r] = 0.0f;
    const int kl = 8 * (l >> 5);
    load(0);
    for (int k0 = 0; k0 < Kp; k0 += kRK) {
        store();
        __syncthreads();
        if (k0 + kRK < Kp) load(k0 + kRK);
        // B operand: A^T (k x point), lane supplies A[point mb + l % 32][k]
        const bf16x8 bh = frag_p(sAh, mb + (l & 31), kl), blo = frag_p(sAl, mb + (l & 31), kl);
#pragma unroll
        for (int j = 0; j < kRNB; ++j) {
            if (j < nblk) {
                // A operand: Q^T (column x k), lane supplies Q[k][column]
                const int cn = (ns * kRNB + j) * 32 + (l & 31);
                const bf16x8 qhf = frag_p(sQh, cn, kl), qlf = frag_p(sQl, cn, kl);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qhf, bh, acc[j], 0, 0, 0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qhf, blo, acc[j], 0, 0, 0);
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qlf, bh, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[column][point]: lane l holds column 8 (r / 4) + 4 (l / 32) + r % 4
    // of the block and point l % 32
    const long m = m0 + mb + (l & 31);
#pragma unroll
    for (int j = 0; j < kRNB; ++j) {
        if (j >= nblk) continue;
        const int cb = n0 + (ns * kRNB + j) * 32;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int n = cb + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
            if (m < M && n < N) {
                if constexpr (DOUT) static_cast<double *>(Cv)[m + (long)n * ldc] = (double)acc[j][r];
                else static_cast<float *>(Cv)[m + (long)n * ldc] = acc[j][r];
            }
        }
    }
}

}  // namespace

size_t rotate_bf16x3_work(int K, int N)
{
    const int Kp = (K + kRK - 1) / kRK * kRK, Npad = (N + 31) / 32 * 32;
    return (size_t)Kp * Npad * 2 * sizeof(unsigned short);
}

void rotate_bf16x3(const float *A, long lda, long M, const double *Q, int ldq, int K, int N, void *C, long ldc,
                   bool c_double, void *work, hipStream_t st)
{
    if (M < 1 || K < 1 || N < 1) throw std::runtime_error("rotate_bf16x3: empty product");
    const int Kp = (K + kRK - 1) / kRK * kRK, Npad = (N + 31) / 32 * 32;
    unsigned short *qh = static_cast<unsigned short *>(work);
    unsigned short *ql = qh + (size_t)Kp * Npad;
    const long nq = (long)Kp * Npad;
    k_split_q<<<(unsigned)((nq + 255) / 256), 256, 0, st>>>(Q, ldq, K, N, Kp, Npad, qh, ql);
    const int mtiles = (int)((M + kRB - 1) / kRB), per_xcd = (mtiles + 7) / 8;
    // column strips of 160 per workgroup (one to three), more workgroup rows beyond 480
    const int ns = std::min(3, (Npad + 159) / 160);
    const dim3 grid(8 * per_xcd, (Npad + 160 * ns - 1) / (160 * ns));
    auto go = [&](auto kern, int nt) {
        hipLaunchKernelGGL(kern, grid, dim3(nt), 0, st, A, lda, M, K, qh, ql, Kp, N, C, ldc, mtiles, per_xcd);
    };
    if (ns == 1) c_double ? go(k_rotate_bf16x3<1, true>, 256) : go(k_rotate_bf16x3<1, false>, 256);
    else if (ns == 2) c_double ? go(k_rotate_bf16x3<2, true>, 512) : go(k_rotate_bf16x3<2, false>, 512);
    else c_double ? go(k_rotate_bf16x3<3, true>, 768) : go(k_rotate_bf16x3<3, false>, 768);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("k_rotate_bf16x3: ") + hipGetErrorString(e));
}

int gram_bf16x3_splits(long K, int N)
{
    const int tiles = (N + kTile - 1) / kTile;
    const int pairs = tiles * (tiles + 1) / 2;   // the symmetric product's tile count
    // about three resident workgroups per CU on 256 CUs, chunks of at
    // least one LDS block
    const long want = std::max(1L, (768L + pairs - 1) / pairs);
    return (int)std::max(1L, std::min(want, (K + kKB - 1) / kKB));
}

int gram_bf16x3(const float *A, long lda, const float *B, long ldb, long K, int N, float *part, int nsplit,
                hipStream_t st)
{
    if (N < 1 || K < 1 || nsplit < 1) throw std::runtime_error("gram_bf16x3: empty product");
    if ((lda % 4) || (ldb % 4) || (reinterpret_cast<uintptr_t>(A) % 16) || (reinterpret_cast<uintptr_t>(B) % 16))
        throw std::runtime_error("gram_bf16x3: columns must be 16-byte aligned");
    const int tiles = (N + kTile - 1) / kTile;
    const int same = A == B && lda == ldb ? 1 : 0;
    const int pairs = same ? tiles * (tiles + 1) / 2 : tiles * tiles;
    // chunk length: a multiple of the LDS block, so 16-byte loads stay aligned
    long kch = (K + nsplit - 1) / nsplit;
    kch = (kch + kKB - 1) / kKB * kKB;
    const int used = (int)((K + kch - 1) / kch);
    const int items = pairs * used, per_xcd = (items + 7) / 8;
    k_gram_bf16x3<<<8 * per_xcd, kThreads, 0, st>>>(A, lda, B, ldb, K, N, kch, tiles, same, pairs, items, per_xcd,
                                                    part);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("k_gram_bf16x3: ") + hipGetErrorString(e));
    return used;
}

int gram_f64_splits(long K, int N)
{
    const int tiles = (N + kDT - 1) / kDT;
    const int pairs = tiles * (tiles + 1) / 2;
    const long want = std::max(1L, (768L + pairs - 1) / pairs);
    return (int)std::max(1L, std::min(want, (K + kDKB - 1) / kDKB));
}

bool gram_f64_fits(const double *A, long lda, const double *B, long ldb)
{
    return lda % 2 == 0 && ldb % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 &&
           reinterpret_cast<uintptr_t>(B) % 16 == 0;
}

int gram_f64(const double *A, long lda, const double *B, long ldb, long K, int N, double *part, int nsplit,
             bool sym, hipStream_t st)
{
    if (N < 1 || K < 1 || nsplit < 1) throw std::runtime_error("gram_f64: empty product");
    if (!gram_f64_fits(A, lda, B, ldb)) throw std::runtime_error("gram_f64: columns must be 16-byte aligned");
    const int tiles = (N + kDT - 1) / kDT;
    const int same = A == B && lda == ldb ? 1 : 0;
    const int upper = sym || same ? 1 : 0;
    const int pairs = upper ? tiles * (tiles + 1) / 2 : tiles * tiles;
    // chunk length: a multiple of the LDS block, so 16-byte loads stay aligned
    long kch = (K + nsplit - 1) / nsplit;
    kch = (kch + kDKB - 1) / kDKB * kDKB;
    const int used = (int)((K + kch - 1) / kch);
    const int items = pairs * used, per_xcd = (items + 7) / 8;
    k_gram_f64<<<8 * per_xcd, 256, 0, st>>>(A, lda, B, ldb, K, N, kch, tiles, same, upper, pairs, items, per_xcd,
                                            part);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string("k_gram_f64: ") + hipGetErrorString(e));
    return used;
}

}  // namespace gpu
}  // namespace lynx
