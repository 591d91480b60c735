// Complex64 trailing products of the unpivoted LU on fp16 MFMA (v_mfma_f32_16x16x32_f16).
//
// Each fp32 value x of a row of L21 (or a column of U12) is held as x = 2^-e (hi + lo) with
// hi = fp16(x 2^e) and lo = fp16(x 2^e - hi), e chosen per row (per column) so that the row's
// largest |re| or |im| lands in [2^14, 2^15).  A real product is then hi hi + hi lo + lo hi: the
// three fp16 products are exact in the fp32 accumulator, and the dropped lo lo term and the
// rounding of lo leave about 3 * 2^-24 of |a||b| per product, below the error the fp32 accumulation
// of a K = 128 sum already carries.  Entries 2^17 below their row's largest lose the low bits of
// lo to fp16's subnormal range; their absolute error stays under 2^-39 of the row's scale.
//
// Complex products use four real ones (re = Ur Lr - Ui Li, im = Ur Li + Ui Lr), twelve fp16
// MFMAs per 16 x 16 block and 32-deep k chunk.  The product is taken transposed (D = U^T L^T), so
// that 16 consecutive lanes of an accumulator value hold 16 consecutive rows of C and the
// read-modify-write of C goes down its columns.
//
// Lane layout of v_mfma_f32_16x16x32_f16 (measured on gfx950 with one-hot operands): A[i][k] in
// lane i + 16 (k / 8), slot k % 8; B[k][j] in lane j + 16 (k / 8), slot k % 8; D[i][j] in lane
// j + 16 (i / 4), value i % 4.
#include "common.hpp"
#include "lu.hpp"

#include <cstdlib>

namespace frico {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------------------------------------
// Split: S[k + r ld] (r < R rows of K <= 128 contiguous values each) -> four planes
// P[p][r][k] (p = re_hi, re_lo, im_hi, im_lo; HSPLIT_K halves a row, zero beyond K and for the
// padding rows R <= r < Rp) and the row exponents e[r].  16 lanes a row, 8 values a lane.
// ----------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void split_f16(const c64* __restrict__ S, int64_t ld, int64_t R, int K, int64_t Rp,
                                                 _Float16* __restrict__ P, int64_t cap, int* __restrict__ E)
{
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t r = gid >> 4;
    const int g = (int)(gid & 15);
    if (r >= Rp)
        return;
    float vr[8], vi[8];
    float mx = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int k = 8 * g + q;
        c64 v{0.0f, 0.0f};
        if (r < R && k < K)
            v = S[k + r * ld];
        vr[q] = v.re;
        vi[q] = v.im;
        mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
    }
    // the row's largest magnitude, over its 16 lanes
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
        mx = fmaxf(mx, __shfl_xor(mx, o, 16));
    int e = 0;
    if (mx > 0.0f) {
        e = 14 - ilogbf(mx); // mx 2^e in [2^14, 2^15)
        e = e < -100 ? -100 : (e > 200 ? 200 : e);
    }
    half8 rh, rl, ih, il;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float xr = ldexpf(vr[q], e), xi = ldexpf(vi[q], e);
        const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
        rh[q] = hr;
        rl[q] = (_Float16)(xr - (float)hr);
        ih[q] = hi;
        il[q] = (_Float16)(xi - (float)hi);
    }
    const int64_t off = r * HSPLIT_K + 8 * g, ps = cap * HSPLIT_K;
    *reinterpret_cast<half8*>(P + off) = rh;
    *reinterpret_cast<half8*>(P + ps + off) = rl;
    *reinterpret_cast<half8*>(P + 2 * ps + off) = ih;
    *reinterpret_cast<half8*>(P + 3 * ps + off) = il;
    if (g == 0)
        E[r] = e;
}

// ----------------------------------------------------------------------------------------------
// C[M x N] -= L U with L given as split rows (La, ea: M rows) and U as split columns (Ub, eb:
// N columns), K <= 128 (kc = K rounded up to 32).  Tile BM x BN = 32 RB x 32 CB per 256-thread
// workgroup, 2 x 2 waves of RB x CB blocks of 16 x 16, 32-deep k chunks through LDS.  The split
// planes are padded to whole 128-row tiles, so the loads need no bounds; only C is masked.  Small
// tiles (RB CB <= 4) would load their part of C before the products; the shapes used read it in
// the epilogue.
// ----------------------------------------------------------------------------------------------
constexpr int HKC = 32; // k chunk
// LDS row of one plane: 32 halves = four 16-byte segments, stored at segment s ^ ((row >> 1) & 3)
// so that eight consecutive rows read by one quarter-wave hit distinct banks.
__device__ inline int hseg(int row, int s) { return s ^ ((row >> 1) & 3); }

// Tile order.  Workgroups go to the eight XCDs in turn (workgroup b to XCD b % 8), each XCD with its
// own L2.  Workgroup b takes tile s = (b % 8) ceil(T / 8) + b / 8 of a sequence that runs through
// groups of HG column tiles, all row tiles of a group (HG tiles a row) before the next group: each
// XCD works through a contiguous stretch of groups, keeping its columns of U in its L2 and reusing
// each row block of L across the group.
constexpr int HG = 4;
__device__ inline bool tile_of(int64_t b, int64_t gx, int64_t gy, int64_t& tm, int64_t& tn)
{
    const int64_t T = gx * gy, per = (T + 7) / 8;
    const int64_t s = (b % 8) * per + b / 8;
    if (s >= T)
        return false;
    const int64_t full = gy / HG, infull = full * gx * HG;
    int64_t g, si, width;
    if (s < infull) {
        g = s / (gx * HG);
        si = s % (gx * HG);
        width = HG;
    } else {
        g = full;
        si = s - infull;
        width = gy - full * HG;
    }
    tm = si / width;
    tn = g * HG + si % width;
    return true;
}

// A second k half (TWO): La1/ea1 and Ub1/eb1 with kc1 more k, each half split with its own
// exponents.  After the first half the accumulators are rescaled to the second half's exponents
// (a power of two per entry: D = 2^-(e0) D0 + 2^-(e1) D1 = 2^-(e1) (2^(e1-e0) D0 + D1)), so that one
// read-modify-write of C takes both: two steps' trailing updates at the C traffic of one.
struct F16Half {
    const _Float16* La;
    const int* ea;
    const _Float16* Ub;
    const int* eb;
    int kc;
};

template <int RB, int CB, int OCC, bool TWO = false>
__global__ __launch_bounds__(256, OCC) void cgemm_sub_f16(int64_t M, int64_t N, int kc, const _Float16* __restrict__ La,
                                                         const int* __restrict__ ea, int64_t acap,
                                                         const _Float16* __restrict__ Ub, const int* __restrict__ eb,
                                                         int64_t bcap, c64* __restrict__ C, int64_t ldc, int64_t gx,
                                                         int64_t gy, F16Half h1, Batch bt)
{
    // batched: the L planes (both halves) += z s0, their exponents += z s1, the U planes += z s2,
    // their exponents += z s3, C += z s4
    {
        const int64_t z = blockIdx.z;
        La += z * bt.s0;
        ea += z * bt.s1;
        Ub += z * bt.s2;
        eb += z * bt.s3;
        C += z * bt.s4;
        if constexpr (TWO) {
            h1.La += z * bt.s0;
            h1.ea += z * bt.s1;
            h1.Ub += z * bt.s2;
            h1.eb += z * bt.s3;
        }
    }
    constexpr int BM = 32 * RB, BN = 32 * CB;
    constexpr bool PREC = RB * CB <= 4;
    __shared__ half8 sA[4][BM][4]; // [plane][row][segment]
    __shared__ half8 sB[4][BN][4]; // [plane][column][segment]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
    int64_t tm, tn;
    if (!tile_of(blockIdx.x, gx, gy, tm, tn))
        return;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const int64_t aps = acap * HSPLIT_K, bps = bcap * HSPLIT_K;
    // lane's rows: m0 + wm 16 RB + rb 16 + (lane & 15); columns: n0 + wn 16 CB + cb 16 + 4 (lane >> 4) + r
    const int64_t rbase = m0 + wm * 16 * RB + (lane & 15), cbase = n0 + wn * 16 * CB + 4 * (lane >> 4);

    c64 cv[PREC ? RB : 1][PREC ? CB : 1][4];
    if constexpr (PREC) {
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int cb = 0; cb < CB; cb++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t row = rbase + rb * 16, col = cbase + cb * 16 + r;
                    const int64_t ok = row < M && col < N;
                    cv[rb][cb][r] = C[ok * (row + col * ldc)];
                }
    }

    f32x4 are[RB][CB], aim[RB][CB];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < CB; b++) {
            are[a][b] = f32x4{0, 0, 0, 0};
            aim[a][b] = f32x4{0, 0, 0, 0};
        }

    constexpr int SA = BM / 16, SB = BN / 16; // 16-byte segments per thread per chunk
    auto kloop = [&](const _Float16* __restrict__ La, const _Float16* __restrict__ Ub, int kc) {
    for (int k0 = 0; k0 < kc; k0 += HKC) {
        // global -> registers: thread e takes segment e % 4 of row (e / 4) % BM of plane e / (4 BM)
        half8 ra[SA], rb_[SB];
#pragma unroll
        for (int s = 0; s < SA; s++) {
            const int e = tid + 256 * s, p = e / (4 * BM), row = (e >> 2) % BM, sg = e & 3;
            ra[s] = *reinterpret_cast<const half8*>(La + p * aps + (m0 + row) * HSPLIT_K + k0 + 8 * sg);
        }
#pragma unroll
        for (int s = 0; s < SB; s++) {
            const int e = tid + 256 * s, p = e / (4 * BN), row = (e >> 2) % BN, sg = e & 3;
            rb_[s] = *reinterpret_cast<const half8*>(Ub + p * bps + (n0 + row) * HSPLIT_K + k0 + 8 * sg);
        }
        __syncthreads(); // the previous chunk's fragments have been read
#pragma unroll
        for (int s = 0; s < SA; s++) {
            const int e = tid + 256 * s, p = e / (4 * BM), row = (e >> 2) % BM, sg = e & 3;
            sA[p][row][hseg(row, sg)] = ra[s];
        }
#pragma unroll
        for (int s = 0; s < SB; s++) {
            const int e = tid + 256 * s, p = e / (4 * BN), row = (e >> 2) % BN, sg = e & 3;
            sB[p][row][hseg(row, sg)] = rb_[s];
        }
        __syncthreads();
        // this wave's column blocks of U (the MFMA's A operand: U^T, 16 columns x 32 k)
        const int sgl = lane >> 4;
        half8 bf[CB][4];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int col = wn * 16 * CB + cb * 16 + (lane & 15);
#pragma unroll
            for (int p = 0; p < 4; p++)
                bf[cb][p] = sB[p][col][hseg(col, sgl)];
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int row = wm * 16 * RB + rb * 16 + (lane & 15);
            half8 af[4];
#pragma unroll
            for (int p = 0; p < 4; p++)
                af[p] = sA[p][row][hseg(row, sgl)];
            const half8 nih = -af[2], nil = -af[3];
            // re += Ur Lr - Ui Li ; im += Ur Li + Ui Lr (hi hi + lo hi + hi lo of each)
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[0], are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[2], aim[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][1], af[0], are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][1], af[2], aim[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[1], are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[3], aim[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], nih, are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], af[0], aim[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][3], nih, are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][3], af[0], aim[rb][cb], 0, 0, 0);
            }
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], nil, are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], af[1], aim[rb][cb], 0, 0, 0);
            }
        }
    }

    };
    kloop(La, Ub, kc);
    if constexpr (TWO) {
        // rescale the first half's sums to the second half's exponents, then the second half
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int64_t row = rbase + rb * 16;
            const int64_t rr = row < M ? row : 0;
            const int dr = h1.ea[rr] - ea[rr];
#pragma unroll
            for (int cb = 0; cb < CB; cb++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int64_t col = cbase + cb * 16 + r;
                    const int64_t cc = col < N ? col : 0;
                    int d = dr + h1.eb[cc] - eb[cc];
                    d = d < -200 ? -200 : (d > 100 ? 100 : d);
                    are[rb][cb][r] = ldexpf(are[rb][cb][r], d);
                    aim[rb][cb][r] = ldexpf(aim[rb][cb][r], d);
                }
        }
        kloop(h1.La, h1.Ub, h1.kc);
        ea = h1.ea;
        eb = h1.eb;
    }
    // C -= 2^-(ea + eb) D^T: lane holds row (lane & 15) of each row block and columns
    // 4 (lane >> 4) + r of each column block
    int ecol[CB][4];
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            ecol[cb][r] = eb[cbase + cb * 16 + r];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int64_t row = rbase + rb * 16;
        if (row >= M)
            continue;
        const int erow = ea[row];
        c64 cr[CB][4];
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if constexpr (PREC) {
                    cr[cb][r] = cv[rb][cb][r];
                } else {
                    const int64_t col = cbase + cb * 16 + r;
                    const int64_t ok = col < N;
                    cr[cb][r] = C[ok * (row + col * ldc)];
                }
            }
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t col = cbase + cb * 16 + r;
                if (col < N) {
                    const int sc = -(erow + ecol[cb][r]);
                    c64 v = cr[cb][r];
                    v.re -= ldexpf(are[rb][cb][r], sc);
                    v.im -= ldexpf(aim[rb][cb][r], sc);
                    C[row + col * ldc] = v;
                }
            }
    }
}

// ----------------------------------------------------------------------------------------------
// Products of the triangular solves on fp16 MFMA, split as the operands are staged.
//   C[M x N] -= A[M x K] B[K x N], all complex64 and column-major (A a block of the LU factors, B
//   a block of the solution), A's row r scaled by 2^ea[r] and B's column j by 2^eb[j] before the
//   fp16 hi/lo split, C -= 2^-(ea + eb) D in the epilogue.  The exponents put the largest |re| or
//   |im| of the row (column) over the product's K in [2^14, 2^15), as split_f16 does, so the
//   arithmetic is that of the factorisation's split products.
// Exponents: rowexp_blocks gives, for every block of W columns of a matrix, each row's exponent
// over the block (one table for the solves' blocks of the factors); colexp each column's over K
// rows of a block of the solution.
// ----------------------------------------------------------------------------------------------
__device__ inline int split_exponent(float mx)
{
    if (!(mx > 0.0f))
        return 0;
    const int e = 14 - ilogbf(mx);
    return e < -100 ? -100 : (e > 200 ? 200 : e);
}

__global__ __launch_bounds__(256) void rowexp_blocks(const c64* __restrict__ A, int64_t lda, int64_t n, int W,
                                                     int* __restrict__ E, int64_t ldE)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t b = blockIdx.y, c0 = b * W, c1 = (c0 + W) < n ? (c0 + W) : n;
    if (r >= n)
        return;
    float mx = 0.0f;
    for (int64_t j = c0; j < c1; j++) {
        const c64 v = A[r + j * lda];
        mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
    }
    E[b * ldE + r] = split_exponent(mx);
}

__global__ __launch_bounds__(256) void colexp(const c64* __restrict__ B, int64_t ldb, int64_t K, int* __restrict__ E)
{
    __shared__ float red[4];
    const int64_t j = blockIdx.x;
    float mx = 0.0f;
    for (int64_t k = threadIdx.x; k < K; k += 256) {
        const c64 v = B[k + j * ldb];
        mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0)
        E[j] = split_exponent(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3])));
}

// x 2^e = hi + lo for eight values (one 16-byte segment of each of the four planes)
__device__ inline void split8(const c64 (&v)[8], int e, half8& rh, half8& rl, half8& ih, half8& il)
{
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float xr = ldexpf(v[q].re, e), xi = ldexpf(v[q].im, e);
        const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
        rh[q] = hr;
        rl[q] = (_Float16)(xr - (float)hr);
        ih[q] = hi;
        il[q] = (_Float16)(xi - (float)hi);
    }
}

// Tile 128 x 64 (2 x 2 waves of 64 x 32), 32-deep k chunks.  A: thread t stages rows t % 128,
// k segments 2 (t / 128) and 2 (t / 128) + 1 (eight k each; a load instruction reads 64
// consecutive rows of one column); B: column t / 4, k segment t % 4 (64 contiguous bytes).
__global__ __launch_bounds__(256, 2) void cgemm_sub_f16s(int64_t M, int64_t N, int64_t K, const c64* __restrict__ A,
                                                         int64_t lda, const int* __restrict__ ea,
                                                         const c64* __restrict__ B, int64_t ldb,
                                                         const int* __restrict__ eb, c64* __restrict__ C, int64_t ldc,
                                                         int64_t gx, int64_t gy)
{
    constexpr int RB = 4, CB = 2, BM = 128, BN = 64;
    __shared__ half8 sA[4][BM][4];
    __shared__ half8 sB[4][BN][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
    int64_t tm, tn;
    if (!tile_of(blockIdx.x, gx, gy, tm, tn))
        return;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    // staging: A row ar, segments as0, as0 + 1; B column bc, segment bs
    const int ar = tid % BM, as0 = 2 * (tid / BM), bc = tid / 4, bs = tid % 4;
    const int64_t arow = m0 + ar, bcol = n0 + bc;
    const bool aok = arow < M, bok = bcol < N;
    const int eA = aok ? ea[arow] : 0, eB = bok ? eb[bcol] : 0;
    const c64* __restrict__ Ar = A + (aok ? arow : 0);
    const c64* __restrict__ Bc = B + (bok ? bcol : 0) * ldb;

    f32x4 are[RB][CB], aim[RB][CB];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
        for (int b = 0; b < CB; b++) {
            are[a][b] = f32x4{0, 0, 0, 0};
            aim[a][b] = f32x4{0, 0, 0, 0};
        }
    for (int64_t k0 = 0; k0 < K; k0 += HKC) {
        // branch-free loads (out-of-range ones read element 0 and are zeroed after), all in flight
        // together
        c64 va[2][8], vb[8];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int64_t k = k0 + 8 * (as0 + h) + q;
                va[h][q] = Ar[(k < K ? k : 0) * lda];
            }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int64_t k = k0 + 8 * bs + q;
            vb[q] = Bc[k < K ? k : 0];
        }
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (!(aok && k0 + 8 * (as0 + h) + q < K))
                    va[h][q] = c64{0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (!(bok && k0 + 8 * bs + q < K))
                vb[q] = c64{0.0f, 0.0f};
        __syncthreads(); // the previous chunk's fragments have been read
#pragma unroll
        for (int h = 0; h < 2; h++) {
            half8 rh, rl, ih, il;
            split8(va[h], eA, rh, rl, ih, il);
            const int sg = hseg(ar, as0 + h);
            sA[0][ar][sg] = rh;
            sA[1][ar][sg] = rl;
            sA[2][ar][sg] = ih;
            sA[3][ar][sg] = il;
        }
        {
            half8 rh, rl, ih, il;
            split8(vb, eB, rh, rl, ih, il);
            const int sg = hseg(bc, bs);
            sB[0][bc][sg] = rh;
            sB[1][bc][sg] = rl;
            sB[2][bc][sg] = ih;
            sB[3][bc][sg] = il;
        }
        __syncthreads();
        const int sgl = lane >> 4;
        half8 bf[CB][4];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int col = wn * 16 * CB + cb * 16 + (lane & 15);
#pragma unroll
            for (int p = 0; p < 4; p++)
                bf[cb][p] = sB[p][col][hseg(col, sgl)];
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int row = wm * 16 * RB + rb * 16 + (lane & 15);
            half8 af[4];
#pragma unroll
            for (int p = 0; p < 4; p++)
                af[p] = sA[p][row][hseg(row, sgl)];
            const half8 nih = -af[2], nil = -af[3];
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[0], are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[2], aim[rb][cb], 0, 0, 0);
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][1], af[0], are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][1], af[2], aim[rb][cb], 0, 0, 0);
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[1], are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[3], aim[rb][cb], 0, 0, 0);
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], nih, are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], af[0], aim[rb][cb], 0, 0, 0);
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][3], nih, are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][3], af[0], aim[rb][cb], 0, 0, 0);
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], nil, are[rb][cb], 0, 0, 0);
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], af[1], aim[rb][cb], 0, 0, 0);
            }
        }
    }
    // C -= 2^-(ea + eb) D^T: lane holds row (lane & 15) of each row block and columns 4 (lane >> 4) + r
    const int64_t rbase = m0 + wm * 16 * RB + (lane & 15), cbase = n0 + wn * 16 * CB + 4 * (lane >> 4);
    int ecol[CB][4];
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t col = cbase + cb * 16 + r;
            ecol[cb][r] = eb[col < N ? col : 0];
        }
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const int64_t row = rbase + rb * 16;
        if (row >= M)
            continue;
        const int erow = ea[row];
        c64 cr[CB][4];
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t col = cbase + cb * 16 + r;
                const int64_t ok = col < N;
                cr[cb][r] = C[ok * (row + col * ldc)];
            }
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t col = cbase + cb * 16 + r;
                if (col < N) {
                    const int sc = -(erow + ecol[cb][r]);
                    c64 v = cr[cb][r];
                    v.re -= ldexpf(are[rb][cb][r], sc);
                    v.im -= ldexpf(aim[rb][cb][r], sc);
                    C[row + col * ldc] = v;
                }
            }
    }
}

int launch_rowexp_blocks(const c64* A, int64_t lda, int64_t n, int W, int* E, int64_t ldE, hipStream_t st)
{
    if (n <= 0)
        return 0;
    const int64_t nb = (n + W - 1) / W;
    if (nb > 65535 || ldE < n)
        return set_error("row exponents: %lld blocks, table stride %lld", (long long)nb, (long long)ldE), -1;
    launch_kernel(&rowexp_blocks, dim3((unsigned)((n + 255) / 256), (unsigned)nb), dim3(256), 0, st, A, lda, n, W, E, ldE);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int gemm_sub_f16s(int64_t M, int64_t N, int64_t K, const c64* A, int64_t lda, const int* ea, const c64* B, int64_t ldb,
                  int* eb, c64* C, int64_t ldc, hipStream_t st)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    launch_kernel(&colexp, dim3((unsigned)N), dim3(256), 0, st, B, ldb, K, eb);
    const int64_t gx = (M + 127) / 128, gy = (N + 63) / 64;
    const int64_t nblk = (gx * gy + 7) / 8 * 8;
    if (nblk > 0x7fffffff)
        return set_error("fp16 product: too many tiles"), -1;
    launch_kernel(&cgemm_sub_f16s, dim3((unsigned)nblk), dim3(256), 0, st, M, N, K, A, lda, ea, B, ldb, (const int*)eb, C,
                  ldc, gx, gy);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_split_f16(const c64* S, int64_t ld, int64_t R, int K, _Float16* planes, int* e, int64_t cap,
                     hipStream_t st)
{
    if (R <= 0)
        return 0;
    if (K < 1 || K > HSPLIT_K)
        return set_error("fp16 split: K = %d outside 1..%d", K, HSPLIT_K), -1;
    const int64_t Rp = (R + 127) / 128 * 128;
    if (Rp > cap)
        return set_error("fp16 split: %lld rows exceed the buffer's %lld", (long long)Rp, (long long)cap), -1;
    launch_kernel(&split_f16, dim3((unsigned)(Rp * 16 / 256)), dim3(256), 0, st, S, ld, R, K, Rp, planes, cap, e);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// Tile shape of the products: 128 x 128, or 128 x 64 (FRICO_GEMM_F16_TILE=12864 for every product).
static int f16_tile()
{
    static const int t = [] {
        const char* e = std::getenv("FRICO_GEMM_F16_TILE");
        return e && std::atoi(e) == 12864 ? 12864 : 128;
    }();
    return t;
}

int gemm_sub_f16(int64_t M, int64_t N, int K, const _Float16* La, const int* ea, int64_t acap, const _Float16* Ub,
                 const int* eb, int64_t bcap, c64* C, int64_t ldc, hipStream_t st, KernelTimer* kt, const Batch& bt,
                 int nz)
{
    if (nz < 1 || nz > 65535)
        return set_error("fp16 product: batch of %d systems", nz), -1;
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    if (K > HSPLIT_K)
        return set_error("fp16 product: K = %d above %d", K, HSPLIT_K), -1;
    // 128 x 128 tiles for products that fill the GPU several times over, 128 x 64 below
    int tile = f16_tile();
    if (tile == 128 && ((M + 127) / 128) * ((N + 127) / 128) < 1024)
        tile = 12864;
    const int BM = 128, BN = tile == 128 ? 128 : 64;
    const int64_t gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
    // the split planes are padded to 128-row tiles, which cover any smaller tile
    if ((M + 127) / 128 * 128 > acap || (N + 127) / 128 * 128 > bcap)
        return set_error("fp16 product: operands not padded to whole tiles"), -1;
    const int64_t nblk = (gx * gy + 7) / 8 * 8;
    if (nblk > 0x7fffffff)
        return set_error("fp16 product: too many tiles"), -1;
    const int kc = (K + HKC - 1) / HKC * HKC;
    if (kt)
        kt->begin(st);
    const dim3 grid((unsigned)nblk, 1, (unsigned)nz);
    if (tile == 128)
        launch_kernel(&cgemm_sub_f16<4, 4, 2>, grid, dim3(256), 0, st, M, N, kc, La, ea, acap, Ub, eb, bcap, C, ldc, gx, gy,
                      F16Half{}, bt);
    else
        launch_kernel(&cgemm_sub_f16<4, 2, 3>, grid, dim3(256), 0, st, M, N, kc, La, ea, acap, Ub, eb, bcap, C, ldc, gx, gy,
                      F16Half{}, bt);
    // algorithmic bytes: C read and written once (complex64), L and U read once as complex64
    if (kt)
        kt->end(st, 8.0 * (double)M * (double)N * (double)K,
                16.0 * (double)M * (double)N + 8.0 * ((double)M + (double)N) * (double)K);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int gemm_sub_f16_pair(int64_t M, int64_t N, int K0, const _Float16* La0, const int* ea0, const _Float16* Ub0,
                      const int* eb0, int K1, const _Float16* La1, const int* ea1, const _Float16* Ub1, const int* eb1,
                      int64_t acap, int64_t bcap, c64* C, int64_t ldc, hipStream_t st, KernelTimer* kt,
                      const Batch& bt, int nz)
{
    if (M <= 0 || N <= 0)
        return 0;
    if (nz < 1 || nz > 65535)
        return set_error("fp16 product: batch of %d systems", nz), -1;
    if (K0 <= 0 || K0 > HSPLIT_K || K1 <= 0 || K1 > HSPLIT_K)
        return set_error("fp16 product: halves of K = %d + %d outside 1..%d", K0, K1, HSPLIT_K), -1;
    if (K0 % HKC != 0)
        return set_error("fp16 product: first half K = %d not a multiple of %d", K0, HKC), -1;
    int tile = f16_tile();
    if (tile == 128 && ((M + 127) / 128) * ((N + 127) / 128) < 1024)
        tile = 12864;
    const int BM = 128, BN = tile == 128 ? 128 : 64;
    const int64_t gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
    if ((M + 127) / 128 * 128 > acap || (N + 127) / 128 * 128 > bcap)
        return set_error("fp16 product: operands not padded to whole tiles"), -1;
    const int64_t nblk = (gx * gy + 7) / 8 * 8;
    if (nblk > 0x7fffffff)
        return set_error("fp16 product: too many tiles"), -1;
    const F16Half h1{La1, ea1, Ub1, eb1, (K1 + HKC - 1) / HKC * HKC};
    if (kt)
        kt->begin(st);
    const dim3 grid((unsigned)nblk, 1, (unsigned)nz);
    if (tile == 128)
        launch_kernel(&cgemm_sub_f16<4, 4, 2, true>, grid, dim3(256), 0, st, M, N, K0, La0, ea0, acap, Ub0, eb0, bcap, C, ldc,
                      gx, gy, h1, bt);
    else
        launch_kernel(&cgemm_sub_f16<4, 2, 3, true>, grid, dim3(256), 0, st, M, N, K0, La0, ea0, acap, Ub0, eb0, bcap, C, ldc,
                      gx, gy, h1, bt);
    const double K = (double)K0 + (double)K1;
    if (kt)
        kt->end(st, 8.0 * (double)M * (double)N * K, 16.0 * (double)M * (double)N + 8.0 * ((double)M + (double)N) * K);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

} // namespace frico
