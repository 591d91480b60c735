// Dense complex LU with partial pivoting, triangular solves and iterative-refinement helpers
// for gfx950.  Replaces zgetrf/zgetrs (src/lapack_wrappers.cpp:53-133).
//
// Right-looking blocked LU, block width nb:
//   panel_factor   one launch per panel; G workgroups each keep a slab of R rows x nb columns
//                  in LDS and factor the panel column by column (pivot = first max of |re|+|im|,
//                  as izamax).  One grid barrier per column: every workgroup publishes its local
//                  pivot candidate together with that candidate row, so after the barrier every
//                  workgroup can pick the winner and perform the row interchange locally.
//   laswp          applies the panel's interchanges to the columns right of the panel.  Columns
//                  left of the panel are not swapped; the forward solve applies each panel's
//                  interchanges just before using that panel's L, which is equivalent.
//   trsm_block     U12 = L11^{-1} A12 (unit lower; one wavefront per column group, lane per row).
//   cgemm_sub      A22 -= L21 U12, complex GEMM on MFMA (v_mfma_f64_16x16x4f64 or
//                  v_mfma_f32_16x16x4f32), 64x64 complex output tile per workgroup.
#include "../../include/frico_gpu.h"
#include "common.hpp"
#include "lu.hpp"

#include <cstdlib>
#include <type_traits>

namespace frico {

// ----------------------------------------------------------------------------------------------
// MFMA wrappers.  Operand lane layout of both 16x16x4 shapes: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Accumulator value r of lane l holds D[row][l%16] with
//   f32: row = 4*(l/16) + r        f64: row = 4*r + l/16
// (measured on gfx950 with one-hot operands).
// ----------------------------------------------------------------------------------------------
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline acc_t op(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int acc_row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static inline acc_t op(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int acc_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ----------------------------------------------------------------------------------------------
// Complex output tiles on MFMA.  Workgroup = 4 waves (2 x 2), wave tile (WM*16) x (WN*16).
// The MFMA computes the transposed product (B^T A^T) so that the 16 lanes of an accumulator row
// hold 16 consecutive rows of the output: stores are contiguous down a column.
// ----------------------------------------------------------------------------------------------
constexpr int GBK = 16;

template <int WM_, int WN_> struct TileCfg {
    static constexpr int WM = WM_, WN = WN_, BM = 32 * WM_, BN = 32 * WN_;
};
// 64 x 64: the chunk-aligned products of the solves and of W
using Tile64 = TileCfg<2, 2>;
constexpr int GBM = Tile64::BM, GBN = Tile64::BN;
// Trailing updates, residuals and the transposed solve.  Measured at M = N = 12096, K = 64:
// 64 x 64 tiles reach 92 TFLOP/s (complex64) and 40 (complex128); 128 x 128 / 128 x 64 tiles
// reached 39 / 27 at one wave per SIMD and 61 / 37 when held to two, so the products stay 64 x 64.
template <typename T> struct TileBig {
    using type = Tile64;
};
// Minimum workgroups per CU asked of the compiler for the product kernels.  Complex128 tiles held to
// two (166 VGPRs, no accumulation registers: three waves per SIMD instead of two) ran the K = 64
// update at 49.8 TFLOP/s against 39.7 and the K = 12160 residual at 54.7 against 48.7; complex64
// tiles already reach three waves per SIMD.
template <typename T> constexpr int kGemmOcc = sizeof(T) == 8 ? 2 : 1;

template <typename T, typename TL> struct GemmAcc {
    typename Mfma<T>::acc_t re[TL::WM][TL::WN], im[TL::WM][TL::WN];
    // three-product form: re = T1 = sum br ar, im = T2 = sum bi ai, s3 = T3 = sum (br+bi)(ar+ai)
    typename Mfma<T>::acc_t s3[TL::WM][TL::WN];
};

template <typename T, typename TL> struct GemmSmem {
    cx<T> a[GBK][TL::BM + 1];
    cx<T> b[GBK][TL::BN + 1];
};

// acc = A_t[0:mr, 0:K] * B_t[0:K, 0:nc] (entries outside are zero), with
//   A_t(r, kk) = TRANSA ? A[kk + r lda] : A[r + kk lda],   B_t(kk, c) = B[kk + c ldb].
// With BLDS the B operand comes from LDS instead: B_t(kk, c) = Bs[kk][c].
// M3: complex products with three real MFMAs (T1 = br ar, T2 = bi ai, T3 = (br + bi)(ar + ai);
// re = T1 - T2, im = T3 - T1 - T2) instead of four; gemm_store combines them.
template <typename T, typename TA, bool TRANSA, typename TL, bool BLDS = false, bool M3 = false>
__device__ inline void gemm_tile(int64_t mr, int64_t nc, int64_t K, const cx<TA>* __restrict__ A, int64_t lda,
                                 const cx<T>* __restrict__ B, int64_t ldb, GemmAcc<T, TL>& acc, GemmSmem<T, TL>& sm,
                                 const cx<T> (*Bs)[65] = nullptr)
{
    using MF = Mfma<T>;
    using acc_t = typename MF::acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
#pragma unroll
    for (int a = 0; a < TL::WM; a++)
#pragma unroll
        for (int b = 0; b < TL::WN; b++) {
            acc.re[a][b] = acc_t{0, 0, 0, 0};
            acc.im[a][b] = acc_t{0, 0, 0, 0};
            if (M3)
                acc.s3[a][b] = acc_t{0, 0, 0, 0};
        }

    constexpr int LA = TL::BM * GBK / 256, LB = GBK * TL::BN / 256;
    // Raw loaded values and their in-range masks; the mask is applied when the values go to LDS, so
    // that the prefetch of the next step is not waited for before this step's products.  Loads are
    // unconditional (out-of-range lanes read element 0), so a step's loads are in flight together.
    cx<TA> ra[LA];
    cx<T> rb[LB];
    bool oka[LA], okb[LB];
    auto gload = [&](int64_t k0) {
#pragma unroll
        for (int s = 0; s < LA; s++) {
            const int e = tid + 256 * s;
            // consecutive threads walk the contiguous dimension of A in memory
            const int r = TRANSA ? e / GBK : e % TL::BM, kk = TRANSA ? e % GBK : e / TL::BM;
            const int64_t gk = k0 + kk;
            const int64_t ok = r < mr && gk < K;
            oka[s] = ok;
            ra[s] = A[ok * (TRANSA ? gk + r * lda : r + gk * lda)];
        }
#pragma unroll
        for (int s = 0; s < LB; s++) {
            const int e = tid + 256 * s;
            const int kk = e % GBK, c = e / GBK;
            const int64_t gk = k0 + kk;
            const int64_t ok = gk < K && c < nc;
            okb[s] = ok;
            if (BLDS)
                rb[s] = ok ? Bs[gk][c] : cx<T>{0, 0};
            else
                rb[s] = B[ok * (gk + c * ldb)];
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int s = 0; s < LA; s++) {
            const int e = tid + 256 * s;
            const cx<T> v = oka[s] ? cx<T>{(T)ra[s].re, (T)ra[s].im} : cx<T>{0, 0};
            if (TRANSA)
                sm.a[e % GBK][e / GBK] = v;
            else
                sm.a[e / TL::BM][e % TL::BM] = v;
        }
#pragma unroll
        for (int s = 0; s < LB; s++) {
            const int e = tid + 256 * s;
            sm.b[e % GBK][e / GBK] = okb[s] ? rb[s] : cx<T>{0, 0};
        }
    };

    gload(0);
    for (int64_t k0 = 0; k0 < K; k0 += GBK) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + GBK < K)
            gload(k0 + GBK);
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            cx<T> af[TL::WM], bf[TL::WN];
#pragma unroll
            for (int a = 0; a < TL::WM; a++)
                af[a] = sm.a[kk + (lane >> 4)][wm * TL::WM * 16 + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < TL::WN; b++)
                bf[b] = sm.b[kk + (lane >> 4)][wn * TL::WN * 16 + b * 16 + (lane & 15)];
            if (M3) {
                T as[TL::WM], bs[TL::WN];
#pragma unroll
                for (int a = 0; a < TL::WM; a++)
                    as[a] = af[a].re + af[a].im;
#pragma unroll
                for (int b = 0; b < TL::WN; b++)
                    bs[b] = bf[b].re + bf[b].im;
#pragma unroll
                for (int a = 0; a < TL::WM; a++)
#pragma unroll
                    for (int b = 0; b < TL::WN; b++) {
                        acc.re[a][b] = MF::op(bf[b].re, af[a].re, acc.re[a][b]);
                        acc.im[a][b] = MF::op(bf[b].im, af[a].im, acc.im[a][b]);
                        acc.s3[a][b] = MF::op(bs[b], as[a], acc.s3[a][b]);
                    }
            } else {
#pragma unroll
                for (int a = 0; a < TL::WM; a++)
#pragma unroll
                    for (int b = 0; b < TL::WN; b++) {
                        // D^T = B^T A^T:  re += br*ar - bi*ai ; im += br*ai + bi*ar
                        acc.re[a][b] = MF::op(bf[b].re, af[a].re, acc.re[a][b]);
                        acc.re[a][b] = MF::op(-bf[b].im, af[a].im, acc.re[a][b]);
                        acc.im[a][b] = MF::op(bf[b].re, af[a].im, acc.im[a][b]);
                        acc.im[a][b] = MF::op(bf[b].im, af[a].re, acc.im[a][b]);
                    }
            }
        }
    }
}

// The output tile's current values, read before the product so that their latency overlaps it.
template <typename T, typename TL> struct GemmC {
    cx<T> v[TL::WM][TL::WN][4];
};

// Complex128 tiles and the large tiles read C in the epilogue instead: holding it through the
// product would cost the registers of a second wave per SIMD.
template <typename T, typename TL> constexpr bool kPrefetchC = sizeof(T) == 4 && TL::WM * TL::WN <= 4;

template <typename T, typename TL>
__device__ inline void gemm_cload(int64_t mr, int64_t nc, const cx<T>* __restrict__ C, int64_t ldc, GemmC<T, TL>& cv)
{
    if (!kPrefetchC<T, TL>)
        return;
    using MF = Mfma<T>;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
#pragma unroll
    for (int a = 0; a < TL::WM; a++)
#pragma unroll
        for (int b = 0; b < TL::WN; b++) {
            const int row = wm * TL::WM * 16 + a * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = wn * TL::WN * 16 + b * 16 + MF::acc_row(lane, r);
                const int64_t ok = row < mr && col < nc;
                cv.v[a][b][r] = C[ok * (row + (int64_t)col * ldc)];
            }
        }
}

// C_t[0:mr, 0:nc] = cv - acc (cv non-null) or C_t = acc (cv null); C_t(r, c) = C[r + c ldc].
template <typename T, typename TL, bool M3 = false>
__device__ inline void gemm_store(int64_t mr, int64_t nc, cx<T>* __restrict__ C, int64_t ldc,
                                  const GemmAcc<T, TL>& acc, const GemmC<T, TL>* cv)
{
    using MF = Mfma<T>;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
#pragma unroll
    for (int a = 0; a < TL::WM; a++)
#pragma unroll
        for (int b = 0; b < TL::WN; b++) {
            const int row = wm * TL::WM * 16 + a * 16 + (lane & 15);
            if (row >= mr)
                continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int col = wn * TL::WN * 16 + b * 16 + MF::acc_row(lane, r);
                const int64_t ok = col < nc;
                cx<T>* p = C + row + (int64_t)col * ldc;
                const T pre = M3 ? acc.re[a][b][r] - acc.im[a][b][r] : acc.re[a][b][r];
                const T pim = M3 ? acc.s3[a][b][r] - acc.re[a][b][r] - acc.im[a][b][r] : acc.im[a][b][r];
                cx<T> v;
                if (cv) {
                    v = kPrefetchC<T, TL> ? cv->v[a][b][r] : C[ok * (row + (int64_t)col * ldc)];
                    v.re -= pre;
                    v.im -= pim;
                } else {
                    v = cx<T>{pre, pim};
                }
                if (ok)
                    *p = v;
            }
        }
}

// One output tile (rows m0.., columns n0..) of C -= op(A) B.
template <typename T, typename TA, bool TRANSA, bool M3, typename TL>
__device__ inline void gemm_sub_tile(int64_t m0, int64_t n0, int64_t M, int64_t N, int64_t K, const cx<TA>* __restrict__ A,
                                     int64_t lda, const cx<T>* __restrict__ B, int64_t ldb, cx<T>* __restrict__ C,
                                     int64_t ldc, GemmSmem<T, TL>& sm)
{
    GemmAcc<T, TL> acc;
    GemmC<T, TL> cv;
    gemm_cload<T, TL>(M - m0, N - n0, C + m0 + n0 * ldc, ldc, cv);
    gemm_tile<T, TA, TRANSA, TL, false, M3>(M - m0, N - n0, K, TRANSA ? A + m0 * lda : A + m0, lda, B + n0 * ldb, ldb,
                                             acc, sm);
    gemm_store<T, TL, M3>(M - m0, N - n0, C + m0 + n0 * ldc, ldc, acc, &cv);
}

// C[M x N] -= op(A)[M x K] * B[K x N], complex, column-major; op(A) = A or A^T.
template <typename T, typename TA, bool TRANSA = false, bool M3 = false>
__global__ __launch_bounds__(256, kGemmOcc<T>) void cgemm_sub(int64_t M, int64_t N, int64_t K, const cx<TA>* __restrict__ A, int64_t lda,
                                                 const cx<T>* __restrict__ B, int64_t ldb, cx<T>* __restrict__ C,
                                                 int64_t ldc)
{
    using TL = typename TileBig<T>::type;
    __shared__ GemmSmem<T, TL> sm;
    gemm_sub_tile<T, TA, TRANSA, M3, TL>((int64_t)blockIdx.x * TL::BM, (int64_t)blockIdx.y * TL::BN, M, N, K, A, lda, B,
                                          ldb, C, ldc, sm);
}

// Two products of the same M x N shape in one launch, C_s -= A_s B_s for s = blockIdx.z, so that
// the tiles of both share the CUs' last wave (a K sum split over two accumulators).
template <typename T> struct GemmPair {
    const cx<T>* A[2];
    const cx<T>* B[2];
    cx<T>* C[2];
    int64_t K[2];
    int64_t lda, ldb, ldc;
};
//
// Tile order: workgroup b runs on XCD b % 8 (each XCD with its own L2) and takes tile
// s = (b % 8) ceil(T / 8) + b / 8 of the row-major tile sequence (all column tiles of a row tile
// together).  A residual product has few column tiles (361 right-hand sides: six), so each XCD
// works through a contiguous stretch of row tiles and reads each row block of Z from HBM once
// instead of once per column tile.
__device__ inline bool row_major_tile(int64_t b, int64_t gx, int64_t gy, int64_t& tm, int64_t& tn)
{
    const int64_t T = gx * gy, per = (T + 7) / 8;
    const int64_t s = (b % 8) * per + b / 8;
    if (s >= T)
        return false;
    tm = s / gy;
    tn = s % gy;
    return true;
}

template <typename T, bool M3>
__global__ __launch_bounds__(256, kGemmOcc<T>) void cgemm_sub2(int64_t M, int64_t N, int64_t gx, int64_t gy,
                                                                GemmPair<T> p)
{
    using TL = typename TileBig<T>::type;
    __shared__ GemmSmem<T, TL> sm;
    const int s = blockIdx.z;
    int64_t tm, tn;
    if (!row_major_tile(blockIdx.x, gx, gy, tm, tn))
        return;
    gemm_sub_tile<T, T, false, M3, TL>(tm * TL::BM, tn * TL::BN, M, N, p.K[s], p.A[s], p.lda, p.B[s], p.ldb, p.C[s],
                                        p.ldc, sm);
}

// The same product on the tiles (i, j) with i >= j only (row tile i, column tile j; square tiles,
// so with C's first entry on the diagonal of a symmetric matrix these are the tiles that hold its
// lower triangle).  One workgroup per such tile, numbered down the columns: column j holds the
// gx - j tiles i = j .. gx - 1, so it starts at j gx - j (j - 1) / 2.
template <typename T, bool M3 = false>
__global__ __launch_bounds__(256, kGemmOcc<T>) void cgemm_sub_lower(int64_t M, int64_t N, int64_t K, int64_t gx, int64_t gy,
                                                                    const cx<T>* __restrict__ A, int64_t lda,
                                                                    const cx<T>* __restrict__ B, int64_t ldb,
                                                                    cx<T>* __restrict__ C, int64_t ldc)
{
    using TL = typename TileBig<T>::type;
    static_assert(TL::BM == TL::BN, "square tiles");
    __shared__ GemmSmem<T, TL> sm;
    const int64_t t = blockIdx.x;
    auto start = [gx](int64_t j) { return j * gx - j * (j - 1) / 2; };
    const double b = 2.0 * (double)gx + 1.0;
    int64_t j = (int64_t)(0.5 * (b - sqrt(b * b - 8.0 * (double)t)));
    j = j < 0 ? 0 : (j >= gy ? gy - 1 : j);
    while (j > 0 && start(j) > t)
        j--;
    while (j + 1 < gy && start(j + 1) <= t)
        j++;
    const int64_t i = j + (t - start(j));
    gemm_sub_tile<T, T, false, M3, TL>(i * TL::BM, j * TL::BN, M, N, K, A, lda, B, ldb, C, ldc, sm);
}

// W blocks for the forward solves: for row tile t > chunk j (64 x 64 each)
//   W(t, j) = L(t, j) Linv_j,  stored at the same place in WV (leading dimension ldw).
// Grid (row tiles, chunks).  (The backward solve does not use U's diagonal-block inverses: a
// product with an explicit U_jj^{-1} loses accuracy in proportion to cond(U_jj) on EFIE
// matrices, so U_jj is applied by substitution.)
template <typename T>
__global__ __launch_bounds__(256, kGemmOcc<T>) void make_w(int64_t n, const cx<T>* __restrict__ A, int64_t lda,
                                              const cx<T>* __restrict__ inv, cx<T>* __restrict__ WV, int64_t ldw)
{
    const int64_t t = blockIdx.x, j = blockIdx.y;
    if (t <= j)
        return;
    __shared__ GemmSmem<T, Tile64> sm;
    const int64_t m0 = t * 64, k = j * 64;
    const int64_t w = (n - k) < 64 ? (n - k) : 64;
    GemmAcc<T, Tile64> acc;
    gemm_tile<T, T, false, Tile64>(n - m0, w, w, A + m0 + k * lda, lda, inv + (size_t)(2 * j) * 4096, 64, acc, sm);
    gemm_store<T, Tile64>(n - m0, w, WV + m0 + k * ldw, ldw, acc, nullptr);
}

// One chunk of a blocked triangular solve as a single product.  Rows [r0, r0 + M) in 64-row tiles;
// tile ct holds the chunk's own rows and computes Xo = inv * Y (assign), every other tile computes
// C -= WV(rows, chunk) * Y.  Y is the chunk's K x nrhs right-hand side (leading dimension ldy);
// it is only read, and C and Xo do not overlap it.
template <typename T>
__global__ __launch_bounds__(256, kGemmOcc<T>) void solve_chunk(int64_t r0, int64_t M, int64_t nrhs, int64_t K, int64_t ct,
                                                   const cx<T>* __restrict__ inv, const cx<T>* __restrict__ WVc,
                                                   int64_t ldw, const cx<T>* __restrict__ Y, int64_t ldy,
                                                   cx<T>* __restrict__ C, int64_t ldc, cx<T>* __restrict__ Xo,
                                                   int64_t ldx)
{
    __shared__ GemmSmem<T, Tile64> sm;
    const int64_t bx = blockIdx.x, m0 = bx * GBM, n0 = (int64_t)blockIdx.y * GBN;
    const bool own = bx == ct;
    const cx<T>* At = own ? inv : WVc + r0 + m0;
    GemmAcc<T, Tile64> acc;
    GemmC<T, Tile64> cv;
    if (!own)
        gemm_cload<T, Tile64>(M - m0, nrhs - n0, C + r0 + m0 + n0 * ldc, ldc, cv);
    gemm_tile<T, T, false, Tile64>(M - m0, nrhs - n0, K, At, own ? 64 : ldw, Y + n0 * ldy, ldy, acc, sm);
    if (own)
        gemm_store<T, Tile64>(M - m0, nrhs - n0, Xo + r0 + m0 + n0 * ldx, ldx, acc, nullptr);
    else
        gemm_store<T, Tile64>(M - m0, nrhs - n0, C + r0 + m0 + n0 * ldc, ldc, acc, &cv);
}

// ----------------------------------------------------------------------------------------------
// Row interchanges.  The swaps of one panel (rows k .. k+w-1 exchanged in turn with ipiv[k + c])
// are reduced once, when the panel is factored, to the permutation they make of the rows they
// touch: a list of (row, source row) pairs, meaning "row receives the content source held before
// the panel".  Applying a panel's interchanges to any set of columns is then one gather.
//
// Replay (one wavefront): slot s < w is panel row k + s, slots w, w+1, ... are the rows below the
// panel in order of first appearance (held one per lane, found by ballot); slot s keeps the row
// whose content it currently holds in register sr<s / 64> of lane s % 64.  Slot indices are
// wave-uniform, so every access is a readlane and the chain of swaps never waits on memory.
// ----------------------------------------------------------------------------------------------
constexpr int PERM_STRIDE = 1 + 2 * 2 * LU_NB_MAX; // ints per panel: count, then (row, src) pairs

// Slot access for the replay: slot sl lives in sr<sl / 64> of lane sl % 64 (sl wave-uniform).
#define PERM_GET(sl)                                                                                        \
    ((sl) < 64 ? __builtin_amdgcn_readlane(sr0, (sl) & 63)                                                  \
               : (sl) < 128 ? __builtin_amdgcn_readlane(sr1, (sl) & 63)                                     \
                            : (sl) < 192 ? __builtin_amdgcn_readlane(sr2, (sl) & 63)                        \
                                         : __builtin_amdgcn_readlane(sr3, (sl) & 63))
#define PERM_PUT(sl, v)                                                                                     \
    do {                                                                                                    \
        if (lane == ((sl) & 63)) {                                                                          \
            if ((sl) < 64)                                                                                  \
                sr0 = (v);                                                                                  \
            else if ((sl) < 128)                                                                            \
                sr1 = (v);                                                                                  \
            else if ((sl) < 192)                                                                            \
                sr2 = (v);                                                                                  \
            else                                                                                            \
                sr3 = (v);                                                                                  \
        }                                                                                                   \
    } while (0)

// Called by one full wavefront.  piv[c] (LDS or global) = absolute pivot row of panel column c.
// s_row must hold >= 2 w ints of scratch.  Writes the list of moved rows to out (PERM_STRIDE ints).
__device__ inline void replay_panel_swaps(int k, int w, const int* piv, int* s_row, int* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int pv0 = lane < w ? piv[lane] : 0; // pivots, two per lane (w <= 128)
    const int pv1 = lane + 64 < w ? piv[64 + lane] : 0;
    int sr0 = k + lane, sr1 = k + 64 + lane, sr2 = -1, sr3 = -1;
    s_row[lane] = k + lane;
    s_row[64 + lane] = k + 64 + lane;
    int e0 = -1, e1 = -1; // rows of the extra slots w + lane and w + 64 + lane
    int cnt = 0;
    const int kw = k + w;
    for (int c = 0; c < w; c++) {
        const int p = c < 64 ? __builtin_amdgcn_readlane(pv0, c) : __builtin_amdgcn_readlane(pv1, c - 64);
        if (p == k + c)
            continue;
        int sp;
        if (p < kw) {
            sp = p - k;
        } else {
            const unsigned long long m0 = __ballot(e0 == p), m1 = __ballot(e1 == p);
            if (m0) {
                sp = w + __ffsll((long long)m0) - 1;
            } else if (m1) {
                sp = w + 64 + __ffsll((long long)m1) - 1;
            } else {
                sp = w + cnt;
                if (cnt < 64) {
                    if (lane == cnt)
                        e0 = p;
                } else if (lane == cnt - 64) {
                    e1 = p;
                }
                PERM_PUT(sp, p);
                cnt++;
            }
        }
        const int a = PERM_GET(c), b = PERM_GET(sp);
        PERM_PUT(c, b);
        PERM_PUT(sp, a);
    }
    // rows of the extra slots
    if (lane < cnt)
        s_row[w + lane] = e0;
    if (lane + 64 < cnt)
        s_row[w + 64 + lane] = e1;
    // compact the moved slots (source differs from the slot's row) into the list
    const int L = w + cnt;
    int base = 0;
    for (int j = 0; j < 4; j++) {
        const int sl = 64 * j + lane;
        const int src = j == 0 ? sr0 : j == 1 ? sr1 : j == 2 ? sr2 : sr3;
        const int row = sl < L ? s_row[sl] : -1;
        const bool moved = sl < L && src != row;
        const unsigned long long m = __ballot(moved);
        if (moved) {
            const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
            out[1 + 2 * pos] = row;
            out[2 + 2 * pos] = src;
        }
        base += __popcll(m);
    }
    if (lane == 0)
        out[0] = base;
}

// Applies one panel's list (perm, PERM_STRIDE ints) to columns [col0, col0 + ncols).
// One workgroup per LASWP_COLS columns; all reads of a column precede its writes (through LDS).
constexpr int LASWP_COLS = 8;

// Interchange lists of the composite panels (see lu_factor) from their pivots, one wavefront per
// panel: panel p covers columns [p nb, p nb + w), w = min(nb, n - p nb).
__global__ __launch_bounds__(64) void build_perm_lists(int nb, int64_t n, const int* __restrict__ ipiv,
                                                       int* __restrict__ out)
{
    __shared__ int s_row[2 * LU_NB_MAX];
    const int64_t k = (int64_t)blockIdx.x * nb;
    const int w = (int)((n - k) < nb ? (n - k) : nb);
    replay_panel_swaps((int)k, w, ipiv + k, s_row, out + (size_t)blockIdx.x * PERM_STRIDE);
}

// Row interchanges of one panel on columns [col0, col0 + ncols): one gather through LDS.  With a
// second list (perm2, applied after perm) the two are composed first: row r of perm2 receives what
// its source held before perm, i.e. the content of perm's source of that row; rows of perm that
// perm2 also moves are dropped from perm's part.
template <typename T>
__global__ __launch_bounds__(256) void laswp(cx<T>* __restrict__ A, int64_t lda, int64_t col0, int64_t ncols,
                                             const int* __restrict__ perm, const int* __restrict__ perm2)
{
    __shared__ int2 s_mv[2 * LU_NB_MAX];
    __shared__ int s_x2[2 * LU_NB_MAX]; // rows of perm2
    static_assert(2 * LU_NB_MAX <= 256, "one list entry per thread");
    const int tid = threadIdx.x;
    // the count and every entry slot read together (a list occupies PERM_STRIDE ints)
    const int L1 = perm[0];
    const int2 p1{perm[1 + 2 * tid], perm[2 + 2 * tid]};
    int2 a{0, 0};
    int L2 = 0;
    if (perm2) {
        L2 = perm2[0];
        a = int2{perm2[1 + 2 * tid], perm2[2 + 2 * tid]};
    }
    if (tid < L1)
        s_mv[tid] = p1;
    if (tid < L2)
        s_x2[tid] = a.x;
    int L = L1;
    if (perm2) {
        __syncthreads();
        // this thread's entries: a of perm2 (source composed with perm), b of perm (or dropped)
        int2 b = p1;
        if (tid < L2)
            for (int j = 0; j < L1; j++)
                if (s_mv[j].x == a.y)
                    a.y = s_mv[j].y;
        if (tid < L1)
            for (int j = 0; j < L2; j++)
                if (s_x2[j] == b.x)
                    b.x = -1;
        __syncthreads();
        if (tid < L1)
            s_mv[tid] = b;
        if (tid < L2)
            s_mv[L1 + tid] = a;
        L = L1 + L2;
    }
    __syncthreads();
    const int64_t c0 = col0 + (int64_t)blockIdx.x * LASWP_COLS;
    const int nc = (int)((col0 + ncols - c0) < LASWP_COLS ? (col0 + ncols - c0) : LASWP_COLS);
    // all of a thread's gathers in flight at once (unconditional reads of the column's first
    // entry where an element does not move), then the scatters
    constexpr int PER = 2 * LU_NB_MAX * LASWP_COLS / 256;
    T vr[PER], vi[PER];
    int64_t dst[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = tid + 256 * t;
        const int m = e % L, cc = e / L;
        const bool mv = e < L * LASWP_COLS && cc < nc && s_mv[m].x >= 0;
        const cx<T> x = A[(mv ? (int64_t)s_mv[m].y : 0) + (c0 + (mv ? cc : 0)) * lda];
        vr[t] = x.re;
        vi[t] = x.im;
        dst[t] = mv ? (int64_t)s_mv[m].x + (c0 + cc) * lda : -1;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; t++)
        if (dst[t] >= 0)
            A[dst[t]] = cx<T>{vr[t], vi[t]};
}

template <typename T>
static void launch_laswp(cx<T>* A, int64_t lda, int64_t col0, int64_t ncols, const int* perm, hipStream_t st,
                         const int* perm2 = nullptr)
{
    if (ncols <= 0 || !perm) // no list: no interchanges
        return;
    launch_kernel(&laswp<T>, dim3((unsigned)((ncols + LASWP_COLS - 1) / LASWP_COLS)), dim3(256), 0, st, A, lda,
                  col0, ncols, perm, perm2);
}

// ----------------------------------------------------------------------------------------------
// Panel factorisation.  G workgroups, workgroup g owns panel rows [g*R, g*R + R) in LDS, one row
// per thread.  Per column every workgroup publishes its local pivot candidate (value, row index,
// row contents) and, if it owns it, the current diagonal row; every workgroup then reads all G
// candidates, picks the same winner (first max of |re|+|im|), reads the winner's row and performs
// the interchange on its own rows.
//
// Column c's rank-1 update is split: each workgroup first updates column c + 1 alone, finds and
// publishes its candidate for c + 1 (the published rows carry the rest of the update, computed as
// they are written), and only then updates columns c + 2 .. w - 1, while the other workgroups'
// candidates travel.  The pivot sequence and the arithmetic of every entry are those of the
// unsplit right-looking order.
//
// The exchange area holds 64-bit words {sequence number (high 32 bits), payload (low 32 bits)},
// written and read with relaxed agent-scope atomics.  A reader accepts a word only when it carries
// the current column's sequence number (k + c + 1, increasing over the factorisation), so no
// ordering between different words is needed and no fence (which on gfx950 writes back and
// invalidates the XCD's whole L2) is issued.  Two slots per group alternate by column parity: a
// group rewrites slot c & 1 only after every group has published column c + 1, i.e. after every
// reader of column c is done.
// ----------------------------------------------------------------------------------------------
constexpr int LL_HDR = 4;                 // header words per group: value lo, value hi, row
constexpr int LL_ROWW = 4 * LU_NB_MAX;    // words per published row (complex128: 4 per element)

__host__ __device__ inline size_t ll_words(int gmax)
{
    return (size_t)2 * gmax * LL_HDR + (size_t)2 * gmax * LL_ROWW + (size_t)2 * LL_ROWW;
}

struct PanelSync {
    unsigned long long* words; // ll_words(gmax) words, zeroed before the factorisation
    int gmax;
    int* abort_flag; // set if a wait exceeds its spin limit
    int* info;       // first zero pivot (1-based global column), 0 if none
    __device__ unsigned long long* hdr(int par, int g) const { return words + ((size_t)par * gmax + g) * LL_HDR; }
    __device__ unsigned long long* row(int par, int g) const
    {
        return words + (size_t)2 * gmax * LL_HDR + ((size_t)par * gmax + g) * LL_ROWW;
    }
    __device__ unsigned long long* diag(int par) const
    {
        return words + (size_t)2 * gmax * LL_HDR + (size_t)2 * gmax * LL_ROWW + (size_t)par * LL_ROWW;
    }
};

constexpr long SPIN_LIMIT = 40000000; // ~seconds; a wait this long means a missing workgroup

__device__ inline void ll_put(unsigned long long* p, unsigned seq, unsigned payload)
{
    __hip_atomic_store(p, ((unsigned long long)seq << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline unsigned long long ll_load(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait of one lane: returns false (and raises the abort flag) after SPIN_LIMIT polls, or
// when another workgroup has raised it.
__device__ inline bool ll_spin(long& spins, int* abort_flag)
{
    if (++spins > SPIN_LIMIT) {
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    if ((spins & 255) == 0 && __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return false;
    __builtin_amdgcn_s_sleep(1);
    return true;
}

// Pivot-search key of an entry: |re| + |im| (izamax's measure), with NaN taken as +inf so that a
// NaN is chosen and published as a pivot like any other entry and spreads through the factors
// (the refinement then reports it), instead of leaving a column with no published candidate.
template <typename T> __device__ inline double pivot_key(cx<T> a)
{
    const double v = (double)cabs1(a);
    return v != v ? __builtin_huge_val() : v;
}

__device__ inline void argmax_merge(double& v, int& i, double ov, int oi)
{
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

// Wavefront reductions with DPP (quad_perm, row_ror, row_bcast): the result is read from lane 63
// and returned to every lane.
template <int CTRL, int RMASK> __device__ inline double dpp_f64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, RMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, RMASK, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ inline double wave_max(double v)
{
    v = fmax(v, dpp_f64<0xb1, 0xf>(v));  // quad_perm [1,0,3,2]
    v = fmax(v, dpp_f64<0x4e, 0xf>(v));  // quad_perm [2,3,0,1]
    v = fmax(v, dpp_f64<0x124, 0xf>(v)); // row_ror:4
    v = fmax(v, dpp_f64<0x128, 0xf>(v)); // row_ror:8
    v = fmax(v, dpp_f64<0x142, 0xa>(v)); // row_bcast:15
    v = fmax(v, dpp_f64<0x143, 0xc>(v)); // row_bcast:31
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ inline int wave_min(int v)
{
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xb1, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4e, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xa, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// First maximum over a wavefront of (value, index) pairs: the largest value, and among equal values
// the smallest index; also returns the lane that holds it.
__device__ inline void wave_argmax(double v, int vi, double& bv, int& bi, int& blane)
{
    bv = wave_max(v);
    bi = wave_min(v == bv ? vi : 0x7fffffff);
    const unsigned long long m = __ballot(v == bv && vi == bi);
    blane = __ffsll((long long)m) - 1;
}

// 32-bit word q of a complex value, in memory order (re low word first).
__device__ inline unsigned cx_word(cx<float> x, int q) { return __float_as_uint(q == 0 ? x.re : x.im); }
__device__ inline unsigned cx_word(cx<double> x, int q)
{
    const long long b = __double_as_longlong(q < 2 ? x.re : x.im);
    return (q & 1) ? (unsigned)((unsigned long long)b >> 32) : (unsigned)b;
}
template <typename T>
__global__ __launch_bounds__(256) void panel_factor(cx<T>* __restrict__ A, int64_t lda, int64_t k, int64_t m, int w,
                                                    int R, int G, int* __restrict__ ipiv, PanelSync ps,
                                                    int* __restrict__ perm)
{
    extern __shared__ unsigned char smem_raw[];
    const int RS = R + 1; // slab column stride: a row of the slab spans every LDS bank
    cx<T>* slab = reinterpret_cast<cx<T>*>(smem_raw); // [w][RS] column-major
    cx<T>* prow = slab + (size_t)w * RS;               // [w] pivot row
    cx<T>* drow = prow + w;                            // [w] previous content of row c
    __shared__ double rv[4];
    __shared__ int ri[4];
    __shared__ int s_piv, s_gw;
    __shared__ double s_pval;
    __shared__ int s_pivots[LU_NB_MAX]; // absolute pivot rows of the panel (workgroup 0)
    __shared__ int s_scratch[2 * LU_NB_MAX];
    constexpr int CW = (int)(sizeof(cx<T>) / 4); // 32-bit words per complex element

    const int g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r0 = g * R;                                      // first panel row of this workgroup
    const int nr = (int)((m - r0) < R ? (m - r0) : R);         // rows owned (one per thread)
    cx<T>* Ap = A + k + k * lda;                               // panel origin
    if (tid < nr) {
        // sixteen columns of loads in flight per thread
        const cx<T>* src = Ap + r0 + tid;
        int c0 = 0;
        for (; c0 + 16 <= w; c0 += 16) {
            cx<T> v[16];
#pragma unroll
            for (int q = 0; q < 16; q++)
                v[q] = src[(int64_t)(c0 + q) * lda];
#pragma unroll
            for (int q = 0; q < 16; q++)
                slab[(size_t)(c0 + q) * RS + tid] = v[q];
        }
        for (; c0 < w; c0++)
            slab[(size_t)c0 * RS + tid] = src[(int64_t)c0 * lda];
    }
    __syncthreads();

    const bool own = tid < nr;
    const int grow = r0 + tid; // panel-relative row of this thread
    const int RW = w * CW;     // words of a published row

    // Local pivot candidate of column c: the first maximum of v over the workgroup (v < 0 for rows
    // that do not take part); every thread returns with the same (lb, li).
    double lb;
    int li;
    auto candidate = [&](double v, int vi) {
        double bv;
        int bi, bl;
        wave_argmax(v, vi, bv, bi, bl);
        if (lane == 0) {
            rv[wv] = bv;
            ri[wv] = bi;
        }
        __syncthreads();
        lb = rv[0];
        li = ri[0];
        for (int x = 1; x < 4; x++)
            argmax_merge(lb, li, rv[x], ri[x]);
    };
    // Entry (j, r) of the slab with the rank-1 update of column cp applied when that update is
    // still pending on column j (pend: the update of column cp has reached only column cp + 1).
    bool pend = false;
    int cp = 0;
    auto current = [&](int j, int r) {
        cx<T> x = slab[(size_t)j * RS + r];
        if (pend && j > cp + 1) {
            const cx<T> l = slab[(size_t)cp * RS + r], u = prow[j];
            x.re -= l.re * u.re - l.im * u.im;
            x.im -= l.re * u.im + l.im * u.re;
        }
        return x;
    };
    // Publish column cn: header, candidate row, and row cn itself if this workgroup holds it.
    auto publish = [&](int cn) {
        const int pb = cn & 1;
        const unsigned sq = (unsigned)(k + cn + 1);
        if (tid == 0) {
            const unsigned long long vb = (unsigned long long)__double_as_longlong(lb);
            unsigned long long* h = ps.hdr(pb, g);
            ll_put(h + 0, sq, (unsigned)vb);
            ll_put(h + 1, sq, (unsigned)(vb >> 32));
            ll_put(h + 2, sq, (unsigned)li);
        }
        if (lb >= 0.0)
            for (int t = tid; t < RW; t += 256)
                ll_put(ps.row(pb, g) + t, sq, cx_word(current(t / CW, li - r0), t % CW));
        if (cn >= r0 && cn < r0 + nr)
            for (int t = tid; t < RW; t += 256)
                ll_put(ps.diag(pb) + t, sq, cx_word(current(t / CW, cn - r0), t % CW));
    };

    candidate(own ? pivot_key(slab[tid]) : -1.0, own ? grow : 0x7fffffff);
    if (G > 1)
        publish(0);

    for (int c = 0; c < w; c++) {
        const int par = c & 1;
        int p;
        double pval;
        if (G > 1) {
            const unsigned seq = (unsigned)(k + c + 1);
            // ---- all candidates (wave 0, up to 4 groups per lane, all words in flight together)
            bool ok = true;
            if (wv == 0) {
                unsigned long long hv[4][3];
                long spins = 0;
                for (;;) {
                    bool ready = true;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int x = lane + 64 * j;
                        if (x < G) {
                            const unsigned long long* h = ps.hdr(par, x);
#pragma unroll
                            for (int q = 0; q < 3; q++) {
                                hv[j][q] = ll_load(h + q);
                                ready = ready && (unsigned)(hv[j][q] >> 32) == seq;
                            }
                        }
                    }
                    if (ready || !(ok = ll_spin(spins, ps.abort_flag)))
                        break;
                }
                double b = -1.0;
                int bi = 0x7fffffff, bg = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int x = lane + 64 * j;
                    if (x < G) {
                        const double cv = __longlong_as_double(
                            (long long)(((hv[j][1] & 0xffffffffull) << 32) | (hv[j][0] & 0xffffffffull)));
                        const int ci = (int)(unsigned)hv[j][2];
                        if (cv > b || (cv == b && ci < bi)) {
                            b = cv;
                            bi = ci;
                            bg = x;
                        }
                    }
                }
                double wb;
                int wi, wl;
                wave_argmax(b, bi, wb, wi, wl);
                const int wg = __builtin_amdgcn_readlane(bg, wl);
                if (lane == 0) {
                    s_pval = wb;
                    s_piv = wi;
                    s_gw = wg;
                }
            }
            if (__syncthreads_or(!ok))
                return;
            p = s_piv;
            pval = s_pval;
            // ---- winner row, and the previous row c when row p is ours (both words in flight)
            const bool need_diag = pval > 0.0 && p != c && p >= r0 && p < r0 + nr;
            const unsigned long long* wrow = ps.row(par, s_gw);
            for (int t = tid; t < RW && ok; t += 256) {
                long spins = 0;
                for (;;) {
                    const unsigned long long a = ll_load(wrow + t);
                    const unsigned long long d = need_diag ? ll_load(ps.diag(par) + t) : ((unsigned long long)seq << 32);
                    if ((unsigned)(a >> 32) == seq && (unsigned)(d >> 32) == seq) {
                        reinterpret_cast<unsigned*>(prow)[t] = (unsigned)a;
                        if (need_diag)
                            reinterpret_cast<unsigned*>(drow)[t] = (unsigned)d;
                        break;
                    }
                    if (!(ok = ll_spin(spins, ps.abort_flag)))
                        break;
                }
            }
            if (__syncthreads_or(!ok))
                return;
            if (pval > 0.0) {
                if (c >= r0 && c < r0 + nr && tid < w)
                    slab[(size_t)tid * RS + (c - r0)] = prow[tid];
                if (need_diag && tid < w)
                    slab[(size_t)tid * RS + (p - r0)] = drow[tid];
            }
            __syncthreads();
        } else {
            __syncthreads(); // the previous column's update is complete
            p = li;
            pval = lb;
            if (pval > 0.0 && p != c && tid < w) {
                const cx<T> t = slab[(size_t)tid * RS + c];
                slab[(size_t)tid * RS + c] = slab[(size_t)tid * RS + p];
                slab[(size_t)tid * RS + p] = t;
            }
            __syncthreads();
            if (tid < w)
                prow[tid] = slab[(size_t)tid * RS + c];
            __syncthreads();
        }
        if (g == 0 && tid == 0) {
            ipiv[k + c] = (int)(k + (pval > 0.0 ? p : c));
            s_pivots[c] = (int)(k + (pval > 0.0 ? p : c));
            if (!(pval > 0.0) && *ps.info == 0)
                *ps.info = (int)(k + c + 1);
        }
        // ---- multiplier of this thread's row and the update of column c + 1 only; the next
        // column's candidate is published before the rest of the rank-1 update, which then runs
        // while the other workgroups' candidates are on their way
        const bool mine = pval > 0.0 && own && grow > c;
        cx<T> l{};
        if (mine) {
            const cx<T> rp = cdiv(cx<T>{1, 0}, prow[c]);
            l = slab[(size_t)c * RS + tid] * rp;
            slab[(size_t)c * RS + tid] = l;
        }
        if (c + 1 < w) {
            double v = -1.0;
            if (own && grow > c) {
                cx<T> a1 = slab[(size_t)(c + 1) * RS + tid];
                if (mine) {
                    const cx<T> u = prow[c + 1];
                    a1.re -= l.re * u.re - l.im * u.im;
                    a1.im -= l.re * u.im + l.im * u.re;
                    slab[(size_t)(c + 1) * RS + tid] = a1;
                }
                v = pivot_key(a1);
            }
            candidate(v, v >= 0.0 ? grow : 0x7fffffff);
            if (G > 1) {
                pend = pval > 0.0;
                cp = c;
                publish(c + 1);
                __syncthreads(); // published rows are read before the update below rewrites them
            }
        }
        if (R <= 128) {
            // two threads a row (the slab has at most 128 rows): the row's owner and thread
            // owner + 128 take alternate groups of eight columns, the group after the last pair
            // and the remainder going to the side with fewer groups.  The multiplier is read from
            // the slab, where the owner wrote it before the candidate's barrier.
            const int rr = tid & 127, side = tid >> 7;
            if (pval > 0.0 && rr < nr && r0 + rr > c) {
                const cx<T> lr = slab[(size_t)c * RS + rr];
                const int j0 = c + 2, ng = (w - j0) > 0 ? (w - j0) / 8 : 0;
                for (int gi = side; gi < ng; gi += 2) {
                    const int j = j0 + 8 * gi;
                    cx<T> a[8], u[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        a[q] = slab[(size_t)(j + q) * RS + rr];
                        u[q] = prow[j + q];
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        a[q].re -= lr.re * u[q].re - lr.im * u[q].im;
                        a[q].im -= lr.re * u[q].im + lr.im * u[q].re;
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        slab[(size_t)(j + q) * RS + rr] = a[q];
                }
                if (side == (ng & 1)) {
                    for (int j = j0 + 8 * ng; j < w; j++) {
                        const cx<T> u = prow[j];
                        cx<T>& a = slab[(size_t)j * RS + rr];
                        a.re -= lr.re * u.re - lr.im * u.im;
                        a.im -= lr.re * u.im + lr.im * u.re;
                    }
                }
            }
        } else if (mine) {
            int j = c + 2;
            // eight columns at a time, all LDS reads of a group in flight together
            for (; j + 8 <= w; j += 8) {
                cx<T> a[8], u[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    a[q] = slab[(size_t)(j + q) * RS + tid];
                    u[q] = prow[j + q];
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    a[q].re -= l.re * u[q].re - l.im * u[q].im;
                    a[q].im -= l.re * u[q].im + l.im * u[q].re;
                }
#pragma unroll
                for (int q = 0; q < 8; q++)
                    slab[(size_t)(j + q) * RS + tid] = a[q];
            }
            for (; j < w; j++) {
                const cx<T> u = prow[j];
                cx<T>& a = slab[(size_t)j * RS + tid];
                a.re -= l.re * u.re - l.im * u.im;
                a.im -= l.re * u.im + l.im * u.re;
            }
        }
    }
    __syncthreads();
    if (g == 0 && wv == 0)
        replay_panel_swaps((int)k, w, s_pivots, s_scratch, perm);
#pragma unroll 16
    for (int c = 0; c < w; c++)
        if (tid < nr)
            Ap[(r0 + tid) + (int64_t)c * lda] = slab[(size_t)c * RS + tid];
}

// Row interchanges of one panel undone (swaps applied in reverse order) on columns [col0, col0 + ncols);
// used by the transposed solve.
template <typename T>
__global__ __launch_bounds__(256) void laswp_reverse(cx<T>* __restrict__ A, int64_t lda, int64_t col0, int64_t ncols,
                                                     int64_t k, int w, const int* __restrict__ ipiv)
{
    const int64_t j = col0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= col0 + ncols)
        return;
    cx<T>* col = A + j * lda;
    for (int c = w - 1; c >= 0; c--) {
        const int64_t p = ipiv[k + c];
        if (p != k + c) {
            cx<T> t = col[k + c];
            col[k + c] = col[p];
            col[p] = t;
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Triangular solve with a w x w diagonal block (w <= 64) on a w x ncols block: Bout = E^{-1} Bin
// (Bout may equal Bin), E = D, or D^T with TRANSD; UPPER: E upper (backward substitution), else
// lower (forward); UNIT takes the diagonal as one.
//
// One wavefront per NC columns, lane j holding row j of each: the row steps run in registers with
// no barrier.  Step i reads x_i = b_i / E_ii from lane i (readlane) and subtracts E_ji x_i from
// every row j that depends on it; the coefficients sit in LDS, sE[i][j] = E_ji for the dependent
// rows and zero elsewhere (the diagonal included), so lane i's b_i is left alone after its step
// and every lane scales by 1 / E_jj once at the end.  Row j of sE is stored with its columns
// permuted (slot j ^ i), so both the coalesced load and the transposed one are free of bank
// conflicts.
// ----------------------------------------------------------------------------------------------
constexpr int TRSM_WAVES = 4;

__device__ inline cx<float> read_lane(cx<float> v, int i)
{
    return {__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.re), i)),
            __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.im), i))};
}
__device__ inline double read_lane(double v, int i)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, i), hi = __builtin_amdgcn_readlane((int)(b >> 32), i);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline cx<double> read_lane(cx<double> v, int i) { return {read_lane(v.re, i), read_lane(v.im, i)}; }

template <typename T, bool UPPER, bool TRANSD, bool UNIT, int NC>
__global__ __launch_bounds__(256) void trsm_block(int w, const cx<T>* __restrict__ D, int64_t ldd, const cx<T>* Bin,
                                                  int64_t ldi, cx<T>* Bout, int64_t ldo, int64_t ncols)
{
    __shared__ cx<T> sE[64][64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t cb = ((int64_t)blockIdx.x * TRSM_WAVES + wv) * NC; // first column of this wavefront
    // every load in flight at once: the block (16 entries per thread, unconditional reads of an
    // in-range entry, masked at the LDS store), the right-hand sides and the diagonal
    cx<T> dv[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int e = tid + 256 * t, r = e & 63, c = e >> 6; // D(r, c); r fastest, coalesced
        dv[t] = D[(r < w ? r : 0) + (int64_t)(c < w ? c : 0) * ldd];
    }
    cx<T> b[NC];
#pragma unroll
    for (int q = 0; q < NC; q++)
        b[q] = Bin[(lane < w ? lane : 0) + (cb + q < ncols ? cb + q : 0) * ldi];
    const cx<T> dd = D[(lane < w ? lane : 0) * (ldd + 1)];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        const int e = tid + 256 * t, r = e & 63, c = e >> 6;
        const int i = TRANSD ? r : c, j = TRANSD ? c : r; // E_ji
        const bool dep = UPPER ? j < i : j > i;
        sE[i][j ^ i] = (dep && r < w && c < w) ? dv[t] : cx<T>{0, 0};
    }
#pragma unroll
    for (int q = 0; q < NC; q++)
        if (!(lane < w && cb + q < ncols))
            b[q] = cx<T>{0, 0};
    cx<T> rd{1, 0};
    if (!UNIT && lane < w)
        rd = cdiv(cx<T>{1, 0}, dd);
    __syncthreads();
    // coefficients of the next PF steps in flight while a step's readlane chain runs
    constexpr int PF = 4;
    cx<T> ep[PF];
#pragma unroll
    for (int s = 0; s < PF; s++) {
        const int i = UPPER ? 63 - s : s;
        ep[s] = sE[i][lane ^ i];
    }
#pragma unroll
    for (int s = 0; s < 64; s++) {
        const int i = UPPER ? 63 - s : s;
        const cx<T> e = ep[s % PF];
        if (s + PF < 64) {
            const int i2 = UPPER ? 63 - (s + PF) : s + PF;
            ep[s % PF] = sE[i2][lane ^ i2];
        }
        if (i < w) {
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const cx<T> x = read_lane(UNIT ? b[q] : b[q] * rd, i);
                b[q].re -= e.re * x.re - e.im * x.im;
                b[q].im -= e.re * x.im + e.im * x.re;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NC; q++)
        if (lane < w && cb + q < ncols)
            Bout[lane + (cb + q) * ldo] = UNIT ? b[q] : b[q] * rd;
}

template <typename T, bool UPPER, bool TRANSD = false, bool UNIT = !UPPER>
static void launch_trsm(int w, const cx<T>* D, int64_t ldd, const cx<T>* Bin, int64_t ldi, cx<T>* Bout, int64_t ldo,
                        int64_t ncols, hipStream_t st)
{
    if (ncols <= 0)
        return;
    // wide blocks (the factorisation's row of U) share each coefficient load over four columns a
    // wavefront; narrow ones (the solves' right-hand sides) take one column a wavefront
    if (ncols > 1024) {
        constexpr int NC = 4;
        launch_kernel(&trsm_block<T, UPPER, TRANSD, UNIT, NC>,
                      dim3((unsigned)((ncols + TRSM_WAVES * NC - 1) / (TRSM_WAVES * NC))), dim3(256), 0, st, w, D, ldd,
                      Bin, ldi, Bout, ldo, ncols);
    } else {
        launch_kernel(&trsm_block<T, UPPER, TRANSD, UNIT, 1>, dim3((unsigned)((ncols + TRSM_WAVES - 1) / TRSM_WAVES)),
                      dim3(256), 0, st, w, D, ldd, Bin, ldi, Bout, ldo, ncols);
    }
}

// Inverses of the 64 x 64 diagonal chunks of the factors, for the solves: chunk ch covers rows and
// columns [64 ch, 64 ch + 64).  Linv = L_cc^{-1} (unit lower), Uinv = U_cc^{-1} (upper).  Stored
// dense column-major, 64 x 64 per chunk: inv[(2 ch + 0) * 4096] = Linv, [(2 ch + 1) * 4096] = Uinv.
template <typename T>
__global__ __launch_bounds__(256) void diag_inverse(const cx<T>* __restrict__ A, int64_t n, int64_t lda,
                                                    cx<T>* __restrict__ inv)
{
    __shared__ cx<T> sD[64][65];
    __shared__ cx<T> sX[64][65];
    const int64_t ch = blockIdx.x, k = ch * 64;
    const int w = (int)((n - k) < 64 ? (n - k) : 64);
    const int tid = threadIdx.x, col = tid & 63, rg = tid >> 6;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e % 64, c = e / 64;
        sD[r][c] = (r < w && c < w) ? A[(k + r) + (k + c) * lda] : cx<T>{0, 0};
        sX[r][c] = cx<T>{(T)(r == c && r < w ? 1 : 0), 0};
    }
    __syncthreads();
    for (int r = 0; r < w; r++) {
        const cx<T> x = sX[r][col];
        for (int rr = r + 1 + rg; rr < w; rr += 4) {
            const cx<T> l = sD[rr][r];
            sX[rr][col].re -= l.re * x.re - l.im * x.im;
            sX[rr][col].im -= l.re * x.im + l.im * x.re;
        }
        __syncthreads();
    }
    cx<T>* Li = inv + (size_t)(2 * ch) * 4096;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e % 64, c = e / 64;
        Li[e] = sX[r][c];
        sX[r][c] = cx<T>{(T)(r == c && r < w ? 1 : 0), 0};
    }
    __syncthreads();
    for (int r = w - 1; r >= 0; r--) {
        if (rg == 0)
            sX[r][col] = cdiv(sX[r][col], sD[r][r]);
        __syncthreads();
        const cx<T> x = sX[r][col];
        for (int rr = rg; rr < r; rr += 4) {
            const cx<T> u = sD[rr][r];
            sX[rr][col].re -= u.re * x.re - u.im * x.im;
            sX[rr][col].im -= u.re * x.im + u.im * x.re;
        }
        __syncthreads();
    }
    cx<T>* Ui = inv + (size_t)(2 * ch + 1) * 4096;
    for (int e = tid; e < 64 * 64; e += 256)
        Ui[e] = sX[e % 64][e / 64];
}

// B (w x ncols, leading dimension ldb) <- Tm B with a dense 64 x 64 Tm (column-major), in place;
// one workgroup per 64 columns, every thread computes 16 rows of one column.
template <typename T>
__global__ __launch_bounds__(256) void trmm_block(int w, const cx<T>* __restrict__ Tm, cx<T>* __restrict__ B, int64_t ldb,
                                                  int64_t ncols)
{
    __shared__ cx<T> sT[64][65];
    __shared__ cx<T> sX[64][65];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int nc = (int)((ncols - c0) < 64 ? (ncols - c0) : 64);
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e % 64, c = e / 64;
        sT[r][c] = (r < w && c < w) ? Tm[r + 64 * c] : cx<T>{0, 0};
        sX[r][c] = (r < w && c < nc) ? B[r + (c0 + c) * ldb] : cx<T>{0, 0};
    }
    __syncthreads();
    const int col = tid & 63, rg = tid >> 6;
    cx<T> acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        acc[i] = cx<T>{0, 0};
    for (int j = 0; j < w; j++) {
        const cx<T> x = sX[j][col];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const cx<T> t = sT[rg + 4 * i][j];
            acc[i].re += t.re * x.re - t.im * x.im;
            acc[i].im += t.re * x.im + t.im * x.re;
        }
    }
    if (col < nc)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int r = rg + 4 * i;
            if (r < w)
                B[r + (c0 + col) * ldb] = acc[i];
        }
}

__global__ void zero_u32(unsigned* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

__global__ void copy_u32(const unsigned* __restrict__ src, unsigned* __restrict__ dst, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

// Clears / copies device memory with kernels rather than memset / memcpy nodes, so that the
// sequences that use them can be captured into HIP graphs.  Sizes are multiples of 4 bytes.
int launch_zero_bytes(void* p, size_t bytes, hipStream_t st)
{
    const int64_t n = (int64_t)(bytes / 4);
    if (n > 0)
        launch_kernel(&zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (unsigned*)p, n);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_copy_bytes(const void* src, void* dst, size_t bytes, hipStream_t st)
{
    const int64_t n = (int64_t)(bytes / 4);
    if (n > 0)
        launch_kernel(&copy_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const unsigned*)src,
                           (unsigned*)dst, n);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// ----------------------------------------------------------------------------------------------
// Elementwise helpers for mixed precision and refinement.
// ----------------------------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void convert_mat(const cx<TI>* __restrict__ X, int64_t ldx, cx<TO>* __restrict__ Y,
                                                   int64_t ldy, int64_t rows)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r < rows) {
        const cx<TI> v = X[r + c * ldx];
        Y[r + c * ldy] = cx<TO>{(TO)v.re, (TO)v.im};
    }
}

// X64 += D32
__global__ __launch_bounds__(256) void add_correction(c128* __restrict__ X, int64_t ldx, const c64* __restrict__ D,
                                                      int64_t ldd, int64_t rows)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (r < rows) {
        const c64 d = D[r + c * ldd];
        X[r + c * ldx].re += (double)d.re;
        X[r + c * ldx].im += (double)d.im;
    }
}

// out[c] = sum_r |X[r,c]|^2, one workgroup per column.
// R += R2 (rows x cols, leading dimension ld).
__global__ __launch_bounds__(256) void add_into(c128* __restrict__ R, const c128* __restrict__ R2, int64_t ld,
                                                int64_t rows)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (r < rows) {
        const c128 v = R2[r + c * ld];
        R[r + c * ld].re += v.re;
        R[r + c * ld].im += v.im;
    }
}

__global__ __launch_bounds__(256) void col_norm2(const c128* __restrict__ X, int64_t ldx, int64_t rows,
                                                 double* __restrict__ out)
{
    const int64_t c = blockIdx.x;
    double s = 0;
    for (int64_t r = threadIdx.x; r < rows; r += 256) {
        const c128 v = X[r + c * ldx];
        s += v.re * v.re + v.im * v.im;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        out[c] = red[0] + red[1] + red[2] + red[3];
}

// ----------------------------------------------------------------------------------------------
// Host drivers
// ----------------------------------------------------------------------------------------------
// Complex products with three real MFMAs per complex multiply-add instead of four.  Normwise as
// stable as the four-product form, with a larger constant in the imaginary part.  Used unless
// FRICO_GEMM_3M=0 or the caller asks for the exact form: always for complex64 (the LU of the
// refinement form; at N = 12160 the refinement still ends after two sweeps and the update runs
// 1.13x faster), and for complex128 from K = 128 (two waves per SIMD with 196 VGPRs: the K = 128
// update at 64.4 TFLOP/s against 60.9; at K = 64 the four-product tile is faster, 49.7 against
// 45.2).  The FP64 residuals stay exact: with three products (65.5 TFLOP/s against 54.9 at
// K = 12160) a refinement to 1e-13 at n = 700 stalled at 1.2e-13.
static bool gemm_three_products()
{
    static const bool on = [] {
        const char* e = std::getenv("FRICO_GEMM_3M");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

// C[:, q] -= A B[:, q] for q < ncol <= GEMV_MAXN (A: M x K, column-major), the products of solves
// and residuals with one to four right-hand sides.  The tile product would run a 64 x 64 tile per 64
// rows with all but one to four of its columns empty, a K step at a time (~380 us for a 3500 x
// 3500 residual); here the product is a read of A at HBM speed.  Workgroup: 64 rows (one a lane),
// its sixteen waves a contiguous sixteenth of K each, sixteen columns of A in flight a lane; the
// waves' partial sums are added through LDS in a fixed order, so the result does not depend on
// scheduling.  The four real products of each complex one are kept (no three-product form).
constexpr int GEMV_MAXN = 4;
constexpr int GEMV_WAVES = 16;

template <typename TA, typename T>
__global__ __launch_bounds__(64 * GEMV_WAVES) void gemv_sub(int64_t M, int64_t K, int ncol, const cx<TA>* __restrict__ A,
                                                           int64_t lda, const cx<T>* __restrict__ B, int64_t ldb,
                                                           cx<T>* __restrict__ C, int64_t ldc, Batch bt)
{
    // batched: A += z s0, B += z s1, C += z s2
    A += blockIdx.z * bt.s0;
    B += blockIdx.z * bt.s1;
    C += blockIdx.z * bt.s2;
    __shared__ cx<T> sp[GEMV_WAVES][64][GEMV_MAXN];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t row = (int64_t)blockIdx.x * 64 + lane;
    const int64_t arow = row < M ? row : M - 1;
    const int64_t kq = (K + GEMV_WAVES - 1) / GEMV_WAVES, k0 = wv * kq;
    const int64_t k1 = (k0 + kq) < K ? (k0 + kq) : K;
    cx<T> acc[GEMV_MAXN];
#pragma unroll
    for (int q = 0; q < GEMV_MAXN; q++)
        acc[q] = cx<T>{0, 0};
    constexpr int U = 16;
    for (int64_t k = k0; k < k1; k += U) {
        cx<TA> a[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t kk = (k + u) < k1 ? (k + u) : k1 - 1;
            a[u] = A[arow + kk * lda];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (k + u >= k1)
                break;
#pragma unroll
            for (int q = 0; q < GEMV_MAXN; q++)
                if (q < ncol) {
                    const cx<T> b = B[(k + u) + q * ldb];
                    const T ar = (T)a[u].re, ai = (T)a[u].im;
                    acc[q].re += ar * b.re - ai * b.im;
                    acc[q].im += ar * b.im + ai * b.re;
                }
        }
    }
#pragma unroll
    for (int q = 0; q < GEMV_MAXN; q++)
        sp[wv][lane][q] = acc[q];
    __syncthreads();
    if (wv == 0 && row < M) {
        for (int q = 0; q < ncol; q++) {
            cx<T> t = sp[0][lane][q];
#pragma unroll
            for (int w = 1; w < GEMV_WAVES; w++) {
                t.re += sp[w][lane][q].re;
                t.im += sp[w][lane][q].im;
            }
            cx<T> c = C[row + q * ldc];
            c.re -= t.re;
            c.im -= t.im;
            C[row + q * ldc] = c;
        }
    }
}

template <typename TA, typename T>
static int launch_gemv_sub(int64_t M, int64_t N, int64_t K, const cx<TA>* A, int64_t lda, const cx<T>* B, int64_t ldb,
                           cx<T>* C, int64_t ldc, hipStream_t st, const Batch& bt = Batch{}, int nz = 1)
{
    const int64_t g = (M + 63) / 64;
    if (g > INT32_MAX || nz < 1 || nz > 65535)
        return set_error("gemv: %lld row blocks, %d systems", (long long)g, nz), -1;
    launch_kernel(&gemv_sub<TA, T>, dim3((unsigned)g, 1, (unsigned)nz), dim3(64 * GEMV_WAVES), 0, st, M, K, (int)N, A, lda,
                  B, ldb, C, ldc, bt);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int gemv_sub_mixed_batched(int64_t M, int64_t N, int64_t K, const c64* A, int64_t lda, int64_t sa, const c128* B,
                           int64_t ldb, int64_t sb, c128* C, int64_t ldc, int64_t sc, int nz, hipStream_t st)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    if (N > GEMV_MAXN)
        return set_error("gemv: %lld columns", (long long)N), -1;
    Batch bt;
    bt.s0 = sa;
    bt.s1 = sb;
    bt.s2 = sc;
    return launch_gemv_sub<float, double>(M, N, K, A, lda, B, ldb, C, ldc, st, bt, nz);
}

template <typename T>
int gemm_sub(int64_t M, int64_t N, int64_t K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb, cx<T>* C,
             int64_t ldc, hipStream_t st, KernelTimer* kt, bool exact)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    using TL = typename TileBig<T>::type;
    const int64_t gx = (M + TL::BM - 1) / TL::BM, gy = (N + TL::BN - 1) / TL::BN;
    if (gy > 65535)
        return set_error("gemm: too many column tiles"), -1;
    if (kt)
        kt->begin(st);
    if (N <= GEMV_MAXN) {
        const int rc = launch_gemv_sub<T, T>(M, N, K, A, lda, B, ldb, C, ldc, st);
        if (kt)
            kt->end(st, 8.0 * (double)M * (double)N * (double)K,
                    (double)sizeof(cx<T>) * (2.0 * (double)M * (double)N + ((double)M + (double)N) * (double)K));
        return rc;
    }
    const dim3 grid((unsigned)gx, (unsigned)gy);
    if (!exact && gemm_three_products() && (sizeof(T) == 4 || K >= 128))
        launch_kernel(&cgemm_sub<T, T, false, true>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc);
    else
        launch_kernel(&cgemm_sub<T, T>, grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc);
    if (kt)
        kt->end(st, 8.0 * (double)M * (double)N * (double)K,
                (double)sizeof(cx<T>) * (2.0 * (double)M * (double)N + ((double)M + (double)N) * (double)K));
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// C -= A B on the tiles on and below C's diagonal (C square-tiled from a diagonal entry; N <= M).
template <typename T>
static int gemm_sub_lower(int64_t M, int64_t N, int64_t K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb,
                          cx<T>* C, int64_t ldc, hipStream_t st, KernelTimer* kt)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    if (N > M)
        return set_error("lower product: more columns than rows"), -1;
    using TL = typename TileBig<T>::type;
    const int64_t gx = (M + TL::BM - 1) / TL::BM, gy = (N + TL::BN - 1) / TL::BN;
    const int64_t tiles = gy * gx - gy * (gy - 1) / 2;
    if (kt)
        kt->begin(st);
    if (gemm_three_products() && (sizeof(T) == 4 || K >= 128))
        launch_kernel(&cgemm_sub_lower<T, true>, dim3((unsigned)tiles), dim3(256), 0, st, M, N, K, gx, gy, A, lda, B, ldb,
                      C, ldc);
    else
        launch_kernel(&cgemm_sub_lower<T>, dim3((unsigned)tiles), dim3(256), 0, st, M, N, K, gx, gy, A, lda, B, ldb, C,
                      ldc);
    // flops of the tiles computed (the diagonal tiles in full)
    const double cols = (double)N, full = (double)M * cols - cols * (cols - TL::BN) / 2.0;
    if (kt)
        kt->end(st, 8.0 * full * (double)K);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// C_s -= A_s B_s, s = 0, 1 (M x K_s times K_s x N), in one launch.
static int gemm_sub_pair(int64_t M, int64_t N, const GemmPair<double>& p, hipStream_t st, bool exact)
{
    if (M <= 0 || N <= 0 || p.K[0] <= 0 || p.K[1] <= 0)
        return set_error("gemm pair: empty product"), -1;
    using TL = typename TileBig<double>::type;
    const int64_t gx = (M + TL::BM - 1) / TL::BM, gy = (N + TL::BN - 1) / TL::BN;
    if (gy > 65535)
        return set_error("gemm: too many column tiles"), -1;
    const int64_t nwg = 8 * ((gx * gy + 7) / 8);
    if (nwg > INT32_MAX)
        return set_error("gemm: too many tiles"), -1;
    const dim3 grid((unsigned)nwg, 1, 2);
    if (!exact && gemm_three_products() && p.K[0] >= 128 && p.K[1] >= 128)
        launch_kernel(&cgemm_sub2<double, true>, grid, dim3(256), 0, st, M, N, gx, gy, p);
    else
        launch_kernel(&cgemm_sub2<double, false>, grid, dim3(256), 0, st, M, N, gx, gy, p);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int gemm_sub_mixed(int64_t M, int64_t N, int64_t K, const c64* A, int64_t lda, const c128* B, int64_t ldb, c128* C,
                   int64_t ldc, hipStream_t st, KernelTimer* kt)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    using TL = typename TileBig<double>::type;
    const int64_t gx = (M + TL::BM - 1) / TL::BM, gy = (N + TL::BN - 1) / TL::BN;
    if (gy > 65535)
        return set_error("gemm: too many column tiles"), -1;
    if (kt)
        kt->begin(st);
    if (N <= GEMV_MAXN) {
        if (int rc = launch_gemv_sub<float, double>(M, N, K, A, lda, B, ldb, C, ldc, st))
            return rc;
    } else {
        launch_kernel(&cgemm_sub<double, float>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, M, N, K, A, lda, B,
                      ldb, C, ldc);
    }
    if (kt)
        kt->end(st, 8.0 * (double)M * (double)N * (double)K);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// Z = (Z + Z^T) / 2 in place (src/emw_solver.h:118-120): one thread per pair j > i.
__global__ __launch_bounds__(256) void symmetrize_k(c128* __restrict__ Z, int64_t n, int64_t ldz)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= n || j <= i)
        return;
    const c128 a = Z[j + i * ldz], b = Z[i + j * ldz];
    const c128 s{(a.re + b.re) / 2.0, (a.im + b.im) / 2.0};
    Z[j + i * ldz] = s;
    Z[i + j * ldz] = s;
}

int launch_symmetrize(c128* Z, int64_t n, int64_t ldz, hipStream_t st)
{
    if (n > 65535)
        return set_error("symmetrize: n too large"), -1;
    launch_kernel(&symmetrize_k, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, st, Z, n, ldz);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// Rows per panel workgroup: one row per thread (at most 256), slab of R rows x w complex plus two
// rows in LDS.  With shared_cu the slab stays within ~72 KB, which leaves a CU room for the
// trailing-update workgroups it already holds, so that the panel runs beside the look-ahead update
// instead of waiting for CUs to drain; otherwise it may use ~144 KB.
template <typename T> int panel_rows(int w, bool shared_cu)
{
    const size_t budget = (size_t)(shared_cu ? 72 : 144) * 1024;
    int R = (int)(budget / ((size_t)w * sizeof(cx<T>))) - 3; // slab padding, pivot and previous rows
    R = R >= 256 ? 256 : (R >= 128 ? 128 : (R >= 64 ? 64 : 32));
    return R;
}

// Update of columns [c0, c0 + nc) by step k (panel of width w): row interchanges, U12 = L11^{-1} A12
// and A22 -= L21 U12.
// perm2: a second interchange list applied after perm (the two halves of a composite panel).
// U12 = L11^{-1} A12 on columns [c0, c0 + nc) of the block row [k, k + w) (L11 unit lower, in
// 64-row chunks: substitution, then the chunk's product with the rows below it in the block).
template <typename T>
static int row_block_solve(cx<T>* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, hipStream_t st)
{
    for (int s = 0; s < w; s += 64) {
        const int ws = (w - s) < 64 ? (w - s) : 64;
        cx<T>* D = A + (k + s) + (k + s) * lda;
        cx<T>* Bb = A + (k + s) + c0 * lda;
        launch_trsm<T, false>(ws, D, lda, Bb, lda, Bb, lda, nc, st);
        if (s + ws < w) {
            int rc = gemm_sub<T>(w - s - ws, nc, ws, A + (k + s + ws) + (k + s) * lda, lda, Bb, lda,
                                 A + (k + s + ws) + c0 * lda, lda, st, nullptr);
            if (rc)
                return rc;
        }
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
static int step_update(cx<T>* A, int64_t n, int64_t lda, const int* perm, const int* perm2, int64_t k, int w,
                       int64_t c0, int64_t nc, hipStream_t st, KernelTimer* kt)
{
    if (nc <= 0)
        return 0;
    launch_laswp<T>(A, lda, c0, nc, perm, st, perm2);
    if (int rc = row_block_solve<T>(A, lda, k, w, c0, nc, st))
        return rc;
    return gemm_sub<T>(n - k - w, nc, w, A + (k + w) + k * lda, lda, A + k + c0 * lda, lda, A + (k + w) + c0 * lda, lda,
                       st, kt);
}

template <typename T>
static int solve_tables(LuWork& wk, const cx<T>* A, int64_t n, int64_t lda, int nb, hipStream_t st);

// Right-looking blocked LU.  With a side stream the update of step p is split: the columns of panel
// p+1 are updated on the main stream, which then factors panel p+1 while the side stream updates
// the remaining columns (look-ahead of depth one).  Dependencies are carried by events:
//   main: panel p -> [P_p] -> (wait W_{p-1}) -> narrow update p -> panel p+1
//   side: (wait P_p) -> wide update p -> [W_p]
template <typename T>
int lu_factor(LuWork& wk, cx<T>* A, int64_t n, int64_t lda, int* ipiv, int nb, hipStream_t st, KernelTimer* gemm_timer,
              int64_t naug)
{
    if (nb > LU_NB_MAX || nb < 1)
        return set_error("lu: block size out of range"), -1;
    const int64_t npan = (n + nb - 1) / nb;
    const bool lookahead = wk.side != nullptr && wk.events != nullptr && wk.nevents >= 2 * npan + 1;
    // Panels wider than 64 columns are factored as 64-column sub-panels (see panel() below): the
    // panel kernel's per-column cost does not depend on its width, and the trailing products then
    // run with K = nb.
    const int pw = nb < 64 ? nb : 64;
    int R = panel_rows<T>(pw, lookahead);
    if ((n + R - 1) / R > wk.max_groups)
        R = panel_rows<T>(pw, false);
    const size_t sh_max = ((size_t)pw * (R + 1) + 2 * pw) * sizeof(cx<T>);
    if (sh_max > 64 * 1024)
        FRICO_HIP_CHECK(allow_lds(reinterpret_cast<const void*>(&panel_factor<T>), sh_max));
    const int gmax = wk.max_groups;
    if (!wk.ll || wk.nll < ll_words(gmax))
        return set_error("lu: exchange workspace too small"), -1;
    const bool composite = nb > pw && n > pw;
    if (!wk.perm || wk.nperm < npan * (composite ? 3 : 1))
        return set_error("lu: interchange-list workspace too small"), -1;
    // lists of the two halves of composite panel p: sub_perm + (2 p, 2 p + 1) * PERM_STRIDE
    int* sub_perm = wk.perm + npan * PERM_STRIDE;
    // interchange lists the trailing updates of panel p apply
    auto lists = [&](int64_t p, const int*& l1, const int*& l2) {
        const int64_t k = p * nb;
        if (composite && n - k > pw) {
            l1 = sub_perm + 2 * p * PERM_STRIDE;
            l2 = l1 + PERM_STRIDE;
        } else {
            l1 = wk.perm + p * PERM_STRIDE;
            l2 = nullptr;
        }
    };
    const int64_t nz = 2 * (int64_t)ll_words(gmax);
    launch_kernel(&zero_u32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, (unsigned*)wk.ll, nz);
    launch_kernel(&zero_u32, dim3(1), dim3(256), 0, st, (unsigned*)wk.flags, (int64_t)2);
    FRICO_HIP_CHECK(hipGetLastError());
    PanelSync ps{};
    ps.words = wk.ll;
    ps.gmax = gmax;
    ps.abort_flag = wk.flags;
    ps.info = wk.flags + 1;
    // One panel kernel: columns [k, k + w) (w <= 64), rows [k, n); its interchange list to plist.
    auto panel64 = [&](int64_t k, int w, int* plist) -> int {
        const int64_t m = n - k;
        const int G = (int)((m + R - 1) / R);
        if (G > gmax)
            return set_error("lu: panel needs %d workgroups, only %d can be co-resident", G, gmax), FRICO_ERR_LIMIT;
        const size_t sh = ((size_t)w * (R + 1) + 2 * w) * sizeof(cx<T>);
        launch_kernel(&panel_factor<T>, dim3(G), dim3(256), sh, st, A, lda, k, m, w, R, G, ipiv, ps, plist);
        FRICO_HIP_CHECK(hipGetLastError());
        return 0;
    };
    // Panel p.  A composite panel (w > 64) is factored as [P1 | P2]: P1 by the panel kernel; P2
    // given P1's interchanges, U12 = L11^{-1} A12 and A22 -= L21 U12 (the step update restricted to
    // P2's columns); P2 by the panel kernel; then P2's interchanges applied to P1's columns, so that
    // the w columns of L share one row order and the panel behaves as one of width w.  Its
    // two lists are applied by the trailing updates in one pass, and the panel's own list (for the
    // solves) is built from the w pivots after the factorisation.  Partial pivoting picks the same
    // rows as it would over the whole w columns: column by column, each among the fully updated rows.
    auto panel = [&](int64_t p) -> int {
        const int64_t k = p * nb;
        const int w = (int)((n - k) < nb ? (n - k) : nb);
        int* plist = wk.perm + p * PERM_STRIDE;
        if (w <= pw)
            return panel64(k, w, plist);
        const int w1 = pw, w2 = w - pw;
        int* l1 = sub_perm + 2 * p * PERM_STRIDE;
        int rc;
        if ((rc = panel64(k, w1, l1)))
            return rc;
        cx<T>* A12 = A + k + (k + w1) * lda;
        launch_laswp<T>(A, lda, k + w1, w2, l1, st);
        launch_trsm<T, false>(w1, A + k + k * lda, lda, A12, lda, A12, lda, w2, st);
        if ((rc = gemm_sub<T>(n - k - w1, w2, w1, A + (k + w1) + k * lda, lda, A12, lda, A + (k + w1) + (k + w1) * lda,
                              lda, st, nullptr)))
            return rc;
        if ((rc = panel64(k + w1, w2, l1 + PERM_STRIDE)))
            return rc;
        launch_laswp<T>(A, lda, k, w1, l1 + PERM_STRIDE, st);
        FRICO_HIP_CHECK(hipGetLastError());
        return 0;
    };
    int rc;
    if (!lookahead) {
        for (int64_t p = 0; p < npan; p++) {
            const int64_t k = p * nb;
            const int w = (int)((n - k) < nb ? (n - k) : nb);
            const int *l1, *l2;
            lists(p, l1, l2);
            if ((rc = panel(p)) || (rc = step_update<T>(A, n, lda, l1, l2, k, w, k + w, n + naug - k - w, st, gemm_timer)))
                return rc;
        }
    } else {
        hipEvent_t* evP = wk.events;
        hipEvent_t* evW = wk.events + npan;
        hipEvent_t ev0 = wk.events[2 * npan];
        FRICO_HIP_CHECK(hipEventRecord(ev0, st));
        FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, ev0, 0));
        if ((rc = panel(0)))
            return rc;
        bool prev_wide = false;
        for (int64_t p = 0; p < npan; p++) {
            const int64_t k = p * nb;
            const int w = (int)((n - k) < nb ? (n - k) : nb);
            const int64_t nright = n - k - w;
            const int *l1, *l2;
            lists(p, l1, l2);
            if (nright <= 0) {
                // the last panel's update of the appended columns
                if (naug > 0) {
                    if (prev_wide)
                        FRICO_HIP_CHECK(hipStreamWaitEvent(st, evW[p - 1], 0));
                    if ((rc = step_update<T>(A, n, lda, l1, l2, k, w, n, naug, st, gemm_timer)))
                        return rc;
                }
                break;
            }
            const int64_t nnext = nright < nb ? nright : nb;
            const int64_t nrest = nright + naug - nnext; // the appended columns go with the wide part
            FRICO_HIP_CHECK(hipEventRecord(evP[p], st));
            if (nrest > 0) {
                FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, evP[p], 0));
                if ((rc = step_update<T>(A, n, lda, l1, l2, k, w, k + w + nnext, nrest, wk.side, gemm_timer)))
                    return rc;
                FRICO_HIP_CHECK(hipEventRecord(evW[p], wk.side));
            }
            if (prev_wide)
                FRICO_HIP_CHECK(hipStreamWaitEvent(st, evW[p - 1], 0));
            if ((rc = step_update<T>(A, n, lda, l1, l2, k, w, k + w, nnext, st, gemm_timer)) || (rc = panel(p + 1)))
                return rc;
            prev_wide = nrest > 0;
        }
        // the main stream continues only after all side work
        FRICO_HIP_CHECK(hipEventRecord(ev0, wk.side));
        FRICO_HIP_CHECK(hipStreamWaitEvent(st, ev0, 0));
    }
    if (composite) {
        // the composite panels' own lists, for the solves (a panel kernel wrote the others)
        launch_kernel(&build_perm_lists, dim3((unsigned)npan), dim3(64), 0, st, nb, n, (const int*)ipiv, wk.perm);
        FRICO_HIP_CHECK(hipGetLastError());
    }
    return solve_tables<T>(wk, A, n, lda, nb, st);
}

// The solves' tables of the factors: the diagonal-chunk inverses and the W products (LuWork).
template <typename T>
static int solve_tables(LuWork& wk, const cx<T>* A, int64_t n, int64_t lda, int nb, hipStream_t st)
{
    if (wk.dinv) {
        const int64_t nch = (n + 63) / 64;
        launch_kernel(&diag_inverse<T>, dim3((unsigned)nch), dim3(256), 0, st, A, n, lda, (cx<T>*)wk.dinv);
        FRICO_HIP_CHECK(hipGetLastError());
        if (wk.wv && nb % 64 == 0 && nch > 1) {
            launch_kernel(&make_w<T>, dim3((unsigned)nch, (unsigned)nch), dim3(256), 0, st, n, A, lda,
                               (const cx<T>*)wk.dinv, (cx<T>*)wk.wv, n);
            FRICO_HIP_CHECK(hipGetLastError());
        }
    }
    return 0;
}

// ----------------------------------------------------------------------------------------------
// Symmetric factorisation Z = L D L^T of a complex symmetric (not Hermitian) Z, without pivoting:
// the factorisation behind solve_symmetric (zsytrf, src/lapack_wrappers.cpp:135-177) for the
// refinement form.  Only the lower triangle of Z is read, and only the tiles holding the lower
// triangle of each trailing matrix are updated: half the trailing products of the LU.  The factors
// are stored as the LU of Z with no interchanges would be, L strictly below the diagonal and
// U = D L^T on and above it, so the LU solves run on them unchanged with empty interchange lists.
// Block column [k, k + w):
//   lu_diag<T, true> L11 and U11 = D1 L11^T of the w x w diagonal block from its lower triangle
//                    (the blocked diagonal-block LU below, one workgroup)
//   sym_transpose    A12 = A21^T: the block row, taken from the block column
//   row_block_solve  U12 = L11^{-1} A12 (= D1 L21^T)
//   sym_transpose    L21 = U12^T D1^{-1}
//   cgemm_sub_lower  A22 -= L21 U12 on the tiles on and below A22's diagonal
// Without interchanges the factors can grow: a zero pivot is reported like zgetrf's info, and the
// accuracy of the answer rests on the refinement against the kept Z (the caller returns to the
// pivoted LU when the refinement stalls).
// ----------------------------------------------------------------------------------------------
template <typename T, bool SYM>
__global__ void lu_diag(cx<T>* __restrict__ A, int64_t lda, int64_t k, int w, int* info, Batch bt);
// block-row columns per transpose workgroup
template <typename T> constexpr int kSymCols = sizeof(T) == 4 ? 32 : 16;

// Block row <-> block column of step k (w columns, m = n - k - w rows below), kSymCols values of c a
// workgroup, every load of a thread in flight before its LDS stores:
//   TO_LOWER false: A(k + i, k + w + c) = A(k + w + c, k + i)                        (A12 = A21^T)
//   TO_LOWER true:  A(k + w + c, k + i) = A(k + i, k + w + c) / A(k + i, k + i)     (L21 = U12^T D^{-1})
template <typename T, bool TO_LOWER>
__global__ __launch_bounds__(256) void sym_transpose(cx<T>* __restrict__ A, int64_t lda, int64_t k, int w, int64_t m)
{
    constexpr int TC = kSymCols<T>;
    constexpr int NE = LU_NB_MAX * TC / 256;
    __shared__ cx<T> S[LU_NB_MAX][TC + 1];
    __shared__ cx<T> sinv[LU_NB_MAX];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * TC;
    const int nc = (int)((m - c0) < TC ? (m - c0) : TC);
    cx<T>* A21 = A + (k + w + c0) + k * lda; // entry (c, i) at A21[c + i lda]
    cx<T>* A12 = A + k + (k + w + c0) * lda; // entry (i, c) at A12[i + c lda]
    cx<T> v[NE];
    if (TO_LOWER) {
        // block row, down each column (i fastest)
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, i = e % LU_NB_MAX, c = e / LU_NB_MAX;
            const bool ok = i < w && c < nc;
            v[s] = A12[ok ? i + (int64_t)c * lda : 0];
        }
        if (tid < w) {
            const cx<T> d = A[(k + tid) * (lda + 1)];
            sinv[tid] = (d.re == (T)0 && d.im == (T)0) ? cx<T>{0, 0} : cdiv(cx<T>{1, 0}, d);
        }
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, i = e % LU_NB_MAX, c = e / LU_NB_MAX;
            S[i][c] = v[s];
        }
        __syncthreads();
        // block column, down each of its columns (c fastest)
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, c = e % TC, i = e / TC;
            if (i < w && c < nc)
                A21[c + (int64_t)i * lda] = S[i][c] * sinv[i];
        }
    } else {
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, c = e % TC, i = e / TC;
            const bool ok = i < w && c < nc;
            v[s] = A21[ok ? c + (int64_t)i * lda : 0];
        }
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, c = e % TC, i = e / TC;
            S[i][c] = v[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, i = e % LU_NB_MAX, c = e / LU_NB_MAX;
            if (i < w && c < nc)
                A12[i + (int64_t)c * lda] = S[i][c];
        }
    }
}

__global__ void iota_pivots(int* ipiv, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        ipiv[i] = (int)(i + 1);
}

// Blocked L D L^T (see above).  With a side stream, as lu_factor: the update of step p is split
// into the next block column (main stream, which then factors it) and the rest (side stream);
// the appended right-hand-side columns are carried on the side stream.
template <typename T>
int ldlt_factor(LuWork& wk, cx<T>* A, int64_t n, int64_t lda, int* ipiv, int nb, hipStream_t st, KernelTimer* gemm_timer,
                int64_t naug)
{
    if (nb > LU_NB_MAX || nb < 1)
        return set_error("ldlt: block size out of range"), -1;
    const int64_t npan = (n + nb - 1) / nb;
    if (!wk.perm || wk.nperm < npan)
        return set_error("ldlt: interchange-list workspace too small"), -1;
    const bool lookahead = wk.side != nullptr && wk.events != nullptr && wk.nevents >= 2 * npan + 1;
    // no interchanges: empty lists for the solves, the identity in ipiv
    const int64_t nz = npan * PERM_STRIDE;
    launch_kernel(&zero_u32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, (unsigned*)wk.perm, nz);
    launch_kernel(&zero_u32, dim3(1), dim3(256), 0, st, (unsigned*)wk.flags, (int64_t)2);
    launch_kernel(&iota_pivots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ipiv, n);
    FRICO_HIP_CHECK(hipGetLastError());
    // diagonal block, block row and block column of step k
    auto column = [&](int64_t k, int w) -> int {
        const int64_t m = n - k - w;
        launch_kernel(&lu_diag<T, true>, dim3(1), dim3(1024), 0, st, A, lda, k, w, wk.flags + 1, Batch{});
        if (m > 0) {
            const dim3 g((unsigned)((m + kSymCols<T> - 1) / kSymCols<T>));
            launch_kernel(&sym_transpose<T, false>, g, dim3(256), 0, st, A, lda, k, w, m);
            if (int rc = row_block_solve<T>(A, lda, k, w, k + w, m, st))
                return rc;
            launch_kernel(&sym_transpose<T, true>, g, dim3(256), 0, st, A, lda, k, w, m);
        }
        FRICO_HIP_CHECK(hipGetLastError());
        return 0;
    };
    // the appended columns: their block row solved with L11, then the rows below updated
    auto appended = [&](int64_t k, int w, hipStream_t s) -> int {
        if (naug <= 0)
            return 0;
        int rc = row_block_solve<T>(A, lda, k, w, n, naug, s);
        return rc ? rc
                  : gemm_sub<T>(n - k - w, naug, w, A + (k + w) + k * lda, lda, A + k + n * lda, lda,
                                A + (k + w) + n * lda, lda, s, nullptr);
    };
    // lower triangle of the trailing matrix in columns [c0, c0 + nc) (rows [c0, n))
    auto update = [&](int64_t k, int w, int64_t c0, int64_t nc, hipStream_t s) -> int {
        return gemm_sub_lower<T>(n - c0, nc, w, A + c0 + k * lda, lda, A + k + c0 * lda, lda, A + c0 + c0 * lda, lda, s,
                                 gemm_timer);
    };
    auto width = [&](int64_t k) { return (int)((n - k) < nb ? (n - k) : nb); };
    int rc;
    if (!lookahead) {
        for (int64_t p = 0; p < npan; p++) {
            const int64_t k = p * nb;
            const int w = width(k);
            if ((rc = column(k, w)) || (rc = appended(k, w, st)) || (rc = update(k, w, k + w, n - k - w, st)))
                return rc;
        }
    } else {
        hipEvent_t* evP = wk.events;
        hipEvent_t* evW = wk.events + npan;
        hipEvent_t ev0 = wk.events[2 * npan];
        FRICO_HIP_CHECK(hipEventRecord(ev0, st));
        FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, ev0, 0));
        if ((rc = column(0, width(0))))
            return rc;
        for (int64_t p = 0; p < npan; p++) {
            const int64_t k = p * nb;
            const int w = width(k);
            const int64_t nright = n - k - w;
            const int64_t nnext = nright < nb ? nright : nb;
            // The side stream's products start after the narrow update, so that the next diagonal
            // block (one 1024-thread workgroup) is dispatched before they fill the CUs.
            if (nright > 0) {
                // the next block column was last updated by the wide part of the previous step
                if (p > 0)
                    FRICO_HIP_CHECK(hipStreamWaitEvent(st, evW[p - 1], 0));
                if ((rc = update(k, w, k + w, nnext, st)))
                    return rc;
            }
            FRICO_HIP_CHECK(hipEventRecord(evP[p], st));
            FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, evP[p], 0));
            if ((rc = appended(k, w, wk.side)) || (rc = update(k, w, k + w + nnext, nright - nnext, wk.side)))
                return rc;
            FRICO_HIP_CHECK(hipEventRecord(evW[p], wk.side));
            if (nright <= 0)
                break;
            if ((rc = column(k + w, width(k + w))))
                return rc;
        }
        FRICO_HIP_CHECK(hipEventRecord(ev0, wk.side));
        FRICO_HIP_CHECK(hipStreamWaitEvent(st, ev0, 0));
    }
    return solve_tables<T>(wk, A, n, lda, nb, st);
}

// ----------------------------------------------------------------------------------------------
// LU without pivoting, for the refinement form: the refinement against the kept FP64 Z corrects
// what the complex64 factors lose, and the caller returns to the pivoted LU when a pivot is zero
// or the refinement stalls.  Without interchanges a block column needs no grid-wide exchange per
// column: one workgroup factors the diagonal block and the rows below are triangular solves.
// Block column [k, k + w):
//   lu_diag           L11 U11 of the w x w diagonal block (one workgroup, 16-column sub-panels)
//   col_transpose     T = A21^T into a workspace
//   trsm (U11^T)      T = U11^{-T} T, i.e. L21 = A21 U11^{-1}
//   col_transpose     A21 = T^T
// and the trailing updates are the LU's (U12 = L11^{-1} A12, A22 -= L21 U12) with no interchanges.
// ----------------------------------------------------------------------------------------------

// Blocked right-looking LU of the w x w (w <= 128) diagonal block at (k, k), in 16-column
// sub-panels, one workgroup.  Each thread keeps a 4 x 4 tile of the block in registers for the
// trailing updates (wavefronts of 16 rows by 64 columns).  Sub-panel j, with D the dj x dj block
// at s0 = 16 j and e = s0 + dj:
//   * the panel column (rows >= s0) and panel row (columns >= e) are staged in LDS;
//   * row wavefronts hold D's rows in lanes 0-15 and up to 48 rows below it in lanes 16-63, one
//     row a lane, and run D's column steps with the pivot row read by readlane (no barrier):
//     lanes 0-15 end with D's L\U, the others with their rows of L21 = A21 U_D^-1;
//   * column wavefronts do the same on the transposed strip (D's columns in lanes 0-15, up to 48
//     columns of A12 in lanes 16-63) and end with U12 = L_D^-1 A12;
//   * both write their results to A and to LDS, and every thread takes the rank-16 product
//     L21 U12 off its tile of the trailing block.
// Two barriers a sub-panel; the staging buffers alternate.  A zero pivot leaves its column of L
// zero and sets *info (1-based column) if unset.  SYM: the block is symmetric and only its lower
// triangle is read (entry (r, c), r < c, is taken from (c, r)); the factors are then L and U = D L^T.
constexpr int LUD_SUB = 16;   // sub-panel width
constexpr int LUD_CHAIN = 48; // rows (columns) past D per chain wavefront

__device__ inline float lane_value(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ inline double lane_value(double v, int l)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <typename T> __device__ inline cx<T> lane_value(cx<T> v, int l) { return {lane_value(v.re, l), lane_value(v.im, l)}; }

// 1 / d by Smith's scaling (no overflow for any finite nonzero d), with hardware reciprocals for
// complex64: the pivot inverse is on the dependent chain of every column step.
__device__ inline cx<float> pivot_inverse(cx<float> d)
{
    const bool big_re = fabsf(d.re) >= fabsf(d.im);
    const float p = big_re ? d.re : d.im, q = big_re ? d.im : d.re;
    const float r = q * __builtin_amdgcn_rcpf(p);
    const float rd = __builtin_amdgcn_rcpf(p + q * r);
    return big_re ? cx<float>{rd, -r * rd} : cx<float>{r * rd, -rd};
}
__device__ inline cx<double> pivot_inverse(cx<double> d) { return cdiv(cx<double>{1, 0}, d); }

template <typename T, bool SYM = false>
__global__ __launch_bounds__(1024) void lu_diag(cx<T>* __restrict__ A, int64_t lda, int64_t k, int w, int* info, Batch bt)
{
    // batched: A += z s0, info += z s1
    A += blockIdx.z * bt.s0;
    info += blockIdx.z * bt.s1;
    constexpr int SB = LUD_SUB, CR = LUD_CHAIN;
    __shared__ __align__(16) cx<T> Pc[2][SB][LU_NB_MAX];      // panel column: Pc[b][t][r - s0], r >= s0
    __shared__ __align__(16) cx<T> Pr[2][SB][LU_NB_MAX - SB]; // panel row: Pr[b][i][c - e], c >= e
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    // thread (bi, bj) holds rows 4 bi .. 4 bi + 3 of columns 4 bj .. 4 bj + 3; sixteen consecutive
    // lanes take consecutive row blocks, so that the block's loads run down its columns
    const int bi = 16 * (wv & 1) + (lane & 15), bj = 4 * (wv >> 1) + (lane >> 4);
    const int r0 = 4 * bi, j0 = 4 * bj;
    cx<T> a[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
            const int gr = r0 + i, gc = j0 + jj;
            const bool own = gr < w && gc < w, up = SYM && gc > gr;
            a[i][jj] = A[(own ? k + (up ? gc : gr) : k) + (own ? k + (up ? gr : gc) : k) * lda];
        }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
            if (!(r0 + i < w && j0 + jj < w))
                a[i][jj] = cx<T>{0, 0};
    // the tiles of sub-panel j's panel column and panel row into buffer b
    auto stage = [&](int j, int b) {
        const int s0 = SB * j, e = s0 + (w - s0 < SB ? w - s0 : SB);
        if (j0 >= s0 && j0 < e && r0 >= s0 && r0 < w)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    Pc[b][j0 + jj - s0][r0 + i - s0] = a[i][jj];
        if (r0 >= s0 && r0 < e && j0 >= e && j0 < w)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    Pr[b][r0 + i - s0][j0 + jj - e] = a[i][jj];
    };
    stage(0, 0);
    __syncthreads();
    const int nsp = (w + SB - 1) / SB;
    for (int j = 0; j < nsp; j++) {
        const int b = j & 1, s0 = SB * j, dj = w - s0 < SB ? w - s0 : SB, e = s0 + dj, R = w - e;
        const int q = wv >> 1;
        const int nrw = R > 0 ? (R + CR - 1) / CR : 1, ncw = (R + CR - 1) / CR;
        if ((wv & 1) == 0 && q < nrw) {
            // row wavefront q: lane -> row s0 + idx
            const int idx = lane < SB ? lane : dj + CR * q + (lane - SB);
            const bool valid = lane < SB ? lane < dj : idx < w - s0;
            const int ci = valid ? idx : 0;
            cx<T> x[SB];
#pragma unroll
            for (int t = 0; t < SB; t++)
                x[t] = (valid && t < dj) ? Pc[b][t][ci] : cx<T>{0, 0};
            // D's column steps; the pivot is wave-uniform (read by readlane), and so is the first
            // zero pivot's column zc, recorded after the steps.  A full sub-panel (dj == SB, all
            // but possibly the last) runs as one straight-line sequence.
            int zc = -1;
            auto steps = [&](auto full) {
#pragma unroll
                for (int c = 0; c < SB; c++) {
                    if (full || c < dj) {
                        const cx<T> d = lane_value(x[c], c);
                        const bool zero = d.re == (T)0 && d.im == (T)0;
                        const cx<T> inv = zero ? cx<T>{0, 0} : pivot_inverse(d);
                        zc = (zero && zc < 0) ? c : zc;
                        const bool below = lane > c;
                        const cx<T> l = below ? x[c] * inv : cx<T>{0, 0};
                        if (below)
                            x[c] = l;
#pragma unroll
                        for (int jj = c + 1; jj < SB; jj++) {
                            const cx<T> u = lane_value(x[jj], c);
                            x[jj].re -= l.re * u.re - l.im * u.im;
                            x[jj].im -= l.re * u.im + l.im * u.re;
                        }
                    }
                }
            };
            if (dj == SB)
                steps(std::true_type{});
            else
                steps(std::false_type{});
            if (zc >= 0 && q == 0 && lane == 0 && *info == 0)
                *info = (int)(k + s0 + zc + 1);
            if (valid && (q == 0 || lane >= SB)) {
                cx<T>* dst = A + (k + s0 + idx) + (k + s0) * lda;
#pragma unroll
                for (int t = 0; t < SB; t++)
                    if (t < dj)
                        dst[t * lda] = x[t];
                if (lane >= SB)
#pragma unroll
                    for (int t = 0; t < SB; t++)
                        Pc[b][t][idx] = x[t];
            }
        } else if ((wv & 1) == 1 && q < ncw) {
            // column wavefront q: lanes 0-15 -> D's column s0 + lane, the others -> column e + idx
            const int idx = lane < SB ? lane : CR * q + (lane - SB);
            const bool valid = lane < SB ? lane < dj : idx < R;
            const int pi = (lane >= SB && valid) ? idx : 0;
            cx<T> y[SB];
#pragma unroll
            for (int i = 0; i < SB; i++)
                y[i] = (valid && i < dj) ? (lane < SB ? Pc[b][lane & (SB - 1)][i] : Pr[b][i][pi]) : cx<T>{0, 0};
            auto steps = [&](auto full) {
#pragma unroll
                for (int c = 0; c < SB; c++) {
                    if (full || c < dj) {
                        const cx<T> d = lane_value(y[c], c);
                        const bool zero = d.re == (T)0 && d.im == (T)0;
                        const cx<T> inv = zero ? cx<T>{0, 0} : pivot_inverse(d);
                        const cx<T> u = lane > c ? y[c] * inv : cx<T>{0, 0};
#pragma unroll
                        for (int i = c + 1; i < SB; i++) {
                            const cx<T> l = lane_value(y[i], c);
                            y[i].re -= l.re * u.re - l.im * u.im;
                            y[i].im -= l.re * u.im + l.im * u.re;
                        }
                    }
                }
            };
            if (dj == SB)
                steps(std::true_type{});
            else
                steps(std::false_type{});
            if (valid && lane >= SB) {
                cx<T>* dst = A + (k + s0) + (k + e + idx) * lda;
#pragma unroll
                for (int i = 0; i < SB; i++)
                    if (i < dj) {
                        dst[i] = y[i];
                        Pr[b][i][idx] = y[i];
                    }
            }
        }
        __syncthreads();
        if (R > 0) {
            // dj == SB here: a partial sub-panel is the last one
            if (r0 >= e && j0 >= e && r0 < w && j0 < w) {
#pragma unroll 4
                for (int t = 0; t < SB; t++) {
                    cx<T> l[4], u[4];
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        l[i] = Pc[b][t][r0 + i - s0];
#pragma unroll
                    for (int jj = 0; jj < 4; jj++)
                        u[jj] = Pr[b][t][j0 + jj - e];
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int jj = 0; jj < 4; jj++) {
                            a[i][jj].re -= l[i].re * u[jj].re - l[i].im * u[jj].im;
                            a[i][jj].im -= l[i].re * u[jj].im + l[i].im * u[jj].re;
                        }
                }
            }
            stage(j + 1, b ^ 1);
        }
        __syncthreads();
    }
}

// T (w x m, leading dimension w) <-> the block column A21 = A(k + w + c, k + i), 32 values of c a
// workgroup through LDS, all loads of a thread in flight before its stores:
//   BACK false: T(i, c) = A21(c, i);   BACK true: A21(c, i) = T(i, c).
template <typename T, bool BACK>
__global__ __launch_bounds__(256) void col_transpose(cx<T>* __restrict__ A, int64_t lda, int64_t k, int w, int64_t m,
                                                     cx<T>* __restrict__ Tb)
{
    constexpr int TC = kSymCols<T>;
    constexpr int NE = LU_NB_MAX * TC / 256;
    __shared__ cx<T> S[LU_NB_MAX][TC + 1];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * TC;
    const int nc = (int)((m - c0) < TC ? (m - c0) : TC);
    cx<T>* A21 = A + (k + w + c0) + k * lda; // entry (c, i) at A21[c + i lda]
    cx<T>* Tc = Tb + c0 * w;                 // entry (i, c) at Tc[i + c w]
    cx<T> v[NE];
    if (!BACK) {
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, c = e % TC, i = e / TC;
            v[s] = A21[(i < w && c < nc) ? c + (int64_t)i * lda : 0];
        }
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, c = e % TC, i = e / TC;
            S[i][c] = v[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, i = e % LU_NB_MAX, c = e / LU_NB_MAX;
            if (i < w && c < nc)
                Tc[i + (int64_t)c * w] = S[i][c];
        }
    } else {
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, i = e % LU_NB_MAX, c = e / LU_NB_MAX;
            v[s] = Tc[(i < w && c < nc) ? i + (int64_t)c * w : 0];
        }
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, i = e % LU_NB_MAX, c = e / LU_NB_MAX;
            S[i][c] = v[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NE; s++) {
            const int e = tid + 256 * s, c = e % TC, i = e / TC;
            if (i < w && c < nc)
                A21[c + (int64_t)i * lda] = S[i][c];
        }
    }
}

// C[M x N] -= A^T B: the transposed-operand product (A^T(r, kk) = A[kk + r lda]).
template <typename T>
static int gemm_sub_ta(int64_t M, int64_t N, int64_t K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb,
                       cx<T>* C, int64_t ldc, hipStream_t st)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    using TL = typename TileBig<T>::type;
    const int64_t gx = (M + TL::BM - 1) / TL::BM, gy = (N + TL::BN - 1) / TL::BN;
    if (gy > 65535)
        return set_error("gemm: too many column tiles"), -1;
    launch_kernel(&cgemm_sub<T, T, true>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, M, N, K, A, lda, B, ldb, C,
                  ldc);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// Widest block solved by tri_solve_split (FRICO_LU_FUSED_MAX overrides).
static int64_t fused_step_max_cols()
{
    static const int64_t v = [] {
        const char* e = std::getenv("FRICO_LU_FUSED_MAX");
        return e ? (int64_t)std::atoll(e) : (int64_t)4096;
    }();
    return v;
}

static bool solve_blocked();

// ----------------------------------------------------------------------------------------------
// Paired steps of the unpivoted complex64 LU (fp16 products, look-ahead).  Steps a (even) and b
// = a + 1 share their trailing product: after both block columns are factored, the rows of both
// blocks are solved for the columns right of the pair (U12(a) = L11(a)^{-1} A(a, :);
// A(b, :) -= L(b, a) U12(a); U12(b) = L11(b)^{-1} A(b, :)) and the rows below take
// L21(a) U12(a) + L21(b) U12(b) in one K = 256 pass over C.  The next pair's 256 columns are
// updated first, on the main stream, so that its first diagonal block can be factored while the
// side stream updates the remaining columns:
//   main: narrow(a -> b), column(b), [wait W(prev)] next-pair block rows + product, column(a + 2)
//   side: [wait column(b)] block rows of the remaining columns, [wait next-pair product] product
// L21 of step s is split into ha[s mod 4]; U12 of the main stream's products into hb[0..1], of
// the side stream's into hb[2..3].
// ----------------------------------------------------------------------------------------------
static bool pairs_on(const LuWork& wk, bool f16, bool direct, int nb)
{
    static const bool on = [] {
        const char* e = std::getenv("FRICO_LU_PAIRS");
        return !(e && std::atoi(e) == 0);
    }();
    return on && f16 && !direct && wk.nsplit >= 4 && nb == 128 && wk.fused_step;
}

static int lu_nopiv_pairs(LuWork& wk, c64* A, int64_t n, int64_t lda, int nb, hipStream_t st, bool lookahead,
                          KernelTimer* gemm_timer, int64_t naug)
{
    const int64_t npan = (n + nb - 1) / nb;
    // without the look-ahead the side stream's work runs on the main stream, in the same order and
    // with the same arithmetic, and no event is used
    hipStream_t side = lookahead ? wk.side : st;
    auto record = [&](hipEvent_t e, hipStream_t s) -> hipError_t { return lookahead ? hipEventRecord(e, s) : hipSuccess; };
    auto wait = [&](hipStream_t s, hipEvent_t e) -> hipError_t { return lookahead ? hipStreamWaitEvent(s, e, 0) : hipSuccess; };
    c64* Tb = (c64*)wk.tbuf;
    const int64_t fused_max = fused_step_max_cols();
    auto width = [&](int64_t k) { return (int)((n - k) < nb ? (n - k) : nb); };
    // diagonal block and L21 of the step at k, L21 split into ha[ai]
    auto col_step = [&](int64_t k, int w, int ai) -> int {
        const int64_t m = n - k - w;
        launch_kernel(&lu_diag<float>, dim3(1), dim3(1024), 0, st, A, lda, k, w, wk.flags + 1, Batch{});
        if (m <= 0)
            return 0;
        if (m <= fused_max)
            return launch_tri_solve_split(A, lda, k, w, 0, m, true, wk.ha[ai], wk.hea[ai], wk.hacap, st);
        const dim3 g((unsigned)((m + kSymCols<float> - 1) / kSymCols<float>));
        launch_kernel(&col_transpose<float, false>, g, dim3(256), 0, st, A, lda, k, w, m, Tb);
        for (int s = 0; s < w; s += 64) {
            const int ws = (w - s) < 64 ? (w - s) : 64;
            launch_trsm<float, false, true, false>(ws, A + (k + s) + (k + s) * lda, lda, Tb + s, w, Tb + s, w, m, st);
            if (s + ws < w) {
                int rc = gemm_sub_ta<float>(w - s - ws, m, ws, A + (k + s) + (k + s + ws) * lda, lda, Tb + s, w,
                                            Tb + s + ws, w, st);
                if (rc)
                    return rc;
            }
        }
        launch_kernel(&col_transpose<float, true>, g, dim3(256), 0, st, A, lda, k, w, m, Tb);
        return launch_split_f16(Tb, w, m, w, wk.ha[ai], wk.hea[ai], wk.hacap, st);
    };
    // U12 = L11^{-1} A(k .. k+w, c0 .. c0+nc) of the step at k, split into hb[bi]
    auto usolve = [&](int64_t k, int w, int64_t c0, int64_t nc, hipStream_t s, int bi) -> int {
        if (nc <= 0)
            return 0;
        if (nc <= fused_max)
            return launch_tri_solve_split(A, lda, k, w, c0, nc, false, wk.hb[bi], wk.heb[bi], wk.hbcap, s);
        int rc = row_block_solve<float>(A, lda, k, w, c0, nc, s);
        return rc ? rc : launch_split_f16(A + k + c0 * lda, lda, nc, w, wk.hb[bi], wk.heb[bi], wk.hbcap, s);
    };
    // C (M x nc at A(r0, c0)) -= L21 rows [off, off + M) of step ai (K = w) times U12 in hb[bi]
    auto prod1 = [&](int64_t M, int64_t nc, int w, int ai, int64_t off, int bi, int64_t r0, int64_t c0,
                     hipStream_t s) -> int {
        if (M <= 0 || nc <= 0)
            return 0;
        return gemm_sub_f16(M, nc, w, wk.ha[ai] + off * HSLAB, wk.hea[ai] + off, wk.hacap, wk.hb[bi], wk.heb[bi],
                            wk.hbcap, A + r0 + c0 * lda, lda, s, gemm_timer);
    };
    // the pair's product on rows [r0, n) x columns [c0, c0 + nc): L21(a) from its row off_a, L21(b)
    auto prod2 = [&](int64_t nc, int wa, int ia, int64_t off_a, int b0, int wb, int ib, int b1, int64_t r0, int64_t c0,
                     hipStream_t s) -> int {
        const int64_t M = n - r0;
        if (M <= 0 || nc <= 0)
            return 0;
        return gemm_sub_f16_pair(M, nc, wa, wk.ha[ia] + off_a * HSLAB, wk.hea[ia] + off_a, wk.hb[b0], wk.heb[b0], wb,
                                 wk.ha[ib], wk.hea[ib], wk.hb[b1], wk.heb[b1], wk.hacap, wk.hbcap, A + r0 + c0 * lda, lda,
                                 s, gemm_timer);
    };
    hipEvent_t* evP = lookahead ? wk.events : nullptr;
    hipEvent_t* evW = lookahead ? wk.events + npan : nullptr;
    hipEvent_t ev0 = lookahead ? wk.events[2 * npan] : nullptr;
    FRICO_HIP_CHECK(record(ev0, st));
    FRICO_HIP_CHECK(wait(side, ev0));
    int rc;
    if ((rc = col_step(0, width(0), 0)))
        return rc;
    int64_t prev = -1; // pair whose side-stream product the next narrow update waits for
    for (int64_t p = 0; p < npan; p += 2) {
        const int64_t ka = p * nb;
        const int wa = width(ka), ia = (int)(p % 4);
        if (p + 1 >= npan) {
            // a last single block: its rows of the appended columns
            if (naug > 0) {
                if (prev >= 0)
                    FRICO_HIP_CHECK(wait(st, (lookahead ? evW[prev] : nullptr)));
                if ((rc = usolve(ka, wa, n, naug, st, 0)))
                    return rc;
            }
            break;
        }
        const int64_t kb = ka + nb;
        const int wb = width(kb), ib = (int)((p + 1) % 4);
        // step a's update of block b, then block b's factorisation
        if ((rc = usolve(ka, wa, kb, wb, st, 0)) || (rc = prod1(n - kb, wb, wa, ia, 0, 0, kb, kb, st)) ||
            (rc = col_step(kb, wb, ib)))
            return rc;
        FRICO_HIP_CHECK(record((lookahead ? evP[p] : nullptr), st));
        const int64_t c2 = kb + wb;                             // first column right of the pair
        const int64_t nnext = (n - c2) < 2 * nb ? (n - c2) : 2 * nb; // the next pair's matrix columns
        const int64_t cw = c2 + nnext, nw = n + naug - cw;       // the rest, with the appended columns
        // side: the pair's block rows of the remaining columns
        if (nw > 0) {
            FRICO_HIP_CHECK(wait(side, (lookahead ? evP[p] : nullptr)));
            if ((rc = usolve(ka, wa, cw, nw, side, 2)) || (rc = prod1(wb, nw, wa, ia, 0, 2, kb, cw, side)) ||
                (rc = usolve(kb, wb, cw, nw, side, 3)))
                return rc;
        }
        // main: the next pair's columns, which the previous pair's side product updated
        if (nnext > 0) {
            if (prev >= 0)
                FRICO_HIP_CHECK(wait(st, (lookahead ? evW[prev] : nullptr)));
            if ((rc = usolve(ka, wa, c2, nnext, st, 0)) || (rc = prod1(wb, nnext, wa, ia, 0, 0, kb, c2, st)) ||
                (rc = usolve(kb, wb, c2, nnext, st, 1)) || (rc = prod2(nnext, wa, ia, wb, 0, wb, ib, 1, c2, c2, st)))
                return rc;
        }
        // side: the pair's product on the remaining columns, after the next pair's (so that the
        // next diagonal block is dispatched before this product fills the CUs)
        if (nw > 0) {
            FRICO_HIP_CHECK(record(ev0, st));
            FRICO_HIP_CHECK(wait(side, ev0));
            if ((rc = prod2(nw, wa, ia, wb, 2, wb, ib, 3, c2, cw, side)))
                return rc;
            FRICO_HIP_CHECK(record((lookahead ? evW[p] : nullptr), side));
            prev = p;
        }
        if (c2 < n && (rc = col_step(c2, width(c2), (int)((p + 2) % 4))))
            return rc;
    }
    FRICO_HIP_CHECK(record(ev0, side));
    FRICO_HIP_CHECK(wait(st, ev0));
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// ----------------------------------------------------------------------------------------------
// Lockstep batch of the paired-step factorisation: nz systems of one size, each launch of
// lu_nopiv_pairs' sequence (without the look-ahead stream) taking all of them through blockIdx.z,
// so that a diagonal block, a block solve or a product is one launch for the batch instead of one
// per system.  System z: A + z sa; its split operands at w.ha[i] + z w.spa (planes) and
// w.hea[i] + z w.sea (exponents), the same for hb; its first zero pivot in w.info[z].  Blocks are
// solved by tri_solve_split only, so n + naug <= the fused width (4096).
// ----------------------------------------------------------------------------------------------
int lu_nopiv_batched(const BatchLu& w, c64* A, int64_t sa, int64_t n, int64_t lda, int64_t naug, int nz,
                     hipStream_t st)
{
    constexpr int nb = 128;
    if (nz < 1 || nz > 65535)
        return set_error("batched lu: %d systems", nz), -1;
    if (n < 1 || n + naug > fused_step_max_cols())
        return set_error("batched lu: %lld + %lld columns outside 1..%lld", (long long)n, (long long)naug,
                         (long long)fused_step_max_cols()), -1;
    if ((n + 127) / 128 * 128 > w.hacap || (n + naug + 127) / 128 * 128 > w.hbcap)
        return set_error("batched lu: split buffers too small"), -1;
    const int64_t npan = (n + nb - 1) / nb;
    launch_kernel(&zero_u32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, (unsigned*)w.info, (int64_t)nz);
    Batch bd;
    bd.s0 = sa;
    bd.s1 = 1;
    Batch ba, bb, bp;
    ba.s0 = sa;
    ba.s1 = w.spa;
    ba.s2 = w.sea;
    bb.s0 = sa;
    bb.s1 = w.spb;
    bb.s2 = w.seb;
    bp.s0 = w.spa;
    bp.s1 = w.sea;
    bp.s2 = w.spb;
    bp.s3 = w.seb;
    bp.s4 = sa;
    auto width = [&](int64_t k) { return (int)((n - k) < nb ? (n - k) : nb); };
    auto col_step = [&](int64_t k, int wd, int ai) -> int {
        const int64_t m = n - k - wd;
        launch_kernel(&lu_diag<float>, dim3(1, 1, (unsigned)nz), dim3(1024), 0, st, A, lda, k, wd, w.info, bd);
        if (m <= 0)
            return 0;
        return launch_tri_solve_split(A, lda, k, wd, 0, m, true, w.ha[ai], w.hea[ai], w.hacap, st, ba, nz);
    };
    auto usolve = [&](int64_t k, int wd, int64_t c0, int64_t nc, int bi) -> int {
        if (nc <= 0)
            return 0;
        return launch_tri_solve_split(A, lda, k, wd, c0, nc, false, w.hb[bi], w.heb[bi], w.hbcap, st, bb, nz);
    };
    auto prod1 = [&](int64_t M, int64_t nc, int wd, int ai, int64_t off, int bi, int64_t r0, int64_t c0) -> int {
        if (M <= 0 || nc <= 0)
            return 0;
        return gemm_sub_f16(M, nc, wd, w.ha[ai] + off * HSLAB, w.hea[ai] + off, w.hacap, w.hb[bi], w.heb[bi], w.hbcap,
                            A + r0 + c0 * lda, lda, st, nullptr, bp, nz);
    };
    // the pair's product on rows [r0, n): L21(a) from its row off_a (L21(a) starts at block b's first
    // row), L21(b) from its first
    auto prod2 = [&](int64_t nc, int wa, int ia, int64_t off_a, int b0, int wb, int ib, int b1, int64_t r0,
                     int64_t c0) -> int {
        const int64_t M = n - r0;
        if (M <= 0 || nc <= 0)
            return 0;
        return gemm_sub_f16_pair(M, nc, wa, w.ha[ia] + off_a * HSLAB, w.hea[ia] + off_a, w.hb[b0], w.heb[b0], wb,
                                 w.ha[ib], w.hea[ib], w.hb[b1], w.heb[b1], w.hacap, w.hbcap, A + r0 + c0 * lda, lda, st,
                                 nullptr, bp, nz);
    };
    int rc;
    if ((rc = col_step(0, width(0), 0)))
        return rc;
    for (int64_t p = 0; p < npan; p += 2) {
        const int64_t ka = p * nb;
        const int wa = width(ka), ia = (int)(p % 4);
        if (p + 1 >= npan) {
            if (naug > 0 && (rc = usolve(ka, wa, n, naug, 0)))
                return rc;
            break;
        }
        const int64_t kb = ka + nb;
        const int wb = width(kb), ib = (int)((p + 1) % 4);
        if ((rc = usolve(ka, wa, kb, wb, 0)) || (rc = prod1(n - kb, wb, wa, ia, 0, 0, kb, kb)) ||
            (rc = col_step(kb, wb, ib)))
            return rc;
        const int64_t c2 = kb + wb;
        const int64_t nnext = (n - c2) < 2 * nb ? (n - c2) : 2 * nb;
        const int64_t cw = c2 + nnext, nw = n + naug - cw;
        if (nw > 0 && ((rc = usolve(ka, wa, cw, nw, 2)) || (rc = prod1(wb, nw, wa, ia, 0, 2, kb, cw)) ||
                       (rc = usolve(kb, wb, cw, nw, 3))))
            return rc;
        if (nnext > 0 && ((rc = usolve(ka, wa, c2, nnext, 0)) || (rc = prod1(wb, nnext, wa, ia, 0, 0, kb, c2)) ||
                          (rc = usolve(kb, wb, c2, nnext, 1)) ||
                          (rc = prod2(nnext, wa, ia, wb, 0, wb, ib, 1, c2, c2))))
            return rc;
        if (nw > 0 && (rc = prod2(nw, wa, ia, wb, 2, wb, ib, 3, c2, cw)))
            return rc;
        if (c2 < n && (rc = col_step(c2, width(c2), (int)((p + 2) % 4))))
            return rc;
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

template <typename T>
int lu_nopiv_factor(LuWork& wk, cx<T>* A, int64_t n, int64_t lda, int* ipiv, int nb, hipStream_t st,
                    KernelTimer* gemm_timer, int64_t naug)
{
    if (nb > LU_NB_MAX || nb < 1)
        return set_error("lu (no pivoting): block size out of range"), -1;
    const int64_t npan = (n + nb - 1) / nb;
    if (!wk.perm || wk.nperm < npan)
        return set_error("lu (no pivoting): interchange-list workspace too small"), -1;
    if (!wk.tbuf || wk.ntbuf < (int64_t)nb * n)
        return set_error("lu (no pivoting): transpose workspace too small"), -1;
    cx<T>* Tb = (cx<T>*)wk.tbuf;
    const bool lookahead = wk.side != nullptr && wk.events != nullptr && wk.nevents >= 2 * npan + 1;
    // complex64 trailing products on fp16 MFMA with split operands (gemm_f16.hip)
    constexpr bool c64f = std::is_same<T, float>::value;
    const bool f16 = c64f && wk.ha[0] && wk.ha[1] && wk.hb[0] && wk.hb[1];
    if (f16 && (nb > HSPLIT_K || (n + 127) / 128 * 128 > wk.hacap || (n + naug + 127) / 128 * 128 > wk.hbcap))
        return set_error("lu (no pivoting): fp16 split buffers too small"), -1;
    const int64_t nz = npan * PERM_STRIDE;
    launch_kernel(&zero_u32, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, (unsigned*)wk.perm, nz);
    launch_kernel(&zero_u32, dim3(1), dim3(256), 0, st, (unsigned*)wk.flags, (int64_t)2);
    launch_kernel(&iota_pivots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ipiv, n);
    FRICO_HIP_CHECK(hipGetLastError());
    // diagonal block and block column of step k
    const bool direct = c64f && wk.direct_solves;
    // The fused block solves (one launch with the split) pay where a step is short and the host's
    // launches set its pace; wide blocks keep the throughput-oriented solves and a separate split.
    const bool fused_any = f16 && wk.fused_step && !direct;
    const int64_t fused_max = fused_step_max_cols();
    auto fused = [&](int64_t cols) { return fused_any && cols <= fused_max; };
    auto column = [&](int64_t k, int w) -> int {
        const int64_t m = n - k - w;
        launch_kernel(&lu_diag<T>, dim3(1), dim3(1024), 0, st, A, lda, k, w, wk.flags + 1, Batch{});
        if constexpr (c64f) {
            if (fused(m)) {
                // L21 = A21 U11^{-1} and its fp16 split in one launch
                const int par = (int)((k / nb) & 1);
                int rc = m > 0 ? launch_tri_solve_split(reinterpret_cast<c64*>(A), lda, k, w, 0, m, true, wk.ha[par],
                                                        wk.hea[par], wk.hacap, st)
                               : 0;
                if (rc)
                    return rc;
                FRICO_HIP_CHECK(hipGetLastError());
                return 0;
            }
            if (direct) {
                // L21 = A21 U11^{-1} a row per lane, then its fp16 split straight from A
                int rc = launch_solve_l21(reinterpret_cast<c64*>(A), lda, k, w, m, st);
                if (!rc && f16 && m > 0) {
                    const int par = (int)((k / nb) & 1);
                    rc = launch_split_f16_cm(reinterpret_cast<const c64*>(A + (k + w) + k * lda), lda, m, w, wk.ha[par],
                                             wk.hea[par], wk.hacap, st);
                }
                if (rc)
                    return rc;
                FRICO_HIP_CHECK(hipGetLastError());
                return 0;
            }
        }
        if (m > 0) {
            const dim3 g((unsigned)((m + kSymCols<T> - 1) / kSymCols<T>));
            launch_kernel(&col_transpose<T, false>, g, dim3(256), 0, st, A, lda, k, w, m, Tb);
            for (int s = 0; s < w; s += 64) {
                const int ws = (w - s) < 64 ? (w - s) : 64;
                launch_trsm<T, false, true, false>(ws, A + (k + s) + (k + s) * lda, lda, Tb + s, w, Tb + s, w, m, st);
                if (s + ws < w) {
                    int rc = gemm_sub_ta<T>(w - s - ws, m, ws, A + (k + s) + (k + s + ws) * lda, lda, Tb + s, w,
                                            Tb + s + ws, w, st);
                    if (rc)
                        return rc;
                }
            }
            launch_kernel(&col_transpose<T, true>, g, dim3(256), 0, st, A, lda, k, w, m, Tb);
            // L21 (held transposed in Tb, k contiguous per row) split once for the step's products
            if constexpr (c64f) {
                if (f16) {
                    const int par = (int)((k / nb) & 1);
                    int rc = launch_split_f16(reinterpret_cast<const c64*>(Tb), w, m, w, wk.ha[par], wk.hea[par],
                                              wk.hacap, st);
                    if (rc)
                        return rc;
                }
            }
        }
        FRICO_HIP_CHECK(hipGetLastError());
        return 0;
    };
    // update of columns [c0, c0 + nc) by step k: U12 = L11^{-1} A12, A22 -= L21 U12
    // parts: 1 = the block-row solve, 2 = the product, 3 = both
    auto update = [&](int64_t k, int w, int64_t c0, int64_t nc, hipStream_t s, int parts) -> int {
        if (nc <= 0)
            return 0;
        int rc = 0;
        if constexpr (c64f) {
            if (fused(nc)) {
                // U12 = L11^{-1} A12 with its split (per column, on this stream's buffer), then the product
                const int64_t m = n - k - w;
                const int par = (int)((k / nb) & 1), sb = s == st ? 0 : 1;
                if (parts & 1)
                    rc = launch_tri_solve_split(reinterpret_cast<c64*>(A), lda, k, w, c0, nc, false, wk.hb[sb], wk.heb[sb],
                                                wk.hbcap, s);
                if (rc || !(parts & 2) || m <= 0)
                    return rc;
                return gemm_sub_f16(m, nc, w, wk.ha[par], wk.hea[par], wk.hacap, wk.hb[sb], wk.heb[sb], wk.hbcap,
                                    reinterpret_cast<c64*>(A + (k + w) + c0 * lda), lda, s, gemm_timer);
            }
        }
        if (parts & 1) {
            if constexpr (c64f) {
                rc = direct ? launch_solve_u12(reinterpret_cast<c64*>(A), lda, k, w, c0, nc, s)
                            : row_block_solve<T>(A, lda, k, w, c0, nc, s);
            } else {
                rc = row_block_solve<T>(A, lda, k, w, c0, nc, s);
            }
        }
        if (rc || !(parts & 2))
            return rc;
        if constexpr (c64f) {
            if (f16) {
                const int64_t m = n - k - w;
                if (m <= 0)
                    return 0;
                // U12 of these columns split on this stream's buffer, then the product with the step's L21
                const int par = (int)((k / nb) & 1), sb = s == st ? 0 : 1;
                rc = launch_split_f16(reinterpret_cast<const c64*>(A + k + c0 * lda), lda, nc, w, wk.hb[sb],
                                      wk.heb[sb], wk.hbcap, s);
                return rc ? rc
                          : gemm_sub_f16(m, nc, w, wk.ha[par], wk.hea[par], wk.hacap, wk.hb[sb], wk.heb[sb], wk.hbcap,
                                         reinterpret_cast<c64*>(A + (k + w) + c0 * lda), lda, s, gemm_timer);
            }
        }
        return gemm_sub<T>(n - k - w, nc, w, A + (k + w) + k * lda, lda, A + k + c0 * lda, lda, A + (k + w) + c0 * lda,
                           lda, s, gemm_timer);
    };
    auto width = [&](int64_t k) { return (int)((n - k) < nb ? (n - k) : nb); };
    int rc;
    if constexpr (c64f) {
        if (pairs_on(wk, f16, direct, nb)) {
            if ((rc = lu_nopiv_pairs(wk, reinterpret_cast<c64*>(A), n, lda, nb, st, lookahead, gemm_timer, naug)))
                return rc;
            // the blocked solves apply the factors directly: no inverse or W tables
            return solve_blocked() ? 0 : solve_tables<T>(wk, A, n, lda, nb, st);
        }
    }
    if (!lookahead) {
        for (int64_t p = 0; p < npan; p++) {
            const int64_t k = p * nb;
            const int w = width(k);
            if ((rc = column(k, w)) || (rc = update(k, w, k + w, n + naug - k - w, st, 3)))
                return rc;
        }
    } else {
        hipEvent_t* evP = wk.events;
        hipEvent_t* evW = wk.events + npan;
        hipEvent_t ev0 = wk.events[2 * npan];
        FRICO_HIP_CHECK(hipEventRecord(ev0, st));
        FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, ev0, 0));
        if ((rc = column(0, width(0))))
            return rc;
        bool prev_wide = false;
        for (int64_t p = 0; p < npan; p++) {
            const int64_t k = p * nb;
            const int w = width(k);
            const int64_t nright = n - k - w;
            if (nright <= 0) {
                // the last block's update of the appended columns
                if (naug > 0) {
                    if (prev_wide)
                        FRICO_HIP_CHECK(hipStreamWaitEvent(st, evW[p - 1], 0));
                    if ((rc = update(k, w, n, naug, st, 3)))
                        return rc;
                }
                break;
            }
            const int64_t nnext = nright < nb ? nright : nb;
            const int64_t nrest = nright + naug - nnext; // the appended columns go with the wide part
            // The wide part's block-row solve starts with the step; its product waits for the
            // narrow update, so that the next diagonal block (one 1024-thread workgroup, which
            // does not fit beside a CU's three product workgroups) is dispatched before the product
            // fills the CUs, and factors beside it instead of after it.
            FRICO_HIP_CHECK(hipEventRecord(evP[p], st));
            if (nrest > 0) {
                FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, evP[p], 0));
                if ((rc = update(k, w, k + w + nnext, nrest, wk.side, 1)))
                    return rc;
            }
            if (prev_wide)
                FRICO_HIP_CHECK(hipStreamWaitEvent(st, evW[p - 1], 0));
            if ((rc = update(k, w, k + w, nnext, st, 3)))
                return rc;
            if (nrest > 0) {
                FRICO_HIP_CHECK(hipEventRecord(ev0, st));
                FRICO_HIP_CHECK(hipStreamWaitEvent(wk.side, ev0, 0));
                if ((rc = update(k, w, k + w + nnext, nrest, wk.side, 2)))
                    return rc;
                FRICO_HIP_CHECK(hipEventRecord(evW[p], wk.side));
            }
            if ((rc = column(k + w, width(k + w))))
                return rc;
            prev_wide = nrest > 0;
        }
        FRICO_HIP_CHECK(hipEventRecord(ev0, wk.side));
        FRICO_HIP_CHECK(hipStreamWaitEvent(st, ev0, 0));
    }
    if constexpr (c64f) {
        if (solve_blocked())
            return 0; // the blocked solves apply the factors directly: no inverse or W tables
    }
    return solve_tables<T>(wk, A, n, lda, nb, st);
}

constexpr int RESID_ROWS = 512; // rows of the solution per residual product beside the substitution

// ----------------------------------------------------------------------------------------------
// Blocked triangular solves (complex64 factors without interchanges to apply).  The rows are taken
// in blocks of BSOLVE_ROWS = RESID_ROWS.  One launch of block_solve solves a block's diagonal
// block against the right-hand sides, and one product carries the block's solution to the rows
// beyond it, so a block costs two launches instead of two per 64-row chunk.
//
// block_solve: workgroup = BSOLVE_COLS right-hand sides, whose rows of the block it keeps in LDS,
// 256 threads (one wavefront a SIMD) under 168 registers a lane.  The backward solve runs beside
// the FP64 residual products (two workgroups a CU at 196 registers); at this size a block_solve
// workgroup fits on a CU as soon as one of them finishes, where a 1024-thread form (128
// registers, all of a CU's register file) waited for a CU with no other work: 530 us against 79
// alone, every second block.  Per 64-row chunk of the block, in solve order:
//   * every thread of the update issues the loads of its first factor columns (they do not depend
//     on the chunk's solution), then the chunk's diagonal block, loaded during the previous chunk,
//     goes to LDS;
//   * substitution, one wavefront per right-hand side, one row per lane: the rows' finished values
//     are broadcast by readlane, the coefficients come from LDS a few steps ahead (U_jj is applied
//     by substitution, not through an explicit inverse: an inverse loses accuracy in proportion to
//     cond(U_jj) on EFIE matrices);
//   * the rows of the block not yet solved take the chunk off, y_r -= D(r, chunk) x_chunk: each
//     thread owns two rows and streams their factor entries BSOLVE_G columns at a time.
// UPPER: U (non-unit diagonal), chunks from the last, rows above; else L (unit diagonal), chunks
// from the first, rows below.
// ----------------------------------------------------------------------------------------------
constexpr int BSOLVE_ROWS = RESID_ROWS;
constexpr int BSOLVE_COLS = 4;
constexpr int BSOLVE_PF = 4; // coefficient loads in flight ahead of the substitution step

// Substitution of one chunk (w rows, w == 64 when FULL) for one right-hand side held as b (lane =
// row); sd[c][r] = D(r, c); rd = 1 / D(lane, lane) (UPPER).  Returns x (lane = row).
template <bool UPPER, bool FULL>
__device__ inline c64 chunk_substitute(c64 b, c64 rd, int w, const c64 (*sd)[65], int lane)
{
    if constexpr (FULL) {
        constexpr int W = 64;
        c64 ring[BSOLVE_PF];
#pragma unroll
        for (int p = 0; p < BSOLVE_PF; p++) {
            const int i = UPPER ? W - 1 - p : p;
            const bool dep = UPPER ? lane < i : lane > i;
            ring[p] = dep ? sd[i][lane] : c64{0.0f, 0.0f};
        }
#pragma unroll
        for (int t = 0; t < W; t++) {
            const int i = UPPER ? W - 1 - t : t;
            const c64 e = ring[t % BSOLVE_PF];
            if (t + BSOLVE_PF < W) {
                const int i2 = UPPER ? W - 1 - (t + BSOLVE_PF) : t + BSOLVE_PF;
                const bool dep2 = UPPER ? lane < i2 : lane > i2;
                ring[t % BSOLVE_PF] = dep2 ? sd[i2][lane] : c64{0.0f, 0.0f};
            }
            const c64 xi = lane_value(UPPER ? b * rd : b, i);
            b.re -= e.re * xi.re - e.im * xi.im;
            b.im -= e.re * xi.im + e.im * xi.re;
        }
    } else {
        // a short chunk (the last rows of the matrix): one coefficient read a step
        for (int t = 0; t < w; t++) {
            const int i = UPPER ? w - 1 - t : t;
            const bool dep = UPPER ? lane < i : lane > i;
            const c64 e = dep ? sd[i][lane] : c64{0.0f, 0.0f};
            const c64 xi = lane_value(UPPER ? b * rd : b, i);
            b.re -= e.re * xi.re - e.im * xi.im;
            b.im -= e.re * xi.im + e.im * xi.re;
        }
    }
    return UPPER ? b * rd : b;
}

// Loads are branch-free: inactive rows read row kb and columns past a short chunk read its last
// column; the result of inactive rows is dropped, and columns past a short chunk meet x = 0.
constexpr int BSOLVE_NT = 256;
constexpr int BSOLVE_G = 4;    // factor columns of the update a load group
constexpr int BSOLVE_RING = 2; // a group's loads are issued this many groups before its use

template <bool UPPER>
__global__ __launch_bounds__(BSOLVE_NT, 3) void block_solve(const c64* __restrict__ A, int64_t lda, int64_t kb, int bw,
                                                            const c64* Y, int64_t ldy, c64* X, int64_t ldx,
                                                            int64_t nrhs)
{
    constexpr int RPT = BSOLVE_ROWS / BSOLVE_NT; // update rows a thread
    constexpr int NG = 64 / BSOLVE_G;
    constexpr int DPT = 64 * 64 / BSOLVE_NT; // diagonal-block entries a thread
    __shared__ c64 sy[BSOLVE_ROWS + 1][BSOLVE_COLS]; // + a spare row for inactive threads' stores
    __shared__ c64 sd[64][65]; // sd[c][r] = D(r, c) of the chunk's diagonal block
    __shared__ c64 sx[64][BSOLVE_COLS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t q0 = (int64_t)blockIdx.x * BSOLVE_COLS;
    for (int e = tid; e < BSOLVE_ROWS * BSOLVE_COLS; e += BSOLVE_NT) {
        const int r = e % BSOLVE_ROWS, q = e / BSOLVE_ROWS;
        sy[r][q] = (r < bw && q0 + q < nrhs) ? Y[(kb + r) + (q0 + q) * ldy] : c64{0.0f, 0.0f};
    }
    const int nch = (bw + 63) / 64;
    // Loads address a column through a wave-uniform pointer and the row through a 32-bit byte
    // offset, so that a load costs one register of address, not a 64-bit pointer a load.
    const int wvu = __builtin_amdgcn_readfirstlane(wv);
    auto at = [&](int64_t col, unsigned boff) {
        return *reinterpret_cast<const c64*>(reinterpret_cast<const char*>(A + col * lda) + boff);
    };
    auto dload = [&](int j, c64* dv) {
        const int c0 = 64 * j, w = (bw - c0) < 64 ? (bw - c0) : 64;
        const unsigned boff = (unsigned)(kb + c0 + (lane < w ? lane : w - 1)) * (unsigned)sizeof(c64);
#pragma unroll
        for (int t = 0; t < DPT; t++) {
            const int c = wvu + (BSOLVE_NT / 64) * t, cc = c < w ? c : w - 1; // e = tid + 256 t: row lane
            dv[t] = at(kb + c0 + cc, boff);
        }
    };
    c64 dpre[DPT];
    dload(UPPER ? nch - 1 : 0, dpre);
    for (int s = 0; s < nch; s++) {
        const int j = UPPER ? nch - 1 - s : s;
        const int c0 = 64 * j, w = (bw - c0) < 64 ? (bw - c0) : 64;
        const int rb = UPPER ? 0 : c0 + w, re = UPPER ? c0 : bw; // rows of the update
        unsigned roff[RPT];
        bool act[RPT];
#pragma unroll
        for (int h = 0; h < RPT; h++) {
            const int row = rb + tid + BSOLVE_NT * h;
            act[h] = row < re;
            roff[h] = (unsigned)(kb + (act[h] ? row : 0)) * (unsigned)sizeof(c64);
        }
        c64 ring[BSOLVE_RING][BSOLVE_G][RPT];
        auto gload = [&](int g, c64 (*buf)[RPT]) {
#pragma unroll
            for (int t = 0; t < BSOLVE_G; t++) {
                const int c = BSOLVE_G * g + t, cc = c < w ? c : w - 1;
#pragma unroll
                for (int h = 0; h < RPT; h++)
                    buf[t][h] = at(kb + c0 + cc, roff[h]);
            }
        };
#pragma unroll
        for (int g = 0; g < BSOLVE_RING; g++)
            gload(g, ring[g]);
        __syncthreads(); // the previous chunk is done with sd and sx
        {
            const int c0n = 64 * j, wn = (bw - c0n) < 64 ? (bw - c0n) : 64;
#pragma unroll
            for (int t = 0; t < DPT; t++) {
                const int e = tid + BSOLVE_NT * t, r = e & 63, c = e >> 6;
                sd[c][r] = (r < wn && c < wn) ? dpre[t] : c64{0.0f, 0.0f};
            }
        }
        if (s + 1 < nch)
            dload(UPPER ? j - 1 : j + 1, dpre);
        __syncthreads();
        {
            const c64 b = lane < w ? sy[c0 + lane][wv] : c64{0.0f, 0.0f};
            c64 rd{1.0f, 0.0f};
            if (UPPER && lane < w)
                rd = cdiv(c64{1.0f, 0.0f}, sd[lane][lane]);
            const c64 x = w == 64 ? chunk_substitute<UPPER, true>(b, rd, w, sd, lane)
                                  : chunk_substitute<UPPER, false>(b, rd, w, sd, lane);
            sx[lane][wv] = x;
            if (lane < w && q0 + wv < nrhs)
                X[(kb + c0 + lane) + (q0 + wv) * ldx] = x;
        }
        __syncthreads();
        c64 acc[RPT][BSOLVE_COLS];
#pragma unroll
        for (int h = 0; h < RPT; h++)
#pragma unroll
            for (int q = 0; q < BSOLVE_COLS; q++)
                acc[h][q] = c64{0.0f, 0.0f};
#pragma unroll
        for (int g = 0; g < NG; g++) {
            c64(*cur)[RPT] = ring[g % BSOLVE_RING];
#pragma unroll
            for (int t = 0; t < BSOLVE_G; t++) {
                c64 x[BSOLVE_COLS];
#pragma unroll
                for (int q = 0; q < BSOLVE_COLS; q++)
                    x[q] = sx[BSOLVE_G * g + t][q];
#pragma unroll
                for (int h = 0; h < RPT; h++)
#pragma unroll
                    for (int q = 0; q < BSOLVE_COLS; q++) {
                        acc[h][q].re += cur[t][h].re * x[q].re - cur[t][h].im * x[q].im;
                        acc[h][q].im += cur[t][h].re * x[q].im + cur[t][h].im * x[q].re;
                    }
            }
            if (g + BSOLVE_RING < NG)
                gload(g + BSOLVE_RING, cur);
        }
        // unconditional stores (inactive rows into the spare row): a store under act[h] lets the
        // compiler sink the whole accumulation into the branch and keep every operand live to it
#pragma unroll
        for (int h = 0; h < RPT; h++) {
            const int row = act[h] ? rb + tid + BSOLVE_NT * h : BSOLVE_ROWS;
#pragma unroll
            for (int q = 0; q < BSOLVE_COLS; q++) {
                sy[row][q].re -= acc[h][q].re;
                sy[row][q].im -= acc[h][q].im;
            }
        }
    }
}

template <bool UPPER>
static void launch_block_solve(const c64* A, int64_t lda, int64_t kb, int bw, const c64* Y, int64_t ldy, c64* X,
                               int64_t ldx, int64_t nrhs, hipStream_t st)
{
    launch_kernel(&block_solve<UPPER>, dim3((unsigned)((nrhs + BSOLVE_COLS - 1) / BSOLVE_COLS)), dim3(BSOLVE_NT), 0, st,
                  A, lda, kb, bw, Y, ldy, X, ldx, nrhs);
}

// FRICO_SOLVE_BLOCKED=0 returns the complex64 solves to one substitution and one product per chunk.
static bool solve_blocked()
{
    static const bool on = [] {
        const char* e = std::getenv("FRICO_SOLVE_BLOCKED");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

bool lu_blocked_solves() { return solve_blocked(); }
int lu_solve_block_rows() { return BSOLVE_ROWS; }

// A product of a blocked solve, C -= A(factors) X(block of the solution, K rows): on fp16 MFMA
// with the factors' exponents of block kblk when the refinement provides them, else fp32.
static int solve_product(const ResidStream* rs, bool second, int64_t kblk, int64_t row0, int64_t M, int64_t N,
                         int64_t K, const c64* A, int64_t lda, const c64* X, int64_t ldx, c64* C, int64_t ldc,
                         hipStream_t st)
{
    if (rs && rs->aexp)
        return gemm_sub_f16s(M, N, K, A, lda, rs->aexp + kblk * rs->ldae + row0, X, ldx, second ? rs->bexp2 : rs->bexp,
                             C, ldc, st);
    return gemm_sub<float>(M, N, K, A, lda, X, ldx, C, ldc, st, nullptr);
}

// Residual product of one finished block of the solution beside the substitution (see lu_backward).
static int resid_block(const ResidStream* rs, int64_t n, int64_t kb, int64_t bw, const c64* X, int64_t ldx,
                       int64_t nrhs, hipStream_t st)
{
    const int64_t b = kb / RESID_ROWS;
    FRICO_HIP_CHECK(hipEventRecord(rs->ev[b], st));
    FRICO_HIP_CHECK(hipStreamWaitEvent(rs->side, rs->ev[b], 0));
    int rc = launch_add_correction(rs->S + kb, rs->lds, X + kb, ldx, bw, nrhs, rs->side);
    if (!rc)
        rc = gemm_sub<double>(n, nrhs, bw, rs->Z + kb * rs->ldz, rs->ldz, rs->S + kb, rs->lds, rs->R, rs->ldr, rs->side,
                              nullptr, rs->exact);
    return rc;
}

// Blocked backward half (complex64): per block from the last, X(block) = U(block)^{-1} Y(block),
// then Y(above) -= U(above, block) X(block).
static int lu_backward_blocked(const c64* A, int64_t n, int64_t lda, c64* Y, int64_t ldy, c64* X, int64_t ldx,
                               int64_t nrhs, hipStream_t st, const ResidStream* rs)
{
    const int64_t nblk = (n + BSOLVE_ROWS - 1) / BSOLVE_ROWS;
    const bool pair = rs && rs->R2;
    int64_t pend = -1, pend_bw = 0; // with R2: a finished block whose residual product waits for the next
    for (int64_t b = nblk - 1; b >= 0; b--) {
        const int64_t kb = b * BSOLVE_ROWS, bw = (n - kb) < BSOLVE_ROWS ? (n - kb) : BSOLVE_ROWS;
        launch_block_solve<true>(A, lda, kb, (int)bw, Y, ldy, X, ldx, nrhs, st);
        FRICO_HIP_CHECK(hipGetLastError());
        int rc = 0;
        if (rs && !pair) {
            rc = resid_block(rs, n, kb, bw, X, ldx, nrhs, st);
        } else if (pair) {
            FRICO_HIP_CHECK(hipEventRecord(rs->ev[b], st));
            FRICO_HIP_CHECK(hipStreamWaitEvent(rs->side, rs->ev[b], 0));
            rc = launch_add_correction(rs->S + kb, rs->lds, X + kb, ldx, bw, nrhs, rs->side);
            if (!rc && pend < 0) {
                pend = kb;
                pend_bw = bw;
            } else if (!rc) {
                GemmPair<double> gp;
                gp.A[0] = rs->Z + pend * rs->ldz;
                gp.B[0] = rs->S + pend;
                gp.C[0] = rs->R;
                gp.K[0] = pend_bw;
                gp.A[1] = rs->Z + kb * rs->ldz;
                gp.B[1] = rs->S + kb;
                gp.C[1] = rs->R2;
                gp.K[1] = bw;
                gp.lda = rs->ldz;
                gp.ldb = rs->lds;
                gp.ldc = rs->ldr;
                rc = gemm_sub_pair(n, nrhs, gp, rs->side, rs->exact);
                pend = -1;
            }
        }
        if (!rc && kb > 0)
            rc = solve_product(rs, false, b, 0, kb, nrhs, bw, A + kb * lda, lda, X + kb, ldx, Y, ldy, st);
        if (rc)
            return rc;
    }
    if (pair) {
        int rc = 0;
        if (pend >= 0)
            rc = gemm_sub<double>(n, nrhs, pend_bw, rs->Z + pend * rs->ldz, rs->ldz, rs->S + pend, rs->lds, rs->R,
                                  rs->ldr, rs->side, nullptr, rs->exact);
        if (rc)
            return rc;
        launch_kernel(&add_into, dim3((unsigned)((n + 255) / 256), (unsigned)nrhs), dim3(256), 0, rs->side, rs->R,
                      (const c128*)rs->R2, rs->ldr, n);
        FRICO_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// Blocked forward half (complex64, unit L, no interchanges), in place: per block from the first,
// B(block) = L(block)^{-1} B(block), then B(below) -= L(below, block) B(block).  With a second
// stream (rs), the update of the next block's rows stays on the main stream, so that the next
// block's substitution starts after it, while the rows below the next block are updated on the
// second stream; the main stream's update of a block's rows waits for the previous block's update
// of the rows below (both write those rows).
static int lu_forward_blocked(const c64* A, int64_t n, int64_t lda, c64* B, int64_t ldb, int64_t nrhs, hipStream_t st,
                              const ResidStream* rs)
{
    const int64_t nblk = (n + BSOLVE_ROWS - 1) / BSOLVE_ROWS;
    const bool ahead = rs && rs->side && rs->nev >= nblk + 1;
    bool pending = false; // a rows-below update of the previous block on the second stream
    for (int64_t b = 0; b < nblk; b++) {
        const int64_t kb = b * BSOLVE_ROWS, bw = (n - kb) < BSOLVE_ROWS ? (n - kb) : BSOLVE_ROWS;
        launch_block_solve<false>(A, lda, kb, (int)bw, B, ldb, B, ldb, nrhs, st);
        FRICO_HIP_CHECK(hipGetLastError());
        const int64_t below = n - kb - bw;
        if (below <= 0)
            break;
        if (!ahead) {
            int rc = solve_product(rs, false, b, kb + bw, below, nrhs, bw, A + (kb + bw) + kb * lda, lda, B + kb, ldb,
                                   B + kb + bw, ldb, st);
            if (rc)
                return rc;
            continue;
        }
        const int64_t nx = below < BSOLVE_ROWS ? below : BSOLVE_ROWS, rest = below - nx;
        if (pending) {
            FRICO_HIP_CHECK(hipEventRecord(rs->ev[b], rs->side));
            FRICO_HIP_CHECK(hipStreamWaitEvent(st, rs->ev[b], 0));
            pending = false;
        }
        int rc = solve_product(rs, false, b, kb + bw, nx, nrhs, bw, A + (kb + bw) + kb * lda, lda, B + kb, ldb,
                               B + kb + bw, ldb, st);
        if (rc)
            return rc;
        if (rest > 0) {
            // the second stream needs this block's solution, which the main stream has at this point
            FRICO_HIP_CHECK(hipEventRecord(rs->ev[nblk], st));
            FRICO_HIP_CHECK(hipStreamWaitEvent(rs->side, rs->ev[nblk], 0));
            rc = solve_product(rs, true, b, kb + bw + nx, rest, nrhs, bw, A + (kb + bw + nx) + kb * lda, lda, B + kb, ldb,
                               B + kb + bw + nx, ldb, rs->side);
            if (rc)
                return rc;
            pending = true;
        }
    }
    if (pending) {
        FRICO_HIP_CHECK(hipEventRecord(rs->ev[0], rs->side));
        FRICO_HIP_CHECK(hipStreamWaitEvent(st, rs->ev[0], 0));
    }
    return 0;
}

// Backward half of the solve: per 64-row chunk j from the last, x_j = U_jj^{-1} y_j by substitution
// (from Y into X, or in place when Y == X), then y_above -= U(above, j) x_j.  With rs (complex64
// solves of a refinement), each RESID_ROWS-row block of X, once final, is added to the FP64 solution and
// taken off the residual on a second stream while the substitution goes on:
// S(block) += X(block), R -= Z(:, block) S(block); with R = B beforehand, R ends as B - Z S.
// ----------------------------------------------------------------------------------------------
// Chained triangular solves for a few right-hand sides (complex64 factors without interchanges).
// With one to four right-hand sides a blocked solve is a chain of launches, each a single
// workgroup's substitution followed by a product with one useful column, so a solve of a few
// thousand rows costs about a millisecond of launch and wait latency.  Here one launch solves the
// whole triangle (of each of nz systems): a workgroup takes 64-row chunks from a counter in solve
// order (chunk 0 first for L, forward; the last for U, backward; the systems interleaved), so it
// depends only on chunks taken before its own by running workgroups: neither co-residency nor the
// dispatch order is assumed, and any grid size is safe.  The dependency chunks i are taken in
// the order they finish, up to CHAIN_WIN a step: wave 0 waits for the first one's x_i and takes the
// following ones already published in the same poll, and every thread takes D(j, i) x_i off its
// partial sums (row = lane, sixteen columns of the chunk = wave), the blocks of D loaded before
// the wait.  A workgroup far behind the front so takes four chunks per round trip of the exchange.
// Then the four waves' partial sums are reduced, wave q substitutes right-hand side q with the
// diagonal block (from LDS, as in block_solve) and publishes x_j.
// The exchange area holds, per chunk, row and right-hand side, two 64-bit words {sequence, re} and
// {sequence, im} written and read with relaxed agent-scope atomics (as the panel factorisation's
// exchange): a reader takes a word only when it carries the launch's sequence number, so the area
// is cleared (by a kernel, which a recorded graph replays) before each solve and the forward and
// backward halves use sequence numbers 1 and 2.  Word 0 is the abort flag of the bounded waits; a
// workgroup whose wait gives up writes NaN into its rows, so the refinement reports the solve.
// ----------------------------------------------------------------------------------------------
constexpr int CHAIN_MAXR = 4;

// per system: word 0 the abort flag (low half), word 1 the forward and backward launches' item
// counters (only system 0's are used), then the x words
size_t lu_chain_words(int64_t n, int64_t nrhs) { return 2 + (size_t)((n + 63) / 64) * 64 * (size_t)nrhs * 2; }

constexpr int CHAIN_WIN = 4; // dependency chunks taken per wait

template <bool UPPER>
__global__ __launch_bounds__(256) void chain_solve(const c64* __restrict__ A0, int64_t lda, int64_t n, c64* B0,
                                                   int64_t ldb, int nr, unsigned long long* words0, unsigned seq,
                                                   Batch bt, int nz)
{
    __shared__ c64 sd[64][65];                         // sd[c][r] = D(r, c), the chunk's diagonal block
    __shared__ c64 sx[2][CHAIN_WIN][64][CHAIN_MAXR];   // x of the chunks of a step (two steps)
    __shared__ c64 sp[4][64][CHAIN_MAXR];              // the four waves' partial sums
    __shared__ int s_ok, s_nav[2], s_item;
    // Items (system z, chunk number t) are taken from a counter in the order t nz + z, so that every
    // item's dependencies (the same system's earlier chunks) are taken before it, by workgroups that
    // are running: a workgroup waits only on items held by running workgroups, whatever the grid
    // size, the dispatch order or the other kernels on the GPU.  System z at A0 + z s0, B0 + z s1,
    // its exchange area at words0 + z s2.
    const int nch = (int)((n + 63) / 64);
    int* counter = reinterpret_cast<int*>(words0 + 1) + (UPPER ? 1 : 0);
    const int64_t nitems = (int64_t)nch * nz;
    for (;;) {
    if (threadIdx.x == 0)
        s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int64_t item = s_item;
    if (item >= nitems)
        break;
    const int zs = (int)(item % nz), jj = (int)(item / nz);
    const c64* __restrict__ A = A0 + zs * bt.s0;
    c64* B = B0 + zs * bt.s1;
    unsigned long long* words = words0 + zs * bt.s2;
    int* abort_flag = reinterpret_cast<int*>(words);
    unsigned long long* xw = words + 2;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int j = UPPER ? nch - 1 - jj : jj;
    const int64_t r0 = 64 * (int64_t)j;
    const int wj = (n - r0) < 64 ? (int)(n - r0) : 64;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        sd[c][r] = (r < wj && c < wj) ? A[(r0 + r) + (r0 + c) * lda] : c64{0.0f, 0.0f};
    }
    if (tid == 0)
        s_ok = 1;
    // this thread's row of D(j, i) (inactive rows read row r0) and its sixteen columns
    const int64_t arow = r0 + (lane < wj ? lane : 0);
    const int ndep = UPPER ? nch - 1 - j : j;
    auto lload = [&](int t, c64* dst) {
        const int tt = t < ndep ? t : ndep - 1; // past the last dependency: a valid block, unused
        const int i = UPPER ? nch - 1 - tt : tt;
        const int64_t c0 = 64 * (int64_t)i + 16 * wv;
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const int64_t cc = (c0 + c) < n ? (c0 + c) : n - 1; // past a short chunk: x = 0 there
            dst[c] = A[arow + cc * lda];
        }
    };
    c64 acc[CHAIN_MAXR];
#pragma unroll
    for (int q = 0; q < CHAIN_MAXR; q++)
        acc[q] = c64{0.0f, 0.0f};
    // the chunk's right-hand side (wave q: column q), read now rather than on the chain's path
    const c64 bq = (wv < nr && lane < wj) ? B[(r0 + lane) + wv * ldb] : c64{0.0f, 0.0f};
    __syncthreads(); // sd and s_ok written
    // Per step: the blocks of the next CHAIN_WIN dependency chunks are loaded, wave 0 waits for the
    // first one's x and takes every following one already published in the same poll (up to
    // CHAIN_WIN), and every thread takes the chunks found off its partial sums.
    int t = 0, par = 0;
    while (t < ndep && s_ok) {
        c64 lv[CHAIN_WIN][16];
#pragma unroll
        for (int u = 0; u < CHAIN_WIN; u++)
            lload(t + u, lv[u]);
        if (wv == 0) {
            bool ok = true;
            int nav = 0;
            unsigned long long v[CHAIN_WIN][2 * CHAIN_MAXR];
            long spins = 0;
            for (;;) {
                nav = 0;
                bool run = true;
#pragma unroll
                for (int u = 0; u < CHAIN_WIN; u++) {
                    if (t + u < ndep) {
                        const int i = UPPER ? nch - 1 - (t + u) : t + u;
                        const unsigned long long* src = xw + ((size_t)i * 64 + lane) * (size_t)nr * 2;
                        bool ready = true;
#pragma unroll
                        for (int q = 0; q < 2 * CHAIN_MAXR; q++)
                            if (q < 2 * nr) {
                                v[u][q] = ll_load(src + q);
                                ready = ready && (unsigned)(v[u][q] >> 32) == seq;
                            }
                        ready = __all(ready);
                        run = run && ready;
                        nav += run ? 1 : 0;
                    }
                }
                if (nav > 0 || !(ok = ll_spin(spins, abort_flag)))
                    break;
            }
#pragma unroll
            for (int u = 0; u < CHAIN_WIN; u++)
                if (u < nav)
#pragma unroll
                    for (int q = 0; q < CHAIN_MAXR; q++)
                        sx[par][u][lane][q] = q < nr ? c64{__uint_as_float((unsigned)v[u][2 * q]),
                                                           __uint_as_float((unsigned)v[u][2 * q + 1])}
                                                     : c64{0.0f, 0.0f};
            if (lane == 0) {
                s_nav[par] = nav;
                if (!ok)
                    s_ok = 0;
            }
        }
        __syncthreads();
        const int nav = s_nav[par];
#pragma unroll
        for (int u = 0; u < CHAIN_WIN; u++)
            if (u < nav)
#pragma unroll
                for (int c = 0; c < 16; c++)
#pragma unroll
                    for (int q = 0; q < CHAIN_MAXR; q++) {
                        const c64 x = sx[par][u][16 * wv + c][q];
                        acc[q].re += lv[u][c].re * x.re - lv[u][c].im * x.im;
                        acc[q].im += lv[u][c].re * x.im + lv[u][c].im * x.re;
                    }
        t += nav;
        par ^= 1;
    }
    __syncthreads(); // every wave done with sx; s_ok final
#pragma unroll
    for (int q = 0; q < CHAIN_MAXR; q++)
        sp[wv][lane][q] = acc[q];
    __syncthreads();
    if (wv < nr) {
        const int q = wv;
        c64 b = c64{0.0f, 0.0f};
        if (lane < wj) {
            b = bq;
#pragma unroll
            for (int w = 0; w < 4; w++) {
                b.re -= sp[w][lane][q].re;
                b.im -= sp[w][lane][q].im;
            }
        }
        c64 rd{1.0f, 0.0f};
        if (UPPER && lane < wj)
            rd = cdiv(c64{1.0f, 0.0f}, sd[lane][lane]);
        c64 x = wj == 64 ? chunk_substitute<UPPER, true>(b, rd, wj, sd, lane)
                         : chunk_substitute<UPPER, false>(b, rd, wj, sd, lane);
        if (!s_ok)
            x = c64{__builtin_nanf(""), __builtin_nanf("")};
        if (lane < wj)
            B[(r0 + lane) + q * ldb] = x;
        if (s_ok) {
            unsigned long long* dst = xw + ((size_t)j * 64 + lane) * (size_t)nr * 2 + 2 * q;
            ll_put(dst, seq, __float_as_uint(lane < wj ? x.re : 0.0f));
            ll_put(dst + 1, seq, __float_as_uint(lane < wj ? x.im : 0.0f));
        }
    }
    __syncthreads(); // LDS (and s_item) reused by the next item
    }
}

// FRICO_SOLVE_CHAIN=0 keeps the blocked solves for a few right-hand sides too.
static bool chain_on()
{
    static const bool on = [] {
        const char* e = std::getenv("FRICO_SOLVE_CHAIN");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

static bool use_chain(const ChainSolve* cs, int64_t n, int64_t nrhs)
{
    return cs && cs->words && chain_on() && nrhs >= 1 && nrhs <= CHAIN_MAXR && n > 64 &&
           cs->nwords >= lu_chain_words(n, nrhs);
}

// Both halves (forward = true) or the backward half of a solve in place by the chained kernels; with
// a residual stream, X += D and the residual product follow as one block on the second stream.
// Both halves (forward) or the backward half of nz systems' solves by the chained kernels, system z
// at A + z sa, B + z sb, its exchange area at words + z lu_chain_words(n, nrhs).
int lu_chain_solve_batched(const c64* A, int64_t sa, int64_t n, int64_t lda, c64* B, int64_t sb, int64_t ldb,
                           int64_t nrhs, int nz, unsigned long long* words, bool forward, hipStream_t st)
{
    const int64_t nch = (n + 63) / 64;
    if (nch > INT32_MAX || nz < 1 || nz > 65535)
        return set_error("chained solve: %lld chunks, %d systems", (long long)nch, nz), -1;
    if (nrhs < 1 || nrhs > CHAIN_MAXR)
        return set_error("chained solve: %lld right-hand sides", (long long)nrhs), -1;
    const int64_t nw = (int64_t)lu_chain_words(n, nrhs);
    int rc = launch_zero_bytes(words, (size_t)(nw * nz) * sizeof(unsigned long long), st);
    if (rc)
        return rc;
    Batch bt;
    bt.s0 = sa;
    bt.s1 = sb;
    bt.s2 = nw;
    // the items are taken from a counter: any grid size is safe; one workgroup an item up to about
    // one a CU
    const int64_t items = nch * nz;
    const dim3 grid((unsigned)(items < 240 ? items : 240));
    if (forward)
        launch_kernel(&chain_solve<false>, grid, dim3(256), 0, st, A, lda, n, B, ldb, (int)nrhs, words, 1u, bt, nz);
    launch_kernel(&chain_solve<true>, grid, dim3(256), 0, st, A, lda, n, B, ldb, (int)nrhs, words, 2u, bt, nz);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

static int lu_chain_solve(const c64* A, int64_t n, int64_t lda, c64* B, int64_t ldb, int64_t nrhs, hipStream_t st,
                          const ChainSolve* cs, bool forward, const ResidStream* rs)
{
    int rc = lu_chain_solve_batched(A, 0, n, lda, B, 0, ldb, nrhs, 1, cs->words, forward, st);
    if (rc)
        return rc;
    if (!rs)
        return 0;
    if (rs->nev < 2)
        return set_error("chained solve: the residual stream needs two events"), -1;
    FRICO_HIP_CHECK(hipEventRecord(rs->ev[0], st));
    FRICO_HIP_CHECK(hipStreamWaitEvent(rs->side, rs->ev[0], 0));
    if ((rc = launch_add_correction(rs->S, rs->lds, B, ldb, n, nrhs, rs->side)) ||
        (rc = gemm_sub<double>(n, nrhs, n, rs->Z, rs->ldz, rs->S, rs->lds, rs->R, rs->ldr, rs->side, nullptr, rs->exact)))
        return rc;
    FRICO_HIP_CHECK(hipEventRecord(rs->ev[1], rs->side));
    FRICO_HIP_CHECK(hipStreamWaitEvent(st, rs->ev[1], 0));
    return 0;
}

template <typename T>
static int lu_backward(const cx<T>* A, int64_t n, int64_t lda, cx<T>* Y, int64_t ldy, cx<T>* X, int64_t ldx,
                       int64_t nrhs, hipStream_t st, const ResidStream* rs)
{
    const int64_t nch = (n + 63) / 64;
    const int64_t nblk = (n + RESID_ROWS - 1) / RESID_ROWS;
    if (rs && (!std::is_same<T, float>::value || rs->nev < nblk + 1))
        return set_error("solve: residual stream needs complex64 factors and %lld events", (long long)(nblk + 1)), -1;
    if constexpr (std::is_same<T, float>::value) {
        if (solve_blocked()) {
            int rc = lu_backward_blocked(A, n, lda, Y, ldy, X, ldx, nrhs, st, rs);
            if (rc)
                return rc;
            if (rs) {
                FRICO_HIP_CHECK(hipEventRecord(rs->ev[nblk], rs->side));
                FRICO_HIP_CHECK(hipStreamWaitEvent(st, rs->ev[nblk], 0));
            }
            FRICO_HIP_CHECK(hipGetLastError());
            return 0;
        }
    }
    for (int64_t ch = nch - 1; ch >= 0; ch--) {
        const int64_t kc = ch * 64, ws = (n - kc) < 64 ? (n - kc) : 64;
        launch_trsm<T, true>((int)ws, A + kc + kc * lda, lda, Y + kc, ldy, X + kc, ldx, nrhs, st);
        if (kc > 0) {
            int rc = gemm_sub<T>(kc, nrhs, ws, A + kc * lda, lda, X + kc, ldx, Y, ldy, st, nullptr);
            if (rc)
                return rc;
        }
        if constexpr (std::is_same<T, float>::value) {
            if (rs && kc % RESID_ROWS == 0) {
                const int64_t kb = kc, bw = (n - kb) < RESID_ROWS ? (n - kb) : RESID_ROWS, b = kb / RESID_ROWS;
                FRICO_HIP_CHECK(hipEventRecord(rs->ev[b], st));
                FRICO_HIP_CHECK(hipStreamWaitEvent(rs->side, rs->ev[b], 0));
                int rc = launch_add_correction(rs->S + kb, rs->lds, X + kb, ldx, bw, nrhs, rs->side);
                if (!rc)
                    rc = gemm_sub<double>(n, nrhs, bw, rs->Z + kb * rs->ldz, rs->ldz, rs->S + kb, rs->lds, rs->R, rs->ldr,
                                          rs->side, nullptr, rs->exact);
                if (rc)
                    return rc;
            }
        }
    }
    if (rs) {
        // the caller's stream continues after the last residual product
        FRICO_HIP_CHECK(hipEventRecord(rs->ev[nblk], rs->side));
        FRICO_HIP_CHECK(hipStreamWaitEvent(st, rs->ev[nblk], 0));
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// Forward half of the fused solve: per 64-row chunk j (k = 64 j, width w), after the panel's
// interchanges, y_j = Linv_j b_j into Y and b_below -= W(:, j) b_j; one launch per chunk plus
// one laswp per panel.
template <typename T>
static int lu_forward_fused(const cx<T>* A, int64_t n, int64_t lda, const int* perm, int nb, cx<T>* B, int64_t ldb,
                            int64_t nrhs, hipStream_t st, const cx<T>* inv, const cx<T>* WV, cx<T>* Y)
{
    const int64_t npan = (n + nb - 1) / nb;
    const unsigned gy = (unsigned)((nrhs + GBN - 1) / GBN);
    for (int64_t p = 0; p < npan; p++) {
        const int64_t k = p * nb;
        const int w = (int)((n - k) < nb ? (n - k) : nb);
        launch_laswp<T>(B, ldb, 0, nrhs, perm ? perm + p * PERM_STRIDE : nullptr, st);
        for (int s = 0; s < w; s += 64) {
            const int64_t kc = k + s, ws = (n - kc) < 64 ? (n - kc) : 64, ch = kc / 64;
            const unsigned gx = (unsigned)((n - kc + GBM - 1) / GBM);
            launch_kernel(&solve_chunk<T>, dim3(gx, gy), dim3(256), 0, st, kc, n - kc, nrhs, ws, (int64_t)0,
                          inv + (size_t)(2 * ch) * 4096, WV + kc * n, n, B + kc, ldb, B, ldb, Y, n);
        }
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// Fused form of the solve: the forward half above, then the backward half from Y into B.
template <typename T>
static int lu_solve_fused(const cx<T>* A, int64_t n, int64_t lda, const int* perm, int nb, cx<T>* B, int64_t ldb,
                          int64_t nrhs, hipStream_t st, const cx<T>* inv, const cx<T>* WV, cx<T>* Y,
                          const ResidStream* rs = nullptr, const ChainSolve* cs = nullptr)
{
    int rc = lu_forward_fused<T>(A, n, lda, perm, nb, B, ldb, nrhs, st, inv, WV, Y);
    if (rc)
        return rc;
    if constexpr (std::is_same<T, float>::value) {
        // a few right-hand sides: the backward half by the chained kernel, in place in B
        if (use_chain(cs, n, nrhs)) {
            for (int64_t q = 0; q < nrhs; q++)
                if ((rc = launch_copy_bytes(Y + q * n, B + q * ldb, (size_t)n * sizeof(cx<T>), st)))
                    return rc;
            return lu_chain_solve(A, n, lda, B, ldb, nrhs, st, cs, false, rs);
        }
    }
    return lu_backward<T>(A, n, lda, Y, n, B, ldb, nrhs, st, rs);
}

// Backward half of the solve in place (see lu_backward).
template <typename T>
int lu_solve_upper(const cx<T>* A, int64_t n, int64_t lda, cx<T>* B, int64_t ldb, int64_t nrhs, hipStream_t st,
                   const ResidStream* rs, const ChainSolve* cs)
{
    if constexpr (std::is_same<T, float>::value)
        if (use_chain(cs, n, nrhs))
            return lu_chain_solve(A, n, lda, B, ldb, nrhs, st, cs, false, rs);
    return lu_backward<T>(A, n, lda, B, ldb, B, ldb, nrhs, st, rs);
}

template <typename T>
int lu_solve(const cx<T>* A, int64_t n, int64_t lda, const int* perm, int nb, cx<T>* B, int64_t ldb, int64_t nrhs,
             hipStream_t st, const void* dinv, const void* wv, void* ybuf, const ResidStream* rs, const ChainSolve* cs)
{
    const int64_t npan = (n + nb - 1) / nb;
    const cx<T>* inv = (const cx<T>*)dinv;
    if constexpr (std::is_same<T, float>::value) {
        // no interchanges: a few right-hand sides by the chained kernels, else both halves blocked
        if (!perm && use_chain(cs, n, nrhs))
            return lu_chain_solve(A, n, lda, B, ldb, nrhs, st, cs, true, rs);
        if (!perm && solve_blocked()) {
            int rc = lu_forward_blocked(A, n, lda, B, ldb, nrhs, st, rs);
            return rc ? rc : lu_backward<T>(A, n, lda, B, ldb, B, ldb, nrhs, st, rs);
        }
    }
    if (inv && wv && ybuf && nb % 64 == 0 && n > 64)
        return lu_solve_fused<T>(A, n, lda, perm, nb, B, ldb, nrhs, st, inv, (const cx<T>*)wv, (cx<T>*)ybuf, rs, cs);
    if (rs)
        return set_error("solve: the residual stream needs the fused solve"), -1;
    const unsigned gcols = (unsigned)((nrhs + 63) / 64);
    for (int64_t p = 0; p < npan; p++) {
        const int64_t k = p * nb;
        const int w = (int)((n - k) < nb ? (n - k) : nb);
        launch_laswp<T>(B, ldb, 0, nrhs, perm ? perm + p * PERM_STRIDE : nullptr, st);
        for (int s = 0; s < w; s += 64) {
            const int ws = (w - s) < 64 ? (w - s) : 64;
            // the diagonal sub-blocks of a chunk inverse are the inverses of the chunk's diagonal sub-blocks
            const int64_t ch = (k + s) / 64, off = (k + s) % 64;
            if (inv)
                launch_kernel(&trmm_block<T>, dim3(gcols), dim3(256), 0, st, ws,
                                   inv + (size_t)(2 * ch) * 4096 + off + off * 64, B + (k + s), ldb, nrhs);
            else
                launch_trsm<T, false>(ws, A + (k + s) + (k + s) * lda, lda, B + (k + s), ldb, B + (k + s), ldb, nrhs, st);
            const int64_t below = n - (k + s + ws);
            if (below > 0) {
                int rc = gemm_sub<T>(below, nrhs, ws, A + (k + s + ws) + (k + s) * lda, lda, B + (k + s), ldb,
                                     B + (k + s + ws), ldb, st, nullptr);
                if (rc)
                    return rc;
            }
        }
    }
    for (int64_t p = npan - 1; p >= 0; p--) {
        const int64_t k = p * nb;
        const int w = (int)((n - k) < nb ? (n - k) : nb);
        for (int s = ((w - 1) / 64) * 64; s >= 0; s -= 64) {
            const int ws = (w - s) < 64 ? (w - s) : 64;
            const int64_t ch = (k + s) / 64, off = (k + s) % 64;
            if (inv)
                launch_kernel(&trmm_block<T>, dim3(gcols), dim3(256), 0, st, ws,
                                   inv + (size_t)(2 * ch + 1) * 4096 + off + off * 64, B + (k + s), ldb, nrhs);
            else
                launch_trsm<T, true>(ws, A + (k + s) + (k + s) * lda, lda, B + (k + s), ldb, B + (k + s), ldb, nrhs, st);
            const int64_t above = k + s;
            if (above > 0) {
                int rc = gemm_sub<T>(above, nrhs, ws, A + (k + s) * lda, lda, B + (k + s), ldb, B, ldb, st, nullptr);
                if (rc)
                    return rc;
            }
        }
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

// X = Z^{-T} B with the factors of lu_factor (deferred left interchanges):
//   Z^{-1} = U^{-1} E_last S_last ... E_1 S_1,  E_p: [B_p <- L11^{-1} B_p ; B_below -= L21 B_p]
//   Z^{-T} = S_1^T E_1^T ... S_last^T E_last^T U^{-T}
//   U^T y = b forward (lower, non-unit), then per panel p descending:
//   B_p <- L11^{-T} (B_p - L21^T B_below), then undo panel p's interchanges in reverse order.
template <typename T>
int lu_solve_transposed(const cx<T>* A, int64_t n, int64_t lda, const int* ipiv, int nb, cx<T>* B, int64_t ldb,
                        int64_t nrhs, hipStream_t st)
{
    const int64_t npan = (n + nb - 1) / nb;
    // U^T y = b: block rows ascending; y_p = U_pp^{-T} (b_p - U(0:k, p)^T y(0:k))
    for (int64_t p = 0; p < npan; p++) {
        const int64_t k = p * nb;
        const int w = (int)((n - k) < nb ? (n - k) : nb);
        for (int s = 0; s < w; s += 64) {
            const int ws = (w - s) < 64 ? (w - s) : 64;
            const int64_t above = k + s;
            if (above > 0) {
                using TL = typename TileBig<T>::type;
                const int64_t gx = (ws + TL::BM - 1) / TL::BM, gy = (nrhs + TL::BN - 1) / TL::BN;
                launch_kernel(&cgemm_sub<T, T, true>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, (int64_t)ws,
                                   nrhs, above, A + (k + s) * lda, lda, B, ldb, B + (k + s), ldb);
            }
            launch_trsm<T, false, true, false>(ws, A + (k + s) + (k + s) * lda, lda, B + (k + s), ldb, B + (k + s), ldb, nrhs, st);
        }
    }
    for (int64_t p = npan - 1; p >= 0; p--) {
        const int64_t k = p * nb;
        const int w = (int)((n - k) < nb ? (n - k) : nb);
        for (int s = ((w - 1) / 64) * 64; s >= 0; s -= 64) {
            const int ws = (w - s) < 64 ? (w - s) : 64;
            const int64_t below = n - (k + s + ws);
            if (below > 0) {
                using TL = typename TileBig<T>::type;
                const int64_t gx = (ws + TL::BM - 1) / TL::BM, gy = (nrhs + TL::BN - 1) / TL::BN;
                launch_kernel(&cgemm_sub<T, T, true>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, (int64_t)ws,
                                   nrhs, below, A + (k + s + ws) + (k + s) * lda, lda, B + (k + s + ws), ldb, B + (k + s),
                                   ldb);
            }
            launch_trsm<T, true, true, true>(ws, A + (k + s) + (k + s) * lda, lda, B + (k + s), ldb, B + (k + s), ldb, nrhs, st);
        }
        launch_kernel(&laswp_reverse<T>, dim3((unsigned)((nrhs + 255) / 256)), dim3(256), 0, st, B, ldb, (int64_t)0,
                           nrhs, k, w, ipiv);
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_convert_c128_c64(const c128* X, int64_t ldx, c64* Y, int64_t ldy, int64_t rows, int64_t cols, hipStream_t st)
{
    for (int64_t c0 = 0; c0 < cols; c0 += 65535) {
        const int64_t nc = (cols - c0) < 65535 ? (cols - c0) : 65535;
        launch_kernel(&convert_mat<double, float>, dim3((unsigned)((rows + 255) / 256), (unsigned)nc), dim3(256), 0, st,
                           X + c0 * ldx, ldx, Y + c0 * ldy, ldy, rows);
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_add_correction(c128* X, int64_t ldx, const c64* D, int64_t ldd, int64_t rows, int64_t cols, hipStream_t st)
{
    launch_kernel(&add_correction, dim3((unsigned)((rows + 255) / 256), (unsigned)cols), dim3(256), 0, st, X, ldx, D,
                       ldd, rows);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_col_norm2(const c128* X, int64_t ldx, int64_t rows, int64_t cols, double* out, hipStream_t st)
{
    launch_kernel(&col_norm2, dim3((unsigned)cols), dim3(256), 0, st, X, ldx, rows, out);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

template int gemm_sub<double>(int64_t, int64_t, int64_t, const c128*, int64_t, const c128*, int64_t, c128*, int64_t,
                              hipStream_t, KernelTimer*, bool);
template int gemm_sub<float>(int64_t, int64_t, int64_t, const c64*, int64_t, const c64*, int64_t, c64*, int64_t,
                             hipStream_t, KernelTimer*, bool);
template int lu_factor<double>(LuWork&, c128*, int64_t, int64_t, int*, int, hipStream_t, KernelTimer*, int64_t);
template int lu_factor<float>(LuWork&, c64*, int64_t, int64_t, int*, int, hipStream_t, KernelTimer*, int64_t);
template int ldlt_factor<double>(LuWork&, c128*, int64_t, int64_t, int*, int, hipStream_t, KernelTimer*, int64_t);
template int ldlt_factor<float>(LuWork&, c64*, int64_t, int64_t, int*, int, hipStream_t, KernelTimer*, int64_t);
template int lu_nopiv_factor<float>(LuWork&, c64*, int64_t, int64_t, int*, int, hipStream_t, KernelTimer*, int64_t);
template int lu_solve_upper<double>(const c128*, int64_t, int64_t, c128*, int64_t, int64_t, hipStream_t,
                                    const ResidStream*, const ChainSolve*);
template int lu_solve_upper<float>(const c64*, int64_t, int64_t, c64*, int64_t, int64_t, hipStream_t, const ResidStream*,
                                   const ChainSolve*);
template int lu_solve<double>(const c128*, int64_t, int64_t, const int*, int, c128*, int64_t, int64_t, hipStream_t,
                              const void*, const void*, void*, const ResidStream*, const ChainSolve*);
template int lu_solve<float>(const c64*, int64_t, int64_t, const int*, int, c64*, int64_t, int64_t, hipStream_t,
                             const void*, const void*, void*, const ResidStream*, const ChainSolve*);
template int lu_solve_transposed<double>(const c128*, int64_t, int64_t, const int*, int, c128*, int64_t, int64_t,
                                         hipStream_t);
template int lu_solve_transposed<float>(const c64*, int64_t, int64_t, const int*, int, c64*, int64_t, int64_t,
                                        hipStream_t);
int lu_perm_stride() { return PERM_STRIDE; }
size_t lu_exchange_words(int gmax) { return ll_words(gmax); }
template int panel_rows<double>(int, bool);
template int panel_rows<float>(int, bool);

} // namespace frico
