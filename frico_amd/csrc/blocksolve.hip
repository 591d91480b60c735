// Block-column and block-row solves of the unpivoted complex64 LU (lu_nopiv_factor), one lane per
// row (per column), and the split of L21 for the fp16 trailing products straight from A.
//
// After the diagonal block A11 = L11 U11 of a step (w <= 128 columns from k) is factored:
//   L21 = A21 U11^{-1}   (the m rows below the block):       solve_l21, lane = row of A21
//   U12 = L11^{-1} A12   (the block row, columns c0 ..):    solve_u12, lane = column of A12
// Each 64-lane workgroup keeps its 64 rows (columns) in LDS as s[index][lane], so that every lane
// reads its own values without bank conflicts, and computes four outputs at a time: the inner sum
// over the finished outputs takes one LDS read of x and four broadcast reads of the triangular
// factor, whose four needed columns (rows) are staged in LDS per step of four.  The arithmetic is
// forward substitution in the textbook order, in complex64, with U11's diagonal applied through
// Smith's reciprocal.
#include "common.hpp"
#include "lu.hpp"

namespace frico {

constexpr int BSW = 128;     // largest block width
constexpr int BSP = 65;      // LDS row pitch (complex values) of s: 64 lanes + 1

__device__ inline c64 cfms(c64 acc, c64 a, c64 b) // acc - a b
{
    acc.re -= a.re * b.re - a.im * b.im;
    acc.im -= a.re * b.im + a.im * b.re;
    return acc;
}

// L21 = A21 U11^{-1} in place.  Grid: ceil(m / 64) workgroups of 64 lanes.
__global__ __launch_bounds__(64) void solve_l21(c64* __restrict__ A, int64_t lda, int64_t k, int w, int64_t m)
{
    __shared__ c64 s[BSW][BSP];
    __shared__ c64 u[BSW][4]; // u[kk][j] = U11(kk, c + j)
    const int lane = threadIdx.x;
    const int64_t r = (int64_t)blockIdx.x * 64 + lane;
    const bool ok = r < m;
    c64* A21 = A + (k + w) + k * lda;
    const c64* U = A + k + k * lda;
    for (int c = 0; c < w; c++)
        s[c][lane] = ok ? A21[r + c * lda] : c64{0.0f, 0.0f};
    for (int c = 0; c < w; c += 4) {
        const int nq = (w - c) < 4 ? (w - c) : 4;
        __syncthreads(); // the previous step's u has been read
        for (int j = 0; j < nq; j++)
            for (int kk = lane; kk < c + nq; kk += 64)
                u[kk][j] = U[kk + (int64_t)(c + j) * lda];
        __syncthreads();
        c64 acc[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[j] = j < nq ? s[c + j][lane] : c64{0.0f, 0.0f};
        for (int kk = 0; kk < c; kk++) {
            const c64 x = s[kk][lane];
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[j] = cfms(acc[j], x, u[kk][j]);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j < nq) {
                acc[j] = acc[j] * cdiv(c64{1.0f, 0.0f}, u[c + j][j]);
#pragma unroll
                for (int jj = j + 1; jj < 4; jj++)
                    if (jj < nq)
                        acc[jj] = cfms(acc[jj], acc[j], u[c + j][jj]);
                s[c + j][lane] = acc[j];
            }
        }
    }
    if (ok)
        for (int c = 0; c < w; c++)
            A21[r + c * lda] = s[c][lane];
}

// U12 = L11^{-1} A12 in place (L11 unit lower).  Grid: ceil(nc / 64) workgroups of 64 lanes.
__global__ __launch_bounds__(64) void solve_u12(c64* __restrict__ A, int64_t lda, int64_t k, int w, int64_t c0,
                                                int64_t nc)
{
    __shared__ c64 s[BSW][BSP];
    __shared__ c64 l[4][BSW]; // l[j][kk] = L11(r + j, kk)
    const int lane = threadIdx.x;
    const int64_t cb = (int64_t)blockIdx.x * 64;
    const int ncol = (nc - cb) < 64 ? (int)(nc - cb) : 64;
    c64* A12 = A + k + (c0 + cb) * lda;
    const c64* L = A + k + k * lda;
    // columns of this workgroup into LDS, each column's rows read together
    for (int cc = 0; cc < ncol; cc++)
        for (int r = lane; r < w; r += 64)
            s[r][cc] = A12[r + (int64_t)cc * lda];
    for (int r = 0; r < w; r += 4) {
        const int nq = (w - r) < 4 ? (w - r) : 4;
        __syncthreads(); // the loads above, and the previous step's l, are done
        // rows r .. r + nq - 1 of L11, columns 0 .. r + nq - 1: four consecutive values per column
        for (int kk = lane; kk < r + nq; kk += 64)
            for (int j = 0; j < nq; j++)
                l[j][kk] = L[(r + j) + (int64_t)kk * lda];
        __syncthreads();
        c64 acc[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[j] = j < nq ? s[r + j][lane] : c64{0.0f, 0.0f};
        for (int kk = 0; kk < r; kk++) {
            const c64 x = s[kk][lane];
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[j] = cfms(acc[j], l[j][kk], x);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j < nq) {
#pragma unroll
                for (int jj = j + 1; jj < 4; jj++)
                    if (jj < nq)
                        acc[jj] = cfms(acc[jj], l[jj][r + j], acc[j]);
                s[r + j][lane] = acc[j];
            }
        }
    }
    __syncthreads();
    for (int cc = 0; cc < ncol; cc++)
        for (int r = lane; r < w; r += 64)
            A12[r + (int64_t)cc * lda] = s[r][cc];
}

// The fp16 split of S[r + k ld] (r < R, k < K <= 128; rows contiguous, as L21 lies in A) into the
// planes P[p][r][k] and exponents of split_f16 (gemm_f16.hip).  1024 threads: 64 rows x 16 groups
// of eight k; the row's largest magnitude is reduced over its 16 groups through LDS.
__global__ __launch_bounds__(1024) void split_f16_cm(const c64* __restrict__ S, int64_t ld, int64_t R, int K,
                                                     _Float16* __restrict__ P, int64_t cap, int* __restrict__ E)
{
    __shared__ float red[16][64];
    const int rl = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 64 + rl;
    float vr[8], vi[8];
    float mx = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int kk = 8 * g + q;
        c64 v{0.0f, 0.0f};
        if (r < R && kk < K)
            v = S[r + kk * ld];
        vr[q] = v.re;
        vi[q] = v.im;
        mx = fmaxf(mx, fmaxf(fabsf(v.re), fabsf(v.im)));
    }
    red[g][rl] = mx;
    __syncthreads();
    mx = 0.0f;
#pragma unroll
    for (int x = 0; x < 16; x++)
        mx = fmaxf(mx, red[x][rl]);
    int e = 0;
    if (mx > 0.0f) {
        e = 14 - ilogbf(mx); // mx 2^e in [2^14, 2^15)
        e = e < -100 ? -100 : (e > 200 ? 200 : e);
    }
    typedef _Float16 half8 __attribute__((ext_vector_type(8)));
    half8 rh, rlo, ih, il;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float xr = ldexpf(vr[q], e), xi = ldexpf(vi[q], e);
        const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
        rh[q] = hr;
        rlo[q] = (_Float16)(xr - (float)hr);
        ih[q] = hi;
        il[q] = (_Float16)(xi - (float)hi);
    }
    const int64_t off = r * HSPLIT_K + 8 * g, ps = cap * HSPLIT_K;
    *reinterpret_cast<half8*>(P + off) = rh;
    *reinterpret_cast<half8*>(P + ps + off) = rlo;
    *reinterpret_cast<half8*>(P + 2 * ps + off) = ih;
    *reinterpret_cast<half8*>(P + 3 * ps + off) = il;
    if (g == 0)
        E[r] = e;
}

int launch_solve_l21(c64* A, int64_t lda, int64_t k, int w, int64_t m, hipStream_t st)
{
    if (m <= 0)
        return 0;
    if (w < 1 || w > BSW)
        return set_error("block solve: width %d outside 1..%d", w, BSW), -1;
    launch_kernel(&solve_l21, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, st, A, lda, k, w, m);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_solve_u12(c64* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, hipStream_t st)
{
    if (nc <= 0)
        return 0;
    if (w < 1 || w > BSW)
        return set_error("block solve: width %d outside 1..%d", w, BSW), -1;
    launch_kernel(&solve_u12, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, st, A, lda, k, w, c0, nc);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_split_f16_cm(const c64* S, int64_t ld, int64_t R, int K, _Float16* planes, int* e, int64_t cap,
                        hipStream_t st)
{
    if (R <= 0)
        return 0;
    if (K < 1 || K > HSPLIT_K)
        return set_error("fp16 split: K = %d outside 1..%d", K, HSPLIT_K), -1;
    const int64_t Rp = (R + 127) / 128 * 128;
    if (Rp > cap)
        return set_error("fp16 split: %lld rows exceed the buffer's %lld", (long long)Rp, (long long)cap), -1;
    launch_kernel(&split_f16_cm, dim3((unsigned)(Rp / 64)), dim3(1024), 0, st, S, ld, R, K, planes, cap, e);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

} // namespace frico
