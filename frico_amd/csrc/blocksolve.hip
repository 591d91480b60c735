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

#include <cstdlib>

namespace frico {

constexpr int BSW = 128;     // largest block width
constexpr int BSP = 65;      // LDS row pitch (complex values) of s: 64 lanes + 1

__device__ inline c64 lane_bcast(c64 v, int l) // v of lane l (l wave-uniform)
{
    return {__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.re), l)),
            __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v.im), l))};
}

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
// planes (k slabs, hsplit_at) and exponents of split_f16 (gemm_f16.hip).  1024 threads: 64 rows x 16 groups
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
    const int64_t off = hsplit_at(r, 8 * g, cap), ps = cap * HSPLIT_K;
    *reinterpret_cast<half8*>(P + off) = rh;
    *reinterpret_cast<half8*>(P + ps + off) = rlo;
    *reinterpret_cast<half8*>(P + 2 * ps + off) = ih;
    *reinterpret_cast<half8*>(P + 3 * ps + off) = il;
    if (g == 0)
        E[r] = e;
}

// ----------------------------------------------------------------------------------------------
// tri_solve_split: the block solves of a step of the unpivoted LU in one launch each, with the
// fp16 split of their result (the operand format of cgemm_sub_f16) written beside it.
//   X = M^{-1} B, M the w x w (w <= 128) lower triangle of the step's diagonal block:
//   !TRANS: M = L11 (unit), B = the block row A(k .. k+w, c0 + j): U12 = L11^{-1} A12 in place;
//           split per column j of U12 (row j of the planes).
//   TRANS:  M = U11^T, B(:, j) = row c0 + j of A21 = A(k + w + c0 + j, k .. k+w), so that
//           X^T = A21 U11^{-1} = L21, written back to A21; split per row of L21.
// Workgroup: 4 NC columns of B, NC a wavefront, rows lane and lane + 64 of each column in a lane.
// The coefficients never touch LDS: each lane holds its own rows of M for 16 columns at a time
// (the next 16 loading while these are used), and a step c broadcasts the finished row c by
// readlane and takes M(r, c) x_c off the rows below.  For U11^T the rows are carried undivided
// (y_c = x_c M(c, c)), the coefficients are scaled by 1 / M(c, c) as they are used and the rows
// divided by their diagonal at the end.  A21 goes through LDS both ways so that its reads and
// writes run down its columns.  Planes and exponents: rows [0, roundup(nc, 128)) of the split
// buffer; columns past nc give zero rows (the product reads whole 128-row tiles).
// ----------------------------------------------------------------------------------------------
constexpr int TS_STRIP = 16; // columns of M a lane holds at a time

template <bool TRANS, int NC>
__global__ __launch_bounds__(256) void tri_solve_split(c64* __restrict__ A, int64_t lda, int64_t k, int w, int64_t c0,
                                                       int64_t nc, _Float16* __restrict__ P, int* __restrict__ E,
                                                       int64_t cap, Batch bt)
{
    // batched: A += z s0, P += z s1, E += z s2
    A += blockIdx.z * bt.s0;
    P += blockIdx.z * bt.s1;
    E += blockIdx.z * bt.s2;
    constexpr int COLS = 4 * NC;
    __shared__ c64 sT[TRANS ? COLS : 1][BSW + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t j0 = (int64_t)blockIdx.x * COLS;
    const c64* D = A + k + k * lda;
    // M(r, c) for r > c (zero elsewhere and past w): L11(r, c) = D(r, c), or U11^T(r, c) = D(c, r)
    auto mval = [&](int r, int c) -> c64 {
        if (!(r > c && r < w))
            return c64{0.0f, 0.0f};
        return TRANS ? D[c + (int64_t)r * lda] : D[r + (int64_t)c * lda];
    };
    // the strip of columns [16 s, 16 s + 16): rows lane (needed while the columns are < 64) and lane + 64
    auto strip = [&](int sidx, c64* ml, c64* mh) {
#pragma unroll
        for (int t = 0; t < TS_STRIP; t++) {
            const int c = TS_STRIP * sidx + t;
            if (TS_STRIP * sidx < 64)
                ml[t] = mval(lane, c);
            mh[t] = mval(lane + 64, c);
        }
    };
    c64 rl{1.0f, 0.0f}, rh{1.0f, 0.0f}; // TRANS: 1 / M(r, r) of rows lane, lane + 64
    if constexpr (TRANS) {
        if (lane < w)
            rl = cdiv(c64{1.0f, 0.0f}, D[lane + (int64_t)lane * lda]);
        if (lane + 64 < w)
            rh = cdiv(c64{1.0f, 0.0f}, D[(lane + 64) + (int64_t)(lane + 64) * lda]);
    }
    c64 lo[NC], hi[NC];
    if constexpr (TRANS) {
        c64 v[COLS / 2];
#pragma unroll
        for (int t = 0; t < COLS / 2; t++) {
            const int e = tid + 256 * t, jj = e % COLS, kk = e / COLS;
            const int64_t j = j0 + jj;
            v[t] = (j < nc && kk < w) ? A[(k + w + c0 + j) + (k + kk) * lda] : c64{0.0f, 0.0f};
        }
#pragma unroll
        for (int t = 0; t < COLS / 2; t++) {
            const int e = tid + 256 * t;
            sT[e % COLS][e / COLS] = v[t];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NC; q++) {
            lo[q] = sT[NC * wv + q][lane];
            hi[q] = sT[NC * wv + q][lane + 64];
        }
    } else {
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const int64_t j = j0 + NC * wv + q;
            const c64* col = A + k + (c0 + j) * lda;
            lo[q] = (j < nc && lane < w) ? col[lane] : c64{0.0f, 0.0f};
            hi[q] = (j < nc && lane + 64 < w) ? col[lane + 64] : c64{0.0f, 0.0f};
        }
    }
    c64 cl[TS_STRIP], ch[TS_STRIP], nl[TS_STRIP], nh[TS_STRIP];
#pragma unroll
    for (int t = 0; t < TS_STRIP; t++)
        cl[t] = nl[t] = c64{0.0f, 0.0f};
    strip(0, cl, ch);
#pragma unroll
    for (int sidx = 0; sidx < BSW / TS_STRIP; sidx++) {
        if (sidx + 1 < BSW / TS_STRIP)
            strip(sidx + 1, nl, nh);
#pragma unroll
        for (int t = 0; t < TS_STRIP; t++) {
            const int c = TS_STRIP * sidx + t;
            c64 ml = cl[t], mh = ch[t];
            if constexpr (TRANS) {
                const c64 rc = lane_bcast(c < 64 ? rl : rh, c & 63);
                ml = ml * rc;
                mh = mh * rc;
            }
#pragma unroll
            for (int q = 0; q < NC; q++) {
                if (c < 64) {
                    const c64 x = lane_bcast(lo[q], c);
                    lo[q] = cfms(lo[q], ml, x);
                    hi[q] = cfms(hi[q], mh, x);
                } else {
                    const c64 x = lane_bcast(hi[q], c - 64);
                    hi[q] = cfms(hi[q], mh, x);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < TS_STRIP; t++) {
            cl[t] = nl[t];
            ch[t] = nh[t];
        }
    }
    if constexpr (TRANS) {
#pragma unroll
        for (int q = 0; q < NC; q++) {
            lo[q] = lo[q] * rl;
            hi[q] = hi[q] * rh;
        }
    }
    // the split of each column (row of the planes)
    const int64_t ps = cap * HSPLIT_K;
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const int64_t j = j0 + NC * wv + q;
        float mx = fmaxf(fmaxf(fabsf(lo[q].re), fabsf(lo[q].im)), fmaxf(fabsf(hi[q].re), fabsf(hi[q].im)));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
            mx = fmaxf(mx, __shfl_xor(mx, o));
        int e = 0;
        if (mx > 0.0f) {
            e = 14 - ilogbf(mx); // mx 2^e in [2^14, 2^15)
            e = e < -100 ? -100 : (e > 200 ? 200 : e);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const c64 v = h ? hi[q] : lo[q];
            _Float16* pr = P + hsplit_at(j, lane + 64 * h, cap);
            const float xr = ldexpf(v.re, e), xi = ldexpf(v.im, e);
            const _Float16 hr = (_Float16)xr, hm = (_Float16)xi;
            pr[0] = hr;
            pr[ps] = (_Float16)(xr - (float)hr);
            pr[2 * ps] = hm;
            pr[3 * ps] = (_Float16)(xi - (float)hm);
        }
        if (lane == 0)
            E[j] = e;
        if constexpr (!TRANS) {
            c64* col = A + k + (c0 + j) * lda;
            if (j < nc && lane < w)
                col[lane] = lo[q];
            if (j < nc && lane + 64 < w)
                col[lane + 64] = hi[q];
        }
    }
    if constexpr (TRANS) {
        __syncthreads(); // every wavefront has read its columns of sT
#pragma unroll
        for (int q = 0; q < NC; q++) {
            sT[NC * wv + q][lane] = lo[q];
            sT[NC * wv + q][lane + 64] = hi[q];
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < COLS / 2; t++) {
            const int e = tid + 256 * t, jj = e % COLS, kk = e / COLS;
            const int64_t j = j0 + jj;
            if (j < nc && kk < w)
                A[(k + w + c0 + j) + (k + kk) * lda] = sT[jj][kk];
        }
    }
}

// The same solve with M staged in LDS once (for launches on the critical path, where the strips'
// load latency would set the pace): the strictly lower triangle, packed by columns (column c holds
// rows c+1 .. 127 from tri_off(c)), read with every load of a thread in flight at once, in an
// order that runs down D's columns (folded so that the 8128 entries split evenly); for U11^T the
// columns come scaled by 1 / M(c, c).  The steps read M four steps ahead.
constexpr int TS_TRI = BSW * (BSW - 1) / 2; // 8128
__device__ inline int tri_off(int c) { return (BSW - 1) * c - c * (c - 1) / 2; }

template <bool TRANS, int NC>
__global__ __launch_bounds__(256) void tri_solve_split_lds(c64* __restrict__ A, int64_t lda, int64_t k, int w,
                                                           int64_t c0, int64_t nc, _Float16* __restrict__ P,
                                                           int* __restrict__ E, int64_t cap, Batch bt)
{
    // batched: A += z s0, P += z s1, E += z s2
    A += blockIdx.z * bt.s0;
    P += blockIdx.z * bt.s1;
    E += blockIdx.z * bt.s2;
    constexpr int COLS = 4 * NC;
    constexpr int PER = (TS_TRI + 255) / 256; // 32
    constexpr int PF = 4;
    __shared__ c64 sM[TS_TRI];
    __shared__ c64 sR[TRANS ? BSW : 1];
    __shared__ c64 sT[TRANS ? COLS : 1][BSW + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t j0 = (int64_t)blockIdx.x * COLS;
    const c64* D = A + k + k * lda;
    // entry e of the enumeration -> (r, c) of M, r > c, and its place D(a, b) in memory
    auto entry = [](int e, int& r, int& c, int& a, int& b) {
        if constexpr (!TRANS) {
            // M = D's strictly lower part, column c = D's column c: fold columns f and 127 - f
            const int f = e / (BSW - 1), i = e % (BSW - 1);
            if (i < BSW - 1 - f) {
                c = f;
                r = f + 1 + i;
            } else {
                c = BSW - 1 - f;
                r = i + 1;
            }
            a = r;
            b = c;
        } else {
            // M(r, c) = D(c, r): D's strictly upper part, read down D's column b = r (rows a < b):
            // fold columns f + 1 and 127 - f (128 entries), then column 64 alone
            if (e < 63 * BSW) {
                const int f = e / BSW, i = e % BSW;
                if (i < f + 1) {
                    b = f + 1;
                    a = i;
                } else {
                    b = BSW - 1 - f;
                    a = i - (f + 1);
                }
            } else {
                b = 64;
                a = e - 63 * BSW;
            }
            r = b;
            c = a;
        }
    };
    if constexpr (TRANS) {
        if (tid < BSW)
            sR[tid] = tid < w ? cdiv(c64{1.0f, 0.0f}, D[tid + (int64_t)tid * lda]) : c64{1.0f, 0.0f};
    }
    c64 v[PER];
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = tid + 256 * t;
        int r = 0, c = 0, a = 0, b = 0;
        if (e < TS_TRI)
            entry(e, r, c, a, b);
        v[t] = (e < TS_TRI && r < w) ? D[a + (int64_t)b * lda] : c64{0.0f, 0.0f};
    }
    c64 bcol[TRANS ? COLS / 2 : 1];
    if constexpr (TRANS) {
#pragma unroll
        for (int t = 0; t < COLS / 2; t++) {
            const int e = tid + 256 * t, jj = e % COLS, kk = e / COLS;
            const int64_t j = j0 + jj;
            bcol[t] = (j < nc && kk < w) ? A[(k + w + c0 + j) + (k + kk) * lda] : c64{0.0f, 0.0f};
        }
    }
    __syncthreads(); // sR
#pragma unroll
    for (int t = 0; t < PER; t++) {
        const int e = tid + 256 * t;
        if (e < TS_TRI) {
            int r, c, a, b;
            entry(e, r, c, a, b);
            sM[tri_off(c) + (r - c - 1)] = TRANS ? v[t] * sR[c] : v[t];
        }
    }
    if constexpr (TRANS) {
#pragma unroll
        for (int t = 0; t < COLS / 2; t++) {
            const int e = tid + 256 * t;
            sT[e % COLS][e / COLS] = bcol[t];
        }
    }
    __syncthreads();
    const c64 rl = TRANS ? sR[lane] : c64{1.0f, 0.0f}, rh = TRANS ? sR[lane + 64] : c64{1.0f, 0.0f};
    c64 lo[NC], hi[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const int64_t j = j0 + NC * wv + q;
        if constexpr (TRANS) {
            lo[q] = sT[NC * wv + q][lane];
            hi[q] = sT[NC * wv + q][lane + 64];
        } else {
            const c64* col = A + k + (c0 + j) * lda;
            lo[q] = (j < nc && lane < w) ? col[lane] : c64{0.0f, 0.0f};
            hi[q] = (j < nc && lane + 64 < w) ? col[lane + 64] : c64{0.0f, 0.0f};
        }
    }
    // coefficient of step c for the rows lane (low) and lane + 64 (high)
    auto coef = [&](int c, c64& ml, c64& mh) {
        ml = (c < 64 && lane > c) ? sM[tri_off(c) + (lane - c - 1)] : c64{0.0f, 0.0f};
        mh = (lane + 64 > c) ? sM[tri_off(c) + (lane + 63 - c)] : c64{0.0f, 0.0f};
    };
    c64 rlo[PF], rhi[PF];
#pragma unroll
    for (int p = 0; p < PF; p++)
        coef(p, rlo[p], rhi[p]);
#pragma unroll
    for (int c = 0; c < BSW; c++) {
        const c64 ml = rlo[c % PF], mh = rhi[c % PF];
        if (c + PF < BSW)
            coef(c + PF, rlo[c % PF], rhi[c % PF]);
#pragma unroll
        for (int q = 0; q < NC; q++) {
            if (c < 64) {
                const c64 x = lane_bcast(lo[q], c);
                lo[q] = cfms(lo[q], ml, x);
                hi[q] = cfms(hi[q], mh, x);
            } else {
                const c64 x = lane_bcast(hi[q], c - 64);
                hi[q] = cfms(hi[q], mh, x);
            }
        }
    }
    if constexpr (TRANS) {
#pragma unroll
        for (int q = 0; q < NC; q++) {
            lo[q] = lo[q] * rl;
            hi[q] = hi[q] * rh;
        }
    }
    const int64_t ps = cap * HSPLIT_K;
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const int64_t j = j0 + NC * wv + q;
        float mx = fmaxf(fmaxf(fabsf(lo[q].re), fabsf(lo[q].im)), fmaxf(fabsf(hi[q].re), fabsf(hi[q].im)));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
            mx = fmaxf(mx, __shfl_xor(mx, o));
        int e = 0;
        if (mx > 0.0f) {
            e = 14 - ilogbf(mx); // mx 2^e in [2^14, 2^15)
            e = e < -100 ? -100 : (e > 200 ? 200 : e);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const c64 vv = h ? hi[q] : lo[q];
            _Float16* pr = P + hsplit_at(j, lane + 64 * h, cap);
            const float xr = ldexpf(vv.re, e), xi = ldexpf(vv.im, e);
            const _Float16 hr = (_Float16)xr, hm = (_Float16)xi;
            pr[0] = hr;
            pr[ps] = (_Float16)(xr - (float)hr);
            pr[2 * ps] = hm;
            pr[3 * ps] = (_Float16)(xi - (float)hm);
        }
        if (lane == 0)
            E[j] = e;
        if constexpr (!TRANS) {
            c64* col = A + k + (c0 + j) * lda;
            if (j < nc && lane < w)
                col[lane] = lo[q];
            if (j < nc && lane + 64 < w)
                col[lane + 64] = hi[q];
        }
    }
    if constexpr (TRANS) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NC; q++) {
            sT[NC * wv + q][lane] = lo[q];
            sT[NC * wv + q][lane + 64] = hi[q];
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < COLS / 2; t++) {
            const int e = tid + 256 * t, jj = e % COLS, kk = e / COLS;
            const int64_t j = j0 + jj;
            if (j < nc && kk < w)
                A[(k + w + c0 + j) + (k + kk) * lda] = sT[jj][kk];
        }
    }
}

template <bool TRANS, int NC>
static void launch_tss_lds(c64* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, int64_t Rp, _Float16* planes,
                           int* e, int64_t cap, hipStream_t st, const Batch& bt, int nz)
{
    launch_kernel(&tri_solve_split_lds<TRANS, NC>, dim3((unsigned)(Rp / (4 * NC)), 1, (unsigned)nz), dim3(256), 0, st, A,
                  lda, k, w, c0, nc, planes, e, cap, bt);
}

template <bool TRANS, int NC>
static void launch_tss(c64* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, int64_t Rp, _Float16* planes,
                       int* e, int64_t cap, hipStream_t st, const Batch& bt, int nz)
{
    launch_kernel(&tri_solve_split<TRANS, NC>, dim3((unsigned)(Rp / (4 * NC)), 1, (unsigned)nz), dim3(256), 0, st, A, lda,
                  k, w, c0, nc, planes, e, cap, bt);
}

// FRICO_TSS_LDS: 0 never stages M in LDS, 1 (default) as chosen below, 2 always
static bool tss_lds(bool trans, int64_t nc)
{
    static const int mode = [] {
        const char* e = std::getenv("FRICO_TSS_LDS");
        return e ? std::atoi(e) : 1;
    }();
    if (mode == 0 || mode == 2)
        return mode == 2 && !(trans && nc > 6144);
    return !trans || nc <= 4096;
}

int launch_tri_solve_split(c64* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, bool trans, _Float16* planes,
                           int* e, int64_t cap, hipStream_t st, const Batch& bt, int nz)
{
    if (nz < 1 || nz > 65535)
        return set_error("block solve: batch of %d systems", nz), -1;
    if (nc <= 0)
        return 0;
    if (w < 1 || w > BSW)
        return set_error("block solve: width %d outside 1..%d", w, BSW), -1;
    const int64_t Rp = (nc + 127) / 128 * 128;
    if (Rp > cap)
        return set_error("fp16 split: %lld rows exceed the buffer's %lld", (long long)Rp, (long long)cap), -1;
    // M staged in LDS except for L21 with many rows, which runs beside the wide product and must fit
    // next to its workgroups; columns a wavefront: one where the launch is short and on the critical
    // path, four for wide blocks
    const int ncw = nc <= 2048 ? 1 : (nc <= 6144 ? 2 : 4);
    if (tss_lds(trans, nc)) {
        if (trans) {
            if (ncw == 1)
                launch_tss_lds<true, 1>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
            else
                launch_tss_lds<true, 2>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
        } else {
            if (ncw == 1)
                launch_tss_lds<false, 1>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
            else if (ncw == 2)
                launch_tss_lds<false, 2>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
            else
                launch_tss_lds<false, 4>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
        }
        FRICO_HIP_CHECK(hipGetLastError());
        return 0;
    }
    if (trans) {
        if (ncw == 1)
            launch_tss<true, 1>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
        else if (ncw == 2)
            launch_tss<true, 2>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
        else
            launch_tss<true, 4>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
    } else {
        if (ncw == 1)
            launch_tss<false, 1>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
        else if (ncw == 2)
            launch_tss<false, 2>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
        else
            launch_tss<false, 4>(A, lda, k, w, c0, nc, Rp, planes, e, cap, st, bt, nz);
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
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
