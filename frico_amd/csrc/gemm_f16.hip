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
// N columns), K <= 128 (kc = K rounded up to 32).  128 x 128 tile per 256-thread workgroup, 64 x 64
// per wave (4 x 4 blocks of 16 x 16), 32-deep k chunks through LDS.  The split planes are padded
// to whole tiles, so the loads need no bounds; only C is masked.
// ----------------------------------------------------------------------------------------------
constexpr int HB = 128;   // tile rows / columns
constexpr int HKC = 32;   // k chunk
// LDS row of one plane: 32 halves = four 16-byte segments, stored at segment s ^ ((row >> 1) & 3)
// so that eight consecutive rows read by one quarter-wave hit distinct banks.
__device__ inline int hseg(int row, int s) { return s ^ ((row >> 1) & 3); }

__global__ __launch_bounds__(256, 2) void cgemm_sub_f16(int64_t M, int64_t N, int kc, const _Float16* __restrict__ La,
                                                       const int* __restrict__ ea, int64_t acap,
                                                       const _Float16* __restrict__ Ub, const int* __restrict__ eb,
                                                       int64_t bcap, c64* __restrict__ C, int64_t ldc)
{
    __shared__ half8 sA[4][HB][4]; // [plane][row][segment]
    __shared__ half8 sB[4][HB][4]; // [plane][column][segment]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
    const int64_t m0 = (int64_t)blockIdx.x * HB, n0 = (int64_t)blockIdx.y * HB;
    const int64_t aps = acap * HSPLIT_K, bps = bcap * HSPLIT_K;

    f32x4 are[4][4], aim[4][4]; // [row block][column block]
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            are[a][b] = f32x4{0, 0, 0, 0};
            aim[a][b] = f32x4{0, 0, 0, 0};
        }

    for (int k0 = 0; k0 < kc; k0 += HKC) {
        // global -> registers: 2048 16-byte segments of each operand, eight per thread; thread e
        // takes segment e % 4 of row (e / 4) % 128 of plane e / 512
        half8 ra[8], rb[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int e = tid + 256 * s, p = e >> 9, row = (e >> 2) & 127, sg = e & 3;
            ra[s] = *reinterpret_cast<const half8*>(La + p * aps + (m0 + row) * HSPLIT_K + k0 + 8 * sg);
            rb[s] = *reinterpret_cast<const half8*>(Ub + p * bps + (n0 + row) * HSPLIT_K + k0 + 8 * sg);
        }
        __syncthreads(); // the previous chunk's fragments have been read
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int e = tid + 256 * s, p = e >> 9, row = (e >> 2) & 127, sg = e & 3;
            sA[p][row][hseg(row, sg)] = ra[s];
            sB[p][row][hseg(row, sg)] = rb[s];
        }
        __syncthreads();
        // this wave's four column blocks of U (the MFMA's A operand: U^T, 16 columns x 32 k)
        const int sgl = lane >> 4;
        half8 bf[4][4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const int col = wn * 64 + cb * 16 + (lane & 15);
#pragma unroll
            for (int p = 0; p < 4; p++)
                bf[cb][p] = sB[p][col][hseg(col, sgl)];
        }
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const int row = wm * 64 + rb * 16 + (lane & 15);
            half8 af[4];
#pragma unroll
            for (int p = 0; p < 4; p++)
                af[p] = sA[p][row][hseg(row, sgl)];
            const half8 nih = -af[2], nil = -af[3];
            // re += Ur Lr - Ui Li
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[0], are[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][1], af[0], are[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[1], are[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], nih, are[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][3], nih, are[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                are[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], nil, are[rb][cb], 0, 0, 0);
            // im += Ur Li + Ui Lr
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[2], aim[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][1], af[2], aim[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][0], af[3], aim[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], af[0], aim[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][3], af[0], aim[rb][cb], 0, 0, 0);
#pragma unroll
            for (int cb = 0; cb < 4; cb++)
                aim[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(bf[cb][2], af[1], aim[rb][cb], 0, 0, 0);
        }
    }

    // C -= 2^-(ea + eb) D^T: lane holds row (lane & 15) of each row block and columns
    // 4 (lane >> 4) + r of each column block
    int ecol[4][4];
#pragma unroll
    for (int cb = 0; cb < 4; cb++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            ecol[cb][r] = eb[n0 + wn * 64 + cb * 16 + 4 * (lane >> 4) + r];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        const int64_t row = m0 + wm * 64 + rb * 16 + (lane & 15);
        if (row >= M)
            continue;
        const int erow = ea[row];
        c64 cv[4][4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t col = n0 + wn * 64 + cb * 16 + 4 * (lane >> 4) + r;
                const int64_t ok = col < N;
                cv[cb][r] = C[ok * (row + col * ldc)];
            }
#pragma unroll
        for (int cb = 0; cb < 4; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t col = n0 + wn * 64 + cb * 16 + 4 * (lane >> 4) + r;
                if (col < N) {
                    const int sc = -(erow + ecol[cb][r]);
                    c64 v = cv[cb][r];
                    v.re -= ldexpf(are[rb][cb][r], sc);
                    v.im -= ldexpf(aim[rb][cb][r], sc);
                    C[row + col * ldc] = v;
                }
            }
    }
}

int launch_split_f16(const c64* S, int64_t ld, int64_t R, int K, _Float16* planes, int* e, int64_t cap,
                     hipStream_t st)
{
    if (R <= 0)
        return 0;
    if (K < 1 || K > HSPLIT_K)
        return set_error("fp16 split: K = %d outside 1..%d", K, HSPLIT_K), -1;
    const int64_t Rp = (R + HB - 1) / HB * HB;
    if (Rp > cap)
        return set_error("fp16 split: %lld rows exceed the buffer's %lld", (long long)Rp, (long long)cap), -1;
    launch_kernel(&split_f16, dim3((unsigned)(Rp * 16 / 256)), dim3(256), 0, st, S, ld, R, K, Rp, planes, cap, e);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int gemm_sub_f16(int64_t M, int64_t N, int K, const _Float16* La, const int* ea, int64_t acap, const _Float16* Ub,
                 const int* eb, int64_t bcap, c64* C, int64_t ldc, hipStream_t st, KernelTimer* kt)
{
    if (M <= 0 || N <= 0 || K <= 0)
        return 0;
    if (K > HSPLIT_K)
        return set_error("fp16 product: K = %d above %d", K, HSPLIT_K), -1;
    const int64_t gx = (M + HB - 1) / HB, gy = (N + HB - 1) / HB;
    if (gx * HB > acap || gy * HB > bcap)
        return set_error("fp16 product: operands not padded to whole tiles"), -1;
    if (gy > 65535)
        return set_error("fp16 product: too many column tiles"), -1;
    const int kc = (K + HKC - 1) / HKC * HKC;
    if (kt)
        kt->begin(st);
    launch_kernel(&cgemm_sub_f16, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, M, N, kc, La, ea, acap, Ub, eb,
                  bcap, C, ldc);
    if (kt)
        kt->end(st, 8.0 * (double)M * (double)N * (double)K);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

} // namespace frico
