// Complex64 trailing products of the unpivoted LU on fp16 MFMA (v_mfma_f32_32x32x16_f16), and the
// products of its triangular solves (v_mfma_f32_16x16x32_f16).
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
// MFMAs per block and k step.  The product is taken transposed (D = U^T L^T), so that consecutive
// lanes of an accumulator value hold consecutive rows of C and the read-modify-write of C goes down
// its columns.
//
// Lane layouts (checked on gfx950 against FP64 products of random operands):
//   v_mfma_f32_16x16x32_f16: A[i][k] in lane i + 16 (k / 8), slot k % 8; B[k][j] in lane
//   j + 16 (k / 8), slot k % 8; D[i][j] in lane j + 16 (i / 4), value i % 4.
//   v_mfma_f32_32x32x16_f16: A[i][k] in lane i + 32 (k / 8), slot k % 8; B[k][j] in lane
//   j + 32 (k / 8), slot k % 8; D[i][j] in lane j + 32 ((i / 4) % 2), value 4 (i / 8) + i % 4.
#include "common.hpp"
#include "lu.hpp"

#include <cstdlib>

namespace frico {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------------------------------------
// Split: S[k + r ld] (r < R rows of K <= 128 contiguous values each) -> four planes
// P[p] (p = re_hi, re_lo, im_hi, im_lo; element (r, k) at hsplit_at(r, k, cap), HSPLIT_K halves a
// row, zero beyond K and for the padding rows R <= r < Rp) and the row exponents e[r].  16 lanes a row, 8 values a lane.
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
    const int64_t off = hsplit_at(r, 8 * g, cap), ps = cap * HSPLIT_K;
    *reinterpret_cast<half8*>(P + off) = rh;
    *reinterpret_cast<half8*>(P + ps + off) = rl;
    *reinterpret_cast<half8*>(P + 2 * ps + off) = ih;
    *reinterpret_cast<half8*>(P + 3 * ps + off) = il;
    if (g == 0)
        E[r] = e;
}

// ----------------------------------------------------------------------------------------------
// C[M x N] -= L U with L given as split rows (La, ea: M rows) and U as split columns (Ub, eb:
// N columns), K <= 128 (kc = K rounded up to 16).  Tile 128 x 64 CBW per 256-thread workgroup,
// 2 x 2 waves of 64 x 32 CBW, each wave 2 x CBW blocks of 32 x 32 (v_mfma_f32_32x32x16_f16).  K
// runs in 16-deep stages through two LDS buffers, loaded straight from the slab layout into LDS
// (global_load_lds_dwordx4): each wave reads its fragments of stage s, issues the loads of stage
// s + 1 into the other buffer, runs its MFMAs of stage s, and one barrier ends the stage.  The split
// planes are padded to whole 128-row tiles, so the loads need no bounds; only C is masked.
// ----------------------------------------------------------------------------------------------
constexpr int HKC = 32; // k chunk of the solve products
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

typedef float f32x16 __attribute__((ext_vector_type(16)));

// 16-byte slot of segment s (k 8 s .. 8 s + 7 of a stage) of row r within one plane of a stage:
// rows are 32 bytes apart, and the two segments swap places on rows whose bit 2 differs from bit
// 3, so that the eight (sixteen) consecutive rows of a quarter-wave (half-wave) read of one
// segment fall on distinct banks with 128 (256) bytes of banks.
__device__ inline int xslot(int r, int s) { return 2 * r + (s ^ (((r >> 2) ^ (r >> 3)) & 1)); }

// One lane's part of a direct global -> LDS load: 16 bytes from g to lds_wave + 16 lane, where
// lds_wave is the same for the whole wave.
__device__ inline void load_lds16(const _Float16* g, char* lds_wave)
{
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)g,
                                     (__attribute__((address_space(3))) void*)lds_wave, 16, 0, 0);
}

// s_waitcnt vmcnt(0) lgkmcnt(0): this wave's vector memory loads (the LDS DMA included) and LDS
// reads have landed
__device__ inline void wait_loads() { __builtin_amdgcn_s_waitcnt(0x70); }

template <int CBW, int OCC, bool TWO = false>
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
    constexpr int BM = 128, BN = 64 * CBW;
    constexpr int ASL = 4 * BM * 2, BSL = 4 * BN * 2; // 16-byte slots of a stage: L, U (4 planes, 2 a row)
    constexpr int NA = ASL / 256, NB = BSL / 256;      // loads a thread a stage
    constexpr int STAGE = (ASL + BSL) * 16;            // bytes of a stage
    extern __shared__ __attribute__((aligned(16))) char lds[]; // two stages
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = wv & 1, wn = wv >> 1;
    int64_t tm, tn;
    if (!tile_of(blockIdx.x, gx, gy, tm, tn))
        return;
    const int64_t m0 = tm * BM, n0 = tn * BN;

    // this thread's loads of a stage (slot tid + 256 i of the L and U regions): plane p, row r and
    // segment s of slot w = (tid + 256 i) % (2 rows) -> element offset within the k slab 0 of the
    // operand; the slab of stage ks adds ks cap HSLAB
    // (byte offsets; the tile's row and column offsets are in the wave-uniform bases)
    uint32_t aoff[NA], boff[NB];
#pragma unroll
    for (int i = 0; i < NA; i++) {
        const int q = tid + 256 * i, p = q / (2 * BM), w = q % (2 * BM), r = w >> 1;
        const int s = (w & 1) ^ (((r >> 2) ^ (r >> 3)) & 1);
        aoff[i] = (uint32_t)(2 * (p * acap * HSPLIT_K + r * HSLAB + 8 * s));
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
        const int q = tid + 256 * i, p = q / (2 * BN), w = q % (2 * BN), r = w >> 1;
        const int s = (w & 1) ^ (((r >> 2) ^ (r >> 3)) & 1);
        boff[i] = (uint32_t)(2 * (p * bcap * HSPLIT_K + r * HSLAB + 8 * s));
    }
    const int S0 = kc / HSLAB, S = S0 + (TWO ? h1.kc / HSLAB : 0);
    auto issue = [&](int st, int buf) {
        const bool second = TWO && st >= S0;
        const int ks = second ? st - S0 : st;
        const char* A = reinterpret_cast<const char*>((second ? h1.La : La) + (int64_t)ks * acap * HSLAB + m0 * HSLAB);
        const char* B = reinterpret_cast<const char*>((second ? h1.Ub : Ub) + (int64_t)ks * bcap * HSLAB + n0 * HSLAB);
        char* dst = lds + buf * STAGE + 64 * 16 * wv;
#pragma unroll
        for (int i = 0; i < NA; i++)
            load_lds16(reinterpret_cast<const _Float16*>(A + aoff[i]), dst + 256 * 16 * i);
#pragma unroll
        for (int i = 0; i < NB; i++)
            load_lds16(reinterpret_cast<const _Float16*>(B + boff[i]), dst + ASL * 16 + 256 * 16 * i);
    };

    f32x16 are[2][CBW], aim[2][CBW];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < CBW; b++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                are[a][b][v] = 0.0f;
                aim[a][b][v] = 0.0f;
            }
    // lane's rows: m0 + wm 64 + rb 32 + (lane & 31); columns: n0 + wn 32 CBW + cb 32 + 8 (v / 4) +
    // 4 (lane >> 5) + v % 4 for accumulator value v
    const int64_t rbase = m0 + wm * 64 + (lane & 31), cbase = n0 + wn * 32 * CBW + 4 * (lane >> 5);
    const int sgl = lane >> 5;

    issue(0, 0);
    wait_loads();
    __syncthreads();
    for (int st = 0; st < S; st++) {
        const char* buf = lds + (st & 1) * STAGE;
        if (TWO && st == S0) {
            // rescale the first half's sums to the second half's exponents (the exponent tables
            // cover the padded rows and columns of whole tiles, so their reads need no bounds)
            int dr[2];
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
                dr[rb] = h1.ea[rbase + 32 * rb] - ea[rbase + 32 * rb];
#pragma unroll
            for (int cb = 0; cb < CBW; cb++) {
                const int* e0 = eb + cbase + 32 * cb;
                const int* e1 = h1.eb + cbase + 32 * cb;
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int o = 8 * (v >> 2) + (v & 3), dc = e1[o] - e0[o];
#pragma unroll
                    for (int rb = 0; rb < 2; rb++) {
                        int d = dr[rb] + dc;
                        d = d < -200 ? -200 : (d > 100 ? 100 : d);
                        are[rb][cb][v] = ldexpf(are[rb][cb][v], d);
                        aim[rb][cb][v] = ldexpf(aim[rb][cb][v], d);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // fragments of this stage: U columns (the MFMA's A operand, U^T) and L rows (B operand, L^T)
        // (read in the order the MFMAs below consume them)
        half8 uf[CBW][4], lf[2][4];
        auto rd_u = [&](int p) {
#pragma unroll
            for (int cb = 0; cb < CBW; cb++) {
                const int c = wn * 32 * CBW + cb * 32 + (lane & 31);
                uf[cb][p] = *reinterpret_cast<const half8*>(buf + (ASL + p * 2 * BN + xslot(c, sgl)) * 16);
            }
        };
        auto rd_l = [&](int p) {
#pragma unroll
            for (int rb = 0; rb < 2; rb++) {
                const int r = wm * 64 + rb * 32 + (lane & 31);
                lf[rb][p] = *reinterpret_cast<const half8*>(buf + (p * 2 * BM + xslot(r, sgl)) * 16);
            }
        };
        rd_u(0);
        rd_l(0);
        rd_l(2);
        rd_u(1);
        rd_l(1);
        rd_l(3);
        rd_u(2);
        rd_u(3);
        // the loads of stage st + 1; past the last stage the last one again into the free buffer
        // (no branch, so that the scheduler can interleave all of this stage's work)
        issue(st + 1 < S ? st + 1 : S - 1, (st + 1) & 1);
        half8 nh[2], nl[2];
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            nh[rb] = -lf[rb][2];
            nl[rb] = -lf[rb][3];
        }
        // re += Ur Lr - Ui Li ; im += Ur Li + Ui Lr (hi hi + lo hi + hi lo of each); the blocks
        // interleaved so that consecutive MFMAs write different accumulators
#define FRICO_MF(acc, u, l)                                                                                   \
    _Pragma("unroll") for (int rb = 0; rb < 2; rb++) _Pragma("unroll") for (int cb = 0; cb < CBW; cb++) acc[rb][cb] = \
        __builtin_amdgcn_mfma_f32_32x32x16_f16(u, l, acc[rb][cb], 0, 0, 0);
        FRICO_MF(are, uf[cb][0], lf[rb][0])
        FRICO_MF(aim, uf[cb][0], lf[rb][2])
        FRICO_MF(are, uf[cb][1], lf[rb][0])
        FRICO_MF(aim, uf[cb][1], lf[rb][2])
        FRICO_MF(are, uf[cb][0], lf[rb][1])
        FRICO_MF(aim, uf[cb][0], lf[rb][3])
        FRICO_MF(are, uf[cb][2], nh[rb])
        FRICO_MF(aim, uf[cb][2], lf[rb][0])
        FRICO_MF(are, uf[cb][3], nh[rb])
        FRICO_MF(aim, uf[cb][3], lf[rb][0])
        FRICO_MF(are, uf[cb][2], nl[rb])
        FRICO_MF(aim, uf[cb][2], lf[rb][1])
        // schedule: the LDS reads spread over the first MFMAs, then the stage loads over the next
        constexpr int MB = 2 * CBW; // MFMAs a term
#pragma unroll
        for (int g = 0; g < 8; g++) {
            __builtin_amdgcn_sched_group_barrier(0x100, (2 * CBW + 4) / 4, 0); // DS reads
            __builtin_amdgcn_sched_group_barrier(0x008, MB / 2, 0);            // MFMA
        }
#pragma unroll
        for (int g = 0; g < NA + NB; g++) {
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0); // VMEM (LDS DMA)
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
#undef FRICO_MF
        __builtin_amdgcn_sched_barrier(0); // the wait below stays after the MFMAs
        wait_loads();                      // this wave's loads of the next stage have landed
        __syncthreads();                   // everyone's have, and this stage's buffer is free
    }
    if constexpr (TWO) {
        ea = h1.ea;
        eb = h1.eb;
    }
    // C -= 2^-(ea + eb) D^T, a row of 32 x 32 blocks at a time (its C loads all in flight together),
    // addressed as the tile's corner plus 32-bit byte offsets (a tile spans < 4 GiB of C)
    int ec[CBW][16];
#pragma unroll
    for (int cb = 0; cb < CBW; cb++)
#pragma unroll
        for (int v = 0; v < 16; v++)
            ec[cb][v] = eb[cbase + 32 * cb + 8 * (v >> 2) + (v & 3)];
    char* Ct = reinterpret_cast<char*>(C + m0 + n0 * ldc);
    const uint32_t ldb8 = (uint32_t)ldc * 8u;
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        const int64_t row = rbase + 32 * rb;
        if (row < M) {
            const int erow = ea[row];
            const uint32_t o0 = (uint32_t)(row - m0) * 8u + (uint32_t)(cbase - n0) * ldb8;
            c64 cr[CBW][16];
#pragma unroll
            for (int cb = 0; cb < CBW; cb++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int dc = 32 * cb + 8 * (v >> 2) + (v & 3);
                    const bool ok = cbase + dc < N;
                    cr[cb][v] = *reinterpret_cast<const c64*>(Ct + (ok ? o0 + (uint32_t)dc * ldb8 : 0u));
                }
#pragma unroll
            for (int cb = 0; cb < CBW; cb++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const int dc = 32 * cb + 8 * (v >> 2) + (v & 3);
                    if (cbase + dc < N) {
                        const int sc = -(erow + ec[cb][v]);
                        c64 x = cr[cb][v];
                        x.re -= ldexpf(are[rb][cb][v], sc);
                        x.im -= ldexpf(aim[rb][cb][v], sc);
                        *reinterpret_cast<c64*>(Ct + o0 + (uint32_t)dc * ldb8) = x;
                    }
                }
        }
        __builtin_amdgcn_sched_barrier(0);
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
    // branch-free loads of a chunk (out-of-range ones read element 0 and are zeroed when the chunk
    // is split), all in flight together; the next chunk's are issued before this chunk's MFMAs
    c64 va[2][8], vb[8];
    auto gload = [&](int64_t k0) {
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
    };
    gload(0);
    for (int64_t k0 = 0; k0 < K; k0 += HKC) {
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
        if (k0 + HKC < K)
            gload(k0 + HKC);
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

template <int CBW, bool TWO>
static hipError_t launch_f16_prod(dim3 grid, hipStream_t st, int64_t M, int64_t N, int kc, const _Float16* La,
                                  const int* ea, int64_t acap, const _Float16* Ub, const int* eb, int64_t bcap, c64* C,
                                  int64_t ldc, int64_t gx, int64_t gy, F16Half h1, const Batch& bt)
{
    // two 16-deep stages of the four planes of 128 rows of L and 64 CBW columns of U
    constexpr size_t lds = 2 * (size_t)(4 * 128 * 2 + 4 * 64 * CBW * 2) * 16;
    auto kern = &cgemm_sub_f16<CBW, CBW == 2 ? 2 : 3, TWO>;
    const hipError_t e = allow_lds(reinterpret_cast<const void*>(kern), lds);
    if (e != hipSuccess)
        return e;
    launch_kernel(kern, grid, dim3(256), lds, st, M, N, kc, La, ea, acap, Ub, eb, bcap, C, ldc, gx, gy, h1, bt);
    return hipSuccess;
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
    // 128 x 64 tiles (with one k half, the 128 x 128 tiles' epilogue outweighs their fewer loads)
    const int BM = 128, BN = 64;
    const int64_t gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
    // the split planes are padded to 128-row tiles, which cover any smaller tile
    if ((M + 127) / 128 * 128 > acap || (N + 127) / 128 * 128 > bcap)
        return set_error("fp16 product: operands not padded to whole tiles"), -1;
    const int64_t nblk = (gx * gy + 7) / 8 * 8;
    if (nblk > 0x7fffffff)
        return set_error("fp16 product: too many tiles"), -1;
    const int kc = (K + HSLAB - 1) / HSLAB * HSLAB;
    if (kt)
        kt->begin(st);
    const dim3 grid((unsigned)nblk, 1, (unsigned)nz);
    FRICO_HIP_CHECK((launch_f16_prod<1, false>(grid, st, M, N, kc, La, ea, acap, Ub, eb, bcap, C, ldc, gx, gy,
                                               F16Half{}, bt)));
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
    if (K0 % HSLAB != 0)
        return set_error("fp16 product: first half K = %d not a multiple of %d", K0, HSLAB), -1;
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
    const F16Half h1{La1, ea1, Ub1, eb1, (K1 + HSLAB - 1) / HSLAB * HSLAB};
    if (kt)
        kt->begin(st);
    const dim3 grid((unsigned)nblk, 1, (unsigned)nz);
    if (tile == 128)
        FRICO_HIP_CHECK((launch_f16_prod<2, true>(grid, st, M, N, K0, La0, ea0, acap, Ub0, eb0, bcap, C, ldc, gx, gy, h1,
                                                  bt)));
    else
        FRICO_HIP_CHECK((launch_f16_prod<1, true>(grid, st, M, N, K0, La0, ea0, acap, Ub0, eb0, bcap, C, ldc, gx, gy, h1,
                                                  bt)));
    const double K = (double)K0 + (double)K1;
    if (kt)
        kt->end(st, 8.0 * (double)M * (double)N * K, 16.0 * (double)M * (double)N + 8.0 * ((double)M + (double)N) * K);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

} // namespace frico
