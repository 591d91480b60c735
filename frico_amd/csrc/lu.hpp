// Dense complex LU / solve / refinement drivers (device side in lu.hip).
#pragma once

#include "common.hpp"

#include <vector>

namespace frico {

constexpr int LU_NB_MAX = 128;
constexpr int LU_MAX_GROUPS = 512;

// Device workspace of the panel factorisation.
struct LuWork {
    // panel exchange area (lu_exchange_words(max_groups) 64-bit words), zeroed by lu_factor
    unsigned long long* ll = nullptr;
    size_t nll = 0;
    int* flags = nullptr; // [0] abort, [1] info
    int max_groups = 0;   // co-resident workgroup limit of the panel kernel
    void* dinv = nullptr; // [ceil(n/64)][2][64*64] diagonal-chunk inverses, written at the end of lu_factor
    // W = L(t, j) Linv_j below each diagonal chunk (n x n, leading dimension n), written at the end of
    // lu_factor when non-null (needs dinv and nb % 64 == 0)
    void* wv = nullptr;
    // per panel: the row-interchange list written by the panel kernel (lu_perm_stride() ints each)
    int* perm = nullptr;
    int64_t nperm = 0; // panels it can hold
    hipStream_t side = nullptr;    // second stream (default priority) for the look-ahead, or null
    hipEvent_t* events = nullptr;  // >= 2 * npanels + 1 events (timing disabled)
    int nevents = 0;
    // lu_nopiv_factor: nb x n complex values (the transposed block column), or null
    void* tbuf = nullptr;
    int64_t ntbuf = 0;
    // lu_nopiv_factor (complex64): fp16 split operands of the trailing products (see gemm_f16.hip),
    // or null for the fp32 products.  ha[parity]: L21 of the steps of that parity (hacap rows);
    // hb[0] / hb[1]: U12 of the products on the main / side stream (hbcap columns).  Each holds
    // four planes of cap x HSPLIT_K halves; hea / heb their row / column exponents.
    // With four of each (nsplit = 4), the paired steps keep L21 of four steps (step mod 4) and two
    // U12 operands a stream (hb[0..1] main, hb[2..3] side).
    _Float16* ha[4] = {nullptr, nullptr, nullptr, nullptr};
    int* hea[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t hacap = 0;
    _Float16* hb[4] = {nullptr, nullptr, nullptr, nullptr};
    int* heb[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t hbcap = 0;
    int nsplit = 0;
    // lu_nopiv_factor (complex64): block-column and block-row solves by the lane-per-row kernels
    // (false: transposed block column through tbuf and the 64-row triangular-solve kernels)
    bool fused_step = true; // complex64 unpivoted LU with fp16 products: tri_solve_split for the block solves
    bool direct_solves = true;
};

// fp16 split products (gemm_f16.hip).  Rows of a split operand: HSPLIT_K halves each, stored in
// slabs of HSLAB consecutive k: element (r, k) of a plane of cap rows sits at hsplit_at(r, k, cap),
// so that the k slab of a run of rows is one contiguous stretch (32 bytes a row).  A view that starts
// at row off of a plane begins at element off * HSLAB.
constexpr int HSPLIT_K = 128;
constexpr int HSLAB = 16;
__host__ __device__ inline int64_t hsplit_at(int64_t r, int k, int64_t cap)
{
    return (int64_t)(k / HSLAB) * cap * HSLAB + r * HSLAB + (k % HSLAB);
}
// Device bytes of a split operand of cap rows (cap a multiple of 128): planes and exponents.
inline size_t split_f16_bytes(int64_t cap) { return (size_t)cap * (4 * HSPLIT_K * 2 + sizeof(int)); }
// S[k + r ld], r < R, k < K <= HSPLIT_K -> planes (cap rows each) and exponents of the R rows, padded
// with zero rows to a multiple of 128.
int launch_split_f16(const c64* S, int64_t ld, int64_t R, int K, _Float16* planes, int* e, int64_t cap,
                     hipStream_t st);
// The same split of S[r + k ld] (rows contiguous, as a block column of a matrix lies in memory).
int launch_split_f16_cm(const c64* S, int64_t ld, int64_t R, int K, _Float16* planes, int* e, int64_t cap,
                        hipStream_t st);
// Block solves of the unpivoted complex64 LU after A11 = L11 U11 (w <= 128 columns from k):
// L21 = A21 U11^{-1} for the m rows below the block, and U12 = L11^{-1} A12 for columns c0 .. c0 + nc.
// The step's block solve with the fp16 split of its result (blocksolve.hip): trans false: U12 =
// L11^{-1} A(k .. k+w, c0 .. c0+nc) in place, split per column; trans true: L21 = A21 U11^{-1} for
// rows k+w+c0 .. k+w+c0+nc of A21, in place, split per row.
int launch_tri_solve_split(c64* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, bool trans, _Float16* planes,
                           int* e, int64_t cap, hipStream_t st, const Batch& bt = Batch{}, int nz = 1);
int launch_solve_l21(c64* A, int64_t lda, int64_t k, int w, int64_t m, hipStream_t st);
int launch_solve_u12(c64* A, int64_t lda, int64_t k, int w, int64_t c0, int64_t nc, hipStream_t st);
// C[M x N] -= L U, L as split rows (La, ea; acap rows), U as split columns (Ub, eb; bcap columns).
// Two steps' products in one pass over C: C -= L0 U0 + L1 U1 (K0 = 128 and K1 <= 128, each pair of
// operands split with its own exponents; L0, L1 with acap rows, U0, U1 with bcap columns).
int gemm_sub_f16_pair(int64_t M, int64_t N, int K0, const _Float16* La0, const int* ea0, const _Float16* Ub0,
                      const int* eb0, int K1, const _Float16* La1, const int* ea1, const _Float16* Ub1, const int* eb1,
                      int64_t acap, int64_t bcap, c64* C, int64_t ldc, hipStream_t st, KernelTimer* kt,
                      const Batch& bt = Batch{}, int nz = 1);
int gemm_sub_f16(int64_t M, int64_t N, int K, const _Float16* La, const int* ea, int64_t acap, const _Float16* Ub,
                 const int* eb, int64_t bcap, c64* C, int64_t ldc, hipStream_t st, KernelTimer* kt,
                 const Batch& bt = Batch{}, int nz = 1);

// Products of the triangular solves on fp16 MFMA (gemm_f16.hip): C -= A B, complex64 and
// column-major, A's rows scaled by 2^ea and B's columns by 2^eb (computed here into eb[N], over B's
// K rows) before the hi/lo split.  launch_rowexp_blocks fills E[b ldE + r] with row r's exponent
// over columns [b W, (b + 1) W) of A.
int gemm_sub_f16s(int64_t M, int64_t N, int64_t K, const c64* A, int64_t lda, const int* ea, const c64* B, int64_t ldb,
                  int* eb, c64* C, int64_t ldc, hipStream_t st);
int launch_rowexp_blocks(const c64* A, int64_t lda, int64_t n, int W, int* E, int64_t ldE, hipStream_t st);

// C -= A B.  exact: four real products per complex multiply-add even where three would be used
// (residuals, whose rounding floor bounds the refinement).
template <typename T>
int gemm_sub(int64_t M, int64_t N, int64_t K, const cx<T>* A, int64_t lda, const cx<T>* B, int64_t ldb, cx<T>* C,
             int64_t ldc, hipStream_t st, KernelTimer* kt, bool exact = false);
// Residual GEMM with a complex64-stored A and FP64 arithmetic: C -= double(A) * B.
int gemm_sub_mixed(int64_t M, int64_t N, int64_t K, const c64* A, int64_t lda, const c128* B, int64_t ldb, c128* C,
                   int64_t ldc, hipStream_t st, KernelTimer* kt);
template <typename T> int panel_rows(int w, bool shared_cu = false);
// naug > 0: columns [n, n + naug) of A (same leading dimension) are right-hand sides carried
// through the factorisation; they end as L^{-1} P B, ready for lu_solve_upper.
template <typename T>
int lu_factor(LuWork& wk, cx<T>* A, int64_t n, int64_t lda, int* ipiv, int nb, hipStream_t st, KernelTimer* gemm_timer,
              int64_t naug = 0);
// Z = L D L^T without pivoting, for complex symmetric Z (only its lower triangle is read).  The
// factors are stored as lu_factor's would be with no interchanges (L below the diagonal, U = D L^T
// on and above it; empty interchange lists in LuWork::perm, ipiv the identity), so lu_solve and
// lu_solve_upper take them as they are.  naug as for lu_factor.  nb <= 128.
template <typename T>
int ldlt_factor(LuWork& wk, cx<T>* A, int64_t n, int64_t lda, int* ipiv, int nb, hipStream_t st,
                KernelTimer* gemm_timer, int64_t naug = 0);
// LU without pivoting (for the refinement form, which corrects its rounding): the same layout as
// lu_factor's factors with no interchanges (empty lists in LuWork::perm, ipiv the identity), so the
// solves take them as they are.  Needs LuWork::tbuf.  naug as for lu_factor.  A zero pivot is
// reported like zgetrf's info.  Instantiated for complex64 only.
template <typename T>
int lu_nopiv_factor(LuWork& wk, cx<T>* A, int64_t n, int64_t lda, int* ipiv, int nb, hipStream_t st,
                    KernelTimer* gemm_timer, int64_t naug = 0);
int launch_zero_bytes(void* p, size_t bytes, hipStream_t st);
int launch_copy_bytes(const void* src, void* dst, size_t bytes, hipStream_t st);
// ints per panel of LuWork::perm
int lu_perm_stride();
// 64-bit words of LuWork::ll for gmax workgroups
size_t lu_exchange_words(int gmax);
// perm: the interchange lists written by lu_factor (LuWork::perm).
// dinv: the 64 x 64 diagonal-chunk inverses written by lu_factor (LuWork::dinv), or null to use
// in-LDS triangular solves.  wv: the W/V products of lu_factor (LuWork::wv), or null; with wv the
// solve runs one product per 64-row chunk and needs ybuf, an n x nrhs scratch (leading dimension n).
// A second stream on which the backward half of a complex64 refinement solve adds each finished
// 512-row block of the correction X to the FP64 solution S and takes it off the residual:
// S(block) += X(block, :), R -= Z(:, block) S(block, :).  With R = B beforehand, R ends as
// B - Z S, computed as the full product would (the K sum in blocks).  ev: at least
// ceil(n / 128) + 1 events.  With R2 (zero beforehand, same layout as R), the blocked solve takes
// two blocks a launch, the second into R2, and adds R2 to R at the end.
struct ResidStream {
    hipStream_t side = nullptr;
    hipEvent_t* ev = nullptr;
    int nev = 0;
    const c128* Z = nullptr;
    int64_t ldz = 0;
    c128* S = nullptr;
    int64_t lds = 0;
    c128* R = nullptr;
    int64_t ldr = 0;
    c128* R2 = nullptr;
    bool exact = true; // four real products per complex multiply-add (see gemm_sub)
    // The solves' own products on fp16 MFMA (gemm_sub_f16s) when aexp is set: row r's exponent
    // over the factors' block b of lu_solve_block_rows() columns at aexp[b * ldae + r]; bexp, bexp2:
    // nrhs-long scratch for the column exponents of the main and the second stream's products.
    const int* aexp = nullptr;
    int64_t ldae = 0;
    int* bexp = nullptr;
    int* bexp2 = nullptr;
};
// Split operands and zero-pivot words of a lockstep batch of unpivoted complex64 factorisations
// (lu_nopiv_batched): ha[i] / hea[i] hold system z's planes at + z spa and exponents at + z sea
// (hacap rows a system), hb likewise; info[z] the first zero pivot (1-based) of system z, 0 if none.
struct BatchLu {
    _Float16* ha[4] = {nullptr, nullptr, nullptr, nullptr};
    int* hea[4] = {nullptr, nullptr, nullptr, nullptr};
    _Float16* hb[4] = {nullptr, nullptr, nullptr, nullptr};
    int* heb[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t hacap = 0, hbcap = 0, spa = 0, sea = 0, spb = 0, seb = 0;
    int* info = nullptr;
};
// The paired-step unpivoted complex64 LU (fp16 products) of nz systems of size n in lockstep, system
// z at A + z sa with naug carried columns; n + naug <= 4096.
int lu_nopiv_batched(const BatchLu& w, c64* A, int64_t sa, int64_t n, int64_t lda, int64_t naug, int nz,
                     hipStream_t st);
// Chained complex64 solves (one launch a triangle) for one to four right-hand sides of factors
// without interchanges: the exchange area of lu_chain_words(n, nrhs) 64-bit words.
struct ChainSolve {
    unsigned long long* words = nullptr;
    size_t nwords = 0;
};
size_t lu_chain_words(int64_t n, int64_t nrhs);
// Both halves (forward) or the backward half of nz systems' solves with factors without interchanges,
// in place: system z's factors at A + z sa, its right-hand sides at B + z sb, its exchange area at
// words + z lu_chain_words(n, nrhs) (cleared here).  1 <= nrhs <= 4.
int lu_chain_solve_batched(const c64* A, int64_t sa, int64_t n, int64_t lda, c64* B, int64_t sb, int64_t ldb,
                           int64_t nrhs, int nz, unsigned long long* words, bool forward, hipStream_t st);
// C_z -= A_z B_z for nz systems (A complex64, B and C complex128; N <= 4 columns), system z's
// operands at A + z sa, B + z sb, C + z sc.
int gemv_sub_mixed_batched(int64_t M, int64_t N, int64_t K, const c64* A, int64_t lda, int64_t sa, const c128* B,
                           int64_t ldb, int64_t sb, c128* C, int64_t ldc, int64_t sc, int nz, hipStream_t st);
// Rows of a block of the blocked complex64 solves (their substitutions and products).
int lu_solve_block_rows();
// Whether complex64 solves without interchanges take the blocked path (which needs neither the
// diagonal-block inverses nor the W products).
bool lu_blocked_solves();
// perm: the factorisation's interchange lists, or null for factors without interchanges.
template <typename T>
int lu_solve(const cx<T>* A, int64_t n, int64_t lda, const int* perm, int nb, cx<T>* B, int64_t ldb, int64_t nrhs,
             hipStream_t st, const void* dinv, const void* wv = nullptr, void* ybuf = nullptr,
             const ResidStream* rs = nullptr, const ChainSolve* cs = nullptr);
// B = U^{-1} B in place (the backward half of the solve), U the upper factor of lu_factor.
template <typename T>
int lu_solve_upper(const cx<T>* A, int64_t n, int64_t lda, cx<T>* B, int64_t ldb, int64_t nrhs, hipStream_t st,
                   const ResidStream* rs = nullptr, const ChainSolve* cs = nullptr);
template <typename T>
int lu_solve_transposed(const cx<T>* A, int64_t n, int64_t lda, const int* ipiv, int nb, cx<T>* B, int64_t ldb,
                        int64_t nrhs, hipStream_t st);
int launch_convert_c128_c64(const c128* X, int64_t ldx, c64* Y, int64_t ldy, int64_t rows, int64_t cols, hipStream_t st);
int launch_add_correction(c128* X, int64_t ldx, const c64* D, int64_t ldd, int64_t rows, int64_t cols, hipStream_t st);
int launch_col_norm2(const c128* X, int64_t ldx, int64_t rows, int64_t cols, double* out, hipStream_t st);
int launch_symmetrize(c128* Z, int64_t n, int64_t ldz, hipStream_t st);

} // namespace frico
