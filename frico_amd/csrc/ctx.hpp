// Context, host mesh and device-buffer definitions shared by the C-ABI sources.
#pragma once

#include "../../include/frico_gpu.h"

#include "adjoint.hpp"
#include "common.hpp"
#include "lu.hpp"
#include "mesh_dev.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "dunavant_rules.inc"

namespace frico {

template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int ensure(size_t count)
    {
        if (count <= n && p)
            return 0;
        release();
        if (count == 0)
            return 0;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            n = 0;
            set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
            return FRICO_ERR_NOMEM;
        }
        n = count;
        return 0;
    }
    int upload(const T* h, size_t count, hipStream_t st)
    {
        int rc = ensure(count);
        if (rc)
            return rc;
        if (count)
            FRICO_HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
        return 0;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

namespace geom {

struct vec3 {
    double x, y, z;
};
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator*(vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator/(vec3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3 cross(vec3 a, vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(vec3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

struct QP {
    vec3 p;
    double w;
};

// Heron's formula in Kahan's stable ordering (src/quadratures.cpp:41-57).
inline double kahan_area(vec3 p0, vec3 p1, vec3 p2)
{
    std::array<double, 3> l{norm(p1 - p0), norm(p2 - p0), norm(p2 - p1)};
    std::sort(l.begin(), l.end());
    const double a = l[2], b = l[1], c = l[0];
    return 0.25 * std::sqrt((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c)));
}

// Symmetric Dunavant rule on a triangle: each orbit expands to 1, 3 or 6 barycentric
// permutations in a fixed order; weights are scaled by the triangle area
// (src/quadratures.cpp:321-372).
inline void dunavant(int degree, vec3 p0, vec3 p1, vec3 p2, std::vector<QP>& out)
{
    const int rule = degree == 0 ? 0 : degree - 1;
    const double area = kahan_area(p0, p1, p2);
    auto at = [&](double a, double b, double c) { return p0 * a + p1 * b + p2 * c; };
    for (int o = frico_dunavant_orbit_begin[rule]; o < frico_dunavant_orbit_begin[rule + 1]; o++) {
        const auto& d = frico_dunavant_orbits[o];
        const double a = d.a, b = d.b, c = d.c, w = d.w * area;
        if (d.orbit == 1) {
            out.push_back({at(a, b, c), w});
        } else if (d.orbit == 3) {
            out.push_back({at(a, b, c), w});
            out.push_back({at(b, c, a), w});
            out.push_back({at(b, a, c), w});
        } else {
            out.push_back({at(a, b, c), w});
            out.push_back({at(a, c, b), w});
            out.push_back({at(b, a, c), w});
            out.push_back({at(b, c, a), w});
            out.push_back({at(c, a, b), w});
            out.push_back({at(c, b, a), w});
        }
    }
}

// Barycentric coordinates of the rule points, in the expansion order of dunavant().
inline void dunavant_bary(int degree, std::vector<std::array<double, 3>>& out)
{
    const int rule = degree == 0 ? 0 : degree - 1;
    for (int o = frico_dunavant_orbit_begin[rule]; o < frico_dunavant_orbit_begin[rule + 1]; o++) {
        const auto& d = frico_dunavant_orbits[o];
        const double a = d.a, b = d.b, c = d.c;
        if (d.orbit == 1) {
            out.push_back({a, b, c});
        } else if (d.orbit == 3) {
            out.push_back({a, b, c});
            out.push_back({b, c, a});
            out.push_back({b, a, c});
        } else {
            out.push_back({a, b, c});
            out.push_back({a, c, b});
            out.push_back({b, a, c});
            out.push_back({b, c, a});
            out.push_back({c, a, b});
            out.push_back({c, b, a});
        }
    }
}

struct Neighbours {
    int64_t tm = -1, tp = -1;
    int lm = -1, lp = -1;
    bool has_iface = false;
    int iface = 0;
};

} // namespace geom
using namespace geom;

struct HostMesh {
    int64_t nv = 0, nt = 0;
    std::vector<vec3> V;
    std::vector<std::array<int64_t, 3>> T;
    std::vector<int> tag;
    std::vector<std::pair<int64_t, int64_t>> edges;
    std::vector<frico_rwg> bf;
    int degree = -1, q = 0;
    // device-side arrays prepared on the host
    std::vector<double> cen, area, sqrtA, alpha, tri_sgn, rwg_len, rwg_Am, rwg_Ap;
    std::vector<double4> qreg, qsub;
    std::vector<int> tri_rwg, rwg_t, rwg_le;
    std::vector<int2> near;
    // adjoint: triangle vertices and barycentric weights of the rule points
    std::vector<int> tri_v;
    std::vector<double> qbeta, qsub_beta;
};

// Resolves the device address of frico_fields_noalloc's scratch on `device` (current), once per
// device (noalloc.hip).
int noalloc_init(int device);

} // namespace frico

using namespace frico;

struct frico_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    hipStream_t side = nullptr;      // look-ahead stream of the LU (default priority)
    std::vector<hipEvent_t> lu_events;
    bool lookahead = true; // FRICO_LU_LOOKAHEAD=0 disables the look-ahead stream
    int nb = 0; // LU block width; 0: chosen from N at each factorisation
    int max_groups = 0;
    bool profiling = false;
    bool have_mesh = false;
    HostMesh hm;
    int64_t N = 0;
    // device mesh
    DevBuf<double> d_cen, d_area, d_sqrtA, d_alpha, d_tri_sgn, d_rwg_len, d_rwg_Am, d_rwg_Ap;
    DevBuf<double4> d_qreg, d_qsub;
    DevBuf<int> d_tri_rwg, d_rwg_t, d_rwg_le;
    DevBuf<int2> d_near;
    DevMesh dm{};
    // matrix state
    DevBuf<c128> Z64;
    DevBuf<c64> Z32, LU32;
    DevBuf<unsigned char> Cbuf;
    int zprec = -1;        // precision of the current Z (-1: none)
    bool z64_factored = false;
    int factored = -1;     // form of the current factorisation (-1: none)
    int nb_factored = 64;  // block width the factors were computed with
    DevBuf<int> ipiv;
    // lu workspace
    DevBuf<unsigned long long> llw; // panel exchange area
    DevBuf<int> flags;
    DevBuf<unsigned char> dinv;
    DevBuf<int> perm;             // row-interchange lists of the last factorisation
    DevBuf<unsigned char> wv;     // W products of the last factorisation (fused solves)
    DevBuf<unsigned char> ysolve; // n x nrhs scratch of the fused solves
    DevBuf<c64> tbuf;             // transposed block column of the unpivoted LU
    DevBuf<unsigned char> hsplit; // fp16 split operands of the unpivoted LU's trailing products
    // Split products of the complex64 solves (FRICO_SOLVE_F16=0: fp32 products): the factors' row
    // exponents over each block of lu_solve_block_rows() columns (valid after an unpivoted or
    // symmetric complex64 factorisation), and two scratch vectors of the solution's column exponents
    DevBuf<int> sexp, bexp;
    // exchange area of the chained complex64 solves (one to four right-hand sides)
    DevBuf<unsigned long long> chain;
    bool sexp_ok = false;
    bool solve_f16 = true;
    bool gemm_f16 = true;         // FRICO_GEMM_F16=0: fp32 trailing products in the unpivoted LU
    bool fused_step = true;       // FRICO_LU_FUSED_STEP=0: the unpivoted LU's block solves and fp16 splits as separate launches
    bool direct_solves = false;   // FRICO_LU_DIRECT_SOLVES=1: lane-per-row block solves in the unpivoted LU
    bool wv_ok = false;           // wv holds the products of the current factors
    bool fused_solve = true;
    // Right-hand sides carried through the complex64 factorisation (FRICO_LU_AUGMENT=0 disables):
    // LU32 then holds aug_cols more columns, L^{-1} P B for the B of rhs_version aug_version.
    bool augment = true;
    // Refinement residuals updated beside the backward substitution on the second stream
    // (FRICO_STREAM_RESIDUAL=0: one full residual product after each solve).
    bool stream_residual = true;
    bool resid_pair = true; // two residual blocks a launch, the second into R2 (FRICO_RESID_PAIR=0: one)
    // Residual products of the current solve in the exact four-product form: taken when the
    // tolerance is within 1e3 of the three-product form's rounding floor (~1e-13 relative, where a
    // refinement to 1e-13 stalled), i.e. below 1e-10.
    bool resid_exact = true;
    int64_t aug_cols = 0;
    bool aug_valid = false;
    uint64_t rhs_version = 0, aug_version = 0;
    // rhs / solution
    int64_t nrhs = 0;
    bool have_x = false;
    DevBuf<c128> B, X, R, R2;
    DevBuf<c64> D32;
    DevBuf<double> norms;
    // fields
    DevBuf<c128> TS, EH;
    DevBuf<double> pts;
    DevBuf<int> pcols;
    DevBuf<PlaneWaveDev> waves;
    // profiling
    KernelTimer t_fill, t_rhs, t_factor, t_solve, t_refine, t_gemm, t_tripair;
    // adjoint
    DevBuf<int> d_tri_v;
    DevBuf<double> d_qbeta, d_qsub_beta, d_gvert;
    // fill mode of the adjoint pairs (FRICO_FILL_QUAD or FRICO_FILL_APPROX) and the approximate
    // mode's one-point rule workspaces
    int adj_fill = FRICO_FILL_QUAD;
    DevBuf<double4> adj_cent;
    DevBuf<double> adj_third;
    DevBuf<c128> adjW, aw, bw;
    AdjMesh am{};
    struct AdjointState {
        bool forward_done = false, grads_done = false;
        bool approx = false; // the forward fill was the approximate mode's
        double freq = 0;
        std::vector<int32_t> ifaces;
        std::complex<double> Vg = 0, Iport = 0, zin = 0;
        double obs[3] = {0, 0, 0};
        std::vector<double> signs;
        std::vector<std::complex<double>> I, V, c, E, H;
        double absS = 0, P = 0, gain = 0;
        std::vector<std::complex<double>> dzin; // 3 nv
        std::vector<double> dgain;              // 3 nv
        // RCS pair (frico_adjoint_rcs_*): the plane wave, the observation and the two components
        bool rcs_done = false;
        double theta_i = 0, phi_i = 0, theta_s = 0, phi_s = 0, R = 0;
        std::complex<double> eth = 0, eph = 0, a_vv = 0, a_hh = 0;
    } adj;
};

