// Impedance-matrix fill, right-hand sides and field sums for gfx950.
//
// Quadrature-mode fill (reference: compute_matrix, src/emw_solver.cpp:45-142) is split into
// three kernels that together produce Z(j, i) for every RWG pair:
//
//   tripair_*   one thread per ordered triangle pair (P, Q).  It runs the regular x regular
//               Dunavant double sum of exp(-jkR)/R once per triangle pair and reduces it to
//               eight moments (S0, Su, Sv, Suv, with points taken relative to the centroids);
//               from them it forms the nine "local-edge" contributions
//               C[P][m][Q][n] = sum w w (0.25 (r-f_m).(r'-f_n) - 1/k^2) g / (A_P A_Q)
//               for every choice of free vertices f_m of P and f_n of Q.
//   assemble_*  one thread per Z entry: Z(j,i) = l_i l_j/(4 pi) (C-- - C-+ - C+- + C++),
//               gathering the four triangle pairs of the two RWG functions.
//   near_*      one wavefront per RWG pair whose row function touches a triangle of the
//               column function (at most 5 per column).  Those entries are recomputed in
//               FP64 exactly as the reference does, with the sub-triangle rule on the row
//               function's coincident triangles (src/emw_solver.cpp:67-78), and overwrite
//               what the assembly wrote.
//
// Each RWG pair's four triangle pairs are therefore evaluated once per *triangle* pair
// instead of once per RWG pair, which removes ~N^2/T^2 * 4 of the Green's-function work.
#include "common.hpp"
#include "mesh_dev.hpp"

namespace frico {

// ----------------------------------------------------------------------------------------------
// Triangle-pair moments, FP64.
// ----------------------------------------------------------------------------------------------
constexpr int TP_TILE = 16; // P triangles per workgroup
constexpr int TQ_TILE = 16; // Q triangles per workgroup

template <int QC> struct QN {
    __device__ static int get(int q) { return QC; }
};
template <> struct QN<0> {
    __device__ static int get(int q) { return q; }
};

// (sin x, cos x) for the small per-point phase of a triangle pair: Taylor series through x^17 and
// x^18 for |x| <= 1 (truncation below 1e-17), the library sincos otherwise.
__device__ inline void small_sincos(double x, double& sn, double& cs)
{
    if (fabs(x) <= 1.0) {
        const double x2 = x * x;
        double s = 1.0 / 355687428096000.0;        // 1/17!
        s = fma(s, -x2, 1.0 / 1307674368000.0);     // 1/15!
        s = fma(s, -x2, 1.0 / 6227020800.0);        // 1/13!
        s = fma(s, -x2, 1.0 / 39916800.0);          // 1/11!
        s = fma(s, -x2, 1.0 / 362880.0);            // 1/9!
        s = fma(s, -x2, 1.0 / 5040.0);              // 1/7!
        s = fma(s, -x2, 1.0 / 120.0);               // 1/5!
        s = fma(s, -x2, 1.0 / 6.0);                 // 1/3!
        s = fma(s, -x2, 1.0);
        double c = 1.0 / 6402373705728000.0;       // 1/18!
        c = fma(c, -x2, 1.0 / 20922789888000.0);    // 1/16!
        c = fma(c, -x2, 1.0 / 87178291200.0);       // 1/14!
        c = fma(c, -x2, 1.0 / 479001600.0);         // 1/12!
        c = fma(c, -x2, 1.0 / 3628800.0);           // 1/10!
        c = fma(c, -x2, 1.0 / 40320.0);             // 1/8!
        c = fma(c, -x2, 1.0 / 720.0);               // 1/6!
        c = fma(c, -x2, 1.0 / 24.0);                // 1/4!
        c = fma(c, -x2, 0.5);                       // 1/2!
        sn = x * s;
        cs = fma(c, -x2, 1.0);
    } else {
        sincos(x, &sn, &cs);
    }
}

template <int QC, typename OutT>
__global__ __launch_bounds__(256) void tripair_f64(DevMesh m, double k, double inv_ksq, cx<OutT>* __restrict__ C)
{
    constexpr int QMAX = QC ? QC : 79;
    extern __shared__ double4 smem_tp[];
    const int q = QN<QC>::get(m.q);
    double4* sP = smem_tp;           // [TP_TILE][q]
    double4* sQ = smem_tp + TP_TILE * q; // [TQ_TILE][q]
    // The moments are symmetric, C(P, m, Q, n) = C(Q, n, P, m) (swapping the triangles swaps the
    // roles of Su and Sv): tiles below the diagonal are written by their mirror tile.
    if (blockIdx.y > blockIdx.x)
        return;
    const bool mirror = blockIdx.y < blockIdx.x;
    const int P0 = blockIdx.y * TP_TILE, Q0 = blockIdx.x * TQ_TILE;
    for (int i = threadIdx.x; i < TP_TILE * q; i += blockDim.x) {
        int t = P0 + i / q;
        sP[i] = t < m.T ? m.qreg[(size_t)t * q + i % q] : double4{0, 0, 0, 0};
        int s = Q0 + i / q;
        sQ[i] = s < m.T ? m.qreg[(size_t)s * q + i % q] : double4{0, 0, 0, 0};
    }
    __syncthreads();
    const int tq = threadIdx.x % TQ_TILE, tp = threadIdx.x / TQ_TILE;
    const int P = P0 + tp, Q = Q0 + tq;
    if (P >= m.T || Q >= m.T)
        return;
    const double dx = m.cen[3 * P] - m.cen[3 * Q];
    const double dy = m.cen[3 * P + 1] - m.cen[3 * Q + 1];
    const double dz = m.cen[3 * P + 2] - m.cen[3 * Q + 2];
    // exp(-jkR) = exp(-jk R0) exp(-jk (R - R0)) with R0 the centroid distance: the per-point factor
    // has a small argument, the pair factor is applied once to the moments.
    const double R0 = sqrt(dx * dx + dy * dy + dz * dz);

    // Q's points in registers when the rule size is a compile-time constant.
    double4 vq[QC ? QC : 1];
    if constexpr (QC > 0) {
#pragma unroll
        for (int r = 0; r < QC; r++)
            vq[r] = sQ[tq * QC + r];
    }
    double s0r = 0, s0i = 0;
    double sur[3] = {0, 0, 0}, sui[3] = {0, 0, 0};
    double svr[3] = {0, 0, 0}, svi[3] = {0, 0, 0};
    double suvr = 0, suvi = 0;
    (void)QMAX;
#pragma unroll 1
    for (int p = 0; p < q; p++) {
        const double4 u = sP[tp * q + p];
        const double ex = dx + u.x, ey = dy + u.y, ez = dz + u.z;
        double hr = 0, hi = 0, vxr = 0, vxi = 0, vyr = 0, vyi = 0, vzr = 0, vzi = 0;
#pragma unroll(QC ? QC : 1)
        for (int r = 0; r < q; r++) {
            double4 v;
            if constexpr (QC > 0)
                v = vq[r];
            else
                v = sQ[tq * q + r];
            const double rx = ex - v.x, ry = ey - v.y, rz = ez - v.z;
            const double R2 = rx * rx + ry * ry + rz * rz;
            const double rinv = R2 > 0 ? rsqrt(R2) : 0.0;
            const double R = R2 * rinv;
            double sn, cs;
            small_sincos(k * (R - R0), sn, cs);
            const double f = v.w * rinv;
            const double gr = cs * f, gi = -sn * f;
            hr += gr;
            hi += gi;
            vxr += gr * v.x;
            vxi += gi * v.x;
            vyr += gr * v.y;
            vyi += gi * v.y;
            vzr += gr * v.z;
            vzi += gi * v.z;
        }
        const double w = u.w;
        s0r += w * hr;
        s0i += w * hi;
        sur[0] += w * u.x * hr;
        sui[0] += w * u.x * hi;
        sur[1] += w * u.y * hr;
        sui[1] += w * u.y * hi;
        sur[2] += w * u.z * hr;
        sui[2] += w * u.z * hi;
        svr[0] += w * vxr;
        svi[0] += w * vxi;
        svr[1] += w * vyr;
        svi[1] += w * vyi;
        svr[2] += w * vzr;
        svi[2] += w * vzi;
        suvr += w * (u.x * vxr + u.y * vyr + u.z * vzr);
        suvi += w * (u.x * vxi + u.y * vyi + u.z * vzi);
    }
    {
        // rotate the moments by the pair phase exp(-jk R0)
        double ps, pc;
        sincos(k * R0, &ps, &pc);
        auto rot = [&](double& a, double& b) {
            const double t = a * pc + b * ps;
            b = b * pc - a * ps;
            a = t;
        };
        rot(s0r, s0i);
        rot(suvr, suvi);
        for (int c = 0; c < 3; c++) {
            rot(sur[c], sui[c]);
            rot(svr[c], svi[c]);
        }
    }
    const double invAA = 1.0 / (m.area[P] * m.area[Q]);
    const double* aP = m.alpha + 9 * (size_t)P;
    const double* aQ = m.alpha + 9 * (size_t)Q;
#pragma unroll
    for (int mm = 0; mm < 3; mm++) {
        const double ax = aP[3 * mm], ay = aP[3 * mm + 1], az = aP[3 * mm + 2];
        // alpha_m . Sv
        const double avr = ax * svr[0] + ay * svr[1] + az * svr[2];
        const double avi = ax * svi[0] + ay * svi[1] + az * svi[2];
        cx<OutT>* out = C + ((size_t)(3 * P + mm) * m.T + Q) * 3;
#pragma unroll
        for (int nn = 0; nn < 3; nn++) {
            const double bx = aQ[3 * nn], by = aQ[3 * nn + 1], bz = aQ[3 * nn + 2];
            const double bur = bx * sur[0] + by * sur[1] + bz * sur[2];
            const double bui = bx * sui[0] + by * sui[1] + bz * sui[2];
            const double ab = ax * bx + ay * by + az * bz;
            const double cr = (0.25 * (suvr + avr + bur + ab * s0r) - inv_ksq * s0r) * invAA;
            const double ci = (0.25 * (suvi + avi + bui + ab * s0i) - inv_ksq * s0i) * invAA;
            out[nn] = cx<OutT>{(OutT)cr, (OutT)ci};
            if (mirror)
                C[((size_t)(3 * Q + nn) * m.T + P) * 3 + mm] = cx<OutT>{(OutT)cr, (OutT)ci};
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Triangle-pair moments, FP32 bulk with an FP64 pair phase.
// exp(-jkR) = exp(-jk R0) exp(-jk (R - R0)), R0 = |c_P - c_Q| in FP64; the per-point factor has
// a small argument and is evaluated in FP32, the pair factor once in FP64.
// ----------------------------------------------------------------------------------------------
template <int QC>
__global__ __launch_bounds__(256) void tripair_f32(DevMesh m, double k, double inv_ksq, c64* __restrict__ C)
{
    extern __shared__ float4 smem_tp32[];
    const int q = QN<QC>::get(m.q);
    float4* sP = smem_tp32;
    float4* sQ = smem_tp32 + TP_TILE * q;
    // symmetric moments: tiles below the diagonal are written by their mirror tile
    if (blockIdx.y > blockIdx.x)
        return;
    const bool mirror = blockIdx.y < blockIdx.x;
    const int P0 = blockIdx.y * TP_TILE, Q0 = blockIdx.x * TQ_TILE;
    for (int i = threadIdx.x; i < TP_TILE * q; i += blockDim.x) {
        int t = P0 + i / q;
        double4 a = t < m.T ? m.qreg[(size_t)t * q + i % q] : double4{0, 0, 0, 0};
        sP[i] = float4{(float)a.x, (float)a.y, (float)a.z, (float)a.w};
        int s = Q0 + i / q;
        double4 b = s < m.T ? m.qreg[(size_t)s * q + i % q] : double4{0, 0, 0, 0};
        sQ[i] = float4{(float)b.x, (float)b.y, (float)b.z, (float)b.w};
    }
    __syncthreads();
    const int tq = threadIdx.x % TQ_TILE, tp = threadIdx.x / TQ_TILE;
    const int P = P0 + tp, Q = Q0 + tq;
    if (P >= m.T || Q >= m.T)
        return;
    const double dxd = m.cen[3 * P] - m.cen[3 * Q];
    const double dyd = m.cen[3 * P + 1] - m.cen[3 * Q + 1];
    const double dzd = m.cen[3 * P + 2] - m.cen[3 * Q + 2];
    const double R0d = sqrt(dxd * dxd + dyd * dyd + dzd * dzd);
    double ps, pc;
    sincos(k * R0d, &ps, &pc);
    const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd, R0 = (float)R0d;
    const float R0sq = (float)(R0d * R0d);
    const float kf = (float)k;

    float4 vq[QC ? QC : 1];
    if constexpr (QC > 0) {
#pragma unroll
        for (int r = 0; r < QC; r++)
            vq[r] = sQ[tq * QC + r];
    }
    float s0r = 0, s0i = 0, sur[3] = {0, 0, 0}, sui[3] = {0, 0, 0}, svr[3] = {0, 0, 0},
          svi[3] = {0, 0, 0}, suvr = 0, suvi = 0;
#pragma unroll 1
    for (int p = 0; p < q; p++) {
        const float4 u = sP[tp * q + p];
        float hr = 0, hi = 0, vxr = 0, vxi = 0, vyr = 0, vyi = 0, vzr = 0, vzi = 0;
#pragma unroll(QC ? QC : 1)
        for (int r = 0; r < q; r++) {
            float4 v;
            if constexpr (QC > 0)
                v = vq[r];
            else
                v = sQ[tq * q + r];
            // delta = u - v; R^2 = R0^2 + 2 d.delta + |delta|^2
            const float ex = u.x - v.x, ey = u.y - v.y, ez = u.z - v.z;
            const float num = 2.0f * (dx * ex + dy * ey + dz * ez) + (ex * ex + ey * ey + ez * ez);
            const float R2 = R0sq + num;
            const float rinv = R2 > 0.0f ? rsqrtf(R2) : 0.0f;
            const float R = R2 * rinv;
            const float dR = R2 > 0.0f ? num / (R + R0) : 0.0f;
            float sn, cs;
            __sincosf(kf * dR, &sn, &cs);
            const float f = v.w * rinv;
            const float gr = cs * f, gi = -sn * f;
            hr += gr;
            hi += gi;
            vxr += gr * v.x;
            vxi += gi * v.x;
            vyr += gr * v.y;
            vyi += gi * v.y;
            vzr += gr * v.z;
            vzi += gi * v.z;
        }
        const float w = u.w;
        s0r += w * hr;
        s0i += w * hi;
        sur[0] += w * u.x * hr;
        sui[0] += w * u.x * hi;
        sur[1] += w * u.y * hr;
        sui[1] += w * u.y * hi;
        sur[2] += w * u.z * hr;
        sui[2] += w * u.z * hi;
        svr[0] += w * vxr;
        svi[0] += w * vxi;
        svr[1] += w * vyr;
        svi[1] += w * vyi;
        svr[2] += w * vzr;
        svi[2] += w * vzi;
        suvr += w * (u.x * vxr + u.y * vyr + u.z * vzr);
        suvi += w * (u.x * vxi + u.y * vyi + u.z * vzi);
    }
    // rotate by the pair phase exp(-j k R0) = (pc, -ps)
    const float er = (float)pc, ei = (float)(-ps);
    auto rot = [&](float& a, float& b) {
        float t = a * er - b * ei;
        b = a * ei + b * er;
        a = t;
    };
    rot(s0r, s0i);
    rot(suvr, suvi);
    for (int c = 0; c < 3; c++) {
        rot(sur[c], sui[c]);
        rot(svr[c], svi[c]);
    }
    const float invAA = (float)(1.0 / (m.area[P] * m.area[Q]));
    const float ikk = (float)inv_ksq;
    const double* aP = m.alpha + 9 * (size_t)P;
    const double* aQ = m.alpha + 9 * (size_t)Q;
#pragma unroll
    for (int mm = 0; mm < 3; mm++) {
        const float ax = (float)aP[3 * mm], ay = (float)aP[3 * mm + 1], az = (float)aP[3 * mm + 2];
        const float avr = ax * svr[0] + ay * svr[1] + az * svr[2];
        const float avi = ax * svi[0] + ay * svi[1] + az * svi[2];
        c64* out = C + ((size_t)(3 * P + mm) * m.T + Q) * 3;
#pragma unroll
        for (int nn = 0; nn < 3; nn++) {
            const float bx = (float)aQ[3 * nn], by = (float)aQ[3 * nn + 1], bz = (float)aQ[3 * nn + 2];
            const float bur = bx * sur[0] + by * sur[1] + bz * sur[2];
            const float bui = bx * sui[0] + by * sui[1] + bz * sui[2];
            const float ab = ax * bx + ay * by + az * bz;
            const c64 v{(0.25f * (suvr + avr + bur + ab * s0r) - ikk * s0r) * invAA,
                        (0.25f * (suvi + avi + bui + ab * s0i) - ikk * s0i) * invAA};
            out[nn] = v;
            if (mirror)
                C[((size_t)(3 * Q + nn) * m.T + P) * 3 + mm] = v;
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Assembly: Z(j,i) = prodl * (C[i-][j-] - C[i-][j+] - C[i+][j-] + C[i+][j+]).
// One workgroup per column i, sweeping the rows j (contiguous in the column-major Z): the two
// C rows of column i are then read by one CU, whose L2 serves both the j- and the j+ entries
// (nearby in C for nearby j), instead of by workgroups on every XCD.  Two rows per thread per
// pass keep eight gathers in flight.
// ----------------------------------------------------------------------------------------------
template <typename TC, typename TZ>
__global__ __launch_bounds__(256) void assemble(DevMesh m, const cx<TC>* __restrict__ C, cx<TZ>* __restrict__ Z,
                                                int64_t ldz, int64_t col0, int64_t ncols)
{
    const int64_t ic = blockIdx.x;
    if (ic >= ncols)
        return;
    const int64_t i = col0 + ic;
    const int im = m.rwg_t[2 * i], ip = m.rwg_t[2 * i + 1];
    const int lim = m.rwg_le[2 * i], lip = m.rwg_le[2 * i + 1];
    const cx<TC>* rowm = C + (size_t)(3 * im + lim) * m.T * 3;
    const cx<TC>* rowp = C + (size_t)(3 * ip + lip) * m.T * 3;
    const double li = 0.25 * m.rwg_len[i];
    auto entry = [&](int64_t j, cx<TC> a, cx<TC> b, cx<TC> c, cx<TC> d) {
        const double prodl = li * m.rwg_len[j] * INV_PI;
        const TC pl = (TC)prodl;
        TC re = ((a.re - b.re) - c.re) + d.re;
        TC imv = ((a.im - b.im) - c.im) + d.im;
        Z[j + i * ldz] = cx<TZ>{(TZ)(pl * re), (TZ)(pl * imv)};
    };
    for (int64_t j0 = threadIdx.x; j0 < m.N; j0 += 512) {
        const int64_t j1 = j0 + 256;
        const bool has1 = j1 < m.N;
        const int64_t jj1 = has1 ? j1 : j0;
        const int jm0 = m.rwg_t[2 * j0], jp0 = m.rwg_t[2 * j0 + 1];
        const int ljm0 = m.rwg_le[2 * j0], ljp0 = m.rwg_le[2 * j0 + 1];
        const int jm1 = m.rwg_t[2 * jj1], jp1 = m.rwg_t[2 * jj1 + 1];
        const int ljm1 = m.rwg_le[2 * jj1], ljp1 = m.rwg_le[2 * jj1 + 1];
        const cx<TC> a0 = rowm[3 * jm0 + ljm0], b0 = rowm[3 * jp0 + ljp0], c0 = rowp[3 * jm0 + ljm0],
                     d0 = rowp[3 * jp0 + ljp0];
        const cx<TC> a1 = rowm[3 * jm1 + ljm1], b1 = rowm[3 * jp1 + ljp1], c1 = rowp[3 * jm1 + ljm1],
                     d1 = rowp[3 * jp1 + ljp1];
        entry(j0, a0, b0, c0, d0);
        if (has1)
            entry(j1, a1, b1, c1, d1);
    }
}

// ----------------------------------------------------------------------------------------------
// Near pairs, FP64, the reference's per-entry formula with its sub-triangle rule.
// One 64-lane wavefront per (i, j) listed in pairs[].
// ----------------------------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

template <typename TZ>
__global__ __launch_bounds__(64) void near_f64(DevMesh m, double k, double inv_ksq, const int2* __restrict__ pairs,
                                               int64_t npairs, cx<TZ>* __restrict__ Z, int64_t ldz)
{
    const int64_t pidx = blockIdx.x;
    if (pidx >= npairs)
        return;
    const int i = pairs[pidx].x, j = pairs[pidx].y;
    const int lane = threadIdx.x;
    const int ti[2] = {m.rwg_t[2 * i], m.rwg_t[2 * i + 1]};
    const int li[2] = {m.rwg_le[2 * i], m.rwg_le[2 * i + 1]};
    const int tj[2] = {m.rwg_t[2 * j], m.rwg_t[2 * j + 1]};
    const int lj[2] = {m.rwg_le[2 * j], m.rwg_le[2 * j + 1]};
    double er = 0, ei = 0;
    for (int a = 0; a < 2; a++) {
        const int P = ti[a];
        const double* cP = m.cen + 3 * (size_t)P;
        const double* al = m.alpha + 9 * (size_t)P + 3 * li[a];
        for (int b = 0; b < 2; b++) {
            const int Q = tj[b];
            const bool split = (Q == ti[0]) || (Q == ti[1]);
            const double4* qq = split ? m.qsub + (size_t)Q * 3 * m.q : m.qreg + (size_t)Q * m.q;
            const int nq = split ? 3 * m.q : m.q;
            const double* cQ = m.cen + 3 * (size_t)Q;
            const double* be = m.alpha + 9 * (size_t)Q + 3 * lj[b];
            const double invAA = 1.0 / (m.area[P] * m.area[Q]);
            const double sg = (a == b) ? 1.0 : -1.0;
            const int total = m.q * nq;
            double br = 0, bi = 0;
            for (int t = lane; t < total; t += 64) {
                const int p = t / nq, r = t % nq;
                const double4 u = m.qreg[(size_t)P * m.q + p];
                const double4 v = qq[r];
                // rho_i = u + alpha, rho_j = v + beta; R = |(cP + u) - (cQ + v)|
                const double rpx = cP[0] + u.x, rpy = cP[1] + u.y, rpz = cP[2] + u.z;
                const double rqx = cQ[0] + v.x, rqy = cQ[1] + v.y, rqz = cQ[2] + v.z;
                const double vv = 0.25 * ((u.x + al[0]) * (v.x + be[0]) + (u.y + al[1]) * (v.y + be[1]) +
                                          (u.z + al[2]) * (v.z + be[2]));
                const double w = u.w * v.w * invAA;
                const double dx = rpx - rqx, dy = rpy - rqy, dz = rpz - rqz;
                const double R = sqrt(dx * dx + dy * dy + dz * dz);
                double sn, cs;
                sincos(k * R, &sn, &cs);
                const double f = w * (vv - inv_ksq) / R;
                br += f * cs;
                bi -= f * sn;
            }
            er += sg * br;
            ei += sg * bi;
        }
    }
    er = wave_sum(er);
    ei = wave_sum(ei);
    if (lane == 0) {
        const double prodl = 0.25 * m.rwg_len[i] * m.rwg_len[j] * INV_PI;
        Z[(int64_t)j + (int64_t)i * ldz] = cx<TZ>{(TZ)(prodl * er), (TZ)(prodl * ei)};
    }
}

// ----------------------------------------------------------------------------------------------
// Approximate (MOM3D centroid) fill: src/emw_solver.cpp:144-220.  One thread per entry.
// ----------------------------------------------------------------------------------------------
template <typename TZ>
__global__ __launch_bounds__(256) void fill_approx(DevMesh m, double k, double inv_ksq, cx<TZ>* __restrict__ Z,
                                                   int64_t ldz, int64_t col0)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = col0 + blockIdx.y;
    if (j >= m.N || i >= m.N)
        return;
    const double sqrt_4pi = sqrt(4 * PI);
    double er = 0, ei = 0;
    for (int a = 0; a < 2; a++) {
        const int P = m.rwg_t[2 * i + a];
        const double* cP = m.cen + 3 * (size_t)P;
        const double* al = m.alpha + 9 * (size_t)P + 3 * m.rwg_le[2 * i + a];
        for (int b = 0; b < 2; b++) {
            const int Q = m.rwg_t[2 * j + b];
            const double* cQ = m.cen + 3 * (size_t)Q;
            const double* be = m.alpha + 9 * (size_t)Q + 3 * m.rwg_le[2 * j + b];
            const double rho_c = al[0] * be[0] + al[1] * be[1] + al[2] * be[2];
            double gr, gi;
            if (P == Q) {
                // 0.25 * (sqrt(4 pi)/sqrt(A_P) - jk) / pi
                gr = 0.25 * (sqrt_4pi / m.sqrtA[P]) * INV_PI;
                gi = -0.25 * k * INV_PI;
            } else {
                const double dx = cP[0] - cQ[0], dy = cP[1] - cQ[1], dz = cP[2] - cQ[2];
                const double R = sqrt(dx * dx + dy * dy + dz * dz);
                double sn, cs;
                sincos(k * R, &sn, &cs);
                const double f = 1.0 / (4.0 * PI * R);
                gr = cs * f;
                gi = -sn * f;
            }
            const double s = 0.25 * rho_c - inv_ksq;
            const double sg = (a == b) ? 1.0 : -1.0;
            er += sg * s * gr;
            ei += sg * s * gi;
        }
    }
    const double ll = m.rwg_len[i] * m.rwg_len[j];
    Z[j + i * ldz] = cx<TZ>{(TZ)(ll * er), (TZ)(ll * ei)};
}

// ----------------------------------------------------------------------------------------------
// Plane-wave right-hand sides, one column per incidence: src/emw_solver.cpp:380-416 with
// plane_wave::compute (src/sources.h:67-74).  B[n + a*ldb] = V_n for incidence a.
// ----------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rhs_planewave(DevMesh m, double k, double omega_mu0, const PlaneWaveDev* __restrict__ waves,
                                                     int nang, c128* __restrict__ B, int64_t ldb)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (n >= m.N || a >= nang)
        return;
    const PlaneWaveDev w = waves[a];
    double vr = 0, vi = 0;
    for (int s = 0; s < 2; s++) {
        const int t = m.rwg_t[2 * n + s];
        const double* c = m.cen + 3 * (size_t)t;
        const double* al = m.alpha + 9 * (size_t)t + 3 * m.rwg_le[2 * n + s];
        const double A = s == 0 ? m.rwg_Am[n] : m.rwg_Ap[n];
        const double wf = 0.5 * m.rwg_len[n] / A;
        double sr = 0, si = 0;
        for (int p = 0; p < m.q; p++) {
            const double4 u = m.qreg[(size_t)t * m.q + p];
            const double rx = c[0] + u.x, ry = c[1] + u.y, rz = c[2] + u.z;
            double sn, cs;
            sincos(k * (w.dir[0] * rx + w.dir[1] * ry + w.dir[2] * rz), &sn, &cs);
            // rho = r - free vertex = u + alpha
            const double px = u.x + al[0], py = u.y + al[1], pz = u.z + al[2];
            const double dr = px * w.e0r[0] + py * w.e0r[1] + pz * w.e0r[2];
            const double di = px * w.e0i[0] + py * w.e0i[1] + pz * w.e0i[2];
            // (dr + j di) * (cs + j sn)
            const double f = wf * u.w;
            sr += f * (dr * cs - di * sn);
            si += f * (dr * sn + di * cs);
        }
        if (s == 0) {
            vr -= sr;
            vi -= si;
        } else {
            vr += sr;
            vi += si;
        }
    }
    // v / (0, -omega mu0) = j v / (omega mu0)
    B[n + a * ldb] = c128{-vi / omega_mu0, vr / omega_mu0};
}

// ----------------------------------------------------------------------------------------------
// RWG -> triangle scatter (src/emw_solver.h:137-164) for nrhs current vectors.
// TS[(rhs * T + t)] = {AJx, AJy, AJz, AdivJ}.  Gathered per triangle, no atomics.
// ----------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void scatter_currents(DevMesh m, const c128* __restrict__ I, int64_t ldi,
                                                        int nrhs, c128* __restrict__ TS)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (t >= m.T || r >= nrhs)
        return;
    double ajr[3] = {0, 0, 0}, aji[3] = {0, 0, 0}, adr = 0, adi = 0;
    const double* c = m.cen + 3 * t;
    for (int e = 0; e < 3; e++) {
        const int n = m.tri_rwg[3 * t + e];
        if (n < 0)
            continue;
        const double sg = m.tri_sgn[3 * t + e];
        const c128 In = I[n + r * ldi];
        const double l = m.rwg_len[n];
        // A f(c) = sign * l * (c - free vertex) / 2 = sign * l * alpha / 2
        const double* al = m.alpha + 9 * t + 3 * e;
        for (int d = 0; d < 3; d++) {
            const double f = sg * l * al[d] * 0.5;
            ajr[d] += f * In.re;
            aji[d] += f * In.im;
        }
        adr += sg * l * In.re;
        adi += sg * l * In.im;
    }
    c128* out = TS + ((size_t)r * m.T + t) * 4;
    out[0] = c128{ajr[0], aji[0]};
    out[1] = c128{ajr[1], aji[1]};
    out[2] = c128{ajr[2], aji[2]};
    out[3] = c128{adr, adi};
}

// ----------------------------------------------------------------------------------------------
// Field sums E, H at a point (src/emw_postpro_common.cpp:31-66): one workgroup per point, which
// uses current vector cols[point] (column 0 when cols is null); threads stride over triangles,
// block reduction of 12 doubles.  H keeps the reference's exp(+jkR) factor unless h_sign = -1.
// A point whose column is outside [0, ncols) gets NaN fields (the callers that take device-side
// column lists cannot check them on the host).  E and H: npts x 3 complex, either may be null.
// ----------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void eval_fields(DevMesh m, double k, double omega_mu0, const c128* __restrict__ TS,
                                                   int ncols, const double* __restrict__ pts,
                                                   const int* __restrict__ cols, int npts, double h_sign,
                                                   c128* __restrict__ E, c128* __restrict__ H)
{
    const int ip = blockIdx.x;
    if (ip >= npts)
        return;
    const int r = cols ? cols[ip] : 0;
    if (r < 0 || r >= ncols) {
        if (threadIdx.x < 3) {
            const double nan = __builtin_nan("");
            if (E)
                E[3 * (size_t)ip + threadIdx.x] = c128{nan, nan};
            if (H)
                H[3 * (size_t)ip + threadIdx.x] = c128{nan, nan};
        }
        return;
    }
    const double px = pts[3 * ip], py = pts[3 * ip + 1], pz = pts[3 * ip + 2];
    double acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const c128* ts = TS + (size_t)r * m.T * 4;
    for (int t = threadIdx.x; t < m.T; t += blockDim.x) {
        const double vx = px - m.cen[3 * t], vy = py - m.cen[3 * t + 1], vz = pz - m.cen[3 * t + 2];
        const double R2 = vx * vx + vy * vy + vz * vz;
        const double R = sqrt(R2);
        const double kR = k * R;
        double sn, cs;
        sincos(kR, &sn, &cs);
        const double inv4piR = 1.0 / (4 * PI * R);
        // g = exp(-jkR)/(4 pi R)
        const double gr = cs * inv4piR, gi = -sn * inv4piR;
        const c128 J0 = ts[4 * t], J1 = ts[4 * t + 1], J2 = ts[4 * t + 2], dJ = ts[4 * t + 3];
        // q = divJ (1 + jkR) / (kR)^2
        const double ikr2 = 1.0 / (kR * kR);
        const double qr = (dJ.re - dJ.im * kR) * ikr2, qi = (dJ.im + dJ.re * kR) * ikr2;
        const double v[3] = {vx, vy, vz};
        const c128 J[3] = {J0, J1, J2};
        for (int d = 0; d < 3; d++) {
            // t = J - q v ; E += -j omega mu0 t g
            const double tr = J[d].re - qr * v[d], ti = J[d].im - qi * v[d];
            const double pr = tr * gr - ti * gi, pi = tr * gi + ti * gr;
            acc[2 * d] += omega_mu0 * pi;
            acc[2 * d + 1] += -omega_mu0 * pr;
        }
        // gradg = -(1 + jkR) / (4 pi R^3) * exp(+-jkR)
        const double inv4piR3 = inv4piR / R2;
        const double er = cs, ei = h_sign * sn;
        const double ggr = -(er - kR * ei) * inv4piR3, ggi = -(ei + kR * er) * inv4piR3;
        const c128 RxJ[3] = {c128{vy * J2.re - vz * J1.re, vy * J2.im - vz * J1.im},
                             c128{vz * J0.re - vx * J2.re, vz * J0.im - vx * J2.im},
                             c128{vx * J1.re - vy * J0.re, vx * J1.im - vy * J0.im}};
        for (int d = 0; d < 3; d++) {
            acc[6 + 2 * d] += RxJ[d].re * ggr - RxJ[d].im * ggi;
            acc[6 + 2 * d + 1] += RxJ[d].re * ggi + RxJ[d].im * ggr;
        }
    }
    __shared__ double red[12][4];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    for (int c = 0; c < 12; c++) {
        double v = wave_sum(acc[c]);
        if (lane == 0)
            red[c][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++)
            s += red[threadIdx.x][w];
        c128* out = threadIdx.x < 6 ? E : H;
        if (out)
            reinterpret_cast<double*>(out + (size_t)ip * 3)[threadIdx.x % 6] = s;
    }
}

// ----------------------------------------------------------------------------------------------
// Host launchers
// ----------------------------------------------------------------------------------------------
int launch_fill_quad(const DevMesh& m, double k, double inv_ksq, int prec, void* Cbuf, void* Z, int64_t ldz,
                     const int2* near_pairs, int64_t nnear, hipStream_t st, KernelTimer* kt)
{
    dim3 g1((m.T + TQ_TILE - 1) / TQ_TILE, (m.T + TP_TILE - 1) / TP_TILE);
    if (g1.y > 65535)
        return set_error("too many triangles for the triangle-pair launch"), -1;
    if (kt)
        kt->begin(st);
    if (prec == 0) {
        size_t sh = 2 * TP_TILE * m.q * sizeof(double4);
        if (m.q == 7) {
            hipLaunchKernelGGL((tripair_f64<7, double>), g1, dim3(256), sh, st, m, k, inv_ksq, (c128*)Cbuf);
        } else {
            if (sh > 64 * 1024)
                FRICO_HIP_CHECK(allow_lds(reinterpret_cast<const void*>(&tripair_f64<0, double>), sh));
            hipLaunchKernelGGL((tripair_f64<0, double>), g1, dim3(256), sh, st, m, k, inv_ksq, (c128*)Cbuf);
        }
    } else {
        size_t sh = 2 * TP_TILE * m.q * sizeof(float4);
        if (m.q == 7) {
            hipLaunchKernelGGL((tripair_f32<7>), g1, dim3(256), sh, st, m, k, inv_ksq, (c64*)Cbuf);
        } else {
            if (sh > 64 * 1024)
                FRICO_HIP_CHECK(allow_lds(reinterpret_cast<const void*>(&tripair_f32<0>), sh));
            hipLaunchKernelGGL((tripair_f32<0>), g1, dim3(256), sh, st, m, k, inv_ksq, (c64*)Cbuf);
        }
    }
    if (kt)
        kt->end(st, (double)m.T * (double)m.T);
    FRICO_HIP_CHECK(hipGetLastError());
    // Assembly in column slabs so that grid.y stays within limits.
    const int64_t slab = 65535;
    for (int64_t c0 = 0; c0 < m.N; c0 += slab) {
        int64_t nc = m.N - c0 < slab ? m.N - c0 : slab;
        dim3 g2((unsigned)nc);
        if (prec == 0)
            hipLaunchKernelGGL((assemble<double, double>), g2, dim3(256), 0, st, m, (const c128*)Cbuf, (c128*)Z, ldz, c0, nc);
        else
            hipLaunchKernelGGL((assemble<float, float>), g2, dim3(256), 0, st, m, (const c64*)Cbuf, (c64*)Z, ldz, c0, nc);
    }
    FRICO_HIP_CHECK(hipGetLastError());
    if (nnear > 0) {
        if (prec == 0)
            hipLaunchKernelGGL((near_f64<double>), dim3((unsigned)nnear), dim3(64), 0, st, m, k, inv_ksq, near_pairs, nnear,
                               (c128*)Z, ldz);
        else
            hipLaunchKernelGGL((near_f64<float>), dim3((unsigned)nnear), dim3(64), 0, st, m, k, inv_ksq, near_pairs, nnear,
                               (c64*)Z, ldz);
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_fill_approx(const DevMesh& m, double k, double inv_ksq, int prec, void* Z, int64_t ldz, hipStream_t st)
{
    const int64_t slab = 65535;
    for (int64_t c0 = 0; c0 < m.N; c0 += slab) {
        int64_t nc = m.N - c0 < slab ? m.N - c0 : slab;
        dim3 g((m.N + 255) / 256, (unsigned)nc);
        if (prec == 0)
            hipLaunchKernelGGL((fill_approx<double>), g, dim3(256), 0, st, m, k, inv_ksq, (c128*)Z, ldz, c0);
        else
            hipLaunchKernelGGL((fill_approx<float>), g, dim3(256), 0, st, m, k, inv_ksq, (c64*)Z, ldz, c0);
    }
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_rhs_planewave(const DevMesh& m, double k, double omega_mu0, const PlaneWaveDev* waves, int nang, c128* B,
                         int64_t ldb, hipStream_t st)
{
    dim3 g((m.N + 255) / 256, nang);
    hipLaunchKernelGGL(rhs_planewave, g, dim3(256), 0, st, m, k, omega_mu0, waves, nang, B, ldb);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_scatter(const DevMesh& m, const c128* I, int64_t ldi, int nrhs, c128* TS, hipStream_t st)
{
    dim3 g((m.T + 255) / 256, nrhs);
    hipLaunchKernelGGL(scatter_currents, g, dim3(256), 0, st, m, I, ldi, nrhs, TS);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_fields(const DevMesh& m, double k, double omega_mu0, const c128* TS, int ncols, const double* pts,
                  const int* cols, int npts, double h_sign, c128* E, c128* H, hipStream_t st)
{
    hipLaunchKernelGGL(eval_fields, dim3(npts), dim3(256), 0, st, m, k, omega_mu0, TS, ncols, pts, cols, npts, h_sign,
                       E, H);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

} // namespace frico
