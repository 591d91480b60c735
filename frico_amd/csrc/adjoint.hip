// Adjoint of the quadrature-mode (and approximate-mode) impedance fill: the contraction
//     g(x) = sum_{i,j} W_j I_i dZ(j,i)/dx          (x = vertex coordinates, lengths held fixed)
// for a forward current I and an adjoint vector W.  It has no counterpart in the reference; it
// differentiates compute_matrix (src/emw_solver.cpp:45-142).
//
// Decomposition (same as the forward fill):
//   tripair_adjoint  every ordered triangle pair with the regular x regular rule, one thread per
//                    pair.  With a_m = s I_i l_i (RWG i on local edge m of P) and
//                    b_n = s W_j l_j / (4 pi), the pair contributes
//                    S = sum_mn a_m b_n C[P,m;Q,n] = 1/(A_P A_Q) sum_pr w_p w_r h_pr g(R_pr),
//                    h = 0.25 (A0 r_p - Af).(B0 r_r - Bf) - A0 B0 / k^2.  Its derivative with
//                    respect to the 9 + 9 vertex coordinates of P and Q is accumulated per vertex.
//   near_adjoint     RWG pairs with a sub-triangle block: adds d(S_split - S_regular) of that
//                    block, so the sum equals the derivative of the entries the fill produces.
// The quadrature weights are w = (rule weight) x (triangle or sub-triangle area) and the
// prefactor is 1/(A_P A_Q); the area dependence cancels exactly (each sub-triangle is a third of
// the triangle), so only point positions and free vertices carry derivatives.  The edge-length
// dependence of l_i l_j is added by the host from Z I and Z^T W.
#include "adjoint.hpp"


namespace frico {

__device__ inline void cadd3(double* acc, double sr, double si, double x, double y, double z)
{
    // acc[6] (re/im interleaved per component) += (sr + j si) * (x, y, z)
    acc[0] += sr * x;
    acc[1] += si * x;
    acc[2] += sr * y;
    acc[3] += si * y;
    acc[4] += sr * z;
    acc[5] += si * z;
}

// c * (vr + j vi) for complex scalar c and complex 3-vector v, accumulated
__device__ inline void caddc3(double* acc, double cr, double ci, const double* v)
{
    for (int d = 0; d < 3; d++) {
        acc[2 * d] += cr * v[2 * d] - ci * v[2 * d + 1];
        acc[2 * d + 1] += cr * v[2 * d + 1] + ci * v[2 * d];
    }
}

template <int QC>
__global__ __launch_bounds__(256) void tripair_adjoint(DevMesh m, AdjMesh am, double k, double inv_ksq,
                                                       const c128* __restrict__ aw, const c128* __restrict__ bw,
                                                       double* __restrict__ gvert)
{
    constexpr int QM = QC;
    extern __shared__ double4 smem_adj[];
    double4* sP = smem_adj;
    double4* sQ = smem_adj + 16 * QM;
    __shared__ double red[4][16][18];
    const int P0 = blockIdx.y * 16, Q0 = blockIdx.x * 16;
    for (int i = threadIdx.x; i < 16 * QM; i += blockDim.x) {
        const int t = P0 + i / QM, s = Q0 + i / QM;
        sP[i] = t < m.T ? m.qreg[(size_t)t * QM + i % QM] : double4{0, 0, 0, 0};
        sQ[i] = s < m.T ? m.qreg[(size_t)s * QM + i % QM] : double4{0, 0, 0, 0};
    }
    __syncthreads();
    const int tq = threadIdx.x & 15, tp = threadIdx.x >> 4;
    const int P = P0 + tp, Q = Q0 + tq;
    const bool valid = P < m.T && Q < m.T;
    double gP[18], gQ[18];
#pragma unroll
    for (int x = 0; x < 18; x++)
        gP[x] = gQ[x] = 0;
    if (valid) {
        // A0 = sum a_m, Aal = sum a_m alpha_m (complex 3-vector), same for B
        double a0r = 0, a0i = 0, b0r = 0, b0i = 0, Aal[6] = {0, 0, 0, 0, 0, 0}, Bbe[6] = {0, 0, 0, 0, 0, 0};
        double ar[3], ai[3], br[3], bi[3];
        for (int mm = 0; mm < 3; mm++) {
            const c128 a = aw[3 * P + mm], b = bw[3 * Q + mm];
            ar[mm] = a.re;
            ai[mm] = a.im;
            br[mm] = b.re;
            bi[mm] = b.im;
            a0r += a.re;
            a0i += a.im;
            b0r += b.re;
            b0i += b.im;
            const double* al = m.alpha + 9 * (size_t)P + 3 * mm;
            const double* be = m.alpha + 9 * (size_t)Q + 3 * mm;
            cadd3(Aal, a.re, a.im, al[0], al[1], al[2]);
            cadd3(Bbe, b.re, b.im, be[0], be[1], be[2]);
        }
        const double dx = m.cen[3 * P] - m.cen[3 * Q], dy = m.cen[3 * P + 1] - m.cen[3 * Q + 1],
                     dz = m.cen[3 * P + 2] - m.cen[3 * Q + 2];
        const double ab_r = (a0r * b0r - a0i * b0i) * inv_ksq, ab_i = (a0r * b0i + a0i * b0r) * inv_ksq;
        // V_r = B0 v_r + Bbe for each point of Q
        double Vr[QM][6];
        double4 vq[QM];
#pragma unroll
        for (int r = 0; r < QM; r++) {
            vq[r] = sQ[tq * QM + r];
            const double v[3] = {vq[r].x, vq[r].y, vq[r].z};
            for (int d = 0; d < 3; d++) {
                Vr[r][2 * d] = b0r * v[d] - b0i * 0.0 + Bbe[2 * d];
                Vr[r][2 * d + 1] = b0i * v[d] + Bbe[2 * d + 1];
            }
        }
        double gr[QM][6];
#pragma unroll
        for (int r = 0; r < QM; r++)
#pragma unroll
            for (int x = 0; x < 6; x++)
                gr[r][x] = 0;
        double YP[6] = {0, 0, 0, 0, 0, 0}, YQ[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int p = 0; p < QM; p++) {
            const double4 u = sP[tp * QM + p];
            const double uu[3] = {u.x, u.y, u.z};
            double Up[6];
            for (int d = 0; d < 3; d++) {
                Up[2 * d] = a0r * uu[d] + Aal[2 * d];
                Up[2 * d + 1] = a0i * uu[d] + Aal[2 * d + 1];
            }
            const double ex = dx + u.x, ey = dy + u.y, ez = dz + u.z;
            double gp[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < QM; r++) {
                const double rx = ex - vq[r].x, ry = ey - vq[r].y, rz = ez - vq[r].z;
                const double R2 = rx * rx + ry * ry + rz * rz;
                if (R2 > 0) {
                    const double R = sqrt(R2), iR = 1.0 / R;
                    double sn, cs;
                    sincos(k * R, &sn, &cs);
                    const double gre = cs * iR, gim = -sn * iR;
                    // g'/R = -(jk + 1/R) g / R
                    const double gpr = -(iR * gre - k * gim) * iR, gpi = -(iR * gim + k * gre) * iR;
                    // h = 0.25 U.V - A0 B0 / k^2
                    double hr = 0, hi = 0;
                    for (int d = 0; d < 3; d++) {
                        hr += Up[2 * d] * Vr[r][2 * d] - Up[2 * d + 1] * Vr[r][2 * d + 1];
                        hi += Up[2 * d] * Vr[r][2 * d + 1] + Up[2 * d + 1] * Vr[r][2 * d];
                    }
                    hr = 0.25 * hr - ab_r;
                    hi = 0.25 * hi - ab_i;
                    const double wr = vq[r].w, wp = u.w;
                    // h g'/R
                    const double hgr = hr * gpr - hi * gpi, hgi = hr * gpi + hi * gpr;
                    // dS/dr_p (/ w_p): w_r [0.25 A0 g V_r + h g'/R d]
                    const double cAr = 0.25 * (a0r * gre - a0i * gim) * wr, cAi = 0.25 * (a0r * gim + a0i * gre) * wr;
                    caddc3(gp, cAr, cAi, Vr[r]);
                    cadd3(gp, hgr * wr, hgi * wr, rx, ry, rz);
                    // dS/dr_r (/ w_r): w_p [0.25 B0 g U_p - h g'/R d]
                    const double cBr = 0.25 * (b0r * gre - b0i * gim) * wp, cBi = 0.25 * (b0r * gim + b0i * gre) * wp;
                    caddc3(gr[r], cBr, cBi, Up);
                    cadd3(gr[r], -hgr * wp, -hgi * wp, rx, ry, rz);
                    // Y_P = sum w w g V_r, Y_Q = sum w w g U_p
                    caddc3(YP, wp * wr * gre, wp * wr * gim, Vr[r]);
                    caddc3(YQ, wp * wr * gre, wp * wr * gim, Up);
                }
            }
            for (int kk = 0; kk < 3; kk++) {
                const double f = am.qbeta[3 * p + kk] * u.w;
                for (int x = 0; x < 6; x++)
                    gP[6 * kk + x] += f * gp[x];
            }
        }
#pragma unroll
        for (int r = 0; r < QM; r++)
            for (int kk = 0; kk < 3; kk++) {
                const double f = am.qbeta[3 * r + kk] * vq[r].w;
                for (int x = 0; x < 6; x++)
                    gQ[6 * kk + x] += f * gr[r][x];
            }
        // free vertices: vertex k of P is the free vertex of local edge minv[k] = {1,2,0}[k];
        // dS/df_m = -a_m * 0.25 * Y_P
        const int minv[3] = {1, 2, 0};
        for (int kk = 0; kk < 3; kk++) {
            const int mm = minv[kk];
            caddc3(gP + 6 * kk, -0.25 * ar[mm], -0.25 * ai[mm], YP);
            caddc3(gQ + 6 * kk, -0.25 * br[mm], -0.25 * bi[mm], YQ);
        }
        const double invAA = 1.0 / (m.area[P] * m.area[Q]);
        for (int x = 0; x < 18; x++) {
            gP[x] *= invAA;
            gQ[x] *= invAA;
        }
    }
    // P part: reduce over the 16 lanes sharing P (contiguous lanes)
#pragma unroll
    for (int x = 0; x < 18; x++) {
        double v = gP[x];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        gP[x] = v;
    }
    if (tq == 0 && P < m.T) {
        for (int kk = 0; kk < 3; kk++) {
            const int vtx = am.tri_v[3 * P + kk];
            for (int x = 0; x < 6; x++)
                atomicAdd(gvert + 6 * (size_t)vtx + x, gP[6 * kk + x]);
        }
    }
    // Q part: lanes with equal tq are 16 apart; reduce inside the wave, then across waves in LDS
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int x = 0; x < 18; x++) {
        double v = gQ[x];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        gQ[x] = v;
    }
    if ((threadIdx.x & 63) < 16)
        for (int x = 0; x < 18; x++)
            red[wv][tq][x] = gQ[x];
    __syncthreads();
    if (threadIdx.x < 16 && Q0 + (int)threadIdx.x < m.T) {
        const int q = Q0 + threadIdx.x;
        for (int kk = 0; kk < 3; kk++) {
            const int vtx = am.tri_v[3 * q + kk];
            for (int x = 0; x < 6; x++) {
                const double s = red[0][threadIdx.x][6 * kk + x] + red[1][threadIdx.x][6 * kk + x] +
                                 red[2][threadIdx.x][6 * kk + x] + red[3][threadIdx.x][6 * kk + x];
                atomicAdd(gvert + 6 * (size_t)vtx + x, s);
            }
        }
    }
}

__device__ inline double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// One wavefront per near RWG pair (i, j): d(S_split - S_regular) of the blocks whose row triangle
// is a triangle of i.
__global__ __launch_bounds__(64) void near_adjoint(DevMesh m, AdjMesh am, double k, double inv_ksq,
                                                   const int2* __restrict__ pairs, int64_t npairs,
                                                   const c128* __restrict__ I, const c128* __restrict__ W,
                                                   double* __restrict__ gvert)
{
    const int64_t pidx = blockIdx.x;
    if (pidx >= npairs)
        return;
    const int i = pairs[pidx].x, j = pairs[pidx].y, lane = threadIdx.x;
    const int ti[2] = {m.rwg_t[2 * i], m.rwg_t[2 * i + 1]};
    const int li[2] = {m.rwg_le[2 * i], m.rwg_le[2 * i + 1]};
    const int tj[2] = {m.rwg_t[2 * j], m.rwg_t[2 * j + 1]};
    const int lj[2] = {m.rwg_le[2 * j], m.rwg_le[2 * j + 1]};
    const c128 Ii = I[i], Wj = W[j];
    const double li_len = m.rwg_len[i], lj_len = m.rwg_len[j];
    for (int a = 0; a < 2; a++) {
        for (int b = 0; b < 2; b++) {
            const int P = ti[a], Q = tj[b];
            if (!(Q == ti[0] || Q == ti[1]))
                continue;
            const double sa = a == 0 ? -1.0 : 1.0, sb = b == 0 ? -1.0 : 1.0;
            // a = s I_i l_i, b = s W_j l_j / (4 pi) for this block's single (m, n)
            const double a0r = sa * li_len * Ii.re, a0i = sa * li_len * Ii.im;
            const double b0r = sb * lj_len * Wj.re * 0.25 * INV_PI, b0i = sb * lj_len * Wj.im * 0.25 * INV_PI;
            const double* al = m.alpha + 9 * (size_t)P + 3 * li[a];
            const double* be = m.alpha + 9 * (size_t)Q + 3 * lj[b];
            const double dx = m.cen[3 * P] - m.cen[3 * Q], dy = m.cen[3 * P + 1] - m.cen[3 * Q + 1],
                         dz = m.cen[3 * P + 2] - m.cen[3 * Q + 2];
            const double ab_r = (a0r * b0r - a0i * b0i) * inv_ksq, ab_i = (a0r * b0i + a0i * b0r) * inv_ksq;
            double gP[18], gQ[18], YP[6] = {0, 0, 0, 0, 0, 0}, YQ[6] = {0, 0, 0, 0, 0, 0};
            for (int x = 0; x < 18; x++)
                gP[x] = gQ[x] = 0;
            const int q = m.q, nsub = 3 * m.q;
            const int total = q * (nsub + q);
            for (int t = lane; t < total; t += 64) {
                const int p = t / (nsub + q), rr = t % (nsub + q);
                const bool split = rr < nsub;
                const double sgn = split ? 1.0 : -1.0;
                const double4 u = m.qreg[(size_t)P * q + p];
                const double4 v = split ? m.qsub[(size_t)Q * nsub + rr] : m.qreg[(size_t)Q * q + (rr - nsub)];
                const double* bq = split ? am.qsub_beta + 3 * rr : am.qbeta + 3 * (rr - nsub);
                const double rx = dx + u.x - v.x, ry = dy + u.y - v.y, rz = dz + u.z - v.z;
                const double R2 = rx * rx + ry * ry + rz * rz;
                if (!(R2 > 0))
                    continue;
                const double R = sqrt(R2), iR = 1.0 / R;
                double sn, cs;
                sincos(k * R, &sn, &cs);
                const double gre = cs * iR, gim = -sn * iR;
                const double gpr = -(iR * gre - k * gim) * iR, gpi = -(iR * gim + k * gre) * iR;
                double Up[6], Vr[6];
                const double uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
                for (int d = 0; d < 3; d++) {
                    Up[2 * d] = a0r * (uu[d] + al[d]);
                    Up[2 * d + 1] = a0i * (uu[d] + al[d]);
                    Vr[2 * d] = b0r * (vv[d] + be[d]);
                    Vr[2 * d + 1] = b0i * (vv[d] + be[d]);
                }
                double hr = 0, hi = 0;
                for (int d = 0; d < 3; d++) {
                    hr += Up[2 * d] * Vr[2 * d] - Up[2 * d + 1] * Vr[2 * d + 1];
                    hi += Up[2 * d] * Vr[2 * d + 1] + Up[2 * d + 1] * Vr[2 * d];
                }
                hr = 0.25 * hr - ab_r;
                hi = 0.25 * hi - ab_i;
                const double ww = sgn * u.w * v.w;
                const double hgr = hr * gpr - hi * gpi, hgi = hr * gpi + hi * gpr;
                double gp[6] = {0, 0, 0, 0, 0, 0}, gq[6] = {0, 0, 0, 0, 0, 0};
                caddc3(gp, 0.25 * (a0r * gre - a0i * gim) * ww, 0.25 * (a0r * gim + a0i * gre) * ww, Vr);
                cadd3(gp, hgr * ww, hgi * ww, rx, ry, rz);
                caddc3(gq, 0.25 * (b0r * gre - b0i * gim) * ww, 0.25 * (b0r * gim + b0i * gre) * ww, Up);
                cadd3(gq, -hgr * ww, -hgi * ww, rx, ry, rz);
                caddc3(YP, ww * gre, ww * gim, Vr);
                caddc3(YQ, ww * gre, ww * gim, Up);
                for (int kk = 0; kk < 3; kk++) {
                    const double fp = am.qbeta[3 * p + kk], fq = bq[kk];
                    for (int x = 0; x < 6; x++) {
                        gP[6 * kk + x] += fp * gp[x];
                        gQ[6 * kk + x] += fq * gq[x];
                    }
                }
            }
            // free vertex of local edge m is vertex {2,0,1}[m]: df = -(0.25) * (coefficient) * Y
            const int evmap[3] = {2, 0, 1};
            const double invAA = 1.0 / (m.area[P] * m.area[Q]);
            // dS/df_P = -0.25 a0 Y_P / a0 ... Up already carries a0: the derivative of U = a0 (u + al)
            // w.r.t. f (al = c - f) is -a0, so df_P += -0.25 * ww * g * V_r * a0 -> use YP with a0
            double fP[6] = {0, 0, 0, 0, 0, 0}, fQ[6] = {0, 0, 0, 0, 0, 0};
            caddc3(fP, -0.25 * a0r, -0.25 * a0i, YP);
            caddc3(fQ, -0.25 * b0r, -0.25 * b0i, YQ);
            for (int x = 0; x < 18; x++) {
                gP[x] = wsum(gP[x]) * invAA;
                gQ[x] = wsum(gQ[x]) * invAA;
            }
            for (int x = 0; x < 6; x++) {
                fP[x] = wsum(fP[x]) * invAA;
                fQ[x] = wsum(fQ[x]) * invAA;
            }
            if (lane == 0) {
                for (int kk = 0; kk < 3; kk++) {
                    const int vp = am.tri_v[3 * P + kk], vq = am.tri_v[3 * Q + kk];
                    for (int x = 0; x < 6; x++) {
                        atomicAdd(gvert + 6 * (size_t)vp + x, gP[6 * kk + x]);
                        atomicAdd(gvert + 6 * (size_t)vq + x, gQ[6 * kk + x]);
                    }
                }
                const int fpv = am.tri_v[3 * P + evmap[li[a]]], fqv = am.tri_v[3 * Q + evmap[lj[b]]];
                for (int x = 0; x < 6; x++) {
                    atomicAdd(gvert + 6 * (size_t)fpv + x, fP[x]);
                    atomicAdd(gvert + 6 * (size_t)fqv + x, fQ[x]);
                }
            }
        }
    }
}

// a[3T], b[3T] from I and W: a[t][e] = sgn * l_n * I_n, b[t][e] = sgn * l_n * W_n / (4 pi)
__global__ __launch_bounds__(256) void adjoint_weights(DevMesh m, const c128* __restrict__ I, const c128* __restrict__ W,
                                                       c128* __restrict__ aw, c128* __restrict__ bw)
{
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x >= 3 * (int64_t)m.T)
        return;
    const int n = m.tri_rwg[x];
    if (n < 0) {
        aw[x] = c128{0, 0};
        bw[x] = c128{0, 0};
        return;
    }
    const double s = m.tri_sgn[x] * m.rwg_len[n];
    aw[x] = c128{s * I[n].re, s * I[n].im};
    bw[x] = c128{s * W[n].re * 0.25 * INV_PI, s * W[n].im * 0.25 * INV_PI};
}

// ----------------------------------------------------------------------------------------------
// Approximate mode (compute_matrix_approx, src/emw_solver.cpp:144-220).  Its entry for a triangle
// pair P != Q is the quadrature form's with the one-point rule at the barycenter (weight 1, so
// w_P w_Q / (A_P A_Q) = 1): tripair_adjoint<1> on the centroid rule below.  A pair of equal
// triangles takes G = 0.25 (sqrt(4 pi) / sqrt(A) - jk) / pi instead of the singular 1 / R, whose
// pair tripair_adjoint skips (R = 0): approx_self_adjoint adds it.  With a_m, b_m as above (b
// carrying 1 / (4 pi)), S_PP = h g, h = 0.25 U.V - A0 B0 / k^2, U = sum a_m (c - f_m), V = sum
// b_m (c - f_m), g = 4 pi G = sqrt(4 pi / A) - j k; dS = g dh + h (dg/dA) dA, dh/dc =
// 0.25 (A0 V + B0 U) (a third to each vertex), dh/df_m = -0.25 (a_m V + b_m U), dg/dA =
// -0.5 sqrt(4 pi) A^(-3/2), dA/dv the cross-product area gradient.
// ----------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void centroid_rule(DevMesh m, double4* __restrict__ out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < m.T)
        out[t] = double4{0.0, 0.0, 0.0, m.area[t]};
}

__global__ __launch_bounds__(256) void approx_self_adjoint(DevMesh m, AdjMesh am, double k, double inv_ksq,
                                                           const c128* __restrict__ aw, const c128* __restrict__ bw,
                                                           double* __restrict__ gvert)
{
    const int P = blockIdx.x * 256 + threadIdx.x;
    if (P >= m.T)
        return;
    double a0r = 0, a0i = 0, b0r = 0, b0i = 0, U[6] = {0, 0, 0, 0, 0, 0}, V[6] = {0, 0, 0, 0, 0, 0};
    double ar[3], ai[3], br[3], bi[3];
    for (int mm = 0; mm < 3; mm++) {
        const c128 a = aw[3 * P + mm], b = bw[3 * P + mm];
        ar[mm] = a.re;
        ai[mm] = a.im;
        br[mm] = b.re;
        bi[mm] = b.im;
        a0r += a.re;
        a0i += a.im;
        b0r += b.re;
        b0i += b.im;
        const double* al = m.alpha + 9 * (size_t)P + 3 * mm; // c - f_m
        cadd3(U, a.re, a.im, al[0], al[1], al[2]);
        cadd3(V, b.re, b.im, al[0], al[1], al[2]);
    }
    const double A = m.area[P];
    const double s4pi = sqrt(4.0 * PI);
    const double gre = s4pi / sqrt(A), gim = -k;
    double hr = 0, hi = 0;
    for (int d = 0; d < 3; d++) {
        hr += U[2 * d] * V[2 * d] - U[2 * d + 1] * V[2 * d + 1];
        hi += U[2 * d] * V[2 * d + 1] + U[2 * d + 1] * V[2 * d];
    }
    hr = 0.25 * hr - (a0r * b0r - a0i * b0i) * inv_ksq;
    hi = 0.25 * hi - (a0r * b0i + a0i * b0r) * inv_ksq;
    // g dh/dc, a third to each vertex
    double dc[6] = {0, 0, 0, 0, 0, 0}, gc[6] = {0, 0, 0, 0, 0, 0};
    caddc3(dc, 0.25 * a0r, 0.25 * a0i, V);
    caddc3(dc, 0.25 * b0r, 0.25 * b0i, U);
    caddc3(gc, gre, gim, dc);
    // vertex k = c - alpha of local edge minv[k] (vertex k is the free vertex of that edge)
    const int minv[3] = {1, 2, 0};
    const double* c = m.cen + 3 * (size_t)P;
    double v[3][3];
    for (int kk = 0; kk < 3; kk++)
        for (int d = 0; d < 3; d++)
            v[kk][d] = c[d] - m.alpha[9 * (size_t)P + 3 * minv[kk] + d];
    // cross-product area gradient: d/dv0 = 0.5 n x (v2 - v1), cyclic
    const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double dgA = -0.5 * s4pi / (A * sqrt(A));
    for (int kk = 0; kk < 3; kk++) {
        double g[6];
        for (int x = 0; x < 6; x++)
            g[x] = gc[x] / 3.0;
        // free vertex of local edge minv[kk]: g dh/df = -0.25 g (a_m V + b_m U)
        const int mm = minv[kk];
        double df[6] = {0, 0, 0, 0, 0, 0};
        caddc3(df, -0.25 * ar[mm], -0.25 * ai[mm], V);
        caddc3(df, -0.25 * br[mm], -0.25 * bi[mm], U);
        caddc3(g, gre, gim, df);
        if (nn > 0) {
            const double* p1 = v[(kk + 1) % 3];
            const double* p2 = v[(kk + 2) % 3];
            const double e[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
            const double ga[3] = {0.5 * (n[1] * e[2] - n[2] * e[1]) / nn, 0.5 * (n[2] * e[0] - n[0] * e[2]) / nn,
                                  0.5 * (n[0] * e[1] - n[1] * e[0]) / nn};
            cadd3(g, hr * dgA, hi * dgA, ga[0], ga[1], ga[2]);
        }
        const int vtx = am.tri_v[3 * P + kk];
        for (int x = 0; x < 6; x++)
            atomicAdd(gvert + 6 * (size_t)vtx + x, g[x]);
    }
}

int launch_adjoint_contract_approx(const DevMesh& m, const AdjMesh& am, double k, double inv_ksq, const c128* I,
                                   const c128* W, c128* aw, c128* bw, double4* cent_rule, const double* third,
                                   double* gvert, int64_t nv, hipStream_t st)
{
    FRICO_HIP_CHECK(hipMemsetAsync(gvert, 0, (size_t)nv * 6 * sizeof(double), st));
    const unsigned gt = (unsigned)((m.T + 255) / 256);
    hipLaunchKernelGGL(adjoint_weights, dim3((unsigned)((3 * (int64_t)m.T + 255) / 256)), dim3(256), 0, st, m, I, W, aw, bw);
    hipLaunchKernelGGL(centroid_rule, dim3(gt), dim3(256), 0, st, m, cent_rule);
    DevMesh mc = m;
    mc.q = 1;
    mc.qreg = cent_rule;
    AdjMesh amc = am;
    amc.qbeta = third;
    dim3 g((m.T + 15) / 16, (m.T + 15) / 16);
    hipLaunchKernelGGL((tripair_adjoint<1>), g, dim3(256), 2 * 16 * sizeof(double4), st, mc, amc, k, inv_ksq, aw, bw,
                       gvert);
    hipLaunchKernelGGL(approx_self_adjoint, dim3(gt), dim3(256), 0, st, m, am, k, inv_ksq, aw, bw, gvert);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_adjoint_contract(const DevMesh& m, const AdjMesh& am, double k, double inv_ksq, const int2* near_pairs,
                            int64_t nnear, const c128* I, const c128* W, c128* aw, c128* bw, double* gvert, int64_t nv,
                            hipStream_t st)
{
    FRICO_HIP_CHECK(hipMemsetAsync(gvert, 0, (size_t)nv * 6 * sizeof(double), st));
    hipLaunchKernelGGL(adjoint_weights, dim3((unsigned)((3 * (int64_t)m.T + 255) / 256)), dim3(256), 0, st, m, I, W, aw, bw);
    dim3 g((m.T + 15) / 16, (m.T + 15) / 16);
    const size_t sh = 2 * 16 * (size_t)m.q * sizeof(double4);
    switch (m.q) {
    case 1:
        hipLaunchKernelGGL((tripair_adjoint<1>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    case 3:
        hipLaunchKernelGGL((tripair_adjoint<3>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    case 4:
        hipLaunchKernelGGL((tripair_adjoint<4>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    case 12:
        hipLaunchKernelGGL((tripair_adjoint<12>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    case 13:
        hipLaunchKernelGGL((tripair_adjoint<13>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    case 6:
        hipLaunchKernelGGL((tripair_adjoint<6>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    case 7:
        hipLaunchKernelGGL((tripair_adjoint<7>), g, dim3(256), sh, st, m, am, k, inv_ksq, aw, bw, gvert);
        break;
    default:
        return set_error("adjoint: Dunavant degree with %d points is not instantiated (degrees 0 to 7 are)", m.q), -1;
    }
    FRICO_HIP_CHECK(hipGetLastError());
    if (nnear > 0)
        hipLaunchKernelGGL(near_adjoint, dim3((unsigned)nnear), dim3(64), 0, st, m, am, k, inv_ksq, near_pairs, nnear, I, W,
                           gvert);
    FRICO_HIP_CHECK(hipGetLastError());
    return 0;
}

} // namespace frico
