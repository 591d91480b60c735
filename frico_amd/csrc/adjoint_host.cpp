// Adjoint forward/backward pairs of the antenna analysis and of the RCS (no counterpart in the
// reference).
//
// forward:  Z(x) I = V(x) with the delta-gap source, Zin = V / I_port, I_port = sum_gap l s I,
//           G = 4 pi |p|^2 |S| / P at the observation point p, S = E x conj(H) / 2,
//           P = Re(V conj(I_port) / 2)   (src/emw_ana_antenna.cpp:33-105).
// backward: two solves with Z^T on the same LU factors, Z^T lam = c (c = l s on the gap) and
//           Z^T mu = d|S|/dI (Wirtinger), the device contraction of dZ/dx with (I, lam) and
//           (I, mu), and the O(N + T) host terms (edge lengths of V, of the port weights and of
//           Z's l_i l_j prefactor; derivative of the field sums at fixed currents).
#include "ctx.hpp"

namespace {

using cd = std::complex<double>;

struct cvec {
    cd x, y, z;
};

// d hypot(dx, dy) / d vertices of edge e, scaled by s, accumulated into g (3 nv complex)
void add_length_grad(const HostMesh& hm, int64_t edge, double len, cd s, std::vector<cd>& g)
{
    if (len == 0)
        return;
    const auto& e = hm.edges[edge];
    const vec3 d = hm.V[e.second] - hm.V[e.first];
    g[3 * e.second] += s * (d.x / len);
    g[3 * e.second + 1] += s * (d.y / len);
    g[3 * e.first] -= s * (d.x / len);
    g[3 * e.first + 1] -= s * (d.y / len);
}

// Vector-Jacobian product of E, H at point pt (centroid-dipole sums, src/emw_postpro_common.cpp:31-66,
// fed by the scatter of src/emw_solver.h:137-164): gx (3 nv) and gI (N) for cotangents eE, eH.
void field_vjp(const HostMesh& hm, double freq, const std::vector<cd>& I, const double* pt, const cd eE[3],
               const cd eH[3], std::vector<cd>& gx, std::vector<cd>& gI)
{
    const size_t nv = hm.V.size(), T = hm.T.size(), N = hm.bf.size();
    gx.assign(3 * nv, cd(0));
    gI.assign(N, cd(0));
    const double omega = 2.0 * PI * freq, k = omega * std::sqrt(MU0 * EPS0);
    const cd jwm{0.0, omega * MU0}, jk{0.0, k};
    const int evmap[3] = {2, 0, 1};
    std::vector<cd> AJ(3 * T, cd(0)), D(T, cd(0));
    for (size_t t = 0; t < T; t++)
        for (int e = 0; e < 3; e++) {
            const int n = hm.tri_rwg[3 * t + e];
            if (n < 0)
                continue;
            const double s = hm.tri_sgn[3 * t + e] * hm.rwg_len[n];
            for (int d = 0; d < 3; d++)
                AJ[3 * t + d] += 0.5 * s * hm.alpha[9 * t + 3 * e + d] * I[n];
            D[t] += s * I[n];
        }
    std::vector<cd> cAJ(3 * T), cD(T), gc(3 * T, cd(0));
    for (size_t t = 0; t < T; t++) {
        const double v[3] = {pt[0] - hm.cen[3 * t], pt[1] - hm.cen[3 * t + 1], pt[2] - hm.cen[3 * t + 2]};
        const double R = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const cd e = std::exp(cd{0, -k * R}), ep = std::exp(cd{0, k * R});
        const cd g = e / (4 * PI * R), gp = -(jk + 1.0 / R) * g;
        const cd vphi = e / (4 * PI * k * k * R * R * R);
        const cd phi = (1.0 + jk * R) * vphi, phip = vphi * (-3.0 / R + k * k * R - 3.0 * jk);
        const cd vpsi = ep / (4 * PI * R * R * R);
        const cd psi = -(1.0 + jk * R) * vpsi, psip = vpsi * (3.0 / R + k * k * R + jk);
        const cd* A = &AJ[3 * t];
        const cd eEv = eE[0] * v[0] + eE[1] * v[1] + eE[2] * v[2];
        const cd eEA = eE[0] * A[0] + eE[1] * A[1] + eE[2] * A[2];
        const cd eHxv[3] = {eH[1] * v[2] - eH[2] * v[1], eH[2] * v[0] - eH[0] * v[2], eH[0] * v[1] - eH[1] * v[0]};
        const cd vxA[3] = {v[1] * A[2] - v[2] * A[1], v[2] * A[0] - v[0] * A[2], v[0] * A[1] - v[1] * A[0]};
        const cd eHvxA = eH[0] * vxA[0] + eH[1] * vxA[1] + eH[2] * vxA[2];
        const cd Axe[3] = {A[1] * eH[2] - A[2] * eH[1], A[2] * eH[0] - A[0] * eH[2], A[0] * eH[1] - A[1] * eH[0]};
        for (int d = 0; d < 3; d++)
            cAJ[3 * t + d] = -jwm * g * eE[d] + psi * eHxv[d];
        cD[t] = jwm * phi * eEv;
        for (int d = 0; d < 3; d++) {
            const cd dE = -jwm * (eEA * gp * v[d] / R - D[t] * (phip * eEv * v[d] / R + phi * eE[d]));
            const cd dH = psi * Axe[d] + psip * v[d] / R * eHvxA;
            gc[3 * t + d] -= dE + dH; // v = pt - centroid
        }
    }
    for (size_t t = 0; t < T; t++)
        for (int e = 0; e < 3; e++) {
            const int n = hm.tri_rwg[3 * t + e];
            if (n < 0)
                continue;
            const double sg = hm.tri_sgn[3 * t + e], l = hm.rwg_len[n];
            const double* al = &hm.alpha[9 * t + 3 * e];
            const cd ca = cAJ[3 * t] * al[0] + cAJ[3 * t + 1] * al[1] + cAJ[3 * t + 2] * al[2];
            gI[n] += sg * l * 0.5 * ca + sg * l * cD[t];
            // d/dl of AJ and D
            add_length_grad(hm, hm.bf[n].edge_index, l, sg * 0.5 * ca * I[n] + sg * I[n] * cD[t], gx);
            // alpha = centroid - free vertex: the free vertex gets -coefficient, the centroid +
            const int64_t fv = hm.T[t][evmap[e]];
            for (int d = 0; d < 3; d++) {
                const cd coef = sg * l * 0.5 * I[n] * cAJ[3 * t + d];
                gc[3 * t + d] += coef;
                gx[3 * fv + d] -= coef;
            }
        }
    for (size_t t = 0; t < T; t++)
        for (int kk = 0; kk < 3; kk++)
            for (int d = 0; d < 3; d++)
                gx[3 * hm.T[t][kk] + d] += gc[3 * t + d] / 3.0;
}

// Reference-triangle Dunavant rule: (beta_0, beta_1, beta_2, reference weight) of each point, in the
// expansion order of dunavant() (src/quadratures.cpp:321-372); the point is sum_k beta_k v_k and its
// weight the reference weight times the triangle's area.
void rule_ref(int degree, std::vector<std::array<double, 4>>& out)
{
    out.clear();
    std::vector<std::array<double, 3>> bary;
    dunavant_bary(degree, bary);
    const int rule = degree == 0 ? 0 : degree - 1;
    size_t idx = 0;
    for (int o = frico_dunavant_orbit_begin[rule]; o < frico_dunavant_orbit_begin[rule + 1]; o++)
        for (int s = 0; s < frico_dunavant_orbits[o].orbit; s++, idx++)
            out.push_back({bary[idx][0], bary[idx][1], bary[idx][2], frico_dunavant_orbits[o].w});
}

// Gradient of the area |(b - a) x (c - a)| / 2 with respect to a, b and c.
void area_grad(vec3 a, vec3 b, vec3 c, vec3 g[3])
{
    const vec3 n = cross(b - a, c - a);
    const double nn = norm(n);
    if (nn == 0) {
        g[0] = g[1] = g[2] = vec3{0, 0, 0};
        return;
    }
    const vec3 u = n / nn;
    g[0] = cross(u, c - b) * 0.5;
    g[1] = cross(u, a - c) * 0.5;
    g[2] = cross(u, b - a) * 0.5;
}

// g (3 nv) += sum_n W_n dV_n/dx for the plane-wave moments (update_rhs(plane_wave),
// src/emw_solver.cpp:380-416): V_n = sum over T-, T+ of -+ (l / 2A) sum_p w_p (r_p - p_T) . E(r_p)
// / (-j omega mu0), E(r) = E0 exp(+j k dir . r).  Every vertex dependence is differentiated: l,
// A, the weights w_p = w_ref area, the points r_p = sum_k beta_pk v_k, the free vertex p_T and the
// field's phase (dE/dr = j k dir E^T).
void planewave_vjp(const HostMesh& hm, double freq, double theta, double phi, cd eth, cd eph, const std::vector<cd>& W,
                   std::vector<cd>& g)
{
    const double omega = 2.0 * PI * freq, k = omega * std::sqrt(MU0 * EPS0);
    const vec3 dir{std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
    const vec3 th{std::cos(theta) * std::cos(phi), std::cos(theta) * std::sin(phi), -std::sin(theta)};
    const vec3 ph{-std::sin(phi), std::cos(phi), 0.0};
    const cd E0[3] = {eth * th.x + eph * ph.x, eth * th.y + eph * ph.y, eth * th.z + eph * ph.z};
    const cd jk{0.0, k}, c0 = 1.0 / cd{0.0, -omega * MU0};
    const int evmap[3] = {2, 0, 1};
    std::vector<std::array<double, 4>> rule;
    rule_ref(hm.degree, rule);
    auto addv = [&](int64_t v, vec3 d, cd s) {
        g[3 * v] += d.x * s;
        g[3 * v + 1] += d.y * s;
        g[3 * v + 2] += d.z * s;
    };
    for (size_t n = 0; n < hm.bf.size(); n++) {
        const frico_rwg& b = hm.bf[n];
        if (W[n] == cd(0))
            continue;
        for (int side = 0; side < 2; side++) {
            const int64_t t = side == 0 ? b.itminus : b.itplus;
            const double sgn = side == 0 ? -1.0 : 1.0, A = side == 0 ? b.Aminus : b.Aplus;
            const int le = side == 0 ? b.loc_eminus : b.loc_eplus;
            const int64_t fv = hm.T[t][evmap[le]];
            const vec3 P[3] = {hm.V[hm.T[t][0]], hm.V[hm.T[t][1]], hm.V[hm.T[t][2]]};
            const double area = kahan_area(P[0], P[1], P[2]);
            vec3 ga[3];
            area_grad(P[0], P[1], P[2], ga);
            const vec3 pf = P[evmap[le]];
            cd F = 0;
            cd dF[3][3] = {}, dFree[3] = {};
            for (const auto& r : rule) {
                const vec3 rp = P[0] * r[0] + P[1] * r[1] + P[2] * r[2];
                const double w = r[3] * area;
                const cd e = std::exp(jk * dot(dir, rp));
                const cd E[3] = {E0[0] * e, E0[1] * e, E0[2] * e};
                const vec3 rho = rp - pf;
                const cd rE = rho.x * E[0] + rho.y * E[1] + rho.z * E[2];
                F += w * rE;
                const double dd[3] = {dir.x, dir.y, dir.z};
                for (int kk = 0; kk < 3; kk++) {
                    const double gw[3] = {ga[kk].x * r[3], ga[kk].y * r[3], ga[kk].z * r[3]};
                    for (int d = 0; d < 3; d++)
                        dF[kk][d] += gw[d] * rE + w * r[kk] * (E[d] + rE * jk * dd[d]);
                }
                for (int d = 0; d < 3; d++)
                    dFree[d] -= w * E[d];
            }
            const cd pre = W[n] * sgn * c0;
            const double lA = b.length / (2.0 * A);
            // d(l / 2A) F = (dl / 2A - l dA / 2A^2) F (A = |e0 x e1| / 2, as the moments use)
            add_length_grad(hm, b.edge_index, b.length, pre * F / (2.0 * A), g);
            for (int kk = 0; kk < 3; kk++)
                addv(hm.T[t][kk], ga[kk], -pre * F * b.length / (2.0 * A * A));
            for (int kk = 0; kk < 3; kk++)
                for (int d = 0; d < 3; d++)
                    g[3 * hm.T[t][kk] + d] += pre * lA * dF[kk][d];
            for (int d = 0; d < 3; d++)
                g[3 * fv + d] += pre * lA * dFree[d];
        }
    }
}

// The device contraction sum W_j I_i dZ(j,i)/dx at fixed edge lengths for the forward fill's mode.
int contract(frico_ctx* c, const c128* I, const c128* W, double k, double inv_ksq)
{
    const HostMesh& hm = c->hm;
    if (!c->adj.approx)
        return launch_adjoint_contract(c->dm, c->am, k, inv_ksq, c->d_near.p, (int64_t)hm.near.size(), I, W, c->aw.p,
                                       c->bw.p, c->d_gvert.p, hm.nv, c->st);
    int rc;
    if ((rc = c->adj_cent.ensure((size_t)hm.nt)))
        return rc;
    if (!c->adj_third.p) {
        const double third[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
        if ((rc = c->adj_third.upload(third, 3, c->st)))
            return rc;
    }
    return launch_adjoint_contract_approx(c->dm, c->am, k, inv_ksq, I, W, c->aw.p, c->bw.p, c->adj_cent.p,
                                          c->adj_third.p, c->d_gvert.p, hm.nv, c->st);
}

int compute_gradients(frico_ctx* c)
{
    auto& A = c->adj;
    const HostMesh& hm = c->hm;
    const int64_t N = c->N, nv = hm.nv;
    hipStream_t st = c->st;
    int rc;
    // Wirtinger derivatives of |S| with respect to E and H
    const cd E[3] = {A.E[0], A.E[1], A.E[2]}, H[3] = {A.H[0], A.H[1], A.H[2]};
    const cd Hc[3] = {std::conj(H[0]), std::conj(H[1]), std::conj(H[2])};
    const cd S[3] = {0.5 * (E[1] * Hc[2] - E[2] * Hc[1]), 0.5 * (E[2] * Hc[0] - E[0] * Hc[2]), 0.5 * (E[0] * Hc[1] - E[1] * Hc[0])};
    const cd Sc[3] = {std::conj(S[0]), std::conj(S[1]), std::conj(S[2])};
    const cd Ec[3] = {std::conj(E[0]), std::conj(E[1]), std::conj(E[2])};
    const double f2 = 1.0 / (2.0 * A.absS);
    const cd eE[3] = {0.5 * (Hc[1] * Sc[2] - Hc[2] * Sc[1]) * f2, 0.5 * (Hc[2] * Sc[0] - Hc[0] * Sc[2]) * f2,
                      0.5 * (Hc[0] * Sc[1] - Hc[1] * Sc[0]) * f2};
    const cd eH[3] = {0.5 * (S[1] * Ec[2] - S[2] * Ec[1]) * f2, 0.5 * (S[2] * Ec[0] - S[0] * Ec[2]) * f2,
                      0.5 * (S[0] * Ec[1] - S[1] * Ec[0]) * f2};
    std::vector<cd> gxf, gI;
    field_vjp(hm, A.freq, A.I, A.obs, eE, eH, gxf, gI);

    // Z^T [lam mu] = [c gI] on the complex128 factors
    std::vector<cd> rhs(2 * N);
    for (int64_t n = 0; n < N; n++) {
        rhs[n] = A.c[n];
        rhs[N + n] = gI[n];
    }
    if ((rc = c->adjW.ensure((size_t)(2 * N))) || (rc = c->aw.ensure((size_t)(3 * hm.nt))) ||
        (rc = c->bw.ensure((size_t)(3 * hm.nt))) || (rc = c->d_gvert.ensure((size_t)(6 * nv))))
        return rc;
    FRICO_HIP_CHECK(hipMemcpyAsync(c->adjW.p, rhs.data(), rhs.size() * sizeof(c128), hipMemcpyHostToDevice, st));
    if ((rc = lu_solve_transposed<double>(c->Z64.p, N, N, c->ipiv.p, c->nb_factored, c->adjW.p, N, 2, st)))
        return rc;
    std::vector<cd> W(2 * N);
    FRICO_HIP_CHECK(hipMemcpyAsync(W.data(), c->adjW.p, W.size() * sizeof(c128), hipMemcpyDeviceToHost, st));
    FRICO_HIP_CHECK(hipStreamSynchronize(st));

    const double omega = 2.0 * PI * A.freq;
    const double k = omega * std::sqrt(MU0 * EPS0), inv_ksq = 1. / (omega * omega * MU0 * EPS0);
    std::vector<cd> C[2];
    for (int w = 0; w < 2; w++) {
        if ((rc = contract(c, c->X.p, c->adjW.p + w * N, k, inv_ksq)))
            return rc;
        C[w].assign(3 * nv, cd(0));
        FRICO_HIP_CHECK(hipMemcpyAsync(C[w].data(), c->d_gvert.p, (size_t)(6 * nv) * sizeof(double), hipMemcpyDeviceToHost, st));
        FRICO_HIP_CHECK(hipStreamSynchronize(st));
        // l_i l_j prefactor: d/dl_n = (I_n (Z^T W)_n + W_n (Z I)_n) / l_n with Z^T W = c or gI, Z I = V
        for (int64_t n = 0; n < N; n++) {
            const double l = hm.bf[n].length;
            if (l == 0)
                continue;
            const cd ztw = w == 0 ? A.c[n] : gI[n];
            add_length_grad(hm, hm.bf[n].edge_index, l, (A.I[n] * ztw + W[w * N + n] * A.V[n]) / l, C[w]);
        }
    }
    // edge-length dependence of V (delta gap) and of the port weights c = l s
    std::vector<cd> dcI(3 * nv, cd(0)), lam_dV(3 * nv, cd(0)), mu_dV(3 * nv, cd(0));
    for (int64_t n = 0; n < N; n++) {
        if (A.c[n] == cd(0) && !hm.bf[n].has_interface)
            continue;
        bool gap = false;
        for (int32_t f : A.ifaces)
            gap |= hm.bf[n].has_interface && hm.bf[n].interface_tag == f;
        if (!gap)
            continue;
        const double s = A.signs.empty() ? 1.0 : A.signs[n];
        const cd dV = cd(0, -1) * A.Vg * s / (omega * MU0); // dV_n / dl_n
        add_length_grad(hm, hm.bf[n].edge_index, hm.bf[n].length, s * A.I[n], dcI);
        add_length_grad(hm, hm.bf[n].edge_index, hm.bf[n].length, W[n] * dV, lam_dV);
        add_length_grad(hm, hm.bf[n].edge_index, hm.bf[n].length, W[N + n] * dV, mu_dV);
    }
    const double r2 = A.obs[0] * A.obs[0] + A.obs[1] * A.obs[1] + A.obs[2] * A.obs[2];
    A.dzin.assign(3 * nv, cd(0));
    A.dgain.assign(3 * nv, 0.0);
    for (int64_t x = 0; x < 3 * nv; x++) {
        const cd dIport = dcI[x] + lam_dV[x] - C[0][x];
        A.dzin[x] = -A.Vg / (A.Iport * A.Iport) * dIport;
        const double dP = (0.5 * A.Vg * std::conj(dIport)).real();
        const double dabsS = 2.0 * gxf[x].real() + 2.0 * (mu_dV[x] - C[1][x]).real();
        A.dgain[x] = 4 * PI * r2 / A.P * dabsS - A.gain / A.P * dP;
    }
    A.grads_done = true;
    return 0;
}

} // namespace

extern "C" {

int frico_adjoint_set_fill(frico_ctx* c, int mode)
{
    if (!c)
        return set_error("null context"), -1;
    if (mode != FRICO_FILL_QUAD && mode != FRICO_FILL_APPROX)
        return set_error("fill mode must be FRICO_FILL_QUAD or FRICO_FILL_APPROX"), -2;
    c->adj_fill = mode;
    return 0;
}

int frico_adjoint_forward(frico_ctx* c, double freq, const int32_t* ifaces, int nif, double v_re, double v_im,
                          const double* obs, double* out)
{
    if (!c)
        return set_error("null context"), -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!ifaces || nif < 1)
        return set_error("no delta-gap interfaces"), -3;
    if (!obs || !out)
        return set_error("null observation point or output"), -7;
    const bool approx = c->adj_fill == FRICO_FILL_APPROX;
    if (!approx && c->hm.degree > 7)
        return set_error("quadrature-mode adjoint supports Dunavant degrees 0 to 7"), FRICO_ERR_LIMIT;
    auto& A = c->adj;
    A = frico_ctx::AdjointState{};
    A.approx = approx;
    const int64_t N = c->N;
    std::vector<double> signs((size_t)N);
    int applied = 0;
    int rc = frico_reorient_deltagap(c, ifaces, nif, 0, signs.data(), &applied);
    if (rc)
        return rc;
    if (applied)
        A.signs = signs;
    if ((rc = frico_fill(c, freq, c->adj_fill, FRICO_C128)) ||
        (rc = frico_rhs_deltagap(c, freq, ifaces, nif, v_re, v_im, applied ? signs.data() : nullptr)))
        return rc;
    A.V.resize((size_t)N);
    if ((rc = frico_rhs_get(c, reinterpret_cast<double*>(A.V.data()), N)))
        return rc;
    if ((rc = frico_factor(c, FRICO_SOLVE_C128)) || (rc = frico_solve(c, 0.0, 1, nullptr)))
        return rc;
    A.I.resize((size_t)N);
    if ((rc = frico_solution_get(c, reinterpret_cast<double*>(A.I.data()), N)))
        return rc;
    A.freq = freq;
    A.ifaces.assign(ifaces, ifaces + nif);
    A.Vg = cd(v_re, v_im);
    A.c.assign((size_t)N, cd(0));
    for (int64_t n = 0; n < N; n++) {
        const frico_rwg& b = c->hm.bf[n];
        bool gap = false;
        for (int f = 0; f < nif; f++)
            gap |= b.has_interface && b.interface_tag == ifaces[f];
        if (gap)
            A.c[n] = b.length * (A.signs.empty() ? 1.0 : A.signs[n]);
    }
    for (int64_t n = 0; n < N; n++)
        A.Iport += A.c[n] * A.I[n];
    A.zin = A.Vg / A.Iport;
    for (int d = 0; d < 3; d++)
        A.obs[d] = obs[d];
    A.E.resize(3);
    A.H.resize(3);
    if ((rc = frico_fields(c, freq, obs, 1, nullptr, 1.0, reinterpret_cast<double*>(A.E.data()),
                           reinterpret_cast<double*>(A.H.data()))))
        return rc;
    const cd Hc[3] = {std::conj(A.H[0]), std::conj(A.H[1]), std::conj(A.H[2])};
    const cd S[3] = {0.5 * (A.E[1] * Hc[2] - A.E[2] * Hc[1]), 0.5 * (A.E[2] * Hc[0] - A.E[0] * Hc[2]),
                     0.5 * (A.E[0] * Hc[1] - A.E[1] * Hc[0])};
    A.absS = std::sqrt(std::norm(S[0]) + std::norm(S[1]) + std::norm(S[2]));
    A.P = (0.5 * A.Vg * std::conj(A.Iport)).real();
    const double r2 = obs[0] * obs[0] + obs[1] * obs[1] + obs[2] * obs[2];
    A.gain = 4 * PI * r2 * A.absS / A.P;
    out[0] = A.zin.real();
    out[1] = A.zin.imag();
    out[2] = A.gain;
    A.forward_done = true;
    return 0;
}

int frico_adjoint_rcs_forward(frico_ctx* c, double freq, double theta_i, double phi_i, const double* eth,
                              const double* eph, double theta_s, double phi_s, double R, double* out)
{
    if (!c)
        return set_error("null context"), -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!eth || !eph)
        return set_error("null polarisation amplitudes"), -5;
    if (!(R > 0))
        return set_error("observation distance must be positive"), -9;
    if (!out)
        return set_error("null output"), -10;
    const bool approx = c->adj_fill == FRICO_FILL_APPROX;
    if (!approx && c->hm.degree > 7)
        return set_error("quadrature-mode adjoint supports Dunavant degrees 0 to 7"), FRICO_ERR_LIMIT;
    auto& A = c->adj;
    A = frico_ctx::AdjointState{};
    A.approx = approx;
    const int64_t N = c->N;
    int rc;
    if ((rc = frico_fill(c, freq, c->adj_fill, FRICO_C128)) ||
        (rc = frico_rhs_planewave(c, freq, 1, &theta_i, &phi_i, eth, eph)))
        return rc;
    A.V.resize((size_t)N);
    if ((rc = frico_rhs_get(c, reinterpret_cast<double*>(A.V.data()), N)))
        return rc;
    if ((rc = frico_factor(c, FRICO_SOLVE_C128)) || (rc = frico_solve(c, 0.0, 1, nullptr)))
        return rc;
    A.I.resize((size_t)N);
    if ((rc = frico_solution_get(c, reinterpret_cast<double*>(A.I.data()), N)))
        return rc;
    A.freq = freq;
    A.theta_i = theta_i;
    A.phi_i = phi_i;
    A.theta_s = theta_s;
    A.phi_s = phi_s;
    A.R = R;
    A.eth = cd(eth[0], eth[1]);
    A.eph = cd(eph[0], eph[1]);
    // the observation point sph2rect(R, theta_s, phi_s) (src/utils.h:53-60)
    A.obs[0] = R * std::cos(phi_s) * std::sin(theta_s);
    A.obs[1] = R * std::sin(phi_s) * std::sin(theta_s);
    A.obs[2] = R * std::cos(theta_s);
    A.E.resize(3);
    A.H.resize(3);
    if ((rc = frico_fields(c, freq, A.obs, 1, nullptr, 1.0, reinterpret_cast<double*>(A.E.data()),
                           reinterpret_cast<double*>(A.H.data()))))
        return rc;
    // bistatic_RCS_unit_Ei (src/emw_ana_bistatic_rcs.cpp:62-82)
    const double th[3] = {std::cos(theta_s) * std::cos(phi_s), std::cos(theta_s) * std::sin(phi_s), -std::sin(theta_s)};
    const double ph[3] = {-std::sin(phi_s), std::cos(phi_s), 0.0};
    A.a_vv = th[0] * A.E[0] + th[1] * A.E[1] + th[2] * A.E[2];
    A.a_hh = ph[0] * A.E[0] + ph[1] * A.E[1] + ph[2] * A.E[2];
    out[0] = 4.0 * PI * R * R * std::norm(A.a_vv);
    out[1] = 4.0 * PI * R * R * std::norm(A.a_hh);
    A.rcs_done = true;
    return 0;
}

int frico_adjoint_rcs_backward(frico_ctx* c, double g_vv, double g_hh, double* grad)
{
    if (!c)
        return set_error("null context"), -1;
    if (!grad)
        return set_error("null output"), -4;
    auto& A = c->adj;
    if (!A.rcs_done)
        return set_error("backward before forward"), FRICO_ERR_STATE;
    if (!c->have_x || c->factored != FRICO_SOLVE_C128)
        return set_error("the forward factors were replaced: run forward again"), FRICO_ERR_STATE;
    const HostMesh& hm = c->hm;
    const int64_t N = c->N, nv = hm.nv;
    hipStream_t st = c->st;
    int rc;
    // d f / d E (Wirtinger) of f = g_vv sigma_VV + g_hh sigma_HH, sigma = 4 pi R^2 |component|^2
    const double th[3] = {std::cos(A.theta_s) * std::cos(A.phi_s), std::cos(A.theta_s) * std::sin(A.phi_s),
                          -std::sin(A.theta_s)};
    const double ph[3] = {-std::sin(A.phi_s), std::cos(A.phi_s), 0.0};
    const double s4 = 4.0 * PI * A.R * A.R;
    cd eE[3], eH[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++)
        eE[d] = s4 * (g_vv * std::conj(A.a_vv) * th[d] + g_hh * std::conj(A.a_hh) * ph[d]);
    std::vector<cd> gx, gI;
    field_vjp(hm, A.freq, A.I, A.obs, eE, eH, gx, gI);
    // Z^T mu = gI on the complex128 factors
    if ((rc = c->adjW.ensure((size_t)N)) || (rc = c->aw.ensure((size_t)(3 * hm.nt))) ||
        (rc = c->bw.ensure((size_t)(3 * hm.nt))) || (rc = c->d_gvert.ensure((size_t)(6 * nv))))
        return rc;
    FRICO_HIP_CHECK(hipMemcpyAsync(c->adjW.p, gI.data(), (size_t)N * sizeof(c128), hipMemcpyHostToDevice, st));
    if ((rc = lu_solve_transposed<double>(c->Z64.p, N, N, c->ipiv.p, c->nb_factored, c->adjW.p, N, 1, st)))
        return rc;
    std::vector<cd> mu((size_t)N);
    FRICO_HIP_CHECK(hipMemcpyAsync(mu.data(), c->adjW.p, (size_t)N * sizeof(c128), hipMemcpyDeviceToHost, st));
    FRICO_HIP_CHECK(hipStreamSynchronize(st));
    // mu^T dZ I: the device contraction at fixed edge lengths, then the l_i l_j prefactor
    const double omega = 2.0 * PI * A.freq;
    const double k = omega * std::sqrt(MU0 * EPS0), inv_ksq = 1. / (omega * omega * MU0 * EPS0);
    if ((rc = contract(c, c->X.p, c->adjW.p, k, inv_ksq)))
        return rc;
    std::vector<cd> C((size_t)(3 * nv));
    FRICO_HIP_CHECK(hipMemcpyAsync(C.data(), c->d_gvert.p, (size_t)(6 * nv) * sizeof(double), hipMemcpyDeviceToHost, st));
    FRICO_HIP_CHECK(hipStreamSynchronize(st));
    for (int64_t n = 0; n < N; n++) {
        const double l = hm.bf[n].length;
        if (l != 0)
            add_length_grad(hm, hm.bf[n].edge_index, l, (A.I[n] * gI[n] + mu[n] * A.V[n]) / l, C);
    }
    // mu^T dV of the plane-wave moments
    std::vector<cd> dV((size_t)(3 * nv), cd(0));
    planewave_vjp(hm, A.freq, A.theta_i, A.phi_i, A.eth, A.eph, mu, dV);
    for (int64_t x = 0; x < 3 * nv; x++)
        grad[x] = 2.0 * (gx[x] + dV[x] - C[x]).real();
    return 0;
}

int frico_adjoint_backward(frico_ctx* c, double g_zin_re, double g_zin_im, double g_gain, double* grad)
{
    if (!c)
        return set_error("null context"), -1;
    if (!grad)
        return set_error("null output"), -5;
    auto& A = c->adj;
    if (!A.forward_done)
        return set_error("backward before forward"), FRICO_ERR_STATE;
    if (!c->have_x || c->factored != FRICO_SOLVE_C128)
        return set_error("the forward factors were replaced: run forward again"), FRICO_ERR_STATE;
    int rc;
    if (!A.grads_done && (rc = compute_gradients(c)))
        return rc;
    for (size_t x = 0; x < A.dgain.size(); x++)
        grad[x] = g_zin_re * A.dzin[x].real() + g_zin_im * A.dzin[x].imag() + g_gain * A.dgain[x];
    return 0;
}

} // extern "C"
