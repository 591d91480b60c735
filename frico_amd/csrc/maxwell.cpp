// C++ host mirror of the reference's analysis drivers on top of the C ABI.
//
//   frico::maxwell::run_context   <- run_context<Source> (src/emw_solver.h:92-167)
//   frico::maxwell::do_sweep      <- do_sweep(simulation&, antenna_analysis&)
//                                    (src/emw_ana_antenna.cpp:236-295), do_sweep(simulation&,
//                                    const bistatic_rcs_analysis&) (src/emw_ana_bistatic_rcs.cpp:120-151)
//   compute_gain, compute_port_values, bistatic_RCS_unit_Ei mirror src/emw_ana_antenna.cpp:33-105
//   and src/emw_ana_bistatic_rcs.cpp:62-82.
// Differences from the reference: the solve status is surfaced (the reference ignores it,
// src/emw_solver.h:126-131), and a monostatic multi-incidence sweep is added (the reference's
// -m flag is a no-op, src/main.cpp:92-93).
#include "../../include/frico_gpu.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace frico::maxwell {

using cd = std::complex<double>;

struct config {
    int fill_mode = FRICO_FILL_QUAD;
    bool force_symmetry = false;
    bool force_reorient_deltagap = false;
    double Z0 = 50.0;
    int form = FRICO_SOLVE_C128;
    double tol = 1e-12;
    int maxit = 10;
};

struct simulation {
    config cfg;
    frico_ctx* ctx = nullptr;
    std::vector<double> delta_gap_signs;
};

struct delta_gap {
    std::vector<int> interfaces;
    cd voltage = 1.0;
};

struct port_values {
    cd voltage = 0.0, current = 0.0, power = 0.0, impedance = 0.0, gamma = 0.0;
    double swr = 0.0;
};

static cd gamma_of(cd Z, double Z0) { return (Z - Z0) / (Z + Z0); }
static double swr_of(cd g) { return (1.0 + std::abs(g)) / (1.0 - std::abs(g)); }

// run_context solves approximate-mode and symmetrised systems with solve_symmetric
// (src/emw_solver.h:124-128): the refinement form takes the symmetric factorisation there.
static int solve_form(int form, int fill_mode, bool force_symmetry)
{
    return form == FRICO_SOLVE_C64_IR && (fill_mode == FRICO_FILL_APPROX || force_symmetry) ? FRICO_SOLVE_C64_SYM_IR
                                                                                           : form;
}

// Delta-gap RHS, the solve in the configured form, and the solve status.
static int run_context(simulation& sim, double freq, const delta_gap& dg, frico_solve_info* info)
{
    int rc = frico_fill(sim.ctx, freq, sim.cfg.fill_mode, FRICO_C128);
    if (rc)
        return rc;
    if (sim.cfg.force_symmetry && (rc = frico_symmetrize(sim.ctx)))
        return rc;
    std::vector<int32_t> ifa(dg.interfaces.begin(), dg.interfaces.end());
    rc = frico_rhs_deltagap(sim.ctx, freq, ifa.data(), (int)ifa.size(), dg.voltage.real(), dg.voltage.imag(),
                            sim.delta_gap_signs.empty() ? nullptr : sim.delta_gap_signs.data());
    if (rc)
        return rc;
    if ((rc = frico_factor(sim.ctx, solve_form(sim.cfg.form, sim.cfg.fill_mode, sim.cfg.force_symmetry))))
        return rc;
    return frico_solve(sim.ctx, sim.cfg.tol, sim.cfg.maxit, info);
}

// compute_port_values: src/emw_ana_antenna.cpp:76-105
static port_values compute_port_values(const simulation& sim, const std::vector<frico_rwg>& bf, const cd* I,
                                       const delta_gap& dg)
{
    port_values pv;
    for (size_t n = 0; n < bf.size(); n++) {
        if (!bf[n].has_interface)
            continue;
        for (int itf : dg.interfaces)
            if (bf[n].interface_tag == itf) {
                const double sign = sim.delta_gap_signs.empty() ? 1.0 : sim.delta_gap_signs[n];
                const cd Ie = bf[n].length * I[n] * sign;
                pv.current += Ie;
                pv.power += 0.5 * dg.voltage * std::conj(Ie);
            }
    }
    pv.voltage = dg.voltage;
    pv.impedance = dg.voltage / pv.current;
    pv.gamma = gamma_of(pv.impedance, sim.cfg.Z0);
    pv.swr = swr_of(pv.gamma);
    return pv;
}

// compute_gain on the xy, yz and xz cuts (src/emw_ana_antenna.cpp:33-65), all 1080 points in one
// field launch.  G = 4 pi |P|^2 |S| / pwr with S = E x conj(H) / 2.
static int compute_gain(simulation& sim, double freq, double pwr, double R, double* gains)
{
    std::vector<double> pts(3 * 1080);
    for (int plane = 0; plane < 3; plane++)
        for (int deg = 0; deg < 360; deg++) {
            const double th = deg * M_PI / 180;
            const double c = std::cos(th), s = std::sin(th);
            double* P = &pts[3 * (plane * 360 + deg)];
            if (plane == 0) {
                P[0] = R * c;
                P[1] = R * s;
                P[2] = 0.0;
            } else if (plane == 1) {
                P[0] = 0.0;
                P[1] = R * c;
                P[2] = R * s;
            } else {
                P[0] = -R * c;
                P[1] = 0.0;
                P[2] = R * s;
            }
        }
    std::vector<cd> E(3 * 1080), H(3 * 1080);
    int rc = frico_fields(sim.ctx, freq, pts.data(), 1080, nullptr, 1.0, reinterpret_cast<double*>(E.data()),
                          reinterpret_cast<double*>(H.data()));
    if (rc)
        return rc;
    for (int i = 0; i < 1080; i++) {
        const cd* e = &E[3 * i];
        const cd hc[3] = {std::conj(H[3 * i]), std::conj(H[3 * i + 1]), std::conj(H[3 * i + 2])};
        const cd S[3] = {0.5 * (e[1] * hc[2] - e[2] * hc[1]), 0.5 * (e[2] * hc[0] - e[0] * hc[2]),
                         0.5 * (e[0] * hc[1] - e[1] * hc[0])};
        const cd SS = std::conj(S[0]) * S[0] + std::conj(S[1]) * S[1] + std::conj(S[2]) * S[2];
        const double Prad = std::real(std::sqrt(SS));
        const double* P = &pts[3 * i];
        const double Rn = std::sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2]);
        gains[i] = 4 * M_PI * Rn * Rn * Prad / pwr;
    }
    return 0;
}

// bistatic_RCS_unit_Ei: src/emw_ana_bistatic_rcs.cpp:62-82
static void rcs_unit_Ei(const cd* Es, double theta_s, double phi_s, double R, double* vv, double* hh)
{
    const double th[3] = {std::cos(theta_s) * std::cos(phi_s), std::cos(theta_s) * std::sin(phi_s), -std::sin(theta_s)};
    const double ph[3] = {-std::sin(phi_s), std::cos(phi_s), 0.0};
    const cd Et = th[0] * Es[0] + th[1] * Es[1] + th[2] * Es[2];
    const cd Ep = ph[0] * Es[0] + ph[1] * Es[1] + ph[2] * Es[2];
    *vv = 4.0 * M_PI * R * R * std::norm(Et);
    *hh = 4.0 * M_PI * R * R * std::norm(Ep);
}

static void sph2rect(double R, double theta, double phi, double* p)
{
    p[0] = R * std::cos(phi) * std::sin(theta);
    p[1] = R * std::sin(phi) * std::sin(theta);
    p[2] = R * std::cos(theta);
}

} // namespace frico::maxwell

using namespace frico::maxwell;

extern "C" {

int frico_sweep_frequencies(double start, double end, double step, double* freqs, int64_t max_freqs, int64_t* count)
{
    if (!count)
        return -6;
    if (!(step > 0))
        return -3;
    // init_sweep (src/emw_solver.cpp:460-471): the frequency is accumulated, freq += step, and the
    // loop runs while freq <= end, so rounding can drop the last point of a sweep
    int64_t n = 0;
    for (double freq = start; freq <= end; freq += step) {
        if (freqs && n < max_freqs)
            freqs[n] = freq;
        n++;
        if (n > (int64_t)1 << 30)
            return FRICO_ERR_LIMIT;
    }
    *count = n;
    return (freqs && n > max_freqs) ? -5 : 0;
}

int frico_antenna_sweep(frico_ctx* ctx, const double* freqs, int nfreq, int mode, const int32_t* ifaces, int nif,
                        double v_re, double v_im, double R, double Z0, int force_symmetry, int force_reorient, int form,
                        double tol, int maxit, double* port_out, double* gain_out)
{
    if (!ctx || !freqs || nfreq < 1 || !ifaces || nif < 1)
        return -1;
    simulation sim;
    sim.ctx = ctx;
    sim.cfg.fill_mode = mode;
    sim.cfg.force_symmetry = force_symmetry != 0;
    sim.cfg.force_reorient_deltagap = force_reorient != 0;
    sim.cfg.Z0 = Z0;
    sim.cfg.form = form;
    sim.cfg.tol = tol;
    sim.cfg.maxit = maxit;
    const int64_t N = frico_num_unknowns(ctx);
    if (N < 1)
        return FRICO_ERR_STATE;
    delta_gap dg;
    dg.interfaces.assign(ifaces, ifaces + nif);
    dg.voltage = cd(v_re, v_im);
    // reorient_deltagap_edges once per sweep (src/emw_ana_antenna.cpp:238)
    std::vector<double> signs((size_t)N);
    int applied = 0;
    int rc = frico_reorient_deltagap(ctx, ifaces, nif, force_reorient, signs.data(), &applied);
    if (rc)
        return rc;
    if (applied)
        sim.delta_gap_signs = signs;
    std::vector<frico_rwg> bf((size_t)N);
    if ((rc = frico_rwg_table(ctx, bf.data())))
        return rc;
    std::vector<cd> I((size_t)N);
    for (int f = 0; f < nfreq; f++) {
        frico_solve_info info{};
        rc = run_context(sim, freqs[f], dg, &info);
        if (rc)
            return rc;
        if ((rc = frico_solution_get(ctx, reinterpret_cast<double*>(I.data()), N)))
            return rc;
        const port_values pv = compute_port_values(sim, bf, I.data(), dg);
        if (port_out) {
            double* o = port_out + 10 * f;
            o[0] = pv.current.real();
            o[1] = pv.current.imag();
            o[2] = pv.power.real();
            o[3] = pv.power.imag();
            o[4] = pv.impedance.real();
            o[5] = pv.impedance.imag();
            o[6] = pv.gamma.real();
            o[7] = pv.gamma.imag();
            o[8] = pv.swr;
            o[9] = info.max_rel_residual;
        }
        if (gain_out && (rc = compute_gain(sim, freqs[f], pv.power.real(), R, gain_out + 1080 * f)))
            return rc;
    }
    return 0;
}

int frico_bistatic_sweep(frico_ctx* ctx, const double* freqs, int nfreq, int mode, double theta_r, double phi_r, double eth,
                         double eph, double R, int form, double tol, int maxit, double* table_out)
{
    if (!ctx || !freqs || nfreq < 1 || !table_out)
        return -1;
    const double deg2rad = M_PI / 180.0;
    std::vector<double> pts(3 * 718);
    for (int i = 0; i < 359; i++) {
        sph2rect(R, deg2rad * i, 0.0, &pts[3 * (2 * i)]);
        sph2rect(R, 0.0, deg2rad * i, &pts[3 * (2 * i + 1)]);
    }
    const double th[1] = {theta_r}, ph[1] = {phi_r}, et[2] = {eth, 0.0}, ep[2] = {eph, 0.0};
    std::vector<cd> E(3 * 718);
    for (int f = 0; f < nfreq; f++) {
        int rc = frico_fill(ctx, freqs[f], mode, FRICO_C128);
        if (rc || (rc = frico_rhs_planewave(ctx, freqs[f], 1, th, ph, et, ep)) || (rc = frico_factor(ctx, frico::maxwell::solve_form(form, mode, false))) ||
            (rc = frico_solve(ctx, tol, maxit, nullptr)))
            return rc;
        if ((rc = frico_fields(ctx, freqs[f], pts.data(), 718, nullptr, 1.0, reinterpret_cast<double*>(E.data()), nullptr)))
            return rc;
        double* tab = table_out + 359 * 4 * f;
        for (int i = 0; i < 359; i++) {
            rcs_unit_Ei(&E[3 * (2 * i)], deg2rad * i, 0.0, R, &tab[4 * i], &tab[4 * i + 1]);
            rcs_unit_Ei(&E[3 * (2 * i + 1)], 0.0, deg2rad * i, R, &tab[4 * i + 2], &tab[4 * i + 3]);
        }
    }
    return 0;
}

int frico_monostatic_sweep(frico_ctx* ctx, const double* freqs, int nfreq, int mode, int64_t nang, const double* theta,
                           const double* phi, double eth, double eph, double R, int form, double tol, int maxit,
                           double* sigma_out, frico_solve_info* info_out)
{
    if (!ctx || !freqs || nfreq < 1 || nang < 1 || !theta || !phi || !sigma_out)
        return -1;
    std::vector<double> et(2 * nang), ep(2 * nang), pts(3 * nang);
    std::vector<int64_t> cols(nang);
    for (int64_t a = 0; a < nang; a++) {
        et[2 * a] = eth;
        et[2 * a + 1] = 0;
        ep[2 * a] = eph;
        ep[2 * a + 1] = 0;
        sph2rect(R, theta[a], phi[a], &pts[3 * a]);
        cols[a] = a;
    }
    std::vector<cd> E(3 * nang);
    for (int f = 0; f < nfreq; f++) {
        frico_solve_info info{};
        int rc = frico_fill(ctx, freqs[f], mode, FRICO_C128);
        if (rc || (rc = frico_rhs_planewave(ctx, freqs[f], nang, theta, phi, et.data(), ep.data())) ||
            (rc = frico_factor(ctx, frico::maxwell::solve_form(form, mode, false))) ||
            (rc = frico_solve(ctx, tol, maxit, &info)))
            return rc;
        if (info_out)
            info_out[f] = info;
        if ((rc = frico_fields(ctx, freqs[f], pts.data(), nang, cols.data(), 1.0, reinterpret_cast<double*>(E.data()),
                               nullptr)))
            return rc;
        double* s = sigma_out + 2 * nang * f;
        for (int64_t a = 0; a < nang; a++)
            rcs_unit_Ei(&E[3 * a], theta[a], phi[a], R, &s[2 * a], &s[2 * a + 1]);
    }
    return 0;
}

} // extern "C"
