"""Pins the FP64 CPU adjoint (oracle.antenna_gradients / oracle.adjoint_contract) by central finite
differences of the oracle's forward path.  The adjoint has no counterpart in the reference; its
forward pieces are the reference's (src/emw_solver.cpp:45-142, src/emw_ana_antenna.cpp:33-105)."""
import math

import numpy as np
import pytest

import oracle
from frico_amd import meshes

F0 = 300e6


def _perturbed(sm, V):
    return meshes.SurfaceMesh(V, sm.triangles, sm.tags, sm.bedges, sm.btags, feed=sm.feed)


@pytest.mark.parametrize("sm", [meshes.strip_dipole(0.3, 0.04, 4, 2), meshes.icosphere(0.4, 1)],
                         ids=lambda s: s.name)
def test_contraction_matches_central_differences(sm):
    om = oracle.OracleMesh(sm)
    rng = np.random.default_rng(1)
    I = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    W = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    g = oracle.adjoint_contract(om, F0, 5, I, W)
    h = 1e-6
    coords = [(v, c) for v in range(0, sm.nv, max(1, sm.nv // 5)) for c in range(3)]
    for v, c in coords:
        Vp = sm.vertices.copy()
        Vp[v, c] += h
        Vm = sm.vertices.copy()
        Vm[v, c] -= h
        Zp = oracle.OracleMesh(_perturbed(sm, Vp)).fill(F0, 5)
        Zm = oracle.OracleMesh(_perturbed(sm, Vm)).fill(F0, 5)
        fd = W @ ((Zp - Zm) / (2 * h)) @ I
        assert abs(fd - g[v, c]) <= 1e-7 * np.max(np.abs(g))


def test_impedance_and_gain_gradients_match_central_differences():
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 4, 3)
    r = oracle.antenna_gradients(sm, F0, 5, obs=(0.0, 0.0, 10.0))
    h = 1e-6 * lam

    def fwd(V):
        a = oracle.antenna_analysis(_perturbed(sm, V), F0, degree=5)
        m = a["mesh"]
        AJ, AdivJ, _ = m.scatter(a["I"])
        E, H = m.fields(F0, AJ, AdivJ, np.array([0.0, 0.0, 10.0]))
        S = 0.5 * np.cross(E[0], np.conj(H[0]))
        return a["port"]["impedance"], 4 * math.pi * 100.0 * math.sqrt(np.real(np.vdot(S, S))) / a["port"]["power"].real

    z0, g0 = fwd(sm.vertices)
    assert abs(z0 - r["zin"]) <= 1e-12 * abs(z0)
    assert abs(g0 - r["gain"]) <= 1e-12 * abs(g0)
    for v in range(0, sm.nv, 3):
        for c in range(3):
            Vp = sm.vertices.copy()
            Vp[v, c] += h
            Vm = sm.vertices.copy()
            Vm[v, c] -= h
            zp, gp = fwd(Vp)
            zm, gm = fwd(Vm)
            assert abs((zp - zm) / (2 * h) - r["dzin"][v, c]) <= 1e-6 * np.max(np.abs(r["dzin"]))
            assert abs((gp - gm) / (2 * h) - r["dgain"][v, c]) <= 1e-6 * np.max(np.abs(r["dgain"]))


@pytest.mark.parametrize("th,ph", [(0.0, 0.0), (0.6, 0.4), (2.2, -1.3)])
def test_planewave_rhs_vjp_matches_central_differences(th, ph):
    """sum_n W_n dV_n/dx of the plane-wave moments (rule points, weights, areas, edge lengths, free
    vertices and the field's phase all move with the vertices) against central differences of the
    oracle's own update_rhs(plane_wave)."""
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 4, 3)
    om = oracle.OracleMesh(sm)
    rng = np.random.default_rng(3)
    W = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    eth, eph = 1.0 - 0.3j, 0.4 + 0.2j
    g = oracle.rhs_planewave_vjp(om, F0, 5, th, ph, eth, eph, W)
    h = 1e-6 * lam
    for v in range(0, sm.nv, 2):
        for c in range(3):
            Vp, Vm = sm.vertices.copy(), sm.vertices.copy()
            Vp[v, c] += h
            Vm[v, c] -= h
            fp = W @ oracle.OracleMesh(_perturbed(sm, Vp)).rhs_planewave(F0, th, ph, eth, eph, 5)
            fm = W @ oracle.OracleMesh(_perturbed(sm, Vm)).rhs_planewave(F0, th, ph, eth, eph, 5)
            assert abs((fp - fm) / (2 * h) - g[v, c]) <= 1e-7 * np.max(np.abs(g))


@pytest.mark.parametrize("bistatic", [False, True])
def test_rcs_gradients_match_central_differences(bistatic):
    """sigma_VV and sigma_HH of a plate lit obliquely (theta-phi mixed polarisation), monostatic and
    bistatic, and their vertex gradients against central differences (step 1e-6 wavelengths) of the
    oracle's forward RCS."""
    lam = meshes.wavelength(F0)
    sm = meshes.plate(0.6 * lam, 3)
    thi, phi_ = 0.5, 0.3
    ths, phs = (1.1, -0.7) if bistatic else (thi, phi_)
    eth, eph = 1.0, 0.5j
    r = oracle.rcs_gradients(sm, F0, thi, phi_, eth, eph, ths, phs, degree=5, R=10.0)

    def fwd(V):
        s2 = _perturbed(sm, V)
        m = oracle.OracleMesh(s2)
        I, _ = oracle.lu_solve(m.fill(F0, 5), m.rhs_planewave(F0, thi, phi_, eth, eph, 5))
        AJ, AdivJ, _ = m.scatter(I)
        E, _ = m.fields(F0, AJ, AdivJ, oracle.sph2rect(10.0, ths, phs))
        return oracle.rcs_unit_Ei(E[0], ths, phs, 10.0)

    s0 = fwd(sm.vertices)
    assert np.allclose(s0, r["sigma"], rtol=1e-12)
    h = 1e-6 * lam
    for v in range(0, sm.nv, 3):
        for c in range(3):
            Vp, Vm = sm.vertices.copy(), sm.vertices.copy()
            Vp[v, c] += h
            Vm[v, c] -= h
            d = (fwd(Vp) - fwd(Vm)) / (2 * h)
            assert abs(d[0] - r["dsigma_vv"][v, c]) <= 1e-6 * np.max(np.abs(r["dsigma_vv"]))
            assert abs(d[1] - r["dsigma_hh"][v, c]) <= 1e-6 * np.max(np.abs(r["dsigma_hh"]))


# ---- approximate mode (compute_matrix_approx, src/emw_solver.cpp:144-220: barycenters, free
# vertices, the self terms' sqrt(area), hypot(dx, dy) lengths) -------------------------------------

@pytest.mark.parametrize("sm", [meshes.strip_dipole(0.3, 0.04, 4, 2), meshes.icosphere(0.4, 1),
                                meshes.bowtie(0.25, 60.0, 3, 2)], ids=lambda s: s.name)
def test_approx_contraction_matches_central_differences(sm):
    om = oracle.OracleMesh(sm)
    rng = np.random.default_rng(2)
    I = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    W = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    g = oracle.adjoint_contract(om, F0, 5, I, W, approx=True)
    h = 1e-6
    for v in range(0, sm.nv, max(1, sm.nv // 6)):
        for c in range(3):
            Vp, Vm = sm.vertices.copy(), sm.vertices.copy()
            Vp[v, c] += h
            Vm[v, c] -= h
            Zp = oracle.OracleMesh(_perturbed(sm, Vp)).fill(F0, 5, approx=True)
            Zm = oracle.OracleMesh(_perturbed(sm, Vm)).fill(F0, 5, approx=True)
            fd = W @ ((Zp - Zm) / (2 * h)) @ I
            assert abs(fd - g[v, c]) <= 1e-7 * np.max(np.abs(g))


def test_approx_contraction_is_not_the_quadrature_one():
    """The two modes' contractions differ (a mode switch that is ignored fails here)."""
    sm = meshes.strip_dipole(0.3, 0.04, 4, 2)
    om = oracle.OracleMesh(sm)
    rng = np.random.default_rng(4)
    I = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    W = rng.normal(size=om.N) + 1j * rng.normal(size=om.N)
    ga = oracle.adjoint_contract(om, F0, 5, I, W, approx=True)
    gq = oracle.adjoint_contract(om, F0, 5, I, W)
    assert np.max(np.abs(ga - gq)) > 1e-3 * np.max(np.abs(gq))


def test_approx_impedance_gradient_matches_central_differences():
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 4, 3)
    r = oracle.antenna_gradients(sm, F0, 5, obs=(0.0, 0.0, 10.0), approx=True)
    h = 1e-6 * lam

    def zin(V):
        s2 = _perturbed(sm, V)
        m = oracle.OracleMesh(s2)
        signs = m.reorient(s2.feed)
        I, _ = oracle.lu_solve(m.fill(F0, 5, approx=True), m.rhs_deltagap(F0, s2.feed, 1.0, signs))
        s = np.ones(m.N) if signs is None else signs
        gap = np.isin(m.iface, s2.feed) & (m.has_iface != 0)
        return 1.0 / (np.where(gap, m.length * s, 0.0) @ I)

    assert abs(zin(sm.vertices) - r["zin"]) <= 1e-12 * abs(r["zin"])
    for v in range(0, sm.nv, 3):
        for c in range(3):
            Vp, Vm = sm.vertices.copy(), sm.vertices.copy()
            Vp[v, c] += h
            Vm[v, c] -= h
            d = (zin(Vp) - zin(Vm)) / (2 * h)
            assert abs(d - r["dzin"][v, c]) <= 1e-6 * np.max(np.abs(r["dzin"]))
