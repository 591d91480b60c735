"""ctypes front-end of the FP64 CPU oracle (oracle/cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg.  The product package (frico_amd) never imports it.

How each function here is pinned (tests/test_oracle_*.py):
  OracleMesh (fill, rhs, scatter, fields, port values, gain, bistatic): see cpu_ref.cpp's header;
  lu_solve: scipy's zgetrf/zgetrs on random matrices, normwise backward error, the izamax pivot
      rule and the first-zero-pivot info;
  sy_solve: scipy's LAPACK zsytrf/zsytrs themselves (the routines solve_symmetric calls), held to a
      general solve on complex symmetric matrices, on a zero-diagonal matrix that forces 2 x 2
      pivots, and on the approximate-mode EFIE matrix;
  antenna_analysis / bistatic_analysis: the reference's dipole impedance table and physics pins.
  The vendor LAPACK's own rounding, which the reference inherits, stays unpinned (absent here).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfrico_oracle.so")
_lib = None

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.orc_mesh_build.argtypes = [_d, C.c_int64, _i64, _i32, C.c_int64, _i64, _i32, C.c_int64,
                                     C.POINTER(C.c_void_p)]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_mesh_num_rwg.argtypes = [C.c_void_p]
        L.orc_mesh_num_rwg.restype = C.c_int64
        L.orc_mesh_num_edges.argtypes = [C.c_void_p]
        L.orc_mesh_num_edges.restype = C.c_int64
        L.orc_mesh_rwg.argtypes = [C.c_void_p, _i64, _i64, _i32, _i32, _i64, _i32, _i32, _d, _d, _d,
                                   _d, _d]
        L.orc_mesh_edges.argtypes = [C.c_void_p, _i64]
        L.orc_dunavant.argtypes = [C.c_int, _d, _d, _d, _d, _d]
        L.orc_subtri.argtypes = [C.c_void_p, C.c_int64, C.c_int, _d, _d]
        L.orc_fill_quad.argtypes = [C.c_void_p, C.c_double, C.c_int, _d]
        L.orc_fill_approx.argtypes = [C.c_void_p, C.c_double, _d]
        L.orc_reorient_deltagap.argtypes = [C.c_void_p, _i32, C.c_int, C.c_int, _d,
                                            C.POINTER(C.c_int)]
        L.orc_rhs_deltagap.argtypes = [C.c_void_p, C.c_double, _i32, C.c_int, C.c_double,
                                       C.c_double, _d, _d]
        L.orc_rhs_planewave.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double, _d]
        L.orc_lu_solve.argtypes = [_d, C.c_int64, _d, C.c_int64]
        L.orc_scatter.argtypes = [C.c_void_p, _d, _d, _d, _d]
        L.orc_eval_fields.argtypes = [C.c_void_p, C.c_double, _d, _d, _d, C.c_int64, _d, _d]
        L.orc_port_values.argtypes = [C.c_void_p, _d, _i32, C.c_int, C.c_double, C.c_double, _d,
                                      C.c_double, _d]
        L.orc_gain.argtypes = [C.c_void_p, C.c_double, _d, _d, C.c_double, C.c_double, _d]
        L.orc_rcs_unit_Ei.argtypes = [_d, C.c_double, C.c_double, C.c_double, _d]
        L.orc_bistatic.argtypes = [C.c_void_p, C.c_double, _d, _d, C.c_double, _d]
        L.orc_sph2rect.argtypes = [C.c_double, C.c_double, C.c_double, _d]
        L.orc_plane_wave_field.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, _d, C.c_int64, _d]
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t)


class OracleMesh:
    """Mesh + RWG function space built by the oracle's restatement of the reference."""

    def __init__(self, sm):
        L = lib()
        self.sm = sm
        h = C.c_void_p()
        rc = L.orc_mesh_build(_p(sm.vertices), sm.nv, _p(sm.triangles, _i64), _p(sm.tags, _i32),
                              sm.nt, _p(sm.bedges, _i64), _p(sm.btags, _i32), sm.bedges.shape[0],
                              C.byref(h))
        if rc != 0:
            raise ValueError(f"oracle mesh build failed ({rc})")
        self.h = h
        self.N = L.orc_mesh_num_rwg(h)
        N = self.N
        self.itm = np.zeros(N, np.int64)
        self.itp = np.zeros(N, np.int64)
        self.lem = np.zeros(N, np.int32)
        self.lep = np.zeros(N, np.int32)
        self.edge = np.zeros(N, np.int64)
        self.has_iface = np.zeros(N, np.int32)
        self.iface = np.zeros(N, np.int32)
        self.length = np.zeros(N)
        self.Am = np.zeros(N)
        self.Ap = np.zeros(N)
        self.pm = np.zeros((N, 3))
        self.pp = np.zeros((N, 3))
        L.orc_mesh_rwg(h, _p(self.itm, _i64), _p(self.itp, _i64), _p(self.lem, _i32),
                       _p(self.lep, _i32), _p(self.edge, _i64), _p(self.has_iface, _i32),
                       _p(self.iface, _i32), _p(self.length), _p(self.Am), _p(self.Ap), _p(self.pm),
                       _p(self.pp))
        ne = L.orc_mesh_num_edges(h)
        self.edges = np.zeros((ne, 2), np.int64)
        L.orc_mesh_edges(h, _p(self.edges, _i64))

    def __del__(self):
        try:
            lib().orc_mesh_free(self.h)
        except Exception:
            pass

    def fill(self, freq, degree=5, approx=False):
        Z = np.zeros((self.N, self.N), np.complex128, order="F")
        if approx:
            rc = lib().orc_fill_approx(self.h, freq, _p(Z))
        else:
            rc = lib().orc_fill_quad(self.h, freq, degree, _p(Z))
        if rc:
            raise ValueError(f"fill failed ({rc})")
        return Z

    def reorient(self, ifaces, force=False):
        signs = np.ones(self.N)
        applied = C.c_int(0)
        ifa = np.asarray(ifaces, np.int32)
        rc = lib().orc_reorient_deltagap(self.h, _p(ifa, _i32), len(ifa), int(force), _p(signs),
                                         C.byref(applied))
        if rc:
            raise ValueError("delta-gap edges are not connected")
        return signs if applied.value else None

    def rhs_deltagap(self, freq, ifaces, voltage=1.0, signs=None):
        V = np.zeros(self.N, np.complex128)
        ifa = np.asarray(ifaces, np.int32)
        lib().orc_rhs_deltagap(self.h, freq, _p(ifa, _i32), len(ifa), complex(voltage).real,
                               complex(voltage).imag, None if signs is None else _p(signs),
                               _p(V))
        return V

    def rhs_planewave(self, freq, theta, phi, eth=1.0, eph=0.0, degree=5):
        V = np.zeros(self.N, np.complex128)
        rc = lib().orc_rhs_planewave(self.h, freq, degree, theta, phi, complex(eth).real,
                                     complex(eth).imag, complex(eph).real, complex(eph).imag,
                                     _p(V))
        if rc:
            raise ValueError("bad degree")
        return V

    def scatter(self, I):
        I = np.ascontiguousarray(I, np.complex128)
        T = self.sm.nt
        AJ = np.zeros((T, 3), np.complex128)
        AdivJ = np.zeros(T, np.complex128)
        J = np.zeros((T, 3), np.complex128)
        lib().orc_scatter(self.h, _p(I), _p(AJ),
                          _p(AdivJ), _p(J))
        return AJ, AdivJ, J

    def fields(self, freq, AJ, AdivJ, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        E = np.zeros((pts.shape[0], 3), np.complex128)
        H = np.zeros((pts.shape[0], 3), np.complex128)
        AJ = np.ascontiguousarray(AJ, np.complex128)
        AdivJ = np.ascontiguousarray(AdivJ, np.complex128)
        lib().orc_eval_fields(self.h, freq, _p(AJ),
                              _p(AdivJ), _p(pts), pts.shape[0],
                              _p(E), _p(H))
        return E, H

    def port_values(self, I, ifaces, voltage=1.0, signs=None, Z0=50.0):
        I = np.ascontiguousarray(I, np.complex128)
        ifa = np.asarray(ifaces, np.int32)
        out = np.zeros(10)
        lib().orc_port_values(self.h, _p(I), _p(ifa, _i32), len(ifa),
                              complex(voltage).real, complex(voltage).imag,
                              None if signs is None else _p(signs), Z0, _p(out))
        return {"current": complex(out[0], out[1]), "power": complex(out[2], out[3]),
                "impedance": complex(out[4], out[5]), "gamma": complex(out[6], out[7]),
                "swr": out[8]}

    def gain(self, freq, AJ, AdivJ, pwr, R=10.0):
        g = np.zeros(1080)
        AJ = np.ascontiguousarray(AJ, np.complex128)
        AdivJ = np.ascontiguousarray(AdivJ, np.complex128)
        lib().orc_gain(self.h, freq, _p(AJ), _p(AdivJ),
                       pwr, R, _p(g))
        return g.reshape(3, 360)

    def bistatic(self, freq, AJ, AdivJ, R=10.0):
        tab = np.zeros((359, 4))
        AJ = np.ascontiguousarray(AJ, np.complex128)
        AdivJ = np.ascontiguousarray(AdivJ, np.complex128)
        lib().orc_bistatic(self.h, freq, _p(AJ),
                           _p(AdivJ), R, _p(tab.reshape(-1)))
        return tab


def dunavant(degree, p0, p1, p2):
    pts = np.zeros((100, 3))
    w = np.zeros(100)
    n = lib().orc_dunavant(degree, _p(np.asarray(p0, np.float64)), _p(np.asarray(p1, np.float64)),
                           _p(np.asarray(p2, np.float64)), _p(pts), _p(w))
    if n < 0:
        raise ValueError("degree out of range")
    return pts[:n], w[:n]


def lu_solve(A, B):
    """zgetrf/zgetrs restatement; returns (X, info)."""
    A = np.array(A, np.complex128, order="F", copy=True)
    B2 = np.array(B, np.complex128, order="F", copy=True)
    if B2.ndim == 1:
        B2 = B2.reshape(-1, 1, order="F")
    info = lib().orc_lu_solve(_p(A), A.shape[0],
                              _p(B2), B2.shape[1])
    return (B2[:, 0] if np.ndim(B) == 1 else B2), info


def sy_solve(A, B):
    """solve_symmetric (src/lapack_wrappers.cpp:135-177): LAPACK zsytrf('U') + zsytrs on a complex
    symmetric A, through scipy's LAPACK (the reference links MKL or a system LAPACK; the routine is
    the same published Bunch-Kaufman factorisation).  Returns (X, info)."""
    from scipy.linalg import lapack
    A = np.array(A, np.complex128, order="F", copy=True)
    B2 = np.array(B, np.complex128, order="F", copy=True)
    if B2.ndim == 1:
        B2 = B2.reshape(-1, 1, order="F")
    n = A.shape[0]
    # workspace query, then the factorisation, as the reference does
    _, _, info = lapack.zsytrf(A, lower=0, lwork=-1)
    if info != 0:
        return (B2[:, 0] if np.ndim(B) == 1 else B2), info
    ldu, ipiv, info = lapack.zsytrf(A, lower=0, lwork=max(1, 64 * n))
    if info != 0:
        return (B2[:, 0] if np.ndim(B) == 1 else B2), info
    X, info = lapack.zsytrs(ldu, ipiv, B2, lower=0)
    return (X[:, 0] if np.ndim(B) == 1 else X), info


def plane_wave_field(freq, theta, phi, eth, eph, pts):
    """plane_wave's incident field (src/sources.h:34-75) at pts (n x 3): returns E (n x 3 complex)."""
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    E = np.zeros((pts.shape[0], 3), np.complex128)
    eth, eph = complex(eth), complex(eph)
    lib().orc_plane_wave_field(freq, theta, phi, eth.real, eth.imag, eph.real, eph.imag, _p(pts), pts.shape[0], _p(E))
    return E


def sph2rect(R, theta, phi):
    p = np.zeros(3)
    lib().orc_sph2rect(R, theta, phi, _p(p))
    return p


def rcs_unit_Ei(Es, theta_s, phi_s, R):
    out = np.zeros(2)
    Es = np.ascontiguousarray(Es, np.complex128)
    lib().orc_rcs_unit_Ei(_p(Es), theta_s, phi_s, R, _p(out))
    return out


# ---- end-to-end reference flows (restating do_sweep for one frequency) ------------------------

def antenna_analysis(sm, freq, degree=5, approx=False, voltage=1.0, R=10.0, Z0=50.0,
                     force_reorient=False, force_symmetry=False):
    """One frequency of do_sweep(sim, antenna_analysis&): src/emw_ana_antenna.cpp:236-272,
    run_context src/emw_solver.h:92-167, postpro_context src/emw_ana_antenna.cpp:107-149."""
    m = OracleMesh(sm)
    signs = m.reorient(sm.feed, force_reorient)
    Z = m.fill(freq, degree, approx)
    V = m.rhs_deltagap(freq, sm.feed, voltage, signs)
    if force_symmetry:
        Z = (Z + Z.T) / 2.0
    # run_context's dispatch (src/emw_solver.h:118-132): the symmetric solve for the approximate
    # matrix or a forced symmetry, the general one otherwise
    I, info = (sy_solve if approx or force_symmetry else lu_solve)(Z, V)
    AJ, AdivJ, _ = m.scatter(I)
    port = m.port_values(I, sm.feed, voltage, signs, Z0)
    gains = m.gain(freq, AJ, AdivJ, port["power"].real, R)
    return {"mesh": m, "I": I, "V": V, "port": port, "gain": gains, "signs": signs, "info": info}


def bistatic_analysis(sm, freq, theta, phi, eth=1.0, eph=1.0, degree=5, approx=False, R=10.0):
    """One frequency of do_sweep(sim, bistatic_rcs_analysis&): src/emw_ana_bistatic_rcs.cpp:120-151."""
    m = OracleMesh(sm)
    Z = m.fill(freq, degree, approx)
    V = m.rhs_planewave(freq, theta, phi, eth, eph, degree)
    I, info = (sy_solve if approx else lu_solve)(Z, V)
    AJ, AdivJ, _ = m.scatter(I)
    tab = m.bistatic(freq, AJ, AdivJ, R)
    return {"mesh": m, "I": I, "V": V, "table": tab, "info": info}


# ---- adjoint reference (no counterpart in the reference code) -------------------------------

def _adj_sigs():
    L = lib()
    L.orc_adjoint_contract.argtypes = [C.c_void_p, C.c_double, C.c_int, _d, _d, _d]
    L.orc_adjoint_contract_approx.argtypes = [C.c_void_p, C.c_double, _d, _d, _d]
    L.orc_rhs_planewave_vjp.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_double, _d, _d]
    L.orc_fields_vjp.argtypes = [C.c_void_p, C.c_double, _d, _d, _d, _d, _d, _d]
    L.orc_length_grad.argtypes = [C.c_void_p, C.c_int64, _d]
    return L


def adjoint_contract(om, freq, degree, I, W, approx=False):
    """g[v, c] = sum_ij W_j I_i dZ(j,i)/dx_{v,c} (complex), quadrature-mode Z, or the
    approximate-mode Z of compute_matrix_approx (degree unused) when approx."""
    L = _adj_sigs()
    I = np.ascontiguousarray(I, np.complex128)
    W = np.ascontiguousarray(W, np.complex128)
    g = np.zeros(om.sm.nv * 3, np.complex128)
    if approx:
        rc = L.orc_adjoint_contract_approx(om.h, freq, _p(I), _p(W), _p(g))
    else:
        rc = L.orc_adjoint_contract(om.h, freq, degree, _p(I), _p(W), _p(g))
    if rc:
        raise ValueError("bad degree")
    return g.reshape(-1, 3)


def antenna_gradients(sm, freq, degree=5, obs=(0.0, 0.0, 10.0), voltage=1.0, approx=False):
    """FP64 CPU reference of the adjoint: input impedance Zin and gain G at the observation point
    `obs`, with dZin/dx (complex) and dG/dx (real) for every vertex coordinate.

    Forward: Z I = V (delta gap), Zin = V / sum_gap l s I, G = 4 pi |obs|^2 |S| / P with
    S = E x conj(H) / 2 (src/emw_ana_antenna.cpp:33-105).  Adjoint: Z^T lam = c (c = l s on the
    gap), Z^T mu = d|S|/dI (Wirtinger); dI_port = (dc) I + lam^T (dV - dZ I).  approx: the
    approximate-mode Z (the delta-gap moments do not depend on it)."""
    L = _adj_sigs()
    m = OracleMesh(sm)
    signs = m.reorient(sm.feed)
    s = np.ones(m.N) if signs is None else signs
    Z = m.fill(freq, degree, approx)
    V = m.rhs_deltagap(freq, sm.feed, voltage, signs)
    I, info = lu_solve(Z, V)
    gap = np.isin(m.iface, sm.feed) & (m.has_iface != 0)
    c = np.where(gap, m.length * s, 0.0).astype(np.complex128)
    Iport = c @ I
    Vg = complex(voltage)
    zin = Vg / Iport
    omega = 2 * math.pi * freq
    mu0 = 1.256637061435917e-06
    # fields at the observation point and |S|
    AJ, AdivJ, _ = m.scatter(I)
    obs = np.asarray(obs, np.float64)
    E, H = m.fields(freq, AJ, AdivJ, obs)
    E, H = E[0], H[0]
    S = 0.5 * np.cross(E, np.conj(H))
    absS = math.sqrt(float(np.real(np.vdot(S, S))))
    P = (0.5 * Vg * np.conj(Iport)).real
    G = 4 * math.pi * float(obs @ obs) * absS / P
    # Wirtinger derivatives of |S| w.r.t. E and H
    eE = 0.5 * np.cross(np.conj(H), np.conj(S)) / (2 * absS)
    eH = 0.5 * np.cross(S, np.conj(E)) / (2 * absS)
    gx_f = np.zeros(sm.nv * 3, np.complex128)
    gI = np.zeros(m.N, np.complex128)
    L.orc_fields_vjp(m.h, freq, _p(np.ascontiguousarray(I)), _p(obs), _p(np.ascontiguousarray(eE)),
                     _p(np.ascontiguousarray(eH)), _p(gx_f), _p(gI))
    lam = np.linalg.solve(Z.T, c)
    mu = np.linalg.solve(Z.T, gI)
    Clam = adjoint_contract(m, freq, degree, I, lam, approx).reshape(-1)
    Cmu = adjoint_contract(m, freq, degree, I, mu, approx).reshape(-1)
    # edge-length terms of V and of the port weights
    dc_I = np.zeros(sm.nv * 3, np.complex128)
    lam_dV = np.zeros(sm.nv * 3, np.complex128)
    mu_dV = np.zeros(sm.nv * 3, np.complex128)
    gl = np.zeros(sm.nv * 3)
    for n in np.nonzero(gap)[0]:
        L.orc_length_grad(m.h, int(n), _p(gl))
        dVn = -1j * Vg * s[n] / (omega * mu0)  # dV_n / dl_n
        dc_I += s[n] * I[n] * gl
        lam_dV += lam[n] * dVn * gl
        mu_dV += mu[n] * dVn * gl
    dIport = dc_I + lam_dV - Clam
    dzin = -Vg / Iport ** 2 * dIport
    dP = (0.5 * Vg * np.conj(dIport)).real
    dabsS = 2 * (gx_f.real) + 2 * (mu_dV - Cmu).real
    dG = 4 * math.pi * float(obs @ obs) / P * dabsS - G / P * dP
    return {"zin": zin, "gain": G, "dzin": dzin.reshape(-1, 3), "dgain": dG.reshape(-1, 3), "I": I,
            "lam": lam, "mu": mu, "N": m.N}


def rhs_planewave_vjp(om, freq, degree, theta, phi, eth, eph, W):
    """g[v, c] = sum_n W_n dV_n/dx_{v,c} (complex) for the plane-wave moments of rhs_planewave."""
    L = _adj_sigs()
    W = np.ascontiguousarray(W, np.complex128)
    g = np.zeros(om.sm.nv * 3, np.complex128)
    eth, eph = complex(eth), complex(eph)
    rc = L.orc_rhs_planewave_vjp(om.h, freq, degree, theta, phi, eth.real, eth.imag, eph.real, eph.imag, _p(W), _p(g))
    if rc:
        raise ValueError("bad degree")
    return g.reshape(-1, 3)


def rcs_gradients(sm, freq, theta_i, phi_i, eth=1.0, eph=0.0, theta_s=None, phi_s=None, degree=5, R=10.0,
                  approx=False):
    """FP64 CPU reference of the RCS adjoint: sigma_VV and sigma_HH at (R, theta_s, phi_s) for the
    plane wave from (theta_i, phi_i) (bistatic_RCS_unit_Ei, src/emw_ana_bistatic_rcs.cpp:62-82, of
    the currents of update_rhs(plane_wave), src/emw_solver.cpp:380-416, and compute_matrix), with
    their gradients with respect to every vertex coordinate.  Monostatic when theta_s, phi_s are
    omitted.

    sigma = 4 pi R^2 |a|^2 with a the theta (or phi) component of E_s; with Z^T mu = dE-cotangent
    pulled back to the currents, d sigma = 2 Re(g_x + mu^T dV - mu^T dZ I), g_x the explicit
    vertex dependence of the field sums.  approx: the approximate-mode Z (the plane-wave moments
    keep the degree's rule, as update_rhs does)."""
    L = _adj_sigs()
    theta_s = theta_i if theta_s is None else theta_s
    phi_s = phi_i if phi_s is None else phi_s
    m = OracleMesh(sm)
    Z = m.fill(freq, degree, approx)
    V = m.rhs_planewave(freq, theta_i, phi_i, eth, eph, degree)
    I, info = lu_solve(Z, V)
    AJ, AdivJ, _ = m.scatter(I)
    obs = sph2rect(R, theta_s, phi_s)
    E, _ = m.fields(freq, AJ, AdivJ, obs)
    E = E[0]
    th = np.array([math.cos(theta_s) * math.cos(phi_s), math.cos(theta_s) * math.sin(phi_s), -math.sin(theta_s)])
    ph = np.array([-math.sin(phi_s), math.cos(phi_s), 0.0])
    a, b = th @ E, ph @ E
    sig = 4 * math.pi * R * R * np.array([abs(a) ** 2, abs(b) ** 2])
    out = {"sigma": sig, "I": I, "N": m.N}
    for key, vec, amp in (("dsigma_vv", th, a), ("dsigma_hh", ph, b)):
        eE = 4 * math.pi * R * R * np.conj(amp) * vec.astype(np.complex128)
        gx = np.zeros(sm.nv * 3, np.complex128)
        gI = np.zeros(m.N, np.complex128)
        L.orc_fields_vjp(m.h, freq, _p(np.ascontiguousarray(I)), _p(np.ascontiguousarray(obs)),
                         _p(np.ascontiguousarray(eE)), _p(np.zeros(3, np.complex128)), _p(gx), _p(gI))
        mu = np.linalg.solve(Z.T, gI)
        dV = rhs_planewave_vjp(m, freq, degree, theta_i, phi_i, eth, eph, mu).reshape(-1)
        Cmu = adjoint_contract(m, freq, degree, I, mu, approx).reshape(-1)
        out[key] = (2 * (gx + dV - Cmu).real).reshape(-1, 3)
    return out
