"""Pins the oracle's dense solve (zgetrf/zgetrs semantics, src/lapack_wrappers.cpp:53-133), its field
sums (src/emw_postpro_common.cpp:31-66) by the Hertzian-dipole closed form, and the end-to-end
physics against the reference-held dipole table (share/examples/dipole/comparison/dipole_frico.txt)."""
import math

import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from frico_amd import meshes

MU0 = 1.256637061435917e-06
EPS0 = 8.8541878188e-12
ETA0 = math.sqrt(MU0 / EPS0)
F0 = 300e6


@pytest.mark.parametrize("n", [1, 2, 17, 90])
def test_lu_solve_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    B = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    X, info = oracle.lu_solve(A, B)
    assert info == 0
    Xs = sla.lu_solve(sla.lu_factor(A), B)
    assert np.max(np.abs(X - Xs)) <= 1e-12 * np.max(np.abs(Xs))
    # normwise backward error of a stable LU
    r = np.linalg.norm(B - A @ X) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert r <= 1e-14


@pytest.mark.parametrize("n", [1, 7, 64, 200])
def test_symmetric_solve_matches_a_general_solve(n):
    """sy_solve (zsytrf('U') + zsytrs, the reference's solve_symmetric) on complex symmetric
    (not Hermitian) matrices agrees with a general solve."""
    rng = np.random.default_rng(100 + n)
    M = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    A = M + M.T + 2 * np.sqrt(n) * np.eye(n)
    B = rng.uniform(-1, 1, (n, 3)) + 1j * rng.uniform(-1, 1, (n, 3))
    X, info = oracle.sy_solve(A, B)
    assert info == 0
    assert np.max(np.abs(X - np.linalg.solve(A, B))) <= 1e-11 * np.max(np.abs(X))


def test_symmetric_solve_takes_two_by_two_pivots():
    """A zero diagonal leaves Bunch-Kaufman no 1x1 pivot at the first step; the factorisation takes
    2x2 blocks (negative pivot entries) and the solution still matches."""
    from scipy.linalg import lapack
    n = 40
    rng = np.random.default_rng(3)
    M = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n))
    A = M + M.T
    A[np.arange(n), np.arange(n)] = 0.0
    _, ipiv, info = lapack.zsytrf(A, lower=0)
    assert info == 0 and np.any(ipiv < 0)
    b = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    x, info = oracle.sy_solve(A, b)
    assert info == 0
    assert np.max(np.abs(A @ x - b)) <= 1e-11 * np.max(np.abs(b)) * np.linalg.cond(A)


def test_symmetric_and_general_solves_agree_on_the_approximate_efie_matrix():
    """The approximate fill's matrix is symmetric; run_context solves it with solve_symmetric.  The
    symmetric and the general factorisations give the same currents to the conditioning."""
    sm = meshes.strip_dipole(0.5, 0.005, 60, 1)
    om = oracle.OracleMesh(sm)
    Z = om.fill(300e6, approx=True)
    assert np.max(np.abs(Z - Z.T)) <= 1e-13 * np.max(np.abs(Z))
    signs = om.reorient(sm.feed)
    V = om.rhs_deltagap(300e6, sm.feed, 1.0, signs)
    Is, i1 = oracle.sy_solve(Z, V)
    Ig, i2 = oracle.lu_solve(Z, V)
    assert i1 == 0 and i2 == 0
    assert np.max(np.abs(Is - Ig)) <= 1e-11 * np.max(np.abs(Ig))


def test_lu_pivots_on_first_max_of_abs_re_plus_abs_im():
    """izamax uses |re|+|im|: with column [1+1j, 1.9, ...] the first row is chosen (2 > 1.9) even
    though |1.9| > |1+1j|; a modulus-based pivot would pick row 1.  Both pivots solve the system,
    the factors differ -- visible in the rounding of an ill-conditioned case."""
    A = np.array([[1 + 1j, 2.0], [1.9, 1.0]])
    b = np.array([1.0, 0.0])
    x, info = oracle.lu_solve(A, b)
    assert info == 0
    assert np.allclose(A @ x, b, atol=1e-15)
    lu, piv = sla.lu_factor(A)
    assert piv[0] == 0  # LAPACK zgetrf keeps row 0 as well


def test_singular_matrix_reports_first_zero_pivot():
    A = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 0.0, 1.0]], dtype=complex)
    _, info = oracle.lu_solve(A, np.ones(3))
    assert info == 2


def _point_source_mesh(points):
    """Tiny triangles whose centroids sit exactly at `points` (field sums use centroids)."""
    V, T = [], []
    for p in points:
        a = np.array(p) + np.array([-1e-9, -1e-9, 0.0])
        b = np.array(p) + np.array([2e-9, -1e-9, 0.0])
        c = np.array(p) + np.array([-1e-9, 2e-9, 0.0])
        base = len(V)
        V += [a, b, c]
        T.append([base, base + 1, base + 2])
    return meshes.SurfaceMesh(np.array(V), np.array(T), np.ones(len(T)), np.zeros((0, 2)), np.zeros(0))


def test_fields_of_a_hertzian_dipole():
    """Current element I dl z at the origin with charges q = +-I/(j omega) at +-dl/2: the
    reference's centroid-dipole sums reproduce the closed-form Hertzian dipole E field, and H up
    to the reference's exp(+jkR) factor (src/emw_postpro_common.cpp:60-61)."""
    lam = meshes.wavelength(F0)
    k = 2 * math.pi * F0 * math.sqrt(MU0 * EPS0)
    dl, I = 1e-5 * lam, 1.0
    sm = _point_source_mesh([(0, 0, 0), (0, 0, dl / 2), (0, 0, -dl / 2)])
    om = oracle.OracleMesh.__new__(oracle.OracleMesh)
    # build only the mesh handle (no RWG needed for the field sum)
    import ctypes as C
    h = C.c_void_p()
    rc = oracle.lib().orc_mesh_build(oracle._p(sm.vertices), sm.nv, oracle._p(sm.triangles, oracle._i64),
                                     oracle._p(sm.tags, oracle._i32), sm.nt, oracle._p(sm.bedges, oracle._i64),
                                     oracle._p(sm.btags, oracle._i32), 0, C.byref(h))
    assert rc == 0
    om.h, om.sm = h, sm
    AJ = np.zeros((3, 3), complex)
    AJ[0, 2] = I * dl
    AdivJ = np.array([0.0, -I, I], complex)
    for r, th in ((0.7 * lam, 0.4), (3 * lam, 1.3), (10 * lam, 2.2)):
        p = np.array([r * math.sin(th), 0.0, r * math.cos(th)])
        E, H = om.fields(F0, AJ, AdivJ, p)
        ph = np.exp(-1j * k * r)
        Er = ETA0 * I * dl * math.cos(th) / (2 * math.pi * r ** 2) * (1 + 1 / (1j * k * r)) * ph
        Et = 1j * ETA0 * k * I * dl * math.sin(th) / (4 * math.pi * r) * (1 + 1 / (1j * k * r) - 1 / (k * r) ** 2) * ph
        Hp = 1j * k * I * dl * math.sin(th) / (4 * math.pi * r) * (1 + 1 / (1j * k * r)) * ph
        rhat = p / r
        that = np.array([math.cos(th), 0.0, -math.sin(th)])
        E_cf = Er * rhat + Et * that
        assert np.max(np.abs(E[0] - E_cf)) <= 1e-6 * np.max(np.abs(E_cf))
        H_cf = np.array([0.0, Hp, 0.0]) * np.exp(2j * k * r)  # reference: exp(+jkR) in H
        assert np.max(np.abs(H[0] - H_cf)) <= 1e-6 * np.max(np.abs(H_cf))


def test_reference_dipole_impedance_and_gain():
    """Strip dipole of share/examples/dipole/dipole.geo (arm 0.47 lambda/2, width 10 mm, cells of
    5 mm) in approximate mode: FRICO reports 72.203395 + j0.045661 Ohm at 300 MHz
    (dipole_frico.txt:8) and 'about 2.15 dB' of gain (README.md:25).  The gmsh mesh differs
    from the structured one here, so this is a ~1% physics pin, not parity."""
    lam = 299.792458 / 300.0
    arm = 0.47 * lam / 2
    nx = 2 * int(round(arm / 0.005))
    sm = meshes.strip_dipole(2 * arm, 0.01, nx, 2)
    r = oracle.antenna_analysis(sm, F0, approx=True)
    z = r["port"]["impedance"]
    assert abs(z.real - 72.203395) / 72.203395 < 0.02
    assert abs(z.imag - 0.045661) < 2.0
    gmax_db = 10 * math.log10(r["gain"].max())
    assert abs(gmax_db - 2.15) < 0.1
    # the pattern of a dipole along x: null along the axis (xy cut at 0 deg), max broadside
    assert r["gain"][0, 0] < 1e-2 * r["gain"].max()


def test_plate_broadside_rcs_near_physical_optics():
    """1 x 1 wavelength plate, normal incidence, observer far away: sigma ~ 4 pi A^2 / lambda^2
    (physical optics) within the ~30% that edge diffraction allows at this size."""
    lam = meshes.wavelength(F0)
    sm = meshes.plate(lam, 10)
    r = oracle.bistatic_analysis(sm, F0, 0.0, 0.0, 1.0, 0.0, degree=5, R=1e4)
    sigma_vv = r["table"][0, 0]
    po = 4 * math.pi * lam ** 2
    assert 0.7 * po < sigma_vv < 1.3 * po
