"""Writes the small golden fixtures of tests/golden/ with the FP64 CPU oracle (oracle/cpu_ref.cpp).

    python tests/golden/make_golden.py          # the small fixtures
    python tests/golden/make_golden.py --c3     # the N = 12160 headline fixture (minutes)
    python tests/golden/make_golden.py --c4     # the N = 1983 bowtie gradient fixture (a minute)

The reference cannot be built or run here (C++23 <print>, Eigen, gmsh, SILO and HDF5 are absent),
so the fixtures are produced by the oracle, whose functions are pinned separately by the tests in
tests/test_oracle_*.py.  Each fixture stores the mesh generator arguments and the outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from frico_amd import meshes  # noqa: E402

F0 = 300e6


def main():
    # 1. impedance matrices of a small strip (quadrature degree 5 and 2, approximate mode)
    sm = meshes.strip_dipole(0.5, 0.01, 16, 2)
    om = oracle.OracleMesh(sm)
    np.savez_compressed(os.path.join(HERE, "strip16x2_Z.npz"), length=0.5, width=0.01, nx=16, ny=2,
                        freq=F0, Zq5=om.fill(F0, 5), Zq2=om.fill(F0, 2), Za=om.fill(F0, approx=True))
    # 2. strip dipole N=399 antenna analysis (quadrature, degree 5)
    sm = meshes.workload_dipole_c1()
    r = oracle.antenna_analysis(sm, F0, degree=5)
    ra = oracle.antenna_analysis(sm, F0, approx=True)
    np.savez_compressed(os.path.join(HERE, "dipole_N399_antenna.npz"), freq=F0,
                        zin=r["port"]["impedance"], swr=r["port"]["swr"], gain=r["gain"], I=r["I"],
                        zin_approx=ra["port"]["impedance"], gain_approx=ra["gain"])
    # 3. bistatic RCS of a 1 x 1 wavelength plate, 8 x 8 cells, incidence (30 deg, 10 deg)
    lam = meshes.wavelength(F0)
    sm = meshes.plate(lam, 8)
    r = oracle.bistatic_analysis(sm, F0, np.radians(30.0), np.radians(10.0), 1.0, 1.0, degree=5)
    np.savez_compressed(os.path.join(HERE, "plate8_bistatic.npz"), size=lam, n=8, freq=F0,
                        theta=np.radians(30.0), phi=np.radians(10.0), table=r["table"])
    print("fixtures written to", HERE)


def c3_fixture():
    """The headline workload: PEC plate 3 x 3 wavelengths, 64 x 64 squares (N = 12160), Dunavant
    degree 5, 361 plane waves theta = 0..90 deg in 0.25 deg steps (phi = 0, E_theta = 1), solved in
    FP64 by the oracle's fill and LAPACK zgetrf/zgetrs (scipy).  Stored: the monostatic sigma_VV and
    sigma_HH at R = 10 m for every incidence (bistatic_RCS_unit_Ei, src/emw_ana_bistatic_rcs.cpp:62-82,
    at the back-scattering point), a seeded sample of Z entries (every 24th diagonal entry and 4096
    random (j, i) pairs), the solution at 48 seeded rows for every incidence, and the oracle's own
    relative residual per column.  About 12 GB of memory and a few minutes on 8 cores."""
    import time

    import scipy.linalg as sla
    t0 = time.perf_counter()
    sm = meshes.workload_plate(F0, 64)
    om = oracle.OracleMesh(sm)
    N = om.N
    th = np.radians(np.arange(361) * 0.25)
    Z = om.fill(F0, 5)
    t_fill = time.perf_counter() - t0
    rng = np.random.default_rng(20261204)
    diag = np.arange(0, N, 24)
    rj, ri = rng.integers(0, N, 4096), rng.integers(0, N, 4096)
    zj = np.concatenate([diag, rj])
    zi = np.concatenate([diag, ri])
    zval = Z[zj, zi].copy()
    B = np.stack([om.rhs_planewave(F0, t, 0.0, 1.0, 0.0, 5) for t in th], axis=1)
    lu = sla.lu_factor(Z, overwrite_a=False, check_finite=False)
    X = sla.lu_solve(lu, B, check_finite=False)
    del lu
    R = B - Z @ X
    res = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    del Z, R
    sig = np.zeros((361, 2))
    for a, t in enumerate(th):
        AJ, AdivJ, _ = om.scatter(X[:, a])
        pt = 10.0 * np.array([np.sin(t), 0.0, np.cos(t)])
        E, _ = om.fields(F0, AJ, AdivJ, pt[None, :])
        sig[a] = oracle.rcs_unit_Ei(E[0], t, 0.0, 10.0)
    xrows = np.sort(rng.choice(N, 48, replace=False))
    np.savez_compressed(os.path.join(HERE, "plate64_c3_monostatic.npz"), freq=F0, n=64, degree=5, N=N,
                        theta=th, R=10.0, sigma=sig, z_rows=zj, z_cols=zi, z_vals=zval, x_rows=xrows,
                        x_vals=X[xrows, :], oracle_rel_residual=res)
    print(f"C3 fixture: N={N}, fill {t_fill:.1f} s, total {time.perf_counter() - t0:.1f} s, "
          f"max oracle residual {res.max():.2e}")


def c4_gradient_fixture():
    """The adjoint workload: planar bowtie (arm 0.25 wavelength, flare 60 deg, the C4 mesh of
    meshes.workload_bowtie, N = 1983) at 300 MHz, Dunavant degree 5, 1 V delta gap: the oracle's
    FP64 input impedance, gain at (0, 0, 10 m) and their gradients with respect to every vertex
    coordinate (oracle.antenna_gradients).  About a minute on 8 cores."""
    import time
    t0 = time.perf_counter()
    sm = meshes.workload_bowtie(F0)
    r = oracle.antenna_gradients(sm, F0, 5, obs=(0.0, 0.0, 10.0))
    np.savez_compressed(os.path.join(HERE, "bowtie_c4_gradients.npz"), freq=F0, degree=5, N=r["N"],
                        obs=np.array([0.0, 0.0, 10.0]), zin=r["zin"], gain=r["gain"], dzin=r["dzin"],
                        dgain=r["dgain"])
    print(f"C4 gradient fixture: N={r['N']}, {time.perf_counter() - t0:.1f} s")


if __name__ == "__main__":
    if "--c4" in sys.argv:
        c4_gradient_fixture()
    elif "--c3" in sys.argv:
        c3_fixture()
    else:
        main()
        if "--all" in sys.argv:
            c3_fixture()
            c4_gradient_fixture()
