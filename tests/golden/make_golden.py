"""Writes the small golden fixtures of tests/golden/ with the FP64 CPU oracle (oracle/cpu_ref.cpp).

    python tests/golden/make_golden.py

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


if __name__ == "__main__":
    main()
