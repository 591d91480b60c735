"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py with the oracle):
the oracle must keep reproducing them (CPU), and the HIP library must match them (GPU)."""
import os

import numpy as np
import pytest

import oracle
from frico_amd import meshes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name))


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_oracle_reproduces_matrix_fixture():
    g = load("strip16x2_Z.npz")
    om = oracle.OracleMesh(meshes.strip_dipole(float(g["length"]), float(g["width"]), int(g["nx"]), int(g["ny"])))
    assert rel(om.fill(float(g["freq"]), 5), g["Zq5"]) <= 1e-14
    assert rel(om.fill(float(g["freq"]), 2), g["Zq2"]) <= 1e-14
    assert rel(om.fill(float(g["freq"]), approx=True), g["Za"]) <= 1e-14


def test_oracle_reproduces_antenna_fixture():
    g = load("dipole_N399_antenna.npz")
    r = oracle.antenna_analysis(meshes.workload_dipole_c1(), float(g["freq"]), degree=5)
    assert abs(r["port"]["impedance"] - complex(g["zin"])) <= 1e-12 * abs(complex(g["zin"]))
    assert rel(r["gain"], g["gain"]) <= 1e-12


@pytest.mark.gpu
def test_gpu_matches_matrix_fixture(gpu_ctx):
    g = load("strip16x2_Z.npz")
    sm = meshes.strip_dipole(float(g["length"]), float(g["width"]), int(g["nx"]), int(g["ny"]))
    for degree, key in ((5, "Zq5"), (2, "Zq2")):
        gpu_ctx.load_mesh(sm, degree)
        gpu_ctx.fill(float(g["freq"]), "quad", "c128")
        assert rel(gpu_ctx.matrix(), g[key]) <= 1e-10
    gpu_ctx.fill(float(g["freq"]), "approx", "c128")
    assert rel(gpu_ctx.matrix(), g["Za"]) <= 1e-10


@pytest.mark.gpu
def test_gpu_matches_antenna_fixture(gpu_ctx):
    g = load("dipole_N399_antenna.npz")
    sm = meshes.workload_dipole_c1()
    gpu_ctx.load_mesh(sm, 5)
    out = gpu_ctx.antenna_sweep([float(g["freq"])], sm.feed)[0]
    assert abs(out["impedance"] - complex(g["zin"])) <= 1e-10 * abs(complex(g["zin"]))
    assert rel(out["gain"], g["gain"]) <= 1e-10
    outa = gpu_ctx.antenna_sweep([float(g["freq"])], sm.feed, mode="approx")[0]
    assert abs(outa["impedance"] - complex(g["zin_approx"])) <= 1e-10 * abs(complex(g["zin_approx"]))


@pytest.mark.gpu
def test_gpu_matches_bistatic_fixture(gpu_ctx):
    g = load("plate8_bistatic.npz")
    gpu_ctx.load_mesh(meshes.plate(float(g["size"]), int(g["n"])), 5)
    tab = gpu_ctx.bistatic_sweep([float(g["freq"])], float(g["theta"]), float(g["phi"]), 1.0, 1.0)[0]
    assert rel(tab, g["table"]) <= 1e-10
