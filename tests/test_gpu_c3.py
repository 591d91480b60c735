"""GPU: the headline workload at full size against the FP64 oracle's fixture.

PEC plate 3 x 3 wavelengths at 300 MHz, 64 x 64 squares (N = 12160, T = 8192), Dunavant degree 5,
361 plane waves theta = 0..90 deg in 0.25 deg steps, solved as one block at the default block width
(128 from N = 6144) with the look-ahead stream.  tests/golden/plate64_c3_monostatic.npz holds the
oracle's answer (its fill, LAPACK zgetrf/zgetrs through scipy): the monostatic sigma of every
incidence (bistatic_RCS_unit_Ei, src/emw_ana_bistatic_rcs.cpp:62-82), a sample of Z and of the
solution.  The complex128 form is held to 1e-10; the refinement forms (residual <= 1e-10) report
their error against the oracle and are held to the error that residual allows.
"""
import os

import numpy as np
import pytest

from frico_amd import meshes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "plate64_c3_monostatic.npz")


@pytest.fixture(scope="module")
def c3():
    import frico_amd
    g = np.load(FIX)
    sm = meshes.workload_plate(float(g["freq"]), int(g["n"]))
    ctx = frico_amd.Context(0)
    ctx.load_mesh(sm, int(g["degree"]))
    assert ctx.N == int(g["N"]) == 12160
    yield ctx, g
    ctx.close()


def _sigma_err(sig, ref):
    """max |sigma - sigma_ref| / max sigma_ref over the 361 incidences (VV and HH)."""
    return float(np.max(np.abs(sig - ref)) / np.max(np.abs(ref)))


def test_c3_impedance_matrix_sample(c3):
    ctx, g = c3
    ctx.fill(float(g["freq"]), "quad", "c128")
    Z = ctx.matrix()
    zs = Z[g["z_rows"], g["z_cols"]]
    del Z
    err = float(np.max(np.abs(zs - g["z_vals"])) / np.max(np.abs(g["z_vals"])))
    print(f"C3 Z sample ({zs.size} entries): max error / max |Z| = {err:.2e}")
    assert err <= 1e-10


def test_c3_complex128_form_matches_the_oracle(c3):
    ctx, g = c3
    f, th = float(g["freq"]), g["theta"]
    sig, _ = ctx.monostatic_sweep([f], th, 0.0, form="c128", R=float(g["R"]))
    X = ctx.solution()
    es = _sigma_err(sig[0], g["sigma"])
    ex = float(np.max(np.abs(X[g["x_rows"], :] - g["x_vals"])) / np.max(np.abs(g["x_vals"])))
    print(f"C3 complex128: sigma error {es:.2e}, solution-sample error {ex:.2e}")
    assert es <= 1e-10 and ex <= 1e-10


@pytest.mark.parametrize("form", ["c64_np_ir", "c64_ir"])
def test_c3_refinement_forms_against_the_oracle(c3, form):
    """The headline form (complex64 LU without pivoting) and the pivoted complex64 LU, both refined
    to a 1e-10 FP64 residual: converged, the unpivoted factors kept (no fall-back), and within 1e-8
    of the oracle in sigma and in the solution sample (the residual bound times the conditioning the
    complex128 solve shows)."""
    import frico_amd
    ctx, g = c3
    f, th = float(g["freq"]), g["theta"]
    sig, info = ctx.monostatic_sweep([f], th, 0.0, form=form, tol=1e-10, R=float(g["R"]))
    X = ctx.solution()
    assert info[0].converged and info[0].max_rel_residual <= 1e-10
    want = frico_amd.SOLVE_C64_NP_IR if form == "c64_np_ir" else frico_amd.SOLVE_C64_IR
    assert ctx.factored_form == want
    es = _sigma_err(sig[0], g["sigma"])
    ex = float(np.max(np.abs(X[g["x_rows"], :] - g["x_vals"])) / np.max(np.abs(g["x_vals"])))
    print(f"C3 {form}: {info[0].iterations} sweeps, residual {info[0].max_rel_residual:.2e}, "
          f"sigma error {es:.2e}, solution-sample error {ex:.2e}")
    assert es <= 1e-8 and ex <= 1e-8
