"""GPU: fields on a near-field sampling plane (eval_fields(mesh), src/emw_postpro_common.cpp:77-90,
as write_fields samples it, :180-189) against the oracle's eval_fields at the same vertices."""
import numpy as np
import pytest

import oracle
from frico_amd import meshes

pytestmark = pytest.mark.gpu

F0 = 300e6


def test_fields_on_a_plane_match_the_oracle(gpu_ctx):
    """Approximate-mode strip dipole; plane "xz" (y = 0) through the strip's axis, 80 x 40 cells,
    3321 vertices, several on the strip's own line at 0.83 mm from the nearest centroid.  Held to 1e-12
    of the plane's largest field, and to 1e-10 point by point."""
    sm = meshes.workload_dipole_c1()
    gpu_ctx.load_mesh(sm, 5)
    gpu_ctx.antenna_sweep([F0], sm.feed, mode="approx")
    I = gpu_ctx.solution()[:, 0]
    om = oracle.OracleMesh(sm)
    AJ, AdivJ, _ = om.scatter(I)
    V, T, E, H = gpu_ctx.fields_on_plane(F0, "xz:0.8:0.4:0.01")
    assert V.shape == (3321, 3) and np.max(np.abs(V[:, 1])) == 0.0
    Eo, Ho = om.fields(F0, AJ, AdivJ, V)
    # normwise over the plane: on the strip's own axis H vanishes by symmetry and holds only rounding
    e = np.max(np.abs(E - Eo)) / np.max(np.abs(Eo))
    h = np.max(np.abs(H - Ho)) / np.max(np.abs(Ho))
    # and point by point where the field is not a cancellation (above 1e-6 of the plane's maximum)
    big = np.max(np.abs(Ho), axis=1) > 1e-6 * np.max(np.abs(Ho))
    hp = np.max(np.abs(H - Ho)[big], axis=1) / np.max(np.abs(Ho)[big], axis=1)
    ep = np.max(np.abs(E - Eo), axis=1) / np.max(np.abs(Eo), axis=1)
    print(f"plane: E {e:.2e}, H {h:.2e}; worst point: E {ep.max():.2e}, H {hp.max():.2e} ({big.sum()} points)")
    assert e <= 1e-12 and h <= 1e-12
    assert ep.max() <= 1e-10 and hp.max() <= 1e-10


def test_fields_on_planes_of_each_orientation(gpu_ctx):
    """A random current on the small plate, the three named planes and a tilted one, 1e-10."""
    lam = meshes.wavelength(F0)
    sm = meshes.plate(0.6 * lam, 5)
    om = oracle.OracleMesh(sm)
    gpu_ctx.load_mesh(sm, 5)
    rng = np.random.default_rng(3)
    I = rng.uniform(-1, 1, om.N) + 1j * rng.uniform(-1, 1, om.N)
    gpu_ctx.set_rhs(I)
    gpu_ctx.set_solution(I)
    AJ, AdivJ, _ = om.scatter(I)
    for spec in ("xy:3:3:0.3", "yz:2:1:0.1", "xz:2:1:0.1", "1:1:1:2:2:0.2"):
        V, _, E, H = gpu_ctx.fields_on_plane(F0, spec)
        # planes through the plate's own plane may pass within rounding of a centroid: keep points
        # at least 1e-6 m from every triangle centroid, as the reference's 1/R would otherwise blow up
        cen = om.sm.vertices[om.sm.triangles].mean(axis=1)
        dmin = np.min(np.linalg.norm(V[:, None, :] - cen[None, :, :], axis=2), axis=1)
        keep = dmin > 1e-6
        Eo, Ho = om.fields(F0, AJ, AdivJ, V[keep])
        assert np.max(np.abs(E[keep] - Eo)) <= 1e-10 * np.max(np.abs(Eo))
        assert np.max(np.abs(H[keep] - Ho)) <= 1e-10 * np.max(np.abs(Ho))
