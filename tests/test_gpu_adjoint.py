"""GPU adjoint: vertex gradients of input impedance and gain (no counterpart in the reference).

Checked against the FP64 CPU adjoint of the oracle (same mesh, same forward) and against central
finite differences of the GPU's own complex128 forward path (step 1e-6 wavelengths), to 1e-6
relative to the largest gradient entry."""
import numpy as np
import pytest

import oracle
from frico_amd import meshes

pytestmark = pytest.mark.gpu

F0 = 300e6
OBS = (0.0, 0.0, 10.0)


def _perturbed(sm, V):
    return meshes.SurfaceMesh(V, sm.triangles, sm.tags, sm.bedges, sm.btags, feed=sm.feed)


@pytest.mark.parametrize("degree", [5, 2, 1, 3, 7])
def test_gradients_match_the_cpu_adjoint(gpu_ctx, degree):
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 4, 3)
    ref = oracle.antenna_gradients(sm, F0, degree, obs=OBS)
    gpu_ctx.load_mesh(sm, degree)
    g = gpu_ctx.antenna_gradients(F0, sm.feed, 1.0, OBS)
    assert abs(g["zin"] - ref["zin"]) <= 1e-10 * abs(ref["zin"])
    assert abs(g["gain"] - ref["gain"]) <= 1e-10 * abs(ref["gain"])
    ez = np.max(np.abs(g["dzin"] - ref["dzin"])) / np.max(np.abs(ref["dzin"]))
    eg = np.max(np.abs(g["dgain"] - ref["dgain"])) / np.max(np.abs(ref["dgain"]))
    print(f"degree {degree}: dZin rel err {ez:.2e}, dG rel err {eg:.2e}")
    assert ez <= 1e-9 and eg <= 1e-9


def test_gradients_with_128_column_panels():
    """The transposed solves of the adjoint with factors from 128-column panels (each factored as
    two 64-column panel launches)."""
    import frico_amd
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 8, 5)
    ref = oracle.antenna_gradients(sm, F0, 5, obs=OBS)
    with frico_amd.Context(0) as c:
        c.set_block(128)
        c.load_mesh(sm, 5)
        assert c.N > 128
        g = c.antenna_gradients(F0, sm.feed, 1.0, OBS)
    assert abs(g["zin"] - ref["zin"]) <= 1e-10 * abs(ref["zin"])
    assert np.max(np.abs(g["dzin"] - ref["dzin"])) <= 1e-9 * np.max(np.abs(ref["dzin"]))
    assert np.max(np.abs(g["dgain"] - ref["dgain"])) <= 1e-9 * np.max(np.abs(ref["dgain"]))


def test_gradients_on_a_strip_with_a_multi_edge_gap(gpu_ctx):
    sm = meshes.strip_dipole(0.4, 0.03, 10, 3)
    ref = oracle.antenna_gradients(sm, F0, 5, obs=(0.0, 3.0, 4.0))
    gpu_ctx.load_mesh(sm, 5)
    g = gpu_ctx.antenna_gradients(F0, sm.feed, 1.0, (0.0, 3.0, 4.0))
    assert np.max(np.abs(g["dzin"] - ref["dzin"])) <= 1e-9 * np.max(np.abs(ref["dzin"]))
    assert np.max(np.abs(g["dgain"] - ref["dgain"])) <= 1e-9 * np.max(np.abs(ref["dgain"]))


def test_bowtie_c4_gradients_match_the_oracle_fixture(gpu_ctx):
    """The adjoint workload at full size (bowtie arm 0.25 lambda, flare 60 degrees, N = 1983,
    degree 5) against the oracle's FP64 gradients stored in tests/golden/bowtie_c4_gradients.npz
    (tests/golden/make_golden.py --c4): Zin and gain to 1e-10, every vertex gradient entry to 1e-9
    of the largest."""
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "bowtie_c4_gradients.npz"))
    assert float(d["freq"]) == F0 and int(d["degree"]) == 5
    sm = meshes.workload_bowtie(F0)
    gpu_ctx.load_mesh(sm, 5)
    assert gpu_ctx.N == int(d["N"]) == 1983
    g = gpu_ctx.antenna_gradients(F0, sm.feed, 1.0, tuple(d["obs"]))
    assert g["dzin"].shape == d["dzin"].shape == (sm.nv, 3)
    zin, gain = complex(d["zin"]), float(d["gain"])
    assert abs(g["zin"] - zin) <= 1e-10 * abs(zin)
    assert abs(g["gain"] - gain) <= 1e-10 * abs(gain)
    ez = np.max(np.abs(g["dzin"] - d["dzin"])) / np.max(np.abs(d["dzin"]))
    eg = np.max(np.abs(g["dgain"] - d["dgain"])) / np.max(np.abs(d["dgain"]))
    print(f"C4: dZin rel err {ez:.2e}, dG rel err {eg:.2e}")
    assert ez <= 1e-9 and eg <= 1e-9


def test_bowtie_c4_gradients_match_central_differences(gpu_ctx):
    """Bowtie with arm 0.25 lambda and 60 degree flare, N = 1983: adjoint vs central differences
    with a step of 1e-6 wavelengths on a spread of vertex coordinates."""
    lam = meshes.wavelength(F0)
    sm = meshes.workload_bowtie(F0)
    gpu_ctx.load_mesh(sm, 5)
    assert gpu_ctx.N == 1983
    g = gpu_ctx.antenna_gradients(F0, sm.feed, 1.0, OBS)
    h = 1e-6 * lam
    rng = np.random.default_rng(4)
    # the feed vertices (x = 0) and a random sample elsewhere
    feed_v = np.nonzero(sm.vertices[:, 0] == 0.0)[0][:3]
    verts = list(feed_v) + list(rng.choice(sm.nv, 5, replace=False))
    scale_z, scale_g = np.max(np.abs(g["dzin"])), np.max(np.abs(g["dgain"]))
    for v in verts:
        for c in (0, 1, 2):
            vals = []
            for sgn in (1, -1):
                V = sm.vertices.copy()
                V[v, c] += sgn * h
                gpu_ctx.load_mesh(_perturbed(sm, V), 5)
                vals.append(gpu_ctx.adjoint_forward(F0, sm.feed, 1.0, OBS))
            fz = (vals[0][0] - vals[1][0]) / (2 * h)
            fg = (vals[0][1] - vals[1][1]) / (2 * h)
            assert abs(fz - g["dzin"][v, c]) <= 1e-6 * scale_z, (v, c, fz, g["dzin"][v, c])
            assert abs(fg - g["dgain"][v, c]) <= 1e-6 * scale_g, (v, c, fg, g["dgain"][v, c])


def test_forward_backward_pair_shape():
    import frico_amd
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 3, 2)
    with frico_amd.Context(0) as ctx:
        f = frico_amd.AntennaResponse(ctx, sm, F0)
        y = f.forward(sm.vertices)
        assert y.shape == (3,)
        gv = f.backward([1.0, 0.0, 0.0])
        assert gv.shape == sm.vertices.shape
        # linear in the cotangent
        g2 = f.backward([2.0, -1.0, 0.5])
        g_re, g_im, g_g = f.backward([1, 0, 0]), f.backward([0, 1, 0]), f.backward([0, 0, 1])
        assert np.allclose(g2, 2 * g_re - g_im + 0.5 * g_g, rtol=0, atol=1e-12 * np.max(np.abs(g2)))


def test_backward_before_forward_is_an_error():
    import frico_amd
    sm = meshes.bowtie(0.2, 60.0, 2, 2)
    with frico_amd.Context(0) as ctx:
        ctx.load_mesh(sm, 5)
        with pytest.raises(frico_amd.FricoError):
            ctx.adjoint_backward(1.0, 0.0, 0.0)


@pytest.mark.parametrize("degree", [5, 2])
@pytest.mark.parametrize("bistatic", [False, True])
def test_rcs_gradients_match_the_cpu_adjoint(gpu_ctx, degree, bistatic):
    """sigma_VV, sigma_HH of an obliquely lit plate (mixed polarisation) and their vertex gradients
    against the oracle's FP64 RCS adjoint (oracle.rcs_gradients)."""
    lam = meshes.wavelength(F0)
    sm = meshes.plate(0.6 * lam, 3)
    thi, phi_ = 0.5, 0.3
    ths, phs = (1.1, -0.7) if bistatic else (thi, phi_)
    ref = oracle.rcs_gradients(sm, F0, thi, phi_, 1.0, 0.5j, ths, phs, degree=degree, R=10.0)
    gpu_ctx.load_mesh(sm, degree)
    g = gpu_ctx.rcs_gradients(F0, thi, phi_, 1.0, 0.5j, ths, phs, R=10.0)
    assert np.max(np.abs(g["sigma"] - ref["sigma"]) / ref["sigma"]) <= 1e-10
    for key in ("dsigma_vv", "dsigma_hh"):
        e = np.max(np.abs(g[key] - ref[key])) / np.max(np.abs(ref[key]))
        print(f"degree {degree} bistatic {bistatic}: {key} rel err {e:.2e}")
        assert e <= 1e-9


def test_monostatic_rcs_gradient_matches_central_differences(gpu_ctx):
    """Plate 1.5 x 1.5 wavelengths (20 x 20 squares, N = 1160), monostatic sigma_VV at theta = 20 deg:
    the adjoint gradient against central differences of the GPU's own forward (step 1e-6
    wavelengths) on a spread of vertex coordinates, to 1e-6 of the largest gradient entry."""
    lam = meshes.wavelength(F0)
    sm = meshes.plate(1.5 * lam, 20)
    th = np.radians(20.0)
    gpu_ctx.load_mesh(sm, 5)
    s0 = gpu_ctx.rcs_forward(F0, th, 0.0)
    grad = gpu_ctx.rcs_backward(1.0, 0.0)
    scale = np.max(np.abs(grad))
    h = 1e-6 * lam
    rng = np.random.default_rng(4)
    picks = [(int(v), int(c)) for v, c in zip(rng.integers(0, sm.nv, 10), rng.integers(0, 3, 10))]
    picks.append((int(np.argmax(np.max(np.abs(grad), axis=1))), int(np.argmax(np.max(np.abs(grad), axis=0)))))
    for v, c in picks:
        Vp, Vm = sm.vertices.copy(), sm.vertices.copy()
        Vp[v, c] += h
        Vm[v, c] -= h
        gpu_ctx.load_mesh(_perturbed(sm, Vp), 5)
        sp = gpu_ctx.rcs_forward(F0, th, 0.0)[0]
        gpu_ctx.load_mesh(_perturbed(sm, Vm), 5)
        sm_ = gpu_ctx.rcs_forward(F0, th, 0.0)[0]
        fd = (sp - sm_) / (2 * h)
        print(f"vertex {v} coord {c}: adjoint {grad[v, c]:.6e} central difference {fd:.6e}")
        assert abs(fd - grad[v, c]) <= 1e-6 * scale
    assert s0[0] > 0


# ---- approximate mode (the reference's -A: compute_matrix_approx, src/emw_solver.cpp:144-220) ---

@pytest.mark.parametrize("degree", [5, 3, 8])
def test_approx_gradients_match_the_cpu_adjoint(gpu_ctx, degree):
    """Zin and gain gradients with the approximate-mode Z against the oracle's FP64 approximate
    adjoint, at degrees the quadrature-mode adjoint has no kernel for (the approximate Z does not
    use the rule)."""
    lam = meshes.wavelength(F0)
    sm = meshes.bowtie(0.25 * lam, 60.0, 4, 3)
    ref = oracle.antenna_gradients(sm, F0, degree, obs=OBS, approx=True)
    gpu_ctx.load_mesh(sm, degree)
    g = gpu_ctx.antenna_gradients(F0, sm.feed, 1.0, OBS, mode="approx")
    assert abs(g["zin"] - ref["zin"]) <= 1e-10 * abs(ref["zin"])
    assert abs(g["gain"] - ref["gain"]) <= 1e-10 * abs(ref["gain"])
    ez = np.max(np.abs(g["dzin"] - ref["dzin"])) / np.max(np.abs(ref["dzin"]))
    eg = np.max(np.abs(g["dgain"] - ref["dgain"])) / np.max(np.abs(ref["dgain"]))
    print(f"approx degree {degree}: dZin rel err {ez:.2e}, dG rel err {eg:.2e}")
    assert ez <= 1e-9 and eg <= 1e-9
    # the quadrature-mode gradients of the same mesh differ (the mode is not ignored)
    if degree == 5:
        q = gpu_ctx.antenna_gradients(F0, sm.feed, 1.0, OBS)
        assert np.max(np.abs(q["dzin"] - g["dzin"])) > 1e-3 * np.max(np.abs(g["dzin"]))


@pytest.mark.parametrize("bistatic", [False, True])
def test_approx_rcs_gradients_match_the_cpu_adjoint(gpu_ctx, bistatic):
    lam = meshes.wavelength(F0)
    sm = meshes.plate(0.6 * lam, 3)
    thi, phi_ = 0.5, 0.3
    ths, phs = (1.1, -0.7) if bistatic else (thi, phi_)
    ref = oracle.rcs_gradients(sm, F0, thi, phi_, 1.0, 0.5j, ths, phs, degree=5, R=10.0, approx=True)
    gpu_ctx.load_mesh(sm, 5)
    g = gpu_ctx.rcs_gradients(F0, thi, phi_, 1.0, 0.5j, ths, phs, R=10.0, mode="approx")
    assert np.allclose(g["sigma"], ref["sigma"], rtol=1e-10)
    for key in ("dsigma_vv", "dsigma_hh"):
        e = np.max(np.abs(g[key] - ref[key])) / np.max(np.abs(ref[key]))
        print(f"approx rcs bistatic={bistatic} {key} rel err {e:.2e}")
        assert e <= 1e-9


def test_approx_gradient_matches_central_differences():
    """The GPU's approximate-mode dZin against central differences (1e-6 wavelengths) of its own
    approximate-mode forward, on a strip with a 2-edge gap."""
    import frico_amd
    lam = meshes.wavelength(F0)
    sm = meshes.strip_dipole(0.5 * lam, 0.02 * lam, 12, 2)
    h = 1e-6 * lam
    with frico_amd.Context(0) as c:
        c.load_mesh(sm, 1)
        g = c.antenna_gradients(F0, sm.feed, 1.0, OBS, mode="approx")
        for v in range(0, sm.nv, 5):
            for d in range(3):
                Vp, Vm = sm.vertices.copy(), sm.vertices.copy()
                Vp[v, d] += h
                Vm[v, d] -= h
                c.load_mesh(_perturbed(sm, Vp), 1)
                zp, _ = c.adjoint_forward(F0, sm.feed, 1.0, OBS, mode="approx")
                c.load_mesh(_perturbed(sm, Vm), 1)
                zm, _ = c.adjoint_forward(F0, sm.feed, 1.0, OBS, mode="approx")
                fd = (zp - zm) / (2 * h)
                assert abs(fd - g["dzin"][v, d]) <= 1e-6 * np.max(np.abs(g["dzin"]))
