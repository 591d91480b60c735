"""GPU: the sweeps' results through the reference's text outputs (port_sweep.txt, polar_<n>.txt,
bistatic_rcs_<n>.txt) read back exactly."""
import numpy as np
import pytest

import frico_amd as fa
from frico_amd import meshes

pytestmark = pytest.mark.gpu


def test_antenna_sweep_text_outputs_round_trip(gpu_ctx, tmp_path):
    sm = meshes.workload_dipole_c1()
    gpu_ctx.load_mesh(sm, 5)
    freqs = [2.5e8, 3e8]
    res = gpu_ctx.antenna_sweep(freqs, sm.feed, mode="approx")
    fa.write_antenna_text_outputs(tmp_path, freqs, res, Z0=50.0)
    port = np.loadtxt(tmp_path / "port_sweep.txt")
    assert port.shape == (2, 5)
    for n, r in enumerate(res):
        assert port[n, 0] == n and port[n, 1] == freqs[n]
        assert port[n, 2] == r["impedance"].real and port[n, 3] == r["impedance"].imag and port[n, 4] == r["swr"]
        polar = np.loadtxt(tmp_path / f"polar_{n}.txt")
        assert polar.shape == (360, 7)
        np.testing.assert_array_equal(polar[:, [1, 3, 5]].T, r["gain"])
        assert (tmp_path / f"polar_{n}.txt").read_text().splitlines()[0] == \
            f"# FRICO radiation patterns. Frequency = {fa.format_double(freqs[n])}"


def test_bistatic_text_output_round_trip(gpu_ctx, tmp_path):
    lam = meshes.wavelength(3e8)
    sm = meshes.plate(lam, 8)
    gpu_ctx.load_mesh(sm, 5)
    tab = gpu_ctx.bistatic_sweep([3e8], 0.0, 0.0, mode="approx")[0]
    fa.write_bistatic_rcs(tmp_path / "bistatic_rcs_0.txt", tab)
    back = np.loadtxt(tmp_path / "bistatic_rcs_0.txt")
    np.testing.assert_array_equal(back[:, 0], np.arange(359))
    np.testing.assert_array_equal(back[:, 1:], tab)
