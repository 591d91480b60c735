"""GPU: a mesh read from a gmsh .msh file (frico_msh_read) runs the antenna analysis exactly as the
mesh it was written from, with the feed taken from its physical group, as FRICO's source option
names it (src/main.cpp:183-199)."""
import numpy as np
import pytest

import frico_amd as fa
import oracle
from frico_amd import meshes

pytestmark = pytest.mark.gpu

F0 = 300e6


def test_antenna_from_msh_file(gpu_ctx, tmp_path):
    sm = meshes.bowtie(0.25 * meshes.wavelength(F0), 60.0, 6, 4)
    p = tmp_path / "bowtie.msh"
    meshes.write_msh(sm, p, groups={"source": (1, list(sm.feed))})
    rd, _ = fa.read_msh(p, feed_group="source")
    ref = oracle.antenna_analysis(sm, F0, degree=5)
    gpu_ctx.load_mesh(rd, 5)
    out = gpu_ctx.antenna_sweep([F0], rd.feed)[0]
    zo = ref["port"]["impedance"]
    assert abs(out["impedance"] - zo) / abs(zo) <= 1e-10
    assert np.max(np.abs(out["gain"] - ref["gain"])) <= 1e-10 * np.max(ref["gain"])
