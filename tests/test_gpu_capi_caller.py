"""GPU: the C ABI called from a compiled C++ program (tests/capi/run_context.cpp, built with g++
against include/frico_gpu.h by frico_amd/build.py): the run_context binding of a FRICO maintainer
(src/emw_solver.h:92-167) on the strip dipole C1 (N = 399, Dunavant degree 5), read from a gmsh
file with its feed in a physical group.  Its input impedance and gain match the FP64 oracle to
1e-10, and the library's own frico_antenna_sweep."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from frico_amd import meshes

pytestmark = pytest.mark.gpu

F0 = 300e6
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "capi", "run_context")


def _run(tmp_path, sm, approx):
    p = tmp_path / "mesh.msh"
    meshes.write_msh(sm, p, groups={"source": (1, list(sm.feed))})
    out = tmp_path / "out.txt"
    assert os.path.exists(BIN), "tests/capi/run_context is built by frico_amd/build.py"
    print(f"run_context: N-mesh {sm.nt} triangles, approx={approx}", flush=True)
    r = subprocess.run([BIN, str(p), "source", repr(F0), "5", "1" if approx else "0", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vals = {}
    for line in out.read_text().splitlines():
        k, *v = line.split()
        if k in ("zin", "zin_sweep", "gain", "gain_sweep"):
            vals[k] = [float(x) for x in v]
    return vals


@pytest.mark.parametrize("approx", [False, True])
def test_compiled_run_context_binding_matches_the_oracle(tmp_path, approx):
    sm = meshes.workload_dipole_c1()
    vals = _run(tmp_path, sm, approx)
    ref = oracle.antenna_analysis(sm, F0, degree=5, approx=approx)
    zo = ref["port"]["impedance"]
    zin = complex(*vals["zin"])
    zsw = complex(*vals["zin_sweep"])
    gain = np.array(vals["gain"]).reshape(3, 360)
    gsw = np.array(vals["gain_sweep"]).reshape(3, 360)
    print(f"compiled caller: Zin {zin:.10f} oracle {zo:.10f}")
    assert abs(zin - zo) <= 1e-10 * abs(zo)
    assert abs(zsw - zo) <= 1e-10 * abs(zo)
    assert np.max(np.abs(gain - ref["gain"])) <= 1e-10 * np.max(ref["gain"])
    assert np.max(np.abs(gsw - ref["gain"])) <= 1e-10 * np.max(ref["gain"])
