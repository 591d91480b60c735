"""HDF5 outputs: the reader checked on the reference's own antenna-analysis file
(share/hdf5/antenna_analysis/biquad.h5, copied as tests/golden/biquad_antenna.h5), then the
writer (frico_write_antenna_h5, frico_write_linear_system_h5) checked against it and by round trip."""
import os

import numpy as np
import pytest

from frico_amd.h5 import H5File

HERE = os.path.dirname(__file__)
REF = os.path.join(HERE, "golden", "biquad_antenna.h5")
BASE = "/frico/antenna_analysis"


@pytest.fixture(scope="module")
def ref():
    return H5File(REF)


def test_reference_file_layout(ref):
    assert ref.keys("/") == ["frico"] and ref.keys("/frico") == ["antenna_analysis"]
    n = int(ref[BASE + "/sweep_steps"])
    assert n == 11
    names = ref.keys(BASE)
    # symbol-table order (strcmp) over more entries than one symbol node holds
    assert names == sorted(names) and len(names) == n + 3
    assert set(names) == {str(i) for i in range(n)} | {"Z0", "rdiag_dist", "sweep_steps"}
    assert float(ref[BASE + "/Z0"]) == 50.0 and float(ref[BASE + "/rdiag_dist"]) == 10.0
    assert ref.keys(BASE + "/0") == ["frequency", "gain_XY", "gain_XZ", "gain_YZ", "port"]
    assert ref.keys(BASE + "/0/port") == ["current", "gamma", "impedance", "power", "swr", "voltage"]


def test_reference_port_values_satisfy_the_port_relations(ref):
    """compute_port_values and the SWR helpers (src/emw_ana_antenna.cpp:76-105): Z = V / I,
    P = V conj(I) / 2, Gamma = (Z - Z0) / (Z + Z0), SWR = (1 + |Gamma|) / (1 - |Gamma|)."""
    Z0 = float(ref[BASE + "/Z0"])
    freqs = []
    for s in range(int(ref[BASE + "/sweep_steps"])):
        p = f"{BASE}/{s}/port/"
        V, I, P = complex(ref[p + "voltage"]), complex(ref[p + "current"]), complex(ref[p + "power"])
        Z, G, S = complex(ref[p + "impedance"]), complex(ref[p + "gamma"]), float(ref[p + "swr"])
        assert V == 1.0
        assert abs(Z - V / I) <= 1e-12 * abs(Z)
        assert abs(P - 0.5 * V * np.conj(I)) <= 1e-15 + 1e-12 * abs(P)
        assert abs(G - (Z - Z0) / (Z + Z0)) <= 1e-14
        assert abs(S - (1 + abs(G)) / (1 - abs(G))) <= 1e-13 * S
        g = [ref[f"{BASE}/{s}/gain_{pl}"] for pl in ("XY", "YZ", "XZ")]
        assert all(x.shape == (360,) and np.all(x > 0) for x in g)
        freqs.append(float(ref[f"{BASE}/{s}/frequency"]))
    assert np.all(np.diff(freqs) > 0) and freqs[0] == 2.4e9


def _msgs(f, path):
    o = f._object(f._resolve(path))
    return {t: m for t, fl, m in o.msgs}


def _synthetic(nsteps, seed=0):
    rng = np.random.default_rng(seed)
    freqs = 2.4e9 + 1e7 * np.arange(nsteps)
    port = rng.standard_normal((nsteps, 10))
    gain = rng.uniform(0.01, 3.0, (nsteps, 1080))
    return freqs, port, gain


def test_written_antenna_file_matches_the_reference_structures(tmp_path, ref):
    import frico_amd
    freqs, port, gain = _synthetic(3)
    path = tmp_path / "ant.h5"
    frico_amd.write_antenna_h5(path, freqs, port, gain, rdiag_dist=10.0, Z0=50.0, voltage=1.0)
    f = H5File(str(path))
    assert f.keys("/") == ["frico"] and f.keys(BASE) == ["0", "1", "2", "Z0", "rdiag_dist", "sweep_steps"]
    assert f.keys(BASE + "/0") == ref.keys(BASE + "/0")
    assert f.keys(BASE + "/0/port") == ref.keys(BASE + "/0/port")
    # same superblock parameters and the same datatype, dataspace and fill value encodings
    assert (f.leaf_k, f.internal_k) == (ref.leaf_k, ref.internal_k)
    for p in ["/sweep_steps", "/Z0", "/0/gain_XY", "/0/port/impedance", "/0/port/swr", "/0/frequency"]:
        mine, theirs = _msgs(f, BASE + p), _msgs(ref, BASE + p)
        for t in (0x0001, 0x0003, 0x0005):
            assert mine[t] == theirs[t], (p, t)
        assert mine[0x0008][:2] == theirs[0x0008][:2] and mine[0x0008][10:18] == theirs[0x0008][10:18]
    assert len(f.b) == f.eof
    # values
    assert int(f[BASE + "/sweep_steps"]) == 3 and float(f[BASE + "/rdiag_dist"]) == 10.0
    for s in range(3):
        p = f"{BASE}/{s}/"
        assert float(f[p + "frequency"]) == freqs[s]
        for k, pl in enumerate(("XY", "YZ", "XZ")):
            assert np.array_equal(f[p + "gain_" + pl], gain[s, 360 * k:360 * (k + 1)])
        po = port[s]
        assert complex(f[p + "port/voltage"]) == 1.0
        for k, name in enumerate(("current", "power", "impedance", "gamma")):
            assert complex(f[p + "port/" + name]) == complex(po[2 * k], po[2 * k + 1])
        assert float(f[p + "port/swr"]) == po[8]


@pytest.mark.parametrize("nsteps", [8, 9, 300])
def test_large_sweeps_span_several_symbol_nodes(tmp_path, nsteps):
    import frico_amd
    freqs, port, gain = _synthetic(nsteps, seed=nsteps)
    path = tmp_path / "big.h5"
    frico_amd.write_antenna_h5(path, freqs, port, gain)
    f = H5File(str(path))
    names = f.keys(BASE)
    assert len(names) == nsteps + 3 and names == sorted(names)
    assert int(f[BASE + "/sweep_steps"]) == nsteps
    s = nsteps - 1
    assert float(f[f"{BASE}/{s}/frequency"]) == freqs[s]
    assert np.array_equal(f[f"{BASE}/{s}/gain_XZ"], gain[s, 720:])


@pytest.mark.parametrize("nrhs,with_i", [(1, True), (3, True), (1, False)])
def test_linear_system_dump_round_trip(tmp_path, nrhs, with_i):
    import frico_amd
    rng = np.random.default_rng(nrhs)
    n = 7
    Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    V = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    I = np.linalg.solve(Z, V)
    path = tmp_path / "frico_linear_system_0.h5"
    frico_amd.write_linear_system_h5(path, Z, V[:, 0] if nrhs == 1 else V, I if with_i else None)
    f = H5File(str(path))
    assert f.keys("/frico") == (["I", "V", "Z"] if with_i else ["V", "Z"])
    assert f.describe("/frico/Z")[1] == (n, n)
    assert np.array_equal(f["/frico/Z"], Z)  # entry [i][j] = Z[i, j]
    assert np.array_equal(f["/frico/V"], V)
    if with_i:
        assert np.array_equal(f["/frico/I"], I)


def test_reader_rejects_what_it_does_not_read(tmp_path):
    from frico_amd.h5 import H5Error
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file at all")
    with pytest.raises(H5Error):
        H5File(str(p))
    b = bytearray(open(REF, "rb").read())
    b[8] = 2  # superblock version 2
    with pytest.raises(H5Error):
        H5File(bytes(b))
