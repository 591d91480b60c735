"""Pins the oracle's connectivity and RWG construction (src/input_gmsh.cpp:120-322,
src/rwg_basis.cpp:27-68, src/geom_mesh.cpp:46-78) by invariants of the function space."""
import numpy as np
import pytest

import oracle
from frico_amd import meshes

F0 = 300e6


def all_meshes():
    lam = meshes.wavelength(F0)
    return [meshes.strip_dipole(0.5, 0.005, 20, 1), meshes.strip_dipole(0.4, 0.05, 8, 4),
            meshes.plate(lam, 6), meshes.icosphere(lam, 3), meshes.bowtie(0.25 * lam, 60, 5, 4),
            meshes.strip_loop(0.16, 0.01, 30)]


@pytest.mark.parametrize("sm", all_meshes(), ids=lambda s: s.name)
def test_counts_and_ordering(sm):
    om = oracle.OracleMesh(sm)
    assert om.N == sm.num_internal_edges()
    assert np.all(om.itm < om.itp)  # T- is the first visit (src/input_gmsh.cpp:259-260)
    # matrix order = sorted unique edge order
    assert np.all(np.diff(om.edge) > 0)
    e = om.edges
    assert np.all(e[:, 0] < e[:, 1])
    assert np.all((e[1:, 0] > e[:-1, 0]) | ((e[1:, 0] == e[:-1, 0]) & (e[1:, 1] > e[:-1, 1])))


def test_closed_sphere_euler():
    for nu in (1, 2, 4):
        sm = meshes.icosphere(1.0, nu)
        om = oracle.OracleMesh(sm)
        V, T = sm.nv, sm.nt
        E = om.edges.shape[0]
        assert V - E + T == 2
        assert om.N == E == 3 * T // 2 == 30 * nu * nu


@pytest.mark.parametrize("sm", all_meshes(), ids=lambda s: s.name)
def test_rwg_geometry(sm):
    om = oracle.OracleMesh(sm)
    V, Tr = sm.vertices, sm.triangles
    for n in range(om.N):
        tm, tp = Tr[om.itm[n]], Tr[om.itp[n]]
        shared = set(tm) & set(tp)
        assert len(shared) == 2
        a, b = sorted(shared)
        assert (om.edges[om.edge[n]] == [a, b]).all()
        # free vertices are the vertices not on the edge
        (fm,) = set(tm) - shared
        (fp,) = set(tp) - shared
        np.testing.assert_array_equal(om.pm[n], V[fm])
        np.testing.assert_array_equal(om.pp[n], V[fp])
        d = V[b] - V[a]
        assert om.length[n] == np.hypot(d[0], d[1])  # xy-only length (src/geom_mesh.cpp:51)
        Am = 0.5 * np.linalg.norm(np.cross(V[tm[1]] - V[tm[0]], V[tm[2]] - V[tm[1]]))
        assert om.Am[n] == pytest.approx(Am, rel=1e-15)
        # normal component of f- and f+ on the shared edge: rho . n_edge = height = 2A/l  (planar, z=0)
        if np.all(V[:, 2] == 0):
            mid = 0.5 * (V[a] + V[b])
            t = d / np.linalg.norm(d)
            nrm = np.array([t[1], -t[0], 0.0])
            hm = abs(np.dot(mid - om.pm[n], nrm))
            hp = abs(np.dot(mid - om.pp[n], nrm))
            assert hm == pytest.approx(2 * om.Am[n] / np.linalg.norm(d), rel=1e-12)
            assert hp == pytest.approx(2 * om.Ap[n] / np.linalg.norm(d), rel=1e-12)


def test_interface_tags_only_on_internal_feed_edges():
    sm = meshes.strip_dipole(0.5, 0.01, 20, 3)
    om = oracle.OracleMesh(sm)
    feed = om.has_iface.astype(bool)
    assert feed.sum() == 3
    assert np.all(om.iface[feed] == meshes.FEED_TAG)
    xs = sm.vertices[om.edges[om.edge[feed]]][:, :, 0]
    assert np.all(xs == 0.0)


def test_lonely_edge_and_nonmanifold_errors():
    sm = meshes.plate(1.0, 2)
    bad = meshes.SurfaceMesh(sm.vertices, sm.triangles, sm.tags, np.array([[0, 8]]), np.array([7]))
    with pytest.raises(ValueError, match=r"\(-1\)"):
        oracle.OracleMesh(bad)
    # a third triangle on an existing edge
    V = np.vstack([sm.vertices, [[0.0, 0.0, 1.0]]])
    T = np.vstack([sm.triangles, [[0, 4, len(V) - 1]]])  # diagonal (0,4) already has two
    nm = meshes.SurfaceMesh(V, T, np.ones(len(T)), np.zeros((0, 2)), np.zeros(0))
    with pytest.raises(ValueError, match=r"\(-2\)"):
        oracle.OracleMesh(nm)


def test_deltagap_reorientation_rules():
    # x>0 arm first: T- in surface 1, T+ in surface 2 -> no reorientation
    sm = meshes.strip_dipole(0.5, 0.01, 20, 3)
    om = oracle.OracleMesh(sm)
    assert om.reorient(sm.feed) is None
    # forced: signs of a straight feed line are all +1
    s = om.reorient(sm.feed, force=True)
    feed = np.nonzero(om.has_iface)[0]
    assert np.all(s[feed] == 1.0)
    # swap surface tags so that T+.tag <= T-.tag: reorientation triggers, still consistent (+1)
    sm2 = meshes.SurfaceMesh(sm.vertices, sm.triangles, 3 - sm.tags, sm.bedges, sm.btags, feed=sm.feed)
    om2 = oracle.OracleMesh(sm2)
    s2 = om2.reorient(sm.feed)
    assert s2 is not None and np.all(np.abs(s2[feed]) == 1.0)
