"""gmsh .msh 4.1 ingestion (frico_msh_read) against load_from_gmsh's rules (src/input_gmsh.cpp:36-406):
what gmsh's API returns for a loaded file is restated here in plain Python (nodes by entity
(dim, tag) then file order, triangles and lines by entity tag then file order, unused vertices
dropped in order) and compared with the library's reader; the RWG numbering of a mesh read back
from a file equals that of the mesh itself (oracle, CPU)."""
import numpy as np
import pytest

import frico_amd as fa
import oracle
from frico_amd import meshes

F0 = 300e6


def gen():
    lam = meshes.wavelength(F0)
    return {
        "strip": meshes.strip_dipole(0.5, 0.005, 40, 1),
        "plate": meshes.plate(0.6 * lam, 5),
        "sphere": meshes.icosphere(0.3 * lam, 2),
        "bowtie": meshes.bowtie(0.25 * lam, 60.0, 4, 3),
        "loop": meshes.strip_loop(0.16, 0.01, 24),
    }


def by_tag(sm):
    """Triangles and curve edges in the order gmsh's API lists them: by entity tag, then file order."""
    ot = np.argsort(sm.tags, kind="stable")
    ob = np.argsort(sm.btags, kind="stable")
    return sm.triangles[ot], sm.tags[ot], sm.bedges[ob], sm.btags[ob]


@pytest.mark.parametrize("name", list(gen()))
def test_round_trip_and_rwg_numbering(tmp_path, name):
    sm = gen()[name]
    p = tmp_path / f"{name}.msh"
    meshes.write_msh(sm, p, groups={"feed": (1, list(sm.feed))} if sm.feed else None)
    rd, groups = fa.read_msh(p, feed_group="feed" if sm.feed else None)
    T, tg, B, bt = by_tag(sm)
    np.testing.assert_array_equal(rd.vertices, sm.vertices)
    np.testing.assert_array_equal(rd.triangles, T)
    np.testing.assert_array_equal(rd.tags, tg)
    np.testing.assert_array_equal(rd.bedges, B)
    np.testing.assert_array_equal(rd.btags, bt)
    assert rd.feed == list(sm.feed)
    if np.array_equal(T, sm.triangles):
        a, b = oracle.OracleMesh(sm), oracle.OracleMesh(rd)
        assert a.N == b.N
        np.testing.assert_array_equal(a.itm, b.itm)
        np.testing.assert_array_equal(a.itp, b.itp)
        np.testing.assert_array_equal(a.edge, b.edge)
        np.testing.assert_array_equal(a.length, b.length)


def test_node_blocks_sparse_tags_and_unused_nodes(tmp_path):
    """Node blocks written out of (dim, tag) order, node tags offset by 1000, and two nodes that no
    triangle uses: the reader lists nodes block by block in (dim, tag) order and drops the unused
    ones, keeping the order of the rest."""
    sm = meshes.plate(1.0, 3)
    nv = sm.vertices.shape[0]
    extra = np.array([[5.0, 5.0, 5.0], [6.0, 6.0, 6.0]])
    sm2 = meshes.SurfaceMesh(np.vstack([sm.vertices, extra]), sm.triangles, sm.tags, sm.bedges, sm.btags)
    rng = np.random.default_rng(4)
    perm = rng.permutation(nv + 2)
    st = int(sm.tags.min())
    blocks = [(2, st + 7, list(perm[:6])), (0, 3, list(perm[6:9])), (2, st, list(perm[9:20])),
              (1, 2, list(perm[20:]))]
    p = tmp_path / "blocks.msh"
    meshes.write_msh(sm2, p, node_blocks=blocks, tag_offset=1000)
    rd, _ = fa.read_msh(p)
    order = [i for _, _, idx in sorted(blocks, key=lambda b: (b[0], b[1])) for i in idx]
    pos = {v: k for k, v in enumerate(order)}
    used = sorted({pos[int(v)] for v in sm.triangles.ravel()})
    new = {o: k for k, o in enumerate(used)}
    want_V = sm2.vertices[[order[o] for o in used]]
    want_T = np.vectorize(lambda v: new[pos[int(v)]])(sm.triangles)
    np.testing.assert_array_equal(rd.vertices, want_V)
    np.testing.assert_array_equal(rd.triangles, want_T)
    assert rd.vertices.shape[0] == nv


def test_second_order_triangles_use_their_corners(tmp_path):
    """6-node triangles (type 9) and 3-node lines (type 8): the first three (two) nodes are the
    corners; the mid-edge nodes are then unused and dropped."""
    txt = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
1
1 5 "gap"
$EndPhysicalNames
$Entities
0 1 1 0
3 0 0 0 1 1 0 1 5 0
1 0 0 0 1 1 0 0 0
$EndEntities
$Nodes
1 9 1 9
2 1 0 9
1
2
3
4
5
6
7
8
9
0 0 0
1 0 0
1 1 0
0 1 0
0.5 0 0
1 0.5 0
0.5 0.5 0
0 0.5 0
0.5 1 0
$EndNodes
$Elements
2 3 1 3
1 3 8 1
1 1 3 7
2 1 9 2
2 1 2 3 5 6 7
3 1 3 4 7 9 8
$EndElements
"""
    p = tmp_path / "o2.msh"
    p.write_text(txt)
    rd, groups = fa.read_msh(p, feed_group="gap")
    np.testing.assert_array_equal(rd.vertices, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(rd.triangles, [[0, 1, 2], [0, 2, 3]])
    np.testing.assert_array_equal(rd.bedges, [[0, 2]])
    assert groups == {"gap": (1, 5, [3])} and rd.feed == [3]


def test_groups_are_listed_by_name(tmp_path):
    sm = meshes.strip_dipole(0.5, 0.005, 10, 1)
    p = tmp_path / "g.msh"
    meshes.write_msh(sm, p, groups={"zeta": (2, [1]), "alpha": (2, [2]), "feed": (1, list(sm.feed))})
    _, groups = fa.read_msh(p)
    assert list(groups) == ["alpha", "feed", "zeta"]
    assert groups["alpha"][2] == [2] and groups["zeta"][2] == [1]
    with pytest.raises(fa.FricoError):
        fa.read_msh(p, feed_group="nope")


@pytest.mark.parametrize("mutate,msg", [
    (lambda t: t.replace("4.1 0 8", "2.2 0 8"), "format"),
    (lambda t: t.replace("4.1 0 8", "4.1 1 8"), "binary"),
    (lambda t: t.replace("$EndNodes", "$EndNodez"), "EndNodes"),
    (lambda t: t.replace("2 1 2 1\n", "2 1 99 1\n"), "element type"),
    (lambda t: t.replace("\n1 1 2 3\n", "\n1 1 2 77\n"), "node 77"),
])
def test_read_errors(tmp_path, mutate, msg):
    txt = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Nodes
1 3 1 3
2 1 0 3
1
2
3
0 0 0
1 0 0
0 1 0
$EndNodes
$Elements
1 1 1 1
2 1 2 1
1 1 2 3
$EndElements
"""
    p = tmp_path / "e.msh"
    p.write_text(txt)
    rd, _ = fa.read_msh(p)
    assert rd.triangles.tolist() == [[0, 1, 2]]
    p.write_text(mutate(txt))
    with pytest.raises(fa.FricoError, match=msg):
        fa.read_msh(p)


def test_lonely_curve_edge_is_an_error(tmp_path):
    """A line element on nodes no triangle uses is the reference's lonely edge (:177-185)."""
    sm = meshes.plate(1.0, 2)
    V = np.vstack([sm.vertices, [[9.0, 9.0, 9.0], [9.0, 8.0, 9.0]]])
    nv = sm.vertices.shape[0]
    bad = meshes.SurfaceMesh(V, sm.triangles, sm.tags, np.array([[nv, nv + 1]]), np.array([7], np.int32))
    p = tmp_path / "lonely.msh"
    meshes.write_msh(bad, p)
    with pytest.raises(fa.FricoError, match="lonely edge on curve 7"):
        fa.read_msh(p)
