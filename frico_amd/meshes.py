"""Deterministic surface-mesh generators for the EFIE workloads.

The reference meshes `.geo` files with the gmsh SDK (src/input_gmsh.cpp:324-462), which is not
available here.  These generators produce the same kind of input that `load_from_gmsh` hands to
`compute_connectivity`: compacted vertices, triangles tagged with their surface entity, and the
1D-entity edges that become delta-gap interfaces (src/input_gmsh.cpp:143-199, 264-269).  The
connectivity and RWG ordering rules are applied afterwards by the library, exactly as the
reference applies them to a gmsh mesh.

Every generator is pure numpy and fully determined by its arguments (seeded where random).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# src/constants.h:23-24
MU0 = 1.256637061435917e-06
EPS0 = 8.8541878188e-12
C0 = 1.0 / math.sqrt(MU0 * EPS0)

FEED_TAG = 4  # 1D entity tag of the feed curve (share/examples/dipole/dipole.geo:39)


def wavelength(freq: float) -> float:
    return C0 / freq


@dataclass
class SurfaceMesh:
    """Triangulated PEC surface, as produced by gmsh ingestion."""

    vertices: np.ndarray  # (nv, 3) float64
    triangles: np.ndarray  # (nt, 3) int64, vertex indices
    tags: np.ndarray  # (nt,) int32 surface entity tag
    bedges: np.ndarray  # (nb, 2) int64, 1D-entity edges (interfaces when internal)
    btags: np.ndarray  # (nb,) int32 1D entity tag
    name: str = "mesh"
    feed: list = field(default_factory=list)  # interface tags of the delta-gap source

    def __post_init__(self):
        self.vertices = np.ascontiguousarray(self.vertices, dtype=np.float64)
        self.triangles = np.ascontiguousarray(self.triangles, dtype=np.int64)
        self.tags = np.ascontiguousarray(self.tags, dtype=np.int32)
        self.bedges = np.ascontiguousarray(self.bedges, dtype=np.int64).reshape(-1, 2)
        self.btags = np.ascontiguousarray(self.btags, dtype=np.int32)

    @property
    def nv(self) -> int:
        return self.vertices.shape[0]

    @property
    def nt(self) -> int:
        return self.triangles.shape[0]

    def num_internal_edges(self) -> int:
        e = np.sort(np.concatenate([self.triangles[:, [0, 1]], self.triangles[:, [1, 2]],
                                    self.triangles[:, [2, 0]]]), axis=1)
        _, counts = np.unique(e, axis=0, return_counts=True)
        return int(np.sum(counts == 2))


def _grid_cells(nx: int, ny: int, vid, tag: int):
    """Two triangles per cell, split along the (i,j)-(i+1,j+1) diagonal."""
    tris = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    return tris, [tag] * len(tris)


def strip_dipole(length: float = 0.5, width: float = 0.005, nx: int = 200, ny: int = 1,
                 name: str = "strip_dipole") -> SurfaceMesh:
    """Flat PEC strip along x, centred at the origin in z=0, fed at x=0.

    `nx` cells along the whole length (even), `ny` across.  The x>0 arm is surface 1 and its
    triangles come first, so T- of every feed edge lies in the x>0 arm (as with
    share/examples/dipole/dipole.geo:31-39).  N = 3*nx*ny - nx - ny.
    """
    if nx % 2:
        raise ValueError("nx must be even so that a vertex column sits on the feed")
    xs = np.linspace(-length / 2, length / 2, nx + 1)
    xs[nx // 2] = 0.0
    ys = np.linspace(-width / 2, width / 2, ny + 1)
    verts = np.array([(x, y, 0.0) for x in xs for y in ys])

    def vid(i, j):
        return i * (ny + 1) + j

    h = nx // 2
    t1, g1 = _grid_cells(h, ny, lambda i, j: vid(h + i, j), 1)  # x > 0 arm first
    t2, g2 = _grid_cells(h, ny, lambda i, j: vid(i, j), 2)
    feed = [(vid(h, j), vid(h, j + 1)) for j in range(ny)]
    return SurfaceMesh(verts, np.array(t1 + t2), np.array(g1 + g2), np.array(feed),
                       np.full(len(feed), FEED_TAG), name=name, feed=[FEED_TAG])


def plate(size: float, n: int, name: str = "plate") -> SurfaceMesh:
    """Square PEC plate [-s/2, s/2]^2 in z=0, n x n cells -> T = 2n^2, N = 3n^2 - 2n."""
    xs = np.linspace(-size / 2, size / 2, n + 1)
    verts = np.array([(x, y, 0.0) for y in xs for x in xs])

    def vid(i, j):
        return j * (n + 1) + i

    tris, tags = _grid_cells(n, n, vid, 1)
    return SurfaceMesh(verts, np.array(tris), np.array(tags), np.zeros((0, 2)), np.zeros(0),
                       name=name)


def _rotation(a: float, b: float, c: float) -> np.ndarray:
    ca, sa, cb, sb, cc, sc = math.cos(a), math.sin(a), math.cos(b), math.sin(b), math.cos(c), math.sin(c)
    rz1 = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rz2 = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return rz2 @ ry @ rz1


def icosphere(radius: float, nu: int, rotation=(0.3141, 0.7071, 1.1213),
              name: str = "icosphere") -> SurfaceMesh:
    """Geodesic sphere: every icosahedron face split into nu^2 triangles, projected radially.

    T = 20 nu^2, N = 30 nu^2, V = 10 nu^2 + 2.  A generic rotation is applied so that no edge is
    parallel to z (the reference's edge length is hypot(dx, dy), src/geom_mesh.cpp:51).
    """
    p = (1 + math.sqrt(5)) / 2
    base = np.array([(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p),
                     (0, -1, -p), (0, 1, -p), (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)],
                    dtype=np.float64)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    index = {}
    verts = []

    def vkey(wts):
        return tuple(sorted((v, w) for v, w in wts if w))

    def vertex(wts):
        key = vkey(wts)
        if key not in index:
            pt = sum(base[v] * w for v, w in wts) / nu
            index[key] = len(verts)
            verts.append(pt / np.linalg.norm(pt) * radius)
        return index[key]

    tris = []
    for (A, B, C) in faces:
        def vid(i, j):
            return vertex(((A, nu - i - j), (B, i), (C, j)))
        for i in range(nu):
            for j in range(nu - i):
                tris.append((vid(i, j), vid(i + 1, j), vid(i, j + 1)))
                if i + j < nu - 1:
                    tris.append((vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)))
    R = _rotation(*rotation)
    verts = np.array(verts) @ R.T
    tris = np.array(tris)
    return SurfaceMesh(verts, tris, np.ones(len(tris), dtype=np.int32), np.zeros((0, 2)),
                       np.zeros(0), name=name)


def bowtie(arm: float, flare_deg: float = 60.0, nx: int = 20, ny: int = 17,
           feed_width: float | None = None, name: str = "bowtie") -> SurfaceMesh:
    """Planar bowtie in z=0: two trapezoidal arms along +/-x joined at the feed line x=0.

    Arm half-width at |x| is feed_width/2 + |x| tan(flare/2).  The ny edges on x=0 form the
    delta-gap interface.  x>0 arm is surface 1 (first), x<0 arm surface 2.
    N = 6 nx ny - 2 nx - ny.
    """
    if feed_width is None:
        feed_width = arm / 20
    t = math.tan(math.radians(flare_deg) / 2)
    xs = np.linspace(-arm, arm, 2 * nx + 1)
    xs[nx] = 0.0
    verts = []
    for x in xs:
        h = feed_width / 2 + abs(x) * t
        for j in range(ny + 1):
            verts.append((x, h * (2.0 * j / ny - 1.0), 0.0))
    verts = np.array(verts)

    def vid(i, j):
        return i * (ny + 1) + j

    t1, g1 = _grid_cells(nx, ny, lambda i, j: vid(nx + i, j), 1)
    t2, g2 = _grid_cells(nx, ny, lambda i, j: vid(i, j), 2)
    feed = [(vid(nx, j), vid(nx, j + 1)) for j in range(ny)]
    return SurfaceMesh(verts, np.array(t1 + t2), np.array(g1 + g2), np.array(feed),
                       np.full(len(feed), FEED_TAG), name=name, feed=[FEED_TAG])


def strip_loop(radius: float, width: float, nphi: int, name: str = "strip_loop") -> SurfaceMesh:
    """Flat annular strip loop in z=0 (one radial cell), fed on the radial edge at phi=0.

    N = 2 nphi.  The first half of the loop is surface 1 and comes first.
    """
    verts = []
    for i in range(nphi):
        ph = 2 * math.pi * i / nphi
        c, s = math.cos(ph), math.sin(ph)
        verts.append(((radius - width / 2) * c, (radius - width / 2) * s, 0.0))
        verts.append(((radius + width / 2) * c, (radius + width / 2) * s, 0.0))
    verts = np.array(verts)
    tris, tags = [], []
    for i in range(nphi):
        a, b = 2 * i, 2 * i + 1
        c, d = 2 * ((i + 1) % nphi), 2 * ((i + 1) % nphi) + 1
        tg = 1 if i < nphi // 2 else 2
        tris += [(a, c, d), (a, d, b)]
        tags += [tg, tg]
    feed = [(0, 1)]
    return SurfaceMesh(verts, np.array(tris), np.array(tags), np.array(feed),
                       np.full(1, FEED_TAG), name=name, feed=[FEED_TAG])


def batch_workload(count: int = 32, seed0: int = 20261203, nmin: int = 500, nmax: int = 4000):
    """Uneven batch: `count` strip dipoles / strip loops with N ~ U[nmin, nmax].

    Mesh k uses numpy's PCG64 seeded with seed0 + k.  Returns a list of SurfaceMesh.
    """
    out = []
    for k in range(count):
        rng = np.random.default_rng(seed0 + k)
        kind = int(rng.integers(0, 2))
        n = int(rng.integers(nmin, nmax + 1))
        if kind == 0:
            nx = 2 * max(1, (n + 1) // 4)  # N = 2 nx - 1 for one cell across
            out.append(strip_dipole(0.5, 0.005, nx, 1, name=f"dipole{k}"))
        else:
            out.append(strip_loop(0.16, 0.005, max(8, n // 2), name=f"loop{k}"))
    return out


def batch_frequencies(n: int = 16, f0: float = 100e6, f1: float = 500e6):
    """f_i = f0 + i (f1 - f0)/(n-1), computed by index rather than accumulation."""
    return [f0 + i * (f1 - f0) / (n - 1) for i in range(n)]


# The benchmark/test workloads named by their shapes.
def workload_dipole_c1() -> SurfaceMesh:
    """Strip dipole 0.5 m x 5 mm, 200 x 1 cells: T=400, V=402, N=399."""
    return strip_dipole(0.5, 0.005, 200, 1, name="strip_dipole_N399")


def workload_sphere(freq: float = 300e6, nu: int = 14) -> SurfaceMesh:
    """Sphere of radius one wavelength, nu=14: T=3920, N=5880."""
    return icosphere(wavelength(freq), nu, name=f"sphere_N{30 * nu * nu}")


def workload_plate(freq: float = 300e6, n: int = 64) -> SurfaceMesh:
    """3 x 3 wavelength plate, 64 x 64 squares: T=8192, V=4225, N=12160."""
    return plate(3 * wavelength(freq), n, name=f"plate_N{3 * n * n - 2 * n}")


def workload_bowtie(freq: float = 300e6, nx: int = 20, ny: int = 17) -> SurfaceMesh:
    """Bowtie, arm 0.25 wavelength, flare 60 degrees: N = 6 nx ny - 2 nx - ny (1983)."""
    return bowtie(0.25 * wavelength(freq), 60.0, nx, ny, name="bowtie")


def write_msh(sm: SurfaceMesh, path, groups=None, node_blocks=None, tag_offset: int = 0) -> None:
    """Writes a SurfaceMesh as a gmsh 4.1 ASCII .msh file (for meshes made elsewhere and for tests).

    Surfaces are the distinct triangle tags, curves the distinct edge tags; groups maps a physical
    name to (dim, [entity tags]).  node_blocks: optional list of (dim, entity tag, vertex indices)
    giving the $Nodes blocks in file order (default: every vertex in one block on the first
    surface).  Node tags are vertex index + 1 + tag_offset.
    """
    groups = groups or {}
    stags = sorted(set(int(t) for t in sm.tags))
    ctags = sorted(set(int(t) for t in sm.btags))
    pnum = {}  # (dim, name) -> physical tag
    for i, (name, (dim, _)) in enumerate(sorted(groups.items()), start=1):
        pnum[(dim, name)] = i

    def phys(dim, etag):
        return [pnum[(d, n)] for n, (d, ents) in sorted(groups.items()) if d == dim and etag in ents]

    V = sm.vertices
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    if groups:
        lines.append("$PhysicalNames")
        lines.append(str(len(groups)))
        for (dim, name), t in sorted(pnum.items(), key=lambda kv: kv[1]):
            lines.append(f'{dim} {t} "{name}"')
        lines.append("$EndPhysicalNames")
    lines.append("$Entities")
    lines.append(f"0 {len(ctags)} {len(stags)} 0")
    lo, hi = V.min(axis=0), V.max(axis=0)
    box = " ".join(repr(float(x)) for x in (*lo, *hi))
    for t in ctags:
        p = phys(1, t)
        lines.append(f"{t} {box} {len(p)} {' '.join(map(str, p))} 0".replace("  ", " "))
    for t in stags:
        p = phys(2, t)
        lines.append(f"{t} {box} {len(p)} {' '.join(map(str, p))} 0".replace("  ", " "))
    lines.append("$EndEntities")
    if node_blocks is None:
        node_blocks = [(2, stags[0], list(range(V.shape[0])))]
    nn = sum(len(b[2]) for b in node_blocks)
    alltags = [i + 1 + tag_offset for b in node_blocks for i in b[2]]
    lines.append("$Nodes")
    lines.append(f"{len(node_blocks)} {nn} {min(alltags)} {max(alltags)}")
    for dim, etag, idx in node_blocks:
        lines.append(f"{dim} {etag} 0 {len(idx)}")
        lines += [str(i + 1 + tag_offset) for i in idx]
        lines += [" ".join(repr(float(x)) for x in V[i]) for i in idx]
    lines.append("$EndNodes")
    eblocks = []
    for t in ctags:
        e = sm.bedges[sm.btags == t]
        eblocks.append((1, t, 1, e))
    for t in stags:
        e = sm.triangles[sm.tags == t]
        eblocks.append((2, t, 2, e))
    ne = sum(len(b[3]) for b in eblocks)
    lines.append("$Elements")
    lines.append(f"{len(eblocks)} {ne} 1 {ne}")
    et = 1
    for dim, t, typ, e in eblocks:
        lines.append(f"{dim} {t} {typ} {len(e)}")
        for row in e:
            lines.append(f"{et} " + " ".join(str(int(v) + 1 + tag_offset) for v in row))
            et += 1
    lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
