"""Near-field sampling planes on the host: the -P parser (src/sampling_planes.cpp:30-106) and the
plane's mesh (make_sampling_plane_mesh, :108-156).  The reference meshes the rectangle with gmsh,
absent here, so vertex placement is this library's structured mesh ("parity unpinned" against
gmsh); what is pinned is the rectangle, its rotation (against scipy's Rotation, an independent
implementation), the element size, and the parser's acceptance rules."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import frico_amd
from frico_amd import FricoError, parse_sampling_plane, sampling_plane_mesh


@pytest.mark.parametrize("name,axis", [("xy", (0, 0, 1)), ("yz", (0, 1, 0)), ("xz", (1, 0, 0))])
def test_short_form(name, axis):
    sp = parse_sampling_plane(f"{name}:2.5:1.5:0.25")
    assert sp.axis == axis and (sp.width, sp.height, sp.h) == (2.5, 1.5, 0.25)


def test_long_form_normalises_the_axis():
    sp = parse_sampling_plane("3:0:4:1:2:0.1")
    assert np.allclose(sp.axis, (0.6, 0.0, 0.8), atol=0, rtol=1e-15)
    assert (sp.width, sp.height, sp.h) == (1.0, 2.0, 0.1)


def test_empty_tokens_are_dropped():
    """split() skips runs of separators (src/utils.cpp:30-55)."""
    assert parse_sampling_plane("::xy:::1:2:0.5:") == parse_sampling_plane("xy:1:2:0.5")


def test_numbers_take_their_numeric_prefix():
    """std::stod reads the longest numeric prefix and ignores the rest."""
    sp = parse_sampling_plane("xz:1.5e0m: 2:0.5abc")
    assert (sp.width, sp.height, sp.h) == (1.5, 2.0, 0.5)


@pytest.mark.parametrize("spec", ["ab:1:1:1", "xy:1:1", "xy:1:1:1:1", "xy:a:1:1", "1:2:3:4:5:x", "0:0:0:1:1:1", ""])
def test_invalid_specifications(spec):
    with pytest.raises(FricoError):
        parse_sampling_plane(spec)


@pytest.mark.parametrize("spec", ["xy:2:1:0.3", "yz:2:1:0.3", "xz:2:1:0.3", "1:-2:0.5:1.2:0.8:0.1"])
def test_mesh_is_the_rectangle_turned_about_the_axis(spec):
    sp = parse_sampling_plane(spec)
    V, T = sampling_plane_mesh(sp)
    nx, ny = math.ceil(sp.width / sp.h), math.ceil(sp.height / sp.h)
    assert V.shape == ((nx + 1) * (ny + 1), 3) and T.shape == (2 * nx * ny, 3)
    # the unrotated grid, turned by +90 degrees about the axis
    x = -sp.width / 2 + sp.width * np.arange(nx + 1) / nx
    y = -sp.height / 2 + sp.height * np.arange(ny + 1) / ny
    X, Y = np.meshgrid(x, y)
    flat = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    rot = Rotation.from_rotvec(np.asarray(sp.axis) * math.pi / 2)
    np.testing.assert_allclose(V, rot.apply(flat), rtol=0, atol=1e-14)
    # every triangle is a right triangle with legs at most h, the cells tile the rectangle, and the
    # normals all point along the turned z axis
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    cr = np.cross(b - a, c - a)
    assert abs(0.5 * np.linalg.norm(cr, axis=1).sum() - sp.width * sp.height) <= 1e-12
    nz = rot.apply([0.0, 0.0, 1.0])
    assert np.allclose(cr / np.linalg.norm(cr, axis=1)[:, None], nz, atol=1e-12)
    edges = np.concatenate([np.linalg.norm(b - a, axis=1), np.linalg.norm(c - b, axis=1), np.linalg.norm(a - c, axis=1)])
    assert edges.max() <= math.hypot(sp.h, sp.h) + 1e-12


def test_named_planes_contain_their_axes():
    """The names mean what they say once the reference's rotation is applied: "xy" lies in z = 0,
    "yz" in x = 0, "xz" in y = 0 (the `normal` of the short form is the rotation axis)."""
    for name, zero in (("xy", 2), ("yz", 0), ("xz", 1)):
        V, _ = sampling_plane_mesh(parse_sampling_plane(f"{name}:2:1:0.5"))
        assert np.max(np.abs(V[:, zero])) <= 1e-15


def test_mesh_size_queries_and_errors():
    sp = frico_amd.SamplingPlane((0.0, 0.0, 1.0), 1.0, 1.0, 0.3)
    V, T = sampling_plane_mesh(sp)
    assert V.shape == (25, 3) and T.shape == (32, 3)
    with pytest.raises(FricoError):
        sampling_plane_mesh(frico_amd.SamplingPlane((0.0, 0.0, 1.0), 1.0, 1.0, 0.0))
    with pytest.raises(FricoError):
        sampling_plane_mesh(frico_amd.SamplingPlane((0.0, 0.0, 0.0), 1.0, 1.0, 0.1))
