"""Pins the oracle's Dunavant rules (src/quadratures.cpp:41-57, 321-372, 2685-2702) by mathematics:
polynomial exactness on the reference triangle, weight sums, Kahan area, sub-triangle rule."""
import math

import numpy as np
import pytest

import oracle

P0, P1, P2 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
NPTS = [1, 3, 4, 6, 7, 12, 13, 16, 19, 25, 27, 33, 37, 42, 48, 52, 61, 70, 73, 79]


def monomial_integral(a, b):
    # int over the unit right triangle of x^a y^b = a! b! / (a+b+2)!
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("degree", range(1, 21))
def test_point_counts_and_weight_sum(degree):
    pts, w = oracle.dunavant(degree, P0, P1, P2)
    assert len(w) == NPTS[degree - 1]
    # 15-digit tables: the weights sum to the area up to ~1e-14 (rule 3 has a negative weight)
    assert abs(w.sum() - 0.5) <= 5e-14


@pytest.mark.parametrize("degree", range(1, 21))
def test_polynomial_exactness(degree):
    pts, w = oracle.dunavant(degree, P0, P1, P2)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            q = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b)
            assert abs(q - monomial_integral(a, b)) <= 2e-14, (degree, a, b)


@pytest.mark.parametrize("degree", [2, 5, 9])
def test_not_exact_one_degree_higher(degree):
    """A mistyped table that happened to be more accurate would be suspicious too: the rule of
    degree d integrates some monomial of degree d+1 inexactly."""
    pts, w = oracle.dunavant(degree, P0, P1, P2)
    errs = []
    d = degree + 1
    for a in range(d + 1):
        b = d - a
        errs.append(abs(np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b) - monomial_integral(a, b)))
    assert max(errs) > 1e-12


def test_degree_zero_is_rule_one_and_out_of_range_raises():
    p, w = oracle.dunavant(0, P0, P1, P2)
    p1, w1 = oracle.dunavant(1, P0, P1, P2)
    np.testing.assert_array_equal(p, p1)
    with pytest.raises(ValueError):
        oracle.dunavant(21, P0, P1, P2)


def test_weights_scale_with_kahan_area_on_a_general_triangle():
    rng = np.random.default_rng(5)
    for _ in range(20):
        a, b, c = rng.normal(size=(3, 3))
        area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        pts, w = oracle.dunavant(5, a, b, c)
        assert abs(w.sum() - area) <= 1e-13 * max(area, 1)
        # degree-5 exactness in the affine frame: integral of a linear function = area * value at centroid
        f = lambda p: 2.0 * p[..., 0] - 3.0 * p[..., 1] + 0.5 * p[..., 2] + 1.0
        assert abs(np.sum(w * f(pts)) - area * f((a + b + c) / 3)) <= 1e-12 * max(area, 1)


def test_orbit_expansion_order_of_six_point_orbits():
    """Degree 6 starts with two 3-orbits then one 6-orbit (src/quadratures.cpp:93-97):
    the 6-orbit points come in the order (a,b,c),(a,c,b),(b,a,c),(b,c,a),(c,a,b),(c,b,a)."""
    pts, w = oracle.dunavant(6, np.array([1.0, 0, 0]), np.array([0.0, 1, 0]), np.array([0.0, 0, 1]))
    six = pts[6:12]
    a, b, c = 0.053145049844817, 0.310352451033784, 0.636502499121399
    expect = [(a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)]
    np.testing.assert_allclose(six, np.array(expect), atol=1e-15)


def test_subtri_rule():
    from frico_amd import meshes
    sm = meshes.plate(1.0, 1)
    om = oracle.OracleMesh(sm)
    pts = np.zeros((300, 3))
    w = np.zeros(300)
    import ctypes as C
    n = oracle.lib().orc_subtri(om.h, 0, 5, oracle._p(pts), oracle._p(w))
    assert n == 21
    assert abs(w[:n].sum() - 0.5) <= 1e-14
    # exact for quadratics on the whole triangle (piecewise rule of degree 5)
    tri = sm.vertices[sm.triangles[0]]
    A = 0.5 * np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    x = pts[:n]
    q = np.sum(w[:n] * x[:, 0] * x[:, 1])
    p, ww = oracle.dunavant(5, tri[0], tri[1], tri[2])
    assert abs(q - np.sum(ww * p[:, 0] * p[:, 1])) <= 1e-15
    assert A == pytest.approx(0.5)
