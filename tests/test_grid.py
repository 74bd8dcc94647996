"""Icosahedral grid builder: closed forms and invariants (CPU).

No shipped reference fixture pins the grid (THOR's grid_test references are
written at run time, src/test/grid_test.cpp:138-154); these checks pin what
the construction fixes: point counts and ordering (grid.cu:113-141, 631-656),
the sphere-area normalisation (grid.cu:2410-2419), neighbour topology, and
the consistency of the discrete operators with closed-form fields.
"""
import numpy as np
import pytest

from tests.common import grid

A = 6.371e6


@pytest.fixture(scope="module", params=[4, 5])
def g(request, built):
    return grid(built, request.param, 8)


def test_counts_and_pentagons(g):
    n = 2 ** g.glevel
    assert g.P == 2 + 10 * n * n
    assert g.nr == (g.P - 2) // 256
    nfaces = 4 ** (g.glevel - 4)
    assert list(g.pent_ind[:10]) == [k * nfaces * 256 for k in range(10)]
    assert list(g.pent_ind[10:]) == [g.P - 2, g.P - 1]
    # centroid relocation (grid.cu:1873-1921) moves even the poles by < 1 mm
    assert np.allclose(g.xyz[g.P - 2], [0, 0, A], atol=1e-3) and np.allclose(g.xyz[g.P - 1], [0, 0, -A], atol=1e-3)
    # every pentagon is a vertex of the base icosahedron: |z| in {A, A/sqrt(5)}
    zabs = np.sort(np.abs(g.xyz[g.pent_ind, 2]) / A)
    assert np.allclose(zabs[:10], 1 / np.sqrt(5), atol=1e-9) and np.allclose(zabs[10:], 1.0)


def test_area_normalisation(g):
    assert abs(g.areasT.sum() / (4 * np.pi * A * A) - 1) < 1e-12
    # each triangle counted twice per vertex pair... sub-triangles sum to the triangle
    sub = g.areas.sum(axis=2)
    hexa = np.ones(g.P, bool)
    hexa[g.pent_ind] = False
    assert (g.areas[hexa] > 0).all()
    # triangle and sub-triangle normalisations differ by a factor close to 1
    ratio = sub[hexa] / g.areasTr[hexa]
    assert np.allclose(ratio, ratio.mean(), rtol=1e-12) and abs(ratio.mean() - 1) < 1e-3


def test_neighbour_topology(g):
    pl = g.point_local
    pent = set(int(i) for i in g.pent_ind)
    for p in range(g.P):
        n = 5 if p in pent else 6
        ring = pl[p, :n]
        assert len(set(ring)) == n and (ring >= 0).all()
        for q in ring:  # symmetric adjacency
            assert p in pl[q, : (5 if int(q) in pent else 6)]
        # consecutive ring members are mutually adjacent (a closed fan of triangles)
        for k in range(n):
            a, b = int(ring[k]), int(ring[(k + 1) % n])
            assert b in pl[a, : (5 if a in pent else 6)]


def test_ring_is_rhombus_order(g):
    # pt1..pt6 = (x, y+1), (x+1, y+1), (x+1, y), (x, y-1), (x-1, y-1), (x-1, y)
    # inside a 16x16 tile (grid.cu:1495-1621, thor_fastmodes.h:122-127)
    for r in range(g.nr):
        for y in range(1, 15):
            for x in range(1, 15):
                p = r * 256 + y * 16 + x
                want = [p + 16, p + 17, p + 1, p - 16, p - 17, p - 1]
                assert list(g.point_local[p]) == want


def test_maps_consistent_with_rings(g):
    m = g.maps
    for r in range(g.nr):
        for y in range(16):
            for x in range(16):
                p = r * 256 + y * 16 + x
                assert m[r, y + 1, x + 1] == p
        pent = (r * 256) in set(int(i) for i in g.pent_ind)
        assert (m[r, 0, 0] == -1) == pent
        for y in range(16):
            p = r * 256 + y * 16
            assert m[r, y + 1, 0] == g.point_local[p, 4 if (pent and y == 0) else 5]
            assert m[r, y + 1, 17] == g.point_local[p + 15, 2]
        for x in range(16):
            assert m[r, 0, x + 1] == g.point_local[r * 256 + x, 3]
            assert m[r, 17, x + 1] == g.point_local[r * 256 + 15 * 16 + x, 0]


def test_gradient_of_constant_is_zero(g):
    s = g.grad.sum(axis=1)
    assert np.abs(s).max() <= 1e-12 * np.abs(g.grad).max()


def test_operators_on_closed_form_fields(g):
    """div of solid-body rotation ~ 0; grad of z is the tangential projection of e_z."""
    P = g.P
    r = g.func_r
    slots = np.concatenate([np.arange(P)[:, None], g.point_local[:, :5],
                            np.where(g.point_local[:, 5:6] < 0, g.point_local[:, 4:5], g.point_local[:, 5:6])],
                           axis=1)
    omega = np.array([0.3, -0.2, 1.0])
    u = np.cross(omega, g.xyz)                     # tangent, divergence-free
    divu = np.einsum("pjk,pjk->p", g.div, u[slots])
    scale = np.abs(u).max() / (A * np.sqrt(4 * np.pi / g.P))
    assert np.abs(divu).max() < 2e-3 * scale
    f = g.xyz[:, 2] / A                            # f = sin(lat)
    gradf = np.einsum("pjk,pj->pk", g.grad, f[slots])
    exact = (np.array([0, 0, 1.0])[None, :] - r[:, 2:3] * r) / A
    err = np.abs(gradf - exact).max() / np.abs(exact).max()
    assert err < 5e-3


def test_uniform_vertical_grid(g):
    top = 36000.0
    assert np.allclose(g.Altitudeh, np.linspace(0, top, g.nv + 1), rtol=0, atol=1e-9)
    assert np.allclose(g.Altitude, 0.5 * (g.Altitudeh[1:] + g.Altitudeh[:-1]))


def test_spring_dynamics_converged(built):
    g4 = grid(built, 4, 4)
    assert np.allclose(np.linalg.norm(g4.xyz, axis=1), A, rtol=1e-14)
    # pentagons stay fixed at the icosahedron's vertices
    zp = np.abs(g4.xyz[g4.pent_ind[:10], 2]) / A
    assert np.allclose(zp, 1 / np.sqrt(5), atol=1e-9)


def _rotation(angle, flip=False):
    c, s = np.cos(angle), np.sin(angle)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, -1.0 if flip else 1.0]])


@pytest.mark.parametrize("angle,flip", [(2 * np.pi / 5, False), (np.pi / 5, True)])
def test_icosahedral_symmetry(g, angle, flip):
    """The construction (lattice bisection, spring dynamics, centroid
    relocation, control areas) commutes with the icosahedron's symmetries
    about the polar axis: a 72-degree rotation and a 36-degree rotation
    combined with z -> -z map the point set onto itself and carry every
    control area with it, to round-off. A wrong neighbour, lattice index or
    force term breaks the symmetry at the size of the grid spacing."""
    from scipy.spatial import cKDTree

    x = g.xyz / A
    dist, idx = cKDTree(x).query(x @ _rotation(angle, flip).T)
    assert len(set(idx.tolist())) == g.P
    assert dist.max() < 1e-12
    assert np.abs(g.areasT[idx] - g.areasT).max() < 1e-10 * g.areasT.max()


def _normproj(v1, v2):
    """Great-circle arc from v2 to v1 as (unit normal) x (angle)."""
    c = np.cross(v1, v2)
    l2 = np.linalg.norm(c, axis=-1, keepdims=True)
    return c / l2 * np.arctan2(l2[..., 0], np.einsum("...k,...k->...", v1, v2))[..., None]


def test_centres_are_control_volume_centroids(g):
    """After spring dynamics every point is moved to the spherical centroid of
    its control volume (grid.cu:1838-1900): the normalised sum of the arcs
    between consecutive corners xyzq, in corner order."""
    q = g.xyzq / A
    pent = np.zeros(g.P, bool)
    pent[g.pent_ind] = True
    s = np.zeros((g.P, 3))
    for j in range(1, 6):
        with np.errstate(invalid="ignore"):  # a pentagon's unused sixth corner is the zero vector
            arc = _normproj(q[:, j], q[:, j - 1])
        s += np.where((pent & (j == 5))[:, None], 0.0, arc)
    last = np.where(pent, 4, 5)
    s += _normproj(q[:, 0], q[np.arange(g.P), last])
    want = s / np.linalg.norm(s, axis=1, keepdims=True)
    # (both arrays were scaled by A and back, and the short arcs cancel in
    # their cross products: round-off; a wrong corner moves it by ~1e-3)
    assert np.abs(g.xyz / A - want).max() < 1e-12
