"""Pins of the oracle's membrane bending chain (sem_kernel.cu:724-1128,
1499-1638; topology gpuSEM_LB.c:846-936):

- mean curvature from the quadratic 6-node triangles against the closed form
  of a sphere (km = -1/R for the generator's winding, +1/R for the opposite
  winding) and of an ellipsoid;
- the ring Laplacian on planar patches: zero for linear functions and exactly
  4 for x^2 + y^2 on acute triangulations (the cotangent weights and the
  Voronoi area), and the regular hexagonal cell area sqrt(3)/2;
- the bending force is independent of the winding (curvature and normal
  change sign together)."""
import ctypes as C

import numpy as np
import pytest

from helpers import ModelArrays


def ellipsoid(pfg, n, a, b, c):
    L = pfg.lib()
    x, y, z = (np.empty(n, np.float32) for _ in range(3))
    t = np.empty(3 * (2 * n - 4), np.int32)
    nt = C.c_int32()
    st = L.pfg_geom_ellipsoid(n, a, b, c, 0.0, 0.0, 0.0, 0.0, 1,
                              *[q.ctypes.data_as(C.POINTER(C.c_float)) for q in (x, y, z)],
                              t.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nt))
    assert st == 0
    return np.stack([x, y, z], 1).astype(np.float64), t.reshape(-1, 3)


def oracle_for(pfg, po, P, tri, kbend=0.0, shift=40.0):
    p = pfg.default_params(nx=96, ny=96, nz=96)
    p.kbend = kbend
    n = len(P)
    ma = ModelArrays(P[:, 0] + shift, P[:, 1] + shift, P[:, 2] + shift, plt_elem_start=[0, n],
                     plt_tri_start=[0, len(tri)], tri=tri.ravel(), tri_S0=np.ones(len(tri)),
                     plt_S0=[1.0], plt_V0=[1.0])
    o = po.Oracle(p, model=ma.model)
    o._keep = ma
    return o


@pytest.mark.parametrize("n", [1000, 4000])
def test_sphere_mean_curvature(pfg, po, n):
    R = 10.0
    P, tri = ellipsoid(pfg, n, R, R, R)
    km, fb, nrm = oracle_for(pfg, po, P, tri).bend()
    # the generator winds e01 x e02 outward, so n_elem and the node normals
    # point inward and K = grad n gives km = -1/R
    assert np.mean(np.sum(nrm * P, 1) / np.linalg.norm(P, axis=1)) < -0.999
    assert abs(km.mean() * R + 1.0) < (0.025 if n == 1000 else 0.006)
    flipped = tri[:, [0, 2, 1]]
    km2, _, nrm2 = oracle_for(pfg, po, P, flipped).bend()
    assert np.allclose(km2, -km, rtol=1e-9, atol=1e-12)
    assert np.allclose(nrm2, -nrm, atol=1e-12)


def test_ellipsoid_mean_curvature(pfg, po):
    a, b, c = 8.0, 6.0, 4.0
    P, tri = ellipsoid(pfg, 4000, a, b, c)
    km, _, _ = oracle_for(pfg, po, P, tri).bend()
    X, Y, Z = P.T
    H = (X * X + Y * Y + Z * Z - a * a - b * b - c * c) / (
        2 * a * a * b * b * c * c * (X * X / a ** 4 + Y * Y / b ** 4 + Z * Z / c ** 4) ** 1.5)
    rel = np.abs(km - H) / np.abs(H)
    assert np.mean(rel) < 0.02 and np.median(rel) < 0.01


def test_bending_force_independent_of_winding(pfg, po):
    P, tri = ellipsoid(pfg, 1000, 8.0, 6.0, 4.0)
    _, fb, _ = oracle_for(pfg, po, P, tri, kbend=0.7).bend()
    _, fb2, _ = oracle_for(pfg, po, P, tri[:, [0, 2, 1]], kbend=0.7).bend()
    assert np.abs(fb).max() > 0
    assert np.allclose(fb, fb2, rtol=1e-8, atol=1e-12 * np.abs(fb).max())
    _, fb0, _ = oracle_for(pfg, po, P, tri, kbend=0.0).bend()
    assert np.abs(fb0).max() == 0


def hex_patch(rows, cols, jitter, seed):
    """Planar triangulated patch of a hexagonal lattice, unit spacing,
    counter-clockwise triangles, optional seeded jitter of the vertices."""
    rng = np.random.default_rng(seed)
    P = []
    for r in range(rows):
        for q in range(cols):
            P.append([q + 0.5 * (r % 2), r * np.sqrt(3) / 2, 0.0])
    P = np.array(P)
    P[:, :2] += jitter * rng.uniform(-1, 1, (len(P), 2))
    tri = []
    for r in range(rows - 1):
        for q in range(cols - 1):
            a, b = r * cols + q, r * cols + q + 1
            c, d = (r + 1) * cols + q, (r + 1) * cols + q + 1
            if r % 2 == 0:
                tri += [(a, b, c), (b, d, c)]
            else:
                tri += [(a, b, d), (a, d, c)]
    return P, np.array(tri, np.int32)


@pytest.mark.parametrize("jitter", [0.0, 0.08])
def test_ring_laplacian_planar(pfg, po, jitter):
    P, tri = hex_patch(9, 9, jitter, 3)
    o = oracle_for(pfg, po, P, tri)
    Pm = P + 40.0  # the positions the oracle holds (float32 of the shifted values)
    Pm = Pm.astype(np.float32).astype(np.float64)
    lap1, area = o.ring_laplacian(2.0 * Pm[:, 0] - 3.0 * Pm[:, 1] + 1.0)
    inner = area > 0
    assert inner.sum() == 7 * 7  # every vertex off the border has a closed ring
    assert np.abs(lap1[inner]).max() < 1e-9
    q = (Pm[:, 0] - 40.0) ** 2 + (Pm[:, 1] - 40.0) ** 2
    lap2, _ = o.ring_laplacian(q)
    assert np.abs(lap2[inner] - 4.0).max() < 1e-9
    if jitter == 0.0:  # the hexagonal Voronoi cell (positions held in FP32)
        assert np.allclose(area[inner], np.sqrt(3) / 2, rtol=2e-6)
