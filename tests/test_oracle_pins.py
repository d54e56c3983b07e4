"""Pins of three oracle functions against what the mathematics fixes, each
check independent of the formula the oracle codes:

* wall repulsion on the lowest element (sem_kernel.cu:409-430): the force is
  minus the gradient of the potential U(h) = -a ln(1 - e^{-kappa h}) (finite
  differences), acts in +z on the lowest element of each platelet only, is
  zero above the cutoff and constant below hc/100;
* the receptor bond spring and its split over the triangle's vertices
  (sem_kernel.cu:2039-2065): the vertex forces sum to minus the gradient of the
  spring energy k (|d| - l0)^2 / 2 under a rigid translation (finite
  differences), a stretched bond pulls toward its ligand, the shares are
  inversely proportional to the receptor-vertex distances, and the special
  positions (a vertex, the centroid of an equilateral triangle) give the
  shares that fixes;
* the marker update x += v dt (sem_kernel.cu:2087-2148): in a uniform flow,
  an exact fixed point of the lattice update, free markers read the flow
  velocity exactly (partition of unity) and move by u dt per step, anchored
  ones not at all.
"""
import numpy as np

from helpers import ModelArrays


def _tetra(z0):
    """Four elements of one platelet: a tetrahedron whose lowest vertex
    (index 2) sits at height z0."""
    v = np.array([[8.0, 8.0, z0 + 1.0], [9.0, 8.0, z0 + 1.0], [8.5, 8.5, z0],
                  [8.5, 8.3, z0 + 1.8]])
    tri = np.array([[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]], dtype=np.int32)
    return v, tri


def _tetra_model(po, v, tri):
    S = np.array([0.5 * np.linalg.norm(np.cross(v[b] - v[a], v[c] - v[a])) for a, b, c in tri])
    return ModelArrays(v[:, 0], v[:, 1], v[:, 2], plt_elem_start=[0, 4], plt_tri_start=[0, 4],
                       tri=tri.ravel(), tri_S0=S, plt_S0=[S.sum()], plt_V0=[-1.0])


def _wall_params(pfg, a, kappa, hc):
    p = pfg.default_params(nx=16, ny=16, nz=16)
    p.wall[2] = pfg.WALL_BOUNCE
    p.wall_z = -0.5
    p.morse_u0 = 0.0
    p.wall_rep_a, p.wall_rep_kappa, p.wall_rep_hc = a, kappa, hc
    return p


def _forces(po, p, M):
    o = po.Oracle(p, model=M.model)
    for st in (0, 1, 2, 3):  # cells, membrane, bonds, forces
        o.stage(st)
    F = np.stack([o.marker(k).copy() for k in ("fx", "fy", "fz")], 1)
    o.close()
    return F


def test_wall_repulsion_is_the_gradient_of_its_potential(pfg, po):
    a, kappa, hc = 0.3, 4.0, 1.5
    U = lambda h: -a * np.log1p(-np.exp(-kappa * h))  # noqa: E731
    p = _wall_params(pfg, a, kappa, hc)
    for h in (0.05, 0.2, 0.7, 1.4):
        v, tri = _tetra(p.wall_z + h)
        F = _forces(po, p, _tetra_model(po, v, tri))
        # the height the oracle sees is that of the FP32 position
        hf = float(np.float32(v[2, 2])) - float(np.float32(p.wall_z))
        dh = 1e-6
        fd = -(U(hf + dh) - U(hf - dh)) / (2 * dh)
        assert F[2, 2] > 0
        assert abs(F[2, 2] - fd) < 1e-6 * abs(fd)
        assert np.all(F[:, :2] == 0) and np.all(F[[0, 1, 3], 2] == 0)
    # small gaps: a/h (the potential's -a ln(kappa h) limit)
    v, tri = _tetra(p.wall_z + 0.02)
    F = _forces(po, p, _tetra_model(po, v, tri))
    hf = float(np.float32(v[2, 2])) - float(np.float32(p.wall_z))
    assert abs(F[2, 2] * hf / a - 1) < 0.05
    # above the cutoff: nothing
    v, tri = _tetra(p.wall_z + hc + 0.1)
    assert np.all(_forces(po, p, _tetra_model(po, v, tri)) == 0)
    # below hc / 100 the gap is clamped: the same force at any smaller gap
    v1, tri = _tetra(p.wall_z + 0.004)
    v2, _ = _tetra(p.wall_z + 0.001)
    F1 = _forces(po, p, _tetra_model(po, v1, tri))
    F2 = _forces(po, p, _tetra_model(po, v2, tri))
    fd = -(U(0.015 + 1e-7) - U(0.015 - 1e-7)) / 2e-7
    assert F1[2, 2] == F2[2, 2] and abs(F1[2, 2] - fd) < 1e-5 * fd
    # the force follows the lowest element: lift vertex 2 above vertex 0
    v, tri = _tetra(p.wall_z + 0.3)
    v[2, 2] += 2.0
    F = _forces(po, p, _tetra_model(po, v, tri))
    low = int(np.argmin(v[:, 2]))
    assert F[low, 2] > 0 and np.count_nonzero(F) == 1


def _bond_setup(pfg, po, verts, r1, r2, lig, k=0.7, l0=0.3):
    p = pfg.default_params(nx=16, ny=16, nz=16)
    p.morse_u0 = 0.0
    p.k_on = 0.0
    p.k_off0 = 0.0  # the held bond never ruptures
    p.bond_k, p.bond_l0, p.bond_reach = k, l0, 4.0
    v = np.asarray(verts, dtype=np.float32).astype(np.float64)
    tri = np.array([0, 1, 2], dtype=np.int32)
    S = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
    M = ModelArrays(v[:, 0], v[:, 1], v[:, 2], plt_elem_start=[0, 3], plt_tri_start=[0, 1],
                    tri=tri, tri_S0=[S], plt_S0=[S], plt_V0=[-1.0], rec_tri=[0],
                    rec_r1=[r1], rec_r2=[r2], lig_x=[lig[0]], lig_y=[lig[1]], lig_z=[lig[2]])
    o = po.Oracle(p, model=M.model)
    o.set_bonds(np.array([0], np.int32))
    for st in (0, 1, 2, 3):
        o.stage(st)
    F = np.stack([o.marker(q).copy() for q in ("fx", "fy", "fz")], 1)
    _, _, rf = o.bonds()
    o.close()
    return v, F, rf[0], p


def test_bond_spring_is_the_gradient_of_its_energy(pfg, po):
    # equilateral triangle, receptor at the centroid (r1 = 2/3, r2 = 1/2)
    s3 = np.sqrt(3.0)
    verts = [[6.0, 6.0, 3.0], [7.0, 6.0, 3.0], [6.5, 6.0 + s3 / 2, 3.0]]
    lig = [6.9, 7.4, 1.2]
    v, F, rf, p = _bond_setup(pfg, po, verts, 2.0 / 3.0, 0.5, lig)
    ligv = np.asarray(lig, np.float32).astype(np.float64)
    r1, r2 = float(np.float32(2.0 / 3.0)), float(np.float32(0.5))
    rec = (1 - r1) * v[0] + r1 * (1 - r2) * v[1] + r1 * r2 * v[2]
    k, l0 = float(np.float32(p.bond_k)), float(np.float32(p.bond_l0))
    E = lambda t: 0.5 * k * (np.linalg.norm(ligv - (rec + t)) - l0) ** 2  # noqa: E731
    h = 1e-6
    grad = np.array([(E(h * e) - E(-h * e)) / (2 * h) for e in np.eye(3)])
    total = F.sum(0)
    assert np.allclose(total, -grad, rtol=0, atol=1e-8 * np.abs(grad).max())
    assert np.allclose(rf, total, rtol=1e-12, atol=0)
    # stretched (|d| = 2.0 > l0): pulls the receptor toward the ligand
    assert np.dot(total, ligv - rec) > 0
    # equal distances to the three vertices: equal shares
    assert np.allclose(F[0], F[1], rtol=1e-6) and np.allclose(F[1], F[2], rtol=1e-6)


def test_bond_split_is_inverse_distance(pfg, po):
    verts = [[5.0, 5.0, 2.0], [7.5, 5.4, 2.3], [5.8, 7.9, 2.6]]
    lig = [6.0, 6.5, 0.4]
    r1, r2 = 0.55, 0.3
    v, F, rf, _ = _bond_setup(pfg, po, verts, r1, r2, lig)
    a, b = float(np.float32(r1)), float(np.float32(r2))
    rec = (1 - a) * v[0] + a * (1 - b) * v[1] + a * b * v[2]
    dist = np.linalg.norm(v - rec, axis=1)
    # each vertex carries the bond force times its share w_k, with
    # w_k proportional to 1 / r_k and summing to one
    w = np.array([np.dot(F[q], rf) / np.dot(rf, rf) for q in range(3)])
    assert abs(w.sum() - 1) < 1e-12
    assert np.allclose(w * dist, (w * dist)[0], rtol=1e-9)
    assert np.allclose(F, w[:, None] * rf[None, :], rtol=1e-12, atol=1e-15)
    # a receptor on a vertex (r1 = 0) puts the whole force there
    v, F, rf, _ = _bond_setup(pfg, po, verts, 0.0, 0.4, lig)
    assert np.allclose(F[0], rf, rtol=1e-12) and np.all(np.abs(F[1:]) < 1e-20)
    # a compressed bond (|d| < l0) pushes the receptor away
    v, F, rf, _ = _bond_setup(pfg, po, verts, 0.0, 0.4, [5.0, 5.0, 1.85])
    assert np.dot(F.sum(0), np.array([5.0, 5.0, 1.85]) - v[0]) < 0


def test_marker_update_in_uniform_flow(pfg, po):
    u0 = np.array([0.011, -0.0045, 0.0023])
    p = pfg.default_params(nx=16, ny=12, nz=10, tau=0.9)
    p.morse_u0 = 0.0
    p.marker_dt = 0.75
    x = np.array([3.3, 7.1, 12.9, 5.0])
    y = np.array([2.2, 9.7, 0.4, 6.0])
    z = np.array([4.4, 1.3, 8.8, 5.0])
    kind = np.array([pfg.MARKER_FIBER] * 3 + [pfg.MARKER_FIXED], np.int32)
    M = ModelArrays(x, y, z, kind=kind, body=-np.ones(4))
    o = po.Oracle(p, model=M.model)
    N = p.nx * p.ny * p.nz
    L = po.lattice(p)
    o.set_f(po.init_eq(L, None, *[np.full(N, c) for c in u0]))
    X0 = np.stack([o.marker(q).copy() for q in "xyz"], 1)
    n = 7
    o.step(n)
    X = np.stack([o.marker(q).copy() for q in "xyz"], 1)
    V = np.stack([o.marker(q).copy() for q in ("vx", "vy", "vz")], 1)
    o.close()
    dt = float(np.float32(0.75))
    assert np.allclose(V[:3], u0[None, :], rtol=0, atol=1e-14)
    assert np.allclose(X[:3], X0[:3] + n * dt * u0[None, :], rtol=0, atol=1e-12)
    assert np.all(X[3] == X0[3]) and np.all(V[3] == 0)


def _bond_rates(pfg, po, law, r, steps=50, n_side=20):
    """Formation and rupture rates per step of a receptor with one ligand at
    distance r: n_side^2 receptors (one per triangle, on its first vertex),
    each with its own ligand straight below and no other within reach,
    decided afresh at each of `steps` steps (bonds reset before each), so
    the fractions estimate p = 1 - e^{-k} over n_side^2 * steps draws."""
    p = pfg.default_params(nx=64, ny=64, nz=8)
    p.morse_u0 = 0.0
    p.k_on, p.k_off0 = 0.3, 0.2
    p.bond_k, p.bond_l0, p.bond_reach = 1.0, 0.3, 1.4
    p.bond_law, p.bell_xb_kT, p.bell_kT = law, 0.5, 0.5
    n = n_side * n_side
    g = np.arange(n)
    cx, cy = 1.0 + 3.0 * (g % n_side), 1.0 + 3.0 * (g // n_side)
    v = np.zeros((3 * n, 3))
    v[0::3] = np.stack([cx, cy, np.full(n, 4.0)], 1)
    v[1::3] = np.stack([cx + 0.8, cy, np.full(n, 4.0)], 1)
    v[2::3] = np.stack([cx, cy + 0.8, np.full(n, 4.0)], 1)
    tri = np.arange(3 * n, dtype=np.int32)
    S = np.full(n, 0.32)
    M = ModelArrays(v[:, 0], v[:, 1], v[:, 2], plt_elem_start=[0, 3 * n],
                    plt_tri_start=[0, n], tri=tri, tri_S0=S, plt_S0=[S.sum()], plt_V0=[-1.0],
                    rec_tri=g, rec_r1=np.zeros(n), rec_r2=np.zeros(n),
                    lig_x=cx, lig_y=cy, lig_z=np.full(n, 4.0 - r))
    o = po.Oracle(p, model=M.model)
    formed = ruptured = 0
    for s in range(steps):
        for bonded in (False, True):
            o.set_step(1000 + s)
            o.set_bonds(g.astype(np.int32) if bonded else np.full(n, -1, np.int32))
            for st in (0, 1, 2):
                o.stage(st)
            rb = o.bonds()[0]
            if bonded:
                ruptured += int((rb < 0).sum())
            else:
                formed += int((rb == g).sum())
                assert np.all((rb == g) | (rb < 0))  # only its own ligand is in reach
    o.close()
    w = abs(float(np.float32(4.0)) - float(np.float32(4.0 - r)) - p.bond_l0)
    k_on = -np.log1p(-formed / (n * steps))
    k_off = -np.log1p(-ruptured / (n * steps))
    return w, k_on, k_off, p


def test_reference_bond_law_detailed_balance(pfg, po):
    """The reference's force-dependent rates (sem_kernel.cu:1852-1863) obey
    detailed balance with the bond's spring energy E = k w^2 / 2, w = |l0 - r|:
    k_on / k_off = (Kf0 / Kr0) e^{-E / kT}, and the off rate follows Bell's
    law in the bond force f = k w: k_off = Kr0 e^{x_b f / kT}, for stretched
    and compressed bonds alike.  Under the Bell law option the on rate does
    not depend on the distance and a compressed bond feels no force.  Each
    rate is estimated from 20 000 draws (log-rate errors ~0.02)."""
    for r in (0.05, 0.7, 1.1):  # compressed (w = 0.25), stretched (w = 0.4, 0.8)
        w, kon, koff, p = _bond_rates(pfg, po, pfg.BOND_REFERENCE, r)
        E = 0.5 * p.bond_k * w * w
        lhs = np.log(kon / koff) - np.log(p.k_on / p.k_off0)
        assert abs(lhs + E / p.bell_kT) < 0.08, (r, lhs, -E / p.bell_kT)
        assert abs(np.log(koff / p.k_off0) - p.bell_xb_kT * p.bond_k * w) < 0.08, (r, koff)
    for r in (0.05, 1.1):
        w, kon, koff, p = _bond_rates(pfg, po, pfg.BOND_BELL, r)
        assert abs(np.log(kon / p.k_on)) < 0.08, (r, kon)
        f = p.bond_k * w if r > p.bond_l0 else 0.0
        assert abs(np.log(koff / p.k_off0) - p.bell_xb_kT * f) < 0.08, (r, koff)
