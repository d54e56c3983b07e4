"""Pins of the FP64 lattice-Boltzmann oracle against what the mathematics fixes.

Closed forms (equilibrium moments, Guo source moments, Poiseuille and Couette
profiles), exact invariants (mass, momentum gain N*F per step) and the
rest-state fixed point.  A wrong weight, direction, bounce-back partner, sign
or factor in the oracle fails one of these.
"""
import numpy as np
import pytest

Q = 19


def tables(po):
    L = po.load()
    c = np.array([[L.orc_c(d, a) for a in range(3)] for d in range(Q)])
    w = np.array([L.orc_w(d) for d in range(Q)])
    bb = np.array([L.orc_bb(d) for d in range(Q)])
    return c, w, bb


def params(pfg, nx, ny, nz, walls=(0, 0, 0), tau=0.8, g=(0, 0, 0), **kw):
    p = pfg.default_params(nx=nx, ny=ny, nz=nz, tau=tau)
    for a in range(3):
        p.wall[a] = walls[a]
        p.body_force[a] = g[a]
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_d3q19_tables(po):
    c, w, bb = tables(po)
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w @ c, 0)
    # isotropy: sum w c_a c_b = delta/3, 4th moment sum w c_a^2 c_b^2 = 1/9 (a!=b)
    assert np.allclose(np.einsum("d,da,db->ab", w, c, c), np.eye(3) / 3)
    assert np.isclose((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum(), 1 / 9)
    assert np.isclose((w * c[:, 0] ** 4).sum(), 1 / 3)
    assert np.array_equal(c[bb], -c)
    # the reference's order (LB_kernel.cu:43-46)
    assert list(c[:, 0]) == [0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0]


def test_equilibrium_moments(pfg, po):
    c, w, _ = tables(po)
    p = params(pfg, 4, 3, 2)
    L = po.lattice(p)
    rng = np.random.default_rng(1)
    N = 24
    rho = 1 + 0.1 * rng.standard_normal(N)
    u = 0.05 * rng.standard_normal((3, N))
    f = po.init_eq(L, rho, *u)
    assert np.allclose(f.sum(0), rho, atol=1e-14)
    assert np.allclose(c.T @ f, rho * u, atol=1e-14)
    Pi = np.einsum("da,db,dn->abn", c, c, f)
    expect = np.einsum("n,an,bn->abn", rho, u, u) + np.eye(3)[:, :, None] * rho / 3
    assert np.allclose(Pi, expect, atol=1e-14)


@pytest.mark.parametrize("walls", [(0, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_rest_state_fixed_point(pfg, po, walls):
    _, w, _ = tables(po)
    p = params(pfg, 6, 5, 4, walls=walls)
    L = po.lattice(p)
    f = np.repeat(w[:, None], 120, axis=1)
    f1 = po.lb_step(L, f)
    assert np.abs(f1 - f).max() < 1e-16


def test_mass_conserved_with_walls_and_solids(pfg, po):
    rng = np.random.default_rng(2)
    p = params(pfg, 6, 5, 7, walls=(0, 1, 1), use_flags=1)
    solid = (rng.random(6 * 5 * 7) < 0.15).astype(np.uint8)
    L = po.lattice(p, solid)
    _, w, _ = tables(po)
    f = w[:, None] * (1 + 0.05 * rng.standard_normal((Q, 210)))
    f[:, solid == 1] = 0  # solid cells hold nothing
    m0 = po.conservation(L, f)[0]
    for _ in range(5):
        f = po.lb_step(L, f)
    assert abs(po.conservation(L, f)[0] - m0) < 1e-12 * m0


def test_momentum_gain_is_N_times_F(pfg, po):
    rng = np.random.default_rng(3)
    g = (1e-4, -2e-4, 3e-4)
    p = params(pfg, 5, 4, 3, g=g, tau=0.7)
    L = po.lattice(p)
    _, w, _ = tables(po)
    f = w[:, None] * (1 + 0.02 * rng.standard_normal((Q, 60)))
    j0 = po.conservation(L, f)[1:]
    f = po.lb_step(L, f)
    j1 = po.conservation(L, f)[1:]
    g32 = np.array(g, dtype=np.float32).astype(np.float64)  # parameters are FP32
    assert np.allclose(j1 - j0, 60 * g32, rtol=0, atol=1e-15)
    # a local force field adds its own total as well
    Fx = np.zeros((3, 60))
    Fx[:, 7] = (1e-3, 2e-3, -1e-3)
    j2 = po.conservation(L, po.lb_step(L, f, Fu=Fx, Fs=Fx))[1:]
    assert np.allclose(j2 - j1, 60 * g32 + Fx[:, 7], rtol=0, atol=1e-15)


def test_guo_source_moments(pfg, po):
    """Collision of f = feq(rho, u) with force F: the post-collision state has
    moments rho, rho u + F/2 (sum c S = (1 - 1/2tau) F plus the shifted feq)."""
    c, w, _ = tables(po)
    tau = 0.9
    p = params(pfg, 1, 1, 1, tau=tau)
    L = po.lattice(p)
    rho, u = 1.03, np.array([0.02, -0.01, 0.015])
    F = np.array([1e-3, 2e-3, -3e-3])
    f = po.init_eq(L, [rho], *u[:, None])
    # moments of f give u_m = (rho u + F/2)/rho = u + F/(2 rho)
    f_in = f.copy()
    fo = po.lb_step(L, f_in, Fu=F[:, None], Fs=F[:, None])
    assert np.isclose(fo.sum(), rho, atol=1e-15)
    # post-collision momentum = j + F
    assert np.allclose(c.T @ fo[:, 0], rho * u + F, atol=1e-15)


def test_poiseuille_between_plates(pfg, po):
    H, g, tau = 12, 1e-5, 0.8
    nu = (tau - 0.5) / 3
    p = params(pfg, 2, 2, H, walls=(0, 0, 1), tau=tau, g=(g, 0, 0))
    L = po.lattice(p)
    f = po.init_eq(L)
    for _ in range(3000):
        f = po.lb_step(L, f)
    rho, ux, _, _ = po.fields(L, f)
    z = np.arange(H) + 0.5
    exact = g / (2 * nu) * z * (H - z)
    prof = ux.reshape(H, 2, 2)[:, 0, 0]
    assert np.abs(prof - exact).max() < 5e-3 * exact.max()


def test_couette_linear_profile(pfg, po):
    H, U, tau = 10, 0.02, 0.7
    p = params(pfg, 2, 2, H, walls=(0, 0, 1), tau=tau)
    p.wall_u_hi[2][0] = U
    L = po.lattice(p)
    f = po.init_eq(L)
    for _ in range(3000):
        f = po.lb_step(L, f)
    _, ux, uy, uz = po.fields(L, f)
    prof = ux.reshape(H, 2, 2)[:, 0, 0]
    exact = U * (np.arange(H) + 0.5) / H
    assert np.abs(prof - exact).max() < 1e-6
    assert np.abs(uy).max() < 1e-12 and np.abs(uz).max() < 1e-12


def test_reference_order_differs_only_through_collision_velocity(pfg, po):
    """Reference order collides with u from the previous force, F(n-1); the
    fused order with F(n).  Mass is identical; post-collision momentum differs
    by exactly (F(n) - F(n-1))/(2 tau) per cell."""
    rng = np.random.default_rng(4)
    tau = 0.75
    p = params(pfg, 4, 4, 4, tau=tau)
    L = po.lattice(p)
    c, w, _ = tables(po)
    f = w[:, None] * (1 + 0.01 * rng.standard_normal((Q, 64)))
    F = 1e-4 * rng.standard_normal((3, 64))
    Fprev = 0.5 * F
    a = po.lb_step(L, f, Fu=F, Fs=F)
    b = po.lb_step(L, f, Fu=Fprev, Fs=F)
    assert np.allclose(a.sum(0), b.sum(0), atol=1e-15)
    assert np.allclose(c.T @ a - c.T @ b, (F - Fprev) / (2 * tau), atol=1e-15)
    assert np.array_equal(po.lb_step(L, f, Fu=F, Fs=F), a)


def test_results_do_not_depend_on_the_thread_count(pfg, po):
    """The oracle's lattice loops and interpolation run under OpenMP; every
    cell and marker is computed on its own, so a coupled run gives the same
    bits on one thread and on several."""
    sc = pfg.Scene(pfg.SCENE_CHANNEL)
    before = po.set_threads(0)
    out = []
    try:
        for threads in (1, 4):
            po.set_threads(threads)
            o = po.Oracle(sc.params, scene=sc)
            o.step(3)
            out.append((o.f().copy(), o.markers()))
            o.close()
    finally:
        po.set_threads(before)
    assert np.array_equal(out[0][0], out[1][0])
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(out[0][1][k], out[1][1][k])
