"""Pins of the immersed-boundary oracle: kernel identities, the reference's
own node-loops-over-markers form, conservation of spread force and
spread/interpolation adjointness."""
import numpy as np


def test_cosine_kernel_identities(po):
    rng = np.random.default_rng(0)
    for x in rng.random(50) * 10:
        j = np.arange(int(np.floor(x)) - 3, int(np.floor(x)) + 4)
        phi = np.array([po.phi(x - k) for k in j])
        assert abs(phi.sum() - 1.0) < 1e-14          # partition of unity
        assert abs((phi ** 2).sum() - 3.0 / 8.0) < 1e-14
    assert po.phi(2.0) == 0.0 and po.phi(-2.0) == 0.0
    assert abs(po.phi(0.0) - 0.5) < 1e-16
    # the first moment is not zero for this kernel (do not assume it)
    j = np.arange(-3, 5)
    assert abs(sum((0.25 - k) * po.phi(0.25 - k) for k in j)) > 1e-3


def _lattice(pfg, po, n=(8, 7, 6), walls=(0, 0, 0)):
    p = pfg.default_params(nx=n[0], ny=n[1], nz=n[2])
    for a in range(3):
        p.wall[a] = walls[a]
    return p, po.lattice(p)


def test_spread_matches_reference_scan_form(pfg, po):
    rng = np.random.default_rng(1)
    for walls in [(0, 0, 0), (0, 1, 1)]:
        p, L = _lattice(pfg, po, walls=walls)
        n = 9
        x = rng.random(n) * 8 - 0.5
        y = rng.random(n) * 7 - 0.5
        z = rng.random(n) * 5.6 - 0.3
        f = rng.standard_normal((3, n))
        a = po.ib_spread(L, x, y, z, *f)
        b = po.ib_spread(L, x, y, z, *f, scan=True)
        assert np.abs(a - b).max() < 1e-14


def test_spread_conserves_total_force_periodic(pfg, po):
    rng = np.random.default_rng(2)
    p, L = _lattice(pfg, po)
    n = 20
    x, y, z = rng.random(n) * 30 - 10, rng.random(n) * 7, rng.random(n) * 6
    f = rng.standard_normal((3, n))
    F = po.ib_spread(L, x, y, z, *f, scale=0.5)
    assert np.allclose(F.sum(1), 0.5 * f.sum(1), atol=1e-13)


def test_spread_interp_adjoint(pfg, po):
    rng = np.random.default_rng(3)
    p, L = _lattice(pfg, po, walls=(0, 1, 0))
    N = 8 * 7 * 6
    n = 11
    x, y, z = rng.random(n) * 8, rng.random(n) * 7 - 0.5, rng.random(n) * 6
    f = rng.standard_normal((3, n))
    u = rng.standard_normal((3, N))
    F = po.ib_spread(L, x, y, z, *f)
    v = np.array(po.ib_interp(L, x, y, z, u))
    assert abs((F * u).sum() - (f * v).sum()) < 1e-12


def test_interp_constant_field_periodic(pfg, po):
    p, L = _lattice(pfg, po)
    N = 8 * 7 * 6
    u = np.array([0.3, -0.2, 0.1])[:, None] * np.ones((3, N))
    rng = np.random.default_rng(4)
    x, y, z = rng.random(5) * 8, rng.random(5) * 7, rng.random(5) * 6
    v = np.array(po.ib_interp(L, x, y, z, u))
    assert np.allclose(v, u[:, :5], atol=1e-14)
