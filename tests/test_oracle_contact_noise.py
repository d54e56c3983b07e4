"""Pins of the oracle's platelet-platelet contact (sem_kernel.cu:436-509) and
of the thermal noise of the fibrin kick (sem_kernel.cu.orig:251-266):

- contact: the force acts on exactly the closest pair of elements of two
  platelets (found here by brute force), with the closed-form magnitude
  a kappa e^{-kappa h}/(1 - e^{-kappa h}), opposite on the two, and vanishes
  beyond the cutoff;
- noise: Box-Muller draws have mean 0, variance 1 and kurtosis 3, and the
  velocity of free kicked nodes relaxes to the equipartition variance T
  (unit mass) with the exact Ornstein-Uhlenbeck step."""
import ctypes as C

import numpy as np
import pytest

from helpers import ModelArrays


def two_platelets(pfg, gap, n=200):
    L = pfg.lib()
    P, T = [], []
    for k, cx in enumerate((20.0, 20.0 + 8.0 + gap)):
        x, y, z = (np.empty(n, np.float32) for _ in range(3))
        t = np.empty(3 * (2 * n - 4), np.int32)
        nt = C.c_int32()
        L.pfg_geom_ellipsoid(n, 4.0, 3.0, 2.0, cx, 20.0, 20.0, 0.05, 7 + k,
                             *[q.ctypes.data_as(C.POINTER(C.c_float)) for q in (x, y, z)],
                             t.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nt))
        P.append(np.stack([x, y, z], 1))
        T.append(t.reshape(-1, 3) + k * n)
    return np.concatenate(P), np.concatenate(T), n


def contact_oracle(pfg, po, P, tri, n, a=0.01, kappa=10.0, hc=1.0):
    p = pfg.default_params(nx=64, ny=40, nz=40)
    p.morse_u0 = 0.0
    p.contact_a, p.contact_kappa, p.contact_hc = a, kappa, hc
    ma = ModelArrays(P[:, 0], P[:, 1], P[:, 2], body=np.repeat([0, 1], n),
                     plt_elem_start=[0, n, 2 * n], plt_tri_start=[0, len(tri) // 2, len(tri)],
                     tri=tri.ravel(), tri_S0=np.ones(len(tri)), plt_S0=[1.0, 1.0],
                     plt_V0=[1.0, 1.0])
    o = po.Oracle(p, model=ma.model)
    o._keep = ma
    for st in (pfg.STAGE_CELLS, pfg.STAGE_MEMBRANE, pfg.STAGE_BONDS, pfg.STAGE_FORCES):
        o.stage(st)
    return np.stack([o.marker(k) for k in ("fx", "fy", "fz")], 1)


def test_contact_closest_pair_closed_form(pfg, po):
    P, tri, n = two_platelets(pfg, 0.4)
    a, kappa, hc = 0.01, 10.0, 1.0
    F = contact_oracle(pfg, po, P, tri, n, a, kappa, hc)
    Pd = P.astype(np.float64)
    D = np.linalg.norm(Pd[:n, None, :] - Pd[None, n:, :], axis=2)
    i, j = np.unravel_index(np.argmin(D), D.shape)
    j += n
    d = D.min()
    assert d < hc
    nz = np.flatnonzero(np.abs(F).sum(1))
    assert sorted(nz.tolist()) == sorted([i, j])
    e = np.exp(-kappa * max(d, 0.01 * hc))
    mag = a * kappa * e / (1 - e)
    u = (Pd[i] - Pd[j]) / d
    assert np.allclose(F[i], mag * u, rtol=1e-12)
    assert np.allclose(F[j], -mag * u, rtol=1e-12)


def test_contact_beyond_cutoff(pfg, po):
    P, tri, n = two_platelets(pfg, 1.6)
    F = contact_oracle(pfg, po, P, tri, n, hc=1.0)
    assert np.abs(F).max() == 0


def test_box_muller_moments(pfg, po):
    """Free kicked nodes after one step from rest: v = sigma z per component."""
    z = kick_velocities(pfg, po, n=20000, steps=1, gamma=0.5, T=2.0, dt=1.0)
    sig = np.sqrt(2.0 * (1 - np.exp(-2 * 0.5 * 1.0)))
    zz = (z / sig).ravel()
    m = zz.size
    assert abs(zz.mean()) < 4 / np.sqrt(m)
    assert abs(zz.var() - 1.0) < 4 * np.sqrt(2.0 / m)
    kurt = np.mean(zz ** 4) / zz.var() ** 2
    assert abs(kurt - 3.0) < 0.1


def test_kick_noise_equipartition(pfg, po):
    v = kick_velocities(pfg, po, n=3000, steps=200, gamma=0.05, T=0.5, dt=1.0)
    var = v.var()
    assert abs(var / 0.5 - 1.0) < 0.04
    zero = kick_velocities(pfg, po, n=50, steps=5, gamma=0.05, T=0.0, dt=1.0)
    assert np.abs(zero).max() == 0


def kick_velocities(pfg, po, n, steps, gamma, T, dt):
    p = pfg.default_params(nx=16, ny=16, nz=16)
    p.fiber_move, p.fiber_gamma, p.kick_temp, p.marker_dt = pfg.FIBER_KICK, gamma, T, dt
    rng = np.random.default_rng(0)
    X = rng.uniform(2, 14, (n, 3))
    ma = ModelArrays(X[:, 0], X[:, 1], X[:, 2], kind=np.full(n, pfg.MARKER_FIBER),
                     body=np.full(n, -1))
    o = po.Oracle(p, model=ma.model)
    o._keep = ma
    for _ in range(steps):
        o.stage(pfg.STAGE_UPDATE)
    return np.stack([o.marker(k) for k in ("vx", "vy", "vz")], 1)
