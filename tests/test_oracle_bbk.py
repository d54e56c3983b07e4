"""Pins of the oracle's Langevin-leapfrog (BBK) element update
(verlet_1 / verlet_3, sem_kernel.cu.orig:224-568) against what the
mathematics fixes:

* free decay: with no force and no noise each step multiplies the velocity
  by e^{-g dt} exactly and the drift sums the geometric series;
* without friction it is the leapfrog: two elements on a spring follow the
  discrete harmonic oscillator r_{n+1} - 2 r_n + r_{n-1} = -(w0 dt)^2 r_n
  exactly (closed form with cos(w' dt) = 1 - (w0 dt)^2 / 2);
* with friction and noise free elements relax to the equipartition
  variance T/m per velocity component (each half step is the exact
  Ornstein-Uhlenbeck step, which keeps that variance);
* the drift weight: with w = 1 the elements move with the fluid only;
* energy: without friction or noise a free Morse cluster keeps its energy to
  O(dt^2) (halving dt cuts the error fourfold) with no drift, and its
  momentum exactly.
"""
import numpy as np

from helpers import ModelArrays


def _params(pfg, **kw):
    p = pfg.default_params(nx=24, ny=24, nz=24)
    p.morse_u0 = 0.0
    p.ib_force_scale = 0.0  # the elements do not drive the fluid
    p.sem_integrator = pfg.SEM_BBK
    p.sem_fluid_weight = 0.0
    p.sem_temp = 0.0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _platelet_model(x, y, z, edges=None, l0=None):
    n = len(x)
    e = None if edges is None else np.asarray(edges, np.int32).ravel()
    return ModelArrays(x, y, z, plt_elem_start=[0, n], plt_tri_start=[0, 0], plt_S0=[1.0],
                       plt_V0=[-1.0], edge=e, edge_l0=l0)


def _run(po, p, M, steps, v0=None):
    o = po.Oracle(p, model=M.model)
    if v0 is not None:
        for a, k in enumerate(("vx", "vy", "vz")):
            o.marker(k)[:] = v0[:, a]
    X0 = np.stack([o.marker(k).copy() for k in "xyz"], 1)
    traj = []
    for _ in range(steps):
        o.step(1)
        traj.append((np.stack([o.marker(k).copy() for k in "xyz"], 1),
                     np.stack([o.marker(k).copy() for k in ("vx", "vy", "vz")], 1)))
    o.close()
    return X0, traj


def test_free_decay_closed_form(pfg, po):
    g, dt, n = 0.3, 0.7, 12
    p = _params(pfg, sem_gamma=g, sem_mass=2.5, marker_dt=dt)
    x = np.array([5.0, 9.0, 14.0])
    M = _platelet_model(x, x + 1.0, x + 2.0)
    v0 = np.array([[0.02, -0.01, 0.005], [0.0, 0.03, -0.02], [-0.015, 0.0, 0.01]])
    X0, traj = _run(po, p, M, n, v0)
    dtf = float(np.float32(dt))
    decay = np.exp(-float(np.float32(g)) * dtf)
    for k, (X, V) in enumerate(traj, start=1):
        assert np.allclose(V, v0 * decay ** k, rtol=1e-12, atol=1e-18)
        series = sum(decay ** j for j in range(1, k + 1))
        assert np.allclose(X, X0 + dtf * series * v0, rtol=0, atol=1e-12)


def test_frictionless_leapfrog_is_the_discrete_oscillator(pfg, po):
    k_s, m, dt, L = 0.8, 1.7, 0.5, 0.9
    p = _params(pfg, sem_gamma=0.0, sem_mass=m, marker_dt=dt, k_spring=k_s)
    stretch = 0.2
    x = np.array([10.0, 10.0 + L + stretch])
    M = _platelet_model(x, [12.0, 12.0], [12.0, 12.0], edges=[[0, 1]], l0=[L])
    X0, traj = _run(po, p, M, 60)
    s = np.array([X0[1, 0] - X0[0, 0]] + [X[1, 0] - X[0, 0] for X, _ in traj])
    # relative coordinate: reduced mass m/2, so w0^2 = 2 k / m; the loaded
    # velocity is the half-step velocity v_{-1/2} = 0
    kf, mf, dtf = (float(np.float32(v)) for v in (k_s, m, dt))
    Lf = float(np.float32(L))
    w2 = 2.0 * kf / mf * dtf * dtf
    r = s - Lf
    # the recursion itself, exactly (up to rounding)
    assert np.allclose(r[2:] - 2 * r[1:-1] + r[:-2], -w2 * r[1:-1], rtol=0, atol=1e-12)
    # and its closed form: r_n = r_0 cos(n t) + B sin(n t) with cos t = 1 - w2/2
    t = np.arccos(1.0 - 0.5 * w2)
    B = (r[1] - r[0] * np.cos(t)) / np.sin(t)
    n = np.arange(r.size)
    assert np.allclose(r, r[0] * np.cos(n * t) + B * np.sin(n * t), rtol=0, atol=1e-10)
    # first step from rest: r_1 = r_0 (1 - w2)  (one full kick of dt / m
    # from the two half kicks, then the drift)
    assert abs(r[1] - r[0] * (1.0 - w2)) < 1e-12
    # momentum: the two elements move symmetrically
    mid = np.array([0.5 * (X[0, 0] + X[1, 0]) for X, _ in traj])
    assert np.allclose(mid, 0.5 * (X0[0, 0] + X0[1, 0]), rtol=0, atol=1e-12)


def test_equipartition(pfg, po):
    g, T, m = 0.8, 0.02, 2.0
    p = _params(pfg, sem_gamma=g, sem_temp=T, sem_mass=m, marker_dt=1.0)
    rng = np.random.default_rng(3)
    n = 3000
    P = rng.random((n, 3)) * 20 + 2
    M = _platelet_model(P[:, 0], P[:, 1], P[:, 2])
    o = po.Oracle(p, model=M.model)
    o.step(60)  # 1 - e^{-2 g t}: relaxed
    V = np.stack([o.marker(k).copy() for k in ("vx", "vy", "vz")], 1)
    o.step(1)
    V2 = np.stack([o.marker(k).copy() for k in ("vx", "vy", "vz")], 1)
    o.close()
    var = V.var()
    Tm = float(np.float32(T)) / float(np.float32(m))
    assert abs(var / Tm - 1.0) < 0.05
    assert abs(V.mean()) < 5 * np.sqrt(Tm / V.size)
    # one step later: correlation e^{-g dt} (two half steps of e^{-g dt/2})
    corr = (V * V2).mean() / np.sqrt((V * V).mean() * (V2 * V2).mean())
    assert abs(corr - np.exp(-float(np.float32(g)))) < 0.03


def test_drift_weight_one_follows_the_fluid(pfg, po):
    """w = 1: the drift is the fluid velocity alone, whatever the element's
    own velocity; in a uniform flow (an exact fixed point of the lattice)
    the elements move by u dt per step, as with advection."""
    u0 = np.array([0.013, -0.004, 0.006])
    p = _params(pfg, sem_gamma=0.4, sem_temp=0.01, sem_fluid_weight=1.0, marker_dt=0.5)
    x = np.array([4.2, 11.9, 17.3])
    M = _platelet_model(x, x[::-1], x + 0.5)
    o = po.Oracle(p, model=M.model)
    N = p.nx * p.ny * p.nz
    o.set_f(po.init_eq(po.lattice(p), None, *[np.full(N, c) for c in u0]))
    X0 = np.stack([o.marker(k).copy() for k in "xyz"], 1)
    o.step(9)
    X = np.stack([o.marker(k).copy() for k in "xyz"], 1)
    V = np.stack([o.marker(k).copy() for k in ("vx", "vy", "vz")], 1)
    o.close()
    assert np.allclose(X, X0 + 9 * 0.5 * u0[None, :], rtol=0, atol=1e-12)
    assert np.abs(V).max() > 0  # the elements' own (thermal) velocity is kept apart


def _morse_cluster_energy(pfg, po, dt, steps):
    m = 1.3
    # a cutoff no pair reaches (a smooth potential), so the leapfrog's energy
    # error is its own O(dt^2)
    p = _params(pfg, sem_gamma=0.0, sem_mass=m, marker_dt=dt, morse_u0=1.0, morse_rho=2.0,
                morse_req=0.8, morse_cutoff=4.0, morse_taper=0.0, bin_cells=4)
    rng = np.random.default_rng(7)
    n = 8  # a 2 x 2 x 2 cube of side 0.85 (near r_eq), shaken by 0.08
    g = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], float)
    X = 11.0 + 0.85 * g + 0.08 * rng.standard_normal((n, 3))
    M = _platelet_model(X[:, 0], X[:, 1], X[:, 2])
    o = po.Oracle(p, model=M.model)
    rho, re, rc = 2.0, float(np.float32(0.8)), float(np.float32(4.0))

    def U(Y):
        d = np.linalg.norm(Y[:, None, :] - Y[None, :, :], axis=2)[np.triu_indices(n, 1)]
        assert d.max() < rc
        return float((np.exp(-2 * rho * (d - re)) - 2 * np.exp(-rho * (d - re))).sum())

    mf = float(np.float32(m))
    E, KE, P = [], [], []
    for _ in range(steps):
        Xn = np.stack([o.marker(k).copy() for k in "xyz"], 1)
        v0 = np.stack([o.marker(k).copy() for k in ("vx", "vy", "vz")], 1)
        o.step(1)
        v1 = np.stack([o.marker(k).copy() for k in ("vx", "vy", "vz")], 1)
        v = 0.5 * (v0 + v1)
        KE.append(0.5 * mf * (v * v).sum())
        E.append(KE[-1] + U(Xn))
        P.append(np.abs(v1.sum(0)).max())
    o.close()
    E = np.array(E)
    return np.abs(E - E[0]).max(), abs(E[-steps // 6:].mean() - E[:steps // 6].mean()), max(KE), max(P)


def test_free_morse_cluster_conserves_energy(pfg, po):
    """Without friction or noise the element update is the leapfrog, a
    symplectic integrator: a free cluster of elements bound by the Morse
    potential keeps its energy E = sum m v^2 / 2 + U (v at whole steps the
    mean of the two half-step velocities) within an error that falls as
    dt^2, with no drift, and its total momentum exactly."""
    dev1, drift1, ke, p1 = _morse_cluster_energy(pfg, po, 0.04, 2000)
    dev2, drift2, _, p2 = _morse_cluster_energy(pfg, po, 0.02, 4000)
    print(f"free Morse cluster: kinetic peak {ke:.3e}; max |E - E0| {dev1:.2e} (dt 0.04), "
          f"{dev2:.2e} (dt 0.02); drift {drift1:.2e}, {drift2:.2e}")
    assert ke > 1.0  # the cluster moves: energy is exchanged
    assert dev1 < 2e-2 * ke
    assert 3.0 < dev1 / dev2 < 5.0  # second order
    # the mean energy of the last sixth against the first: no secular drift
    # beyond the bounded O(dt^2) oscillation
    assert drift2 < 0.1 * dev1 and drift1 > drift2
    assert max(p1, p2) < 1e-12  # no net force on a free cluster
