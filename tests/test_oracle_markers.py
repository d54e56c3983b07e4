"""Pins of the marker-side oracle: cell lists against brute force, membrane
area/volume against closed forms, forces against finite differences of their
energies, the fibrin forces, kick and graph against the reference's own
network.cpp (compiled into oracle/_ref), Philox against published vectors."""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import ModelArrays, area_volume_np, cube_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libnetref.so")


def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32_10."""
    src = os.path.join(ROOT, "platelet_fiber_gpu_amd", "csrc", "pfg_rng.h")
    import subprocess
    import tempfile
    prog = r'''
#include <stdio.h>
#include "%s"
int main(void){
 uint32_t c[3][4]={{0,0,0,0},{0xffffffffu,0xffffffffu,0xffffffffu,0xffffffffu},
   {0x243f6a88u,0x85a308d3u,0x13198a2eu,0x03707344u}};
 uint32_t k[3][2]={{0,0},{0xffffffffu,0xffffffffu},{0xa4093822u,0x299f31d0u}};
 for(int i=0;i<3;++i){pfg_u32x4 r=pfg_philox4x32_10(c[i],k[i]);
  printf("%%08x %%08x %%08x %%08x\n",r.v[0],r.v[1],r.v[2],r.v[3]);}
 return 0;}
''' % src
    with tempfile.TemporaryDirectory() as d:
        cfile = os.path.join(d, "kat.c")
        open(cfile, "w").write(prog)
        exe = os.path.join(d, "kat")
        subprocess.run(["gcc", "-std=c99", "-O2", cfile, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "6627e8d5 e169c58d bc57ac4c 9b00dbd8"
    assert out[1] == "408f276d 41c83b0e a20bc7c6 6d5451fd"
    assert out[2] == "d16cfe09 94fdcceb 5001e420 24126ea1"


def test_cell_list_brute_force(po):
    rng = np.random.default_rng(0)
    dims = (20, 13, 11)
    periodic = (1, 1, 0)
    cells = 3
    nb = [max(1, d // cells) for d in dims]
    n = 400
    x = (rng.random(n) * 60 - 20).astype(np.float32)
    y = (rng.random(n) * 13).astype(np.float32)
    z = (rng.random(n) * 12 - 0.5).astype(np.float32)
    start, items = po.cell_list(x, y, z, dims, periodic, nb)
    assert start[0] == 0 and start[-1] == n and np.all(np.diff(start) >= 0)
    assert sorted(items) == list(range(n))
    bins = np.empty(n, int)
    for b in range(len(start) - 1):
        seg = items[start[b]:start[b + 1]]
        assert np.all(np.diff(seg) > 0)  # ascending marker order
        bins[seg] = b
    bxyz = np.stack([bins % nb[0], (bins // nb[0]) % nb[1], bins // (nb[0] * nb[1])], 1)
    # every pair closer than the bin width sits in neighbouring bins
    P = np.stack([x, y, z], 1).astype(np.float64)
    for i in range(n):
        d = P - P[i]
        for a in range(3):
            if periodic[a]:
                d[:, a] -= dims[a] * np.round(d[:, a] / dims[a])
        close = np.where((d ** 2).sum(1) < cells ** 2)[0]
        for j in close:
            for a in range(3):
                db = abs(int(bxyz[i, a]) - int(bxyz[j, a]))
                if periodic[a]:
                    db = min(db, nb[a] - db)
                assert db <= 1


def test_membrane_closed_forms(pfg, po):
    v, tri = cube_mesh()
    S, Sall, V = area_volume_np(v, tri)
    assert np.isclose(Sall, 6.0) and np.isclose(V, -1.0)  # reference sign
    # the library's host routine and the oracle agree with the closed form
    tS = np.zeros(12)
    Sa, Va = C.c_double(), C.c_double()
    x, y, z = [np.ascontiguousarray(v[:, k], np.float32) for k in range(3)]
    t = np.ascontiguousarray(tri.ravel(), np.int32)
    pfg.lib().pfg_geom_area_volume(12, t.ctypes.data_as(C.POINTER(C.c_int32)),
                                   *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in (x, y, z)],
                                   tS.ctypes.data_as(C.POINTER(C.c_double)), C.byref(Sa),
                                   C.byref(Va))
    assert np.isclose(Sa.value, 6.0) and np.isclose(Va.value, -1.0)
    assert np.allclose(tS, 0.5)
    ma = ModelArrays(v[:, 0] + 3, v[:, 1] + 3, v[:, 2] + 3, plt_elem_start=[0, 8],
                     plt_tri_start=[0, 12], tri=tri.ravel(), tri_S0=np.full(12, 0.5),
                     plt_S0=[6.0], plt_V0=[-1.0])
    p = pfg.default_params(nx=8, ny=8, nz=8)
    o = po.Oracle(p, model=ma.model)
    o.stage(pfg.STAGE_MEMBRANE)
    S_o, V_o, tS_o = o.membrane()
    assert np.isclose(S_o[0], 6.0, atol=1e-12) and np.isclose(V_o[0], -1.0, atol=1e-12)


def test_ellipsoid_mesh_volume(pfg):
    s = pfg.Scene(pfg.SCENE_CHANNEL, elements=1024)
    a, b, c = 8.0, 8.0, 3.0  # (4, 4, 1.5) * sqrt(1024/256)
    vol = 4.0 / 3.0 * np.pi * a * b * c
    assert -s.plt_V0[0] < vol and -s.plt_V0[0] > 0.97 * vol
    tri = s.tri.reshape(-1, 3)
    assert tri.shape[0] == 2 * 1024 - 4
    # closed 2-manifold: every directed edge appears once, with its twin
    e = set()
    for t in tri:
        for k in range(3):
            e.add((t[k], t[(k + 1) % 3]))
    assert len(e) == 3 * len(tri) and all((j, i) in e for i, j in e)


def _energy(x, tri, S0, S0all, V0, ks, kv):
    S, Sall, V = area_volume_np(x, tri)
    # V and V0 carry the reference's sign (negative for outward-wound meshes),
    # so the volume energy is taken with |V0|: the force restores V towards V0
    return (ks * ((S - S0) ** 2 / (2 * S0)).sum() + ks * (Sall - S0all) ** 2 / (2 * S0all)
            + kv * (V - V0) ** 2 / (2 * abs(V0)))


def test_area_volume_forces_are_energy_gradients(pfg, po):
    """The reference's constraint forces (sem_kernel.cu:1640-1749) equal -grad E
    with E = Ks[sum (S_t-S0_t)^2/2S0_t + (S-S0)^2/2S0] + Kv (V-V0)^2/2|V0|."""
    s = pfg.Scene(pfg.SCENE_CHANNEL, elements=64)
    rng = np.random.default_rng(1)
    X = np.stack([s.x, s.y, s.z], 1).astype(np.float64)
    X += 0.05 * rng.standard_normal(X.shape)
    X = X.astype(np.float32).astype(np.float64)
    tri = s.tri.reshape(-1, 3)
    p = pfg.default_params(nx=64, ny=32, nz=32, ks=0.7, kv=0.3)
    ma = ModelArrays(X[:, 0], X[:, 1], X[:, 2], plt_elem_start=s.plt_elem_start,
                     plt_tri_start=s.plt_tri_start, tri=s.tri, tri_S0=s.tri_S0,
                     plt_S0=s.plt_S0, plt_V0=s.plt_V0)
    o = po.Oracle(p, model=ma.model)
    o.stage(pfg.STAGE_MEMBRANE)
    o.stage(pfg.STAGE_FORCES)
    F = np.stack([o.marker(k) for k in ("fx", "fy", "fz")], 1)
    h = 1e-6
    ks, kv = np.float32(0.7), np.float32(0.3)
    G = np.zeros_like(X)
    for i in range(X.shape[0]):
        for a in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[i, a] += h
            Xm[i, a] -= h
            G[i, a] = (_energy(Xp, tri, s.tri_S0, s.plt_S0[0], s.plt_V0[0], ks, kv) -
                       _energy(Xm, tri, s.tri_S0, s.plt_S0[0], s.plt_V0[0], ks, kv)) / (2 * h)
    assert np.abs(F + G).max() < 1e-6 * max(1.0, np.abs(G).max())


def test_morse_force_is_energy_gradient(pfg, po):
    rng = np.random.default_rng(2)
    n = 12
    X = (rng.random((n, 3)) * 2.0 + 5).astype(np.float32).astype(np.float64)
    p = pfg.default_params(nx=16, ny=16, nz=16, morse_u0=1.0, morse_rho=2.0, morse_req=0.8,
                           morse_cutoff=1.5, morse_taper=0.4)
    ma = ModelArrays(X[:, 0], X[:, 1], X[:, 2], plt_elem_start=[0, n], plt_tri_start=[0, 0],
                     plt_S0=[1.0], plt_V0=[1.0])
    o = po.Oracle(p, model=ma.model)
    o.stage(pfg.STAGE_FORCES)
    F = np.stack([o.marker(k) for k in ("fx", "fy", "fz")], 1)
    # energy with a cut pair potential (pairs counted once)
    rho, re, U0, rc, tp = 2.0, np.float32(0.8), 1.0, np.float32(1.5), np.float32(0.4)

    def E(Y):
        e = 0.0
        for i in range(n):
            for j in range(i + 1, n):
                r = np.linalg.norm(Y[i] - Y[j])
                if r < rc:
                    t = max(0.0, (r - (rc - tp)) / tp)
                    S = 1 - t * t * (3 - 2 * t)
                    e += S * U0 * (np.exp(-2 * rho * (r - re)) - 2 * np.exp(-rho * (r - re)))
        return e

    h = 1e-6
    for i in range(n):
        for a in range(3):
            Yp, Ym = X.copy(), X.copy()
            Yp[i, a] += h
            Ym[i, a] -= h
            g = (E(Yp) - E(Ym)) / (2 * h)
            assert abs(F[i, a] + g) < 1e-5
    # zero force at r_eq for an isolated pair (0.75: exact in binary)
    p.morse_req = 0.75
    ma2 = ModelArrays([5.0, 5.75], [5.0, 5.0], [5.0, 5.0], plt_elem_start=[0, 2],
                      plt_tri_start=[0, 0], plt_S0=[1.0], plt_V0=[1.0])
    o2 = po.Oracle(p, model=ma2.model)
    o2.stage(pfg.STAGE_FORCES)
    assert abs(o2.marker("fx")[0]) < 1e-14


def _ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = C.CDLL(REF_LIB)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    L.ref_spring_force.argtypes = [C.c_int, dp, dp, C.c_int, ip, C.c_double, C.c_double, dp]
    L.ref_solverkick.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double]
    L.ref_definelinks.argtypes = [C.c_int, dp, C.c_double, ip, ip, C.c_int]
    L.ref_definelinks.restype = C.c_int
    return L


def _fibrin(rng, n=30):
    rest = rng.random((n, 3)) * 6
    links = [(i, j) for i in range(n) for j in range(i + 1, n)
             if np.linalg.norm(rest[i] - rest[j]) < 2.5]
    links = np.array(links, dtype=np.int32)
    cur = rest + 0.15 * rng.standard_normal(rest.shape)
    return rest, cur, links


def test_wlc_against_reference_network(pfg, po):
    L = _ref()
    rng = np.random.default_rng(3)
    rest, cur, links = _fibrin(rng)
    rest = rest.astype(np.float32).astype(np.float64)
    cur = cur.astype(np.float32).astype(np.float64)
    n = len(rest)
    l0 = np.linalg.norm(rest[links[:, 1]] - rest[links[:, 0]], axis=1)
    kT = 1.3806503e-8 * 10  # kB * mTemp of network.cpp:388-391
    Fref = np.zeros(3 * n)
    dp = C.POINTER(C.c_double)
    L.ref_spring_force(n, np.ascontiguousarray(rest.ravel()).ctypes.data_as(dp),
                       np.ascontiguousarray(cur.ravel()).ctypes.data_as(dp), len(links),
                       np.ascontiguousarray(links.ravel()).ctypes.data_as(C.POINTER(C.c_int)),
                       1.0, 1.0, Fref.ctypes.data_as(dp))
    Fref = Fref.reshape(n, 3)
    p = pfg.default_params(nx=16, ny=16, nz=16, fiber_k=kT, fiber_ref_sign=1,
                           fiber_xmax=10.0)
    ma = ModelArrays(cur[:, 0], cur[:, 1], cur[:, 2], kind=np.ones(n), body=-np.ones(n),
                     link=links.ravel(), link_l0=l0)
    # rest lengths pass through FP32 in the model; give the reference the same
    l0_32 = ma.a["link_l0"].astype(np.float64)
    assert np.allclose(l0_32, l0, rtol=1e-6)
    o = po.Oracle(p, model=ma.model)
    o.stage(pfg.STAGE_FORCES)
    F = np.stack([o.marker(k) for k in ("fx", "fy", "fz")], 1)
    scale = np.abs(Fref).max()
    assert scale > 0
    assert np.abs(F - Fref).max() < 1e-5 * scale
    assert np.allclose(F.sum(0), 0, atol=1e-12 * scale * n)  # Newton's third law
    # restoring sign (default) is the exact negative
    p.fiber_ref_sign = 0
    o2 = po.Oracle(p, model=ma.model)
    o2.stage(pfg.STAGE_FORCES)
    F2 = np.stack([o2.marker(k) for k in ("fx", "fy", "fz")], 1)
    assert np.allclose(F2, -F, atol=1e-20)


def test_linear_link_closed_form(pfg, po):
    p = pfg.default_params(nx=8, ny=8, nz=8, fiber_k=2.0, fiber_law=pfg.FIBER_LINEAR)
    ma = ModelArrays([1.0, 3.5], [1.0, 1.0], [1.0, 1.0], kind=[1, 1], body=[-1, -1],
                     link=[0, 1], link_l0=[2.0])
    o = po.Oracle(p, model=ma.model)
    o.stage(pfg.STAGE_FORCES)
    # tension Y A (l - l0)/l0 = 2 * 0.5 / 2 = 0.5, pulling the ends together
    assert np.isclose(o.marker("fx")[0], 0.5) and np.isclose(o.marker("fx")[1], -0.5)


def test_kick_against_reference(pfg, po):
    L = _ref()
    rng = np.random.default_rng(4)
    v = rng.standard_normal(6)
    f = rng.standard_normal(6)
    dt, gam = 1.0, np.float32(0.3)
    vr = v.copy()
    dp = C.POINTER(C.c_double)
    L.ref_solverkick(6, vr.ctypes.data_as(dp), f.ctypes.data_as(dp), dt, float(gam))
    p = pfg.default_params(nx=8, ny=8, nz=8, fiber_move=pfg.FIBER_KICK, fiber_gamma=gam)
    ma = ModelArrays([1.0, 2.0], [1.0, 1.0], [1.0, 1.0], kind=[1, 1], body=[-1, -1])
    o = po.Oracle(p, model=ma.model)
    o.marker("vx")[:] = v[0::3]
    o.marker("vy")[:] = v[1::3]
    o.marker("vz")[:] = v[2::3]
    o.marker("fx")[:] = f[0::3]
    o.marker("fy")[:] = f[1::3]
    o.marker("fz")[:] = f[2::3]
    x0 = o.marker("x").copy()
    o.stage(pfg.STAGE_UPDATE)
    assert np.allclose(o.marker("vx"), vr[0::3], atol=1e-15)
    assert np.allclose(o.marker("vz"), vr[2::3], atol=1e-15)
    assert np.allclose(o.marker("x"), x0 + vr[0::3] * dt, atol=1e-15)


def test_fibrin_graph_matches_reference_definelinks(pfg):
    L = _ref()
    n = 300
    lo = np.array([0, 0, 0], np.float32)
    hi = np.array([20, 20, 20], np.float32)
    x, y, z = (np.zeros(n, np.float32) for _ in range(3))
    link = np.zeros(2 * 100000, np.int32)
    l0 = np.zeros(100000, np.float32)
    nl = C.c_int32()
    fp = C.POINTER(C.c_float)
    pfg.lib().pfg_geom_fibrin(n, lo.ctypes.data_as(fp), hi.ctypes.data_as(fp), 3.0, 100000,
                              7, x.ctypes.data_as(fp), y.ctypes.data_as(fp),
                              z.ctypes.data_as(fp), link.ctypes.data_as(C.POINTER(C.c_int32)),
                              l0.ctypes.data_as(fp), C.byref(nl))
    mine = set(map(tuple, link[:2 * nl.value].reshape(-1, 2)))
    xyz = np.ascontiguousarray(np.stack([x, y, z], 1).astype(np.float64).ravel())
    counts = np.zeros(n, np.int32)
    lists = np.zeros(n * n, np.int32)
    tot = L.ref_definelinks(n, xyz.ctypes.data_as(C.POINTER(C.c_double)), 9.0,
                            counts.ctypes.data_as(C.POINTER(C.c_int)),
                            lists.ctypes.data_as(C.POINTER(C.c_int)), n * n)
    assert tot > 0
    ref = set()
    k = 0
    for i in range(n):
        for j in lists[k:k + counts[i]]:
            if j > i:
                ref.add((i, int(j)))
        k += counts[i]
    assert mine == ref


def test_bell_probabilities_closed_form(pfg, po):
    """One bound receptor at zero tension ruptures with p = 1 - exp(-k_off0);
    a free one in reach binds with p = 1 - exp(-k_on).  Over many steps the
    oracle's event counts follow those rates."""
    rng = np.random.default_rng(5)
    n_rec = 400
    # a flat sheet of triangles at z = 0.3 above ligands at z = 0
    X, tri = [], []
    for r in range(n_rec):
        b = len(X)
        X += [[r * 3.0, 0.0, 0.3], [r * 3.0 + 0.1, 0.0, 0.3], [r * 3.0, 0.1, 0.3]]
        tri += [b, b + 1, b + 2]
    X = np.array(X)
    p = pfg.default_params(nx=1200, ny=4, nz=4, k_on=0.05, k_off0=0.2, bond_k=0.0,
                           bond_reach=1.0, wall_z=0.0, bin_cells=2)
    p.wall[2] = pfg.WALL_BOUNCE
    ma = ModelArrays(X[:, 0], X[:, 1], X[:, 2], body=np.repeat(np.arange(n_rec), 3),
                     plt_elem_start=np.arange(n_rec + 1) * 3, plt_tri_start=np.arange(n_rec + 1),
                     tri=tri, tri_S0=np.ones(n_rec), plt_S0=np.ones(n_rec), plt_V0=np.ones(n_rec),
                     rec_tri=np.arange(n_rec), rec_r1=np.zeros(n_rec), rec_r2=np.zeros(n_rec),
                     lig_x=np.arange(n_rec) * 3.0, lig_y=np.zeros(n_rec), lig_z=np.zeros(n_rec))
    o = po.Oracle(p, model=ma.model)
    o.stage(pfg.STAGE_CELLS)
    o.stage(pfg.STAGE_BONDS)
    rb, _, _ = o.bonds()
    frac_on = (rb >= 0).mean()
    p_on = 1 - np.exp(-np.float32(0.05))
    assert abs(frac_on - p_on) < 4 * np.sqrt(p_on * (1 - p_on) / n_rec)
    bound = rb >= 0
    o.set_step(1)
    o.stage(pfg.STAGE_CELLS)
    o.stage(pfg.STAGE_BONDS)
    rb2, _, _ = o.bonds()
    p_off = 1 - np.exp(-np.float32(0.2))
    frac_off = (rb2[bound] < 0).mean()
    nb = bound.sum()
    assert abs(frac_off - p_off) < 4 * np.sqrt(p_off * (1 - p_off) / nb)
