"""Whole coupled steps on the device against the FP64 oracle run in the same
(fused) order from the same seeded scene."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
Q = 19


def run_both(pfg, po, sc, steps, graph=True, staging=0):
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        s.set_graph(graph)
        s.set_ib_staging(staging)
        info = s.step(steps)
        f = s.f().astype(np.float64)
        mk = s.markers()
        rho, ux, uy, uz = s.fields()
        bonds = s.bonds()[0] if s.n_rec else None
    o = po.Oracle(sc.params, scene=sc)
    o.step(steps)
    return dict(info=info, f=f, mk=mk, rho=rho, u=(ux, uy, uz), bonds=bonds), o


def rel_l2(a, b):
    return np.sqrt(((a - b) ** 2).sum()) / max(np.sqrt((b ** 2).sum()), 1e-300)


def receptor_platelets(sc):
    """Platelet of each receptor (through its triangle)."""
    return np.searchsorted(sc.plt_tri_start, sc.rec_tri, side="right") - 1


def run_stepwise(pfg, po, sc, steps):
    """Both runs one step at a time, keeping the start-of-step positions of
    every step (FP32 on the device, FP64 in the oracle)."""
    dev_pos, orc_pos = [], []
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        for _ in range(steps):
            mk = s.markers()
            dev_pos.append(np.stack([mk[k] for k in ("x", "y", "z")], 1))
            info = s.step(1)
        f = s.f().astype(np.float64)
        mk = s.markers()
        rho, ux, uy, uz = s.fields()
        bonds = s.bonds()[0] if s.n_rec else None
    o = po.Oracle(sc.params, scene=sc)
    for _ in range(steps):
        orc_pos.append(np.stack([o.marker(k).copy() for k in ("x", "y", "z")], 1))
        o.step(1)
    g = dict(info=info, f=f, mk=mk, rho=rho, u=(ux, uy, uz), bonds=bonds)
    return g, o, dev_pos, orc_pos


def _fma32(a, b, c):
    """float32 a*b + c with one rounding (the product is exact in FP64)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def morse_switches(sc, dev_pos, orc_pos):
    """Platelets with a Morse pair that the device and the oracle decide
    differently in some step: the decision of both is the squared distance
    of the start-of-step FP32 positions, in FP32 in a fixed order, against
    the FP32 squared cutoff, so it can only differ where their positions
    differ in the last bits.  Returns the per-platelet flags and the largest
    |r - cutoff| over the pairs that switched."""
    from scipy.spatial import cKDTree
    p = sc.params
    cut2 = np.float32(p.morse_cutoff) * np.float32(p.morse_cutoff)
    body = sc.body
    flags = np.zeros(sc.plt_S0.size, bool)
    worst = 0.0
    if p.morse_u0 == 0 or not flags.size:
        return flags, worst
    el = np.flatnonzero(body >= 0)
    for Pd, Po in zip(dev_pos, orc_pos):
        Pd = Pd.astype(np.float32)
        Po = Po.astype(np.float32)
        pairs = set()
        for P in (Pd, Po):
            t = cKDTree(P[el].astype(np.float64))
            pr = t.query_pairs(float(p.morse_cutoff) + 1e-3, output_type="ndarray")
            pairs.update(map(tuple, el[pr]))
        if not pairs:
            continue
        pr = np.array(sorted(pairs))
        i, j = pr[:, 0], pr[:, 1]
        same = body[i] == body[j]
        i, j = i[same], j[same]

        def inside(P):
            d = P[j] - P[i]  # FP32 differences
            r2 = _fma32(d[:, 2], d[:, 2], _fma32(d[:, 1], d[:, 1], d[:, 0] * d[:, 0]))
            return r2 < cut2
        sw = inside(Pd) != inside(Po)
        if sw.any():
            flags[np.unique(body[i[sw]])] = True
            r = np.linalg.norm((Po[j[sw]] - Po[i[sw]]).astype(np.float64), axis=1)
            worst = max(worst, float(np.abs(r - p.morse_cutoff).max()))
    return flags, worst


def excluded_platelets(sc, o, gpu_bonds=None, switched=None):
    """Platelets whose elements are held to the loose tolerance: those with
    a receptor whose bond differs between the runs (a decision within
    epsilon of its threshold, counted) and those with a Morse pair that the
    two runs decided differently (a pair within rounding of the cutoff)."""
    npl = sc.plt_S0.size
    out = np.zeros(npl, bool) if switched is None else switched.copy()
    if gpu_bonds is not None and sc.rec_tri.size:
        bad = gpu_bonds != o.bonds()[0]
        out[np.unique(receptor_platelets(sc)[bad])] = True
    return out


def compared_markers(sc, excl):
    """Markers held to the tight tolerance: fibrin nodes and the elements of
    the platelets not excluded."""
    body = sc.body
    keep = np.ones(body.size, bool)
    el = body >= 0
    keep[el] = ~excl[body[el]]
    return keep


def check_positions(pos, opos, keep, tight=1e-4, loose=3e-3):
    """Every marker is compared: those in `keep` within `tight`, the others
    (platelets with a counted near-threshold event, where one switched pair
    force moves the platelet by up to ~1e-3 within a few steps) within
    `loose`.  Returns the two maxima."""
    err = np.abs(pos - opos).max(axis=0)
    e_keep = float(err[keep].max()) if keep.any() else 0.0
    e_rest = float(err[~keep].max()) if (~keep).any() else 0.0
    assert e_keep < tight, e_keep
    assert e_rest < loose, e_rest
    return e_keep, e_rest


def cells_away_from(sc, pts, radius):
    """Boolean mask over the lattice cells farther than `radius` (per axis,
    periodic where the lattice is) from every point of pts (n x 3)."""
    p = sc.params
    dims = (p.nx, p.ny, p.nz)
    near = np.zeros((p.nz, p.ny, p.nx), bool)
    for q in pts:
        idx = []
        for a in range(3):
            c = int(np.floor(q[a]))
            r = np.arange(c - radius, c + radius + 2)
            r = r % dims[a] if p.wall[a] == 0 else r[(r >= 0) & (r < dims[a])]
            idx.append(r)
        near[np.ix_(idx[2], idx[1], idx[0])] = True
    return ~near.ravel()


def test_channel_platelet_short_horizon(gpu, po):
    """64x32x32 channel with the 256-element Morse platelet, 50 steps: the
    FP32 trajectory follows the FP64 one element by element."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_CHANNEL)
    g, o = run_both(pfg, po, sc, 50)
    ofx = po.fields(po.lattice(sc.params), o.f())
    err_u = rel_l2(np.stack(g["u"]), np.stack(ofx[1:]))
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    err_x = np.abs(pos - opos).max()
    print(f"channel 50 steps: rel L2 u {err_u:.2e}, max position error {err_x:.2e}")
    assert err_u < 1e-3
    assert err_x < 1e-4


def test_headline_config_full_size_short_horizon(gpu, po):
    """The headline configuration at full size (256x128x128 Couette, tau 0.7,
    wall velocity 0.05, one 2048-element platelet with 512 receptors over
    4096 wall ligands, FP32 storage), 5 coupled steps against the FP64
    oracle: fields, every marker and every bond."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR)
    assert sc.N == 256 * 128 * 128 and sc.n_markers == 2048
    g, o, dpos, opos_steps = run_stepwise(pfg, po, sc, 5)
    switched, worst = morse_switches(sc, dpos, opos_steps)
    assert worst < 1e-4
    ofx = po.fields(po.lattice(sc.params), o.f())
    err_u = rel_l2(np.stack(g["u"]), np.stack(ofx[1:]))
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    err_x = np.abs(pos - opos).max()
    ob = o.bonds()[0]
    mism = int((g["bonds"] != ob).sum())
    near = o.counters()["near_threshold"]
    excl = excluded_platelets(sc, o, g["bonds"], switched)
    keep = compared_markers(sc, excl)
    print(f"headline config, 5 steps: rel L2 u {err_u:.2e}, max position error {err_x:.2e}, "
          f"bonds {int((ob >= 0).sum())}, mismatches {mism} (near threshold {near}), "
          f"platelets with a near-threshold event {int(excl.sum())} "
          f"(near-cutoff pair visits {o.counters()['near_cutoff']}, device "
          f"{g['info']['near_cutoff']})")
    assert err_u < 1e-5
    assert mism <= near
    check_positions(pos, opos, keep)  # coordinates near 100: a few FP32 ulps


def test_headline_config_variant_fast_binding_full_size(gpu, po):
    """Variant of the headline configuration with a binding rate of 0.5 per
    step instead of 1e-3 (everything else as configured: 256x128x128
    Couette, one 2048-element platelet with 512 receptors half a cell above
    4096 wall ligands, FP32 storage), so that bonds form within the short
    horizon: 6 coupled steps against the FP64 oracle, every bond (exact
    except decisions within epsilon of their threshold, counted), the bond
    forces through the element positions, and the flow."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR)
    sc.params.k_on = 0.5
    g, o, dpos, opos_steps = run_stepwise(pfg, po, sc, 6)
    switched, worst = morse_switches(sc, dpos, opos_steps)
    assert worst < 1e-4
    ob = o.bonds()[0]
    nb = int((ob >= 0).sum())
    mism = int((g["bonds"] != ob).sum())
    near = o.counters()["near_threshold"]
    ofx = po.fields(po.lattice(sc.params), o.f())
    err_u = rel_l2(np.stack(g["u"]), np.stack(ofx[1:]))
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    keep = compared_markers(sc, excluded_platelets(sc, o, g["bonds"], switched))
    assert nb >= 5
    assert mism <= near
    assert err_u < 1e-5
    e_keep, e_rest = check_positions(pos, opos, keep)
    print(f"fast-binding headline variant, 6 steps: bonds {nb} (events "
          f"{o.counters()['bond_events']}), mismatches {mism} (near threshold {near}), "
          f"rel L2 u {err_u:.2e}, position errors {e_keep:.2e} (tight set, "
          f"{int(keep.sum())} of {keep.size}) / {e_rest:.2e}")


def test_channel_platelet_config(gpu, po):
    """The CPU-runnable configuration in full: 64x32x32 channel, tau 0.8,
    F = 1e-6, one 256-element Morse platelet, 4-point IB, 500 steps.
    The Morse cutoff makes single-element trajectories sensitive to rounding
    (a pair crossing the cutoff switches a force on or off), so the long run
    is compared through the platelet centroid, its FP64 membrane volume and
    the flow rate."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_CHANNEL)
    g, o = run_both(pfg, po, sc, 500)
    ofx = po.fields(po.lattice(sc.params), o.f())
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    dcen = np.abs(pos.mean(1) - opos.mean(1)).max()
    disp = np.abs(opos.mean(1) - np.array([sc.x.mean(), sc.y.mean(), sc.z.mean()])).max()
    q_g, q_o = g["u"][0].mean(), ofx[1].mean()
    print(f"channel 500 steps: centroid error {dcen:.2e} (displacement {disp:.2e}), "
          f"flow rate {q_g:.6e} vs {q_o:.6e}")
    assert dcen < 1e-3 * max(disp, 1.0)
    assert abs(q_g - q_o) < 1e-3 * abs(q_o)
    assert pos[0].mean() > sc.x.mean()  # the platelet moved with the flow


@pytest.mark.parametrize("kbend", [0.0, 0.5])
def test_reference_mesh_platelet(gpu, po, kbend):
    """The reference's own platelet mesh (Platelet_Mesh_958.neu, 958 nodes,
    1912 triangles) in the channel, with mesh springs and the area and
    volume constraints on, without and with membrane bending, 50 steps
    against the oracle.  Morse is off: its cutoff switches pair forces on
    and off on rounding, which the element-wise comparison here would see
    (the channel tests above cover it through centroids)."""
    import os
    pfg = gpu
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures",
                        "Platelet_Mesh_958.neu")
    sc = pfg.Scene(pfg.SCENE_CHANNEL, mesh_path=path)
    sc.params.k_spring, sc.params.ks, sc.params.kv = 0.3, 0.5, 0.4
    sc.params.morse_u0 = 0.0
    sc.params.kbend = kbend
    g, o = run_both(pfg, po, sc, 50)
    ofx = po.fields(po.lattice(sc.params), o.f())
    err_u = rel_l2(np.stack(g["u"]), np.stack(ofx[1:]))
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    err_x = np.abs(pos - opos).max()
    print(f"mesh platelet 50 steps: rel L2 u {err_u:.2e}, max position error {err_x:.2e}")
    assert sc.n_markers == 958
    assert err_u < 1e-3
    assert err_x < 1e-4


def test_shear_with_bonds_short(gpu, po):
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=128,
                   ligands=800)
    sc.params.k_on = 0.2
    g, o = run_both(pfg, po, sc, 20)
    ob = o.bonds()[0]
    mism = int((g["bonds"] != ob).sum())
    near = o.counters()["near_threshold"]
    print(f"shear: bonds gpu {int((g['bonds'] >= 0).sum())} oracle {int((ob >= 0).sum())}, "
          f"mismatches {mism}, near-threshold {near}")
    assert mism <= near
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    check_positions(pos, opos, compared_markers(sc, excluded_platelets(sc, o, g["bonds"])))


@pytest.mark.parametrize("staging", [1, 2])
def test_thrombus_small_flags_fibrin(gpu, po, staging):
    """Tube flags, three platelets, fibrin nodes spread through the lattice
    (their chunks' boxes exceed LDS and fall back to direct spreading), with
    both spreading forms."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=3, elements=150,
                   receptors=16, ligands=400, fibrin_nodes=500, fibrin_links=800)
    g, o = run_both(pfg, po, sc, 15, staging=staging)
    fluid = sc.flags == 0
    w = np.array([po.load().orc_w(d) for d in range(Q)])[:, None]
    d = np.abs((g["f"] - w) - (o.f() - w))[:, fluid]
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    print(f"thrombus: max f-dev error {d.max():.2e}, max position error "
          f"{np.abs(pos - opos).max():.2e}")
    assert d.max() < 1e-5
    assert np.abs(pos - opos).max() < 1e-4


def test_thrombus_config_full_size_two_steps(gpu, po):
    """The tube configuration at full size with its own physics (512x256x256
    flag tube of radius 120, body force, 500 platelets x 1000 elements with
    the hard Morse cutoff, fibrin 20k nodes / 30k links, FP32 storage), 2
    coupled steps against the FP64 oracle.  Both take the Morse cutoff
    decision from the same FP32 arithmetic on their start-of-step positions,
    so they take the same pairs while the positions agree; where the
    positions differ in the last bits a pair within rounding of the cutoff
    can switch.  The decisions of both runs are replayed here from their
    positions of every step: the platelets with a pair decided differently
    (each within 1e-4 of the cutoff) and those with a bond decision that
    differs are held to a looser tolerance, and the fluid cells within four
    cells of them are left out.  Everything else is compared tightly: every
    other element and fibrin node, every other fluid cell, every bond."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS)
    assert sc.N == 512 * 256 * 256 and sc.n_markers == 520000
    assert sc.params.morse_taper == 0.0
    g, o, dpos, opos_steps = run_stepwise(pfg, po, sc, 2)
    switched, worst = morse_switches(sc, dpos, opos_steps)
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    ob = o.bonds()[0]
    mism = int((g["bonds"] != ob).sum())
    near = o.counters()["near_threshold"]
    excl = excluded_platelets(sc, o, g["bonds"], switched)
    keep = compared_markers(sc, excl)
    away = (sc.flags == 0) & cells_away_from(sc, opos[:, ~keep].T, 4)
    w = np.array([po.load().orc_w(d) for d in range(Q)])[:, None]
    d = 0.0
    for q in range(Q):  # one direction at a time keeps the host memory low
        d = max(d, float(np.abs((g["f"][q] - w[q]) - (o.f()[q] - w[q]))[away].max()))
    print(f"tube config, 2 steps: max f-dev error {d:.2e} over {int(away.sum())} fluid "
          f"cells, bonds {int((ob >= 0).sum())}, mismatches {mism} (near threshold {near}), "
          f"platelets with a switched Morse pair or bond {int(excl.sum())} of {excl.size} "
          f"(largest |r - cutoff| of a switched pair {worst:.1e}; near-cutoff pair visits: "
          f"oracle {o.counters()['near_cutoff']}, device {g['info']['near_cutoff']})")
    assert mism <= near
    assert worst < 1e-4  # only pairs within rounding of the cutoff switch
    assert excl.sum() <= 0.05 * excl.size
    assert d < 1e-5
    e_keep, e_rest = check_positions(pos, opos, keep)
    print(f"  position errors {e_keep:.2e} ({int(keep.sum())} markers) / {e_rest:.2e}")


def test_thrombus_fibrin_from_file(gpu, po):
    """The reference's fibrin network (HemoThickInputCor.txt, 1175 nodes, 1422
    links) fitted into the tube's fibrin box with two platelets, 15 steps."""
    import os
    pfg = gpu
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures",
                        "HemoThickInputCor.txt")
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=96, ny=64, nz=64, n_platelets=2, elements=150,
                   receptors=16, ligands=400, fibrin_path=path)
    assert sc.link.size // 2 == 1422
    g, o = run_both(pfg, po, sc, 15)
    pos = np.stack([g["mk"][k] for k in ("x", "y", "z")])
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")])
    err = np.abs(pos - opos).max()
    print(f"fibrin file: max position error {err:.2e}")
    assert err < 1e-4


def test_large_config_subbox_128_fp16_hard_cutoff(gpu, po):
    """The 1024x512x512 configuration's own physics (hard Morse cutoff, FP16
    deviation storage) on a 128^3 sub-box at its densities (16 platelets,
    fibrin, wall ligands), 50 coupled steps against the FP64 oracle:
    velocity relative L2 within the FP16 tolerance 1e-3, density, platelet
    centroids, and the same with FP32 storage.  50 steps: with the hard
    cutoff a pair that crosses it a step apart in two runs moves its
    platelet, and such differences grow along the trajectory -- an FP64 run
    of this scene with its positions perturbed by 1e-6 departs from the
    unperturbed one by ~1e-4 in velocity within 50 steps and ~5e-3 within
    75; measured here, the FP16 and FP32 stores give 6.6e-4 and 5.9e-4 at
    50 steps and both ~2.5e-3 at 200, so beyond ~100 steps the comparison
    measures the trajectory, not the storage.  Near-cutoff pair visits are
    counted on both sides."""
    pfg = gpu
    sc = pfg.Scene.large_subbox(128)
    assert sc.params.store == pfg.STORE_F16 and sc.params.morse_taper == 0.0
    npl = int(sc.plt_S0.size)
    steps = 50
    dev = {}
    for store in (pfg.STORE_F16, pfg.STORE_F32):
        sc.params.store = store
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            info = s.step(steps)
            dev[store] = (s.fields(), s.markers(), info)
    sc.params.store = pfg.STORE_F16
    o = po.Oracle(sc.params, scene=sc)
    o.step(steps)
    orho, ox, oy, oz = po.fields(po.lattice(sc.params), o.f())
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")], 1)
    body = sc.body
    ocen = np.array([opos[body == b].mean(0) for b in range(npl)])
    res = {}
    for store, tag in ((pfg.STORE_F16, "FP16"), (pfg.STORE_F32, "FP32")):
        (rho, ux, uy, uz), mk, info = dev[store]
        pos = np.stack([mk[k] for k in ("x", "y", "z")], 1).astype(np.float64)
        cen = np.array([pos[body == b].mean(0) for b in range(npl)])
        res[tag] = (rel_l2(np.stack([ux, uy, uz]), np.stack([ox, oy, oz])),
                    rel_l2(rho, orho), float(np.linalg.norm(cen - ocen, axis=1).max()),
                    info["near_cutoff"])
        print(f"128^3 sub-box, {steps} steps, {tag} store: rel L2 u {res[tag][0]:.2e}, "
              f"rho {res[tag][1]:.2e}, centroid max {res[tag][2]:.2e}, device near-cutoff "
              f"pair visits {res[tag][3]}")
    print(f"  oracle near-cutoff pair visits {o.counters()['near_cutoff']} over "
          f"{int((o.platelets_near_cutoff() > 0).sum())} of {npl} platelets")
    for tag in ("FP16", "FP32"):
        err_u, err_rho, err_cen, _ = res[tag]
        assert err_u < 1e-3, tag
        assert err_rho < 2e-5, tag
        assert err_cen < 5e-4, tag


def test_large_config_full_size_conservation_and_profile(gpu):
    """The 1024x512x512 configuration at full size on the device (268M cells,
    FP16 storage, 2000 platelets x 1000 elements, fibrin 80k / 120k, 65536
    wall ligands), 20 coupled steps: FP64 totals of mass (rounding to FP16
    is unbiased: the relative drift stays near 1e-9), the Couette profile of
    the plane-averaged velocity, and every marker finite and inside the
    lattice."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_LARGE)
    p = sc.params
    assert sc.N == 1024 * 512 * 512 and p.store == pfg.STORE_F16
    with pfg.Solver(p) as s:
        s.load(sc)
        c0 = s.conservation()
        info = s.step(20)
        c1 = s.conservation()
        _, ux, _, _ = s.fields()
        mk = s.markers()
    drift = abs(c1[0] - c0[0]) / c0[0]
    prof = ux.reshape(p.nz, p.ny * p.nx).mean(axis=1, dtype=np.float64)
    exact = 0.05 * (np.arange(p.nz) + 0.5) / p.nz
    dev_prof = np.abs(prof - exact).max()
    pos = np.stack([mk[k] for k in ("x", "y", "z")])
    print(f"1024x512x512 FP16, 20 steps: relative mass drift {drift:.2e}, Couette profile "
          f"max deviation {dev_prof:.2e} (wall velocity 0.05), bonds {info['total_bonds']}, "
          f"near-cutoff pair visits {info['near_cutoff']}")
    assert drift < 1e-7
    assert dev_prof < 1e-3
    assert np.isfinite(pos).all()
    assert pos[2].min() > -0.5 and pos[2].max() < p.nz - 0.5


def test_forces_with_spreading_equals_separate_passes(gpu):
    """With a wavefront per marker and direct spreading the force pass also
    spreads; halting at detach turns that off (the test never fires here:
    detach rate 0).  Same steps both ways, up to the FP32 order of the
    force-field atomics."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=64,
                   ligands=400)
    sc.params.detach_rate = 0.0
    out = []
    for halt in (False, True):
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            s.set_halt_on_detach(halt)
            s.step(20)
            mk = s.markers()
            out.append((np.stack([mk[k] for k in ("x", "y", "z")]), s.fields()[1]))
    assert np.abs(out[0][0] - out[1][0]).max() < 1e-5
    assert np.abs(out[0][1] - out[1][1]).max() < 1e-6


def test_autotune_restores_the_state(gpu):
    """Coupled shear scene with bonds: the autotuner times every shape from a
    device snapshot and restores it, so the run continues from where it was
    (same step index, bonds and trajectory as an untuned run, up to the FP32
    summation order that the chosen shapes may change)."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=128,
                   ligands=800)
    sc.params.k_on = 0.2
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        ref_info = s.step(20)
        ref = s.markers()
        ref_bonds = s.bonds()[0]
        _, ref_u, _, _ = s.fields()
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        s.step(10)
        before = s.markers()
        t = s.autotune(8)
        print(f"autotune: {t}")
        assert t["shapes_timed"] >= 7
        assert t["forces_lanes"] in (0, 1, 2, 4, 8, 16, 32, 64) and t["ib_staging"] in (0, 1, 2, 3, 4)
        after = s.markers()
        for k in ("x", "y", "z"):
            assert np.array_equal(after[k], before[k])
        info = s.step(10)
        assert info["step"] == 20
        got = s.markers()
        _, u, _, _ = s.fields()
        mism = int((s.bonds()[0] != ref_bonds).sum())
    # a bond decision within epsilon of its threshold may flip with the
    # summation order; those are counted
    assert mism <= max(info["near_threshold"], ref_info["near_threshold"])
    # one platelet: a flipped bond moves it (its bond force switches), so the
    # tolerance widens only then
    err = max(np.abs(got[k] - ref[k]).max() for k in ("x", "y", "z"))
    assert err < (1e-4 if mism == 0 else 3e-3)
    assert np.abs(u - ref_u).max() < (1e-6 if mism == 0 else 1e-4)


@pytest.mark.parametrize("graph", [True, False])
def test_halt_on_detach_freezes_the_run(gpu, po, graph):
    """With no ligands no bond can form, so the detach test runs every step
    (hazard 0.02; with this scene's seed the first draw below it comes at
    step 14).  A run halted at detach ends in the state of a plain run of
    exactly the steps that completed, and the detach step is the oracle's
    (same counter-RNG draws)."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=32,
                   ligands=-1)
    sc.params.detach_rate = 0.02
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        s.set_graph(graph)
        s.set_halt_on_detach(True)
        info = s.step(30)
        assert info["detached"] == 1
        done = info["step"]
        assert done > 0
        assert done == info["detach_step"]
        f_halt = s.f()
        mk_halt = s.markers()
        info2 = s.step(5)  # further steps change nothing
        assert info2["step"] == done
        assert np.array_equal(s.f(), f_halt)
        assert np.array_equal(s.markers()["x"], mk_halt["x"])
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        s.step(done)
        f_plain = s.f()
        mk_plain = s.markers()
    o = po.Oracle(sc.params, scene=sc)
    o.step(done + 1)
    print(f"halt: detached after {done} steps (oracle {o.counters()['detach_step']})")
    assert o.counters()["detach_step"] == done
    assert np.abs(f_halt - f_plain).max() < 1e-6
    for k in ("x", "y", "z"):
        assert np.abs(mk_halt[k] - mk_plain[k]).max() < 1e-5


@pytest.mark.parametrize("graph", [True, False])
def test_overlap_equals_serial(gpu, graph):
    """The LB rows no IB stencil touches update on a second stream while the
    marker pipeline runs; the result equals the serial step (the spread force
    field sums atomics in arrival order, hence the FP32 tolerance).  A
    platelet near a periodic face makes its stencil rows wrap."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=64,
                   ligands=600)
    sc.x[:] += 30.0  # across the periodic x face
    sc.params.k_on = 0.2
    res = []
    for overlap in (False, True):
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            s.set_graph(graph)
            s.set_overlap(overlap)
            s.step(24)
            res.append((s.f(), s.markers(), s.bonds()[0]))
    (f0, m0, b0), (f1, m1, b1) = res
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, np.float32)[:, None]
    assert np.abs((f0 - w) - (f1 - w)).max() < 1e-6
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.abs(m0[k] - m1[k]).max() < 1e-5, k
    assert (b0 != b1).sum() <= 1
