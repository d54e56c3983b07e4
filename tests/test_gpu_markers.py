"""Stage-level parity of the marker kernels with the FP64 oracle on identical
inputs: cell lists (exact), membrane sums (FP64), forces, IB spreading and
interpolation, bond decisions (exact except counted near-threshold cases)."""
import os

import numpy as np
import pytest

from helpers import ModelArrays

pytestmark = pytest.mark.gpu


def scene_small(pfg, **kw):
    base = dict(nx=48, ny=32, nz=24, elements=256, receptors=64, ligands=600)
    base.update(kw)
    return pfg.Scene(pfg.SCENE_SHEAR, **base)


def perturbed(sc, amp=0.08, seed=3):
    rng = np.random.default_rng(seed)
    x = (sc.x + amp * rng.standard_normal(sc.x.size)).astype(np.float32)
    y = (sc.y + amp * rng.standard_normal(sc.x.size)).astype(np.float32)
    z = (sc.z + amp * rng.standard_normal(sc.x.size)).astype(np.float32)
    return x, y, z


def pair(pfg, po, sc, params=None):
    p = params or sc.params
    s = pfg.Solver(p)
    s.load(sc)
    o = po.Oracle(p, scene=sc)
    return s, o


def set_pos(s, o, x, y, z):
    s.set_markers(x, y, z)
    o.marker("x")[:] = x
    o.marker("y")[:] = y
    o.marker("z")[:] = z


def test_cell_list_exact(gpu, po):
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=4, elements=200,
                   receptors=8, ligands=100, fibrin_nodes=500, fibrin_links=800)
    s, o = pair(pfg, po, sc)
    with s:
        x, y, z = perturbed(sc, 0.3)
        x[:10] += 64.0  # unwrapped positions beyond the periodic box
        x[10:20] -= 64.0
        set_pos(s, o, x, y, z)
        s.stage(pfg.STAGE_CELLS)
        o.stage(pfg.STAGE_CELLS)
        dims, start, items = s.cells()
        onb, ostart, oitems = o.cells()
    assert np.array_equal(dims, onb)
    assert np.array_equal(start, ostart)
    assert np.array_equal(items, oitems)


def test_membrane_sums_fp64(gpu, po):
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=3, elements=300,
                   receptors=4, ligands=10, fibrin_nodes=-1)
    s, o = pair(pfg, po, sc)
    with s:
        set_pos(s, o, *perturbed(sc, 0.1))
        s.stage(pfg.STAGE_MEMBRANE)
        o.stage(pfg.STAGE_MEMBRANE)
        S, V, tS = s.membrane()
        oS, oV, otS = o.membrane()
    assert np.allclose(S, oS, rtol=1e-12)
    assert np.allclose(V, oV, rtol=1e-12)
    assert np.allclose(tS, otS, rtol=1e-12)


@pytest.mark.parametrize("lanes", [0, 8, 4, 1, -1])
@pytest.mark.parametrize("what", ["morse", "springs_area_volume", "wall", "fibrin_wlc",
                                  "fibrin_linear_refsign", "bending", "bending_mesh"])
def test_forces(gpu, po, what, lanes):
    """Marker forces against the oracle with a wavefront per marker (0: the
    automatic choice at this size), with 8, 4 or 1 lanes per marker in
    cell-list order, and (-1) by platelet: each platelet's Morse pairs from
    a cell list of its own in LDS, the other markers with 4 lanes each."""
    pfg = gpu
    if what.startswith("fibrin"):
        sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=2, elements=150,
                       receptors=-1, ligands=-1, fibrin_nodes=600, fibrin_links=900)
    elif what == "bending_mesh":  # the reference's platelet mesh, two copies
        sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=96, ny=64, nz=64, n_platelets=2, receptors=-1,
                       ligands=-1, fibrin_nodes=-1, mesh_path=os.path.join(
                           os.path.dirname(os.path.abspath(__file__)), "fixtures",
                           "Platelet_Mesh_958.neu"))
    else:
        sc = scene_small(pfg, receptors=-1, ligands=-1)
    p = sc.params
    p.morse_u0 = 1.0 if what == "morse" else 0.0
    if what == "springs_area_volume":
        p.k_spring, p.ks, p.kv = 0.3, 0.5, 0.4
    if what == "wall":
        p.wall_rep_a, p.wall_rep_kappa, p.wall_rep_hc = 0.01, 10.0, 3.0
    if what == "fibrin_linear_refsign":
        p.fiber_law, p.fiber_ref_sign, p.fiber_k = pfg.FIBER_LINEAR, 1, 0.3
    if what.startswith("bending"):  # the curvature chain alone (FP64 on both sides)
        p.kbend = 2.0
    s, o = pair(pfg, po, sc, p)
    with s:
        if lanes < 0:
            s.set_forces_by_body(1)
            s.set_forces_lanes(4)
        else:
            s.set_forces_by_body(2)
            s.set_forces_lanes(lanes)
        set_pos(s, o, *perturbed(sc, 0.1))
        for st in (pfg.STAGE_CELLS, pfg.STAGE_MEMBRANE, pfg.STAGE_BONDS, pfg.STAGE_FORCES):
            s.stage(st)
            o.stage(st)
        mk = s.markers()
    F = np.stack([mk["fx"], mk["fy"], mk["fz"]])
    oF = np.stack([o.marker("fx"), o.marker("fy"), o.marker("fz")])
    scale = np.abs(oF).max()
    assert scale > 0
    assert np.abs(F - oF).max() < 2e-5 * scale, (np.abs(F - oF).max(), scale)


@pytest.mark.parametrize("staging", [1, 2, 3, 4, -4])
def test_spread_interp_parity(gpu, po, staging):
    """The four spreading forms (1: LDS-staged chunks, 2: direct atomics,
    3: tile gather over the cell lists, 4: a workgroup per platelet storing
    the tiles the force pass claimed for it alone; -4: form 4 without the
    claims, so every node is added with atomics) against the oracle's
    stencil form; a seventh of the markers sit one periodic image away, so
    staged chunks, platelet boxes and gathered tiles straddle the periodic
    boundary.  Forms 1, 3 and 4 also stage the interpolation by chunk."""
    pfg = gpu
    sc = scene_small(pfg, receptors=-1, ligands=-1)
    rng = np.random.default_rng(5)
    n = sc.n_markers
    f = (1e-3 * rng.standard_normal((3, n))).astype(np.float32)
    s, o = pair(pfg, po, sc)
    with s:
        s.set_ib_staging(abs(staging))
        x, y, z = perturbed(sc, 0.1)
        x[::7] += sc.params.nx  # wrap on the periodic axis
        set_pos(s, o, x, y, z)
        if staging == 4:
            s.stage(pfg.STAGE_FORCES)  # the tile claims of these positions
        s.set_marker_forces(*f)
        for k, name in enumerate(("fx", "fy", "fz")):
            o.marker(name)[:] = f[k]
        if staging == 3:
            s.stage(pfg.STAGE_CELLS)  # the gather finds the markers on the cell lists
        s.stage(pfg.STAGE_SPREAD)
        o.stage(pfg.STAGE_SPREAD)
        Fg = s.aux()
        Fo = o.force_field()
        scale = np.abs(Fo).max() * sc.params.ib_force_scale / sc.params.ib_force_scale
        assert np.abs(Fg - Fo).max() < 1e-6 * max(scale, 1e-12) + 1e-12
        s.stage(pfg.STAGE_LB)
        o.stage(pfg.STAGE_LB)
        s.stage(pfg.STAGE_INTERP)
        o.stage(pfg.STAGE_INTERP)
        mk = s.markers()
        s.stage(pfg.STAGE_UPDATE)
        o.stage(pfg.STAGE_UPDATE)
        s.stage(pfg.STAGE_CLEAR)
        o.stage(pfg.STAGE_CLEAR)
        assert np.abs(s.aux()).max() == 0.0  # cleared tiles
        mk2 = s.markers()
    for k in ("vx", "vy", "vz"):
        assert np.abs(mk[k] - o.marker(k)).max() < 1e-6, k
    for k in ("x", "y", "z"):
        assert np.abs(mk2[k] - o.marker(k)).max() < 1e-5, k


def test_bond_decisions(gpu, po):
    pfg = gpu
    sc = scene_small(pfg, receptors=256, ligands=2000)
    p = sc.params
    p.k_on, p.k_off0 = 0.3, 0.1  # many events in few steps
    p.bond_reach = 1.5
    s, o = pair(pfg, po, sc, p)
    mism = 0
    with s:
        x, y, z = perturbed(sc, 0.05)
        set_pos(s, o, x, y, z)
        for step in range(6):
            s.set_step(step)
            o.set_step(step)
            for st in (pfg.STAGE_CELLS, pfg.STAGE_MEMBRANE, pfg.STAGE_BONDS):
                s.stage(st)
                o.stage(st)
            rb, lo, rf = s.bonds()
            orb, olo, orf = o.bonds()
            mism += int((rb != orb).sum())
            # continue both from the oracle's decisions
            s.set_bonds(orb)
        info = s.step(0)
    near = o.counters()["near_threshold"]
    print(f"bond decision mismatches {mism}, near-threshold cases {near}")
    assert mism <= near
    assert (orb >= 0).sum() > 10


@pytest.mark.parametrize("gap", [0.35, 0.7])
def test_contact_parity(gpu, po, gap):
    """Platelet-platelet contact on the cell lists against the oracle's brute
    force: the second of two platelets is moved so that the closest element
    pair is `gap` apart; only that pair feels the repulsion."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=2, elements=200,
                   receptors=-1, ligands=-1, fibrin_nodes=-1)
    p = sc.params
    p.morse_u0 = 0.0
    p.contact_a, p.contact_kappa, p.contact_hc = 0.01, 10.0, 1.0
    n0 = int(sc.plt_elem_start[1])
    P = np.stack([sc.x, sc.y, sc.z], 1).astype(np.float64)
    A, B = P[:n0], P[n0:]
    # put B next to A along x, then set the closest distance to `gap`
    B = B - B.mean(0) + A.mean(0) + np.array([A[:, 0].max() - A[:, 0].min() + 1.0, 0.3, 0.2])
    D = np.linalg.norm(A[:, None] - B[None], axis=2)
    B[:, 0] += gap - D.min()
    x, y, z = (np.concatenate([A[:, k], B[:, k]]).astype(np.float32) for k in range(3))
    s, o = pair(pfg, po, sc, p)
    with s:
        set_pos(s, o, x, y, z)
        for st in (pfg.STAGE_CELLS, pfg.STAGE_MEMBRANE, pfg.STAGE_BONDS, pfg.STAGE_FORCES):
            s.stage(st)
            o.stage(st)
        mk = s.markers()
        # a second evaluation finds the same pair (the slots were emptied)
        s.stage(pfg.STAGE_CELLS)
        s.stage(pfg.STAGE_FORCES)
        mk2 = s.markers()
    F = np.stack([mk["fx"], mk["fy"], mk["fz"]], 1)
    oF = np.stack([o.marker(k) for k in ("fx", "fy", "fz")], 1)
    nz = np.flatnonzero(np.abs(oF).sum(1))
    assert len(nz) == 2
    assert np.flatnonzero(np.abs(F).sum(1)).tolist() == nz.tolist()
    assert np.abs(F - oF).max() < 1e-5 * np.abs(oF).max()
    F2 = np.stack([mk2["fx"], mk2["fy"], mk2["fz"]], 1)
    assert np.array_equal(F2, F)


def test_kick_noise_parity(gpu, po):
    """Fibrin nodes moved by the kick with thermal noise: the device draws
    the oracle's normals (same counter RNG), two steps of the update."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=1, elements=150,
                   receptors=-1, ligands=-1, fibrin_nodes=600, fibrin_links=900)
    p = sc.params
    p.fiber_move, p.fiber_gamma, p.kick_temp = pfg.FIBER_KICK, 0.2, 1e-4
    s, o = pair(pfg, po, sc, p)
    with s:
        for _ in range(2):
            for st in (pfg.STAGE_CELLS, pfg.STAGE_MEMBRANE, pfg.STAGE_BONDS,
                       pfg.STAGE_FORCES, pfg.STAGE_SPREAD, pfg.STAGE_LB, pfg.STAGE_INTERP,
                       pfg.STAGE_UPDATE):
                s.stage(st)
                o.stage(st)
        mk = s.markers()
    fib = sc.kind == pfg.MARKER_FIBER
    v = np.stack([mk["vx"], mk["vy"], mk["vz"]], 1)[fib]
    ov = np.stack([o.marker(k) for k in ("vx", "vy", "vz")], 1)[fib]
    assert np.abs(ov).max() > 0
    assert np.abs(v - ov).max() < 1e-4 * np.abs(ov).max()


@pytest.mark.parametrize("staged", [True, False])
def test_bbk_elements_parity(gpu, po, staged):
    """Platelet elements on the Langevin leapfrog (BBK half kicks with
    thermal noise from the counter RNG, drift with half the fluid velocity,
    as verlet_1 couples them): 12 steps against the oracle, stage by stage
    (the fluid velocity kept apart from the elements' own) or as whole
    graph steps.  Morse is off so that no pair switches at its cutoff."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_CHANNEL)
    p = sc.params
    p.morse_u0 = 0.0
    p.k_spring = 0.3
    p.sem_integrator, p.sem_mass, p.sem_gamma = pfg.SEM_BBK, 1.5, 0.4
    p.sem_temp, p.sem_fluid_weight = 1e-5, 0.5
    steps = 12
    s, o = pair(pfg, po, sc, p)
    with s:
        if staged:
            for _ in range(steps):
                for st in range(pfg.STAGE_CLEAR + 1):  # a whole step, force clear included
                    s.stage(st)
                    o.stage(st)
        else:
            s.step(steps)
            o.step(steps)
        mk = s.markers()
    pos = np.stack([mk[k] for k in ("x", "y", "z")], 1)
    opos = np.stack([o.marker(k) for k in ("x", "y", "z")], 1)
    v = np.stack([mk[k] for k in ("vx", "vy", "vz")], 1)
    ov = np.stack([o.marker(k) for k in ("vx", "vy", "vz")], 1)
    x0 = np.stack([sc.x, sc.y, sc.z], 1)
    print(f"BBK {steps} steps ({'stages' if staged else 'graph'}): max position error "
          f"{np.abs(pos - opos).max():.2e} (moved {np.abs(opos - x0).max():.2e}), max "
          f"velocity error {np.abs(v - ov).max():.2e} of {np.abs(ov).max():.2e}")
    assert np.abs(ov).max() > 1e-4  # thermal and spring velocities, not the fluid's
    # positions near 32 carry FP32 steps of 2e-6 each; the spring and membrane
    # forces turn those into velocity differences of about 1e-6 a step, so the
    # velocities agree to 1e-3 of their size (a wrong or missing kick, noise
    # draw or weight is of the order of the velocity itself)
    assert np.abs(pos - opos).max() < 2e-5
    assert np.abs(v - ov).max() < 1e-3 * np.abs(ov).max()


def test_contact_dense_cluster_parity(gpu, po):
    """Seven platelets packed around a central one, each within the contact
    cutoff of it: an element of the centre can have more than four other
    platelets within reach, and every closest pair must still be found (the
    oracle by brute force over all pairs)."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=64, nz=64, n_platelets=7, elements=120,
                   receptors=-1, ligands=-1, fibrin_nodes=-1)
    p = sc.params
    p.morse_u0 = 0.0
    p.contact_a, p.contact_kappa, p.contact_hc = 0.01, 6.0, 2.0
    P = np.stack([sc.x, sc.y, sc.z], 1).astype(np.float64)
    starts = sc.plt_elem_start
    centre = np.array([32.0, 32.0, 32.0])
    dirs = [np.zeros(3)] + [np.array(d, float) for d in
                            ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                             [0.6, 0.6, -0.5])]
    out = []
    for q in range(7):
        A = P[starts[q]:starts[q + 1]]
        A = (A - A.mean(0)) * 0.35  # small platelets: all within reach of the centre
        out.append(A + centre + 2.2 * dirs[q])
    Q = np.concatenate(out)
    x, y, z = (Q[:, k].astype(np.float32) for k in range(3))
    s, o = pair(pfg, po, sc, p)
    with s:
        set_pos(s, o, x, y, z)
        for st in (pfg.STAGE_CELLS, pfg.STAGE_MEMBRANE, pfg.STAGE_BONDS, pfg.STAGE_FORCES):
            s.stage(st)
            o.stage(st)
        mk = s.markers()
        info = s.step(0)
    F = np.stack([mk["fx"], mk["fy"], mk["fz"]], 1)
    oF = np.stack([o.marker(k) for k in ("fx", "fy", "fz")], 1)
    pairs = int(np.count_nonzero(np.abs(oF).sum(1)))
    print(f"dense contact: {pairs} elements with a contact force, candidates proposed "
          f"beyond four platelets {info['contact_overflow']}")
    assert pairs >= 12
    assert info["contact_overflow"] > 0  # the case the fifth-platelet path covers
    assert np.flatnonzero(np.abs(F).sum(1)).tolist() == np.flatnonzero(np.abs(oF).sum(1)).tolist()
    assert np.abs(F - oF).max() < 1e-5 * np.abs(oF).max()


def test_spreading_forms_agree_on_rotated_platelets(gpu):
    """The 512x256x256 tube's 500 platelets each turned by 45 degrees about x
    around their centroids, so that the axis-aligned boxes of their chunks
    (cut at load from the flat orientation) grow past the LDS box: chunks up
    to twice its nodes are summed by one wavefront in both boxes' memory,
    larger ones spread with atomics.  One spreading stage of the same forces
    with the staged chunks (form 1) and by platelet (form 4, with the force
    launch's tile claims) against the direct form (2): the same field up to
    the FP32 order of the sums."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_THROMBUS)
    x, y, z = sc.x.copy(), sc.y.copy(), sc.z.copy()
    c, sn = np.cos(np.pi / 4), np.sin(np.pi / 4)
    st = sc.plt_elem_start
    for p in range(len(st) - 1):
        a, b = st[p], st[p + 1]
        cy, cz = y[a:b].mean(), z[a:b].mean()
        dy, dz = y[a:b] - cy, z[a:b] - cz
        y[a:b] = (cy + c * dy - sn * dz).astype(np.float32)
        z[a:b] = (cz + sn * dy + c * dz).astype(np.float32)
    rng = np.random.default_rng(17)
    f = (1e-3 * rng.standard_normal((3, sc.n_markers))).astype(np.float32)
    fields = {}
    for form in (2, 1, 4):
        with pfg.Solver(sc.params) as s:
            s.set_ib_staging(form)
            s.load(sc)
            s.set_markers(x, y, z)
            s.stage(pfg.STAGE_FORCES)  # the tile claims of these positions
            s.set_marker_forces(*f)
            s.stage(pfg.STAGE_SPREAD)
            fields[form] = s.aux()
    scale = np.abs(fields[2]).max()
    assert scale > 0
    for form in (1, 4):
        err = float(np.abs(fields[form] - fields[2]).max())
        print(f"form {form}: max difference {err:.2e} of {scale:.2e}")
        assert err < 1e-5 * scale


def test_force_pass_by_body_request_with_spreading_force_pass(gpu):
    """The force pass by body requested where the step's force pass spreads
    each marker's force itself (direct spreading, a wavefront per marker):
    that pass sums the Morse pairs from the cell lists, which must then hold
    the elements; the steps follow the same run without the request."""
    pfg = gpu
    sc = scene_small(pfg, receptors=-1, ligands=-1)
    sc.params.morse_u0 = 1.0
    runs = []
    for mode in (2, 1):
        with pfg.Solver(sc.params) as s:
            s.set_ib_staging(2)
            s.set_forces_lanes(64)
            s.set_forces_by_body(mode)
            s.load(sc)
            s.step(10)
            mk = s.markers()
            runs.append(np.stack([mk[k] for k in ("x", "y", "z")]))
    assert np.abs(runs[0] - runs[1]).max() <= 2 * np.spacing(np.float32(sc.params.nx))


@pytest.mark.parametrize("name", ["thrombus_small", "tube_full"])
def test_force_pass_by_body_follows_the_per_marker_pass(gpu, name):
    """Coupled steps with the force pass by platelet (Morse pairs from each
    platelet's neighbour lists, rebuilt when an element has moved half the
    skin away from its platelet's mean motion) against the per-marker pass
    over the global cell lists, on a scene with platelets, bonds and a
    fibrin network: the same markers and fields up to the FP32 order of the
    force sums, with the hard Morse cutoff; the near-cutoff pair visits the
    two passes count agree (both visit every pair within the cutoff from
    both ends; the trajectories differ in the last bits, which may move a
    pair across the edge of that window).  Midway the positions are
    uploaded again, which makes every platelet build its lists again."""
    pfg = gpu
    if name == "thrombus_small":
        sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=96, ny=64, nz=64, n_platelets=6, elements=300,
                       receptors=16, ligands=400, fibrin_nodes=800, fibrin_links=1200)
        steps = 30
    else:
        sc = pfg.Scene(pfg.SCENE_THROMBUS)
        steps = 4
    runs = []
    for mode in (2, 1):
        with pfg.Solver(sc.params) as s:
            s.set_forces_by_body(mode)
            s.load(sc)
            s.step(steps // 2)
            mk = s.markers()
            s.set_markers(mk["x"], mk["y"], mk["z"])
            info = s.step(steps - steps // 2)
            mk = s.markers()
            runs.append((np.stack([mk[k] for k in ("x", "y", "z")]), s.fields(), info))
    dx = float(np.abs(runs[0][0] - runs[1][0]).max())
    du = max(float(np.abs(a - b).max()) for a, b in zip(runs[0][1], runs[1][1]))
    print(f"{name}: {steps} steps, max position difference {dx:.2e}, max field difference "
          f"{du:.2e}, near-cutoff visits {runs[0][2]['near_cutoff']} / {runs[1][2]['near_cutoff']}")
    xmax = max(sc.params.nx, sc.params.ny, sc.params.nz)
    assert dx <= 4 * np.spacing(np.float32(xmax))
    assert du < 5e-6
    n0, n1 = runs[0][2]["near_cutoff"], runs[1][2]["near_cutoff"]
    assert abs(n0 - n1) <= max(2, n0 // 100)


@pytest.mark.parametrize("name", ["thrombus_small", "tube_full"])
def test_body_spreading_follows_the_staged_form(gpu, name):
    """Coupled steps with the body spreading (form 4: a workgroup per
    platelet, plain stores into the tiles the force pass claimed for it
    alone, atomics elsewhere; the fibrin nodes as in form 1) against the
    LDS-staged chunks (form 1) on the same scene, with platelets, receptors,
    bonds and a fibrin network (at full size the 512x256x256 flag tube with
    500 platelets and 20000 fibrin nodes): the same markers and fields up to
    the FP32 order of the force sums."""
    pfg = gpu
    if name == "thrombus_small":
        sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=96, ny=64, nz=64, n_platelets=6, elements=300,
                       receptors=16, ligands=400, fibrin_nodes=800, fibrin_links=1200)
        steps = 20
    else:
        sc = pfg.Scene(pfg.SCENE_THROMBUS)
        steps = 2
    sc.params.morse_taper = 0.2  # no pair switches at the cutoff on a rounding difference
    runs = []
    for form in (1, 4):
        with pfg.Solver(sc.params) as s:
            s.set_ib_staging(form)
            s.load(sc)
            s.step(steps)
            mk = s.markers()
            runs.append((np.stack([mk[k] for k in ("x", "y", "z")]), s.fields()))
    dx = float(np.abs(runs[0][0] - runs[1][0]).max())
    du = max(float(np.abs(a - b).max()) for a, b in zip(runs[0][1], runs[1][1]))
    print(f"{name}: {steps} steps, max position difference {dx:.2e}, max field difference "
          f"{du:.2e}")
    xmax = max(sc.params.nx, sc.params.ny, sc.params.nz)
    assert dx <= 2 * np.spacing(np.float32(xmax))
    assert du < 5e-6


@pytest.mark.parametrize("name", ["thrombus_small", "tube_full"])
def test_tile_gather_follows_the_staged_form(gpu, name):
    """Coupled steps with the tile-gather spreading (form 3) against the
    LDS-staged chunks (form 1) on the same scene: platelets with receptors
    and bonds and a fibrin network (and, at full size, the 512x256x256 flag
    tube with 500 platelets): the same markers and fields up to the FP32
    order of the force sums.  The gather sums in cell-list order, so two
    runs of form 3 are bitwise equal."""
    pfg = gpu
    if name == "thrombus_small":
        sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=96, ny=64, nz=64, n_platelets=6, elements=300,
                       receptors=16, ligands=400, fibrin_nodes=800, fibrin_links=1200)
        sc.params.morse_taper = 0.2
        steps = 20
    else:
        sc = pfg.Scene(pfg.SCENE_THROMBUS)
        sc.params.morse_taper = 0.2  # no pair switches at the cutoff on a rounding difference
        steps = 2
    runs = []
    for form in (1, 3, 3):
        with pfg.Solver(sc.params) as s:
            s.set_ib_staging(form)
            s.load(sc)
            s.step(steps)
            mk = s.markers()
            runs.append((np.stack([mk[k] for k in ("x", "y", "z")]), s.fields()))
    dx = float(np.abs(runs[0][0] - runs[1][0]).max())
    du = max(float(np.abs(a - b).max()) for a, b in zip(runs[0][1], runs[1][1]))
    print(f"{name}: {steps} steps, max position difference {dx:.2e}, max field difference "
          f"{du:.2e}")
    # positions: two FP32 ulps of the largest coordinate; fields: the force
    # sums' FP32 order carried through the steps
    xmax = max(sc.params.nx, sc.params.ny, sc.params.nz)
    assert dx <= 2 * np.spacing(np.float32(xmax))
    assert du < 5e-6
    assert np.array_equal(runs[1][0], runs[2][0])
    assert all(np.array_equal(a, b) for a, b in zip(runs[1][1], runs[2][1]))
