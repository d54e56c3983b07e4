"""Batches (pfg_batch_*): every lattice of a batch advances exactly as the
same lattice in a handle of its own, which the single-lattice parity tests
tie to the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def lattice_params(pfg, nx, ny, nz, walls, g=(0.0, 0.0, 0.0), tau=0.8, u_top=0.0, store=0):
    p = pfg.default_params(nx=nx, ny=ny, nz=nz, tau=tau, store=store)
    for k in range(3):
        p.wall[k] = walls[k]
    for k in range(3):
        p.body_force[k] = g[k]
    p.wall_u_hi[2][0] = u_top
    return p


def fluid_set(pfg, store):
    P, B = pfg.WALL_PERIODIC, pfg.WALL_BOUNCE
    return [lattice_params(pfg, 40, 24, 20, (P, P, P), g=(2e-5, 0, 1e-5), store=store),
            lattice_params(pfg, 70, 16, 18, (P, P, B), tau=0.7, u_top=0.04, store=store),
            lattice_params(pfg, 130, 9, 5, (P, B, B), g=(1e-5, 0, 0), tau=1.1, store=store)]


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("cpt", [1, 2, 4])
@pytest.mark.parametrize("store", [0, 1])
def test_batch_fluid_equals_single(gpu, store, cpt, persistent):
    """Three lattices of different shapes, walls, tau and forcing: the batched
    LB launch gives bitwise the fields of three separate handles, with one,
    two or four z-cells per lane, as one workgroup per x-block row or as the
    persistent launch that takes the rows from a work queue (25 launches in
    a row: the queue resets itself after each)."""
    pfg = gpu
    ps = fluid_set(pfg, store)
    steps = 25
    with pfg.Batch(ps) as b:
        b.set_fluid_only(True)
        b.set_lb_tuning(cpt)
        b.set_lb_persistent(persistent)
        b.step(steps)
        got = [b.fields_of(k) for k in range(len(ps))]
        cons = [b.conservation_of(k) for k in range(len(ps))]
    for k, p in enumerate(ps):
        with pfg.Solver(p) as s:
            s.set_fluid_only(True)
            s.step(steps)
            want = s.fields()
            c = s.conservation()
        for a, w in zip(got[k], want):
            assert np.array_equal(a, w), k
        assert np.array_equal(cons[k], c), k


def test_batch_single_lattice_calls_are_refused(gpu):
    pfg = gpu
    ps = fluid_set(pfg, 0)
    with pfg.Batch(ps) as b:
        with pytest.raises(pfg.PfgError) as e:
            b.fields()
        assert e.value.code == 5  # PFG_ERR_STATE
        with pytest.raises(pfg.PfgError):
            b.init_equilibrium()
    with pfg.Solver(ps[0]) as s:
        with pytest.raises(pfg.PfgError) as e:
            s._c("pfg_set_lb_persistent", 1)
        assert e.value.code == 5


@pytest.mark.parametrize("graph", [0, 1])
def test_batch_persistent_coupled_equals_per_row_launch(gpu, graph):
    """Coupled shear lattices, eagerly and as captured steps replayed with
    host reads in between: the persistent queue launch gives the markers and
    fields of the one-workgroup-per-row launch, up to the FP32 order of the
    force-field atomics."""
    pfg = gpu
    scenes = [pfg.Scene(pfg.SCENE_BATCH, nx=48, ny=32, nz=24, elements=128, receptors=-1,
                        ligands=-1, seed=11),
              pfg.Scene(pfg.SCENE_BATCH, nx=40, ny=40, nz=28, n_platelets=2, elements=96,
                        receptors=-1, ligands=-1, seed=12)]
    runs = []
    for persistent in (0, 1):
        with pfg.Batch([s.params for s in scenes]) as b:
            b.set_graph(graph)
            b.set_lb_persistent(persistent)
            b.load_scenes(scenes)
            got = []
            for _ in range(4):
                b.step(3)
                got.append([(b.markers_of(k), b.fields_of(k)) for k in range(len(scenes))])
            runs.append(got)
    for r in range(4):
        for k in range(len(scenes)):
            for name in ("x", "y", "z", "vx", "vy", "vz"):
                d = np.abs(runs[0][r][k][0][name] - runs[1][r][k][0][name]).max()
                assert d < 1e-5, (r, k, name)
            for a, w in zip(runs[0][r][k][1], runs[1][r][k][1]):
                assert np.abs(a - w).max() < 1e-6, (r, k)


@pytest.mark.parametrize("staging", [1, 2, 3, 4])
def test_batch_coupled_equals_single(gpu, staging):
    """Shear lattices of different sizes with one or two platelets each (no
    receptors: the bond RNG is keyed by the batch-wide receptor index): each
    lattice's markers and fields follow its separate run."""
    pfg = gpu
    scenes = [pfg.Scene(pfg.SCENE_BATCH, nx=48, ny=32, nz=24, elements=128, receptors=-1,
                        ligands=-1, seed=11),
              pfg.Scene(pfg.SCENE_BATCH, nx=40, ny=40, nz=28, n_platelets=2, elements=96,
                        receptors=-1, ligands=-1, seed=12),
              pfg.Scene(pfg.SCENE_BATCH, nx=64, ny=24, nz=20, elements=160, receptors=-1,
                        ligands=-1, seed=13)]
    steps = 12
    with pfg.Batch([s.params for s in scenes]) as b:
        b.set_ib_staging(staging)
        b.load_scenes(scenes)
        b.step(steps)
        got = [(b.markers_of(k), b.fields_of(k)) for k in range(len(scenes))]
    for k, sc in enumerate(scenes):
        with pfg.Solver(sc.params) as s:
            s.set_ib_staging(staging)
            s.load(sc)
            s.step(steps)
            mk = s.markers()
            fl = s.fields()
        for name in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.abs(got[k][0][name] - mk[name]).max() < 1e-5, (k, name)
        for a, w in zip(got[k][1], fl):
            assert np.abs(a - w).max() < 1e-6, k


def test_batch_autotune_restores_the_state(gpu):
    """The autotuner on a batch (it also rebuilds the batched LB launch for
    each cells-per-lane shape) restores the lattices and markers it stepped:
    the batch continues as an untuned one, up to the FP32 order of the
    force-field atomics."""
    pfg = gpu
    scenes = [pfg.Scene(pfg.SCENE_BATCH, nx=48, ny=32, nz=24, elements=128, receptors=-1,
                        ligands=-1, seed=11),
              pfg.Scene(pfg.SCENE_BATCH, nx=64, ny=24, nz=20, elements=160, receptors=-1,
                        ligands=-1, seed=13)]
    runs = []
    for tune in (False, True):
        with pfg.Batch([s.params for s in scenes]) as b:
            b.load_scenes(scenes)
            b.step(6)
            if tune:
                t = b.autotune(4)
                assert t["shapes_timed"] >= 7 and t["lb_cells_per_lane"] in (1, 2, 4)
            info = b.step(6)
            assert info["step"] == 12
            runs.append([(b.markers_of(k), b.fields_of(k)) for k in range(len(scenes))])
    for k in range(len(scenes)):
        for name in ("x", "y", "z"):
            assert np.abs(runs[0][k][0][name] - runs[1][k][0][name]).max() < 1e-5, (k, name)
        for a, w in zip(runs[0][k][1], runs[1][k][1]):
            assert np.abs(a - w).max() < 1e-6, k


def test_full_batch_workload_equals_separate_handles(gpu):
    """The benchmark's batch at full size (256 vessel segments of uneven
    size and shear, 1-8 platelets of 256 elements each), 4 coupled steps:
    every lattice's markers and fields follow the same lattice stepped in a
    handle of its own."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    pfg = gpu
    scenes = bench.batch_scenes(pfg)
    assert len(scenes) == 256
    steps = 4
    with pfg.Batch([sc.params for sc in scenes]) as b:
        b.load_scenes(scenes)
        b.step(steps)
        got = [(b.markers_of(k), b.fields_of(k)) for k in range(len(scenes))]
    worst_x = worst_u = 0.0
    for k, sc in enumerate(scenes):
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            s.step(steps)
            mk = s.markers()
            fl = s.fields()
        for name in ("x", "y", "z"):
            worst_x = max(worst_x, float(np.abs(got[k][0][name] - mk[name]).max()))
        for a, w in zip(got[k][1], fl):
            worst_u = max(worst_u, float(np.abs(a - w).max()))
    print(f"full batch, {steps} steps: max position difference {worst_x:.2e}, "
          f"max field difference {worst_u:.2e}")
    assert worst_x < 3e-5  # coordinates up to 96: a few FP32 ulps
    assert worst_u < 1e-6


def test_batch_bonds_follow_single_handles(gpu):
    """Shear lattices with receptors over wall ligands and fast binding: each
    lattice of the batch keys its bond draws by its own seed and its own
    receptor indices, so it forms and breaks the bonds it does in a handle
    of its own (decisions within the near-threshold band counted on both
    sides), and its markers follow."""
    pfg = gpu
    specs = [(48, 32, 24, 11), (56, 40, 24, 12), (40, 32, 28, 13)]
    scenes = []
    for nx, ny, nz, seed in specs:
        sc = pfg.Scene(pfg.SCENE_SHEAR, nx=nx, ny=ny, nz=nz, elements=256, receptors=256,
                       ligands=2000, seed=seed)
        sc.params.k_on, sc.params.k_off0, sc.params.bond_reach = 0.3, 0.1, 1.5
        scenes.append(sc)
    steps = 8
    with pfg.Batch([s.params for s in scenes]) as b:
        b.load_scenes(scenes)
        info_b = b.step(steps)
        rb_all = b.bonds()[0]
        got = [b.markers_of(k) for k in range(len(scenes))]
        nrec, nlig = b.n_rec_k, b.n_wall_lig_k
    r0 = np.concatenate([[0], np.cumsum(nrec)])
    l0 = np.concatenate([[0], np.cumsum(nlig)])
    mism = near = formed = 0
    for k, sc in enumerate(scenes):
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            info = s.step(steps)
            rb = s.bonds()[0]
            mk = s.markers()
        mine = rb_all[r0[k]:r0[k + 1]]
        mine = np.where(mine >= 0, mine - l0[k], -1)
        mism += int((mine != rb).sum())
        near += info["near_threshold"]
        formed += int((rb >= 0).sum())
        for name in ("x", "y", "z"):
            assert np.abs(got[k][name] - mk[name]).max() < 1e-4, (k, name)
    print(f"batch bonds: {formed} bonds held in the separate runs, {mism} differ "
          f"(near threshold: separate {near}, batch {info_b['near_threshold']})")
    assert formed > 20
    assert mism <= near + info_b["near_threshold"]
