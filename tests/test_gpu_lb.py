"""Parity of the fused HIP LB kernel with the FP64 oracle, and full-size
properties (conservation, steady Couette) on the production configuration."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
Q = 19


def w_col(po):
    return np.array([po.load().orc_w(d) for d in range(Q)])[:, None]


def make_params(pfg, nx, ny, nz, walls, store=0, tau=0.8, g=(0, 0, 0), u_hi=None,
                use_flags=0):
    p = pfg.default_params(nx=nx, ny=ny, nz=nz, tau=tau, store=store, use_flags=use_flags)
    for a in range(3):
        p.wall[a] = walls[a]
        p.body_force[a] = g[a]
    if u_hi is not None:
        for k in range(3):
            p.wall_u_hi[2][k] = u_hi[k]
    return p


CASES = [
    dict(name="periodic", walls=(0, 0, 0), g=(1e-5, 0, 0)),
    dict(name="channel_yz", walls=(0, 1, 1), g=(1e-5, 2e-6, 0)),
    dict(name="couette_z", walls=(0, 0, 1), u_hi=(0.05, 0.01, 0)),
    dict(name="box_flags", walls=(1, 1, 1), flags=True),
    # inflow/outflow faces: duct profile over bounded x and y, one over y only
    dict(name="velocity_faces_duct", walls=(1, 1, 2), u_hi=(0.0, 0.0, 0.03),
         u_lo=(0.01, -0.005, 0.04)),
    dict(name="velocity_faces_slab", walls=(0, 1, 2), u_hi=(0.0, 0.0, 0.02),
         u_lo=(0.0, 0.0, 0.02), g=(0, 0, 1e-6)),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_single_step_parity(gpu, po, case):
    pfg = gpu
    nx, ny, nz = 70, 12, 10  # nx not a multiple of 64: exercises the edge guard
    N = nx * ny * nz
    rng = np.random.default_rng(11)
    flags = None
    if case.get("flags"):
        flags = (rng.random(N) < 0.1).astype(np.uint8)
    p = make_params(pfg, nx, ny, nz, case["walls"], g=case.get("g", (0, 0, 0)),
                    u_hi=case.get("u_hi"), use_flags=int(flags is not None))
    if case.get("u_lo") is not None:
        for k in range(3):
            p.wall_u_lo[2][k] = case["u_lo"][k]
    w = w_col(po)
    f = (w * (1 + 0.05 * rng.standard_normal((Q, N)))).astype(np.float32)
    with pfg.Solver(p) as s:
        if flags is not None:
            s.upload_flags(flags)
        s.set_f(f)
        s.set_fluid_only(True)
        s.step(1)
        fg = s.f().astype(np.float64)
        rho_g, ux_g, uy_g, uz_g = s.fields()
    L = po.lattice(p, flags)
    fo = po.lb_step(L, f.astype(np.float64))
    fluid = np.ones(N, bool) if flags is None else flags == 0
    d = np.abs((fg - w) - (fo - w))[:, fluid]
    assert d.max() < 2e-7, d.max()
    rho_o, ux_o, uy_o, uz_o = po.fields(L, fo)
    assert np.abs(ux_g[fluid] - ux_o[fluid]).max() < 1e-6
    assert np.abs(rho_g[fluid] - rho_o[fluid]).max() < 1e-6


@pytest.mark.parametrize("store,tol", [(0, 1e-5), (1, 1e-3)])
def test_channel_flow_multistep_l2(gpu, po, store, tol):
    """Body-force channel 64x32x32, tau 0.8, F = 1e-6 (the CPU-runnable
    configuration's flow), 300 steps: relative L2 velocity error against the
    FP64 oracle (FP32 storage < 1e-5, FP16 deviation storage < 1e-3)."""
    pfg = gpu
    p = make_params(pfg, 64, 32, 32, (0, 1, 1), store=store, g=(1e-6, 0, 0))
    steps = 300
    with pfg.Solver(p) as s:
        s.set_fluid_only(True)
        s.step(steps)
        _, ux, uy, uz = s.fields()
    L = po.lattice(p)
    f = po.init_eq(L)
    for _ in range(steps):
        f = po.lb_step(L, f)
    _, oux, ouy, ouz = po.fields(L, f)
    num = np.sqrt(((ux - oux) ** 2 + (uy - ouy) ** 2 + (uz - ouz) ** 2).sum())
    den = np.sqrt((oux ** 2 + ouy ** 2 + ouz ** 2).sum())
    print(f"store {store}: relative L2 velocity error {num / den:.3e}")
    assert num / den < tol


def test_bf16_storage_runs_and_conserves(gpu):
    pfg = gpu
    p = make_params(pfg, 64, 32, 32, (0, 1, 1), store=2, g=(1e-5, 0, 0))
    with pfg.Solver(p) as s:
        s.set_fluid_only(True)
        m0 = s.conservation()[0]
        s.step(50)
        m1 = s.conservation()[0]
    assert abs(m1 - m0) < 1e-2 * m0


def test_full_size_couette_conservation_and_profile(gpu):
    """256x128x128 Couette (wall velocity 0.05, tau 0.7): the linear profile is
    steady, mass is conserved in FP64 totals, momentum gain matches the wall
    shear exchange direction (x only)."""
    pfg = gpu
    sc = pfg.Scene(pfg.SCENE_SHEAR, n_platelets=-1, receptors=-1, ligands=-1)
    p = sc.params
    with pfg.Solver(p) as s:
        s.init_equilibrium(None, *sc.u0)
        s.set_fluid_only(True)
        c0 = s.conservation()
        s.step(200)
        c1 = s.conservation()
        _, ux, uy, uz = s.fields()
    N = p.nx * p.ny * p.nz
    # FP32 rounding of the collision: ~1e-8 relative drift over 200 steps
    assert abs(c1[0] - c0[0]) < 1e-7 * N
    prof = ux.reshape(p.nz, p.ny, p.nx).mean(axis=(1, 2))
    exact = 0.05 * (np.arange(p.nz) + 0.5) / p.nz
    assert np.abs(prof - exact).max() < 2e-5
    assert np.abs(uy).max() < 1e-5 and np.abs(uz).max() < 1e-5


def test_rest_state_is_fixed_point_full_size(gpu):
    pfg = gpu
    for store in (0, 1, 2):
        p = make_params(pfg, 256, 128, 128, (0, 1, 1), store=store)
        with pfg.Solver(p) as s:
            s.set_fluid_only(True)
            s.step(3)
            c = s.conservation()
        assert abs(c[0] - 256 * 128 * 128) < 1e-6 * 256 * 128 * 128
        assert np.abs(c[1:]).max() < 1e-6


def test_graph_and_eager_agree(gpu):
    pfg = gpu
    p = make_params(pfg, 64, 32, 32, (0, 1, 1), g=(1e-5, 0, 0))
    out = []
    for graph in (True, False):
        with pfg.Solver(p) as s:
            s.set_graph(graph)
            s.set_fluid_only(True)
            s.step(7)
            out.append(s.f())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("xcd", [0, 1])
@pytest.mark.parametrize("cpt", [1, 2, 4])
@pytest.mark.parametrize("store", [0, 1])
def test_launch_shapes_agree(gpu, cpt, store, xcd):
    """z-cells per lane and the workgroup order change only the launch, never
    the result (96 cells along x: one full and one partial 64-cell block)."""
    pfg = gpu
    p = make_params(pfg, 96, 20, 18, (0, 1, 1), g=(1e-5, 0, 0))
    p.store = store
    ref = None
    with pfg.Solver(p) as s:
        s.set_fluid_only(True)
        s.step(3)
        ref = s.f()
    with pfg.Solver(p) as s:
        s.set_lb_tuning(cpt)
        s.set_lb_xcd_rows(xcd)
        s.set_fluid_only(True)
        s.step(3)
        assert np.array_equal(s.f(), ref)


def test_autotune_leaves_the_lattice_unchanged(gpu):
    """The autotuner's trial steps are undone: after it, the lattice steps on
    exactly as an untuned one (cells per lane change only the launch)."""
    pfg = gpu
    p = make_params(pfg, 96, 20, 18, (0, 1, 1), g=(1e-5, 0, 0))
    with pfg.Solver(p) as s:
        s.set_fluid_only(True)
        s.step(5)
        ref = s.f()
    with pfg.Solver(p) as s:
        s.set_fluid_only(True)
        s.step(2)
        t = s.autotune(8)
        assert t["lb_cells_per_lane"] in (1, 2, 4)
        assert t["shapes_timed"] == 4 and t["ms_per_step"] > 0
        assert t["lb_xcd_rows"] in (0, 1)
        assert s.step(3)["step"] == 5
        assert np.array_equal(s.f(), ref)


@pytest.mark.parametrize("store,tol", [(0, 2e-6), (1, 2e-4)])
def test_velocity_faces_multistep(gpu, po, store, tol):
    """Inflow/outflow channel (bounce-back in y, periodic x), 300 steps from
    rest against the oracle, FP32 and FP16 storage."""
    pfg = gpu
    nx, ny, nz = 8, 16, 24
    p = make_params(pfg, nx, ny, nz, (0, 1, 2), store=store, tau=1.0, u_hi=(0, 0, 0.02))
    for k in range(3):
        p.wall_u_lo[2][k] = (0.0, 0.0, 0.02)[k]
    with pfg.Solver(p) as s:
        s.init_equilibrium()
        s.set_fluid_only(True)
        s.step(300)
        rho_g, ux_g, uy_g, uz_g = s.fields()
    L = po.lattice(p)
    f = po.init_eq(L)
    for _ in range(300):
        f = po.lb_step(L, f)
    rho_o, ux_o, uy_o, uz_o = po.fields(L, f)
    assert np.abs(uz_o).max() > 0.01
    assert np.abs(uz_g - uz_o).max() < tol
    assert np.abs(rho_g - rho_o).max() < 10 * tol
