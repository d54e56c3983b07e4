"""Host driver runs on the device: output files in the reference's formats,
frames equal to the same run through the Python binding, and the loop ending
at detach with the pause time appended."""
import os
import subprocess

import numpy as np
import pytest

from test_driver import DRIVER, KEYS, read_fixture, write_par

pytestmark = pytest.mark.gpu


def driver(tmp_path, vals, *extra):
    par = tmp_path / "parameter.dat"
    write_par(par, vals)
    cmd = [DRIVER, "-p", str(par), "-o", str(tmp_path), "--lattice", "48", "32", "24", *extra]
    print("driver:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return r.stdout


def small_vals(**kw):
    v = read_fixture()
    v.update(MAX_ELEM="200", SURF_ELEM="396", RECEPTOR="1584", OUT_FREQ="10",
             DELTA_T="1e-8", SIM_TIME="3e-7")
    v.update({k: str(x) for k, x in kw.items()})
    return v


def test_driver_frames_match_the_binding(gpu, tmp_path):
    """With the scene's own physics (--scene-physics), the driver's frames
    are the binding's run of the same scene."""
    pfg = gpu
    out = driver(tmp_path, small_vals(), "--scene-physics")
    od = tmp_path / "Output"
    n, frames = 200, 1 + 3  # initial frame + 30 steps / OUT_FREQ 10
    pos = np.fromfile(od / "cellsPos.data", dtype=np.float64)
    assert pos.size == frames * n * 4
    pos = pos.reshape(frames, n, 4)
    assert np.array_equal(pos[..., 2], pos[..., 3])  # z written twice
    assert (np.fromfile(od / "cellsVel.data", dtype=np.float64).size == n * 4)
    assert (np.fromfile(od / "cellsForce.data", dtype=np.float64).size == n * 4)
    for name in ("cellsPosASCII.dat", "cellsVelASCII.dat", "cellsForceASCII.dat"):
        lines = open(od / name).read().split("\n")
        assert lines[0] == "Timestep       ElemID        X         Y         Z"
        rows = [l.split() for l in lines[1:] if l.strip()]
        assert len(rows) == frames * n
        assert [int(r[0]) for r in rows[::n]] == [0, 10, 20, 30]
        assert [int(r[1]) for r in rows[:n]] == list(range(n))
    assert "Time Step = 30" in out
    asc = np.array([[float(x) for x in r[2:]] for r in
                    [l.split() for l in open(od / "cellsPosASCII.dat").read().split("\n")[1:]
                     if l.strip()]]).reshape(frames, n, 3)
    assert np.allclose(asc, pos[..., :3], rtol=1e-5, atol=1e-6)

    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=48, ny=32, nz=24, n_platelets=1, elements=200,
                   receptors=int(1584 / 396 / 2) * 396, seed=1018)
    assert np.allclose(pos[0, :, 0], sc.x) and np.allclose(pos[0, :, 1], sc.y)
    with pfg.Solver(sc.params) as s:
        s.load(sc)
        for k in range(1, frames):
            s.step(10)
            mk = s.markers()
            got = np.stack([mk["x"], mk["y"], mk["z"]], 1)
            assert np.abs(got - pos[k, :, :3]).max() < 1e-5, k


def test_driver_stops_at_detach(gpu, tmp_path):
    """Rates from the file with DELTA_T 0.01 s and SHEAR_RATE 30/s: detach
    hazard 0.3 per step and almost no bond formation, so the run ends at
    detach well inside 400 steps and appends the pause time."""
    vals = small_vals(DELTA_T="0.01", SIM_TIME="4", SHEAR_RATE="30")
    out = driver(tmp_path, vals, "--scene-physics", "--rates-from-file")
    od = tmp_path / "Output"
    line = open(od / "cellsPauseTime.dat").read().strip().split("\n")[-1]
    t_pause, seed = line.split(", ")
    assert int(seed) == 1018
    steps = round(float(t_pause) / 0.01)
    assert 0 <= steps < 400
    assert f"detached after {steps} steps" in out
    rows = [l.split() for l in open(od / "cellsPosASCII.dat").read().split("\n")[1:] if l.strip()]
    assert int(rows[-1][0]) == steps  # the last frame is the state at detach


def test_driver_file_physics_matches_the_binding(gpu, tmp_path):
    """With the physics of the parameter file (the default), the driver's
    frames are the binding's run of the same scene with the parameters
    mapped from the file's block as the driver documents (pfg_params_from_sem
    for the physics, the top wall at SHEAR_RATE x DELTA_T per lattice length,
    no Morse, the linear shear start)."""
    import ctypes as C
    import json
    pfg = gpu
    vals = small_vals()
    par = tmp_path / "parameter.dat"
    write_par(par, vals)
    mesh = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures",
                        "Platelet_Mesh_958.neu")
    info = json.loads(subprocess.run([DRIVER, "-p", str(par), "--lattice", "48", "32", "24",
                                      "--no-mesh", "--parse-only"], capture_output=True,
                                     text=True, timeout=120).stdout)
    assert info["physics"] == "file" and mesh
    driver(tmp_path, vals, "--no-mesh")
    od = tmp_path / "Output"
    n, frames = 200, 4
    pos = np.fromfile(od / "cellsPos.data", dtype=np.float64).reshape(frames, n, 4)
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=48, ny=32, nz=24, n_platelets=1, elements=200,
                   receptors=int(1584 / 396 / 2) * 396, seed=1018)
    p = sc.params
    q = pfg.Params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    sem = np.array(info["sem"], np.float32)
    dx, dt = float(vals["DELTA_X"]), float(vals["DELTA_T"])
    assert pfg.lib().pfg_params_from_sem(sem.ctypes.data_as(C.POINTER(C.c_float)), dx, dt,
                                        p.nx, p.ny, p.nz, C.byref(q)) == 0
    for k in ("tau", "ib_force_scale", "marker_dt", "k_spring", "ks", "kv", "kbend", "k_on",
              "k_off0", "detach_rate", "bond_law", "bond_k", "bond_l0", "bell_xb_kT", "bell_kT",
              "wall_rep_a", "wall_rep_kappa", "wall_rep_hc"):
        setattr(p, k, getattr(q, k))
    p.morse_u0 = 0.0
    p.seed = 1018
    u_top = float(vals["SHEAR_RATE"]) * dt * p.nz
    scale = u_top / p.wall_u_hi[2][0]
    p.wall_u_hi[2][0] = u_top
    sc.u0_z[0][:] *= scale
    assert np.isclose(p.tau, info["tau"], rtol=1e-6)
    with pfg.Solver(p) as s:
        s.load(sc)
        for k in range(1, frames):
            s.step(10)
            mk = s.markers()
            got = np.stack([mk["x"], mk["y"], mk["z"]], 1)
            err = np.abs(got - pos[k, :, :3]).max()
            assert err < 1e-5, (k, err)
    # the platelet rides the shear: it moved along x, by about u(z) per step
    assert np.isfinite(pos).all()
    assert pos[-1, :, 0].mean() > pos[0, :, 0].mean()
