"""Host driver (platelet_fiber_gpu_amd/pfg_driver): parameter file, derived
parameter block and error behaviour, on the CPU.

The parameter file is the reference's own Input/parameter.dat (from its
gpu_20151203 tarball), kept as a data fixture."""
import json
import math
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "platelet_fiber_gpu_amd", "pfg_driver")
FIXTURE = os.path.join(ROOT, "tests", "fixtures", "parameter.dat")

KEYS = ["PDB_PATH", "PSF_PATH", "NEU_PATH", "DELTA_X", "DELTA_T", "SIM_TIME", "OUT_FREQ",
        "MAX_CELL", "MAX_ELEM", "SURF_ELEM", "RECEPTOR", "RANDOM_SEED", "K_SPRING", "Kr0",
        "SHEAR_RATE", "INIT_RHO"]


def run(*args, cwd=None):
    if not os.path.exists(DRIVER):
        pytest.fail("pfg_driver is not built (make -C platelet_fiber_gpu_amd)")
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=cwd, timeout=120)


def read_fixture():
    vals = {}
    for line in open(FIXTURE):
        k, v = line.split()
        vals[k] = v
    return vals


def write_par(path, vals, order=KEYS):
    with open(path, "w") as f:
        for k in order:
            f.write(f"{k}      {vals[k]}\n")


def f32(v):
    import numpy as np
    return float(np.float32(v))


def test_fixture_keyword_order():
    assert list(read_fixture()) == KEYS


def test_parse_only_fills_the_reference_parameter_block():
    """gpuSEM_LB.c:1206-1233 with the reference's parameter file: each entry
    from its definition, on the 256x128x128 shear lattice (lz = 128) and the
    generated 958-element mesh (Lm from the driver's own mesh, checked through
    the two entries that use it)."""
    r = run("-p", FIXTURE, "--parse-only")
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout)
    v = {k: float(x) for k, x in read_fixture().items() if k not in KEYS[:3]}
    sem = out["sem"]
    shear, rho0, dx, dt, lz = v["SHEAR_RATE"], v["INIT_RHO"], v["DELTA_X"], v["DELTA_T"], 128
    umax = shear * (lz - 1) * dx
    nu = 1.2e6 / rho0
    expect = {0: v["RECEPTOR"], 1: v["RANDOM_SEED"], 2: v["K_SPRING"],
              3: 0.0165 * math.exp(0.00065 * shear), 4: v["Kr0"], 5: shear, 6: rho0,
              7: 1.0, 8: 1.0 / umax, 9: rho0, 10: 1.0, 11: umax, 12: nu,
              13: 3.0 * f32(nu) * dt / dx ** 2 + 0.5, 14: rho0 * umax ** 2, 15: dx / dt / umax,
              16: 1e12, 19: 500 * 1.3806488e-8 * 300 * 1e-6}
    for k, e in expect.items():
        assert sem[k] == pytest.approx(e, rel=2e-6), k
    # [17] = 600 kT / 2 / Lm^2 and [18] = [17] / Lm share one mean edge length
    kt = 1.3806488e-8 * 300
    lm = sem[17] / sem[18]
    assert sem[17] == pytest.approx(600 * kt / 2 / lm ** 2, rel=1e-5)
    assert 0.05 < lm < 2.0  # microns: a lattice-unit edge length times dx
    # counts: MAX_CELL platelets of MAX_ELEM elements, RECEPTOR/SURF_ELEM/2
    # (truncated) receptors per triangle over 2n-4 triangles
    assert out["platelets"] == 1 and out["markers"] == 958
    assert out["receptors"] == int(10688 / 1912 / 2) * (2 * 958 - 4)
    assert out["max_steps"] == 1 and out["out_freq"] == 1000
    assert out["seed"] == 1018 and out["seeds"] == [1018, 1018]


def test_rates_from_file():
    r = run("-p", FIXTURE, "--parse-only", "--rates-from-file")
    out = json.loads(r.stdout)
    sem = out["sem"]
    assert out["k_on"] == pytest.approx(sem[3] * 1e-8, rel=1e-6)
    assert out["k_off0"] == pytest.approx(sem[4] * 1e-8, rel=1e-6)
    assert out["detach_rate"] == pytest.approx(300 * 1e-8, rel=1e-6)


def test_keyword_out_of_order_is_an_error(tmp_path):
    vals = read_fixture()
    order = KEYS.copy()
    order[3], order[4] = order[4], order[3]  # DELTA_T before DELTA_X
    par = tmp_path / "parameter.dat"
    write_par(par, vals, order)
    r = run("-p", str(par), "--parse-only")
    assert r.returncode == 1
    assert "Parameter Read Failed for DELTA_X" in r.stdout


def test_missing_file_and_bad_arguments(tmp_path):
    r = run("-p", str(tmp_path / "nope.dat"), "--parse-only")
    assert r.returncode == 1 and "Parameter File was not open" in r.stdout
    r = run("-p", FIXTURE, "--bogus")
    assert r.returncode == 1 and "unknown argument" in r.stderr
    r = run("-p", FIXTURE, "--seeds", "5", "2")
    assert r.returncode == 1


def test_mesh_from_parameter_file():
    """NEU_PATH (Input/Platelet_Mesh_958.neu) resolves next to the parameter
    file: 958 elements, 1912 triangles and RECEPTOR / SURF_ELEM / 2 = 2
    receptors per triangle (gpuSEM_LB.c:723), as the reference sets up."""
    r = run("-p", FIXTURE, "--parse-only")
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout)
    assert out["mesh"].endswith("Platelet_Mesh_958.neu")
    assert out["markers"] == 958 and out["triangles"] == 1912
    assert out["receptors"] == 2 * 1912
    r = run("-p", FIXTURE, "--parse-only", "--no-mesh")
    assert json.loads(r.stdout)["mesh"] == ""


def test_mesh_missing_file_is_an_io_error(tmp_path):
    r = run("-p", FIXTURE, "--parse-only", "--mesh", str(tmp_path / "none.neu"))
    assert r.returncode == 2
    assert "cannot read mesh" in r.stderr


def test_fibrin_network_from_file(tmp_path):
    """--fibrin places the reference's network file (1422 links) in the scene."""
    net = os.path.join(ROOT, "tests", "fixtures", "HemoThickInputCor.txt")
    r = run("-p", FIXTURE, "--parse-only", "--scene", "thrombus", "--lattice", "128", "96",
            "96", "--fibrin", net)
    assert r.returncode == 0, r.stdout + r.stderr
    assert json.loads(r.stdout)["links"] == 1422
    r = run("-p", FIXTURE, "--parse-only", "--fibrin", str(tmp_path / "none.txt"))
    assert r.returncode == 2


def test_physics_follows_the_parameter_file(tmp_path):
    """The run's physics comes from the file's block (pfg_params_from_sem):
    a changed parameter.dat changes the handle's parameters accordingly --
    the spring stiffness with K_SPRING, tau with INIT_RHO (nu = 1.2e6 /
    rho0, gpuSEM_LB.c:1220-1221), the top wall and the detach hazard with
    SHEAR_RATE, the IB force scale with both -- while --scene-physics keeps
    the scene's lattice-unit physics."""
    base = json.loads(run("-p", FIXTURE, "--parse-only").stdout)
    v = read_fixture()
    dx, dt = float(v["DELTA_X"]), float(v["DELTA_T"])
    assert base["physics"] == "file"
    assert base["tau"] == pytest.approx(base["sem"][13], rel=1e-6)
    # a marker at lattice velocity u moves u lattice lengths per step
    assert base["marker_dt"] == pytest.approx(1.0, rel=1e-6)
    assert base["k_spring"] == pytest.approx(float(v["K_SPRING"]) * dx, rel=1e-6)
    assert base["wall_u_top"] == pytest.approx(float(v["SHEAR_RATE"]) * dt * 128, rel=1e-6)
    assert base["morse_u0"] == 0.0 and base["bond_law"] == 1

    v2 = dict(v)
    v2["K_SPRING"] = "2.5e3"
    v2["INIT_RHO"] = "2.0"
    v2["SHEAR_RATE"] = "6e2"
    par = tmp_path / "parameter.dat"
    write_par(par, v2)
    # the mesh path of the file resolves relative to the file: keep the
    # fixture's mesh
    mesh = os.path.join(ROOT, "tests", "fixtures", "Platelet_Mesh_958.neu")
    new = json.loads(run("-p", str(par), "--parse-only", "--mesh", mesh).stdout)
    assert new["k_spring"] == pytest.approx(2.5e3 * dx, rel=1e-6)
    nu = 1.2e6 / 2.0
    assert new["tau"] == pytest.approx(3 * nu * dt / dx ** 2 + 0.5, rel=1e-5)
    assert new["tau"] != pytest.approx(base["tau"], rel=1e-3)
    assert new["wall_u_top"] == pytest.approx(600 * dt * 128, rel=1e-6)
    assert new["detach_rate"] == pytest.approx(600 * dt, rel=1e-6)
    # F0 UNIT_FACTOR dt / (rho0 UMAX C) = F0 UF dt^2 / (rho0 dx): rho0 only
    assert new["ib_force_scale"] == pytest.approx(base["ib_force_scale"] * 1.0239 / 2.0,
                                                  rel=1e-5)
    # the scene's own physics on request
    sc = json.loads(run("-p", FIXTURE, "--parse-only", "--scene-physics").stdout)
    assert sc["physics"] == "scene"
    assert sc["tau"] == pytest.approx(0.7) and sc["wall_u_top"] == pytest.approx(0.05)
