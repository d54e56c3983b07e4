"""The C-ABI library loads without a GPU, exports every entry point declared
in include/*.h, and its structs have the sizes the ctypes bindings assume.
Also: host-only entry points (parameters, scenes) behave and reject bad input."""
import ctypes as C
import glob
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(pfg_[a-z0-9_]+|(?:fluid|sem|fiber)_[A-Za-z_]+GPUKernel"
                             r"[A-Za-z_]*)\s*\(", src):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported_and_bound(pfg):
    from platelet_fiber_gpu_amd import _ffi
    lib = pfg.lib()
    names = declared_functions()
    assert len(names) > 40
    # the reference's own entry points (semLB.cu:111-1005), all thirteen
    ref = {n for n in names if not n.startswith("pfg_")}
    assert len(ref) == 13, sorted(ref)
    for n in sorted(names):
        assert hasattr(lib, n), n
        assert n in _ffi.SIGNATURES, n


def test_struct_sizes_match(pfg):
    from platelet_fiber_gpu_amd import _ffi
    lib = pfg.lib()
    assert lib.pfg_struct_size(0) == C.sizeof(_ffi.Params)
    assert lib.pfg_struct_size(1) == C.sizeof(_ffi.Model)
    assert lib.pfg_struct_size(2) == C.sizeof(_ffi.StepInfo)
    assert lib.pfg_struct_size(3) == C.sizeof(_ffi.SceneOpts)
    assert lib.pfg_struct_size(4) == C.sizeof(_ffi.Tuning)
    assert lib.pfg_struct_size(99) == 0


def test_status_strings_and_errors(pfg):
    lib = pfg.lib()
    assert lib.pfg_status_string(0) == b"ok"
    assert lib.pfg_status_string(1) == b"invalid argument"
    # null arguments are rejected, never dereferenced
    assert lib.pfg_create(None, 0, None) == 1
    assert lib.pfg_destroy(None) == 1
    assert lib.pfg_step(None, 1, None) == 1


def test_create_without_device_reports_device_error(pfg):
    if pfg.device_count() > 0:
        return
    p = pfg.default_params()
    h = C.c_void_p()
    assert pfg.lib().pfg_create(C.byref(p), 0, C.byref(h)) == 3
    p.tau = 0.4  # invalid relaxation time is an argument error before any device call
    assert pfg.lib().pfg_create(C.byref(p), 0, C.byref(h)) == 1


def test_params_from_reference_block(pfg):
    """hostSEMparameters[0..19] layout of gpuSEM_LB.c:1206-1233 and the
    unit mapping: the device, in lattice lengths with forces in F0, must
    give the reference's forces and displacements."""
    s = np.zeros(20, np.float32)
    dx, dt = 0.2, 1e-8
    rate, umax = 1000.0, 1000.0 * 20 * 0.2
    C_ = dx / dt / umax
    s[1], s[2], s[3], s[4], s[5] = 1101, 0.5, 0.02, 3.21, rate
    s[6], s[7], s[8], s[10], s[11], s[13] = 1.0, 1.0, 1.0 / umax, 1.0, umax, 1.379
    s[15], s[16], s[17], s[18], s[19] = C_, 1e12, 0.1, 0.2, 2e-4
    p = pfg.default_params()
    rc = pfg.lib().pfg_params_from_sem(s.ctypes.data_as(C.POINTER(C.c_float)), dx, dt,
                                      81, 301, 21, C.byref(p))
    assert rc == 0
    assert (p.nx, p.ny, p.nz) == (81, 301, 21)
    assert p.seed == 1101 and np.isclose(p.tau, 1.379)
    assert list(np.array(p.sem)) == list(s)
    # a spring stretched by 0.3 um: the reference's force k (r - L) / F0
    # (sem_kernel.cu:374-389) from lattice lengths
    assert np.isclose(p.k_spring * (0.3 / dx), 0.5 * 0.3, rtol=1e-6)
    # area and volume stiffness carry one and two more lengths
    assert np.isclose(p.ks / dx, 0.1, rtol=1e-6) and np.isclose(p.kv / dx ** 2, 0.2, rtol=1e-6)
    # bending: the mean-curvature Laplacian loses three lengths
    assert np.isclose(p.kbend * dx ** 3, 2e-4, rtol=1e-5)
    # a marker at lattice velocity u moves u * marker_dt lattice lengths per
    # step; in the reference V = C u (UMAX units) moves V dt / T0 * L0 um
    # (sem_kernel.cu:2111-2137)
    u = 1e-3
    assert np.isclose(u * p.marker_dt * dx, C_ * u * dt / s[8] * 1.0, rtol=1e-5)
    # rates per second become per step
    assert np.isclose(p.k_on, 0.02 * dt) and np.isclose(p.k_off0, 3.21 * dt)
    assert np.isclose(p.detach_rate, rate * dt)
    # the vWF bond of sem_kernel.cu:79-85: 10 nN/um over 0.078 um
    assert p.bond_law == pfg.BOND_REFERENCE
    assert np.isclose(p.bond_k * (0.1 / dx), 10.0 * 0.1, rtol=1e-6)
    assert np.isclose(p.bond_l0 * dx, 0.078, rtol=1e-6)
    # the wall repulsion a kappa e^{-kappa h}/(1 - e^{-kappa h}) at h = 0.01 um
    h = 0.01
    ref = 5e-4 * 2000 * np.exp(-2000 * h) / (1 - np.exp(-2000 * h))
    k_ = p.wall_rep_kappa
    dev = p.wall_rep_a * k_ * np.exp(-k_ * h / dx) / (1 - np.exp(-k_ * h / dx))
    assert np.isclose(dev, ref, rtol=1e-5) and np.isclose(p.wall_rep_hc * dx, 0.02, rtol=1e-6)
    # the IB force scale: F0 UNIT_FACTOR dt / (rho0 UMAX) / C (LB_kernel.cu:1158-1163)
    assert np.isclose(p.ib_force_scale, 1.0 * 1e12 * dt / (1.0 * umax) / C_, rtol=1e-5)


def test_scenes_are_seeded_and_well_formed(pfg):
    a = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=40,
                  ligands=300)
    b = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=40,
                  ligands=300)
    c = pfg.Scene(pfg.SCENE_SHEAR, nx=64, ny=32, nz=32, elements=256, receptors=40,
                  ligands=300, seed=7)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.lig_x, b.lig_x)
    assert not np.array_equal(a.x, c.x)
    assert a.rec_tri.size == 40 and a.lig_x.size == 300
    assert a.z.min() > a.params.wall_z  # above the ligand wall
    assert np.all((a.rec_r1 >= 0) & (a.rec_r1 <= 1))
    t = pfg.Scene(pfg.SCENE_THROMBUS, nx=64, ny=48, nz=48, n_platelets=3, elements=100,
                  receptors=8, ligands=100, fibrin_nodes=400, fibrin_links=600)
    assert t.flags is not None and t.flags.sum() > 0
    assert t.fluid_cells == int((t.flags == 0).sum())
    assert (t.kind == pfg.MARKER_ELEMENT).sum() == 300
    assert ((t.kind == pfg.MARKER_FIBER) | (t.kind == pfg.MARKER_FIXED)).sum() == 400
    assert t.link.size // 2 <= 600 and t.link.size > 0
    # every platelet centre sits inside the tube, clear of the wall
    ny, nz = 48, 48
    for q in range(3):
        e0, e1 = t.plt_elem_start[q], t.plt_elem_start[q + 1]
        cy, cz = t.y[e0:e1].mean() - (ny - 1) / 2, t.z[e0:e1].mean() - (nz - 1) / 2
        assert np.hypot(cy, cz) < t.params.body_force[0] * 0 + 120 * 48 / 256
