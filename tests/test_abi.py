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
        for m in re.finditer(r"\b(pfg_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported_and_bound(pfg):
    from platelet_fiber_gpu_amd import _ffi
    lib = pfg.lib()
    names = declared_functions()
    assert len(names) > 40
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
    """hostSEMparameters[0..19] layout of gpuSEM_LB.c:1206-1233."""
    s = np.zeros(20, np.float32)
    s[1], s[2], s[3], s[4], s[5] = 1101, 0.5, 0.02, 3.21, 100.0
    s[6], s[8], s[10], s[11], s[13], s[15], s[16] = 1.0, 1.0, 1.0, 2.0, 1.379, 1.0, 1.0
    s[17], s[18] = 0.1, 0.2
    p = pfg.default_params()
    rc = pfg.lib().pfg_params_from_sem(s.ctypes.data_as(C.POINTER(C.c_float)), 0.2, 1e-2,
                                      81, 301, 21, C.byref(p))
    assert rc == 0
    assert (p.nx, p.ny, p.nz) == (81, 301, 21)
    assert p.seed == 1101 and np.isclose(p.tau, 1.379) and np.isclose(p.k_spring, 0.5)
    assert np.isclose(p.ks, 0.1) and np.isclose(p.kv, 0.2)
    assert np.isclose(p.k_on, 0.02 * 1e-2) and np.isclose(p.k_off0, 3.21 * 1e-2)
    assert np.isclose(p.detach_rate, 100.0 * 1e-2)
    assert list(np.array(p.sem)) == list(s)


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
