"""ctypes view of the C ABI in include/pfg.h and include/pfg_geometry.h.

The structures below mirror the C declarations field for field; the tests
check ``ctypes.sizeof`` against the C ``sizeof`` reported by the library.
Loading fails loudly when libpfg.so has not been built: there is no fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfg.so")

PFG_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "allocation failed", 3: "no usable HIP device",
          4: "kernel launch or HIP runtime error", 5: "call out of order", 6: "file I/O error"}

STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
WALL_PERIODIC, WALL_BOUNCE, WALL_VELOCITY = 0, 1, 2
FIBER_WLC, FIBER_LINEAR = 0, 1
FIBER_IB, FIBER_KICK = 0, 1
SEM_ADVECT, SEM_BBK = 0, 1
BOND_BELL, BOND_REFERENCE = 0, 1
MARKER_ELEMENT, MARKER_FIBER, MARKER_FIXED = 0, 1, 2
(STAGE_CELLS, STAGE_MEMBRANE, STAGE_BONDS, STAGE_FORCES, STAGE_SPREAD, STAGE_LB,
 STAGE_INTERP, STAGE_UPDATE, STAGE_CLEAR) = range(9)
SCENE_CHANNEL, SCENE_SHEAR, SCENE_THROMBUS, SCENE_LARGE, SCENE_BATCH = 1, 2, 3, 4, 5

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("store", C.c_int32), ("f16_scale_log2", C.c_int32),
        ("wall", C.c_int32 * 3),
        ("wall_u_lo", (C.c_float * 3) * 3), ("wall_u_hi", (C.c_float * 3) * 3),
        ("tau", C.c_float), ("body_force", C.c_float * 3), ("use_flags", C.c_int32),
        ("ib_force_scale", C.c_float), ("marker_dt", C.c_float),
        ("morse_u0", C.c_float), ("morse_rho", C.c_float), ("morse_req", C.c_float),
        ("morse_cutoff", C.c_float), ("morse_taper", C.c_float), ("k_spring", C.c_float), ("ks", C.c_float),
        ("kv", C.c_float), ("wall_rep_a", C.c_float), ("wall_rep_kappa", C.c_float),
        ("wall_rep_hc", C.c_float), ("wall_z", C.c_float),
        ("k_on", C.c_float), ("k_off0", C.c_float), ("bond_k", C.c_float),
        ("bond_l0", C.c_float), ("bond_reach", C.c_float), ("bell_xb_kT", C.c_float),
        ("detach_rate", C.c_float), ("seed", C.c_uint64),
        ("fiber_law", C.c_int32), ("fiber_move", C.c_int32), ("fiber_k", C.c_float),
        ("fiber_xmax", C.c_float), ("fiber_gamma", C.c_float),
        ("fiber_ref_sign", C.c_int32), ("bin_cells", C.c_int32),
        ("sem", C.c_float * 20), ("kbend", C.c_float),
        ("contact_a", C.c_float), ("contact_kappa", C.c_float), ("contact_hc", C.c_float),
        ("kick_temp", C.c_float),
        ("sem_integrator", C.c_int32), ("sem_mass", C.c_float), ("sem_gamma", C.c_float),
        ("sem_temp", C.c_float), ("sem_fluid_weight", C.c_float),
        ("bond_law", C.c_int32), ("bell_kT", C.c_float),
    ]


class Model(C.Structure):
    _fields_ = [
        ("n_markers", C.c_int32), ("x", f32p), ("y", f32p), ("z", f32p),
        ("kind", i32p), ("body", i32p),
        ("n_platelets", C.c_int32), ("plt_elem_start", i32p), ("plt_tri_start", i32p),
        ("n_tri", C.c_int32), ("tri", i32p), ("tri_S0", f64p), ("plt_S0", f64p),
        ("plt_V0", f64p),
        ("n_edges", C.c_int32), ("edge", i32p), ("edge_l0", f32p),
        ("n_receptors", C.c_int32), ("rec_tri", i32p), ("rec_r1", f32p), ("rec_r2", f32p),
        ("n_wall_ligands", C.c_int32), ("lig_x", f32p), ("lig_y", f32p), ("lig_z", f32p),
        ("fibrin_ligands", C.c_int32),
        ("n_links", C.c_int32), ("link", i32p), ("link_l0", f32p),
    ]


class StepInfo(C.Structure):
    _fields_ = [("step", C.c_int64), ("total_bonds", C.c_int32), ("detached", C.c_int32),
                ("detach_step", C.c_int64), ("near_threshold", C.c_int64),
                ("bond_events", C.c_int64), ("near_cutoff", C.c_int64),
                ("contact_overflow", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("lb_cells_per_lane", C.c_int32), ("lb_xcd_rows", C.c_int32),
                ("forces_lanes", C.c_int32),
                ("ib_staging", C.c_int32), ("overlap", C.c_int32),
                ("ms_per_step", C.c_double), ("ms_default", C.c_double),
                ("shapes_timed", C.c_int32), ("lb_persistent", C.c_int32),
                ("forces_by_body", C.c_int32)]


class SceneOpts(C.Structure):
    _fields_ = [("config", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("n_platelets", C.c_int32), ("elements", C.c_int32),
                ("receptors", C.c_int32), ("ligands", C.c_int32),
                ("fibrin_nodes", C.c_int32), ("fibrin_links", C.c_int32),
                ("tube_radius", C.c_float), ("store", C.c_int32), ("seed", C.c_uint64),
                ("mesh_path", C.c_char_p), ("mesh_scale", C.c_float),
                ("fibrin_path", C.c_char_p), ("fibrin_scale", C.c_float)]


# name -> (restype, argtypes); every entry point declared in include/*.h
SIGNATURES = {
    "pfg_version": (C.c_char_p, []),
    "pfg_status_string": (C.c_char_p, [C.c_int]),
    "pfg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pfg_struct_size": (C.c_size_t, [C.c_int]),
    "pfg_params_default": (None, [C.POINTER(Params)]),
    "pfg_params_from_sem": (C.c_int, [f32p, C.c_float, C.c_float, C.c_int, C.c_int,
                                      C.c_int, C.POINTER(Params)]),
    "pfg_create": (C.c_int, [C.POINTER(Params), C.c_int, C.POINTER(C.c_void_p)]),
    "pfg_destroy": (C.c_int, [C.c_void_p]),
    "pfg_get_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "pfg_device_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "pfg_init_equilibrium": (C.c_int, [C.c_void_p, f32p, f32p, f32p, f32p]),
    "pfg_init_equilibrium_z": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "pfg_upload_f": (C.c_int, [C.c_void_p, f32p]),
    "pfg_download_f": (C.c_int, [C.c_void_p, f32p]),
    "pfg_upload_flags": (C.c_int, [C.c_void_p, u8p]),
    "pfg_download_fields": (C.c_int, [C.c_void_p, f32p, f32p, f32p, f32p]),
    "pfg_conservation": (C.c_int, [C.c_void_p, f64p]),
    "pfg_load_model": (C.c_int, [C.c_void_p, C.POINTER(Model)]),
    "pfg_upload_markers": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "pfg_download_markers": (C.c_int, [C.c_void_p] + [f32p] * 9),
    "pfg_upload_marker_forces": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "pfg_upload_marker_velocities": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "pfg_download_membrane": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
    "pfg_download_bonds": (C.c_int, [C.c_void_p, i32p, i32p, f32p]),
    "pfg_upload_bonds": (C.c_int, [C.c_void_p, i32p]),
    "pfg_cell_grid": (C.c_int, [C.c_void_p, i32p]),
    "pfg_download_cells": (C.c_int, [C.c_void_p, i32p, i32p]),
    "pfg_download_aux": (C.c_int, [C.c_void_p, f32p]),
    "pfg_step": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(StepInfo)]),
    "pfg_run_stage": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "pfg_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_fluid_only": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_halt_on_detach": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_ib_staging": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_forces_lanes": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_forces_by_body": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_batch_create": (C.c_int, [C.POINTER(Params), C.c_int, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    "pfg_batch_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pfg_batch_init_equilibrium": (C.c_int, [C.c_void_p, C.c_int, f32p, f32p, f32p, f32p]),
    "pfg_batch_init_equilibrium_z": (C.c_int, [C.c_void_p, C.c_int, f32p, f32p, f32p]),
    "pfg_batch_upload_flags": (C.c_int, [C.c_void_p, C.c_int, u8p]),
    "pfg_batch_load_models": (C.c_int, [C.c_void_p, C.POINTER(Model)]),
    "pfg_batch_download_fields": (C.c_int, [C.c_void_p, C.c_int, f32p, f32p, f32p, f32p]),
    "pfg_batch_download_markers": (C.c_int, [C.c_void_p, C.c_int] + [f32p] * 9),
    "pfg_batch_conservation": (C.c_int, [C.c_void_p, C.c_int, f64p]),
    "pfg_synchronize": (C.c_int, [C.c_void_p]),
    "pfg_profile": (C.c_int, [C.c_void_p, C.c_int, f64p, f64p]),
    "pfg_time_lb": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "pfg_copy_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "pfg_set_lb_tuning": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Tuning)]),
    "pfg_set_lb_xcd_rows": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_set_lb_persistent": (C.c_int, [C.c_void_p, C.c_int]),
    "pfg_download_fields_noalloc": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_float)] * 4),
    "pfg_scene_opts_init": (None, [C.POINTER(SceneOpts), C.c_int]),
    "pfg_scene_build": (C.c_int, [C.POINTER(SceneOpts), C.POINTER(Params),
                                  C.POINTER(C.c_void_p)]),
    "pfg_scene_destroy": (C.c_int, [C.c_void_p]),
    "pfg_scene_model": (C.c_int, [C.c_void_p, C.POINTER(Model)]),
    "pfg_scene_flags": (u8p, [C.c_void_p]),
    "pfg_scene_velocity": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "pfg_scene_velocity_z": (C.c_int, [C.c_void_p, f32p, f32p, f32p]),
    "pfg_scene_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "pfg_scene_fluid_cells": (C.c_int64, [C.c_void_p]),
    "pfg_geom_ellipsoid": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_float, C.c_uint64, f32p,
                                     f32p, f32p, i32p, i32p]),
    "pfg_geom_area_volume": (C.c_int, [C.c_int, i32p, f32p, f32p, f32p, f64p, f64p, f64p]),
    "pfg_geom_read_neu": (C.c_int, [C.c_char_p, C.c_int, C.c_int, f32p, f32p, f32p, i32p,
                                    i32p, i32p]),
    "pfg_geom_read_fibrin": (C.c_int, [C.c_char_p, C.c_int, C.c_int, f32p, f32p, f32p, i32p,
                                       f32p, i32p, i32p]),
    "pfg_geom_fibrin": (C.c_int, [C.c_int, f32p, f32p, C.c_float, C.c_int, C.c_uint64,
                                  f32p, f32p, f32p, i32p, f32p, i32p]),
    # the reference's own entry points (include/pfg_ref.h)
    "pfg_ref_last_status": (C.c_int, []),
    "pfg_ref_handle": (C.c_void_p, [C.c_void_p]),
    "fluid_allocGPUKernel": (C.c_void_p, [C.c_void_p, C.c_float, C.c_float, C.c_int, C.c_int,
                                          C.c_int]),
    "fluid_initGPUKernel": (None, [C.c_void_p, C.c_void_p, C.c_int, f32p,
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)] +
                            [C.c_void_p] * 9),
    "fluid_invokeGPUKernel": (None, [C.c_void_p, C.c_void_p, C.c_void_p, f64p, C.c_void_p,
                                     C.c_int]),
    "fluid_releaseGPUKernel": (None, [C.c_void_p, C.c_void_p]),
    "sem_allocGPUKernel": (C.c_void_p, [C.c_void_p] + [C.c_int] * 5 +
                           [C.c_float, C.c_double, f32p]),
    "sem_initGPUKernel": (None, [C.c_void_p, C.c_void_p, C.c_int, i32p, C.c_int, C.c_int] +
                          [C.c_void_p] * 23 + [C.c_double, C.c_void_p, C.c_void_p]),
    "sem_invokeGPUKernel_Force": (None, [C.c_void_p, C.c_void_p, C.c_int, i32p, i32p,
                                         C.c_void_p, C.c_double]),
    "sem_copyGPUKernel": (None, [C.c_void_p, C.c_void_p] + [C.c_void_p] * 10 + [C.c_int]),
    "sem_releaseGPUKernel": (None, [C.c_void_p, C.c_void_p]),
    "fiber_allocGPUKernel": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                          f64p]),
    "fiber_init_GPUKernel": (None, [C.c_void_p, C.c_void_p] + [C.c_void_p] * 19),
    "fiber_copy_GPUKernel": (None, [C.c_void_p, C.c_void_p] + [C.c_void_p] * 12),
    "fiber_release_GPUKernel": (None, [C.c_void_p, C.c_void_p]),
}

_lib = None


def load():
    """The built library; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C "
            "platelet_fiber_gpu_amd) first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PfgError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} failed: {STATUS.get(code, code)} ({code})")
        self.code = code


def check(fn, code):
    if code != PFG_OK:
        raise PfgError(fn, code)
    return code
