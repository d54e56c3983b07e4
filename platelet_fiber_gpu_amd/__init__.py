"""MI355X-native coupled platelet-fibrin step: Python view of the C ABI.

The compute path is the HIP library ``libpfg.so`` (include/pfg.h).  This module
only moves host arrays in and out of it with ctypes; it never computes the
physics itself and raises if the library is missing.

Objects mirror the reference's host-side plumbing (gpuSEM_LB.c): ``Scene`` builds
the seeded inputs that runGPUSimulation assembles (gpuSEM_LB.c:649-1258),
``Solver`` owns the device state that fluid_allocGPUKernel / sem_allocGPUKernel
allocate (semLB.cu:111, 749) and runs the per-step calls of gpuSEM_LB.c:1441-1442.
"""
import ctypes as C
import os

import numpy as np

from . import _ffi
from ._ffi import (BOND_BELL, BOND_REFERENCE, SEM_ADVECT, SEM_BBK)  # noqa: F401
from ._ffi import (FIBER_IB, FIBER_KICK, FIBER_LINEAR, FIBER_WLC, MARKER_ELEMENT,
                   MARKER_FIBER, MARKER_FIXED, SCENE_BATCH, SCENE_CHANNEL, SCENE_LARGE,
                   SCENE_SHEAR, SCENE_THROMBUS, STAGE_BONDS, STAGE_CELLS, STAGE_CLEAR,
                   STAGE_FORCES, STAGE_INTERP, STAGE_LB, STAGE_MEMBRANE, STAGE_SPREAD,
                   STAGE_UPDATE, STORE_BF16, STORE_F16, STORE_F32, WALL_BOUNCE, WALL_VELOCITY,
                   WALL_PERIODIC, Model, Params, PfgError, SceneOpts, StepInfo)

__all__ = [
    "Scene", "Solver", "Batch", "Params", "Model", "StepInfo", "PfgError", "lib", "default_params",
    "device_count", "STORE_F32", "STORE_F16", "STORE_BF16", "WALL_PERIODIC", "WALL_BOUNCE",
    "WALL_VELOCITY",
    "MARKER_ELEMENT", "MARKER_FIBER", "MARKER_FIXED", "FIBER_WLC", "FIBER_LINEAR",
    "FIBER_IB", "FIBER_KICK", "SCENE_CHANNEL", "SCENE_SHEAR", "SCENE_THROMBUS",
    "SCENE_LARGE", "SCENE_BATCH", "STAGE_CELLS", "STAGE_MEMBRANE", "STAGE_BONDS",
    "STAGE_FORCES", "STAGE_SPREAD", "STAGE_LB", "STAGE_INTERP", "STAGE_UPDATE",
    "STAGE_CLEAR", "read_neu", "read_fibrin", "SEM_ADVECT", "SEM_BBK", "BOND_BELL",
    "BOND_REFERENCE",
]

Q = 19


def lib():
    return _ffi.load()


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def default_params(**kw):
    p = Params()
    lib().pfg_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def read_neu(path):
    """A Gambit .neu surface mesh as the library reads it (NEUReader,
    gpuSEM_LB.c:346-424): x, y, z (float32) and triangles (int32, n x 3, from 0)."""
    L = lib()
    nn, nt = C.c_int32(0), C.c_int32(0)
    p = os.fsencode(path)
    _ffi.check("pfg_geom_read_neu", L.pfg_geom_read_neu(p, 0, 0, None, None, None, None,
                                                        C.byref(nn), C.byref(nt)))
    x, y, z = (np.empty(nn.value, np.float32) for _ in range(3))
    tri = np.empty((nt.value, 3), np.int32)
    _ffi.check("pfg_geom_read_neu", L.pfg_geom_read_neu(
        p, nn.value, nt.value, _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(z, C.c_float),
        _ptr(tri, C.c_int32), C.byref(nn), C.byref(nt)))
    return x, y, z, tri


def read_fibrin(path):
    """A fibrin network text file as the library reads it (init_fibrin_network,
    fiber.cpp:145-217): node positions x, y, z (float32), links (int32, m x 2,
    from 0, the pairs i < j of each node's list) and their thicknesses."""
    L = lib()
    nn, nl = C.c_int32(0), C.c_int32(0)
    p = os.fsencode(path)
    _ffi.check("pfg_geom_read_fibrin", L.pfg_geom_read_fibrin(
        p, 0, 0, None, None, None, None, None, C.byref(nn), C.byref(nl)))
    x, y, z = (np.empty(nn.value, np.float32) for _ in range(3))
    link = np.empty((nl.value, 2), np.int32)
    th = np.empty(nl.value, np.float32)
    _ffi.check("pfg_geom_read_fibrin", L.pfg_geom_read_fibrin(
        p, nn.value, nl.value, _ptr(x, C.c_float), _ptr(y, C.c_float), _ptr(z, C.c_float),
        _ptr(link, C.c_int32), _ptr(th, C.c_float), C.byref(nn), C.byref(nl)))
    return x, y, z, link, th


def device_count():
    n = C.c_int(0)
    lib().pfg_device_count(C.byref(n))
    return n.value


def _view(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Scene:
    """A seeded workload built by the library's host-side generators."""

    def __init__(self, config, seed=0, nx=0, ny=0, nz=0, n_platelets=0, elements=0,
                 receptors=0, ligands=0, fibrin_nodes=0, fibrin_links=0, tube_radius=0.0,
                 store=-1, mesh_path=None, mesh_scale=0.0, fibrin_path=None,
                 fibrin_scale=0.0):
        """mesh_path: a Gambit .neu surface mesh used for every platelet
        (scaled by mesh_scale lattice units per mesh unit, 0: mean edge 0.75)
        instead of the generated ellipsoids.  fibrin_path: a fibrin network
        text file (see read_fibrin) placed in the scene's fibrin box (scaled by
        fibrin_scale, 0: fitted) instead of a generated network."""
        L = lib()
        o = SceneOpts()
        L.pfg_scene_opts_init(C.byref(o), config)
        o.seed, o.nx, o.ny, o.nz = seed, nx, ny, nz
        o.n_platelets, o.elements, o.receptors, o.ligands = n_platelets, elements, receptors, ligands
        o.fibrin_nodes, o.fibrin_links = fibrin_nodes, fibrin_links
        o.tube_radius, o.store = tube_radius, store
        o.mesh_path = os.fsencode(mesh_path) if mesh_path else None
        o.mesh_scale = mesh_scale
        o.fibrin_path = os.fsencode(fibrin_path) if fibrin_path else None
        o.fibrin_scale = fibrin_scale
        self.params = Params()
        h = C.c_void_p()
        _ffi.check("pfg_scene_build", L.pfg_scene_build(C.byref(o), C.byref(self.params),
                                                        C.byref(h)))
        self._h = h
        m = Model()
        _ffi.check("pfg_scene_model", L.pfg_scene_model(h, C.byref(m)))
        p = self.params
        self.N = p.nx * p.ny * p.nz
        n = m.n_markers
        self.x, self.y, self.z = (_view(m.x, n, np.float32), _view(m.y, n, np.float32),
                                  _view(m.z, n, np.float32))
        self.kind, self.body = _view(m.kind, n, np.int32), _view(m.body, n, np.int32)
        npl = m.n_platelets
        self.plt_elem_start = _view(m.plt_elem_start, npl + 1, np.int32)
        self.plt_tri_start = _view(m.plt_tri_start, npl + 1, np.int32)
        self.tri = _view(m.tri, 3 * m.n_tri, np.int32)
        self.tri_S0 = _view(m.tri_S0, m.n_tri, np.float64)
        self.plt_S0 = _view(m.plt_S0, npl, np.float64)
        self.plt_V0 = _view(m.plt_V0, npl, np.float64)
        self.edge = _view(m.edge, 2 * m.n_edges, np.int32)
        self.edge_l0 = _view(m.edge_l0, m.n_edges, np.float32)
        self.rec_tri = _view(m.rec_tri, m.n_receptors, np.int32)
        self.rec_r1 = _view(m.rec_r1, m.n_receptors, np.float32)
        self.rec_r2 = _view(m.rec_r2, m.n_receptors, np.float32)
        nl = m.n_wall_ligands
        self.lig_x, self.lig_y = _view(m.lig_x, nl, np.float32), _view(m.lig_y, nl, np.float32)
        self.lig_z = _view(m.lig_z, nl, np.float32)
        self.fibrin_ligands = m.fibrin_ligands
        self.link = _view(m.link, 2 * m.n_links, np.int32)
        self.link_l0 = _view(m.link_l0, m.n_links, np.float32)
        fp = L.pfg_scene_flags(h)
        self.flags = (np.ctypeslib.as_array(fp, shape=(self.N,)).copy() if fp else None)
        self.fluid_cells = int(L.pfg_scene_fluid_cells(h))
        uz_ = [np.zeros(p.nz, np.float32) for _ in range(3)]
        _ffi.check("pfg_scene_velocity_z", L.pfg_scene_velocity_z(
            h, *[_ptr(a, C.c_float) for a in uz_]))
        # initial flow: one value per z-plane (the scenes' flow varies along z
        # only); change it here before loading
        self.u0_z = tuple(uz_)
        L.pfg_scene_destroy(h)
        self._h = None

    @property
    def u0(self):
        """Initial velocity per cell ([nz][ny][nx] flattened, per component),
        expanded from u0_z (a fresh copy on every access)."""
        p = self.params
        return tuple(np.repeat(a, p.nx * p.ny) for a in self.u0_z)

    @property
    def n_markers(self):
        return int(self.x.size)

    @classmethod
    def large_subbox(cls, n=128, seed=0):
        """The 1024x512x512 configuration's physics on an n^3 sub-box with the
        same per-cell parameters: tau 0.7, Couette at the shear rate of a 0.05
        wall velocity over 512 cells, FP16 deviation storage, and the same
        densities of platelets (2000 x 1000 elements per 1024x512x512),
        fibrin (80k nodes, 120k links) and wall ligands (65536 per
        1024x512); the hard Morse cutoff of the configuration."""
        frac = n ** 3 / (1024 * 512 * 512)
        sc = cls(SCENE_LARGE, seed=seed, nx=n, ny=n, nz=n,
                 n_platelets=max(1, round(2000 * frac)),
                 fibrin_nodes=max(8, round(80000 * frac)),
                 fibrin_links=max(8, round(120000 * frac)),
                 ligands=max(64, round(65536 * n * n / (1024 * 512))))
        uw = 0.05 * n / 512
        sc.params.wall_u_hi[2][0] = uw
        sc.params.detach_rate = 0.05 / 512
        sc.u0_z[0][:] = uw * (np.arange(n) + 0.5) / n
        return sc

    def model(self):
        """A Model struct pointing at this scene's numpy arrays (keep self alive)."""
        m = Model()
        n = self.n_markers
        m.n_markers = n
        self._keep = []

        def P(a, ct):
            a = np.ascontiguousarray(a)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct)) if a.size else None

        m.x, m.y, m.z = P(self.x, C.c_float), P(self.y, C.c_float), P(self.z, C.c_float)
        m.kind, m.body = P(self.kind, C.c_int32), P(self.body, C.c_int32)
        m.n_platelets = int(self.plt_S0.size)
        m.plt_elem_start = P(self.plt_elem_start, C.c_int32)
        m.plt_tri_start = P(self.plt_tri_start, C.c_int32)
        m.n_tri = int(self.tri_S0.size)
        m.tri, m.tri_S0 = P(self.tri, C.c_int32), P(self.tri_S0, C.c_double)
        m.plt_S0, m.plt_V0 = P(self.plt_S0, C.c_double), P(self.plt_V0, C.c_double)
        m.n_edges = int(self.edge_l0.size)
        m.edge, m.edge_l0 = P(self.edge, C.c_int32), P(self.edge_l0, C.c_float)
        m.n_receptors = int(self.rec_tri.size)
        m.rec_tri = P(self.rec_tri, C.c_int32)
        m.rec_r1, m.rec_r2 = P(self.rec_r1, C.c_float), P(self.rec_r2, C.c_float)
        m.n_wall_ligands = int(self.lig_x.size)
        m.lig_x, m.lig_y = P(self.lig_x, C.c_float), P(self.lig_y, C.c_float)
        m.lig_z = P(self.lig_z, C.c_float)
        m.fibrin_ligands = int(self.fibrin_ligands)
        m.n_links = int(self.link_l0.size)
        m.link, m.link_l0 = P(self.link, C.c_int32), P(self.link_l0, C.c_float)
        return m


class Solver:
    """Device state of one coupled simulation (one handle, one HIP stream)."""

    def __init__(self, params, device=0):
        self._lib = lib()
        self.params = params
        h = C.c_void_p()
        _ffi.check("pfg_create", self._lib.pfg_create(C.byref(params), device, C.byref(h)))
        self._h = h
        p = params
        self.shape = (p.nz, p.ny, p.nx)
        self.N = p.nx * p.ny * p.nz
        self.n_markers = 0
        self.n_rec = 0
        self.n_lig = 0
        self.n_tri = 0
        self.n_plt = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pfg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, name, *args):
        return _ffi.check(name, getattr(self._lib, name)(self._h, *args))

    # ---- lattice ----
    def init_equilibrium(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else f32(a).ravel() for a in (rho, ux, uy, uz)]
        self._c("pfg_init_equilibrium", *[_ptr(a, C.c_float) for a in arrs])

    def init_equilibrium_z(self, ux_z=None, uy_z=None, uz_z=None):
        """f = f_eq(1, u(z)) from one velocity value per z-plane."""
        arrs = [None if a is None else f32(a).ravel() for a in (ux_z, uy_z, uz_z)]
        for a in arrs:
            if a is not None and a.size != self.params.nz:
                raise ValueError("one value per z-plane expected")
        self._c("pfg_init_equilibrium_z", *[_ptr(a, C.c_float) for a in arrs])

    def upload_flags(self, flags):
        a = np.ascontiguousarray(flags, dtype=np.uint8).ravel()
        self._c("pfg_upload_flags", _ptr(a, C.c_uint8))

    def set_f(self, f):
        a = f32(f).reshape(Q, self.N)
        self._c("pfg_upload_f", _ptr(a, C.c_float))

    def f(self):
        out = np.empty((Q, self.N), np.float32)
        self._c("pfg_download_f", _ptr(out, C.c_float))
        return out

    def fields(self):
        out = [np.empty(self.N, np.float32) for _ in range(4)]
        self._c("pfg_download_fields", *[_ptr(a, C.c_float) for a in out])
        return out

    def conservation(self):
        out = np.zeros(4, np.float64)
        self._c("pfg_conservation", _ptr(out, C.c_double))
        return out

    def aux(self):
        out = np.empty((3, self.N), np.float32)
        self._c("pfg_download_aux", _ptr(out, C.c_float))
        return out

    # ---- markers ----
    def load(self, scene_or_model):
        m = scene_or_model.model() if isinstance(scene_or_model, Scene) else scene_or_model
        self._c("pfg_load_model", C.byref(m))
        self.n_markers = m.n_markers
        self.n_rec = m.n_receptors
        self.n_lig = m.n_wall_ligands + (m.n_markers if m.fibrin_ligands else 0)
        self.n_tri = m.n_tri
        self.n_plt = m.n_platelets
        if isinstance(scene_or_model, Scene):
            s = scene_or_model
            if s.flags is not None and self.params.use_flags:
                self.upload_flags(s.flags)
            self.init_equilibrium_z(*s.u0_z)

    def set_markers(self, x, y, z):
        a = [f32(v) for v in (x, y, z)]
        self._c("pfg_upload_markers", *[_ptr(v, C.c_float) for v in a])

    def set_marker_forces(self, fx, fy, fz):
        a = [f32(v) for v in (fx, fy, fz)]
        self._c("pfg_upload_marker_forces", *[_ptr(v, C.c_float) for v in a])

    def set_marker_velocities(self, vx, vy, vz):
        a = [f32(v) for v in (vx, vy, vz)]
        self._c("pfg_upload_marker_velocities", *[_ptr(v, C.c_float) for v in a])

    def markers(self):
        n = self.n_markers
        out = {k: np.empty(n, np.float32) for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}
        self._c("pfg_download_markers", *[_ptr(out[k], C.c_float) for k in
                                          ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")])
        return out

    def membrane(self):
        S = np.empty(self.n_plt, np.float64)
        V = np.empty(self.n_plt, np.float64)
        tS = np.empty(self.n_tri, np.float64)
        self._c("pfg_download_membrane", _ptr(S, C.c_double), _ptr(V, C.c_double),
                _ptr(tS, C.c_double))
        return S, V, tS

    def bonds(self):
        rb = np.empty(self.n_rec, np.int32)
        lo = np.empty(self.n_lig, np.int32)
        rf = np.empty(3 * self.n_rec, np.float32)
        self._c("pfg_download_bonds", _ptr(rb, C.c_int32), _ptr(lo, C.c_int32),
                _ptr(rf, C.c_float))
        return rb, lo, rf.reshape(-1, 3)

    def set_bonds(self, rec_bond):
        a = np.ascontiguousarray(rec_bond, dtype=np.int32)
        self._c("pfg_upload_bonds", _ptr(a, C.c_int32))

    def cells(self):
        dims = np.zeros(3, np.int32)
        self._c("pfg_cell_grid", _ptr(dims, C.c_int32))
        nb = int(np.prod(dims))
        start = np.empty(nb + 1, np.int32)
        items = np.empty(self.n_markers, np.int32)
        self._c("pfg_download_cells", _ptr(start, C.c_int32), _ptr(items, C.c_int32))
        return dims, start, items

    # ---- stepping ----
    def step(self, n=1):
        info = StepInfo()
        self._c("pfg_step", int(n), C.byref(info))
        return {k: getattr(info, k) for k, _ in StepInfo._fields_}

    def stage(self, s):
        self._c("pfg_run_stage", int(s))

    def set_step(self, step):
        self._c("pfg_set_step", int(step))

    def set_graph(self, on):
        self._c("pfg_set_graph", int(bool(on)))

    def set_fluid_only(self, on):
        self._c("pfg_set_fluid_only", int(bool(on)))

    def set_ib_staging(self, mode):
        """IB spreading form: 0 automatic, 1 LDS-staged chunks, 2 direct,
        3 tile gather (no device-wide atomics), 4 by platelet (plain stores
        into the tiles a platelet meets alone)."""
        self._c("pfg_set_ib_staging", int(mode))

    def set_forces_lanes(self, lanes):
        """Lanes per marker in the force pass: 0 automatic, or 1, 4, 16, 64."""
        self._c("pfg_set_forces_lanes", int(lanes))

    def set_forces_by_body(self, mode):
        """Force pass by platelet (Morse pairs from an LDS cell list of each
        platelet): 0 automatic, 1 on, 2 off."""
        self._c("pfg_set_forces_by_body", int(mode))

    def set_overlap(self, on):
        """LB rows away from IB stencils update concurrently with the marker
        pipeline (default off)."""
        self._c("pfg_set_overlap", int(bool(on)))

    def set_halt_on_detach(self, on):
        """End the run at detach: later steps change nothing and step()
        reports the steps that ran."""
        self._c("pfg_set_halt_on_detach", int(bool(on)))

    def synchronize(self):
        self._c("pfg_synchronize")

    def profile(self, n):
        tot = C.c_double()
        lb = C.c_double()
        self._c("pfg_profile", int(n), C.byref(tot), C.byref(lb))
        return tot.value, lb.value

    def time_lb(self, n):
        """Mean LB-kernel duration in ms over n back-to-back launches."""
        ms = C.c_double()
        self._c("pfg_time_lb", int(n), C.byref(ms))
        return ms.value

    def set_lb_tuning(self, cells_per_lane):
        self._c("pfg_set_lb_tuning", int(cells_per_lane))

    def set_lb_xcd_rows(self, on):
        self._c("pfg_set_lb_xcd_rows", int(bool(on)))

    def autotune(self, steps=32):
        """Time `steps` steps per launch shape, keep the fastest on the
        handle (the state is restored afterwards); returns the choice."""
        t = _ffi.Tuning()
        self._c("pfg_autotune", int(steps), C.byref(t))
        return {k: getattr(t, k) for k, _ in _ffi.Tuning._fields_}

    def device_bytes(self):
        b = C.c_size_t()
        self._c("pfg_device_bytes", C.byref(b))
        return b.value


class Batch(Solver):
    """Many lattices stepped together (pfg_batch_*): lattice k has its own
    size, walls, tau and body force; storage and marker parameters come from
    the first.  step(), set_graph(), markers() etc. act on the whole batch."""

    def __init__(self, params_list, device=0):
        self._lib = lib()
        self.params_list = list(params_list)
        self.params = self.params_list[0]
        arr = (Params * len(self.params_list))(*self.params_list)
        h = C.c_void_p()
        _ffi.check("pfg_batch_create", self._lib.pfg_batch_create(
            arr, len(self.params_list), device, C.byref(h)))
        self._h = h
        self.K = len(self.params_list)
        self.Ns = [p.nx * p.ny * p.nz for p in self.params_list]
        self.N = sum(self.Ns)
        self.n_markers = 0
        self.n_markers_k = [0] * self.K
        self.n_rec = self.n_lig = self.n_tri = self.n_plt = 0

    def load_scenes(self, scenes):
        """Models, flags and initial flow of one seeded scene per lattice."""
        models = (Model * self.K)(*[s.model() for s in scenes])
        self._c("pfg_batch_load_models", models)
        self.n_markers_k = [m.n_markers for m in models]
        self.n_markers = sum(self.n_markers_k)
        self.n_rec = sum(m.n_receptors for m in models)
        self.n_rec_k = [m.n_receptors for m in models]
        self.n_wall_lig_k = [m.n_wall_ligands for m in models]
        self.n_lig = sum(self.n_wall_lig_k) + (self.n_markers if models[0].fibrin_ligands else 0)
        self.n_tri = sum(m.n_tri for m in models)
        self.n_plt = sum(m.n_platelets for m in models)
        for k, s in enumerate(scenes):
            if s.flags is not None and self.params_list[k].use_flags:
                a = np.ascontiguousarray(s.flags, dtype=np.uint8).ravel()
                self._c("pfg_batch_upload_flags", int(k), _ptr(a, C.c_uint8))
            arrs = [f32(v).ravel() for v in s.u0_z]
            self._c("pfg_batch_init_equilibrium_z", int(k), *[_ptr(a, C.c_float) for a in arrs])

    def set_lb_persistent(self, on):
        """One persistent launch of resident workgroups taking every lattice's
        x-block rows from a device work queue (pfg_set_lb_persistent)."""
        self._c("pfg_set_lb_persistent", int(bool(on)))

    def init_lattice(self, k, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else f32(a).ravel() for a in (rho, ux, uy, uz)]
        self._c("pfg_batch_init_equilibrium", int(k), *[_ptr(a, C.c_float) for a in arrs])

    def fields_of(self, k):
        out = [np.empty(self.Ns[k], np.float32) for _ in range(4)]
        self._c("pfg_batch_download_fields", int(k), *[_ptr(a, C.c_float) for a in out])
        return out

    def conservation_of(self, k):
        out = np.zeros(4, np.float64)
        self._c("pfg_batch_conservation", int(k), _ptr(out, C.c_double))
        return out

    def markers_of(self, k):
        n = self.n_markers_k[k]
        names = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")
        out = {name: np.empty(n, np.float32) for name in names}
        self._c("pfg_batch_download_markers", int(k), *[_ptr(out[nm], C.c_float) for nm in names])
        return out
