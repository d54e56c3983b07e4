"""ctypes wrapper of the FP64 oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from platelet_fiber_gpu_amd._ffi import Model, Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")
PORT_LIB = os.path.join(_HERE, "_build", "libcpuport.so")
Q = 19

f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int)
f32p = C.POINTER(C.c_float)


class Lattice(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("periodic", C.c_int * 3),
                ("wall_u_lo", (C.c_double * 3) * 3), ("wall_u_hi", (C.c_double * 3) * 3),
                ("tau", C.c_double), ("g", C.c_double * 3),
                ("solid", C.POINTER(C.c_uint8)), ("zvel", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    sig = {
        "orc_lattice_from_params": (None, [C.POINTER(Params), C.POINTER(C.c_uint8), C.POINTER(Lattice)]),
        "orc_w": (C.c_double, [C.c_int]),
        "orc_c": (C.c_int, [C.c_int, C.c_int]),
        "orc_bb": (C.c_int, [C.c_int]),
        "orc_init_eq": (None, [C.POINTER(Lattice), f64p, f64p, f64p, f64p, f64p]),
        "orc_lb_step": (None, [C.POINTER(Lattice), f64p, f64p, f64p, f64p, f64p]),
        "orc_fields": (None, [C.POINTER(Lattice), f64p, f64p, f64p, f64p, f64p]),
        "orc_conservation": (None, [C.POINTER(Lattice), f64p, f64p]),
        "orc_phi": (C.c_double, [C.c_double]),
        "orc_ib_spread": (None, [C.POINTER(Lattice), C.c_int, f64p, f64p, f64p, f64p, f64p,
                                 f64p, i32p, C.c_double, f64p]),
        "orc_ib_spread_scan": (None, [C.POINTER(Lattice), C.c_int, f64p, f64p, f64p, f64p,
                                      f64p, f64p, i32p, C.c_double, f64p]),
        "orc_ib_interp": (None, [C.POINTER(Lattice), C.c_int, f64p, f64p, f64p, f64p, i32p,
                                 f64p, f64p, f64p]),
        "orc_cell_list": (None, [C.c_int, f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, i32p,
                                 i32p, i32p, i32p]),
        "orc_create": (C.c_void_p, [C.POINTER(Params), C.POINTER(Model), C.POINTER(C.c_uint8)]),
        "orc_destroy": (None, [C.c_void_p]),
        "orc_set_order": (None, [C.c_void_p, C.c_int]),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_step": (None, [C.c_void_p]),
        "orc_stage": (None, [C.c_void_p, C.c_int]),
        "orc_set_step": (None, [C.c_void_p, C.c_int64]),
        "orc_f": (f64p, [C.c_void_p]),
        "orc_force_field": (f64p, [C.c_void_p]),
        "orc_vel_field": (f64p, [C.c_void_p]),
        "orc_marker": (f64p, [C.c_void_p, C.c_int]),
        "orc_membrane": (f64p, [C.c_void_p, C.c_int]),
        "orc_lowest": (i32p, [C.c_void_p]),
        "orc_plt_near_cutoff": (C.POINTER(C.c_int64), [C.c_void_p]),
        "orc_rec_bond": (i32p, [C.c_void_p]),
        "orc_lig_owner": (i32p, [C.c_void_p]),
        "orc_rec_force": (f64p, [C.c_void_p]),
        "orc_cells": (i32p, [C.c_void_p, C.c_int, i32p]),
        "orc_counters": (None, [C.c_void_p, C.POINTER(C.c_int64)]),
        "orc_bend": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
        "orc_ring_laplacian": (C.c_int, [C.c_void_p, f64p, f64p, f64p]),
    }
    for k, (r, a) in sig.items():
        fn = getattr(L, k)
        fn.restype = r
        fn.argtypes = a
    _lib = L
    return L


def set_threads(n=0):
    """OpenMP threads of the oracle's lattice loops (0 keeps the current
    count); returns the count in use."""
    return int(load().orc_set_threads(int(n)))


def P(a, ct=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


def d64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lattice(params, solid=None):
    L = Lattice()
    s = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8)
    load().orc_lattice_from_params(C.byref(params), P(s, C.c_uint8), C.byref(L))
    L._keep = s
    return L


def init_eq(L, rho=None, ux=None, uy=None, uz=None):
    N = L.nx * L.ny * L.nz
    f = np.zeros((Q, N))
    args = [None if a is None else d64(a).ravel() for a in (rho, ux, uy, uz)]
    load().orc_init_eq(C.byref(L), *[P(a) for a in args], P(f))
    return f


def lb_step(L, fin, Fu=None, Fs=None, want_u=False):
    N = L.nx * L.ny * L.nz
    fin = d64(fin).reshape(Q, N)
    fout = np.zeros_like(fin)
    u = np.zeros((3, N)) if want_u else None
    Fu = None if Fu is None else d64(Fu).reshape(3, N)
    Fs = None if Fs is None else d64(Fs).reshape(3, N)
    load().orc_lb_step(C.byref(L), P(fin), P(fout), P(Fu), P(Fs), P(u))
    return (fout, u) if want_u else fout


def fields(L, f):
    N = L.nx * L.ny * L.nz
    out = [np.zeros(N) for _ in range(4)]
    load().orc_fields(C.byref(L), P(d64(f).reshape(Q, N)), *[P(a) for a in out])
    return out


def conservation(L, f):
    out = np.zeros(4)
    N = L.nx * L.ny * L.nz
    load().orc_conservation(C.byref(L), P(d64(f).reshape(Q, N)), P(out))
    return out


def phi(r):
    return load().orc_phi(float(r))


def ib_spread(L, x, y, z, fx, fy, fz, coupled=None, scale=1.0, scan=False):
    N = L.nx * L.ny * L.nz
    F = np.zeros((3, N))
    a = [d64(v) for v in (x, y, z, fx, fy, fz)]
    c = None if coupled is None else np.ascontiguousarray(coupled, dtype=np.int32)
    fn = load().orc_ib_spread_scan if scan else load().orc_ib_spread
    fn(C.byref(L), len(a[0]), *[P(v) for v in a], P(c, C.c_int), float(scale), P(F))
    return F


def ib_interp(L, x, y, z, u, coupled=None):
    n = len(x)
    a = [d64(v) for v in (x, y, z)]
    out = [np.zeros(n) for _ in range(3)]
    c = None if coupled is None else np.ascontiguousarray(coupled, dtype=np.int32)
    N = L.nx * L.ny * L.nz
    load().orc_ib_interp(C.byref(L), n, *[P(v) for v in a], P(d64(u).reshape(3, N)),
                         P(c, C.c_int), *[P(v) for v in out])
    return out


def cell_list(x, y, z, dims, periodic, nb):
    n = len(x)
    xs = [np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z)]
    per = np.ascontiguousarray(periodic, dtype=np.int32)
    nbv = np.ascontiguousarray(nb, dtype=np.int32)
    start = np.zeros(int(np.prod(nbv)) + 1, np.int32)
    items = np.zeros(n, np.int32)
    load().orc_cell_list(n, *[P(v, C.c_float) for v in xs], int(dims[0]), int(dims[1]),
                         int(dims[2]), P(per, C.c_int), P(nbv, C.c_int), P(start, C.c_int),
                         P(items, C.c_int))
    return start, items


class Oracle:
    """Coupled FP64 state, stepped stage by stage like the device."""

    def __init__(self, params, scene=None, model=None, flags=None):
        self._lib = load()
        self.params = params
        m = scene.model() if scene is not None else model
        self._scene = scene
        fl = None
        if scene is not None and scene.flags is not None:
            fl = scene.flags
        if flags is not None:
            fl = flags
        self._flags = None if fl is None else np.ascontiguousarray(fl, dtype=np.uint8)
        self._h = self._lib.orc_create(C.byref(params), C.byref(m) if m is not None else None,
                                       P(self._flags, C.c_uint8))
        self.N = params.nx * params.ny * params.nz
        self.n = m.n_markers if m is not None else 0
        self.n_rec = m.n_receptors if m is not None else 0
        self.n_lig = (m.n_wall_ligands + (self.n if m.fibrin_ligands else 0)) if m is not None else 0
        self.n_tri = m.n_tri if m is not None else 0
        self.n_plt = m.n_platelets if m is not None else 0
        if scene is not None:
            L = lattice(params, self._flags)
            self.set_f(init_eq(L, None, *scene.u0))

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def f(self):
        return self._arr(self._lib.orc_f(self._h), Q * self.N).reshape(Q, self.N)

    def set_f(self, f):
        self.f()[...] = np.asarray(f, dtype=np.float64).reshape(Q, self.N)

    def force_field(self):
        return self._arr(self._lib.orc_force_field(self._h), 3 * self.N).reshape(3, self.N)

    def vel_field(self):
        return self._arr(self._lib.orc_vel_field(self._h), 3 * self.N).reshape(3, self.N)

    def marker(self, which):
        names = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")
        k = names.index(which) if isinstance(which, str) else which
        return self._arr(self._lib.orc_marker(self._h, k), self.n)

    def markers(self):
        return {k: self.marker(k).copy() for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")}

    def membrane(self):
        S = self._arr(self._lib.orc_membrane(self._h, 0), self.n_plt).copy()
        V = self._arr(self._lib.orc_membrane(self._h, 1), self.n_plt).copy()
        tS = self._arr(self._lib.orc_membrane(self._h, 2), self.n_tri).copy()
        return S, V, tS

    def bend(self):
        """Mean curvature, bending force (n x 3) and node normals (n x 3) of
        the vertices from the current positions; None for open meshes."""
        km = np.zeros(self.n)
        fb = np.zeros((self.n, 3))
        nrm = np.zeros((self.n, 3))
        if not self._lib.orc_bend(self._h, P(km), P(fb), P(nrm)):
            return None
        return km, fb, nrm

    def ring_laplacian(self, f):
        f = d64(f)
        out = np.zeros(self.n)
        area = np.zeros(self.n)
        if not self._lib.orc_ring_laplacian(self._h, P(f), P(out), P(area)):
            return None
        return out, area

    def platelets_near_cutoff(self):
        """Per platelet: Morse pair visits whose distance lay within 1e-4 of
        the cutoff so far (a rounding difference can switch those pairs)."""
        if self.n_plt == 0:
            return np.zeros(0, np.int64)
        return self._arr(self._lib.orc_plt_near_cutoff(self._h), self.n_plt).copy()

    def lowest(self):
        return self._arr(self._lib.orc_lowest(self._h), self.n_plt).copy()

    def bonds(self):
        rb = self._arr(self._lib.orc_rec_bond(self._h), self.n_rec).copy() if self.n_rec else np.zeros(0, np.int32)
        lo = self._arr(self._lib.orc_lig_owner(self._h), self.n_lig).copy() if self.n_lig else np.zeros(0, np.int32)
        rf = (self._arr(self._lib.orc_rec_force(self._h), 3 * self.n_rec).reshape(-1, 3).copy()
              if self.n_rec else np.zeros((0, 3)))
        return rb, lo, rf

    def set_bonds(self, rec_bond):
        rb = self._arr(self._lib.orc_rec_bond(self._h), self.n_rec)
        lo = self._arr(self._lib.orc_lig_owner(self._h), self.n_lig)
        rb[:] = rec_bond
        lo[:] = -1
        for r, l in enumerate(rec_bond):
            if l >= 0:
                lo[l] = r

    def cells(self):
        nb = np.zeros(3, np.int32)
        start_p = self._lib.orc_cells(self._h, 0, P(nb, C.c_int))
        nbins = int(np.prod(nb))
        start = self._arr(start_p, nbins + 1).copy()
        items = self._arr(self._lib.orc_cells(self._h, 1, None), self.n).copy() if self.n else np.zeros(0, np.int32)
        return nb, start, items

    def counters(self):
        out = (C.c_int64 * 6)()
        self._lib.orc_counters(self._h, out)
        return dict(zip(("total_bonds", "detached", "detach_step", "near_threshold", "bond_events",
                         "near_cutoff"), list(out)))

    def set_order(self, order):
        self._lib.orc_set_order(self._h, int(order))

    def set_step(self, s):
        self._lib.orc_set_step(self._h, int(s))

    def step(self, n=1):
        for _ in range(n):
            self._lib.orc_step(self._h)

    def stage(self, s):
        self._lib.orc_stage(self._h, int(s))


_port = None


def load_port():
    """The FP32 OpenMP CPU port (the CPU baseline), built by the Makefile."""
    global _port
    if _port is not None:
        return _port
    if not os.path.exists(PORT_LIB):
        build()
    L = C.CDLL(PORT_LIB)
    f32 = C.POINTER(C.c_float)
    sig = {
        "cport_create": (C.c_void_p, [C.POINTER(Params), C.POINTER(Model), f32, f32, f32, C.c_int]),
        "cport_destroy": (None, [C.c_void_p]),
        "cport_step": (None, [C.c_void_p, C.c_int]),
        "cport_markers": (None, [C.c_void_p, f32, f32, f32]),
        "cport_velocity": (None, [C.c_void_p, f32, f32, f32]),
        "cport_bonds": (C.c_int64, [C.c_void_p]),
    }
    for k, (r, a) in sig.items():
        fn = getattr(L, k)
        fn.restype = r
        fn.argtypes = a
    _port = L
    return L


class CpuPort:
    """The coupled step in FP32 on the host's cores (OpenMP): the CPU
    baseline.  Raises ValueError when the scene uses a term the port does
    not carry."""

    def __init__(self, params, scene, threads=0):
        self._lib = load_port()
        m = scene.model()
        self._scene = scene
        u = [np.ascontiguousarray(a, dtype=np.float32) for a in scene.u0_z]
        self._keep = u
        f32 = C.POINTER(C.c_float)
        self._h = self._lib.cport_create(C.byref(params), C.byref(m),
                                         *[a.ctypes.data_as(f32) for a in u], int(threads))
        if not self._h:
            raise ValueError("the scene uses a term the CPU port does not carry")
        self.n = m.n_markers
        self.N = params.nx * params.ny * params.nz

    def step(self, n=1):
        self._lib.cport_step(self._h, int(n))

    def markers(self):
        out = [np.empty(self.n, np.float32) for _ in range(3)]
        self._lib.cport_markers(self._h, *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in out])
        return out

    def velocity(self):
        out = [np.empty(self.N, np.float32) for _ in range(3)]
        self._lib.cport_velocity(self._h, *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in out])
        return out

    def close(self):
        if self._h:
            self._lib.cport_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
