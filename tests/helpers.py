"""Shared test helpers: build pfg_model structs from numpy arrays."""
import ctypes as C

import numpy as np

from platelet_fiber_gpu_amd._ffi import Model


class ModelArrays:
    """Owns numpy arrays and exposes a Model struct pointing at them."""

    def __init__(self, x, y, z, kind=None, body=None, plt_elem_start=None,
                 plt_tri_start=None, tri=None, tri_S0=None, plt_S0=None, plt_V0=None,
                 edge=None, edge_l0=None, rec_tri=None, rec_r1=None, rec_r2=None,
                 lig_x=None, lig_y=None, lig_z=None, fibrin_ligands=0, link=None,
                 link_l0=None):
        n = len(x)
        self.a = {}

        def keep(name, arr, dt):
            if arr is None:
                self.a[name] = None
                return None
            v = np.ascontiguousarray(arr, dtype=dt)
            self.a[name] = v
            return v

        keep("x", x, np.float32)
        keep("y", y, np.float32)
        keep("z", z, np.float32)
        keep("kind", np.zeros(n) if kind is None else kind, np.int32)
        keep("body", np.zeros(n) if body is None else body, np.int32)
        nplt = 0 if plt_S0 is None else len(plt_S0)
        keep("plt_elem_start", [0] if plt_elem_start is None else plt_elem_start, np.int32)
        keep("plt_tri_start", [0] if plt_tri_start is None else plt_tri_start, np.int32)
        keep("tri", tri, np.int32)
        keep("tri_S0", tri_S0, np.float64)
        keep("plt_S0", plt_S0, np.float64)
        keep("plt_V0", plt_V0, np.float64)
        keep("edge", edge, np.int32)
        keep("edge_l0", edge_l0, np.float32)
        keep("rec_tri", rec_tri, np.int32)
        keep("rec_r1", rec_r1, np.float32)
        keep("rec_r2", rec_r2, np.float32)
        keep("lig_x", lig_x, np.float32)
        keep("lig_y", lig_y, np.float32)
        keep("lig_z", lig_z, np.float32)
        keep("link", link, np.int32)
        keep("link_l0", link_l0, np.float32)
        m = Model()
        P = self._p
        m.n_markers = n
        m.x, m.y, m.z = P("x", C.c_float), P("y", C.c_float), P("z", C.c_float)
        m.kind, m.body = P("kind", C.c_int32), P("body", C.c_int32)
        m.n_platelets = nplt
        m.plt_elem_start = P("plt_elem_start", C.c_int32)
        m.plt_tri_start = P("plt_tri_start", C.c_int32)
        m.n_tri = 0 if tri is None else len(tri) // 3
        m.tri, m.tri_S0 = P("tri", C.c_int32), P("tri_S0", C.c_double)
        m.plt_S0, m.plt_V0 = P("plt_S0", C.c_double), P("plt_V0", C.c_double)
        m.n_edges = 0 if edge_l0 is None else len(edge_l0)
        m.edge, m.edge_l0 = P("edge", C.c_int32), P("edge_l0", C.c_float)
        m.n_receptors = 0 if rec_tri is None else len(rec_tri)
        m.rec_tri = P("rec_tri", C.c_int32)
        m.rec_r1, m.rec_r2 = P("rec_r1", C.c_float), P("rec_r2", C.c_float)
        m.n_wall_ligands = 0 if lig_x is None else len(lig_x)
        m.lig_x, m.lig_y, m.lig_z = P("lig_x", C.c_float), P("lig_y", C.c_float), P("lig_z", C.c_float)
        m.fibrin_ligands = int(fibrin_ligands)
        m.n_links = 0 if link_l0 is None else len(link_l0)
        m.link, m.link_l0 = P("link", C.c_int32), P("link_l0", C.c_float)
        self.model = m

    def _p(self, name, ct):
        v = self.a[name]
        if v is None or v.size == 0:
            return None
        return v.ctypes.data_as(C.POINTER(ct))


def cube_mesh():
    """Unit cube, 8 vertices, 12 triangles wound counter-clockwise seen from
    outside (e01 x e02 outward)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=float)
    faces = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
             (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return v, np.array(faces, dtype=np.int32)


def area_volume_np(v, tri):
    """FP64 restatement of the triangle area and reference-signed volume."""
    p0, p1, p2 = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    ln = np.linalg.norm(n, axis=1)
    S = 0.5 * ln
    nin = -n / ln[:, None]
    cen = (p0 + p1 + p2) / 3
    V = (np.einsum("ij,ij->i", cen, nin) * S / 3).sum()
    return S, S.sum(), V
