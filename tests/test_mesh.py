"""The Gambit .neu mesh reader (NEUReader, gpuSEM_LB.c:346-424) and platelets
built from a mesh file.

Fixtures: cube.neu, 1triangle.neu and Platelet_Mesh_958.neu are input files
of the reference (its Input/ directory, gpu_20151203.tar.gz), copied as data.
The cube (side 2, each face split into four triangles about its centre)
pins the reader together with the area/volume routine through closed forms:
S = 24, V = 8."""
import collections
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "fixtures")


def mesh(pfg, name):
    return pfg.read_neu(os.path.join(FIX, name))


def area_volume(pfg, x, y, z, tri):
    """The library's FP64 area/volume (sem_kernel.cu:793-808 sign)."""
    tS = np.empty(len(tri), np.float64)
    S, V = C.c_double(), C.c_double()
    t = np.ascontiguousarray(tri, np.int32)
    pfg.lib().pfg_geom_area_volume(
        len(tri), t.ctypes.data_as(C.POINTER(C.c_int32)),
        *[np.ascontiguousarray(a, np.float32).ctypes.data_as(C.POINTER(C.c_float))
          for a in (x, y, z)],
        tS.ctypes.data_as(C.POINTER(C.c_double)), C.byref(S), C.byref(V))
    return tS, S.value, V.value


def test_cube_closed_forms(pfg):
    x, y, z, tri = mesh(pfg, "cube.neu")
    assert x.size == 14 and tri.shape == (24, 3)
    # node 8 of the file (1-based) is (1, 1, -1); triangle 1 is "1 8 7"
    assert (x[7], y[7], z[7]) == (1.0, 1.0, -1.0)
    assert tri[0].tolist() == [0, 7, 6]
    tS, S, V = area_volume(pfg, x, y, z, tri)
    assert np.allclose(tS, 1.0, rtol=0, atol=1e-12)
    assert S == pytest.approx(24.0, abs=1e-12)
    assert V == pytest.approx(8.0, abs=1e-12)


def test_single_triangle_values(pfg):
    x, y, z, tri = mesh(pfg, "1triangle.neu")
    assert tri.tolist() == [[0, 1, 2]]
    assert x.tolist() == pytest.approx([-0.3215051, -0.4505655, -0.3339725], abs=1e-7)
    assert z.tolist() == pytest.approx([-0.1856425, -0.1533348, -0.1526267], abs=1e-7)


def test_reference_platelet_mesh(pfg):
    """958 nodes and 1912 = 2 x 958 - 4 triangles: a closed surface, every
    directed edge once and its reverse once; area and volume match an
    independent numpy evaluation."""
    x, y, z, tri = mesh(pfg, "Platelet_Mesh_958.neu")
    assert x.size == 958 and tri.shape == (1912, 3)
    e = collections.Counter()
    for t in tri:
        for k in range(3):
            e[(t[k], t[(k + 1) % 3])] += 1
    assert max(e.values()) == 1
    assert all(e.get((b, a), 0) == 1 for (a, b) in e)
    P = np.stack([x, y, z], 1).astype(np.float64)
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    cr = np.cross(b - a, c - a)
    S_np = 0.5 * np.linalg.norm(cr, axis=1).sum()
    V_np = abs(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0  # divergence theorem
    _, S, V = area_volume(pfg, x, y, z, tri)
    assert S == pytest.approx(S_np, rel=1e-12)
    assert V == pytest.approx(V_np, rel=1e-10)
    assert V > 0


@pytest.mark.parametrize("text, code", [
    ("3 1\n1 0 0 0\n2 1 0 0\n3 0 1 0\n1 3 3 1 2 4\n", 1),      # node index past the count
    ("3 1\n1 0 0 0\n1 1 0 0\n3 0 1 0\n1 3 3 1 2 3\n", 1),      # node 1 twice, node 2 missing
    ("3 1\n1 0 0 0\n2 1 0 0\n3 0 1 0\n1 3 3 1 2\n", 1),        # truncated triangle line
    ("2 1\n1 0 0 0\n2 1 0 0\n1 3 3 1 2 2\n", 1),               # fewer than 3 nodes
    ("3 1\n1 0 0 0\n2 1 0 nan\n3 0 1 0\n1 3 3 1 2 3\n", 1),    # non-finite coordinate
])
def test_reader_rejects_malformed(pfg, tmp_path, text, code):
    p = tmp_path / "bad.neu"
    p.write_text(text)
    with pytest.raises(pfg.PfgError) as e:
        pfg.read_neu(str(p))
    assert e.value.code == code


def test_reader_missing_file(pfg, tmp_path):
    with pytest.raises(pfg.PfgError) as e:
        pfg.read_neu(str(tmp_path / "none.neu"))
    assert e.value.code == 6


def test_reader_buffer_too_small(pfg):
    L = pfg.lib()
    nn, nt = C.c_int32(), C.c_int32()
    buf = [np.empty(14, np.float32) for _ in range(3)]
    tri = np.empty((23, 3), np.int32)
    st = L.pfg_geom_read_neu(os.fsencode(os.path.join(FIX, "cube.neu")), 14, 23,
                             *[b.ctypes.data_as(C.POINTER(C.c_float)) for b in buf],
                             tri.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nn), C.byref(nt))
    assert st == 1


def test_mesh_scene(pfg):
    """Every platelet of the scene is the file mesh, translated and scaled so
    that its mean edge is 0.75: rest area and volume scale as s^2 and s^3."""
    path = os.path.join(FIX, "Platelet_Mesh_958.neu")
    x, y, z, tri = mesh(pfg, "Platelet_Mesh_958.neu")
    P = np.stack([x, y, z], 1).astype(np.float64)
    edges = np.concatenate([P[tri[:, k]] - P[tri[:, (k + 1) % 3]] for k in range(3)])
    s = 0.75 / np.linalg.norm(edges, axis=1).mean()
    _, S, V = area_volume(pfg, x, y, z, tri)
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=96, ny=64, nz=64, n_platelets=3, receptors=16,
                   ligands=-1, fibrin_nodes=-1, mesh_path=path)
    assert sc.n_markers == 3 * 958 and sc.tri.size == 3 * 3 * 1912
    assert np.allclose(sc.plt_S0, S * s * s, rtol=1e-5)
    assert np.allclose(sc.plt_V0, V * s ** 3, rtol=1e-5)
    first = np.stack([sc.x[:958], sc.y[:958], sc.z[:958]], 1)
    d = (first - first.mean(0)) / s - (P - P.mean(0))
    assert np.abs(d).max() < 1e-4
    assert sc.tri[:3].tolist() == tri[0].tolist()
    # an explicit scale
    sc2 = pfg.Scene(pfg.SCENE_CHANNEL, mesh_path=path, mesh_scale=6.0)
    assert np.allclose(sc2.plt_S0, S * 36.0, rtol=1e-5)
