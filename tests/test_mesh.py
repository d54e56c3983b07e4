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


# ---- fibrin network files (init_fibrin_network, fiber.cpp:145-217) ----
# HemoThickInputCor.txt is the reference's own network input (its Input/
# directory), copied as data.

def parse_fibrin_np(path):
    """Independent restatement of the line format for the checks below."""
    rows = {}
    for line in open(path):
        v = line.split()
        if not v:
            continue
        i, n = int(v[0]), int(float(v[4]))
        rows[i] = ([float(q) for q in v[1:4]], [int(float(q)) for q in v[5:5 + n]],
                   [float(q) for q in v[5 + n:5 + 2 * n]])
    return rows


def test_fibrin_file(pfg):
    path = os.path.join(FIX, "HemoThickInputCor.txt")
    x, y, z, link, th = pfg.read_fibrin(path)
    rows = parse_fibrin_np(path)
    assert x.size == len(rows) == 1175
    for i, (p, nb, t) in rows.items():
        assert (x[i - 1], y[i - 1], z[i - 1]) == tuple(np.float32(p))
    want = [(i - 1, c - 1, rows[i][2][k]) for i in sorted(rows)
            for k, c in enumerate(rows[i][1]) if c > i]
    assert len(want) == len(link) == 1422
    assert link.tolist() == [[a, b] for a, b, _ in want]
    assert np.allclose(th, [t for _, _, t in want], rtol=1e-6)
    assert np.all(link[:, 0] < link[:, 1])


@pytest.mark.parametrize("text", [
    "1 0 0 0 1 2 5.0\n2 1 0 0 0\n",                # 1 lists 2, 2 does not list 1
    "1 0 0 0 1 2 5.0\n2 1 0 0 1 1\n",              # thickness missing
    "1 0 0 0 1 3 5.0\n2 1 0 0 1 1 5.0\n",          # neighbour past the node count
    "1 0 0 0 1 2 5.0\n1 1 0 0 1 1 5.0\n",          # node 1 twice, node 2 missing
    "1 0 0 0\n",                                   # no connection count
])
def test_fibrin_file_rejects_malformed(pfg, tmp_path, text):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(pfg.PfgError) as e:
        pfg.read_fibrin(str(p))
    assert e.value.code == 1


def test_fibrin_file_self_links_and_repeats(pfg, tmp_path):
    """Self links are dropped; a repeated neighbour is a repeated link."""
    p = tmp_path / "net.txt"
    p.write_text("1 0 0 0 3 1 2 2 9 4 4\n2 3 0 0 2 1 1 4 4\n")
    x, y, z, link, th = pfg.read_fibrin(str(p))
    assert link.tolist() == [[0, 1], [0, 1]]
    assert th.tolist() == [4.0, 4.0]


def test_fibrin_file_scene(pfg):
    path = os.path.join(FIX, "HemoThickInputCor.txt")
    x, y, z, link, th = pfg.read_fibrin(path)
    sc = pfg.Scene(pfg.SCENE_THROMBUS, nx=128, ny=96, nz=96, n_platelets=2, elements=256,
                   receptors=8, ligands=64, fibrin_path=path)
    nf = x.size
    fx, fy, fz = sc.x[-nf:], sc.y[-nf:], sc.z[-nf:]
    assert sc.link.size // 2 == len(link)
    base = sc.n_markers - nf
    assert np.array_equal(sc.link.reshape(-1, 2) - base, link)
    P = np.stack([fx, fy, fz], 1).astype(np.float64)
    d = np.linalg.norm(P[link[:, 1]] - P[link[:, 0]], axis=1)
    assert np.allclose(sc.link_l0, d, rtol=1e-6)
    # one uniform scale: distances scale alike
    Q = np.stack([x, y, z], 1).astype(np.float64)
    dq = np.linalg.norm(Q[link[:, 1]] - Q[link[:, 0]], axis=1)
    ok = dq > 0
    s = d[ok] / dq[ok]
    assert np.allclose(s, s[0], rtol=1e-4)
    kinds = sc.kind[-nf:]
    assert set(np.unique(kinds)) <= {pfg.MARKER_FIBER, pfg.MARKER_FIXED}
    assert (kinds == pfg.MARKER_FIXED).any()
