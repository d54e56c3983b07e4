"""The CPU baseline (oracle/cpu_port.cpp: the coupled step in FP32 with
OpenMP) computes what the FP64 oracle computes: the large configuration's
physics on a small sub-box, 15 coupled steps, collision velocity and marker
positions; and it refuses terms it does not carry."""
import numpy as np
import pytest


def test_cpu_port_follows_the_oracle(pfg, po):
    sc = pfg.Scene.large_subbox(40)
    sc.params.store = pfg.STORE_F32
    assert sc.n_markers > 0 and sc.rec_tri.size > 0 and sc.link.size > 0
    c = po.CpuPort(sc.params, sc, threads=4)
    c.step(15)
    o = po.Oracle(sc.params, scene=sc)
    o.step(15)
    ou = o.vel_field()
    u = np.stack(c.velocity())
    err_u = np.sqrt(((u - ou) ** 2).sum() / (ou ** 2).sum())
    P = np.stack(c.markers(), 1)
    OP = np.stack([o.marker(k) for k in "xyz"], 1)
    err_x = np.abs(P - OP).max()
    c.close()
    o.close()
    assert err_u < 1e-4
    assert err_x < 1e-4


def test_cpu_port_refuses_other_terms(pfg, po):
    sc = pfg.Scene(pfg.SCENE_SHEAR, nx=32, ny=16, nz=16, elements=64, receptors=8, ligands=16)
    sc.params.k_spring = 0.2
    with pytest.raises(ValueError):
        po.CpuPort(sc.params, sc)
