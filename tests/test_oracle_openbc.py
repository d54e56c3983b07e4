"""Pins of the oracle's velocity (inflow/outflow) faces
(fluid3d_in_out_flow_boundary_kernel, LB_kernel_old_08102011.cu:586-1090):

- after a step without forcing, the populations of every face cell carry
  exactly the prescribed velocity (Zou/He reconstructs them for it and the
  collision keeps mass and momentum);
- a channel between two bounce-back walls fed at both z faces with the same
  duct profile relaxes to plane Poiseuille flow: the parabola everywhere and
  the pressure gradient dp/dz = -8 nu U / L^2 of the closed form."""
import numpy as np
import pytest


def channel(pfg, po, nx=2, ny=12, nz=24, tau=1.0, U=0.02):
    p = pfg.default_params(nx=nx, ny=ny, nz=nz)
    p.tau = tau
    p.wall[0], p.wall[1], p.wall[2] = pfg.WALL_PERIODIC, pfg.WALL_BOUNCE, pfg.WALL_VELOCITY
    p.wall_u_lo[2][2] = U
    p.wall_u_hi[2][2] = U
    return p, po.lattice(p)


def moments(f, Q=19):
    from pyoracle import load
    L = load()
    c = np.array([[L.orc_c(d, a) for a in range(3)] for d in range(Q)], dtype=float)
    rho = f.sum(0)
    j = c.T @ f
    return rho, j / rho


def test_face_cells_carry_the_prescribed_velocity(pfg, po):
    p, L = channel(pfg, po)
    rng = np.random.default_rng(1)
    f = po.init_eq(L) * (1 + 0.01 * rng.standard_normal((19, p.nx * p.ny * p.nz)))
    fout = po.lb_step(L, f)
    rho, u = moments(fout)
    y = np.arange(p.ny) + 0.5
    shape = 4 * y * (p.ny - y) / p.ny ** 2
    for z in (0, p.nz - 1):
        for yy in range(p.ny):
            for x in range(p.nx):
                i = x + p.nx * (yy + p.ny * z)
                want = float(np.float32(0.02)) * shape[yy]  # the FP32 parameter
                assert u[0, i] == pytest.approx(0.0, abs=1e-15)
                assert u[1, i] == pytest.approx(0.0, abs=1e-15)
                assert u[2, i] == pytest.approx(want, rel=1e-12, abs=1e-15)
                if yy in (0, p.ny - 1):  # next to a wall: feq(1, u)
                    assert rho[i] == pytest.approx(1.0, abs=1e-14)


def test_channel_relaxes_to_plane_poiseuille(pfg, po):
    """The cells next to the walls on the faces are reset to equilibrium,
    which lowers the flow rate by a few per cent at 12 cells across and by
    0.2 % at 24 (checked here)."""
    U, tau, ny = 0.01, 1.4, 24
    p, L = channel(pfg, po, ny=ny, tau=tau, U=U)
    f = po.init_eq(L)
    for _ in range(5000):
        f = po.lb_step(L, f)
    rho, ux, uy, uz = po.fields(L, f)
    nx, nz = p.nx, p.nz
    uz = uz.reshape(nz, ny, nx)
    rho = rho.reshape(nz, ny, nx)
    y = np.arange(ny) + 0.5
    U = float(np.float32(U))
    para = 4 * U * y * (ny - y) / ny ** 2
    mid = uz[nz // 2, :, 0]
    assert np.abs(mid - para).max() < 0.01 * U
    nu = (tau - 0.5) / 3.0
    pz = rho[:, ny // 2, 0] / 3.0
    slope = np.polyfit(np.arange(4, nz - 4), pz[4:nz - 4], 1)[0]
    assert slope == pytest.approx(-8 * nu * U / ny ** 2, rel=0.05)
