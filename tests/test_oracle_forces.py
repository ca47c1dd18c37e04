"""Pins the oracle's force path and equilibrium boundaries with mathematics (CPU only).

- update_force_field + object_force (src/kernel.cpp:1873-1884, 1940-1955): in a periodic
  channel driven by a body force to steady state, the momentum the walls take per step is
  the momentum the force puts in, fx per fluid cell, split evenly between the two walls.
  A wrong factor in F = 2 sum c_i f_i, loading the solid cell's slots with the wrong
  parity, or a sign slip fails it.
- object_center_of_mass (:1909-1930): the mean cell position of a sphere or box is its
  centre, in the position() convention (src/kernel.cpp:824-826).
- object_torque (:1956-1970): the torque about a point p equals the torque about the
  centre c plus (c - p) x F_total, and vanishes about the centre of a mirror-symmetric
  body in a mirror-symmetric flow.  A flipped cross product or swapped operand fails it.
- TYPE_E emission (src/kernel.cpp:1484-1493, 1603, 1631): an equilibrium cell emits the
  textbook equilibrium of its preset (rho, u), which its neighbours load one step later,
  also after the presets change.
"""
import numpy as np
import pytest

from tests import scenes
from tests.test_oracle_physics import VS, feq_textbook


def channel(oracle, q, trt, storage, fx, Nx=4, Ny=16, Nz=4):
    """Periodic channel, TYPE_S wall at y=0 and TYPE_S|TYPE_X wall at y=Ny-1 (two markers,
    so each wall's force is summed on its own), body force fx."""
    L = oracle.OracleLBM(Nx, Ny, Nz, nu=1.0 / 6.0, q=q, trt=trt, storage=storage, ext=oracle.EXT_VOLUME_FORCE, fx=fx)
    _, y, _ = scenes.coords(Nx, Ny, Nz)
    L.flags[y == 0] = oracle.TYPE_S
    L.flags[y == Ny - 1] = oracle.TYPE_S | oracle.TYPE_X
    return L, int(((y > 0) & (y < Ny - 1)).sum())


@pytest.mark.parametrize("q,trt", [(19, False), (19, True), (27, False), (27, True)])
def test_steady_channel_wall_force_equals_the_body_force(oracle, q, trt):
    fx = 1e-5
    L, fluid = channel(oracle, q, trt, "FP64", fx)
    L.initialize()
    L.step(4000)  # slowest mode decays as exp(-nu (pi/14)^2 t): 3e-15 after 4000 steps
    bottom = L.object_force(oracle.TYPE_S)
    top = L.object_force(oracle.TYPE_S | oracle.TYPE_X)
    total = fx * fluid
    assert abs(bottom[0] + top[0] - total) < 1e-9 * total, (bottom, top, total)
    assert abs(bottom[0] - top[0]) < 1e-9 * total
    # wall-normal forces are equal and opposite, nothing along z
    assert abs(bottom[1] + top[1]) < 1e-12 and abs(bottom[2]) < 1e-15 and abs(top[2]) < 1e-15


@pytest.mark.parametrize("storage,rtol", [("FP32", 2e-3), ("FP16S", 2e-2)])
def test_steady_channel_wall_force_in_single_precision(oracle, storage, rtol):
    """The reference's FP32 arithmetic (the path the GPU kernels are compared with bit for
    bit) balances to the rounding of the stored DDFs: the wall-shear signal is ~1e-4 of
    the DDFs here, so FP32 storage is within 0.2 % and FP16S (11-bit mantissa) within 2 %."""
    fx = 1e-5
    L, fluid = channel(oracle, 19, False, storage, fx)
    L.initialize()
    L.step(4000)
    total = fx * fluid
    s = L.object_force(oracle.TYPE_S)[0] + L.object_force(oracle.TYPE_S | oracle.TYPE_X)[0]
    assert abs(s - total) < rtol * total, (s, total)


@pytest.mark.parametrize("q,trt", [(19, False), (19, True), (27, False), (27, True)])
def test_wall_force_is_the_momentum_the_lattice_loses_each_step(oracle, q, trt):
    """Exact momentum budget of one step, at any time (transient flow past an off-centre
    sphere in an oblique body force).  Let P(t) be the momentum of the populations every
    cell loads at step t (solid cells load their mailboxes).  During step t the fluid
    cells' populations gain f per fluid cell and move on; the populations in the solid
    mailboxes are loaded reversed by the fluid at step t+1 (full-way bounce-back), so
    P(t+1) = P(t) + N_fluid f - 2 P_solid(t), and the object force at step t is
    2 P_solid(t) = P(t) + N_fluid f - P(t+1).  Loading the solid cells' slots with the
    other parity, or any factor on the momentum, breaks it."""
    force = np.array([1e-5, 4e-6, -3e-6])
    dims = (24, 16, 16)
    L = oracle.OracleLBM(*dims, nu=0.05, q=q, trt=trt, storage="FP64", ext=oracle.EXT_VOLUME_FORCE,
                         fx=force[0], fy=force[1], fz=force[2])
    x, y, z = scenes.coords(*dims)
    solid = (x - 9) ** 2 + (y - 6) ** 2 + (z - 10) ** 2 <= 16
    L.flags[solid] = oracle.TYPE_S
    rho, u, _ = scenes.random_scene(*dims, seed=12, solid_fraction=0.0, u_amp=0.02, rho_amp=0.01)
    L.rho[:] = np.where(solid, 1.0, rho)  # solid cells at rest with rho = 1: zero shifted DDFs
    L.u[:] = u
    L.initialize()
    c = VS[q]

    def momentum():
        pops = L.populations()
        return np.array([(c[a][:, None] * pops).sum() for a in range(3)])
    checked = 0
    for t in range(40):
        if t in (0, 1, 5, 17, 39):
            P0 = momentum()
            F = L.object_force(oracle.TYPE_S)
            L.step(1)
            budget = P0 + int((~solid).sum()) * force - momentum()
            assert np.allclose(F, budget, rtol=0, atol=1e-13), (t, F, budget)
            assert np.abs(F).max() > 1e-4
            checked += 1
        else:
            L.step(1)
    assert checked == 5


def test_transient_wall_force_is_not_the_body_force(oracle):
    """Control for the steady-state test: early on, the walls have taken almost nothing."""
    fx = 1e-5
    L, fluid = channel(oracle, 19, False, "FP64", fx)
    L.initialize()
    L.step(20)
    s = L.object_force(oracle.TYPE_S)[0] + L.object_force(oracle.TYPE_S | oracle.TYPE_X)[0]
    assert s < 0.5 * fx * fluid


@pytest.mark.parametrize("storage", ["FP32", "FP64"])
def test_center_of_mass_is_the_body_centre(oracle, storage):
    dims = (20, 18, 16)
    L = oracle.OracleLBM(*dims, nu=0.1, storage=storage)
    x, y, z = scenes.coords(*dims)
    sphere = (x - 10) ** 2 + (y - 9) ** 2 + (z - 8) ** 2 <= 20.25  # R = 4.5, exact in integers
    L.flags[sphere] = oracle.TYPE_S | oracle.TYPE_X
    box = (x >= 2) & (x < 5) & (y >= 1) & (y < 8) & (z >= 11) & (z < 15)
    L.flags[box] = oracle.TYPE_S
    com, cells = L.object_center_of_mass(oracle.TYPE_S | oracle.TYPE_X)
    assert cells == int(sphere.sum())
    assert np.array_equal(com, [10 + 0.5 - 10, 9 + 0.5 - 9, 8 + 0.5 - 8])  # position(10, 9, 8)
    com, cells = L.object_center_of_mass(oracle.TYPE_S)
    assert cells == 3 * 7 * 4
    centre = [(2 + 4) / 2 + 0.5 - 10, (1 + 7) / 2 + 0.5 - 9, (11 + 14) / 2 + 0.5 - 8]
    assert np.allclose(com, centre, rtol=0, atol=1e-12)


def sphere_in_flow(oracle, storage, force, centre, steps):
    dims = (24, 16, 16)
    L = oracle.OracleLBM(*dims, nu=0.05, storage=storage, ext=oracle.EXT_VOLUME_FORCE, fx=force[0], fy=force[1], fz=force[2])
    x, y, z = scenes.coords(*dims)
    L.flags[(x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2 <= 16] = oracle.TYPE_S
    L.initialize()
    L.step(steps)
    pos = [c + 0.5 - d / 2 for c, d in zip(centre, dims)]  # position() of the centre cell
    return L, np.array(pos)


@pytest.mark.parametrize("storage,rtol", [("FP64", 1e-12), ("FP32", 1e-5)])
def test_torque_about_an_offset_point(oracle, storage, rtol):
    """tau(p) = tau(c) + (c - p) x F with an off-centre sphere in an oblique flow, so that
    F, tau(c) and (c - p) x F are all nonzero and not parallel."""
    L, c = sphere_in_flow(oracle, storage, (1e-5, 4e-6, -3e-6), (9, 6, 10), 150)
    F = L.object_force(oracle.TYPE_S)
    tc = L.object_torque(c, oracle.TYPE_S)
    p = c + np.array([1.5, -2.0, 3.0])
    tp = L.object_torque(p, oracle.TYPE_S)
    expect = tc + np.cross(c - p, F)
    scale = np.linalg.norm(F) * 4.0
    assert np.linalg.norm(np.cross(c - p, F)) > 0.5 * scale  # the check has teeth
    assert np.allclose(tp, expect, rtol=0, atol=rtol * scale), (tp, expect)


def test_torque_vanishes_for_a_symmetric_body(oracle):
    """Sphere at cell (12, 8, 8) of a periodic 24x16x16 box, flow along x: the flow is
    mirror symmetric in y and z about the sphere centre, so the torque about it is zero
    while the drag is not."""
    L, c = sphere_in_flow(oracle, "FP64", (1e-5, 0.0, 0.0), (12, 8, 8), 150)
    F = L.object_force(oracle.TYPE_S)
    t = L.object_torque(c, oracle.TYPE_S)
    assert F[0] > 0 and abs(F[1]) < 1e-12 * F[0] and abs(F[2]) < 1e-12 * F[0]
    assert np.linalg.norm(t) < 1e-12 * 4.0 * F[0]
    # about p = c + (0, 2, 0) the body lies at r_y = -2: tau_z = r_x F_y - r_y F_x = 2 F_x
    t_off = L.object_torque(c + np.array([0.0, 2.0, 0.0]), oracle.TYPE_S)
    assert abs(t_off[2] - 2.0 * F[0]) < 1e-12 * F[0] and abs(t_off[0]) < 1e-12 * F[0]


@pytest.mark.parametrize("q,trt", [(19, False), (19, True), (27, False), (27, True)])
@pytest.mark.parametrize("storage,atol", [("FP64", 1e-16), ("FP32", 1e-8)])
def test_equilibrium_cells_emit_the_equilibrium_of_their_preset(oracle, q, trt, storage, atol):
    """After a step, the neighbour at n + e_i of a TYPE_E cell n loads as population i
    exactly feq_i(rho_n, u_n) of the preset; the presets are changed after initialize, so
    what is checked is what stream_collide emits, not what initialize stored."""
    dims = (12, 8, 6)
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    rho, u, flags = scenes.random_scene(*dims, seed=8, solid_fraction=0.05, equilibrium_faces=True)
    L = oracle.OracleLBM(*dims, nu=0.05, q=q, trt=trt, storage=storage, ext=oracle.EXT_EQUILIBRIUM_BOUNDARIES)
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.initialize()
    rng = np.random.default_rng(9)
    e = np.flatnonzero(flags == oracle.TYPE_E)
    c = VS[q]
    for step in range(3):
        L.rho[e] = 1.0 + 0.02 * rng.uniform(-1, 1, e.size)
        for d in range(3):
            L.u[d * N + e] = 0.08 * rng.uniform(-1, 1, e.size)
        L.step(1)
        pops = L.populations()
        expect = feq_textbook(q, L.rho.astype(np.float64), L.u.astype(np.float64))
        x, y, z = scenes.coords(*dims)
        checked = 0
        for i in range(1, q):
            m = ((x[e] + c[0][i]) % Nx) + (((y[e] + c[1][i]) % Ny) + ((z[e] + c[2][i]) % Nz) * Ny) * Nx
            live = (flags[m] & oracle.TYPE_S) == 0
            got = pops[i][m[live]].astype(np.float64)
            want = expect[i][e[live]]
            if storage == "FP32":
                want = want.astype(np.float32).astype(np.float64)
            assert np.allclose(got, want, rtol=0, atol=atol), (step, i, np.abs(got - want).max())
            checked += int(live.sum())
        assert checked > 0.8 * e.size * (q - 1)
