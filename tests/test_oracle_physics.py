"""Pins the oracle's LBM arithmetic and Esoteric-Pull bookkeeping with mathematics (CPU only).

Each check is chosen so that a plausible mistake fails it: a swapped slot or parity
breaks the streaming permutation, a wrong bounce-back mapping breaks the 2-step echo,
a wrong equilibrium or forcing moment breaks the uniform-state closed forms, a wrong
viscosity breaks the Taylor-Green decay rate and the Poiseuille profile.
"""
import numpy as np
import pytest

from tests import scenes

VS = {
    19: (np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0]),
         np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1]),
         np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1])),
    27: (np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1]),
         np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1]),
         np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1])),
}


def weights(q):
    c = VS[q]
    c2 = c[0] ** 2 + c[1] ** 2 + c[2] ** 2
    if q == 19:
        table = {0: 1 / 3, 1: 1 / 18, 2: 1 / 36}
    else:
        table = {0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}
    return np.array([table[int(v)] for v in c2])


def feq_textbook(q, rho, u):
    """Shifted second-order equilibrium w_i*(rho*(1 + 3cu + 4.5cu^2 - 1.5u^2) - 1), float64."""
    c, w = VS[q], weights(q)
    N = rho.size
    ux, uy, uz = u[:N], u[N:2 * N], u[2 * N:]
    out = np.empty((q, N))
    usq = ux * ux + uy * uy + uz * uz
    for i in range(q):
        cu = c[0][i] * ux + c[1][i] * uy + c[2][i] * uz
        out[i] = w[i] * (rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq) - 1)
    return out


def grid_roll(a, shift, dims):
    Nx, Ny, Nz = dims
    return np.roll(a.reshape(Nz, Ny, Nx), shift=(shift[2], shift[1], shift[0]), axis=(0, 1, 2)).ravel()


@pytest.mark.parametrize("q", [19, 27])
def test_weights_are_the_lattice_weights(q):
    c, w = VS[q], weights(q)
    assert abs(w.sum() - 1) < 1e-15
    for a in range(3):
        assert abs((w * c[a]).sum()) < 1e-15
        for b in range(3):
            assert abs((w * c[a] * c[b]).sum() - (1 / 3 if a == b else 0)) < 1e-15
    # opposite pairs i, i+1
    for i in range(1, q, 2):
        assert all(c[a][i] == -c[a][i + 1] for a in range(3))


@pytest.mark.parametrize("q,trt,storage", [(q, t, s) for q in (19, 27) for t in (False, True) for s in ("FP32", "FP16S", "FP16C", "FP64")])
def test_rest_state_is_a_fixed_point(oracle, q, trt, storage):
    L = oracle.OracleLBM(10, 8, 6, nu=0.05, q=q, trt=trt, storage=storage)
    L.flags[::7] = oracle.TYPE_S
    L.initialize()
    L.step(9)
    assert np.all(L.ddf_values() == 0)


@pytest.mark.parametrize("q", [19, 27])
@pytest.mark.parametrize("storage", ["FP32", "FP64"])
def test_uniform_flow_is_invariant(oracle, q, storage):
    """A uniform state stays exactly uniform in every slot (every cell computes the same
    thing and streaming only permutes equal values) and keeps its velocity (a swapped slot
    pair would reverse it)."""
    L = oracle.OracleLBM(12, 10, 8, nu=1.0 / 6.0, q=q, storage=storage)
    assert L.w == 1.0
    L.u[:L.N] = 0.03
    L.u[L.N:2 * L.N] = -0.02
    L.u[2 * L.N:] = 0.01
    L.initialize()
    p0 = L.populations()
    for _ in range(5):
        L.step(1)
        per_slot = L.fi.reshape(q, L.N)
        assert np.all(per_slot == per_slot[:, :1])
    assert np.allclose(L.populations(), p0, rtol=1e-5, atol=1e-9)
    L.update_fields()
    assert np.allclose(L.u[:L.N], 0.03, rtol=1e-6) and np.allclose(L.rho, 1.0, atol=1e-6)


@pytest.mark.parametrize("q", [19, 27])
def test_pure_streaming_moves_each_population_by_its_velocity(oracle, q):
    """w=0 makes the collision the identity; then population i that cell n loads at step t
    must equal the initial population i of cell n - (t+1)*e_i (periodic), for both
    parities (initialize stores with the odd-step convention, src/kernel.cpp:1429, so the
    first load is already one streaming step)."""
    dims = (9, 7, 5)
    N = 9 * 7 * 5
    rho, u, flags = scenes.random_scene(*dims, seed=3, solid_fraction=0.0, u_amp=0.1, rho_amp=0.05)
    L = oracle.OracleLBM(*dims, w=0.0, q=q, storage="FP64")
    L.rho[:] = rho
    L.u[:] = u
    L.initialize()
    f0 = feq_textbook(q, rho.astype(np.float64), u.astype(np.float64))
    c = VS[q]
    for t in range(0, 6):
        pops = L.populations()
        for i in range(q):
            k = t + 1
            expect = grid_roll(f0[i], (c[0][i] * k, c[1][i] * k, c[2][i] * k), dims)
            assert np.allclose(pops[i], expect, rtol=0, atol=1e-15), (t, i)
        L.step(1)


def test_full_way_bounce_back_timing(oracle):
    """Solid cells are mailboxes.  With w=0, a population arriving at fluid cell n from a
    solid neighbour is: at t=0 the solid cell's own initial equilibrium, at t=1 cell n's own
    initial opposite population, and for t>=2 cell n's opposite population of step t-2
    (the population sent toward the wall two steps earlier, reversed)."""
    q = 19
    dims = (10, 6, 4)
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    rho, u, flags = scenes.random_scene(*dims, seed=5, solid_fraction=0.15, u_amp=0.08, rho_amp=0.05)
    L = oracle.OracleLBM(*dims, w=0.0, q=q, storage="FP64")
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.initialize()
    solid = (flags & 1) == 1
    u0 = u.astype(np.float64).copy()
    u0[np.concatenate([solid, solid, solid])] = 0.0  # initialize zeroes u of solid cells
    f0 = feq_textbook(q, rho.astype(np.float64), u0)
    c = VS[q]
    opp = [0] + [i + 1 if i % 2 else i - 1 for i in range(1, q)]
    history = []
    for t in range(7):
        pops = L.populations()
        history.append(pops)
        for i in range(1, q):
            src = grid_roll(np.arange(N), (c[0][i], c[1][i], c[2][i]), dims)  # index of n - e_i
            from_solid = solid[src] & ~solid
            from_fluid = ~solid[src] & ~solid
            if t == 0:
                expect = f0[i][src]
            else:
                expect = history[t - 1][i][src]
            assert np.allclose(pops[i][from_fluid], expect[from_fluid], atol=1e-15), (t, i)
            if t == 0:
                e_s = f0[i][src]
            elif t == 1:
                e_s = f0[opp[i]]
            else:
                e_s = history[t - 2][opp[i]]
            assert np.allclose(pops[i][from_solid], e_s[from_solid], atol=1e-15), (t, i)
        L.step(1)


@pytest.mark.parametrize("q,trt", [(19, False), (19, True), (27, False), (27, True)])
def test_mass_and_momentum_conservation(oracle, q, trt):
    dims = (16, 12, 10)
    N = 16 * 12 * 10
    rho, u, flags = scenes.random_scene(*dims, seed=9, solid_fraction=0.1)
    L = oracle.OracleLBM(*dims, nu=0.02, q=q, trt=trt, storage="FP64")
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.initialize()
    m0 = L.mass()
    L.step(50)
    assert abs(L.mass() - m0) < 1e-12 * m0
    # all-fluid periodic box: total momentum is conserved too
    P = oracle.OracleLBM(*dims, nu=0.02, q=q, trt=trt, storage="FP64")
    P.rho[:] = rho
    P.u[:] = u
    P.initialize()
    c = VS[q]

    def momentum():
        pops = P.populations()
        return np.array([(c[a][:, None] * pops).sum() for a in range(3)])
    p0 = momentum()
    P.step(40)
    assert np.allclose(momentum(), p0, atol=1e-12)


@pytest.mark.parametrize("q,trt", [(19, False), (27, True)])
def test_uniform_body_force_adds_momentum_exactly(oracle, q, trt):
    """Guo forcing: the post-collision momentum of a uniform state is j + F per step."""
    F = np.array([2e-5, -1e-5, 5e-6])
    L = oracle.OracleLBM(6, 5, 4, nu=0.04, q=q, trt=trt, storage="FP64", ext=oracle.EXT_VOLUME_FORCE, fx=F[0], fy=F[1], fz=F[2])
    L.u[:L.N] = 0.01
    L.initialize()
    c = VS[q]
    for t in range(1, 30):
        L.step(1)
        pops = L.populations()
        j = np.array([(c[a][:, None] * pops).sum(axis=0) for a in range(3)])
        assert np.allclose(j, (np.array([0.01, 0, 0]) + t * F)[:, None], rtol=0, atol=1e-14)


@pytest.mark.parametrize("q,trt", [(19, False), (19, True), (27, False)])
def test_uniform_brinkman_decay_closed_form(oracle, q, trt):
    """Porosity alpha on a uniform state: j_{t+1} = j_t (1 - alpha/2)/(1 + alpha/2)."""
    a = 0.3
    L = oracle.OracleLBM(6, 5, 4, nu=0.1, q=q, trt=trt, storage="FP64", ext=oracle.EXT_POROSITY)
    L.alpha[:] = a
    L.u[:L.N] = 0.04
    L.u[L.N:2 * L.N] = -0.02
    L.initialize()
    c = VS[q]
    r = (1 - a / 2) / (1 + a / 2)
    for t in range(1, 20):
        L.step(1)
        pops = L.populations()
        jx = (c[0][:, None] * pops).sum(axis=0)
        jy = (c[1][:, None] * pops).sum(axis=0)
        assert np.allclose(jx, 0.04 * r ** t, rtol=1e-12), t
        assert np.allclose(jy, -0.02 * r ** t, rtol=1e-12), t


def test_porosity_zero_is_bit_identical_to_volume_force(oracle):
    dims = (12, 10, 8)
    rho, u, flags = scenes.random_scene(*dims, seed=2)
    out = []
    for ext in (oracle.EXT_VOLUME_FORCE, oracle.EXT_POROSITY):
        L = oracle.OracleLBM(*dims, nu=0.03, storage="FP32", ext=ext, fx=np.float32(1e-5), fy=np.float32(-2e-6), fz=0.0)
        L.rho[:] = rho
        L.u[:] = u
        L.flags[:] = flags
        L.initialize()
        L.step(5)
        out.append(L.fi.copy())
    assert np.array_equal(out[0], out[1])


def test_equilibrium_boundary_cells_keep_their_state(oracle):
    dims = (12, 8, 6)
    rho, u, flags = scenes.random_scene(*dims, seed=4, solid_fraction=0.0, equilibrium_faces=True)
    L = oracle.OracleLBM(*dims, nu=0.05, storage="FP64", ext=oracle.EXT_EQUILIBRIUM_BOUNDARIES)
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.initialize()
    L.step(5)
    L.update_fields()
    e = flags == oracle.TYPE_E
    assert np.array_equal(L.rho[e], rho[e].astype(np.float64))
    assert np.array_equal(L.u[np.concatenate([e, e, e])], u[np.concatenate([e, e, e])].astype(np.float64))


@pytest.mark.parametrize("trt", [False, True])
def test_poiseuille_channel_profile(oracle, refhelpers, trt):
    """Plane Poiseuille flow between TYPE_S walls at y=0 and y=Ny-1 (walls half-way to the
    first fluid cell), body force from the reference's closed form
    f = 2*u*rho*nu/R^2 (src/units.hpp:110)."""
    Nx, Ny, Nz = 4, 34, 4
    R = (Ny - 2) / 2.0
    nu = 1.0 / 6.0
    umax = 0.01
    f = refhelpers.ref_f_from_u_Poiseuille_2D(umax, 1.0, nu, R)
    L = oracle.OracleLBM(Nx, Ny, Nz, nu=nu, trt=trt, storage="FP64", ext=oracle.EXT_VOLUME_FORCE, fx=f)
    _, y, _ = scenes.coords(Nx, Ny, Nz)
    L.flags[(y == 0) | (y == Ny - 1)] = oracle.TYPE_S
    L.initialize()
    L.step(30000)
    L.update_fields()
    fluid = (y > 0) & (y < Ny - 1)
    yy = y[fluid] + 0.5 - Ny / 2.0  # distance from the channel centre
    uref = umax * (R * R - yy * yy) / (R * R)
    unum = L.u[:L.N][fluid]
    err = np.sqrt(np.sum((unum - uref) ** 2) / np.sum(uref ** 2))
    assert err < (1e-3 if trt else 3e-3), err


def test_taylor_green_decay_rate(oracle):
    N = 32
    rho, u, _ = scenes.taylor_green(N, N, N, u0=0.02)
    L = oracle.OracleLBM(N, N, N, nu=0.05, storage="FP64")
    L.rho[:] = rho
    L.u[:] = u
    L.initialize()
    L.step(50)
    L.update_fields()
    ke1 = scenes.kinetic_energy(L.rho, L.u, N ** 3)
    L.step(150)
    L.update_fields()
    ke2 = scenes.kinetic_energy(L.rho, L.u, N ** 3)
    rate = np.log(ke1 / ke2) / 150
    expect = 2 * 0.05 * 3 * (2 * np.pi / N) ** 2
    assert abs(rate / expect - 1) < 0.02
