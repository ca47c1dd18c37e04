"""Pins the oracle's velocity clamp (src/kernel.cpp:1553-1560 in stream_collide,
1853-1855 in update_fields) with a numpy restatement of the moments (CPU only).

A coherent fast flow, u ~ (0.6, -0.58, 0) +- 0.1, drives many cells past def_c =
0.57735027 = 1/sqrt(3) (src/lbm.cpp:366) in every step.
- update_fields: u = clamp(j/rho, -def_c, def_c) per component of the loaded populations.
- stream_collide: with SRT at w = 1 and no force the collision output is the equilibrium
  itself, so the population a neighbour loads next is feq(rho, clamp(u)) of the sender.
Dropping the clamp, clamping the magnitude instead of each component, or clamping at
another bound fails one of them.
"""
import numpy as np
import pytest

from tests import scenes
from tests.test_oracle_physics import VS, feq_textbook

DEF_C = 0.57735027  # the FP64 oracle evaluates the scheme with the decimal constant


def fast_scene(Nx, Ny, Nz, seed=21, equilibrium_faces=False):
    rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=seed, u_amp=0.1, rho_amp=0.01, equilibrium_faces=equilibrium_faces)
    N = Nx * Ny * Nz
    u[:N] += np.float32(0.6)
    u[N:2 * N] -= np.float32(0.58)
    return rho, u, flags


def moments(q, pops):
    c = VS[q]
    rho = pops.sum(axis=0) + 1.0
    return rho, [(c[a][:, None] * pops).sum(axis=0) / rho for a in range(3)]


@pytest.mark.parametrize("q", [19, 27])
def test_update_fields_clamps_each_component(oracle, q):
    dims = (16, 8, 6)
    rho, u, flags = fast_scene(*dims)
    L = oracle.OracleLBM(*dims, nu=0.1, q=q, storage="FP64")
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.initialize()
    fluid = (flags & oracle.TYPE_S) == 0
    for _ in range(3):
        L.step(1)
        r, uu = moments(q, L.populations())
        L.update_fields()
        N = L.N
        clamped = 0
        for a in range(3):
            want = np.clip(uu[a], -DEF_C, DEF_C)
            got = L.u[a * N:(a + 1) * N]
            assert np.allclose(got[fluid], want[fluid], rtol=0, atol=1e-15)
            clamped += int((np.abs(uu[a][fluid]) > DEF_C).sum())
        assert np.allclose(L.rho[fluid], r[fluid], rtol=0, atol=1e-15)
        assert clamped > 10


@pytest.mark.parametrize("q", [19, 27])
def test_collision_uses_the_clamped_velocity(oracle, q):
    dims = (16, 8, 6)
    Nx, Ny, Nz = dims
    rho, u, flags = fast_scene(*dims)
    flags[:] = 0
    L = oracle.OracleLBM(*dims, nu=1.0 / 6.0, q=q, storage="FP64")
    assert L.w == 1.0
    L.rho[:] = rho
    L.u[:] = u
    L.initialize()
    L.step(1)
    c = VS[q]
    x, y, z = scenes.coords(*dims)
    for _ in range(3):
        pre = L.populations()
        r, uu = moments(q, pre)
        assert (np.abs(uu[0]) > DEF_C).sum() > 50
        feq = feq_textbook(q, r, np.concatenate([np.clip(v, -DEF_C, DEF_C) for v in uu]))
        L.step(1)
        post = L.populations()
        for i in range(q):
            m = ((x + c[0][i]) % Nx) + (((y + c[1][i]) % Ny) + ((z + c[2][i]) % Nz) * Ny) * Nx
            assert np.allclose(post[i][m], feq[i], rtol=0, atol=2e-15), i
