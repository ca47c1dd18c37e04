"""Halo transfer (fx_lbm_transfer_extract / fx_lbm_transfer_insert) through the
multi-domain driver fluidx3d_amd.domains.DomainLBM: a lattice split into domains with
one-cell halos, all on one GPU, gives the undivided lattice's DDFs and rho/u on every
cell, bit for bit (the reference's multi-GPU decomposition, src/lbm.cpp:721-773,
881-953, 1355-1383, with the transfer kernels of src/kernel.cpp:2049-2183)."""
import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu

VARIANTS = [(19, "SRT", "FP32"), (19, "TRT", "FP16S"), (27, "SRT", "FP16C"), (27, "TRT", "FP32")]
SPLITS = [(2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 2, 2), (3, 1, 1)]
EXT = ("EQUILIBRIUM_BOUNDARIES", "VOLUME_FORCE", "FORCE_FIELD")


@pytest.mark.parametrize("split", SPLITS, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("q,collision,storage", VARIANTS)
def test_domains_match_undivided_lattice(q, collision, storage, split):
    import fluidx3d_amd as fx
    from fluidx3d_amd.domains import DomainLBM
    Nx, Ny, Nz = 24, 18, 12
    N = Nx * Ny * Nz
    rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=q + sum(split), solid_fraction=0.08, equilibrium_faces=True)
    F = (1e-4 * np.random.default_rng(3).uniform(-1, 1, 3 * N)).astype(np.float32)
    f = (1e-5, -2e-5, 5e-6)
    one = fx.LBM(Nx, Ny, Nz, 0.05, *f, velocity_set=q, collision=collision, storage=storage, extensions=EXT)
    many = DomainLBM(Nx, Ny, Nz, *split, 0.05, *f, velocity_set=q, collision=collision, storage=storage, extensions=EXT)
    try:
        one.rho.host[:] = rho; one.u.host[:] = u; one.flags.host[:] = flags; one.F.host[:] = F
        many.rho[:] = rho; many.u[:] = u; many.flags[:] = flags; many.F[:] = F
        one.initialize()
        many.initialize()
        assert np.array_equal(one.read_ddfs(), many.global_ddfs()), "DDFs differ after initialize"
        for step in range(9):  # both parities, several round trips across every face
            one.run(1)
            many.run(1)
            a, b = one.read_ddfs(), many.global_ddfs()
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"step {step + 1}: {bad.size} DDF values differ, first slot/cell {divmod(int(bad[0]), N)}"
        one.rho.read_from_device(); one.u.read_from_device()
        r, v = many.read_fields()
        assert np.array_equal(one.rho.host, r) and np.array_equal(one.u.host, v)
    finally:
        one.close()
        many.close()
