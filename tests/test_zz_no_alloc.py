"""fx_lbm_mass_sum_noalloc (the FP64 total-mass sum over a static device scratch array, for
callers that cannot allocate) against fx_lbm_mass_sum on the same lattices: bit-identical
results on two calls in a row, the second on a smaller lattice than the first, and a clear
FX_ERR_CONFIG for a lattice whose grid needs more partials than the array holds."""
import ctypes

import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu


def sums(lbm):
    from fluidx3d_amd import _abi
    L = lbm.ctx.lib
    plain, static = ctypes.c_double(), ctypes.c_double()
    _abi.check(L.fx_lbm_mass_sum(lbm.ctx.h, lbm.fi, lbm.flags.device, lbm.t, ctypes.byref(plain)))
    _abi.check(L.fx_lbm_mass_sum_noalloc(lbm.ctx.h, lbm.fi, lbm.flags.device, lbm.t, ctypes.byref(static)))
    return plain.value, static.value


@pytest.mark.parametrize("storage", ["FP32", "FP16S"])
def test_noalloc_mass_sum_equals_mass_sum_on_two_calls(storage):
    import fluidx3d_amd as fx
    results = []
    # first a larger lattice (2400 partials), then a smaller one (135): the second call must
    # not see the first call's partials left in the shared array
    for shape, seed in (((200, 48, 50), 3), ((40, 27, 5), 4)):
        Nx, Ny, Nz = shape
        rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=seed)
        lbm = fx.LBM(Nx, Ny, Nz, 0.05, storage=storage)
        try:
            lbm.rho.host[:] = rho
            lbm.u.host[:] = u
            lbm.flags.host[:] = flags
            lbm.run(7)
            plain, static = sums(lbm)
            # the same call again on the same state gives the same bits
            plain2, static2 = sums(lbm)
        finally:
            lbm.close()
        print(f"\n{storage} {Nx}x{Ny}x{Nz}: mass {plain!r} (static scratch {static!r})")
        assert np.isfinite(plain) and plain > 0.0
        assert static == plain and static2 == plain2 == plain
        results.append(plain)
    assert results[0] != results[1]


def test_noalloc_mass_sum_rejects_grids_beyond_the_scratch():
    from fluidx3d_amd import _abi
    L = _abi.load()
    cfg = _abi.Config(64, 1024, 1025, 19, _abi.FX_SRT, _abi.FX_FP32, 0, np.float32(1.0))
    ctx = ctypes.c_void_p()
    _abi.check(L.fx_ctx_create(0, ctypes.byref(cfg), ctypes.byref(ctx)))
    try:
        buf = ctypes.c_void_p()
        _abi.check(L.fx_buf_alloc(ctx, 4096, ctypes.byref(buf)))
        out = ctypes.c_double(-1.0)
        # 64-cell rows: one block per row, 1024*1025 = 1049600 blocks > 2^20 partials
        rc = L.fx_lbm_mass_sum_noalloc(ctx, buf, buf, 0, ctypes.byref(out))
        msg = L.fx_last_error().decode()
        assert rc == _abi.FX_ERR_CONFIG, (rc, msg)
        assert "1049600" in msg and "1048576" in msg, msg
        assert out.value == -1.0
    finally:
        L.fx_ctx_destroy(ctx)
