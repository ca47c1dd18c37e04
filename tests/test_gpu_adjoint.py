"""GPU adjoint (fx_lbm_adjoint_step / fx_lbm_forward / fx_lbm_backward) against the FP64
oracle adjoint and finite differences.

The oracle's adjoint step takes its derivatives by forward-mode dual numbers through the
oracle's own collision code; the GPU kernel uses the hand-derived transposed collision
(weighted moment sums).  Agreement of the two pins both.  FP32 (GPU) against FP64
(oracle): tolerances are stated per test.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXT = {"VOLUME_FORCE": 1, "FORCE_FIELD": 2, "EQUILIBRIUM_BOUNDARIES": 4, "UPDATE_FIELDS": 8, "POROSITY": 16,
       "PRECISE_DENSITY": 32}
# The gradient runs use PRECISE_DENSITY: at body forces of 1e-5..1e-6 per step from rest,
# the reference's FP32 density rounding alone moves the objective by 0.05 %..3 % against
# FP64 (tests/test_oracle_precision.py), which would swamp a 1e-3 gradient check.
ADJ_EXT = ("VOLUME_FORCE", "POROSITY", "PRECISE_DENSITY")


def bits(names):
    return sum(EXT[n] for n in names)


# 20x12x10: rows of one block; 300x4x3: rows of five 64-cell blocks (XCD-mapped grid)
@pytest.mark.parametrize("grid", [(20, 12, 10), (300, 4, 3)], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("q,collision", [(19, "SRT"), (19, "TRT"), (27, "SRT"), (27, "TRT")])
@pytest.mark.parametrize("variant", ["porosity", "all", "clamped"])
def test_adjoint_step_matches_fp64_oracle(oracle, q, collision, variant, grid):
    """One record pass and one adjoint step against the FP64 oracle's adjoint step; the
    record pass advances the DDFs exactly as a plain stream_collide does.  "clamped" runs
    a fast flow (u ~ (0.6, -0.58, 0) +- 0.1) in which many cells take the velocity clamp,
    so the adjoint's clamp mask is 0 there (the clamped components do not depend on the
    populations; the oracle's dual numbers carry a zero derivative through the clamp)."""
    import fluidx3d_amd as fx
    from fluidx3d_amd import _abi
    Nx, Ny, Nz = grid
    N = Nx * Ny * Nz
    names = ("VOLUME_FORCE", "POROSITY") if variant == "porosity" else ("VOLUME_FORCE", "POROSITY", "FORCE_FIELD", "EQUILIBRIUM_BOUNDARIES", "PRECISE_DENSITY")
    if variant == "clamped":
        from tests.test_oracle_clamp import fast_scene
        rho, u, flags = fast_scene(Nx, Ny, Nz, seed=5)
    else:
        rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=5, solid_fraction=0.08, equilibrium_faces=variant == "all")
    rng = np.random.default_rng(11)
    alpha = rng.uniform(0, 1, N).astype(np.float32)
    if variant == "clamped":  # weak damping keeps the flow fast
        alpha *= np.float32(0.02)
    F = (1e-4 * rng.uniform(-1, 1, 3 * N)).astype(np.float32)
    f = (2e-5, -1e-5, 5e-6)
    lbm = fx.LBM(Nx, Ny, Nz, 0.05, *f, velocity_set=q, collision=collision, extensions=names)
    plain = fx.LBM(Nx, Ny, Nz, 0.05, *f, velocity_set=q, collision=collision, extensions=names)
    for L_ in (lbm, plain):
        L_.rho.host[:] = rho; L_.u.host[:] = u; L_.flags.host[:] = flags; L_.porosity.host[:] = alpha
        if L_.F is not None:
            L_.F.host[:] = F
    try:
        plain.run(4)
        lbm.run(3)  # state before step t=3
        t = 3
        state = lbm.read_ddfs().astype(np.float64)
        L = lbm.ctx.lib
        moments = lbm.ctx.alloc(16 * N)
        _abi.check(L.fx_lbm_stream_collide_record(lbm.ctx.h, lbm.fi, lbm.rho.device, lbm.u.device, lbm.flags.device, t, *f,
                                                  lbm._ptr(lbm.F), lbm.porosity.device, moments))
        assert np.array_equal(lbm.read_ddfs(), plain.read_ddfs()), "record pass differs from stream_collide"
        lam1 = rng.standard_normal(q * N).astype(np.float32)
        lam = lbm.ctx.alloc(4 * q * N)
        grad = lbm.ctx.alloc(4 * N)
        lbm.ctx.write(lam, lam1)
        lbm.ctx.fill(grad, 4 * N, 0)
        plane_x, wobj = Nx - 2, 0.25
        _abi.check(L.fx_lbm_adjoint_step(lbm.ctx.h, lam, lbm.rho.device, lbm.u.device, lbm.flags.device, t, *f,
                                         lbm._ptr(lbm.F), lbm.porosity.device, moments, plane_x, wobj, grad))
        g_lam = np.empty(q * N, np.float32); g_grad = np.empty(N, np.float32)
        lbm.ctx.read(lam, g_lam); lbm.ctx.read(grad, g_grad)
        ref = oracle.OracleLBM(Nx, Ny, Nz, w=lbm.w, q=q, trt=collision == "TRT", storage="FP64", ext=bits(names), fx=f[0], fy=f[1], fz=f[2])
        ref.rho[:] = rho; ref.u[:] = u; ref.flags[:] = flags; ref.alpha[:] = alpha
        if ref.F is not None:
            ref.F[:] = F
        o_lam = lam1.astype(np.float64)
        o_grad = np.zeros(N)
        rc = oracle.lib().fxo_adjoint_step(ctypes.byref(ref.cfg), oracle._ptr(state), oracle._ptr(o_lam), ref.flags_ptr, t, f[0], f[1], f[2],
                                           oracle._ptr(ref.F), oracle._ptr(ref.alpha), plane_x, wobj, oracle._ptr(o_grad), oracle._ptr(ref.rho), oracle._ptr(ref.u))
        assert rc == 0
        err = np.max(np.abs(g_lam - o_lam)) / np.max(np.abs(o_lam))
        gerr = np.max(np.abs(g_grad - o_grad)) / np.max(np.abs(o_grad))
        print(f"\nq={q} {collision} {variant}: max rel lambda err {err:.2e}, grad err {gerr:.2e}")
        assert err < 2e-6 and gerr < 2e-5
    finally:
        lbm.close()
        plain.close()


def small_channel(fx_value=1e-5, dims=(32, 16, 12)):
    Nx, Ny, Nz = dims
    rho, u, flags, alpha = scenes.porous_channel(Nx, Ny, Nz)
    return (Nx, Ny, Nz), rho, u, flags, alpha


def gpu_gradient(dims, rho, u, flags, alpha, steps, window, plane, k, fxv=1e-5):
    import fluidx3d_amd as fx
    from fluidx3d_amd.adjoint import Adjoint
    lbm = fx.LBM(*dims, 1.0 / 6.0, fxv, 0.0, 0.0, extensions=ADJ_EXT)
    lbm.rho.host[:] = rho; lbm.u.host[:] = u; lbm.flags.host[:] = flags; lbm.porosity.host[:] = alpha
    adj = Adjoint(lbm, steps, window, plane, k)
    J = adj.forward()
    g = adj.backward()
    adj.close()
    lbm.close()
    return J, g


# 32x16x12: rows of one block; 260x8x6: rows of five 64-cell blocks (XCD-mapped grid)
@pytest.mark.parametrize("grid", [(32, 16, 12), (260, 8, 6)], ids=lambda g: "x".join(map(str, g)))
def test_gradient_matches_fp64_adjoint(oracle, grid):
    """Porous channel, 300 steps, objective over the last 100: checkpointed GPU gradient vs
    the FP64 oracle adjoint over the stored trajectory."""
    dims, rho, u, flags, alpha = small_channel(dims=grid)
    steps, window, plane = 300, 200, dims[0] - 1
    J, g = gpu_gradient(dims, rho, u, flags, alpha, steps, window, plane, 30)
    ref = oracle.OracleLBM(*dims, nu=1.0 / 6.0, storage="FP64", ext=bits(ADJ_EXT), fx=np.float32(1e-5))
    ref.rho[:] = rho; ref.u[:] = u; ref.flags[:] = flags; ref.alpha[:] = alpha
    ref.initialize()
    J64, g64 = ref.adjoint_gradient(steps, window, plane)
    porous = alpha > 0
    rel_l2 = np.linalg.norm(g[porous] - g64[porous]) / np.linalg.norm(g64[porous])
    print(f"\nJ gpu {J:.10e} fp64 {J64:.10e}; gradient relative L2 error {rel_l2:.2e}")
    assert abs(J - J64) < 1e-5 * abs(J64)
    assert rel_l2 < 5e-5
    big = np.abs(g64) > 1e-2 * np.abs(g64).max()
    assert np.max(np.abs(g[big] - g64[big]) / np.abs(g64[big])) < 1e-3


def test_checkpoint_interval_does_not_change_the_gradient():
    """Recomputation replays the forward bit for bit, so every schedule gives the same
    gradient: k=7, k=ceil(sqrt(T)), one segment."""
    dims, rho, u, flags, alpha = small_channel()
    steps, window, plane = 150, 100, dims[0] - 1
    grads = [gpu_gradient(dims, rho, u, flags, alpha, steps, window, plane, k)[1] for k in (7, 0, steps)]
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_channel_128_gradient_matches_finite_differences():
    """128x64x64 channel, 2000 steps, outlet momentum averaged over the last 500 steps,
    checkpoints every 100: GPU adjoint vs central differences of the FP64 oracle on 32
    seeded cells (tests/golden/adjoint_channel_fd.json, made by tests/golden/make_adjoint_fd.py)."""
    path = os.path.join(ROOT, "tests", "golden", "adjoint_channel_fd.json")
    if not os.path.exists(path):
        pytest.skip("tests/golden/adjoint_channel_fd.json not generated")
    gold = json.load(open(path))
    c = gold["config"]
    rho, u, flags, alpha = scenes.porous_channel(c["Nx"], c["Ny"], c["Nz"])
    J, g = gpu_gradient((c["Nx"], c["Ny"], c["Nz"]), rho, u, flags, alpha, c["steps"], c["window_begin"], c["plane_x"],
                        c["checkpoint_interval"], c["fx"])
    fd = np.array(gold["fd"])
    ad = g[np.array(gold["cells"])].astype(np.float64)
    rel = np.abs(ad - fd) / np.abs(fd)
    print(f"\nJ gpu {J:.10e} fp64 {gold['J']:.10e}; FD relative error max {rel.max():.2e} median {np.median(rel):.2e}")
    assert abs(J - gold["J"]) < 1e-4 * abs(gold["J"])
    assert rel.max() < 1e-3


def test_autograd_wrapper_gives_the_abi_gradient():
    """The torch.autograd wrapper (harness tests/torch_gradcheck.py, own process so that
    torch starts its GPU runtime first): same gradient as fx_lbm_backward, and within 1e-3
    of FP64 central differences."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "torch_gradcheck.py")], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"\nautograd == ABI: {out['autograd_equals_abi']}, FD relative error max {out['fd_rel_err_max']:.2e}")
    assert out["autograd_equals_abi"]
    assert out["fd_rel_err_max"] < 1e-3
