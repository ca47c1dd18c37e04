"""Full-size workloads on the GPU: the 256^3 Taylor-Green vortex in every storage mode
against the FP64 oracle, the 512x256x256 D3Q27 sphere over its full 20000 steps, the
1536x768x768 (906 M cells) D3Q19 FP16S sphere, and the checkpointed adjoint of the
384x192x192 porous channel over 5000 steps.  Each test prints the figures it checks."""
import os

import numpy as np
import pytest

from tests import scenes
from tests.test_gpu_parity import compare_state
from tests.test_host_cpp import sphere_run

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN_TGV256 = os.path.join(ROOT, "tests", "golden", "tgv256_fp64.npz")

# Error bounds against the FP64 oracle of the same scheme, relative to u0 for u and absolute
# for rho, for each DDF storage mode (the FP32 arithmetic is the same in all three).  With
# FP32 arithmetic the reference's scheme forms rho = sum(f) + 1 in float, which rounds the
# density deviation (~3 u0^2 / 16 = 5e-4 here) to 2^-24, about 1e-4 of it: the pressure
# field, and through it u, carries errors of order 1e-5 u0 after 1000 steps (measured
# 1.5e-5 u0 at 256^3, 2.6e-6 u0 at 64^3).  The 16-bit formats round each stored DDF to
# 2^-11 (FP16S, 10 stored mantissa bits) and 2^-12 (FP16C, 11 bits) of the shifted DDF
# f - w, whose size is set by the flow (here |f - w| <= w * 3 u0 ~ 1e-2).  Every mode
# also rounds each collision's outputs, which lets the total density drift slowly against
# FP64 (FP32: 4e-6 of it over 10000 steps, a uniform offset that dominates the rho error
# there; FP16C 5e-6, FP16S 1e-5).
TGV_BOUNDS = {"FP32": (5e-5, 1e-5), "FP16S": (2e-3, 2e-5), "FP16C": (1e-3, 1e-5)}
TGV_MASS_DRIFT = {"FP32": 1e-5, "FP16S": 1e-4, "FP16C": 1e-4}


@pytest.mark.parametrize("storage", ["FP32", "FP16S", "FP16C"])
def test_taylor_green_256_storage_error_against_fp64(storage):
    """D3Q19 SRT 256^3 Taylor-Green vortex (nu = 1/6, u0 = 0.05), 10000 steps, in each DDF
    storage mode.  rho and u at 4096 fixed cells, the kinetic energy and the mass after
    1000 and 10000 steps against the FP64 oracle's run of the same inputs
    (tests/golden/tgv256_fp64.npz, made by tests/golden/make_tgv256_fp64.py).  States the
    error of each mode and checks it against TGV_BOUNDS."""
    import fluidx3d_amd as fx
    if not os.path.exists(GOLDEN_TGV256):
        pytest.skip("tests/golden/tgv256_fp64.npz not generated")
    gold = np.load(GOLDEN_TGV256)
    n = int(gold["size"])
    N = n ** 3
    u0 = float(gold["u0"])
    cells = gold["cells"]
    rho, u, flags = scenes.taylor_green(n, n, n, u0=u0)
    lbm = fx.LBM(n, n, n, 1.0 / 6.0, storage=storage)
    try:
        assert lbm.w == np.float32(gold["w"])
        lbm.rho.host[:] = rho
        lbm.u.host[:] = u
        lbm.flags.host[:] = flags
        lbm.initialize()
        m_init = lbm.mass()
        bu, brho = TGV_BOUNDS[storage]
        for s in (int(v) for v in gold["steps"]):
            lbm.run(s - lbm.t)
            lbm.rho.read_from_device()
            lbm.u.read_from_device()
            r = lbm.rho.host.astype(np.float64)
            v = lbm.u.host.astype(np.float64).reshape(3, N)
            assert np.isfinite(v).all()
            err_u = np.abs(v[:, cells] - gold[f"u_{s}"]).max()
            err_rho = np.abs(r[cells] - gold[f"rho_{s}"]).max()
            ke = 0.5 * np.sum(r * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
            err_ke = abs(ke / float(gold[f"ke_{s}"]) - 1.0)
            err_mass = abs(np.sum(r) / float(gold[f"mass_{s}"]) - 1.0)
            u_max = np.abs(gold[f"u_{s}"]).max()
            print(f"\n{storage} t={s}: max|u-u64| = {err_u:.3e} ({err_u / u0:.2e} u0, {err_u / u_max:.2e} of max|u64| "
                  f"{u_max:.3e}), max|rho-rho64| = {err_rho:.3e}, kinetic energy rel. error {err_ke:.2e}, "
                  f"sum(rho) rel. error {err_mass:.2e}")
            assert err_u < bu * u0, (err_u, bu * u0)
            assert err_rho < brho, (err_rho, brho)
        # the DDF mass invariant holds in every storage mode up to the format's rounding
        drift = abs(lbm.mass() - m_init) / m_init
        print(f"{storage}: FP64 DDF mass drift over 10000 steps {drift:.2e}")
        assert drift < TGV_MASS_DRIFT[storage]
    finally:
        lbm.close()


# Drag coefficient of a sphere at Re = 100 in unbounded flow: C_d = 24/Re (1 + 0.1935
# Re^0.6305) = 1.087 (Clift, Grace and Weber, "Bubbles, Drops and Particles", 1978, the
# standard drag curve for 20 <= Re <= 260).  The 512x256x256 domain confines the sphere
# (diameter 64 in a periodic 256x256 cross-section, 4.9 % of its area) and places the
# equilibrium inlet two diameters upstream; both raise the drag, never lower it.  The band
# admits up to +25 % for them and fails a factor of 2 in the momentum exchange, a sign, or a
# missing term of the force sum.
CD_UNBOUNDED_RE100 = 1.087
CD_BAND = (CD_UNBOUNDED_RE100, 1.25 * CD_UNBOUNDED_RE100)


def test_sphere_512x256x256_full_run_drag(tmp_path):
    """D3Q27 TRT flow past a sphere (R = 32 at (128,128,128), TYPE_E inflow/outflow u = 0.075
    at x = 0 and 511, periodic in y and z, Re = 100, 1e-4 seeded perturbation), the whole
    20000 steps in FP16C storage (the C++ scene fluidx3d_amd/host/bin/sphere_20k) and the
    same run in FP32 storage (sphere_20k_fp32).  The FP64 drag: C_d inside CD_BAND, settled
    (the last two reports, 1000 steps apart, within 0.5 %), no lateral force beyond 1e-3 of
    the drag, and FP16C within 1 % of FP32."""
    results = {}
    for name in ("sphere_20k", "sphere_20k_fp32"):
        drags, mass, out = sphere_run(tmp_path, name, timeout=900)
        assert drags[-1][0] == 20000 and len(drags) == 20
        t, Fx, Fy, Fz, Cd = drags[-1]
        Cd_prev = drags[-2][4]
        print(f"\n{name}: t={t} F=({Fx:.6e}, {Fy:.3e}, {Fz:.3e}) C_d={Cd:.5f} (t={drags[-2][0]}: {Cd_prev:.5f}); "
              f"C_d every 1000 steps: {' '.join(f'{d[4]:.4f}' for d in drags)}; mass relative change {mass[2]:.2e}")
        assert mass[3] == 0
        assert CD_BAND[0] < Cd < CD_BAND[1], (Cd, CD_BAND)
        assert abs(Cd / Cd_prev - 1.0) < 5e-3
        assert abs(Fy) < 1e-3 * Fx and abs(Fz) < 1e-3 * Fx
        assert abs(mass[2]) < 1e-3
        results[name] = Cd
    rel = abs(results["sphere_20k"] / results["sphere_20k_fp32"] - 1.0)
    print(f"C_d FP16C {results['sphere_20k']:.6f} against FP32 {results['sphere_20k_fp32']:.6f}: {rel:.2e}")
    assert rel < 1e-2


def sphere_scene(Nx, Ny, Nz, R, u_in=0.075):
    """Flags and velocity of the sphere scenes, without the seeded perturbation: TYPE_S|TYPE_X
    sphere() of radius R at (4R, 4R, 4R), TYPE_E inflow/outflow planes x = 0 and Nx-1."""
    c = 4.0 * R
    flags = scenes.sphere_flags(Nx, Ny, Nz, (c, c, c), R, scenes.TYPE_S | scenes.TYPE_X)
    x, _, _ = scenes.coords(Nx, Ny, Nz)
    edge = ((x == 0) | (x == Nx - 1)) & (flags == 0)
    flags[edge] = scenes.TYPE_E
    N = Nx * Ny * Nz
    u = np.zeros(3 * N, dtype=np.float32)
    u[:N] = np.where(flags == 0, u_in, 0.0)
    u[:N][edge] = u_in
    return np.ones(N, dtype=np.float32), u, flags


@pytest.mark.parametrize("storage", ["FP32", "FP16S", "FP16C"])
def test_sphere_128x64x64_drag_series_matches_oracle(oracle, storage):
    """The sphere scene at a quarter of the 512x256x256 lattice (R = 8, D3Q19 SRT,
    EQUILIBRIUM_BOUNDARIES + FORCE_FIELD, nu for Re = 100): every 50 steps up to 200, the
    GPU's FP64 drag sum equals the oracle's to 1e-12 (the force fields are identical; only
    the order of the FP64 sums differs), and the DDFs are identical at the end."""
    import fluidx3d_amd as fx
    Nx, Ny, Nz, R, u_in = 128, 64, 64, 8.0, 0.075
    rho, u, flags = sphere_scene(Nx, Ny, Nz, R, u_in)
    lbm = fx.LBM(Nx, Ny, Nz, u_in * 2.0 * R / 100.0, storage=storage, extensions=("EQUILIBRIUM_BOUNDARIES", "FORCE_FIELD"))
    ref = oracle.OracleLBM(Nx, Ny, Nz, w=lbm.w, storage=storage, ext=oracle.EXT_EQUILIBRIUM_BOUNDARIES | oracle.EXT_FORCE_FIELD)
    marker = scenes.TYPE_S | scenes.TYPE_X
    try:
        for t_ in (lbm, ref):
            (t_.rho.host if hasattr(t_.rho, "host") else t_.rho)[:] = rho
            (t_.u.host if hasattr(t_.u, "host") else t_.u)[:] = u
            (t_.flags.host if hasattr(t_.flags, "host") else t_.flags)[:] = flags
        lbm.initialize()
        ref.initialize()
        series = []
        for _ in range(4):
            lbm.run(50)
            ref.step(50)
            ref.update_force_field()
            F, F_ref = lbm.object_force(marker), ref.object_force(marker)
            assert np.allclose(F, F_ref, rtol=1e-12, atol=1e-12 * abs(F_ref[0])), (lbm.t, F, F_ref)
            series.append(F[0] / (0.5 * u_in ** 2 * np.pi * R * R))
        compare_state(lbm, ref, oracle, storage, f"sphere 128x64x64 {storage} after 200 steps")
        print(f"\n{storage}: C_d every 50 steps {' '.join(f'{v:.6f}' for v in series)} (GPU == oracle)")
    finally:
        lbm.close()


# FP16S against FP32 on the 906 M-cell sphere: FP16S rounds a shifted DDF f - w of about
# w * 3 u = 1.25e-2 (u = 0.075) to half a unit of 2^-17, 3.8e-6, while the non-equilibrium
# part that carries the wall stress is only about w * 3 tau du/dx = 6e-5 at R = 96 (tau =
# 0.93, du/dx ~ u/D = 3.9e-4): the format's rounding is some 6 % of it.  The drag of the
# transient start therefore differs from FP32 by more than on the smaller spheres of the
# same scene (FP16S against FP32, largest difference over the first reports: 8.5e-4 at
# R = 16, 1.3e-3 at R = 32, 2.8e-3 at R = 64, 1.7e-2 at R = 96; FP16C, with one more mantissa
# bit and a binade-aligned exponent, 3.9e-3 at R = 96), while the kernels equal the oracle
# bit for bit (test above).  Bound: 3 %.
FP16S_DRAG_BOUND_R96 = 3e-2


def test_sphere_1536x768x768_forward(tmp_path):
    """D3Q19 SRT flow past a sphere on 1536x768x768 (906 M cells; R = 96 at (384,384,384),
    TYPE_E inflow/outflow u = 0.075, Re = 100), 1000 steps in FP16S storage (the C++ scene
    fluidx3d_amd/host/bin/sphere_scale3) and in FP32 storage (sphere_scale3_fp32): every
    velocity finite, the FP64 DDF mass within 1e-3 of its initial value (the open inflow and
    outflow exchange mass), the drag positive with lateral components below 1 % of it, and
    the FP16S drag within FP16S_DRAG_BOUND_R96 of the FP32 one."""
    results = {}
    for name in ("sphere_scale3", "sphere_scale3_fp32"):
        drags, mass, out = sphere_run(tmp_path, name, timeout=900)
        t, Fx, Fy, Fz, Cd = drags[-1]
        print(f"\n{name}: t={t} F=({Fx:.6e}, {Fy:.3e}, {Fz:.3e}) C_d={Cd:.5f}; mass {mass[0]:.9e} -> {mass[1]:.9e} "
              f"(relative change {mass[2]:.2e}); non-finite u components {mass[3]}")
        print(out.strip().splitlines()[-2])
        assert t == 1000
        assert mass[3] == 0
        assert abs(mass[2]) < 1e-3
        assert Fx > 0 and abs(Fy) < 1e-2 * Fx and abs(Fz) < 1e-2 * Fx
        results[name] = Cd
    rel = abs(results["sphere_scale3"] / results["sphere_scale3_fp32"] - 1.0)
    print(f"C_d FP16S {results['sphere_scale3']:.6f} against FP32 {results['sphere_scale3_fp32']:.6f}: {rel:.2e}")
    assert rel < FP16S_DRAG_BOUND_R96


def test_porous_channel_384x192x192_adjoint():
    """Checkpointed adjoint of the porous channel scaled to 384x192x192 (14.2 M cells; walls
    at y = 0 and 191, body force 1e-6 in x, Brinkman porosity from smoothed noise on the
    middle third of x), 5000 steps, outlet x-momentum averaged over the last 1250 steps.
    (1) The gradient is bit-identical with ceil(sqrt(T)) = 71-step and 250-step checkpoint
    intervals (recomputation replays the forward exactly).  (2) The directional derivative
    along a smooth seeded field d, <dJ/dporosity, d>, matches the fourth-order central
    difference of four GPU forward runs J(a + k h d), k = -2, -1, 1, 2, within 1e-3."""
    import fluidx3d_amd as fx
    from fluidx3d_amd.adjoint import Adjoint
    from tests.test_gpu_adjoint import ADJ_EXT
    dims = (384, 192, 192)
    steps, window, plane = 5000, 3750, dims[0] - 1
    rho, u, flags, alpha = scenes.porous_channel(*dims)

    def run(porosity, k, gradient):
        lbm = fx.LBM(*dims, 1.0 / 6.0, 1e-6, 0.0, 0.0, extensions=ADJ_EXT)
        try:
            lbm.rho.host[:] = rho; lbm.u.host[:] = u; lbm.flags.host[:] = flags; lbm.porosity.host[:] = porosity
            adj = Adjoint(lbm, steps, window, plane, k)
            try:
                J = adj.forward()
                g = adj.backward() if gradient else None
            finally:
                adj.close()
        finally:
            lbm.close()
        return J, g

    J, g = run(alpha, 0, True)
    J2, g2 = run(alpha, 250, True)
    assert J2 == J
    assert np.array_equal(g, g2), "checkpoint interval changed the gradient"
    porous = (alpha > 0) & ((flags & scenes.TYPE_S) == 0)
    d = (scenes.porosity_noise(*dims, dims[0] // 3, (2 * dims[0]) // 3, 4.0, seed=7) - 0.5) * porous
    g64 = g.astype(np.float64)
    # J is a smooth function of the porosity but not a linear one at the 10 % level, and each
    # FP32 forward carries rounding of ~1e-7 J: a fourth-order central difference with
    # h = 0.1 (porosity moved by at most 0.1 of its [0, 1] range) keeps the truncation and
    # the rounding error both below 1e-4 of the derivative.  Each J(a + k h d) is compared
    # with the derivative along the float32 field actually applied.
    h = 0.1
    J_at, diff = {}, {}
    for k in (-2, -1, 1, 2):
        a_k = (alpha + k * h * d).astype(np.float32)
        J_at[k], _ = run(a_k, 0, False)
        diff[k] = a_k.astype(np.float64) - alpha
    fd4 = (-J_at[2] + 8.0 * J_at[1] - 8.0 * J_at[-1] + J_at[-2]) / 12.0
    ad4 = float(np.dot(g64, -diff[2] + 8.0 * diff[1] - 8.0 * diff[-1] + diff[-2])) / 12.0
    rel = abs(fd4 - ad4) / abs(ad4)
    fd2 = (J_at[1] - J_at[-1]) / 2.0
    ad2 = float(np.dot(g64, diff[1] - diff[-1])) / 2.0
    print(f"\nJ = {J:.10e}; <g,d> = {np.dot(g64, d):.6e}; h = {h}: fourth-order difference {fd4:.6e} against the "
          f"adjoint's {ad4:.6e} (relative error {rel:.2e}); second-order {fd2:.6e} against {ad2:.6e} "
          f"({abs(fd2 - ad2) / abs(ad2):.2e})")
    assert rel < 1e-3
