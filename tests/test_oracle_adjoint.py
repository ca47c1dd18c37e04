"""The FP64 oracle adjoint (fxo_adjoint_step, forward-mode dual numbers through the
oracle's own collision) against central finite differences of the FP64 oracle forward:
pins the adjoint that the GPU gradient is tested against, on the CPU, for every
collision/velocity-set variant and the extensions the adjoint supports."""
import numpy as np
import pytest

from tests import scenes


def channel(oracle, q, trt, ext, alpha_override=None):
    Nx, Ny, Nz = 14, 8, 6
    rho, u, flags, alpha = scenes.porous_channel(Nx, Ny, Nz)
    rng = np.random.default_rng(3)
    alpha = np.where(alpha > 0, alpha, 0.0) + 0.2 * rng.uniform(0, 1, alpha.size) * ((flags & 1) == 0)
    if ext & oracle.EXT_EQUILIBRIUM_BOUNDARIES:  # inflow plane with a preset velocity
        x, _, _ = scenes.coords(Nx, Ny, Nz)
        flags[(x == 0) & ((flags & 1) == 0)] = 2
        u[:Nx * Ny * Nz][(x == 0) & (flags == 2)] = 0.02
    L = oracle.OracleLBM(Nx, Ny, Nz, nu=0.1, q=q, trt=trt, storage="FP64", ext=ext, fx=2e-5, fy=-3e-6, threads=4)
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.alpha[:] = alpha if alpha_override is None else alpha_override
    if L.F is not None:
        L.F[:] = 1e-5 * rng.uniform(-1, 1, 3 * L.N)
    return L, (rho, u, flags, alpha)


STEPS, WINDOW = 90, 60


def objective(oracle, q, trt, ext, alpha):
    L, _ = channel(oracle, q, trt, ext, alpha)
    L.initialize()
    J, _ = L.run_objective(STEPS, WINDOW, L.Nx - 1)
    return J


@pytest.mark.parametrize("q,trt,variant", [(19, False, "porosity"), (19, True, "porosity"), (27, False, "porosity"),
                                           (27, True, "porosity"), (19, False, "all"), (27, True, "all")])
def test_oracle_adjoint_matches_central_differences(oracle, q, trt, variant):
    ext = oracle.EXT_VOLUME_FORCE | oracle.EXT_POROSITY
    if variant == "all":
        ext |= oracle.EXT_FORCE_FIELD | oracle.EXT_EQUILIBRIUM_BOUNDARIES | oracle.EXT_PRECISE_DENSITY
    L, (_, _, flags, alpha) = channel(oracle, q, trt, ext)
    L.initialize()
    J, grad = L.adjoint_gradient(STEPS, WINDOW, L.Nx - 1)
    fluid = np.flatnonzero((flags & 3) == 0)
    cells = np.random.default_rng(q + trt).choice(fluid, 6, replace=False)
    eps = 1e-4
    for n in cells:
        a_p = L.alpha.copy(); a_p[n] += eps
        a_m = L.alpha.copy(); a_m[n] -= eps
        fd = (objective(oracle, q, trt, ext, a_p) - objective(oracle, q, trt, ext, a_m)) / (2 * eps)
        scale = max(abs(fd), 1e-3 * np.abs(grad).max())
        assert abs(grad[n] - fd) <= 1e-6 * scale, f"cell {n}: adjoint {grad[n]!r} central difference {fd!r}"
    # a plausible slip (a dropped forcing term, a sign) moves the gradient far beyond this
    assert np.abs(grad).max() > 0


def test_oracle_adjoint_gradient_of_a_solid_cell_is_zero(oracle):
    """Porosity acts only where the force block runs: solid cells get no gradient."""
    L, (_, _, flags, _) = channel(oracle, 19, False, oracle.EXT_VOLUME_FORCE | oracle.EXT_POROSITY)
    L.initialize()
    _, grad = L.adjoint_gradient(STEPS, WINDOW, L.Nx - 1)
    assert np.all(grad[(flags & 1) == 1] == 0.0)
