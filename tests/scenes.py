"""Scene builders shared by the tests and the benchmark: synthetic inputs in the
reference layout (rho[n], u[d*N+n], flags[n]) as float32/uint8 numpy arrays."""
import numpy as np

TYPE_S = 0x01
TYPE_E = 0x02
TYPE_X = 0x40


def coords(Nx, Ny, Nz):
    n = np.arange(Nx * Ny * Nz, dtype=np.int64)
    return n % Nx, (n // Nx) % Ny, n // (Nx * Ny)


def taylor_green(Nx, Ny, Nz, u0=0.05, periods=1):
    """3D Taylor-Green vortex on cell centres (lbm.position, src/lbm.hpp:501-503):
    u = (u0 sin kx cos ky cos kz, -u0 cos kx sin ky cos kz, 0), k = 2*pi*periods/Nx,
    rho = 1 + 3*(u0^2/16)*(cos 2kx + cos 2ky)*(cos 2kz + 2)."""
    x, y, z = coords(Nx, Ny, Nz)
    px, py, pz = x + 0.5 - 0.5 * Nx, y + 0.5 - 0.5 * Ny, z + 0.5 - 0.5 * Nz
    kx, ky, kz = 2 * np.pi * periods / Nx, 2 * np.pi * periods / Ny, 2 * np.pi * periods / Nz
    N = Nx * Ny * Nz
    u = np.zeros(3 * N, dtype=np.float64)
    u[:N] = u0 * np.sin(kx * px) * np.cos(ky * py) * np.cos(kz * pz)
    u[N:2 * N] = -u0 * np.cos(kx * px) * np.sin(ky * py) * np.cos(kz * pz)
    rho = 1.0 + 3.0 * (u0 ** 2 / 16.0) * (np.cos(2 * kx * px) + np.cos(2 * ky * py)) * (np.cos(2 * kz * pz) + 2.0)
    return rho.astype(np.float32), u.astype(np.float32), np.zeros(N, dtype=np.uint8)


def random_scene(Nx, Ny, Nz, seed=1, solid_fraction=0.1, equilibrium_faces=False, u_amp=0.05, rho_amp=0.01):
    """Random rho/u, scattered TYPE_S cells and optional TYPE_E inflow/outflow planes at
    x=0 and x=Nx-1: exercises bounce-back, equilibrium boundaries and all directions."""
    rng = np.random.default_rng(seed)
    N = Nx * Ny * Nz
    rho = (1.0 + rho_amp * rng.uniform(-1, 1, N)).astype(np.float32)
    u = (u_amp * rng.uniform(-1, 1, 3 * N)).astype(np.float32)
    flags = np.zeros(N, dtype=np.uint8)
    flags[rng.uniform(0, 1, N) < solid_fraction] = TYPE_S
    if equilibrium_faces:
        x, _, _ = coords(Nx, Ny, Nz)
        flags[(x == 0) | (x == Nx - 1)] = TYPE_E
    return rho, u, flags


def sphere_flags(Nx, Ny, Nz, center, radius, marker=TYPE_S):
    """sphere() of src/shapes.cpp:5-8: (x,y,z)-p squared distance <= r^2, float arithmetic."""
    x, y, z = coords(Nx, Ny, Nz)
    c = np.asarray(center, dtype=np.float32)
    dx = x.astype(np.float32) - c[0]
    dy = y.astype(np.float32) - c[1]
    dz = z.astype(np.float32) - c[2]
    inside = (dx * dx + dy * dy) + dz * dz <= np.float32(radius) * np.float32(radius)
    flags = np.zeros(Nx * Ny * Nz, dtype=np.uint8)
    flags[inside] = marker
    return flags


def kinetic_energy(rho, u, N):
    ux, uy, uz = u[:N].astype(np.float64), u[N:2 * N].astype(np.float64), u[2 * N:].astype(np.float64)
    return 0.5 * float(np.sum(rho.astype(np.float64) * (ux * ux + uy * uy + uz * uz)))


def porosity_noise(Nx, Ny, Nz, x0, x1, sigma=4.0, seed=42):
    """Seeded Gaussian-smoothed noise in [0,1] on the x-slab [x0, x1), 0 elsewhere: normal
    noise (numpy PCG64, seed) smoothed with a periodic Gaussian of width sigma cells (FFT,
    float64), then min-max normalised over the slab."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((Nz, Ny, Nx))
    k = [np.fft.fftfreq(m) for m in (Nz, Ny, Nx)]
    kz, ky, kx = np.meshgrid(*k, indexing="ij")
    gauss = np.exp(-2.0 * np.pi ** 2 * sigma ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    smooth = np.real(np.fft.ifftn(np.fft.fftn(noise) * gauss))
    out = np.zeros((Nz, Ny, Nx))
    slab = smooth[:, :, x0:x1]
    out[:, :, x0:x1] = (slab - slab.min()) / (slab.max() - slab.min())
    return out.ravel()


def porous_channel(Nx=128, Ny=64, Nz=64, sigma=4.0, seed=42):
    """Adjoint channel: TYPE_S walls at y=0 and y=Ny-1, periodic in x and z, fluid at rest,
    Brinkman porosity in [0,1] from porosity_noise on the middle third of x (zero on the
    walls).  Returns (rho, u, flags, porosity) as float32/uint8 arrays."""
    N = Nx * Ny * Nz
    _, y, _ = coords(Nx, Ny, Nz)
    flags = np.zeros(N, dtype=np.uint8)
    flags[(y == 0) | (y == Ny - 1)] = TYPE_S
    alpha = porosity_noise(Nx, Ny, Nz, Nx // 3, (2 * Nx) // 3, sigma, seed)
    alpha[flags == TYPE_S] = 0.0
    return (np.ones(N, dtype=np.float32), np.zeros(3 * N, dtype=np.float32), flags, alpha.astype(np.float32))


# adjoint configuration of the 128x64x64 channel: nu=1/6, body force 1e-6 in x, 2000 steps,
# objective = outlet-plane (x = Nx-1) x-momentum averaged over the last 500 steps
CHANNEL_128 = {"Nx": 128, "Ny": 64, "Nz": 64, "nu": 1.0 / 6.0, "fx": 1e-6, "steps": 2000, "window_begin": 1500,
      "plane_x": 127, "checkpoint_interval": 100, "fd_cells": 32, "fd_seed": 7, "fd_eps": 1e-3}


def fd_cells(flags, alpha, count, seed):
    """Seeded choice of porous fluid cells for the finite-difference check."""
    candidates = np.flatnonzero((alpha > 0.0) & ((flags & TYPE_S) == 0))
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(candidates, count, replace=False))
