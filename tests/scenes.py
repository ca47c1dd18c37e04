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
