"""The C++ host layer (fluidx3d_amd/host): scenes written against the reference's LBM API,
built with g++ against libfx_lbm.so."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "fluidx3d_amd", "host", "bin")


def binary(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (make -C fluidx3d_amd/host)")
    return path


def read_vtk(path):
    """Binary legacy VTK STRUCTURED_POINTS written by Memory_Container::write_vtk."""
    raw = open(path, "rb").read()
    head_end = raw.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    header = raw[:head_end].decode()
    dims = [int(v) for v in re.search(r"DIMENSIONS (\d+) (\d+) (\d+)", header).groups()]
    comps = int(re.search(r"SCALARS data float (\d+)", header).group(1))
    data = np.frombuffer(raw[head_end:], dtype=">f4").astype(np.float32)
    return dims, data.reshape(-1, comps)


def test_scene_without_gpu_fails_like_reference(tmp_path):
    """No device: the scene prints an error and exits with status 1 (print_error, src/utilities.hpp:4074-4086)."""
    import fluidx3d_amd as fx
    try:
        if fx.device_count() > 0:
            pytest.skip("a GPU is present")
    except Exception:
        pass
    r = subprocess.run([binary("tgv64")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "Error" in r.stderr


@pytest.mark.gpu
def test_cpp_taylor_green_matches_python_path(tmp_path):
    """64^3 TGV, 1000 steps through the C++ LBM class == the same run through the Python
    mirror (both call libfx_lbm.so).  Both evaluate the initial fields in double and round
    to float; libm and numpy sin/cos may differ in the last double bit, which can flip a
    float rounding, so agreement is required to 1e-7 u0 rather than bitwise."""
    import fluidx3d_amd as fx
    from tests import scenes
    r = subprocess.run([binary("tgv64")], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    m = re.search(r"mass drift=([-0-9.e+]+)", r.stdout)
    assert m and abs(float(m.group(1))) < 1e-6
    dims, u_cpp = read_vtk(os.path.join(tmp_path, "export", "u-000001000.vtk"))
    assert dims == [64, 64, 64]
    # LBM::write_status (src/lbm.cpp:1047-1072): the reference's lines and number text
    status = open(os.path.join(tmp_path, "export", "status-000001000.txt")).read().splitlines()
    assert status[0] == "Grid Resolution = 64 x 64 x 64 = 262144"
    assert status[1] == "Grid Domains = 1 x 1 x 1 = 1"
    assert status[2] == "LBM Type = D3Q19 SRT"
    assert status[4] == "Maximum Allocation Size = 19 MB"
    assert status[5] == "Time Steps = 1000 / 1000"
    # to_string(1.0f/6.0f) of src/utilities.hpp:2745-2754: the float is scaled by 10 in float
    # arithmetic before its digits are taken, which rounds 1.66666671... up to 1.66666675...
    # (the reference's own helper prints the same, tests/test_host_text.py)
    assert status[8] == "Kinematic Viscosity = 1.66666676E-1"
    assert status[9] == "Relaxation Time = 1.00000000"
    assert re.fullmatch(r"Average MLUPs/s = \d+", status[7]) and int(status[7].split()[-1]) > 1000
    N = 64 ** 3
    lbm = fx.LBM(64, 64, 64, 1.0 / 6.0)
    rho, u, flags = scenes.taylor_green(64, 64, 64)
    lbm.rho.host[:] = rho
    lbm.u.host[:] = u
    lbm.run(1000)
    lbm.u.read_from_device()
    u_py = lbm.u.host.reshape(3, N).T
    lbm.close()
    assert np.max(np.abs(u_cpp - u_py)) < 5e-9
    exact = np.mean(u_cpp == u_py)
    print(f"\nC++ vs Python u: {100 * exact:.4f}% bit-identical, max |diff| {np.max(np.abs(u_cpp - u_py)):.2e}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["benchmark_fp32", "benchmark_fp16s", "benchmark_fp16c"])
def test_cpp_benchmark_scene(tmp_path, name):
    """The reference's benchmark scene (256^3, nu=1, 1000 x run(10), peak smoothed MLUPs/s)."""
    r = subprocess.run([binary(name)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    mlups = int(re.search(r"Peak MLUPs/s = (\d+)", r.stdout).group(1))
    print(f"\n{name}: peak {mlups} MLUPs/s")
    floor = {"benchmark_fp32": 30380, "benchmark_fp16s": 54494, "benchmark_fp16c": 40157}[name]  # reference, README.md:1137
    assert mlups > floor


DRAG = re.compile(r"t=(\d+) drag F=\(([-0-9.e+naif]+), ([-0-9.e+naif]+), ([-0-9.e+naif]+)\) Cd=([-0-9.e+naif]+)")
MASS = re.compile(r"mass ([-0-9.e+naif]+) -> ([-0-9.e+naif]+) \(relative change ([-0-9.e+naif]+)\), non-finite u components (\d+)")


def sphere_run(tmp_path, name, timeout=600):
    """Runs a sphere scene binary; returns its drag reports [(t, Fx, Fy, Fz, Cd)], the mass
    line (m0, m1, relative change, non-finite u components) and stdout."""
    r = subprocess.run([binary(name)], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    drags = [(int(m.group(1)),) + tuple(float(v) for v in m.groups()[1:]) for m in DRAG.finditer(r.stdout)]
    m = MASS.search(r.stdout)
    assert m, r.stdout[-2000:]
    mass = tuple(float(v) for v in m.groups()[:3]) + (int(m.group(4)),)
    return drags, mass, r.stdout


@pytest.mark.gpu
def test_cpp_sphere_scene(tmp_path):
    """D3Q27 TRT FP16C sphere with TYPE_E inflow/outflow: 1000 steps, finite positive drag,
    finite velocities, FP64 mass within 1e-3 of its initial value."""
    drags, mass, out = sphere_run(tmp_path, "sphere")
    t, Fx, Fy, Fz, Cd = drags[-1]
    print("\n" + "\n".join(out.strip().splitlines()[-2:]))
    assert t == 1000
    assert np.isfinite([Fx, Fy, Fz, Cd]).all() and Fx > 0 and abs(Fy) < 0.1 * Fx and abs(Fz) < 0.1 * Fx
    assert mass[3] == 0 and abs(mass[2]) < 1e-3


def lcg_channel_porosity(Nx, Ny, Nz, seed=42):
    """The porous channel scene's porosity: the C99 LCG of utilities.hpp (the reference's
    random(), src/utilities.hpp:244-250) drawn in cell order over the fluid cells of the
    middle third of x; walls at y=0 and y=Ny-1."""
    N = Nx * Ny * Nz
    n = np.arange(N)
    x, y = n % Nx, (n // Nx) % Ny
    wall = (y == 0) | (y == Ny - 1)
    porous = ~wall & (x >= Nx // 3) & (x < 2 * Nx // 3)
    draws = np.empty(int(porous.sum()), np.float32)
    s = seed
    for i in range(draws.size):
        s = (1103515245 * s + 12345) % 2147483648
        draws[i] = np.float32(s)
    alpha = np.zeros(N, np.float32)
    alpha[porous] = np.float32(1.0) * draws * np.float32(4.65661287e-10)
    flags = np.where(wall, 1, 0).astype(np.uint8)
    return flags, alpha


@pytest.mark.gpu
def test_cpp_porous_channel_backward_matches_python_path(tmp_path):
    """LBM::forward/LBM::backward from C++ (128x64x64 porous channel, 2000 steps, objective
    over the last 500, sqrt(T) checkpoints) == fluidx3d_amd.adjoint.Adjoint on the same
    inputs: same objective to the printed digits and a bit-identical gradient field."""
    import fluidx3d_amd as fx
    from fluidx3d_amd.adjoint import Adjoint
    r = subprocess.run([binary("porous_channel")], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    m = re.search(r"J=([-0-9.e+]+) \|dJ/dporosity\|=([-0-9.e+]+) forward ([0-9.]+) ms backward ([0-9.]+) ms", r.stdout)
    assert m, r.stdout
    J_cpp = float(m.group(1))
    print("\n" + m.group(0))
    dims, g_cpp = read_vtk(os.path.join(tmp_path, "export", "porosity_gradient-000002000.vtk"))
    Nx, Ny, Nz = 128, 64, 64
    assert dims == [Nx, Ny, Nz]
    flags, alpha = lcg_channel_porosity(Nx, Ny, Nz)
    lbm = fx.LBM(Nx, Ny, Nz, 1.0 / 6.0, 1e-6, 0.0, 0.0, extensions=("VOLUME_FORCE", "POROSITY", "PRECISE_DENSITY"))
    lbm.flags.host[:] = flags
    lbm.porosity.host[:] = alpha
    adj = Adjoint(lbm, 2000, 1500, Nx - 1, 0)
    J = adj.forward()
    g = adj.backward()
    adj.close()
    lbm.close()
    assert abs(J - J_cpp) <= 1e-11 * abs(J)
    assert np.array_equal(g_cpp[:, 0], g)
    assert np.linalg.norm(g) > 0


@pytest.mark.gpu
def test_cpp_lid_driven_cavity_matches_python_path(tmp_path):
    """The reference's lid-driven cavity scene (src/setup.cpp:236-251, graphics lines left
    out) compiled unchanged with MOVING_BOUNDARIES: 48^3, Re=1000, 1000 steps.  The C++ run
    equals the Python mirror on the same inputs bit for bit, and the lid drags the fluid
    below it along +y."""
    import fluidx3d_amd as fx
    r = subprocess.run([binary("cavity48")], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    dims, u_cpp = read_vtk(os.path.join(tmp_path, "export", "u-000001000.vtk"))
    L = 48
    assert dims == [L, L, L]
    nu = float(np.float32(np.float32(L - 2) * np.float32(0.1)) / np.float32(1000.0))  # units.nu_from_Re in float
    N = L ** 3
    n = np.arange(N)
    x, y, z = n % L, (n // L) % L, n // (L * L)
    lbm = fx.LBM(L, L, L, nu, extensions=("MOVING_BOUNDARIES",))
    try:
        lbm.u.host[N:2 * N][z == L - 1] = 0.1
        wall = (x == 0) | (x == L - 1) | (y == 0) | (y == L - 1) | (z == 0) | (z == L - 1)
        lbm.flags.host[wall] = 1
        lbm.run(1000)
        lbm.u.read_from_device()
        u_py = lbm.u.host.reshape(3, N).T
    finally:
        lbm.close()
    assert np.array_equal(u_cpp, u_py)
    below_lid = (z == L - 2) & (x > 0) & (x < L - 1) & (y > 2) & (y < L - 3)
    assert u_py[below_lid, 1].mean() > 0.02


@pytest.mark.gpu
def test_cpp_box_transfers_move_the_reference_elements(tmp_path):
    """Memory_Container::read_from_device_1d/2d/3d and write_to_device_1d/2d/3d (the
    reference's Memory<T> box transfers, src/opencl.hpp:532-607) on u (3 components, each
    and all) and flags: every element that the reference's row loops move, and no other,
    changes (the scene checks each element itself)."""
    r = subprocess.run([binary("box_io")], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "box transfers: 30 cases, 0 wrong elements" in r.stdout, r.stdout
    # LBM::index(float3) / coordinates(float3) / mirror_position (src/lbm.hpp:484-515)
    assert "positions: 0 wrong cells" in r.stdout, r.stdout


@pytest.mark.gpu
def test_cpp_stokes_drag_scene(tmp_path):
    """The reference's Stokes drag validation scene (src/setup.cpp:148-187), compiled against
    this build's LBM API with FORCE_FIELD and EQUILIBRIUM_BOUNDARIES: sphere R = 32 in a
    256^3 box at Re = 0.01, all faces TYPE_E preset to the Stokes solution; it runs until the
    drag error has converged (slope and curvature below 1e-4 of the error per 100 steps).
    The converged drag of the FP64 object sum must match F_Stokes = 6 pi rho nu R u within
    the bound below, point against the far-field velocity (-u, 0, 0) and have no lateral
    component (the scene is mirror symmetric in y and z)."""
    r = subprocess.run([binary("stokes")], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    m = re.search(r"t=(\d+) F=\(([-0-9.e+]+), ([-0-9.e+]+), ([-0-9.e+]+)\) F_Stokes=([-0-9.e+]+)", r.stdout)
    assert m, r.stdout[-2000:]
    t = int(m.group(1))
    Fx, Fy, Fz, F_theo = (float(v) for v in m.groups()[1:])
    err = abs(np.sqrt(Fx * Fx + Fy * Fy + Fz * Fz) - F_theo) / F_theo
    print(f"\nStokes drag: t={t}, F=({Fx:.6e}, {Fy:.3e}, {Fz:.3e}), F_Stokes={F_theo:.6e}, error {100 * err:.2f}%")
    assert abs(F_theo - 6 * np.pi * 1.0 * 1.0 * 32.0 * (0.01 * 1.0 / 64.0)) < 1e-6 * F_theo
    assert Fx < 0 and abs(Fy) < 1e-3 * abs(Fx) and abs(Fz) < 1e-3 * abs(Fx)
    assert err < 0.1
