"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on the same inputs.

Bar: FP32 arithmetic is the reference's operation order on both sides, so DDFs and
rho/u agree exactly (as float values; the sign of an exact zero is not compared, see
fluidx3d_amd/csrc/lbm_device.hpp).  FP16S/FP16C DDFs agree code-for-code up to the sign
of zero.  Each FP32 result's error against the FP64 oracle is stated in the TGV test.
"""
import ctypes
import itertools

import numpy as np
import pytest

from tests import scenes

pytestmark = pytest.mark.gpu

EXT_SETS = {
    "plain": (),
    "forces": ("VOLUME_FORCE", "FORCE_FIELD", "EQUILIBRIUM_BOUNDARIES", "UPDATE_FIELDS"),
    "porosity": ("POROSITY",),
    "precise": ("VOLUME_FORCE", "POROSITY", "PRECISE_DENSITY", "EQUILIBRIUM_BOUNDARIES"),
    # subsets with a specialised stream_collide of their own
    "equilibrium": ("EQUILIBRIUM_BOUNDARIES",),
    "volume_force": ("VOLUME_FORCE",),
    "drag": ("EQUILIBRIUM_BOUNDARIES", "FORCE_FIELD"),
    # solid cells keep their random velocities: moving walls
    "moving": ("MOVING_BOUNDARIES", "EQUILIBRIUM_BOUNDARIES", "UPDATE_FIELDS"),
    "subgrid": ("SUBGRID", "VOLUME_FORCE"),
    "moving_subgrid": ("MOVING_BOUNDARIES", "SUBGRID", "FORCE_FIELD", "VOLUME_FORCE"),
}
EXT_BITS = {"VOLUME_FORCE": 1, "FORCE_FIELD": 2, "EQUILIBRIUM_BOUNDARIES": 4, "UPDATE_FIELDS": 8, "POROSITY": 16,
            "PRECISE_DENSITY": 32, "MOVING_BOUNDARIES": 64, "SUBGRID": 128}


def ext_bits(names):
    b = 0
    for n in names:
        b |= EXT_BITS[n]
    return b


def decoded(storage, ddf, oracle):
    if storage == "FP32":
        return ddf.astype(np.float32)
    return oracle.decode(storage, ddf)


def setup_pair(oracle, Nx, Ny, Nz, nu, q, collision, storage, ext_names, rho, u, flags, F=None, alpha=None, f=(0.0, 0.0, 0.0)):
    import fluidx3d_amd as fx
    lbm = fx.LBM(Nx, Ny, Nz, nu, *f, velocity_set=q, collision=collision, storage=storage, extensions=ext_names)
    ref = oracle.OracleLBM(Nx, Ny, Nz, w=lbm.w, q=q, trt=(collision == "TRT"), storage=storage, ext=ext_bits(ext_names),
                           fx=np.float32(f[0]), fy=np.float32(f[1]), fz=np.float32(f[2]))
    for tgt in (lbm, ref):
        r = tgt.rho.host if hasattr(tgt.rho, "host") else tgt.rho
        v = tgt.u.host if hasattr(tgt.u, "host") else tgt.u
        fl = tgt.flags.host if hasattr(tgt.flags, "host") else tgt.flags
        r[:] = rho
        v[:] = u
        fl[:] = flags
    if F is not None:
        lbm.F.host[:] = F
        ref.F[:] = F
    if alpha is not None:
        lbm.porosity.host[:] = alpha
        ref.alpha[:] = alpha
    return lbm, ref


def compare_state(lbm, ref, oracle, storage, tag):
    g = decoded(storage, lbm.read_ddfs(), oracle)
    o = decoded(storage, ref.fi, oracle)
    bad = np.flatnonzero(g != o)
    assert bad.size == 0, f"{tag}: {bad.size} DDF values differ, first slot/cell {divmod(int(bad[0]), lbm.N)}: gpu {g[bad[0]]!r} oracle {o[bad[0]]!r}"


# 20x12x10, 150x6x5, 520x3x3: rows off the 128-byte lines (cell-linear grid for 16-bit, and
# for FP32 where Nx*4 misses them); 64x5x4: rows on lines, one block; 512x4x3: one block per
# row, and two cells per lane for the D3Q19 16-bit kernels that take it; 1024x3x2: rows of
# several blocks on the XCD-mapped grid, with both kernel forms
GRIDS = {"20x12x10": (20, 12, 10), "150x6x5": (150, 6, 5), "520x3x3": (520, 3, 3), "64x5x4": (64, 5, 4), "512x4x3": (512, 4, 3),
         "1024x3x2": (1024, 3, 2)}


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("q,collision,storage,ext", list(itertools.product((19, 27), ("SRT", "TRT"), ("FP32", "FP16S", "FP16C"), EXT_SETS)))
def test_stream_collide_matches_oracle(oracle, q, collision, storage, ext, grid):
    """Random fields with scattered solids (and TYPE_E planes, forces, porosity): 7 steps,
    both parities, ragged x-rows: Nx=20 in one 64-lane block; Nx=150, which with two cells
    per lane leaves a block with a full wave, a partly live wave and dead second cells;
    Nx=520, rows of several blocks on the 1-D XCD-mapped grid, whose last group of 8 rows
    (9 rows here) is mostly padding.  These rows do not start on 128-byte lines, so the
    kernels run on the cell-linear grid, whose blocks cross rows and planes (Nx=20: 12.8
    rows per block); Nx=64 (one block per row) and Nx=512 (two blocks, XCD-mapped grid)
    start every row on a line and keep the row grid."""
    Nx, Ny, Nz = GRIDS[grid]
    names = EXT_SETS[ext]
    rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=q + len(names), equilibrium_faces="EQUILIBRIUM_BOUNDARIES" in names)
    N = Nx * Ny * Nz
    rng = np.random.default_rng(7)
    F = (1e-4 * rng.uniform(-1, 1, 3 * N)).astype(np.float32) if "FORCE_FIELD" in names else None
    alpha = rng.uniform(0, 1, N).astype(np.float32) if "POROSITY" in names else None
    if "PRECISE_DENSITY" in names:  # weak forcing near rest, where the DDF-sum density matters
        u *= np.float32(0.02)
    f = (2e-5, -1e-5, 3e-6) if ("VOLUME_FORCE" in names or "POROSITY" in names) else (0.0, 0.0, 0.0)
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 0.048, q, collision, storage, names, rho, u, flags, F, alpha, f)
    try:
        lbm.initialize()
        ref.initialize()
        compare_state(lbm, ref, oracle, storage, "after initialize")
        if "MOVING_BOUNDARIES" in names:  # TYPE_MS marks written by initialize
            lbm.flags.read_from_device()
            assert np.array_equal(lbm.flags.host, ref.flags) and (ref.flags & 3 == 3).any()
        for s in range(7):
            lbm.run(1)
            ref.step(1)
            compare_state(lbm, ref, oracle, storage, f"after step {s + 1}")
        lbm.u.read_from_device()
        lbm.rho.read_from_device()
        if "UPDATE_FIELDS" not in names:  # with UPDATE_FIELDS rho/u come from the last stream_collide
            ref.update_fields()           # (src/lbm.cpp:184-191, src/lbm.hpp:390-393)
        assert np.array_equal(lbm.rho.host, ref.rho), "rho differs after update_fields"
        assert np.array_equal(lbm.u.host, ref.u), "u differs after update_fields"
        assert abs(lbm.mass() - ref.mass()) <= 1e-12 * abs(ref.mass())
    finally:
        lbm.close()


PAIR_SETS = {"equilibrium": ("EQUILIBRIUM_BOUNDARIES",), "porous": ("VOLUME_FORCE", "POROSITY", "PRECISE_DENSITY")}


@pytest.mark.parametrize("collision,storage,ext", list(itertools.product(("SRT", "TRT"), ("FP16S", "FP16C"), PAIR_SETS)))
def test_pair_kernel_dead_waves_and_lanes(oracle, collision, storage, ext):
    """Two cells per lane (D3Q19 16-bit, rows of whole waves of 128 cells): rows whose first
    or last wave is all solid (the wave returns before its exchanges), rows with both cells
    of every fourth lane solid (the lane still passes codes between the x-shifted runs and
    the cells' lanes and writes its own back unchanged), scattered single solid cells and,
    with EQUILIBRIUM_BOUNDARIES, TYPE_E planes at the row ends, on 256x5x4 with rows of one
    block; 6 steps, both parities, against the oracle."""
    Nx, Ny, Nz = 256, 5, 4
    names = PAIR_SETS[ext]
    rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=23, solid_fraction=0.05, equilibrium_faces="EQUILIBRIUM_BOUNDARIES" in names)
    if "PRECISE_DENSITY" in names:
        u *= np.float32(0.02)
    x, y, z = scenes.coords(Nx, Ny, Nz)
    flags[(y == 1) & (x < 128)] = scenes.TYPE_S               # a row's first wave
    flags[(y == 2) & (z == 3) & (x >= 128)] = scenes.TYPE_S   # a row's last wave (the +x wrap's source)
    flags[(y == 3) & (x % 8 < 2)] = scenes.TYPE_S             # lanes with both cells solid
    N = Nx * Ny * Nz
    alpha = np.random.default_rng(5).uniform(0, 1, N).astype(np.float32) if "POROSITY" in names else None
    f = (2e-5, -1e-5, 3e-6) if "VOLUME_FORCE" in names else (0.0, 0.0, 0.0)
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 0.048, 19, collision, storage, names, rho, u, flags, None, alpha, f)
    try:
        lbm.initialize()
        ref.initialize()
        for s in range(6):
            lbm.run(1)
            ref.step(1)
            compare_state(lbm, ref, oracle, storage, f"after step {s + 1}")
        assert abs(lbm.mass() - ref.mass()) <= 1e-12 * abs(ref.mass())
    finally:
        lbm.close()


CLAMP_SETS = {
    "plain": (),
    "volume_force": ("VOLUME_FORCE",),
    "equilibrium": ("EQUILIBRIUM_BOUNDARIES",),  # TYPE_E presets above def_c, clamped in the emission
    "forces": ("VOLUME_FORCE", "FORCE_FIELD", "EQUILIBRIUM_BOUNDARIES", "UPDATE_FIELDS"),
}


@pytest.mark.parametrize("grid", ["64x6x5", "20x12x10"])
@pytest.mark.parametrize("q,collision,storage,ext", list(itertools.product((19, 27), ("SRT", "TRT"), ("FP32", "FP16S", "FP16C"), CLAMP_SETS)))
def test_velocity_clamp_matches_oracle(oracle, q, collision, storage, ext, grid):
    """A coherent fast flow, u ~ (0.6, -0.58, 0) +- 0.1 (TYPE_E presets too), with a strong
    body force where enabled, so that hundreds of cells per step have velocity components
    beyond def_c = 0.57735027 and take the clamp (src/kernel.cpp:1553-1560, 1853-1855):
    6 steps, every DDF compared after every step, then rho/u after update_fields."""
    from tests.test_oracle_clamp import fast_scene
    dims = {"64x6x5": (64, 6, 5), "20x12x10": (20, 12, 10)}[grid]
    Nx, Ny, Nz = dims
    N = Nx * Ny * Nz
    names = CLAMP_SETS[ext]
    rho, u, flags = fast_scene(Nx, Ny, Nz, seed=q + len(names), equilibrium_faces="EQUILIBRIUM_BOUNDARIES" in names)
    rng = np.random.default_rng(3)
    F = (1e-3 * rng.uniform(-1, 1, 3 * N)).astype(np.float32) if "FORCE_FIELD" in names else None
    f = (0.01, -0.01, 0.005) if "VOLUME_FORCE" in names else (0.0, 0.0, 0.0)
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 0.1, q, collision, storage, names, rho, u, flags, F, None, f)
    try:
        lbm.initialize()
        ref.initialize()
        compare_state(lbm, ref, oracle, storage, "after initialize")
        clamped = 0
        for s in range(6):
            lbm.run(1)
            ref.step(1)
            compare_state(lbm, ref, oracle, storage, f"after step {s + 1}")
            if "UPDATE_FIELDS" not in names:
                ref.update_fields()
            clamped += int((np.abs(ref.u) == np.float32(0.57735027)).sum())
        assert clamped > 100, clamped
        lbm.u.read_from_device()
        lbm.rho.read_from_device()
        assert np.array_equal(lbm.rho.host, ref.rho), "rho differs after update_fields"
        assert np.array_equal(lbm.u.host, ref.u), "u differs after update_fields"
    finally:
        lbm.close()


def wide_codes(storage, rng, n):
    """Random finite 16-bit codes, half of them below 2^-6 in magnitude, half of any exponent."""
    sign = rng.integers(0, 2, n).astype(np.uint16) << 15
    if storage == "FP16S":  # IEEE binary16 of x*2^15: exponent fields 0..30 (31 is inf/NaN)
        e = np.where(rng.uniform(0, 1, n) < 0.5, rng.integers(0, 9, n), rng.integers(0, 31, n)).astype(np.uint16)
        return sign | (e << 10) | rng.integers(0, 1 << 10, n).astype(np.uint16)
    e = np.where(rng.uniform(0, 1, n) < 0.5, rng.integers(0, 9, n), rng.integers(0, 16, n)).astype(np.uint16)
    return sign | (e << 11) | rng.integers(0, 1 << 11, n).astype(np.uint16)


WIDE_SETS = {"plain": (), "volume_force": ("VOLUME_FORCE",), "drag": ("EQUILIBRIUM_BOUNDARIES", "FORCE_FIELD")}


@pytest.mark.parametrize("grid", ["512x4x3", "20x12x10"])
@pytest.mark.parametrize("q,collision,storage,ext", list(itertools.product((19, 27), ("SRT", "TRT"), ("FP16S", "FP16C"), WIDE_SETS)))
def test_codes_of_every_magnitude(oracle, q, collision, storage, ext, grid):
    """One step from DDFs that are random 16-bit codes of every exponent: densities
    rho = 1 + sum f from below -1 to above 2, velocities j/rho of any size (clamped), outputs
    beyond the formats' range.  Planted cells load populations that give rho = 0 with j != 0
    (u = +-inf before the clamp), rho = the smallest code (2^-39 FP16S, 2^-25 FP16C),
    rho = -2^-8 and rho = 2 - 2^-23 (an all-ones mantissa, the divisor for which a
    reciprocal one ulp off would misround the quotient); TYPE_E cells carry preset
    densities 1e-30, 1e30 and 3e-20 (the 0.5/rho of the force block far outside the usual
    range).  Zero, non-finite and extreme densities take the division's general path where
    the kernels otherwise share one reciprocal of rho between u = j/rho and 0.5/rho.
    Every code is compared with the oracle's."""
    Nx, Ny, Nz = GRIDS[grid]
    N = Nx * Ny * Nz
    names = WIDE_SETS[ext]
    rho, u, flags = scenes.random_scene(Nx, Ny, Nz, seed=11 + q, equilibrium_faces="EQUILIBRIUM_BOUNDARIES" in names)
    e_cells = np.flatnonzero(flags == 2)
    if e_cells.size:
        rho[e_cells[:3]] = np.float32([1e-30, 1e30, 3e-20])
    rng = np.random.default_rng(q + 5 * len(names))
    F = (1e-3 * rng.uniform(-1, 1, 3 * N)).astype(np.float32) if "FORCE_FIELD" in names else None
    f = (1e-3, 2e-3, 5e-4) if "VOLUME_FORCE" in names else (0.0, 0.0, 0.0)
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 0.1, q, collision, storage, names, rho, u, flags, F, None, f)
    try:
        lbm.initialize()
        ref.initialize()
        codes = wide_codes(storage, rng, q * N)
        # where each cell loads population i at the next step: an FP32 lattice of indices
        probe = oracle.OracleLBM(Nx, Ny, Nz, w=lbm.w, q=q, storage="FP32")
        probe.fi[:] = np.arange(q * N, dtype=np.float32)
        probe.t = ref.t
        at = probe.populations().astype(np.int64)  # (q, N)
        minus_one = np.uint16(0xF800)  # -1 in both formats (FP16S: -2^15 in binary16; FP16C: e=15)
        quarter = np.uint16(0x7000 if storage == "FP16S" else 0x6800)
        smallest = np.uint16(1)
        fluid = np.flatnonzero((flags & 3) != 1)
        plant = {"zero": fluid[::97], "tiny": fluid[31::97], "negative": fluid[61::97], "ones": fluid[79::97]}
        for kind, cells in plant.items():
            for n in cells:
                codes[at[:, n]] = 0
                codes[at[0, n]] = minus_one
                if kind == "zero":  # +-x, +-y, +-z pairs cancel in rho; j = (1/2, 1/2, 1/2), u = +inf
                    for i in range(1, 7):
                        codes[at[i, n]] = quarter if i % 2 else quarter | np.uint16(0x8000)
                    if F is not None:  # forces that keep u = +inf + f*0.5/rho at +inf
                        F[[n, N + n, 2 * N + n]] = np.abs(F[[n, N + n, 2 * N + n]])
                elif kind == "tiny":
                    codes[at[1, n]] = smallest
                elif kind == "negative":  # rho = -2^-8
                    codes[at[3, n]] = np.uint16(0xD800 if storage == "FP16S" else 0xB800)
                else:  # rho = 2 - 2^-23 (all-ones mantissa): 1 - 2^-11, 2^-11 - 2^-22, 2^-23
                    codes[at[0, n]] = 0
                    for i, c in zip((1, 3, 5), (0x77FF, 0x4BFF, 0x1C00) if storage == "FP16S" else (0x77FE, 0x1FFE, 0x0004)):
                        codes[at[i, n]] = np.uint16(c)
        lbm.write_ddfs(codes)
        ref.fi[:] = codes
        if F is not None:
            lbm.F.host[:] = F
            lbm.F.write_to_device()
            ref.F[:] = F
        pops = oracle.decode(storage, codes)[at.ravel()].reshape(q, N).astype(np.float64)
        dens = 1.0 + pops.sum(axis=0)
        assert (dens[plant["zero"]] == 0).all() and (dens[plant["tiny"]] > 0).all() and (dens[plant["tiny"]] < 1e-7).all()
        assert (dens[plant["negative"]] == -2.0 ** -8).all() and (dens[plant["ones"]] == 2.0 - 2.0 ** -23).all()
        assert dens[fluid].min() < -1 and dens[fluid].max() > 2
        lbm.run(1)
        ref.step(1)
        compare_state(lbm, ref, oracle, storage, "after one step")
    finally:
        lbm.close()


@pytest.mark.parametrize("storage", ["FP32", "FP16S", "FP16C"])
def test_tgv_64_1000_steps(oracle, storage):
    """D3Q19 SRT 64^3 Taylor-Green, nu=1/6, u0=0.05, 1000 steps.
    GPU == FP32 oracle (same storage) exactly; the error against the FP64 oracle is stated
    and bounded relative to u0; the kinetic energy decays at the viscous rate 2*nu*|k|^2."""
    Nx = Ny = Nz = 64
    N = Nx * Ny * Nz
    u0 = 0.05
    rho, u, flags = scenes.taylor_green(Nx, Ny, Nz, u0=u0)
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 1.0 / 6.0, 19, "SRT", storage, (), rho, u, flags)
    ref64 = oracle.OracleLBM(Nx, Ny, Nz, w=lbm.w, storage="FP64")
    ref64.rho[:] = rho
    ref64.u[:] = u
    try:
        lbm.initialize(); ref.initialize(); ref64.initialize()
        lbm.run(200)
        lbm.u.read_from_device(); lbm.rho.read_from_device()
        ke200 = scenes.kinetic_energy(lbm.rho.host, lbm.u.host, N)
        lbm.run(800); ref.step(1000); ref64.step(1000)
        compare_state(lbm, ref, oracle, storage, "TGV 1000 steps")
        lbm.u.read_from_device(); lbm.rho.read_from_device()
        ref.update_fields(); ref64.update_fields()
        assert np.array_equal(lbm.u.host, ref.u) and np.array_equal(lbm.rho.host, ref.rho)
        err_u = np.abs(lbm.u.host.astype(np.float64) - ref64.u).max()
        err_rho = np.abs(lbm.rho.host.astype(np.float64) - ref64.rho).max()
        print(f"\n{storage}: max|u-u64| = {err_u:.3e} ({err_u / u0:.2e} u0), max|rho-rho64| = {err_rho:.3e}")
        bound = {"FP32": 1e-5, "FP16S": 5e-5, "FP16C": 5e-5}[storage]
        assert err_u < bound * u0 and err_rho < 1e-5
        ke1000 = scenes.kinetic_energy(lbm.rho.host, lbm.u.host, N)
        rate = np.log(ke200 / ke1000) / 800.0
        expect = 2.0 * (1.0 / 6.0) * 3.0 * (2.0 * np.pi / 64.0) ** 2
        assert abs(rate / expect - 1.0) < 0.01, (rate, expect)
    finally:
        lbm.close()


def test_full_size_256_two_steps(oracle):
    """256^3 D3Q19 SRT FP32 TGV, 2 steps, every cell compared with the oracle."""
    Nx = Ny = Nz = 256
    rho, u, flags = scenes.taylor_green(Nx, Ny, Nz)
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 1.0 / 6.0, 19, "SRT", "FP32", (), rho, u, flags)
    try:
        lbm.initialize(); ref.initialize()
        lbm.run(2); ref.step(2)
        compare_state(lbm, ref, oracle, "FP32", "256^3 2 steps")
    finally:
        lbm.close()


@pytest.mark.parametrize("storage,reps,q,collision,ext,n0", [("FP32", (4, 8, 8), 19, "SRT", (), 64), ("FP16S", (8, 8, 8), 19, "SRT", (), 64),
                                                         ("FP16C", (6, 8, 8), 27, "TRT", ("EQUILIBRIUM_BOUNDARIES", "FORCE_FIELD"), 64),
                                                         ("FP16C", (7, 8, 8), 19, "TRT", ("EQUILIBRIUM_BOUNDARIES", "FORCE_FIELD"), 64),
                                                         ("FP16S", (23, 25, 25), 19, "SRT", ("VOLUME_FORCE",), 20)])
def test_lattice_beyond_4_gib_matches_its_tile(storage, reps, q, collision, ext, n0):
    """64-bit addressing at full scale: a 512x512x256 FP32 / 512^3 FP16S D3Q19 SRT lattice
    (5.1 GB of DDFs, slot offsets up to 4.8 GB; FP32 above 4 GiB on the cell-linear grid,
    FP16S on the XCD-mapped row grid, the 64^3 lattices on the row grid) filled with a
    Taylor-Green vortex of period 64 is the 64^3 vortex repeated, so after 3 steps every
    64^3 tile of slots 0, 9, 17 and 18 (and 26 for D3Q27) equals the 64^3 lattice's DDFs bit
    for bit (the 64^3 lattice is oracle-checked above).  The D3Q27 TRT FP16C kernel with
    EQUILIBRIUM_BOUNDARIES + FORCE_FIELD (no TYPE_E cells, zero force field) takes buffer
    instructions at 64^3 and, with 5.4 GB of DDFs at 512x512x384, 32-bit slot offsets; the
    D3Q19 TRT FP16C two-cells-per-lane kernel takes buffer instructions at 64^3 and global
    addressing with 4.5 GB at 512x512x448; the 20^3 vortex (cell-linear grid, rows off the
    lines) tiled to 500x500x460 (4.4 GB of FP16S DDFs) leaves the cell-linear grid's buffer
    instructions for its 32-bit slot offsets.  Every large lattice here also takes its
    blocks in reverse order on odd steps."""
    import fluidx3d_amd as fx
    rho0, u0, _ = scenes.taylor_green(n0, n0, n0)
    opts = dict(storage=storage, velocity_set=q, collision=collision, extensions=ext)
    small = fx.LBM(n0, n0, n0, 1.0 / 6.0, **opts)
    big = fx.LBM(n0 * reps[2], n0 * reps[1], n0 * reps[0], 1.0 / 6.0, **opts)
    try:
        small.rho.host[:] = rho0; small.u.host[:] = u0
        small.run(3)
        ref = small.read_ddfs().reshape(q, n0, n0, n0)
        big.rho.host[:] = np.tile(rho0.reshape(n0, n0, n0), reps).ravel()
        for d in range(3):
            big.u.host[d * big.N:(d + 1) * big.N] = np.tile(u0[d * n0 ** 3:(d + 1) * n0 ** 3].reshape(n0, n0, n0), reps).ravel()
        big.run(3)
        esize = 4 if storage == "FP32" else 2
        assert q * big.N * esize > 4 << 30
        for i in (0, 9, 17, 18) + ((26,) if q == 27 else ()):
            slot = np.empty(big.N, np.float32 if storage == "FP32" else np.uint16)
            big.ctx.read(big.fi, slot, offset=i * big.N * esize)
            tiles = slot.reshape(reps[0], n0, reps[1], n0, reps[2], n0)
            assert np.array_equal(tiles, np.broadcast_to(ref[i][None, :, None, :, None, :], tiles.shape)), f"slot {i}"
    finally:
        small.close()
        big.close()


@pytest.mark.parametrize("storage", ["FP16S", "FP16C"])
def test_codec_matches_oracle(oracle, storage):
    """Device encode/decode against the oracle: every 16-bit code that encodes a finite
    value round-trips, and 2M random floats across all magnitudes encode identically
    (FP16C: also every float of the four binades where the encoding rule changes form).
    The device encodes odd elements with the second-cell form of the two-cells-per-lane
    kernel (encode_high), even ones with the plain encoder."""
    import fluidx3d_amd as fx
    from fluidx3d_amd import _abi
    cfg = fx.make_config(4, 4, 4, 1.0)
    ctx = _abi.Context(cfg)
    try:
        codes = np.arange(65536, dtype=np.uint16)
        vals = oracle.decode(storage, codes)
        finite = np.isfinite(vals)
        rng = np.random.default_rng(3)
        rnd = rng.integers(0, 2 ** 32, 2 ** 21, dtype=np.uint64).astype(np.uint32).view(np.float32)
        rnd = rnd[np.isfinite(rnd) & (np.abs(rnd) < (2.0 if storage == "FP16C" else 1.99))]
        parts = [vals[finite], rnd, (rng.uniform(-1, 1, 2 ** 20) * 10.0 ** rng.uniform(-12, 0, 2 ** 20)).astype(np.float32)]
        if storage == "FP16C":
            # every float of the binades where the reference's rule changes form: zero below
            # exponent 101, the denormal codes up to 112, the normal codes from 113 on
            m = np.arange(2 ** 23, dtype=np.uint32)
            for e in (100, 101, 112, 113):
                parts.append(((np.uint32(e) << np.uint32(23)) | m).view(np.float32))
                parts.append(((np.uint32(256 + e) << np.uint32(23)) | m).view(np.float32))
        x = np.concatenate(parts)
        n = x.size
        dx = ctx.alloc(4 * n); dc = ctx.alloc(2 * n); dd = ctx.alloc(4 * n)
        ctx.write(dx, x)
        _abi.check(ctx.lib.fx_codec_roundtrip(ctx.h, _abi.STORAGE[storage], dx, dc, dd, n))
        gc = np.empty(n, dtype=np.uint16); gd = np.empty(n, dtype=np.float32)
        ctx.read(dc, gc); ctx.read(dd, gd)
        oc = oracle.encode(storage, x)
        od = oracle.decode(storage, oc)
        mism = np.flatnonzero(gc != oc)
        # only signed zeros may differ
        assert np.all(od[mism] == 0) and np.all(gd[mism] == 0), f"{mism.size} codes differ, e.g. x={x[mism[:3]]}"
        assert np.array_equal(gd, od)
        m = finite.sum()
        assert np.array_equal(gd[:m], vals[finite]), "decode(encode(decode(code))) != decode(code)"
    finally:
        ctx.close()


def test_object_force_and_mass(oracle):
    """FORCE_FIELD: momentum-exchange force field equal to the oracle's, FP64 object sums
    equal to the oracle's FP64 sums to 1e-12."""
    Nx, Ny, Nz = 48, 32, 32
    flags = scenes.sphere_flags(Nx, Ny, Nz, (16, 16, 16), 6, marker=scenes.TYPE_S | scenes.TYPE_X)
    x, _, _ = scenes.coords(Nx, Ny, Nz)
    flags[(x == 0) | (x == Nx - 1)] = scenes.TYPE_E
    N = Nx * Ny * Nz
    rho = np.ones(N, dtype=np.float32)
    u = np.zeros(3 * N, dtype=np.float32)
    u[:N][(flags & scenes.TYPE_S) == 0] = 0.05
    names = ("FORCE_FIELD", "EQUILIBRIUM_BOUNDARIES", "VOLUME_FORCE")
    lbm, ref = setup_pair(oracle, Nx, Ny, Nz, 0.05, 27, "TRT", "FP32", names, rho, u, flags, F=np.zeros(3 * N, np.float32))
    try:
        lbm.initialize(); ref.initialize()
        lbm.run(50); ref.step(50)
        marker = scenes.TYPE_S | scenes.TYPE_X
        g = lbm.object_force(marker)
        o = ref.object_force(marker)
        lbm.F.read_from_device()
        assert np.array_equal(lbm.F.host, ref.F)
        assert np.allclose(g, o, rtol=1e-12, atol=1e-15), (g, o)
        assert g[0] > 0  # drag along the flow
        c = (0.5 * Nx - 0.5, 0.5 * Ny - 0.5, 0.5 * Nz - 0.5)
        assert np.allclose(lbm.object_torque(c, marker), ref.object_torque(c, marker), rtol=1e-10, atol=1e-12)
        gc = lbm.object_center_of_mass(marker)
        oc, cells = ref.object_center_of_mass(marker)
        assert np.allclose(gc, oc, rtol=0, atol=1e-9)
        assert abs(lbm.mass() - ref.mass()) <= 1e-12 * ref.mass()
    finally:
        lbm.close()
