"""Host-side launch logic of the HIP library (fluidx3d_amd/csrc/lbm_launch.hpp), compiled
with hipcc into a small host program and run on the CPU: the residency cap reserves LDS in
whole granules of 1/128 of the CU's LDS so that exactly the intended number of blocks fits
(a 256-byte rounding let 54 528-byte blocks through, of which only 2, not 3, fit per CU);
block shapes and the XCD-mapped grid cover every cell of a row; the row grid is kept
exactly for rows that start on 128-byte lines, the cell-linear grid covers the others."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

PROGRAM = r'''
#include "lbm_launch.hpp"
#include <cstdio>
#include <cstring>
#include <vector>
int main() {
  using namespace fx;
  for (unsigned bt : {64u, 128u, 256u}) for (unsigned w : {2u, 3u, 4u, 5u, 6u, 7u, 8u})
    printf("cap %u %u %zu\n", bt, w, occupancy_lds(w, bt, 163840, 65536));
  for (unsigned nx : {1u, 20u, 64u, 150u, 256u, 300u, 384u, 520u, 1030u, 1536u}) {
    LaunchShape s = lattice_shape(nx, 7, 5), m = stream_shape(nx, 7, 5, true);
    printf("shape %u %u %u %u %u %u %u\n", nx, s.block.x, s.grid.x, s.grid.y, s.grid.z, m.grid.x, m.xchunks);
  }
  printf("tune %u %u %u %u %u\n", tune_waves_cap(19, ST_FP32, 0), tune_waves_cap(27, ST_FP32, 0), tune_waves_cap(19, ST_FP16S, 0),
         tune_waves_cap(19, ST_FP32, EXT_EQUILIBRIUM_BOUNDARIES), tune_waves_cap(19, ST_FP32, EXT_ALL));
  printf("large %d %d %d\n", (int)tune_linear_large(19, ST_FP32, 0), (int)tune_linear_large(27, ST_FP32, EXT_ALL), (int)tune_linear_large(19, ST_FP16S, 0));
  LbmKernels k{};
  k.stream_collide_linear_large = true;
  // N*q*4 bytes at and just above the 4 GiB limit (q = 16 keeps the arithmetic exact)
  const uint64_t n_lim = LARGE_DDF_BYTES / 64;
  printf("capsel %d %d %d %d\n", (int)use_linear_grid(k, 256, 4, n_lim, 16), (int)use_linear_grid(k, 256, 4, n_lim + 1, 16),
         (int)use_linear_grid(k, 200, 4, 1000, 19), (int)use_linear_grid(k, 256, 4, 1000, 19));
  k.stream_collide_linear_large = false;
  // 16-bit: rows off the lines take the cell-linear grid unless a slot reaches 4 GiB
  printf("capsel16 %d %d %d %d\n", (int)use_linear_grid(k, 256, 2, 1ull << 30, 19), (int)use_linear_grid(k, 96, 2, 1000, 19),
         (int)use_linear_grid(k, 96, 2, (1ull << 31) - 1, 19), (int)use_linear_grid(k, 96, 2, 1ull << 31, 19));
  for (unsigned long long n : {1ull, 255ull, 256ull, 257ull, 8000000ull, 906000000ull}) for (int xcd = 0; xcd < 2; xcd++) {
    LaunchShape l = linear_shape(n, xcd != 0);
    printf("linear %llu %d %u %u %u %u %d\n", n, xcd, l.grid.x, l.grid.y, l.grid.z, l.block.x, (int)l.lin);
  }
  // segment order: every segment exactly once, each XCD (L mod 8) on runs of 8 consecutive
  // segments, for padded grids; the identity for unpadded ones
  for (unsigned long long blocks : {64ull, 128ull, 31296ull, 100ull}) {
    std::vector<int> seen(blocks, 0);
    unsigned long long runs_bad = 0, moved = 0;
    for (unsigned long long L = 0; L < blocks; L++) {
      const unsigned long long g = linear_segment(L, blocks, true);
      if (g < blocks) seen[g]++;
      moved += g != L;
      if (blocks % 64 == 0 && linear_segment(L, blocks, true) / 8 != (L / 64) * 8 + L % 8) runs_bad++;
    }
    int once = 1;
    for (int v : seen) once &= v == 1;
    printf("segment %llu %d %llu %llu\n", blocks, once, runs_bad, moved);
  }
  for (unsigned nx : {20u, 32u, 64u, 150u, 200u, 256u, 300u, 320u, 384u, 520u})
    printf("lines %u %d %d\n", nx, (int)rows_on_lines(nx, 4), (int)rows_on_lines(nx, 2));
  // lattices of up to ~5.2 G cells (D3Q19 FP16 in 288 GB): every 1-D grid within 2^32-1
  // work-items, else the 3-D row grid
  const unsigned big[][3] = {{1536u, 768u, 768u}, {2048u, 1536u, 1536u}, {4096u, 1024u, 1024u}, {256u, 4096u, 4096u},
                             {1024u, 2048u, 2048u}, {512u, 65536u, 128u}};
  for (const auto& d : big) {
    LaunchShape m = stream_shape(d[0], d[1], d[2], true);
    printf("big %u %u %u %u %u %u %u %u\n", d[0], d[1], d[2], m.grid.x, m.grid.y, m.grid.z, m.block.x, m.xchunks);
    const unsigned long long n = (unsigned long long)d[0] * d[1] * d[2];
    LbmKernels kl{};
    kl.stream_collide_linear_large = true;
    kl.stream_collide_linear_xcd = true;
    // rows off the lines (Nx+1) and the large-FP32 rule, with the 1-D linear grid's size
    printf("biglin %u %u %u %d %d %llu\n", d[0], d[1], d[2], (int)use_linear_grid(kl, d[0] + 1u, 4, n, 19),
           (int)use_linear_grid(kl, d[0], 4, n, 19), (unsigned long long)linear_blocks(n, true));
  }
  // two cells per lane: Nx/2 lanes per row filling whole waves (Nx a multiple of 128), also
  // the lattices above 4.29 G cells
  for (unsigned nx : {64u, 128u, 192u, 256u, 512u, 1536u, 4096u}) printf("pairrows %u %d\n", nx, (int)pair_rows(nx));
  for (unsigned nx : {128u, 256u, 384u, 512u, 1536u, 4096u}) {
    LaunchShape p = pair_shape(nx, 7, 5);
    printf("pair %u 7 5 %u %u %u %u %u %d\n", nx, p.block.x, p.grid.x, p.grid.y, p.grid.z, p.xchunks, (int)p.pairs);
  }
  for (const auto& d : big) {
    LaunchShape p = pair_shape(d[0], d[1], d[2]);
    printf("pair %u %u %u %u %u %u %u %u %d\n", d[0], d[1], d[2], p.block.x, p.grid.x, p.grid.y, p.grid.z, p.xchunks, (int)p.pairs);
  }
  {
    constexpr uint32_t E = EXT_EQUILIBRIUM_BOUNDARIES, VF = EXT_VOLUME_FORCE, FF = EXT_FORCE_FIELD, POR = EXT_VOLUME_FORCE | EXT_POROSITY | EXT_PRECISE_DENSITY;
    const uint32_t exts[6] = {0u, E, VF, E | FF, POR, EXT_ALL};
    for (int st = 0; st < 3; st++) for (int trt = 0; trt < 2; trt++) for (int q : {19, 27}) {
      printf("off32 %d %d %d", q, st, trt);
      for (uint32_t e : exts) printf(" %d", (int)tune_offsets32(q, (uint32_t)st, trt != 0, e));
      printf("\n");
      printf("tunepairs %d %d %d", q, st, trt);
      for (uint32_t e : exts) printf(" %d", (int)tune_pairs(q, (uint32_t)st, trt != 0, e));
      printf("\n");
      printf("tunebuf %d %d %d", q, st, trt);
      for (uint32_t e : exts) printf(" %d", (int)tune_buffer(q, (uint32_t)st, trt != 0, e));
      printf("\n");
      printf("tunepairbuf %d %d %d", q, st, trt);
      for (uint32_t e : exts) printf(" %d", (int)tune_pair_buffer(q, (uint32_t)st, trt != 0, e));
      printf("\n");
      printf("tunelinbuf %d %d %d", q, st, trt);
      for (uint32_t e : exts) printf(" %d", (int)tune_linear_buffer(q, (uint32_t)st, trt != 0, e));
      printf("\n");
    }
  }
  printf("fits %d %d %d\n", (int)grid_fits(16777215ull, 256u), (int)grid_fits(16777216ull, 256u), (int)grid_fits(4294967295ull, 1u));
  // TRT antisymmetric rate as the host forms it, for 2000 rates across (0, 2)
  for (int k = 1; k <= 2000; k++) {
    const float w = (float)k * 0.000999f;
    const float wm = trt_minus_rate(w);
    unsigned wb, mb; memcpy(&wb, &w, 4); memcpy(&mb, &wm, 4);
    printf("trt %08x %08x\n", wb, mb);
  }
  return 0;
}
'''


@pytest.fixture(scope="module")
def launch_facts(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("launch")
    src, exe = d / "launch.hip", d / "launch"
    src.write_text(PROGRAM)
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "fluidx3d_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    caps, shapes, tune, lines = {}, {}, None, {}
    for line in out.splitlines():
        p = line.split()
        if p[0] == "cap":
            caps[(int(p[1]), int(p[2]))] = int(p[3])
        elif p[0] == "shape":
            shapes[int(p[1])] = tuple(int(v) for v in p[2:])
        elif p[0] in ("tune", "large", "capsel", "capsel16"):
            tune = (tune or {}) | {p[0]: tuple(int(v) for v in p[1:])}
        elif p[0] == "lines":
            lines[int(p[1])] = (bool(int(p[2])), bool(int(p[3])))
        elif p[0] == "linear":
            lines[("linear", int(p[1]), int(p[2]))] = tuple(int(v) for v in p[3:])
        elif p[0] == "trt":
            lines.setdefault("trt", []).append((int(p[1], 16), int(p[2], 16)))
        elif p[0] == "segment":
            lines[("segment", int(p[1]))] = tuple(int(v) for v in p[2:])
        elif p[0] in ("big", "biglin", "pair", "tunepairs", "pairrows", "off32", "tunebuf", "tunepairbuf", "tunelinbuf"):
            lines.setdefault(p[0], []).append(tuple(int(v) for v in p[1:]))
        elif p[0] == "fits":
            lines["fits"] = tuple(int(v) for v in p[1:])
    return caps, shapes, tune, lines


def test_residency_cap_reserves_whole_granules(launch_facts):
    caps, _, _, _ = launch_facts
    lds_per_cu, granule = 163840, 163840 // 128
    assert caps[(256, 3)] == 53760  # 3 blocks of 4 waves: 3 waves/SIMD
    # 2 blocks of 4 waves: the exact share (80 KiB) exceeds the 64 KiB per-block limit, the
    # largest whole number of granules below it still leaves room for only 2 blocks
    assert caps[(256, 2)] == 65280
    for (threads, waves), lds in caps.items():
        if lds == 0:
            continue
        blocks = waves * 4 // (threads // 64)
        assert lds % granule == 0 and lds <= 65536
        assert lds_per_cu // lds == blocks, (threads, waves, lds)
    # 64-thread blocks at 5 waves/SIMD would need 20 blocks, which no whole number of
    # granules gives exactly (7680 B: 21 blocks): no reservation rather than a wrong one
    assert caps[(64, 5)] == 0


def test_block_shapes_cover_the_rows(launch_facts):
    _, shapes, _, _ = launch_facts
    Ny, Nz = 7, 5
    for nx, (bx, gx, gy, gz, mx, chunks) in shapes.items():
        assert bx in (64, 128, 256) and gx * bx >= nx and (gx - 1) * bx < nx and (gy, gz) == (Ny, Nz)
        assert bx == 64 or gx * bx - nx < 64  # less than one idle wave per row
        if gx > 1:  # XCD-mapped 1-D grid: groups of 8 rows, every chunk of every row
            assert chunks == gx and mx == ((Ny * Nz + 7) // 8) * 8 * gx
        else:
            assert chunks == 0


def test_tuning_table(launch_facts):
    _, _, tune, _ = launch_facts
    # plain FP32 D3Q19/D3Q27 at 3 waves/SIMD, FP16 uncapped, EQUILIBRIUM_BOUNDARIES at 4, generic uncapped
    assert tune["tune"] == (3, 3, 0, 4, 0)
    # above 4 GiB of DDFs the FP32 kernels take the cell-linear grid, the 16-bit ones keep rows
    assert tune["large"] == (1, 1, 0)
    # strictly above the limit, or wherever the rows miss the 128-byte lines
    assert tune["capsel"] == (0, 1, 1, 0)
    assert tune["capsel16"] == (0, 1, 1, 0)


def test_grid_follows_row_alignment(launch_facts):
    """Rows on 128-byte lines keep the row grid; the others take the cell-linear grid, whose
    blocks of 256 consecutive cells cover the lattice exactly once."""
    _, _, _, lines = launch_facts
    for key, v in lines.items():
        if key in ("trt", "big", "biglin", "fits", "pair", "tunepairs", "pairrows", "off32", "tunebuf", "tunepairbuf", "tunelinbuf"):
            continue
        if isinstance(key, tuple) and key[0] == "linear":
            n, xcd = key[1], key[2]
            gx, gy, gz, bx, lin = v
            need = (n + 255) // 256
            assert lin == 1 and bx == 256 and (gy, gz) == (1, 1), key
            if xcd:  # padded to whole groups of 64 blocks for the XCD-aware segment order
                assert gx % 64 == 0 and need <= gx < need + 64, key
            else:
                assert gx == need, key
        elif isinstance(key, tuple) and key[0] == "segment":
            once, runs_bad, moved = v
            assert once == 1 and runs_bad == 0, key  # a permutation; XCD k runs segments 8k..8k+7 of each group
            assert (moved == 0) == (key[1] % 64 != 0), key  # identity exactly for unpadded grids
        else:
            fp32, fp16 = v
            assert fp32 == (key * 4 % 128 == 0) and fp16 == (key * 2 % 128 == 0), key
    assert lines[256] == (True, True) and lines[32] == (True, False) and lines[200] == (False, False)


def test_trt_rate_formed_on_the_host(launch_facts):
    """The kernels take the TRT antisymmetric rate 1/(0.1875/(1/w - 0.5) + 0.5) from the host
    (src/kernel.cpp:1605-1633, Lambda = 3/16): the host's float operations must give the
    bits that the same IEEE binary32 operations give, here evaluated with numpy float32."""
    import numpy as np
    pairs = launch_facts[3]["trt"]
    assert len(pairs) == 2000
    w = np.array([a for a, _ in pairs], np.uint32).view(np.float32)
    f = np.float32
    expect = (f(1.0) / (f(0.1875) / (f(1.0) / w - f(0.5)) + f(0.5))).astype(np.float32)
    got = np.array([b for _, b in pairs], np.uint32).view(np.float32)
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32))


def test_large_lattices_stay_within_the_launch_limit(launch_facts):
    """A launch holds at most 2^32-1 work-items per grid dimension.  Above that the XCD-mapped
    1-D grid falls back to the 3-D row grid (chunks, Ny, Nz), which covers every cell, and the
    cell-linear grid is not taken; below it both are kept (906M cells: XCD-mapped)."""
    lines = launch_facts[3]
    assert lines["fits"] == (1, 0, 1)
    limit = 2 ** 32 - 1
    for Nx, Ny, Nz, gx, gy, gz, bx, chunks in lines["big"]:
        row_chunks = -(-Nx // bx)
        assert gx * bx <= limit and gy <= 65536 and gz <= 65536
        if chunks:  # XCD-mapped 1-D grid over every row
            assert (gy, gz) == (1, 1) and chunks == row_chunks and gx == -(-(Ny * Nz) // 8) * 8 * chunks
        else:
            assert (gx, gy, gz) == (row_chunks, Ny, Nz)
        if Nx * Ny * Nz <= 906_000_000:
            assert chunks, "lattices below the limit keep the XCD-mapped grid"
    assert sum(1 for b in lines["big"] if b[7] == 0) >= 3  # the ones above 4.29 G cells fall back
    for Nx, Ny, Nz, off_lines, large_fp32, blocks in lines["biglin"]:
        fits = blocks * 256 <= limit
        assert off_lines == int(fits) and large_fp32 == int(fits), (Nx, Ny, Nz)


def test_pair_grid_covers_every_lane(launch_facts):
    """k_stream_collide_pair: Nx/2 lanes per row (two cells each), blocks of 64/128/256
    lanes, every chunk of every row once (XCD-mapped when a row takes several blocks, the
    3-D row grid above the launch limit), within 2^32-1 work-items; selected only for the
    D3Q19 16-bit kernels where it measured faster: SRT and TRT without extensions, with
    EQUILIBRIUM_BOUNDARIES (+ FORCE_FIELD); TRT porous; SRT VOLUME_FORCE for FP16S and SRT
    porous for FP16C; never with MOVING_BOUNDARIES (the all-extension kernels)."""
    lines = launch_facts[3]
    limit = 2 ** 32 - 1
    assert len(lines["pair"]) == 12
    for Nx, Ny, Nz, bx, gx, gy, gz, chunks, pairs in lines["pair"]:
        lanes = Nx // 2
        row_chunks = -(-lanes // bx)
        assert pairs == 1 and Nx % 128 == 0 and bx in (64, 128, 256) and lanes % 64 == 0
        assert row_chunks * bx >= lanes and (row_chunks - 1) * bx < lanes
        assert gx * bx <= limit and gy <= 65536 and gz <= 65536
        if chunks:
            assert (gy, gz) == (1, 1) and chunks == row_chunks and gx == -(-(Ny * Nz) // 8) * 8 * chunks
        else:
            assert (gx, gy, gz) == (row_chunks, Ny, Nz)
    table = {(q, st, trt): flags for q, st, trt, *flags in lines["tunepairs"]}
    assert len(table) == 12
    for (q, st, trt), flags in table.items():   # none, E, VF, E+FF, porous, all
        if q == 27 or st == 0:
            assert flags == [0, 0, 0, 0, 0, 0], (q, st, trt)
        elif trt:
            assert flags == [1, 1, 0, 1, 1, 0], (q, st, trt)
        else:
            assert flags == [1, 1, int(st == 1), 1, int(st == 2), 0], (q, st, trt)
    assert dict((r[0], r[1]) for r in lines["pairrows"]) == {64: 0, 128: 1, 192: 0, 256: 1, 512: 1, 1536: 1, 4096: 1}


def test_offsets32_selection(launch_facts):
    """32-bit slot offsets on the row grid only for the 16-bit kernels where they measured
    faster: D3Q27 every SRT subset but the all-extension one, TRT with
    EQUILIBRIUM_BOUNDARIES (+ FORCE_FIELD) and the porous subset; D3Q19 the TRT subsets
    other than VOLUME_FORCE and the all-extension one, where rows of whole waves of lanes
    do not take two cells per lane instead."""
    table = {(q, st, trt): flags for q, st, trt, *flags in launch_facts[3]["off32"]}
    assert len(table) == 12
    for (q, st, trt), flags in table.items():   # none, E, VF, E+FF, porous, all
        if st == 0:
            assert flags == [0, 0, 0, 0, 0, 0], (q, st, trt)
        elif q == 27 and trt == 0:
            assert flags == [1, 1, 1, 1, 1, 0], (q, st, trt)
        elif q == 27:
            assert flags == [0, 1, 0, 1, 1, 0], (q, st, trt)
        elif trt == 0:
            assert flags == [0, 0, 0, 0, 0, 0], (q, st, trt)
        else:
            assert flags == [1, 1, 0, 1, 1, 0], (q, st, trt)


def test_buffer_selection(launch_facts):
    """Buffer instructions on the row grid only for the 16-bit kernels that keep one cell per
    lane and measured faster with them: D3Q27 every subset but SRT none and SRT VOLUME_FORCE
    with FP16S; D3Q19 the TRT VOLUME_FORCE and all-extension kernels, SRT porous (FP16S) and
    SRT VOLUME_FORCE (FP16C); never FP32 and never a kernel that takes two cells per lane."""
    lines = launch_facts[3]
    buf = {(q, st, trt): flags for q, st, trt, *flags in lines["tunebuf"]}
    pairs = {(q, st, trt): flags for q, st, trt, *flags in lines["tunepairs"]}
    assert len(buf) == 12
    for (q, st, trt), flags in buf.items():   # none, E, VF, E+FF, porous, all
        if st == 0:
            assert flags == [0, 0, 0, 0, 0, 0], (q, st, trt)
        elif q == 27:
            assert flags == ([1] * 6 if trt else [0, 1, int(st == 2), 1, 1, 1]), (q, st, trt)
        else:
            assert not any(b and p for b, p in zip(flags, pairs[(q, st, trt)])), (q, st, trt)
            expect = [0, 0, 1, 0, 0, 1] if trt else [0, 0, int(st == 2), 0, int(st == 1), 0]
            assert flags == expect, (q, st, trt)


def test_pair_buffer_selection(launch_facts):
    """The two-cells-per-lane kernel takes buffer instructions wherever it runs, except the
    FP16S SRT kernels without forcing (none, EQUILIBRIUM_BOUNDARIES, + FORCE_FIELD), which
    measured 1 % faster with global addressing."""
    lines = launch_facts[3]
    pb = {(q, st, trt): flags for q, st, trt, *flags in lines["tunepairbuf"]}
    pairs = {(q, st, trt): flags for q, st, trt, *flags in lines["tunepairs"]}
    assert len(pb) == 12
    for key, flags in pb.items():   # none, E, VF, E+FF, porous, all
        q, st, trt = key
        expect = [p and not (st == 1 and not trt and j in (0, 1, 3)) for j, p in enumerate(pairs[key])]
        assert flags == [int(e) for e in expect], key


def test_linear_buffer_selection(launch_facts):
    """On the cell-linear grid every 16-bit kernel takes buffer instructions and no FP32 one."""
    table = {(q, st, trt): flags for q, st, trt, *flags in launch_facts[3]["tunelinbuf"]}
    assert len(table) == 12
    for (q, st, trt), flags in table.items():
        assert flags == [int(st != 0)] * 6, (q, st, trt)
