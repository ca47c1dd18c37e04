"""Benchmark: D3Q19 SRT stream_collide on a 256^3 periodic Taylor-Green vortex, one MI355X.

One step = one stream_collide launch over all 256^3 cells (Esoteric-Pull, in place).
Prints ONE JSON line.  `value` is MLUPs/s of the FP32-storage run (the reference's
headline configuration, src/setup.cpp:5-36 with README.md:1137); FP16S and FP16C are
measured the same way and reported under "modes".

Timing, per storage mode: W untimed warm-up steps, then the reference-style figure (one
launch and one blocking synchronisation per step, src/lbm.cpp:951, 964-973), which also
brings the clocks to steady state, then `--repeats` batches of exactly K steps launched
back to back on the context stream and bracketed by HIP events on that stream
(fx_lbm_run), with a device synchronisation before and after each.  MLUPs/s = N*K/t of
the median batch (min and max reported beside it).  roofline.achieved = algorithmic bytes
per launch (bandwidth_bytes_per_cell_device, src/lbm.cpp:51-72: 153 B FP32, 77 B FP16
per cell) / the median batch's average launch duration; peak = 8000 GB/s (MI355X).
The device copy peak is measured first.  Beside the headline: the two sphere-flow
lattices (512x256x256 D3Q27 TRT FP16C with equilibrium boundaries and the force field,
1536x768x768 D3Q19 SRT FP16S), the checkpointed adjoint against the plain forward step,
and the CPU baseline.
"""
import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBPS = 8000.0
REFERENCE_MLUPS = {"FP32": 30380.0, "FP16S": 54494.0, "FP16C": 40157.0}  # README.md:1137, 1x MI355X


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=100)
    p.add_argument("--repeats", type=int, default=3, help="event-timed batches of K steps per mode (median reported)")
    p.add_argument("--size", type=int, default=256)
    p.add_argument("--modes", default="FP32,FP16S,FP16C")
    p.add_argument("--sync-steps", type=int, default=200, help="steps of the per-step-synchronised figure")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--no-spheres", action="store_true", help="skip the sphere-flow legs (512x256x256 D3Q27, 1536x768x768)")
    return p.parse_args()


def read_traffic(mode, size):
    """HBM bytes per stream_collide launch from the committed rocprofv3 --pmc summary, which
    holds them for one lattice size (lines `traffic_bytes_per_launch[MODE@S^3] = bytes`);
    None for a size it does not cover."""
    path = os.path.join(ROOT, "profiles", "stream_collide_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch\[(\w+)@(\d+)\^3\]\s*=\s*([0-9.eE+]+)", line)
        if m and m.group(1) == mode and int(m.group(2)) == size:
            return float(m.group(3))
    return None


def timed_batches(L, args, t, steps, repeats, fx_=0.0, F=None, alpha=None):
    """`repeats` batches of exactly `steps` stream_collide launches back to back on the
    context stream, each bracketed by HIP events on that stream (fx_lbm_run); returns the
    per-batch milliseconds and the next t."""
    import ctypes
    from fluidx3d_amd import _abi
    ms = ctypes.c_float(0.0)
    out = []
    for _ in range(repeats):
        _abi.check(L.fx_lbm_run(*args, t, steps, fx_, 0.0, 0.0, F, alpha, ctypes.byref(ms)))
        out.append(ms.value)
        t += steps
    return out, t


def sync_leg(L, ctx, args, t, steps, F=None):
    """Reference-style figure: one launch + one blocking synchronisation per step
    (src/lbm.cpp:951, 964-973); seconds per step and the next t."""
    from fluidx3d_amd import _abi
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        _abi.check(L.fx_lbm_stream_collide(*args, t, 0.0, 0.0, 0.0, F, None))
        ctx.sync()
        t += 1
    return (time.perf_counter() - t0) / max(steps, 1), t


def rate_summary(N, bpc, batch_ms, steps):
    """MLUPs/s of the median batch (min and max beside it) and its roofline figures."""
    per_launch = sorted(b / steps for b in batch_ms)
    med = per_launch[len(per_launch) // 2]
    gbps = bpc * N / (med * 1e-3) * 1e-9
    return {
        "mlups": N / (med * 1e-3) * 1e-6,
        "mlups_min": N / (per_launch[-1] * 1e-3) * 1e-6,
        "mlups_max": N / (per_launch[0] * 1e-3) * 1e-6,
        "ms_per_step": med,
        "batches": len(per_launch),
        "bytes_per_cell": bpc,
        "achieved_GBps": gbps,
        "frac": gbps / PEAK_HBM_GBPS,
    }


def gpu_mode(fx, size, storage, steps, warmup, sync_steps, repeats):
    """One storage mode of the 256^3 Taylor-Green benchmark.  Order: initialize, W warm-up
    steps, the per-step-synchronised leg (which also brings the clocks to steady state),
    then `repeats` event-timed batches of exactly K steps; the median batch is the rate."""
    from tests import scenes
    lbm = fx.LBM(size, size, size, 1.0 / 6.0, storage=storage)
    rho, u, flags = scenes.taylor_green(size, size, size)
    lbm.rho.host[:] = rho
    lbm.u.host[:] = u
    lbm.flags.host[:] = flags
    lbm.initialize()
    m0 = lbm.mass()
    L = lbm.ctx.lib
    args = (lbm.ctx.h, lbm.fi, lbm.rho.device, lbm.u.device, lbm.flags.device)
    _, t = timed_batches(L, args, 0, warmup, 1)
    sync_s, t = sync_leg(L, lbm.ctx, args, t, sync_steps)
    lbm.ctx.sync()
    wall0 = time.perf_counter()
    batch_ms, t = timed_batches(L, args, t, steps, repeats)
    lbm.ctx.sync()
    wall = time.perf_counter() - wall0
    lbm.t = t
    N = lbm.N
    out = rate_summary(N, fx.bandwidth_bytes_per_cell(lbm.cfg), batch_ms, steps)
    out["wall_mlups"] = N * steps * repeats / wall * 1e-6
    out["per_step_sync_mlups"] = N / sync_s * 1e-6
    out["mass_rel_drift"] = f"{abs(lbm.mass() - m0) / m0:.3e}"
    lbm.u.read_from_device()
    out["finite"] = bool(np.isfinite(lbm.u.host).all())
    lbm.close()
    return out


def sphere_leg(fx, name, dims, radius, velocity_set, collision, storage, extensions, steps, warmup, sync_steps, repeats,
               perturb):
    """Flow past a sphere (the C++ scene of fluidx3d_amd/host/setup.cpp): TYPE_S|TYPE_X
    sphere() of radius R at (R*4, R*4, R*4), TYPE_E inflow/outflow u = (0.075, 0, 0) at
    x = 0 and x = Nx-1, periodic in y and z, nu for Re = 100 on the diameter; `perturb`
    adds a seeded uniform perturbation of amplitude 1e-4 (numpy PCG64, seed 42) to u of the
    other cells.  Timed like the benchmark (the drag after 20000 steps is a test, not a
    bench figure)."""
    from tests import scenes
    Nx, Ny, Nz = dims
    u_in = 0.075
    nu = u_in * 2.0 * radius / 100.0
    lbm = fx.LBM(Nx, Ny, Nz, nu, velocity_set=velocity_set, collision=collision, storage=storage,
                 extensions=extensions)
    N = lbm.N
    c = 4.0 * radius
    plane = Nx * Ny
    rng = np.random.default_rng(42) if perturb else None
    x = np.arange(Nx, dtype=np.float32) - np.float32(c)
    y = np.arange(Ny, dtype=np.float32) - np.float32(c)
    dxy = (x * x)[None, :] + (y * y)[:, None]
    r2 = np.float32(radius) * np.float32(radius)
    lbm.u.x[:] = u_in
    for z in range(Nz):  # plane by plane: 906M-cell lattices stay within a few GB of temporaries
        sl = slice(z * plane, (z + 1) * plane)
        dz = np.float32(z) - np.float32(c)
        solid = (dxy + dz * dz <= r2).ravel()
        f = np.zeros(plane, dtype=np.uint8)
        f[solid] = scenes.TYPE_S | scenes.TYPE_X
        edge = np.zeros((Ny, Nx), dtype=bool)
        edge[:, 0] = edge[:, -1] = True
        edge = edge.ravel() & ~solid
        f[edge] = scenes.TYPE_E
        lbm.flags.host[sl] = f
        if perturb:
            p = rng.uniform(-1e-4, 1e-4, (3, plane)).astype(np.float32)
            keep = ~(solid | edge)
            lbm.u.x[sl][keep] += p[0][keep]
            lbm.u.y[sl][keep] = p[1][keep]
            lbm.u.z[sl][keep] = p[2][keep]
        lbm.u.x[sl][solid] = 0.0
    lbm.initialize()
    L = lbm.ctx.lib
    args = (lbm.ctx.h, lbm.fi, lbm.rho.device, lbm.u.device, lbm.flags.device)
    F = lbm.F.device if lbm.F is not None else None
    _, t = timed_batches(L, args, 0, warmup, 1, F=F)
    sync_s, t = sync_leg(L, lbm.ctx, args, t, sync_steps, F=F)
    batch_ms, t = timed_batches(L, args, t, steps, repeats, F=F)
    lbm.t = t
    out = {"workload": name}
    out.update(rate_summary(N, fx.bandwidth_bytes_per_cell(lbm.cfg), batch_ms, steps))
    if "FORCE_FIELD" in extensions and "VOLUME_FORCE" not in extensions:
        # the byte count above includes 12 B/cell of force field (src/lbm.cpp:60-62), but
        # without VOLUME_FORCE stream_collide never uses F, so neither kernel reads it
        # (src/kernel.cpp:1497-1503 adds it to a force that only VOLUME_FORCE applies)
        out["moved_bytes_per_cell"] = out["bytes_per_cell"] - 12
        out["moved_GBps"] = out["achieved_GBps"] * out["moved_bytes_per_cell"] / out["bytes_per_cell"]
    out["per_step_sync_mlups"] = N / sync_s * 1e-6
    out["steps_run"] = t
    lbm.close()
    return out


def adjoint_timing(fx, dims, steps, window, interval):
    """Checkpointed adjoint of the porous-channel objective: GPU time of fx_lbm_forward
    (checkpoints + objective) and fx_lbm_backward (recompute + adjoint) per step, against
    the plain forward step (stream_collide with the same extensions) on the same grid."""
    import ctypes
    from fluidx3d_amd import _abi
    from fluidx3d_amd.adjoint import Adjoint
    from tests import scenes
    rho, u, flags, alpha = scenes.porous_channel(*dims)
    lbm = fx.LBM(*dims, 1.0 / 6.0, 1e-6, 0.0, 0.0, extensions=("VOLUME_FORCE", "POROSITY", "PRECISE_DENSITY"))
    lbm.rho.host[:] = rho; lbm.u.host[:] = u; lbm.flags.host[:] = flags; lbm.porosity.host[:] = alpha
    lbm.initialize()
    ms = ctypes.c_float()
    args = (lbm.ctx.h, lbm.fi, lbm.rho.device, lbm.u.device, lbm.flags.device)
    plain_steps = min(steps, 500)
    _abi.check(lbm.ctx.lib.fx_lbm_run(*args, 0, 20, lbm.fx, 0.0, 0.0, None, lbm.porosity.device, ctypes.byref(ms)))
    _abi.check(lbm.ctx.lib.fx_lbm_run(*args, 20, plain_steps, lbm.fx, 0.0, 0.0, None, lbm.porosity.device, ctypes.byref(ms)))
    plain = ms.value / plain_steps
    adj = Adjoint(lbm, steps, window, dims[0] - 1, interval)
    mem = Adjoint.memory_bytes(lbm, steps, window, dims[0] - 1, interval)
    J = adj.forward()
    grad = adj.backward()
    fwd, bwd = adj.forward_ms / steps, adj.backward_ms / steps
    adj.close()
    lbm.close()
    return {"workload": f"{dims[0]}x{dims[1]}x{dims[2]} porous channel, {steps} steps, checkpoint every {adj.cfg.checkpoint_interval or 'ceil(sqrt(T))'} steps",
            "plain_forward_ms_per_step": round(plain, 5), "forward_ms_per_step": round(fwd, 5), "backward_ms_per_step": round(bwd, 5),
            "backward_over_plain_forward": round(bwd / plain, 3), "forward_plus_backward_over_plain": round((fwd + bwd) / plain, 3),
            "workspace_GB": round(mem / 1e9, 2), "objective": J, "grad_finite": bool(np.isfinite(grad).all())}


def measured_copy_peak(fx):
    from fluidx3d_amd import _abi
    import ctypes
    ctx = _abi.Context(fx.make_config(4, 4, 4, 1.0))
    nbytes = 4 << 30
    a = ctx.alloc(nbytes)
    b = ctx.alloc(nbytes)
    ctx.fill(a, nbytes, 0)
    ms = ctypes.c_float()
    _abi.check(ctx.lib.fx_bench_copy(ctx.h, b, a, nbytes, 2, ctypes.byref(ms)))
    _abi.check(ctx.lib.fx_bench_copy(ctx.h, b, a, nbytes, 10, ctypes.byref(ms)))
    ctx.close()
    return 2.0 * nbytes * 10 / (ms.value * 1e-3) * 1e-9


def cpu_opencl_devices():
    """CPU OpenCL devices visible to clinfo (the reference's CPU path needs one)."""
    import shutil
    import subprocess
    if not shutil.which("clinfo"):
        return "clinfo not installed"
    try:
        out = subprocess.run(["clinfo"], capture_output=True, text=True, timeout=60).stdout
    except Exception as e:  # noqa: BLE001
        return f"clinfo failed: {e}"
    cpus = [l for l in out.splitlines() if "Device Type" in l and "CPU" in l]
    return f"{len(cpus)} CPU OpenCL device(s)"


def cpu_baseline(size, seconds):
    """The oracle's FP32 restatement (a port, not the reference: the reference's CPU path
    needs a CPU OpenCL runtime, which the GPU host lacks), multithreaded, timed on a
    bounded number of steps."""
    from oracle import oracle as O
    from tests import scenes
    threads = min(16, os.cpu_count() or 1)
    ref = O.OracleLBM(size, size, size, nu=1.0 / 6.0, threads=threads)
    rho, u, flags = scenes.taylor_green(size, size, size)
    ref.rho[:] = rho
    ref.u[:] = u
    ref.initialize()
    ref.step(1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        ref.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 200:
            break
    return {"value": ref.N * steps / el * 1e-6, "unit": "MLUPs/s", "cores": threads, "kind": "port",
            "sample": f"{size}^3 D3Q19 SRT FP32 Taylor-Green, {steps} steps after 1 warm-up step, oracle FP32 "
                      f"restatement on {threads} threads of {platform.processor() or platform.machine()} "
                      f"({os.cpu_count()} logical CPUs); reference CPU OpenCL baseline unavailable "
                      f"(clinfo: {cpu_opencl_devices()}; the reference source is not shipped to the GPU host)"}


def main():
    a = parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {a.gpus} is not supported (multi-GPU domain decomposition "
              f"is out of scope for this build)", file=sys.stderr)
        sys.exit(2)
    import fluidx3d_amd as fx
    modes = [m.strip() for m in a.modes.split(",") if m.strip()]
    copy_gbps = measured_copy_peak(fx)
    print(f"# device copy: {copy_gbps:.0f} GB/s", file=sys.stderr, flush=True)
    results = {}
    for m in modes:
        results[m] = gpu_mode(fx, a.size, m, a.steps, a.warmup, a.sync_steps, a.repeats)
        r = results[m]
        print(f"# {m}: {r['mlups']:.0f} MLUPs/s (batches {r['mlups_min']:.0f}-{r['mlups_max']:.0f}), "
              f"{r['achieved_GBps']:.0f} GB/s ({100 * r['frac']:.1f}% of 8 TB/s); per-step-synchronised "
              f"{r['per_step_sync_mlups']:.0f}", file=sys.stderr, flush=True)
    spheres = None
    if not a.no_spheres:
        spheres = {}
        legs = [("sphere_512x256x256", "D3Q27 TRT FP16C flow past a sphere, 512x256x256, R=32, TYPE_E inflow/outflow "
                 "u=0.075, Re=100, EQUILIBRIUM_BOUNDARIES+FORCE_FIELD", (512, 256, 256), 32.0, 27, "TRT", "FP16C",
                 ("EQUILIBRIUM_BOUNDARIES", "FORCE_FIELD"), min(a.steps, 1000), True),
                ("sphere_1536x768x768", "D3Q19 SRT FP16S flow past a sphere, 1536x768x768 (906M cells), R=96, TYPE_E "
                 "inflow/outflow u=0.075, Re=100, EQUILIBRIUM_BOUNDARIES", (1536, 768, 768), 96.0, 19, "SRT", "FP16S",
                 ("EQUILIBRIUM_BOUNDARIES",), min(a.steps, 200), False)]
        for key, name, dims, radius, q, coll, st, ext, steps, perturb in legs:
            r = sphere_leg(fx, name, dims, radius, q, coll, st, ext, steps, min(a.warmup, 20), min(a.sync_steps, 50),
                           a.repeats, perturb)
            spheres[key] = {k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in r.items()}
            moved = f"; {r['moved_GBps']:.0f} GB/s moved ({r['moved_bytes_per_cell']} B/cell)" if "moved_GBps" in r else ""
            print(f"# {key}: {r['mlups']:.0f} MLUPs/s, {r['achieved_GBps']:.0f} GB/s ({100 * r['frac']:.1f}% of 8 TB/s){moved}",
                  file=sys.stderr, flush=True)
    adjoint = None
    if not a.no_adjoint:
        adjoint = {"scaled": adjoint_timing(fx, (384, 192, 192), 5000, 3750, 0),
                   "channel_128x64x64": adjoint_timing(fx, (128, 64, 64), 2000, 1500, 100)}
        print(f"# adjoint: backward/plain-forward {adjoint['scaled']['backward_over_plain_forward']}x at 384x192x192",
              file=sys.stderr, flush=True)
    primary = "FP32" if "FP32" in results else modes[0]
    r = results[primary]
    traffic = read_traffic(primary, a.size)
    out = {
        "metric": "MLUPs/s + %HBM roofline, D3Q19 FP32/FP16S 256^3; adjoint/forward time",
        "value": round(r["mlups"], 1),
        "unit": "MLUPs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(r["mlups"] / REFERENCE_MLUPS[primary], 4) if a.size == 256 else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"D3Q19 SRT {primary} stream_collide, {a.size}^3 periodic Taylor-Green vortex, nu=1/6",
                   "Nx": a.size, "Ny": a.size, "Nz": a.size, "velocity_set": "D3Q19", "collision": "SRT",
                   "storage": primary},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_GBps"], 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(r["frac"], 4), "traffic": traffic},
        "measured_copy_GBps": round(copy_gbps, 1),
        "frac_of_measured_copy": round(r["achieved_GBps"] / copy_gbps, 4),
        "modes": {m: {k: (float(f"{v:.6g}") if isinstance(v, float) else v) for k, v in results[m].items()} for m in results},
        "reference_mlups_mi355x": REFERENCE_MLUPS,
        "spheres": spheres,
        "adjoint": adjoint,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(min(a.size, 256), a.cpu_seconds)
        cb = out["cpu_baseline"]["value"]
        cb = round(cb, 2)
        out["cpu_baseline"]["value"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
