"""Benchmark: D3Q19 SRT stream_collide on a 256^3 periodic Taylor-Green vortex, one MI355X.

One step = one stream_collide launch over all 256^3 cells (Esoteric-Pull, in place).
Prints ONE JSON line.  `value` is MLUPs/s of the FP32-storage run (the reference's
headline configuration, src/setup.cpp:5-36 with README.md:1137); FP16S and FP16C are
measured the same way and reported under "modes".

Timing: W untimed warm-up steps, then exactly K steps launched back to back on the
context stream and bracketed by HIP events on that stream (fx_lbm_run), with a device
synchronisation before and after.  MLUPs/s = N*K/t.  roofline.achieved = algorithmic
bytes per launch (bandwidth_bytes_per_cell_device, src/lbm.cpp:51-72: 153 B FP32, 77 B
FP16 per cell) / average launch duration; peak = 8000 GB/s (MI355X).
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
    p.add_argument("--steps", type=int, default=10000)
    p.add_argument("--warmup", type=int, default=200)
    p.add_argument("--size", type=int, default=256)
    p.add_argument("--modes", default="FP32,FP16S,FP16C")
    p.add_argument("--sync-steps", type=int, default=200, help="steps of the per-step-synchronised figure")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-adjoint", action="store_true")
    return p.parse_args()


def read_traffic(mode):
    """HBM bytes per stream_collide launch from the committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", "stream_collide_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch\[(\w+)\]\s*=\s*([0-9.eE+]+)", line)
        if m and m.group(1) == mode:
            return float(m.group(2))
    return None


def gpu_mode(fx, size, storage, steps, warmup, sync_steps):
    from tests import scenes
    from fluidx3d_amd import _abi
    lbm = fx.LBM(size, size, size, 1.0 / 6.0, storage=storage)
    rho, u, flags = scenes.taylor_green(size, size, size)
    lbm.rho.host[:] = rho
    lbm.u.host[:] = u
    lbm.flags.host[:] = flags
    lbm.initialize()
    L = lbm.ctx.lib
    import ctypes
    ms = ctypes.c_float(0.0)
    args = (lbm.ctx.h, lbm.fi, lbm.rho.device, lbm.u.device, lbm.flags.device)
    _abi.check(L.fx_lbm_run(*args, 0, warmup, 0.0, 0.0, 0.0, None, None, ctypes.byref(ms)))
    t = warmup
    lbm.ctx.sync()
    wall0 = time.perf_counter()
    _abi.check(L.fx_lbm_run(*args, t, steps, 0.0, 0.0, 0.0, None, None, ctypes.byref(ms)))
    lbm.ctx.sync()
    wall = time.perf_counter() - wall0
    t += steps
    event_s = ms.value * 1e-3
    N = lbm.N
    bpc = fx.bandwidth_bytes_per_cell(lbm.cfg)
    # reference-style figure: one launch + one blocking synchronisation per step (src/lbm.cpp:951, 964-973)
    t0 = time.perf_counter()
    for _ in range(sync_steps):
        _abi.check(L.fx_lbm_stream_collide(*args, t, 0.0, 0.0, 0.0, None, None))
        lbm.ctx.sync()
        t += 1
    sync_s = (time.perf_counter() - t0) / sync_steps
    mass = lbm.mass()
    lbm.u.read_from_device()
    finite = bool(np.isfinite(lbm.u.host).all())
    lbm.close()
    per_launch = event_s / steps
    return {
        "mlups": N * steps / event_s * 1e-6,
        "ms_per_step": per_launch * 1e3,
        "wall_mlups": N * steps / wall * 1e-6,
        "bytes_per_cell": bpc,
        "achieved_GBps": bpc * N / per_launch * 1e-9,
        "frac": bpc * N / per_launch * 1e-9 / PEAK_HBM_GBPS,
        "per_step_sync_mlups": N / sync_s * 1e-6,
        "mass_rel_drift": abs(mass - N) / N,
        "finite": finite,
    }


def adjoint_timing(fx, dims, steps, window, interval):
    """Checkpointed adjoint of the porous-channel objective: GPU time of fx_lbm_forward
    (checkpoints + objective) and fx_lbm_backward (recompute + adjoint) per step, against
    the plain forward step (stream_collide with the same extensions) on the same grid."""
    import ctypes
    from fluidx3d_amd import _abi
    from fluidx3d_amd.adjoint import Adjoint
    from tests import scenes
    rho, u, flags, alpha = scenes.channel_c4(*dims)
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
    results = {}
    for m in modes:
        results[m] = gpu_mode(fx, a.size, m, a.steps, a.warmup, a.sync_steps)
        print(f"# {m}: {results[m]['mlups']:.0f} MLUPs/s, {results[m]['achieved_GBps']:.0f} GB/s "
              f"({100 * results[m]['frac']:.1f}% of 8 TB/s)", file=sys.stderr, flush=True)
    copy_gbps = measured_copy_peak(fx)
    adjoint = None
    if not a.no_adjoint:
        adjoint = {"scaled": adjoint_timing(fx, (384, 192, 192), 5000, 3750, 0),
                   "channel_128x64x64": adjoint_timing(fx, (128, 64, 64), 2000, 1500, 100)}
        print(f"# adjoint: backward/plain-forward {adjoint['scaled']['backward_over_plain_forward']}x at 384x192x192",
              file=sys.stderr, flush=True)
    primary = "FP32" if "FP32" in results else modes[0]
    r = results[primary]
    traffic = read_traffic(primary)
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
        "modes": {m: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in results[m].items()} for m in results},
        "reference_mlups_mi355x": REFERENCE_MLUPS,
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
