"""Writes tests/golden/adjoint_channel_fd.json: central finite differences of the adjoint
objective of the 128x64x64 porous channel (tests/scenes.py CHANNEL_128) with respect to the
porosity of 32 seeded cells, from the FP64 oracle (double arithmetic and storage).

    python tests/golden/make_adjoint_fd.py [--threads 8] [--scale 1]

--scale 2 halves every grid dimension and the step counts (a quick check of the script).
The GPU adjoint gradient is compared with these values in tests/test_gpu_adjoint.py.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from tests import scenes  # noqa: E402


def objective(cfg, rho, u, flags, alpha, threads):
    L = O.OracleLBM(cfg["Nx"], cfg["Ny"], cfg["Nz"], nu=cfg["nu"], storage="FP64", ext=O.EXT_VOLUME_FORCE | O.EXT_POROSITY,
                    fx=cfg["fx"], threads=threads)
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    L.alpha[:] = alpha.astype(np.float64)
    L.initialize()
    J, _ = L.run_objective(cfg["steps"], cfg["window_begin"], cfg["plane_x"])
    return J


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--scale", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "adjoint_channel_fd.json"))
    a = ap.parse_args()
    cfg = dict(scenes.CHANNEL_128)
    if a.scale > 1:
        for k in ("Nx", "Ny", "Nz", "steps", "window_begin", "checkpoint_interval"):
            cfg[k] //= a.scale
        cfg["plane_x"] = cfg["Nx"] - 1
    rho, u, flags, alpha = scenes.porous_channel(cfg["Nx"], cfg["Ny"], cfg["Nz"])
    cells = scenes.fd_cells(flags, alpha, cfg["fd_cells"], cfg["fd_seed"])
    eps = cfg["fd_eps"]
    t0 = time.time()
    J0 = objective(cfg, rho, u, flags, alpha, a.threads)
    fd = []
    for k, n in enumerate(cells):
        ap_, am_ = alpha.astype(np.float64).copy(), alpha.astype(np.float64).copy()
        ap_[n] += eps
        am_[n] -= eps
        Jp = objective(cfg, rho, u, flags, ap_, a.threads)
        Jm = objective(cfg, rho, u, flags, am_, a.threads)
        fd.append((Jp - Jm) / (2 * eps))
        print(f"cell {k + 1}/{len(cells)} n={n}: dJ/dalpha = {fd[-1]:.12e}  ({time.time() - t0:.0f} s)", flush=True)
    out = {"config": cfg, "J": J0, "cells": [int(c) for c in cells], "fd": fd, "eps": eps,
           "method": "central differences of the FP64 oracle objective, porosity perturbed in double"}
    json.dump(out, open(a.out, "w"), indent=1)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
