"""Writes tests/golden/tgv256_fp64.npz: the FP64 oracle's 256^3 Taylor-Green vortex.

D3Q19 SRT, nu = 1/6, u0 = 0.05, k = 2 pi / 256, fully periodic, the inputs of
tests.scenes.taylor_green (float32 rho/u, the same values the GPU starts from).  The FP64
oracle (same scheme and operation order as the reference's kernels, double storage and
arithmetic) runs 10000 steps; after 1000 and after 10000 steps the file holds:

* rho and u (3 components) at 4096 cells drawn with a fixed seed (`cells`),
* the kinetic energy sum over all cells of rho |u|^2 / 2 (`ke`), and the total mass
  sum over all cells of rho (`mass`), both in FP64.

The GPU test runs the same lattice in FP32, FP16S and FP16C storage and reports each
mode's error against these values.  Takes about an hour on 8 CPU threads:

    python tests/golden/make_tgv256_fp64.py [threads]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from tests import scenes  # noqa: E402

SIZE = 256
U0 = 0.05
SNAPSHOTS = (1000, 10000)
SAMPLES = 4096
SEED = 256


def sample_cells(N):
    return np.sort(np.random.default_rng(SEED).choice(N, SAMPLES, replace=False)).astype(np.int64)


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else min(8, os.cpu_count() or 1)
    n = SIZE
    N = n ** 3
    rho, u, flags = scenes.taylor_green(n, n, n, u0=U0)
    ref = O.OracleLBM(n, n, n, nu=1.0 / 6.0, storage="FP64", threads=threads)
    ref.rho[:] = rho
    ref.u[:] = u
    ref.flags[:] = flags
    ref.initialize()
    cells = sample_cells(N)
    out = {"cells": cells, "steps": np.array(SNAPSHOTS, np.int64), "size": np.int64(n), "u0": np.float64(U0),
           "w": np.float64(ref.w)}
    t0 = time.time()
    done = 0
    for s in SNAPSHOTS:
        while done < s:
            k = min(500, s - done)
            ref.step(k)
            done += k
            print(f"step {done}: {time.time() - t0:.0f} s", flush=True)
        ref.update_fields()
        r = np.asarray(ref.rho, np.float64)
        v = np.asarray(ref.u, np.float64).reshape(3, N)
        out[f"rho_{s}"] = r[cells]
        out[f"u_{s}"] = v[:, cells]
        out[f"ke_{s}"] = np.float64(0.5 * np.sum(r * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)))
        out[f"mass_{s}"] = np.float64(np.sum(r))
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tgv256_fp64.npz"), **out)
    print(f"wrote tgv256_fp64.npz after {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
