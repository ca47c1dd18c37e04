"""Writes tests/golden/lattice_fixtures.json: regression fixtures of the CPU oracle.

* 8^3 (and 16^3) lattices with a TYPE_S block and, with EQUILIBRIUM_BOUNDARIES, a TYPE_E
  face at x=0: D3Q19/D3Q27 x SRT/TRT x FP32/FP16S/FP16C x {plain, E+VOLUME_FORCE+
  FORCE_FIELD}.  For the oracle in the reference's FP32 arithmetic: SHA-256 of the DDF
  array after 1, 2, 3 and 10 steps (stored bits; -0 counted as +0).  For the FP64 oracle:
  rho and u of 8 sample cells after 10 steps.
* the 64^3 Taylor-Green vortex (nu = 1/6, u0 = 0.05, 1000 steps): kinetic energy every 10
  steps from the FP64 oracle.
* FP16C: SHA-256 of the decoded values of all 65 536 codes and of the codes of a fixed
  sample of floats.

    python tests/golden/make_lattice_fixtures.py

The values come from this repository's oracle (the reference has no test vectors); the
tests check the oracle and the GPU against them.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from tests import scenes  # noqa: E402

EXT_SETS = {"plain": 0, "forces": O.EXT_EQUILIBRIUM_BOUNDARIES | O.EXT_VOLUME_FORCE | O.EXT_FORCE_FIELD}
FORCE = (2e-5, -1e-5, 5e-6)
NU = 0.06
STEPS = (1, 2, 3, 10)
SAMPLES = 8


def scene(n, ext, seed):
    """Random rho/u, a 2^3 solid block in the middle, TYPE_E face at x=0 with forces."""
    rho, u, flags = scenes.random_scene(n, n, n, seed=seed, solid_fraction=0.0, equilibrium_faces=False)
    x, y, z = scenes.coords(n, n, n)
    c = n // 2
    flags[(x >= c - 1) & (x <= c) & (y >= c - 1) & (y <= c) & (z >= c - 1) & (z <= c)] = scenes.TYPE_S
    if ext & O.EXT_EQUILIBRIUM_BOUNDARIES:
        flags[x == 0] = scenes.TYPE_E
    F = (1e-4 * np.random.default_rng(seed + 1).uniform(-1, 1, 3 * n ** 3)).astype(np.float32)
    return rho, u, flags, F


def ddf_hash(fi):
    a = np.array(fi, copy=True)
    if a.dtype == np.float32:
        a = a + np.float32(0.0)  # -0 -> +0
    else:
        a[a == 0x8000] = 0
    return hashlib.sha256(a.tobytes()).hexdigest()


def make_lattice(n, q, collision, storage, ext_name, seed, storage_real=None):
    ext = EXT_SETS[ext_name]
    rho, u, flags, F = scene(n, ext, seed)
    f = FORCE if ext & O.EXT_VOLUME_FORCE else (0.0, 0.0, 0.0)
    L = O.OracleLBM(n, n, n, nu=NU, q=q, trt=collision == "TRT", storage=storage_real or storage, ext=ext,
                    fx=np.float32(f[0]), fy=np.float32(f[1]), fz=np.float32(f[2]), threads=4)
    L.rho[:] = rho
    L.u[:] = u
    L.flags[:] = flags
    if L.F is not None:
        L.F[:] = F
    return L


def lattice_entry(n, q, collision, storage, ext_name, seed):
    L = make_lattice(n, q, collision, storage, ext_name, seed)
    L.initialize()
    hashes = {}
    done = 0
    for s in STEPS:
        L.step(s - done)
        done = s
        hashes[str(s)] = ddf_hash(L.fi)
    R = make_lattice(n, q, collision, storage, ext_name, seed, storage_real="FP64")
    R.initialize()
    R.step(STEPS[-1])
    R.update_fields()
    cells = np.random.default_rng(seed).choice(n ** 3, SAMPLES, replace=False)
    N = n ** 3
    return {"n": n, "q": q, "collision": collision, "storage": storage, "ext": ext_name, "seed": seed,
            "hash": hashes, "fp64_cells": [int(c) for c in cells],
            "fp64_rho": [float(R.rho[c]) for c in cells],
            "fp64_u": [[float(R.u[d * N + c]) for d in range(3)] for c in cells]}


def tgv_ke(steps=1000, every=10):
    n = 64
    rho, u, flags = scenes.taylor_green(n, n, n, u0=0.05)
    L = O.OracleLBM(n, n, n, nu=1.0 / 6.0, storage="FP64")
    L.rho[:] = rho
    L.u[:] = u
    L.initialize()
    ke = []
    for t in range(0, steps + 1, every):
        if t:
            L.step(every)
        L.update_fields()
        ke.append(float(scenes.kinetic_energy(L.rho, L.u, n ** 3)))
    return ke


def fp16c_tables():
    codes = np.arange(65536, dtype=np.uint16)
    dec = O.decode("FP16C", codes).astype(np.float32)
    x = np.random.default_rng(5).standard_normal(100000).astype(np.float32) * np.float32(0.01)
    enc = O.encode("FP16C", x).astype(np.uint16)
    return {"decode_sha256": hashlib.sha256(dec.tobytes()).hexdigest(), "encode_sample_seed": 5,
            "encode_sample_sha256": hashlib.sha256(enc.tobytes()).hexdigest()}


def main():
    out = {"nu": NU, "force": FORCE, "steps": list(STEPS), "lattices": []}
    seed = 100
    for q in (19, 27):
        for collision in ("SRT", "TRT"):
            for storage in ("FP32", "FP16S", "FP16C"):
                for ext_name in EXT_SETS:
                    out["lattices"].append(lattice_entry(8, q, collision, storage, ext_name, seed))
                    seed += 1
    for ext_name in EXT_SETS:
        out["lattices"].append(lattice_entry(16, 19, "SRT", "FP32", ext_name, seed))
        seed += 1
    out["tgv64_ke_every_10"] = tgv_ke()
    out["fp16c"] = fp16c_tables()
    path = os.path.join(ROOT, "tests", "golden", "lattice_fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
