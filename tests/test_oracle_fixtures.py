"""The oracle against its committed regression fixtures (tests/golden/lattice_fixtures.json,
made by tests/golden/make_lattice_fixtures.py): DDF hashes of the FP32-arithmetic oracle
after 1, 2, 3 and 10 steps and FP64 rho/u samples for every velocity set, collision,
storage and extension set on 8^3 and 16^3 lattices with a solid block and an equilibrium
face; the FP64 kinetic-energy curve of the 64^3 Taylor-Green vortex; the FP16C tables."""
import hashlib
import json
import os

import numpy as np
import pytest

from tests import scenes
from tests.golden import make_lattice_fixtures as M

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "lattice_fixtures.json")))


def ids(e):
    return f"{e['n']}^3-D3Q{e['q']}-{e['collision']}-{e['storage']}-{e['ext']}"


@pytest.mark.parametrize("entry", GOLD["lattices"], ids=ids)
def test_oracle_reproduces_lattice_fixture(oracle, entry):
    L = M.make_lattice(entry["n"], entry["q"], entry["collision"], entry["storage"], entry["ext"], entry["seed"])
    L.initialize()
    done = 0
    for s in GOLD["steps"]:
        L.step(s - done)
        done = s
        assert M.ddf_hash(L.fi) == entry["hash"][str(s)], f"step {s}"
    R = M.make_lattice(entry["n"], entry["q"], entry["collision"], entry["storage"], entry["ext"], entry["seed"], storage_real="FP64")
    R.initialize()
    R.step(GOLD["steps"][-1])
    R.update_fields()
    N = entry["n"] ** 3
    for c, r, u in zip(entry["fp64_cells"], entry["fp64_rho"], entry["fp64_u"]):
        assert R.rho[c] == r and [R.u[d * N + c] for d in range(3)] == u


def test_oracle_reproduces_tgv_kinetic_energy_curve(oracle):
    """First 200 steps recomputed; the whole 1000-step curve decays at the viscous rate
    KE ~ exp(-6 nu k^2 t) (k = 2 pi/64, nu = 1/6) within 2 %."""
    ke = M.tgv_ke(steps=200, every=10)
    assert ke == GOLD["tgv64_ke_every_10"][:len(ke)]
    curve = np.array(GOLD["tgv64_ke_every_10"])
    t = 10.0 * np.arange(curve.size)
    rate = -np.polyfit(t[20:], np.log(curve[20:]), 1)[0]
    k = 2.0 * np.pi / 64.0
    assert abs(rate / (6.0 * (1.0 / 6.0) * k * k) - 1.0) < 0.02


def test_oracle_reproduces_fp16c_tables(oracle):
    assert M.fp16c_tables() == GOLD["fp16c"]
