"""Gradient-check harness of the torch.autograd wrapper (run as its own process:
`python tests/torch_gradcheck.py`; torch must start its GPU runtime before libfx_lbm.so is
loaded).  A porous channel 32x16x12 is run for 300 steps with the objective averaged over
the last 100; the autograd gradient is compared with the C ABI's fx_lbm_backward called
directly and with FP64 central differences of the oracle on 6 seeded cells.  Prints one
JSON line."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import fluidx3d_amd.autograd as fxa  # noqa: E402  (first: starts torch's GPU runtime)
import numpy as np  # noqa: E402
import torch  # noqa: E402
import fluidx3d_amd as fx  # noqa: E402
from fluidx3d_amd.adjoint import Adjoint  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests import scenes  # noqa: E402

EXT = ("VOLUME_FORCE", "POROSITY", "PRECISE_DENSITY")


def make_lbm(dims, rho, u, flags, fxv):
    lbm = fx.LBM(*dims, 1.0 / 6.0, fxv, 0.0, 0.0, extensions=EXT)
    lbm.rho.host[:] = rho; lbm.u.host[:] = u; lbm.flags.host[:] = flags
    return lbm


def main():
    dims = (32, 16, 12)
    steps, window, plane, fxv = 300, 200, dims[0] - 1, 1e-5
    rho, u, flags, alpha = scenes.porous_channel(*dims)
    lbm = make_lbm(dims, rho, u, flags, fxv)
    adj = Adjoint(lbm, steps, window, plane, 30)
    p = torch.tensor(alpha, device="cuda", requires_grad=True)
    J = fxa.OutletFlux.apply(p, adj)
    J.backward()
    g_autograd = p.grad.detach().cpu().numpy()
    lbm.porosity.host[:] = alpha
    adj.forward()
    g_abi = adj.backward()
    adj.close(); lbm.close()
    cells = scenes.fd_cells(flags, alpha, 6, 3)
    eps, fd = 1e-3, []
    for n in cells:
        Js = []
        for s in (+1, -1):
            a = alpha.astype(np.float64).copy()
            a[n] += s * eps
            L = O.OracleLBM(*dims, nu=1.0 / 6.0, storage="FP64", ext=O.EXT_VOLUME_FORCE | O.EXT_POROSITY | O.EXT_PRECISE_DENSITY, fx=float(np.float32(fxv)))
            L.rho[:] = rho; L.u[:] = u; L.flags[:] = flags; L.alpha[:] = a
            L.initialize()
            Js.append(L.run_objective(steps, window, plane)[0])
        fd.append((Js[0] - Js[1]) / (2 * eps))
    fd = np.array(fd)
    rel = np.abs(g_autograd[cells] - fd) / np.abs(fd)
    print(json.dumps({"J": float(J.item()), "autograd_equals_abi": bool(np.array_equal(g_autograd, g_abi)),
                      "fd_cells": [int(c) for c in cells], "fd": fd.tolist(), "autograd": g_autograd[cells].astype(float).tolist(),
                      "fd_rel_err_max": float(rel.max())}))


if __name__ == "__main__":
    main()
