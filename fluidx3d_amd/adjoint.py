"""Checkpointed adjoint of an LBM run (LBM::backward) through the C ABI.

Objective: J = mean over steps t in [window_begin, steps) of the x-momentum of the
populations loaded at step t, summed over the non-solid cells of the plane x = plane_x.
Gradient: dJ/dporosity (FX_POROSITY, per-cell Brinkman porosity).

    lbm = LBM(..., extensions=("VOLUME_FORCE", "POROSITY"))
    ... set lbm.rho/u/flags/porosity host fields ...
    adj = Adjoint(lbm, steps=2000, window_begin=1500, plane_x=lbm.Nx-1, checkpoint_interval=100)
    J = adj.forward()          # initialize + steps, DDF checkpoints every k steps
    grad = adj.backward()      # numpy float32[N]
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check


class Adjoint:
    def __init__(self, lbm, steps, window_begin, plane_x, checkpoint_interval=0):
        self.lbm = lbm
        self.cfg = _abi.AdjointConfig(int(steps), int(checkpoint_interval), int(window_begin), int(plane_x))
        L = lbm.ctx.lib
        h = ctypes.c_void_p()
        check(L.fx_adjoint_create(lbm.ctx.h, ctypes.byref(self.cfg), ctypes.byref(h)))
        self.h = h
        self.grad_device = lbm.ctx.alloc(4 * lbm.N)
        self.forward_ms = None
        self.backward_ms = None

    @staticmethod
    def memory_bytes(lbm, steps, window_begin, plane_x, checkpoint_interval=0):
        cfg = _abi.AdjointConfig(int(steps), int(checkpoint_interval), int(window_begin), int(plane_x))
        b = ctypes.c_uint64()
        check(lbm.ctx.lib.fx_adjoint_memory(lbm.ctx.h, ctypes.byref(cfg), ctypes.byref(b)))
        return b.value

    def _fields(self):
        lbm = self.lbm
        return (lbm.rho.device, lbm.u.device, lbm.flags.device)

    def forward(self):
        """Writes the host fields, runs initialize (t=1 convention) and `steps` steps from t=0
        with checkpoints; returns J.  The device fields then hold the state at t=steps."""
        lbm = self.lbm
        lbm.initialize()
        J = ctypes.c_double()
        ms = ctypes.c_float()
        check(lbm.ctx.lib.fx_lbm_forward(lbm.ctx.h, self.h, lbm.fi, *self._fields(), lbm.fx, lbm.fy, lbm.fz,
                                         lbm._ptr(lbm.F), lbm._ptr(lbm.porosity), ctypes.byref(J), ctypes.byref(ms)))
        lbm.t = int(self.cfg.steps)
        self.forward_ms = ms.value
        return J.value

    def backward_device(self, grad_ptr=None):
        """Adjoint sweep; writes dJ/dporosity to grad_ptr (device, N floats) or to the
        handle's own buffer.  Returns the device pointer."""
        lbm = self.lbm
        dst = self.grad_device if grad_ptr is None else grad_ptr
        ms = ctypes.c_float()
        check(lbm.ctx.lib.fx_lbm_backward(lbm.ctx.h, self.h, *self._fields(), lbm.fx, lbm.fy, lbm.fz,
                                          lbm._ptr(lbm.F), lbm._ptr(lbm.porosity), dst, ctypes.byref(ms)))
        self.backward_ms = ms.value
        return dst

    def backward(self):
        """Adjoint sweep; returns dJ/dporosity as numpy float32[N]."""
        self.backward_device()
        g = np.empty(self.lbm.N, dtype=np.float32)
        self.lbm.ctx.read(self.grad_device, g)
        return g

    def close(self):
        if self.h:
            self.lbm.ctx.lib.fx_adjoint_destroy(self.h)
            self.h = None
