"""Python mirror of the reference's LBM host class (src/lbm.hpp:205-611, src/lbm.cpp:700-1016)
on top of the C ABI.

    lbm = LBM(256, 256, 256, nu=1/6)            # LBM(Nx,Ny,Nz,nu,fx,fy,fz)
    lbm.u.x[:] = ...; lbm.rho[:] = ...; lbm.flags[n] = TYPE_S
    lbm.run(1000)                               # initialize on first call, then steps
    lbm.u.read_from_device()                    # update_fields at the current t, copy back

Compile-time switches of the reference (src/defines.hpp) are constructor keywords here:
velocity_set (19/27), collision ("SRT"/"TRT"), storage ("FP32"/"FP16S"/"FP16C") and
extensions (names from src/defines.hpp plus "POROSITY").  Host buffers are numpy arrays
in the reference layout (u[d*N+n]); device buffers belong to the C ABI context.
"""
import numpy as np

from . import _abi
from ._abi import check

TYPE_S = 0x01
TYPE_E = 0x02
TYPE_T = 0x04
TYPE_F = 0x08
TYPE_I = 0x10
TYPE_G = 0x20
TYPE_X = 0x40
TYPE_Y = 0x80


class Field:
    """Host+device field, the single-domain Memory_Container (src/lbm.hpp:239-408)."""

    def __init__(self, lbm, name, dims, dtype, fill=0):
        self.lbm = lbm
        self.name = name
        self.dims = dims
        self.N = lbm.N
        self.host = np.full(lbm.N * dims, fill, dtype=dtype)
        self.device = lbm.ctx.alloc(self.host.nbytes)
        if dims > 1:
            self.x = self.host[0:self.N]
            self.y = self.host[self.N:2 * self.N]
            self.z = self.host[2 * self.N:3 * self.N] if dims > 2 else None

    def __getitem__(self, i):
        return self.host[i]

    def __setitem__(self, i, v):
        self.host[i] = v

    def __len__(self):
        return self.host.size

    def reset(self, value=0):
        self.host[:] = value

    def write_to_device(self):
        self.lbm.ctx.write(self.device, self.host)

    def read_from_device(self):
        """update_fields first if rho/u are stale (src/lbm.hpp:390-396), then copy back."""
        self._refresh()
        self.lbm.ctx.read(self.device, self.host)

    def _refresh(self):
        if self.name in ("rho", "u") and self.lbm.initialized:
            self.lbm._update_fields_if_stale()

    # Box transfers into/out of the same offsets of the host array, one component
    # (dimension >= 0) or all; Memory<T>::read_from_device_1d/2d/3d and
    # write_to_device_1d/2d/3d (src/opencl.hpp:532-607).
    def read_from_device_1d(self, x0, x1, dimension=-1, blocking=True):
        self._refresh()
        self.lbm.ctx.read_box(self.device, self.host, self.dims, self.N, 1, x0, x1, dimension=dimension, blocking=blocking)

    def write_to_device_1d(self, x0, x1, dimension=-1, blocking=True):
        self.lbm.ctx.write_box(self.device, self.host, self.dims, self.N, 1, x0, x1, dimension=dimension, blocking=blocking)

    def read_from_device_2d(self, x0, x1, y0, y1, Nx, Ny, dimension=-1, blocking=True):
        self._refresh()
        self.lbm.ctx.read_box(self.device, self.host, self.dims, Nx, Ny, x0, x1, y0, y1, dimension=dimension,
                              blocking=blocking)

    def write_to_device_2d(self, x0, x1, y0, y1, Nx, Ny, dimension=-1, blocking=True):
        self.lbm.ctx.write_box(self.device, self.host, self.dims, Nx, Ny, x0, x1, y0, y1, dimension=dimension,
                               blocking=blocking)

    def read_from_device_3d(self, x0, x1, y0, y1, z0, z1, Nx, Ny, Nz, dimension=-1, blocking=True):
        self._refresh()
        self.lbm.ctx.read_box(self.device, self.host, self.dims, Nx, Ny, x0, x1, y0, y1, z0, z1, dimension, blocking)

    def write_to_device_3d(self, x0, x1, y0, y1, z0, z1, Nx, Ny, Nz, dimension=-1, blocking=True):
        self.lbm.ctx.write_box(self.device, self.host, self.dims, Nx, Ny, x0, x1, y0, y1, z0, z1, dimension, blocking)


class LBM:
    def __init__(self, Nx, Ny, Nz, nu, fx=0.0, fy=0.0, fz=0.0, *, velocity_set=19, collision="SRT",
                 storage="FP32", extensions=(), device=0, w=None):
        ext = 0
        for name in extensions:
            ext |= _abi.EXTENSIONS[name]
        if nu == 0.0:
            raise ValueError("Viscosity cannot be 0.")  # src/lbm.cpp:802-803
        if nu < 0.0:
            raise ValueError("Viscosity cannot be negative.")
        if (fx != 0.0 or fy != 0.0 or fz != 0.0) and not ext & (_abi.FX_VOLUME_FORCE | _abi.FX_POROSITY):
            raise ValueError("Volume force is set, but VOLUME_FORCE is not enabled.")  # src/lbm.cpp:811-813
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.N = self.Nx * self.Ny * self.Nz
        self.nu = float(nu)
        self.w = _abi.relaxation_rate(nu) if w is None else float(w)
        self.fx, self.fy, self.fz = float(fx), float(fy), float(fz)
        self.velocity_set = int(velocity_set)
        self.storage = storage
        self.extensions = ext
        self.cfg = _abi.make_config(Nx, Ny, Nz, self.w, velocity_set, collision, storage, ext)
        self.ctx = _abi.Context(self.cfg, device)
        ddf_bytes = 4 if storage == "FP32" else 2
        self.fi = self.ctx.alloc(self.velocity_set * self.N * ddf_bytes)
        self.rho = Field(self, "rho", 1, np.float32, 1.0)
        self.u = Field(self, "u", 3, np.float32)
        self.flags = Field(self, "flags", 1, np.uint8)
        self.F = Field(self, "F", 3, np.float32) if ext & _abi.FX_FORCE_FIELD else None
        self.porosity = Field(self, "porosity", 1, np.float32) if ext & _abi.FX_POROSITY else None
        self.t = 0
        self.t_last_update_fields = None
        self.t_last_force_field = None
        self.initialized = False

    # -- helpers --------------------------------------------------------------------
    def _ptr(self, field):
        return None if field is None else field.device

    def get_t(self):
        return self.t

    def get_tau(self):
        return 3.0 * self.nu + 0.5

    def coordinates(self, n):
        return n % self.Nx, (n // self.Nx) % self.Ny, n // (self.Nx * self.Ny)

    def index(self, x, y, z):
        return x + (y + z * self.Ny) * self.Nx

    def position(self, x, y, z):
        return (x + 0.5 - 0.5 * self.Nx, y + 0.5 - 0.5 * self.Ny, z + 0.5 - 0.5 * self.Nz)

    def center(self):
        return (0.5 * self.Nx - 0.5, 0.5 * self.Ny - 0.5, 0.5 * self.Nz - 0.5)

    # -- simulation -------------------------------------------------------------------
    def initialize(self):
        """Write fields to the device, run initialize with t=1, reset t=0 (src/lbm.cpp:881-922)."""
        flags = self.flags.host
        if (flags & TYPE_BO_MASK == TYPE_E).any() and not self.extensions & _abi.FX_EQUILIBRIUM_BOUNDARIES:
            raise ValueError("Some cells are TYPE_E, but EQUILIBRIUM_BOUNDARIES is not enabled.")  # src/lbm.cpp:866-867
        self.rho.write_to_device()
        self.u.write_to_device()
        self.flags.write_to_device()
        if self.F is not None:
            self.F.write_to_device()
        if self.porosity is not None:
            self.porosity.write_to_device()
        check(self.ctx.lib.fx_lbm_initialize(self.ctx.h, self.fi, self.rho.device, self.u.device, self.flags.device))
        self.ctx.sync()
        self.t = 0
        self.t_last_update_fields = None
        self.t_last_force_field = None
        self.initialized = True

    def _step_args(self):
        return (self.ctx.h, self.fi, self.rho.device, self.u.device, self.flags.device)

    def run(self, steps=1, sync_each_step=False):
        """LBM::run (src/lbm.cpp:955-975).  Steps are queued back to back on the stream;
        sync_each_step=True reproduces the reference's per-step finish_queue (src/lbm.cpp:951)."""
        if not self.initialized:
            self.initialize()
        for _ in range(int(steps)):
            check(self.ctx.lib.fx_lbm_stream_collide(*self._step_args(), self.t, self.fx, self.fy, self.fz,
                                                     self._ptr(self.F), self._ptr(self.porosity)))
            if sync_each_step:
                self.ctx.sync()
            self.t += 1
        if self.extensions & _abi.FX_UPDATE_FIELDS:
            self.t_last_update_fields = self.t
        self.ctx.sync()

    def _update_fields_if_stale(self):
        if self.t_last_update_fields != self.t:
            check(self.ctx.lib.fx_lbm_update_fields(*self._step_args(), self.t, self.fx, self.fy, self.fz,
                                                    self._ptr(self.F), self._ptr(self.porosity)))
            self.t_last_update_fields = self.t

    def update_fields(self):
        """LBM::update_fields (src/lbm.cpp:977-980): refresh rho/u on the device."""
        self._update_fields_if_stale()
        self.ctx.sync()

    def update_moving_boundaries(self):
        """LBM::update_moving_boundaries (src/lbm.cpp:1019-1027): after wall velocities
        changed (u of TYPE_S cells written to the device), re-mark TYPE_MS."""
        if not self.extensions & _abi.FX_MOVING_BOUNDARIES:
            raise ValueError("MOVING_BOUNDARIES is not enabled")
        check(self.ctx.lib.fx_lbm_update_moving_boundaries(self.ctx.h, self.u.device, self.flags.device))
        self.ctx.sync()

    def reset(self):
        self.initialized = False

    def set_f(self, fx, fy, fz):
        self.fx, self.fy, self.fz = float(fx), float(fy), float(fz)

    # -- forces (FORCE_FIELD) ---------------------------------------------------------
    def update_force_field(self):
        if self.F is None:
            raise ValueError("FORCE_FIELD is not enabled")
        if self.t_last_force_field != self.t:
            check(self.ctx.lib.fx_lbm_update_force_field(self.ctx.h, self.fi, self.flags.device, self.t, self.F.device))
            self.t_last_force_field = self.t
        self.ctx.sync()

    def object_force(self, flag_marker=TYPE_S):
        import ctypes
        self.update_force_field()
        out = (ctypes.c_double * 3)()
        check(self.ctx.lib.fx_lbm_object_force(self.ctx.h, self.F.device, self.flags.device, flag_marker, out))
        return np.array(out[:])

    def object_torque(self, rotation_center, flag_marker=TYPE_S):
        import ctypes
        self.update_force_field()
        out = (ctypes.c_double * 3)()
        c = [float(v) for v in rotation_center]
        check(self.ctx.lib.fx_lbm_object_torque(self.ctx.h, self.F.device, self.flags.device, flag_marker, c[0], c[1], c[2], out))
        return np.array(out[:])

    def object_center_of_mass(self, flag_marker=TYPE_S):
        import ctypes
        out = (ctypes.c_double * 4)()
        check(self.ctx.lib.fx_lbm_object_center_of_mass(self.ctx.h, self.flags.device, flag_marker, out))
        return np.array(out[:3]) / max(out[3], 1.0)

    def mass(self):
        import ctypes
        out = ctypes.c_double()
        check(self.ctx.lib.fx_lbm_mass_sum(self.ctx.h, self.fi, self.flags.device, self.t, ctypes.byref(out)))
        return out.value

    def read_ddfs(self):
        """Raw DDF array (float32 or uint16 codes) in the slot layout fi[i*N+n]."""
        dt = np.float32 if self.storage == "FP32" else np.uint16
        host = np.empty(self.velocity_set * self.N, dtype=dt)
        self.ctx.read(self.fi, host)
        return host

    def write_ddfs(self, host):
        self.ctx.write(self.fi, host)

    def close(self):
        self.ctx.close()


TYPE_BO_MASK = 0x03
