"""ctypes/numpy front end of the CPU oracle (oracle/lbm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker.  The product package never imports it.

`OracleLBM` reproduces the reference's host sequencing around the kernels
(src/lbm.cpp:881-975): `initialize` runs the initialize kernel with the odd-step
convention t=1 and then resets t to 0; every `step` launches stream_collide with the
current t and increments it; `update_fields` recomputes rho/u from the DDFs at the
current t (src/lbm.cpp:184-191).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STORAGE = {"FP32": 0, "FP16S": 1, "FP16C": 2, "FP64": 3}
EXT_VOLUME_FORCE = 1
EXT_FORCE_FIELD = 2
EXT_EQUILIBRIUM_BOUNDARIES = 4
EXT_UPDATE_FIELDS = 8
EXT_POROSITY = 16
EXT_PRECISE_DENSITY = 32
EXT_MOVING_BOUNDARIES = 64
EXT_SUBGRID = 128

TYPE_S = 0x01
TYPE_E = 0x02
TYPE_MS = 0x03
TYPE_X = 0x40


class Cfg(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_uint32), ("Ny", ctypes.c_uint32), ("Nz", ctypes.c_uint32),
                ("q", ctypes.c_uint32), ("trt", ctypes.c_uint32), ("storage", ctypes.c_uint32),
                ("ext", ctypes.c_uint32), ("threads", ctypes.c_uint32), ("w", ctypes.c_double)]


def lib():
    """Load oracle/liboracle.so (built by `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        vp, u8p, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_double)
        cp = ctypes.POINTER(Cfg)
        L.fxo_relaxation_rate.restype = ctypes.c_float
        L.fxo_relaxation_rate.argtypes = [ctypes.c_float]
        L.fxo_format_float.argtypes = [ctypes.c_float, ctypes.c_char_p, ctypes.c_int]
        for name in ("fxo_fp16s_encode", "fxo_fp16c_encode", "fxo_f16_rne"):
            getattr(L, name).restype = ctypes.c_uint16
            getattr(L, name).argtypes = [ctypes.c_float]
        for name in ("fxo_fp16s_decode", "fxo_fp16c_decode", "fxo_f16_to_f32"):
            getattr(L, name).restype = ctypes.c_float
            getattr(L, name).argtypes = [ctypes.c_uint16]
        L.fxo_f16_selfcheck.restype = ctypes.c_uint64
        L.fxo_f16_selfcheck.argtypes = [ctypes.c_uint32]
        L.fxo_encode_array.argtypes = [ctypes.c_int, vp, vp, ctypes.c_uint64]
        L.fxo_decode_array.argtypes = [ctypes.c_int, vp, vp, ctypes.c_uint64]
        L.fxo_initialize.argtypes = [cp, vp, vp, vp, u8p]
        L.fxo_update_moving_boundaries.argtypes = [cp, vp, u8p]
        step_args = [cp, vp, vp, vp, u8p, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp, vp]
        L.fxo_stream_collide.argtypes = step_args
        L.fxo_update_fields.argtypes = step_args
        L.fxo_update_force_field.argtypes = [cp, vp, u8p, ctypes.c_uint64, vp]
        L.fxo_load_populations.argtypes = [cp, vp, ctypes.c_uint64, vp]
        L.fxo_mass_sum.argtypes = [cp, vp, u8p, ctypes.c_uint64, dp]
        L.fxo_object_force.argtypes = [cp, vp, u8p, ctypes.c_uint8, dp]
        L.fxo_object_torque.argtypes = [cp, vp, u8p, ctypes.c_uint8, ctypes.c_float, ctypes.c_float, ctypes.c_float, dp]
        L.fxo_object_center_of_mass.argtypes = [cp, u8p, ctypes.c_uint8, dp]
        L.fxo_adjoint_step.argtypes = [cp, vp, vp, u8p, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, vp, vp, ctypes.c_uint32, ctypes.c_double, vp, vp, vp]
        L.fxo_objective_plane.argtypes = [cp, vp, u8p, ctypes.c_uint64, ctypes.c_uint32, dp]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def relaxation_rate(nu):
    """w = 1/tau as the reference's kernels see it (printed and re-parsed, src/lbm.cpp:367)."""
    return float(lib().fxo_relaxation_rate(ctypes.c_float(nu)))


def format_float(x):
    buf = ctypes.create_string_buffer(64)
    n = lib().fxo_format_float(ctypes.c_float(x), buf, 64)
    assert n > 0
    return buf.value.decode()


def encode(storage, x):
    """float32 array -> uint16 codes (storage 'FP16S', 'FP16C' or 'F16' = plain IEEE half RNE)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    code = {"FP16S": 1, "FP16C": 2, "F16": 9}[storage]
    lib().fxo_encode_array(code, _ptr(x), _ptr(out), x.size)
    return out


def decode(storage, h):
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    code = {"FP16S": 1, "FP16C": 2, "F16": 9}[storage]
    lib().fxo_decode_array(code, _ptr(h), _ptr(out), h.size)
    return out


class OracleLBM:
    """One lattice on the CPU with the reference's memory layout and host sequencing."""

    def __init__(self, Nx, Ny, Nz, nu=None, w=None, q=19, trt=False, storage="FP32", ext=0,
                 fx=0.0, fy=0.0, fz=0.0, threads=None):
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.N = self.Nx * self.Ny * self.Nz
        self.q = int(q)
        self.storage = storage
        if w is None:
            w = relaxation_rate(nu)
        self.w = float(w)
        self.ext = int(ext)
        self.fx, self.fy, self.fz = fx, fy, fz
        self.threads = int(threads or min(16, os.cpu_count() or 1))
        self.cfg = Cfg(self.Nx, self.Ny, self.Nz, self.q, 1 if trt else 0, STORAGE[storage], self.ext,
                       self.threads, self.w)
        real = np.float64 if storage == "FP64" else np.float32
        self.real = real
        ddf = {"FP32": np.float32, "FP16S": np.uint16, "FP16C": np.uint16, "FP64": np.float64}[storage]
        self.fi = np.zeros(self.q * self.N, dtype=ddf)
        self.rho = np.ones(self.N, dtype=real)
        self.u = np.zeros(3 * self.N, dtype=real)
        self.flags = np.zeros(self.N, dtype=np.uint8)
        self.F = np.zeros(3 * self.N, dtype=real) if self.ext & EXT_FORCE_FIELD else None
        self.alpha = np.zeros(self.N, dtype=real) if self.ext & EXT_POROSITY else None
        self.t = 0

    @property
    def flags_ptr(self):
        return self.flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))

    def initialize(self):
        rc = lib().fxo_initialize(ctypes.byref(self.cfg), _ptr(self.fi), _ptr(self.rho), _ptr(self.u), self.flags_ptr)
        assert rc == 0, rc
        self.t = 0

    def update_moving_boundaries(self):
        """TYPE_MS marks of the cells next to moving solid cells (src/kernel.cpp:1433-1449)."""
        assert lib().fxo_update_moving_boundaries(ctypes.byref(self.cfg), _ptr(self.u), self.flags_ptr) == 0

    def step(self, steps=1):
        for _ in range(steps):
            rc = lib().fxo_stream_collide(ctypes.byref(self.cfg), _ptr(self.fi), _ptr(self.rho), _ptr(self.u),
                                          self.flags_ptr, self.t, self.fx, self.fy, self.fz, _ptr(self.F),
                                          _ptr(self.alpha))
            assert rc == 0, rc
            self.t += 1

    def update_fields(self):
        rc = lib().fxo_update_fields(ctypes.byref(self.cfg), _ptr(self.fi), _ptr(self.rho), _ptr(self.u),
                                     self.flags_ptr, self.t, self.fx, self.fy, self.fz, _ptr(self.F), _ptr(self.alpha))
        assert rc == 0, rc

    def update_force_field(self):
        if self.F is None:
            self.F = np.zeros(3 * self.N, dtype=self.real)
        rc = lib().fxo_update_force_field(ctypes.byref(self.cfg), _ptr(self.fi), self.flags_ptr, self.t, _ptr(self.F))
        assert rc == 0, rc

    def object_force(self, marker=TYPE_S):
        self.update_force_field()
        out = (ctypes.c_double * 3)()
        assert lib().fxo_object_force(ctypes.byref(self.cfg), _ptr(self.F), self.flags_ptr, marker, out) == 0
        return np.array(out[:])

    def object_torque(self, center, marker=TYPE_S):
        self.update_force_field()
        out = (ctypes.c_double * 3)()
        assert lib().fxo_object_torque(ctypes.byref(self.cfg), _ptr(self.F), self.flags_ptr, marker,
                                       center[0], center[1], center[2], out) == 0
        return np.array(out[:])

    def object_center_of_mass(self, marker=TYPE_S):
        out = (ctypes.c_double * 4)()
        assert lib().fxo_object_center_of_mass(ctypes.byref(self.cfg), self.flags_ptr, marker, out) == 0
        return np.array(out[:3]) / out[3], int(out[3])

    def mass(self):
        out = ctypes.c_double()
        assert lib().fxo_mass_sum(ctypes.byref(self.cfg), _ptr(self.fi), self.flags_ptr, self.t, ctypes.byref(out)) == 0
        return out.value

    def populations(self):
        """Pre-collision populations f_i(n) the next step (time t) will load, shape (q, N)."""
        out = np.empty(self.q * self.N, dtype=self.real)
        assert lib().fxo_load_populations(ctypes.byref(self.cfg), _ptr(self.fi), self.t, _ptr(out)) == 0
        return out.reshape(self.q, self.N)

    # -- adjoint (FP64 storage only) --------------------------------------------------
    def objective_plane(self, plane_x):
        """x-momentum of the populations loaded at the current step, summed over the
        non-solid cells of the plane x = plane_x."""
        out = ctypes.c_double()
        assert lib().fxo_objective_plane(ctypes.byref(self.cfg), _ptr(self.fi), self.flags_ptr, self.t, plane_x, ctypes.byref(out)) == 0
        return out.value

    def run_objective(self, steps, window_begin, plane_x, keep_states=False):
        """Step `steps` times from t=0; J = mean over steps t >= window_begin of the plane
        momentum loaded at step t.  Optionally keeps the DDF array before every step."""
        states = []
        acc = 0.0
        for t in range(steps):
            if keep_states:
                states.append(self.fi.copy())
            if t >= window_begin:
                acc += self.objective_plane(plane_x)
            self.step(1)
        return acc / (steps - window_begin), states

    def adjoint_gradient(self, steps, window_begin, plane_x):
        """dJ/dporosity by the FP64 adjoint over the stored forward states (small grids).
        Call after initialize(); returns (J, grad)."""
        assert self.storage == "FP64"
        J, states = self.run_objective(steps, window_begin, plane_x, keep_states=True)
        lam = np.zeros_like(self.fi)
        grad = np.zeros(self.N, dtype=np.float64)
        wobj = 1.0 / (steps - window_begin)
        for t in range(steps - 1, -1, -1):
            rc = lib().fxo_adjoint_step(ctypes.byref(self.cfg), _ptr(states[t]), _ptr(lam), self.flags_ptr, t,
                                        self.fx, self.fy, self.fz, _ptr(self.F), _ptr(self.alpha), plane_x,
                                        wobj if t >= window_begin else 0.0, _ptr(grad), _ptr(self.rho), _ptr(self.u))
            assert rc == 0, rc
        return J, grad

    def ddf_values(self):
        """DDF array decoded to float (shifted populations, slot layout as stored)."""
        if self.storage in ("FP32", "FP64"):
            return self.fi.copy()
        return decode(self.storage, self.fi)

    # convenience views
    def coords(self):
        n = np.arange(self.N)
        return n % self.Nx, (n // self.Nx) % self.Ny, n // (self.Nx * self.Ny)
