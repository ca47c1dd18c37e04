"""ctypes binding of libfx_lbm.so (C ABI declared in include/fx_lbm.h).

The library is loaded from fluidx3d_amd/lib/.  There is no fallback: if the HIP library
is missing or no GPU is present, calls raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfx_lbm.so")

FX_OK = 0
FX_ERR_ARG, FX_ERR_HIP, FX_ERR_CONFIG, FX_ERR_MEMORY, FX_ERR_STATE = -1, -2, -3, -4, -5
FX_D3Q19, FX_D3Q27 = 19, 27
FX_SRT, FX_TRT = 0, 1
FX_FP32, FX_FP16S, FX_FP16C = 0, 1, 2
FX_VOLUME_FORCE = 1
FX_FORCE_FIELD = 2
FX_EQUILIBRIUM_BOUNDARIES = 4
FX_UPDATE_FIELDS = 8
FX_POROSITY = 16
FX_PRECISE_DENSITY = 32
FX_MOVING_BOUNDARIES = 64
FX_SUBGRID = 128
FX_FIELD_FI, FX_FIELD_RHO_U_FLAGS, FX_FIELD_FLAGS, FX_FIELD_F = 0, 1, 2, 3

STORAGE = {"FP32": FX_FP32, "FP16S": FX_FP16S, "FP16C": FX_FP16C}
COLLISION = {"SRT": FX_SRT, "TRT": FX_TRT}
EXTENSIONS = {"VOLUME_FORCE": FX_VOLUME_FORCE, "FORCE_FIELD": FX_FORCE_FIELD,
              "EQUILIBRIUM_BOUNDARIES": FX_EQUILIBRIUM_BOUNDARIES, "UPDATE_FIELDS": FX_UPDATE_FIELDS,
              "POROSITY": FX_POROSITY, "PRECISE_DENSITY": FX_PRECISE_DENSITY,
              "MOVING_BOUNDARIES": FX_MOVING_BOUNDARIES, "SUBGRID": FX_SUBGRID}

# every symbol include/fx_lbm.h declares
EXPORTS = [
    "fx_last_error", "fx_device_count", "fx_device_info_get", "fx_relaxation_rate", "fx_memory_bytes_per_cell",
    "fx_bandwidth_bytes_per_cell", "fx_ctx_create", "fx_ctx_destroy", "fx_ctx_config", "fx_sync", "fx_buf_alloc",
    "fx_buf_free", "fx_buf_write", "fx_buf_read", "fx_buf_fill", "fx_buf_copy", "fx_lbm_initialize",
    "fx_lbm_stream_collide", "fx_lbm_run", "fx_lbm_update_fields", "fx_lbm_update_force_field",
    "fx_lbm_reset_force_field", "fx_lbm_object_force", "fx_lbm_object_torque", "fx_lbm_object_center_of_mass",
    "fx_lbm_mass_sum", "fx_codec_roundtrip", "fx_bench_copy", "fx_adjoint_memory", "fx_adjoint_create",
    "fx_adjoint_destroy", "fx_lbm_forward", "fx_lbm_backward", "fx_lbm_stream_collide_record", "fx_lbm_adjoint_step",
    "fx_lbm_objective_plane", "fx_transfer_bytes", "fx_lbm_transfer_extract", "fx_lbm_transfer_insert",
    "fx_lbm_update_moving_boundaries", "fx_buf_read_box", "fx_buf_write_box", "fx_lbm_mass_sum_noalloc",
]


class Config(ctypes.Structure):
    _fields_ = [("Nx", ctypes.c_uint32), ("Ny", ctypes.c_uint32), ("Nz", ctypes.c_uint32),
                ("velocity_set", ctypes.c_uint32), ("collision", ctypes.c_uint32), ("storage", ctypes.c_uint32),
                ("extensions", ctypes.c_uint32), ("w", ctypes.c_float)]


class AdjointConfig(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("checkpoint_interval", ctypes.c_uint64), ("window_begin", ctypes.c_uint64),
                ("plane_x", ctypes.c_uint32)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 64), ("memory_bytes", ctypes.c_uint64),
                ("memory_free", ctypes.c_uint64), ("compute_units", ctypes.c_uint32), ("clock_mhz", ctypes.c_uint32)]


class FxError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"fx error {code}: {message}")
        self.code = code


_lib = None


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or `make -C fluidx3d_amd/csrc`")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, u32, f32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_float, ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    L.fx_last_error.restype = ctypes.c_char_p
    L.fx_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.fx_device_info_get.argtypes = [i32, ctypes.POINTER(DeviceInfo)]
    L.fx_relaxation_rate.restype = f32
    L.fx_relaxation_rate.argtypes = [f32]
    for name in ("fx_memory_bytes_per_cell", "fx_bandwidth_bytes_per_cell"):
        getattr(L, name).restype = u64
        getattr(L, name).argtypes = [ctypes.POINTER(Config)]
    L.fx_ctx_create.argtypes = [i32, ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.fx_ctx_destroy.argtypes = [vp]
    L.fx_ctx_config.argtypes = [vp, ctypes.POINTER(Config)]
    L.fx_sync.argtypes = [vp]
    L.fx_buf_alloc.argtypes = [vp, u64, ctypes.POINTER(vp)]
    L.fx_buf_free.argtypes = [vp, vp]
    L.fx_buf_write.argtypes = [vp, vp, u64, u64, vp, i32]
    L.fx_buf_read.argtypes = [vp, vp, u64, u64, vp, i32]
    L.fx_buf_fill.argtypes = [vp, vp, u64, ctypes.c_uint8]
    L.fx_buf_copy.argtypes = [vp, vp, vp, u64]
    box = [vp, vp, vp, u32, u64, u32, u64, u64, u64, u64, u64, u64, u64, u64, i32, i32]
    L.fx_buf_read_box.argtypes = box
    L.fx_buf_write_box.argtypes = box
    L.fx_lbm_initialize.argtypes = [vp, vp, vp, vp, vp]
    L.fx_lbm_update_moving_boundaries.argtypes = [vp, vp, vp]
    step = [vp, vp, vp, vp, vp, u64, f32, f32, f32, vp, vp]
    L.fx_lbm_stream_collide.argtypes = step
    L.fx_lbm_update_fields.argtypes = step
    L.fx_lbm_run.argtypes = [vp, vp, vp, vp, vp, u64, u64, f32, f32, f32, vp, vp, ctypes.POINTER(f32)]
    L.fx_lbm_update_force_field.argtypes = [vp, vp, vp, u64, vp]
    L.fx_lbm_reset_force_field.argtypes = [vp, vp]
    L.fx_lbm_object_force.argtypes = [vp, vp, vp, ctypes.c_uint8, dp]
    L.fx_lbm_object_torque.argtypes = [vp, vp, vp, ctypes.c_uint8, f32, f32, f32, dp]
    L.fx_lbm_object_center_of_mass.argtypes = [vp, vp, ctypes.c_uint8, dp]
    L.fx_lbm_mass_sum.argtypes = [vp, vp, vp, u64, dp]
    L.fx_lbm_mass_sum_noalloc.argtypes = [vp, vp, vp, u64, dp]
    L.fx_codec_roundtrip.argtypes = [vp, u32, vp, vp, vp, u64]
    acp = ctypes.POINTER(AdjointConfig)
    fp = ctypes.POINTER(f32)
    L.fx_adjoint_memory.argtypes = [vp, acp, ctypes.POINTER(u64)]
    L.fx_adjoint_create.argtypes = [vp, acp, ctypes.POINTER(vp)]
    L.fx_adjoint_destroy.argtypes = [vp]
    L.fx_lbm_forward.argtypes = [vp, vp, vp, vp, vp, vp, f32, f32, f32, vp, vp, dp, fp]
    L.fx_lbm_backward.argtypes = [vp, vp, vp, vp, vp, f32, f32, f32, vp, vp, vp, fp]
    L.fx_lbm_stream_collide_record.argtypes = [vp, vp, vp, vp, vp, u64, f32, f32, f32, vp, vp, vp]
    L.fx_lbm_adjoint_step.argtypes = [vp, vp, vp, vp, vp, u64, f32, f32, f32, vp, vp, vp, u32, f32, vp]
    L.fx_lbm_objective_plane.argtypes = [vp, vp, vp, u64, u32, dp]
    L.fx_bench_copy.argtypes = [vp, vp, vp, u64, u32, ctypes.POINTER(f32)]
    L.fx_transfer_bytes.restype = u64
    L.fx_transfer_bytes.argtypes = [ctypes.POINTER(Config), u32, u32]
    L.fx_lbm_transfer_extract.argtypes = [vp, u32, u32, u64, vp, vp, vp, vp, vp, vp, vp]
    L.fx_lbm_transfer_insert.argtypes = [vp, u32, u32, u64, vp, vp, vp, vp, vp, vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != FX_OK:
        raise FxError(rc, load().fx_last_error().decode(errors="replace"))
    return rc


def relaxation_rate(nu):
    return float(load().fx_relaxation_rate(ctypes.c_float(nu)))


def device_count():
    c = ctypes.c_int(0)
    check(load().fx_device_count(ctypes.byref(c)))
    return c.value


def device_info(device=0):
    info = DeviceInfo()
    check(load().fx_device_info_get(device, ctypes.byref(info)))
    return {"name": info.name.decode(), "arch": info.arch.decode(), "memory_bytes": info.memory_bytes,
            "memory_free": info.memory_free, "compute_units": info.compute_units, "clock_mhz": info.clock_mhz}


def make_config(Nx, Ny, Nz, w, velocity_set=19, collision="SRT", storage="FP32", extensions=0):
    if not isinstance(extensions, int):
        e = 0
        for name in extensions:
            e |= EXTENSIONS[name]
        extensions = e
    return Config(int(Nx), int(Ny), int(Nz), int(velocity_set), COLLISION.get(collision, collision),
                  STORAGE.get(storage, storage), int(extensions), float(w))


def bandwidth_bytes_per_cell(cfg):
    return int(load().fx_bandwidth_bytes_per_cell(ctypes.byref(cfg)))


def memory_bytes_per_cell(cfg):
    return int(load().fx_memory_bytes_per_cell(ctypes.byref(cfg)))


class Context:
    """Owns an fx_ctx (device, HIP stream, selected kernel variant)."""

    def __init__(self, cfg, device=0):
        self.lib = load()
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(self.lib.fx_ctx_create(device, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.N = cfg.Nx * cfg.Ny * cfg.Nz

    def close(self):
        if self.h:
            self.lib.fx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        check(self.lib.fx_buf_alloc(self.h, int(nbytes), ctypes.byref(p)))
        return p.value

    def free(self, p):
        check(self.lib.fx_buf_free(self.h, p))

    def write(self, dst, host_array, offset=0, blocking=True):
        check(self.lib.fx_buf_write(self.h, dst, offset, host_array.nbytes, host_array.ctypes.data, 1 if blocking else 0))

    def read(self, src, host_array, offset=0, blocking=True):
        check(self.lib.fx_buf_read(self.h, src, offset, host_array.nbytes, host_array.ctypes.data, 1 if blocking else 0))

    def read_box(self, src, host_array, dims, Nx, Ny, x0, x1, y0=0, y1=1, z0=0, z1=1, dimension=-1, blocking=True):
        """Box [x0,x1) x [y0,y1) x [z0,z1) of component `dimension` (all if < 0) of a field of
        `dims` components into the same offsets of host_array (fx_buf_read_box)."""
        N = host_array.size // dims
        check(self.lib.fx_buf_read_box(self.h, src, host_array.ctypes.data, host_array.itemsize, N, dims, Nx, Ny, x0, x1,
                                       y0, y1, z0, z1, dimension, 1 if blocking else 0))

    def write_box(self, dst, host_array, dims, Nx, Ny, x0, x1, y0=0, y1=1, z0=0, z1=1, dimension=-1, blocking=True):
        N = host_array.size // dims
        check(self.lib.fx_buf_write_box(self.h, dst, host_array.ctypes.data, host_array.itemsize, N, dims, Nx, Ny, x0, x1,
                                        y0, y1, z0, z1, dimension, 1 if blocking else 0))

    def fill(self, dst, nbytes, value=0):
        check(self.lib.fx_buf_fill(self.h, dst, int(nbytes), value))

    def copy(self, dst, src, nbytes):
        check(self.lib.fx_buf_copy(self.h, dst, src, int(nbytes)))

    def sync(self):
        check(self.lib.fx_sync(self.h))
