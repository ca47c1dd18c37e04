/*
 * fx_lbm.h - C ABI of the MI355X-native lattice Boltzmann core (libfx_lbm.so).
 *
 * This layer replaces FluidX3D's OpenCL-Wrapper device/buffer/kernel layer
 * (src/opencl.hpp:284-692) for the LBM hot path.  Everything is plain C: opaque context
 * handles, device pointers as void*, sizes as uint64_t, int status codes.  No torch
 * types.  Citations are paths in the FluidX3D v3.7 reference repository.
 *
 * Conventions
 *   - Every function returns FX_OK (0) or a negative FX_ERR_* code.  The message of the
 *     last failure on the calling thread is returned by fx_last_error().  The reference
 *     calls print_error()+exit(1) instead (src/utilities.hpp:4074-4086); the C++ host
 *     layer (fluidx3d_amd/host) maps a nonzero status back to that behaviour.
 *   - Device arrays use the reference's layout: fi[i*N+n] (SoA DDFs, i = slot, n =
 *     x+(y+z*Ny)*Nx), rho[n], u[d*N+n], flags[n] (uint8 TYPE_* bits), F[d*N+n],
 *     alpha[n] (src/kernel.cpp:861-863, 960-967, src/lbm.cpp:121-176).
 *   - Kernel entry points are asynchronous on the context's HIP stream, in submission
 *     order, like the reference's in-order command queue (src/opencl.hpp:311).  Entry
 *     points that return host values (fx_lbm_object_*, fx_lbm_mass_sum, timing) wait
 *     for their own work.  A context is not re-entrant; separate contexts are
 *     independent.
 */
#ifndef FX_LBM_H
#define FX_LBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_OK 0
#define FX_ERR_ARG (-1)      /* invalid argument (null pointer, bad size, bad enum) */
#define FX_ERR_HIP (-2)      /* HIP runtime error; fx_last_error() has the HIP message */
#define FX_ERR_CONFIG (-3)   /* configuration rejected (grid size 0, nu <= 0, ...) */
#define FX_ERR_MEMORY (-4)   /* allocation failed or grid does not fit device memory */
#define FX_ERR_STATE (-5)    /* call not valid in the current state */

/* velocity sets (src/defines.hpp:5-8), collision operators (:10-11), DDF storage (:13-14) */
#define FX_D3Q19 19
#define FX_D3Q27 27
#define FX_SRT 0
#define FX_TRT 1
#define FX_FP32 0
#define FX_FP16S 1
#define FX_FP16C 2

/* extension bits (src/defines.hpp:18-25); FX_POROSITY is new: a per-cell Brinkman
 * porosity alpha[n] >= 0 folded into the Guo force (u = (j/rho + f/(2 rho))/(1+alpha/2),
 * F = f - alpha*rho*u).  With alpha == 0 it is bit-identical to FX_VOLUME_FORCE. */
#define FX_VOLUME_FORCE 1u
#define FX_FORCE_FIELD 2u
#define FX_EQUILIBRIUM_BOUNDARIES 4u
#define FX_UPDATE_FIELDS 8u
#define FX_POROSITY 16u
/* FX_PRECISE_DENSITY is new too: the equilibrium takes rho-1 as the sum of the shifted
 * DDFs instead of forming rho = sum+1 and subtracting 1 again (src/kernel.cpp:1005,
 * 1063-1066), which in FP32 quantises rho-1 to ~6e-8 per cell and step.  Off by default
 * (the reference's arithmetic); the adjoint configurations use it because at body forces
 * of 1e-6 per step that rounding is several percent of the flow's momentum input. */
#define FX_PRECISE_DENSITY 32u
/* MOVING_BOUNDARIES (src/kernel.cpp:1104-1113, 1433-1449): solid cells with a velocity
 * move (Ladd's correction on the populations bounced from them, wall density 1); cells next
 * to them carry TYPE_MS.  SUBGRID (src/kernel.cpp:1579-1593): Smagorinsky-Lilly eddy
 * viscosity. */
#define FX_MOVING_BOUNDARIES 64u
#define FX_SUBGRID 128u

/* cell type bits (src/defines.hpp:53-60) */
#define FX_TYPE_S 0x01u
#define FX_TYPE_E 0x02u
#define FX_TYPE_MS 0x03u /* next to a moving solid cell (src/lbm.cpp:402) */
#define FX_TYPE_T 0x04u
#define FX_TYPE_F 0x08u
#define FX_TYPE_I 0x10u
#define FX_TYPE_G 0x20u
#define FX_TYPE_X 0x40u
#define FX_TYPE_Y 0x80u

/* One lattice configuration: what the reference bakes into the OpenCL source as
 * #defines (src/lbm.cpp:334-468).  Here all variants are compiled ahead of time for
 * gfx950 and the context selects one. */
typedef struct fx_lbm_config {
    uint32_t Nx, Ny, Nz;     /* grid size; N = Nx*Ny*Nz cells */
    uint32_t velocity_set;   /* FX_D3Q19 or FX_D3Q27 */
    uint32_t collision;      /* FX_SRT or FX_TRT */
    uint32_t storage;        /* FX_FP32, FX_FP16S or FX_FP16C */
    uint32_t extensions;     /* OR of FX_VOLUME_FORCE, ... */
    float w;                 /* relaxation rate 1/tau as the kernels use it; use
                                fx_relaxation_rate(nu) for the reference's value */
} fx_lbm_config;

typedef struct fx_device_info {
    char name[256];          /* marketing name */
    char arch[64];           /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    uint64_t memory_bytes;   /* global memory */
    uint64_t memory_free;    /* free global memory now */
    uint32_t compute_units;
    uint32_t clock_mhz;
} fx_device_info;

typedef struct fx_ctx fx_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char* fx_last_error(void);

/* Device discovery; replaces Device_Info / get_devices (src/opencl.hpp:89-252). */
int fx_device_count(int* count);
int fx_device_info_get(int device, fx_device_info* out);

/* Relaxation rate w = 1/tau, tau = 3*nu+1/2 (src/lbm.hpp:132), after the reference's
 * decimal round trip: to_string(1.0f/tau) is printed with 9 significant digits and
 * re-parsed as a float literal by the device compiler (src/lbm.cpp:367,
 * src/utilities.hpp:2599-2621, 2745-2754). */
float fx_relaxation_rate(float nu);

/* Byte accounting per cell: device memory footprint (bytes_per_cell_device,
 * src/lbm.cpp:38-50) and algorithmic HBM traffic per cell and time step
 * (bandwidth_bytes_per_cell_device, src/lbm.cpp:51-72).  0 for an invalid config. */
uint64_t fx_memory_bytes_per_cell(const fx_lbm_config* cfg);
uint64_t fx_bandwidth_bytes_per_cell(const fx_lbm_config* cfg);

/* Context = device + one in-order HIP stream + the selected precompiled kernel variant.
 * Replaces Device(info, opencl_c_code) (src/opencl.hpp:284-340) and the per-domain
 * kernel set of LBM_Domain (src/lbm.cpp:96-176).  Validates the configuration like
 * LBM::sanity_checks_constructor (src/lbm.cpp:783-838) but returns an error code. */
int fx_ctx_create(int device, const fx_lbm_config* cfg, fx_ctx** out);
int fx_ctx_destroy(fx_ctx* ctx);
int fx_ctx_config(const fx_ctx* ctx, fx_lbm_config* out);
/* Wait for all work on the context stream; replaces Device::finish_queue (src/opencl.hpp:335). */
int fx_sync(fx_ctx* ctx);

/* Device buffers; replace the device half of Memory<T> (src/opencl.hpp:342-614).
 * Host memory stays with the caller.  fx_buf_write/read copy between a host pointer and
 * dst/src+offset; blocking != 0 waits for the copy (enqueue_write_to_device /
 * enqueue_read_from_device, src/opencl.hpp:510-531).  fx_buf_fill sets every byte.
 * fx_buf_copy is device to device on the context stream. */
int fx_buf_alloc(fx_ctx* ctx, uint64_t bytes, void** dptr);
int fx_buf_free(fx_ctx* ctx, void* dptr);
int fx_buf_write(fx_ctx* ctx, void* dst, uint64_t offset, uint64_t bytes, const void* host, int blocking);
int fx_buf_read(fx_ctx* ctx, const void* src, uint64_t offset, uint64_t bytes, void* host, int blocking);
int fx_buf_fill(fx_ctx* ctx, void* dst, uint64_t bytes, uint8_t value);
int fx_buf_copy(fx_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* Box transfers of a field of d components with N elements of elem_bytes each (component c
 * starts at element c*N; cell n = x + (y + z*Ny)*Nx).  The rows x0..x1-1 of the box
 * [x0,x1) x [y0,y1) x [z0,z1) of one component (every component when component < 0) move
 * between the device buffer and the SAME element offsets of the host array, which has the
 * full N*d layout.  A row that runs past element N*d is cut there, as in
 * Memory<T>::read_from_device_1d/2d/3d and write_to_device_1d/2d/3d
 * (src/opencl.hpp:532-607); the 1-D form is Nx = N, Ny = 1, y0 = z0 = 0, y1 = z1 = 1,
 * the 2-D form z0 = 0, z1 = 1.  Inverted ranges and component >= d are argument errors. */
int fx_buf_read_box(fx_ctx* ctx, const void* src, void* host, uint32_t elem_bytes, uint64_t N, uint32_t d,
                    uint64_t Nx, uint64_t Ny, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, uint64_t z0,
                    uint64_t z1, int component, int blocking);
int fx_buf_write_box(fx_ctx* ctx, void* dst, const void* host, uint32_t elem_bytes, uint64_t N, uint32_t d,
                     uint64_t Nx, uint64_t Ny, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, uint64_t z0,
                     uint64_t z1, int component, int blocking);

/* ---- LBM kernels (typed launches; replace Kernel::enqueue_run, src/opencl.hpp:616-692) ---- */

/* initialize (src/kernel.cpp:1358-1430): zero u of TYPE_S cells (with
 * FX_MOVING_BOUNDARIES: of solid cells with only solid neighbours, and mark TYPE_MS in
 * flags), store feq(rho,u) of every cell with the odd-step Esoteric-Pull convention (t=1).
 * Bound as kernel_initialize(fi,rho,u,flags) in src/lbm.cpp:127; the host sequencing
 * (t=1 then t=0, src/lbm.cpp:881-922) is the caller's. */
int fx_lbm_initialize(fx_ctx* ctx, void* fi, const float* rho, float* u, uint8_t* flags);

/* update_moving_boundaries (src/kernel.cpp:1433-1449): re-mark TYPE_MS after wall
 * velocities changed (FX_MOVING_BOUNDARIES).  Replaces LBM_Domain::
 * enqueue_update_moving_boundaries (src/lbm.cpp:241-243). */
int fx_lbm_update_moving_boundaries(fx_ctx* ctx, const float* u, uint8_t* flags);

/* stream_collide (src/kernel.cpp:1454-1636): one time step t, in place (Esoteric-Pull).
 * F (3N floats) is read only with FX_FORCE_FIELD, alpha (N floats) only with
 * FX_POROSITY, rho/u only for TYPE_E cells with FX_EQUILIBRIUM_BOUNDARIES and written
 * with FX_UPDATE_FIELDS, u of solid cells (wall velocities) with FX_MOVING_BOUNDARIES.  Replaces LBM_Domain::enqueue_stream_collide (src/lbm.cpp:181-183). */
int fx_lbm_stream_collide(fx_ctx* ctx, void* fi, float* rho, float* u, const uint8_t* flags,
                          uint64_t t, float fx, float fy, float fz, const float* F, const float* alpha);

/* steps consecutive stream_collide launches for t = t0 .. t0+steps-1, bracketed by HIP
 * events on the context stream; waits and returns the elapsed milliseconds in *ms
 * (may be NULL: then it does not wait).  Replaces the loop of LBM::run
 * (src/lbm.cpp:955-975) without the per-step host synchronisation of src/lbm.cpp:951. */
int fx_lbm_run(fx_ctx* ctx, void* fi, float* rho, float* u, const uint8_t* flags, uint64_t t0, uint64_t steps,
               float fx, float fy, float fz, const float* F, const float* alpha, float* ms);

/* update_fields (src/kernel.cpp:1794-1870): rho/u of non-solid cells from the DDFs at
 * time t (TYPE_E cells keep theirs with FX_EQUILIBRIUM_BOUNDARIES).  Replaces
 * LBM_Domain::enqueue_update_fields (src/lbm.cpp:184-191). */
int fx_lbm_update_fields(fx_ctx* ctx, const void* fi, float* rho, float* u, const uint8_t* flags,
                         uint64_t t, float fx, float fy, float fz, const float* F, const float* alpha);

/* update_force_field (src/kernel.cpp:1873-1884): F = momentum exchange on TYPE_S cells at
 * time t.  reset_force_field (:1885-1889) zeroes F.  Replaces
 * LBM_Domain::enqueue_update_force_field (src/lbm.cpp:207-212). */
int fx_lbm_update_force_field(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, float* F);
int fx_lbm_reset_force_field(fx_ctx* ctx, float* F);

/* Object sums over cells with flags[n]==marker (src/kernel.cpp:1901-1959; host
 * src/lbm.cpp:213-238, 986-1016).  Per-cell terms in FP32 as in the reference, summed
 * in FP64 in a fixed order (deterministic, unlike the reference's float atomics).
 * Synchronous.  center_of_mass writes (sum x, sum y, sum z, cell count). */
int fx_lbm_object_force(fx_ctx* ctx, const float* F, const uint8_t* flags, uint8_t marker, double out[3]);
int fx_lbm_object_torque(fx_ctx* ctx, const float* F, const uint8_t* flags, uint8_t marker,
                         float cx, float cy, float cz, double out[3]);
int fx_lbm_object_center_of_mass(fx_ctx* ctx, const uint8_t* flags, uint8_t marker, double out[4]);

/* Total mass: sum over all cells (solid ones included) of 1 + the shifted DDFs stored in
 * the cell's own slots, FP64, deterministic.  This is exactly invariant under
 * stream_collide (values only move inside disjoint Esoteric-Pull footprints and the
 * collision conserves sum_i f_i); populations in flight toward a wall are held in the
 * solid cell's slots.  t is accepted for symmetry and not used.  No reference
 * counterpart; the conservation check. */
int fx_lbm_mass_sum(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, double* out);

/* fx_lbm_mass_sum without any allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The partial sums and the result live in a static
 * device array of the library (8 MiB per device, held until the process ends), so the
 * lattice's grid may have at most 2^20 blocks (FX_ERR_CONFIG beyond; about 268 M cells
 * with 256-cell rows).  Same result as fx_lbm_mass_sum, bit for bit.  All calls on one
 * device share the array: calls that overlap on one device (other contexts' streams,
 * other threads) race and must be serialised by the caller. */
int fx_lbm_mass_sum_noalloc(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, double* out);

/* ---- differentiation (no reference counterpart; the reference has no adjoint) ----
 *
 * Objective J = mean over steps t in [window_begin, steps) of the x-momentum
 * sum_i c_ix v_i of the populations loaded at step t, summed over the non-solid cells of
 * the plane x = plane_x (the outlet flux of a channel).  fx_lbm_forward runs the steps
 * from t = 0 (fi initialised by fx_lbm_initialize), stores a copy of fi every
 * checkpoint_interval steps and returns J (FP64 sums).  fx_lbm_backward sweeps the
 * checkpoint segments backward: each segment is recomputed forward recording the
 * pre-collision moments (rho, j) of every step (16 B/cell/step), then the adjoint step
 * (reverse Esoteric-Pull with the transposed collision, FP32) runs backward through it.
 * grad_alpha[n] (N floats, device) receives dJ/dalpha[n] for FX_POROSITY. */
typedef struct fx_adjoint_config {
    uint64_t steps;               /* forward steps T, t = 0 .. T-1 */
    uint64_t checkpoint_interval; /* DDF checkpoint every k steps; 0 selects ceil(sqrt(T)) */
    uint64_t window_begin;        /* first step of the objective's averaging window (< steps) */
    uint32_t plane_x;             /* objective plane, 0 <= plane_x < Nx */
} fx_adjoint_config;

typedef struct fx_adjoint fx_adjoint;

/* device bytes fx_adjoint_create will allocate for this context and configuration */
int fx_adjoint_memory(const fx_ctx* ctx, const fx_adjoint_config* cfg, uint64_t* bytes);
int fx_adjoint_create(fx_ctx* ctx, const fx_adjoint_config* cfg, fx_adjoint** out);
int fx_adjoint_destroy(fx_adjoint* adj);
/* forward pass with checkpoints; *objective = J; *ms = GPU time (either may be NULL) */
int fx_lbm_forward(fx_ctx* ctx, fx_adjoint* adj, void* fi, float* rho, float* u, const uint8_t* flags,
                   float fx, float fy, float fz, const float* F, const float* alpha, double* objective, float* ms);
/* checkpointed adjoint sweep after fx_lbm_forward; rho/u hold the TYPE_E presets */
int fx_lbm_backward(fx_ctx* ctx, fx_adjoint* adj, const float* rho, const float* u, const uint8_t* flags,
                    float fx, float fy, float fz, const float* F, const float* alpha, float* grad_alpha, float* ms);
/* single-step primitives of the two passes (tests, custom schedules):
 * stream_collide that also writes moments[c*N+n] = (rho, jx, jy, jz) of step t; */
int fx_lbm_stream_collide_record(fx_ctx* ctx, void* fi, float* rho, float* u, const uint8_t* flags, uint64_t t,
                                 float fx, float fy, float fz, const float* F, const float* alpha, float* moments);
/* one adjoint step t on lambda (FP32, q*N, in place) with the moments of step t; adds
 * obj_weight*c_ix on the plane x = plane_x and accumulates grad_alpha (FX_POROSITY) */
int fx_lbm_adjoint_step(fx_ctx* ctx, float* lambda, const float* rho, const float* u, const uint8_t* flags, uint64_t t,
                        float fx, float fy, float fz, const float* F, const float* alpha, const float* moments,
                        uint32_t plane_x, float obj_weight, float* grad_alpha);
/* plane x-momentum of the populations loaded at step t (FP64 sum), synchronous */
int fx_lbm_objective_plane(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, uint32_t plane_x, double* out);

/* ---- halo transfer between the domains of a split lattice ----
 *
 * A lattice split into domains along an axis gives every domain one halo layer on each
 * split side (src/lbm.cpp:721-733: domain size N/D + 2 along a split axis).  After each
 * stream_collide at step t, fx_lbm_transfer_extract of FX_FIELD_FI packs the DDF values
 * that cross the two faces normal to `direction` (0 = x, 1 = y, 2 = z) into buffer_p
 * (+ face) and buffer_m (- face); the neighbours then fx_lbm_transfer_insert them:
 * domain d inserts as buffer_p the buffer_m of its + neighbour and as buffer_m the
 * buffer_p of its - neighbour.  Directions are communicated one after another (x, y, z),
 * each extract finished before the matching insert.  Field data (FX_FIELD_RHO_U_FLAGS:
 * 17 B per face cell, FX_FIELD_FLAGS: 1 B, FX_FIELD_F: 12 B) moves from the interior
 * layers 1 and N-2 into the halo layers 0 and N-1 the same way.  Buffers are device
 * memory of fx_transfer_bytes each; values move as stored (no FP16 decode).  Replace the
 * transfer kernels of src/kernel.cpp:2049-2183 and LBM_Domain::enqueue_transfer_*_field
 * (src/lbm.cpp:1343-1354); the exchange itself (LBM::communicate_field,
 * src/lbm.cpp:1355-1383, host pointer swaps over PCIe) is the caller's, device to
 * device. */
enum { FX_FIELD_FI = 0, FX_FIELD_RHO_U_FLAGS = 1, FX_FIELD_FLAGS = 2, FX_FIELD_F = 3 };
/* bytes of one side's buffer for `field` across the faces normal to `direction` */
uint64_t fx_transfer_bytes(const fx_lbm_config* cfg, uint32_t field, uint32_t direction);
int fx_lbm_transfer_extract(fx_ctx* ctx, uint32_t field, uint32_t direction, uint64_t t, void* buffer_p, void* buffer_m,
                            const void* fi, const float* rho, const float* u, const uint8_t* flags, const float* F);
int fx_lbm_transfer_insert(fx_ctx* ctx, uint32_t field, uint32_t direction, uint64_t t, const void* buffer_p, const void* buffer_m,
                           void* fi, float* rho, float* u, uint8_t* flags, float* F);

/* ---- diagnostics ---- */

/* Storage codec round trip on the device: codes[k] = encode(x[k]), decoded[k] =
 * decode(codes[k]) for storage FX_FP16S or FX_FP16C, exactly as the kernels store and
 * load DDFs (src/lbm.cpp:410-414, src/kernel.cpp:848-859).  Device pointers. */
int fx_codec_roundtrip(fx_ctx* ctx, uint32_t storage, const float* x, uint16_t* codes, float* decoded, uint64_t n);

/* Device-to-device copy kernel over bytes (multiple of 16) repeated reps times,
 * timed with HIP events: the measured HBM peak and the counter calibration. */
int fx_bench_copy(fx_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint32_t reps, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* FX_LBM_H */
