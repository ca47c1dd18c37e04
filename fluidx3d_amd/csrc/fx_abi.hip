// C ABI of libfx_lbm.so (declared in include/fx_lbm.h): device discovery, contexts,
// buffers, typed kernel launches and FP64 reductions.  Replaces the OpenCL-Wrapper
// device/buffer/kernel layer of the reference (src/opencl.hpp:284-692).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <mutex>
#include <set>
#include "fx_lbm.h"
#include "lbm_launch.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& message) {
	g_last_error = message;
	return code;
}
int fail_hip(hipError_t e, const char* what) {
	return fail(FX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define FX_HIP(call) do { const hipError_t fx_e_ = (call); if(fx_e_ != hipSuccess) return fail_hip(fx_e_, #call); } while(0)

bool valid_enums(const fx_lbm_config* c) {
	return c && (c->velocity_set == FX_D3Q19 || c->velocity_set == FX_D3Q27) && c->collision <= FX_TRT && c->storage <= FX_FP16C &&
	       (c->extensions & ~(FX_VOLUME_FORCE | FX_FORCE_FIELD | FX_EQUILIBRIUM_BOUNDARIES | FX_UPDATE_FIELDS | FX_POROSITY | FX_PRECISE_DENSITY |
	                          FX_MOVING_BOUNDARIES | FX_SUBGRID)) == 0u;
}

} // namespace

struct fx_ctx {
	int device = 0;
	fx_lbm_config cfg{};
	uint64_t N = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	const fx::LbmKernels* k = nullptr;
	fx::LaunchShape shape{};    // one lane per cell: every kernel but stream_collide
	fx::LaunchShape shape_sc{}; // stream_collide (residency cap, XCD-mapped grid)
	fx::LaunchShape shape_adj{}; // record pass and adjoint step (XCD-mapped grid)
	double* partial = nullptr; // reduction partials (device)
	uint64_t partial_cap = 0;  // in doubles
	double* result = nullptr;  // 4 doubles (device)
	double* scratch = nullptr; // this device's copy of the static scratch array (fx_lbm_mass_sum_noalloc)
	std::set<void*> buffers;   // allocations owned by this context
};

struct fx_adjoint {
	fx_ctx* ctx = nullptr;
	fx_adjoint_config cfg{};
	uint64_t k = 0, checkpoints = 0, fi_bytes = 0, obj_blocks = 0;
	char* ckpt = nullptr;       // checkpoints * fi_bytes
	void* work_fi = nullptr;    // recomputation state
	float* moments = nullptr;   // k * 4 * N
	float* lambda = nullptr;    // q * N
	double* partial = nullptr;  // obj_blocks
	double* acc = nullptr;      // 1
	bool forward_done = false;
};

namespace {

int set_device(const fx_ctx* ctx) {
	FX_HIP(hipSetDevice(ctx->device));
	return FX_OK;
}

// Scratch memory of fx_lbm_mass_sum_noalloc, for callers that cannot allocate: the FP64
// per-block partials (at most NOALLOC_MASS_PARTIALS) followed by the one-double result.
// Every device the library runs on holds its own copy (8 MiB) until the process ends, and
// every call on a device shares it.  Host code never passes the array's name to a kernel or
// copy (on the host it is a stand-in, not a device address): it uses the device address
// that resolve_scratch obtains once per device with that device current.
constexpr uint64_t NOALLOC_MASS_PARTIALS = 1ull << 20;
static __device__ __align__(256) double scratch[NOALLOC_MASS_PARTIALS + 1];

std::mutex g_scratch_mutex;
double* g_scratch_address[256] = {}; // per device ordinal

// Device address of `scratch` on `device` (which must be current), resolved on the first
// call for that device.
int resolve_scratch(int device, double** out) {
	if(device < 0 || device >= 256) return fail(FX_ERR_ARG, "device ordinal out of range");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if(!g_scratch_address[device]) {
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
		g_scratch_address[device] = (double*)p;
	}
	*out = g_scratch_address[device];
	return FX_OK;
}

int ensure_partial(fx_ctx* ctx, uint64_t doubles) {
	if(ctx->partial_cap >= doubles) return FX_OK;
	if(ctx->partial) FX_HIP(hipFree(ctx->partial));
	ctx->partial = nullptr;
	ctx->partial_cap = 0;
	FX_HIP(hipMalloc(&ctx->partial, doubles * sizeof(double)));
	ctx->partial_cap = doubles;
	return FX_OK;
}

fx::LbmArgs make_args(const fx_ctx* ctx, const void* fi, const float* rho, const float* u, const uint8_t* flags, uint64_t t,
                      float fx_, float fy_, float fz_, const float* F, const float* alpha) {
	fx::LbmArgs a;
	a.fi = const_cast<void*>(fi);
	a.rho = const_cast<float*>(rho);
	a.u = const_cast<float*>(u);
	a.flags = flags;
	a.F = F;
	a.alpha = alpha;
	a.N = ctx->N;
	a.Nx = ctx->cfg.Nx; a.Ny = ctx->cfg.Ny; a.Nz = ctx->cfg.Nz;
	a.ext = ctx->cfg.extensions;
	a.odd = (uint32_t)(t & 1ull);
	a.xchunks = 0u; // set by the stream_collide launcher from its shape
	a.reverse = 0u;
	a.w = ctx->cfg.w;
	a.wm = fx::trt_minus_rate(a.w);
	a.fx = fx_; a.fy = fy_; a.fz = fz_;
	return a;
}

int check_step_args(const fx_ctx* ctx, const void* fi, const float* rho, const float* u, const uint8_t* flags, const float* F, const float* alpha) {
	if(!ctx) return fail(FX_ERR_ARG, "null context");
	if(!fi || !flags) return fail(FX_ERR_ARG, "fi and flags must be device pointers");
	const uint32_t ext = ctx->cfg.extensions;
	if((ext & FX_FORCE_FIELD) && !F) return fail(FX_ERR_ARG, "FORCE_FIELD is enabled but F is null");
	if((ext & FX_POROSITY) && !alpha) return fail(FX_ERR_ARG, "POROSITY is enabled but alpha is null");
	if((ext & (FX_EQUILIBRIUM_BOUNDARIES | FX_UPDATE_FIELDS)) && (!rho || !u)) return fail(FX_ERR_ARG, "rho and u are required with EQUILIBRIUM_BOUNDARIES or UPDATE_FIELDS");
	if((ext & FX_MOVING_BOUNDARIES) && !u) return fail(FX_ERR_ARG, "u (the wall velocities) is required with MOVING_BOUNDARIES");
	return FX_OK;
}

int check_launch(const char* what) {
	const hipError_t e = hipGetLastError();
	if(e != hipSuccess) return fail_hip(e, what);
	return FX_OK;
}

// Relaxation-rate text exactly as the reference prints it: split_float scales x into
// [1,10) with float multiplications, keeps 8 decimals of the fraction (rounded half up
// in float arithmetic) and prints "I.DDDDDDDD[E<exp>]"
// (src/utilities.hpp:2599-2621, 2669-2675, 2745-2754).
std::string reference_float_text(float x) {
	std::string text = x < 0.0f ? "-" : "";
	x = std::fabs(x);
	if(std::isnan(x)) return text + "NaN";
	if(std::isinf(x)) return text + "Inf";
	int e10 = 0;
	static const float up[6] = {1E32f, 1E16f, 1E8f, 1E4f, 1E2f, 1E1f};
	static const float down[6] = {1E-32f, 1E-16f, 1E-8f, 1E-4f, 1E-2f, 1E-1f};
	static const float small[6] = {1E-31f, 1E-15f, 1E-7f, 1E-3f, 1E-1f, 1E0f};
	static const int step[6] = {32, 16, 8, 4, 2, 1};
	if(x >= 10.0f) {
		for(int k = 0; k < 6; k++) if(x >= up[k]) { x *= down[k]; e10 += step[k]; }
	}
	if(x > 0.0f && x <= 1.0f) {
		for(int k = 0; k < 6; k++) if(x < small[k]) { x *= up[k]; e10 -= step[k]; }
	}
	unsigned int whole = (unsigned int)x;
	const float frac8 = (x - (float)whole) * 1E8f;
	unsigned int digits = (unsigned int)frac8;
	if(frac8 - (float)digits >= 0.5f && ++digits >= 100000000u) {
		digits = 0u;
		if(++whole >= 10u) { whole = 1u; e10++; }
	}
	char buf[48];
	std::snprintf(buf, sizeof(buf), "%u.%08u", whole, digits);
	text += buf;
	if(e10 != 0) text += "E" + std::to_string(e10);
	return text;
}

} // namespace

// ---------------------------------------------------------------------------------------
// Device kernels that do not depend on the lattice variant
// ---------------------------------------------------------------------------------------
namespace fx {

constexpr int RED_BLOCK = 256;
constexpr int RED_GRID = 1024;

// Object sums (src/kernel.cpp:1901-1959): mode 0 force, 1 torque about (cx,cy,cz),
// 2 centre of mass (positions and cell count).  Per-cell terms in FP32 as the reference
// computes them, accumulated in FP64; one partial (4 doubles) per block.
__global__ __launch_bounds__(RED_BLOCK) void k_object_partial(const float* F, const uint8_t* flags, uint32_t marker, uint64_t N,
                                                              uint32_t Nx, uint32_t Ny, uint32_t Nz, int mode, float cx, float cy, float cz,
                                                              double* partial) {
	__shared__ double red[4][RED_BLOCK];
	double acc[4] = {0.0, 0.0, 0.0, 0.0};
	const uint64_t NxNy = (uint64_t)Nx * Ny;
	for(uint64_t n = (uint64_t)blockIdx.x * RED_BLOCK + threadIdx.x; n < N; n += (uint64_t)gridDim.x * RED_BLOCK) {
		if(flags[n] != marker) continue;
		if(mode == 0) {
			acc[0] += (double)F[n]; acc[1] += (double)F[N + n]; acc[2] += (double)F[2ull * N + n];
		} else {
			const uint64_t tt = n % NxNy;
			const uint32_t x = (uint32_t)(tt % Nx), y = (uint32_t)(tt / Nx), z = (uint32_t)(n / NxNy);
			const float px = (float)x + 0.5f - 0.5f * (float)Nx; // position(), src/kernel.cpp:824-826
			const float py = (float)y + 0.5f - 0.5f * (float)Ny;
			const float pz = (float)z + 0.5f - 0.5f * (float)Nz;
			if(mode == 1) {
				const float rx = px - cx, ry = py - cy, rz = pz - cz;
				const float Fx = F[n], Fy = F[N + n], Fz = F[2ull * N + n];
				acc[0] += (double)(ry * Fz - rz * Fy);
				acc[1] += (double)(rz * Fx - rx * Fz);
				acc[2] += (double)(rx * Fy - ry * Fx);
			} else {
				acc[0] += (double)px; acc[1] += (double)py; acc[2] += (double)pz; acc[3] += 1.0;
			}
		}
	}
	for(int k = 0; k < 4; k++) red[k][threadIdx.x] = acc[k];
	__syncthreads();
	for(int s = RED_BLOCK / 2; s > 0; s >>= 1) {
		if((int)threadIdx.x < s) for(int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
		__syncthreads();
	}
	if(threadIdx.x < 4) partial[(uint64_t)blockIdx.x * 4u + threadIdx.x] = red[threadIdx.x][0];
}

// Sum `count` partials of `stride` doubles each, component by component, in a fixed
// order (single block): out[k] = sum_b partial[b*stride+k].
__global__ __launch_bounds__(RED_BLOCK) void k_reduce_partials(const double* partial, uint64_t count, uint32_t stride, double* out) {
	__shared__ double red[RED_BLOCK];
	for(uint32_t k = 0; k < stride; k++) {
		double s = 0.0;
		for(uint64_t b = threadIdx.x; b < count; b += RED_BLOCK) s += partial[b * stride + k];
		red[threadIdx.x] = s;
		__syncthreads();
		for(int h = RED_BLOCK / 2; h > 0; h >>= 1) {
			if((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
			__syncthreads();
		}
		if(threadIdx.x == 0) out[k] = red[0];
		__syncthreads();
	}
}

// acc[0] += sum of count partials (single block, fixed order)
__global__ __launch_bounds__(RED_BLOCK) void k_reduce_add(const double* partial, uint64_t count, double* acc) {
	__shared__ double red[RED_BLOCK];
	double s = 0.0;
	for(uint64_t b = threadIdx.x; b < count; b += RED_BLOCK) s += partial[b];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int h = RED_BLOCK / 2; h > 0; h >>= 1) {
		if((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
		__syncthreads();
	}
	if(threadIdx.x == 0) acc[0] += red[0];
}

// Element-wise kernels run a grid-stride loop over a grid of at most STRIDE_BLOCKS blocks,
// so that any element count stays within a launch's 2^32-1 work-items (2^20 blocks of 256
// lanes cover 2^28 elements per pass: a 4 GiB copy in one pass).
constexpr uint64_t STRIDE_BLOCKS = 1u << 20;
inline unsigned stride_grid(uint64_t n, uint32_t block) {
	const uint64_t b = (n + block - 1u) / block;
	return (unsigned)(b < 1u ? 1u : b < STRIDE_BLOCKS ? b : STRIDE_BLOCKS);
}

__global__ void k_reset3(float* F, uint64_t n3) {
	for(uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (uint64_t)gridDim.x * blockDim.x) F[i] = 0.0f;
}

// Odd elements go through encode_high, the form the two-cells-per-lane kernel packs its
// second cell's codes with, so one round trip checks both encoders.
template<uint32_t ST> __global__ void k_codec(const float* x, uint16_t* codes, float* decoded, uint64_t n) {
	for(uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		const uint16_t c = (i & 1u) ? (uint16_t)(Storage<ST>::encode_high(x[i], 0u) >> 16) : Storage<ST>::encode(x[i]);
		codes[i] = c;
		decoded[i] = Storage<ST>::decode(c);
	}
}

// Plain copy, 16 bytes per lane, non-temporal loads and stores (the fastest copy measured:
// 6.48 TB/s over 4 GiB against 6.20 TB/s with cached accesses).
typedef float copy_vec __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const copy_vec* __restrict__ src, copy_vec* __restrict__ dst, uint64_t n16) {
	for(uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * 256u)
		__builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

} // namespace fx

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
// Field halo transfer (rho/u/flags, flags, F): interior layers 1 and N-2 -> halo layers 0
// and N-1 of the neighbours.  Face cell a as in k_transfer_fi; layout per field: rho/u/flags
// = 4 float planes of A then A flag bytes; flags = A bytes; F = 3 float planes of A.
template<int AXIS, bool INSERT>
__global__ void k_transfer_field(uint32_t field, char* buffer_p, char* buffer_m, float* rho, float* u, uint8_t* flags, float* F,
                                 uint32_t Nx, uint32_t Ny, uint32_t Nz) {
	const uint32_t Nax = AXIS == 0 ? Nx : AXIS == 1 ? Ny : Nz;
	const uint64_t A = AXIS == 0 ? (uint64_t)Ny * Nz : AXIS == 1 ? (uint64_t)Nx * Nz : (uint64_t)Nx * Ny;
	const uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if(a >= A) return;
	const uint64_t N = (uint64_t)Nx * Ny * Nz;
	for(int sd = 0; sd < 2; sd++) {
		char* buffer = sd == 0 ? buffer_p : buffer_m;
		const uint32_t layer = INSERT ? (sd == 0 ? Nax - 1u : 0u) : (sd == 0 ? Nax - 2u : 1u);
		uint32_t x, y, z;
		fx::face_cell<AXIS>((uint32_t)a, layer, Nx, Ny, x, y, z);
		const uint64_t n = ((uint64_t)z * Ny + y) * Nx + x;
		float* bf = (float*)buffer;
		if(field == FX_FIELD_RHO_U_FLAGS) {
			uint8_t* bb = (uint8_t*)buffer + 16u * A;
			if(INSERT) { rho[n] = bf[a]; u[n] = bf[A + a]; u[N + n] = bf[2u * A + a]; u[2u * N + n] = bf[3u * A + a]; flags[n] = bb[a]; }
			else { bf[a] = rho[n]; bf[A + a] = u[n]; bf[2u * A + a] = u[N + n]; bf[3u * A + a] = u[2u * N + n]; bb[a] = flags[n]; }
		} else if(field == FX_FIELD_FLAGS) {
			uint8_t* bb = (uint8_t*)buffer;
			if(INSERT) flags[n] = bb[a]; else bb[a] = flags[n];
		} else { // FX_FIELD_F
			if(INSERT) { F[n] = bf[a]; F[N + n] = bf[A + a]; F[2u * N + n] = bf[2u * A + a]; }
			else { bf[a] = F[n]; bf[A + a] = F[N + n]; bf[2u * A + a] = F[2u * N + n]; }
		}
	}
}

// Box copy between a device buffer and a host array of the same layout (elements e at byte
// e*eb on both sides).  For every component i and slice z, the rows y0..y1-1 start at
// element i*N + x0 + (y + z*Ny)*Nx and hold x1-x0 elements; rows that end past N*d are cut
// at N*d and rows that start there are skipped.  The rows of a slice that lie whole below
// N*d go in one pitched copy (pitch Nx elements) when a row is no wider than the pitch.
static int box_copy(fx_ctx* ctx, char* device, char* host, bool to_device, uint32_t eb, uint64_t N, uint32_t d, uint64_t Nx,
                    uint64_t Ny, uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, uint64_t z0, uint64_t z1, int component,
                    int blocking) {
	if(!ctx || !device || !host) return fail(FX_ERR_ARG, "null argument");
	if(eb == 0u || N == 0ull || d == 0u || Nx == 0ull || Ny == 0ull) return fail(FX_ERR_ARG, "elem_bytes, N, d, Nx and Ny must be positive");
	if(x1 < x0 || y1 < y0 || z1 < z0) return fail(FX_ERR_ARG, "box ranges must satisfy x0 <= x1, y0 <= y1, z0 <= z1");
	if(component >= (int)d) return fail(FX_ERR_ARG, "component must be below the number of components d (or negative for all)");
	const uint64_t range = N * d, w = x1 - x0;
	const hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
	auto copy = [&](uint64_t e, uint64_t width, uint64_t pitch, uint64_t rows) -> hipError_t {
		char* dv = device + e * eb;
		char* hs = host + e * eb;
		if(rows == 1ull) return hipMemcpyAsync(to_device ? dv : hs, to_device ? hs : dv, width * eb, kind, ctx->stream);
		return hipMemcpy2DAsync(to_device ? dv : hs, pitch * eb, to_device ? hs : dv, pitch * eb, width * eb, rows, kind, ctx->stream);
	};
	const uint32_t i0 = component < 0 ? 0u : (uint32_t)component, i1 = component < 0 ? d : i0 + 1u;
	if(w > 0ull && y1 > y0) {
		for(uint32_t i = i0; i < i1; i++) {
			for(uint64_t z = z0; z < z1; z++) {
				const uint64_t first = (uint64_t)i * N + x0 + (y0 + z * Ny) * Nx; // element of row y0
				if(first >= range) break; // later slices start further on
				uint64_t whole = 0ull;      // rows from y0 on that end at or below N*d
				if(first + w <= range) whole = std::min<uint64_t>(y1 - y0, (range - w - first) / Nx + 1ull);
				if(whole > 0ull) {
					if(w <= Nx) FX_HIP(copy(first, w, Nx, whole));
					else for(uint64_t r = 0ull; r < whole; r++) FX_HIP(copy(first + r * Nx, w, Nx, 1ull));
				}
				if(whole < y1 - y0) { // the row that N*d cuts, if it starts below N*d
					const uint64_t e = first + whole * Nx;
					if(e < range) FX_HIP(copy(e, range - e, Nx, 1ull));
					break;
				}
			}
		}
	}
	if(blocking) FX_HIP(hipStreamSynchronize(ctx->stream));
	return FX_OK;
}

extern "C" {

const char* fx_last_error(void) { return g_last_error.c_str(); }

int fx_device_count(int* count) {
	if(!count) return fail(FX_ERR_ARG, "null count");
	int c = 0;
	const hipError_t e = hipGetDeviceCount(&c);
	if(e == hipErrorNoDevice) { *count = 0; return FX_OK; }
	if(e != hipSuccess) return fail_hip(e, "hipGetDeviceCount");
	*count = c;
	return FX_OK;
}

int fx_device_info_get(int device, fx_device_info* out) {
	if(!out) return fail(FX_ERR_ARG, "null output");
	hipDeviceProp_t p;
	FX_HIP(hipGetDeviceProperties(&p, device));
	std::memset(out, 0, sizeof(*out));
	std::snprintf(out->name, sizeof(out->name), "%s", p.name);
	std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
	out->memory_bytes = p.totalGlobalMem;
	out->compute_units = (uint32_t)p.multiProcessorCount;
	out->clock_mhz = (uint32_t)(p.clockRate / 1000);
	int current = 0;
	FX_HIP(hipGetDevice(&current));
	FX_HIP(hipSetDevice(device));
	size_t free_b = 0, total_b = 0;
	const hipError_t e = hipMemGetInfo(&free_b, &total_b);
	(void)hipSetDevice(current);
	if(e != hipSuccess) return fail_hip(e, "hipMemGetInfo");
	out->memory_free = free_b;
	return FX_OK;
}

float fx_relaxation_rate(float nu) {
	const float tau = 3.0f * nu + 0.5f;
	return std::strtof(reference_float_text(1.0f / tau).c_str(), nullptr);
}

uint64_t fx_memory_bytes_per_cell(const fx_lbm_config* c) {
	if(!valid_enums(c)) return 0;
	uint64_t b = (uint64_t)c->velocity_set * (c->storage == FX_FP32 ? 4u : 2u) + 17u; // fi, rho, u, flags
	if(c->extensions & FX_FORCE_FIELD) b += 12u;
	if(c->extensions & FX_POROSITY) b += 4u;
	return b;
}

uint64_t fx_bandwidth_bytes_per_cell(const fx_lbm_config* c) {
	if(!valid_enums(c)) return 0;
	uint64_t b = (uint64_t)c->velocity_set * 2u * (c->storage == FX_FP32 ? 4u : 2u) + 1u; // q loads + q stores + flags
	if(c->extensions & FX_UPDATE_FIELDS) b += 16u;
	if(c->extensions & FX_FORCE_FIELD) b += 12u;
	if(c->extensions & FX_POROSITY) b += 4u;
	if(c->extensions & FX_MOVING_BOUNDARIES) b += (uint64_t)c->velocity_set - 1u; // neighbour flags (src/lbm.cpp:62-64)
	return b;
}

int fx_ctx_create(int device, const fx_lbm_config* cfg, fx_ctx** out) {
	if(!out) return fail(FX_ERR_ARG, "null output handle");
	*out = nullptr;
	if(!cfg) return fail(FX_ERR_ARG, "null config");
	if(!valid_enums(cfg)) return fail(FX_ERR_CONFIG, "unsupported velocity set / collision / storage / extension bits");
	const uint64_t N = (uint64_t)cfg->Nx * cfg->Ny * cfg->Nz;
	if(N == 0ull) return fail(FX_ERR_CONFIG, "Grid point number is 0");
	if(cfg->Ny > 65536u || cfg->Nz > 65536u) return fail(FX_ERR_CONFIG, "Ny and Nz must be at most 65536");
	if(!(cfg->w > 0.0f && cfg->w < 2.0f)) return fail(FX_ERR_CONFIG, "relaxation rate w must be in (0,2): viscosity must be positive");
	int count = 0;
	const int rc = fx_device_count(&count);
	if(rc != FX_OK) return rc;
	if(device < 0 || device >= count) return fail(FX_ERR_ARG, "no such device");
	fx_ctx* ctx = new fx_ctx();
	ctx->device = device;
	ctx->cfg = *cfg;
	ctx->N = N;
	ctx->k = fx::select_kernels(cfg->velocity_set, cfg->storage, cfg->collision == FX_TRT, cfg->extensions);
	ctx->shape = fx::lattice_shape(cfg->Nx, cfg->Ny, cfg->Nz);
	ctx->shape_sc = fx::stream_shape(cfg->Nx, cfg->Ny, cfg->Nz, ctx->k->stream_collide_xcd_map);
	ctx->shape_adj = fx::adjoint_shape(cfg->Nx, cfg->Ny, cfg->Nz);
	// cell-linear grid for rows that miss the 128-byte lines and for large FP32 lattices
	// (the adjoint's FP32 lambda rows are aligned whenever the DDF rows are)
	const uint32_t eb = cfg->storage == FX_FP32 ? 4u : 2u;
	if(fx::use_linear_grid(*ctx->k, cfg->Nx, eb, N, cfg->velocity_set)) {
		ctx->shape_sc = fx::linear_shape(N, ctx->k->stream_collide_linear_xcd);
		ctx->shape_sc.buf = ctx->k->stream_collide_linear_buffer && N * cfg->velocity_set * eb <= 0xFFFFFFFFull;
		ctx->shape_adj = fx::linear_shape(N, false);
	} else if(ctx->k->stream_collide_pairs && fx::pair_rows(cfg->Nx)) {
		ctx->shape_sc = fx::pair_shape(cfg->Nx, cfg->Ny, cfg->Nz);
		ctx->shape_sc.buf = ctx->k->stream_collide_pair_buffer && N * cfg->velocity_set * eb <= 0xFFFFFFFFull;
	} else if(ctx->k->stream_collide_buffer && N * cfg->velocity_set * eb <= 0xFFFFFFFFull) {
		ctx->shape_sc.buf = true; // the whole DDF array below 4 GiB: 32-bit buffer offsets
	} else if(ctx->k->stream_collide_offsets32 && N * eb <= 0xFFFFFFFFull) {
		ctx->shape_sc.off32 = true; // every slot below 4 GiB: 32-bit byte offsets
	}
	ctx->shape_sc.alternate = N * cfg->velocity_set * eb >= fx::ALTERNATE_ORDER_BYTES;
	ctx->shape_adj.alternate = ctx->shape_sc.alternate;
	const uint32_t cap = ctx->k->stream_collide_waves_cap;
	hipError_t e = hipSetDevice(device);
	if(e == hipSuccess) {
		hipDeviceProp_t prop;
		e = hipGetDeviceProperties(&prop, device);
		if(e == hipSuccess && cap)
			ctx->shape_sc.lds = fx::occupancy_lds(cap, ctx->shape_sc.block.x, prop.maxSharedMemoryPerMultiProcessor, prop.sharedMemPerBlock);
	}
	if(e == hipSuccess) {
		size_t free_b = 0, total_b = 0;
		e = hipMemGetInfo(&free_b, &total_b);
		if(e == hipSuccess && N * fx_memory_bytes_per_cell(cfg) > (uint64_t)free_b) {
			delete ctx;
			return fail(FX_ERR_MEMORY, "Grid resolution is too large for the free device memory");
		}
	}
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipEventCreate(&ctx->ev0);
	if(e == hipSuccess) e = hipEventCreate(&ctx->ev1);
	if(e == hipSuccess) e = hipMalloc(&ctx->result, 4 * sizeof(double));
	if(e != hipSuccess) {
		fx_ctx_destroy(ctx);
		return fail_hip(e, "fx_ctx_create");
	}
	if(resolve_scratch(device, &ctx->scratch) != FX_OK) { // device is current here
		const std::string msg = g_last_error;
		fx_ctx_destroy(ctx);
		return fail(FX_ERR_HIP, msg);
	}
	*out = ctx;
	return FX_OK;
}

int fx_ctx_destroy(fx_ctx* ctx) {
	if(!ctx) return FX_OK;
	(void)hipSetDevice(ctx->device);
	if(ctx->stream) (void)hipStreamSynchronize(ctx->stream);
	for(void* p : ctx->buffers) (void)hipFree(p);
	if(ctx->partial) (void)hipFree(ctx->partial);
	if(ctx->result) (void)hipFree(ctx->result);
	if(ctx->ev0) (void)hipEventDestroy(ctx->ev0);
	if(ctx->ev1) (void)hipEventDestroy(ctx->ev1);
	if(ctx->stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
	return FX_OK;
}

int fx_ctx_config(const fx_ctx* ctx, fx_lbm_config* out) {
	if(!ctx || !out) return fail(FX_ERR_ARG, "null argument");
	*out = ctx->cfg;
	return FX_OK;
}

int fx_sync(fx_ctx* ctx) {
	if(!ctx) return fail(FX_ERR_ARG, "null context");
	FX_HIP(hipStreamSynchronize(ctx->stream));
	return FX_OK;
}

int fx_buf_alloc(fx_ctx* ctx, uint64_t bytes, void** dptr) {
	if(!ctx || !dptr) return fail(FX_ERR_ARG, "null argument");
	*dptr = nullptr;
	if(bytes == 0ull) return fail(FX_ERR_ARG, "Memory size must be larger than 0");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	void* p = nullptr;
	const hipError_t e = hipMalloc(&p, bytes);
	if(e != hipSuccess) return fail(FX_ERR_MEMORY, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
	ctx->buffers.insert(p);
	*dptr = p;
	return FX_OK;
}

int fx_buf_free(fx_ctx* ctx, void* dptr) {
	if(!ctx) return fail(FX_ERR_ARG, "null context");
	if(!dptr) return FX_OK;
	auto it = ctx->buffers.find(dptr);
	if(it == ctx->buffers.end()) return fail(FX_ERR_ARG, "pointer was not allocated by this context");
	FX_HIP(hipStreamSynchronize(ctx->stream));
	FX_HIP(hipFree(dptr));
	ctx->buffers.erase(it);
	return FX_OK;
}

int fx_buf_write(fx_ctx* ctx, void* dst, uint64_t offset, uint64_t bytes, const void* host, int blocking) {
	if(!ctx || !dst || (!host && bytes)) return fail(FX_ERR_ARG, "null argument");
	FX_HIP(hipMemcpyAsync((char*)dst + offset, host, bytes, hipMemcpyHostToDevice, ctx->stream));
	if(blocking) FX_HIP(hipStreamSynchronize(ctx->stream));
	return FX_OK;
}

int fx_buf_read(fx_ctx* ctx, const void* src, uint64_t offset, uint64_t bytes, void* host, int blocking) {
	if(!ctx || !src || (!host && bytes)) return fail(FX_ERR_ARG, "null argument");
	FX_HIP(hipMemcpyAsync(host, (const char*)src + offset, bytes, hipMemcpyDeviceToHost, ctx->stream));
	if(blocking) FX_HIP(hipStreamSynchronize(ctx->stream));
	return FX_OK;
}

int fx_buf_read_box(fx_ctx* ctx, const void* src, void* host, uint32_t elem_bytes, uint64_t N, uint32_t d, uint64_t Nx, uint64_t Ny,
                    uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, uint64_t z0, uint64_t z1, int component, int blocking) {
	return box_copy(ctx, (char*)src, (char*)host, false, elem_bytes, N, d, Nx, Ny, x0, x1, y0, y1, z0, z1, component, blocking);
}

int fx_buf_write_box(fx_ctx* ctx, void* dst, const void* host, uint32_t elem_bytes, uint64_t N, uint32_t d, uint64_t Nx, uint64_t Ny,
                     uint64_t x0, uint64_t x1, uint64_t y0, uint64_t y1, uint64_t z0, uint64_t z1, int component, int blocking) {
	return box_copy(ctx, (char*)dst, (char*)host, true, elem_bytes, N, d, Nx, Ny, x0, x1, y0, y1, z0, z1, component, blocking);
}

int fx_buf_fill(fx_ctx* ctx, void* dst, uint64_t bytes, uint8_t value) {
	if(!ctx || !dst) return fail(FX_ERR_ARG, "null argument");
	FX_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
	return FX_OK;
}

int fx_buf_copy(fx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
	if(!ctx || !dst || !src) return fail(FX_ERR_ARG, "null argument");
	FX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
	return FX_OK;
}

int fx_lbm_initialize(fx_ctx* ctx, void* fi, const float* rho, float* u, uint8_t* flags) {
	if(!ctx) return fail(FX_ERR_ARG, "null context");
	if(!fi || !rho || !u || !flags) return fail(FX_ERR_ARG, "fi, rho, u and flags must be device pointers");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx::LbmArgs a = make_args(ctx, fi, rho, u, flags, 1ull, 0.0f, 0.0f, 0.0f, nullptr, nullptr);
	ctx->k->initialize(a, ctx->shape, ctx->stream);
	return check_launch("initialize");
}

int fx_lbm_update_moving_boundaries(fx_ctx* ctx, const float* u, uint8_t* flags) {
	if(!ctx) return fail(FX_ERR_ARG, "null context");
	if(!u || !flags) return fail(FX_ERR_ARG, "u and flags must be device pointers");
	if(!(ctx->cfg.extensions & FX_MOVING_BOUNDARIES)) return fail(FX_ERR_CONFIG, "MOVING_BOUNDARIES is not enabled");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx::LbmArgs a = make_args(ctx, nullptr, nullptr, u, flags, 0ull, 0.0f, 0.0f, 0.0f, nullptr, nullptr);
	ctx->k->update_moving_boundaries(a, ctx->shape, ctx->stream);
	return check_launch("update_moving_boundaries");
}

int fx_lbm_stream_collide(fx_ctx* ctx, void* fi, float* rho, float* u, const uint8_t* flags, uint64_t t, float fx_, float fy_, float fz_,
                          const float* F, const float* alpha) {
	const int rc = check_step_args(ctx, fi, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx::LbmArgs a = make_args(ctx, fi, rho, u, flags, t, fx_, fy_, fz_, F, alpha);
	ctx->k->stream_collide(a, ctx->shape_sc, ctx->stream);
	return check_launch("stream_collide");
}

int fx_lbm_run(fx_ctx* ctx, void* fi, float* rho, float* u, const uint8_t* flags, uint64_t t0, uint64_t steps, float fx_, float fy_, float fz_,
               const float* F, const float* alpha, float* ms) {
	const int rc = check_step_args(ctx, fi, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	FX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
	fx::LbmArgs a = make_args(ctx, fi, rho, u, flags, t0, fx_, fy_, fz_, F, alpha);
	for(uint64_t s = 0; s < steps; s++) {
		a.odd = (uint32_t)((t0 + s) & 1ull);
		ctx->k->stream_collide(a, ctx->shape_sc, ctx->stream);
	}
	const int lc = check_launch("stream_collide");
	if(lc != FX_OK) return lc;
	FX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
	if(ms) {
		FX_HIP(hipEventSynchronize(ctx->ev1));
		FX_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
	}
	return FX_OK;
}

int fx_lbm_update_fields(fx_ctx* ctx, const void* fi, float* rho, float* u, const uint8_t* flags, uint64_t t, float fx_, float fy_, float fz_,
                         const float* F, const float* alpha) {
	const int rc = check_step_args(ctx, fi, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if(!rho || !u) return fail(FX_ERR_ARG, "rho and u must be device pointers");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx::LbmArgs a = make_args(ctx, fi, rho, u, flags, t, fx_, fy_, fz_, F, alpha);
	ctx->k->update_fields(a, ctx->shape, ctx->stream);
	return check_launch("update_fields");
}

int fx_lbm_update_force_field(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, float* F) {
	if(!ctx) return fail(FX_ERR_ARG, "null context");
	if(!fi || !flags || !F) return fail(FX_ERR_ARG, "fi, flags and F must be device pointers");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx::LbmArgs a = make_args(ctx, fi, nullptr, nullptr, flags, t, 0.0f, 0.0f, 0.0f, F, nullptr);
	ctx->k->update_force_field(a, ctx->shape, ctx->stream);
	return check_launch("update_force_field");
}

int fx_lbm_reset_force_field(fx_ctx* ctx, float* F) {
	if(!ctx || !F) return fail(FX_ERR_ARG, "null argument");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const uint64_t n3 = 3ull * ctx->N;
	fx::k_reset3<<<fx::stride_grid(n3, 256u), 256, 0, ctx->stream>>>(F, n3);
	return check_launch("reset_force_field");
}

static int object_sum(fx_ctx* ctx, const float* F, const uint8_t* flags, uint8_t marker, int mode, float cx, float cy, float cz, double* out, int count) {
	if(!ctx || !flags || !out || (mode != 2 && !F)) return fail(FX_ERR_ARG, "null argument");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	int rc = ensure_partial(ctx, (uint64_t)fx::RED_GRID * 4u);
	if(rc != FX_OK) return rc;
	fx::k_object_partial<<<fx::RED_GRID, fx::RED_BLOCK, 0, ctx->stream>>>(F, flags, marker, ctx->N, ctx->cfg.Nx, ctx->cfg.Ny, ctx->cfg.Nz, mode, cx, cy, cz, ctx->partial);
	fx::k_reduce_partials<<<1, fx::RED_BLOCK, 0, ctx->stream>>>(ctx->partial, fx::RED_GRID, 4u, ctx->result);
	rc = check_launch("object sum");
	if(rc != FX_OK) return rc;
	double h[4];
	FX_HIP(hipMemcpyAsync(h, ctx->result, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
	FX_HIP(hipStreamSynchronize(ctx->stream));
	for(int k = 0; k < count; k++) out[k] = h[k];
	return FX_OK;
}

int fx_lbm_object_force(fx_ctx* ctx, const float* F, const uint8_t* flags, uint8_t marker, double out[3]) {
	return object_sum(ctx, F, flags, marker, 0, 0.0f, 0.0f, 0.0f, out, 3);
}
int fx_lbm_object_torque(fx_ctx* ctx, const float* F, const uint8_t* flags, uint8_t marker, float cx, float cy, float cz, double out[3]) {
	return object_sum(ctx, F, flags, marker, 1, cx, cy, cz, out, 3);
}
int fx_lbm_object_center_of_mass(fx_ctx* ctx, const uint8_t* flags, uint8_t marker, double out[4]) {
	return object_sum(ctx, nullptr, flags, marker, 2, 0.0f, 0.0f, 0.0f, out, 4);
}

// One partial per block of the lattice grid, then one fixed-order sum into result[0]; every
// partial the sum reads is written by this call's partial kernel, so nothing needs clearing.
static int mass_sum_with(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, double* partial, double* result, double* out) {
	const uint64_t blocks = (uint64_t)ctx->shape.grid.x * ctx->shape.grid.y * ctx->shape.grid.z;
	const fx::LbmArgs a = make_args(ctx, fi, nullptr, nullptr, flags, t, 0.0f, 0.0f, 0.0f, nullptr, nullptr);
	ctx->k->mass_partial(a, ctx->shape, ctx->stream, partial);
	fx::k_reduce_partials<<<1, fx::RED_BLOCK, 0, ctx->stream>>>(partial, blocks, 1u, result);
	const int rc = check_launch("mass sum");
	if(rc != FX_OK) return rc;
	FX_HIP(hipMemcpyAsync(out, result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
	FX_HIP(hipStreamSynchronize(ctx->stream));
	return FX_OK;
}

int fx_lbm_mass_sum(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, double* out) {
	if(!ctx || !fi || !flags || !out) return fail(FX_ERR_ARG, "null argument");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const uint64_t blocks = (uint64_t)ctx->shape.grid.x * ctx->shape.grid.y * ctx->shape.grid.z;
	const int rc = ensure_partial(ctx, blocks);
	if(rc != FX_OK) return rc;
	return mass_sum_with(ctx, fi, flags, t, ctx->partial, ctx->result, out);
}

// The same sum over the static scratch array: no allocation, no clearing needed (every
// partial the sum reads and the result are written by this call's kernels).  Calls that
// overlap on one device (contexts on different streams, or threads) race on the array.
int fx_lbm_mass_sum_noalloc(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, double* out) {
	if(!ctx || !fi || !flags || !out) return fail(FX_ERR_ARG, "null argument");
	if(!ctx->scratch) return fail(FX_ERR_STATE, "the scratch array's device address is not resolved");
	const uint64_t blocks = (uint64_t)ctx->shape.grid.x * ctx->shape.grid.y * ctx->shape.grid.z;
	if(blocks > NOALLOC_MASS_PARTIALS)
		return fail(FX_ERR_CONFIG, "fx_lbm_mass_sum_noalloc: the lattice needs " + std::to_string(blocks) + " partial sums, the static scratch holds " +
		                           std::to_string(NOALLOC_MASS_PARTIALS) + "; use fx_lbm_mass_sum");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	return mass_sum_with(ctx, fi, flags, t, ctx->scratch, ctx->scratch + NOALLOC_MASS_PARTIALS, out);
}

static uint64_t ddf_bytes(const fx_ctx* ctx) {
	return ctx->N * (uint64_t)ctx->cfg.velocity_set * (ctx->cfg.storage == FX_FP32 ? 4u : 2u);
}
static int adjoint_layout(const fx_ctx* ctx, const fx_adjoint_config* cfg, uint64_t& k, uint64_t& checkpoints, uint64_t& bytes) {
	if(!ctx || !cfg) return fail(FX_ERR_ARG, "null argument");
	if(ctx->cfg.extensions & (FX_MOVING_BOUNDARIES | FX_SUBGRID))
		return fail(FX_ERR_CONFIG, "the adjoint does not differentiate MOVING_BOUNDARIES or SUBGRID");
	if(cfg->steps == 0u || cfg->window_begin >= cfg->steps) return fail(FX_ERR_ARG, "need 0 <= window_begin < steps");
	if(cfg->plane_x >= ctx->cfg.Nx) return fail(FX_ERR_ARG, "plane_x out of range");
	k = cfg->checkpoint_interval;
	if(k == 0u) { k = (uint64_t)std::ceil(std::sqrt((double)cfg->steps)); if(k == 0u) k = 1u; }
	if(k > cfg->steps) k = cfg->steps;
	checkpoints = (cfg->steps + k - 1u) / k;
	const uint64_t N = ctx->N, q = ctx->cfg.velocity_set;
	const uint64_t obj_blocks = ((uint64_t)ctx->cfg.Ny * ctx->cfg.Nz + 255u) / 256u;
	bytes = checkpoints * ddf_bytes(ctx) + ddf_bytes(ctx) + k * 4u * N * 4u + q * N * 4u + (obj_blocks + 1u) * 8u;
	return FX_OK;
}

int fx_adjoint_memory(const fx_ctx* ctx, const fx_adjoint_config* cfg, uint64_t* bytes) {
	uint64_t k, c, b;
	const int rc = adjoint_layout(ctx, cfg, k, c, b);
	if(rc != FX_OK) return rc;
	if(!bytes) return fail(FX_ERR_ARG, "null output");
	*bytes = b;
	return FX_OK;
}

int fx_adjoint_create(fx_ctx* ctx, const fx_adjoint_config* cfg, fx_adjoint** out) {
	if(!out) return fail(FX_ERR_ARG, "null output handle");
	*out = nullptr;
	uint64_t k, checkpoints, bytes;
	int rc = adjoint_layout(ctx, cfg, k, checkpoints, bytes);
	if(rc != FX_OK) return rc;
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	size_t free_b = 0, total_b = 0;
	FX_HIP(hipMemGetInfo(&free_b, &total_b));
	if(bytes > (uint64_t)free_b) return fail(FX_ERR_MEMORY, "adjoint workspace of " + std::to_string(bytes) + " bytes does not fit the free device memory");
	fx_adjoint* adj = new fx_adjoint();
	adj->ctx = ctx; adj->cfg = *cfg; adj->k = k; adj->checkpoints = checkpoints; adj->fi_bytes = ddf_bytes(ctx);
	adj->obj_blocks = ((uint64_t)ctx->cfg.Ny * ctx->cfg.Nz + 255u) / 256u;
	hipError_t e = hipMalloc(&adj->ckpt, checkpoints * adj->fi_bytes);
	if(e == hipSuccess) e = hipMalloc(&adj->work_fi, adj->fi_bytes);
	if(e == hipSuccess) e = hipMalloc(&adj->moments, k * 4u * ctx->N * sizeof(float));
	if(e == hipSuccess) e = hipMalloc(&adj->lambda, (uint64_t)ctx->cfg.velocity_set * ctx->N * sizeof(float));
	if(e == hipSuccess) e = hipMalloc(&adj->partial, adj->obj_blocks * sizeof(double));
	if(e == hipSuccess) e = hipMalloc(&adj->acc, sizeof(double));
	if(e != hipSuccess) { fx_adjoint_destroy(adj); return fail(FX_ERR_MEMORY, std::string("adjoint workspace: ") + hipGetErrorString(e)); }
	*out = adj;
	return FX_OK;
}

int fx_adjoint_destroy(fx_adjoint* adj) {
	if(!adj) return FX_OK;
	(void)hipSetDevice(adj->ctx->device);
	(void)hipStreamSynchronize(adj->ctx->stream);
	if(adj->ckpt) (void)hipFree(adj->ckpt);
	if(adj->work_fi) (void)hipFree(adj->work_fi);
	if(adj->moments) (void)hipFree(adj->moments);
	if(adj->lambda) (void)hipFree(adj->lambda);
	if(adj->partial) (void)hipFree(adj->partial);
	if(adj->acc) (void)hipFree(adj->acc);
	delete adj;
	return FX_OK;
}

int fx_lbm_stream_collide_record(fx_ctx* ctx, void* fi, float* rho, float* u, const uint8_t* flags, uint64_t t, float fx_, float fy_, float fz_,
                                 const float* F, const float* alpha, float* moments) {
	const int rc = check_step_args(ctx, fi, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if(!moments) return fail(FX_ERR_ARG, "null moments");
	if(ctx->cfg.extensions & (FX_MOVING_BOUNDARIES | FX_SUBGRID))
		return fail(FX_ERR_CONFIG, "the adjoint does not differentiate MOVING_BOUNDARIES or SUBGRID");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	ctx->k->stream_collide_record(make_args(ctx, fi, rho, u, flags, t, fx_, fy_, fz_, F, alpha), ctx->shape_adj, ctx->stream, moments);
	return check_launch("stream_collide_record");
}

static fx::AdjointArgs adjoint_args(const fx_ctx* ctx, float* lambda, const float* rho, const float* u, const uint8_t* flags, uint64_t t, float fx_, float fy_, float fz_,
                                    const float* F, const float* alpha, const float* moments, uint32_t plane_x, float obj_weight, float* grad) {
	fx::AdjointArgs g;
	g.a = make_args(ctx, lambda, rho, u, flags, t, fx_, fy_, fz_, F, alpha);
	g.moments = moments;
	g.grad = grad;
	g.plane_x = plane_x;
	g.obj_weight = obj_weight;
	return g;
}

int fx_lbm_adjoint_step(fx_ctx* ctx, float* lambda, const float* rho, const float* u, const uint8_t* flags, uint64_t t, float fx_, float fy_, float fz_,
                        const float* F, const float* alpha, const float* moments, uint32_t plane_x, float obj_weight, float* grad_alpha) {
	int rc = check_step_args(ctx, lambda, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if(!moments) return fail(FX_ERR_ARG, "null moments");
	if(ctx->cfg.extensions & (FX_MOVING_BOUNDARIES | FX_SUBGRID))
		return fail(FX_ERR_CONFIG, "the adjoint does not differentiate MOVING_BOUNDARIES or SUBGRID");
	if((ctx->cfg.extensions & FX_POROSITY) && !grad_alpha) return fail(FX_ERR_ARG, "POROSITY needs grad_alpha");
	if((ctx->cfg.extensions & FX_EQUILIBRIUM_BOUNDARIES) && (!rho || !u)) return fail(FX_ERR_ARG, "TYPE_E presets rho/u required");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	ctx->k->adjoint_step(adjoint_args(ctx, lambda, rho, u, flags, t, fx_, fy_, fz_, F, alpha, moments, plane_x, obj_weight, grad_alpha), ctx->shape_adj, ctx->stream);
	return check_launch("adjoint_step");
}

static int objective_accumulate(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, uint32_t plane_x, double* partial, uint64_t blocks, double* acc) {
	ctx->k->objective_partial(make_args(ctx, fi, nullptr, nullptr, flags, t, 0.0f, 0.0f, 0.0f, nullptr, nullptr), plane_x, ctx->stream, partial);
	fx::k_reduce_add<<<1, fx::RED_BLOCK, 0, ctx->stream>>>(partial, blocks, acc);
	return check_launch("objective");
}

int fx_lbm_objective_plane(fx_ctx* ctx, const void* fi, const uint8_t* flags, uint64_t t, uint32_t plane_x, double* out) {
	if(!ctx || !fi || !flags || !out) return fail(FX_ERR_ARG, "null argument");
	if(plane_x >= ctx->cfg.Nx) return fail(FX_ERR_ARG, "plane_x out of range");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const uint64_t blocks = ((uint64_t)ctx->cfg.Ny * ctx->cfg.Nz + 255u) / 256u;
	int rc = ensure_partial(ctx, blocks);
	if(rc != FX_OK) return rc;
	FX_HIP(hipMemsetAsync(ctx->result, 0, sizeof(double), ctx->stream));
	rc = objective_accumulate(ctx, fi, flags, t, plane_x, ctx->partial, blocks, ctx->result);
	if(rc != FX_OK) return rc;
	FX_HIP(hipMemcpyAsync(out, ctx->result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
	FX_HIP(hipStreamSynchronize(ctx->stream));
	return FX_OK;
}

int fx_lbm_forward(fx_ctx* ctx, fx_adjoint* adj, void* fi, float* rho, float* u, const uint8_t* flags, float fx_, float fy_, float fz_,
                   const float* F, const float* alpha, double* objective, float* ms) {
	int rc = check_step_args(ctx, fi, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if(!adj || adj->ctx != ctx) return fail(FX_ERR_ARG, "adjoint handle does not belong to this context");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx_adjoint_config& c = adj->cfg;
	FX_HIP(hipMemsetAsync(adj->acc, 0, sizeof(double), ctx->stream));
	FX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
	fx::LbmArgs a = make_args(ctx, fi, rho, u, flags, 0u, fx_, fy_, fz_, F, alpha);
	for(uint64_t t = 0; t < c.steps; t++) {
		if(t % adj->k == 0u) FX_HIP(hipMemcpyAsync(adj->ckpt + (t / adj->k) * adj->fi_bytes, fi, adj->fi_bytes, hipMemcpyDeviceToDevice, ctx->stream));
		if(t >= c.window_begin) {
			rc = objective_accumulate(ctx, fi, flags, t, c.plane_x, adj->partial, adj->obj_blocks, adj->acc);
			if(rc != FX_OK) return rc;
		}
		a.odd = (uint32_t)(t & 1u);
		ctx->k->stream_collide(a, ctx->shape_sc, ctx->stream);
	}
	rc = check_launch("forward");
	if(rc != FX_OK) return rc;
	FX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
	double sum = 0.0;
	FX_HIP(hipMemcpyAsync(&sum, adj->acc, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
	FX_HIP(hipStreamSynchronize(ctx->stream));
	if(ms) FX_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
	if(objective) *objective = sum / (double)(c.steps - c.window_begin);
	adj->forward_done = true;
	return FX_OK;
}

int fx_lbm_backward(fx_ctx* ctx, fx_adjoint* adj, const float* rho, const float* u, const uint8_t* flags, float fx_, float fy_, float fz_,
                    const float* F, const float* alpha, float* grad_alpha, float* ms) {
	if(!adj || adj->ctx != ctx) return fail(FX_ERR_ARG, "adjoint handle does not belong to this context");
	if(!adj->forward_done) return fail(FX_ERR_STATE, "fx_lbm_forward must run before fx_lbm_backward");
	int rc = check_step_args(ctx, adj->work_fi, rho, u, flags, F, alpha);
	if(rc != FX_OK) return rc;
	if((ctx->cfg.extensions & FX_POROSITY) && !grad_alpha) return fail(FX_ERR_ARG, "POROSITY needs grad_alpha");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const fx_adjoint_config& c = adj->cfg;
	const uint64_t N = ctx->N;
	// rho/u are only read (TYPE_E presets); UPDATE_FIELDS writes of the recomputation go to scratch
	float* rho_w = const_cast<float*>(rho);
	float* u_w = const_cast<float*>(u);
	FX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
	FX_HIP(hipMemsetAsync(adj->lambda, 0, (uint64_t)ctx->cfg.velocity_set * N * sizeof(float), ctx->stream));
	if(grad_alpha) FX_HIP(hipMemsetAsync(grad_alpha, 0, N * sizeof(float), ctx->stream));
	const float wobj = (float)(1.0 / (double)(c.steps - c.window_begin));
	fx::LbmArgs a = make_args(ctx, adj->work_fi, rho_w, u_w, flags, 0u, fx_, fy_, fz_, F, alpha);
	a.ext &= ~FX_UPDATE_FIELDS; // the recomputation must not overwrite rho/u
	for(uint64_t s = adj->checkpoints; s-- > 0u;) {
		const uint64_t ts = s * adj->k, te = std::min<uint64_t>(c.steps, ts + adj->k);
		FX_HIP(hipMemcpyAsync(adj->work_fi, adj->ckpt + s * adj->fi_bytes, adj->fi_bytes, hipMemcpyDeviceToDevice, ctx->stream));
		for(uint64_t t = ts; t < te; t++) {
			a.odd = (uint32_t)(t & 1u);
			ctx->k->stream_collide_record(a, ctx->shape_adj, ctx->stream, adj->moments + (t - ts) * 4u * N);
		}
		for(uint64_t t = te; t-- > ts;) {
			fx::AdjointArgs g = adjoint_args(ctx, adj->lambda, rho, u, flags, t, fx_, fy_, fz_, F, alpha, adj->moments + (t - ts) * 4u * N,
			                                 c.plane_x, t >= c.window_begin ? wobj : 0.0f, grad_alpha);
			g.a.ext &= ~FX_UPDATE_FIELDS;
			ctx->k->adjoint_step(g, ctx->shape_adj, ctx->stream);
		}
		rc = check_launch("backward");
		if(rc != FX_OK) return rc;
	}
	FX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
	FX_HIP(hipEventSynchronize(ctx->ev1));
	if(ms) FX_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
	return FX_OK;
}

static uint64_t face_area(const fx_lbm_config* cfg, uint32_t direction) {
	return direction == 0u ? (uint64_t)cfg->Ny * cfg->Nz : direction == 1u ? (uint64_t)cfg->Nx * cfg->Nz : (uint64_t)cfg->Nx * cfg->Ny;
}

uint64_t fx_transfer_bytes(const fx_lbm_config* cfg, uint32_t field, uint32_t direction) {
	if(!cfg || direction > 2u || !valid_enums(cfg)) return 0u;
	const uint64_t A = face_area(cfg, direction);
	switch(field) {
		case FX_FIELD_FI: return A * (cfg->velocity_set == FX_D3Q19 ? 5u : 9u) * (cfg->storage == FX_FP32 ? 4u : 2u);
		case FX_FIELD_RHO_U_FLAGS: return A * 17u;
		case FX_FIELD_FLAGS: return A;
		case FX_FIELD_F: return A * 12u;
		default: return 0u;
	}
}

static int transfer(fx_ctx* ctx, uint32_t field, uint32_t direction, uint64_t t, void* bp, void* bm, void* fi, float* rho, float* u,
                    uint8_t* flags, float* F, bool insert) {
	if(!ctx || !bp || !bm) return fail(FX_ERR_ARG, "null argument");
	if(direction > 2u) return fail(FX_ERR_ARG, "direction must be 0 (x), 1 (y) or 2 (z)");
	const fx_lbm_config& c = ctx->cfg;
	const uint32_t Nax = direction == 0u ? c.Nx : direction == 1u ? c.Ny : c.Nz;
	if(Nax < 3u) return fail(FX_ERR_CONFIG, "a split axis needs at least 3 cells (two halo layers and an interior)");
	const uint64_t A = face_area(&c, direction);
	if(A >= 0xFFFFFFFFull) return fail(FX_ERR_CONFIG, "face too large");
	if(field == FX_FIELD_FI && !fi) return fail(FX_ERR_ARG, "null fi");
	if(field == FX_FIELD_RHO_U_FLAGS && (!rho || !u || !flags)) return fail(FX_ERR_ARG, "null rho/u/flags");
	if(field == FX_FIELD_FLAGS && !flags) return fail(FX_ERR_ARG, "null flags");
	if(field == FX_FIELD_F && !F) return fail(FX_ERR_ARG, "null F");
	if(field > FX_FIELD_F) return fail(FX_ERR_ARG, "unknown field");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	if(field == FX_FIELD_FI) {
		ctx->k->transfer_fi(fi, bp, bm, c.Nx, c.Ny, c.Nz, direction, insert, (uint32_t)(t & 1ull), ctx->stream);
	} else {
		const unsigned blocks = (unsigned)((A + 255u) / 256u);
		char* p = (char*)bp; char* m = (char*)bm;
		if(direction == 0u) { if(insert) k_transfer_field<0, true><<<blocks, 256, 0, ctx->stream>>>(field, p, m, rho, u, flags, F, c.Nx, c.Ny, c.Nz);
		                      else k_transfer_field<0, false><<<blocks, 256, 0, ctx->stream>>>(field, p, m, rho, u, flags, F, c.Nx, c.Ny, c.Nz); }
		else if(direction == 1u) { if(insert) k_transfer_field<1, true><<<blocks, 256, 0, ctx->stream>>>(field, p, m, rho, u, flags, F, c.Nx, c.Ny, c.Nz);
		                           else k_transfer_field<1, false><<<blocks, 256, 0, ctx->stream>>>(field, p, m, rho, u, flags, F, c.Nx, c.Ny, c.Nz); }
		else { if(insert) k_transfer_field<2, true><<<blocks, 256, 0, ctx->stream>>>(field, p, m, rho, u, flags, F, c.Nx, c.Ny, c.Nz);
		       else k_transfer_field<2, false><<<blocks, 256, 0, ctx->stream>>>(field, p, m, rho, u, flags, F, c.Nx, c.Ny, c.Nz); }
	}
	return check_launch(insert ? "transfer insert" : "transfer extract");
}

int fx_lbm_transfer_extract(fx_ctx* ctx, uint32_t field, uint32_t direction, uint64_t t, void* buffer_p, void* buffer_m,
                            const void* fi, const float* rho, const float* u, const uint8_t* flags, const float* F) {
	return transfer(ctx, field, direction, t, buffer_p, buffer_m, const_cast<void*>(fi), const_cast<float*>(rho), const_cast<float*>(u),
	                const_cast<uint8_t*>(flags), const_cast<float*>(F), false);
}
int fx_lbm_transfer_insert(fx_ctx* ctx, uint32_t field, uint32_t direction, uint64_t t, const void* buffer_p, const void* buffer_m,
                           void* fi, float* rho, float* u, uint8_t* flags, float* F) {
	return transfer(ctx, field, direction, t, const_cast<void*>(buffer_p), const_cast<void*>(buffer_m), fi, rho, u, flags, F, true);
}

int fx_codec_roundtrip(fx_ctx* ctx, uint32_t storage, const float* x, uint16_t* codes, float* decoded, uint64_t n) {
	if(!ctx || !x || !codes || !decoded) return fail(FX_ERR_ARG, "null argument");
	if(storage != FX_FP16S && storage != FX_FP16C) return fail(FX_ERR_ARG, "storage must be FX_FP16S or FX_FP16C");
	if(n == 0) return FX_OK;
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const unsigned blocks = fx::stride_grid(n, 256u);
	if(storage == FX_FP16S) fx::k_codec<fx::ST_FP16S><<<blocks, 256, 0, ctx->stream>>>(x, codes, decoded, n);
	else fx::k_codec<fx::ST_FP16C><<<blocks, 256, 0, ctx->stream>>>(x, codes, decoded, n);
	return check_launch("codec");
}

int fx_bench_copy(fx_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint32_t reps, float* ms) {
	if(!ctx || !dst || !src) return fail(FX_ERR_ARG, "null argument");
	if(bytes % 16u != 0u || bytes == 0u) return fail(FX_ERR_ARG, "bytes must be a positive multiple of 16");
	if(set_device(ctx) != FX_OK) return FX_ERR_HIP;
	const uint64_t n16 = bytes / 16u;
	const unsigned blocks = fx::stride_grid(n16, 256u);
	FX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
	for(uint32_t r = 0; r < reps; r++) fx::k_copy16<<<(unsigned)blocks, 256, 0, ctx->stream>>>((const fx::copy_vec*)src, (fx::copy_vec*)dst, n16);
	const int rc = check_launch("copy");
	if(rc != FX_OK) return rc;
	FX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
	FX_HIP(hipEventSynchronize(ctx->ev1));
	float t = 0.0f;
	FX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
	if(ms) *ms = t;
	return FX_OK;
}

} // extern "C"
