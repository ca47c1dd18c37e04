// Launch table of the precompiled kernel variants.  One table per (velocity set,
// storage, collision); filled by the per-velocity-set translation units.
#pragma once
#include "lbm_device.hpp"

namespace fx {

struct LaunchShape {
	dim3 grid, block;
	size_t lds = 0; // dynamic LDS reserved per block (occupancy limit), see occupancy_lds()
};
// Block of bx consecutive x cells of one row (bx = 64, 128 or 256); grid over (x chunk, y, z).
inline LaunchShape lattice_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz) {
	const uint32_t bx = Nx <= 64u ? 64u : Nx <= 128u ? 128u : 256u;
	LaunchShape s;
	s.block = dim3(bx, 1, 1);
	s.grid = dim3((Nx + bx - 1u) / bx, Ny, Nz);
	return s;
}

// Dynamic LDS per block that caps residency at `waves_per_simd` waves per SIMD (4 SIMDs
// per CU), for a kernel that uses no LDS itself.  0 when the cap would not bind.
inline size_t occupancy_lds(uint32_t waves_per_simd, uint32_t block_threads, size_t lds_per_cu, size_t lds_per_block_max) {
	const uint32_t waves_per_block = (block_threads + 63u) / 64u;
	const uint32_t blocks = (waves_per_simd * 4u) / waves_per_block;
	if(blocks == 0u || lds_per_cu == 0u) return 0;
	size_t lds = (lds_per_cu / blocks) & ~(size_t)255;
	while(lds > 0 && lds_per_cu / lds < blocks) lds -= 256; // at least `blocks` fit
	return lds <= lds_per_block_max ? lds : 0;
}

typedef void (*lattice_fn)(const LbmArgs&, const LaunchShape&, hipStream_t);
typedef void (*mass_fn)(const LbmArgs&, const LaunchShape&, hipStream_t, double*);
typedef void (*record_fn)(const LbmArgs&, const LaunchShape&, hipStream_t, float*);
typedef void (*adjoint_fn)(const AdjointArgs&, const LaunchShape&, hipStream_t);
typedef void (*objective_fn)(const LbmArgs&, uint32_t, hipStream_t, double*);

struct LbmKernels {
	lattice_fn initialize;
	lattice_fn stream_collide;
	lattice_fn update_fields;
	lattice_fn update_force_field;
	mass_fn mass_partial;
	uint32_t stream_collide_waves_cap;
	record_fn stream_collide_record;
	adjoint_fn adjoint_step;
	objective_fn objective_partial; // grid of ceil(Ny*Nz/256) blocks, one partial each
};

const LbmKernels* kernels_q19(uint32_t storage, bool trt, bool plain);
const LbmKernels* kernels_q27(uint32_t storage, bool trt, bool plain);

constexpr uint32_t EXT_ALL = EXT_VOLUME_FORCE | EXT_FORCE_FIELD | EXT_EQUILIBRIUM_BOUNDARIES | EXT_UPDATE_FIELDS | EXT_POROSITY | EXT_PRECISE_DENSITY;

// Kernel set of one (velocity set, storage, collision) variant; EXT = 0 instantiates the
// extension-free kernels, EXT_ALL the ones that read the runtime extension mask.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT> struct KernelSet {
	static void initialize(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_initialize<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void stream_collide(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_stream_collide<Q, ST, TRT, EXT><<<s.grid, s.block, s.lds, st>>>(a); }
	static void update_fields(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_fields<Q, ST, EXT><<<s.grid, s.block, 0, st>>>(a); }
	static void update_force_field(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_force_field<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void mass_partial(const LbmArgs& a, const LaunchShape& s, hipStream_t st, double* p) { k_mass_partial<Q, ST><<<s.grid, s.block, 0, st>>>(a, p); }
	static void stream_collide_record(const LbmArgs& a, const LaunchShape& s, hipStream_t st, float* m) { k_stream_collide_record<Q, ST, TRT><<<s.grid, s.block, 0, st>>>(a, m); }
	static void adjoint_step(const AdjointArgs& g, const LaunchShape& s, hipStream_t st) { k_adjoint_step<Q, TRT><<<s.grid, s.block, 0, st>>>(g); }
	static void objective_partial(const LbmArgs& a, uint32_t plane_x, hipStream_t st, double* p) {
		const uint64_t cells = (uint64_t)a.Ny * a.Nz;
		k_objective_partial<Q, ST><<<(unsigned)((cells + 255u) / 256u), 256, 0, st>>>(a, plane_x, p);
	}
	// Residency cap for stream_collide, in waves per SIMD (0 = none).  Measured on MI355X
	// at 256^3: the plain FP32 kernel streams fastest at 3 waves/SIMD (fewer concurrent
	// DDF streams); the 2-byte formats need full occupancy to keep enough bytes in flight.
	static constexpr uint32_t waves_cap = (EXT == 0u && ST == ST_FP32) ? 3u : 0u;
	static const LbmKernels* table() {
		static const LbmKernels k = {initialize, stream_collide, update_fields, update_force_field, mass_partial, waves_cap,
		                             stream_collide_record, adjoint_step, objective_partial};
		return &k;
	}
};

template<int Q, uint32_t EXT> const LbmKernels* select_kernels_ext(uint32_t storage, bool trt) {
	switch(storage * 2u + (trt ? 1u : 0u)) {
		case ST_FP32 * 2u: return KernelSet<Q, ST_FP32, false, EXT>::table();
		case ST_FP32 * 2u + 1u: return KernelSet<Q, ST_FP32, true, EXT>::table();
		case ST_FP16S * 2u: return KernelSet<Q, ST_FP16S, false, EXT>::table();
		case ST_FP16S * 2u + 1u: return KernelSet<Q, ST_FP16S, true, EXT>::table();
		case ST_FP16C * 2u: return KernelSet<Q, ST_FP16C, false, EXT>::table();
		case ST_FP16C * 2u + 1u: return KernelSet<Q, ST_FP16C, true, EXT>::table();
		default: return nullptr;
	}
}
template<int Q> const LbmKernels* select_kernels(uint32_t storage, bool trt, bool plain) {
	return plain ? select_kernels_ext<Q, 0u>(storage, trt) : select_kernels_ext<Q, EXT_ALL>(storage, trt);
}

} // namespace fx
