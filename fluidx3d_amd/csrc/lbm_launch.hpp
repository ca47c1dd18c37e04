// Launch table of the precompiled kernel variants.  One table per (velocity set,
// storage, collision); filled by the per-velocity-set translation units.
#pragma once
#include "lbm_device.hpp"

namespace fx {

struct LaunchShape {
	dim3 grid, block;
	size_t lds = 0; // dynamic LDS reserved per block (occupancy limit), see occupancy_lds()
};
// Block of bx lanes (bx = 64, 128 or 256) over bx*cells consecutive x cells of one row;
// grid over (x chunk, y, z).  bx is the smallest width that covers the row, at most 256.
inline LaunchShape lattice_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t cells = 1u) {
	const uint32_t lanes = (Nx + cells - 1u) / cells;
	const uint32_t bx = lanes <= 64u ? 64u : lanes <= 128u ? 128u : 256u;
	LaunchShape s;
	s.block = dim3(bx, 1, 1);
	s.grid = dim3((Nx + bx * cells - 1u) / (bx * cells), Ny, Nz);
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
	uint32_t stream_collide_cells;  // cells per lane of stream_collide (its launch shape)
};

constexpr uint32_t EXT_ALL = EXT_VOLUME_FORCE | EXT_FORCE_FIELD | EXT_EQUILIBRIUM_BOUNDARIES | EXT_UPDATE_FIELDS | EXT_POROSITY | EXT_PRECISE_DENSITY;

// Measured launch tuning (see KernelSet).  The plain FP32 kernel streams fastest with
// fewer concurrent DDF streams (3 waves/SIMD); the 2-byte formats need full occupancy;
// D3Q19 FP16C, the most ALU-heavy variant, gains from two cells per lane.
constexpr uint32_t tune_cells(int Q, uint32_t ST, uint32_t EXT) {
	return (ST == ST_FP16C && Q == 19) ? 2u : 1u;
}
constexpr uint32_t tune_waves_cap(int Q, uint32_t ST, uint32_t EXT) {
	return (EXT == 0u && ST == ST_FP32) ? 3u : 0u;
}

// Kernel set of one (velocity set, storage, collision) variant; EXT = 0 instantiates the
// extension-free kernels, EXT_ALL the ones that read the runtime extension mask.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT> struct KernelSet {
	static void initialize(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_initialize<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void stream_collide(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_stream_collide<Q, ST, TRT, EXT, cells><<<s.grid, s.block, s.lds, st>>>(a); }
	static void update_fields(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_fields<Q, ST, (EXT == 0u ? 0u : EXT_ALL)><<<s.grid, s.block, 0, st>>>(a); }
	static void update_force_field(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_force_field<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void mass_partial(const LbmArgs& a, const LaunchShape& s, hipStream_t st, double* p) { k_mass_partial<Q, ST><<<s.grid, s.block, 0, st>>>(a, p); }
	static void stream_collide_record(const LbmArgs& a, const LaunchShape& s, hipStream_t st, float* m) { k_stream_collide_record<Q, ST, TRT><<<s.grid, s.block, 0, st>>>(a, m); }
	static void adjoint_step(const AdjointArgs& g, const LaunchShape& s, hipStream_t st) { k_adjoint_step<Q, TRT><<<s.grid, s.block, 0, st>>>(g); }
	static void objective_partial(const LbmArgs& a, uint32_t plane_x, hipStream_t st, double* p) {
		const uint64_t cells = (uint64_t)a.Ny * a.Nz;
		k_objective_partial<Q, ST><<<(unsigned)((cells + 255u) / 256u), 256, 0, st>>>(a, plane_x, p);
	}
	// Launch tuning of stream_collide, measured on MI355X at 256^3 (D3Q19 and D3Q27, SRT,
	// every storage; sweep of cells per lane x block width x residency cap):
	// cells per lane, and a residency cap in waves per SIMD (0 = none).
	static constexpr uint32_t cells = tune_cells(Q, ST, EXT);
	static constexpr uint32_t waves_cap = tune_waves_cap(Q, ST, EXT);
	static const LbmKernels* table() {
		static const LbmKernels k = {initialize, stream_collide, update_fields, update_force_field, mass_partial, waves_cap,
		                             stream_collide_record, adjoint_step, objective_partial, cells};
		return &k;
	}
};

// Extension subsets with a stream_collide of their own, tried in this order; the first
// that covers the requested extensions is used (EXT_ALL covers every combination).
// Compiling the unused paths out saves registers: measured on MI355X at 256^3, the
// specialised kernel streams 3-13 % faster than the generic one.
template<int Q, uint32_t ST, uint32_t EXT> const LbmKernels* pick(bool trt) {
	return trt ? KernelSet<Q, ST, true, EXT>::table() : KernelSet<Q, ST, false, EXT>::table();
}
template<int Q, uint32_t ST> const LbmKernels* kernels_for(bool trt, uint32_t ext) {
	constexpr uint32_t E = EXT_EQUILIBRIUM_BOUNDARIES, VF = EXT_VOLUME_FORCE, FF = EXT_FORCE_FIELD;
	constexpr uint32_t POROUS = EXT_VOLUME_FORCE | EXT_POROSITY | EXT_PRECISE_DENSITY;
	if(ext == 0u) return pick<Q, ST, 0u>(trt);
	if((ext & ~E) == 0u) return pick<Q, ST, E>(trt);
	if((ext & ~VF) == 0u) return pick<Q, ST, VF>(trt);
	if((ext & ~(E | FF)) == 0u) return pick<Q, ST, E | FF>(trt);
	if((ext & ~POROUS) == 0u) return pick<Q, ST, POROUS>(trt);
	return pick<Q, ST, EXT_ALL>(trt);
}
// One translation unit per (velocity set, storage) instantiates each table.
extern template const LbmKernels* kernels_for<19, ST_FP32>(bool, uint32_t);
extern template const LbmKernels* kernels_for<19, ST_FP16S>(bool, uint32_t);
extern template const LbmKernels* kernels_for<19, ST_FP16C>(bool, uint32_t);
extern template const LbmKernels* kernels_for<27, ST_FP32>(bool, uint32_t);
extern template const LbmKernels* kernels_for<27, ST_FP16S>(bool, uint32_t);
extern template const LbmKernels* kernels_for<27, ST_FP16C>(bool, uint32_t);

inline const LbmKernels* select_kernels(uint32_t q, uint32_t storage, bool trt, uint32_t ext) {
	switch(q * 4u + storage) {
		case 19u * 4u + ST_FP32: return kernels_for<19, ST_FP32>(trt, ext);
		case 19u * 4u + ST_FP16S: return kernels_for<19, ST_FP16S>(trt, ext);
		case 19u * 4u + ST_FP16C: return kernels_for<19, ST_FP16C>(trt, ext);
		case 27u * 4u + ST_FP32: return kernels_for<27, ST_FP32>(trt, ext);
		case 27u * 4u + ST_FP16S: return kernels_for<27, ST_FP16S>(trt, ext);
		case 27u * 4u + ST_FP16C: return kernels_for<27, ST_FP16C>(trt, ext);
		default: return nullptr;
	}
}

} // namespace fx
