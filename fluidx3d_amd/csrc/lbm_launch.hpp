// Launch table of the precompiled kernel variants.  One table per (velocity set,
// storage, collision); filled by the per-velocity-set translation units.
#pragma once
#include "lbm_device.hpp"

namespace fx {

struct LaunchShape {
	dim3 grid, block;
};
// Block of bx consecutive x cells of one row (bx = 64, 128 or 256); grid over (x chunk, y, z).
inline LaunchShape lattice_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz) {
	const uint32_t bx = Nx <= 64u ? 64u : Nx <= 128u ? 128u : 256u;
	LaunchShape s;
	s.block = dim3(bx, 1, 1);
	s.grid = dim3((Nx + bx - 1u) / bx, Ny, Nz);
	return s;
}

typedef void (*lattice_fn)(const LbmArgs&, const LaunchShape&, hipStream_t);
typedef void (*mass_fn)(const LbmArgs&, const LaunchShape&, hipStream_t, double*);

struct LbmKernels {
	lattice_fn initialize;
	lattice_fn stream_collide;
	lattice_fn update_fields;
	lattice_fn update_force_field;
	mass_fn mass_partial;
};

const LbmKernels* kernels_q19(uint32_t storage, bool trt);
const LbmKernels* kernels_q27(uint32_t storage, bool trt);

template<int Q, uint32_t ST, bool TRT> struct KernelSet {
	static void initialize(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_initialize<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void stream_collide(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_stream_collide<Q, ST, TRT><<<s.grid, s.block, 0, st>>>(a); }
	static void update_fields(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_fields<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void update_force_field(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_force_field<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void mass_partial(const LbmArgs& a, const LaunchShape& s, hipStream_t st, double* p) { k_mass_partial<Q, ST><<<s.grid, s.block, 0, st>>>(a, p); }
	static const LbmKernels* table() {
		static const LbmKernels k = {initialize, stream_collide, update_fields, update_force_field, mass_partial};
		return &k;
	}
};

template<int Q> const LbmKernels* select_kernels(uint32_t storage, bool trt) {
	switch(storage * 2u + (trt ? 1u : 0u)) {
		case ST_FP32 * 2u: return KernelSet<Q, ST_FP32, false>::table();
		case ST_FP32 * 2u + 1u: return KernelSet<Q, ST_FP32, true>::table();
		case ST_FP16S * 2u: return KernelSet<Q, ST_FP16S, false>::table();
		case ST_FP16S * 2u + 1u: return KernelSet<Q, ST_FP16S, true>::table();
		case ST_FP16C * 2u: return KernelSet<Q, ST_FP16C, false>::table();
		case ST_FP16C * 2u + 1u: return KernelSet<Q, ST_FP16C, true>::table();
		default: return nullptr;
	}
}

} // namespace fx
