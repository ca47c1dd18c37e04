#pragma once
// Run-time metrics of the current LBM (reference interface: src/info.hpp): per-step time,
// its smoothed value used by the benchmark scene, MLUPs/s and bandwidth.
#include "utilities.hpp"

class LBM;
struct Info {
	LBM* lbm = nullptr;
	double runtime_lbm = 0.0, runtime_total = 0.0;
	double runtime_lbm_timestep_last = 1.0, runtime_lbm_timestep_smooth = 1.0;
	ulong steps = max_ulong, steps_last = 0ull;
	string collision;
	uint cpu_mem_required = 0u, gpu_mem_required = 0u;
	Clock clock;
	void append(const ulong steps, const ulong total_steps, const ulong t);
	void update(const double dt); // one time step took dt seconds
	void print_initialize(LBM* lbm);
	void print_update() const;
	void print_finalize();
};
extern Info info;
