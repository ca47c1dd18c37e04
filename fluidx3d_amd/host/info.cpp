#include "info.hpp"
#include "lbm.hpp"

Info info;
Units units;

void Info::append(const ulong s, const ulong total_steps, const ulong t) {
	steps = total_steps == max_ulong ? s : total_steps;
	steps_last = t;
}
// exponential smoothing of the step time with the reference's constant 0.3 s
// (src/info.cpp:16-21): the benchmark scene reports the peak of N/smooth
void Info::update(const double dt) {
	runtime_lbm_timestep_last = dt;
	runtime_lbm_timestep_smooth = (dt + 0.3) / (0.3 / runtime_lbm_timestep_smooth + 1.0);
	runtime_lbm += dt;
	runtime_total = clock.stop();
}
void Info::print_initialize(LBM* l) {
	lbm = l;
	collision = string(
#ifdef SRT
		"SRT"
#else
		"TRT"
#endif
	);
	const string storage =
#if defined(FP16S)
		"FP32/FP16S";
#elif defined(FP16C)
		"FP32/FP16C";
#else
		"FP32/FP32";
#endif
	fx_device_info d;
	string device = "?";
	if(fx_device_info_get(0, &d) == FX_OK) device = string(d.name) + " " + string(d.arch) + ", " + to_string(d.compute_units) + " CUs, " + to_string((ulong)(d.memory_bytes >> 20)) + " MB";
	gpu_mem_required = (uint)(l->get_N() * (ulong)bytes_per_cell_device() / 1048576ull);
	cpu_mem_required = (uint)(l->get_N() * (ulong)bytes_per_cell_host() / 1048576ull);
	println("| Device  | " + device);
	println("| Grid    | " + to_string(l->get_Nx()) + " x " + to_string(l->get_Ny()) + " x " + to_string(l->get_Nz()) + " = " + to_string(l->get_N()) + " cells");
	println("| LBM     | D3Q" + to_string(l->get_velocity_set()) + " " + collision + " (" + storage + "), nu = " + to_string(l->get_nu()) + ", tau = " + to_string(l->get_tau()) + ", w = " + to_string(l->get_w()));
	println("| Memory  | GPU " + to_string(gpu_mem_required) + " MB, CPU " + to_string(cpu_mem_required) + " MB");
}
void Info::print_update() const {
	if(!lbm) return;
	const double mlups = (double)lbm->get_N() * 1E-6 / runtime_lbm_timestep_smooth;
	println("| t = " + to_string(lbm->get_t()) + " | " + to_string((ulong)mlups) + " MLUPs/s | " + to_string((ulong)(mlups * 1E-3 * (double)bandwidth_bytes_per_cell_device())) + " GB/s |");
}
void Info::print_finalize() {
	if(!lbm) return;
	if(runtime_lbm > 0.0) println("| finished t = " + to_string(lbm->get_t()) + ", average " + to_string((ulong)((double)lbm->get_N() * (double)lbm->get_t() * 1E-6 / runtime_lbm)) + " MLUPs/s |");
	lbm = nullptr;
}
