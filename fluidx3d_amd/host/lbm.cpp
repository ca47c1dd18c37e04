#include "lbm.hpp"

#if defined(D3Q27)
static const uint velocity_set = 27u;
#else
static const uint velocity_set = 19u;
#endif

void fx_check(const int status, const string& what) {
	if(status != FX_OK) print_error(what + " failed: " + string(fx_last_error()));
}

uint bytes_per_cell_host() {
	uint b = 17u; // rho, u, flags
#ifdef FORCE_FIELD
	b += 12u;
#endif
#ifdef POROSITY
	b += 4u;
#endif
	return b;
}
uint bytes_per_cell_device() {
	uint b = velocity_set * (uint)sizeof(fpxx) + 17u;
#ifdef FORCE_FIELD
	b += 12u;
#endif
#ifdef POROSITY
	b += 4u;
#endif
	return b;
}
uint bandwidth_bytes_per_cell_device() {
	uint b = velocity_set * 2u * (uint)sizeof(fpxx) + 1u;
#ifdef UPDATE_FIELDS
	b += 16u;
#endif
#ifdef FORCE_FIELD
	b += 12u;
#endif
#ifdef POROSITY
	b += 4u;
#endif
	return b;
}
uint3 resolution(const float3 box_aspect_ratio, const uint memory) {
	const float required = box_aspect_ratio.x * box_aspect_ratio.y * box_aspect_ratio.z * (float)bytes_per_cell_device() / 1048576.0f;
	const float scaling = std::cbrt((float)memory / required);
	return uint3(to_uint(scaling * box_aspect_ratio.x), to_uint(scaling * box_aspect_ratio.y), to_uint(scaling * box_aspect_ratio.z));
}

fx_lbm_config lbm_config(const uint Nx, const uint Ny, const uint Nz, const float nu) {
	fx_lbm_config c{};
	c.Nx = Nx; c.Ny = Ny; c.Nz = Nz;
	c.velocity_set = velocity_set;
#ifdef TRT
	c.collision = FX_TRT;
#else
	c.collision = FX_SRT;
#endif
#if defined(FP16S)
	c.storage = FX_FP16S;
#elif defined(FP16C)
	c.storage = FX_FP16C;
#else
	c.storage = FX_FP32;
#endif
	c.extensions = 0u;
#ifdef VOLUME_FORCE
	c.extensions |= FX_VOLUME_FORCE;
#endif
#ifdef FORCE_FIELD
	c.extensions |= FX_FORCE_FIELD;
#endif
#ifdef EQUILIBRIUM_BOUNDARIES
	c.extensions |= FX_EQUILIBRIUM_BOUNDARIES;
#endif
#ifdef UPDATE_FIELDS
	c.extensions |= FX_UPDATE_FIELDS;
#endif
#ifdef POROSITY
	c.extensions |= FX_POROSITY;
#endif
#ifdef PRECISE_DENSITY
	c.extensions |= FX_PRECISE_DENSITY;
#endif
#ifdef MOVING_BOUNDARIES
	c.extensions |= FX_MOVING_BOUNDARIES;
#endif
#ifdef SUBGRID
	c.extensions |= FX_SUBGRID;
#endif
	c.w = fx_relaxation_rate(nu);
	return c;
}

LBM::LBM(const uint Nx, const uint Ny, const uint Nz, const uint Dx, const uint Dy, const uint Dz, const float nu, const float fx, const float fy, const float fz, const float sigma, const float alpha, const float beta, const uint particles_N, const float particles_rho) {
	construct(Nx, Ny, Nz, Dx, Dy, Dz, nu, fx, fy, fz, sigma, alpha, beta, particles_N, particles_rho);
}
LBM::LBM(const uint Nx, const uint Ny, const uint Nz, const float nu, const float fx, const float fy, const float fz, const float sigma, const float alpha, const float beta, const uint particles_N, const float particles_rho) {
	construct(Nx, Ny, Nz, 1u, 1u, 1u, nu, fx, fy, fz, sigma, alpha, beta, particles_N, particles_rho);
}
LBM::LBM(const uint Nx, const uint Ny, const uint Nz, const float nu, const uint particles_N, const float particles_rho) {
	construct(Nx, Ny, Nz, 1u, 1u, 1u, nu, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, particles_N, particles_rho);
}
LBM::LBM(const uint Nx, const uint Ny, const uint Nz, const float nu, const float fx, const float fy, const float fz, const uint particles_N, const float particles_rho) {
	construct(Nx, Ny, Nz, 1u, 1u, 1u, nu, fx, fy, fz, 0.0f, 0.0f, 0.0f, particles_N, particles_rho);
}
LBM::LBM(const uint3 N, const uint Dx, const uint Dy, const uint Dz, const float nu, const float fx, const float fy, const float fz, const float sigma, const float alpha, const float beta, const uint particles_N, const float particles_rho) {
	construct(N.x, N.y, N.z, Dx, Dy, Dz, nu, fx, fy, fz, sigma, alpha, beta, particles_N, particles_rho);
}
LBM::LBM(const uint3 N, const float nu, const float fx, const float fy, const float fz, const float sigma, const float alpha, const float beta, const uint particles_N, const float particles_rho) {
	construct(N.x, N.y, N.z, 1u, 1u, 1u, nu, fx, fy, fz, sigma, alpha, beta, particles_N, particles_rho);
}
LBM::LBM(const uint3 N, const float nu, const uint particles_N, const float particles_rho) {
	construct(N.x, N.y, N.z, 1u, 1u, 1u, nu, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, particles_N, particles_rho);
}
LBM::LBM(const uint3 N, const float nu, const float fx, const float fy, const float fz, const uint particles_N, const float particles_rho) {
	construct(N.x, N.y, N.z, 1u, 1u, 1u, nu, fx, fy, fz, 0.0f, 0.0f, 0.0f, particles_N, particles_rho);
}

// Constructor checks of src/lbm.cpp:783-838 (the ones that apply to this build), then
// the C ABI context, the DDF buffer and the host+device fields (src/lbm.cpp:121-176).
void LBM::construct(const uint nx, const uint ny, const uint nz, const uint dx, const uint dy, const uint dz, const float viscosity, const float f_x, const float f_y, const float f_z,
                    const float sigma, const float alpha, const float beta, const uint particles_N, const float particles_rho) {
	(void)particles_rho;
	if((ulong)nx * (ulong)ny * (ulong)nz == 0ull) print_error("Grid point number is 0: " + to_string(nx) + "x" + to_string(ny) + "x" + to_string(nz) + " = 0.");
	if(dx * dy * dz == 0u) print_error("You specified 0 LBM grid domains (" + to_string(dx) + "x" + to_string(dy) + "x" + to_string(dz) + ").");
	if(dx * dy * dz != 1u) print_error("This build runs one lattice on one GPU; multi-domain decomposition (" + to_string(dx) + "x" + to_string(dy) + "x" + to_string(dz) + ") is not available.");
	if(viscosity == 0.0f) print_error("Viscosity cannot be 0. Change it in setup.cpp.");
	if(viscosity < 0.0f) print_error("Viscosity cannot be negative. Remove the \"-\" in setup.cpp.");
#if !defined(VOLUME_FORCE) && !defined(POROSITY)
	if(f_x != 0.0f || f_y != 0.0f || f_z != 0.0f) print_error("Volume force is set in LBM constructor in main_setup(), but VOLUME_FORCE is not enabled. Uncomment \"#define VOLUME_FORCE\" in defines.hpp.");
#endif
	if(sigma != 0.0f) print_error("Surface tension is set in LBM constructor in main_setup(), but SURFACE is not available in this build.");
	if(alpha != 0.0f || beta != 0.0f) print_error("Thermal diffusion/expansion coefficients are set in LBM constructor in main_setup(), but TEMPERATURE is not available in this build.");
	if(particles_N > 0u) print_error("The number of particles is set to " + to_string(particles_N) + ">0, but PARTICLES is not available in this build.");
	Nx = nx; Ny = ny; Nz = nz; Dx = dx; Dy = dy; Dz = dz;
	nu = viscosity; fx = f_x; fy = f_y; fz = f_z;
	cfg = lbm_config(Nx, Ny, Nz, nu);
	int devices = 0;
	fx_check(fx_device_count(&devices), "device discovery");
	if(devices == 0) print_error("There are no GPU devices available.");
	fx_check(fx_ctx_create(0, &cfg, &ctx), "creating the LBM context");
	const ulong N = get_N();
	fx_check(fx_buf_alloc(ctx, N * (ulong)velocity_set * sizeof(fpxx), &fi), "allocating DDFs");
	rho.allocate(this, N, 1u, "rho", 1.0f);
	u.allocate(this, N, 3u, "u", 0.0f);
	flags.allocate(this, N, 1u, "flags", (uchar)0u);
#ifdef FORCE_FIELD
	F.allocate(this, N, 3u, "F", 0.0f);
#endif
#ifdef POROSITY
	porosity.allocate(this, N, 1u, "porosity", 0.0f);
	porosity_gradient.allocate(this, N, 1u, "porosity_gradient", 0.0f);
#endif
}

LBM::~LBM() {
	info.print_finalize();
	rho.release(); u.release(); flags.release();
#ifdef FORCE_FIELD
	F.release();
#endif
#ifdef POROSITY
	porosity.release();
	porosity_gradient.release();
	if(adjoint) fx_adjoint_destroy(adjoint);
#endif
	if(ctx) fx_ctx_destroy(ctx); // frees fi with the context
	ctx = nullptr;
}

// flags used vs extensions enabled, src/lbm.cpp:840-879
void LBM::sanity_checks_initialization() {
	bool equilibrium_used = false, moving_used = false, unsupported = false;
	const ulong N = get_N();
	for(ulong n = 0ull; n < N; n++) {
		const uchar f = flags[n];
		if((f & (TYPE_S | TYPE_E)) == TYPE_E) equilibrium_used = true;
		if((f & (TYPE_S | TYPE_E)) == TYPE_S && (u.x[n] != 0.0f || u.y[n] != 0.0f || u.z[n] != 0.0f)) moving_used = true;
		if(f & (TYPE_T | TYPE_F | TYPE_I | TYPE_G)) unsupported = true;
	}
#ifndef MOVING_BOUNDARIES
	if(moving_used) print_warning("Some boundary cells have non-zero velocity, but MOVING_BOUNDARIES is not enabled. If you intend to use moving boundaries, uncomment \"#define MOVING_BOUNDARIES\" in defines.hpp.");
#else
	if(!moving_used) print_warning("The MOVING_BOUNDARIES extension is enabled but no moving boundary cells (TYPE_S flag and velocity unequal to zero) are placed in the simulation box.");
#endif
#ifndef EQUILIBRIUM_BOUNDARIES
	if(equilibrium_used) print_error("Some cells are set as equilibrium boundaries with the TYPE_E flag, but EQUILIBRIUM_BOUNDARIES is not enabled. Uncomment \"#define EQUILIBRIUM_BOUNDARIES\" in defines.hpp.");
#else
	if(!equilibrium_used) print_warning("The EQUILIBRIUM_BOUNDARIES extension is enabled but no equilibrium boundary cells (TYPE_E flag) are placed in the simulation box.");
#endif
	if(unsupported) print_error("Some cells use TYPE_T/TYPE_F/TYPE_I/TYPE_G, whose extensions (TEMPERATURE, SURFACE) are not available in this build.");
}

// write fields, run initialize with the odd-step convention, reset t (src/lbm.cpp:881-922)
void LBM::initialize() {
#ifndef BENCHMARK
	sanity_checks_initialization();
#endif
	rho.write_to_device();
	u.write_to_device();
	flags.write_to_device();
#ifdef FORCE_FIELD
	F.write_to_device();
#endif
#ifdef POROSITY
	porosity.write_to_device();
#endif
	fx_check(fx_lbm_initialize(ctx, fi, (const float*)rho.device_buffer(), (float*)u.device_buffer(), (uint8_t*)flags.device_buffer()), "initialize");
	fx_check(fx_sync(ctx), "initialize");
	t = 0ull;
	t_last_update_fields = max_ulong;
	t_last_force_field = max_ulong;
	initialized = true;
}

#ifdef MOVING_BOUNDARIES
// mark/unmark cells next to moving solid cells with TYPE_MS after the host changed wall
// velocities (u written to the device), src/lbm.cpp:1019-1027
void LBM::update_moving_boundaries() {
	fx_check(fx_lbm_update_moving_boundaries(ctx, (const float*)u.device_buffer(), (uint8_t*)flags.device_buffer()), "update_moving_boundaries");
	fx_check(fx_sync(ctx), "update_moving_boundaries");
}
#endif

// LBM::run (src/lbm.cpp:955-975).  Steps are launched back to back in batches on the
// context stream; the per-step time given to Info is the batch average from HIP events.
void LBM::run(const ulong steps, const ulong total_steps) {
	info.append(steps, total_steps, t);
	if(!initialized) {
		initialize();
		info.print_initialize(this);
	}
	const float* Fp =
#ifdef FORCE_FIELD
		(const float*)F.device_buffer();
#else
		nullptr;
#endif
	const float* ap =
#ifdef POROSITY
		(const float*)porosity.device_buffer();
#else
		nullptr;
#endif
	ulong done = 0ull;
	while(done < steps) {
		const ulong batch = std::min<ulong>(steps - done, 1000ull);
		float ms = 0.0f;
		fx_check(fx_lbm_run(ctx, fi, (float*)rho.device_buffer(), (float*)u.device_buffer(), (const uint8_t*)flags.device_buffer(), t, batch, fx, fy, fz, Fp, ap, &ms), "stream_collide");
		const double dt = 1E-3 * (double)ms / (double)batch;
		for(ulong s = 0ull; s < batch; s++) info.update(dt);
		t += batch;
		done += batch;
	}
#ifdef UPDATE_FIELDS
	t_last_update_fields = t;
#endif
}

void LBM::update_fields() {
	if(!initialized) return;
	rho.read_from_device(); // runs update_fields if stale
	u.read_from_device();
}

void LBM::reset() { initialized = false; }

double LBM::total_mass() {
	double m = 0.0;
	fx_check(fx_lbm_mass_sum(ctx, fi, (const uint8_t*)flags.device_buffer(), t, &m), "mass sum");
	return m;
}

#ifdef FORCE_FIELD
void LBM::update_force_field() {
	if(t_last_force_field != t) {
		fx_check(fx_lbm_update_force_field(ctx, fi, (const uint8_t*)flags.device_buffer(), t, (float*)F.device_buffer()), "update_force_field");
		t_last_force_field = t;
	}
	fx_check(fx_sync(ctx), "update_force_field");
}
float3 LBM::object_center_of_mass(const uchar flag_marker) {
	double s[4];
	fx_check(fx_lbm_object_center_of_mass(ctx, (const uint8_t*)flags.device_buffer(), flag_marker, s), "object_center_of_mass");
	return float3((float)(s[0] / s[3]), (float)(s[1] / s[3]), (float)(s[2] / s[3]));
}
float3 LBM::object_force(const uchar flag_marker) {
	update_force_field();
	double s[3];
	fx_check(fx_lbm_object_force(ctx, (const float*)F.device_buffer(), (const uint8_t*)flags.device_buffer(), flag_marker, s), "object_force");
	return float3((float)s[0], (float)s[1], (float)s[2]);
}
float3 LBM::object_torque(const float3& rotation_center, const uchar flag_marker) {
	update_force_field();
	double s[3];
	fx_check(fx_lbm_object_torque(ctx, (const float*)F.device_buffer(), (const uint8_t*)flags.device_buffer(), flag_marker, rotation_center.x, rotation_center.y, rotation_center.z, s), "object_torque");
	return float3((float)s[0], (float)s[1], (float)s[2]);
}
#endif

#ifdef POROSITY
double LBM::forward(const ulong steps, const ulong window_begin, const uint plane_x, const ulong checkpoint_interval) {
	if(adjoint) { fx_adjoint_destroy(adjoint); adjoint = nullptr; }
	fx_adjoint_config c{steps, checkpoint_interval, window_begin, plane_x};
	fx_check(fx_adjoint_create(ctx, &c, &adjoint), "adjoint workspace");
	initialize();
	double J = 0.0;
	fx_check(fx_lbm_forward(ctx, adjoint, fi, (float*)rho.device_buffer(), (float*)u.device_buffer(), (const uint8_t*)flags.device_buffer(), fx, fy, fz,
#ifdef FORCE_FIELD
	                        (const float*)F.device_buffer(),
#else
	                        nullptr,
#endif
	                        (const float*)porosity.device_buffer(), &J, &forward_ms), "forward");
	t = steps;
	return J;
}
void LBM::backward() {
	if(!adjoint) print_error("LBM::backward() needs a preceding LBM::forward()");
	fx_check(fx_lbm_backward(ctx, adjoint, (const float*)rho.device_buffer(), (const float*)u.device_buffer(), (const uint8_t*)flags.device_buffer(), fx, fy, fz,
#ifdef FORCE_FIELD
	                         (const float*)F.device_buffer(),
#else
	                         nullptr,
#endif
	                         (const float*)porosity.device_buffer(), (float*)porosity_gradient.device_buffer(), &backward_ms), "backward");
	fx_check(fx_buf_read(ctx, porosity_gradient.device_buffer(), 0ull, porosity_gradient.capacity(), porosity_gradient.data(), 1), "reading the gradient");
}
#endif

// LBM status report in the reference's format (src/lbm.cpp:1047-1072), written to
// <path>status-<t as 9 digits>.txt (default path: the executable's export/ folder)
void LBM::write_status(const string& path) {
	string s;
	s += "Grid Resolution = " + to_string(Nx) + " x " + to_string(Ny) + " x " + to_string(Nz) + " = " + to_string(get_N()) + "\n";
	s += "Grid Domains = " + to_string(Dx) + " x " + to_string(Dy) + " x " + to_string(Dz) + " = " + to_string(get_D()) + "\n";
	s += "LBM Type = D3Q" + to_string(get_velocity_set()) + " " + info.collision + "\n";
	s += "Memory Usage = CPU " + to_string(info.cpu_mem_required) + " MB, GPU " + to_string(get_D()) + "x " + to_string(info.gpu_mem_required) + " MB\n";
	s += "Maximum Allocation Size = " + to_string((uint)(get_N() / (ulong)get_D() * (ulong)(get_velocity_set() * sizeof(fpxx)) / 1048576ull)) + " MB\n";
	s += "Time Steps = " + to_string(get_t()) + " / " + (info.steps == max_ulong ? string("infinite") : to_string(info.steps)) + "\n";
	s += "Runtime = " + print_time(info.runtime_total) + " (total) = " + print_time(info.runtime_lbm) + " (LBM) + " + print_time(info.runtime_total - info.runtime_lbm) + " (rendering and data evaluation)\n";
	s += "Average MLUPs/s = " + to_string(to_uint(1E-6 * (double)get_N() * (double)get_t() / info.runtime_lbm)) + "\n";
	s += "Kinematic Viscosity = " + to_string(get_nu()) + "\n";
	s += "Relaxation Time = " + to_string(get_tau()) + "\n";
	s += "Maximum Reynolds Number = " + to_string(get_Re_max()) + "\n";
#ifdef VOLUME_FORCE
	s += "Volume Force = (" + to_string(get_fx()) + ", " + to_string(get_fy()) + ", " + to_string(get_fz()) + ")\n";
#endif
	write_file(default_filename(path, "status", ".txt", get_t()), s);
}

// every member of the two field types, including those no scene calls
template class Memory_Container<float>;
template class Memory_Container<uchar>;
