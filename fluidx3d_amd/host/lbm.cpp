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
	if(viscosity == 0.0f) print_error("Viscosity cannot be 0. Change it in setup.cpp.");
	if(viscosity < 0.0f) print_error("Viscosity cannot be negative. Remove the \"-\" in setup.cpp.");
#if !defined(VOLUME_FORCE) && !defined(POROSITY)
	if(f_x != 0.0f || f_y != 0.0f || f_z != 0.0f) print_error("Volume force is set in LBM constructor in main_setup(), but VOLUME_FORCE is not enabled. Uncomment \"#define VOLUME_FORCE\" in defines.hpp.");
#endif
	if(sigma != 0.0f) print_error("Surface tension is set in LBM constructor in main_setup(), but SURFACE is not available in this build.");
	if(alpha != 0.0f || beta != 0.0f) print_error("Thermal diffusion/expansion coefficients are set in LBM constructor in main_setup(), but TEMPERATURE is not available in this build.");
	if(particles_N > 0u) print_error("The number of particles is set to " + to_string(particles_N) + ">0, but PARTICLES is not available in this build.");
	// resolution divisible by the domains (src/lbm.cpp:721-725)
	const uint ndx = (nx / dx) * dx, ndy = (ny / dy) * dy, ndz = (nz / dz) * dz;
	if(ndx != nx || ndy != ny || ndz != nz) print_warning("LBM grid (" + to_string(nx) + "x" + to_string(ny) + "x" + to_string(nz) + ") is not equally divisible in domains (" + to_string(dx) + "x" + to_string(dy) + "x" + to_string(dz) + "). Changing resolution to (" + to_string(ndx) + "x" + to_string(ndy) + "x" + to_string(ndz) + ").");
	if((ulong)ndx * (ulong)ndy * (ulong)ndz == 0ull) print_error("Grid point number is 0 after dividing into domains.");
	Nx = ndx; Ny = ndy; Nz = ndz; Dx = dx; Dy = dy; Dz = dz;
	nu = viscosity; fx = f_x; fy = f_y; fz = f_z;
	int devices = 0;
	fx_check(fx_device_count(&devices), "device discovery");
	if(devices == 0) print_error("There are no GPU devices available.");
	// domains d = x + (y + z*Dy)*Dx of Nx/Dx (+2 halo cells along a split axis), all on
	// device 0 with a context each (src/lbm.cpp:727-733)
	const uint hx = Dx > 1u, hy = Dy > 1u, hz = Dz > 1u, D = get_D();
	domain.assign(D, LBM_Domain());
	for(uint d = 0u; d < D; d++) {
		LBM_Domain& dom = domain[d];
		const uint px = d % Dx, py = (d / Dx) % Dy, pz = d / (Dx * Dy);
		dom.Nx = Nx / Dx + 2u * hx; dom.Ny = Ny / Dy + 2u * hy; dom.Nz = Nz / Dz + 2u * hz;
		dom.ox = (int)(px * (Nx / Dx)) - (int)hx; dom.oy = (int)(py * (Ny / Dy)) - (int)hy; dom.oz = (int)(pz * (Nz / Dz)) - (int)hz;
		dom.cfg = lbm_config(dom.Nx, dom.Ny, dom.Nz, nu);
		fx_check(fx_ctx_create(0, &dom.cfg, &dom.ctx), "creating the LBM context of domain " + to_string(d));
		fx_check(fx_buf_alloc(dom.ctx, dom.get_N() * (ulong)velocity_set * sizeof(fpxx), &dom.fi), "allocating DDFs");
		const uint split[3] = {Dx, Dy, Dz};
		for(uint axis = 0u; axis < 3u; axis++) {
			if(split[axis] == 1u) continue;
			ulong bytes = 0ull;
			for(uint field = FX_FIELD_FI; field <= FX_FIELD_F; field++) bytes = max(bytes, (ulong)fx_transfer_bytes(&dom.cfg, field, axis));
			fx_check(fx_buf_alloc(dom.ctx, bytes, &dom.transfer_p[axis]), "allocating transfer buffers");
			fx_check(fx_buf_alloc(dom.ctx, bytes, &dom.transfer_m[axis]), "allocating transfer buffers");
		}
	}
	ctx = domain[0].ctx; fi = domain[0].fi; cfg = domain[0].cfg;
	const ulong N = get_N();
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
	for(LBM_Domain& dom : domain) if(dom.ctx) fx_ctx_destroy(dom.ctx); // frees fi and the transfer buffers with the context
	domain.clear();
	ctx = nullptr;
}

void LBM::require_single_domain(const string& what) const {
	if(get_D() != 1u) print_error(what + " addresses one domain's buffers and is not available on a lattice split into " + to_string(Dx) + "x" + to_string(Dy) + "x" + to_string(Dz) + " domains.");
}
void LBM::sync_domains() {
	for(LBM_Domain& dom : domain) fx_check(fx_sync(dom.ctx), "synchronising the domains");
}
// domain next to d along axis in direction step (+1/-1), periodic
uint LBM::neighbour(const uint d, const uint axis, const int step) const {
	uint c[3] = {d % Dx, (d / Dx) % Dy, d / (Dx * Dy)};
	const uint n[3] = {Dx, Dy, Dz};
	c[axis] = (uint)(((int)c[axis] + step + (int)n[axis]) % (int)n[axis]);
	return c[0] + (c[1] + c[2] * Dy) * Dx;
}
// Per split axis: every domain extracts the values crossing its two faces into its
// buffers (kernel on its own stream), all finish, then every domain inserts the buffers
// of its neighbours (device to device: all domains share the device), all finish.
void LBM::communicate(const uint field) {
	const uint split[3] = {Dx, Dy, Dz};
	for(uint axis = 0u; axis < 3u; axis++) {
		if(split[axis] == 1u) continue;
		for(uint d = 0u; d < get_D(); d++) {
			LBM_Domain& dom = domain[d];
			fx_check(fx_lbm_transfer_extract(dom.ctx, field, axis, t, dom.transfer_p[axis], dom.transfer_m[axis], dom.fi, (const float*)rho.device_buffer(d), (const float*)u.device_buffer(d), (const uint8_t*)flags.device_buffer(d),
#ifdef FORCE_FIELD
			                                 (const float*)F.device_buffer(d)),
#else
			                                 nullptr),
#endif
			         "halo extract");
		}
		sync_domains();
		for(uint d = 0u; d < get_D(); d++) {
			LBM_Domain& dom = domain[d];
			const void* from_p = domain[neighbour(d, axis, +1)].transfer_m[axis]; // the + neighbour's - side
			const void* from_m = domain[neighbour(d, axis, -1)].transfer_p[axis]; // the - neighbour's + side
			fx_check(fx_lbm_transfer_insert(dom.ctx, field, axis, t, from_p, from_m, dom.fi, (float*)rho.device_buffer(d), (float*)u.device_buffer(d), (uint8_t*)flags.device_buffer(d),
#ifdef FORCE_FIELD
			                                (float*)F.device_buffer(d)),
#else
			                                nullptr),
#endif
			         "halo insert");
		}
		sync_domains();
	}
}
void LBM::update_fields_on_device() {
	if(!initialized || t_last_update_fields == t) return;
	for(uint d = 0u; d < get_D(); d++) {
		fx_check(fx_lbm_update_fields(domain[d].ctx, domain[d].fi, (float*)rho.device_buffer(d), (float*)u.device_buffer(d), (const uint8_t*)flags.device_buffer(d), t, fx, fy, fz,
#ifdef FORCE_FIELD
		                              (const float*)F.device_buffer(d),
#else
		                              nullptr,
#endif
#ifdef POROSITY
		                              (const float*)porosity.device_buffer(d)),
#else
		                              nullptr),
#endif
		         "update_fields");
	}
	t_last_update_fields = t;
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
	if(get_D() == 1u) {
		fx_check(fx_lbm_initialize(ctx, fi, (const float*)rho.device_buffer(), (float*)u.device_buffer(), (uint8_t*)flags.device_buffer()), "initialize");
		fx_check(fx_sync(ctx), "initialize");
	} else { // halos as the reference fills them (src/lbm.cpp:884-920); t odd for the transfers
#ifdef FORCE_FIELD
		communicate(FX_FIELD_F);
#endif
		t = 1ull;
		communicate(FX_FIELD_RHO_U_FLAGS);
		for(uint d = 0u; d < get_D(); d++) fx_check(fx_lbm_initialize(domain[d].ctx, domain[d].fi, (const float*)rho.device_buffer(d), (float*)u.device_buffer(d), (uint8_t*)flags.device_buffer(d)), "initialize");
		communicate(FX_FIELD_RHO_U_FLAGS);
		communicate(FX_FIELD_FI);
		sync_domains();
	}
	t = 0ull;
	t_last_update_fields = max_ulong;
	t_last_force_field = max_ulong;
	initialized = true;
}

#ifdef MOVING_BOUNDARIES
// mark/unmark cells next to moving solid cells with TYPE_MS after the host changed wall
// velocities (u written to the device), src/lbm.cpp:1019-1027
void LBM::update_moving_boundaries() {
	for(uint d = 0u; d < get_D(); d++) fx_check(fx_lbm_update_moving_boundaries(domain[d].ctx, (const float*)u.device_buffer(d), (uint8_t*)flags.device_buffer(d)), "update_moving_boundaries");
	if(get_D() > 1u) communicate(FX_FIELD_FLAGS);
	sync_domains();
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
	if(get_D() == 1u) {
		while(done < steps) {
			const ulong batch = std::min<ulong>(steps - done, 1000ull);
			float ms = 0.0f;
			fx_check(fx_lbm_run(ctx, fi, (float*)rho.device_buffer(), (float*)u.device_buffer(), (const uint8_t*)flags.device_buffer(), t, batch, fx, fy, fz, Fp, ap, &ms), "stream_collide");
			const double dt = 1E-3 * (double)ms / (double)batch;
			for(ulong s = 0ull; s < batch; s++) info.update(dt);
			t += batch;
			done += batch;
		}
	} else { // do_time_step (src/lbm.cpp:924-953): every domain, then the DDF halos
		Clock clock;
		for(; done < steps; done++) {
			clock.start();
			for(uint d = 0u; d < get_D(); d++) {
				fx_check(fx_lbm_stream_collide(domain[d].ctx, domain[d].fi, (float*)rho.device_buffer(d), (float*)u.device_buffer(d), (const uint8_t*)flags.device_buffer(d), t, fx, fy, fz,
#ifdef FORCE_FIELD
				                               (const float*)F.device_buffer(d),
#else
				                               nullptr,
#endif
#ifdef POROSITY
				                               (const float*)porosity.device_buffer(d)),
#else
				                               nullptr),
#endif
				         "stream_collide");
			}
			communicate(FX_FIELD_FI);
			info.update(clock.stop());
			t++;
		}
	}
	(void)Fp; (void)ap;
#ifdef UPDATE_FIELDS
	t_last_update_fields = t;
#endif
}

void LBM::update_fields() {
	if(!initialized) return;
	update_fields_on_device();
	sync_domains();
}

void LBM::reset() { initialized = false; }

// DDF codes as the kernels store them (src/lbm.cpp:410-414, src/kernel.cpp:848-859), for
// the host-side sum over the interior cells of split lattices
static inline float ddf_value(const float v) { return v; }
static inline float ddf_value(const ushort c) {
#if defined(FP16S) // IEEE binary16 times 2^-15
	const uint sign = (uint)(c & 0x8000u) << 16, e = (c >> 10) & 0x1Fu, m = c & 0x3FFu;
	float v;
	if(e == 0u) v = std::ldexp((float)m, -24); // subnormal
	else if(e == 31u) v = m ? NAN : INFINITY;
	else v = as_float(((e + 112u) << 23) | (m << 13));
	return std::ldexp(as_float(as_uint(v) | sign), -15);
#else // 1-4-11 with bias 15: exponent and mantissa at the binary32 positions, times 2^112
	return std::ldexp(as_float(((uint)(c & 0x8000u) << 16) | ((uint)(c & 0x7FFFu) << 12)), 112);
#endif
}
double LBM::total_mass() {
	double m = 0.0;
	if(get_D() == 1u) {
		fx_check(fx_lbm_mass_sum(ctx, fi, (const uint8_t*)flags.device_buffer(), t, &m), "mass sum");
		return m;
	}
	// split lattice: 1 plus the DDFs in the own slots of every interior cell, as the
	// undivided lattice's sum; the halo cells are copies and are left out
	const uint hx = Dx > 1u, hy = Dy > 1u, hz = Dz > 1u;
	for(const LBM_Domain& dom : domain) {
		const ulong NL = dom.get_N();
		vector<fpxx> local(NL * (ulong)velocity_set);
		fx_check(fx_buf_read(dom.ctx, dom.fi, 0ull, local.size() * sizeof(fpxx), local.data(), 1), "reading DDFs");
		for(uint lz = hz; lz < dom.Nz - hz; lz++) for(uint ly = hy; ly < dom.Ny - hy; ly++) for(uint lx = hx; lx < dom.Nx - hx; lx++) {
			const ulong l = (ulong)lx + ((ulong)ly + (ulong)lz * dom.Ny) * dom.Nx;
			m += 1.0;
			for(uint i = 0u; i < velocity_set; i++) m += (double)ddf_value(local[(ulong)i * NL + l]);
		}
	}
	return m;
}

#ifdef FORCE_FIELD
void LBM::update_force_field() {
	if(t_last_force_field != t) {
		for(uint d = 0u; d < get_D(); d++) fx_check(fx_lbm_update_force_field(domain[d].ctx, domain[d].fi, (const uint8_t*)flags.device_buffer(d), t, (float*)F.device_buffer(d)), "update_force_field");
		t_last_force_field = t;
	}
	sync_domains();
}
// Object sums.  One domain: FP64 reductions on the device.  A split lattice: the interior
// cells' F and flags come to the host and are summed there in FP64 in cell order, per-cell
// terms in FP32 as on the device, so the result is the undivided lattice's to FP64 rounding.
float3 LBM::object_center_of_mass(const uchar flag_marker) {
	double s[4] = {0.0, 0.0, 0.0, 0.0};
	if(get_D() == 1u) {
		fx_check(fx_lbm_object_center_of_mass(ctx, (const uint8_t*)flags.device_buffer(), flag_marker, s), "object_center_of_mass");
	} else {
		flags.read_from_device();
		for(ulong n = 0ull; n < get_N(); n++) {
			if(flags[n] != flag_marker) continue;
			const float3 p = position(n);
			s[0] += (double)p.x; s[1] += (double)p.y; s[2] += (double)p.z; s[3] += 1.0;
		}
	}
	return float3((float)(s[0] / s[3]), (float)(s[1] / s[3]), (float)(s[2] / s[3]));
}
float3 LBM::object_force(const uchar flag_marker) {
	update_force_field();
	double s[3] = {0.0, 0.0, 0.0};
	if(get_D() == 1u) {
		fx_check(fx_lbm_object_force(ctx, (const float*)F.device_buffer(), (const uint8_t*)flags.device_buffer(), flag_marker, s), "object_force");
	} else {
		F.read_from_device();
		flags.read_from_device();
		for(ulong n = 0ull; n < get_N(); n++) {
			if(flags[n] != flag_marker) continue;
			s[0] += (double)F.x[n]; s[1] += (double)F.y[n]; s[2] += (double)F.z[n];
		}
	}
	return float3((float)s[0], (float)s[1], (float)s[2]);
}
float3 LBM::object_torque(const float3& rotation_center, const uchar flag_marker) {
	update_force_field();
	double s[3] = {0.0, 0.0, 0.0};
	if(get_D() == 1u) {
		fx_check(fx_lbm_object_torque(ctx, (const float*)F.device_buffer(), (const uint8_t*)flags.device_buffer(), flag_marker, rotation_center.x, rotation_center.y, rotation_center.z, s), "object_torque");
	} else {
		F.read_from_device();
		flags.read_from_device();
		for(ulong n = 0ull; n < get_N(); n++) {
			if(flags[n] != flag_marker) continue;
			const float3 r = position(n) - rotation_center;
			const float Fx = F.x[n], Fy = F.y[n], Fz = F.z[n];
			s[0] += (double)(r.y * Fz - r.z * Fy); s[1] += (double)(r.z * Fx - r.x * Fz); s[2] += (double)(r.x * Fy - r.y * Fx);
		}
	}
	return float3((float)s[0], (float)s[1], (float)s[2]);
}
#endif

#ifdef POROSITY
double LBM::forward(const ulong steps, const ulong window_begin, const uint plane_x, const ulong checkpoint_interval) {
	require_single_domain("forward()/backward()");
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
