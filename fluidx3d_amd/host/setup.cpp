// Scenes written against the reference's LBM API.  Select one with -DSCENE_<NAME> (the
// reference selects by commenting blocks in src/setup.cpp); BENCHMARK selects the benchmark.
#include "setup.hpp"

#if defined(BENCHMARK)
// Peak MLUPs/s of an empty periodic 256^3 box at nu=1 over 1000 x run(10)
// (same scene as the reference's benchmark, src/setup.cpp:5-36).
void main_setup() {
	uint mlups = 0u;
	{
		LBM lbm(256u, 256u, 256u, 1.0f);
		for(uint i = 0u; i < 1000u; i++) {
			lbm.run(10u, 1000u * 10u);
			mlups = max(mlups, to_uint((double)lbm.get_N() * 1E-6 / info.runtime_lbm_timestep_smooth));
		}
	}
	print_info("Peak MLUPs/s = " + to_string(mlups));
}

#elif defined(SCENE_TAYLOR_GREEN)
#ifndef TGV_N
#define TGV_N 64u
#endif
#ifndef TGV_STEPS
#define TGV_STEPS 1000u
#endif
// 3D Taylor-Green vortex, u0 = 0.05, nu = 1/6, periodic (workload of the 64^3 and 256^3
// configurations); prints the centre-cell u.x like the reference's binary-identity line
// (src/setup.cpp:58) and writes u as VTK when FX_VTK is set.
void main_setup() {
	LBM lbm(TGV_N, TGV_N, TGV_N, 1.0f / 6.0f);
	const uint Nx = lbm.get_Nx(), Ny = lbm.get_Ny(), Nz = lbm.get_Nz();
	const double u0 = 0.05; // fields evaluated in double on cell centres, then rounded to float
	parallel_for(lbm.get_N(), [&](ulong n) {
		uint x = 0u, y = 0u, z = 0u; lbm.coordinates(n, x, y, z);
		const double px = (double)x + 0.5 - 0.5 * (double)Nx, py = (double)y + 0.5 - 0.5 * (double)Ny, pz = (double)z + 0.5 - 0.5 * (double)Nz;
		const double kx = 2.0 * pi / (double)Nx, ky = 2.0 * pi / (double)Ny, kz = 2.0 * pi / (double)Nz;
		lbm.u.x[n] = (float)(u0 * std::sin(kx * px) * std::cos(ky * py) * std::cos(kz * pz));
		lbm.u.y[n] = (float)(-u0 * std::cos(kx * px) * std::sin(ky * py) * std::cos(kz * pz));
		lbm.u.z[n] = 0.0f;
		lbm.rho[n] = (float)(1.0 + 3.0 * (u0 * u0 / 16.0) * (std::cos(2.0 * kx * px) + std::cos(2.0 * ky * py)) * (std::cos(2.0 * kz * pz) + 2.0));
	});
	const double m0 = (lbm.run(0u), lbm.total_mass());
	lbm.run(TGV_STEPS);
	lbm.u.read_from_device();
	lbm.rho.read_from_device();
	double ke = 0.0;
	for(ulong n = 0ull; n < lbm.get_N(); n++) ke += 0.5 * (double)lbm.rho[n] * ((double)sq(lbm.u.x[n]) + (double)sq(lbm.u.y[n]) + (double)sq(lbm.u.z[n]));
	char line[256];
	std::snprintf(line, sizeof(line), "t=%llu centre u.x=%.9e kinetic energy=%.12e mass drift=%.3e", (unsigned long long)lbm.get_t(), (double)lbm.u.x[lbm.index(Nx / 2u, Ny / 2u, Nz / 2u)], ke, (lbm.total_mass() - m0) / m0);
	println(line);
#ifdef FX_VTK
	lbm.u.write_host_to_vtk("export/", false); // export/u-<t>.vtk
#endif
	lbm.write_status("export/"); // export/status-<t>.txt
}

#elif defined(SCENE_SPHERE)
#ifndef SPHERE_STEPS
#define SPHERE_STEPS 1000u
#endif
#ifndef SPHERE_SCALE
#define SPHERE_SCALE 1u
#endif
#ifndef SPHERE_REPORT
#define SPHERE_REPORT 1000u
#endif
// Flow past a sphere: 512x256x256 (times SPHERE_SCALE), radius 32*scale at
// (128,128,128)*scale, TYPE_E inflow/outflow u=(0.075,0,0) at x=0 and x=Nx-1, periodic in y
// and z, Re=100 on the diameter, seeded 1e-4 perturbation of u.  Needs
// EQUILIBRIUM_BOUNDARIES.  With FORCE_FIELD the drag (FP64 object force sum) is printed
// every SPHERE_REPORT steps (and at the end); after the run, the FP64 total mass against its
// value after initialization and the number of non-finite velocity components.
void main_setup() {
	const uint s = SPHERE_SCALE;
	const float R = 32.0f * (float)s, u_in = 0.075f, Re = 100.0f;
	const float nu = units.nu_from_Re(Re, 2.0f * R, u_in);
	LBM lbm(512u * s, 256u * s, 256u * s, nu);
	const uint Nx = lbm.get_Nx();
	const float3 c = float3(128.0f * (float)s, 128.0f * (float)s, 128.0f * (float)s);
	uint seed = 42u;
	for(ulong n = 0ull; n < lbm.get_N(); n++) { // sequential: the LCG perturbation is seeded and ordered
		uint x = 0u, y = 0u, z = 0u; lbm.coordinates(n, x, y, z);
		const float px = random_symmetric(seed, 1E-4f), py = random_symmetric(seed, 1E-4f), pz = random_symmetric(seed, 1E-4f);
		if(sphere(x, y, z, c, R)) { lbm.flags[n] = TYPE_S | TYPE_X; continue; }
		lbm.u.x[n] = u_in + px; lbm.u.y[n] = py; lbm.u.z[n] = pz;
		if(x == 0u || x == Nx - 1u) { lbm.flags[n] = TYPE_E; lbm.u.x[n] = u_in; lbm.u.y[n] = 0.0f; lbm.u.z[n] = 0.0f; }
	}
	lbm.run(0u); // initialize
	const double mass0 = lbm.total_mass();
	char line[256];
	while(lbm.get_t() < (ulong)SPHERE_STEPS) {
		lbm.run(std::min<ulong>((ulong)SPHERE_REPORT, (ulong)SPHERE_STEPS - lbm.get_t()));
#ifdef FORCE_FIELD
		const float3 F = lbm.object_force(TYPE_S | TYPE_X);
		const double Cd = (double)F.x / (0.5 * (double)sq(u_in) * pi * (double)sq(R));
		std::snprintf(line, sizeof(line), "t=%llu drag F=(%.9e, %.9e, %.9e) Cd=%.6f MLUPs/s=%.0f", (unsigned long long)lbm.get_t(), (double)F.x, (double)F.y, (double)F.z, Cd, (double)lbm.get_N() * (double)lbm.get_t() * 1E-6 / info.runtime_lbm);
		println(line);
#endif
	}
	const double mass1 = lbm.total_mass();
	lbm.u.read_from_device();
	ulong nonfinite = 0ull;
	for(ulong i = 0ull; i < 3ull * lbm.get_N(); i++) nonfinite += !std::isfinite(lbm.u.data()[i]);
	std::snprintf(line, sizeof(line), "mass %.12e -> %.12e (relative change %.3e), non-finite u components %llu", mass0, mass1, (mass1 - mass0) / mass0, (unsigned long long)nonfinite);
	println(line);
}

#elif defined(SCENE_STOKES)
// Stokes drag validation, the reference's scene (src/setup.cpp:148-187): a sphere of
// radius 32 in a 256^3 box at Re = 0.01, nu = 1; all six faces are TYPE_E cells preset to
// the Stokes solution u_Stokes/rho_Stokes with far-field velocity (-u, 0, 0), as is the
// initial fluid; every 100 steps the FP64 object force of the TYPE_S|TYPE_X sphere is
// compared with F_Stokes = 6 pi rho nu R u until the error has converged.  Needs
// FORCE_FIELD and EQUILIBRIUM_BOUNDARIES.
void main_setup() {
	const ulong dt = 100ull; // check error every dt time steps
	const float R = 32.0f; // sphere radius
	const float Re = 0.01f; // Reynolds number
	const float nu = 1.0f; // kinematic shear viscosity
	const float rho = 1.0f; // density
	const uint L = to_uint(8.0f*R); // simulation box size
	const float u = units.u_from_Re(Re, 2.0f*R, nu); // velocity
	LBM lbm(L, L, L, nu); // flow driven by equilibrium boundaries
	const uint Nx=lbm.get_Nx(), Ny=lbm.get_Ny(), Nz=lbm.get_Nz(); parallel_for(lbm.get_N(), [&](ulong n) { uint x=0u, y=0u, z=0u; lbm.coordinates(n, x, y, z);
		if(x==0u||x==Nx-1u||y==0u||y==Ny-1u||z==0u||z==Nz-1u) lbm.flags[n] = TYPE_E;
		if(sphere(x, y, z, lbm.center(), R)) {
			lbm.flags[n] = TYPE_S|TYPE_X; // flag boundary cells for force summation additionally with TYPE_X
		} else {
			lbm.rho[n] = units.rho_Stokes(lbm.position(x, y, z), float3(-u, 0.0f, 0.0f), R, rho, nu);
			const float3 un = units.u_Stokes(lbm.position(x, y, z), float3(-u, 0.0f, 0.0f), R);
			lbm.u.x[n] = un.x;
			lbm.u.y[n] = un.y;
			lbm.u.z[n] = un.z;
		}
	});
	double E1=1000.0, E2=1000.0;
	while(true) { // main simulation loop
		lbm.run(dt);
		const float3 force = lbm.object_force(TYPE_S|TYPE_X);
		const double F_theo = units.F_Stokes(rho, u, nu, R);
		const double F_sim = (double)length(force);
		const double E0 = fabs(F_sim-F_theo)/F_theo;
		print_info(to_string(lbm.get_t())+", expected: "+to_string(F_theo, 6u)+", measured: "+to_string(F_sim, 6u)+", error = "+to_string((float)(100.0*E0), 1u)+"%");
		if(converged(E2, E1, E0, 1E-4)) { // stop when error has sufficiently converged
			print_info("Error converged after "+to_string(lbm.get_t())+" steps to "+to_string(100.0*E0, 1u)+"%");
			wait();
			break;
		}
#ifdef STOKES_MAX_STEPS
		if(lbm.get_t() >= STOKES_MAX_STEPS) print_error("Stokes drag did not converge within " + to_string(lbm.get_t()) + " steps");
#endif
		E2 = E1;
		E1 = E0;
	}
	const float3 F = lbm.object_force(TYPE_S|TYPE_X); // no step since the last check: the same force
	char line[256];
	std::snprintf(line, sizeof(line), "t=%llu F=(%.9e, %.9e, %.9e) F_Stokes=%.9e", (unsigned long long)lbm.get_t(), (double)F.x, (double)F.y, (double)F.z, (double)units.F_Stokes(rho, u, nu, R));
	println(line);
}

#elif defined(SCENE_DOMAINS)
#ifndef DOMAINS_X
#define DOMAINS_X 1u
#endif
#ifndef DOMAINS_Y
#define DOMAINS_Y 1u
#endif
#ifndef DOMAINS_Z
#define DOMAINS_Z 1u
#endif
// A 24x18x12 lattice split into DOMAINS_X x DOMAINS_Y x DOMAINS_Z domains on one device
// (the reference's LBM(Nx,Ny,Nz,Dx,Dy,Dz,...) constructor, src/lbm.cpp:721-773): seeded
// random rho/u/F, scattered TYPE_S cells, TYPE_E planes at x=0 and x=Nx-1, a volume force.
// Writes the initial fields and rho/u after 9 steps as VTK to export/, prints the FP64
// sums.  Needs EQUILIBRIUM_BOUNDARIES, VOLUME_FORCE and FORCE_FIELD.
void main_setup() {
	LBM lbm(24u, 18u, 12u, DOMAINS_X, DOMAINS_Y, DOMAINS_Z, 0.05f, 1E-5f, -2E-5f, 5E-6f);
	const uint Nx = lbm.get_Nx();
	uint seed = 7u;
	for(ulong n = 0ull; n < lbm.get_N(); n++) { // sequential: the LCG draws are ordered
		uint x = 0u, y = 0u, z = 0u; lbm.coordinates(n, x, y, z);
		lbm.rho[n] = 1.0f + random_symmetric(seed, 0.01f);
		const float ux = random_symmetric(seed, 0.05f), uy = random_symmetric(seed, 0.05f), uz = random_symmetric(seed, 0.05f);
		lbm.F.x[n] = random_symmetric(seed, 1E-4f); lbm.F.y[n] = random_symmetric(seed, 1E-4f); lbm.F.z[n] = random_symmetric(seed, 1E-4f);
		if(random(seed, 1.0f) < 0.08f) { lbm.flags[n] = TYPE_S; continue; }
		lbm.u.x[n] = ux; lbm.u.y[n] = uy; lbm.u.z[n] = uz;
		if(x == 0u || x == Nx - 1u) lbm.flags[n] = TYPE_E;
	}
	lbm.rho.write_host_to_vtk("export/", false);
	lbm.u.write_host_to_vtk("export/", false);
	lbm.flags.write_host_to_vtk("export/", false);
	lbm.F.write_host_to_vtk("export/", false);
	lbm.run(9u);
	lbm.rho.write_device_to_vtk("export/", false);
	lbm.u.write_device_to_vtk("export/", false);
	const double mass = lbm.total_mass();
	const float3 F = lbm.object_force(TYPE_S), M = lbm.object_torque(float3(1.5f, -2.0f, 0.5f), TYPE_S), c = lbm.object_center_of_mass(TYPE_S);
	char line[512];
	std::snprintf(line, sizeof(line), "domains %ux%ux%u t=%llu mass=%.17g force=(%.9e, %.9e, %.9e) torque=(%.9e, %.9e, %.9e) com=(%.9e, %.9e, %.9e)", lbm.get_Dx(), lbm.get_Dy(), lbm.get_Dz(),
	              (unsigned long long)lbm.get_t(), mass, (double)F.x, (double)F.y, (double)F.z, (double)M.x, (double)M.y, (double)M.z, (double)c.x, (double)c.y, (double)c.z);
	println(line);
}

#elif defined(SCENE_CAVITY)
#ifndef CAVITY_L
#define CAVITY_L 128u
#endif
#ifndef CAVITY_STEPS
#define CAVITY_STEPS 10000u
#endif
// Lid-driven cavity, the reference's scene (src/setup.cpp:236-251) without graphics: the
// lid z = Nz-1 moves with u = 0.1 in y, all six faces solid.  Needs MOVING_BOUNDARIES.
void main_setup() {
	const uint L = CAVITY_L;
	const float Re = 1000.0f;
	const float u = 0.1f;
	LBM lbm(L, L, L, units.nu_from_Re(Re, (float)(L-2u), u));
	const uint Nx=lbm.get_Nx(), Ny=lbm.get_Ny(), Nz=lbm.get_Nz(); parallel_for(lbm.get_N(), [&](ulong n) { uint x=0u, y=0u, z=0u; lbm.coordinates(n, x, y, z);
		if(z==Nz-1) lbm.u.y[n] = u;
		if(x==0u||x==Nx-1u||y==0u||y==Ny-1u||z==0u||z==Nz-1u) lbm.flags[n] = TYPE_S; // all non periodic
	});
	lbm.run(CAVITY_STEPS);
	lbm.u.read_from_device();
	char line[256];
	std::snprintf(line, sizeof(line), "t=%llu centre u=(%.9e, %.9e, %.9e)", (unsigned long long)lbm.get_t(), (double)lbm.u.x[lbm.index(Nx/2u, Ny/2u, Nz/2u)], (double)lbm.u.y[lbm.index(Nx/2u, Ny/2u, Nz/2u)], (double)lbm.u.z[lbm.index(Nx/2u, Ny/2u, Nz/2u)]);
	println(line);
#ifdef FX_VTK
	lbm.u.write_host_to_vtk("export/", false);
#endif
}

#elif defined(SCENE_POROUS_CHANNEL)
#ifndef CHANNEL_SCALE
#define CHANNEL_SCALE 1u
#endif
// Porous channel of the adjoint: walls at y=0 and y=Ny-1, periodic x/z, body force 1e-6,
// porosity (Brinkman) in the middle third of x; J = outlet x-momentum over the last quarter
// of the run; prints J and dJ/dporosity at a few cells.  Needs VOLUME_FORCE and POROSITY.
void main_setup() {
	const uint s = CHANNEL_SCALE;
	LBM lbm(128u * s, 64u * s, 64u * s, 1.0f / 6.0f, 1E-6f, 0.0f, 0.0f);
	const uint Nx = lbm.get_Nx(), Ny = lbm.get_Ny();
	uint seed = 42u;
	for(ulong n = 0ull; n < lbm.get_N(); n++) {
		uint x = 0u, y = 0u, z = 0u; lbm.coordinates(n, x, y, z);
		if(y == 0u || y == Ny - 1u) { lbm.flags[n] = TYPE_S; continue; }
		lbm.porosity[n] = (x >= Nx / 3u && x < 2u * Nx / 3u) ? random(seed, 1.0f) : 0.0f;
	}
	const ulong steps = 2000ull * s;
	const double J = lbm.forward(steps, steps - steps / 4ull, Nx - 1u, 0ull);
	lbm.backward();
	double norm = 0.0;
	for(ulong n = 0ull; n < lbm.get_N(); n++) norm += sq((double)lbm.porosity_gradient[n]);
	char line[256];
	std::snprintf(line, sizeof(line), "J=%.12e |dJ/dporosity|=%.6e forward %.1f ms backward %.1f ms", J, std::sqrt(norm), (double)lbm.forward_ms, (double)lbm.backward_ms);
	println(line);
#ifdef FX_VTK
	lbm.porosity_gradient.write_host_to_vtk("export/", false);
#endif
}

#elif defined(SCENE_BOX_IO)
// Box transfers of the field containers (read_from_device_1d/2d/3d, write_to_device_*):
// the elements moved must be those of the reference's row loops (src/opencl.hpp:532-607),
// per component and row: offset min(i*N+n, N*d), length min(x1-x0, N*d-offset).
template<typename T> static ulong box_case(LBM& lbm, Memory_Container<T>& m, const bool write, const int form, const int dimension) {
	const ulong N = m.length(), D = m.dimensions(), R = m.range(), Nx = lbm.get_Nx(), Ny = lbm.get_Ny(), Nz = lbm.get_Nz();
	std::vector<T> device(R), host(R);
	for(ulong e = 0ull; e < R; e++) { device[e] = (T)(e % 251ull + 1ull); host[e] = (T)((e * 7ull) % 249ull + 2ull); }
	for(ulong e = 0ull; e < R; e++) m[e] = device[e];
	m.write_to_device();
	for(ulong e = 0ull; e < R; e++) m[e] = host[e];
	// the box of each form (1-D rows run over component ends and the buffer end)
	ulong x0 = 2ull, x1 = 9ull, y0 = 1ull, y1 = 6ull, z0 = 1ull, z1 = 4ull, nx = Nx, ny = Ny;
	if(form == 1) { x0 = N - 20ull; x1 = N + 35ull; y0 = 0ull; y1 = 1ull; z0 = 0ull; z1 = 1ull; nx = N; ny = 1ull; }
	if(form == 2) { x0 = 1ull; x1 = 11ull; y0 = 3ull; y1 = Ny * Nz + 1ull; z0 = 0ull; z1 = 1ull; ny = Ny * Nz; }
	if(form == 1) write ? m.write_to_device_1d(x0, x1, dimension) : m.read_from_device_1d(x0, x1, dimension);
	else if(form == 2) write ? m.write_to_device_2d(x0, x1, y0, y1, nx, ny, dimension) : m.read_from_device_2d(x0, x1, y0, y1, nx, ny, dimension);
	else write ? m.write_to_device_3d(x0, x1, y0, y1, z0, z1, Nx, Ny, Nz, dimension) : m.read_from_device_3d(x0, x1, y0, y1, z0, z1, Nx, Ny, Nz, dimension);
	std::vector<char> moved(R, 0);
	for(ulong z = z0; z < z1; z++) for(ulong y = y0; y < y1; y++) {
		const ulong n = x0 + (y + z * ny) * nx;
		for(ulong i = dimension < 0 ? 0ull : (ulong)dimension; i < (dimension < 0 ? D : (ulong)dimension + 1ull); i++) {
			const ulong off = min(i * N + n, R), len = min(x1 - x0, R - off);
			for(ulong e = off; e < off + len; e++) moved[e] = 1;
		}
	}
	ulong bad = 0ull;
	if(write) {
		for(ulong e = 0ull; e < R; e++) m[e] = (T)0;
		m.read_from_device();
		for(ulong e = 0ull; e < R; e++) bad += m[e] != (moved[e] ? host[e] : device[e]);
	} else {
		for(ulong e = 0ull; e < R; e++) bad += m[e] != (moved[e] ? device[e] : host[e]);
	}
	return bad;
}
void main_setup() {
	LBM lbm(12u, 7u, 5u, 0.1f);
	ulong bad = 0ull, cases = 0ull;
	for(int write = 0; write < 2; write++) for(int form = 0; form < 3; form++) {
		for(int dimension = -1; dimension < 3; dimension++) { bad += box_case(lbm, lbm.u, write != 0, form, dimension); cases++; }
		bad += box_case(lbm, lbm.flags, write != 0, form, -1); cases++;
	}
	println("box transfers: " + to_string(cases) + " cases, " + to_string(bad) + " wrong elements");
	// position -> cell: every cell centre maps back to its cell, also shifted by whole periods
	ulong wrong = 0ull;
	for(ulong n = 0ull; n < lbm.get_N(); n++) {
		const float3 p = lbm.position(n);
		wrong += lbm.index(p) != n;
		wrong += lbm.index(float3(p.x + (float)lbm.get_Nx(), p.y - 2.0f * (float)lbm.get_Ny(), p.z + 3.0f * (float)lbm.get_Nz())) != n;
	}
	println("positions: " + to_string(wrong) + " wrong cells");
}

#else
void main_setup() {
	print_error("no scene selected: compile with -DBENCHMARK, -DSCENE_TAYLOR_GREEN, -DSCENE_SPHERE, -DSCENE_CAVITY or -DSCENE_POROUS_CHANNEL");
}
#endif
