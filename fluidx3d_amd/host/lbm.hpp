#pragma once
// The LBM host class with the reference's interface (src/lbm.hpp:205-611) on top of the
// C ABI in include/fx_lbm.h.  A lattice may be split into Dx*Dy*Dz domains with one halo
// layer per split side, as in the reference (src/lbm.cpp:721-773); all domains run on
// device 0, each with its own context (stream), and exchange halos device to device.
#include "defines.hpp"
#include "utilities.hpp"
#include "units.hpp"
#include "shapes.hpp"
#include "info.hpp"
#include "fx_lbm.h"

uint bytes_per_cell_host();             // rho, u, flags (+F, +porosity), src/lbm.cpp:25-37
uint bytes_per_cell_device();           // src/lbm.cpp:38-50
uint bandwidth_bytes_per_cell_device(); // src/lbm.cpp:51-72
uint3 resolution(const float3 box_aspect_ratio, const uint memory); // grid for a VRAM budget in MB
fx_lbm_config lbm_config(const uint Nx, const uint Ny, const uint Nz, const float nu); // the compiled switches

class LBM;
void fx_check(const int status, const string& what); // nonzero status -> print_error, like the reference

// One box of a lattice split into Dx*Dy*Dz domains (LBM_Domain, src/lbm.cpp:96-176,
// 721-773): its own context (device, stream, kernel variant) and DDFs; Nx/Dx + 2 cells
// along a split axis (one halo layer per side), Nx along an unsplit one; (ox, oy, oz) are
// the global coordinates of local cell (0, 0, 0), -1 along a split axis.  The halo transfer
// buffers hold one side's values of the largest field per split axis.
struct LBM_Domain {
	fx_ctx* ctx = nullptr;
	void* fi = nullptr;
	fx_lbm_config cfg{};
	uint Nx = 1u, Ny = 1u, Nz = 1u;
	int ox = 0, oy = 0, oz = 0;
	void* transfer_p[3] = {nullptr, nullptr, nullptr};
	void* transfer_m[3] = {nullptr, nullptr, nullptr};
	ulong get_N() const { return (ulong)Nx * (ulong)Ny * (ulong)Nz; }
};

// Host buffer + device buffer of one field; the single-domain Memory_Container of the
// reference (src/lbm.hpp:239-408).  Components are stored one after another (u.x[n] at
// n, u.y[n] at N+n, u.z[n] at 2N+n).
template<typename T> class Memory_Container {
	LBM* lbm = nullptr;
	ulong N = 0ull;
	uint d = 1u;
	string name;
	vector<T> host;
	vector<void*> devices; // one buffer per domain (local cells, halos included)
	void* device = nullptr; // devices[0]
	void write_vtk(const string& path, const bool convert_to_si_units);
	void refresh(); // update_fields when rho/u on the device are stale
	// global host array <-> one domain's local layout: every local cell (halos included,
	// wrapped periodically) from the host, interior cells back to it
	void gather(const LBM_Domain& dom, vector<T>& local) const;
	void scatter(const LBM_Domain& dom, const vector<T>& local);
	void box(const bool to_device, const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong z0, const ulong z1, const ulong Nx, const ulong Ny, const int dimension, const bool blocking);
public:
	class Pointer {
		Memory_Container* m = nullptr;
		uint dimension = 0u;
	public:
		Pointer() {}
		Pointer(Memory_Container* m, const uint dimension) : m(m), dimension(dimension) {}
		T& operator[](const ulong i) { return m->host[i + (ulong)dimension * m->N]; }
		const T& operator[](const ulong i) const { return m->host[i + (ulong)dimension * m->N]; }
	};
	Pointer x, y, z;

	Memory_Container() {}
	Memory_Container(const Memory_Container&) = delete;
	Memory_Container& operator=(const Memory_Container&) = delete;
	void allocate(LBM* lbm, const ulong N, const uint d, const string& name, const T fill);
	void release();

	T& operator[](const ulong i) { return host[i]; }
	const T& operator[](const ulong i) const { return host[i]; }
	T operator()(const ulong i) const { return host[i]; }
	T operator()(const ulong i, const uint dimension) const { return host[i + (ulong)dimension * N]; }
	ulong length() const { return N; }
	uint dimensions() const { return d; }
	ulong range() const { return N * (ulong)d; }
	ulong capacity() const { return N * (ulong)d * sizeof(T); }
	T* data() { return host.data(); }
	void* device_buffer() const { return device; }
	void* device_buffer(const uint domain) const { return devices[domain]; }
	void reset(const T value = (T)0) { std::fill(host.begin(), host.end(), value); }
	void read_from_device();  // refreshes rho/u on the device first when stale (src/lbm.hpp:390-396)
	void write_to_device();
	// box transfers to/from the same offsets of the host array, one component (dimension >= 0)
	// or all; Memory<T>::read_from_device_1d/2d/3d, write_to_device_1d/2d/3d
	// (src/opencl.hpp:532-607)
	void read_from_device_1d(const ulong x0, const ulong x1, const int dimension = -1, const bool blocking = true);
	void write_to_device_1d(const ulong x0, const ulong x1, const int dimension = -1, const bool blocking = true);
	void read_from_device_2d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong Nx, const ulong Ny, const int dimension = -1, const bool blocking = true);
	void write_to_device_2d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong Nx, const ulong Ny, const int dimension = -1, const bool blocking = true);
	void read_from_device_3d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong z0, const ulong z1, const ulong Nx, const ulong Ny, const ulong Nz, const int dimension = -1, const bool blocking = true);
	void write_to_device_3d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong z0, const ulong z1, const ulong Nx, const ulong Ny, const ulong Nz, const int dimension = -1, const bool blocking = true);
	void write_host_to_vtk(const string& path = "", const bool convert_to_si_units = true);
	void write_device_to_vtk(const string& path = "", const bool convert_to_si_units = true);
};

class LBM {
	uint Nx = 1u, Ny = 1u, Nz = 1u;
	uint Dx = 1u, Dy = 1u, Dz = 1u;
	float nu = 1.0f / 6.0f;
	float fx = 0.0f, fy = 0.0f, fz = 0.0f;
	ulong t = 0ull;
	bool initialized = false;
	ulong t_last_update_fields = max_ulong;
	ulong t_last_force_field = max_ulong;
	fx_ctx* ctx = nullptr; // domain[0].ctx
	void* fi = nullptr;    // domain[0].fi: DDFs, device only
	fx_lbm_config cfg{};   // domain[0].cfg
	fx_adjoint* adjoint = nullptr;
	vector<LBM_Domain> domain; // d = x + (y + z*Dy)*Dx

	void construct(const uint Nx, const uint Ny, const uint Nz, const uint Dx, const uint Dy, const uint Dz, const float nu, const float fx, const float fy, const float fz,
	               const float sigma, const float alpha, const float beta, const uint particles_N, const float particles_rho);
	void sanity_checks_initialization();
	void initialize();
	void update_fields_on_device(); // update_fields on every domain if stale
	void sync_domains();
	// halo exchange of one field after the extract kernels, axis by axis (x, y, z):
	// LBM::communicate_field (src/lbm.cpp:1355-1383), device to device
	void communicate(const uint field);
	uint neighbour(const uint d, const uint axis, const int step) const;
	void require_single_domain(const string& what) const;
	template<typename T> friend class Memory_Container;

public:
	Memory_Container<float> rho;   // density
	Memory_Container<float> u;     // velocity
	Memory_Container<uchar> flags; // cell type bits
#ifdef FORCE_FIELD
	Memory_Container<float> F;     // force per cell
#endif
#ifdef POROSITY
	Memory_Container<float> porosity; // Brinkman porosity alpha per cell (not in the reference)
	Memory_Container<float> porosity_gradient; // dJ/dporosity after backward()
#endif

	LBM(const uint Nx, const uint Ny, const uint Nz, const uint Dx, const uint Dy, const uint Dz, const float nu, const float fx = 0.0f, const float fy = 0.0f, const float fz = 0.0f, const float sigma = 0.0f, const float alpha = 0.0f, const float beta = 0.0f, const uint particles_N = 0u, const float particles_rho = 0.0f);
	LBM(const uint Nx, const uint Ny, const uint Nz, const float nu, const float fx = 0.0f, const float fy = 0.0f, const float fz = 0.0f, const float sigma = 0.0f, const float alpha = 0.0f, const float beta = 0.0f, const uint particles_N = 0u, const float particles_rho = 1.0f);
	LBM(const uint Nx, const uint Ny, const uint Nz, const float nu, const uint particles_N, const float particles_rho = 1.0f);
	LBM(const uint Nx, const uint Ny, const uint Nz, const float nu, const float fx, const float fy, const float fz, const uint particles_N, const float particles_rho = 1.0f);
	LBM(const uint3 N, const uint Dx, const uint Dy, const uint Dz, const float nu, const float fx = 0.0f, const float fy = 0.0f, const float fz = 0.0f, const float sigma = 0.0f, const float alpha = 0.0f, const float beta = 0.0f, const uint particles_N = 0u, const float particles_rho = 0.0f);
	LBM(const uint3 N, const float nu, const float fx = 0.0f, const float fy = 0.0f, const float fz = 0.0f, const float sigma = 0.0f, const float alpha = 0.0f, const float beta = 0.0f, const uint particles_N = 0u, const float particles_rho = 1.0f);
	LBM(const uint3 N, const float nu, const uint particles_N, const float particles_rho = 1.0f);
	LBM(const uint3 N, const float nu, const float fx, const float fy, const float fz, const uint particles_N, const float particles_rho = 1.0f);
	LBM(const LBM&) = delete;
	LBM& operator=(const LBM&) = delete;
	~LBM();

	void run(const ulong steps = max_ulong, const ulong total_steps = max_ulong); // initialize on first call, then steps
	void update_fields(); // rho/u from the DDFs at the current step
#ifdef MOVING_BOUNDARIES
	void update_moving_boundaries(); // mark/unmark cells next to TYPE_S cells with velocity!=0 with TYPE_MS
#endif
	void reset();         // next run() re-initializes from the host fields
#ifdef FORCE_FIELD
	void update_force_field();
	float3 object_center_of_mass(const uchar flag_marker = TYPE_S);
	float3 object_force(const uchar flag_marker = TYPE_S);
	float3 object_torque(const float3& rotation_center, const uchar flag_marker = TYPE_S);
#endif
	double total_mass(); // FP64 mass invariant of the DDFs (not in the reference)
#ifdef POROSITY
	// Differentiation (not in the reference).  forward() re-initializes from the host fields,
	// runs `steps` steps from t=0 with DDF checkpoints every checkpoint_interval steps
	// (0 = ceil(sqrt(steps))) and returns J = mean over t >= window_begin of the x-momentum
	// through the plane x = plane_x.  backward() runs the checkpointed adjoint and leaves
	// dJ/dporosity in porosity_gradient (host and device).
	double forward(const ulong steps, const ulong window_begin, const uint plane_x, const ulong checkpoint_interval = 0ull);
	void backward();
	float backward_ms = 0.0f, forward_ms = 0.0f; // GPU time of the last forward()/backward()
#endif

	uint get_Nx() const { return Nx; }
	uint get_Ny() const { return Ny; }
	uint get_Nz() const { return Nz; }
	ulong get_N() const { return (ulong)Nx * (ulong)Ny * (ulong)Nz; }
	uint get_Dx() const { return Dx; }
	uint get_Dy() const { return Dy; }
	uint get_Dz() const { return Dz; }
	uint get_D() const { return Dx * Dy * Dz; }
	float get_nu() const { return nu; }
	float get_tau() const { return 3.0f * nu + 0.5f; }
	float get_w() const { return cfg.w; } // relaxation rate the kernels use
	float get_Re_max() const { return 0.57735027f * std::sqrt((float)(sq(Nx) + sq(Ny) + sq(Nz))) / nu; }
	float get_fx() const { return fx; }
	float get_fy() const { return fy; }
	float get_fz() const { return fz; }
	ulong get_t() const { return t; }
	uint get_velocity_set() const { return cfg.velocity_set; }
	void set_fx(const float v) { fx = v; }
	void set_fy(const float v) { fy = v; }
	void set_fz(const float v) { fz = v; }
	void set_f(const float x, const float y, const float z) { fx = x; fy = y; fz = z; }
	fx_ctx* context() const { return ctx; }      // the C ABI context (device, stream, kernels)
	void* ddf_buffer() const { return fi; }      // device DDF array fi[i*N+n]

	void coordinates(const ulong n, uint& x, uint& y, uint& z) const {
		const ulong r = n % ((ulong)Nx * (ulong)Ny);
		x = (uint)(r % (ulong)Nx);
		y = (uint)(r / (ulong)Nx);
		z = (uint)(n / ((ulong)Nx * (ulong)Ny));
	}
	ulong index(const uint x, const uint y, const uint z) const { return (ulong)x + ((ulong)y + (ulong)z * (ulong)Ny) * (ulong)Nx; }
	ulong index(const uint3 xyz) const { return index(xyz.x, xyz.y, xyz.z); }
	// position -> closest grid cell, positions wrapped into the periodic box
	// (src/lbm.hpp:484-500, 509-515)
	float3 mirror_position(const float3& p) const {
		auto wrap = [](const float v, const float n) { return (v >= 0.0f ? 1.0f : -1.0f) * (std::fmod(std::fabs(v) + 0.5f * n, n) - 0.5f * n); };
		return float3(wrap(p.x, (float)Nx), wrap(p.y, (float)Ny), wrap(p.z, (float)Nz));
	}
	void coordinates(const float3& p, uint& x, uint& y, uint& z) const {
		const float3 mp = mirror_position(p);
		x = (uint)(mp.x + 1.5f * (float)Nx) % Nx;
		y = (uint)(mp.y + 1.5f * (float)Ny) % Ny;
		z = (uint)(mp.z + 1.5f * (float)Nz) % Nz;
	}
	ulong index(const float3& p) const { uint x = 0u, y = 0u, z = 0u; coordinates(p, x, y, z); return index(x, y, z); }
	float3 position(const uint x, const uint y, const uint z) const { return float3((float)x + 0.5f - 0.5f * (float)Nx, (float)y + 0.5f - 0.5f * (float)Ny, (float)z + 0.5f - 0.5f * (float)Nz); }
	float3 position(const ulong n) const { uint x, y, z; coordinates(n, x, y, z); return position(x, y, z); }
	float3 relative_position(const uint x, const uint y, const uint z) const { return float3(((float)x + 0.5f) / (float)Nx - 0.5f, ((float)y + 0.5f) / (float)Ny - 0.5f, ((float)z + 0.5f) / (float)Nz - 0.5f); }
	float3 relative_position(const ulong n) const { uint x, y, z; coordinates(n, x, y, z); return relative_position(x, y, z); }
	float3 size() const { return float3((float)Nx, (float)Ny, (float)Nz); }
	float3 center() const { return float3(0.5f * (float)Nx - 0.5f, 0.5f * (float)Ny - 0.5f, 0.5f * (float)Nz - 0.5f); }
	uint smallest_side_length() const { return min(min(Nx, Ny), Nz); }
	uint largest_side_length() const { return max(max(Nx, Ny), Nz); }
	void write_status(const string& path = "");
};

template<typename T> void Memory_Container<T>::allocate(LBM* l, const ulong n, const uint dims, const string& field_name, const T fill) {
	lbm = l; N = n; d = dims; name = field_name;
	if(N * (ulong)d == 0ull) print_error("Memory size must be larger than 0.");
	host.assign(N * (ulong)d, fill);
	devices.assign(lbm->domain.size(), nullptr);
	for(size_t k = 0u; k < lbm->domain.size(); k++)
		fx_check(fx_buf_alloc(lbm->domain[k].ctx, lbm->domain[k].get_N() * (ulong)d * sizeof(T), &devices[k]), "allocating " + name);
	device = devices[0];
	x = Pointer(this, 0u);
	if(d > 1u) y = Pointer(this, 1u);
	if(d > 2u) z = Pointer(this, 2u);
}
template<typename T> void Memory_Container<T>::release() {
	if(lbm) for(size_t k = 0u; k < devices.size(); k++) if(devices[k] && lbm->domain[k].ctx) fx_buf_free(lbm->domain[k].ctx, devices[k]);
	devices.clear();
	device = nullptr;
}
template<typename T> void Memory_Container<T>::gather(const LBM_Domain& dom, vector<T>& local) const {
	const ulong Nx = lbm->Nx, Ny = lbm->Ny, Nz = lbm->Nz, NL = dom.get_N();
	local.resize(NL * (ulong)d);
	parallel_for((ulong)dom.Nz, [&](ulong lz) {
		const ulong gz = (ulong)(((long long)lz + dom.oz + (long long)Nz) % (long long)Nz);
		for(uint ly = 0u; ly < dom.Ny; ly++) {
			const ulong gy = (ulong)(((long long)ly + dom.oy + (long long)Ny) % (long long)Ny);
			for(uint lx = 0u; lx < dom.Nx; lx++) {
				const ulong gx = (ulong)(((long long)lx + dom.ox + (long long)Nx) % (long long)Nx);
				const ulong g = gx + (gy + gz * Ny) * Nx, l = (ulong)lx + ((ulong)ly + lz * (ulong)dom.Ny) * (ulong)dom.Nx;
				for(uint c = 0u; c < d; c++) local[l + (ulong)c * NL] = host[g + (ulong)c * N];
			}
		}
	});
}
template<typename T> void Memory_Container<T>::scatter(const LBM_Domain& dom, const vector<T>& local) {
	const ulong Nx = lbm->Nx, Ny = lbm->Ny, NL = dom.get_N();
	const uint hx = lbm->Dx > 1u, hy = lbm->Dy > 1u, hz = lbm->Dz > 1u;
	parallel_for((ulong)(dom.Nz - 2u * hz), [&](ulong iz) {
		const ulong lz = iz + hz, gz = (ulong)((long long)lz + dom.oz);
		for(uint ly = hy; ly < dom.Ny - hy; ly++) {
			const ulong gy = (ulong)((long long)ly + dom.oy);
			for(uint lx = hx; lx < dom.Nx - hx; lx++) {
				const ulong gx = (ulong)((long long)lx + dom.ox);
				const ulong g = gx + (gy + gz * Ny) * Nx, l = (ulong)lx + ((ulong)ly + lz * (ulong)dom.Ny) * (ulong)dom.Nx;
				for(uint c = 0u; c < d; c++) host[g + (ulong)c * N] = local[l + (ulong)c * NL];
			}
		}
	});
}
template<typename T> void Memory_Container<T>::read_from_device() {
	refresh();
	if(lbm->domain.size() == 1u) {
		fx_check(fx_buf_read(lbm->ctx, device, 0ull, capacity(), host.data(), 1), "reading " + name);
		return;
	}
	vector<T> local;
	for(size_t k = 0u; k < lbm->domain.size(); k++) {
		const LBM_Domain& dom = lbm->domain[k];
		local.resize(dom.get_N() * (ulong)d);
		fx_check(fx_buf_read(dom.ctx, devices[k], 0ull, local.size() * sizeof(T), local.data(), 1), "reading " + name);
		scatter(dom, local);
	}
}
template<typename T> void Memory_Container<T>::refresh() {
#ifndef UPDATE_FIELDS
	lbm->update_fields_on_device();
#endif
}
template<typename T> void Memory_Container<T>::box(const bool to_device, const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong z0, const ulong z1, const ulong Nx, const ulong Ny, const int dimension, const bool blocking) {
	lbm->require_single_domain("box transfers of " + name);
	if(to_device) fx_check(fx_buf_write_box(lbm->ctx, device, host.data(), (uint)sizeof(T), N, d, Nx, Ny, x0, x1, y0, y1, z0, z1, dimension, blocking ? 1 : 0), "writing " + name);
	else {
		refresh();
		fx_check(fx_buf_read_box(lbm->ctx, device, host.data(), (uint)sizeof(T), N, d, Nx, Ny, x0, x1, y0, y1, z0, z1, dimension, blocking ? 1 : 0), "reading " + name);
	}
}
template<typename T> void Memory_Container<T>::read_from_device_1d(const ulong x0, const ulong x1, const int dimension, const bool blocking) { box(false, x0, x1, 0ull, 1ull, 0ull, 1ull, N, 1ull, dimension, blocking); }
template<typename T> void Memory_Container<T>::write_to_device_1d(const ulong x0, const ulong x1, const int dimension, const bool blocking) { box(true, x0, x1, 0ull, 1ull, 0ull, 1ull, N, 1ull, dimension, blocking); }
template<typename T> void Memory_Container<T>::read_from_device_2d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong Nx, const ulong Ny, const int dimension, const bool blocking) { box(false, x0, x1, y0, y1, 0ull, 1ull, Nx, Ny, dimension, blocking); }
template<typename T> void Memory_Container<T>::write_to_device_2d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong Nx, const ulong Ny, const int dimension, const bool blocking) { box(true, x0, x1, y0, y1, 0ull, 1ull, Nx, Ny, dimension, blocking); }
template<typename T> void Memory_Container<T>::read_from_device_3d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong z0, const ulong z1, const ulong Nx, const ulong Ny, const ulong Nz, const int dimension, const bool blocking) { (void)Nz; box(false, x0, x1, y0, y1, z0, z1, Nx, Ny, dimension, blocking); }
template<typename T> void Memory_Container<T>::write_to_device_3d(const ulong x0, const ulong x1, const ulong y0, const ulong y1, const ulong z0, const ulong z1, const ulong Nx, const ulong Ny, const ulong Nz, const int dimension, const bool blocking) { (void)Nz; box(true, x0, x1, y0, y1, z0, z1, Nx, Ny, dimension, blocking); }
template<typename T> void Memory_Container<T>::write_to_device() {
	if(lbm->domain.size() == 1u) {
		fx_check(fx_buf_write(lbm->ctx, device, 0ull, capacity(), host.data(), 1), "writing " + name);
		return;
	}
	vector<T> local;
	for(size_t k = 0u; k < lbm->domain.size(); k++) {
		gather(lbm->domain[k], local);
		fx_check(fx_buf_write(lbm->domain[k].ctx, devices[k], 0ull, local.size() * sizeof(T), local.data(), 1), "writing " + name);
	}
}
// `path` is a prefix, by default the executable's export/ folder: the file is
// <path><name>-<t as 9 digits>.vtk (default_filename, src/lbm.hpp:398-404)
template<typename T> void Memory_Container<T>::write_host_to_vtk(const string& path, const bool convert_to_si_units) {
	write_vtk(default_filename(path, name, ".vtk", lbm->get_t()), convert_to_si_units);
}
template<typename T> void Memory_Container<T>::write_device_to_vtk(const string& path, const bool convert_to_si_units) {
	read_from_device();
	write_host_to_vtk(path, convert_to_si_units);
}
// Binary legacy VTK STRUCTURED_POINTS, big-endian, components interleaved per point, the
// format of the reference's write_vtk (src/lbm.hpp:298-340).
template<typename T> void Memory_Container<T>::write_vtk(const string& path, const bool convert_to_si_units) {
	const uint nx = lbm->get_Nx(), ny = lbm->get_Ny(), nz = lbm->get_Nz();
	float spacing = 1.0f, factor = 1.0f;
	if(convert_to_si_units) {
		spacing = units.si_x(1.0f);
		if(name == "rho") factor = units.si_rho(1.0f);
		if(name == "u") factor = units.si_u(1.0f);
		if(name == "F") factor = units.si_F(1.0f);
	}
	const char* type = std::is_same<T, float>::value ? "float" : std::is_same<T, uchar>::value ? "unsigned_char" : "double";
	std::ostringstream header;
	header << "# vtk DataFile Version 3.0\nFluidX3D " << path.substr(path.rfind('/') + 1) << "\nBINARY\nDATASET STRUCTURED_POINTS\n"
	       << "DIMENSIONS " << nx << " " << ny << " " << nz << "\n"
	       << "ORIGIN " << to_string(spacing * (0.5f - 0.5f * (float)nx)) << " " << to_string(spacing * (0.5f - 0.5f * (float)ny)) << " " << to_string(spacing * (0.5f - 0.5f * (float)nz)) << "\n"
	       << "SPACING " << to_string(spacing) << " " << to_string(spacing) << " " << to_string(spacing) << "\n"
	       << "POINT_DATA " << N << "\nSCALARS data " << type << " " << d << "\nLOOKUP_TABLE default\n";
	create_folder(path);
	FILE* file = std::fopen(path.c_str(), "wb");
	if(!file) print_error("cannot open \"" + path + "\" for writing");
	const string h = header.str();
	std::fwrite(h.data(), 1, h.size(), file);
	vector<T> chunk;
	const ulong per_chunk = 1ull << 20;
	for(ulong a = 0ull; a < N; a += per_chunk) {
		const ulong b = std::min(N, a + per_chunk);
		chunk.resize((b - a) * d);
		for(ulong n = a; n < b; n++) {
			for(uint c = 0u; c < d; c++) {
				T v = host[n + (ulong)c * N];
				if(!std::is_same<T, uchar>::value) v = (T)(factor * (float)v);
				uchar* bytes = (uchar*)&v;
				std::reverse(bytes, bytes + sizeof(T)); // big-endian
				chunk[(n - a) * d + c] = v;
			}
		}
		std::fwrite(chunk.data(), sizeof(T), chunk.size(), file);
	}
	std::fclose(file);
	print_info("File \"" + path + "\" saved.");
}
