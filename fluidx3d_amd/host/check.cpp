// C exports of the host layer's number text, convergence test, file names and Units closed
// forms, so that they can be compared with the reference's own host helpers on the CPU
// (tests/test_host_text.py).  Built as bin/libhost_check.so; no GPU code.
#include "utilities.hpp"
#include "units.hpp"
#include <cstring>

Units units;

extern "C" {
static int copy_text(const string& s, char* buf, int size) {
	if((int)s.size() + 1 > size) return -1;
	std::memcpy(buf, s.c_str(), s.size() + 1);
	return (int)s.size();
}
int host_to_string_float(float x, char* buf, int size) { return copy_text(to_string(x), buf, size); }
int host_to_string_double(double x, char* buf, int size) { return copy_text(to_string(x), buf, size); }
int host_to_string_float_decimals(float x, unsigned int decimals, char* buf, int size) { return copy_text(to_string(x, decimals), buf, size); }
int host_to_string_double_decimals(double x, unsigned int decimals, char* buf, int size) { return copy_text(to_string(x, decimals), buf, size); }
int host_print_time(double t, char* buf, int size) { return copy_text(print_time(t), buf, size); }
int host_converged_double(double a, double b, double c, double threshold) { return converged(a, b, c, threshold) ? 1 : 0; }
int host_converged_float(float a, float b, float c, float threshold) { return converged(a, b, c, threshold) ? 1 : 0; }
int host_default_filename(const char* path, const char* name, const char* extension, unsigned long long t, char* buf, int size) {
	return copy_text(default_filename(path, name, extension, (ulong)t), buf, size);
}
void host_u_Stokes(float x, float y, float z, float ux, float uy, float uz, float R, float* out) {
	const float3 v = units.u_Stokes(float3(x, y, z), float3(ux, uy, uz), R);
	out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
float host_rho_Stokes(float x, float y, float z, float ux, float uy, float uz, float R, float rho, float nu) { return units.rho_Stokes(float3(x, y, z), float3(ux, uy, uz), R, rho, nu); }
float host_F_Stokes(float rho, float u, float nu, float R) { return units.F_Stokes(rho, u, nu, R); }
float host_f_Stokes(float rho, float u, float nu, float R, float V) { return units.f_Stokes(rho, u, nu, R, V); }
float host_f_from_u_rectangular_duct(float w, float h, float rho, float nu, float u) { return units.f_from_u_rectangular_duct(w, h, rho, nu, u); }
float host_u_from_Re(float Re, float x, float nu) { return units.u_from_Re(Re, x, nu); }
float host_f_from_u_Poiseuille_2D(float u, float rho, float nu, float R) { return units.f_from_u_Poiseuille_2D(u, rho, nu, R); }
}
