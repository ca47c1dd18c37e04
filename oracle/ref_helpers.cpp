// Exposes a few of the reference's own host helpers through a C ABI so the tests can pin
// the oracle's restatements against them.  TEST INFRASTRUCTURE ONLY; compiled from the
// reference's sources where they lie (see oracle/Makefile, target "ref"), output only into
// oracle/_ref/.  Nothing here is part of the product.
//   src/utilities.hpp:121-144  host FP16 / FP16C conversions
//   src/utilities.hpp:244-250  LCG random / random_symmetric
//   src/utilities.hpp:2745-2754 to_string(float) (the def_w text, src/lbm.cpp:367)
//   src/shapes.cpp:5-8          sphere()
//   src/units.hpp:108-134       Poiseuille and Stokes closed forms
#include "utilities.hpp"
#include "shapes.hpp"
#include "units.hpp"
#include <cstring>

extern "C" {
int ref_to_string_float(float x, char* buf, int size) {
	const string s = to_string(x);
	if((int)s.size() + 1 > size) return -1;
	std::memcpy(buf, s.c_str(), s.size() + 1);
	return (int)s.size();
}
float ref_tau(float nu) { return 3.0f * nu + 0.5f; } // src/lbm.hpp:132, same host flags
float ref_half_to_float_custom(unsigned short x) { return half_to_float_custom(x); }
unsigned short ref_float_to_half_custom(float x) { return float_to_half_custom(x); }
float ref_half_to_float(unsigned short x) { return half_to_float(x); }
float ref_random(unsigned int* seed, float x) { uint s = *seed; const float r = random(s, x); *seed = s; return r; }
float ref_random_symmetric(unsigned int* seed, float x) { uint s = *seed; const float r = random_symmetric(s, x); *seed = s; return r; }
int ref_sphere(unsigned int x, unsigned int y, unsigned int z, float px, float py, float pz, float r) { return sphere(x, y, z, float3(px, py, pz), r) ? 1 : 0; }
float ref_f_from_u_Poiseuille_2D(float u, float rho, float nu, float R) { Units un; return un.f_from_u_Poiseuille_2D(u, rho, nu, R); }
float ref_F_Stokes(float rho, float u, float nu, float R) { Units un; return un.F_Stokes(rho, u, nu, R); }
}

// more of the reference's host helpers, for the host layer's restatements
//   src/utilities.hpp:2755-2786 to_string(double), to_string(float/double, decimals)
//   src/utilities.hpp:208-213, 330-335 converged(), 2876-2888 print_time()
//   src/units.hpp:114-134 rectangular duct, Stokes velocity/density/force
extern "C" {
static int copy_text(const string& s, char* buf, int size) {
	if((int)s.size() + 1 > size) return -1;
	std::memcpy(buf, s.c_str(), s.size() + 1);
	return (int)s.size();
}
int ref_to_string_double(double x, char* buf, int size) { return copy_text(to_string(x), buf, size); }
int ref_to_string_float_decimals(float x, unsigned int decimals, char* buf, int size) { return copy_text(to_string(x, decimals), buf, size); }
int ref_to_string_double_decimals(double x, unsigned int decimals, char* buf, int size) { return copy_text(to_string(x, decimals), buf, size); }
int ref_print_time(double t, char* buf, int size) { return copy_text(print_time(t), buf, size); }
int ref_converged_double(double a, double b, double c, double threshold) { return converged(a, b, c, threshold) ? 1 : 0; }
int ref_converged_float(float a, float b, float c, float threshold) { return converged(a, b, c, threshold) ? 1 : 0; }
void ref_u_Stokes(float x, float y, float z, float ux, float uy, float uz, float R, float* out) {
	Units un;
	const float3 v = un.u_Stokes(float3(x, y, z), float3(ux, uy, uz), R);
	out[0] = v.x; out[1] = v.y; out[2] = v.z;
}
float ref_rho_Stokes(float x, float y, float z, float ux, float uy, float uz, float R, float rho, float nu) { Units un; return un.rho_Stokes(float3(x, y, z), float3(ux, uy, uz), R, rho, nu); }
float ref_f_Stokes(float rho, float u, float nu, float R, float V) { Units un; return un.f_Stokes(rho, u, nu, R, V); }
float ref_f_from_u_rectangular_duct(float w, float h, float rho, float nu, float u) { Units un; return un.f_from_u_rectangular_duct(w, h, rho, nu, u); }
float ref_u_from_Re(float Re, float x, float nu) { Units un; return un.u_from_Re(Re, x, nu); }
}
