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
