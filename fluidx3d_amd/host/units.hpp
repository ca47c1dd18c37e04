#pragma once
// Unit conversion between SI and lattice units and the closed forms used by validation
// scenes (reference interface: src/units.hpp).  Lattice units: dx = dt = rho0 = 1.
#include "utilities.hpp"

class Units {
public:
	float m = 1.0f, kg = 1.0f, s = 1.0f; // one lattice length/mass/time unit in SI
	// the SI values of lattice length x, velocity u and density rho fix m, kg and s
	void set_m_kg_s(const float x, const float u, const float rho, const float si_x, const float si_u, const float si_rho) {
		m = si_x / x;
		kg = si_rho / rho * cb(m);
		s = u / si_u * m;
	}
	float x(const float si_x) const { return si_x / m; }
	float u(const float si_u) const { return si_u * s / m; }
	float rho(const float si_rho) const { return si_rho * cb(m) / kg; }
	float nu(const float si_nu) const { return si_nu * s / sq(m); }
	float t(const float si_t) const { return si_t / s; }
	float f(const float si_f) const { return si_f * sq(m) * sq(s) / kg; } // force per volume
	float F(const float si_F) const { return si_F * sq(s) / (kg * m); }
	float si_x(const float x) const { return x * m; }
	float si_u(const float u) const { return u * m / s; }
	float si_rho(const float rho) const { return rho * kg / cb(m); }
	float si_nu(const float nu) const { return nu * sq(m) / s; }
	float si_t(const ulong t) const { return (float)t * s; }
	float si_F(const float F) const { return F * kg * m / sq(s); }
	float si_T(const float T) const { return T; }

	float nu_from_tau(const float tau) const { return (tau - 0.5f) / 3.0f; }
	float tau_from_nu(const float nu) const { return 3.0f * nu + 0.5f; }
	float nu_from_Re(const float Re, const float x, const float u) const { return x * u / Re; }
	float u_from_Re(const float Re, const float x, const float nu) const { return Re * nu / x; }
	float si_Re(const float si_x, const float si_u, const float si_nu) const { return si_x * si_u / si_nu; }
	// plane / pipe Poiseuille: centre velocity u for body force f, half width (radius) R
	float f_from_u_Poiseuille_2D(const float u, const float rho, const float nu, const float R) const { return 2.0f * u * rho * nu / sq(R); }
	float f_from_u_Poiseuille_3D(const float u, const float rho, const float nu, const float R) const { return 4.0f * u * rho * nu / sq(R); }
	float u_from_f_Poiseuille_2D(const float f, const float rho, const float nu, const float R) const { return f * sq(R) / (2.0f * rho * nu); }
	float u_from_f_Poiseuille_3D(const float f, const float rho, const float nu, const float R) const { return f * sq(R) / (4.0f * rho * nu); }
	// Stokes drag on a sphere, F = 6 pi rho nu R u
	float F_Stokes(const float rho, const float u, const float nu, const float R) const { return 6.0f * pif * rho * nu * R * u; }
};
extern Units units;
