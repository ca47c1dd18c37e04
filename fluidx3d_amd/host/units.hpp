#pragma once
// Unit conversion between SI and lattice units, dimensionless numbers and the closed forms
// used by validation scenes.  The method names, signatures and formulas are those of the
// reference's Units class (src/units.hpp:5-160) so that its scenes compile and compute the
// same values; lattice units are dx = dt = rho0 = 1, c = 1/sqrt(3).
#include "utilities.hpp"

class Units {
	// one lattice unit of length / mass / time / temperature in SI
	float unit_m = 1.0f, unit_kg = 1.0f, unit_s = 1.0f, unit_K = 1.0f;
	void announce() const { print_info("Unit Conversion: 1 cell = " + to_string(1000.0f * si_x(1.0f), 3u) + " mm, 1 s = " + to_string(t(1.0f)) + " time steps"); }
public:
	// the SI values of a lattice length x, velocity u and density rho fix m, kg and s
	void set_m_kg_s(const float x, const float u, const float rho, const float si_x, const float si_u, const float si_rho) {
		unit_m = si_x / x;
		unit_kg = si_rho / rho * cb(unit_m);
		unit_s = u / si_u * unit_m;
		announce();
	}
	void set_m_kg_s(const float m, const float kg, const float s) { unit_m = m; unit_kg = kg; unit_s = s; announce(); }
	void set_m_kg_s_K(const float x, const float u, const float rho, const float T, const float si_x, const float si_u, const float si_rho, const float si_T) {
		unit_m = si_x / x;
		unit_kg = si_rho / rho * cb(unit_m);
		unit_s = u / si_u * unit_m;
		unit_K = si_T / T;
		announce();
	}
	void set_m_kg_s_K(const float m, const float kg, const float s, const float K) { unit_m = m; unit_kg = kg; unit_s = s; unit_K = K; announce(); }

	// SI -> lattice units
	float x(const float si_x) const { return si_x / unit_m; }
	float m(const float si_m) const { return si_m / unit_kg; }
	ulong t(const float si_t) const { return to_ulong(si_t / unit_s); }
	float frequency(const float si_frequency) const { return si_frequency * unit_s; }
	float omega(const float si_omega) const { return si_omega * unit_s; }
	float u(const float si_u) const { return si_u * unit_s / unit_m; }
	float rho(const float si_rho) const { return si_rho * cb(unit_m) / unit_kg; }
	float Q(const float si_Q) const { return si_Q * unit_s / cb(unit_m); }
	float nu(const float si_nu) const { return si_nu * unit_s / sq(unit_m); }
	float mu(const float si_mu) const { return si_mu * unit_s * unit_m / unit_kg; }
	float g(const float si_g) const { return si_g / unit_m * sq(unit_s); }
	float f(const float si_f) const { return si_f * sq(unit_m * unit_s) / unit_kg; } // force per volume
	float f(const float si_rho, const float si_g) const { return si_rho * si_g * sq(unit_m * unit_s) / unit_kg; }
	float F(const float si_F) const { return si_F * sq(unit_s) / (unit_kg * unit_m); }
	float M(const float si_M) const { return si_M * sq(unit_s) / (unit_kg * sq(unit_m)); } // torque
	float sigma(const float si_sigma) const { return si_sigma * sq(unit_s) / unit_kg; }
	float T(const float si_T) { return si_T / unit_K; }
	float alpha(const float si_alpha) { return si_alpha * unit_s / sq(unit_m); } // thermal diffusion coefficient
	float beta(const float si_beta) { return si_beta * unit_K; } // thermal expansion coefficient

	// lattice -> SI units
	float si_x(const uint x) const { return (float)x * unit_m; }
	float si_x(const float x) const { return x * unit_m; }
	float si_m(const float m) const { return m * unit_kg; }
	float si_t(const ulong t) const { return (float)t * unit_s; }
	float si_frequency(const float frequency) const { return frequency / unit_s; }
	float si_V(const float V) const { return V * cb(unit_m); }
	float si_u(const float u) const { return u * unit_m / unit_s; }
	float si_rho(const float rho) const { return rho * unit_kg / cb(unit_m); }
	float si_p(const float p) const { return p * unit_kg / (unit_m * sq(unit_s)); }
	float si_Q(const float Q) const { return Q * cb(unit_m) / unit_s; }
	float si_nu(const float nu) const { return nu * sq(unit_m) / unit_s; }
	float si_g(const float g) const { return g * unit_m / sq(unit_s); }
	float si_f(const float f) const { return f * unit_kg / sq(unit_m * unit_s); }
	float si_F(const float F) const { return F * unit_kg * unit_m / sq(unit_s); }
	float si_M(const float M) const { return M * unit_kg * sq(unit_m) / sq(unit_s); }
	float si_sigma(const float sigma) const { return sigma * unit_kg / sq(unit_s); }
	float si_T(const float T) { return T * unit_K; }
	float si_alpha(const float alpha) { return alpha * sq(unit_m) / unit_s; }
	float si_beta(const float beta) { return beta / unit_K; }

	// dimensionless numbers and relations in lattice units
	float Re(const float si_Re) const { return si_Re; }
	float Re(const float x, const float u, const float nu) const { return x * u / nu; }
	float Re(const float x, const float u, const float mu, const float rho) const { return x * u * rho / mu; }
	float We(const float x, const float u, const float rho, const float sigma) const { return x * sq(u) * rho / sigma; }
	float Fr(const float x, const float u, const float g) const { return u / std::sqrt(x * g); }
	float Ca(const float u, const float mu, const float sigma) const { return u * mu / sigma; }
	float Ca(const float u, const float rho, const float nu, const float sigma) const { return u * rho * nu / sigma; }
	float Bo(const float x, const float rho, const float g, const float sigma) const { return sq(x) * rho * g / sigma; }
	float Mo(const float rho, const float delta_rho, const float g, const float sigma, const float mu) const { return g * delta_rho * sq(sq(mu)) / (cb(sigma) * sq(rho)); }
	float Ga(const float x, const float nu, const float g) { return cb(x) * g / sq(nu); }
	float Ga(const float x, const float rho, const float nu, const float f) { return cb(x) * f / (sq(nu) * rho); }
	float Ma(const float u) const { return u / 0.57735027f; }
	float p_from_rho(const float rho) const { return (rho - 1.0f) / 3.0f; }
	float rho_from_p(const float p) const { return 1.0f + 3.0f * p; }
	float rho_laplace(const float sigma, const float R) { return 6.0f * sigma / R; }
	float rho_hydrostatic(const float f, const float z, const float h) { return 3.0f * f * (h - z) + 1.0f; }
	float nu_from_mu(const float mu, const float rho) const { return mu / rho; }
	float nu_from_tau(const float tau) const { return (tau - 0.5f) / 3.0f; }
	float tau_from_nu(const float nu) const { return 3.0f * nu + 0.5f; }
	float nu_from_Re(const float Re, const float x, const float u) const { return x * u / Re; }
	float f_from_F(const float F, const float V) const { return F / V; }
	float f_from_g(const float g, const float rho) const { return rho * g; }
	float g_from_f(const float f, const float rho) const { return f / rho; }
	float u_from_Re(const float Re, const float x, const float nu) const { return Re * nu / x; }
	float u_from_Re(const float Re, const float x, const float mu, const float rho) const { return Re * mu / (x * rho); }
	float u_from_Ma(const float Ma) const { return 0.57735027f * Ma; }
	float u_from_We(const float We, const float x, const float sigma, const float rho) const { return std::sqrt(We * sigma / (x * rho)); }
	float u_from_Fr(const float Fr, const float x, const float g) const { return Fr * std::sqrt(x * g); }
	float u_from_Ca(const float Ca, const float sigma, const float nu, const float rho) const { return Ca * sigma / (rho * nu); }
	float u_from_Ca(const float Ca, const float sigma, const float mu) const { return Ca * sigma / mu; }

	// plane (2D) and pipe (3D) Poiseuille flow: centre velocity u, body force f per volume,
	// half width / radius R, flow rate Q
	float u_from_f_Poiseuille_2D(const float f, const float rho, const float nu, const float R) const { return f * sq(R) / (2.0f * rho * nu); }
	float u_from_f_Poiseuille_3D(const float f, const float rho, const float nu, const float R) const { return f * sq(R) / (4.0f * rho * nu); }
	float u_from_f_Poiseuille_2D(const float Q, const float R) const { return 0.75f * Q / R; }
	float u_from_f_Poiseuille_3D(const float Q, const float R) const { return 2.0f / pif * Q / sq(R); }
	float f_from_u_Poiseuille_2D(const float u, const float rho, const float nu, const float R) const { return 2.0f * u * rho * nu / sq(R); }
	float f_from_u_Poiseuille_3D(const float u, const float rho, const float nu, const float R) const { return 4.0f * u * rho * nu / sq(R); }
	// rectangular duct -w/2 <= y <= w/2, -h/2 <= z <= h/2: body force for centre velocity u
	// from the Fourier series of the duct profile, 23 odd terms summed smallest first (FP64)
	float f_from_u_rectangular_duct(const float w, const float h, const float rho, const float nu, const float u) {
		const double scale = cb(pi) * (double)rho * (double)nu * (double)u / (4.0 * sq((double)h));
		const double aspect = pi * 0.5 * (double)w / (double)h, half_pi = pi * 0.5;
		double series = 0.0;
		for(int k = 45; k >= 1; k -= 2) {
			const double n = (double)k;
			series += (1.0 - 1.0 / std::cosh(n * aspect)) / cb(n) * std::sin(n * half_pi);
		}
		return (float)(scale / series);
	}
	// Stokes flow with far-field velocity u0 past a sphere of radius R at the origin:
	// velocity and density at point x (inside the sphere: 0 and rho), drag F = 6 pi rho nu R u
	float3 u_Stokes(const float3& x, const float3& u0, const float R) const {
		const float ir = 1.0f / length(x), ir3 = cb(ir), ir5 = sq(ir) * ir3, R2 = sq(R);
		if(!(1.0f / ir > R)) return u0 - u0;
		return u0 - (u0 * (ir + R2 * ir3 / 3.0f) + x * dot(u0, x) * (ir3 - R2 * ir5)) * 0.75f * R;
	}
	float rho_Stokes(const float3& x, const float3& u0, const float R, const float rho, const float nu) const {
		const float r = length(x);
		return r > R ? rho - 4.5f * rho * nu * R * dot(u0, x) / cb(r) : rho;
	}
	float f_Stokes(const float rho, const float u, const float nu, const float R, const float V) const { return 6.0f * pif * rho * nu * R * u / V; }
	float F_Stokes(const float rho, const float u, const float nu, const float R) const { return 6.0f * pif * rho * nu * R * u; }

	// SI dimensionless numbers and relations (usable before set_m_kg_s)
	float si_Re(const float Re) const { return Re; }
	float si_Re(const float si_x, const float si_u, const float si_nu) const { return si_x * si_u / si_nu; }
	float si_Re(const float si_x, const float si_u, const float si_mu, const float si_rho) const { return si_x * si_u * si_rho / si_mu; }
	float si_We(const float si_x, const float si_u, const float si_rho, const float si_sigma) const { return si_x * sq(si_u) * si_rho / si_sigma; }
	float si_Fr(const float si_x, const float si_u, const float si_g) const { return si_u / std::sqrt(si_x * si_g); }
	float si_Ca(const float si_u, const float si_mu, const float si_sigma) const { return si_u * si_mu / si_sigma; }
	float si_Ca(const float si_u, const float si_rho, const float si_nu, const float si_sigma) const { return si_u * si_rho * si_nu / si_sigma; }
	float si_Bo(const float si_x, const float si_rho, const float si_g, const float si_sigma) const { return sq(si_x) * si_rho * si_g / si_sigma; }
	float si_Mo(const float si_rho, const float si_delta_rho, const float si_g, const float si_sigma, const float si_mu) const { return Mo(si_rho, si_delta_rho, si_g, si_sigma, si_mu); }
	float si_Ga(const float si_x, const float si_nu, const float si_g) { return cb(si_x) * si_g / sq(si_nu); }
	float si_Ga(const float si_x, const float si_rho, const float si_nu, const float si_f) { return cb(si_x) * si_f / (sq(si_nu) * si_rho); }
	float si_nu_from_si_mu(const float si_mu, const float si_rho) const { return si_mu / si_rho; }
	float si_nu_from_si_Re(const float si_Re, const float si_x, const float si_u) const { return si_x * si_u / si_Re; }
	float si_mu_from_si_nu(const float si_nu, const float si_rho) const { return si_nu * si_rho; }
	float si_f_from_si_g(const float si_g, const float si_rho) const { return si_rho * si_g; }
	float si_g_from_si_f(const float si_f, const float si_rho) const { return si_f / si_rho; }
	float si_u_from_si_Re(const float si_Re, const float si_x, const float si_nu) const { return si_Re * si_nu / si_x; }
	float si_u_from_si_Re(const float si_Re, const float si_x, const float si_mu, const float si_rho) const { return si_Re * si_mu / (si_x * si_rho); }
	float si_u_from_si_We(const float si_We, const float si_x, const float si_sigma, const float si_rho) const { return std::sqrt(si_We * si_sigma / (si_x * si_rho)); }
	float si_u_from_si_Fr(const float si_Fr, const float si_x, const float si_g) const { return si_Fr * std::sqrt(si_x * si_g); }
	float si_u_from_si_h(const float si_h, const float si_g) const { return std::sqrt(2.0f * si_h * si_g); }
	float si_u_Poiseuille_2D(const float si_Q, const float si_R) const { return 0.75f * si_Q / si_R; }
	float si_u_Poiseuille_3D(const float si_Q, const float si_R) const { return 2.0f / pif * si_Q / sq(si_R); }
};
extern Units units;
