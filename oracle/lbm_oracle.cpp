// CPU oracle for the FluidX3D lattice Boltzmann hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain CPU restatement of the reference's
// initialize / stream_collide / update_fields kernels, its force-field and object-sum
// kernels, its FP16S/FP16C storage formats and its relaxation-rate formatting.  It is
// used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
// only there, as the checker.  The product (fluidx3d_amd) never links or calls it.
//
// Every function cites the reference lines it restates (paths relative to the reference
// repository root, FluidX3D v3.7).  Two arithmetic types are provided:
//   * float  : the reference's FP32 arithmetic, operation by operation.  Built with
//              -ffp-contract=off, so a*b+c is two roundings and fma() is one, exactly as
//              written in the reference source.  Division is IEEE correctly rounded.
//   * double : the FP64 oracle of the same scheme (same operation order, double storage,
//              no FP16 quantisation).  Used to state each storage mode's error.
// The FP64 adjoint of the forward step (transposed per-cell Jacobian applied in the
// reverse Esoteric-Pull order) lives at the end of this file.
//
// Memory layout matches the reference (src/kernel.cpp:861-863, 960-967):
//   fi[i*N+n]  (SoA DDFs), rho[n], u[d*N+n], flags[n], F[d*N+n], n = x+(y+z*Ny)*Nx.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <thread>
#include <vector>
#include <algorithm>
#include <string>
#include <type_traits>
#include <atomic>
#include <immintrin.h>

namespace {

enum : uint32_t {
	EXT_VOLUME_FORCE = 1u, EXT_FORCE_FIELD = 2u, EXT_EQUILIBRIUM_BOUNDARIES = 4u,
	EXT_UPDATE_FIELDS = 8u, EXT_POROSITY = 16u, EXT_PRECISE_DENSITY = 32u, EXT_MOVING_BOUNDARIES = 64u,
	EXT_SUBGRID = 128u
};
enum : uint32_t { ST_FP32 = 0u, ST_FP16S = 1u, ST_FP16C = 2u, ST_FP64 = 3u };
// cell type bits, src/defines.hpp:53-60 and src/lbm.cpp:399-411
constexpr uint8_t TYPE_S = 0x01, TYPE_E = 0x02, TYPE_MS = 0x03, TYPE_BO = 0x03, TYPE_T = 0x04, TYPE_SU = 0x38, TYPE_G = 0x20;

inline uint32_t fbits(float x) { uint32_t b; std::memcpy(&b, &x, 4); return b; }
inline float bitsf(uint32_t b) { float x; std::memcpy(&x, &b, 4); return x; }

// ---------------------------------------------------------------------------------------
// Storage formats
// ---------------------------------------------------------------------------------------

// IEEE-754 binary32 -> binary16, round to nearest, ties to even (the semantics of OpenCL
// vstore_half_rte, used by the FP16S store macro src/lbm.cpp:413).  Written from the
// IEEE-754 definition; overflow goes to infinity, NaN stays NaN.
uint16_t f32_to_f16_rne(float x) {
	const uint32_t b = fbits(x);
	const uint32_t sign = (b >> 16) & 0x8000u;
	const uint32_t a = b & 0x7FFFFFFFu;
	if(a > 0x7F800000u) return (uint16_t)(sign | 0x7E00u | ((a >> 13) & 0x3FFu)); // NaN
	if(a >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u); // >= 65520 rounds to infinity
	if(a >= 0x38800000u) { // result is a normal half (|x| >= 2^-14)
		const uint32_t r = a + 0x0FFFu + ((a >> 13) & 1u); // round mantissa to 10 bits, ties to even
		return (uint16_t)(sign | ((r - 0x38000000u) >> 13)); // rebias exponent 127 -> 15
	}
	// half subnormal k*2^-24: k = RNE(M * 2^(e-126)), M = 24-bit significand
	const uint32_t e = a >> 23;
	if(e < 102u) return (uint16_t)sign; // |x| < 2^-25: rounds to zero (2^-25 itself is a tie to 0)
	const uint32_t M = (a & 0x7FFFFFu) | 0x800000u;
	const uint32_t s = 126u - e; // 14..24
	const uint32_t q = M >> s, rem = M & ((1u << s) - 1u), half = 1u << (s - 1u);
	const uint32_t k = q + ((rem > half || (rem == half && (q & 1u))) ? 1u : 0u);
	return (uint16_t)(sign | k);
}
// binary16 -> binary32 (exact), the semantics of OpenCL vload_half (src/lbm.cpp:412).
float f16_to_f32(uint16_t h) {
	const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
	const uint32_t e = ((uint32_t)h >> 10) & 0x1Fu, m = (uint32_t)h & 0x3FFu;
	if(e == 0u) { // zero or subnormal: m*2^-24, exact in binary32
		const float v = (float)m * 5.9604644775390625e-8f;
		return bitsf(fbits(v) | sign);
	}
	if(e == 31u) return bitsf(sign | 0x7F800000u | (m << 13)); // inf / NaN
	return bitsf(sign | ((e + 112u) << 23) | (m << 13));
}

// FP16S storage: range-shifted IEEE half, load = half*2^-15, store = RNE(x*2^15)
// (src/lbm.cpp:410-414; 3.0517578E-5f is exactly 2^-15).  The lattice loops use the x86
// F16C conversions (IEEE binary16, round to nearest even), which fxo_f16_selfcheck()
// compares against f32_to_f16_rne / f16_to_f32 above for every binary32 input.
inline float fp16s_load(uint16_t h) { return _cvtsh_ss(h) * 3.0517578E-5f; }
inline uint16_t fp16s_store(float x) { return (uint16_t)_cvtss_sh(x * 32768.0f, _MM_FROUND_TO_NEAREST_INT); }

// FP16C storage: custom 1-4-11 format with exponent bias 15, device version
// src/kernel.cpp:848-859 (encode adds 0x800 before truncation, no saturation).
float fp16c_load(uint16_t x) {
	const uint32_t e = ((uint32_t)x & 0x7800u) >> 11;
	const uint32_t m = ((uint32_t)x & 0x07FFu) << 12;
	const uint32_t v = fbits((float)m) >> 23; // 127+floor(log2(m)) for m!=0
	const uint32_t sign = ((uint32_t)x & 0x8000u) << 16;
	const uint32_t normal = (e != 0u) ? (((e + 112u) << 23) | m) : 0u;
	const uint32_t denormal = (e == 0u && m != 0u) ? (((v - 37u) << 23) | ((m << (150u - v)) & 0x007FF000u)) : 0u;
	return bitsf(sign | normal | denormal);
}
uint16_t fp16c_store(float xf) {
	const uint32_t b = fbits(xf) + 0x00000800u;
	const uint32_t e = (b & 0x7F800000u) >> 23;
	const uint32_t m = b & 0x007FFFFFu;
	const uint32_t sign = (b & 0x80000000u) >> 16;
	const uint32_t normal = (e > 112u) ? ((((e - 112u) << 11) & 0x7800u) | (m >> 12)) : 0u;
	const uint32_t denormal = (e < 113u && e > 100u) ? ((((0x007FF800u + m) >> (124u - e)) + 1u) >> 1) : 0u;
	return (uint16_t)(sign | normal | denormal);
}

// ---------------------------------------------------------------------------------------
// Relaxation rate as the reference's kernels see it: to_string(1.0f/tau) printed with
// split_float's 9 significant digits (src/utilities.hpp:2599-2621, 2669-2675, 2745-2754)
// and parsed back as a float literal by the device compiler (src/lbm.cpp:367).
// ---------------------------------------------------------------------------------------
std::string format_float_like_reference(float x) {
	std::string s;
	if(x < 0.0f) { s += "-"; x = -x; }
	if(std::isnan(x)) return s + "NaN";
	if(std::isinf(x)) return s + "Inf";
	int exponent = 0;
	if(x >= 10.0f) {
		if(x >= 1E32f) { x *= 1E-32f; exponent += 32; }
		if(x >= 1E16f) { x *= 1E-16f; exponent += 16; }
		if(x >= 1E8f) { x *= 1E-8f; exponent += 8; }
		if(x >= 1E4f) { x *= 1E-4f; exponent += 4; }
		if(x >= 1E2f) { x *= 1E-2f; exponent += 2; }
		if(x >= 1E1f) { x *= 1E-1f; exponent += 1; }
	}
	if(x > 0.0f && x <= 1.0f) {
		if(x < 1E-31f) { x *= 1E32f; exponent -= 32; }
		if(x < 1E-15f) { x *= 1E16f; exponent -= 16; }
		if(x < 1E-7f) { x *= 1E8f; exponent -= 8; }
		if(x < 1E-3f) { x *= 1E4f; exponent -= 4; }
		if(x < 1E-1f) { x *= 1E2f; exponent -= 2; }
		if(x < 1E0f) { x *= 1E1f; exponent -= 1; }
	}
	uint32_t integral = (uint32_t)x;
	const float remainder = (x - (float)integral) * 1E8f;
	uint32_t decimal = (uint32_t)remainder;
	if(remainder - (float)decimal >= 0.5f) {
		decimal++;
		if(decimal >= 100000000u) {
			decimal = 0u;
			integral++;
			if(integral >= 10u) { integral = 1u; exponent++; }
		}
	}
	char digits[9];
	for(int k = 7; k >= 0; k--) { digits[k] = (char)('0' + decimal % 10u); decimal /= 10u; }
	digits[8] = 0;
	s += std::to_string(integral) + "." + digits;
	if(exponent != 0) s += "E" + std::to_string(exponent);
	return s;
}
float relaxation_rate_like_reference(float nu) {
	const float tau = 3.0f * nu + 0.5f; // src/lbm.hpp:132
	const std::string text = format_float_like_reference(1.0f / tau);
	return std::strtof(text.c_str(), nullptr);
}

// ---------------------------------------------------------------------------------------
// Velocity sets (src/kernel.cpp:864-903, src/lbm.cpp:376-385): the direction order and
// weights must be identical to the reference because the DDF slot index i is part of the
// memory layout.  Directions come in opposite pairs (i odd, i+1 = -e_i).
// ---------------------------------------------------------------------------------------
template<int Q> struct VS;
template<> struct VS<19> {
	static constexpr int cx[19] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0};
	static constexpr int cy[19] = {0, 0, 0, 1,-1, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1};
	static constexpr int cz[19] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0,-1, 1,-1, 1};
	template<typename T> static T w0() { return T(1) / T(3); }
	template<typename T> static T ws() { return T(1) / T(18); }
	template<typename T> static T we() { return T(1) / T(36); }
	template<typename T> static T wc() { return T(0); }
};
template<> struct VS<27> {
	static constexpr int cx[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1};
	static constexpr int cy[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1, 1,-1};
	static constexpr int cz[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0,-1, 1,-1, 1, 1,-1,-1, 1, 1,-1, 1,-1};
	template<typename T> static T w0() { return T(1) / T(3.375); }
	template<typename T> static T ws() { return T(1) / T(13.5); }
	template<typename T> static T we() { return T(1) / T(54); }
	template<typename T> static T wc() { return T(1) / T(216); }
};
template<int Q, typename T> T weight_of(int i) {
	if(i == 0) return VS<Q>::template w0<T>();
	const int c2 = VS<Q>::cx[i] * VS<Q>::cx[i] + VS<Q>::cy[i] * VS<Q>::cy[i] + VS<Q>::cz[i] * VS<Q>::cz[i];
	return c2 == 1 ? VS<Q>::template ws<T>() : c2 == 2 ? VS<Q>::template we<T>() : VS<Q>::template wc<T>();
}
template<int Q, typename T> struct WTable {
	T w[Q];
	WTable() { for(int i = 0; i < Q; i++) w[i] = weight_of<Q, T>(i); }
};
template<int Q, typename T> inline T weight(int i) { static const WTable<Q, T> table; return table.w[i]; }

// constants that appear as literals in the reference kernels
template<typename T> struct Lit;
template<> struct Lit<float> {
	static float c() { return 0.57735027f; } // def_c, src/lbm.cpp:366
	static float third() { return 0.33333334f; } // src/kernel.cpp:1093-1099
	static float smagorinsky() { return 0.76421222f; } // 18*sqrt(2)*(C*Delta)^2, src/kernel.cpp:1591
};
template<> struct Lit<double> {
	static double c() { return 0.57735027; }
	static double third() { return 1.0 / 3.0; }
	static double smagorinsky() { return 0.76421222; }
};

// Dual number (value, derivative along one direction) for forward-mode derivatives of the
// FP64 collision; used only by the adjoint oracle.
struct Dual {
	double v = 0.0, d = 0.0;
	Dual() {}
	Dual(double value, double deriv = 0.0) : v(value), d(deriv) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
template<> struct Lit<Dual> {
	static Dual c() { return Dual(0.57735027); }
	static Dual third() { return Dual(1.0 / 3.0); }
	static Dual smagorinsky() { return Dual(0.76421222); }
};
inline Dual sqrtT(Dual a) { const double r = std::sqrt(a.v); return Dual(r, a.d / (2.0 * r)); }
inline Dual fmaT(Dual a, Dual b, Dual c) { return Dual(std::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d); }
inline Dual clampT(Dual x, Dual lo, Dual hi) { return x.v < lo.v ? lo : x.v > hi.v ? hi : x; }

inline float fmaT(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fmaT(double a, double b, double c) { return std::fma(a, b, c); }
inline float sqrtT(float a) { return std::sqrt(a); }   // correctly rounded
inline double sqrtT(double a) { return std::sqrt(a); }
template<typename T> inline T clampT(T x, T lo, T hi) { return std::fmin(std::fmax(x, lo), hi); } // OpenCL clamp

// c_i.v evaluated like the reference's hand-written projections (u0=ux+uy, u3=ux-uy,
// u9=-ux+uy+uz, ...): left to right over the non-zero components, src/kernel.cpp:1029-1058.
template<int Q, typename T> inline T project(int i, T x, T y, T z) {
	const int c[3] = {VS<Q>::cx[i], VS<Q>::cy[i], VS<Q>::cz[i]};
	const T v[3] = {x, y, z};
	bool first = true; T s = T(0);
	#pragma GCC unroll 3
	for(int a = 0; a < 3; a++) {
		if(c[a] == 0) continue;
		if(first) { s = c[a] > 0 ? v[a] : -v[a]; first = false; }
		else s = c[a] > 0 ? s + v[a] : s - v[a];
	}
	return s;
}

// calculate_f_eq, src/kernel.cpp:1004-1061 (3D branch; DDF-shifted equilibrium).  rhom1 is
// rho-1: the reference forms it as rho - 1 (src/kernel.cpp:1005); with PRECISE_DENSITY the
// caller passes the DDF sum itself, which keeps its relative precision.
template<int Q, typename T> void calc_feq(const T rho, const T rhom1, T ux, T uy, T uz, T* feq) {
	const T c3 = T(-3) * ((ux * ux + uy * uy) + uz * uz);
	uz *= T(3); ux *= T(3); uy *= T(3);
	feq[0] = VS<Q>::template w0<T>() * fmaT(rho, T(0.5) * c3, rhom1);
	#pragma GCC unroll 32
	for(int i = 1; i < Q; i += 2) {
		const T wi = weight<Q, T>(i);
		const T rhoi = wi * rho, rhom1i = wi * rhom1;
		const T uk = project<Q, T>(i, ux, uy, uz);
		feq[i] = fmaT(rhoi, fmaT(T(0.5), fmaT(uk, uk, c3), uk), rhom1i);
		feq[i + 1] = fmaT(rhoi, fmaT(T(0.5), fmaT(uk, uk, c3), -uk), rhom1i);
	}
}

// calculate_rho_u, src/kernel.cpp:1063-1088: rho summed in slot order then +1 last;
// each velocity component summed pair by pair, positive member first.
template<int Q, typename T> void calc_rho_u(const T* f, T* rhon, T* uxn, T* uyn, T* uzn) {
	T rho = f[0];
	#pragma GCC unroll 32
	for(int i = 1; i < Q; i++) rho += f[i];
	rho += T(1);
	const int* cc[3] = {VS<Q>::cx, VS<Q>::cy, VS<Q>::cz};
	T u[3];
	#pragma GCC unroll 3
	for(int a = 0; a < 3; a++) {
		bool first = true; T s = T(0);
		#pragma GCC unroll 32
		for(int i = 1; i < Q; i += 2) {
			const int ca = cc[a][i];
			if(ca == 0) continue;
			const T fp = ca > 0 ? f[i] : f[i + 1], fm = ca > 0 ? f[i + 1] : f[i];
			if(first) { s = fp; first = false; } else s = s + fp;
			s = s - fm;
		}
		u[a] = s;
	}
	*rhon = rho;
	*uxn = u[0] / rho;
	*uyn = u[1] / rho;
	*uzn = u[2] / rho;
}

// calculate_forcing_terms, src/kernel.cpp:1090-1101 (Guo forcing)
template<int Q, typename T> void calc_forcing(T ux, T uy, T uz, T fx, T fy, T fz, T* Fin) {
	const T uF = -Lit<T>::third() * fmaT(ux, fx, fmaT(uy, fy, uz * fz));
	Fin[0] = (T(9) * VS<Q>::template w0<T>()) * uF;
	#pragma GCC unroll 32
	for(int i = 1; i < Q; i++) {
		const T cx = (T)VS<Q>::cx[i], cy = (T)VS<Q>::cy[i], cz = (T)VS<Q>::cz[i];
		Fin[i] = (T(9) * weight<Q, T>(i)) * fmaT((cx * fx + cy * fy) + cz * fz, ((cx * ux + cy * uy) + cz * uz) + Lit<T>::third(), uF);
	}
}

struct Cfg {
	uint32_t Nx, Ny, Nz, q, trt, storage, ext, threads;
	double w;
};

// The force block of stream_collide / update_fields, src/kernel.cpp:1549-1563 and
// 1847-1859, plus the Brinkman porosity term (not in the reference):
//   u = (j/rho + f*0.5/rho)/(1+alpha/2), F = f - alpha*rho*u  (implicit Guo-Brinkman).
// With alpha = 0 the result is bit-identical to the reference's VOLUME_FORCE path.
template<typename T> void force_block(uint32_t ext, T rhon, T& uxn, T& uyn, T& uzn, T& fxn, T& fyn, T& fzn, T a) {
	const T cmax = Lit<T>::c();
	if(ext & (EXT_VOLUME_FORCE | EXT_POROSITY)) {
		const T rho2 = T(0.5) / rhon;
		uxn = fmaT(fxn, rho2, uxn);
		uyn = fmaT(fyn, rho2, uyn);
		uzn = fmaT(fzn, rho2, uzn);
		T ar = T(0);
		if(ext & EXT_POROSITY) {
			const T k = T(1) / (T(1) + T(0.5) * a);
			uxn = uxn * k; uyn = uyn * k; uzn = uzn * k;
			ar = a * rhon;
		}
		uxn = clampT(uxn, -cmax, cmax);
		uyn = clampT(uyn, -cmax, cmax);
		uzn = clampT(uzn, -cmax, cmax);
		if(ext & EXT_POROSITY) {
			fxn = fmaT(-ar, uxn, fxn);
			fyn = fmaT(-ar, uyn, fyn);
			fzn = fmaT(-ar, uzn, fzn);
		}
	} else {
		uxn = clampT(uxn, -cmax, cmax);
		uyn = clampT(uyn, -cmax, cmax);
		uzn = clampT(uzn, -cmax, cmax);
	}
}

// SUBGRID, src/kernel.cpp:1579-1593: Smagorinsky-Lilly eddy viscosity.  The local
// relaxation rate follows from the non-equilibrium stress H = sum_i c_i c_i (f_i - feq_i)
// (i >= 1): w = 2/(tau0 + sqrt(tau0^2 + 0.76421222*sqrt(H:H)/rho)), tau0 = 1/w.  Terms
// whose lattice coefficient is 0 are left out (they add exact zeros).
template<int Q, typename T> T subgrid_rate(T w, const T* fhn, const T* feq, T rhon) {
	const T tau0 = T(1) / w;
	T Hxx = T(0), Hyy = T(0), Hzz = T(0), Hxy = T(0), Hxz = T(0), Hyz = T(0);
	for(int i = 1; i < Q; i++) {
		const T fneq = fhn[i] - feq[i];
		const int cx = VS<Q>::cx[i], cy = VS<Q>::cy[i], cz = VS<Q>::cz[i];
		if(cx != 0) Hxx += fneq;
		if(cx * cy != 0) Hxy = cx * cy > 0 ? Hxy + fneq : Hxy - fneq;
		if(cy != 0) Hyy += fneq;
		if(cx * cz != 0) Hxz = cx * cz > 0 ? Hxz + fneq : Hxz - fneq;
		if(cy * cz != 0) Hyz = cy * cz > 0 ? Hyz + fneq : Hyz - fneq;
		if(cz != 0) Hzz += fneq;
	}
	const T Qs = ((Hxx * Hxx + Hyy * Hyy) + Hzz * Hzz) + T(2) * ((Hxy * Hxy + Hxz * Hxz) + Hyz * Hyz);
	return T(2) / (tau0 + sqrtT(tau0 * tau0 + Lit<T>::smagorinsky() * sqrtT(Qs) / rhon));
}

// Collision of one cell, src/kernel.cpp:1474-1633: moments (or the TYPE_E presets), force
// block, Guo forcing terms, equilibrium, SRT or TRT, in place on the loaded populations
// fhn.  fxn.. is the external force of the cell (volume force plus force field).  If
// fields is not null it receives rho, ux, uy, uz (the UPDATE_FIELDS values).
template<int Q, typename T> void collide_cell(const Cfg& c, T* fhn, bool isE, T rhoE, T uxE, T uyE, T uzE,
                                              T fxn, T fyn, T fzn, T alpha, T w, T* fields) {
	const bool vf = c.ext & (EXT_VOLUME_FORCE | EXT_POROSITY);
	T feq[Q], Fin[Q];
	T rhon, uxn, uyn, uzn;
	T rhom1;
	if(isE) { rhon = rhoE; uxn = uxE; uyn = uyE; uzn = uzE; rhom1 = rhon - T(1); }
	else {
		calc_rho_u<Q, T>(fhn, &rhon, &uxn, &uyn, &uzn);
		if(c.ext & EXT_PRECISE_DENSITY) { // rho-1 = sum of the shifted DDFs, same order as calc_rho_u
			rhom1 = fhn[0];
			for(int i = 1; i < Q; i++) rhom1 += fhn[i];
		} else rhom1 = rhon - T(1);
	}
	force_block<T>(c.ext, rhon, uxn, uyn, uzn, fxn, fyn, fzn, alpha);
	if(vf) calc_forcing<Q, T>(uxn, uyn, uzn, fxn, fyn, fzn, Fin);
	else for(int i = 0; i < Q; i++) Fin[i] = T(0);
	if(fields) { fields[0] = rhon; fields[1] = uxn; fields[2] = uyn; fields[3] = uzn; }
	calc_feq<Q, T>(rhon, rhom1, uxn, uyn, uzn, feq);
	if(c.ext & EXT_SUBGRID) w = subgrid_rate<Q, T>(w, fhn, feq, rhon);
	if(!c.trt) { // SRT, src/kernel.cpp:1595-1604
		if(vf) { const T c_tau = fmaT(w, T(-0.5), T(1)); for(int i = 0; i < Q; i++) Fin[i] *= c_tau; }
		#pragma GCC unroll 32
		for(int i = 0; i < Q; i++) fhn[i] = isE ? feq[i] : fmaT(T(1) - w, fhn[i], fmaT(w, feq[i], Fin[i]));
	} else { // TRT, src/kernel.cpp:1605-1633
		const T wp = w;
		const T wm = T(1) / (T(0.1875) / (T(1) / w - T(0.5)) + T(0.5));
		if(vf) {
			const T c_taup = fmaT(wp, T(-0.25), T(0.5)), c_taum = fmaT(wm, T(-0.25), T(0.5));
			T Fib[Q]; Fib[0] = Fin[0];
			#pragma GCC unroll 32
			for(int i = 1; i < Q; i += 2) { Fib[i] = Fin[i + 1]; Fib[i + 1] = Fin[i]; }
			#pragma GCC unroll 32
			for(int i = 0; i < Q; i++) Fin[i] = fmaT(c_taup, Fin[i] + Fib[i], c_taum * (Fin[i] - Fib[i]));
		}
		T fhb[Q], feb[Q];
		fhb[0] = fhn[0]; feb[0] = feq[0];
		#pragma GCC unroll 32
		for(int i = 1; i < Q; i += 2) { fhb[i] = fhn[i + 1]; fhb[i + 1] = fhn[i]; feb[i] = feq[i + 1]; feb[i + 1] = feq[i]; }
		#pragma GCC unroll 32
		for(int i = 0; i < Q; i++) {
			fhn[i] = isE ? feq[i] : fmaT(T(0.5) * wp, ((feq[i] - fhn[i]) + feb[i]) - fhb[i],
			                             fmaT(T(0.5) * wm, ((feq[i] - feb[i]) - fhn[i]) + fhb[i], fhn[i] + Fin[i]));
		}
	}
}

template<typename T, int ST> struct Store;
template<> struct Store<float, ST_FP32> {
	typedef float E;
	static float load(const E* p, uint64_t o) { return p[o]; }
	static void store(E* p, uint64_t o, float x) { p[o] = x; }
};
template<> struct Store<float, ST_FP16S> {
	typedef uint16_t E;
	static float load(const E* p, uint64_t o) { return fp16s_load(p[o]); }
	static void store(E* p, uint64_t o, float x) { p[o] = fp16s_store(x); }
};
template<> struct Store<float, ST_FP16C> {
	typedef uint16_t E;
	static float load(const E* p, uint64_t o) { return fp16c_load(p[o]); }
	static void store(E* p, uint64_t o, float x) { p[o] = fp16c_store(x); }
};
template<> struct Store<double, ST_FP64> {
	typedef double E;
	static double load(const E* p, uint64_t o) { return p[o]; }
	static void store(E* p, uint64_t o, double x) { p[o] = x; }
};

template<typename F> void parallel_cells(uint64_t N, uint32_t threads, F&& body) {
	if(threads <= 1u || N < 4096u) { body(0ull, N); return; }
	std::vector<std::thread> pool;
	const uint64_t chunk = (N + threads - 1u) / threads;
	for(uint32_t k = 0; k < threads; k++) {
		const uint64_t a = (uint64_t)k * chunk, b = std::min<uint64_t>(N, a + chunk);
		if(a >= b) break;
		pool.emplace_back([&body, a, b]() { body(a, b); });
	}
	for(auto& th : pool) th.join();
}

template<typename T, int Q, int ST> struct Lattice {
	typedef typename Store<T, ST>::E E;
	const Cfg& c;
	uint64_t N;
	explicit Lattice(const Cfg& cfg) : c(cfg), N((uint64_t)cfg.Nx * cfg.Ny * cfg.Nz) {}

	// coordinates / calculate_indices / neighbors, src/kernel.cpp:817-820, 904-958
	void neighbors(uint64_t n, uint64_t* j) const {
		const uint64_t NxNy = (uint64_t)c.Nx * c.Ny;
		const uint64_t tt = n % NxNy;
		const uint32_t x = (uint32_t)(tt % c.Nx), y = (uint32_t)(tt / c.Nx), z = (uint32_t)(n / NxNy);
		const uint64_t X[3] = {(uint64_t)((x + c.Nx - 1u) % c.Nx), (uint64_t)x, (uint64_t)((x + 1u) % c.Nx)};
		const uint64_t Y[3] = {(uint64_t)((y + c.Ny - 1u) % c.Ny) * c.Nx, (uint64_t)y * c.Nx, (uint64_t)((y + 1u) % c.Ny) * c.Nx};
		const uint64_t Z[3] = {(uint64_t)((z + c.Nz - 1u) % c.Nz) * NxNy, (uint64_t)z * NxNy, (uint64_t)((z + 1u) % c.Nz) * NxNy};
		#pragma GCC unroll 32
		for(int i = 0; i < Q; i++) j[i] = X[VS<Q>::cx[i] + 1] + Y[VS<Q>::cy[i] + 1] + Z[VS<Q>::cz[i] + 1];
	}
	uint64_t idx(uint64_t n, int i) const { return (uint64_t)i * N + n; } // index_f, src/kernel.cpp:861
	// Esoteric-Pull load_f / store_f, src/kernel.cpp:1326-1339
	void load_f(uint64_t n, T* f, const E* fi, const uint64_t* j, uint64_t t) const {
		f[0] = Store<T, ST>::load(fi, idx(n, 0));
		#pragma GCC unroll 32
		for(int i = 1; i < Q; i += 2) {
			f[i] = Store<T, ST>::load(fi, idx(n, (t & 1u) ? i : i + 1));
			f[i + 1] = Store<T, ST>::load(fi, idx(j[i], (t & 1u) ? i + 1 : i));
		}
	}
	void store_f(uint64_t n, const T* f, E* fi, const uint64_t* j, uint64_t t) const {
		Store<T, ST>::store(fi, idx(n, 0), f[0]);
		#pragma GCC unroll 32
		for(int i = 1; i < Q; i += 2) {
			Store<T, ST>::store(fi, idx(j[i], (t & 1u) ? i + 1 : i), f[i]);
			Store<T, ST>::store(fi, idx(n, (t & 1u) ? i : i + 1), f[i + 1]);
		}
	}

	// A non-solid cell next to a solid cell with non-zero velocity, src/kernel.cpp:1395-1398
	// and 1441-1444 (TYPE_MS marking)
	bool next_to_moving_solid(const uint64_t* j, const T* u, const uint8_t* flags) const {
		bool m = false;
		for(int i = 1; i < Q; i++) {
			const uint64_t k = j[i];
			m = m || ((flags[k] & TYPE_BO) == TYPE_S && (u[k] != T(0) || u[N + k] != T(0) || u[2 * N + k] != T(0)));
		}
		return m;
	}
	// initialize kernel, src/kernel.cpp:1358-1430 (no SURFACE / TEMPERATURE).  Without
	// MOVING_BOUNDARIES every solid cell's velocity is zeroed; with it only solid cells
	// whose neighbours are all solid, and the other non-TYPE_E cells are marked or unmarked
	// TYPE_MS.  New flags are formed from the old ones first and written afterwards (the
	// reference updates in place; the marks never change the solid test it reads).
	void initialize(E* fi, const T* rho, T* u, uint8_t* flags) const {
		const bool mb = c.ext & EXT_MOVING_BOUNDARIES;
		std::vector<uint8_t> nf(flags, flags + N);
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q];
			for(uint64_t n = a; n < b; n++) {
				const uint8_t bo = flags[n] & TYPE_BO;
				if(bo == TYPE_S) {
					if(!mb) { u[n] = T(0); u[N + n] = T(0); u[2 * N + n] = T(0); continue; }
					neighbors(n, j);
					bool only_s = true;
					for(int i = 1; i < Q; i++) only_s = only_s && (flags[j[i]] & TYPE_BO) == TYPE_S;
					if(only_s) { u[n] = T(0); u[N + n] = T(0); u[2 * N + n] = T(0); }
				} else if(mb && bo != TYPE_E) {
					neighbors(n, j);
					nf[n] = next_to_moving_solid(j, u, flags) ? (uint8_t)(flags[n] | TYPE_MS) : (uint8_t)(flags[n] & ~TYPE_MS);
				}
			}
		});
		std::memcpy(flags, nf.data(), N);
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q]; T feq[Q];
			for(uint64_t n = a; n < b; n++) {
				neighbors(n, j);
				calc_feq<Q, T>(rho[n], rho[n] - T(1), u[n], u[N + n], u[2 * N + n], feq);
				store_f(n, feq, fi, j, 1u);
			}
		});
	}
	// update_moving_boundaries kernel, src/kernel.cpp:1433-1449 (not on solid, TYPE_E or
	// TYPE_T cells)
	void update_moving_boundaries(const T* u, uint8_t* flags) const {
		std::vector<uint8_t> nf(flags, flags + N);
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q];
			for(uint64_t n = a; n < b; n++) {
				const uint8_t bo = flags[n] & TYPE_BO;
				if(bo == TYPE_S || bo == TYPE_E || (flags[n] & TYPE_T)) continue;
				neighbors(n, j);
				nf[n] = next_to_moving_solid(j, u, flags) ? (uint8_t)(flags[n] | TYPE_MS) : (uint8_t)(flags[n] & ~TYPE_MS);
			}
		});
		std::memcpy(flags, nf.data(), N);
	}
	// apply_moving_boundaries, src/kernel.cpp:1104-1113: a population that arrives from a
	// solid neighbour (n - e_i for f_i) gains -6 w_i (c_opposite . u_wall), wall density 1
	// (Krueger et al. p. 180)
	void apply_moving_boundaries(T* fhn, const uint64_t* j, const T* u, const uint8_t* flags) const {
		for(int i = 1; i < Q; i += 2) {
			const T w6 = T(-6) * weight<Q, T>(i);
			uint64_t k = j[i + 1];
			if((flags[k] & TYPE_BO) == TYPE_S) fhn[i] = fmaT(w6, project<Q, T>(i + 1, u[k], u[N + k], u[2 * N + k]), fhn[i]);
			k = j[i];
			if((flags[k] & TYPE_BO) == TYPE_S) fhn[i + 1] = fmaT(w6, project<Q, T>(i, u[k], u[N + k], u[2 * N + k]), fhn[i + 1]);
		}
	}

	T relaxation_rate() const { return (T)(std::is_same<T, float>::value ? (double)(float)c.w : c.w); }

	// stream_collide kernel, src/kernel.cpp:1454-1636
	void stream_collide(E* fi, T* rho, T* u, const uint8_t* flags, uint64_t t, T fx, T fy, T fz, const T* F, const T* alpha) const {
		const bool eq = c.ext & EXT_EQUILIBRIUM_BOUNDARIES;
		const T w = relaxation_rate();
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q]; T fhn[Q], fields[4];
			for(uint64_t n = a; n < b; n++) {
				const uint8_t flagsn = flags[n];
				const uint8_t bo = flagsn & TYPE_BO, su = flagsn & TYPE_SU;
				if(bo == TYPE_S || su == TYPE_G) continue;
				neighbors(n, j);
				load_f(n, fhn, fi, j, t);
				if((c.ext & EXT_MOVING_BOUNDARIES) && bo == TYPE_MS) apply_moving_boundaries(fhn, j, u, flags);
				const bool isE = eq && bo == TYPE_E;
				T fxn = fx, fyn = fy, fzn = fz;
				if(c.ext & EXT_FORCE_FIELD) { fxn += F[n]; fyn += F[N + n]; fzn += F[2 * N + n]; }
				const T an = (c.ext & EXT_POROSITY) ? alpha[n] : T(0);
				collide_cell<Q, T>(c, fhn, isE, isE ? rho[n] : T(0), isE ? u[n] : T(0), isE ? u[N + n] : T(0), isE ? u[2 * N + n] : T(0),
				                   fxn, fyn, fzn, an, w, fields);
				if((c.ext & EXT_UPDATE_FIELDS) && !isE) { rho[n] = fields[0]; u[n] = fields[1]; u[N + n] = fields[2]; u[2 * N + n] = fields[3]; }
				store_f(n, fhn, fi, j, t);
			}
		});
	}

	// update_fields kernel, src/kernel.cpp:1794-1870
	void update_fields(const E* fi, T* rho, T* u, const uint8_t* flags, uint64_t t, T fx, T fy, T fz, const T* F, const T* alpha) const {
		const bool eq = c.ext & EXT_EQUILIBRIUM_BOUNDARIES;
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q]; T fhn[Q];
			for(uint64_t n = a; n < b; n++) {
				const uint8_t flagsn = flags[n];
				const uint8_t bo = flagsn & TYPE_BO, su = flagsn & TYPE_SU;
				if(bo == TYPE_S || su == TYPE_G) continue;
				neighbors(n, j);
				load_f(n, fhn, fi, j, t);
				if((c.ext & EXT_MOVING_BOUNDARIES) && bo == TYPE_MS) apply_moving_boundaries(fhn, j, u, flags);
				T rhon, uxn, uyn, uzn;
				calc_rho_u<Q, T>(fhn, &rhon, &uxn, &uyn, &uzn);
				T fxn = fx, fyn = fy, fzn = fz;
				if(c.ext & EXT_FORCE_FIELD) { fxn += F[n]; fyn += F[N + n]; fzn += F[2 * N + n]; }
				force_block<T>(c.ext, rhon, uxn, uyn, uzn, fxn, fyn, fzn, (c.ext & EXT_POROSITY) ? alpha[n] : T(0));
				if(!(eq && bo == TYPE_E)) { rho[n] = rhon; u[n] = uxn; u[N + n] = uyn; u[2 * N + n] = uzn; }
			}
		});
	}

	// update_force_field kernel, src/kernel.cpp:1873-1884 (momentum exchange on TYPE_S cells)
	void update_force_field(const E* fi, const uint8_t* flags, uint64_t t, T* F) const {
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q]; T fhn[Q];
			for(uint64_t n = a; n < b; n++) {
				if((flags[n] & TYPE_BO) != TYPE_S) continue;
				neighbors(n, j);
				load_f(n, fhn, fi, j, t);
				T Fb, fxv, fyv, fzv;
				calc_rho_u<Q, T>(fhn, &Fb, &fxv, &fyv, &fzv);
				const T s = T(2) * Fb;
				F[n] = s * fxv; F[N + n] = s * fyv; F[2 * N + n] = s * fzv;
			}
		});
	}

	// Objective of the adjoint: x-momentum of the populations loaded at step t, summed over
	// the non-solid cells of the plane x = plane_x (FP64).
	double objective_plane(const E* fi, const uint8_t* flags, uint64_t t, uint32_t plane_x) const {
		double s = 0.0;
		uint64_t j[Q]; T fhn[Q];
		for(uint32_t z = 0; z < c.Nz; z++) for(uint32_t y = 0; y < c.Ny; y++) {
			const uint64_t n = (uint64_t)plane_x + ((uint64_t)y + (uint64_t)z * c.Ny) * c.Nx;
			if((flags[n] & TYPE_BO) == TYPE_S) continue;
			neighbors(n, j);
			load_f(n, fhn, fi, j, t);
			for(int i = 1; i < Q; i++) s += (double)VS<Q>::cx[i] * (double)fhn[i];
		}
		return s;
	}

	// Adjoint of stream_collide at step t (FP64 oracle).  The forward step maps the
	// populations v loaded at the cell's footprint locations L_k to v' = C(v), stored at
	// S_m = L_{opposite(m)}.  Its transpose reads mu_m = lambda[S_m], forms
	// nu_k = sum_m mu_m dv'_m/dv_k (+ the objective's c_kx weight on the plane) and writes
	// lambda[L_k] = nu_k; solid cells are the identity.  The derivatives are taken by
	// forward-mode (dual numbers) through collide_cell, the same code as the forward step,
	// one direction per population and one for the porosity; grad[n] += sum_m mu_m dv'_m/dalpha.
	void adjoint_step(const E* f_state, double* lambda, const uint8_t* flags, uint64_t t, double fx, double fy, double fz,
	                  const double* F, const double* alpha, const double* rho, const double* u, uint32_t plane_x, double obj_weight, double* grad) const {
		const bool eq = c.ext & EXT_EQUILIBRIUM_BOUNDARIES;
		const Dual w = Dual(c.w);
		parallel_cells(N, c.threads, [&](uint64_t a, uint64_t b) {
			uint64_t j[Q]; double v[Q], lam[Q], mu[Q], nu[Q], out[Q];
			Dual vd[Q];
			for(uint64_t n = a; n < b; n++) {
				const uint8_t flagsn = flags[n];
				const uint8_t bo = flagsn & TYPE_BO, su = flagsn & TYPE_SU;
				if(bo == TYPE_S || su == TYPE_G) continue;
				neighbors(n, j);
				load_f(n, v, f_state, j, t);
				load_f(n, lam, lambda, j, t); // lam[k] = lambda[L_k]
				mu[0] = lam[0];
				for(int i = 1; i < Q; i += 2) { mu[i] = lam[i + 1]; mu[i + 1] = lam[i]; } // mu[m] = lambda[S_m]
				const bool isE = eq && bo == TYPE_E;
				double fxn = fx, fyn = fy, fzn = fz;
				if(c.ext & EXT_FORCE_FIELD) { fxn += F[n]; fyn += F[N + n]; fzn += F[2 * N + n]; }
				const double an = (c.ext & EXT_POROSITY) ? alpha[n] : 0.0;
				const Dual rE = isE ? Dual(rho[n]) : Dual(0.0), uxE = isE ? Dual(u[n]) : Dual(0.0), uyE = isE ? Dual(u[N + n]) : Dual(0.0), uzE = isE ? Dual(u[2 * N + n]) : Dual(0.0);
				for(int k = 0; k < Q; k++) {
					for(int i = 0; i < Q; i++) vd[i] = Dual(v[i], i == k ? 1.0 : 0.0);
					collide_cell<Q, Dual>(c, vd, isE, rE, uxE, uyE, uzE, Dual(fxn), Dual(fyn), Dual(fzn), Dual(an), w, nullptr);
					double s = 0.0;
					for(int m = 0; m < Q; m++) s += mu[m] * vd[m].d;
					nu[k] = s;
				}
				if(c.ext & EXT_POROSITY) {
					for(int i = 0; i < Q; i++) vd[i] = Dual(v[i], 0.0);
					collide_cell<Q, Dual>(c, vd, isE, rE, uxE, uyE, uzE, Dual(fxn), Dual(fyn), Dual(fzn), Dual(an, 1.0), w, nullptr);
					double s = 0.0;
					for(int m = 0; m < Q; m++) s += mu[m] * vd[m].d;
					grad[n] += s;
				}
				if(obj_weight != 0.0 && (uint32_t)(n % c.Nx) == plane_x) {
					for(int k = 1; k < Q; k++) nu[k] += obj_weight * (double)VS<Q>::cx[k];
				}
				// lambda[L_k] = nu[k]: store_f writes out[i] to S_i = L_opposite(i)
				out[0] = nu[0];
				for(int i = 1; i < Q; i += 2) { out[i] = nu[i + 1]; out[i + 1] = nu[i]; }
				store_f(n, out, lambda, j, t);
			}
		});
	}

	// populations every cell would load at step t (load_f for all cells), out[i*N+n]
	void load_populations(const E* fi, uint64_t t, T* out) const {
		uint64_t j[Q]; T fhn[Q];
		for(uint64_t n = 0; n < N; n++) {
			neighbors(n, j);
			load_f(n, fhn, fi, j, t);
			for(int i = 0; i < Q; i++) out[(uint64_t)i * N + n] = fhn[i];
		}
	}

	// total mass: sum over every cell (solid ones included) of 1 + the shifted DDFs it
	// stores.  Each step only permutes stored values inside disjoint footprints and the
	// collision conserves sum_i f_i, so this is invariant; populations in flight toward a
	// wall sit in the solid cell's slots and are counted there.  FP64, cell order.
	double mass_sum(const E* fi) const {
		double s = 0.0;
		for(uint64_t n = 0; n < N; n++) {
			double c = 1.0;
			for(int i = 0; i < Q; i++) c += (double)Store<T, ST>::load(fi, idx(n, i));
			s += c;
		}
		return s;
	}
};

// position(coordinates(n)), src/kernel.cpp:824-826 (FP32 cell-centred position)
inline void position_f32(const Cfg& c, uint64_t n, float* p) {
	const uint64_t NxNy = (uint64_t)c.Nx * c.Ny;
	const uint64_t tt = n % NxNy;
	const uint32_t x = (uint32_t)(tt % c.Nx), y = (uint32_t)(tt / c.Nx), z = (uint32_t)(n / NxNy);
	p[0] = (float)x + 0.5f - 0.5f * (float)c.Nx;
	p[1] = (float)y + 0.5f - 0.5f * (float)c.Ny;
	p[2] = (float)z + 0.5f - 0.5f * (float)c.Nz;
}

template<template<typename, int, int> class Body, typename... A> int dispatch(const Cfg& c, A&&... args) {
	switch(c.q * 10u + c.storage) {
		case 190u + ST_FP32: return Body<float, 19, ST_FP32>::run(c, args...);
		case 190u + ST_FP16S: return Body<float, 19, ST_FP16S>::run(c, args...);
		case 190u + ST_FP16C: return Body<float, 19, ST_FP16C>::run(c, args...);
		case 190u + ST_FP64: return Body<double, 19, ST_FP64>::run(c, args...);
		case 270u + ST_FP32: return Body<float, 27, ST_FP32>::run(c, args...);
		case 270u + ST_FP16S: return Body<float, 27, ST_FP16S>::run(c, args...);
		case 270u + ST_FP16C: return Body<float, 27, ST_FP16C>::run(c, args...);
		case 270u + ST_FP64: return Body<double, 27, ST_FP64>::run(c, args...);
		default: return -1;
	}
}

template<typename T, int Q, int ST> struct DoInit {
	static int run(const Cfg& c, void* fi, const void* rho, void* u, uint8_t* flags) {
		Lattice<T, Q, ST>(c).initialize((typename Store<T, ST>::E*)fi, (const T*)rho, (T*)u, flags);
		return 0;
	}
};
template<typename T, int Q, int ST> struct DoMovingBoundaries {
	static int run(const Cfg& c, const void* u, uint8_t* flags) {
		Lattice<T, Q, ST>(c).update_moving_boundaries((const T*)u, flags);
		return 0;
	}
};
template<typename T, int Q, int ST> struct DoStep {
	static int run(const Cfg& c, void* fi, void* rho, void* u, const uint8_t* flags, uint64_t t, double fx, double fy, double fz, const void* F, const void* alpha) {
		Lattice<T, Q, ST>(c).stream_collide((typename Store<T, ST>::E*)fi, (T*)rho, (T*)u, flags, t, (T)fx, (T)fy, (T)fz, (const T*)F, (const T*)alpha);
		return 0;
	}
};
template<typename T, int Q, int ST> struct DoFields {
	static int run(const Cfg& c, const void* fi, void* rho, void* u, const uint8_t* flags, uint64_t t, double fx, double fy, double fz, const void* F, const void* alpha) {
		Lattice<T, Q, ST>(c).update_fields((const typename Store<T, ST>::E*)fi, (T*)rho, (T*)u, flags, t, (T)fx, (T)fy, (T)fz, (const T*)F, (const T*)alpha);
		return 0;
	}
};
template<typename T, int Q, int ST> struct DoForceField {
	static int run(const Cfg& c, const void* fi, const uint8_t* flags, uint64_t t, void* F) {
		Lattice<T, Q, ST>(c).update_force_field((const typename Store<T, ST>::E*)fi, flags, t, (T*)F);
		return 0;
	}
};
template<typename T, int Q, int ST> struct DoPopulations {
	static int run(const Cfg& c, const void* fi, uint64_t t, void* out) {
		Lattice<T, Q, ST>(c).load_populations((const typename Store<T, ST>::E*)fi, t, (T*)out);
		return 0;
	}
};
template<typename T, int Q, int ST> struct DoMass {
	static int run(const Cfg& c, const void* fi, const uint8_t* flags, uint64_t t, double* out) {
		(void)flags; (void)t;
		*out = Lattice<T, Q, ST>(c).mass_sum((const typename Store<T, ST>::E*)fi);
		return 0;
	}
};

bool valid(const Cfg* c) {
	return c && c->Nx && c->Ny && c->Nz && (c->q == 19u || c->q == 27u) && c->storage <= ST_FP64 && c->trt <= 1u;
}

} // namespace

extern "C" {

typedef Cfg fxo_cfg;

// relaxation rate w exactly as the reference's kernels receive it (def_w round trip)
float fxo_relaxation_rate(float nu) { return relaxation_rate_like_reference(nu); }
// writes the reference-formatted decimal text of x into buf (size >= 32); returns length
int fxo_format_float(float x, char* buf, int size) {
	const std::string s = format_float_like_reference(x);
	if((int)s.size() + 1 > size) return -1;
	std::memcpy(buf, s.c_str(), s.size() + 1);
	return (int)s.size();
}

uint16_t fxo_fp16s_encode(float x) { return fp16s_store(x); }
// Compares the software binary16 conversions with the F16C hardware conversions over all
// 2^32 binary32 bit patterns (encode; NaN payloads excluded) and all 2^16 codes (decode).
// Returns the number of mismatches.
uint64_t fxo_f16_selfcheck(uint32_t threads) {
	std::atomic<uint64_t> bad{0};
	if(threads == 0u) threads = 1u;
	std::vector<std::thread> pool;
	for(uint32_t k = 0; k < threads; k++) {
		pool.emplace_back([k, threads, &bad]() {
			uint64_t local = 0;
			const uint64_t lo = (0x100000000ull * k) / threads, hi = (0x100000000ull * (k + 1u)) / threads;
			for(uint64_t b = lo; b < hi; b++) {
				const float x = bitsf((uint32_t)b);
				if(std::isnan(x)) continue;
				if(f32_to_f16_rne(x) != (uint16_t)_cvtss_sh(x, _MM_FROUND_TO_NEAREST_INT)) local++;
			}
			bad += local;
		});
	}
	for(auto& th : pool) th.join();
	uint64_t nbad = bad.load();
	for(uint32_t h = 0; h < 65536u; h++) {
		const float a = f16_to_f32((uint16_t)h), b = _cvtsh_ss((uint16_t)h);
		if(std::isnan(a) != std::isnan(b) || (!std::isnan(a) && fbits(a) != fbits(b))) nbad++;
	}
	return nbad;
}
float fxo_fp16s_decode(uint16_t h) { return fp16s_load(h); }
uint16_t fxo_f16_rne(float x) { return f32_to_f16_rne(x); }
float fxo_f16_to_f32(uint16_t h) { return f16_to_f32(h); }
uint16_t fxo_fp16c_encode(float x) { return fp16c_store(x); }
float fxo_fp16c_decode(uint16_t h) { return fp16c_load(h); }
void fxo_encode_array(int storage, const float* x, uint16_t* out, uint64_t n) {
	for(uint64_t k = 0; k < n; k++) out[k] = storage == (int)ST_FP16S ? fp16s_store(x[k]) : storage == (int)ST_FP16C ? fp16c_store(x[k]) : f32_to_f16_rne(x[k]);
}
void fxo_decode_array(int storage, const uint16_t* h, float* out, uint64_t n) {
	for(uint64_t k = 0; k < n; k++) out[k] = storage == (int)ST_FP16S ? fp16s_load(h[k]) : storage == (int)ST_FP16C ? fp16c_load(h[k]) : f16_to_f32(h[k]);
}

// flags is written with MOVING_BOUNDARIES (TYPE_MS marks)
int fxo_initialize(const fxo_cfg* c, void* fi, const void* rho, void* u, uint8_t* flags) {
	if(!valid(c)) return -1;
	return dispatch<DoInit>(*c, fi, rho, u, flags);
}
int fxo_update_moving_boundaries(const fxo_cfg* c, const void* u, uint8_t* flags) {
	if(!valid(c)) return -1;
	return dispatch<DoMovingBoundaries>(*c, u, flags);
}
int fxo_stream_collide(const fxo_cfg* c, void* fi, void* rho, void* u, const uint8_t* flags, uint64_t t, double fx, double fy, double fz, const void* F, const void* alpha) {
	if(!valid(c)) return -1;
	if((c->ext & EXT_FORCE_FIELD) && !F) return -2;
	if((c->ext & EXT_POROSITY) && !alpha) return -2;
	return dispatch<DoStep>(*c, fi, rho, u, flags, t, fx, fy, fz, F, alpha);
}
int fxo_update_fields(const fxo_cfg* c, const void* fi, void* rho, void* u, const uint8_t* flags, uint64_t t, double fx, double fy, double fz, const void* F, const void* alpha) {
	if(!valid(c)) return -1;
	if((c->ext & EXT_FORCE_FIELD) && !F) return -2;
	if((c->ext & EXT_POROSITY) && !alpha) return -2;
	return dispatch<DoFields>(*c, fi, rho, u, flags, t, fx, fy, fz, F, alpha);
}
int fxo_update_force_field(const fxo_cfg* c, const void* fi, const uint8_t* flags, uint64_t t, void* F) {
	if(!valid(c)) return -1;
	return dispatch<DoForceField>(*c, fi, flags, t, F);
}
// FP64 only: adjoint of stream_collide at step t (see Lattice::adjoint_step).
int fxo_adjoint_step(const fxo_cfg* c, const void* f_state, void* lambda, const uint8_t* flags, uint64_t t, double fx, double fy, double fz,
                     const void* F, const void* alpha, uint32_t plane_x, double obj_weight, void* grad, const void* rho, const void* u) {
	if(!valid(c) || c->storage != ST_FP64) return -1;
	if(c->ext & (EXT_MOVING_BOUNDARIES | EXT_SUBGRID)) return -3; // not differentiated
	if((c->ext & EXT_POROSITY) && (!alpha || !grad)) return -2;
	if((c->ext & EXT_FORCE_FIELD) && !F) return -2;
	if((c->ext & EXT_EQUILIBRIUM_BOUNDARIES) && (!rho || !u)) return -2;
	if(c->q == 19u) Lattice<double, 19, ST_FP64>(*c).adjoint_step((const double*)f_state, (double*)lambda, flags, t, fx, fy, fz, (const double*)F, (const double*)alpha, (const double*)rho, (const double*)u, plane_x, obj_weight, (double*)grad);
	else Lattice<double, 27, ST_FP64>(*c).adjoint_step((const double*)f_state, (double*)lambda, flags, t, fx, fy, fz, (const double*)F, (const double*)alpha, (const double*)rho, (const double*)u, plane_x, obj_weight, (double*)grad);
	return 0;
}
// FP64 only: objective plane sum at step t
int fxo_objective_plane(const fxo_cfg* c, const void* fi, const uint8_t* flags, uint64_t t, uint32_t plane_x, double* out) {
	if(!valid(c) || c->storage != ST_FP64 || plane_x >= c->Nx) return -1;
	*out = c->q == 19u ? Lattice<double, 19, ST_FP64>(*c).objective_plane((const double*)fi, flags, t, plane_x)
	                   : Lattice<double, 27, ST_FP64>(*c).objective_plane((const double*)fi, flags, t, plane_x);
	return 0;
}

// pre-collision populations of every cell at step t (float, or double for FP64)
int fxo_load_populations(const fxo_cfg* c, const void* fi, uint64_t t, void* out) {
	if(!valid(c)) return -1;
	return dispatch<DoPopulations>(*c, fi, t, out);
}
int fxo_mass_sum(const fxo_cfg* c, const void* fi, const uint8_t* flags, uint64_t t, double* out) {
	if(!valid(c)) return -1;
	return dispatch<DoMass>(*c, fi, flags, t, out);
}
// object_force / object_torque / object_center_of_mass, src/kernel.cpp:1901-1959:
// per-cell values in the reference's FP32 arithmetic, summed in FP64 in cell order.
// F is float for FP32/FP16 storage and double for FP64.
int fxo_object_force(const fxo_cfg* c, const void* F, const uint8_t* flags, uint8_t marker, double* out) {
	if(!valid(c)) return -1;
	const uint64_t N = (uint64_t)c->Nx * c->Ny * c->Nz;
	double s[3] = {0, 0, 0};
	for(uint64_t n = 0; n < N; n++) {
		if(flags[n] != marker) continue;
		for(int d = 0; d < 3; d++) s[d] += c->storage == ST_FP64 ? ((const double*)F)[d * N + n] : (double)((const float*)F)[d * N + n];
	}
	for(int d = 0; d < 3; d++) out[d] = s[d];
	return 0;
}
int fxo_object_torque(const fxo_cfg* c, const void* F, const uint8_t* flags, uint8_t marker, float cx, float cy, float cz, double* out) {
	if(!valid(c)) return -1;
	const uint64_t N = (uint64_t)c->Nx * c->Ny * c->Nz;
	double s[3] = {0, 0, 0};
	for(uint64_t n = 0; n < N; n++) {
		if(flags[n] != marker) continue;
		float p[3]; position_f32(*c, n, p);
		if(c->storage == ST_FP64) {
			const double* Fd = (const double*)F;
			const double r[3] = {(double)p[0] - cx, (double)p[1] - cy, (double)p[2] - cz};
			const double f[3] = {Fd[n], Fd[N + n], Fd[2 * N + n]};
			s[0] += r[1] * f[2] - r[2] * f[1]; s[1] += r[2] * f[0] - r[0] * f[2]; s[2] += r[0] * f[1] - r[1] * f[0];
		} else {
			const float* Ff = (const float*)F;
			const float r[3] = {p[0] - cx, p[1] - cy, p[2] - cz};
			const float f[3] = {Ff[n], Ff[N + n], Ff[2 * N + n]};
			s[0] += (double)(r[1] * f[2] - r[2] * f[1]);
			s[1] += (double)(r[2] * f[0] - r[0] * f[2]);
			s[2] += (double)(r[0] * f[1] - r[1] * f[0]);
		}
	}
	for(int d = 0; d < 3; d++) out[d] = s[d];
	return 0;
}
int fxo_object_center_of_mass(const fxo_cfg* c, const uint8_t* flags, uint8_t marker, double* out) {
	if(!valid(c)) return -1;
	const uint64_t N = (uint64_t)c->Nx * c->Ny * c->Nz;
	double s[3] = {0, 0, 0}; double cells = 0.0;
	for(uint64_t n = 0; n < N; n++) {
		if(flags[n] != marker) continue;
		float p[3]; position_f32(*c, n, p);
		for(int d = 0; d < 3; d++) s[d] += (double)p[d];
		cells += 1.0;
	}
	for(int d = 0; d < 3; d++) out[d] = s[d];
	out[3] = cells;
	return 0;
}

} // extern "C"
