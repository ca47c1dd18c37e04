// Device code of the LBM hot path for gfx950: storage codecs, per-cell LBM arithmetic
// and the Esoteric-Pull in-place access pattern.  Included by the per-velocity-set
// translation units (lbm_q19.hip, lbm_q27.hip).
//
// Arithmetic follows the reference kernels operation by operation (src/kernel.cpp
// :1004-1101, 1454-1636, 1794-1870): explicit fmaf() where the reference writes fma(),
// everything else uncontracted (built with -ffp-contract=off), IEEE division (in the
// 16-bit kernels u = j/rho through SharedDivisor, which gives the same quotients).  Terms
// whose lattice coefficient is 0 are left out, as the reference's compiler does under
// -cl-no-signed-zeros (src/opencl.hpp:316); results agree with the oracle up to the
// sign of zero.
//
// Thread mapping: one lane per cell.  A block covers BX consecutive cells of one x-row;
// grid = (ceil(Nx/BX), Ny, Nz).  Everything that depends only on (y, z) - the periodic
// row offsets of the 9 neighbour rows and the slot base pointers - is wave-uniform and
// lives in scalar registers; per lane only x, x+1, x-1 (with wrap) remain, so every DDF
// access is a scalar-base + 32-bit lane offset load/store, coalesced along x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fx {

enum : uint32_t {
	EXT_VOLUME_FORCE = 1u, EXT_FORCE_FIELD = 2u, EXT_EQUILIBRIUM_BOUNDARIES = 4u,
	EXT_UPDATE_FIELDS = 8u, EXT_POROSITY = 16u, EXT_PRECISE_DENSITY = 32u, EXT_MOVING_BOUNDARIES = 64u,
	EXT_SUBGRID = 128u
};
enum : uint32_t { ST_FP32 = 0u, ST_FP16S = 1u, ST_FP16C = 2u };
constexpr uint32_t TYPE_S = 0x01u, TYPE_E = 0x02u, TYPE_MS = 0x03u, TYPE_BO = 0x03u, TYPE_T = 0x04u, TYPE_SU = 0x38u, TYPE_G = 0x20u;
constexpr float DEF_C = 0.57735027f; // lattice speed of sound, src/lbm.cpp:366
constexpr float THIRD = 0.33333334f; // src/kernel.cpp:1093-1099
constexpr float SMAGORINSKY = 0.76421222f; // 18*sqrt(2)*(C*Delta)^2, src/kernel.cpp:1591

// Kernel arguments, passed by value (kernarg segment -> scalar registers).
struct LbmArgs {
	void* fi;
	float* rho;
	float* u;
	const uint8_t* flags;
	const float* F;
	const float* alpha;
	uint64_t N;
	uint32_t Nx, Ny, Nz;
	uint32_t ext;
	uint32_t odd;   // parity of the time step t
	float w;
	float wm;       // TRT antisymmetric rate trt_minus_rate(w), formed once by the host
	float fx, fy, fz;
	uint32_t xchunks; // stream_collide on a 1-D XCD-mapped grid: x chunks per row (0: 3-D grid)
	uint32_t reverse; // stream_collide takes the blocks in reverse order (block_coords, cell_coords)
};
// TRT antisymmetric relaxation rate for magic parameter Lambda = 3/16 (src/kernel.cpp:
// 1605-1633): the same IEEE float operations on the host and the device give the same
// value, so the kernels take it from LbmArgs instead of three divisions per lane (SUBGRID,
// whose rate varies per cell, forms it in the kernel).
__host__ __device__ inline float trt_minus_rate(float w) { return 1.0f / (0.1875f / (1.0f / w - 0.5f) + 0.5f); }

// ------------------------------------------------------------------------------------
// DDF storage formats
// ------------------------------------------------------------------------------------
// Element o of the array at the wave-uniform base p, addressed as base + 32-bit byte
// offset: the form global_load/global_store take with a scalar base register (saddr) and
// one 32-bit VGPR offset per lane.  Callers keep o*sizeof(E) below 2^32 (rows of at most
// 2^26 cells).
template<typename E> __device__ __forceinline__ const E* elem(const E* p, uint32_t o) {
	return (const E*)((const char*)p + (uint32_t)(o * (uint32_t)sizeof(E)));
}
template<typename E> __device__ __forceinline__ E* elem(E* p, uint32_t o) {
	return (E*)((char*)p + (uint32_t)(o * (uint32_t)sizeof(E)));
}
// A row base that is the same in every lane of the wave, passed through readfirstlane: the
// compiler then keeps it in scalar registers, and the load and the later store of the same
// element do not share a 64-bit per-lane address held in vector registers across the
// collision.  Only for bases built from the block's row (never from a per-lane cell index).
template<typename E> __device__ __forceinline__ E* uniform(E* p) {
	const uint64_t v = (uint64_t)p;
	const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
	return (E*)(((uint64_t)hi << 32) | lo);
}
template<bool U, typename E> __device__ __forceinline__ E* row_base(E* p) { if constexpr(U) return uniform(p); else return p; }

template<uint32_t ST> struct Storage;
// DDFs are streamed with non-temporal loads and stores: no DDF value is touched twice in
// a step.  Slots at x-shifted neighbour positions are the exception: a wave's access there
// straddles a cache line that the neighbouring wave shares, and ordinary (cached) stores
// let the two halves meet in L2 instead of being written back in two non-temporal pieces.
// Measured at 256^3: FP32 D3Q19 non-temporal against cached at the unshifted slots +11 %,
// cached x-shifted stores +4.7 %; D3Q27 also keeps cached x-shifted loads (+3 %); FP16S
// +9 % from non-temporal loads and stores together at the unshifted slots (either alone
// gains nothing).  With rows that do not start on 128-byte lines (Nx times the element
// size not a multiple of 128 bytes) the kernels run on the cell-linear grid (see
// cell_coords): every wave then covers 64 consecutive cells, so its unshifted slots are
// whole lines again and stay non-temporal, while the slots at shifted positions (offsets
// of 1, Nx and Nx*Ny cells) straddle lines that neighbouring waves share and take cached
// accesses (raw_cached, store_cached).
template<> struct Storage<ST_FP32> {
	typedef float E;
	__device__ static __forceinline__ float raw(const E* p, uint32_t o) { return __builtin_nontemporal_load(elem(p, o)); }
	__device__ static __forceinline__ float decode(float x) { return x; }
	__device__ static __forceinline__ float load(const E* p, uint32_t o) { return raw(p, o); }
	__device__ static __forceinline__ void store(E* p, uint32_t o, float x) { __builtin_nontemporal_store(x, elem(p, o)); }
	__device__ static __forceinline__ float raw_cached(const E* p, uint32_t o) { return *elem(p, o); }
	__device__ static __forceinline__ void store_cached(E* p, uint32_t o, float x) { *elem(p, o) = x; }
	template<int Q> __device__ static __forceinline__ float raw_shifted(const E* p, uint32_t o) { return Q == 27 ? *elem(p, o) : raw(p, o); }
	template<int Q> __device__ static __forceinline__ void store_shifted(E* p, uint32_t o, float x) { *elem(p, o) = x; }
};
// FP16S (src/lbm.cpp:410-414): IEEE binary16 of x*2^15, round to nearest even
// (v_cvt_f16_f32 in the default rounding mode), load = binary16 * 2^-15.
template<> struct Storage<ST_FP16S> {
	typedef uint16_t E;
	__device__ static __forceinline__ float decode(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h) * 3.0517578E-5f; }
	__device__ static __forceinline__ uint16_t encode(float x) { return __builtin_bit_cast(uint16_t, (_Float16)(x * 32768.0f)); }
	__device__ static __forceinline__ uint32_t encode_word(float x) { return encode(x); }
	// (code << 16) | (low & 0xFFFF)
	__device__ static __forceinline__ uint32_t encode_high(float x, uint32_t low) { return __builtin_amdgcn_perm(encode_word(x), low, 0x05040100u); }
	__device__ static __forceinline__ uint16_t raw(const E* p, uint32_t o) { return __builtin_nontemporal_load(elem(p, o)); }
	__device__ static __forceinline__ float load(const E* p, uint32_t o) { return decode(raw(p, o)); }
	__device__ static __forceinline__ void store(E* p, uint32_t o, float x) { __builtin_nontemporal_store(encode(x), elem(p, o)); }
	__device__ static __forceinline__ uint16_t raw_cached(const E* p, uint32_t o) { return *elem(p, o); }
	__device__ static __forceinline__ void store_cached(E* p, uint32_t o, float x) { *elem(p, o) = encode(x); }
	template<int Q> __device__ static __forceinline__ uint16_t raw_shifted(const E* p, uint32_t o) { return *elem(p, o); }
	template<int Q> __device__ static __forceinline__ void store_shifted(E* p, uint32_t o, float x) { *elem(p, o) = encode(x); }
};
// FP16C: 1 sign, 4 exponent (bias 15), 11 mantissa bits (src/kernel.cpp:848-859).
// Decode: the code's exponent and mantissa fields placed at the binary32 positions give
// 2^-112 times the value, for normal codes and (as a binary32 denormal) for denormal
// codes alike; one exact multiplication by 2^112 finishes it (FP32 denormals are kept).
// Encode is the reference's device rule (add 0x800, truncate; no saturation) in fewer
// operations; it matches the literal rule on all 2^32 floats.
template<> struct Storage<ST_FP16C> {
	typedef uint16_t E;
	__device__ static __forceinline__ float decode(uint16_t h) {
		// sign-extended code << 12 puts the magnitude at bits 12..26 and the sign at 27..31
		const uint32_t bits = ((uint32_t)(int32_t)(int16_t)h << 12) & 0x87FFF000u;
		return __uint_as_float(bits) * 0x1p112f;
	}
	// decode of the low half of a two-cell word: one shift that takes the half
	// sign-extended as its operand (SDWA) instead of a shift up and an arithmetic shift down
	__device__ static __forceinline__ float decode_low(uint32_t w) {
		uint32_t t;
		asm("v_lshlrev_b32_sdwa %0, %1, sext(%2) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
		    : "=v"(t) : "v"(12u), "v"(w));
		return __uint_as_float(t & 0x87FFF000u) * 0x1p112f;
	}
	// With b = bits(x) + 0x800: the normal code's exponent field is (e_b - 112) mod 16 =
	// e_b mod 16, so the normal magnitude is bits 12..26 of b.  The reference's denormal
	// code ((0x7FF800 + m_b) >> (124 - e_b) + 1) >> 1 equals floor(|x|*2^25 + 1/2) for every
	// |x| below 2^-14*(1 - 2^-12) (e_b <= 112; zero from e_b <= 100 on), with or without a
	// carry out of the mantissa in b.  It is taken as the truncating conversion of
	// fma(|x|, 2^25, 1/2 - 2^-25): the single rounding of the fma turns the ties
	// |x|*2^25 = k - 1/2 into k (ties to even) and keeps every value below an integer below
	// it (checked on all 2^32 floats; the addend 1/2 itself fails at |x|*2^25 = 1/2 - 2^-25).
	// A NaN fails the comparison and takes the normal path, as in the reference.
	// The code in a 32-bit word, not truncated to 16 bits: it is below 2^16 for every input
	// (the normal code is a 15-bit field and the denormal one is taken only below 2^-14,
	// where it is below 2^11), but the compiler cannot see that and would mask the denormal
	// branch to 16 bits, one VALU operation per DDF; 16-bit stores and the two-cell packing
	// drop the upper bits themselves.
	// magnitude: the code without its sign, below 2^15; b = bits(x) + 0x800
	__device__ static __forceinline__ uint32_t magnitude(float x, uint32_t b) {
		const uint32_t normal = (b >> 12) & 0x7FFFu;
		const float ax = fabsf(x);
		const uint32_t denormal = __float2uint_rz(__builtin_fmaf(ax, 0x1p25f, 0x1.fffffep-2f));
		return ax < __uint_as_float(0x387FF800u) ? denormal : normal;
	}
	__device__ static __forceinline__ uint32_t encode_word(float x) {
		const uint32_t b = __float_as_uint(x) + 0x00000800u;
		return ((b >> 16) & 0x8000u) | magnitude(x, b);
	}
	// (code << 16) | (low & 0xFFFF), the code of the second cell of a two-cell word: the
	// magnitude goes to the high half with one byte permute and the sign (bit 31 of b, which
	// is bit 15 of the code shifted by 16) is inserted in place, one operation fewer than
	// shifting the sign down and the code up.
	__device__ static __forceinline__ uint32_t encode_high(float x, uint32_t low) {
		const uint32_t b = __float_as_uint(x) + 0x00000800u;
		const uint32_t t = __builtin_amdgcn_perm(magnitude(x, b), low, 0x05040100u);
		return (t & 0x7FFFFFFFu) | (b & 0x80000000u);
	}
	__device__ static __forceinline__ uint16_t encode(float x) { return (uint16_t)encode_word(x); }
	__device__ static __forceinline__ uint16_t raw(const E* p, uint32_t o) { return __builtin_nontemporal_load(elem(p, o)); }
	__device__ static __forceinline__ float load(const E* p, uint32_t o) { return decode(raw(p, o)); }
	__device__ static __forceinline__ void store(E* p, uint32_t o, float x) { __builtin_nontemporal_store(encode(x), elem(p, o)); }
	__device__ static __forceinline__ uint16_t raw_cached(const E* p, uint32_t o) { return *elem(p, o); }
	__device__ static __forceinline__ void store_cached(E* p, uint32_t o, float x) { *elem(p, o) = encode(x); }
	template<int Q> __device__ static __forceinline__ uint16_t raw_shifted(const E* p, uint32_t o) { return *elem(p, o); }
	template<int Q> __device__ static __forceinline__ void store_shifted(E* p, uint32_t o, float x) { *elem(p, o) = encode(x); }
};

// ------------------------------------------------------------------------------------
// Velocity sets (src/kernel.cpp:864-903; weights src/lbm.cpp:376-385).  The slot order
// is part of the memory layout, so it is the reference's.
// ------------------------------------------------------------------------------------
template<int Q> struct VS;
template<> struct VS<19> {
	static constexpr int8_t cx[19] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0};
	static constexpr int8_t cy[19] = {0, 0, 0, 1,-1, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1};
	static constexpr int8_t cz[19] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0,-1, 1,-1, 1};
	static constexpr float w0 = 1.0f / 3.0f, ws = 1.0f / 18.0f, we = 1.0f / 36.0f, wc = 0.0f;
};
template<> struct VS<27> {
	static constexpr int8_t cx[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1};
	static constexpr int8_t cy[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1, 1,-1};
	static constexpr int8_t cz[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0,-1, 1,-1, 1, 1,-1,-1, 1, 1,-1, 1,-1};
	static constexpr float w0 = 1.0f / 3.375f, ws = 1.0f / 13.5f, we = 1.0f / 54.0f, wc = 1.0f / 216.0f;
};
template<int Q> __device__ __forceinline__ constexpr float weight(int i) {
	return i == 0 ? VS<Q>::w0 : (VS<Q>::cx[i] * VS<Q>::cx[i] + VS<Q>::cy[i] * VS<Q>::cy[i] + VS<Q>::cz[i] * VS<Q>::cz[i]) == 1 ? VS<Q>::ws
	     : (VS<Q>::cx[i] * VS<Q>::cx[i] + VS<Q>::cy[i] * VS<Q>::cy[i] + VS<Q>::cz[i] * VS<Q>::cz[i]) == 2 ? VS<Q>::we : VS<Q>::wc;
}

// ------------------------------------------------------------------------------------
// Per-cell arithmetic
// ------------------------------------------------------------------------------------
__device__ __forceinline__ float clampc(float x) { return fminf(fmaxf(x, -DEF_C), DEF_C); } // OpenCL clamp

// calculate_f_eq (src/kernel.cpp:1004-1061), DDF-shifted equilibrium; rhom1 = rho-1 (the
// reference forms it as rho - 1.0f; with PRECISE_DENSITY it is the DDF sum itself).  The
// shared terms are formed once; pair(k) gives the two populations of the opposite pair
// (k, k+1), so a collision can take them one pair at a time.
template<int Q> struct Equilibrium {
	float c3, ux, uy, uz, f0, rs, re, rc, rm1s, rm1e, rm1c;
	__device__ __forceinline__ Equilibrium(float rho, float rhom1, float vx, float vy, float vz) {
		c3 = -3.0f * ((vx * vx + vy * vy) + vz * vz);
		uz = vz * 3.0f; ux = vx * 3.0f; uy = vy * 3.0f;
		f0 = VS<Q>::w0 * __builtin_fmaf(rho, 0.5f * c3, rhom1);
		rs = VS<Q>::ws * rho; re = VS<Q>::we * rho; rm1s = VS<Q>::ws * rhom1; rm1e = VS<Q>::we * rhom1;
		rc = VS<Q>::wc * rho; rm1c = VS<Q>::wc * rhom1;
	}
	__device__ __forceinline__ void pair(int k, float& a, float& b) const {
		float r = re, m = rm1e, uk;
		switch(k) {
			case 1: r = rs; m = rm1s; uk = ux; break;
			case 3: r = rs; m = rm1s; uk = uy; break;
			case 5: r = rs; m = rm1s; uk = uz; break;
			case 7: uk = ux + uy; break;
			case 9: uk = ux + uz; break;
			case 11: uk = uy + uz; break;
			case 13: uk = ux - uy; break;
			case 15: uk = ux - uz; break;
			case 17: uk = uy - uz; break;
			case 19: r = rc; m = rm1c; uk = (ux + uy) + uz; break;
			case 21: r = rc; m = rm1c; uk = (ux + uy) - uz; break;
			case 23: r = rc; m = rm1c; uk = (ux - uy) + uz; break;
			default: r = rc; m = rm1c; uk = (-ux + uy) + uz; break; // 25
		}
		const float t = __builtin_fmaf(uk, uk, c3);
		a = __builtin_fmaf(r, __builtin_fmaf(0.5f, t, uk), m);
		b = __builtin_fmaf(r, __builtin_fmaf(0.5f, t, -uk), m);
	}
};
template<int Q> __device__ __forceinline__ void calc_feq(float rho, float rhom1, float ux, float uy, float uz, float* feq) {
	const Equilibrium<Q> eq(rho, rhom1, ux, uy, uz);
	feq[0] = eq.f0;
#pragma unroll
	for(int k = 1; k < Q; k += 2) eq.pair(k, feq[k], feq[k + 1]);
}

// calculate_rho_u (src/kernel.cpp:1063-1088): rho summed in slot order, +1 last (DDF
// shifting); first moment j in the reference's alternating order; u = j/rho.
template<int Q> __device__ __forceinline__ void calc_rho_j(const float* f, float& rho, float& jx, float& jy, float& jz) {
	float r = f[0];
#pragma unroll
	for(int i = 1; i < Q; i++) r += f[i];
	rho = r + 1.0f;
	if constexpr(Q == 19) {
		jx = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[13] - f[14] + f[15] - f[16];
		jy = f[3] - f[4] + f[7] - f[8] + f[11] - f[12] + f[14] - f[13] + f[17] - f[18];
		jz = f[5] - f[6] + f[9] - f[10] + f[11] - f[12] + f[16] - f[15] + f[18] - f[17];
	} else {
		jx = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[13] - f[14] + f[15] - f[16] + f[19] - f[20] + f[21] - f[22] + f[23] - f[24] + f[26] - f[25];
		jy = f[3] - f[4] + f[7] - f[8] + f[11] - f[12] + f[14] - f[13] + f[17] - f[18] + f[19] - f[20] + f[21] - f[22] + f[24] - f[23] + f[25] - f[26];
		jz = f[5] - f[6] + f[9] - f[10] + f[11] - f[12] + f[16] - f[15] + f[18] - f[17] + f[19] - f[20] + f[22] - f[21] + f[23] - f[24] + f[25] - f[26];
	}
}
template<int Q> __device__ __forceinline__ void calc_rho_u(const float* f, float& rho, float& ux, float& uy, float& uz) {
	float jx, jy, jz;
	calc_rho_j<Q>(f, rho, jx, jy, jz);
	ux = jx / rho;
	uy = jy / rho;
	uz = jz / rho;
}

// c_i.v with the zero-coefficient terms left out, summed x, y, z in order
template<int Q> __device__ __forceinline__ float cdot(int i, float x, float y, float z) {
	const int a = VS<Q>::cx[i], b = VS<Q>::cy[i], c = VS<Q>::cz[i];
	float s;
	bool any = false;
	s = 0.0f;
	if(a != 0) { s = a > 0 ? x : -x; any = true; }
	if(b != 0) { s = any ? (b > 0 ? s + y : s - y) : (b > 0 ? y : -y); any = true; }
	if(c != 0) { s = any ? (c > 0 ? s + z : s - z) : (c > 0 ? z : -z); }
	return s;
}
// calculate_forcing_terms (src/kernel.cpp:1090-1101), Guo forcing; at(i) is population i's
// term, so a collision can take the terms one pair at a time.
template<int Q> struct Forcing {
	float uF, ux, uy, uz, fx, fy, fz;
	__device__ __forceinline__ Forcing(float vx, float vy, float vz, float gx, float gy, float gz)
		: ux(vx), uy(vy), uz(vz), fx(gx), fy(gy), fz(gz) {
		uF = -THIRD * __builtin_fmaf(ux, fx, __builtin_fmaf(uy, fy, uz * fz));
	}
	__device__ __forceinline__ float at(int i) const {
		if(i == 0) return (9.0f * VS<Q>::w0) * uF;
		return (9.0f * weight<Q>(i)) * __builtin_fmaf(cdot<Q>(i, fx, fy, fz), cdot<Q>(i, ux, uy, uz) + THIRD, uF);
	}
	// terms of the opposite pair (k, k+1), k odd: c_{k+1} = -c_k, and rounding to nearest is
	// symmetric, so the dot products of k+1 are exactly the negated ones of k
	__device__ __forceinline__ void pair(int k, float& a, float& b) const {
		const float cf = cdot<Q>(k, fx, fy, fz), cu = cdot<Q>(k, ux, uy, uz), wk = 9.0f * weight<Q>(k);
		a = wk * __builtin_fmaf(cf, cu + THIRD, uF);
		b = wk * __builtin_fmaf(-cf, -cu + THIRD, uF);
	}
};
template<int Q> __device__ __forceinline__ void calc_forcing(float ux, float uy, float uz, float fx, float fy, float fz, float* Fin) {
	const Forcing<Q> fo(ux, uy, uz, fx, fy, fz);
	Fin[0] = fo.at(0);
#pragma unroll
	for(int k = 1; k < Q; k += 2) fo.pair(k, Fin[k], Fin[k + 1]);
}

// Quotients a/rho with one reciprocal of rho shared between them, each equal to the IEEE
// division a/rho.  The compiler's division takes about 11 operations per quotient: a
// reciprocal r = rcp(d) refined by one Newton step, q = n*r, two corrections q += (n-d*q)*r,
// with v_div_scale scaling n and d beforehand when the operands could under- or overflow
// in between, and v_div_fixup for zero, infinite and NaN operands.  Where neither can
// apply, that arithmetic runs as it is, and r, formed once, serves every quotient:
// 3 operations once, 5 per quotient.  The scaling applies to none of these operands when
// d is normal with 2^-64 <= |d| <= 2^64 and a is 0 or 2^-39 <= |a| < 2^7: then 1/d, a/d and
// the residuals are normal, and the exponents of a and d differ by less than 96.
// Numerators that qualify: 0.5, and the momentum j of a 16-bit kernel without
// MOVING_BOUNDARIES, whose populations are all decoded codes, i.e. multiples of 2^-39
// (FP16S) or 2^-25 (FP16C) below 2 in magnitude, so that every partial sum is one too and
// |j| < 2*27.  Such a sum as the density, rho = 1 + sum f, is 0, non-finite (an FP16S code
// of inf/NaN) or normal with 2^-39 <= |rho| < 2^6, so there a normal rho (plain_sum, one
// class test) is enough; a preset density needs the range test (plain_any).  A lane whose
// rho fails takes the IEEE division.
struct SharedDivisor {
	float d, r;
	__device__ __forceinline__ explicit SharedDivisor(float dn) : d(dn) {
		const float r0 = __builtin_amdgcn_rcpf(dn);
		r = __builtin_fmaf(__builtin_fmaf(-dn, r0, 1.0f), r0, r0);
	}
	__device__ __forceinline__ float quotient(float a) const {
		float q = a * r;
		q = __builtin_fmaf(__builtin_fmaf(-d, q, a), r, q);
		return __builtin_fmaf(__builtin_fmaf(-d, q, a), r, q);
	}
	__device__ __forceinline__ bool plain_sum() const { return __builtin_amdgcn_classf(d, 0x108); } // -normal, +normal
	// |d| in [2^-64, 2^64]: (bits << 1) drops the sign; exponent fields 63..191
	__device__ __forceinline__ bool plain_any() const {
		return (__float_as_uint(d) << 1) - (0x1F800000u << 1) <= ((0x5F800000u - 0x1F800000u) << 1);
	}
};
// The kernels that may use SharedDivisor for u = j/rho (see there).
constexpr bool shared_division(uint32_t ST, uint32_t EXT) { return ST != ST_FP32 && !(EXT & EXT_MOVING_BOUNDARIES); }
// D3Q19 TRT kernels whose whole-array collision forms each opposite pair's equilibrium
// difference once (9 VALU instructions fewer per cell): those through buffer instructions
// and the all-extension ones, which drop 10 VGPRs with it (5 -> 6 waves/SIMD; 256^3 FP16S
// +8 %, FP16C +6.5 %).  Others held one VGPR more with it, which at 64 VGPRs costs a wave
// per SIMD; in D3Q27 the all-extension kernel measured 2 % slower and the C3 sphere kernel
// unchanged, so D3Q27 keeps the reference's form.
constexpr bool shared_pair_difference(int Q, uint32_t EXT, bool BUF) { return Q == 19 && (BUF || (EXT & EXT_SUBGRID)); }

// Force block of stream_collide/update_fields (src/kernel.cpp:1549-1563, 1847-1859)
// plus the Brinkman porosity term: u = (u + f*0.5/rho)/(1+alpha/2), F = f - alpha*rho*u.
// SD: 0.5/rho through SharedDivisor (its quotient is the IEEE one); PRESET: rhon may be
// a preset density (EQUILIBRIUM_BOUNDARIES) rather than a sum of decoded populations.
template<bool SD = false, bool PRESET = true>
__device__ __forceinline__ void force_block(uint32_t ext, float rhon, float& ux, float& uy, float& uz,
                                            float& fxn, float& fyn, float& fzn, float alphan) {
	if(ext & (EXT_VOLUME_FORCE | EXT_POROSITY)) {
		float rho2;
		if constexpr(SD) { // rhon is the populations' sum or, with EQUILIBRIUM_BOUNDARIES, a preset
			const SharedDivisor dv(rhon);
			rho2 = dv.quotient(0.5f);
			if(!(PRESET ? dv.plain_any() : dv.plain_sum())) rho2 = 0.5f / rhon;
		} else rho2 = 0.5f / rhon;
		ux = __builtin_fmaf(fxn, rho2, ux);
		uy = __builtin_fmaf(fyn, rho2, uy);
		uz = __builtin_fmaf(fzn, rho2, uz);
		if(ext & EXT_POROSITY) {
			const float k = 1.0f / (1.0f + 0.5f * alphan);
			ux = ux * k; uy = uy * k; uz = uz * k;
		}
		ux = clampc(ux); uy = clampc(uy); uz = clampc(uz);
		if(ext & EXT_POROSITY) {
			const float ar = -(alphan * rhon);
			fxn = __builtin_fmaf(ar, ux, fxn);
			fyn = __builtin_fmaf(ar, uy, fyn);
			fzn = __builtin_fmaf(ar, uz, fzn);
		}
	} else {
		ux = clampc(ux); uy = clampc(uy); uz = clampc(uz);
	}
}

// ------------------------------------------------------------------------------------
// Lattice addressing for one lane: x of this lane, the uniform (y,z) row offsets.
// ------------------------------------------------------------------------------------
struct Site {
	uint32_t X[3];   // x-1, x, x+1 (periodic), lane values
	uint64_t R[9];   // row offsets (y+dy)*Nx + (z+dz)*Nx*Ny for dy,dz in {-1,0,1}, uniform
	uint32_t x;
	__device__ __forceinline__ uint64_t row(int dy, int dz) const { return R[(dy + 1) * 3 + (dz + 1)]; }
	__device__ __forceinline__ uint64_t n() const { return R[4] + x; }
};
__device__ __forceinline__ Site make_site(uint32_t x, uint32_t y, uint32_t z, uint32_t Nx, uint32_t Ny, uint32_t Nz) {
	Site s;
	s.x = x;
	s.X[0] = x == 0u ? Nx - 1u : x - 1u;
	s.X[1] = x;
	s.X[2] = x + 1u == Nx ? 0u : x + 1u;
	const uint32_t Y[3] = {y == 0u ? Ny - 1u : y - 1u, y, y + 1u == Ny ? 0u : y + 1u};
	const uint32_t Z[3] = {z == 0u ? Nz - 1u : z - 1u, z, z + 1u == Nz ? 0u : z + 1u};
	const uint64_t NxNy = (uint64_t)Nx * Ny;
#pragma unroll
	for(int a = 0; a < 3; a++)
#pragma unroll
		for(int b = 0; b < 3; b++) s.R[a * 3 + b] = (uint64_t)Y[a] * Nx + (uint64_t)Z[b] * NxNy;
	return s;
}

// Esoteric-Pull load_f / store_f (src/kernel.cpp:1326-1339).  For odd i the cell owns
// slot pair (n, i|i+1) and (n+e_i, i+1|i); which one holds which population flips with
// the parity of t.  Loads and stores touch the same locations, so the update is in place
// and race-free.
// U: the row offsets in s are the block's (wave-uniform); false where a lane has a row of
// its own (the objective plane's one lane per (y, z)).
// SC: cached accesses at every shifted slot (the cell-linear grid); false: non-temporal
// at the slots shifted in y/z only, the format's rule at the x-shifted ones.
// With SC the 16-bit formats address each slot as its uniform base plus a 32-bit per-lane
// element offset (a slot is below 4 GiB wherever the cell-linear grid runs them,
// use_linear_grid): one register per address instead of a 64-bit pair, one for all
// unshifted slots; 69-71 instead of 79 VGPRs for the plain kernels, measured +0.4 to
// +7.6 % on 14 of the 16 D3Q19 16-bit variants at 200^3, D3Q27 -4.8 to +18 % (11 of 16
// up); FP32 -2.3 to +1 %, not taken.
// O32: the same 32-bit offsets on the row grid (the row grid's access rule), for lattices
// whose slots are below 4 GiB; tune_offsets32 selects the kernels that gain from it.
template<int Q, uint32_t ST> constexpr bool slot_offsets32(bool SC) { return SC && ST != ST_FP32; }
// BUF: the row grid through buffer instructions, for a DDF array below 4 GiB.  The address
// of every access is the array's base (a buffer resource in scalar registers) plus a scalar
// offset (slot and row, wave-uniform) plus a 32-bit lane offset, of which there are three
// per lane (x-1, x, x+1): no per-access vector address arithmetic, and no 64-bit addresses
// held in vector registers from the loads to the stores.
// The resource: base fi, stride 0 (raw buffer), the largest record count (the range check
// never clips an offset of a DDF array below 4 GiB), and descriptor word 3 = 0x00020000,
// data format 32 bits, which the untyped buffer_load/store_short/dword of the gfx9 family
// need to be a valid format (the value AMD's composable_kernel uses for gfx90a/gfx94x/gfx950).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ddf_buffer(const void* fi) {
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(fi), (short)0, (int)0xFFFFFFFFu, 0x00020000);
}
// uniform_bytes is the same in every lane; readfirstlane says so to the compiler, which
// otherwise may not see it through the parity selects and wraps the access in a loop over
// the distinct values.  NT: non-temporal (cache policy bit 1).
template<typename E, bool NT> __device__ __forceinline__ E buffer_elem(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t uniform_bytes) {
	const uint32_t u = __builtin_amdgcn_readfirstlane(uniform_bytes);
	if constexpr(sizeof(E) == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, lane_bytes, u, NT ? 2 : 0);
	else return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(r, lane_bytes, u, NT ? 2 : 0));
}
template<typename E, bool NT> __device__ __forceinline__ void buffer_elem(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t uniform_bytes, E v) {
	const uint32_t u = __builtin_amdgcn_readfirstlane(uniform_bytes);
	if constexpr(sizeof(E) == 2) __builtin_amdgcn_raw_buffer_store_b16(v, r, lane_bytes, u, NT ? 2 : 0);
	else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, lane_bytes, u, NT ? 2 : 0);
}
// the stored element of a population value
template<uint32_t ST> __device__ __forceinline__ typename Storage<ST>::E stored(float x) {
	if constexpr(ST == ST_FP32) return x;
	else return Storage<ST>::encode(x);
}
template<int Q, uint32_t ST, bool U = false, bool SC = false, bool O32 = false, bool BUF = false>
__device__ __forceinline__ void load_raw(const typename Storage<ST>::E* fi, uint64_t N, const Site& s, uint32_t odd, typename Storage<ST>::E* r) {
	typedef Storage<ST> S;
	if constexpr(BUF && SC) { // the cell-linear grid: rows per lane, every shifted slot cached
		typedef typename Storage<ST>::E E;
		constexpr uint32_t eb = sizeof(E);
		const __amdgpu_buffer_rsrc_t rs = ddf_buffer(fi);
		const uint32_t slot = (uint32_t)N * eb, c = ((uint32_t)s.R[4] + s.x) * eb;
		r[0] = buffer_elem<E, true>(rs, c, 0u);
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			const uint32_t sa = odd ? (uint32_t)i : (uint32_t)(i + 1), sb = odd ? (uint32_t)(i + 1) : (uint32_t)i;
			r[i] = buffer_elem<E, true>(rs, c, sa * slot);
			const uint32_t o = ((uint32_t)s.row(VS<Q>::cy[i], VS<Q>::cz[i]) + s.X[VS<Q>::cx[i] + 1]) * eb;
			r[i + 1] = buffer_elem<E, false>(rs, o, sb * slot);
		}
		return;
	} else if constexpr(BUF) {
		typedef typename Storage<ST>::E E;
		constexpr uint32_t eb = sizeof(E);
		constexpr bool shifted_nt = ST == ST_FP32 && Q == 19; // x-shifted loads: raw_shifted's rule
		const __amdgpu_buffer_rsrc_t rs = ddf_buffer(fi);
		const uint32_t slot = (uint32_t)N * eb, n0 = (uint32_t)s.R[4] * eb;
		const uint32_t v[3] = {s.X[0] * eb, s.x * eb, s.X[2] * eb};
		r[0] = buffer_elem<E, true>(rs, v[1], n0);
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			const uint32_t sa = odd ? (uint32_t)i : (uint32_t)(i + 1), sb = odd ? (uint32_t)(i + 1) : (uint32_t)i;
			r[i] = buffer_elem<E, true>(rs, v[1], sa * slot + n0);
			const uint32_t ob = sb * slot + (uint32_t)s.row(VS<Q>::cy[i], VS<Q>::cz[i]) * eb;
			if(VS<Q>::cx[i] != 0) r[i + 1] = buffer_elem<E, shifted_nt>(rs, v[VS<Q>::cx[i] + 1], ob);
			else r[i + 1] = buffer_elem<E, true>(rs, v[1], ob);
		}
		return;
	}
	if constexpr(O32 || slot_offsets32<Q, ST>(SC)) {
		const uint32_t c = (uint32_t)s.R[4] + s.x; // the cell's index
		r[0] = S::raw(fi, c);
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			const uint64_t sa = odd ? (uint64_t)i : (uint64_t)(i + 1);
			const uint64_t sb = odd ? (uint64_t)(i + 1) : (uint64_t)i;
			r[i] = S::raw(fi + sa * N, c);
			const uint32_t o = (uint32_t)s.row(VS<Q>::cy[i], VS<Q>::cz[i]) + s.X[VS<Q>::cx[i] + 1];
			if(SC || VS<Q>::cx[i] != 0) r[i + 1] = S::raw_cached(fi + sb * N, o);
			else r[i + 1] = S::raw(fi + sb * N, o);
		}
		return;
	}
	const uint64_t n0 = s.R[4];
	r[0] = S::raw(row_base<U>(fi + n0), s.x);
#pragma unroll
	for(int i = 1; i < Q; i += 2) {
		const uint64_t sa = odd ? (uint64_t)i : (uint64_t)(i + 1);
		const uint64_t sb = odd ? (uint64_t)(i + 1) : (uint64_t)i;
		r[i] = S::raw(row_base<U>(fi + sa * N + n0), s.x);
		const typename Storage<ST>::E* pb = row_base<U>(fi + sb * N + s.row(VS<Q>::cy[i], VS<Q>::cz[i]));
		const uint32_t xb = s.X[VS<Q>::cx[i] + 1];
		if(SC) r[i + 1] = S::raw_cached(pb, xb);
		else r[i + 1] = VS<Q>::cx[i] != 0 ? S::template raw_shifted<Q>(pb, xb) : S::raw(pb, xb);
	}
}
template<int Q, uint32_t ST, bool U = false, bool SC = false, bool O32 = false, bool BUF = false>
__device__ __forceinline__ void load_f(const typename Storage<ST>::E* fi, uint64_t N, const Site& s, uint32_t odd, float* f) {
	typename Storage<ST>::E r[Q];
	load_raw<Q, ST, U, SC, O32, BUF>(fi, N, s, odd, r);
#pragma unroll
	for(int i = 0; i < Q; i++) f[i] = Storage<ST>::decode(r[i]);
}
template<int Q, uint32_t ST, bool U = false, bool SC = false, bool O32 = false, bool BUF = false>
__device__ __forceinline__ void store_f(typename Storage<ST>::E* fi, uint64_t N, const Site& s, uint32_t odd, const float* f) {
	typedef Storage<ST> S;
	if constexpr(BUF && SC) {
		typedef typename Storage<ST>::E E;
		constexpr uint32_t eb = sizeof(E);
		const __amdgpu_buffer_rsrc_t rs = ddf_buffer(fi);
		const uint32_t slot = (uint32_t)N * eb, c = ((uint32_t)s.R[4] + s.x) * eb;
		buffer_elem<E, true>(rs, c, 0u, stored<ST>(f[0]));
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			const uint32_t sa = odd ? (uint32_t)i : (uint32_t)(i + 1), sb = odd ? (uint32_t)(i + 1) : (uint32_t)i;
			const uint32_t o = ((uint32_t)s.row(VS<Q>::cy[i], VS<Q>::cz[i]) + s.X[VS<Q>::cx[i] + 1]) * eb;
			buffer_elem<E, false>(rs, o, sb * slot, stored<ST>(f[i]));
			buffer_elem<E, true>(rs, c, sa * slot, stored<ST>(f[i + 1]));
		}
		return;
	} else if constexpr(BUF) {
		typedef typename Storage<ST>::E E;
		constexpr uint32_t eb = sizeof(E);
		const __amdgpu_buffer_rsrc_t rs = ddf_buffer(fi);
		const uint32_t slot = (uint32_t)N * eb, n0 = (uint32_t)s.R[4] * eb;
		const uint32_t v[3] = {s.X[0] * eb, s.x * eb, s.X[2] * eb};
		buffer_elem<E, true>(rs, v[1], n0, stored<ST>(f[0]));
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			const uint32_t sa = odd ? (uint32_t)i : (uint32_t)(i + 1), sb = odd ? (uint32_t)(i + 1) : (uint32_t)i;
			const uint32_t ob = sb * slot + (uint32_t)s.row(VS<Q>::cy[i], VS<Q>::cz[i]) * eb;
			if(VS<Q>::cx[i] != 0) buffer_elem<E, false>(rs, v[VS<Q>::cx[i] + 1], ob, stored<ST>(f[i]));
			else buffer_elem<E, true>(rs, v[1], ob, stored<ST>(f[i]));
			buffer_elem<E, true>(rs, v[1], sa * slot + n0, stored<ST>(f[i + 1]));
		}
		return;
	}
	if constexpr(O32 || slot_offsets32<Q, ST>(SC)) {
		const uint32_t c = (uint32_t)s.R[4] + s.x;
		S::store(fi, c, f[0]);
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			const uint64_t sa = odd ? (uint64_t)i : (uint64_t)(i + 1);
			const uint64_t sb = odd ? (uint64_t)(i + 1) : (uint64_t)i;
			const uint32_t o = (uint32_t)s.row(VS<Q>::cy[i], VS<Q>::cz[i]) + s.X[VS<Q>::cx[i] + 1];
			if(SC || VS<Q>::cx[i] != 0) S::store_cached(fi + sb * N, o, f[i]);
			else S::store(fi + sb * N, o, f[i]);
			S::store(fi + sa * N, c, f[i + 1]);
		}
		return;
	}
	const uint64_t n0 = s.R[4];
	S::store(row_base<U>(fi + n0), s.x, f[0]);
#pragma unroll
	for(int i = 1; i < Q; i += 2) {
		const uint64_t sa = odd ? (uint64_t)i : (uint64_t)(i + 1);
		const uint64_t sb = odd ? (uint64_t)(i + 1) : (uint64_t)i;
		typename Storage<ST>::E* pb = row_base<U>(fi + sb * N + s.row(VS<Q>::cy[i], VS<Q>::cz[i]));
		const uint32_t xb = s.X[VS<Q>::cx[i] + 1];
		if(VS<Q>::cx[i] != 0) S::template store_shifted<Q>(pb, xb, f[i]);
		else if(SC) S::store_cached(pb, xb, f[i]);
		else S::store(pb, xb, f[i]);
		S::store(row_base<U>(fi + sa * N + n0), s.x, f[i + 1]);
	}
}

// ------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------
constexpr int BLOCK = 256;

// initialize (src/kernel.cpp:1358-1430), non-moving boundaries
// neighbour cell (dx,dy,dz) of a site
template<int Q> __device__ __forceinline__ uint64_t neighbour(const Site& s, int i) {
	return s.row(VS<Q>::cy[i], VS<Q>::cz[i]) + s.X[VS<Q>::cx[i] + 1];
}
// true if a neighbour of the cell is a solid cell with non-zero velocity (src/kernel.cpp
// :1395-1398, 1441-1444)
template<int Q> __device__ __forceinline__ bool next_to_moving_solid(const LbmArgs& a, const Site& s) {
	bool m = false;
#pragma unroll
	for(int i = 1; i < Q; i++) {
		const uint64_t j = neighbour<Q>(s, i);
		m = m || ((a.flags[j] & TYPE_BO) == TYPE_S && (a.u[j] != 0.0f || a.u[a.N + j] != 0.0f || a.u[2ull * a.N + j] != 0.0f));
	}
	return m;
}
// initialize (src/kernel.cpp:1358-1430), no SURFACE/TEMPERATURE.  Without MOVING_BOUNDARIES
// every solid cell's velocity is zeroed; with it only that of solid cells whose neighbours
// are all solid, and every other non-TYPE_E cell is marked (or unmarked) TYPE_MS.  A cell
// writes only its own flags and velocity, and the marks never change what the other cells
// test (a solid type), so the in-place update is race-free.
template<int Q, uint32_t ST>
__global__ __launch_bounds__(BLOCK) void k_initialize(LbmArgs a) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if(x >= a.Nx) return;
	const Site s = make_site(x, blockIdx.y, blockIdx.z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	float ux = a.u[n], uy = a.u[a.N + n], uz = a.u[2ull * a.N + n];
	const uint32_t flagsn = a.flags[n], bo = flagsn & TYPE_BO;
	bool zero = bo == TYPE_S;
	if(a.ext & EXT_MOVING_BOUNDARIES) {
		if(bo == TYPE_S) {
#pragma unroll
			for(int i = 1; i < Q; i++) zero = zero && (a.flags[neighbour<Q>(s, i)] & TYPE_BO) == TYPE_S;
		} else if(bo != TYPE_E) {
			((uint8_t*)a.flags)[n] = (uint8_t)(next_to_moving_solid<Q>(a, s) ? flagsn | TYPE_MS : flagsn & ~TYPE_MS);
		}
	}
	if(zero) {
		ux = 0.0f; uy = 0.0f; uz = 0.0f;
		a.u[n] = 0.0f; a.u[a.N + n] = 0.0f; a.u[2ull * a.N + n] = 0.0f;
	}
	float feq[Q];
	calc_feq<Q>(a.rho[n], a.rho[n] - 1.0f, ux, uy, uz, feq);
	store_f<Q, ST>((typename Storage<ST>::E*)a.fi, a.N, s, 1u, feq);
}

// apply_moving_boundaries (src/kernel.cpp:1104-1113): a population that arrives from a
// solid neighbour (from n - e_i for f_i) gains -6 w_i (c_opposite . u_wall), wall density 1.
template<int Q> __device__ __forceinline__ void moving_boundaries(const LbmArgs& a, const Site& s, float* fhn) {
#pragma unroll
	for(int i = 1; i < Q; i += 2) {
		const float w6 = -6.0f * weight<Q>(i);
		uint64_t j = neighbour<Q>(s, i + 1);
		if((a.flags[j] & TYPE_BO) == TYPE_S) fhn[i] = __builtin_fmaf(w6, cdot<Q>(i + 1, a.u[j], a.u[a.N + j], a.u[2ull * a.N + j]), fhn[i]);
		j = neighbour<Q>(s, i);
		if((a.flags[j] & TYPE_BO) == TYPE_S) fhn[i + 1] = __builtin_fmaf(w6, cdot<Q>(i, a.u[j], a.u[a.N + j], a.u[2ull * a.N + j]), fhn[i + 1]);
	}
}
// update_moving_boundaries (src/kernel.cpp:1433-1449): re-mark TYPE_MS after wall velocities
// changed; not on solid, TYPE_E or TYPE_T cells.
template<int Q>
__global__ __launch_bounds__(BLOCK) void k_update_moving_boundaries(LbmArgs a) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if(x >= a.Nx) return;
	const Site s = make_site(x, blockIdx.y, blockIdx.z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	const uint32_t flagsn = a.flags[n], bo = flagsn & TYPE_BO;
	if(bo == TYPE_S || bo == TYPE_E || (flagsn & TYPE_T)) return;
	((uint8_t*)a.flags)[n] = (uint8_t)(next_to_moving_solid<Q>(a, s) ? flagsn | TYPE_MS : flagsn & ~TYPE_MS);
}
// SUBGRID (src/kernel.cpp:1579-1593), Smagorinsky-Lilly: local relaxation rate from the
// non-equilibrium stress H = sum_i c_i c_i (f_i - feq_i), i >= 1 (zero-coefficient terms
// left out), w = 2/(tau0 + sqrt(tau0^2 + 0.76421222 sqrt(H:H)/rho)), tau0 = 1/w.
template<int Q> __device__ __forceinline__ float subgrid_rate(float w, const float* fhn, const float* feq, float rhon) {
	const float tau0 = 1.0f / w;
	float Hxx = 0.0f, Hyy = 0.0f, Hzz = 0.0f, Hxy = 0.0f, Hxz = 0.0f, Hyz = 0.0f;
#pragma unroll
	for(int i = 1; i < Q; i++) {
		const float fneq = fhn[i] - feq[i];
		const int ax = VS<Q>::cx[i], ay = VS<Q>::cy[i], az = VS<Q>::cz[i];
		if(ax != 0) Hxx += fneq;
		if(ax * ay != 0) Hxy = ax * ay > 0 ? Hxy + fneq : Hxy - fneq;
		if(ay != 0) Hyy += fneq;
		if(ax * az != 0) Hxz = ax * az > 0 ? Hxz + fneq : Hxz - fneq;
		if(ay * az != 0) Hyz = ay * az > 0 ? Hyz + fneq : Hyz - fneq;
		if(az != 0) Hzz += fneq;
	}
	const float Qs = ((Hxx * Hxx + Hyy * Hyy) + Hzz * Hzz) + 2.0f * ((Hxy * Hxy + Hxz * Hxz) + Hyz * Hyz);
	return 2.0f / (tau0 + sqrtf(tau0 * tau0 + SMAGORINSKY * sqrtf(Qs) / rhon));
}

// stream_collide (src/kernel.cpp:1454-1636).  EXT is a compile-time mask of the
// extensions this instantiation may use; the runtime mask a.ext selects among them.
// EXT = 0 is the plain scheme with every extension path compiled out.
//
// collide_cell: everything between load_f and store_f for the fluid or TYPE_E cell n;
// fhn holds the streamed-in DDFs on entry and the post-collision DDFs on return.
// SRT/TRT collision (src/kernel.cpp:1595-1633) one opposite pair at a time: each pair's
// equilibrium and forcing terms are formed right where the pair collides, the same
// operations as with whole arrays, so only a pair's values are live at once.  Used where it
// measured faster (256^3, same process): the forced kernels (VOLUME_FORCE, POROSITY) of
// D3Q19 and the forced TRT kernels of D3Q27, +10-14 % for the 16-bit TRT ones; elsewhere
// the whole-array form was as fast or faster (D3Q27 FP16C TRT EQUILIBRIUM_BOUNDARIES: -7 %
// by pairs).  SUBGRID always takes whole arrays.
constexpr bool collide_by_pairs(int Q, bool TRT, uint32_t EXT) {
	return (EXT & EXT_SUBGRID) == 0u && (EXT & (EXT_VOLUME_FORCE | EXT_POROSITY)) != 0u && (TRT || Q == 19);
}
// kernels_for selects the VOLUME_FORCE subset only when exactly that extension is on, so
// the mask can be a compile-time constant there (the unforced paths compile out).  Measured
// at 256^3: D3Q27 FP16S/FP16C +9 to +18 %; on the D3Q19 row grid, with plain row bases
// (uniform_rows), FP16S TRT +5.7 %, FP16C SRT +3.1 %, FP16C TRT +8.4 %, FP16S SRT -2 % (it
// keeps the runtime mask there); on the D3Q19 cell-linear grid (200^3) SRT +26 % (FP16S)
// and +21.5 % (FP16C, 98 -> 75-79 VGPRs), TRT -1.3 to +0.6 %; FP32 -5 to +1.5 %.
template<bool TRT> constexpr bool constant_vf_mask(int Q, uint32_t ST, uint32_t EXT, bool LIN) {
	if(ST == ST_FP32 || EXT != EXT_VOLUME_FORCE) return false;
	return Q == 27 || LIN || TRT || ST == ST_FP16C;
}
// Uniform row bases (readfirstlane) for stream_collide: the forced D3Q19 16-bit kernels
// drop from 96-99 to 58-66 VGPRs and gain 2-20 %, the unforced D3Q19 FP32 TRT kernels
// 2-4 %; D3Q27 (27 slot bases in scalar registers) and the unforced 16-bit kernels lost up
// to 30 %, the forced FP32 ones up to 1 %, so they keep the plain bases.  The 16-bit
// VOLUME_FORCE kernels with a constant extension mask (constant_vf_mask) hold 59-63 VGPRs
// with plain bases as well and need fewer instructions that way.
template<bool TRT> constexpr bool uniform_rows(int Q, uint32_t ST, uint32_t EXT) {
	const bool forced = (EXT & (EXT_VOLUME_FORCE | EXT_POROSITY)) != 0u;
	if(Q != 19) return false;
	return ST == ST_FP32 ? TRT && !forced : forced && !constant_vf_mask<TRT>(Q, ST, EXT, false);
}
// FP16S kernels that issue their DDF loads at raised wave priority (s_setprio 1, back to 0
// once all are issued).  Measured at 256^3 against the same kernels without: plain D3Q19
// +1.3 % (SRT) / +1.4 % (TRT), D3Q27 +1.3 % / +3.4 %; TRT VOLUME_FORCE +3.6 % (D3Q19), +8.5 %
// (D3Q27); EQUILIBRIUM_BOUNDARIES (+FORCE_FIELD) -0.2 to +0.6 %; the porous and generic
// kernels -1 to -5 % and SRT VOLUME_FORCE -0.8 % (D3Q19) go without; FP32 and FP16C lost
// 3.5 % and 0.5 % with it.
template<bool TRT> constexpr bool load_priority(uint32_t EXT) {
	return EXT == 0u || EXT == EXT_EQUILIBRIUM_BOUNDARIES || EXT == (EXT_EQUILIBRIUM_BOUNDARIES | EXT_FORCE_FIELD) ||
	       (TRT && EXT == EXT_VOLUME_FORCE);
}
template<int Q, bool TRT>
__device__ __forceinline__ void collide_pairs(float w, float wm, bool vf, bool isE, const Equilibrium<Q>& eq, const Forcing<Q>& fo, float* fhn) {
	if constexpr(!TRT) { // src/kernel.cpp:1595-1604
		const float c_tau = __builtin_fmaf(w, -0.5f, 1.0f);
		const float omw = 1.0f - w;
		// the forcing terms one population at a time: the pairwise form (Forcing::pair)
		// measured 1.5-8 % slower here for the D3Q19 16-bit kernels
		auto one = [&](int i, float fe) {
			const float Fi = vf ? fo.at(i) * c_tau : 0.0f;
			const float c = __builtin_fmaf(omw, fhn[i], __builtin_fmaf(w, fe, Fi));
			fhn[i] = isE ? fe : c;
		};
		one(0, eq.f0);
#pragma unroll
		for(int k = 1; k < Q; k += 2) {
			float ea, eb;
			eq.pair(k, ea, eb);
			one(k, ea);
			one(k + 1, eb);
		}
	} else { // src/kernel.cpp:1605-1633, Lambda = 3/16
		const float wp = w;
		const float c_taup = __builtin_fmaf(wp, -0.25f, 0.5f), c_taum = __builtin_fmaf(wm, -0.25f, 0.5f);
		const float hwp = 0.5f * wp, hwm = 0.5f * wm;
		// dm = fe - feb.  D3Q19: the pair's two populations take dm and -dm, exactly the
		// reference's feq[i]-feb[i] for both except that an exact zero difference becomes -0
		// for one of them, which reaches the result only as the sign of an exact zero.
		auto trt = [&](float fe, float feb, float fh, float fhb, float dm, float F) {
			return __builtin_fmaf(hwp, ((fe - fh) + feb) - fhb, __builtin_fmaf(hwm, (dm - fh) + fhb, fh + F));
		};
		{
			float F0 = 0.0f;
			if(vf) {
				const float g = fo.at(0);
				F0 = __builtin_fmaf(c_taup, g + g, c_taum * (g - g));
			}
			const float c = trt(eq.f0, eq.f0, fhn[0], fhn[0], eq.f0 - eq.f0, F0);
			fhn[0] = isE ? eq.f0 : c;
		}
#pragma unroll
		for(int k = 1; k < Q; k += 2) {
			float ea, eb;
			eq.pair(k, ea, eb);
			float Fa = 0.0f, Fb = 0.0f;
			if(vf) {
				float ga, gb;
				fo.pair(k, ga, gb);
				Fa = __builtin_fmaf(c_taup, ga + gb, c_taum * (ga - gb));
				// c_taum*(gb-ga) as -(c_taum*(ga-gb)), equal but for the sign of a zero
				Fb = __builtin_fmaf(c_taup, gb + ga, -(c_taum * (ga - gb)));
			}
			const float d = ea - eb;
			const float ca = trt(ea, eb, fhn[k], fhn[k + 1], d, Fa);
			const float cb = trt(eb, ea, fhn[k + 1], fhn[k], Q == 19 ? -d : eb - ea, Fb);
			fhn[k] = isE ? ea : ca;
			fhn[k + 1] = isE ? eb : cb;
		}
	}
}
// SD: u = j/rho and 0.5/rho through SharedDivisor; only where shared_division allows it.
// SHARE: the whole-array TRT collision forms each opposite pair's equilibrium difference
// once (shared_pair_difference).
template<int Q, bool TRT, uint32_t EXT, bool RECORD, bool SD = false, bool SHARE = false>
__device__ __forceinline__ void collide_cell(const LbmArgs& a, uint64_t n, uint32_t bo, float* fhn, float* moments) {
	const bool isE = (a.ext & EXT_EQUILIBRIUM_BOUNDARIES) && bo == TYPE_E;
	float rhon, uxn, uyn, uzn, rhom1;
	{
		float jx, jy, jz;
		calc_rho_j<Q>(fhn, rhon, jx, jy, jz);
		if(a.ext & EXT_PRECISE_DENSITY) {
			rhom1 = fhn[0];
#pragma unroll
			for(int i = 1; i < Q; i++) rhom1 += fhn[i];
		} else rhom1 = rhon - 1.0f;
		if constexpr(RECORD) { // pre-collision moments for the adjoint sweep
			moments[n] = rhon; moments[a.N + n] = jx; moments[2ull * a.N + n] = jy; moments[3ull * a.N + n] = jz;
		}
		if constexpr(SD) {
			const SharedDivisor dv(rhon);
			uxn = dv.quotient(jx); uyn = dv.quotient(jy); uzn = dv.quotient(jz);
			if(!dv.plain_sum()) { uxn = jx / rhon; uyn = jy / rhon; uzn = jz / rhon; }
		} else {
			uxn = jx / rhon; uyn = jy / rhon; uzn = jz / rhon;
		}
	}
	if(isE) { // preset density/velocity of equilibrium boundary cells
		rhon = a.rho[n]; uxn = a.u[n]; uyn = a.u[a.N + n]; uzn = a.u[2ull * a.N + n];
		rhom1 = rhon - 1.0f;
	}
	float fxn = a.fx, fyn = a.fy, fzn = a.fz;
	if(a.ext & EXT_FORCE_FIELD) { fxn += a.F[n]; fyn += a.F[a.N + n]; fzn += a.F[2ull * a.N + n]; }
	const float alphan = (a.ext & EXT_POROSITY) ? a.alpha[n] : 0.0f;
	force_block<SD, (EXT & EXT_EQUILIBRIUM_BOUNDARIES) != 0u>(a.ext, rhon, uxn, uyn, uzn, fxn, fyn, fzn, alphan);
	const bool vf = a.ext & (EXT_VOLUME_FORCE | EXT_POROSITY);
	if((a.ext & EXT_UPDATE_FIELDS) && !isE) {
		a.rho[n] = rhon; a.u[n] = uxn; a.u[a.N + n] = uyn; a.u[2ull * a.N + n] = uzn;
	}
	if constexpr(collide_by_pairs(Q, TRT, EXT)) {
		collide_pairs<Q, TRT>(a.w, a.wm, vf, isE, Equilibrium<Q>(rhon, rhom1, uxn, uyn, uzn), Forcing<Q>(uxn, uyn, uzn, fxn, fyn, fzn), fhn);
		return;
	}
	// whole arrays (SUBGRID needs every equilibrium population before the collision)
	float Fin[Q];
	if(vf) calc_forcing<Q>(uxn, uyn, uzn, fxn, fyn, fzn, Fin);
	else {
#pragma unroll
		for(int i = 0; i < Q; i++) Fin[i] = 0.0f;
	}
	float feq[Q];
	calc_feq<Q>(rhon, rhom1, uxn, uyn, uzn, feq);
	float w = a.w;
	if(a.ext & EXT_SUBGRID) w = subgrid_rate<Q>(w, fhn, feq, rhon);
	if constexpr(!TRT) { // src/kernel.cpp:1595-1604
		if(vf) {
			const float c_tau = __builtin_fmaf(w, -0.5f, 1.0f);
#pragma unroll
			for(int i = 0; i < Q; i++) Fin[i] *= c_tau;
		}
		const float omw = 1.0f - w;
#pragma unroll
		for(int i = 0; i < Q; i++) {
			const float c = __builtin_fmaf(omw, fhn[i], __builtin_fmaf(w, feq[i], Fin[i]));
			fhn[i] = isE ? feq[i] : c;
		}
	} else { // src/kernel.cpp:1605-1633, Lambda = 3/16
		const float wp = w;
		const float wm = (a.ext & EXT_SUBGRID) ? trt_minus_rate(w) : a.wm;
		if(vf) {
			const float c_taup = __builtin_fmaf(wp, -0.25f, 0.5f), c_taum = __builtin_fmaf(wm, -0.25f, 0.5f);
			float Fn[Q];
			Fn[0] = __builtin_fmaf(c_taup, Fin[0] + Fin[0], c_taum * (Fin[0] - Fin[0]));
#pragma unroll
			for(int i = 1; i < Q; i += 2) {
				Fn[i] = __builtin_fmaf(c_taup, Fin[i] + Fin[i + 1], c_taum * (Fin[i] - Fin[i + 1]));
				Fn[i + 1] = __builtin_fmaf(c_taup, Fin[i + 1] + Fin[i], c_taum * (Fin[i + 1] - Fin[i]));
			}
#pragma unroll
			for(int i = 0; i < Q; i++) Fin[i] = Fn[i];
		}
		const float hwp = 0.5f * wp, hwm = 0.5f * wm;
		// Without VOLUME_FORCE the forcing terms are zero; the reference still adds them
		// (fhn[i] + 0.0f), which changes only the sign of an exact zero, so that addition
		// is left out (one VALU operation per population)
		constexpr bool forced = (EXT & (EXT_VOLUME_FORCE | EXT_POROSITY)) != 0u;
		float out[Q];
#pragma unroll
		for(int i = 0; i < Q; i++) {
			const int b = i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); // opposite slot
			// feq[i]-feq[b] of an even slot as the negated difference of its odd partner (the
			// same value, but -0 for +0: only the sign of an exact zero in the result)
			const float dm = !SHARE || (i & 1) || i == 0 ? feq[i] - feq[b] : -(feq[b] - feq[i]);
			const float c = __builtin_fmaf(hwp, ((feq[i] - fhn[i]) + feq[b]) - fhn[b],
			                               __builtin_fmaf(hwm, (dm - fhn[i]) + fhn[b], forced ? fhn[i] + Fin[i] : fhn[i]));
			out[i] = isE ? feq[i] : c;
		}
#pragma unroll
		for(int i = 0; i < Q; i++) fhn[i] = out[i];
	}
}

// Block -> (x chunk, y, z).  3-D grid: the block indices.  1-D grid (xchunks > 0): the
// hardware hands consecutive blocks to the 8 XCDs in turn, each with its own L2; the
// mapping gives every chunk of a row the same XCD (block L: group g = L/(8*xchunks) of 8
// rows, chunk = (L mod 8*xchunks)/8, row = 8g + L mod 8), so the cache lines that the
// +-1 x offsets share between neighbouring chunks, stores included, meet in one L2 instead
// of being written back partially by two.  Returns false for the padding blocks of the
// last group of 8 rows.
template<bool XCD> __device__ __forceinline__ bool block_coords(const LbmArgs& a, uint32_t& chunk, uint32_t& y, uint32_t& z) {
	if(!XCD || a.xchunks == 0u) {
		chunk = blockIdx.x; y = blockIdx.y; z = blockIdx.z;
		if(a.reverse) { y = gridDim.y - 1u - y; z = gridDim.z - 1u - z; }
		return true;
	}
	// the grid is whole groups of 8 rows: reversing keeps every row's chunks on one XCD
	const uint32_t L = a.reverse ? gridDim.x - 1u - blockIdx.x : blockIdx.x, span = 8u * a.xchunks;
	const uint32_t g = L / span, rem = L - g * span;
	chunk = rem >> 3;
	const uint32_t row = g * 8u + (rem & 7u);
	if(row >= a.Ny * a.Nz) return false;
	z = row / a.Ny;
	y = row - z * a.Ny;
	return true;
}
// Segment (block of consecutive cells) of block L on the cell-linear grid.  The hardware
// hands consecutive blocks to the 8 XCDs in turn; with xcd and a grid of whole groups of
// 64 blocks, group g's block 8j + k takes segment 64g + 8k + j, so each XCD runs 8
// consecutive segments and the cache lines that neighbouring segments share at the shifted
// slots mostly stay in one L2.  A permutation of each group; the identity otherwise.
__host__ __device__ inline uint64_t linear_segment(uint64_t L, uint64_t blocks, bool xcd) {
	if(!xcd || (blocks & 63u) != 0u) return L;
	return (L & ~(uint64_t)63u) | ((L & 7u) << 3) | ((L >> 3) & 7u);
}
// Cell of this lane.  LIN: the cell-linear grid, block L covering segment
// linear_segment(L) of blockDim.x consecutive cells of the whole lattice, rows and planes
// crossed wherever they end (every wave starts on a multiple of 64 cells, so its unshifted
// slot accesses are whole 128-byte lines whatever Nx is); XCD there selects the XCD-aware
// segment order.  Otherwise the row grid of block_coords.  False for lanes past the lattice.
template<bool XCD, bool LIN> __device__ __forceinline__ bool cell_coords(const LbmArgs& a, uint32_t& x, uint32_t& y, uint32_t& z) {
	if constexpr(LIN) {
		const uint32_t L = a.reverse ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
		const uint64_t c0 = linear_segment(L, gridDim.x, XCD) * blockDim.x; // first cell of the block
		if(c0 + threadIdx.x >= a.N) return false;
		const uint64_t r0 = c0 / a.Nx; // its row y + z*Ny
		const uint32_t z0 = (uint32_t)(r0 / a.Ny);
		x = (uint32_t)(c0 - r0 * a.Nx) + threadIdx.x;
		y = (uint32_t)(r0 - (uint64_t)z0 * a.Ny);
		z = z0;
		while(x >= a.Nx) { // rows crossed within the block (at most blockDim.x/Nx + 1)
			x -= a.Nx;
			if(++y == a.Ny) { y = 0u; z++; }
		}
		return true;
	} else {
		uint32_t chunk;
		if(!block_coords<XCD>(a, chunk, y, z)) return false;
		x = chunk * blockDim.x + threadIdx.x;
		return x < a.Nx;
	}
}
// One lane per cell.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT, bool RECORD, bool XCD, bool LIN, bool O32 = false, bool BUF = false>
__device__ __forceinline__ void stream_collide_cell(LbmArgs a, float* moments) {
	a.ext &= EXT;
	if constexpr(constant_vf_mask<TRT>(Q, ST, EXT, LIN)) a.ext = EXT_VOLUME_FORCE;
	uint32_t x, y, z;
	if(!cell_coords<XCD, LIN>(a, x, y, z)) return;
	constexpr bool U = !LIN && uniform_rows<TRT>(Q, ST, EXT); // rows are per lane on the cell-linear grid
	const Site s = make_site(x, y, z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	const uint32_t flagsn = a.flags[n];
	const uint32_t bo = flagsn & TYPE_BO;
	if(bo == TYPE_S || (flagsn & TYPE_SU) == TYPE_G) return; // solid cells are mailboxes
	typedef typename Storage<ST>::E E;
	E* fi = (E*)a.fi;
	float f[Q];
	if constexpr(ST == ST_FP16S) {
		// every load is issued before the first decode: left alone, the scheduler may
		// convert each 16-bit value right after its load and wait for it, one at a time
		E r[Q];
		if constexpr(load_priority<TRT>(EXT)) __builtin_amdgcn_s_setprio(1);
		load_raw<Q, ST, U, LIN, O32, BUF>(fi, a.N, s, a.odd, r);
		__builtin_amdgcn_sched_barrier(0);
		if constexpr(load_priority<TRT>(EXT)) __builtin_amdgcn_s_setprio(0);
#pragma unroll
		for(int i = 0; i < Q; i++) f[i] = Storage<ST>::decode(r[i]);
	} else {
		load_f<Q, ST, U, LIN, O32, BUF>(fi, a.N, s, a.odd, f);
	}
	if((a.ext & EXT_MOVING_BOUNDARIES) && bo == TYPE_MS) moving_boundaries<Q>(a, s, f);
	collide_cell<Q, TRT, EXT, RECORD, shared_division(ST, EXT), shared_pair_difference(Q, EXT, BUF)>(a, n, bo, f, moments);
	store_f<Q, ST, U, LIN, O32, BUF>(fi, a.N, s, a.odd, f);
}
// XCD: the kernel may run on the 1-D XCD-mapped grid (a.xchunks > 0); false compiles the
// mapping out.  LIN: the cell-linear grid (cell_coords), for lattices whose rows do not
// start on 128-byte lines.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT, bool XCD = false, bool LIN = false, bool O32 = false, bool BUF = false>
__global__ __launch_bounds__(BLOCK) void k_stream_collide(LbmArgs a) {
	stream_collide_cell<Q, ST, TRT, EXT, false, XCD, LIN, O32, BUF>(a, nullptr);
}
// stream_collide with two adjacent cells per lane, for the 16-bit formats on rows of whole
// waves of lanes (Nx a multiple of 128, pair_rows).  A wave takes the 128 cells c0 .. c0+127
// of one row, lane l the cells xa = c0 + 2l and xa+1.  Each unshifted slot and each slot
// shifted only in y/z is one 32-bit word per lane holding both cells' codes (low half: cell
// xa), so those accesses are 4 bytes per lane instead of 2.  The slots shifted in x are
// accessed as two contiguous runs of 64 codes, lane m taking the neighbour of cell c0+m and
// then of c0+64+m (periodic at the row ends), and the codes move between those lanes and the
// cells' lanes with ds_bpermute: each run access touches the two lines the 64 codes span,
// where a per-cell 16-bit access at a stride of two codes (xa's neighbour and xa+1's are not
// one aligned word) would span three.  The per-lane words are decoded, collided and
// re-encoded one cell at a time, so only one cell's floats are live.  A cell that does not
// collide (solid or gas) writes each of its codes back to where it was loaded from:
// Esoteric-Pull stores a population to the other location of its pair, so the codes of each
// pair are swapped first, and the cell's memory is left as it was (a wave with no live cell
// returns at once).  Same results as k_stream_collide, bit for bit.
//
// The one-cell 16-bit kernels keep the texture address unit 81 % (FP16S) and 91 % (FP16C)
// busy at about 10 cycles per wave access, 2 bytes per lane.  Here the 32-bit words cost
// about 13 cycles; with one 16-bit access per cell at the x-shifted slots (about 25 cycles
// each, two codes apart) the address work per cell stayed that of the one-cell kernel and
// the gain over it was small (D3Q19 SRT at 256^3: FP16S +2.5 %, FP16C +1.2 %, EQUILIBRIUM_
// BOUNDARIES +5 %; the scalar instructions per cell fell from 3.2 to 2.0, one flag load and
// one set of row offsets per two cells); the contiguous runs added +6 to +7 % (FP16S) and
// +6 % (FP16C) at 256^3 and +6 to +7 % (FP16S) at 1024x128x96, same process, bit-exact.
// Earlier forms that moved halves between lanes from aligned words needed single-lane
// accesses at the wave's edges, which cost about as much as full ones (10 % slower).
typedef __attribute__((address_space(1))) uint16_t gcode;
typedef __attribute__((address_space(1))) uint32_t gword;
// Wave-uniform base of a slot row, in scalar registers and in the global address space
// (readfirstlane returns an integer; without the address space the accesses would be
// generic flat ones).
__device__ __forceinline__ gcode* global_row(const uint16_t* p) {
	const uint64_t v = (uint64_t)p;
	const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
	return (gcode*)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ gcode* code_at(gcode* p, uint32_t x) { return (gcode*)((__attribute__((address_space(1))) char*)p + x * 2u); }
__device__ __forceinline__ gword* word_at(gcode* p, uint32_t x) { return (gword*)code_at(p, x); } // x even
// BUF: every access through buffer instructions (load_raw's BUF; a DDF array below 4 GiB):
// no 64-bit store addresses rebuilt in vector registers after the collision.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT, bool BUF = false>
__global__ __launch_bounds__(BLOCK) void k_stream_collide_pair(LbmArgs a) {
	static_assert(ST != ST_FP32 && !(EXT & EXT_MOVING_BOUNDARIES), "16-bit codes only; populations as decoded");
	a.ext &= EXT;
	typedef Storage<ST> S;
	uint32_t chunk, y, z;
	if(!block_coords<true>(a, chunk, y, z)) return;
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t c0 = 2u * (chunk * blockDim.x + (threadIdx.x & ~63u)); // the wave's first cell
	if(c0 >= a.Nx) return; // whole waves: Nx/2 lanes per row, a multiple of 64
	const uint32_t xa = c0 + 2u * lane;
	const Site s = make_site(xa, y, z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	const uint32_t fl = *(const uint16_t*)(a.flags + n); // n is even: flags[n] low, flags[n+1] high
	const uint32_t flA = fl & 0xFFu, flB = fl >> 8;
	const bool liveA = !((flA & TYPE_BO) == TYPE_S || (flA & TYPE_SU) == TYPE_G);
	const bool liveB = !((flB & TYPE_BO) == TYPE_S || (flB & TYPE_SU) == TYPE_G);
	if(__ballot(liveA || liveB) == 0ull) return; // wave-uniform: every lane takes part in the exchanges
	// run positions: lane m's cells c0+m and c0+64+m, their x-1 / x+1 neighbours (the row
	// wraps only at c0+m-1 for c0 = 0 and at c0+65+m for the row's last wave)
	const uint32_t r1 = c0 + lane, r2 = r1 + 64u;
	const uint32_t r1m = r1 == 0u ? a.Nx - 1u : r1 - 1u, r2p = r2 + 1u == a.Nx ? 0u : r2 + 1u;
	uint16_t* fi = (uint16_t*)a.fi;
	const uint64_t N = a.N, n0 = s.R[4];
	// a slot row's first element (uniform): its word at xa, its codes at q
	const auto word_load = [&](uint64_t row) -> uint32_t {
		if constexpr(BUF) return buffer_elem<uint32_t, true>(ddf_buffer(fi), xa * 2u, (uint32_t)row * 2u);
		else return __builtin_nontemporal_load(word_at(global_row(fi + row), xa));
	};
	const auto word_store = [&](uint64_t row, uint32_t w) {
		if constexpr(BUF) buffer_elem<uint32_t, true>(ddf_buffer(fi), xa * 2u, (uint32_t)row * 2u, w);
		else __builtin_nontemporal_store(w, word_at(global_row(fi + row), xa));
	};
	const auto code_load = [&](uint64_t row, uint32_t q) -> uint32_t {
		if constexpr(BUF) return buffer_elem<uint16_t, false>(ddf_buffer(fi), q * 2u, (uint32_t)row * 2u);
		else return *code_at(global_row(fi + row), q);
	};
	const auto code_store = [&](uint64_t row, uint32_t q, uint16_t h) {
		if constexpr(BUF) buffer_elem<uint16_t, false>(ddf_buffer(fi), q * 2u, (uint32_t)row * 2u, h);
		else *code_at(global_row(fi + row), q) = h;
	};
	uint32_t P[Q]; // population i of both cells, as loaded (load_raw's order): low half cell xa
	uint32_t X[Q]; // x-shifted slots: the codes of the runs, low half cell c0+m's, high c0+64+m's
	__builtin_amdgcn_s_setprio(1);
	P[0] = word_load(n0);
#pragma unroll
	for(int i = 1; i < Q; i += 2) {
		const uint64_t sa = a.odd ? (uint64_t)i : (uint64_t)(i + 1);
		const uint64_t sb = a.odd ? (uint64_t)(i + 1) : (uint64_t)i;
		P[i] = word_load(sa * N + n0);
		const uint64_t rb = sb * N + s.row(VS<Q>::cy[i], VS<Q>::cz[i]);
		if(VS<Q>::cx[i] == 0) P[i + 1] = word_load(rb);
		else {
			const uint32_t q1 = VS<Q>::cx[i] > 0 ? r1 + 1u : r1m, q2 = VS<Q>::cx[i] > 0 ? r2p : r2 - 1u;
			X[i + 1] = code_load(rb, q1) | (code_load(rb, q2) << 16);
		}
	}
	__builtin_amdgcn_sched_barrier(0);
	__builtin_amdgcn_s_setprio(0);
	{ // runs -> cells: lane l's cells 2l, 2l+1 sit in lanes 2l, 2l+1 (mod 64), low halves for l < 32
		const int ia = (int)(((2u * lane) & 63u) * 4u), ib = (int)(((2u * lane + 1u) & 63u) * 4u);
		const uint32_t sel = lane < 32u ? 0x05040100u : 0x07060302u;
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			if(VS<Q>::cx[i] == 0) continue;
			const uint32_t ca = (uint32_t)__builtin_amdgcn_ds_bpermute(ia, (int)X[i + 1]);
			const uint32_t cb = (uint32_t)__builtin_amdgcn_ds_bpermute(ib, (int)X[i + 1]);
			P[i + 1] = __builtin_amdgcn_perm(cb, ca, sel); // (cb half << 16) | ca half
		}
	}
	float f[Q];
	if(liveA) {
#pragma unroll
		for(int i = 0; i < Q; i++) {
			if constexpr(ST == ST_FP16C) f[i] = S::decode_low(P[i]);
			else f[i] = S::decode((uint16_t)P[i]);
		}
		collide_cell<Q, TRT, EXT, false, shared_division(ST, EXT), shared_pair_difference(Q, EXT, BUF)>(a, n, flA & TYPE_BO, f, nullptr);
#pragma unroll
		for(int i = 0; i < Q; i++) P[i] = (P[i] & 0xFFFF0000u) | S::encode_word(f[i]);
	} else {
#pragma unroll
		for(int i = 1; i < Q; i += 2) { const uint32_t t = (P[i] ^ P[i + 1]) & 0x0000FFFFu; P[i] ^= t; P[i + 1] ^= t; }
	}
	__builtin_amdgcn_sched_barrier(0); // one cell's floats at a time
	if(liveB) {
#pragma unroll
		for(int i = 0; i < Q; i++) f[i] = S::decode((uint16_t)(P[i] >> 16));
		collide_cell<Q, TRT, EXT, false, shared_division(ST, EXT), shared_pair_difference(Q, EXT, BUF)>(a, n + 1u, flB & TYPE_BO, f, nullptr);
#pragma unroll
		for(int i = 0; i < Q; i++) P[i] = S::encode_high(f[i], P[i]); // (code << 16) | (P & 0xFFFF)
	} else {
#pragma unroll
		for(int i = 1; i < Q; i += 2) { const uint32_t t = (P[i] ^ P[i + 1]) & 0xFFFF0000u; P[i] ^= t; P[i + 1] ^= t; }
	}
	{ // cells -> runs: cell c0+m sits in lane m/2, cell c0+64+m in lane 32+m/2, high halves for odd m
		const int ja = (int)((lane >> 1) * 4u), jb = (int)((32u + (lane >> 1)) * 4u);
		const uint32_t sel = (lane & 1u) ? 0x07060302u : 0x05040100u;
#pragma unroll
		for(int i = 1; i < Q; i += 2) {
			if(VS<Q>::cx[i] == 0) continue;
			const uint32_t ca = (uint32_t)__builtin_amdgcn_ds_bpermute(ja, (int)P[i]);
			const uint32_t cb = (uint32_t)__builtin_amdgcn_ds_bpermute(jb, (int)P[i]);
			X[i] = __builtin_amdgcn_perm(cb, ca, sel);
		}
	}
	// store_f's order: population i to the shifted location, i+1 to the cell's own slot
	word_store(n0, P[0]);
#pragma unroll
	for(int i = 1; i < Q; i += 2) {
		const uint64_t sa = a.odd ? (uint64_t)i : (uint64_t)(i + 1);
		const uint64_t sb = a.odd ? (uint64_t)(i + 1) : (uint64_t)i;
		const uint64_t rb = sb * N + s.row(VS<Q>::cy[i], VS<Q>::cz[i]);
		if(VS<Q>::cx[i] == 0) word_store(rb, P[i]);
		else {
			const uint32_t q1 = VS<Q>::cx[i] > 0 ? r1 + 1u : r1m, q2 = VS<Q>::cx[i] > 0 ? r2p : r2 - 1u;
			code_store(rb, q1, (uint16_t)X[i]);
			code_store(rb, q2, (uint16_t)(X[i] >> 16));
		}
		word_store(sa * N + n0, P[i + 1]);
	}
}

// stream_collide that also writes the pre-collision (rho, jx, jy, jz) of every fluid cell to
// moments[c*N+n], c = 0..3 (16 B/cell): what the adjoint step needs of the forward state.
// EXT is the context's extension subset; MOVING_BOUNDARIES and SUBGRID are not
// differentiated and compiled out.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT, bool XCD = false, bool LIN = false>
__global__ __launch_bounds__(BLOCK) void k_stream_collide_record(LbmArgs a, float* moments) {
	stream_collide_cell<Q, ST, TRT, EXT & ~(EXT_MOVING_BOUNDARIES | EXT_SUBGRID), true, XCD, LIN>(a, moments);
}

// update_fields (src/kernel.cpp:1794-1870)
template<int Q, uint32_t ST, uint32_t EXT>
__global__ __launch_bounds__(BLOCK) void k_update_fields(LbmArgs a) {
	a.ext &= EXT;
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if(x >= a.Nx) return;
	const Site s = make_site(x, blockIdx.y, blockIdx.z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	const uint32_t flagsn = a.flags[n];
	const uint32_t bo = flagsn & TYPE_BO;
	if(bo == TYPE_S || (flagsn & TYPE_SU) == TYPE_G) return;
	float fhn[Q];
	load_f<Q, ST>((const typename Storage<ST>::E*)a.fi, a.N, s, a.odd, fhn);
	if((a.ext & EXT_MOVING_BOUNDARIES) && bo == TYPE_MS) moving_boundaries<Q>(a, s, fhn);
	float rhon, uxn, uyn, uzn;
	calc_rho_u<Q>(fhn, rhon, uxn, uyn, uzn);
	float fxn = a.fx, fyn = a.fy, fzn = a.fz;
	if(a.ext & EXT_FORCE_FIELD) { fxn += a.F[n]; fyn += a.F[a.N + n]; fzn += a.F[2ull * a.N + n]; }
	const float alphan = (a.ext & EXT_POROSITY) ? a.alpha[n] : 0.0f;
	force_block(a.ext, rhon, uxn, uyn, uzn, fxn, fyn, fzn, alphan);
	if(!((a.ext & EXT_EQUILIBRIUM_BOUNDARIES) && bo == TYPE_E)) {
		a.rho[n] = rhon; a.u[n] = uxn; a.u[a.N + n] = uyn; a.u[2ull * a.N + n] = uzn;
	}
}

// update_force_field (src/kernel.cpp:1873-1884): momentum exchange on solid cells
template<int Q, uint32_t ST>
__global__ __launch_bounds__(BLOCK) void k_update_force_field(LbmArgs a) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	if(x >= a.Nx) return;
	const Site s = make_site(x, blockIdx.y, blockIdx.z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	if((a.flags[n] & TYPE_BO) != TYPE_S) return;
	float fhn[Q];
	load_f<Q, ST>((const typename Storage<ST>::E*)a.fi, a.N, s, a.odd, fhn);
	float Fb, fxv, fyv, fzv;
	calc_rho_u<Q>(fhn, Fb, fxv, fyv, fzv);
	const float sc = 2.0f * Fb;
	float* F = (float*)a.F;
	F[n] = sc * fxv; F[a.N + n] = sc * fyv; F[2ull * a.N + n] = sc * fzv;
}

// Total mass, per-row partial sums (FP64): every cell, solid or not, contributes 1 plus
// the shifted DDFs stored in its own slots fi[i*N+n].  That sum is invariant under
// stream_collide (values only move inside disjoint footprints, collisions conserve
// sum_i f_i); one partial per block, summed later in a fixed order.
template<int Q, uint32_t ST>
__global__ __launch_bounds__(BLOCK) void k_mass_partial(LbmArgs a, double* partial) {
	__shared__ double red[BLOCK];
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
	double v = 0.0;
	if(x < a.Nx) {
		const uint64_t n = ((uint64_t)blockIdx.z * a.Ny + blockIdx.y) * a.Nx + x;
		const typename Storage<ST>::E* fi = (const typename Storage<ST>::E*)a.fi;
		v = 1.0;
#pragma unroll
		for(int i = 0; i < Q; i++) v += (double)Storage<ST>::load(fi + (uint64_t)i * a.N + n, 0u);
	}
	red[threadIdx.x] = v;
	__syncthreads();
	for(uint32_t k = blockDim.x / 2u; k > 0u; k >>= 1) { // blockDim.x is 64, 128 or 256
		if(threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
		__syncthreads();
	}
	if(threadIdx.x == 0) partial[((uint64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// ------------------------------------------------------------------------------------
// Halo transfer between domains of a split lattice (the role of src/kernel.cpp:2049-2124,
// transfer_extract_fi / transfer__insert_fi, and 2126-2183 for rho/u/flags, flags, F).  A
// domain split along an axis has one halo layer on each split side; after every
// stream_collide, the DDF values that cross each face move to the neighbour:
//  - extract side +: interior layer N-2, the directions with c_axis(i) = +1;
//  - extract side -: interior layer 1, the directions with c_axis(i) = -1;
//  - insert side +: halo layer N-1, the directions with c_axis(i) = -1 (what the + neighbour
//    extracted on its side -);
//  - insert side -: halo layer 0, the directions with c_axis(i) = +1.
// For cell n and direction i the value is the population f_i that stream_collide of step t
// stored: at (n + e_i, slot t odd ? i+1 : i) for odd i, at (n, slot t odd ? i-1 : i) for
// even i (store_f).  Extract reads it relative to the interior cell, insert writes the same
// physical location relative to the receiving domain's halo cell.  The values are copied as
// stored (no decode).  Buffer layout b*A + a: a = face cell (x faces: y + z*Ny, y faces:
// x + z*Nx, z faces: x + y*Nx), b = position of i in the increasing direction list.
// ------------------------------------------------------------------------------------
template<int Q> struct TransferList {
	static constexpr int T = Q == 19 ? 5 : 9; // directions crossing one face (src/lbm.cpp:18,22)
	int8_t i[3][2][9];                        // [axis][side: 0 = +, 1 = -][b]
	constexpr TransferList() : i{} {
		for(int ax = 0; ax < 3; ax++)
			for(int sd = 0; sd < 2; sd++) {
				int b = 0;
				for(int k = 1; k < Q; k++) {
					const int c = ax == 0 ? VS<Q>::cx[k] : ax == 1 ? VS<Q>::cy[k] : VS<Q>::cz[k];
					if(c == (sd == 0 ? 1 : -1)) i[ax][sd][b++] = (int8_t)k;
				}
			}
	}
};
__device__ __forceinline__ uint32_t wrap(int v, uint32_t n) { return v < 0 ? n - 1u : (uint32_t)v >= n ? 0u : (uint32_t)v; }
// Face cell a of the layer `layer` normal to AXIS -> (x, y, z)
template<int AXIS> __device__ __forceinline__ void face_cell(uint32_t a, uint32_t layer, uint32_t Nx, uint32_t Ny, uint32_t& x, uint32_t& y, uint32_t& z) {
	if constexpr(AXIS == 0) { x = layer; y = a % Ny; z = a / Ny; }
	else if constexpr(AXIS == 1) { x = a % Nx; y = layer; z = a / Nx; }
	else { x = a % Nx; y = a / Nx; z = layer; }
}
// DDF location of direction i stored by cell (x,y,z) at parity odd
template<int Q> __device__ __forceinline__ uint64_t transfer_location(int i, uint32_t x, uint32_t y, uint32_t z, uint32_t Nx, uint32_t Ny, uint32_t Nz, uint64_t N, uint32_t odd) {
	uint64_t slot;
	if(i & 1) {
		x = wrap((int)x + VS<Q>::cx[i], Nx); y = wrap((int)y + VS<Q>::cy[i], Ny); z = wrap((int)z + VS<Q>::cz[i], Nz);
		slot = odd ? (uint64_t)(i + 1) : (uint64_t)i;
	} else {
		slot = odd ? (uint64_t)(i - 1) : (uint64_t)i;
	}
	return slot * N + ((uint64_t)z * Ny + y) * Nx + x;
}
template<int Q, typename E, int AXIS, bool INSERT>
__global__ __launch_bounds__(BLOCK) void k_transfer_fi(E* fi, E* buffer_p, E* buffer_m, uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t odd) {
	constexpr TransferList<Q> L{};
	constexpr int T = TransferList<Q>::T;
	const uint32_t Nax = AXIS == 0 ? Nx : AXIS == 1 ? Ny : Nz;
	const uint64_t A = AXIS == 0 ? (uint64_t)Ny * Nz : AXIS == 1 ? (uint64_t)Nx * Nz : (uint64_t)Nx * Ny;
	const uint64_t a = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if(a >= A) return;
	const uint64_t N = (uint64_t)Nx * Ny * Nz;
#pragma unroll
	for(int sd = 0; sd < 2; sd++) { // side +, side -
		E* buffer = sd == 0 ? buffer_p : buffer_m;
		const uint32_t layer = INSERT ? (sd == 0 ? Nax - 1u : 0u) : (sd == 0 ? Nax - 2u : 1u);
		uint32_t x, y, z;
		face_cell<AXIS>((uint32_t)a, layer, Nx, Ny, x, y, z);
#pragma unroll
		for(int b = 0; b < T; b++) {
			const int i = L.i[AXIS][INSERT ? 1 - sd : sd][b];
			const uint64_t m = transfer_location<Q>(i, x, y, z, Nx, Ny, Nz, N, odd);
			if constexpr(INSERT) fi[m] = buffer[(uint64_t)b * A + a];
			else buffer[(uint64_t)b * A + a] = fi[m];
		}
	}
}

// ------------------------------------------------------------------------------------
// Adjoint (reverse) step.  The forward step maps the populations v loaded from the cell's
// footprint locations L_k to v' = C(v) stored at S_m = L_opposite(m).  Its transpose reads
// mu_m = lambda[S_m], forms nu = J_C^T mu (plus the objective weight on the outlet plane)
// and writes lambda[L_k] = nu_k: the same footprint walk with loads and stores exchanged,
// in place and race-free.  Solid cells are the identity.  J_C is the identity part of the
// collision plus a correction through (rho, j) only, so it is evaluated from the
// recorded forward moments with weighted moment sums of mu:
//   SRT: v'_i = (1-w) v_i + w feq_i(rho,u) + (1-w/2) Fin_i(u,F)
//   TRT: v'_i = v_i + (w+/2)(feq_i+feq_-i - v_i - v_-i) + (w-/2)(feq_i-feq_-i - v_i + v_-i) + Fin'_i
// with u = clamp(k (j/rho + f/(2 rho))), k = 1/(1+alpha/2), F = f - alpha rho u (porosity).
// The porosity derivative sum_m mu_m dv'_m/dalpha is accumulated into grad[n] (FP32).
// ------------------------------------------------------------------------------------
struct AdjointArgs {
	LbmArgs a;               // a.fi = lambda (FP32, q*N), a.rho/a.u = TYPE_E presets
	const float* moments;    // (rho, jx, jy, jz) of step t, 4*N
	float* grad;             // d objective / d porosity, N
	uint32_t plane_x;        // objective plane
	float obj_weight;        // objective weight of step t (0 outside the averaging window)
};

template<int Q, bool TRT, bool XCD = false, bool LIN = false>
__global__ __launch_bounds__(BLOCK) void k_adjoint_step(AdjointArgs g) {
	const LbmArgs& a = g.a;
	uint32_t x, y, z;
	if(!cell_coords<XCD, LIN>(a, x, y, z)) return;
	const Site s = make_site(x, y, z, a.Nx, a.Ny, a.Nz);
	const uint64_t n = s.n();
	const uint32_t flagsn = a.flags[n];
	const uint32_t bo = flagsn & TYPE_BO;
	if(bo == TYPE_S || (flagsn & TYPE_SU) == TYPE_G) return;
	float* lambda = (float*)a.fi;
	float lam[Q], mu[Q];
	load_f<Q, ST_FP32, false, LIN>(lambda, a.N, s, a.odd, lam); // lam[k] = lambda[L_k]
	mu[0] = lam[0];
#pragma unroll
	for(int i = 1; i < Q; i += 2) { mu[i] = lam[i + 1]; mu[i + 1] = lam[i]; } // mu[m] = lambda[S_m]

	const uint32_t ext = a.ext;
	const bool isE = (ext & EXT_EQUILIBRIUM_BOUNDARIES) && bo == TYPE_E;
	const bool vf = ext & (EXT_VOLUME_FORCE | EXT_POROSITY);
	const bool por = ext & EXT_POROSITY;
	float fxn = a.fx, fyn = a.fy, fzn = a.fz;
	if(ext & EXT_FORCE_FIELD) { fxn += a.F[n]; fyn += a.F[a.N + n]; fzn += a.F[2ull * a.N + n]; }
	const float alphan = por ? a.alpha[n] : 0.0f;
	// forward quantities of step t, recomputed with the forward kernel's arithmetic
	float rho, u0x, u0y, u0z;
	if(isE) { rho = a.rho[n]; u0x = a.u[n]; u0y = a.u[a.N + n]; u0z = a.u[2ull * a.N + n]; }
	else {
		rho = g.moments[n];
		u0x = g.moments[a.N + n] / rho; u0y = g.moments[2ull * a.N + n] / rho; u0z = g.moments[3ull * a.N + n] / rho;
	}
	float ubx = u0x, uby = u0y, ubz = u0z, k = 1.0f;
	if(vf) {
		const float rho2 = 0.5f / rho;
		ubx = __builtin_fmaf(fxn, rho2, u0x); uby = __builtin_fmaf(fyn, rho2, u0y); ubz = __builtin_fmaf(fzn, rho2, u0z);
		if(por) { k = 1.0f / (1.0f + 0.5f * alphan); ubx *= k; uby *= k; ubz *= k; }
	}
	const float ux = clampc(ubx), uy = clampc(uby), uz = clampc(ubz);
	const float mx = (ubx > -DEF_C && ubx < DEF_C) ? 1.0f : 0.0f;
	const float my = (uby > -DEF_C && uby < DEF_C) ? 1.0f : 0.0f;
	const float mz = (ubz > -DEF_C && ubz < DEF_C) ? 1.0f : 0.0f;
	float Fx = fxn, Fy = fyn, Fz = fzn;
	if(por) { const float ar = -(alphan * rho); Fx = __builtin_fmaf(ar, ux, fxn); Fy = __builtin_fmaf(ar, uy, fyn); Fz = __builtin_fmaf(ar, uz, fzn); }

	// weights of feq and Fin in sum_m mu_m v'_m
	const float w = a.w;
	float hp = 0.0f, hm = 0.0f, ctp = 0.0f, ctm = 0.0f;
	if constexpr(TRT) {
		const float wm = a.wm;
		hp = 0.5f * w; hm = 0.5f * wm;
		ctp = __builtin_fmaf(w, -0.25f, 0.5f); ctm = __builtin_fmaf(wm, -0.25f, 0.5f);
	}
	const float ctau = __builtin_fmaf(w, -0.5f, 1.0f);
	// weighted moments A (equilibrium weights) and B (forcing weights)
	float A0 = 0.0f, Ax = 0.0f, Ay = 0.0f, Az = 0.0f, Axx = 0.0f, Ayy = 0.0f, Azz = 0.0f, Axy = 0.0f, Axz = 0.0f, Ayz = 0.0f;
	float B0 = 0.0f, Bx = 0.0f, By = 0.0f, Bz = 0.0f, Bxx = 0.0f, Byy = 0.0f, Bzz = 0.0f, Bxy = 0.0f, Bxz = 0.0f, Byz = 0.0f;
#pragma unroll
	for(int i = 0; i < Q; i++) {
		const int b = i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1);
		float leq, lF;
		if(isE) { leq = mu[i]; lF = 0.0f; }
		else if constexpr(TRT) { leq = hp * (mu[i] + mu[b]) + hm * (mu[i] - mu[b]); lF = ctp * (mu[i] + mu[b]) + ctm * (mu[i] - mu[b]); }
		else { leq = w * mu[i]; lF = ctau * mu[i]; }
		const float we = weight<Q>(i) * leq, wf = weight<Q>(i) * lF;
		const int cx = VS<Q>::cx[i], cy = VS<Q>::cy[i], cz = VS<Q>::cz[i];
		A0 += we; B0 += wf;
		if(cx) { Ax += cx * we; Bx += cx * wf; Axx += we; Bxx += wf; }
		if(cy) { Ay += cy * we; By += cy * wf; Ayy += we; Byy += wf; }
		if(cz) { Az += cz * we; Bz += cz * wf; Azz += we; Bzz += wf; }
		if(cx && cy) { Axy += cx * cy * we; Bxy += cx * cy * wf; }
		if(cx && cz) { Axz += cx * cz * we; Bxz += cx * cz * wf; }
		if(cy && cz) { Ayz += cy * cz * we; Byz += cy * cz * wf; }
	}
	// d/du and d/drho of sum leq_i feq_i(rho,u) = rho (3 A1.u + 4.5 u.A2.u - 1.5 A0 u.u) + A0 (rho-1)
	const float A2ux = Axx * ux + Axy * uy + Axz * uz, A2uy = Axy * ux + Ayy * uy + Ayz * uz, A2uz = Axz * ux + Ayz * uy + Azz * uz;
	const float usq = ux * ux + uy * uy + uz * uz;
	float gux = rho * (3.0f * Ax + 9.0f * A2ux - 3.0f * A0 * ux);
	float guy = rho * (3.0f * Ay + 9.0f * A2uy - 3.0f * A0 * uy);
	float guz = rho * (3.0f * Az + 9.0f * A2uz - 3.0f * A0 * uz);
	float grho = 3.0f * (Ax * ux + Ay * uy + Az * uz) + 4.5f * (ux * A2ux + uy * A2uy + uz * A2uz) - 1.5f * A0 * usq + A0;
	float galpha = 0.0f;
	if(vf && !isE) { // sum lF_i Fin_i = 9 (F.B2.u + B1.F/3 - B0 u.F/3)
		const float B2ux = Bxx * ux + Bxy * uy + Bxz * uz, B2uy = Bxy * ux + Byy * uy + Byz * uz, B2uz = Bxz * ux + Byz * uy + Bzz * uz;
		const float B2Fx = Bxx * Fx + Bxy * Fy + Bxz * Fz, B2Fy = Bxy * Fx + Byy * Fy + Byz * Fz, B2Fz = Bxz * Fx + Byz * Fy + Bzz * Fz;
		const float dFx = 9.0f * B2ux + 3.0f * Bx - 3.0f * B0 * ux; // d/dF at fixed u
		const float dFy = 9.0f * B2uy + 3.0f * By - 3.0f * B0 * uy;
		const float dFz = 9.0f * B2uz + 3.0f * Bz - 3.0f * B0 * uz;
		gux += 9.0f * B2Fx - 3.0f * B0 * Fx; // d/du at fixed F
		guy += 9.0f * B2Fy - 3.0f * B0 * Fy;
		guz += 9.0f * B2Fz - 3.0f * B0 * Fz;
		if(por) { // F = f - alpha rho u
			const float ar = alphan * rho;
			gux -= ar * dFx; guy -= ar * dFy; guz -= ar * dFz;
			const float udF = ux * dFx + uy * dFy + uz * dFz;
			grho -= alphan * udF;
			galpha -= rho * udF;
		}
	}
	// through u = clamp(ub), ub = k (j/rho + f/(2 rho)): dub/dj = k/rho, dub/drho = -ub/rho, dub/dalpha = -k ub/2
	const float gbx = mx * gux, gby = my * guy, gbz = mz * guz;
	const float gbub = gbx * ubx + gby * uby + gbz * ubz;
	if(por) galpha -= 0.5f * k * gbub;
	float nu[Q];
	if(isE) {
#pragma unroll
		for(int i = 0; i < Q; i++) nu[i] = 0.0f; // v' = feq(presets): no dependence on v
	} else {
		const float kr = k / rho;
		const float djx = gbx * kr, djy = gby * kr, djz = gbz * kr;
		const float drho = grho - gbub / rho;
		const float lin_a = TRT ? 1.0f - hp - hm : 1.0f - w, lin_b = TRT ? hm - hp : 0.0f;
#pragma unroll
		for(int i = 0; i < Q; i++) {
			const int b = i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1);
			float v = lin_a * mu[i] + drho;
			if constexpr(TRT) v += lin_b * mu[b];
			if(VS<Q>::cx[i]) v += VS<Q>::cx[i] * djx;
			if(VS<Q>::cy[i]) v += VS<Q>::cy[i] * djy;
			if(VS<Q>::cz[i]) v += VS<Q>::cz[i] * djz;
			nu[i] = v;
		}
	}
	if(x == g.plane_x && g.obj_weight != 0.0f) {
#pragma unroll
		for(int i = 1; i < Q; i++) if(VS<Q>::cx[i]) nu[i] += VS<Q>::cx[i] * g.obj_weight;
	}
	if(por) g.grad[n] += galpha;
	float out[Q]; // lambda[L_k] = nu_k through store_f, which writes out[i] to S_i = L_opposite(i)
	out[0] = nu[0];
#pragma unroll
	for(int i = 1; i < Q; i += 2) { out[i] = nu[i + 1]; out[i + 1] = nu[i]; }
	store_f<Q, ST_FP32, false, LIN>(lambda, a.N, s, a.odd, out);
}

// Objective of the adjoint: x-momentum sum_i c_ix v_i of the populations loaded at step t,
// per non-solid cell of the plane x = plane_x, summed in FP64 per block (one thread per
// (y,z) cell of the plane; grid.x*256 >= Ny*Nz).
template<int Q, uint32_t ST>
__global__ __launch_bounds__(256) void k_objective_partial(LbmArgs a, uint32_t plane_x, double* partial) {
	__shared__ double red[256];
	const uint64_t m = (uint64_t)blockIdx.x * 256u + threadIdx.x;
	double v = 0.0;
	if(m < (uint64_t)a.Ny * a.Nz) {
		const uint32_t y = (uint32_t)(m % a.Ny), z = (uint32_t)(m / a.Ny);
		const Site s = make_site(plane_x, y, z, a.Nx, a.Ny, a.Nz);
		if((a.flags[s.n()] & TYPE_BO) != TYPE_S) {
			float f[Q];
			load_f<Q, ST>((const typename Storage<ST>::E*)a.fi, a.N, s, a.odd, f);
			float r, jx, jy, jz;
			calc_rho_j<Q>(f, r, jx, jy, jz);
			v = (double)jx;
		}
	}
	red[threadIdx.x] = v;
	__syncthreads();
	for(uint32_t k = 128u; k > 0u; k >>= 1) {
		if(threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
		__syncthreads();
	}
	if(threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

} // namespace fx
