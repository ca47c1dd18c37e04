#pragma once
// Host utilities used by LBM scenes: integer typedefs, small vector types, printing,
// parallel_for, Clock and the C99 LCG.  The names follow the reference's
// src/utilities.hpp so scene code compiles unchanged; the implementations are this
// build's own and cover what the LBM API and its scenes use.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>
#include <thread>
#include <chrono>
#include <functional>
#include <algorithm>
#include <iostream>
#include <sstream>
#include <iomanip>
#include <type_traits>

typedef unsigned char uchar;
typedef unsigned short ushort;
typedef unsigned int uint;
typedef int64_t slong;
typedef uint64_t ulong;
using std::string;
using std::vector;
using std::thread;

#define pif 3.1415927f
#define pi 3.141592653589793
#define max_uint 4294967295u
#define max_ulong 18446744073709551615ull
#define max_double 1.7976931348623158E308
#define max_float 3.402823466E38f

template<typename T> inline T sq(const T x) { return x * x; }
template<typename T> inline T cb(const T x) { return x * x * x; }
template<typename T> inline T clamp(const T x, const T a, const T b) { return std::min(std::max(x, a), b); }
using std::min;
using std::max;
inline uint to_uint(const float x) { return (uint)std::max(x + 0.5f, 0.5f); }
inline uint to_uint(const double x) { return (uint)std::max(x + 0.5, 0.5); }
inline int to_int(const float x) { return (int)(x + 0.5f - (float)(x < 0.0f)); }
inline float as_float(const uint x) { float f; std::memcpy(&f, &x, 4); return f; }
inline uint as_uint(const float x) { uint u; std::memcpy(&u, &x, 4); return u; }

struct uint3 {
	uint x = 0u, y = 0u, z = 0u;
	uint3() = default;
	uint3(uint a, uint b, uint c) : x(a), y(b), z(c) {}
	explicit uint3(uint a) : x(a), y(a), z(a) {}
	uint3 operator*(uint s) const { return uint3(x * s, y * s, z * s); }
	uint3 operator/(uint s) const { return uint3(x / s, y / s, z / s); }
};
struct float3 {
	float x = 0.0f, y = 0.0f, z = 0.0f;
	float3() = default;
	float3(float a, float b, float c) : x(a), y(b), z(c) {}
	explicit float3(float a) : x(a), y(a), z(a) {}
	float3(uint a, uint b, uint c) : x((float)a), y((float)b), z((float)c) {}
	float3 operator+(const float3& v) const { return float3(x + v.x, y + v.y, z + v.z); }
	float3 operator-(const float3& v) const { return float3(x - v.x, y - v.y, z - v.z); }
	float3 operator-() const { return float3(-x, -y, -z); }
	float3 operator*(float s) const { return float3(x * s, y * s, z * s); }
	float3 operator/(float s) const { return float3(x / s, y / s, z / s); }
	float3 operator*(const float3& v) const { return float3(x * v.x, y * v.y, z * v.z); }
	float3& operator+=(const float3& v) { x += v.x; y += v.y; z += v.z; return *this; }
	float3& operator-=(const float3& v) { x -= v.x; y -= v.y; z -= v.z; return *this; }
	float3& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
	float3& operator/=(float s) { x /= s; y /= s; z /= s; return *this; }
};
inline float3 operator*(float s, const float3& v) { return v * s; }
inline float dot(const float3& a, const float3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float3 cross(const float3& a, const float3& b) { return float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(const float3& v) { return std::sqrt(sq(v.x) + sq(v.y) + sq(v.z)); }
inline float3 normalize(const float3& v) { const float l = length(v); return l > 0.0f ? v / l : v; }

inline string to_string(const string& s) { return s; }
inline string to_string(const char* s) { return string(s); }
inline string to_string(int x) { return std::to_string(x); }
inline string to_string(uint x) { return std::to_string(x); }
inline string to_string(slong x) { return std::to_string(x); }
inline string to_string(ulong x) { return std::to_string(x); }
inline string to_string(long long x) { return std::to_string(x); }
inline string to_string(unsigned long long x) { return std::to_string(x); }
inline string to_string(double x, uint decimals) { std::ostringstream o; o << std::fixed << std::setprecision((int)decimals) << x; return o.str(); }
inline string to_string(float x, uint decimals) { return to_string((double)x, decimals); }
inline string to_string(double x) { std::ostringstream o; o << std::setprecision(16) << x; return o.str(); }
inline string to_string(float x) { std::ostringstream o; o << std::setprecision(9) << x; return o.str(); }

inline void print(const string& s) { std::cout << s << std::flush; }
inline void println(const string& s = "") { std::cout << s << std::endl; }
template<typename T> inline void println(const T& x) { println(to_string(x)); }
inline void print_info(const string& s) { println("Info: " + s); }
inline void print_warning(const string& s) { println("Warning: " + s); }
[[noreturn]] inline void print_error(const string& s) { std::cerr << "Error: " << s << std::endl; std::exit(1); }
inline void wait() {} // the reference waits for a key press; this build runs headless
inline void sleep(const double seconds) { std::this_thread::sleep_for(std::chrono::duration<double>(seconds)); }
inline string get_exe_path() { return ""; }

// parallel_for over [0,N) on all hardware threads; the body gets n (and the thread id)
inline void parallel_for(const ulong N, const uint threads, std::function<void(ulong, uint)> body) {
	const uint T = std::max(1u, std::min(threads, (uint)std::max<ulong>(1ull, N)));
	vector<thread> pool;
	for(uint t = 0; t < T; t++) {
		pool.emplace_back([=, &body]() {
			const ulong a = N * t / T, b = N * (t + 1u) / T;
			for(ulong n = a; n < b; n++) body(n, t);
		});
	}
	for(auto& th : pool) th.join();
}
inline void parallel_for(const ulong N, const uint threads, std::function<void(ulong)> body) {
	parallel_for(N, threads, std::function<void(ulong, uint)>([&](ulong n, uint) { body(n); }));
}
inline void parallel_for(const ulong N, std::function<void(ulong)> body) {
	parallel_for(N, std::max(1u, std::thread::hardware_concurrency()), body);
}

class Clock {
	std::chrono::steady_clock::time_point t0;
public:
	Clock() { start(); }
	void start() { t0 = std::chrono::steady_clock::now(); }
	double stop() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

// C99 linear congruential generator with the reference's scaling (src/utilities.hpp:244-250)
inline float random(uint& seed, const float x = 1.0f) {
	seed = (1103515245u * seed + 12345u) % 2147483648u;
	return x * (float)seed * 4.65661287E-10f;
}
inline float random_symmetric(uint& seed, const float x = 1.0f) { return 2.0f * x * (random(seed, 1.0f) - 0.5f); }

// true once the last three errors differ by less than tolerance (relative)
inline bool converged(const double e2, const double e1, const double e0, const double tolerance) {
	return std::fabs(e1 - e0) < tolerance * std::fabs(e0) && std::fabs(e2 - e1) < tolerance * std::fabs(e1);
}
