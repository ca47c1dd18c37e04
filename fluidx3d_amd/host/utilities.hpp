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
#include <filesystem>
#include <fstream>
#include <unistd.h>

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
// Float and double text in the reference's format (src/utilities.hpp:2599-2666,
// 2745-2786): the value is scaled into [1,10) by multiplications in its own precision
// with the decade steps 10^(2^k), then printed as "I.<8 or 16 decimals>[E<exp>]" with the
// last decimal rounded half up in that precision; "NaN"/"Inf" with a leading '-' for
// negative values.  Scenes print their results with these, so text matches the reference.
template<typename T, int DIGITS, int STEPS> inline string scientific_text(T x, const T (&up)[STEPS], const T (&down)[STEPS], const T (&small)[STEPS], const int (&step)[STEPS]) {
	string text = x < (T)0 ? "-" : "";
	x = std::fabs(x);
	if(std::isnan(x)) return text + "NaN";
	if(std::isinf(x)) return text + "Inf";
	int e10 = 0;
	if(x >= (T)10) for(int k = 0; k < STEPS; k++) if(x >= up[k]) { x *= down[k]; e10 += step[k]; }
	if(x > (T)0 && x <= (T)1) for(int k = 0; k < STEPS; k++) if(x < small[k]) { x *= up[k]; e10 -= step[k]; }
	typedef typename std::conditional<DIGITS == 8, uint, ulong>::type D;
	const T scale = DIGITS == 8 ? (T)1E8 : (T)1E16;
	const D limit = DIGITS == 8 ? (D)100000000u : (D)10000000000000000ull;
	uint whole = (uint)x;
	const T frac = (x - (T)whole) * scale;
	D digits = (D)frac;
	if(frac - (T)digits >= (T)0.5 && ++digits >= limit) {
		digits = 0;
		if(++whole >= 10u) { whole = 1u; e10++; }
	}
	string dec(DIGITS, '0');
	for(int k = DIGITS - 1; k >= 0; k--) { dec[k] = (char)('0' + digits % 10u); digits /= 10u; }
	return text + std::to_string(whole) + "." + dec + (e10 != 0 ? "E" + std::to_string(e10) : "");
}
inline string to_string(float x) {
	static const float up[6] = {1E32f, 1E16f, 1E8f, 1E4f, 1E2f, 1E1f};
	static const float down[6] = {1E-32f, 1E-16f, 1E-8f, 1E-4f, 1E-2f, 1E-1f};
	static const float small[6] = {1E-31f, 1E-15f, 1E-7f, 1E-3f, 1E-1f, 1E0f};
	static const int step[6] = {32, 16, 8, 4, 2, 1};
	return scientific_text<float, 8, 6>(x, up, down, small, step);
}
inline string to_string(double x) {
	static const double up[9] = {1E256, 1E128, 1E64, 1E32, 1E16, 1E8, 1E4, 1E2, 1E1};
	static const double down[9] = {1E-256, 1E-128, 1E-64, 1E-32, 1E-16, 1E-8, 1E-4, 1E-2, 1E-1};
	static const double small[9] = {1E-255, 1E-127, 1E-63, 1E-31, 1E-15, 1E-7, 1E-3, 1E-1, 1E0};
	static const int step[9] = {256, 128, 64, 32, 16, 8, 4, 2, 1};
	return scientific_text<double, 16, 9>(x, up, down, small, step);
}
// Fixed number of decimals (at most 8 for float, 16 for double), rounded by adding half
// a unit of the last decimal in the value's own precision, then truncated.
template<typename T> inline string fixed_text(T x, const uint decimals, const uint max_decimals) {
	string text = x < (T)0 ? "-" : "";
	x = std::fabs(x);
	if(std::isnan(x)) return text + "NaN";
	if(std::isinf(x) || x > (T)max_ulong) return text + "Inf";
	const uint d = std::min(decimals, max_decimals);
	const T power = std::pow((T)10, (T)d);
	x += (T)0.5 / power;
	const ulong whole = (ulong)x;
	ulong digits = (ulong)((x - (T)whole) * power);
	text += std::to_string(whole);
	if(decimals == 0u) return text;
	string dec(d, '0');
	for(int k = (int)d - 1; k >= 0; k--) { dec[k] = (char)('0' + digits % 10ull); digits /= 10ull; }
	return text + "." + dec;
}
inline string to_string(double x, uint decimals) { return fixed_text<double>(x, decimals, 16u); }
inline string to_string(float x, uint decimals) { return fixed_text<float>(x, decimals, 8u); }

inline void print(const string& s) { std::cout << s << std::flush; }
inline void println(const string& s = "") { std::cout << s << std::endl; }
template<typename T> inline void println(const T& x) { println(to_string(x)); }
inline void print_info(const string& s) { println("Info: " + s); }
inline void print_warning(const string& s) { println("Warning: " + s); }
[[noreturn]] inline void print_error(const string& s) { std::cerr << "Error: " << s << std::endl; std::exit(1); }
inline void wait() {} // the reference waits for a key press; this build runs headless
inline void sleep(const double seconds) { std::this_thread::sleep_for(std::chrono::duration<double>(seconds)); }
// directory of the running executable, ending with '/' (src/utilities.hpp:3204-3228)
inline string get_exe_path() {
	char buffer[4096];
	const ssize_t n = readlink("/proc/self/exe", buffer, sizeof(buffer) - 1u);
	if(n <= 0) return "";
	const string path(buffer, (size_t)n);
	return path.substr(0, path.rfind('/') + 1u);
}
// the folder part of a file path, created if missing (src/utilities.hpp:4197-4204)
inline void create_folder(const string& path) {
	const size_t slash = path.rfind('/');
	if(slash == string::npos) return;
	const string folder = path.substr(0, slash);
	if(!folder.empty() && !std::filesystem::is_directory(folder)) std::filesystem::create_directories(folder);
}
// replaces the file's extension (from its last '.') by `extension` (src/utilities.hpp:4205-4207)
inline string create_file_extension(const string& filename, const string& extension) {
	return filename.substr(0, filename.rfind('.')) + (extension.at(0) != '.' ? "." : "") + extension;
}
inline void write_file(const string& filename, const string& content = "") {
	create_folder(filename);
	std::ofstream file(filename, std::ios::out);
	file.write(content.c_str(), (std::streamsize)content.length());
}
// file name of exported data: `path` is a prefix (a folder ending in '/', default the
// executable's export/ folder), then <name>-<t as 9 digits><extension>
// (src/lbm.cpp:85-92)
inline string default_filename(const string& path, const string& name, const string& extension, const ulong t) {
	const string digits = std::to_string(t);
	const string stamp = digits.size() >= 9u ? digits.substr(digits.size() - 9u) : string(9u - digits.size(), '0') + digits;
	return (path == "" ? get_exe_path() + "export/" : path) + create_file_extension((name == "" ? "file" : name) + "-" + stamp, extension);
}
inline string default_filename(const string& name, const string& extension, const ulong t) { return default_filename("", name, extension, t); }
// seconds as [y][ddd d][hh h][mm m]ss s, leading units only where reached (src/utilities.hpp:2876-2888)
inline string print_time(const double seconds) {
	uint s = (uint)(seconds + 0.5);
	const uint y = s / 31556925u; s %= 31556925u;
	const uint d = s / 86400u; s %= 86400u;
	const uint h = s / 3600u; s %= 3600u;
	const uint m = s / 60u; s %= 60u;
	const bool show_y = seconds > 31556924.5, show_d = seconds > 86399.5, show_h = seconds > 3599.5, show_m = seconds > 59.5;
	string r;
	if(show_y) r += std::to_string(y) + "y";
	if(show_d) r += string(show_y && d < 10u ? "00" : show_y && d < 100u ? "0" : "") + std::to_string(d) + "d";
	if(show_h) r += string(show_d && h < 10u ? "0" : "") + std::to_string(h) + "h";
	if(show_m) r += string(show_h && m < 10u ? "0" : "") + std::to_string(m) + "m";
	return r + string(show_m && s < 10u ? "0" : "") + std::to_string(s) + "s";
}

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

// Convergence of a sequence (src/utilities.hpp:208-213, 330-335): the last slope
// |x - x_last| and curvature |x - 2 x_last + x_penultimate| are both below threshold times
// the mean of the three values.
inline bool converged(const double x_penultimate, const double x_last, const double x, const double threshold = 1E-5) {
	const double bound = (x_penultimate + x_last + x) / 3.0 * threshold;
	return std::fabs(x - x_last) < bound && std::fabs(x - 2.0 * x_last + x_penultimate) < bound;
}
inline bool converged(const float x_penultimate, const float x_last, const float x, const float threshold = 1E-5f) {
	const float bound = (x_penultimate + x_last + x) / 3.0f * threshold;
	return std::fabs(x - x_last) < bound && std::fabs(x - 2.0f * x_last + x_penultimate) < bound;
}
inline ulong to_ulong(const float x) { return (ulong)std::max(x + 0.5f, 0.5f); }
inline ulong to_ulong(const double x) { return (ulong)std::max(x + 0.5, 0.5); }
