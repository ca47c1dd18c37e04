#include "shapes.hpp"

// Each predicate tests the cell position (x,y,z) against a body centred at p.
bool sphere(const uint x, const uint y, const uint z, const float3& p, const float r) {
	const float3 d = float3(x, y, z) - p;
	return sq(d.x) + sq(d.y) + sq(d.z) <= sq(r);
}
bool ellipsoid(const uint x, const uint y, const uint z, const float3& p, const float3& r) {
	const float3 d = float3(x, y, z) - p;
	return sq(d.x / r.x) + sq(d.y / r.y) + sq(d.z / r.z) <= 1.0f;
}
bool cube(const uint x, const uint y, const uint z, const float3& p, const float l) {
	return cuboid(x, y, z, p, float3(l));
}
bool cuboid(const uint x, const uint y, const uint z, const float3& p, const float3& l) {
	const float3 d = float3(x, y, z) - p;
	return std::fabs(d.x) <= 0.5f * l.x && std::fabs(d.y) <= 0.5f * l.y && std::fabs(d.z) <= 0.5f * l.z;
}
bool cylinder(const uint x, const uint y, const uint z, const float3& p, const float3& n, const float r) {
	const float3 d = float3(x, y, z) - p;
	const float along = dot(normalize(n), d);
	const float radial2 = sq(d.x) + sq(d.y) + sq(d.z) - sq(along);
	return radial2 <= sq(r) && sq(along) <= sq(0.5f * length(n));
}
