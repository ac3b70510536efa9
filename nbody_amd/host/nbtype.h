// nbtype.h -- scalar and vector types of the host layer.
//
// Same meaning as the reference's nbody/nbtype.h:16-26 (double precision
// build) and nbody/vertex.h:7-163: nbcoord_t is IEEE binary64, nbvertex_t a
// 3-vector whose dot product and norm sum x, y, z left to right.
#pragma once

#include <cmath>
#include <cstddef>

typedef double nbcoord_t;

struct nbvertex_t
{
	nbcoord_t x, y, z;

	nbvertex_t() : x(0), y(0), z(0) {}
	nbvertex_t(nbcoord_t x_, nbcoord_t y_, nbcoord_t z_) : x(x_), y(y_), z(z_) {}

	nbvertex_t operator+(const nbvertex_t& v) const { return nbvertex_t(x + v.x, y + v.y, z + v.z); }
	nbvertex_t operator-(const nbvertex_t& v) const { return nbvertex_t(x - v.x, y - v.y, z - v.z); }
	nbvertex_t operator*(nbcoord_t a) const { return nbvertex_t(x * a, y * a, z * a); }
	nbvertex_t operator/(nbcoord_t a) const { return nbvertex_t(x / a, y / a, z / a); }
	nbvertex_t operator-() const { return nbvertex_t(-x, -y, -z); }
	nbvertex_t& operator+=(const nbvertex_t& v) { x += v.x; y += v.y; z += v.z; return *this; }
	nbvertex_t& operator-=(const nbvertex_t& v) { x -= v.x; y -= v.y; z -= v.z; return *this; }
	nbvertex_t& operator*=(nbcoord_t a) { x *= a; y *= a; z *= a; return *this; }
	nbvertex_t& operator/=(nbcoord_t a) { x /= a; y /= a; z /= a; return *this; }
	// dot product
	nbcoord_t operator*(const nbvertex_t& v) const { return x * v.x + y * v.y + z * v.z; }
	// cross product
	nbvertex_t operator^(const nbvertex_t& v) const
	{
		return nbvertex_t(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x);
	}
	nbcoord_t norm() const { return x * x + y * y + z * z; }
	nbcoord_t length() const { return std::sqrt((*this) * (*this)); }
	void normalize() { *this = *this / length(); }
};

// Display colour of a body, RGBA in [0, 1] (nbcolor_t of the reference,
// vertex4<float>, nbody/nbtype.h:31): 16 bytes, as the .col stream file holds it.
struct nbcolor_t
{
	float r, g, b, a;

	nbcolor_t() : r(1), g(1), b(1), a(1) {}
	nbcolor_t(float r_, float g_, float b_, float a_) : r(r_), g(g_), b(b_), a(a_) {}
};

#define NBODY_DATA_BLOCK_SIZE 64

namespace nbody {
constexpr nbcoord_t MinDistance = 1e-8;                 // on r^2, nbody/nbtype.h:79
constexpr nbcoord_t GravityConstSI = 6.67259e-11;       // m^3/(kg s^2)
constexpr nbcoord_t Au = 149597870700.0;                // m
constexpr nbcoord_t Day = 86400.0;                      // s
constexpr nbcoord_t GravityConstAuDayKg = (GravityConstSI * Day * Day) / (Au * Au * Au);
}

// Unit systems of the initial-state loaders.
enum e_units_type
{
	eut_G1,
	eut_SI,
	eut_au_day_kg
};

// Kahan step: returns a + b, carrying the lost low part in `correction`
// (same three operations as nbody/summation.h:7-14).
template<class T>
inline T summation_k(const T& a, const T& b, T& correction)
{
	T corrected = b - correction;
	T new_sum = a + corrected;
	correction = (new_sum - a) - corrected;
	return new_sum;
}
