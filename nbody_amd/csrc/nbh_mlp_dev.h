// nbh_mlp_dev.h -- device side of the correction network: a 16-64-64-9 MLP with
// tanh activations, BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate), one wave per tile of 16 samples.
//
// Register layouts of v_mfma_f32_16x16x16bf16_1k for a wave of 64 lanes
// (lane l, e = 0..3):
//   A (16x16, rows i, columns k):  lane l holds A[l % 16][4 (l / 16) + e]
//   B (16x16, rows k, columns j):  lane l holds B[4 (l / 16) + e][l % 16]
//   D (16x16 FP32):                lane l holds D[4 (l / 16) + e][l % 16]
// A D fragment is therefore already the B fragment of the same matrix, and
// the A fragment of its transpose; only D -> A of the same matrix goes
// through LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nbh_net {

constexpr int kIn = 16;      // input features (padded)
constexpr int kHid = 64;     // hidden width (two hidden layers)
constexpr int kOutP = 16;    // output columns computed (first kOut used)
constexpr int kOut = 9;     // residual (3) and its symmetric Jacobian (6)
// The Jacobian outputs are J d / scale times kJac: a body sits within about a
// quarter of d from the group centre, so this weighs a Jacobian error in the
// training loss as it shows in the correction of a body.
constexpr double kJac = 0.25;
// FP32 master parameters, row-major W[in][out]
constexpr int offW1 = 0;
constexpr int offB1 = offW1 + kIn * kHid;        // 1024
constexpr int offW2 = offB1 + kHid;              // 1088
constexpr int offB2 = offW2 + kHid * kHid;       // 5184
constexpr int offW3 = offB2 + kHid;              // 5248
constexpr int offB3 = offW3 + kHid * kOutP;      // 6272
constexpr int kParams = offB3 + kOutP;           // 6288
// packed BF16 fragments (one uint2 per lane per fragment)
constexpr int fragW1 = 0;                        // 4  : k-chunk 0, n-tiles 0..3
constexpr int fragW2 = 4;                        // 16 : k-chunk kc, n-tile nt -> 4 + 4 kc + nt
constexpr int fragW3 = 20;                       // 4  : k-chunk kc -> 20 + kc
constexpr int fragW2T = 24;                      // 16 : (W2^T) k-chunk kc, n-tile nt -> 24 + 4 kc + nt
constexpr int fragW3T = 40;                      // 4  : (W3^T) n-tile nt -> 40 + nt
constexpr int kFrags = 44;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__device__ __host__ inline uint16_t to_bf16(float f)
{
	uint32_t u = __builtin_bit_cast(uint32_t, f);
	u += 0x7FFFu + ((u >> 16) & 1u);
	return static_cast<uint16_t>(u >> 16);
}

__device__ __host__ inline float from_bf16(uint16_t h)
{
	return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

__device__ inline bf16x4 pack4(float a, float b, float c, float d)
{
	bf16x4 r;
	r[0] = static_cast<short>(to_bf16(a));
	r[1] = static_cast<short>(to_bf16(b));
	r[2] = static_cast<short>(to_bf16(c));
	r[3] = static_cast<short>(to_bf16(d));
	return r;
}

__device__ inline bf16x4 pack4(const f32x4& v)
{
	return pack4(v[0], v[1], v[2], v[3]);
}

__device__ inline bf16x4 frag_load(const uint2* __restrict__ frags, int f, int lane)
{
	uint2 u = frags[f * 64 + lane];
	bf16x4 r;
	r[0] = static_cast<short>(u.x & 0xFFFFu);
	r[1] = static_cast<short>(u.x >> 16);
	r[2] = static_cast<short>(u.y & 0xFFFFu);
	r[3] = static_cast<short>(u.y >> 16);
	return r;
}

// tanh(x) = 1 - 2 / (exp(2x) + 1) with the hardware exp2 and reciprocal
// (absolute error ~1e-7, far below the BF16 rounding of the activations).
__device__ inline float tanh_fast(float x)
{
	const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);   // 2 / ln 2
	return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ inline f32x4 mfma(const bf16x4& a, const bf16x4& b, const f32x4& c)
{
	return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// A fragment of rows [16] x columns [kc*16, kc*16+16) of a row-major BF16
// matrix in LDS with `ld` columns.
__device__ inline bf16x4 lds_afrag(const uint16_t* m, int ld, int kc, int lane)
{
	const uint16_t* p = m + (lane & 15) * ld + kc * 16 + 4 * (lane >> 4);
	const uint2 u = *reinterpret_cast<const uint2*>(p);
	bf16x4 r;
	r[0] = static_cast<short>(u.x & 0xFFFFu);
	r[1] = static_cast<short>(u.x >> 16);
	r[2] = static_cast<short>(u.y & 0xFFFFu);
	r[3] = static_cast<short>(u.y >> 16);
	return r;
}

// Store a D fragment (n-tile nt of a 16 x ld matrix) to LDS as BF16.
__device__ inline void lds_store_d(uint16_t* m, int ld, int nt, int lane, const f32x4& v)
{
	const int col = nt * 16 + (lane & 15);
	const int row = 4 * (lane >> 4);
	for(int e = 0; e < 4; ++e)
	{
		m[(row + e) * ld + col] = to_bf16(v[e]);
	}
}

// Forward pass of one tile of 16 samples.  x: the lane's A fragment of the
// inputs.  h1/h2 (optional outputs): tanh activations in D layout; out: the
// output D fragment (columns 0..kOut-1 meaningful).  lds: 2 * 16 * 64 BF16.
__device__ inline f32x4 forward_tile(const bf16x4& x, const uint2* __restrict__ frags, const float* __restrict__ bias,
                                     uint16_t* lds, int lane, f32x4* h1, f32x4* h2)
{
	uint16_t* H1 = lds;
	uint16_t* H2 = lds + 16 * kHid;
	const int col = lane & 15;
	f32x4 a1[4], a2[4];
	for(int nt = 0; nt < 4; ++nt)
	{
		f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
		acc = mfma(x, frag_load(frags, fragW1 + nt, lane), acc);
		const float b = bias[offB1 + nt * 16 + col];
		for(int e = 0; e < 4; ++e) { acc[e] = tanh_fast(acc[e] + b); }
		a1[nt] = acc;
		lds_store_d(H1, kHid, nt, lane, acc);
	}
	__syncthreads();   // callers run work-groups of exactly one wave
	bf16x4 ah[4];
	for(int kc = 0; kc < 4; ++kc) { ah[kc] = lds_afrag(H1, kHid, kc, lane); }
	for(int nt = 0; nt < 4; ++nt)
	{
		f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
		for(int kc = 0; kc < 4; ++kc) { acc = mfma(ah[kc], frag_load(frags, fragW2 + 4 * kc + nt, lane), acc); }
		const float b = bias[offB2 + nt * 16 + col];
		for(int e = 0; e < 4; ++e) { acc[e] = tanh_fast(acc[e] + b); }
		a2[nt] = acc;
		lds_store_d(H2, kHid, nt, lane, acc);
	}
	__syncthreads();
	f32x4 out = {0.0f, 0.0f, 0.0f, 0.0f};
	for(int kc = 0; kc < 4; ++kc) { out = mfma(lds_afrag(H2, kHid, kc, lane), frag_load(frags, fragW3 + kc, lane), out); }
	const float b3 = bias[offB3 + col];
	for(int e = 0; e < 4; ++e) { out[e] += b3; }
	if(h1 != nullptr)
	{
		for(int nt = 0; nt < 4; ++nt) { h1[nt] = a1[nt]; h2[nt] = a2[nt]; }
	}
	return out;
}

// Input features of one (evaluation point p, cell) pair and the scale of the
// target.  u = cm - p, d = |u|, r = bounding radius of the cell (the radius of
// its opening test, nbody/nbody_space_heap.cpp:130-133), q = Q / (M r^2) with Q
// the second central moments of the cell.  The network predicts the residual
// R(p) = exact cell field - monopole field at p, divided by scale = M (r/d)^2 / d^2,
// and its Jacobian dR/dp (symmetric: xx yy zz xy xz yz) multiplied by kJac d / scale,
// so that R(p + delta) ~ scale * (out[0:3] + J(out[3:9]) delta / (kJac d)).
// f: [uhat(3), rho, qxx, qyy, qzz, qxy, qxz, qyz, (q uhat)(3), uhat q uhat, rho^2, tr q]
__device__ __host__ inline float cell_features(double px, double py, double pz, double cx, double cy, double cz,
                                               double M, const double* bmin, const double* bmax, const double* Q,
                                               float* f)
{
	const double ux = cx - px, uy = cy - py, uz = cz - pz;
	const double d2 = ux * ux + uy * uy + uz * uz;
	const double d = sqrt(d2);
	const double ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
	const double gx = 0.5 * (bmax[0] + bmin[0]) - cx, gy = 0.5 * (bmax[1] + bmin[1]) - cy,
	             gz = 0.5 * (bmax[2] + bmin[2]) - cz;
	double r = 0.5 * sqrt(ex * ex + ey * ey + ez * ez) + sqrt(gx * gx + gy * gy + gz * gz);
	if(!(r > 0.0)) { r = 1e-30; }
	const double inv = 1.0 / (M * r * r);
	const double qxx = Q[0] * inv, qyy = Q[1] * inv, qzz = Q[2] * inv;
	const double qxy = Q[3] * inv, qxz = Q[4] * inv, qyz = Q[5] * inv;
	const double hx = ux / d, hy = uy / d, hz = uz / d;
	const double wx = qxx * hx + qxy * hy + qxz * hz;
	const double wy = qxy * hx + qyy * hy + qyz * hz;
	const double wz = qxz * hx + qyz * hy + qzz * hz;
	const double rho = r / d;
	f[0] = static_cast<float>(hx); f[1] = static_cast<float>(hy); f[2] = static_cast<float>(hz);
	f[3] = static_cast<float>(rho);
	f[4] = static_cast<float>(qxx); f[5] = static_cast<float>(qyy); f[6] = static_cast<float>(qzz);
	f[7] = static_cast<float>(qxy); f[8] = static_cast<float>(qxz); f[9] = static_cast<float>(qyz);
	f[10] = static_cast<float>(wx); f[11] = static_cast<float>(wy); f[12] = static_cast<float>(wz);
	f[13] = static_cast<float>(hx * wx + hy * wy + hz * wz);
	f[14] = static_cast<float>(rho * rho);
	f[15] = static_cast<float>(qxx + qyy + qzz);
	return static_cast<float>(M * rho * rho / d2);
}

}  // namespace nbh_net
