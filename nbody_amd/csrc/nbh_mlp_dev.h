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
constexpr int kOut = 9;     // basis coefficients: residual (3) and its symmetric Jacobian (6)
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

// Features, basis and target scale of one (evaluation point p, cell) pair.
//   u = cm - p, d = |u|, uh = u / d; r = bounding radius of the cell (the radius
//   of its opening test, nbody/nbody_space_heap.cpp:130-133), rho = r / d;
//   q = Q / (M r^2) with Q the second central moments; w = q uh; o = (box
//   centre - cm) / r.
// The network sees only rotation invariants (f[16]); its 9 outputs are the
// coefficients of a basis built from uh, w, o and q (b[16] = uh, w, o, q as
// xx yy zz xy xz yz, 0), so the predicted field turns with the cell:
//   R / scale              = c0 uh + c1 w + c2 o
//   kJac J d / scale       = c3 q + c4 (uh w' + w uh') + c5 I + c6 uh uh'
//                            + c7 (uh o' + o uh') + c8 o o'
// with R(p) the residual field (exact field of the cell's bodies - monopole)
// and J = dR/dp, scale = M rho^2 / d^2.  The quadrupole term is exactly of
// this form (c0 = 7.5 uh.w - 1.5 tr q, c1 = -3, c3 = 3 kJac, c4 = -15 kJac,
// c5 = (1.5 tr q - 7.5 uh.w) kJac, c6 = (52.5 uh.w - 7.5 tr q) kJac).
constexpr int kBasis = 16;

// Core of cell_features in arithmetic T: u = cm - p, e = box extents,
// g = box centre - cm, M, Q.
template <typename T>
__device__ __host__ inline float cell_features_t(T ux, T uy, T uz, T M, T ex, T ey, T ez, T gx, T gy, T gz,
                                                 const T* Q, float* f, float* b)
{
	const T d2 = ux * ux + uy * uy + uz * uz;
	const T d = sqrt(d2);
	T r = T(0.5) * sqrt(ex * ex + ey * ey + ez * ez) + sqrt(gx * gx + gy * gy + gz * gz);
	if(!(r > T(0))) { r = T(1e-30); }
	const T inv = T(1) / (M * r * r);
	const T qxx = Q[0] * inv, qyy = Q[1] * inv, qzz = Q[2] * inv;
	const T qxy = Q[3] * inv, qxz = Q[4] * inv, qyz = Q[5] * inv;
	const T id = T(1) / d;
	const T hx = ux * id, hy = uy * id, hz = uz * id;
	const T wx = qxx * hx + qxy * hy + qxz * hz;
	const T wy = qxy * hx + qyy * hy + qyz * hz;
	const T wz = qxz * hx + qyz * hy + qzz * hz;
	const T ir = T(1) / r;
	const T ox = gx * ir, oy = gy * ir, oz = gz * ir;
	const T vx = qxx * ox + qxy * oy + qxz * oz;   // q o
	const T vy = qxy * ox + qyy * oy + qyz * oz;
	const T vz = qxz * ox + qyz * oy + qzz * oz;
	const T rho = r * id;
	const T trq = qxx + qyy + qzz;
	const T uqu = hx * wx + hy * wy + hz * wz;
	const T uo = hx * ox + hy * oy + hz * oz;
	f[0] = static_cast<float>(rho);
	f[1] = static_cast<float>(rho * rho);
	f[2] = static_cast<float>(trq);
	f[3] = static_cast<float>(uqu);
	f[4] = static_cast<float>(wx * wx + wy * wy + wz * wz);
	f[5] = static_cast<float>(qxx * qxx + qyy * qyy + qzz * qzz + T(2) * (qxy * qxy + qxz * qxz + qyz * qyz));
	f[6] = static_cast<float>(uo);
	f[7] = static_cast<float>(ox * ox + oy * oy + oz * oz);
	f[8] = static_cast<float>(wx * ox + wy * oy + wz * oz);
	f[9] = static_cast<float>(ox * vx + oy * vy + oz * vz);
	f[10] = static_cast<float>(rho * uqu);
	f[11] = static_cast<float>(rho * trq);
	f[12] = static_cast<float>(uo * uo);
	f[13] = static_cast<float>(rho * uo);
	f[14] = static_cast<float>(vx * vx + vy * vy + vz * vz);
	f[15] = static_cast<float>(uqu * uqu);
	b[0] = static_cast<float>(hx); b[1] = static_cast<float>(hy); b[2] = static_cast<float>(hz);
	b[3] = static_cast<float>(wx); b[4] = static_cast<float>(wy); b[5] = static_cast<float>(wz);
	b[6] = static_cast<float>(ox); b[7] = static_cast<float>(oy); b[8] = static_cast<float>(oz);
	b[9] = static_cast<float>(qxx); b[10] = static_cast<float>(qyy); b[11] = static_cast<float>(qzz);
	b[12] = static_cast<float>(qxy); b[13] = static_cast<float>(qxz); b[14] = static_cast<float>(qyz);
	b[15] = 0.0f;
	return static_cast<float>(M * rho * rho / d2);
}

// FP64 features of (p, cell): used for the training samples.
__device__ __host__ inline float cell_features(double px, double py, double pz, double cx, double cy, double cz,
                                               double M, const double* bmin, const double* bmax, const double* Q,
                                               float* f, float* b)
{
	return cell_features_t<double>(cx - px, cy - py, cz - pz, M, bmax[0] - bmin[0], bmax[1] - bmin[1],
	                               bmax[2] - bmin[2], 0.5 * (bmax[0] + bmin[0]) - cx, 0.5 * (bmax[1] + bmin[1]) - cy,
	                               0.5 * (bmax[2] + bmin[2]) - cz, Q, f, b);
}

// FP32 features for the walk: the differences are taken in FP64 and rounded
// once, the rest in FP32 (the inputs are rounded to BF16 anyway).
__device__ inline float cell_features_fast(float ux, float uy, float uz, double cx, double cy, double cz, double M,
                                           const double* bmin, const double* bmax, const float* Q, float* f,
                                           float* b)
{
	const float q[6] = {Q[0], Q[1], Q[2], Q[3], Q[4], Q[5]};
	return cell_features_t<float>(ux, uy, uz, static_cast<float>(M), static_cast<float>(bmax[0] - bmin[0]),
	                              static_cast<float>(bmax[1] - bmin[1]), static_cast<float>(bmax[2] - bmin[2]),
	                              static_cast<float>(0.5 * (bmax[0] + bmin[0]) - cx),
	                              static_cast<float>(0.5 * (bmax[1] + bmin[1]) - cy),
	                              static_cast<float>(0.5 * (bmax[2] + bmin[2]) - cz), q, f, b);
}

// Basis element of output column col as its 9 target components (R: 3, then
// the symmetric Jacobian as xx yy zz xy xz yz); zero for col >= kOut.
__device__ __host__ inline void basis_col(const float* b, int col, float* v)
{
	for(int k = 0; k < kOut; ++k) { v[k] = 0.0f; }
	const float ux = b[0], uy = b[1], uz = b[2];
	const float wx = b[3], wy = b[4], wz = b[5];
	const float ox = b[6], oy = b[7], oz = b[8];
	switch(col)
	{
	case 0: v[0] = ux; v[1] = uy; v[2] = uz; break;
	case 1: v[0] = wx; v[1] = wy; v[2] = wz; break;
	case 2: v[0] = ox; v[1] = oy; v[2] = oz; break;
	case 3: for(int k = 0; k < 6; ++k) { v[3 + k] = b[9 + k]; } break;
	case 4:
		v[3] = 2.0f * ux * wx; v[4] = 2.0f * uy * wy; v[5] = 2.0f * uz * wz;
		v[6] = ux * wy + uy * wx; v[7] = ux * wz + uz * wx; v[8] = uy * wz + uz * wy;
		break;
	case 5: v[3] = 1.0f; v[4] = 1.0f; v[5] = 1.0f; break;
	case 6:
		v[3] = ux * ux; v[4] = uy * uy; v[5] = uz * uz; v[6] = ux * uy; v[7] = ux * uz; v[8] = uy * uz;
		break;
	case 7:
		v[3] = 2.0f * ux * ox; v[4] = 2.0f * uy * oy; v[5] = 2.0f * uz * oz;
		v[6] = ux * oy + uy * ox; v[7] = ux * oz + uz * ox; v[8] = uy * oz + uz * oy;
		break;
	case 8:
		v[3] = ox * ox; v[4] = oy * oy; v[5] = oz * oz; v[6] = ox * oy; v[7] = ox * oz; v[8] = oy * oz;
		break;
	default: break;
	}
}

}  // namespace nbh_net
