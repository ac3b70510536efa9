// nbh_direct.hip -- direct all-pairs accelerations on gfx950.
//
// FP64 kernel: two targets per work-item, sources staged in LDS tiles, the
// source loop runs in index order and each pair is evaluated with exactly the
// operations of nbody_data::force followed by the division by m1 of
// nbody_engine_simple::fcompute (nbody/nbody_data.cpp:35-44,
// nbody/nbody_engine_simple.cpp:152-171).  This file is compiled with
// -ffp-contract=off, so the FP64 result equals the CPU restatement bit for bit
// (IEEE sqrt and division are correctly rounded on both sides).
//
// FP32 kernel (the throughput path): the bodies are sorted along a Morton curve
// and copied once into records (x, y, z, m) in that order; each work-group owns
// 256*TGT consecutive targets and a slice of the sources.  Sources are converted
// once per work-group from FP64 to FP32 coordinates relative to the
// work-group's origin (the FP64 position of its first target), staged in LDS as
// float4 (x, y, z, m) and read back as wave-wide broadcasts; each work-item
// reuses every source for TGT targets held in registers.  Per pair: 3 sub,
// 3 mul/fma for r^2, 1 v_rsq_f32 whose output clamp is the reference's
// r^2 >= 1e-8 clamp (in scaled units), 3 mul, 3 fma.  Partial sums of
// one LDS tile stay in FP32 and are folded into a Kahan (summation_k) FP32
// accumulator per tile; the source slices are then added in FP64 in a fixed
// order by a second kernel, which also copies the velocities into f[0:3n].

#include <cmath>
#include <mutex>

#include "nbh_common.h"
#include "nbh_f64.h"

namespace {

constexpr double kMinDistance = 1e-8;   // nbody/nbtype.h:79

// ---------------------------------------------------------------------------
// FP64, reference operation order
// ---------------------------------------------------------------------------
// Pair laws of the FP64 kernel.  kLawSimple: one term of nbody_data::force
// (nbody/nbody_data.cpp:35-44) with the mass factor -m1 precomputed,
// t += d * ((gm1 * m2) / (r2 * sqrt(r2))), divided by m1 at the end
// (nbody_engine_simple.cpp:169-171).  kLawBlock: nbody_engine_block's term
// (nbody/nbody_engine_block.cpp:96-112), t -= d * (m2 / (sqrt(r2) * r2)),
// which is the acceleration itself.  d = r1 - r2 in both, r2 clamped at 1e-8.
constexpr int kLawSimple = 0;
constexpr int kLawBlock = 1;

template <int kLaw, bool kFast>
__device__ inline void pair64(double x1, double y1, double z1, double gm1, const double4& s, double& tx, double& ty,
                              double& tz)
{
	const double dx = x1 - s.x;
	const double dy = y1 - s.y;
	const double dz = z1 - s.z;
	double r2 = dx * dx + dy * dy + dz * dz;
	double coeff;
	if constexpr(kFast)
	{
		r2 = fmax(r2, kMinDistance);     // r2 is finite here: the same as the branch below
		coeff = kLaw == kLawSimple ? div_rn(gm1 * s.w, r2 * sqrt_rn(r2)) : div_rn(s.w, sqrt_rn(r2) * r2);
	}
	else
	{
		if(r2 < kMinDistance)
		{
			r2 = kMinDistance;
		}
		coeff = kLaw == kLawSimple ? (gm1 * s.w) / (r2 * sqrt(r2)) : s.w / (sqrt(r2) * r2);
	}
	if constexpr(kLaw == kLawSimple)
	{
		tx += dx * coeff;
		ty += dy * coeff;
		tz += dz * coeff;
	}
	else
	{
		tx -= dx * coeff;
		ty -= dy * coeff;
		tz -= dz * coeff;
	}
}

// kF64Tgt targets per work-item (i0 + k * blockDim.x), sources staged in LDS as
// double4 (x, y, z, m) tiles of blockDim.x and read back as wave-wide
// broadcasts, each reused for every target of the work-item.  The source loop
// runs in index order.  The simple law skips the self pair in the reference;
// here, on the fast path, it is summed: dx = dy = dz = +0 there, so its term is
// -0 and adding -0 leaves every value (+0 included) unchanged.  The block law
// sums it in the reference too (its term subtracts +0).  Two targets per
// work-item when that still leaves two waves per SIMD (n >= 2^18 on 256 CUs),
// else one, so that smaller n still spreads over every SIMD.
template <int kLaw, int kF64Tgt>
__global__ void __launch_bounds__(1024)
k_direct_fp64(const double* __restrict__ y, const double* __restrict__ mass, double* __restrict__ f, size_t n)
{
	extern __shared__ double4 src[];
	const int bs = blockDim.x;
	const size_t i0 = static_cast<size_t>(blockIdx.x) * (kF64Tgt * bs) + threadIdx.x;
	double x1[kF64Tgt], y1[kF64Tgt], z1[kF64Tgt], m1[kF64Tgt], gm1[kF64Tgt];
	double tx[kF64Tgt], ty[kF64Tgt], tz[kF64Tgt];
	bool safe = true;
	for(int k = 0; k < kF64Tgt; ++k)
	{
		const size_t i = i0 + static_cast<size_t>(k) * bs;
		const bool active = i < n;
		x1[k] = active ? y[i] : 0.0;
		y1[k] = active ? y[n + i] : 0.0;
		z1[k] = active ? y[2 * n + i] : 0.0;
		m1[k] = active ? mass[i] : 1.0;
		gm1[k] = -1.0 * m1[k];            // (-GravityConst * mass1), G = 1
		tx[k] = ty[k] = tz[k] = 0.0;
		safe = safe && coord_in_range(x1[k]) && coord_in_range(y1[k]) && coord_in_range(z1[k]) &&
		       (!active || kLaw == kLawBlock || mass_in_range(m1[k]));
	}
	for(size_t base = 0; base < n; base += bs)
	{
		const size_t j = base + threadIdx.x;
		__syncthreads();
		double4 s = make_double4(0.0, 0.0, 0.0, 0.0);
		bool ok = true;
		if(j < n)
		{
			s = make_double4(y[j], y[n + j], y[2 * n + j], mass[j]);
			ok = coord_in_range(s.x) && coord_in_range(s.y) && coord_in_range(s.z) && mass_in_range(s.w);
		}
		src[threadIdx.x] = s;
		const bool fast = __syncthreads_and(ok && safe) != 0;
		const int cnt = n - base < static_cast<size_t>(bs) ? static_cast<int>(n - base) : bs;
		if(fast)
		{
			for(int jj = 0; jj < cnt; ++jj)
			{
				const double4 sj = src[jj];
				for(int k = 0; k < kF64Tgt; ++k)
				{
					pair64<kLaw, true>(x1[k], y1[k], z1[k], gm1[k], sj, tx[k], ty[k], tz[k]);
				}
			}
		}
		else
		{
			for(int jj = 0; jj < cnt; ++jj)
			{
				const double4 sj = src[jj];
				for(int k = 0; k < kF64Tgt; ++k)
				{
					if(kLaw == kLawBlock || base + jj != i0 + static_cast<size_t>(k) * bs)
					{
						pair64<kLaw, false>(x1[k], y1[k], z1[k], gm1[k], sj, tx[k], ty[k], tz[k]);
					}
				}
			}
		}
	}
	for(int k = 0; k < kF64Tgt; ++k)
	{
		const size_t i = i0 + static_cast<size_t>(k) * bs;
		if(i < n)
		{
			f[i] = y[3 * n + i];
			f[n + i] = y[4 * n + i];
			f[2 * n + i] = y[5 * n + i];
			f[3 * n + i] = kLaw == kLawSimple ? tx[k] / m1[k] : tx[k];
			f[4 * n + i] = kLaw == kLawSimple ? ty[k] / m1[k] : ty[k];
			f[5 * n + i] = kLaw == kLawSimple ? tz[k] / m1[k] : tz[k];
		}
	}
}

// Same sum for the targets idx[0..m): acc[3t..3t+2] = acceleration of body idx[t].
__global__ void __launch_bounds__(256)
k_direct_sample64(const double* __restrict__ y, const double* __restrict__ mass, size_t n,
                  const uint32_t* __restrict__ idx, size_t m, double* __restrict__ acc)
{
	__shared__ double sx[256], sy[256], sz[256], sm[256];
	const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	const bool active = t < m;
	const size_t i = active ? idx[t] : 0;
	const double x1 = active ? y[i] : 0.0;
	const double y1 = active ? y[n + i] : 0.0;
	const double z1 = active ? y[2 * n + i] : 0.0;
	const double m1 = active ? mass[i] : 1.0;
	const double gm1 = -1.0 * m1;
	double tx = 0.0, ty = 0.0, tz = 0.0;
	for(size_t base = 0; base < n; base += 256)
	{
		const size_t j = base + threadIdx.x;
		__syncthreads();
		sx[threadIdx.x] = j < n ? y[j] : 0.0;
		sy[threadIdx.x] = j < n ? y[n + j] : 0.0;
		sz[threadIdx.x] = j < n ? y[2 * n + j] : 0.0;
		sm[threadIdx.x] = j < n ? mass[j] : 0.0;
		__syncthreads();
		const int cnt = n - base < 256 ? static_cast<int>(n - base) : 256;
		const int self = (active && i >= base && i < base + cnt) ? static_cast<int>(i - base) : -1;
		for(int jj = 0; jj < cnt; ++jj)
		{
			if(jj == self)
			{
				continue;
			}
			const double dx = x1 - sx[jj];
			const double dy = y1 - sy[jj];
			const double dz = z1 - sz[jj];
			double r2 = dx * dx + dy * dy + dz * dz;
			if(r2 < kMinDistance)
			{
				r2 = kMinDistance;
			}
			const double coeff = (gm1 * sm[jj]) / (r2 * sqrt(r2));
			tx += dx * coeff;
			ty += dy * coeff;
			tz += dz * coeff;
		}
	}
	if(active)
	{
		acc[3 * t] = tx / m1;
		acc[3 * t + 1] = ty / m1;
		acc[3 * t + 2] = tz / m1;
	}
}

// ---------------------------------------------------------------------------
// FP32 tiles
// ---------------------------------------------------------------------------
constexpr int kF32Threads = 256;   // work-items per work-group (4 waves)
constexpr int kF32Tgt = 4;         // targets per work-item
constexpr int kF32Tile = 256;      // sources per LDS tile (one per work-item)
constexpr int kF32TgtPerGroup = kF32Threads * kF32Tgt;

// The tiles work in scaled units: offsets x' = 1e4 x and masses m' = 1e8 m (both
// formed in FP64 and rounded once).  Then r'^2 = 1e8 r^2, and the reference's
// clamp r^2 >= 1e-8 (nbody/nbody_data.cpp:35-44) becomes r'^2 >= 1: with
// q = min(1/r', 1), the rsq instruction's own [0, 1] output clamp,
//   q = 1e-4 / sqrt(max(r^2, 1e-8)),
// and m' x' q^3 = m x / max(r^2, 1e-8)^{3/2} with no rescaling of the sums.
// A coincident pair (r' = 0) gets q = 1 and contributes m' * 0 = 0.
constexpr double kF32Len = 1e4;
constexpr double kF32Mass = 1e8;

__device__ inline float rsq_clamped(float r2)
{
	return __builtin_amdgcn_fmed3f(__builtin_amdgcn_rsqf(r2), 0.0f, 1.0f);   // folds into v_rsq_f32 clamp
}

__device__ inline void kahan_f32(float& s, float& c, float v)
{
	float corrected = v - c;
	float ns = s + corrected;
	c = (ns - s) - corrected;
	s = ns;
}

// The state in Morton order as one record per body: sorted[p] = (x, y, z, m) of
// body perm[p].  Written once per force evaluation, so that the tile kernel
// reads its targets and sources as contiguous 32-byte records instead of
// gathering three coordinates through the permutation.
__global__ void __launch_bounds__(256)
k_sorted_copy(const double* __restrict__ y, const double* __restrict__ mass, const uint32_t* __restrict__ perm,
              double4* __restrict__ sorted, size_t n)
{
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride)
	{
		const size_t b = perm[p];
		sorted[p] = make_double4(y[b], y[n + b], y[2 * n + b], mass[b]);
	}
}

// part: [nsplit][3][n] doubles by Morton position, slice blockIdx.y of the
// sources (also in Morton order).  A work-group's 1024 targets are consecutive
// along the curve, so they are close together and their offsets from its
// origin (its first target) -- and those of the sources near them -- are
// small in FP32.
__global__ void __launch_bounds__(kF32Threads)
k_direct_fp32(const double4* __restrict__ sorted, double* __restrict__ part, size_t n, size_t jchunk)
{
	__shared__ float4 tile[kF32Tile];

	const size_t tbase = static_cast<size_t>(blockIdx.x) * kF32TgtPerGroup;
	const double4 org = sorted[tbase];
	const double ox = org.x;
	const double oy = org.y;
	const double oz = org.z;

	float xi[kF32Tgt], yi[kF32Tgt], zi[kF32Tgt];
	float sx[kF32Tgt], sy[kF32Tgt], sz[kF32Tgt];
	float cx[kF32Tgt], cy[kF32Tgt], cz[kF32Tgt];
#pragma unroll
	for(int k = 0; k < kF32Tgt; ++k)
	{
		size_t i = tbase + static_cast<size_t>(k) * kF32Threads + threadIdx.x;
		bool ok = i < n;
		const double4 t = ok ? sorted[i] : org;
		xi[k] = ok ? static_cast<float>((t.x - ox) * kF32Len) : 0.0f;
		yi[k] = ok ? static_cast<float>((t.y - oy) * kF32Len) : 0.0f;
		zi[k] = ok ? static_cast<float>((t.z - oz) * kF32Len) : 0.0f;
		sx[k] = sy[k] = sz[k] = 0.0f;
		cx[k] = cy[k] = cz[k] = 0.0f;
	}

	const size_t jbeg = static_cast<size_t>(blockIdx.y) * jchunk;
	const size_t jend = jbeg + jchunk < n ? jbeg + jchunk : n;

	for(size_t base = jbeg; base < jend; base += kF32Tile)
	{
		const size_t j = base + threadIdx.x;
		__syncthreads();
		if(j < jend)
		{
			const double4 sj = sorted[j];
			tile[threadIdx.x] = make_float4(static_cast<float>((sj.x - ox) * kF32Len),
			                                static_cast<float>((sj.y - oy) * kF32Len),
			                                static_cast<float>((sj.z - oz) * kF32Len),
			                                static_cast<float>(sj.w * kF32Mass));
		}
		else
		{
			tile[threadIdx.x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // zero mass: no force
		}
		__syncthreads();

		float px[kF32Tgt], py[kF32Tgt], pz[kF32Tgt];
#pragma unroll
		for(int k = 0; k < kF32Tgt; ++k) { px[k] = py[k] = pz[k] = 0.0f; }

#pragma unroll 4
		for(int jj = 0; jj < kF32Tile; ++jj)
		{
			const float4 s = tile[jj];
#pragma unroll
			for(int k = 0; k < kF32Tgt; ++k)
			{
				const float dx = s.x - xi[k];
				const float dy = s.y - yi[k];
				const float dz = s.z - zi[k];
				const float q = rsq_clamped(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
				const float w = (s.w * q) * (q * q);
				px[k] = fmaf(dx, w, px[k]);
				py[k] = fmaf(dy, w, py[k]);
				pz[k] = fmaf(dz, w, pz[k]);
			}
		}
#pragma unroll
		for(int k = 0; k < kF32Tgt; ++k)
		{
			kahan_f32(sx[k], cx[k], px[k]);
			kahan_f32(sy[k], cy[k], py[k]);
			kahan_f32(sz[k], cz[k], pz[k]);
		}
	}

	double* out = part + static_cast<size_t>(blockIdx.y) * 3 * n;
#pragma unroll
	for(int k = 0; k < kF32Tgt; ++k)
	{
		size_t i = tbase + static_cast<size_t>(k) * kF32Threads + threadIdx.x;
		if(i < n)
		{
			out[i] = static_cast<double>(sx[k]) - static_cast<double>(cx[k]);
			out[n + i] = static_cast<double>(sy[k]) - static_cast<double>(cy[k]);
			out[2 * n + i] = static_cast<double>(sz[k]) - static_cast<double>(cz[k]);
		}
	}
}

// FP64 correction pass of NBH_FP32_CORR: for the target at Morton position p,
// the sources at positions p - w .. p + w (its neighbours along the curve, where
// the FP32 rounding of the offsets weighs most) are recomputed in FP64 (force
// law of nbody/nbody_data.cpp:35-44) and in the FP32 tiles' own arithmetic (same
// work-group origin, scaled units and clamped rsq), and f gets the difference.
constexpr int kCorrWindow = 32;

__global__ void __launch_bounds__(256)
k_direct_window(const double4* __restrict__ sorted, const uint32_t* __restrict__ perm, double* __restrict__ f,
                size_t n, int w)
{
	const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if(p >= n) { return; }
	const uint32_t i = perm[p];
	const double4 org = sorted[p / kF32TgtPerGroup * kF32TgtPerGroup];     // the tile kernel's origin
	const double ox = org.x, oy = org.y, oz = org.z;
	const double4 me = sorted[p];
	const double xi = me.x, yi = me.y, zi = me.z;
	const float txf = static_cast<float>((xi - ox) * kF32Len);
	const float tyf = static_cast<float>((yi - oy) * kF32Len);
	const float tzf = static_cast<float>((zi - oz) * kF32Len);
	double cx = 0.0, cy = 0.0, cz = 0.0;
	const size_t q0 = p > static_cast<size_t>(w) ? p - w : 0;
	const size_t q1 = p + w < n - 1 ? p + w : n - 1;
	for(size_t q = q0; q <= q1; ++q)
	{
		if(q == p) { continue; }
		const double4 sj = sorted[q];
		const double xj = sj.x, yj = sj.y, zj = sj.z, mj = sj.w;
		const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
		double r2 = dx * dx + dy * dy + dz * dz;
		if(r2 < kMinDistance) { r2 = kMinDistance; }
		const double c = mj / (r2 * sqrt(r2));
		const float fx = static_cast<float>((xj - ox) * kF32Len) - txf;
		const float fy = static_cast<float>((yj - oy) * kF32Len) - tyf;
		const float fz = static_cast<float>((zj - oz) * kF32Len) - tzf;
		const float qf = rsq_clamped(fmaf(fz, fz, fmaf(fy, fy, fx * fx)));
		const float wf = (static_cast<float>(mj * kF32Mass) * qf) * (qf * qf);
		cx += dx * c - static_cast<double>(fx * wf);
		cy += dy * c - static_cast<double>(fy * wf);
		cz += dz * c - static_cast<double>(fz * wf);
	}
	f[3 * n + i] += cx;
	f[4 * n + i] += cy;
	f[5 * n + i] += cz;
}

// f[0:3n] = v, f[3n:6n] = sum over slices (fixed order).
__global__ void __launch_bounds__(256)
k_direct_fp32_finish(const double* __restrict__ y, const uint32_t* __restrict__ perm,
                     const double* __restrict__ part, double* __restrict__ f, size_t n, int nsplit)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		double ax = 0.0, ay = 0.0, az = 0.0;
		for(int s = 0; s < nsplit; ++s)
		{
			const double* p = part + static_cast<size_t>(s) * 3 * n;
			ax += p[i];
			ay += p[n + i];
			az += p[2 * n + i];
		}
		f[i] = y[3 * n + i];
		f[n + i] = y[4 * n + i];
		f[2 * n + i] = y[5 * n + i];
		const size_t b = perm[i];
		f[3 * n + b] = ax;
		f[4 * n + b] = ay;
		f[5 * n + b] = az;
	}
}

// Bounding box of the positions: per work-group minima and maxima ...
__global__ void __launch_bounds__(256)
k_bbox_partial(const double* __restrict__ y, size_t n, double* __restrict__ part)
{
	__shared__ double sh[6][256];
	double lo[3] = {INFINITY, INFINITY, INFINITY};
	double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		for(int d = 0; d < 3; ++d)
		{
			const double v = y[d * n + i];
			lo[d] = fmin(lo[d], v);
			hi[d] = fmax(hi[d], v);
		}
	}
	for(int d = 0; d < 3; ++d) { sh[d][threadIdx.x] = lo[d]; sh[3 + d][threadIdx.x] = hi[d]; }
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1)
	{
		if(static_cast<int>(threadIdx.x) < off)
		{
			for(int d = 0; d < 3; ++d)
			{
				sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + off]);
				sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + off]);
			}
		}
		__syncthreads();
	}
	if(threadIdx.x < 6) { part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0]; }
}

// ... folded by one work-group into box[6] = min xyz, max xyz.
__global__ void __launch_bounds__(256)
k_bbox_final(const double* __restrict__ part, int g, double* __restrict__ box)
{
	__shared__ double sh[6][256];
	double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
	for(int b = threadIdx.x; b < g; b += blockDim.x)
	{
		for(int d = 0; d < 3; ++d)
		{
			v[d] = fmin(v[d], part[6 * b + d]);
			v[3 + d] = fmax(v[3 + d], part[6 * b + 3 + d]);
		}
	}
	for(int d = 0; d < 6; ++d) { sh[d][threadIdx.x] = v[d]; }
	__syncthreads();
	for(int off = 128; off > 0; off >>= 1)
	{
		if(static_cast<int>(threadIdx.x) < off)
		{
			for(int d = 0; d < 3; ++d)
			{
				sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + off]);
				sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + off]);
			}
		}
		__syncthreads();
	}
	if(threadIdx.x < 6) { box[threadIdx.x] = sh[threadIdx.x][0]; }
}

__device__ inline uint32_t spread10(uint32_t x)
{
	x = (x | (x << 16)) & 0x030000FFu;
	x = (x | (x << 8)) & 0x0300F00Fu;
	x = (x | (x << 4)) & 0x030C30C3u;
	x = (x | (x << 2)) & 0x09249249u;
	return x;
}

// 30-bit Morton keys of the positions on a 1024^3 grid over the box (values:
// body indices).  Non-finite coordinates map to cell 0.
__global__ void __launch_bounds__(256)
k_morton_keys(const double* __restrict__ y, size_t n, const double* __restrict__ box, uint64_t* __restrict__ keys,
              uint32_t* __restrict__ vals)
{
	double lo[3], sc[3];
	for(int d = 0; d < 3; ++d)
	{
		lo[d] = box[d];
		const double ext = box[3 + d] - box[d];
		sc[d] = ext > 0.0 ? 1024.0 / ext : 0.0;
	}
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		uint32_t q[3];
		for(int d = 0; d < 3; ++d)
		{
			const double t = (y[d * n + i] - lo[d]) * sc[d];
			q[d] = t >= 1023.0 ? 1023u : (t > 0.0 ? static_cast<uint32_t>(t) : 0u);
		}
		keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
		vals[i] = static_cast<uint32_t>(i);
	}
}

// Launch shape and scratch layout of the FP32 path for n bodies: the sources
// are split so that the grid holds ~8 work-groups per CU; the scratch holds
// the split partial sums, the Morton-ordered records, the sort keys / values
// (two halves each), sort counts, scan chunk sums, bounding-box partials and
// the box, each at a 256-byte boundary.
struct fp32_layout {
	size_t gx, nsplit, jchunk;
	unsigned gbox;
	size_t off_sorted, off_keys, off_vals, off_counts, off_scan, off_bpart, off_box, bytes;
};

fp32_layout fp32_plan(const nbh_ctx* ctx, size_t n)
{
	fp32_layout l;
	l.gx = (n + kF32TgtPerGroup - 1) / kF32TgtPerGroup;
	const size_t want = static_cast<size_t>(ctx->cu_count > 0 ? ctx->cu_count : 256) * 8;
	size_t nsplit = (want + l.gx - 1) / l.gx;
	const size_t max_split = (n + kF32Tile - 1) / kF32Tile;
	if(nsplit > max_split) { nsplit = max_split; }
	if(nsplit > 64) { nsplit = 64; }
	if(nsplit < 1) { nsplit = 1; }
	size_t jchunk = (n + nsplit - 1) / nsplit;
	l.jchunk = (jchunk + kF32Tile - 1) / kF32Tile * kF32Tile;
	l.nsplit = (n + l.jchunk - 1) / l.jchunk;
	l.gbox = nbh::grid_for(ctx, n, 256);
	auto align = [](size_t b) { return (b + 255) / 256 * 256; };
	l.off_sorted = align(l.nsplit * 3 * n * sizeof(double));
	l.off_keys = l.off_sorted + align(n * sizeof(double4));
	l.off_vals = l.off_keys + align(2 * n * sizeof(uint64_t));
	l.off_counts = l.off_vals + align(2 * n * sizeof(uint32_t));
	l.off_scan = l.off_counts + align(nbh::sort_counts_size(n) * sizeof(uint32_t));
	l.off_bpart = l.off_scan + align(nbh::sort_scan_size() * sizeof(uint32_t));
	l.off_box = l.off_bpart + align(6 * static_cast<size_t>(l.gbox) * sizeof(double));
	l.bytes = l.off_box + align(6 * sizeof(double));
	return l;
}

// The FP32 path with its scratch at `base` (l.bytes bytes).  Every scratch
// word it reads was written earlier in the same call, so the scratch needs no
// clearing.  timed: bracket the tile kernel with the context's timer.
nbh_status fp32_run(nbh_ctx* ctx, const double* y, const double* mass, double* f, size_t n, nbh_precision precision,
                    const fp32_layout& l, char* base, bool timed)
{
	double* part = reinterpret_cast<double*>(base);
	double4* sorted = reinterpret_cast<double4*>(base + l.off_sorted);
	uint64_t* keys = reinterpret_cast<uint64_t*>(base + l.off_keys);
	uint32_t* vals = reinterpret_cast<uint32_t*>(base + l.off_vals);
	uint32_t* counts = reinterpret_cast<uint32_t*>(base + l.off_counts);
	uint32_t* scanp = reinterpret_cast<uint32_t*>(base + l.off_scan);
	double* bpart = reinterpret_cast<double*>(base + l.off_bpart);
	double* box = reinterpret_cast<double*>(base + l.off_box);
	hipLaunchKernelGGL(k_bbox_partial, dim3(l.gbox), dim3(256), 0, ctx->stream, y, n, bpart);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, ctx->stream, bpart, static_cast<int>(l.gbox), box);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_morton_keys, dim3(l.gbox), dim3(256), 0, ctx->stream, y, n, box, keys, vals);
	NBH_LAUNCHED();
	uint32_t* perm = nullptr;
	NBH_TRY(nbh::sort_pairs(ctx, keys, vals, keys + n, vals + n, n, 30, counts, scanp, &perm));
	hipLaunchKernelGGL(k_sorted_copy, dim3(l.gbox), dim3(256), 0, ctx->stream, y, mass, perm, sorted, n);
	NBH_LAUNCHED();
	{
		nbh::scoped_timer timer(timed ? ctx : nullptr, "direct_fp32");
		hipLaunchKernelGGL(k_direct_fp32, dim3(static_cast<unsigned>(l.gx), static_cast<unsigned>(l.nsplit)),
		                   dim3(kF32Threads), 0, ctx->stream, sorted, part, n, l.jchunk);
		NBH_LAUNCHED();
	}
	hipLaunchKernelGGL(k_direct_fp32_finish, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0,
	                   ctx->stream, y, perm, part, f, n, static_cast<int>(l.nsplit));
	NBH_LAUNCHED();
	if(precision == NBH_FP32_CORR)
	{
		hipLaunchKernelGGL(k_direct_window, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, ctx->stream,
		                   sorted, perm, f, n, kCorrWindow);
		NBH_LAUNCHED();
	}
	return NBH_OK;
}

}  // namespace

extern "C" nbh_status nbh_direct_sample(nbh_ctx* ctx, const double* y, const double* mass, size_t n,
                                        const uint32_t* idx, size_t m, double* acc)
{
	NBH_TRY(nbh::enter(ctx));
	if(m == 0 || n == 0) { return NBH_OK; }
	NBH_CHECK(y != nullptr && mass != nullptr && idx != nullptr && acc != nullptr, "null buffer");
	hipLaunchKernelGGL(k_direct_sample64, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, ctx->stream,
	                   y, mass, n, idx, m, acc);
	NBH_LAUNCHED();
	return NBH_OK;
}

extern "C" nbh_status nbh_fcompute_direct(nbh_ctx* ctx, const double* y, const double* mass,
                                          double* f, size_t n, nbh_precision precision,
                                          int block_size)
{
	NBH_TRY(nbh::enter(ctx));
	if(n == 0) { return NBH_OK; }
	NBH_CHECK(y != nullptr && mass != nullptr && f != nullptr, "null buffer");
	NBH_CHECK(y != f, "y and f must not alias");
	NBH_CHECK(n < (static_cast<size_t>(1) << 31), "n = %zu too large", n);
	if(precision == NBH_FP64 || precision == NBH_FP64_BLOCK)
	{
		int bs = block_size == 0 ? 256 : block_size;
		NBH_CHECK(bs >= 64 && bs <= 1024 && bs % 64 == 0,
		          "block_size %d: must be a multiple of 64 in [64, 1024]", block_size);
		// two targets per work-item when the grid still holds two waves per SIMD
		const size_t simds = static_cast<size_t>(ctx->cu_count > 0 ? ctx->cu_count : 256) * 4;
		const int tgt = n >= 2 * 64 * 2 * simds ? 2 : 1;
		unsigned grid = static_cast<unsigned>((n + tgt * bs - 1) / (tgt * bs));
		nbh::scoped_timer timer(ctx, precision == NBH_FP64 ? "direct_fp64" : "direct_fp64_block");
		const size_t lds = bs * sizeof(double4);
		if(precision == NBH_FP64)
		{
			if(tgt == 2)
			{
				hipLaunchKernelGGL((k_direct_fp64<kLawSimple, 2>), dim3(grid), dim3(bs), lds, ctx->stream, y, mass, f, n);
			}
			else
			{
				hipLaunchKernelGGL((k_direct_fp64<kLawSimple, 1>), dim3(grid), dim3(bs), lds, ctx->stream, y, mass, f, n);
			}
		}
		else
		{
			if(tgt == 2)
			{
				hipLaunchKernelGGL((k_direct_fp64<kLawBlock, 2>), dim3(grid), dim3(bs), lds, ctx->stream, y, mass, f, n);
			}
			else
			{
				hipLaunchKernelGGL((k_direct_fp64<kLawBlock, 1>), dim3(grid), dim3(bs), lds, ctx->stream, y, mass, f, n);
			}
		}
		NBH_LAUNCHED();
		return NBH_OK;
	}
	NBH_CHECK(precision == NBH_FP32 || precision == NBH_FP32_CORR, "unknown precision %d",
	          static_cast<int>(precision));
	const fp32_layout lay = fp32_plan(ctx, n);
	void* sp = nullptr;
	NBH_TRY(nbh::scratch(ctx, lay.bytes, &sp));
	return fp32_run(ctx, y, mass, f, n, precision, lay, static_cast<char*>(sp), true);
}

// ---------------------------------------------------------------------------
// nbh_fcompute_direct_noalloc: the same computation with its device scratch in
// a fixed array of this library, for callers that cannot allocate.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocBytes = static_cast<size_t>(64) << 20;
constexpr size_t kNoAllocMaxBodies = 131072;
constexpr int kNoAllocDevices = 64;

}  // namespace

// One copy per device, held from the loading of this library until the process
// ends.  The name is a host-side handle: kernels get the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kNoAllocDevices] = {};

// Device address of `scratch` on ctx's device (current after nbh::enter),
// looked up on the first call for that device.
nbh_status noalloc_scratch(nbh_ctx* ctx, char** out)
{
	NBH_CHECK(ctx->device >= 0 && ctx->device < kNoAllocDevices, "device %d out of range", ctx->device);
	std::lock_guard<std::mutex> lock(g_noalloc_mutex);
	if(g_noalloc_ptr[ctx->device] == nullptr)
	{
		void* p = nullptr;
		NBH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		NBH_CHECK(p != nullptr, "no device address for the scratch array");
		g_noalloc_ptr[ctx->device] = p;
	}
	*out = static_cast<char*>(g_noalloc_ptr[ctx->device]);
	return NBH_OK;
}

}  // namespace

extern "C" nbh_status nbh_fcompute_direct_noalloc(nbh_ctx* ctx, const double* y, const double* mass, double* f,
                                                  size_t n, nbh_precision precision, int block_size)
{
	NBH_TRY(nbh::enter(ctx));
	if(n == 0) { return NBH_OK; }
	NBH_CHECK(y != nullptr && mass != nullptr && f != nullptr, "null buffer");
	NBH_CHECK(y != f, "y and f must not alias");
	NBH_CHECK(n <= kNoAllocMaxBodies, "n = %zu exceeds the %zu bodies of the allocation-free direct sum", n,
	          kNoAllocMaxBodies);
	if(precision == NBH_FP64 || precision == NBH_FP64_BLOCK)
	{
		// the FP64 kernels use no scratch: the allocating entry point is allocation-free here
		return nbh_fcompute_direct(ctx, y, mass, f, n, precision, block_size);
	}
	NBH_CHECK(precision == NBH_FP32 || precision == NBH_FP32_CORR, "unknown precision %d",
	          static_cast<int>(precision));
	const fp32_layout lay = fp32_plan(ctx, n);
	NBH_CHECK(lay.bytes <= kNoAllocBytes, "n = %zu needs %zu scratch bytes, more than the %zu of the fixed array", n,
	          lay.bytes, kNoAllocBytes);
	char* base = nullptr;
	NBH_TRY(noalloc_scratch(ctx, &base));
	return fp32_run(ctx, y, mass, f, n, precision, lay, base, false);
}
