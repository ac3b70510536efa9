// nbh_ah.hip -- Ahmad-Cohen neighbour scheme on gfx950 (the `hip_ah` engine).
//
// Reference: nbody_engine_ah (nbody/nbody_engine_ah.cpp:16-160).  Every
// full_recompute_rate-th step a full pass sums every pair in FP64, classifies
// body2 as a neighbour of body1 when |r1 - r2|^2 < max_dist^2 or |F12|^2 >
// min_force^2, and freezes the sum of the other bodies' forces (the far force);
// the steps in between sum the frozen far force and the neighbours' forces at
// the current positions.  The pair force is nbody_data::force
// (nbody/nbody_data.cpp:35-44) operation for operation, and every target sums
// its pairs in the reference's order (sources in index order in the full pass,
// list order -- also index order -- in the sparse pass), so the results equal
// the CPU restatement bit for bit.
//
// Layout in HBM: far force univ[3][n] (F, not F/m), neighbour counts [n],
// CSR offsets [n + 1] (uint64: lists can hold more than 2^32 entries in all),
// the lists themselves (uint32 body numbers), grown to the largest total seen.
//
// Full pass: one target per work-item, sources staged in LDS as double4 tiles
// and read back as wave-wide broadcasts.  It runs twice: the first launch sums
// the forces and counts the neighbours, a device scan turns the counts into
// offsets (its total is the one value read back to the host, to size the
// lists), and the second launch repeats the classification -- the same
// arithmetic, so the same decisions -- and writes the lists.  Sparse pass: one
// target per work-item gathering its neighbours' records.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "nbh_common.h"
#include "nbh_f64.h"

struct nbh_ah {
	nbh_ctx*  ctx = nullptr;
	size_t    n = 0;
	double*   univ = nullptr;     // [3][n]
	uint32_t* count = nullptr;    // [n]
	uint64_t* offs = nullptr;     // [n + 1]
	uint64_t* part = nullptr;     // [kAhParts + 1] scan partials
	uint32_t* list = nullptr;     // [cap]
	size_t    cap = 0;
	size_t    total = 0;
	bool      valid = false;      // a full pass has run
};

namespace {

constexpr double kMinDistance = 1e-8;   // nbody/nbtype.h:79
constexpr int kAhThreads = 256;
constexpr int kAhParts = 1024;          // chunks of the offset scan

// nbody_data::force (nbody/nbody_data.cpp:35-44) for target (x1, y1, z1) with
// gm1 = -m1 and the source record s = (x, y, z, m); d2 = |r1 - r2|^2 before the
// clamp (vertex3::norm, vertex.h:151-154).  kFast: the bare quotient and root,
// valid for operands in the ranges of coord_in_range / mass_in_range.
template <bool kFast>
__device__ inline void ah_force(double x1, double y1, double z1, double gm1, const double4& s, double& fx,
                                double& fy, double& fz, double& d2)
{
	const double dx = x1 - s.x;
	const double dy = y1 - s.y;
	const double dz = z1 - s.z;
	d2 = dx * dx + dy * dy + dz * dz;
	double r2 = d2;
	double coeff;
	if constexpr(kFast)
	{
		r2 = fmax(r2, kMinDistance);
		coeff = div_rn(gm1 * s.w, r2 * sqrt_rn(r2));
	}
	else
	{
		if(r2 < kMinDistance)
		{
			r2 = kMinDistance;
		}
		coeff = (gm1 * s.w) / (r2 * sqrt(r2));
	}
	fx = dx * coeff;
	fy = dy * coeff;
	fz = dz * coeff;
}

// Full pass over all sources for one target per work-item.  kFill = false:
// forces, far force and neighbour counts; kFill = true: the neighbour lists at
// offs (same classification arithmetic as the counting launch).
template <bool kFill>
__global__ void __launch_bounds__(kAhThreads)
k_ah_full(const double* __restrict__ y, const double* __restrict__ mass, size_t n, double dist2, double force2,
          double* __restrict__ f, double* __restrict__ univ, uint32_t* __restrict__ count,
          const uint64_t* __restrict__ offs, uint32_t* __restrict__ list)
{
	__shared__ double4 src[kAhThreads];
	const size_t i = static_cast<size_t>(blockIdx.x) * kAhThreads + threadIdx.x;
	const bool active = i < n;
	const double x1 = active ? y[i] : 0.0;
	const double y1 = active ? y[n + i] : 0.0;
	const double z1 = active ? y[2 * n + i] : 0.0;
	const double m1 = active ? mass[i] : 1.0;
	const double gm1 = -1.0 * m1;     // (-GravityConst * mass1), G = 1
	const bool safe = coord_in_range(x1) && coord_in_range(y1) && coord_in_range(z1) && (!active || mass_in_range(m1));
	double tx = 0.0, ty = 0.0, tz = 0.0;   // all pairs
	double ux = 0.0, uy = 0.0, uz = 0.0;   // far pairs
	uint32_t nb = 0;
	uint64_t pos = (kFill && active) ? offs[i] : 0;
	const uint64_t end = (kFill && active) ? offs[i + 1] : 0;
	for(size_t base = 0; base < n; base += kAhThreads)
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
		const int cnt = n - base < static_cast<size_t>(kAhThreads) ? static_cast<int>(n - base) : kAhThreads;
		if(!active)
		{
			continue;
		}
		for(int jj = 0; jj < cnt; ++jj)
		{
			const size_t b2 = base + jj;
			if(b2 == i)
			{
				continue;    // nbody_engine_ah.cpp:81-84
			}
			double fx, fy, fz, d2;
			if(fast) { ah_force<true>(x1, y1, z1, gm1, src[jj], fx, fy, fz, d2); }
			else { ah_force<false>(x1, y1, z1, gm1, src[jj], fx, fy, fz, d2); }
			const bool near = d2 < dist2 || (fx * fx + fy * fy + fz * fz) > force2;
			if constexpr(kFill)
			{
				if(near && pos < end)
				{
					list[pos++] = static_cast<uint32_t>(b2);
				}
			}
			else
			{
				if(near)
				{
					++nb;
				}
				else
				{
					ux += fx;
					uy += fy;
					uz += fz;
				}
				tx += fx;
				ty += fy;
				tz += fz;
			}
		}
	}
	if constexpr(!kFill)
	{
		if(active)
		{
			f[i] = y[3 * n + i];
			f[n + i] = y[4 * n + i];
			f[2 * n + i] = y[5 * n + i];
			f[3 * n + i] = tx / m1;
			f[4 * n + i] = ty / m1;
			f[5 * n + i] = tz / m1;
			univ[i] = ux;
			univ[n + i] = uy;
			univ[2 * n + i] = uz;
			count[i] = nb;
		}
	}
}

// Sparse pass (nbody_engine_ah.cpp:116-160): far force + neighbours in list order.
__global__ void __launch_bounds__(kAhThreads)
k_ah_sparse(const double* __restrict__ y, const double* __restrict__ mass, size_t n,
            const double* __restrict__ univ, const uint64_t* __restrict__ offs, const uint32_t* __restrict__ list,
            double* __restrict__ f)
{
	const size_t i = static_cast<size_t>(blockIdx.x) * kAhThreads + threadIdx.x;
	if(i >= n)
	{
		return;
	}
	const double x1 = y[i];
	const double y1 = y[n + i];
	const double z1 = y[2 * n + i];
	const double m1 = mass[i];
	const double gm1 = -1.0 * m1;
	const bool safe = coord_in_range(x1) && coord_in_range(y1) && coord_in_range(z1) && mass_in_range(m1);
	double tx = univ[i], ty = univ[n + i], tz = univ[2 * n + i];
	const uint64_t e = offs[i + 1];
	for(uint64_t k = offs[i]; k < e; ++k)
	{
		const uint32_t b2 = list[k];
		const double4 s = make_double4(y[b2], y[n + b2], y[2 * n + b2], mass[b2]);
		double fx, fy, fz, d2;
		if(safe && coord_in_range(s.x) && coord_in_range(s.y) && coord_in_range(s.z) && mass_in_range(s.w))
		{
			ah_force<true>(x1, y1, z1, gm1, s, fx, fy, fz, d2);
		}
		else
		{
			ah_force<false>(x1, y1, z1, gm1, s, fx, fy, fz, d2);
		}
		tx += fx;
		ty += fy;
		tz += fz;
	}
	f[i] = y[3 * n + i];
	f[n + i] = y[4 * n + i];
	f[2 * n + i] = y[5 * n + i];
	f[3 * n + i] = tx / m1;
	f[4 * n + i] = ty / m1;
	f[5 * n + i] = tz / m1;
}

// Exclusive scan of the n neighbour counts into 64-bit offsets (offs[n] = the
// total): chunk sums, one work-group scanning the <= kAhParts chunk sums, then
// every chunk from its offset.
__device__ inline uint64_t block_sum(uint64_t v, uint64_t* sh)
{
	for(int off = 32; off > 0; off >>= 1)
	{
		v += __shfl_xor(v, off, 64);
	}
	const int w = threadIdx.x >> 6;
	__syncthreads();
	if((threadIdx.x & 63) == 0) { sh[w] = v; }
	__syncthreads();
	uint64_t t = 0;
	for(int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) { t += sh[k]; }
	return t;
}

__global__ void __launch_bounds__(kAhThreads)
k_ah_scan_reduce(const uint32_t* __restrict__ c, size_t n, size_t chunk, uint64_t* __restrict__ part)
{
	__shared__ uint64_t sh[kAhThreads / 64];
	const size_t b = static_cast<size_t>(blockIdx.x) * chunk;
	const size_t e = b + chunk < n ? b + chunk : n;
	uint64_t v = 0;
	for(size_t k = b + threadIdx.x; k < e; k += kAhThreads) { v += c[k]; }
	v = block_sum(v, sh);
	if(threadIdx.x == 0) { part[blockIdx.x] = v; }
}

__global__ void __launch_bounds__(kAhParts)
k_ah_scan_parts(uint64_t* __restrict__ part, int g, uint64_t* __restrict__ offs, size_t n)
{
	__shared__ uint64_t sh[kAhParts];
	const int t = threadIdx.x;
	const uint64_t mine = t < g ? part[t] : 0;
	sh[t] = mine;
	__syncthreads();
	for(int off = 1; off < kAhParts; off <<= 1)
	{
		const uint64_t add = t >= off ? sh[t - off] : 0;
		__syncthreads();
		sh[t] += add;
		__syncthreads();
	}
	if(t < g) { part[t] = sh[t] - mine; }          // exclusive
	if(t == kAhParts - 1)
	{
		part[kAhParts] = sh[t];                    // total
		offs[n] = sh[t];
	}
}

// Work-item k of a chunk scans its run of `per` consecutive counts.
__global__ void __launch_bounds__(kAhThreads)
k_ah_scan_apply(const uint32_t* __restrict__ c, size_t n, size_t chunk, const uint64_t* __restrict__ part,
                uint64_t* __restrict__ offs)
{
	__shared__ uint64_t sh[kAhThreads];
	const size_t b = static_cast<size_t>(blockIdx.x) * chunk;
	const size_t e = b + chunk < n ? b + chunk : n;
	const size_t per = (chunk + kAhThreads - 1) / kAhThreads;
	const size_t rb = b + threadIdx.x * per;
	const size_t re = rb + per < e ? rb + per : e;
	uint64_t v = 0;
	for(size_t k = rb; k < re; ++k) { v += c[k]; }
	const int t = threadIdx.x;
	sh[t] = v;
	__syncthreads();
	for(int off = 1; off < kAhThreads; off <<= 1)
	{
		const uint64_t add = t >= off ? sh[t - off] : 0;
		__syncthreads();
		sh[t] += add;
		__syncthreads();
	}
	uint64_t run = part[blockIdx.x] + sh[t] - v;
	for(size_t k = rb; k < re; ++k)
	{
		offs[k] = run;
		run += c[k];
	}
}

void ah_release(nbh_ah* a)
{
	void* ptrs[] = { a->univ, a->count, a->offs, a->part, a->list };
	for(void* p : ptrs)
	{
		if(p != nullptr) { (void)hipFree(p); }
	}
	delete a;
}

}  // namespace

extern "C" {

nbh_status nbh_ah_create(nbh_ctx* ctx, size_t n, nbh_ah** out)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(out != nullptr, "null output pointer");
	*out = nullptr;
	NBH_CHECK(n > 0 && n < (static_cast<size_t>(1) << 31), "body count %zu out of range", n);
	nbh_ah* a = new nbh_ah();
	a->ctx = ctx;
	a->n = n;
	hipError_t e = hipMalloc(reinterpret_cast<void**>(&a->univ), 3 * n * sizeof(double));
	if(e == hipSuccess) { e = hipMalloc(reinterpret_cast<void**>(&a->count), n * sizeof(uint32_t)); }
	if(e == hipSuccess) { e = hipMalloc(reinterpret_cast<void**>(&a->offs), (n + 1) * sizeof(uint64_t)); }
	if(e == hipSuccess) { e = hipMalloc(reinterpret_cast<void**>(&a->part), (kAhParts + 1) * sizeof(uint64_t)); }
	if(e != hipSuccess)
	{
		ah_release(a);
		return nbh::hip_fail(e, "neighbour-scheme buffers", __FILE__, __LINE__);
	}
	*out = a;
	return NBH_OK;
}

nbh_status nbh_ah_destroy(nbh_ah* a)
{
	if(a == nullptr)
	{
		return NBH_OK;
	}
	NBH_TRY(nbh::enter(a->ctx));
	(void)hipStreamSynchronize(a->ctx->stream);
	ah_release(a);
	return NBH_OK;
}

nbh_status nbh_ah_fcompute_full(nbh_ctx* ctx, nbh_ah* a, const double* y, const double* mass, double* f,
                                double max_dist, double min_force)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(a != nullptr && a->ctx == ctx, "neighbour scheme does not belong to this context");
	NBH_CHECK(y != nullptr && mass != nullptr && f != nullptr, "null buffer");
	NBH_CHECK(y != f, "y and f must not alias");
	const size_t n = a->n;
	const double dist2 = max_dist * max_dist;      // nbody_engine_ah.cpp:3-9
	const double force2 = min_force * min_force;
	const unsigned grid = static_cast<unsigned>((n + kAhThreads - 1) / kAhThreads);
	nbh::scoped_timer timer(ctx, "ah_full");
	a->valid = false;
	hipLaunchKernelGGL(k_ah_full<false>, dim3(grid), dim3(kAhThreads), 0, ctx->stream, y, mass, n, dist2, force2, f,
	                   a->univ, a->count, a->offs, a->list);
	NBH_LAUNCHED();
	// offsets of the lists
	const size_t g = std::min(static_cast<size_t>(kAhParts), (n + 1023) / 1024);
	const size_t chunk = (n + g - 1) / g;
	hipLaunchKernelGGL(k_ah_scan_reduce, dim3(static_cast<unsigned>(g)), dim3(kAhThreads), 0, ctx->stream, a->count,
	                   n, chunk, a->part);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_ah_scan_parts, dim3(1), dim3(kAhParts), 0, ctx->stream, a->part, static_cast<int>(g), a->offs,
	                   n);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_ah_scan_apply, dim3(static_cast<unsigned>(g)), dim3(kAhThreads), 0, ctx->stream, a->count, n,
	                   chunk, a->part, a->offs);
	NBH_LAUNCHED();
	NBH_TRY(nbh::pinned(ctx, sizeof(uint64_t)));
	NBH_HIP(hipMemcpyAsync(ctx->pinned, a->part + kAhParts, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	uint64_t total = 0;
	std::memcpy(&total, ctx->pinned, sizeof(total));
	if(total > a->cap)
	{
		if(a->list != nullptr)
		{
			NBH_HIP(hipFree(a->list));
			a->list = nullptr;
			a->cap = 0;
		}
		const size_t cap = total + total / 8 + 1024;
		NBH_HIP(hipMalloc(reinterpret_cast<void**>(&a->list), cap * sizeof(uint32_t)));
		a->cap = cap;
	}
	a->total = total;
	if(total > 0)
	{
		hipLaunchKernelGGL(k_ah_full<true>, dim3(grid), dim3(kAhThreads), 0, ctx->stream, y, mass, n, dist2, force2,
		                   f, a->univ, a->count, a->offs, a->list);
		NBH_LAUNCHED();
	}
	a->valid = true;
	return NBH_OK;
}

nbh_status nbh_ah_fcompute_sparse(nbh_ctx* ctx, nbh_ah* a, const double* y, const double* mass, double* f)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(a != nullptr && a->ctx == ctx, "neighbour scheme does not belong to this context");
	NBH_CHECK(a->valid, "no full pass has built the neighbour lists");
	NBH_CHECK(y != nullptr && mass != nullptr && f != nullptr, "null buffer");
	NBH_CHECK(y != f, "y and f must not alias");
	const size_t n = a->n;
	nbh::scoped_timer timer(ctx, "ah_sparse");
	hipLaunchKernelGGL(k_ah_sparse, dim3(static_cast<unsigned>((n + kAhThreads - 1) / kAhThreads)), dim3(kAhThreads),
	                   0, ctx->stream, y, mass, n, a->univ, a->offs, a->list, f);
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status nbh_ah_lists(nbh_ctx* ctx, const nbh_ah* a, size_t* total, uint64_t* offs, uint32_t* list,
                        double* univ)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(a != nullptr && a->ctx == ctx, "neighbour scheme does not belong to this context");
	NBH_CHECK(total != nullptr, "null output");
	NBH_CHECK(a->valid, "no full pass has built the neighbour lists");
	*total = a->total;
	if(offs != nullptr)
	{
		NBH_HIP(hipMemcpyAsync(offs, a->offs, (a->n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
	}
	if(list != nullptr && a->total > 0)
	{
		NBH_HIP(hipMemcpyAsync(list, a->list, a->total * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
	}
	if(univ != nullptr)
	{
		NBH_HIP(hipMemcpyAsync(univ, a->univ, 3 * a->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
	}
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

}  // extern "C"
