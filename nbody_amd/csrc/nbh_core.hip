// nbh_core.hip -- context, memory, FP64 vector ops and conservation checks.
//
// Vector ops are HBM-bound streams: each work-item handles two consecutive
// doubles (16-byte loads) in a grid-stride loop sized to fill all CUs.
// Built with -ffp-contract=off; every fused multiply-add below is an explicit
// fma(), every other operation rounds where it is written.

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "nbh_common.h"

namespace {
thread_local std::string g_last_error;
constexpr int kVecThreads = 256;
constexpr int kMaxTerms = 48;   // terms per fused launch; longer lists are chained
}

namespace nbh {

void set_error(const char* fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_last_error = buf;
}

nbh_status hip_fail(hipError_t e, const char* what, const char* file, int line)
{
	set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
	return e == hipErrorOutOfMemory ? NBH_ERR_NOMEM : NBH_ERR_HIP;
}

nbh_status enter(nbh_ctx* ctx)
{
	NBH_CHECK(ctx != nullptr, "null context");
	NBH_HIP(hipSetDevice(ctx->device));
	return NBH_OK;
}

nbh_status scratch(nbh_ctx* ctx, size_t bytes, void** out)
{
	if(bytes > ctx->scratch_bytes)
	{
		if(ctx->scratch != nullptr)
		{
			NBH_HIP(hipStreamSynchronize(ctx->stream));
			NBH_HIP(hipFree(ctx->scratch));
			ctx->scratch = nullptr;
			ctx->scratch_bytes = 0;
		}
		size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
		NBH_HIP(hipMalloc(&ctx->scratch, want));
		ctx->scratch_bytes = want;
	}
	*out = ctx->scratch;
	return NBH_OK;
}

nbh_status pinned(nbh_ctx* ctx, size_t bytes)
{
	if(bytes > ctx->pinned_bytes)
	{
		if(ctx->pinned != nullptr)
		{
			NBH_HIP(hipHostFree(ctx->pinned));
			ctx->pinned = nullptr;
			ctx->pinned_bytes = 0;
		}
		NBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->pinned), bytes, hipHostMallocDefault));
		ctx->pinned_bytes = bytes;
	}
	return NBH_OK;
}

namespace {

hipEvent_t take_event(nbh_ctx* ctx)
{
	if(!ctx->event_pool.empty())
	{
		hipEvent_t e = ctx->event_pool.back();
		ctx->event_pool.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	if(hipEventCreate(&e) != hipSuccess)
	{
		(void)hipGetLastError();
		return nullptr;
	}
	return e;
}

// Sum the elapsed time of every pending pair (the stream must be idle past them).
void fold_timers(nbh_ctx* ctx)
{
	for(auto& kv : ctx->timers)
	{
		for(auto& pr : kv.second.pending)
		{
			float ms = 0.0f;
			if(hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess)
			{
				kv.second.total_ms += ms;
				kv.second.launches += 1;
			}
			else
			{
				(void)hipGetLastError();
			}
			ctx->event_pool.push_back(pr.first);
			ctx->event_pool.push_back(pr.second);
		}
		kv.second.pending.clear();
	}
}

}  // namespace

scoped_timer::scoped_timer(nbh_ctx* ctx, const char* name) : m_ctx(ctx), m_name(name), m_start(nullptr)
{
	if(ctx != nullptr && ctx->timing)
	{
		m_start = take_event(ctx);
		if(m_start != nullptr && hipEventRecord(m_start, ctx->stream) != hipSuccess)
		{
			(void)hipGetLastError();
			ctx->event_pool.push_back(m_start);
			m_start = nullptr;
		}
	}
}

scoped_timer::~scoped_timer()
{
	if(m_start == nullptr)
	{
		return;
	}
	hipEvent_t stop = take_event(m_ctx);
	if(stop == nullptr || hipEventRecord(stop, m_ctx->stream) != hipSuccess)
	{
		(void)hipGetLastError();
		m_ctx->event_pool.push_back(m_start);
		if(stop != nullptr) { m_ctx->event_pool.push_back(stop); }
		return;
	}
	auto& cls = m_ctx->timers[m_name];
	cls.pending.emplace_back(m_start, stop);
	if(cls.pending.size() >= 4096 && hipStreamSynchronize(m_ctx->stream) == hipSuccess)
	{
		fold_timers(m_ctx);
	}
}

}  // namespace nbh

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

struct term_list {
	const double* src[kMaxTerms];
	double        coef[kMaxTerms];
};

__global__ void __launch_bounds__(kVecThreads)
k_fill(double* __restrict__ a, double v, size_t n)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		a[i] = v;
	}
}

// Streams larger than the 256 MB last-level cache (MALL) are read and written
// with nontemporal hints: measured 5.0-5.3 -> 5.8-5.9 TB/s for 12 terms over
// 6 x 2^22 doubles; smaller ones keep the cache, which the next force
// evaluation reads from.
constexpr size_t kStreamBytes = static_cast<size_t>(256) << 20;

template <bool kNT>
__device__ inline double vld(const double* p)
{
	if constexpr(kNT) { return __builtin_nontemporal_load(p); }
	else { return *p; }
}

template <bool kNT>
__device__ inline void vst(double* p, double v)
{
	if constexpr(kNT) { __builtin_nontemporal_store(v, p); }
	else { *p = v; }
}

// acc = init (or 0); acc = fma(src[t], coef[t], acc) for t in order; a = acc.
// Two doubles per work-item per pass; `init` and any src may alias a.
template <bool kNT>
__global__ void __launch_bounds__(kVecThreads)
k_fmaddn(double* a, const double* init, term_list terms, int k, size_t n)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	size_t npair = n / 2;
	for(size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < npair; p += stride)
	{
		size_t i = 2 * p;
		double2 acc = make_double2(0.0, 0.0);
		if(init != nullptr)
		{
			if constexpr(kNT) { acc = make_double2(vld<true>(init + i), vld<true>(init + i + 1)); }
			else { acc = *reinterpret_cast<const double2*>(init + i); }
		}
		for(int t = 0; t < k; ++t)
		{
			double2 s;
			if constexpr(kNT) { s = make_double2(vld<true>(terms.src[t] + i), vld<true>(terms.src[t] + i + 1)); }
			else { s = *reinterpret_cast<const double2*>(terms.src[t] + i); }
			double c = terms.coef[t];
			acc.x = fma(s.x, c, acc.x);
			acc.y = fma(s.y, c, acc.y);
		}
		if constexpr(kNT)
		{
			vst<true>(a + i, acc.x);
			vst<true>(a + i + 1, acc.y);
		}
		else
		{
			*reinterpret_cast<double2*>(a + i) = acc;
		}
	}
	// odd tail element
	if((n & 1) && blockIdx.x == 0 && threadIdx.x == 0)
	{
		size_t i = n - 1;
		double acc = init != nullptr ? init[i] : 0.0;
		for(int t = 0; t < k; ++t)
		{
			acc = fma(terms.src[t][i], terms.coef[t], acc);
		}
		a[i] = acc;
	}
}

// Scalar variant for pointers that are not 16-byte aligned.
__global__ void __launch_bounds__(kVecThreads)
k_fmaddn_scalar(double* a, const double* init, term_list terms, int k, size_t n)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		double acc = init != nullptr ? init[i] : 0.0;
		for(int t = 0; t < k; ++t)
		{
			acc = fma(terms.src[t][i], terms.coef[t], acc);
		}
		a[i] = acc;
	}
}

// Kahan step of nbody/summation.h:7-14; the product is rounded on its own.
__device__ inline double kahan(double a, double term, double& corr)
{
	double corrected = term - corr;
	double new_sum = a + corrected;
	corr = (new_sum - a) - corrected;
	return new_sum;
}

template <bool kNT>
__global__ void __launch_bounds__(kVecThreads)
k_fmaddn_corr(double* a, double* corr, term_list terms, int k, size_t n)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		double s = vld<kNT>(a + i);
		double c = vld<kNT>(corr + i);
		for(int t = 0; t < k; ++t)
		{
			double term = vld<kNT>(terms.src[t] + i) * terms.coef[t];
			s = kahan(s, term, c);
		}
		vst<kNT>(a + i, s);
		vst<kNT>(corr + i, c);
	}
}

// max |a| with NaN ignored (the reference's `v > result` never takes a NaN).
__device__ inline double max_abs_step(double m, double v)
{
	double av = fabs(v);
	return av > m ? av : m;
}

template<int BS>
__device__ inline double block_max(double v, double* sh)
{
	for(int off = 32; off > 0; off >>= 1)
	{
		double o = __shfl_down(v, off, 64);
		v = o > v ? o : v;
	}
	int lane = threadIdx.x & 63;
	int wave = threadIdx.x >> 6;
	if(lane == 0) { sh[wave] = v; }
	__syncthreads();
	if(threadIdx.x == 0)
	{
		for(int w = 1; w < BS / 64; ++w) { v = sh[w] > v ? sh[w] : v; }
	}
	return v;
}

__global__ void __launch_bounds__(kVecThreads)
k_maxabs_partial(const double* __restrict__ a, size_t n, double* __restrict__ part)
{
	__shared__ double sh[kVecThreads / 64];
	double m = 0.0;
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		m = max_abs_step(m, a[i]);
	}
	m = block_max<kVecThreads>(m, sh);
	if(threadIdx.x == 0) { part[blockIdx.x] = m; }
}

// Final pass: max of the partials; a NaN in a[0] propagates as in
// nbody_engine_simple::fmaxabs (result seeded with fabs(a[0])).
__global__ void __launch_bounds__(kVecThreads)
k_maxabs_final(const double* __restrict__ part, int nparts, const double* __restrict__ a,
               double* __restrict__ out)
{
	__shared__ double sh[kVecThreads / 64];
	double m = 0.0;
	for(int i = threadIdx.x; i < nparts; i += blockDim.x)
	{
		m = part[i] > m ? part[i] : m;
	}
	m = block_max<kVecThreads>(m, sh);
	if(threadIdx.x == 0)
	{
		double a0 = a[0];
		*out = (a0 != a0) ? a0 : m;
	}
}

// nbody_engine_simple.cpp:175-199 on the first 3*nb entries.
__global__ void __launch_bounds__(kVecThreads)
k_clamp(double* __restrict__ y, size_t n3, double b)
{
	double diam = 2.0 * b;
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n3; i += stride)
	{
		double r = y[i];
		if(r > +b) { r -= diam; }
		if(r < -b) { r += diam; }
		y[i] = r;
	}
}

// ---- compensated reductions for the conservation checks -------------------
// Each running value is a (sum, err) pair updated with Knuth's TwoSum, so the
// result is accurate to a few ulps independent of the summation order.
struct csum {
	double s;
	double e;
};

__device__ inline void cadd(csum& a, double v)
{
	double s = a.s + v;
	double bp = s - a.s;
	double err = (a.s - (s - bp)) + (v - bp);
	a.s = s;
	a.e += err;
}

__device__ inline void cmerge(csum& a, const csum& b)
{
	cadd(a, b.s);
	a.e += b.e;
}

constexpr int kStatQ = 11;   // P(3) L(3) KE CM(3) M
constexpr int kStatThreads = 256;

template<int Q>
__device__ inline void block_reduce_csum(csum* v, csum (*sh)[Q])
{
	// wave-level tree, then across waves through LDS
	for(int q = 0; q < Q; ++q)
	{
		for(int off = 32; off > 0; off >>= 1)
		{
			csum o;
			o.s = __shfl_down(v[q].s, off, 64);
			o.e = __shfl_down(v[q].e, off, 64);
			cmerge(v[q], o);
		}
	}
	int lane = threadIdx.x & 63;
	int wave = threadIdx.x >> 6;
	if(lane == 0)
	{
		for(int q = 0; q < Q; ++q) { sh[wave][q] = v[q]; }
	}
	__syncthreads();
	if(threadIdx.x == 0)
	{
		for(int w = 1; w < static_cast<int>(blockDim.x) / 64; ++w)
		{
			for(int q = 0; q < Q; ++q) { cmerge(v[q], sh[w][q]); }
		}
	}
}

__global__ void __launch_bounds__(kStatThreads)
k_stats_linear(const double* __restrict__ y, const double* __restrict__ mass, size_t n,
               csum* __restrict__ part)
{
	__shared__ csum sh[kStatThreads / 64][kStatQ];
	csum acc[kStatQ];
	for(int q = 0; q < kStatQ; ++q) { acc[q].s = 0.0; acc[q].e = 0.0; }
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		double m = mass[i];
		double rx = y[i], ry = y[n + i], rz = y[2 * n + i];
		double vx = y[3 * n + i], vy = y[4 * n + i], vz = y[5 * n + i];
		// impulce_proxy / impulce_moment_proxy / kinetic_energy_proxy /
		// mass_center_proxy (nbody/summation_proxy.h:6-66)
		double px = vx * m, py = vy * m, pz = vz * m;
		cadd(acc[0], px);
		cadd(acc[1], py);
		cadd(acc[2], pz);
		cadd(acc[3], ry * pz - rz * py);
		cadd(acc[4], rz * px - rx * pz);
		cadd(acc[5], rx * py - ry * px);
		cadd(acc[6], (vx * vx + vy * vy + vz * vz) * m);
		cadd(acc[7], rx * m);
		cadd(acc[8], ry * m);
		cadd(acc[9], rz * m);
		cadd(acc[10], m);
	}
	block_reduce_csum<kStatQ>(acc, sh);
	if(threadIdx.x == 0)
	{
		for(int q = 0; q < kStatQ; ++q) { part[blockIdx.x * kStatQ + q] = acc[q]; }
	}
}

// Potential energy: per target i the sum over j != i of -(m_i m_j)/r, with
// pairs closer than sqrt(1e-8) contributing 0 (nbody/nbody_data.cpp:46-55).
constexpr int kPeTile = 256;

__global__ void __launch_bounds__(kStatThreads)
k_stats_pe(const double* __restrict__ y, const double* __restrict__ mass, size_t n,
           csum* __restrict__ part)
{
	__shared__ double sx[kPeTile], sy[kPeTile], sz[kPeTile], sm[kPeTile];
	__shared__ csum sh[kStatThreads / 64][1];
	size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	bool active = i < n;
	double x1 = active ? y[i] : 0.0;
	double y1 = active ? y[n + i] : 0.0;
	double z1 = active ? y[2 * n + i] : 0.0;
	double m1 = active ? mass[i] : 0.0;
	csum acc[1];
	acc[0].s = 0.0;
	acc[0].e = 0.0;
	for(size_t base = 0; base < n; base += kPeTile)
	{
		size_t j = base + threadIdx.x;
		__syncthreads();
		sx[threadIdx.x] = j < n ? y[j] : 0.0;
		sy[threadIdx.x] = j < n ? y[n + j] : 0.0;
		sz[threadIdx.x] = j < n ? y[2 * n + j] : 0.0;
		sm[threadIdx.x] = j < n ? mass[j] : 0.0;
		__syncthreads();
		size_t cnt = n - base < kPeTile ? n - base : kPeTile;
		if(active)
		{
			for(size_t jj = 0; jj < cnt; ++jj)
			{
				if(base + jj == i) { continue; }
				double dx = x1 - sx[jj];
				double dy = y1 - sy[jj];
				double dz = z1 - sz[jj];
				double r2 = dx * dx + dy * dy + dz * dz;
				if(r2 < 1e-8) { continue; }
				cadd(acc[0], -((1.0 * m1) * sm[jj]) / sqrt(r2));
			}
		}
	}
	block_reduce_csum<1>(acc, sh);
	if(threadIdx.x == 0) { part[blockIdx.x] = acc[0]; }
}

template<int Q>
__global__ void __launch_bounds__(kStatThreads)
k_csum_final(const csum* __restrict__ part, int nparts, double* __restrict__ out)
{
	__shared__ csum sh[kStatThreads / 64][Q];
	csum acc[Q];
	for(int q = 0; q < Q; ++q) { acc[q].s = 0.0; acc[q].e = 0.0; }
	for(int b = threadIdx.x; b < nparts; b += blockDim.x)
	{
		for(int q = 0; q < Q; ++q) { cmerge(acc[q], part[b * Q + q]); }
	}
	block_reduce_csum<Q>(acc, sh);
	if(threadIdx.x == 0)
	{
		for(int q = 0; q < Q; ++q) { out[q] = acc[q].s + acc[q].e; }
	}
}

bool aligned16(const void* p)
{
	return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* nbh_last_error(void)
{
	return g_last_error.c_str();
}

int nbh_version(void)
{
	return 1;
}

nbh_status nbh_device_count(int* count)
{
	NBH_CHECK(count != nullptr, "null count");
	int c = 0;
	hipError_t e = hipGetDeviceCount(&c);
	if(e != hipSuccess)
	{
		c = 0;
		(void)hipGetLastError();
	}
	*count = c;
	return NBH_OK;
}

nbh_status nbh_ctx_create(int device, nbh_ctx** out)
{
	NBH_CHECK(out != nullptr, "null output pointer");
	*out = nullptr;
	int count = 0;
	NBH_HIP(hipGetDeviceCount(&count));
	NBH_CHECK(device >= 0 && device < count, "device %d out of range (%d visible)", device, count);
	NBH_HIP(hipSetDevice(device));
	hipDeviceProp_t prop;
	NBH_HIP(hipGetDeviceProperties(&prop, device));
	nbh_ctx* ctx = new nbh_ctx();
	ctx->device = device;
	ctx->cu_count = prop.multiProcessorCount;
	std::snprintf(ctx->name, sizeof(ctx->name), "%s", prop.gcnArchName);
	hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
	if(e != hipSuccess)
	{
		delete ctx;
		return nbh::hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
	}
	*out = ctx;
	return NBH_OK;
}

nbh_status nbh_ctx_destroy(nbh_ctx* ctx)
{
	if(ctx == nullptr)
	{
		return NBH_OK;
	}
	NBH_TRY(nbh::enter(ctx));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	nbh::fold_timers(ctx);
	for(hipEvent_t e : ctx->event_pool) { (void)hipEventDestroy(e); }
	ctx->event_pool.clear();
	if(ctx->scratch != nullptr) { NBH_HIP(hipFree(ctx->scratch)); }
	if(ctx->pinned != nullptr) { NBH_HIP(hipHostFree(ctx->pinned)); }
	NBH_HIP(hipStreamDestroy(ctx->stream));
	delete ctx;
	return NBH_OK;
}

nbh_status nbh_ctx_synchronize(nbh_ctx* ctx)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

nbh_status nbh_timing_enable(nbh_ctx* ctx, int enable)
{
	NBH_TRY(nbh::enter(ctx));
	ctx->timing = enable != 0;
	return NBH_OK;
}

nbh_status nbh_timing_reset(nbh_ctx* ctx)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	nbh::fold_timers(ctx);
	ctx->timers.clear();
	return NBH_OK;
}

nbh_status nbh_timing_read(nbh_ctx* ctx, const char* kernel_class, double* total_ms, uint64_t* launches)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(kernel_class != nullptr, "null kernel class");
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	nbh::fold_timers(ctx);
	auto it = ctx->timers.find(kernel_class);
	if(total_ms != nullptr) { *total_ms = it == ctx->timers.end() ? 0.0 : it->second.total_ms; }
	if(launches != nullptr) { *launches = it == ctx->timers.end() ? 0 : it->second.launches; }
	return NBH_OK;
}

nbh_status nbh_ctx_info(nbh_ctx* ctx, char* name, size_t name_len, int* cu_count)
{
	NBH_CHECK(ctx != nullptr, "null context");
	if(name != nullptr && name_len > 0)
	{
		std::snprintf(name, name_len, "%s", ctx->name);
	}
	if(cu_count != nullptr)
	{
		*cu_count = ctx->cu_count;
	}
	return NBH_OK;
}

nbh_status nbh_malloc(nbh_ctx* ctx, size_t bytes, void** dptr)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(dptr != nullptr, "null output pointer");
	*dptr = nullptr;
	if(bytes == 0)
	{
		return NBH_OK;
	}
	NBH_HIP(hipMalloc(dptr, bytes));
	return NBH_OK;
}

nbh_status nbh_free(nbh_ctx* ctx, void* dptr)
{
	NBH_TRY(nbh::enter(ctx));
	if(dptr == nullptr)
	{
		return NBH_OK;
	}
	// Work queued on the stream may still read the buffer.
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	NBH_HIP(hipFree(dptr));
	return NBH_OK;
}

nbh_status nbh_memcpy_h2d(nbh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
	NBH_TRY(nbh::enter(ctx));
	if(bytes == 0) { return NBH_OK; }
	NBH_CHECK(dst != nullptr && src != nullptr, "null pointer");
	// The source is caller memory that may change as soon as we return.
	NBH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

nbh_status nbh_memcpy_d2h(nbh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
	NBH_TRY(nbh::enter(ctx));
	if(bytes == 0) { return NBH_OK; }
	NBH_CHECK(dst != nullptr && src != nullptr, "null pointer");
	NBH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

nbh_status nbh_memcpy_d2d(nbh_ctx* ctx, void* dst, const void* src, size_t bytes)
{
	NBH_TRY(nbh::enter(ctx));
	if(bytes == 0 || dst == src) { return NBH_OK; }
	NBH_CHECK(dst != nullptr && src != nullptr, "null pointer");
	NBH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
	return NBH_OK;
}

nbh_status nbh_fill(nbh_ctx* ctx, double* a, double value, size_t n)
{
	NBH_TRY(nbh::enter(ctx));
	if(n == 0) { return NBH_OK; }
	NBH_CHECK(a != nullptr, "null buffer");
	hipLaunchKernelGGL(k_fill, dim3(nbh::grid_for(ctx, n, kVecThreads)), dim3(kVecThreads), 0,
	                   ctx->stream, a, value, n);
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status nbh_fmaddn(nbh_ctx* ctx, double* a, const double* init,
                      const double* const* src, const double* coef, size_t k, size_t n)
{
	NBH_TRY(nbh::enter(ctx));
	if(n == 0) { return NBH_OK; }
	NBH_CHECK(a != nullptr, "null output buffer");
	NBH_CHECK(k == 0 || (src != nullptr && coef != nullptr), "null term list");
	for(size_t t = 0; t < k; ++t)
	{
		NBH_CHECK(src[t] != nullptr, "null source buffer %zu", t);
	}
	// Chain launches of at most kMaxTerms terms; each continues from `a`.
	nbh::scoped_timer timer(ctx, "fmaddn");
	size_t done = 0;
	const double* cur_init = init;
	do
	{
		size_t kk = k - done < static_cast<size_t>(kMaxTerms) ? k - done : kMaxTerms;
		term_list tl;
		std::memset(&tl, 0, sizeof(tl));
		bool al = aligned16(a) && (cur_init == nullptr || aligned16(cur_init));
		for(size_t t = 0; t < kk; ++t)
		{
			tl.src[t] = src[done + t];
			tl.coef[t] = coef[done + t];
			al = al && aligned16(tl.src[t]);
		}
		if(al && 8 * n * (kk + 2) > kStreamBytes)
		{
			hipLaunchKernelGGL(k_fmaddn<true>, dim3(nbh::grid_for(ctx, n / 2 + 1, kVecThreads)),
			                   dim3(kVecThreads), 0, ctx->stream, a, cur_init, tl,
			                   static_cast<int>(kk), n);
		}
		else if(al)
		{
			hipLaunchKernelGGL(k_fmaddn<false>, dim3(nbh::grid_for(ctx, n / 2 + 1, kVecThreads)),
			                   dim3(kVecThreads), 0, ctx->stream, a, cur_init, tl,
			                   static_cast<int>(kk), n);
		}
		else
		{
			hipLaunchKernelGGL(k_fmaddn_scalar, dim3(nbh::grid_for(ctx, n, kVecThreads)),
			                   dim3(kVecThreads), 0, ctx->stream, a, cur_init, tl,
			                   static_cast<int>(kk), n);
		}
		NBH_LAUNCHED();
		done += kk;
		cur_init = a;
	}
	while(done < k);
	return NBH_OK;
}

nbh_status nbh_fmaddn_corr(nbh_ctx* ctx, double* a, double* corr,
                           const double* const* src, const double* coef, size_t k, size_t n)
{
	NBH_TRY(nbh::enter(ctx));
	if(n == 0 || k == 0) { return NBH_OK; }
	NBH_CHECK(a != nullptr && corr != nullptr, "null output buffer");
	NBH_CHECK(src != nullptr && coef != nullptr, "null term list");
	nbh::scoped_timer timer(ctx, "fmaddn");
	size_t done = 0;
	while(done < k)
	{
		size_t kk = k - done < static_cast<size_t>(kMaxTerms) ? k - done : kMaxTerms;
		term_list tl;
		std::memset(&tl, 0, sizeof(tl));
		for(size_t t = 0; t < kk; ++t)
		{
			NBH_CHECK(src[done + t] != nullptr, "null source buffer %zu", done + t);
			tl.src[t] = src[done + t];
			tl.coef[t] = coef[done + t];
		}
		if(8 * n * (kk + 4) > kStreamBytes)
		{
			hipLaunchKernelGGL(k_fmaddn_corr<true>, dim3(nbh::grid_for(ctx, n, kVecThreads)),
			                   dim3(kVecThreads), 0, ctx->stream, a, corr, tl, static_cast<int>(kk), n);
		}
		else
		{
			hipLaunchKernelGGL(k_fmaddn_corr<false>, dim3(nbh::grid_for(ctx, n, kVecThreads)),
			                   dim3(kVecThreads), 0, ctx->stream, a, corr, tl, static_cast<int>(kk), n);
		}
		NBH_LAUNCHED();
		done += kk;
	}
	return NBH_OK;
}

nbh_status nbh_maxabs(nbh_ctx* ctx, const double* a, size_t n, double* result)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(result != nullptr, "null result pointer");
	if(n == 0)
	{
		*result = 0.0;
		return NBH_OK;
	}
	NBH_CHECK(a != nullptr, "null buffer");
	unsigned grid = nbh::grid_for(ctx, n, kVecThreads * 4);
	void* sp = nullptr;
	NBH_TRY(nbh::scratch(ctx, (grid + 1) * sizeof(double), &sp));
	double* part = static_cast<double*>(sp);
	NBH_TRY(nbh::pinned(ctx, 64));
	hipLaunchKernelGGL(k_maxabs_partial, dim3(grid), dim3(kVecThreads), 0, ctx->stream, a, n, part);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_maxabs_final, dim3(1), dim3(kVecThreads), 0, ctx->stream, part,
	                   static_cast<int>(grid), a, part + grid);
	NBH_LAUNCHED();
	NBH_HIP(hipMemcpyAsync(ctx->pinned, part + grid, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	*result = ctx->pinned[0];
	return NBH_OK;
}

nbh_status nbh_clamp(nbh_ctx* ctx, double* y, size_t nbodies, double b)
{
	NBH_TRY(nbh::enter(ctx));
	if(nbodies == 0) { return NBH_OK; }
	NBH_CHECK(y != nullptr, "null buffer");
	size_t n3 = 3 * nbodies;
	hipLaunchKernelGGL(k_clamp, dim3(nbh::grid_for(ctx, n3, kVecThreads)), dim3(kVecThreads), 0,
	                   ctx->stream, y, n3, b);
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status nbh_stats(nbh_ctx* ctx, const double* y, const double* mass, size_t n,
                     int with_energy, double out[12])
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(out != nullptr, "null output");
	std::memset(out, 0, 12 * sizeof(double));
	if(n == 0) { return NBH_OK; }
	NBH_CHECK(y != nullptr && mass != nullptr, "null buffer");
	unsigned grid_lin = nbh::grid_for(ctx, n, kStatThreads);
	unsigned grid_pe = static_cast<unsigned>((n + kStatThreads - 1) / kStatThreads);
	size_t lin_bytes = static_cast<size_t>(grid_lin) * kStatQ * sizeof(csum);
	size_t pe_bytes = with_energy ? static_cast<size_t>(grid_pe) * sizeof(csum) : 0;
	void* sp = nullptr;
	NBH_TRY(nbh::scratch(ctx, lin_bytes + pe_bytes + 16 * sizeof(double), &sp));
	csum* lin = static_cast<csum*>(sp);
	csum* pe = reinterpret_cast<csum*>(static_cast<char*>(sp) + lin_bytes);
	double* res = reinterpret_cast<double*>(static_cast<char*>(sp) + lin_bytes + pe_bytes);
	NBH_TRY(nbh::pinned(ctx, 16 * sizeof(double)));
	nbh::scoped_timer timer(ctx, "stats");
	hipLaunchKernelGGL(k_stats_linear, dim3(grid_lin), dim3(kStatThreads), 0, ctx->stream,
	                   y, mass, n, lin);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_csum_final<kStatQ>, dim3(1), dim3(kStatThreads), 0, ctx->stream,
	                   lin, static_cast<int>(grid_lin), res);
	NBH_LAUNCHED();
	if(with_energy)
	{
		hipLaunchKernelGGL(k_stats_pe, dim3(grid_pe), dim3(kStatThreads), 0, ctx->stream,
		                   y, mass, n, pe);
		NBH_LAUNCHED();
		hipLaunchKernelGGL(k_csum_final<1>, dim3(1), dim3(kStatThreads), 0, ctx->stream,
		                   pe, static_cast<int>(grid_pe), res + kStatQ);
		NBH_LAUNCHED();
	}
	NBH_HIP(hipMemcpyAsync(ctx->pinned, res, (kStatQ + 1) * sizeof(double), hipMemcpyDeviceToHost,
	                       ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	const double* r = ctx->pinned;
	out[0] = r[0]; out[1] = r[1]; out[2] = r[2];     // P
	out[3] = r[3]; out[4] = r[4]; out[5] = r[5];     // L
	out[6] = r[6] / 2.0;                            // KE
	out[7] = with_energy ? r[kStatQ] / 2.0 : 0.0;   // PE
	out[11] = r[10];                                // total mass
	out[8] = r[7] / r[10];                          // mass centre
	out[9] = r[8] / r[10];
	out[10] = r[9] / r[10];
	return NBH_OK;
}

}  // extern "C"
