// nbh_common.h -- shared internals of libnbody_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "nbody_hip.h"

// Device time of one kernel class: event pairs not yet read, plus the totals
// of those already folded in.
struct nbh_timer_class {
	std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
	double   total_ms = 0.0;
	uint64_t launches = 0;
};

// One device and one in-order stream.  Scratch memory for reductions and the
// split-j partial sums of the FP32 direct kernel is owned by the context and
// grown on demand (never shrunk until nbh_ctx_destroy).
struct nbh_ctx {
	int          device = 0;
	hipStream_t  stream = nullptr;
	int          cu_count = 0;
	char         name[128] = {0};
	void*        scratch = nullptr;       // device scratch
	size_t       scratch_bytes = 0;
	double*      pinned = nullptr;        // small pinned host buffer for results
	size_t       pinned_bytes = 0;
	bool         timing = false;
	std::map<std::string, nbh_timer_class> timers;
	std::vector<hipEvent_t> event_pool;
};

namespace nbh {

void set_error(const char* fmt, ...);

// Record a HIP failure and return NBH_ERR_HIP.
nbh_status hip_fail(hipError_t e, const char* what, const char* file, int line);

// Make ctx->device current; validates ctx.
nbh_status enter(nbh_ctx* ctx);

// Ensure at least `bytes` of device scratch; returns the pointer via *out.
nbh_status scratch(nbh_ctx* ctx, size_t bytes, void** out);

// Ensure at least `bytes` of pinned host memory.
nbh_status pinned(nbh_ctx* ctx, size_t bytes);

// Brackets the launches made during its lifetime with events on ctx->stream
// when timing is enabled (kernel classes of nbh_timing_read).
class scoped_timer {
	nbh_ctx*    m_ctx;
	const char* m_name;
	hipEvent_t  m_start;
public:
	scoped_timer(nbh_ctx* ctx, const char* name);
	~scoped_timer();
};

// Device buffers of a correction network (defined in nbh_mlp.hip).
const uint2* mlp_frags(const nbh_mlp* m);
const float* mlp_params(const nbh_mlp* m);
const nbh_ctx* mlp_ctx(const nbh_mlp* m);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// keys, 8 bits per pass, ping-ponging between (k0, v0) and (k1, v1); *sorted_vals
// points at the buffer holding the result.  counts: sort_counts_size(n) words,
// scan_part: sort_scan_size() words of device scratch.
size_t sort_counts_size(size_t n);
size_t sort_scan_size();
nbh_status sort_pairs(nbh_ctx* ctx, uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n, int bits,
                      uint32_t* counts, uint32_t* scan_part, uint32_t** sorted_vals);

// Number of work-groups for a grid-stride kernel over `work` items with
// `per_group` items per work-group pass: enough to fill every CU several times.
inline unsigned grid_for(const nbh_ctx* ctx, size_t work, size_t per_group)
{
	size_t want = (work + per_group - 1) / per_group;
	size_t cap = static_cast<size_t>(ctx->cu_count > 0 ? ctx->cu_count : 256) * 8;
	if(want > cap) { want = cap; }
	if(want == 0) { want = 1; }
	return static_cast<unsigned>(want);
}

}  // namespace nbh

#define NBH_HIP(expr)                                                          \
	do {                                                                       \
		hipError_t nbh_e_ = (expr);                                            \
		if(nbh_e_ != hipSuccess)                                               \
		{                                                                      \
			return nbh::hip_fail(nbh_e_, #expr, __FILE__, __LINE__);           \
		}                                                                      \
	} while(0)

#define NBH_LAUNCHED()                                                         \
	do {                                                                       \
		hipError_t nbh_e_ = hipGetLastError();                                 \
		if(nbh_e_ != hipSuccess)                                               \
		{                                                                      \
			return nbh::hip_fail(nbh_e_, "kernel launch", __FILE__, __LINE__); \
		}                                                                      \
	} while(0)

#define NBH_CHECK(cond, ...)                                                   \
	do {                                                                       \
		if(!(cond))                                                            \
		{                                                                      \
			nbh::set_error(__VA_ARGS__);                                       \
			return NBH_ERR_INVALID;                                            \
		}                                                                      \
	} while(0)

#define NBH_TRY(expr)                                                          \
	do {                                                                       \
		nbh_status nbh_s_ = (expr);                                            \
		if(nbh_s_ != NBH_OK) { return nbh_s_; }                                \
	} while(0)
