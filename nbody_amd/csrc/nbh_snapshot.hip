// nbh_snapshot.hip -- asynchronous state snapshots for frame dumps.
//
// The reference dumps a frame by copying the state to the host and writing it
// before the next step starts (nbody/nbody_solver.cpp:93-103,
// nbody_engine_cuda.cpp:141-165).  Here a snapshot is two copies: device to a
// device staging slot on the context's stream (ordered after the step that
// produced the state, and before the next one overwrites it; ~0.1 ms for 8M
// bodies), then staging to pinned host memory on a second stream, which runs
// while the context's stream goes on with the next steps.  The caller waits for
// the second copy only when it needs the host data (a writer thread).
#include <cstring>

#include "nbh_common.h"

struct nbh_snapshot {
	nbh_ctx*    ctx = nullptr;
	size_t      bytes = 0;
	int         slots = 0;
	hipStream_t copy = nullptr;
	void*       stage[2] = {nullptr, nullptr};   // device
	void*       host[2] = {nullptr, nullptr};    // pinned
	hipEvent_t  ready[2] = {nullptr, nullptr};   // staging copy done (context stream)
	hipEvent_t  done[2] = {nullptr, nullptr};    // host copy done (copy stream)
	bool        pending[2] = {false, false};
};

namespace {

void release(nbh_snapshot* s)
{
	if(s->copy != nullptr) { (void)hipStreamSynchronize(s->copy); }
	for(int k = 0; k < 2; ++k)
	{
		if(s->stage[k] != nullptr) { (void)hipFree(s->stage[k]); }
		if(s->host[k] != nullptr) { (void)hipHostFree(s->host[k]); }
		if(s->ready[k] != nullptr) { (void)hipEventDestroy(s->ready[k]); }
		if(s->done[k] != nullptr) { (void)hipEventDestroy(s->done[k]); }
	}
	if(s->copy != nullptr) { (void)hipStreamDestroy(s->copy); }
	delete s;
}

}  // namespace

extern "C" {

nbh_status nbh_snapshot_create(nbh_ctx* ctx, size_t bytes, int slots, nbh_snapshot** out)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(out != nullptr, "null output pointer");
	*out = nullptr;
	NBH_CHECK(bytes > 0, "empty snapshot");
	NBH_CHECK(slots == 1 || slots == 2, "slots %d: must be 1 or 2", slots);
	nbh_snapshot* s = new nbh_snapshot();
	s->ctx = ctx;
	s->bytes = bytes;
	s->slots = slots;
	hipError_t e = hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking);
	for(int k = 0; k < slots && e == hipSuccess; ++k)
	{
		e = hipMalloc(&s->stage[k], bytes);
		if(e == hipSuccess) { e = hipHostMalloc(&s->host[k], bytes, hipHostMallocDefault); }
		if(e == hipSuccess) { e = hipEventCreateWithFlags(&s->ready[k], hipEventDisableTiming); }
		if(e == hipSuccess) { e = hipEventCreateWithFlags(&s->done[k], hipEventDisableTiming); }
	}
	if(e != hipSuccess)
	{
		release(s);
		return nbh::hip_fail(e, "snapshot allocation", __FILE__, __LINE__);
	}
	*out = s;
	return NBH_OK;
}

nbh_status nbh_snapshot_destroy(nbh_snapshot* s)
{
	if(s == nullptr)
	{
		return NBH_OK;
	}
	NBH_TRY(nbh::enter(s->ctx));
	release(s);
	return NBH_OK;
}

nbh_status nbh_snapshot_begin(nbh_ctx* ctx, nbh_snapshot* s, int slot, const void* src)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(s != nullptr && s->ctx == ctx, "snapshot does not belong to this context");
	NBH_CHECK(slot >= 0 && slot < s->slots, "slot %d out of range", slot);
	NBH_CHECK(src != nullptr, "null source");
	NBH_CHECK(!s->pending[slot], "slot %d still holds a snapshot nobody waited for", slot);
	NBH_HIP(hipMemcpyAsync(s->stage[slot], src, s->bytes, hipMemcpyDeviceToDevice, ctx->stream));
	NBH_HIP(hipEventRecord(s->ready[slot], ctx->stream));
	NBH_HIP(hipStreamWaitEvent(s->copy, s->ready[slot], 0));
	NBH_HIP(hipMemcpyAsync(s->host[slot], s->stage[slot], s->bytes, hipMemcpyDeviceToHost, s->copy));
	NBH_HIP(hipEventRecord(s->done[slot], s->copy));
	s->pending[slot] = true;
	return NBH_OK;
}

nbh_status nbh_snapshot_wait(nbh_ctx* ctx, nbh_snapshot* s, int slot, const void** host)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(s != nullptr && s->ctx == ctx, "snapshot does not belong to this context");
	NBH_CHECK(slot >= 0 && slot < s->slots, "slot %d out of range", slot);
	NBH_CHECK(host != nullptr, "null output pointer");
	NBH_CHECK(s->pending[slot], "slot %d holds no snapshot", slot);
	NBH_HIP(hipEventSynchronize(s->done[slot]));
	s->pending[slot] = false;
	*host = s->host[slot];
	return NBH_OK;
}

}  // extern "C"
