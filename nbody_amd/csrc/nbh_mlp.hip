// nbh_mlp.hip -- the correction network: parameters, BF16-MFMA forward,
// hand-written backward and Adam, training driver (gfx950).
//
// Forward and backward of a 16-sample tile run in one wave on the matrix cores
// (nbh_mlp_dev.h).  A training step over a batch:
//   k_mlp_grad   each wave takes every nw-th tile of the batch, runs forward and
//                backward and writes its partial gradient (FP32) and loss;
//   k_mlp_adam   sums the partials of each parameter in a fixed order and
//                applies Adam to the FP32 master weights;
//   k_mlp_pack   re-packs the BF16 fragments.
// Batches are strided over the sample array (batch b holds samples b, b+nb,
// b+2nb, ...), so samples generated group by group are mixed without a shuffle.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "nbh_common.h"
#include "nbh_mlp_dev.h"

using namespace nbh_net;

struct nbh_mlp {
	nbh_ctx* ctx = nullptr;
	float*   params = nullptr;     // [kParams] FP32 master weights
	float*   adam_m = nullptr;
	float*   adam_v = nullptr;
	uint2*   frags = nullptr;      // [kFrags * 64] BF16 fragments
	float*   partial = nullptr;    // [max_waves * (kParams + 1)]
	float*   grad = nullptr;       // [kParams + 1] last gradient, loss at [kParams]
	int      max_waves = 0;
	long     step = 0;
};

namespace {

constexpr int kGradWaves = 32;      // waves per gradient launch

// B-fragment element of lane l, component e, of fragment f (see nbh_mlp_dev.h)
__device__ inline float frag_value(const float* __restrict__ p, int f, int lane, int e)
{
	const int kk = 4 * (lane >> 4) + e;   // row of B within the 16x16 block
	const int jj = lane & 15;             // column of B
	if(f < fragW2)                        // W1: B = W1[k][nt*16 + j]
	{
		const int nt = f - fragW1;
		return p[offW1 + kk * kHid + nt * 16 + jj];
	}
	if(f < fragW3)                        // W2: B = W2[kc*16 + k][nt*16 + j]
	{
		const int kc = (f - fragW2) / 4, nt = (f - fragW2) % 4;
		return p[offW2 + (kc * 16 + kk) * kHid + nt * 16 + jj];
	}
	if(f < fragW2T)                       // W3: B = W3[kc*16 + k][j]
	{
		const int kc = f - fragW3;
		return p[offW3 + (kc * 16 + kk) * kOutP + jj];
	}
	if(f < fragW3T)                       // W2^T: B = W2[nt*16 + j][kc*16 + k]
	{
		const int kc = (f - fragW2T) / 4, nt = (f - fragW2T) % 4;
		return p[offW2 + (nt * 16 + jj) * kHid + kc * 16 + kk];
	}
	const int nt = f - fragW3T;           // W3^T: B = W3[nt*16 + j][k]
	return p[offW3 + (nt * 16 + jj) * kOutP + kk];
}

__device__ inline void pack_all(const float* __restrict__ p, uint2* __restrict__ frags)
{
	for(int i = blockIdx.x * blockDim.x + threadIdx.x; i < kFrags * 64; i += gridDim.x * blockDim.x)
	{
		const int f = i / 64, lane = i % 64;
		uint32_t h[4];
		for(int e = 0; e < 4; ++e) { h[e] = to_bf16(frag_value(p, f, lane, e)); }
		frags[i] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
	}
}

__global__ void __launch_bounds__(256) k_mlp_pack(const float* __restrict__ p, uint2* __restrict__ frags)
{
	pack_all(p, frags);
}

// x: [m][16] features, out: [m][3]
__global__ void __launch_bounds__(64)
k_mlp_forward(const float* __restrict__ x, size_t m, const uint2* __restrict__ frags,
              const float* __restrict__ p, float* __restrict__ out)
{
	__shared__ __align__(16) uint16_t lds[2 * 16 * kHid];
	const int lane = threadIdx.x;
	const size_t ntiles = (m + 15) / 16;
	for(size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
	{
		const size_t row = tile * 16 + (lane & 15);
		const int k0 = 4 * (lane >> 4);
		float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
		if(row < m) { for(int e = 0; e < 4; ++e) { v[e] = x[row * kIn + k0 + e]; } }
		const f32x4 o = forward_tile(pack4(v[0], v[1], v[2], v[3]), frags, p, lds, lane, nullptr, nullptr);
		const int c = lane & 15;
		for(int e = 0; e < 4; ++e)
		{
			const size_t r = tile * 16 + 4 * (lane >> 4) + e;
			if(r < m && c < kOut) { out[r * kOut + c] = o[e]; }
		}
		__syncthreads();
	}
}

// One wave (block) per partial: tiles w, w + nw, ... of the batch whose
// samples are base + stride * s, s < count.
__global__ void __launch_bounds__(64)
k_mlp_grad(const float* __restrict__ x, const float* __restrict__ bs, const float* __restrict__ t, size_t base,
           size_t stride, size_t count, float inv_batch, const uint2* __restrict__ frags, const float* __restrict__ p,
           float* __restrict__ partial)
{
	__shared__ __align__(16) uint16_t lds[2 * 16 * kHid];
	__shared__ __align__(16) uint16_t zt[16 * kHid];
	__shared__ __align__(16) uint16_t dt[16 * 16];
	const int lane = threadIdx.x;
	const int col = lane & 15;
	const int grp = lane >> 4;
	f32x4 gW1[4], gW2[4][4], gW3[4];
	const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
	for(int i = 0; i < 4; ++i)
	{
		gW1[i] = zero;
		gW3[i] = zero;
		for(int j = 0; j < 4; ++j) { gW2[i][j] = zero; }
	}
	float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.0f, loss = 0.0f;
	const size_t ntiles = (count + 15) / 16;
	for(size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
	{
		const size_t s0 = tile * 16;
		// A fragment of X (rows = samples) and of X^T (rows = features)
		float xv[4], xtv[4];
		{
			const size_t r = s0 + col;
			for(int e = 0; e < 4; ++e) { xv[e] = r < count ? x[(base + stride * r) * kIn + 4 * grp + e] : 0.0f; }
			for(int e = 0; e < 4; ++e)
			{
				const size_t rs = s0 + 4 * grp + e;
				xtv[e] = rs < count ? x[(base + stride * rs) * kIn + col] : 0.0f;
			}
		}
		f32x4 h1[4], h2[4];
		const f32x4 out = forward_tile(pack4(xv[0], xv[1], xv[2], xv[3]), frags, p, lds, lane, h1, h2);
		// dL/dout, D layout (rows = samples 4 grp + e, column = output col): the
		// prediction of a row is the sum over its 16 lanes of c_col * basis_col
		f32x4 dout;
		for(int e = 0; e < 4; ++e)
		{
			const size_t r = s0 + 4 * grp + e;
			const bool ok = r < count;             // uniform over the row's 16 lanes
			float v[kOut];
			if(ok)
			{
				float bv[kBasis];
				const float* br = bs + (base + stride * r) * kBasis;
				for(int k = 0; k < kBasis; ++k) { bv[k] = br[k]; }
				basis_col(bv, col, v);
			}
			else
			{
				for(int k = 0; k < kOut; ++k) { v[k] = 0.0f; }
			}
			float P[kOut];
			for(int k = 0; k < kOut; ++k) { P[k] = out[e] * v[k]; }
			for(int sh = 1; sh < 16; sh <<= 1)
			{
				for(int k = 0; k < kOut; ++k) { P[k] += __shfl_xor(P[k], sh, 64); }
			}
			float gsum = 0.0f;
			for(int k = 0; k < kOut; ++k)
			{
				const float diff = ok ? P[k] - t[(base + stride * r) * kOut + k] : 0.0f;
				gsum += diff * v[k];
				if(col == 0) { loss += diff * diff; }
			}
			dout[e] = 2.0f * gsum * inv_batch;
			gb3 += dout[e];
		}
		const bf16x4 bdout = pack4(dout);
		for(int rt = 0; rt < 4; ++rt) { gW3[rt] = mfma(pack4(h2[rt]), bdout, gW3[rt]); }
		// dH2 = dOut W3^T: dOut as an A fragment through LDS
		for(int e = 0; e < 4; ++e) { dt[(4 * grp + e) * 16 + col] = to_bf16(dout[e]); }
		__syncthreads();
		const bf16x4 adout = lds_afrag(dt, 16, 0, lane);
		f32x4 dz2[4];
		for(int nt = 0; nt < 4; ++nt)
		{
			f32x4 d = mfma(adout, frag_load(frags, fragW3T + nt, lane), zero);
			for(int e = 0; e < 4; ++e) { d[e] *= 1.0f - h2[nt][e] * h2[nt][e]; gb2[nt] += d[e]; }
			dz2[nt] = d;
		}
		for(int rt = 0; rt < 4; ++rt)
		{
			const bf16x4 ah1 = pack4(h1[rt]);
			for(int nt = 0; nt < 4; ++nt) { gW2[rt][nt] = mfma(ah1, pack4(dz2[nt]), gW2[rt][nt]); }
		}
		// dH1 = dZ2 W2^T: dZ2 as A fragments through LDS
		for(int nt = 0; nt < 4; ++nt) { lds_store_d(zt, kHid, nt, lane, dz2[nt]); }
		__syncthreads();
		bf16x4 az[4];
		for(int kc = 0; kc < 4; ++kc) { az[kc] = lds_afrag(zt, kHid, kc, lane); }
		const bf16x4 axt = pack4(xtv[0], xtv[1], xtv[2], xtv[3]);
		for(int nt = 0; nt < 4; ++nt)
		{
			f32x4 d = zero;
			for(int kc = 0; kc < 4; ++kc) { d = mfma(az[kc], frag_load(frags, fragW2T + 4 * kc + nt, lane), d); }
			for(int e = 0; e < 4; ++e) { d[e] *= 1.0f - h1[nt][e] * h1[nt][e]; gb1[nt] += d[e]; }
			gW1[nt] = mfma(axt, pack4(d), gW1[nt]);
		}
		__syncthreads();
	}
	// partial gradient of this wave
	float* g = partial + static_cast<size_t>(blockIdx.x) * (kParams + 1);
	for(int nt = 0; nt < 4; ++nt)
	{
		for(int e = 0; e < 4; ++e)
		{
			const int f = 4 * grp + e;
			g[offW1 + f * kHid + nt * 16 + col] = gW1[nt][e];
		}
	}
	for(int rt = 0; rt < 4; ++rt)
	{
		for(int nt = 0; nt < 4; ++nt)
		{
			for(int e = 0; e < 4; ++e)
			{
				const int k = rt * 16 + 4 * grp + e;
				g[offW2 + k * kHid + nt * 16 + col] = gW2[rt][nt][e];
			}
		}
		for(int e = 0; e < 4; ++e)
		{
			const int h = rt * 16 + 4 * grp + e;
			g[offW3 + h * kOutP + col] = gW3[rt][e];
		}
	}
	// biases: sum the four lane groups of each column
	for(int nt = 0; nt < 4; ++nt)
	{
		float v1 = gb1[nt], v2 = gb2[nt];
		v1 += __shfl_xor(v1, 16, 64); v1 += __shfl_xor(v1, 32, 64);
		v2 += __shfl_xor(v2, 16, 64); v2 += __shfl_xor(v2, 32, 64);
		if(lane < 16)
		{
			g[offB1 + nt * 16 + lane] = v1;
			g[offB2 + nt * 16 + lane] = v2;
		}
	}
	float v3 = gb3;
	v3 += __shfl_xor(v3, 16, 64); v3 += __shfl_xor(v3, 32, 64);
	if(lane < 16) { g[offB3 + lane] = v3; }
	for(int off = 32; off > 0; off >>= 1) { loss += __shfl_down(loss, off, 64); }
	if(lane == 0) { g[kParams] = loss; }
}

// Sum the nw partial gradients of each parameter in order (and the loss);
// with apply, take an Adam step on the FP32 master weights.  The BF16
// fragments are re-packed by k_mlp_pack afterwards.
__global__ void __launch_bounds__(256)
k_mlp_adam(const float* __restrict__ partial, int nw, float* __restrict__ grad, float* __restrict__ p,
           float* __restrict__ am, float* __restrict__ av, float lr, float b1t, float b2t, int apply)
{
	const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i > kParams) { return; }
	float s = 0.0f;
	for(int w = 0; w < nw; ++w) { s += partial[static_cast<size_t>(w) * (kParams + 1) + i]; }
	grad[i] = s;
	if(apply && i < kParams)
	{
		const float m = beta1 * am[i] + (1.0f - beta1) * s;
		const float v = beta2 * av[i] + (1.0f - beta2) * s * s;
		am[i] = m;
		av[i] = v;
		const float mh = m / (1.0f - b1t);
		const float vh = v / (1.0f - b2t);
		p[i] -= lr * mh / (sqrtf(vh) + eps);
	}
}

constexpr int kAdamBlocks = (kParams + 1 + 255) / 256;
constexpr int kPackBlocks = (kFrags * 64 + 255) / 256;

}  // namespace

extern "C" {

nbh_status nbh_mlp_create(nbh_ctx* ctx, uint64_t seed, nbh_mlp** out)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(out != nullptr, "null output pointer");
	*out = nullptr;
	nbh_mlp* m = new nbh_mlp();
	m->ctx = ctx;
	m->max_waves = kGradWaves;
	hipError_t e = hipSuccess;
	auto alloc = [&](void** ptr, size_t bytes) {
		if(e == hipSuccess) { e = hipMalloc(ptr, bytes); }
	};
	alloc(reinterpret_cast<void**>(&m->params), kParams * sizeof(float));
	alloc(reinterpret_cast<void**>(&m->adam_m), kParams * sizeof(float));
	alloc(reinterpret_cast<void**>(&m->adam_v), kParams * sizeof(float));
	alloc(reinterpret_cast<void**>(&m->frags), kFrags * 64 * sizeof(uint2));
	alloc(reinterpret_cast<void**>(&m->partial), static_cast<size_t>(kGradWaves) * (kParams + 1) * sizeof(float));
	alloc(reinterpret_cast<void**>(&m->grad), (kParams + 1) * sizeof(float));
	if(e != hipSuccess)
	{
		nbh_mlp_destroy(m);
		return nbh::hip_fail(e, "mlp allocation", __FILE__, __LINE__);
	}
	// Glorot-uniform weights, zero biases; unused output columns stay zero.
	std::vector<float> h(kParams, 0.0f);
	std::mt19937_64 rng(seed);
	auto fill = [&](int off, int fan_in, int fan_out, int used_out) {
		const float a = std::sqrt(6.0f / static_cast<float>(fan_in + used_out));
		for(int i = 0; i < fan_in; ++i)
		{
			for(int j = 0; j < fan_out; ++j)
			{
				float u = static_cast<float>((rng() >> 11) * (1.0 / 9007199254740992.0));
				h[off + i * fan_out + j] = j < used_out ? (2.0f * u - 1.0f) * a : 0.0f;
			}
		}
	};
	fill(offW1, kIn, kHid, kHid);
	fill(offW2, kHid, kHid, kHid);
	fill(offW3, kHid, kOutP, kOut);
	nbh_status s = nbh_mlp_set_params(ctx, m, h.data());
	if(s != NBH_OK)
	{
		nbh_mlp_destroy(m);
		return s;
	}
	*out = m;
	return NBH_OK;
}

nbh_status nbh_mlp_destroy(nbh_mlp* m)
{
	if(m == nullptr)
	{
		return NBH_OK;
	}
	if(m->ctx != nullptr)
	{
		(void)hipSetDevice(m->ctx->device);
		(void)hipStreamSynchronize(m->ctx->stream);
	}
	void* ptrs[] = { m->params, m->adam_m, m->adam_v, m->frags, m->partial, m->grad };
	for(void* p : ptrs)
	{
		if(p != nullptr) { (void)hipFree(p); }
	}
	delete m;
	return NBH_OK;
}

nbh_status nbh_mlp_param_count(int* count)
{
	NBH_CHECK(count != nullptr, "null output");
	*count = kParams;
	return NBH_OK;
}

nbh_status nbh_mlp_set_params(nbh_ctx* ctx, nbh_mlp* m, const float* params)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(m != nullptr && m->ctx == ctx && params != nullptr, "bad arguments");
	NBH_HIP(hipMemcpyAsync(m->params, params, kParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
	NBH_HIP(hipMemsetAsync(m->adam_m, 0, kParams * sizeof(float), ctx->stream));
	NBH_HIP(hipMemsetAsync(m->adam_v, 0, kParams * sizeof(float), ctx->stream));
	hipLaunchKernelGGL(k_mlp_pack, dim3(kPackBlocks), dim3(256), 0, ctx->stream, m->params, m->frags);
	NBH_LAUNCHED();
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	m->step = 0;
	return NBH_OK;
}

nbh_status nbh_mlp_get_params(nbh_ctx* ctx, const nbh_mlp* m, float* params)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(m != nullptr && m->ctx == ctx && params != nullptr, "bad arguments");
	NBH_HIP(hipMemcpyAsync(params, m->params, kParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

nbh_status nbh_mlp_forward(nbh_ctx* ctx, const nbh_mlp* m, const float* x, size_t count, float* out)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(m != nullptr && m->ctx == ctx, "network does not belong to this context");
	if(count == 0) { return NBH_OK; }
	NBH_CHECK(x != nullptr && out != nullptr, "null buffer");
	unsigned grid = nbh::grid_for(ctx, (count + 15) / 16, 1);
	hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(64), 0, ctx->stream, x, count, m->frags, m->params, out);
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status nbh_mlp_gradient(nbh_ctx* ctx, nbh_mlp* m, const float* x, const float* basis, const float* t,
                            size_t count, float* grad, float* loss)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(m != nullptr && m->ctx == ctx, "network does not belong to this context");
	NBH_CHECK(count > 0 && x != nullptr && basis != nullptr && t != nullptr && grad != nullptr, "bad arguments");
	const int nw = static_cast<int>(std::min<size_t>(kGradWaves, (count + 15) / 16));
	hipLaunchKernelGGL(k_mlp_grad, dim3(nw), dim3(64), 0, ctx->stream, x, basis, t, static_cast<size_t>(0),
	                   static_cast<size_t>(1), count, 1.0f / static_cast<float>(count), m->frags, m->params,
	                   m->partial);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_mlp_adam, dim3(kAdamBlocks), dim3(256), 0, ctx->stream, m->partial, nw, m->grad,
	                   m->params, m->adam_m, m->adam_v, 0.0f, 0.0f, 0.0f, 0);
	NBH_LAUNCHED();
	std::vector<float> g(kParams + 1);
	NBH_HIP(hipMemcpyAsync(g.data(), m->grad, (kParams + 1) * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	std::memcpy(grad, g.data(), kParams * sizeof(float));
	if(loss != nullptr) { *loss = g[kParams] / static_cast<float>(count); }
	return NBH_OK;
}

nbh_status nbh_mlp_train(nbh_ctx* ctx, nbh_mlp* m, const float* x, const float* basis, const float* t, size_t count,
                         int epochs, size_t batch, float lr, float* epoch_loss)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(m != nullptr && m->ctx == ctx, "network does not belong to this context");
	NBH_CHECK(count > 0 && x != nullptr && basis != nullptr && t != nullptr && epochs > 0 && batch >= 16,
	          "bad arguments");
	const size_t nb = (count + batch - 1) / batch;
	const long total = static_cast<long>(nb) * epochs;
	for(int ep = 0; ep < epochs; ++ep)
	{
		double lsum = 0.0;
		for(size_t b = 0; b < nb; ++b)
		{
			// strided batch b: samples b, b + nb, ...
			const size_t cnt = (count - b + nb - 1) / nb;
			const int nw = static_cast<int>(std::min<size_t>(kGradWaves, (cnt + 15) / 16));
			++m->step;
			const float frac = static_cast<float>(m->step) / static_cast<float>(total);
			const float lr_t = lr * 0.5f * (1.0f + std::cos(3.14159265f * std::min(frac, 1.0f)));
			const float b1t = std::pow(0.9f, static_cast<float>(m->step));
			const float b2t = std::pow(0.999f, static_cast<float>(m->step));
			hipLaunchKernelGGL(k_mlp_grad, dim3(nw), dim3(64), 0, ctx->stream, x, basis, t, b, nb, cnt,
			                   1.0f / static_cast<float>(cnt), m->frags, m->params, m->partial);
			NBH_LAUNCHED();
			hipLaunchKernelGGL(k_mlp_adam, dim3(kAdamBlocks), dim3(256), 0, ctx->stream, m->partial, nw, m->grad,
			                   m->params, m->adam_m, m->adam_v, lr_t, b1t, b2t, 1);
			NBH_LAUNCHED();
			hipLaunchKernelGGL(k_mlp_pack, dim3(kPackBlocks), dim3(256), 0, ctx->stream, m->params, m->frags);
			NBH_LAUNCHED();
			if(epoch_loss != nullptr && b + 1 == nb)
			{
				float l = 0.0f;
				NBH_HIP(hipMemcpyAsync(&l, m->grad + kParams, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
				NBH_HIP(hipStreamSynchronize(ctx->stream));
				lsum = l / static_cast<double>(cnt);
			}
		}
		if(epoch_loss != nullptr) { epoch_loss[ep] = static_cast<float>(lsum); }
	}
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

}  // extern "C"

namespace nbh {

const uint2* mlp_frags(const nbh_mlp* m) { return m->frags; }
const float* mlp_params(const nbh_mlp* m) { return m->params; }
const nbh_ctx* mlp_ctx(const nbh_mlp* m) { return m->ctx; }

}  // namespace nbh

