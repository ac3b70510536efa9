// Issue-rate probe for FP32 vector instructions on gfx950 (no packed FP32).
//
// Each kernel runs 8 independent chains of one instruction form per work-item
// (the pair probe: 4 targets x 4 sources), 4 waves per work-group, 32
// work-groups per CU, and reports wave
// instructions per SIMD per clock at the nominal 2.4 GHz.  The force kernels'
// roofline is stated against these rates: the FP32 peak of 157.3 TFLOP/s counts
// packed FP32 (two lanes' worth per instruction), which this build does not
// issue, and the single-lane forms issue at different rates by encoding.
//
//   fma_vop3      v_fma_f32 (64-bit encoding, one SGPR operand)
//   fmac_vop2     v_fmac_f32_e32, VGPR operands
//   mul_vop2      v_mul_f32_e32, VGPR operands
//   rsq           v_rsq_f32
//   pair_regs     the direct kernel's pair arithmetic (13 instructions per
//                 pair) with the sources held in registers
//   pair_regs_v   the same with targets that differ in every coordinate per lane
//   pair_mfma     the pair arithmetic up to q^3 (8 VALU + v_rsq_f32), the sum
//                 over sources on v_mfma_f32_4x4x1_16b_f32 (one per 64 pairs)
//   pair_mfma_v   the same with the per-lane targets of pair_regs_v
// and checks the lane layout of v_mfma_f32_4x4x1_16b_f32 (A lane 4b+i is row i
// of block b, B lane 4b+j column j, D lane 4b+j register i).
//
// Usage: valu_issue  (prints one line per kernel: best of 3 launches)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define PROBE_CHECK(x)                                                                      \
	do                                                                                      \
	{                                                                                       \
		hipError_t e_ = (x);                                                                \
		if(e_ != hipSuccess)                                                                \
		{                                                                                   \
			std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                    \
			std::exit(1);                                                                   \
		}                                                                                   \
	} while(0)

namespace {

constexpr int kThreads = 256;
constexpr int kIters = 4096;

__global__ void __launch_bounds__(kThreads) k_fma_vop3(float* out, float a, float b)
{
	float x[8];
	for(int k = 0; k < 8; ++k) { x[k] = threadIdx.x * 1e-3f + k; }
	for(int i = 0; i < kIters; ++i)
	{
#pragma unroll
		for(int k = 0; k < 8; ++k) { x[k] = fmaf(x[k], a, b); }
	}
	float s = 0.0f;
	for(int k = 0; k < 8; ++k) { s += x[k]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) k_fmac_vop2(float* out, float a, float b)
{
	float x[8], y[8];
	const float z = a + threadIdx.x * 1e-9f;
	for(int k = 0; k < 8; ++k)
	{
		x[k] = threadIdx.x * 1e-3f + k;
		y[k] = b + k * 1e-6f;
	}
	for(int i = 0; i < kIters; ++i)
	{
#pragma unroll
		for(int k = 0; k < 8; ++k) { x[k] = fmaf(y[k], z, x[k]); }
	}
	float s = 0.0f;
	for(int k = 0; k < 8; ++k) { s += x[k]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) k_mul_vop2(float* out, float a, float b)
{
	float x[8], y[8];
	for(int k = 0; k < 8; ++k)
	{
		x[k] = threadIdx.x * 1e-3f + k;
		y[k] = a + k * 1e-7f + threadIdx.x * 1e-9f;
	}
	for(int i = 0; i < kIters; ++i)
	{
#pragma unroll
		for(int k = 0; k < 8; ++k) { x[k] = x[k] * y[k]; }
	}
	float s = 0.0f;
	for(int k = 0; k < 8; ++k) { s += x[k]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s + b;
}

__global__ void __launch_bounds__(kThreads) k_rsq(float* out, float a, float b)
{
	float x[8];
	for(int k = 0; k < 8; ++k) { x[k] = threadIdx.x * 1e-3f + k + 1.0f; }
	for(int i = 0; i < kIters; ++i)
	{
#pragma unroll
		for(int k = 0; k < 8; ++k) { x[k] = __builtin_amdgcn_rsqf(x[k]); }
	}
	float s = 0.0f;
	for(int k = 0; k < 8; ++k) { s += x[k]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s * a + b;
}

// 4 targets per work-item against 4 register-held sources per iteration; the
// sources rotate each iteration so the loop cannot be folded.
constexpr int kPairIters = kIters / 4;

__global__ void __launch_bounds__(kThreads) k_pair_regs(float* out, float a, float b)
{
	float xi[4], yi[4], zi[4], px[4], py[4], pz[4];
	for(int k = 0; k < 4; ++k)
	{
		xi[k] = threadIdx.x * 1e-3f + k;
		yi[k] = 0.5f * k;
		zi[k] = -0.25f * k;
		px[k] = py[k] = pz[k] = 0.0f;
	}
	float4 s4[4];
	for(int j = 0; j < 4; ++j) { s4[j] = make_float4(a * (j + 1), b * j - 1.0f, 0.5f * j, 0.1f * (j + 1)); }
	for(int i = 0; i < kPairIters; ++i)
	{
#pragma unroll
		for(int j = 0; j < 4; ++j)
		{
			const float4 s = s4[j];
#pragma unroll
			for(int k = 0; k < 4; ++k)
			{
				const float dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
				const float q = __builtin_amdgcn_fmed3f(
				    __builtin_amdgcn_rsqf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))), 0.0f, 1.0f);
				const float w = (s.w * q) * (q * q);
				px[k] = fmaf(dx, w, px[k]);
				py[k] = fmaf(dy, w, py[k]);
				pz[k] = fmaf(dz, w, pz[k]);
			}
		}
		const float4 t = s4[0];
		s4[0] = s4[1];
		s4[1] = s4[2];
		s4[2] = s4[3];
		s4[3] = t;
	}
	float s = 0.0f;
	for(int k = 0; k < 4; ++k) { s += px[k] + py[k] + pz[k]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s;
}

template <bool kVary>
__global__ void __launch_bounds__(kThreads) k_pair_regs_t(float* out, float a, float b)
{
	float xi[4], yi[4], zi[4], px[4], py[4], pz[4];
	for(int k = 0; k < 4; ++k)
	{
		xi[k] = threadIdx.x * 1e-3f + k;
		yi[k] = kVary ? threadIdx.x * 0.5f * k : 0.5f * k;
		zi[k] = kVary ? -0.25f * k * threadIdx.x : -0.25f * k;
		px[k] = py[k] = pz[k] = 0.0f;
	}
	float4 s4[4];
	for(int j = 0; j < 4; ++j) { s4[j] = make_float4(a * (j + 1), b * j - 1.0f, 0.5f * j, 0.1f * (j + 1)); }
	for(int i = 0; i < kPairIters; ++i)
	{
#pragma unroll
		for(int j = 0; j < 4; ++j)
		{
			const float4 s = s4[j];
#pragma unroll
			for(int k = 0; k < 4; ++k)
			{
				const float dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
				const float q = __builtin_amdgcn_fmed3f(
				    __builtin_amdgcn_rsqf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))), 0.0f, 1.0f);
				const float w = (s.w * q) * (q * q);
				px[k] = fmaf(dx, w, px[k]);
				py[k] = fmaf(dy, w, py[k]);
				pz[k] = fmaf(dz, w, pz[k]);
			}
		}
		const float4 t = s4[0];
		s4[0] = s4[1];
		s4[1] = s4[2];
		s4[2] = s4[3];
		s4[3] = t;
	}
	float s = 0.0f;
	for(int k = 0; k < 4; ++k) { s += px[k] + py[k] + pz[k]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s;
}

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// Sum over sources as a contraction: acc[k] += q^3 (this lane's target k) x B,
// B = one component of the source per lane (as m x, m y, m z, m would be).
template <bool kVary>
__global__ void __launch_bounds__(kThreads) k_pair_mfma_t(float* out, float a, float b)
{
	float xi[4], yi[4], zi[4];
	mfma_f4 acc[4];
	for(int k = 0; k < 4; ++k)
	{
		xi[k] = threadIdx.x * 1e-3f + k;
		yi[k] = kVary ? threadIdx.x * 0.5f * k : 0.5f * k;
		zi[k] = kVary ? -0.25f * k * threadIdx.x : -0.25f * k;
		acc[k] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
	}
	float4 s4[4];
	float bb[4];
	for(int j = 0; j < 4; ++j)
	{
		s4[j] = make_float4(a * (j + 1), b * j - 1.0f, 0.5f * j, 0.1f * (j + 1));
		bb[j] = 0.3f * j + threadIdx.x * 1e-4f;
	}
	for(int i = 0; i < kPairIters; ++i)
	{
#pragma unroll
		for(int j = 0; j < 4; ++j)
		{
			const float4 s = s4[j];
#pragma unroll
			for(int k = 0; k < 4; ++k)
			{
				const float dx = s.x - xi[k], dy = s.y - yi[k], dz = s.z - zi[k];
				const float q = __builtin_amdgcn_fmed3f(
				    __builtin_amdgcn_rsqf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))), 0.0f, 1.0f);
				acc[k] = __builtin_amdgcn_mfma_f32_4x4x1f32((q * q) * q, bb[j], acc[k], 0, 0, 0);
			}
		}
		const float4 t = s4[0];
		s4[0] = s4[1];
		s4[1] = s4[2];
		s4[2] = s4[3];
		s4[3] = t;
		const float tb = bb[0];
		bb[0] = bb[1];
		bb[1] = bb[2];
		bb[2] = bb[3];
		bb[3] = tb;
	}
	float s = 0.0f;
	for(int k = 0; k < 4; ++k) { s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3]; }
	out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// Work-group w (one wave): A = 1 on lane w only, B = lane + 1; out[w][lane][4] = D.
__global__ void __launch_bounds__(64) k_mfma_layout(float* out)
{
	const int l = threadIdx.x;
	const int w = blockIdx.x;
	const mfma_f4 d = __builtin_amdgcn_mfma_f32_4x4x1f32(l == w ? 1.0f : 0.0f, static_cast<float>(l + 1),
	                                                     mfma_f4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
	for(int v = 0; v < 4; ++v) { out[(w * 64 + l) * 4 + v] = d[v]; }
}

} // namespace

int main()
{
	hipDeviceProp_t prop;
	PROBE_CHECK(hipGetDeviceProperties(&prop, 0));
	const int cus = prop.multiProcessorCount;
	const int grid = cus * 32;
	float* out = nullptr;
	PROBE_CHECK(hipMalloc(&out, sizeof(float) * static_cast<size_t>(grid) * kThreads));
	hipEvent_t e0, e1;
	PROBE_CHECK(hipEventCreate(&e0));
	PROBE_CHECK(hipEventCreate(&e1));
	struct Probe
	{
		const char* name;
		void (*kernel)(float*, float, float);
		double instr_per_item;    // wave instructions of the probed form per wave
	};
	const Probe probes[] = {
	    {"fma_vop3", k_fma_vop3, 8.0 * kIters},
	    {"fmac_vop2", k_fmac_vop2, 8.0 * kIters},
	    {"mul_vop2", k_mul_vop2, 8.0 * kIters},
	    {"rsq", k_rsq, 8.0 * kIters},
	    {"pair_regs", k_pair_regs, 16.0 * kPairIters * 13.0},
	    {"pair_regs_v", k_pair_regs_t<true>, 16.0 * kPairIters * 13.0},
	    {"pair_mfma", k_pair_mfma_t<false>, 16.0 * kPairIters * 13.0},
	    {"pair_mfma_v", k_pair_mfma_t<true>, 16.0 * kPairIters * 13.0},
	};
	std::printf("device %s, %d CUs, clock %d MHz\n", prop.gcnArchName, cus, prop.clockRate / 1000);
	for(const Probe& p : probes)
	{
		float best = 1e30f;
		for(int rep = 0; rep < 3; ++rep)
		{
			PROBE_CHECK(hipEventRecord(e0));
			hipLaunchKernelGGL(p.kernel, dim3(grid), dim3(kThreads), 0, 0, out, 0.999f, 0.001f);
			PROBE_CHECK(hipGetLastError());
			PROBE_CHECK(hipEventRecord(e1));
			PROBE_CHECK(hipEventSynchronize(e1));
			float ms = 0.0f;
			PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
			if(ms < best) { best = ms; }
		}
		const double waves = static_cast<double>(grid) * (kThreads / 64);
		const double rate = p.instr_per_item * waves / (best * 1e-3);
		std::printf("%-10s %8.3f ms  %.3e wave-instr/s  %.3f per SIMD per clock at 2.4 GHz\n", p.name, best, rate,
		            rate / (4.0 * cus) / 2.4e9);
	}
	// the pair_* lines are normalised to 13 instructions per pair for every form:
	// compare their times (equal pair counts)
	static float lay[64 * 64 * 4];
	hipLaunchKernelGGL(k_mfma_layout, dim3(64), dim3(64), 0, 0, out);
	PROBE_CHECK(hipGetLastError());
	PROBE_CHECK(hipMemcpy(lay, out, sizeof(lay), hipMemcpyDeviceToHost));
	int bad = 0;
	for(int w = 0; w < 64; ++w)
	{
		for(int l = 0; l < 64; ++l)
		{
			for(int v = 0; v < 4; ++v)
			{
				const bool hit = l / 4 == w / 4 && v == w % 4;
				const float want = hit ? static_cast<float>(l + 1) : 0.0f;
				if(lay[(w * 64 + l) * 4 + v] != want) { ++bad; }
			}
		}
	}
	std::printf("mfma_f32_4x4x1 layout (A lane 4b+i row i, B lane 4b+j column j, D lane 4b+j register i): %s\n",
	            bad == 0 ? "ok" : "MISMATCH");
	PROBE_CHECK(hipFree(out));
	return bad == 0 ? 0 : 2;
}
