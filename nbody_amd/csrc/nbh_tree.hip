// nbh_tree.hip -- Barnes-Hut heap kd-tree on the GPU: exact build, refit and
// traversal (gfx950).
//
// Layout (nbody/nbody_space_heap.cpp:13-37): n = 2^L bodies, heap indices
// 1..2n-1 (root 1, children 2i and 2i+1, leaves at [n, 2n)).
//
// Build = the reference's recursive median split (nbody_space_heap.cpp:136-191)
// done level by level for every segment at once:
//   1. three stable LSD radix sorts give the bodies in x, y and z order (keys:
//      order-preserving 64-bit images of the FP64 coordinates; equal
//      coordinates keep index order);
//   2. at level l every segment of S = n >> l consecutive list entries holds the
//      bodies of one node; along axis a = l % 3 the first S/2 entries of the
//      a-sorted list are exactly the count/2 smallest, i.e. std::nth_element's
//      left part.  Each body is tagged left/right, and the two other lists are
//      stably partitioned inside every segment (segmented scan + scatter), so
//      every child segment again holds its bodies sorted along all three axes;
//   3. after L levels entry s of any list is the body of leaf n + s;
//   4. moments bottom-up with the operations of nbody_space_heap::update
//      (:119-134), in FP64 with contraction off: bit-identical node records.
// Traversal: one target per work-item, targets in heap-leaf order (spatially
// coherent), nodes as 32-byte AoS records {cm.xyz, crit_r2} plus mass.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "nbh_common.h"
#include "nbh_mlp_dev.h"

namespace {

constexpr int kRoot = 1;               // NBODY_HEAP_ROOT_INDEX
constexpr int kMaxStack = 64;          // MAX_STACK_SIZE of the reference traversal
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;   // 4096 keys per work-group
constexpr int kPartThreads = 256;
constexpr int kPartItems = 4;
constexpr int kPartTile = kPartThreads * kPartItems;   // 1024 entries per work-group
constexpr int kScanThreads = 256;
constexpr int kScanParts = 1024;                     // chunks of one scan
constexpr int kScanOneThreads = 1024;
constexpr int kScanOneItems = 16;                    // scans of up to 16384 counts in one work-group
constexpr int kLocalSeg = 4096;                      // segments split in LDS, one work-group each
constexpr int kLocalThreads = 1024;
constexpr int kBoxParts = 2048;                      // partial boxes of the Morton tree's bounding box
constexpr int kRadixLevels = 128;                    // prefix lengths of index-extended 63-bit keys: 0..127

}  // namespace

// Node record of the group walks: what a walk step reads for one node, in 64
// bytes, so that a sibling pair (2i, 2i + 1) is one 128-byte cache line.
struct __align__(64) nbh_walk_rec {
	double4 c;      // cm.x, cm.y, cm.z, crit_r2 (as xyzr)
	double  m;      // mass
	float   q[6];   // second central moments xx yy zz xy xz yz (0 on leaves), rounded to float
};
static_assert(sizeof(nbh_walk_rec) == 64, "walk record must be 64 bytes");

struct nbh_tree {
	nbh_ctx*  ctx = nullptr;
	size_t    n = 0;
	int       levels = 0;
	// per-axis sorted body lists and scratch
	uint32_t* list[3] = {nullptr, nullptr, nullptr};
	uint32_t* tmp = nullptr;
	uint8_t*  side = nullptr;
	uint64_t* keys = nullptr;      // [2][n] radix sort ping-pong
	uint32_t* vals = nullptr;      // [2][n]
	uint32_t* counts = nullptr;    // radix / partition block counts
	uint32_t* scan_part = nullptr; // [kScanParts] chunk sums of a scan
	// nodes, heap index 0..2n-1 (0 unused)
	double4*  xyzr = nullptr;      // cm.x, cm.y, cm.z, crit_r2
	double*   mass = nullptr;
	double*   bmin = nullptr;      // [2n*3]
	double*   bmax = nullptr;      // [2n*3]
	double*   quad = nullptr;      // [2n*6] second central moments (0 on leaves)
	nbh_walk_rec* wrec = nullptr;  // [2n] packed copy for the group walks
	int32_t*  body_n = nullptr;    // -1 on internal nodes
	unsigned long long* counters = nullptr;   // [5] interactions, visits, stack spills, far cells, near cells
	// two-level local walk (allocated on first use): per super-group node lists,
	// their lengths and the super-group expansions
	int       two_cap = 0;           // list capacity per super-group (0: kTwoCap)
	int*      two_lists = nullptr;
	int*      two_count = nullptr;
	float*    two_coef = nullptr;
	// the super-groups' far cells for the learned correction (allocated on first use)
	int*      far_list = nullptr;
	int*      far_count = nullptr;
	// Morton radix tree (nbh_tree_build_radix; allocated on its first build)
	int2*     child = nullptr;     // [n] children of internal ids 1..n-1
	int*      parent = nullptr;    // [2n] parent id of every non-root node
	int*      level = nullptr;     // [n] level of internal ids: the common prefix length of their keys
	int*      level_list = nullptr;   // [n] internal ids level by level
	int*      level_info = nullptr;   // [3 * kRadixLevels + 3] level counts, offsets, cursors; the group counts
	int2*     gmark = nullptr;     // [n] group marks at their first sorted leaf (build scratch)
	int2*     smark = nullptr;     // [n] super-group marks
	int2*     wmark = nullptr;     // [n] walk group marks
	uint32_t* flags = nullptr;     // [3n] build group, super-group and walk group flags, scanned in place
	int4*     grp = nullptr;       // [n] build groups (node, first leaf, leaves)
	int4*     wgrp = nullptr;      // [n] walk groups (node, first leaf, leaves, super-group)
	int2*     sup = nullptr;       // [n / kRadixWalk + 1] super-groups (node, first leaf)
	double*   box = nullptr;       // [6 * (kBoxParts + 1)] partial boxes, then the bounding box
	double4*  pk = nullptr;        // [n] position and mass of each body, packed for the sorted gather
	int       ngroups = 0;         // build groups, walk groups and super-groups of the last Morton build
	int       nwalk = 0;
	int       nsupers = 0;
	int       level_count[kRadixLevels] = {};   // host copy of the level counts (refit)
	int       two_slots = 0;       // super-groups the two-level buffers hold
	bool      radix = false;       // the last build made the Morton tree
	bool      radix_lean = false;  // Morton node arrays only where the build reads them back
	double    ratio = 0.0;         // opening ratio of the last build / refit
	bool      built = false;
};

namespace {

// ---------------------------------------------------------------------------
// radix sort of (64-bit key, 32-bit value), stable, 8 bits per pass
// ---------------------------------------------------------------------------
__device__ inline uint64_t order_key(double v)
{
	v = v + 0.0;                                   // -0 -> +0: equal values compare equal
	uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
	return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__global__ void __launch_bounds__(256)
k_make_keys(const double* __restrict__ coord, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, size_t n)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		keys[i] = order_key(coord[i]);
		vals[i] = static_cast<uint32_t>(i);
	}
}

// counts[d * nblocks + b] = number of keys of tile b with digit d
__global__ void __launch_bounds__(kSortThreads)
k_radix_hist(const uint64_t* __restrict__ keys, size_t n, int shift, uint32_t* __restrict__ counts, int nblocks)
{
	__shared__ uint32_t hist[256];
	hist[threadIdx.x] = 0;
	__syncthreads();
	size_t base = static_cast<size_t>(blockIdx.x) * kSortTile;
	for(int k = 0; k < kSortItems; ++k)
	{
		size_t i = base + static_cast<size_t>(k) * kSortThreads + threadIdx.x;
		if(i < n)
		{
			atomicAdd(&hist[(keys[i] >> shift) & 0xFF], 1u);
		}
	}
	__syncthreads();
	counts[static_cast<size_t>(threadIdx.x) * nblocks + blockIdx.x] = hist[threadIdx.x];
}

// Exclusive scan in place of `m` uint32 counts: chunk sums, a scan of the (at
// most kScanParts) chunk sums, then a scan of every chunk from its offset.
__global__ void __launch_bounds__(kScanThreads)
k_scan_reduce(const uint32_t* __restrict__ c, size_t m, size_t chunk, uint32_t* __restrict__ part)
{
	__shared__ uint32_t ws[kScanThreads / 64];
	const size_t b0 = static_cast<size_t>(blockIdx.x) * chunk;
	const size_t b1 = b0 + chunk < m ? b0 + chunk : m;
	uint32_t s = 0;
	for(size_t i = b0 + threadIdx.x; i < b1; i += kScanThreads) { s += c[i]; }
	for(int off = 32; off > 0; off >>= 1) { s += __shfl_down(s, off, 64); }
	if((threadIdx.x & 63) == 0) { ws[threadIdx.x >> 6] = s; }
	__syncthreads();
	if(threadIdx.x == 0)
	{
		uint32_t t = 0;
		for(int w = 0; w < kScanThreads / 64; ++w) { t += ws[w]; }
		part[blockIdx.x] = t;
	}
}

__global__ void __launch_bounds__(kScanParts)
k_scan_parts(uint32_t* __restrict__ part, int g)
{
	__shared__ uint32_t sh[kScanParts];
	const uint32_t v = static_cast<int>(threadIdx.x) < g ? part[threadIdx.x] : 0;
	sh[threadIdx.x] = v;
	__syncthreads();
	for(int off = 1; off < kScanParts; off <<= 1)
	{
		const uint32_t u = threadIdx.x >= static_cast<unsigned>(off) ? sh[threadIdx.x - off] : 0;
		__syncthreads();
		sh[threadIdx.x] += u;
		__syncthreads();
	}
	if(static_cast<int>(threadIdx.x) < g) { part[threadIdx.x] = sh[threadIdx.x] - v; }
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_apply(uint32_t* __restrict__ c, size_t m, size_t chunk, const uint32_t* __restrict__ part)
{
	constexpr int kWaves = kScanThreads / 64;
	__shared__ uint32_t ws[kWaves];
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	uint32_t carry = part[blockIdx.x];
	const size_t b0 = static_cast<size_t>(blockIdx.x) * chunk;
	const size_t b1 = b0 + chunk < m ? b0 + chunk : m;
	for(size_t t0 = b0; t0 < b1; t0 += 4 * kScanThreads)
	{
		const size_t i0 = t0 + 4 * static_cast<size_t>(threadIdx.x);
		uint32_t v[4];
		for(int k = 0; k < 4; ++k) { v[k] = i0 + k < b1 ? c[i0 + k] : 0; }
		const uint32_t tsum = v[0] + v[1] + v[2] + v[3];
		uint32_t x = tsum;
		for(int off = 1; off < 64; off <<= 1)
		{
			const uint32_t u = __shfl_up(x, off, 64);
			if(lane >= off) { x += u; }
		}
		if(lane == 63) { ws[wave] = x; }
		__syncthreads();
		uint32_t wbase = 0, total = 0;
		for(int w = 0; w < kWaves; ++w)
		{
			wbase += w < wave ? ws[w] : 0;
			total += ws[w];
		}
		uint32_t run = carry + wbase + x - tsum;
		for(int k = 0; k < 4; ++k)
		{
			if(i0 + k < b1) { c[i0 + k] = run; }
			run += v[k];
		}
		carry += total;
		__syncthreads();
	}
}

// Exclusive scan in place of m <= kScanOneThreads * kScanOneItems counts by one
// work-group: each work-item scans a run of consecutive counts.
__global__ void __launch_bounds__(kScanOneThreads)
k_scan_one(uint32_t* __restrict__ c, int m)
{
	__shared__ uint32_t ws[kScanOneThreads / 64];
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const int per = (m + kScanOneThreads - 1) / kScanOneThreads;
	const int i0 = static_cast<int>(threadIdx.x) * per;
	uint32_t v[kScanOneItems];
	uint32_t s = 0;
#pragma unroll
	for(int k = 0; k < kScanOneItems; ++k)
	{
		v[k] = k < per && i0 + k < m ? c[i0 + k] : 0;
		s += v[k];
	}
	uint32_t x = s;
	for(int off = 1; off < 64; off <<= 1)
	{
		const uint32_t u = __shfl_up(x, off, 64);
		if(lane >= off) { x += u; }
	}
	if(lane == 63) { ws[wave] = x; }
	__syncthreads();
	uint32_t run = x - s;
	for(int w = 0; w < wave; ++w) { run += ws[w]; }
#pragma unroll
	for(int k = 0; k < kScanOneItems; ++k)
	{
		if(k < per && i0 + k < m) { c[i0 + k] = run; }
		run += v[k];
	}
}

// Stable scatter of one 4096-key tile: rows of 256 keys are ranked in index
// order, wave by wave, with a ballot match on the 8 digit bits, and placed in
// LDS in digit order; the tile then leaves as one contiguous run per digit.
// `offsets` is the exclusive scan (digit-major) of the per-tile digit counts,
// so a tile's count of digit d is the next entry minus its own (n after the
// last).
__global__ void __launch_bounds__(kSortThreads)
k_radix_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n, int shift,
                const uint32_t* __restrict__ offsets, int nblocks)
{
	constexpr int kWaves = kSortThreads / 64;
	__shared__ uint64_t skey[kSortTile];
	__shared__ uint32_t sval[kSortTile];
	__shared__ uint32_t goff[256];     // first output position of digit d for this tile
	__shared__ uint32_t loff[256];     // first LDS position of digit d
	__shared__ uint32_t running[256];  // keys of digit d placed so far
	__shared__ uint32_t wcount[kWaves][256];
	__shared__ uint32_t wsum[kWaves];
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	const size_t base = static_cast<size_t>(blockIdx.x) * kSortTile;
	const int tcount = n - base < static_cast<size_t>(kSortTile) ? static_cast<int>(n - base) : kSortTile;
	{
		const size_t idx = static_cast<size_t>(threadIdx.x) * nblocks + blockIdx.x;
		const uint32_t g0 = offsets[idx];
		const size_t m = static_cast<size_t>(nblocks) * 256;
		const uint32_t cnt = (idx + 1 < m ? offsets[idx + 1] : static_cast<uint32_t>(n)) - g0;
		goff[threadIdx.x] = g0;
		running[threadIdx.x] = 0;
		uint32_t x = cnt;
		for(int off = 1; off < 64; off <<= 1)
		{
			const uint32_t u = __shfl_up(x, off, 64);
			if(lane >= off) { x += u; }
		}
		if(lane == 63) { wsum[wave] = x; }
		__syncthreads();
		uint32_t ex = x - cnt;
		for(int w = 0; w < wave; ++w) { ex += wsum[w]; }
		loff[threadIdx.x] = ex;
	}
	const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
	for(int k = 0; k < kSortItems; ++k)
	{
		for(int w = 0; w < kWaves; ++w) { wcount[w][threadIdx.x] = 0; }
		__syncthreads();
		const int r = k * kSortThreads + threadIdx.x;
		const bool ok = r < tcount;
		const uint64_t key = ok ? kin[base + r] : 0;
		const uint32_t val = ok ? vin[base + r] : 0;
		const uint32_t d = static_cast<uint32_t>((key >> shift) & 0xFF);
		uint64_t peers = __ballot(ok);
		for(int bit = 0; bit < 8; ++bit)
		{
			const uint64_t bal = __ballot((d >> bit) & 1u);
			peers &= ((d >> bit) & 1u) ? bal : ~bal;
		}
		const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt_mask));
		const uint32_t cnt = static_cast<uint32_t>(__popcll(peers));
		if(ok && (peers & lt_mask) == 0) { wcount[wave][d] = cnt; }
		__syncthreads();
		if(ok)
		{
			uint32_t pos = loff[d] + running[d] + rank;
			for(int w = 0; w < wave; ++w) { pos += wcount[w][d]; }
			skey[pos] = key;
			sval[pos] = val;
		}
		__syncthreads();
		uint32_t add = 0;
		for(int w = 0; w < kWaves; ++w) { add += wcount[w][threadIdx.x]; }
		running[threadIdx.x] += add;
	}
	__syncthreads();
	for(int i = threadIdx.x; i < tcount; i += kSortThreads)
	{
		const uint64_t key = skey[i];
		const uint32_t d = static_cast<uint32_t>((key >> shift) & 0xFF);
		const size_t pos = static_cast<size_t>(goff[d]) + static_cast<uint32_t>(i - static_cast<int>(loff[d]));
		kout[pos] = key;
		vout[pos] = sval[i];
	}
}

// ---------------------------------------------------------------------------
// per-level split
// ---------------------------------------------------------------------------
// side[list_a[p]] = 1 for the upper half of p's segment
__global__ void __launch_bounds__(256)
k_mark_side(const uint32_t* __restrict__ list_a, uint8_t* __restrict__ side, size_t n, size_t seg)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	size_t half = seg >> 1;
	for(size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride)
	{
		side[list_a[p]] = (p & (seg - 1)) >= half ? 1 : 0;
	}
}

// rights[b] = number of upper-half bodies in tile b of `list`
__global__ void __launch_bounds__(kPartThreads)
k_part_count(const uint32_t* __restrict__ list, const uint8_t* __restrict__ side, size_t n,
             uint32_t* __restrict__ rights)
{
	__shared__ uint32_t wsum[kPartThreads / 64];
	size_t base = static_cast<size_t>(blockIdx.x) * kPartTile + static_cast<size_t>(threadIdx.x) * kPartItems;
	uint32_t c = 0;
	for(int k = 0; k < kPartItems; ++k)
	{
		size_t p = base + k;
		if(p < n) { c += side[list[p]]; }
	}
	for(int off = 32; off > 0; off >>= 1) { c += __shfl_down(c, off, 64); }
	if((threadIdx.x & 63) == 0) { wsum[threadIdx.x >> 6] = c; }
	__syncthreads();
	if(threadIdx.x == 0)
	{
		uint32_t t = 0;
		for(int w = 0; w < kPartThreads / 64; ++w) { t += wsum[w]; }
		rights[blockIdx.x] = t;
	}
}

// Stable partition of every segment of `list` into lower-half then upper-half
// bodies (flags in `side`), written to `out`.  `tile_off` holds, per tile, the
// number of upper-half bodies in all tiles before it (exclusive scan over the
// whole list); for segments of at least one tile the count inside the segment
// is that minus the value at the segment's first tile.  Smaller segments are
// handled inside the tile.
__global__ void __launch_bounds__(kPartThreads)
k_part_scatter(const uint32_t* __restrict__ list, const uint8_t* __restrict__ side, size_t n, size_t seg,
               const uint32_t* __restrict__ tile_off, uint32_t* __restrict__ out)
{
	__shared__ uint32_t excl[kPartTile];
	__shared__ uint32_t tsum[kPartThreads];
	const size_t tile = static_cast<size_t>(blockIdx.x) * kPartTile;
	const int tb = threadIdx.x * kPartItems;
	uint32_t body[kPartItems];
	uint8_t fl[kPartItems];
	uint32_t c = 0;
	for(int k = 0; k < kPartItems; ++k)
	{
		size_t p = tile + tb + k;
		body[k] = p < n ? list[p] : 0;
		fl[k] = p < n ? side[body[k]] : 0;
		excl[tb + k] = c;
		c += fl[k];
	}
	tsum[threadIdx.x] = c;
	__syncthreads();
	for(int off = 1; off < kPartThreads; off <<= 1)
	{
		uint32_t v = threadIdx.x >= static_cast<unsigned>(off) ? tsum[threadIdx.x - off] : 0;
		__syncthreads();
		tsum[threadIdx.x] += v;
		__syncthreads();
	}
	uint32_t add = threadIdx.x == 0 ? 0 : tsum[threadIdx.x - 1];
	for(int k = 0; k < kPartItems; ++k) { excl[tb + k] += add; }
	__syncthreads();
	const size_t half = seg >> 1;
	for(int k = 0; k < kPartItems; ++k)
	{
		size_t p = tile + tb + k;
		if(p >= n) { continue; }
		size_t seg_start = p & ~(seg - 1);
		uint32_t before;
		if(seg >= static_cast<size_t>(kPartTile))
		{
			before = tile_off[blockIdx.x] - tile_off[seg_start / kPartTile] + excl[tb + k];
		}
		else
		{
			before = excl[tb + k] - excl[seg_start - tile];
		}
		size_t pos = fl[k] ? seg_start + half + before : p - before;
		out[pos] = body[k];
	}
}

// Levels first_level .. levels-1 of the split (the steps of k_mark_side and
// k_part_scatter) for segments of seg0 <= kLocalSeg bodies, one work-group per
// segment, in LDS.  Bodies get local numbers (their position in list0's
// segment, through the global table g2l) so that the side flags live in LDS.
// Only the final order is written back, to list0: after the last level every
// segment holds one body and the three lists agree.
__global__ void __launch_bounds__(kLocalThreads)
k_split_local(uint32_t* __restrict__ list0, const uint32_t* __restrict__ list1, const uint32_t* __restrict__ list2,
              uint32_t* __restrict__ g2l, int seg0, int first_level, int levels)
{
	__shared__ uint16_t ls[3][kLocalSeg];
	__shared__ uint16_t tmp[kLocalSeg];
	__shared__ uint16_t excl[kLocalSeg];
	__shared__ uint32_t l2g[kLocalSeg];
	__shared__ uint8_t side[kLocalSeg];
	__shared__ uint32_t ws[kLocalThreads / 64];
	const size_t base = static_cast<size_t>(blockIdx.x) * seg0;
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	for(int p = threadIdx.x; p < seg0; p += kLocalThreads)
	{
		const uint32_t b = list0[base + p];
		l2g[p] = b;
		ls[0][p] = static_cast<uint16_t>(p);
		g2l[b] = static_cast<uint32_t>(p);
	}
	__syncthreads();
	for(int p = threadIdx.x; p < seg0; p += kLocalThreads)
	{
		ls[1][p] = static_cast<uint16_t>(g2l[list1[base + p]]);
		ls[2][p] = static_cast<uint16_t>(g2l[list2[base + p]]);
	}
	__syncthreads();
	// work-item t owns positions [t * per, t * per + per) of a list
	const int per = (seg0 + kLocalThreads - 1) / kLocalThreads;
	const int p0 = static_cast<int>(threadIdx.x) * per;
	for(int level = first_level; level < levels; ++level)
	{
		const int seg = seg0 >> (level - first_level);
		const int half = seg >> 1;
		const int a = level % 3;
		for(int p = threadIdx.x; p < seg0; p += kLocalThreads)
		{
			side[ls[a][p]] = (p & (seg - 1)) >= half ? 1 : 0;
		}
		__syncthreads();
		for(int b = 0; b < 3; ++b)
		{
			if(b == a) { continue; }
			uint32_t c = 0;
			for(int k = 0; k < per; ++k)
			{
				const int p = p0 + k;
				if(p < seg0)
				{
					excl[p] = static_cast<uint16_t>(c);
					c += side[ls[b][p]];
				}
			}
			uint32_t x = c;
			for(int off = 1; off < 64; off <<= 1)
			{
				const uint32_t u = __shfl_up(x, off, 64);
				if(lane >= off) { x += u; }
			}
			if(lane == 63) { ws[wave] = x; }
			__syncthreads();
			uint32_t add = x - c;
			for(int w = 0; w < wave; ++w) { add += ws[w]; }
			for(int k = 0; k < per; ++k)
			{
				const int p = p0 + k;
				if(p < seg0) { excl[p] = static_cast<uint16_t>(excl[p] + add); }
			}
			__syncthreads();
			for(int k = 0; k < per; ++k)
			{
				const int p = p0 + k;
				if(p < seg0)
				{
					const uint16_t id = ls[b][p];
					const int seg_start = p & ~(seg - 1);
					const int before = excl[p] - excl[seg_start];
					tmp[side[id] ? seg_start + half + before : p - before] = id;
				}
			}
			__syncthreads();
			for(int k = 0; k < per; ++k)
			{
				const int p = p0 + k;
				if(p < seg0) { ls[b][p] = tmp[p]; }
			}
			__syncthreads();
		}
	}
	for(int p = threadIdx.x; p < seg0; p += kLocalThreads) { list0[base + p] = l2g[ls[0][p]]; }
}

// ---------------------------------------------------------------------------
// moments (nbody/nbody_space_heap.cpp:119-134), FP64, contraction off
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_leaves(const uint32_t* __restrict__ order, const double* __restrict__ y, const double* __restrict__ mass,
         size_t n, double4* __restrict__ xyzr, double* __restrict__ nmass, double* __restrict__ bmin,
         double* __restrict__ bmax, int32_t* __restrict__ body_n, int set_body)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t s = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < n; s += stride)
	{
		size_t idx = n + s;
		uint32_t b = set_body ? order[s] : static_cast<uint32_t>(body_n[idx]);
		double x = y[b], yy = y[n + b], z = y[2 * n + b];
		xyzr[idx] = make_double4(x, yy, z, 0.0);
		if(set_body)
		{
			body_n[idx] = static_cast<int32_t>(b);
			nmass[idx] = mass[b];
		}
		bmin[3 * idx] = x; bmin[3 * idx + 1] = yy; bmin[3 * idx + 2] = z;
		bmax[3 * idx] = x; bmax[3 * idx + 1] = yy; bmax[3 * idx + 2] = z;
	}
}

__device__ inline double min_ref(double a, double b) { return (b < a) ? b : a; }   // std::min
__device__ inline double max_ref(double a, double b) { return (a < b) ? b : a; }   // std::max

// A node's moments from its children's (mass centre cl / cr, masses ml / mr,
// boxes lol..hil / lor..hir): mass, mass centre, box and crit_r2 = (radius *
// ratio)^2 with the heap's operation order.
__device__ inline void combine_node(const double4& cl, double ml, const double* lol, const double* hil,
                                    const double4& cr, double mr, const double* lor, const double* hir,
                                    double ratio_sqr, double4& c, double& m, double* lo, double* hi)
{
	m = ml + mr;
	double cx = (cl.x * ml + cr.x * mr) / m;
	double cy = (cl.y * ml + cr.y * mr) / m;
	double cz = (cl.z * ml + cr.z * mr) / m;
	for(int d = 0; d < 3; ++d)
	{
		lo[d] = min_ref(lol[d], lor[d]);
		hi[d] = max_ref(hil[d], hir[d]);
	}
	double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
	double gx = (hi[0] + lo[0]) / 2.0 - cx;
	double gy = (hi[1] + lo[1]) / 2.0 - cy;
	double gz = (hi[2] + lo[2]) / 2.0 - cz;
	double rad = sqrt(ex * ex + ey * ey + ez * ez) * 0.5 + sqrt(gx * gx + gy * gy + gz * gz);
	c = make_double4(cx, cy, cz, (rad * rad) * ratio_sqr);
}

// Node idx from its children l and r (2 idx and 2 idx + 1 in the heap).
__device__ inline void update_node(size_t idx, size_t l, size_t r, double4* __restrict__ xyzr,
                                   double* __restrict__ nmass, double* __restrict__ bmin, double* __restrict__ bmax,
                                   double ratio_sqr)
{
	double4 c;
	double m, lo[3], hi[3];
	combine_node(xyzr[l], nmass[l], bmin + 3 * l, bmax + 3 * l, xyzr[r], nmass[r], bmin + 3 * r, bmax + 3 * r,
	             ratio_sqr, c, m, lo, hi);
	for(int d = 0; d < 3; ++d)
	{
		bmin[3 * idx + d] = lo[d];
		bmax[3 * idx + d] = hi[d];
	}
	nmass[idx] = m;
	xyzr[idx] = c;
}

// Second central moments Q = sum m (x - cm)(x - cm)^T of node idx from its
// children (parallel-axis theorem); components xx yy zz xy xz yz.  Not part of
// the reference tree: input of the learned correction.
__device__ inline void combine_quad(const double4& c, const double4& cl, double ml, const double* ql,
                                    const double4& cr, double mr, const double* qr, double* q)
{
	for(int k = 0; k < 6; ++k) { q[k] = 0.0; }
	for(int side = 0; side < 2; ++side)
	{
		const double4 cc = side == 0 ? cl : cr;
		const double m = side == 0 ? ml : mr;
		const double* qc = side == 0 ? ql : qr;
		const double dx = cc.x - c.x, dy = cc.y - c.y, dz = cc.z - c.z;
		q[0] += qc[0] + m * dx * dx;
		q[1] += qc[1] + m * dy * dy;
		q[2] += qc[2] + m * dz * dz;
		q[3] += qc[3] + m * dx * dy;
		q[4] += qc[4] + m * dx * dz;
		q[5] += qc[5] + m * dy * dz;
	}
}

__device__ inline void update_quad(size_t idx, size_t l, size_t r, const double4* __restrict__ xyzr,
                                   const double* __restrict__ nmass, double* __restrict__ quad)
{
	double q[6];
	combine_quad(xyzr[idx], xyzr[l], nmass[l], quad + 6 * l, xyzr[r], nmass[r], quad + 6 * r, q);
	for(int k = 0; k < 6; ++k) { quad[6 * idx + k] = q[k]; }
}

// nodes [lo, hi) of one level
__global__ void __launch_bounds__(256)
k_level(size_t lo, size_t hi, double4* __restrict__ xyzr, double* __restrict__ nmass, double* __restrict__ bmin,
        double* __restrict__ bmax, double* __restrict__ quad, double ratio_sqr)
{
	size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t idx = lo + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < hi; idx += stride)
	{
		update_node(idx, 2 * idx, 2 * idx + 1, xyzr, nmass, bmin, bmax, ratio_sqr);
		update_quad(idx, 2 * idx, 2 * idx + 1, xyzr, nmass, quad);
	}
}

// all levels with fewer than 2 * 256 nodes in one work-group, bottom-up
__global__ void __launch_bounds__(256)
k_levels_top(size_t top, double4* __restrict__ xyzr, double* __restrict__ nmass, double* __restrict__ bmin,
             double* __restrict__ bmax, double* __restrict__ quad, double ratio_sqr)
{
	for(size_t lo = top; lo >= 1; lo >>= 1)
	{
		for(size_t idx = lo + threadIdx.x; idx < 2 * lo; idx += blockDim.x)
		{
			update_node(idx, 2 * idx, 2 * idx + 1, xyzr, nmass, bmin, bmax, ratio_sqr);
			update_quad(idx, 2 * idx, 2 * idx + 1, xyzr, nmass, quad);
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// Morton radix tree (the fast path of the FP32 group walks; no counterpart in
// the reference, whose heap kd-tree stays the tree of the parity walks)
// ---------------------------------------------------------------------------
// Bodies sorted by 63-bit Morton codes (21 bits per axis over the bounding
// cube), ties broken by body index; internal nodes by the binary radix tree
// construction of Karras (2012, "Maximizing parallelism in the construction of
// BVHs, octrees, and k-d trees"): internal node i covers a key range with i at
// one end and splits it at the highest bit where its keys differ.  Node ids
// follow the heap's conventions where the walks rely on them: the root is id 1
// (internal node i is id i + 1), the body at sorted position j is leaf id n + j,
// and ids >= n are leaves.  Cells are bounded by the boxes of their bodies,
// with the heap's opening radius, so a cell of the Morton tree is compact where
// a count-balanced heap node of a centrally concentrated state spans from the
// core to the outskirts.

// Spread the low 21 bits of v to every third bit.
__device__ inline uint64_t morton_spread(uint64_t v)
{
	v &= 0x1fffffull;
	v = (v | (v << 32)) & 0x1f00000000ffffull;
	v = (v | (v << 16)) & 0x1f0000ff0000ffull;
	v = (v | (v << 8)) & 0x100f00f00f00f00full;
	v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
	v = (v | (v << 2)) & 0x1249249249249249ull;
	return v;
}

// Cell coordinate of x on a grid of 2^21 cells over [lo, lo + 2^21 / scale).
__device__ inline uint64_t morton_cell(double x, double lo, double scale)
{
	double c = floor((x - lo) * scale);
	c = c < 0.0 ? 0.0 : c;
	c = c > 2097151.0 ? 2097151.0 : c;
	return static_cast<uint64_t>(c);
}

// box[6] = bounding box lo.xyz hi.xyz of the bodies: scale = 2^21 / longest side
__global__ void __launch_bounds__(256)
k_morton_codes(const double* __restrict__ y, const double* __restrict__ mass, size_t n, const double* __restrict__ box,
               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, double4* __restrict__ pk)
{
	const double side = fmax(fmax(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
	const double scale = side > 0.0 ? 2097152.0 / side : 0.0;
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		const double x = y[i], yy = y[n + i], z = y[2 * n + i];
		const uint64_t cx = morton_cell(x, box[0], scale);
		const uint64_t cy = morton_cell(yy, box[1], scale);
		const uint64_t cz = morton_cell(z, box[2], scale);
		keys[i] = (morton_spread(cx) << 2) | (morton_spread(cy) << 1) | morton_spread(cz);
		vals[i] = static_cast<uint32_t>(i);
		pk[i] = make_double4(x, yy, z, mass[i]);
	}
}

// Bounding box partials of the positions, then the box (lo.xyz hi.xyz).
__global__ void __launch_bounds__(256)
k_box_partial(const double* __restrict__ y, size_t n, double* __restrict__ part)
{
	__shared__ double sh[6][256];
	double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
	{
		for(int d = 0; d < 3; ++d)
		{
			const double v = y[static_cast<size_t>(d) * n + i];
			lo[d] = fmin(lo[d], v);
			hi[d] = fmax(hi[d], v);
		}
	}
	for(int d = 0; d < 3; ++d) { sh[d][threadIdx.x] = lo[d]; sh[3 + d][threadIdx.x] = hi[d]; }
	__syncthreads();
	for(int w = 128; w > 0; w >>= 1)
	{
		if(static_cast<int>(threadIdx.x) < w)
		{
			for(int d = 0; d < 3; ++d)
			{
				sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + w]);
				sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + w]);
			}
		}
		__syncthreads();
	}
	if(threadIdx.x < 6) { part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0]; }
}

// The partials are taken strided by 256 threads, then reduced in LDS (min and
// max are exact, so the order does not change the box).
__global__ void __launch_bounds__(256)
k_box_final(const double* __restrict__ part, int count, double* __restrict__ box)
{
	__shared__ double sh[6][256];
	double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	for(int k = threadIdx.x; k < count; k += blockDim.x)
	{
		for(int d = 0; d < 3; ++d)
		{
			lo[d] = fmin(lo[d], part[6 * k + d]);
			hi[d] = fmax(hi[d], part[6 * k + 3 + d]);
		}
	}
	for(int d = 0; d < 3; ++d) { sh[d][threadIdx.x] = lo[d]; sh[3 + d][threadIdx.x] = hi[d]; }
	__syncthreads();
	for(int w = 128; w > 0; w >>= 1)
	{
		if(static_cast<int>(threadIdx.x) < w)
		{
			for(int d = 0; d < 3; ++d)
			{
				sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + w]);
				sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + w]);
			}
		}
		__syncthreads();
	}
	if(threadIdx.x < 6) { box[threadIdx.x] = sh[threadIdx.x][0]; }
}

// Length of the common prefix of sorted keys i and j (with the index as a
// tie-breaking suffix), -1 outside [0, n).
__device__ inline int radix_delta(const uint64_t* __restrict__ k, long long n, long long i, long long j)
{
	if(j < 0 || j >= n) { return -1; }
	const uint64_t a = k[i], b = k[j];
	if(a == b)
	{
		return 64 + __clzll(static_cast<unsigned long long>(i ^ j));
	}
	return __clzll(a ^ b);
}

// Groups of the Morton tree, all maximal subtrees: build groups of at most
// kRadixGroup leaves (one wave computes their nodes' moments), walk groups of at
// most kRadixWalk leaves (one wave walks for them, two bodies per lane), and
// super-groups of at most kRadixSuper leaves that hold more than one walk group
// (a walk group in no super-group walks in one level).  Their leaves are
// consecutive in Morton order and their boxes are node boxes, so they never
// straddle the core of the state the way a fixed run of the Morton order can.
constexpr int kRadixGroup = 64;
constexpr int kRadixWalk = 128;
constexpr int kRadixSuper = 2048;
constexpr int kRadixSuperMin = kRadixWalk + 1;   // at least two walk groups

// Internal node i of [0, n - 1): its key range, split and children (ids), the
// parent links of the children, its level (the common prefix length of its
// keys, which grows strictly from a node to its children, so nodes of one level
// are independent; -1 inside a walk group, whose nodes k_radix_group_moments
// takes), the count of each level, and the marks of walk groups and
// super-groups at the first sorted position of their leaves: (node id, leaves).
__global__ void __launch_bounds__(256)
k_radix_nodes(const uint64_t* __restrict__ k, long long n, int2* __restrict__ child, int* __restrict__ parent,
              int* __restrict__ level, int* __restrict__ level_count, int2* __restrict__ gmark,
              int2* __restrict__ smark, int2* __restrict__ wmark)
{
	__shared__ int hist[kRadixLevels];
	for(int b = threadIdx.x; b < kRadixLevels; b += blockDim.x) { hist[b] = 0; }
	__syncthreads();
	const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
	for(long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n - 1; i += stride)
	{
		const int dir = radix_delta(k, n, i, i + 1) - radix_delta(k, n, i, i - 1) > 0 ? 1 : -1;
		const int dmin = radix_delta(k, n, i, i - dir);
		long long lmax = 2;
		while(radix_delta(k, n, i, i + lmax * dir) > dmin) { lmax <<= 1; }
		long long l = 0;
		for(long long t = lmax >> 1; t >= 1; t >>= 1)
		{
			if(radix_delta(k, n, i, i + (l + t) * dir) > dmin) { l += t; }
		}
		const long long j = i + l * dir;
		const int dnode = radix_delta(k, n, i, j);
		long long s = 0;
		for(long long t = (l + 1) >> 1;; t = (t + 1) >> 1)
		{
			if(s + t < l + 1 && radix_delta(k, n, i, i + (s + t) * dir) > dnode) { s += t; }
			if(t <= 1) { break; }
		}
		const long long gamma = i + s * dir + (dir < 0 ? -1 : 0);
		const long long lo = i < j ? i : j, hi = i < j ? j : i;
		const int left = lo == gamma ? static_cast<int>(n + gamma) : static_cast<int>(gamma + 1);
		const int right = hi == gamma + 1 ? static_cast<int>(n + gamma + 1) : static_cast<int>(gamma + 2);
		const int id = static_cast<int>(i + 1);
		child[id] = make_int2(left, right);
		parent[left] = id;
		parent[right] = id;
		const long long cnt = hi - lo + 1;
		level[id] = cnt > kRadixGroup ? dnode : -1;
		if(cnt > kRadixGroup) { atomicAdd(&hist[dnode], 1); }
		const long long cl = gamma - lo + 1, cr = hi - gamma;
		if(id == kRoot && cnt <= kRadixGroup) { gmark[lo] = make_int2(id, static_cast<int>(cnt)); }
		if(id == kRoot && cnt <= kRadixSuper && cnt >= kRadixSuperMin) { smark[lo] = make_int2(id, static_cast<int>(cnt)); }
		if(cnt > kRadixGroup)
		{
			if(cl <= kRadixGroup) { gmark[lo] = make_int2(left, static_cast<int>(cl)); }
			if(cr <= kRadixGroup) { gmark[gamma + 1] = make_int2(right, static_cast<int>(cr)); }
		}
		if(id == kRoot && cnt <= kRadixWalk) { wmark[lo] = make_int2(id, static_cast<int>(cnt)); }
		if(cnt > kRadixWalk)
		{
			if(cl <= kRadixWalk) { wmark[lo] = make_int2(left, static_cast<int>(cl)); }
			if(cr <= kRadixWalk) { wmark[gamma + 1] = make_int2(right, static_cast<int>(cr)); }
		}
		if(cnt > kRadixSuper)
		{
			if(cl <= kRadixSuper && cl >= kRadixSuperMin) { smark[lo] = make_int2(left, static_cast<int>(cl)); }
			if(cr <= kRadixSuper && cr >= kRadixSuperMin) { smark[gamma + 1] = make_int2(right, static_cast<int>(cr)); }
		}
	}
	__syncthreads();
	for(int b = threadIdx.x; b < kRadixLevels; b += blockDim.x)
	{
		if(hist[b] != 0) { atomicAdd(&level_count[b], hist[b]); }
	}
}

__global__ void __launch_bounds__(256)
k_fill_int(int* __restrict__ a, size_t count, int v)
{
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) { a[i] = v; }
}

// 1 where a mark starts a group / super-group / walk group, for the scans that
// number them.
__global__ void __launch_bounds__(256)
k_radix_flags(const int2* __restrict__ gmark, const int2* __restrict__ smark, const int2* __restrict__ wmark,
              size_t n, uint32_t* __restrict__ gflag, uint32_t* __restrict__ sflag, uint32_t* __restrict__ wflag)
{
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride)
	{
		gflag[p] = gmark[p].x != 0 ? 1u : 0u;
		sflag[p] = smark[p].x != 0 ? 1u : 0u;
		wflag[p] = wmark[p].x != 0 ? 1u : 0u;
	}
}

// Build groups grp[g] and walk groups wgrp[w] = (node id, first sorted leaf,
// leaves, super-group or -1), super-groups sup[s] = (node id, first sorted
// leaf), all in Morton order (gidx / sidx / widx: exclusive scans of the flags);
// info[0] = build groups, info[1] = super-groups, info[2] = walk groups.
__global__ void __launch_bounds__(256)
k_radix_compact(const int2* __restrict__ gmark, const int2* __restrict__ smark, const int2* __restrict__ wmark,
                const uint32_t* __restrict__ gidx, const uint32_t* __restrict__ sidx,
                const uint32_t* __restrict__ widx, size_t n, int4* __restrict__ grp, int2* __restrict__ sup,
                int4* __restrict__ wgrp, int* __restrict__ info)
{
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < n; p += stride)
	{
		const int2 g = gmark[p];
		if(g.x != 0) { grp[gidx[p]] = make_int4(g.x, static_cast<int>(p), g.y, -1); }
		const int2 sm = smark[p];
		if(sm.x != 0) { sup[sidx[p]] = make_int2(sm.x, static_cast<int>(p)); }
		const int2 w = wmark[p];
		if(w.x != 0) { wgrp[widx[p]] = make_int4(w.x, static_cast<int>(p), w.y, -1); }
		if(p == n - 1)
		{
			info[0] = static_cast<int>(gidx[p] + (g.x != 0 ? 1u : 0u));
			info[1] = static_cast<int>(sidx[p] + (sm.x != 0 ? 1u : 0u));
			info[2] = static_cast<int>(widx[p] + (w.x != 0 ? 1u : 0u));
		}
	}
}

// The super-group of each walk group: the last super-group starting at or
// before the group's first leaf, if its node's leaves reach the group (a walk
// group lies inside at most one super-group: both are subtrees).
__global__ void __launch_bounds__(256)
k_radix_group_super(int4* __restrict__ grp, const int2* __restrict__ sup, const int* __restrict__ info,
                    const int2* __restrict__ child, size_t n)
{
	const int ng = info[2], ns = info[1];
	for(int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x)
	{
		const int first = grp[g].y;
		int lo = 0, hi = ns - 1, s = -1;
		while(lo <= hi)
		{
			const int mid = (lo + hi) >> 1;
			if(sup[mid].y <= first) { s = mid; lo = mid + 1; }
			else { hi = mid - 1; }
		}
		if(s >= 0)
		{
			// leaves of the super-group's node: its first leaf and its size, found
			// by walking its rightmost branch down to a leaf
			int node = sup[s].x;
			while(node < static_cast<int>(n)) { node = child[node].y; }
			const int last = node - static_cast<int>(n);
			if(first > last) { s = -1; }
		}
		grp[g].w = s;
	}
}

// Offsets of the levels (exclusive scan of kRadixLevels counts, one work-group).
__global__ void __launch_bounds__(kRadixLevels)
k_radix_level_offsets(const int* __restrict__ count, int* __restrict__ offset, int* __restrict__ cursor)
{
	__shared__ int sh[kRadixLevels];
	const int t = threadIdx.x;
	sh[t] = count[t];
	__syncthreads();
	if(t == 0)
	{
		int run = 0;
		for(int b = 0; b < kRadixLevels; ++b)
		{
			const int c = sh[b];
			sh[b] = run;
			run += c;
		}
	}
	__syncthreads();
	offset[t] = sh[t];
	cursor[t] = 0;
}

// Internal ids above the walk groups listed level by level, kRadixListTile ids
// per work-group: counts per level in LDS, one global reservation per level
// present, then ranks from LDS counters.
constexpr int kRadixListTile = 4096;
__global__ void __launch_bounds__(256)
k_radix_level_lists(const int* __restrict__ level, size_t n, const int* __restrict__ offset, int* __restrict__ cursor,
                    int* __restrict__ list)
{
	__shared__ int cnt[kRadixLevels];
	__shared__ int base[kRadixLevels];
	for(int b = threadIdx.x; b < kRadixLevels; b += blockDim.x) { cnt[b] = 0; }
	__syncthreads();
	const size_t lo = 1 + static_cast<size_t>(blockIdx.x) * kRadixListTile;
	const size_t hi = lo + kRadixListTile < n ? lo + kRadixListTile : n;
	for(size_t id = lo + threadIdx.x; id < hi; id += blockDim.x)
	{
		const int lv = level[id];
		if(lv >= 0) { atomicAdd(&cnt[lv], 1); }
	}
	__syncthreads();
	for(int b = threadIdx.x; b < kRadixLevels; b += blockDim.x)
	{
		base[b] = cnt[b] != 0 ? offset[b] + atomicAdd(&cursor[b], cnt[b]) : 0;
		cnt[b] = 0;
	}
	__syncthreads();
	for(size_t id = lo + threadIdx.x; id < hi; id += blockDim.x)
	{
		const int lv = level[id];
		if(lv >= 0) { list[base[lv] + atomicAdd(&cnt[lv], 1)] = static_cast<int>(id); }
	}
}

// Walk record of node id from its moments (as k_pack_records).
__device__ inline void put_record(int id, const double4* __restrict__ xyzr, const double* __restrict__ nmass,
                                  const double* __restrict__ quad, nbh_walk_rec* __restrict__ wr)
{
	nbh_walk_rec r;
	r.c = xyzr[id];
	r.m = nmass[id];
	for(int k = 0; k < 6; ++k) { r.q[k] = static_cast<float>(quad[6 * static_cast<size_t>(id) + k]); }
	wr[id] = r;
}

// Moments of the internal ids list[0, count) of one level, from their children
// (the heap's update_node and update_quad); the levels run from the longest
// prefix to the root, one launch each.
__global__ void __launch_bounds__(256)
k_radix_level(const int* __restrict__ list, int count, const int2* __restrict__ child, double4* __restrict__ xyzr,
              double* __restrict__ nmass, double* __restrict__ bmin, double* __restrict__ bmax,
              double* __restrict__ quad, double ratio_sqr, nbh_walk_rec* __restrict__ wr)
{
	const int stride = gridDim.x * blockDim.x;
	for(int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride)
	{
		const int id = list[k];
		const int2 c = child[id];
		update_node(id, c.x, c.y, xyzr, nmass, bmin, bmax, ratio_sqr);
		update_quad(id, c.x, c.y, xyzr, nmass, quad);
		put_record(id, xyzr, nmass, quad, wr);
	}
}

// Leaves and internal nodes of each walk group (a subtree of at most
// kRadixGroup leaves), one wave per group, with their walk records.  The
// group's leaves (as k_leaves; kBuild: the body at sorted position s is
// order[s], read from its packed record; refit: the body numbers of the build,
// positions from y and the masses of the leaf records) and node moments are held in LDS: leaf
// first + j in slot j, internal id k in slot kRadixGroup + k - first - 1 (the
// internal ids of a subtree over the sorted leaves [a, b] lie in [a + 1, b + 1]:
// each internal node's index is one end of its key range).  The internal nodes
// are listed breadth-first, then taken depth by depth from the deepest (the
// heap's update_node and update_quad on the LDS slots), and written out.
// lean: every node's walk record is written (and each leaf's body number), the
// FP64 node arrays only for the group's root, which the levels above read, or
// for a group of one leaf.
template <bool kBuild>
__global__ void __launch_bounds__(kRadixGroup)
k_radix_group_moments(const int4* __restrict__ grp, int n, const int2* __restrict__ child,
                      const uint32_t* __restrict__ order, const double4* __restrict__ pk, const double* __restrict__ y,
                      int32_t* __restrict__ body_n, double4* __restrict__ xyzr, double* __restrict__ nmass,
                      double* __restrict__ bmin, double* __restrict__ bmax, double* __restrict__ quad,
                      double ratio_sqr, nbh_walk_rec* __restrict__ wr, int lean)
{
	// leaves: position (box and mass centre) and mass; internal slots: all moments
	__shared__ double4 s_lc[kRadixGroup];
	__shared__ double s_lm[kRadixGroup];
	__shared__ double4 s_c[kRadixGroup];
	__shared__ double s_m[kRadixGroup];
	__shared__ double s_lo[3 * kRadixGroup];
	__shared__ double s_hi[3 * kRadixGroup];
	__shared__ double s_q[6 * kRadixGroup];
	__shared__ int2 s_child[kRadixGroup];
	__shared__ int q[kRadixGroup];          // internal nodes, breadth-first (at most kRadixGroup - 1)
	__shared__ int start[kRadixGroup + 1];  // first entry of each depth
	__shared__ double s_zero[6];            // the second moments of a leaf
	const int lane = threadIdx.x;
	const int4 gr = grp[blockIdx.x];
	const int root = gr.x, first = gr.y, cnt = gr.z;
	if(lane < 6) { s_zero[lane] = 0.0; }
	if(lane < cnt)
	{
		const size_t sidx = static_cast<size_t>(first) + lane;
		const size_t idx = static_cast<size_t>(n) + sidx;
		double x, yy, z, m;
		if constexpr(kBuild)
		{
			const uint32_t b = order[sidx];
			const double4 p = pk[b];
			x = p.x;
			yy = p.y;
			z = p.z;
			m = p.w;
			body_n[idx] = static_cast<int32_t>(b);
			if(!lean || root >= n) { nmass[idx] = m; }
		}
		else
		{
			// the mass from the leaf's walk record, which every build writes
			const size_t b = static_cast<uint32_t>(body_n[idx]);
			x = y[b];
			yy = y[n + b];
			z = y[2 * static_cast<size_t>(n) + b];
			m = wr[idx].m;
		}
		const double4 c = make_double4(x, yy, z, 0.0);
		if(!lean || root >= n)
		{
			xyzr[idx] = c;
			bmin[3 * idx] = x; bmin[3 * idx + 1] = yy; bmin[3 * idx + 2] = z;
			bmax[3 * idx] = x; bmax[3 * idx + 1] = yy; bmax[3 * idx + 2] = z;
		}
		nbh_walk_rec r;
		r.c = c;
		r.m = m;
		for(int k = 0; k < 6; ++k) { r.q[k] = 0.0f; }
		wr[idx] = r;
		s_lc[lane] = c;
		s_lm[lane] = m;
		const int id = first + 1 + lane;
		s_child[lane] = id < n ? child[id] : make_int2(0, 0);
	}
	if(root >= n) { return; }              // a group of one leaf
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	if(lane == 0)
	{
		q[0] = root;
		start[0] = 0;
	}
	int head = 0, tail = 1, depth = 0;
	__syncthreads();
	while(head < tail)
	{
		const int width = tail - head;
		const bool act = lane < width;
		const int2 c = act ? s_child[q[head + lane] - first - 1] : make_int2(n, n);
		const bool li = c.x < n, ri = c.y < n;
		const uint64_t bl = __ballot(li), br = __ballot(ri);
		int pos = tail + __popcll(bl & lt) + __popcll(br & lt);
		if(li) { q[pos++] = c.x; }
		if(ri) { q[pos] = c.y; }
		head = tail;
		tail += __popcll(bl) + __popcll(br);
		++depth;
		if(lane == 0) { start[depth] = head; }
		__syncthreads();
	}
	// depth d holds q[start[d], start[d + 1]); start[depth] == tail.  A child is
	// leaf first + j (id n + first + j) or internal slot id - first - 1.
	for(int d = depth - 1; d >= 0; --d)
	{
		const int i = start[d] + lane;
		if(i < start[d + 1])
		{
			const int id = q[i];
			const int k = id - first - 1;
			const int2 ch = s_child[k];
			double4 cc[2];
			double mm[2];
			double lo2[2][3], hi2[2][3];
			const double* qq[2];
			for(int side = 0; side < 2; ++side)
			{
				const int cid = side == 0 ? ch.x : ch.y;
				if(cid >= n)
				{
					const int j = cid - n - first;
					cc[side] = s_lc[j];
					mm[side] = s_lm[j];
					lo2[side][0] = hi2[side][0] = cc[side].x;
					lo2[side][1] = hi2[side][1] = cc[side].y;
					lo2[side][2] = hi2[side][2] = cc[side].z;
					qq[side] = s_zero;
				}
				else
				{
					const int j = cid - first - 1;
					cc[side] = s_c[j];
					mm[side] = s_m[j];
					for(int e = 0; e < 3; ++e)
					{
						lo2[side][e] = s_lo[3 * j + e];
						hi2[side][e] = s_hi[3 * j + e];
					}
					qq[side] = s_q + 6 * j;
				}
			}
			double4 c;
			double m, lo[3], hi[3], qv[6];
			combine_node(cc[0], mm[0], lo2[0], hi2[0], cc[1], mm[1], lo2[1], hi2[1], ratio_sqr, c, m, lo, hi);
			combine_quad(c, cc[0], mm[0], qq[0], cc[1], mm[1], qq[1], qv);
			s_c[k] = c;
			s_m[k] = m;
			const size_t gid = static_cast<size_t>(id);
			nbh_walk_rec rec;
			rec.c = c;
			rec.m = m;
			for(int e = 0; e < 3; ++e)
			{
				s_lo[3 * k + e] = lo[e];
				s_hi[3 * k + e] = hi[e];
			}
			for(int e = 0; e < 6; ++e)
			{
				s_q[6 * k + e] = qv[e];
				rec.q[e] = static_cast<float>(qv[e]);
			}
			wr[gid] = rec;
			if(!lean || id == root)
			{
				xyzr[gid] = c;
				nmass[gid] = m;
				for(int e = 0; e < 3; ++e)
				{
					bmin[3 * gid + e] = lo[e];
					bmax[3 * gid + e] = hi[e];
				}
				for(int e = 0; e < 6; ++e) { quad[6 * gid + e] = qv[e]; }
			}
		}
		__syncthreads();
	}
}

// Levels lv_hi down to lv_lo, each too small to fill the device, in one
// work-group (info: the level counts, then their offsets in list).
__global__ void __launch_bounds__(1024)
k_radix_levels_small(const int* __restrict__ list, const int* __restrict__ info, int lv_hi, int lv_lo,
                     const int2* __restrict__ child, double4* __restrict__ xyzr, double* __restrict__ nmass,
                     double* __restrict__ bmin, double* __restrict__ bmax, double* __restrict__ quad,
                     double ratio_sqr, nbh_walk_rec* __restrict__ wr)
{
	for(int lv = lv_hi; lv >= lv_lo; --lv)
	{
		const int count = info[lv], off = info[kRadixLevels + lv];
		for(int k = threadIdx.x; k < count; k += blockDim.x)
		{
			const int id = list[off + k];
			const int2 c = child[id];
			update_node(id, c.x, c.y, xyzr, nmass, bmin, bmax, ratio_sqr);
			update_quad(id, c.x, c.y, xyzr, nmass, quad);
			put_record(id, xyzr, nmass, quad, wr);
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------
// heap index algebra (nbody/nbody_space_heap_func_priv.h:4-72)
__device__ inline int next_down(int idx)
{
	idx = idx >> __builtin_ctz(~static_cast<unsigned>(idx));
	return idx + 1;
}

__device__ inline int next_up(int idx, int tree_size)
{
	int left = idx << 1;
	return left < tree_size ? left : next_down(idx);
}

// Counters are cleared by a kernel rather than a memset node, so a captured
// sequence of calls replays correctly.
__global__ void k_zero_counters(unsigned long long* counters)
{
	if(threadIdx.x < 5) { counters[threadIdx.x] = 0; }
}

__device__ inline void count_add(unsigned long long* counters, unsigned long long inter, unsigned long long vis)
{
	for(int off = 32; off > 0; off >>= 1)
	{
		inter += __shfl_down(inter, off, 64);
		vis += __shfl_down(vis, off, 64);
	}
	if((threadIdx.x & 63) == 0)
	{
		atomicAdd(&counters[0], inter);
		atomicAdd(&counters[1], vis);
	}
}

// FP64 monopole term of nbody_data::force: tot += (v1 - cm) * (-(m1 M) / r^3)
__device__ inline void add_force64(double x1, double y1, double z1, double gm1, const double4& nd, double M,
                                   double& tx, double& ty, double& tz)
{
	double dx = x1 - nd.x;
	double dy = y1 - nd.y;
	double dz = z1 - nd.z;
	double r2 = dx * dx + dy * dy + dz * dz;
	if(r2 < 1e-8) { r2 = 1e-8; }
	double coeff = (gm1 * M) / (r2 * sqrt(r2));
	tx += dx * coeff;
	ty += dy * coeff;
	tz += dz * coeff;
}

__device__ inline void write_f(const double* __restrict__ y, const double* __restrict__ mass, double* __restrict__ f,
                               size_t n, int b, double tx, double ty, double tz)
{
	double m1 = mass[b];
	f[b] = y[3 * n + b];
	f[n + b] = y[4 * n + b];
	f[2 * n + b] = y[5 * n + b];
	f[3 * n + b] = tx / m1;
	f[4 * n + b] = ty / m1;
	f[5 * n + b] = tz / m1;
}

// nbody_space_heap_stackless::traverse (nbody_space_heap_stackless.cpp:3-26)
__global__ void __launch_bounds__(256)
k_bh_stackless64(const double4* __restrict__ xyzr, const double* __restrict__ nmass,
                 const int32_t* __restrict__ body_n, const double* __restrict__ y, const double* __restrict__ mass,
                 double* __restrict__ f, int n, unsigned long long* __restrict__ counters)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int tree_size = 2 * n;
	unsigned long long inter = 0, vis = 0;
	if(t < n)
	{
		const int tn = n + t;
		const double4 me = xyzr[tn];
		const double m1 = nmass[tn];
		const double gm1 = -1.0 * m1;
		double tx = 0.0, ty = 0.0, tz = 0.0;
		int curr = kRoot;
		do
		{
			const double4 nd = xyzr[curr];
			const double dx = me.x - nd.x;
			const double dy = me.y - nd.y;
			const double dz = me.z - nd.z;
			const double d2 = dx * dx + dy * dy + dz * dz;
			++vis;
			if(d2 > nd.w)
			{
				add_force64(me.x, me.y, me.z, gm1, nd, nmass[curr], tx, ty, tz);
				++inter;
				curr = next_down(curr);
			}
			else
			{
				curr = next_up(curr, tree_size);
			}
		}
		while(curr != kRoot);
		write_f(y, mass, f, n, body_n[tn], tx, ty, tz);
	}
	count_add(counters, inter, vis);
}

// nbody_space_heap::traverse (nbody_space_heap.cpp:64-97): explicit stack, the
// right child pushed first so the left subtree is walked first.
__global__ void __launch_bounds__(256)
k_bh_stack64(const double4* __restrict__ xyzr, const double* __restrict__ nmass,
             const int32_t* __restrict__ body_n, const double* __restrict__ y, const double* __restrict__ mass,
             double* __restrict__ f, int n, unsigned long long* __restrict__ counters)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int tree_size = 2 * n;
	unsigned long long inter = 0, vis = 0;
	if(t < n)
	{
		const int tn = n + t;
		const double4 me = xyzr[tn];
		const double m1 = nmass[tn];
		const double gm1 = -1.0 * m1;
		double tx = 0.0, ty = 0.0, tz = 0.0;
		int stack[kMaxStack];
		int sp = 0;
		stack[sp++] = kRoot;
		while(sp != 0)
		{
			const int curr = stack[--sp];
			const double4 nd = xyzr[curr];
			const double dx = me.x - nd.x;
			const double dy = me.y - nd.y;
			const double dz = me.z - nd.z;
			const double d2 = dx * dx + dy * dy + dz * dz;
			++vis;
			if(d2 > nd.w)
			{
				add_force64(me.x, me.y, me.z, gm1, nd, nmass[curr], tx, ty, tz);
				++inter;
			}
			else
			{
				const int left = curr << 1;
				const int rght = left + 1;
				if(rght < tree_size && sp < kMaxStack) { stack[sp++] = rght; }
				if(left < tree_size && sp < kMaxStack) { stack[sp++] = left; }
			}
		}
		write_f(y, mass, f, n, body_n[tn], tx, ty, tz);
	}
	count_add(counters, inter, vis);
}

// FP32 interaction math: FP64 offsets to the node centre are rounded to FP32,
// the opening test and the monopole term run in FP32 with v_rsq_f32; sums in
// FP32 with a Kahan carry.
__global__ void __launch_bounds__(256)
k_bh_stackless32(const double4* __restrict__ xyzr, const double* __restrict__ nmass,
                 const int32_t* __restrict__ body_n, const double* __restrict__ y, const double* __restrict__ mass,
                 double* __restrict__ f, int n, unsigned long long* __restrict__ counters)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int tree_size = 2 * n;
	unsigned long long inter = 0, vis = 0;
	if(t < n)
	{
		const int tn = n + t;
		const double4 me = xyzr[tn];
		float ax = 0.0f, ay = 0.0f, az = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
		int curr = kRoot;
		do
		{
			const double4 nd = xyzr[curr];
			const float dx = static_cast<float>(nd.x - me.x);
			const float dy = static_cast<float>(nd.y - me.y);
			const float dz = static_cast<float>(nd.z - me.z);
			const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
			++vis;
			if(d2 > static_cast<float>(nd.w))
			{
				const float r2 = fmaxf(d2, 1e-8f);
				const float ri = __builtin_amdgcn_rsqf(r2);
				const float w = (static_cast<float>(nmass[curr]) * ri) * (ri * ri);
				float v = dx * w - cx; float s = ax + v; cx = (s - ax) - v; ax = s;
				v = dy * w - cy; s = ay + v; cy = (s - ay) - v; ay = s;
				v = dz * w - cz; s = az + v; cz = (s - az) - v; az = s;
				++inter;
				curr = next_down(curr);
			}
			else
			{
				curr = next_up(curr, tree_size);
			}
		}
		while(curr != kRoot);
		const int b = body_n[tn];
		f[b] = y[3 * n + b];
		f[n + b] = y[4 * n + b];
		f[2 * n + b] = y[5 * n + b];
		f[3 * n + b] = static_cast<double>(ax) - static_cast<double>(cx);
		f[4 * n + b] = static_cast<double>(ay) - static_cast<double>(cy);
		f[5 * n + b] = static_cast<double>(az) - static_cast<double>(cz);
	}
	count_add(counters, inter, vis);
}

// ---------------------------------------------------------------------------
// group walk: one wave per subtree of 64 leaves
// ---------------------------------------------------------------------------
// The 64 targets of a wave are the leaves of one heap node g, whose tight box
// bounds them all.  The wave walks the tree once for the whole group, 64 nodes
// per step from an LDS stack: a leaf is always taken as a body; an internal node
// is taken as a monopole when the distance from the group box to its mass
// centre exceeds its opening radius (then every target of the group accepts it
// too, so the cell set refines each target's own walk); otherwise its children
// are pushed.  Accepted entries are buffered in LDS as FP32 offsets from the
// box centre and swept by all 64 lanes, the same pair arithmetic as the direct
// FP32 kernel.  No node is dropped: a child push that would overflow the stack
// takes the node as a monopole instead (counted in counters[2]).
constexpr int kGroup = 64;
constexpr int kWalkStack = 1024;   // entries of sibling pairs; (levels + 2) * 32 suffice
constexpr int kWalkBuf = 128;

// Group handled by this work-group.  Work-groups are dispatched round-robin
// over the 8 XCDs, each with its own L2; consecutive groups are spatial
// neighbours that read mostly the same tree nodes, so the ids are permuted to
// give each XCD one contiguous run of groups (when their count divides by 8).
constexpr int kXcds = 8;

__device__ inline int xcd_group(int block, int nblocks)
{
	if(nblocks % kXcds != 0) { return block; }
	return (block % kXcds) * (nblocks / kXcds) + block / kXcds;
}


// Half the diagonal of a box.
__device__ inline double half_diagonal(double x0, double y0, double z0, double x1, double y1, double z1)
{
	const double ex = x1 - x0, ey = y1 - y0, ez = z1 - z0;
	return 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
}

// The walk stack holds sibling pairs: an even entry e stands for the nodes e and
// e + 1, the odd entry kRoot for the root alone.  A step pops up to 32 entries,
// one node per lane (lane 2i takes the left node of entry i, lane 2i + 1 the
// right one); every node the step opens pushes one entry, its two children.
// Popped in this order, the stack never holds more than (levels + 2) * 32
// entries.  In the Morton tree (kRadix) an entry encodes the children of the
// opened node (radix_kids), so that popping reads no memory, and 0 stands for
// the root alone.
template <bool kRadix = false>
__device__ inline int walk_pop(const int* stack, int& sp, int lane, bool& valid, int n = 0)
{
	const int cnt = sp < kGroup / 2 ? sp : kGroup / 2;
	const int i = lane >> 1;
	int node = kRoot;
	valid = false;
	if(i < cnt)
	{
		const int e = stack[sp - cnt + i];
		if constexpr(kRadix)
		{
			if(e == 0)
			{
				valid = (lane & 1) == 0;
			}
			else
			{
				// g = split + 1; bit 30: the left child is leaf n + g - 1 (else
				// internal id g), bit 31: the right child is leaf n + g (else g + 1)
				const unsigned u = static_cast<unsigned>(e);
				const int g = static_cast<int>(u & 0x3fffffffu);
				node = (lane & 1) ? ((u >> 31) ? n + g : g + 1) : (((u >> 30) & 1u) ? n + g - 1 : g);
				valid = true;
			}
		}
		else
		{
			node = e == kRoot ? kRoot : e + (lane & 1);
			valid = e != kRoot || (lane & 1) == 0;
		}
	}
	sp -= cnt;
	return node;
}

// Stack entry of an opened node (in the Morton tree: from its children kids,
// which are leaf n + s / internal id s + 1 and leaf n + s + 1 / internal id s + 2
// for the split s), and the entry the walk starts from.
template <bool kRadix>
__device__ inline int walk_entry(int node, int2 kids = make_int2(0, 0), int n = 0)
{
	if constexpr(kRadix)
	{
		const unsigned g = static_cast<unsigned>(kids.x < n ? kids.x : kids.x - n + 1);
		return static_cast<int>(g | (kids.x >= n ? 1u << 30 : 0u) | (kids.y >= n ? 1u << 31 : 0u));
	}
	return 2 * node;
}
template <bool kRadix>
__device__ inline int walk_start() { return kRadix ? 0 : kRoot; }

// The group walks sweep in scaled units: offsets x' = 1e4 x, masses m' = 1e8 m
// and second moments Q' = 1e16 Q (mass x length^2), each formed in FP64 and
// rounded once.  r'^2 >= 1 is then the reference's clamp r^2 >= 1e-8
// (nbody/nbody_data.cpp:35-44), which the [0, 1] output clamp of v_rsq_f32
// applies at no cost: q = min(1/r', 1) = 1e-4 / sqrt(max(r^2, 1e-8)), and the
// monopole term m' x' q^3 and the quadrupole term below come out in unscaled
// acceleration units.
constexpr double kSweepLen = 1e4;
constexpr double kSweepMass = 1e8;
constexpr double kSweepQuad = 1e16;

__device__ inline float rsq_clamped(float r2)
{
	return __builtin_amdgcn_fmed3f(__builtin_amdgcn_rsqf(r2), 0.0f, 1.0f);   // folds into v_rsq_f32 clamp
}

// Sweep entry (scaled offset from the group centre, scaled mass) of a node.
__device__ inline float4 sweep_entry(const double4& rec, double m, double ccx, double ccy, double ccz)
{
	return make_float4(static_cast<float>((rec.x - ccx) * kSweepLen), static_cast<float>((rec.y - ccy) * kSweepLen),
	                   static_cast<float>((rec.z - ccz) * kSweepLen), static_cast<float>(m * kSweepMass));
}

// Pop the next chunk of the walk stack (walk_pop) and request its node records;
// returns the number of stack entries taken.
// In the Morton tree the children of internal nodes are requested with the
// records (kids), for the stack entries of the nodes the step opens; kQ: the
// second moments too (q), from the same 64-byte record.
template <bool kRadix = false, bool kQ = false>
__device__ inline int walk_fetch(const int* stack, int& sp, int lane, const nbh_walk_rec* __restrict__ wr, int& node,
                                 bool& valid, double4& rec, double& m, int2& kids, const int2* __restrict__ child = nullptr,
                                 int n = 0, float* q = nullptr)
{
	const int cnt = sp < kGroup / 2 ? sp : kGroup / 2;
	node = walk_pop<kRadix>(stack, sp, lane, valid, n);
	rec = make_double4(0.0, 0.0, 0.0, 0.0);
	m = 0.0;
	kids = make_int2(0, 0);
	if constexpr(kQ)
	{
		for(int k = 0; k < 6; ++k) { q[k] = 0.0f; }
	}
	if(valid)
	{
		rec = wr[node].c;
		m = wr[node].m;
		if constexpr(kQ)
		{
			for(int k = 0; k < 6; ++k) { q[k] = wr[node].q[k]; }
		}
		if(kRadix && node < n) { kids = child[node]; }
	}
	return cnt;
}

__device__ inline void sweep32(const float4* __restrict__ buf, int count, float tx, float ty, float tz,
                               float& ax, float& ay, float& az, float& cx, float& cy, float& cz)
{
	float px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll 4
	for(int j = 0; j < count; ++j)
	{
		const float4 s = buf[j];
		const float dx = s.x - tx;
		const float dy = s.y - ty;
		const float dz = s.z - tz;
		const float ri = rsq_clamped(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
		const float w = (s.w * ri) * (ri * ri);
		px = fmaf(dx, w, px);
		py = fmaf(dy, w, py);
		pz = fmaf(dz, w, pz);
	}
	float v = px - cx; float t = ax + v; cx = (t - ax) - v; ax = t;
	v = py - cy; t = ay + v; cy = (t - ay) - v; ay = t;
	v = pz - cz; t = az + v; cz = (t - az) - v; az = t;
}

// sweep32 for kT bodies of one lane at once: each buffer entry is read once
// for all of them, and their independent chains interleave.  Every body's sums
// take the same operations in the same order as sweep32's.
template <int kT>
__device__ inline void sweep32_n(const float4* __restrict__ buf, int count, const float* tx, const float* ty,
                                 const float* tz, float* ax, float* ay, float* az, float* cx, float* cy, float* cz)
{
	float px[kT], py[kT], pz[kT];
#pragma unroll
	for(int t = 0; t < kT; ++t) { px[t] = py[t] = pz[t] = 0.0f; }
#pragma unroll 2
	for(int j = 0; j < count; ++j)
	{
		const float4 s = buf[j];
#pragma unroll
		for(int t = 0; t < kT; ++t)
		{
			const float dx = s.x - tx[t];
			const float dy = s.y - ty[t];
			const float dz = s.z - tz[t];
			const float ri = rsq_clamped(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
			const float w = (s.w * ri) * (ri * ri);
			px[t] = fmaf(dx, w, px[t]);
			py[t] = fmaf(dy, w, py[t]);
			pz[t] = fmaf(dz, w, pz[t]);
		}
	}
#pragma unroll
	for(int t = 0; t < kT; ++t)
	{
		float v = px[t] - cx[t]; float u = ax[t] + v; cx[t] = (u - ax[t]) - v; ax[t] = u;
		v = py[t] - cy[t]; u = ay[t] + v; cy[t] = (u - ay[t]) - v; ay[t] = u;
		v = pz[t] - cz[t]; u = az[t] + v; cz[t] = (u - az[t]) - v; az[t] = u;
	}
}

// Cells waiting for the monopole + quadrupole sweep, 40 bytes each in LDS:
// e0 = (mass-centre offset from the group centre, M), e1 = 2.5 x (Q'xx Q'yy
// Q'zz Q'xy), e2 = 2.5 x (Q'xz Q'yz), with Q' = 3 Q - tr(Q) I the traceless
// second moment about the mass centre (all scaled as the sweeps are).
template <int kN>
struct quad_lds {
	float4 e0[kN];
	float4 e1[kN];
	float2 e2[kN];
};

// Monopole + quadrupole field of the cells in qb at the lane's body; with
// r = cm - x, d = |r|, u = 1/d:
//   a = M r / d^3 + (2.5 (r.Q'.r) / d^2 r - Q' r) / d^5
//     = (M + s u^4) u^3 r - 0.4 u^5 (2.5 Q' r),   s = r.(2.5 Q').r,
// all in the scaled units of sweep32 (Q' times kSweepQuad).  Kahan per sweep.
template <int kN>
__device__ inline void sweep_mq(const quad_lds<kN>& qb, int count, float tx, float ty, float tz, float& ax, float& ay,
                                float& az, float& cx, float& cy, float& cz)
{
	float px = 0.0f, py = 0.0f, pz = 0.0f;
	for(int j = 0; j < count; ++j)
	{
		const float4 e0 = qb.e0[j], e1 = qb.e1[j];
		const float2 e2 = qb.e2[j];
		const float rx = e0.x - tx, ry = e0.y - ty, rz = e0.z - tz;
		const float u = rsq_clamped(fmaf(rz, rz, fmaf(ry, ry, rx * rx)));
		const float u2 = u * u;
		const float u3 = u2 * u;
		const float u4 = u2 * u2;
		const float Qx = fmaf(e1.x, rx, fmaf(e1.w, ry, e2.x * rz));
		const float Qy = fmaf(e1.w, rx, fmaf(e1.y, ry, e2.y * rz));
		const float Qz = fmaf(e2.x, rx, fmaf(e2.y, ry, e1.z * rz));
		const float s = fmaf(rx, Qx, fmaf(ry, Qy, rz * Qz));
		const float k = fmaf(s, u4, e0.w) * u3;
		const float g = -0.4f * (u4 * u);
		px = fmaf(rx, k, fmaf(Qx, g, px));
		py = fmaf(ry, k, fmaf(Qy, g, py));
		pz = fmaf(rz, k, fmaf(Qz, g, pz));
	}
	float v = px - cx; float t = ax + v; cx = (t - ax) - v; ax = t;
	v = py - cy; t = ay + v; cy = (t - ay) - v; ay = t;
	v = pz - cz; t = az + v; cz = (t - az) - v; az = t;
}

// sweep_mq for kT bodies of one lane at once (as sweep32_n: one read of each
// cell, the same operations per body as sweep_mq).
template <int kT, int kN>
__device__ inline void sweep_mq_n(const quad_lds<kN>& qb, int count, const float* tx, const float* ty, const float* tz,
                                  float* ax, float* ay, float* az, float* cx, float* cy, float* cz)
{
	float px[kT], py[kT], pz[kT];
#pragma unroll
	for(int t = 0; t < kT; ++t) { px[t] = py[t] = pz[t] = 0.0f; }
	for(int j = 0; j < count; ++j)
	{
		const float4 e0 = qb.e0[j], e1 = qb.e1[j];
		const float2 e2 = qb.e2[j];
#pragma unroll
		for(int t = 0; t < kT; ++t)
		{
			const float rx = e0.x - tx[t], ry = e0.y - ty[t], rz = e0.z - tz[t];
			const float u = rsq_clamped(fmaf(rz, rz, fmaf(ry, ry, rx * rx)));
			const float u2 = u * u;
			const float u3 = u2 * u;
			const float u4 = u2 * u2;
			const float Qx = fmaf(e1.x, rx, fmaf(e1.w, ry, e2.x * rz));
			const float Qy = fmaf(e1.w, rx, fmaf(e1.y, ry, e2.y * rz));
			const float Qz = fmaf(e2.x, rx, fmaf(e2.y, ry, e1.z * rz));
			const float s = fmaf(rx, Qx, fmaf(ry, Qy, rz * Qz));
			const float k = fmaf(s, u4, e0.w) * u3;
			const float g = -0.4f * (u4 * u);
			px[t] = fmaf(rx, k, fmaf(Qx, g, px[t]));
			py[t] = fmaf(ry, k, fmaf(Qy, g, py[t]));
			pz[t] = fmaf(rz, k, fmaf(Qz, g, pz[t]));
		}
	}
#pragma unroll
	for(int t = 0; t < kT; ++t)
	{
		float v = px[t] - cx[t]; float w = ax[t] + v; cx[t] = (w - ax[t]) - v; ax[t] = w;
		v = py[t] - cy[t]; w = ay[t] + v; cy[t] = (w - ay[t]) - v; ay[t] = w;
		v = pz[t] - cz[t]; w = az[t] + v; cz[t] = (w - az[t]) - v; az[t] = w;
	}
}

// Cells waiting for the monopole + quadrupole sweep; a sweep runs before a
// step whenever more than kQuadBuf - 64 wait.
constexpr int kQuadBuf = 88;

// Entry of a node (scaled mass-centre offset and mass in ent, second moments Q)
// at slot `slot` of qb, with the 2.5 x Q' that sweep_mq expects.
template <int kN>
__device__ inline void put_quad(quad_lds<kN>& qb, int slot, const float4& ent, const float* __restrict__ Q)
{
	const float q0 = Q[0], q1 = Q[1], q2 = Q[2], q3 = Q[3], q4 = Q[4], q5 = Q[5];
	const float tr = q0 + q1 + q2;
	const float k = static_cast<float>(2.5 * kSweepQuad);
	qb.e0[slot] = ent;
	qb.e1[slot] = make_float4((3.0f * q0 - tr) * k, (3.0f * q1 - tr) * k, (3.0f * q2 - tr) * k, 3.0f * q3 * k);
	qb.e2[slot] = make_float2(3.0f * q4 * k, 3.0f * q5 * k);
}

// Quadrupole field at the group centre of a cell whose mass centre sits at r
// from it, and its Jacobian, added to acc[9] (field, then J as xx yy zz xy xz
// yz).  With Q' = 3 Q - tr(Q) I, s = r.Q'.r and d = |r|:
//   g = (2.5 s / d^2 r - Q' r) / d^5,
//   G = dg/dr = (5 (r (Q'r)' + (Q'r) r') + 2.5 s I) / d^7 - 17.5 s r r' / d^9 - Q' / d^5,
// and since r = cm - x, the field at centre + delta is g - G delta: J = -G.
__device__ inline void quad_expansion(float rx, float ry, float rz, float qxx, float qyy, float qzz, float qxy,
                                      float qxz, float qyz, float* acc)
{
	const float Qx = fmaf(qxx, rx, fmaf(qxy, ry, qxz * rz));
	const float Qy = fmaf(qxy, rx, fmaf(qyy, ry, qyz * rz));
	const float Qz = fmaf(qxz, rx, fmaf(qyz, ry, qzz * rz));
	const float sq = fmaf(rx, Qx, fmaf(ry, Qy, rz * Qz));
	const float ri = __builtin_amdgcn_rsqf(fmaf(rz, rz, fmaf(ry, ry, rx * rx)));
	const float ri2 = ri * ri;
	const float ri5 = (ri2 * ri2) * ri;
	const float ri7 = ri5 * ri2;
	const float c = 2.5f * sq * ri7;
	acc[0] += fmaf(c, rx, -Qx * ri5);
	acc[1] += fmaf(c, ry, -Qy * ri5);
	acc[2] += fmaf(c, rz, -Qz * ri5);
	const float a = 5.0f * ri7;
	const float e = 17.5f * sq * (ri7 * ri2);
	acc[3] -= a * (2.0f * rx * Qx) + c - e * rx * rx - qxx * ri5;
	acc[4] -= a * (2.0f * ry * Qy) + c - e * ry * ry - qyy * ri5;
	acc[5] -= a * (2.0f * rz * Qz) + c - e * rz * rz - qzz * ri5;
	acc[6] -= a * (rx * Qy + ry * Qx) - e * rx * ry - qxy * ri5;
	acc[7] -= a * (rx * Qz + rz * Qx) - e * rx * rz - qxz * ri5;
	acc[8] -= a * (ry * Qz + rz * Qy) - e * ry * rz - qyz * ri5;
}

// Monopole field of a cell of mass M whose mass centre sits at r from the group
// centre, expanded about the centre to third order, added to acc[34]: field
// acc[0..2], Jacobian acc[3..8] (xx yy zz xy xz yz, as quad_expansion), the sum
// T = M r r r / d^7 in acc[9..18] (xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz) and
// W = M r r r r / d^9 in acc[19..33] (xxxx xxxy xxxz xxyy xxyz xxzz xyyy xyyz
// xyzz xzzz yyyy yyyz yyzz yzzz zzzz).  With f(r) = M r / d^3, the field at
// centre + delta is f(r - delta):
//   a(delta) = f - D1 delta + 1/2 D2[delta^2] - 1/6 D3[delta^3] + O(|delta|^4),
//   D1 = M (I / d^3 - 3 r r' / d^5)            (J = -D1),
//   1/2 D2[delta^2] = -3 (V.delta) delta - 1.5 |delta|^2 V + 7.5 T[delta^2],
//   -1/6 D3[delta^3] = 1.5 s |delta|^2 delta - 7.5 (delta.A.delta) delta
//                      - 7.5 |delta|^2 A delta + 17.5 W[delta^3],
// where V = sum M r / d^5 = T_ikk, A = sum M r r' / d^7 = W_ijkk and
// s = sum M / d^5 = A_kk, so only T and W are accumulated (local_terms
// evaluates the last two lines per body).
__device__ inline void local_expansion(float rx, float ry, float rz, float M, float* acc)
{
	const float u = __builtin_amdgcn_rsqf(fmaf(rz, rz, fmaf(ry, ry, rx * rx)));
	const float u2 = u * u;
	const float w3 = (M * u) * u2;
	const float w5 = w3 * u2;
	const float w7 = w5 * u2;
	acc[0] = fmaf(w3, rx, acc[0]);
	acc[1] = fmaf(w3, ry, acc[1]);
	acc[2] = fmaf(w3, rz, acc[2]);
	const float t5 = 3.0f * w5;
	const float tx = t5 * rx, ty = t5 * ry;
	acc[3] += fmaf(tx, rx, -w3);
	acc[4] += fmaf(ty, ry, -w3);
	acc[5] += fmaf(t5 * rz, rz, -w3);
	acc[6] = fmaf(tx, ry, acc[6]);
	acc[7] = fmaf(tx, rz, acc[7]);
	acc[8] = fmaf(ty, rz, acc[8]);
	const float px = w7 * rx, py = w7 * ry, pz = w7 * rz;
	const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
	acc[9] = fmaf(px, xx, acc[9]);
	acc[10] = fmaf(px * rx, ry, acc[10]);
	acc[11] = fmaf(px * rx, rz, acc[11]);
	acc[12] = fmaf(px, yy, acc[12]);
	acc[13] = fmaf(px * ry, rz, acc[13]);
	acc[14] = fmaf(px, zz, acc[14]);
	acc[15] = fmaf(py, yy, acc[15]);
	acc[16] = fmaf(py * ry, rz, acc[16]);
	acc[17] = fmaf(py, zz, acc[17]);
	acc[18] = fmaf(pz, zz, acc[18]);
	const float w9 = w7 * u2;
	const float a = w9 * xx, b = w9 * yy, c = w9 * zz;
	const float xy = rx * ry, xz = rx * rz, yz = ry * rz;
	acc[19] = fmaf(a, xx, acc[19]);
	acc[20] = fmaf(a, xy, acc[20]);
	acc[21] = fmaf(a, xz, acc[21]);
	acc[22] = fmaf(a, yy, acc[22]);
	acc[23] = fmaf(a, yz, acc[23]);
	acc[24] = fmaf(a, zz, acc[24]);
	acc[25] = fmaf(b, xy, acc[25]);
	acc[26] = fmaf(b, xz, acc[26]);
	acc[27] = fmaf(c, xy, acc[27]);
	acc[28] = fmaf(c, xz, acc[28]);
	acc[29] = fmaf(b, yy, acc[29]);
	acc[30] = fmaf(b, yz, acc[30]);
	acc[31] = fmaf(b, zz, acc[31]);
	acc[32] = fmaf(c, yz, acc[32]);
	acc[33] = fmaf(c, zz, acc[33]);
}

// Second- and third-order terms of the local expansion (T = acc[9..18] and
// W = acc[19..33] of local_expansion, summed over the cells) at offset
// (dx, dy, dz) from the group centre.
__device__ inline void local_terms(const float* T, float dx, float dy, float dz, float& ox, float& oy, float& oz)
{
	const float vx = T[0] + T[3] + T[5], vy = T[1] + T[6] + T[8], vz = T[2] + T[7] + T[9];
	const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
	const float xy2 = 2.0f * dx * dy, xz2 = 2.0f * dx * dz, yz2 = 2.0f * dy * dz;
	const float tdx = T[0] * xx + T[3] * yy + T[5] * zz + T[1] * xy2 + T[2] * xz2 + T[4] * yz2;
	const float tdy = T[1] * xx + T[6] * yy + T[8] * zz + T[3] * xy2 + T[4] * xz2 + T[7] * yz2;
	const float tdz = T[2] * xx + T[7] * yy + T[9] * zz + T[4] * xy2 + T[5] * xz2 + T[8] * yz2;
	const float vd = vx * dx + vy * dy + vz * dz;
	const float dd = xx + yy + zz;
	ox = -3.0f * vd * dx - 1.5f * dd * vx + 7.5f * tdx;
	oy = -3.0f * vd * dy - 1.5f * dd * vy + 7.5f * tdy;
	oz = -3.0f * vd * dz - 1.5f * dd * vz + 7.5f * tdz;
	const float* W = T + 10;
	// monomials of delta^3 times their multiplicities
	const float m0 = xx * dx, m1 = 3.0f * xx * dy, m2 = 3.0f * xx * dz, m3 = 3.0f * dx * yy, m4 = 3.0f * xy2 * dz;
	const float m5 = 3.0f * dx * zz, m6 = yy * dy, m7 = 3.0f * yy * dz, m8 = 3.0f * dy * zz, m9 = zz * dz;
	const float wx = W[0] * m0 + W[1] * m1 + W[2] * m2 + W[3] * m3 + W[4] * m4 + W[5] * m5 + W[6] * m6 + W[7] * m7 +
	                 W[8] * m8 + W[9] * m9;
	const float wy = W[1] * m0 + W[3] * m1 + W[4] * m2 + W[6] * m3 + W[7] * m4 + W[8] * m5 + W[10] * m6 +
	                 W[11] * m7 + W[12] * m8 + W[13] * m9;
	const float wz = W[2] * m0 + W[4] * m1 + W[5] * m2 + W[7] * m3 + W[8] * m4 + W[9] * m5 + W[11] * m6 +
	                 W[12] * m7 + W[13] * m8 + W[14] * m9;
	const float axx = W[0] + W[3] + W[5], ayy = W[3] + W[10] + W[12], azz = W[5] + W[12] + W[14];
	const float axy = W[1] + W[6] + W[8], axz = W[2] + W[7] + W[9], ayz = W[4] + W[11] + W[13];
	const float adx = axx * dx + axy * dy + axz * dz;
	const float ady = axy * dx + ayy * dy + ayz * dz;
	const float adz = axz * dx + ayz * dy + azz * dz;
	const float dad = dx * adx + dy * ady + dz * adz;
	const float s5 = 1.5f * (axx + ayy + azz) * dd - 7.5f * dad;
	ox += s5 * dx - 7.5f * dd * adx + 17.5f * wx;
	oy += s5 * dy - 7.5f * dd * ady + 17.5f * wy;
	oz += s5 * dz - 7.5f * dd * adz + 17.5f * wz;
}

// Far cells wait in an LDS ring of kFarRing node numbers and are expanded 64 at
// a time, one per lane, so the expansion runs on full waves; the expansion
// reads the node's mass centre, mass and second moments from its record again
// (one cache line the walk step has just read).  A walk step adds at most 64
// entries and a batch is taken as soon as 64 wait, so at most 127 are ever
// pending.
constexpr int kFarRing = 128;

// Expand `count` (<= 64) ring entries from `head`, one per lane, about the
// centre (ccx, ccy, ccz): offsets of the mass centres from it, unscaled.
template <bool kLocal>
__device__ inline void far_batch(const int* fn, const nbh_walk_rec* __restrict__ wr, int head, int count, int lane,
                                 double ccx, double ccy, double ccz, float* acc)
{
	if(lane < count)
	{
		const int s = (head + lane) & (kFarRing - 1);
		const nbh_walk_rec* r = wr + fn[s];
		const double4 c = r->c;
		const float4 e = make_float4(static_cast<float>(c.x - ccx), static_cast<float>(c.y - ccy),
		                             static_cast<float>(c.z - ccz), static_cast<float>(r->m));
		const float* Q = r->q;
		const float q0 = Q[0], q1 = Q[1], q2 = Q[2], q3 = Q[3], q4 = Q[4], q5 = Q[5];
		const float tr = q0 + q1 + q2;
		quad_expansion(e.x, e.y, e.z, 3.0f * q0 - tr, 3.0f * q1 - tr, 3.0f * q2 - tr, 3.0f * q3, 3.0f * q4, 3.0f * q5,
		               acc);
		if constexpr(kLocal) { local_expansion(e.x, e.y, e.z, e.w, acc); }
	}
}

// Two-level local walk.  A super-group is the heap node of 2^kSuperShift
// consecutive groups (1024 bodies).  One wave walks the tree once for the
// super-group's box: a node it accepts is accepted by every group inside that
// box too (the distance from a smaller box inside is no shorter), so accepted
// cells at least near_k half-diagonals of the super box from its centre go into
// one local expansion about that centre (far_batch, third order).  Every other
// accepted node, and every rejected node whose bodies all lie within that
// distance, goes into a list instead of being opened; other rejected nodes are
// opened.  Each group then walks from the listed nodes instead of the
// root, with its own acceptance test (so near its box the cells are those of
// its own walk), sweeping and expanding about its own centre as the one-level
// walk does; each body adds both expansions.  A list longer than `cap`, or a
// stack spill in the super walk, makes the groups of that super-group walk from
// the root instead.  The super-group's near distance is max(near_k,
// kSuperNearMin) half-diagonals of its box: rejected nodes are then listed
// higher up the tree, for each group to open with its own test.  Measured at
// 2^22 Plummer bodies, theta 0.65, near_k 4: walk 52.7 -> 51.8 ms, RMS 1.0e-3 ->
// 9.9e-4; at theta 0.6, near_k 3 no list overflows where near_k alone had some.
constexpr int kSuperShift = 4;
constexpr float kSuperNearMin = 5.0f;

// The tree a group walk runs on.  Heap (child == nullptr): children of node k
// are 2k and 2k + 1, and the walk groups and super-groups are the heap nodes
// of 64 and 1024 bodies.  Morton tree: the children of internal id k are
// child[k]; group g is grp[g] = (node id, first sorted leaf, leaves (<= 64),
// super-group or -1) and super-group s is sup[s] = (node id, first leaf).  In
// both, a group's or super-group's box is its node's box.
struct WalkTree {
	const int2* child = nullptr;
	const int4* grp = nullptr;
	const int2* sup = nullptr;
	int         ngroups = 0;
	int         nsupers = 0;
};

struct TwoLevel {
	int*   lists = nullptr;   // [super][cap] listed nodes
	int*   count = nullptr;   // [super] list length (> cap: overflowed)
	float* coef = nullptr;    // [super][34] expansion about the super box centre (local_expansion layout)
	int    cap = 0;
	// optional: the super-group's far cells, for the learned correction
	int*   far_list = nullptr;   // [super][far_cap] far nodes, in walk order
	int*   far_count = nullptr;  // [super] far cells found (> far_cap: the rest are not listed)
	int    far_cap = 0;
};

// Next 64 listed nodes from position pos and their records; `entry` holds this
// lane's list entry, read one list step ahead, and takes that of the next step.
template <bool kRadix = false, bool kQ = false>
__device__ inline void list_fetch(const int* __restrict__ list, int pos, int count, int lane,
                                  const nbh_walk_rec* __restrict__ wr, int& entry, int& node, bool& valid,
                                  double4& rec, double& m, int2& kids, const int2* __restrict__ child = nullptr,
                                  int n = 0, float* q = nullptr)
{
	const int i = pos + lane;
	valid = i < count;
	node = valid ? entry : kRoot;
	rec = make_double4(0.0, 0.0, 0.0, 0.0);
	m = 0.0;
	kids = make_int2(0, 0);
	if constexpr(kQ)
	{
		for(int k = 0; k < 6; ++k) { q[k] = 0.0f; }
	}
	if(valid)
	{
		rec = wr[node].c;
		m = wr[node].m;
		if constexpr(kQ)
		{
			for(int k = 0; k < 6; ++k) { q[k] = wr[node].q[k]; }
		}
		if(kRadix && node < n) { kids = child[node]; }
	}
	entry = i + kGroup < count ? list[i + kGroup] : kRoot;
}

// kQuad: internal cells taken as monopoles also contribute their quadrupole
// term (second central moments `quad`).  A cell whose mass centre lies within
// `near` of the group centre gets it per body (fused monopole + quadrupole
// sweep); a farther one gets the monopole per body and its quadrupole field
// expanded to first order about the group centre, once per group.
// kLocal (with kQuad): a far cell's monopole field is expanded too, to third
// order (local_expansion), instead of being swept per body.  kTwo (with kLocal):
// the group takes its super-group's list instead of walking (TwoLevel).
template <bool kQuad, bool kLocal, bool kTwo, bool kRadix>
__device__ inline void group_walk_one(int gi, int* __restrict__ stack, float4* __restrict__ buf,
                                      quad_lds<kQuad ? kQuadBuf : 1>& qbuf, int* __restrict__ farn,
                                      const nbh_walk_rec* __restrict__ wr, const double* __restrict__ bmin,
                                      const double* __restrict__ bmax, const int32_t* __restrict__ body_n,
                                      const double* __restrict__ y, double* __restrict__ f, int n, float near_k,
                                      unsigned long long* __restrict__ counters, const TwoLevel& two,
                                      const WalkTree& tree)
{
	const int lane = threadIdx.x;
	// the group's node, first leaf, size and super-group
	int gnode = n / kGroup + gi, gfirst = gi * kGroup, gcount = kGroup, gsuper = gi >> kSuperShift;
	if constexpr(kRadix)
	{
		const int4 gr = tree.grp[gi];
		gnode = gr.x;
		gfirst = gr.y;
		gcount = gr.z;
		gsuper = gr.w;
	}
	const double* glo = bmin + 3 * static_cast<size_t>(gnode);
	const double* ghi = bmax + 3 * static_cast<size_t>(gnode);
	const double gx0 = glo[0], gy0 = glo[1], gz0 = glo[2];
	const double gx1 = ghi[0], gy1 = ghi[1], gz1 = ghi[2];
	const double ccx = 0.5 * (gx0 + gx1), ccy = 0.5 * (gy0 + gy1), ccz = 0.5 * (gz0 + gz1);
	// this lane's bodies: leaf lane of the group, and leaf lane + 64 in the Morton
	// tree's walk groups of up to 128 bodies (kT = 2); lanes beyond a smaller
	// group's size repeat its first body and write nothing
	constexpr int kT = kRadix ? 2 : 1;
	bool body_lane[kT];
	int tn[kT];
	float tx[kT], ty[kT], tz[kT];                   // scaled offsets from the group centre, for the sweeps
	float ax[kT], ay[kT], az[kT], cx[kT], cy[kT], cz[kT];
#pragma unroll
	for(int t = 0; t < kT; ++t)
	{
		const int j = lane + kGroup * t;
		body_lane[t] = j < gcount;
		tn[t] = n + gfirst + (body_lane[t] ? j : 0);
		const double4 me = wr[tn[t]].c;
		tx[t] = static_cast<float>((me.x - ccx) * kSweepLen);
		ty[t] = static_cast<float>((me.y - ccy) * kSweepLen);
		tz[t] = static_cast<float>((me.z - ccz) * kSweepLen);
		ax[t] = ay[t] = az[t] = cx[t] = cy[t] = cz[t] = 0.0f;
	}
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	unsigned vis = 0, taken = 0, spill = 0, nfar = 0, nnear = 0;
	constexpr int kAcc = kLocal ? 34 : 9;
	float acc[kAcc];
	for(int k = 0; k < kAcc; ++k) { acc[k] = 0.0f; }
	const float near_d = static_cast<float>(static_cast<double>(near_k) * kSweepLen *
	                                        half_diagonal(gx0, gy0, gz0, gx1, gy1, gz1));   // scaled

	int si = 0, lcnt = 0, lpos = 0;
	bool use_list = false;
	if constexpr(kTwo)
	{
		si = gsuper;
		use_list = si >= 0 && two.count[si] <= two.cap;
		lcnt = use_list ? two.count[si] : 0;
	}
	const int* list = kTwo ? two.lists + static_cast<size_t>(si) * two.cap : nullptr;
	int lentry = kTwo && lane < lcnt ? list[lane] : kRoot;   // list entry of the next list step
	int sp = use_list ? 0 : 1;
	int nbuf = 0;
	int nq = 0;
	int fhead = 0, ftail = 0;     // far-cell ring: entries [fhead, ftail) pending
	if(lane == 0) { stack[0] = walk_start<kRadix>(); }
	__syncthreads();
	while(sp > 0 || lpos < lcnt)
	{
		bool valid;
		int node;
		double4 rec;
		double m;
		int2 kids;
		float nq6[kQuad ? 6 : 1];     // second moments, read with the record
		// issue this step's node loads, then sweep the buffers the previous steps
		// filled while they are in flight; listed nodes are taken when the stack
		// below them is done
		if(sp == 0)
		{
			list_fetch<kRadix, kQuad>(list, lpos, lcnt, lane, wr, lentry, node, valid, rec, m, kids, tree.child, n,
			                          nq6);
			lpos += kGroup;
		}
		else
		{
			walk_fetch<kRadix, kQuad>(stack, sp, lane, wr, node, valid, rec, m, kids, tree.child, n, nq6);
		}
		if(nbuf > kWalkBuf - kGroup)
		{
			sweep32_n<kT>(buf, nbuf, tx, ty, tz, ax, ay, az, cx, cy, cz);
			nbuf = 0;
		}
		if constexpr(kQuad)
		{
			if(nq > kQuadBuf - kGroup)
			{
				sweep_mq_n<kT>(qbuf, nq, tx, ty, tz, ax, ay, az, cx, cy, cz);
				nq = 0;
			}
			if(ftail - fhead >= kGroup)
			{
				far_batch<kLocal>(farn, wr, fhead, kGroup, lane, ccx, ccy, ccz, acc);
				fhead += kGroup;
			}
		}
		__syncthreads();
		bool accept = false, open = false, cell = false;
		float4 ent = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		if(valid)
		{
			ent = sweep_entry(rec, m, ccx, ccy, ccz);
			if(node >= n)
			{
				accept = true;
			}
			else
			{
				// distance from the group box to the node's mass centre
				const double ex = fmax(fmax(gx0 - rec.x, rec.x - gx1), 0.0);
				const double ey = fmax(fmax(gy0 - rec.y, rec.y - gy1), 0.0);
				const double ez = fmax(fmax(gz0 - rec.z, rec.z - gz1), 0.0);
				accept = ex * ex + ey * ey + ez * ez > rec.w;
				open = !accept;
				cell = kQuad && accept;     // taken internal cell (stack spills stay monopoles)
			}
		}
		const uint64_t bopen = __ballot(open);
		const int nopen = __popcll(bopen);
		if(sp + nopen > kWalkStack)
		{
			// no room for the children: take these nodes as monopoles
			spill += open ? 1 : 0;
			accept = accept || open;
			open = false;
		}
		const uint64_t bacc = __ballot(accept);
		const uint64_t bopen2 = __ballot(open);
		if constexpr(kQuad)
		{
			// near taken internal cells: monopole + quadrupole sweep; the rest:
			// monopole sweep, far cells adding their expansion
			const bool far = cell && !(ent.x * ent.x + ent.y * ent.y + ent.z * ent.z < near_d * near_d);
			const bool nearc = cell && !far;
			const uint64_t bnear = __ballot(nearc);
			const uint64_t bfar = __ballot(far);
			const uint64_t bmono = kLocal ? bacc & ~bnear & ~bfar : bacc & ~bnear;
			if(nearc) { put_quad(qbuf, nq + __popcll(bnear & lt), ent, nq6); }
			else if(accept && !(kLocal && far)) { buf[nbuf + __popcll(bmono & lt)] = ent; }
			if(far)
			{
				farn[(ftail + __popcll(bfar & lt)) & (kFarRing - 1)] = node;
			}
			ftail += __popcll(bfar);
			nq += __popcll(bnear);
			nbuf += __popcll(bmono);
			nfar += far ? 1 : 0;
			nnear += nearc ? 1 : 0;
		}
		else
		{
			if(accept) { buf[nbuf + __popcll(bacc & lt)] = ent; }
			nbuf += __popcll(bacc);
		}
		if(open) { stack[sp + __popcll(bopen2 & lt)] = walk_entry<kRadix>(node, kids, n); }
		sp += __popcll(bopen2);
		vis += valid ? 1 : 0;
		taken += accept ? 1 : 0;
		__syncthreads();
	}
	if(nbuf > 0)
	{
		sweep32_n<kT>(buf, nbuf, tx, ty, tz, ax, ay, az, cx, cy, cz);
	}
	float kx[kT], ky[kT], kz[kT];
	for(int t = 0; t < kT; ++t) { kx[t] = ky[t] = kz[t] = 0.0f; }
	if constexpr(kQuad)
	{
		if(nq > 0) { sweep_mq_n<kT>(qbuf, nq, tx, ty, tz, ax, ay, az, cx, cy, cz); }
		for(; fhead < ftail; fhead += kGroup)
		{
			far_batch<kLocal>(farn, wr, fhead, ftail - fhead < kGroup ? ftail - fhead : kGroup, lane, ccx, ccy, ccz,
			                  acc);
		}
		for(int sh = 1; sh < 64; sh <<= 1)
		{
			for(int k = 0; k < kAcc; ++k) { acc[k] += __shfl_xor(acc[k], sh, 64); }
		}
#pragma unroll
		for(int t = 0; t < kT; ++t)
		{
			const double4 me = wr[tn[t]].c;
			const float ux = static_cast<float>(me.x - ccx);   // unscaled offset for the expansion
			const float uy = static_cast<float>(me.y - ccy);
			const float uz = static_cast<float>(me.z - ccz);
			kx[t] = acc[0] + acc[3] * ux + acc[6] * uy + acc[7] * uz;
			ky[t] = acc[1] + acc[6] * ux + acc[4] * uy + acc[8] * uz;
			kz[t] = acc[2] + acc[7] * ux + acc[8] * uy + acc[5] * uz;
			if constexpr(kLocal)
			{
				float sx, sy, sz;
				local_terms(acc + 9, ux, uy, uz, sx, sy, sz);
				kx[t] += sx;
				ky[t] += sy;
				kz[t] += sz;
			}
		}
		if constexpr(kTwo)
		{
			if(use_list)
			{
				// the super-group's expansion about its own box centre
				const size_t sg = kRadix ? static_cast<size_t>(tree.sup[si].x)
				                         : (static_cast<size_t>(n) >> (6 + kSuperShift)) + si;
				const double* slo = bmin + 3 * sg;
				const double* shi = bmax + 3 * sg;
				const double scx = 0.5 * (slo[0] + shi[0]);
				const double scy = 0.5 * (slo[1] + shi[1]);
				const double scz = 0.5 * (slo[2] + shi[2]);
				const float* c = two.coef + 34 * static_cast<size_t>(si);
				float e[34];
				for(int k = 0; k < 34; ++k) { e[k] = c[k]; }
#pragma unroll
				for(int t = 0; t < kT; ++t)
				{
					const double4 me = wr[tn[t]].c;
					const float vx = static_cast<float>(me.x - scx), vy = static_cast<float>(me.y - scy);
					const float vz = static_cast<float>(me.z - scz);
					float sx, sy, sz;
					local_terms(e + 9, vx, vy, vz, sx, sy, sz);
					kx[t] += e[0] + e[3] * vx + e[6] * vy + e[7] * vz + sx;
					ky[t] += e[1] + e[6] * vx + e[4] * vy + e[8] * vz + sy;
					kz[t] += e[2] + e[7] * vx + e[8] * vy + e[5] * vz + sz;
				}
			}
		}
	}
#pragma unroll
	for(int t = 0; t < kT; ++t)
	{
		if(body_lane[t])
		{
			const int b = body_n[tn[t]];
			f[b] = y[3 * n + b];
			f[n + b] = y[4 * n + b];
			f[2 * n + b] = y[5 * n + b];
			f[3 * n + b] = static_cast<double>(ax[t]) - static_cast<double>(cx[t]) + static_cast<double>(kx[t]);
			f[4 * n + b] = static_cast<double>(ay[t]) - static_cast<double>(cy[t]) + static_cast<double>(ky[t]);
			f[5 * n + b] = static_cast<double>(az[t]) - static_cast<double>(cz[t]) + static_cast<double>(kz[t]);
		}
	}
	// counters: [0] pair interactions (entries x the group's bodies), [1] node
	// visits, [2] spills
	for(int off = 32; off > 0; off >>= 1)
	{
		vis += __shfl_down(vis, off, 64);
		taken += __shfl_down(taken, off, 64);
		spill += __shfl_down(spill, off, 64);
		nfar += __shfl_down(nfar, off, 64);
		nnear += __shfl_down(nnear, off, 64);
	}
	if(lane == 0)
	{
		atomicAdd(&counters[0], static_cast<unsigned long long>(taken) * gcount);
		atomicAdd(&counters[1], static_cast<unsigned long long>(vis));
		if(spill) { atomicAdd(&counters[2], static_cast<unsigned long long>(spill)); }
		if(nfar) { atomicAdd(&counters[3], static_cast<unsigned long long>(nfar)); }
		if(nnear) { atomicAdd(&counters[4], static_cast<unsigned long long>(nnear)); }
	}
}

// The group walk kernel, one work-group per group: the heap's groups in
// XCD-contiguous order, the Morton tree's in launch order.  (Measured at 4M
// Plummer bodies: persistent work-groups that take groups from counters, per
// XCD or in one queue, and chunked XCD orders all ran the Morton walk 9-19%
// slower than this.)
template <bool kQuad, bool kLocal = false, bool kTwo = false, bool kRadix = false>
__global__ void __launch_bounds__(kGroup)
k_bh_group32(const nbh_walk_rec* __restrict__ wr, const double* __restrict__ bmin, const double* __restrict__ bmax,
             const int32_t* __restrict__ body_n, const double* __restrict__ y, double* __restrict__ f, int n,
             float near_k, unsigned long long* __restrict__ counters, TwoLevel two = TwoLevel(),
             WalkTree tree = WalkTree())
{
	__shared__ int stack[kWalkStack];
	__shared__ float4 buf[kWalkBuf];
	__shared__ quad_lds<kQuad ? kQuadBuf : 1> qbuf;
	__shared__ int farn[kQuad ? kFarRing : 1];
	const int gi = kRadix ? static_cast<int>(blockIdx.x) : xcd_group(blockIdx.x, gridDim.x);
	group_walk_one<kQuad, kLocal, kTwo, kRadix>(gi, stack, buf, qbuf, farn, wr, bmin, bmax, body_n, y, f, n, near_k,
	                                             counters, two, tree);
}

// Super-group walk of the two-level local walk (TwoLevel): the group walk's
// loop with the super box, far cells into the super expansion, every other
// accepted node into the list.
template <bool kRadix = false>
__global__ void __launch_bounds__(kGroup)
k_bh_super(const nbh_walk_rec* __restrict__ wr, const double* __restrict__ bmin, const double* __restrict__ bmax,
           int n, float near_k, double inv_ratio, TwoLevel two, unsigned long long* __restrict__ counters,
           int stack_limit, WalkTree tree = WalkTree())
{
	__shared__ int stack[kWalkStack];
	__shared__ int farn[kFarRing];
	const int lane = threadIdx.x;
	const int si = kRadix ? static_cast<int>(blockIdx.x) : xcd_group(blockIdx.x, gridDim.x);
	// node of the super-group
	const size_t sg = kRadix ? static_cast<size_t>(tree.sup[si].x) : (static_cast<size_t>(n) >> (6 + kSuperShift)) + si;
	const double* glo = bmin + 3 * sg;
	const double* ghi = bmax + 3 * sg;
	const double gx0 = glo[0], gy0 = glo[1], gz0 = glo[2];
	const double gx1 = ghi[0], gy1 = ghi[1], gz1 = ghi[2];
	const double ccx = 0.5 * (gx0 + gx1), ccy = 0.5 * (gy0 + gy1), ccz = 0.5 * (gz0 + gz1);
	const double super_hd64 = half_diagonal(gx0, gy0, gz0, gx1, gy1, gz1);
	const float near_d = static_cast<float>(static_cast<double>(near_k) * super_hd64);
	const float super_hd = static_cast<float>(super_hd64);
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	int* list = two.lists + static_cast<size_t>(si) * two.cap;
	float acc[34];
	for(int k = 0; k < 34; ++k) { acc[k] = 0.0f; }
	unsigned long long vis = 0, nfar = 0;
	int lcount = 0;
	int nfar_listed = 0;
	bool overflow = false;
	int sp = 1;
	int fhead = 0, ftail = 0;
	if(lane == 0) { stack[0] = walk_start<kRadix>(); }
	__syncthreads();
	while(sp > 0)
	{
		bool valid;
		int node;
		double4 rec;
		double m;
		int2 kids;
		walk_fetch<kRadix>(stack, sp, lane, wr, node, valid, rec, m, kids, tree.child, n);
		if(ftail - fhead >= kGroup)
		{
			far_batch<true>(farn, wr, fhead, kGroup, lane, ccx, ccy, ccz, acc);
			fhead += kGroup;
		}
		__syncthreads();
		bool accept = false, open = false, far = false;
		float ux = 0.0f, uy = 0.0f, uz = 0.0f;
		if(valid)
		{
			ux = static_cast<float>(rec.x - ccx);
			uy = static_cast<float>(rec.y - ccy);
			uz = static_cast<float>(rec.z - ccz);
			if(node >= n)
			{
				accept = true;
			}
			else
			{
				const double ex = fmax(fmax(gx0 - rec.x, rec.x - gx1), 0.0);
				const double ey = fmax(fmax(gy0 - rec.y, rec.y - gy1), 0.0);
				const double ez = fmax(fmax(gz0 - rec.z, rec.z - gz1), 0.0);
				accept = ex * ex + ey * ey + ez * ez > rec.w;
				const float du = sqrtf(ux * ux + uy * uy + uz * uz);
				far = accept && !(du < near_d);
				// a rejected node whose bodies all lie within near_d of the centre
				// (ball of the cell radius sqrt(crit_r2) / ratio about its mass
				// centre) is left to the groups; others are opened here
				// (in the Morton tree also one wider than half the super box: it is
				// opened here once rather than by every walk group, which lowers the
				// error at the same cost)
				const float rad = static_cast<float>(sqrt(rec.w) * inv_ratio);
				open = !accept && (!(du + rad < near_d) || (kRadix && rad > 0.5f * super_hd));
			}
		}
		const bool tolist = valid && !far && !open;
		if(sp + __popcll(__ballot(open)) > stack_limit)
		{
			overflow = true;           // the groups of this super-group will walk for themselves
			open = false;
		}
		const uint64_t bopen = __ballot(open);
		const uint64_t bfar = __ballot(far);
		const uint64_t blist = __ballot(tolist);
		if(far)
		{
			farn[(ftail + __popcll(bfar & lt)) & (kFarRing - 1)] = node;
			if(two.far_list != nullptr)
			{
				const int fslot = nfar_listed + __popcll(bfar & lt);
				if(fslot < two.far_cap) { two.far_list[static_cast<size_t>(si) * two.far_cap + fslot] = node; }
			}
		}
		nfar_listed += __popcll(bfar);
		if(tolist)
		{
			const int slot = lcount + __popcll(blist & lt);
			if(slot < two.cap) { list[slot] = node; }
		}
		if(open) { stack[sp + __popcll(bopen & lt)] = walk_entry<kRadix>(node, kids, n); }
		ftail += __popcll(bfar);
		lcount += __popcll(blist);
		sp += __popcll(bopen);
		vis += valid ? 1 : 0;
		nfar += far ? 1 : 0;
		__syncthreads();
	}
	for(; fhead < ftail; fhead += kGroup)
	{
		far_batch<true>(farn, wr, fhead, ftail - fhead < kGroup ? ftail - fhead : kGroup, lane, ccx, ccy, ccz, acc);
	}
	for(int sh = 1; sh < 64; sh <<= 1)
	{
		for(int k = 0; k < 34; ++k) { acc[k] += __shfl_xor(acc[k], sh, 64); }
	}
	overflow = __ballot(overflow) != 0;
	if(lane == 0)
	{
		float* c = two.coef + 34 * static_cast<size_t>(si);
		for(int k = 0; k < 34; ++k) { c[k] = acc[k]; }
		two.count[si] = overflow || lcount > two.cap ? two.cap + 1 : lcount;
		if(two.far_count != nullptr) { two.far_count[si] = nfar_listed; }
	}
	for(int off = 32; off > 0; off >>= 1)
	{
		vis += __shfl_down(vis, off, 64);
		nfar += __shfl_down(nfar, off, 64);
	}
	if(lane == 0)
	{
		atomicAdd(&counters[1], vis);
		if(nfar) { atomicAdd(&counters[3], nfar); }
		if(overflow || lcount > two.cap) { atomicAdd(&counters[2], 1ull); }
	}
}

// ---------------------------------------------------------------------------
// learned correction: samples, and the group walk with the network
// ---------------------------------------------------------------------------
// Internal cells the group walk takes as monopoles are split by the distance
// dc from the group's box centre to their mass centre, in units of the box's
// half-diagonal h:
//   near (dc < near_k h): the cell's second moments are applied to every body
//        of the group (quadrupole term), since its field varies too much over
//        the group for one expansion about the centre;
//   far  (otherwise), with bounding radius r > rho_min dc: the network predicts
//        the cell's residual field at the centre and its Jacobian.
enum { kCellNone = 0, kCellNear = 1, kCellFar = 2 };

__device__ inline double cell_radius(const double* lo, const double* hi, double cx, double cy, double cz)
{
	const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
	const double gx = 0.5 * (hi[0] + lo[0]) - cx, gy = 0.5 * (hi[1] + lo[1]) - cy, gz = 0.5 * (hi[2] + lo[2]) - cz;
	return 0.5 * sqrt(ex * ex + ey * ey + ez * ez) + sqrt(gx * gx + gy * gy + gz * gz);
}

__device__ inline int classify_cell(const double* lo, const double* hi, const double4& rec, double ccx, double ccy,
                                    double ccz, double half_diag, float near_k, float rho_min, double* dc_out)
{
	const double dx = rec.x - ccx, dy = rec.y - ccy, dz = rec.z - ccz;
	const double dc = sqrt(dx * dx + dy * dy + dz * dz);
	*dc_out = dc;
	if(dc < static_cast<double>(near_k) * half_diag) { return kCellNear; }
	const double r = cell_radius(lo, hi, rec.x, rec.y, rec.z);
	return r > static_cast<double>(rho_min) * dc ? kCellFar : kCellNone;
}

// Group walk without forces: emit (group ordinal, node) for every far cell.
__global__ void __launch_bounds__(kGroup)
k_band_collect(const double4* __restrict__ xyzr, const double* __restrict__ bmin, const double* __restrict__ bmax,
               int n, const uint32_t* __restrict__ groups, float rho_min, float near_k, uint32_t* __restrict__ pairs,
               unsigned long long capacity, unsigned long long* __restrict__ counter)
{
	__shared__ int stack[kWalkStack];
	const int lane = threadIdx.x;
	const uint32_t go = groups[blockIdx.x];
	const int g = n / kGroup + static_cast<int>(go);
	const double gx0 = bmin[3 * g], gy0 = bmin[3 * g + 1], gz0 = bmin[3 * g + 2];
	const double gx1 = bmax[3 * g], gy1 = bmax[3 * g + 1], gz1 = bmax[3 * g + 2];
	const double ccx = 0.5 * (gx0 + gx1), ccy = 0.5 * (gy0 + gy1), ccz = 0.5 * (gz0 + gz1);
	const double hd = half_diagonal(gx0, gy0, gz0, gx1, gy1, gz1);
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	int sp = 1;
	if(lane == 0) { stack[0] = kRoot; }
	__syncthreads();
	while(sp > 0)
	{
		bool valid;
		const int node = walk_pop(stack, sp, lane, valid);
		__syncthreads();
		bool open = false, band = false;
		if(valid && node < n)
		{
			const double4 rec = xyzr[node];
			const double ex = fmax(fmax(gx0 - rec.x, rec.x - gx1), 0.0);
			const double ey = fmax(fmax(gy0 - rec.y, rec.y - gy1), 0.0);
			const double ez = fmax(fmax(gz0 - rec.z, rec.z - gz1), 0.0);
			open = !(ex * ex + ey * ey + ez * ez > rec.w);
			if(!open)
			{
				double dc;
				band = classify_cell(bmin + 3 * node, bmax + 3 * node, rec, ccx, ccy, ccz, hd, near_k, rho_min, &dc) ==
				       kCellFar;
			}
		}
		const int nopen = __popcll(__ballot(open));
		if(sp + nopen > kWalkStack) { open = false; }
		const uint64_t bopen = __ballot(open);
		if(open) { stack[sp + __popcll(bopen & lt)] = 2 * node; }
		sp += __popcll(bopen);
		const uint64_t bband = __ballot(band);
		const int nband = __popcll(bband);
		if(nband > 0)
		{
			unsigned long long base = 0;
			if(lane == 0) { base = atomicAdd(counter, static_cast<unsigned long long>(nband)); }
			base = __shfl(base, 0, 64);
			const unsigned long long slot = base + __popcll(bband & lt);
			if(band && slot < capacity)
			{
				pairs[2 * slot] = go;
				pairs[2 * slot + 1] = static_cast<uint32_t>(node);
			}
		}
		__syncthreads();
	}
}

// Quadrupole term of a cell's field at p and its Jacobian, FP64: Q the second
// central moments (xx yy zz xy xz yz) about the mass centre cm; with
// r = cm - p, d = |r|, Q' = 3 Q - tr(Q) I and s = r.Q'.r,
//   g = (2.5 s / d^2 r - Q' r) / d^5,
//   J = dg/dp = -(5 (r (Q'r)' + (Q'r) r') + 2.5 s I) / d^7 + 17.5 s r r' / d^9 + Q' / d^5
// (the terms quad_expansion evaluates in FP32); out[9] = g (3), J (6).
__device__ inline void quad_field64(double rx, double ry, double rz, const double* Q, double* out)
{
	const double tr = Q[0] + Q[1] + Q[2];
	const double qxx = 3.0 * Q[0] - tr, qyy = 3.0 * Q[1] - tr, qzz = 3.0 * Q[2] - tr;
	const double qxy = 3.0 * Q[3], qxz = 3.0 * Q[4], qyz = 3.0 * Q[5];
	const double Qx = qxx * rx + qxy * ry + qxz * rz;
	const double Qy = qxy * rx + qyy * ry + qyz * rz;
	const double Qz = qxz * rx + qyz * ry + qzz * rz;
	const double sq = rx * Qx + ry * Qy + rz * Qz;
	const double d2 = rx * rx + ry * ry + rz * rz;
	const double d = sqrt(d2);
	const double i5 = 1.0 / (d2 * d2 * d);
	const double i7 = i5 / d2;
	const double c = 2.5 * sq * i7;
	out[0] = c * rx - Qx * i5;
	out[1] = c * ry - Qy * i5;
	out[2] = c * rz - Qz * i5;
	const double a = 5.0 * i7;
	const double e = 17.5 * sq * i7 / d2;
	out[3] = -(a * (2.0 * rx * Qx) + c - e * rx * rx - qxx * i5);
	out[4] = -(a * (2.0 * ry * Qy) + c - e * ry * ry - qyy * i5);
	out[5] = -(a * (2.0 * rz * Qz) + c - e * rz * rz - qzz * i5);
	out[6] = -(a * (rx * Qy + ry * Qx) - e * rx * ry - qxy * i5);
	out[7] = -(a * (rx * Qz + rz * Qx) - e * rx * rz - qxz * i5);
	out[8] = -(a * (ry * Qz + rz * Qy) - e * ry * rz - qyz * i5);
}

// One wave per sample (centre node centre_base + pairs[2s], cell pairs[2s+1]):
// the exact FP64 field of the cell's bodies at the centre node's box centre p
// and its Jacobian, minus those of the monopole (order 1) or of the monopole
// and the quadrupole term (order 2), written as features x[16], basis b[16] and
// target t[9] = [R / w, kJac J d / w] (J as xx yy zz xy xz yz), where w is the
// scale of cell_features (M rho^2 / d^2) times rho^(order - 1).
__global__ void __launch_bounds__(64)
k_band_residual(const double4* __restrict__ xyzr, const double* __restrict__ nmass, const double* __restrict__ bmin,
                const double* __restrict__ bmax, const double* __restrict__ quad, int n, int levels, int centre_base,
                int order, const uint32_t* __restrict__ pairs, size_t count, float* __restrict__ x,
                float* __restrict__ bs, float* __restrict__ t)
{
	const int lane = threadIdx.x;
	for(size_t s = blockIdx.x; s < count; s += gridDim.x)
	{
		const int g = centre_base + static_cast<int>(pairs[2 * s]);
		const int c = static_cast<int>(pairs[2 * s + 1]);
		const double px = 0.5 * (bmin[3 * g] + bmax[3 * g]);
		const double py = 0.5 * (bmin[3 * g + 1] + bmax[3 * g + 1]);
		const double pz = 0.5 * (bmin[3 * g + 2] + bmax[3 * g + 2]);
		const int depth = 31 - __builtin_clz(static_cast<unsigned>(c));
		const int span = levels - depth;
		const int leaf0 = c << span, leaf1 = (c + 1) << span;
		double a[9];
		for(int k = 0; k < 9; ++k) { a[k] = 0.0; }
		for(int leaf = leaf0 + lane; leaf < leaf1; leaf += 64)
		{
			const double4 b = xyzr[leaf];
			const double dx = b.x - px, dy = b.y - py, dz = b.z - pz;
			double r2 = dx * dx + dy * dy + dz * dz;
			if(r2 < 1e-8) { r2 = 1e-8; }
			const double w = nmass[leaf] / (r2 * sqrt(r2));
			const double w5 = 3.0 * w / r2;
			a[0] += dx * w;
			a[1] += dy * w;
			a[2] += dz * w;
			a[3] += w5 * dx * dx - w;
			a[4] += w5 * dy * dy - w;
			a[5] += w5 * dz * dz - w;
			a[6] += w5 * dx * dy;
			a[7] += w5 * dx * dz;
			a[8] += w5 * dy * dz;
		}
		for(int off = 32; off > 0; off >>= 1)
		{
			for(int k = 0; k < 9; ++k) { a[k] += __shfl_down(a[k], off, 64); }
		}
		if(lane == 0)
		{
			const double4 rec = xyzr[c];
			const double M = nmass[c];
			const double dx = rec.x - px, dy = rec.y - py, dz = rec.z - pz;
			const double d2 = dx * dx + dy * dy + dz * dz;
			const double d = sqrt(d2);
			const double w = M / (d2 * d);
			const double w5 = 3.0 * w / d2;
			a[0] -= dx * w;
			a[1] -= dy * w;
			a[2] -= dz * w;
			a[3] -= w5 * dx * dx - w;
			a[4] -= w5 * dy * dy - w;
			a[5] -= w5 * dz * dz - w;
			a[6] -= w5 * dx * dy;
			a[7] -= w5 * dx * dz;
			a[8] -= w5 * dy * dz;
			if(order >= 2)
			{
				double q[9];
				quad_field64(dx, dy, dz, quad + 6 * c, q);
				for(int k = 0; k < 9; ++k) { a[k] -= q[k]; }
			}
			float f[nbh_net::kIn];
			float bv[nbh_net::kBasis];
			double scale = nbh_net::cell_features(px, py, pz, rec.x, rec.y, rec.z, M, bmin + 3 * c, bmax + 3 * c,
			                                      quad + 6 * c, f, bv);
			if(order >= 2) { scale *= static_cast<double>(f[0]); }    // f[0] = rho
			for(int k = 0; k < nbh_net::kIn; ++k) { x[s * nbh_net::kIn + k] = f[k]; }
			for(int k = 0; k < nbh_net::kBasis; ++k) { bs[s * nbh_net::kBasis + k] = bv[k]; }
			for(int k = 0; k < 3; ++k) { t[s * nbh_net::kOut + k] = static_cast<float>(a[k] / scale); }
			for(int k = 3; k < 9; ++k) { t[s * nbh_net::kOut + k] = static_cast<float>(a[k] * nbh_net::kJac * d / scale); }
		}
	}
}

// Learned correction of the two-level walk's shared far field: one wave per
// super-group evaluates the network on the super-group's far cells (listed by
// k_bh_super), 16 at a time on the matrix cores, and adds the predicted
// remainder after the monopole and quadrupole terms -- field R and Jacobian J
// at the super box centre, R / w = c0 uh + c1 w + c2 o and kJac J d / w =
// (basis of nbh_mlp_dev.h), w = M rho^3 / d^2 -- to the super-group's
// expansion coefficients (field and Jacobian slots), which every body of the
// super-group applies.  Features are formed in FP64 and rounded once.
__global__ void __launch_bounds__(64)
k_super_net(const double4* __restrict__ xyzr, const double* __restrict__ nmass, const double* __restrict__ bmin,
            const double* __restrict__ bmax, const double* __restrict__ quad, int n, TwoLevel two,
            const uint2* __restrict__ frags, const float* __restrict__ mlp_p)
{
	__shared__ __align__(16) uint16_t tf[16 * nbh_net::kIn];
	__shared__ float tb[16 * nbh_net::kBasis];
	__shared__ float tw[2 * 16];
	__shared__ __align__(16) uint16_t mlds[2 * 16 * nbh_net::kHid];
	const int lane = threadIdx.x;
	const int col = lane & 15;
	const int si = blockIdx.x;
	const int sg = (n >> (6 + kSuperShift)) + si;
	const double px = 0.5 * (bmin[3 * sg] + bmax[3 * sg]);
	const double py = 0.5 * (bmin[3 * sg + 1] + bmax[3 * sg + 1]);
	const double pz = 0.5 * (bmin[3 * sg + 2] + bmax[3 * sg + 2]);
	const int rows = min(two.far_count[si], two.far_cap);
	const int* list = two.far_list + static_cast<size_t>(si) * two.far_cap;
	float acc[nbh_net::kOut];
	for(int k = 0; k < nbh_net::kOut; ++k) { acc[k] = 0.0f; }
	for(int base = 0; base < rows; base += 16)
	{
		if(lane < 16)
		{
			const int r = base + lane;
			float f[nbh_net::kIn];
			float bv[nbh_net::kBasis];
			float w0 = 0.0f, w1 = 0.0f;
			if(r < rows)
			{
				const int c = list[r];
				const double4 rec = xyzr[c];
				const double scale = nbh_net::cell_features(px, py, pz, rec.x, rec.y, rec.z, nmass[c], bmin + 3 * c,
				                                            bmax + 3 * c, quad + 6 * c, f, bv) *
				                     static_cast<double>(f[0]);
				const double dx = rec.x - px, dy = rec.y - py, dz = rec.z - pz;
				const double d = sqrt(dx * dx + dy * dy + dz * dz);
				w0 = static_cast<float>(scale);
				w1 = static_cast<float>(scale / (nbh_net::kJac * d));
			}
			else
			{
				for(int k = 0; k < nbh_net::kIn; ++k) { f[k] = 0.0f; }
				for(int k = 0; k < nbh_net::kBasis; ++k) { bv[k] = 0.0f; }
			}
			for(int k = 0; k < nbh_net::kIn; ++k) { tf[lane * nbh_net::kIn + k] = nbh_net::to_bf16(f[k]); }
			for(int k = 0; k < nbh_net::kBasis; ++k) { tb[lane * nbh_net::kBasis + k] = bv[k]; }
			tw[2 * lane] = w0;
			tw[2 * lane + 1] = w1;
		}
		__syncthreads();
		const nbh_net::bf16x4 xa = nbh_net::lds_afrag(tf, nbh_net::kIn, 0, lane);
		const nbh_net::f32x4 o = nbh_net::forward_tile(xa, frags, mlp_p, mlds, lane, nullptr, nullptr);
		for(int e = 0; e < 4; ++e)
		{
			const int row = 4 * (lane >> 4) + e;
			float v[nbh_net::kOut];
			nbh_net::basis_col(tb + row * nbh_net::kBasis, col, v);
			const float c0 = o[e] * tw[2 * row], c1 = o[e] * tw[2 * row + 1];
			for(int k = 0; k < 3; ++k) { acc[k] = fmaf(c0, v[k], acc[k]); }
			for(int k = 3; k < nbh_net::kOut; ++k) { acc[k] = fmaf(c1, v[k], acc[k]); }
		}
		__syncthreads();
	}
	for(int sh = 1; sh < 64; sh <<= 1)
	{
		for(int k = 0; k < nbh_net::kOut; ++k) { acc[k] += __shfl_xor(acc[k], sh, 64); }
	}
	if(lane == 0)
	{
		float* c = two.coef + 34 * static_cast<size_t>(si);
		for(int k = 0; k < nbh_net::kOut; ++k) { c[k] += acc[k]; }
	}
}

// Training pairs (super-group ordinal, far cell) of the super-groups sel[g]:
// block g copies the far list of super-group sel[g] to pairs[off[g] ...].
__global__ void __launch_bounds__(64)
k_super_pairs(TwoLevel two, const uint32_t* __restrict__ sel, const unsigned long long* __restrict__ off,
              uint32_t* __restrict__ pairs, unsigned long long capacity)
{
	const uint32_t si = sel[blockIdx.x];
	const int rows = min(two.far_count[si], two.far_cap);
	const int* list = two.far_list + static_cast<size_t>(si) * two.far_cap;
	for(int r = threadIdx.x; r < rows; r += blockDim.x)
	{
		const unsigned long long slot = off[blockIdx.x] + static_cast<unsigned long long>(r);
		if(slot < capacity)
		{
			pairs[2 * slot] = si;
			pairs[2 * slot + 1] = static_cast<uint32_t>(list[r]);
		}
	}
}

constexpr int kBand = 80;        // far-cell rows waiting for the network (< 16 + 64)
// k_bh_group32 plus the correction.  Near cells enter the monopole +
// quadrupole sweep, every other taken node the monopole sweep; far cells are
// fed to the network in tiles of 16, and the predicted residual field of each,
// expanded to first order about the group centre, is summed over the group's
// far cells into C (3) and J (6), the correction of a body at offset delta from
// the centre being C + J delta.
__global__ void __launch_bounds__(kGroup)
k_bh_group32_mlp(const nbh_walk_rec* __restrict__ wr, const double* __restrict__ bmin,
                 const double* __restrict__ bmax, const int32_t* __restrict__ body_n,
                 const double* __restrict__ y, double* __restrict__ f, int n, const uint2* __restrict__ frags,
                 const float* __restrict__ mlp_p, float rho_min, float near_k,
                 unsigned long long* __restrict__ counters)
{
	__shared__ int stack[kWalkStack];
	__shared__ float4 buf[kWalkBuf];
	__shared__ quad_lds<kQuadBuf> qbuf;
	__shared__ __align__(16) uint16_t bandf[kBand * nbh_net::kIn];
	__shared__ float bandw[2 * kBand];      // per far row: scale, scale / (kJac d)
	__shared__ float bandb[kBand * nbh_net::kBasis];
	__shared__ __align__(16) uint16_t mlds[2 * 16 * nbh_net::kHid];
	const int lane = threadIdx.x;
	const int col = lane & 15;
	const int gi = xcd_group(blockIdx.x, gridDim.x);
	const int g = n / kGroup + gi;
	const double gx0 = bmin[3 * g], gy0 = bmin[3 * g + 1], gz0 = bmin[3 * g + 2];
	const double gx1 = bmax[3 * g], gy1 = bmax[3 * g + 1], gz1 = bmax[3 * g + 2];
	const double ccx = 0.5 * (gx0 + gx1), ccy = 0.5 * (gy0 + gy1), ccz = 0.5 * (gz0 + gz1);
	const double hd = half_diagonal(gx0, gy0, gz0, gx1, gy1, gz1);
	const int tn = n + gi * kGroup + lane;
	const double4 me = wr[tn].c;
	const float tx = static_cast<float>((me.x - ccx) * kSweepLen);   // scaled, for the sweeps
	const float ty = static_cast<float>((me.y - ccy) * kSweepLen);
	const float tz = static_cast<float>((me.z - ccz) * kSweepLen);
	float ax = 0.0f, ay = 0.0f, az = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
	float acc[nbh_net::kOut];               // this lane's share of the far-cell correction (R, J)
	for(int k = 0; k < nbh_net::kOut; ++k) { acc[k] = 0.0f; }
	const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	unsigned long long vis = 0, taken = 0, spill = 0, corrected = 0, nearq = 0;

	// Evaluate full tiles of 16 far rows (and, with flush, the zero-padded
	// remainder); rows not consumed move to the front.
	auto run_tiles = [&](int& nb, bool flush) {
		while(nb >= 16 || (flush && nb > 0))
		{
			if(nb < 16)
			{
				for(int r = nb + (lane >> 4); r < 16; r += 4)
				{
					bandf[r * nbh_net::kIn + col] = 0;
					bandb[r * nbh_net::kBasis + col] = 0.0f;
					if(col < 2) { bandw[2 * r + col] = 0.0f; }
				}
				__syncthreads();
			}
			const nbh_net::bf16x4 xa = nbh_net::lds_afrag(bandf, nbh_net::kIn, 0, lane);
			const nbh_net::f32x4 o = nbh_net::forward_tile(xa, frags, mlp_p, mlds, lane, nullptr, nullptr);
			for(int e = 0; e < 4; ++e)
			{
				const int row = 4 * (lane >> 4) + e;
				float v[nbh_net::kOut];
				nbh_net::basis_col(bandb + row * nbh_net::kBasis, col, v);
				const float c0 = o[e] * bandw[2 * row], c1 = o[e] * bandw[2 * row + 1];
				for(int k = 0; k < 3; ++k) { acc[k] = fmaf(c0, v[k], acc[k]); }
				for(int k = 3; k < nbh_net::kOut; ++k) { acc[k] = fmaf(c1, v[k], acc[k]); }
			}
			const int rest = nb > 16 ? nb - 16 : 0;
			__syncthreads();
			for(int r = lane >> 4; r < rest; r += 4)
			{
				bandf[r * nbh_net::kIn + col] = bandf[(r + 16) * nbh_net::kIn + col];
				bandb[r * nbh_net::kBasis + col] = bandb[(r + 16) * nbh_net::kBasis + col];
				if(col < 2) { bandw[2 * r + col] = bandw[2 * (r + 16) + col]; }
			}
			__syncthreads();
			nb = rest;
		}
	};

	int sp = 1;
	int nbuf = 0;
	int nband = 0;
	int nq = 0;
	if(lane == 0) { stack[0] = kRoot; }
	__syncthreads();
	while(sp > 0)
	{
		bool valid;
		int node;
		double4 rec;
		double m;
		// node loads first, then the pending sweeps (same order as k_bh_group32)
		int2 kids;
		walk_fetch(stack, sp, lane, wr, node, valid, rec, m, kids);
		if(nbuf > kWalkBuf - kGroup)
		{
			sweep32(buf, nbuf, tx, ty, tz, ax, ay, az, cx, cy, cz);
			nbuf = 0;
		}
		if(nq > kQuadBuf - kGroup)
		{
			sweep_mq(qbuf, nq, tx, ty, tz, ax, ay, az, cx, cy, cz);
			nq = 0;
		}
		__syncthreads();
		bool accept = false, open = false;
		int cls = kCellNone;
		float4 ent = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		double dc = 0.0;
		if(valid)
		{
			ent = sweep_entry(rec, m, ccx, ccy, ccz);
			if(node >= n)
			{
				accept = true;
			}
			else
			{
				const double ex = fmax(fmax(gx0 - rec.x, rec.x - gx1), 0.0);
				const double ey = fmax(fmax(gy0 - rec.y, rec.y - gy1), 0.0);
				const double ez = fmax(fmax(gz0 - rec.z, rec.z - gz1), 0.0);
				accept = ex * ex + ey * ey + ez * ez > rec.w;
				open = !accept;
				if(accept)
				{
					cls = classify_cell(bmin + 3 * node, bmax + 3 * node, rec, ccx, ccy, ccz, hd, near_k, rho_min, &dc);
				}
			}
		}
		const int nopen = __popcll(__ballot(open));
		if(sp + nopen > kWalkStack)
		{
			spill += open ? 1 : 0;
			accept = accept || open;
			open = false;
		}
		const bool band = cls == kCellFar;
		const bool isnear = cls == kCellNear;
		const uint64_t bacc = __ballot(accept);
		const uint64_t bopen = __ballot(open);
		const uint64_t bband = __ballot(band);
		const uint64_t bnear = __ballot(isnear);
		const uint64_t bmono = bacc & ~bnear;     // near cells go to the monopole + quadrupole sweep
		if(accept && !isnear) { buf[nbuf + __popcll(bmono & lt)] = ent; }
		if(open) { stack[sp + __popcll(bopen & lt)] = 2 * node; }
		if(band)
		{
			const int slot = nband + __popcll(bband & lt);
			float fe[nbh_net::kIn];
			float be[nbh_net::kBasis];
			const float scale = nbh_net::cell_features_fast(
			    static_cast<float>(rec.x - ccx), static_cast<float>(rec.y - ccy), static_cast<float>(rec.z - ccz),
			    rec.x, rec.y, rec.z, m, bmin + 3 * node, bmax + 3 * node, wr[node].q, fe, be);
			for(int k = 0; k < nbh_net::kBasis; ++k) { bandb[slot * nbh_net::kBasis + k] = be[k]; }
			bandw[2 * slot] = scale;
			bandw[2 * slot + 1] = scale / (static_cast<float>(nbh_net::kJac) * static_cast<float>(dc));
			for(int k = 0; k < nbh_net::kIn; ++k) { bandf[slot * nbh_net::kIn + k] = nbh_net::to_bf16(fe[k]); }
		}
		if(isnear)
		{
			put_quad(qbuf, nq + __popcll(bnear & lt), ent, wr[node].q);
		}
		nbuf += __popcll(bmono);
		sp += __popcll(bopen);
		nband += __popcll(bband);
		nq += __popcll(bnear);
		corrected += band ? 1 : 0;
		nearq += isnear ? 1 : 0;
		vis += valid ? 1 : 0;
		taken += accept ? 1 : 0;
		__syncthreads();
		if(nband >= 16) { run_tiles(nband, false); }
	}
	if(nbuf > 0)
	{
		sweep32(buf, nbuf, tx, ty, tz, ax, ay, az, cx, cy, cz);
	}
	if(nq > 0)
	{
		sweep_mq(qbuf, nq, tx, ty, tz, ax, ay, az, cx, cy, cz);
	}
	__syncthreads();
	run_tiles(nband, true);
	for(int sh = 1; sh < 64; sh <<= 1)
	{
		for(int k = 0; k < nbh_net::kOut; ++k) { acc[k] += __shfl_xor(acc[k], sh, 64); }
	}
	const float* S = acc;
	const float ux = static_cast<float>(me.x - ccx);   // unscaled offset for the correction
	const float uy = static_cast<float>(me.y - ccy);
	const float uz = static_cast<float>(me.z - ccz);
	const float kx = S[0] + S[3] * ux + S[6] * uy + S[7] * uz;
	const float ky = S[1] + S[6] * ux + S[4] * uy + S[8] * uz;
	const float kz = S[2] + S[7] * ux + S[8] * uy + S[5] * uz;
	const int b = body_n[tn];
	f[b] = y[3 * n + b];
	f[n + b] = y[4 * n + b];
	f[2 * n + b] = y[5 * n + b];
	f[3 * n + b] = static_cast<double>(ax) - static_cast<double>(cx) + static_cast<double>(kx);
	f[4 * n + b] = static_cast<double>(ay) - static_cast<double>(cy) + static_cast<double>(ky);
	f[5 * n + b] = static_cast<double>(az) - static_cast<double>(cz) + static_cast<double>(kz);
	for(int off = 32; off > 0; off >>= 1)
	{
		vis += __shfl_down(vis, off, 64);
		taken += __shfl_down(taken, off, 64);
		spill += __shfl_down(spill, off, 64);
		corrected += __shfl_down(corrected, off, 64);
		nearq += __shfl_down(nearq, off, 64);
	}
	if(lane == 0)
	{
		atomicAdd(&counters[0], taken * kGroup);
		atomicAdd(&counters[1], vis);
		if(spill) { atomicAdd(&counters[2], spill); }
		atomicAdd(&counters[3], corrected);
		atomicAdd(&counters[4], nearq);
	}
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
nbh_status scan_exclusive(nbh_ctx* ctx, uint32_t* c, size_t m, uint32_t* scan_part)
{
	if(m <= static_cast<size_t>(kScanOneThreads) * kScanOneItems)
	{
		hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneThreads), 0, ctx->stream, c, static_cast<int>(m));
		NBH_LAUNCHED();
		return NBH_OK;
	}
	size_t chunk = (m + kScanParts - 1) / kScanParts;
	chunk = (chunk + 4 * kScanThreads - 1) / (4 * kScanThreads) * (4 * kScanThreads);
	const unsigned g = static_cast<unsigned>((m + chunk - 1) / chunk);
	hipLaunchKernelGGL(k_scan_reduce, dim3(g), dim3(kScanThreads), 0, ctx->stream, c, m, chunk, scan_part);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(kScanParts), 0, ctx->stream, scan_part, static_cast<int>(g));
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_scan_apply, dim3(g), dim3(kScanThreads), 0, ctx->stream, c, m, chunk, scan_part);
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status radix_sort(nbh_ctx* ctx, nbh_tree* t, const double* coord, uint32_t* out_vals)
{
	const size_t n = t->n;
	hipLaunchKernelGGL(k_make_keys, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, coord, t->keys,
	                   t->vals, n);
	NBH_LAUNCHED();
	uint32_t* sorted = nullptr;
	NBH_TRY(nbh::sort_pairs(ctx, t->keys, t->vals, t->keys + n, t->vals + n, n, 64, t->counts, t->scan_part,
	                        &sorted));
	NBH_HIP(hipMemcpyAsync(out_vals, sorted, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
	return NBH_OK;
}

// Packed walk records of nodes 1 .. count-1.
__global__ void __launch_bounds__(256)
k_pack_records(const double4* __restrict__ xyzr, const double* __restrict__ mass, const double* __restrict__ quad,
               nbh_walk_rec* __restrict__ wr, size_t count)
{
	const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
	for(size_t i = 1 + static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
	{
		nbh_walk_rec r;
		r.c = xyzr[i];
		r.m = mass[i];
		for(int k = 0; k < 6; ++k) { r.q[k] = static_cast<float>(quad[6 * i + k]); }
		wr[i] = r;
	}
}

nbh_status moments(nbh_ctx* ctx, nbh_tree* t, const double* y, const double* mass, double ratio, bool build)
{
	t->ratio = ratio;
	const size_t n = t->n;
	hipLaunchKernelGGL(k_leaves, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, t->list[0], y, mass,
	                   n, t->xyzr, t->mass, t->bmin, t->bmax, t->body_n, build ? 1 : 0);
	NBH_LAUNCHED();
	const double ratio_sqr = ratio * ratio;
	size_t lo = n / 2;
	for(; lo >= 512; lo >>= 1)
	{
		hipLaunchKernelGGL(k_level, dim3(nbh::grid_for(ctx, lo, 256)), dim3(256), 0, ctx->stream, lo, 2 * lo,
		                   t->xyzr, t->mass, t->bmin, t->bmax, t->quad, ratio_sqr);
		NBH_LAUNCHED();
	}
	if(lo >= 1)
	{
		hipLaunchKernelGGL(k_levels_top, dim3(1), dim3(256), 0, ctx->stream, lo, t->xyzr, t->mass, t->bmin, t->bmax,
		                   t->quad, ratio_sqr);
		NBH_LAUNCHED();
	}
	hipLaunchKernelGGL(k_pack_records, dim3(nbh::grid_for(ctx, 2 * n, 256)), dim3(256), 0, ctx->stream, t->xyzr,
	                   t->mass, t->quad, t->wrec, 2 * n);
	NBH_LAUNCHED();
	return NBH_OK;
}

}  // namespace

namespace nbh {

size_t sort_counts_size(size_t n)
{
	return ((n + kSortTile - 1) / kSortTile) * 256 + 1;
}

size_t sort_scan_size() { return kScanParts; }

nbh_status sort_pairs(nbh_ctx* ctx, uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t n, int bits,
                      uint32_t* counts, uint32_t* scan_part, uint32_t** sorted_vals)
{
	const int nblocks = static_cast<int>((n + kSortTile - 1) / kSortTile);
	for(int shift = 0; shift < bits; shift += 8)
	{
		hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kSortThreads), 0, ctx->stream, k0, n, shift, counts,
		                   nblocks);
		NBH_LAUNCHED();
		NBH_TRY(scan_exclusive(ctx, counts, static_cast<size_t>(nblocks) * 256, scan_part));
		hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kSortThreads), 0, ctx->stream, k0, v0, k1, v1, n,
		                   shift, counts, nblocks);
		NBH_LAUNCHED();
		std::swap(k0, k1);
		std::swap(v0, v1);
	}
	*sorted_vals = v0;
	return NBH_OK;
}

}  // namespace nbh

extern "C" {

// The buffers a tree allocates at creation, with their sizes in bytes.
constexpr int kTreeBuffers = 17;
static void tree_buffers(nbh_tree* t, void** ptrs[kTreeBuffers], size_t sizes[kTreeBuffers])
{
	const size_t n = t->n;
	const size_t nb = (n + kSortTile - 1) / kSortTile;
	const size_t ncounts = std::max(nb * 256, (n + kPartTile - 1) / kPartTile) + 1;
	int k = 0;
	auto add = [&](void* p, size_t bytes) {
		ptrs[k] = static_cast<void**>(p);
		sizes[k] = bytes;
		++k;
	};
	for(int a = 0; a < 3; ++a) { add(&t->list[a], n * sizeof(uint32_t)); }
	add(&t->tmp, n * sizeof(uint32_t));
	add(&t->side, n);
	add(&t->keys, 2 * n * sizeof(uint64_t));
	add(&t->vals, 2 * n * sizeof(uint32_t));
	add(&t->counts, ncounts * sizeof(uint32_t));
	add(&t->scan_part, kScanParts * sizeof(uint32_t));
	add(&t->xyzr, 2 * n * sizeof(double4));
	add(&t->mass, 2 * n * sizeof(double));
	add(&t->bmin, 2 * n * 3 * sizeof(double));
	add(&t->bmax, 2 * n * 3 * sizeof(double));
	add(&t->quad, 2 * n * 6 * sizeof(double));
	add(&t->wrec, 2 * n * sizeof(nbh_walk_rec));
	add(&t->body_n, 2 * n * sizeof(int32_t));
	add(&t->counters, 5 * sizeof(unsigned long long));
}

static size_t tree_base_bytes(size_t n)
{
	nbh_tree probe;
	probe.n = n;
	void** ptrs[kTreeBuffers];
	size_t sizes[kTreeBuffers];
	tree_buffers(&probe, ptrs, sizes);
	size_t total = 0;
	for(size_t b : sizes) { total += b; }
	return total;
}

nbh_status nbh_tree_create(nbh_ctx* ctx, size_t n, nbh_tree** out)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(out != nullptr, "null output pointer");
	*out = nullptr;
	NBH_CHECK(n >= 1, "empty tree");
	NBH_CHECK(n <= (static_cast<size_t>(1) << 28), "body count %zu too large", n);
	nbh_tree* t = new nbh_tree();
	t->ctx = ctx;
	t->n = n;
	while((static_cast<size_t>(1) << t->levels) < n) { ++t->levels; }
	void** ptrs[kTreeBuffers];
	size_t sizes[kTreeBuffers];
	tree_buffers(t, ptrs, sizes);
	hipError_t e = hipSuccess;
	for(int k = 0; k < kTreeBuffers && e == hipSuccess; ++k)
	{
		e = hipMalloc(ptrs[k], sizes[k]);
	}
	if(e == hipSuccess) { e = hipMemsetAsync(t->xyzr, 0, 2 * n * sizeof(double4), ctx->stream); }
	if(e == hipSuccess) { e = hipMemsetAsync(t->mass, 0, 2 * n * sizeof(double), ctx->stream); }
	if(e == hipSuccess) { e = hipMemsetAsync(t->quad, 0, 2 * n * 6 * sizeof(double), ctx->stream); }
	if(e == hipSuccess) { e = hipMemsetAsync(t->wrec, 0, 2 * n * sizeof(nbh_walk_rec), ctx->stream); }
	if(e == hipSuccess) { e = hipMemsetAsync(t->body_n, 0xff, 2 * n * sizeof(int32_t), ctx->stream); }
	if(e == hipSuccess) { e = hipMemsetAsync(t->counters, 0, 5 * sizeof(unsigned long long), ctx->stream); }
	if(e == hipSuccess) { e = hipStreamSynchronize(ctx->stream); }
	if(e != hipSuccess)
	{
		nbh_tree_destroy(t);
		return nbh::hip_fail(e, "tree allocation", __FILE__, __LINE__);
	}
	*out = t;
	return NBH_OK;
}

// The Morton tree's buffers, all together or none.
static nbh_status radix_alloc(nbh_tree* t)
{
	if(t->child != nullptr)
	{
		return NBH_OK;
	}
	const size_t n = t->n;
	constexpr int kBufs = 14;
	void* p[kBufs] = {};
	const size_t sizes[kBufs] = { n * sizeof(int2), 2 * n * sizeof(int), n * sizeof(int), n * sizeof(int),
	                              (3 * kRadixLevels + 3) * sizeof(int), n * sizeof(int2), n * sizeof(int2),
	                              3 * n * sizeof(uint32_t), n * sizeof(int4),
	                              (n / kRadixWalk + 1) * sizeof(int2), 6 * (kBoxParts + 1) * sizeof(double),
	                              n * sizeof(double4), n * sizeof(int2), n * sizeof(int4) };
	hipError_t e = hipSuccess;
	for(int k = 0; k < kBufs && e == hipSuccess; ++k)
	{
		e = hipMalloc(&p[k], sizes[k]);
	}
	if(e != hipSuccess)
	{
		for(void* q : p)
		{
			if(q != nullptr) { (void)hipFree(q); }
		}
		return nbh::hip_fail(e, "Morton tree allocation", __FILE__, __LINE__);
	}
	t->child = static_cast<int2*>(p[0]);
	t->parent = static_cast<int*>(p[1]);
	t->level = static_cast<int*>(p[2]);
	t->level_list = static_cast<int*>(p[3]);
	t->level_info = static_cast<int*>(p[4]);
	t->gmark = static_cast<int2*>(p[5]);
	t->smark = static_cast<int2*>(p[6]);
	t->flags = static_cast<uint32_t*>(p[7]);
	t->grp = static_cast<int4*>(p[8]);
	t->sup = static_cast<int2*>(p[9]);
	t->box = static_cast<double*>(p[10]);
	t->pk = static_cast<double4*>(p[11]);
	t->wmark = static_cast<int2*>(p[12]);
	t->wgrp = static_cast<int4*>(p[13]);
	return NBH_OK;
}

// Levels with fewer internal nodes than this run together in one work-group.
constexpr int kRadixSmallLevel = 512;

// Moments and walk records of the Morton tree (its structure, groups and level
// lists already built): the leaves and the nodes inside the walk groups one wave
// per group, then the nodes above them level by level from the longest prefix
// up to the root.  build: leaves from `order` (sorted body numbers) and the
// packed bodies, else from the body numbers of the last build and the
// positions y.
static nbh_status radix_moments(nbh_ctx* ctx, nbh_tree* t, const double* y, const uint32_t* order, double ratio,
                                bool build)
{
	const size_t n = t->n;
	t->ratio = ratio;
	const double ratio_sqr = ratio * ratio;
	// the leaves and the nodes inside the walk groups, then the levels above them
	if(build)
	{
		hipLaunchKernelGGL(k_radix_group_moments<true>, dim3(static_cast<unsigned>(t->ngroups)), dim3(kRadixGroup), 0,
		                   ctx->stream, t->grp, static_cast<int>(n), t->child, order, t->pk, y, t->body_n, t->xyzr,
		                   t->mass, t->bmin, t->bmax, t->quad, ratio_sqr, t->wrec, t->radix_lean ? 1 : 0);
	}
	else
	{
		hipLaunchKernelGGL(k_radix_group_moments<false>, dim3(static_cast<unsigned>(t->ngroups)), dim3(kRadixGroup),
		                   0, ctx->stream, t->grp, static_cast<int>(n), t->child, order, t->pk, y, t->body_n,
		                   t->xyzr, t->mass, t->bmin, t->bmax, t->quad, ratio_sqr, t->wrec, t->radix_lean ? 1 : 0);
	}
	NBH_LAUNCHED();
	int offset[kRadixLevels];
	int run = 0;
	for(int lv = 0; lv < kRadixLevels; ++lv)
	{
		offset[lv] = run;
		run += t->level_count[lv];
	}
	int lv = kRadixLevels - 1;
	while(lv >= 0)
	{
		if(t->level_count[lv] == 0)
		{
			--lv;
			continue;
		}
		if(t->level_count[lv] >= kRadixSmallLevel)
		{
			hipLaunchKernelGGL(k_radix_level, dim3(nbh::grid_for(ctx, t->level_count[lv], 256)), dim3(256), 0,
			                   ctx->stream, t->level_list + offset[lv], t->level_count[lv], t->child, t->xyzr, t->mass,
			                   t->bmin, t->bmax, t->quad, ratio_sqr, t->wrec);
			NBH_LAUNCHED();
			--lv;
			continue;
		}
		int lo = lv;
		while(lo > 0 && t->level_count[lo - 1] < kRadixSmallLevel) { --lo; }
		hipLaunchKernelGGL(k_radix_levels_small, dim3(1), dim3(1024), 0, ctx->stream, t->level_list, t->level_info, lv,
		                   lo, t->child, t->xyzr, t->mass, t->bmin, t->bmax, t->quad, ratio_sqr, t->wrec);
		NBH_LAUNCHED();
		lv = lo - 1;
	}
	return NBH_OK;
}

nbh_status nbh_tree_build_radix(nbh_ctx* ctx, nbh_tree* t, const double* y, const double* mass, double ratio)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(y != nullptr && mass != nullptr, "null buffer");
	NBH_CHECK(t->n >= 2, "the Morton tree needs at least 2 bodies");
	NBH_TRY(radix_alloc(t));
	nbh::scoped_timer timer(ctx, "tree_build");
	const size_t n = t->n;
	t->built = false;
	const unsigned parts = std::min<unsigned>(kBoxParts, nbh::grid_for(ctx, n, 256));
	hipLaunchKernelGGL(k_box_partial, dim3(parts), dim3(256), 0, ctx->stream, y, n, t->box + 6);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_box_final, dim3(1), dim3(256), 0, ctx->stream, t->box + 6, static_cast<int>(parts), t->box);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_morton_codes, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, y, mass, n,
	                   t->box, t->keys, t->vals, t->pk);
	NBH_LAUNCHED();
	uint32_t* order = nullptr;
	// 63-bit keys: 8 passes, so the sorted keys end where they started
	NBH_TRY(nbh::sort_pairs(ctx, t->keys, t->vals, t->keys + n, t->vals + n, n, 63, t->counts, t->scan_part, &order));
	const unsigned gn = nbh::grid_for(ctx, 2 * n, 256);
	hipLaunchKernelGGL(k_fill_int, dim3(gn), dim3(256), 0, ctx->stream, t->parent, 2 * n, 0);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_fill_int, dim3(gn), dim3(256), 0, ctx->stream, reinterpret_cast<int*>(t->gmark), 2 * n, 0);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_fill_int, dim3(gn), dim3(256), 0, ctx->stream, reinterpret_cast<int*>(t->smark), 2 * n, 0);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_fill_int, dim3(gn), dim3(256), 0, ctx->stream, reinterpret_cast<int*>(t->wmark), 2 * n, 0);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, ctx->stream, t->level_info,
	                   static_cast<size_t>(3 * kRadixLevels + 3), 0);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_radix_nodes, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, t->keys,
	                   static_cast<long long>(n), t->child, t->parent, t->level, t->level_info, t->gmark, t->smark,
	                   t->wmark);
	NBH_LAUNCHED();
	// groups and super-groups numbered in Morton order
	hipLaunchKernelGGL(k_radix_flags, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, t->gmark,
	                   t->smark, t->wmark, n, t->flags, t->flags + n, t->flags + 2 * n);
	NBH_LAUNCHED();
	NBH_TRY(scan_exclusive(ctx, t->flags, n, t->scan_part));
	NBH_TRY(scan_exclusive(ctx, t->flags + n, n, t->scan_part));
	NBH_TRY(scan_exclusive(ctx, t->flags + 2 * n, n, t->scan_part));
	int* info = t->level_info + 3 * kRadixLevels;
	hipLaunchKernelGGL(k_radix_compact, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, t->gmark,
	                   t->smark, t->wmark, t->flags, t->flags + n, t->flags + 2 * n, n, t->grp, t->sup, t->wgrp, info);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_radix_group_super, dim3(nbh::grid_for(ctx, n / 16 + 1, 256)), dim3(256), 0, ctx->stream,
	                   t->wgrp, t->sup, info, t->child, n);
	NBH_LAUNCHED();
	// internal ids listed by level
	hipLaunchKernelGGL(k_radix_level_offsets, dim3(1), dim3(kRadixLevels), 0, ctx->stream, t->level_info,
	                   t->level_info + kRadixLevels, t->level_info + 2 * kRadixLevels);
	NBH_LAUNCHED();
	hipLaunchKernelGGL(k_radix_level_lists, dim3(static_cast<unsigned>((n - 1 + kRadixListTile - 1) / kRadixListTile)),
	                   dim3(256), 0, ctx->stream, t->level, n, t->level_info + kRadixLevels,
	                   t->level_info + 2 * kRadixLevels, t->level_list);
	NBH_LAUNCHED();
	// the level counts and the group counts drive the launches that follow
	constexpr size_t kInfoBytes = (3 * kRadixLevels + 3) * sizeof(int);
	NBH_TRY(nbh::pinned(ctx, kInfoBytes));
	const int* host_info = reinterpret_cast<const int*>(ctx->pinned);
	NBH_HIP(hipMemcpyAsync(ctx->pinned, t->level_info, kInfoBytes, hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	for(int lv = 0; lv < kRadixLevels; ++lv) { t->level_count[lv] = host_info[lv]; }
	t->ngroups = host_info[3 * kRadixLevels];
	t->nsupers = host_info[3 * kRadixLevels + 1];
	t->nwalk = host_info[3 * kRadixLevels + 2];
	NBH_CHECK(t->ngroups >= 1 && static_cast<size_t>(t->ngroups) <= n, "bad group count %d", t->ngroups);
	NBH_CHECK(t->nwalk >= 1 && t->nwalk <= t->ngroups, "bad walk group count %d", t->nwalk);
	NBH_CHECK(t->nsupers >= 0 && static_cast<size_t>(t->nsupers) <= n / kRadixWalk + 1,
	          "bad super-group count %d", t->nsupers);
	NBH_TRY(radix_moments(ctx, t, y, order, ratio, true));
	t->radix = true;
	t->built = true;
	return NBH_OK;
}

nbh_status nbh_tree_get_children(nbh_ctx* ctx, const nbh_tree* t, int32_t* child, int32_t* parent)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(t->built && t->radix, "no Morton tree has been built");
	if(child != nullptr)
	{
		NBH_HIP(hipMemcpyAsync(child, t->child, t->n * sizeof(int2), hipMemcpyDeviceToHost, ctx->stream));
	}
	if(parent != nullptr)
	{
		NBH_HIP(hipMemcpyAsync(parent, t->parent, 2 * t->n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	}
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

nbh_status nbh_tree_destroy(nbh_tree* t)
{
	if(t == nullptr)
	{
		return NBH_OK;
	}
	if(t->ctx != nullptr)
	{
		(void)hipSetDevice(t->ctx->device);
		(void)hipStreamSynchronize(t->ctx->stream);
	}
	void* ptrs[] = { t->list[0], t->list[1], t->list[2], t->tmp, t->side, t->keys, t->vals, t->counts, t->scan_part,
	                 t->xyzr, t->mass, t->bmin, t->bmax, t->quad, t->wrec, t->body_n, t->counters,
	                 t->two_lists, t->two_count, t->two_coef, t->far_list, t->far_count,
	                 t->child, t->parent, t->level, t->level_list, t->level_info, t->gmark, t->smark, t->flags,
	                 t->grp, t->sup, t->box, t->pk, t->wmark, t->wgrp };
	for(void* p : ptrs)
	{
		if(p != nullptr) { (void)hipFree(p); }
	}
	delete t;
	return NBH_OK;
}

nbh_status nbh_tree_build(nbh_ctx* ctx, nbh_tree* t, const double* y, const double* mass, double ratio)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(y != nullptr && mass != nullptr, "null buffer");
	NBH_CHECK((t->n & (t->n - 1)) == 0, "body count %zu is not a power of two (the heap tree needs 2^k leaves)", t->n);
	nbh::scoped_timer timer(ctx, "tree_build");
	const size_t n = t->n;
	t->radix = false;
	for(int a = 0; a < 3; ++a)
	{
		NBH_TRY(radix_sort(ctx, t, y + static_cast<size_t>(a) * n, t->list[a]));
	}
	const size_t ntiles = (n + kPartTile - 1) / kPartTile;
	// the last levels (segments of at most kLocalSeg bodies) run in LDS
	const int local_levels = std::min(t->levels, 12);
	static_assert(kLocalSeg == 1 << 12, "kLocalSeg is 2^12");
	const int global_levels = t->levels - local_levels;
	for(int level = 0; level < global_levels; ++level)
	{
		const size_t seg = n >> level;
		const int a = level % 3;
		hipLaunchKernelGGL(k_mark_side, dim3(nbh::grid_for(ctx, n, 256)), dim3(256), 0, ctx->stream, t->list[a],
		                   t->side, n, seg);
		NBH_LAUNCHED();
		for(int b = 0; b < 3; ++b)
		{
			if(b == a) { continue; }
			if(seg >= static_cast<size_t>(kPartTile))
			{
				hipLaunchKernelGGL(k_part_count, dim3(static_cast<unsigned>(ntiles)), dim3(kPartThreads), 0,
				                   ctx->stream, t->list[b], t->side, n, t->counts);
				NBH_LAUNCHED();
				NBH_TRY(scan_exclusive(ctx, t->counts, ntiles, t->scan_part));
			}
			hipLaunchKernelGGL(k_part_scatter, dim3(static_cast<unsigned>(ntiles)), dim3(kPartThreads), 0,
			                   ctx->stream, t->list[b], t->side, n, seg, t->counts, t->tmp);
			NBH_LAUNCHED();
			std::swap(t->list[b], t->tmp);
		}
	}
	if(local_levels > 0)
	{
		const size_t seg0 = n >> global_levels;
		hipLaunchKernelGGL(k_split_local, dim3(static_cast<unsigned>(n / seg0)), dim3(kLocalThreads), 0, ctx->stream,
		                   t->list[0], t->list[1], t->list[2], t->tmp, static_cast<int>(seg0), global_levels,
		                   t->levels);
		NBH_LAUNCHED();
	}
	NBH_TRY(moments(ctx, t, y, mass, ratio, true));
	t->built = true;
	return NBH_OK;
}

nbh_status nbh_tree_set_radix_lean(nbh_ctx* ctx, nbh_tree* t, int lean)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	t->radix_lean = lean != 0;
	return NBH_OK;
}

nbh_status nbh_tree_refit(nbh_ctx* ctx, nbh_tree* t, const double* y, double ratio)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(y != nullptr, "null buffer");
	nbh::scoped_timer timer(ctx, "tree_refit");
	if(t->radix)
	{
		return radix_moments(ctx, t, y, nullptr, ratio, false);
	}
	return moments(ctx, t, y, nullptr, ratio, false);
}

nbh_status nbh_bh_fcompute(nbh_ctx* ctx, const nbh_tree* t, const double* y, const double* mass, double* f,
                           nbh_tree_layout layout, nbh_precision precision)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(!t->radix, "this walk runs on the heap tree only (the last build made the Morton tree)");
	NBH_CHECK(y != nullptr && mass != nullptr && f != nullptr && y != f, "bad buffers");
	NBH_CHECK(layout == NBH_LAYOUT_HEAP || layout == NBH_LAYOUT_HEAP_STACKLESS, "unknown layout %d",
	          static_cast<int>(layout));
	NBH_CHECK(precision == NBH_FP64 || precision == NBH_FP32, "unknown precision %d", static_cast<int>(precision));
	const int n = static_cast<int>(t->n);
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, t->counters);
	NBH_LAUNCHED();
	nbh::scoped_timer timer(ctx, "bh_traverse");
	dim3 grid((n + 255) / 256);
	if(precision == NBH_FP32)
	{
		hipLaunchKernelGGL(k_bh_stackless32, grid, dim3(256), 0, ctx->stream, t->xyzr, t->mass, t->body_n, y, mass,
		                   f, n, t->counters);
	}
	else if(layout == NBH_LAYOUT_HEAP)
	{
		hipLaunchKernelGGL(k_bh_stack64, grid, dim3(256), 0, ctx->stream, t->xyzr, t->mass, t->body_n, y, mass, f,
		                   n, t->counters);
	}
	else
	{
		hipLaunchKernelGGL(k_bh_stackless64, grid, dim3(256), 0, ctx->stream, t->xyzr, t->mass, t->body_n, y, mass,
		                   f, n, t->counters);
	}
	NBH_LAUNCHED();
	return NBH_OK;
}

enum GroupMode { kGroupMonopole, kGroupQuad, kGroupLocal };

static WalkTree walk_tree(const nbh_tree* t)
{
	WalkTree wt;
	if(t->radix)
	{
		wt.child = t->child;
		wt.grp = t->wgrp;
		wt.sup = t->sup;
		wt.ngroups = t->nwalk;
		wt.nsupers = t->nsupers;
	}
	return wt;
}

static nbh_status group_walk(nbh_ctx* ctx, const nbh_tree* t, const double* y, double* f, GroupMode mode,
                             float near_k)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(y != nullptr && f != nullptr && y != f, "bad buffers");
	NBH_CHECK(t->radix || t->n >= static_cast<size_t>(kGroup), "group walk needs at least %d bodies", kGroup);
	NBH_CHECK(t->radix || (t->levels + 2) * (kGroup / 2) <= kWalkStack, "tree too deep for the walk stack");
	const int n = static_cast<int>(t->n);
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, t->counters);
	NBH_LAUNCHED();
	nbh::scoped_timer timer(ctx, "bh_traverse");
	if(t->radix)
	{
		// the Morton tree's depth is not bounded by log2(n): a step whose children
		// would not fit on the stack takes the nodes as monopoles (counters[2])
		const WalkTree wt = walk_tree(t);
		if(mode == kGroupQuad)
		{
			const unsigned grid = static_cast<unsigned>(t->nwalk);
			hipLaunchKernelGGL((k_bh_group32<true, false, false, true>), dim3(grid), dim3(kGroup), 0,
			                   ctx->stream, t->wrec, t->bmin, t->bmax, t->body_n, y, f, n, near_k, t->counters,
			                   TwoLevel(), wt);
		}
		else if(mode == kGroupLocal)
		{
			const unsigned grid = static_cast<unsigned>(t->nwalk);
			hipLaunchKernelGGL((k_bh_group32<true, true, false, true>), dim3(grid), dim3(kGroup), 0,
			                   ctx->stream, t->wrec, t->bmin, t->bmax, t->body_n, y, f, n, near_k, t->counters,
			                   TwoLevel(), wt);
		}
		else
		{
			const unsigned grid = static_cast<unsigned>(t->nwalk);
			hipLaunchKernelGGL((k_bh_group32<false, false, false, true>), dim3(grid), dim3(kGroup), 0,
			                   ctx->stream, t->wrec, t->bmin, t->bmax, t->body_n, y, f, n, 0.0f, t->counters,
			                   TwoLevel(), wt);
		}
		NBH_LAUNCHED();
		return NBH_OK;
	}
	if(mode == kGroupQuad)
	{
		const unsigned grid = static_cast<unsigned>(n / kGroup);
		hipLaunchKernelGGL((k_bh_group32<true>), dim3(grid), dim3(kGroup), 0, ctx->stream, t->wrec, t->bmin,
		                   t->bmax, t->body_n, y, f, n, near_k, t->counters);
	}
	else if(mode == kGroupLocal)
	{
		const unsigned grid = static_cast<unsigned>(n / kGroup);
		hipLaunchKernelGGL((k_bh_group32<true, true>), dim3(grid), dim3(kGroup), 0, ctx->stream, t->wrec,
		                   t->bmin, t->bmax, t->body_n, y, f, n, near_k, t->counters);
	}
	else
	{
		const unsigned grid = static_cast<unsigned>(n / kGroup);
		hipLaunchKernelGGL((k_bh_group32<false>), dim3(grid), dim3(kGroup), 0, ctx->stream, t->wrec, t->bmin,
		                   t->bmax, t->body_n, y, f, n, 0.0f, t->counters);
	}
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status nbh_bh_fcompute_group(nbh_ctx* ctx, const nbh_tree* t, const double* y, double* f)
{
	return group_walk(ctx, t, y, f, kGroupMonopole, 0.0f);
}

nbh_status nbh_bh_fcompute_group_quad(nbh_ctx* ctx, const nbh_tree* t, const double* y, double* f, float near_k)
{
	NBH_CHECK(near_k >= 0.0f, "near_k must be >= 0");
	return group_walk(ctx, t, y, f, kGroupQuad, near_k);
}

nbh_status nbh_bh_fcompute_group_local(nbh_ctx* ctx, const nbh_tree* t, const double* y, double* f, float near_k)
{
	NBH_CHECK(near_k >= 1.0f, "near_k must be >= 1 (the expansion needs the cell outside the group)");
	return group_walk(ctx, t, y, f, kGroupLocal, near_k);
}

// List capacity per super-group of the two-level walk (4 bytes per entry).
constexpr int kTwoCap = 32768;

// Bytes of the two-level walk's buffers: per super-group a list of `cap` node
// indices, its length and the 34 coefficients of its expansion.
static size_t two_bytes(size_t supers, int cap)
{
	return supers * (static_cast<size_t>(cap) * sizeof(int) + sizeof(int) + 34 * sizeof(float));
}

// Super-groups of the tree's last build: 1024-body heap nodes, or the Morton
// tree's maximal subtrees of 129 to 2048 bodies.
static size_t super_count(const nbh_tree* t)
{
	return t->radix ? static_cast<size_t>(t->nsupers) : t->n >> (6 + kSuperShift);
}

// Frees the two-level walk's buffers.  Every pointer is null afterwards, also
// when a hipFree fails (the first failure is reported).
static nbh_status two_free(nbh_ctx* ctx, nbh_tree* t)
{
	if(t->two_lists == nullptr && t->two_count == nullptr && t->two_coef == nullptr)
	{
		return NBH_OK;
	}
	hipError_t e = hipStreamSynchronize(ctx->stream);
	void* ptrs[3] = { t->two_lists, t->two_count, t->two_coef };
	t->two_lists = nullptr;
	t->two_count = nullptr;
	t->two_coef = nullptr;
	t->two_slots = 0;
	for(void* p : ptrs)
	{
		if(p != nullptr)
		{
			hipError_t ef = hipFree(p);
			if(e == hipSuccess) { e = ef; }
		}
	}
	if(e != hipSuccess)
	{
		return nbh::hip_fail(e, "freeing the two-level walk lists", __FILE__, __LINE__);
	}
	return NBH_OK;
}

// Allocates the two-level walk's buffers all together or not at all: on any
// failure the ones already allocated are freed and every pointer stays null,
// so a later call allocates afresh instead of launching with a missing buffer.
// NBH_TEST_FAIL_ALLOC=k (k = 1, 2, 3) in the environment makes the k-th of the
// three allocations fail as out of memory (a test hook for this path).
static nbh_status two_alloc(nbh_ctx* ctx, nbh_tree* t, int cap, size_t supers)
{
	NBH_TRY(two_free(ctx, t));
	const size_t sizes[3] = { supers * static_cast<size_t>(cap) * sizeof(int), supers * sizeof(int),
	                          supers * 34 * sizeof(float) };
	int fail_at = 0;
	if(const char* env = getenv("NBH_TEST_FAIL_ALLOC"))
	{
		fail_at = atoi(env);
	}
	void* p[3] = { nullptr, nullptr, nullptr };
	hipError_t e = hipSuccess;
	for(int k = 0; k < 3 && e == hipSuccess; ++k)
	{
		e = (k + 1 == fail_at) ? hipErrorOutOfMemory : hipMalloc(&p[k], sizes[k]);
	}
	if(e != hipSuccess)
	{
		for(void* q : p)
		{
			if(q != nullptr) { (void)hipFree(q); }
		}
		return nbh::hip_fail(e, "two-level walk list allocation", __FILE__, __LINE__);
	}
	t->two_lists = static_cast<int*>(p[0]);
	t->two_count = static_cast<int*>(p[1]);
	t->two_coef = static_cast<float*>(p[2]);
	t->two_slots = static_cast<int>(supers);
	return NBH_OK;
}

// Far cells listed per super-group for the learned correction (4 bytes each);
// cells beyond the cap are left to the plain expansion.
constexpr int kFarCap = 8192;

// The super-groups' far-cell lists, all or nothing (as two_alloc).
static nbh_status far_alloc(nbh_ctx* ctx, nbh_tree* t)
{
	if(t->far_list != nullptr && t->far_count != nullptr)
	{
		return NBH_OK;
	}
	const size_t supers = t->n >> (6 + kSuperShift);
	void* a = nullptr;
	void* b = nullptr;
	hipError_t e = hipMalloc(&a, supers * static_cast<size_t>(kFarCap) * sizeof(int));
	if(e == hipSuccess) { e = hipMalloc(&b, supers * sizeof(int)); }
	if(e != hipSuccess)
	{
		if(a != nullptr) { (void)hipFree(a); }
		if(b != nullptr) { (void)hipFree(b); }
		return nbh::hip_fail(e, "far-cell list allocation", __FILE__, __LINE__);
	}
	(void)ctx;
	if(t->far_list != nullptr) { (void)hipFree(t->far_list); }
	if(t->far_count != nullptr) { (void)hipFree(t->far_count); }
	t->far_list = static_cast<int*>(a);
	t->far_count = static_cast<int*>(b);
	return NBH_OK;
}

// Super-group walk (k_bh_super) with the lists of t; with far lists when
// want_far.  Allocates the lists when they are missing.
static nbh_status super_walk(nbh_ctx* ctx, nbh_tree* t, float near_k, bool want_far, TwoLevel* two)
{
	const size_t supers = super_count(t);
	const int cap = t->two_cap > 0 ? t->two_cap : kTwoCap;
	if(t->two_lists == nullptr || t->two_count == nullptr || t->two_coef == nullptr ||
	   static_cast<size_t>(t->two_slots) < supers)
	{
		NBH_TRY(two_alloc(ctx, t, cap, supers));
	}
	if(want_far)
	{
		NBH_TRY(far_alloc(ctx, t));
	}
	two->lists = t->two_lists;
	two->count = t->two_count;
	two->coef = t->two_coef;
	two->cap = cap;
	two->far_list = want_far ? t->far_list : nullptr;
	two->far_count = want_far ? t->far_count : nullptr;
	two->far_cap = want_far ? kFarCap : 0;
	const int n = static_cast<int>(t->n);
	const float super_near = fmaxf(near_k, kSuperNearMin);
	// NBH_TEST_SUPER_STACK=k (64 <= k <= the stack size) in the environment lets
	// the super-group walk use only k stack entries, so that its overflow path
	// (the groups of a super-group that ran out walk from the root) can be tested
	int stack_limit = kWalkStack;
	if(const char* env = getenv("NBH_TEST_SUPER_STACK"))
	{
		const int k = atoi(env);
		if(k >= kGroup && k <= kWalkStack) { stack_limit = k; }
	}
	if(t->radix)
	{
		hipLaunchKernelGGL(k_bh_super<true>, dim3(static_cast<unsigned>(supers)), dim3(kGroup), 0, ctx->stream,
		                   t->wrec, t->bmin, t->bmax, n, super_near, 1.0 / t->ratio, *two, t->counters, stack_limit,
		                   walk_tree(t));
	}
	else
	{
		hipLaunchKernelGGL(k_bh_super<false>, dim3(static_cast<unsigned>(supers)), dim3(kGroup), 0, ctx->stream,
		                   t->wrec, t->bmin, t->bmax, n, super_near, 1.0 / t->ratio, *two, t->counters, stack_limit,
		                   WalkTree());
	}
	NBH_LAUNCHED();
	return NBH_OK;
}

// Two-level local walk, with the network's correction of the super-groups'
// far fields when net is not null.
static nbh_status two_level(nbh_ctx* ctx, nbh_tree* t, const nbh_mlp* net, const double* y, double* f, float near_k)
{
	NBH_CHECK(near_k >= 1.0f, "near_k must be >= 1 (the expansion needs the cell outside the group)");
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(net == nullptr || nbh::mlp_ctx(net) == ctx, "network does not belong to this context");
	const size_t supers = super_count(t);
	if(supers < 1)
	{
		NBH_CHECK(net == nullptr, "the learned two-level walk needs at least %d bodies", kGroup << kSuperShift);
		return group_walk(ctx, t, y, f, kGroupLocal, near_k);
	}
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(y != nullptr && f != nullptr && y != f, "bad buffers");
	NBH_CHECK(t->radix || (t->levels + 2) * (kGroup / 2) <= kWalkStack, "tree too deep for the walk stack");
	NBH_CHECK(net == nullptr || !t->radix, "the learned two-level walk runs on the heap tree only");
	const int n = static_cast<int>(t->n);
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, t->counters);
	NBH_LAUNCHED();
	nbh::scoped_timer timer(ctx, "bh_traverse");
	TwoLevel two;
	NBH_TRY(super_walk(ctx, t, near_k, net != nullptr, &two));
	if(net != nullptr)
	{
		hipLaunchKernelGGL(k_super_net, dim3(static_cast<unsigned>(supers)), dim3(64), 0, ctx->stream, t->xyzr,
		                   t->mass, t->bmin, t->bmax, t->quad, n, two, nbh::mlp_frags(net), nbh::mlp_params(net));
		NBH_LAUNCHED();
	}
	{
		// the group walks alone, inside bh_traverse
		nbh::scoped_timer group_timer(ctx, "bh_group_walk");
		if(t->radix)
		{
			const unsigned grid = static_cast<unsigned>(t->nwalk);
			hipLaunchKernelGGL((k_bh_group32<true, true, true, true>), dim3(grid), dim3(kGroup), 0,
			                   ctx->stream, t->wrec, t->bmin, t->bmax, t->body_n, y, f, n, near_k, t->counters, two,
			                   walk_tree(t));
		}
		else
		{
			const unsigned grid = static_cast<unsigned>(n / kGroup);
			hipLaunchKernelGGL((k_bh_group32<true, true, true>), dim3(grid), dim3(kGroup), 0, ctx->stream,
			                   t->wrec, t->bmin, t->bmax, t->body_n, y, f, n, near_k, t->counters, two);
		}
		NBH_LAUNCHED();
	}
	return NBH_OK;
}

nbh_status nbh_bh_fcompute_group_local2(nbh_ctx* ctx, nbh_tree* t, const double* y, double* f, float near_k)
{
	return two_level(ctx, t, nullptr, y, f, near_k);
}

nbh_status nbh_bh_fcompute_group_local2_mlp(nbh_ctx* ctx, nbh_tree* t, const nbh_mlp* net, const double* y,
                                            double* f, float near_k)
{
	NBH_CHECK(net != nullptr, "null network");
	return two_level(ctx, t, net, y, f, near_k);
}

nbh_status nbh_bh_super_samples(nbh_ctx* ctx, nbh_tree* t, float near_k, const uint32_t* supers, size_t n_supers,
                                float* x, float* basis, float* target, uint32_t* cells, size_t capacity,
                                size_t* count)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(!t->radix, "this walk runs on the heap tree only (the last build made the Morton tree)");
	NBH_CHECK(near_k >= 1.0f, "near_k must be >= 1");
	NBH_CHECK(supers != nullptr && count != nullptr, "null argument");
	const size_t nsup = t->n >> (6 + kSuperShift);
	NBH_CHECK(nsup >= 1, "no super-group: at least %d bodies are needed", kGroup << kSuperShift);
	NBH_CHECK(n_supers > 0 && n_supers <= nsup, "bad super-group count %zu", n_supers);
	NBH_CHECK(capacity == 0 || (x != nullptr && basis != nullptr && target != nullptr), "null sample buffer");
	NBH_CHECK(capacity <= (static_cast<size_t>(1) << 31), "capacity %zu too large", capacity);
	NBH_CHECK((t->levels + 2) * (kGroup / 2) <= kWalkStack, "tree too deep for the walk stack");
	for(size_t i = 0; i < n_supers; ++i)
	{
		NBH_CHECK(supers[i] < nsup, "super-group %u out of range (%zu)", supers[i], nsup);
	}
	*count = 0;
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, t->counters);
	NBH_LAUNCHED();
	TwoLevel two;
	NBH_TRY(super_walk(ctx, t, near_k, true, &two));
	std::vector<int> far_count(nsup);
	NBH_HIP(hipMemcpyAsync(far_count.data(), t->far_count, nsup * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	std::vector<unsigned long long> off(n_supers);
	unsigned long long total = 0;
	for(size_t i = 0; i < n_supers; ++i)
	{
		off[i] = total;
		total += static_cast<unsigned long long>(std::min(far_count[supers[i]], kFarCap));
	}
	*count = static_cast<size_t>(total);
	const size_t m = std::min(static_cast<size_t>(total), capacity);
	if(m == 0)
	{
		return NBH_OK;
	}
	// scratch: selected super-groups, their offsets, then the pairs
	const size_t off_off = (n_supers * sizeof(uint32_t) + 255) / 256 * 256;
	const size_t off_pairs = off_off + (n_supers * sizeof(unsigned long long) + 255) / 256 * 256;
	void* sp = nullptr;
	NBH_TRY(nbh::scratch(ctx, off_pairs + m * 2 * sizeof(uint32_t), &sp));
	char* base = static_cast<char*>(sp);
	uint32_t* dsel = reinterpret_cast<uint32_t*>(base);
	unsigned long long* doff = reinterpret_cast<unsigned long long*>(base + off_off);
	uint32_t* pairs = reinterpret_cast<uint32_t*>(base + off_pairs);
	NBH_HIP(hipMemcpyAsync(dsel, supers, n_supers * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
	NBH_HIP(hipMemcpyAsync(doff, off.data(), n_supers * sizeof(unsigned long long), hipMemcpyHostToDevice,
	                       ctx->stream));
	hipLaunchKernelGGL(k_super_pairs, dim3(static_cast<unsigned>(n_supers)), dim3(64), 0, ctx->stream, two, dsel, doff,
	                   pairs, static_cast<unsigned long long>(m));
	NBH_LAUNCHED();
	const int n = static_cast<int>(t->n);
	hipLaunchKernelGGL(k_band_residual, dim3(nbh::grid_for(ctx, m, 1)), dim3(64), 0, ctx->stream, t->xyzr, t->mass,
	                   t->bmin, t->bmax, t->quad, n, t->levels, n >> (6 + kSuperShift), 2, pairs, m, x, basis,
	                   target);
	NBH_LAUNCHED();
	if(cells != nullptr)
	{
		NBH_HIP(hipMemcpyAsync(cells, pairs, m * 2 * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
	}
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	return NBH_OK;
}

nbh_status nbh_tree_set_list_cap(nbh_ctx* ctx, nbh_tree* t, int cap)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(cap >= 64 && cap <= (1 << 24), "list capacity %d: must be in [64, 2^24]", cap);
	NBH_TRY(two_free(ctx, t));
	t->two_cap = cap;
	const size_t supers = t->built && t->radix ? super_count(t) : t->n >> (6 + kSuperShift);
	if(supers < 1)
	{
		return NBH_OK;   // no super-group: the walk runs in one level and needs no lists
	}
	return two_alloc(ctx, t, cap, supers);
}

nbh_status nbh_tree_bytes(nbh_ctx* ctx, const nbh_tree* t, size_t* bytes)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(bytes != nullptr, "null output");
	*bytes = tree_base_bytes(t->n) +
	         (t->two_lists != nullptr ? two_bytes(t->two_slots, t->two_cap > 0 ? t->two_cap : kTwoCap) : 0) +
	         (t->far_list != nullptr ? (t->n >> (6 + kSuperShift)) * (static_cast<size_t>(kFarCap) + 1) * sizeof(int) : 0);
	return NBH_OK;
}

nbh_status nbh_bh_band_samples(nbh_ctx* ctx, const nbh_tree* t, const uint32_t* groups, size_t n_groups,
                               float rho_min, float near_k, float* x, float* basis, float* target,
                               uint32_t* cells, size_t capacity, size_t* count)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(!t->radix, "this walk runs on the heap tree only (the last build made the Morton tree)");
	NBH_CHECK(t->n >= static_cast<size_t>(kGroup), "group walk needs at least %d bodies", kGroup);
	NBH_CHECK((t->levels + 2) * (kGroup / 2) <= kWalkStack, "tree too deep for the walk stack");
	NBH_CHECK(groups != nullptr && count != nullptr, "null argument");
	NBH_CHECK(n_groups > 0 && n_groups <= (static_cast<size_t>(1) << 30), "bad group count %zu", n_groups);
	NBH_CHECK(capacity == 0 || (x != nullptr && basis != nullptr && target != nullptr), "null sample buffer");
	NBH_CHECK(capacity <= (static_cast<size_t>(1) << 31), "capacity %zu too large", capacity);
	NBH_CHECK(rho_min >= 0.0f && near_k >= 0.0f, "rho_min and near_k must be >= 0");
	const size_t ngroups_tree = t->n / kGroup;
	for(size_t i = 0; i < n_groups; ++i)
	{
		NBH_CHECK(groups[i] < ngroups_tree, "group %u out of range (%zu groups)", groups[i], ngroups_tree);
	}
	*count = 0;
	// scratch: [0, 256) counter, then the group list, then the (group, cell) pairs
	const size_t off_groups = 256;
	const size_t off_pairs = (off_groups + n_groups * sizeof(uint32_t) + 255) / 256 * 256;
	void* sp = nullptr;
	NBH_TRY(nbh::scratch(ctx, off_pairs + capacity * 2 * sizeof(uint32_t), &sp));
	char* base = static_cast<char*>(sp);
	unsigned long long* counter = reinterpret_cast<unsigned long long*>(base);
	uint32_t* dgroups = reinterpret_cast<uint32_t*>(base + off_groups);
	uint32_t* pairs = reinterpret_cast<uint32_t*>(base + off_pairs);
	NBH_HIP(hipMemcpyAsync(dgroups, groups, n_groups * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, counter);
	NBH_LAUNCHED();
	const int n = static_cast<int>(t->n);
	hipLaunchKernelGGL(k_band_collect, dim3(static_cast<unsigned>(n_groups)), dim3(kGroup), 0, ctx->stream, t->xyzr,
	                   t->bmin, t->bmax, n, dgroups, rho_min, near_k, pairs, static_cast<unsigned long long>(capacity),
	                   counter);
	NBH_LAUNCHED();
	unsigned long long found = 0;
	NBH_HIP(hipMemcpyAsync(&found, counter, sizeof(found), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	*count = static_cast<size_t>(found);
	const size_t m = std::min(static_cast<size_t>(found), capacity);
	if(m > 0)
	{
		hipLaunchKernelGGL(k_band_residual, dim3(nbh::grid_for(ctx, m, 1)), dim3(64), 0, ctx->stream, t->xyzr,
		                   t->mass, t->bmin, t->bmax, t->quad, n, t->levels, n / kGroup, 1, pairs, m, x, basis,
		                   target);
		NBH_LAUNCHED();
		if(cells != nullptr)
		{
			NBH_HIP(hipMemcpyAsync(cells, pairs, m * 2 * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
		}
		NBH_HIP(hipStreamSynchronize(ctx->stream));
	}
	return NBH_OK;
}

nbh_status nbh_bh_fcompute_group_mlp(nbh_ctx* ctx, const nbh_tree* t, const nbh_mlp* net, const double* y,
                                     double* f, float rho_min, float near_k)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(net != nullptr && nbh::mlp_ctx(net) == ctx, "network does not belong to this context");
	NBH_CHECK(t->built, "tree has not been built");
	NBH_CHECK(!t->radix, "this walk runs on the heap tree only (the last build made the Morton tree)");
	NBH_CHECK(y != nullptr && f != nullptr && y != f, "bad buffers");
	NBH_CHECK(t->n >= static_cast<size_t>(kGroup), "group walk needs at least %d bodies", kGroup);
	NBH_CHECK((t->levels + 2) * (kGroup / 2) <= kWalkStack, "tree too deep for the walk stack");
	NBH_CHECK(rho_min >= 0.0f && near_k >= 0.0f, "rho_min and near_k must be >= 0");
	const int n = static_cast<int>(t->n);
	hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, ctx->stream, t->counters);
	NBH_LAUNCHED();
	nbh::scoped_timer timer(ctx, "bh_traverse");
	hipLaunchKernelGGL(k_bh_group32_mlp, dim3(n / kGroup), dim3(kGroup), 0, ctx->stream, t->wrec, t->bmin, t->bmax,
	                   t->body_n, y, f, n, nbh::mlp_frags(net), nbh::mlp_params(net), rho_min, near_k, t->counters);
	NBH_LAUNCHED();
	return NBH_OK;
}

nbh_status nbh_bh_walk_counters(nbh_ctx* ctx, const nbh_tree* t, uint64_t out[5])
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	NBH_CHECK(out != nullptr, "null output");
	unsigned long long c[5] = {0, 0, 0, 0, 0};
	NBH_HIP(hipMemcpyAsync(c, t->counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	for(int k = 0; k < 5; ++k) { out[k] = c[k]; }
	return NBH_OK;
}

nbh_status nbh_tree_get(nbh_ctx* ctx, const nbh_tree* t, int32_t* body_n, double* cm, double* mass,
                        double* crit_r2, double* bmin, double* bmax)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	const size_t s = 2 * t->n;
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	if(body_n != nullptr) { NBH_HIP(hipMemcpy(body_n, t->body_n, s * sizeof(int32_t), hipMemcpyDeviceToHost)); }
	if(mass != nullptr) { NBH_HIP(hipMemcpy(mass, t->mass, s * sizeof(double), hipMemcpyDeviceToHost)); }
	if(bmin != nullptr) { NBH_HIP(hipMemcpy(bmin, t->bmin, 3 * s * sizeof(double), hipMemcpyDeviceToHost)); }
	if(bmax != nullptr) { NBH_HIP(hipMemcpy(bmax, t->bmax, 3 * s * sizeof(double), hipMemcpyDeviceToHost)); }
	if(cm != nullptr || crit_r2 != nullptr)
	{
		std::vector<double> rec(4 * s);
		NBH_HIP(hipMemcpy(rec.data(), t->xyzr, s * sizeof(double4), hipMemcpyDeviceToHost));
		for(size_t i = 0; i < s; ++i)
		{
			if(cm != nullptr)
			{
				cm[3 * i] = rec[4 * i];
				cm[3 * i + 1] = rec[4 * i + 1];
				cm[3 * i + 2] = rec[4 * i + 2];
			}
			if(crit_r2 != nullptr) { crit_r2[i] = rec[4 * i + 3]; }
		}
	}
	return NBH_OK;
}

nbh_status nbh_bh_counters(nbh_ctx* ctx, const nbh_tree* t, uint64_t* interactions, uint64_t* visits)
{
	NBH_TRY(nbh::enter(ctx));
	NBH_CHECK(t != nullptr && t->ctx == ctx, "tree does not belong to this context");
	unsigned long long c[3] = {0, 0, 0};
	NBH_HIP(hipMemcpyAsync(c, t->counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
	NBH_HIP(hipStreamSynchronize(ctx->stream));
	if(interactions != nullptr) { *interactions = c[0]; }
	if(visits != nullptr) { *visits = c[1]; }
	return NBH_OK;
}

}  // extern "C"
