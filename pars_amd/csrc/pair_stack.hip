// pair_stack.hip — R6 for balanced documents without a sort.
//
// model::pair_brackets (include/pars/model/nest.hpp:45-92) pairs brackets by a
// stable sort on key (depth - 1 for an open, depth for a close) and pairs the
// sorted neighbours (2i, 2i+1).  Along the +-1 depth walk, the brackets with
// key k are the crossings of the edge (k, k+1) in document order, alternately
// up (an open) and down (a close).  When the document is balanced (the depth
// never drops below 0 and ends at 0) every key's run starts with an up-crossing
// and has an even length, so the sorted neighbours are each open and the next
// bracket at its key: exactly the partner a stack matcher gives.  The sort is
// then unnecessary, and the stack runs in three levels:
//
//   thread   16 consecutive brackets, the stack a bit mask of unmatched opens
//            (the partner of a close is the highest set bit) and a shift
//            register of their expected closing bytes;
//   chunk    1024 threads = 16384 brackets: a thread's unmatched closes sit at
//            the levels [m, e) below its entry depth e, its unmatched opens at
//            [m, x) below its exit depth x (m = its minimum depth).  The close
//            at level k pairs with the open at level k of the last earlier
//            thread whose minimum is <= k, found by binary lifting over a
//            sparse table of the threads' minima in LDS;
//   document what no thread of a chunk matches is the chunk's residue - one
//            close per level below its entry depth, one open per level below
//            its exit depth - matched across chunks the same way, through a
//            sparse table of the chunks' minima in global memory.
//
// An unbalanced document (min depth < 0 or final depth != 0) takes the sorting
// path (pair_kernels.hip) instead: these kernels exit at once, and the host
// runs that path after reading the summary.  Outputs and flags are the radix
// path's: pair[] in the CPU convention (-1 for a pair whose bytes do not
// match, nest.hpp:80-88), pair_pos[] in the structural-index convention, and
// the summary's mismatch flag.
#include "dev_common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <climits>

namespace pars_hip {

namespace {

constexpr int kStkThreads = 1024;
constexpr int kStkItems	  = 16;
constexpr int kStkLevels  = 10;						 // 2^9 = 512: lifting and range minima over 1024 threads
constexpr int16_t kNoMin  = INT16_MAX;				 // a thread without brackets
constexpr int kResolveBlock = 256;					 // four chunks per block in the resolve pass

__device__ __forceinline__ uint32_t stack_nb(const StackArgs& a) {
	const uint32_t n = a.sum->n_brackets;
	return (uint64_t)n > a.cap_br ? (uint32_t)a.cap_br : n;
}

__device__ __forceinline__ bool stack_applies(const DevSummary* s) {
	return s->min_bal >= 0 && s->last_bal == 0;
}

__device__ __forceinline__ uint32_t nib(uint64_t w, uint32_t j) {
	return (uint32_t)(w >> (4 * j)) & 15u;
}

// The last index u < t with m[u] <= k, or -1 (st[j][i] = min m[i - 2^j + 1 .. i]).
__device__ __forceinline__ int last_at_most(const int16_t (*st)[kStkThreads], int t, int k) {
	int pos = t - 1;
#pragma unroll
	for (int j = kStkLevels - 1; j >= 0; --j)
		if (pos - (1 << j) + 1 >= 0 && (int)st[j][pos] > k)
			pos -= 1 << j;
	return pos;
}

// min m[l .. r] for 0 <= l <= r < kStkThreads: two blocks of 2^j that cover
// the range (the top level, 512, covers any range of up to 1024 twice over).
__device__ __forceinline__ int range_min(const int16_t (*st)[kStkThreads], int l, int r) {
	const int j = min(31 - __clz(r - l + 1), kStkLevels - 1);
	return min((int)st[j][l + (1 << j) - 1], (int)st[j][r]);
}

__device__ __forceinline__ uint32_t byte_of(const uint32_t (&w)[4], uint32_t q) {
	const uint32_t x = (q & 8u) ? ((q & 4u) ? w[3] : w[2]) : ((q & 4u) ? w[1] : w[0]);
	return (x >> (8 * (q & 3u))) & 0xFFu;
}

__device__ __forceinline__ void write_pair(const StackArgs& a, uint32_t i, int32_t v) {
	a.pair[i] = v;
	if (a.pair_pos)
		a.pair_pos[a.bsidx[i]] = v >= 0 ? (int32_t)a.bsidx[v] : -1;
}

__global__ __launch_bounds__(kStkThreads) void k_stack_local(StackArgs a) {
	__shared__ int16_t	s_st[kStkLevels][kStkThreads];	// thread minima relative to the chunk's entry depth
	__shared__ uint64_t s_ro[kStkThreads];	// a thread's unmatched opens: local indices, ascending level
	__shared__ uint8_t	s_flags[256], s_cof[256];
	__shared__ int32_t	s_wsum[kStkThreads / 64];
	__shared__ int32_t	s_entry;
	if (!stack_applies(a.sum))
		return;
	const int	   tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	const uint32_t nb  = stack_nb(a);
	const uint32_t nch = (nb + kStkChunk - 1) / kStkChunk;
	if (tid < 256) {
		s_flags[tid] = a.plan->br_flags[tid];
		s_cof[tid]	 = a.plan->br_close_of[tid];
	}
	bool bad = false, lost = false;	 // lost: a residue outside its chunk's range (cannot happen)
	for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
		__syncthreads();  // LDS of the previous chunk is free; tables loaded
		const uint32_t base = c * kStkChunk, i0 = base + (uint32_t)tid * kStkItems;
		const uint32_t cnt	= i0 < nb ? min(nb - i0, (uint32_t)kStkItems) : 0u;
		uint32_t	   w[4] = {0u, 0u, 0u, 0u};
		if (cnt == (uint32_t)kStkItems) {
			const uint4 v = *reinterpret_cast<const uint4*>(a.bchar + i0);
			w[0] = v.x, w[1] = v.y, w[2] = v.z, w[3] = v.w;
		} else {
			for (uint32_t q = 0; q < cnt; ++q) w[q >> 2] |= (uint32_t)a.bchar[i0 + q] << (8 * (q & 3));
		}
		// ---- the thread's stack ----
		uint32_t mask = 0, hp = 0, okm = 0, rc = 0;
		uint64_t pk = 0, rcl = 0, tlo = 0, thi = 0;	 // partners; unmatched closes; expected closes (LIFO)
#pragma unroll
		for (uint32_t q = 0; q < (uint32_t)kStkItems; ++q) {
			if (q < cnt) {
				const uint32_t ch = (w[q >> 2] >> (8 * (q & 3))) & 0xFFu;
				if (s_flags[ch] & 1u) {
					mask |= 1u << q;
					thi = (thi << 8) | (tlo >> 56);
					tlo = (tlo << 8) | s_cof[ch];
				} else if (mask) {
					const uint32_t p = 31u - (uint32_t)__clz(mask);
					mask ^= 1u << p;
					const uint32_t e = (uint32_t)(tlo & 0xFFu);
					tlo				 = (tlo >> 8) | (thi << 56);
					thi >>= 8;
					pk |= ((uint64_t)q << (4 * p)) | ((uint64_t)p << (4 * q));
					hp |= (1u << p) | (1u << q);
					if (e == ch)
						okm |= (1u << p) | (1u << q);
				} else {
					rcl |= (uint64_t)q << (4 * rc);
					++rc;
				}
			}
		}
		const uint32_t ro = (uint32_t)__popc(mask);
		uint64_t	   rol = 0;	 // unmatched opens, ascending
		{
			uint32_t mm = mask;
			for (uint32_t j = 0; mm; ++j) {
				rol |= (uint64_t)(__ffs(mm) - 1) << (4 * j);
				mm &= mm - 1;
			}
		}
		// ---- depths relative to the chunk's entry: e = exclusive prefix of deltas ----
		const int32_t delta = (int32_t)ro - (int32_t)rc;
		const int32_t winc	= (int32_t)wave_incl_sum((uint32_t)delta);
		if (lane == 63)
			s_wsum[wid] = winc;
		if (tid == 0)
			s_entry = nb ? a.bkey[base] + ((s_flags[w[0] & 0xFFu] & 1u) ? 0 : 1) : 0;
		__syncthreads();
		int32_t e = winc - delta, xc = 0;
#pragma unroll
		for (int q = 0; q < kStkThreads / 64; ++q) {
			if (q < wid)
				e += s_wsum[q];
			xc += s_wsum[q];
		}
		const int32_t m = e - (int32_t)rc;
		s_st[0][tid]	= cnt ? (int16_t)m : kNoMin;
		s_ro[tid]		= rol;
		__syncthreads();
#pragma unroll
		for (int j = 1; j < kStkLevels; ++j) {
			const int h = 1 << (j - 1);
			s_st[j][tid] = tid >= h ? min(s_st[j - 1][tid], s_st[j - 1][tid - h]) : s_st[j - 1][tid];
			__syncthreads();
		}
		const int32_t mc	= range_min(s_st, 0, kStkThreads - 1);	// <= 0: thread 0 enters at 0
		const int32_t entry = s_entry;
		// ---- pairs inside the thread ----
#pragma unroll
		for (uint32_t q = 0; q < (uint32_t)kStkItems; ++q) {
			if (q < cnt && ((hp >> q) & 1u)) {
				const bool ok = (okm >> q) & 1u;
				bad			  = bad || !ok;
				write_pair(a, i0 + q, ok ? (int32_t)(i0 + nib(pk, q)) : -1);
			}
		}
		// ---- the thread's unmatched closes: an earlier thread's open, or the chunk's residue ----
		for (uint32_t j = 0; j < rc; ++j) {
			const int32_t  k  = e - 1 - (int32_t)j;
			const uint32_t gc = i0 + nib(rcl, j);
			const int	   u  = last_at_most(s_st, tid, k);
			if (u >= 0) {
				const uint32_t go = base + (uint32_t)u * kStkItems + nib(s_ro[u], (uint32_t)(k - s_st[0][u]));
				if (go >= nb || go >= gc) {	 // cannot happen: an earlier thread's open
					lost = true;
					continue;
				}
				const bool ok = s_cof[a.bchar[go]] == byte_of(w, nib(rcl, j));
				bad				  = bad || !ok;
				write_pair(a, gc, ok ? (int32_t)go : -1);
				write_pair(a, go, ok ? (int32_t)gc : -1);
			} else {
				const uint32_t idx = (uint32_t)(-1 - k);  // closes: level entry-1, entry-2, ...
				if (k < 0 && base + idx < nb)
					a.res[base + idx] = gc;
				else
					lost = true;
			}
		}
		// ---- the thread's unmatched opens no later thread closes: the chunk's residue ----
		const int32_t suf = tid + 1 < kStkThreads ? range_min(s_st, tid + 1, kStkThreads - 1) : INT_MAX;
		for (uint32_t j = 0; j < ro; ++j) {
			const int32_t k = m + (int32_t)j;
			if (k < suf) {
				const uint32_t idx = (uint32_t)(-mc) + (uint32_t)(k - mc);	// opens: level min, min+1, ...
				if (base + idx < nb)
					a.res[base + idx] = i0 + nib(rol, j);
				else
					lost = true;
			}
		}
		if (tid == 0) {
			a.crec[3 * (uint64_t)c + 0] = entry;
			a.crec[3 * (uint64_t)c + 1] = entry + mc;
			a.crec[3 * (uint64_t)c + 2] = entry + xc;
		}
	}
	if (__ballot(bad) && lane == 0)
		atomicOr(&a.sum->mismatched, 1u);
	if (__ballot(lost) && lane == 0)
		atomicOr(&a.sum->overflow, kStackLost);
}

// The chunks' minimum depths as a sparse table: cst[j * max_chunks + c] =
// min over chunks c - 2^j + 1 .. c.  One block; levels in turn.
__global__ __launch_bounds__(1024) void k_stack_table(StackArgs a) {
	if (!stack_applies(a.sum))
		return;
	const uint32_t nch = (stack_nb(a) + kStkChunk - 1) / kStkChunk;
	for (uint32_t c = threadIdx.x; c < nch; c += blockDim.x) a.cst[c] = a.crec[3 * (uint64_t)c + 1];
	__syncthreads();
	for (uint32_t j = 1; j < a.st_levels && (1u << (j - 1)) < nch; ++j) {
		const uint32_t h	= 1u << (j - 1);
		const int32_t* prev = a.cst + (uint64_t)(j - 1) * a.max_chunks;
		int32_t*	   cur	= a.cst + (uint64_t)j * a.max_chunks;
		for (uint32_t c = threadIdx.x; c < nch; c += blockDim.x) cur[c] = c >= h ? min(prev[c], prev[c - h]) : prev[c];
		__syncthreads();
	}
}

// Every chunk's residue closes, one wave per chunk, one lane per level: the
// open at level k of the last earlier chunk whose minimum is <= k.
__global__ __launch_bounds__(kResolveBlock) void k_stack_resolve(StackArgs a) {
	if (!stack_applies(a.sum))
		return;
	const uint32_t nb	 = stack_nb(a);
	const uint32_t nch	 = (nb + kStkChunk - 1) / kStkChunk;
	const int	   lane	 = threadIdx.x & 63;
	const uint32_t waves = gridDim.x * (kResolveBlock / 64);
	uint32_t	   L	 = 1;  // table levels in use
	while (L < a.st_levels && (1u << (L - 1)) < nch) ++L;
	const DevPlan* P   = a.plan;
	bool		   bad = false, lost = false;
	for (uint32_t c = blockIdx.x * (kResolveBlock / 64) + (threadIdx.x >> 6); c < nch; c += waves) {
		const int32_t E = a.crec[3 * (uint64_t)c], M = a.crec[3 * (uint64_t)c + 1];
		for (int32_t j = lane; j < E - M; j += 64) {
			const int32_t k = E - 1 - j;
			if ((uint64_t)c * kStkChunk + (uint32_t)j >= nb) {	// cannot happen: more residues than brackets
				lost = true;
				continue;
			}
			const uint32_t gc = a.res[(uint64_t)c * kStkChunk + (uint32_t)j];
			if (gc >= nb) {
				lost = true;
				continue;
			}
			int64_t		   pos = (int64_t)c - 1;
			for (int jj = (int)L - 1; jj >= 0; --jj)
				if (pos - (1ll << jj) + 1 >= 0 && a.cst[(uint64_t)jj * a.max_chunks + (uint64_t)pos] > k)
					pos -= 1ll << jj;
			if (pos < 0) {	// cannot happen on a balanced document
				lost = true;
				write_pair(a, gc, -1);
				continue;
			}
			const int32_t Eo = a.crec[3 * (uint64_t)pos], Mo = a.crec[3 * (uint64_t)pos + 1],
						  Xo = a.crec[3 * (uint64_t)pos + 2];
			const uint64_t ro = (uint64_t)pos * kStkChunk + (uint32_t)(Eo - Mo) + (uint32_t)(k - Mo);
			if (k < Mo || k >= Xo || ro >= nb) {  // cannot happen on a balanced document
				lost = true;
				write_pair(a, gc, -1);
				continue;
			}
			const uint32_t go = a.res[ro];
			if (go >= nb) {
				lost = true;
				write_pair(a, gc, -1);
				continue;
			}
			const bool	   ok = P->br_close_of[a.bchar[go]] == a.bchar[gc];
			bad				  = bad || !ok;
			write_pair(a, gc, ok ? (int32_t)go : -1);
			write_pair(a, go, ok ? (int32_t)gc : -1);
		}
	}
	if (__ballot(bad) && lane == 0)
		atomicOr(&a.sum->mismatched, 1u);
	if (__ballot(lost) && lane == 0)
		atomicOr(&a.sum->overflow, kStackLost);
}

}  // namespace

uint32_t stack_levels(uint64_t max_chunks) {
	uint32_t L = 1;
	while ((1ull << (L - 1)) < max_chunks) ++L;
	return L;
}

hipError_t launch_stack_pairing(const StackArgs& a, hipStream_t s) {
	if (a.cap_br == 0)
		return hipSuccess;
	int dev = 0, cus = 0;
	(void)hipGetDevice(&dev);
	(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	cus						= std::max(cus, 1);
	const unsigned lgrid	= (unsigned)std::min<uint64_t>(std::max<uint64_t>(a.max_chunks, 1), 2ull * cus);
	const unsigned rgrid	= (unsigned)std::min<uint64_t>(
		   std::max<uint64_t>((a.max_chunks + kResolveBlock / 64 - 1) / (kResolveBlock / 64), 1), 8ull * cus);
	hipLaunchKernelGGL(k_stack_local, dim3(lgrid), dim3(kStkThreads), 0, s, a);
	hipLaunchKernelGGL(k_stack_table, dim3(1), dim3(1024), 0, s, a);
	hipLaunchKernelGGL(k_stack_resolve, dim3(rgrid), dim3(kResolveBlock), 0, s, a);
	return hipGetLastError();
}

}  // namespace pars_hip
