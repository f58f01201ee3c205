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
//   thread   32 consecutive brackets, the stack a bit mask of unmatched opens
//            (the partner of a close is the highest set bit);
//   chunk    256 threads = 8192 brackets, staged in LDS as match keys (an
//            open: a flag and the id of its closing byte; a close: its id): a
//            thread's unmatched closes sit at the levels [m, e) below its entry
//            depth e, one per level, its unmatched opens at [m, x) below its
//            exit depth x (m = its minimum depth).  The close at level k pairs
//            with the open at level k of the last earlier thread whose minimum
//            is <= k.  Each thread finds its previous smaller minimum once
//            (binary lifting over a sparse table of the minima in LDS); a
//            thread's closes descend in level, so they walk that chain back
//            from its neighbour, one LDS read per step.  The chunk's unmatched
//            closes (0 to 32 per thread) are listed and spread evenly over the
//            threads, each taking a run of the list.  Partners are collected
//            in LDS, checked (keys equal but for the flag) and stored
//            coalesced, four per thread and store for one document;
//   document what no thread of a chunk matches is the chunk's residue - one
//            close per level below its entry depth, one open per level below
//            its exit depth - matched across chunks by one wave per chunk that
//            walks back to the nearest chunk whose minimum is below the levels
//            still open, skipping groups of 64 and 4096 chunks by their minima.
//
// An unbalanced document (min depth < 0 or final depth != 0) takes the sorting
// path (pair_kernels.hip) instead: these kernels exit at once, and the host
// runs that path after reading the summary.  A batch whose documents are all
// balanced pairs here too (document-relative outputs, mismatches marked per
// document); otherwise these kernels exit and the batch's sorting chain, always
// launched after them, runs - without a host round trip, so a batch stays one
// captured graph.  Outputs and flags are the radix
// path's: pair[] in the CPU convention (-1 for a pair whose bytes do not
// match, nest.hpp:80-88), pair_pos[] in the structural-index convention, and
// the summary's mismatch flag.
#include "dev_common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <climits>

namespace pars_hip {

namespace {

constexpr int	  kStkThreads = 256;
constexpr int	  kStkItems	  = 32;
constexpr int	  kStkLevels  = 8;	// 2^7 = 128: lifting and range minima over 256 threads
constexpr int16_t kNoMin	  = INT16_MAX;	// a thread without brackets
constexpr uint16_t kPartLater = 0xFFFEu;	// a chunk residue: the resolve pass writes it
constexpr int	  kResolveBlock = 256;
constexpr uint32_t kListCap = kStkLevels * kStkThreads / 2;	// unmatched closes listed per chunk		// four chunks per block in the resolve pass
static_assert(kStkThreads * kStkItems == (int)kStkChunk, "chunk = threads x items");

__device__ __forceinline__ uint32_t stack_nb(const StackArgs& a) {
	const uint32_t n = a.sum->n_brackets;
	return (uint64_t)n > a.cap_br ? (uint32_t)a.cap_br : n;
}

// One document: balanced.  A batch: every document balanced - then the
// concatenated bracket stream is balanced too and returns to depth 0 at every
// document boundary, so no stack pair crosses a document and each document's
// pairs are those of its own depth sort.
__device__ __forceinline__ bool stack_applies(const StackArgs& a) {
	const DevSummary* s = a.sum;
	return s->min_bal >= 0 && (a.bdoc ? s->doc_unbal == 0 : s->last_bal == 0);
}

// Bracket i's partner v (a bracket index, or -1 for none or a mismatch) in the
// output conventions: pair[] by bracket index, pair_pos[] by structural index
// (when `pos`), both document-relative in a batch.
__device__ __forceinline__ void put_pair(const StackArgs& a, uint32_t i, int32_t v, bool pos) {
	uint32_t b0 = 0, s0 = 0;
	if (a.bdoc) {
		const uint32_t d = a.bdoc[i];
		b0				 = a.br_start[d];
		s0				 = a.st_start[d];
	}
	a.pair[i] = v >= 0 ? v - (int32_t)b0 : -1;
	if (pos && a.pair_pos)
		a.pair_pos[a.bsidx[i]] = v >= 0 ? (int32_t)(a.bsidx[v] - s0) : -1;
}

// The j-th set bit (from 0) of m, which has more than j set bits.
__device__ __forceinline__ uint32_t select_bit(uint32_t m, uint32_t j) {
	uint32_t pos = 0, c;
	c = (uint32_t)__popc(m & 0xFFFFu);
	if (j >= c) j -= c, m >>= 16, pos += 16;
	c = (uint32_t)__popc(m & 0xFFu);
	if (j >= c) j -= c, m >>= 8, pos += 8;
	c = (uint32_t)__popc(m & 0xFu);
	if (j >= c) j -= c, m >>= 4, pos += 4;
	c = (uint32_t)__popc(m & 0x3u);
	if (j >= c) j -= c, m >>= 2, pos += 2;
	return pos + ((j >= (m & 1u)) ? 1u : 0u);
}

// Group minima tagged with the launch epoch, kept by atomicMax: the newest
// epoch wins, and within it the smallest minimum (stored inverted).
__device__ __forceinline__ unsigned long long tag_min(uint32_t epoch, int32_t m) {
	return ((unsigned long long)epoch << 32) | (uint32_t)(0x7FFFFFFF - m);
}

__device__ __forceinline__ int32_t read_min(unsigned long long v, uint32_t epoch) {
	return (uint32_t)(v >> 32) == epoch ? 0x7FFFFFFF - (int32_t)(uint32_t)v : INT_MAX;
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
// the range (the top level, 128, covers any range of up to 256 twice over).
__device__ __forceinline__ int range_min(const int16_t (*st)[kStkThreads], int l, int r) {
	const int j = min(31 - __clz(r - l + 1), kStkLevels - 1);
	return min((int)st[j][l + (1 << j) - 1], (int)st[j][r]);
}

__global__ __launch_bounds__(kStkThreads) void k_stack_local(StackArgs a) {
	__shared__ __align__(16) int16_t s_st[kStkLevels][kStkThreads];	// thread minima relative to the chunk's entry depth
	__shared__ int16_t	s_e[kStkThreads];	// entry depths, for the list of unmatched closes
	__shared__ uint32_t s_ro[kStkThreads];				// a thread's unmatched opens (bit per item)
	__shared__ __align__(16) uint8_t s_ch[kStkChunk];	// the chunk's brackets as match keys
	__shared__ __align__(16) uint16_t s_part[kStkChunk];				// the chunk's partners (chunk-local) or kPartLater
	__shared__ uint8_t	s_flags[256], s_cof[256], s_kid[256], s_kt[256];
	__shared__ int32_t	s_wsum[kStkThreads / 64];
	__shared__ int32_t	s_entry;
	__shared__ uint32_t s_mp[kStkThreads];	// per thread: its minimum (int16) | (previous smaller + 1) << 16
	if (!stack_applies(a))
		return;
	const int	   tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	if (a.bdoc && blockIdx.x == 0 && tid == 0)	// the batch's sorting chain, launched next, has nothing to do
		a.sum->small_done = 1u;
	const uint32_t nb  = stack_nb(a);
	const uint32_t nch = (nb + kStkChunk - 1) / kStkChunk;
	const uint32_t epoch = a.sum->epoch;
	s_flags[tid] = a.plan->br_flags[tid];
	s_cof[tid]	 = a.plan->br_close_of[tid];
	{
		// a small id per closing byte value (its rank among them), then each
		// byte's key: an open -> 0x80 | id of its close, a close -> its id;
		// a pair matches iff the low 7 bits of its keys agree
		const bool	   isc = (a.plan->br_flags[tid] & 2u) != 0;
		const uint64_t bal = __ballot(isc);
		if (lane == 0)
			s_wsum[wid] = __popcll(bal);
		__syncthreads();
		uint32_t id = (uint32_t)__popcll(bal & lanemask_lt());
		for (int q = 0; q < wid; ++q) id += (uint32_t)s_wsum[q];
		s_kid[tid] = (uint8_t)min(id, 127u);
		__syncthreads();
		const uint32_t f = s_flags[tid];
		s_kt[tid]		 = (f & 1u) ? (uint8_t)(0x80u | s_kid[s_cof[tid]]) : s_kid[tid];
	}
	bool bad = false, lost = false;	 // lost: a residue outside its chunk's range (cannot happen)
	for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
		const uint32_t base = c * kStkChunk, lim = min(nb - base, (uint32_t)kStkChunk);
		int32_t		   key0 = 0;
		if (tid == 0)
			key0 = a.bkey[base];  // in flight while the bytes load
		__syncthreads();		  // LDS of the previous chunk is free; tables loaded
		// ---- the chunk's bytes into LDS, coalesced ----
		for (uint32_t o = (uint32_t)tid * 16u; o < lim; o += kStkThreads * 16u) {
			if (o + 16u <= lim) {
				const uint4 v = *reinterpret_cast<const uint4*>(a.bchar + base + o);
				uint32_t	x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
				for (int q = 0; q < 4; ++q)
					x[q] = (uint32_t)s_kt[x[q] & 0xFFu] | (uint32_t)s_kt[(x[q] >> 8) & 0xFFu] << 8 |
						   (uint32_t)s_kt[(x[q] >> 16) & 0xFFu] << 16 | (uint32_t)s_kt[x[q] >> 24] << 24;
				*reinterpret_cast<uint4*>(s_ch + o) = make_uint4(x[0], x[1], x[2], x[3]);
			} else {
				for (uint32_t q = o; q < lim; ++q) s_ch[q] = s_kt[a.bchar[base + q]];
			}
		}
		__syncthreads();
		const uint32_t l0  = (uint32_t)tid * kStkItems;	 // chunk-local index of the thread's first bracket
		const uint32_t cnt = l0 < lim ? min(lim - l0, (uint32_t)kStkItems) : 0u;
		uint32_t	   w[8];
#pragma unroll
		for (int q = 0; q < 8; ++q) w[q] = reinterpret_cast<const uint32_t*>(s_ch + l0)[q];
		// ---- the thread's stack: the opens' bits first (independent lookups), then bit
		// operations only; each pair's bytes are checked when the chunk's partners are stored ----
		uint32_t opn = 0;
#pragma unroll
		for (uint32_t q = 0; q < (uint32_t)kStkItems; ++q)
			opn |= ((w[q >> 2] >> (8 * (q & 3) + 7)) & 1u) << q;
		const uint32_t valid = cnt >= 32u ? ~0u : ((1u << cnt) - 1u);
		opn &= valid;
		const uint32_t cls	= valid & ~opn;
		uint32_t	   mask = 0, rcm = 0;  // unmatched opens; unmatched closes
#pragma unroll
		for (uint32_t q = 0; q < (uint32_t)kStkItems; ++q) {
			if ((opn >> q) & 1u) {
				mask |= 1u << q;
			} else if ((cls >> q) & 1u) {
				if (mask) {
					const uint32_t p = 31u - (uint32_t)__clz(mask);
					mask ^= 1u << p;
					s_part[l0 + p] = (uint16_t)(l0 + q);
					s_part[l0 + q] = (uint16_t)(l0 + p);
				} else {
					rcm |= 1u << q;
				}
			}
		}
		const uint32_t ro = (uint32_t)__popc(mask), rc = (uint32_t)__popc(rcm);
		// ---- depths relative to the chunk's entry: e = exclusive prefix of deltas ----
		const int32_t delta = (int32_t)ro - (int32_t)rc;
		const int32_t winc	= (int32_t)wave_incl_sum((uint32_t)delta);
		if (lane == 63)
			s_wsum[wid] = winc;
		if (tid == 0)
			s_entry = key0 + ((w[0] & 0x80u) ? 0 : 1);	// depth before the first bracket
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
		s_ro[tid]		= mask;
		__syncthreads();
#pragma unroll
		for (int j = 1; j < kStkLevels; ++j) {
			const int h = 1 << (j - 1);
			s_st[j][tid] = tid >= h ? min(s_st[j - 1][tid], s_st[j - 1][tid - h]) : s_st[j - 1][tid];
			__syncthreads();
		}
		const int32_t mc	= range_min(s_st, 0, kStkThreads - 1);	// <= 0: thread 0 enters at 0
		const int32_t suf	= tid + 1 < kStkThreads ? range_min(s_st, tid + 1, kStkThreads - 1) : INT_MAX;
		const int32_t entry = s_entry;
		{
			// the last earlier thread whose minimum is below this one's (all threads at
			// once): the closes below walk this chain instead of searching per level
			const int32_t mt  = (int32_t)s_st[0][tid];
			const int	  psm = last_at_most(s_st, tid, mt - 1);
			s_mp[tid]		  = (uint32_t)(uint16_t)(int16_t)mt | (uint32_t)(psm + 1) << 16;
		}
		const uint32_t rinc = wave_incl_sum(rc);
		if (lane == 63)
			s_wsum[wid] = (int32_t)rinc;
		__syncthreads();  // the chain is complete and the sparse table free
		uint32_t roff = rinc - rc, rtot = 0;
#pragma unroll
		for (int q = 0; q < kStkThreads / 64; ++q) {
			if (q < wid)
				roff += (uint32_t)s_wsum[q];
			rtot += (uint32_t)s_wsum[q];
		}
		// The close at level k (entry depth e, its j-th unmatched close, k = e - 1 - j)
		// pairs with the open at level k of the last thread before t whose minimum is
		// <= k, or is the chunk's residue.  u, wu: the candidate and its chain word;
		// levels only descend from one close to the next of a thread, so the answer
		// only moves back along the chain.
		auto close_at = [&](int32_t k, uint32_t lc, int& u, uint32_t& wu) {
			while (u >= 0 && (int32_t)(int16_t)(wu & 0xFFFFu) > k) {
				u  = (int)(wu >> 16) - 1;
				wu = u >= 0 ? s_mp[u] : 0u;
			}
			if (u >= 0) {
				const int32_t jo = k - (int32_t)(int16_t)(wu & 0xFFFFu);
				if (jo < 0 || jo >= __popc(s_ro[u])) {	// cannot happen: the open at level k
					lost = true;
					return;
				}
				const uint32_t lo = (uint32_t)u * kStkItems + select_bit(s_ro[u], (uint32_t)jo);
				s_part[lc]		  = (uint16_t)lo;
				s_part[lo]		  = (uint16_t)lc;
			} else {
				const uint32_t idx = (uint32_t)(-1 - k);  // closes: level entry-1, entry-2, ...
				if (k < 0 && idx < lim)
					a.res[base + idx] = base + lc;
				else
					lost = true;
				s_part[lc] = kPartLater;
			}
		};
		// ---- the unmatched closes, spread evenly over the threads through a list in
		// the sparse table's space (a thread's own can number 0 to 32) ----
		uint32_t* s_list = reinterpret_cast<uint32_t*>(&s_st[0][0]);
		const bool listed = rtot <= kListCap;  // the same in every lane
		if (listed) {
			uint32_t r = rcm;
			for (uint32_t j = 0; r; ++j, r &= r - 1)
				s_list[roff + j] = (l0 + (uint32_t)(__ffs(r) - 1)) | j << 13 | (uint32_t)tid << 18;
			s_e[tid] = (int16_t)e;
		}
		__syncthreads();
		if (listed) {
			const uint32_t per = (rtot + kStkThreads - 1) / kStkThreads;
			const uint32_t i0 = min(rtot, (uint32_t)tid * per), i1 = min(rtot, i0 + per);
			int			   t_cur = -1, u = -1;
			uint32_t	   wu = 0;
			int32_t		   et = 0;
			for (uint32_t i = i0; i < i1; ++i) {
				const uint32_t w = s_list[i];
				const int	   t = (int)(w >> 18);
				if (t != t_cur) {
					t_cur = t;
					et	  = s_e[t];
					u	  = t - 1;
					wu	  = u >= 0 ? s_mp[u] : 0u;
				}
				close_at(et - 1 - (int32_t)((w >> 13) & 31u), w & 0x1FFFu, u, wu);
			}
		} else {
			uint32_t r	= rcm;
			int		 u	= tid - 1;
			uint32_t wu = u >= 0 ? s_mp[u] : 0u;
			for (uint32_t j = 0; r; ++j, r &= r - 1)
				close_at(e - 1 - (int32_t)j, l0 + (uint32_t)(__ffs(r) - 1), u, wu);
		}
		// ---- the thread's unmatched opens no later thread closes: the chunk's residue ----
		{
			uint32_t r = mask;
			for (uint32_t j = 0; r; ++j, r &= r - 1) {
				const int32_t k = m + (int32_t)j;
				if (k < suf) {
					const uint32_t lo  = l0 + (uint32_t)(__ffs(r) - 1);
					const uint32_t idx = (uint32_t)(-mc) + (uint32_t)(k - mc);	// opens: level min, min+1, ...
					if (idx < lim)
						a.res[base + idx] = base + lo;
					else
						lost = true;
					s_part[lo] = kPartLater;
				}
			}
		}
		if (tid == 0) {
			a.crec[3 * (uint64_t)c + 0] = entry;
			a.crec[3 * (uint64_t)c + 1] = entry + mc;
			a.crec[3 * (uint64_t)c + 2] = entry + xc;
			atomicMax(a.gmin1 + (c >> 6), tag_min(epoch, entry + mc));
			atomicMax(a.gmin2 + (c >> 12), tag_min(epoch, entry + mc));
		}
		__syncthreads();
		// ---- the chunk's partners, checked and stored coalesced: a pair whose bytes do
		// not match is -1 on both sides (nest.hpp:80-88); residues get -1 until the
		// resolve pass ----
		uint32_t l_from = (uint32_t)tid;
		if (!a.bdoc && !a.pair_pos) {
			// one document without pair_pos: four partners per thread, one 16-byte store
			const uint32_t l4 = lim & ~3u;
#pragma unroll 2
			for (uint32_t l = (uint32_t)tid * 4u; l < l4; l += kStkThreads * 4u) {
				const uint2	   pv2 = *reinterpret_cast<const uint2*>(s_part + l);
				const uint32_t kl  = *reinterpret_cast<const uint32_t*>(s_ch + l);
				const uint32_t v4[4] = {pv2.x & 0xFFFFu, pv2.x >> 16, pv2.y & 0xFFFFu, pv2.y >> 16};
				int32_t		   o[4];
#pragma unroll
				for (int q = 0; q < 4; ++q) {
					o[q] = -1;
					if (v4[q] < kStkChunk) {
						const bool ok = (((kl >> (8 * q)) ^ s_ch[v4[q]]) & 0x7Fu) == 0u;
						bad			  = bad || !ok;
						o[q]		  = ok ? (int32_t)(base + v4[q]) : -1;
					}
				}
				*reinterpret_cast<int4*>(a.pair + base + l) = make_int4(o[0], o[1], o[2], o[3]);
			}
			l_from = l4 + (uint32_t)tid;
		}
#pragma unroll 4
		for (uint32_t l = l_from; l < lim; l += kStkThreads) {
			const uint16_t v  = s_part[l];
			int32_t		   pv = -1;
			if (v < (uint16_t)kStkChunk) {
				const bool ok = ((s_ch[l] ^ s_ch[v]) & 0x7Fu) == 0u;
				bad				  = bad || !ok;
				pv				  = ok ? (int32_t)(base + v) : -1;
				if (!ok && a.doc_bad)
					a.doc_bad[a.bdoc[base + l]] = 1u;
			}
			put_pair(a, base + l, pv, v != kPartLater);
		}
	}
	if (__ballot(bad) && lane == 0)
		atomicOr(&a.sum->mismatched, 1u);
	if (__ballot(lost) && lane == 0)
		atomicOr(&a.sum->overflow, kStackLost);
}

// The nearest chunk p <= cur whose minimum is below thr, or -1: one wave,
// 64 chunks per step, groups of 64 chunks (gmin1) and of 4096 (gmin2) skipped
// by their minima and entered where one is below thr.
__device__ int64_t prev_below(const StackArgs& a, uint32_t epoch, int64_t cur, int32_t thr) {
	const int lane = lane_id();
	auto	  m0   = [&](int64_t c) -> int32_t { return a.crec[3 * (uint64_t)c + 1]; };
	auto	  m1   = [&](int64_t g) -> int32_t { return read_min(a.gmin1[g], epoch); };
	auto	  m2   = [&](int64_t h) -> int32_t { return read_min(a.gmin2[h], epoch); };
	// the highest index <= hi of a 64-aligned window where f(index) < thr, or -1
	auto scan64 = [&](int64_t hi, auto f) -> int64_t {
		const int64_t lo = hi & ~63ll, i = lo + lane;
		const uint64_t b = __ballot(i <= hi && f(i) < thr);
		return b ? lo + 63 - __clzll((long long)b) : -1;
	};
	while (cur >= 0) {
		int64_t p = scan64(cur, m0);  // the rest of cur's group of 64
		if (p >= 0)
			return p;
		int64_t g = (cur >> 6) - 1;	// groups before it
		if (g < 0)
			return -1;
		int64_t G = scan64(g, m1);	// the rest of its group of 4096
		if (G < 0) {
			int64_t h = (g >> 6) - 1;  // groups of 4096 before it, 64 at a time
			int64_t H = -1;
			while (h >= 0 && (H = scan64(h, m2)) < 0) h = (h & ~63ll) - 1;
			if (H < 0)
				return -1;
			G = scan64(H * 64 + 63, m1);
			if (G < 0)	// a stale minimum (cannot happen): continue below that group
			{
				cur = H * 4096 - 1;
				continue;
			}
		}
		p = scan64(G * 64 + 63, m0);
		if (p >= 0)
			return p;
		cur = G * 64 - 1;  // a stale minimum (cannot happen)
	}
	return -1;
}

// Every chunk's residue closes, one wave per chunk: walk back to the nearest
// chunk whose minimum is below the highest level still open; its opens close
// the levels from max(its minimum, ours) up; repeat below it.
__global__ __launch_bounds__(kResolveBlock) void k_stack_resolve(StackArgs a) {
	if (!stack_applies(a))
		return;
	const uint32_t nb	 = stack_nb(a);
	const uint32_t nch	 = (nb + kStkChunk - 1) / kStkChunk;
	const uint32_t epoch = a.sum->epoch;
	const int	   lane	 = threadIdx.x & 63;
	const uint32_t waves = gridDim.x * (kResolveBlock / 64);
	const DevPlan* P	 = a.plan;
	bool		   bad = false, lost = false;
	for (uint32_t c = blockIdx.x * (kResolveBlock / 64) + (threadIdx.x >> 6); c < nch; c += waves) {
		const int32_t E = a.crec[3 * (uint64_t)c], M = a.crec[3 * (uint64_t)c + 1];
		int32_t		  rm  = E;	// levels [M, rm) still need their opens
		int64_t		  cur = (int64_t)c - 1;
		while (rm > M) {
			const int64_t p = prev_below(a, epoch, cur, rm);
			if (p < 0) {  // cannot happen on a balanced document
				lost = true;
				break;
			}
			const int32_t Ep = a.crec[3 * (uint64_t)p], Mp = a.crec[3 * (uint64_t)p + 1],
						  Xp = a.crec[3 * (uint64_t)p + 2];
			const int32_t lo = max(Mp, M);
			for (int32_t k = lo + lane; k < rm; k += 64) {
				const uint64_t ic = (uint64_t)c * kStkChunk + (uint32_t)(E - 1 - k);
				const uint64_t io = (uint64_t)p * kStkChunk + (uint32_t)(Ep - Mp) + (uint32_t)(k - Mp);
				if (k >= Xp || ic >= nb || io >= nb) {	// cannot happen on a balanced document
					lost = true;
					continue;
				}
				const uint32_t gc = a.res[ic], go = a.res[io];
				if (gc >= nb || go >= nb) {
					lost = true;
					continue;
				}
				const bool ok = P->br_close_of[a.bchar[go]] == a.bchar[gc];
				bad			  = bad || !ok;
				if (!ok && a.doc_bad)
					a.doc_bad[a.bdoc[gc]] = 1u;
				put_pair(a, gc, ok ? (int32_t)go : -1, true);
				put_pair(a, go, ok ? (int32_t)gc : -1, true);
			}
			rm	= lo;
			cur = p - 1;
		}
	}
	if (__ballot(bad) && lane == 0)
		atomicOr(&a.sum->mismatched, 1u);
	if (__ballot(lost) && lane == 0)
		atomicOr(&a.sum->overflow, kStackLost);
}

}  // namespace

hipError_t launch_stack_pairing(const StackArgs& a, hipStream_t s) {
	if (a.cap_br == 0)
		return hipSuccess;
	int dev = 0, cus = 0;
	(void)hipGetDevice(&dev);
	(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	cus					 = std::max(cus, 1);
	// as many blocks as stay resident, so every block takes the same share of chunks
	int occ = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stack_local, kStkThreads, 0) != hipSuccess || occ < 1)
		occ = 4;
	const unsigned lgrid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(a.max_chunks, 1), (uint64_t)occ * cus);
	const unsigned rgrid = (unsigned)std::min<uint64_t>(
		std::max<uint64_t>((a.max_chunks + kResolveBlock / 64 - 1) / (kResolveBlock / 64), 1), 8ull * cus);
	hipLaunchKernelGGL(k_stack_local, dim3(lgrid), dim3(kStkThreads), 0, s, a);
	hipLaunchKernelGGL(k_stack_resolve, dim3(rgrid), dim3(kResolveBlock), 0, s, a);
	return hipGetLastError();
}

}  // namespace pars_hip
