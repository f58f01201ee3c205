// pair_kernels.hip — R6: depth-sort bracket pairing, sized entirely on device.
//
// Restates model::pair_brackets (include/pars/model/nest.hpp:45-92), the CPU
// semantics the results must equal bit for bit: key = depth-1 for an open,
// depth for a close; a STABLE sort by key; sorted neighbours (2i, 2i+1) pair
// when they are a matching open/close.  The reference GPU chain
// (pipeline.cuh:932-996) clamps keys at 0 and sorts unstably through Thrust,
// which diverges from the CPU on malformed input; here the full key range is
// sorted stably, so the two agree on every input.
//
// Layout: the scan kernel leaves bkey/bchar/bsidx per bracket (keys of a single
// document only at chunk starts when it was scanned for the stack path: the
// key pass below rebuilds the rest before a sort).  The radix sort
// runs LSD passes over u = key - kmin with a digit width fitted to the key
// range found on device (one 6-bit pass for every document up to depth 63,
// at most 8 bits per pass - or a tuned narrower cap - and 1..4 passes): per-chunk digit histograms (LDS
// atomics — counts are order-free), one device-wide exclusive scan of the
// digit-major histogram, then a stable scatter that ranks equal digits in
// input order with wave ballots.  The last pass also records each bracket's
// sorted rank, so pairing walks brackets in document order, gathers the
// sorted neighbour, and writes every output sequentially.
#include "dev_common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace pars_hip {

namespace {

constexpr int kSortBlock = 256;
constexpr int kSortRows	 = 16;
constexpr int kChunk	 = kSortBlock * kSortRows;	// brackets per sort chunk
constexpr int kScanGrid	 = 256;						// blocks of the histogram scan

__device__ __forceinline__ uint32_t n_brackets(const PairArgs& a) {
	const uint32_t n = a.sum->n_brackets;
	return (uint64_t)n > a.cap_br ? (uint32_t)a.cap_br : n;
}

__global__ void k_sort_prep(PairArgs a) {
	if (threadIdx.x != 0 || blockIdx.x != 0)
		return;
	DevSummary* s = a.sum;
	if (s->small_done) {  // the one-block path paired everything: no pass runs
		s->sort_passes = 0;
		s->n_chunks	   = 0;
		return;
	}
	const uint32_t nb	= n_brackets(a);
	const int32_t  kmin = s->min_bal - 1;  // open keys are depth-1, depths >= min_bal
	const int32_t  kmax = s->max_bal;
	// batch: the key is a document's own key range, placed after the ranges of
	// the documents before it (k_kbase_*)
	const uint64_t range = a.bdoc ? (s->key_total ? s->key_total - 1u : 0u) : (uint64_t)(kmax - kmin);
	uint32_t	   bits	 = range ? 64u - (uint32_t)__clzll((long long)range) : 1u;
	if (bits > 32u) {  // more than one 32-bit sort key can hold
		s->overflow |= 4u;
		bits = 32u;
	}
	const uint32_t cap = a.max_digit_bits >= 1u && a.max_digit_bits <= 8u ? a.max_digit_bits : 8u;
	const uint32_t p   = (bits + cap - 1u) / cap;
	if (p > 4u)	 // four passes are launched: wider digits
		s->overflow |= 8u;
	s->kmin			 = kmin;
	s->sort_passes		 = p;
	s->digit_bits		 = (bits + p - 1) / p;
	s->n_chunks			 = (nb + kChunk - 1) / kChunk;
	if ((uint64_t)s->n_chunks > a.max_chunks)
		s->overflow |= 2u;
}

// ---- keys of a document scanned with sparse keys ----------------------------
// The scan kernel left the key of every kStkChunk-th bracket only.  One block
// per chunk rebuilds the others from the bracket bytes: the depth before the
// chunk's first bracket follows from its key (an open's key is the depth before
// it, a close's the depth after it), a block scan of the +1 / -1 steps gives
// each thread's entry depth, and each thread walks its 32 brackets.
__global__ __launch_bounds__(256) void k_key_fill(PairArgs a) {
	__shared__ uint8_t s_open[256];
	__shared__ int32_t s_w[4];
	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	s_open[tid]	  = a.plan->br_flags[tid] & 1u;
	__syncthreads();
	const uint32_t nb  = n_brackets(a);
	const uint32_t nch = (nb + kStkChunk - 1) / kStkChunk;
	for (uint32_t c = blockIdx.x; c < nch; c += gridDim.x) {
		const uint32_t base = c * kStkChunk, lim = min(nb - base, kStkChunk);
		const uint32_t l0 = (uint32_t)tid * 32u, cnt = l0 < lim ? min(lim - l0, 32u) : 0u;
		const uint8_t* ch = a.bchar + base + l0;
		uint32_t	   om = 0;	// bit q: bracket q of the thread opens
		if (cnt == 32u) {
			uint4 v[2];
			v[0] = reinterpret_cast<const uint4*>(ch)[0];
			v[1] = reinterpret_cast<const uint4*>(ch)[1];
			const uint32_t* w = reinterpret_cast<const uint32_t*>(v);
#pragma unroll
			for (int q = 0; q < 32; ++q) om |= (uint32_t)s_open[(w[q >> 2] >> (8 * (q & 3))) & 0xFFu] << q;
		} else {
			for (uint32_t q = 0; q < cnt; ++q) om |= (uint32_t)s_open[ch[q]] << q;
		}
		const int32_t  delta = 2 * __popc(om) - (int32_t)cnt;
		const uint32_t inc	 = wave_incl_sum((uint32_t)delta);
		if (lane == 63)
			s_w[wid] = (int32_t)inc;
		const int32_t key0 = a.key_fill[base];
		const int32_t d0   = s_open[a.bchar[base]] ? key0 : key0 + 1;	// depth before the chunk
		__syncthreads();
		int32_t d = d0 + (int32_t)inc - delta;
		for (int w2 = 0; w2 < wid; ++w2) d += s_w[w2];
		int32_t* out = a.key_fill + base + l0;
		if (cnt == 32u) {
			int32_t k[32];
#pragma unroll
			for (int q = 0; q < 32; ++q) {
				const bool op = (om >> q) & 1u;
				k[q]		  = op ? d : d - 1;
				d += op ? 1 : -1;
			}
#pragma unroll
			for (int q = 0; q < 8; ++q)
				reinterpret_cast<int4*>(out)[q] = make_int4(k[4 * q], k[4 * q + 1], k[4 * q + 2], k[4 * q + 3]);
		} else {
			for (uint32_t q = 0; q < cnt; ++q) {
				const bool op = (om >> q) & 1u;
				out[q]		  = op ? d : d - 1;
				d += op ? 1 : -1;
			}
		}
		__syncthreads();  // s_w is reused by the next chunk
	}
}

// ---- batch key bases -------------------------------------------------------
// Document d's keys lie in [doc_min[d] - 1, doc_max[d]] (doc_min <= 0 <= doc_max:
// depths after its brackets, and 0): span_d = doc_max[d] - doc_min[d] + 2 values.
// doc_kbase[d] = the sum of the spans of the documents before d, so the batch's
// (document, key) order is the order of doc_kbase[d] + key - (doc_min[d] - 1),
// within sum(span_d) <= n_brackets + 2 n_docs values.
__device__ __forceinline__ uint64_t doc_span(const PairArgs& a, uint64_t d) {
	return (uint64_t)((int64_t)a.doc_max[d] - (int64_t)a.doc_min[d]) + 2u;
}

__device__ __forceinline__ uint32_t doc_key(const PairArgs& a, uint32_t i) {
	const uint32_t d = a.bdoc[i];
	return a.doc_kbase[d] + (uint32_t)(a.bkey[i] - (a.doc_min[d] - 1));
}

// Exclusive block scan of 64-bit values (blockDim.x a multiple of 64, <= 1024).
__device__ uint64_t block_excl_sum64(uint64_t v, uint64_t* s_w, uint64_t& total) {
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
	uint64_t  inc  = v;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint64_t t = __shfl_up(inc, o, 64);
		if (lane >= o)
			inc += t;
	}
	if (lane == 63)
		s_w[wid] = inc;
	__syncthreads();
	uint64_t p = 0, t = 0;
	for (int w = 0; w < nw; ++w) {
		if (w < wid)
			p += s_w[w];
		t += s_w[w];
	}
	__syncthreads();
	total = t;
	return p + inc - v;
}

__device__ __forceinline__ void store_kbase(const PairArgs& a, uint64_t d, uint64_t base) {
	a.doc_kbase[d] = (uint32_t)base;  // exact whenever key_total fits 32 bits (else the sort refuses)
}

// Up to kKbaseOneBlock documents: one block.
constexpr uint32_t kKbaseOneBlock = 16384;

__global__ __launch_bounds__(1024) void k_kbase_one(PairArgs a) {
	if (a.sum->small_done)  // a batch paired by stack matching needs no key bases
		return;
	__shared__ uint64_t s_w[16];
	const uint64_t		nd	= a.n_docs;
	uint64_t			run = 0;
	for (uint64_t d0 = 0; d0 < nd; d0 += blockDim.x) {
		const uint64_t d = d0 + threadIdx.x;
		const uint64_t v = d < nd ? doc_span(a, d) : 0u;
		uint64_t	   tot;
		const uint64_t ex = block_excl_sum64(v, s_w, tot);
		if (d < nd)
			store_kbase(a, d, run + ex);
		run += tot;
	}
	if (threadIdx.x == 0)
		a.sum->key_total = run;
}

__device__ __forceinline__ void kb_range(uint64_t len, uint32_t b, uint64_t& lo, uint64_t& hi) {
	const uint64_t per = (len + kScanGrid - 1) / kScanGrid;
	lo				   = min<uint64_t>((uint64_t)b * per, len);
	hi				   = min<uint64_t>(lo + per, len);
}

__global__ __launch_bounds__(kSortBlock) void k_kbase_partial(PairArgs a) {
	if (a.sum->small_done)  // a batch paired by stack matching needs no key bases
		return;
	__shared__ uint64_t s_w[4];
	uint64_t			lo, hi;
	kb_range(a.n_docs, blockIdx.x, lo, hi);
	uint64_t acc = 0;
	for (uint64_t d = lo + threadIdx.x; d < hi; d += kSortBlock) acc += doc_span(a, d);
	uint64_t tot;
	(void)block_excl_sum64(acc, s_w, tot);
	if (threadIdx.x == 0)
		a.kb_part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kSortBlock) void k_kbase_top(PairArgs a) {
	if (a.sum->small_done)  // a batch paired by stack matching needs no key bases
		return;
	__shared__ uint64_t s_w[4];
	uint64_t			tot;
	const uint64_t		v  = a.kb_part[threadIdx.x];  // kScanGrid == kSortBlock
	const uint64_t		ex = block_excl_sum64(v, s_w, tot);
	a.kb_part[threadIdx.x] = ex;
	if (threadIdx.x == 0)
		a.sum->key_total = tot;
}

__global__ __launch_bounds__(kSortBlock) void k_kbase_apply(PairArgs a) {
	if (a.sum->small_done)  // a batch paired by stack matching needs no key bases
		return;
	__shared__ uint64_t s_w[4];
	uint64_t			lo, hi;
	kb_range(a.n_docs, blockIdx.x, lo, hi);
	uint64_t run = a.kb_part[blockIdx.x];
	for (uint64_t d0 = lo; d0 < hi; d0 += kSortBlock) {
		const uint64_t d = d0 + threadIdx.x;
		const uint64_t v = d < hi ? doc_span(a, d) : 0u;
		uint64_t	   tot;
		const uint64_t ex = block_excl_sum64(v, s_w, tot);
		if (d < hi)
			store_kbase(a, d, run + ex);
		run += tot;
	}
}

__device__ __forceinline__ uint32_t key_at(const PairArgs& a, int pass, uint32_t i, int32_t kmin) {
	if (pass != 0)
		return a.ukeys[(pass - 1) & 1][i];
	return a.bdoc ? doc_key(a, i) : (uint32_t)(a.bkey[i] - kmin);
}

template<int PASS>
__global__ __launch_bounds__(kSortBlock) void k_sort_hist(PairArgs a) {
	if ((uint32_t)PASS >= a.sum->sort_passes)
		return;
	__shared__ uint32_t h[256];
	const uint32_t		nb = n_brackets(a), nc = a.sum->n_chunks;
	const int32_t		kmin = a.sum->kmin;
	const uint32_t		w = a.sum->digit_bits, B = 1u << w, mask = B - 1u, shift = PASS * w;
	const int			tid = threadIdx.x;
	for (uint32_t c = blockIdx.x; c < nc; c += gridDim.x) {
		h[tid] = 0;
		__syncthreads();
#pragma unroll 4
		for (int r = 0; r < kSortRows; ++r) {
			const uint32_t i = c * kChunk + r * kSortBlock + tid;
			if (i < nb)
				atomicAdd(&h[(key_at(a, PASS, i, kmin) >> shift) & mask], 1u);
		}
		__syncthreads();
		if ((uint32_t)tid < B)
			a.hist[(uint64_t)tid * nc + c] = h[tid];
		__syncthreads();
	}
}

// Device-wide exclusive scan of hist[0 .. B*nc), three phases, fixed grid.
__device__ __forceinline__ void scan_range(uint32_t len, uint32_t b, uint32_t& lo, uint32_t& hi) {
	const uint64_t per = ((uint64_t)len + kScanGrid - 1) / kScanGrid;
	lo				   = (uint32_t)min<uint64_t>((uint64_t)b * per, len);
	hi				   = (uint32_t)min<uint64_t>((uint64_t)lo + per, len);
}

__device__ uint32_t block_excl_sum(uint32_t v, uint32_t* s_w, uint32_t& total) {
	const int	   lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const uint32_t inc	= wave_incl_sum(v);
	if (lane == 63)
		s_w[wid] = inc;
	__syncthreads();
	uint32_t p = 0, t = 0;
#pragma unroll
	for (int w = 0; w < kSortBlock / 64; ++w) {
		if (w < wid)
			p += s_w[w];
		t += s_w[w];
	}
	__syncthreads();
	total = t;
	return p + inc - v;
}

__device__ __forceinline__ uint32_t hist_len(const PairArgs& a) {
	return (1u << a.sum->digit_bits) * a.sum->n_chunks;
}

template<int PASS>
__global__ __launch_bounds__(kSortBlock) void k_scan_partial(PairArgs a) {
	if ((uint32_t)PASS >= a.sum->sort_passes)
		return;
	__shared__ uint32_t s_w[4];
	uint32_t			lo, hi;
	scan_range(hist_len(a), blockIdx.x, lo, hi);
	uint32_t acc = 0;
	for (uint32_t i = lo + threadIdx.x; i < hi; i += kSortBlock) acc += a.hist[i];
	uint32_t tot;
	(void)block_excl_sum(acc, s_w, tot);
	if (threadIdx.x == 0)
		a.hist_part[blockIdx.x] = tot;
}

template<int PASS>
__global__ __launch_bounds__(kSortBlock) void k_scan_top(PairArgs a) {
	if ((uint32_t)PASS >= a.sum->sort_passes)
		return;
	__shared__ uint32_t s_w[4];
	uint32_t			tot;
	const uint32_t		v = a.hist_part[threadIdx.x];  // kScanGrid == kSortBlock
	a.hist_part[threadIdx.x] = block_excl_sum(v, s_w, tot);
}

template<int PASS>
__global__ __launch_bounds__(kSortBlock) void k_scan_apply(PairArgs a) {
	if ((uint32_t)PASS >= a.sum->sort_passes)
		return;
	__shared__ uint32_t s_w[4];
	uint32_t			lo, hi;
	scan_range(hist_len(a), blockIdx.x, lo, hi);
	uint32_t run = a.hist_part[blockIdx.x];
	for (uint32_t i0 = lo; i0 < hi; i0 += kSortBlock) {
		const uint32_t i = i0 + threadIdx.x;
		const uint32_t v = i < hi ? a.hist[i] : 0u;
		uint32_t	   tot;
		const uint32_t ex = block_excl_sum(v, s_w, tot);
		if (i < hi)
			a.hist[i] = run + ex;
		run += tot;
	}
}

template<int PASS>
__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(PairArgs a) {
	const uint32_t passes = a.sum->sort_passes;
	if ((uint32_t)PASS >= passes)
		return;
	__shared__ uint32_t s_base[256];
	__shared__ uint32_t s_wc[4][256];
	__shared__ uint32_t s_wo[4][256];
	const uint32_t		nb = n_brackets(a), nc = a.sum->n_chunks;
	const int32_t		kmin = a.sum->kmin;
	const uint32_t		w = a.sum->digit_bits, B = 1u << w, mask = B - 1u, shift = PASS * w;
	const bool			last = (uint32_t)PASS + 1 == passes;
	const int			tid = threadIdx.x, wid = tid >> 6;
	uint32_t*			okey = a.ukeys[PASS & 1];
	uint32_t*			oval = a.order[PASS & 1];
	for (uint32_t c = blockIdx.x; c < nc; c += gridDim.x) {
		s_base[tid] = (uint32_t)tid < B ? a.hist[(uint64_t)tid * nc + c] : 0u;
#pragma unroll
		for (int q = 0; q < 4; ++q) s_wc[q][tid] = 0;
		__syncthreads();
		for (int r = 0; r < kSortRows; ++r) {
			const uint32_t i	 = c * kChunk + r * kSortBlock + tid;
			const bool	   valid = i < nb;
			uint32_t	   u = 0, v = 0, d = 0;
			if (valid) {
				u = key_at(a, PASS, i, kmin);
				v = PASS == 0 ? i : a.order[(PASS - 1) & 1][i];
				d = (u >> shift) & mask;
			}
			// lanes holding the same digit, by ballots over the digit's bits
			uint64_t peers = __ballot(valid);
			for (uint32_t b = 0; b < w; ++b) {
				const bool	   bit = (d >> b) & 1u;
				const uint64_t bb  = __ballot(valid && bit);
				peers &= bit ? bb : ~bb;
			}
			const uint32_t wrank = (uint32_t)__popcll(peers & lanemask_lt());
			if (valid && wrank == 0)
				s_wc[wid][d] = (uint32_t)__popcll(peers);
			__syncthreads();
			if ((uint32_t)tid < B) {
				uint32_t run = s_base[tid];
#pragma unroll
				for (int q = 0; q < 4; ++q) {
					const uint32_t t = s_wc[q][tid];
					s_wo[q][tid]	 = run;
					run += t;
					s_wc[q][tid] = 0;
				}
				s_base[tid] = run;
			}
			__syncthreads();
			if (valid) {
				const uint32_t dst = s_wo[wid][d] + wrank;
				oval[dst]		   = v;
				if (last)
					a.rank[v] = dst;
				else
					okey[dst] = u;
			}
		}
		__syncthreads();
	}
}

// nest.hpp:78-90: sorted neighbours (2i, 2i+1) pair if open/close and matching.
// Walked in document order: bracket i sits at sorted position rank[i]; its
// neighbour is at rank[i] ^ 1.
__global__ __launch_bounds__(256) void k_pair(PairArgs a) {
	if (a.sum->small_done)
		return;
	const uint32_t	nb	  = n_brackets(a);
	const uint32_t	pass  = a.sum->sort_passes;
	const uint32_t* order = a.order[(pass - 1) & 1];
	const DevPlan*	P	  = a.plan;
	bool			bad	  = false;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
		// sorted positions count from the document's first bracket (batch) or 0
		uint32_t d = 0, b0 = 0, n = nb, s0 = 0;
		if (a.bdoc) {
			d  = a.bdoc[i];
			b0 = a.br_start[d];
			n  = a.br_start[d + 1] - b0;
			s0 = a.st_start[d];
		}
		const uint32_t p = a.rank[i] - b0, q = p ^ 1u;
		if (a.close) {	// partners only (the tape reads no pairs and no mismatch report)
			if (P->br_flags[a.bchar[i]] & 1u) {
				// the bracket after this open in sorted order, when it has the same
				// key: its close (after an open at key d only a close brings the depth
				// back to d, so the next bracket at key d in document order closes)
				uint32_t tc = ~0u;
				if (p + 1u < n) {
					const uint32_t y = order[b0 + p + 1u];
					if (a.bkey[y] == a.bkey[i])
						tc = a.bsidx[y];
				}
				a.close[i] = tc;
			}
			continue;
		}
		int32_t		   partner = -1;
		if (q < n) {
			const uint32_t j  = order[b0 + q];
			const uint32_t ia = (p & 1u) ? j : i, ib = (p & 1u) ? i : j;
			const uint8_t  ca = a.bchar[ia], cb = a.bchar[ib];
			const bool ok = (P->br_flags[ca] & 1u) && (P->br_flags[cb] & 2u) && P->br_close_of[ca] == cb;
			if (ok)
				partner = (int32_t)j;
			else if (!(p & 1u)) {
				bad = true;	 // one report per sorted pair
				if (a.doc_bad)
					a.doc_bad[d] = 1u;
			}
		}
		if (a.pair)
			a.pair[i] = partner >= 0 ? partner - (int32_t)b0 : -1;
		if (a.pair_pos)
			a.pair_pos[a.bsidx[i]] = partner >= 0 ? (int32_t)(a.bsidx[partner] - s0) : -1;
	}
	if (__ballot(bad) && lane_id() == 0)
		atomicOr(&a.sum->mismatched, 1u);
}

// R6 in one block for at most kSmallBrackets brackets whose (document, key)
// range fits 19 bits (batch keys as in k_kbase_*): every bracket becomes
// (key << 13) | index, so sorting
// these distinct values is the stable sort by key; a bitonic sort in LDS, then
// sorted neighbours pair exactly as in k_pair.  Small batches (the serving
// case) and small documents then cost one launch instead of the radix passes.
// Sets small_done when it applied; otherwise leaves everything to them.
constexpr int kSmallBlock = 1024;

__global__ __launch_bounds__(kSmallBlock) void k_pair_small(PairArgs a) {
	__shared__ uint32_t s_v[kSmallBrackets];
	__shared__ uint32_t s_ok;
	DevSummary*			sum = a.sum;
	const uint32_t		nb	= n_brackets(a);
	const int			tid = threadIdx.x;
	const int32_t		kmin = sum->min_bal - 1, kmax = sum->max_bal;
	const uint64_t		range = a.bdoc ? (sum->key_total ? sum->key_total - 1u : 0u) : (uint64_t)(kmax - kmin);
	if (nb > kSmallBrackets || nb == 0 || range >= (1ull << 19))
		return;	 // not this path (nb == 0: nothing to pair, flags need no sort)
	uint32_t N = 1;
	while (N < nb) N <<= 1;
	for (uint32_t i = tid; i < N; i += kSmallBlock) {
		uint32_t v = 0xFFFFFFFFu;  // padding sorts last
		if (i < nb)
			v = ((a.bdoc ? doc_key(a, i) : (uint32_t)(a.bkey[i] - kmin)) << 13) | i;
		s_v[i] = v;
	}
	__syncthreads();
	for (uint32_t k = 2; k <= N; k <<= 1)
		for (uint32_t j = k >> 1; j > 0; j >>= 1) {
			for (uint32_t i = tid; i < N; i += kSmallBlock) {
				const uint32_t l = i ^ j;
				if (l > i) {
					const uint32_t x = s_v[i], y = s_v[l];
					if ((x > y) == ((i & k) == 0)) {
						s_v[i] = y;
						s_v[l] = x;
					}
				}
			}
			__syncthreads();
		}
	const DevPlan* P   = a.plan;
	bool		   bad = false;
	for (uint32_t m = tid; m < nb; m += kSmallBlock) {
		const uint32_t v = s_v[m] & 0x1FFFu;
		uint32_t	   d = 0, b0 = 0, n = nb, s0 = 0;
		if (a.bdoc) {
			d  = a.bdoc[v];
			b0 = a.br_start[d];
			n  = a.br_start[d + 1] - b0;
			s0 = a.st_start[d];
		}
		const uint32_t p = m - b0, q = p ^ 1u;
		int32_t		   partner = -1;
		if (q < n) {
			const uint32_t j  = s_v[b0 + q] & 0x1FFFu;
			const uint32_t ia = (p & 1u) ? j : v, ib = (p & 1u) ? v : j;
			const uint8_t  ca = a.bchar[ia], cb = a.bchar[ib];
			const bool ok = (P->br_flags[ca] & 1u) && (P->br_flags[cb] & 2u) && P->br_close_of[ca] == cb;
			if (ok)
				partner = (int32_t)j;
			else if (!(p & 1u)) {
				bad = true;
				if (a.doc_bad)
					a.doc_bad[d] = 1u;
			}
		}
		if (a.pair)
			a.pair[v] = partner >= 0 ? partner - (int32_t)b0 : -1;
		if (a.pair_pos)
			a.pair_pos[a.bsidx[v]] = partner >= 0 ? (int32_t)(a.bsidx[partner] - s0) : -1;
	}
	if (tid == 0)
		s_ok = 0;
	__syncthreads();
	if (bad)
		s_ok = 1;
	__syncthreads();
	if (tid == 0) {
		if (s_ok)
			atomicOr(&sum->mismatched, 1u);
		sum->small_done = 1u;
	}
}

template<int PASS>
hipError_t launch_pass(const PairArgs& a, unsigned grid, hipStream_t s) {
	hipLaunchKernelGGL((k_sort_hist<PASS>), dim3(grid), dim3(kSortBlock), 0, s, a);
	hipLaunchKernelGGL((k_scan_partial<PASS>), dim3(kScanGrid), dim3(kSortBlock), 0, s, a);
	hipLaunchKernelGGL((k_scan_top<PASS>), dim3(1), dim3(kSortBlock), 0, s, a);
	hipLaunchKernelGGL((k_scan_apply<PASS>), dim3(kScanGrid), dim3(kSortBlock), 0, s, a);
	hipLaunchKernelGGL((k_sort_scatter<PASS>), dim3(grid), dim3(kSortBlock), 0, s, a);
	return hipGetLastError();
}

}  // namespace

hipError_t launch_pairing(const PairArgs& a, hipStream_t s) {
	if (a.cap_br == 0)
		return hipSuccess;
	if (a.key_fill) {
		const uint64_t fc = (a.cap_br + kStkChunk - 1) / kStkChunk;
		hipLaunchKernelGGL(k_key_fill, dim3((unsigned)std::min<uint64_t>(fc, 2048)), dim3(256), 0, s, a);
	}
	if (a.bdoc) {  // per-document key bases first
		if (a.n_docs <= kKbaseOneBlock) {
			hipLaunchKernelGGL(k_kbase_one, dim3(1), dim3(1024), 0, s, a);
		} else {
			hipLaunchKernelGGL(k_kbase_partial, dim3(kScanGrid), dim3(kSortBlock), 0, s, a);
			hipLaunchKernelGGL(k_kbase_top, dim3(1), dim3(kSortBlock), 0, s, a);
			hipLaunchKernelGGL(k_kbase_apply, dim3(kScanGrid), dim3(kSortBlock), 0, s, a);
		}
	}
	if (a.small) {
		hipLaunchKernelGGL(k_pair_small, dim3(1), dim3(kSmallBlock), 0, s, a);
		if (a.small == 2)
			return hipGetLastError();
	}
	const uint64_t chunks = (a.cap_br + kChunk - 1) / kChunk;
	const unsigned grid	  = (unsigned)std::min<uint64_t>(std::max<uint64_t>(chunks, 1), 2048);
	hipLaunchKernelGGL(k_sort_prep, dim3(1), dim3(64), 0, s, a);
	hipError_t e;
	if ((e = launch_pass<0>(a, grid, s)) != hipSuccess)
		return e;
	if ((e = launch_pass<1>(a, grid, s)) != hipSuccess)
		return e;
	if ((e = launch_pass<2>(a, grid, s)) != hipSuccess)
		return e;
	if ((e = launch_pass<3>(a, grid, s)) != hipSuccess)
		return e;
	const unsigned pgrid = (unsigned)std::min<uint64_t>(std::max<uint64_t>((a.cap_br + 255) / 256, 1), 4096);
	hipLaunchKernelGGL(k_pair, dim3(pgrid), dim3(256), 0, s, a);
	return hipGetLastError();
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
	for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
		p[i] = v;
}

hipError_t launch_fill_i32(int32_t* p, uint64_t n, int32_t v, hipStream_t s) {
	if (!n)
		return hipSuccess;
	const uint64_t g = std::min<uint64_t>((n + 255) / 256, 4096);
	hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)g), dim3(256), 0, s, p, n, v);
	return hipGetLastError();
}

uint64_t pairing_max_chunks(uint64_t cap_br) {
	return (cap_br + kChunk - 1) / kChunk;
}

}  // namespace pars_hip
