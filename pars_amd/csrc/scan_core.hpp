// scan_core.hpp — the single-pass R1-R5 tile kernel, shared by every lane
// evaluator (general descriptor walk, JSON-style escape and quote-parity fast
// paths).
//
// Per tile (W rows of 256 lanes x 64 contiguous bytes):
//   1. the evaluator walks each lane's bytes once and yields, per descriptor,
//      the lane's state map, plus enough to produce the lane's masked bits for
//      any lane-entry state without a second walk;
//   2. block scans of the lane maps per row; decoupled look-back over tiles on
//      tagged 64-bit words gives the tile-entry states and output offsets;
//   3. masks -> unmasked structural / bracket bits; counts and the +-1 depth
//      delta; block scans give each lane's output offsets inside the tile;
//   4. row-local positions and bracket records staged in LDS per row and
//      written with coalesced stores (four consecutive records per thread:
//      positions and bracket positions as 16-byte stores, bracket bytes as
//      4-byte stores, each from one LDS read); the bitmap word per lane.
//
// Chains.  A plan with one descriptor (JSON, CSV, every single-automaton
// grammar) runs ONE chain: a tile publishes its state map together with its
// (structural, bracket, depth) counts for every possible tile-entry state, so
// a single look-back yields both the entry state and the output offsets.
// Plans with several descriptors have too many joint entry states for that and
// run two chains: states first, then counts.
//
// Replaces the reference's classify_kernel, desc_wordt64_kernel,
// thrust::inclusive_scan(TCompose64), desc_replay64_kernel,
// structural_out_kernel, popcount/exclusive_scan x2, pack_counts_kernel and
// scatter_kernel (include/pars/backend/cuda/pipeline.cuh:203-265, 292-317,
// 394-443, 503-511, 784-930) and their two blocking host syncs.
#pragma once

#include "dev_common.hpp"
#include "kernels.hpp"

#include <limits.h>
#include <type_traits>

namespace pars_hip {

// A look-back waits only on lower tiles, which were all dispatched earlier and
// publish before they wait, so it always ends; the spin bound is a guard so a
// defect can never hang the device: past it the kernel flags the summary and
// returns.
constexpr uint32_t kSpinLimit	= 1u << 20;
constexpr uint32_t kErrLookback = 0x80000000u;
// Predecessor records read per look-back round trip (the newest lanes of the
// wave) unless tuned; the nearest inclusive prefix is typically 10-20 tiles
// back.
constexpr int kLbWindow = 32;

__device__ __forceinline__ int lb_window(uint32_t w) {
	return w >= 8u && w <= 64u ? (int)w : kLbWindow;
}
// Bit 31 of an aggregate's map word (one chain) or structural-count word (two
// chains): the tile holds a document head (batch mode).
constexpr uint32_t kHeadFlag = 0x80000000u;

// 4x4 byte transpose: out[k] byte g = in[g] byte k.
__device__ __forceinline__ void transpose4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t& t0,
										   uint32_t& t1, uint32_t& t2, uint32_t& t3) {
	const uint32_t u0 = perm(a1, a0, 0x05010400u), u1 = perm(a1, a0, 0x07030602u);
	const uint32_t u2 = perm(a3, a2, 0x05010400u), u3 = perm(a3, a2, 0x07030602u);
	t0				  = perm(u2, u0, 0x05040100u);
	t1				  = perm(u2, u0, 0x07060302u);
	t2				  = perm(u3, u1, 0x05040100u);
	t3				  = perm(u3, u1, 0x07060302u);
}

__device__ __forceinline__ uint32_t popc64(uint64_t x) {
	return (uint32_t)__popcll(x);
}

// Set bits of x strictly below bit p (p < 64).
__device__ __forceinline__ uint32_t popc_below(uint64_t x, int p) {
	return popc64(x & ((1ull << p) - 1ull));
}

// v[s] for s < N, as an OR of masked values: a chain of selects over an
// array is turned into a dynamically indexed load from scratch memory.
template<int N>
__device__ __forceinline__ uint32_t pick(const uint32_t (&v)[N], uint32_t s) {
	uint32_t r = 0;
#pragma unroll
	for (int q = 0; q < N; ++q) r |= v[q] & (0u - (uint32_t)(s == (uint32_t)q));
	return r;
}

// ---------------------------------------------------------------------------
// Look-backs run in one full wave.  Lane L >= 64 - window reads the record
// of tile jb - 63 + L, so lane 63 holds the nearest predecessor and folding a
// window is a forward wave scan (lanes below the window are neutral); the
// newest lane holding an inclusive prefix (kp) ends the walk.  While a
// predecessor has not published, only the missing records are read again.
// ---------------------------------------------------------------------------
struct Window {
	int		 kp;	  // newest lane with an inclusive prefix, -1 if none
	uint64_t b_none;  // lanes whose record is not published yet
};

__device__ __forceinline__ Window window_extent(bool incl, bool none) {
	const uint64_t b_incl = __ballot(incl);
	Window		   w;
	w.kp	 = b_incl ? 63 - (int)__clzll((long long)b_incl) : -1;
	w.b_none = __ballot(none);
	return w;
}

// Every lane newer than kp must have published before the window can fold.
__device__ __forceinline__ bool window_ready(const Window& w) {
	const uint64_t newer = w.kp >= 63 ? 0ull : (~0ull << (w.kp + 1));
	return (w.b_none & newer) == 0;
}

// Two-chain mode, chain 1: packed per-descriptor states entering tile > 0.
template<int ND, bool WIDE>
__device__ uint32_t lookback_states(const uint64_t* recs, uint32_t tile, uint32_t epoch, uint32_t in_states,
									int win, DevSummary* sum) {
	using M	  = MapT<WIDE>;
	using Map = typename M::type;
	const int  lane	  = lane_id();
	const bool active = lane >= kWave - win;
	Map		   acc[ND];	 // windows folded so far: their entry -> the tile's entry
#pragma unroll
	for (int d = 0; d < ND; ++d) acc[d] = M::id();
	int64_t	 jb	   = (int64_t)tile - 1;
	uint32_t spins = 0;
	bool	 load  = true, incl = false, none = false;
	uint32_t ival  = 0;
	Map		 agg[ND];
	while (true) {
		const int64_t j = jb - 63 + lane;
		if (!active || j < 0) {	 // outside the window: neutral; before the input: the carry
			incl = active;
			none = false;
			ival = in_states;
#pragma unroll
			for (int d = 0; d < ND; ++d) agg[d] = M::id();
		} else if (load) {
			// one round trip: the inclusive and every aggregate word together
			const uint64_t* r = recs + (uint64_t)j * kRecWords;
			const uint64_t	w = peek(r + kRecSIncl);
			uint64_t		wa[ND];
#pragma unroll
			for (int d = 0; d < ND; ++d) wa[d] = peek(r + kRecSAgg + d);
			incl = word_is(w, epoch, kIncl);
			none = false;
			ival = (uint32_t)w;
#pragma unroll
			for (int d = 0; d < ND; ++d) {
				if (!word_is(wa[d], epoch, kAgg))
					none = !incl;
				agg[d] = M::unpack((uint32_t)wa[d]);
			}
		}
		const Window w = window_extent(incl, none);
		if (!window_ready(w)) {
			if (++spins > kSpinLimit) {
				if (lane == 0)
					atomicOr(&sum->overflow, kErrLookback);
				return in_states;
			}
			__builtin_amdgcn_s_sleep(1);
			load = none;  // only records still missing are read again
			continue;
		}
		Map wmap[ND];  // m_63 . m_62 . ... . m_{kp+1}
#pragma unroll
		for (int d = 0; d < ND; ++d)
			wmap[d] = M::readlane(wave_scan_maps<WIDE>(lane > w.kp ? agg[d] : M::id()), 63);
		if (w.kp >= 0) {
			const uint32_t S   = readlane(ival, w.kp);
			uint32_t	   out = 0;
#pragma unroll
			for (int d = 0; d < ND; ++d) {
				uint32_t sd = (S >> (3 * d)) & 7u;
				sd			= M::apply(acc[d], M::apply(wmap[d], sd));
				out |= sd << (3 * d);
			}
			return out;
		}
#pragma unroll
		for (int d = 0; d < ND; ++d) acc[d] = M::compose(acc[d], wmap[d]);
		jb -= win;
		load = true;
	}
}

// Two-chain mode, chain 2: (structural, bracket, depth) offsets entering tile > 0.
// A count aggregate whose structural word carries kHeadFlag holds a document
// head: its depth word is the depth after its last head, and older depths do
// not reach past it (structural and bracket offsets always do).
__device__ inline void lookback_counts(const uint64_t* recs, uint32_t tile, uint32_t epoch, const uint32_t base[3],
									   int win, uint32_t out[3], DevSummary* sum) {
	const int  lane	  = lane_id();
	const bool active = lane >= kWave - win;
	uint32_t   acc[3] = {0u, 0u, 0u};
	int64_t	   jb	  = (int64_t)tile - 1;
	uint32_t   spins  = 0;
	bool	   load = true, incl = false, none = false, head = false, closed = false;
	uint32_t   v[3] = {0u, 0u, 0u};
	while (true) {
		const int64_t j = jb - 63 + lane;
		if (!active || j < 0) {
			incl = active;
			none = false;
			head = false;
#pragma unroll
			for (int f = 0; f < 3; ++f) v[f] = active ? base[f] : 0u;
		} else if (load) {
			const uint64_t* r = recs + (uint64_t)j * kRecWords;
			uint64_t		wi[3], wa[3];
#pragma unroll
			for (int f = 0; f < 3; ++f) {
				wi[f] = peek(r + kRecCIncl + f);
				wa[f] = peek(r + kRecCAgg + f);
			}
			bool inc = true, agg = true;
#pragma unroll
			for (int f = 0; f < 3; ++f) {
				inc = inc && word_is(wi[f], epoch, kIncl);
				agg = agg && word_is(wa[f], epoch, kAgg);
			}
			incl = inc;
			none = !inc && !agg;
			head = !inc && (wa[0] & kHeadFlag);
#pragma unroll
			for (int f = 0; f < 3; ++f) v[f] = (uint32_t)(inc ? wi[f] : wa[f]);
			if (!inc)
				v[0] &= ~kHeadFlag;
		}
		const Window w = window_extent(incl, none);
		if (!window_ready(w)) {
			if (++spins > kSpinLimit) {
				if (lane == 0)
					atomicOr(&sum->overflow, kErrLookback);
				out[0] = base[0];
				out[1] = base[1];
				out[2] = base[2];
				return;
			}
			__builtin_amdgcn_s_sleep(1);
			load = none;
			continue;
		}
		// depth restarts at the newest document head (H) newer than the inclusive
		const uint64_t hb = __ballot(head && lane > w.kp);
		const int	   H  = hb ? 63 - (int)__clzll((long long)hb) : -1;
		acc[0] += wave_sum(lane >= w.kp ? v[0] : 0u);
		acc[1] += wave_sum(lane >= w.kp ? v[1] : 0u);
		const bool own_d = H >= 0 ? lane >= H : lane >= w.kp;
		const uint32_t wd = wave_sum(own_d ? v[2] : 0u);
		if (!closed)
			acc[2] += wd;
		closed = closed || hb;
		if (w.kp >= 0) {
			out[0] = acc[0];
			out[1] = acc[1];
			out[2] = acc[2];
			return;
		}
		jb -= win;
		load = true;
	}
}

// One-chain mode: the state entering tile > 0 (`st`) and the counts before it
// (`cnt`); a map word with kHeadFlag marks a tile holding a document head,
// whose depth counts are the depth after its last head.  A tile's aggregate is its map m plus v[f][e], its counts when
// entered in state e.  Folding the aggregates newer than an inclusive (S, C):
// lane L enters in s_L = (m_{L-1} . ... . m_{kp+1})(S) and adds v_L[f][s_L].
// A window without an inclusive is folded into `acc` as a function of its
// entry state (for every entry state) and the walk moves on.
template<bool WIDE, int NS>
__device__ void lookback_fused(const uint64_t* recs, uint32_t tile, uint32_t epoch, uint32_t in_states,
							   const uint32_t base[3], int win, uint32_t& st, uint32_t cnt[3], DevSummary* sum) {
	using M	  = MapT<WIDE>;
	using Map = typename M::type;
	const int  lane	  = lane_id();
	const bool active = lane >= kWave - win;
	uint32_t   accS[NS], accC[NS][3];  // entry state e of the folded windows -> (tile entry state, counts)
#pragma unroll
	for (int e = 0; e < NS; ++e) {
		accS[e]	   = (uint32_t)e;
		accC[e][0] = accC[e][1] = accC[e][2] = 0u;
	}
	int64_t	 jb	   = (int64_t)tile - 1;
	uint32_t spins = 0;
	bool	 load = true, incl = false, none = false, head = false, closed = false;
	uint32_t iS = 0, iC[3] = {0u, 0u, 0u};
	Map		 m = M::id();
	uint32_t v[3][NS];
	while (true) {
		const int64_t j = jb - 63 + lane;
		if (!active || j < 0) {	 // outside the window: neutral; before the input: the carry
			incl = active;
			none = false;
			head = false;
			iS	 = in_states;
			m	 = M::id();
#pragma unroll
			for (int f = 0; f < 3; ++f) {
				iC[f] = active ? base[f] : 0u;
#pragma unroll
				for (int e = 0; e < NS; ++e) v[f][e] = 0u;
			}
		} else if (load) {
			// one round trip: every word of the record
			const uint64_t* r = recs + (uint64_t)j * kRecWords;
			uint64_t		wi[4], wv[3 * NS];
#pragma unroll
			for (int q = 0; q < 4; ++q) wi[q] = peek(r + kRecFIncl + q);
			const uint64_t wm = peek(r + kRecFMap);
#pragma unroll
			for (int q = 0; q < 3 * NS; ++q) wv[q] = peek(r + kRecFCnt + q);
			bool inc = true, agg = word_is(wm, epoch, kAgg);
#pragma unroll
			for (int q = 0; q < 4; ++q) inc = inc && word_is(wi[q], epoch, kIncl);
#pragma unroll
			for (int q = 0; q < 3 * NS; ++q) agg = agg && word_is(wv[q], epoch, kAgg);
			incl  = inc;
			none  = !inc && !agg;
			head  = !inc && ((uint32_t)wm & kHeadFlag);
			iS	  = (uint32_t)wi[0];
			iC[0] = (uint32_t)wi[1];
			iC[1] = (uint32_t)wi[2];
			iC[2] = (uint32_t)wi[3];
			m	  = M::unpack((uint32_t)wm);
#pragma unroll
			for (int e = 0; e < NS; ++e)
#pragma unroll
				for (int f = 0; f < 3; ++f) v[f][e] = (uint32_t)wv[3 * e + f];
		}
		const Window w = window_extent(incl, none);
		if (!window_ready(w)) {
			if (++spins > kSpinLimit) {
				if (lane == 0)
					atomicOr(&sum->overflow, kErrLookback);
				st	   = in_states;
				cnt[0] = base[0];
				cnt[1] = base[1];
				cnt[2] = base[2];
				return;
			}
			__builtin_amdgcn_s_sleep(1);
			load = none;  // only records still missing are read again
			continue;
		}
		const bool	   agg_lane = lane > w.kp;
		const Map	   inc		= wave_scan_maps<WIDE>(agg_lane ? m : M::id());
		const Map	   exc		= wave_shift_maps<WIDE>(inc);
		const Map	   wmap		= M::readlane(inc, 63);
		// depth restarts at the newest document head (H) newer than the inclusive
		const uint64_t hb	 = __ballot(head && agg_lane);
		const int	   H	 = hb ? 63 - (int)__clzll((long long)hb) : -1;
		const bool	   own_d = H >= 0 ? lane >= H : agg_lane;
		if (w.kp >= 0) {
			const uint32_t S   = readlane(iS, w.kp) & 7u;
			const uint32_t s_l = M::apply(exc, S);
			const uint32_t so  = M::apply(wmap, S);
			st				   = pick(accS, so);
#pragma unroll
			for (int f = 0; f < 3; ++f) {
				uint32_t own;
				if (f < 2)
					own = agg_lane ? pick(v[f], s_l) : (lane == w.kp ? iC[f] : 0u);
				else
					own = own_d ? pick(v[f], s_l) : (lane == w.kp && H < 0 ? iC[f] : 0u);
				uint32_t a[NS];
#pragma unroll
				for (int e = 0; e < NS; ++e) a[e] = accC[e][f];
				const uint32_t ws = wave_sum(own);
				cnt[f]			  = (f == 2 && closed) ? pick(a, so) : ws + pick(a, so);
			}
			return;
		}
		// no inclusive in this window: fold its function of the entry state
		uint32_t nS[NS], nC[NS][3];
#pragma unroll
		for (int e = 0; e < NS; ++e) {
			const uint32_t s_l = M::apply(exc, (uint32_t)e);
			const uint32_t we  = M::apply(wmap, (uint32_t)e);
			nS[e]			   = pick(accS, we);
#pragma unroll
			for (int f = 0; f < 3; ++f) {
				uint32_t a[NS];
#pragma unroll
				for (int q = 0; q < NS; ++q) a[q] = accC[q][f];
				const bool	   take = f < 2 ? agg_lane : own_d;
				const uint32_t ws	= wave_sum(take ? pick(v[f], s_l) : 0u);
				nC[e][f]			= (f == 2 && closed) ? pick(a, we) : ws + pick(a, we);
			}
		}
#pragma unroll
		for (int e = 0; e < NS; ++e) {
			accS[e] = nS[e];
#pragma unroll
			for (int f = 0; f < 3; ++f) accC[e][f] = nC[e][f];
		}
		closed = closed || hb;
		jb -= win;
		load = true;
	}
}

// One-chain mode, single document: the chaser.  One wave of a block that draws
// no tiles follows the tiles' aggregates in order, 64 tiles per round trip,
// and publishes their inclusive words as soon as a prefix of them is ready -
// ahead of the tiles' own look-backs, which then find an inclusive close by
// instead of waiting for one to propagate tile by tile.  A tile publishes the
// same inclusive itself (the values are a function of the input), and no tile
// waits for the chaser, so it only ever shortens waits.  Lane L reads tile
// f + L; the window's ready prefix (every tile with an aggregate or an
// inclusive) is folded with a forward wave scan from its newest inclusive.
template<bool WIDE, int NS>
__device__ void chase_fused(uint64_t* recs, uint32_t n_tiles, uint32_t epoch, uint32_t in_states,
							const uint32_t base[3], DevSummary* sum) {
	using M	  = MapT<WIDE>;
	using Map = typename M::type;
	const int lane = lane_id();
	uint32_t  S	   = in_states & 7u;  // state and counts before tile f
	uint32_t  C[3] = {base[0], base[1], base[2]};
	uint32_t  f = 0, spins = 0;
	while (f < n_tiles) {
		const uint32_t t	= f + (uint32_t)lane;
		const bool	   live = t < n_tiles;
		bool		   incl = false, agg = false;
		uint32_t	   iS = 0, iC[3] = {0u, 0u, 0u};
		Map			   m = M::id();
		uint32_t	   v[3][NS];
#pragma unroll
		for (int e = 0; e < NS; ++e) v[0][e] = v[1][e] = v[2][e] = 0u;
		if (live) {
			const uint64_t* r = recs + (uint64_t)t * kRecWords;
			uint64_t		wi[4], wv[3 * NS];
#pragma unroll
			for (int q = 0; q < 4; ++q) wi[q] = peek(r + kRecFIncl + q);
			const uint64_t wm = peek(r + kRecFMap);
#pragma unroll
			for (int q = 0; q < 3 * NS; ++q) wv[q] = peek(r + kRecFCnt + q);
			incl = true;
#pragma unroll
			for (int q = 0; q < 4; ++q) incl = incl && word_is(wi[q], epoch, kIncl);
			agg = word_is(wm, epoch, kAgg);
#pragma unroll
			for (int q = 0; q < 3 * NS; ++q) agg = agg && word_is(wv[q], epoch, kAgg);
			iS	  = (uint32_t)wi[0];
			iC[0] = (uint32_t)wi[1];
			iC[1] = (uint32_t)wi[2];
			iC[2] = (uint32_t)wi[3];
			m	  = M::unpack((uint32_t)wm);
#pragma unroll
			for (int e = 0; e < NS; ++e)
#pragma unroll
				for (int g = 0; g < 3; ++g) v[g][e] = (uint32_t)wv[3 * e + g];
		}
		const uint64_t ready = __ballot(live && (incl || agg));
		const int	   P	 = ~ready ? (int)__builtin_ctzll(~ready) : 64;  // ready prefix
		if (P == 0) {
			if (++spins > kSpinLimit)
				return;	 // an accelerator only: the tiles finish without it
			__builtin_amdgcn_s_sleep(2);
			continue;
		}
		spins			  = 0;
		const bool	   in = lane < P;
		const uint64_t bi = __ballot(in && incl);
		const int	   kp = bi ? 63 - (int)__clzll((long long)bi) : -1;	 // newest inclusive
		uint32_t	   S0 = S, B[3] = {C[0], C[1], C[2]};
		if (kp >= 0) {
			S0	 = readlane(iS, kp) & 7u;
			B[0] = readlane(iC[0], kp);
			B[1] = readlane(iC[1], kp);
			B[2] = readlane(iC[2], kp);
		}
		const bool	   fold = in && lane > kp;	// tiles folded from their aggregates
		const Map	   inc	= wave_scan_maps<WIDE>(fold ? m : M::id());
		const Map	   exc	= wave_shift_maps<WIDE>(inc);
		const uint32_t s_l	= M::apply(exc, S0);
		uint32_t	   eS	= M::apply(inc, S0), eC[3];
#pragma unroll
		for (int g = 0; g < 3; ++g) eC[g] = B[g] + wave_incl_sum(fold ? pick(v[g], s_l) : 0u);
		if (fold) {
			uint64_t* r = recs + (uint64_t)t * kRecWords;
			publish(r + kRecFIncl, make_word(epoch, kIncl, eS));
#pragma unroll
			for (int g = 0; g < 3; ++g) publish(r + kRecFIncl + 1 + g, make_word(epoch, kIncl, eC[g]));
		}
		if (lane == kp) {
			eS = iS & 7u;
#pragma unroll
			for (int g = 0; g < 3; ++g) eC[g] = iC[g];
		}
		S = readlane(eS, P - 1) & 7u;
#pragma unroll
		for (int g = 0; g < 3; ++g) C[g] = readlane(eC[g], P - 1);
		f += (uint32_t)P;
	}
	(void)sum;
}

// ---------------------------------------------------------------------------
// The tile kernel.  Persistent: every block loads its LDS tables once, then
// draws tiles from an atomic ticket in dispatch order until none is left (so a
// tile only ever waits on tiles held by running blocks).  A tile is W rows of
// 16 KiB; in row r, lane l owns bytes [r*16K + l*64, +64), so every row's
// outputs are one contiguous run: they are staged in LDS (row-local 16-bit
// offsets) and written back with coalesced stores (scattered 4-byte stores
// amplify HBM writes several-fold).
//
// BATCH: the input is a batch of documents (ScanArgs::b).  A lane belongs to
// one document (documents start on lane boundaries); at a head lane the state
// maps become constant maps to the document's start-state image, so ordinary
// map scans and look-backs restart there, and depth is scanned per document.
// Positions and bracket keys come out document-relative; output arrays stay
// compacted across the batch.
//
// E (the lane evaluator) provides:
//   kND, kWide, Map        descriptors and their map type
//   kBitParallel           a bit-parallel lane evaluator (no per-byte state walk)
//   Lds, load(Lds&, plan)  LDS tables (cooperative load, all threads)
//   walk(Lds, plan, x, valid)   one pass over a lane's 64 bytes
//   T[d]                    the lane's map after walk
//   masked(s0[])            masked bits for lane-entry states s0
//   masked_all(m[])         (one descriptor) masked bits for every entry state
//   classes(S, OC, OP)      structural / bracket / open bits
// ---------------------------------------------------------------------------
constexpr uint32_t kRowBytes = kBlock * kBPL;  // 16 KiB
// Outputs of one row staged in LDS before the coalesced write-back (a denser
// row stores directly, lane by lane).  Single documents at three or four rows
// per tile stage denser rows (small-record JSON reaches ~4900 structural and
// ~1700 bracket bytes per 16 KiB row): the extra 7.5 KB still leaves room for
// the three blocks per CU their registers allow.  The other shapes keep the
// block count their registers allow with the smaller staging.
template<int W, bool BATCH>
constexpr int kCapP = (!BATCH && W >= 3) ? 6144 : 4096;  // staged structural outputs per row
template<int W, bool BATCH>
constexpr int kCapB = (!BATCH && W >= 3) ? 2048 : 1536;  // staged bracket outputs per row

// Waves per SIMD the register allocation must leave room for.  Many-descriptor
// plans with wide descriptors need most of the register file; they are held to
// two rather than left to fall to one.  The bit-parallel evaluators at three
// and four rows per tile fit three (four rows: 168 registers, a few spilled)
// and run 15 % faster there than at two, where the look-back and scatter
// latencies stay exposed.
template<class E, int W>
constexpr int kMinWaves = (E::kND >= 3 && E::kWide) ? 2 : (E::kBitParallel && W >= 3) ? 3 : 1;

template<class E, int W, bool BATCH, int NWV = kBlock / kWave>
__global__ __launch_bounds__(NWV * kWave) __attribute__((amdgpu_waves_per_eu(kMinWaves<E, W>))) void k_scan(ScanArgs a) {
	constexpr int	   NT	= NWV * kWave;	// threads per block
	constexpr uint32_t RowB = NT * kBPL;	// bytes per row
	constexpr int  ND	 = E::kND;
	constexpr int  NDA	 = ND > 0 ? ND : 1;
	constexpr bool FUSED = ND == 1;
	constexpr int  NS	 = E::kWide ? 8 : 4;  // entry states a one-descriptor tile is summarised for
	constexpr int  NPK	 = NS / 4;			  // count dwords per field, four states each
	constexpr int  NL	 = W * NT;			  // lanes per tile
	constexpr int  NHW	 = NL / 64;			  // head-bitmap words per tile
	using M				 = MapT<E::kWide>;
	using Map			 = typename M::type;
	using KeyT			 = typename std::conditional<BATCH, int32_t, int16_t>::type;
	constexpr uint64_t kTileBytes = (uint64_t)W * RowB;
	constexpr int	   CAPP		  = kCapP<W, BATCH> * NWV / 4;	// staging scales with the row
	constexpr int	   CAPB		  = kCapB<W, BATCH> * NWV / 4;

	__shared__ typename E::Lds s_lds;
	__shared__ Map			   s_wagg[W][NWV][NDA];
	__shared__ uint32_t		   s_misc[8];
	__shared__ uint32_t		   s_wcnt[W][NWV][3];
	__shared__ uint32_t		   s_vsum[NWV][3][2 * NPK];	// one-chain: per-wave 16-bit-pair count sums
	__shared__ int32_t		   s_vpost[NWV];		   // one-chain batch: per-wave depth after the last head
	__shared__ uint32_t		   s_v[3][NS];			   // one-chain: tile counts per entry state
	__shared__ int32_t		   s_mm[NWV][2];
	__shared__ __align__(16) uint16_t s_pos[CAPP];	 // row-local byte offset
	__shared__ __align__(16) uint16_t s_bpos[CAPB];	 // row-local byte offset
	__shared__ uint16_t		   s_bsidx[CAPB];	 // row-local structural index
	__shared__ KeyT			   s_bkey[CAPB];	 // key: relative to the row's entry depth / document-relative
	// bit-parallel evaluators: each lane's bracket ordinal bits per row (from the
	// walk) and the staged bracket bytes rebuilt from them
	__shared__ uint64_t		   s_bo[E::kBitParallel ? W : 1][E::kBitParallel ? NT : 1];
	__shared__ __align__(16) uint8_t s_bch[E::kBitParallel ? CAPB : 1];
	__shared__ uint64_t		   s_hw[BATCH ? NHW : 1];		// the tile's head bits
	__shared__ uint32_t		   s_hpre[BATCH ? NHW : 1];		// heads in the words before
	__shared__ int32_t		   s_dhead[BATCH ? NL : 1];		// tile-relative depth at each head, by slot
	__shared__ uint32_t		   s_lpos0[BATCH ? NL : 1];		// document-local offset of each lane's first byte

	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	a.epoch		  = a.sum->epoch;  // set by the summary reset that precedes every launch
	if (a.carry) {	// a chained chunk: where the previous one left off
		a.in_states	 = a.carry->states;
		a.out_base	 = a.carry->out_base;
		a.br_base	 = a.carry->br_base;
		a.depth_base = a.carry->depth;
	}
	E::load(s_lds, a.plan);
	const bool aligned	 = ((uintptr_t)a.data & 15u) == 0;
	const bool want_sidx = a.bsidx != nullptr;
	const bool all_keys	 = BATCH || !a.key_sparse;
	// bracket bytes from the walk's ordinal bits instead of the input
	const bool	   bit_chr = E::kBitParallel && a.plan->br_bit_ok;
	const uint32_t brc	   = (uint32_t)a.plan->br_chr[0][0] | ((uint32_t)a.plan->br_chr[0][1] << 8)
						 | ((uint32_t)a.plan->br_chr[1][0] << 16) | ((uint32_t)a.plan->br_chr[1][1] << 24);

	// batch: lane L of the tile (row-major) -> its document slot (document d0 + slot)
	auto slot_of = [&](int L) -> uint32_t {
		const uint64_t hw = s_hw[L >> 6];
		return s_hpre[L >> 6] + popc64(hw & (~0ull >> (63 - (L & 63)))) - (uint32_t)(s_hw[0] & 1ull);
	};

	// One chain, one document: the grid's last block draws no tiles and runs
	// the chaser instead when the launch asks for it.
	bool chaser = false;
	if constexpr (FUSED && !BATCH) {
		chaser = a.chase && blockIdx.x == gridDim.x - 1;
		if (chaser && wid == 0) {
			const uint32_t base[3] = {a.out_base, a.br_base, (uint32_t)a.depth_base};
			chase_fused<E::kWide, NS>(a.recs, a.n_tiles, a.epoch, a.in_states, base, a.sum);
		}
	}

	// A block takes its next tile only when it is ready to start it: a ticket
	// drawn early leaves a hole that every later tile's look-back waits on.
	while (!chaser) {
		__syncthreads();  // LDS of the previous tile is free
		if (tid == 0)
			s_misc[0] = atomicAdd(a.ticket, 1u);
		__syncthreads();
		const uint32_t tile = s_misc[0];
		if (tile >= a.n_tiles)
			break;
		const uint64_t tbase = (uint64_t)tile * kTileBytes;
		uint64_t*	   rec	 = a.recs + (uint64_t)tile * kRecWords;

		// ---- 0. batch: documents of the tile's lanes ----
		int32_t	 d0 = 0;
		int32_t	 ldoc[W];		  // document of the lane (-1: none)
		bool	 lhead[W];		  // the lane starts its document
		bool	 llast[W];		  // the lane holds its document's last byte
		uint32_t vdoc[W];		  // the lane's bytes inside its document
		int		 last_head = -1;  // last head lane of the tile
		if constexpr (BATCH) {
			if (tid < NHW) {
				const uint64_t wi = (uint64_t)tile * NHW + tid;
				s_hw[tid]		  = wi < a.b.n_head_words ? a.b.head_bits[wi] : 0ull;
			}
			if (tid == 0)
				s_misc[6] = (uint32_t)a.b.tile_doc[tile];
			__syncthreads();
			if (wid == 0) {	 // exclusive prefix of the per-word head counts
				const uint32_t c   = lane < NHW ? popc64(s_hw[lane]) : 0u;
				const uint32_t inc = wave_incl_sum(c);
				if (lane < NHW)
					s_hpre[lane] = inc - c;
			}
			__syncthreads();
			d0 = (int32_t)s_misc[6];
#pragma unroll
			for (int q = NHW - 1; q >= 0; --q)
				if (last_head < 0 && s_hw[q])
					last_head = q * 64 + 63 - (int)__clzll((long long)s_hw[q]);
#pragma unroll
			for (int r = 0; r < W; ++r) {
				const int	   L	  = r * NT + tid;
				const uint64_t lstart = tbase + (uint64_t)r * RowB + (uint64_t)tid * kBPL;
				lhead[r]			  = (s_hw[L >> 6] >> (L & 63)) & 1ull;
				ldoc[r]				  = d0 + (int32_t)slot_of(L);
				vdoc[r]				  = 0;
				llast[r]			  = false;
				if (ldoc[r] >= 0 && (uint32_t)ldoc[r] < a.b.n_docs) {
					const uint64_t off = a.b.doc_off[ldoc[r]], end = off + a.b.doc_len[ldoc[r]];
					if (lstart < end) {
						vdoc[r]	 = (uint32_t)min<uint64_t>((uint64_t)kBPL, end - lstart);
						llast[r] = lstart + kBPL >= end;
					}
					s_lpos0[L] = (uint32_t)(lstart - off);
				} else {
					ldoc[r] = -1;
				}
			}
		}

		// ---- 1. one walk per row ----
		// Every row's loads are issued before the first walk.
		E		 ev[W];
		uint32_t valid[W];
		uint32_t xb[W][16];
		auto	 load_row = [&](int r, uint32_t (&x)[16]) {
			const uint64_t base = tbase + (uint64_t)r * RowB + (uint64_t)tid * kBPL;
			valid[r]			= base >= a.n ? 0u : (uint32_t)min<uint64_t>((uint64_t)kBPL, a.n - base);
			if (valid[r] == kBPL && aligned) {
				const uint4* p = reinterpret_cast<const uint4*>(a.data + base);
#pragma unroll
				for (int k = 0; k < 4; ++k) {
					const uint4 v = p[k];
					x[4 * k + 0]  = v.x;
					x[4 * k + 1]  = v.y;
					x[4 * k + 2]  = v.z;
					x[4 * k + 3]  = v.w;
				}
			} else {
#pragma unroll
				for (int k = 0; k < 16; ++k) {
					uint32_t w = 0;
#pragma unroll
					for (int b = 0; b < 4; ++b)
						if ((uint32_t)(4 * k + b) < valid[r])
							w |= (uint32_t)a.data[base + 4 * k + b] << (8 * b);
					x[k] = w;
				}
			}
			if constexpr (BATCH)
				valid[r] = min(valid[r], vdoc[r]);
		};
#pragma unroll
		for (int r = 0; r < W; ++r) load_row(r, xb[r]);
#pragma unroll
		for (int r = 0; r < W; ++r) {
			ev[r].walk(s_lds, a.plan, xb[r], valid[r]);
			if constexpr (E::kBitParallel)
				s_bo[r][tid] = ev[r].BO;  // read back by this thread only, at the scatter
			if constexpr (BATCH && ND > 0)
				if (lhead[r])  // restart: whatever came before, the lane enters in the start states
#pragma unroll
					for (int d = 0; d < ND; ++d)
						ev[r].T[d] = M::constant(M::apply(ev[r].T[d], (a.in_states >> (3 * d)) & 7u));
		}

		// ---- 2. entry states (and, one-chain, output offsets) ----
		uint32_t	   tile_in	= a.in_states;
		const uint32_t cbase[3] = {a.out_base, a.br_base, (uint32_t)a.depth_base};
		uint32_t	   cex[3]	= {cbase[0], cbase[1], cbase[2]};
		Map			   lpre[W][NDA];
		uint32_t	   row_in[W];
		uint32_t	   post = 0;  // batch: depth after the tile's last head
		if constexpr (ND > 0) {
			Map exc[W][ND];
#pragma unroll
			for (int r = 0; r < W; ++r)
#pragma unroll
				for (int d = 0; d < ND; ++d) {
					const Map inc = wave_scan_maps<E::kWide>(ev[r].T[d]);
					exc[r][d]	  = wave_shift_maps<E::kWide>(inc);
					if (lane == 63)
						s_wagg[r][wid][d] = inc;
				}
			__syncthreads();
			Map ragg[W][ND], tagg[ND];
#pragma unroll
			for (int d = 0; d < ND; ++d) {
				tagg[d] = M::id();
#pragma unroll
				for (int r = 0; r < W; ++r) {
					Map wp = M::id();
#pragma unroll
					for (int w2 = 0; w2 < NWV; ++w2) {
						if (w2 == wid)
							lpre[r][d] = M::compose(exc[r][d], wp);
						wp = M::compose(s_wagg[r][w2][d], wp);
					}
					ragg[r][d] = wp;
					tagg[d]	   = M::compose(wp, tagg[d]);
				}
			}
			if constexpr (BATCH)
#pragma unroll
				for (int r = 0; r < W; ++r)
					if (lhead[r])
#pragma unroll
						for (int d = 0; d < ND; ++d) lpre[r][d] = M::constant((a.in_states >> (3 * d)) & 7u);

			if constexpr (FUSED) {
				// The lane's counts for every lane-entry state, one byte per state
				// (depth delta biased by +64), gathered per tile-entry state through
				// the lane's entry map, then summed as 16-bit pairs over rows and
				// lanes (at most 4 x 128 x 64 = 32768 per field per wave).
				uint32_t acc16[3][2 * NPK];
				int32_t	 dpost = 0;	 // batch: depth deltas from the tile's last head on
#pragma unroll
				for (int f = 0; f < 3; ++f)
#pragma unroll
					for (int q = 0; q < 2 * NPK; ++q) acc16[f][q] = 0u;
				Map rowpre = M::id();  // tile entry -> row entry
#pragma unroll
				for (int r = 0; r < W; ++r) {
					uint64_t ms[NS];
					ev[r].masked_all(ms);
					uint64_t S, OC, OP;
					ev[r].classes(S, OC, OP);
					uint32_t pk[3][NPK];
#pragma unroll
					for (int f = 0; f < 3; ++f)
#pragma unroll
						for (int h = 0; h < NPK; ++h) pk[f][h] = 0u;
#pragma unroll
					for (int s = 0; s < NS; ++s) {
						const uint32_t ns = popc64(S & ~ms[s]), nb = popc64(OC & ~ms[s]);
						const uint32_t dd = 2u * popc64(OP & ~ms[s]) - nb + 64u;
						pk[0][s >> 2] |= ns << (8 * (s & 3));
						pk[1][s >> 2] |= nb << (8 * (s & 3));
						pk[2][s >> 2] |= dd << (8 * (s & 3));
					}
					const Map lmap = M::compose(lpre[r][0], rowpre);  // tile entry -> lane entry
#pragma unroll
					for (int f = 0; f < 3; ++f) {
						uint32_t g[NPK];
						if constexpr (E::kWide) {
							g[0]	   = perm(pk[f][NPK - 1], pk[f][0], lmap.lo);
							g[NPK - 1] = perm(pk[f][NPK - 1], pk[f][0], lmap.hi);
						} else {
							g[0] = perm(0u, pk[f][0], lmap.lo);
						}
#pragma unroll
						for (int h = 0; h < NPK; ++h) {
							acc16[f][2 * h] += g[h] & 0x00FF00FFu;
							acc16[f][2 * h + 1] += (g[h] >> 8) & 0x00FF00FFu;
						}
						if constexpr (BATCH)
							if (f == 2 && r * NT + tid >= last_head && last_head >= 0)
								dpost += (int32_t)(g[0] & 0xFFu) - 64;	// entry-independent after a head
					}
					rowpre = M::compose(ragg[r][0], rowpre);
				}
#pragma unroll
				for (int f = 0; f < 3; ++f)
#pragma unroll
					for (int q = 0; q < 2 * NPK; ++q) {
						const uint32_t t = wave_sum(acc16[f][q]);
						if (lane == 0)
							s_vsum[wid][f][q] = t;
					}
				if constexpr (BATCH) {
					const uint32_t t = wave_sum((uint32_t)dpost);
					if (lane == 0)
						s_vpost[wid] = (int32_t)t;
				}
				__syncthreads();
				if constexpr (BATCH)
					for (int w2 = 0; w2 < NWV; ++w2) post += (uint32_t)s_vpost[w2];
				if (tid < 3 * NS) {	 // thread (f, e): field e of the four wave sums
					const int f = tid / NS, e = tid % NS;
					const int q = 2 * (e >> 2) + (e & 1), sh = (e & 2) ? 16 : 0;
					uint32_t  t = 0;
#pragma unroll
					for (int w2 = 0; w2 < NWV; ++w2) t += (s_vsum[w2][f][q] >> sh) & 0xFFFFu;
					if (f == 2)
						t -= 64u * (uint32_t)(NT * W);
					if (BATCH && f == 2 && last_head >= 0)
						t = post;
					s_v[f][e] = t;
					if (tile != 0)
						publish(rec + kRecFCnt + 3 * e + f, make_word(a.epoch, kAgg, t));
				}
				if (tile != 0) {
					if (tid == 0)
						publish(rec + kRecFMap, make_word(a.epoch, kAgg,
														  M::pack(tagg[0]) | (last_head >= 0 ? kHeadFlag : 0u)));
					if (wid == 0) {
						uint32_t st, o[3];
						lookback_fused<E::kWide, NS>(a.recs, tile, a.epoch, a.in_states, cbase, lb_window(a.lb_window), st,
													 o, a.sum);
						if (lane == 0) {
							s_misc[1] = st;
							s_misc[2] = o[0];
							s_misc[3] = o[1];
							s_misc[4] = o[2];
						}
					}
				}
				__syncthreads();
				if (tile != 0) {
					tile_in = s_misc[1];
					cex[0]	= s_misc[2];
					cex[1]	= s_misc[3];
					cex[2]	= s_misc[4];
				}
			} else if (tile != 0) {
				if (tid == 0)
#pragma unroll
					for (int d = 0; d < ND; ++d)
						publish(rec + kRecSAgg + d, make_word(a.epoch, kAgg, M::pack(tagg[d])));
				if (wid == 0) {
					const uint32_t st =
						lookback_states<ND, E::kWide>(a.recs, tile, a.epoch, a.in_states, lb_window(a.lb_window), a.sum);
					if (lane == 0)
						s_misc[1] = st;
				}
				__syncthreads();
				tile_in = s_misc[1];
			}
			uint32_t cur = tile_in;
#pragma unroll
			for (int r = 0; r < W; ++r) {
				row_in[r]	 = cur;
				uint32_t nxt = 0;
#pragma unroll
				for (int d = 0; d < ND; ++d) nxt |= M::apply(ragg[r][d], (cur >> (3 * d)) & 7u) << (3 * d);
				cur = nxt;
			}
			if (tid == 0) {
				if constexpr (FUSED) {
					const uint32_t e = tile_in & 7u;
					publish(rec + kRecFIncl, make_word(a.epoch, kIncl, cur));
#pragma unroll
					for (int f = 0; f < 3; ++f) {
						const uint32_t v = (BATCH && f == 2 && last_head >= 0) ? post : cex[f] + s_v[f][e];
						publish(rec + kRecFIncl + 1 + f, make_word(a.epoch, kIncl, v));
					}
				} else {
					publish(rec + kRecSIncl, make_word(a.epoch, kIncl, cur));
				}
				if (tile == a.n_tiles - 1)
					a.sum->final_states = cur;
			}
		} else {
#pragma unroll
			for (int r = 0; r < W; ++r) row_in[r] = 0;
			if (tid == 0 && tile == a.n_tiles - 1)
				a.sum->final_states = 0;
		}

		// ---- 3. masks and in-tile offsets ----
		uint64_t sm[W], om[W], opm[W];
		uint32_t cnt[W][3];
		uint32_t lexit[W];	// batch: packed states after the lane
#pragma unroll
		for (int r = 0; r < W; ++r) {
			uint32_t s0[NDA];
#pragma unroll
			for (int d = 0; d < ND; ++d) s0[d] = M::apply(lpre[r][d], (row_in[r] >> (3 * d)) & 7u);
			if constexpr (BATCH) {
				lexit[r] = 0;
#pragma unroll
				for (int d = 0; d < ND; ++d) lexit[r] |= M::apply(ev[r].T[d], s0[d]) << (3 * d);
			}
			const uint64_t masked = ND > 0 ? ev[r].masked(s_lds, s0) : 0ull;
			uint64_t	   S, OC, OP;
			ev[r].classes(S, OC, OP);
			sm[r]  = S & ~masked;
			om[r]  = OC & ~masked;
			opm[r] = OP & ~masked;
			if (a.bitmap && (tbase + (uint64_t)r * RowB + (uint64_t)tid * kBPL) < a.n)
				a.bitmap[(tbase + (uint64_t)r * RowB) / 64 + tid] = sm[r];
			cnt[r][0] = popc64(sm[r]);
			cnt[r][1] = popc64(om[r]);
			cnt[r][2] = 2u * popc64(opm[r]) - cnt[r][1];
		}
		uint32_t cpre[W][3], rtot[W][3], ctot[3] = {0u, 0u, 0u};
		{
			uint32_t cinc[W][3];
#pragma unroll
			for (int r = 0; r < W; ++r)
#pragma unroll
				for (int f = 0; f < 3; ++f) {
					cinc[r][f] = wave_incl_sum(cnt[r][f]);
					if (lane == 63)
						s_wcnt[r][wid][f] = cinc[r][f];
				}
			__syncthreads();
#pragma unroll
			for (int r = 0; r < W; ++r)
#pragma unroll
				for (int f = 0; f < 3; ++f) {
					uint32_t p = 0, t = 0;
#pragma unroll
					for (int w2 = 0; w2 < NWV; ++w2) {
						const uint32_t v = s_wcnt[r][w2][f];
						if (w2 < wid)
							p += v;
						t += v;
					}
					cpre[r][f] = p + cinc[r][f] - cnt[r][f];
					rtot[r][f] = readfirstlane(t);	// block-uniform
					ctot[f] += rtot[r][f];
				}
		}
		// batch: tile-relative depth at each head; the depth after the last head
		if constexpr (BATCH) {
			int32_t rd = 0;
#pragma unroll
			for (int r = 0; r < W; ++r) {
				if (lhead[r])
					s_dhead[ldoc[r] - d0] = rd + (int32_t)cpre[r][2];
				rd += (int32_t)rtot[r][2];
			}
			__syncthreads();
			if constexpr (!FUSED)
				if (last_head >= 0)
					post = (uint32_t)((int32_t)ctot[2] - s_dhead[slot_of(last_head)]);
		}

		// two-chain mode, chain 2: tile output offsets
		if constexpr (!FUSED) {
			const uint32_t dtot = (BATCH && last_head >= 0) ? post : ctot[2];
			if (tile != 0) {
				if (tid == 0) {
					publish(rec + kRecCAgg + 0,
							make_word(a.epoch, kAgg, ctot[0] | (BATCH && last_head >= 0 ? kHeadFlag : 0u)));
					publish(rec + kRecCAgg + 1, make_word(a.epoch, kAgg, ctot[1]));
					publish(rec + kRecCAgg + 2, make_word(a.epoch, kAgg, dtot));
				}
				if (wid == 0) {
					uint32_t o[3];
					lookback_counts(a.recs, tile, a.epoch, cbase, lb_window(a.lb_window), o, a.sum);
					if (lane == 0) {
						s_misc[2] = o[0];
						s_misc[3] = o[1];
						s_misc[4] = o[2];
					}
				}
				__syncthreads();
				cex[0] = s_misc[2];
				cex[1] = s_misc[3];
				cex[2] = s_misc[4];
			}
			if (tid == 0) {
				publish(rec + kRecCIncl + 0, make_word(a.epoch, kIncl, cex[0] + ctot[0]));
				publish(rec + kRecCIncl + 1, make_word(a.epoch, kIncl, cex[1] + ctot[1]));
				publish(rec + kRecCIncl + 2,
						make_word(a.epoch, kIncl, (BATCH && last_head >= 0) ? post : cex[2] + ctot[2]));
			}
		}
		if (tid == 0 && tile == a.n_tiles - 1) {
			a.sum->n_struct	  = cex[0] + ctot[0];
			a.sum->n_brackets = cex[1] + ctot[1];
			a.sum->last_bal	  = (int32_t)(cex[2] + ctot[2]);
			if constexpr (BATCH) {
				a.b.st_start[a.b.n_docs] = cex[0] + ctot[0];
				a.b.br_start[a.b.n_docs] = cex[1] + ctot[1];
			}
		}
		if (a.pair_pos)	 // -1 for every structural index of the tile; R6 overwrites brackets
			for (uint32_t i = tid; i < ctot[0]; i += NT)
				if ((uint64_t)cex[0] + i < a.cap_struct)
					a.pair_pos[cex[0] + i] = -1;

		// ---- 4. scatter, row by row ----
		// Depths are relative to the tile's entry depth until the end; in batch
		// mode each lane adds its document's offset (doc-relative depths).
		uint32_t rb[3] = {cex[0], cex[1], cex[2]};
		int32_t	 rowd = 0;	// the row's entry depth
		int32_t	 lmin = INT_MAX, lmax = INT_MIN;
#pragma unroll
		for (int r = 0; r < W; ++r) {
			const uint64_t rowb	 = tbase + (uint64_t)r * RowB;	 // data index of the row
			const uint32_t loff	 = (uint32_t)tid * kBPL;			 // lane's offset in the row
			const uint64_t pbase = a.byte_base + rowb;
			int32_t		   dl	 = rowd + (int32_t)cpre[r][2];	// the lane's entry depth
			if constexpr (BATCH) {
				if (ldoc[r] >= 0) {
					const uint32_t sl	  = (uint32_t)(ldoc[r] - d0);
					const bool	   inside = sl > 0 || (s_hw[0] & 1ull);	// the document's head is in the tile
					dl += inside ? -s_dhead[sl] : (int32_t)cex[2];
					if (lhead[r]) {
						a.b.st_start[ldoc[r]] = rb[0] + cpre[r][0];
						a.b.br_start[ldoc[r]] = rb[1] + cpre[r][1];
					}
					if (llast[r]) {
						const int32_t fin	   = dl + (int32_t)cnt[r][2];
						a.b.doc_final[ldoc[r]] = lexit[r];
						a.b.doc_last[ldoc[r]]  = fin;
						if (fin != 0)  // the batch cannot pair by stack matching
							atomicOr(&a.sum->doc_unbal, 1u);
					}
				}
			}
			int32_t		   rmin = INT_MAX, rmax = INT_MIN;	// this lane, this row
			const bool	   staged = rtot[r][0] <= (uint32_t)CAPP && rtot[r][1] <= (uint32_t)CAPB;
			if (staged) {  // block-uniform
				{  // both 32-bit halves of the mask at once: half the iterations
					uint32_t mlo = (uint32_t)sm[r], mhi = (uint32_t)(sm[r] >> 32);
					uint32_t ilo = cpre[r][0], ihi = ilo + (uint32_t)__popc(mlo);
					while (mlo | mhi) {
						if (mlo) {
							s_pos[ilo++] = (uint16_t)(loff + (uint32_t)__builtin_ctz(mlo));
							mlo &= mlo - 1;
						}
						if (mhi) {
							s_pos[ihi++] = (uint16_t)(loff + 32u + (uint32_t)__builtin_ctz(mhi));
							mhi &= mhi - 1;
						}
					}
				}
				uint64_t bm = om[r];
				uint32_t lb = cpre[r][1];
				int32_t	 bal = dl;	// brackets come in order: a running depth
				uint64_t bo = 0;
				if constexpr (E::kBitParallel)
					bo = s_bo[r][tid];
				while (bm) {
					const int	  p	   = __builtin_ctzll(bm);
					const bool	  op   = (opm[r] >> p) & 1ull;
					const int32_t bal0 = bal;
					bal += op ? 1 : -1;
					rmin		= min(rmin, bal);
					rmax		= max(rmax, bal);
					s_bpos[lb] = (uint16_t)(loff + (uint32_t)p);
					if (want_sidx)
						s_bsidx[lb] = (uint16_t)(cpre[r][0] + popc_below(sm[r], p));
					s_bkey[lb] = (KeyT)((op ? bal0 : bal) - (BATCH ? 0 : rowd));
					if constexpr (E::kBitParallel)
						s_bch[lb] = (uint8_t)(brc >> (8 * ((op ? 2 : 0) + (int)((bo >> p) & 1ull))));
					++lb;
					bm &= bm - 1;
				}
				__syncthreads();
				const uint32_t np = rtot[r][0], nb = rtot[r][1];
				const uint64_t lim_p = a.cap_struct > rb[0] ? a.cap_struct - rb[0] : 0;
				const uint64_t lim_b = a.cap_br > rb[1] ? a.cap_br - rb[1] : 0;
				const uint32_t mp = (uint32_t)min<uint64_t>(np, lim_p), mb = (uint32_t)min<uint64_t>(nb, lim_b);
				// input bytes for chars are gathered four per thread at once, so
				// their latencies overlap
				constexpr int U = 4;
				auto		  out_pos = [&](uint32_t off) -> uint32_t {
					 if constexpr (BATCH)
						 return s_lpos0[r * NT + (off >> 6)] + (off & 63u);
					 else
						 return (uint32_t)(pbase + off);
				};
				if (!a.chars) {
					// four consecutive positions per thread: one 8-byte LDS read, one
					// 16-byte store (dword-aligned: the output need not be 16-byte
					// aligned); a tail one by one
					const uint64_t g0 = rb[0];
					for (uint32_t i = 4u * (uint32_t)tid; i + 4u <= mp; i += 4u * NT) {
						const uint2	   p2 = *reinterpret_cast<const uint2*>(s_pos + i);
						const uint32_t v[4] = {out_pos(p2.x & 0xFFFFu), out_pos(p2.x >> 16), out_pos(p2.y & 0xFFFFu),
											   out_pos(p2.y >> 16)};
						__builtin_memcpy(a.positions + g0 + i, v, 16);
					}
					const uint32_t body = mp & ~3u;
					if ((uint32_t)tid < mp - body)
						a.positions[g0 + body + tid] = out_pos(s_pos[body + tid]);
				} else
				for (uint32_t i0 = 0; i0 < mp; i0 += U * NT) {
					uint8_t ch[U];
					if (a.chars)
#pragma unroll
						for (int u = 0; u < U; ++u) {
							const uint32_t i = i0 + u * NT + tid;
							ch[u]			 = i < mp ? a.data[rowb + s_pos[i]] : 0;
						}
#pragma unroll
					for (int u = 0; u < U; ++u) {
						const uint32_t i = i0 + u * NT + tid;
						if (i < mp) {
							a.positions[rb[0] + i] = out_pos(s_pos[i]);
							if (a.chars)
								a.chars[rb[0] + i] = ch[u];
						}
					}
				}
				const int32_t kb = BATCH ? 0 : (int32_t)rb[2];
				bool		  bdone = false;
				if constexpr (E::kBitParallel) {
					if (bit_chr && !want_sidx) {
						// four consecutive bracket records per thread: positions from one
						// 8-byte LDS read in one 16-byte store, bytes from one 4-byte LDS
						// read in one 4-byte store (both only element-aligned in the
						// output), keys only at the stack chunks' starts; a tail one by one
						const uint64_t g0  = rb[1];
						auto		   one = [&](uint32_t i) {
							  const uint64_t g = g0 + i;
							  a.bpos[g]		   = out_pos(s_bpos[i]);
							  if (all_keys || (g & (kStkChunk - 1u)) == 0)
								  a.bkey[g] = kb + (int32_t)s_bkey[i];
							  a.bchar[g] = s_bch[i];
							  if constexpr (BATCH)
								  a.b.bdoc[g] = (uint32_t)d0 + slot_of(r * NT + (int)(s_bpos[i] >> 6));
						};
						for (uint32_t i = 4u * (uint32_t)tid; i + 4u <= mb; i += 4u * NT) {
							const uint64_t g	= g0 + i;
							const uint2	   p2	= *reinterpret_cast<const uint2*>(s_bpos + i);
							const uint32_t o[4] = {p2.x & 0xFFFFu, p2.x >> 16, p2.y & 0xFFFFu, p2.y >> 16};
							const uint32_t v[4] = {out_pos(o[0]), out_pos(o[1]), out_pos(o[2]), out_pos(o[3])};
							__builtin_memcpy(a.bpos + g, v, 16);
							const uint32_t c4 = *reinterpret_cast<const uint32_t*>(s_bch + i);
							__builtin_memcpy(a.bchar + g, &c4, 4);
							if (all_keys) {
								int32_t k[4];
#pragma unroll
								for (int u = 0; u < 4; ++u) k[u] = kb + (int32_t)s_bkey[i + u];
								__builtin_memcpy(a.bkey + g, k, 16);
							} else {  // the chunk start in [g, g + 4), if there is one
								const uint64_t c = (g + 3u) & ~(uint64_t)(kStkChunk - 1u);
								if (c >= g)
									a.bkey[c] = kb + (int32_t)s_bkey[i + (uint32_t)(c - g)];
							}
							if constexpr (BATCH) {
								uint32_t d[4];
#pragma unroll
								for (int u = 0; u < 4; ++u) d[u] = (uint32_t)d0 + slot_of(r * NT + (int)(o[u] >> 6));
								__builtin_memcpy(a.b.bdoc + g, d, 16);
							}
						}
						const uint32_t body = mb & ~3u;
						if ((uint32_t)tid < mb - body)
							one(body + (uint32_t)tid);
						bdone = true;
					}
				}
				if (!bdone)
				for (uint32_t i0 = 0; i0 < mb; i0 += U * NT) {
					uint8_t ch[U];
#pragma unroll
					for (int u = 0; u < U; ++u) {
						const uint32_t i = i0 + u * NT + tid;
						if constexpr (E::kBitParallel)
							ch[u] = i >= mb ? 0 : bit_chr ? s_bch[i] : a.data[rowb + s_bpos[i]];
						else
							ch[u] = i < mb ? a.data[rowb + s_bpos[i]] : 0;
					}
#pragma unroll
					for (int u = 0; u < U; ++u) {
						const uint32_t i = i0 + u * NT + tid;
						if (i < mb) {
							const uint32_t off = s_bpos[i];
							const uint32_t g   = rb[1] + i;
							a.bpos[g]		   = out_pos(off);
							if (want_sidx)
								a.bsidx[g] = rb[0] + s_bsidx[i];
							if (all_keys || (g & (kStkChunk - 1u)) == 0)
								a.bkey[g] = kb + (int32_t)s_bkey[i];
							a.bchar[g] = ch[u];
							if constexpr (BATCH)
								a.b.bdoc[rb[1] + i] = (uint32_t)d0 + slot_of(r * NT + (int)(off >> 6));
						}
					}
				}
				__syncthreads();
			} else {  // a row too dense to stage: direct stores
				uint64_t m	= sm[r];
				uint32_t li = cpre[r][0], lb = cpre[r][1];
				int32_t	 bal = dl;
				while (m) {
					const int	   p   = __builtin_ctzll(m);
					uint32_t	   pos = (uint32_t)(pbase + loff + (uint64_t)p);
					if constexpr (BATCH)
						pos = s_lpos0[r * NT + tid] + (uint32_t)p;
					const uint8_t ch = a.data[rowb + loff + p];
					if ((uint64_t)rb[0] + li < a.cap_struct) {
						a.positions[rb[0] + li] = pos;
						if (a.chars)
							a.chars[rb[0] + li] = ch;
					}
					if ((om[r] >> p) & 1ull) {
						const bool op = (opm[r] >> p) & 1ull;
						bal += op ? 1 : -1;
						rmin = min(rmin, bal);
						rmax = max(rmax, bal);
						if ((uint64_t)rb[1] + lb < a.cap_br) {
							const uint32_t g = rb[1] + lb;
							a.bpos[g]		 = pos;
							if (want_sidx)
								a.bsidx[g] = rb[0] + li;
							if (all_keys || (g & (kStkChunk - 1u)) == 0)
								a.bkey[g] = (BATCH ? 0 : (int32_t)cex[2]) + (op ? bal - 1 : bal);
							a.bchar[g] = ch;
							if constexpr (BATCH)
								a.b.bdoc[g] = (uint32_t)ldoc[r];
						}
						++lb;
					}
					++li;
					m &= m - 1;
				}
			}
			if constexpr (BATCH) {
				// per-document extremes: a segmented wave scan over the lanes of
				// one document (a contiguous run), then one atomic per run
				int32_t		   dmin = rmin, dmax = rmax;
				const uint32_t doc	= (uint32_t)ldoc[r];
				auto		   seg	= [&](uint32_t od, int32_t omin, int32_t omax) {
					   if (od == doc) {
						   dmin = min(dmin, omin);
						   dmax = max(dmax, omax);
					   }
				};
#define PARS_SEG_STEP(CTRL, RM)                                                                          \
	seg(dpp<CTRL, RM>(0xFFFFFFFFu, doc), (int32_t)dpp<CTRL, RM>((uint32_t)INT_MAX, (uint32_t)dmin),      \
		(int32_t)dpp<CTRL, RM>((uint32_t)INT_MIN, (uint32_t)dmax))
				PARS_SEG_STEP(kDppRowShr1, 0xF);
				PARS_SEG_STEP(kDppRowShr2, 0xF);
				PARS_SEG_STEP(kDppRowShr4, 0xF);
				PARS_SEG_STEP(kDppRowShr8, 0xF);
				PARS_SEG_STEP(kDppBcast15, 0xA);
				PARS_SEG_STEP(kDppBcast31, 0xC);
#undef PARS_SEG_STEP
				const uint64_t hw	= s_hw[r * 4 + wid];  // heads of this wave's lanes
				const bool	   tail = lane == 63 || ((hw >> (lane + 1)) & 1ull);
				if (tail && ldoc[r] >= 0 && dmin != INT_MAX) {	// document-relative already
					atomicMin(&a.b.doc_min[ldoc[r]], dmin);
					atomicMax(&a.b.doc_max[ldoc[r]], dmax);
				}
			}
			lmin = min(lmin, rmin);
			lmax = max(lmax, rmax);
			rowd += (int32_t)rtot[r][2];
#pragma unroll
			for (int f = 0; f < 3; ++f) rb[f] += rtot[r][f];
		}
		lmin = wave_min(lmin);
		lmax = wave_max(lmax);
		if (lane == 0) {
			s_mm[wid][0] = lmin;
			s_mm[wid][1] = lmax;
		}
		__syncthreads();
		if (tid == 0) {
			int32_t mn = s_mm[0][0], mx = s_mm[0][1];
#pragma unroll
			for (int w2 = 1; w2 < NWV; ++w2) {
				mn = min(mn, s_mm[w2][0]);
				mx = max(mx, s_mm[w2][1]);
			}
			const int32_t off = BATCH ? 0 : (int32_t)cex[2];
			if (mn != INT_MAX)
				atomicMin(&a.sum->min_bal, off + mn);
			if (mx != INT_MIN)
				atomicMax(&a.sum->max_bal, off + mx);
		}
	}
	// every block has drawn a ticket past the end: the last one out resets
	if (tid == 0 && atomicAdd(a.done, 1u) == gridDim.x - 1) {
		atomicExch(a.ticket, 0u);
		atomicExch(a.done, 0u);
	}
}

}  // namespace pars_hip
