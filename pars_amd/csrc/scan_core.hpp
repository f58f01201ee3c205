// scan_core.hpp — the single-pass R1-R5 tile kernel, shared by every lane
// evaluator (general descriptor walk, JSON-style escape and quote-parity fast
// paths).
//
// Per tile (W rows of 256 lanes x 64 contiguous bytes):
//   1. the evaluator walks each lane's bytes once and yields, per descriptor,
//      the lane's state map, plus enough to produce the lane's masked bits for
//      any lane-entry state without a second walk;
//   2. block scans of the lane maps per row; decoupled look-back over tiles on
//      tagged 64-bit words gives the tile-entry states (chain 1);
//   3. masks -> unmasked structural / bracket bits; counts and the +-1 depth
//      delta; block scans + a second look-back (chain 2) give output offsets;
//   4. uint32 positions (and bytes, bracket keys for R6) staged in LDS per row
//      and written with coalesced stores; the bitmap word per lane.
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

namespace pars_hip {

// A look-back waits only on lower tiles, which were all dispatched earlier and
// publish before they wait, so it always ends; the spin bound is a guard so a
// defect can never hang the device: past it the kernel flags the summary and
// returns.
constexpr uint32_t kSpinLimit	= 1u << 21;
constexpr uint32_t kErrLookback = 0x80000000u;

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, kWave);
	return v;
}

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

// Chain 1 look-back, run by one full wave of tile `tile` > 0.  Returns the
// packed per-descriptor states entering the tile.
template<int ND, bool WIDE>
__device__ uint32_t lookback_states(const uint64_t* recs, uint32_t tile, uint32_t epoch, uint32_t in_states,
									DevSummary* sum) {
	using M	  = MapT<WIDE>;
	using Map = typename M::type;
	const int lane = lane_id();
	Map		  acc[ND];
#pragma unroll
	for (int d = 0; d < ND; ++d) acc[d] = M::id();
	int64_t	 jb	   = (int64_t)tile - 1;
	uint32_t spins = 0;
	while (true) {
		const int64_t j		 = jb - lane;
		uint32_t	  status = 0, ival = 0;
		Map			  agg[ND];
#pragma unroll
		for (int d = 0; d < ND; ++d) agg[d] = M::id();
		if (j < 0) {
			status = kIncl;
			ival   = in_states;
		} else {
			// one round trip: the inclusive and every aggregate word together
			const uint64_t* r = recs + (uint64_t)j * kRecWords;
			const uint64_t	w = peek(r + kRecSIncl);
			uint64_t		wa[ND];
#pragma unroll
			for (int d = 0; d < ND; ++d) wa[d] = peek(r + kRecSAgg + d);
			if (word_is(w, epoch, kIncl)) {
				status = kIncl;
				ival   = (uint32_t)w;
			} else {
				status = kAgg;
#pragma unroll
				for (int d = 0; d < ND; ++d) {
					if (!word_is(wa[d], epoch, kAgg))
						status = 0;
					else
						agg[d] = M::unpack((uint32_t)wa[d]);
				}
			}
		}
		const uint64_t b_incl = __ballot(status == kIncl);
		const uint64_t b_none = __ballot(status == 0);
		const int	   k	  = b_incl ? (__ffsll((long long)b_incl) - 1) : 64;
		const uint64_t need	  = k >= 64 ? ~0ull : ((1ull << k) - 1ull);
		if (b_none & need) {
			if (++spins > kSpinLimit) {
				if (lane == 0)
					atomicOr(&sum->overflow, kErrLookback);
				return in_states;
			}
			__builtin_amdgcn_s_sleep(1);
			continue;
		}
		// window map: agg_0 . agg_1 . ... . agg_{k-1} (lane 0 = newest tile, outermost)
		Map win[ND];
#pragma unroll
		for (int d = 0; d < ND; ++d) {
			Map m = lane < k ? agg[d] : M::id();
#pragma unroll
			for (int off = 1; off < kWave; off <<= 1) {
				Map o = M::shfl_up(m, off);
				if (lane >= off)
					m = M::compose(o, m);
			}
			win[d] = k == 0 ? M::id() : M::shfl(m, k - 1);
		}
		if (k < 64) {
			const uint32_t S   = (uint32_t)__shfl((int)ival, k, kWave);
			uint32_t	   out = 0;
#pragma unroll
			for (int d = 0; d < ND; ++d) {
				uint32_t sd = (S >> (3 * d)) & 7u;
				sd			= M::apply(win[d], sd);
				sd			= M::apply(acc[d], sd);
				out |= sd << (3 * d);
			}
			return out;
		}
#pragma unroll
		for (int d = 0; d < ND; ++d) acc[d] = M::compose(acc[d], win[d]);
		jb -= 64;
	}
}

// Chain 2 look-back (counts), run by one full wave of tile `tile` > 0.
__device__ inline void lookback_counts(const uint64_t* recs, uint32_t tile, uint32_t epoch, const uint32_t base[3],
									   uint32_t out[3], DevSummary* sum) {
	const int lane	 = lane_id();
	uint32_t  acc[3] = {0u, 0u, 0u};
	int64_t	  jb	 = (int64_t)tile - 1;
	uint32_t  spins	 = 0;
	while (true) {
		const int64_t j		 = jb - lane;
		uint32_t	  status = 0, v[3] = {0u, 0u, 0u};
		if (j < 0) {
			status = kIncl;
			v[0]   = base[0];
			v[1]   = base[1];
			v[2]   = base[2];
		} else {
			// one round trip: the inclusive and aggregate words together
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
			status = inc ? kIncl : (agg ? kAgg : 0u);
#pragma unroll
			for (int f = 0; f < 3; ++f) v[f] = (uint32_t)(inc ? wi[f] : wa[f]);
		}
		const uint64_t b_incl = __ballot(status == kIncl);
		const uint64_t b_none = __ballot(status == 0);
		const int	   k	  = b_incl ? (__ffsll((long long)b_incl) - 1) : 64;
		const uint64_t need	  = k >= 64 ? ~0ull : ((1ull << k) - 1ull);
		if (b_none & need) {
			if (++spins > kSpinLimit) {
				if (lane == 0)
					atomicOr(&sum->overflow, kErrLookback);
				out[0] = base[0];
				out[1] = base[1];
				out[2] = base[2];
				return;
			}
			__builtin_amdgcn_s_sleep(1);
			continue;
		}
#pragma unroll
		for (int f = 0; f < 3; ++f) acc[f] += wave_sum_u32(lane <= k ? v[f] : 0u);
		if (k < 64) {
			out[0] = acc[0];
			out[1] = acc[1];
			out[2] = acc[2];
			return;
		}
		jb -= 64;
	}
}

// ---------------------------------------------------------------------------
// The tile kernel.  Persistent: every block loads its LDS tables once, then
// draws tiles from an atomic ticket in dispatch order until none is left (so a
// tile only ever waits on tiles held by running blocks).  A tile is W rows of
// 16 KiB; in row r, lane l owns bytes [r*16K + l*64, +64), so every row's
// outputs are one contiguous run: they are staged in LDS and written back with
// coalesced stores (scattered 4-byte stores amplify HBM writes several-fold).
//
// E (the lane evaluator) provides:
//   kND, kWide, Map        descriptors and their map type
//   Lds, load(Lds&, plan)  LDS tables (cooperative load, all threads)
//   walk(Lds, plan, x, valid)   one pass over a lane's 64 bytes
//   T[d]                    the lane's map after walk
//   masked(s0[])            masked bits for lane-entry states s0
//   classes(S, OC, OP)      structural / bracket / open bits
// ---------------------------------------------------------------------------
constexpr uint32_t kRowBytes = kBlock * kBPL;  // 16 KiB
constexpr int	   kCapP	 = 3072;		   // staged structural outputs per row
constexpr int	   kCapB	 = 768;			   // staged bracket outputs per row

template<class E, int W>
__global__ __launch_bounds__(kBlock) void k_scan(ScanArgs a) {
	constexpr int ND  = E::kND;
	constexpr int NDA = ND > 0 ? ND : 1;
	using M			  = MapT<E::kWide>;
	using Map		  = typename M::type;
	constexpr uint64_t kTileBytes = (uint64_t)W * kRowBytes;

	__shared__ typename E::Lds s_lds;
	__shared__ Map			   s_wagg[W][4][NDA];
	__shared__ uint32_t		   s_misc[8];
	__shared__ uint32_t		   s_wcnt[W][4][3];
	__shared__ int32_t		   s_mm[4][2];
	__shared__ uint32_t		   s_pos[kCapP];
	__shared__ uint32_t		   s_bpos[kCapB];
	__shared__ uint32_t		   s_bsidx[kCapB];
	__shared__ int32_t		   s_bkey[kCapB];
	__shared__ uint8_t		   s_bch[kCapB];

	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	E::load(s_lds, a.plan);
	const bool aligned = ((uintptr_t)a.data & 15u) == 0;

	while (true) {
		__syncthreads();  // LDS of the previous tile is free
		if (tid == 0)
			s_misc[0] = atomicAdd(a.ticket, 1u);
		__syncthreads();
		const uint32_t tile = s_misc[0];
		if (tile >= a.n_tiles)
			break;
		const uint64_t tbase = (uint64_t)tile * kTileBytes;
		uint64_t*	   rec	 = a.recs + (uint64_t)tile * kRecWords;

		// ---- 1. one walk per row ----
		E		 ev[W];
		uint32_t valid[W];
#pragma unroll
		for (int r = 0; r < W; ++r) {
			const uint64_t base = tbase + (uint64_t)r * kRowBytes + (uint64_t)tid * kBPL;
			valid[r]			= base >= a.n ? 0u : (uint32_t)min<uint64_t>((uint64_t)kBPL, a.n - base);
			uint32_t x[16];
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
			ev[r].walk(s_lds, a.plan, x, valid[r]);
		}

		// ---- 2. chain 1: entry states ----
		uint32_t tile_in = a.in_states;
		Map		 lpre[W][NDA];
		uint32_t row_in[W];
		if constexpr (ND > 0) {
			Map exc[W][ND];
#pragma unroll
			for (int r = 0; r < W; ++r)
#pragma unroll
				for (int d = 0; d < ND; ++d) {
					const Map inc = wave_scan_maps<E::kWide>(ev[r].T[d]);
					exc[r][d]	  = M::shfl_up(inc, 1);
					if (lane == 0)
						exc[r][d] = M::id();
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
					for (int w2 = 0; w2 < 4; ++w2) {
						if (w2 == wid)
							lpre[r][d] = M::compose(exc[r][d], wp);
						wp = M::compose(s_wagg[r][w2][d], wp);
					}
					ragg[r][d] = wp;
					tagg[d]	   = M::compose(wp, tagg[d]);
				}
			}
			if (tile != 0) {
				if (tid == 0)
#pragma unroll
					for (int d = 0; d < ND; ++d)
						publish(rec + kRecSAgg + d, make_word(a.epoch, kAgg, M::pack(tagg[d])));
				if (wid == 0) {
					const uint32_t st = lookback_states<ND, E::kWide>(a.recs, tile, a.epoch, a.in_states, a.sum);
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
				publish(rec + kRecSIncl, make_word(a.epoch, kIncl, cur));
				if (tile == a.n_tiles - 1)
					a.sum->final_states = cur;
			}
		} else {
#pragma unroll
			for (int r = 0; r < W; ++r) row_in[r] = 0;
			if (tid == 0 && tile == a.n_tiles - 1)
				a.sum->final_states = 0;
		}

		// ---- 3. masks and counts ----
		uint64_t sm[W], om[W], opm[W];
		uint32_t cnt[W][3];
#pragma unroll
		for (int r = 0; r < W; ++r) {
			uint32_t s0[NDA];
#pragma unroll
			for (int d = 0; d < ND; ++d) s0[d] = M::apply(lpre[r][d], (row_in[r] >> (3 * d)) & 7u);
			const uint64_t masked = ND > 0 ? ev[r].masked(s0) : 0ull;
			uint64_t	   S, OC, OP;
			ev[r].classes(S, OC, OP);
			sm[r]  = S & ~masked;
			om[r]  = OC & ~masked;
			opm[r] = OP & ~masked;
			if (a.bitmap && valid[r])
				a.bitmap[(tbase + (uint64_t)r * kRowBytes) / 64 + tid] = sm[r];
			cnt[r][0] = (uint32_t)__popcll(sm[r]);
			cnt[r][1] = (uint32_t)__popcll(om[r]);
			cnt[r][2] = 2u * (uint32_t)__popcll(opm[r]) - cnt[r][1];
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
					for (int w2 = 0; w2 < 4; ++w2) {
						const uint32_t v = s_wcnt[r][w2][f];
						if (w2 < wid)
							p += v;
						t += v;
					}
					cpre[r][f] = p + cinc[r][f] - cnt[r][f];
					rtot[r][f] = t;
					ctot[f] += t;
				}
		}

		// chain 2: tile output offsets
		const uint32_t cbase[3] = {a.out_base, a.br_base, (uint32_t)a.depth_base};
		uint32_t	   cex[3]	= {cbase[0], cbase[1], cbase[2]};
		if (tile != 0) {
			if (tid == 0)
#pragma unroll
				for (int f = 0; f < 3; ++f) publish(rec + kRecCAgg + f, make_word(a.epoch, kAgg, ctot[f]));
			if (wid == 0) {
				uint32_t o[3];
				lookback_counts(a.recs, tile, a.epoch, cbase, o, a.sum);
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
#pragma unroll
			for (int f = 0; f < 3; ++f) publish(rec + kRecCIncl + f, make_word(a.epoch, kIncl, cex[f] + ctot[f]));
			if (tile == a.n_tiles - 1) {
				a.sum->n_struct	  = cex[0] + ctot[0];
				a.sum->n_brackets = cex[1] + ctot[1];
				a.sum->last_bal	  = (int32_t)(cex[2] + ctot[2]);
			}
		}
		if (a.pair_pos)	 // -1 for every structural index of the tile; R6 overwrites brackets
			for (uint32_t i = tid; i < ctot[0]; i += kBlock)
				if ((uint64_t)cex[0] + i < a.cap_struct)
					a.pair_pos[cex[0] + i] = -1;

		// ---- 4. scatter, row by row ----
		uint32_t rb[3] = {cex[0], cex[1], cex[2]};
		int32_t	 lmin = INT_MAX, lmax = INT_MIN;
#pragma unroll
		for (int r = 0; r < W; ++r) {
			const uint64_t rowb	  = tbase + (uint64_t)r * kRowBytes;
			const uint64_t lbase  = rowb + (uint64_t)tid * kBPL;
			const bool	   staged = rtot[r][0] <= (uint32_t)kCapP && rtot[r][1] <= (uint32_t)kCapB;
			uint32_t	   li = cpre[r][0], lb = cpre[r][1];
			int32_t		   bal	= (int32_t)(rb[2] + cpre[r][2]);
			const uint64_t pbase = a.byte_base + lbase;
			uint64_t	   m	 = sm[r];
			while (m) {
				const int	   p   = __builtin_ctzll(m);
				const uint32_t pos = (uint32_t)(pbase + (uint64_t)p);
				const bool	   br  = (om[r] >> p) & 1ull;
				uint8_t		   ch  = 0;
				if (br || a.chars)
					ch = a.data[lbase + p];
				if (a.chars && (uint64_t)rb[0] + li < a.cap_struct)
					a.chars[rb[0] + li] = ch;
				if (staged)
					s_pos[li] = pos;
				else if ((uint64_t)rb[0] + li < a.cap_struct)
					a.positions[rb[0] + li] = pos;
				if (br) {
					const bool op = (opm[r] >> p) & 1ull;
					bal += op ? 1 : -1;
					lmin			  = min(lmin, bal);
					lmax			  = max(lmax, bal);
					const int32_t key = op ? bal - 1 : bal;
					if (staged) {
						s_bpos[lb]	= pos;
						s_bsidx[lb] = rb[0] + li;
						s_bkey[lb]	= key;
						s_bch[lb]	= ch;
					} else if ((uint64_t)rb[1] + lb < a.cap_br) {
						const uint32_t g = rb[1] + lb;
						a.bpos[g]		 = pos;
						a.bsidx[g]		 = rb[0] + li;
						a.bkey[g]		 = key;
						a.bchar[g]		 = ch;
					}
					++lb;
				}
				++li;
				m &= m - 1;
			}
			if (staged) {  // block-uniform
				__syncthreads();
				const uint32_t np = rtot[r][0], nb = rtot[r][1];
				const uint64_t lim_p = a.cap_struct > rb[0] ? a.cap_struct - rb[0] : 0;
				const uint64_t lim_b = a.cap_br > rb[1] ? a.cap_br - rb[1] : 0;
				for (uint32_t i = tid; i < np && i < lim_p; i += kBlock) a.positions[rb[0] + i] = s_pos[i];
				for (uint32_t i = tid; i < nb && i < lim_b; i += kBlock) {
					a.bpos[rb[1] + i]  = s_bpos[i];
					a.bsidx[rb[1] + i] = s_bsidx[i];
					a.bkey[rb[1] + i]  = s_bkey[i];
					a.bchar[rb[1] + i] = s_bch[i];
				}
				__syncthreads();
			}
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
			for (int w2 = 1; w2 < 4; ++w2) {
				mn = min(mn, s_mm[w2][0]);
				mx = max(mx, s_mm[w2][1]);
			}
			if (mn != INT_MAX)
				atomicMin(&a.sum->min_bal, mn);
			if (mx != INT_MIN)
				atomicMax(&a.sum->max_bal, mx);
		}
	}
	// every block has drawn a ticket past the end: the last one out resets
	if (tid == 0 && atomicAdd(a.done, 1u) == gridDim.x - 1) {
		atomicExch(a.ticket, 0u);
		atomicExch(a.done, 0u);
	}
}

}  // namespace pars_hip
