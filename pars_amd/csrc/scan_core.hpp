// scan_core.hpp — the single-pass R1-R5 tile kernel, shared by every lane
// evaluator (general descriptor walk, JSON-style escape fast path).
//
// Per 16 KiB tile (256 lanes x 64 contiguous bytes):
//   1. the evaluator walks the lane's bytes once and yields, per descriptor, the
//      lane's state map, plus enough to produce the lane's masked bits for any
//      lane-entry state without a second walk;
//   2. block scan of the lane maps; decoupled look-back over tiles on tagged
//      64-bit words gives the tile-entry states (chain 1);
//   3. masks -> unmasked structural / bracket bits; counts and the +-1 depth
//      delta; block scan + a second look-back (chain 2) give output offsets;
//   4. scatter of uint32 positions (and chars, bitmap, bracket keys for R6).
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
			const uint64_t* r = recs + (uint64_t)j * kRecWords;
			const uint64_t	w = peek(r + kRecSIncl);
			if (word_is(w, epoch, kIncl)) {
				status = kIncl;
				ival   = (uint32_t)w;
			} else {
				status = kAgg;
#pragma unroll
				for (int d = 0; d < ND; ++d) {
					const uint64_t wa = peek(r + kRecSAgg + d);
					if (!word_is(wa, epoch, kAgg))
						status = 0;
					else
						agg[d] = M::unpack((uint32_t)wa);
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
			const uint64_t* r	= recs + (uint64_t)j * kRecWords;
			bool			inc = true;
#pragma unroll
			for (int f = 0; f < 3; ++f) {
				const uint64_t w = peek(r + kRecCIncl + f);
				if (!word_is(w, epoch, kIncl))
					inc = false;
				v[f] = (uint32_t)w;
			}
			if (inc) {
				status = kIncl;
			} else {
				status = kAgg;
#pragma unroll
				for (int f = 0; f < 3; ++f) {
					const uint64_t w = peek(r + kRecCAgg + f);
					if (!word_is(w, epoch, kAgg))
						status = 0;
					v[f] = (uint32_t)w;
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
// The tile kernel.  E (the lane evaluator) provides:
//   kND, kWide, Map        descriptors and their map type
//   Lds, load(Lds&, plan)  LDS tables (cooperative load, all threads)
//   walk(Lds, plan, x, valid)         one pass over the lane's bytes
//   T[d]                    lane maps after walk
//   masked(s0[])            masked bits for lane-entry states s0
//   classes(S, OC, OP)      structural / bracket / open bits
// ---------------------------------------------------------------------------
template<class E>
__global__ __launch_bounds__(kBlock) void k_scan(ScanArgs a) {
	constexpr int ND  = E::kND;
	constexpr int NDA = ND > 0 ? ND : 1;
	using M			  = MapT<E::kWide>;
	using Map		  = typename M::type;

	__shared__ typename E::Lds s_lds;
	__shared__ Map			   s_wagg[4][NDA];
	__shared__ uint32_t		   s_misc[8];
	__shared__ uint32_t		   s_wcnt[4][3];
	__shared__ int32_t		   s_mm[4][2];

	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	if (tid == 0) {
		const uint32_t t = atomicAdd(a.ticket, 1u);
		if (t == a.n_tiles - 1)
			atomicExch(a.ticket, 0u);  // every block of this launch has drawn its ticket
		s_misc[0] = t;
	}
	E::load(s_lds, a.plan);
	__syncthreads();

	const uint32_t tile	 = s_misc[0];
	const uint64_t base	 = (uint64_t)tile * kTile + (uint64_t)tid * kBPL;
	const uint32_t valid = base >= a.n ? 0u : (uint32_t)min<uint64_t>((uint64_t)kBPL, a.n - base);
	uint64_t*	   rec	 = a.recs + (uint64_t)tile * kRecWords;

	// ---- load 64 bytes (the tail lane reads only what exists) ----
	uint32_t   x[16];
	const bool aligned = ((uintptr_t)a.data & 15u) == 0;
	if (valid == kBPL && aligned) {
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
				if ((uint32_t)(4 * k + b) < valid)
					w |= (uint32_t)a.data[base + 4 * k + b] << (8 * b);
			x[k] = w;
		}
	}

	// ---- 1. one walk ----
	E ev;
	ev.walk(s_lds, a.plan, x, valid);

	// ---- 2. chain 1: entry states ----
	uint32_t tile_in = a.in_states;
	Map		 lpre[NDA];
	if constexpr (ND > 0) {
		Map inc[ND], exc[ND];
#pragma unroll
		for (int d = 0; d < ND; ++d) {
			inc[d] = wave_scan_maps<E::kWide>(ev.T[d]);
			exc[d] = M::shfl_up(inc[d], 1);
			if (lane == 0)
				exc[d] = M::id();
		}
		if (lane == 63)
#pragma unroll
			for (int d = 0; d < ND; ++d) s_wagg[wid][d] = inc[d];
		__syncthreads();
		Map tagg[ND];
#pragma unroll
		for (int d = 0; d < ND; ++d) {
			Map wp = M::id();
#pragma unroll
			for (int w2 = 0; w2 < 4; ++w2) {
				if (w2 == wid)
					lpre[d] = M::compose(exc[d], wp);
				wp = M::compose(s_wagg[w2][d], wp);
			}
			tagg[d] = wp;
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
		if (tid == 0) {
			uint32_t out = 0;
#pragma unroll
			for (int d = 0; d < ND; ++d) out |= M::apply(tagg[d], (tile_in >> (3 * d)) & 7u) << (3 * d);
			publish(rec + kRecSIncl, make_word(a.epoch, kIncl, out));
			if (tile == a.n_tiles - 1)
				a.sum->final_states = out;
		}
	} else {
		if (tid == 0 && tile == a.n_tiles - 1)
			a.sum->final_states = 0;
	}

	// ---- 3. masks and counts ----
	uint32_t s0[NDA];
#pragma unroll
	for (int d = 0; d < ND; ++d) s0[d] = M::apply(lpre[d], (tile_in >> (3 * d)) & 7u);
	const uint64_t masked = ND > 0 ? ev.masked(s0) : 0ull;
	uint64_t	   S, OC, OP;
	ev.classes(S, OC, OP);
	const uint64_t sm = S & ~masked, om = OC & ~masked, opm = OP & ~masked;
	if (a.bitmap && valid)
		a.bitmap[base >> 6] = sm;

	const uint32_t cnt[3] = {(uint32_t)__popcll(sm), (uint32_t)__popcll(om),
							 2u * (uint32_t)__popcll(opm) - (uint32_t)__popcll(om)};
	uint32_t	   cinc[3];
#pragma unroll
	for (int f = 0; f < 3; ++f) cinc[f] = wave_incl_sum(cnt[f]);
	if (lane == 63)
#pragma unroll
		for (int f = 0; f < 3; ++f) s_wcnt[wid][f] = cinc[f];
	__syncthreads();
	uint32_t cpre[3], ctot[3];
#pragma unroll
	for (int f = 0; f < 3; ++f) {
		uint32_t p = 0, t = 0;
#pragma unroll
		for (int w2 = 0; w2 < 4; ++w2) {
			const uint32_t v = s_wcnt[w2][f];
			if (w2 < wid)
				p += v;
			t += v;
		}
		cpre[f] = p + cinc[f] - cnt[f];
		ctot[f] = t;
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

	// ---- 4. scatter ----
	uint32_t	   si = cex[0] + cpre[0], bi = cex[1] + cpre[1];
	int32_t		   bal	= (int32_t)(cex[2] + cpre[2]);
	int32_t		   lmin = INT_MAX, lmax = INT_MIN;
	const uint64_t pbase = a.byte_base + base;
#pragma unroll
	for (int q = 0; q < 16; ++q) {
		uint32_t m = (uint32_t)(sm >> (4 * q)) & 0xFu;
		while (m) {
			const int	   b   = __builtin_ctz(m);
			const int	   p   = 4 * q + b;
			const uint32_t pos = (uint32_t)(pbase + (uint64_t)p);
			const uint8_t  ch  = (uint8_t)(x[q] >> (8 * b));
			if (si < a.cap_struct) {
				a.positions[si] = pos;
				if (a.chars)
					a.chars[si] = ch;
				if (a.pair_pos)
					a.pair_pos[si] = -1;
			}
			if ((om >> p) & 1ull) {
				const bool op = (opm >> p) & 1ull;
				bal += op ? 1 : -1;
				lmin = min(lmin, bal);
				lmax = max(lmax, bal);
				if (bi < a.cap_br) {
					a.bpos[bi]	= pos;
					a.bsidx[bi] = si;
					a.bkey[bi]	= op ? bal - 1 : bal;
					a.bchar[bi] = ch;
				}
				++bi;
			}
			++si;
			m &= m - 1;
		}
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

}  // namespace pars_hip
