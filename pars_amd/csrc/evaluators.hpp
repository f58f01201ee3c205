// evaluators.hpp — lane evaluators plugged into the tile kernel (scan_core.hpp).
//
//   GeneralEval<ND, WIDE>  any plan: every descriptor stepped byte by byte as
//                          a byte-map composition (v_perm), with the masked bit
//                          kept for every possible lane-entry state.
//   EscapeEval             one affine descriptor with a prefix escape (JSON
//                          strings): bit-parallel over the lane's 64 bytes, exact
//                          for every input including a backslash outside a string
//                          (which does NOT escape there, model/scan.hpp semantics).
//   ParityEval             one quote-parity descriptor (CSV doubled quotes).
#pragma once

#include "scan_core.hpp"

namespace pars_hip {

void set_last_scan_occupancy(int per_cu);

namespace {

template<bool WIDE>
struct DTEntry;

template<>
struct DTEntry<false> {
	using type = uint2;	 // {next lo, masked lo}
};

template<>
struct DTEntry<true> {
	using type = uint4;	 // {next lo, next hi, masked lo, masked hi}
};

// Class bytes of eight accumulated groups -> three 64-bit masks
// (byte 0 structural, 1 open|close, 2 open).
__device__ __forceinline__ void gather_classes(const uint32_t (&c)[8], uint64_t& S, uint64_t& OC, uint64_t& OP,
											   uint64_t& B3) {
	uint32_t s0, o0, p0, q0, s1, o1, p1, q1;
	transpose4(c[0], c[1], c[2], c[3], s0, o0, p0, q0);
	transpose4(c[4], c[5], c[6], c[7], s1, o1, p1, q1);
	S  = ((uint64_t)s1 << 32) | s0;
	OC = ((uint64_t)o1 << 32) | o0;
	OP = ((uint64_t)p1 << 32) | p0;
	B3 = ((uint64_t)q1 << 32) | q0;
}

// Class lookups of a lane's 64 bytes through the position-shifted two-dword
// table: group g ORs the entries of bytes 8g..8g+7 (a balanced OR tree, so the
// compiler forms three-input ORs).  Every lane looks up all 64 bytes (no
// divergent tail path); callers clear the bits at and past `valid` in the
// gathered masks.
__device__ __forceinline__ void lookup_groups(const uint2* cls, const uint32_t (&x)[16], uint32_t (&clo)[8],
											  uint32_t (&chi)[8]) {
#pragma unroll
	for (int g = 0; g < 8; ++g) {
		uint2 e[8];
#pragma unroll
		for (int b = 0; b < 8; ++b) {
			const int	   j = 8 * g + b;
			const uint32_t c = (x[j >> 2] >> ((j & 3) * 8)) & 0xFFu;
			e[b]			 = cls[b * 256 + c];
		}
		clo[g] = ((e[0].x | e[1].x) | (e[2].x | e[3].x)) | ((e[4].x | e[5].x) | (e[6].x | e[7].x));
		chi[g] = ((e[0].y | e[1].y) | (e[2].y | e[3].y)) | ((e[4].y | e[5].y) | (e[6].y | e[7].y));
	}
}

// Bits below `valid` (valid <= 64).
__device__ __forceinline__ uint64_t valid_mask(uint32_t valid) {
	return valid >= 64 ? ~0ull : ((1ull << valid) - 1ull);
}

// Byte 0 (byte 1) of eight accumulated groups -> a 64-bit mask (byte g = group g).
__device__ __forceinline__ uint64_t gather_byte0(const uint32_t (&c)[8]) {
	const uint32_t lo = perm(c[1], c[0], 0x0C0C0400u) | perm(c[3], c[2], 0x04000C0Cu);
	const uint32_t hi = perm(c[5], c[4], 0x0C0C0400u) | perm(c[7], c[6], 0x04000C0Cu);
	return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t gather_byte1(const uint32_t (&c)[8]) {
	const uint32_t lo = perm(c[1], c[0], 0x0C0C0501u) | perm(c[3], c[2], 0x05010C0Cu);
	const uint32_t hi = perm(c[5], c[4], 0x0C0C0501u) | perm(c[7], c[6], 0x05010C0Cu);
	return ((uint64_t)hi << 32) | lo;
}

// ---------------------------------------------------------------------------
template<int ND, bool WIDE>
struct GeneralEval {
	static constexpr int  kND		   = ND;
	static constexpr bool kWide		   = WIDE;
	static constexpr bool kBitParallel = false;
	static constexpr int  NDA		   = ND > 0 ? ND : 1;
	using M						= MapT<WIDE>;
	using Map					= typename M::type;
	using DT					= typename DTEntry<WIDE>::type;

	struct Lds {
		uint32_t cls[8 * 256];
		DT		 dt[NDA * 256];
	};

	__device__ static void load(Lds& L, const DevPlan* P) {
		for (int i = threadIdx.x; i < 8 * 256; i += (int)blockDim.x) L.cls[i] = (&P->cls[0][0])[i];
		for (int i = threadIdx.x; i < ND * 256; i += (int)blockDim.x) {
			const uint32_t* e = P->dt[i >> 8][i & 255];
			if constexpr (WIDE)
				L.dt[i] = make_uint4(e[0], e[1], e[2], e[3]);
			else
				L.dt[i] = make_uint2(e[0], e[2]);
		}
	}

	Map		 T[NDA];
	uint32_t accl[NDA][8];
	uint32_t acch[WIDE ? NDA : 1][8];
	uint32_t cacc[8];

	__device__ void walk(const Lds& L, const DevPlan*, const uint32_t (&x)[16], uint32_t valid) {
#pragma unroll
		for (int g = 0; g < 8; ++g) {
			cacc[g] = 0;
#pragma unroll
			for (int d = 0; d < NDA; ++d) {
				accl[d][g] = 0;
				if constexpr (WIDE)
					acch[d][g] = 0;
			}
		}
#pragma unroll
		for (int d = 0; d < NDA; ++d) T[d] = M::id();
		auto step = [&](int j) {
			const uint32_t c = (x[j >> 2] >> ((j & 3) * 8)) & 0xFFu;
			cacc[j >> 3] |= L.cls[(j & 7) * 256 + c];
#pragma unroll
			for (int d = 0; d < ND; ++d) {
				const DT e = L.dt[d * 256 + c];
				if constexpr (WIDE) {
					accl[d][j >> 3] |= perm(e.w, e.z, T[d].lo) << (j & 7);
					acch[d][j >> 3] |= perm(e.w, e.z, T[d].hi) << (j & 7);
					T[d] = M::compose(Map {e.x, e.y}, T[d]);
				} else {
					accl[d][j >> 3] |= perm(0u, e.y, T[d].lo) << (j & 7);
					T[d] = M::compose(Map {e.x}, T[d]);
				}
			}
		};
		if (valid == kBPL) {
#pragma unroll
			for (int j = 0; j < kBPL; ++j) step(j);
		} else {
#pragma unroll
			for (int j = 0; j < kBPL; ++j)
				if ((uint32_t)j < valid)
					step(j);
		}
	}

	__device__ uint64_t masked(const Lds&, const uint32_t (&s0)[NDA]) const { return masked(s0); }

	__device__ uint64_t masked(const uint32_t (&s0)[NDA]) const {
		uint64_t m = 0;
#pragma unroll
		for (int d = 0; d < ND; ++d) {
			const uint32_t s = s0[d];
#pragma unroll
			for (int g = 0; g < 8; ++g) {
				uint32_t byte;
				if constexpr (WIDE)
					byte = ((s < 4 ? accl[d][g] : acch[d][g]) >> (8 * (s & 3))) & 0xFFu;
				else
					byte = (accl[d][g] >> (8 * s)) & 0xFFu;
				m |= (uint64_t)byte << (8 * g);
			}
		}
		return m;
	}

	// One descriptor: masked bits for every entry state (state s = byte s).
	__device__ void masked_all(uint64_t (&m)[WIDE ? 8 : 4]) const {
		uint32_t t0[4], t1[4];
		transpose4(accl[0][0], accl[0][1], accl[0][2], accl[0][3], t0[0], t0[1], t0[2], t0[3]);
		transpose4(accl[0][4], accl[0][5], accl[0][6], accl[0][7], t1[0], t1[1], t1[2], t1[3]);
#pragma unroll
		for (int s = 0; s < 4; ++s) m[s] = ((uint64_t)t1[s] << 32) | t0[s];
		if constexpr (WIDE) {
			transpose4(acch[0][0], acch[0][1], acch[0][2], acch[0][3], t0[0], t0[1], t0[2], t0[3]);
			transpose4(acch[0][4], acch[0][5], acch[0][6], acch[0][7], t1[0], t1[1], t1[2], t1[3]);
#pragma unroll
			for (int s = 0; s < 4; ++s) m[4 + s] = ((uint64_t)t1[s] << 32) | t0[s];
		}
	}

	__device__ void classes(uint64_t& S, uint64_t& OC, uint64_t& OP) const {
		uint64_t unused;
		gather_classes(cacc, S, OC, OP, unused);
	}
};

// ---------------------------------------------------------------------------
// EscapeEval: the automaton out -delim-> in, in -delim-> closed, in -escape->
// esc, esc -any-> in, closed behaving as out; masked exactly in `in`/`esc`.
// A lane evaluates it bit-parallel for each possible entry state:
//   escapers  = backslashes not themselves escaped (odd positions within a
//               run's even/odd parity class, one carry-add per word);
//   U / X     = delimiters not escaped / escaped;
//   in-string = exclusive prefix-xor of the real toggles, where U always
//               toggles and X toggles only when outside (an escape outside a
//               string is inert); X events are resolved in order, rarely more
//               than one per lane, none on valid JSON for the true entry state.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t prefix_xor(uint64_t x) {
	x ^= x << 1;
	x ^= x << 2;
	x ^= x << 4;
	x ^= x << 8;
	x ^= x << 16;
	x ^= x << 32;
	return x;
}

__device__ __forceinline__ uint64_t escapers(uint64_t bs, uint64_t carry) {
	constexpr uint64_t EVEN		 = 0x5555555555555555ull;
	const uint64_t	   b		 = bs & ~carry;	 // an escaped first byte escapes nothing
	const uint64_t	   starts	 = b & ~(b << 1);
	const uint64_t	   even_runs = b & ~(b + (starts & EVEN));
	const uint64_t	   odd_runs	 = b & ~even_runs;
	return (even_runs & EVEN) | (odd_runs & ~EVEN);
}

enum : uint32_t { kCodeOut = 0, kCodeIn = 1, kCodeClosed = 2, kCodeEsc = 3 };

struct EscHyp {
	uint64_t P;		   // bytes masked (state before the byte is in or esc)
	uint64_t R;		   // delimiters that toggle
	uint64_t escaped;  // bytes consumed by an escape
};

__device__ __forceinline__ EscHyp esc_eval(uint64_t Q, uint64_t BS, bool in0, bool esc0) {
	const uint64_t carry   = esc0 ? 1ull : 0ull;
	const uint64_t escaped = (escapers(BS, carry) << 1) | carry;
	const uint64_t U = Q & ~escaped, X = Q & escaped;
	uint64_t	   R = U;
	uint64_t	   P = prefix_xor(U) ^ U;
	if (in0)
		P = ~P;
	uint64_t pend = X & ~P;
	while (pend) {
		const uint64_t b	 = pend & (~pend + 1ull);
		const uint64_t above = ~((b << 1) - 1ull);
		P ^= above;
		R |= b;
		pend = X & ~P & above;
	}
	return EscHyp {P, R, escaped};
}

__device__ __forceinline__ uint32_t esc_code(const EscHyp& h, uint64_t BS, uint64_t last) {
	const bool pl = (h.P & last) != 0, rl = (h.R & last) != 0;
	if (pl == rl)
		return pl ? kCodeClosed : kCodeOut;
	return (pl && (BS & last) && !(h.escaped & last)) ? kCodeEsc : kCodeIn;
}

struct EscapeEval {
	static constexpr int  kND		   = 1;
	static constexpr bool kWide		   = false;
	static constexpr bool kBitParallel = true;
	using Map					= Map4;

	struct Lds {
		uint2 cls[8 * 256];	 // {S|OC|OP|delim bytes, escape byte}, pre-shifted by position
	};

	__device__ static void load(Lds& L, const DevPlan* P) {
		for (int i = threadIdx.x; i < 8 * 256; i += (int)blockDim.x) {
			const uint32_t* e = P->cls8[i >> 8][i & 255];
			L.cls[i]		  = make_uint2(e[0], e[1]);
		}
	}

	Map		 T[1];
	uint64_t S, OC, OP;
	uint64_t BO;  // bracket ordinal bits (plan br_chr); stored to LDS right after the walk
	uint64_t mO, mI, mE;
	uint32_t st_in, st_esc;

	__device__ void walk(const Lds& L, const DevPlan* plan, const uint32_t (&x)[16], uint32_t valid) {
		uint32_t clo[8], chi[8];
		lookup_groups(L.cls, x, clo, chi);
		uint64_t Q;
		gather_classes(clo, S, OC, OP, Q);
		uint64_t	   BS = gather_byte0(chi);
		BO				  = gather_byte1(chi);
		const uint64_t vm = valid_mask(valid);	// bytes past the lane's end have no class
		S &= vm;
		OC &= vm;
		OP &= vm;
		Q &= vm;
		BS &= vm;

		const uint32_t f0 = plan->fk_state[0], f1 = plan->fk_state[1], f2 = plan->fk_state[2],
					   f3 = plan->fk_state[3];
		st_in			  = f1;
		st_esc			  = f3;
		if (valid == 0) {
			T[0].lo = kIdLo;
			mO = mI = mE = 0;
			return;
		}
		const uint64_t last = 1ull << (valid - 1);
		const EscHyp   hI	= esc_eval(Q, BS, true, false);
		EscHyp		   hO, hE;
		if ((Q & hI.escaped) == 0) {  // no escaped delimiter: entry out is entry in, inverted
			hO = EscHyp {~hI.P, hI.R, hI.escaped};
		} else {
			hO = esc_eval(Q, BS, false, false);
		}
		if (((Q | BS) & 1ull) == 0) {  // first byte plain: entry esc only consumes it
			hE = EscHyp {hI.P, hI.R, hI.escaped | 1ull};
		} else {
			hE = esc_eval(Q, BS, true, true);
		}
		mO = hO.P;
		mI = hI.P;
		mE = hE.P;
		const uint32_t fk = f0 | (f1 << 8) | (f2 << 16) | (f3 << 24);  // code -> plan state
		const uint32_t tO = (fk >> (8 * esc_code(hO, BS, last))) & 0xFFu;
		const uint32_t tI = (fk >> (8 * esc_code(hI, BS, last))) & 0xFFu;
		const uint32_t tE = (fk >> (8 * esc_code(hE, BS, last))) & 0xFFu;
		T[0].lo = (tO << (8 * f0)) | (tI << (8 * f1)) | (tO << (8 * f2)) | (tE << (8 * f3));
	}

	__device__ uint64_t masked(const Lds&, const uint32_t (&s0)[1]) const { return masked(s0); }

	__device__ uint64_t masked(const uint32_t (&s0)[1]) const {
		// blend with masks, not a select (a select of members lowers to scratch)
		const uint64_t kI = 0ull - (uint64_t)(s0[0] == st_in), kE = 0ull - (uint64_t)(s0[0] == st_esc);
		return (mO & ~(kI | kE)) | (mI & kI) | (mE & kE);
	}

	__device__ void masked_all(uint64_t (&m)[4]) const {
#pragma unroll
		for (int s = 0; s < 4; ++s) {
			const uint32_t s0[1] = {(uint32_t)s};
			m[s]				 = masked(s0);
		}
	}

	__device__ void classes(uint64_t& s, uint64_t& oc, uint64_t& op) const {
		s  = S;
		oc = OC;
		op = OP;
	}

};

// ---------------------------------------------------------------------------
// ParityEval: the three-state quote-parity automaton (CSV doubled quotes,
// escape-less strings).  The state before byte i is `in` exactly when an odd
// number of delimiters precede it; `closed` when the parity is even and byte
// i-1 is a delimiter.  One prefix-xor ladder serves every entry state.
// ---------------------------------------------------------------------------
struct ParityEval {
	static constexpr int  kND		   = 1;
	static constexpr bool kWide		   = false;
	static constexpr bool kBitParallel = true;
	using Map					= Map4;
	using Lds					= EscapeEval::Lds;

	__device__ static void load(Lds& L, const DevPlan* P) { EscapeEval::load(L, P); }

	Map		 T[1];
	uint64_t S, OC, OP, Q, Pe;
	uint64_t BO;  // bracket ordinal bits (plan br_chr); stored to LDS right after the walk
	uint32_t f_in, f_cl, doubled;

	__device__ void walk(const Lds& L, const DevPlan* plan, const uint32_t (&x)[16], uint32_t valid) {
		uint32_t clo[8], chi[8];
		lookup_groups(L.cls, x, clo, chi);
		gather_classes(clo, S, OC, OP, Q);
		BO				  = gather_byte1(chi);
		const uint64_t vm = valid_mask(valid);	// bytes past the lane's end have no class
		S &= vm;
		OC &= vm;
		OP &= vm;
		Q &= vm;
		const uint32_t f_out = plan->fk_state[0];
		f_in				 = plan->fk_state[1];
		f_cl				 = plan->fk_state[2];
		doubled				 = plan->fk_doubled;
		Pe					 = prefix_xor(Q) ^ Q;  // parity before each byte, entered even
		if (valid == 0) {
			T[0].lo = kIdLo;
			return;
		}
		const uint64_t last	  = 1ull << (valid - 1);
		const bool	   ql	  = (Q & last) != 0;
		const bool	   odd_ev = ((Pe & last) != 0) != ql;  // parity after the lane, entered even
		const uint32_t fe	  = odd_ev ? f_in : (ql ? f_cl : f_out);
		const uint32_t fo	  = !odd_ev ? f_in : (ql ? f_cl : f_out);
		T[0].lo = (fe << (8 * f_out)) | (fo << (8 * f_in)) | (fe << (8 * f_cl)) | (3u << 24);
	}

	__device__ uint64_t masked(const Lds&, const uint32_t (&s0)[1]) const { return masked(s0); }

	__device__ uint64_t masked(const uint32_t (&s0)[1]) const {
		const uint64_t odd = s0[0] == f_in ? ~Pe : Pe;
		if (!doubled)
			return odd;
		const uint64_t prevq = (Q << 1) | (s0[0] == f_cl ? 1ull : 0ull);
		return odd | (Q & prevq);
	}

	__device__ void masked_all(uint64_t (&m)[4]) const {
#pragma unroll
		for (int s = 0; s < 4; ++s) {
			const uint32_t s0[1] = {(uint32_t)s};
			m[s]				 = masked(s0);
		}
	}

	__device__ void classes(uint64_t& s, uint64_t& oc, uint64_t& op) const {
		s  = S;
		oc = OC;
		op = OP;
	}

};

// ---------------------------------------------------------------------------
// EventEval<ND, WIDE>: plans of two or more descriptors whose background runs
// act as a single byte (the DevPlan::ev_* tables).  A lane finds its event
// bytes (a byte that is not background for some descriptor) among its class
// lookups, then
//   walk:   composes each descriptor's map at the events only (the background
//           run before an event is folded into the event's map, a run to the
//           lane's end is one more step): the lane's maps for every entry
//           state, as the per-byte walk would give them;
//   masked: once the lane's entry states are known, steps the joint state of
//           all descriptors, one table lookup per event, and sets the masked
//           bits of each background run (its first byte, the rest) and of the
//           event byte.
// Exact for every input: inside a run the per-byte walk would apply the same
// idempotent map again and again, so every run byte after the first sees the
// same states.
// ---------------------------------------------------------------------------
template<int ND, bool WIDE>
struct EventEval {
	static constexpr int  kND		   = ND;
	static constexpr bool kWide		   = WIDE;
	static constexpr bool kBitParallel = false;
	using M							   = MapT<WIDE>;
	using Map						   = typename M::type;

	struct Lds {
		uint2	 cls[8 * 256];			   // DevPlan::evc
		uint2	 map[kEvIds * 2 * ND];	   // [(id * 2 + g) * ND + d]: ev_map lo, hi
		uint16_t joint[kMaxJoint * kEvIds * 2];  // [(J * kEvIds + id) * 2 + g]: ev_joint
		uint16_t stride[ND];
	};

	__device__ static void load(Lds& L, const DevPlan* P) {
		for (int i = threadIdx.x; i < 8 * 256; i += (int)blockDim.x) {
			const uint32_t* e = P->evc[i >> 8][i & 255];
			L.cls[i]		  = make_uint2(e[0], e[1]);
		}
		for (int i = threadIdx.x; i < kEvIds * 2 * ND; i += (int)blockDim.x) {
			const int d = i % ND, g = (i / ND) & 1, id = i / (2 * ND);
			L.map[i]	= make_uint2(P->ev_map[id][g][d][0], P->ev_map[id][g][d][1]);
		}
		const int nj = (int)P->ev_n_joint * kEvIds * 2;
		for (int i = threadIdx.x; i < nj; i += (int)blockDim.x) L.joint[i] = (&P->ev_joint[0][0][0])[i];
		if (threadIdx.x < ND)
			L.stride[threadIdx.x] = P->ev_stride[threadIdx.x];
	}

	Map		 T[ND];
	uint64_t S, OC, OP;
	uint64_t E0, E1, E2;  // event id bit planes
	uint32_t nv;		  // valid bytes of the lane

	__device__ uint32_t event_id(int p) const {
		return ((uint32_t)(E0 >> p) & 1u) | (((uint32_t)(E1 >> p) & 1u) << 1) | (((uint32_t)(E2 >> p) & 1u) << 2);
	}

	__device__ static Map map_at(const uint2& e) {
		if constexpr (WIDE)
			return Map {e.x, e.y};
		else
			return Map {e.x};
	}

	__device__ void walk(const Lds& L, const DevPlan*, const uint32_t (&x)[16], uint32_t valid) {
		uint32_t clo[8], chi[8];
		lookup_groups(L.cls, x, clo, chi);
		uint64_t e0;
		gather_classes(clo, S, OC, OP, e0);
		const uint64_t vm = valid_mask(valid);	// bytes past the lane's end have no class
		S &= vm;
		OC &= vm;
		OP &= vm;
		E0 = e0 & vm;
		E1 = gather_byte0(chi) & vm;
		E2 = gather_byte1(chi) & vm;
		nv = valid;
#pragma unroll
		for (int d = 0; d < ND; ++d) T[d] = M::id();
		uint64_t u = E0 | E1 | E2;
		int		 q = -1;  // the last event (or -1)
		while (u) {
			const int p = __builtin_ctzll(u);
			u &= u - 1ull;
			const uint2* mp = &L.map[(event_id(p) * 2 + (p > q + 1 ? 1 : 0)) * ND];
#pragma unroll
			for (int d = 0; d < ND; ++d) T[d] = M::compose(map_at(mp[d]), T[d]);
			q = p;
		}
		if (q < (int)valid - 1) {  // a background run to the lane's end
			const uint2* mp = &L.map[1 * ND];
#pragma unroll
			for (int d = 0; d < ND; ++d) T[d] = M::compose(map_at(mp[d]), T[d]);
		}
	}

	// bits a .. b-1 (0 <= a < b <= 64)
	__device__ static uint64_t span(int a, int b) {
		const uint64_t hi = b >= 64 ? ~0ull : ((1ull << b) - 1ull);
		return hi & ~((1ull << a) - 1ull);
	}

	// the masked bits of the background run q+1 .. end-1 from its table entry
	__device__ static uint64_t run_bits(uint32_t e, int q, int end) {
		uint64_t m = 0;
		if (end > q + 1) {
			if (e & (1u << 12))
				m |= 1ull << (q + 1);
			if ((e & (1u << 13)) && end > q + 2)
				m |= span(q + 2, end);
		}
		return m;
	}

	__device__ uint64_t masked(const Lds& L, const uint32_t (&s0)[ND]) const {
		uint32_t J = 0;
#pragma unroll
		for (int d = 0; d < ND; ++d) J += s0[d] * (uint32_t)L.stride[d];
		uint64_t m = 0, u = E0 | E1 | E2;
		int		 q = -1;
		while (u) {
			const int p = __builtin_ctzll(u);
			u &= u - 1ull;
			const uint32_t e = L.joint[(J * kEvIds + event_id(p)) * 2 + (p > q + 1 ? 1u : 0u)];
			m |= run_bits(e, q, p);
			m |= (uint64_t)((e >> 14) & 1u) << p;
			J = e & 0xFFFu;
			q = p;
		}
		if (q < (int)nv - 1)
			m |= run_bits(L.joint[(J * kEvIds) * 2 + 1], q, (int)nv);
		return m;
	}

	__device__ void classes(uint64_t& s, uint64_t& oc, uint64_t& op) const {
		s  = S;
		oc = OC;
		op = OP;
	}
};

// ---------------------------------------------------------------------------
// Persistent launch: as many blocks as fit on the device at once (never more
// than there are tiles); each draws tiles until none is left.
template<class E, int W, bool BATCH, int NWV = kBlock / kWave>
hipError_t launch_eval(const ScanArgs& in, hipStream_t s) {
	constexpr int NT  = NWV * kWave;  // threads per block
	static int	  cus = 0, per_cu = 0;	// per kernel instance
	if (!cus) {
		int		   dev = 0, c = 0, p = 0;
		hipError_t e = hipGetDevice(&dev);
		if (e == hipSuccess)
			e = hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
		if (e == hipSuccess)
			e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k_scan<E, W, BATCH, NWV>, NT, 0);
		if (e != hipSuccess)
			return e;
		per_cu = p > 0 ? p : 1;
		cus	   = c > 0 ? c : 1;
	}
	set_last_scan_occupancy(per_cu);
	const int resident = cus * (in.blocks_per_cu > 0 && (int)in.blocks_per_cu < per_cu ? (int)in.blocks_per_cu : per_cu);
	ScanArgs  a		   = in;
	const uint64_t tbytes = (uint64_t)W * NT * kBPL;
	a.n_tiles			  = (uint32_t)((a.n + tbytes - 1) / tbytes);
	if (a.n_tiles == 0)
		return hipSuccess;
	if constexpr (BATCH) {
		if (a.b.tile_rows != (uint32_t)W) {	 // the batch prep looked up another tile height
			const hipError_t e = launch_tile_doc(a.b, a.n_tiles, tbytes, s);
			if (e != hipSuccess)
				return e;
		}
	}
	const uint32_t grid = a.n_tiles < (uint32_t)resident ? a.n_tiles : (uint32_t)resident;
	// a chaser pays off once tiles outnumber the resident blocks several times
	a.chase = E::kND == 1 && !BATCH && in.chase != 0 && grid >= 2 && a.n_tiles >= 4 * grid;
	hipLaunchKernelGGL((k_scan<E, W, BATCH, NWV>), dim3(grid), dim3(NT), 0, s, a);
	return hipGetLastError();
}

// rows = 16 KiB rows per tile; 0 picks `def` (measured per evaluator: the
// bit-parallel paths amortise the look-back best over 64 KiB tiles).
template<class E, bool BATCH = false>
hipError_t launch_rows(const ScanArgs& a, int rows, int def, hipStream_t s) {
	switch (rows ? rows : def) {
	case 1: return launch_eval<E, 1, BATCH>(a, s);
	case 3: return launch_eval<E, 3, BATCH>(a, s);
	case 4: return launch_eval<E, 4, BATCH>(a, s);
	default: return launch_eval<E, 2, BATCH>(a, s);
	}
}

// One document through a bit-parallel evaluator with `waves` waves per block
// (0 or 4: the four-wave block; 6 or 8: rows of 24 or 32 KiB - the same tile
// geometry scaled, two blocks of six waves or one of eight per CU at the
// evaluators' register budget).  Rows per tile as launch_rows; the wider
// blocks take two or four rows.
template<class E>
hipError_t launch_waves(const ScanArgs& a, int rows, int def, int waves, hipStream_t s) {
	const int w = rows ? rows : def;
	if (waves == 6)
		return w >= 3 ? launch_eval<E, 4, false, 6>(a, s) : launch_eval<E, 2, false, 6>(a, s);
	if (waves == 8)
		return w >= 3 ? launch_eval<E, 4, false, 8>(a, s) : launch_eval<E, 2, false, 8>(a, s);
	return launch_rows<E>(a, rows, def, s);
}

template<bool BATCH>
hipError_t launch_general(const ScanArgs& a, int nd, bool wide, int rows, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	if (!wide && nd == 1)
		return launch_rows<GeneralEval<1, false>, BATCH>(a, rows, 2, s);
	if (!wide) {
		switch (nd) {
		case 0: return launch_eval<GeneralEval<0, false>, 2, BATCH>(a, s);
		case 2: return launch_eval<GeneralEval<2, false>, 1, BATCH>(a, s);
		case 3: return launch_eval<GeneralEval<3, false>, 1, BATCH>(a, s);
		case 4: return launch_eval<GeneralEval<4, false>, 1, BATCH>(a, s);
		case 5: return launch_eval<GeneralEval<5, false>, 1, BATCH>(a, s);
		case 6: return launch_eval<GeneralEval<6, false>, 1, BATCH>(a, s);
		case 7: return launch_eval<GeneralEval<7, false>, 1, BATCH>(a, s);
		case 8: return launch_eval<GeneralEval<8, false>, 1, BATCH>(a, s);
		}
	} else {
		switch (nd) {
		case 1: return launch_eval<GeneralEval<1, true>, 1, BATCH>(a, s);
		case 2: return launch_eval<GeneralEval<2, true>, 1, BATCH>(a, s);
		case 3: return launch_eval<GeneralEval<3, true>, 1, BATCH>(a, s);
		case 4: return launch_eval<GeneralEval<4, true>, 1, BATCH>(a, s);
		case 5: return launch_eval<GeneralEval<5, true>, 1, BATCH>(a, s);
		case 6: return launch_eval<GeneralEval<6, true>, 1, BATCH>(a, s);
		case 7: return launch_eval<GeneralEval<7, true>, 1, BATCH>(a, s);
		case 8: return launch_eval<GeneralEval<8, true>, 1, BATCH>(a, s);
		}
	}
	return hipErrorInvalidValue;
}

}  // namespace
}  // namespace pars_hip
