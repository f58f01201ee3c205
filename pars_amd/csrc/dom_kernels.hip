// dom_kernels.hip — the JSON tape: the reference's JSON DOM hook
// (bench/bench_parse.cu:41-348), which replays the structural index
// (include/pars/model/visit.hpp:35-48) through a sequential stack parser,
// computed here in parallel from the index and R6's depth-sorted brackets.
//
// The hook's state before any event depends on two things only: the innermost
// open container (its byte, or none at the top level) and the kind of the
// previous event in that container (none, key, ':', value, ','), which fixes
// the container's phase (bench_parse.cu:48-53).  Both are local once each
// bracket knows the container that encloses its pair:
//   * R6 sorts brackets stably by depth key; within one key, the children of
//     one parent form a contiguous run, and the run starts at an open whose
//     previous bracket (in document order) is an open: the parent.  A max-scan
//     of run starts over the sorted order gives every bracket its parent.
//   * The innermost container before structural t is then the last bracket
//     before t if it opens, else the parent of that (closing) bracket.
// This is the stack the hook keeps, exactly, up to the hook's first rejected
// event — and nothing after that event is consumed.  Every event is checked
// in parallel, the first rejection is a minimum, and the tape is written in
// document order from prefix counts: each gap contributes at most one value
// node, each opening bracket one container node; a container's end is the
// position of the close that follows its open at the same key, when that
// close is accepted.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace pars_hip {

namespace {

constexpr int kDomBlock = 256;
constexpr int kDomRows	= (int)(kDomTile / kDomBlock);

// symbol class of structural t (the last token has none)
enum : uint32_t { SYM_END = 0, SYM_OOPEN, SYM_AOPEN, SYM_OCLOSE, SYM_ACLOSE, SYM_COMMA, SYM_COLON, SYM_OTHER };
// what the gap before structural t holds, as the hook parses a value
// (bench_parse.cu:219-229); STR..NULL equal the tape kinds
enum : uint32_t { G_EMPTY = 0, G_BAD = 1, G_STR = 2, G_NUM = 3, G_TRUE = 4, G_FALSE = 5, G_NULL = 6 };
// the previous event in the innermost container
enum : uint32_t { P_NONE = 0, P_KEY, P_COLON, P_VALUE, P_COMMA };

// tinfo: bits [0,3) symbol class, [3,6) gap kind, bit 6 the gap parsed with
// only whitespace after it
__device__ __forceinline__ uint32_t t_sym(uint32_t w) {
	return w & 7u;
}
__device__ __forceinline__ uint32_t t_gap(uint32_t w) {
	return (w >> 3) & 7u;
}
__device__ __forceinline__ bool t_gok(uint32_t w) {
	return (w >> 6) & 1u;
}
__device__ __forceinline__ uint32_t t_nodes(uint32_t w) {
	const uint32_t s = t_sym(w);
	return (t_gap(w) >= G_STR ? 1u : 0u) + (s == SYM_OOPEN || s == SYM_AOPEN ? 1u : 0u);
}
__device__ __forceinline__ bool t_bracket(uint32_t w) {
	const uint32_t s = t_sym(w);
	return s >= SYM_OOPEN && s <= SYM_ACLOSE;
}

__device__ __forceinline__ uint32_t sym_class(uint32_t c) {
	switch (c) {
	case '{': return SYM_OOPEN;
	case '[': return SYM_AOPEN;
	case '}': return SYM_OCLOSE;
	case ']': return SYM_ACLOSE;
	case ',': return SYM_COMMA;
	case ':': return SYM_COLON;
	default: return SYM_OTHER;
	}
}

struct Gap {
	uint32_t kind, ok, beg, end, dec;
};

// The input read through a 32-byte register window: two vector loads issued
// together per aligned 32 bytes, so a typical gap (mean ~10 bytes) costs one
// load round trip, not one per 16-byte boundary it crosses (a wave waits for
// every lane's reload, and lanes cross boundaries at different bytes).  Byte
// loads serve the buffer's edges, where a window would reach outside it.
struct Bytes {
	const uint8_t* d;
	uint32_t	   lo, lim;	 // windows [a, a + 32) with lo <= a <= lim lie in the buffer
	uint32_t	   base;	 // the window's first index
	uint4		   v0, v1;

	// Indexes are 32-bit (documents are under 4 GiB).  `from` is the first
	// index read: every read is at or after it, so base = from - 64 (mod 2^32)
	// stands for "no window yet" (i - base >= 64 for every such i).
	__host__ __device__ Bytes(const uint8_t* d_, uint64_t n_, uint32_t from)
		: d(d_), lo(n_ >= 32u ? (16u - (uint32_t)((uintptr_t)d_ & 15u)) & 15u : ~0u),
		  lim(n_ >= 32u ? (uint32_t)(n_ - 32u) : 0u), base(from - 64u), v0(make_uint4(0, 0, 0, 0)),
		  v1(make_uint4(0, 0, 0, 0)) {}

	__host__ __device__ __forceinline__ uint32_t operator[](uint32_t i) {
		if (i - base >= 32u) {	// outside the window
			const uint32_t a = i - (uint32_t)(((uintptr_t)d + i) & 15u);
			if (i < lo || a > lim)
				return d[i];
			base = a;
			v0	 = *reinterpret_cast<const uint4*>(d + a);
			v1	 = *reinterpret_cast<const uint4*>(d + a + 16u);
		}
		// selects of values, not of fields (a select of fields can become a
		// select of addresses, which keeps the window in scratch)
		const uint32_t k  = i - base;
		const uint64_t h0 = (k & 8u) ? ((uint64_t)v0.w << 32 | v0.z) : ((uint64_t)v0.y << 32 | v0.x);
		const uint64_t h1 = (k & 8u) ? ((uint64_t)v1.w << 32 | v1.z) : ((uint64_t)v1.y << 32 | v1.x);
		return (uint32_t)(((k & 16u) ? h1 : h0) >> (8u * (k & 7u))) & 0xFFu;
	}
};

// The value in [b, e) as the hook parses it (bench_parse.cu:86-130, 219-332)
// — leading whitespace, one string / number / literal, trailing whitespace —
// as one DFA over byte classes, so every lane of a wave runs the same loop body
// whatever its gap holds (a branch per value kind made a wave run each kind's
// loop in turn).  Transition entries: next state in bits 0-4; bit 5: a value
// completes here; bit 6: the completing byte belongs to the value (a closing
// quote, a literal's last letter) rather than following it (a number ends at
// the first byte that cannot continue it).  The gap's end is one more class,
// K_END, stepped once after the last byte: a number ends there, anything else
// unfinished fails (the end is handled inside the automaton; a separate
// fix-up after the loop lost its store of the end offset in device code).
enum : uint8_t {
	K_SP, K_WSC, K_CTL, K_QUOTE, K_BSL, K_MINUS, K_PLUS, K_DOT, K_ZERO, K_D19, K_e, K_E, K_a, K_b, K_cd, K_f, K_HEXU,
	K_l, K_n, K_r, K_s, K_t, K_u, K_SLASH, K_OTHER, K_END, K_N
};
enum : uint8_t {
	S_LEAD, S_STR, S_ESC, S_U1, S_U2, S_U3, S_U4, S_NMIN, S_NZERO, S_NINT, S_NDOT, S_NFRAC, S_NE, S_NESIGN, S_NEXP,
	S_T1, S_T2, S_T3, S_F1, S_F2, S_F3, S_F4, S_N1, S_N2, S_N3, S_TRAIL, S_BADVAL, S_BADTRAIL, S_N
};
constexpr uint8_t kDone = 32, kIncl = 64;

struct GapTables {
	uint8_t cls[256];
	uint8_t tr[S_N * K_N];
	uint8_t esc[256];  // byte an escape decodes to (0 = not an escape letter)
};

constexpr GapTables make_gap_tables() {
	GapTables g {};
	for (int c = 0; c < 256; ++c) {
		uint8_t k = K_OTHER;
		if (c == ' ')
			k = K_SP;
		else if (c == '\t' || c == '\n' || c == '\r')
			k = K_WSC;
		else if (c < 0x20)
			k = K_CTL;
		else if (c == '"')
			k = K_QUOTE;
		else if (c == '\\')
			k = K_BSL;
		else if (c == '-')
			k = K_MINUS;
		else if (c == '+')
			k = K_PLUS;
		else if (c == '.')
			k = K_DOT;
		else if (c == '0')
			k = K_ZERO;
		else if (c >= '1' && c <= '9')
			k = K_D19;
		else if (c == 'e')
			k = K_e;
		else if (c == 'E')
			k = K_E;
		else if (c == 'a')
			k = K_a;
		else if (c == 'b')
			k = K_b;
		else if (c == 'c' || c == 'd')
			k = K_cd;
		else if (c == 'f')
			k = K_f;
		else if (c == 'A' || c == 'B' || c == 'C' || c == 'D' || c == 'F')
			k = K_HEXU;
		else if (c == 'l')
			k = K_l;
		else if (c == 'n')
			k = K_n;
		else if (c == 'r')
			k = K_r;
		else if (c == 's')
			k = K_s;
		else if (c == 't')
			k = K_t;
		else if (c == 'u')
			k = K_u;
		else if (c == '/')
			k = K_SLASH;
		g.cls[c] = k;
	}
	auto set = [&](int s, int k, int v) { g.tr[s * K_N + k] = (uint8_t)v; };
	for (int s = 0; s < S_N; ++s)
		for (int k = 0; k < K_N; ++k) set(s, k, s == S_BADTRAIL ? S_BADTRAIL : S_BADVAL);
	auto ws		= [](int k) { return k == K_SP || k == K_WSC; };
	auto digit	= [](int k) { return k == K_ZERO || k == K_D19; };
	auto hex	= [](int k) {
		   return k == K_ZERO || k == K_D19 || k == K_a || k == K_b || k == K_cd || k == K_e || k == K_E || k == K_f ||
				  k == K_HEXU;
	};
	for (int k = 0; k < K_N; ++k) {
		// before the value
		set(S_LEAD, k, ws(k) ? S_LEAD
					   : k == K_QUOTE ? S_STR
					   : k == K_MINUS ? S_NMIN
					   : k == K_ZERO  ? S_NZERO
					   : k == K_D19   ? S_NINT
					   : k == K_t	  ? S_T1
					   : k == K_f	  ? S_F1
					   : k == K_n	  ? S_N1
									  : S_BADVAL);
		// strings (bench_parse.cu:231-284)
		set(S_STR, k, k == K_QUOTE ? (S_TRAIL | kDone | kIncl)
				  : k == K_BSL	   ? S_ESC
				  : (k == K_WSC || k == K_CTL) ? S_BADVAL
											   : S_STR);
		const bool esc_letter = k == K_QUOTE || k == K_BSL || k == K_SLASH || k == K_b || k == K_f || k == K_n ||
								k == K_r || k == K_t;
		set(S_ESC, k, esc_letter ? S_STR : k == K_u ? S_U1 : S_BADVAL);
		set(S_U1, k, hex(k) ? S_U2 : S_BADVAL);
		set(S_U2, k, hex(k) ? S_U3 : S_BADVAL);
		set(S_U3, k, hex(k) ? S_U4 : S_BADVAL);
		set(S_U4, k, hex(k) ? S_STR : S_BADVAL);
		// numbers (bench_parse.cu:286-306): a number ends before the first byte
		// that cannot continue it; that byte must then be whitespace
		const int after = ws(k) ? (S_TRAIL | kDone) : (S_BADTRAIL | kDone);
		set(S_NMIN, k, k == K_ZERO ? S_NZERO : k == K_D19 ? S_NINT : S_BADVAL);
		set(S_NZERO, k, k == K_DOT ? S_NDOT : (k == K_e || k == K_E) ? S_NE : after);
		set(S_NINT, k, digit(k) ? S_NINT : k == K_DOT ? S_NDOT : (k == K_e || k == K_E) ? S_NE : after);
		set(S_NDOT, k, digit(k) ? S_NFRAC : S_BADVAL);
		set(S_NFRAC, k, digit(k) ? S_NFRAC : (k == K_e || k == K_E) ? S_NE : after);
		set(S_NE, k, (k == K_PLUS || k == K_MINUS) ? S_NESIGN : digit(k) ? S_NEXP : S_BADVAL);
		set(S_NESIGN, k, digit(k) ? S_NEXP : S_BADVAL);
		set(S_NEXP, k, digit(k) ? S_NEXP : after);
		// literals (bench_parse.cu:308-332)
		set(S_T1, k, k == K_r ? S_T2 : S_BADVAL);
		set(S_T2, k, k == K_u ? S_T3 : S_BADVAL);
		set(S_T3, k, k == K_e ? (S_TRAIL | kDone | kIncl) : S_BADVAL);
		set(S_F1, k, k == K_a ? S_F2 : S_BADVAL);
		set(S_F2, k, k == K_l ? S_F3 : S_BADVAL);
		set(S_F3, k, k == K_s ? S_F4 : S_BADVAL);
		set(S_F4, k, k == K_e ? (S_TRAIL | kDone | kIncl) : S_BADVAL);
		set(S_N1, k, k == K_u ? S_N2 : S_BADVAL);
		set(S_N2, k, k == K_l ? S_N3 : S_BADVAL);
		set(S_N3, k, k == K_l ? (S_TRAIL | kDone | kIncl) : S_BADVAL);
		// after the value: whitespace only
		set(S_TRAIL, k, ws(k) ? S_TRAIL : S_BADTRAIL);
	}
	// the gap's end
	for (int st = 0; st < S_N; ++st) set(st, K_END, st == S_BADTRAIL ? S_BADTRAIL : S_BADVAL);
	set(S_LEAD, K_END, S_LEAD);
	set(S_TRAIL, K_END, S_TRAIL);
	for (int st : {S_NZERO, S_NINT, S_NFRAC, S_NEXP}) set(st, K_END, S_TRAIL | kDone);
	g.esc['"']	= '"';
	g.esc['\\'] = '\\';
	g.esc['/']	= '/';
	g.esc['b']	= '\b';
	g.esc['f']	= '\f';
	g.esc['n']	= '\n';
	g.esc['r']	= '\r';
	g.esc['t']	= '\t';
	return g;
}

__constant__ GapTables kGapTables = make_gap_tables();

// The tables in LDS (every block copies them first).
__device__ __forceinline__ void load_gap_tables(GapTables& L) {
	const uint32_t* src = reinterpret_cast<const uint32_t*>(&kGapTables);
	uint32_t*		dst = reinterpret_cast<uint32_t*>(&L);
	for (uint32_t i = threadIdx.x; i < sizeof(GapTables) / 4; i += blockDim.x) dst[i] = src[i];
	__syncthreads();
}

// Bytes of a dword that end a string's plain run - '"', '\\' or a byte below
// 0x20 - as bit 7 of each byte.  Exact for the lowest such byte (a borrow can
// flag a byte only above a truly flagged one).
__host__ __device__ __forceinline__ uint32_t str_stops(uint32_t x) {
	const uint32_t q = x ^ 0x22222222u, b = x ^ 0x5C5C5C5Cu;
	return (((q - 0x01010101u) & ~q) | ((b - 0x01010101u) & ~b) | ((x - 0x20202020u) & ~x)) & 0x80808080u;
}

// How many bytes from j on (before e) continue a string unchanged (S_STR to
// S_STR, one decoded byte each), as far as the window that holds j reaches; 0
// when j is not in a window (the buffer's edges, read byte by byte).
__host__ __device__ __forceinline__ uint32_t str_run(Bytes& d, uint32_t j, uint32_t e) {
	(void)d[j];	 // the window holding j, when there is one
	const uint32_t k = j - d.base;
	if (k >= 32u)
		return 0u;
	const uint32_t w[8] = {d.v0.x, d.v0.y, d.v0.z, d.v0.w, d.v1.x, d.v1.y, d.v1.z, d.v1.w};
	uint32_t	   m	= 0;  // bit i: window byte i ends a run
#pragma unroll
	for (int q = 0; q < 8; ++q) m |= (((str_stops(w[q]) >> 7) * 0x00204081u) >> 21 & 0xFu) << (4 * q);
	m &= ~0u << k;
	const uint32_t stop = m ? (uint32_t)__builtin_ctz(m) : 32u;	 // window-relative
	const uint32_t lim	= e - d.base < 32u ? e - d.base : 32u;
	return (stop < lim ? stop : lim) - k;
}

// kind, whether it parsed with only whitespace after it, its first byte, one
// past its last byte (0 for a string that did not close), and the decoded
// bytes a string mixed into the hash.  EMIT writes those bytes to `out`.
template<bool EMIT>
__host__ __device__ __forceinline__ Gap parse_gap(const GapTables& L, Bytes& d, uint32_t b, uint32_t e, uint8_t* out) {
	Gap		 g {G_EMPTY, 0u, 0u, 0u, 0u};
	uint32_t s = S_LEAD, dec = 0, v = 0, first = K_OTHER;
	bool	 done = false;
	for (uint32_t j = b;; ++j) {  // through j == e, the gap's end class
		if constexpr (!EMIT) {
			// inside a string, a run of plain bytes only counts decoded bytes:
			// skip to the run's end (the window's end at most)
			if (s == S_STR && j < e) {
				const uint32_t r = str_run(d, j, e);
				j += r;
				dec += r;
			}
		}
		const bool	   end = j == e;
		const uint32_t c   = end ? 0u : d[j];
		const uint32_t k   = end ? (uint32_t)K_END : (uint32_t)L.cls[c];
		const uint32_t x = L.tr[s * K_N + k];
		const uint32_t t = x & 31u;
		// every step computes every case and selects (no branches: the lanes
		// of a wave are in different states)
		const bool lead = s == S_LEAD && t != S_LEAD;
		g.beg			= lead ? j : g.beg;
		first			= lead ? k : first;
		// decoded bytes: a plain string byte or an escape letter (one), a \u
		// unit (1-3 UTF-8 bytes)
		uint32_t nw = (t == S_STR && (s == S_STR || s == S_ESC)) ? 1u : 0u;
		// \u units: four hex digits accumulate from S_U1 on; the fourth
		// (S_U4 → S_STR) decodes to 1-3 UTF-8 bytes.  Rare: a wave skips this
		// when none of its lanes is inside one.
		uint32_t vn = 0;
		bool	 u4 = false;
		if (s >= S_U1 && s <= S_U4) {
			vn = (v << 4) | (c <= '9' ? c - '0' : (c | 0x20u) - ('a' - 10u));
			u4 = s == S_U4 && t == S_STR;
			if (u4)
				nw = vn < 0x80u ? 1u : vn < 0x800u ? 2u : 3u;
			v = vn;
		}
		if constexpr (EMIT) {
			uint32_t w = s == S_ESC ? (uint32_t)L.esc[c] : c;
			if (u4) {
				if (vn < 0x80u)
					w = vn;
				else if (vn < 0x800u)
					w = (0xC0u | (vn >> 6)) | ((0x80u | (vn & 0x3Fu)) << 8);
				else
					w = (0xE0u | (vn >> 12)) | ((0x80u | ((vn >> 6) & 0x3Fu)) << 8) | ((0x80u | (vn & 0x3Fu)) << 16);
			}
			for (uint32_t q = 0; q < nw; ++q) out[dec + q] = (uint8_t)(w >> (8 * q));
		}
		v = t == S_U1 ? 0u : v;
		dec += nw;
		const bool fin = (x & kDone) != 0u;
		done		   = done || fin;
		g.end		   = fin ? ((x & kIncl) ? j + 1u : j) : g.end;
		s			   = t;
		if (end || s >= S_BADVAL)
			break;
	}
	if (first == K_OTHER && s == S_LEAD)
		return g;  // whitespace only
	switch (first) {
	case K_QUOTE: g.kind = G_STR; break;
	case K_MINUS:
	case K_ZERO:
	case K_D19: g.kind = done ? G_NUM : G_BAD; break;
	case K_t: g.kind = done ? G_TRUE : G_BAD; break;
	case K_f: g.kind = done ? G_FALSE : G_BAD; break;
	case K_n: g.kind = done ? G_NULL : G_BAD; break;
	default: g.kind = G_BAD; break;
	}
	if (!done)
		g.end = 0;
	g.dec = dec;
	g.ok  = done && s == S_TRAIL;
	return g;
}

__device__ __forceinline__ void gap_bounds(const DomArgs& a, uint64_t t, uint64_t& b, uint64_t& e) {
	b = t ? (uint64_t)a.positions[t - 1] + 1u : 0u;
	e = t < a.n_struct ? (uint64_t)a.positions[t] : a.n;
}

// Exclusive scan of three fields over the block (row by row); returns totals.
__device__ __forceinline__ void block_excl3(uint32_t (&v)[3], uint32_t (&tot)[3], uint32_t (*s)[3]) {
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint32_t  inc[3];
#pragma unroll
	for (int f = 0; f < 3; ++f) inc[f] = wave_incl_sum(v[f]);
	if (lane == 63)
#pragma unroll
		for (int f = 0; f < 3; ++f) s[w][f] = inc[f];
	__syncthreads();
#pragma unroll
	for (int f = 0; f < 3; ++f) {
		uint32_t pre = 0, t = 0;
#pragma unroll
		for (int q = 0; q < kDomBlock / 64; ++q) {
			const uint32_t x = s[q][f];
			pre += q < w ? x : 0u;
			t += x;
		}
		v[f]   = pre + inc[f] - v[f];
		tot[f] = t;
	}
	__syncthreads();
}

__device__ __forceinline__ int32_t wave_incl_max(int32_t v) {
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const int32_t y = __shfl_up(v, o, 64);
		if (lane >= o)
			v = max(v, y);
	}
	return v;
}

// Inclusive max-scan over the block; `carry` is folded in and updated to the
// block's maximum.
__device__ __forceinline__ int32_t block_incl_max(int32_t v, int32_t& carry, int32_t* s) {
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int32_t inc = wave_incl_max(v);
	if (lane == 63)
		s[w] = inc;
	__syncthreads();
	int32_t pre = carry, all = carry;
#pragma unroll
	for (int q = 0; q < kDomBlock / 64; ++q) {
		const int32_t x = s[q];
		if (q < w)
			pre = max(pre, x);
		all = max(all, x);
	}
	__syncthreads();
	carry = all;
	return max(pre, inc);
}

// Exclusive max-scan over the block (-1 for the first thread).
__device__ __forceinline__ int32_t block_excl_max(int32_t v, int32_t* s) {
	const int	  lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const int32_t inc  = wave_incl_max(v);
	if (lane == 63)
		s[w] = inc;
	__syncthreads();
	int32_t pre = -1;
#pragma unroll
	for (int q = 0; q < kDomBlock / 64; ++q)
		if (q < w)
			pre = max(pre, s[q]);
	__syncthreads();
	int32_t wex = __shfl_up(inc, 1, 64);
	if (lane == 0)
		wex = -1;
	return max(pre, wex);
}

__device__ __forceinline__ bool br_open(const DomArgs& a, uint64_t b) {
	return a.plan->br_flags[a.bchar[b]] & 1u;
}

// ---------------------------------------------------------------------------
__global__ void k_dom_reset(DomSummary* ds) {
	if (threadIdx.x == 0) {
		ds->fail_key   = kDomNone;
		ds->e1		   = kDomNone;
		ds->e2		   = kDomNone;
		ds->fail_event = kDomNone;
		ds->n_nodes	   = 0;
		ds->arena_bytes = 0;
		ds->ok		   = 0;
		ds->partial	   = 0;
	}
}

// The per-gap kernels take kSortTok tokens per block and hand them to lanes
// in order of their work (gap bytes to walk), so the 64 lanes of a wave walk
// gaps of about one length: a wave lasts as long as its longest gap, and in
// document order that is ~4.5× the mean on JSON.
constexpr int kSortTok = (int)kDomSub;
constexpr int kSortPer = kSortTok / kDomBlock;
constexpr int kBins	   = 64;
static_assert(kDomTile % kSortTok == 0, "a block's tokens lie in one sum group");

// work → bin: exact below 32 bytes, 8-byte bins to 280, one bin beyond
__device__ __forceinline__ uint32_t work_bin(uint32_t w) {
	return w < 32u ? w : (w < 32u + 31u * 8u ? 32u + (w - 32u) / 8u : 63u);
}

// Counting sort of the block's tokens (local index i * kDomBlock + thread)
// by bin: perm[slot] = local index.  Order within a bin is arbitrary, which
// changes no result: every token's outputs depend on that token alone.
__device__ __forceinline__ void sort_by_work(const uint32_t (&work)[kSortPer], uint16_t* perm, uint32_t* bins) {
	if (threadIdx.x < kBins)
		bins[threadIdx.x] = 0;
	__syncthreads();
	uint32_t bin[kSortPer], rank[kSortPer];
#pragma unroll
	for (int i = 0; i < kSortPer; ++i) {
		bin[i]	= work_bin(work[i]);
		rank[i] = atomicAdd(&bins[bin[i]], 1u);
	}
	__syncthreads();
	if (threadIdx.x < kBins) {	// wave 0: kBins == 64 lanes
		const uint32_t c = bins[threadIdx.x];
		bins[threadIdx.x] = wave_incl_sum(c) - c;
	}
	__syncthreads();
#pragma unroll
	for (int i = 0; i < kSortPer; ++i) perm[bins[bin[i]] + rank[i]] = (uint16_t)(i * kDomBlock + threadIdx.x);
	__syncthreads();
}

// 1. Per token: symbol class and the gap's value; the counts summed per
// 4096-token group (the group sums start at zero).  Gaps are parsed in work
// order and their results go back to document order through LDS.
__global__ __launch_bounds__(kDomBlock) void k_dom_tok(DomArgs a) {
	__shared__ uint32_t	 s[kDomBlock / 64][3];
	__shared__ GapTables L;
	__shared__ uint16_t	 perm[kSortTok];
	__shared__ uint32_t	 bins[kBins];
	__shared__ uint32_t	 r_beg[kSortTok], r_end[kSortTok], r_dec[kSortTok];
	__shared__ uint8_t	 r_gap[kSortTok];
	load_gap_tables(L);
	const uint64_t ntok = a.n_struct + 1;
	const uint64_t base = (uint64_t)blockIdx.x * kSortTok;
	uint32_t	   work[kSortPer];
#pragma unroll
	for (int i = 0; i < kSortPer; ++i) {
		const uint32_t j = (uint32_t)(i * kDomBlock) + threadIdx.x;
		const uint64_t t = base + j;
		work[i]			 = 0;
		if (t < ntok) {
			uint64_t b, e;
			gap_bounds(a, t, b, e);
			work[i]	 = (uint32_t)(e - b);
			r_beg[j] = (uint32_t)b;	 // the gap's bounds, replaced by the results
			r_end[j] = (uint32_t)e;
		}
	}
	sort_by_work(work, perm, bins);
#pragma unroll 1
	for (int i = 0; i < kSortPer; ++i) {
		const uint32_t j = perm[i * kDomBlock + threadIdx.x];
		if (base + j < ntok) {
			Bytes	  rd(a.data, a.n, r_beg[j]);
			const Gap g = parse_gap<false>(L, rd, r_beg[j], r_end[j], nullptr);
			r_beg[j]	= g.beg;
			r_end[j]	= g.end;
			r_dec[j]	= g.dec;
			r_gap[j]	= (uint8_t)((g.kind << 3) | (g.ok << 6));
		}
	}
	__syncthreads();
	uint32_t sum[3] = {0u, 0u, 0u};
#pragma unroll
	for (int i = 0; i < kSortPer; ++i) {
		const uint32_t j = (uint32_t)(i * kDomBlock) + threadIdx.x;
		const uint64_t t = base + j;
		if (t < ntok) {
			const uint32_t sym = t < a.n_struct ? sym_class(a.data[a.positions[t]]) : SYM_END;
			const uint32_t w   = sym | r_gap[j];
			a.tinfo[t]		   = (uint8_t)w;
			a.gbeg[t]		   = r_beg[j];
			a.gend[t]		   = r_end[j];
			a.gdec[t]		   = r_dec[j];
			sum[0] += t_nodes(w);
			sum[1] += r_dec[j];
			sum[2] += t_bracket(w) ? 1u : 0u;
		}
	}
	uint32_t tot[3];
	block_excl3(sum, tot, s);
	const uint64_t grp = base / kDomTile;
	if (threadIdx.x == 0) {
#pragma unroll
		for (int f = 0; f < 3; ++f)
			if (tot[f])
				atomicAdd(&a.tblk[grp * 3 + f], tot[f]);
		a.sub[blockIdx.x * 3u + 1u] = tot[1];  // this block's decoded bytes (the check pass makes it a prefix)
	}
}

// 2. Exclusive scan of the block sums (one block; each thread a contiguous run).
__global__ __launch_bounds__(kDomBlock) void k_dom_tok_top(DomArgs a) {
	__shared__ uint32_t s[kDomBlock / 64][3];
	const uint64_t		nb	= a.n_tblk;
	const uint64_t		per = (nb + kDomBlock - 1) / kDomBlock;
	const uint64_t		lo = threadIdx.x * per, hi = lo + per < nb ? lo + per : nb;
	uint32_t			v[3] = {0u, 0u, 0u};
	for (uint64_t k = lo; k < hi; ++k)
#pragma unroll
		for (int f = 0; f < 3; ++f) v[f] += a.tblk[k * 3 + f];
	uint32_t tot[3];
	block_excl3(v, tot, s);
	for (uint64_t k = lo; k < hi; ++k)
#pragma unroll
		for (int f = 0; f < 3; ++f) {
			const uint32_t x  = a.tblk[k * 3 + f];
			a.tblk[k * 3 + f] = v[f];
			v[f] += x;
		}
}

// 3. Run starts over the sorted brackets: block maxima.
__device__ __forceinline__ int32_t head_at(const DomArgs& a, uint64_t p) {
	const uint32_t x = a.order[p];
	return (br_open(a, x) && (x == 0 || br_open(a, x - 1u))) ? (int32_t)p : -1;
}

__global__ __launch_bounds__(kDomBlock) void k_dom_heads(DomArgs a) {
	__shared__ int32_t s[kDomBlock / 64];
	const uint64_t	   base = (uint64_t)blockIdx.x * kDomTile;
	int32_t			   m	= -1;
	for (int r = 0; r < kDomRows; ++r) {
		const uint64_t p = base + (uint64_t)r * kDomBlock + threadIdx.x;
		if (p < a.nb)
			m = max(m, head_at(a, p));
	}
	int32_t carry = -1;
	(void)block_incl_max(m, carry, s);
	if (threadIdx.x == 0)
		a.hblk[blockIdx.x] = carry;
}

// 4. Exclusive max over the block maxima (one block).
__global__ __launch_bounds__(kDomBlock) void k_dom_heads_top(DomArgs a) {
	__shared__ int32_t s[kDomBlock / 64];
	const uint64_t	   nb  = a.n_hblk;
	const uint64_t	   per = (nb + kDomBlock - 1) / kDomBlock;
	const uint64_t	   lo = threadIdx.x * per, hi = lo + per < nb ? lo + per : nb;
	int32_t			   m  = -1;
	for (uint64_t k = lo; k < hi; ++k) m = max(m, a.hblk[k]);
	int32_t run = block_excl_max(m, s);
	for (uint64_t k = lo; k < hi; ++k) {
		const int32_t x = a.hblk[k];
		a.hblk[k]		= run;
		run				= max(run, x);
	}
}

// 5. Every bracket's parent: the bracket before its run's start.
__global__ __launch_bounds__(kDomBlock) void k_dom_parent(DomArgs a) {
	__shared__ int32_t s[kDomBlock / 64];
	const uint64_t	   base	 = (uint64_t)blockIdx.x * kDomTile;
	int32_t			   carry = a.hblk[blockIdx.x];
	for (int r = 0; r < kDomRows; ++r) {
		const uint64_t p  = base + (uint64_t)r * kDomBlock + threadIdx.x;
		const bool	   in = p < a.nb;
		const int32_t  h  = block_incl_max(in ? head_at(a, p) : -1, carry, s);
		if (in)
			a.parent[a.order[p]] = h >= 0 ? (int32_t)a.order[h] - 1 : -1;
	}
}

__device__ __forceinline__ uint32_t prev_of_symbol(uint32_t sym) {
	switch (sym) {
	case SYM_OCLOSE:
	case SYM_ACLOSE: return P_VALUE;
	case SYM_COMMA: return P_COMMA;
	case SYM_COLON: return P_COLON;
	default: return P_NONE;	 // document start, an opening bracket (and a rejected symbol)
	}
}

__device__ __forceinline__ void wave_atomic_min(unsigned long long* p, unsigned long long v) {
#pragma unroll
	for (int o = 32; o >= 1; o >>= 1) {
		const unsigned long long y = __shfl_xor(v, o, 64);
		v						   = y < v ? y : v;
	}
	if ((threadIdx.x & 63) == 0 && v != kDomNone)
		atomicMin(p, v);
}

// The two smallest of a set of distinct events, kDomNone where missing.
struct Two {
	unsigned long long m1, m2;
	__device__ __forceinline__ void add(unsigned long long v) {	 // v above every value added so far
		if (m1 == kDomNone)
			m1 = v;
		else if (m2 == kDomNone)
			m2 = v;
	}
	__device__ __forceinline__ void merge(unsigned long long b1, unsigned long long b2) {
		const unsigned long long lo = m1 < b1 ? m1 : b1, hi = m1 < b1 ? b1 : m1, r2 = m2 < b2 ? m2 : b2;
		m1							= lo;
		m2							= hi < r2 ? hi : r2;
	}
};

// 6. Per token exclusive prefixes (nodes, decoded bytes, brackets) and every
// event against the hook's rules (bench_parse.cu:86-194), in one pass over a
// 4096-token group per block: the check needs the bracket prefix the scan
// has just produced.  Top-level value events are collected: every one after
// the first is rejected (the hook's top_done), so the final step needs the
// two smallest; each group keeps its own two.
__device__ __forceinline__ void check_token(const DomArgs& a, uint64_t t, uint32_t w, uint32_t brk,
											unsigned long long& fk, Two& top) {
	// the innermost open container before structural t
	int64_t		   cont = -1;
	const int64_t  b	= (int64_t)brk - 1;
	if (b >= 0)
		cont = br_open(a, (uint64_t)b) ? b : (int64_t)a.parent[b];
	const uint32_t cc	= cont >= 0 ? a.bchar[cont] : 0u;
	const uint32_t pt	= t ? prev_of_symbol(t_sym(a.tinfo[t - 1])) : P_NONE;
	const uint32_t gk	= t_gap(w);
	uint32_t	   p	= pt;
	uint8_t		   flag = 0;
	auto reject = [&](unsigned long long ev, uint32_t partial) {
		const unsigned long long k = (ev << 1) | partial;
		fk						   = k < fk ? k : fk;
	};
	if (gk != G_EMPTY) {
		const unsigned long long eg = 2ull * t;
		const uint32_t			 node = gk >= G_STR ? 1u : 0u;
		if (cc == '{' && (pt == P_NONE || pt == P_COMMA)) {	 // a key
			if (gk != G_STR)
				reject(eg, 0u);
			else if (!t_gok(w))
				reject(eg, 1u);
			p = P_KEY;
		} else {
			if (cc == 0u)
				flag |= 1u;
			else if (!(cc == '{' ? pt == P_COLON : (pt == P_NONE || pt == P_COMMA)))
				reject(eg, 0u);
			if (gk == G_BAD)
				reject(eg, 0u);
			else if (!t_gok(w))
				reject(eg, node);
			p = P_VALUE;
		}
	}
	if (t < a.n_struct) {
		const unsigned long long es = 2ull * t + 1ull;
		bool					 ok = true;
		switch (t_sym(w)) {
		case SYM_OOPEN:
		case SYM_AOPEN:
			if (cc == 0u)
				flag |= 2u;
			else
				ok = cc == '{' ? p == P_COLON : (p == P_NONE || p == P_COMMA);
			break;
		case SYM_OCLOSE: ok = cc == '{' && (p == P_NONE || p == P_COMMA || p == P_VALUE); break;
		case SYM_ACLOSE: ok = cc == '[' && (p == P_NONE || p == P_COMMA || p == P_VALUE); break;
		case SYM_COLON: ok = cc == '{' && p == P_KEY; break;
		case SYM_COMMA: ok = cc != 0u && p == P_VALUE; break;
		default: ok = false; break;
		}
		if (!ok)
			reject(es, 0u);
	}
	if (flag & 1u)
		top.add(2ull * t);
	if (flag & 2u)
		top.add(2ull * t + 1ull);
}

__global__ __launch_bounds__(kDomBlock) void k_dom_apply_check(DomArgs a) {
	__shared__ uint32_t s[kDomBlock / 64][3];
	const uint64_t		ntok = a.n_struct + 1;
	const uint64_t		base = (uint64_t)blockIdx.x * kDomTile;
	uint32_t			carry[3];
#pragma unroll
	for (int f = 0; f < 3; ++f) carry[f] = a.tblk[blockIdx.x * 3 + f];
	// the group's kDomSub-token blocks' decoded-byte totals (left by k_dom_tok
	// in the table this pass turns into prefixes): the per-token counts are
	// not read again
	constexpr int		SUBS = (int)(kDomTile / kDomSub);
	__shared__ uint32_t s_dsub[SUBS];
	if (threadIdx.x < SUBS) {
		const uint64_t sb	= (uint64_t)blockIdx.x * SUBS + threadIdx.x;
		s_dsub[threadIdx.x] = sb * kDomSub < ntok ? a.sub[sb * 3u + 1u] : 0u;
	}
	__syncthreads();
	unsigned long long fk  = kDomNone;
	Two				   top = {kDomNone, kDomNone};	// a thread's tokens come in increasing order
	for (int r = 0; r < kDomRows; ++r) {
		const uint64_t t  = base + (uint64_t)r * kDomBlock + threadIdx.x;
		const bool	   in = t < ntok;
		if (r % kSortPer == 0 && threadIdx.x == 0 && in)  // a kDomSub-token block starts here
#pragma unroll
			for (int f = 0; f < 3; ++f) a.sub[(t / kDomSub) * 3 + f] = carry[f];
		uint32_t v[3] = {0u, 0u, 0u}, w = 0;
		if (in) {
			w	 = a.tinfo[t];
			v[0] = t_nodes(w);
			v[2] = t_bracket(w) ? 1u : 0u;
		}
		uint32_t tot[3];
		block_excl3(v, tot, s);
		if (in)
			check_token(a, t, w, carry[2] + v[2], fk, top);
		carry[0] += tot[0];
		carry[2] += tot[2];
		if (r % kSortPer == kSortPer - 1)
			carry[1] += s_dsub[r / kSortPer];
	}
	wave_atomic_min(&a.ds->fail_key, fk);
	__shared__ unsigned long long s_top[kDomBlock / 64][2];
#pragma unroll
	for (int o = 32; o >= 1; o >>= 1) top.merge(__shfl_xor(top.m1, o, 64), __shfl_xor(top.m2, o, 64));
	if ((threadIdx.x & 63) == 0) {
		s_top[threadIdx.x >> 6][0] = top.m1;
		s_top[threadIdx.x >> 6][1] = top.m2;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int q = 1; q < kDomBlock / 64; ++q) top.merge(s_top[q][0], s_top[q][1]);
		a.top2[2u * blockIdx.x]		 = top.m1;
		a.top2[2u * blockIdx.x + 1u] = top.m2;
	}
}

// 7. Where the hook stopped, how many nodes and decoded bytes it produced:
// the first two top-level value events over every group, then one thread.
__global__ __launch_bounds__(kDomBlock) void k_dom_final(DomArgs a) {
	__shared__ unsigned long long s_top[kDomBlock / 64][2];
	Two							  top = {kDomNone, kDomNone};
	for (uint64_t g = threadIdx.x; g < a.n_tblk; g += kDomBlock) top.merge(a.top2[2u * g], a.top2[2u * g + 1u]);
#pragma unroll
	for (int o = 32; o >= 1; o >>= 1) top.merge(__shfl_xor(top.m1, o, 64), __shfl_xor(top.m2, o, 64));
	if ((threadIdx.x & 63) == 0) {
		s_top[threadIdx.x >> 6][0] = top.m1;
		s_top[threadIdx.x >> 6][1] = top.m2;
	}
	__syncthreads();
	__shared__ unsigned long long s_fk;
	__shared__ uint64_t			  s_pre[kDomBlock / 64][2];
	DomSummary*					  ds = a.ds;
	if (threadIdx.x == 0) {
		for (int q = 1; q < kDomBlock / 64; ++q) top.merge(s_top[q][0], s_top[q][1]);
		ds->e1						= top.m1;
		ds->e2						= top.m2;
		unsigned long long		 fk = ds->fail_key;
		const unsigned long long e2 = top.m2;
		if (e2 != kDomNone && (e2 << 1) < fk)
			fk = e2 << 1;  // a second top-level value is rejected before it is parsed
		s_fk = fk;
	}
	__syncthreads();
	// the token the hook stopped at (the last one when every event is
	// accepted), and the node and decoded-byte prefixes before it: its
	// block's entry in the table plus the tokens of that block before it
	const unsigned long long fk = s_fk;
	const uint64_t			 t	= fk == kDomNone ? a.n_struct : (fk >> 2);
	const uint64_t			 t0 = (t / kDomSub) * kDomSub;
	uint64_t				 pn = 0, pd = 0;
	for (uint64_t u = t0 + threadIdx.x; u < t; u += kDomBlock) {
		pn += t_nodes(a.tinfo[u]);
		pd += a.gdec[u];
	}
#pragma unroll
	for (int o = 32; o >= 1; o >>= 1) {
		pn += __shfl_xor(pn, o, 64);
		pd += __shfl_xor(pd, o, 64);
	}
	if ((threadIdx.x & 63) == 0) {
		s_pre[threadIdx.x >> 6][0] = pn;
		s_pre[threadIdx.x >> 6][1] = pd;
	}
	__syncthreads();
	if (threadIdx.x != 0)
		return;
	uint64_t nodes_x = a.sub[(t / kDomSub) * 3], arena_x = a.sub[(t / kDomSub) * 3 + 1];
	for (int q = 0; q < kDomBlock / 64; ++q) {
		nodes_x += s_pre[q][0];
		arena_x += s_pre[q][1];
	}
	const uint32_t w = a.tinfo[t];
	if (fk == kDomNone) {  // the last token: every event accepted
		ds->fail_event	= kDomNone;
		ds->partial		= 0;
		ds->n_nodes		= nodes_x + t_nodes(w);
		ds->arena_bytes = arena_x + a.gdec[t];
		ds->ok			= a.sum->last_bal == 0 && ds->e1 != kDomNone;  // bench_parse.cu:196-201
		return;
	}
	const unsigned long long ev		 = fk >> 1;
	const uint32_t			 partial = (uint32_t)(fk & 1u);
	ds->fail_event					 = ev;
	ds->partial						 = partial;
	ds->ok							 = 0;
	if ((ev & 1u) == 0) {  // the gap before structural t
		ds->n_nodes		= nodes_x + partial;
		ds->arena_bytes = arena_x + (partial ? a.gdec[t] : 0u);
	} else {  // structural t, after its accepted gap
		ds->n_nodes		= nodes_x + (t_gap(w) >= G_STR ? 1u : 0u);
		ds->arena_bytes = arena_x + a.gdec[t];
	}
}

// 8. The tape, in document order, and the decoded string bytes.
// Node records in document order; strings' decoded bytes afterwards, in work
// order (bytes to copy, or gap bytes to decode again).  A block's strings own
// one contiguous run of the arena: the lanes write their bytes into LDS and
// the block stores the run with aligned 16-byte stores (byte stores from 64
// lanes into 64 strings cost one memory request per byte).  A run longer
// than the staging buffer is written directly.
enum : uint8_t { H_NONE = 0, H_COPY, H_DECODE };
constexpr uint32_t kStage = 8192;	// staged decoded bytes per block

__global__ __launch_bounds__(kDomBlock) void k_dom_tape(DomArgs a) {
	__shared__ GapTables L;
	__shared__ uint16_t	 perm[kSortTok];
	__shared__ uint32_t	 bins[kBins];
	__shared__ uint8_t	 s_how[kSortTok];
	__shared__ __align__(16) uint8_t stage[kStage];
	__shared__ uint32_t	 s_arx[kSortTok];  // each token's decoded-byte prefix
	__shared__ uint32_t	 s3[kDomBlock / 64][3];
	load_gap_tables(L);
	const uint64_t			 ntok = a.n_struct + 1;
	const unsigned long long fe	  = a.ds->fail_event;  // kDomNone: none
	const uint32_t			 part = a.ds->partial;
	const uint64_t			 base = (uint64_t)blockIdx.x * kSortTok;
	uint32_t				 carry[3];	// node, decoded-byte and bracket prefixes, row by row
#pragma unroll
	for (int f = 0; f < 3; ++f) carry[f] = a.sub[blockIdx.x * 3u + f];
	const uint64_t a0 = carry[1];  // the block's run of the arena starts here
	uint32_t	   work[kSortPer];
#pragma unroll
	for (int i = 0; i < kSortPer; ++i) {
		const uint32_t j  = (uint32_t)(i * kDomBlock) + threadIdx.x;
		const uint64_t t  = base + j;
		const bool	   in = t < ntok;
		const uint32_t w  = in ? (uint32_t)a.tinfo[t] : 0u;
		const uint32_t gd = in ? a.gdec[t] : 0u;
		uint32_t	   v[3] = {in ? t_nodes(w) : 0u, gd, in && t_bracket(w) ? 1u : 0u}, tot[3];
		block_excl3(v, tot, s3);  // every thread, before any skips its token
		const uint32_t nodes_x = carry[0] + v[0], arena_x = carry[1] + v[1], brk_x = carry[2] + v[2];
#pragma unroll
		for (int f = 0; f < 3; ++f) carry[f] += tot[f];
		work[i]		 = 0;
		s_how[j]	 = H_NONE;
		s_arx[j]	 = arena_x;
		const unsigned long long eg = 2ull * t;
		if (!in || eg > fe)
			continue;
		const uint32_t gk  = t_gap(w);
		uint64_t	   idx = nodes_x;
		if (gk >= G_STR) {
			if (eg < fe || part) {
				if (idx < a.cap_nodes) {
					a.kind[idx] = (uint8_t)gk;
					a.nbeg[idx] = a.gbeg[t];
					a.nend[idx] = a.gend[t];
					a.soff[idx] = arena_x;
				}
				const uint32_t dec = gd;
				if (gk == G_STR && (uint64_t)arena_x + dec <= a.cap_arena) {
					const uint32_t gb = a.gbeg[t], ge = a.gend[t];
					// closed, and as long decoded as between its quotes: no escape
					// (every escape decodes shorter), the bytes as they are
					if (ge && ge - gb - 2u == dec) {
						work[i]	 = dec + 1u;
						s_how[j] = H_COPY;
					} else {
						uint64_t b, e;
						gap_bounds(a, t, b, e);
						work[i]	 = (uint32_t)(e - b) + 1u;
						s_how[j] = H_DECODE;
					}
				}
			}
			++idx;
		}
		const uint32_t sym = t_sym(w);
		if ((sym == SYM_OOPEN || sym == SYM_AOPEN) && eg + 1ull < fe && idx < a.cap_nodes) {
			// the close that follows this open at the same key is its partner
			// while the hook accepts it
			const uint32_t tc  = a.close[brk_x];
			const uint32_t end = tc != ~0u && 2ull * tc + 1ull < fe ? a.positions[tc] : 0u;
			a.kind[idx] = sym == SYM_OOPEN ? 0u : 1u;
			a.nbeg[idx] = a.positions[t];
			a.nend[idx] = end;
			a.soff[idx] = arena_x + gd;
		}
	}
	// the block's run of the arena, cut at the tape's end (strings after the
	// hook's first rejected event are not written)
	const uint64_t lim = a.ds->arena_bytes, ae = carry[1];
	const uint64_t a1  = ae < lim ? (ae > a0 ? ae : a0) : (lim > a0 ? lim : a0);
	sort_by_work(work, perm, bins);	 // (its barriers also publish s_how and s_arx)
	const uint64_t c0 = a0 & ~15ull;  // stage[g - c0] holds arena byte g: LDS and HBM agree mod 16
	const bool	   staged = a1 - c0 <= kStage;	// block-uniform
	// one loop per destination, so the staged one stores to LDS directly
	// rather than through generic addresses
	auto strings = [&](auto dst) {
#pragma unroll 1
		for (int i = 0; i < kSortPer; ++i) {
			const uint32_t j   = perm[i * kDomBlock + threadIdx.x];
			const uint32_t how = s_how[j];
			if (how == H_NONE)
				continue;
			const uint64_t t   = base + j;
			uint8_t*	   out = dst(s_arx[j]);
			if (how == H_COPY) {
				const uint32_t src = a.gbeg[t] + 1u;
				const uint32_t dec = a.gdec[t];
				Bytes		   rd(a.data, a.n, src);
				for (uint32_t q = 0; q < dec; ++q) out[q] = (uint8_t)rd[src + q];
			} else {
				uint64_t b, e;
				gap_bounds(a, t, b, e);
				Bytes rd(a.data, a.n, (uint32_t)b);
				(void)parse_gap<true>(L, rd, (uint32_t)b, (uint32_t)e, out);
			}
		}
	};
	if (!staged) {
		strings([&](uint64_t off) { return a.arena + off; });
	} else {
		strings([&](uint64_t off) { return stage + (off - c0); });
		__syncthreads();
		const uint64_t c1 = (a1 + 15u) & ~15ull;
		for (uint64_t c = c0 + 16u * threadIdx.x; c < c1; c += 16u * kDomBlock) {
			const uint32_t so = (uint32_t)(c - c0);
			if (c >= a0 && c + 16u <= a1)
				*reinterpret_cast<uint4*>(a.arena + c) = *reinterpret_cast<const uint4*>(stage + so);
			else  // the run's ends share 16 bytes with the neighbouring blocks' runs
				for (uint32_t q = 0; q < 16u; ++q)
					if (c + q >= a0 && c + q < a1)
						a.arena[c + q] = stage[so + q];
		}
	}
}


}  // namespace

hipError_t launch_json_tape(const DomArgs& a, hipStream_t s) {
	const uint64_t ntok = a.n_struct + 1;
	const unsigned gs	= (unsigned)((ntok + kSortTok - 1) / kSortTok);		// kSortTok tokens per block
	hipLaunchKernelGGL(k_dom_reset, dim3(1), dim3(64), 0, s, a.ds);
	hipError_t e = launch_fill_i32(reinterpret_cast<int32_t*>(a.tblk), a.n_tblk * 3, 0, s);
	if (e != hipSuccess)
		return e;
	hipLaunchKernelGGL(k_dom_tok, dim3(gs), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_tok_top, dim3(1), dim3(kDomBlock), 0, s, a);
	if (a.nb) {
		hipLaunchKernelGGL(k_dom_heads, dim3((unsigned)a.n_hblk), dim3(kDomBlock), 0, s, a);
		hipLaunchKernelGGL(k_dom_heads_top, dim3(1), dim3(kDomBlock), 0, s, a);
		hipLaunchKernelGGL(k_dom_parent, dim3((unsigned)a.n_hblk), dim3(kDomBlock), 0, s, a);
	}
	hipLaunchKernelGGL(k_dom_apply_check, dim3((unsigned)a.n_tblk), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_final, dim3(1), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_tape, dim3(gs), dim3(kDomBlock), 0, s, a);
	return hipGetLastError();
}

}  // namespace pars_hip
