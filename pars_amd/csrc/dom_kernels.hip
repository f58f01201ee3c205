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

__device__ __forceinline__ bool is_ws(uint32_t c) {
	return c == ' ' || c == '\t' || c == '\n' || c == '\r';
}
__device__ __forceinline__ bool is_digit(uint32_t c) {
	return c - '0' < 10u;
}

struct Gap {
	uint32_t kind, ok, beg, end, dec;
};

// The input read through a 16-byte register window: one vector load per
// aligned 16 bytes, byte loads where a window would reach past the buffer.
// A lane walks its gap forwards, so most bytes come from registers.
struct Bytes {
	const uint8_t* d;
	uint64_t	   n;
	uint64_t	   base;
	uint4		   v;

	__device__ Bytes(const uint8_t* d_, uint64_t n_) : d(d_), n(n_), base(~0ull), v(make_uint4(0, 0, 0, 0)) {}

	__device__ __forceinline__ uint32_t operator[](uint64_t i) {
		const uint64_t mis = (uint64_t)((uintptr_t)(d + i) & 15u);
		const uint64_t a   = i - mis;
		if (a != base) {
			if (i < mis || a + 16u > n)
				return d[i];
			base = a;
			v	 = *reinterpret_cast<const uint4*>(d + a);
		}
		const uint32_t k = (uint32_t)mis;
		const uint32_t w = k < 8u ? (k < 4u ? v.x : v.y) : (k < 12u ? v.z : v.w);
		return (w >> (8u * (k & 3u))) & 0xFFu;
	}

	// Bytes of a string body that decode to themselves - not '"', '\\' or a
	// control byte - starting at i, counted to the end of i's window or to e.
	// Byte masks by the zero-byte test on 32-bit words: every real match is
	// flagged; a false flag can only sit above a real one, so the first flag
	// is never late (a false one just hands that byte to the caller's check).
	__device__ __forceinline__ uint32_t plain_run(uint64_t i, uint64_t e) {
		const uint64_t mis = (uint64_t)((uintptr_t)(d + i) & 15u);
		const uint64_t a   = i - mis;
		if (i < mis || a + 16u > n)
			return 0;
		if (a != base) {
			base = a;
			v	 = *reinterpret_cast<const uint4*>(d + a);
		}
		auto special = [](uint32_t w) {
			const uint32_t q = w ^ 0x22222222u, b = w ^ 0x5C5C5C5Cu;
			const uint32_t m = ((q - 0x01010101u) & ~q) | ((b - 0x01010101u) & ~b) | ((w - 0x20202020u) & ~w);
			const uint32_t t = m & 0x80808080u;
			return ((t >> 7) & 1u) | ((t >> 14) & 2u) | ((t >> 21) & 4u) | ((t >> 28) & 8u);
		};
		uint32_t mask = special(v.x) | (special(v.y) << 4) | (special(v.z) << 8) | (special(v.w) << 12);
		const uint32_t k = (uint32_t)mis;
		mask &= ~((1u << k) - 1u);
		uint32_t stop = mask ? (uint32_t)__builtin_ctz(mask) : 16u;
		const uint64_t lim = e - a;	 // bytes of the window inside [a, e)
		if (lim < stop)
			stop = (uint32_t)lim;
		return stop > k ? stop - k : 0u;
	}
};

// The value in [b, e) as the hook parses it (bench_parse.cu:86-130, 219-332):
// kind, whether it parsed with only whitespace after it, its first byte, one
// past its last byte (0 for a string that did not close), and the decoded
// bytes a string mixed into the hash.  EMIT writes those bytes to `out`.
template<bool EMIT>
__device__ __forceinline__ Gap parse_gap(Bytes& d, uint64_t b, uint64_t e, uint8_t* out) {
	Gap		 g {G_EMPTY, 0u, 0u, 0u, 0u};
	uint64_t i = b;
	while (i < e && is_ws(d[i])) ++i;
	if (i >= e)
		return g;
	g.beg			 = (uint32_t)i;
	const uint32_t c = d[i];
	bool		   parsed = false;
	if (c == '"') {
		g.kind		 = G_STR;
		uint64_t j	 = i + 1;
		uint32_t dec = 0;
		auto put = [&](uint32_t v) {
			if constexpr (EMIT)
				out[dec] = (uint8_t)v;
			++dec;
		};
		bool fail = false;
		while (j < e && !fail && !parsed) {
			const uint32_t run = d.plain_run(j, e);
			if (run) {
				if constexpr (EMIT)
					for (uint32_t k = 0; k < run; ++k) out[dec + k] = (uint8_t)d[j + k];
				dec += run;
				j += run;
				continue;
			}
			const uint32_t x = d[j++];
			if (x == '"') {
				g.end  = (uint32_t)j;
				parsed = true;
			} else if (x == '\\') {
				if (j >= e) {
					fail = true;
					break;
				}
				const uint32_t esc = d[j++];
				if (esc == 'u') {
					if (j + 4 > e) {
						fail = true;
						break;
					}
					uint32_t v = 0;
					for (int k = 0; k < 4 && !fail; ++k) {
						const uint32_t h = d[j++];
						v <<= 4;
						if (is_digit(h))
							v |= h - '0';
						else if (h - 'a' < 6u)
							v |= h - 'a' + 10u;
						else if (h - 'A' < 6u)
							v |= h - 'A' + 10u;
						else
							fail = true;
					}
					if (fail)
						break;
					if (v < 0x80u) {
						put(v);
					} else if (v < 0x800u) {
						put(0xC0u | (v >> 6));
						put(0x80u | (v & 0x3Fu));
					} else {
						put(0xE0u | (v >> 12));
						put(0x80u | ((v >> 6) & 0x3Fu));
						put(0x80u | (v & 0x3Fu));
					}
				} else {
					uint32_t v;
					switch (esc) {
					case '"': v = '"'; break;
					case '\\': v = '\\'; break;
					case '/': v = '/'; break;
					case 'b': v = '\b'; break;
					case 'f': v = '\f'; break;
					case 'n': v = '\n'; break;
					case 'r': v = '\r'; break;
					case 't': v = '\t'; break;
					default: v = 256u; break;
					}
					if (v > 255u) {
						fail = true;
						break;
					}
					put(v);
				}
			} else if (x < 0x20u) {
				fail = true;
			} else {
				put(x);
			}
		}
		g.dec = dec;
	} else if (c == '-' || is_digit(c)) {
		uint64_t j	 = i + (c == '-' ? 1 : 0);
		bool	 bad = j >= e;
		if (!bad) {
			if (d[j] == '0')
				++j;
			else if (d[j] - '1' < 9u)
				while (j < e && is_digit(d[j])) ++j;
			else
				bad = true;
		}
		if (!bad && j < e && d[j] == '.') {
			++j;
			if (j >= e || !is_digit(d[j]))
				bad = true;
			else
				while (j < e && is_digit(d[j])) ++j;
		}
		if (!bad && j < e && (d[j] == 'e' || d[j] == 'E')) {
			++j;
			if (j < e && (d[j] == '+' || d[j] == '-'))
				++j;
			if (j >= e || !is_digit(d[j]))
				bad = true;
			else
				while (j < e && is_digit(d[j])) ++j;
		}
		g.kind = bad ? G_BAD : G_NUM;
		if (!bad) {
			g.end  = (uint32_t)j;
			parsed = true;
		}
	} else if (c == 't' || c == 'f' || c == 'n') {
		// "true", "false", "null" (bench_parse.cu:308-332)
		const uint64_t avail = e - i;
		uint32_t	   w4	  = 0;	// the first four bytes, little endian
		for (uint32_t k = 0; k < 4u && k < avail; ++k) w4 |= d[i + k] << (8u * k);
		g.kind = G_BAD;
		if (avail >= 4u && w4 == 0x65757274u) {	 // "true"
			g.kind = G_TRUE;
			g.end  = (uint32_t)(i + 4);
		} else if (avail >= 5u && w4 == 0x736C6166u && d[i + 4] == 'e') {	// "fals" "e"
			g.kind = G_FALSE;
			g.end  = (uint32_t)(i + 5);
		} else if (avail >= 4u && w4 == 0x6C6C756Eu) {  // "null"
			g.kind = G_NULL;
			g.end  = (uint32_t)(i + 4);
		}
		parsed = g.kind != G_BAD;
	} else {
		g.kind = G_BAD;
	}
	if (parsed) {
		uint64_t k = g.end;
		while (k < e && is_ws(d[k])) ++k;
		g.ok = k == e;
	}
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

// 1. Per token (one per thread): symbol class and the gap's value; the counts
// summed per 4096-token group (the group sums start at zero).
__global__ __launch_bounds__(kDomBlock) void k_dom_tok(DomArgs a) {
	__shared__ uint32_t s[kDomBlock / 64][3];
	const uint64_t		ntok = a.n_struct + 1;
	const uint64_t		t	 = (uint64_t)blockIdx.x * kDomBlock + threadIdx.x;
	uint32_t			sum[3] = {0u, 0u, 0u};
	if (t < ntok) {
		uint64_t b, e;
		gap_bounds(a, t, b, e);
		Bytes		   rd(a.data, a.n);
		const Gap	   g   = parse_gap<false>(rd, b, e, nullptr);
		const uint32_t sym = t < a.n_struct ? sym_class(a.data[a.positions[t]]) : SYM_END;
		const uint32_t w   = sym | (g.kind << 3) | (g.ok << 6);
		a.tinfo[t]		   = w;
		a.gbeg[t]		   = g.beg;
		a.gend[t]		   = g.end;
		a.gdec[t]		   = g.dec;
		sum[0]			   = t_nodes(w);
		sum[1]			   = g.dec;
		sum[2]			   = t_bracket(w) ? 1u : 0u;
	}
	uint32_t tot[3];
	block_excl3(sum, tot, s);
	const uint64_t grp = ((uint64_t)blockIdx.x * kDomBlock) / kDomTile;
	if (threadIdx.x == 0)
#pragma unroll
		for (int f = 0; f < 3; ++f)
			if (tot[f])
				atomicAdd(&a.tblk[grp * 3 + f], tot[f]);
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

// 3. Per token exclusive prefixes: nodes, decoded bytes, brackets.
__global__ __launch_bounds__(kDomBlock) void k_dom_tok_apply(DomArgs a) {
	__shared__ uint32_t s[kDomBlock / 64][3];
	const uint64_t		ntok = a.n_struct + 1;
	const uint64_t		base = (uint64_t)blockIdx.x * kDomTile;
	uint32_t			carry[3];
#pragma unroll
	for (int f = 0; f < 3; ++f) carry[f] = a.tblk[blockIdx.x * 3 + f];
	for (int r = 0; r < kDomRows; ++r) {
		const uint64_t t  = base + (uint64_t)r * kDomBlock + threadIdx.x;
		const bool	   in = t < ntok;
		uint32_t	   v[3] = {0u, 0u, 0u};
		if (in) {
			const uint32_t w = a.tinfo[t];
			v[0]			 = t_nodes(w);
			v[1]			 = a.gdec[t];
			v[2]			 = t_bracket(w) ? 1u : 0u;
		}
		uint32_t tot[3];
		block_excl3(v, tot, s);
		if (in) {
			a.nodes_x[t] = carry[0] + v[0];
			a.arena_x[t] = carry[1] + v[1];
			a.brk_x[t]	 = carry[2] + v[2];
		}
#pragma unroll
		for (int f = 0; f < 3; ++f) carry[f] += tot[f];
	}
}

// 4. Run starts over the sorted brackets: block maxima.
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

// 5. Exclusive max over the block maxima (one block).
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

// 6. Every bracket's parent: the bracket before its run's start.
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

// 7. Every event against the hook's rules (bench_parse.cu:86-194).  Top-level
// values are flagged: every one after the first is rejected (the hook's
// top_done), resolved by k_dom_top2.
__global__ __launch_bounds__(kDomBlock) void k_dom_check(DomArgs a) {
	const uint64_t ntok = a.n_struct + 1;
	{
		const uint64_t	   t  = (uint64_t)blockIdx.x * kDomBlock + threadIdx.x;
		unsigned long long fk = kDomNone, top = kDomNone;
		if (t < ntok) {
			const uint32_t w = a.tinfo[t];
			// the innermost open container before structural t
			int64_t		   cont = -1;
			const int64_t  b	= (int64_t)a.brk_x[t] - 1;
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
			a.topflag[t] = flag;
			if (flag)
				top = (flag & 1u) ? 2ull * t : 2ull * t + 1ull;
		}
		wave_atomic_min(&a.ds->fail_key, fk);
		wave_atomic_min(&a.ds->e1, top);
	}
}

// 8. The second top-level value event.
__global__ __launch_bounds__(kDomBlock) void k_dom_top2(DomArgs a) {
	const uint64_t			 ntok = a.n_struct + 1;
	const unsigned long long e1	  = a.ds->e1;
	{
		const uint64_t	   t = (uint64_t)blockIdx.x * kDomBlock + threadIdx.x;
		unsigned long long v = kDomNone;
		if (t < ntok && e1 != kDomNone) {
			const uint32_t f = a.topflag[t];
			if ((f & 1u) && 2ull * t > e1)
				v = 2ull * t;
			else if ((f & 2u) && 2ull * t + 1ull > e1)
				v = 2ull * t + 1ull;
		}
		wave_atomic_min(&a.ds->e2, v);
	}
}

// 9. Where the hook stopped, how many nodes and decoded bytes it produced.
__global__ void k_dom_final(DomArgs a) {
	if (threadIdx.x != 0)
		return;
	DomSummary*				 ds = a.ds;
	unsigned long long		 fk = ds->fail_key;
	const unsigned long long e2 = ds->e2;
	if (e2 != kDomNone && (e2 << 1) < fk)
		fk = e2 << 1;  // a second top-level value is rejected before it is parsed
	if (fk == kDomNone) {
		const uint64_t t = a.n_struct;	// the last token: every event accepted
		const uint32_t w = a.tinfo[t];
		ds->fail_event	 = kDomNone;
		ds->partial		 = 0;
		ds->n_nodes		 = (uint64_t)a.nodes_x[t] + t_nodes(w);
		ds->arena_bytes	 = (uint64_t)a.arena_x[t] + a.gdec[t];
		ds->ok			 = a.sum->last_bal == 0 && ds->e1 != kDomNone;  // bench_parse.cu:196-201
		return;
	}
	const unsigned long long ev = fk >> 1;
	const uint32_t			 partial = (uint32_t)(fk & 1u);
	const uint64_t			 t		 = ev >> 1;
	const uint32_t			 w		 = a.tinfo[t];
	ds->fail_event					 = ev;
	ds->partial						 = partial;
	ds->ok							 = 0;
	if ((ev & 1u) == 0) {  // the gap before structural t
		ds->n_nodes		= (uint64_t)a.nodes_x[t] + partial;
		ds->arena_bytes = (uint64_t)a.arena_x[t] + (partial ? a.gdec[t] : 0u);
	} else {  // structural t, after its accepted gap
		ds->n_nodes		= (uint64_t)a.nodes_x[t] + (t_gap(w) >= G_STR ? 1u : 0u);
		ds->arena_bytes = (uint64_t)a.arena_x[t] + a.gdec[t];
	}
}

// 10. The tape, in document order, and the decoded string bytes.
__global__ __launch_bounds__(kDomBlock) void k_dom_tape(DomArgs a) {
	const uint64_t			 ntok = a.n_struct + 1;
	const unsigned long long fe	  = a.ds->fail_event;  // kDomNone: none
	const uint32_t			 part = a.ds->partial;
	const uint64_t			 t	  = (uint64_t)blockIdx.x * kDomBlock + threadIdx.x;
	if (t < ntok) {
		const unsigned long long eg = 2ull * t;
		if (eg > fe)
			return;
		const uint32_t w  = a.tinfo[t];
		const uint32_t gk = t_gap(w);
		uint64_t	   idx = a.nodes_x[t];
		if (gk >= G_STR) {
			if (eg < fe || part) {
				if (idx < a.cap_nodes) {
					a.kind[idx] = (uint8_t)gk;
					a.nbeg[idx] = a.gbeg[t];
					a.nend[idx] = a.gend[t];
					a.soff[idx] = a.arena_x[t];
				}
				if (gk == G_STR && (uint64_t)a.arena_x[t] + a.gdec[t] <= a.cap_arena) {
					uint64_t b, e;
					gap_bounds(a, t, b, e);
					Bytes rd(a.data, a.n);
					(void)parse_gap<true>(rd, b, e, a.arena + a.arena_x[t]);
				}
			}
			++idx;
		}
		const uint32_t sym = t_sym(w);
		if ((sym == SYM_OOPEN || sym == SYM_AOPEN) && eg + 1ull < fe && idx < a.cap_nodes) {
			// the close that follows this open at the same key is its partner
			// while the hook accepts it
			const uint32_t b   = a.brk_x[t];
			const uint32_t r   = a.rank[b];
			uint32_t	   end = 0;
			if ((uint64_t)r + 1u < a.nb) {
				const uint32_t x = a.order[r + 1u];
				if (a.bkey[x] == a.bkey[b] && (a.plan->br_flags[a.bchar[x]] & 2u)) {
					const uint32_t tc = a.bsidx[x];
					if (2ull * tc + 1ull < fe)
						end = a.positions[tc];
				}
			}
			a.kind[idx] = sym == SYM_OOPEN ? 0u : 1u;
			a.nbeg[idx] = a.positions[t];
			a.nend[idx] = end;
			a.soff[idx] = a.arena_x[t] + a.gdec[t];
		}
	}
}


}  // namespace

hipError_t launch_json_tape(const DomArgs& a, hipStream_t s) {
	const uint64_t ntok = a.n_struct + 1;
	const unsigned gt	= (unsigned)((ntok + kDomBlock - 1) / kDomBlock);	// one token per thread
	hipLaunchKernelGGL(k_dom_reset, dim3(1), dim3(64), 0, s, a.ds);
	hipError_t e = launch_fill_i32(reinterpret_cast<int32_t*>(a.tblk), a.n_tblk * 3, 0, s);
	if (e != hipSuccess)
		return e;
	hipLaunchKernelGGL(k_dom_tok, dim3(gt), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_tok_top, dim3(1), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_tok_apply, dim3((unsigned)a.n_tblk), dim3(kDomBlock), 0, s, a);
	if (a.nb) {
		hipLaunchKernelGGL(k_dom_heads, dim3((unsigned)a.n_hblk), dim3(kDomBlock), 0, s, a);
		hipLaunchKernelGGL(k_dom_heads_top, dim3(1), dim3(kDomBlock), 0, s, a);
		hipLaunchKernelGGL(k_dom_parent, dim3((unsigned)a.n_hblk), dim3(kDomBlock), 0, s, a);
	}
	hipLaunchKernelGGL(k_dom_check, dim3(gt), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_top2, dim3(gt), dim3(kDomBlock), 0, s, a);
	hipLaunchKernelGGL(k_dom_final, dim3(1), dim3(64), 0, s, a);
	hipLaunchKernelGGL(k_dom_tape, dim3(gt), dim3(kDomBlock), 0, s, a);
	return hipGetLastError();
}

}  // namespace pars_hip
