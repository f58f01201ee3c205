// kv_kernels.hip — the INI and TOML key/value tables: the reference's stage-2
// hooks for these formats (bench/bench_parse_host.cu: each_token :34-56,
// ini_parse :69-134, toml_parse :202-248), which replay the structural index
// (include/pars/model/visit.hpp:35-48) through a sequential hook, computed
// here from the document's bytes and structural bitmap in parallel.
//
// Two small automata carry everything sequential in those hooks:
//   * the gap tokenizer (each_token): whitespace separates tokens; a token that
//     starts with a quote runs to the same quote (a backslash skips the next
//     byte) or to the gap's end.  Eight states - between tokens with none yet in
//     this gap (G0), between tokens (G1), just after a closing quote (CQ), in a
//     plain token (PL), in a "..." token (QD), after its backslash (ED), in a
//     '...' token (QS), after its backslash (ES) - and every structural byte
//     sends every state to G0.  Per 64-byte lane the tokenizer's map is composed
//     with v_perm (8 states = two dwords), scanned over lanes and tiles, and the
//     lane walked once from its entry state; the walk's states give bit masks of
//     token starts, first starts of a gap and token ends.
//   * the hook (ini_parse / toml_parse): (comment flag, expect / mode), eight
//     states.  Only the first token of a gap and the structural bytes change it
//     (later tokens of a gap leave it as the first one did), so a lane's hook
//     map is the composition over those events; maps are scanned over lanes and
//     tiles like the tokenizer's, and every token's role follows from the state
//     its start meets.
// A TOML value's key (the hook's `key`, the last key token before it, kept
// across lines) is a max-scan over the table.
//
// Passes: k_kv_tokmap (lane tokenizer maps) -> k_kv_chain (tile entry states)
// -> k_kv_walk (masks, counts, lane hook maps) -> k_kv_chain, two blocks
// (tile hook states, token offsets) | host sizes the table | k_kv_emit (roles, begins,
// ends) -> TOML: k_kv_chain (last key before each tile) -> k_kv_keys.
#include "dev_common.hpp"
#include "kernels.hpp"

namespace pars_hip {

namespace {

using M8 = MapT<true>;

constexpr uint32_t kWaves = kKvLanes / kWave;
// a lane's 16 dwords staged in LDS at a stride of 17, so the lanes of a wave
// spread over all banks
constexpr int kStage = 17;

// tokenizer states
enum : uint32_t { T_G0 = 0, T_G1 = 1, T_CQ = 2, T_PL = 3, T_QD = 4, T_ED = 5, T_QS = 6, T_ES = 7 };

// The tokenizer's transition per byte class, as a v_perm map (byte s = next
// state of s; lo: states 0-3, hi: 4-7) and packed by nibbles (nibble s).
//            G0  G1  CQ  PL | QD  ED  QS  ES
//   space:   G0  G1  G1  G1 | QD  QD  QS  QS
//   '"':     QD  QD  QD  PL | CQ  QD  QS  QS
//   '\'':    QS  QS  QS  PL | QD  QD  CQ  QS
//   '\\':    PL  PL  PL  PL | ED  QD  ES  QS
//   other:   PL  PL  PL  PL | QD  QD  QS  QS
//   structural byte: every state -> G0
struct TokClass {
	uint32_t lo, hi, nib;
};
constexpr TokClass kTokWs	 = {0x01010100u, 0x06060404u, 0x66441110u};
constexpr TokClass kTokDq	 = {0x03040404u, 0x06060402u, 0x66423444u};
constexpr TokClass kTokSq	 = {0x03060606u, 0x06020404u, 0x62443666u};
constexpr TokClass kTokBs	 = {0x03030303u, 0x06070405u, 0x67453333u};
constexpr TokClass kTokOther = {0x03030303u, 0x06060404u, 0x66443333u};

__device__ __forceinline__ TokClass tok_class(uint32_t c) {
	if (c == ' ' || c == '\t' || c == '\n' || c == '\r')
		return kTokWs;
	if (c == '"')
		return kTokDq;
	if (c == '\'')
		return kTokSq;
	if (c == '\\')
		return kTokBs;
	return kTokOther;
}

// Hook states: bits 0-1 expect (INI: 0 key, 1 section, 2 value) or mode (TOML:
// 0 none, 1 section, 2 value, 3 array), bit 2 in a comment.  Maps as v_perm
// byte maps (states 4-7 are the comment states: identity unless a newline).
constexpr uint32_t kHookIdHi = 0x07060504u;
// the first token of a gap: expect / mode -> 0, except a TOML array (3) stays
constexpr Map8 kHookFirst = {0x03000000u, kHookIdHi};

// A structural byte's hook map (ini_parse::symbol :107-121, toml_parse::symbol
// :235-247).
template<bool TOML>
__device__ __forceinline__ Map8 hook_symbol(uint32_t c) {
	if (c == '\n')
		return Map8 {0u, 0u};
	if (c == '#' || (!TOML && c == ';'))
		return Map8 {0x07060504u, 0x07060504u};
	if (c == '[')
		return Map8 {TOML ? 0x01030101u : 0x01010101u, kHookIdHi};
	if (c == ']')
		return Map8 {0u, kHookIdHi};
	if (c == '=')
		return Map8 {0x02020202u, kHookIdHi};
	return Map8 {kIdLo, kIdHi};
}

// The role a token gets from the hook state its start meets: in a comment the
// hook does not see it (SKIPPED); else by expect / mode.
template<bool TOML>
__device__ __forceinline__ uint32_t hook_role(uint32_t h) {
	// roles of states 0-3 by nibble: SECTION 0, KEY 1, VALUE 2, ELEMENT 3
	constexpr uint32_t kTab = TOML ? 0x3201u : 0x1201u;
	return (h & 4u) ? 4u : (kTab >> (4u * h)) & 0xFu;
}

// A lane's 64 bytes (fewer at the document's end), as 16 little-endian dwords,
// and its structural bits.  Returns how many bytes are in the document.
__device__ __forceinline__ uint32_t load_lane(const KvArgs& a, uint64_t base, uint32_t (&x)[16], uint64_t& sb) {
	const uint32_t valid = base >= a.n ? 0u : (uint32_t)min<uint64_t>(64, a.n - base);
	if (valid == 64 && (((uintptr_t)(a.data + base)) & 15u) == 0) {
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
	sb = valid ? a.bitmap[base >> 6] : 0ull;
	if (valid < 64)
		sb &= valid ? (~0ull >> (64 - valid)) : 0ull;
	return valid;
}

__device__ __forceinline__ uint32_t byte_of(const uint32_t (&x)[16], int j) {
	return (x[j >> 2] >> (8 * (j & 3))) & 0xFFu;
}

// Exclusive scan of lane maps over the block (lane order), the block's
// aggregate in *agg.  `s_w` holds kWaves maps.
__device__ __forceinline__ Map8 block_excl_maps(Map8 m, Map8* s_w, Map8* agg) {
	const int	tid	 = threadIdx.x, w = tid / kWave;
	const Map8	incl = wave_scan_maps<true>(m);
	const Map8	excl = wave_shift_maps<true>(incl);
	if (lane_id() == kWave - 1)
		s_w[w] = incl;
	__syncthreads();
	Map8 pre = M8::id();
	for (int k = 0; k < w; ++k) pre = M8::compose(s_w[k], pre);
	Map8 all = pre;
	for (int k = w; k < (int)kWaves; ++k) all = M8::compose(s_w[k], all);
	*agg = all;
	return M8::compose(excl, pre);
}

// Exclusive sum over the block, the total in *tot.  `s_w` holds kWaves words.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* s_w, uint32_t* tot) {
	const int	   tid	= threadIdx.x, w = tid / kWave;
	const uint32_t incl = wave_incl_sum(v);
	if (lane_id() == kWave - 1)
		s_w[w] = incl;
	__syncthreads();
	uint32_t pre = 0, all = 0;
	for (int k = 0; k < (int)kWaves; ++k) {
		if (k < w)
			pre += s_w[k];
		all += s_w[k];
	}
	*tot = all;
	return pre + incl - v;
}

// ---------------------------------------------------------------------------
// Pass 1: each lane's tokenizer map; the block's exclusive scan gives each
// lane its map from the tile's start, the tile's own map goes to t_tmap.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kKvLanes) void k_kv_tokmap(KvArgs a) {
	__shared__ uint2 s_map[512];  // [byte value | structural << 8]
	__shared__ Map8	 s_w[kWaves];
	const int		 tid = threadIdx.x;
	{
		const TokClass t = tok_class((uint32_t)tid);
		s_map[tid]		 = make_uint2(t.lo, t.hi);
		s_map[256 + tid] = make_uint2(0u, 0u);
	}
	if (blockIdx.x == 0 && tid == 0) {
		a.sum->open_end = 0;
		a.sum->bad		= 0;
		for (int r = 0; r < 5; ++r) a.sum->roles[r] = 0;
	}
	__syncthreads();
	const uint64_t lane = (uint64_t)blockIdx.x * kKvLanes + tid, base = lane * 64;
	uint32_t	   x[16];
	uint64_t	   sb;
	const uint32_t valid = load_lane(a, base, x, sb);
	Map8		   m	 = M8::id();
	if (valid == 64) {
#pragma unroll
		for (int j = 0; j < 64; ++j) {
			const uint2 t = s_map[byte_of(x, j) | ((uint32_t)(sb >> j) & 1u) << 8];
			m			  = M8::compose(Map8 {t.x, t.y}, m);
		}
	} else {
		for (int j = 0; j < (int)valid; ++j) {
			const uint32_t c = (uint32_t)a.data[base + j];
			const uint2	   t = s_map[c | ((uint32_t)(sb >> j) & 1u) << 8];
			m				 = M8::compose(Map8 {t.x, t.y}, m);
		}
	}
	Map8	   agg;
	const Map8 entry = block_excl_maps(m, s_w, &agg);
	a.l_tmap[lane]	 = M8::pack(entry);
	if (tid == 0)
		a.t_tmap[blockIdx.x] = M8::pack(agg);
}

// ---------------------------------------------------------------------------
// Pass 2 (one block): a scan over the tiles, in one of three kinds -
//   maps:  out[t] = the state entering tile t (maps packed by 3 bits), from init
//   sums:  out[t] = the sum before tile t, the total into *total
//   max:   out[t] = the maximum (int32) before tile t, init -1
// Each thread folds a run of tiles, the block scans the folds, and each thread
// walks its run again with what came before it.
// ---------------------------------------------------------------------------
constexpr int kChainThreads = 1024;
enum { CHAIN_MAPS = 0, CHAIN_SUMS = 1, CHAIN_MAX = 2 };

// f(t, in[t]) for t in [b, e) in order, 32 loads issued per round trip.
template<class F>
__device__ __forceinline__ void each32(const uint32_t* in, uint32_t b, uint32_t e, F&& f) {
	constexpr uint32_t K = 32;
	for (uint32_t t = b; t < e; t += K) {
		uint32_t v[K];
#pragma unroll
		for (uint32_t u = 0; u < K; ++u) v[u] = t + u < e ? in[t + u] : 0u;
#pragma unroll
		for (uint32_t u = 0; u < K; ++u)
			if (t + u < e)
				f(t + u, v[u]);
	}
}

// One scan over the tiles: what k_kv_chain computes (see there).
struct ChainJob {
	const uint32_t* in;
	uint32_t*		out;
	uint32_t		T;
	int				kind;
	uint32_t		init;
	uint64_t*		total;
};

// Independent scans run as the blocks of one launch (block i: job i).
__global__ __launch_bounds__(kChainThreads) void k_kv_chain(ChainJob j0, ChainJob j1) {
	const bool		   b0	 = blockIdx.x == 0;
	const uint32_t*	   in	 = b0 ? j0.in : j1.in;
	uint32_t*		   out	 = b0 ? j0.out : j1.out;
	const uint32_t	   T	 = b0 ? j0.T : j1.T, init = b0 ? j0.init : j1.init;
	const int		   kind	 = b0 ? j0.kind : j1.kind;
	uint64_t*		   total = b0 ? j0.total : j1.total;
	constexpr int	   NW = kChainThreads / kWave;
	__shared__ Map8	   s_m[NW];
	__shared__ uint32_t s_v[NW];
	const int		   tid = threadIdx.x, w = tid / kWave;
	const uint32_t	   per = (T + kChainThreads - 1) / kChainThreads;
	const uint32_t	   b = min(T, (uint32_t)tid * per), e = min(T, b + per);
	if (kind == CHAIN_MAPS) {
		Map8 f = M8::id();
		each32(in, b, e, [&](uint32_t, uint32_t v) { f = M8::compose(M8::unpack(v), f); });
		const Map8 incl = wave_scan_maps<true>(f);
		const Map8 excl = wave_shift_maps<true>(incl);
		if (lane_id() == kWave - 1)
			s_m[w] = incl;
		__syncthreads();
		Map8 pre = M8::id();
		for (int k = 0; k < w; ++k) pre = M8::compose(s_m[k], pre);
		uint32_t s = M8::apply(M8::compose(excl, pre), init);
		each32(in, b, e, [&](uint32_t t, uint32_t v) {
			out[t] = s;
			s	   = M8::apply(M8::unpack(v), s);
		});
	} else if (kind == CHAIN_SUMS) {
		uint64_t f = 0;
		each32(in, b, e, [&](uint32_t, uint32_t v) { f += v; });
		// the document has fewer than 2^32 bytes, so fewer tokens
		const uint32_t incl = wave_incl_sum((uint32_t)f);
		if (lane_id() == kWave - 1)
			s_v[w] = incl;
		__syncthreads();
		uint32_t pre = 0, all = 0;
		for (int k = 0; k < NW; ++k) {
			if (k < w)
				pre += s_v[k];
			all += s_v[k];
		}
		uint32_t s = pre + incl - (uint32_t)f;
		each32(in, b, e, [&](uint32_t t, uint32_t v) {
			out[t] = s;
			s += v;
		});
		if (tid == 0)
			*total = all;
	} else {
		int f = -1;
		each32(in, b, e, [&](uint32_t, uint32_t v) { f = max(f, (int)v); });
		int incl = f;
#pragma unroll
		for (int o = 1; o < kWave; o <<= 1) {
			const int u = __shfl_up(incl, o, kWave);
			if (lane_id() >= o)
				incl = max(incl, u);
		}
		if (lane_id() == kWave - 1)
			s_v[w] = (uint32_t)incl;
		__syncthreads();
		int pre = -1;
		for (int k = 0; k < w; ++k) pre = max(pre, (int)s_v[k]);
		int excl = __shfl_up(incl, 1, kWave);
		if (lane_id() == 0)
			excl = -1;
		int s = max(pre, excl);
		each32(in, b, e, [&](uint32_t t, uint32_t v) {
			out[t] = (uint32_t)s;
			s	   = max(s, (int)v);
		});
	}
}

// The 64 nibbles of a walk (one state per byte, nibble j of dword j / 8) to
// bit masks: byte j's bit set where its state satisfies the test.  In the
// SWAR tests no nibble carries into the next (every state is < 8).
__device__ __forceinline__ uint32_t nib_bits(uint32_t t) {	 // bits at 4i+3 -> bits i
	uint32_t x = t >> 3;
	x		   = (x | (x >> 3)) & 0x03030303u;
	x		   = (x | (x >> 6)) & 0x000F000Fu;
	x		   = (x | (x >> 12)) & 0xFFu;
	return x;
}

struct WalkMasks {
	uint64_t in_tok;  // state after byte j is inside a token (PL .. ES)
	uint64_t gap0;	  // ... is G0 (no token yet in this gap)
	uint64_t cq;	  // ... is CQ (byte j closed a quoted token)
};

__device__ __forceinline__ WalkMasks nib_masks(const uint32_t (&nb)[8]) {
	WalkMasks m {0, 0, 0};
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		const uint32_t v = nb[k];
		m.in_tok |= (uint64_t)nib_bits((v + 0x55555555u) & 0x88888888u) << (8 * k);
		m.gap0 |= (uint64_t)nib_bits(~(v + 0x77777777u) & 0x88888888u) << (8 * k);
		m.cq |= (uint64_t)nib_bits(~((v ^ 0x22222222u) + 0x77777777u) & 0x88888888u) << (8 * k);
	}
	return m;
}

// ---------------------------------------------------------------------------
// Pass 3: walk each lane from its tokenizer entry state; token start / first
// start / end masks, the lane's token count and hook map; block scans give each
// lane its offsets from the tile's start, tiles their totals.
// ---------------------------------------------------------------------------
template<bool TOML>
__global__ __launch_bounds__(kKvLanes) void k_kv_walk(KvArgs a) {
	__shared__ uint32_t s_nib[512];	 // [byte value | structural << 8]
	__shared__ uint2	s_hm[256];
	__shared__ uint32_t s_b[kKvLanes * kStage];	 // the tile's bytes, for the symbols' values
	__shared__ Map8		s_w[kWaves];
	__shared__ uint32_t s_c[kWaves];
	const int			tid = threadIdx.x;
	{
		s_nib[tid]		 = tok_class((uint32_t)tid).nib;
		s_nib[256 + tid] = 0u;
		const Map8 h = hook_symbol<TOML>((uint32_t)tid);
		s_hm[tid]	 = make_uint2(h.lo, h.hi);
	}
	const uint64_t lane = (uint64_t)blockIdx.x * kKvLanes + tid, base = lane * 64;
	uint32_t	   x[16];
	uint64_t	   sb;
	const uint32_t valid = load_lane(a, base, x, sb);
#pragma unroll
	for (int k = 0; k < 16; ++k) s_b[tid * kStage + k] = x[k];
	const uint32_t s_in = M8::apply(M8::unpack(a.l_tmap[lane]), a.t_tin[blockIdx.x]);
	__syncthreads();

	uint32_t nb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	uint32_t s	   = s_in;
	if (valid == 64) {
#pragma unroll
		for (int j = 0; j < 64; ++j) {
			const uint32_t w = s_nib[byte_of(x, j) | ((uint32_t)(sb >> j) & 1u) << 8];
			s				 = __builtin_amdgcn_ubfe(w, 4u * s, 4u);
			nb[j >> 3] |= s << (4 * (j & 7));
		}
	} else {
#pragma unroll
		for (int j = 0; j < 64; ++j)
			if ((uint32_t)j < valid) {
				const uint32_t w = s_nib[byte_of(x, j) | ((uint32_t)(sb >> j) & 1u) << 8];
				s				 = __builtin_amdgcn_ubfe(w, 4u * s, 4u);
				nb[j >> 3] |= s << (4 * (j & 7));
			}
	}
	const uint64_t vmask  = valid == 64 ? ~0ull : (valid ? ~0ull >> (64 - valid) : 0ull);
	const WalkMasks wm	  = nib_masks(nb);
	const uint64_t	in_e  = s_in >= T_PL ? 1ull : 0ull;
	const uint64_t	prevI = (wm.in_tok << 1) | in_e;
	const uint64_t	prevZ = (wm.gap0 << 1) | (s_in == T_G0 ? 1ull : 0ull);
	const uint64_t	start = wm.in_tok & ~prevI & vmask;
	const uint64_t	first = start & prevZ;
	const uint64_t	endq  = wm.cq & vmask;
	const uint64_t	end	  = ((prevI & ~wm.in_tok & ~wm.cq) | wm.cq) & vmask;
	a.l_start[lane]		  = start;
	a.l_first[lane]		  = first;
	a.l_end[lane]		  = end;
	a.l_endq[lane]		  = endq;
	if (valid && base + valid == a.n && s >= T_PL)
		a.sum->open_end = 1u;

	// the lane's hook map: first tokens of a gap and structural bytes, in order
	Map8	 h	= M8::id();
	uint64_t ev = (sb | first) & vmask;
	while (ev) {
		const int j = __builtin_ctzll(ev);
		ev &= ev - 1;
		Map8 e = kHookFirst;
		if ((sb >> j) & 1u) {
			const uint32_t c = (s_b[tid * kStage + (j >> 2)] >> (8 * (j & 3))) & 0xFFu;
			const uint2	   t = s_hm[c];
			e				 = Map8 {t.x, t.y};
		}
		h = M8::compose(e, h);
	}
	uint32_t	   tot;
	const uint32_t cnt = (uint32_t)__builtin_popcountll(start);
	const uint32_t off = block_excl_sum(cnt, s_c, &tot);
	Map8		   hagg;
	const Map8	   hentry = block_excl_maps(h, s_w, &hagg);
	a.l_off[lane]		  = off | (uint32_t)in_e << 31;
	a.l_hmap[lane]		  = M8::pack(hentry);
	if (tid == 0) {
		a.t_cnt[blockIdx.x]	 = tot;
		a.t_hmap[blockIdx.x] = M8::pack(hagg);
	}
}

// ---------------------------------------------------------------------------
// Pass 4: the table.  Each lane replays its structural bytes and token starts
// in order from its hook entry state: a start takes its role from the state,
// a first start then updates it, a structural byte updates it.  The k-th end
// in a lane closes the token started k-th in it, or the open one before it.
// ---------------------------------------------------------------------------
// Tiles with at most this many token starts stage their rows in LDS and write
// them out coalesced; denser tiles write each row from its lane.
constexpr uint32_t kEmitStage = 3072;

template<bool TOML>
__global__ __launch_bounds__(kKvLanes) void k_kv_emit(KvArgs a) {
	__shared__ uint2	s_hm[256];
	__shared__ uint32_t s_b[kKvLanes * kStage];
	__shared__ uint32_t s_r[5][kWaves];
	__shared__ int		s_k[kWaves];
	__shared__ uint8_t	s_role[kEmitStage];
	__shared__ uint32_t s_beg[kEmitStage], s_end[kEmitStage];
	const int			tid = threadIdx.x;
	const uint64_t		t0 = a.t_off[blockIdx.x];
	const uint32_t		tcnt = a.t_cnt[blockIdx.x];
	const bool			staged = tcnt <= kEmitStage;  // the same in every lane
	{
		const Map8 h = hook_symbol<TOML>((uint32_t)tid);
		s_hm[tid]	 = make_uint2(h.lo, h.hi);
	}
	const uint64_t lane = (uint64_t)blockIdx.x * kKvLanes + tid, base = lane * 64;
	uint32_t	   x[16];
	uint64_t	   sb;
	const uint32_t valid = load_lane(a, base, x, sb);
#pragma unroll
	for (int k = 0; k < 16; ++k) s_b[tid * kStage + k] = x[k];
	__syncthreads();
	const uint64_t start = valid ? a.l_start[lane] : 0ull, first = valid ? a.l_first[lane] : 0ull;
	const uint64_t end = valid ? a.l_end[lane] : 0ull, endq = valid ? a.l_endq[lane] : 0ull;
	const uint32_t offw = valid ? a.l_off[lane] : 0u;
	uint32_t	   h	= valid ? M8::apply(M8::unpack(a.l_hmap[lane]), a.t_hin[blockIdx.x]) : 0u;
	const uint64_t tb	= t0 + (offw & 0x7FFFFFFFu);
	uint64_t	   nr = 0;	// tokens per role, a byte each (a lane starts at most 64)
	int			   lastkey = -1;
	const uint64_t ntok	   = a.sum->n_tokens;
	uint64_t	   k	   = tb;
	uint64_t	   ev	   = sb | start;
	while (ev) {
		const int j = __builtin_ctzll(ev);
		ev &= ev - 1;
		if ((sb >> j) & 1u) {
			const uint32_t c = (s_b[tid * kStage + (j >> 2)] >> (8 * (j & 3))) & 0xFFu;
			const uint2	   t = s_hm[c];
			h				 = M8::apply(Map8 {t.x, t.y}, h);
		} else {
			const uint32_t r = hook_role<TOML>(h);
			if (k >= ntok || k - t0 >= tcnt) {	// the passes disagree on a count: refuse, write nothing out of bounds
				a.sum->bad = 1u;
				break;
			}
			if (staged) {
				s_role[k - t0] = (uint8_t)r;
				s_beg[k - t0]  = (uint32_t)(base + j);
			} else {
				a.role[k] = (uint8_t)r;
				a.beg[k]  = (uint32_t)(base + j);
			}
			if (TOML && r == 1u)
				lastkey = (int)k;
			nr += 1ull << (8 * r);
			++k;
			if ((first >> j) & 1u)
				h = M8::apply(kHookFirst, h);
		}
	}
	uint64_t ke = tb - (offw >> 31);
	for (uint64_t e = end; e; e &= e - 1) {
		const int j = __builtin_ctzll(e);
		if (ke >= ntok || (ke >= t0 && ke - t0 >= tcnt)) {
			a.sum->bad = 1u;
			break;
		}
		const uint32_t e1 = (uint32_t)(base + j + ((endq >> j) & 1u));
		if (staged && ke >= t0)
			s_end[ke - t0] = e1;
		else  // a token started in an earlier tile, or a dense tile
			a.end[ke] = e1;
		++ke;
	}
	if (valid && base + valid == a.n && a.sum->open_end && ntok)
		a.end[ntok - 1] = (uint32_t)a.n;

	const int w = tid / kWave;
#pragma unroll
	for (int r = 0; r < 5; ++r) {
		const uint32_t t = wave_sum((uint32_t)(nr >> (8 * r)) & 0xFFu);
		if (lane_id() == 0)
			s_r[r][w] = t;
	}
	if (TOML) {
		const int m = wave_max(lastkey);
		if (lane_id() == 0)
			s_k[w] = m;
	}
	__syncthreads();
	if (staged) {
		// the tile's last token may run on into the next tile (or to the
		// document's end): its end is written by whoever meets it
		const bool	   open_out = blockIdx.x + 1 < a.n_tiles ? a.t_tin[blockIdx.x + 1] >= T_PL : a.sum->open_end != 0;
		const uint32_t n_end	= (open_out && tcnt) ? tcnt - 1 : tcnt;
		const uint32_t n_row	= (uint32_t)min<uint64_t>(tcnt, ntok > t0 ? ntok - t0 : 0);
		for (uint32_t i = tid; i < n_row; i += kKvLanes) {
			a.role[t0 + i] = s_role[i];
			a.beg[t0 + i]  = s_beg[i];
			if (i < n_end)
				a.end[t0 + i] = s_end[i];
		}
	}
	if (tid < 5) {
		uint32_t t = 0;
		for (int q = 0; q < (int)kWaves; ++q) t += s_r[tid][q];
		if (t)
			atomicAdd(&a.sum->roles[tid], t);
	}
	if (TOML && tid == 0) {
		int m = -1;
		for (int q = 0; q < (int)kWaves; ++q) m = max(m, s_k[q]);
		a.t_key[blockIdx.x] = m;
	}
}

// ---------------------------------------------------------------------------
// Pass 5 (TOML): each token's key - for a VALUE, the last KEY token before it
// (toml_parse keeps `key` across lines, :216-224); -1 for other roles or when
// no key came yet.  One block per tile's tokens, a max-scan in chunks.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kKvLanes) void k_kv_keys(KvArgs a) {
	constexpr uint32_t PER = 4;	 // consecutive tokens per thread
	__shared__ int	   s_w[kWaves];
	const int		   tid = threadIdx.x, w = tid / kWave;
	const uint64_t	   b = a.t_off[blockIdx.x], e = b + a.t_cnt[blockIdx.x];
	int				   carry = (int)a.t_kin[blockIdx.x];
	for (uint64_t i0 = b; i0 < e; i0 += (uint64_t)kKvLanes * PER) {
		const uint64_t i = i0 + (uint64_t)tid * PER;
		uint32_t	   r[PER];
		int			   mine = -1;  // this thread's last KEY
#pragma unroll
		for (uint32_t u = 0; u < PER; ++u) {
			r[u] = i + u < e ? a.role[i + u] : 4u;
			if (r[u] == 1u)
				mine = (int)(i + u);
		}
		int v = mine;
#pragma unroll
		for (int o = 1; o < kWave; o <<= 1) {
			const int q = __shfl_up(v, o, kWave);
			if (lane_id() >= o)
				v = max(v, q);
		}
		if (lane_id() == kWave - 1)
			s_w[w] = v;
		__syncthreads();
		int pre = carry, all = carry;
		for (int q = 0; q < (int)kWaves; ++q) {
			if (q < w)
				pre = max(pre, s_w[q]);
			all = max(all, s_w[q]);
		}
		int ex = __shfl_up(v, 1, kWave);  // last KEY of the wave's earlier threads
		if (lane_id() == 0)
			ex = -1;
		int last = max(pre, ex);
#pragma unroll
		for (uint32_t u = 0; u < PER; ++u) {
			if (i + u < e)
				a.key[i + u] = r[u] == 2u ? last : -1;
			if (r[u] == 1u)
				last = (int)(i + u);
		}
		carry = all;
		__syncthreads();
	}
}

}  // namespace

hipError_t launch_kv_count(const KvArgs& a, hipStream_t s) {
	if (a.n_tiles == 0)
		return hipSuccess;
	hipLaunchKernelGGL(k_kv_tokmap, dim3(a.n_tiles), dim3(kKvLanes), 0, s, a);
	const ChainJob none {};
	hipLaunchKernelGGL(k_kv_chain, dim3(1), dim3(kChainThreads), 0, s,
					   ChainJob {a.t_tmap, a.t_tin, a.n_tiles, (int)CHAIN_MAPS, (uint32_t)T_G0, nullptr}, none);
	if (a.toml)
		hipLaunchKernelGGL(k_kv_walk<true>, dim3(a.n_tiles), dim3(kKvLanes), 0, s, a);
	else
		hipLaunchKernelGGL(k_kv_walk<false>, dim3(a.n_tiles), dim3(kKvLanes), 0, s, a);
	// the tiles' hook states and token offsets: two blocks of one launch
	hipLaunchKernelGGL(k_kv_chain, dim3(2), dim3(kChainThreads), 0, s,
					   ChainJob {a.t_hmap, a.t_hin, a.n_tiles, (int)CHAIN_MAPS, 0u, nullptr},
					   ChainJob {a.t_cnt, a.t_off, a.n_tiles, (int)CHAIN_SUMS, 0u, &a.sum->n_tokens});
	return hipGetLastError();
}

hipError_t launch_kv_emit(const KvArgs& a, hipStream_t s) {
	if (a.n_tiles == 0)
		return hipSuccess;
	if (a.toml) {
		hipLaunchKernelGGL(k_kv_emit<true>, dim3(a.n_tiles), dim3(kKvLanes), 0, s, a);
		hipLaunchKernelGGL(k_kv_chain, dim3(1), dim3(kChainThreads), 0, s,
						   ChainJob {(const uint32_t*)a.t_key, (uint32_t*)a.t_kin, a.n_tiles, (int)CHAIN_MAX, 0u, nullptr},
						   ChainJob {});
		hipLaunchKernelGGL(k_kv_keys, dim3(a.n_tiles), dim3(kKvLanes), 0, s, a);
	} else {
		hipLaunchKernelGGL(k_kv_emit<false>, dim3(a.n_tiles), dim3(kKvLanes), 0, s, a);
	}
	return hipGetLastError();
}

}  // namespace pars_hip
