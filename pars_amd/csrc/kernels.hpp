// kernels.hpp — launch interface between the host orchestration and the HIP
// kernels (all of which live in this directory).
#pragma once

#include "plan_dev.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pars_hip {

constexpr int kBlock = 256;			   // threads per scan block (4 waves)
constexpr int kBPL	 = 64;			   // bytes per lane and row
constexpr int kTile	 = kBlock * kBPL;  // bytes per tile row (16 KiB); a tile holds 1..4 rows
constexpr int kRecWords = 32;		   // look-back words per tile record (256 bytes)

// Record layout, two-chain mode (u64 tagged words): [0..7] state-chain
// aggregates, one per descriptor (3-bit packed map); [8] state-chain inclusive
// (packed states); [9..11] count aggregates (structural, brackets, depth);
// [12..14] count inclusives.
enum : int { kRecSAgg = 0, kRecSIncl = 8, kRecCAgg = 9, kRecCIncl = 12 };
// One-chain mode (one descriptor): [0] aggregate map; [1 + 3e + f] count f of
// the tile entered in state e (e < 8); [28] inclusive state; [29..31]
// inclusive counts.
enum : int { kRecFMap = 0, kRecFCnt = 1, kRecFIncl = 28 };

// Device-side result summary of one scan (written by the kernels, read back by
// the host once at the end of a call).
struct DevSummary {
	uint32_t n_struct;	   // total structural count (may exceed the output capacity)
	uint32_t n_brackets;   // total bracket count
	int32_t	 last_bal;	   // depth after the last bracket
	uint32_t final_states; // packed per-descriptor states after the last byte
	int32_t	 min_bal;	   // min over bracket prefix depths and 0
	int32_t	 max_bal;	   // max over bracket prefix depths and 0
	uint32_t mismatched;   // some sorted pair is not a matching open/close
	uint32_t sort_passes;  // radix passes the pairing sort needs
	int32_t	 kmin;		   // smallest pairing key
	uint32_t n_chunks;	   // pairing sort chunks
	uint32_t overflow;	   // an output exceeded its capacity
	uint32_t digit_bits;   // radix digit width of the pairing sort
	// Look-back epoch of the current launch: advanced by every summary reset
	// (never reset itself), so records of earlier launches - also of earlier
	// replays of one captured graph - read as unpublished.
	uint32_t epoch;
	uint32_t small_done;  // R6 is complete before the sort: the one-block path or the stack path paired it
	uint32_t doc_unbal;	  // batch: some document ends at a depth other than 0
	// batch pairing: distinct (document, key) values, the sum over documents
	// of (max depth - min depth + 2)
	uint64_t key_total;
};

// Batch mode: the input holds n_docs documents in ascending order, each
// starting on a 64-byte boundary (a lane's first byte); bytes between
// documents belong to none.  A document head resets the state carry to the
// start states and the depth to 0; structural and bracket outputs stay
// compacted across the whole batch (documents in order), with each document's
// first index recorded.  All null / 0 outside batch mode.
struct BatchArgs {
	const uint64_t* doc_off;	  // [n_docs] first byte of each document
	const uint32_t* doc_len;	  // [n_docs] bytes (>= 1)
	int32_t*		tile_doc;	  // [n_tiles] last document starting at or before the tile (-1: none)
	const uint64_t* head_bits;	  // bit g set: a document starts at byte 64 g
	uint64_t		n_head_words;
	uint32_t		n_docs;
	uint32_t*		st_start;	  // [n_docs + 1] index of each document's first structural
	uint32_t*		br_start;	  // [n_docs + 1] index of each document's first bracket
	uint32_t*		doc_final;	  // [n_docs] packed states after the document's last byte
	int32_t*		doc_min;	  // [n_docs] min depth after a bracket (starts at 0)
	int32_t*		doc_max;	  // [n_docs] max depth after a bracket (starts at 0)
	int32_t*		doc_last;	  // [n_docs] depth after the document
	uint32_t*		bdoc;		  // [per bracket] document index
	uint32_t		tile_rows;	  // tile_doc was filled for tiles of this many rows (0: not yet)
	uint32_t*		pack;		  // optional: per-document results packed for one copy (k_batch_finish)
};

// The state a chunk of a longer document starts from, kept on the device when
// chunks are chained without the host (pars_hip_scan_stream).
struct DevCarry {
	uint32_t states;	// packed per-descriptor states before the chunk
	uint32_t out_base;	// structurals before the chunk
	uint32_t br_base;	// brackets before the chunk
	int32_t	 depth;		// depth before the chunk
};

// R6's stack path pairs brackets in chunks of this many (pair_stack.hip); the
// scan kernel can leave keys only at these chunks' first brackets.
constexpr uint32_t kStkChunk = 8192;  // brackets per chunk (256 threads x 32)

struct ScanArgs {
	const uint8_t*	data;
	uint64_t		n;
	uint32_t		n_tiles;  // set by the launcher from the evaluator's rows per tile
	uint32_t		epoch;	  // filled in by the kernel from sum->epoch
	const DevPlan*	plan;
	uint64_t*		recs;	  // one record per tile; sized for one-row tiles
	uint32_t*		ticket;	  // [0] next tile, [1] blocks finished; both 0 between launches
	uint32_t*		done;
	DevSummary*		sum;
	uint32_t*		positions;
	uint8_t*		chars;	   // optional
	uint64_t*		bitmap;	   // optional, ceil(n/64) words
	int32_t*		pair_pos;  // optional: -1 written for every structural index
	uint32_t*		bpos;
	uint32_t*		bsidx;	   // optional: only pair_pos and the tape's partners read it
	int32_t*		bkey;
	uint8_t*		bchar;
	// one document: keys only for bracket indices that are multiples of
	// kStkChunk (what the stack path reads); a sort of the document first fills
	// the rest (PairArgs::key_fill).  Batches always write every key.
	uint32_t		key_sparse;
	uint64_t		cap_struct;
	uint64_t		cap_br;
	uint64_t		byte_base;	 // added to every emitted position (chunked mode)
	uint32_t		out_base;	 // structural index of the first output (chunked mode)
	uint32_t		br_base;	 // bracket index of the first bracket (chunked mode)
	int32_t			depth_base;	 // depth before the first byte (chunked mode)
	uint32_t		in_states;	 // packed per-descriptor states before the first byte
	// Tuning (launch shape): predecessor records read per look-back round trip
	// (8..64), and resident blocks per CU of the persistent grid (0 = as many
	// as fit).
	uint32_t		lb_window;
	uint32_t		blocks_per_cu;
	uint32_t		chase;		 // set by the launcher: the last block chases aggregates (one chain)
	const DevCarry* carry;		 // when set, in_states / out_base / br_base / depth_base come from here
	BatchArgs		b;			 // batch mode only
};

// Blocks per CU the most recent scan launch on this thread could keep
// resident (its occupancy limit), so a tuner can sweep below it.
int last_scan_occupancy();

// R1-R5: one pass over the input, general descriptor path.  `rows` = 16 KiB
// rows per tile (1..4; 0 = the default); tuned per plan.
hipError_t launch_scan_general(const ScanArgs& a, int n_descs, bool wide, int rows, hipStream_t s);
// R1-R5 for plans whose one descriptor is an affine escape automaton (JSON).
hipError_t launch_scan_escape(const ScanArgs& a, int rows, int waves, hipStream_t s);
// R1-R5 for plans whose one descriptor is quote parity (CSV doubled quotes).
hipError_t launch_scan_parity(const ScanArgs& a, int rows, int waves, hipStream_t s);
// The escape / parity evaluators with six- or eight-wave blocks (waves 6, 8;
// one document): scan_wide_kernels.hip.
hipError_t launch_scan_wide(const ScanArgs& a, bool parity, int rows, int waves, hipStream_t s);
// R1-R5 for plans of two to five descriptors whose background runs act as one
// byte (DevPlan::ev_ok): the lanes step at their event bytes only.
bool	   event_path_applies(const DevPlan& p);
uint32_t   event_tile_rows(int rows);	// 16 KiB rows per tile of the event launchers (1 or 2)
hipError_t launch_scan_event(const ScanArgs& a, int n_descs, bool wide, int rows, hipStream_t s);
hipError_t launch_batch_event(const ScanArgs& a, int n_descs, bool wide, int rows, hipStream_t s);
// The same three over a batch (a.b filled in); the document heads must be set.
hipError_t launch_batch_general(const ScanArgs& a, int n_descs, bool wide, int rows, hipStream_t s);
hipError_t launch_batch_escape(const ScanArgs& a, int rows, hipStream_t s);
hipError_t launch_batch_parity(const ScanArgs& a, int rows, hipStream_t s);
// Batch preparation in two launches: reset the summary (as launch_summary_reset
// with a fresh document), clear the head bitmap, reset the per-document
// statistics and doc_bad, fill tile_doc for tiles of `tile_rows` rows; then set
// the head bits.
hipError_t launch_batch_prep(const BatchArgs& b, uint64_t n, DevSummary* sum, uint32_t* doc_bad, uint32_t tile_rows,
							 hipStream_t s);
// Rows per tile the batch launchers will use for this plan and options (the
// launch recomputes tile_doc itself if it differs).
uint32_t batch_tile_rows(const DevPlan& p, int path, int rows);
// tile_doc for tiles of `tile_bytes` (called by the batch launchers).
hipError_t launch_tile_doc(const BatchArgs& b, uint32_t n_tiles, uint64_t tile_bytes, hipStream_t s);
// Per-document flags once R6 has run: bit0 balanced, bit1 well_formed, bit2
// unclosed_mode; start_packed = the plan's start states.  With b.pack, also the
// per-document results in one block of u32 (n = n_docs): [0, n] structural
// starts, [n+1, 2n+1] bracket starts, [2n+2, 3n+1] max depths, [3n+2, 4n+1]
// flags, [4n+2] the summary's overflow word.
hipError_t launch_batch_finish(const BatchArgs& b, const uint32_t* doc_bad, uint32_t start_mask,
							   uint32_t start_packed, int run_pairs, uint8_t* doc_flags, const DevSummary* sum,
							   hipStream_t s);
inline uint64_t batch_pack_words(uint64_t n_docs) {
	return 4 * n_docs + 3;	// batch_pack_bytes (internal.hpp) / 4
}
// Initial summary values: a fresh document starts all at 0; a chunk of a
// longer document starts from the running values its carry holds.
struct SummaryInit {
	uint32_t n_struct	= 0;
	uint32_t n_brackets = 0;
	int32_t	 depth		= 0;
	int32_t	 min_depth	= 0;
	int32_t	 max_depth	= 0;
};

hipError_t launch_summary_reset(DevSummary* sum, const SummaryInit& in, hipStream_t s);
// Between two chained chunks: the next chunk's carry from the summary the last
// one left (its running counts, depth extremes and error bits carry on), and
// a new look-back epoch.
hipError_t launch_carry_next(DevSummary* sum, DevCarry* carry, hipStream_t s);

// R6: stable sort of bracket indices by depth key, then adjacent pairing.
struct PairArgs {
	DevSummary*	   sum;
	const DevPlan* plan;
	const int32_t* bkey;
	const uint8_t* bchar;
	const uint32_t* bsidx;
	uint32_t	   br_base;	  // bsidx values are relative to this structural base
	uint64_t	   cap_br;
	uint32_t*	   ukeys[2];  // radix ping-pong keys
	uint32_t*	   order[2];  // radix ping-pong values (bracket indices)
	uint32_t*	   rank;	  // sorted position of each bracket (last pass)
	uint32_t*	   hist;	  // 256 * max_chunks
	uint32_t*	   hist_part; // scan partials
	uint64_t	   max_chunks;
	int32_t*	   pair;	  // CPU convention, per bracket index (optional)
	int32_t*	   pair_pos;  // device convention, per structural index (optional)
	uint32_t*	   close;	  // per opening bracket: the structural index of the next bracket at its
							  // key in sorted order, ~0u if none (optional; one document).  When set,
							  // the pair pass writes only this: no pairs, no mismatch report
	uint32_t	   out_base;  // structural index base of pair_pos entries
	// batch mode: sort by (document, key); pair within a document; indices in
	// the outputs are document-local.  Document d's keys k map to
	// doc_kbase[d] + k - (doc_min[d] - 1), doc_kbase being the exclusive prefix
	// of the per-document key spans doc_max[d] - doc_min[d] + 2
	const uint32_t* bdoc;
	const uint32_t* br_start;
	const uint32_t* st_start;
	uint32_t		n_docs;
	uint32_t*		doc_bad;  // [n_docs] a sorted pair of the document does not match
	const int32_t*	doc_min;
	const int32_t*	doc_max;
	uint32_t*		doc_kbase;	// [n_docs]
	uint64_t*		kb_part;	// [256] partial sums of the key-base scan
	uint32_t		max_digit_bits;	 // radix digit width cap per pass (1..8; 0 = 8)
	// One-block R6 for at most kSmallBrackets brackets: 0 = not tried, 1 = tried
	// first (the radix path then runs only if it did not apply), 2 = the only
	// path (the caller guarantees it applies: see small_pairing_mode)
	uint32_t		small;
	// one document scanned with sparse keys: the key of every bracket is
	// rebuilt here first (the same buffer as bkey) from the bracket bytes and
	// the keys at the chunk starts
	int32_t*		key_fill;
};

constexpr uint32_t kSmallBrackets = 8192;
// The R6 path for n input bytes in n_docs documents (1 for a single document):
// the one-block path is certain when n bounds the bracket count to
// kSmallBrackets (the (document, key) range is then at most n + 2 n_docs + 1,
// within its 19 bits), worth trying up to 1 MiB, and skipped above.
inline uint32_t small_pairing_mode(uint64_t n, uint64_t n_docs) {
	if (n <= kSmallBrackets && n + 2 * n_docs + 2 < (1ull << 19))
		return 2;
	return n <= (1ull << 20) ? 1u : 0u;
}

hipError_t launch_pairing(const PairArgs& a, hipStream_t s);

// R6 for balanced documents by stack matching (pair_stack.hip): the same pairs
// the sort gives when the depth never drops below 0 and ends at 0; the kernels
// exit at once otherwise and the sorting path runs instead.
constexpr uint32_t kStackLost = 0x10u;		  // summary overflow bit: a residue close found no open (internal)
struct StackArgs {
	DevSummary*		sum;
	const DevPlan*	plan;
	const uint8_t*	bchar;
	const int32_t*	bkey;
	const uint32_t* bsidx;
	uint64_t		cap_br;
	int32_t*		pair;	   // CPU convention, per bracket index
	int32_t*		pair_pos;  // optional: structural-index convention
	uint32_t*		res;	   // [cap_br] per chunk: unmatched closes, then unmatched opens (bracket indices)
	int32_t*		crec;	   // [3 max_chunks] per chunk: entry, minimum and exit depth
	// minima of groups of 64 and of 4096 chunks, tagged with the launch epoch
	unsigned long long* gmin1;	// [max_chunks / 64 + 1]
	unsigned long long* gmin2;	// [max_chunks / 4096 + 1]
	uint64_t		max_chunks;
	// batch mode (null for one document): pairs and pair_pos values are
	// document-relative, a mismatched pair marks its document in doc_bad, and
	// the path applies when every document is balanced; it then marks the
	// summary's small_done so the sorting chain launched after it exits at once
	const uint32_t* bdoc;
	const uint32_t* br_start;
	const uint32_t* st_start;
	uint32_t*		doc_bad;
};
hipError_t launch_stack_pairing(const StackArgs& a, hipStream_t s);

// Fill n int32 values (graph-safe replacement for hipMemsetAsync).
hipError_t launch_fill_i32(int32_t* p, uint64_t n, int32_t v, hipStream_t s);

// ---------------------------------------------------------------------------
// The JSON tape (dom_kernels.hip): the reference's replay-driven JSON DOM hook
// (bench/bench_parse.cu:41-348) computed from a document's structural index
// and R6's depth-sorted brackets.  Events: the gap before structural i is 2i,
// structural i is 2i+1, the gap after the last structural is 2 n_struct.
constexpr uint64_t kDomNone = ~0ull;

struct DomSummary {
	unsigned long long fail_key;  // min over failing events of (event << 1 | partial node)
	unsigned long long e1, e2;	  // the first and second top-level value events
	unsigned long long fail_event;  // resolved: the event the hook stopped at, or kDomNone
	uint64_t		   n_nodes, arena_bytes;
	uint32_t		   ok, partial;
};

struct DomArgs {
	const uint8_t*	data;
	uint64_t		n;
	const uint32_t* positions;
	uint64_t		n_struct;  // tokens = n_struct + 1 (the last one has only a gap)
	uint64_t		nb;		   // brackets
	const DevPlan*	plan;
	const DevSummary* sum;	   // R1-R6 summary of the same document (final depth)
	const uint8_t*	bchar;
	const int32_t*	bkey;
	const uint32_t* bsidx;
	const uint32_t* order;	   // R6: bracket at each sorted position
	// per token
	uint8_t*  tinfo;	// symbol class, gap kind, gap parsed whole (7 bits)
	uint32_t* gbeg;
	uint32_t* gend;
	uint32_t* gdec;
	// per kDomSub-token block: exclusive prefixes of nodes, decoded bytes and
	// brackets at its first token (the kernels scan within a block themselves)
	uint32_t* sub;
	// per 4096-token group: its two smallest top-level value events
	unsigned long long* top2;
	uint32_t* tblk;	  // per token block: sums, then exclusive prefixes (3 each)
	uint64_t  n_tblk;
	// per sorted bracket position / per bracket
	int32_t*  hblk;	  // per sorted block: last head, then the exclusive max
	uint64_t  n_hblk;
	int32_t*  parent;  // per bracket: the container that encloses its pair, or -1
	uint32_t* close;   // per opening bracket: the structural index of its partner, or ~0u
	DomSummary* ds;
	// outputs
	uint8_t*  kind;
	uint32_t* nbeg;
	uint32_t* nend;
	uint32_t* soff;
	uint8_t*  arena;
	uint64_t  cap_nodes, cap_arena;
};

constexpr uint64_t kDomTile = 4096;	 // tokens or sorted brackets per block
constexpr uint64_t kDomSub = 1024;	// tokens per block of the per-gap kernels

hipError_t launch_json_tape(const DomArgs& a, hipStream_t s);

// ---------------------------------------------------------------------------
// The INI / TOML key-value tables (kv_kernels.hip): the reference's stage-2
// hooks for these formats (bench/bench_parse_host.cu:30-134, 202-248)
// computed from a document's bytes and structural bitmap.  A tile is 256
// lanes of 64 bytes; per-lane arrays are indexed tile * 256 + lane.
constexpr uint32_t kKvLanes = 256;
constexpr uint32_t kKvTile	= kKvLanes * 64;

struct KvSummary {
	uint64_t n_tokens;
	uint32_t open_end;	// a token runs to the document's end
	uint32_t bad;		// the table pass met more tokens than the count (internal error)
	uint32_t roles[5];	// tokens per role
	uint32_t pad2;
};

struct KvArgs {
	const uint8_t*	data;
	uint64_t		n;
	const uint64_t* bitmap;	 // structural bitmap, ceil(n / 64) words
	uint32_t		n_tiles;
	int				toml;	 // 0 INI, 1 TOML
	// per tile (n_tiles each)
	uint32_t* t_tmap;	// tokenizer map of the tile (packed)
	uint32_t* t_tin;	// tokenizer state entering the tile
	uint32_t* t_hmap;	// hook map of the tile (packed)
	uint32_t* t_hin;	// hook state entering the tile
	uint32_t* t_cnt;	// tokens starting in the tile
	uint32_t* t_off;	// tokens starting before the tile
	int32_t*  t_key;	// TOML: the tile's last KEY token, -1 if none
	int32_t*  t_kin;	// TOML: the last KEY token before the tile, -1 if none
	// per lane
	uint32_t* l_tmap;  // tokenizer map from the tile's start to the lane's (packed)
	uint32_t* l_hmap;  // hook map from the tile's start to the lane's (packed)
	uint32_t* l_off;   // tokens starting earlier in the tile | (a token is open at the lane's start) << 31
	uint64_t* l_start; // bit j: a token starts at byte j
	uint64_t* l_first; // ... and is the first of its gap
	uint64_t* l_end;   // bit j: a token ends at byte j (before it) or with it (closing quote)
	uint64_t* l_endq;  // bit j: the token ending at j ends with it (a closing quote)
	KvSummary* sum;
	// outputs
	uint8_t*  role;
	uint32_t* beg;
	uint32_t* end;
	int32_t*  key;	// TOML only
};

// The tokenizer and hook passes up to the token count (sum->n_tokens is ready
// once they are done) ...
hipError_t launch_kv_count(const KvArgs& a, hipStream_t s);
// ... and the passes that write the table (role, beg, end and, TOML, key).
hipError_t launch_kv_emit(const KvArgs& a, hipStream_t s);

}  // namespace pars_hip
