/* pars_hip.h — C ABI of the MI355X (gfx950) scan_plan backend.
 *
 * Drop-in for the reference CUDA backend's boundary,
 * include/pars/backend/cuda/pipeline.cuh:1164-1281 (class gpu::scanner<S> and
 * the free functions gpu::scan_device / scan_host / scan_chunk).  The plan is a
 * runtime value (pars_plan, byte-compatible with pars::scan_plan) instead of a
 * template argument; include/pars_hip.hpp restores the template API on top of
 * this header.  Plain pointers and sizes only.
 *
 * Conventions (differences from the reference are deliberate and listed):
 *   - every function returns a pars_status; pars_hip_last_error() describes
 *     the most recent failure on the calling thread (the reference calls
 *     exit(1), pipeline.cuh:520-555);
 *   - positions are uint32 byte offsets (reference: int32, which overflows
 *     past 2^31, pipeline.cuh:828,840); the values are identical below 2^31;
 *   - device outputs alias the scanner's arena and stay valid until the next
 *     call on the same scanner (pipeline.cuh:1013-1014, same rule);
 *   - a scanner is not thread-safe; use one per thread (pipeline.cuh:1264);
 *   - the input is read only within [0, n): no padding is required and the
 *     input is never written (the reference reads ceil(n/32)*32 bytes and
 *     scan_chunk zeroes the pad through a const_cast, pipeline.cuh:1042-1043).
 */
#ifndef PARS_HIP_H
#define PARS_HIP_H

#include "pars_plan.h"

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t (`struct ihipStream_t*`); NULL = the null stream. */
typedef struct ihipStream_t* pars_stream;

typedef struct pars_scanner pars_scanner;

typedef enum {
	PARS_OK			  = 0,
	PARS_ERR_ARGUMENT = 1, /* bad pointer, size or option */
	PARS_ERR_PLAN	  = 2, /* the plan violates a plan.hpp cap */
	PARS_ERR_MEMORY	  = 3, /* device or pinned allocation failed */
	PARS_ERR_DEVICE	  = 4, /* a HIP call failed */
	PARS_ERR_BUDGET	  = 5, /* serving: the request does not fit the HBM budget */
	PARS_ERR_BUSY	  = 6  /* serving: no free slot right now, retry */
} pars_status;

/* pipeline.cuh:745-748: per-stage milliseconds measured with events on the
 * call's own stream.  p1 = R1-R5 (one fused kernel: the stages have no
 * separate launches), p2 = the stage-2 kernels (the JSON tape's,
 * pars_hip_json_tape, or the key-value table's, pars_hip_kv_parse; else 0),
 * p3 = 0, p4 = R6 pairing (all of its kernels; a kernel trace splits them). */
typedef struct {
	double p1, p2, p3, p4;
} pars_stage_times;

/* pipeline.cuh:722-728 plus the extra outputs this backend produces. */
typedef struct {
	const uint32_t* structural;	  /* device: byte offsets of unmasked structural bytes */
	const int32_t*	pair_pos;	  /* device: partner structural index or -1, or NULL */
	uint32_t		n_struct;
	int32_t			balanced;
	int32_t			well_formed;
	int32_t			max_depth;
	/* extensions (not in the reference device_result) */
	const uint8_t*	chars;		  /* device: structural bytes, or NULL unless enabled */
	const uint64_t* struct_bitmap; /* device: ceil(n/64) words, bit i = byte i emitted */
	const uint32_t* bracket_pos;  /* device: byte offsets of unmasked brackets */
	const uint8_t*	bracket_char; /* device */
	const int32_t*	bracket_pair; /* device: CPU convention (nest.hpp:18), or NULL */
	uint32_t		n_brackets;
	int32_t			unclosed_mode; /* scan.hpp:66-70 */
} pars_device_result;

/* pipeline.cuh:730-735; buffers are owned by the scanner (pinned host memory),
 * valid until the next call on it. */
typedef struct {
	const uint32_t* structural;
	const int32_t*	pair_pos; /* NULL unless want_pairs */
	uint32_t		n_struct;
	int32_t			balanced;
	int32_t			well_formed;
	int32_t			max_depth;
	int32_t			unclosed_mode;
} pars_host_result;

/* pipeline.cuh:737-743, widened: `scan` holds the per-descriptor states after
 * the chunk (3 bits each, plus bit 31 set = valid; 0 = document start, so a
 * zero-initialised carry starts a document exactly as the reference's `{}`).
 * depth/bracket_base/min_depth/max_depth carry R6 state across chunks. */
typedef struct {
	uint64_t scan;
	uint32_t out_base;
	uint32_t bracket_base;
	int32_t	 depth;
	int32_t	 min_depth;
	int32_t	 max_depth;
	uint32_t reserved;
} pars_chunk_carry;

/* Options (pars_hip_set_option). */
typedef enum {
	PARS_OPT_WANT_CHARS	 = 1, /* also emit structural bytes (default 0) */
	PARS_OPT_WANT_BITMAP = 2, /* also emit the structural bitmap (default 1) */
	PARS_OPT_PATH		 = 3, /* 0 = auto, 1 = force the general descriptor path */
	PARS_OPT_ROWS_PER_TILE = 4, /* 16 KiB rows per scan tile: 0 = default, 1..4 (bytes per lane: 64 x rows) */
	PARS_OPT_GRAPH		   = 5, /* batches: capture the launch chain once per shape and replay it */
	PARS_OPT_LOOKBACK_WINDOW = 6, /* predecessor records per look-back round trip: 0 = default (32), 8..64 */
	PARS_OPT_BLOCKS_PER_CU	 = 7, /* resident scan blocks per CU (4 waves each): 0 = as many as fit */
	PARS_OPT_RADIX_BITS		 = 8, /* pairing sort: digit width cap per pass, 0 = default (8), 1..8 */
	PARS_OPT_CHASER			 = 9, /* one-descriptor plans: one block publishes inclusive prefixes ahead (default 0) */
	PARS_OPT_LAST_PAIRING	 = 10, /* read-only (pars_hip_get_option): how R6 of the last scan_device /
									* scan_host / scan_batch call ran, a pars_pairing value */
	PARS_OPT_WAVES_PER_BLOCK = 11  /* scan blocks of one document under the bit-parallel evaluators
									* (JSON- and CSV-style plans): 0 or 4 (default), 6, 8 waves; the
									* row grows with the block (64 bytes per lane); other paths and
									* batches keep four */
} pars_option;

/* How R6 paired brackets (PARS_OPT_LAST_PAIRING). */
typedef enum {
	PARS_PAIRING_NONE	   = 0, /* no pairing ran */
	PARS_PAIRING_ONE_BLOCK = 1, /* one block, bitonic sort in LDS (small inputs) */
	PARS_PAIRING_STACK	   = 2, /* stack matching: a balanced document, or a batch of balanced documents */
	PARS_PAIRING_SORT	   = 3	/* stable radix sort by (document,) depth */
} pars_pairing;

/* One document of a batch: its bytes are d_in[offset, offset + size). */
typedef struct {
	uint64_t offset; /* a multiple of 64; documents in ascending order, not overlapping */
	uint64_t size;	 /* >= 1 */
} pars_doc_span;

/* Results of pars_hip_scan_batch; every pointer is device memory owned by the
 * scanner and valid until its next call.  Per-document runs are concatenated in
 * document order: document d's structural positions are
 * structural[struct_start[d] .. struct_start[d + 1]), byte offsets inside the
 * document (what scanning the document alone yields); its brackets likewise
 * through bracket_start, and bracket_pair holds document-local bracket
 * indices (nest.hpp pair_table convention), -1 when unpaired. */
typedef struct {
	uint64_t		n_docs;
	uint64_t		n_struct;		/* all documents */
	uint64_t		n_brackets;
	const uint32_t* structural;
	const uint32_t* struct_start;	/* [n_docs + 1] */
	const uint32_t* bracket_pos;	/* document-local byte offsets */
	const uint8_t*	bracket_char;
	const uint32_t* bracket_start;	/* [n_docs + 1] */
	const int32_t*	bracket_pair;	/* with run_pairs; else NULL */
	const int32_t*	pair_pos;		/* with want_pairs: document-local structural index of the partner, -1 */
	const uint8_t*	doc_flags;		/* [n_docs] bit0 balanced, bit1 well_formed, bit2 unclosed_mode */
	const int32_t*	max_depth;		/* [n_docs] */
	const uint64_t* struct_bitmap;	/* bit per byte of d_in (bytes outside documents are 0), or NULL */
} pars_batch_result;

/* ---- device memory: an injectable allocator and a shared budget ------------
 * Reference: DeviceCache (pipeline.cuh:562-714) cudaMallocs its buffers and
 * exits the process when an allocation fails (:520-555).  Here a scanner gets
 * its buffers from the caller's allocator (default: hipMalloc / hipFree),
 * charged against an optional budget that several scanners and servers can
 * share; an allocation that fails or would pass the budget returns
 * PARS_ERR_MEMORY, and the scanner stays usable for later calls. */
typedef struct {
	/* device memory of at least `bytes`, or NULL when none can be had; called on
	 * the calling host thread, outside any kernel, never during graph capture
	 * of a replayed batch (buffers are sized before capture) */
	void* (*alloc)(void* ctx, size_t bytes);
	/* returns what alloc gave (same pointer and size) */
	void (*release)(void* ctx, void* p, size_t bytes);
	void* ctx;
} pars_allocator;

typedef struct pars_budget pars_budget;

/* A limit of `limit_bytes` on the device memory its users hold at once. */
int pars_hip_budget_create(uint64_t limit_bytes, pars_budget** out);
/* PARS_ERR_ARGUMENT while a scanner or server still uses it. */
int pars_hip_budget_destroy(pars_budget* b);
/* Bytes charged now, the most ever charged at once, and the limit (any may be NULL). */
int pars_hip_budget_usage(const pars_budget* b, uint64_t* in_use, uint64_t* peak, uint64_t* limit);

/* ---- scanner lifetime (reference: gpu::scanner<S> constructor/destructor) ---- */
int pars_hip_scanner_create(const pars_plan* plan, pars_scanner** out);
/* The same with the scanner's device memory from `alloc` (NULL: hipMalloc)
 * and charged to `budget` (NULL: no limit). */
int pars_hip_scanner_create_ex(const pars_plan* plan, const pars_allocator* alloc, pars_budget* budget,
							   pars_scanner** out);
int pars_hip_scanner_destroy(pars_scanner* sc);
int pars_hip_set_option(pars_scanner* sc, int option, int64_t value);

/* pipeline.cuh:1167-1172 scanner::scan_device.  d_in: device bytes [0, n),
 * n < 2^32.  want_pairs: fill pair_pos (structural-index convention) and
 * bracket_pair.  run_pairs: run R6 (flags stay as computed by the depth scan
 * otherwise); pairs are int32 bracket indices, so a document with more than
 * 2^31-1 brackets is refused with pairing (PARS_ERR_ARGUMENT).  st: optional
 * stage times.  Blocks until results are ready (like the reference, which
 * returns counts by value). */
int pars_hip_scan_device(
	pars_scanner* sc, const uint8_t* d_in, size_t n, pars_stream stream, int want_pairs, int run_pairs,
	pars_stage_times* st, pars_device_result* out
);

/* No counterpart in the reference: pars_hip_scan_device for callers that
 * cannot allocate device memory - code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The caller passes the input and
 * every output buffer (positions[cap_struct]; optionally struct_bitmap,
 * ceil(n/64) words; bracket_pos / bracket_char [cap_brackets]; bracket_pair
 * [cap_brackets] when run_pairs); all scratch (look-back records, summary, the
 * bracket hand-off, R6's residues, keys and histograms) lives in a static
 * device array of about 34 MiB that each GPU holds until the process ends, so
 * the call performs no hipMalloc / hipFree or other allocation.  Limits:
 * n <= 32 MiB, cap_brackets <= 2^20; a document with more structural or
 * bracket bytes than the capacities returns PARS_ERR_ARGUMENT.  The scanner
 * supplies its plan and launch options; chars and pair_pos are not emitted.
 * Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (two streams, two threads) race, so
 * serialise them.  Blocks until the results are ready. */
int pars_hip_scan_device_noalloc(
	pars_scanner* sc, const uint8_t* d_in, size_t n, pars_stream stream, int run_pairs, uint32_t* positions,
	size_t cap_struct, uint64_t* struct_bitmap, uint32_t* bracket_pos, uint8_t* bracket_char, int32_t* bracket_pair,
	size_t cap_brackets, pars_device_result* out
);

/* pipeline.cuh:1176-1212 scanner::scan_host: pinned staging, H2D, scan,
 * D2H of the structural index (and pair_pos when want_pairs). */
int pars_hip_scan_host(
	pars_scanner* sc, const uint8_t* h_in, size_t n, pars_stream stream, int want_pairs,
	pars_host_result* out
);

/* pipeline.cuh:1216 scanner::reserve_struct: size the structural output for a
 * chunked document of up to `cap` structural bytes. */
int pars_hip_reserve_struct(pars_scanner* sc, size_t cap);

/* pipeline.cuh:1218-1223 scanner::scan_chunk: stage 1 of one chunk, writing at
 * in->out_base, positions offset by byte_base; works for every plan (the
 * reference's carry covers one packed 16-slot transform).  Blocks until done. */
int pars_hip_scan_chunk(
	pars_scanner* sc, const uint8_t* d_chunk, size_t n, const pars_chunk_carry* in, pars_stream stream,
	size_t byte_base, pars_chunk_carry* out
);

/* Streaming host boundary (the reference's scan_host, pipeline.cuh:1176-1212,
 * copies and uploads the whole input before scanning): the host bytes are
 * staged chunk by chunk (chunk_bytes, default 64 MiB, rounded to 64 KiB) into
 * two pinned buffers by copy_threads threads (default 8), uploaded on a copy
 * stream, and scanned on `stream` with the state carried from chunk to chunk
 * on the device, so host copies, PCIe transfers and kernels overlap; R6 then
 * pairs the whole document.  Results as pars_hip_scan_device (device memory,
 * whole-document positions); st->p1 = first chunk to last scan, p4 = R6. */
int pars_hip_scan_stream(
	pars_scanner* sc, const uint8_t* h_in, size_t n, size_t chunk_bytes, int copy_threads, pars_stream stream,
	int want_pairs, pars_stage_times* st, pars_device_result* out
);

/* R6 over everything the chunks of one document emitted (the reference has
 * no pairing across chunks, docs/model/model.md:355-359). */
int pars_hip_pair_chunks(
	pars_scanner* sc, const pars_chunk_carry* last, pars_stream stream, int want_pairs,
	pars_device_result* out
);

/* Many documents in one persistent launch (no reference counterpart: the
 * reference runs pipeline.cuh:752-1017 once per document).  Segment heads
 * reset the state carry and the depth; each document's results equal those of
 * scanning it alone.  docs: host array of n_docs spans into d_in[0, n), n <
 * 2^32.  Blocks until results are ready.  With PARS_OPT_GRAPH set and a
 * non-default stream, the launch chain is captured into a HIP graph the first
 * time a shape (d_in, n, n_docs, flags) is seen and replayed after that. */
int pars_hip_scan_batch(
	pars_scanner* sc, const uint8_t* d_in, size_t n, const pars_doc_span* docs, size_t n_docs, pars_stream stream,
	int want_pairs, int run_pairs, pars_stage_times* st, pars_batch_result* out
);

/* ---- serving (no reference counterpart) ------------------------------------
 * A server accepts parse requests as they arrive, stages their bytes in a
 * pinned host ring, copies them asynchronously into a device ring buffer, and
 * runs them through the batched scan with continuous batching: whatever is
 * pending when a scanner slot frees up forms the next launch, so requests that
 * arrive while others are on the GPU join the next batch while the previous
 * one still runs.  Admission is gated by an HBM budget (inputs in the ring
 * plus worst-case outputs of every request not yet released).  Results stay on
 * the device until released.  Each result equals scanning the request alone. */
typedef struct pars_server pars_server;

typedef struct {
	const pars_plan* plans;			 /* n_plans plans; requests name one by index */
	uint32_t		 n_plans;
	uint32_t		 slots_per_plan;  /* scanners per plan (batches in flight), default 2 */
	uint64_t		 ring_bytes;	  /* device input ring, all plans (default 4 GiB) */
	uint64_t		 staging_bytes;	  /* pinned host staging, all plans (default 1 GiB) */
	uint64_t		 hbm_budget;	  /* bytes the server may hold on the device (default 64 GiB) */
	uint64_t		 max_batch_bytes; /* per launch (default 64 MiB; also the largest request) */
	uint32_t		 max_batch_docs;  /* per launch (default 16384) */
	uint32_t		 run_pairs;		  /* run R6 (default 1) */
	uint32_t		 submit_threads;  /* threads pars_hip_server_replay submits from (default 4) */
	uint32_t		 reserved;
	/* optional, shared with other scanners and servers: charged with the rings,
	 * the scanners' buffers and, from a batch's launch until release, each
	 * request's worst-case footprint (18 bytes per input byte); a batch waits
	 * for budget like for hbm_budget */
	pars_budget*	 budget;
} pars_server_config;

typedef struct {
	uint64_t		id;
	int32_t			status;	   /* PARS_OK, or the error that failed the request's batch */
	uint32_t		plan;
	uint64_t		n_bytes;
	uint64_t		n_struct;
	uint64_t		n_brackets;
	const uint32_t* structural;	   /* device; document-local byte offsets */
	const uint32_t* bracket_pos;   /* device */
	const uint8_t*	bracket_char;  /* device */
	const int32_t*	bracket_pair;  /* device, document-local indices (NULL without R6) */
	uint32_t		flags;		   /* bit0 balanced, bit1 well_formed, bit2 unclosed_mode */
	int32_t			max_depth;
	uint64_t		t_submit_ns;   /* host steady clock */
	uint64_t		t_launch_ns;   /* its batch was enqueued */
	uint64_t		t_done_ns;	   /* its batch's outputs were complete on the device */
	uint32_t		batch_docs;	   /* documents in its batch */
	uint32_t		reserved;
	const uint8_t*	input;		   /* device: the request's bytes in the input ring, valid until release */
} pars_request_result;

typedef struct {
	uint64_t submitted, completed, released;
	uint64_t batches;
	uint64_t bytes_completed;
	uint64_t busy_rejects;		 /* submits refused for lack of staging space */
	uint64_t budget_waits;		 /* batch formations cut short by the HBM budget */
	uint64_t device_bytes;		 /* held on the device (rings + scanner arenas) */
	uint64_t inflight_reserved;	 /* budget held by admitted, unreleased requests */
	uint64_t ring_wraps;		 /* device ring allocations that wrapped to offset 0 past live data */
	uint64_t staging_wraps;		 /* the same for the pinned staging ring */
	uint64_t worker_launch_ns;	 /* host time the worker spent forming and enqueueing batches */
	uint64_t worker_retire_ns;	 /* host time the worker spent retiring batches */
	uint64_t submit_ns;			 /* host time inside submit, summed over the submitting threads */
	uint64_t submit_lock_ns;	 /* of which waiting for the server lock */
} pars_server_stats;

int pars_hip_server_create(const pars_server_config* cfg, pars_server** out);
int pars_hip_server_destroy(pars_server* sv);
/* Copies n bytes into the pinned staging ring; PARS_ERR_BUSY when it is full
 * (retry after polling and releasing).  Thread-safe: concurrent submitters
 * copy in parallel (only the ring reservation is serialised). */
int pars_hip_server_submit(pars_server* sv, uint32_t plan, const uint8_t* data, size_t n, uint64_t* id);
/* Up to cap completed requests not returned before; waits up to wait_us for
 * the first one. */
int pars_hip_server_poll(pars_server* sv, pars_request_result* out, size_t cap, size_t* n_out, int64_t wait_us);
/* The request's outputs are no longer needed (frees its budget; a scanner
 * slot, and its batch's region of the device input ring, are reused once all
 * of the batch's requests are released). */
int pars_hip_server_release(pars_server* sv, uint64_t id);
int pars_hip_server_stats(pars_server* sv, pars_server_stats* out);
/* Load driver: submits n requests (host bytes data[i], sizes[i], plan[i]) at
 * arrival_ns[i] after the start (all 0: as fast as admission allows) from
 * submit_threads threads (request i from thread i mod submit_threads, each in
 * arrival order), while the calling thread polls and releases as they
 * complete; latency_ns[i] = completion - scheduled arrival.  *elapsed_ns =
 * first arrival to last completion. */
int pars_hip_server_replay(
	pars_server* sv, size_t n, const uint32_t* plan, const uint8_t* const* data, const uint64_t* sizes,
	const uint64_t* arrival_ns, uint64_t* latency_ns, uint64_t* elapsed_ns
);

/* Current value of an option. */
int pars_hip_get_option(const pars_scanner* sc, int option, int64_t* value);

/* ---- autotuner (no reference counterpart) ---------------------------------
 * Times the scanner on the caller's input - one document, or a batch when docs
 * is non-NULL - over its launch knobs and keeps the fastest setting on the
 * scanner: rows per tile (bytes per lane and tile = 64 x rows), look-back
 * window, resident blocks per CU (waves per SIMD), waves per block (one
 * document, bit-parallel evaluators), and the pairing sort's radix digit cap.  Coordinate descent from the current setting; a candidate
 * must beat the best so far by 2 %; each is the median of `reps` runs (1..15)
 * after a warm-up.  Overwrites the scanner's outputs; blocks. */
typedef struct {
	int32_t	 rows, lb_window, blocks_per_cu, radix_bits;  /* chosen option values (0 = default) */
	float	 scan_ms, pair_ms;							  /* with the chosen setting */
	float	 base_scan_ms, base_pair_ms;				  /* with the setting before tuning */
	uint32_t trials;									  /* runs timed */
	uint32_t max_blocks_per_cu;							  /* the scan kernel's occupancy limit */
	int32_t	 waves_per_block;							  /* chosen option value (0 = default) */
} pars_tune_result;

int pars_hip_autotune(
	pars_scanner* sc, const uint8_t* d_in, size_t n, const pars_doc_span* docs, size_t n_docs, pars_stream stream,
	int reps, pars_tune_result* out
);

/* pipeline.cuh:1225 scanner::struct_ptr. */
const uint32_t* pars_hip_struct_ptr(const pars_scanner* sc);

/* The JSON tape: one node per value the reference's JSON DOM hook
 * (bench/bench_parse.cu:41-333, driven through model::replay, visit.hpp:35-48,
 * by pars_parse, bench_parse.cu:336-348) pushes, in the same order with the
 * same (kind, begin, end), plus the bytes each string decodes to.  All device
 * memory, valid until the scanner's next call. */
typedef struct {
	const uint8_t*	kind;	   /* 0 object, 1 array, 2 string, 3 number, 4 true, 5 false, 6 null */
	const uint32_t* begin;	   /* byte offset of the value's first byte (a container: its open) */
	const uint32_t* end;	   /* one past its last byte; a container: its close's offset; 0 while unclosed */
	const uint32_t* str_off;   /* decoded bytes so far: node k's string is arena[str_off[k], str_off[k+1]) */
	const uint8_t*	arena;	   /* every string's decoded bytes, in node order */
	const uint32_t* structural; /* the structural index the tape was built from */
	uint64_t		n_nodes;
	uint64_t		arena_bytes;
	uint64_t		n_struct;
	int64_t			fail_event; /* the first event the hook rejected, -1 if none: 2i = the run of
								 * bytes before structural i, 2i+1 = structural i, 2 n_struct =
								 * the run after the last one */
	int32_t			ok;		   /* the hook's verdict after its final checks (bench_parse.cu:196-201) */
	int32_t			pad;
} pars_json_tape;

/* Scan a JSON document on the device (R1-R6) and build its tape there: the
 * reference's stage-2 JSON hook (bench_parse.cu:336-348: scan_host, replay into
 * json_parser, finish) without the host replay.  The scanner's plan must be a
 * JSON plan: punctuation within ',' ':', brackets within {} [] (else
 * PARS_ERR_ARGUMENT).  Blocks until done; st->p1 = R1-R5, p4 = R6, p2 = the
 * tape kernels (ms), when st is not NULL. */
int pars_hip_json_tape(
	pars_scanner* sc, const uint8_t* d_in, size_t n, pars_stream stream, pars_stage_times* st, pars_json_tape* out
);

/* The INI / TOML key-value tables: every token the reference's stage-2 hooks
 * for these formats (bench/bench_parse_host.cu: ini_parse :69-134, toml_parse
 * :202-248, splitting gaps with each_token :34-56, driven through
 * model::replay, visit.hpp:35-48) meet, in document order, with the role the
 * hook gives it.  A token is a byte range of the input (quoted tokens keep
 * their quotes, as the hooks' string_views do). */
typedef enum { PARS_KV_INI = 0, PARS_KV_TOML = 1 } pars_kv_format;

typedef enum {
	PARS_KV_SECTION = 0, /* INI: sections; TOML: sections */
	PARS_KV_KEY		= 1, /* INI: keys (ini_parse::add(tok, false)); TOML: a key token (toml_parse::key) */
	PARS_KV_VALUE	= 2, /* INI: vals; TOML: the value of an entry (entries: (its key, it)) */
	PARS_KV_ELEMENT = 3, /* TOML: an array element (toml_parse::elems) */
	PARS_KV_SKIPPED = 4	 /* inside a comment the hook tracks (';' / '#' as a structural byte): not seen */
} pars_kv_role;

typedef struct {
	const uint8_t*	role;		/* pars_kv_role per token */
	const uint32_t* begin;		/* byte offset of its first byte */
	const uint32_t* end;		/* one past its last byte */
	const int32_t*	key;		/* TOML: a VALUE's key, the last KEY token before it (-1: none yet);
								 * -1 for the other roles.  NULL for INI */
	const uint32_t* structural; /* the structural index the table was built from */
	uint64_t		n_tokens;
	uint64_t		n_struct;
	uint64_t		n_role[5];	/* tokens per role */
} pars_kv_table;

/* Scan an INI or TOML document on the device (R1-R5; the hooks need no
 * pairing) and build its key-value table there: the reference's full-parse
 * benchmark for these formats (bench_parse_host.cu:250-333: scan_host, replay
 * into the hook) without the host replay.  Every array is device memory, valid
 * until the scanner's next call.  At most 2^32-1 bytes (TOML: and 2^31-1
 * tokens, as keys are int32 token indices).  Blocks until done;
 * st->p1 = R1-R5, p2 = the table's kernels (ms), when st is not NULL. */
int pars_hip_kv_parse(
	pars_scanner* sc, const uint8_t* d_in, size_t n, int format, pars_stream stream, pars_stage_times* st,
	pars_kv_table* out
);

/* Message for the last failing call on this thread ("" if none). */
const char* pars_hip_last_error(void);

/* Library version string. */
const char* pars_hip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PARS_HIP_H */
