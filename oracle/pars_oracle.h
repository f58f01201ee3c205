/* pars_oracle.h — CPU restatement of the reference's scalar backend.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this; it is the checker, never the
 * product path.  Functions and the reference code they restate:
 *
 *   oracle_classifier  include/pars/model/classify.hpp:39-55
 *   oracle_scan        include/pars/model/scan.hpp:27-73
 *   oracle_pair        include/pars/model/nest.hpp:45-92
 *   oracle_json_tape   bench/bench_parse.cu:41-348 through visit.hpp:35-48
 *                      (json_dom.c; the C twin of oracle/json_dom.py)
 *
 * Pinning (what checks each function against something other than itself):
 *   - golden vectors written by the reference compiled from its own sources
 *     (oracle/ref_driver.cpp, built by oracle/Makefile into oracle/_ref/),
 *     committed under tests/golden/ and compared in tests/test_oracle.py;
 *   - the reference's known-answer tests (tests/test_model.cpp:152-257)
 *     re-run against this file in tests/test_oracle.py;
 *   - brute force: every string up to length 8 over each descriptor's class
 *     alphabet against a textbook DFA walk, and a stack matcher for pairing on
 *     balanced input (tests/test_oracle.py).
 */
#ifndef PARS_ORACLE_H
#define PARS_ORACLE_H

#include "../include/pars_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* classify.hpp:39-55: byte -> class bits (PARS_CBIT_*). */
void oracle_classifier(const pars_plan* S, uint32_t bits[256]);

/* scan.hpp:27-73.  Writes at most `cap` entries into each output array (any
 * may be NULL) and returns the true counts in counts[0] (structural) and
 * counts[1] (brackets).  `masked_bits` (NULL or ceil(n/64) words) receives the
 * per-byte OR of the descriptors' masked bits; `struct_bits` (NULL or
 * ceil(n/64) words) the unmasked structural bytes.  Returns unclosed_mode. */
int oracle_scan(
	const pars_plan* S, const uint8_t* data, size_t n, uint32_t* positions, uint8_t* chars,
	uint32_t* bracket_pos, uint8_t* bracket_char, size_t cap, uint64_t counts[2],
	uint64_t* masked_bits, uint64_t* struct_bits
);

/* nest.hpp:45-92.  pair[nb] receives the partner bracket index or -1;
 * flags[0..2] = balanced, well_formed, max_depth.  Returns 0, or -1 when out
 * of memory. */
int oracle_pair(const pars_plan* S, const uint8_t* bchar, size_t nb, int32_t* pair, int32_t flags[3]);

/* Final per-descriptor state after scanning `data` from `in_states` (NULL =
 * start states): the carry a chunked scan threads between chunks. */
void oracle_final_states(
	const pars_plan* S, const uint8_t* data, size_t n, const uint8_t* in_states, uint8_t* out_states
);

/* 64-bit FNV-1a over a byte range, chained from `h` (start with
 * 0xcbf29ce484222325). */
uint64_t oracle_fnv1a(uint64_t h, const void* p, size_t n);

/* baseline.c: the restatement timed over documents buf[offs[d], +sizes[d]).
 * Per-document outputs (any array may be NULL): structural and bracket
 * counts, flags (bit0 balanced, bit1 well_formed, bit2 unclosed_mode) and max
 * depth.  Returns 0, or -1 when out of memory. */
typedef struct {
	double	 seconds;	  /* scan + pair only */
	uint64_t bytes, n_struct, n_brackets;
	uint64_t hash_pos;	  /* FNV-1a 64 chain over every document's positions, in order */
	uint64_t hash_pair;	  /* the same over the pair tables (document-local bracket indices) */
	int32_t	 balanced, well_formed, max_depth, unclosed; /* of a single document */
} oracle_run;

int oracle_time_docs(const pars_plan* S, const uint8_t* buf, const uint64_t* offs, const uint64_t* sizes, size_t nd,
					 uint32_t* doc_struct, uint32_t* doc_brackets, uint8_t* doc_flags, int32_t* doc_depth,
					 oracle_run* out);

/* The JSON hook's results (json_dom.c). */
typedef struct {
	uint64_t n_nodes;	 /* nodes pushed (the arrays hold the first cap_nodes) */
	uint64_t dec_bytes;	 /* decoded string bytes (dec holds the first cap_dec) */
	uint64_t hash;		 /* the hook's FNV-1a 64 */
	int64_t	 fail_event; /* 2i: the run before structural i; 2i+1: structural i; -1: none */
	int32_t	 ok;
	int32_t	 pad;
} oracle_tape_info;

/* bench_parse.cu:41-348: replay the structural index into the JSON hook.
 * Returns 0, or -1 when its stack cannot be allocated. */
int oracle_json_tape(const uint8_t* data, size_t n, const uint32_t* positions, size_t n_struct, uint8_t* kind,
					 uint32_t* begin, uint32_t* end, size_t cap_nodes, uint8_t* dec, size_t cap_dec,
					 oracle_tape_info* info);

#ifdef __cplusplus
}
#endif

#endif
