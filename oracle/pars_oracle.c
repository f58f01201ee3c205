/* pars_oracle.c — CPU restatement of the reference's scalar backend.
 *
 * TEST INFRASTRUCTURE ONLY (see pars_oracle.h): the checker for the HIP path
 * and the "port" CPU baseline of bench.py; never linked into the product.
 *
 * Same algorithm as the reference, one byte at a time, single-threaded:
 *   scan: include/pars/model/scan.hpp:33-72
 *   pair: include/pars/model/nest.hpp:45-92 (std::stable_sort is restated as a
 *         bottom-up merge sort, which is stable and therefore yields the same
 *         permutation for the same keys).
 */
#include "pars_oracle.h"

#include <stdlib.h>
#include <string.h>

void oracle_classifier(const pars_plan* S, uint32_t bits[256]) {
	memset(bits, 0, 256 * sizeof(uint32_t));
	/* classify.hpp:41-42 */
	for (unsigned i = 0; i < S->n_punctuation; ++i)
		bits[(unsigned char)S->punctuation[i]] |= PARS_CBIT_STRUCTURAL;
	/* classify.hpp:43-46 */
	for (unsigned i = 0; i < S->n_brackets; ++i) {
		bits[(unsigned char)S->brackets[i].open] |= PARS_CBIT_STRUCTURAL | PARS_CBIT_OPEN;
		bits[(unsigned char)S->brackets[i].close] |= PARS_CBIT_STRUCTURAL | PARS_CBIT_CLOSE;
	}
	/* classify.hpp:47-51 */
	for (unsigned i = 0; i < S->n_modes; ++i) {
		bits[(unsigned char)S->modes[i].delim] |= PARS_CBIT_MODE_DELIM(i);
		if (S->modes[i].escape && S->modes[i].escape != S->modes[i].delim)
			bits[(unsigned char)S->modes[i].escape] |= PARS_CBIT_MODE_ESCAPE(i);
	}
	/* classify.hpp:52-53 */
	for (unsigned i = 0; i < S->n_comments; ++i)
		bits[(unsigned char)S->comments[i]] |= PARS_CBIT_COMMENT(i);
}

int oracle_scan(
	const pars_plan* S, const uint8_t* data, size_t n, uint32_t* positions, uint8_t* chars,
	uint32_t* bracket_pos, uint8_t* bracket_char, size_t cap, uint64_t counts[2],
	uint64_t* masked_bits, uint64_t* struct_bits
) {
	uint32_t cls[256];
	oracle_classifier(S, cls);
	const unsigned nd = S->n_descs;
	uint8_t		   st[PARS_MAX_MODES];
	for (unsigned d = 0; d < nd; ++d) st[d] = S->descs[d].start; /* scan.hpp:38 */
	if (masked_bits)
		memset(masked_bits, 0, ((n + 63) / 64) * sizeof(uint64_t));
	if (struct_bits)
		memset(struct_bits, 0, ((n + 63) / 64) * sizeof(uint64_t));

	size_t k = 0, kb = 0;
	for (size_t i = 0; i < n; ++i) {
		const unsigned char c = data[i];
		int					masked = 0;
		/* scan.hpp:46-52: every descriptor steps on every byte; masks are ORed */
		for (unsigned d = 0; d < nd; ++d) {
			const pars_lex_desc* desc = &S->descs[d];
			const uint8_t		 cl	  = desc->rep_of[c];
			if (desc->masked[st[d]][cl])
				masked = 1;
			st[d] = desc->trans[st[d]][cl];
		}
		if (masked) {
			if (masked_bits)
				masked_bits[i >> 6] |= 1ull << (i & 63);
			continue; /* scan.hpp:53-54 */
		}
		if (cls[c] & PARS_CBIT_STRUCTURAL) { /* scan.hpp:56-63 */
			if (struct_bits)
				struct_bits[i >> 6] |= 1ull << (i & 63);
			if (k < cap) {
				if (positions)
					positions[k] = (uint32_t)i;
				if (chars)
					chars[k] = c;
			}
			++k;
			if (cls[c] & (PARS_CBIT_OPEN | PARS_CBIT_CLOSE)) {
				if (kb < cap) {
					if (bracket_pos)
						bracket_pos[kb] = (uint32_t)i;
					if (bracket_char)
						bracket_char[kb] = c;
				}
				++kb;
			}
		}
	}
	counts[0] = k;
	counts[1] = kb;
	/* scan.hpp:66-70 */
	for (unsigned d = 0; d < nd; ++d)
		if (st[d] != S->descs[d].start)
			return 1;
	return 0;
}

void oracle_final_states(
	const pars_plan* S, const uint8_t* data, size_t n, const uint8_t* in_states, uint8_t* out_states
) {
	const unsigned nd = S->n_descs;
	uint8_t		   st[PARS_MAX_MODES];
	for (unsigned d = 0; d < nd; ++d) st[d] = in_states ? in_states[d] : S->descs[d].start;
	for (size_t i = 0; i < n; ++i)
		for (unsigned d = 0; d < nd; ++d)
			st[d] = S->descs[d].trans[st[d]][S->descs[d].rep_of[data[i]]];
	for (unsigned d = 0; d < nd; ++d) out_states[d] = st[d];
}

/* nest.hpp:26-40 */
static int is_open_char(char c, const pars_plan* S) {
	for (unsigned i = 0; i < S->n_brackets; ++i)
		if (S->brackets[i].open == c)
			return 1;
	return 0;
}

static int is_close_char(char c, const pars_plan* S) {
	for (unsigned i = 0; i < S->n_brackets; ++i)
		if (S->brackets[i].close == c)
			return 1;
	return 0;
}

static int match_brackets(char o, char c, const pars_plan* S) {
	for (unsigned i = 0; i < S->n_brackets; ++i)
		if (S->brackets[i].open == o)
			return S->brackets[i].close == c;
	return 0;
}

int oracle_pair(const pars_plan* S, const uint8_t* bchar, size_t n, int32_t* pair, int32_t flags[3]) {
	int balanced = 1, well_formed = 1, max_depth = 0;
	for (size_t i = 0; i < n; ++i) pair[i] = -1; /* nest.hpp:49 */
	if (n == 0) {
		flags[0] = 1;
		flags[1] = 1;
		flags[2] = 0;
		return 0;
	}
	int32_t* key   = (int32_t*)malloc(n * sizeof(int32_t));
	int32_t* order = (int32_t*)malloc(n * sizeof(int32_t));
	int32_t* tmp   = (int32_t*)malloc(n * sizeof(int32_t));
	uint8_t* open  = (uint8_t*)malloc(n);
	if (!key || !order || !tmp || !open) {
		free(key);
		free(order);
		free(tmp);
		free(open);
		return -1;
	}
	/* nest.hpp:53-66: inclusive prefix of +-1, max_depth from 0, min from 0 */
	int run = 0, min_bal = 0;
	for (size_t i = 0; i < n; ++i) {
		open[i] = (uint8_t)is_open_char((char)bchar[i], S);
		run += open[i] ? 1 : -1;
		if (run > max_depth)
			max_depth = run;
		if (run < min_bal)
			min_bal = run;
		key[i]	 = open[i] ? run - 1 : run; /* nest.hpp:71-75 */
		order[i] = (int32_t)i;
	}
	balanced = (min_bal >= 0) && (run == 0);

	/* nest.hpp:70-76: stable sort of indices by key (bottom-up merge sort) */
	int32_t* src = order;
	int32_t* dst = tmp;
	for (size_t w = 1; w < n; w <<= 1) {
		for (size_t lo = 0; lo < n; lo += 2 * w) {
			size_t mid = lo + w < n ? lo + w : n;
			size_t hi  = lo + 2 * w < n ? lo + 2 * w : n;
			size_t a = lo, b = mid, o = lo;
			while (a < mid && b < hi) dst[o++] = (key[src[b]] < key[src[a]]) ? src[b++] : src[a++];
			while (a < mid) dst[o++] = src[a++];
			while (b < hi) dst[o++] = src[b++];
		}
		int32_t* t = src;
		src		   = dst;
		dst		   = t;
	}

	/* nest.hpp:78-88: adjacent pairing in sorted order */
	for (size_t k = 0; k + 1 < n; k += 2) {
		const int32_t a = src[k], b = src[k + 1];
		if (open[a] && is_close_char((char)bchar[b], S) && match_brackets((char)bchar[a], (char)bchar[b], S)) {
			pair[a] = b;
			pair[b] = a;
		} else {
			well_formed = 0;
		}
	}
	if (n % 2 != 0)
		well_formed = 0;
	well_formed = well_formed && balanced; /* nest.hpp:90 */
	flags[0]	= balanced;
	flags[1]	= well_formed;
	flags[2]	= max_depth;
	free(key);
	free(order);
	free(tmp);
	free(open);
	return 0;
}

uint64_t oracle_fnv1a(uint64_t h, const void* p, size_t n) {
	const uint8_t* b = (const uint8_t*)p;
	for (size_t i = 0; i < n; ++i) {
		h ^= b[i];
		h *= 0x100000001b3ull;
	}
	return h;
}
