/* json_dom.c — the reference's JSON DOM hook (bench/bench_parse.cu:41-333),
 * driven by model::replay (include/pars/model/visit.hpp:35-48) over a
 * structural index, restated in C: the same restatement as oracle/json_dom.py
 * (which the tests pin against hand-worked cases and Python's json parser),
 * fast enough to time as a CPU baseline and to check large documents.
 *
 * TEST INFRASTRUCTURE ONLY (tests/ and bench.py's cpu_baseline leg).
 *
 * Events are numbered as the device tape reports them: the run before
 * structural i is 2i, structural i is 2i+1, the run after the last is 2n. */
#include "pars_oracle.h"

#include <stdlib.h>
#include <string.h>

enum { J_OBJ, J_ARR, J_STR, J_NUM, J_TRUE, J_FALSE, J_NULL };

typedef struct {
	const uint8_t* b;
	uint8_t*	   kind;
	uint32_t *	   beg, *end;
	size_t		   cap_nodes, n_nodes;
	uint8_t*	   dec;
	size_t		   cap_dec, n_dec;
	uint64_t	   hash;
	/* the parser stack: open byte, phase; and the tape index of each open node */
	uint8_t*	   st_open;
	uint8_t*	   st_phase;
	size_t*		   st_node;
	size_t		   depth;
	int			   ok, top_done;
	int64_t		   fail_event, ev;
} hook;

static int is_ws(uint8_t c) {
	return c == ' ' || c == '\t' || c == '\n' || c == '\r';
}
static int is_digit(uint8_t c) {
	return c >= '0' && c <= '9';
}

static void mix(hook* h, uint8_t x) {
	h->hash ^= x;
	h->hash *= 1099511628211ull;
}

static void put(hook* h, uint8_t x) { /* a decoded string byte */
	if (h->n_dec < h->cap_dec)
		h->dec[h->n_dec] = x;
	h->n_dec++;
	mix(h, x);
}

static void push_node(hook* h, int k, size_t b) {
	if (h->n_nodes < h->cap_nodes) {
		h->kind[h->n_nodes] = (uint8_t)k;
		h->beg[h->n_nodes]	= (uint32_t)b;
		h->end[h->n_nodes]	= 0;
	}
	h->n_nodes++;
}

static void set_end(hook* h, size_t node, size_t e) {
	if (node < h->cap_nodes)
		h->end[node] = (uint32_t)e;
}

static void reject(hook* h) {
	if (h->ok) {
		h->ok		  = 0;
		h->fail_event = h->ev;
	}
}

static int in_value(const hook* h) {
	if (!h->depth)
		return !h->top_done;
	return h->st_open[h->depth - 1] == '{' ? h->st_phase[h->depth - 1] == 2 : h->st_phase[h->depth - 1] == 0;
}

static void value_done(hook* h) {
	if (!h->depth)
		h->top_done = 1;
	else
		h->st_phase[h->depth - 1] = h->st_open[h->depth - 1] == '{' ? 3 : 1;
}

#define NPOS ((size_t)-1)

static size_t parse_string(hook* h, size_t i, size_t e) { /* bench_parse.cu:231-284 */
	const uint8_t* b = h->b;
	push_node(h, J_STR, i);
	mix(h, 'S');
	size_t j = i + 1;
	while (j < e) {
		const uint8_t c = b[j++];
		if (c == '"') {
			set_end(h, h->n_nodes - 1, j);
			mix(h, 0);
			return j;
		}
		if (c == '\\') {
			if (j >= e)
				return NPOS;
			const uint8_t esc = b[j++];
			uint8_t		  d;
			switch (esc) {
			case '"': d = '"'; break;
			case '\\': d = '\\'; break;
			case '/': d = '/'; break;
			case 'b': d = '\b'; break;
			case 'f': d = '\f'; break;
			case 'n': d = '\n'; break;
			case 'r': d = '\r'; break;
			case 't': d = '\t'; break;
			case 'u': {
				if (j + 4 > e)
					return NPOS;
				unsigned v = 0;
				for (int k = 0; k < 4; ++k) {
					const uint8_t x = b[j++];
					v <<= 4;
					if (x >= '0' && x <= '9')
						v |= (unsigned)(x - '0');
					else if (x >= 'a' && x <= 'f')
						v |= (unsigned)(x - 'a' + 10);
					else if (x >= 'A' && x <= 'F')
						v |= (unsigned)(x - 'A' + 10);
					else
						return NPOS;
				}
				if (v < 0x80) {
					put(h, (uint8_t)v);
				} else if (v < 0x800) {
					put(h, (uint8_t)(0xC0 | (v >> 6)));
					put(h, (uint8_t)(0x80 | (v & 0x3F)));
				} else {
					put(h, (uint8_t)(0xE0 | (v >> 12)));
					put(h, (uint8_t)(0x80 | ((v >> 6) & 0x3F)));
					put(h, (uint8_t)(0x80 | (v & 0x3F)));
				}
				continue;
			}
			default: return NPOS;
			}
			put(h, d);
		} else if (c < 0x20) {
			return NPOS;
		} else {
			put(h, c);
		}
	}
	return NPOS;
}

static size_t parse_number(hook* h, size_t i, size_t e) { /* bench_parse.cu:286-306 */
	const uint8_t* b = h->b;
	size_t		   j = i;
	if (b[j] == '-')
		++j;
	if (j >= e)
		return NPOS;
	if (b[j] == '0')
		++j;
	else if (b[j] >= '1' && b[j] <= '9')
		while (j < e && is_digit(b[j])) ++j;
	else
		return NPOS;
	if (j < e && b[j] == '.') {
		++j;
		if (j >= e || !is_digit(b[j]))
			return NPOS;
		while (j < e && is_digit(b[j])) ++j;
	}
	if (j < e && (b[j] == 'e' || b[j] == 'E')) {
		++j;
		if (j < e && (b[j] == '+' || b[j] == '-'))
			++j;
		if (j >= e || !is_digit(b[j]))
			return NPOS;
		while (j < e && is_digit(b[j])) ++j;
	}
	push_node(h, J_NUM, i);
	set_end(h, h->n_nodes - 1, j);
	mix(h, 'N');
	for (size_t k = i; k < j; ++k) mix(h, b[k]);
	return j;
}

static size_t parse_lit(hook* h, size_t i, size_t e) { /* bench_parse.cu:308-332 */
	static const struct {
		const char* w;
		size_t		len;
		int			kind;
		uint8_t		tag;
	} L[3] = {{"true", 4, J_TRUE, 'T'}, {"false", 5, J_FALSE, 'F'}, {"null", 4, J_NULL, '0'}};
	for (int k = 0; k < 3; ++k)
		if (i + L[k].len <= e && memcmp(h->b + i, L[k].w, L[k].len) == 0) {
			push_node(h, L[k].kind, i);
			set_end(h, h->n_nodes - 1, i + L[k].len);
			mix(h, L[k].tag);
			return i + L[k].len;
		}
	return NPOS;
}

static size_t parse_value(hook* h, size_t i, size_t e) {
	const uint8_t c = h->b[i];
	if (c == '"')
		return parse_string(h, i, e);
	if (c == '-' || is_digit(c))
		return parse_number(h, i, e);
	if (c == 't' || c == 'f' || c == 'n')
		return parse_lit(h, i, e);
	return NPOS;
}

static void scalar(hook* h, size_t b, size_t e) { /* bench_parse.cu:86-130 */
	if (!h->ok)
		return;
	size_t i = b;
	while (i < e && is_ws(h->b[i])) ++i;
	if (i >= e)
		return;
	const int as_key = h->depth && h->st_open[h->depth - 1] == '{' && h->st_phase[h->depth - 1] == 0;
	size_t	  j;
	if (as_key) {
		if (h->b[i] != '"') {
			reject(h);
			return;
		}
		j = parse_string(h, i, e);
	} else {
		if (!in_value(h)) {
			reject(h);
			return;
		}
		j = parse_value(h, i, e);
	}
	if (j == NPOS) {
		reject(h);
		return;
	}
	while (j < e && is_ws(h->b[j])) ++j;
	if (j != e) {
		reject(h);
		return;
	}
	if (as_key)
		h->st_phase[h->depth - 1] = 1;
	else
		value_done(h);
}

static void symbol(hook* h, size_t p, uint8_t c) { /* bench_parse.cu:132-194 */
	if (!h->ok)
		return;
	size_t d = h->depth;
	switch (c) {
	case '{':
	case '[':
		if (!in_value(h)) {
			reject(h);
			return;
		}
		push_node(h, c == '{' ? J_OBJ : J_ARR, p);
		h->st_node[d]  = h->n_nodes - 1;
		h->st_open[d]  = c;
		h->st_phase[d] = 0;
		h->depth	   = d + 1;
		break;
	case '}':
	case ']': {
		const uint8_t o = c == '}' ? '{' : '[';
		if (!d || h->st_open[d - 1] != o) {
			reject(h);
			return;
		}
		const uint8_t ph = h->st_phase[d - 1];
		if (!(ph == 0 || ph == (c == '}' ? 3 : 1))) {
			reject(h);
			return;
		}
		set_end(h, h->st_node[d - 1], p);
		h->depth = d - 1;
		value_done(h);
		break;
	}
	case ':':
		if (!d || h->st_open[d - 1] != '{' || h->st_phase[d - 1] != 1) {
			reject(h);
			return;
		}
		h->st_phase[d - 1] = 2;
		break;
	case ',':
		if (!d || h->st_phase[d - 1] != (h->st_open[d - 1] == '{' ? 3 : 1)) {
			reject(h);
			return;
		}
		h->st_phase[d - 1] = 0;
		break;
	default: reject(h); break;
	}
}

int oracle_json_tape(const uint8_t* data, size_t n, const uint32_t* positions, size_t n_struct, uint8_t* kind,
					 uint32_t* begin, uint32_t* end, size_t cap_nodes, uint8_t* dec, size_t cap_dec,
					 oracle_tape_info* info) {
	hook h;
	memset(&h, 0, sizeof h);
	h.b			= data;
	h.kind		= kind;
	h.beg		= begin;
	h.end		= end;
	h.cap_nodes = kind && begin && end ? cap_nodes : 0;
	h.dec		= dec;
	h.cap_dec	= dec ? cap_dec : 0;
	h.hash		= 1469598103934665603ull;
	h.ok		= 1;
	h.fail_event = -1;
	const size_t sd = n_struct + 1;
	h.st_open		= (uint8_t*)malloc(sd);
	h.st_phase		= (uint8_t*)malloc(sd);
	h.st_node		= (size_t*)malloc(sd * sizeof(size_t));
	if (!h.st_open || !h.st_phase || !h.st_node) {
		free(h.st_open);
		free(h.st_phase);
		free(h.st_node);
		return -1;
	}
	size_t prev = 0; /* visit.hpp:35-48 */
	for (size_t i = 0; i < n_struct && h.ok; ++i) {
		const size_t p = positions[i];
		if (p > prev) {
			h.ev = 2 * (int64_t)i;
			scalar(&h, prev, p);
		}
		h.ev = 2 * (int64_t)i + 1;
		symbol(&h, p, data[p]);
		prev = p + 1;
	}
	if (h.ok && n > prev) {
		h.ev = 2 * (int64_t)n_struct;
		scalar(&h, prev, n);
	}
	if (h.depth || !h.top_done) /* bench_parse.cu:196-201 */
		h.ok = 0;
	info->n_nodes	 = h.n_nodes;
	info->dec_bytes	 = h.n_dec;
	info->hash		 = h.hash;
	info->fail_event = h.fail_event;
	info->ok		 = h.ok;
	free(h.st_open);
	free(h.st_phase);
	free(h.st_node);
	return 0;
}
