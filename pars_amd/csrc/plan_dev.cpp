// plan_dev.cpp — lowering of a pars_plan into DevPlan tables (host side).
#include "plan_dev.hpp"

#include <string.h>

namespace pars_hip {

namespace {

// classify.hpp:39-55 restricted to the bits R3/R5/R6 consume.
void class_bits(const pars_plan& S, uint8_t st[256], uint8_t oc[256], uint8_t op[256]) {
	memset(st, 0, 256);
	memset(oc, 0, 256);
	memset(op, 0, 256);
	for (unsigned i = 0; i < S.n_punctuation; ++i) st[(uint8_t)S.punctuation[i]] = 1;
	for (unsigned i = 0; i < S.n_brackets; ++i) {
		const uint8_t o = (uint8_t)S.brackets[i].open, c = (uint8_t)S.brackets[i].close;
		st[o] = st[c] = 1;
		oc[o] = oc[c] = 1;
		op[o]		  = 1;  // nest.hpp:26-31: a byte that opens any bracket counts as open
	}
}

// One affine descriptor with a prefix escape, exactly the four-state automaton
// compile/plan.hpp:495-558 yields for `"` (\\ any | [^"\\])* `"`:
//   out  --delim--> in,  in --delim--> closed, in --escape--> esc, esc --any--> in,
//   closed behaves as out; masked exactly in `in` and `esc`.
bool is_escape_desc(const pars_lex_desc& d, uint8_t& dl, uint8_t& es, uint8_t st[4]) {
	if (d.n_states != 4 || d.n_classes != 3 || d.kind != PARS_LEX_AFFINE || !d.delim || !d.escape)
		return false;
	const uint8_t q = d.rep_of[(uint8_t)d.delim], e = d.rep_of[(uint8_t)d.escape];
	if (q == e)
		return false;
	uint8_t o = 0xFF;
	for (int b = 0; b < 256; ++b) {
		const uint8_t c = d.rep_of[b];
		if (b == (uint8_t)d.delim || b == (uint8_t)d.escape) {
			if (c != (b == (uint8_t)d.delim ? q : e))
				return false;
			continue;
		}
		if (o == 0xFF)
			o = c;
		if (c != o)
			return false;
	}
	if (o == 0xFF || o == q || o == e)
		return false;
	const uint8_t s_out = d.start, s_in = d.trans[s_out][q];
	const uint8_t s_cl = d.trans[s_in][q], s_esc = d.trans[s_in][e];
	auto row = [&](uint8_t s, uint8_t to, uint8_t tq, uint8_t te, uint8_t m) {
		return d.trans[s][o] == to && d.trans[s][q] == tq && d.trans[s][e] == te && d.masked[s][o] == m
			&& d.masked[s][q] == m && d.masked[s][e] == m;
	};
	if (!(row(s_out, s_out, s_in, s_out, 0) && row(s_in, s_in, s_cl, s_esc, 1)
		  && row(s_cl, s_out, s_in, s_out, 0) && row(s_esc, s_in, s_in, s_in, 1)))
		return false;
	if (s_out == s_in || s_out == s_cl || s_out == s_esc || s_in == s_cl || s_in == s_esc || s_cl == s_esc)
		return false;
	dl	  = (uint8_t)d.delim;
	es	  = (uint8_t)d.escape;
	st[0] = s_out;
	st[1] = s_in;
	st[2] = s_cl;
	st[3] = s_esc;
	return true;
}

// One descriptor that is quote parity, exactly the three-state automaton the
// compiler yields for a doubled delimiter (`"` ("" | [^"])* `"`, CSV) or an
// escape-less one (`'` [^']* `'`):
//   out --delim--> in --delim--> closed --delim--> in, closed --other--> out,
//   in --other--> in; masked in `in`, and in `closed` on a delimiter when doubled.
bool is_parity_desc(const pars_lex_desc& d, uint8_t& dl, uint8_t st[4], uint8_t& doubled) {
	if (d.n_states != 3 || d.n_classes != 2)
		return false;
	int			  nq = 0, qb = -1;
	const uint8_t c0 = d.rep_of[0];
	for (int b = 0; b < 256; ++b)
		if (d.rep_of[b] != c0) {
			++nq;
			qb = b;
		}
	if (nq != 1) {	// maybe byte 0 itself is the delimiter
		if (nq != 255)
			return false;
		qb = 0;
	}
	const uint8_t q = d.rep_of[qb], o = q == 0 ? 1 : 0;
	if (d.rep_of[qb] == d.rep_of[(qb + 1) & 255])
		return false;
	const uint8_t s_out = d.start, s_in = d.trans[s_out][q], s_cl = d.trans[s_in][q];
	if (s_out == s_in || s_out == s_cl || s_in == s_cl)
		return false;
	const bool ok = d.trans[s_out][o] == s_out && d.masked[s_out][o] == 0 && d.masked[s_out][q] == 0
				 && d.trans[s_in][o] == s_in && d.masked[s_in][o] == 1 && d.masked[s_in][q] == 1
				 && d.trans[s_cl][o] == s_out && d.trans[s_cl][q] == s_in && d.masked[s_cl][o] == 0;
	if (!ok)
		return false;
	dl		= (uint8_t)qb;
	st[0]	= s_out;
	st[1]	= s_in;
	st[2]	= s_cl;
	st[3]	= s_cl;
	doubled = d.masked[s_cl][q];
	return true;
}

// The event path's tables (DevPlan::evc, ev_map, ev_joint); false when the
// plan does not fit it (fewer than two descriptors, a background map that is
// not idempotent, more than 7 event ids or kMaxJoint joint states).
bool lower_events(const pars_plan& S, DevPlan& P) {
	const unsigned nd = S.n_descs;
	if (nd < 2)
		return false;
	uint8_t bg[kMaxDescs];
	for (unsigned d = 0; d < nd; ++d) {
		const pars_lex_desc& D = S.descs[d];
		int					 cnt[PARS_MAX_DFA_CLASSES] = {0};
		for (int c = 0; c < 256; ++c) cnt[D.rep_of[c]]++;
		bg[d] = 0;
		for (int k = 1; k < D.n_classes; ++k)
			if (cnt[k] > cnt[bg[d]])
				bg[d] = (uint8_t)k;
		for (int s = 0; s < D.n_states; ++s) {
			const uint8_t f = D.trans[s][bg[d]];
			if (D.trans[f][bg[d]] != f)
				return false;
		}
	}
	// event ids: distinct class tuples of the event bytes, numbered from 1 in
	// byte order
	uint8_t ev_id[256];
	uint8_t cls_of[kEvIds][kMaxDescs];	// per id, each descriptor's class
	int		n_ids = 0;
	for (unsigned d = 0; d < nd; ++d) cls_of[0][d] = bg[d];
	for (int c = 0; c < 256; ++c) {
		bool event = false;
		for (unsigned d = 0; d < nd; ++d) event = event || S.descs[d].rep_of[c] != bg[d];
		if (!event) {
			ev_id[c] = 0;
			continue;
		}
		int id = -1;
		for (int i = 1; i <= n_ids && id < 0; ++i) {
			bool same = true;
			for (unsigned d = 0; d < nd; ++d) same = same && cls_of[i][d] == S.descs[d].rep_of[c];
			if (same)
				id = i;
		}
		if (id < 0) {
			if (n_ids == kEvIds - 1)
				return false;
			id = ++n_ids;
			for (unsigned d = 0; d < nd; ++d) cls_of[id][d] = S.descs[d].rep_of[c];
		}
		ev_id[c] = (uint8_t)id;
	}
	uint32_t n_joint = 1;
	for (unsigned d = 0; d < nd; ++d) {
		P.ev_stride[d] = (uint16_t)n_joint;
		n_joint *= S.descs[d].n_states;
		if (n_joint > (uint32_t)kMaxJoint)
			return false;
	}
	for (int j = 0; j < 8; ++j)
		for (int c = 0; c < 256; ++c) {
			const uint32_t id = ev_id[c];
			P.evc[j][c][0]	  = P.cls[j][c] | ((id & 1u) << (24 + j));
			P.evc[j][c][1]	  = (((id >> 1) & 1u) << j) | (((id >> 2) & 1u) << (8 + j));
		}
	for (int i = 0; i < kEvIds; ++i)
		for (int g = 0; g < 2; ++g)
			for (unsigned d = 0; d < nd; ++d) {
				const pars_lex_desc& D = S.descs[d];
				uint8_t				 m[8];
				for (int s = 0; s < 8; ++s) {
					if (s >= D.n_states || i > n_ids) {
						m[s] = (uint8_t)s;
						continue;
					}
					const uint8_t s1 = g ? D.trans[s][bg[d]] : (uint8_t)s;
					m[s]			 = D.trans[s1][cls_of[i][d]];
				}
				P.ev_map[i][g][d][0] = (uint32_t)m[0] | ((uint32_t)m[1] << 8) | ((uint32_t)m[2] << 16) | ((uint32_t)m[3] << 24);
				P.ev_map[i][g][d][1] = (uint32_t)m[4] | ((uint32_t)m[5] << 8) | ((uint32_t)m[6] << 16) | ((uint32_t)m[7] << 24);
			}
	for (uint32_t J = 0; J < n_joint; ++J)
		for (int i = 0; i < kEvIds; ++i)
			for (int g = 0; g < 2; ++g) {
				uint32_t next = 0, first = 0, rest = 0, ev = 0;
				for (unsigned d = 0; d < nd && i <= n_ids; ++d) {
					const pars_lex_desc& D = S.descs[d];
					const uint32_t		 s = (J / P.ev_stride[d]) % D.n_states;
					uint32_t			 t = s;	 // state before the event byte
					if (g) {
						first |= D.masked[s][bg[d]];
						t = D.trans[s][bg[d]];
						rest |= D.masked[t][bg[d]];
					}
					ev |= D.masked[t][cls_of[i][d]];
					next += D.trans[t][cls_of[i][d]] * P.ev_stride[d];
				}
				P.ev_joint[J][i][g] = (uint16_t)(next | (first << 12) | (rest << 13) | (ev << 14));
			}
	P.ev_n_joint = (uint16_t)n_joint;
	P.ev_n_ids	 = (uint8_t)n_ids;
	return true;
}

}  // namespace

int lower_plan(const pars_plan& S, DevPlan& P) {
	memset(&P, 0, sizeof(P));
	if (S.n_descs > kMaxDescs || S.n_punctuation > PARS_MAX_PUNCT || S.n_brackets > PARS_MAX_BRACKETS)
		return -1;
	for (unsigned d = 0; d < S.n_descs; ++d) {
		const pars_lex_desc& D = S.descs[d];
		if (D.n_states < 1 || D.n_states > PARS_MAX_DFA_STATES || D.n_classes < 1
			|| D.n_classes > PARS_MAX_DFA_CLASSES || D.start >= D.n_states)
			return -2;
		for (int b = 0; b < 256; ++b)
			if (D.rep_of[b] >= D.n_classes)
				return -3;
		for (int s = 0; s < D.n_states; ++s)
			for (int c = 0; c < D.n_classes; ++c)
				if (D.trans[s][c] >= D.n_states || D.masked[s][c] > 1)
					return -4;
	}

	uint8_t st[256], oc[256], op[256];
	class_bits(S, st, oc, op);
	for (int j = 0; j < 8; ++j)
		for (int c = 0; c < 256; ++c)
			P.cls[j][c] = ((uint32_t)st[c] << j) | ((uint32_t)oc[c] << (8 + j)) | ((uint32_t)op[c] << (16 + j));

	P.n_descs = S.n_descs;
	for (unsigned d = 0; d < S.n_descs; ++d) {
		const pars_lex_desc& D = S.descs[d];
		P.start[d]			   = D.start;
		P.n_states[d]		   = D.n_states;
		P.start_packed |= (uint32_t)D.start << (3 * d);
		if (D.n_states > 4)
			P.wide = 1;
		for (int c = 0; c < 256; ++c) {
			const uint8_t cl = D.rep_of[c];
			uint8_t		  nx[8], mk[8];
			for (int s = 0; s < 8; ++s) {
				nx[s] = s < D.n_states ? D.trans[s][cl] : (uint8_t)s;
				mk[s] = s < D.n_states ? D.masked[s][cl] : 0;
			}
			auto pack4 = [](const uint8_t* v) {
				return (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
			};
			P.dt[d][c][0] = pack4(nx);
			P.dt[d][c][1] = pack4(nx + 4);
			P.dt[d][c][2] = pack4(mk);
			P.dt[d][c][3] = pack4(mk + 4);
		}
	}

	// nest.hpp:26-48: is_open_char / is_close_char / match_brackets
	for (unsigned i = 0; i < S.n_brackets; ++i) {
		const uint8_t o = (uint8_t)S.brackets[i].open, c = (uint8_t)S.brackets[i].close;
		if (!(P.br_flags[o] & 1))
			P.br_close_of[o] = c;  // the first spec opening with o decides its partner
		P.br_flags[o] |= 1;
		P.br_flags[c] |= 2;
	}

	P.fast_kind = kFastNone;
	if (S.n_descs == 1 && is_escape_desc(S.descs[0], P.fk_delim, P.fk_escape, P.fk_state))
		P.fast_kind = kFastEscape;
	else if (S.n_descs == 1 && is_parity_desc(S.descs[0], P.fk_delim, P.fk_state, P.fk_doubled))
		P.fast_kind = kFastParity;
	// one-bit ordinals of the bracket bytes: among the bytes that open, and
	// among the close-only bytes
	uint8_t ord[256] = {0};
	int		n_open = 0, n_close = 0;
	for (int c = 0; c < 256; ++c) {
		if (op[c]) {
			if (n_open < 2)
				P.br_chr[1][n_open] = (uint8_t)c;
			ord[c] = (uint8_t)(n_open++ & 1);
		} else if (oc[c]) {
			if (n_close < 2)
				P.br_chr[0][n_close] = (uint8_t)c;
			ord[c] = (uint8_t)(n_close++ & 1);
		}
	}
	P.br_bit_ok = n_open <= 2 && n_close <= 2;
	P.ev_ok		= lower_events(S, P) ? 1 : 0;
	for (int j = 0; j < 8; ++j)
		for (int c = 0; c < 256; ++c) {
			const uint32_t q = P.fast_kind != kFastNone && c == P.fk_delim;
			const uint32_t e = P.fast_kind == kFastEscape && c == P.fk_escape;
			P.cls8[j][c][0]	 = P.cls[j][c] | (q << (24 + j));
			P.cls8[j][c][1]	 = (e << j) | ((uint32_t)ord[c] << (8 + j));
		}
	return 0;
}

}  // namespace pars_hip
