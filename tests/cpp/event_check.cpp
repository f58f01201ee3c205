// event_check.cpp — host check of the event path's tables and stepping
// (DevPlan::ev_*, built by pars_amd/csrc/plan_dev.cpp; stepped on the device
// by EventEval in pars_amd/csrc/evaluators.hpp).  For every plan file given,
// random 64-byte lanes (event bytes over-represented, random valid lengths)
// are stepped two ways:
//   - byte by byte through each descriptor's DFA (the semantics of
//     model/scan.hpp: masked[state before][class], then trans);
//   - the way EventEval does: event ids from the position-shifted class
//     table, each descriptor's map composed at the events only (ev_map, the
//     background run before an event folded in, a run to the end as one
//     step), and the masked bits from the joint-state table (ev_joint).
// The maps must agree for every entry state, the masks for random entry
// states.  Besides the plan files, random plans whose background runs'
// first and later bytes mask differently.  Prints "ok <plans checked>" or the
// first difference.
#include "../../pars_amd/csrc/plan_dev.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

using namespace pars_hip;

static bool read_plan(const char* path, pars_plan& p) {
	FILE* f = std::fopen(path, "rb");
	if (!f)
		return false;
	const size_t n = std::fread(&p, 1, sizeof p, f);
	std::fclose(f);
	return n == sizeof p;
}

static uint32_t event_id(const DevPlan& P, int j, uint8_t b) {
	const int k = j & 7;
	return ((P.evc[k][b][0] >> (24 + k)) & 1u) | (((P.evc[k][b][1] >> k) & 1u) << 1)
		 | (((P.evc[k][b][1] >> (8 + k)) & 1u) << 2);
}

static uint8_t map_byte(const uint32_t (&m)[2], uint32_t s) {
	return (uint8_t)(m[s >> 2] >> (8 * (s & 3)));
}

static uint64_t span(int a, int b) {
	const uint64_t hi = b >= 64 ? ~0ull : ((1ull << b) - 1ull);
	return hi & ~((1ull << a) - 1ull);
}

static uint64_t run_bits(uint32_t e, int q, int end) {
	uint64_t m = 0;
	if (end > q + 1) {
		if (e & (1u << 12))
			m |= 1ull << (q + 1);
		if ((e & (1u << 13)) && end > q + 2)
			m |= span(q + 2, end);
	}
	return m;
}

// A random plan of 2..5 descriptors in which the event path applies: each
// descriptor's class 0 (most bytes) maps states idempotently (every state to
// one of a random set of fixed points), the other classes and the masked bits
// are random, so a background run's first byte and its later bytes can differ.
static void random_plan(std::mt19937_64& rng, pars_plan& S) {
	std::memset(&S, 0, sizeof S);
	S.n_descs	   = (uint8_t)(2 + rng() % 4);
	uint32_t joint = 1;
	uint8_t	 evb[7];  // the plan's event bytes
	for (int k = 0; k < 7; ++k) evb[k] = (uint8_t)(k == 0 ? rng() % 256 : (evb[k - 1] + 1 + rng() % 30) % 256);
	for (int d = 0; d < S.n_descs; ++d) {
		pars_lex_desc& D = S.descs[d];
		const uint32_t room = (uint32_t)kMaxJoint / joint;	// joint stays within kMaxJoint
		if (room < 2) {
			S.n_descs = (uint8_t)d;
			break;
		}
		const int ns = 2 + (int)(rng() % (std::min<uint32_t>(room, 8) - 1));
		joint *= ns;
		D.n_states	= (uint8_t)ns;
		D.n_classes = (uint8_t)(2 + rng() % 3);
		D.kind		= PARS_LEX_TABLE;
		// at most 7 distinct class tuples among the event bytes: only evb's bytes
		// leave class 0, and every class has a byte
		for (int c = 0; c < 256; ++c) D.rep_of[c] = 0;
		for (int k = 0; k < 7; ++k) D.rep_of[evb[k]] = (uint8_t)(rng() % D.n_classes);
		D.rep_of[evb[rng() % 7]] = (uint8_t)(D.n_classes - 1);
		uint8_t fixed[8];
		int		nf = 0;
		for (int s = 0; s < ns; ++s)
			if (nf == 0 || rng() % 2)
				fixed[nf++] = (uint8_t)s;
		for (int s = 0; s < ns; ++s) {
			bool is_fixed = false;
			for (int k = 0; k < nf; ++k) is_fixed = is_fixed || fixed[k] == s;
			D.trans[s][0] = is_fixed ? (uint8_t)s : fixed[rng() % nf];
			for (int c = 1; c < D.n_classes; ++c) D.trans[s][c] = (uint8_t)(rng() % ns);
			for (int c = 0; c < D.n_classes; ++c) D.masked[s][c] = (uint8_t)(rng() % 2);
		}
	}
}

int main(int argc, char** argv) {
	int				checked = 0;
	std::mt19937_64 prng(99);
	const int		n_random = 300;
	for (int a = 1; a < argc + n_random; ++a) {
		pars_plan	S;
		const char* name = a < argc ? argv[a] : "a random plan";
		if (a < argc) {
			if (!read_plan(argv[a], S)) {
				std::printf("cannot read %s\n", argv[a]);
				return 1;
			}
		} else {
			random_plan(prng, S);
		}
		static DevPlan P;
		if (lower_plan(S, P) != 0) {
			std::printf("lower_plan failed: %s\n", name);
			return 1;
		}
		if (!P.ev_ok) {
			if (a >= argc) {
				std::printf("the event path refused %s\n", name);
				return 1;
			}
			continue;
		}
		const int trials = a < argc ? 200000 : 3000;
		const int nd = S.n_descs;
		std::vector<uint8_t> ev_bytes;
		for (int c = 0; c < 256; ++c)
			if (event_id(P, 0, (uint8_t)c))
				ev_bytes.push_back((uint8_t)c);
		std::mt19937_64 rng(1234 + a);
		for (int trial = 0; trial < trials; ++trial) {
			uint8_t		   x[64];
			const double   pe = (trial % 4) * 0.25;	 // event byte density
			for (int j = 0; j < 64; ++j)
				x[j] = (std::uniform_real_distribution<double>(0, 1)(rng) < pe && !ev_bytes.empty())
						 ? ev_bytes[rng() % ev_bytes.size()]
						 : (uint8_t)(rng() & 255);
			const int valid = (trial % 3) ? 64 : (int)(rng() % 65);
			// event stepping: maps
			uint8_t T[kMaxDescs][8];
			for (int d = 0; d < nd; ++d)
				for (int s = 0; s < 8; ++s) T[d][s] = (uint8_t)s;
			int q = -1;
			for (int p = 0; p < valid; ++p) {
				const uint32_t id = event_id(P, p, x[p]);
				if (!id)
					continue;
				const int g = p > q + 1;
				for (int d = 0; d < nd; ++d)
					for (int s = 0; s < 8; ++s) T[d][s] = map_byte(P.ev_map[id][g][d], T[d][s]);
				q = p;
			}
			if (q < valid - 1)
				for (int d = 0; d < nd; ++d)
					for (int s = 0; s < 8; ++s) T[d][s] = map_byte(P.ev_map[0][1][d], T[d][s]);
			// byte stepping for every entry state of every descriptor
			uint64_t mk_all[kMaxDescs][8];
			for (int d = 0; d < nd; ++d) {
				const pars_lex_desc& D = S.descs[d];
				for (int s0 = 0; s0 < D.n_states; ++s0) {
					uint32_t s = s0;
					uint64_t m = 0;
					for (int j = 0; j < valid; ++j) {
						const uint8_t c = D.rep_of[x[j]];
						m |= (uint64_t)D.masked[s][c] << j;
						s = D.trans[s][c];
					}
					mk_all[d][s0] = m;
					if (T[d][s0] != s) {
						std::printf("%s: map differs (descriptor %d, entry %d, valid %d, trial %d)\n", name, d, s0,
									valid, trial);
						return 1;
					}
				}
			}
			// masks for a random entry state tuple
			uint32_t s0[kMaxDescs], J = 0;
			uint64_t want = 0;
			for (int d = 0; d < nd; ++d) {
				s0[d] = (uint32_t)(rng() % S.descs[d].n_states);
				J += s0[d] * P.ev_stride[d];
				want |= mk_all[d][s0[d]];
			}
			uint64_t got = 0;
			q			 = -1;
			for (int p = 0; p < valid; ++p) {
				const uint32_t id = event_id(P, p, x[p]);
				if (!id)
					continue;
				const uint32_t e = P.ev_joint[J][id][p > q + 1 ? 1 : 0];
				got |= run_bits(e, q, p) | ((uint64_t)((e >> 14) & 1u) << p);
				J = e & 0xFFFu;
				q = p;
			}
			if (q < valid - 1)
				got |= run_bits(P.ev_joint[J][0][1], q, valid);
			if (got != want) {
				std::printf("%s: masked bits differ (valid %d, trial %d): %016llx vs %016llx\n", name, valid, trial,
							(unsigned long long)got, (unsigned long long)want);
				return 1;
			}
		}
		++checked;
	}
	std::printf("ok %d\n", checked);
	return 0;
}
