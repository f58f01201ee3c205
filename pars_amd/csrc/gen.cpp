// gen.cpp — seeded synthetic documents for the benchmark workloads and tests.
//
// Deterministic: SplitMix64 (Steele, Lea & Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014), integer arithmetic for every
// choice, so the same (seed, size) gives the same bytes on every machine with
// this image.  Every generator writes exactly n bytes, padding with spaces
// (never structural in any grammar here, so padding never changes outputs).
//
// Workloads (sizes are the caller's):
//   json   top-level array of objects with 4-12 keys [a-z]{3,12}; values uniform
//          over strings 0-40 chars (5% with \" or \\, 2% with \uXXXX, some
//          UTF-8), int64, %.6g doubles, true/false/null, nested arrays/objects;
//          single ' ' / '\n' whitespace; optional string-byte-fraction target
//          and maximum depth (deep chains reach it).
//   csv    16 columns, RFC-4180 quoting, 10% quoted fields holding ',', CRLF or
//          doubled quotes, CRLF record endings.
//   clike  six-mode C-like source: identifiers, numbers, 40 operators and
//          punctuators, strings (5% escapes), char literals, // lines (10%),
//          block comments of 1-20 lines (1% with '*' runs), # lines, balanced
//          () [] {} nesting to a maximum depth.
//   alpha  uniform bytes over a caller alphabet (adversarial inputs).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <thread>
#include <vector>

namespace {

struct Rng {
	uint64_t s;

	uint64_t next() {
		uint64_t z = (s += 0x9e3779b97f4a7c15ull);
		z		   = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
		z		   = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
		return z ^ (z >> 31);
	}

	uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }

	uint32_t range(uint32_t lo, uint32_t hi) { return lo + below(hi - lo + 1); }  // inclusive

	bool pct(uint32_t p) { return below(100) < p; }	 // p percent

	bool permille(uint32_t p) { return below(1000) < p; }
};

uint64_t mix(uint64_t seed, uint64_t index) {
	Rng r {seed ^ (index * 0xd1b54a32d192ed03ull)};
	return r.next();
}

// ---------------------------------------------------------------------------
// JSON
// ---------------------------------------------------------------------------
struct JsonGen {
	Rng			 r;
	std::string& o;
	int			 max_depth;
	int			 str_target_pm;	 // string-byte fraction target in permille, <0 = none
	uint64_t	 str_bytes = 0;
	uint64_t	 base	   = 0;	 // bytes emitted before the current object buffer

	JsonGen(uint64_t seed, std::string& out, int md, int st) : r {seed}, o(out), max_depth(md), str_target_pm(st) {}

	void ws() {
		if (r.pct(20))
			o.push_back(r.below(3) ? ' ' : '\n');
	}

	void lower_run(int n) {
		for (int i = 0; i < n; ++i) o.push_back((char)('a' + r.below(26)));
	}

	void key() {
		const size_t s = o.size();
		o.push_back('"');
		lower_run((int)r.range(3, 12));
		o.push_back('"');
		str_bytes += o.size() - s;
	}

	void string_value() {
		const size_t s	 = o.size();
		const int	 len = (int)r.range(0, 40);
		const bool	 esc = r.pct(5), uni = r.pct(2);
		o.push_back('"');
		static const char kPunct[] = ",:{}[]-_.!?/ ";
		for (int i = 0; i < len; ++i) {
			const uint32_t k = r.below(100);
			if (k < 60)
				o.push_back((char)('a' + r.below(26)));
			else if (k < 70)
				o.push_back((char)('A' + r.below(26)));
			else if (k < 78)
				o.push_back((char)('0' + r.below(10)));
			else if (k < 92)
				o.push_back(kPunct[r.below(sizeof(kPunct) - 1)]);
			else if (k < 96) {	// two-byte UTF-8 (U+00C0..U+017F)
				const uint32_t cp = 0xC0 + r.below(0xC0);
				o.push_back((char)(0xC0 | (cp >> 6)));
				o.push_back((char)(0x80 | (cp & 0x3F)));
			} else if (k < 98) {  // three-byte UTF-8 (U+4E00..)
				const uint32_t cp = 0x4E00 + r.below(0x5000);
				o.push_back((char)(0xE0 | (cp >> 12)));
				o.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
				o.push_back((char)(0x80 | (cp & 0x3F)));
			} else
				o.push_back(' ');
			if (esc && r.pct(12)) {
				o.push_back('\\');
				o.push_back(r.below(2) ? '"' : '\\');
			}
			if (uni && r.pct(15)) {
				static const char kHex[] = "0123456789abcdef";
				o.push_back('\\');
				o.push_back('u');
				for (int h = 0; h < 4; ++h) o.push_back(kHex[r.below(16)]);
			}
		}
		if (esc && len == 0) {
			o.push_back('\\');
			o.push_back('"');
		}
		o.push_back('"');
		str_bytes += o.size() - s;
	}

	void int_value() {
		if (r.pct(30))
			o.push_back('-');
		const int digits = (int)r.range(1, 18);
		o.push_back((char)('1' + r.below(9)));
		for (int i = 1; i < digits; ++i) o.push_back((char)('0' + r.below(10)));
	}

	void double_value() {
		const double m = ((double)(int64_t)(r.next() >> 11) / 9007199254740992.0) * 2000.0 - 1000.0;
		const int	 e = (int)r.range(0, 20) - 10;
		char		 buf[40];
		snprintf(buf, sizeof buf, "%.6g", m * pow(10.0, e));
		o += buf;
	}

	void literal() {
		static const char* kLit[] = {"true", "false", "null"};
		o += kLit[r.below(3)];
	}

	// -1: no target; 0: string share at/above target; 1: below target
	int string_deficit() const {
		if (str_target_pm < 0)
			return -1;
		const uint64_t total = base + o.size();
		const uint64_t pm	 = total ? str_bytes * 1000 / total : 0;
		return pm < (uint64_t)str_target_pm ? 1 : 0;
	}

	void value(int depth) {
		// Kind weights: string, int, double, literal, array, object.  Nesting
		// thins with depth so a top-level object stays ~0.5 KiB.  With a
		// string-share target, strings are chosen while the share is short of
		// it, and key-free content (number arrays, numbers) while it is above.
		const int	   deficit = string_deficit();
		const bool	   can	   = depth < max_depth;
		uint32_t	   w[6];
		if (deficit < 0) {
			const uint32_t wn = can ? 12 / (uint32_t)(depth + 1) : 0;
			w[0] = w[1] = w[2] = w[3] = 12;
			w[4] = w[5] = wn;
		} else if (deficit == 1) {
			w[0] = 40, w[1] = 6, w[2] = 6, w[3] = 4, w[4] = can ? 2 : 0, w[5] = can ? 2 : 0;
		} else {  // above the target: numbers and key-free number arrays dilute the keys' strings
			w[0] = 0, w[1] = 12, w[2] = 12, w[3] = 4, w[4] = can ? 32 / (uint32_t)depth : 0, w[5] = 0;
		}
		uint32_t tot = 0;
		for (uint32_t v : w) tot += v;
		uint32_t k = r.below(tot);
		int		 kind = 0;
		while (k >= w[kind]) k -= w[kind++];
		switch (kind) {
		case 0: return string_value();
		case 1: return int_value();
		case 2: return double_value();
		case 3: return literal();
		case 4: return array(depth + 1, -1);
		default: return object(depth + 1, -1);
		}
	}

	// chain_to: when >= 0, one member continues a chain of nested containers
	// down to that depth.
	void object(int depth, int chain_to) {
		o.push_back('{');
		ws();
		const int nk	= (int)r.range(4, 12);
		const int chain = chain_to > depth ? (int)r.below((uint32_t)nk) : -1;
		for (int i = 0; i < nk; ++i) {
			if (i) {
				o.push_back(',');
				ws();
			}
			key();
			ws();
			o.push_back(':');
			ws();
			if (i == chain)
				(r.below(2) ? (void)array(depth + 1, chain_to) : (void)object(depth + 1, chain_to));
			else
				value(depth);
			ws();
		}
		o.push_back('}');
	}

	void array(int depth, int chain_to) {
		o.push_back('[');
		ws();
		const int ne	= (int)r.range(chain_to > depth ? 1 : 0, 8);
		const int chain = chain_to > depth ? (int)r.below((uint32_t)ne) : -1;
		for (int i = 0; i < ne; ++i) {
			if (i) {
				o.push_back(',');
				ws();
			}
			if (i == chain)
				(r.below(2) ? (void)array(depth + 1, chain_to) : (void)object(depth + 1, chain_to));
			else
				value(depth);
			ws();
		}
		o.push_back(']');
	}
};

void pad_to(uint8_t* out, uint64_t have, uint64_t n) {
	if (have < n)
		memset(out + have, ' ', n - have);
}

// ---------------------------------------------------------------------------
// CSV
// ---------------------------------------------------------------------------
void csv_record(Rng& r, std::string& o) {
	for (int f = 0; f < 16; ++f) {
		if (f)
			o.push_back(',');
		if (r.pct(10)) {  // quoted, with at least one special
			o.push_back('"');
			const int len = (int)r.range(1, 24);
			bool	  any = false;
			for (int i = 0; i < len; ++i) {
				const uint32_t k = r.below(100);
				if (k < 8 || (i == len - 1 && !any)) {
					const uint32_t w = r.below(3);
					if (w == 0)
						o.push_back(',');
					else if (w == 1)
						o += "\r\n";
					else
						o += "\"\"";
					any = true;
				} else if (k < 70)
					o.push_back((char)('a' + r.below(26)));
				else if (k < 85)
					o.push_back((char)('0' + r.below(10)));
				else
					o.push_back(' ');
			}
			o.push_back('"');
		} else {
			const uint32_t kind = r.below(3);
			if (kind == 0) {
				const int len = (int)r.range(0, 12);
				for (int i = 0; i < len; ++i) o.push_back((char)('a' + r.below(26)));
			} else if (kind == 1) {
				const int len = (int)r.range(1, 9);
				for (int i = 0; i < len; ++i) o.push_back((char)('0' + r.below(10)));
			} else {
				char buf[32];
				snprintf(buf, sizeof buf, "%u.%02u", r.below(100000), r.below(100));
				o += buf;
			}
		}
	}
	o += "\r\n";
}

// ---------------------------------------------------------------------------
// C-like (six modes)
// ---------------------------------------------------------------------------
struct CGen {
	Rng			 r;
	std::string& o;
	int			 max_depth;
	int			 depth = 0;	 // current bracket depth
	int			 deep_budget = 0;

	CGen(uint64_t seed, std::string& out, int md) : r {seed}, o(out), max_depth(md) {}

	void ident() {
		static const char kFirst[] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ_";
		static const char kRest[]  = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ_0123456789";
		const int		  len	   = (int)r.range(1, 16);
		o.push_back(kFirst[r.below(sizeof(kFirst) - 1)]);
		for (int i = 1; i < len; ++i) o.push_back(kRest[r.below(sizeof(kRest) - 1)]);
	}

	void number() {
		char buf[40];
		switch (r.below(4)) {
		case 0: snprintf(buf, sizeof buf, "%u", r.below(1000000)); break;
		case 1: snprintf(buf, sizeof buf, "0x%x", r.below(0x7fffffff)); break;
		case 2: snprintf(buf, sizeof buf, "%u.%uf", r.below(1000), r.below(1000)); break;
		default: snprintf(buf, sizeof buf, "%ue%u", r.range(1, 9), r.below(30)); break;
		}
		o += buf;
	}

	// 40 operators / punctuators; ; and , are the structural ones in the plan.
	void op() {
		static const char* kOps[] = {"+",  "-",	 "*",  "/",	 "%",  "=",	 "==", "!=", "<",  ">",
									 "<=", ">=", "&&", "||", "!",  "&",	 "|",  "^",	 "~",  "<<",
									 ">>", "+=", "-=", "*=", "/=", "%=", "&=", "|=", "^=", "<<=",
									 ">>=", "++", "--", "->", ".",	"?",  ":",	",",  ";",	"::"};
		o += kOps[r.below(40)];
	}

	void string_lit() {
		o.push_back('"');
		const int len = (int)r.range(0, 60);
		const bool esc = r.pct(5);
		for (int i = 0; i < len; ++i) {
			const uint32_t k = r.below(100);
			if (k < 70)
				o.push_back((char)('a' + r.below(26)));
			else if (k < 80)
				o.push_back((char)('0' + r.below(10)));
			else if (k < 92)
				o.push_back(' ');
			else {
				static const char kP[] = ",;:(){}[].!?-+=<>%&";
				o.push_back(kP[r.below(sizeof(kP) - 1)]);
			}
			if (esc && r.pct(10)) {
				static const char kE[] = "\"\\nt0r";
				o.push_back('\\');
				o.push_back(kE[r.below(sizeof(kE) - 1)]);
			}
		}
		o.push_back('"');
	}

	void char_lit() {
		o.push_back('\'');
		if (r.pct(20)) {
			static const char kE[] = "nt0\\'r";
			o.push_back('\\');
			o.push_back(kE[r.below(sizeof(kE) - 1)]);
		} else {
			static const char kC[] = "abcxyzABC019 ;,(){}[]+-*";
			o.push_back(kC[r.below(sizeof(kC) - 1)]);
		}
		o.push_back('\'');
	}

	void comment_text(int len) {
		static const char kT[] = "abcdefghijklmnopqrstuvwxyz  ,;:(){}[]+-=<>.";
		for (int i = 0; i < len; ++i) o.push_back(kT[r.below(sizeof(kT) - 1)]);
	}

	void line_comment() {
		o += "//";
		comment_text((int)r.range(0, 60));
		o.push_back('\n');
	}

	void block_comment() {
		const bool stars = r.pct(1);
		o += stars ? "/*****" : "/*";
		const int lines = (int)r.range(1, 20);
		for (int l = 0; l < lines; ++l) {
			o.push_back(' ');
			comment_text((int)r.range(0, 50));
			if (stars && r.pct(50))
				o += " ** * ***";
			if (l + 1 < lines)
				o.push_back('\n');
		}
		o += stars ? " *****/\n" : " */\n";
	}

	void pp_line() {
		static const char* kPP[] = {"#include <", "#define ", "#ifdef ", "#endif", "#pragma once", "#if "};
		const uint32_t	   k	 = r.below(6);
		o += kPP[k];
		if (k == 0) {
			ident();
			o += ".h>";
		} else if (k == 1 || k == 2 || k == 5) {
			ident();
			if (k == 1) {
				o.push_back(' ');
				number();
			}
		}
		o.push_back('\n');
	}

	void sp() { o.push_back(' '); }

	// expression with bracket nesting
	void expr(int budget) {
		const int terms = (int)r.range(1, 5);
		for (int t = 0; t < terms; ++t) {
			if (t) {
				sp();
				op();
				sp();
			}
			const uint32_t k	  = r.below(100);
			const bool	   nest	  = depth < max_depth && (deep_budget > 0 || k < (uint32_t)(budget > 0 ? 22 : 0));
			if (nest) {
				const uint32_t w = r.below(3);
				if (w == 0) {  // ( expr )
					o.push_back('(');
					++depth;
					--deep_budget;
					expr(budget - 1);
					o.push_back(')');
					--depth;
				} else if (w == 1) {  // call f(a, b)
					ident();
					o.push_back('(');
					++depth;
					--deep_budget;
					const int args = (int)r.range(0, 3);
					for (int a = 0; a < args; ++a) {
						if (a)
							o += ", ";
						expr(budget - 1);
					}
					o.push_back(')');
					--depth;
				} else {  // a[expr]
					ident();
					o.push_back('[');
					++depth;
					--deep_budget;
					expr(budget - 1);
					o.push_back(']');
					--depth;
				}
			} else if (k < 50)
				ident();
			else if (k < 70)
				number();
			else if (k < 85)
				string_lit();
			else
				char_lit();
		}
	}

	void indent() {
		for (int i = 0; i < depth && i < 8; ++i) o += "  ";
	}

	void statement(int budget) {
		indent();
		const uint32_t k = r.below(100);
		if (k < 10) {
			line_comment();
			return;
		}
		if (k < 12) {
			block_comment();
			return;
		}
		if (k < 14) {
			pp_line();
			return;
		}
		if (k < 24 && depth < max_depth && budget > 0) {  // nested block
			o += r.below(2) ? "if (" : "while (";
			++depth;
			expr(1);
			--depth;
			o += ") {\n";
			++depth;
			const int n = (int)r.range(1, 5);
			for (int i = 0; i < n; ++i) statement(budget - 1);
			--depth;
			indent();
			o += "}\n";
			return;
		}
		expr(3);
		o += ";\n";
	}

	void function() {
		if (r.permille(20))
			deep_budget = max_depth;  // occasionally drive one expression to the maximum depth
		o += "static int ";
		ident();
		o.push_back('(');
		++depth;
		o += "int a, char *b";
		--depth;
		o += ") {\n";
		++depth;
		const int n = (int)r.range(2, 12);
		for (int i = 0; i < n; ++i) statement(4);
		--depth;
		o += "}\n\n";
		deep_budget = 0;
	}
};

// ---------------------------------------------------------------------------
// Segmented drivers: a document of n bytes is cut into `nseg` fixed regions
// (nseg = the caller's count, lowered so regions stay >= 1 MiB); each region is
// generated from its own seed, independently, so threads can fill regions in
// any order and the bytes never depend on the thread count.  A region holds
// whole units (objects, records, functions) followed by spaces.
// ---------------------------------------------------------------------------
template<class F>
void run_regions(uint64_t nseg, F&& fill_region) {
	if (nseg == 1) {  // small documents: no threads
		fill_region(0);
		return;
	}
	unsigned nt = std::thread::hardware_concurrency();
	nt			= nt ? (nt > 32 ? 32 : nt) : 1;
	if (nt > nseg)
		nt = (unsigned)nseg;
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nt; ++t)
		th.emplace_back([&, t] {
			for (uint64_t i = t; i < nseg; i += nt) fill_region(i);
		});
	for (auto& x : th) x.join();
}

uint64_t region_count(uint64_t n, uint64_t nseg) {
	const uint64_t by_size = n >> 20;
	if (nseg > by_size)
		nseg = by_size;
	return nseg ? nseg : 1;
}

}  // namespace

extern "C" {

// A top-level JSON array of objects, exactly n bytes.  max_depth bounds the
// bracket depth (the array is depth 1); str_target_permille < 0 leaves the
// string share to the value mix.  Returns the string-byte count (quotes
// included) through *str_bytes when non-null.
int pars_gen_json(uint64_t seed, uint64_t n, int max_depth, int str_target_permille, uint64_t nseg, uint8_t* out,
				  uint64_t* str_bytes) {
	if (max_depth < 2)
		max_depth = 2;
	if (n < 2) {
		pad_to(out, 0, n);
		if (str_bytes)
			*str_bytes = 0;
		return 0;
	}
	nseg			   = region_count(n, nseg);
	const uint64_t	 B = (n - 2) / nseg;
	std::vector<uint64_t> sb(nseg, 0);
	out[0] = '[';
	run_regions(nseg, [&](uint64_t i) {
		uint8_t* reg = out + 1 + i * B;
		std::string buf;
		buf.reserve(1 << 16);
		JsonGen	 g(nseg == 1 ? seed : mix(seed, i), buf, max_depth, str_target_permille);
		uint64_t have = 0;
		bool	 first = true;
		while (true) {
			buf.clear();
			if (!first || i > 0) {
				buf.push_back(',');
				g.ws();
			}
			const uint64_t s0 = g.str_bytes;
			g.base			  = have;
			const int chain	  = (max_depth > 8 && g.r.permille(4)) ? max_depth : -1;
			g.object(2, chain);
			if (have + buf.size() > B) {
				g.str_bytes = s0;
				break;
			}
			memcpy(reg + have, buf.data(), buf.size());
			have += buf.size();
			first = false;
		}
		pad_to(reg, have, B);
		sb[i] = g.str_bytes;
	});
	// a region with no object would leave a dangling separator: never happens
	// for regions >= 1 MiB; the single-region case never writes a leading comma.
	out[1 + nseg * B] = ']';
	pad_to(out, 2 + nseg * B, n);
	if (str_bytes) {
		uint64_t t = 0;
		for (uint64_t v : sb) t += v;
		*str_bytes = t;
	}
	return 0;
}

int pars_gen_csv(uint64_t seed, uint64_t n, uint64_t nseg, uint8_t* out) {
	nseg			 = region_count(n, nseg);
	const uint64_t B = n / nseg;
	run_regions(nseg, [&](uint64_t i) {
		Rng			r {nseg == 1 ? seed : mix(seed, i)};
		std::string buf;
		uint8_t*	reg	 = out + i * B;
		uint64_t	have = 0;
		while (true) {
			buf.clear();
			csv_record(r, buf);
			if (have + buf.size() > B)
				break;
			memcpy(reg + have, buf.data(), buf.size());
			have += buf.size();
		}
		pad_to(reg, have, B);
	});
	pad_to(out, nseg * B, n);
	return 0;
}

int pars_gen_clike(uint64_t seed, uint64_t n, int max_depth, uint64_t nseg, uint8_t* out) {
	nseg			 = region_count(n, nseg);
	const uint64_t B = n / nseg;
	run_regions(nseg, [&](uint64_t i) {
		std::string buf;
		CGen		g(nseg == 1 ? seed : mix(seed, i), buf, max_depth < 1 ? 1 : max_depth);
		uint8_t*	reg	 = out + i * B;
		uint64_t	have = 0;
		while (true) {
			buf.clear();
			g.depth = 0;
			if (g.r.pct(20))
				g.pp_line();
			g.function();
			if (have + buf.size() > B)
				break;
			memcpy(reg + have, buf.data(), buf.size());
			have += buf.size();
		}
		pad_to(reg, have, B);
	});
	pad_to(out, nseg * B, n);
	return 0;
}

int pars_gen_alpha(uint64_t seed, uint64_t n, const char* alphabet, uint64_t alen, uint8_t* out) {
	if (!alen)
		return -1;
	Rng r {seed};
	for (uint64_t i = 0; i < n; ++i) out[i] = (uint8_t)alphabet[r.below((uint32_t)alen)];
	return 0;
}

// Log-normal document sizes: median * exp(sigma * z), z ~ N(0,1) by the
// Box-Muller transform, clipped to [lo, hi].
int pars_gen_sizes(uint64_t seed, uint64_t count, double median, double sigma, uint64_t lo, uint64_t hi,
				   uint64_t* out) {
	Rng r {seed};
	for (uint64_t i = 0; i < count; ++i) {
		const double u1 = ((double)(r.next() >> 11) + 1.0) / 9007199254740993.0;
		const double u2 = (double)(r.next() >> 11) / 9007199254740992.0;
		const double z	= sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
		double		 v	= median * exp(sigma * z);
		if (v < (double)lo)
			v = (double)lo;
		if (v > (double)hi)
			v = (double)hi;
		out[i] = (uint64_t)v;
	}
	return 0;
}

// Many documents into one buffer: document i (kind 0 JSON with depth <= 8,
// 1 CSV, 2 C-like with depth <= 24) of sizes[i] bytes from seeds[i] at
// out + offsets[i]; documents are generated in parallel, each exactly as the
// single-document generator makes it.
int pars_gen_docs(uint64_t count, const uint8_t* kinds, const uint64_t* seeds, const uint64_t* sizes,
				  const uint64_t* offsets, uint8_t* out) {
	unsigned nt = std::thread::hardware_concurrency();
	nt			= nt ? (nt > 32 ? 32 : nt) : 1;
	std::vector<std::thread> th;
	std::vector<int>		 rc(nt, 0);
	for (unsigned t = 0; t < nt; ++t)
		th.emplace_back([&, t] {
			for (uint64_t i = t; i < count; i += nt) {
				uint8_t* o = out + offsets[i];
				int		 r = 0;
				switch (kinds[i]) {
				case 0: r = pars_gen_json(seeds[i], sizes[i], 8, -1, 1, o, nullptr); break;
				case 1: r = pars_gen_csv(seeds[i], sizes[i], 1, o); break;
				case 2: r = pars_gen_clike(seeds[i], sizes[i], 24, 1, o); break;
				default: r = -1;
				}
				if (r)
					rc[t] = r;
			}
		});
	for (auto& x : th) x.join();
	for (int r : rc)
		if (r)
			return r;
	return 0;
}

// Per-document seed (splitmix(seed ^ index)).
uint64_t pars_gen_doc_seed(uint64_t seed, uint64_t index) {
	return mix(seed, index);
}

}  // extern "C"
