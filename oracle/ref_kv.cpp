// oracle/ref_kv.cpp — the reference's own INI and TOML stage-2 hooks
// (bench/bench_parse_host.cu:30-248: is_ws, each_token, fnv, ini_parse,
// toml_parse), fed as the reference's full-parse benchmark feeds them
// (:337-360, 361-384): its CPU scanner gives the structural index
// (model/scan.hpp), model::replay (model/visit.hpp:35-48) drives the hook.
// Writes golden vectors for the key/value restatement (oracle/kv_hooks.py).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ (git
// ignored), where parse_host.inc holds the hooks' lines taken unchanged from
// the reference's source file at build time; used by
// tests/golden/make_kv_golden.py.  Nothing in the product links it.
//
//   ref_kv ini|toml <in>     prints one JSON object: the hook's hash and, per
//                            list it fills, the item count and an FNV-1a 64
//                            over (u32 length, bytes) of each item in order
#include "pars/pars.hpp"

#include "grammars/ini.hpp"
#include "grammars/toml.hpp"

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

namespace ref_hooks {
using namespace pars;
#include "parse_host.inc"
}  // namespace ref_hooks

namespace {

constexpr auto kIni	 = pars::compile::make_scan_plan<pars::bench::ini_grammar>();
constexpr auto kToml = pars::compile::make_scan_plan<pars::bench::toml_grammar>();

std::string read_file(const char* path) {
	std::FILE* f = std::fopen(path, "rb");
	if (!f) {
		std::perror(path);
		std::exit(2);
	}
	std::string s;
	char		buf[1 << 16];
	std::size_t r;
	while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, r);
	std::fclose(f);
	return s;
}

struct List {
	std::uint64_t n = 0, h = 1469598103934665603ull;
	void		  add(std::string_view s) {
		 const std::uint32_t len = (std::uint32_t)s.size();
		 for (int k = 0; k < 4; ++k) mix((std::uint8_t)(len >> (8 * k)));
		 for (unsigned char c : s) mix(c);
		 ++n;
	}
	void mix(std::uint8_t b) {
		h ^= b;
		h *= 1099511628211ull;
	}
};

void print_list(const char* name, const List& l, bool last) {
	std::printf("\"%s\": {\"n\": %llu, \"fnv\": %llu}%s", name, (unsigned long long)l.n, (unsigned long long)l.h,
				last ? "" : ", ");
}

template<pars::scan_plan S>
std::vector<std::int32_t> structural(const std::string& s) {
	pars::model::scanner<S>		   sc;
	const pars::model::scan_result r = sc.scan(s);
	return std::vector<std::int32_t>(r.positions.begin(), r.positions.end());
}

}  // namespace

int main(int argc, char** argv) {
	if (argc != 3) {
		std::fprintf(stderr, "usage: ref_kv ini|toml <in>\n");
		return 2;
	}
	const std::string	s	 = read_file(argv[2]);
	const std::uint8_t* base = reinterpret_cast<const std::uint8_t*>(s.data());
	if (std::strcmp(argv[1], "ini") == 0) {
		const std::vector<std::int32_t> pos = structural<kIni>(s);
		pars::model::structural_view	sv {pos.data(), nullptr, pos.size()};
		ref_hooks::ini_parse			p {};
		p.base = base;
		pars::model::replay(base, s.size(), sv, p);
		List sec, keys, vals;
		for (const auto& x : p.sections) sec.add(x);
		for (const auto& k : p.keys) keys.add(std::string_view(p.arena).substr(k.first, k.second));
		for (const auto& v : p.vals) vals.add(std::string_view(p.arena).substr(v.first, v.second));
		std::printf("{\"hash\": %llu, ", (unsigned long long)p.hash);
		print_list("sections", sec, false);
		print_list("keys", keys, false);
		print_list("vals", vals, true);
		std::printf("}\n");
		return 0;
	}
	const std::vector<std::int32_t> pos = structural<kToml>(s);
	pars::model::structural_view	sv {pos.data(), nullptr, pos.size()};
	ref_hooks::toml_parse			p {};
	p.base = base;
	pars::model::replay(base, s.size(), sv, p);
	List sec, ent;
	for (const auto& x : p.sections) sec.add(x);
	for (const auto& e : p.entries) {
		ent.add(e.first);
		ent.add(e.second);
	}
	std::printf("{\"hash\": %llu, ", (unsigned long long)p.hash);
	print_list("sections", sec, false);
	print_list("entries", ent, false);
	std::printf("\"elements\": {\"n\": %llu}}\n", (unsigned long long)p.elems);
	return 0;
}
