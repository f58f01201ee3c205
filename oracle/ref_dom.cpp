// oracle/ref_dom.cpp — the reference's own JSON DOM hook (bench/bench_parse.cu:
// 34-354, `json_parser`), fed exactly as the reference's pars_parse feeds it
// (bench_parse.cu:357-371): the reference's CPU scanner gives the structural
// index (model/scan.hpp), model::replay (model/visit.hpp:35-48) drives the
// hook, finish() gives the verdict.  Writes golden vectors for the stage-2
// restatement (oracle/json_dom.py, oracle/json_dom.c) on malformed input too,
// where Python's json module cannot serve as the check.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ (git
// ignored), where json_parser.inc holds the hook's lines taken unchanged from
// the reference's source file at build time; used by
// tests/golden/make_dom_golden.py.  Nothing in the product links it.
//
//   ref_dom <in> <out.bin>
//
// Output (little endian): "PRSDOM01", u8 ok, u64 nodes, u64 hash, u64 tape
// length, then per tape node u8 kind, u32 begin, u32 end.
#include "pars/pars.hpp"

#include "grammars/json.hpp"

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace ref_hook {
using namespace pars;
#include "json_parser.inc"
}  // namespace ref_hook

namespace {

constexpr auto kJson = pars::compile::make_scan_plan<pars::bench::json_grammar>();

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

}  // namespace

int main(int argc, char** argv) {
	if (argc != 3) {
		std::fprintf(stderr, "usage: ref_dom <in> <out.bin>\n");
		return 2;
	}
	const std::string s = read_file(argv[1]);
	pars::model::scanner<kJson>		 sc;
	const pars::model::scan_result	 r = sc.scan(s);
	const std::vector<std::int32_t> pos(r.positions.begin(), r.positions.end());
	ref_hook::json_parser			 p;
	p.base = reinterpret_cast<const std::uint8_t*>(s.data());
	p.n	   = s.size();
	pars::model::structural_view sv;
	sv.pos	= pos.data();
	sv.pair = nullptr;
	sv.n	= pos.size();
	pars::model::replay(p.base, p.n, sv, p);
	p.finish();
	std::FILE* f = std::fopen(argv[2], "wb");
	if (!f)
		return 2;
	const std::uint8_t	ok	  = p.ok ? 1 : 0;
	const std::uint64_t nodes = p.nodes, hash = p.hash, len = p.tape.size();
	std::fwrite("PRSDOM01", 1, 8, f);
	std::fwrite(&ok, 1, 1, f);
	std::fwrite(&nodes, 8, 1, f);
	std::fwrite(&hash, 8, 1, f);
	std::fwrite(&len, 8, 1, f);
	for (const auto& t : p.tape) {
		const std::uint8_t	k = t.kind;
		const std::uint32_t b = t.begin, e = t.end;
		std::fwrite(&k, 1, 1, f);
		std::fwrite(&b, 4, 1, f);
		std::fwrite(&e, 4, 1, f);
	}
	std::fclose(f);
	return 0;
}
