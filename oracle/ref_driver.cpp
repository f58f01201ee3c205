// oracle/ref_driver.cpp — runs the reference's own CPU backend, compiled from
// its sources where they lie (the reference tree's include/ and bench/grammars/),
// to write golden vectors and to time the reference CPU path.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ (git
// ignored); used by tests/golden/make_golden.py and by bench.py's cpu_baseline
// leg.  Nothing in the product links it.
//
//   ref_driver plan <grammar> <out.plan>         the compiled scan_plan bytes
//   ref_driver scan <grammar> <in> <out.bin>     model::scanner<S>::scan (scan.hpp:27-73)
//                                                 + model::pair_brackets<S> (nest.hpp:45-92)
//   ref_driver time <grammar> <in> <max_bytes>   scan + pair on the first max_bytes, one core
//
// Output of `scan` (little endian): "PRSREF01", u64 n_pos, u64 n_br, i32
// unclosed, balanced, well_formed, max_depth, then u32 positions[n_pos], u8
// chars[n_pos], u32 bracket_pos[n_br], u8 bracket_char[n_br], i32 pair[n_br].
#include "pars/pars.hpp"

#include "grammars/clike.hpp"
#include "grammars/csv.hpp"
#include "grammars/ini.hpp"
#include "grammars/json.hpp"
#include "grammars/toml.hpp"
#include "grammars/xml.hpp"

#include "grammars/clike6.hpp"
#include "grammars/compiler_cases.hpp"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <vector>

using namespace pars;

namespace {

constexpr auto kJson   = compile::make_scan_plan<bench::json_grammar>();
constexpr auto kCsv	   = compile::make_scan_plan<bench::csv_grammar>();
constexpr auto kClike  = compile::make_scan_plan<bench::clike_grammar>();
constexpr auto kToml   = compile::make_scan_plan<bench::toml_grammar>();
constexpr auto kIni	   = compile::make_scan_plan<bench::ini_grammar>();
constexpr auto kXml	   = compile::make_scan_plan<bench::xml_grammar>();
constexpr auto kClike6 = compile::make_scan_plan<pars_amd::grammars::clike6_grammar>();

// the compiler probe grammars (grammars/compiler_cases.hpp)
namespace cc	  = pars_amd::compiler_cases;
constexpr auto kSqlCase	  = compile::make_scan_plan<cc::sql_case>();
constexpr auto kEscCase	  = compile::make_scan_plan<cc::esc_case>();
constexpr auto kManyCase  = compile::make_scan_plan<cc::many_case>();
constexpr auto kCapsCase  = compile::make_scan_plan<cc::caps_case>();
constexpr auto kTinyCase  = compile::make_scan_plan<cc::tiny_case>();
constexpr auto kUnsupCase = compile::make_scan_plan<cc::unsup_case>();
constexpr auto kPropCase  = compile::make_scan_plan<cc::prop_case>();

static_assert(sizeof(scan_plan) == 3177, "plan layout changed");

std::string read_file(const char* path, std::size_t max_bytes = SIZE_MAX) {
	std::FILE* f = std::fopen(path, "rb");
	if (!f) {
		std::perror(path);
		std::exit(2);
	}
	std::string s;
	char		buf[1 << 16];
	std::size_t r;
	while (s.size() < max_bytes && (r = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, r);
	std::fclose(f);
	if (s.size() > max_bytes)
		s.resize(max_bytes);
	return s;
}

template<scan_plan S>
int do_plan(const char* out) {
	std::FILE* f = std::fopen(out, "wb");
	if (!f)
		return 2;
	std::fwrite(&S, 1, sizeof(S), f);
	std::fclose(f);
	return 0;
}

template<scan_plan S>
int do_scan(const char* in, const char* out) {
	const std::string		  s = read_file(in);
	model::scanner<S>		  sc;
	const model::scan_result  r = sc.scan(s);
	const model::pair_table	  t = model::pair_brackets<S>(r.bracket_char, r.bracket_pos);
	std::FILE*				  f = std::fopen(out, "wb");
	if (!f)
		return 2;
	const std::uint64_t np = r.positions.size(), nb = r.bracket_pos.size();
	const std::int32_t	hdr[4] = {r.unclosed_mode ? 1 : 0, t.balanced ? 1 : 0, t.well_formed ? 1 : 0,
								  t.max_depth};
	std::fwrite("PRSREF01", 1, 8, f);
	std::fwrite(&np, 8, 1, f);
	std::fwrite(&nb, 8, 1, f);
	std::fwrite(hdr, 4, 4, f);
	std::fwrite(r.positions.data(), 4, np, f);
	std::fwrite(r.chars.data(), 1, np, f);
	std::fwrite(r.bracket_pos.data(), 4, nb, f);
	std::fwrite(r.bracket_char.data(), 1, nb, f);
	std::fwrite(t.pair.data(), 4, nb, f);
	std::fclose(f);
	return 0;
}

template<scan_plan S>
int do_time(const char* in, std::size_t max_bytes) {
	const std::string s = read_file(in, max_bytes);
	model::scanner<S> sc;
	auto			  t0 = std::chrono::steady_clock::now();
	const auto		  r	 = sc.scan(s);
	auto			  t1 = std::chrono::steady_clock::now();
	const auto		  p	 = model::pair_brackets<S>(r.bracket_char, r.bracket_pos);
	auto			  t2 = std::chrono::steady_clock::now();
	const double	  a	 = std::chrono::duration<double>(t1 - t0).count();
	const double	  b	 = std::chrono::duration<double>(t2 - t1).count();
	std::printf(
		"{\"bytes\": %zu, \"scan_s\": %.6f, \"pair_s\": %.6f, \"n_struct\": %zu, \"n_brackets\": %zu, "
		"\"balanced\": %d, \"max_depth\": %d}\n",
		s.size(), a, b, r.positions.size(), r.bracket_pos.size(), p.balanced ? 1 : 0, p.max_depth
	);
	return 0;
}

template<scan_plan S>
int dispatch(int argc, char** argv) {
	const std::string cmd = argv[1];
	if (cmd == "plan" && argc >= 4)
		return do_plan<S>(argv[3]);
	if (cmd == "scan" && argc >= 5)
		return do_scan<S>(argv[3], argv[4]);
	if (cmd == "time" && argc >= 5)
		return do_time<S>(argv[3], (std::size_t)std::strtoull(argv[4], nullptr, 10));
	std::fprintf(stderr, "bad command\n");
	return 2;
}

}  // namespace

int main(int argc, char** argv) {
	if (argc < 3) {
		std::fprintf(stderr, "usage: ref_driver plan|scan|time <grammar> ...\n");
		return 2;
	}
	const std::string g = argv[2];
	if (g == "json")
		return dispatch<kJson>(argc, argv);
	if (g == "csv")
		return dispatch<kCsv>(argc, argv);
	if (g == "clike")
		return dispatch<kClike>(argc, argv);
	if (g == "toml")
		return dispatch<kToml>(argc, argv);
	if (g == "ini")
		return dispatch<kIni>(argc, argv);
	if (g == "xml")
		return dispatch<kXml>(argc, argv);
	if (g == "clike6")
		return dispatch<kClike6>(argc, argv);
	if (g == "sql_case")
		return dispatch<kSqlCase>(argc, argv);
	if (g == "esc_case")
		return dispatch<kEscCase>(argc, argv);
	if (g == "many_case")
		return dispatch<kManyCase>(argc, argv);
	if (g == "caps_case")
		return dispatch<kCapsCase>(argc, argv);
	if (g == "tiny_case")
		return dispatch<kTinyCase>(argc, argv);
	if (g == "unsup_case")
		return dispatch<kUnsupCase>(argc, argv);
	if (g == "prop_case")
		return dispatch<kPropCase>(argc, argv);
	std::fprintf(stderr, "unknown grammar %s\n", g.c_str());
	return 2;
}
