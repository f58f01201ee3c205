// oracle/ref_plans.cpp — writes the scan_plan bytes the reference compiler
// (include/pars/compile/plan.hpp:711-748 make_scan_plan) produces for the
// compiler probe grammars of grammars/compiler_cases.hpp, and whether the
// reference's support gate (plan.hpp:237-238 is_supported) admits each.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/ (git
// ignored); run by tests/golden/make_plans.py.
//
//   ref_plans <out_dir>     writes <out_dir>/<name>.plan, prints one
//                           "<name> <supported 0|1>" line per grammar
#include "pars/pars.hpp"

#include "grammars/compiler_cases.hpp"

#include <cstdio>
#include <string>

using namespace pars;
namespace cc = pars_amd::compiler_cases;

namespace {

static_assert(sizeof(scan_plan) == 3177, "plan layout changed");

template<class G>
int emit(const std::string& dir, const char* name) {
	constexpr scan_plan S = compile::make_scan_plan<G>();
	const std::string	path = dir + "/" + name + ".plan";
	std::FILE*			f	 = std::fopen(path.c_str(), "wb");
	if (!f) {
		std::perror(path.c_str());
		return 1;
	}
	std::fwrite(&S, 1, sizeof(S), f);
	std::fclose(f);
	std::printf("%s %d\n", name, compile::is_supported<G> ? 1 : 0);
	return 0;
}

}  // namespace

int main(int argc, char** argv) {
	if (argc < 2) {
		std::fprintf(stderr, "usage: ref_plans <out_dir>\n");
		return 2;
	}
	const std::string d = argv[1];
	int				  rc = 0;
	rc |= emit<cc::sql_case>(d, "sql_case");
	rc |= emit<cc::esc_case>(d, "esc_case");
	rc |= emit<cc::many_case>(d, "many_case");
	rc |= emit<cc::caps_case>(d, "caps_case");
	rc |= emit<cc::tiny_case>(d, "tiny_case");
	rc |= emit<cc::unsup_case>(d, "unsup_case");
	rc |= emit<cc::prop_case>(d, "prop_case");
	return rc;
}
