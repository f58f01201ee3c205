// dropin_check — compile-only proof that the binding INTEGRATION.md gives a
// maintainer (tests/cpp/dropin/pars/backend/hip/pipeline.hpp, the same text)
// is a drop-in for the reference's pars::gpu (pipeline.cuh:1164-1281): it is
// compiled against the reference's own headers (-I<pars>/include,
// -I<pars>/bench), with plans from the reference compiler
// (compile::make_scan_plan<G>()), and every call below is written the way the
// reference's bench drivers write it (bench/bench_boundary.cu:109-121,
// 134-189, 225-296).  Nothing here runs; the object file is the check.
#include "pars/pars.hpp"
#include "pars/backend/hip/pipeline.hpp"

#include "grammars/clike.hpp"
#include "grammars/csv.hpp"
#include "grammars/ini.hpp"
#include "grammars/json.hpp"
#include "grammars/toml.hpp"
#include "grammars/xml.hpp"

#include <cstdint>
#include <string>
#include <vector>

using namespace pars;

// bench_boundary.cu:109-121
template<scan_plan S>
static bool validate(const std::string& s) {
	model::scanner<S> cpu;
	auto			  c = cpu.scan(s);
	gpu::scanner<S>	  g;
	auto			  r = g.scan_host(reinterpret_cast<const std::uint8_t*>(s.data()), s.size());
	if (c.positions.size() != r.structural.size())
		return false;
	for (std::size_t i = 0; i < c.positions.size(); ++i)
		if ((std::int32_t)c.positions[i] != r.structural[i])
			return false;
	return true;
}

// bench_boundary.cu:134-189: chunks with the carry, the structural index
// read back through struct_ptr()
template<scan_plan S>
static std::uint32_t chunked(gpu::scanner<S>& sc, const std::uint8_t* const dev[2], std::size_t n,
							 std::size_t chunk, hipStream_t s0) {
	sc.reserve_struct(n);
	gpu::chunk_carry carry {};
	std::size_t		 off = 0;
	for (int i = 0; off < n; ++i) {
		const std::size_t len = std::min(chunk, n - off);
		carry				  = sc.scan_chunk(dev[i & 1], len, carry, s0, off);
		off += len;
	}
	const std::int32_t* idx = sc.struct_ptr();
	(void)idx;
	return carry.out_base;
}

// bench_boundary.cu:225-296: the tiers, stage times, device and host results
template<scan_plan S>
static double tiers(const std::uint8_t* dp, const std::string& payload) {
	const std::size_t n = payload.size();
	gpu::scanner<S>	  g;
	gpu::stage_times  st {};
	auto			  dr	 = g.scan_device(dp, n, 0, false, true, &st);
	const std::uint32_t tokens = dr.n_struct;
	(void)g.scan_device(dp, n, 0, false, false);
	(void)g.scan_device(dp, n, 0, false, true);
	auto hr = g.scan_host(reinterpret_cast<const std::uint8_t*>(payload.data()), n, 0, false);
	const bool ok = dr.balanced && dr.well_formed && hr.balanced && dr.max_depth >= 0;
	const std::int32_t* pp = dr.pair_pos;
	(void)pp;
	return st.p1 + st.p2 + st.p3 + st.p4 + tokens + (ok ? 1 : 0) + (double)hr.structural.size();
}

// the free functions (pipeline.cuh:1260-1281)
template<scan_plan S>
static std::size_t free_functions(const std::uint8_t* dp, const std::string& s) {
	auto d = gpu::scan_device<S>(dp, s.size(), nullptr, true);
	auto h = gpu::scan_host<S>(reinterpret_cast<const std::uint8_t*>(s.data()), s.size(), true);
	auto c = gpu::scan_chunk<S>(dp, s.size(), gpu::chunk_carry {}, nullptr, 0);
	return d.n_struct + h.structural.size() + h.pair_pos.size() + c.out_base;
}

template<scan_plan S>
static double all(const std::uint8_t* const dev[2], const std::string& s) {
	gpu::scanner<S> sc;
	return validate<S>(s) + chunked<S>(sc, dev, s.size(), 1 << 20, nullptr) + tiers<S>(dev[0], s)
		 + (double)free_functions<S>(dev[0], s);
}

// bench_boundary.cu:339-344: the six plans, compiled by the reference compiler
double dropin_check(const std::uint8_t* const dev[2], const std::string& s) {
	constexpr auto kJ = compile::make_scan_plan<bench::json_grammar>();
	constexpr auto kC = compile::make_scan_plan<bench::csv_grammar>();
	constexpr auto kT = compile::make_scan_plan<bench::toml_grammar>();
	constexpr auto kL = compile::make_scan_plan<bench::clike_grammar>();
	constexpr auto kI = compile::make_scan_plan<bench::ini_grammar>();
	constexpr auto kX = compile::make_scan_plan<bench::xml_grammar>();
	return all<kJ>(dev, s) + all<kC>(dev, s) + all<kT>(dev, s) + all<kL>(dev, s) + all<kI>(dev, s)
		 + all<kX>(dev, s);
}
