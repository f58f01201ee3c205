// pars_hip.hpp — C++ facade over the C ABI that restores the reference
// backend's template API (include/pars/backend/cuda/pipeline.cuh:1164-1281):
//
//   pars::hip::scanner<S>      scan_device / scan_host / reserve_struct /
//                              scan_chunk / struct_ptr, json_tape (the
//                              reference's stage-2 JSON hook, bench_parse.cu:
//                              336-348, built on the device) and kv_parse
//                              (its INI / TOML hooks, bench_parse_host.cu)
//   pars::hip::scan_device<S>, scan_host<S>, scan_chunk<S>
//                              one scanner per plan per thread (:1260-1281)
//   pars::hip::replay, structural_view, structural_cursor, trim
//                              the consumer side over a host result
//                              (include/pars/model/visit.hpp, cursor.hpp)
//
// S is the caller's compiled plan (a `pars::scan_plan` constant from
// compile::make_scan_plan<G>(), or a `pars_plan`); its bytes are handed over
// as-is, which is why pars_plan keeps scan_plan's exact layout.
// `hipStream_t` replaces `cudaStream_t`; errors throw pars::hip::error (the
// reference prints and calls exit(1), pipeline.cuh:520-555).
#pragma once

#include "pars_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace pars::hip {

class error : public std::runtime_error {
public:
	error(int status, const char* what) : std::runtime_error(std::string(what)), status_(status) {}

	int status() const noexcept { return status_; }

private:
	int status_;
};

inline void check(int status) {
	if (status != PARS_OK)
		throw error(status, pars_hip_last_error());
}

// pipeline.cuh:722-728 (positions are handed out as int32 like the reference;
// the values are identical below 2^31, the uint32 view is `positions`).
struct device_result {
	const std::int32_t*	 structural = nullptr;
	const std::int32_t*	 pair_pos	= nullptr;
	std::uint32_t		 n_struct	= 0;
	bool				 balanced = true, well_formed = true;
	int					 max_depth	= 0;
	// extensions
	const std::uint32_t* positions	   = nullptr;
	const std::uint8_t*	 chars		   = nullptr;
	const std::uint64_t* struct_bitmap = nullptr;
	const std::uint32_t* bracket_pos   = nullptr;
	const std::uint8_t*	 bracket_char  = nullptr;
	const std::int32_t*	 bracket_pair  = nullptr;
	std::uint32_t		 n_brackets	   = 0;
	bool				 unclosed_mode = false;
};

// pipeline.cuh:730-735
struct host_result {
	std::vector<std::int32_t> structural;
	std::vector<std::int32_t> pair_pos;
	bool					  balanced = true, well_formed = true;
	int						  max_depth		= 0;
	bool					  unclosed_mode = false;
};

// pipeline.cuh:737-743 (zero-initialised = document start, as in the reference)
using chunk_carry = pars_chunk_carry;

// pipeline.cuh:745-748
using stage_times = pars_stage_times;

// The JSON tape (pars_hip_json_tape): device pointers, valid until the
// scanner's next call.
using json_tape = pars_json_tape;

// The INI / TOML key-value table (pars_hip_kv_parse): device pointers, valid
// until the scanner's next call.
using kv_table = pars_kv_table;

// A scanner for a plan known at run time.
class runtime_scanner {
public:
	explicit runtime_scanner(const pars_plan& plan) { check(pars_hip_scanner_create(&plan, &h_)); }

	runtime_scanner(const runtime_scanner&)			   = delete;
	runtime_scanner& operator=(const runtime_scanner&) = delete;

	~runtime_scanner() { pars_hip_scanner_destroy(h_); }

	device_result scan_device(const std::uint8_t* d_data, std::size_t n, hipStream_t s = nullptr,
							  bool want_pairs = false, bool run_pairs = true, stage_times* st = nullptr) const {
		pars_device_result r {};
		check(pars_hip_scan_device(h_, d_data, n, s, want_pairs, run_pairs, st, &r));
		device_result o;
		o.structural	= reinterpret_cast<const std::int32_t*>(r.structural);
		o.pair_pos		= r.pair_pos;
		o.n_struct		= r.n_struct;
		o.balanced		= r.balanced != 0;
		o.well_formed	= r.well_formed != 0;
		o.max_depth		= r.max_depth;
		o.positions		= r.structural;
		o.chars			= r.chars;
		o.struct_bitmap = r.struct_bitmap;
		o.bracket_pos	= r.bracket_pos;
		o.bracket_char	= r.bracket_char;
		o.bracket_pair	= r.bracket_pair;
		o.n_brackets	= r.n_brackets;
		o.unclosed_mode = r.unclosed_mode != 0;
		return o;
	}

	host_result scan_host(const std::uint8_t* h_data, std::size_t n, hipStream_t s = nullptr,
						  bool want_pairs = false) const {
		pars_host_result r {};
		check(pars_hip_scan_host(h_, h_data, n, s, want_pairs, &r));
		host_result o;
		o.structural.assign(reinterpret_cast<const std::int32_t*>(r.structural),
							reinterpret_cast<const std::int32_t*>(r.structural) + r.n_struct);
		if (want_pairs && r.pair_pos)
			o.pair_pos.assign(r.pair_pos, r.pair_pos + r.n_struct);
		o.balanced		= r.balanced != 0;
		o.well_formed	= r.well_formed != 0;
		o.max_depth		= r.max_depth;
		o.unclosed_mode = r.unclosed_mode != 0;
		return o;
	}

	void reserve_struct(std::size_t cap) const { check(pars_hip_reserve_struct(h_, cap)); }

	chunk_carry scan_chunk(const std::uint8_t* d_chunk, std::size_t n, chunk_carry in, hipStream_t s = nullptr,
						   std::size_t byte_base = 0) const {
		chunk_carry out {};
		check(pars_hip_scan_chunk(h_, d_chunk, n, &in, s, byte_base, &out));
		return out;
	}

	const std::int32_t* struct_ptr() const { return reinterpret_cast<const std::int32_t*>(pars_hip_struct_ptr(h_)); }

	// R1-R6 and the JSON tape of a device-resident document: what the
	// reference's pars_parse (bench_parse.cu:336-348) computes on the host by
	// replaying the structural index into its JSON DOM hook.
	json_tape json_tape_device(const std::uint8_t* d_data, std::size_t n, hipStream_t s = nullptr,
								stage_times* st = nullptr) const {
		json_tape t {};
		check(pars_hip_json_tape(h_, d_data, n, s, st, &t));
		return t;
	}

	// R1-R5 and the key-value table of a device-resident INI or TOML document:
	// the reference's full-parse hooks (bench_parse_host.cu:69-134, 202-248)
	// built on the device instead of replayed on the host.
	kv_table kv_parse_device(const std::uint8_t* d_data, std::size_t n, pars_kv_format format,
							 hipStream_t s = nullptr, stage_times* st = nullptr) const {
		kv_table t {};
		check(pars_hip_kv_parse(h_, d_data, n, format, s, st, &t));
		return t;
	}

	pars_scanner* handle() const noexcept { return h_; }

private:
	pars_scanner* h_ = nullptr;
};

// gpu::scanner<S> for a compile-time plan S (a pars::scan_plan or pars_plan).
template<auto S>
class scanner : public runtime_scanner {
	static_assert(sizeof(S) == PARS_PLAN_BYTES, "S must be a scan_plan (3177 bytes)");

public:
	scanner() : runtime_scanner(*reinterpret_cast<const pars_plan*>(&S)) {}
};

template<auto S>
device_result scan_device(const std::uint8_t* d_data, std::size_t n, hipStream_t s = nullptr,
						  bool want_pairs = false) {
	static thread_local scanner<S> sc;
	return sc.scan_device(d_data, n, s, want_pairs);
}

template<auto S>
host_result scan_host(const std::uint8_t* h_data, std::size_t n, bool want_pairs = false) {
	static thread_local scanner<S> sc;
	return sc.scan_host(h_data, n, nullptr, want_pairs);
}

template<auto S>
chunk_carry scan_chunk(const std::uint8_t* d_chunk, std::size_t n, chunk_carry in, hipStream_t s = nullptr,
					   std::size_t byte_base = 0) {
	static thread_local scanner<S> sc;
	return sc.scan_chunk(d_chunk, n, in, s, byte_base);
}

// ---------------------------------------------------------------------------
// Consumer side (host): the structural index of a scan_host result handed to
// user code, as the reference's model::replay (include/pars/model/visit.hpp:
// 28-48) and model::structural_cursor (include/pars/model/cursor.hpp:16-56)
// hand it.  Format-agnostic: byte offsets and bracket partners only.
// ---------------------------------------------------------------------------

// visit.hpp:28-32: offsets of the structural bytes, and per structural the
// index of its partner bracket (-1: none), or no partners at all.
struct structural_view {
	const std::int32_t* pos	 = nullptr;
	const std::int32_t* pair = nullptr;
	std::size_t			n	 = 0;
};

// A view of a scan_host result (pairs when it was asked for them).
inline structural_view view_of(const host_result& r) {
	return {r.structural.data(), r.pair_pos.empty() ? nullptr : r.pair_pos.data(), r.structural.size()};
}

// visit.hpp:35-48: one pass over the n input bytes as a covering,
// non-overlapping event stream - v.scalar(begin, end) for every non-empty run
// of bytes between structural bytes (and before the first / after the last),
// v.symbol(offset, byte, partner offset or -1) for every structural byte.
template<class Visitor>
void replay(const std::uint8_t* base, std::size_t n, const structural_view& sv, Visitor&& v) {
	std::size_t next = 0;  // first byte not yet reported
	for (std::size_t i = 0; i < sv.n; ++i) {
		const auto at = static_cast<std::size_t>(sv.pos[i]);
		if (at > next)
			v.scalar(next, at);
		std::int32_t partner = -1;
		if (sv.pair) {
			const std::int32_t j = sv.pair[i];
			if (j >= 0)
				partner = sv.pos[j];
		}
		v.symbol(at, static_cast<char>(base[at]), partner);
		next = at + 1;
	}
	if (n > next)
		v.scalar(next, n);
}

// cursor.hpp:16-56: random access to the index - offset, byte and partner of
// structural i, and the run of bytes just before it (its gap).
struct structural_cursor {
	const std::int32_t* pos	 = nullptr;
	const std::int32_t* pair = nullptr;
	const std::uint8_t* base = nullptr;
	std::size_t			n	 = 0;

	bool		 valid(std::size_t i) const noexcept { return i < n; }
	std::int32_t off(std::size_t i) const noexcept { return pos[i]; }
	char		 ch(std::size_t i) const noexcept { return static_cast<char>(base[pos[i]]); }
	std::int32_t match_index(std::size_t i) const noexcept { return pair ? pair[i] : -1; }
	std::size_t	 gap_begin(std::size_t i) const noexcept { return i == 0 ? 0 : static_cast<std::size_t>(pos[i - 1]) + 1; }
	std::size_t	 gap_end(std::size_t i) const noexcept { return static_cast<std::size_t>(pos[i]); }

	std::string_view gap(std::size_t i) const noexcept {
		const std::size_t b = gap_begin(i);
		return std::string_view(reinterpret_cast<const char*>(base) + b, gap_end(i) - b);
	}
};

// cursor.hpp:58-66: ASCII whitespace (space, tab, newline, carriage return)
// removed from both ends.
inline std::string_view trim(std::string_view s) noexcept {
	auto ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
	std::size_t b = 0, e = s.size();
	while (b < e && ws(s[b])) ++b;
	while (e > b && ws(s[e - 1])) --e;
	return s.substr(b, e - b);
}

}  // namespace pars::hip
