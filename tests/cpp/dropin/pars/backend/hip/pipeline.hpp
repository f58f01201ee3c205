// include/pars/backend/hip/pipeline.hpp  (in the pars tree)
#pragma once
#include "pars/plan.hpp"
#include <pars_hip.hpp>  // this repository's include/

namespace pars::gpu {
using pars::hip::chunk_carry;
using pars::hip::device_result;
using pars::hip::host_result;
using pars::hip::stage_times;

template<scan_plan S>
using scanner = pars::hip::scanner<S>;

template<scan_plan S>
device_result scan_device(const std::uint8_t* d, std::size_t n, hipStream_t s = nullptr, bool want_pairs = false) {
	return pars::hip::scan_device<S>(d, n, s, want_pairs);
}

template<scan_plan S>
host_result scan_host(const std::uint8_t* h, std::size_t n, bool want_pairs = false) {
	return pars::hip::scan_host<S>(h, n, want_pairs);
}

template<scan_plan S>
chunk_carry scan_chunk(const std::uint8_t* d, std::size_t n, chunk_carry in, hipStream_t s = nullptr,
					   std::size_t byte_base = 0) {
	return pars::hip::scan_chunk<S>(d, n, in, s, byte_base);
}
}  // namespace pars::gpu
