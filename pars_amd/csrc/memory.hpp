// memory.hpp — where scanners and servers get their device memory: the
// caller's allocator (or hipMalloc), charged against an optional byte budget
// that several scanners and servers can share.
//
// Reference: the CUDA backend's DeviceCache (include/pars/backend/cuda/
// pipeline.cuh:562-714) grows every buffer with cudaMalloc and exits the
// process when an allocation fails (:520-555).  Here a failed or refused
// allocation returns PARS_ERR_MEMORY and the scanner stays usable: every
// buffer tracks its own capacity and simply grows again on the next call.
#pragma once

#include "../../include/pars_hip.h"

#include <atomic>
#include <stdint.h>

// A limit on the device bytes its users hold at once.  Charges are taken
// before an allocation and returned after its free; a charge that would pass
// the limit is refused.  Scanner buffers are charged while they exist; a
// server charges its rings and, per batch it launches, the worst-case
// footprint of the batch's requests until they are released.
struct pars_budget {
	uint64_t			  limit = 0;
	std::atomic<uint64_t> in_use {0}, peak {0};
	std::atomic<uint32_t> users {0};  // scanners and servers drawing on it

	bool charge(uint64_t b) {
		uint64_t cur = in_use.load(std::memory_order_relaxed);
		do {
			if (b > limit || cur > limit - b)
				return false;
		} while (!in_use.compare_exchange_weak(cur, cur + b, std::memory_order_relaxed));
		uint64_t p = peak.load(std::memory_order_relaxed);
		while (cur + b > p && !peak.compare_exchange_weak(p, cur + b, std::memory_order_relaxed)) {
		}
		return true;
	}

	void uncharge(uint64_t b) { in_use.fetch_sub(b, std::memory_order_relaxed); }
};

namespace pars_hip {

// A scanner's memory source: the caller's allocator when one was given
// (alloc.alloc set), else hipMalloc / hipFree; the budget when one is attached.
struct MemSource {
	pars_allocator alloc {};
	pars_budget*   budget = nullptr;

	// `bytes` of device memory into *p, charged to the budget; on failure *p is
	// null and the status (PARS_ERR_MEMORY) carries the reason.
	int	 get(void** p, size_t bytes);
	void put(void* p, size_t bytes);
};

}  // namespace pars_hip
