// ring.hpp — the byte ring the serving runtime stages requests in (pinned host
// staging and the device input ring).  Plain C++ with no HIP dependency, so
// its allocation and release order is checked on the host.
#pragma once

#include <cstdint>
#include <deque>

namespace pars_hip {

// A byte ring: allocations are contiguous and placed in allocation order; live
// data spans [tail, head), possibly wrapped.  A region can be released out of
// order; its space is reused once every older region is released too (the tail
// only moves past a released prefix).
struct Ring {
	struct Region {
		uint64_t off;
		bool	 released;
	};
	uint64_t		   cap = 0, head = 0, tail = 0;
	uint64_t		   wraps = 0;  // allocations placed at 0 while older data was still live past them
	std::deque<Region> live;	   // in allocation order

	bool alloc(uint64_t bytes, uint64_t& off) {
		if (bytes == 0 || bytes > cap)
			return false;
		if (live.empty())
			head = tail = 0;
		uint64_t at;
		if (live.empty() || head > tail) {	// free: [head, cap) and [0, tail)
			if (head + bytes <= cap)
				at = head;
			else if (bytes <= tail) {
				at = 0;	 // wrap; [head, cap) idles until the tail passes it
				++wraps;
			} else
				return false;
		} else if (head < tail && head + bytes <= tail) {  // free: [head, tail)
			at = head;
		} else {
			return false;
		}
		off	 = at;
		head = at + bytes;
		live.push_back({at, false});
		return true;
	}

	// The oldest live region.
	void release_oldest() {
		if (live.empty())
			return;
		live.front().released = true;
		collapse();
	}

	// The live region that starts at `off`; false if there is none.
	bool release(uint64_t off) {
		for (Region& g : live)
			if (g.off == off && !g.released) {
				g.released = true;
				collapse();
				return true;
			}
		return false;
	}

private:
	void collapse() {
		while (!live.empty() && live.front().released) live.pop_front();
		if (live.empty())
			head = tail = 0;
		else
			tail = live.front().off;
	}
};

}  // namespace pars_hip
