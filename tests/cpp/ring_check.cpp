// ring_check — host check of the serving runtime's byte ring
// (pars_amd/csrc/ring.hpp) under the order the server uses it: every batch,
// the ones that fail to enqueue included, is retired in launch order and frees
// its own region then.  A shadow list of live regions asserts that no
// allocation ever overlaps a region whose batch has not been retired, through
// many wraps, with random sizes and random failures; regions released out of
// order (a batch's input region is held until its slot frees) likewise.  Prints "ok" and the
// number of wraps, or the first violation.
#include "../../pars_amd/csrc/ring.hpp"

#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>

using pars_hip::Ring;

struct Region {
	uint64_t off, len;
	bool	 failed;
};

static bool overlaps(const Region& a, uint64_t off, uint64_t len) {
	return off < a.off + a.len && a.off < off + len;
}

int main() {
	// the sequence the advisor named: alloc A, alloc B, B fails, retire A, retire B
	{
		Ring r;
		r.cap = 1000;
		uint64_t a = 0, b = 0, c = 0;
		if (!r.alloc(400, a) || !r.alloc(400, b)) {
			std::puts("FAIL: setup");
			return 1;
		}
		std::deque<Region> running = {{a, 400, false}, {b, 400, true}};
		r.release_oldest();	 // A retires first (B failed, but waits its turn)
		running.pop_front();
		if (r.alloc(300, c) && overlaps(running.front(), c, 300)) {
			std::puts("FAIL: an allocation overlaps the failed batch's live region");
			return 1;
		}
		r.release_oldest();	 // B
		running.pop_front();
		if (!r.live.empty() && r.live.size() != 1) {
			std::puts("FAIL: live regions after retiring both");
			return 1;
		}
	}
	std::mt19937_64 rng(12345);
	for (int trial = 0; trial < 200; ++trial) {
		Ring r;
		r.cap = 1 + rng() % 5000;
		std::deque<Region> running;	 // launch order
		for (int step = 0; step < 5000; ++step) {
			const bool retire = !running.empty() && (rng() % 3 == 0 || running.size() > 8);
			if (retire) {
				if (rng() % 2) {  // the oldest (staging: retirement order)
					r.release_oldest();
					running.pop_front();
				} else {  // any live region (device inputs: slots free in any order)
					const size_t k = rng() % running.size();
					if (!r.release(running[k].off)) {
						std::printf("FAIL: live region at %llu not found\n", (unsigned long long)running[k].off);
						return 1;
					}
					running.erase(running.begin() + (long)k);
				}
				continue;
			}
			const uint64_t len = 1 + rng() % (r.cap / 2 + 1);
			uint64_t	   off = 0;
			if (!r.alloc(len, off))
				continue;
			if (off + len > r.cap) {
				std::printf("FAIL: region [%llu, +%llu) past the capacity %llu\n", (unsigned long long)off,
							(unsigned long long)len, (unsigned long long)r.cap);
				return 1;
			}
			for (const Region& g : running)
				if (overlaps(g, off, len)) {
					std::printf("FAIL: trial %d step %d: [%llu, +%llu) overlaps live [%llu, +%llu)%s\n", trial, step,
								(unsigned long long)off, (unsigned long long)len, (unsigned long long)g.off,
								(unsigned long long)g.len, g.failed ? " of a failed batch" : "");
					return 1;
				}
			running.push_back({off, len, rng() % 7 == 0});
		}
	}
	Ring w;
	w.cap = 100;
	uint64_t o;
	w.alloc(60, o);
	w.alloc(30, o);
	w.release_oldest();
	w.alloc(50, o);	 // must wrap to 0 past the live [60, 90)
	std::printf("ok wraps=%llu first_wrap_off=%llu\n", (unsigned long long)w.wraps, (unsigned long long)o);
	return (w.wraps == 1 && o == 0) ? 0 : 1;
}
