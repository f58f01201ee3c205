// tests/cpp/facade_check.cpp — the C++ facade (include/pars_hip.hpp) used the
// way the reference's callers use gpu::scanner<S>: a compile-time plan as a
// template argument, and a runtime plan loaded from a compiled plan file.
//
//   facade_check <plan file> <input file>   prints "n_struct balanced well_formed max_depth"
//                                           then the structural positions, one per line
//   facade_check <plan file> <input file> tape
//                                           prints "n_nodes ok fail_event arena_bytes" of the
//                                           JSON tape, then "kind begin end" per node
#include "pars_hip.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace {

// A hand-built compile-time plan: punctuation ',' ':', brackets [] {}, no
// lexical descriptors (no masking).
constexpr pars_plan make_tiny() {
	pars_plan p {};
	p.punctuation[0] = ',';
	p.punctuation[1] = ':';
	p.n_punctuation	 = 2;
	p.brackets[0]	 = {'[', ']'};
	p.brackets[1]	 = {'{', '}'};
	p.n_brackets	 = 2;
	return p;
}

constexpr pars_plan kTiny = make_tiny();

std::vector<unsigned char> slurp(const char* path) {
	std::vector<unsigned char> v;
	std::FILE*				   f = std::fopen(path, "rb");
	if (!f)
		return v;
	unsigned char buf[1 << 14];
	std::size_t	  r;
	while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) v.insert(v.end(), buf, buf + r);
	std::fclose(f);
	return v;
}

}  // namespace

int main(int argc, char** argv) {
	if (argc < 3) {
		std::fprintf(stderr, "usage: facade_check <plan> <input>\n");
		return 2;
	}
	const auto plan_bytes = slurp(argv[1]);
	const auto input	  = slurp(argv[2]);
	if (plan_bytes.size() != PARS_PLAN_BYTES)
		return 2;
	try {
		// template API over a compile-time plan (gpu::scanner<S> style)
		pars::hip::scanner<kTiny> tiny;
		const unsigned char		  probe[] = "[1,{\"a\":2}]";
		const auto				  t		  = tiny.scan_host(probe, sizeof probe - 1);
		if (t.structural.size() != 6 || !t.balanced || t.max_depth != 2) {
			std::fprintf(stderr, "tiny plan gave %zu structural\n", t.structural.size());
			return 1;
		}
		// runtime plan (the reference compiler's output, loaded from a file)
		pars_plan p;
		std::memcpy(&p, plan_bytes.data(), sizeof p);
		pars::hip::runtime_scanner sc(p);
		if (argc > 3 && std::strcmp(argv[3], "tape") == 0) {
			void* d = nullptr;
			if (hipMalloc(&d, input.size() + 1) != hipSuccess ||
				hipMemcpy(d, input.data(), input.size(), hipMemcpyHostToDevice) != hipSuccess)
				return 1;
			const auto t = sc.json_tape_device(static_cast<const std::uint8_t*>(d), input.size());
			std::vector<std::uint8_t>  kind(t.n_nodes);
			std::vector<std::uint32_t> b(t.n_nodes), e(t.n_nodes);
			if (t.n_nodes && (hipMemcpy(kind.data(), t.kind, t.n_nodes, hipMemcpyDeviceToHost) != hipSuccess ||
							  hipMemcpy(b.data(), t.begin, 4 * t.n_nodes, hipMemcpyDeviceToHost) != hipSuccess ||
							  hipMemcpy(e.data(), t.end, 4 * t.n_nodes, hipMemcpyDeviceToHost) != hipSuccess))
				return 1;
			(void)hipFree(d);
			std::printf("%llu %d %lld %llu\n", (unsigned long long)t.n_nodes, t.ok, (long long)t.fail_event,
						(unsigned long long)t.arena_bytes);
			for (std::size_t k = 0; k < t.n_nodes; ++k) std::printf("%u %u %u\n", kind[k], b[k], e[k]);
			return 0;
		}
		const auto				   r = sc.scan_host(input.data(), input.size(), nullptr, true);
		std::printf("%zu %d %d %d\n", r.structural.size(), r.balanced ? 1 : 0, r.well_formed ? 1 : 0, r.max_depth);
		for (std::int32_t v : r.structural) std::printf("%d\n", v);
	} catch (const pars::hip::error& e) {
		std::fprintf(stderr, "pars error %d: %s\n", e.status(), e.what());
		return 1;
	}
	return 0;
}
