// tests/cpp/replay_check.cpp — the consumer side of the C++ facade
// (pars::hip::replay, structural_cursor, trim in include/pars_hip.hpp) over a
// given structural index, host only.
//
//   replay_check <input> <positions.i32> [<pairs.i32>]
//
// Prints the replay's event stream, one event a line ("S begin end" for a
// scalar run, "Y offset byte partner" for a structural byte), then per
// structural i the cursor's view: "C i off byte match gap_begin gap_end
// trimmed_begin trimmed_length".
#include "pars_hip.hpp"

#include <cstdio>
#include <vector>

namespace {

template<class T>
std::vector<T> slurp(const char* path) {
	std::vector<T> v;
	std::FILE*	   f = std::fopen(path, "rb");
	if (!f)
		return v;
	T x;
	while (std::fread(&x, sizeof x, 1, f) == 1) v.push_back(x);
	std::fclose(f);
	return v;
}

struct printer {
	void scalar(std::size_t b, std::size_t e) { std::printf("S %zu %zu\n", b, e); }
	void symbol(std::size_t p, char c, std::int32_t m) { std::printf("Y %zu %d %d\n", p, (int)(unsigned char)c, m); }
};

}  // namespace

int main(int argc, char** argv) {
	if (argc < 3)
		return 2;
	const auto in	= slurp<std::uint8_t>(argv[1]);
	const auto pos	= slurp<std::int32_t>(argv[2]);
	const auto pair = argc > 3 ? slurp<std::int32_t>(argv[3]) : std::vector<std::int32_t> {};
	pars::hip::structural_view sv {pos.data(), pair.empty() ? nullptr : pair.data(), pos.size()};
	printer					   p;
	pars::hip::replay(in.data(), in.size(), sv, p);
	const pars::hip::structural_cursor c {sv.pos, sv.pair, in.data(), sv.n};
	for (std::size_t i = 0; c.valid(i); ++i) {
		const std::string_view g = c.gap(i), t = pars::hip::trim(g);
		std::printf("C %zu %d %d %d %zu %zu %zu %zu\n", i, c.off(i), (int)(unsigned char)c.ch(i), c.match_index(i),
					c.gap_begin(i), c.gap_end(i), c.gap_begin(i) + (t.data() - g.data()),
					t.size());
	}
	return c.valid(sv.n) ? 1 : 0;
}
