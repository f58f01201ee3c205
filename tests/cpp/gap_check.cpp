// tests/cpp/gap_check.cpp — the JSON tape's gap parser (parse_gap in
// pars_amd/csrc/dom_kernels.hip, the device code) built for the host, so its
// automaton can be checked against the oracle on the CPU.
//
//   gap_check <gaps file>   gaps as (u32 length, bytes) records; prints per gap
//                           "kind ok begin end decoded_len decoded_hex" with
//                           begin/end relative to the gap (end -1: none)
#include "../../pars_amd/csrc/dom_kernels.hip"

#include <cstdio>
#include <cstring>
#include <vector>

namespace pars_hip {
// the tape launcher references the fill kernel of pair_kernels.hip; nothing
// here launches a kernel
hipError_t launch_fill_i32(int32_t*, uint64_t, int32_t, hipStream_t) {
	return hipSuccess;
}
}  // namespace pars_hip

int main(int argc, char** argv) {
	using namespace pars_hip;
	if (argc < 2)
		return 2;
	std::FILE* f = std::fopen(argv[1], "rb");
	if (!f)
		return 2;
	std::vector<unsigned char> in;
	unsigned char			   tmp[1 << 16];
	size_t					   r;
	while ((r = std::fread(tmp, 1, sizeof tmp, f)) > 0) in.insert(in.end(), tmp, tmp + r);
	std::fclose(f);
	static const GapTables	   T = make_gap_tables();
	std::vector<unsigned char> out(1 << 20);
	size_t					   off = 0;
	while (off + 4 <= in.size()) {
		uint32_t len;
		std::memcpy(&len, &in[off], 4);
		off += 4;
		// the gap at a varying alignment inside a larger buffer, as in a document
		const size_t			   shift = off % 16;
		std::vector<unsigned char> buf(len + 64, 'Z');
		std::memcpy(&buf[shift], &in[off], len);
		const uint32_t b = (uint32_t)shift, e = (uint32_t)(shift + len);
		Bytes		   rd(buf.data(), buf.size(), b), rd2(buf.data(), buf.size(), b);
		const Gap	   g  = parse_gap<false>(T, rd, b, e, nullptr);
		const Gap	   g2 = parse_gap<true>(T, rd2, b, e, out.data());
		if (g.dec != g2.dec || g.kind != g2.kind || g.end != g2.end || g.ok != g2.ok)
			return 3;
		// the gap alone in a buffer of its own size at an odd address: every
		// read near the buffer's edges goes through the byte path
		std::vector<unsigned char> tight(len + 1), out3(g2.dec + 1);
		std::memcpy(&tight[1], &in[off], len);
		Bytes	  rd3(tight.data() + 1, len, 0u);
		const Gap g3 = parse_gap<true>(T, rd3, 0, len, out3.data());
		if (g3.dec != g2.dec || g3.kind != g2.kind || g3.ok != g2.ok || (g3.end ? g3.end + shift : 0) != g2.end ||
			(g3.kind && g3.beg + shift != g2.beg) || std::memcmp(out3.data(), out.data(), g2.dec) != 0)
			return 4;
		std::printf("%u %u %lld %lld %u ", g.kind, g.ok, g.kind ? (long long)g.beg - (long long)shift : 0LL,
					g.end ? (long long)g.end - (long long)shift : -1LL, g.dec);
		for (uint32_t q = 0; q < g2.dec; ++q) std::printf("%02x", out[q]);
		std::printf("\n");
		off += len;
	}
	return 0;
}
