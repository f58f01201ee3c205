// dev_common.hpp — device building blocks shared by the scan, pairing and batch
// kernels: byte-map composition with v_perm, wave scans, and self-validating
// look-back words.
//
// Written for gfx950 (CDNA4): 64-lane wavefronts, v_perm_b32 byte permutes,
// relaxed agent-scope atomics (plain `sc1` loads/stores; acquire/release would
// emit an L2 write-back/invalidate on this multi-XCD part, so look-back never
// relies on them: every published value carries its own tag in one 64-bit word).
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

namespace pars_hip {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// State maps.  A descriptor with <= 4 states keeps its map s -> m[s] in one
// dword (byte s holds m[s]); one with <= 8 states in two (lo: states 0-3, hi:
// states 4-7).  Composition is a byte permute: (outer . inner)[s] = outer[inner[s]]
// = v_perm({outer.hi, outer.lo}, inner).  Selector bytes are always < 8.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
	return __builtin_amdgcn_perm(hi, lo, sel);
}

// DPP lane moves (GFX9 encodings): row_shr:n = 0x110 + n (within 16-lane
// rows), wave_shr:1 = 0x138, row_bcast:15 = 0x142 (lane 15 of each row to the
// next row), row_bcast:31 = 0x143 (lane 31 to rows 2 and 3).  Lanes whose source
// is outside the row, or whose row is masked off, keep `old`.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShr1 = 0x138, kDppBcast15 = 0x142, kDppBcast31 = 0x143;

template<int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t src) {
	return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, ROW_MASK, 0xF, false);
}

__device__ __forceinline__ uint32_t readlane(uint32_t v, int lane) {
	return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// A value known to be the same in every active lane, moved to a scalar register.
__device__ __forceinline__ uint32_t readfirstlane(uint32_t v) {
	return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

struct Map4 {
	uint32_t lo;
};

struct Map8 {
	uint32_t lo, hi;
};

constexpr uint32_t kIdLo = 0x03020100u;
constexpr uint32_t kIdHi = 0x07060504u;

template<bool WIDE>
struct MapT;

template<>
struct MapT<false> {
	using type = Map4;

	__device__ static Map4 id() { return Map4 {kIdLo}; }

	__device__ static Map4 compose(Map4 outer, Map4 inner) { return Map4 {perm(0u, outer.lo, inner.lo)}; }

	__device__ static uint32_t apply(Map4 m, uint32_t s) { return (m.lo >> (8u * s)) & 0xFFu; }

	__device__ static Map4 shfl_up(Map4 m, int off) { return Map4 {(uint32_t)__shfl_up((int)m.lo, off, kWave)}; }

	__device__ static Map4 shfl(Map4 m, int src) { return Map4 {(uint32_t)__shfl((int)m.lo, src, kWave)}; }

	template<int CTRL, int ROW_MASK = 0xF>
	__device__ static Map4 dpp(Map4 m) {
		return Map4 {pars_hip::dpp<CTRL, ROW_MASK>(kIdLo, m.lo)};
	}

	__device__ static Map4 readlane(Map4 m, int lane) { return Map4 {pars_hip::readlane(m.lo, lane)}; }

	// every state -> v
	__device__ static Map4 constant(uint32_t v) { return Map4 {v * 0x01010101u}; }

	// 3-bit packed form used in look-back words (24 bits for 8 states)
	__device__ static uint32_t pack(Map4 m) {
		uint32_t r = 0;
#pragma unroll
		for (int s = 0; s < 4; ++s) r |= ((m.lo >> (8 * s)) & 7u) << (3 * s);
		r |= (4u << 12) | (5u << 15) | (6u << 18) | (7u << 21);
		return r;
	}

	__device__ static Map4 unpack(uint32_t p) {
		uint32_t lo = 0;
#pragma unroll
		for (int s = 0; s < 4; ++s) lo |= ((p >> (3 * s)) & 7u) << (8 * s);
		return Map4 {lo};
	}
};

template<>
struct MapT<true> {
	using type = Map8;

	__device__ static Map8 id() { return Map8 {kIdLo, kIdHi}; }

	__device__ static Map8 compose(Map8 outer, Map8 inner) {
		return Map8 {perm(outer.hi, outer.lo, inner.lo), perm(outer.hi, outer.lo, inner.hi)};
	}

	__device__ static uint32_t apply(Map8 m, uint32_t s) {
		const uint32_t w = s < 4 ? m.lo : m.hi;
		return (w >> (8u * (s & 3u))) & 0xFFu;
	}

	__device__ static Map8 shfl_up(Map8 m, int off) {
		return Map8 {(uint32_t)__shfl_up((int)m.lo, off, kWave), (uint32_t)__shfl_up((int)m.hi, off, kWave)};
	}

	__device__ static Map8 shfl(Map8 m, int src) {
		return Map8 {(uint32_t)__shfl((int)m.lo, src, kWave), (uint32_t)__shfl((int)m.hi, src, kWave)};
	}

	template<int CTRL, int ROW_MASK = 0xF>
	__device__ static Map8 dpp(Map8 m) {
		return Map8 {pars_hip::dpp<CTRL, ROW_MASK>(kIdLo, m.lo), pars_hip::dpp<CTRL, ROW_MASK>(kIdHi, m.hi)};
	}

	__device__ static Map8 readlane(Map8 m, int lane) {
		return Map8 {pars_hip::readlane(m.lo, lane), pars_hip::readlane(m.hi, lane)};
	}

	__device__ static Map8 constant(uint32_t v) { return Map8 {v * 0x01010101u, v * 0x01010101u}; }

	__device__ static uint32_t pack(Map8 m) {
		uint32_t r = 0;
#pragma unroll
		for (int s = 0; s < 4; ++s) {
			r |= ((m.lo >> (8 * s)) & 7u) << (3 * s);
			r |= ((m.hi >> (8 * s)) & 7u) << (3 * (s + 4));
		}
		return r;
	}

	__device__ static Map8 unpack(uint32_t p) {
		uint32_t lo = 0, hi = 0;
#pragma unroll
		for (int s = 0; s < 4; ++s) {
			lo |= ((p >> (3 * s)) & 7u) << (8 * s);
			hi |= ((p >> (3 * (s + 4))) & 7u) << (8 * s);
		}
		return Map8 {lo, hi};
	}
};

// ---------------------------------------------------------------------------
// Tagged look-back words: high 32 bits = (epoch << 2) | status, low 32 bits =
// value.  status 1 = tile aggregate, 2 = inclusive prefix.  A word from an
// older launch carries an older epoch and reads as "not ready", so records
// never need clearing between launches.
// ---------------------------------------------------------------------------
constexpr uint32_t kAgg	 = 1u;
constexpr uint32_t kIncl = 2u;

__device__ __forceinline__ uint64_t make_word(uint32_t epoch, uint32_t status, uint32_t value) {
	return ((uint64_t)((epoch << 2) | status) << 32) | value;
}

__device__ __forceinline__ void publish(uint64_t* p, uint64_t w) {
	__hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t peek(const uint64_t* p) {
	return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool word_is(uint64_t w, uint32_t epoch, uint32_t status) {
	return (uint32_t)(w >> 32) == ((epoch << 2) | status);
}

// ---------------------------------------------------------------------------
// Wave-level helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() {
	return __lane_id();
}

__device__ __forceinline__ uint64_t lanemask_lt() {
	const int l = lane_id();
	return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave scans on DPP: four row_shr steps scan each 16-lane row, two
// row broadcasts carry rows 0..2 into the rows above.  `old` of every move is
// the operation's identity, so lanes without a source combine with it.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
	v += dpp<kDppRowShr1>(0u, v);
	v += dpp<kDppRowShr2>(0u, v);
	v += dpp<kDppRowShr4>(0u, v);
	v += dpp<kDppRowShr8>(0u, v);
	v += dpp<kDppBcast15, 0xA>(0u, v);
	v += dpp<kDppBcast31, 0xC>(0u, v);
	return v;
}

// Sum over the wave, the same value in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
	return readlane(wave_incl_sum(v), kWave - 1);
}

__device__ __forceinline__ int wave_min(int v) {
	uint32_t u = (uint32_t)v;
	u = (uint32_t)min((int)u, (int)dpp<kDppRowShr1>((uint32_t)INT_MAX, u));
	u = (uint32_t)min((int)u, (int)dpp<kDppRowShr2>((uint32_t)INT_MAX, u));
	u = (uint32_t)min((int)u, (int)dpp<kDppRowShr4>((uint32_t)INT_MAX, u));
	u = (uint32_t)min((int)u, (int)dpp<kDppRowShr8>((uint32_t)INT_MAX, u));
	u = (uint32_t)min((int)u, (int)dpp<kDppBcast15, 0xA>((uint32_t)INT_MAX, u));
	u = (uint32_t)min((int)u, (int)dpp<kDppBcast31, 0xC>((uint32_t)INT_MAX, u));
	return (int)readlane(u, kWave - 1);
}

__device__ __forceinline__ int wave_max(int v) {
	uint32_t u = (uint32_t)v;
	u = (uint32_t)max((int)u, (int)dpp<kDppRowShr1>((uint32_t)INT_MIN, u));
	u = (uint32_t)max((int)u, (int)dpp<kDppRowShr2>((uint32_t)INT_MIN, u));
	u = (uint32_t)max((int)u, (int)dpp<kDppRowShr4>((uint32_t)INT_MIN, u));
	u = (uint32_t)max((int)u, (int)dpp<kDppRowShr8>((uint32_t)INT_MIN, u));
	u = (uint32_t)max((int)u, (int)dpp<kDppBcast15, 0xA>((uint32_t)INT_MIN, u));
	u = (uint32_t)max((int)u, (int)dpp<kDppBcast31, 0xC>((uint32_t)INT_MIN, u));
	return (int)readlane(u, kWave - 1);
}

// Inclusive wave scan of maps in lane order: lane i ends with
// m_i . m_{i-1} . ... . m_0 (the lowest lane is applied first).
template<bool WIDE>
__device__ __forceinline__ typename MapT<WIDE>::type wave_scan_maps(typename MapT<WIDE>::type m) {
	using M = MapT<WIDE>;
	m		= M::compose(m, M::template dpp<kDppRowShr1>(m));
	m		= M::compose(m, M::template dpp<kDppRowShr2>(m));
	m		= M::compose(m, M::template dpp<kDppRowShr4>(m));
	m		= M::compose(m, M::template dpp<kDppRowShr8>(m));
	m		= M::compose(m, M::template dpp<kDppBcast15, 0xA>(m));
	m		= M::compose(m, M::template dpp<kDppBcast31, 0xC>(m));
	return m;
}

// The previous lane's value (identity in lane 0).
template<bool WIDE>
__device__ __forceinline__ typename MapT<WIDE>::type wave_shift_maps(typename MapT<WIDE>::type m) {
	return MapT<WIDE>::template dpp<kDppWaveShr1>(m);
}

}  // namespace pars_hip
