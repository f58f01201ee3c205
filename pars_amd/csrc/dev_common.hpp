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

	__device__ static Map4 shfl_down(Map4 m, int off) {
		return Map4 {(uint32_t)__shfl_down((int)m.lo, off, kWave)};
	}

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

	__device__ static Map8 shfl_down(Map8 m, int off) {
		return Map8 {(uint32_t)__shfl_down((int)m.lo, off, kWave), (uint32_t)__shfl_down((int)m.hi, off, kWave)};
	}

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

template<class T>
__device__ __forceinline__ T wave_incl_sum(T v) {
	const int l = lane_id();
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		T o = __shfl_up(v, off, kWave);
		if (l >= off)
			v += o;
	}
	return v;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
	return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
	return v;
}

// Inclusive wave scan of maps in lane order: lane i ends with
// m_i . m_{i-1} . ... . m_0 (the earliest lane is applied first).
template<bool WIDE>
__device__ __forceinline__ typename MapT<WIDE>::type wave_scan_maps(typename MapT<WIDE>::type m) {
	using M		= MapT<WIDE>;
	const int l = lane_id();
#pragma unroll
	for (int off = 1; off < kWave; off <<= 1) {
		auto o = M::shfl_up(m, off);
		if (l >= off)
			m = M::compose(m, o);
	}
	return m;
}

}  // namespace pars_hip
