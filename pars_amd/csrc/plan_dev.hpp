// plan_dev.hpp — a pars_plan lowered to the tables the kernels read.
//
// The reference rebuilds its GPU tables on the host on every call
// (include/pars/backend/cuda/pipeline.cuh:51-85 make_gpu_spec, :132-170
// build_gpu_plan) and packs all descriptors into one 16-slot transform, which
// caps the plan at 16 states in total (:758-762).  Here the plan is lowered
// once per scanner, and every descriptor keeps its own byte map, so any plan
// within the plan.hpp caps (8 descriptors x 8 states x 8 classes) runs.
#pragma once

#include "../../include/pars_plan.h"

#include <stdint.h>

namespace pars_hip {

constexpr int kMaxDescs = 8;

// Per-byte class bits, pre-shifted by the byte's position within a group of 8
// so a lane ORs eight lookups into one dword: byte 0 = structural, byte 1 =
// open|close, byte 2 = open.  cls[j][c] = (S(c) << j) | (OC(c) << 8+j) | (O(c) << 16+j).
struct DevPlan {
	uint32_t cls[8][256];
	// The same with two more class bytes for the escape fast path: entry [0] =
	// bytes {structural, open|close, open, delimiter}, [1] = byte {escape}.
	uint32_t cls8[8][256][2];
	// Per descriptor d and byte c, the step taken on c, as maps over the 8 state
	// slots: {next-state map lo (states 0-3), hi (4-7), masked-bit map lo, hi}.
	// Unused slots map to themselves and never mask.
	uint32_t dt[kMaxDescs][256][4];
	uint8_t	 start[kMaxDescs];
	uint8_t	 n_states[kMaxDescs];
	uint8_t	 br_flags[256];	   // bit0 open, bit1 close (nest.hpp:26-40)
	uint8_t	 br_close_of[256]; // for an open byte: the close of the first spec opening with it
	uint32_t start_packed;	   // start states, 3 bits per descriptor
	uint8_t	 n_descs;
	uint8_t	 wide;			   // some descriptor has more than 4 states
	uint8_t	 fast_kind;		   // a specialised evaluator applies (0 = none)
	uint8_t	 pad;
	// affine-with-escape / parity fast paths: the descriptor's class of each
	// event byte and its state numbering (filled when fast_kind != 0)
	uint8_t fk_delim, fk_escape;
	uint8_t fk_state[4];  // escape: out, in, closed, esc; parity: out, in, closed, closed
	uint8_t fk_doubled;	  // parity: a delimiter right after a closing one is masked (CSV "")
	// Bracket bytes without re-reading the input (bit-parallel evaluators):
	// byte 1 of cls8[j][c][1] carries bit j = the ordinal of bracket byte c
	// among the bytes that open (c opens) or among the close-only bytes, and
	// br_chr[opens][ordinal] gives c back.  Only when each set has at most two
	// bytes (br_bit_ok); otherwise bracket bytes are read from the input.
	uint8_t br_bit_ok;
	uint8_t br_chr[2][2];
};

enum FastKind : uint8_t {
	kFastNone	= 0,
	kFastEscape = 1, // one affine descriptor with a prefix escape (JSON strings)
	kFastParity = 2	 // one doubled or escape-less affine descriptor (CSV quotes)
};

// Returns 0 and fills `out`, or a negative code naming the first violated cap.
int lower_plan(const pars_plan& S, DevPlan& out);

}  // namespace pars_hip
