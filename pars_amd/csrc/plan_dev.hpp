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
constexpr int kMaxJoint = 512;	// joint states (product of the descriptors' state counts) of the event path
constexpr int kEvIds	= 8;	// event ids 1..7 (0 = a byte every descriptor treats as background)

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
	// Event path (EventEval, plans of two or more descriptors).  Each
	// descriptor's most common class is its background, and a run of
	// background bytes acts as one (its map f_d is idempotent: f_d(f_d(s)) =
	// f_d(s)).  A byte that is not background for some descriptor is an event;
	// its event id (1..7) numbers the distinct tuples of per-descriptor classes
	// among event bytes.  A lane then steps only at its events.
	//   evc[j][c]: cls[j][c] | (id bit 0) << (24 + j) in [0], (id bit 1) << j |
	//     (id bit 2) << (8 + j) in [1] (pre-shifted like cls).
	//   ev_map[i][g][d]: descriptor d's map over event id i, after a background
	//     run when g = 1 (trans[.][class] after f_d), packed as dt's lo/hi.
	//   ev_joint[J][i][g]: for the joint state J = sum_d s_d ev_stride[d]: the
	//     joint state after the step (bits 0-11), whether the run's first byte
	//     is masked (bit 12), its later bytes (bit 13), the event byte (bit 14);
	//     id 0 with g = 1 is a background run to the lane's end.
	uint32_t evc[8][256][2];
	uint32_t ev_map[kEvIds][2][kMaxDescs][2];
	uint16_t ev_joint[kMaxJoint][kEvIds][2];
	uint16_t ev_stride[kMaxDescs];
	uint16_t ev_n_joint;
	uint8_t	 ev_ok;	 // the event path applies
	uint8_t	 ev_n_ids;
};

enum FastKind : uint8_t {
	kFastNone	= 0,
	kFastEscape = 1, // one affine descriptor with a prefix escape (JSON strings)
	kFastParity = 2	 // one doubled or escape-less affine descriptor (CSV quotes)
};

// Returns 0 and fills `out`, or a negative code naming the first violated cap.
int lower_plan(const pars_plan& S, DevPlan& out);

}  // namespace pars_hip
