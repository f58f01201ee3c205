// internal.hpp — host-side entry points the serving runtime (server.hip) uses
// on a scanner without going through the blocking C ABI calls.
#pragma once

#include "../../include/pars_hip.h"

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pars_hip {

// The calling thread's last error, which pars_hip_last_error() returns; every
// part of the library (scanners and servers) records its failures here.
int				   set_error(int code, const std::string& msg);
const std::string& last_error();

// Device pointers of one enqueued batch (valid until the scanner's next batch).
struct BatchPtrs {
	const uint32_t* structural;
	const uint32_t* st_start;	  // [n_docs + 1]
	const uint32_t* bracket_pos;
	const uint8_t*	bracket_char;
	const uint32_t* br_start;	  // [n_docs + 1]
	const int32_t*	bracket_pair; // with run_pairs
	const uint8_t*	doc_flags;	  // [n_docs]
	const int32_t*	max_depth;	  // [n_docs]
	const void*		summary;	  // device summary (DevSummary)
	void*			h_summary;	  // its pinned host copy (filled by the caller's D2H)
	size_t			summary_bytes;
	// the per-document results packed in one block for one copy (u32, n = n_docs):
	// [0, n] structural starts, [n+1, 2n+1] bracket starts, [2n+2, 3n+1] max
	// depths, [3n+2, 4n+1] flags, [4n+2] the summary's overflow word
	const uint32_t* pack;
	size_t			pack_bytes;
};

// Size every buffer for batches of up to max_bytes / max_docs documents with
// the worst-case output count (no regrow is ever needed afterwards).
int batch_reserve(pars_scanner* sc, uint64_t max_bytes, uint64_t max_docs);

// Enqueue one batch on stream s without any host synchronisation: document
// spans are (offsets, lengths) into d_in[0, n).  The caller orders later work
// after it on the same stream.
int batch_enqueue(pars_scanner* sc, const uint8_t* d_in, uint64_t n, const uint64_t* offsets,
				  const uint32_t* lengths, uint64_t n_docs, hipStream_t s, bool run_pairs, BatchPtrs* out);

// True when the summary copied to h_summary reports a failed look-back or an
// output past its capacity.
bool batch_failed(const void* h_summary);
// The same from the overflow word of a packed result block.
bool batch_failed_word(uint32_t overflow);

// Bytes of the packed result block of a batch of n_docs documents.
inline uint64_t batch_pack_bytes(uint64_t n_docs) {
	return (4 * n_docs + 3) * 4;
}

}  // namespace pars_hip
