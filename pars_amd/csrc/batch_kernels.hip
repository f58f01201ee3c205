// batch_kernels.hip — the batched form of the scan: many documents in one
// persistent launch.  Documents sit in one device buffer, each starting on a
// 64-byte boundary; a head bitmap marks the 64-byte lanes where documents
// start, and every tile looks up the document it starts in.  The tile kernel
// itself (scan_core.hpp, BATCH = true) resets state and depth at heads.
//
// No counterpart in the reference backend, which runs one launch chain per
// document (include/pars/backend/cuda/pipeline.cuh:752-1017); per-document
// results equal those of separate per-document scans.
#include "evaluators.hpp"

namespace pars_hip {

namespace {

// tile_doc[t] = the last document starting at or before byte t * tile_bytes.
__device__ __forceinline__ int32_t doc_at(const BatchArgs& b, uint64_t x) {
	uint32_t lo = 0, hi = b.n_docs;	 // first document starting after x
	while (lo < hi) {
		const uint32_t mid = (lo + hi) >> 1;
		if (b.doc_off[mid] <= x)
			lo = mid + 1;
		else
			hi = mid;
	}
	return (int32_t)lo - 1;
}

// Everything a batch needs reset or looked up before the heads are set.
__global__ void k_batch_prep(BatchArgs b, DevSummary* sum, uint32_t* doc_bad, uint32_t n_tiles, uint64_t tile_bytes) {
	const uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, st = (uint64_t)gridDim.x * blockDim.x;
	if (i0 == 0) {	// a fresh document's summary; the look-back epoch moves on
		sum->n_struct	  = 0;
		sum->n_brackets	  = 0;
		sum->last_bal	  = 0;
		sum->final_states = 0;
		sum->min_bal	  = 0;
		sum->max_bal	  = 0;
		sum->mismatched	  = 0;
		sum->sort_passes  = 0;
		sum->kmin		  = 0;
		sum->n_chunks	  = 0;
		sum->overflow	  = 0;
		sum->small_done	  = 0;
		sum->doc_unbal	  = 0;
		sum->key_total	  = 0;
		const uint32_t e  = (sum->epoch + 1u) & 0x3FFFFFFFu;
		sum->epoch		  = e ? e : 1u;
	}
	for (uint64_t i = i0; i < b.n_head_words; i += st) const_cast<uint64_t*>(b.head_bits)[i] = 0ull;
	for (uint64_t i = i0; i < b.n_docs; i += st) {
		b.doc_min[i]   = 0;
		b.doc_max[i]   = 0;
		b.doc_last[i]  = 0;
		b.doc_final[i] = 0;
		if (doc_bad)
			doc_bad[i] = 0;
	}
	for (uint64_t t = i0; t < n_tiles; t += st) b.tile_doc[t] = doc_at(b, t * tile_bytes);
}

__global__ void k_batch_heads(BatchArgs b) {
	for (uint64_t d = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; d < b.n_docs;
		 d += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t lane = b.doc_off[d] >> 6;
		atomicOr(reinterpret_cast<unsigned long long*>(const_cast<uint64_t*>(b.head_bits)) + (lane >> 6),
				 1ull << (lane & 63));
	}
}

__global__ void k_tile_doc(BatchArgs b, uint32_t n_tiles, uint64_t tile_bytes) {
	for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles; t += gridDim.x * blockDim.x)
		b.tile_doc[t] = doc_at(b, (uint64_t)t * tile_bytes);
}

// nest.hpp:45-92 per document: balanced = depth never below 0 and back to 0;
// well_formed = balanced, an even bracket count and every sorted pair matching.
__global__ void k_batch_finish(BatchArgs b, const uint32_t* doc_bad, uint32_t start_mask, uint32_t start_packed,
							   int run_pairs, uint8_t* flags, const DevSummary* sum) {
	const uint64_t nd = b.n_docs;
	uint32_t*	   pk = b.pack;
	for (uint64_t d = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; d < nd; d += (uint64_t)gridDim.x * blockDim.x) {
		const bool	   bal = b.doc_min[d] >= 0 && b.doc_last[d] == 0;
		const uint32_t nb  = b.br_start[d + 1] - b.br_start[d];
		const bool	   wf  = bal && (nb % 2 == 0) && !(run_pairs && doc_bad[d]);
		const bool	   unc = (b.doc_final[d] & start_mask) != (start_packed & start_mask);
		const uint8_t  f   = (uint8_t)((bal ? 1u : 0u) | (wf ? 2u : 0u) | (unc ? 4u : 0u));
		flags[d]		   = f;
		if (pk) {
			pk[d]				= b.st_start[d];
			pk[nd + 1 + d]		= b.br_start[d];
			pk[2 * nd + 2 + d]	= (uint32_t)b.doc_max[d];
			pk[3 * nd + 2 + d]	= f;
			if (d == 0) {
				pk[nd]			= b.st_start[nd];
				pk[2 * nd + 1]	= b.br_start[nd];
				pk[4 * nd + 2]	= sum->overflow;
			}
		}
	}
}

unsigned grid_for(uint64_t n) {
	return (unsigned)std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 4096);
}

}  // namespace

hipError_t launch_batch_prep(const BatchArgs& b, uint64_t n, DevSummary* sum, uint32_t* doc_bad, uint32_t tile_rows,
							 hipStream_t s) {
	const uint64_t tbytes  = (uint64_t)tile_rows * kTile;
	const uint32_t n_tiles = (uint32_t)((n + tbytes - 1) / tbytes);
	const uint64_t work	   = std::max<uint64_t>(std::max<uint64_t>(b.n_head_words, b.n_docs), n_tiles);
	hipLaunchKernelGGL(k_batch_prep, dim3(grid_for(work)), dim3(256), 0, s, b, sum, doc_bad, n_tiles, tbytes);
	hipLaunchKernelGGL(k_batch_heads, dim3(grid_for(b.n_docs)), dim3(256), 0, s, b);
	return hipGetLastError();
}

hipError_t launch_tile_doc(const BatchArgs& b, uint32_t n_tiles, uint64_t tile_bytes, hipStream_t s) {
	hipLaunchKernelGGL(k_tile_doc, dim3(grid_for(n_tiles)), dim3(256), 0, s, b, n_tiles, tile_bytes);
	return hipGetLastError();
}

hipError_t launch_batch_finish(const BatchArgs& b, const uint32_t* doc_bad, uint32_t start_mask,
							   uint32_t start_packed, int run_pairs, uint8_t* doc_flags, const DevSummary* sum,
							   hipStream_t s) {
	hipLaunchKernelGGL(k_batch_finish, dim3(grid_for(b.n_docs)), dim3(256), 0, s, b, doc_bad, start_mask,
					   start_packed, run_pairs, doc_flags, sum);
	return hipGetLastError();
}

uint32_t batch_tile_rows(const DevPlan& p, int path, int rows) {
	auto pick = [](int r, int def) -> uint32_t {	// launch_rows: 1..4, anything else 2
		const int w = r ? r : def;
		return (w >= 1 && w <= 4) ? (uint32_t)w : 2u;
	};
	if (path == 0 && (p.fast_kind == kFastEscape || p.fast_kind == kFastParity))
		return pick(rows, 4);
	if (path == 0 && event_path_applies(p))
		return event_tile_rows(rows);
	if (!p.wide && p.n_descs == 1)
		return pick(rows, 2);
	return p.n_descs == 0 && !p.wide ? 2u : 1u;	 // launch_general's fixed shapes
}

hipError_t launch_batch_escape(const ScanArgs& a, int rows, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	return launch_rows<EscapeEval, true>(a, rows, 4, s);
}

hipError_t launch_batch_parity(const ScanArgs& a, int rows, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	return launch_rows<ParityEval, true>(a, rows, 4, s);
}

hipError_t launch_batch_general(const ScanArgs& a, int nd, bool wide, int rows, hipStream_t s) {
	return launch_general<true>(a, nd, wide, rows, s);
}

}  // namespace pars_hip
