// api.hip — host orchestration behind the C ABI (include/pars_hip.h).
//
// Reference boundary: include/pars/backend/cuda/pipeline.cuh:1164-1281.  The
// reference sizes its outputs after a blocking count readback (:890-895) and
// syncs again for the R6 flags (:972-981).  Here the whole chain is enqueued
// without a host round trip — every kernel reads its sizes from the device
// summary — and one synchronisation at the end reads the summary back.
// Outputs are sized from the input length (structural <= n) and an optimistic
// bracket capacity; if the device reports more brackets than fit, the scan is
// rerun once with the exact capacity.
#include "../../include/pars_hip.h"
#include "internal.hpp"
#include "kernels.hpp"
#include "memory.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>
#include <new>
#include <string>

namespace pars_hip {
uint64_t pairing_max_chunks(uint64_t cap_br);
}

using namespace pars_hip;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
	char	buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_err = buf;
	return code;
}

int hip_fail(hipError_t e, const char* what) {
	return fail(PARS_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

// One device buffer of a scanner, from the scanner's memory source.
struct DevBuf {
	void*	   p   = nullptr;
	size_t	   cap = 0;
	MemSource* mem = nullptr;  // set for every buffer when the scanner is made

	void release() {
		if (p)
			mem->put(p, cap);
		p	= nullptr;
		cap = 0;
	}

	// Grow-only; `zero` clears a fresh allocation (look-back records, tickets).
	// A failed growth leaves the buffer empty; the next call grows it again.
	int ensure(size_t bytes, bool zero = false) {
		if (bytes <= cap && p)
			return PARS_OK;
		if (!mem)
			return fail(PARS_ERR_DEVICE, "a scanner buffer without a memory source (internal error)");
		release();
		const size_t want = bytes ? bytes : 64;
		int			 r	  = mem->get(&p, want);
		if (r != PARS_OK)
			return r;
		cap = want;
		if (zero) {
			const hipError_t e = hipMemset(p, 0, want);
			if (e != hipSuccess)
				return hip_fail(e, "hipMemset");
		}
		return PARS_OK;
	}

	template<class T>
	T* as() const {
		return static_cast<T*>(p);
	}
};

}  // namespace

namespace pars_hip {
int MemSource::get(void** p, size_t bytes) {
	*p = nullptr;
	if (budget && !budget->charge(bytes))
		return fail(PARS_ERR_MEMORY,
					"device memory budget exceeded: %zu more bytes wanted, %llu of %llu in use", bytes,
					(unsigned long long)budget->in_use.load(), (unsigned long long)budget->limit);
	if (alloc.alloc) {
		*p = alloc.alloc(alloc.ctx, bytes);
		if (!*p) {
			if (budget)
				budget->uncharge(bytes);
			return fail(PARS_ERR_MEMORY, "the caller's allocator returned no memory for %zu bytes", bytes);
		}
		return PARS_OK;
	}
	const hipError_t e = hipMalloc(p, bytes);
	if (e != hipSuccess) {
		*p = nullptr;
		if (budget)
			budget->uncharge(bytes);
		return fail(PARS_ERR_MEMORY, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
	}
	return PARS_OK;
}

void MemSource::put(void* p, size_t bytes) {
	if (alloc.alloc)
		alloc.release(alloc.ctx, p, bytes);
	else
		(void)hipFree(p);
	if (budget)
		budget->uncharge(bytes);
}

int set_error(int code, const std::string& msg) {
	g_err = msg;
	return code;
}

const std::string& last_error() {
	return g_err;
}
}  // namespace pars_hip

struct pars_scanner {
	pars_plan plan;
	DevPlan	  hplan;
	DevBuf	  dplan;
	DevBuf	  recs, ticket, summary;
	DevBuf	  positions, chars, bitmap, pair_pos;
	DevBuf	  bpos, bsidx, bkey, bchar, ukeys0, ukeys1, order0, order1, rank, pair, hist, hist_part;
	DevBuf	  stk_res, stk_rec, stk_g1, stk_g2;	 // stack pairing: chunk residues, chunk depths, group minima
	bool	  stack_pending = false;	 // the last enqueued R6 was the stack path (sort path if unbalanced)
	bool	  pairs_ran		 = false;	 // the last scan_device / scan_host / scan_batch call ran R6
	bool	  stack_launched = false;	 // ... and launched the stack path first
	// the batch chain enqueued last (also what its captured graph replays) tried
	// the stack path and left the sort to the host, if that path refused
	bool	  batch_defers_sort = false;
	bool	  keys_sparse	= false;	 // the document's keys are only at stack-chunk starts (a sort fills them)
	uint64_t  cap_struct = 0, cap_br = 0, br_hint = 0;
	int		  want_chars = 0, want_bitmap = 1, path = 0, rows = 0;
	int		  lb_window = 0, blocks_per_cu = 0, radix_bits = 0, waves = 0;	// tuning; 0 = default
	int		  chase		= 0;
	DevSummary* h_sum	 = nullptr;	 // pinned
	hipEvent_t	ev[3]	 = {nullptr, nullptr, nullptr};
	// scan_host staging
	uint8_t*  h_in		  = nullptr;
	size_t	  h_in_cap	  = 0;
	DevBuf	  d_in;
	uint32_t* h_struct	  = nullptr;
	size_t	  h_struct_cap = 0;
	int32_t*  h_pairs	  = nullptr;
	size_t	  h_pairs_cap = 0;
	// batches: the document table (pinned staging + device), per-document results
	// the document table: n offsets (u64) then n lengths (u32), one copy up
	DevBuf	  doc_tab, tile_doc, head_bits, st_start, br_start, doc_final, doc_min, doc_max, doc_last, bdoc, doc_bad,
		doc_flags, doc_pack, doc_kbase, kb_part;
	uint8_t*  h_doctab	 = nullptr;
	size_t	  h_tab_cap	 = 0;
	int		  use_graph	 = 0;
	// streaming host boundary: two pinned staging buffers, two device chunks,
	// a copy stream, their events, and the chained carry
	uint8_t*	  h_chunk[2]  = {nullptr, nullptr};
	size_t		  h_chunk_cap = 0;
	DevBuf		  d_chunk[2], carry;
	hipStream_t	  copy_stream = nullptr;
	hipEvent_t	  ev_h2d[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr};
	hipGraphExec_t			gexec = nullptr;
	std::vector<uintptr_t>	gkey;  // shape and buffers the captured graph was made for
	// the JSON tape: per-token and per-bracket scratch, the tape, its summary
	int			force_radix = 0;  // R6 through the radix passes (their sorted order is kept)
	DevBuf		dom_tinfo, dom_gbeg, dom_gend, dom_gdec, dom_sub, dom_top, dom_tblk, dom_hblk,
		dom_parent, dom_close, dom_sum, tape_kind, tape_beg, tape_end, tape_soff, tape_arena;
	DomSummary* h_dom = nullptr;  // pinned
	size_t		h_dom_cap = 0;
	// the INI / TOML key-value tables: per-tile and per-lane scratch, the table
	DevBuf		kv_tile, kv_lane, kv_mask, kv_sum, kv_role, kv_beg, kv_end, kv_key;
	KvSummary*	h_kv	 = nullptr;	 // pinned
	size_t		h_kv_cap = 0;
	MemSource	mem;  // every DevBuf above grows through it

	// Every device buffer of the scanner.
	template<class F>
	void each_buf(F&& f) {
		for (DevBuf* b : {&dplan, &recs, &ticket, &summary, &positions, &chars, &bitmap, &pair_pos, &bpos, &bsidx,
						  &bkey, &bchar, &ukeys0, &ukeys1, &order0, &order1, &rank, &pair, &hist, &hist_part, &stk_res,
						  &stk_rec, &stk_g1, &stk_g2, &d_in, &doc_tab, &tile_doc, &head_bits, &st_start, &br_start,
						  &doc_final, &doc_min, &doc_max, &doc_last, &bdoc, &doc_bad, &doc_flags, &doc_pack, &doc_kbase,
						  &kb_part, &d_chunk[0], &d_chunk[1], &carry, &dom_tinfo, &dom_gbeg, &dom_gend, &dom_gdec,
						  &dom_sub, &dom_top, &dom_tblk, &dom_hblk, &dom_parent, &dom_close, &dom_sum, &tape_kind,
						  &tape_beg, &tape_end, &tape_soff, &tape_arena, &kv_tile, &kv_lane, &kv_mask, &kv_sum,
						  &kv_role, &kv_beg, &kv_end, &kv_key})
			f(*b);
	}

	pars_scanner(const pars_allocator* a, pars_budget* b) {
		if (a)
			mem.alloc = *a;
		mem.budget = b;
		if (b)
			b->users.fetch_add(1);
		each_buf([&](DevBuf& d) { d.mem = &mem; });
	}

	~pars_scanner() {
		if (gexec)
			(void)hipGraphExecDestroy(gexec);
		each_buf([](DevBuf& d) { d.release(); });
		if (mem.budget)
			mem.budget->users.fetch_sub(1);
		if (h_dom)
			(void)hipHostFree(h_dom);
		if (h_kv)
			(void)hipHostFree(h_kv);
		if (h_doctab)
			(void)hipHostFree(h_doctab);
		for (int i = 0; i < 2; ++i) {
			if (h_chunk[i])
				(void)hipHostFree(h_chunk[i]);
			if (ev_h2d[i])
				(void)hipEventDestroy(ev_h2d[i]);
			if (ev_done[i])
				(void)hipEventDestroy(ev_done[i]);
		}
		if (copy_stream)
			(void)hipStreamDestroy(copy_stream);
		if (h_sum)
			(void)hipHostFree(h_sum);
		if (h_in)
			(void)hipHostFree(h_in);
		if (h_struct)
			(void)hipHostFree(h_struct);
		if (h_pairs)
			(void)hipHostFree(h_pairs);
		for (hipEvent_t& e : ev)
			if (e)
				(void)hipEventDestroy(e);
	}
};

namespace {

int ensure_pinned(void** p, size_t& cap, size_t bytes) {
	if (bytes <= cap && *p)
		return PARS_OK;
	if (*p)
		(void)hipHostFree(*p);
	*p				  = nullptr;
	cap				  = 0;
	const size_t want = std::max<size_t>(bytes, 64);
	hipError_t	 e	  = hipHostMalloc(p, want, hipHostMallocDefault);
	if (e != hipSuccess) {
		*p = nullptr;
		return fail(PARS_ERR_MEMORY, "hipHostMalloc(%zu): %s", want, hipGetErrorString(e));
	}
	cap = want;
	return PARS_OK;
}

// Capacity for per-structural and per-bracket outputs of a scan of n bytes.
int ensure_outputs(pars_scanner* sc, uint64_t n_tiles, uint64_t cap_struct, uint64_t cap_br, bool want_pairs) {
	int r;
	if ((r = sc->recs.ensure(n_tiles * kRecWords * sizeof(uint64_t), true)))
		return r;
	if ((r = sc->positions.ensure(cap_struct * 4)))
		return r;
	if (sc->want_chars && (r = sc->chars.ensure(cap_struct)))
		return r;
	if (want_pairs && (r = sc->pair_pos.ensure(cap_struct * 4)))
		return r;
	for (DevBuf* b : {&sc->bpos, &sc->bsidx, &sc->bkey, &sc->ukeys0, &sc->ukeys1, &sc->order0, &sc->order1, &sc->rank,
					  &sc->pair})
		if ((r = b->ensure(cap_br * 4)))
			return r;
	if ((r = sc->bchar.ensure(cap_br)))
		return r;
	if (sc->force_radix && (r = sc->dom_close.ensure(cap_br * 4)))  // the tape's bracket partners
		return r;
	const uint64_t chunks = pairing_max_chunks(cap_br);
	if ((r = sc->hist.ensure(256 * std::max<uint64_t>(chunks, 1) * 4)))
		return r;
	const uint64_t schunks = (cap_br + kStkChunk - 1) / kStkChunk;
	// group minima start zeroed: epoch 0 is never a launch's, so they read as unset
	if ((r = sc->stk_res.ensure(cap_br * 4)) || (r = sc->stk_rec.ensure(std::max<uint64_t>(schunks, 1) * 12))
		|| (r = sc->stk_g1.ensure((schunks / 64 + 1) * 8, true)) || (r = sc->stk_g2.ensure((schunks / 4096 + 1) * 8, true)))
		return r;
	if ((r = sc->hist_part.ensure(256 * 4)))
		return r;
	sc->cap_struct = cap_struct;
	sc->cap_br	   = cap_br;
	return PARS_OK;
}

ScanArgs make_scan_args(pars_scanner* sc, const uint8_t* d_in, uint64_t n, bool want_pairs) {
	ScanArgs a {};
	a.data		 = d_in;
	a.n			 = n;
	a.n_tiles	 = (uint32_t)((n + kTile - 1) / kTile);
	a.plan		 = sc->dplan.as<DevPlan>();
	a.recs		 = sc->recs.as<uint64_t>();
	a.ticket	 = sc->ticket.as<uint32_t>();
	a.done		 = sc->ticket.as<uint32_t>() + 1;
	a.sum		 = sc->summary.as<DevSummary>();
	a.positions	 = sc->positions.as<uint32_t>();
	a.chars		 = sc->want_chars ? sc->chars.as<uint8_t>() : nullptr;
	a.bitmap	 = sc->want_bitmap ? sc->bitmap.as<uint64_t>() : nullptr;
	a.pair_pos	 = want_pairs ? sc->pair_pos.as<int32_t>() : nullptr;
	a.bpos		 = sc->bpos.as<uint32_t>();
	// structural indices of brackets: read only for pair_pos and the tape's partners
	a.bsidx		 = (want_pairs || sc->force_radix) ? sc->bsidx.as<uint32_t>() : nullptr;
	a.bkey		 = sc->bkey.as<int32_t>();
	a.bchar		 = sc->bchar.as<uint8_t>();
	a.cap_struct = sc->cap_struct;
	a.cap_br	 = sc->cap_br;
	a.in_states	 = sc->hplan.start_packed;
	a.lb_window	 = (uint32_t)sc->lb_window;
	a.blocks_per_cu = (uint32_t)sc->blocks_per_cu;
	a.chase			= (uint32_t)sc->chase;
	return a;
}

PairArgs make_pair_args(pars_scanner* sc, bool want_pairs) {
	PairArgs p {};
	p.sum		 = sc->summary.as<DevSummary>();
	p.plan		 = sc->dplan.as<DevPlan>();
	p.bkey		 = sc->bkey.as<int32_t>();
	p.bchar		 = sc->bchar.as<uint8_t>();
	p.bsidx		 = sc->bsidx.as<uint32_t>();
	p.cap_br	 = sc->cap_br;
	p.ukeys[0]	 = sc->ukeys0.as<uint32_t>();
	p.ukeys[1]	 = sc->ukeys1.as<uint32_t>();
	p.order[0]	 = sc->order0.as<uint32_t>();
	p.order[1]	 = sc->order1.as<uint32_t>();
	p.rank		 = sc->rank.as<uint32_t>();
	p.hist		 = sc->hist.as<uint32_t>();
	p.hist_part	 = sc->hist_part.as<uint32_t>();
	p.max_chunks = pairing_max_chunks(sc->cap_br);
	p.pair		 = sc->pair.as<int32_t>();
	p.pair_pos	 = want_pairs ? sc->pair_pos.as<int32_t>() : nullptr;
	p.close		 = sc->force_radix ? sc->dom_close.as<uint32_t>() : nullptr;
	p.max_digit_bits = (uint32_t)sc->radix_bits;
	return p;
}

StackArgs make_stack_args(pars_scanner* sc, bool want_pairs) {
	StackArgs a {};
	a.sum		 = sc->summary.as<DevSummary>();
	a.plan		 = sc->dplan.as<DevPlan>();
	a.bchar		 = sc->bchar.as<uint8_t>();
	a.bkey		 = sc->bkey.as<int32_t>();
	a.bsidx		 = sc->bsidx.as<uint32_t>();
	a.cap_br	 = sc->cap_br;
	a.pair		 = sc->pair.as<int32_t>();
	a.pair_pos	 = want_pairs ? sc->pair_pos.as<int32_t>() : nullptr;
	a.res		 = sc->stk_res.as<uint32_t>();
	a.crec		 = sc->stk_rec.as<int32_t>();
	a.gmin1		 = sc->stk_g1.as<unsigned long long>();
	a.gmin2		 = sc->stk_g2.as<unsigned long long>();
	a.max_chunks = std::max<uint64_t>((sc->cap_br + kStkChunk - 1) / kStkChunk, 1);
	return a;
}

// Whether R6 of a document of n bytes starts on the stack path (enqueue_pairs);
// its scan then writes keys only at the stack chunks' starts.
bool stack_first(const pars_scanner* sc, uint64_t n) {
	return !sc->radix_bits && !sc->force_radix && small_pairing_mode(n, 1) != 2;
}

// R6 of one document (or of the chunks of one): the stack path unless a
// digit cap, a consumer of the sorted order or the one-block path asks for the
// sort.  The stack path's kernels exit on an unbalanced document, which then
// takes the sort once the summary shows it (finish_pairs).
hipError_t enqueue_pairs(pars_scanner* sc, uint64_t n, bool want_pairs, hipStream_t s) {
	const uint32_t small = small_pairing_mode(n, 1);
	if (stack_first(sc, n)) {
		sc->stack_pending  = true;
		sc->stack_launched = true;
		return launch_stack_pairing(make_stack_args(sc, want_pairs), s);
	}
	sc->stack_pending = false;
	PairArgs p		  = make_pair_args(sc, want_pairs);
	p.small			  = (sc->radix_bits || sc->force_radix) ? 0u : small;
	p.key_fill		  = sc->keys_sparse ? sc->bkey.as<int32_t>() : nullptr;
	return launch_pairing(p, s);
}

int read_summary(pars_scanner* sc, hipStream_t s);

// After the summary of a stack-path R6 is read: an unbalanced document runs the
// sort (its pairs differ from a stack's), then the summary is read again.
int finish_pairs(pars_scanner* sc, bool want_pairs, hipStream_t s, double* ms) {
	if (!sc->stack_pending)
		return PARS_OK;
	sc->stack_pending = false;
	const DevSummary& h = *sc->h_sum;
	if (h.overflow & kStackLost)
		return fail(PARS_ERR_DEVICE, "stack pairing found an unmatched residue on a balanced document (internal error)");
	if (h.min_bal >= 0 && h.last_bal == 0)
		return PARS_OK;
	hipError_t e;
	if (ms && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	PairArgs p = make_pair_args(sc, want_pairs);
	p.small	   = 0;
	p.key_fill = sc->keys_sparse ? sc->bkey.as<int32_t>() : nullptr;
	if ((e = launch_pairing(p, s)) != hipSuccess)
		return hip_fail(e, "pairing launch");
	if (ms && (e = hipEventRecord(sc->ev[2], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	int r;
	if ((r = read_summary(sc, s)))
		return r;
	float t = 0;
	if (ms && hipEventElapsedTime(&t, sc->ev[1], sc->ev[2]) == hipSuccess)
		*ms += t;
	return PARS_OK;
}

int read_summary(pars_scanner* sc, hipStream_t s) {
	hipError_t e = hipMemcpyAsync(sc->h_sum, sc->summary.p, sizeof(DevSummary), hipMemcpyDeviceToHost, s);
	if (e != hipSuccess)
		return hip_fail(e, "summary readback");
	e = hipStreamSynchronize(s);
	if (e != hipSuccess)
		return hip_fail(e, "hipStreamSynchronize");
	if (sc->h_sum->overflow & 0x80000000u)
		return fail(PARS_ERR_DEVICE, "scan look-back exceeded its spin bound (internal error)");
	if (sc->h_sum->overflow & 8u)
		return fail(PARS_ERR_ARGUMENT, "the radix digit cap needs more than four pairing passes for this key range");
	return PARS_OK;
}

void fill_result(pars_scanner* sc, uint64_t n, bool want_pairs, bool run_pairs, pars_device_result* out) {
	const DevSummary& h = *sc->h_sum;
	std::memset(out, 0, sizeof *out);
	out->structural	   = sc->positions.as<uint32_t>();
	out->pair_pos	   = want_pairs ? sc->pair_pos.as<int32_t>() : nullptr;
	out->n_struct	   = h.n_struct;
	out->balanced	   = (h.min_bal >= 0) && (h.last_bal == 0);
	out->well_formed   = out->balanced && !h.mismatched && (h.n_brackets % 2 == 0);
	out->max_depth	   = h.max_bal;
	out->chars		   = sc->want_chars ? sc->chars.as<uint8_t>() : nullptr;
	out->struct_bitmap = sc->want_bitmap ? sc->bitmap.as<uint64_t>() : nullptr;
	out->bracket_pos   = sc->bpos.as<uint32_t>();
	out->bracket_char  = sc->bchar.as<uint8_t>();
	out->bracket_pair  = run_pairs ? sc->pair.as<int32_t>() : nullptr;
	out->n_brackets	   = h.n_brackets;
	const uint32_t nd  = sc->hplan.n_descs;
	const uint32_t m   = nd ? ((1u << (3 * nd)) - 1u) : 0u;
	out->unclosed_mode = n ? ((h.final_states & m) != (sc->hplan.start_packed & m)) : 0;
	if (!run_pairs)
		out->well_formed = out->balanced && (h.n_brackets % 2 == 0);
}

// R1-R5 through the fastest evaluator that is exact for the plan.
hipError_t launch_scan(pars_scanner* sc, const ScanArgs& a, hipStream_t s) {
	if (sc->path == 0 && sc->hplan.fast_kind == kFastEscape)
		return launch_scan_escape(a, sc->rows, sc->waves, s);
	if (sc->path == 0 && sc->hplan.fast_kind == kFastParity)
		return launch_scan_parity(a, sc->rows, sc->waves, s);
	if (sc->path == 0 && event_path_applies(sc->hplan))
		return launch_scan_event(a, sc->hplan.n_descs, sc->hplan.wide, sc->rows, s);
	return launch_scan_general(a, sc->hplan.n_descs, sc->hplan.wide, sc->rows, s);
}

// One full document: reset, R1-R5, optional R6; no host sync inside.
int enqueue_document(pars_scanner* sc, const uint8_t* d_in, uint64_t n, hipStream_t s, bool want_pairs,
					 bool run_pairs, bool timed) {
	hipError_t e;
	if ((e = launch_summary_reset(sc->summary.as<DevSummary>(), SummaryInit {}, s)) != hipSuccess)
		return hip_fail(e, "summary reset");
	if (timed && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	ScanArgs a		= make_scan_args(sc, d_in, n, want_pairs);
	a.key_sparse	= !run_pairs || stack_first(sc, n);
	sc->keys_sparse = a.key_sparse;
	if ((e = launch_scan(sc, a, s)) != hipSuccess)
		return hip_fail(e, "scan kernel launch");
	if (timed && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	sc->stack_pending  = false;
	sc->pairs_ran	   = run_pairs;
	sc->stack_launched = false;
	if (run_pairs && (e = enqueue_pairs(sc, n, want_pairs, s)) != hipSuccess)
		return hip_fail(e, "pairing launch");
	if (timed && (e = hipEventRecord(sc->ev[2], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	return PARS_OK;
}

// ---------------------------------------------------------------------------
// Batches.
// ---------------------------------------------------------------------------
// The host document table: offsets at [0, 8n), lengths at [8n, 12n).
uint64_t* tab_off(pars_scanner* sc) {
	return reinterpret_cast<uint64_t*>(sc->h_doctab);
}

uint32_t* tab_len(pars_scanner* sc, uint64_t n_docs) {
	return reinterpret_cast<uint32_t*>(sc->h_doctab + 8 * n_docs);
}

BatchArgs make_batch_args(pars_scanner* sc, uint64_t n, uint64_t n_docs) {
	BatchArgs b {};
	b.doc_off	   = sc->doc_tab.as<uint64_t>();
	b.doc_len	   = reinterpret_cast<const uint32_t*>(sc->doc_tab.as<uint8_t>() + 8 * n_docs);
	b.tile_doc	   = sc->tile_doc.as<int32_t>();
	b.head_bits	   = sc->head_bits.as<uint64_t>();
	b.n_head_words = (((n + 63) / 64) + 63) / 64;
	b.n_docs	   = (uint32_t)n_docs;
	b.st_start	   = sc->st_start.as<uint32_t>();
	b.br_start	   = sc->br_start.as<uint32_t>();
	b.doc_final	   = sc->doc_final.as<uint32_t>();
	b.doc_min	   = sc->doc_min.as<int32_t>();
	b.doc_max	   = sc->doc_max.as<int32_t>();
	b.doc_last	   = sc->doc_last.as<int32_t>();
	b.bdoc		   = sc->bdoc.as<uint32_t>();
	return b;
}

int ensure_batch(pars_scanner* sc, uint64_t n, uint64_t n_docs, uint64_t cap_br, bool want_pairs) {
	int			   r;
	const uint64_t n_tiles = (n + kTile - 1) / kTile;
	if ((r = ensure_outputs(sc, n_tiles, std::max<uint64_t>(n, sc->cap_struct), cap_br, want_pairs)))
		return r;
	if (sc->want_bitmap && (r = sc->bitmap.ensure(((n + 63) / 64) * 8)))
		return r;
	const uint64_t words = (((n + 63) / 64) + 63) / 64;
	if ((r = sc->doc_tab.ensure(n_docs * 12)) || (r = sc->doc_pack.ensure(batch_pack_words(n_docs) * 4))
		|| (r = sc->tile_doc.ensure(n_tiles * 4)) || (r = sc->head_bits.ensure(std::max<uint64_t>(words, 1) * 8))
		|| (r = sc->st_start.ensure((n_docs + 1) * 4)) || (r = sc->br_start.ensure((n_docs + 1) * 4))
		|| (r = sc->doc_final.ensure(n_docs * 4)) || (r = sc->doc_min.ensure(n_docs * 4))
		|| (r = sc->doc_max.ensure(n_docs * 4)) || (r = sc->doc_last.ensure(n_docs * 4))
		|| (r = sc->doc_bad.ensure(n_docs * 4)) || (r = sc->doc_flags.ensure(n_docs))
		|| (r = sc->doc_kbase.ensure(n_docs * 4)) || (r = sc->kb_part.ensure(256 * 8))
		|| (r = sc->bdoc.ensure(sc->cap_br * 4)))
		return r;
	if ((r = ensure_pinned((void**)&sc->h_doctab, sc->h_tab_cap, n_docs * 12)))
		return r;
	return PARS_OK;
}

hipError_t launch_batch_scan(pars_scanner* sc, const ScanArgs& a, hipStream_t s) {
	if (sc->path == 0 && sc->hplan.fast_kind == kFastEscape)
		return launch_batch_escape(a, sc->rows, s);
	if (sc->path == 0 && sc->hplan.fast_kind == kFastParity)
		return launch_batch_parity(a, sc->rows, s);
	if (sc->path == 0 && event_path_applies(sc->hplan))
		return launch_batch_event(a, sc->hplan.n_descs, sc->hplan.wide, sc->rows, s);
	return launch_batch_general(a, sc->hplan.n_descs, sc->hplan.wide, sc->rows, s);
}

PairArgs make_batch_pair_args(pars_scanner* sc, const BatchArgs& b, bool want_pairs) {
	PairArgs p	= make_pair_args(sc, want_pairs);
	p.bdoc		= b.bdoc;
	p.br_start	= b.br_start;
	p.st_start	= b.st_start;
	p.n_docs	= b.n_docs;
	p.doc_bad	= sc->doc_bad.as<uint32_t>();
	p.doc_min	= b.doc_min;
	p.doc_max	= b.doc_max;
	p.doc_kbase = sc->doc_kbase.as<uint32_t>();
	p.kb_part	= sc->kb_part.as<uint64_t>();
	return p;
}

hipError_t launch_finish(pars_scanner* sc, const BatchArgs& b, bool run_pairs, hipStream_t s) {
	const uint32_t nd = sc->hplan.n_descs;
	const uint32_t m  = nd ? ((1u << (3 * nd)) - 1u) : 0u;
	return launch_batch_finish(b, sc->doc_bad.as<uint32_t>(), m, sc->hplan.start_packed, run_pairs ? 1 : 0,
							   sc->doc_flags.as<uint8_t>(), sc->summary.as<DevSummary>(), s);
}

// The whole batch: no host sync inside, so it can be captured into a graph.
// A batch over 1 MiB pairs by stack matching when every document is
// balanced; whether it is shows on the device only.  `defer_sort` (the
// scan_batch call, which reads the summary back anyway): the sorting chain is
// not launched, and finish_batch_pairs runs it after the read-back if the
// stack path refused - one more round trip for a batch holding an unbalanced
// document, ~20 launches fewer for every other.  Otherwise (the serving
// runtime, which never waits) the sorting chain follows the stack kernels and
// exits at once when they paired the batch.
int enqueue_batch(pars_scanner* sc, const uint8_t* d_in, uint64_t n, uint64_t n_docs, hipStream_t s,
				  bool want_pairs, bool run_pairs, bool timed, bool packed = false) {
	const bool defer_sort = !packed;
	hipError_t e;
	if ((e = hipMemcpyAsync(sc->doc_tab.p, sc->h_doctab, n_docs * 12, hipMemcpyHostToDevice, s)) != hipSuccess)
		return hip_fail(e, "document table H2D");
	BatchArgs b = make_batch_args(sc, n, n_docs);
	b.pack		= packed ? sc->doc_pack.as<uint32_t>() : nullptr;
	b.tile_rows = batch_tile_rows(sc->hplan, sc->path, sc->rows);
	if ((e = launch_batch_prep(b, n, sc->summary.as<DevSummary>(), run_pairs ? sc->doc_bad.as<uint32_t>() : nullptr,
							   b.tile_rows, s))
		!= hipSuccess)
		return hip_fail(e, "batch prep");
	if (timed && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	ScanArgs a = make_scan_args(sc, d_in, n, want_pairs);
	a.b		   = b;
	if ((e = launch_batch_scan(sc, a, s)) != hipSuccess)
		return hip_fail(e, "batch scan launch");
	if (timed && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	sc->pairs_ran		  = run_pairs;
	sc->stack_launched	  = false;
	sc->batch_defers_sort = false;
	if (run_pairs) {
		PairArgs p = make_batch_pair_args(sc, b, want_pairs);
		p.small	   = sc->radix_bits ? 0u : small_pairing_mode(n, n_docs);
		// every document balanced: stack matching pairs the batch and marks the
		// summary so the sorting chain after it exits at once; otherwise the
		// stack kernels exit and the sort pairs it (decided on the device)
		if (p.small == 0 && !sc->radix_bits) {
			StackArgs k = make_stack_args(sc, want_pairs);
			k.bdoc		= b.bdoc;
			k.br_start	= b.br_start;
			k.st_start	= b.st_start;
			k.doc_bad	= sc->doc_bad.as<uint32_t>();
			if ((e = launch_stack_pairing(k, s)) != hipSuccess)
				return hip_fail(e, "stack pairing launch");
			sc->stack_launched	  = true;
			sc->batch_defers_sort = defer_sort;
		}
		if (!sc->batch_defers_sort && (e = launch_pairing(p, s)) != hipSuccess)
			return hip_fail(e, "pairing launch");
	}
	if ((e = launch_finish(sc, b, run_pairs, s)) != hipSuccess)
		return hip_fail(e, "batch finish");
	if (timed && (e = hipEventRecord(sc->ev[2], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	return PARS_OK;
}

std::vector<uintptr_t> batch_graph_key(pars_scanner* sc, const uint8_t* d_in, uint64_t n, uint64_t n_docs,
									   hipStream_t s, bool want_pairs, bool run_pairs) {
	std::vector<uintptr_t> k = {(uintptr_t)d_in, (uintptr_t)n, (uintptr_t)n_docs, (uintptr_t)s,
								(uintptr_t)want_pairs, (uintptr_t)run_pairs, (uintptr_t)sc->want_chars,
								(uintptr_t)sc->want_bitmap, (uintptr_t)sc->path, (uintptr_t)sc->rows,
								(uintptr_t)sc->lb_window, (uintptr_t)sc->blocks_per_cu, (uintptr_t)sc->radix_bits,
								(uintptr_t)sc->chase,
								(uintptr_t)sc->cap_struct, (uintptr_t)sc->cap_br, (uintptr_t)sc->h_doctab};
	for (DevBuf* b : {&sc->recs, &sc->ticket, &sc->summary, &sc->positions, &sc->chars, &sc->bitmap, &sc->pair_pos,
					  &sc->bpos, &sc->bsidx, &sc->bkey, &sc->bchar, &sc->ukeys0, &sc->ukeys1, &sc->order0,
					  &sc->order1, &sc->rank, &sc->pair, &sc->hist, &sc->hist_part, &sc->doc_tab, &sc->tile_doc, &sc->head_bits, &sc->st_start, &sc->br_start, &sc->doc_final, &sc->doc_min,
					  &sc->doc_max, &sc->doc_last, &sc->bdoc, &sc->doc_bad, &sc->doc_flags, &sc->doc_kbase, &sc->kb_part})
		k.push_back((uintptr_t)b->p);
	return k;
}

// Replay the captured chain for this shape.  A new shape runs eagerly once
// (which also settles every lazily initialised launch parameter), is captured
// on its second call and replayed from then on.
int run_batch_graph(pars_scanner* sc, const uint8_t* d_in, uint64_t n, uint64_t n_docs, hipStream_t s,
					bool want_pairs, bool run_pairs) {
	const std::vector<uintptr_t> key = batch_graph_key(sc, d_in, n, n_docs, s, want_pairs, run_pairs);
	hipError_t					 e;
	if (sc->gkey != key) {
		if (sc->gexec) {
			(void)hipGraphExecDestroy(sc->gexec);
			sc->gexec = nullptr;
		}
		sc->gkey = key;
		return enqueue_batch(sc, d_in, n, n_docs, s, want_pairs, run_pairs, false);
	}
	if (!sc->gexec) {
		if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
			return hip_fail(e, "hipStreamBeginCapture");
		const int	r = enqueue_batch(sc, d_in, n, n_docs, s, want_pairs, run_pairs, false);
		hipGraph_t	g = nullptr;
		e			  = hipStreamEndCapture(s, &g);
		if (r)
			return r;
		if (e != hipSuccess)
			return hip_fail(e, "hipStreamEndCapture");
		e = hipGraphInstantiate(&sc->gexec, g, nullptr, nullptr, 0);
		(void)hipGraphDestroy(g);
		if (e != hipSuccess) {
			sc->gexec = nullptr;
			return hip_fail(e, "hipGraphInstantiate");
		}
	}
	if ((e = hipGraphLaunch(sc->gexec, s)) != hipSuccess)
		return hip_fail(e, "hipGraphLaunch");
	return PARS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Internal entry points for the serving runtime (internal.hpp).
// ---------------------------------------------------------------------------
namespace pars_hip {

int batch_reserve(pars_scanner* sc, uint64_t max_bytes, uint64_t max_docs) {
	const uint64_t n = std::max<uint64_t>(max_bytes, 64);
	sc->br_hint		 = n;
	return ensure_batch(sc, n, std::max<uint64_t>(max_docs, 1), n, false);
}

int batch_enqueue(pars_scanner* sc, const uint8_t* d_in, uint64_t n, const uint64_t* offsets,
				  const uint32_t* lengths, uint64_t n_docs, hipStream_t s, bool run_pairs, BatchPtrs* out) {
	if (!n_docs || n > sc->cap_struct || n > sc->cap_br || n_docs * 12 > sc->h_tab_cap
		|| batch_pack_words(n_docs) * 4 > sc->doc_pack.cap)
		return fail(PARS_ERR_ARGUMENT, "batch exceeds the reserved capacity");
	std::memcpy(tab_off(sc), offsets, n_docs * 8);
	std::memcpy(tab_len(sc, n_docs), lengths, n_docs * 4);
	int r;
	if ((r = enqueue_batch(sc, d_in, n, n_docs, s, false, run_pairs, false, true)))
		return r;
	out->pack		   = sc->doc_pack.as<uint32_t>();
	out->pack_bytes	   = batch_pack_words(n_docs) * 4;
	out->structural	   = sc->positions.as<uint32_t>();
	out->st_start	   = sc->st_start.as<uint32_t>();
	out->bracket_pos   = sc->bpos.as<uint32_t>();
	out->bracket_char  = sc->bchar.as<uint8_t>();
	out->br_start	   = sc->br_start.as<uint32_t>();
	out->bracket_pair  = run_pairs ? sc->pair.as<int32_t>() : nullptr;
	out->doc_flags	   = sc->doc_flags.as<uint8_t>();
	out->max_depth	   = sc->doc_max.as<int32_t>();
	out->summary	   = sc->summary.p;
	out->h_summary	   = sc->h_sum;
	out->summary_bytes = sizeof(DevSummary);
	return PARS_OK;
}

bool batch_failed_word(uint32_t overflow) {
	return (overflow & (0x8000000Du | kStackLost)) != 0;
}

bool batch_failed(const void* h_summary) {
	return batch_failed_word(static_cast<const DevSummary*>(h_summary)->overflow);
}

}  // namespace pars_hip


// ---------------------------------------------------------------------------
// Scanning without allocating (pars_hip_scan_device_noalloc).  Every scratch
// buffer of the single-document chain - the tile look-back records, the
// ticket, the summary, the bracket hand-off R6 reads (structural index, key),
// the stack path's residues, chunk depths and group minima, and the sorting
// path's keys, order, ranks and histograms - lives in one file-scope device
// array, `scratch`, sized for documents of up to kNaMaxBytes bytes and
// kNaMaxBrackets brackets.  Each GPU holds its own copy until the process
// ends; its address is looked up once per device.  Nothing needs clearing
// between calls: look-back records and group minima are tagged with the
// summary's epoch (the array starts at zero, epoch 0 is never a launch's), the
// summary is reset by a kernel, the scan kernel leaves its ticket at zero.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kNaMaxBytes	  = 32ull << 20;
constexpr uint64_t kNaMaxBrackets = 1ull << 20;
constexpr uint64_t na_up(uint64_t x) {
	return (x + 255) / 256 * 256;
}
constexpr uint64_t kNaTiles		 = (kNaMaxBytes + kTile - 1) / kTile;  // one-row tiles need the most records
constexpr uint64_t kNaStkChunks	 = (kNaMaxBrackets + kStkChunk - 1) / kStkChunk;
constexpr uint64_t kNaSortChunks = kNaMaxBrackets / 1024;  // at least the sort's chunks (4096 brackets each)
// byte offsets into the array
constexpr uint64_t kNaRecs	 = 0;
constexpr uint64_t kNaTicket = kNaRecs + na_up(kNaTiles * kRecWords * 8);
constexpr uint64_t kNaSum	 = kNaTicket + 256;
constexpr uint64_t kNaBr	 = kNaSum + na_up(sizeof(DevSummary));	// 8 arrays of 4 B per bracket
constexpr uint64_t kNaBrArr	 = na_up(kNaMaxBrackets * 4);
enum : int { kNaBsidx, kNaBkey, kNaUk0, kNaUk1, kNaOrd0, kNaOrd1, kNaRank, kNaRes, kNaBrCount };
constexpr uint64_t kNaHist	   = kNaBr + kNaBrCount * kNaBrArr;
constexpr uint64_t kNaHistPart = kNaHist + na_up(256 * 4 * kNaSortChunks);
constexpr uint64_t kNaCrec	   = kNaHistPart + na_up(256 * 4);
constexpr uint64_t kNaG1	   = kNaCrec + na_up(kNaStkChunks * 12);
constexpr uint64_t kNaG2	   = kNaG1 + na_up((kNaStkChunks / 64 + 1) * 8);
constexpr uint64_t kNaBytes	   = kNaG2 + na_up((kNaStkChunks / 4096 + 1) * 8);

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex	  g_na_mu;
uint8_t*	  g_na_ptr[kNaMaxDevices] = {};

// The calling thread's device's copy of `scratch`: its address, looked up the
// first time a call runs on that device (with it current).  The array's name
// in host code is not a device address; only this pointer is passed on.
int na_scratch(uint8_t** out) {
	int		   dev = 0;
	hipError_t e   = hipGetDevice(&dev);
	if (e != hipSuccess)
		return hip_fail(e, "hipGetDevice");
	if (dev < 0 || dev >= kNaMaxDevices)
		return fail(PARS_ERR_DEVICE, "device %d is beyond the %d this entry point serves", dev, kNaMaxDevices);
	std::lock_guard<std::mutex> lk(g_na_mu);
	if (!g_na_ptr[dev]) {
		void* p = nullptr;
		if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess)
			return hip_fail(e, "hipGetSymbolAddress(scratch)");
		if (!p)
			return fail(PARS_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
		g_na_ptr[dev] = static_cast<uint8_t*>(p);
	}
	*out = g_na_ptr[dev];
	return PARS_OK;
}

template<class T>
T* na_at(uint8_t* base, uint64_t off) {
	return reinterpret_cast<T*>(base + off);
}

PairArgs na_pair_args(pars_scanner* sc, uint8_t* S, uint64_t cap_br, const uint8_t* bchar, int32_t* pair) {
	PairArgs p {};
	p.sum			 = na_at<DevSummary>(S, kNaSum);
	p.plan			 = sc->dplan.as<DevPlan>();
	p.bkey			 = na_at<int32_t>(S, kNaBr + kNaBkey * kNaBrArr);
	p.bchar			 = bchar;
	p.bsidx			 = na_at<uint32_t>(S, kNaBr + kNaBsidx * kNaBrArr);
	p.cap_br		 = cap_br;
	p.ukeys[0]		 = na_at<uint32_t>(S, kNaBr + kNaUk0 * kNaBrArr);
	p.ukeys[1]		 = na_at<uint32_t>(S, kNaBr + kNaUk1 * kNaBrArr);
	p.order[0]		 = na_at<uint32_t>(S, kNaBr + kNaOrd0 * kNaBrArr);
	p.order[1]		 = na_at<uint32_t>(S, kNaBr + kNaOrd1 * kNaBrArr);
	p.rank			 = na_at<uint32_t>(S, kNaBr + kNaRank * kNaBrArr);
	p.hist			 = na_at<uint32_t>(S, kNaHist);
	p.hist_part		 = na_at<uint32_t>(S, kNaHistPart);
	p.max_chunks	 = pairing_max_chunks(cap_br);
	p.pair			 = pair;
	p.max_digit_bits = (uint32_t)sc->radix_bits;
	return p;
}

int na_read_summary(pars_scanner* sc, DevSummary* d_sum, hipStream_t s) {
	hipError_t e = hipMemcpyAsync(sc->h_sum, d_sum, sizeof(DevSummary), hipMemcpyDeviceToHost, s);
	if (e != hipSuccess)
		return hip_fail(e, "summary readback");
	if ((e = hipStreamSynchronize(s)) != hipSuccess)
		return hip_fail(e, "hipStreamSynchronize");
	const uint32_t o = sc->h_sum->overflow;
	if (o & 0x80000000u)
		return fail(PARS_ERR_DEVICE, "scan look-back exceeded its spin bound (internal error)");
	if (o & 8u)
		return fail(PARS_ERR_ARGUMENT, "the radix digit cap needs more than four pairing passes for this key range");
	if (o & kStackLost)
		return fail(PARS_ERR_DEVICE, "stack pairing found an unmatched residue on a balanced document (internal error)");
	return PARS_OK;
}

}  // namespace

extern "C" {

const char* pars_hip_version(void) {
	return "pars-hip 0.1 (gfx950)";
}

const char* pars_hip_last_error(void) {
	return g_err.c_str();
}

int pars_hip_budget_create(uint64_t limit_bytes, pars_budget** out) {
	if (!out)
		return fail(PARS_ERR_ARGUMENT, "null argument");
	*out = new (std::nothrow) pars_budget();
	if (!*out)
		return fail(PARS_ERR_MEMORY, "host allocation failed");
	(*out)->limit = limit_bytes;
	return PARS_OK;
}

int pars_hip_budget_destroy(pars_budget* b) {
	if (!b)
		return PARS_OK;
	if (b->users.load())
		return fail(PARS_ERR_ARGUMENT, "the budget is still used by %u scanners or servers", b->users.load());
	delete b;
	return PARS_OK;
}

int pars_hip_budget_usage(const pars_budget* b, uint64_t* in_use, uint64_t* peak, uint64_t* limit) {
	if (!b)
		return fail(PARS_ERR_ARGUMENT, "null budget");
	if (in_use)
		*in_use = b->in_use.load();
	if (peak)
		*peak = b->peak.load();
	if (limit)
		*limit = b->limit;
	return PARS_OK;
}

int pars_hip_scanner_create(const pars_plan* plan, pars_scanner** out) {
	return pars_hip_scanner_create_ex(plan, nullptr, nullptr, out);
}

int pars_hip_scanner_create_ex(const pars_plan* plan, const pars_allocator* alloc, pars_budget* budget,
							   pars_scanner** out) {
	if (!plan || !out)
		return fail(PARS_ERR_ARGUMENT, "null argument");
	*out = nullptr;
	if (alloc && (!alloc->alloc || !alloc->release))
		return fail(PARS_ERR_ARGUMENT, "an allocator needs both alloc and release");
	pars_scanner* sc = new (std::nothrow) pars_scanner(alloc, budget);
	if (!sc)
		return fail(PARS_ERR_MEMORY, "host allocation failed");
	std::memcpy(&sc->plan, plan, sizeof(pars_plan));
	const int lr = lower_plan(sc->plan, sc->hplan);
	if (lr != 0) {
		delete sc;
		return fail(PARS_ERR_PLAN, "plan rejected (code %d): descriptor tables exceed the plan.hpp caps", lr);
	}
	int r;
	if ((r = sc->dplan.ensure(sizeof(DevPlan))) || (r = sc->ticket.ensure(64, true))
		|| (r = sc->summary.ensure(sizeof(DevSummary), true))) {
		delete sc;
		return r;
	}
	hipError_t e = hipMemcpy(sc->dplan.p, &sc->hplan, sizeof(DevPlan), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = hipHostMalloc((void**)&sc->h_sum, sizeof(DevSummary), hipHostMallocDefault);
	for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreate(&sc->ev[i]);
	if (e != hipSuccess) {
		delete sc;
		return hip_fail(e, "scanner setup");
	}
	*out = sc;
	return PARS_OK;
}

int pars_hip_scanner_destroy(pars_scanner* sc) {
	delete sc;
	return PARS_OK;
}

int pars_hip_set_option(pars_scanner* sc, int option, int64_t value) {
	if (!sc)
		return fail(PARS_ERR_ARGUMENT, "null scanner");
	switch (option) {
	case PARS_OPT_WANT_CHARS: sc->want_chars = value != 0; return PARS_OK;
	case PARS_OPT_WANT_BITMAP: sc->want_bitmap = value != 0; return PARS_OK;
	case PARS_OPT_PATH:
		if (value < 0 || value > 1)
			return fail(PARS_ERR_ARGUMENT, "path must be 0 (auto) or 1 (general)");
		sc->path = (int)value;
		return PARS_OK;
	case PARS_OPT_ROWS_PER_TILE:
		if (value < 0 || value > 4)
			return fail(PARS_ERR_ARGUMENT, "rows per tile must be 0 (default), 1, 2, 3 or 4");
		sc->rows = (int)value;
		return PARS_OK;
	case PARS_OPT_GRAPH: sc->use_graph = value != 0; return PARS_OK;
	case PARS_OPT_LOOKBACK_WINDOW:
		if (value != 0 && (value < 8 || value > 64))
			return fail(PARS_ERR_ARGUMENT, "look-back window must be 0 (default) or 8..64");
		sc->lb_window = (int)value;
		return PARS_OK;
	case PARS_OPT_BLOCKS_PER_CU:
		if (value < 0 || value > 64)
			return fail(PARS_ERR_ARGUMENT, "blocks per CU must be 0 (as many as fit) .. 64");
		sc->blocks_per_cu = (int)value;
		return PARS_OK;
	case PARS_OPT_RADIX_BITS:
		if (value < 0 || value > 8)
			return fail(PARS_ERR_ARGUMENT, "radix digit bits must be 0 (default 8) .. 8");
		sc->radix_bits = (int)value;
		return PARS_OK;
	case PARS_OPT_CHASER: sc->chase = value != 0; return PARS_OK;
	case PARS_OPT_LAST_PAIRING: return fail(PARS_ERR_ARGUMENT, "option %d is read-only", option);
	case PARS_OPT_WAVES_PER_BLOCK:
		if (value != 0 && value != 4 && value != 6 && value != 8)
			return fail(PARS_ERR_ARGUMENT, "waves per block must be 0 (default 4), 4, 6 or 8");
		sc->waves = (int)value;
		return PARS_OK;
	}
	return fail(PARS_ERR_ARGUMENT, "unknown option %d", option);
}

int pars_hip_scan_device(pars_scanner* sc, const uint8_t* d_in, size_t n, pars_stream stream, int want_pairs,
						 int run_pairs, pars_stage_times* st, pars_device_result* out) {
	if (!sc || !out || (n && !d_in))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	if ((uint64_t)n > 0xFFFFFFFFull)
		return fail(PARS_ERR_ARGUMENT, "documents are limited to 2^32-1 bytes (uint32 positions)");
	hipStream_t s = (hipStream_t)stream;
	if (st)
		*st = pars_stage_times {0, 0, 0, 0};
	if (n == 0) {
		std::memset(out, 0, sizeof *out);
		out->balanced = out->well_formed = 1;
		return PARS_OK;
	}
	const uint64_t n_tiles = (n + kTile - 1) / kTile;
	if (!sc->br_hint)
		sc->br_hint = std::max<uint64_t>(n / 8, 1u << 16);
	uint64_t cap_br = std::min<uint64_t>(std::max<uint64_t>(sc->br_hint, sc->cap_br), n);
	int		 r;
	if ((r = ensure_outputs(sc, n_tiles, std::max<uint64_t>(n, sc->cap_struct), cap_br, want_pairs != 0)))
		return r;
	if (sc->want_bitmap && (r = sc->bitmap.ensure(((n + 63) / 64) * 8)))
		return r;
	for (int attempt = 0; attempt < 2; ++attempt) {
		if ((r = enqueue_document(sc, d_in, n, s, want_pairs != 0, run_pairs != 0, st != nullptr)))
			return r;
		if ((r = read_summary(sc, s)))
			return r;
		if (sc->h_sum->n_brackets <= sc->cap_br)
			break;
		// more brackets than the optimistic capacity: grow exactly and rerun
		sc->br_hint = sc->h_sum->n_brackets;
		if ((r = ensure_outputs(sc, n_tiles, sc->cap_struct, sc->h_sum->n_brackets, want_pairs != 0)))
			return r;
	}
	if (run_pairs && sc->h_sum->n_brackets > (uint32_t)INT32_MAX)	// pairs are int32 bracket indices
		return fail(PARS_ERR_ARGUMENT, "pairing holds at most 2^31-1 brackets (%u here): scan without pairs",
					sc->h_sum->n_brackets);
	if (st) {
		float ms = 0;
		if (hipEventElapsedTime(&ms, sc->ev[0], sc->ev[1]) == hipSuccess)
			st->p1 = ms;
		if (hipEventElapsedTime(&ms, sc->ev[1], sc->ev[2]) == hipSuccess)
			st->p4 = ms;
	}
	if ((r = finish_pairs(sc, want_pairs != 0, s, st ? &st->p4 : nullptr)))
		return r;
	fill_result(sc, n, want_pairs != 0, run_pairs != 0, out);
	return PARS_OK;
}

int pars_hip_scan_host(pars_scanner* sc, const uint8_t* h_in, size_t n, pars_stream stream, int want_pairs,
					   pars_host_result* out) {
	if (!sc || !out || (n && !h_in))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	hipStream_t s = (hipStream_t)stream;
	int			r;
	if ((r = ensure_pinned((void**)&sc->h_in, sc->h_in_cap, n)) || (r = sc->d_in.ensure(std::max<size_t>(n, 64))))
		return r;
	if (n) {
		std::memcpy(sc->h_in, h_in, n);
		hipError_t e = hipMemcpyAsync(sc->d_in.p, sc->h_in, n, hipMemcpyHostToDevice, s);
		if (e != hipSuccess)
			return hip_fail(e, "H2D");
	}
	pars_device_result dr;
	if ((r = pars_hip_scan_device(sc, sc->d_in.as<uint8_t>(), n, stream, want_pairs, 1, nullptr, &dr)))
		return r;
	std::memset(out, 0, sizeof *out);
	out->n_struct	   = dr.n_struct;
	out->balanced	   = dr.balanced;
	out->well_formed   = dr.well_formed;
	out->max_depth	   = dr.max_depth;
	out->unclosed_mode = dr.unclosed_mode;
	if ((r = ensure_pinned((void**)&sc->h_struct, sc->h_struct_cap, (size_t)dr.n_struct * 4)))
		return r;
	out->structural = sc->h_struct;
	hipError_t e	= hipSuccess;
	if (dr.n_struct)
		e = hipMemcpyAsync(sc->h_struct, dr.structural, (size_t)dr.n_struct * 4, hipMemcpyDeviceToHost, s);
	if (e == hipSuccess && want_pairs) {
		if ((r = ensure_pinned((void**)&sc->h_pairs, sc->h_pairs_cap, (size_t)dr.n_struct * 4)))
			return r;
		out->pair_pos = sc->h_pairs;
		if (dr.n_struct)
			e = hipMemcpyAsync(sc->h_pairs, dr.pair_pos, (size_t)dr.n_struct * 4, hipMemcpyDeviceToHost, s);
	}
	if (e == hipSuccess)
		e = hipStreamSynchronize(s);
	return e == hipSuccess ? PARS_OK : hip_fail(e, "D2H");
}

int pars_hip_reserve_struct(pars_scanner* sc, size_t cap) {
	if (!sc)
		return fail(PARS_ERR_ARGUMENT, "null scanner");
	const uint64_t c = std::max<uint64_t>(cap, 1);
	int			   r = ensure_outputs(sc, std::max<uint64_t>(sc->recs.cap / (kRecWords * 8), 1),
								  std::max<uint64_t>(c, sc->cap_struct), std::max<uint64_t>(c, sc->cap_br), false);
	if (r)
		return r;
	if (sc->want_chars && (r = sc->chars.ensure(sc->cap_struct)))
		return r;
	return PARS_OK;
}

int pars_hip_scan_batch(pars_scanner* sc, const uint8_t* d_in, size_t n, const pars_doc_span* docs, size_t n_docs,
						pars_stream stream, int want_pairs, int run_pairs, pars_stage_times* st,
						pars_batch_result* out) {
	if (!sc || !out || !docs || !n_docs || !d_in)
		return fail(PARS_ERR_ARGUMENT, "null argument or empty batch");
	if ((uint64_t)n > 0xFFFFFFFFull)
		return fail(PARS_ERR_ARGUMENT, "a batch is limited to 2^32-1 bytes (uint32 output indices)");
	uint64_t prev_end = 0;
	for (size_t d = 0; d < n_docs; ++d) {
		const pars_doc_span& x = docs[d];
		if (x.offset % 64 || x.size == 0 || x.offset < prev_end || x.offset + x.size > n)
			return fail(PARS_ERR_ARGUMENT,
						"document %zu: spans must start on 64-byte boundaries, be non-empty, ascending, "
						"non-overlapping and inside the buffer",
						d);
		prev_end = x.offset + x.size;
	}
	hipStream_t s = (hipStream_t)stream;
	if (st)
		*st = pars_stage_times {0, 0, 0, 0};
	if (!sc->br_hint)
		sc->br_hint = std::max<uint64_t>(n / 8, 1u << 16);
	uint64_t cap_br = std::min<uint64_t>(std::max<uint64_t>(sc->br_hint, sc->cap_br), std::max<uint64_t>(n, 1));
	int		 r;
	if ((r = ensure_batch(sc, n, n_docs, cap_br, want_pairs != 0)))
		return r;
	for (size_t d = 0; d < n_docs; ++d) {
		tab_off(sc)[d]			= docs[d].offset;
		tab_len(sc, n_docs)[d] = (uint32_t)docs[d].size;
	}
	const bool graph = sc->use_graph && s != nullptr;
	for (int attempt = 0; attempt < 2; ++attempt) {
		if (graph) {
			hipError_t e;
			if (st && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
				return hip_fail(e, "hipEventRecord");
			if ((r = run_batch_graph(sc, d_in, n, n_docs, s, want_pairs != 0, run_pairs != 0)))
				return r;
			if (st && (e = hipEventRecord(sc->ev[2], s)) != hipSuccess)
				return hip_fail(e, "hipEventRecord");
		} else if ((r = enqueue_batch(sc, d_in, n, n_docs, s, want_pairs != 0, run_pairs != 0, st != nullptr))) {
			return r;
		}
		if ((r = read_summary(sc, s)))
			return r;
		if (sc->h_sum->n_brackets <= sc->cap_br)
			break;
		sc->br_hint = sc->h_sum->n_brackets;  // grow exactly and rerun
		if ((r = ensure_batch(sc, n, n_docs, sc->h_sum->n_brackets, want_pairs != 0)))
			return r;
	}
	if (sc->h_sum->overflow & 4u)  // more than 2^32 distinct (document, depth) keys
		return fail(PARS_ERR_ARGUMENT, "batch pairing needs more than 2^32 (document, depth) keys; split the batch");
	if (sc->h_sum->overflow & kStackLost)
		return fail(PARS_ERR_DEVICE, "stack pairing found an unmatched residue on a balanced batch (internal error)");
	if (st) {
		float ms = 0;
		if (graph) {
			if (hipEventElapsedTime(&ms, sc->ev[0], sc->ev[2]) == hipSuccess)
				st->p1 = ms;
		} else {
			if (hipEventElapsedTime(&ms, sc->ev[0], sc->ev[1]) == hipSuccess)
				st->p1 = ms;
			if (hipEventElapsedTime(&ms, sc->ev[1], sc->ev[2]) == hipSuccess)
				st->p4 = ms;
		}
	}
	float sort_ms = 0;
	if (sc->batch_defers_sort) {  // the stack path was tried on the device: did it pair the batch?
		if (!sc->h_sum->small_done) {  // some document is unbalanced: the sort pairs the batch
			BatchArgs b = make_batch_args(sc, n, n_docs);
			b.tile_rows = batch_tile_rows(sc->hplan, sc->path, sc->rows);
			PairArgs  p = make_batch_pair_args(sc, b, want_pairs != 0);
			hipError_t e;
			if (st && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
				return hip_fail(e, "hipEventRecord");
			if ((e = launch_pairing(p, s)) != hipSuccess)
				return hip_fail(e, "pairing launch");
			if ((e = launch_finish(sc, b, true, s)) != hipSuccess)
				return hip_fail(e, "batch finish");
			if (st && (e = hipEventRecord(sc->ev[2], s)) != hipSuccess)
				return hip_fail(e, "hipEventRecord");
			if ((r = read_summary(sc, s)))
				return r;
			if (sc->h_sum->overflow & 4u)
				return fail(PARS_ERR_ARGUMENT,
							"batch pairing needs more than 2^32 (document, depth) keys; split the batch");
			if (st)
				(void)hipEventElapsedTime(&sort_ms, sc->ev[1], sc->ev[2]);
		}
	}
	if (st && sort_ms > 0) {  // the deferred sort belongs to R6 (and to the chain's wall time)
		st->p4 += sort_ms;
		if (graph)
			st->p1 += sort_ms;
	}
	std::memset(out, 0, sizeof *out);
	out->n_docs		   = n_docs;
	out->n_struct	   = sc->h_sum->n_struct;
	out->n_brackets	   = sc->h_sum->n_brackets;
	out->structural	   = sc->positions.as<uint32_t>();
	out->struct_start  = sc->st_start.as<uint32_t>();
	out->bracket_pos   = sc->bpos.as<uint32_t>();
	out->bracket_char  = sc->bchar.as<uint8_t>();
	out->bracket_start = sc->br_start.as<uint32_t>();
	out->bracket_pair  = run_pairs ? sc->pair.as<int32_t>() : nullptr;
	out->pair_pos	   = want_pairs ? sc->pair_pos.as<int32_t>() : nullptr;
	out->doc_flags	   = sc->doc_flags.as<uint8_t>();
	out->max_depth	   = sc->doc_max.as<int32_t>();
	out->struct_bitmap = sc->want_bitmap ? sc->bitmap.as<uint64_t>() : nullptr;
	return PARS_OK;
}


int pars_hip_scan_device_noalloc(pars_scanner* sc, const uint8_t* d_in, size_t n, pars_stream stream, int run_pairs,
								 uint32_t* positions, size_t cap_struct, uint64_t* struct_bitmap,
								 uint32_t* bracket_pos, uint8_t* bracket_char, int32_t* bracket_pair,
								 size_t cap_brackets, pars_device_result* out) {
	if (!sc || !out || (n && !d_in))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	if ((uint64_t)n > kNaMaxBytes)
		return fail(PARS_ERR_ARGUMENT, "pars_hip_scan_device_noalloc takes documents of at most %llu bytes (%zu given)",
					(unsigned long long)kNaMaxBytes, n);
	if ((uint64_t)cap_brackets > kNaMaxBrackets)
		return fail(PARS_ERR_ARGUMENT, "pars_hip_scan_device_noalloc takes at most %llu brackets (cap_brackets %zu)",
					(unsigned long long)kNaMaxBrackets, cap_brackets);
	if (n && (!positions || !bracket_pos || !bracket_char || (run_pairs && !bracket_pair)))
		return fail(PARS_ERR_ARGUMENT, "positions, bracket_pos, bracket_char (and bracket_pair with run_pairs) are required");
	std::memset(out, 0, sizeof *out);
	if (n == 0) {
		out->balanced = out->well_formed = 1;
		return PARS_OK;
	}
	uint8_t* S = nullptr;
	int		 r;
	if ((r = na_scratch(&S)))
		return r;
	hipStream_t s	  = (hipStream_t)stream;
	DevSummary* d_sum = na_at<DevSummary>(S, kNaSum);
	ScanArgs	a {};
	a.data			= d_in;
	a.n				= n;
	a.n_tiles		= (uint32_t)((n + kTile - 1) / kTile);
	a.plan			= sc->dplan.as<DevPlan>();
	a.recs			= na_at<uint64_t>(S, kNaRecs);
	a.ticket		= na_at<uint32_t>(S, kNaTicket);
	a.done			= a.ticket + 1;
	a.sum			= d_sum;
	a.positions		= positions;
	a.bitmap		= struct_bitmap;
	a.bpos			= bracket_pos;
	a.bsidx			= na_at<uint32_t>(S, kNaBr + kNaBsidx * kNaBrArr);
	a.bkey			= na_at<int32_t>(S, kNaBr + kNaBkey * kNaBrArr);
	a.bchar			= bracket_char;
	a.cap_struct	= cap_struct;
	a.cap_br		= cap_brackets;
	a.in_states		= sc->hplan.start_packed;
	a.lb_window		= (uint32_t)sc->lb_window;
	a.blocks_per_cu = (uint32_t)sc->blocks_per_cu;
	a.chase			= (uint32_t)sc->chase;
	hipError_t e;
	if ((e = launch_summary_reset(d_sum, SummaryInit {}, s)) != hipSuccess)
		return hip_fail(e, "summary reset");
	if ((e = launch_scan(sc, a, s)) != hipSuccess)
		return hip_fail(e, "scan kernel launch");
	bool stack = false;
	if (run_pairs && cap_brackets) {
		const uint32_t small = small_pairing_mode(n, 1);
		if (!sc->radix_bits && small != 2) {
			StackArgs k {};
			k.sum		 = d_sum;
			k.plan		 = sc->dplan.as<DevPlan>();
			k.bchar		 = bracket_char;
			k.bkey		 = a.bkey;
			k.bsidx		 = a.bsidx;
			k.cap_br	 = cap_brackets;
			k.pair		 = bracket_pair;
			k.res		 = na_at<uint32_t>(S, kNaBr + kNaRes * kNaBrArr);
			k.crec		 = na_at<int32_t>(S, kNaCrec);
			k.gmin1		 = na_at<unsigned long long>(S, kNaG1);
			k.gmin2		 = na_at<unsigned long long>(S, kNaG2);
			k.max_chunks = std::max<uint64_t>((cap_brackets + kStkChunk - 1) / kStkChunk, 1);
			e			 = launch_stack_pairing(k, s);
			stack		 = true;
		} else {
			PairArgs p = na_pair_args(sc, S, cap_brackets, bracket_char, bracket_pair);
			p.small	   = sc->radix_bits ? 0u : small;
			e		   = launch_pairing(p, s);
		}
		if (e != hipSuccess)
			return hip_fail(e, "pairing launch");
	}
	if ((r = na_read_summary(sc, d_sum, s)))
		return r;
	const DevSummary& h = *sc->h_sum;
	if (h.n_struct > cap_struct || h.n_brackets > cap_brackets)
		return fail(PARS_ERR_ARGUMENT,
					"outputs too small: the document has %u structural and %u bracket bytes (capacities %zu, %zu)",
					h.n_struct, h.n_brackets, cap_struct, cap_brackets);
	if (stack && !(h.min_bal >= 0 && h.last_bal == 0)) {  // unbalanced: the sort's pairs
		PairArgs p = na_pair_args(sc, S, cap_brackets, bracket_char, bracket_pair);
		if ((e = launch_pairing(p, s)) != hipSuccess)
			return hip_fail(e, "pairing launch");
		if ((r = na_read_summary(sc, d_sum, s)))
			return r;
	}
	out->structural	   = positions;
	out->pair_pos	   = nullptr;
	out->n_struct	   = h.n_struct;
	out->balanced	   = (h.min_bal >= 0) && (h.last_bal == 0);
	out->well_formed   = out->balanced && (h.n_brackets % 2 == 0) && !(run_pairs && h.mismatched);
	out->max_depth	   = h.max_bal;
	out->chars		   = nullptr;
	out->struct_bitmap = struct_bitmap;
	out->bracket_pos   = bracket_pos;
	out->bracket_char  = bracket_char;
	out->bracket_pair  = run_pairs ? bracket_pair : nullptr;
	out->n_brackets	   = h.n_brackets;
	const uint32_t nd  = sc->hplan.n_descs;
	const uint32_t m   = nd ? ((1u << (3 * nd)) - 1u) : 0u;
	out->unclosed_mode = (h.final_states & m) != (sc->hplan.start_packed & m);
	return PARS_OK;
}

const uint32_t* pars_hip_struct_ptr(const pars_scanner* sc) {
	return sc ? static_cast<const uint32_t*>(sc->positions.p) : nullptr;
}

int pars_hip_scan_chunk(pars_scanner* sc, const uint8_t* d_chunk, size_t n, const pars_chunk_carry* in,
						pars_stream stream, size_t byte_base, pars_chunk_carry* out) {
	if (!sc || !in || !out || (n && !d_chunk))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	if ((uint64_t)n > 0xFFFFFFFFull || (uint64_t)byte_base + n > 0xFFFFFFFFull)
		return fail(PARS_ERR_ARGUMENT, "chunks of a document end at or below 2^32-1 bytes (uint32 positions)");
	hipStream_t		 s	 = (hipStream_t)stream;
	pars_chunk_carry c	 = *in;
	const bool		 pad = (c.scan >> 31) & 1u;
	if (!pad) {	 // document start
		c.depth = c.min_depth = c.max_depth = 0;
	}
	if (n == 0) {
		*out = c;
		return PARS_OK;
	}
	const uint64_t n_tiles = (n + kTile - 1) / kTile;
	const uint64_t need	   = (uint64_t)c.out_base + n;
	if (need > sc->cap_struct)
		return fail(PARS_ERR_ARGUMENT, "chunk output past reserve_struct capacity (%llu > %llu)",
					(unsigned long long)need, (unsigned long long)sc->cap_struct);
	int r;
	if ((r = ensure_outputs(sc, n_tiles, sc->cap_struct, sc->cap_br, false)))
		return r;
	ScanArgs a	 = make_scan_args(sc, d_chunk, n, false);
	a.bitmap	 = nullptr;	 // the per-document bitmap is not chunk-addressed
	a.bsidx		 = sc->bsidx.as<uint32_t>();  // pair_chunks may be asked for pair_pos
	// R6 over the chunks starts on the stack path unless a digit cap or the tape
	// asks for the sort; a sort after sparse keys fills them first
	a.key_sparse	= !sc->radix_bits && !sc->force_radix;
	sc->keys_sparse = (c.bracket_base != 0 && sc->keys_sparse) || a.key_sparse;
	a.byte_base	 = byte_base;
	a.out_base	 = c.out_base;
	a.br_base	 = c.bracket_base;
	a.depth_base = c.depth;
	a.in_states	 = pad ? (uint32_t)(c.scan & 0xFFFFFFu) : sc->hplan.start_packed;
	SummaryInit si;
	si.n_struct	  = c.out_base;
	si.n_brackets = c.bracket_base;
	si.depth	  = c.depth;
	si.min_depth  = c.min_depth;
	si.max_depth  = c.max_depth;
	hipError_t e  = launch_summary_reset(a.sum, si, s);
	if (e != hipSuccess)
		return hip_fail(e, "summary reset");
	if ((e = launch_scan(sc, a, s)) != hipSuccess)
		return hip_fail(e, "scan kernel launch");
	if ((r = read_summary(sc, s)))
		return r;
	const DevSummary& h = *sc->h_sum;
	out->scan			= (1ull << 31) | (uint64_t)(h.final_states & 0xFFFFFFu);
	out->out_base		= h.n_struct;
	out->bracket_base	= h.n_brackets;
	out->depth			= h.last_bal;
	out->min_depth		= h.min_bal;
	out->max_depth		= h.max_bal;
	out->reserved		= 0;
	return PARS_OK;
}

int pars_hip_pair_chunks(pars_scanner* sc, const pars_chunk_carry* last, pars_stream stream, int want_pairs,
						 pars_device_result* out) {
	if (!sc || !last || !out)
		return fail(PARS_ERR_ARGUMENT, "null argument");
	if (last->bracket_base > sc->cap_br || last->out_base > sc->cap_struct)
		return fail(PARS_ERR_ARGUMENT, "carry exceeds the reserved capacity");
	hipStream_t s = (hipStream_t)stream;
	int			r;
	if (want_pairs && (r = sc->pair_pos.ensure((size_t)sc->cap_struct * 4)))
		return r;
	SummaryInit si;
	si.n_struct	  = last->out_base;
	si.n_brackets = last->bracket_base;
	si.depth	  = last->depth;
	si.min_depth  = last->min_depth;
	si.max_depth  = last->max_depth;
	hipError_t e  = launch_summary_reset(sc->summary.as<DevSummary>(), si, s);
	if (e != hipSuccess)
		return hip_fail(e, "summary reset");
	if (want_pairs && last->out_base) {
		if ((e = launch_fill_i32(sc->pair_pos.as<int32_t>(), last->out_base, -1, s)) != hipSuccess)
			return hip_fail(e, "pair_pos fill");
	}
	sc->stack_pending = false;
	if ((e = enqueue_pairs(sc, std::max<uint64_t>(last->bracket_base, (uint64_t)kSmallBrackets + 1), want_pairs != 0, s))
		!= hipSuccess)
		return hip_fail(e, "pairing launch");
	if ((r = read_summary(sc, s)) || (r = finish_pairs(sc, want_pairs != 0, s, nullptr)))
		return r;
	fill_result(sc, 1, want_pairs != 0, true, out);
	out->unclosed_mode = ((last->scan & 0xFFFFFFu) & ((1u << (3 * sc->hplan.n_descs)) - 1u))
					  != (sc->hplan.start_packed & ((1u << (3 * sc->hplan.n_descs)) - 1u));
	out->struct_bitmap = nullptr;
	return PARS_OK;
}


int pars_hip_get_option(const pars_scanner* sc, int option, int64_t* value) {
	if (!sc || !value)
		return fail(PARS_ERR_ARGUMENT, "null argument");
	switch (option) {
	case PARS_OPT_WANT_CHARS: *value = sc->want_chars; return PARS_OK;
	case PARS_OPT_WANT_BITMAP: *value = sc->want_bitmap; return PARS_OK;
	case PARS_OPT_PATH: *value = sc->path; return PARS_OK;
	case PARS_OPT_ROWS_PER_TILE: *value = sc->rows; return PARS_OK;
	case PARS_OPT_GRAPH: *value = sc->use_graph; return PARS_OK;
	case PARS_OPT_LOOKBACK_WINDOW: *value = sc->lb_window; return PARS_OK;
	case PARS_OPT_BLOCKS_PER_CU: *value = sc->blocks_per_cu; return PARS_OK;
	case PARS_OPT_RADIX_BITS: *value = sc->radix_bits; return PARS_OK;
	case PARS_OPT_CHASER: *value = sc->chase; return PARS_OK;
	case PARS_OPT_WAVES_PER_BLOCK: *value = sc->waves; return PARS_OK;
	case PARS_OPT_LAST_PAIRING: {
		// read from the last call's final summary: a sort that ran set its pass
		// count; the one-block path (and a batch's stack path) set small_done
		const DevSummary& h = *sc->h_sum;
		*value				= !sc->pairs_ran			? PARS_PAIRING_NONE
							  : h.sort_passes > 0		? PARS_PAIRING_SORT
							  : sc->stack_launched		? PARS_PAIRING_STACK
							  : h.small_done			? PARS_PAIRING_ONE_BLOCK
														: PARS_PAIRING_NONE;
		return PARS_OK;
	}
	}
	return fail(PARS_ERR_ARGUMENT, "unknown option %d", option);
}

// Coordinate descent over the launch knobs, each candidate timed on the
// caller's input with the scanner's own stage events (scan kernel for the
// scan knobs, the pairing chain for the digit cap).
int pars_hip_autotune(pars_scanner* sc, const uint8_t* d_in, size_t n, const pars_doc_span* docs, size_t n_docs,
					  pars_stream stream, int reps, pars_tune_result* out) {
	if (!sc || !out || !d_in || !n || (docs && !n_docs))
		return fail(PARS_ERR_ARGUMENT, "autotune needs a scanner, an input and a result");
	reps = std::min(std::max(reps, 1), 15);
	std::memset(out, 0, sizeof *out);
	const int saved_graph = sc->use_graph;
	sc->use_graph		  = 0;	// stage times split scan and pairing only eagerly
	uint64_t n_brackets	  = 0;
	// median stage times of `reps` runs after one warm-up; failing settings
	// (a digit cap too narrow for the key range) report a status
	auto measure = [&](float& scan_ms, float& pair_ms) -> int {
		std::vector<float> a, b;
		for (int i = 0; i <= reps; ++i) {
			pars_stage_times st {};
			int				 r;
			if (docs) {
				pars_batch_result br;
				r			= pars_hip_scan_batch(sc, d_in, n, docs, n_docs, stream, 0, 1, &st, &br);
				n_brackets = br.n_brackets;
			} else {
				pars_device_result dr;
				r			= pars_hip_scan_device(sc, d_in, n, stream, 0, 1, &st, &dr);
				n_brackets = sc->h_sum->n_brackets;
			}
			if (r)
				return r;
			++out->trials;
			if (i) {
				a.push_back(st.p1);
				b.push_back(st.p4);
			}
		}
		std::sort(a.begin(), a.end());
		std::sort(b.begin(), b.end());
		scan_ms = a[a.size() / 2];
		pair_ms = b[b.size() / 2];
		return PARS_OK;
	};
	int	  r;
	float best_scan = 0, best_pair = 0;
	if ((r = measure(best_scan, best_pair))) {
		sc->use_graph = saved_graph;
		return r;
	}
	out->base_scan_ms = best_scan;
	out->base_pair_ms = best_pair;
	// one knob at a time; `field` keeps the best value seen
	auto sweep = [&](int& field, const std::vector<int>& values, bool pairing) {
		const int keep = field;
		int		  best = keep;
		for (int v : values) {
			if (v == keep)
				continue;
			field = v;
			float sm = 0, pm = 0;
			if (measure(sm, pm))
				continue;  // not usable for this input
			if (pairing ? pm < best_pair * 0.98f : sm < best_scan * 0.98f) {
				best	  = v;
				best_scan = pairing ? best_scan : sm;
				best_pair = pairing ? pm : best_pair;
			}
		}
		field = best;
	};
	const bool one_desc = sc->hplan.n_descs == 1;  // several descriptors: one row per tile always
	if (one_desc)
		sweep(sc->rows, {1, 2, 3, 4}, false);
	sweep(sc->lb_window, {16, 32, 48, 64}, false);
	// resident blocks per CU below the kernel's occupancy limit
	float sm = 0, pm = 0;
	if ((r = measure(sm, pm)) == PARS_OK) {
		const int occ			= std::max(last_scan_occupancy(), 1);
		out->max_blocks_per_cu = (uint32_t)occ;
		std::vector<int> bpc;
		for (int b = 1; b < occ; ++b) bpc.push_back(b);
		sweep(sc->blocks_per_cu, bpc, false);
	}
	// waves per block: one document through a bit-parallel evaluator (batches
	// and the descriptor walks keep four); the launcher caps blocks per CU at
	// what the wider blocks can keep resident
	if (!docs && sc->path == 0 && sc->hplan.fast_kind != kFastNone)
		sweep(sc->waves, {6, 8}, false);
	if (n_brackets)
		sweep(sc->radix_bits, {8, 7, 6, 5, 4}, true);
	if ((r = measure(sm, pm)) == PARS_OK) {
		out->scan_ms = sm;
		out->pair_ms = pm;
	}
	sc->use_graph		= saved_graph;
	out->rows			= sc->rows;
	out->lb_window		= sc->lb_window;
	out->blocks_per_cu	= sc->blocks_per_cu;
	out->radix_bits		= sc->radix_bits;
	out->waves_per_block = sc->waves;
	return r;
}


// Streaming host boundary.  Chunk k's bytes go host -> pinned staging (k % 2)
// by `copy_threads` threads, then up to device chunk (k % 2) on the copy
// stream, and are scanned on the call's stream with the carry chained on the
// device; the host only waits for staging buffer k % 2 to be read (chunk
// k - 2's upload), so host copies, PCIe transfers and kernels overlap.  R6
// runs once over the whole document at the end.
int pars_hip_scan_stream(pars_scanner* sc, const uint8_t* h_in, size_t n, size_t chunk_bytes, int copy_threads,
						 pars_stream stream, int want_pairs, pars_stage_times* st, pars_device_result* out) {
	if (!sc || !out || (n && !h_in))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	if ((uint64_t)n > 0xFFFFFFFFull)
		return fail(PARS_ERR_ARGUMENT, "documents are limited to 2^32-1 bytes (uint32 positions)");
	if (st)
		*st = pars_stage_times {0, 0, 0, 0};
	if (n == 0) {
		std::memset(out, 0, sizeof *out);
		out->balanced = out->well_formed = 1;
		return PARS_OK;
	}
	hipStream_t	   s	 = (hipStream_t)stream;
	const uint64_t unit	 = 64 << 10;  // chunks start on lane and bitmap-word boundaries
	uint64_t	   chunk = chunk_bytes ? chunk_bytes : (64ull << 20);
	chunk				 = std::max<uint64_t>(unit, (chunk + unit - 1) / unit * unit);
	chunk				 = std::min<uint64_t>(chunk, (n + unit - 1) / unit * unit);
	const int	   thr	 = std::min(std::max(copy_threads ? copy_threads : 8, 1), 16);
	const uint64_t n_chunks = (n + chunk - 1) / chunk;
	int			   r;
	hipError_t	   e;
	if (!sc->copy_stream && (e = hipStreamCreateWithFlags(&sc->copy_stream, hipStreamNonBlocking)) != hipSuccess)
		return hip_fail(e, "copy stream");
	for (int i = 0; i < 2; ++i) {
		if (!sc->ev_h2d[i] && (e = hipEventCreateWithFlags(&sc->ev_h2d[i], hipEventDisableTiming)) != hipSuccess)
			return hip_fail(e, "event");
		if (!sc->ev_done[i] && (e = hipEventCreateWithFlags(&sc->ev_done[i], hipEventDisableTiming)) != hipSuccess)
			return hip_fail(e, "event");
	}
	if (chunk > sc->h_chunk_cap) {
		for (int i = 0; i < 2; ++i) {
			if (sc->h_chunk[i])
				(void)hipHostFree(sc->h_chunk[i]);
			sc->h_chunk[i] = nullptr;
		}
		sc->h_chunk_cap = 0;
		for (int i = 0; i < 2; ++i)
			if ((e = hipHostMalloc((void**)&sc->h_chunk[i], chunk, hipHostMallocDefault)) != hipSuccess)
				return fail(PARS_ERR_MEMORY, "hipHostMalloc(%llu): %s", (unsigned long long)chunk, hipGetErrorString(e));
		sc->h_chunk_cap = chunk;
	}
	if ((r = sc->d_chunk[0].ensure(chunk)) || (r = sc->d_chunk[1].ensure(chunk)) || (r = sc->carry.ensure(64)))
		return r;
	const uint64_t n_tiles = (chunk + kTile - 1) / kTile;
	if (!sc->br_hint)
		sc->br_hint = std::max<uint64_t>(n / 8, 1u << 16);
	uint64_t cap_br = std::min<uint64_t>(std::max<uint64_t>(sc->br_hint, sc->cap_br), n);
	if ((r = ensure_outputs(sc, n_tiles, std::max<uint64_t>(n, sc->cap_struct), cap_br, want_pairs != 0)))
		return r;
	if (sc->want_bitmap && (r = sc->bitmap.ensure(((n + 63) / 64) * 8)))
		return r;
	// host copy of one chunk into pinned staging, split over threads
	auto stage = [&](uint8_t* dst, const uint8_t* src, uint64_t len) {
		const uint64_t per = ((len + thr - 1) / thr + 4095) / 4096 * 4096;
		if (thr == 1 || len <= per) {
			std::memcpy(dst, src, len);
			return;
		}
		std::vector<std::thread> ts;
		for (uint64_t off = per; off < len; off += per)
			ts.emplace_back([=] { std::memcpy(dst + off, src + off, std::min<uint64_t>(per, len - off)); });
		std::memcpy(dst, src, std::min<uint64_t>(per, len));
		for (std::thread& t : ts) t.join();
	};
	for (int attempt = 0; attempt < 2; ++attempt) {
		if ((e = launch_summary_reset(sc->summary.as<DevSummary>(), SummaryInit {}, s)) != hipSuccess)
			return hip_fail(e, "summary reset");
		if (st && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
			return hip_fail(e, "hipEventRecord");
		for (uint64_t k = 0; k < n_chunks; ++k) {
			const int	   slot = (int)(k & 1);
			const uint64_t off = k * chunk, len = std::min<uint64_t>(chunk, n - off);
			if (k >= 2 && (e = hipEventSynchronize(sc->ev_h2d[slot])) != hipSuccess)  // staging slot read
				return hip_fail(e, "staging wait");
			stage(sc->h_chunk[slot], h_in + off, len);
			if (k >= 2 && (e = hipStreamWaitEvent(sc->copy_stream, sc->ev_done[slot], 0)) != hipSuccess)
				return hip_fail(e, "chunk wait");
			if ((e = hipMemcpyAsync(sc->d_chunk[slot].p, sc->h_chunk[slot], len, hipMemcpyHostToDevice,
									sc->copy_stream))
					!= hipSuccess
				|| (e = hipEventRecord(sc->ev_h2d[slot], sc->copy_stream)) != hipSuccess
				|| (e = hipStreamWaitEvent(s, sc->ev_h2d[slot], 0)) != hipSuccess)
				return hip_fail(e, "chunk upload");
			if (k > 0 && (e = launch_carry_next(sc->summary.as<DevSummary>(), sc->carry.as<DevCarry>(), s)) != hipSuccess)
				return hip_fail(e, "carry");
			ScanArgs a		= make_scan_args(sc, sc->d_chunk[slot].as<uint8_t>(), len, want_pairs != 0);
			a.key_sparse	= stack_first(sc, n);
			sc->keys_sparse = a.key_sparse;
			a.byte_base		= off;
			a.bitmap	= sc->want_bitmap ? sc->bitmap.as<uint64_t>() + off / 64 : nullptr;
			a.carry		= k > 0 ? sc->carry.as<DevCarry>() : nullptr;
			if ((e = launch_scan(sc, a, s)) != hipSuccess)
				return hip_fail(e, "scan kernel launch");
			if ((e = hipEventRecord(sc->ev_done[slot], s)) != hipSuccess)
				return hip_fail(e, "hipEventRecord");
		}
		if (st && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
			return hip_fail(e, "hipEventRecord");
		if ((e = enqueue_pairs(sc, n, want_pairs != 0, s)) != hipSuccess)
			return hip_fail(e, "pairing launch");
		if (st && (e = hipEventRecord(sc->ev[2], s)) != hipSuccess)
			return hip_fail(e, "hipEventRecord");
		if ((r = read_summary(sc, s)))
			return r;
		if (sc->h_sum->n_brackets <= sc->cap_br)
			break;
		sc->br_hint = sc->h_sum->n_brackets;  // more brackets than the optimistic capacity: grow, stream again
		if ((r = ensure_outputs(sc, n_tiles, sc->cap_struct, sc->h_sum->n_brackets, want_pairs != 0)))
			return r;
	}
	if (sc->h_sum->n_brackets > (uint32_t)INT32_MAX)	// pairs are int32 bracket indices
		return fail(PARS_ERR_ARGUMENT, "pairing holds at most 2^31-1 brackets (%u here): scan without pairs",
					sc->h_sum->n_brackets);
	if (st) {
		float ms = 0;
		if (hipEventElapsedTime(&ms, sc->ev[0], sc->ev[1]) == hipSuccess)
			st->p1 = ms;
		if (hipEventElapsedTime(&ms, sc->ev[1], sc->ev[2]) == hipSuccess)
			st->p4 = ms;
	}
	if ((r = finish_pairs(sc, want_pairs != 0, s, st ? &st->p4 : nullptr)))
		return r;
	fill_result(sc, n, want_pairs != 0, true, out);
	return PARS_OK;
}

// ---------------------------------------------------------------------------
// The JSON tape.
// ---------------------------------------------------------------------------
int pars_hip_json_tape(pars_scanner* sc, const uint8_t* d_in, size_t n, pars_stream stream, pars_stage_times* st,
					   pars_json_tape* out) {
	if (!sc || !out || (n && !d_in))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	for (unsigned i = 0; i < sc->plan.n_punctuation; ++i)
		if (sc->plan.punctuation[i] != ',' && sc->plan.punctuation[i] != ':')
			return fail(PARS_ERR_ARGUMENT, "the JSON tape needs a JSON plan (punctuation within ',' ':')");
	for (unsigned i = 0; i < sc->plan.n_brackets; ++i) {
		const char o = sc->plan.brackets[i].open, c = sc->plan.brackets[i].close;
		if (!((o == '{' && c == '}') || (o == '[' && c == ']')))
			return fail(PARS_ERR_ARGUMENT, "the JSON tape needs a JSON plan (brackets within {} [])");
	}
	std::memset(out, 0, sizeof *out);
	out->fail_event = -1;
	if (st)
		*st = pars_stage_times {0, 0, 0, 0};
	if (n == 0)
		return PARS_OK;	 // no value: the hook's final check fails (ok = 0)
	hipStream_t		   s = (hipStream_t)stream;
	pars_device_result res;
	sc->force_radix = 1;
	int r			= pars_hip_scan_device(sc, d_in, n, stream, 0, 1, st, &res);
	sc->force_radix = 0;
	if (r)
		return r;
	const DevSummary& h = *sc->h_sum;
	if (h.small_done || h.sort_passes < 1 || h.sort_passes > 4)
		return fail(PARS_ERR_DEVICE, "the bracket sort did not run through the radix passes (internal error)");
	const uint64_t ns = res.n_struct, nb = res.n_brackets, ntok = ns + 1;
	// sorted bracket positions and parents are int32 in the tape kernels
	if (nb > (uint64_t)INT32_MAX)
		return fail(PARS_ERR_ARGUMENT, "the JSON tape handles at most 2^31-1 brackets (%llu here)",
					(unsigned long long)nb);
	const uint64_t n_tblk = (ntok + kDomTile - 1) / kDomTile, n_hblk = std::max<uint64_t>((nb + kDomTile - 1) / kDomTile, 1);
	const uint64_t cap_nodes = std::min<uint64_t>(n, 2 * ns + 1);
	for (DevBuf* b : {&sc->dom_gbeg, &sc->dom_gend, &sc->dom_gdec})
		if ((r = b->ensure(ntok * 4)))
			return r;
	if ((r = sc->dom_sub.ensure(((ntok + kDomSub - 1) / kDomSub) * 3 * 4)))
		return r;
	if ((r = sc->dom_tinfo.ensure(ntok)))
		return r;
	if ((r = sc->dom_top.ensure(n_tblk * 2 * sizeof(unsigned long long))) || (r = sc->dom_tblk.ensure(n_tblk * 3 * 4)) ||
		(r = sc->dom_hblk.ensure(n_hblk * 4)) || (r = sc->dom_parent.ensure(std::max<uint64_t>(nb, 1) * 4)) ||
		(r = sc->dom_sum.ensure(sizeof(DomSummary))) || (r = sc->tape_kind.ensure(cap_nodes)) ||
		(r = sc->tape_beg.ensure(cap_nodes * 4)) || (r = sc->tape_end.ensure(cap_nodes * 4)) ||
		(r = sc->tape_soff.ensure(cap_nodes * 4)) || (r = sc->tape_arena.ensure(n)))
		return r;
	if (!sc->h_dom && (r = ensure_pinned(reinterpret_cast<void**>(&sc->h_dom), sc->h_dom_cap, sizeof(DomSummary))))
		return r;
	DomArgs a {};
	a.data		= d_in;
	a.n			= n;
	a.positions = sc->positions.as<uint32_t>();
	a.n_struct	= ns;
	a.nb		= nb;
	a.plan		= sc->dplan.as<DevPlan>();
	a.sum		= sc->summary.as<DevSummary>();
	a.bchar		= sc->bchar.as<uint8_t>();
	a.bkey		= sc->bkey.as<int32_t>();
	a.bsidx		= sc->bsidx.as<uint32_t>();
	a.order		= ((h.sort_passes - 1) & 1) ? sc->order1.as<uint32_t>() : sc->order0.as<uint32_t>();
	a.tinfo		= sc->dom_tinfo.as<uint8_t>();
	a.gbeg		= sc->dom_gbeg.as<uint32_t>();
	a.gend		= sc->dom_gend.as<uint32_t>();
	a.gdec		= sc->dom_gdec.as<uint32_t>();
	a.sub		= sc->dom_sub.as<uint32_t>();
	a.top2		= sc->dom_top.as<unsigned long long>();
	a.tblk		= sc->dom_tblk.as<uint32_t>();
	a.n_tblk	= n_tblk;
	a.hblk		= sc->dom_hblk.as<int32_t>();
	a.n_hblk	= n_hblk;
	a.parent	= sc->dom_parent.as<int32_t>();
	a.close		= sc->dom_close.as<uint32_t>();
	a.ds		= sc->dom_sum.as<DomSummary>();
	a.kind		= sc->tape_kind.as<uint8_t>();
	a.nbeg		= sc->tape_beg.as<uint32_t>();
	a.nend		= sc->tape_end.as<uint32_t>();
	a.soff		= sc->tape_soff.as<uint32_t>();
	a.arena		= sc->tape_arena.as<uint8_t>();
	a.cap_nodes = cap_nodes;
	a.cap_arena = n;
	hipError_t e;
	if (st && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	if ((e = launch_json_tape(a, s)) != hipSuccess)
		return hip_fail(e, "tape launch");
	if (st && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	if ((e = hipMemcpyAsync(sc->h_dom, a.ds, sizeof(DomSummary), hipMemcpyDeviceToHost, s)) != hipSuccess)
		return hip_fail(e, "tape summary readback");
	if ((e = hipStreamSynchronize(s)) != hipSuccess)
		return hip_fail(e, "hipStreamSynchronize");
	const DomSummary& d = *sc->h_dom;
	if (d.n_nodes > cap_nodes || d.arena_bytes > n)
		return fail(PARS_ERR_DEVICE, "tape larger than its bound (internal error)");
	if (st) {
		float ms = 0;
		if (hipEventElapsedTime(&ms, sc->ev[0], sc->ev[1]) == hipSuccess)
			st->p2 = ms;
	}
	out->kind		 = a.kind;
	out->begin		 = a.nbeg;
	out->end		 = a.nend;
	out->str_off	 = a.soff;
	out->arena		 = a.arena;
	out->structural	 = a.positions;
	out->n_nodes	 = d.n_nodes;
	out->arena_bytes = d.arena_bytes;
	out->n_struct	 = ns;
	out->fail_event	 = d.fail_event == kDomNone ? -1 : (int64_t)d.fail_event;
	out->ok			 = (int32_t)d.ok;
	return PARS_OK;
}

int pars_hip_kv_parse(pars_scanner* sc, const uint8_t* d_in, size_t n, int format, pars_stream stream,
					  pars_stage_times* st, pars_kv_table* out) {
	if (!sc || !out || (n && !d_in))
		return fail(PARS_ERR_ARGUMENT, "null argument");
	if (format != PARS_KV_INI && format != PARS_KV_TOML)
		return fail(PARS_ERR_ARGUMENT, "unknown key-value format %d (PARS_KV_INI or PARS_KV_TOML)", format);
	if (n > 0xFFFFFFFFull)
		return fail(PARS_ERR_ARGUMENT, "the key-value table keeps 32-bit offsets: at most 2^32-1 bytes (%llu here)",
					(unsigned long long)n);
	std::memset(out, 0, sizeof *out);
	if (st)
		*st = pars_stage_times {0, 0, 0, 0};
	if (n == 0)
		return PARS_OK;
	hipStream_t		   s = (hipStream_t)stream;
	pars_device_result res;
	const int		   wb = sc->want_bitmap;
	sc->want_bitmap		  = 1;	// the passes read the structural bitmap
	int r				  = pars_hip_scan_device(sc, d_in, n, stream, 0, 0, st, &res);
	sc->want_bitmap		  = wb;
	if (r)
		return r;
	const bool	   toml	 = format == PARS_KV_TOML;
	const uint64_t tiles = (n + kKvTile - 1) / kKvTile, lanes = tiles * kKvLanes;
	if ((r = sc->kv_tile.ensure(tiles * 8 * 4)) || (r = sc->kv_lane.ensure(lanes * 3 * 4)) ||
		(r = sc->kv_mask.ensure(lanes * 4 * 8)) || (r = sc->kv_sum.ensure(sizeof(KvSummary))))
		return r;
	if (!sc->h_kv && (r = ensure_pinned(reinterpret_cast<void**>(&sc->h_kv), sc->h_kv_cap, sizeof(KvSummary))))
		return r;
	KvArgs a {};
	a.data			 = d_in;
	a.n				 = n;
	a.bitmap		 = sc->bitmap.as<uint64_t>();
	a.n_tiles		 = (uint32_t)tiles;
	a.toml			 = toml ? 1 : 0;
	uint32_t* tile	 = sc->kv_tile.as<uint32_t>();
	a.t_tmap		 = tile;
	a.t_tin			 = tile + tiles;
	a.t_hmap		 = tile + 2 * tiles;
	a.t_hin			 = tile + 3 * tiles;
	a.t_cnt			 = tile + 4 * tiles;
	a.t_off			 = tile + 5 * tiles;
	a.t_key			 = reinterpret_cast<int32_t*>(tile + 6 * tiles);
	a.t_kin			 = reinterpret_cast<int32_t*>(tile + 7 * tiles);
	uint32_t* lane32 = sc->kv_lane.as<uint32_t>();
	a.l_tmap		 = lane32;
	a.l_hmap		 = lane32 + lanes;
	a.l_off			 = lane32 + 2 * lanes;
	uint64_t* mask	 = sc->kv_mask.as<uint64_t>();
	a.l_start		 = mask;
	a.l_first		 = mask + lanes;
	a.l_end			 = mask + 2 * lanes;
	a.l_endq		 = mask + 3 * lanes;
	a.sum			 = sc->kv_sum.as<KvSummary>();
	hipError_t e;
	float	   ms = 0, ms2 = 0;
	if (st && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	if ((e = launch_kv_count(a, s)) != hipSuccess)
		return hip_fail(e, "key-value count launch");
	if (st && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	if ((e = hipMemcpyAsync(sc->h_kv, a.sum, sizeof(KvSummary), hipMemcpyDeviceToHost, s)) != hipSuccess)
		return hip_fail(e, "key-value summary readback");
	if ((e = hipStreamSynchronize(s)) != hipSuccess)
		return hip_fail(e, "hipStreamSynchronize");
	if (st && hipEventElapsedTime(&ms, sc->ev[0], sc->ev[1]) != hipSuccess)
		ms = 0;
	const uint64_t ntok = sc->h_kv->n_tokens;
	if (ntok > n)
		return fail(PARS_ERR_DEVICE, "more tokens than bytes (internal error)");
	if (toml && ntok > (uint64_t)INT32_MAX)	 // keys are int32 token indices
		return fail(PARS_ERR_ARGUMENT, "a TOML key-value table holds at most 2^31-1 tokens (%llu here)",
					(unsigned long long)ntok);
	const uint64_t cap = std::max<uint64_t>(ntok, 1);
	if ((r = sc->kv_role.ensure(cap)) || (r = sc->kv_beg.ensure(cap * 4)) || (r = sc->kv_end.ensure(cap * 4)) ||
		(toml && (r = sc->kv_key.ensure(cap * 4))))
		return r;
	a.role = sc->kv_role.as<uint8_t>();
	a.beg  = sc->kv_beg.as<uint32_t>();
	a.end  = sc->kv_end.as<uint32_t>();
	a.key  = toml ? sc->kv_key.as<int32_t>() : nullptr;
	if (st && (e = hipEventRecord(sc->ev[0], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	if ((e = launch_kv_emit(a, s)) != hipSuccess)
		return hip_fail(e, "key-value table launch");
	if (st && (e = hipEventRecord(sc->ev[1], s)) != hipSuccess)
		return hip_fail(e, "hipEventRecord");
	if ((e = hipMemcpyAsync(sc->h_kv, a.sum, sizeof(KvSummary), hipMemcpyDeviceToHost, s)) != hipSuccess)
		return hip_fail(e, "key-value summary readback");
	if ((e = hipStreamSynchronize(s)) != hipSuccess)
		return hip_fail(e, "hipStreamSynchronize");
	if (st) {
		if (hipEventElapsedTime(&ms2, sc->ev[0], sc->ev[1]) != hipSuccess)
			ms2 = 0;
		st->p2 = ms + ms2;
	}
	const KvSummary& k = *sc->h_kv;
	if (k.bad)
		return fail(PARS_ERR_DEVICE, "the key-value table passes disagree on the token count (internal error)");
	uint64_t		 tot = 0;
	for (int i = 0; i < 5; ++i) tot += k.roles[i];
	if (tot != ntok)
		return fail(PARS_ERR_DEVICE, "token roles do not add up to the token count (internal error)");
	out->role		= a.role;
	out->begin		= a.beg;
	out->end		= a.end;
	out->key		= a.key;
	out->structural = res.structural;
	out->n_tokens	= ntok;
	out->n_struct	= res.n_struct;
	for (int i = 0; i < 5; ++i) out->n_role[i] = k.roles[i];
	return PARS_OK;
}

}  // extern "C"
