// server.hip — the serving runtime: parse requests arrive one by one, are
// staged in pinned host memory, copied asynchronously into a device ring, and
// scanned by the batched kernel with continuous batching.
//
// Per plan: a pinned staging ring, a region of the device ring, a copy stream,
// a compute stream and `slots` scanners.  A submit reserves its staging space
// under the server lock and copies its bytes outside it, so submitters copy in
// parallel; the request becomes batchable when its copy is done.  One worker
// thread per plan (batch enqueueing is host-bound: a batch is ~20 launches)
// loops:
//   - retires batches whose completion event fired (their per-document counts
//     and flags were copied to pinned memory at the end of the batch), marks
//     their requests done and frees their ring and staging space;
//   - frees a slot once every request of its batch is released;
//   - for every plan with pending requests and a free slot, takes the longest
//     run of staged pending requests that is contiguous in staging and fits the
//     batch limits, the device ring and the HBM budget, and enqueues H2D copy ->
//     batched scan -> D2H of the per-document results -> completion event.
// Batches of one plan run in order on its compute stream (so ring space frees
// in allocation order); plans run on their own streams, and a batch's copy
// overlaps the previous batch's kernels.  No counterpart in the reference.
#include "../../include/pars_hip.h"
#include "internal.hpp"
#include "memory.hpp"
#include "ring.hpp"

#include <hip/hip_runtime.h>

#include <emmintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <memory>
#include <vector>

using namespace pars_hip;

namespace {

uint64_t now_ns() {
	return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
			   std::chrono::steady_clock::now().time_since_epoch())
		.count();
}

constexpr uint64_t kAlign = 64;

uint64_t align_up(uint64_t x) {
	return (x + kAlign - 1) / kAlign * kAlign;
}

// A request's bytes into staging, padded with spaces to the next 64 bytes.
// Staging is read only by the copy engine, so the stores bypass the caches
// (non-temporal: no read-for-ownership of the destination lines); the fence
// orders them before the request is marked staged.  dst is 64-byte aligned.
void stage_copy(uint8_t* dst, const uint8_t* src, size_t n) {
	const size_t body = n & ~(size_t)63;
	for (size_t o = 0; o < body; o += 64) {
		const __m128i a = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + o));
		const __m128i b = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + o + 16));
		const __m128i c = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + o + 32));
		const __m128i d = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + o + 48));
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + o), a);
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + o + 16), b);
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + o + 32), c);
		_mm_stream_si128(reinterpret_cast<__m128i*>(dst + o + 48), d);
	}
	if (body < n) {
		alignas(64) uint8_t tail[64];
		std::memset(tail, ' ', sizeof tail);
		std::memcpy(tail, src + body, n - body);
		for (int q = 0; q < 4; ++q)
			_mm_stream_si128(reinterpret_cast<__m128i*>(dst + body + 16 * q),
							 _mm_load_si128(reinterpret_cast<const __m128i*>(tail + 16 * q)));
	}
	_mm_sfence();
}

// The server lock guards short critical sections (ring reservations, queue
// pushes, a request record) taken by every submitter, the drain and the plan
// workers at up to ~10^6 per second: a spin lock that pauses, then yields,
// hands it over in tens of nanoseconds where a futex mutex would put waiters
// to sleep.
struct SpinLock {
	std::atomic<bool> held {false};
	void			  lock() {
		 for (int spins = 0; held.exchange(true, std::memory_order_acquire);)
			 while (held.load(std::memory_order_relaxed))
				 if (++spins < 256)
					 _mm_pause();
				 else
					 std::this_thread::yield();
	}
	bool try_lock() { return !held.load(std::memory_order_relaxed) && !held.exchange(true, std::memory_order_acquire); }
	void unlock() { held.store(false, std::memory_order_release); }
};

enum ReqState : int { kPending = 0, kInBatch = 1, kDone = 2 };

struct Request {
	uint64_t id		   = 0;	 // (generation << 32) | index in the request pool; 0 while free
	uint32_t plan	   = 0;
	uint64_t n		   = 0, stage_off = 0, reserve = 0;
	uint64_t t_submit  = 0, t_launch = 0, t_done = 0;
	uint64_t tag	   = 0;	 // the replay driver's request index
	int		 state	   = kPending;
	int		 slot	   = -1;
	uint32_t doc	   = 0;
	bool	 polled	   = false;
	std::atomic<bool> staged {false};  // its bytes are in staging (set after the copy, outside the lock)
	pars_request_result res {};
};

enum SlotState : int { kFree = 0, kRunning = 1, kHeld = 2 };

struct Slot {
	pars_scanner*		  sc = nullptr;
	hipEvent_t			  h2d = nullptr, done = nullptr;
	int					  state = kFree;
	int					  fail_status = PARS_OK;  // the batch could not be enqueued: retire it as failed
	uint64_t			  ring_off	  = 0;		  // its inputs in the device ring, held until the slot frees
	bool				  ring_live	  = false;
	std::vector<Request*> reqs;	 // stable: the request table never moves its elements
	uint32_t			  outstanding = 0;
	uint64_t			  launch_ns	  = 0;
	BatchPtrs			  ptrs {};
	// pinned per-document results of the batch, packed (BatchPtrs::pack)
	uint32_t* h_pack = nullptr;
	std::vector<uint64_t> offs;
	std::vector<uint32_t> lens;
};

struct PlanState {
	uint8_t*		  stage = nullptr;	// pinned
	Ring			  stage_ring;
	uint8_t*		  ring = nullptr;  // device
	Ring			  dev_ring;
	hipStream_t		  copy = nullptr, compute = nullptr;
	std::deque<Request*> pending;  // submitted requests in staging order (stable pointers into the table)
	std::condition_variable_any cv;  // wakes this plan's idle worker
	std::atomic<bool>		idle {false};
	std::vector<Slot>	 slots;
	std::deque<int>		 running;  // slots in launch order
};

}  // namespace

struct pars_server {
	pars_server_config			  cfg {};
	std::vector<PlanState>		  plans;
	SpinLock					  mu;
	std::condition_variable_any	  cv_done;
	std::atomic<uint64_t>		  work_seq {0};
	// request pool: chunks of kPoolChunk records that never move, a free list of
	// indices; an id carries its slot's generation, so a stale id finds nothing
	static constexpr uint32_t			  kPoolChunk = 4096;
	std::vector<std::unique_ptr<Request[]>> pool;
	std::vector<uint32_t>				  pool_gen;
	std::vector<uint32_t>				  pool_free;
	Request*							  req_alloc();			 // mu held
	Request*							  req_find(uint64_t id);	 // mu held; nullptr if not live
	void								  req_free(Request& q);	 // mu held
	std::deque<uint64_t>		  completed;
	uint64_t					  reserved = 0;	 // budget held by admitted, unreleased requests
	uint64_t					  fixed	   = 0;	 // device bytes held for rings and arenas
	uint64_t					  ring_charged = 0;	// ring bytes charged to the shared budget
	pars_server_stats			  st {};
	std::vector<std::thread>	  workers;	// one per plan
	std::atomic<uint64_t>		  launch_ns {0}, retire_ns {0};
	std::atomic<bool>			  stop {false};
	std::atomic<uint64_t>		  submit_ns {0}, submit_lock_ns {0};
	std::string					  error;
	int							  dev = 0;

	~pars_server();
	void run(uint32_t p);
	bool retire(uint32_t p);
	bool launch(uint32_t p);
	int	 submit(uint32_t plan, const uint8_t* data, size_t n, uint64_t tag, uint64_t* id);
	void release_locked(Request& q);  // mu held
	void wake_idle() {				  // a release frees slots, ring space and budget: any plan may launch
		for (PlanState& p : plans)
			if (p.idle.load())
				p.cv.notify_one();
	}
};

namespace {

int sfail(int code, const char* msg) {
	return set_error(code, msg);
}

int shipfail(hipError_t e, const char* what) {
	return set_error(PARS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

pars_server::~pars_server() {
	{
		std::lock_guard<SpinLock> lk(mu);
		stop = true;
	}
	wake_idle();
	for (std::thread& w : workers)
		if (w.joinable())
			w.join();
	for (PlanState& p : plans) {
		if (p.compute)
			(void)hipStreamSynchronize(p.compute);
		for (Slot& s : p.slots) {
			if (s.sc)
				pars_hip_scanner_destroy(s.sc);
			if (s.h2d)
				(void)hipEventDestroy(s.h2d);
			if (s.done)
				(void)hipEventDestroy(s.done);
			if (s.h_pack)
				(void)hipHostFree(s.h_pack);
		}
		if (p.stage)
			(void)hipHostFree(p.stage);
		if (p.ring)
			(void)hipFree(p.ring);
	}
	if (pars_budget* b = cfg.budget) {	// rings, and requests still held by their callers
		b->uncharge(ring_charged + reserved);
		b->users.fetch_sub(1);
	}
	for (PlanState& p : plans) {
		if (p.copy)
			(void)hipStreamDestroy(p.copy);
		if (p.compute)
			(void)hipStreamDestroy(p.compute);
	}
}

// Batches whose completion event fired: fill their requests' results.
bool pars_server::retire(uint32_t p) {
	bool any = false;
	{
		PlanState& ps = plans[p];
		while (!ps.running.empty()) {
			Slot& s = ps.slots[ps.running.front()];
			// a batch that failed to enqueue never recorded its event; it is
			// retired in launch order like the others, so the ring regions freed
			// below are always the oldest batch's own
			const hipError_t e = s.fail_status ? hipSuccess : hipEventQuery(s.done);
			if (e == hipErrorNotReady)
				break;
			const uint64_t	t	   = now_ns();
			const size_t	nd	   = s.reqs.size();
			const uint32_t* pk	   = s.h_pack;	// packed: starts, bracket starts, depths, flags, overflow
			const bool		failed = s.fail_status || e != hipSuccess || batch_failed_word(pk[4 * nd + 2]);
			const int		status = s.fail_status ? s.fail_status : PARS_ERR_DEVICE;
			// results are filled outside the lock: while a request is in a batch
			// only this thread touches it; publishing it under the lock below
			// orders these writes before any poll that returns it
			for (size_t d = 0; d < nd; ++d) {
				Request&			 q = *s.reqs[d];
				pars_request_result& r = q.res;
				r					   = pars_request_result {};
				r.id				   = q.id;
				r.status			   = failed ? status : PARS_OK;
				r.plan				   = p;
				r.n_bytes			   = q.n;
				if (!failed) {
					const uint32_t st0 = pk[d], br0 = pk[nd + 1 + d];
					r.n_struct		   = pk[d + 1] - st0;
					r.n_brackets	   = pk[nd + 2 + d] - br0;
					r.structural	   = s.ptrs.structural + st0;
					r.bracket_pos	   = s.ptrs.bracket_pos + br0;
					r.bracket_char	   = s.ptrs.bracket_char + br0;
					r.bracket_pair	   = s.ptrs.bracket_pair ? s.ptrs.bracket_pair + br0 : nullptr;
					r.input			   = ps.ring + s.ring_off + s.offs[d];
					r.flags			   = pk[3 * nd + 2 + d];
					r.max_depth		   = (int32_t)pk[2 * nd + 2 + d];
				}
				r.t_submit_ns		   = q.t_submit;
				r.t_launch_ns		   = q.t_launch;
				r.t_done_ns			   = t;
				r.batch_docs		   = (uint32_t)s.reqs.size();
				q.t_done			   = t;
			}
			std::lock_guard<SpinLock> lk(mu);
			for (size_t d = 0; d < nd; ++d) {
				Request& q = *s.reqs[d];
				q.state	   = kDone;
				completed.push_back(q.id);
				st.completed++;
				if (!failed)
					st.bytes_completed += q.n;
			}
			// staging is free once the inputs are on the device; the device ring
			// region stays until the slot frees (the request's input is readable
			// beside its results until release)
			for (size_t d = 0; d < s.reqs.size(); ++d) ps.stage_ring.release_oldest();
			s.state		  = kHeld;
			s.fail_status = PARS_OK;
			ps.running.pop_front();
			any = true;
		}
	}
	if (any)
		cv_done.notify_all();
	return any;
}

// Form and enqueue one batch for plan p if it has pending work and a free slot.
bool pars_server::launch(uint32_t p) {
	PlanState& ps	= plans[p];
	int		   slot = -1;
	uint64_t   bytes = 0, first_off = 0, ring_off = 0;
	const uint64_t t = now_ns();
	{
		std::lock_guard<SpinLock> lk(mu);
		for (size_t i = 0; i < ps.slots.size(); ++i)
			if (ps.slots[i].state == kFree || (ps.slots[i].state == kHeld && ps.slots[i].outstanding == 0)) {
				slot = (int)i;
				break;
			}
		if (slot >= 0 && ps.slots[slot].ring_live) {
			ps.dev_ring.release(ps.slots[slot].ring_off);
			ps.slots[slot].ring_live = false;
		}
		if (slot < 0 || ps.pending.empty())
			return false;
		Slot&		   s	 = ps.slots[slot];
		const uint64_t avail = cfg.hbm_budget > fixed ? cfg.hbm_budget - fixed : 0;
		uint64_t	   next = 0, held = 0;
		s.reqs.clear();
		for (Request* qp : ps.pending) {
			Request& q = *qp;
			if (!q.staged.load(std::memory_order_acquire))	// its copy is still running
				break;
			if (s.reqs.empty())
				first_off = next = q.stage_off;
			else if (q.stage_off != next)  // staging wrapped: the batch ends here
				break;
			const uint64_t span = align_up(q.n);
			if (s.reqs.size() >= cfg.max_batch_docs || bytes + span > cfg.max_batch_bytes)
				break;
			if (reserved + held + q.reserve > avail) {
				st.budget_waits++;
				break;
			}
			held += q.reserve;
			s.reqs.push_back(&q);
			bytes += span;
			next = q.stage_off + span;
		}
		if (s.reqs.empty())
			return false;
		if (cfg.budget && !cfg.budget->charge(held)) {	// the shared budget: wait like for hbm_budget
			st.budget_waits++;
			s.reqs.clear();
			return false;
		}
		if (!ps.dev_ring.alloc(bytes, ring_off)) {	// ring full: retry once a slot frees
			if (cfg.budget)
				cfg.budget->uncharge(held);
			s.reqs.clear();
			return false;
		}
		reserved += held;
		s.ring_off	= ring_off;
		s.ring_live = true;
		s.offs.resize(s.reqs.size());
		s.lens.resize(s.reqs.size());
		for (size_t d = 0; d < s.reqs.size(); ++d) {
			Request& q = *s.reqs[d];
			s.offs[d]  = q.stage_off - first_off;
			s.lens[d]  = (uint32_t)q.n;
			q.state	   = kInBatch;
			q.slot	   = slot;
			q.doc	   = (uint32_t)d;
			q.t_launch = t;
			ps.pending.pop_front();
		}
		s.outstanding = (uint32_t)s.reqs.size();
		st.batches++;
	}
	Slot& s = ps.slots[slot];
	uint8_t*   d_in = ps.ring + ring_off;
	hipError_t e	= hipMemcpyAsync(d_in, ps.stage + first_off, bytes, hipMemcpyHostToDevice, ps.copy);
	if (e == hipSuccess)
		e = hipEventRecord(s.h2d, ps.copy);
	if (e == hipSuccess)
		e = hipStreamWaitEvent(ps.compute, s.h2d, 0);
	int r = e == hipSuccess ? PARS_OK : PARS_ERR_DEVICE;
	if (r == PARS_OK)
		r = batch_enqueue(s.sc, d_in, bytes, s.offs.data(), s.lens.data(), s.reqs.size(), ps.compute,
						  cfg.run_pairs != 0, &s.ptrs);
	if (r == PARS_OK) {	 // one copy brings every per-request result back
		e = hipMemcpyAsync(s.h_pack, s.ptrs.pack, s.ptrs.pack_bytes, hipMemcpyDeviceToHost, ps.compute);
		if (e != hipSuccess)
			r = PARS_ERR_DEVICE;
	}
	if (r == PARS_OK && (e = hipEventRecord(s.done, ps.compute)) != hipSuccess)
		r = PARS_ERR_DEVICE;
	if (r != PARS_OK) {
		// The device is in trouble.  Whatever was enqueued for this batch is
		// drained, and the batch joins the running queue as failed: retire()
		// fails its requests and frees its ring and staging regions in launch
		// order, after the batches still running ahead of it (their regions are
		// older and stay untouched).
		(void)hipStreamSynchronize(ps.copy);
		(void)hipStreamSynchronize(ps.compute);
		std::lock_guard<SpinLock> lk(mu);
		error		  = std::string("batch launch failed: ") + last_error();
		s.fail_status = r;
	}
	s.state		= kRunning;
	s.launch_ns = t;
	ps.running.push_back(slot);
	return true;
}

// Plan p's worker: retire its finished batches, launch what can go, then poll
// the running batches' events (yielding for a while after the last progress,
// then in short sleeps), or sleep until a submit or release wakes it.
void pars_server::run(uint32_t p) {
	(void)hipSetDevice(dev);
	PlanState& ps	  = plans[p];
	uint64_t   t_prog = now_ns();
	while (!stop) {
		const uint64_t seen		= work_seq.load();
		const uint64_t t0		= now_ns();
		const bool	   retired	= retire(p);
		const uint64_t t1		= now_ns();
		bool		   launched = false;
		while (launch(p)) launched = true;
		if (retired)
			retire_ns.fetch_add(t1 - t0, std::memory_order_relaxed);
		if (launched)
			launch_ns.fetch_add(now_ns() - t1, std::memory_order_relaxed);
		if (retired || launched) {
			t_prog = now_ns();
			continue;
		}
		if (!ps.running.empty()) {	// batches on the device: poll their events
			if (now_ns() - t_prog < 200000)
				std::this_thread::yield();
			else
				std::this_thread::sleep_for(std::chrono::microseconds(20));
			continue;
		}
		// idle: sleep until a submit or a release changes what can launch (the
		// submitter bumps work_seq before it looks at the idle flag)
		std::unique_lock<SpinLock> lk(mu);
		ps.idle.store(true);
		ps.cv.wait_for(lk, std::chrono::microseconds(500), [this, seen] { return stop || work_seq.load() != seen; });
		ps.idle.store(false);
	}
}

// The staging space and the request record are taken under the lock; the copy
// runs outside it, so concurrent submitters copy in parallel.  The batch
// former stops at the first pending request whose copy has not finished.
int pars_server::submit(uint32_t plan, const uint8_t* data, size_t n, uint64_t tag, uint64_t* id) {
	if (n == 0 || n > cfg.max_batch_bytes || n > 0xFFFFFFFFull)
		return sfail(PARS_ERR_ARGUMENT, "request size must be 1 .. max_batch_bytes");
	if (cfg.budget && n * 18 > cfg.budget->limit)
		return sfail(PARS_ERR_BUDGET, "the request's worst-case footprint exceeds the shared budget");
	PlanState&	   ps = plans[plan];
	Request*	   q  = nullptr;
	uint64_t	   off;
	const uint64_t t0 = now_ns();
	{
		std::lock_guard<SpinLock> lk(mu);
		submit_lock_ns.fetch_add(now_ns() - t0, std::memory_order_relaxed);
		if (!ps.stage_ring.alloc(align_up(n), off)) {
			st.busy_rejects++;
			return sfail(PARS_ERR_BUSY, "staging ring full");
		}
		q = req_alloc();
		if (!q)
			return sfail(PARS_ERR_MEMORY, "host allocation failed");
		const uint64_t rid = q->id;
		q->plan			   = plan;
		q->n			   = n;
		q->stage_off	   = off;
		q->reserve		   = n * 18;  // the request's input plus its worst-case outputs
		q->tag			   = tag;
		q->t_submit		   = now_ns();
		ps.pending.push_back(q);
		st.submitted++;
		*id = rid;
	}
	stage_copy(ps.stage + off, data, n);
	q->staged.store(true, std::memory_order_release);
	work_seq++;
	if (ps.idle.load())
		ps.cv.notify_one();
	submit_ns.fetch_add(now_ns() - t0, std::memory_order_relaxed);
	return PARS_OK;
}

Request* pars_server::req_alloc() {
	if (pool_free.empty()) {
		const uint32_t base = (uint32_t)pool_gen.size();
		if (base > 0xFFFFFFFFu - kPoolChunk)
			return nullptr;
		std::unique_ptr<Request[]> c(new (std::nothrow) Request[kPoolChunk]);
		if (!c)
			return nullptr;
		pool.push_back(std::move(c));
		pool_gen.resize(base + kPoolChunk, 0u);
		for (uint32_t i = kPoolChunk; i-- > 0;) pool_free.push_back(base + i);
	}
	const uint32_t idx = pool_free.back();
	pool_free.pop_back();
	Request& q = pool[idx / kPoolChunk][idx % kPoolChunk];
	const uint32_t g = ++pool_gen[idx] ? pool_gen[idx] : ++pool_gen[idx];	// never 0: id 0 means free
	q.id			 = ((uint64_t)g << 32) | idx;
	q.t_launch = q.t_done = 0;
	q.state				 = kPending;
	q.slot				 = -1;
	q.doc				 = 0;
	q.polled			 = false;
	q.staged.store(false, std::memory_order_relaxed);
	q.res = pars_request_result {};
	return &q;
}

Request* pars_server::req_find(uint64_t id) {
	const uint32_t idx = (uint32_t)id;
	if (id == 0 || idx >= pool_gen.size())
		return nullptr;
	Request& q = pool[idx / kPoolChunk][idx % kPoolChunk];
	return q.id == id ? &q : nullptr;
}

void pars_server::req_free(Request& q) {
	const uint32_t idx = (uint32_t)q.id;
	q.id			   = 0;
	pool_free.push_back(idx);
}

void pars_server::release_locked(Request& q) {
	if (q.slot >= 0) {
		Slot& s = plans[q.plan].slots[q.slot];
		if (s.outstanding)
			s.outstanding--;
		if (s.outstanding == 0 && s.state == kHeld) {
			s.state = kFree;
			if (s.ring_live) {	// its inputs are no longer needed
				plans[q.plan].dev_ring.release(s.ring_off);
				s.ring_live = false;
			}
		}
	}
	const uint64_t back = std::min(reserved, q.reserve);
	reserved -= back;
	if (cfg.budget)
		cfg.budget->uncharge(back);
	st.released++;
	req_free(q);
	work_seq++;
}

extern "C" {

int pars_hip_server_create(const pars_server_config* cfg_in, pars_server** out) {
	if (!cfg_in || !out || !cfg_in->plans || !cfg_in->n_plans)
		return sfail(PARS_ERR_ARGUMENT, "server needs at least one plan");
	*out			   = nullptr;
	pars_server_config cfg = *cfg_in;
	if (!cfg.slots_per_plan)
		cfg.slots_per_plan = 2;
	if (!cfg.ring_bytes)
		cfg.ring_bytes = 4ull << 30;
	if (!cfg.staging_bytes)
		cfg.staging_bytes = 1ull << 30;
	if (!cfg.hbm_budget)
		cfg.hbm_budget = 64ull << 30;
	if (!cfg.max_batch_bytes)
		cfg.max_batch_bytes = 64ull << 20;
	if (!cfg.max_batch_docs)
		cfg.max_batch_docs = 16384;
	if (!cfg.submit_threads)
		cfg.submit_threads = 4;
	cfg.submit_threads = std::min<uint32_t>(cfg.submit_threads, 64);
	cfg.max_batch_bytes = align_up(cfg.max_batch_bytes);
	const uint64_t ring_each  = cfg.ring_bytes / cfg.n_plans / kAlign * kAlign;
	const uint64_t stage_each = cfg.staging_bytes / cfg.n_plans / kAlign * kAlign;
	if (ring_each < cfg.max_batch_bytes || stage_each < cfg.max_batch_bytes)
		return sfail(PARS_ERR_ARGUMENT, "ring and staging must hold one largest batch per plan");
	pars_server* sv = new (std::nothrow) pars_server();
	if (!sv)
		return sfail(PARS_ERR_MEMORY, "host allocation failed");
	sv->cfg	 = cfg;
	if (cfg.budget)
		cfg.budget->users.fetch_add(1);
	(void)hipGetDevice(&sv->dev);
	sv->plans = std::vector<PlanState>(cfg.n_plans);  // constructed in place (PlanState does not move)
	int r = PARS_OK;
	for (uint32_t p = 0; p < cfg.n_plans && r == PARS_OK; ++p) {
		PlanState& ps = sv->plans[p];
		hipError_t e  = hipHostMalloc((void**)&ps.stage, stage_each, hipHostMallocDefault);
		if (e == hipSuccess)
			e = hipMalloc((void**)&ps.ring, ring_each);
		if (e == hipSuccess)
			e = hipStreamCreateWithFlags(&ps.copy, hipStreamNonBlocking);
		if (e == hipSuccess)
			e = hipStreamCreateWithFlags(&ps.compute, hipStreamNonBlocking);
		if (e != hipSuccess) {
			r = shipfail(e, "server rings");
			break;
		}
		// map every staging page now: the kernel would otherwise fault them in on
		// the first submits' copies (a minor fault per 4 KiB page)
		for (uint64_t o = 0; o < stage_each; o += 4096) ps.stage[o] = 0;
		ps.stage_ring.cap = stage_each;
		ps.dev_ring.cap	  = ring_each;
		sv->fixed += ring_each;
		if (cfg.budget) {
			if (!cfg.budget->charge(ring_each)) {
				r = sfail(PARS_ERR_BUDGET, "the server's device ring exceeds the shared budget");
				break;
			}
			sv->ring_charged += ring_each;
		}
		ps.slots.resize(cfg.slots_per_plan);
		for (Slot& s : ps.slots) {
			if ((r = pars_hip_scanner_create_ex(&cfg.plans[p], nullptr, cfg.budget, &s.sc)))
				break;
			if ((r = pars_hip_set_option(s.sc, PARS_OPT_WANT_BITMAP, 0)))
				break;
			if ((r = batch_reserve(s.sc, cfg.max_batch_bytes, cfg.max_batch_docs)))
				break;
			const uint64_t nd = cfg.max_batch_docs;
			e				  = hipEventCreateWithFlags(&s.h2d, hipEventDisableTiming);
			if (e == hipSuccess)
				e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
			if (e == hipSuccess)
				e = hipHostMalloc((void**)&s.h_pack, batch_pack_bytes(nd), hipHostMallocDefault);
			if (e != hipSuccess) {
				r = shipfail(e, "server slot");
				break;
			}
			// worst-case outputs of a full batch: positions 4 B, bracket records
			// (position, index, key, byte, document, sort keys and ranks, pair) 45 B
			sv->fixed += cfg.max_batch_bytes * 49;
		}
	}
	if (r == PARS_OK && sv->fixed > cfg.hbm_budget)
		r = sfail(PARS_ERR_BUDGET, "rings and scanner arenas alone exceed the HBM budget");
	if (r != PARS_OK) {
		delete sv;
		return r;
	}
	sv->st.device_bytes = sv->fixed;
	for (uint32_t p = 0; p < cfg.n_plans; ++p) sv->workers.emplace_back([sv, p] { sv->run(p); });
	*out				= sv;
	return PARS_OK;
}

int pars_hip_server_destroy(pars_server* sv) {
	delete sv;
	return PARS_OK;
}

int pars_hip_server_submit(pars_server* sv, uint32_t plan, const uint8_t* data, size_t n, uint64_t* id) {
	if (!sv || !id || (n && !data) || plan >= sv->plans.size())
		return sfail(PARS_ERR_ARGUMENT, "bad submit arguments");
	return sv->submit(plan, data, n, 0, id);
}

int pars_hip_server_poll(pars_server* sv, pars_request_result* out, size_t cap, size_t* n_out, int64_t wait_us) {
	if (!sv || !n_out || (cap && !out))
		return sfail(PARS_ERR_ARGUMENT, "bad poll arguments");
	std::unique_lock<SpinLock> lk(sv->mu);
	if (sv->completed.empty() && wait_us > 0)
		sv->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [sv] { return !sv->completed.empty(); });
	size_t k = 0;
	while (k < cap && !sv->completed.empty()) {
		Request* qp = sv->req_find(sv->completed.front());
		if (!qp) {	// released without being polled
			sv->completed.pop_front();
			continue;
		}
		Request& q = *qp;
		sv->completed.pop_front();
		q.polled = true;
		out[k++] = q.res;
	}
	*n_out = k;
	return PARS_OK;
}

int pars_hip_server_release(pars_server* sv, uint64_t id) {
	if (!sv)
		return sfail(PARS_ERR_ARGUMENT, "null server");
	{
		std::lock_guard<SpinLock> lk(sv->mu);
		Request*					q = sv->req_find(id);
		if (!q || q->state != kDone)
			return sfail(PARS_ERR_ARGUMENT, "unknown or unfinished request");
		sv->release_locked(*q);
	}
	sv->wake_idle();
	return PARS_OK;
}

int pars_hip_server_stats(pars_server* sv, pars_server_stats* out) {
	if (!sv || !out)
		return sfail(PARS_ERR_ARGUMENT, "null argument");
	std::lock_guard<SpinLock> lk(sv->mu);
	*out				   = sv->st;
	out->inflight_reserved = sv->reserved;
	out->ring_wraps		   = 0;
	out->staging_wraps	   = 0;
	out->submit_ns		   = sv->submit_ns.load();
	out->worker_launch_ns  = sv->launch_ns.load();
	out->worker_retire_ns  = sv->retire_ns.load();
	out->submit_lock_ns	   = sv->submit_lock_ns.load();
	for (const PlanState& p : sv->plans) {
		out->ring_wraps += p.dev_ring.wraps;
		out->staging_wraps += p.stage_ring.wraps;
	}
	return PARS_OK;
}

int pars_hip_server_replay(pars_server* sv, size_t n, const uint32_t* plan, const uint8_t* const* data,
						   const uint64_t* sizes, const uint64_t* arrival_ns, uint64_t* latency_ns,
						   uint64_t* elapsed_ns) {
	if (!sv || !plan || !data || !sizes || !arrival_ns || !latency_ns)
		return sfail(PARS_ERR_ARGUMENT, "null argument");
	if (n == 0) {
		if (elapsed_ns)
			*elapsed_ns = 0;
		return PARS_OK;
	}
	for (size_t i = 0; i < n; ++i)
		if (plan[i] >= sv->plans.size() || (sizes[i] && !data[i]))
			return sfail(PARS_ERR_ARGUMENT, "bad replay request");
	const uint32_t		 nt = (uint32_t)std::min<size_t>(sv->cfg.submit_threads, n);
	std::atomic<int>	 err {PARS_OK};
	std::atomic<size_t>	 submitted {0};
	std::string			 err_msg;
	std::mutex			 err_mu;
	auto				 fail_with = [&](int code) {
		int expect = PARS_OK;
		if (err.compare_exchange_strong(expect, code)) {
			std::lock_guard<std::mutex> lk(err_mu);
			err_msg = last_error();
		}
	};
	const uint64_t t0 = now_ns();
	// submitters: thread k takes requests k, k + nt, ... in arrival order
	std::vector<std::thread> th;
	th.reserve(nt);
	for (uint32_t k = 0; k < nt; ++k)
		th.emplace_back([&, k] {
			for (size_t i = k; i < n && err.load(std::memory_order_relaxed) == PARS_OK; i += nt) {
				// sleep through long waits (spinning submitters would take CPU from
				// the worker), spin through the last stretch
				for (uint64_t now = now_ns(); now < t0 + arrival_ns[i]; now = now_ns()) {
					if (err.load(std::memory_order_relaxed) != PARS_OK)
						return;
					const uint64_t wait = t0 + arrival_ns[i] - now;
					if (wait > 200000)
						std::this_thread::sleep_for(std::chrono::nanoseconds(wait - 150000));
					else
						std::this_thread::yield();
				}
				uint64_t id;
				int		 r;
				while ((r = sv->submit(plan[i], data[i], sizes[i], i, &id)) == PARS_ERR_BUSY) {
					if (err.load(std::memory_order_relaxed) != PARS_OK)
						return;
					std::this_thread::sleep_for(std::chrono::microseconds(10));
				}
				if (r) {
					fail_with(r);
					return;
				}
				submitted.fetch_add(1, std::memory_order_relaxed);
			}
		});
	// this thread: completions, released in groups under one lock
	size_t	 done = 0;
	uint64_t last = 0, progress = now_ns();
	while (done < n && err.load() == PARS_OK) {
		size_t got = 0;
		{
			std::unique_lock<SpinLock> lk(sv->mu);
			if (sv->completed.empty())
				sv->cv_done.wait_for(lk, std::chrono::microseconds(200), [sv] { return !sv->completed.empty(); });
			while (!sv->completed.empty()) {
				Request* qp = sv->req_find(sv->completed.front());
				sv->completed.pop_front();
				if (!qp)
					continue;
				Request& q = *qp;
				if (q.res.status != PARS_OK) {
					err.store(q.res.status);
					std::lock_guard<std::mutex> ek(err_mu);
					err_msg = "a request failed";
					break;
				}
				const size_t i = (size_t)q.tag;
				latency_ns[i]  = q.res.t_done_ns - (t0 + arrival_ns[i]);
				last		   = std::max(last, q.res.t_done_ns);
				sv->release_locked(q);
				++got;
			}
		}
		if (got) {
			done += got;
			progress = now_ns();
			sv->wake_idle();
		} else if (now_ns() - progress > 30ull * 1000000000ull) {
			int expect = PARS_OK;
			if (err.compare_exchange_strong(expect, PARS_ERR_DEVICE)) {
				std::lock_guard<std::mutex> ek(err_mu);
				err_msg = "server made no progress for 30 s";
			}
		}
	}
	for (std::thread& t : th) t.join();
	if (err.load() != PARS_OK)
		return sfail(err.load(), err_msg.c_str());
	if (elapsed_ns)
		*elapsed_ns = last - (t0 + arrival_ns[0]);
	return PARS_OK;
}

}  // extern "C"
