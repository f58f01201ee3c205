// scan_kernels.hip — single-document launchers of the tile kernel over each
// lane evaluator (evaluators.hpp), and the summary reset.
#include "evaluators.hpp"

namespace pars_hip {

namespace {

__global__ void k_summary_reset(DevSummary* s, SummaryInit in) {
	if (threadIdx.x == 0) {
		s->n_struct		= in.n_struct;
		s->n_brackets	= in.n_brackets;
		s->last_bal		= in.depth;
		s->final_states = 0;
		s->min_bal		= in.min_depth;
		s->max_bal		= in.max_depth;
		s->mismatched	= 0;
		s->sort_passes	= 0;
		s->kmin			= 0;
		s->n_chunks		= 0;
		s->overflow		= 0;
		s->small_done	= 0;
		s->doc_unbal	= 0;
		s->key_total	= 0;
		const uint32_t e = (s->epoch + 1u) & 0x3FFFFFFFu;
		s->epoch		 = e ? e : 1u;
	}
}

__global__ void k_carry_next(DevSummary* s, DevCarry* c) {
	if (threadIdx.x == 0) {
		c->states		 = s->final_states;
		c->out_base		 = s->n_struct;
		c->br_base		 = s->n_brackets;
		c->depth		 = s->last_bal;
		const uint32_t e = (s->epoch + 1u) & 0x3FFFFFFFu;
		s->epoch		 = e ? e : 1u;
	}
}

}  // namespace

hipError_t launch_carry_next(DevSummary* sum, DevCarry* carry, hipStream_t s) {
	hipLaunchKernelGGL(k_carry_next, dim3(1), dim3(64), 0, s, sum, carry);
	return hipGetLastError();
}

namespace {
thread_local int t_last_occupancy = 0;
}

void set_last_scan_occupancy(int per_cu) {
	t_last_occupancy = per_cu;
}

int last_scan_occupancy() {
	return t_last_occupancy;
}

hipError_t launch_summary_reset(DevSummary* sum, const SummaryInit& in, hipStream_t s) {
	hipLaunchKernelGGL(k_summary_reset, dim3(1), dim3(64), 0, s, sum, in);
	return hipGetLastError();
}

hipError_t launch_scan_escape(const ScanArgs& a, int rows, int waves, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	if (waves == 6 || waves == 8)
		return launch_scan_wide(a, false, rows, waves, s);
	return launch_rows<EscapeEval>(a, rows, 4, s);
}

hipError_t launch_scan_parity(const ScanArgs& a, int rows, int waves, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	if (waves == 6 || waves == 8)
		return launch_scan_wide(a, true, rows, waves, s);
	return launch_rows<ParityEval>(a, rows, 4, s);
}

hipError_t launch_scan_general(const ScanArgs& a, int nd, bool wide, int rows, hipStream_t s) {
	return launch_general<false>(a, nd, wide, rows, s);
}

}  // namespace pars_hip
