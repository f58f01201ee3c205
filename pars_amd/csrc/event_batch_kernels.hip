// event_batch_kernels.hip — the batch launchers of the event evaluator (see
// event_kernels.hip), compiled apart so the two sets build in parallel.
#include "evaluators.hpp"

namespace pars_hip {

namespace {

template<int W>
hipError_t launch_batch_event_rows(const ScanArgs& a, int nd, bool wide, hipStream_t s) {
	if (wide) {
		switch (nd) {
		case 2: return launch_eval<EventEval<2, true>, W, true>(a, s);
		case 3: return launch_eval<EventEval<3, true>, W, true>(a, s);
		case 4: return launch_eval<EventEval<4, true>, W, true>(a, s);
		case 5: return launch_eval<EventEval<5, true>, W, true>(a, s);
		}
	} else {
		switch (nd) {
		case 2: return launch_eval<EventEval<2, false>, W, true>(a, s);
		case 3: return launch_eval<EventEval<3, false>, W, true>(a, s);
		case 4: return launch_eval<EventEval<4, false>, W, true>(a, s);
		case 5: return launch_eval<EventEval<5, false>, W, true>(a, s);
		}
	}
	return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_batch_event(const ScanArgs& a, int nd, bool wide, int rows, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	return event_tile_rows(rows) == 1 ? launch_batch_event_rows<1>(a, nd, wide, s)
									  : launch_batch_event_rows<2>(a, nd, wide, s);
}

}  // namespace pars_hip
