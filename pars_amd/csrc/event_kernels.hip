// event_kernels.hip — launchers of the tile kernel over the event evaluator
// (EventEval, evaluators.hpp): plans of two to five descriptors whose
// background runs act as one byte, single documents and batches.
#include "evaluators.hpp"

namespace pars_hip {

namespace {

template<bool BATCH>
hipError_t launch_event(const ScanArgs& a, int nd, bool wide, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	if (wide) {
		switch (nd) {
		case 2: return launch_eval<EventEval<2, true>, 2, BATCH>(a, s);
		case 3: return launch_eval<EventEval<3, true>, 2, BATCH>(a, s);
		case 4: return launch_eval<EventEval<4, true>, 2, BATCH>(a, s);
		case 5: return launch_eval<EventEval<5, true>, 2, BATCH>(a, s);
		}
	} else {
		switch (nd) {
		case 2: return launch_eval<EventEval<2, false>, 2, BATCH>(a, s);
		case 3: return launch_eval<EventEval<3, false>, 2, BATCH>(a, s);
		case 4: return launch_eval<EventEval<4, false>, 2, BATCH>(a, s);
		case 5: return launch_eval<EventEval<5, false>, 2, BATCH>(a, s);
		}
	}
	return hipErrorInvalidValue;
}

}  // namespace

bool event_path_applies(const DevPlan& p) {
	return p.ev_ok && p.n_descs >= 2 && p.n_descs <= 5;
}

hipError_t launch_scan_event(const ScanArgs& a, int nd, bool wide, hipStream_t s) {
	return launch_event<false>(a, nd, wide, s);
}

hipError_t launch_batch_event(const ScanArgs& a, int nd, bool wide, hipStream_t s) {
	return launch_event<true>(a, nd, wide, s);
}

}  // namespace pars_hip
