// event_kernels.hip — single-document launchers of the tile kernel over the
// event evaluator (EventEval, evaluators.hpp): plans of two to five
// descriptors whose background runs act as one byte; one or two 16 KiB rows
// per tile.
#include "evaluators.hpp"

namespace pars_hip {

namespace {

template<int W, bool BATCH>
hipError_t launch_event_rows(const ScanArgs& a, int nd, bool wide, hipStream_t s) {
	if (wide) {
		switch (nd) {
		case 2: return launch_eval<EventEval<2, true>, W, BATCH>(a, s);
		case 3: return launch_eval<EventEval<3, true>, W, BATCH>(a, s);
		case 4: return launch_eval<EventEval<4, true>, W, BATCH>(a, s);
		case 5: return launch_eval<EventEval<5, true>, W, BATCH>(a, s);
		}
	} else {
		switch (nd) {
		case 2: return launch_eval<EventEval<2, false>, W, BATCH>(a, s);
		case 3: return launch_eval<EventEval<3, false>, W, BATCH>(a, s);
		case 4: return launch_eval<EventEval<4, false>, W, BATCH>(a, s);
		case 5: return launch_eval<EventEval<5, false>, W, BATCH>(a, s);
		}
	}
	return hipErrorInvalidValue;
}

}  // namespace

bool event_path_applies(const DevPlan& p) {
	return p.ev_ok && p.n_descs >= 2 && p.n_descs <= 5;
}

uint32_t event_tile_rows(int rows) {
	return rows == 1 ? 1u : 2u;
}

hipError_t launch_scan_event(const ScanArgs& a, int nd, bool wide, int rows, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	return event_tile_rows(rows) == 1 ? launch_event_rows<1, false>(a, nd, wide, s)
									  : launch_event_rows<2, false>(a, nd, wide, s);
}

}  // namespace pars_hip
