// scan_wide_kernels.hip — the single-document launchers of the tile kernel
// with six- or eight-wave blocks (the autotuner's waves-per-block knob) for
// the bit-parallel evaluators; the four-wave launchers are in scan_kernels.hip.
#include "evaluators.hpp"

namespace pars_hip {

hipError_t launch_scan_wide(const ScanArgs& a, bool parity, int rows, int waves, hipStream_t s) {
	if (a.n == 0)
		return hipSuccess;
	return parity ? launch_waves<ParityEval>(a, rows, 4, waves, s) : launch_waves<EscapeEval>(a, rows, 4, waves, s);
}

}  // namespace pars_hip
