/* baseline.c — the timed CPU baseline: the C restatement of the reference's
 * CPU path (model::scanner<S>::scan, include/pars/model/scan.hpp:27-73, then
 * model::pair_brackets<S>, include/pars/model/nest.hpp:45-92; pars_oracle.c)
 * run over a set of documents, one after the other on the calling thread.
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline legs time it and compare
 * what it produced with the GPU's outputs; nothing in the product calls it.
 *
 * Only scan + pair are timed (CLOCK_MONOTONIC around each document's two
 * calls); output buffers are reused across documents, as a caller reusing
 * its vectors would.  Per-document counts and flags, and an FNV-1a 64 chain
 * over every document's positions and pairs (in document order, outside the
 * timed region), let the caller check the GPU's outputs against these. */
#include "pars_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

int oracle_time_docs(const pars_plan* S, const uint8_t* buf, const uint64_t* offs, const uint64_t* sizes, size_t nd,
					 uint32_t* doc_struct, uint32_t* doc_brackets, uint8_t* doc_flags, int32_t* doc_depth,
					 oracle_run* out) {
	memset(out, 0, sizeof *out);
	uint64_t cap = 0;
	for (size_t d = 0; d < nd; ++d)
		if (sizes[d] > cap)
			cap = sizes[d];
	uint32_t* pos  = (uint32_t*)malloc((cap ? cap : 1) * sizeof(uint32_t));
	uint32_t* bpos = (uint32_t*)malloc((cap ? cap : 1) * sizeof(uint32_t));
	uint8_t*  ch   = (uint8_t*)malloc(cap ? cap : 1);
	uint8_t*  bch  = (uint8_t*)malloc(cap ? cap : 1);
	int32_t*  pair = (int32_t*)malloc((cap ? cap : 1) * sizeof(int32_t));
	if (!pos || !bpos || !ch || !bch || !pair) {
		free(pos), free(bpos), free(ch), free(bch), free(pair);
		return -1;
	}
	uint64_t hp = 0xcbf29ce484222325ull, hq = 0xcbf29ce484222325ull;
	double	 secs = 0;
	int		 rc	  = 0;
	for (size_t d = 0; d < nd && rc == 0; ++d) {
		uint64_t	   counts[2];
		int32_t		   flags[3];
		const double   t0 = now_s();
		const int	   unc = oracle_scan(S, buf + offs[d], (size_t)sizes[d], pos, ch, bpos, bch, (size_t)cap, counts,
										 NULL, NULL);
		if (oracle_pair(S, bch, (size_t)counts[1], pair, flags) != 0)
			rc = -1;
		secs += now_s() - t0;
		out->n_struct += counts[0];
		out->n_brackets += counts[1];
		out->bytes += sizes[d];
		hp = oracle_fnv1a(hp, pos, (size_t)counts[0] * sizeof(uint32_t));
		hq = oracle_fnv1a(hq, pair, (size_t)counts[1] * sizeof(int32_t));
		const uint8_t f = (uint8_t)((flags[0] ? 1 : 0) | (flags[1] ? 2 : 0) | (unc ? 4 : 0));
		if (doc_struct)
			doc_struct[d] = (uint32_t)counts[0];
		if (doc_brackets)
			doc_brackets[d] = (uint32_t)counts[1];
		if (doc_flags)
			doc_flags[d] = f;
		if (doc_depth)
			doc_depth[d] = flags[2];
		if (nd == 1) {
			out->balanced	 = flags[0];
			out->well_formed = flags[1];
			out->max_depth	 = flags[2];
			out->unclosed	 = unc;
		}
	}
	out->seconds   = secs;
	out->hash_pos  = hp;
	out->hash_pair = hq;
	free(pos), free(bpos), free(ch), free(bch), free(pair);
	return rc;
}
