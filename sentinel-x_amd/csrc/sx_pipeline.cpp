// sx_pipeline.cpp -- device filter cascade and autotuning entry points.
#include "sx_pipeline.h"

#include "../../include/sx_hmmer.h"

extern "C" int sx_run_cascade(sx_ctx*, int, sx_cascade_stats*, sx_pair_scores*, uint64_t, uint64_t*) {
    return SX_ERR_STATE;
}
