// Kernel launch interface of the pileup / scoring device stages.
#pragma once
#include <hip/hip_runtime.h>

#include "rv_pileup_body.h"

namespace rv {

hipError_t launch_prep(const ReadsDev& R, const Window& W, const Config& cfg, PrepRec* prep, uint32_t* cig,
                       WalkCounters* ctr, int* max_span, hipStream_t s);
hipError_t launch_fill_rows(DenseRow* rows, int n, hipStream_t s);
hipError_t launch_walk(const ReadsDev& R, const Window& W, const Config& cfg, const PrepRec* prep, const uint32_t* cig,
                       const WalkSink& sink, hipStream_t s);
hipError_t launch_anchor(const Anchor* a, const unsigned* n_anchors, unsigned cap, DenseRow* rows, int dense_lo,
                         int dense_n, const Config& cfg, hipStream_t s);

}  // namespace rv
