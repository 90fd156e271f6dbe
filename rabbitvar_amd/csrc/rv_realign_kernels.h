// Device matching kernels of the realigner (phase A): candidate scoring that depends only on the
// candidate's own strings, computed for a batch of candidates before the host's ordered pass
// (phase B) consumes the results in the reference's order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rv {

// find35match (reference src/VariationRealigner.cpp:1396-1428) for a batch of (5' consensus,
// 3' consensus) pairs.  pair[4 * k .. 4 * k + 3] = offset and length of seq5, offset and length of
// seq3 in `arena`; out[3 * k .. 3 * k + 2] = b5, b3, score.  Strings longer than F35_MAX_LEN are
// not scored on the device (out = -1, -1, -1: the host computes those).
constexpr int F35_MAX_LEN = 1024;
struct F35Result {
    int32_t b5, b3, score;
};
hipError_t launch_find35(const int32_t* pair, int n, const char* arena, F35Result* out, hipStream_t s);

}  // namespace rv
