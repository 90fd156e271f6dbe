// Kernel launch interface of the pileup / scoring device stages.
#pragma once
#include <hip/hip_runtime.h>

#include "rv_pileup_body.h"
#include "rv_score.h"

namespace rv {

hipError_t launch_prep(const WalkArgs* A, int64_t n, PrepRec* prep, uint32_t* cig, WalkCounters* ctr, int* max_span,
                       bool dedup, hipStream_t s);
hipError_t launch_fill_rows(DenseRow* rows, int n, hipStream_t s);
hipError_t launch_walk(const WalkArgs* A, int64_t n, const PrepRec* prep, const uint32_t* cig, int nreads,
                       hipStream_t s);
hipError_t launch_no_passing(const WalkArgs* A, int64_t n, int start, int end, unsigned* count, hipStream_t s);
hipError_t launch_anchor(const Anchor* a, const unsigned* n_anchors, unsigned cap, DenseRow* rows, int dense_lo,
                         int dense_n, const Config& cfg, hipStream_t s);

static_assert(sizeof(DenseRow) % 16 == 0, "dense rows are copied in 16-byte words");
hipError_t launch_gather_rows(const DenseRow* rows, int dense_n, const int32_t* idx, int n, DenseRow* out,
                              hipStream_t s);

hipError_t launch_key_list(const DenseRow* rows, int dense_n, int dense_lo, unsigned long long* keys, int32_t* pos,
                           unsigned* count, hipStream_t s);
hipError_t launch_slot_keys(const int32_t* keys, int n, int cap_log2, int idx_bits, unsigned long long* comp,
                            int32_t* val, hipStream_t s);
// temp == nullptr: sets temp_bytes only
hipError_t sort_key_pairs(void* temp, size_t& temp_bytes, const unsigned long long* kin, unsigned long long* kout,
                          const int32_t* vin, int32_t* vout, int n, int end_bit, hipStream_t s);

hipError_t launch_score(const ScoreIn& S, hipStream_t s);
hipError_t launch_fisher_table(const int* tab, int n, double* out, const double* lgt, int lgt_n, hipStream_t s);
// the allocation-free variant's whole call (host tables in, host p-values out; rv_score.hip)
constexpr int FISHER_NOALLOC_MAX_TABLES = 1 << 16;
hipError_t fisher_tables_noalloc(const int* tab, int n, double* out, const char** err);

}  // namespace rv
