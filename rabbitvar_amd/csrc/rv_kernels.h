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

// ---- robin_hood slot order by growth epochs (rv_slotorder.hip)
constexpr int SO_EV_MAX = 1024;    // events of one simulated run (more: too_big)
constexpr int SO_LEN_MAX = 1024;   // slots of one simulated run
constexpr int SO_GROUP_MAX = 64;   // keys of one home
struct SoEpoch {
    const int32_t* prev;  // the previous table's keys in slot order (re-inserted first)
    int np;
    const int32_t* ops;  // the epoch's operations (keys) and transient flags
    const uint8_t* tr;
    int S;       // np + operations
    int n_keep;  // keys present at the epoch's end
    uint32_t mask;
    int lcap;
};
struct SoRuns {  // per run of the epoch's final packing (sized for n_keep runs)
    int32_t *lo, *hi, *a, *b, *maxd, *tlo, *tn;
    int* n_runs;
};
struct SoHalve {  // steps before which the info increment halves
    int nh;
    uint32_t at[4];
};
// alignedVariants' operations: the entries of `proc` (n positions, the processing order) whose
// status is nonzero, in order, after the record-list positions' status overrides their dense rows'
// (status holds dense_n rows then ncx record-list positions).  bcount/boff: n / SO_CB + 2 entries;
// boff[n / SO_CB rounded up] = the number of operations.
constexpr int SO_CB = 2048;
hipError_t launch_so_ops(const int32_t* proc, int n, uint8_t* status, int dense_lo, int dense_n, const int32_t* cx_pos,
                         int ncx, int32_t* bcount, int32_t* boff, void* tmp, size_t tmp_bytes, int32_t* ops, uint8_t* tr,
                         hipStream_t s);
hipError_t so_ops_scan_bytes(int nb, size_t& tmp_bytes);
hipError_t launch_so_epoch_keys(const SoEpoch& E, uint32_t* skey, uint32_t* sval, hipStream_t s);
// record sort keys for the printer's order: (index of the record's position in `order`, or last;
// position - pos_base; allele order), with idx = record index; *bad set when a field does not fit
// (rank_of: dense_n entries, overwritten)
hipError_t launch_so_rank(const int32_t* order, int n_keep, const OutRec* recs, int n, int dense_lo, int dense_n,
                          long long pos_base, int32_t* rank_of, unsigned long long* key, int32_t* idx, unsigned* bad,
                          hipStream_t s);
// slots, runs (maxd, transient ranges) of the sorted steps
hipError_t launch_so_runs(const uint32_t* skey, int n_keep, int S, uint32_t mask, int32_t* v, int32_t* pm,
                          int32_t* flag, int32_t* rid, const SoRuns& R, void* tmp, size_t tmp_bytes, hipStream_t s);
hipError_t launch_so_simulate(const uint32_t* skey, const uint32_t* sval, uint32_t mask, const SoRuns& R, int max_runs,
                              unsigned* first, unsigned* too_big, hipStream_t s);
// the halving steps from `first` (into *H, on the device), then each key's slot order into out
hipError_t launch_so_final(const uint32_t* skey, const uint32_t* sval, const SoEpoch& E, const unsigned* first, SoHalve* H,
                           int32_t* out, unsigned* too_big, unsigned* declined, hipStream_t s);
hipError_t so_sort_pairs(void* tmp, size_t& tmp_bytes, const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                         uint32_t* vout, int n, int end_bit, hipStream_t s);
hipError_t so_scan_bytes(int n, size_t& tmp_bytes);

// the largest displacement of n keys laid out in slot order in a table of 2^cap_log2 slots, into
// *out (which must hold 0 beforehand); v, pm: n entries; temp == nullptr sets temp_bytes only
hipError_t launch_max_displacement(const int32_t* keys, int n, int cap_log2, int32_t* v, int32_t* pm, int* out,
                                   void* temp, size_t& temp_bytes, hipStream_t s);

hipError_t launch_score(const ScoreIn& S, hipStream_t s);
// k_score_prepick into S.pick_list / S.pick_count (dense_n entries, then the count); S needs the
// rows, the reference window and the configuration only
hipError_t launch_score_prepick(const ScoreIn& S, hipStream_t s);
hipError_t launch_fisher_table(const int* tab, int n, double* out, const double* lgt, int lgt_n, hipStream_t s);
// the allocation-free variant's whole call (host tables in, host p-values out; rv_score.hip)
constexpr int FISHER_NOALLOC_MAX_TABLES = 1 << 16;
hipError_t fisher_tables_noalloc(const int* tab, int n, double* out, const char** err);

}  // namespace rv
