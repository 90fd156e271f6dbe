// Per-position scoring (device): restates ToVarsBuilder::process for one position
// (reference src/ToVarsBuilder.cpp:97-205 and the helpers it calls, :213-1116), the simple-mode
// output filter (src/modes/simpleMode.cpp:144-208 with include/Variant.h:99-248) and the two-sided
// Fisher exact test the printer adds with --fisher (src/modes/simpleMode.cpp:96-108; htslib's
// kt_fisher_exact, whose published algorithm is restated in fisher_two_sided below).
//
// One thread scores one position.  Positions whose alleles are all single-base A/C/G/T rows read
// the dense pileup table; positions the host touched (non-dense descriptors, insertions,
// realignment results) read a CSR record list.  Printed variants are appended to an output
// buffer with their strings in an arena; the host only formats numbers.
//
// This header holds the layouts shared with the host; the device code is in rv_score_body.h.
#pragma once
#include <hip/hip_runtime.h>

#include "rv_common.h"

namespace rv {

struct HVarD {
    int32_t cnt, fwd, rev, hicnt, locnt, extracnt;
    double pmean, qmean, mqmean, nm;
    int32_t pstd, qstd;
};

struct CxEntry {
    int32_t kind;  // 0 nonInsertionVariants, 1 insertionVariants
    uint32_t desc_off;
    int32_t desc_len;
    int32_t pad;
    HVarD v;
};

struct OutRec {
    int32_t position, order, start, end;
    int32_t tpc, poscov, rfc, rrc, fwd, rev, shift3, msint, hicnt, hicov, pstd, qstd, vartype;
    int32_t nfields;  // 5 for a pileup line without a reference variant
    double freq, pmean, qmean, mq, hlr, hifreq, extrafreq, msi, nm, fisher_p, fisher_ratio;
    uint64_t str_off;  // byte offset into the string arena (the arena can exceed 4 GiB)
    uint16_t len[7];  // refallele, varallele, genotype, bias, leftseq, rightseq, descriptor
    int32_t kind;     // OUT_VARIANT, or (somatic passes) OUT_REF / OUT_REF_ONLY / OUT_VARIANT_RAW
    int32_t start_rank;  // host: the position's rank in the printer's order
};
static_assert(sizeof(OutRec) == 192, "OutRec layout (records are copied to the host as they are)");

// OutRec kinds.  Simple mode emits printed variants only.  The somatic passes emit a position's
// Vars: every non-reference variant in CMP_VARI order plus its reference variant (OUT_REF, as
// createVariant left it), or, on request, the reference-only form of a position without variants
// (OUT_REF_ONLY, ToVarsBuilder.cpp:1074-1097), and the variants a failed reference lookup left as
// createVariant made them (OUT_VARIANT_RAW).
enum : int { OUT_VARIANT = 0, OUT_REF = 1, OUT_REF_ONLY = 2, OUT_VARIANT_RAW = 3 };
// ScoreIn::somatic
enum : int { SCORE_SIMPLE = 0, SCORE_SOMATIC_VARS = 1, SCORE_SOMATIC_REF = 2 };

enum : int { VT_SNV = 0, VT_INS = 1, VT_DEL = 2, VT_COMPLEX = 3, VT_EMPTY = 4, VT_SV = 5 };

struct ScoreIn {
    const DenseRow* rows;
    int dense_lo, dense_n;
    const int32_t* cx_pos;
    int n_cx;
    const int32_t* cx_off;
    const CxEntry* cx;
    const char* cx_arena;
    const int32_t* cx_cov;
    const uint8_t* cx_key;
    const char* ref;
    int ref_lo, ref_hi, chr_len;
    int region_start, region_end;
    Config cfg;
    OutRec* out;
    unsigned* n_out;
    unsigned max_out;
    char* arena;
    unsigned long long* arena_used;
    unsigned long long arena_cap;
    uint8_t* status;  // per dense row, then per record-list position: 1 = alignedVariants entry kept,
                      // 2 = added and erased again (nothing above -f), 0 = not added; the somatic
                      // passes: 1 = kept with variants, 3 = kept without (reference only)
    int somatic;                // SCORE_*
    const uint8_t* want;        // SCORE_SOMATIC_REF: per status slot, emit the reference-only record
    unsigned long long* counters;  // [0] capacity overflows, [1] insertion side effects on p+1,
                                   // [2] sorts above 16 elements, [3] qstd ambiguous rows
    const double* lgamma_tab;      // host-built kf_lgamma of the integers (fisher_lgamma_table)
    int lgamma_n;
    // dense positions to score, picked by k_score_pick (dense_n entries, then the count); nullptr:
    // the dense kernel visits every dense position
    int32_t* pick_list;
    unsigned* pick_count;
    // the dense positions that pass every check but the record-list one, listed by
    // k_score_prepick while the realigner runs (those checks read only the pileup's rows and the
    // reference); with it, k_score_filter drops the record-list positions from this list instead
    // of k_score_pick reading every row again.  nullptr: k_score_pick
    const int32_t* pre_list;
    const unsigned* pre_count;
};

}  // namespace rv
