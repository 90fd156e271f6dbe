// One region through the whole path: pileup (device) -> realignment -> scoring (device) -> text.
#pragma once
#include <string>
#include <vector>

#include "rabbitvar_amd.h"
#include "rv_engine.h"

namespace rv {


struct CallSet {
    std::vector<rv_call> calls;
    std::vector<std::string> string_arenas;  // backing storage for the rv_call string pointers
    std::string text;
};

rv_status run_region(Engine& e, const rv_region& region, const rv_reads& reads, const rv_contig& contig, CallSet& out);

// Somatic mode: the first (tumor) and second (normal) samples of one region, classified and
// printed as somaticMode.cpp does (one_region_run_somt :83-128, output :311-735).  The call set
// carries the text only.
rv_status run_region_somatic(Engine& e, const rv_region& region, const rv_reads& first, const rv_reads& second,
                             const rv_contig& contig, CallSet& out);

// One launch of the scoring kernels over `in` (record list already uploaded), retried with larger
// record buffers when they overflow; copies back the records, the string arena and, when asked,
// the per-slot status.  hc receives the kernel counters.
// With `pinned` set, the records and the arena go to the engine's pinned host memory instead of
// the vectors (valid until its pinned_reset()).
struct RecSpan {
    const OutRec* p = nullptr;
    size_t n = 0;
    const char* arena = nullptr;
    const OutRec* begin() const { return p; }
    const OutRec* end() const { return p + n; }
};
rv_status score_pass(Engine& e, ScoreIn& in, size_t n_status, std::vector<OutRec>& recs, std::vector<char>& arena,
                     std::vector<uint8_t>* status, unsigned long long hc[8], float& score_ms,
                     RecSpan* pinned = nullptr);

}  // namespace rv
