// BAM queries for several threads at once (the engine pool): each caller brings its own arrays.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "rabbitvar_amd.h"

namespace rv {

struct BamQuery {
    std::vector<int32_t> pos0, nm, mpos0, tid, mtid;
    std::vector<uint16_t> flag;
    std::vector<uint8_t> mapq, has_nm, bases, quals;
    std::vector<int64_t> seq_off, cigar_off;
    std::vector<uint32_t> cigar;
};

// rv_bam_query into `q` (read-only on the open file, so concurrent calls with distinct `q` are safe)
rv_status bam_query_into(const rv_bam* b, const char* chr, int32_t start, int32_t end, BamQuery& q, rv_reads* out,
                         std::string& err);

}  // namespace rv
