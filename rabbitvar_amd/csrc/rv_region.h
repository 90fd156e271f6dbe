// One region through the whole path: pileup (device) -> realignment -> scoring (device) -> text.
#pragma once
#include <string>
#include <vector>

#include "rabbitvar_amd.h"
#include "rv_engine.h"

namespace rv {

struct CallStrings {
    std::string refallele, varallele, genotype, bias, leftseq, rightseq, vartype;
};

struct CallSet {
    std::vector<rv_call> calls;
    std::vector<CallStrings> strings;  // backing storage for the rv_call string pointers
    std::string text;
};

rv_status run_region(Engine& e, const rv_region& region, const rv_reads& reads, const rv_contig& contig, CallSet& out);

}  // namespace rv
