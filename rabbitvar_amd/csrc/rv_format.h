// Simple-mode output lines (src/modes/simpleMode.cpp:18-64): the 36 tab-separated fields of
// print_output_variant_simple, 38 with --fisher.
#pragma once
#include <string>

#include "rabbitvar_amd.h"
#include "rv_score.h"

namespace rv {

struct LineContext {
    const char* sample;
    const char* gene;
    const char* chr;
    int region_start, region_end;
    bool fisher;
};

// the printed variant type (Variant.h:75-97) of a scored record
std::string vartype_name(const OutRec& o, const char* arena);

// appends v as printf's "%f" (std::to_string of a double)
void put_f(std::string& out, double v);

// appends one output line for a scored record
void format_line(std::string& out, const OutRec& o, const char* arena, const LineContext& ctx);

}  // namespace rv
