// Output line order of simple mode.
//
// The reference prints a region's calls walking alignedVariants in robin_hood slot order
// (src/modes/simpleMode.cpp:146-207); alignedVariants is filled walking nonInsertionVariants in
// slot order (src/ToVarsBuilder.cpp:97-205), and nonInsertionVariants received its keys in read
// order during the CIGAR walk (src/parseCigar.cpp, getVariation in include/VariationUtils.h:148-216),
// then a few insertions and erasures from the realigner.  The engine records, per position, when
// its key was created (the walk's first touch), and the realigner logs its key changes; replaying
// those operations through RHOrder reproduces both maps' slot orders exactly.
#pragma once
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "rv_rhorder.h"

namespace rv {

struct KeyOp {
    int32_t pos;
    bool insert;  // false: erase
};

// nonInsertionVariants after realignment, in slot order.
// created: keys in creation order (walk); log: the realigner's key insertions and erasures.
std::vector<int32_t> noninsertion_order(const std::vector<int32_t>& created, const std::vector<KeyOp>& log);
// the same in two steps: the walk's keys (may run beside the realigner), then the realigner's log
void replay_created(RHOrder& m, const std::vector<int32_t>& created);
std::vector<int32_t> replay_finish(RHOrder& m, const std::vector<KeyOp>& log);

// alignedVariants in slot order, given the positions in processing order and, per position, what
// ToVarsBuilder did with it: 0 nothing, 1 added, 2 added then erased.
template <class StatusOf>
std::vector<int32_t> aligned_order(const std::vector<int32_t>& processing, StatusOf status_of) {
    RHOrder m;
    for (int32_t p : processing) {
        const int st = status_of(p);
        if (st == 0) continue;
        m.insert(p);
        if (st == 2) m.erase(p);
    }
    std::vector<int32_t> out;
    out.reserve(m.size());
    m.for_each([&](int32_t k) { out.push_back(k); });
    return out;
}

}  // namespace rv
