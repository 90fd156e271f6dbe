// Output line order of simple mode.
//
// The reference prints a region's calls walking alignedVariants in robin_hood slot order
// (src/modes/simpleMode.cpp:146-207); alignedVariants is filled walking nonInsertionVariants in
// slot order (src/ToVarsBuilder.cpp:97-205), and nonInsertionVariants received its keys in read
// order during the CIGAR walk (src/parseCigar.cpp, getVariation in include/VariationUtils.h:148-216),
// then a few insertions and erasures from the realigner.  The engine records, per position, when
// its key was created (the walk's first touch), and the realigner logs its key changes.
//
// Two ways to the same order:
//  * replay: every operation through RHOrder (exact in all cases);
//  * sort: when the table never halved its info increment while at its final capacity (no key's
//    displacement reached the overflow threshold), robin_hood's slot order is exactly the keys
//    sorted by (home slot, info hash bits descending, insertion sequence): keys with different home
//    slots are laid out by home, equal homes are ordered by the info byte's hash bits at every
//    comparison, and equal hash bits by insertion (each rehash re-inserts in slot order).  The
//    sort is used when a displacement check on the packed result proves the premise, and the
//    replay otherwise.
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

// capacity of a robin_hood table after `max_count` keys were present at once (8 slots, doubled
// whenever an insertion finds the 80% load reached)
size_t rh_capacity(size_t max_count);
int rh_log2(size_t cap);
// largest displacement (slot - home) of keys laid out in slot order in a table of `cap` slots
int rh_max_displacement(const int32_t* slot_ordered, size_t n, size_t cap);
// keys (in insertion order) sorted by (home, hash bits desc, insertion index): a host radix sort
std::vector<int32_t> rh_sort_keys(const std::vector<int32_t>& keys, size_t cap);
// displacement at or above which an info byte overflows at the initial info increment
constexpr int RH_OVERFLOW_DISP = 6;

// nonInsertionVariants after realignment, in slot order (replay).
// created: keys in creation order (walk); log: the realigner's key insertions and erasures.
std::vector<int32_t> noninsertion_order(const std::vector<int32_t>& created, const std::vector<KeyOp>& log);
// the same in two steps: the walk's keys (may run beside the realigner), then the realigner's log
void replay_created(RHOrder& m, const std::vector<int32_t>& created);
std::vector<int32_t> replay_finish(RHOrder& m, const std::vector<KeyOp>& log);

// nonInsertionVariants after realignment by sorting, given the walk's keys already in slot order
// for rh_capacity(n_walk) (`walk_sorted`, e.g. from the device).  Returns false when the premise
// cannot be proven; `out` is then unspecified.
// `walk_maxd` = rh_max_displacement of walk_sorted at rh_capacity(walk_sorted.size()).
bool noninsertion_order_sorted(const std::vector<int32_t>& walk_sorted, int walk_maxd, const std::vector<KeyOp>& log,
                               std::vector<int32_t>& out);
bool noninsertion_order_sorted(const std::vector<int32_t>& walk_sorted, const std::vector<KeyOp>& log,
                               std::vector<int32_t>& out);

// alignedVariants in slot order, given the positions in processing order and, per position, what
// ToVarsBuilder did with it: 0 nothing, 1 added, 2 added then erased.
template <class StatusOf>
std::vector<int32_t> aligned_order(const std::vector<int32_t>& processing, StatusOf status_of) {
    // sorted form first
    std::vector<int32_t> kept;
    size_t count = 0, max_count = 0;
    bool transient = false;
    for (int32_t p : processing) {
        const int st = status_of(p);
        if (st == 0) continue;
        if (count + 1 > max_count) max_count = count + 1;
        if (st == 2) {
            transient = true;
            continue;
        }
        kept.push_back(p);
        ++count;
    }
    const size_t cap = rh_capacity(max_count);
    std::vector<int32_t> sorted = rh_sort_keys(kept, cap);
    // a transient key can lengthen one cluster by one slot while it is present
    const int limit = RH_OVERFLOW_DISP - (transient ? 1 : 0);
    if (rh_max_displacement(sorted.data(), sorted.size(), cap) < limit) return sorted;
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
