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
// the same for any order of `keys`: the largest displacement once they are packed by home
int rh_packed_displacement(const int32_t* keys, size_t n, size_t cap);
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

// alignedVariants as a table, given the positions in processing order and, per position, what
// ToVarsBuilder did with it: status[p - status_lo] = 0 nothing, 1 added, 2 added then erased.  The
// positions are filtered on `threads` threads, then the operations are replayed.
void aligned_table(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo, int status_n,
                   RHOrder& out, unsigned threads = 1);
// the same as a list in slot order
std::vector<int32_t> aligned_order(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo,
                                   int status_n);

// Replays insert (and insert-then-erase) operations on a robin_hood<int> table from empty and
// returns its slot order.  ops[i] = key, transient[i] = erased right after its insertion.
std::vector<int32_t> replay_ops(const std::vector<int32_t>& ops, const std::vector<uint8_t>& transient);

// The same slot order without replaying, for operations on distinct keys.  Between two growths
// (an epoch) the table re-inserts the previous table's keys in its slot order, then takes the
// epoch's operations; robin_hood keeps keys sorted by home slot (no wrap-around) and, within a
// home, by the info byte's hash bits descending, equal bytes in placement order.  An info byte
// that would overflow (displacement 6, 14, 30, 62 at increments 32, 16, 8, 4) makes the next
// insertion halve the increment, dropping a hash bit.  Two keys of one home keep the order fixed
// when the later one was placed: by the hash bits at the level then in force, ties by placement.
// So the epoch's slot order is a sort by (home, hash bits of the final level descending, then per
// earlier level l: placed after halving l+1, else bit l descending; then placement index).
// The halving steps depend only on displacements, which depend only on how many keys share each
// home: every set present in an epoch is a subset of its final keys plus at most one transient
// key, so each run of the final packing is simulated on its own, and only runs that can reach
// a displacement of 6.  Returns false (`out` unspecified) when a displacement of 126 is reached
// (the table would grow early; the replay covers that case).
bool sorted_ops_order(const int32_t* ops, const uint8_t* transient, size_t n, std::vector<int32_t>& out,
                      unsigned threads = 1);

// The operations ToVarsBuilder makes on alignedVariants (see aligned_table): keys in order, and
// whether each is erased right after its insertion.
void aligned_op_list(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo, int status_n,
                     unsigned threads, std::vector<int32_t>& ops, std::vector<uint8_t>& transient);

// alignedVariants in slot order (as aligned_table), by sorted_ops_order where it applies and the
// replay otherwise.  `sorted` reports which.
std::vector<int32_t> aligned_slot_order(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo,
                                        int status_n, unsigned threads, bool* sorted = nullptr);

}  // namespace rv
