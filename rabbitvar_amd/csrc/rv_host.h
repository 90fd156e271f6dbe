// Host-side bookkeeping of the engine: ordered reduction of the walk's event and clip records
// into per-allele records, the soft-clip aggregates, and text helpers.  None of this re-runs a
// device stage; it consumes what the device produced.
#pragma once
#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "rv_common.h"

namespace rv {

struct HVar {
    int cnt = 0, fwd = 0, rev = 0, hicnt = 0, locnt = 0, extracnt = 0;
    double pmean = 0, qmean = 0, mqmean = 0, nm = 0;
    bool pstd = false, qstd = false;
    int pp = 0;
    double pq = 0;
};

// Iteration order of a robin_hood::unordered_map<char, int> (robin_hood 3.4.3, as vendored by the
// reference in include/robin_hood.h) that only ever received insertions, given the keys in
// insertion order.  The reference's soft-clip consensus breaks ties in that order
// (include/VariationUtils.h:431-443).
std::vector<char> rh_char_iteration_order(const std::vector<char>& inserted);
// the same for at most 7 distinct keys, without allocating (out holds n entries)
int rh_char_order_small(const char* inserted, int n, char* out);

// One clip index of a soft-clip accumulator: the reference's Sclip.nt[i] (base counts, a
// robin_hood map kept here in insertion order) and Sclip.seq[i] (per-base variations, an ordered
// map kept here sorted by base).
struct ClipColumn {
    std::vector<std::pair<char, int>> nt;   // (base, count) in first-seen order
    std::vector<std::pair<char, HVar>> seq;  // sorted by base
    HVar& seq_at(char b) {  // seq[i][b], created when absent
        auto it = std::lower_bound(seq.begin(), seq.end(), b,
                                   [](const std::pair<char, HVar>& e, char k) { return e.first < k; });
        if (it == seq.end() || it->first != b) it = seq.insert(it, {b, HVar{}});
        return it->second;
    }
    const HVar* seq_find(char b) const {
        for (auto& e : seq)
            if (e.first == b) return &e.second;
        return nullptr;
    }
};

// Soft-clip accumulator at one position (the reference's Sclip).  Clip records add indices
// 0 .. len - 1, so the reference's per-index maps are dense here: col[i] is index i.
struct HSclip {
    HVar v;
    std::vector<ClipColumn> col;
    std::string sequence;
    bool used = false;
    // the consensus as the reduction computed it (findconseq takes it over on its first call)
    std::string pre;
    bool pre_valid = false, pre_used = false;
};

// VariationUtils.h:367-394 islowcomplexseq
bool low_complexity(const std::string& s);
// findconseq's computation without its side effects (see rv_realign.h findconseq)
std::string clip_consensus(const HSclip& sc, bool& mark_used);

struct HostPileup {
    int dense_lo = 0, dense_n = 0;
    std::vector<DenseRow> rows;                           // dense single-base rows (window)
    std::map<int, std::map<std::string, HVar>> nd;        // other nonInsertionVariants entries
    std::map<int, std::map<std::string, HVar>> ins;       // insertionVariants
    std::map<int, HSclip> sc5, sc3;
    // the clusters in realignlgins30's order (count descending, position ascending), built by the
    // reduction; valid while sc5_of / sc3_of name these maps (a copied pileup rebuilds them)
    std::vector<std::pair<int, HSclip*>> sc5_order, sc3_order;
    const void* sc5_of = nullptr;
    const void* sc3_of = nullptr;
    std::map<int, std::map<std::string, int>> delCount, insCount, mnp;
    int maxReadLength = 0;
};

// Reduces the walk's records in read order (ord, seq) per (map, position, descriptor).
void reduce_events(const Event* ev, size_t n, const char* arena, double goodq, HostPileup& P);
// side = 5 or 3: only that side's records (its map is written, the other is untouched); -1: both
void reduce_clips(const ClipRec* cl, size_t n, const char* clip_arena, double goodq, HostPileup& P, int side = -1);

// the clusters of a clip map in count-descending, position-ascending order
void clip_order(std::map<int, HSclip>& m, std::vector<std::pair<int, HSclip*>>& ord);

// HVar view of a dense row's base (pileup-stage values)
HVar dense_to_hvar(const DenseBase& d);

// Text dump of the pileup-stage state, in the oracle's dump layout (tests compare the two).
std::string dump_pileup(const HostPileup& P);

std::string fmt_f(double v);  // std::to_string(double)

}  // namespace rv
