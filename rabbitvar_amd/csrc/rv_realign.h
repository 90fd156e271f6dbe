// Local realignment of one region (host side of the engine).
//
// The realignment stage of the reference (src/VariationRealigner.cpp:135-161) rewrites the pileup
// around candidate indels, MNVs and soft-clip clusters: a few thousand positions of a region whose
// pileup has millions.  It runs here on the host over a position store that materializes the
// device's dense rows lazily (block fetches through RowSource) and records which positions it
// changed; the scoring kernels then read changed positions from a record list and every other
// position straight from the dense rows still resident in HBM.
#pragma once
#include <climits>
#include <cstdint>
#include <deque>
#include <functional>
#include <iterator>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rv_host.h"
#include "rv_order.h"
#include "rv_score.h"
#include "rv_screen.h"

namespace rv {

// Source of dense rows (row i holds position dense_lo + i).  fetch() returns a copy of the listed
// rows that stays valid as long as the source does; nullptr on failure.
class RowSource {
  public:
    virtual ~RowSource() {}
    virtual const DenseRow* fetch(const std::vector<int>& rows) = 0;
};

// Host rows already downloaded (tests, diagnostics).
class HostRowSource : public RowSource {
  public:
    explicit HostRowSource(const HostPileup& P) : P_(P) {}
    const DenseRow* fetch(const std::vector<int>& rows) override;

  private:
    const HostPileup& P_;
    std::deque<std::vector<DenseRow>> blocks_;
};

// Rows the realigner has read, by position.  A miss copies a small neighbourhood of rows.
// The row index -> row table can be an external one reused across regions (all null when handed
// over; release_slots() nulls the entries this cache set), or the cache's own.
class RowCache {
  public:
    static constexpr int MISS_RADIUS = 16;
    RowCache(RowSource* src, int dense_lo, int dense_n, std::vector<const DenseRow*>* slots = nullptr)
        : src_(src), lo_(dense_lo), n_(dense_n), slots_(slots ? slots : &own_) {}
    // fetches, in one transfer, every row within `radius` of the given positions and the rows of
    // the `exact` positions
    bool prefetch(const std::vector<int>& positions, int radius, const std::vector<int>& exact = {});
    const DenseRow* row(int pos);  // nullptr outside the dense window
    // row() without fetching: false when the position's row is in the window but not loaded
    bool row_loaded(int pos, const DenseRow*& out) const {
        const int i = pos - lo_;
        out = nullptr;
        if (i < 0 || i >= n_) return true;
        if (!have(i)) return false;
        out = (*slots_)[(size_t)i];
        return true;
    }
    void release_slots();
    int64_t fetches = 0, rows_fetched = 0;
    int64_t miss_fetches = 0;
    double miss_ms = 0;   // time in fetches made on a miss (not by prefetch)
    double fetch_ms = 0;  // time in the source's fetches

  private:
    bool load(std::vector<int>& rows, bool sorted_unique = false);
    bool have(int i) const { return (size_t)i < slots_->size() && (*slots_)[(size_t)i] != nullptr; }
    RowSource* src_;
    int lo_, n_;
    std::vector<const DenseRow*> own_;
    std::vector<const DenseRow*>* slots_;
    std::vector<int> set_;  // entries of *slots_ this cache filled
};

// The reference window as the reference's map: keys [lo, hi]; get() of a missing key creates it
// holding '\0' (robin_hood operator[]), at() of a missing key throws.
struct HostRef {
    const char* seq = nullptr;
    int lo = 1, hi = 0;
    std::set<int> created;
    bool count(int k) const { return (k >= lo && k <= hi) || created.count(k) != 0; }
    char at(int k) const;
    char get(int k);
};

// noPassingReads (reference src/VariationRealigner.cpp:1447-1498) over the region's records: true
// when no record overlapping [start, end] spans it without carrying exactly this deletion.
// find35match's result for one (5' consensus, 3' consensus) pair
struct F35 {
    int b5 = 0, b3 = 0, score = 0;
};
struct F35Job {
    const std::string* s5;
    const std::string* s3;
};

class ReadQuery {
  public:
    virtual ~ReadQuery() = default;
    virtual bool no_passing(int start, int end) = 0;
    // find35match for a batch of pairs on the device (realignlgins30's candidate pairs, scored
    // before its ordered pass); false: not available, the realigner computes each pair itself.
    // A result with score -1 was not computed (the host computes it).
    virtual bool find35_batch(const std::vector<F35Job>& jobs, std::vector<F35>& out) {
        (void)jobs;
        (void)out;
        return false;
    }
};


// Read-only position -> value index over a std::map built once (open addressing, linear probing):
// one or two cache lines per lookup where the map's tree walk takes a dozen.  The map must not gain
// or lose entries while the index is used (values may change).
template <class V>
class PosIndex {
  public:
    // m: an ordered map keyed by position.  With threads > 1 the map's nodes (scattered over the
    // heap) are walked on that many threads, each over its own key range, and the entries then
    // inserted in one pass; lookups do not depend on the insertion order.
    template <class M>
    void build(M& m, unsigned threads = 1) {
        size_t cap = 16;
        while (cap < 2 * m.size() + 1) cap *= 2;
        keys_.assign(cap, EMPTY);
        vals_.assign(cap, nullptr);
        mask_ = cap - 1;
        if (threads <= 1 || m.size() < 4096) {
            for (auto& kv : m) put(kv.first, &kv.second);
            return;
        }
        const int64_t lo = m.begin()->first, span = (int64_t)std::prev(m.end())->first - lo + 1;
        std::vector<std::vector<std::pair<int, V*>>> part(threads);
        auto walk = [&](unsigned t) {
            auto a = m.lower_bound((int)(lo + span * t / threads));
            auto b = t + 1 == threads ? m.end() : m.lower_bound((int)(lo + span * (t + 1) / threads));
            for (; a != b; ++a) part[t].emplace_back(a->first, &a->second);
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < threads; t++) th.emplace_back(walk, t);
        walk(0);
        for (auto& x : th) x.join();
        for (auto& pt : part)
            for (auto& kv : pt) put(kv.first, kv.second);
    }
    V* find(int p) const {
        if (keys_.empty()) return nullptr;
        for (size_t i = slot(p);; i = (i + 1) & mask_) {
            if (keys_[i] == p) return vals_[i];
            if (keys_[i] == EMPTY) return nullptr;
        }
    }

  private:
    void put(int p, V* v) {
        size_t i = slot(p);
        while (keys_[i] != EMPTY) i = (i + 1) & mask_;
        keys_[i] = p;
        vals_[i] = v;
    }
    static constexpr int EMPTY = INT32_MIN;  // never a position
    size_t slot(int p) const { return (size_t)(((uint32_t)p * 0x9E3779B1u) >> 7) & mask_; }
    std::vector<int> keys_;
    std::vector<V*> vals_;
    size_t mask_ = 0;
};

// A realignment candidate: position, descriptor and its count in the candidate map.
struct RealignCand {
    int position;
    std::string desc;
    int count;
};

struct RPos {
    bool loaded = false;
    bool key = false;   // nonInsertionVariants has the position
    int cov = 0;        // refCoverage
    bool dirty = false;
    std::map<std::string, HVar*> v;  // nonInsertionVariants[pos]
};

// Pileup state the realigner reads and writes.  Records live in a pool so that pointers stay valid
// across renames and erasures, as the reference's heap-allocated Variation objects do.
class RState {
  public:
    RState(HostPileup& P, RowCache& rows);
    RPos& at(int pos);
    bool has_key(int pos) { return at(pos).key; }
    HVar* inner(int pos, const std::string& d);
    HVar* get_variation(int pos, const std::string& d);  // getVariation on nonInsertionVariants
    HVar* get_ins(int pos, const std::string& d);        // getVariation on insertionVariants
    HVar* maybe(int pos, char refBase);                  // getVariationMaybe
    // nonInsertionVariants[pos][d] as at() would find it, without materializing the position
    // (valid until the next call); nullptr when absent
    const HVar* peek(int pos, const std::string& d);
    // peek() split in two.  peek_loaded: the position's materialized entries (*hit false when
    // at() has not loaded it).  peek_static: what peek() returns for a position at() has not
    // loaded, from the dense rows and the non-dense map alone, without fetching and without
    // writing anything (safe on several threads): 0 absent, 1 found (copied to *out), 2 unknown
    // (the position's dense row is not loaded).
    const HVar* peek_loaded(int pos, const std::string& d, bool* hit) const;
    int peek_static(int pos, const std::string& d, HVar* out) const;
    void erase_inner(int pos, const std::string& d, bool drop_empty);
    void add_cov(int pos, int k);
    HVar* new_var();

    HostPileup& P;
    RowCache& rows;
    PosIndex<std::map<std::string, HVar>> nd_ix;  // P.nd by position
    PosIndex<HSclip> sc5_ix, sc3_ix;               // P.sc5 / P.sc3 by position
    std::deque<HVar> pool;
    std::unordered_map<int, RPos> pos;
    std::map<int, std::map<std::string, HVar*>> ins;
    std::vector<KeyOp> key_log;  // nonInsertionVariants key insertions / erasures, in order
    // adjustMNP's candidates in its order, sorted while the position indexes are built
    // (mnp_sorted_ready until adjustMNP takes them)
    std::vector<RealignCand> mnp_sorted;
    bool mnp_sorted_ready = false;

    // Change tracking for the screen's labels: between track_begin() and track_end() every
    // position read through at() or track_ins() is snapshotted once; track_end() reports whether
    // any of them now differs (key, coverage, descriptors or their counts and sums).
    void track_begin();
    bool track_end();
    void track_ins(int pos);  // call before reading or changing ins[pos] directly

  private:
    struct Snap {
        bool key = false;
        int cov = 0;
        bool present = false;
        std::vector<std::pair<std::string, HVar>> v;
    };
    HVar peek_tmp_;
    bool track_on_ = false;
    std::unordered_map<int, Snap> track_n_, track_i_;
};

// The realignment screen's plan for one region.  Candidates are listed kind by kind: adjustMNP,
// realigndel and realignins candidates in the order those stages visit them (count descending,
// position ascending, descriptor descending), then realignlgins30's 5' clip clusters by position.
struct ScreenPlan {
    std::vector<ScreenCand> cand;
    int off[SC_NKIND + 1] = {0, 0, 0, 0, 0};
    std::vector<int32_t> tpos[ST_N];  // neighbourhood tables (rv_screen.h)
    std::vector<int64_t> tcum[ST_N];
    std::vector<ScreenCand> lists[SC_NKIND];                  // per kind, while being built
    std::vector<std::pair<int, int64_t>> per_pos[SC_NKIND];   // per kind: (position, candidates there)
    std::vector<uint8_t> keep;     // per candidate; empty: visit every candidate
    bool label = false;            // fill `changed`
    std::vector<uint8_t> changed;  // per candidate: the visit changed a count (unvisited: 0)
    int64_t skipped = 0;
};
void screen_candidates(const HostPileup& P, ScreenPlan& plan);
// the same in two parts: the MNV / deletion / insertion candidates once the event maps are
// reduced, then the clip clusters and the merged tables once the clip maps are
void screen_candidates_events(const HostPileup& P, ScreenPlan& plan);
void screen_candidates_finish(const HostPileup& P, ScreenPlan& plan);

struct RealignStats {
    int64_t candidates = 0, touched = 0, changed = 0;
    double ms_prefetch = 0, ms_mnp = 0, ms_del = 0, ms_ins = 0, ms_lgins = 0;
    double ms_nopass = 0;  // noPassingReads, inside the handlers' times
    int64_t n_nopass = 0;
};

// The realigner's first batch of dense rows, in one transfer: every candidate position and the
// four after it, the non-dense and insertion positions, and the positions the deletion
// candidates' flank mismatches will read.  Reads only the maps the event reduction fills (not the
// soft-clip maps), so it can run while the clips are still being reduced.
bool realign_prefetch(const Config& c, HostRef& ref, int chrLen, const HostPileup& P, RowCache& rows);

// adjustMNP, realigndel, realignins, realignlgins30 in the reference's order (after the prefetch
// above unless `prefetched`).
void realign(const Config& c, int region_start, int region_end, HostRef& ref, int chrLen, ReadQuery* reads,
             RState& S, RealignStats& st, bool prefetched = false,
             ScreenPlan* plan = nullptr);

// Positions the scoring stage reads from a record list instead of the dense rows: every
// non-dense / insertion position and every position the realigner changed, ascending, with the
// nonInsertionVariants entries (key order) followed by the insertionVariants entries (key order).
struct CxHost {
    std::vector<int32_t> pos, off, cov;
    std::vector<uint8_t> key;
    std::vector<CxEntry> ent;
    std::string arena;
};
void build_record_list(RState& S, CxHost& cx);

// soft-clip consensus (include/VariationUtils.h:403-509)
std::string findconseq(HSclip& sc);
// the cluster's consensus, computed once (findconseq stores it in the cluster)
const std::string& consensus(HSclip& sc);
// find35match (reference src/VariationRealigner.cpp:1396-1428): where the 5' clip consensus and the
// reversed 3' clip consensus overlap; score 0 when they do not
void find35match(const std::string& seq5, const std::string& seq3, int& b5, int& b3, int& score);

}  // namespace rv
