// Region driver: pileup (device) -> realignment (host, on the changed positions only) -> scoring
// (device) -> simple-mode text.  Replaces the body of one_region_run and the printer that follows
// it (src/modes/simpleMode.cpp:18-64, :144-208; src/ToVarsBuilder.cpp:97-205).
#include "rv_region.h"
#include "rv_fisher.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "rv_format.h"
#include "rv_order.h"
#include "rv_realign.h"

namespace rv {

namespace {

// The scored records as score_pass downloaded them into this region's pinned host memory.
struct RecView {
    const OutRec* p = nullptr;
    size_t n = 0;
    size_t size() const { return n; }
    const OutRec& operator[](size_t i) const { return p[i]; }
};

// slot[i] = index of recs[i].position in `order` (distinct keys); left as it is when absent.  A
// bitmap of the records' positions filters the scan, which is split over threads.
void slot_ranks(const std::vector<int32_t>& order, const RecView& recs, std::vector<int64_t>& slot,
                unsigned threads) {
    const size_t n = recs.size();
    if (n == 0 || order.empty()) return;
    std::vector<std::pair<int32_t, uint32_t>> by_pos(n);
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (size_t i = 0; i < n; i++) {
        by_pos[i] = {recs[i].position, (uint32_t)i};
        lo = std::min(lo, recs[i].position);
        hi = std::max(hi, recs[i].position);
    }
    std::sort(by_pos.begin(), by_pos.end());
    const size_t span = (size_t)((int64_t)hi - lo + 1);
    std::vector<uint64_t> bits((span + 63) / 64, 0);
    for (const auto& p : by_pos) {
        const size_t k = (size_t)((int64_t)p.first - lo);
        bits[k >> 6] |= 1ull << (k & 63);
    }
    auto scan = [&](size_t a, size_t b) {
        for (size_t r = a; r < b; r++) {
            const int32_t k = order[r];
            if (k < lo || k > hi) continue;
            const size_t x = (size_t)((int64_t)k - lo);
            if (!((bits[x >> 6] >> (x & 63)) & 1u)) continue;
            auto it = std::lower_bound(by_pos.begin(), by_pos.end(), std::make_pair(k, 0u));
            for (; it != by_pos.end() && it->first == k; ++it) slot[it->second] = (int64_t)r;
        }
    };
    const size_t m = order.size();
    const size_t nt = m < (1u << 18) ? 1 : std::max(1u, std::min(threads, 8u));
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(scan, m * t / nt, m * (t + 1) / nt);
    scan(0, m / nt);
    for (auto& th : pool) th.join();
}

class DeviceRowSource : public RowSource {
  public:
    explicit DeviceRowSource(Engine& e) : e_(e) {}
    const DenseRow* fetch(const std::vector<int>& rows) override { return e_.fetch_rows(rows); }

  private:
    Engine& e_;
};

// noPassingReads on the device over the pileup's records (they stay resident in HBM)
class DeviceReadQuery : public ReadQuery {
  public:
    explicit DeviceReadQuery(Engine& e) : e_(e) {}
    bool no_passing(int start, int end) override {
        bool none = true;
        if (!e_.no_passing(start, end, none)) throw std::runtime_error("noPassingReads on the device failed");
        return none;
    }
    bool find35_batch(const std::vector<F35Job>& jobs, std::vector<F35>& out) override { return e_.find35(jobs, out); }

  private:
    Engine& e_;
};

}  // namespace

#define RV_HIPS(x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return e.fail(RV_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

rv_status score_pass(Engine& e, ScoreIn& in, size_t n_status, std::vector<OutRec>& recs, std::vector<char>& arena,
                     std::vector<uint8_t>* status, unsigned long long hc[8], float& score_ms, RecSpan* pinned) {
    unsigned max_out = (unsigned)std::min<int64_t>(65536 + 8 * (int64_t)in.n_cx + in.dense_n / 16, 0x7fffffff);
    max_out = std::max(max_out, e.score_out_hint);
    unsigned long long arena_cap = std::max(160ull * max_out, e.score_arena_hint);
    in.lgamma_tab = e.lgamma_table();
    in.lgamma_n = FISHER_LGT_N;
    if (!in.lgamma_tab) return RV_ERR_DEVICE;
    if (in.dense_n > 0) {
        // the picked dense positions, then the count (64-entry aligned)
        const size_t list_n = ((size_t)in.dense_n + 63) & ~(size_t)63;
        in.pick_list = (int32_t*)e.grow(e.b_pick, sizeof(int32_t) * (list_n + 64));
        if (!in.pick_list) return e.fail(RV_ERR_OOM, "device buffer for the scored positions failed");
        in.pick_count = (unsigned*)(in.pick_list + list_n);
    }
    for (int attempt = 0; attempt < 3; attempt++) {
        in.out = (OutRec*)e.grow(e.b_out, sizeof(OutRec) * max_out);
        in.arena = (char*)e.grow(e.b_out_arena, (size_t)arena_cap);
        unsigned long long* ctr = (unsigned long long*)e.grow(e.b_score_ctr, 64);
        if (!in.out || !in.arena || !ctr) return e.fail(RV_ERR_OOM, "device buffers for the scored records failed");
        RV_HIPS(hipMemsetAsync(ctr, 0, 64, e.stream));
        RV_HIPS(hipMemsetAsync(in.status, 0, n_status, e.stream));
        in.n_out = (unsigned*)ctr;
        in.arena_used = ctr + 1;
        in.counters = ctr + 2;
        in.max_out = max_out;
        in.arena_cap = arena_cap;
        RV_HIPS(hipEventRecord(e.sev[0], e.stream));
        RV_HIPS(launch_score(in, e.stream));
        RV_HIPS(hipEventRecord(e.sev[1], e.stream));
        RV_HIPS(hipMemcpyAsync(hc, ctr, 64, hipMemcpyDeviceToHost, e.stream));
        RV_HIPS(hipStreamSynchronize(e.stream));
        RV_HIPS(hipEventElapsedTime(&score_ms, e.sev[0], e.sev[1]));
        unsigned n_out = (unsigned)(hc[0] & 0xffffffffu);
        unsigned long long used = hc[1];
        if (n_out <= max_out && used <= arena_cap) {
            e.score_out_hint = std::max(e.score_out_hint, max_out);
            e.score_arena_hint = std::max(e.score_arena_hint, arena_cap);
            if (hc[2] != 0) {
                // a position needed more alleles or string space than a scoring thread holds
                return e.fail(RV_ERR_CAPACITY, "scoring: " + std::to_string(hc[2]) +
                                                   " position(s) exceeded the per-thread allele/string capacity");
            }
            OutRec* rdst;
            char* adst;
            if (pinned) {
                rdst = (OutRec*)e.pinned_alloc(sizeof(OutRec) * std::max(n_out, 1u));
                adst = (char*)e.pinned_alloc((size_t)used + 1);
                if (!rdst || !adst) return e.fail(RV_ERR_OOM, "pinned host memory for the scored records failed");
                adst[used] = 0;
                *pinned = RecSpan{rdst, n_out, adst};
            } else {
                recs.resize(n_out);
                arena.resize((size_t)used + 1);
                rdst = recs.data();
                adst = arena.data();
            }
            if (n_out) RV_HIPS(hipMemcpyAsync(rdst, in.out, sizeof(OutRec) * n_out, hipMemcpyDeviceToHost, e.stream));
            if (used) RV_HIPS(hipMemcpyAsync(adst, in.arena, (size_t)used, hipMemcpyDeviceToHost, e.stream));
            if (status) {
                status->resize(n_status);
                RV_HIPS(hipMemcpyAsync(status->data(), in.status, n_status, hipMemcpyDeviceToHost, e.stream));
            }
            RV_HIPS(hipStreamSynchronize(e.stream));
            return RV_OK;
        }
        max_out = std::max<unsigned>(max_out, n_out) * 2;
        arena_cap = std::max(arena_cap, used) * 2;
    }
    return e.fail(RV_ERR_CAPACITY, "scored record buffers kept overflowing");
}

rv_status run_region(Engine& e, const rv_region& region, const rv_reads& reads, const rv_contig& contig, CallSet& out) {
    const auto w0 = std::chrono::steady_clock::now();
    auto ms_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    e.pinned_reset();
    e.f35_device_pairs = 0;
    e.f35_ms = 0;
    PileupResult pr;
    // Work that overlaps the host reductions inside pileup(): the nonInsertionVariants key order
    // (a device sort) once the dense rows are final, and the realigner's first row batch once the
    // event maps are reduced.
    std::vector<int32_t> created, walk_sorted;  // copied from the device only when the host needs them
    int64_t n_walk_keys = 0;
    int walk_maxd_dev = INT_MAX;
    rv_status key_status = RV_OK;
    double key_ms = 0;
    std::thread key_th;
    DeviceRowSource src(e);
    std::unique_ptr<RowCache> rcp;
    HostRef ref;
    bool prefetch_ok = false;
    double prefetch_ms = 0;
    std::exception_ptr prefetch_err;
    std::thread prefetch_th;
    const int32_t* pre_list = nullptr;  // k_score_prepick's list and count, when it was launched
    const unsigned* pre_count = nullptr;
    // realignment screen: its MNV / deletion / insertion candidates are listed while the clips
    // are still being reduced
    const bool screening = e.screen_on || e.screen_collect;
    ScreenPlan plan;
    std::thread screen_th;
    struct Joiner {
        std::thread& t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } key_joiner{key_th}, prefetch_joiner{prefetch_th}, screen_joiner{screen_th};
    // the engine's row table must be all null again when this region ends, on every path
    struct SlotRelease {
        std::unique_ptr<RowCache>& rc;
        ~SlotRelease() {
            if (rc) rc->release_slots();
        }
    } slot_release{rcp};
    PileupHooks hooks;
    hooks.after_walk = [&] {
        key_th = std::thread([&] {
            const auto k0 = std::chrono::steady_clock::now();
            key_status = e.key_order((int)std::min<int64_t>(reads.n_reads, INT32_MAX), n_walk_keys, walk_maxd_dev);
            key_ms = ms_since(k0);
        });
    };
    hooks.after_events = [&](PileupResult& p) {
        ref.seq = p.ref_win.data();
        ref.lo = p.W.ref_lo;
        ref.hi = p.W.ref_lo + (int)p.ref_win.size() - 1;
        rcp = std::make_unique<RowCache>(&src, p.P.dense_lo, p.P.dense_n, &e.row_slots);
        const HostPileup* hp = &p.P;
        if (screening) screen_th = std::thread([&, hp] { screen_candidates_events(*hp, plan); });
        prefetch_th = std::thread([&, hp] {
            try {
                const auto f0 = std::chrono::steady_clock::now();
                prefetch_ok = realign_prefetch(e.cfg, ref, contig.length, *hp, *rcp);
                prefetch_ms = ms_since(f0);
                // the scoring's row checks (the realigner changes no dense row on the device) run
                // on the device while the realigner works on the host
                if (prefetch_ok && hp->dense_n > 0) {
                    ScoreIn q{};
                    q.rows = e.rows;
                    q.dense_lo = hp->dense_lo;
                    q.dense_n = hp->dense_n;
                    q.ref = e.ref_dev;
                    q.ref_lo = ref.lo;
                    q.ref_hi = ref.hi;
                    q.cfg = e.cfg;
                    const size_t list_n = ((size_t)hp->dense_n + 63) & ~(size_t)63;
                    q.pick_list = (int32_t*)e.grow(e.b_pre, sizeof(int32_t) * (list_n + 64));
                    if (q.pick_list) {
                        q.pick_count = (unsigned*)(q.pick_list + list_n);
                        if (launch_score_prepick(q, e.stream) == hipSuccess) {
                            pre_list = q.pick_list;
                            pre_count = q.pick_count;
                        }
                    }
                }
            } catch (...) {
                prefetch_err = std::current_exception();
            }
        });
    };
    rv_status s = e.pileup(region, reads, contig, false, pr, &hooks);
    e.stats.timeline_ms[0] = ms_since(w0);
    if (key_th.joinable()) key_th.join();
    if (prefetch_th.joinable()) prefetch_th.join();
    const bool screen_events_done = screen_th.joinable();
    if (screen_th.joinable()) screen_th.join();
    if (s != RV_OK) return s;
    if (key_status != RV_OK) return key_status;
    if (!rcp) return e.fail(RV_ERR_DEVICE, "pileup finished without reducing its records");
    if (prefetch_err) std::rethrow_exception(prefetch_err);
    if (!prefetch_ok) return e.fail(RV_ERR_DEVICE, "dense row prefetch failed");
    e.stats.timeline_ms[1] = ms_since(w0);
    RowCache& rc = *rcp;
    double order_ms = 0;
    e.stats.detail_ms[6] = key_ms;
    auto h0 = std::chrono::steady_clock::now();
    // a second host thread checks whether the sorted slot order provably holds and otherwise
    // replays the walk's insertions, while this one realigns
    RHOrder nonins;
    int walk_maxd = RH_OVERFLOW_DISP;
    bool replayed = false;
    std::exception_ptr replay_err;
    std::thread replay([&] {
        try {
            walk_maxd = walk_maxd_dev;
            if (walk_maxd >= RH_OVERFLOW_DISP) {
                if (e.download_keys(false, created) != RV_OK) throw std::runtime_error(e.err);
                replay_created(nonins, created);
                replayed = true;
            }
        } catch (...) {
            replay_err = std::current_exception();
        }
    });
    Joiner joiner{replay};
    // ---- realignment screen (screened mode, or feature/label collection)
    ScreenPlan* planp = nullptr;
    e.stats.screen_ms = 0;
    e.stats.screen_list_ms = 0;
    e.stats.screen_candidates = 0;
    e.stats.screen_skipped = 0;
    if (screening) {
        const auto q0 = std::chrono::steady_clock::now();
        plan.label = e.screen_collect;
        s = e.screen_plan(pr, plan, e.screen_collect, screen_events_done);
        if (s != RV_OK) return s;
        planp = &plan;
        e.stats.screen_ms = ms_since(q0);
        e.stats.screen_candidates = (int64_t)plan.cand.size();
    }
    // ---- realignment on the host
    RState S(pr.P, rc);
    DeviceReadQuery rq(e);
    RealignStats rst;
    try {
        realign(e.cfg, region.start, region.end, ref, contig.length, &rq, S, rst, true, planp);
    } catch (const std::exception& ex) {
        return e.fail(RV_ERR_DEVICE, std::string("realignment: ") + ex.what());
    }
    if (planp) {
        e.stats.screen_skipped = plan.skipped;
        if (e.screen_collect) {
            e.collect_label = plan.changed;
            e.collect_kind.resize(plan.cand.size());
            for (size_t i = 0; i < plan.cand.size(); i++) e.collect_kind[i] = plan.cand[i].kind;
        }
    }
    rst.ms_prefetch = prefetch_ms;
    e.stats.detail_ms[0] = rst.ms_prefetch;
    e.stats.detail_ms[1] = rst.ms_mnp;
    e.stats.detail_ms[2] = rst.ms_del;
    e.stats.detail_ms[3] = rst.ms_ins;
    e.stats.detail_ms[4] = rst.ms_lgins;
    e.stats.detail_ms[12] = rst.ms_nopass;
    e.stats.detail_ms[13] = rc.miss_ms;
    e.stats.detail_ms[14] = (double)rst.n_nopass;
    e.stats.detail_ms[15] = (double)rc.miss_fetches;
    const auto c0 = std::chrono::steady_clock::now();
    e.stats.timeline_ms[2] = ms_since(w0);
    CxHost cx;
    try {
        build_record_list(S, cx);
    } catch (const std::exception& ex) {
        return e.fail(RV_ERR_DEVICE, std::string("record list: ") + ex.what());
    }
    e.stats.detail_ms[5] = ms_since(c0);
    double host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    const double realign_ms = host_ms;
    // ---- scoring on the device
    const int ncx = (int)cx.pos.size();
    ScoreIn in{};
    in.rows = e.rows;
    in.dense_lo = pr.P.dense_lo;
    in.dense_n = pr.P.dense_n;
    in.pre_list = pre_list;
    in.pre_count = pre_count;
    in.n_cx = ncx;
    in.ref = e.ref_dev;
    in.ref_lo = ref.lo;
    in.ref_hi = ref.hi;
    in.chr_len = contig.length;
    in.region_start = region.start;
    in.region_end = region.end;
    in.cfg = e.cfg;
    // through pinned staging: from pageable memory each copy would be a synchronous, staged one
    auto up = [&](DevBuf& b, const void* src, size_t bytes) -> void* {
        void* d = e.grow(b, bytes);
        if (d && bytes) {
            void* h = e.pinned_alloc(bytes);
            if (!h) return nullptr;
            std::memcpy(h, src, bytes);
            if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, e.stream) != hipSuccess) return nullptr;
        }
        return d;
    };
    in.cx_pos = (const int32_t*)up(e.b_cx_pos, cx.pos.data(), 4 * cx.pos.size());
    in.cx_off = (const int32_t*)up(e.b_cx_off, cx.off.data(), 4 * cx.off.size());
    in.cx = (const CxEntry*)up(e.b_cx, cx.ent.data(), sizeof(CxEntry) * cx.ent.size());
    in.cx_arena = (const char*)up(e.b_cx_arena, cx.arena.data(), cx.arena.size());
    in.cx_cov = (const int32_t*)up(e.b_cx_cov, cx.cov.data(), 4 * cx.cov.size());
    in.cx_key = (const uint8_t*)up(e.b_cx_key, cx.key.data(), cx.key.size());
    if (!in.cx_pos || !in.cx_off || !in.cx || !in.cx_arena || !in.cx_cov || !in.cx_key)
        return e.fail(RV_ERR_OOM, "device buffers for the record list failed");
    const size_t n_status = (size_t)pr.P.dense_n + (size_t)ncx;
    in.status = (uint8_t*)e.grow(e.b_status, n_status);
    if (!in.status) return e.fail(RV_ERR_OOM, "device buffer for position status failed");
    unsigned long long hc[8];
    float score_ms = 0;
    e.stats.timeline_ms[3] = ms_since(w0);
    // the records and their strings come back into pinned host memory (a copy into pageable
    // memory would be staged); the status stays on the device
    RecSpan rspan;
    {
        std::vector<OutRec> unused_recs;
        std::vector<char> unused_arena;
        s = score_pass(e, in, n_status, unused_recs, unused_arena, nullptr, hc, score_ms, &rspan);
    }
    if (s != RV_OK) return s;
    const RecView recs{rspan.p, rspan.n};
    const char* const arena_p = rspan.arena;
    e.stats.timeline_ms[4] = ms_since(w0);
    // ---- text, in the printer's order: alignedVariants slot order, then the variants' order.  A
    // second thread works out that order while this one formats the lines in record order; the
    // lines are then joined in printing order.
    auto t0 = std::chrono::steady_clock::now();
    replay.join();
    e.stats.detail_ms[7] = ms_since(t0);
    if (replay_err) return e.fail(RV_ERR_DEVICE, "key order replay failed");
    e.stats.order_sorted = replayed ? 0 : 1;  // alignedVariants' bit is added below
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t n = recs.size();
    std::vector<uint32_t> perm(n);  // record indices in printing order
    std::vector<int32_t> aligned;  // alignedVariants' keys in slot order
    int aligned_how = 0;           // 0 replay, 1 per-growth sort on the host, 3 on the device
    std::vector<int32_t> proc;
    std::exception_ptr order_err;
    double order_thread_ms = 0;
    std::thread order_th([&] {
        try {
            const auto q0 = std::chrono::steady_clock::now();
            // nonInsertionVariants' slot order: the device's sorted walk keys as they are when the
            // realigner changed no key and the sorted form provably holds, else rebuilt
            const std::vector<int32_t>* procp = &proc;
            const bool direct = !replayed && S.key_log.empty() && walk_maxd < RH_OVERFLOW_DISP;
            if (!direct) {
                if (!replayed && e.download_keys(true, walk_sorted) != RV_OK) throw std::runtime_error(e.err);
                if (replayed || !noninsertion_order_sorted(walk_sorted, walk_maxd, S.key_log, proc)) {
                    if (!replayed) {
                        if (e.download_keys(false, created) != RV_OK) throw std::runtime_error(e.err);
                        replay_created(nonins, created);
                    }
                    proc = replay_finish(nonins, S.key_log);
                }
            }
            // alignedVariants: its operations listed on the device from the processing order and
            // the scoring status (record-list positions override their dense rows), its slot order
            // there too; else the per-growth sort on the host, else the replay
            const auto l0 = std::chrono::steady_clock::now();
            const int32_t* proc_dev = direct ? e.slot_sorted_dev() : nullptr;
            const size_t n_proc = direct ? (size_t)n_walk_keys : procp->size();
            if (direct && !proc_dev && n_proc) throw std::runtime_error("walk keys missing on the device");
            if (!direct && !procp->empty()) {
                auto* d = (int32_t*)e.grow(e.b_so_proc, 4 * procp->size());
                if (!d || hipMemcpyAsync(d, procp->data(), 4 * procp->size(), hipMemcpyHostToDevice, e.stream) != hipSuccess)
                    throw std::runtime_error("upload of the processing order failed");
                proc_dev = d;
            }
            std::vector<int32_t> aops;
            std::vector<uint8_t> atr;
            bool on_device = false;
            if (e.aligned_slot_order(proc_dev, n_proc, in.status, pr.P.dense_lo, pr.P.dense_n, in.cx_pos, ncx,
                                     aligned, on_device, aops, atr, false) != RV_OK)
                throw std::runtime_error(e.err);
            e.stats.order_ops = (int64_t)atr.size();
            e.stats.order_device_ms = e.slot_order_ms;
            e.stats.order_ops_ms = ms_since(l0) - e.slot_order_ms;
            if (on_device)
                aligned_how = 3;
            else if (sorted_ops_order(aops.data(), atr.data(), aops.size(), aligned, std::min(hw, 8u)))
                aligned_how = 1;
            else
                aligned = replay_ops(aops, atr);
            const auto q1 = std::chrono::steady_clock::now();
            e.stats.detail_ms[8] = std::chrono::duration<double, std::milli>(q1 - q0).count();
            // iteration order is slot order: a record's rank is its start position's index there.
            // With the order on the device, the records (still in HBM as scoring wrote them) are
            // ranked and sorted there; otherwise on the host.
            bool perm_done = false;
            if (on_device) {
                std::vector<uint32_t> dp;
                if (e.record_order(in.out, (int)n, pr.P.dense_lo, pr.P.dense_n, dp, perm_done) != RV_OK)
                    throw std::runtime_error(e.err);
                if (perm_done) perm.swap(dp);
                else if (e.download_slot_order(aligned) != RV_OK)
                    throw std::runtime_error(e.err);
            }
            if (!perm_done) {
                std::vector<int64_t> slot(n, INT64_MAX);
                slot_ranks(aligned, recs, slot, hw);
                for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
                std::sort(perm.begin(), perm.end(), [&](uint32_t x, uint32_t y) {
                    if (slot[x] != slot[y]) return slot[x] < slot[y];
                    if (recs[x].position != recs[y].position) return recs[x].position < recs[y].position;
                    return recs[x].order < recs[y].order;
                });
            }
            e.stats.detail_ms[9] = ms_since(q1);
            order_thread_ms = ms_since(q0);
        } catch (...) {
            order_err = std::current_exception();
        }
    });
    struct OrderJoiner {
        std::thread& t;
        ~OrderJoiner() {
            if (t.joinable()) t.join();
        }
    } order_joiner{order_th};
    const auto f0 = std::chrono::steady_clock::now();
    LineContext ctx{e.sample.c_str(), region.gene ? region.gene : "", region.chr ? region.chr : "", region.start, region.end,
                    e.cfg.fisher != 0};
    // lines are independent: format contiguous chunks of records on several threads
    const size_t nt = n < 4096 ? 1 : std::min<size_t>(std::min(hw, 16u), n / 2048);
    std::vector<std::string> parts(nt);
    std::vector<uint64_t> line_at(n + 1);  // (chunk << 40) | offset of record i's line in its chunk
    // the call records' strings (alleles, genotype, bias, flanks, type), NUL-terminated, in one
    // arena per chunk; str_at[i] = their offsets in record i's chunk arena
    std::vector<std::string> sarenas(nt);
    std::vector<std::array<uint32_t, 7>> str_at(n);
    auto work = [&](size_t t) {
        const size_t a = n * t / nt, b = n * (t + 1) / nt;
        std::string& txt = parts[t];
        std::string& sa = sarenas[t];
        txt.reserve((b - a) * 200);
        sa.reserve((b - a) * 64);
        for (size_t i = a; i < b; i++) {
            const OutRec& o = recs[i];
            line_at[i] = ((uint64_t)t << 40) | (uint64_t)txt.size();
            format_line(txt, o, arena_p, ctx);
            const char* p = arena_p + o.str_off;
            for (int k = 0; k < 6; k++) {
                str_at[i][k] = (uint32_t)sa.size();
                sa.append(p, o.len[k]);
                sa += '\0';
                p += o.len[k];
            }
            str_at[i][6] = (uint32_t)sa.size();
            sa += vartype_name(o, arena_p);
            sa += '\0';
        }
    };
    {
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; t++) pool.emplace_back(work, t);
        if (nt > 0) work(0);
        for (auto& th : pool) th.join();
    }
    const double format_ms = ms_since(f0);
    e.stats.timeline_ms[5] = ms_since(w0);
    // while the order thread finishes: the output's buffers, whose sizes are known now
    out.text.clear();
    out.calls.clear();
    out.string_arenas.clear();
    {
        size_t total = 0;
        for (auto& pt : parts) total += pt.size();
        out.text.resize(total);
        out.calls.resize(n);
    }
    order_th.join();
    e.stats.timeline_ms[6] = ms_since(w0);
    if (order_err) {
        try {
            std::rethrow_exception(order_err);
        } catch (const std::exception& ex) {
            return e.fail(RV_ERR_DEVICE, std::string("output order: ") + ex.what());
        }
    }
    order_ms += order_thread_ms;
    e.stats.order_sorted |= (int64_t)aligned_how << 1;
    // join the lines in printing order
    auto line_len = [&](size_t i) {
        const size_t t = (size_t)(line_at[i] >> 40);
        const size_t off = (size_t)(line_at[i] & ((1ull << 40) - 1));
        const size_t b = n * (t + 1) / nt;  // first record of the next chunk
        size_t end = parts[t].size();
        if (i + 1 < b) end = (size_t)(line_at[i + 1] & ((1ull << 40) - 1));
        return std::make_pair(parts[t].data() + off, end - off);
    };
    // each line's offset in the text (printing order), then lines and call records in parallel; the
    // records' strings stay where formatting put them (the vector moves, its elements do not)
    std::vector<size_t> text_at(n + 1, 0);
    for (size_t k = 0; k < n; k++) text_at[k + 1] = text_at[k] + line_len(perm[k]).second;
    if (text_at[n] != out.text.size()) return e.fail(RV_ERR_DEVICE, "output lines do not add up to the text");
    out.string_arenas = std::move(sarenas);
    auto assemble = [&](size_t a, size_t b) {
        char* txt = n ? &out.text[0] : nullptr;
        for (size_t k = a; k < b; k++) {
            const size_t i = perm[k];
            const auto ln = line_len(i);
            std::memcpy(txt + text_at[k], ln.first, ln.second);
            const OutRec& o = recs[i];
            const char* sb = out.string_arenas[(size_t)(line_at[i] >> 40)].data();
            const std::array<uint32_t, 7>& so = str_at[i];
            rv_call& c = out.calls[k];
            c.position = o.position;
            c.start = o.start;
            c.end = o.end;
            c.total_pos_coverage = o.tpc;
            c.position_coverage = o.poscov;
            c.ref_fwd = o.rfc;
            c.ref_rev = o.rrc;
            c.var_fwd = o.fwd;
            c.var_rev = o.rev;
            c.shift3 = o.shift3;
            c.msint = o.msint;
            c.hicnt = o.hicnt;
            c.hicov = o.hicov;
            c.frequency = o.freq;
            c.mean_position = o.pmean;
            c.mean_quality = o.qmean;
            c.mean_mapq = o.mq;
            c.hi_lo_ratio = o.hlr;
            c.hi_freq = o.hifreq;
            c.extra_freq = o.extrafreq;
            c.msi = o.msi;
            c.nm = o.nm;
            c.fisher_p = o.fisher_p;
            c.fisher_ratio = o.fisher_ratio;
            c.pstd = o.pstd;
            c.qstd = o.qstd;
            c.refallele = sb + so[0];
            c.varallele = sb + so[1];
            c.genotype = sb + so[2];
            c.bias = sb + so[3];
            c.leftseq = sb + so[4];
            c.rightseq = sb + so[5];
            c.vartype = sb + so[6];
        }
    };
    e.stats.timeline_ms[14] = ms_since(w0);
    {
        const size_t na = n < 4096 ? 1 : std::min<size_t>(std::min(hw, 16u), n / 2048);
        std::vector<std::thread> pool;
        for (size_t t = 1; t < na; t++) pool.emplace_back(assemble, n * t / na, n * (t + 1) / na);
        assemble(0, n / na);
        for (auto& th : pool) th.join();
    }
    host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e.stats.realign_ms = realign_ms;
    e.stats.order_ms = order_ms;
    e.stats.format_ms = format_ms;
    e.stats.timeline_ms[15] = ms_since(w0);
    e.stats.wall_ms = ms_since(w0);
    e.stats.realign_device_pairs = e.f35_device_pairs;
    e.stats.realign_device_ms = e.f35_ms;
    e.stats.score_ms = score_ms;
    e.stats.host_ms += host_ms;
    e.stats.total_device_ms = e.stats.prep_ms + e.stats.walk_ms + e.stats.anchor_ms + score_ms;
    e.stats.cx_positions = ncx;
    e.stats.rows_fetched = rc.rows_fetched;
    e.last_cx_positions = ncx;
    e.last_rows_fetched = rc.rows_fetched;
    e.last_insertion_side_effects = (int64_t)hc[3];
    rcp->release_slots();
    // the reference window string goes back to the engine; the largest maps each go to their own
    // release job
    e.ref_spare = std::move(pr.ref_win);
    e.release_later(std::move(S.pos), std::move(S.pool), std::move(S.ins), std::move(pr.P.nd), std::move(pr.P.ins),
                    std::move(pr.P.sc5), std::move(pr.P.sc3), std::move(pr.P.delCount), std::move(pr.P.insCount),
                    std::move(pr.P.mnp), std::move(pr.P.rows));
    e.release_later(std::move(S), std::move(rcp), std::move(pr), std::move(cx), std::move(created),
                    std::move(walk_sorted), std::move(nonins), std::move(proc), std::move(aligned), std::move(parts),
                    std::move(line_at), std::move(str_at), std::move(perm));
    e.stats.timeline_ms[7] = ms_since(w0);
    return RV_OK;
}

}  // namespace rv
