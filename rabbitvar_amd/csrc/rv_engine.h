// Engine: one per device, driven by one host thread.  Owns the stream, the grow-only device
// buffers and the per-stage timers.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "rabbitvar_amd.h"
#include "rv_host.h"
#include "rv_kernels.h"
#include "rv_realign.h"

namespace rv {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct PileupResult {
    HostPileup P;        // reduced non-dense state + dense rows (rows only when requested)
    Window W;
    std::string ref_win;  // the reference window [W.ref_lo, W.ref_hi], upper case
    WalkCounters ctr;
    int64_t n_reads = 0, walked = 0;
};

// Points in pileup() where the caller can start work that overlaps the host reductions.
struct PileupHooks {
    std::function<void()> after_walk;                 // the dense rows are final on the device
    std::function<void(PileupResult&)> after_events;  // the event maps are reduced (clips still running)
};


// A fixed set of threads destroying posted objects (an engine's region state, the call sets it
// returned), shared by the engine and its call sets.
struct Job {
    virtual ~Job() = default;
};
template <class T>
struct Held : Job {
    explicit Held(T&& v) : obj(std::move(v)) {}
    T obj;
};
class ReleasePool {
  public:
    explicit ReleasePool(unsigned n) {
        for (unsigned i = 0; i < n; i++) th_.emplace_back([this] { loop(); });
    }
    ~ReleasePool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void post(std::unique_ptr<Job> j) {
        {
            std::lock_guard<std::mutex> g(m_);
            q_.push_back(std::move(j));
        }
        cv_.notify_one();
    }
    void wait() {
        std::unique_lock<std::mutex> g(m_);
        idle_.wait(g, [&] { return q_.empty() && busy_ == 0; });
    }

  private:
    void loop() {
        std::unique_lock<std::mutex> g(m_);
        for (;;) {
            cv_.wait(g, [&] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;  // stopping with nothing left
            std::unique_ptr<Job> j = std::move(q_.front());
            q_.pop_front();
            busy_++;
            g.unlock();
            j.reset();
            g.lock();
            busy_--;
            if (q_.empty() && busy_ == 0) idle_.notify_all();
        }
    }
    std::mutex m_;
    std::condition_variable cv_, idle_;
    std::deque<std::unique_ptr<Job>> q_;
    unsigned busy_ = 0;
    bool stop_ = false;
    std::vector<std::thread> th_;
};

class Engine {
  public:
    Engine(const rv_config& cfg, int device);
    ~Engine();

    // Runs prep + walk + anchor for one region; reduces the ordered records on the host.
    // When download_rows is set the whole dense window is copied back (diagnostics).
    rv_status pileup(const rv_region& region, const rv_reads& reads, const rv_contig& contig, bool download_rows,
                     PileupResult& out, const PileupHooks* hooks = nullptr);

    Config cfg;
    std::string sample;
    // -Z: the fraction, or a negative value when off.  Masks queued by the caller (the pool, which
    // draws in region order) are used first, one per pileup call, before drawing here.
    double downsampling = -1;
    std::vector<std::vector<uint8_t>> drop_queue;
    size_t drop_queue_pos = 0;
    std::vector<uint8_t> drop_host;
    // scored-record capacity that sufficed before (grow-only): later passes start there instead of
    // overflowing and relaunching
    unsigned score_out_hint = 0;
    unsigned long long score_arena_hint = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    rv_stats stats{};

    // device state of the last pileup (kept for the scoring stage)
    DenseRow* rows = nullptr;
    int rows_n = 0, rows_lo = 0;
    char* ref_dev = nullptr;
    // a finished region's reference window string, handed back for the next pileup to reuse
    std::string ref_spare;

    void* grow(DevBuf& b, size_t bytes);

    // ---- realignment screen (opt-in).  With a model set, the realigner skips the candidates the
    // network scores below the threshold; collect mode records features and change labels instead
    // (exact run).
    bool screen_on = false;
    float screen_threshold = 0.f;
    bool screen_collect = false;
    std::vector<float> collect_feat;     // [n][SCR_F] of the last collected region
    std::vector<uint8_t> collect_label;  // per candidate
    std::vector<int32_t> collect_kind;
    DevBuf b_scr_params;
    // candidate list, features and (when a model is set) keep flags for the last pileup
    // (events_done: screen_candidates_events already ran on pr.P)
    rv_status screen_plan(const PileupResult& pr, struct ScreenPlan& plan, bool want_features,
                          bool events_done = false);
    // copies the listed dense rows (indices into the window) of the last pileup to pinned host
    // memory, valid until pinned_reset(); nullptr on failure
    const DenseRow* fetch_rows(const std::vector<int>& rows);
    // pinned host memory for one region's transfers: blocks are kept and reused by the next region;
    // when a region starts, blocks the previous region did not use are returned beyond PINNED_KEEP
    // bytes in all (a large region's transfers do not stay page-locked while smaller ones follow)
    static constexpr size_t PINNED_KEEP = (size_t)1 << 30;
    void* pinned_alloc(size_t bytes);
    void pinned_reset();
    // row index -> fetched row table of the realigner's row cache, reused across regions (all null
    // between regions)
    std::vector<const DenseRow*> row_slots;
    // the next pileup writes its rows into the other buffer, so the last pileup's rows stay valid
    // (somatic mode keeps the first sample's rows while the second is piled up)
    void swap_row_buffers() { std::swap(b_rows, b_rows2); }
    // noPassingReads over the last pileup's records (device-resident): none = no record passes
    bool no_passing(int start, int end, bool& none);

    // nonInsertionVariants keys of the last pileup in creation order, and the same keys in the
    // slot order they take when sorted (see rv_order.h), both kept on the device; n_keys = their
    // count, max_disp = the largest displacement of the sorted layout (INT_MAX: no sorted layout)
    rv_status key_order(int n_reads, int64_t& n_keys, int& max_disp);
    // copies those keys to the host: creation order, or slot order
    rv_status download_keys(bool slot_order, std::vector<int32_t>& out);

    // The slot order of a robin_hood<int> table after inserting ops[0, m) in order from empty,
    // each key with tr[i] set erased right after its insertion (distinct keys), computed on the
    // device growth by growth (rv_slotorder.hip).  applied = false (out unspecified) when a run is
    // too long for the kernels or the table would grow early; the caller then orders on the host.
    rv_status slot_order(const int32_t* ops, const uint8_t* tr, size_t m, std::vector<int32_t>& out, bool& applied);
    // The same for alignedVariants of the last scoring pass, from device data: the processing
    // order proc_dev (n_proc positions) filtered by the scoring status (status_dev: dense_n rows then
    // ncx record-list positions cx_pos_dev, which override their rows; status 2 = erased after its
    // insertion).  When the device declines, applied = false and ops / tr hold the operations for
    // the host; otherwise tr holds the transient flags only.
    rv_status aligned_slot_order(const int32_t* proc_dev, size_t n_proc, uint8_t* status_dev, int dense_lo, int dense_n,
                                 const int32_t* cx_pos_dev, int ncx, std::vector<int32_t>& out, bool& applied,
                                 std::vector<int32_t>& ops, std::vector<uint8_t>& tr, bool download = true);
    // the epochs over operations already on the device (tr: the same flags on the host); with
    // download unset the order stays on the device for record_order
    rv_status slot_order_dev(const int32_t* ops_d, const uint8_t* tr_d, const uint8_t* tr, size_t m,
                             std::vector<int32_t>& out, bool& applied, bool download = true);
    // The printer's record order from the last applied slot order on the device: the n scored
    // records (recs_dev, as scoring wrote them) by (their position's index in that order, records
    // of positions without a key last; position; allele order).  ok = false when there is no such
    // order or a record's fields do not fit the sort key (the caller sorts on the host).
    rv_status record_order(const OutRec* recs_dev, int n, int dense_lo, int dense_n, std::vector<uint32_t>& perm,
                           bool& ok);
    // copies the last applied slot order to the host
    rv_status download_slot_order(std::vector<int32_t>& out);
    const int32_t* so_final_ = nullptr;  // the last applied slot order (device) and its length
    size_t so_n_ = 0;
    bool so_valid_ = false;
    float slot_order_ms = 0;  // device time of the last slot order's epochs
    // the last key_order's keys in slot order on the device (valid until the next key_order)
    const int32_t* slot_sorted_dev() const { return key_slot_ok_ ? (const int32_t*)b_kpos3.p : nullptr; }
    size_t key_n_ = 0;
    bool key_slot_ok_ = false;

    // the device Fisher test on host tables (n11 n12 n21 n22) -> (two-sided, left, right) per table
    rv_status fisher_tables(const int32_t* tab, int n, double* out);
    // find35match for a batch of consensus pairs on the device (realignlgins30's phase A); false on a
    // device failure (the realigner then scores the pairs itself)
    bool find35(const std::vector<F35Job>& jobs, std::vector<F35>& out);
    int64_t f35_device_pairs = 0;  // pairs scored on the device in the last region
    float f35_ms = 0;              // their kernel time
    // the host-built log-gamma table the device Fisher test reads (FISHER_LGT_N entries), uploaded
    // on first use; nullptr on failure (err set)
    const double* lgamma_table();

    // buffers of the realignment / scoring stages
    DevBuf b_gather_ids, b_gather_out, b_cx_pos, b_cx_off, b_cx, b_cx_arena, b_cx_cov, b_cx_key, b_out, b_out_arena,
        b_score_ctr, b_status, b_keys, b_keys2, b_kpos, b_kpos2, b_sort_tmp, b_kcount, b_kpos3, b_args, b_nopass,
        b_want, b_dup, b_drop, b_scr_cand, b_scr_pos, b_scr_cum, b_scr_feat, b_scr_score, b_fisher_in, b_fisher_out,
        b_lgamma, b_f35_pairs, b_f35_arena, b_f35_out, b_pick, b_pre, b_so_ops, b_so_tr, b_so_ord[2], b_so_k[4], b_so_w[4],
        b_so_runs, b_so_small, b_so_tmp, b_so_proc, b_so_rank, b_so_rk[4];
    WalkArgs* args_dev = nullptr;  // the last pileup's kernel arguments (its reads stay resident)
    int64_t args_n = 0;
    hipEvent_t sev[2] = {nullptr, nullptr};
    hipEvent_t so_ev[2] = {nullptr, nullptr};  // slot_order's own timing (it runs beside other stages)
    // figures of the last region: record-list positions, dense rows fetched for the
    // realigner, insertions whose coverage came from the next position
    int64_t last_cx_positions = 0, last_rows_fetched = 0, last_insertion_side_effects = 0;
    rv_status fail(rv_status s, const std::string& msg) {
        err = msg;
        return s;
    }

    // Destroys the given objects on the engine's release threads, each object a job of its own, so
    // that releasing a region's host working state (maps of hundreds of thousands of nodes)
    // overlaps the caller's next steps and finishes before the next region's reductions allocate.
    // wait_release() returns once every posted object is destroyed; so does the engine's destructor.
    template <class... T>
    void release_later(T&&... objs) {
        (release_->post(std::make_unique<Held<std::decay_t<T>>>(std::forward<T>(objs))), ...);
    }
    void wait_release() { release_->wait(); }
    const std::shared_ptr<ReleasePool>& release_pool() const { return release_; }

  private:
    std::shared_ptr<ReleasePool> release_ = std::make_shared<ReleasePool>(8);
    std::vector<std::pair<void*, size_t>> pinned_;
    size_t pinned_cur_ = 0, pinned_off_ = 0;
    DevBuf b_reads[13], b_prep, b_cig, b_ctr, b_span, b_rows, b_rows2, b_ref;
    DevBuf b_events, b_arena, b_clips, b_clip_arena, b_anchors, b_counts;
    hipEvent_t ev[8];
};

// -Z: draws the C library's rand() for each record that passes the flag filter, in order; drop[i]
// is 1 when the draw is <= RAND_MAX * fraction (recordPreprocessor.cpp:133)
void downsample_draws(const uint16_t* flag, int64_t n, int samfilter, double fraction, uint8_t* drop);

}  // namespace rv
