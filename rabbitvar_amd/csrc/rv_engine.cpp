// Engine implementation: device orchestration of the pileup stage.
#include "rv_engine.h"

#include "rv_order.h"
#include "rv_fisher.h"
#include "rv_realign_kernels.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>

namespace rv {

#define RV_HIP(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(RV_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static Config to_config(const rv_config& c) {
    Config k;
    k.goodq = c.goodq;
    k.bias = c.bias;
    k.freq = c.freq;
    k.qratio = c.qratio;
    k.mapq = c.mapq;
    k.lofreq = c.lofreq;
    k.vext = c.vext;
    k.mismatch = c.mismatch;
    k.trimBasesAfter = c.trimBasesAfter;
    k.performLocalRealignment = c.performLocalRealignment;
    k.indelsize = c.indelsize;
    k.minBiasReads = c.minBiasReads;
    k.minr = c.minr;
    k.moveIndelsTo3 = c.moveIndelsTo3;
    k.samfilter = c.samfilter;
    k.readPosFilter = c.readPosFilter;
    k.doPileup = c.doPileup;
    k.minmatch = c.minmatch;
    k.uniqueModeAlignmentEnabled = c.uniqueModeAlignmentEnabled;
    k.uniqueModeSecondInPairEnabled = c.uniqueModeSecondInPairEnabled;
    k.fisher = c.fisher;
    k.SVMINLEN = c.SVMINLEN;
    k.referenceExtension = c.referenceExtension;
    k.numberNucleotideToExtend = c.numberNucleotideToExtend;
    k.mappingQuality = c.mappingQuality;
    k.removeDuplicatedReads = c.removeDuplicatedReads;
    k.crisprCuttingSite = c.crisprCuttingSite;
    k.hasBam2 = 0;
    return k;
}

// -Z (recordPreprocessor.cpp:133): every record that passes the flag filter draws one value of the
// C library's rand(), in iterator order, and is dropped when the draw is <= RAND_MAX * fraction.
// The reference never seeds rand(), so its draws form one process-wide sequence; calling the same
// rand() here continues that sequence.
void downsample_draws(const uint16_t* flag, int64_t n, int samfilter, double fraction, uint8_t* drop) {
    const double limit = RAND_MAX * fraction;
    for (int64_t i = 0; i < n; i++) drop[i] = ((flag[i] & samfilter) == 0 && rand() <= limit) ? 1 : 0;
}

Engine::Engine(const rv_config& c, int dev) {
    cfg = to_config(c);
    sample = c.sample ? c.sample : "";
    if (c.isDownsampling) downsampling = c.downsampling;
    device = dev;
    if (hipSetDevice(dev) != hipSuccess) {
        err = "hipSetDevice failed";
        return;
    }
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        err = "hipStreamCreate failed";
        return;
    }
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) err = "hipEventCreate failed";
    for (auto& e : sev)
        if (hipEventCreate(&e) != hipSuccess) err = "hipEventCreate failed";
    for (auto& e : so_ev)
        if (hipEventCreate(&e) != hipSuccess) err = "hipEventCreate failed";
}

Engine::~Engine() {
    wait_release();
    for (auto& blk : pinned_) (void)hipHostFree(blk.first);
    DevBuf* all[] = {&b_prep,      &b_cig,     &b_ctr,    &b_span,       &b_rows,     &b_ref,    &b_events,
                     &b_arena,     &b_clips,   &b_clip_arena, &b_anchors, &b_counts,   &b_gather_ids, &b_gather_out,
                     &b_cx_pos,    &b_cx_off,  &b_cx,     &b_cx_arena,   &b_cx_cov,   &b_cx_key, &b_out,
                     &b_out_arena, &b_score_ctr, &b_status, &b_keys, &b_keys2, &b_kpos, &b_kpos2, &b_sort_tmp,
                     &b_kcount, &b_args, &b_kpos3, &b_nopass, &b_rows2, &b_want, &b_dup, &b_drop,
                     &b_scr_cand, &b_scr_pos, &b_scr_cum, &b_scr_feat, &b_scr_score, &b_scr_params,
                     &b_fisher_in, &b_fisher_out, &b_lgamma, &b_f35_pairs, &b_f35_arena, &b_f35_out, &b_pick, &b_pre,
                     &b_so_ops, &b_so_tr, &b_so_ord[0], &b_so_ord[1], &b_so_k[0], &b_so_k[1], &b_so_k[2], &b_so_k[3],
                     &b_so_w[0], &b_so_w[1], &b_so_w[2], &b_so_w[3], &b_so_runs, &b_so_small, &b_so_tmp,
                     &b_so_proc, &b_so_rank, &b_so_rk[0], &b_so_rk[1], &b_so_rk[2], &b_so_rk[3]};
    // teardown: nothing to report a failure to
    for (auto* b : all)
        if (b->p) (void)hipFree(b->p);
    for (auto& b : b_reads)
        if (b.p) (void)hipFree(b.p);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : sev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : so_ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

void* Engine::grow(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return b.p;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = bytes + bytes / 4;
    if (hipMalloc(&b.p, want) != hipSuccess) return nullptr;
    b.cap = want;
    return b.p;
}

// Reference window (recordPreprocessor.cpp:41-78): keys [max(1, s-x-ext), min(len, e+x+ext) - 17]
static void window_bounds(const Config& c, const rv_region& r, int chrLen, int& lo, int& hi) {
    int ext = c.referenceExtension;
    int ss = r.start - c.numberNucleotideToExtend - ext < 1 ? 1 : r.start - c.numberNucleotideToExtend - ext;
    int se = r.end + c.numberNucleotideToExtend + ext > chrLen ? chrLen : r.end + c.numberNucleotideToExtend + ext;
    lo = ss;
    hi = se - 17;
}

rv_status Engine::pileup(const rv_region& region, const rv_reads& rd, const rv_contig& contig, bool download_rows,
                         PileupResult& out, const PileupHooks* hooks) {
    if (cfg.crisprCuttingSite != 0) return fail(RV_ERR_UNSUPPORTED, "CRISPR mode is not implemented");
    auto h0 = std::chrono::steady_clock::now();
    double host_ms = 0;
    const int64_t n = rd.n_reads;
    out.n_reads = n;
    // ---- reads on the device
    ReadsDev R;
    R.n = n;
    if (rd.on_device) {
        R.pos0 = rd.pos0;
        R.flag = rd.flag;
        R.mapq = rd.mapq;
        R.has_nm = rd.has_nm;
        R.nm = rd.nm;
        R.mpos0 = rd.mpos0;
        R.tid = rd.tid;
        R.mtid = rd.mtid;
        R.seq_off = rd.seq_off;
        R.cig_off = rd.cigar_off;
        R.bases = rd.bases;
        R.quals = rd.quals;
        R.cigar = rd.cigar;
    } else {
        int64_t nb = n > 0 ? rd.seq_off[n] : 0, nc = n > 0 ? rd.cigar_off[n] : 0;
        struct Up {
            const void* src;
            size_t bytes;
        } ups[13] = {{rd.pos0, 4 * (size_t)n},        {rd.flag, 2 * (size_t)n},        {rd.mapq, (size_t)n},
                     {rd.has_nm, (size_t)n},          {rd.nm, 4 * (size_t)n},          {rd.mpos0, 4 * (size_t)n},
                     {rd.tid, 4 * (size_t)n},         {rd.mtid, 4 * (size_t)n},        {rd.seq_off, 8 * (size_t)(n + 1)},
                     {rd.cigar_off, 8 * (size_t)(n + 1)}, {rd.bases, (size_t)nb},      {rd.quals, (size_t)nb},
                     {rd.cigar, 4 * (size_t)nc}};
        void* dst[13];
        for (int i = 0; i < 13; i++) {
            dst[i] = grow(b_reads[i], ups[i].bytes);
            if (!dst[i]) return fail(RV_ERR_OOM, "device allocation for reads failed");
            if (ups[i].bytes) RV_HIP(hipMemcpyAsync(dst[i], ups[i].src, ups[i].bytes, hipMemcpyHostToDevice, stream));
        }
        R.pos0 = (const int32_t*)dst[0];
        R.flag = (const uint16_t*)dst[1];
        R.mapq = (const uint8_t*)dst[2];
        R.has_nm = (const uint8_t*)dst[3];
        R.nm = (const int32_t*)dst[4];
        R.mpos0 = (const int32_t*)dst[5];
        R.tid = (const int32_t*)dst[6];
        R.mtid = (const int32_t*)dst[7];
        R.seq_off = (const int64_t*)dst[8];
        R.cig_off = (const int64_t*)dst[9];
        R.bases = (const uint8_t*)dst[10];
        R.quals = (const uint8_t*)dst[11];
        R.cigar = (const uint32_t*)dst[12];
    }
    // ---- -Z drop mask (uploaded before the window's synchronisation below)
    R.drop = nullptr;
    if (downsampling >= 0) {
        if (drop_queue_pos < drop_queue.size()) {
            drop_host.swap(drop_queue[drop_queue_pos++]);
            if ((int64_t)drop_host.size() != n) return fail(RV_ERR_ARG, "queued -Z mask does not match the reads");
        } else {
            drop_host.assign((size_t)n, 0);
            std::vector<uint16_t> fl;
            const uint16_t* flags = rd.flag;
            if (rd.on_device && n > 0) {
                fl.resize((size_t)n);
                RV_HIP(hipMemcpy(fl.data(), rd.flag, 2 * (size_t)n, hipMemcpyDeviceToHost));
                flags = fl.data();
            }
            downsample_draws(flags, n, cfg.samfilter, downsampling, drop_host.data());
        }
        uint8_t* d = (uint8_t*)grow(b_drop, (size_t)std::max<int64_t>(n, 1));
        if (!d) return fail(RV_ERR_OOM, "device allocation failed");
        if (n) RV_HIP(hipMemcpyAsync(d, drop_host.data(), (size_t)n, hipMemcpyHostToDevice, stream));
        R.drop = d;
    }
    // ---- reference window
    Window W;
    W.region_start = region.start;
    W.region_end = region.end;
    W.chr_len = contig.length;
    window_bounds(cfg, region, contig.length, W.ref_lo, W.ref_hi);
    int win_len = W.ref_hi >= W.ref_lo ? W.ref_hi - W.ref_lo + 1 : 0;
    {
        // the window string of an earlier region, when one was handed back: its pages are
        // already mapped
        std::string up = std::move(ref_spare);
        ref_spare.clear();
        up.resize((size_t)std::max(win_len, 1), 'N');
        const unsigned char* src = (const unsigned char*)contig.bases + (W.ref_lo - 1);
        char* dst = &up[0];
        for (int i = 0; i < win_len; i++) {  // branch-free upper case (vectorised)
            const unsigned char ch = src[i];
            dst[i] = (char)(ch - (unsigned char)(((unsigned)(ch - 'a') < 26u) << 5));
        }
        ref_dev = (char*)grow(b_ref, up.size());
        if (!ref_dev) return fail(RV_ERR_OOM, "device allocation for the reference window failed");
        RV_HIP(hipMemcpyAsync(ref_dev, up.data(), up.size(), hipMemcpyHostToDevice, stream));
        RV_HIP(hipStreamSynchronize(stream));  // `up` is pageable: the copy must finish before it moves
        up.resize((size_t)win_len);
        out.ref_win = std::move(up);
    }
    auto since_h0 = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count(); };
    stats.timeline_ms[8] = since_h0();
    W.ref = ref_dev;
    W.dense_lo = 0;
    W.dense_n = 0;
    // ---- prep
    PrepRec* prep = (PrepRec*)grow(b_prep, sizeof(PrepRec) * (size_t)std::max<int64_t>(n, 1));
    uint32_t* cig = (uint32_t*)grow(b_cig, sizeof(uint32_t) * MAXC * (size_t)std::max<int64_t>(n, 1));
    WalkCounters* ctr = (WalkCounters*)grow(b_ctr, sizeof(WalkCounters));
    int* span = (int*)grow(b_span, 4 * sizeof(int));
    if (!prep || !cig || !ctr || !span) return fail(RV_ERR_OOM, "device allocation failed");
    RV_HIP(hipMemsetAsync(ctr, 0, sizeof(WalkCounters), stream));
    RV_HIP(hipMemsetAsync(span, 0, 4 * sizeof(int), stream));
    WalkArgs* dargs = (WalkArgs*)grow(b_args, sizeof(WalkArgs));
    if (!dargs) return fail(RV_ERR_OOM, "device allocation failed");
    // kernel arguments live in device memory (synchronous copies: the stream is idle here)
    WalkArgs hargs{};
    hargs.R = R;
    hargs.W = W;
    hargs.cfg = cfg;
    hargs.dup = nullptr;
    if (cfg.removeDuplicatedReads) {
        hargs.dup = (uint8_t*)grow(b_dup, (size_t)std::max<int64_t>(n, 1));
        if (!hargs.dup) return fail(RV_ERR_OOM, "device allocation failed");
    }
    RV_HIP(hipStreamSynchronize(stream));
    RV_HIP(hipMemcpy(dargs, &hargs, sizeof hargs, hipMemcpyHostToDevice));
    args_dev = dargs;
    args_n = n;
    RV_HIP(hipEventRecord(ev[0], stream));
    RV_HIP(launch_prep(dargs, n, prep, cig, ctr, span, cfg.removeDuplicatedReads != 0, stream));
    RV_HIP(hipEventRecord(ev[1], stream));
    int max_span = 0;
    RV_HIP(hipMemcpyAsync(&max_span, span, sizeof(int), hipMemcpyDeviceToHost, stream));
    RV_HIP(hipStreamSynchronize(stream));
    stats.timeline_ms[9] = since_h0();
    // ---- dense window: every position a region read can touch
    int pad = 2 * std::max(max_span, 64) + 8;
    W.dense_lo = region.start - pad;
    W.dense_n = (region.end + pad) - W.dense_lo + 1;
    rows = (DenseRow*)grow(b_rows, sizeof(DenseRow) * (size_t)W.dense_n);
    if (!rows) return fail(RV_ERR_OOM, "device allocation for the pileup rows failed");
    rows_n = W.dense_n;
    rows_lo = W.dense_lo;
    RV_HIP(launch_fill_rows(rows, W.dense_n, stream));
    // ---- walk with bounded record buffers (grown and re-run on overflow)
    unsigned max_events = (unsigned)std::min<int64_t>(4 * n + 4096, 0x7fffffff);
    unsigned long long arena_cap = 64ull * (unsigned long long)n + 65536;
    unsigned max_clips = (unsigned)std::min<int64_t>(n + 1024, 0x7fffffff);
    unsigned long long clip_cap = 320ull * (unsigned long long)max_clips;
    unsigned max_anchors = (unsigned)std::min<int64_t>(n + 1024, 0x7fffffff);
    WalkCounters hc{};
    unsigned counts[4];
    for (int attempt = 0; attempt < 4; attempt++) {
        WalkSink S;
        S.dense = rows;
        S.dense_lo = W.dense_lo;
        S.dense_n = W.dense_n;
        S.events = (Event*)grow(b_events, sizeof(Event) * (size_t)max_events);
        S.arena = (char*)grow(b_arena, (size_t)arena_cap);
        S.clips = (ClipRec*)grow(b_clips, sizeof(ClipRec) * (size_t)max_clips);
        S.clip_arena = (char*)grow(b_clip_arena, (size_t)clip_cap);
        S.anchors = (Anchor*)grow(b_anchors, sizeof(Anchor) * (size_t)max_anchors);
        unsigned long long* cnt64 = (unsigned long long*)grow(b_counts, 64);
        if (!S.events || !S.arena || !S.clips || !S.clip_arena || !S.anchors || !cnt64)
            return fail(RV_ERR_OOM, "device allocation for walk records failed");
        RV_HIP(hipMemsetAsync(cnt64, 0, 64, stream));
        S.n_events = (unsigned*)cnt64;
        S.n_clips = (unsigned*)cnt64 + 1;
        S.n_anchors = (unsigned*)cnt64 + 2;
        S.arena_used = cnt64 + 2;
        S.clip_used = cnt64 + 3;
        S.max_events = max_events;
        S.arena_cap = arena_cap;
        S.max_clips = max_clips;
        S.clip_cap = clip_cap;
        S.max_anchors = max_anchors;
        S.ctr = ctr;
        hargs.W = W;
        hargs.S = S;
        RV_HIP(hipStreamSynchronize(stream));
        RV_HIP(hipMemcpy(dargs, &hargs, sizeof hargs, hipMemcpyHostToDevice));
        if (attempt > 0) {
            RV_HIP(launch_fill_rows(rows, W.dense_n, stream));
            RV_HIP(hipMemsetAsync(ctr, 0, sizeof(WalkCounters), stream));
            RV_HIP(launch_prep(dargs, n, prep, cig, ctr, span, false, stream));  // duplicate marks kept
        }
        RV_HIP(hipEventRecord(ev[2], stream));
        RV_HIP(launch_walk(dargs, n, prep, cig, walk_reads_for(n, W.dense_n, max_span), stream));
        RV_HIP(hipEventRecord(ev[3], stream));
        RV_HIP(launch_anchor(S.anchors, S.n_anchors, max_anchors, rows, W.dense_lo, W.dense_n, cfg, stream));
        RV_HIP(hipEventRecord(ev[4], stream));
        unsigned long long hcnt[4];
        RV_HIP(hipMemcpyAsync(hcnt, cnt64, 32, hipMemcpyDeviceToHost, stream));
        RV_HIP(hipMemcpyAsync(&hc, ctr, sizeof hc, hipMemcpyDeviceToHost, stream));
        RV_HIP(hipStreamSynchronize(stream));
        stats.timeline_ms[10] = since_h0();
        counts[0] = (unsigned)(hcnt[0] & 0xffffffffu);
        counts[1] = (unsigned)(hcnt[0] >> 32);
        counts[2] = (unsigned)(hcnt[1] & 0xffffffffu);
        unsigned long long aused = hcnt[2], cused = hcnt[3];
        bool over = counts[0] > max_events || aused > arena_cap || counts[1] > max_clips || cused > clip_cap ||
                    counts[2] > max_anchors;
        if (hc.capacity)
            return fail(RV_ERR_CAPACITY, std::to_string(hc.capacity) + " read(s) have a CIGAR longer than " +
                                             std::to_string(MAXC_IN) + " operations (or a rewrite beyond " +
                                             std::to_string(CIGTXT) + " characters); region not processed");
        if (!over && hc.overflow == 0) {
            if (hooks && hooks->after_walk) hooks->after_walk();
            // ---- copy the ordered records back and reduce them
            // into this region's pinned host memory: copies into pageable memory are staged and
            // synchronous
            Event* evs = (Event*)pinned_alloc(sizeof(Event) * std::max(counts[0], 1u));
            char* arena = (char*)pinned_alloc((size_t)aused + 1);
            ClipRec* cls = (ClipRec*)pinned_alloc(sizeof(ClipRec) * std::max(counts[1], 1u));
            char* carena = (char*)pinned_alloc((size_t)cused + 1);
            if (!evs || !arena || !cls || !carena) return fail(RV_ERR_OOM, "pinned host memory for the walk's records failed");
            if (counts[0]) RV_HIP(hipMemcpyAsync(evs, S.events, sizeof(Event) * counts[0], hipMemcpyDeviceToHost, stream));
            if (aused) RV_HIP(hipMemcpyAsync(arena, S.arena, (size_t)aused, hipMemcpyDeviceToHost, stream));
            if (counts[1]) RV_HIP(hipMemcpyAsync(cls, S.clips, sizeof(ClipRec) * counts[1], hipMemcpyDeviceToHost, stream));
            if (cused) RV_HIP(hipMemcpyAsync(carena, S.clip_arena, (size_t)cused, hipMemcpyDeviceToHost, stream));
            if (download_rows) {
                out.P.rows.resize((size_t)W.dense_n);
                RV_HIP(hipMemcpyAsync(out.P.rows.data(), rows, sizeof(DenseRow) * (size_t)W.dense_n, hipMemcpyDeviceToHost,
                                      stream));
            }
            RV_HIP(hipStreamSynchronize(stream));
            stats.timeline_ms[11] = since_h0();
            // the previous region's host state is released while this one waits on the device; wait
            // for that before the reductions allocate
            wait_release();
            stats.timeline_ms[12] = since_h0();
            auto r0 = std::chrono::steady_clock::now();
            out.P.dense_lo = W.dense_lo;
            out.P.dense_n = W.dense_n;
            out.P.maxReadLength = hc.max_read_length;
            out.W = W;
            // the three reductions write disjoint maps: events -> nd / ins / counts, clips -> sc5, sc3
            double clip_ms[2] = {0, 0};
            std::exception_ptr clip_err[2];
            auto clips = [&](int k) {
                try {
                    const auto c0 = std::chrono::steady_clock::now();
                    reduce_clips(cls, counts[1], carena, cfg.goodq, out.P, k == 0 ? 5 : 3);
                    clip_ms[k] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
                } catch (...) {
                    clip_err[k] = std::current_exception();
                }
            };
            std::thread c5(clips, 0), c3(clips, 1);
            try {
                reduce_events(evs, counts[0], arena, cfg.goodq, out.P);
            } catch (...) {
                c5.join();
                c3.join();
                throw;
            }
            const auto r1 = std::chrono::steady_clock::now();
            if (hooks && hooks->after_events) {
                try {
                    hooks->after_events(out);
                } catch (...) {
                    c5.join();
                    c3.join();
                    throw;
                }
            }
            c5.join();
            c3.join();
            for (auto& ce : clip_err)
                if (ce) std::rethrow_exception(ce);
            stats.timeline_ms[13] = since_h0();
            stats.detail_ms[10] = std::chrono::duration<double, std::milli>(r1 - r0).count();
            stats.detail_ms[11] = std::max(clip_ms[0], clip_ms[1]);
            host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - r0).count();
            out.W = W;
            out.ctr = hc;
            float t01 = 0, t23 = 0, t34 = 0;  // stage timings only: a failed query reads as 0
            (void)hipEventElapsedTime(&t01, ev[0], ev[1]);
            (void)hipEventElapsedTime(&t23, ev[2], ev[3]);
            (void)hipEventElapsedTime(&t34, ev[3], ev[4]);
            stats.prep_ms = t01;
            stats.walk_ms = t23;
            stats.anchor_ms = t34;
            stats.reads_in = n;
            stats.events = counts[0];
            stats.clips = counts[1];
            stats.ref_key_creations = (int64_t)hc.ref_insert;
            stats.walk_errors = (int64_t)hc.walk_errors;
            stats.dense_rows = W.dense_n;
            stats.event_bytes = (int64_t)(sizeof(Event) * counts[0] + aused + sizeof(ClipRec) * counts[1] + cused);
            stats.cx_positions = 0;
            stats.rows_fetched = 0;
            stats.host_ms = host_ms;
            stats.reduce_ms = host_ms;
            (void)h0;
            return RV_OK;
        }
        max_events = std::max<unsigned>(max_events, counts[0]) * 2;
        arena_cap = std::max(arena_cap, aused) * 2;
        max_clips = std::max<unsigned>(max_clips, counts[1]) * 2;
        clip_cap = std::max(clip_cap, cused) * 2;
        max_anchors = std::max<unsigned>(max_anchors, counts[2]) * 2;
    }
    return fail(RV_ERR_CAPACITY, "walk record buffers kept overflowing");
}

const double* Engine::lgamma_table() {
    if (b_lgamma.p && b_lgamma.cap >= sizeof(double) * FISHER_LGT_N) return (const double*)b_lgamma.p;
    std::vector<double> t(FISHER_LGT_N);
    fisher_lgamma_table(t.data(), FISHER_LGT_N);
    double* d = (double*)grow(b_lgamma, sizeof(double) * FISHER_LGT_N);
    if (!d) {
        fail(RV_ERR_OOM, "device buffer for the log-gamma table failed");
        return nullptr;
    }
    if (hipMemcpy(d, t.data(), sizeof(double) * FISHER_LGT_N, hipMemcpyHostToDevice) != hipSuccess) {
        fail(RV_ERR_DEVICE, "upload of the log-gamma table failed");
        return nullptr;
    }
    return d;
}

bool Engine::find35(const std::vector<F35Job>& jobs, std::vector<F35>& out) {
    static_assert(sizeof(F35Result) == sizeof(F35) && sizeof(F35) == 12, "F35 layout");
    const size_t n = jobs.size();
    if (n == 0 || n > (size_t)INT32_MAX / 4) return false;
    std::vector<int32_t> pairs(4 * n);
    size_t bytes = 0;
    for (const F35Job& j : jobs) bytes += j.s5->size() + j.s3->size();
    if (bytes > (size_t)INT32_MAX) return false;
    std::string arena;
    arena.reserve(bytes);
    for (size_t k = 0; k < n; k++) {
        pairs[4 * k] = (int32_t)arena.size();
        pairs[4 * k + 1] = (int32_t)jobs[k].s5->size();
        arena += *jobs[k].s5;
        pairs[4 * k + 2] = (int32_t)arena.size();
        pairs[4 * k + 3] = (int32_t)jobs[k].s3->size();
        arena += *jobs[k].s3;
    }
    int32_t* dp = (int32_t*)grow(b_f35_pairs, sizeof(int32_t) * 4 * n);
    char* da = (char*)grow(b_f35_arena, std::max<size_t>(arena.size(), 1));
    F35Result* dout = (F35Result*)grow(b_f35_out, sizeof(F35Result) * n);
    if (!dp || !da || !dout) return false;
    out.resize(n);
    if (hipMemcpyAsync(dp, pairs.data(), sizeof(int32_t) * 4 * n, hipMemcpyHostToDevice, stream) != hipSuccess) return false;
    if (!arena.empty() && hipMemcpyAsync(da, arena.data(), arena.size(), hipMemcpyHostToDevice, stream) != hipSuccess)
        return false;
    if (hipEventRecord(sev[0], stream) != hipSuccess) return false;
    if (launch_find35(dp, (int)n, da, dout, stream) != hipSuccess) return false;
    if (hipEventRecord(sev[1], stream) != hipSuccess) return false;
    if (hipMemcpyAsync(out.data(), dout, sizeof(F35Result) * n, hipMemcpyDeviceToHost, stream) != hipSuccess) return false;
    if (hipStreamSynchronize(stream) != hipSuccess) return false;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, sev[0], sev[1]);
    f35_ms += ms;
    f35_device_pairs += (int64_t)n;
    return true;
}

rv_status Engine::fisher_tables(const int32_t* tab, int n, double* out) {
    const double* lg = lgamma_table();
    if (!lg) return RV_ERR_DEVICE;
    int* din = (int*)grow(b_fisher_in, sizeof(int) * 4 * (size_t)n);
    double* dout = (double*)grow(b_fisher_out, sizeof(double) * 3 * (size_t)n);
    if (!din || !dout) return fail(RV_ERR_OOM, "device buffers for the Fisher tables failed");
    RV_HIP(hipMemcpyAsync(din, tab, sizeof(int) * 4 * (size_t)n, hipMemcpyHostToDevice, stream));
    RV_HIP(launch_fisher_table(din, n, dout, lg, FISHER_LGT_N, stream));
    RV_HIP(hipMemcpyAsync(out, dout, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipStreamSynchronize(stream));
    return RV_OK;
}

rv_status Engine::key_order(int n_reads, int64_t& n_keys, int& max_disp) {
    n_keys = 0;
    max_disp = INT_MAX;
    key_n_ = 0;
    key_slot_ok_ = false;
    if (rows_n <= 0) {
        max_disp = 0;
        key_slot_ok_ = true;
        return RV_OK;
    }
    auto* keys = (unsigned long long*)grow(b_keys, sizeof(unsigned long long) * (size_t)rows_n);
    auto* keys2 = (unsigned long long*)grow(b_keys2, sizeof(unsigned long long) * (size_t)rows_n);
    auto* pos = (int32_t*)grow(b_kpos, sizeof(int32_t) * (size_t)rows_n);
    auto* pos2 = (int32_t*)grow(b_kpos2, sizeof(int32_t) * (size_t)rows_n);
    auto* cnt = (unsigned*)grow(b_kcount, 2 * sizeof(unsigned));
    if (!keys || !keys2 || !pos || !pos2 || !cnt) return fail(RV_ERR_OOM, "device buffers for the key order failed");
    RV_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned), stream));
    RV_HIP(launch_key_list(rows, rows_n, rows_lo, keys, pos, cnt, stream));
    unsigned n = 0;
    RV_HIP(hipMemcpyAsync(&n, cnt, sizeof n, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipStreamSynchronize(stream));
    key_n_ = n;
    n_keys = n;
    if (n == 0) {
        max_disp = 0;
        key_slot_ok_ = true;
        return RV_OK;
    }
    int end_bit = 25;  // touch index (24 bits) + read ordinal
    while (end_bit < 64 && (1ll << (end_bit - 24)) <= (long long)n_reads) end_bit++;
    size_t tb = 0;
    RV_HIP(sort_key_pairs(nullptr, tb, keys, keys2, pos, pos2, (int)n, end_bit, stream));
    void* tmp = grow(b_sort_tmp, tb);
    if (!tmp) return fail(RV_ERR_OOM, "device buffer for the key sort failed");
    RV_HIP(sort_key_pairs(tmp, tb, keys, keys2, pos, pos2, (int)n, end_bit, stream));
    // the same keys in sorted slot order (pos2 holds them in creation order), and the largest
    // displacement of that layout
    auto* pos3 = (int32_t*)grow(b_kpos3, sizeof(int32_t) * (size_t)rows_n);
    if (!pos3) return fail(RV_ERR_OOM, "device buffers for the key order failed");
    const int cap_log2 = rh_log2(rh_capacity(n));
    int idx_bits = 1;
    while (idx_bits < 40 && (1ll << idx_bits) <= (long long)n) idx_bits++;
    const int slot_bits = cap_log2 + 5 + idx_bits;
    if (slot_bits > 64) {
        RV_HIP(hipStreamSynchronize(stream));
        return RV_OK;
    }
    RV_HIP(launch_slot_keys(pos2, (int)n, cap_log2, idx_bits, keys, pos, stream));
    size_t tb2 = 0, tb3 = 0;
    RV_HIP(sort_key_pairs(nullptr, tb2, keys, keys2, pos, pos3, (int)n, slot_bits, stream));
    RV_HIP(launch_max_displacement(pos3, (int)n, cap_log2, nullptr, nullptr, nullptr, nullptr, tb3, stream));
    void* tmp2 = grow(b_sort_tmp, std::max(tb, std::max(tb2, tb3)));
    if (!tmp2) return fail(RV_ERR_OOM, "device buffer for the key sort failed");
    RV_HIP(sort_key_pairs(tmp2, tb2, keys, keys2, pos, pos3, (int)n, slot_bits, stream));
    // keys / keys2 are free again: two int32 arrays of n each
    int32_t* v = (int32_t*)keys;
    int32_t* pm = (int32_t*)keys2;
    int* md = (int*)(cnt + 1);
    RV_HIP(launch_max_displacement(pos3, (int)n, cap_log2, v, pm, md, tmp2, tb3, stream));
    int h = 0;
    RV_HIP(hipMemcpyAsync(&h, md, sizeof h, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipStreamSynchronize(stream));
    max_disp = h;
    key_slot_ok_ = true;
    return RV_OK;
}

rv_status Engine::download_keys(bool slot_order, std::vector<int32_t>& out) {
    out.clear();
    if (slot_order && !key_slot_ok_) return fail(RV_ERR_DEVICE, "no slot-ordered keys on the device");
    if (key_n_ == 0) return RV_OK;
    out.resize(key_n_);
    RV_HIP(hipMemcpyAsync(out.data(), slot_order ? b_kpos3.p : b_kpos2.p, sizeof(int32_t) * key_n_, hipMemcpyDeviceToHost,
                          stream));
    RV_HIP(hipStreamSynchronize(stream));
    return RV_OK;
}

rv_status Engine::slot_order(const int32_t* ops, const uint8_t* tr, size_t m, std::vector<int32_t>& out, bool& applied) {
    applied = false;
    out.clear();
    if (m >= ((size_t)1 << 30)) return RV_OK;
    auto* ops_d = (int32_t*)grow(b_so_ops, 4 * m);
    auto* tr_d = (uint8_t*)grow(b_so_tr, m);
    if (!ops_d || !tr_d) return fail(RV_ERR_OOM, "device buffers for the slot order failed");
    if (m) {
        RV_HIP(hipMemcpyAsync(ops_d, ops, 4 * m, hipMemcpyHostToDevice, stream));
        RV_HIP(hipMemcpyAsync(tr_d, tr, m, hipMemcpyHostToDevice, stream));
    }
    return slot_order_dev(ops_d, tr_d, tr, m, out, applied);
}

rv_status Engine::aligned_slot_order(const int32_t* proc_dev, size_t n_proc, uint8_t* status_dev, int dense_lo,
                                     int dense_n, const int32_t* cx_pos_dev, int ncx, std::vector<int32_t>& out,
                                     bool& applied, std::vector<int32_t>& ops, std::vector<uint8_t>& tr, bool download) {
    applied = false;
    out.clear();
    ops.clear();
    tr.clear();
    if (n_proc >= ((size_t)1 << 30)) return fail(RV_ERR_CAPACITY, "processing order too long for the slot order");
    const int nb = (int)((n_proc + SO_CB - 1) / SO_CB);
    auto* ops_d = (int32_t*)grow(b_so_ops, 4 * std::max<size_t>(n_proc, 1));
    auto* tr_d = (uint8_t*)grow(b_so_tr, std::max<size_t>(n_proc, 1));
    auto* bc = (int32_t*)grow(b_so_w[0], 4 * ((size_t)nb + 2));
    auto* bo = (int32_t*)grow(b_so_w[1], 4 * ((size_t)nb + 2));
    size_t tb = 0;
    RV_HIP(so_ops_scan_bytes(nb, tb));
    void* tmp = grow(b_so_tmp, tb);
    if (!ops_d || !tr_d || !bc || !bo || !tmp) return fail(RV_ERR_OOM, "device buffers for the slot order failed");
    RV_HIP(launch_so_ops(proc_dev, (int)n_proc, status_dev, dense_lo, dense_n, cx_pos_dev, ncx, bc, bo, tmp, tb, ops_d,
                         tr_d, stream));
    int32_t m = 0;
    RV_HIP(hipMemcpyAsync(&m, bo + nb, sizeof m, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipStreamSynchronize(stream));
    tr.resize((size_t)m);
    if (m) {
        RV_HIP(hipMemcpyAsync(tr.data(), tr_d, (size_t)m, hipMemcpyDeviceToHost, stream));
        RV_HIP(hipStreamSynchronize(stream));
    }
    const rv_status st = slot_order_dev(ops_d, tr_d, tr.data(), (size_t)m, out, applied, download);
    if (st != RV_OK || applied) return st;
    // declined: the caller orders on the host
    ops.resize((size_t)m);
    if (m) {
        RV_HIP(hipMemcpyAsync(ops.data(), ops_d, 4 * (size_t)m, hipMemcpyDeviceToHost, stream));
        RV_HIP(hipStreamSynchronize(stream));
    }
    return RV_OK;
}

rv_status Engine::download_slot_order(std::vector<int32_t>& out) {
    out.clear();
    if (!so_valid_) return fail(RV_ERR_DEVICE, "no slot order on the device");
    out.resize(so_n_);
    if (so_n_) {
        RV_HIP(hipMemcpyAsync(out.data(), so_final_, 4 * so_n_, hipMemcpyDeviceToHost, stream));
        RV_HIP(hipStreamSynchronize(stream));
    }
    return RV_OK;
}

rv_status Engine::record_order(const OutRec* recs_dev, int n, int dense_lo, int dense_n, std::vector<uint32_t>& perm,
                               bool& ok) {
    ok = false;
    perm.clear();
    if (!so_valid_ || n < 0) return RV_OK;
    if (n == 0) {
        ok = true;
        return RV_OK;
    }
    auto* rank_of = (int32_t*)grow(b_so_rank, sizeof(int32_t) * (size_t)std::max(dense_n, 1));
    auto* key = (unsigned long long*)grow(b_so_rk[0], 8 * (size_t)n);
    auto* key2 = (unsigned long long*)grow(b_so_rk[1], 8 * (size_t)n);
    auto* idx = (int32_t*)grow(b_so_rk[2], 4 * (size_t)n);
    auto* idx2 = (int32_t*)grow(b_so_rk[3], 4 * (size_t)n);
    auto* bad = (unsigned*)grow(b_so_small, 64) + 15;
    if (!rank_of || !key || !key2 || !idx || !idx2) return fail(RV_ERR_OOM, "device buffers for the record order failed");
    size_t tb = 0;
    RV_HIP(sort_key_pairs(nullptr, tb, key, key2, idx, idx2, n, 64, stream));
    void* tmp = grow(b_so_tmp, tb);
    if (!tmp) return fail(RV_ERR_OOM, "device buffer for the record order failed");
    RV_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned), stream));
    const long long pos_base = (long long)dense_lo - (1ll << 28);
    RV_HIP(launch_so_rank(so_final_, (int)so_n_, recs_dev, n, dense_lo, dense_n, pos_base, rank_of, key, idx, bad,
                          stream));
    RV_HIP(sort_key_pairs(tmp, tb, key, key2, idx, idx2, n, 64, stream));
    perm.resize((size_t)n);
    unsigned hbad = 0;
    RV_HIP(hipMemcpyAsync(perm.data(), idx2, 4 * (size_t)n, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipMemcpyAsync(&hbad, bad, sizeof hbad, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipStreamSynchronize(stream));
    if (hbad) {
        perm.clear();
        return RV_OK;
    }
    ok = true;
    return RV_OK;
}

rv_status Engine::slot_order_dev(const int32_t* ops_d, const uint8_t* tr_d, const uint8_t* tr, size_t m,
                                 std::vector<int32_t>& out, bool& applied, bool download) {
    applied = false;
    out.clear();
    slot_order_ms = 0;
    so_valid_ = false;
    if (m == 0) {
        applied = true;
        so_final_ = nullptr;
        so_n_ = 0;
        so_valid_ = true;
        return RV_OK;
    }
    // growth schedule: an insertion that finds 80% of the slots used doubles the table first
    struct Ep {
        size_t cap, a, b, np, S, keep;
    };
    std::vector<Ep> eps;
    size_t cap = 8, count = 0, a = 0, max_keep = 1, max_S = 1;
    for (size_t i = 0; i < m; i++) {
        if (count >= cap * 80 / 100) {
            eps.push_back({cap, a, i, 0, 0, count});
            a = i;
            cap *= 2;
        }
        count += tr[i] ? 0 : 1;
    }
    eps.push_back({cap, a, m, 0, 0, count});
    if (rh_log2(cap) > 30) return RV_OK;
    for (size_t e = 0; e < eps.size(); e++) {
        eps[e].np = e ? eps[e - 1].keep : 0;
        eps[e].S = eps[e].np + (eps[e].b - eps[e].a);
        max_S = std::max(max_S, eps[e].S);
        max_keep = std::max(max_keep, eps[e].keep);
    }
    int32_t* ord[2] = {(int32_t*)grow(b_so_ord[0], 4 * max_keep), (int32_t*)grow(b_so_ord[1], 4 * max_keep)};
    uint32_t* k[4];
    for (int i = 0; i < 4; i++) k[i] = (uint32_t*)grow(b_so_k[i], 4 * max_S);
    int32_t* w[4];
    for (int i = 0; i < 4; i++) w[i] = (int32_t*)grow(b_so_w[i], 4 * max_keep);
    auto* runs = (int32_t*)grow(b_so_runs, 7 * 4 * max_keep);
    auto* small = (unsigned*)grow(b_so_small, 64);
    if (!ord[0] || !ord[1] || !k[0] || !k[1] || !k[2] || !k[3] || !w[0] || !w[1] || !w[2] || !w[3] || !runs || !small)
        return fail(RV_ERR_OOM, "device buffers for the slot order failed");
    size_t sort_bytes = 0, scan_bytes = 0;
    RV_HIP(so_sort_pairs(nullptr, sort_bytes, k[0], k[1], k[2], k[3], (int)max_S, 31, stream));
    RV_HIP(so_scan_bytes((int)max_keep, scan_bytes));
    const size_t tmp_bytes = std::max(sort_bytes, scan_bytes);
    void* tmp = grow(b_so_tmp, tmp_bytes);
    if (!tmp) return fail(RV_ERR_OOM, "device buffer for the slot order's sorts failed");
    SoRuns R{runs, runs + max_keep, runs + 2 * max_keep, runs + 3 * max_keep, runs + 4 * max_keep,
             runs + 5 * max_keep, runs + 6 * max_keep, (int*)(small + 7)};
    RV_HIP(hipEventRecord(so_ev[0], stream));
    // small: [0, 5) first steps, [5] a run too long to simulate, [6] a home group too large, [7] runs,
    // [8] an early growth, [9, 14) the halving steps; the flags stay set across epochs and are read
    // once at the end (no host synchronisation between epochs)
    RV_HIP(hipMemsetAsync(small, 0, 64, stream));
    SoHalve* H = (SoHalve*)(small + 9);
    int cur = 0;
    for (const Ep& ep : eps) {
        const int lcap = rh_log2(ep.cap);
        SoEpoch E{ord[cur], (int)ep.np, ops_d + ep.a, tr_d + ep.a, (int)ep.S, (int)ep.keep, (uint32_t)(ep.cap - 1), lcap};
        RV_HIP(hipMemsetAsync(small, 0xff, 5 * sizeof(unsigned), stream));
        RV_HIP(hipMemsetAsync(small + 7, 0, sizeof(unsigned), stream));
        RV_HIP(launch_so_epoch_keys(E, k[0], k[1], stream));
        size_t tb = tmp_bytes;
        RV_HIP(so_sort_pairs(tmp, tb, k[0], k[2], k[1], k[3], (int)ep.S, lcap + 1, stream));
        RV_HIP(launch_so_runs(k[2], (int)ep.keep, (int)ep.S, (uint32_t)(ep.cap - 1), w[0], w[1], w[2], w[3], R, tmp,
                              tmp_bytes, stream));
        RV_HIP(launch_so_simulate(k[2], k[3], (uint32_t)(ep.cap - 1), R, (int)ep.keep, small, small + 5, stream));
        RV_HIP(launch_so_final(k[2], k[3], E, small, H, ord[cur ^ 1], small + 6, small + 8, stream));
        cur ^= 1;
    }
    const size_t n_out = eps.back().keep;
    unsigned flags[4] = {0, 0, 0, 0};
    if (download) {
        out.resize(n_out);
        RV_HIP(hipMemcpyAsync(out.data(), ord[cur], 4 * n_out, hipMemcpyDeviceToHost, stream));
    }
    RV_HIP(hipMemcpyAsync(flags, small + 5, sizeof flags, hipMemcpyDeviceToHost, stream));
    RV_HIP(hipEventRecord(so_ev[1], stream));
    RV_HIP(hipStreamSynchronize(stream));
    RV_HIP(hipEventElapsedTime(&slot_order_ms, so_ev[0], so_ev[1]));
    // a run too long to simulate, a home group too large, or an early growth: the host orders
    if (flags[0] || flags[1] || flags[3]) {
        out.clear();
        return RV_OK;
    }
    applied = true;
    so_final_ = ord[cur];
    so_n_ = n_out;
    so_valid_ = true;
    return RV_OK;
}

bool Engine::no_passing(int start, int end, bool& none) {
    none = true;
    if (!args_dev || args_n <= 0) return true;
    unsigned* cnt = (unsigned*)grow(b_nopass, 64);
    if (!cnt) return false;
    unsigned h = 0;
    if (hipMemsetAsync(cnt, 0, sizeof(unsigned), stream) != hipSuccess) return false;
    if (launch_no_passing(args_dev, args_n, start, end, cnt, stream) != hipSuccess) return false;
    if (hipMemcpyAsync(&h, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, stream) != hipSuccess) return false;
    if (hipStreamSynchronize(stream) != hipSuccess) return false;
    none = h == 0;
    return true;
}

void Engine::pinned_reset() {
    // blocks the last region used stay (pinning gigabytes again costs far more than keeping them);
    // unused blocks stay while the total is within PINNED_KEEP
    size_t keep = 0, total = 0;
    while (keep < pinned_.size() && (keep <= pinned_cur_ || total + pinned_[keep].second <= PINNED_KEEP))
        total += pinned_[keep++].second;
    for (size_t i = keep; i < pinned_.size(); i++) (void)hipHostFree(pinned_[i].first);
    pinned_.resize(keep);
    pinned_cur_ = 0;
    pinned_off_ = 0;
}

void* Engine::pinned_alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    while (pinned_cur_ < pinned_.size()) {
        auto& blk = pinned_[pinned_cur_];
        if (pinned_off_ + bytes <= blk.second) {
            void* p = (char*)blk.first + pinned_off_;
            pinned_off_ += bytes;
            return p;
        }
        pinned_cur_++;
        pinned_off_ = 0;
    }
    const size_t sz = std::max(bytes, (size_t)64 << 20);
    void* p = nullptr;
    if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) return nullptr;
    pinned_.push_back({p, sz});
    pinned_cur_ = pinned_.size() - 1;
    pinned_off_ = bytes;
    return p;
}

const DenseRow* Engine::fetch_rows(const std::vector<int>& idx) {
    if (idx.empty()) return (const DenseRow*)pinned_alloc(sizeof(DenseRow));
    for (int i : idx)
        if (i < 0 || i >= rows_n) return nullptr;
    int32_t* ids = (int32_t*)grow(b_gather_ids, sizeof(int32_t) * idx.size());
    DenseRow* dst = (DenseRow*)grow(b_gather_out, sizeof(DenseRow) * idx.size());
    int32_t* h_ids = (int32_t*)pinned_alloc(sizeof(int32_t) * idx.size());
    DenseRow* out = (DenseRow*)pinned_alloc(sizeof(DenseRow) * idx.size());
    if (!ids || !dst || !h_ids || !out) return nullptr;
    std::memcpy(h_ids, idx.data(), sizeof(int32_t) * idx.size());
    if (hipMemcpyAsync(ids, h_ids, sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice, stream) != hipSuccess)
        return nullptr;
    if (launch_gather_rows(rows, rows_n, ids, (int)idx.size(), dst, stream) != hipSuccess) return nullptr;
    if (hipMemcpyAsync(out, dst, sizeof(DenseRow) * idx.size(), hipMemcpyDeviceToHost, stream) != hipSuccess)
        return nullptr;
    if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
    return out;
}

}  // namespace rv
