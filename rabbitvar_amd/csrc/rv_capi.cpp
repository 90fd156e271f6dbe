// extern "C" entry points (declared in include/rabbitvar_amd.h).
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>

#include "rabbitvar_amd.h"
#include "rv_engine.h"
#include "rv_region.h"

struct rv_engine {
    rv::Engine* e;
};

namespace rv {
Engine* engine_of(rv_engine* e) { return e ? e->e : nullptr; }
}  // namespace rv

struct rv_callset {
    rv::CallSet cs;
    // the engine's release threads: freeing a call set (tens of thousands of strings, megabytes
    // of text) hands its contents to them and returns at once
    std::shared_ptr<rv::ReleasePool> pool;
};

namespace {

// No C++ exception crosses the C boundary: allocation failures become RV_ERR_OOM, anything else
// RV_ERR_DEVICE with its message in rv_last_error.
template <class F>
rv_status guarded(rv_engine* e, F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        e->e->err = "host allocation failed";
        return RV_ERR_OOM;
    } catch (const std::exception& ex) {
        e->e->err = ex.what();
        return RV_ERR_DEVICE;
    } catch (...) {
        e->e->err = "unknown error";
        return RV_ERR_DEVICE;
    }
}

}  // namespace

extern "C" {

void rv_config_default(rv_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->goodq = 22.5;
    c->bias = 0.05;
    c->freq = 0.01;
    c->qratio = 1.5;
    c->mapq = 0;
    c->lofreq = 0.05;
    c->vext = 2;
    c->mismatch = 8;
    c->trimBasesAfter = 0;
    c->performLocalRealignment = 1;
    c->indelsize = 50;
    c->minBiasReads = 2;
    c->minr = 2;
    c->moveIndelsTo3 = 0;
    c->samfilter = 0x504;
    c->readPosFilter = 5;
    c->doPileup = 0;
    c->minmatch = 0;
    c->uniqueModeAlignmentEnabled = 0;
    c->uniqueModeSecondInPairEnabled = 0;
    c->fisher = 0;
    c->SVMINLEN = 1000;
    c->referenceExtension = 1200;
    c->numberNucleotideToExtend = 0;
    c->mappingQuality = 0;
    c->removeDuplicatedReads = 0;
    c->crisprCuttingSite = 0;
    c->sample = "";
    c->isDownsampling = 0;
    c->downsampling = 0;
}

rv_status rv_open(const rv_config* cfg, int device, rv_engine** out) {
    if (!cfg || !out) return RV_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count) return RV_ERR_DEVICE;
    rv_engine* h = new (std::nothrow) rv_engine;
    if (!h) return RV_ERR_OOM;
    h->e = new (std::nothrow) rv::Engine(*cfg, device);
    if (!h->e || !h->e->stream || !h->e->err.empty()) {
        delete h->e;
        delete h;
        return RV_ERR_DEVICE;
    }
    *out = h;
    return RV_OK;
}

rv_status rv_run_region(rv_engine* e, const rv_region* region, const rv_reads* reads, const rv_contig* contig,
                        rv_callset** out) {
    if (!e || !region || !reads || !contig || !out) return RV_ERR_ARG;
    *out = nullptr;
    rv_callset* cs = new (std::nothrow) rv_callset;
    if (!cs) return RV_ERR_OOM;
    (void)hipSetDevice(e->e->device);  // the calling thread may be new (engine pools)
    const auto t0 = std::chrono::steady_clock::now();
    rv_status s = guarded(e, [&] { return rv::run_region(*e->e, *region, *reads, *contig, cs->cs); });
    e->e->stats.call_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (s != RV_OK) {
        delete cs;
        return s;
    }
    cs->pool = e->e->release_pool();
    *out = cs;
    return RV_OK;
}

rv_status rv_run_region_somatic(rv_engine* e, const rv_region* region, const rv_reads* tumor,
                                const rv_reads* normal, const rv_contig* contig, rv_callset** out) {
    if (!e || !region || !tumor || !normal || !contig || !out) return RV_ERR_ARG;
    *out = nullptr;
    rv_callset* cs = new (std::nothrow) rv_callset;
    if (!cs) return RV_ERR_OOM;
    (void)hipSetDevice(e->e->device);
    const auto t0 = std::chrono::steady_clock::now();
    rv_status s =
        guarded(e, [&] { return rv::run_region_somatic(*e->e, *region, *tumor, *normal, *contig, cs->cs); });
    e->e->stats.call_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (s != RV_OK) {
        delete cs;
        return s;
    }
    cs->pool = e->e->release_pool();
    *out = cs;
    return RV_OK;
}

size_t rv_callset_count(const rv_callset* cs) { return cs ? cs->cs.calls.size() : 0; }
const rv_call* rv_callset_get(const rv_callset* cs, size_t i) {
    if (!cs || i >= cs->cs.calls.size()) return nullptr;
    return &cs->cs.calls[i];
}
const char* rv_callset_text(const rv_callset* cs, size_t* len) {
    if (!cs) return nullptr;
    if (len) *len = cs->cs.text.size();
    return cs->cs.text.c_str();
}
void rv_callset_free(rv_callset* cs) {
    if (!cs) return;
    if (cs->pool) {
        try {
            cs->pool->post(std::make_unique<rv::Held<rv::CallSet>>(std::move(cs->cs)));
        } catch (...) {
            // no memory for the job: the call set is destroyed here instead
        }
    }
    delete cs;
}

rv_status rv_get_stats(const rv_engine* e, rv_stats* out) {
    if (!e || !out) return RV_ERR_ARG;
    *out = e->e->stats;
    return RV_OK;
}
const char* rv_last_error(const rv_engine* e) { return e ? e->e->err.c_str() : "null engine"; }
void rv_close(rv_engine* e) {
    if (!e) return;
    delete e->e;
    delete e;
}

rv_status rv_debug_pileup(rv_engine* e, const rv_region* region, const rv_reads* reads, const rv_contig* contig,
                          char** text, size_t* len) {
    if (!e || !region || !reads || !contig || !text) return RV_ERR_ARG;
    rv::PileupResult pr;
    rv_status s = guarded(e, [&] {
        e->e->pinned_reset();  // the pileup's host copies live in the engine's pinned memory
        return e->e->pileup(*region, *reads, *contig, true, pr);
    });
    if (s != RV_OK) return s;
    std::string d;
    s = guarded(e, [&] {
        d = rv::dump_pileup(pr.P);
        return RV_OK;
    });
    if (s != RV_OK) return s;
    char* p = (char*)std::malloc(d.size() + 1);
    if (!p) return RV_ERR_OOM;
    std::memcpy(p, d.c_str(), d.size() + 1);
    *text = p;
    if (len) *len = d.size();
    return RV_OK;
}

void rv_free(void* p) { std::free(p); }

rv_status rv_fisher_tables_noalloc(const int32_t* tables, size_t n, double* out) {
    if (n && (!tables || !out)) return RV_ERR_ARG;
    if (n > (size_t)rv::FISHER_NOALLOC_MAX_TABLES) return RV_ERR_CAPACITY;
    if (n == 0) return RV_OK;
    const char* err = nullptr;
    return rv::fisher_tables_noalloc(tables, (int)n, out, &err) == hipSuccess ? RV_OK : RV_ERR_DEVICE;
}

rv_status rv_debug_find35(rv_engine* e, const char* arena, size_t arena_len, const int32_t* pairs, size_t n,
                          int32_t* out) {
    if (!e || (n && (!arena || !pairs || !out))) return RV_ERR_ARG;
    std::vector<std::string> strs(2 * n);
    for (size_t k = 0; k < n; k++)
        for (int h = 0; h < 2; h++) {
            const int64_t off = pairs[4 * k + 2 * h], len = pairs[4 * k + 2 * h + 1];
            if (off < 0 || len < 0 || (size_t)(off + len) > arena_len) return RV_ERR_ARG;
            strs[2 * k + h].assign(arena + off, (size_t)len);
        }
    std::vector<rv::F35Job> jobs(n);
    for (size_t k = 0; k < n; k++) jobs[k] = {&strs[2 * k], &strs[2 * k + 1]};
    std::vector<rv::F35> res;
    if (n && !e->e->find35(jobs, res)) return RV_ERR_DEVICE;
    for (size_t k = 0; k < n; k++) {
        out[3 * k] = res[k].b5;
        out[3 * k + 1] = res[k].b3;
        out[3 * k + 2] = res[k].score;
    }
    return RV_OK;
}

rv_status rv_debug_slot_order(rv_engine* e, const int32_t* keys, const uint8_t* transient, size_t n, int32_t* out,
                              size_t* n_out, int* on_device) {
    if (!e || !n_out || !on_device || (n && (!keys || !transient || !out))) return RV_ERR_ARG;
    *n_out = 0;
    *on_device = 0;
    return guarded(e, [&] {
        std::vector<int32_t> order;
        bool applied = false;
        const rv_status st = e->e->slot_order(keys, transient, n, order, applied);
        if (st != RV_OK || !applied) return st;
        std::copy(order.begin(), order.end(), out);
        *n_out = order.size();
        *on_device = 1;
        return RV_OK;
    });
}

rv_status rv_debug_fisher(rv_engine* e, const int32_t* tables, size_t n, double* out) {
    if (!e || (n && (!tables || !out)) || n > (size_t)0x7fffffff / 4) return RV_ERR_ARG;
    if (n == 0) return RV_OK;
    return guarded(e, [&] { return e->e->fisher_tables(tables, (int)n, out); });
}

}  // extern "C"

// ------------------------------------------------------------------ engine pool (region lists)
#include <atomic>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rv_bam.h"

struct rv_pool {
    std::vector<rv_engine*> engines;
    std::string err;
};

extern "C" {

rv_status rv_pool_open(const rv_config* cfg, int device, int workers, rv_pool** out) {
    if (!cfg || !out || workers <= 0) return RV_ERR_ARG;
    *out = nullptr;
    rv_pool* p = new (std::nothrow) rv_pool;
    if (!p) return RV_ERR_OOM;
    for (int w = 0; w < workers; w++) {
        rv_engine* e = nullptr;
        const rv_status s = rv_open(cfg, device, &e);
        if (s != RV_OK) {
            for (rv_engine* x : p->engines) rv_close(x);
            delete p;
            return s;
        }
        p->engines.push_back(e);
    }
    *out = p;
    return RV_OK;
}

rv_status rv_pool_run(rv_pool* p, const rv_region* regions, size_t n, rv_bam* first, rv_bam* second,
                      const rv_contig* contigs, size_t n_contigs, rv_callset** out) {
    if (!p || (n && (!regions || !first || !contigs || !out))) return RV_ERR_ARG;
    for (size_t i = 0; i < n; i++) out[i] = nullptr;
    std::atomic<size_t> next{0};
    std::atomic<bool> stop{false};
    std::mutex mu;
    rv_status first_err = RV_OK;
    // -Z draws: made in region order (a turn per region), which is the reference's order with one
    // thread; the masks are queued on the engine that runs the region
    const double ds = p->engines.empty() ? -1 : rv::engine_of(p->engines[0])->downsampling;
    size_t draw_turn = 0;
    std::condition_variable draw_cv;
    auto worker = [&](rv_engine* eng) {
        rv::BamQuery q1, q2;
        for (;;) {
            if (stop) return;
            const size_t i = next.fetch_add(1);
            if (i >= n) return;
            const rv_region& rg = regions[i];
            const rv_contig* ct = nullptr;
            for (size_t c = 0; c < n_contigs; c++)
                if (contigs[c].name && rg.chr && std::strcmp(contigs[c].name, rg.chr) == 0) ct = &contigs[c];
            rv_status s = RV_OK;
            std::string msg;
            rv_reads r1{}, r2{};
            if (!ct) {
                s = RV_ERR_ARG;
                msg = std::string("no reference sequence for ") + (rg.chr ? rg.chr : "(null)");
            }
            if (s == RV_OK) s = rv::bam_query_into(first, rg.chr, rg.start, rg.end, q1, &r1, msg);
            if (s == RV_OK && second) s = rv::bam_query_into(second, rg.chr, rg.start, rg.end, q2, &r2, msg);
            rv::Engine* E = rv::engine_of(eng);
            if (ds >= 0) {
                std::unique_lock<std::mutex> lk(mu);
                draw_cv.wait(lk, [&] { return draw_turn == i || stop; });
                if (draw_turn != i) return;
                E->drop_queue.clear();
                E->drop_queue_pos = 0;
                if (s == RV_OK) {
                    const rv_reads* rs[2] = {&r1, second ? &r2 : nullptr};
                    for (const rv_reads* r : rs) {
                        if (!r) continue;
                        E->drop_queue.emplace_back((size_t)r->n_reads);
                        rv::downsample_draws(r->flag, r->n_reads, E->cfg.samfilter, ds, E->drop_queue.back().data());
                    }
                }
                draw_turn++;
                draw_cv.notify_all();
            }
            rv_callset* cs = nullptr;
            if (s == RV_OK) {
                s = second ? rv_run_region_somatic(eng, &rg, &r1, &r2, ct, &cs) : rv_run_region(eng, &rg, &r1, ct, &cs);
                if (s != RV_OK) msg = rv_last_error(eng);
            }
            E->drop_queue.clear();
            E->drop_queue_pos = 0;
            if (s != RV_OK) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_err == RV_OK) {
                    first_err = s;
                    p->err = "region " + std::to_string(i) + ": " + msg;
                }
                stop = true;
                draw_cv.notify_all();
                return;
            }
            out[i] = cs;
        }
    };
    std::vector<std::thread> pool;
    for (size_t w = 1; w < p->engines.size(); w++) pool.emplace_back(worker, p->engines[w]);
    worker(p->engines[0]);
    for (auto& th : pool) th.join();
    if (first_err != RV_OK) {
        for (size_t i = 0; i < n; i++) {
            rv_callset_free(out[i]);
            out[i] = nullptr;
        }
    }
    return first_err;
}

const char* rv_pool_last_error(const rv_pool* p) { return p ? p->err.c_str() : ""; }

rv_status rv_pool_set_screen(rv_pool* p, const float* params, size_t n_params, float threshold) {
    if (!p) return RV_ERR_ARG;
    for (rv_engine* e : p->engines) {
        const rv_status s = rv_set_screen(e, params, n_params, threshold);
        if (s != RV_OK) return s;
    }
    return RV_OK;
}

void rv_pool_close(rv_pool* p) {
    if (!p) return;
    for (rv_engine* e : p->engines) rv_close(e);
    delete p;
}

}  // extern "C"
