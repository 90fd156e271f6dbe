// Slot-order replay and sort of the reference's robin_hood maps (see rv_order.h).
#include "rv_order.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <climits>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_set>

namespace rv {

namespace {

uint64_t hash_int(int32_t k) {
    const uint64_t x = (uint64_t)(int64_t)k;
    const unsigned __int128 p = (unsigned __int128)x * (unsigned __int128)0xde5fb9d2630458e9ULL;
    return (uint64_t)(p >> 64) + (uint64_t)p;
}

int bits_for(size_t n) {
    int b = 1;
    while (b < 63 && ((size_t)1 << b) <= n) b++;
    return b;
}

}  // namespace

size_t rh_capacity(size_t max_count) {
    size_t cap = 8;
    while (cap * 80 / 100 < max_count) cap *= 2;
    return cap;
}

int rh_log2(size_t cap) {
    int l = 0;
    while (((size_t)1 << l) < cap) l++;
    return l;
}

int rh_max_displacement(const int32_t* keys, size_t n, size_t cap) {
    const uint64_t mask = (uint64_t)cap - 1;
    int64_t prev = -1;
    int maxd = 0;
    for (size_t i = 0; i < n; i++) {
        const int64_t home = (int64_t)(hash_int(keys[i]) & mask);
        const int64_t slot = home > prev + 1 ? home : prev + 1;
        if (slot - home > maxd) maxd = (int)(slot - home);
        prev = slot;
    }
    return maxd;
}

int rh_packed_displacement(const int32_t* keys, size_t n, size_t cap) {
    // displacements of a packed, home-ordered layout depend only on how many keys share each home
    std::vector<uint32_t> cnt(cap, 0);
    const uint64_t mask = (uint64_t)cap - 1;
    for (size_t i = 0; i < n; i++) cnt[(size_t)(hash_int(keys[i]) & mask)]++;
    int64_t next = 0;  // first free slot
    int maxd = 0;
    for (size_t h = 0; h < cap; h++) {
        if (!cnt[h]) continue;
        const int64_t first = (int64_t)h > next ? (int64_t)h : next;
        const int64_t last = first + (int64_t)cnt[h] - 1;
        if (last - (int64_t)h > maxd) maxd = (int)(last - (int64_t)h);
        next = last + 1;
    }
    return maxd;
}

std::vector<int32_t> rh_sort_keys(const std::vector<int32_t>& keys, size_t cap) {
    const size_t n = keys.size();
    const int ib = bits_for(n);
    const int cl = rh_log2(cap);
    const int total = cl + 5 + ib;
    std::vector<uint64_t> a(n), b(n);
    std::vector<int32_t> va(n), vb(n);
    for (size_t i = 0; i < n; i++) {
        const uint64_t h = hash_int(keys[i]);
        a[i] = ((((h & ((uint64_t)cap - 1)) << 5) | (31 - (h >> 59))) << ib) | (uint64_t)i;
        va[i] = keys[i];
    }
    // LSD radix sort, 11-bit digits
    constexpr int D = 11;
    std::vector<size_t> cnt((size_t)1 << D);
    for (int sh = 0; sh < total; sh += D) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (size_t i = 0; i < n; i++) cnt[(a[i] >> sh) & ((1u << D) - 1)]++;
        size_t acc = 0;
        for (auto& c : cnt) {
            size_t t = c;
            c = acc;
            acc += t;
        }
        for (size_t i = 0; i < n; i++) {
            const size_t d = (a[i] >> sh) & ((1u << D) - 1);
            b[cnt[d]] = a[i];
            vb[cnt[d]++] = va[i];
        }
        a.swap(b);
        va.swap(vb);
    }
    return va;
}

namespace {

// Applies the operations with the home slots of the next keys prefetched.
void apply_ops(RHOrder& m, const std::vector<int32_t>& ops, const std::vector<uint8_t>* transient) {
    constexpr size_t AHEAD = 24;
    const size_t n = ops.size();
    for (size_t i = 0; i < n; i++) {
        if (i + AHEAD < n) m.prefetch(ops[i + AHEAD]);
        m.insert(ops[i]);
        if (transient && (*transient)[i]) m.erase(ops[i]);
    }
}

}  // namespace

void replay_created(RHOrder& m, const std::vector<int32_t>& created) { apply_ops(m, created, nullptr); }

std::vector<int32_t> replay_ops(const std::vector<int32_t>& ops, const std::vector<uint8_t>& transient) {
    RHOrder m;
    apply_ops(m, ops, &transient);
    std::vector<int32_t> out;
    out.reserve(m.size());
    m.for_each([&](int32_t k) { out.push_back(k); });
    return out;
}

namespace {

// the operations ToVarsBuilder made on alignedVariants, per thread chunk of the processing order
void aligned_ops(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo, int status_n,
                 unsigned threads, std::vector<std::vector<int32_t>>& ops, std::vector<std::vector<uint8_t>>& tr) {
    // the filter reads one status byte per position in a scattered order: split it over threads,
    // keeping each chunk's order
    const size_t n = processing.size();
    const size_t nt = n < 65536 ? 1 : std::max(1u, std::min<unsigned>(threads, 16));
    ops.assign(nt, {});
    tr.assign(nt, {});
    auto filter = [&](size_t t) {
        const size_t a = n * t / nt, b = n * (t + 1) / nt;
        ops[t].reserve((b - a) / 4);
        tr[t].reserve((b - a) / 4);
        for (size_t i = a; i < b; i++) {
            if (i + 32 < b) {
                const int k = processing[i + 32] - status_lo;
                if (k >= 0 && k < status_n) __builtin_prefetch(status + k);
            }
            const int32_t p = processing[i];
            const int k = p - status_lo;
            const int st = (k >= 0 && k < status_n) ? status[k] : 0;
            if (st == 0) continue;
            ops[t].push_back(p);
            tr[t].push_back(st == 2 ? 1 : 0);
        }
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nt; t++) pool.emplace_back(filter, t);
    filter(0);
    for (auto& th : pool) th.join();
}

}  // namespace

void aligned_table(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo, int status_n,
                   RHOrder& out, unsigned threads) {
    std::vector<std::vector<int32_t>> ops;
    std::vector<std::vector<uint8_t>> tr;
    aligned_ops(processing, status, status_lo, status_n, threads, ops, tr);
    out = RHOrder();
    for (size_t t = 0; t < ops.size(); t++) apply_ops(out, ops[t], &tr[t]);
}

void aligned_op_list(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo, int status_n,
                     unsigned threads, std::vector<int32_t>& all, std::vector<uint8_t>& all_tr) {
    std::vector<std::vector<int32_t>> ops;
    std::vector<std::vector<uint8_t>> tr;
    aligned_ops(processing, status, status_lo, status_n, threads, ops, tr);
    size_t n = 0;
    for (auto& v : ops) n += v.size();
    all.clear();
    all_tr.clear();
    all.reserve(n);
    all_tr.reserve(n);
    for (size_t t = 0; t < ops.size(); t++) {
        all.insert(all.end(), ops[t].begin(), ops[t].end());
        all_tr.insert(all_tr.end(), tr[t].begin(), tr[t].end());
    }
}

std::vector<int32_t> aligned_slot_order(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo,
                                        int status_n, unsigned threads, bool* sorted) {
    std::vector<int32_t> all;
    std::vector<uint8_t> all_tr;
    aligned_op_list(processing, status, status_lo, status_n, threads, all, all_tr);
    std::vector<int32_t> out;
    const bool ok = sorted_ops_order(all.data(), all_tr.data(), all.size(), out, threads);
    if (sorted) *sorted = ok;
    if (!ok) out = replay_ops(all, all_tr);
    return out;
}

std::vector<int32_t> aligned_order(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo,
                                   int status_n) {
    // alignedVariants holds a few of the positions in a scattered order, so the sorted form rarely
    // applies (its clusters are long): the replay is as fast as sorting when it would
    RHOrder m;
    aligned_table(processing, status, status_lo, status_n, m, 1);
    std::vector<int32_t> out;
    out.reserve(m.size());
    m.for_each([&](int32_t k) { out.push_back(k); });
    return out;
}

namespace {

// displacement at which an info byte overflows at increments 32, 16, 8, 4, 2
constexpr int RH_LEVEL_DISP[5] = {6, 14, 30, 62, 126};
constexpr uint32_t NEVER = UINT32_MAX;

// A fork-join group of threads kept for one call: run(f) calls f(t) for t in [0, size()) and
// returns when all have finished.  The workers spin (yielding) between tasks: a call makes many
// short parallel steps, and waking sleeping threads would cost more than the steps.
class ForkJoin {
  public:
    explicit ForkJoin(unsigned n) : n_(n ? n : 1) {
        for (unsigned t = 1; t < n_; t++) th_.emplace_back([this, t] { loop(t); });
    }
    ~ForkJoin() {
        stop_.store(true, std::memory_order_relaxed);
        gen_.fetch_add(1, std::memory_order_release);
        for (auto& t : th_) t.join();
    }
    unsigned size() const { return n_; }
    void run(const std::function<void(unsigned)>& f) {
        if (n_ == 1) {
            f(0);
            return;
        }
        task_ = &f;
        pending_.store(n_ - 1, std::memory_order_relaxed);
        gen_.fetch_add(1, std::memory_order_release);
        f(0);
        for (unsigned spins = 0; pending_.load(std::memory_order_acquire) != 0; spins++)
            if (spins > 64) std::this_thread::yield();
    }

  private:
    void loop(unsigned t) {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g;
            for (unsigned spins = 0; (g = gen_.load(std::memory_order_acquire)) == seen; spins++)
                if (spins > 64) std::this_thread::yield();
            seen = g;
            if (stop_.load(std::memory_order_relaxed)) return;
            (*task_)(t);
            pending_.fetch_sub(1, std::memory_order_release);
        }
    }
    unsigned n_;
    std::vector<std::thread> th_;
    const std::function<void(unsigned)>* task_ = nullptr;
    std::atomic<uint64_t> gen_{0};
    std::atomic<unsigned> pending_{0};
    std::atomic<bool> stop_{false};
};

// a kept key of an epoch, in the home-sorted list
struct Placed {
    uint32_t home;
    uint32_t step;  // placement index in the epoch
    int32_t key;
};

struct Run {
    uint32_t lo, hi;  // its kept keys: [lo, hi) of the home-sorted list
    uint32_t a, b;    // first and last slot
    int maxd;         // largest displacement in the final packing
};

// The events of one run ((step, home - run start), in step order; transient keys are removed
// right after their insertion) through a slot model of the run: an insertion goes after the keys
// of smaller or equal home and shifts the rest up to the first free slot.  Records the first step
// at which each level's overflow displacement is reached, up to level `upto`.
void simulate_run(const std::vector<std::pair<uint32_t, int32_t>>& ev, const std::vector<uint8_t>& ev_tr,
                  uint32_t len, int upto, std::vector<int32_t>& occ, uint32_t first[5]) {
    occ.assign(len + 2, -1);  // home of the key in each slot, -1 free
    int maxd = 0, reached = 0;
    for (size_t i = 0; i < ev.size() && reached <= upto; i++) {
        const int32_t u = ev[i].second;
        uint32_t p = (uint32_t)u;
        while (occ[p] >= 0 && occ[p] <= u) p++;
        int md = (int)p - u;
        uint32_t q = p;
        while (occ[q] >= 0) {
            const int d = (int)q + 1 - occ[q];
            md = d > md ? d : md;
            q++;
        }
        md = md > maxd ? md : maxd;
        if (!ev_tr[i]) {
            for (uint32_t j = q; j > p; j--) occ[j] = occ[j - 1];
            occ[p] = u;
            maxd = md;
        }
        while (reached < 5 && md >= RH_LEVEL_DISP[reached]) first[reached++] = ev[i].first;
    }
}

template <class T>
struct Buf {  // uninitialised array: first touched by the threads that fill it
    std::unique_ptr<T[]> p;
    size_t n = 0;
    void reserve(size_t k) {
        if (k > n) {
            p.reset(new T[k]);
            n = k;
        }
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

struct Epoch {
    ForkJoin& pool;
    Buf<Placed> bucket, sorted;
    Buf<int32_t> pmax;  // prefix max of (home - index) within a partition

    explicit Epoch(ForkJoin& p, size_t max_keep) : pool(p) {
        bucket.reserve(max_keep);
        sorted.reserve(max_keep);
        pmax.reserve(max_keep);
    }

    // One epoch: prev[0, np) re-inserted in order, then ops[0, m).  Writes the slot order of the
    // n_keep kept keys to out and returns n_keep, or -1 when the premise fails.
    int64_t run(size_t cap, const int32_t* prev, size_t np, const int32_t* ops, const uint8_t* tr, size_t m,
                int32_t* out);
};

int64_t Epoch::run(size_t cap, const int32_t* prev, size_t np, const int32_t* ops, const uint8_t* tr, size_t m,
                   int32_t* out) {
    const size_t S = np + m;
    const uint64_t mask = (uint64_t)cap - 1;
    const unsigned T = pool.size();
    Placed* const bk = bucket.p.get();
    Placed* const srt = sorted.p.get();
    int32_t* const pm = pmax.p.get();
    // partitions: contiguous home ranges of at most 2^16 homes (a partition's counters stay in
    // cache), at least one per thread
    const int lcap = rh_log2(cap);
    int lp = lcap > 16 ? lcap - 16 : 0;
    while ((1u << lp) < T && lp < lcap) lp++;
    const unsigned P = 1u << lp;
    const int pshift = lcap - lp;
    const size_t width = (size_t)1 << pshift;
    auto key_at = [prev, np, ops](size_t s) { return s < np ? prev[s] : ops[s - np]; };
    auto transient_at = [np, tr](size_t s) { return s >= np && tr[s - np] != 0; };
    // 1. count kept keys per (thread chunk, partition); transient keys as (home, step) per chunk
    std::vector<std::vector<uint32_t>> cnt(T, std::vector<uint32_t>(P, 0));
    std::vector<std::vector<uint64_t>> trs_t(T);
    pool.run([&](unsigned t) {
        const size_t a = S * t / T, b = S * (t + 1) / T;
        uint32_t* c = cnt[t].data();
        std::vector<uint64_t>& tv = trs_t[t];
        for (size_t s = a; s < b; s++) {
            const uint32_t hm = (uint32_t)(hash_int(key_at(s)) & mask);
            if (transient_at(s))
                tv.push_back(((uint64_t)hm << 32) | s);
            else
                c[hm >> pshift]++;
        }
    });
    // 2. bucket offsets (partition-major, thread chunks in step order)
    std::vector<size_t> pstart(P + 1, 0);
    std::vector<std::vector<size_t>> off(T, std::vector<size_t>(P, 0));
    {
        size_t acc = 0;
        for (unsigned p = 0; p < P; p++) {
            pstart[p] = acc;
            for (unsigned t = 0; t < T; t++) {
                off[t][p] = acc;
                acc += cnt[t][p];
            }
        }
        pstart[P] = acc;
    }
    const size_t n_keep = pstart[P];
    // 3. scatter the kept keys to their partitions (stable)
    pool.run([&](unsigned t) {
        const size_t a = S * t / T, b = S * (t + 1) / T;
        size_t* o = off[t].data();
        for (size_t s = a; s < b; s++) {
            if (transient_at(s)) continue;
            const int32_t k = key_at(s);
            const uint32_t hm = (uint32_t)(hash_int(k) & mask);
            bk[o[hm >> pshift]++] = {hm, (uint32_t)s, k};
        }
    });
    // 4. per partition: counting sort by home (stable), then the prefix max of (home - index);
    // a packed slot is index + the max over this and earlier keys of (home - index)
    std::vector<int64_t> part_max(P, INT64_MIN);
    pool.run([&](unsigned t) {
        std::vector<uint32_t> c(width + 1);
        for (unsigned p = t; p < P; p += T) {
            const size_t a = pstart[p], b = pstart[p + 1];
            const uint32_t base = (uint32_t)p << pshift;
            std::fill(c.begin(), c.end(), 0);
            for (size_t i = a; i < b; i++) c[bk[i].home - base + 1]++;
            for (size_t h = 0; h < width; h++) c[h + 1] += c[h];
            for (size_t i = a; i < b; i++) srt[a + c[bk[i].home - base]++] = bk[i];
            int64_t mx = INT64_MIN;
            for (size_t i = a; i < b; i++) {
                const int64_t v = (int64_t)srt[i].home - (int64_t)i;
                mx = v > mx ? v : mx;
                pm[i] = (int32_t)mx;
            }
            part_max[p] = mx;
        }
    });
    std::vector<int64_t> carry(P + 1, INT64_MIN);  // max of (home - index) over earlier partitions
    for (unsigned p = 0; p < P; p++) carry[p + 1] = std::max(carry[p], part_max[p]);
    const int64_t* cy = carry.data();
    // 5. runs (separated by a free slot), each found by the partition holding its first key
    std::vector<std::vector<Run>> runs_t(P);
    pool.run([&](unsigned t) {
        for (unsigned p = t; p < P; p += T) {
            const size_t a = pstart[p], b = pstart[p + 1];
            std::vector<Run>& rs = runs_t[p];
            rs.reserve((b - a) / 4 + 1);
            int64_t prev_slot = a > 0 ? (int64_t)(a - 1) + cy[p] : -2;
            const int64_t cp = cy[p];
            Run* cur = nullptr;
            for (size_t i = a; i < b; i++) {
                const int64_t sl = (int64_t)i + std::max<int64_t>(cp, pm[i]);
                const uint32_t hm = srt[i].home;
                if ((int64_t)hm > prev_slot + 1) {
                    if (cur) cur->hi = (uint32_t)i;
                    rs.push_back({(uint32_t)i, (uint32_t)i, hm, hm, 0});
                    cur = &rs.back();
                }
                if (cur) {
                    const int d = (int)(sl - (int64_t)hm);
                    cur->maxd = d > cur->maxd ? d : cur->maxd;
                    cur->b = (uint32_t)sl;
                }
                prev_slot = sl;
            }
            if (!cur) continue;
            // the last run continues until the next run's start, possibly past this partition
            size_t i = b;
            unsigned q = p;
            while (i < n_keep) {
                while (q + 1 < P && i >= pstart[q + 1]) q++;
                const int64_t sl = (int64_t)i + std::max<int64_t>(cy[q], pm[i]);
                const uint32_t hm = srt[i].home;
                if ((int64_t)hm > prev_slot + 1) break;
                const int d = (int)(sl - (int64_t)hm);
                cur->maxd = d > cur->maxd ? d : cur->maxd;
                cur->b = (uint32_t)sl;
                prev_slot = sl;
                i++;
            }
            cur->hi = (uint32_t)i;
        }
    });
    // a partition's keys before its first run start belong to an earlier partition's run
    std::vector<size_t> run0(P + 1, 0);  // index of partition p's first run
    for (unsigned p = 0; p < P; p++) run0[p + 1] = run0[p] + runs_t[p].size();
    const size_t n_runs = run0[P];
    std::vector<Run> runs(n_runs);
    pool.run([&](unsigned t) {
        for (unsigned p = t; p < P; p += T) std::copy(runs_t[p].begin(), runs_t[p].end(), runs.begin() + (ptrdiff_t)run0[p]);
    });
    // 6. transient keys inside a run (one at a free slot sits at its home and moves nothing),
    // as (run, step) pairs sorted by run then step; runs[r]'s are trs[tr_lo[r], tr_lo[r + 1])
    std::vector<std::vector<uint64_t>> trp(P);
    {
        std::vector<uint64_t> all;
        for (auto& v : trs_t) all.insert(all.end(), v.begin(), v.end());
        std::vector<size_t> tcount(P + 1, 0);
        for (uint64_t x : all) tcount[(size_t)((x >> 32) >> pshift) + 1]++;
        for (unsigned p = 0; p < P; p++) trp[p].reserve(tcount[p + 1]);
        for (uint64_t x : all) trp[(size_t)((x >> 32) >> pshift)].push_back(x);
    }
    pool.run([&](unsigned t) {
        for (unsigned p = t; p < P; p += T) {
            std::vector<uint64_t>& v = trp[p];
            std::sort(v.begin(), v.end());
            // start from the run before the partition's first (it may reach into this one)
            size_t r = run0[p] > 0 ? run0[p] - 1 : 0;
            size_t w = 0;
            for (uint64_t x : v) {
                const uint32_t h = (uint32_t)(x >> 32);
                while (r + 1 < n_runs && runs[r + 1].a <= h) r++;
                if (r < n_runs && runs[r].a <= h && h <= runs[r].b) v[w++] = ((uint64_t)r << 32) | (uint32_t)x;
            }
            v.resize(w);
        }
    });
    std::vector<uint64_t> trs;  // (run << 32) | step
    for (auto& v : trp) trs.insert(trs.end(), v.begin(), v.end());
    std::sort(trs.begin(), trs.end());
    std::vector<uint32_t> tr_lo(n_runs + 1, 0);
    for (uint64_t x : trs) tr_lo[(size_t)(x >> 32) + 1]++;
    for (size_t r = 0; r < n_runs; r++) tr_lo[r + 1] += tr_lo[r];
    auto events = [&](size_t r, std::vector<std::pair<uint32_t, int32_t>>& ev, std::vector<uint8_t>& ev_tr) {
        const Run& R = runs[r];
        ev.clear();
        for (uint32_t i = R.lo; i < R.hi; i++) ev.push_back({srt[i].step, (int32_t)(srt[i].home - R.a)});
        for (uint32_t j = tr_lo[r]; j < tr_lo[r + 1]; j++) {
            const uint32_t s = (uint32_t)trs[j];
            ev.push_back({s, (int32_t)((uint32_t)(hash_int(key_at(s)) & mask) - R.a)});
        }
        std::sort(ev.begin(), ev.end());
        ev_tr.resize(ev.size());
        for (size_t i = 0; i < ev.size(); i++) ev_tr[i] = transient_at(ev[i].first);
    };
    // 7. first step reaching each level's displacement.  Runs that can reach 14 are simulated in
    // full; for 6 alone, runs in order of a lower bound (the step of their seventh event: a
    // displacement of 6 needs seven keys), in batches, until the bound passes the earliest step.
    uint32_t first[5] = {NEVER, NEVER, NEVER, NEVER, NEVER};
    {
        std::vector<std::vector<uint32_t>> deep_t(T), shallow_t(T);
        pool.run([&](unsigned t) {
            const size_t a = n_runs * t / T, b = n_runs * (t + 1) / T;
            for (size_t r = a; r < b; r++) {
                const int reach = runs[r].maxd + (tr_lo[r + 1] > tr_lo[r] ? 1 : 0);
                if (reach >= RH_LEVEL_DISP[1])
                    deep_t[t].push_back((uint32_t)r);
                else if (reach >= RH_LEVEL_DISP[0])
                    shallow_t[t].push_back((uint32_t)r);
            }
        });
        std::vector<uint32_t> deep, shallow;
        for (unsigned t = 0; t < T; t++) {
            deep.insert(deep.end(), deep_t[t].begin(), deep_t[t].end());
            shallow.insert(shallow.end(), shallow_t[t].begin(), shallow_t[t].end());
        }
        std::vector<std::array<uint32_t, 5>> fd(deep.size());
        std::vector<std::pair<uint32_t, uint32_t>> lb(shallow.size());  // (bound, run)
        pool.run([&](unsigned t) {
            std::vector<std::pair<uint32_t, int32_t>> ev;
            std::vector<uint8_t> ev_tr;
            std::vector<int32_t> occ;
            for (size_t i = t; i < deep.size(); i += T) {
                events(deep[i], ev, ev_tr);
                uint32_t f[5] = {NEVER, NEVER, NEVER, NEVER, NEVER};
                simulate_run(ev, ev_tr, runs[deep[i]].b - runs[deep[i]].a + 1, 4, occ, f);
                for (int l = 0; l < 5; l++) fd[i][l] = f[l];
            }
            std::vector<uint32_t> st;
            const size_t a = shallow.size() * t / T, b = shallow.size() * (t + 1) / T;
            for (size_t i = a; i < b; i++) {
                const size_t r = shallow[i];
                const Run& R = runs[r];
                st.clear();
                for (uint32_t k = R.lo; k < R.hi; k++) st.push_back(srt[k].step);
                for (uint32_t j = tr_lo[r]; j < tr_lo[r + 1]; j++) st.push_back((uint32_t)trs[j]);
                const size_t k = (size_t)RH_LEVEL_DISP[0];  // the seventh smallest
                uint32_t bound = NEVER;
                if (st.size() > k) {
                    std::nth_element(st.begin(), st.begin() + (ptrdiff_t)k, st.end());
                    bound = st[k];
                }
                lb[i] = {bound, (uint32_t)r};
            }
        });
        for (size_t i = 0; i < deep.size(); i++)
            for (int l = 0; l < 5; l++) first[l] = std::min(first[l], fd[i][l]);
        std::sort(lb.begin(), lb.end());
        size_t pos = 0, batch = 4 * (size_t)T;
        std::vector<uint32_t> f0;
        while (pos < lb.size() && lb[pos].first < first[0]) {
            const size_t end = std::min(lb.size(), pos + batch);
            const uint32_t limit = first[0];
            f0.assign(end - pos, NEVER);
            pool.run([&](unsigned t) {
                std::vector<std::pair<uint32_t, int32_t>> ev;
                std::vector<uint8_t> ev_tr;
                std::vector<int32_t> occ;
                for (size_t i = pos + t; i < end; i += T) {
                    if (lb[i].first >= limit) continue;
                    events(lb[i].second, ev, ev_tr);
                    uint32_t f[5] = {NEVER, NEVER, NEVER, NEVER, NEVER};
                    simulate_run(ev, ev_tr, runs[lb[i].second].b - runs[lb[i].second].a + 1, 0, occ, f);
                    f0[i - pos] = f[0];
                }
            });
            for (uint32_t v : f0) first[0] = std::min(first[0], v);
            pos = end;
            batch *= 2;
        }
    }
    if (first[4] != NEVER) return -1;
    // a flag raised at step t halves the increment before step t + 1
    uint32_t halve_at[4];
    int nh = 0;
    for (int l = 0; l < 4; l++)
        if (first[l] != NEVER && (size_t)first[l] + 1 < S) halve_at[nh++] = first[l] + 1;
    // 8. per home: a stable sort by the level key (the list is in step order within a home)
    std::atomic<bool> too_deep{false};
    pool.run([&](unsigned t) {
        size_t i = n_keep * t / T, end = n_keep * (t + 1) / T;
        // chunks start and end at home boundaries
        while (i > 0 && i < n_keep && srt[i].home == srt[i - 1].home) i++;
        while (end > 0 && end < n_keep && srt[end].home == srt[end - 1].home) end++;
        uint32_t gk[64];
        int32_t gv[64];
        while (i < end) {
            const uint32_t hm = srt[i].home;
            if (i + 1 == n_keep || srt[i + 1].home != hm) {
                out[i] = srt[i].key;
                i++;
                continue;
            }
            size_t j = i + 1;
            while (j < n_keep && srt[j].home == hm) j++;
            if (j - i > 64) {
                too_deep = true;
                return;
            }
            size_t g = 0;
            for (size_t x = i; x < j; x++, g++) {
                const Placed& e = srt[x];
                const uint32_t hb = (uint32_t)(hash_int(e.key) >> 59);
                int ph = 0;
                while (ph < nh && halve_at[ph] <= e.step) ph++;
                uint32_t k = (~(hb >> nh)) & ((1u << (5 - nh)) - 1);
                for (int l = nh - 1; l >= 0; l--) {
                    const uint32_t late = ph > l;
                    const uint32_t bit = late ? 0u : ((~hb >> l) & 1u);
                    k = (k << 2) | (late << 1) | bit;
                }
                size_t y = g;
                while (y > 0 && gk[y - 1] > k) {
                    gk[y] = gk[y - 1];
                    gv[y] = gv[y - 1];
                    y--;
                }
                gk[y] = k;
                gv[y] = e.key;
            }
            for (size_t x = 0; x < g; x++) out[i + x] = gv[x];
            i = j;
        }
    });
    return too_deep ? -1 : (int64_t)n_keep;
}

}  // namespace

bool sorted_ops_order(const int32_t* ops, const uint8_t* transient, size_t n, std::vector<int32_t>& out,
                      unsigned threads) {
    out.clear();
    if (n == 0) return true;
    // growths: an insertion that finds 80% of the slots used doubles the table first
    std::vector<size_t> growth_at;
    size_t cap = 8, count = 0, max_keep = 0;
    for (size_t i = 0; i < n; i++) {
        if (count >= cap * 80 / 100) {
            growth_at.push_back(i);
            cap *= 2;
        }
        count += transient[i] ? 0 : 1;
        max_keep = std::max(max_keep, count);
    }
    // epochs of tables up to 2^16 slots are replayed (they stay in cache); the rest are sorted
    constexpr size_t SORT_FROM_CAP = (size_t)1 << 17;
    size_t e0 = 0, c = 8;
    while (e0 < growth_at.size() && c < SORT_FROM_CAP) {
        e0++;
        c *= 2;
    }
    const size_t a0 = e0 > 0 ? growth_at[e0 - 1] : 0;  // first operation of epoch e0
    Buf<int32_t> cur, nxt;
    cur.reserve(max_keep);
    nxt.reserve(max_keep);
    size_t n_cur = 0;
    if (c < SORT_FROM_CAP) {
        // the whole sequence stays in small tables
        RHOrder m;
        for (size_t i = 0; i < n; i++) {
            m.insert(ops[i]);
            if (transient[i]) m.erase(ops[i]);
        }
        out.reserve(m.size());
        m.for_each([&](int32_t k) { out.push_back(k); });
        return true;
    }
    {
        RHOrder m;
        for (size_t i = 0; i < a0; i++) {
            m.insert(ops[i]);
            if (transient[i]) m.erase(ops[i]);
        }
        m.for_each([&](int32_t k) { cur[n_cur++] = k; });
    }
    ForkJoin pool(std::max(1u, std::min(threads, 16u)));
    Epoch ep(pool, max_keep);
    size_t a = a0;
    for (size_t e = e0; e <= growth_at.size(); e++) {
        const size_t b = e < growth_at.size() ? growth_at[e] : n;
        const int64_t k = ep.run(c, cur.p.get(), n_cur, ops + a, transient + a, b - a, nxt.p.get());
        if (k < 0) return false;
        std::swap(cur, nxt);
        n_cur = (size_t)k;
        a = b;
        c *= 2;
    }
    out.assign(cur.p.get(), cur.p.get() + n_cur);
    return true;
}

std::vector<int32_t> noninsertion_order(const std::vector<int32_t>& created, const std::vector<KeyOp>& log) {
    RHOrder m;
    replay_created(m, created);
    return replay_finish(m, log);
}

std::vector<int32_t> replay_finish(RHOrder& m, const std::vector<KeyOp>& log) {
    for (const KeyOp& op : log) {
        if (op.insert)
            m.insert(op.pos);
        else
            m.erase(op.pos);
    }
    std::vector<int32_t> out;
    out.reserve(m.size());
    m.for_each([&](int32_t k) { out.push_back(k); });
    return out;
}

bool noninsertion_order_sorted(const std::vector<int32_t>& walk_sorted, const std::vector<KeyOp>& log,
                               std::vector<int32_t>& out) {
    const int maxd = rh_max_displacement(walk_sorted.data(), walk_sorted.size(), rh_capacity(walk_sorted.size()));
    return noninsertion_order_sorted(walk_sorted, maxd, log, out);
}

bool noninsertion_order_sorted(const std::vector<int32_t>& walk_sorted, int maxd, const std::vector<KeyOp>& log,
                               std::vector<int32_t>& out) {
    const size_t n_walk = walk_sorted.size();
    const size_t cap = rh_capacity(n_walk);
    if (log.empty()) {
        if (maxd >= RH_OVERFLOW_DISP) return false;
        out = walk_sorted;
        return true;
    }
    // the realigner's operations: erased walk keys leave the order; inserted keys join it after
    // every earlier key of equal (home, hash bits).  Walk keys and their erasures are bitmaps over
    // the walk keys' span (a few million positions: hash sets of them cost more than the rest).
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int32_t k : walk_sorted) {
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    const size_t span = n_walk ? (size_t)((int64_t)hi - lo + 1) : 0;
    std::vector<uint64_t> is_walk((span + 63) / 64, 0), was_erased((span + 63) / 64, 0);
    auto bit = [&](const std::vector<uint64_t>& b, int32_t k) {
        if (!n_walk || k < lo || k > hi) return false;
        const size_t i = (size_t)((int64_t)k - lo);
        return ((b[i >> 6] >> (i & 63)) & 1u) != 0;
    };
    auto set_bit = [&](std::vector<uint64_t>& b, int32_t k) {
        const size_t i = (size_t)((int64_t)k - lo);
        b[i >> 6] |= 1ull << (i & 63);
    };
    for (int32_t k : walk_sorted) set_bit(is_walk, k);
    std::unordered_set<int32_t> added_set;
    std::vector<int32_t> added;  // in insertion order
    size_t count = n_walk, max_count = n_walk;
    for (const KeyOp& op : log) {
        const bool walk = bit(is_walk, op.pos);
        const bool present = (walk && !bit(was_erased, op.pos)) || added_set.count(op.pos);
        if (op.insert) {
            if (present) continue;
            if (walk) set_bit(was_erased, op.pos);  // re-inserted: now the latest key
            added_set.insert(op.pos);
            added.erase(std::remove(added.begin(), added.end(), op.pos), added.end());
            added.push_back(op.pos);
            ++count;
            if (count > max_count) max_count = count;
        } else {
            if (!present) continue;
            if (added_set.count(op.pos)) {
                added_set.erase(op.pos);
                added.erase(std::remove(added.begin(), added.end(), op.pos), added.end());
            } else {
                set_bit(was_erased, op.pos);
            }
            --count;
        }
    }
    if (rh_capacity(max_count) != cap) return false;
    // the walk's packing already holds every walk key; each added key lengthens a cluster by at
    // most one slot
    if (maxd + (int)added.size() >= RH_OVERFLOW_DISP) return false;
    const uint64_t mask = (uint64_t)cap - 1;
    auto rank = [&](int32_t k) {  // (home, hash bits desc)
        const uint64_t h = hash_int(k);
        return ((h & mask) << 5) | (31 - (h >> 59));
    };
    std::vector<std::pair<uint64_t, int32_t>> extra;
    for (int32_t k : added) extra.push_back({rank(k), k});
    std::stable_sort(extra.begin(), extra.end(),
                     [](const std::pair<uint64_t, int32_t>& x, const std::pair<uint64_t, int32_t>& y) { return x.first < y.first; });
    out.clear();
    out.reserve(n_walk + extra.size());
    size_t e = 0;
    for (size_t i = 0; i < n_walk; i++) {
        const int32_t k = walk_sorted[i];
        if (e < extra.size()) {
            const uint64_t r = rank(k);
            while (e < extra.size() && extra[e].first < r) out.push_back(extra[e++].second);
        }
        if (!bit(was_erased, k)) out.push_back(k);
    }
    while (e < extra.size()) out.push_back(extra[e++].second);
    return true;
}

}  // namespace rv
