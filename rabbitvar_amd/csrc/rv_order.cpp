// Slot-order replay and sort of the reference's robin_hood maps (see rv_order.h).
#include "rv_order.h"

#include <algorithm>
#include <climits>
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

void aligned_table(const std::vector<int32_t>& processing, const uint8_t* status, int status_lo, int status_n,
                   RHOrder& out, unsigned threads) {
    // the filter reads one status byte per position in a scattered order: split it over threads,
    // keeping each chunk's order, then replay the operations in order
    const size_t n = processing.size();
    const size_t nt = n < 65536 ? 1 : std::max(1u, std::min<unsigned>(threads, 16));
    std::vector<std::vector<int32_t>> ops(nt);
    std::vector<std::vector<uint8_t>> tr(nt);
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
    out = RHOrder();
    for (size_t t = 0; t < nt; t++) apply_ops(out, ops[t], &tr[t]);
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
