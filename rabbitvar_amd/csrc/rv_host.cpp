// Host-side reduction of device walk records.  See rv_host.h.
#include "rv_host.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace rv {

std::string fmt_f(double v) {
    char buf[512];
    snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

// robin_hood::unordered_map<char,int> slot placement for an insert-only map (robin_hood 3.4.3:
// 8 initial slots, 5 info bits, max load 80%, hash_int = hi+lo of the 128-bit product with
// 0xde5fb9d2630458e9; insertion walks right while the resident info byte >= ours and shifts the
// run up; iteration visits slots in ascending order).  With at most 16 distinct byte keys the map
// may grow once (8 -> 16 slots), which is modelled by re-inserting in old slot order.
namespace {
struct CharTable {
    int slots = 8;
    unsigned inc = 32, shift = 59;
    size_t max_allowed = 6;
    std::vector<int> key;       // -1 empty
    std::vector<unsigned> info;
    size_t n = 0;

    static uint64_t mix(uint64_t x) {
        unsigned __int128 p = (unsigned __int128)x * (unsigned __int128)0xde5fb9d2630458e9ULL;
        return (uint64_t)(p >> 64) + (uint64_t)p;
    }
    size_t buffer() const {
        size_t m = (size_t)slots * 80 / 100;
        return (size_t)slots + (m < 255 ? m : 255);
    }
    void init(int s) {
        slots = s;
        inc = 32;
        shift = 59;
        max_allowed = (size_t)s * 80 / 100;
        key.assign(buffer() + 8, -1);
        info.assign(buffer() + 8, 0);
        info[buffer()] = 1;
        n = 0;
    }
    void place(int k, bool rehash) {
        uint64_t h = mix((uint64_t)(int64_t)(char)k);
        unsigned inf = inc + (unsigned)(h >> shift);
        size_t idx = (size_t)(h & (uint64_t)(slots - 1));
        if (rehash) {
            while (inf <= info[idx]) {
                idx++;
                inf += inc;
            }
        } else {
            while (inf < info[idx]) {
                idx++;
                inf += inc;
            }
            while (inf == info[idx]) {  // equal-info residents are other keys: go past them
                idx++;
                inf += inc;
            }
        }
        size_t ins = idx;
        unsigned ins_info = rehash ? (inf & 0xffu) : inf;
        if (ins_info + inc > 0xff) max_allowed = 0;
        while (info[idx] != 0) {
            idx++;
            inf += inc;
        }
        while (idx != ins) {
            key[idx] = key[idx - 1];
            info[idx] = (info[idx - 1] + inc) & 0xffu;
            if (info[idx] + inc > 0xff) max_allowed = 0;
            idx--;
        }
        key[ins] = k;
        info[ins] = ins_info & 0xffu;
        n++;
    }
    bool try_increase_info() {
        if (inc <= 2) return false;
        inc >>= 1;
        shift++;
        size_t words = buffer() / 8;
        for (size_t i = 0; i < words * 8; i++) info[i] = (info[i] >> 1) & 0x7fu;
        max_allowed = (size_t)slots * 80 / 100;
        return true;
    }
    void insert(int k) {
        if (n >= max_allowed) {
            if (!(n < (size_t)slots * 80 / 100 && try_increase_info())) {
                std::vector<int> ok = key;
                std::vector<unsigned> oi = info;
                size_t ob = buffer();
                init(slots * 2);
                for (size_t i = 0; i < ob; i++)
                    if (oi[i] != 0) {
                        if (max_allowed == 0 && !try_increase_info()) throw std::runtime_error("rh overflow");
                        place(ok[i], true);
                    }
            }
        }
        place(k, false);
    }
};
}  // namespace

static std::vector<char> char_order_uncached(const std::vector<char>& inserted) {
    CharTable t;
    t.init(8);
    for (char c : inserted) t.insert((int)c);
    std::vector<char> out;
    for (size_t i = 0; i < t.buffer(); i++)
        if (t.info[i] != 0) out.push_back((char)t.key[i]);
    return out;
}

std::vector<char> rh_char_iteration_order(const std::vector<char>& inserted) {
    if (inserted.size() > 7) return char_order_uncached(inserted);
    char buf[8];
    int n = rh_char_order_small(inserted.data(), (int)inserted.size(), buf);
    return std::vector<char>(buf, buf + n);
}

int rh_char_order_small(const char* inserted, int n, char* out) {
    // soft-clip tables see a handful of distinct insertion sequences (first-seen orders of a few
    // bases), so the order is memoised per sequence
    uint64_t key = (uint64_t)n;
    for (int i = 0; i < n; i++) key |= (uint64_t)(uint8_t)inserted[i] << (8 * (i + 1));
    struct Order {
        char c[8];
        int n;
    };
    thread_local std::unordered_map<uint64_t, Order> memo;
    auto it = memo.find(key);
    if (it == memo.end()) {
        std::vector<char> o = char_order_uncached(std::vector<char>(inserted, inserted + n));
        Order v{};
        v.n = (int)o.size();
        for (int i = 0; i < v.n; i++) v.c[i] = o[(size_t)i];
        it = memo.emplace(key, v).first;
    }
    for (int i = 0; i < it->second.n; i++) out[i] = it->second.c[i];
    return it->second.n;
}

HVar dense_to_hvar(const DenseBase& d) {
    HVar v;
    v.cnt = d.cnt;
    v.fwd = d.fwd;
    v.rev = d.rev;
    v.hicnt = d.hi;
    v.locnt = d.lo;
    v.pmean = (double)d.sum_tp;
    v.qmean = (double)d.sum_q;
    v.mqmean = (double)d.sum_mq;
    v.nm = (double)d.sum_nm;
    v.pstd = (d.flags & 2u) != 0;
    v.qstd = (d.flags & 4u) != 0;
    return v;
}

namespace {
// Splits elements into `T` position ranges holding about equal numbers of them (cut points from a
// sample of their positions): out holds range 0's elements, then range 1's, ..., each range in
// input order; seg (T + 1 entries) gives where each range starts.  No position spans two ranges.
template <class E, class F>
void range_scatter(const std::vector<E>& in, F pos_of, unsigned T, std::vector<E>& out, std::vector<size_t>& seg) {
    const size_t n = in.size();
    seg.assign((size_t)T + 1, 0);
    if (T <= 1) {
        out = in;
        seg[1] = n;
        return;
    }
    std::vector<int> smp;
    const size_t step = std::max<size_t>(1, n / (64 * (size_t)T));
    for (size_t i = 0; i < n; i += step) smp.push_back(pos_of(in[i]));
    std::sort(smp.begin(), smp.end());
    std::vector<int> cut(T - 1);
    for (unsigned t = 1; t < T; t++) cut[t - 1] = smp[smp.size() * t / T];
    std::vector<uint8_t> bk(n);
    for (size_t i = 0; i < n; i++) {
        const int b = (int)(std::upper_bound(cut.begin(), cut.end(), pos_of(in[i])) - cut.begin());
        bk[i] = (uint8_t)b;
        seg[(size_t)b + 1]++;
    }
    for (unsigned t = 0; t < T; t++) seg[t + 1] += seg[t];
    std::vector<size_t> at(seg.begin(), seg.end() - 1);
    out.resize(n);
    for (size_t i = 0; i < n; i++) out[at[bk[i]]++] = in[i];
}

unsigned reduce_threads(size_t n) {
    return n < 8192 ? 1u : std::min(8u, std::max(1u, std::thread::hardware_concurrency() / 4));
}

template <class M>
void move_nodes(M& src, M& dst) {  // src's keys all follow dst's
    while (!src.empty()) dst.insert(dst.end(), src.extract(src.begin()));
}
}  // namespace

void reduce_events(const Event* ev, size_t n, const char* arena, double goodq, HostPileup& P) {
    (void)goodq;
    auto cls = [&](const Event& e) { return e.kind == 1 ? 1 : 0; };
    auto cmp = [&](size_t a, size_t b) {
        const Event& x = ev[a];
        const Event& y = ev[b];
        if (cls(x) != cls(y)) return cls(x) < cls(y);
        if (x.pos != y.pos) return x.pos < y.pos;
        int l = std::min(x.desc_len, y.desc_len);
        int c = memcmp(arena + x.desc_off, arena + y.desc_off, (size_t)l);
        if (c != 0) return c < 0;
        if (x.desc_len != y.desc_len) return x.desc_len < y.desc_len;
        if (x.ord != y.ord) return x.ord < y.ord;
        return x.seq < y.seq;
    };
    // A group (map, position, descriptor) never spans positions: position ranges are sorted on
    // separate threads, then reduced in range order on this one (the maps' nodes are allocated
    // by one thread, as the realigner and the release threads prefer).
    const unsigned T = reduce_threads(n);
    std::vector<size_t> all(n), idx, seg;
    std::iota(all.begin(), all.end(), 0);
    range_scatter(all, [&](size_t i) { return ev[i].pos; }, T, idx, seg);
    all = std::vector<size_t>();
    {
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < T; t++)
            pool.emplace_back([&, t] { std::sort(idx.begin() + (long)seg[t], idx.begin() + (long)seg[t + 1], cmp); });
        std::sort(idx.begin() + (long)seg[0], idx.begin() + (long)seg[1], cmp);
        for (auto& th : pool) th.join();
    }
    {
        auto& nd = P.nd;
        auto& ins = P.ins;
        auto& del = P.delCount;
        auto& insc = P.insCount;
        auto& mnp = P.mnp;
        const size_t hi = n;
        size_t i = 0;
        while (i < hi) {
            const Event& h = ev[idx[i]];
            std::string desc(arena + h.desc_off, h.desc_len);
            int c = cls(h);
            size_t j = i;
            HVar v;
            bool created = false;
            int ndel = 0, nins = 0, nmnp = 0;
            for (; j < hi; j++) {
                const Event& e = ev[idx[j]];
                if (cls(e) != c || e.pos != h.pos || e.desc_len != h.desc_len ||
                    memcmp(arena + e.desc_off, arena + h.desc_off, h.desc_len) != 0)
                    break;
                if (e.kind == 3) {  // anchor subtraction against an existing entry
                    if (!created) continue;
                    v.cnt--;
                    if (e.dir)
                        v.rev--;
                    else
                        v.fwd--;
                    v.pmean -= e.tp;
                    v.qmean -= e.q;
                    v.mqmean -= e.mapq;
                    v.nm -= e.nmv;
                    if (e.hiq)
                        v.hicnt--;
                    else
                        v.locnt--;
                    continue;
                }
                created = true;
                if (e.kind == 2) {  // addCnt form
                    v.cnt++;
                    if (e.dir)
                        v.rev++;
                    else
                        v.fwd++;
                    v.pmean += e.rp;
                    v.qmean += e.q;
                    v.mqmean += e.mapq;
                    v.nm += e.nmv;
                    if (e.hiq)
                        v.hicnt++;
                    else
                        v.locnt++;
                    continue;
                }
                // match-run / deletion / insertion form (pp/pq running comparison)
                if (e.dir)
                    v.rev++;
                else
                    v.fwd++;
                v.cnt++;
                if (!v.pstd && v.pp != 0 && e.tp != v.pp) v.pstd = true;
                if (!v.qstd && v.pq != 0 && e.q != v.pq) v.qstd = true;
                v.pmean += e.tp;
                v.qmean += e.q;
                v.mqmean += e.mapq;
                v.pp = e.tp;
                v.pq = e.q;
                v.nm += e.nmv;
                if (e.hiq)
                    v.hicnt++;
                else
                    v.locnt++;
                ndel += e.cntflags & 1;
                nins += (e.cntflags >> 1) & 1;
                nmnp += (e.cntflags >> 2) & 1;
            }
            // the counters grow once per counted record; the group's total is added at its end
            if (ndel) del[h.pos][desc] += ndel;
            if (nins) insc[h.pos][desc] += nins;
            if (nmnp) mnp[h.pos][desc] += nmnp;
            if (created) {
                if (c == 1)
                    ins[h.pos][desc] = v;
                else
                    nd[h.pos][desc] = v;
            }
            i = j;
        }
    }
}

void reduce_clips(const ClipRec* cl, size_t n, const char* clip_arena, double goodq, HostPileup& P, int side) {
    // Records are reduced per position in read order.  Sort keys: (side, position, read ordinal),
    // packed in 64 bits (a read has at most one clip per side, so keys are distinct).
    using Key = std::pair<uint64_t, uint32_t>;
    std::vector<Key> all;
    all.reserve(n);
    for (size_t i = 0; i < n; i++)
        if (side < 0 || cl[i].side == side)
            all.push_back({((uint64_t)(cl[i].side == 5) << 63) | ((uint64_t)((uint32_t)cl[i].pos ^ 0x80000000u) << 31) |
                               (uint64_t)(uint32_t)cl[i].ord,
                           (uint32_t)i});
    const unsigned T = reduce_threads(all.size());
    // position ranges, each sorted and accumulated on its own thread into maps of its own, whose
    // nodes then move into the pileup's maps in order
    std::vector<Key> keys;
    std::vector<size_t> cut;
    range_scatter(all, [](const Key& k) { return (int)((uint32_t)(k.first >> 31) ^ 0x80000000u); }, T, keys, cut);
    all = std::vector<Key>();
    const double gq = (double)(float)goodq;
    auto add = [&](HVar& v, bool dir, int rp, double q, int mapq, int nm) {
        v.cnt++;
        if (dir)
            v.rev++;
        else
            v.fwd++;
        v.pmean += rp;
        v.qmean += q;
        v.mqmean += mapq;
        v.nm += nm;
        if (q >= gq)
            v.hicnt++;
        else
            v.locnt++;
    };
    struct OrdKey {
        int cnt, pos;
        HSclip* sc;
    };
    std::vector<std::map<int, HSclip>> part5(T), part3(T);
    std::vector<std::vector<OrdKey>> ord5(T), ord3(T);
    auto run = [&](unsigned t) {
        std::sort(keys.begin() + (long)cut[t], keys.begin() + (long)cut[t + 1]);
        for (size_t q = cut[t]; q < cut[t + 1]; q++) {
            const ClipRec& r = cl[keys[q].second];
            auto& mp = r.side == 5 ? part5[t] : part3[t];
            auto it = mp.empty() ? mp.end() : std::prev(mp.end());
            if (it == mp.end() || it->first != r.pos) it = mp.emplace_hint(mp.end(), r.pos, HSclip{});
            HSclip& sc = it->second;
            const char* bases = clip_arena + r.bases_off;
            const char* quals = bases + r.nhq;
            if ((int)sc.col.size() < r.nhq) sc.col.resize((size_t)r.nhq);
            for (int i = 0; i < r.nhq; i++) {
                const char ch = bases[i];
                ClipColumn& c = sc.col[(size_t)i];
                bool found = false;
                for (auto& pr : c.nt)
                    if (pr.first == ch) {
                        pr.second++;
                        found = true;
                    }
                if (!found) c.nt.push_back({ch, 1});
                const int rp = r.side == 5 ? r.nhq - 1 - i : r.nhq - i;
                add(c.seq_at(ch), r.dir != 0, rp, (double)(uint8_t)quals[i], r.mapq, r.nm);
            }
            add(sc.v, r.dir != 0, r.cel, r.sumq / (double)r.nhq, r.mapq, r.nm);
        }
        // the clusters are complete: their consensus, here on this range's thread rather than
        // one by one in the realigner, and their order keys (the nodes keep their addresses when
        // they move)
        for (int s5 = 0; s5 < 2; s5++) {
            auto& mp = s5 ? part5[t] : part3[t];
            auto& ol = s5 ? ord5[t] : ord3[t];
            ol.reserve(mp.size());
            for (auto& e : mp) {
                e.second.pre = clip_consensus(e.second, e.second.pre_used);
                e.second.pre_valid = true;
                ol.push_back({e.second.v.cnt, e.first, &e.second});
            }
        }
    };
    {
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < T; t++) pool.emplace_back(run, t);
        run(0);
        for (auto& th : pool) th.join();
    }
    for (unsigned t = 0; t < T; t++) {
        move_nodes(part5[t], P.sc5);
        move_nodes(part3[t], P.sc3);
    }
    for (int s5 = 0; s5 < 2; s5++) {
        if (side >= 0 && side != (s5 ? 5 : 3)) continue;
        auto& m = s5 ? P.sc5 : P.sc3;
        auto& ord = s5 ? P.sc5_order : P.sc3_order;
        // count descending, position ascending (clip_order's order), from the keys
        size_t tot = 0;
        for (auto& ol : s5 ? ord5 : ord3) tot += ol.size();
        if (tot != m.size()) {  // the map held clusters before this call
            clip_order(m, ord);
        } else {
            std::vector<OrdKey> ok;
            ok.reserve(tot);
            for (auto& ol : s5 ? ord5 : ord3) ok.insert(ok.end(), ol.begin(), ol.end());
            std::sort(ok.begin(), ok.end(), [](const OrdKey& a, const OrdKey& b) {
                if (a.cnt != b.cnt) return a.cnt > b.cnt;
                return a.pos < b.pos;
            });
            ord.clear();
            ord.reserve(tot);
            for (auto& x : ok) ord.push_back({x.pos, x.sc});
        }
        (s5 ? P.sc5_of : P.sc3_of) = &m;
    }
}

void clip_order(std::map<int, HSclip>& m, std::vector<std::pair<int, HSclip*>>& ord) {
    ord.clear();
    ord.reserve(m.size());
    for (auto& e : m) ord.push_back({e.first, &e.second});
    std::sort(ord.begin(), ord.end(), [](const std::pair<int, HSclip*>& a, const std::pair<int, HSclip*>& b) {
        if (a.second->v.cnt != b.second->v.cnt) return a.second->v.cnt > b.second->v.cnt;
        return a.first < b.first;
    });
}

bool low_complexity(const std::string& s) {
    int len = (int)s.size();
    if (len == 0) return true;
    int ntcnt = 0;
    const char order[4] = {'A', 'T', 'G', 'C'};
    for (char b : order) {
        int n = (int)std::count(s.begin(), s.end(), b);
        if (n > 0) ntcnt++;
        if (n / (double)len > 0.75) return true;
    }
    return ntcnt < 3;
}

// findconseq's computation (include/VariationUtils.h:403-509) without its side effects: the
// caller stores the sequence and, when mark_used is set, marks the cluster used
std::string clip_consensus(const HSclip& sc, bool& mark_used) {
    mark_used = false;
    int total = 0, match = 0;
    std::string seqq;
    bool flag = false;
    for (size_t idx = 0; idx < sc.col.size(); idx++) {
        const ClipColumn& col = sc.col[idx];
        char firsts[8], order[8];
        std::vector<char> big;
        int nk = (int)col.nt.size(), no;
        const char* ord;
        if (nk <= 7) {
            for (int k = 0; k < nk; k++) firsts[k] = col.nt[(size_t)k].first;
            no = rh_char_order_small(firsts, nk, order);
            ord = order;
        } else {
            for (auto& pr : col.nt) big.push_back(pr.first);
            big = rh_char_iteration_order(big);
            no = (int)big.size();
            ord = big.data();
        }
        int maxCount = 0;
        double maxQuality = 0;
        char chosen = 0;
        int tt = 0;
        for (int oi = 0; oi < no; oi++) {
            const char base = ord[oi];
            int count = 0;
            for (auto& pr : col.nt)
                if (pr.first == base) count = pr.second;
            tt += count;
            const HVar* sv = col.seq_find(base);
            const bool has = sv != nullptr;
            const double bq = has ? sv->qmean : 0;
            if (count > maxCount || (has && bq > maxQuality)) {
                maxCount = count;
                chosen = base;
                maxQuality = has ? bq : 0.0;  // seq[i][base] would be created empty
            }
        }
        if (idx == 3 && sc.col.size() >= 6 && tt / (double)sc.v.cnt < 0.2 && tt <= 2) break;
        if ((tt - maxCount > 2 || maxCount <= tt - maxCount) && maxCount / (double)tt < 0.8) {
            if (flag) break;
            flag = true;
        }
        total += tt;
        match += maxCount;
        if (chosen != 0) seqq += chosen;
    }
    std::string SEQ;
    size_t ntSize = sc.col.size();
    size_t sl = seqq.size();
    if (total != 0 && match / (double)total > 0.9 && sl / 1.5 > (double)(ntSize - sl) &&
        (sl / (double)ntSize > 0.8 || ntSize - sl < 10 || sl > 25)) {
        SEQ = seqq;
    } else {
        SEQ = " ";
    }
    if (!SEQ.empty() && SEQ.size() > 12) {
        bool mm1 = SEQ.size() >= 8 && SEQ.compare(1, 7, "AAAAAAA") == 0;
        bool mm2 = SEQ.size() >= 8 && SEQ.compare(1, 7, "TTTTTTT") == 0;
        if (mm1 || mm2) mark_used = true;
        if (low_complexity(SEQ)) mark_used = true;
    }
    return SEQ;
}


std::string dump_pileup(const HostPileup& P) {
    std::string o;
    char buf[640];
    auto line = [&](const char* tag, int p, const std::string& d, const HVar& v) {
        snprintf(buf, sizeof buf, "%s\t%d\t%s\t%d\t%d\t%d\t%d\t%d\t%.17g\t%.17g\t%.17g\t%.17g\t%d\t%d\t%d\n", tag, p, d.c_str(),
                 v.cnt, v.fwd, v.rev, v.hicnt, v.locnt, v.pmean, v.qmean, v.mqmean, v.nm, (int)v.pstd, (int)v.qstd,
                 v.extracnt);
        o += buf;
    };
    // nonInsertionVariants: positions with any entry (dense or not)
    std::map<int, std::map<std::string, HVar>> all;
    for (int i = 0; i < P.dense_n; i++) {
        const DenseRow& r = P.rows[(size_t)i];
        for (int b = 0; b < 4; b++)
            if (r.b[b].flags & 1u) all[P.dense_lo + i][std::string(1, index_base(b))] = dense_to_hvar(r.b[b]);
    }
    for (auto& e : P.nd)
        for (auto& kv : e.second) all[e.first][kv.first] = kv.second;
    for (auto& e : all)
        for (auto& kv : e.second) line("V", e.first, kv.first, kv.second);
    for (auto& e : P.ins)
        for (auto& kv : e.second) line("I", e.first, kv.first, kv.second);
    for (int i = 0; i < P.dense_n; i++)
        if (P.rows[(size_t)i].cov != 0) {
            snprintf(buf, sizeof buf, "C\t%d\t%d\n", P.dense_lo + i, P.rows[(size_t)i].cov);
            o += buf;
        }
    auto dump_sc = [&](const char* tag, const std::map<int, HSclip>& m) {
        for (auto& e : m) {
            const HSclip& s = e.second;
            snprintf(buf, sizeof buf, "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%.17g\t%.17g\t%.17g\t%.17g\t%zu\n", tag, e.first, s.v.cnt,
                     s.v.fwd, s.v.rev, s.v.hicnt, s.v.locnt, s.v.pmean, s.v.qmean, s.v.mqmean, s.v.nm, s.col.size());
            o += buf;
            for (size_t ci = 0; ci < s.col.size(); ci++) {
                const auto& nt = s.col[ci].nt;
                o += std::string(tag) + "n\t" + std::to_string(e.first) + "\t" + std::to_string(ci);
                std::vector<char> ins;
                for (auto& pr : nt) ins.push_back(pr.first);
                for (char b : rh_char_iteration_order(ins)) {
                    int cnt = 0;
                    for (auto& pr : nt)
                        if (pr.first == b) cnt = pr.second;
                    o += "\t";
                    o += b;
                    o += ":" + std::to_string(cnt);
                }
                o += "\n";
            }
        }
    };
    dump_sc("S5", P.sc5);
    dump_sc("S3", P.sc3);
    for (auto& e : P.delCount)
        for (auto& d : e.second) o += "DC\t" + std::to_string(e.first) + "\t" + d.first + "\t" + std::to_string(d.second) + "\n";
    for (auto& e : P.insCount)
        for (auto& d : e.second) o += "IC\t" + std::to_string(e.first) + "\t" + d.first + "\t" + std::to_string(d.second) + "\n";
    for (auto& e : P.mnp)
        for (auto& d : e.second) o += "MC\t" + std::to_string(e.first) + "\t" + d.first + "\t" + std::to_string(d.second) + "\n";
    o += "L\t" + std::to_string(P.maxReadLength) + "\n";
    return o;
}

}  // namespace rv
