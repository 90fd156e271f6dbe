// Local realignment (host side), restating src/VariationRealigner.cpp over the engine's position
// store:
//   process / stage order               :135-161, :467-482
//   adjustMNP                           :334-465
//   realigndel                          :489-778
//   realignins                          :785-1114
//   realignlgins30                      :1119-1354
//   fillAndSortTmp / CMP_tmp / COMP3    :1362-1386, :44-58, :29-34
//   find35match                         :1396-1428
//   noPassingReads                      :1447-1498
//   ismatch / ismatchref                :1507-1539, :1987-2024
//   adjRefCnt / adjRefFactor            :1753-1822
//   findMM5 / findMM3                   :1857-1976
// with include/VariationUtils.h (adjCnt :291-332, correctCnt :263-280, findconseq :403-509,
// islowcomplexseq :367-394, joinRef :336-349) and include/util.h (vc_substr :27-57, charAt :171-179).
// A candidate whose processing throws (a reference key outside the loaded window, an out-of-range
// substring) is abandoned with the changes made so far, as the reference's per-candidate catch does.
#include "rv_realign.h"
#include "rv_desc.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <stdexcept>
#include <thread>
#include <unordered_map>

namespace rv {

// ------------------------------------------------------------------ rows
const DenseRow* HostRowSource::fetch(const std::vector<int>& rows) {
    blocks_.emplace_back(rows.size(), DenseRow{});
    std::vector<DenseRow>& out = blocks_.back();
    for (size_t k = 0; k < rows.size(); k++) {
        int r = rows[k];
        if (r >= 0 && r < P_.dense_n && (size_t)r < P_.rows.size()) out[k] = P_.rows[(size_t)r];
    }
    return out.data();
}

bool RowCache::load(std::vector<int>& rows, bool sorted_unique) {
    if (!sorted_unique) {
        std::sort(rows.begin(), rows.end());
        rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    }
    if (rows.empty()) return true;
    const auto t0 = std::chrono::steady_clock::now();
    const DenseRow* buf = src_->fetch(rows);
    fetch_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!buf) return false;
    fetches++;
    rows_fetched += (int64_t)rows.size();
    if (slots_->size() < (size_t)n_) slots_->resize((size_t)n_, nullptr);
    for (size_t k = 0; k < rows.size(); k++) {
        (*slots_)[(size_t)rows[k]] = buf + k;
        set_.push_back(rows[k]);
    }
    return true;
}

void RowCache::release_slots() {
    for (int i : set_) (*slots_)[(size_t)i] = nullptr;
    set_.clear();
}

bool RowCache::prefetch(const std::vector<int>& positions, int radius, const std::vector<int>& exact) {
    // the wanted rows as a bitmap over the window: the list comes out sorted and unique
    if (n_ <= 0) return true;
    std::vector<uint64_t> bits(((size_t)n_ + 63) / 64, 0);
    auto mark = [&](int i) {
        if (!have(i)) bits[(size_t)i >> 6] |= 1ull << (i & 63);
    };
    for (int p : positions) {
        const int a = std::max(0, p - radius - lo_), b = std::min(n_ - 1, p + radius - lo_);
        for (int i = a; i <= b; i++) mark(i);
    }
    for (int p : exact) {
        const int i = p - lo_;
        if (i >= 0 && i < n_) mark(i);
    }
    std::vector<int> want;
    for (size_t w = 0; w < bits.size(); w++)
        for (uint64_t x = bits[w]; x; x &= x - 1) want.push_back((int)(w * 64 + (size_t)__builtin_ctzll(x)));
    return load(want, true);
}

const DenseRow* RowCache::row(int pos) {
    int i = pos - lo_;
    if (i < 0 || i >= n_) return nullptr;
    if (!have(i)) {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<int> want;
        for (int k = std::max(0, i - MISS_RADIUS); k <= std::min(n_ - 1, i + MISS_RADIUS); k++)
            if (!have(k)) want.push_back(k);
        if (!load(want)) throw std::runtime_error("dense row fetch failed");
        miss_fetches++;
        miss_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return (*slots_)[(size_t)i];
}

// ------------------------------------------------------------------ reference
char HostRef::at(int k) const {
    if (k >= lo && k <= hi) return seq[k - lo];
    if (created.count(k)) return '\0';
    throw std::out_of_range("reference key");
}
char HostRef::get(int k) {
    if (k >= lo && k <= hi) return seq[k - lo];
    created.insert(k);
    return '\0';
}

// ------------------------------------------------------------------ position store
namespace {
std::vector<RealignCand> sorted_candidates(const std::map<int, std::map<std::string, int>>& m);
}

RState::RState(HostPileup& P_, RowCache& rc) : P(P_), rows(rc) {
    // the three position indexes are independent: the clip maps' on their own threads, and
    // adjustMNP's candidate order (it reads only the MNV map) on a third
    std::thread t5([this] { sc5_ix.build(P.sc5, 5); });
    std::thread t3([this] { sc3_ix.build(P.sc3, 5); });
    std::thread tm([this] {
        mnp_sorted = sorted_candidates(P.mnp);
        mnp_sorted_ready = true;  // seen after the join below
    });
    struct Join {
        std::thread &a, &b, &c;
        ~Join() {
            a.join();
            b.join();
            c.join();
        }
    } join{t5, t3, tm};
    nd_ix.build(P.nd, 5);
    for (auto& e : P.ins)
        for (auto& kv : e.second) {
            HVar* v = new_var();
            *v = kv.second;
            ins[e.first][kv.first] = v;
        }
}

HVar* RState::new_var() {
    pool.emplace_back();
    return &pool.back();
}

RPos& RState::at(int p) {
    RPos& r = pos[p];
    if (r.loaded) {
        if (track_on_ && !track_n_.count(p)) {
            Snap& sn = track_n_[p];
            sn.key = r.key;
            sn.cov = r.cov;
            for (auto& kv : r.v) sn.v.emplace_back(kv.first, *kv.second);
        }
        return r;
    }
    r.loaded = true;
    struct AfterLoad {
        RState& s;
        int p;
        RPos& r;
        ~AfterLoad() {
            if (!s.track_on_ || s.track_n_.count(p)) return;
            Snap& sn = s.track_n_[p];
            sn.key = r.key;
            sn.cov = r.cov;
            for (auto& kv : r.v) sn.v.emplace_back(kv.first, *kv.second);
        }
    } after{*this, p, r};
    if (const DenseRow* d = rows.row(p)) {
        r.key = (d->rowflags & 1u) != 0;
        r.cov = d->cov;
        for (int b = 0; b < 4; b++)
            if (d->b[b].flags & 1u) {
                HVar* v = new_var();
                *v = dense_to_hvar(d->b[b]);
                r.v[std::string(1, index_base(b))] = v;
            }
    }
    if (const auto* nd = nd_ix.find(p)) {
        r.key = true;
        for (auto& kv : *nd) {
            HVar* v = new_var();
            *v = kv.second;
            r.v[kv.first] = v;
        }
    }
    return r;
}

HVar* RState::inner(int p, const std::string& d) {
    RPos& r = at(p);
    if (!r.key) return nullptr;
    auto it = r.v.find(d);
    if (it == r.v.end()) return nullptr;
    r.dirty = true;  // the caller may change it
    return it->second;
}

HVar* RState::get_variation(int p, const std::string& d) {
    RPos& r = at(p);
    r.dirty = true;
    if (!r.key) {
        r.key = true;
        r.v.clear();
        key_log.push_back({p, true});
    }
    HVar*& v = r.v[d];
    if (!v) v = new_var();
    return v;
}

HVar* RState::get_ins(int p, const std::string& d) {
    track_ins(p);
    HVar*& v = ins[p][d];
    if (!v) v = new_var();
    return v;
}

const HVar* RState::peek(int p, const std::string& d) {
    auto it = pos.find(p);
    if (it != pos.end() && it->second.loaded) {
        const RPos& r = it->second;
        if (!r.key) return nullptr;
        auto j = r.v.find(d);
        return j == r.v.end() ? nullptr : j->second;
    }
    // what at() would load: the dense row's key flag and bases, then the non-dense entries
    const DenseRow* dr = rows.row(p);
    const auto* nd = nd_ix.find(p);
    const bool key = (dr && (dr->rowflags & 1u)) || nd != nullptr;
    if (!key) return nullptr;
    if (d.size() == 1 && dr) {
        const int b = base_index(d[0]);
        if (b >= 0 && (dr->b[b].flags & 1u)) {
            peek_tmp_ = dense_to_hvar(dr->b[b]);
            return &peek_tmp_;
        }
    }
    if (nd) {
        auto j = nd->find(d);
        if (j != nd->end()) return &j->second;
    }
    return nullptr;
}

const HVar* RState::peek_loaded(int p, const std::string& d, bool* hit) const {
    auto it = pos.find(p);
    *hit = it != pos.end() && it->second.loaded;
    if (!*hit || !it->second.key) return nullptr;
    auto j = it->second.v.find(d);
    return j == it->second.v.end() ? nullptr : j->second;
}

int RState::peek_static(int p, const std::string& d, HVar* out) const {
    const DenseRow* dr;
    if (!rows.row_loaded(p, dr)) return 2;
    const auto* nd = nd_ix.find(p);
    const bool key = (dr && (dr->rowflags & 1u)) || nd != nullptr;
    if (!key) return 0;
    if (d.size() == 1 && dr) {
        const int b = base_index(d[0]);
        if (b >= 0 && (dr->b[b].flags & 1u)) {
            *out = dense_to_hvar(dr->b[b]);
            return 1;
        }
    }
    if (nd) {
        auto j = nd->find(d);
        if (j != nd->end()) {
            *out = j->second;
            return 1;
        }
    }
    return 0;
}

HVar* RState::maybe(int p, char refBase) {
    if (refBase == 0) return nullptr;
    return inner(p, std::string(1, refBase));
}

void RState::erase_inner(int p, const std::string& d, bool drop_empty) {
    RPos& r = at(p);
    if (!r.key) return;
    r.dirty = true;
    r.v.erase(d);
    if (drop_empty && r.v.empty()) {
        r.key = false;
        key_log.push_back({p, false});
    }
}

void RState::add_cov(int p, int k) {
    RPos& r = at(p);
    r.dirty = true;
    r.cov += k;
}

namespace {
bool same_hvar(const HVar& a, const HVar& b) {
    return a.cnt == b.cnt && a.fwd == b.fwd && a.rev == b.rev && a.hicnt == b.hicnt && a.locnt == b.locnt &&
           a.extracnt == b.extracnt && a.pmean == b.pmean && a.qmean == b.qmean && a.mqmean == b.mqmean &&
           a.nm == b.nm && a.pstd == b.pstd && a.qstd == b.qstd && a.pp == b.pp && a.pq == b.pq;
}
bool same_entries(const std::vector<std::pair<std::string, HVar>>& a, const std::map<std::string, HVar*>& b) {
    if (a.size() != b.size()) return false;
    size_t i = 0;
    for (auto& kv : b) {
        if (a[i].first != kv.first || !same_hvar(a[i].second, *kv.second)) return false;
        i++;
    }
    return true;
}
}  // namespace

void RState::track_begin() {
    track_on_ = true;
    track_n_.clear();
    track_i_.clear();
}

void RState::track_ins(int p) {
    if (!track_on_ || track_i_.count(p)) return;
    Snap& sn = track_i_[p];
    auto it = ins.find(p);
    sn.present = it != ins.end();
    if (sn.present)
        for (auto& kv : it->second) sn.v.emplace_back(kv.first, *kv.second);
}

bool RState::track_end() {
    track_on_ = false;
    bool changed = false;
    for (auto& e : track_n_) {
        const RPos& r = pos[e.first];
        if (r.key != e.second.key || r.cov != e.second.cov || !same_entries(e.second.v, r.v)) {
            changed = true;
            break;
        }
    }
    if (!changed)
        for (auto& e : track_i_) {
            auto it = ins.find(e.first);
            const bool present = it != ins.end();
            if (present != e.second.present || (present && !same_entries(e.second.v, it->second))) {
                changed = true;
                break;
            }
        }
    track_n_.clear();
    track_i_.clear();
    return changed;
}

// ------------------------------------------------------------------ string helpers (util.h)
namespace {

std::string sub1(const std::string& s, int idx) {
    if (idx >= 0) return s.substr((size_t)idx);
    return s.substr((size_t)(s.size() + (size_t)(int64_t)idx));
}
std::string sub2(const std::string& s, int begin, int len) {
    if (begin < 0) begin = (int)(s.size() + (size_t)(int64_t)begin);
    if (len > 0) return s.substr((size_t)begin, (size_t)len);
    if (len == 0) return "";
    len = (int)(s.size() + (size_t)(int64_t)len - (size_t)(int64_t)begin);
    if (len < 0) return "";
    return s.substr((size_t)begin, (size_t)len);
}
char char_at(const std::string& s, int index) {
    if (index < 0) {
        int i = (int)s.size() + index;
        if (i < 0) return (char)-1;
        return s[(size_t)i];
    }
    return s[(size_t)index];
}
std::string join_ref(const HostRef& ref, int from, int to) {
    std::string r;
    for (int i = from; i <= to; i++) r += ref.at(i);
    return r;
}
std::string join_ref_d(const HostRef& ref, int from, double to) {
    std::string r;
    for (int i = from; i < to; i++) r += ref.at(i);
    return r;
}
void replace_first(std::string& s, const std::string& a, const std::string& b) {
    size_t p = s.find(a);
    if (p != std::string::npos) s.replace(p, a.size(), b);
}
bool starts(const std::string& s, char c) { return !s.empty() && s[0] == c; }

// ------------------------------------------------------------------ count transfer
void correct_cnt(HVar* v) {
    if (v->cnt < 0) v->cnt = 0;
    if (v->hicnt < 0) v->hicnt = 0;
    if (v->locnt < 0) v->locnt = 0;
    if (v->pmean < 0) v->pmean = 0;
    if (v->qmean < 0) v->qmean = 0;
    if (v->mqmean < 0) v->mqmean = 0;
    if (v->rev < 0) v->rev = 0;
    if (v->fwd < 0) v->fwd = 0;
}
void adj_cnt(HVar* to, const HVar* v, HVar* refv) {
    to->cnt += v->cnt;
    to->extracnt += v->cnt;
    to->hicnt += v->hicnt;
    to->locnt += v->locnt;
    to->pmean += v->pmean;
    to->qmean += v->qmean;
    to->mqmean += v->mqmean;
    to->nm += v->nm;
    to->pstd = true;
    to->qstd = true;
    to->rev += v->rev;
    to->fwd += v->fwd;
    if (!refv) return;
    refv->cnt -= v->cnt;
    refv->hicnt -= v->hicnt;
    refv->locnt -= v->locnt;
    refv->pmean -= v->pmean;
    refv->qmean -= v->qmean;
    refv->mqmean -= v->mqmean;
    refv->nm -= v->nm;
    refv->rev -= v->rev;
    refv->fwd -= v->fwd;
    correct_cnt(refv);
}
void adj_ref_cnt(const HVar* tv, HVar* refv, int len) {
    if (!refv) return;
    double f = tv->pmean != 0 ? (tv->pmean / (double)tv->cnt - len + 1) / (tv->pmean / (double)tv->cnt) : 0;
    if (f < 0) return;
    if (f > 1) f = 1;
    refv->cnt -= (int)(f * tv->cnt);
    refv->hicnt -= (int)(f * tv->hicnt);
    refv->locnt -= (int)(f * tv->locnt);
    refv->pmean -= f * tv->pmean;
    refv->qmean -= f * tv->qmean;
    refv->mqmean -= f * tv->mqmean;
    refv->nm -= f * tv->nm;
    refv->rev -= (int)(f * tv->rev);
    refv->fwd -= (int)(f * tv->fwd);
    correct_cnt(refv);
}
void adj_ref_factor(HVar* refv, double f) {
    if (!refv) return;
    if (f > 1) f = 1;
    if (f < -1) return;
    int old = refv->cnt;
    refv->cnt -= (int)(f * refv->cnt);
    refv->hicnt -= (int)(f * refv->hicnt);
    refv->locnt -= (int)(f * refv->locnt);
    double fc = old != 0 ? std::abs(refv->cnt - old) / (double)old : 1;
    refv->pmean -= refv->pmean * f * fc;
    refv->qmean -= refv->qmean * f * fc;
    refv->mqmean -= refv->mqmean * f * fc;
    refv->nm -= f * refv->nm;
    refv->fwd -= (int)(f * refv->fwd);
    refv->rev -= (int)(f * refv->rev);
    correct_cnt(refv);
}


}  // namespace

std::string findconseq(HSclip& sc) {
    if (!sc.sequence.empty()) return sc.sequence;
    bool mark = false;
    std::string SEQ;
    if (sc.pre_valid) {  // computed when the clips were reduced
        SEQ.swap(sc.pre);
        sc.pre_valid = false;
        mark = sc.pre_used;
    } else {
        SEQ = clip_consensus(sc, mark);
    }
    if (mark) sc.used = true;
    sc.sequence = SEQ;
    return SEQ;
}

const std::string& consensus(HSclip& sc) {
    if (sc.sequence.empty()) findconseq(sc);
    return sc.sequence;
}

void find35match(const std::string& seq5, const std::string& seq3, int& b5, int& b3, int& score) {
    // The reference tries every start (i, j), i in [0, |seq5| - 9], j in [1, |seq3| - 9] (i outer),
    // and walks n = 0, 1, ... comparing seq5[i + n] with seq3[|seq3| - j - n] (seq5's end counts as
    // a mismatch) until the third mismatch or either end, and returns the first start whose run has
    // more than 8 matches, under 10% mismatches, and reaches an end.  Every start on one
    // anti-diagonal (seq5 index + seq3 index = D) walks the same cells, so each diagonal's
    // mismatches are listed once and each start is evaluated in O(1); the smallest qualifying
    // (i, j) is the reference's answer.
    b5 = b3 = score = 0;
    const int l5 = (int)seq5.size(), l3 = (int)seq3.size();
    if (l5 <= 8 || l3 <= 9) return;
    int bi = INT32_MAX, bj = INT32_MAX, bs = 0;
    std::vector<int> M;
    M.reserve((size_t)l5 + 1);
    // start (i, j) lies on D = i + l3 - j: i <= l5 - 9 and j >= 1 give D <= l5 - 9 + l3 - 1, j <= l3 - 9
    // gives D >= i + 9 >= 9
    for (int D = 9; D <= l5 - 9 + l3 - 1; D++) {
        const int i_lo = std::max(0, D - l3 + 1), i_hi = std::min(l5 - 9, D - 9);
        if (i_lo > i_hi || i_lo > bi) continue;
        const int a_hi = std::min(D, l5);  // last cell: seq3 index 0 or seq5's end
        M.clear();
        for (int a = i_lo; a <= a_hi; a++)
            if (a >= l5 || seq3[(size_t)(D - a)] != seq5[(size_t)a]) M.push_back(a);
        size_t k = 0;
        for (int i = i_lo; i <= i_hi && i <= bi; i++) {
            while (k < M.size() && M[k] < i) k++;
            int n, nm;
            if (k + 2 < M.size()) {
                n = M[k + 2] - i;
                nm = 3;
            } else {
                n = a_hi - i + 1;
                nm = (int)(M.size() - k);
            }
            const int j = i + l3 - D;
            if (n - nm > 8 && nm / (double)n < 0.1 && (n + j >= l3 || i + n >= l5)) {
                if (i < bi || (i == bi && j < bj)) {
                    bi = i;
                    bj = j;
                    bs = n - nm;
                }
                break;  // later starts on this diagonal have larger i
            }
        }
    }
    if (bi != INT32_MAX) {
        b5 = bi;
        b3 = bj;
        score = bs;
    }
}

namespace {

using Cand = RealignCand;
// CMP_tmp: count descending, position ascending, descriptor descending
std::vector<Cand> sorted_candidates(const std::map<int, std::map<std::string, int>>& m) {
    std::vector<Cand> t;
    for (auto& e : m)
        for (auto& v : e.second) t.push_back({e.first, v.first, v.second});
    std::sort(t.begin(), t.end(), [](const Cand& a, const Cand& b) {
        if (a.count != b.count) return a.count > b.count;
        if (a.position != b.position) return a.position < b.position;
        return a.desc.compare(b.desc) > 0;
    });
    return t;
}

struct Mismatch {
    std::string seq;
    int pos;
    int end;
};
struct MM {
    std::vector<Mismatch> mm;
    std::vector<int> scp;
    int nm = 0;
    int misp = 0;
    std::string misnt;
};

std::string strip_marks(std::string s) {
    s.erase(std::remove_if(s.begin(), s.end(), [](char ch) { return ch == '#' || ch == '^'; }), s.end());
    return s;
}


// findMM5 / findMM3: the mismatches next to a candidate's flanks, and the soft-clip positions to
// try.  The soft clips found are marked used in `mark`; with mark == nullptr (a probe for the
// positions a candidate will read) nothing is marked, the clip maps are not read, and no
// reference key is created.
MM findMM5(HostRef& rf, int position, const std::string& wupseq, HostPileup* mark) {
    const bool probe = mark == nullptr;
    std::string seq = strip_marks(wupseq);
    MM r;
    int n = 0, mn = 0, mcnt = 0;
    std::string str;
    auto ne = [&](char ch, int k) { return rf.count(k) && rf.at(k) != ch; };
    auto eq = [&](char ch, int k) { return rf.count(k) && rf.at(k) == ch; };
    while (ne(char_at(seq, -1 - n), position - n) && mcnt < 3) {
        str.insert(str.begin(), char_at(seq, -1 - n));
        r.mm.push_back({str, position - n, 5});
        n++;
        mcnt++;
    }
    r.scp.push_back(position + 1);
    char misnt = 0;
    if (str.size() == 1) {
        while (eq(char_at(seq, -1 - n), position - n)) {
            n++;
            if (n != 0) mn++;
        }
        if (mn > 1) {
            int n2 = 0;
            while (-1 - n - 1 - n2 >= 0 && eq(char_at(seq, -1 - n - 1 - n2), position - n - 1 - n2)) n2++;
            if (n2 > 2) {
                r.scp.push_back(position - n - n2);
                r.misp = position - n;
                misnt = char_at(seq, -1 - n);
                if (!probe) {
                    auto s = mark->sc5.find(position - n - n2);
                    if (s != mark->sc5.end()) s->second.used = true;
                }
                mn += n2;
            } else {
                r.scp.push_back(position - n);
                if (!probe) {
                    auto s = mark->sc5.find(position - n);
                    if (s != mark->sc5.end()) s->second.used = true;
                }
            }
        }
    }
    r.nm = mn;
    r.misnt = misnt == 0 ? "" : std::string(1, misnt);
    return r;
}

MM findMM3(HostRef& rf, int p, const std::string& sanpseq, HostPileup* mark) {
    const bool probe = mark == nullptr;
    // rf.get creates a missing reference key holding '\0'; a probe reads the same value
    auto get = [&](int k) { return probe ? (rf.count(k) ? rf.at(k) : '\0') : rf.get(k); };
    std::string seq = strip_marks(sanpseq);
    MM r;
    int n = 0, mn = 0, mcnt = 0;
    std::string str;
    while (n < (int)seq.size() && rf.count(p + n) && get(p + n) == seq[(size_t)n]) n++;
    r.scp.push_back(p + n);
    int Tbp = p + n;
    while (mcnt <= 3 && n < (int)seq.size() && get(p + n) != seq[(size_t)n]) {
        str += seq[(size_t)n];
        r.mm.push_back({str, Tbp, 3});
        n++;
        mcnt++;
    }
    char misnt = 0;
    auto eq = [&](char ch, int k) { return rf.count(k) && rf.at(k) == ch; };
    if (str.size() == 1) {
        while (n < (int)seq.size() && eq(seq[(size_t)n], p + n)) {
            n++;
            if (n != 0) mn++;
        }
        if (mn > 1) {
            int n2 = 0;
            while (n + n2 + 1 < (int)seq.size() && eq(seq[(size_t)(n + n2 + 1)], p + n + 1 + n2)) n2++;
            if (n2 > 2 && n + n2 + 1 < (int)seq.size()) {
                r.scp.push_back(p + n + n2);
                r.misp = p + n;
                misnt = seq[(size_t)n];
                if (!probe) {
                    auto s = mark->sc3.find(p + n + n2);
                    if (s != mark->sc3.end()) s->second.used = true;
                }
                mn += n2;
            } else {
                r.scp.push_back(p + n);
                if (!probe) {
                    auto s = mark->sc3.find(p + n);
                    if (s != mark->sc3.end()) s->second.used = true;
                }
            }
        }
    }
    r.nm = mn;
    r.misnt = misnt == 0 ? "" : std::string(1, misnt);
    return r;
}

// the descriptor parse and flanking sequences of a deletion candidate (realigndel's set-up)
struct DelSetup {
    int dellen = 0, el = 0;
    std::string wupseq, sanpseq;
};
DelSetup del_setup(HostRef& ref, int chrLen, int p, const std::string& vn) {
    DelSetup d;
    std::string num, rest;
    if (desc::begin_minus_number(vn, num)) d.dellen = std::atoi(num.c_str());
    if (desc::up_number_end(vn, num)) d.dellen += std::atoi(num.c_str());
    std::string extrains, extra, inv5, inv3;
    if (desc::minus_sv(vn, inv5, inv3)) {
    } else if (desc::begin_minus_number_any(vn, rest)) {
        extra = desc::strip_marks3(rest);
        desc::caret_atgnc(vn, extrains);
    }
    int wustart = (p - 200) > 1 ? (p - 200) : 1;
    d.wupseq = join_ref(ref, wustart, p - 1) + extra;
    if (!inv3.empty()) d.wupseq = inv3;
    int sanend = p + 200 > chrLen ? chrLen : p + 200;
    d.el = (int)extra.size() - (int)extrains.size();
    d.sanpseq = extra + join_ref(ref, p + d.dellen + d.el, sanend);
    if (!inv5.empty()) d.sanpseq = inv5;
    return d;
}

// Positions realigndel will read for its candidates' mismatches (found from the reference and
// the descriptors alone), so that they can be fetched with the first batch of rows.
void del_probe_positions(HostRef& ref, int chrLen, const std::map<int, std::map<std::string, int>>& counts,
                         std::vector<int>& out) {
    for (auto& e : counts)
        for (auto& kv : e.second) {
            try {
                const int p = e.first;
                const DelSetup d = del_setup(ref, chrLen, p, kv.first);
                const MM r3 = findMM3(ref, p, d.sanpseq, nullptr);
                const MM r5 = findMM5(ref, p + d.dellen + d.el - 1, d.wupseq, nullptr);
                for (const MM* r : {&r3, &r5}) {
                    for (const Mismatch& m : r->mm) out.push_back(m.pos);
                    if (r->misp != 0) out.push_back(r->misp);
                }
            } catch (const std::exception&) {
            }
        }
}

struct Realigner {
    const Config& c;
    int rstart, rend;
    HostRef& ref;
    int chrLen;
    ReadQuery* reads;
    RState& S;
    RealignStats& st;
    ScreenPlan* plan;

    int maxlen() const { return S.P.maxReadLength; }

    // index of a 5' cluster among the screen's cluster candidates (those are listed by position)
    int lgins_index(int position) const {
        const ScreenCand* b = plan->cand.data() + plan->off[SC_LGINS];
        const ScreenCand* e = plan->cand.data() + plan->off[SC_LGINS + 1];
        const ScreenCand* it =
            std::lower_bound(b, e, position, [](const ScreenCand& x, int v) { return x.position < v; });
        return (it != e && it->position == position) ? (int)(it - b) : -1;
    }

    // ---- screen: candidate i of a kind is skipped when the plan says so; with labels on, a
    // visit's changes are recorded when it ends (every exit of the loop body)
    bool skip(int kind, int i) {
        if (!plan || plan->keep.empty() || i < 0) return false;
        if (plan->keep[(size_t)(plan->off[kind] + i)]) return false;
        plan->skipped++;
        return true;
    }
    struct Visit {
        Realigner& rg;
        int slot;  // -1: not labelled
        Visit(Realigner& r, int kind, int i) : rg(r), slot(-1) {
            if (rg.plan && rg.plan->label && i >= 0) {
                slot = rg.plan->off[kind] + i;
                rg.S.track_begin();
            }
        }
        ~Visit() {
            if (slot >= 0) rg.plan->changed[(size_t)slot] = rg.S.track_end() ? 1 : 0;
        }
    };

    bool ismatch(const std::string& s1, const std::string& s2_in, int dir, int MMx = 3) {
        const bool marked = s2_in.find_first_of("#^") != std::string::npos;
        std::string stripped;
        if (marked) stripped = strip_marks(s2_in);
        const std::string& s2 = marked ? stripped : s2_in;
        int mm = 0;
        const int l2 = (int)s2.size();
        for (int n = 0; n < (int)s1.size() && n < l2; n++) {
            // vc_substr(s2, b, 1)[0] with b = n (forward) or -n-1 (from the end)
            const int b = dir * n - (dir == -1 ? 1 : 0);
            const int at = b < 0 ? l2 + b : b;
            if (s1[(size_t)n] != s2[(size_t)at]) mm++;
        }
        return mm <= MMx && mm / (double)s1.size() < 0.15;
    }
    bool ismatchref(const std::string& s, HostRef& rf, int position, int dir, int MMx = 3) {
        int mm = 0;
        for (int n = 0; n < (int)s.size(); n++) {
            if (!rf.count(position + dir * n)) return false;
            if (char_at(s, dir == 1 ? n : dir * n - 1) != rf.get(position + dir * n)) mm++;
        }
        return mm <= MMx && mm / (double)s.size() < 0.15;
    }

    HSclip* sc5(int p) { return S.sc5_ix.find(p); }
    HSclip* sc3(int p) {
        return S.sc3_ix.find(p);
    }

    // noPassingReads (VariationRealigner.cpp:1447-1498): no record of the region spans [start, end]
    // without carrying exactly this deletion
    bool noPassingReads(int start, int end) {
        const auto t0 = std::chrono::steady_clock::now();
        struct Acc {
            RealignStats& st;
            std::chrono::steady_clock::time_point t0;
            ~Acc() {
                st.n_nopass++;
                st.ms_nopass += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
        } acc{st, t0};
        if (!reads) throw std::runtime_error("region reads unavailable");
        return reads->no_passing(start, end);
    }

    // Every read adjustMNP makes of a position that at() has not loaded comes from the dense rows
    // and the non-dense map, which the realigner never changes.  Those reads (the candidate's
    // descriptor, then each split's left and right piece) and the candidates' clip clusters are
    // resolved on several threads before the ordered walk, which then reads them in order
    // instead of from hash slots and rows scattered over the region.
    struct MnpPre {
        std::vector<size_t> off;      // candidate k's reads: [off[k], off[k+1])
        std::vector<uint8_t> state;   // peek_static's result per read
        std::vector<HVar> v;
        std::vector<HSclip*> s3, s5;  // sc3(position), sc5(position + length)
    };
    void mnp_pre(const std::vector<Cand>& tmp, MnpPre& pre) {
        const size_t n = tmp.size();
        pre.off.assign(n + 1, 0);
        for (size_t k = 0; k < n; k++) {
            const std::string& d = tmp[k].desc;
            const size_t len = d.size() - (d.find('&') != std::string::npos ? 1 : 0);
            pre.off[k + 1] = pre.off[k] + 1 + 2 * (len > 0 ? len - 1 : 0);
        }
        pre.state.resize(pre.off[n]);
        pre.v.resize(pre.off[n]);
        pre.s3.resize(n);
        pre.s5.resize(n);
        auto work = [&](size_t a, size_t b) {
            for (size_t k = a; k < b; k++) {
                const int position = tmp[k].position;
                uint8_t* st = pre.state.data() + pre.off[k];
                HVar* v = pre.v.data() + pre.off[k];
                const size_t nq = pre.off[k + 1] - pre.off[k];
                std::fill(st, st + nq, (uint8_t)2);  // unknown until resolved
                std::string mnt = tmp[k].desc;
                replace_first(mnt, "&", "");
                pre.s3[k] = sc3(position);
                pre.s5[k] = sc5(position + (int)mnt.size());
                try {
                    st[0] = (uint8_t)S.peek_static(position, tmp[k].desc, &v[0]);
                    for (int i = 0; i < (int)mnt.size() - 1; i++) {
                        std::string left = sub2(mnt, 0, i + 1);
                        if (left.size() > 1) left.insert(1, "&");
                        std::string right = sub1(mnt, -(int)(mnt.size() - (size_t)i - 1));
                        if (right.size() > 1) right.insert(1, "&");
                        st[1 + 2 * i] = (uint8_t)S.peek_static(position, left, &v[1 + 2 * i]);
                        st[2 + 2 * i] = (uint8_t)S.peek_static(position + i + 1, right, &v[2 + 2 * i]);
                    }
                } catch (const std::exception&) {  // the ordered walk meets the same error itself
                }
            }
        };
        const size_t nt = n < 2048 ? 1 : std::min<size_t>(8, n / 1024);
        std::vector<std::thread> th;
        for (size_t t = 1; t < nt; t++) th.emplace_back(work, n * t / nt, n * (t + 1) / nt);
        work(0, n / nt);
        for (auto& x : th) x.join();
    }

    void adjustMNP() {
        std::vector<Cand> tmp = S.mnp_sorted_ready ? std::move(S.mnp_sorted) : sorted_candidates(S.P.mnp);
        S.mnp_sorted_ready = false;
        st.candidates += (int64_t)tmp.size();
        MnpPre pre;
        mnp_pre(tmp, pre);
        int ci = -1;
        for (const Cand& t : tmp) {
            ci++;
            if (ci + 8 < (int)tmp.size()) {
                if (const HSclip* f = pre.s3[(size_t)ci + 8]) __builtin_prefetch(f);
                if (const HSclip* f = pre.s5[(size_t)ci + 8]) __builtin_prefetch(f);
            }
            if (skip(SC_MNP, ci)) continue;
            Visit visit(*this, SC_MNP, ci);
            // peek() through the resolved reads: the position's entries once at() has loaded it
            const size_t q0 = pre.off[(size_t)ci];
            auto look = [&](int p, const std::string& d, size_t q) -> const HVar* {
                bool hit;
                const HVar* r = S.peek_loaded(p, d, &hit);
                if (hit) return r;
                if (pre.state[q] == 2) return S.peek(p, d);
                return pre.state[q] ? &pre.v[q] : nullptr;
            };
            try {
                int position = t.position;
                const std::string vn = t.desc;
                // The reference walks a copy of the position's map and erases the left pieces it
                // used from the copy only (VariationRealigner.cpp:350-400).  The position's own map
                // does not change during the walk (the right pieces sit at later positions), so
                // the walk reads it directly and remembers the erased keys.  Positions are read
                // through peek() and materialized only when an entry is taken over: the output
                // does not depend on it (the candidate's position is a non-dense one, which the
                // record list always carries).
                const HVar* v0 = look(position, vn, q0);
                if (!v0) continue;  // no key at the position, or no such descriptor
                HVar vcur = *v0;    // vref's value until the position is materialized
                HVar* vref = nullptr;
                auto take = [&]() {
                    if (vref) return;
                    RPos& rp = S.at(position);
                    rp.dirty = true;
                    vref = rp.v.at(vn);
                };
                auto vcnt = [&]() { return vref ? vref->cnt : vcur.cnt; };
                std::string mnt = vn;
                replace_first(mnt, "&", "");
                std::vector<std::string> used_left;
                for (int i = 0; i < (int)mnt.size() - 1; i++) {
                    std::string left = sub2(mnt, 0, i + 1);
                    if (left.size() > 1) left.insert(1, "&");
                    std::string right = sub1(mnt, -(int)(mnt.size() - (size_t)i - 1));
                    if (right.size() > 1) right.insert(1, "&");
                    const HVar* lt = look(position, left, q0 + 1 + 2 * (size_t)i);
                    if (lt && std::find(used_left.begin(), used_left.end(), left) == used_left.end()) {
                        if (lt->cnt <= 0) continue;
                        if (lt->cnt < vcnt() && lt->pmean / lt->cnt <= i + 1) {
                            take();
                            adj_cnt(vref, S.at(position).v.at(left), nullptr);
                            used_left.push_back(left);
                        }
                    }
                    const HVar* rt = look(position + i + 1, right, q0 + 2 + 2 * (size_t)i);
                    if (rt) {
                        if (rt->cnt < 0) continue;
                        if (rt->cnt < vcnt()) {
                            take();
                            RPos& rr = S.at(position + i + 1);
                            HVar* tref = rr.v.at(right);
                            adj_cnt(vref, tref, nullptr);
                            S.add_cov(position, tref->cnt);
                            rr.dirty = true;
                            rr.v.erase(right);
                        }
                    }
                }
                if (HSclip* s3 = pre.s3[(size_t)ci]) {
                    if (!s3->used) {
                        std::string seq = findconseq(*s3);
                        if (sub2(seq, 0, (int)mnt.size()) == mnt) {
                            if (seq.size() == mnt.size() || ismatchref(seq.substr(mnt.size()), ref, position + (int)mnt.size(), 1)) {
                                RPos& r0 = S.at(position);
                                HVar*& tgt = r0.v[vn];
                                if (!tgt) tgt = S.new_var();
                                r0.dirty = true;
                                adj_cnt(tgt, &s3->v, nullptr);
                                S.add_cov(position, s3->v.cnt);
                                s3->used = true;
                            }
                        }
                    }
                }
                if (HSclip* s5 = pre.s5[(size_t)ci]) {
                    if (!s5->used) {
                        std::string seq = findconseq(*s5);
                        if (seq != " " && seq.size() >= mnt.size()) {
                            std::reverse(seq.begin(), seq.end());
                            if (sub2(seq, (int)(seq.size() - mnt.size()), (int)mnt.size()) == mnt) {
                                if (seq.size() == mnt.size() ||
                                    ismatchref(seq.substr(0, seq.size() - mnt.size()), ref, position - 1, -1)) {
                                    RPos& r0 = S.at(position);
                                    HVar*& tgt = r0.v[vn];
                                    if (!tgt) tgt = S.new_var();
                                    r0.dirty = true;
                                    adj_cnt(tgt, &s5->v, nullptr);
                                    S.add_cov(position, s5->v.cnt);
                                    s5->used = true;
                                }
                            }
                        }
                    }
                }
            } catch (const std::exception&) {
            }
        }
    }

    void realigndel(bool withBams, const std::map<int, std::map<std::string, int>>& counts) {
        std::vector<Cand> tmp = sorted_candidates(counts);
        st.candidates += (int64_t)tmp.size();
        int ci = -1;
        for (const Cand& t : tmp) {
            ci++;
            if (skip(SC_DEL, ci)) continue;
            Visit visit(*this, SC_DEL, ci);
            try {
                int p = t.position;
                std::string vn = t.desc;
                int dcnt = t.count;
                HVar* vref = S.get_variation(p, vn);
                const DelSetup d = del_setup(ref, chrLen, p, vn);
                const int dellen = d.dellen, el = d.el;
                const std::string& wupseq = d.wupseq;
                const std::string& sanpseq = d.sanpseq;
                MM r3 = findMM3(ref, p, sanpseq, &S.P);
                MM r5 = findMM5(ref, p + dellen + el - 1, wupseq, &S.P);
                std::vector<Mismatch> mmm = r3.mm;
                mmm.insert(mmm.end(), r5.mm.begin(), r5.mm.end());
                for (const Mismatch& mis : mmm) {
                    std::string mm = mis.seq;
                    int mp = mis.pos, me = mis.end;
                    if (mm.size() > 1) mm.insert(1, 1, '&');
                    HVar* tv = S.inner(mp, mm);
                    if (!tv) continue;
                    if (tv->cnt == 0) continue;
                    if (tv->qmean / tv->cnt < c.goodq) continue;
                    if (tv->pmean / tv->cnt > (me == 3 ? r3.nm + 4 : r5.nm + 4)) continue;
                    if (tv->cnt >= dcnt + dellen || tv->cnt / dcnt >= 8) continue;
                    if (mp > p && me == 5) {
                        double f = tv->pmean != 0 ? (mp - p) / (tv->pmean / (double)tv->cnt) : 1;
                        if (f > 1) f = 1;
                        S.add_cov(p, (int)(tv->cnt * f));
                        adj_ref_cnt(tv, S.maybe(p, ref.get(p)), dellen);
                    }
                    HVar* lref = nullptr;
                    if (mp > p && me == 3 && S.has_key(p)) lref = S.inner(p, std::string(1, ref.get(p)));
                    adj_cnt(vref, tv, lref);
                    S.erase_inner(mp, mm, true);
                }
                if (r3.misp != 0 && r3.mm.size() == 1) {
                    HVar* v = S.inner(r3.misp, r3.misnt);
                    if (v && v->cnt < dcnt) S.erase_inner(r3.misp, r3.misnt, false);
                }
                if (r5.misp != 0 && r5.mm.size() == 1) {
                    HVar* v = S.inner(r5.misp, r5.misnt);
                    if (v && v->cnt < dcnt) S.erase_inner(r5.misp, r5.misnt, false);
                }
                for (int sc5pp : r5.scp) {
                    HSclip* tv = sc5(sc5pp);
                    if (tv && !tv->used) {
                        std::string seq = findconseq(*tv);
                        if (dcnt <= 2 && tv->v.cnt / dcnt > 5) continue;
                        if (seq != " " && ismatch(seq, wupseq, -1)) {
                            if (sc5pp > p) S.add_cov(p, tv->v.cnt);
                            adj_cnt(vref, &tv->v, nullptr);
                            tv->used = true;
                        }
                    }
                }
                for (int sc3pp : r3.scp) {
                    HSclip* tv = sc3(sc3pp);
                    if (tv && !tv->used) {
                        std::string seq = findconseq(*tv);
                        if (dcnt <= 2 && tv->v.cnt / dcnt > 5) continue;
                        if (seq != " " && ismatch(seq, sub1(sanpseq, sc3pp - p), 1)) {
                            if (sc3pp <= p) S.add_cov(p, tv->v.cnt);
                            HVar* lref = sc3pp <= p ? nullptr : S.maybe(p, ref.get(p));
                            adj_cnt(vref, &tv->v, lref);
                            tv->used = true;
                        }
                    }
                }
                int pe = p + dellen + el;
                HVar* h = S.maybe(p, ref.get(p));
                if (withBams && pe - p >= 5 && pe - p < maxlen() - 10 && h && h->cnt != 0 &&
                    vref->cnt > 2 * h->cnt * (1 - (pe - p) / (double)maxlen()) && noPassingReads(p, pe)) {
                    adj_cnt(vref, h, h);
                }
            } catch (const std::exception&) {
            }
        }
        for (int i = (int)tmp.size() - 1; i > 0; i--) {
            const Cand& t = tmp[(size_t)i];
            int p = t.position;
            const std::string& vn = t.desc;
            RPos& rp = S.at(p);
            if (!rp.key) continue;
            if (!rp.v.count(vn)) continue;
            HVar* vref = rp.v[vn];
            std::string tn;
            if (desc::minus_number_amp_atgcs_end(vn, tn)) {
                if (rp.v.count(tn)) {
                    HVar* tref = rp.v[tn];
                    if (vref->cnt < tref->cnt) {
                        rp.dirty = true;
                        adj_cnt(tref, vref, nullptr);
                        rp.v.erase(vn);
                    }
                }
            }
        }
    }

    void realignins(const std::map<int, std::map<std::string, int>>& counts) {
        std::vector<Cand> tmp = sorted_candidates(counts);
        st.candidates += (int64_t)tmp.size();
        int ci = -1;
        for (const Cand& t : tmp) {
            ci++;
            if (skip(SC_INS, ci)) continue;
            Visit visit(*this, SC_INS, ci);
            try {
                int position = t.position;
                std::string vn = t.desc;
                int icnt = t.count;
                std::string insert1;
                if (!desc::begin_plus_atgc(vn, insert1)) continue;
                std::string ins3, dupn;
                int inslen = (int)insert1.size();
                if (desc::dup_num_atgc(vn, dupn, ins3)) inslen += std::atoi(dupn.c_str()) + (int)ins3.size();
                std::string extra, compm, newins, num;
                desc::amp_atgc(vn, extra);
                desc::hash_atgc(vn, compm);
                desc::caret_atgc_end(vn, newins);
                int newdel = 0;
                if (desc::up_number_end(vn, num)) newdel = std::atoi(num.c_str());
                std::string tn = vn;
                desc::drop_lead_plus(tn);
                replace_first(tn, "&", "");
                replace_first(tn, "#", "");
                desc::drop_caret_digits_end(tn);
                desc::drop_first_caret(tn);
                int wustart = position - 150 > 1 ? (position - 150) : 1;
                std::string wupseq = join_ref(ref, wustart, position) + tn;
                int sanend = position + (int)vn.size() + 100;
                if (chrLen < sanend) sanend = chrLen;
                std::string sanpseq;
                MM r3;
                if (!ins3.empty()) {
                    int p3 = position + inslen - (int)ins3.size() + 50;
                    if (ins3.size() > 50) sanpseq = sub1(ins3, 50 - (int)ins3.size());
                    sanpseq += join_ref(ref, position + 1, position + 101);
                    r3 = findMM3(ref, p3 + 1, sanpseq, &S.P);
                } else {
                    sanpseq = tn + join_ref(ref, position + (int)extra.size() + 1 + (int)compm.size() + newdel, sanend);
                    r3 = findMM3(ref, position + 1, sanpseq, &S.P);
                }
                MM r5 = findMM5(ref, position + (int)extra.size() + (int)compm.size() + newdel, wupseq, &S.P);
                std::vector<Mismatch> mmm = r3.mm;
                mmm.insert(mmm.end(), r5.mm.begin(), r5.mm.end());
                HVar* vref = S.get_ins(position, vn);
                for (const Mismatch& mis : mmm) {
                    std::string mb = mis.seq;
                    int mp = mis.pos, me = mis.end;
                    if (mb.size() > 1) mb = std::string(1, mb[0]) + "&" + mb.substr(1);
                    HVar* v = S.inner(mp, mb);
                    if (!v) continue;
                    if (v->cnt == 0) continue;
                    if (v->qmean / v->cnt < c.goodq) continue;
                    if (v->pmean / v->cnt > (me == 3 ? r3.nm + 4 : r5.nm + 4)) continue;
                    if (v->cnt >= icnt + (int)insert1.size() || v->cnt / icnt >= 8) continue;
                    if (mp > position && me == 5) S.add_cov(position, v->cnt);
                    HVar* lref = nullptr;
                    if (mp > position && me == 3 && S.has_key(position) && ref.count(position))
                        lref = S.inner(position, std::string(1, ref.get(position)));
                    adj_cnt(vref, v, lref);
                    S.erase_inner(mp, mb, true);
                }
                if (r3.misp != 0 && r3.mm.size() == 1) {
                    HVar* v = S.inner(r3.misp, r3.misnt);
                    if (v && v->cnt < icnt) S.erase_inner(r3.misp, r3.misnt, false);
                }
                if (r5.misp != 0 && r5.mm.size() == 1) {
                    HVar* v = S.inner(r5.misp, r5.misnt);
                    if (v && v->cnt < icnt) S.erase_inner(r5.misp, r5.misnt, false);
                }
                for (int sc5pp : r5.scp) {
                    HSclip* tv = sc5(sc5pp);
                    if (!tv) continue;
                    if (!tv->used) {
                        std::string seq = findconseq(*tv);
                        if (seq != " " && ismatch(seq, wupseq, -1)) {
                            if (sc5pp > position) S.add_cov(position, tv->v.cnt);
                            adj_cnt(vref, &tv->v, nullptr);
                            tv->used = true;
                        }
                    }
                }
                for (int sc3pp : r3.scp) {
                    HSclip* tv = sc3(sc3pp);
                    if (!tv) continue;
                    if (!tv->used) {
                        std::string seq = findconseq(*tv);
                        std::string mseq = !ins3.empty() ? sanpseq : sub1(sanpseq, sc3pp - position - 1);
                        if (seq != " " && ismatch(seq, mseq, 1)) {
                            if (sc3pp <= position || (double)insert1.size() > tv->v.pmean / tv->v.cnt)
                                S.add_cov(position, tv->v.cnt);
                            HVar* lref = nullptr;
                            if (sc3pp > position && S.has_key(position) && ref.count(position))
                                lref = S.inner(position, std::string(1, ref.get(position)));
                            if ((double)insert1.size() > tv->v.pmean / tv->v.cnt) lref = nullptr;
                            adj_cnt(vref, &tv->v, lref);
                            tv->used = true;
                            if (insert1.size() + 1 == vn.size() && (int)insert1.size() > maxlen() &&
                                sc3pp >= position + 1 + (int)insert1.size()) {
                                int flag = 0;
                                int off = (sc3pp - position - 1) % (int)insert1.size();
                                std::string tvn = vn;
                                for (int si = 0; si < (int)seq.size() && si + off < (int)insert1.size(); si++) {
                                    if (sub2(seq, si, 1) != sub2(insert1, si + off, 1)) {
                                        flag++;
                                        int shift = si + off + 1;
                                        tvn = tvn.substr(0, (size_t)shift) + sub2(seq, si, 1) + tvn.substr((size_t)shift + 1);
                                    }
                                }
                                if (flag > 0) {
                                    S.track_ins(position);
                                    auto& im = S.ins[position];
                                    HVar* vv = im[vn];
                                    im[tvn] = vv;
                                    im.erase(vn);
                                }
                            }
                        }
                    }
                }
                int first3 = r3.scp[0];
                int first5 = r5.scp[0];
                if (!r3.scp.empty() && !r5.scp.empty() && first3 > first5 + 3 && first3 - first5 < maxlen() * 0.75) {
                    if (ref.count(position) && S.has_key(position)) {
                        HVar* rv = S.inner(position, std::string(1, ref.get(position)));
                        if (rv) adj_ref_factor(rv, (first3 - first5 - 1) / (double)maxlen());
                    }
                    adj_ref_factor(vref, -(first3 - first5 - 1) / (double)maxlen());
                }
            } catch (const std::exception&) {
            }
        }
        for (int i = (int)tmp.size() - 1; i > 0; i--) {
            try {
                const Cand& t = tmp[(size_t)i];
                int p = t.position;
                const std::string& vn = t.desc;
                S.track_ins(p);
                auto mi = S.ins.find(p);
                if (mi == S.ins.end()) continue;
                auto& m = mi->second;
                if (!m.count(vn)) continue;
                HVar* vref = m.at(vn);
                std::string tn;
                if (desc::plus_atgcs_amp_atgcs_end(vn, tn)) {
                    if (m.count(tn)) {
                        HVar* tref = m.at(tn);
                        if (vref->cnt < tref->cnt) {
                            adj_cnt(tref, vref, S.maybe(p, ref.get(p)));
                            m.erase(vn);
                        }
                    }
                }
            } catch (const std::exception&) {
            }
        }
    }

    void realignlgins30() {
        HostRef rf = ref;  // the reference copies the whole map (VariationRealigner.cpp:1120)
        struct SP {
            int position;
            HSclip* sc;
            int count;
            int li;  // index among the screen's cluster candidates (P.sc5 order)
        };
        std::vector<SP> t5, t3;
        // the clusters in COMP3 order (count descending, position ascending), as the reduction
        // sorted them; a pileup copied since then sorts again
        if (S.P.sc5_of != &S.P.sc5) {
            clip_order(S.P.sc5, S.P.sc5_order);
            S.P.sc5_of = &S.P.sc5;
        }
        if (S.P.sc3_of != &S.P.sc3) {
            clip_order(S.P.sc3, S.P.sc3_order);
            S.P.sc3_of = &S.P.sc3;
        }
        t5.reserve(S.P.sc5_order.size());
        for (const auto& e : S.P.sc5_order) {
            if (e.first < rstart - 5000 || e.first > rend + 5000) continue;
            // screened mode: a skipped cluster is dropped here; the kept ones keep their order
            const int li = plan ? lgins_index(e.first) : -1;
            if (skip(SC_LGINS, li)) continue;
            t5.push_back({e.first, e.second, e.second->v.cnt, li});
        }
        if (t5.empty()) return;  // the 3' clusters are only read from a 5' cluster's visit
        t3.reserve(S.P.sc3_order.size());
        for (const auto& e : S.P.sc3_order)
            if (!(e.first < rstart - 5000 || e.first > rend + 5000)) t3.push_back({e.first, e.second, e.second->v.cnt, -1});
        st.candidates += (int64_t)(t5.size() + t3.size());
        // The reference scans every 3' cluster for every 5' cluster and skips those outside
        // [p5 - 2.5 * maxReadLength, p5 + maxReadLength - 10] before anything else; visiting only
        // the clusters inside that window, in the same (COMP3) order, is the same computation.
        std::vector<std::pair<int, int>> by_pos;  // (position, rank in t3)
        by_pos.reserve(t3.size());
        for (size_t i = 0; i < t3.size(); i++) by_pos.push_back({t3[i].position, (int)i});
        std::sort(by_pos.begin(), by_pos.end());
        std::vector<int> ranks;
        auto window = [&](int p5, std::vector<int>& out) {
            out.clear();
            const double lo = p5 - maxlen() * 2.5;
            auto it = std::lower_bound(by_pos.begin(), by_pos.end(), std::make_pair((int)std::floor(lo) - 1, -1));
            for (; it != by_pos.end() && it->first - p5 <= maxlen() - 10; ++it)
                if (!(p5 - it->first > maxlen() * 2.5)) out.push_back(it->second);
            std::sort(out.begin(), out.end());
        };
        // Phase A: every pair the ordered pass below can score (both consensus longer than 10, the
        // distance window and the count ratio; the `used` marks are left to that pass) is scored
        // as one batch on the device.  The consensus strings are read without findconseq's side
        // effects, which the ordered pass applies as the reference does.
        std::unordered_map<uint64_t, int> f35_of;
        std::vector<F35> f35_res;
        if (reads) {
            auto peek_cons = [](const HSclip& sc) -> const std::string* {
                if (!sc.sequence.empty()) return &sc.sequence;
                if (sc.pre_valid) return &sc.pre;
                return nullptr;
            };
            std::vector<F35Job> jobs;
            for (size_t ai = 0; ai < t5.size(); ai++) {
                const std::string* s5 = peek_cons(*t5[ai].sc);
                if (!s5 || s5->size() <= 10) continue;
                window(t5[ai].position, ranks);
                for (int ri : ranks) {
                    const SP& b = t3[(size_t)ri];
                    const std::string* s3 = peek_cons(*b.sc);
                    if (!s3 || s3->size() <= 10) continue;
                    const int cnt5 = t5[ai].count, cnt3 = b.count;
                    if (!(cnt5 / (double)cnt3 >= 0.08 && cnt5 / (double)cnt3 <= 12)) continue;
                    f35_of[((uint64_t)ai << 32) | (uint32_t)ri] = (int)jobs.size();
                    jobs.push_back({s5, s3});
                }
            }
            if (jobs.size() < 8 || !reads->find35_batch(jobs, f35_res) || f35_res.size() != jobs.size()) {
                f35_of.clear();
                f35_res.clear();
            }
        }
        for (size_t ai = 0; ai < t5.size(); ai++) {
            const SP& a = t5[ai];
            Visit visit(*this, SC_LGINS, a.li);
            int p5 = a.position;
            HSclip* sc5v = a.sc;
            int cnt5 = a.count;
            if (sc5v->used) continue;
            const std::string& seq5 = consensus(*sc5v);
            if (seq5.size() <= 10) continue;
            window(p5, ranks);
            for (int ri : ranks) {
                const SP& b = t3[(size_t)ri];
                try {
                    int p3 = b.position;
                    HSclip* sc3v = b.sc;
                    int cnt3 = b.count;
                    if (sc5v->used) break;
                    if (sc3v->used) continue;
                    if (p5 - p3 > maxlen() * 2.5) continue;
                    if (p3 - p5 > maxlen() - 10) continue;
                    const std::string& seq3 = consensus(*sc3v);
                    if (seq3.size() <= 10) continue;
                    if (!(cnt5 / (double)cnt3 >= 0.08 && cnt5 / (double)cnt3 <= 12)) continue;
                    int bp5, bp3, score;
                    auto f = f35_of.find(((uint64_t)ai << 32) | (uint32_t)ri);
                    if (f != f35_of.end() && f35_res[(size_t)f->second].score >= 0) {
                        const F35& r = f35_res[(size_t)f->second];
                        bp5 = r.b5;
                        bp3 = r.b3;
                        score = r.score;
                    } else {
                        find35match(seq5, seq3, bp5, bp3, score);
                    }
                    if (score == 0) continue;
                    int smscore = (int)(score / 2);
                    std::string ins = bp3 + smscore > 1 ? sub2(seq3, 0, -(bp3 + smscore) + 1) : seq3;
                    if (bp5 + smscore > 0) {
                        std::string tmpstr = seq5.substr(0, (size_t)(bp5 + smscore));
                        std::reverse(tmpstr.begin(), tmpstr.end());
                        ins += tmpstr;
                    }
                    if (low_complexity(ins)) continue;
                    int bi = 0;
                    HVar* vref;
                    if (p5 > p3) {
                        if (seq3.size() > ins.size() &&
                            !ismatch(seq3.substr(ins.size()), join_ref(rf, p5, (int)(p5 + seq3.size() - ins.size() + 2)), 1))
                            continue;
                        if (seq5.size() > ins.size() &&
                            !ismatch(seq5.substr(ins.size()), join_ref(rf, (int)(p3 - (seq5.size() - ins.size()) - 2), p3 - 1), -1))
                            continue;
                        std::string tmp = join_ref(rf, p3, p5 - 1);
                        if (tmp.size() > ins.size()) {
                            ins = std::to_string(p3 - p5) + "^" + ins;
                            bi = p3;
                            vref = S.get_variation(p3, ins);
                        } else if (tmp.size() < ins.size()) {
                            ins = sub2(ins, 0, (int)(ins.size() - tmp.size())) + "&" + sub1(ins, p3 - p5);
                            ins = "+" + ins;
                            bi = p3 - 1;
                            vref = S.get_ins(bi, ins);
                        } else {
                            ins = "-" + std::to_string(ins.size()) + "^" + ins;
                            bi = p3;
                            vref = S.get_variation(p3, ins);
                        }
                    } else {
                        if (seq3.size() > ins.size() &&
                            !ismatch(seq3.substr(ins.size()), join_ref(rf, p5, (int)(p5 + seq3.size() - ins.size() + 2)), 1))
                            continue;
                        if (seq5.size() > ins.size() &&
                            !ismatch(seq5.substr(ins.size()), join_ref(rf, (int)(p3 - (seq5.size() - ins.size()) - 2), p3 - 1), -1))
                            continue;
                        std::string tmp;
                        if (ins.size() <= (size_t)(p3 - p5)) {
                            int rpt = 2, tnr = 3;
                            while (((p3 - p5 + ins.size()) / (double)tnr) / (double)ins.size() > 1) {
                                if ((p3 - p5 + ins.size()) % (size_t)tnr == 0) rpt++;
                                tnr++;
                            }
                            tmp += join_ref_d(rf, p5, (p5 + (p3 - p5 + ins.size()) / (double)rpt - ins.size()));
                            ins = "+" + tmp + ins;
                        } else {
                            tmp += join_ref(rf, p5, p3 - 1);
                            if ((ins.size() - tmp.size()) % 2 == 0) {
                                int tex = (int)((ins.size() - tmp.size()) / 2);
                                ins = (tmp + sub2(ins, 0, tex)) == sub1(ins, tex) ? ("+" + sub1(ins, tex)) : "+" + tmp + ins;
                            } else {
                                ins = "+" + tmp + ins;
                            }
                        }
                        bi = p5 - 1;
                        vref = S.get_ins(bi, ins);
                    }
                    sc3v->used = true;
                    sc5v->used = true;
                    vref->pstd = true;
                    vref->qstd = true;
                    S.add_cov(bi, sc5v->v.cnt);
                    if (ins[0] == '+') {
                        HVar* mvref = S.maybe(bi, rf.get(bi));
                        adj_cnt(vref, &sc3v->v, mvref);
                        adj_cnt(vref, &sc5v->v, nullptr);
                        if (p3 - p5 >= 5 && p3 - p5 < maxlen() - 10 && mvref && mvref->cnt != 0 && vref->cnt > 2 * mvref->cnt &&
                            noPassingReads(p5, p3)) {
                            adj_cnt(vref, mvref, mvref);
                        }
                        std::map<int, std::map<std::string, int>> tins;
                        tins[bi][ins] = vref->cnt;
                        realignins(tins);
                    } else if (ins[0] == '-') {
                        adj_cnt(vref, &sc3v->v, S.maybe(bi, rf.get(bi)));
                        adj_cnt(vref, &sc5v->v, nullptr);
                        std::map<int, std::map<std::string, int>> tdel;
                        tdel[bi][ins] = vref->cnt;
                        realigndel(false, tdel);
                    } else {
                        adj_cnt(vref, &sc3v->v, nullptr);
                        adj_cnt(vref, &sc5v->v, nullptr);
                    }
                    break;
                } catch (const std::exception&) {
                }
            }
        }
    }
};

}  // namespace

namespace {

HVarD to_dev(const HVar& v) {
    HVarD d;
    d.cnt = v.cnt;
    d.fwd = v.fwd;
    d.rev = v.rev;
    d.hicnt = v.hicnt;
    d.locnt = v.locnt;
    d.extracnt = v.extracnt;
    d.pmean = v.pmean;
    d.qmean = v.qmean;
    d.mqmean = v.mqmean;
    d.nm = v.nm;
    d.pstd = v.pstd ? 1 : 0;
    d.qstd = v.qstd ? 1 : 0;
    return d;
}

}  // namespace

void build_record_list(RState& S, CxHost& cx) {
    std::vector<int> ps;
    for (auto& e : S.P.nd) ps.push_back(e.first);
    for (auto& e : S.ins) ps.push_back(e.first);
    for (auto& e : S.pos)
        if (e.second.dirty) ps.push_back(e.first);
    std::sort(ps.begin(), ps.end());
    ps.erase(std::unique(ps.begin(), ps.end()), ps.end());
    if (!S.rows.prefetch(ps, 0)) throw std::runtime_error("dense row fetch failed");
    // A position's entries are read, never materialized: one that at() has loaded gives its own;
    // any other gives what at() would load, the dense row's flagged bases and the non-dense
    // map's descriptors merged in the map's (string) order, the non-dense entry taking a key
    // both hold.  Nothing is written to the store, so ranges of positions are listed on several
    // threads and joined in order.
    const size_t n = ps.size();
    const size_t nt = n < 4096 ? 1 : std::min<size_t>(8, n / 2048);
    std::vector<CxHost> part(nt);
    std::vector<char> missing(nt, 0);
    auto work = [&](size_t t) {
        CxHost& o = part[t];
        const size_t a = n * t / nt, b = n * (t + 1) / nt;
        o.pos.reserve(b - a);
        o.cov.reserve(b - a);
        o.key.reserve(b - a);
        o.off.reserve(b - a);
        auto add = [&](int kind, const std::string& d, const HVar& v) {
            CxEntry c;
            c.kind = kind;
            c.desc_off = (uint32_t)o.arena.size();
            c.desc_len = (int32_t)d.size();
            c.pad = 0;
            c.v = to_dev(v);
            o.arena += d;
            o.ent.push_back(c);
        };
        std::string one(1, ' ');
        for (size_t k = a; k < b; k++) {
            const int p = ps[k];
            o.pos.push_back(p);
            auto lp = S.pos.find(p);
            if (lp != S.pos.end() && lp->second.loaded) {
                const RPos& r = lp->second;
                o.cov.push_back(r.cov);
                o.key.push_back(r.key ? 1 : 0);
                if (r.key)
                    for (auto& kv : r.v) add(0, kv.first, *kv.second);
            } else {
                const DenseRow* dr;
                if (!S.rows.row_loaded(p, dr)) {
                    missing[t] = 1;
                    return;
                }
                const auto* nd = S.nd_ix.find(p);
                const bool key = (dr && (dr->rowflags & 1u)) || nd != nullptr;
                o.cov.push_back(dr ? dr->cov : 0);
                o.key.push_back(key ? 1 : 0);
                if (key) {
                    // the flagged bases as single-character keys, in string order
                    char db[4];
                    int nb = 0;
                    if (dr)
                        for (int x = 0; x < 4; x++)
                            if (dr->b[x].flags & 1u) db[nb++] = index_base(x);
                    std::sort(db, db + nb);
                    int i = 0;
                    auto dense_out = [&](char c) {
                        for (int x = 0; x < 4; x++)
                            if (index_base(x) == c) {
                                one[0] = c;
                                add(0, one, dense_to_hvar(dr->b[x]));
                                return;
                            }
                    };
                    if (nd)
                        for (auto& kv : *nd) {
                            for (; i < nb && std::string(1, db[i]) < kv.first; i++) dense_out(db[i]);
                            if (i < nb && kv.first.size() == 1 && kv.first[0] == db[i]) i++;  // the map's
                            add(0, kv.first, kv.second);
                        }
                    for (; i < nb; i++) dense_out(db[i]);
                }
            }
            auto it = S.ins.find(p);
            if (it != S.ins.end())
                for (auto& kv : it->second) add(1, kv.first, *kv.second);
            o.off.push_back((int32_t)o.ent.size());
        }
    };
    {
        std::vector<std::thread> th;
        for (size_t t = 1; t < nt; t++) th.emplace_back(work, t);
        work(0);
        for (auto& x : th) x.join();
    }
    for (char m : missing)
        if (m) throw std::runtime_error("dense row not loaded for the record list");
    size_t ne = 0, na = 0;
    for (auto& o : part) {
        ne += o.ent.size();
        na += o.arena.size();
    }
    cx.pos.reserve(n);
    cx.cov.reserve(n);
    cx.key.reserve(n);
    cx.off.reserve(n + 1);
    cx.ent.reserve(ne);
    cx.arena.reserve(na);
    cx.off.push_back(0);
    for (auto& o : part) {
        const int32_t eb = (int32_t)cx.ent.size();
        const uint32_t ab = (uint32_t)cx.arena.size();
        cx.pos.insert(cx.pos.end(), o.pos.begin(), o.pos.end());
        cx.cov.insert(cx.cov.end(), o.cov.begin(), o.cov.end());
        cx.key.insert(cx.key.end(), o.key.begin(), o.key.end());
        for (int32_t x : o.off) cx.off.push_back(eb + x);
        for (CxEntry c : o.ent) {
            c.desc_off += ab;
            cx.ent.push_back(c);
        }
        cx.arena += o.arena;
    }
}


namespace {

void screen_table(ScreenPlan& plan, int t, const std::vector<std::pair<int, int64_t>>& pc) {
    std::vector<int32_t>& tp = plan.tpos[t];
    std::vector<int64_t>& tc = plan.tcum[t];
    tp.clear();
    tc.assign(1, 0);
    tp.reserve(pc.size());
    tc.reserve(pc.size() + 1);
    for (auto& x : pc) {
        if (!tp.empty() && tp.back() == x.first) {
            tc.back() += x.second;
        } else {
            tp.push_back(x.first);
            tc.push_back(tc.back() + x.second);
        }
    }
}

void screen_counted(ScreenPlan& plan, int kind, int t, const std::map<int, std::map<std::string, int>>& m) {
    std::vector<ScreenCand>& out = plan.lists[kind];
    out.clear();
    for (const Cand& c : sorted_candidates(m)) {
        int len = (int)c.desc.size();
        if (kind == SC_DEL) len = std::max(1, std::atoi(c.desc.c_str() + 1));  // "-N..."
        if (kind == SC_INS) len = std::max(1, len - 1);                         // "+SEQ..."
        out.push_back({kind, c.position, c.count, len});
    }
    std::vector<std::pair<int, int64_t>> sums;
    sums.reserve(m.size());
    plan.per_pos[kind].clear();
    plan.per_pos[kind].reserve(m.size());
    for (auto& e : m) {
        int64_t c = 0;
        for (auto& kv : e.second) c += kv.second;
        sums.push_back({e.first, c});
        plan.per_pos[kind].push_back({e.first, (int64_t)e.second.size()});
    }
    screen_table(plan, t, sums);
}

void screen_clips(ScreenPlan& plan, int t, const std::map<int, HSclip>& m, bool as_candidates) {
    std::vector<std::pair<int, int64_t>> sums;
    sums.reserve(m.size());
    if (as_candidates) {
        plan.lists[SC_LGINS].clear();
        plan.per_pos[SC_LGINS].clear();
    }
    for (auto& e : m) {
        sums.push_back({e.first, (int64_t)e.second.v.cnt});
        if (as_candidates) {
            plan.lists[SC_LGINS].push_back({SC_LGINS, e.first, e.second.v.cnt, (int)e.second.col.size()});
            plan.per_pos[SC_LGINS].push_back({e.first, 1});
        }
    }
    screen_table(plan, t, sums);
}

}  // namespace

void screen_candidates_events(const HostPileup& P, ScreenPlan& plan) {
    std::thread t1([&] { screen_counted(plan, SC_MNP, ST_MNP, P.mnp); });
    screen_counted(plan, SC_DEL, ST_DEL, P.delCount);
    screen_counted(plan, SC_INS, ST_INS, P.insCount);
    t1.join();
}

void screen_candidates_finish(const HostPileup& P, ScreenPlan& plan) {
    std::thread t3([&] { screen_clips(plan, ST_SC3, P.sc3, false); });
    screen_clips(plan, ST_SC5, P.sc5, true);
    t3.join();
    plan.cand.clear();
    size_t total = 0;
    for (auto& l : plan.lists) total += l.size();
    plan.cand.reserve(total);
    for (int k = 0; k < SC_NKIND; k++) {
        plan.off[k] = (int)plan.cand.size();
        plan.cand.insert(plan.cand.end(), plan.lists[k].begin(), plan.lists[k].end());
    }
    plan.off[SC_NKIND] = (int)plan.cand.size();
    // every candidate position: merge of the per-kind position lists (each already sorted)
    std::vector<std::pair<int, int64_t>> all, tmp;
    for (int k = 0; k < SC_NKIND; k++) {
        tmp.clear();
        tmp.reserve(all.size() + plan.per_pos[k].size());
        std::merge(all.begin(), all.end(), plan.per_pos[k].begin(), plan.per_pos[k].end(), std::back_inserter(tmp),
                   [](const std::pair<int, int64_t>& a, const std::pair<int, int64_t>& b) { return a.first < b.first; });
        all.swap(tmp);
    }
    screen_table(plan, ST_ALL, all);
}

void screen_candidates(const HostPileup& P, ScreenPlan& plan) {
    screen_candidates_events(P, plan);
    screen_candidates_finish(P, plan);
}

bool realign_prefetch(const Config& c, HostRef& ref, int chrLen, const HostPileup& P, RowCache& rows) {
    // the rows the candidates' own handlers read first: the candidate positions, the few bases
    // after them, every non-dense / insertion position (read again when the record list is
    // built), and the positions the deletion candidates' flank mismatches will read
    std::vector<int> cand, probe;
    for (auto* m : {&P.mnp, &P.delCount, &P.insCount})
        for (auto& e : *m) cand.push_back(e.first);
    for (auto& e : P.nd) cand.push_back(e.first);
    for (auto& e : P.ins) cand.push_back(e.first);
    std::vector<int> exact;
    if (c.performLocalRealignment) del_probe_positions(ref, chrLen, P.delCount, exact);
    return rows.prefetch(cand, 4, exact);
}

void realign(const Config& c, int region_start, int region_end, HostRef& ref, int chrLen, ReadQuery* reads,
             RState& S, RealignStats& st, bool prefetched, ScreenPlan* plan) {
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    auto t0 = now();
    if (!prefetched && !realign_prefetch(c, ref, chrLen, S.P, S.rows)) throw std::runtime_error("dense row fetch failed");
    auto t1 = now();
    if (plan && plan->label) plan->changed.assign(plan->cand.size(), 0);
    Realigner rg{c, region_start, region_end, ref, chrLen, reads, S, st, plan};
    rg.adjustMNP();
    auto t2 = now();
    st.ms_prefetch = ms(t0, t1);
    st.ms_mnp = ms(t1, t2);
    if (c.performLocalRealignment) {
        rg.realigndel(true, S.P.delCount);
        auto t3 = now();
        rg.realignins(S.P.insCount);
        auto t4 = now();
        rg.realignlgins30();
        auto t5 = now();
        st.ms_del = ms(t2, t3);
        st.ms_ins = ms(t3, t4);
        st.ms_lgins = ms(t4, t5);
    }
    st.touched = (int64_t)S.pos.size();
    for (auto& e : S.pos)
        if (e.second.dirty) st.changed++;
}

}  // namespace rv
