// Device side of the per-position scoring stage (see rv_score.h for the restated reference code).
// Included by rv_score.hip only.
#pragma once
#include "rv_fisher.h"
#include "rv_score.h"

namespace rv {

constexpr int MAXV = 64;  // alleles per position scored on the device

struct Allele {
    const char* d;
    int len;
    const HVarD* v;  // the record (global memory) or the position's dense copy
};

template <int NA>
struct PosView {
    bool key;       // nonInsertionVariants has the position
    int cov;        // refCoverage (0 = absent)
    int n, nins;
    Allele a[NA];
    Allele ins[NA];
    HVarD dv[4];  // dense rows' values
    bool overflow;
    bool complex_;
};

RV_HD bool dstr_less(const char* a, int la, const char* b, int lb) {
    int l = la < lb ? la : lb;
    for (int i = 0; i < l; i++) {
        unsigned char x = (unsigned char)a[i], y = (unsigned char)b[i];
        if (x != y) return x < y;
    }
    return la < lb;
}
RV_HD bool dstr_eq(const char* a, int la, const char* b, int lb) {
    if (la != lb) return false;
    for (int i = 0; i < la; i++)
        if (a[i] != b[i]) return false;
    return true;
}

__device__ HVarD dense_hvar(const DenseBase& b) {
    HVarD v;
    v.cnt = b.cnt;
    v.fwd = b.fwd;
    v.rev = b.rev;
    v.hicnt = b.hi;
    v.locnt = b.lo;
    v.extracnt = 0;
    v.pmean = (double)b.sum_tp;
    v.qmean = (double)b.sum_q;
    v.mqmean = (double)b.sum_mq;
    v.nm = (double)b.sum_nm;
    v.pstd = (b.flags & 2u) ? 1 : 0;
    v.qstd = (b.flags & 4u) ? 1 : 0;
    return v;
}

__device__ int cx_find(const ScoreIn& S, int p) {
    int lo = 0, hi = S.n_cx - 1;
    while (lo <= hi) {
        int m = (lo + hi) >> 1;
        int v = S.cx_pos[m];
        if (v == p) return m;
        if (v < p)
            lo = m + 1;
        else
            hi = m - 1;
    }
    return -1;
}

// Loads the alleles of position p (nonInsertionVariants entries in key order, insertion entries
// in key order).  Dense positions list A,C,G,T rows that exist.
template <int NA>
__device__ void load_pos(const ScoreIn& S, int p, PosView<NA>& v) {
    v.n = v.nins = 0;
    v.overflow = false;
    v.key = false;
    v.cov = 0;
    int k = cx_find(S, p);
    v.complex_ = k >= 0;
    if (k >= 0) {
        v.key = S.cx_key[k] != 0;
        v.cov = S.cx_cov[k];
        for (int e = S.cx_off[k]; e < S.cx_off[k + 1]; e++) {
            const CxEntry& c = S.cx[e];
            Allele al{S.cx_arena + c.desc_off, c.desc_len, &c.v};
            if (c.kind == 1) {
                if (v.nins < NA)
                    v.ins[v.nins++] = al;
                else
                    v.overflow = true;
            } else {
                if (v.n < NA)
                    v.a[v.n++] = al;
                else
                    v.overflow = true;
            }
        }
        return;
    }
    int i = p - S.dense_lo;
    if (i < 0 || i >= S.dense_n) return;
    const DenseRow& r = S.rows[i];
    v.key = (r.rowflags & 1u) != 0;
    v.cov = r.cov;
    for (int b = 0; b < 4; b++)
        if (r.b[b].flags & 1u) {
            if (v.n < NA) {
                v.dv[v.n] = dense_hvar(r.b[b]);
                v.a[v.n] = Allele{"ACGT" + b, 1, &v.dv[v.n]};
                v.n++;
            } else {
                v.overflow = true;
            }
        }
}

// --------------------------------------------------------------------------- strings
template <int CAP>
struct SS : FStr<CAP> {
    __device__ void set(const char* p, int k) {
        this->clear();
        this->append(p, k);
    }
    __device__ bool starts(char ch) const { return this->n > 0 && this->c[0] == ch; }
    __device__ bool replace_first_char(char from, const char* to, int tl) {
        int i = this->find(from);
        if (i < 0) return false;
        SS<CAP> t;
        t.clear();
        t.append(this->c, i);
        t.append(to, tl);
        t.append(this->c + i + 1, this->n - i - 1);
        bool of = this->overflow || t.overflow;
        *this = t;
        this->overflow = of;
        return true;
    }
    __device__ void erase_range(int a, int b) {
        for (int i = a; i + (b - a) < this->n; i++) this->c[i] = this->c[i + (b - a)];
        this->n -= (b - a);
    }
};

struct RefWindow {
    const char* ref;
    int lo, hi;
    __device__ bool has(int k) const { return k >= lo && k <= hi; }
    __device__ char at(int k) const { return ref[k - lo]; }
    // joinRef (include/VariationUtils.h:336-342): false when a key is missing (the reference throws)
    template <int CAP>
    __device__ bool join(SS<CAP>& out, int from, int to) const {
        out.clear();
        for (int i = from; i <= to; i++) {
            if (!has(i)) return false;
            out.push(at(i));
        }
        return true;
    }
};

__device__ inline bool isATGC(char c) { return c == 'A' || c == 'T' || c == 'G' || c == 'C'; }

// findMSI (src/ToVarsBuilder.cpp:656-722)
struct MSIRes {
    double msi;
    int shift3;
    int msint_len;
};
__device__ MSIRes find_msi(const char* t1, int l1, const char* t2, int l2, const char* left, int ll) {
    int nmsi = 1, shift3 = 0, maxlen = 0;
    double msicnt = 0;
    while (nmsi <= l1 && nmsi <= 6) {
        const char* msint = t1 + (l1 - nmsi);
        int lm = ll + l1;
        auto mub = [&](int i) -> char { return i < ll ? left[i] : t1[i - ll]; };
        double cur = 1.0;
        for (int i = lm - nmsi; i > 0; i -= nmsi) {
            bool eq = i > nmsi;
            if (eq)
                for (int k = 0; k < nmsi; k++)
                    if (mub(i - nmsi + k) != msint[k]) {
                        eq = false;
                        break;
                    }
            if (eq)
                cur += 1;
            else
                break;
        }
        for (int i = 0; i < l2; i += nmsi) {
            bool eq = i + nmsi < l2;
            if (eq)
                for (int k = 0; k < nmsi; k++)
                    if (t2[i + k] != msint[k]) {
                        eq = false;
                        break;
                    }
            if (eq)
                cur += 1.0;
            else
                break;
        }
        if (cur > msicnt) {
            maxlen = nmsi;
            msicnt = cur;
        }
        nmsi++;
    }
    auto ts = [&](int i) -> char { return i < l1 ? t1[i] : t2[i - l1]; };
    while (shift3 < l2 && ts(shift3) == t2[shift3]) shift3++;
    return MSIRes{msicnt, shift3, maxlen};
}

// Fisher (FP64): rv_fisher.h

namespace {

// One candidate allele of the position: the record it came from and the coverage figures
// createVariant / createInsertion divide by.  Derived fields are computed when needed.
struct VarW {
    const Allele* al;
    int ttcov, hicov;
    double key;  // meanQuality * positionCoverage, the sort key of CMP_VARI
};

struct VarX {
    int poscov, fwd, rev, bias, pstd, qstd, hicnt, hicov;
    double freq, pmean, qmean, mq, hlr, hifreq, extrafreq, nm;
};

__device__ inline int strand_bias(int f, int r, const Config& c) {
    if (f + r <= 12) return f * r > 0 ? 2 : 0;
    return (f / (double)(f + r) >= c.bias && r / (double)(f + r) >= c.bias && f >= c.minBiasReads && r >= c.minBiasReads)
               ? 2
               : 1;
}

__device__ inline VarW make_var(const Allele& al, int ttcov, int hicov) {
    const HVarD& v = *al.v;
    double qmean = v.qmean / v.cnt;
    return VarW{&al, ttcov, hicov, qmean * v.cnt};
}

// createVariant / createInsertion record (ToVarsBuilder.cpp:419-443, :507-527)
__device__ inline VarX expand(const VarW& w, const Config& c) {
    const HVarD& v = *w.al->v;
    VarX t;
    t.poscov = v.cnt;
    t.fwd = v.fwd;
    t.rev = v.rev;
    t.bias = strand_bias(v.fwd, v.rev, c);
    t.freq = v.cnt / (double)w.ttcov;
    t.pmean = v.pmean / (double)v.cnt;
    t.pstd = v.pstd;
    t.qmean = v.qmean / v.cnt;
    t.qstd = v.qstd;
    t.mq = v.mqmean / (double)v.cnt;
    t.hlr = v.hicnt / (v.locnt != 0 ? (double)v.locnt : 0.5);
    t.hifreq = w.hicov > 0 ? v.hicnt / (double)w.hicov : 0;
    t.extrafreq = v.extracnt != 0 ? v.extracnt / (double)w.ttcov : 0;
    t.nm = v.nm / (double)v.cnt;
    t.hicnt = v.hicnt;
    t.hicov = w.hicov;
    return t;
}

// CMP_VARI (ToVarsBuilder.cpp:20-26)
__device__ inline bool cmp_vari(const VarW& a, const VarW& b) {
    double res = a.key - b.key;
    if (res < 0.00001 && res > -0.00001) return dstr_less(a.al->d, a.al->len, b.al->d, b.al->len);
    return res > 0;
}

// std::sort as libstdc++ performs it: final insertion sort for n <= 16; larger inputs go through
// the median-of-three introsort partition first (bounded here; the comparator is a strict weak
// order except for near-ties within 1e-5).
__device__ void insertion_sort(VarW* v, int lo, int hi) {
    for (int i = lo + 1; i < hi; i++) {
        VarW x = v[i];
        if (cmp_vari(x, v[lo])) {
            for (int k = i; k > lo; k--) v[k] = v[k - 1];
            v[lo] = x;
        } else {
            int k = i;
            while (k - 1 >= lo && cmp_vari(x, v[k - 1])) {
                v[k] = v[k - 1];
                k--;
            }
            v[k] = x;
        }
    }
}
__device__ inline void swapv(VarW& a, VarW& b) {
    VarW t = a;
    a = b;
    b = t;
}
__device__ void introsort_loop(VarW* v, int first, int last, int depth, unsigned long long* ctr) {
    while (last - first > 16) {
        if (depth == 0) {
            atomicAdd(ctr, 1ull);
            insertion_sort(v, first, last);
            return;
        }
        --depth;
        int mid = first + (last - first) / 2;
        int a = first + 1, b = mid, c = last - 1;
        if (cmp_vari(v[a], v[b])) {
            if (cmp_vari(v[b], v[c]))
                swapv(v[first], v[b]);
            else if (cmp_vari(v[a], v[c]))
                swapv(v[first], v[c]);
            else
                swapv(v[first], v[a]);
        } else if (cmp_vari(v[a], v[c]))
            swapv(v[first], v[a]);
        else if (cmp_vari(v[b], v[c]))
            swapv(v[first], v[c]);
        else
            swapv(v[first], v[b]);
        int lo = first + 1, hi = last;
        while (true) {
            while (lo < last && cmp_vari(v[lo], v[first])) ++lo;
            --hi;
            while (hi > first && cmp_vari(v[first], v[hi])) --hi;
            if (!(lo < hi)) break;
            swapv(v[lo], v[hi]);
            ++lo;
        }
        introsort_loop(v, lo, last, depth, ctr);
        last = lo;
    }
}
__device__ void sort_vari(VarW* v, int n, unsigned long long* ctr) {
    if (n < 2) return;
    if (n > 16) {
        atomicAdd(ctr, 1ull);
        int lg = 0;
        for (int m = n; m > 1; m >>= 1) lg++;
        introsort_loop(v, 0, n, 2 * lg, ctr);
        insertion_sort(v, 0, 16);
        for (int i = 16; i < n; i++) {
            VarW x = v[i];
            int k = i;
            while (k - 1 >= 0 && cmp_vari(x, v[k - 1])) {
                v[k] = v[k - 1];
                k--;
            }
            v[k] = x;
        }
        return;
    }
    insertion_sort(v, 0, n);
}

__device__ inline int parse_int(const char* p, int n, int& used) {
    long long v = 0;
    int i = 0;
    while (i < n && p[i] >= '0' && p[i] <= '9') v = v * 10 + (p[i++] - '0');
    used = i;
    return (int)v;
}

template <int CAP>
struct Strs {
    SS<CAP> refallele, varallele, genotype, g1, g2;
    SS<32> bias;
    SS<64> left, right;  // 20-base flanks
};

}  // namespace

// Appends one record and its strings (refallele, varallele, genotype, bias, leftseq, rightseq,
// descriptor); false when the buffers are full (counted).
// The lanes of a wavefront that reach this call together take their slots with one pair of atomics
// (issued by the lowest of them): every scoring thread appends to the same two counters, and one
// atomic per record serialised them.
__device__ inline bool emit_out(const ScoreIn& S, OutRec& o, const char* const pc[7], const int pn[7]) {
    unsigned long long nbytes = 0;
    for (int k = 0; k < 7; k++) nbytes += (unsigned long long)pn[k];
    const unsigned long long active = __builtin_amdgcn_read_exec();
    const int lane = (int)__lane_id();
    unsigned rank = 0, cnt = 0;
    unsigned long long before = 0, total = 0;
    for (unsigned long long m = active; m; m &= m - 1) {
        const int l = __builtin_ctzll(m);  // wave-uniform
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)nbytes, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(nbytes >> 32), l);
        const unsigned long long v = ((unsigned long long)hi << 32) | lo;
        if (l < lane) {
            before += v;
            rank++;
        }
        total += v;
        cnt++;
    }
    const int leader = __builtin_ctzll(active);
    unsigned base = 0;
    unsigned long long abase = 0;
    if (lane == leader) {
        base = atomicAdd(S.n_out, cnt);
        abase = atomicAdd(S.arena_used, total);
    }
    base = (unsigned)__builtin_amdgcn_readlane((int)base, leader);
    abase = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(unsigned)(abase >> 32), leader) << 32) |
            (unsigned)__builtin_amdgcn_readlane((int)(unsigned)abase, leader);
    const unsigned oi = base + rank;
    const unsigned long long off = abase + before;
    if (oi >= S.max_out || off + nbytes > S.arena_cap) {
        atomicAdd(&S.counters[0], 1ull);
        return false;
    }
    char* w = S.arena + off;
    for (int k = 0; k < 7; k++) {
        for (int i = 0; i < pn[k]; i++) *w++ = pc[k][i];
        o.len[k] = (uint16_t)pn[k];
    }
    o.str_off = (uint64_t)off;
    S.out[oi] = o;
    return true;
}

// A position's reference variant as a somatic record: as createVariant left it (OUT_REF), or in the
// reference-only form ToVarsBuilder gives a position without variants (OUT_REF_ONLY,
// ToVarsBuilder.cpp:1074-1097).  x == nullptr: the empty Variant ToVarsBuilder creates when no
// allele has reads (:1098-1100).
// OUT_VARIANT_RAW: a variant collectReferenceVariants never completed (createVariant's fields, and
// the left flank when only the right one failed), in CMP_VARI rank `order`.
__device__ inline void emit_ref(const ScoreIn& S, int p, int kind, const VarX* x, const char* d, int dl, int tpc,
                                int rfc, int rrc, char rbase, int order = -1, const char* lfc = nullptr,
                                int lfn = 0) {
    OutRec o;
    o.position = p;
    o.order = order;
    o.kind = kind;
    o.nfields = 0;
    o.fisher_p = o.fisher_ratio = 0;
    o.vartype = VT_EMPTY;
    o.msi = 0;
    o.start = o.end = 0;
    o.tpc = o.poscov = o.rfc = o.rrc = o.fwd = o.rev = o.shift3 = o.msint = o.hicnt = o.hicov = o.pstd = o.qstd = 0;
    o.freq = o.pmean = o.qmean = o.mq = o.hlr = o.hifreq = o.extrafreq = o.msi = o.nm = 0;
    char bias[8], allele[2] = {0, 0}, gt[4] = {0, 0, 0, 0};
    int nb = 0, na = 0, ng = 0;
    if (x) {
        o.poscov = x->poscov;
        o.fwd = x->fwd;
        o.rev = x->rev;
        o.hicnt = x->hicnt;
        o.hicov = x->hicov;
        o.pstd = x->pstd;
        o.qstd = x->qstd;
        o.freq = x->freq;
        o.pmean = x->pmean;
        o.qmean = x->qmean;
        o.mq = x->mq;
        o.hlr = x->hlr;
        o.hifreq = x->hifreq;
        o.extrafreq = x->extrafreq;
        o.nm = x->nm;
        bias[nb++] = (char)('0' + x->bias);
        if (kind == OUT_REF_ONLY) {
            o.tpc = tpc;
            o.poscov = 0;
            o.freq = 0;
            o.rfc = rfc;
            o.rrc = rrc;
            o.fwd = o.rev = 0;
            o.start = o.end = p;
            bias[nb++] = ';';
            bias[nb++] = '0';
            if (rbase) {
                // validateRefallele maps IUPAC codes to their alphabetically first base
                char v = rbase;
                switch (v) {
                    case 'M': case 'R': case 'W': case 'V': case 'H': case 'D': v = 'A'; break;
                    case 'S': case 'Y': case 'B': v = 'C'; break;
                    case 'K': v = 'G'; break;
                    default: break;
                }
                allele[0] = v;
                na = 1;
                gt[0] = rbase;
                gt[1] = '/';
                gt[2] = rbase;
                ng = 3;
            } else {
                gt[0] = '/';
                ng = 1;
            }
        }
    } else {
        bias[nb++] = '0';
    }
    if (S.cfg.fisher) {
        // the simple printer's Fisher columns (a pileup line of a reference-only position)
        o.fisher_p = fisher_two_sided(o.rfc, o.rrc, o.fwd, o.rev, nullptr, nullptr, LGTable{S.lgamma_tab, S.lgamma_n});
        const double ad = (double)(o.rfc * o.rev), bc = (double)(o.rrc * o.fwd);
        if (bc != 0 && ad != 0) o.fisher_ratio = ad > bc ? (ad / bc) : (bc / ad);
    }
    const char* pc[7] = {allele, allele, gt, bias, lfc ? lfc : allele, allele, d};
    const int pn[7] = {na, na, ng, nb, lfc ? lfn : 0, 0, dl};
    emit_out(S, o, pc, pn);
}

// The checks ToVarsBuilder::process makes before building any variant at p (ToVarsBuilder.cpp:
// 103-136), on the loaded view: false when the position ends there (no key, no allele, no coverage,
// the reference base alone), or when the record-list kernel scores it.  An overflowing view counts
// as work: score_position reports it.
template <int NA>
__device__ __forceinline__ bool score_needed(const ScoreIn& S, int p, bool dense_only, const PosView<NA>& V) {
    if (dense_only && V.complex_) return false;  // scored by the record-list kernel
    if (!V.key) return false;
    if (V.overflow) return true;
    if (V.n == 0 && V.nins == 0) return false;
    if (V.cov == 0) return false;
    const RefWindow rf{S.ref, S.ref_lo, S.ref_hi};
    // isTheSameVariationOnRef (ToVarsBuilder.cpp:213-233)
    if (V.n == 1 && V.nins == 0 && rf.has(p) && V.a[0].len == 1 && V.a[0].d[0] == rf.at(p))
        return S.cfg.doPileup || S.cfg.hasBam2;
    return true;
}

// score_needed for a dense-kernel position, from the row's flags alone (no allele view): what the
// pick pass evaluates for every dense position.
// dense_needed's checks of the row itself: they read only the pileup's rows and the reference
__device__ __forceinline__ bool dense_row_needed(const ScoreIn& S, int p) {
    const int i = p - S.dense_lo;
    if (i < 0 || i >= S.dense_n) return false;
    const DenseRow& r = S.rows[i];
    if (!(r.rowflags & 1u)) return false;
    int n = 0, only = -1;
    for (int b = 0; b < 4; b++)
        if (r.b[b].flags & 1u) {
            n++;
            only = b;
        }
    if (n == 0 || r.cov == 0) return false;
    const RefWindow rf{S.ref, S.ref_lo, S.ref_hi};
    if (n == 1 && rf.has(p) && "ACGT"[only] == rf.at(p)) return S.cfg.doPileup || S.cfg.hasBam2;
    return true;
}

__device__ __forceinline__ bool dense_needed(const ScoreIn& S, int p) {
    if (cx_find(S, p) >= 0) return false;  // scored by the record-list kernel
    return dense_row_needed(S, p);
}

template <int CAP, int NA>
__device__ void score_position(const ScoreIn& S, int p, bool dense_only, int slot) {
    const Config& c = S.cfg;
    PosView<NA> V;
    load_pos<NA>(S, p, V);
    if (!score_needed<NA>(S, p, dense_only, V)) return;
    if (V.overflow) {
        atomicAdd(&S.counters[0], 1ull);
        return;
    }
    RefWindow rf{S.ref, S.ref_lo, S.ref_hi};
    int tpc = V.cov;
    int hicov = 0;
    for (int i = 0; i < V.n; i++) hicov += V.a[i].v->hicnt;
    for (int i = 0; i < V.nins; i++) hicov += V.ins[i].v->hicnt;
    VarW var[2 * NA];
    int nv = 0;
    // createVariant: keys already in ASCII order
    for (int i = 0; i < V.n; i++) {
        const Allele& al = V.a[i];
        if (al.v->cnt == 0) continue;
        int ttcov = tpc;
        if (al.v->cnt > tpc && al.v->extracnt > 0 && al.v->cnt - tpc < al.v->extracnt) ttcov = al.v->cnt;
        var[nv++] = make_var(al, ttcov, hicov);
    }
    // createInsertion
    int hicov_ins = hicov;
    int dec_fwd = 0, dec_rev = 0;  // decrement createInsertion applies to the reference allele at p+1
    for (int i = 0; i < V.nins; i++) {
        const Allele& al = V.ins[i];
        int ttcov = tpc;
        if (al.v->cnt > tpc && al.v->extracnt != 0 && al.v->cnt - tpc < al.v->extracnt) ttcov = al.v->cnt;
        if (ttcov < al.v->cnt) {
            ttcov = al.v->cnt;
            // refCoverage[p+1]; when it wins, the reference allele at p+1 loses this insertion's
            // strand counts (the reference decrements that record in place; counted here)
            int k1 = cx_find(S, p + 1);
            int cov1 = 0;
            if (k1 >= 0)
                cov1 = S.cx_cov[k1];
            else if (p + 1 - S.dense_lo >= 0 && p + 1 - S.dense_lo < S.dense_n)
                cov1 = S.rows[p + 1 - S.dense_lo].cov;
            if (cov1 != 0 && rf.has(p + 1) && ttcov < cov1 - al.v->cnt) {
                ttcov = cov1;
                atomicAdd(&S.counters[1], 1ull);
                dec_fwd += al.v->fwd;
                dec_rev += al.v->rev;
            }
            tpc = ttcov;
        }
        if (hicov_ins < al.v->hicnt) hicov_ins = al.v->hicnt;
        var[nv++] = make_var(al, ttcov, hicov_ins);
    }
    sort_vari(var, nv, &S.counters[2]);
    // collectVarsAtPosition
    int refi = -1;
    double maxfreq = 0;
    int nonref = 0;
    for (int i = 0; i < nv; i++) {
        const Allele& al = *var[i].al;
        if (rf.has(p) && al.len == 1 && al.d[0] == rf.at(p)) {
            refi = i;
        } else {
            nonref++;
            double f = al.v->cnt / (double)var[i].ttcov;
            if (f > maxfreq) maxfreq = f;
        }
    }
    if (nv == 0) {
        // no allele has reads: simple mode creates no Vars entry; with a second sample, or in
        // pileup mode (-p), ToVarsBuilder still creates one, holding an empty reference variant
        // (ToVarsBuilder.cpp:1098-1100), which -p prints (simpleMode.cpp:160-170)
        if (S.somatic == SCORE_SIMPLE) {
            if (!c.doPileup) return;
            S.status[slot] = 1;
            if (!(p < S.region_start || p > S.region_end)) emit_ref(S, p, OUT_REF_ONLY, nullptr, nullptr, 0, 0, 0, 0, 0);
            return;
        }
        S.status[slot] = 3;
        if (S.somatic == SCORE_SOMATIC_REF && S.want[slot] && !(p < S.region_start || p > S.region_end))
            emit_ref(S, p, OUT_REF_ONLY, nullptr, nullptr, 0, 0, 0, 0, 0);
        return;
    }
    // alignedVariants: the position was added, and is erased again when nothing passes -f
    if (!c.doPileup && maxfreq <= c.freq && !c.hasBam2) {
        S.status[slot] = 2;
        return;
    }
    S.status[slot] = (S.somatic != SCORE_SIMPLE && nonref == 0) ? 3 : 1;
    bool in_region = !(p < S.region_start || p > S.region_end);
    // the output filter skips positions outside the region; nothing else reads them
    if (!in_region) return;
    if (S.somatic == SCORE_SOMATIC_REF) {
        // only positions without variants, where the other sample has some
        if (nonref == 0 && S.want[slot]) {
            const VarX rx = expand(var[refi], c);
            const char rb = rf.has(p) ? rf.at(p) : 0;
            emit_ref(S, p, OUT_REF_ONLY, &rx, var[refi].al->d, var[refi].al->len, tpc, rx.fwd, rx.rev, rb);
        }
        return;
    }
    if (S.somatic == SCORE_SIMPLE && c.doPileup && nonref == 0) {
        // -p prints a reference-only position's reference variant in its else-branch form
        // (ToVarsBuilder.cpp:1074-1097; simpleMode.cpp:160-170)
        const VarX rx = expand(var[refi], c);
        const char rb = rf.has(p) ? rf.at(p) : 0;
        emit_ref(S, p, OUT_REF_ONLY, &rx, var[refi].al->d, var[refi].al->len, tpc, rx.fwd, rx.rev, rb);
        return;
    }
    if (S.somatic == SCORE_SOMATIC_VARS) {
        if (nonref == 0) return;
        if (refi >= 0) {
            const VarX rx = expand(var[refi], c);
            emit_ref(S, p, OUT_REF, &rx, var[refi].al->d, var[refi].al->len, 0, 0, 0, 0);
        }
    }
    Strs<CAP>* st;
    Strs<CAP> stv;
    st = &stv;
    const int chr0 = S.chr_len;
    // genotype1 (ToVarsBuilder.cpp:737-751)
    SS<CAP>& g1 = st->g1;
    VarX refx;
    if (refi >= 0) refx = expand(var[refi], c);
    {
        const VarW* gsrc = nullptr;
        if (refi >= 0 && refx.freq >= c.freq)
            gsrc = &var[refi];
        else if (nonref > 0) {
            for (int i = 0; i < nv; i++)
                if (i != refi) {
                    gsrc = &var[i];
                    break;
                }
        } else if (refi >= 0)
            gsrc = &var[refi];
        if (!gsrc) return;
        g1.set(gsrc->al->d, gsrc->al->len);
        if (g1.starts('+')) {
            int l = g1.n - 1;
            g1.clear();
            g1.push('+');
            g1.append_int(l);
        }
    }
    int rfc = 0, rrc = 0;
    if (refi >= 0) {
        rfc = refx.fwd;
        rrc = refx.rev;
    }
    if (tpc > V.cov) {
        // reference allele at p+1 (its strand counts replace the reference's)
        int k1 = cx_find(S, p + 1);
        char rb1 = rf.has(p + 1) ? rf.at(p + 1) : 0;
        if (rb1) {
            if (k1 >= 0) {
                if (S.cx_key[k1])
                    for (int e = S.cx_off[k1]; e < S.cx_off[k1 + 1]; e++) {
                        const CxEntry& ce = S.cx[e];
                        if (ce.kind == 0 && ce.desc_len == 1 && S.cx_arena[ce.desc_off] == rb1) {
                            rfc = ce.v.fwd - dec_fwd;
                            rrc = ce.v.rev - dec_rev;
                        }
                    }
            } else if (p + 1 - S.dense_lo >= 0 && p + 1 - S.dense_lo < S.dense_n) {
                const DenseRow& r1 = S.rows[p + 1 - S.dense_lo];
                int b = base_index(rb1);
                if ((r1.rowflags & 1u) && b >= 0 && (r1.b[b].flags & 1u)) {
                    rfc = r1.b[b].fwd - dec_fwd;
                    rrc = r1.b[b].rev - dec_rev;
                }
            }
        }
    }
    const VarX* refv = refi >= 0 ? &refx : nullptr;
    SS<CAP>& g2 = st->g2;
    int order = 0;
    if (nonref > 0) {
        for (int vi = 0; vi < nv; vi++) {
            if (vi == refi) continue;
            const VarX vr = expand(var[vi], c);
            const char* d = var[vi].al->d;
            const int dl = var[vi].al->len;
            g2.set(d, dl);
            if (g2.starts('+')) {
                int l = g2.n - 1;
                g2.clear();
                g2.push('+');
                g2.append_int(l);
            }
            int dellen = 0;
            if (dl >= 2 && d[0] == '-' && d[1] >= '0' && d[1] <= '9') {
                int used;
                dellen = parse_int(d + 1, dl - 1, used);
            }
            int endPosition = p;
            if (dl > 0 && d[0] == '-') endPosition = p + dellen - 1;
            SS<CAP>& ra = st->refallele;
            SS<CAP>& va = st->varallele;
            ra.clear();
            va.clear();
            int shift3 = 0;
            double msi = 0;
            int msint = 0;
            int startPosition = p;
            bool thrown = false;
            auto has_ch = [&](char ch) {
                for (int i = 0; i < dl; i++)
                    if (d[i] == ch) return true;
                return false;
            };
            bool has_dup = false;
            for (int i = 0; i + 3 < dl; i++)
                if (d[i] == '<' && d[i + 1] == 'd' && d[i + 2] == 'u' && d[i + 3] == 'p') has_dup = true;
            if (dl > 0 && d[0] == '+') {
                if (!has_ch('&') && !has_ch('#') && !has_dup) {
                    // proceedVrefIsInsertion (ToVarsBuilder.cpp:283-317)
                    int a0 = p - 50 > 1 ? p - 50 : 1;
                    int b2 = p + 70 > chr0 ? chr0 : p + 70;
                    bool ok = true;
                    for (int i = a0; i <= p && ok; i++) ok = rf.has(i);
                    for (int i = p + 1; i <= b2 && ok; i++) ok = rf.has(i);
                    if (!ok) {
                        thrown = true;
                    } else {
                        const char* left = S.ref + (a0 - S.ref_lo);
                        int ll = p - a0 + 1;
                        const char* t2 = p + 1 <= b2 ? S.ref + (p + 1 - S.ref_lo) : S.ref;
                        int l2 = b2 - p > 0 ? b2 - p : 0;
                        MSIRes m = find_msi(d + 1, dl - 1, t2, l2, left, ll);
                        MSIRes w = find_msi(left, ll, t2, l2, left, 0);
                        double ms = m.msi;
                        int mlen = m.msint_len;
                        if (ms < w.msi) {
                            ms = w.msi;
                            mlen = w.msint_len;
                        }
                        if (ms <= m.shift3 / (double)(dl - 1)) ms = m.shift3 / (double)(dl - 1);
                        msi = ms;
                        shift3 = m.shift3;
                        msint = mlen;
                    }
                }
                if (!thrown) {
                    if (c.moveIndelsTo3) {
                        startPosition += shift3;
                        endPosition += shift3;
                    }
                    if (rf.has(p)) ra.push(rf.at(p));
                    va.append(ra.c, ra.n);
                    va.append(d + 1, dl - 1);
                }
            } else if (dl > 0 && d[0] == '-') {
                if (dellen < c.SVMINLEN) {
                    // varallele = descriptor without its leading -N
                    int k = 1;
                    while (k < dl && d[k] >= '0' && d[k] <= '9') k++;
                    if (k > 1)
                        va.set(d + k, dl - k);
                    else
                        va.set(d, dl);
                    // proceedVrefIsDeletion (ToVarsBuilder.cpp:242-274)
                    int a0 = p - 70 > 1 ? p - 70 : 1;
                    int b1 = p + dellen + 70 > chr0 ? chr0 : p + dellen + 70;
                    bool ok = true;
                    for (int i = a0; i <= p - 1 && ok; i++) ok = rf.has(i);
                    for (int i = p; i <= b1 && ok; i++) ok = rf.has(i);
                    int tl = b1 - p + 1 > 0 ? b1 - p + 1 : 0;
                    if (ok && dellen > tl) ok = false;  // vc_substr(tseq, dellen) past the end throws
                    if (!ok) {
                        thrown = true;
                    } else {
                        const char* left = S.ref + (a0 - S.ref_lo);
                        int ll = p - 1 - a0 + 1 > 0 ? p - a0 : 0;
                        const char* tseq = S.ref + (p - S.ref_lo);
                        int l1 = dellen < tl ? dellen : tl;
                        MSIRes m = find_msi(tseq, l1, tseq + dellen, tl - dellen, left, ll);
                        MSIRes w = find_msi(left, ll, tseq + dellen, tl - dellen, left, 0);
                        double ms = m.msi;
                        int mlen = m.msint_len;
                        if (ms < w.msi) {
                            ms = w.msi;
                            mlen = w.msint_len;
                        }
                        if (ms <= m.shift3 / (double)dellen) ms = m.shift3 / (double)dellen;
                        msi = ms;
                        shift3 = m.shift3;
                        msint = mlen;
                    }
                }
                if (!thrown) {
                    if (!has_ch('&') && !has_ch('#') && !has_ch('^')) {
                        if (c.moveIndelsTo3) startPosition += shift3;
                        bool isdel = va.n == 5 && va.c[0] == '<' && va.c[1] == 'D' && va.c[2] == 'E' && va.c[3] == 'L' && va.c[4] == '>';
                        if (!isdel) {
                            va.clear();
                            if (rf.has(p - 1)) va.push(rf.at(p - 1));
                        }
                        ra.clear();
                        if (rf.has(p - 1)) ra.push(rf.at(p - 1));
                        startPosition--;
                    }
                    if (dellen < c.SVMINLEN) {
                        for (int i = p; i <= p + dellen - 1 && !thrown; i++) {
                            if (!rf.has(i))
                                thrown = true;
                            else
                                ra.push(rf.at(i));
                        }
                    }
                }
            } else {
                // SNV / MNV: MSI over [p-30, p+1] and [p+2, min(p+70, chr0)]
                int a0 = p - 30 > 1 ? p - 30 : 1;
                int b2 = p + 70 > chr0 ? chr0 : p + 70;
                bool ok = true;
                for (int i = a0; i <= p + 1 && ok; i++) ok = rf.has(i);
                for (int i = p + 2; i <= b2 && ok; i++) ok = rf.has(i);
                if (!ok) {
                    thrown = true;
                } else {
                    const char* t1 = S.ref + (a0 - S.ref_lo);
                    int l1 = p + 1 - a0 + 1;
                    const char* t2 = S.ref + (p + 2 - S.ref_lo);
                    int l2 = b2 - (p + 2) + 1 > 0 ? b2 - (p + 2) + 1 : 0;
                    MSIRes m = find_msi(t1, l1, t2, l2, t1, 0);
                    msi = m.msi;
                    shift3 = m.shift3;
                    msint = m.msint_len;
                    if (rf.has(p)) ra.push(rf.at(p));
                    va.set(d, dl);
                }
            }
            // AMP_ATGC: rightmost '&' followed by [ATGC]+
            if (!thrown) {
                int amp = -1;
                for (int i = dl - 2; i >= 0; i--)
                    if (d[i] == '&' && isATGC(d[i + 1])) {
                        amp = i;
                        break;
                    }
                if (amp >= 0) {
                    int el = 0;
                    while (amp + 1 + el < dl && isATGC(d[amp + 1 + el])) el++;
                    va.replace_first_char('&', "", 0);
                    for (int i = endPosition + 1; i <= endPosition + el && !thrown; i++) {
                        if (!rf.has(i)) {
                            thrown = true;
                            break;
                        }
                        ra.push(rf.at(i));
                        g1.push(rf.at(i));
                    }
                    if (!thrown) {
                        endPosition += el;
                        int amp2 = -1;
                        for (int i = va.n - 2; i >= 0; i--)
                            if (va.c[i] == '&' && isATGC(va.c[i + 1])) {
                                amp2 = i;
                                break;
                            }
                        if (amp2 >= 0) {
                            int vel = 0;
                            while (amp2 + 1 + vel < va.n && isATGC(va.c[amp2 + 1 + vel])) vel++;
                            va.replace_first_char('&', "", 0);
                            for (int i = endPosition + 1; i <= endPosition + vel && !thrown; i++) {
                                if (!rf.has(i)) {
                                    thrown = true;
                                    break;
                                }
                                ra.push(rf.at(i));
                                g1.push(rf.at(i));
                            }
                            if (!thrown) endPosition += vel;
                        }
                        if (!thrown && dl > 0 && d[0] == '+') {
                            if (ra.n > 0) ra.erase_range(0, 1);
                            if (va.n > 0) va.erase_range(0, 1);
                            startPosition++;
                        }
                    }
                }
            }
            // HASH_GROUP_CARET_GROUP "#(.+)\^(.+)"
            if (!thrown) {
                int hi_ = -1, cj = -1;
                for (int i = 0; i < dl && hi_ < 0; i++) {
                    if (d[i] != '#') continue;
                    for (int j = dl - 2; j >= i + 2; j--)
                        if (d[j] == '^') {
                            cj = j;
                            break;
                        }
                    if (cj >= 0) hi_ = i;
                }
                if (hi_ >= 0) {
                    int mlen = cj - hi_ - 1;
                    const char* tail = d + cj + 1;
                    int tl = dl - cj - 1;
                    endPosition += mlen;
                    for (int i = endPosition - mlen + 1; i <= endPosition && !thrown; i++) {
                        if (!rf.has(i))
                            thrown = true;
                        else
                            ra.push(rf.at(i));
                    }
                    if (!thrown && tl > 0 && tail[0] >= '0' && tail[0] <= '9') {
                        int used;
                        int del = parse_int(tail, tl, used);
                        for (int i = endPosition + 1; i <= endPosition + del && !thrown; i++) {
                            if (!rf.has(i))
                                thrown = true;
                            else
                                ra.push(rf.at(i));
                        }
                        if (!thrown) endPosition += del;
                    }
                    if (!thrown) {
                        va.replace_first_char('#', "", 0);
                        // "\^(\d+)?": first '^' and the digits right after it
                        int ci = va.find('^');
                        if (ci >= 0) {
                            int e = ci + 1;
                            while (e < va.n && va.c[e] >= '0' && va.c[e] <= '9') e++;
                            va.erase_range(ci, e);
                        }
                        g1.replace_first_char('#', "m", 1);
                        g1.replace_first_char('^', "i", 1);
                        g2.replace_first_char('#', "m", 1);
                        g2.replace_first_char('^', "i", 1);
                    }
                }
            }
            // CARET_ATGNC "\^([ATGNC]+)"
            if (!thrown) {
                bool caret = false;
                for (int i = 0; i + 1 < dl; i++) {
                    char x = d[i + 1];
                    if (d[i] == '^' && (x == 'A' || x == 'T' || x == 'G' || x == 'N' || x == 'C')) caret = true;
                }
                if (caret) {
                    va.replace_first_char('^', "", 0);
                    g1.replace_first_char('^', "i", 1);
                    g2.replace_first_char('^', "i", 1);
                }
            }
            // flanks
            SS<64>& lf = st->left;
            SS<64>& rt = st->right;
            bool left_set = false;
            if (!thrown) {
                int a = startPosition - 20 < 1 ? 1 : startPosition - 20;
                if (!rf.join(lf, a, startPosition - 1))
                    thrown = true;
                else
                    left_set = true;
            }
            if (!thrown) {
                int b = endPosition + 20 > chr0 ? chr0 : endPosition + 20;
                if (!rf.join(rt, endPosition + 1, b)) thrown = true;
            }
            if (thrown) {
                // A reference lookup outside the window throws inside collectReferenceVariants and
                // the per-position catch (ToVarsBuilder.cpp:105, :183) ends it: this variant and
                // the later ones keep createVariant's fields (no alleles, so simple mode prints none
                // of them); the position's Vars entry stays, which the somatic printer reads.
                if (S.somatic == SCORE_SOMATIC_VARS) {
                    for (int vj = vi; vj < nv; vj++) {
                        if (vj == refi) continue;
                        const VarX xr = expand(var[vj], c);
                        const bool lf_kept = vj == vi && left_set;
                        emit_ref(S, p, OUT_VARIANT_RAW, &xr, var[vj].al->d, var[vj].al->len, 0, 0, 0, 0, order++,
                                 lf_kept ? lf.c : nullptr, lf_kept ? lf.n : 0);
                    }
                }
                return;
            }
            if (ra.overflow || va.overflow || g1.overflow || g2.overflow || lf.overflow || rt.overflow) {
                atomicAdd(&S.counters[0], 1ull);
                return;
            }
            // genotype = genotype1 + "/" + genotype2, minus '&' and '#', '^' -> 'i'
            SS<CAP>& gt = st->genotype;
            gt.clear();
            for (int i = 0; i < g1.n; i++) {
                char x = g1.c[i];
                if (x == '&' || x == '#') continue;
                gt.push(x == '^' ? 'i' : x);
            }
            gt.push('/');
            for (int i = 0; i < g2.n; i++) {
                char x = g2.c[i];
                if (x == '&' || x == '#') continue;
                gt.push(x == '^' ? 'i' : x);
            }
            // validateRefallele: IUPAC codes -> first base alphabetically
            for (int i = 0; i < ra.n; i++) {
                char x = ra.c[i], y = 0;
                switch (x) {
                    case 'M': case 'R': case 'W': case 'V': case 'H': case 'D': y = 'A'; break;
                    case 'S': case 'Y': case 'B': y = 'C'; break;
                    case 'K': y = 'G'; break;
                    default: break;
                }
                if (y) {
                    for (int k = 0; k <= i; k++)
                        if (ra.c[k] == x) {
                            ra.c[k] = y;
                            break;
                        }
                }
            }
            SS<32>& bs = st->bias;
            bs.clear();
            if (refv)
                bs.append_int(refv->bias);
            else
                bs.push('0');
            bs.push(';');
            bs.append_int(vr.bias);
            int fwd = vr.fwd < 0 ? 0 : vr.fwd, rev = vr.rev < 0 ? 0 : vr.rev;
            int orfc = rfc < 0 ? 0 : rfc, orrc = rrc < 0 ? 0 : rrc;
            // ---------------- output filter (simpleMode.cpp:176-201)
            bool hasN = false;
            for (int i = 0; i < ra.n; i++)
                if (ra.c[i] == 'N') hasN = true;
            if (hasN && S.somatic == SCORE_SIMPLE) continue;
            int vt;
            if (ra.n == 1 && va.n == 1)
                vt = VT_SNV;
            else if (va.n == 5 && va.c[0] == '<' && va.c[4] == '>')
                vt = VT_SV;
            else if (ra.n == 0 || va.n == 0)
                vt = VT_COMPLEX;
            else if (ra.c[0] != va.c[0])
                vt = VT_COMPLEX;
            else if (ra.n == 1 && va.n > 1)
                vt = VT_INS;  // varallele starts with refallele (same first base, one-base ref)
            else if (ra.n > 1 && va.n == 1)
                vt = VT_DEL;
            else
                vt = VT_COMPLEX;
            bool good = true;
            if (ra.n == 0) good = false;
            if (good && (vr.freq < c.freq || vr.hicnt < c.minr || vr.pmean < c.readPosFilter || vr.qmean < c.goodq)) good = false;
            if (good && refv && refv->hicnt > c.minr && vr.freq < 0.25) {
                double dd = vr.mq + ra.n + va.n;
                double f = (1 + dd) / (refv->mq + 1);
                if ((dd - 2 < 5 && refv->mq > 20) || f < 0.25) good = false;
            }
            if (good && vr.hlr < c.qratio) good = false;
            if (good && !(vr.freq > 0.30)) {
                if (vr.mq < c.mapq) good = false;
                if (good && bs.n == 3 && bs.c[0] == '2' && bs.c[1] == ';' && bs.c[2] == '1' && vr.freq < 0.20) {
                    if (vt == VT_SNV || (ra.n < 3 && va.n < 3)) good = false;
                }
            }
            if (!good && !c.doPileup && S.somatic == SCORE_SIMPLE) continue;
            // adjComplex (Variant.h:99-132)
            int sp = startPosition, ep = endPosition;
            if (S.somatic == SCORE_SIMPLE && vt == VT_COMPLEX && !(va.n > 0 && va.c[0] == '<')) {
                int n = 0;
                while (ra.n - n > 1 && va.n - n > 1 && ra.c[n] == va.c[n]) n++;
                if (n > 0) {
                    sp += n;
                    // leftseq = (leftseq + ref[0:n])[n:]  (same length as before)
                    for (int i = 0; i < lf.n; i++) {
                        int j = n + i;
                        lf.c[i] = j < lf.n ? lf.c[j] : ra.c[j - lf.n];
                    }
                    ra.erase_range(0, n);
                    va.erase_range(0, n);
                }
                int k = 1;
                while (ra.n - k > 0 && va.n - k > 0 && ra.c[ra.n - k] == va.c[va.n - k]) k++;
                if (k > 1) {
                    ep -= k - 1;
                    // rightseq = ref[-(k-1):] + rightseq[0 : len-(k-1)]
                    int keep = rt.n - (k - 1);
                    if (keep < 0) keep = 0;
                    SS<64> nr;
                    nr.clear();
                    nr.append(ra.c + (ra.n - (k - 1)), k - 1);
                    nr.append(rt.c, keep);
                    bool of = rt.overflow || nr.overflow;
                    rt = nr;
                    if (of) {
                        atomicAdd(&S.counters[0], 1ull);
                        return;
                    }
                    ra.n -= k - 1;
                    va.n -= k - 1;
                }
            }
            // ---------------- append the printed record
            OutRec o;
            o.kind = OUT_VARIANT;
            o.position = p;
            o.order = order++;
            o.start = sp;
            o.end = ep;
            o.tpc = tpc;
            o.poscov = vr.poscov;
            o.rfc = orfc;
            o.rrc = orrc;
            o.fwd = fwd;
            o.rev = rev;
            o.shift3 = shift3;
            o.msint = msint;
            o.hicnt = vr.hicnt;
            o.hicov = vr.hicov;
            o.pstd = vr.pstd;
            o.qstd = vr.qstd;
            o.vartype = vt;
            o.nfields = 36;
            o.freq = vr.freq;
            o.pmean = vr.pmean;
            o.qmean = vr.qmean;
            o.mq = vr.mq;
            o.hlr = vr.hlr;
            o.hifreq = vr.hifreq;
            o.extrafreq = vr.extrafreq;
            o.msi = msi;
            o.nm = vr.nm;
            o.fisher_p = 0;
            o.fisher_ratio = 0;
            if (c.fisher) {
                o.fisher_p = fisher_two_sided(orfc, orrc, fwd, rev, nullptr, nullptr, LGTable{S.lgamma_tab, S.lgamma_n});
                double ad = (double)(orfc * rev), bc = (double)(orrc * fwd);
                double ratio = 0.0;
                if (bc != 0 && ad != 0) ratio = ad > bc ? (ad / bc) : (bc / ad);
                o.fisher_ratio = ratio;
            }
            const char* pc[7] = {ra.c, va.c, gt.c, bs.c, lf.c, rt.c, d};
            const int pn[7] = {ra.n, va.n, gt.n, bs.n, lf.n, rt.n, dl};
            if (!emit_out(S, o, pc, pn)) return;
        }
    }
}

}  // namespace rv
