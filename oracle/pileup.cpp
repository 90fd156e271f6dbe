// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// Read filters and CIGAR walk into the pileup maps, restated from
//   src/recordPreprocessor.cpp:101-180   (next_record filters)
//   src/parseCigar.cpp:410-440           (record loop, per-read exception swallow)
//   src/parseCigar.cpp:497-974           (parseCigar + M-run loop)
//   src/parseCigar.cpp:208-256           (findOffset)
//   src/parseCigar.cpp:1112-1301         (process_softclip)
//   src/parseCigar.cpp:1303-1523         (process_insertion)
//   src/parseCigar.cpp:1525-1727         (process_deletion)
//   src/parseCigar.cpp:1036-1082         (addVariationForDeletion)
//   src/parseCigar.cpp:1756-1815         (appendSegments)
//   src/parseCigar.cpp:1858-1892         (cleanupCigar)
//   include/VariationUtils.h:99-113      (adjInsPos)
//
// Pinned by tests/test_oracle_cases.py (hand-built reads: match runs, substitutions, deletions,
// insertions, coverage, strand and quality sums, read-position statistic, pstd/qstd).
// Pinned by tests/test_oracle_softclip_kat.py: the soft-clip rescue loops and the Sclip consensus
// counts of process_softclip (5' and 3', a low-quality stop, the region check), worked out by hand.
// The MNV join of the M loop is pinned by tests/test_oracle_realign_kat.py.
// Parity unpinned: findOffset / appendSegments on long complex M-run descriptors and the
// insertion-anchor bookkeeping have no known answer of their own; agreement with the engine's
// separate formulation on synthetic data is their only check.
#include <algorithm>
#include <cstdlib>
#include <regex>
#include <stdexcept>

#include "rv_oracle.h"

namespace orc {

bool modify_cigar_record(const RefMap& ref, int& position, std::vector<uint32_t>& rec, int n_orig,
                         int& base, int& n_new, const std::string& seq, const uint8_t* qual, int lseq,
                         int indelLen);

static int nt16_code(char c) {
    static const char* t = "=ACMGRSVTWYHKDBN";
    for (int i = 0; i < 16; i++)
        if (t[i] == c) return i;
    return 15;
}

void preprocess_filter(const Config& c, const std::vector<const Read*>& all, std::vector<const Read*>& out,
                       Counters& ctr) {
    // recordPreprocessor.cpp:121-176.  -Z (:133) draws the C library's rand() once per record that
    // passed the flag filter; the reference never seeds it, so the draws continue one process-wide
    // sequence (this process: one region, the first sample before the second).
    int firstMatchingPosition = 0;
    std::set<std::string> dups;
    for (const Read* rp : all) {
        const Read& r = *rp;
        ctr.reads_in++;
        if (r.flag & c.samfilter) {
            ctr.reads_filtered++;
            continue;
        }
        if (c.isDownsampling && rand() <= RAND_MAX * c.downsampling) {
            ctr.reads_filtered++;
            continue;
        }
        if (r.mapq < c.mappingQuality) {
            ctr.reads_filtered++;
            continue;
        }
        if (r.seq.size() == 1) {
            ctr.reads_filtered++;
            continue;
        }
        if (c.removeDuplicatedReads) {
            int as = r.pos0 + 1, ms = r.mpos0 + 1;
            // getMateReferenceName (:80-90); a contig other than the read's own is named by its id
            std::string mref = !(r.flag & 1) ? "*" : (r.tid == r.mtid ? "=" : "tid" + std::to_string(r.mtid));
            if (as != firstMatchingPosition) dups.clear();
            if (ms < 10) {
                std::string k = std::to_string(as) + "-" + mref + "-" + std::to_string(ms);
                if (dups.count(k)) {
                    ctr.reads_filtered++;
                    continue;
                }
                dups.insert(k);
                firstMatchingPosition = as;
            } else if ((r.flag & 1) && (r.flag & 4)) {
                std::string cs;
                // get_cigar_string (VariationUtils.h:563-570) through htslib's bam_cigar_opchr
                static const char* ops = "MIDNSHP=XB??????";
                for (uint32_t w : r.cigar) cs += std::to_string(clen(w)) + ops[cop(w)];
                std::string k = std::to_string(as) + "-" + cs;
                if (dups.count(k)) {
                    ctr.reads_filtered++;
                    continue;
                }
                dups.insert(k);
                firstMatchingPosition = as;
            }
        }
        out.push_back(&r);
    }
}

namespace {

inline bool isATGC(char ch) { return ch == 'A' || ch == 'T' || ch == 'G' || ch == 'C'; }

// parseCigar.cpp:480-495
bool is_mnp_desc(const std::string& s) {
    if (s.size() > 2 && s[1] == '&' && isATGC(s[0])) {
        for (size_t i = 2; i < s.size(); i++)
            if (!isATGC(s[i])) return false;
        return true;
    }
    return false;
}

struct Offset {
    int offset = 0;
    std::string sequence;
    long q_sum = 0;
    int nm = 0;
};

struct Walker {
    const Config& c;
    const Region& R;
    RefMap& ref;
    int chrLen;
    Pileup& P;

    // per read
    const Read* rd = nullptr;
    std::vector<uint32_t> rec;  // emulated record: CIGAR words then packed SEQ words
    int n_cigar = 0;
    std::string seq;            // decoded SEQ followed by a NUL (see substr())
    std::vector<uint8_t> qual;
    uint8_t qual_before0 = 0;   // byte preceding QUAL in the record (last packed SEQ byte)
    int lseq = 0;
    int start = 0, offset = 0, cel = 0, rpi = 0, rpe = 0;
    uint8_t mapq = 0;
    int nm = 0;
    bool dir = false;
    int position = 0;

    uint32_t cw(int i) const { return (i >= 0 && (size_t)i < rec.size()) ? rec[(size_t)i] : 0u; }
    int op(int i) const { return cop(cw(i)); }
    int len(int i) const { return clen(cw(i)); }
    int q(int i) const {
        if (i >= 0) return qual[(size_t)i];
        if (i == -1) return qual_before0;
        throw std::out_of_range("qual index");
    }
    // std::string(char*, pos, len) on the decoded VLA: a temporary string up to the first NUL,
    // then substr(pos, len)
    std::string sub(int pos, int n) const {
        if (pos < 0 || pos > lseq) throw std::out_of_range("substr");
        return seq.substr((size_t)pos, (size_t)std::max(0, n));
    }
    char sq(int i) const {
        if (i < 0 || i >= (int)seq.size()) return '\0';
        return seq[(size_t)i];
    }

    bool hasEq(int k, char ch) const { return ref.count(k) && ref.at(k) == ch; }
    bool hasNe(int k, char ch) const { return ref.count(k) && ref.at(k) != ch; }

    void addCnt(Var* v, int rp, double bq) {
        v->cnt++;
        if (dir)
            v->rev++;
        else
            v->fwd++;
        v->pmean += rp;
        v->qmean += bq;
        v->mqmean += mapq;
        v->nm += nm;
        if (bq >= (double)(float)c.goodq)
            v->hicnt++;
        else
            v->locnt++;
    }

    bool isNextMatched(int ci) const { return c.performLocalRealignment && n_cigar > ci + 1 && op(ci + 1) == OP_M; }
    bool isNextInsertion(int ci) const {
        return c.performLocalRealignment && n_cigar > ci + 1 && op(ci + 1) == OP_I;
    }
    bool isNextAfterNumMatched(int ci, int number) const { return n_cigar > ci + number && op(ci + 1) == OP_M; }
    bool isIDWithNextMatched(int ci) const {
        return c.performLocalRealignment && n_cigar > ci + 2 && len(ci + 1) <= c.vext && op(ci + 1) == OP_M &&
               (op(ci + 2) == OP_I || op(ci + 2) == OP_D) && op(ci + 3) != OP_I && op(ci + 3) != OP_D;
    }
    int tp_of(int rlen1) const { return rpe < rlen1 - rpe ? rpe + 1 : rlen1 - rpe; }

    Offset findOffset(int refPos, int readPos, int cigLen) {
        Offset o;
        int vsn = 0;
        for (int vi = 0; vsn <= c.vext && vi < cigLen; vi++) {
            if (sq(readPos + vi) == 'N') break;
            if (q(readPos + vi) < c.goodq) break;
            if (ref.count(refPos + vi)) {
                char rc = ref.get(refPos + vi);
                if (sq(readPos + vi) != rc) {
                    o.offset = vi + 1;
                    o.nm++;
                    vsn = 0;
                } else {
                    vsn++;
                }
            }
        }
        if (o.offset > 0) {
            o.sequence = sub(readPos, o.offset);
            for (int i = readPos; i < readPos + o.offset; i++) o.q_sum += q(i);
            for (int k = 0; k < o.offset; k++) P.refCov[refPos + k]++;
        }
        return o;
    }

    void appendSegments(int ci, std::string& desc, long& qs, int& qc, int mLen, int indelLen, int begin,
                        bool isIns) {
        std::string d = "#" + sub(begin, mLen);
        for (int i = begin; i < begin + mLen; i++) qs += q(i);
        if (op(ci + 2) == OP_I)
            d += "^" + sub(begin + mLen, indelLen);
        else
            d += "^" + std::to_string(indelLen);
        if (isIns) {
            if (op(ci + 2) == OP_I) {
                for (int i = begin + mLen; i < begin + mLen + indelLen; i++) qs += q(i);
                qc += indelLen;
            } else {
                qs += q(begin + mLen);
                qc += 1;
            }
        } else {
            if (op(ci + 2) == OP_I) {
                for (int i = begin + mLen; i < begin + mLen + indelLen; i++) qs += q(i);
                qc += indelLen;
            }
        }
        desc += d;
    }

    void softclip(int ci, int totalLen) {
        if (ci == 0) {
            while (cel - 1 >= 0 && start - 1 > 0 && start - 1 <= chrLen && hasEq(start - 1, sq(cel - 1)) &&
                   q(cel - 1) > 10) {
                Var* v = P.getVariation(P.nonIns, start - 1, std::string(1, ref.get(start - 1)));
                addCnt(v, cel, q(cel - 1));
                P.refCov[start - 1]++;
                start--;
                cel--;
            }
            if (cel > 0) {
                int sumq = 0, nhq = 0, nlq = 0;
                for (int si = cel - 1; si >= 0; si--) {
                    if (sq(si) == 'N') break;
                    int bq = q(si);
                    if (bq <= 12) nlq++;
                    if (nlq > 1) break;
                    sumq += bq;
                    nhq++;
                }
                if (nhq >= 1 && nhq > nlq && start >= R.start && start <= R.end) {
                    Sclip** slot = P.sc5.find(start);
                    Sclip* sc;
                    if (!slot) {
                        sc = P.newSclip();
                        P.sc5[start] = sc;
                    } else {
                        sc = *slot;
                    }
                    for (int si = cel - 1; cel - si <= nhq; si--) {
                        char ch = sq(si);
                        int idx = cel - 1 - si;
                        sc->nt[idx][ch]++;
                        Var* sv = seq_var(sc, idx, ch);
                        addCnt(sv, si - (cel - nhq), q(si));
                    }
                    addCnt(sc, cel, sumq / (double)nhq);
                }
            }
            cel = len(ci);
        } else if (ci == n_cigar - 1) {
            while (rpi < lseq && hasEq(start, sq(rpi)) && q(rpi) > 10) {
                Var* v = P.getVariation(P.nonIns, start, std::string(1, ref.at(start)));
                addCnt(v, totalLen - rpe, q(rpi));
                P.refCov[start]++;
                rpi++;
                start++;
                cel--;
                rpe++;
            }
            if (rpi < lseq) {
                int sumq = 0, nhq = 0, nlq = 0;
                for (int si = 0; si < cel; si++) {
                    if (sq(rpi + si) == 'N') break;
                    int bq = q(rpi + si);
                    if (bq <= 12) nlq++;
                    if (nlq > 1) break;
                    sumq += bq;
                    nhq++;
                }
                if (nhq >= 1 && nhq > nlq && start >= R.start && start <= R.end) {
                    Sclip** slot = P.sc3.find(start);
                    Sclip* sc;
                    if (!slot) {
                        sc = P.newSclip();
                        P.sc3.insert(start, sc);
                    } else {
                        sc = *slot;
                    }
                    for (int si = 0; si < nhq; si++) {
                        char ch = sq(rpi + si);
                        sc->nt[si][ch]++;
                        Var* sv = seq_var(sc, si, ch);
                        addCnt(sv, nhq - si, q(rpi + si));
                    }
                    addCnt(sc, cel, sumq / (double)nhq);
                }
            }
        }
        rpi += cel;
        offset = 0;
        start = position;
    }

    // getVariationFromSeq (parseCigar.cpp:333-362)
    Var* seq_var(Sclip* sc, int idx, char ch) {
        auto it = sc->seq.find(idx);
        if (it == sc->seq.end()) {
            Var* v = P.newVar();
            sc->seq[idx][ch] = v;
            return v;
        }
        auto j = it->second.find(ch);
        if (j != it->second.end()) return j->second;
        Var* v = P.newVar();
        it->second[ch] = v;
        return v;
    }

    int insertion(int ci, int rlen1) {
        if ((n_cigar > ci + 1 && op(ci + 1) == OP_N) || (ci > 0 && op(ci - 1) == OP_N)) {
            rpi += cel;
            return ci;
        }
        std::string desc = sub(rpi, cel);
        long qs = 0;
        int qc = 0;
        for (int i = rpi; i < rpi + cel; i++) qs += q(i);
        qc += cel;
        std::string ss;
        int multoffs = 0, multoffp = 0, nmoff = 0;
        if (isIDWithNextMatched(ci)) {
            int mLen = len(ci + 1), indelLen = len(ci + 2);
            int begin = rpi + cel;
            appendSegments(ci, desc, qs, qc, mLen, indelLen, begin, true);
            multoffs += mLen + (op(ci + 2) == OP_D ? indelLen : 0);
            multoffp += mLen + (op(ci + 2) == OP_I ? indelLen : 0);
            int ci6 = n_cigar > ci + 3 ? len(ci + 3) : 0;
            if (ci6 != 0 && op(ci + 3) == OP_M) {
                Offset o = findOffset(start + multoffs, rpi + cel + multoffp, ci6);
                offset = o.offset;
                ss = o.sequence;
                qs += o.q_sum;
                qc += offset;
            }
            ci += 2;
        } else if (isNextMatched(ci)) {
            int vsn = 0;
            for (int vi = 0; vsn <= c.vext && vi < len(ci + 1); vi++) {
                if (sq(rpi + cel + vi) == 'N') break;
                if (q(rpi + cel + vi) < c.goodq) break;
                if (ref.count(start + vi)) {
                    if (sq(rpi + cel + vi) != ref.get(start + vi)) {
                        offset = vi + 1;
                        nmoff++;
                        vsn = 0;
                    } else {
                        vsn++;
                    }
                }
            }
            if (offset != 0) {
                ss += sub(rpi + cel, offset);
                int t0 = rpi + cel;
                for (int i = t0; i < t0 + offset; i++) qs += q(i);
                qc += offset;
                for (int k = 0; k < offset; k++) P.refCov[start + k]++;
            }
        }
        if (offset > 0) desc += "&" + ss;
        if (start - 1 >= R.start && start - 1 <= R.end && desc.find('N') == std::string::npos) {
            int inspos = start - 1;
            bool allATGC = !desc.empty();
            for (char ch : desc)
                if (!isATGC(ch)) allATGC = false;
            if (allATGC) {
                // adjInsPos rotates the caller's string in place
                int bi = start - 1;
                int n = 1;
                int L = (int)desc.size();
                while (ref.get(bi) == desc[(size_t)(L - n)]) {
                    n++;
                    if (n > L) n = 1;
                    bi--;
                }
                if (n > 1) desc = desc.substr((size_t)(L - (n - 1))) + desc.substr(0, (size_t)(L - (n - 1)));
                if (rpi - 1 - (start - 1 - bi) > 0) inspos = bi;
            }
            std::string key = "+" + desc;
            P.insCount[inspos][key]++;
            Var* hv = P.getVariation(P.ins, inspos, key);
            hv->incDir(dir);
            hv->cnt++;
            int tp = tp_of(rlen1);
            double tmpq = (double)qs / qc;
            if (!hv->pstd && hv->pp != 0 && tp != hv->pp) hv->pstd = true;
            if (!hv->qstd && hv->pq != 0 && tmpq != hv->pq) hv->qstd = true;
            hv->pmean += tp;
            hv->qmean += tmpq;
            hv->mqmean += mapq;
            hv->pp = tp;
            hv->pq = tmpq;
            if (tmpq >= c.goodq)
                hv->hicnt++;
            else
                hv->locnt++;
            hv->nm += nm - nmoff;
            int index = rpi - 1 - (start - 1 - inspos);
            if (inspos > position && hasEq(inspos, sq(index))) {
                Var* tv = P.getVariationMaybe(inspos, sq(index));
                if (tv) {
                    tv->cnt--;
                    tv->subDir(dir, 1);
                    tv->pmean -= tp;
                    tv->qmean -= q(index);
                    tv->mqmean -= mapq;
                    tv->nm -= (nm - nmoff);
                    if (q(index) >= c.goodq)
                        tv->hicnt--;
                    else
                        tv->locnt--;
                }
            }
            if (ci == 1 && (op(0) == OP_S || op(0) == OP_H)) {
                Var* tt = P.getVariation(P.nonIns, inspos, std::string(1, ref.get(inspos)));
                tt->incDir(dir);
                tt->cnt++;
                tt->pstd = hv->pstd;
                tt->qstd = hv->qstd;
                tt->pmean += tp;
                tt->qmean += tmpq;
                tt->mqmean += mapq;
                tt->pp = tp;
                tt->pq = tmpq;
                tt->nm += nm - nmoff;
                P.refCov[inspos]++;
            }
        }
        rpi += cel + offset + multoffp;
        rpe += cel + offset + multoffp;
        start += offset + multoffs;
        return ci;
    }

    int deletion(int ci, int rlen1) {
        if ((n_cigar > ci && op(ci + 1) == OP_N) || (ci > 1 && op(ci - 1) == OP_N)) {
            rpi += cel;
            return ci;
        }
        std::string desc = "-" + std::to_string(cel);
        std::string app;
        int qlast = q(rpi - 1);
        long qs = 0;
        int qc = 0;
        int multoffs = 0, multoffp = 0, nmoff = 0;
        if (isIDWithNextMatched(ci)) {
            int mLen = len(ci + 1), indelLen = len(ci + 2);
            appendSegments(ci, desc, qs, qc, mLen, indelLen, rpi, false);
            multoffs += mLen + (op(ci + 2) == OP_D ? indelLen : 0);
            multoffp += mLen + (op(ci + 2) == OP_I ? indelLen : 0);
            if (isNextAfterNumMatched(ci, 3)) {
                int vsn = 0;
                int tn = rpi + multoffp;
                int ts = start + multoffs + cel;
                for (int vi = 0; vsn <= c.vext && vi < len(ci + 3); vi++) {
                    if (sq(tn + vi) == 'N') break;
                    if (q(tn + vi) < c.goodq) break;
                    if (hasEq(ts + vi, 'N')) break;
                    if (ref.count(ts + vi)) {
                        char rc = ref.get(ts + vi);
                        if (sq(tn + vi) != rc) {
                            offset = vi + 1;
                            nmoff++;
                            vsn = 0;
                        } else {
                            vsn++;
                        }
                    }
                }
                if (offset != 0) {
                    app += sub(tn, offset);
                    for (int i = tn; i < tn + offset; i++) qs += q(i);
                    qc += offset;
                }
            }
            ci += 2;
        } else if (isNextInsertion(ci)) {
            int insLen = len(ci + 1);
            desc += "^" + sub(rpi, insLen);
            for (int i = rpi; i < rpi + insLen; i++) qs += q(i);
            qc += insLen;
            multoffp += insLen;
            // isNextAfterNumMatched(ci, 2) tests CIGAR[ci+1] (the insertion) for M: never true
            ci += 1;
        } else if (isNextMatched(ci)) {
            int mLen = len(ci + 1);
            int vsn = 0;
            for (int vi = 0; vsn <= c.vext && vi < mLen; vi++) {
                char sc = sq(rpi + vi);
                if (sc == 'N') break;
                if (q(rpi + vi) < c.goodq) break;
                if (ref.count(start + cel + vi)) {
                    char rc = ref.get(start + cel + vi);
                    if (rc == 'N') break;
                    if (sc != rc) {
                        offset = vi + 1;
                        nmoff++;
                        vsn = 0;
                    } else {
                        vsn++;
                    }
                }
            }
            if (offset != 0) {
                app += sub(rpi, offset);
                for (int i = rpi; i < rpi + offset; i++) qs += q(i);
                qc += offset;
            }
        }
        if (offset > 0) desc += "&" + app;
        if (rpi + offset >= lseq) {
            qs += qlast;
            qc += 1;
        } else {
            int qo = q(rpi + offset);
            qs += qlast > qo ? qlast : qo;
            qc += 1;
        }
        if (start >= R.start && start <= R.end) {
            Var* hv = P.getVariation(P.nonIns, start, desc);
            P.delCount[start][desc]++;
            hv->incDir(dir);
            hv->cnt++;
            int tp = tp_of(rlen1);
            double tmpq = (double)qs / qc;
            if (!hv->pstd && hv->pp != 0 && tp != hv->pp) hv->pstd = true;
            if (!hv->qstd && hv->pq != 0 && tmpq != hv->pq) hv->qstd = true;
            hv->pmean += tp;
            hv->qmean += tmpq;
            hv->mqmean += mapq;
            hv->pp = tp;
            hv->pq = tmpq;
            hv->nm += nm - nmoff;
            if (tmpq >= c.goodq)
                hv->hicnt++;
            else
                hv->locnt++;
            for (int i = 0; i < cel; i++) P.refCov[start + i]++;
        }
        start += cel + offset + multoffs;
        rpi += offset + multoffp;
        rpe += offset + multoffp;
        return ci;
    }

    void match_run(int ci, int c_op, int rlen1, int totalLen) {
        (void)totalLen;
        int nmoff = 0, moffset = 0;
        for (int i = offset; i < cel; i++) {
            bool trim = false;
            if (c.trimBasesAfter != 0) {
                trim = !dir ? rpi > c.trimBasesAfter : totalLen - rpi > c.trimBasesAfter;
            }
            const char ch1 = sq(rpi);
            std::string s(1, ch1);
            bool swd = false;
            if (ch1 == 'N') {
                start++;
                rpi++;
                rpe++;
                continue;
            }
            double qv = q(rpi);
            int qbases = 1, qibases = 0;
            std::string ss;
            while ((start + 1) >= R.start && (start + 1) <= R.end && (i + 1) < cel && qv >= c.goodq &&
                   hasNe(start, sq(rpi)) && ref.get(start) != 'N') {
                if (q(rpi + 1) < c.goodq + 5) break;
                char nuc = sq(rpi + 1);
                if (nuc == 'N') break;
                if (hasEq(start + 1, 'N')) break;
                if (ref.get(start + 1) != nuc) {
                    ss += nuc;
                    qv += q(rpi + 1);
                    qbases++;
                    rpi++;
                    rpe++;
                    i++;
                    start++;
                    nmoff++;
                } else {
                    int ssn = 0;
                    for (int ssi = 1; ssi <= c.vext; ssi++) {
                        if (i + 1 + ssi >= cel) break;
                        if (rpi + 1 + ssi < lseq && hasNe(start + 1 + ssi, sq(rpi + 1 + ssi))) {
                            ssn = ssi + 1;
                            break;
                        }
                    }
                    if (ssn == 0) break;
                    if (q(rpi + ssn) < c.goodq + 5) break;
                    for (int ssi = 1; ssi <= ssn; ssi++) {
                        ss += sq(rpi + ssi);
                        qv += q(rpi + ssi);
                        qbases++;
                    }
                    rpi += ssn;
                    rpe += ssn;
                    i += ssn;
                    start += ssn;
                }
            }
            if (!ss.empty()) s = s + "&" + ss;
            int ddlen = 0;
            if (c.performLocalRealignment && cel - i <= c.vext && ci + 1 < n_cigar && op(ci + 1) == OP_D &&
                ref.count(start) && (!ss.empty() || sq(rpi) != ref.get(start)) && q(rpi) >= c.goodq) {
                while (i + 1 < cel) {
                    s += sq(rpi + 1);
                    qv += q(rpi + 1);
                    qbases++;
                    i++;
                    rpi++;
                    rpe++;
                    start++;
                }
                size_t amp = s.find('&');
                if (amp != std::string::npos) s.erase(amp, 1);
                s = "-" + std::to_string(len(ci + 1)) + "&" + s;
                swd = true;
                ddlen = len(ci + 1);
                ci += 1;
                if (n_cigar > ci + 1 && op(ci + 1) == OP_I) {
                    s += "^" + sub(rpi + 1, len(ci + 1));
                    int next_len = len(ci + 1);
                    for (int qi = 1; qi <= next_len; qi++) {
                        qv += q(rpi + 1 + qi);
                        qibases++;
                    }
                    rpi += next_len;
                    rpe += next_len;
                    ci += 1;
                }
                if (isNextAfterNumMatched(ci, 1)) {
                    Offset o = findOffset(start + ddlen + 1, rpi + 1, len(ci + 1));
                    if (o.offset != 0) {
                        moffset = o.offset;
                        nmoff += o.nm;
                        s += "&" + o.sequence;
                        // the offset bases' qualities are not added (their string is left empty)
                    }
                }
            } else if (c.performLocalRealignment && cel - i <= c.vext && ci + 1 < n_cigar && op(ci + 1) == OP_I &&
                       ref.count(start) && (!ss.empty() || sq(rpi) != ref.get(start)) && q(rpi) >= c.goodq) {
                while (i + 1 < cel) {
                    s += sq(rpi + 1);
                    qv += q(rpi + 1);
                    qbases++;
                    i++;
                    rpi++;
                    rpe++;
                    start++;
                }
                size_t amp = s.find('&');
                if (amp != std::string::npos) s.erase(amp, 1);
                int nextLen = len(ci + 1);
                s += sub(rpi + 1, nextLen);
                s.insert((size_t)nextLen, "&");
                s = "+" + s;
                for (int qi = 1; qi <= nextLen; qi++) {
                    qv += q(rpi + 1 + qi);
                    qibases++;
                }
                rpi += nextLen;
                rpe += nextLen;
                ci += 1;
                qibases--;
                qbases++;
            }
            if (!trim) {
                const int pos = start - qbases + 1;
                if (pos >= R.start && pos <= R.end) {
                    Var* hv = P.getVariation(P.nonIns, pos, s);
                    hv->incDir(dir);
                    if (is_mnp_desc(s)) P.mnp[pos][s]++;
                    hv->cnt++;
                    int tp = tp_of(rlen1);
                    qv = qv / (qbases + qibases);
                    if (!hv->pstd && hv->pp != 0 && tp != hv->pp) hv->pstd = true;
                    if (!hv->qstd && hv->pq != 0 && qv != hv->pq) hv->qstd = true;
                    hv->pmean += tp;
                    hv->qmean += qv;
                    hv->mqmean += mapq;
                    hv->pp = tp;
                    hv->pq = qv;
                    hv->nm += nm - nmoff;
                    if (qv >= c.goodq)
                        hv->hicnt++;
                    else
                        hv->locnt++;
                    for (int qi = 1; qi <= qbases; qi++) P.refCov[start - qi + 1]++;
                    if (swd) {
                        P.delCount[pos][s]++;
                        for (int qi = 1; qi < ddlen; qi++) P.refCov[start + qi]++;
                    }
                }
            }
            if (swd) start += ddlen;
            if (c_op != OP_I) start++;
            if (c_op != OP_D) {
                rpi++;
                rpe++;
            }
        }
        if (moffset != 0) {
            offset = moffset;
            rpi += moffset;
            start += moffset;
            rpe += moffset;
        }
        cur_ci = ci;
    }
    int cur_ci = 0;

    void parse(const Read& r) {
        rd = &r;
        lseq = (int)r.seq.size();
        seq = r.seq;
        qual = r.qual;
        mapq = r.mapq;
        int n0 = (int)r.cigar.size();
        if (n0 <= 0) return;
        // emulated record: CIGAR words, then packed SEQ bytes as little-endian words
        int seq_bytes = (lseq + 1) / 2;
        std::vector<uint8_t> packed((size_t)seq_bytes, 0);
        for (int i = 0; i < lseq; i++) {
            int code = nt16_code(r.seq[(size_t)i]);
            if (i % 2 == 0)
                packed[(size_t)(i / 2)] = (uint8_t)(code << 4);
            else
                packed[(size_t)(i / 2)] |= (uint8_t)code;
        }
        qual_before0 = seq_bytes > 0 ? packed[(size_t)seq_bytes - 1] : 0;
        rec.assign((size_t)n0 + (size_t)(seq_bytes + 3) / 4, 0u);
        for (int i = 0; i < n0; i++) rec[(size_t)i] = r.cigar[(size_t)i];
        for (int b = 0; b < seq_bytes; b++) rec[(size_t)n0 + (size_t)(b / 4)] |= (uint32_t)packed[(size_t)b] << (8 * (b % 4));
        n_cigar = n0;
        int idlen = 0;
        for (int i = 0; i < n0; i++)
            if (cop(r.cigar[(size_t)i]) == OP_I || cop(r.cigar[(size_t)i]) == OP_D) idlen += clen(r.cigar[(size_t)i]);
        int tnm = 0;
        if (r.has_nm) {
            tnm = r.nm - idlen;
            if (tnm > c.mismatch) return;
        } else {
            if (r.flag & 4) return;
        }
        nm = tnm;
        dir = (r.flag & 0x10) != 0;
        position = r.pos0 + 1;
        rpi = rpe = 0;
        if (c.performLocalRealignment) {
            int base = 0, nn = 0;
            int p = position;
            if (modify_cigar_record(ref, p, rec, n0, base, nn, seq, qual.data(), lseq, idlen)) {
                n_cigar = nn;
                position = p;
            }
        }
        // cleanupCigar (parseCigar.cpp:1858-1892); type 3 = M, =, X
        auto ctype3 = [](int o) { return o == OP_M || o == OP_EQ || o == OP_X; };
        {
            bool nm_yet = true;
            for (int i = 0; i < n_cigar && nm_yet; i++) {
                if (op(i) == OP_I)
                    rec[(size_t)i] = (rec[(size_t)i] & ~0xfu) | OP_S;
                else if (op(i) == OP_H) {
                } else if (ctype3(op(i)))
                    nm_yet = false;
            }
            nm_yet = true;
            for (int i = n_cigar - 1; i >= 0 && nm_yet; i--) {
                if (op(i) == OP_I)
                    rec[(size_t)i] = (rec[(size_t)i] & ~0xfu) | OP_S;
                else if (op(i) == OP_H) {
                } else if (ctype3(op(i)))
                    nm_yet = false;
            }
        }
        start = position;
        offset = 0;
        if (op(0) == OP_S && len(0) >= 10 && op(n_cigar - 1) == OP_S && len(n_cigar - 1) >= 10) return;
        int rlen1 = 0, totalLen = 0;
        for (int i = 0; i < n_cigar; i++) {
            int o = op(i);
            if (o == OP_M || o == OP_I) rlen1 += len(i);
            if (o == OP_M || o == OP_I || o == OP_S) totalLen += len(i);
        }
        if (c.minmatch != 0 && rlen1 < c.minmatch) return;
        if (totalLen > P.maxReadLength) P.maxReadLength = totalLen;
        if (r.flag & 2048) return;
        int mateStart = r.mpos0 + 1;
        bool need_break = true;
        for (int ci = 0; ci < n_cigar; ++ci) {
            if (need_break && skip_overlap(mateStart)) break;
            need_break = false;
            int c_op = op(ci);
            cel = len(ci);
            if ((ci == 0 || ci == n_cigar - 1) && c_op == OP_I) c_op = OP_S;
            switch (c_op) {
                case OP_N: {
                    std::string key = std::to_string(start - 1) + "-" + std::to_string(start + cel - 1);
                    P.splice.insert(key);
                    start += cel;
                    offset = 0;
                    continue;
                }
                case OP_S:
                    softclip(ci, totalLen);
                    continue;
                case OP_H:
                    offset = 0;
                    continue;
                case OP_I:
                    offset = 0;
                    ci = insertion(ci, rlen1);
                    continue;
                case OP_D:
                    offset = 0;
                    ci = deletion(ci, rlen1);
                    continue;
                default:
                    break;
            }
            match_run(ci, c_op, rlen1, totalLen);
            ci = cur_ci;
            if (start > R.end) break;
        }
    }

    bool skip_overlap(int mateStart) const {
        bool pairedSame = (rd->flag & 1) && rd->tid == rd->mtid;
        if (c.uniqueModeAlignmentEnabled && pairedSame && !dir && start >= mateStart) return true;
        if (c.uniqueModeSecondInPairEnabled && (rd->flag & 1) && pairedSame) {
            // isReadsOverlap (parseCigar.cpp:196-206)
            int reflen = 0;
            for (uint32_t w : rd->cigar) {
                int o = cop(w);
                if (o == OP_M || o == OP_D || o == OP_N || o == OP_EQ || o == OP_X) reflen += clen(w);
            }
            int endpos = rd->pos0 + (reflen > 0 ? reflen : 1);
            if (position >= mateStart) return start >= mateStart && start <= mateStart + reflen - 1;
            return start >= mateStart && mateStart <= endpos;
        }
        return false;
    }
};

}  // namespace

void pileup_region(const Config& c, const Region& r, RefMap& ref, int chrLen, const std::vector<const Read*>& reads,
                   Pileup& P, Counters& ctr) {
    Walker w{c, r, ref, chrLen, P};
    for (const Read* rd : reads) {
        try {
            w.parse(*rd);
        } catch (const std::exception&) {
            ctr.exceptions++;
        }
        ctr.reads_parsed++;
    }
}

}  // namespace orc
