// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// Scoring and simple-mode output restated from
//   src/ToVarsBuilder.cpp:97-205      (process: per-position loop in nonInsertionVariants order)
//   src/ToVarsBuilder.cpp:213-233     (isTheSameVariationOnRef)
//   src/ToVarsBuilder.cpp:242-317     (proceedVrefIsDeletion / proceedVrefIsInsertion)
//   src/ToVarsBuilder.cpp:327-346     (collectVarsAtPosition)
//   src/ToVarsBuilder.cpp:358-536     (createInsertion / createVariant)
//   src/ToVarsBuilder.cpp:543-587     (adjustVariantCounts / calcHicov)
//   src/ToVarsBuilder.cpp:656-722     (findMSI)
//   src/ToVarsBuilder.cpp:730-1116    (collectReferenceVariants / validateRefallele)
//   include/Variant.h:75-248          (adjComplex, varType, isGoodVar)
//   include/VariationUtils.h:511-521  (strandBias)
//   src/modes/simpleMode.cpp:66-208   (print_output_variant_simple / output)
// and the two-sided Fisher exact test of htslib's kt_fisher_exact (htslib kfunc.c, MIT; the
// reference calls it at src/modes/simpleMode.cpp:98).  htslib is not vendored by the reference and
// its version is unpinned (install.sh clones the default branch); the restatement below follows
// the published algorithm (kf_lgamma: Lanczos/AS245 form; hypergeo_acc: incremental recurrence
// re-anchored every 11 steps; tails walked inward while p < 0.99999999*q, boundary included when
// p < 1.00000001*q).
//
// Pinned: kt_fisher_two_sided by tests/test_oracle_fisher.py (exact rational enumeration, symmetry,
// scipy); the printed fields of SNV / deletion / insertion calls by tests/test_oracle_cases.py.
// findMSI by tests/test_oracle_findmsi_kat.py (hand-worked microsatellite counts and 3' shifts);
// the MNV allele printed from an '&' descriptor by tests/test_oracle_realign_kat.py; a deletion
// followed by an insertion ("-2^XY") and by a mismatch ("-2&T") by tests/test_oracle_allele_kat.py.
// Parity unpinned: the '#' form (an indel, at most vext matched bases, another indel) has no known
// answer: with local realignment on, the CIGAR rewrite merges every such run (M <= 15 between the
// indels) into one deletion plus insertion before the walk, and with it off the walk never builds
// the form, so the default configuration does not reach it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <regex>
#include <set>
#include <stdexcept>

#include "rv_oracle.h"

namespace orc {

std::string vc_substr1(const std::string& s, int idx);
std::string vc_substr2(const std::string& s, int begin, int len);
std::string joinRef(const RefMap& ref, int from, int to);
void replaceFirst(std::string& s, const std::string& a, const std::string& b);
void replaceFirstRe(std::string& s, const std::string& pat, const std::string& b);

std::string to_string_f(double v) {
    char buf[512];
    snprintf(buf, sizeof buf, "%f", v);
    return buf;
}

// ------------------------------------------------------------------ Fisher (htslib kfunc.c)
static double kf_lgamma(double z) {
    double x = 0;
    x += 0.1659470187408462e-06 / (z + 7);
    x += 0.9934937113930748e-05 / (z + 6);
    x -= 0.1385710331296526 / (z + 5);
    x += 12.50734324009056 / (z + 4);
    x -= 176.6150291498386 / (z + 3);
    x += 771.3234287757674 / (z + 2);
    x -= 1259.139216722289 / (z + 1);
    x += 676.5203681218835 / z;
    x += 0.9999999999995183;
    return std::log(x) - 5.58106146679532777 - z + (z - 0.5) * std::log(z + 6.5);
}
static double lbinom(int n, int k) {
    if (k == 0 || n == k) return 0;
    return kf_lgamma(n + 1) - kf_lgamma(k + 1) - kf_lgamma(n - k + 1);
}
static double hypergeo(int n11, int n1_, int n_1, int n) {
    return std::exp(lbinom(n1_, n11) + lbinom(n - n1_, n_1 - n11) - lbinom(n, n_1));
}
struct HGAcc {
    int n11, n1_, n_1, n;
    double p;
};
static double hypergeo_acc(int n11, int n1_, int n_1, int n, HGAcc* a) {
    if (n1_ || n_1 || n) {
        a->n11 = n11;
        a->n1_ = n1_;
        a->n_1 = n_1;
        a->n = n;
    } else {
        if (n11 % 11 && n11 + a->n - a->n1_ - a->n_1) {
            if (n11 == a->n11 + 1) {
                a->p *= (double)(a->n1_ - a->n11) / n11 * (a->n_1 - a->n11) / (n11 + a->n - a->n1_ - a->n_1);
                a->n11 = n11;
                return a->p;
            }
            if (n11 == a->n11 - 1) {
                a->p *= (double)a->n11 / (a->n1_ - n11) * (a->n11 + a->n - a->n1_ - a->n_1) / (a->n_1 - n11);
                a->n11 = n11;
                return a->p;
            }
        }
        a->n11 = n11;
    }
    a->p = hypergeo(a->n11, a->n1_, a->n_1, a->n);
    return a->p;
}
double kt_fisher_two_sided(int n11, int n12, int n21, int n22, double* _left, double* _right) {
    int i, j, max, min;
    double p, q, left, right, two;
    HGAcc aux{0, 0, 0, 0, 0};
    int n1_ = n11 + n12, n_1 = n11 + n21, n = n11 + n12 + n21 + n22;
    max = (n_1 < n1_) ? n_1 : n1_;
    min = n1_ + n_1 - n;
    if (min < 0) min = 0;
    if (_left) *_left = 1.;
    if (_right) *_right = 1.;
    if (min == max) return 1.;
    q = hypergeo_acc(n11, n1_, n_1, n, &aux);
    p = hypergeo_acc(min, 0, 0, 0, &aux);
    for (left = 0., i = min + 1; p < 0.99999999 * q && i <= max; ++i) left += p, p = hypergeo_acc(i, 0, 0, 0, &aux);
    --i;
    if (p < 1.00000001 * q)
        left += p;
    else
        --i;
    p = hypergeo_acc(max, 0, 0, 0, &aux);
    for (right = 0., j = max - 1; p < 0.99999999 * q && j >= 0; --j) right += p, p = hypergeo_acc(j, 0, 0, 0, &aux);
    ++j;
    if (p < 1.00000001 * q)
        right += p;
    else
        ++j;
    two = left + right;
    if (two > 1.) two = 1.;
    if (std::abs(i - n11) < std::abs(j - n11))
        right = 1. - left + q;
    else
        left = 1.0 - right + q;
    if (_left) *_left = left;
    if (_right) *_right = right;
    return two;
}

namespace {

bool starts_with(const std::string& s, const std::string& t) { return s.find(t) == 0; }

int strandBias(int f, int r, const Config& c) {
    if (f + r <= 12) return f * r > 0 ? 2 : 0;
    return (f / (double)(f + r) >= c.bias && r / (double)(f + r) >= c.bias && f >= c.minBiasReads &&
            r >= c.minBiasReads)
               ? 2
               : 1;
}

struct MSI {
    double msi;
    int shift3;
    std::string msint;
};

MSI findMSI(const std::string& tseq1, const std::string& tseq2, const std::string& left) {
    int nmsi = 1, shift3 = 0;
    std::string maxmsi;
    double msicnt = 0;
    while (nmsi <= (int)tseq1.size() && nmsi <= 6) {
        std::string msint = vc_substr2(tseq1, -nmsi, nmsi);
        std::string mubiao = left + tseq1;
        double curmsi = 1.0;
        for (int i = (int)(mubiao.size() - (size_t)nmsi); i > 0; i -= nmsi) {
            if (i > nmsi && mubiao.substr((size_t)(i - nmsi), (size_t)nmsi) == msint)
                curmsi += 1;
            else
                break;
        }
        for (size_t i = 0; i < tseq2.size(); i += (size_t)nmsi) {
            if (i + (size_t)nmsi < tseq2.size() && tseq2.substr(i, (size_t)nmsi) == msint)
                curmsi += 1.0;
            else
                break;
        }
        if (curmsi > msicnt) {
            maxmsi = msint;
            msicnt = curmsi;
        }
        nmsi++;
    }
    std::string tseq = tseq1 + tseq2;
    while ((size_t)shift3 < tseq2.size() && tseq[(size_t)shift3] == tseq2[(size_t)shift3]) shift3++;
    return MSI{msicnt, shift3, maxmsi};
}

struct Scorer {
    const Config& c;
    const Region& R;
    RefMap& ref;
    int chrLen;
    Pileup& P;
    Aligned& A;

    Vars* getOrPut(int pos) {
        Vars** v = A.vars.find(pos);
        if (v) return *v;
        A.varsPool.emplace_back();
        Vars* nv = &A.varsPool.back();
        A.vars[pos] = nv;
        return nv;
    }
    Variant* newVariant() {
        A.vpool.emplace_back();
        return &A.vpool.back();
    }

    MSI vrefIsDeletion(int position, int dellen) {
        std::string leftseq = joinRef(ref, position - 70 > 1 ? position - 70 : 1, position - 1);
        std::string tseq = joinRef(ref, position, position + dellen + 70 > chrLen ? chrLen : position + dellen + 70);
        MSI m = findMSI(vc_substr2(tseq, 0, dellen), vc_substr1(tseq, dellen), leftseq);
        MSI w = findMSI(leftseq, vc_substr1(tseq, dellen), "");
        if (m.msi < w.msi) {
            m.msi = w.msi;
            m.msint = w.msint;
        }
        if (m.msi <= m.shift3 / (double)dellen) m.msi = m.shift3 / (double)dellen;
        return m;
    }
    MSI vrefIsInsertion(int position, const std::string& vn) {
        std::string tseq1 = vn.substr(1);
        std::string leftseq = joinRef(ref, position - 50 > 1 ? position - 50 : 1, position);
        std::string tseq2 = joinRef(ref, position + 1, position + 70 > chrLen ? chrLen : position + 70);
        MSI m = findMSI(tseq1, tseq2, leftseq);
        MSI w = findMSI(leftseq, tseq2, "");
        if (m.msi < w.msi) {
            m.msi = w.msi;
            m.msint = w.msint;
        }
        if (m.msi <= m.shift3 / (double)tseq1.size()) m.msi = m.shift3 / (double)tseq1.size();
        return m;
    }

    Variant* make_variant(const std::string& desc, Var* cnt, int ttcov, int hicov, int fwd, int rev) {
        Variant* t = newVariant();
        t->descriptionString = desc;
        t->positionCoverage = cnt->cnt;
        t->varsCountOnForward = fwd;
        t->varsCountOnReverse = rev;
        t->strandBiasFlag = std::to_string(strandBias(fwd, rev, c));
        t->frequency = cnt->cnt / (double)ttcov;
        t->meanPosition = cnt->pmean / (double)cnt->cnt;
        t->isAtLeastAt2Positions = cnt->pstd;
        t->meanQuality = cnt->qmean / cnt->cnt;
        t->hasAtLeast2DiffQualities = cnt->qstd;
        t->meanMappingQuality = cnt->mqmean / (double)cnt->cnt;
        t->highQualityToLowQualityRatio = cnt->hicnt / (cnt->locnt != 0 ? (double)cnt->locnt : 0.5);
        t->highQualityReadsFrequency = hicov > 0 ? cnt->hicnt / (double)hicov : 0;
        t->extraFrequency = cnt->extracnt != 0 ? cnt->extracnt / (double)ttcov : 0;
        t->shift3 = 0;
        t->msi = 0;
        t->numberOfMismatches = cnt->nm / (double)cnt->cnt;
        t->hicnt = cnt->hicnt;
        t->hicov = hicov;
        t->duprate = 0;
        return t;
    }

    void createVariant(int position, VarMap* vm, int tpc, std::vector<Variant*>& var, int hicov) {
        std::vector<std::string> keys;
        for (auto& kv : *vm) keys.push_back(kv.first);
        std::sort(keys.begin(), keys.end(), [](const std::string& a, const std::string& b) { return a.compare(b) < 0; });
        for (const std::string& d : keys) {
            Var* cnt = vm->at(d);
            if (cnt->cnt == 0) continue;
            int ttcov = tpc;
            if (cnt->cnt > tpc && cnt->extracnt > 0 && cnt->cnt - tpc < cnt->extracnt) ttcov = cnt->cnt;
            var.push_back(make_variant(d, cnt, ttcov, hicov, cnt->fwd, cnt->rev));
        }
        (void)position;
    }

    int createInsertion(int position, int tpc, std::vector<Variant*>& var, int hicov) {
        VarMap** im = P.ins.find(position);
        if (!im) return tpc;
        VarMap copy = **im;
        std::vector<std::string> keys;
        for (auto& kv : copy) keys.push_back(kv.first);
        std::sort(keys.begin(), keys.end(), [](const std::string& a, const std::string& b) { return a.compare(b) < 0; });
        for (const std::string& d : keys) {
            Var* cnt = copy[d];
            int fwd = cnt->fwd, rev = cnt->rev;
            int hicnt = cnt->hicnt;
            int ttcov = tpc;
            if (cnt->cnt > tpc && cnt->extracnt != 0 && cnt->cnt - tpc < cnt->extracnt) ttcov = cnt->cnt;
            if (ttcov < cnt->cnt) {
                ttcov = cnt->cnt;
                auto rc1 = P.refCov.find(position + 1);
                if (rc1 != P.refCov.end() && ref.count(position + 1) && ttcov < rc1->second - cnt->cnt) {
                    ttcov = rc1->second;
                    Var* nx = P.getVariationMaybe(position + 1, ref.at(position + 1));
                    if (nx) {
                        nx->fwd -= fwd;
                        nx->rev -= rev;
                    }
                }
                tpc = ttcov;
            }
            if (hicov < hicnt) hicov = hicnt;
            var.push_back(make_variant(d, cnt, ttcov, hicov, fwd, rev));
        }
        return tpc;
    }

    std::string validateRefallele(std::string r) {
        static const char* from = "MRWSYKVHDB";
        static const char* to = "AAACCGAAAC";
        for (size_t i = 0; i < r.size(); i++) {
            std::string b = vc_substr2(r, (int)i, 1);
            const char* k = b.size() == 1 ? strchr(from, b[0]) : nullptr;
            if (k && b[0] != 0) replaceFirst(r, b, std::string(1, to[k - from]));
        }
        return r;
    }

    void collectReferenceVariants(int position, int tpc, Vars* vp) {
        static const std::regex BEGIN_MINUS_NUMBER("^-(\\d+).*");
        static const std::regex AMP_ATGC(".*&([ATGC]+).*");
        static const std::regex HASH_GROUP_CARET_GROUP("#(.+)\\^(.+)");
        static const std::regex BEGIN_DIGITS("^(\\d+)");
        static const std::regex CARET_ATGNC("\\^([ATGNC]+)");
        int rfc = 0, rrc = 0;
        std::string genotype1;
        if (vp->referenceVariant && vp->referenceVariant->frequency >= c.freq)
            genotype1 = vp->referenceVariant->descriptionString;
        else if (!vp->variants.empty())
            genotype1 = vp->variants[0]->descriptionString;
        else
            genotype1 = vp->referenceVariant->descriptionString;
        if (vp->referenceVariant) {
            rfc = vp->referenceVariant->varsCountOnForward;
            rrc = vp->referenceVariant->varsCountOnReverse;
        }
        if (starts_with(genotype1, "+")) genotype1 = "+" + std::to_string(genotype1.size() - 1);
        std::string genotype2;
        if (tpc > P.refCov[position] && P.nonIns.count(position + 1) && ref.count(position + 1)) {
            VarMap* nm = *P.nonIns.find(position + 1);
            auto it = nm->find(std::string(1, ref.at(position + 1)));
            if (it != nm->end()) {
                rfc = it->second->fwd;
                rrc = it->second->rev;
            }
        }
        if (!vp->variants.empty()) {
            for (Variant* vref : vp->variants) {
                genotype2 = vref->descriptionString;
                if (starts_with(genotype2, "+")) genotype2 = "+" + std::to_string(genotype2.size() - 1);
                std::string d = vref->descriptionString;
                int dellen = 0;
                std::smatch sm;
                if (std::regex_search(d, sm, BEGIN_MINUS_NUMBER)) dellen = std::atoi(sm.str(1).c_str());
                int endPosition = position;
                if (starts_with(d, "-")) endPosition = position + dellen - 1;
                std::string refallele, varallele;
                int shift3 = 0;
                double msi = 0;
                std::string msint;
                int startPosition = position;
                if (starts_with(d, "+")) {
                    if (d.find('&') == std::string::npos && d.find('#') == std::string::npos &&
                        d.find("<dup") == std::string::npos) {
                        MSI m = vrefIsInsertion(position, d);
                        msi = m.msi;
                        shift3 = m.shift3;
                        msint = m.msint;
                    }
                    if (c.moveIndelsTo3) {
                        startPosition += shift3;
                        endPosition += shift3;
                    }
                    refallele = ref.count(position) ? std::string(1, ref.at(position)) : "";
                    varallele = refallele + d.substr(1);
                } else if (starts_with(d, "-")) {
                    if (dellen < c.SVMINLEN) {
                        varallele = d;
                        replaceFirstRe(varallele, "^-\\d+", "");
                        MSI m = vrefIsDeletion(position, dellen);
                        msi = m.msi;
                        shift3 = m.shift3;
                        msint = m.msint;
                    }
                    if (d.find('&') == std::string::npos && d.find('#') == std::string::npos &&
                        d.find('^') == std::string::npos) {
                        if (c.moveIndelsTo3) startPosition += shift3;
                        if (varallele != "<DEL>") varallele = ref.count(position - 1) ? std::string(1, ref.at(position - 1)) : "";
                        refallele = ref.count(position - 1) ? std::string(1, ref.at(position - 1)) : "";
                        startPosition--;
                    }
                    if (dellen < c.SVMINLEN) refallele += joinRef(ref, position, position + dellen - 1);
                } else {
                    std::string tseq1 = joinRef(ref, position - 30 > 1 ? position - 30 : 1, position + 1);
                    std::string tseq2 = joinRef(ref, position + 2, position + 70 > chrLen ? chrLen : position + 70);
                    MSI m = findMSI(tseq1, tseq2, "");
                    msi = m.msi;
                    shift3 = m.shift3;
                    msint = m.msint;
                    refallele = ref.count(position) ? std::string(1, ref.at(position)) : "";
                    varallele = d;
                }
                if (std::regex_search(d, sm, AMP_ATGC)) {
                    std::string extra = sm.str(1);
                    replaceFirst(varallele, "&", "");
                    std::string tch = joinRef(ref, endPosition + 1, endPosition + (int)extra.size());
                    refallele += tch;
                    genotype1 += tch;
                    endPosition += (int)extra.size();
                    std::smatch s2;
                    if (std::regex_search(varallele, s2, AMP_ATGC)) {
                        std::string vextra = s2.str(1);
                        replaceFirst(varallele, "&", "");
                        tch = joinRef(ref, endPosition + 1, endPosition + (int)vextra.size());
                        refallele += tch;
                        genotype1 += tch;
                        endPosition += (int)vextra.size();
                    }
                    if (starts_with(d, "+")) {
                        refallele = refallele.substr(1);
                        varallele = varallele.substr(1);
                        startPosition++;
                    }
                }
                if (std::regex_search(d, sm, HASH_GROUP_CARET_GROUP)) {
                    std::string mseq = sm.str(1);
                    std::string tail = sm.str(2);
                    endPosition += (int)mseq.size();
                    refallele += joinRef(ref, endPosition - (int)mseq.size() + 1, endPosition);
                    std::smatch smq;
                    if (std::regex_search(tail, smq, BEGIN_DIGITS)) {
                        int del = std::atoi(smq.str().c_str());
                        refallele += joinRef(ref, endPosition + 1, endPosition + del);
                        endPosition += del;
                    }
                    replaceFirst(varallele, "#", "");
                    replaceFirstRe(varallele, "\\^(\\d+)?", "");
                    replaceFirst(genotype1, "#", "m");
                    replaceFirst(genotype1, "^", "i");
                    replaceFirst(genotype2, "#", "m");
                    replaceFirst(genotype2, "^", "i");
                }
                if (std::regex_search(d, CARET_ATGNC)) {
                    replaceFirst(varallele, "^", "");
                    replaceFirst(genotype1, "^", "i");
                    replaceFirst(genotype2, "^", "i");
                }
                if (c.crisprCuttingSite != 0) throw std::runtime_error("CRISPR mode is not restated");
                vref->leftseq = joinRef(ref, startPosition - 20 < 1 ? 1 : startPosition - 20, startPosition - 1);
                vref->rightseq = joinRef(ref, endPosition + 1, endPosition + 20 > chrLen ? chrLen : endPosition + 20);
                std::string genotype = genotype1 + "/" + genotype2;
                genotype.erase(std::remove(genotype.begin(), genotype.end(), '&'), genotype.end());
                genotype.erase(std::remove(genotype.begin(), genotype.end(), '#'), genotype.end());
                {
                    std::string g;
                    for (char ch : genotype) {
                        if (ch == '^')
                            g += 'i';
                        else
                            g += ch;
                    }
                    genotype = g;
                }
                vref->msi = msi;
                vref->msint = (int)msint.size();
                vref->shift3 = shift3;
                vref->startPosition = startPosition;
                vref->endPosition = endPosition;
                vref->refallele = validateRefallele(refallele);
                vref->varallele = varallele;
                vref->genotype = genotype;
                vref->totalPosCoverage = tpc;
                vref->refForwardCoverage = rfc;
                vref->refReverseCoverage = rrc;
                if (vp->referenceVariant)
                    vref->strandBiasFlag = vp->referenceVariant->strandBiasFlag + ";" + vref->strandBiasFlag;
                else
                    vref->strandBiasFlag = "0;" + vref->strandBiasFlag;
                if (vref->refForwardCoverage < 0) vref->refForwardCoverage = 0;
                if (vref->refReverseCoverage < 0) vref->refReverseCoverage = 0;
                if (vref->varsCountOnForward < 0) vref->varsCountOnForward = 0;
                if (vref->varsCountOnReverse < 0) vref->varsCountOnReverse = 0;
            }
        } else if (vp->referenceVariant) {
            Variant* vref = vp->referenceVariant;
            vref->totalPosCoverage = tpc;
            vref->positionCoverage = 0;
            vref->frequency = 0;
            vref->refForwardCoverage = rfc;
            vref->refReverseCoverage = rrc;
            vref->varsCountOnForward = 0;
            vref->varsCountOnReverse = 0;
            vref->msi = 0;
            vref->msint = 0;
            vref->strandBiasFlag += ";0";
            vref->shift3 = 0;
            vref->startPosition = position;
            vref->endPosition = position;
            std::string rb = ref.count(position) ? std::string(1, ref.at(position)) : "";
            vref->refallele = validateRefallele(rb);
            vref->varallele = validateRefallele(rb);
            vref->genotype = rb + "/" + rb;
            vref->leftseq = "";
            vref->rightseq = "";
            vref->duprate = 0;
        } else {
            vp->referenceVariant = newVariant();
        }
    }

    void process() {
        P.nonIns.for_each([&](int position, VarMap* vm) {
            try {
                if (vm->empty() && !P.ins.count(position)) return;
                auto rc = P.refCov.find(position);
                if (rc == P.refCov.end() || rc->second == 0) return;
                if (vm->size() == 1 && !P.ins.count(position) && ref.count(position) &&
                    vm->count(std::string(1, ref.at(position)))) {
                    if (!c.doPileup && !c.hasBam2) return;
                }
                int tpc = P.refCov[position];
                int hicov = 0;
                for (auto& kv : *vm) hicov += kv.second->hicnt;
                if (VarMap** im = P.ins.find(position))
                    for (auto& kv : **im) hicov += kv.second->hicnt;
                std::vector<Variant*> var;
                createVariant(position, vm, tpc, var, hicov);
                tpc = createInsertion(position, tpc, var, hicov);
                std::sort(var.begin(), var.end(), [](Variant* a, Variant* b) {
                    double res = a->meanQuality * a->positionCoverage - b->meanQuality * b->positionCoverage;
                    if (res < 0.00001 && res > -0.00001) return a->descriptionString.compare(b->descriptionString) < 0;
                    return res > 0;
                });
                double maxfreq = 0;
                for (Variant* tv : var) {
                    if (ref.count(position) && tv->descriptionString == std::string(1, ref.at(position))) {
                        getOrPut(position)->referenceVariant = tv;
                    } else {
                        Vars* vv = getOrPut(position);
                        vv->variants.push_back(tv);
                        vv->byDesc[tv->descriptionString] = tv;
                        if (tv->frequency > maxfreq) maxfreq = tv->frequency;
                    }
                }
                if (!c.doPileup && maxfreq <= c.freq) {
                    if (!c.hasBam2) {
                        A.vars.erase(position);
                        return;
                    }
                }
                collectReferenceVariants(position, tpc, getOrPut(position));
            } catch (const std::exception&) {
            }
        });
        A.maxReadLength = P.maxReadLength;
    }
};

}  // namespace

// findMSI's result as "msi shift3 msint", for known-answer tests (rvcpu --findmsi)
std::string findmsi_kat(const std::string& t1, const std::string& t2, const std::string& left) {
    MSI m = findMSI(t1, t2, left);
    char b[64];
    snprintf(b, sizeof b, "%.17g %d ", m.msi, m.shift3);
    return std::string(b) + m.msint;
}

void score_region(const Config& c, const Region& r, RefMap& ref, int chrLen, Pileup& P, Aligned& A) {
    Scorer s{c, r, ref, chrLen, P, A};
    s.process();
}

// ------------------------------------------------------------------ output (simpleMode.cpp)
namespace {

std::string varType(const Variant* v) {
    if (v->refallele.size() == 1 && v->varallele.size() == 1) return "SNV";
    if (v->varallele.size() == 5 && v->varallele[0] == '<' && v->varallele[4] == '>') return v->varallele.substr(1, 3);
    if (v->refallele.empty() || v->varallele.empty()) return "Complex";
    if (v->refallele[0] != v->varallele[0]) return "Complex";
    if (v->refallele.size() == 1 && v->varallele.size() > 1 && v->varallele.find(v->refallele) == 0) return "Insertion";
    if (v->refallele.size() > 1 && v->varallele.size() == 1 && v->refallele.find(v->varallele) == 0) return "Deletion";
    return "Complex";
}

bool isGoodVar(const Variant* v, const Variant* rv, const std::string& type, const Config& c) {
    if (v->refallele.empty()) return false;
    if (v->frequency < c.freq || v->hicnt < c.minr || v->meanPosition < c.readPosFilter || v->meanQuality < c.goodq)
        return false;
    if (rv && rv->hicnt > c.minr && v->frequency < 0.25) {
        double d = v->meanMappingQuality + v->refallele.size() + v->varallele.size();
        double f = (1 + d) / (rv->meanMappingQuality + 1);
        if ((d - 2 < 5 && rv->meanMappingQuality > 20) || f < 0.25) return false;
    }
    // the splice-set lookup is only reachable with N CIGAR operations (RNA-seq); DNA inputs keep the
    // set empty, so the Deletion/splice test never rejects here
    if (v->highQualityToLowQualityRatio < c.qratio) return false;
    if (v->frequency > 0.30) return true;
    if (v->meanMappingQuality < c.mapq) return false;
    if (v->strandBiasFlag == "2;1" && v->frequency < 0.20) {
        if (type.empty() || type == "SNV" || (v->refallele.size() < 3 && v->varallele.size() < 3)) return false;
    }
    return true;
}

void adjComplex(Variant* v) {
    std::string ra = v->refallele, va = v->varallele;
    if (va[0] == '<') return;
    size_t n = 0;
    while (ra.size() - n > 1 && va.size() - n > 1 && ra[n] == va[n]) n++;
    if (n > 0) {
        v->startPosition += (int)n;
        v->refallele = vc_substr1(ra, (int)n);
        v->varallele = vc_substr1(va, (int)n);
        v->leftseq += vc_substr2(ra, 0, (int)n);
        v->leftseq = vc_substr1(v->leftseq, (int)n);
    }
    ra = v->refallele;
    va = v->varallele;
    int k = 1;
    while ((int)ra.size() - k > 0 && (int)va.size() - k > 0 && vc_substr2(ra, -k, 1) == vc_substr2(va, -k, 1)) k++;
    if (k > 1) {
        v->endPosition -= k - 1;
        v->refallele = vc_substr2(ra, 0, 1 - k);
        v->varallele = vc_substr2(va, 0, 1 - k);
        v->rightseq = vc_substr2(ra, 1 - k, k - 1) + vc_substr2(v->rightseq, 0, 1 - k);
    }
}

void print_line(std::string& out, const Variant* v, const Region& R, const std::string& sv, int position,
                const Config& c) {
    std::vector<std::string> s;
    s.push_back(c.sample);
    s.push_back(R.gene);
    s.push_back(R.chr);
    if (!v) {
        s.push_back(std::to_string(position));
        s.push_back(std::to_string(position));
    } else {
        s.push_back(std::to_string(v->startPosition));
        s.push_back(std::to_string(v->endPosition));
        s.push_back(v->refallele);
        s.push_back(v->varallele);
        s.push_back(std::to_string(v->totalPosCoverage));
        s.push_back(std::to_string(v->positionCoverage));
        s.push_back(std::to_string(v->refForwardCoverage));
        s.push_back(std::to_string(v->refReverseCoverage));
        s.push_back(std::to_string(v->varsCountOnForward));
        s.push_back(std::to_string(v->varsCountOnReverse));
        s.push_back(v->genotype.empty() ? "0" : v->genotype);
        s.push_back(to_string_f(v->frequency));
        s.push_back(v->strandBiasFlag);
        s.push_back(to_string_f(v->meanPosition));
        s.push_back(v->isAtLeastAt2Positions ? "1" : "0");
        s.push_back(to_string_f(v->meanQuality));
        s.push_back(v->hasAtLeast2DiffQualities ? "1" : "0");
        if (c.fisher) {
            double p = kt_fisher_two_sided(v->refForwardCoverage, v->refReverseCoverage, v->varsCountOnForward,
                                           v->varsCountOnReverse, nullptr, nullptr);
            s.push_back(to_string_f(p));
            double ad = (double)(v->refForwardCoverage * v->varsCountOnReverse);
            double bc = (double)(v->refReverseCoverage * v->varsCountOnForward);
            double ratio = 0.0;
            if (bc != 0 && ad != 0) ratio = ad > bc ? (ad / bc) : (bc / ad);
            s.push_back(to_string_f(ratio));
        }
        s.push_back(to_string_f(v->meanMappingQuality));
        s.push_back(to_string_f(v->highQualityToLowQualityRatio));
        s.push_back(to_string_f(v->highQualityReadsFrequency));
        s.push_back(to_string_f(v->extraFrequency));
        s.push_back(std::to_string(v->shift3));
        s.push_back(to_string_f(v->msi));
        s.push_back(std::to_string(v->msint));
        s.push_back(v->numberOfMismatches > 0 ? to_string_f(v->numberOfMismatches) : "0");
        s.push_back(std::to_string(v->hicnt));
        s.push_back(std::to_string(v->hicov));
        s.push_back(v->leftseq.empty() ? "0" : v->leftseq);
        s.push_back(v->rightseq.empty() ? "0" : v->rightseq);
        s.push_back(R.chr + ":" + std::to_string(R.start) + "-" + std::to_string(R.end));
        s.push_back(v->vartype);
        s.push_back(to_string_f(v->duprate));
        s.push_back(sv.empty() ? "0" : sv);
    }
    for (size_t i = 0; i + 1 < s.size(); i++) {
        out += s[i];
        out += '\t';
    }
    out += s.back();
    out += '\n';
}

}  // namespace

std::string format_simple(const Config& c, const Region& R, Aligned& A) {
    std::string out;
    A.vars.for_each([&](int position, Vars* vv) {
        if (vv->sv.empty() && (position < R.start || position > R.end)) return;
        std::vector<Variant*> vrefs;
        if (vv->variants.empty()) {
            if (!c.doPileup) return;
            Variant* vref = vv->referenceVariant;
            if (!vref) {
                print_line(out, nullptr, R, vv->sv, position, c);
                return;
            }
            vref->vartype = "";
            vrefs.push_back(vref);
        } else {
            for (Variant* vref : vv->variants) {
                if (vref->refallele.find('N') != std::string::npos) continue;
                vref->vartype = varType(vref);
                if (!isGoodVar(vref, vv->referenceVariant, vref->vartype, c)) {
                    if (!c.doPileup) continue;
                }
                vrefs.push_back(vref);
            }
        }
        for (Variant* vref : vrefs) {
            if (vref->vartype == "Complex") adjComplex(vref);
            if (c.crisprCuttingSite == 0) vref->crispr = 0;
            print_line(out, vref, R, vv->sv, position, c);
        }
    });
    return out;
}

}  // namespace orc

// ------------------------------------------------------------------ somatic output (somaticMode.cpp)
namespace orc {
namespace {

const std::string kStrongSomatic = "StrongSomatic", kSampleSpecific = "SampleSpecific", kLikelyLOH = "LikelyLOH",
                  kGermline = "Germline", kStrongLOH = "StrongLOH", kLikelySomatic = "LikelySomatic",
                  kAFDiff = "AFDiff", kComplex = "Complex", kSNV = "SNV";

// Variant::isNoise (include/Variant.h:75-94): clears the variant's counts when it is noise
bool isNoise(Variant* v, const Config& c) {
    const double qual = v->meanQuality;
    if (((qual < 4.5 || (qual < 12 && !v->hasAtLeast2DiffQualities)) && v->positionCoverage <= 3) ||
        (qual < c.goodq && v->frequency < 2 * c.lofreq && v->positionCoverage <= 1)) {
        v->totalPosCoverage -= v->positionCoverage;
        v->positionCoverage = 0;
        v->varsCountOnForward = 0;
        v->varsCountOnReverse = 0;
        v->frequency = 0;
        v->highQualityReadsFrequency = 0;
        return true;
    }
    return false;
}

// MINUS_NUM_NUM (include/patterns.h:18): a '-' followed by two digits
bool minus_num_num(const std::string& s) {
    for (size_t i = 0; i + 2 < s.size(); i++)
        if (s[i] == '-' && isdigit((unsigned char)s[i + 1]) && isdigit((unsigned char)s[i + 2])) return true;
    return false;
}

// put_fisher_ext_and_odds (somaticMode.cpp:130-149): two-sided p, then log(ad/bc + bc/ad) in float
// over the counts plus one half
void put_fisher_odds(std::string& s, int rf, int rr, int af, int ar) {
    s += to_string_f(kt_fisher_two_sided(rf, rr, af, ar, nullptr, nullptr));
    s += '\t';
    const float a = (float)(rf + 0.5), b = (float)(rr + 0.5), cc = (float)(af + 0.5), d = (float)(ar + 0.5);
    const float ad = a * d, bc = b * cc;
    const float ratio = std::log(ad / bc + bc / ad);
    s += to_string_f((double)ratio);
    s += '\t';
}

void sample_fields(std::string& s, const Variant* v, const Variant* genotypeOf) {
    s += std::to_string(v->totalPosCoverage) + "\t" + std::to_string(v->positionCoverage) + "\t" +
         std::to_string(v->refForwardCoverage) + "\t" + std::to_string(v->refReverseCoverage) + "\t" +
         std::to_string(v->varsCountOnForward) + "\t" + std::to_string(v->varsCountOnReverse) + "\t";
    s += v->genotype.empty() ? std::string("0") : genotypeOf->genotype;
    s += "\t" + to_string_f(v->frequency) + "\t" + v->strandBiasFlag + "\t" + to_string_f(v->meanPosition) + "\t" +
         std::to_string(v->isAtLeastAt2Positions ? 1 : 0) + "\t" + to_string_f(v->meanQuality) + "\t" +
         std::to_string(v->hasAtLeast2DiffQualities ? 1 : 0) + "\t" + to_string_f(v->meanMappingQuality) + "\t" +
         to_string_f(v->highQualityToLowQualityRatio) + "\t" + to_string_f(v->highQualityReadsFrequency) + "\t" +
         to_string_f(v->extraFrequency) + "\t" + to_string_f(v->numberOfMismatches) + "\t";
}

// SomaticMode::print_output_variant_simple (somaticMode.cpp:151-309)
std::string somatic_line(const Variant* begin, const Variant* end, const Variant* tumor, const Variant* normal,
                         const Region& R, const std::string& label, const Config& c) {
    static const std::string null18 = "0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t";
    std::string s = c.sample + "\t" + R.gene + "\t" + R.chr + "\t";
    if (begin)
        s += std::to_string(begin->startPosition) + "\t" + std::to_string(begin->endPosition) + "\t" +
             begin->refallele + "\t" + begin->varallele + "\t";
    else
        s += "0\t0\t0\t0\t";
    if (tumor)
        sample_fields(s, tumor, tumor);
    else
        s += null18;
    if (c.fisher) {
        if (tumor)
            put_fisher_odds(s, tumor->refForwardCoverage, tumor->refReverseCoverage, tumor->varsCountOnForward,
                            tumor->varsCountOnReverse);
        else
            s += "0\t0\t";
    }
    if (normal) {
        // the reference prints the tumor variant's genotype here whenever the normal one has one
        if (!normal->genotype.empty() && !tumor) throw std::runtime_error("somatic printer: normal genotype without tumor");
        sample_fields(s, normal, tumor);
    } else {
        s += null18;
    }
    if (c.fisher) {
        if (normal)
            put_fisher_odds(s, normal->refForwardCoverage, normal->refReverseCoverage, normal->varsCountOnForward,
                            normal->varsCountOnReverse);
        else
            s += "0\t0\t";
    }
    if (end)
        s += std::to_string(end->shift3) + "\t" + to_string_f(end->msi) + "\t" + std::to_string(end->msint) + "\t" +
             (end->leftseq.empty() ? std::string("0") : end->leftseq) + "\t" +
             (end->rightseq.empty() ? std::string("0") : end->rightseq) + "\t";
    else
        s += "\t\t\t\t\t";
    s += R.chr + ":" + std::to_string(R.start) + "-" + std::to_string(R.end) + "\t";
    s += label + "\t";
    s += begin ? begin->vartype + "\t" : std::string("\t");
    s += tumor ? to_string_f(tumor->duprate) + "\t" : std::string("0\t");
    s += "0\t";
    s += normal ? to_string_f(normal->duprate) + "\t" : std::string("0\t");
    s += "0\t";
    if (c.fisher) {
        const int t_tot = tumor ? tumor->totalPosCoverage : 0, t_var = tumor ? tumor->positionCoverage : 0;
        const int n_tot = normal ? normal->totalPosCoverage : 0, n_var = normal ? normal->positionCoverage : 0;
        int tref = t_tot - t_var, rref = n_tot - n_var;
        if (tref < 0) tref = 0;
        if (rref < 0) rref = 0;
        put_fisher_odds(s, t_var, tref, n_var, rref);
        const double tv = tumor ? tumor->frequency : 0, nv = normal ? normal->frequency : 0;
        s += to_string_f(std::log(std::max(tv, 0.0001) / std::max(nv, 0.0001))) + "\t";
        s += to_string_f(std::log(((float)t_var + 0.5) / ((float)n_var + 0.5))) + "\t";
    }
    s += "\n";
    return s;
}

struct Somatic {
    const Config& c;
    const Region& R;

    // determinateType (somaticMode.cpp:598-621)
    std::string determinateType(Vars* variants, Variant* standard, Variant* cmp) {
        std::string type;
        if (isGoodVar(cmp, variants->referenceVariant, standard->vartype, c)) {
            if (standard->frequency > (1 - c.lofreq) && cmp->frequency < 0.8 && cmp->frequency > 0.2)
                type = kLikelyLOH;
            else if (cmp->frequency < c.lofreq || cmp->positionCoverage <= 1)
                type = kLikelySomatic;
            else
                type = kGermline;
        } else {
            if (cmp->frequency < c.lofreq || cmp->positionCoverage <= 1)
                type = kLikelySomatic;
            else
                type = kAFDiff;
        }
        if (isNoise(cmp, c) && standard->vartype == kSNV) type = kStrongSomatic;
        return type;
    }

    // callingForOneSample (:358-380)
    std::string one_sample(Vars* vs, bool firstCover, const std::string& label) {
        std::string out;
        for (Variant* v : vs->variants) {
            v->vartype = varType(v);
            if (!isGoodVar(v, vs->referenceVariant, v->vartype, c)) continue;
            if (v->vartype == kComplex) adjComplex(v);
            out += firstCover ? somatic_line(v, v, nullptr, v, R, label, c) : somatic_line(v, v, v, nullptr, R, label, c);
        }
        return out;
    }

    // printVariationsFromFirstSample (:400-518)
    std::string first_sample(Vars* v1, Vars* v2) {
        std::string out;
        size_t k = 0;
        while (k < v1->variants.size()) {
            Variant* vref = v1->variants[k];
            const std::string type0 = vref->vartype.empty() ? varType(vref) : vref->vartype;
            if (!isGoodVar(vref, v1->referenceVariant, type0, c)) break;
            const std::string nt = vref->descriptionString;
            vref->vartype = varType(vref);
            if (vref->vartype == kComplex) adjComplex(vref);
            auto it = v2->byDesc.find(nt);
            if (it != v2->byDesc.end()) {
                Variant* v2nt = it->second;
                const std::string type = determinateType(v2, vref, v2nt);
                out += somatic_line(vref, v2nt, vref, v2nt, R, type, c);
            } else {
                // tumor only: strong somatic
                Variant hold;
                const Variant* forPrint;
                if (!v2->variants.empty()) {
                    const Variant* v2r = v2->variants[0];
                    hold.totalPosCoverage = v2r->totalPosCoverage;
                    hold.refForwardCoverage = v2r->refForwardCoverage;
                    hold.refReverseCoverage = v2r->refReverseCoverage;
                    forPrint = &hold;
                } else {
                    forPrint = v2->referenceVariant;
                }
                if (vref->vartype != kSNV && (nt.size() > 10 || minus_num_num(nt)) &&
                    vref->positionCoverage < c.minr + 3 && nt.find('<') != std::string::npos)
                    throw std::runtime_error("somatic: combineAnalysis (structural descriptors) not restated");
                out += somatic_line(vref, vref, vref, forPrint, R, kStrongSomatic, c);
            }
            k++;
        }
        if (k == 0) {
            if (v2->variants.empty()) return "";
            for (Variant* v2var : v2->variants) {
                v2var->vartype = varType(v2var);
                if (!isGoodVar(v2var, v2->referenceVariant, v2var->vartype, c)) continue;
                const std::string nt = v2var->descriptionString;
                auto it = v1->byDesc.find(nt);
                if (it != v1->byDesc.end()) {
                    Variant* v1nt = it->second;
                    const std::string type = v1nt->frequency < c.lofreq ? kLikelyLOH : kGermline;
                    if (v2var->vartype == kComplex) adjComplex(v1nt);
                    v1nt->vartype = varType(v1nt);
                    out += somatic_line(v1nt, v2var, v1nt, v2var, R, type, c);
                } else {
                    const Variant* v1var = v1->variants.empty() ? nullptr : v1->variants[0];
                    const int tcov = v1var && v1var->totalPosCoverage != 0 ? v1var->totalPosCoverage : 0;
                    const Variant* v1ref = v1->referenceVariant;
                    Variant hold;
                    hold.totalPosCoverage = tcov;
                    hold.refForwardCoverage = v1ref ? v1ref->varsCountOnForward : 0;
                    hold.refReverseCoverage = v1ref ? v1ref->varsCountOnReverse : 0;
                    hold.genotype = v1var ? v1var->genotype
                                          : (v1ref ? v1ref->descriptionString + "/" + v1ref->descriptionString : "N/N");
                    if (v2var->vartype == kComplex) adjComplex(v2var);
                    out += somatic_line(v2var, v2var, &hold, v2var, R, kStrongLOH, c);
                }
            }
        }
        return out;
    }

    // printVariationsFromSecondSample (:525-596)
    std::string second_sample(Vars* v1, Vars* v2) {
        std::string out;
        for (Variant* v2var : v2->variants) {
            v2var->vartype = varType(v2var);
            if (!isGoodVar(v2var, v2->referenceVariant, v2var->vartype, c)) continue;
            const std::string& nt = v2var->descriptionString;
            auto it = v1->byDesc.find(nt);
            Variant fresh;
            Variant* v1nt = it != v1->byDesc.end() ? it->second : &fresh;
            v1nt->positionCoverage = 0;
            if (v2->byDesc.at(nt)->positionCoverage < c.minr + 3 && nt.find('<') != std::string::npos &&
                (nt.size() > 10 || minus_num_num(nt)))
                throw std::runtime_error("somatic: combineAnalysis (structural descriptors) not restated");
            const Variant* forPrint = v1->referenceVariant;
            if (v2var->vartype == kComplex) adjComplex(v2var);
            out += somatic_line(v2var, v2var, forPrint, v2var, R, kStrongLOH, c);
        }
        return out;
    }
};

}  // namespace

std::string format_somatic(const Config& c, const Region& R, Aligned& A1, Aligned& A2) {
    // positions: the first sample's alignedVariants keys, in order (the reference inserts the first
    // map's keys twice, somaticMode.cpp:320-325, so the second map's own positions are not visited)
    std::set<int> positions;
    A1.vars.for_each([&](int p, Vars*) { positions.insert(p); });
    Somatic S{c, R};
    std::string out;
    for (int p : positions) {
        if (p < R.start || p > R.end) continue;
        Vars** v1 = A1.vars.find(p);
        Vars** v2 = A2.vars.find(p);
        if (!v1)
            out += S.one_sample(*v2, true, "Deletion");
        else if (!v2)
            out += S.one_sample(*v1, false, kSampleSpecific);
        else if (!(*v1)->variants.empty())
            out += S.first_sample(*v1, *v2);
        else if (!(*v2)->variants.empty())
            out += S.second_sample(*v1, *v2);
    }
    return out;
}

}  // namespace orc
