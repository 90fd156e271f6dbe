// Somatic mode: two samples of one region, classified and printed.
//
// Restates src/modes/somaticMode.cpp: one_region_run_somt (:83-128: the first sample, then the
// second over the same reference map, starting from the first sample's maximum read length), the
// output walk (:311-356), callingForOneSample (:358-380), callingForBothSamples (:382-393),
// printVariationsFromFirstSample (:400-518), printVariationsFromSecondSample (:525-596),
// determinateType (:598-621) and print_output_variant_simple with its three Fisher tests
// (:130-309); with include/Variant.h (isGoodVar :197-248, varType :167-188, adjComplex :99-132,
// isNoise :75-94).
//
// With a second sample ToVarsBuilder keeps every covered position, also reference-only ones and
// ones with nothing above -f (ToVarsBuilder.cpp:107-155, :213-233).  Each sample is piled up,
// realigned and scored once (every variant of every position that has one), then, once both
// samples' position states are known, scored again for the reference-only positions the printer
// will read: those where the other sample has variants.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "rv_fisher.h"
#include "rv_format.h"
#include "rv_realign.h"
#include "rv_region.h"

namespace rv {

namespace {

// ------------------------------------------------------------------ variant records on the host
struct SV {  // the Variant fields the somatic printer reads (include/Variant.h:34-66)
    std::string desc, refallele, varallele, genotype, bias = "0", leftseq, rightseq, vartype;
    int start = 0, end = 0, tpc = 0, poscov = 0, rfc = 0, rrc = 0, fwd = 0, rev = 0, shift3 = 0, msint = 0;
    int hicnt = 0, hicov = 0;
    bool pstd = false, qstd = false;
    double freq = 0, pmean = 0, qmean = 0, mq = 0, hlr = 0, hifreq = 0, extrafreq = 0, msi = 0, nm = 0,
           duprate = 0;
};

struct SVars {  // Vars (include/Vars.h): reference variant, variants in CMP_VARI order, by descriptor
    bool has_ref = false;
    SV ref;
    std::vector<SV> vars;
    // the variant stored under descriptor d (the first one added with it)
    SV* find(const std::string& d) {
        for (SV& v : vars)
            if (v.desc == d) return &v;
        return nullptr;
    }
};

SV to_sv(const OutRec& o, const char* arena) {
    SV v;
    const char* p = arena + o.str_off;
    std::string* dst[7] = {&v.refallele, &v.varallele, &v.genotype, &v.bias, &v.leftseq, &v.rightseq, &v.desc};
    for (int k = 0; k < 7; k++) {
        dst[k]->assign(p, o.len[k]);
        p += o.len[k];
    }
    v.start = o.start;
    v.end = o.end;
    v.tpc = o.tpc;
    v.poscov = o.poscov;
    v.rfc = o.rfc;
    v.rrc = o.rrc;
    v.fwd = o.fwd;
    v.rev = o.rev;
    v.shift3 = o.shift3;
    v.msint = o.msint;
    v.hicnt = o.hicnt;
    v.hicov = o.hicov;
    v.pstd = o.pstd != 0;
    v.qstd = o.qstd != 0;
    v.freq = o.freq;
    v.pmean = o.pmean;
    v.qmean = o.qmean;
    v.mq = o.mq;
    v.hlr = o.hlr;
    v.hifreq = o.hifreq;
    v.extrafreq = o.extrafreq;
    v.msi = o.msi;
    v.nm = o.nm;
    return v;
}

// ------------------------------------------------------------------ Variant helpers
std::string var_type(const SV& v) {
    if (v.refallele.size() == 1 && v.varallele.size() == 1) return "SNV";
    if (v.varallele.size() == 5 && v.varallele[0] == '<' && v.varallele[4] == '>') return v.varallele.substr(1, 3);
    if (v.refallele.empty() || v.varallele.empty()) return "Complex";
    if (v.refallele[0] != v.varallele[0]) return "Complex";
    if (v.refallele.size() == 1 && v.varallele.size() > 1 && v.varallele.compare(0, 1, v.refallele) == 0) return "Insertion";
    if (v.refallele.size() > 1 && v.varallele.size() == 1 && v.refallele.compare(0, 1, v.varallele) == 0) return "Deletion";
    return "Complex";
}

bool good_var(const SV& v, const SV* ref, std::string type, const Config& c) {
    if (v.refallele.empty()) return false;
    if (type.empty()) type = var_type(v);
    if (v.freq < c.freq || v.hicnt < c.minr || v.pmean < c.readPosFilter || v.qmean < c.goodq) return false;
    if (ref && ref->hicnt > c.minr && v.freq < 0.25) {
        const double d = v.mq + v.refallele.size() + v.varallele.size();
        const double f = (1 + d) / (ref->mq + 1);
        if ((d - 2 < 5 && ref->mq > 20) || f < 0.25) return false;
    }
    // the splice-set test needs N CIGAR operations; DNA regions leave the set empty
    if (v.hlr < c.qratio) return false;
    if (v.freq > 0.30) return true;
    if (v.mq < c.mapq) return false;
    if (v.bias == "2;1" && v.freq < 0.20) {
        if (type.empty() || type == "SNV" || (v.refallele.size() < 3 && v.varallele.size() < 3)) return false;
    }
    return true;
}

// vc_substr (include/util.h:27-57): negative starts count from the end, out-of-range yields ""
std::string vsub(const std::string& s, int b) {
    const int n = (int)s.size();
    if (b < 0) b += n;
    if (b < 0 || b > n) return "";
    return s.substr((size_t)b);
}
std::string vsub(const std::string& s, int b, int len) {
    const int n = (int)s.size();
    if (b < 0) b += n;
    if (b < 0 || b > n) return "";
    if (len < 0) len = n - b + len;
    if (len <= 0) return "";
    return s.substr((size_t)b, (size_t)std::min(len, n - b));
}

void adj_complex(SV& v) {
    std::string ra = v.refallele, va = v.varallele;
    if (va.empty() || va[0] == '<') return;
    size_t n = 0;
    while (ra.size() - n > 1 && va.size() - n > 1 && ra[n] == va[n]) n++;
    if (n > 0) {
        v.start += (int)n;
        v.refallele = vsub(ra, (int)n);
        v.varallele = vsub(va, (int)n);
        v.leftseq += vsub(ra, 0, (int)n);
        v.leftseq = vsub(v.leftseq, (int)n);
    }
    ra = v.refallele;
    va = v.varallele;
    int k = 1;
    while ((int)ra.size() - k > 0 && (int)va.size() - k > 0 && vsub(ra, -k, 1) == vsub(va, -k, 1)) k++;
    if (k > 1) {
        v.end -= k - 1;
        v.refallele = vsub(ra, 0, 1 - k);
        v.varallele = vsub(va, 0, 1 - k);
        v.rightseq = vsub(ra, 1 - k, k - 1) + vsub(v.rightseq, 0, 1 - k);
    }
}

// isNoise clears the counts of a variant it calls noise
bool is_noise(SV& v, const Config& c) {
    const double q = v.qmean;
    if (((q < 4.5 || (q < 12 && !v.qstd)) && v.poscov <= 3) || (q < c.goodq && v.freq < 2 * c.lofreq && v.poscov <= 1)) {
        v.tpc -= v.poscov;
        v.poscov = 0;
        v.fwd = 0;
        v.rev = 0;
        v.freq = 0;
        v.hifreq = 0;
        return true;
    }
    return false;
}

bool minus_num_num(const std::string& s) {  // patterns.h:18, "-\d\d" anywhere
    for (size_t i = 0; i + 2 < s.size(); i++)
        if (s[i] == '-' && s[i + 1] >= '0' && s[i + 1] <= '9' && s[i + 2] >= '0' && s[i + 2] <= '9') return true;
    return false;
}

// ------------------------------------------------------------------ printer
void fisher_odds(std::string& s, int rf, int rr, int af, int ar) {
    s += fmt_f(fisher_two_sided(rf, rr, af, ar, nullptr, nullptr));
    s += '\t';
    // the log of (ad/bc + bc/ad) is taken in single precision over the counts plus one half
    const float a = (float)(rf + 0.5), b = (float)(rr + 0.5), cc = (float)(af + 0.5), d = (float)(ar + 0.5);
    const float ad = a * d, bc = b * cc;
    const float ratio = std::log(ad / bc + bc / ad);
    s += fmt_f((double)ratio);
    s += '\t';
}

void sample_block(std::string& s, const SV& v, const SV* genotype_of) {
    s += std::to_string(v.tpc) + '\t' + std::to_string(v.poscov) + '\t' + std::to_string(v.rfc) + '\t' +
         std::to_string(v.rrc) + '\t' + std::to_string(v.fwd) + '\t' + std::to_string(v.rev) + '\t';
    s += v.genotype.empty() ? std::string("0") : genotype_of->genotype;
    s += '\t' + fmt_f(v.freq) + '\t' + v.bias + '\t' + fmt_f(v.pmean) + '\t' + (v.pstd ? "1" : "0") + '\t' +
         fmt_f(v.qmean) + '\t' + (v.qstd ? "1" : "0") + '\t' + fmt_f(v.mq) + '\t' + fmt_f(v.hlr) + '\t' +
         fmt_f(v.hifreq) + '\t' + fmt_f(v.extrafreq) + '\t' + fmt_f(v.nm) + '\t';
}

struct Printer {
    const Config& c;
    const std::string& sample;
    const rv_region& R;
    std::string out;

    void line(const SV* begin, const SV* end, const SV* tumor, const SV* normal, const std::string& label) {
        static const char* null18 = "0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t";
        std::string& s = out;
        s += sample + '\t' + (R.gene ? R.gene : "") + '\t' + (R.chr ? R.chr : "") + '\t';
        if (begin)
            s += std::to_string(begin->start) + '\t' + std::to_string(begin->end) + '\t' + begin->refallele + '\t' +
                 begin->varallele + '\t';
        else
            s += "0\t0\t0\t0\t";
        if (tumor)
            sample_block(s, *tumor, tumor);
        else
            s += null18;
        if (c.fisher) {
            if (tumor)
                fisher_odds(s, tumor->rfc, tumor->rrc, tumor->fwd, tumor->rev);
            else
                s += "0\t0\t";
        }
        if (normal) {
            // the second sample's genotype column prints the first sample's genotype
            if (!normal->genotype.empty() && !tumor)
                throw std::runtime_error("somatic printer: second-sample genotype without a first-sample variant");
            sample_block(s, *normal, tumor);
        } else {
            s += null18;
        }
        if (c.fisher) {
            if (normal)
                fisher_odds(s, normal->rfc, normal->rrc, normal->fwd, normal->rev);
            else
                s += "0\t0\t";
        }
        if (end)
            s += std::to_string(end->shift3) + '\t' + fmt_f(end->msi) + '\t' + std::to_string(end->msint) + '\t' +
                 (end->leftseq.empty() ? std::string("0") : end->leftseq) + '\t' +
                 (end->rightseq.empty() ? std::string("0") : end->rightseq) + '\t';
        else
            s += "\t\t\t\t\t";
        s += std::string(R.chr ? R.chr : "") + ':' + std::to_string(R.start) + '-' + std::to_string(R.end) + '\t';
        s += label + '\t';
        s += (begin ? begin->vartype : std::string()) + '\t';
        s += (tumor ? fmt_f(tumor->duprate) : std::string("0")) + "\t0\t";
        s += (normal ? fmt_f(normal->duprate) : std::string("0")) + "\t0\t";
        if (c.fisher) {
            const int t_tot = tumor ? tumor->tpc : 0, t_var = tumor ? tumor->poscov : 0;
            const int n_tot = normal ? normal->tpc : 0, n_var = normal ? normal->poscov : 0;
            const int tref = std::max(0, t_tot - t_var), rref = std::max(0, n_tot - n_var);
            fisher_odds(s, t_var, tref, n_var, rref);
            const double tv = tumor ? tumor->freq : 0, nv = normal ? normal->freq : 0;
            s += fmt_f(std::log(std::max(tv, 0.0001) / std::max(nv, 0.0001))) + '\t';
            s += fmt_f(std::log(((float)t_var + 0.5) / ((float)n_var + 0.5))) + '\t';
        }
        s += '\n';
    }

    std::string determinate(SVars& vs, SV& standard, SV& cmp) {
        std::string type;
        const SV* ref = vs.has_ref ? &vs.ref : nullptr;
        if (good_var(cmp, ref, standard.vartype, c)) {
            if (standard.freq > (1 - c.lofreq) && cmp.freq < 0.8 && cmp.freq > 0.2)
                type = "LikelyLOH";
            else if (cmp.freq < c.lofreq || cmp.poscov <= 1)
                type = "LikelySomatic";
            else
                type = "Germline";
        } else {
            type = (cmp.freq < c.lofreq || cmp.poscov <= 1) ? "LikelySomatic" : "AFDiff";
        }
        if (is_noise(cmp, c) && standard.vartype == "SNV") type = "StrongSomatic";
        return type;
    }

    void one_sample(SVars& vs, bool first_cover, const std::string& label) {
        const SV* ref = vs.has_ref ? &vs.ref : nullptr;
        for (SV& v : vs.vars) {
            v.vartype = var_type(v);
            if (!good_var(v, ref, v.vartype, c)) continue;
            if (v.vartype == "Complex") adj_complex(v);
            if (first_cover)
                line(&v, &v, nullptr, &v, label);
            else
                line(&v, &v, &v, nullptr, label);
        }
    }

    void first_sample(SVars& v1, SVars& v2) {
        const SV* ref1 = v1.has_ref ? &v1.ref : nullptr;
        size_t k = 0;
        while (k < v1.vars.size() && good_var(v1.vars[k], ref1, v1.vars[k].vartype, c)) {
            SV& vref = v1.vars[k];
            const std::string nt = vref.desc;
            vref.vartype = var_type(vref);
            if (vref.vartype == "Complex") adj_complex(vref);
            if (SV* v2nt = v2.find(nt)) {
                const std::string type = determinate(v2, vref, *v2nt);
                line(&vref, v2nt, &vref, v2nt, type);
            } else {
                SV hold;
                const SV* for_print;
                if (!v2.vars.empty()) {
                    hold.tpc = v2.vars[0].tpc;
                    hold.rfc = v2.vars[0].rfc;
                    hold.rrc = v2.vars[0].rrc;
                    for_print = &hold;
                } else {
                    for_print = v2.has_ref ? &v2.ref : nullptr;
                }
                if (vref.vartype != "SNV" && (nt.size() > 10 || minus_num_num(nt)) && vref.poscov < c.minr + 3 &&
                    nt.find('<') != std::string::npos)
                    throw std::runtime_error("somatic: structural descriptors (combineAnalysis) are not supported");
                line(&vref, &vref, &vref, for_print, "StrongSomatic");
            }
            k++;
        }
        if (k > 0 || v2.vars.empty()) return;
        const SV* ref2 = v2.has_ref ? &v2.ref : nullptr;
        for (SV& v2var : v2.vars) {
            v2var.vartype = var_type(v2var);
            if (!good_var(v2var, ref2, v2var.vartype, c)) continue;
            const std::string nt = v2var.desc;
            if (SV* v1nt = v1.find(nt)) {
                const std::string type = v1nt->freq < c.lofreq ? "LikelyLOH" : "Germline";
                if (v2var.vartype == "Complex") adj_complex(*v1nt);
                v1nt->vartype = var_type(*v1nt);
                line(v1nt, &v2var, v1nt, &v2var, type);
            } else {
                const SV* v1var = v1.vars.empty() ? nullptr : &v1.vars[0];
                SV hold;
                hold.tpc = v1var && v1var->tpc != 0 ? v1var->tpc : 0;
                hold.rfc = ref1 ? ref1->fwd : 0;
                hold.rrc = ref1 ? ref1->rev : 0;
                hold.genotype = v1var ? v1var->genotype : (ref1 ? ref1->desc + "/" + ref1->desc : "N/N");
                if (v2var.vartype == "Complex") adj_complex(v2var);
                line(&v2var, &v2var, &hold, &v2var, "StrongLOH");
            }
        }
    }

    void second_sample(SVars& v1, SVars& v2) {
        const SV* ref2 = v2.has_ref ? &v2.ref : nullptr;
        for (SV& v2var : v2.vars) {
            v2var.vartype = var_type(v2var);
            if (!good_var(v2var, ref2, v2var.vartype, c)) continue;
            const std::string& nt = v2var.desc;
            if (SV* v1nt = v1.find(nt)) v1nt->poscov = 0;
            if (v2.find(nt)->poscov < c.minr + 3 && nt.find('<') != std::string::npos &&
                (nt.size() > 10 || minus_num_num(nt)))
                throw std::runtime_error("somatic: structural descriptors (combineAnalysis) are not supported");
            const SV* for_print = v1.has_ref ? &v1.ref : nullptr;
            if (v2var.vartype == "Complex") adj_complex(v2var);
            line(&v2var, &v2var, for_print, &v2var, "StrongLOH");
        }
    }
};

// ------------------------------------------------------------------ one sample on the device
struct Sample {
    PileupResult pr;
    std::unique_ptr<RowCache> rc;
    std::unique_ptr<RState> S;
    CxHost cx;
    HostRef ref;
    DenseRow* rows = nullptr;
    int dense_lo = 0, dense_n = 0;
    RecSpan recs;                 // first scoring pass (pinned host memory of the engine)
    std::vector<uint8_t> status;  // per slot: dense rows, then record-list positions
    RecSpan ref_recs;             // reference-only pass

    // status of position p: 1 variants, 3 reference only, 0 not kept
    int state(int p) const {
        auto it = std::lower_bound(cx.pos.begin(), cx.pos.end(), p);
        if (it != cx.pos.end() && *it == p) return status[(size_t)dense_n + (size_t)(it - cx.pos.begin())];
        const int i = p - dense_lo;
        return (i >= 0 && i < dense_n) ? status[(size_t)i] : 0;
    }
};

class DeviceRows : public RowSource {
  public:
    explicit DeviceRows(Engine& e) : e_(e) {}
    const DenseRow* fetch(const std::vector<int>& rows) override { return e_.fetch_rows(rows); }

  private:
    Engine& e_;
};

class DeviceReads : public ReadQuery {
  public:
    explicit DeviceReads(Engine& e) : e_(e) {}
    bool no_passing(int start, int end) override {
        bool none = true;
        if (!e_.no_passing(start, end, none)) throw std::runtime_error("noPassingReads on the device failed");
        return none;
    }
    bool find35_batch(const std::vector<F35Job>& jobs, std::vector<F35>& out) override { return e_.find35(jobs, out); }

  private:
    Engine& e_;
};

// uploads the sample's record list and runs one scoring pass in `mode`
rv_status score_sample(Engine& e, Sample& sm, const rv_region& region, const rv_contig& contig, int mode,
                       const std::vector<uint8_t>* want, RecSpan& recs, std::vector<uint8_t>* status) {
    const CxHost& cx = sm.cx;
    ScoreIn in{};
    in.rows = sm.rows;
    in.dense_lo = sm.dense_lo;
    in.dense_n = sm.dense_n;
    in.n_cx = (int)cx.pos.size();
    in.ref = e.ref_dev;
    in.ref_lo = sm.ref.lo;
    in.ref_hi = sm.ref.hi;
    in.chr_len = contig.length;
    in.region_start = region.start;
    in.region_end = region.end;
    in.cfg = e.cfg;
    in.somatic = mode;
    auto up = [&](DevBuf& b, const void* src, size_t bytes) -> void* {
        void* d = e.grow(b, bytes);
        if (d && bytes && hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, e.stream) != hipSuccess) return nullptr;
        return d;
    };
    in.cx_pos = (const int32_t*)up(e.b_cx_pos, cx.pos.data(), 4 * cx.pos.size());
    in.cx_off = (const int32_t*)up(e.b_cx_off, cx.off.data(), 4 * cx.off.size());
    in.cx = (const CxEntry*)up(e.b_cx, cx.ent.data(), sizeof(CxEntry) * cx.ent.size());
    in.cx_arena = (const char*)up(e.b_cx_arena, cx.arena.data(), cx.arena.size());
    in.cx_cov = (const int32_t*)up(e.b_cx_cov, cx.cov.data(), 4 * cx.cov.size());
    in.cx_key = (const uint8_t*)up(e.b_cx_key, cx.key.data(), cx.key.size());
    const size_t n_status = (size_t)sm.dense_n + cx.pos.size();
    in.status = (uint8_t*)e.grow(e.b_status, n_status);
    if (want) in.want = (const uint8_t*)up(e.b_want, want->data(), want->size());
    if (!in.cx_pos || !in.cx_off || !in.cx || !in.cx_arena || !in.cx_cov || !in.cx_key || !in.status ||
        (want && !in.want))
        return e.fail(RV_ERR_OOM, "device buffers for somatic scoring failed");
    unsigned long long hc[8];
    float ms = 0;
    std::vector<OutRec> unused_recs;
    std::vector<char> unused_arena;
    rv_status s = score_pass(e, in, n_status, unused_recs, unused_arena, status, hc, ms, &recs);
    e.stats.score_ms += ms;
    // the pageable uploads above are read by the kernels before score_pass returns
    return s;
}

// pileup -> realignment -> record list -> first scoring pass of one sample
rv_status run_sample(Engine& e, const rv_region& region, const rv_reads& reads, const rv_contig& contig,
                     int first_max_read_length, const std::set<int>* first_created, Sample& sm) {
    rv_status s = e.pileup(region, reads, contig, false, sm.pr);
    if (s != RV_OK) return s;
    // the second sample starts from the first one's maximum read length, and sees the reference
    // keys the first one's realignment created (they share one reference map)
    sm.pr.P.maxReadLength = std::max(sm.pr.P.maxReadLength, first_max_read_length);
    sm.ref.seq = sm.pr.ref_win.data();
    sm.ref.lo = sm.pr.W.ref_lo;
    sm.ref.hi = sm.pr.W.ref_lo + (int)sm.pr.ref_win.size() - 1;
    if (first_created) sm.ref.created = *first_created;
    sm.rows = e.rows;
    sm.dense_lo = sm.pr.P.dense_lo;
    sm.dense_n = sm.pr.P.dense_n;
    DeviceRows src(e);
    sm.rc = std::make_unique<RowCache>(&src, sm.dense_lo, sm.dense_n, &e.row_slots);
    struct Release {
        RowCache& rc;
        ~Release() { rc.release_slots(); }
    } release{*sm.rc};
    sm.S = std::make_unique<RState>(sm.pr.P, *sm.rc);
    DeviceReads rq(e);
    RealignStats rst;
    // screened mode: each sample's candidates are scored after its own pileup
    ScreenPlan plan;
    ScreenPlan* planp = nullptr;
    if (e.screen_on) {
        const auto q0 = std::chrono::steady_clock::now();
        s = e.screen_plan(sm.pr, plan, false);
        if (s != RV_OK) return s;
        planp = &plan;
        e.stats.screen_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - q0).count();
        e.stats.screen_candidates += (int64_t)plan.cand.size();
    }
    try {
        realign(e.cfg, region.start, region.end, sm.ref, contig.length, &rq, *sm.S, rst, false, planp);
        if (planp) e.stats.screen_skipped += plan.skipped;
        build_record_list(*sm.S, sm.cx);
    } catch (const std::exception& ex) {
        return e.fail(RV_ERR_DEVICE, std::string("realignment: ") + ex.what());
    }
    return score_sample(e, sm, region, contig, SCORE_SOMATIC_VARS, nullptr, sm.recs, &sm.status);
}

// per-position status of a sample over [lo, hi] (Sample::state for every position at once)
std::vector<uint8_t> states_of(const Sample& sm, int lo, int hi) {
    std::vector<uint8_t> st((size_t)(hi - lo + 1), 0);
    for (int i = 0; i < sm.dense_n; i++) {
        const int p = sm.dense_lo + i;
        if (p >= lo && p <= hi) st[(size_t)(p - lo)] = sm.status[(size_t)i];
    }
    for (size_t k = 0; k < sm.cx.pos.size(); k++) {
        const int p = sm.cx.pos[k];
        if (p >= lo && p <= hi) st[(size_t)(p - lo)] = sm.status[(size_t)sm.dense_n + k];
    }
    return st;
}

// a sample's records at the positions the printer reads (`need`, over [lo, hi]), bucketed by
// position: at each position the scoring pass's records in creation order, then the
// reference-only pass's
struct Bucketed {
    struct Item {
        const OutRec* o;
        const char* arena;
        int source;
    };
    std::vector<uint32_t> off;  // per position, into items
    std::vector<Item> items;

    // the position's Vars (include/Vars.h) as the reference's map holds them
    bool vars_at(size_t i, SVars& vs) const {
        if (off[i] == off[i + 1]) return false;
        for (uint32_t k = off[i]; k < off[i + 1]; k++) {
            const Item& it = items[k];
            SV v = to_sv(*it.o, it.arena);
            if (it.o->kind == OUT_VARIANT || it.o->kind == OUT_VARIANT_RAW) {
                vs.vars.push_back(std::move(v));
            } else {
                vs.has_ref = true;
                vs.ref = std::move(v);
            }
        }
        return true;
    }
};

Bucketed bucket(const Sample& sm, int lo, const std::vector<uint8_t>& need) {
    Bucketed b;
    b.off.assign(need.size() + 1, 0);
    auto slot = [&](const OutRec& o) -> long {
        const long i = (long)o.position - lo;
        return (i >= 0 && i < (long)need.size() && need[(size_t)i]) ? i : -1;
    };
    for (const auto* recs : {&sm.recs, &sm.ref_recs})
        for (const OutRec& o : *recs)
            if (const long i = slot(o); i >= 0) b.off[(size_t)i + 1]++;
    for (size_t i = 0; i < need.size(); i++) b.off[i + 1] += b.off[i];
    b.items.resize(b.off.back());
    std::vector<uint32_t> fill(b.off.begin(), b.off.end() - 1);
    for (int src = 0; src < 2; src++) {
        const RecSpan& recs = src ? sm.ref_recs : sm.recs;
        const char* arena = recs.arena;
        for (const OutRec& o : recs)
            if (const long i = slot(o); i >= 0) b.items[fill[(size_t)i]++] = {&o, arena, src};
    }
    for (size_t i = 0; i < need.size(); i++)
        if (b.off[i + 1] - b.off[i] > 1)
            std::sort(b.items.begin() + b.off[i], b.items.begin() + b.off[i + 1],
                      [](const Bucketed::Item& x, const Bucketed::Item& y) {
                          return x.source != y.source ? x.source < y.source : x.o->order < y.o->order;
                      });
    return b;
}

}  // namespace

rv_status run_region_somatic(Engine& e, const rv_region& region, const rv_reads& first, const rv_reads& second,
                             const rv_contig& contig, CallSet& out) {
    const auto w0 = std::chrono::steady_clock::now();
    // with a second sample ToVarsBuilder keeps reference-only and sub-threshold positions
    struct CfgGuard {
        Engine& e;
        int saved;
        ~CfgGuard() { e.cfg.hasBam2 = saved; }
    } guard{e, e.cfg.hasBam2};
    e.cfg.hasBam2 = 1;
    e.pinned_reset();
    e.f35_device_pairs = 0;
    e.f35_ms = 0;
    e.stats.score_ms = 0;
    e.stats.screen_ms = 0;
    e.stats.screen_list_ms = 0;
    e.stats.screen_candidates = 0;
    e.stats.screen_skipped = 0;
    for (double& d : e.stats.detail_ms) d = 0;
    auto lap = [](std::chrono::steady_clock::time_point& t0) {
        const auto t1 = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t0 = t1;
        return ms;
    };
    auto tl = w0;
    Sample t, n;
    rv_status s = run_sample(e, region, first, contig, 0, nullptr, t);
    if (s != RV_OK) return s;
    e.stats.detail_ms[0] = lap(tl);
    e.swap_row_buffers();  // the second pileup must not overwrite the first sample's rows
    s = run_sample(e, region, second, contig, t.pr.P.maxReadLength, &t.ref.created, n);
    if (s != RV_OK) return s;
    e.stats.detail_ms[1] = lap(tl);
    // the reference-only positions the printer will read: where the other sample has variants
    const int lo = region.start, hi = std::max(region.start, region.end);
    const std::vector<uint8_t> s1 = states_of(t, lo, hi), s2 = states_of(n, lo, hi);
    auto wants = [&](const Sample& self, const std::vector<uint8_t>& other) {
        std::vector<uint8_t> w(self.status.size(), 0);
        bool any = false;
        auto mark = [&](size_t slot, int p) {
            if (p < region.start || p > region.end) return;
            if (self.status[slot] == 3 && other[(size_t)(p - lo)] == 1) {
                w[slot] = 1;
                any = true;
            }
        };
        for (int i = 0; i < self.dense_n; i++) mark((size_t)i, self.dense_lo + i);
        for (size_t k = 0; k < self.cx.pos.size(); k++) mark((size_t)self.dense_n + k, self.cx.pos[k]);
        return std::make_pair(w, any);
    };
    auto wt = wants(t, s2), wn = wants(n, s1);
    if (wn.second) {
        s = score_sample(e, n, region, contig, SCORE_SOMATIC_REF, &wn.first, n.ref_recs, nullptr);
        if (s != RV_OK) return s;
    }
    if (wt.second) {
        s = score_sample(e, t, region, contig, SCORE_SOMATIC_REF, &wt.first, t.ref_recs, nullptr);
        if (s != RV_OK) return s;
    }
    e.stats.detail_ms[2] = lap(tl);
    // ---- the printer: the first sample's kept positions, ascending.  Only positions where one
    // sample has variants, or only the first is kept, print anything; the rest (both reference
    // only) are passed over without their records.
    std::vector<uint8_t> need1(s1.size(), 0), need2(s1.size(), 0);
    for (size_t i = 0; i < s1.size(); i++) {
        if (s1[i] == 0) continue;
        need1[i] = (s2[i] == 0 || s1[i] == 1 || s2[i] == 1) ? 1 : 0;
        need2[i] = (s2[i] != 0 && (s1[i] == 1 || s2[i] == 1)) ? 1 : 0;
    }
    Bucketed b1, b2;
    {
        std::exception_ptr berr;
        std::thread tb([&] {
            try {
                b2 = bucket(n, lo, need2);
            } catch (...) {
                berr = std::current_exception();
            }
        });
        try {
            b1 = bucket(t, lo, need1);
        } catch (...) {
            tb.join();
            throw;
        }
        tb.join();
        if (berr) std::rethrow_exception(berr);
    }
    e.stats.detail_ms[3] = lap(tl);
    // -N 'first|second' prints the first name (Launcher.cpp getSampleNamesSomatic)
    const std::string sample = e.sample.substr(0, e.sample.find('|'));
    // positions print independently: chunks of them on host threads, texts joined in order
    const size_t npos = s1.size();
    const size_t n_chunks = std::max<size_t>(1, std::min<size_t>(256, npos / 4096));
    std::vector<std::string> texts(n_chunks);
    std::vector<std::exception_ptr> errs(n_chunks);
    std::atomic<size_t> next_chunk{0};
    auto print_chunks = [&]() {
        for (;;) {
            const size_t ch = next_chunk.fetch_add(1);
            if (ch >= n_chunks) return;
            Printer P{e.cfg, sample, region, {}};
            try {
                for (size_t i = npos * ch / n_chunks; i < npos * (ch + 1) / n_chunks; i++) {
                    const int st1 = s1[i], st2 = s2[i];
                    if (st1 == 0) continue;
                    SVars a, b;
                    const bool has_a = b1.vars_at(i, a);
                    if (st2 == 0) {
                        if (has_a) P.one_sample(a, false, "SampleSpecific");
                        continue;
                    }
                    if (st1 != 1 && st2 != 1) continue;
                    b2.vars_at(i, b);  // a sample without records here takes an empty Vars
                    if (st1 == 1)
                        P.first_sample(a, b);
                    else
                        P.second_sample(a, b);
                }
            } catch (...) {
                errs[ch] = std::current_exception();
            }
            texts[ch] = std::move(P.out);
        }
    };
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>({(size_t)std::min(hw, 16u), n_chunks});
        std::vector<std::thread> th;
        for (size_t k = 1; k < nt; k++) th.emplace_back(print_chunks);
        print_chunks();
        for (auto& x : th) x.join();
    }
    std::string text;
    size_t total = 0;
    for (const auto& x : texts) total += x.size();
    text.reserve(total);
    try {
        for (size_t ch = 0; ch < n_chunks; ch++) {
            if (errs[ch]) std::rethrow_exception(errs[ch]);
            text += texts[ch];
        }
    } catch (const std::exception& ex) {
        return e.fail(RV_ERR_UNSUPPORTED, ex.what());
    }
    e.stats.detail_ms[4] = lap(tl);
    out.text = std::move(text);
    out.calls.clear();
    out.string_arenas.clear();
    e.stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    e.stats.realign_device_pairs = e.f35_device_pairs;
    e.stats.realign_device_ms = e.f35_ms;
    // the row table may hold the second sample's rows no longer: both caches released above
    e.release_later(std::move(t.S), std::move(t.rc), std::move(n.S), std::move(n.rc), std::move(t.pr),
                    std::move(n.pr));
    return RV_OK;
}

}  // namespace rv
