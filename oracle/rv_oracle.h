// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// rvcpu: a CPU restatement, written for this project, of the reference's per-region calling
// path (RabbitVar: RecordPreprocessor -> CigarParser -> VariationRealigner -> ToVarsBuilder ->
// SimpleMode output).  It exists to check the HIP product path; only tests/, bench.py's
// cpu_baseline leg and __graft_entry__.smoke() may load it, and only as the checker.
//
// Language: C++17 built with g++ (the path is string/map heavy: descriptor strings, std::regex
// CIGAR rules, ordered maps).  Data structures follow the reference's semantics, not its code:
//   * int-keyed maps whose iteration order is observable use RHMap (robin_hood slot order);
//   * the reference window reproduces robin_hood operator[]'s insert-'\0' side effect and
//     .at()'s throw (include/VariationUtils.h:336-349, A4 hazards);
//   * CIGAR words are kept in an emulated BAM record buffer so that the reference's reads one
//     word past the CIGAR array see the same bytes (packed SEQ) as the reference does.
#pragma once
#include <cstdint>
#include <deque>
#include <map>
#include <stdexcept>
#include <algorithm>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "rh_map.h"

namespace orc {

// BAM CIGAR op codes (SAM spec)
enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_EQ = 7, OP_X = 8 };
inline int cop(uint32_t w) { return (int)(w & 0xf); }
inline int clen(uint32_t w) { return (int)(w >> 4); }
inline uint32_t cword(int len, int op) { return ((uint32_t)len << 4) | (uint32_t)op; }

// Configuration fields used on the path (reference: include/Configuration.h:98-401 defaults,
// src/Launcher.cpp:373-494 option mapping).
struct Config {
    double goodq = 22.5;             // -q
    int vext = 2;                    // -X
    int mismatch = 8;                // -m
    int trimBasesAfter = 0;          // -T
    bool performLocalRealignment = true;  // -k
    int indelsize = 50;              // -I
    double bias = 0.05;
    int minBiasReads = 2;            // -B
    int minr = 2;                    // -r
    double freq = 0.01;              // -f
    bool moveIndelsTo3 = false;      // -3
    int samfilter = 0x504;           // -F
    int readPosFilter = 5;           // -P
    double qratio = 1.5;             // -o
    double mapq = 0;                 // -O
    bool doPileup = false;           // -p
    double lofreq = 0.05;            // -V
    int minmatch = 0;                // -M
    bool uniqueModeAlignmentEnabled = false;   // -u
    bool uniqueModeSecondInPairEnabled = false;  // --UN
    bool fisher = false;             // --fisher
    bool chimeric = false;           // --chimeric
    int SVMINLEN = 1000;             // -L
    int referenceExtension = 1200;   // -Y
    int numberNucleotideToExtend = 0;  // -x
    int mappingQuality = 0;          // -Q
    bool removeDuplicatedReads = false;  // -t
    bool isDownsampling = false;     // -Z given
    double downsampling = 0;         // -Z fraction
    int crisprCuttingSite = 0;       // -J
    bool hasBam2 = false;            // somatic (-b 'T|N')
    std::string sample;              // -N
};

struct Region {
    std::string chr;
    int start = 0, end = 0;
    std::string gene;
};

// One aligned record as the reference's region iterator delivers it.
struct Read {
    int32_t pos0 = 0;      // core.pos (0-based)
    uint16_t flag = 0;
    uint8_t mapq = 0;
    int32_t nm = 0;
    bool has_nm = true;
    int32_t mpos0 = 0;     // core.mpos
    int32_t tid = 0, mtid = 0;
    std::vector<uint32_t> cigar;
    std::string seq;       // decoded bases (seq_nt16_str)
    std::vector<uint8_t> qual;
};

// Reference window (recordPreprocessor.cpp:41-78): keys [start, end] with end =
// min(len, region.end+x+ext) - 17, upper-cased.  operator[] on a missing key inserts '\0'
// (robin_hood operator[]), .at() throws.
struct RefMap {
    int start = 1, end = 0;          // inclusive window
    std::string win;                 // win[i - start]
    std::unordered_map<int, char> extra;  // keys created by operator[] outside the window
    bool count(int k) const {
        if (k >= start && k <= end) return true;
        return extra.count(k) != 0;
    }
    char at(int k) const {
        if (k >= start && k <= end) return win[(size_t)(k - start)];
        auto it = extra.find(k);
        if (it == extra.end()) throw std::out_of_range("ref.at");
        return it->second;
    }
    char get(int k) {  // operator[]
        if (k >= start && k <= end) return win[(size_t)(k - start)];
        auto it = extra.find(k);
        if (it == extra.end()) {
            extra.emplace(k, '\0');
            return '\0';
        }
        return it->second;
    }
};

struct Var {  // include/Variation.h:12-162
    int cnt = 0, fwd = 0, rev = 0;
    double pmean = 0, qmean = 0, mqmean = 0, nm = 0;
    int locnt = 0, hicnt = 0;
    bool pstd = false, qstd = false;
    int pp = 0;
    double pq = 0;
    int extracnt = 0;
    void incDir(bool d) { d ? rev++ : fwd++; }
    int getDir(bool d) const { return d ? rev : fwd; }
    void addDir(bool d, int a) { d ? rev += a : fwd += a; }
    void subDir(bool d, int a) { d ? rev -= a : fwd -= a; }
};

struct Sclip : Var {  // include/Sclip.h:10-39
    std::map<int, RHMap<char, int>> nt;
    std::map<int, std::map<char, Var*>> seq;
    std::string sequence;
    bool used = false;
};

typedef std::map<std::string, Var*> VarMap;  // VariationMap (inner order never observable)

struct Pileup {
    RHMap<int, VarMap*> nonIns;   // nonInsertionVariants
    RHMap<int, VarMap*> ins;      // insertionVariants
    std::unordered_map<int, int> refCov;
    RHMap<int, Sclip*> sc5, sc3;
    std::map<int, std::map<std::string, int>> insCount, delCount, mnp;
    std::set<std::string> splice;
    int maxReadLength = 0;
    std::deque<Var> pool;
    std::deque<Sclip> sclipPool;
    std::deque<VarMap> mapPool;

    Var* newVar() {
        pool.emplace_back();
        return &pool.back();
    }
    VarMap* newMap() {
        mapPool.emplace_back();
        return &mapPool.back();
    }
    Sclip* newSclip() {
        sclipPool.emplace_back();
        return &sclipPool.back();
    }
    // getVariation (include/VariationUtils.h:148-195)
    Var* getVariation(RHMap<int, VarMap*>& h, int pos, const std::string& d) {
        VarMap** vm = h.find(pos);
        if (!vm) {
            VarMap* m = newMap();
            h.insert(pos, m);
            Var* v = newVar();
            (*m)[d] = v;
            return v;
        }
        auto it = (*vm)->find(d);
        if (it != (*vm)->end()) return it->second;
        Var* v = newVar();
        (**vm)[d] = v;
        return v;
    }
    // getVariationMaybe (include/VariationUtils.h:218-239)
    Var* getVariationMaybe(int pos, char refBase) {
        if (refBase == 0) return nullptr;
        VarMap** vm = nonIns.find(pos);
        if (!vm) return nullptr;
        auto it = (*vm)->find(std::string(1, refBase));
        return it == (*vm)->end() ? nullptr : it->second;
    }
};

// Output record (include/Variant.h:20-67)
struct Variant {
    std::string descriptionString;
    int positionCoverage = 0, varsCountOnForward = 0, varsCountOnReverse = 0;
    std::string strandBiasFlag = "0";
    double frequency = 0, meanPosition = 0;
    bool isAtLeastAt2Positions = false;
    double meanQuality = 0;
    bool hasAtLeast2DiffQualities = false;
    double meanMappingQuality = 0, highQualityToLowQualityRatio = 0, highQualityReadsFrequency = 0,
           extraFrequency = 0;
    int shift3 = 0;
    double msi = 0;
    int msint = 0;
    double numberOfMismatches = 0;
    int hicnt = 0, hicov = 0;
    std::string leftseq, rightseq;
    int startPosition = 0, endPosition = 0;
    int refReverseCoverage = 0, refForwardCoverage = 0, totalPosCoverage = 0;
    double duprate = 0;
    std::string genotype, varallele, refallele, vartype;
    int crispr = 0;
};

struct Vars {
    Variant* referenceVariant = nullptr;
    std::vector<Variant*> variants;
    std::map<std::string, Variant*> byDesc;
    std::string sv;
};

struct Aligned {
    RHMap<int, Vars*> vars;
    std::deque<Variant> vpool;
    std::deque<Vars> varsPool;
    int maxReadLength = 0;
};

struct Counters {
    long reads_in = 0, reads_filtered = 0, reads_parsed = 0, exceptions = 0;
};

// stage entry points
void preprocess_filter(const Config& c, const std::vector<const Read*>& all, std::vector<const Read*>& out,
                       Counters& ctr);
void pileup_region(const Config& c, const Region& r, RefMap& ref, int chrLen,
                   const std::vector<const Read*>& reads, Pileup& P, Counters& ctr);
void realign_region(const Config& c, const Region& r, RefMap& ref, int chrLen,
                    const std::vector<const Read*>& regionReads, Pileup& P);
void score_region(const Config& c, const Region& r, RefMap& ref, int chrLen, Pileup& P, Aligned& A);
std::string format_simple(const Config& c, const Region& r, Aligned& A);
// somatic mode (somaticMode.cpp:311-735): A1 = first (tumor) sample, A2 = second (normal) sample
std::string format_somatic(const Config& c, const Region& r, Aligned& A1, Aligned& A2);

// helpers shared by stages
double kt_fisher_two_sided(int n11, int n12, int n21, int n22, double* left, double* right);
std::string to_string_f(double v);  // std::to_string(double) == "%f"

}  // namespace orc
