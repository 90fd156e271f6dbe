// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// rvcpu command line: the reference's simple-mode region loop (src/modes/simpleMode.cpp:210-387,
// one_region_run :18-64) over the project's read-batch files, with the reference's option names
// (src/Launcher.cpp:295-496) for the options on the calling path.  Region selection follows the
// htslib region iterator (records whose [pos, endpos) overlaps [start-1, end)); the reference
// window follows RecordPreprocessor::makeReference (src/recordPreprocessor.cpp:41-78).
//
//   rvcpu -G ref.fa -b reads.rvb -N sample (-R chr:s-e | -i regions.bed [-c 1 -S 2 -E 3 -g 4])
//         [-f F] [-k 0|1] [--fisher] [-q Q] [-m M] [-X X] [-P P] [-r R] [-o O] [-O MQ] [-B B]
//         [-Q MAPQ] [-F hex] [-T T] [-M M] [-p] [-3] [-t] [-Y ext] [-x ext] [--out FILE]
//         [--dump-pileup FILE] [--dump-realigned FILE] [--dump-order FILE] [--time]
//   rvcpu --fisher-table IN OUT
// --dump-order writes the iteration order of nonInsertionVariants after the pileup and after
// realignment, and of alignedVariants, one line each per region.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "rv_oracle.h"

using namespace orc;

namespace {

struct Fasta {
    std::string name;
    std::string seq;
};

Fasta read_fasta(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open fasta " + path);
    Fasta fa;
    std::string line;
    bool first = true;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (!line.empty() && line[0] == '>') {
            if (!first) break;  // single-contig inputs
            first = false;
            std::string n = line.substr(1);
            size_t sp = n.find_first_of(" \t");
            fa.name = sp == std::string::npos ? n : n.substr(0, sp);
            continue;
        }
        fa.seq += line;
    }
    return fa;
}

template <class T>
void rd(std::ifstream& f, T* p, size_t n) {
    f.read(reinterpret_cast<char*>(p), (std::streamsize)(sizeof(T) * n));
    size_t pad = (8 - (sizeof(T) * n) % 8) % 8;
    char tmp[8];
    if (pad) f.read(tmp, (std::streamsize)pad);
    if (!f) throw std::runtime_error("short read-batch file");
}

// read-batch file (written by rabbitvar_amd/synth): "RVB1", version, counts, then SoA arrays
// each padded to 8 bytes
std::vector<Read> read_rvb(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open reads " + path);
    char magic[4];
    f.read(magic, 4);
    if (std::memcmp(magic, "RVB1", 4) != 0) throw std::runtime_error("bad read-batch magic");
    uint32_t ver;
    f.read(reinterpret_cast<char*>(&ver), 4);
    int64_t n, nb, nc;
    f.read(reinterpret_cast<char*>(&n), 8);
    f.read(reinterpret_cast<char*>(&nb), 8);
    f.read(reinterpret_cast<char*>(&nc), 8);
    std::vector<int32_t> pos((size_t)n), nm((size_t)n), mpos((size_t)n), tid((size_t)n), mtid((size_t)n);
    std::vector<uint16_t> flag((size_t)n);
    std::vector<uint8_t> mapq((size_t)n), hasnm((size_t)n);
    std::vector<int64_t> so((size_t)n + 1), co((size_t)n + 1);
    std::vector<uint8_t> bases((size_t)nb), quals((size_t)nb);
    std::vector<uint32_t> cig((size_t)nc);
    rd(f, pos.data(), (size_t)n);
    rd(f, flag.data(), (size_t)n);
    rd(f, mapq.data(), (size_t)n);
    rd(f, hasnm.data(), (size_t)n);
    rd(f, nm.data(), (size_t)n);
    rd(f, mpos.data(), (size_t)n);
    rd(f, tid.data(), (size_t)n);
    rd(f, mtid.data(), (size_t)n);
    rd(f, so.data(), (size_t)n + 1);
    rd(f, co.data(), (size_t)n + 1);
    rd(f, bases.data(), (size_t)nb);
    rd(f, quals.data(), (size_t)nb);
    rd(f, cig.data(), (size_t)nc);
    std::vector<Read> out((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        Read& r = out[(size_t)i];
        r.pos0 = pos[(size_t)i];
        r.flag = flag[(size_t)i];
        r.mapq = mapq[(size_t)i];
        r.has_nm = hasnm[(size_t)i] != 0;
        r.nm = nm[(size_t)i];
        r.mpos0 = mpos[(size_t)i];
        r.tid = tid[(size_t)i];
        r.mtid = mtid[(size_t)i];
        r.seq.assign(reinterpret_cast<const char*>(bases.data()) + so[(size_t)i], (size_t)(so[(size_t)i + 1] - so[(size_t)i]));
        r.qual.assign(quals.begin() + so[(size_t)i], quals.begin() + so[(size_t)i + 1]);
        r.cigar.assign(cig.begin() + co[(size_t)i], cig.begin() + co[(size_t)i + 1]);
    }
    return out;
}

int ref_len_of(const Read& r) {
    int l = 0;
    for (uint32_t w : r.cigar) {
        int o = cop(w);
        if (o == OP_M || o == OP_D || o == OP_N || o == OP_EQ || o == OP_X) l += clen(w);
    }
    return l > 0 ? l : 1;
}

void dump_pileup(std::ostream& o, Pileup& P) {
    auto dump_map = [&](const char* tag, RHMap<int, VarMap*>& m) {
        std::vector<int> keys = m.keys_in_order();
        std::sort(keys.begin(), keys.end());
        for (int p : keys) {
            VarMap* vm = *m.find(p);
            for (auto& kv : *vm) {
                const Var* v = kv.second;
                char buf[512];
                snprintf(buf, sizeof buf, "%s\t%d\t%s\t%d\t%d\t%d\t%d\t%d\t%.17g\t%.17g\t%.17g\t%.17g\t%d\t%d\t%d\n", tag, p,
                         kv.first.c_str(), v->cnt, v->fwd, v->rev, v->hicnt, v->locnt, v->pmean, v->qmean, v->mqmean,
                         v->nm, (int)v->pstd, (int)v->qstd, v->extracnt);
                o << buf;
            }
        }
    };
    dump_map("V", P.nonIns);
    dump_map("I", P.ins);
    std::vector<std::pair<int, int>> cov(P.refCov.begin(), P.refCov.end());
    std::sort(cov.begin(), cov.end());
    for (auto& kv : cov) o << "C\t" << kv.first << "\t" << kv.second << "\n";
    auto dump_sc = [&](const char* tag, RHMap<int, Sclip*>& m) {
        std::vector<int> keys = m.keys_in_order();
        std::sort(keys.begin(), keys.end());
        for (int p : keys) {
            Sclip* s = *m.find(p);
            char buf[512];
            snprintf(buf, sizeof buf, "%s\t%d\t%d\t%d\t%d\t%d\t%d\t%.17g\t%.17g\t%.17g\t%.17g\t%zu\n", tag, p, s->cnt, s->fwd,
                     s->rev, s->hicnt, s->locnt, s->pmean, s->qmean, s->mqmean, s->nm, s->nt.size());
            o << buf;
            for (auto& e : s->nt) {
                o << tag << "n\t" << p << "\t" << e.first;
                for (char b : e.second.keys_in_order()) o << "\t" << b << ":" << e.second.at(b);
                o << "\n";
            }
        }
    };
    dump_sc("S5", P.sc5);
    dump_sc("S3", P.sc3);
    for (auto& e : P.delCount)
        for (auto& d : e.second) o << "DC\t" << e.first << "\t" << d.first << "\t" << d.second << "\n";
    for (auto& e : P.insCount)
        for (auto& d : e.second) o << "IC\t" << e.first << "\t" << d.first << "\t" << d.second << "\n";
    for (auto& e : P.mnp)
        for (auto& d : e.second) o << "MC\t" << e.first << "\t" << d.first << "\t" << d.second << "\n";
    o << "L\t" << P.maxReadLength << "\n";
}

}  // namespace

// --fisher-table IN OUT: one "n11 n12 n21 n22" table per input line -> "two left right" (%.17g)
static int fisher_table(const std::string& in, const std::string& out) {
    std::ifstream f(in);
    std::ofstream o(out);
    long long a, b, c, d;
    char buf[128];
    while (f >> a >> b >> c >> d) {
        double l = 0, r = 0;
        double two = kt_fisher_two_sided((int)a, (int)b, (int)c, (int)d, &l, &r);
        snprintf(buf, sizeof buf, "%.17g %.17g %.17g\n", two, l, r);
        o << buf;
    }
    return 0;
}

namespace orc {
bool modify_cigar_record(const RefMap& ref, int& position, std::vector<uint32_t>& rec, int n_orig,
                         int& base, int& n_new, const std::string& seq, const uint8_t* qual, int lseq,
                         int indelLen);
std::string cigar_to_text(const uint32_t* w, int n);
std::string findmsi_kat(const std::string& t1, const std::string& t2, const std::string& left);
}  // namespace orc

// --cigar-kat IN OUT: the CIGAR rewrite on single reads, for known-answer tests.  IN: a line
// "REF SEQ" (the reference window, keys 1..len), then one read per line "POS CIGAR SEQ Q" (POS
// 1-based alignment start, Q one Phred value for every base).  OUT, per read:
// "changed position cigar", the CIGAR parseCigar then walks (Q may also be a comma-separated
// list, one value per base): modifyCigar's result read through
// the record's own CIGAR pointer with the returned operation count (parseCigar.cpp:548-574),
// then cleanupCigar (parseCigar.cpp:575, :1858-1892).
static int cigar_kat(const std::string& in, const std::string& out) {
    std::ifstream f(in);
    std::ofstream o(out);
    std::string tag, refseq;
    f >> tag >> refseq;
    orc::RefMap ref;
    ref.start = 1;
    ref.end = (int)refseq.size();
    ref.win = refseq;
    int pos;
    std::string cig, seq, qtok;
    static const std::string codes = "=ACMGRSVTWYHKDBN";
    while (f >> pos >> cig >> seq >> qtok) {
        std::vector<uint32_t> w(64);
        int n = 0;
        {
            std::string s = cig;
            while (!s.empty()) {
                size_t k = 0;
                long v = std::stol(s, &k);
                const std::string ops = "MIDNSHP=X";
                w[(size_t)n++] = ((uint32_t)v << 4) | (uint32_t)ops.find(s[k]);
                s = s.substr(k + 1);
            }
        }
        const int lseq = (int)seq.size();
        std::vector<uint32_t> rec((size_t)n + (size_t)(lseq + 8) / 8 + 1, 0u);
        for (int i = 0; i < n; i++) rec[(size_t)i] = w[(size_t)i];
        for (int i = 0; i < lseq; i++) {
            size_t cpos = codes.find(seq[(size_t)i]);
            uint32_t code = cpos == std::string::npos ? 15u : (uint32_t)cpos;
            int b = i / 2;
            uint32_t v = (i % 2 == 0) ? (code << 4) : code;
            rec[(size_t)n + (size_t)(b / 4)] |= v << (8 * (b % 4));
        }
        std::vector<uint8_t> qual((size_t)lseq, 0);
        if (qtok.find(',') == std::string::npos) {
            for (auto& x : qual) x = (uint8_t)std::stoi(qtok);
        } else {
            std::stringstream qs(qtok);
            std::string one;
            for (size_t i = 0; i < qual.size() && std::getline(qs, one, ','); i++) qual[i] = (uint8_t)std::stoi(one);
        }
        int idlen = 0;
        for (int i = 0; i < n; i++)
            if ((w[(size_t)i] & 15) == 1 || (w[(size_t)i] & 15) == 2) idlen += (int)(w[(size_t)i] >> 4);
        int base = 0, nn = 0, p = pos, n_cigar = n;
        bool changed = orc::modify_cigar_record(ref, p, rec, n, base, nn, seq, qual.data(), lseq, idlen);
        if (changed) n_cigar = nn;
        else p = pos;
        auto op = [&](int i) { return (int)(rec[(size_t)i] & 15); };
        auto m3 = [](int o) { return o == 0 || o == 7 || o == 8; };
        for (int i = 0; i < n_cigar; i++) {
            if (op(i) == 1) rec[(size_t)i] = (rec[(size_t)i] & ~15u) | 4u;
            else if (op(i) == 5) continue;
            else if (m3(op(i))) break;
        }
        for (int i = n_cigar - 1; i >= 0; i--) {
            if (op(i) == 1) rec[(size_t)i] = (rec[(size_t)i] & ~15u) | 4u;
            else if (op(i) == 5) continue;
            else if (m3(op(i))) break;
        }
        o << (changed ? 1 : 0) << " " << p << " " << orc::cigar_to_text(rec.data(), n_cigar) << "\n";
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::string(argv[1]) == "--fisher-table") return fisher_table(argv[2], argv[3]);
    if (argc == 4 && std::string(argv[1]) == "--cigar-kat") return cigar_kat(argv[2], argv[3]);
    // --findmsi TSEQ1 TSEQ2 LEFT ("-" for an empty string): findMSI's "msi shift3 msint"
    if (argc == 5 && std::string(argv[1]) == "--findmsi") {
        auto arg = [&](int i) { return std::string(argv[i]) == "-" ? std::string() : std::string(argv[i]); };
        printf("%s\n", orc::findmsi_kat(arg(2), arg(3), arg(4)).c_str());
        return 0;
    }
    Config c;
    int threads = 1;
    std::string fasta, bam, regionArg, bed, out = "./out.txt", dumpP, dumpR, dumpO;
    int colC = -1, colS = -1, colE = -1, colG = -1;
    bool timing = false;
    c.sample = "";
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto nxt = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
            return argv[++i];
        };
        if (a == "-G") fasta = nxt();
        else if (a == "-b") bam = nxt();
        else if (a == "-N") c.sample = nxt();
        else if (a == "-R") regionArg = nxt();
        else if (a == "-i" || a == "--bed") bed = nxt();
        else if (a == "-c") colC = std::stoi(nxt()) - 1;
        else if (a == "-S") colS = std::stoi(nxt()) - 1;
        else if (a == "-E") colE = std::stoi(nxt()) - 1;
        else if (a == "-g") colG = std::stoi(nxt()) - 1;
        else if (a == "-f") c.freq = std::stod(nxt());
        else if (a == "-k") c.performLocalRealignment = std::stoi(nxt()) == 1;
        else if (a == "--fisher") c.fisher = true;
        else if (a == "-q") c.goodq = std::stod(nxt());
        else if (a == "-m") c.mismatch = std::stoi(nxt());
        else if (a == "-X") c.vext = std::stoi(nxt());
        else if (a == "-P") c.readPosFilter = std::stoi(nxt());
        else if (a == "-r") c.minr = std::stoi(nxt());
        else if (a == "-o") c.qratio = std::stod(nxt());
        else if (a == "-O") c.mapq = std::stod(nxt());
        else if (a == "-B") c.minBiasReads = std::stoi(nxt());
        else if (a == "-Q") c.mappingQuality = std::stoi(nxt());
        else if (a == "-F") c.samfilter = std::stoi(nxt(), nullptr, 16);
        else if (a == "-T") c.trimBasesAfter = std::stoi(nxt());
        else if (a == "-M") c.minmatch = std::stoi(nxt());
        else if (a == "-p") {
            c.doPileup = true;
        } else if (a == "-3") c.moveIndelsTo3 = true;
        else if (a == "-t") c.removeDuplicatedReads = true;
        else if (a == "-Z") {
            c.isDownsampling = true;
            c.downsampling = std::stod(nxt());
        }
        else if (a == "-Y") c.referenceExtension = std::stoi(nxt());
        else if (a == "-x") c.numberNucleotideToExtend = std::stoi(nxt());
        else if (a == "--out") out = nxt();
        else if (a == "--dump-pileup") dumpP = nxt();
        else if (a == "--dump-realigned") dumpR = nxt();
        else if (a == "--dump-order") dumpO = nxt();
        else if (a == "--time") timing = true;
        else if (a == "--th") threads = std::max(1, std::stoi(nxt()));
        else {
            std::cerr << "unknown option " << a << "\n";
            return 2;
        }
    }
    if (c.doPileup) {
        c.freq = -1;
        c.minr = 0;
    }
    // somatic mode: -b 'first|second' (Launcher.cpp:108-109, getSampleNamesSomatic: the printed
    // sample is the first name of -N 'first|second')
    std::string bam2;
    {
        const size_t bar = bam.find('|');
        if (bar != std::string::npos) {
            bam2 = bam.substr(bar + 1);
            bam = bam.substr(0, bar);
            c.hasBam2 = true;  // ToVarsBuilder keeps reference-only and sub-threshold positions
            const size_t nb = c.sample.find('|');
            if (nb != std::string::npos) c.sample = c.sample.substr(0, nb);
        }
    }
    Fasta fa = read_fasta(fasta);
    for (char& ch : fa.seq) ch = (char)toupper((unsigned char)ch);
    int chrLen = (int)fa.seq.size();
    std::vector<Read> reads = read_rvb(bam);
    std::vector<Read> reads2;
    if (!bam2.empty()) reads2 = read_rvb(bam2);
    std::vector<Region> regions;
    if (!regionArg.empty()) {
        Region r;
        size_t colon = regionArg.find(':');
        r.chr = regionArg.substr(0, colon);
        std::string rest = regionArg.substr(colon + 1);
        size_t c2 = rest.find(':');
        std::string range = c2 == std::string::npos ? rest : rest.substr(0, c2);
        r.gene = c2 == std::string::npos ? r.chr : rest.substr(c2 + 1);
        size_t dash = range.find('-');
        auto strip = [](std::string s) {
            s.erase(std::remove(s.begin(), s.end(), ','), s.end());
            return s;
        };
        r.start = std::stoi(strip(range.substr(0, dash)));
        r.end = dash == std::string::npos ? r.start : std::stoi(strip(range.substr(dash + 1)));
        r.start -= c.numberNucleotideToExtend;
        r.end += c.numberNucleotideToExtend;
        if (r.start > r.end) r.start = r.end;
        regions.push_back(r);
    } else {
        std::ifstream bf(bed);
        std::string line;
        while (std::getline(bf, line)) {
            if (line.empty() || line[0] == '#' || line.compare(0, 5, "track") == 0) continue;
            std::vector<std::string> cols;
            std::stringstream ss(line);
            std::string t;
            while (std::getline(ss, t, '\t')) cols.push_back(t);
            Region r;
            r.chr = cols[(size_t)(colC < 0 ? 0 : colC)];
            r.start = std::stoi(cols[(size_t)(colS < 0 ? 1 : colS)]);
            r.end = std::stoi(cols[(size_t)(colE < 0 ? 2 : colE)]);
            r.gene = colG >= 0 && (size_t)colG < cols.size() ? cols[(size_t)colG] : r.chr;
            // 1-based BED starts unless -z (RegionBuilder.cpp:100-103)
            regions.push_back(r);
        }
    }
    int maxRefLen = 1;
    for (const Read& r : reads) maxRefLen = std::max(maxRefLen, ref_len_of(r));
    for (const Read& r : reads2) maxRefLen = std::max(maxRefLen, ref_len_of(r));
    std::ofstream of(out, std::ios::binary);
    std::ofstream dp, dr;
    if (!dumpP.empty()) dp.open(dumpP);
    if (!dumpR.empty()) dr.open(dumpR);
    std::ofstream dord;
    if (!dumpO.empty()) dord.open(dumpO);
    auto dump_keys = [&](const char* tag, const std::vector<int>& keys) {
        dord << tag;
        for (int k : keys) dord << ' ' << k;
        dord << '\n';
    };
    auto t0 = std::chrono::steady_clock::now();
    std::atomic<long> nreads{0};
    double st[5] = {0, 0, 0, 0, 0};
    std::mutex st_mu;
    auto now = [] { return std::chrono::steady_clock::now(); };
    // --th N: regions on N threads, as the reference's OpenMP region loop (simpleMode.cpp:320);
    // the texts are written in region order.  Dumps and -Z (one process-wide rand() sequence) keep
    // one thread.
    if (dp.is_open() || dr.is_open() || dord.is_open() || c.isDownsampling) threads = 1;
    std::vector<std::string> texts(regions.size());
    auto region_text = [&](const Region& R, std::string& text_out) {
        // region iterator: records overlapping [R.start-1, R.end)
        auto region_reads = [&](const std::vector<Read>& rs) {
            std::vector<const Read*> out;
            auto lo = std::lower_bound(rs.begin(), rs.end(), R.start - 1 - maxRefLen,
                                       [](const Read& r, int v) { return r.pos0 < v; });
            for (auto it = lo; it != rs.end() && it->pos0 < R.end; ++it)
                if (it->pos0 + ref_len_of(*it) > R.start - 1) out.push_back(&*it);
            return out;
        };
        std::vector<const Read*> regionReads = region_reads(reads);
        nreads += (long)regionReads.size();
        RefMap ref;
        int ext = c.referenceExtension;
        int ss = R.start - c.numberNucleotideToExtend - ext < 1 ? 1 : R.start - c.numberNucleotideToExtend - ext;
        int se = R.end + c.numberNucleotideToExtend + ext > chrLen ? chrLen : R.end + c.numberNucleotideToExtend + ext;
        ref.start = ss;
        ref.end = se - 17;
        if (ref.end >= ref.start) ref.win = fa.seq.substr((size_t)(ss - 1), (size_t)(ref.end - ref.start + 1));
        if (!bam2.empty()) {
            // one_region_run_somt (somaticMode.cpp:83-128): the first sample, then the second over
            // the same reference map, starting from the first sample's maximum read length
            Aligned A1, A2;
            for (int k = 0; k < 2; k++) {
                std::vector<const Read*> rr = k == 0 ? regionReads : region_reads(reads2);
                if (k == 1) nreads += (long)rr.size();
                Counters ctr;
                std::vector<const Read*> filtered;
                preprocess_filter(c, rr, filtered, ctr);
                Pileup P;
                if (k == 1) P.maxReadLength = A1.maxReadLength;
                pileup_region(c, R, ref, chrLen, filtered, P, ctr);
                realign_region(c, R, ref, chrLen, rr, P);
                score_region(c, R, ref, chrLen, P, k == 0 ? A1 : A2);
            }
            text_out = format_somatic(c, R, A1, A2);
            return;
        }
        Counters ctr;
        std::vector<const Read*> filtered;
        auto a0 = now();
        preprocess_filter(c, regionReads, filtered, ctr);
        Pileup P;
        pileup_region(c, R, ref, chrLen, filtered, P, ctr);
        auto a1 = now();
        if (dp.is_open()) dump_pileup(dp, P);
        if (dord.is_open()) dump_keys("pileup", P.nonIns.keys_in_order());
        auto a2 = now();
        realign_region(c, R, ref, chrLen, regionReads, P);
        auto a3 = now();
        if (dr.is_open()) dump_pileup(dr, P);
        if (dord.is_open()) dump_keys("realigned", P.nonIns.keys_in_order());
        Aligned A;
        auto a4 = now();
        score_region(c, R, ref, chrLen, P, A);
        auto a5 = now();
        std::string text = format_simple(c, R, A);
        if (dord.is_open()) dump_keys("vars", A.vars.keys_in_order());
        auto a6 = now();
        {
            std::lock_guard<std::mutex> lk(st_mu);
            st[0] += std::chrono::duration<double>(a1 - a0).count();
            st[1] += std::chrono::duration<double>(a3 - a2).count();
            st[2] += std::chrono::duration<double>(a5 - a4).count();
            st[3] += std::chrono::duration<double>(a6 - a5).count();
            if (ctr.exceptions) std::cerr << "[rvcpu] " << ctr.exceptions << " reads raised during parse\n";
        }
        text_out = std::move(text);
    };
    if (threads <= 1) {
        for (size_t i = 0; i < regions.size(); i++) {
            region_text(regions[i], texts[i]);
            of << texts[i];
            texts[i].clear();
        }
    } else {
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++)
            pool.emplace_back([&] {
                for (size_t i; (i = next.fetch_add(1)) < regions.size();) region_text(regions[i], texts[i]);
            });
        for (auto& th : pool) th.join();
        for (auto& t : texts) of << t;
    }
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();
    if (timing)
        printf("{\"reads\": %ld, \"seconds\": %.6f, \"reads_per_s\": %.3f, \"pileup_s\": %.6f, \"realign_s\": %.6f, "
               "\"score_s\": %.6f, \"output_s\": %.6f}\n",
               nreads.load(), sec, nreads.load() / sec, st[0], st[1], st[2], st[3]);
    return 0;
}
