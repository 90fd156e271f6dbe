// rvsynth: seeded synthetic inputs for the per-region calling path.
//
// Writes <out>/ref.fa (one contig, chrS), <out>/reads.rvb (aligned records as a structure-of-
// arrays read batch, sorted by position) and <out>/truth.tsv (planted variants).
//
// Model (per workload):
//   reference  iid bases, P(G)=P(C)=0.205, P(A)=P(T)=0.295; optional tandem-repeat tracts of
//              2-6 bp units covering a given fraction of the contig
//   variants   heterozygous (one haplotype): SNVs, short indels (left-normalised), optional long
//              deletions; none within 150 bp of a contig end
//   reads      paired-end, fragment start uniform on the haplotype, insert ~ N(mu, sd) clamped,
//              read 1 forward / read 2 reverse; Phred ~ round(N(35,5)) clamped to [2,41];
//              substitution errors at a fixed rate
//   alignment  the true alignment: M/I/D from the haplotype walk; a read may not start or end with
//              I or D (clipped instead); reads across long deletions are soft-clipped at the
//              breakpoint on the shorter side; NM = mismatches + indel bases; MAPQ 60
//
// Read-batch file layout ("RVB1"): magic, u32 version, i64 n_reads, i64 n_bases, i64 n_cigar, then
// i32 pos0[n], u16 flag[n], u8 mapq[n], u8 has_nm[n], i32 nm[n], i32 mpos0[n], i32 tid[n],
// i32 mtid[n], i64 seq_off[n+1], i64 cigar_off[n+1], u8 bases[nb] (ASCII), u8 quals[nb] (Phred),
// u32 cigar[nc] (BAM op encoding); every array is zero-padded to a multiple of 8 bytes.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#include <thread>

#include <zlib.h>

namespace {

struct Opts {
    std::string out = ".";
    uint64_t seed = 1;
    int len = 200000;
    double depth = 30;
    int readLen = 150;
    double insMu = 350, insSd = 50;
    double snvRate = 1e-3;
    double indelRate = 1e-4;
    int indelMax = 10;
    int longDels = 0;
    int longDelMin = 50, longDelMax = 300;
    double repeatFrac = 0.0;
    double errRate = 0.003;
    double vafLo = 0.5, vafHi = 0.5;  // tumor: per-fragment probability of the somatic haplotype
    double somaticRate = 0;           // somatic SNVs/indels per bp (haplotype 2 = haplotype 0 + these)
    double normalDepth = 0;           // > 0: also write normal.rvb (germline haplotypes only)
    // record-level shapes for the duplicate filter (-t); their own random stream
    double dupRate = 0;        // each record is followed by an identical copy with this probability
    double singleEndFrac = 0;  // records written unpaired (no mate: mate start 0)
    double unmappedFrac = 0;   // paired records flagged unmapped (0x4), alignment kept
    bool bam = false;          // also write the records as BGZF-compressed BAM (reads.bam, normal.bam)
    bool rvb = true;           // write the read batches (reads.rvb, normal.rvb); --no-rvb: BAM only
    // amplicon panel: reads only inside `panel` regions (written to panel.bed, 1-based inclusive),
    // region lengths log-normal (mu, sigma) clamped to [100, 50000], depth per region uniform in
    // [panelDepthLo, panelDepthHi]; each region is tiled by amplicons whose read pairs all start at
    // the amplicon's ends
    int panel = 0;
    double panelMu = 7.71, panelSigma = 1.21;
    double panelDepthLo = 500, panelDepthHi = 5000;
    int ampLen = 200;
};

struct Variant {
    int pos;          // 0-based reference position of the first affected base
    int kind;         // 0 SNV, 1 insertion (before pos), 2 deletion
    int len;          // deleted length / inserted length
    std::string alt;  // SNV base or inserted sequence
    int hap;
};

const char BASES[4] = {'A', 'C', 'G', 'T'};

char draw_base(std::mt19937_64& g) {
    double u = std::uniform_real_distribution<double>(0, 1)(g);
    if (u < 0.295) return 'A';
    if (u < 0.295 + 0.205) return 'C';
    if (u < 0.295 + 0.41) return 'G';
    return 'T';
}

// haplotype: alt sequence and, for each alt base, its reference position (or -1 when inserted)
struct Hap {
    std::string seq;
    std::vector<int32_t> rpos;
};

// haplotype 2 carries haplotype 0's variants and the somatic ones
Hap build_hap(const std::string& ref, const std::vector<Variant>& vars, int h) {
    Hap H;
    H.seq.reserve(ref.size() + 1024);
    H.rpos.reserve(ref.size() + 1024);
    size_t vi = 0;
    int p = 0;
    int n = (int)ref.size();
    std::vector<const Variant*> mine;
    for (const Variant& v : vars)
        if (v.hap == h || (h == 2 && v.hap == 0)) mine.push_back(&v);
    while (p < n) {
        while (vi < mine.size() && mine[vi]->pos < p) vi++;
        if (vi < mine.size() && mine[vi]->pos == p) {
            const Variant& v = *mine[vi];
            vi++;
            if (v.kind == 0) {
                H.seq.push_back(v.alt[0]);
                H.rpos.push_back(p);
                p++;
                continue;
            }
            if (v.kind == 1) {
                for (char ch : v.alt) {
                    H.seq.push_back(ch);
                    H.rpos.push_back(-1);
                }
                H.seq.push_back(ref[(size_t)p]);
                H.rpos.push_back(p);
                p++;
                continue;
            }
            p += v.len;  // deletion
            continue;
        }
        H.seq.push_back(ref[(size_t)p]);
        H.rpos.push_back(p);
        p++;
    }
    return H;
}

struct Rec {
    int32_t pos0;
    uint16_t flag;
    int32_t nm;
    int32_t mpos0;
    std::string bases;
    std::string quals;
    std::vector<uint32_t> cigar;
};

uint32_t cw(int len, int op) { return ((uint32_t)len << 4) | (uint32_t)op; }

// ---------------- BAM output (SAM/BAM specification: BGZF blocks of at most 64 KiB, each a gzip
// member with a "BC" extra field holding the block size; records after the header)
class BgzfWriter {
  public:
    explicit BgzfWriter(const std::string& path) : f_(path, std::ios::binary) {}
    ~BgzfWriter() { close(); }
    void write(const void* p, size_t n) {
        const uint8_t* c = static_cast<const uint8_t*>(p);
        while (n) {
            const size_t k = std::min(n, kMax - buf_.size());
            buf_.insert(buf_.end(), c, c + k);
            c += k;
            n -= k;
            if (buf_.size() == kMax) flush();
        }
    }
    void close() {
        if (!f_.is_open()) return;
        flush();
        drain();
        static const uint8_t eof[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43,
                                        2, 0, 0x1b, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        f_.write(reinterpret_cast<const char*>(eof), 28);
        f_.close();
    }

  private:
    static constexpr size_t kMax = 65280;
    static constexpr size_t kBatch = 512;  // blocks compressed together, on several threads
    void flush() {
        if (buf_.empty()) return;
        pending_.push_back(std::move(buf_));
        buf_.clear();
        if (pending_.size() == kBatch) drain();
    }
    // one BGZF block (header, raw deflate, CRC32, ISIZE) of `in`
    static std::vector<uint8_t> block(const std::vector<uint8_t>& in) {
        std::vector<uint8_t> z(compressBound((uLong)in.size()) + 64);
        z_stream zs;
        memset(&zs, 0, sizeof zs);
        deflateInit2(&zs, 1, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY);  // fast: test data
        zs.next_in = const_cast<uint8_t*>(in.data());
        zs.avail_in = (uInt)in.size();
        zs.next_out = z.data() + 18;
        zs.avail_out = (uInt)(z.size() - 26);
        deflate(&zs, Z_FINISH);
        const size_t clen = zs.total_out;
        deflateEnd(&zs);
        const uint32_t bsize = (uint32_t)(18 + clen + 8 - 1);
        const uint32_t crc = (uint32_t)crc32(0L, in.data(), (uInt)in.size());
        const uint32_t isize = (uint32_t)in.size();
        const uint8_t h[18] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43, 2, 0,
                               (uint8_t)(bsize & 0xff), (uint8_t)(bsize >> 8)};
        memcpy(z.data(), h, 18);
        memcpy(z.data() + 18 + clen, &crc, 4);
        memcpy(z.data() + 22 + clen, &isize, 4);
        z.resize(26 + clen);
        return z;
    }
    void drain() {
        std::vector<std::vector<uint8_t>> out(pending_.size());
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        for (unsigned t = 0; t < nt; t++)
            pool.emplace_back([&] {
                for (size_t i; (i = next.fetch_add(1)) < pending_.size();) out[i] = block(pending_[i]);
            });
        for (auto& th : pool) th.join();
        for (auto& b : out) f_.write(reinterpret_cast<const char*>(b.data()), (std::streamsize)b.size());
        pending_.clear();
    }
    std::ofstream f_;
    std::vector<uint8_t> buf_;
    std::vector<std::vector<uint8_t>> pending_;
};

// reg2bin (SAM specification, section 5.3)
int reg2bin(int beg, int end) {
    --end;
    if (beg >> 14 == end >> 14) return ((1 << 15) - 1) / 7 + (beg >> 14);
    if (beg >> 17 == end >> 17) return ((1 << 12) - 1) / 7 + (beg >> 17);
    if (beg >> 20 == end >> 20) return ((1 << 9) - 1) / 7 + (beg >> 20);
    if (beg >> 23 == end >> 23) return ((1 << 6) - 1) / 7 + (beg >> 23);
    if (beg >> 26 == end >> 26) return ((1 << 3) - 1) / 7 + (beg >> 26);
    return 0;
}

// Build the record for haplotype interval [a, b): returns false when the read has no aligned base.
bool make_record(const std::string& ref, const Hap& H, int a, int b, bool longDelClip,
                 const std::vector<uint8_t>& longDelMask, Rec& r) {
    // ops per base: 'M' (rpos>=0), 'I' (rpos==-1); deletions appear as jumps in rpos
    std::vector<std::pair<int, int>> ops;  // (op, len)
    auto push = [&](int op, int l) {
        if (l <= 0) return;
        if (!ops.empty() && ops.back().first == op)
            ops.back().second += l;
        else
            ops.push_back({op, l});
    };
    int lastR = -1;
    int first = -1;
    for (int i = a; i < b; i++) {
        int rp = H.rpos[(size_t)i];
        if (rp < 0) {
            push(1, 1);
            continue;
        }
        if (lastR >= 0 && rp > lastR + 1) push(2, rp - lastR - 1);
        push(0, 1);
        if (first < 0) first = rp;
        lastR = rp;
    }
    if (first < 0) return false;
    // leading/trailing I become soft clips; a leading/trailing D cannot occur by construction
    if (!ops.empty() && ops.front().first == 1) ops.front().first = 4;
    if (!ops.empty() && ops.back().first == 1) ops.back().first = 4;
    // long deletions: soft-clip the shorter side of the read at the breakpoint
    if (longDelClip) {
        for (size_t k = 0; k < ops.size(); k++) {
            if (ops[k].first != 2) continue;
            // reference position right after the deletion
            int refAt = first;
            for (size_t t = 0; t < k; t++)
                if (ops[t].first == 0 || ops[t].first == 2) refAt += ops[t].second;
            if (!longDelMask[(size_t)refAt]) continue;
            int leftQ = 0, rightQ = 0;
            for (size_t t = 0; t < k; t++)
                if (ops[t].first != 2) leftQ += ops[t].second;
            for (size_t t = k + 1; t < ops.size(); t++)
                if (ops[t].first != 2) rightQ += ops[t].second;
            std::vector<std::pair<int, int>> n2;
            if (leftQ >= rightQ) {
                for (size_t t = 0; t < k; t++) n2.push_back(ops[t]);
                n2.push_back({4, rightQ});
            } else {
                n2.push_back({4, leftQ});
                for (size_t t = k + 1; t < ops.size(); t++) n2.push_back(ops[t]);
                first = refAt + ops[k].second;
            }
            ops.swap(n2);
            break;
        }
    }
    r.cigar.clear();
    for (auto& o : ops)
        if (o.second > 0) r.cigar.push_back(cw(o.second, o.first));
    r.pos0 = first;
    r.bases = H.seq.substr((size_t)a, (size_t)(b - a));
    return true;
}


int usage() {
    fprintf(stderr,
            "rvsynth --out DIR [--seed N] [--len L] [--depth D] [--read-len R] [--snv-rate X]\n"
            "        [--indel-rate X] [--indel-max N] [--long-dels N] [--repeat-frac F]\n"
            "        [--err-rate X] [--insert-mu M] [--insert-sd S] [--vaf LO HI]\n"
            "        [--somatic-rate X] [--normal-depth D] [--dup-rate X] [--single-end-frac F]\n"
            "        [--unmapped-frac F] [--bam] [--no-rvb] [--panel N] [--panel-len MU SIGMA] [--panel-depth LO HI]\n"
            "        [--amp-len L]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto nx = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value");
            return argv[++i];
        };
        if (a == "--out") o.out = nx();
        else if (a == "--seed") o.seed = std::stoull(nx());
        else if (a == "--len") o.len = std::stoi(nx());
        else if (a == "--depth") o.depth = std::stod(nx());
        else if (a == "--read-len") o.readLen = std::stoi(nx());
        else if (a == "--snv-rate") o.snvRate = std::stod(nx());
        else if (a == "--indel-rate") o.indelRate = std::stod(nx());
        else if (a == "--indel-max") o.indelMax = std::stoi(nx());
        else if (a == "--long-dels") o.longDels = std::stoi(nx());
        else if (a == "--repeat-frac") o.repeatFrac = std::stod(nx());
        else if (a == "--err-rate") o.errRate = std::stod(nx());
        else if (a == "--insert-mu") o.insMu = std::stod(nx());
        else if (a == "--insert-sd") o.insSd = std::stod(nx());
        else if (a == "--vaf") {
            o.vafLo = std::stod(nx());
            o.vafHi = std::stod(nx());
        } else if (a == "--somatic-rate") o.somaticRate = std::stod(nx());
        else if (a == "--normal-depth") o.normalDepth = std::stod(nx());
        else if (a == "--dup-rate") o.dupRate = std::stod(nx());
        else if (a == "--single-end-frac") o.singleEndFrac = std::stod(nx());
        else if (a == "--unmapped-frac") o.unmappedFrac = std::stod(nx());
        else if (a == "--bam") o.bam = true;
        else if (a == "--panel") o.panel = std::stoi(nx());
        else if (a == "--no-rvb") o.rvb = false;
        else if (a == "--panel-len") {
            o.panelMu = std::stod(nx());
            o.panelSigma = std::stod(nx());
        } else if (a == "--panel-depth") {
            o.panelDepthLo = std::stod(nx());
            o.panelDepthHi = std::stod(nx());
        } else if (a == "--amp-len") o.ampLen = std::stoi(nx());
        else
            return usage();
    }
    std::mt19937_64 gref(o.seed * 1000003ULL + 1), gvar(o.seed * 1000003ULL + 2), gread(o.seed * 1000003ULL + 3),
        gerr(o.seed * 1000003ULL + 4);
    // ---------------- reference
    std::string ref((size_t)o.len, 'A');
    for (int i = 0; i < o.len; i++) ref[(size_t)i] = draw_base(gref);
    if (o.repeatFrac > 0) {
        long want = (long)(o.repeatFrac * o.len), have = 0;
        std::uniform_int_distribution<int> up(200, o.len - 400), ul(2, 6), tl(12, 60);
        while (have < want) {
            int s = up(gref), u = ul(gref), L = tl(gref);
            std::string unit;
            for (int k = 0; k < u; k++) unit += BASES[gref() % 4];
            for (int k = 0; k < L; k++) ref[(size_t)(s + k)] = unit[(size_t)(k % u)];
            have += L;
        }
    }
    // ---------------- variants
    std::vector<Variant> vars;
    std::vector<uint8_t> busy((size_t)o.len, 0);
    std::vector<uint8_t> longDelMask((size_t)o.len + 1, 0);
    auto mark = [&](int s, int e) {
        for (int k = std::max(0, s); k < std::min(o.len, e); k++) busy[(size_t)k] = 1;
    };
    auto free_span = [&](int s, int e) {
        if (s < 150 || e > o.len - 150) return false;
        for (int k = s; k < e; k++)
            if (busy[(size_t)k]) return false;
        return true;
    };
    std::uniform_real_distribution<double> U(0, 1);
    for (int i = 0; i < o.longDels; i++) {
        int L = std::uniform_int_distribution<int>(o.longDelMin, o.longDelMax)(gvar);
        int p = std::uniform_int_distribution<int>(400, o.len - 400 - L)(gvar);
        if (!free_span(p - 20, p + L + 20)) continue;
        vars.push_back({p, 2, L, "", (int)(gvar() & 1)});
        mark(p - 20, p + L + 20);
        longDelMask[(size_t)p] = 1;
    }
    for (int p = 150; p < o.len - 150; p++) {
        double u = U(gvar);
        if (u < o.snvRate) {
            if (!free_span(p - 3, p + 4)) continue;
            char rb = ref[(size_t)p];
            char alt;
            do {
                alt = BASES[gvar() % 4];
            } while (alt == rb);
            vars.push_back({p, 0, 1, std::string(1, alt), (int)(gvar() & 1)});
            mark(p - 3, p + 4);
        } else if (u < o.snvRate + o.indelRate) {
            int L = std::uniform_int_distribution<int>(1, o.indelMax)(gvar);
            bool ins = (gvar() & 1) != 0;
            if (ins) {
                std::string s;
                for (int k = 0; k < L; k++) s += BASES[gvar() % 4];
                int q = p;
                // left-normalise: move the insertion left while the preceding base equals its last base
                while (q > 150 && ref[(size_t)q - 1] == s.back()) {
                    s = s.back() + s.substr(0, s.size() - 1);
                    q--;
                }
                if (!free_span(q - 5, q + 5)) continue;
                vars.push_back({q, 1, L, s, (int)(gvar() & 1)});
                mark(q - 5, q + 5);
            } else {
                int q = p;
                while (q > 150 && ref[(size_t)q - 1] == ref[(size_t)(q + L - 1)]) q--;
                if (!free_span(q - 5, q + L + 5)) continue;
                vars.push_back({q, 2, L, "", (int)(gvar() & 1)});
                mark(q - 5, q + L + 5);
            }
        }
    }
    if (o.somaticRate > 0) {
        // somatic variants, from their own stream so that the germline set does not change
        std::mt19937_64 gsom(o.seed * 1000003ULL + 5);
        for (int p = 150; p < o.len - 150; p++) {
            double u = U(gsom);
            if (u >= o.somaticRate) continue;
            if (u < o.somaticRate * 5 / 6) {
                if (!free_span(p - 3, p + 4)) continue;
                char rb = ref[(size_t)p], alt;
                do {
                    alt = BASES[gsom() % 4];
                } while (alt == rb);
                vars.push_back({p, 0, 1, std::string(1, alt), 2});
                mark(p - 3, p + 4);
            } else {
                int L = std::uniform_int_distribution<int>(1, std::max(1, o.indelMax))(gsom);
                if (!free_span(p - 5, p + L + 5)) continue;
                vars.push_back({p, 2, L, "", 2});
                mark(p - 5, p + L + 5);
            }
        }
    }
    std::sort(vars.begin(), vars.end(), [](const Variant& a, const Variant& b) { return a.pos < b.pos; });
    Hap H[3] = {build_hap(ref, vars, 0), build_hap(ref, vars, 1), Hap{}};
    if (o.somaticRate > 0) H[2] = build_hap(ref, vars, 2);
    // ---------------- reads
    // tumor (or the only sample): haplotypes 0/1, and haplotype 2 with probability vaf when there
    // are somatic variants; normal: haplotypes 0/1 from its own streams
    // panel regions (1-based inclusive) from their own stream; hap index of every reference
    // position (the next kept base for a deleted one)
    std::vector<std::pair<int, int>> panel_regions;
    std::vector<std::vector<int32_t>> ref2hap(3);
    if (o.panel > 0) {
        std::mt19937_64 gpan(o.seed * 1000003ULL + 9);
        std::normal_distribution<double> ln(o.panelMu, o.panelSigma);
        std::vector<std::pair<int, int>> cand;
        for (int k = 0; k < o.panel; k++) {
            int L = (int)std::lround(std::exp(ln(gpan)));
            L = std::max(100, std::min(50000, std::min(L, o.len - 2)));
            const int s0 = 1 + (int)(gpan() % (uint64_t)(o.len - L));
            cand.push_back({s0, s0 + L - 1});
        }
        std::sort(cand.begin(), cand.end());
        for (auto& r : cand)  // disjoint regions, in order
            if (panel_regions.empty() || r.first > panel_regions.back().second + 300) panel_regions.push_back(r);
        for (int h = 0; h < 3; h++) {
            if (H[h].seq.empty()) continue;
            std::vector<int32_t>& m = ref2hap[(size_t)h];
            m.assign((size_t)o.len + 1, (int32_t)H[h].seq.size());
            for (size_t i = 0; i < H[h].rpos.size(); i++)
                if (H[h].rpos[i] >= 0 && m[(size_t)H[h].rpos[i]] == (int32_t)H[h].seq.size()) m[(size_t)H[h].rpos[i]] = (int32_t)i;
            for (int p = o.len - 1; p >= 0; p--) m[(size_t)p] = std::min(m[(size_t)p], m[(size_t)p + 1]);
        }
        std::ofstream bed(o.out + "/panel.bed");
        int k = 0;
        for (auto& r : panel_regions) bed << "chrS\t" << r.first << "\t" << r.second << "\tamp" << k++ << "\n";
    }
    // Reads are made in chunks (a range of fragments, or one panel region) on several threads; every
    // chunk draws from its own two streams seeded from (stream seed, chunk index), so the output
    // depends on the seed only, not on the thread count.
    auto gen_reads = [&](double depth, bool somatic, uint64_t stream_seed) {
    long npairs = (long)(depth * o.len / (2.0 * o.readLen));
    const long CHUNK = 1 << 15;
    const size_t nchunks = o.panel > 0 ? panel_regions.size() : (size_t)((npairs + CHUNK - 1) / CHUNK);
    std::vector<std::vector<Rec>> parts(nchunks);
    auto make_chunk = [&](size_t ci) {
    std::mt19937_64 gread(stream_seed * 0x9E3779B97F4A7C15ULL + 2 * ci + 1), gerr(stream_seed * 0x9E3779B97F4A7C15ULL + 2 * ci + 2);
    std::vector<Rec>& recs = parts[ci];
    std::normal_distribution<double> ins(o.insMu, o.insSd), qn(35, 5);
    // one read pair of the fragment [a, a + isz) of haplotype hp
    auto emit_pair = [&](const Hap& hp, int a, int isz) {
        Rec r1, r2;
        bool ok1 = make_record(ref, hp, a, a + o.readLen, o.longDels > 0, longDelMask, r1);
        bool ok2 = make_record(ref, hp, a + isz - o.readLen, a + isz, o.longDels > 0, longDelMask, r2);
        if (!ok1 || !ok2) return;
        bool firstIsFwd = (gread() & 1) != 0;
        for (int k = 0; k < 2; k++) {
            Rec& r = k == 0 ? r1 : r2;
            Rec& m = k == 0 ? r2 : r1;
            bool rev = k == 1;
            uint16_t fl = 0x1 | 0x2 | (rev ? 0x10 : 0) | (!rev ? 0x20 : 0);
            fl |= ((k == 0) == firstIsFwd) ? 0x40 : 0x80;
            r.flag = fl;
            r.mpos0 = m.pos0;
            // qualities and errors
            r.quals.resize(r.bases.size());
            for (size_t i = 0; i < r.bases.size(); i++) {
                int q = (int)std::lround(qn(gerr));
                q = std::max(2, std::min(41, q));
                r.quals[i] = (char)q;
                if (U(gerr) < o.errRate) {
                    char b = r.bases[i], nb2;
                    do {
                        nb2 = BASES[gerr() % 4];
                    } while (nb2 == b);
                    r.bases[i] = nb2;
                }
            }
            // NM over aligned bases
            int nmv = 0, rp = r.pos0, qp = 0;
            for (uint32_t w : r.cigar) {
                int op = (int)(w & 0xf), l = (int)(w >> 4);
                if (op == 0) {
                    for (int t = 0; t < l; t++)
                        if (r.bases[(size_t)(qp + t)] != ref[(size_t)(rp + t)]) nmv++;
                    rp += l;
                    qp += l;
                } else if (op == 1) {
                    nmv += l;
                    qp += l;
                } else if (op == 2) {
                    nmv += l;
                    rp += l;
                } else if (op == 4) {
                    qp += l;
                }
            }
            r.nm = nmv;
        }
        recs.push_back(std::move(r1));
        recs.push_back(std::move(r2));
    };
    if (o.panel > 0) {
        // amplicons: fixed start and end per tile, every pair of the tile on one of them
        {
            const auto& reg = panel_regions[ci];
            const double depth = o.panelDepthLo + (o.panelDepthHi - o.panelDepthLo) * U(gread);
            for (int t = reg.first - 1; t < reg.second; t += o.ampLen * 3 / 4) {
                const int t_end = std::min(t + o.ampLen, o.len);
                if (t_end - t < o.readLen) break;
                const long pairs = std::max(1L, (long)(depth * (t_end - t) / (2.0 * o.readLen)));
                for (long f = 0; f < pairs; f++) {
                    int h = (gread() & 1) ? 1 : 0;
                    double vaf = o.vafLo + (o.vafHi - o.vafLo) * U(gread);
                    if (somatic && U(gread) < vaf) h = 2;
                    const Hap& hp = H[h];
                    const int a = ref2hap[h][(size_t)t], b = ref2hap[h][(size_t)t_end];
                    if (b - a >= o.readLen) emit_pair(hp, a, b - a);
                }
            }
        }
    } else {
        for (long f = (long)ci * CHUNK; f < std::min(npairs, (long)(ci + 1) * CHUNK); f++) {
            double vaf = o.vafLo + (o.vafHi - o.vafLo) * U(gread);
            int h = (gread() & 1) ? 1 : 0;
            if (somatic && U(gread) < vaf) h = 2;
            const Hap& hp = H[h];
            int hl = (int)hp.seq.size();
            int isz = (int)std::lround(ins(gread));
            if (isz < o.readLen) isz = o.readLen;
            if (isz > hl) isz = hl;
            int a = std::uniform_int_distribution<int>(0, hl - isz)(gread);
            emit_pair(hp, a, isz);
        }
    }
    };
    {
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        for (unsigned t = 0; t < nt; t++)
            pool.emplace_back([&] {
                for (size_t ci; (ci = next.fetch_add(1)) < nchunks;) make_chunk(ci);
            });
        for (auto& th : pool) th.join();
    }
    if (getenv("RV_SYNTH_TIMING")) fprintf(stderr, "[rvsynth] chunks done\n");
    // records sorted by start, ties in chunk order (a stable sort of the concatenation)
    std::vector<std::pair<int32_t, uint32_t>> key;
    std::vector<std::pair<uint32_t, uint32_t>> at;  // (chunk, index in chunk)
    for (size_t ci = 0; ci < nchunks; ci++)
        for (size_t i = 0; i < parts[ci].size(); i++) {
            key.push_back({parts[ci][i].pos0, (uint32_t)at.size()});
            at.push_back({(uint32_t)ci, (uint32_t)i});
        }
    std::sort(key.begin(), key.end());
    std::vector<Rec> recs;
    recs.reserve(key.size());
    for (const auto& k : key) recs.push_back(std::move(parts[at[k.second].first][at[k.second].second]));
    parts.clear();
    if (o.dupRate > 0 || o.singleEndFrac > 0 || o.unmappedFrac > 0) {
        std::mt19937_64 gdup(stream_seed * 0x9E3779B97F4A7C15ULL + 0x5bd1e995ULL);
        std::vector<Rec> shaped;
        shaped.reserve(recs.size() + recs.size() / 4);
        for (Rec& r : recs) {
            if (U(gdup) < o.singleEndFrac) {
                r.flag &= (uint16_t)~(0x1 | 0x2 | 0x20 | 0x40 | 0x80);
                r.mpos0 = -1;
            } else if (U(gdup) < o.unmappedFrac) {
                r.flag |= 0x4;
            }
            const bool copy = U(gdup) < o.dupRate;
            if (copy) shaped.push_back(r);  // the record and its identical copy
            shaped.push_back(std::move(r));
        }
        recs.swap(shaped);
    }
    return recs;
    };
    auto T0 = std::chrono::steady_clock::now();
    auto tlog = [&](const char* what) {
        if (getenv("RV_SYNTH_TIMING"))
            fprintf(stderr, "[rvsynth] %s %.2f s\n", what, std::chrono::duration<double>(std::chrono::steady_clock::now() - T0).count());
    };
    std::vector<Rec> recs = gen_reads(o.depth, o.somaticRate > 0, o.seed * 1000003ULL + 3);
    tlog("tumor reads");

    // ---------------- write
    {
        std::ofstream fa(o.out + "/ref.fa");
        fa << ">chrS\n";
        for (int i = 0; i < o.len; i += 60) fa << ref.substr((size_t)i, 60) << "\n";
    }
    {
        std::ofstream tv(o.out + "/truth.tsv");
        tv << "pos1\tkind\tlen\talt\thap\n";
        static const char* kn[3] = {"SNV", "INS", "DEL"};
        for (const Variant& v : vars) tv << v.pos + 1 << "\t" << kn[v.kind] << "\t" << v.len << "\t" << v.alt << "\t" << v.hap << "\n";
    }
    auto write_rvb = [&](const std::string& path, const std::vector<Rec>& recs) {
        // streamed field by field through one buffer (the layout of the file header comment)
        int64_t n = (int64_t)recs.size(), nb = 0, nc = 0;
        for (const Rec& r : recs) {
            nb += (int64_t)r.bases.size();
            nc += (int64_t)r.cigar.size();
        }
        std::ofstream f(path, std::ios::binary);
        f.write("RVB1", 4);
        uint32_t ver = 1;
        f.write(reinterpret_cast<const char*>(&ver), 4);
        f.write(reinterpret_cast<const char*>(&n), 8);
        f.write(reinterpret_cast<const char*>(&nb), 8);
        f.write(reinterpret_cast<const char*>(&nc), 8);
        std::vector<char> buf;
        buf.reserve(1 << 22);
        size_t field_bytes = 0;
        auto emit = [&](const void* p, size_t k) {
            const char* c = static_cast<const char*>(p);
            buf.insert(buf.end(), c, c + k);
            field_bytes += k;
            if (buf.size() >= (1 << 22)) {
                f.write(buf.data(), (std::streamsize)buf.size());
                buf.clear();
            }
        };
        auto end_field = [&]() {
            static const char z[8] = {0};
            const size_t pad = (8 - field_bytes % 8) % 8;
            if (pad) emit(z, pad);
            field_bytes = 0;
        };
        for (const Rec& r : recs) emit(&r.pos0, 4);
        end_field();
        for (const Rec& r : recs) emit(&r.flag, 2);
        end_field();
        const uint8_t mq = 60, one = 1;
        for (int64_t i = 0; i < n; i++) emit(&mq, 1);
        end_field();
        for (int64_t i = 0; i < n; i++) emit(&one, 1);
        end_field();
        for (const Rec& r : recs) emit(&r.nm, 4);
        end_field();
        for (const Rec& r : recs) emit(&r.mpos0, 4);
        end_field();
        const int32_t zero = 0;
        for (int64_t i = 0; i < n; i++) emit(&zero, 4);  // tid
        end_field();
        for (int64_t i = 0; i < n; i++) emit(&zero, 4);  // mtid
        end_field();
        int64_t off = 0;
        for (const Rec& r : recs) {
            emit(&off, 8);
            off += (int64_t)r.bases.size();
        }
        emit(&off, 8);
        end_field();
        off = 0;
        for (const Rec& r : recs) {
            emit(&off, 8);
            off += (int64_t)r.cigar.size();
        }
        emit(&off, 8);
        end_field();
        for (const Rec& r : recs) emit(r.bases.data(), r.bases.size());
        end_field();
        for (const Rec& r : recs) emit(r.quals.data(), r.quals.size());
        end_field();
        for (const Rec& r : recs) emit(r.cigar.data(), 4 * r.cigar.size());
        end_field();
        f.write(buf.data(), (std::streamsize)buf.size());
        return n;
    };
    auto write_bam = [&](const std::string& path, const std::vector<Rec>& recs) {
        BgzfWriter w(path);
        const std::string text = "@HD\tVN:1.6\tSO:coordinate\n@SQ\tSN:chrS\tLN:" + std::to_string(o.len) + "\n";
        auto u32 = [&](uint32_t v) { w.write(&v, 4); };
        w.write("BAM\1", 4);
        u32((uint32_t)text.size());
        w.write(text.data(), text.size());
        u32(1);
        u32(5);
        w.write("chrS", 5);
        u32((uint32_t)o.len);
        std::vector<uint8_t> r;
        for (size_t i = 0; i < recs.size(); i++) {
            const Rec& x = recs[i];
            const std::string name = "r" + std::to_string(i);
            int rlen = 0;
            for (uint32_t c : x.cigar) {
                const int op = (int)(c & 15);
                if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rlen += (int)(c >> 4);
            }
            const int l_seq = (int)x.bases.size();
            r.clear();
            auto p32 = [&](uint32_t v) { for (int k = 0; k < 4; k++) r.push_back((uint8_t)(v >> (8 * k))); };
            auto p16 = [&](uint16_t v) { r.push_back((uint8_t)v); r.push_back((uint8_t)(v >> 8)); };
            p32(0);                    // refID
            p32((uint32_t)x.pos0);
            r.push_back((uint8_t)(name.size() + 1));
            r.push_back(60);           // MAPQ
            p16((uint16_t)reg2bin(x.pos0, x.pos0 + (rlen > 0 ? rlen : 1)));
            p16((uint16_t)x.cigar.size());
            p16(x.flag);
            p32((uint32_t)l_seq);
            p32(0);                    // next refID: the same contig (as the RVB records)
            p32((uint32_t)x.mpos0);
            p32(0);                    // TLEN
            r.insert(r.end(), name.begin(), name.end());
            r.push_back(0);
            for (uint32_t c : x.cigar) p32(c);
            static const char* codes = "=ACMGRSVTWYHKDBN";
            for (int q = 0; q < l_seq; q += 2) {
                auto code = [&](int k) -> uint8_t {
                    if (k >= l_seq) return 0;
                    const char* h = strchr(codes, x.bases[(size_t)k]);
                    return h ? (uint8_t)(h - codes) : 15;
                };
                r.push_back((uint8_t)(code(q) << 4 | code(q + 1)));
            }
            for (int q = 0; q < l_seq; q++) r.push_back((uint8_t)x.quals[(size_t)q]);
            r.push_back('N');
            r.push_back('M');
            r.push_back('i');
            p32((uint32_t)x.nm);
            u32((uint32_t)r.size());
            w.write(r.data(), r.size());
        }
        w.close();
    };
    const int64_t n = o.rvb ? write_rvb(o.out + "/reads.rvb", recs) : (int64_t)recs.size();
    tlog("tumor written");
    if (o.bam) write_bam(o.out + "/reads.bam", recs);
    int64_t nn = 0;
    if (o.normalDepth > 0) {
        recs.clear();
        recs.shrink_to_fit();
        const std::vector<Rec> nrecs = gen_reads(o.normalDepth, false, o.seed * 1000003ULL + 6);
        tlog("normal reads");
        nn = o.rvb ? write_rvb(o.out + "/normal.rvb", nrecs) : (int64_t)nrecs.size();
        tlog("normal written");
        if (o.bam) write_bam(o.out + "/normal.bam", nrecs);
    }
    printf("{\"reads\": %lld, \"normal_reads\": %lld, \"variants\": %zu, \"len\": %d}\n", (long long)n, (long long)nn,
           vars.size(), o.len);
    return 0;
}
