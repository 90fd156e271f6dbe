#include <cmath>
#include <cstdlib>
// Test harness only: runs the engine's device stages sequentially on the CPU (prep, walk, anchor,
// then the host reduction and realigner, then one score_position call per position exactly as the
// two scoring kernels map threads to positions) and prints the simple-mode lines.
//
//   rv_cpu_sim REF.fa READS.rvb START END K FISHER OUT      whole path for chrS:START-END
//   rv_cpu_sim --rh-chars B1 B2 ...                         robin_hood <char,int> order of bytes
//   rv_cpu_sim --rh-ints < OPS                              robin_hood <int,V> order after "i K" /
//                                                           "e K" operations (one per line)
#include <algorithm>
#include <chrono>
#include <climits>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "hip/hip_runtime.h"
#include "rv_format.h"
#include "rv_host.h"
#include "rv_pileup_body.h"
#include "rv_order.h"
#include "rv_realign.h"
#include "rv_rhorder.h"
#include "rv_score_body.h"
#include "rv_desc.h"
#include <regex>

using namespace rv;

template <class T>
static void rd(std::ifstream& f, std::vector<T>& v, size_t n) {
    v.resize(n);
    f.read((char*)v.data(), (std::streamsize)(sizeof(T) * n));
    size_t pad = (8 - (sizeof(T) * n) % 8) % 8;
    char t[8];
    if (pad) f.read(t, (std::streamsize)pad);
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "--rh-chars") {
        std::vector<char> ins;
        for (int i = 2; i < argc; i++) ins.push_back((char)atoi(argv[i]));
        for (char c : rh_char_iteration_order(ins)) printf("%d ", (int)(unsigned char)c);
        printf("\n");
        return 0;
    }
    if (argc == 3 && std::string(argv[1]) == "--find35-file") {
        // the realigner's find35match on the pairs of a file ("seq5 seq3" per line): "b5 b3 score"
        std::ifstream f(argv[2]);
        std::string a, b;
        while (f >> a >> b) {
            int b5, b3, sc;
            find35match(a, b, b5, b3, sc);
            printf("%d %d %d\n", b5, b3, sc);
        }
        return 0;
    }
    if (argc == 4 && std::string(argv[1]) == "--desc-fuzz") {
        // the realigner's descriptor matchers (rv_desc.h) against std::regex with the reference's
        // pattern texts, on random strings over the descriptor alphabet and on descriptor shapes.
        // Prints "cases mismatches matches" (matches: cases where some pattern matched).
        const int trials = atoi(argv[2]);
        unsigned seed = (unsigned)atoi(argv[3]);
        auto rnd = [&]() {
            seed = seed * 1664525u + 1013904223u;
            return seed >> 8;
        };
        const std::string alpha = "+-^&#ACGTN0123456789<>dup";
        const char* shapes[] = {"-", "+", "-12", "-3&", "+AC", "#", "^", "&", "^5", "#AG^", "ACGT", "-2^AT", "<dup3>"};
        const std::regex r_bmn{"^-(\\d+).*"}, r_une{"\\^(\\d+)$"}, r_bmna{"^-\\d+(.*)"}, r_ca{"\\^([ATGNC]+)"},
            r_strip{"\\^|&|#"}, r_mnae{"(-\\d+)&[ATGC]+$"}, r_bpa{"^\\+([ATGC]+).*"}, r_amp{".*&([ATGC]+).*"},
            r_hash{"#([ATGC]+).*"}, r_cae{"\\^([ATGC]+)$"}, r_paae{"(\\+[ATGC]+)&[ATGC]+$"}, r_lp{"^\\+"},
            r_cde{"\\^\\d+$"}, r_c{"\\^"}, r_sv{"^-\\d+\\^([ATGNC]+)<...\\d+>([ATGNC]+)$"},
            r_dup{"<dup(\\d+)>([ATGC]+)$"};
        long bad = 0, hits = 0;
        for (int t = 0; t < trials; t++) {
            std::string x;
            const int parts = 1 + (int)(rnd() % 6);
            for (int k = 0; k < parts; k++) {
                if (rnd() % 2) x += shapes[rnd() % (sizeof shapes / sizeof *shapes)];
                const int len = (int)(rnd() % 5);
                for (int q = 0; q < len; q++) x += alpha[rnd() % alpha.size()];
            }
            std::smatch m;
            std::string a, b;
            auto chk = [&](bool got, const std::string& gv, const std::regex& re, int grp) {
                const bool want = std::regex_search(x, m, re);
                hits += want;
                if (got != want || (want && gv != m.str(grp))) {
                    if (bad < 10) fprintf(stderr, "mismatch on '%s'\n", x.c_str());
                    bad++;
                }
            };
            chk(desc::begin_minus_number(x, a), a, r_bmn, 1);
            chk(desc::up_number_end(x, a), a, r_une, 1);
            chk(desc::begin_minus_number_any(x, a), a, r_bmna, 1);
            chk(desc::caret_atgnc(x, a), a, r_ca, 1);
            chk(desc::minus_number_amp_atgcs_end(x, a), a, r_mnae, 1);
            chk(desc::begin_plus_atgc(x, a), a, r_bpa, 1);
            chk(desc::amp_atgc(x, a), a, r_amp, 1);
            chk(desc::hash_atgc(x, a), a, r_hash, 1);
            chk(desc::caret_atgc_end(x, a), a, r_cae, 1);
            chk(desc::plus_atgcs_amp_atgcs_end(x, a), a, r_paae, 1);
            {
                std::string c, d;
                const bool g = desc::minus_sv(x, c, d), w = std::regex_search(x, m, r_sv);
                if (g != w || (w && (c != m.str(1) || d != m.str(2)))) bad++;
                const bool g2 = desc::dup_num_atgc(x, c, d), w2 = std::regex_search(x, m, r_dup);
                if (g2 != w2 || (w2 && (c != m.str(1) || d != m.str(2)))) bad++;
            }
            auto rep = [&](const std::regex& re) { return std::regex_replace(x, re, "", std::regex_constants::format_first_only); };
            if (desc::strip_marks3(x) != std::regex_replace(x, r_strip, "")) bad++;
            a = x;
            desc::drop_lead_plus(a);
            if (a != rep(r_lp)) bad++;
            a = x;
            desc::drop_caret_digits_end(a);
            if (a != rep(r_cde)) bad++;
            a = x;
            desc::drop_first_caret(a);
            if (a != rep(r_c)) bad++;
        }
        printf("%d %ld %ld\n", trials, bad, hits);
        return 0;
    }
    if (argc == 5 && std::string(argv[1]) == "--findmsi") {
        // the scoring kernels' find_msi: "msi shift3 msint" (arguments as rvcpu --findmsi)
        auto arg = [&](int i) { return std::string(argv[i]) == "-" ? std::string() : std::string(argv[i]); };
        std::string t1 = arg(2), t2 = arg(3), left = arg(4);
        MSIRes m = find_msi(t1.c_str(), (int)t1.size(), t2.c_str(), (int)t2.size(), left.c_str(), (int)left.size());
        printf("%.17g %d %s\n", m.msi, m.shift3, t1.substr(t1.size() - (size_t)m.msint_len).c_str());
        return 0;
    }
    if (argc >= 2 && std::string(argv[1]) == "--rh-ints") {
        RHOrder m;
        char op[8];
        long long key;
        while (scanf("%7s %lld", op, &key) == 2) {
            if (op[0] == 'i') m.insert((int32_t)key);
            if (op[0] == 'e') m.erase((int32_t)key);
        }
        m.for_each([](int32_t k) { printf("%d\n", k); });
        return 0;
    }
    if (argc >= 4 && std::string(argv[1]) == "--find35-check") {
        // find35match against a direct restatement of the reference's nested loops, on random
        // clip consensus pairs (half of them built to overlap).  Prints "pairs matched equal".
        const int trials = atoi(argv[2]);
        unsigned seed = (unsigned)atoi(argv[3]);
        auto rnd = [&]() {
            seed = seed * 1664525u + 1013904223u;
            return seed >> 8;
        };
        auto brute = [](const std::string& seq5, const std::string& seq3, int& b5, int& b3, int& score) {
            b5 = b3 = score = 0;
            for (size_t i = 0; i < seq5.size() - 8; i++) {
                for (size_t j = 1; j < seq3.size() - 8; j++) {
                    int nm = 0, n = 0;
                    const int l3 = (int)seq3.size(), l5 = (int)seq5.size();
                    while (n + (int)j <= l3 && (int)i + n <= l5) {
                        if ((int)i + n >= l5 || seq3[(size_t)(l3 - (int)j - n)] != seq5[(size_t)((int)i + n)]) nm++;
                        if (nm > 2) break;
                        n++;
                    }
                    if (n - nm > score && n - nm > 8 && nm / (double)n < 0.1 &&
                        (n + (int)j >= (int)seq3.size() || (int)i + n >= (int)seq5.size())) {
                        score = n - nm;
                        b3 = (int)j;
                        b5 = (int)i;
                        return;
                    }
                }
            }
        };
        const char* AL = "ACGT";
        int matched = 0, equal = 0;
        for (int t = 0; t < trials; t++) {
            const int l5 = 11 + (int)(rnd() % 150), l3 = 11 + (int)(rnd() % 150);
            std::string s5, s3;
            for (int k = 0; k < l5; k++) s5 += AL[rnd() % (t % 7 == 0 ? 2 : 4)];
            if (rnd() % 2) {
                // s3 = reversed stretch of s5 (optionally shifted / mutated), padded to l3
                const int from = (int)(rnd() % (unsigned)l5);
                for (int k = from; k >= 0 && (int)s3.size() < l3; k--) s3 += s5[(size_t)k];
                std::reverse(s3.begin(), s3.end());
                while ((int)s3.size() < l3) s3.insert(s3.begin(), AL[rnd() % 4]);
                const int muts = (int)(rnd() % 4);
                for (int k = 0; k < muts; k++) s3[rnd() % s3.size()] = AL[rnd() % 4];
            } else {
                for (int k = 0; k < l3; k++) s3 += AL[rnd() % (t % 7 == 0 ? 2 : 4)];
            }
            int a5, a3, as, c5, c3, cs;
            brute(s5, s3, a5, a3, as);
            find35match(s5, s3, c5, c3, cs);
            if (as != 0) matched++;
            if (a5 == c5 && a3 == c3 && as == cs) equal++;
            else fprintf(stderr, "differs: %s %s: (%d %d %d) vs (%d %d %d)\n", s5.c_str(), s3.c_str(), a5, a3, as, c5, c3, cs);
        }
        printf("%d %d %d\n", trials, matched, equal);
        return 0;
    }
    if (argc >= 4 && std::string(argv[1]) == "--fmt-check") {
        // put_f against snprintf("%f") on random doubles of every magnitude, exact ties (k / 2^j),
        // signed zeros, subnormals and values just around the rounding points; prints mismatches
        const long n = atol(argv[2]);
        unsigned long long s = (unsigned long long)atoll(argv[3]) | 1ull;
        auto rnd = [&]() {
            s ^= s << 13;
            s ^= s >> 7;
            s ^= s << 17;
            return s;
        };
        long bad = 0;
        auto check = [&](double v) {
            std::string a;
            put_f(a, v);
            char b[512];
            const int k = snprintf(b, sizeof b, "%f", v);
            if (a != std::string(b, (size_t)k)) {
                if (bad < 5) fprintf(stderr, "mismatch %.17g: %s vs %s\n", v, a.c_str(), b);
                bad++;
            }
        };
        for (double v : {0.0, -0.0, 1.0, -1.0, 0.5, 0.0000005, 0.0000015, 0.0000025, 4.9406564584124654e-324,
                         2.2250738585072014e-308, 8796093022207.0, 8796093022208.0, 1e15, -1e-9})
            check(v);
        for (long i = 0; i < n; i++) {
            const uint64_t r = rnd();
            double v;
            switch (i % 4) {
                case 0: {  // any finite bit pattern
                    uint64_t b = r;
                    std::memcpy(&v, &b, 8);
                    if (!std::isfinite(v)) continue;
                    break;
                }
                case 1:  // k / 2^j: exact ties at the sixth decimal
                    v = (double)(r % 100000000) / (double)(1ull << (1 + rnd() % 40));
                    break;
                case 2:  // counts over counts, as the printed fields are
                    v = (double)(r % 100000) / (double)(1 + rnd() % 5000);
                    break;
                default:  // next to a printed rounding point
                    v = std::nextafter((double)(r % 1000000000) / 1e6 + 5e-7, (rnd() & 1) ? 1e300 : -1e300);
                    break;
            }
            if (rnd() & 1) v = -v;
            check(v);
        }
        printf("%ld\n", bad);
        return 0;
    }
    if (argc >= 4 && std::string(argv[1]) == "--aligned-check") {
        // insert / insert-then-erase sequences on distinct keys: the per-growth sort, when it
        // applies, must equal the replay.  Keys are positions in a hashed order (as
        // alignedVariants receives them), some drawn from a few home ranges so that runs grow
        // long and the info increment halves up to the last level.  Prints "applied equal
        // halved bailed" counts (halved: the replayed table ended with a halved increment).
        const int n = atoi(argv[2]);
        unsigned seed = (unsigned)atoi(argv[3]);
        auto rnd = [&]() {
            seed = seed * 1664525u + 1013904223u;
            return seed >> 8;
        };
        int applied = 0, equal = 0, halved = 0, bailed = 0;
        uint32_t min_inc = 32;  // smallest final increment among the sequences the sort applied to
        for (int trial = 0; trial < 12; trial++) {
            std::vector<int32_t> pos;
            for (int i = 0; i < n; i++) pos.push_back(i + 1);
            std::vector<int32_t> proc = noninsertion_order(pos, {});
            const int keep_pct = 5 + (int)(rnd() % 80), tr_pct = (int)(rnd() % 25);
            // crowd: keep only keys whose home at 2^b slots falls in a narrow band for a share of them
            const int crowd_pct = trial % 3 == 0 ? 0 : (int)(rnd() % 60);
            const int bits = 10 + (int)(rnd() % 8);
            const uint64_t band = (1ull << bits) / (8 + rnd() % 24);
            std::vector<int32_t> ops;
            std::vector<uint8_t> tr;
            for (int32_t k : proc) {
                const unsigned u = rnd() % 100;
                const uint64_t x = (uint64_t)(int64_t)k;
                const unsigned __int128 pr = (unsigned __int128)x * (unsigned __int128)0xde5fb9d2630458e9ULL;
                const uint64_t h = (uint64_t)(pr >> 64) + (uint64_t)pr;
                const bool crowded = (h & ((1ull << bits) - 1)) < band;
                if (crowd_pct && !crowded && (int)(rnd() % 100) < crowd_pct) continue;
                if ((int)u < keep_pct) {
                    ops.push_back(k);
                    tr.push_back(0);
                } else if ((int)u < keep_pct + tr_pct) {
                    ops.push_back(k);
                    tr.push_back(1);
                }
            }
            RHOrder m;
            std::vector<int32_t> want, got;
            try {
                for (size_t i = 0; i < ops.size(); i++) {
                    m.insert(ops[i]);
                    if (tr[i]) m.erase(ops[i]);
                }
            } catch (const std::overflow_error&) {
                // robin_hood itself gives up on such a sequence: the sort must not claim it
                try {
                    if (sorted_ops_order(ops.data(), tr.data(), ops.size(), got, 4)) applied++;
                } catch (const std::overflow_error&) {
                }
                bailed++;
                continue;
            }
            m.for_each([&](int32_t k) { want.push_back(k); });
            if (m.info_increment() < 32) halved++;
            if (sorted_ops_order(ops.data(), tr.data(), ops.size(), got, 4)) {
                applied++;
                if (got == want) equal++;
                min_inc = std::min(min_inc, m.info_increment());
            } else {
                bailed++;
            }
        }
        printf("%d %d %d %d %u\n", applied, equal, halved, bailed, min_inc);
        return 0;
    }
    if (argc >= 4 && std::string(argv[1]) == "--order-check") {
        // random walk keys + a random realigner log: the sorted form, when it claims to apply,
        // must equal the replay.  Prints "applied equal" counts.
        const int n = atoi(argv[2]);
        unsigned seed = (unsigned)atoi(argv[3]);
        auto rnd = [&]() {
            seed = seed * 1664525u + 1013904223u;
            return seed >> 8;
        };
        int applied = 0, equal = 0;
        for (int trial = 0; trial < 20; trial++) {
            std::vector<int32_t> created;
            for (int i = 0; i < n; i++) created.push_back((int32_t)(i * 3 + (int)(rnd() % 3)));
            std::vector<KeyOp> log;
            const int nl = (int)(rnd() % 6);
            for (int i = 0; i < nl; i++) {
                const bool ins = rnd() % 2;
                const int32_t k = rnd() % 3 ? created[rnd() % created.size()] : (int32_t)(n * 3 + (int)(rnd() % 1000));
                log.push_back({k, ins});
            }
            std::vector<int32_t> want = noninsertion_order(created, log), got;
            if (noninsertion_order_sorted(rh_sort_keys(created, rh_capacity(created.size())), log, got)) {
                applied++;
                if (got == want) equal++;
            }
        }
        printf("%d %d\n", applied, equal);
        return 0;
    }
    if (argc < 8) return 2;
    std::string fa = argv[1], rvb = argv[2];
    int rs = atoi(argv[3]), re = atoi(argv[4]), k = atoi(argv[5]), fisher = atoi(argv[6]);
    const char* outp = argv[7];
    std::ifstream ff(fa);
    std::string line, seq;
    while (std::getline(ff, line))
        if (!line.empty() && line[0] != '>') seq += line;
    for (char& c : seq) c = (char)toupper((unsigned char)c);
    std::ifstream f(rvb, std::ios::binary);
    char mg[4];
    uint32_t ver;
    int64_t n, nb, nc;
    f.read(mg, 4);
    f.read((char*)&ver, 4);
    f.read((char*)&n, 8);
    f.read((char*)&nb, 8);
    f.read((char*)&nc, 8);
    std::vector<int32_t> pos, nm, mpos, tid, mtid;
    std::vector<uint16_t> flag;
    std::vector<uint8_t> mapq, hasnm, bases, quals;
    std::vector<int64_t> so, co;
    std::vector<uint32_t> cig;
    rd(f, pos, n); rd(f, flag, n); rd(f, mapq, n); rd(f, hasnm, n); rd(f, nm, n); rd(f, mpos, n); rd(f, tid, n);
    rd(f, mtid, n); rd(f, so, n + 1); rd(f, co, n + 1); rd(f, bases, nb); rd(f, quals, nb); rd(f, cig, nc);
    // the region iterator's records
    std::vector<int32_t> P0, NM, MP, TI, MT;
    std::vector<uint16_t> FL;
    std::vector<uint8_t> MQ, HN, B, Q;
    std::vector<int64_t> SO{0}, CO{0};
    std::vector<uint32_t> CG;
    for (int64_t i = 0; i < n; i++) {
        int rl = 0;
        for (int64_t j = co[i]; j < co[i + 1]; j++) {
            int o = cig[j] & 15;
            if (o == 0 || o == 2 || o == 3 || o == 7 || o == 8) rl += (int)(cig[j] >> 4);
        }
        if (rl == 0) rl = 1;
        if (!(pos[i] < re && pos[i] + rl > rs - 1)) continue;
        P0.push_back(pos[i]); FL.push_back(flag[i]); MQ.push_back(mapq[i]); HN.push_back(hasnm[i]); NM.push_back(nm[i]);
        MP.push_back(mpos[i]); TI.push_back(tid[i]); MT.push_back(mtid[i]);
        for (int64_t j = so[i]; j < so[i + 1]; j++) {
            B.push_back(bases[j]);
            Q.push_back(quals[j]);
        }
        SO.push_back((int64_t)B.size());
        for (int64_t j = co[i]; j < co[i + 1]; j++) CG.push_back(cig[j]);
        CO.push_back((int64_t)CG.size());
    }
    const int64_t m = (int64_t)P0.size();
    ReadsDev R{P0.data(), FL.data(), MQ.data(), HN.data(), NM.data(), MP.data(), TI.data(), MT.data(),
               SO.data(), CO.data(), B.data(), Q.data(), CG.data(), m};
    Config cfg{};
    cfg.goodq = 22.5; cfg.bias = 0.05; cfg.freq = 0.01; cfg.qratio = 1.5; cfg.mapq = 0; cfg.lofreq = 0.05;
    cfg.vext = 2; cfg.mismatch = 8; cfg.performLocalRealignment = k; cfg.indelsize = 50; cfg.minBiasReads = 2;
    cfg.minr = 2; cfg.samfilter = 0x504; cfg.readPosFilter = 5; cfg.SVMINLEN = 1000; cfg.referenceExtension = 1200;
    cfg.fisher = fisher;
    if (const char* e = getenv("RV_SIM_DEDUP")) cfg.removeDuplicatedReads = atoi(e);         // -t
    if (const char* e = getenv("RV_SIM_SAMFILTER")) cfg.samfilter = (int)strtol(e, nullptr, 16);  // -F
    if (getenv("RV_SIM_PILEUP")) {  // -p (the reference then also sets -f -1 and -r 0)
        cfg.doPileup = 1;
        cfg.freq = -1;
        cfg.minr = 0;
    }
    const int chrLen = (int)seq.size();
    Window W;
    W.region_start = rs;
    W.region_end = re;
    W.chr_len = chrLen;
    int ss = rs - 1200 < 1 ? 1 : rs - 1200, se = re + 1200 > chrLen ? chrLen : re + 1200;
    W.ref_lo = ss;
    W.ref_hi = se - 17;
    std::string win = seq.substr((size_t)(ss - 1), (size_t)(W.ref_hi - W.ref_lo + 1));
    W.ref = win.c_str();
    std::vector<PrepRec> prep((size_t)std::max<int64_t>(m, 1));
    std::vector<uint32_t> cg((size_t)std::max<int64_t>(m, 1) * MAXC);
    WalkCounters ctr{};
    int span = 0;
    std::vector<uint8_t> dup;
    if (cfg.removeDuplicatedReads) {
        dup.resize((size_t)std::max<int64_t>(m, 1));
        for (int64_t r = 0; r < m; r++) dup[(size_t)r] = is_duplicate(R, cfg, r) ? 1 : 0;
    }
    for (int64_t r = 0; r < m; r++)
        prep_one(r, R, W, cfg, prep.data(), cg.data(), &ctr, &span, dup.empty() ? nullptr : dup.data());
    if (const char* pd = getenv("RV_SIM_PREP_DUMP")) {
        // per read: status, position, CIGAR word count and the words of its record slot
        FILE* fp = fopen(pd, "w");
        for (int64_t r = 0; r < m; r++) {
            const PrepRec& p = prep[(size_t)r];
            fprintf(fp, "%d %d %d", p.status, p.position, p.ncig);
            if (p.status == RD_PARSE)
                for (int i = 0; i < p.ncig; i++) fprintf(fp, " %u", cg[(size_t)r * MAXC + (size_t)i]);
            fprintf(fp, "\n");
        }
        fclose(fp);
    }
    if (ctr.capacity) {
        fprintf(stderr, "capacity: %llu read(s) refused\n", (unsigned long long)ctr.capacity);
        return 4;
    }
    const int pad = 2 * std::max(span, 64) + 8;
    W.dense_lo = rs - pad;
    W.dense_n = (re + pad) - W.dense_lo + 1;
    std::vector<DenseRow> rows((size_t)W.dense_n);
    for (int i = 0; i < W.dense_n; i++) fill_row_one(rows.data(), i);
    std::vector<Event> ev((size_t)(8 * m + 4096));
    std::vector<char> arena((size_t)(128 * m + 65536));
    std::vector<ClipRec> cl((size_t)(m + 1024));
    std::vector<char> carena((size_t)(320 * (m + 1024)));
    std::vector<Anchor> an((size_t)(m + 1024));
    unsigned ne = 0, ncl = 0, na = 0;
    unsigned long long au = 0, cu = 0;
    WalkSink S{rows.data(), W.dense_lo, W.dense_n, ev.data(), &ne, (unsigned)ev.size(), arena.data(), &au,
               arena.size(), cl.data(), &ncl, (unsigned)cl.size(), carena.data(), &cu, carena.size(), an.data(), &na,
               (unsigned)an.size(), &ctr};
    // workgroups of nreads reads sharing one tile, as k_walk runs them
    const int nreads = walk_reads_for(m, W.dense_n, span);
    for (int64_t r0 = 0; r0 < m; r0 += nreads) {
        const int tile_lo = prep[(size_t)r0].position - TILE_MARGIN;
        tile_init(0, 1, W, tile_lo);
        for (int64_t r = r0; r < r0 + nreads && r < m; r++)
            walk_one(r, R, W, cfg, prep.data(), cg.data(), S, 0, 0, tile_lo);
        tile_flush(tile_lo, rows.data(), W.dense_lo, W.dense_n, 0, 1);
    }
    for (unsigned i = 0; i < na; i++) anchor_one(i, an.data(), rows.data(), W.dense_lo, W.dense_n, cfg);
    // nonInsertionVariants keys in creation order
    std::vector<std::pair<uint64_t, int32_t>> ft;
    for (int i = 0; i < W.dense_n; i++)
        if (rows[(size_t)i].rowflags & 1u) ft.push_back({rows[(size_t)i].first_touch, W.dense_lo + i});
    std::sort(ft.begin(), ft.end());
    std::vector<int32_t> created;
    for (auto& e : ft) created.push_back(e.second);
    HostPileup HP;
    HP.dense_lo = W.dense_lo;
    HP.dense_n = W.dense_n;
    HP.rows = rows;
    HP.maxReadLength = ctr.max_read_length;
    reduce_events(ev.data(), ne, arena.data(), cfg.goodq, HP);
    reduce_clips(cl.data(), ncl, carena.data(), cfg.goodq, HP);
    HostRowSource hsrc(HP);
    RowCache rc(&hsrc, HP.dense_lo, HP.dense_n);
    RState RS(HP, rc);
    HostRef ref;
    ref.seq = win.data();
    ref.lo = W.ref_lo;
    ref.hi = W.ref_hi;
    // noPassingReads through the device kernel's per-record predicate, one record at a time
    struct SimReadQuery : ReadQuery {
        const ReadsDev& R;
        explicit SimReadQuery(const ReadsDev& r) : R(r) {}
        bool no_passing(int start, int end) override {
            for (int64_t i = 0; i < R.n; i++)
                if (passes_one(R, i, start, end)) return false;
            return true;
        }
    } rq(R);
    // RV_SIM_REALIGN_REPS=N: realign N extra times on copies of the state first (profiling)
    if (const char* reps = getenv("RV_SIM_REALIGN_REPS")) {
        for (int k = 0; k < atoi(reps); k++) {
            HostPileup hp2 = HP;
            HostRowSource src2(hp2);
            RowCache rc2(&src2, hp2.dense_lo, hp2.dense_n);
            RState rs2(hp2, rc2);
            HostRef ref2;
            ref2.seq = win.data();
            ref2.lo = W.ref_lo;
            ref2.hi = W.ref_hi;
            RealignStats st2;
            realign(cfg, rs, re, ref2, chrLen, &rq, rs2, st2);
        }
    }
    // RV_SIM_SCREEN=1: build the realignment screen's candidate list first (timed) and collect
    // change labels in the realigner
    ScreenPlan plan;
    ScreenPlan* planp = nullptr;
    if (getenv("RV_SIM_SCREEN")) {
        const auto q0 = std::chrono::steady_clock::now();
        screen_candidates(HP, plan);
        const double qms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - q0).count();
        plan.label = true;
        planp = &plan;
        fprintf(stderr, "screen list: %zu candidates (mnp %d del %d ins %d clusters %d), %.2f ms\n", plan.cand.size(),
                plan.off[1] - plan.off[0], plan.off[2] - plan.off[1], plan.off[3] - plan.off[2],
                plan.off[4] - plan.off[3], qms);
    }
    RealignStats rst;
    realign(cfg, rs, re, ref, chrLen, &rq, RS, rst, false, planp);
    if (planp) {
        int64_t pos[4] = {0, 0, 0, 0};
        for (size_t i = 0; i < plan.cand.size(); i++) pos[plan.cand[i].kind] += plan.changed[i];
        fprintf(stderr, "screen labels: changed mnp %lld del %lld ins %lld clusters %lld\n", (long long)pos[0],
                (long long)pos[1], (long long)pos[2], (long long)pos[3]);
    }
    CxHost cx;
    build_record_list(RS, cx);
    ScoreIn in{};
    in.rows = rows.data();
    in.dense_lo = W.dense_lo;
    in.dense_n = W.dense_n;
    in.cx_pos = cx.pos.data();
    in.n_cx = (int)cx.pos.size();
    in.cx_off = cx.off.data();
    in.cx = cx.ent.data();
    in.cx_arena = cx.arena.data();
    in.cx_cov = cx.cov.data();
    in.cx_key = cx.key.data();
    in.ref = win.data();
    in.ref_lo = W.ref_lo;
    in.ref_hi = W.ref_hi;
    in.chr_len = chrLen;
    in.region_start = rs;
    in.region_end = re;
    in.cfg = cfg;
    std::vector<OutRec> out(200000);
    std::vector<char> oarena((size_t)64 << 20);
    unsigned nout = 0;
    unsigned long long aused = 0, sctr[4] = {0, 0, 0, 0};
    in.out = out.data();
    in.n_out = &nout;
    in.max_out = (unsigned)out.size();
    in.arena = oarena.data();
    in.arena_used = &aused;
    in.arena_cap = oarena.size();
    in.counters = sctr;
    std::vector<uint8_t> status((size_t)W.dense_n + cx.pos.size(), 0);
    in.status = status.data();
    for (int i = 0; i < W.dense_n; i++) {
        // the pick pass's flags-only check must agree with the scoring kernel's own checks
        PosView<4> pv;
        load_pos<4>(in, W.dense_lo + i, pv);
        if (dense_needed(in, W.dense_lo + i) != score_needed<4>(in, W.dense_lo + i, true, pv)) {
            fprintf(stderr, "pick check differs at %d\n", W.dense_lo + i);
            return 4;
        }
        score_position<128, 4>(in, W.dense_lo + i, true, i);
    }
    for (int i = 0; i < in.n_cx; i++) score_position<1536, MAXV>(in, cx.pos[(size_t)i], false, W.dense_n + i);
    out.resize(std::min<unsigned>(nout, (unsigned)out.size()));
    // the printer's order: alignedVariants slot order, then the variants' order within a position
    std::vector<int32_t> proc = noninsertion_order(created, RS.key_log);
    {
        // the sorted form must agree with the replay whenever it claims to apply
        std::vector<int32_t> sorted_walk = rh_sort_keys(created, rh_capacity(created.size())), proc2;
        const bool ok = noninsertion_order_sorted(sorted_walk, RS.key_log, proc2);
        fprintf(stderr, "order: keys %zu log %zu sorted-form %s\n", created.size(), RS.key_log.size(),
                ok ? (proc2 == proc ? "applies, equal" : "applies, DIFFERS") : "not provable");
        if (ok && proc2 != proc) return 4;
    }
    auto status_of = [&](int32_t p) -> int {
        auto it = std::lower_bound(cx.pos.begin(), cx.pos.end(), p);
        if (it != cx.pos.end() && *it == p) return status[(size_t)W.dense_n + (size_t)(it - cx.pos.begin())];
        const int i = p - W.dense_lo;
        return (i >= 0 && i < W.dense_n) ? status[(size_t)i] : 0;
    };
    std::vector<uint8_t> merged(status.begin(), status.begin() + W.dense_n);
    for (size_t k = 0; k < cx.pos.size(); k++) {
        const int i = cx.pos[k] - W.dense_lo;
        if (i >= 0 && i < W.dense_n) merged[(size_t)i] = status[(size_t)W.dense_n + k];
        else if (status[(size_t)W.dense_n + k]) {
            fprintf(stderr, "order: status outside the dense window at %d\n", cx.pos[k]);
            return 4;
        }
    }
    const auto ta = std::chrono::steady_clock::now();
    bool aligned_sorted = false;
    std::vector<int32_t> aligned = aligned_slot_order(proc, merged.data(), W.dense_lo, W.dense_n, 4, &aligned_sorted);
    fprintf(stderr, "order: alignedVariants %zu keys, %.1f ms\n", aligned.size(),
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count());
    {
        RHOrder m;
        for (int32_t p : proc) {
            const int st = status_of(p);
            if (st == 0) continue;
            m.insert(p);
            if (st == 2) m.erase(p);
        }
        std::vector<int32_t> ref;
        m.for_each([&](int32_t k) { ref.push_back(k); });
        if (ref != aligned) {
            fprintf(stderr, "order: alignedVariants (%s) DIFFERS from the replay\n",
                    aligned_sorted ? "per-growth sort" : "replay");
            return 4;
        }
    }
    std::unordered_map<int32_t, int32_t> rank;
    for (size_t i = 0; i < aligned.size(); i++) rank[aligned[i]] = (int32_t)i;
    std::sort(out.begin(), out.end(), [&](const OutRec& a, const OutRec& b) {
        const int32_t ra = rank.count(a.position) ? rank[a.position] : INT32_MAX;
        const int32_t rb = rank.count(b.position) ? rank[b.position] : INT32_MAX;
        return ra != rb ? ra < rb : (a.position != b.position ? a.position < b.position : a.order < b.order);
    });
    LineContext ctx{"S1", "chrS", "chrS", rs, re, fisher != 0};
    std::string text;
    for (auto& r : out) format_line(text, r, oarena.data(), ctx);
    std::ofstream o(outp);
    o << text;
    fprintf(stderr, "realign sizes: mnp %zu del %zu ins %zu nd %zu insv %zu sc5 %zu sc3 %zu rows_fetched %lld fetches %lld touched %zu\n",
            HP.mnp.size(), HP.delCount.size(), HP.insCount.size(), HP.nd.size(), HP.ins.size(), HP.sc5.size(),
            HP.sc3.size(), (long long)rc.rows_fetched, (long long)rc.fetches, RS.pos.size());
    fprintf(stderr, "realign ms: noPassingReads %.1f (%lld calls), row misses %.1f (%lld)\n", rst.ms_nopass,
            (long long)rst.n_nopass, rc.miss_ms, (long long)rc.miss_fetches);
    fprintf(stderr, "realign ms: prefetch %.1f mnp %.1f del %.1f ins %.1f lgins %.1f\n", rst.ms_prefetch, rst.ms_mnp,
            rst.ms_del, rst.ms_ins, rst.ms_lgins);
    fprintf(stderr, "reads %lld events %u clips %u walk_errors %llu overflow %llu capacity %llu\n", (long long)m, ne, ncl,
            (unsigned long long)ctr.walk_errors, (unsigned long long)ctr.overflow, sctr[0]);
    return (ctr.walk_errors || ctr.overflow || sctr[0] || nout > 200000) ? 3 : 0;
}
