// Shared definitions for the RabbitVar per-region calling path on MI355X (gfx950).
//
// Everything here is plain data plus small inline helpers usable from HIP device code and from the
// host engine (RV_HD = __host__ __device__).  The host-side uses are bookkeeping on records the
// device produced (descriptor strings, counts); no stage of the path runs on the CPU twice.
#pragma once
#include <cstddef>
// Includers provide the __host__/__device__ qualifiers (hip/hip_runtime.h in the HIP sources).
#include <stddef.h>
#include <stdint.h>

#define RV_HD __host__ __device__ __forceinline__

namespace rv {

// BAM CIGAR op codes (SAM spec; the reference reads them through htslib's bam_cigar_op)
enum : int { OP_M = 0, OP_I = 1, OP_D = 2, OP_N = 3, OP_S = 4, OP_H = 5, OP_P = 6, OP_EQ = 7, OP_X = 8, OP_B = 9 };
RV_HD int cop(uint32_t w) { return (int)(w & 0xfu); }
RV_HD int clen(uint32_t w) { return (int)(w >> 4); }
RV_HD uint32_t cword(int len, int op) { return ((uint32_t)len << 4) | (uint32_t)op; }
RV_HD char op_char(int op) {
    const char t[16] = {'M', 'I', 'D', 'N', 'S', 'H', 'P', '=', 'X', 'B', '?', '?', '?', '?', '?', '?'};
    return t[op & 15];
}
// seq_nt16_str code of a decoded base (inverse of htslib's table "=ACMGRSVTWYHKDBN")
RV_HD int nt16_code(char c) {
    switch (c) {
        case '=': return 0;
        case 'A': return 1;
        case 'C': return 2;
        case 'M': return 3;
        case 'G': return 4;
        case 'R': return 5;
        case 'S': return 6;
        case 'V': return 7;
        case 'T': return 8;
        case 'W': return 9;
        case 'Y': return 10;
        case 'H': return 11;
        case 'K': return 12;
        case 'D': return 13;
        case 'B': return 14;
        default: return 15;
    }
}
// dense allele index for single-base descriptors (A,C,G,T); -1 for anything else
RV_HD int base_index(char c) {
    switch (c) {
        case 'A': return 0;
        case 'C': return 1;
        case 'G': return 2;
        case 'T': return 3;
        default: return -1;
    }
}
RV_HD char index_base(int i) {
    const char t[4] = {'A', 'C', 'G', 'T'};
    return t[i & 3];
}

// Fixed-capacity string used for CIGAR text and allele descriptors on the device.
template <int CAP>
struct FStr {
    int n;
    char c[CAP];
    bool overflow;
    RV_HD void clear() {
        n = 0;
        overflow = false;
    }
    RV_HD void push(char ch) {
        if (n < CAP)
            c[n++] = ch;
        else
            overflow = true;
    }
    RV_HD void append(const char* p, int k) {
        for (int i = 0; i < k; i++) push(p[i]);
    }
    RV_HD void append_int(long long v) {
        char t[24];
        int k = 0;
        bool neg = v < 0;
        unsigned long long u = neg ? (unsigned long long)(-v) : (unsigned long long)v;
        do {
            t[k++] = (char)('0' + (int)(u % 10));
            u /= 10;
        } while (u);
        if (neg) push('-');
        while (k) push(t[--k]);
    }
    template <int C2>
    RV_HD void append(const FStr<C2>& o) {
        append(o.c, o.n);
    }
    RV_HD int find(char ch, int from = 0) const {
        for (int i = from; i < n; i++)
            if (c[i] == ch) return i;
        return -1;
    }
    RV_HD void erase_at(int i) {
        if (i < 0 || i >= n) return;
        for (int k = i; k + 1 < n; k++) c[k] = c[k + 1];
        n--;
    }
    RV_HD void insert_at(int i, char ch) {
        if (i < 0 || i > n) return;
        if (n >= CAP) {
            overflow = true;
            return;
        }
        for (int k = n; k > i; k--) c[k] = c[k - 1];
        c[i] = ch;
        n++;
    }
    template <int C2>
    RV_HD bool eq(const FStr<C2>& o) const {
        if (n != o.n) return false;
        for (int i = 0; i < n; i++)
            if (c[i] != o.c[i]) return false;
        return true;
    }
};

// Configuration fields used on the calling path (reference: include/Configuration.h:98-401).
struct Config {
    double goodq;
    double bias;
    double freq;
    double qratio;
    double mapq;
    double lofreq;
    int vext;
    int mismatch;
    int trimBasesAfter;
    int performLocalRealignment;
    int indelsize;
    int minBiasReads;
    int minr;
    int moveIndelsTo3;
    int samfilter;
    int readPosFilter;
    int doPileup;
    int minmatch;
    int uniqueModeAlignmentEnabled;
    int uniqueModeSecondInPairEnabled;
    int fisher;
    int SVMINLEN;
    int referenceExtension;
    int numberNucleotideToExtend;
    int mappingQuality;
    int removeDuplicatedReads;
    int crisprCuttingSite;
    int hasBam2;
};

// Region and reference window as the device sees them.
struct Window {
    int region_start, region_end;  // 1-based inclusive (Region.start/end)
    int ref_lo, ref_hi;            // reference window keys [ref_lo, ref_hi] (makeReference)
    int chr_len;
    const char* ref;               // ref[k - ref_lo], upper-case
    int dense_lo, dense_n;         // dense position table covers [dense_lo, dense_lo + dense_n)
};

// Per-read prep result (device -> walk kernel)
enum : int {
    RD_PARSE = 0,      // walk the read
    RD_FILTERED = 1,   // removed by the record filters (flag, MAPQ, l_qseq == 1)
    RD_NM = 2,         // removed by NM - indels > mismatch
    RD_NOWALK = 3,     // parsed but returns before the CIGAR walk (both-end clips, -M, supplementary)
};
constexpr int MAXC = 48;        // CIGAR words kept per read (rewritten CIGAR + record tail emulation)
constexpr int MAXC_IN = MAXC - 4;  // longest input CIGAR accepted: the rewrite adds at most 2 ops and
                                   // may read one word past the CIGAR
constexpr int CIGTXT = 192;     // CIGAR text capacity for the rewrite rules

struct PrepRec {
    int32_t status;
    int32_t position;   // 1-based alignment start after the rewrite
    int32_t ncig;       // CIGAR length used by the walk
    int32_t nm_eff;     // NM - indel length
    int32_t rlen1;      // M + I length
    int32_t total_len;  // M + I + S length
    int32_t lseq;
    int32_t flags;      // bit0: rewrite changed the CIGAR; bit1: rewrite text overflow; bit2: negative length
    int32_t reflen;     // M + D + N + = + X length of the record's own CIGAR
    int32_t pad;
};

// Dense single-base allele accumulators, one row per reference position (AoS, 288-byte rows).
// Integer fields are exact; pstd/qstd are reconstructed from the first (tp, q) value and "differs"
// flags, which equals the reference's running comparison whenever every contribution has tp >= 1
// and q >= 1 (has_q0 marks positions where a Phred-0 base broke that premise).
struct DenseBase {
    int32_t cnt, fwd, rev, hi, lo;
    int32_t sum_tp, sum_q, sum_mq, sum_nm;
    uint32_t first_tpq;   // (tp << 8 | q) + 1 of the first contribution that updates pp/pq, 0 = none
    uint32_t flags;       // bit0 exists, bit1 pstd, bit2 qstd, bit3 q==0 seen
    uint32_t min_ord;     // smallest read ordinal that created the entry (0xffffffff = none)
    int32_t last_tp;      // pp of the last contribution in read order (host realigner input)
    uint32_t pad[3];
};
static_assert(sizeof(DenseBase) == 64, "DenseBase layout");
// the walk's flush adds (cnt, fwd), (rev, hi), (lo, sum_tp) and (sum_q, sum_mq) as 64-bit pairs
static_assert(offsetof(DenseBase, cnt) == 0 && offsetof(DenseBase, fwd) == 4 && offsetof(DenseBase, rev) == 8 &&
                  offsetof(DenseBase, hi) == 12 && offsetof(DenseBase, lo) == 16 && offsetof(DenseBase, sum_tp) == 20 &&
                  offsetof(DenseBase, sum_q) == 24 && offsetof(DenseBase, sum_mq) == 28,
              "DenseBase pairs for 64-bit atomics");
struct DenseRow {
    DenseBase b[4];
    int32_t cov;       // refCoverage
    uint32_t rowflags; // bit0: row touched by any getVariation on the position (nonIns key exists)
    unsigned long long first_touch;  // (read ordinal << 24 | touch index within the read) of the key creation
    uint32_t pad[4];
};
static_assert(sizeof(DenseRow) == 288, "DenseRow layout");

// Non-dense variation contribution (multi-character descriptors, non-ACGT single bases,
// insertions).  Reduced on the host in read order, which keeps the FP64 sums bit-exact.
struct Event {
    int32_t pos;
    int32_t ord;        // read ordinal within the region batch
    int32_t seq;        // contribution index within the read
    int32_t tp;
    double q;           // per-read quality value added to meanQuality (already divided)
    int32_t mapq;
    int32_t nmv;        // numberOfMismatches increment
    uint32_t desc_off;  // descriptor bytes in the event arena
    uint16_t desc_len;
    uint8_t kind;       // 0 nonInsertionVariants, 1 insertionVariants, 2 rescue (addCnt form), 3 anchor subtraction
    uint8_t dir;
    uint8_t cntflags;   // bit0 delCount++, bit1 insCount++, bit2 mnp++, bit3 updates pp/pq
    uint8_t hiq;        // 1 when q >= goodq (hi) else lo
    uint8_t pad[2];
    int32_t rp;         // readPosition argument for addCnt-form contributions
};

// Soft-clip record (one per clipped read end that reaches Sclip accumulation)
struct ClipRec {
    int32_t pos;
    int32_t ord;
    int32_t side;       // 5 or 3
    int32_t dir;
    int32_t mapq;
    int32_t nm;
    int32_t cel;        // readPosition argument of the clip-level addCnt
    int32_t nhq;        // clipped high-quality bases kept
    int32_t sumq;
    uint32_t bases_off; // nhq bases then nhq qualities in the clip arena, clip-index order
};

struct WalkCounters {
    unsigned long long events;
    unsigned long long event_bytes;
    unsigned long long clips;
    unsigned long long clip_bytes;
    unsigned long long overflow;     // events/clips/arena overflow (results incomplete)
    unsigned long long ref_insert;   // reads that would create a missing reference key
    unsigned long long walk_errors;  // reads whose walk hit an out-of-range access
    int max_read_length;
    int ttref;                       // reads reaching the insertion-after-clip branch
    unsigned long long capacity;     // reads whose CIGAR or its rewrite exceeds MAXC words / CIGTXT chars
};

}  // namespace rv
