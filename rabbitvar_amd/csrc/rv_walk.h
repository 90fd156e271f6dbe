// Per-read CIGAR walk into the pileup (device).
//
// Restates CigarParser::parseCigar's walk (reference src/parseCigar.cpp:629-972) and its handlers
// process_softclip (:1112-1301), process_insertion (:1303-1523), process_deletion (:1525-1727),
// findOffset (:208-256), appendSegments (:1756-1815), adjInsPos (include/VariationUtils.h:99-113).
//
// Output routing:
//   * single-base A/C/G/T alleles -> DenseRow integer atomics (exact in any order);
//   * every other allele descriptor, insertions, and non-ACGT single bases -> Event records that
//     the engine reduces in read order (their FP64 quality sums are order-sensitive);
//   * soft clips that reach Sclip accumulation -> ClipRec + clip bases/qualities;
//   * the insertion-anchor subtraction (parseCigar.cpp:1471-1490) -> Anchor records, applied after
//     the walk once every entry's creating read ordinal is known.
//
// One wavefront walks one read.  The read's control flow is lane-uniform: every lane keeps the
// same walker state and runs the same steps, and only lane 0 (the leader) issues the side effects.
// Inside an M run the plain bases (neither the start of a multi-base allele, nor the last bases
// before an adjacent indel, nor non-ACGT) are spread over the lanes, one base per lane, so a
// wavefront updates 64 consecutive positions at once; every other base goes through the
// sequential steps.
#pragma once
#include "rv_cigar.h"

namespace rv {

constexpr int DCAP = 640;  // allele descriptor capacity
typedef FStr<DCAP> DStr;

// Workgroup-local accumulator for a window of TILE_POS consecutive positions (LDS).  The reads of
// one workgroup add their single-base contributions here with LDS atomics that return nothing (no
// lane waits on them); the tile is then added to the dense rows once per touched (position, base).
// Every field merges exactly: counts and sums add, creation ordinals take the minimum, flags OR.
// The per-tile fields are packed: TF_CNT holds the forward, reverse and high-quality counts in
// 10-bit fields (cnt = fwd + rev and lo = cnt - hi always hold), TF_QMQ the quality and mapping
// quality sums in 16-bit halves, and the (position, quality) values that decide pstd / qstd are
// kept as their AND and OR (all equal exactly when AND == OR, field by field).  A read adds to a
// (position, base) at most twice (its aligned base and a rescued clip base lie on different
// positions, so in practice once), so a workgroup's sums stay within the packed widths.
constexpr int TILE_POS = 768;
enum : int { TF_CNT, TF_TP, TF_QMQ, TF_NM, TF_MINORD, TF_FLAGS, TF_AND, TF_OR, TF_N };
constexpr unsigned TF_REV1 = 1u << 10, TF_HI1 = 1u << 20;  // TF_CNT increments (forward: 1)
constexpr unsigned TF_HASPQ = 16u;                         // TF_FLAGS: a pp/pq-updating contribution
struct TileAcc {
    uint32_t f[TF_N][4][TILE_POS];  // field, base, position
    int32_t cov[TILE_POS];
    unsigned long long first_touch[TILE_POS];  // ~0: no key created (the row's rowflags bit stays 0)
};
// the walk kernel's tile, addressed directly so that its accesses are LDS instructions
static __shared__ TileAcc walk_tile;

// The walk kernel's workgroup: WALK_THREADS lanes (16 wavefronts, 4 per SIMD: the walk fits in 128
// registers, and the tile, read copies and strings in one CU's LDS) walk a launch-wide number of
// consecutive reads (walk_reads_for) into a tile starting TILE_MARGIN bases before the first read.
constexpr int WALK_THREADS = 1024;
constexpr int WALK_READS_MAX = 128;
constexpr int WALK_READS_MIN = 16;
constexpr int TILE_MARGIN = 64;
static_assert(2 * WALK_READS_MAX < 1024 && 2 * WALK_READS_MAX * 255 < 65536, "packed tile fields hold a workgroup's sums");
// Reads per walk workgroup for one launch: as many as keep a workgroup's reads inside its tile on
// average (their starts spread over n_reads / span_bp per base, each reaching at most max_span
// bases past its start), between WALK_READS_MIN and WALK_READS_MAX.  More reads per workgroup
// means fewer tiles flushed per position; reads that reach past the tile still add correctly,
// through device-scope atomics on the rows.  Host code.
inline int walk_reads_for(long long n_reads, long long span_bp, int max_span) {
    const int reach = max_span < TILE_POS / 2 ? max_span : TILE_POS / 2;
    const double density = span_bp > 0 ? (double)n_reads / (double)span_bp : 0.0;
    const double fit = (double)(TILE_POS - TILE_MARGIN - reach) * density;
    if (fit >= WALK_READS_MAX) return WALK_READS_MAX;
    if (fit <= WALK_READS_MIN) return WALK_READS_MIN;
    return (int)fit;
}

// Per-wavefront copy of the read being walked (bases, qualities, rewritten CIGAR), loaded with one
// coalesced pass so that the lane-uniform steps read LDS instead of waiting on global memory.
// Reads longer than STAGE_LEN are read from global memory.
constexpr int STAGE_LEN = 512;
constexpr int WALK_MAX_WAVES = 16;
struct WaveStage {
    char seq[STAGE_LEN];
    uint8_t qual[STAGE_LEN];
    uint32_t cig[MAXC];
};
static __shared__ WaveStage walk_stage[WALK_MAX_WAVES];

// Per-wavefront allele descriptor strings.  The walker's control flow is lane-uniform, so one copy
// per wavefront serves every lane (per-lane copies would live in scratch memory, 64 times over).
// At most WALK_STRS are live at once on any path: match_run (s, ss, t), insertion (desc, t / key,
// d), deletion (desc), rescue (d).
constexpr int WALK_STRS = 3;
static __shared__ DStr walk_str[WALK_THREADS / 64][WALK_STRS];
static_assert(WALK_THREADS / 64 <= WALK_MAX_WAVES, "one read copy per wavefront");

// Per-workgroup copy of the reference bases around the tile: [tile_lo - REF_BEFORE,
// tile_lo - REF_BEFORE + REF_STAGE).  Positions outside it are read from global memory.
constexpr int REF_BEFORE = 64;
constexpr int REF_STAGE = TILE_POS + 512;
static __shared__ char walk_ref[REF_STAGE];

struct Anchor {
    int32_t pos;
    int32_t ord;
    int32_t base;  // dense index 0..3
    int32_t dir;
    int32_t tp;
    int32_t q;
    int32_t mapq;
    int32_t nmv;
};

struct WalkSink {
    DenseRow* dense;
    int dense_lo, dense_n;
    Event* events;
    unsigned int* n_events;
    unsigned int max_events;
    char* arena;
    unsigned long long* arena_used;
    unsigned long long arena_cap;
    ClipRec* clips;
    unsigned int* n_clips;
    unsigned int max_clips;
    char* clip_arena;
    unsigned long long* clip_used;
    unsigned long long clip_cap;
    Anchor* anchors;
    unsigned int* n_anchors;
    unsigned int max_anchors;
    WalkCounters* ctr;
};

struct RefWin {
    const char* ref;
    int lo, hi;
    WalkCounters* ctr;
    bool leader;
    int slo, sn;  // staged copy: walk_ref[k - slo] for k - slo in [0, sn)
    __device__ __forceinline__ bool has(int k) const { return k >= lo && k <= hi; }
    // The staged copy is read with LDS instructions and the rare position outside it with a global
    // load.  The volatile access keeps the compiler from merging the two loads into one generic
    // (flat) load of a selected pointer, which would put a global-memory wait on every read.
    __device__ __forceinline__ char at(int k) const {
        const unsigned j = (unsigned)(k - slo);
        if (j < (unsigned)sn) return walk_ref[j];
        return *(volatile const char*)(ref + (k - lo));
    }
    // robin_hood operator[]: a missing key would be created with '\0'
    __device__ __forceinline__ char get(int k) const {
        if (has(k)) return at(k);
        if (leader) atomicAdd(&ctr->ref_insert, 1ull);
        return '\0';
    }
    __device__ __forceinline__ bool hasEq(int k, char ch) const { return has(k) && at(k) == ch; }
    __device__ __forceinline__ bool hasNe(int k, char ch) const { return has(k) && at(k) != ch; }
};

// Every member function is force-inlined: one out-of-line call makes the walker live in per-lane
// scratch memory, which puts a memory round trip on every state access of the walk.
struct Walker {
    Config cf;           // by value: no reload through memory the walk writes
    int Rs, Re;          // region
    int chrLen;
    RefWin ref;
    WalkSink S;
    // read
    const char* seq;
    const uint8_t* qual;
    uint8_t qual_before0;
    int lseq;
    int wave;            // wavefront index in the workgroup: its read copy (walk_stage) and
                         // descriptor strings (walk_str)
    bool staged;         // bases and qualities are in stage (else read through seq / qual)
    int rec_n;           // valid emulated words
    int n_cigar;
    int ord;
    int evseq;
    uint32_t tseq;       // nonInsertionVariants key touches so far in this read
    uint8_t mapq;
    int nm;
    bool dir;
    int position;
    int start, offset, cel, rpi, rpe;
    bool bad;
    int lane;      // lane within the wavefront walking this read
    bool leader;   // lane 0: issues the side effects of the lane-uniform steps
    bool use_tile; // accumulate into walk_tile (false: dense rows directly)
    int tile_lo;   // position of tile slot 0

    __device__ __forceinline__ uint32_t cw(int i) const { return (i >= 0 && i < rec_n) ? walk_stage[wave].cig[i] : 0u; }
    // staged: the wavefront's LDS copy; otherwise global memory (volatile for the reason RefWin::at
    // gives)
    __device__ __forceinline__ char base_at(int i) const {
        if (staged) return walk_stage[wave].seq[i];
        return *(volatile const char*)(seq + i);
    }
    __device__ __forceinline__ int qual_at(int i) const {
        if (staged) return walk_stage[wave].qual[i];
        return *(volatile const uint8_t*)(qual + i);
    }
    __device__ __forceinline__ int op(int i) const { return cop(cw(i)); }
    __device__ __forceinline__ int len(int i) const { return clen(cw(i)); }
    __device__ __forceinline__ int q(int i) {
        if (i >= 0 && i < lseq) return qual_at(i);
        if (i == -1) return qual_before0;
        bad = true;
        return 0;
    }
    __device__ __forceinline__ char sq(int i) const { return (i >= 0 && i < lseq) ? base_at(i) : '\0'; }
    // dst = src (distinct strings): the characters are copied lane-parallel
    __device__ __forceinline__ void copy_str(DStr& dst, const DStr& src) {
        for (int i = lane; i < src.n; i += warpSize) dst.c[i] = src.c[i];
        dst.n = src.n;
        dst.overflow = src.overflow;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    __device__ __forceinline__ void sub(DStr& d, int p, int n) {
        if (p < 0 || p > lseq) {
            bad = true;
            return;
        }
        for (int i = p; i < p + n && i < lseq; i++) d.push(base_at(i));
    }

    __device__ __forceinline__ DenseRow* row(int pos) {
        int i = pos - S.dense_lo;
        if (i < 0 || i >= S.dense_n) {
            bad = true;
            return nullptr;
        }
        return &S.dense[i];
    }
    // a getVariation on nonInsertionVariants at this row: the key exists from here on, and the
    // smallest (read, touch) pair records when the reference created it (its hash-map insertion
    // order, replayed on the host)
    __device__ __forceinline__ int tile_slot(int pos) const {
        const int i = pos - tile_lo;
        return (use_tile && i >= 0 && i < TILE_POS) ? i : -1;
    }
    __device__ __forceinline__ void touch_key(DenseRow* r, int pos, unsigned t) {
        const unsigned long long key = ((unsigned long long)(unsigned)ord << 24) | (t & 0xffffffu);
        const int ti = tile_slot(pos);
        if (ti >= 0) {
            atomicMin(&walk_tile.first_touch[ti], key);
            return;
        }
        if (!(r->rowflags & 1u)) atomicOr(&r->rowflags, 1u);
        if (key < r->first_touch) atomicMin(&r->first_touch, key);
    }
    __device__ __forceinline__ void touch_pos(DenseRow* r, int pos) {
        const unsigned t = tseq++;
        if (leader) touch_key(r, pos, t);
    }
    __device__ __forceinline__ void add_cov(DenseRow* r, int pos, int k) {
        const int ti = tile_slot(pos);
        if (ti >= 0)
            atomicAdd(&walk_tile.cov[ti], k);
        else
            atomicAdd(&r->cov, k);
    }
    __device__ __forceinline__ void cov(int pos, int k = 1) {
        DenseRow* r = row(pos);
        if (r && leader) add_cov(r, pos, k);
    }

    // single-base contribution from the match run (updates pp/pq)
    __device__ __forceinline__ void dense_single(int pos, int b, int tp, int qv, int nmv) {
        DenseRow* r = row(pos);
        if (!r) return;
        touch_pos(r, pos);
        if (leader) dense_add(r, pos, b, tp, qv, nmv);
    }
    // the pp/pq-updating contribution to tile slot ti (LDS atomics that return nothing)
    __device__ __forceinline__ void tile_single(int ti, int b, int tp, int qv, int nmv) {
        auto F = [&](int f) -> uint32_t* { return &walk_tile.f[f][b][ti]; };
        atomicAdd(F(TF_CNT), (dir ? TF_REV1 : 1u) + ((double)qv >= cf.goodq ? TF_HI1 : 0u));
        atomicAdd(F(TF_TP), (unsigned)tp);
        atomicAdd(F(TF_QMQ), ((unsigned)qv & 0xffffu) | ((unsigned)mapq << 16));
        atomicAdd(F(TF_NM), (unsigned)nmv);
        atomicMin(F(TF_MINORD), (unsigned)ord);
        atomicOr(F(TF_FLAGS), 1u | TF_HASPQ | ((tp <= 0 || qv <= 0) ? 8u : 0u));
        const unsigned v = (((unsigned)tp & 0xffffffu) << 8) | ((unsigned)qv & 0xffu);
        atomicAnd(F(TF_AND), v);
        atomicOr(F(TF_OR), v);
    }
    // the pp/pq-updating contribution (first (tp, q) value + differs flags)
    __device__ __forceinline__ void dense_add(DenseRow* r, int pos, int b, int tp, int qv, int nmv) {
        const int ti = tile_slot(pos);
        if (ti >= 0) {
            tile_single(ti, b, tp, qv, nmv);
            return;
        }
        DenseBase& d = r->b[b];
        atomicAdd(&d.cnt, 1);
        atomicAdd(dir ? &d.rev : &d.fwd, 1);
        atomicAdd((double)qv >= cf.goodq ? &d.hi : &d.lo, 1);
        atomicAdd(&d.sum_tp, tp);
        atomicAdd(&d.sum_q, qv);
        atomicAdd(&d.sum_mq, (int)mapq);
        atomicAdd(&d.sum_nm, nmv);
        atomicMin(&d.min_ord, (unsigned)ord);
        unsigned want = 1u | ((tp <= 0 || qv <= 0) ? 8u : 0u);
        if ((d.flags & want) != want) atomicOr(&d.flags, want);
        if ((d.flags & 6u) != 6u) {
            unsigned v = ((((unsigned)tp & 0xffffffu) << 8) | ((unsigned)qv & 0xffu)) + 1u;
            unsigned old = atomicCAS(&d.first_tpq, 0u, v);
            if (old != 0u && old != v) {
                unsigned o = old - 1u, nv = v - 1u;
                unsigned add = 0;
                if ((o >> 8) != (nv >> 8)) add |= 2u;
                if ((o & 0xffu) != (nv & 0xffu)) add |= 4u;
                if (add) atomicOr(&d.flags, add);
            }
        }
    }
    // addCnt-form contribution (soft-clip rescue): no pp/pq update
    __device__ __forceinline__ void dense_addcnt(int pos, int b, int rp, int qv) {
        DenseRow* r = row(pos);
        if (!r) return;
        touch_pos(r, pos);
        if (!leader) return;
        const int ti = tile_slot(pos);
        if (ti >= 0) {
            auto F = [&](int f) -> uint32_t* { return &walk_tile.f[f][b][ti]; };
            atomicAdd(F(TF_CNT), (dir ? TF_REV1 : 1u) + ((double)qv >= (double)(float)cf.goodq ? TF_HI1 : 0u));
            atomicAdd(F(TF_TP), (unsigned)rp);
            atomicAdd(F(TF_QMQ), ((unsigned)qv & 0xffffu) | ((unsigned)mapq << 16));
            atomicAdd(F(TF_NM), (unsigned)nm);
            atomicMin(F(TF_MINORD), (unsigned)ord);
            atomicOr(F(TF_FLAGS), 1u);
            return;
        }
        DenseBase& d = r->b[b];
        atomicAdd(&d.cnt, 1);
        atomicAdd(dir ? &d.rev : &d.fwd, 1);
        atomicAdd((double)qv >= (double)(float)cf.goodq ? &d.hi : &d.lo, 1);
        atomicAdd(&d.sum_tp, rp);
        atomicAdd(&d.sum_q, qv);
        atomicAdd(&d.sum_mq, (int)mapq);
        atomicAdd(&d.sum_nm, nm);
        atomicMin(&d.min_ord, (unsigned)ord);
        if (!(d.flags & 1u)) atomicOr(&d.flags, 1u);
    }

    __device__ __forceinline__ void event(int kind, int pos, const DStr& d, int tp, double qv, int nmv, int cntflags, int rp, bool hiq) {
        if (d.overflow) {
            if (leader) atomicAdd(&S.ctr->overflow, 1ull);
            return;
        }
        const int myseq = evseq++;
        if (kind == 0 || kind == 2) {
            DenseRow* r = row(pos);
            if (r) touch_pos(r, pos);
        }
        if (!leader) return;
        unsigned idx = atomicAdd(S.n_events, 1u);
        unsigned long long off = atomicAdd(S.arena_used, (unsigned long long)d.n);
        if (idx >= S.max_events || off + (unsigned long long)d.n > S.arena_cap) {
            atomicAdd(&S.ctr->overflow, 1ull);
            return;
        }
        for (int i = 0; i < d.n; i++) S.arena[off + i] = d.c[i];
        Event e;
        e.pos = pos;
        e.ord = ord;
        e.seq = myseq;
        e.tp = tp;
        e.q = qv;
        e.mapq = mapq;
        e.nmv = nmv;
        e.desc_off = (uint32_t)off;
        e.desc_len = (uint16_t)d.n;
        e.kind = (uint8_t)kind;
        e.dir = dir ? 1 : 0;
        e.cntflags = (uint8_t)cntflags;
        e.hiq = hiq ? 1 : 0;
        e.pad[0] = e.pad[1] = 0;
        e.rp = rp;
        S.events[idx] = e;
    }

    // a nonInsertionVariants contribution from the match run / deletion handlers
    __device__ __forceinline__ void var_contrib(int pos, const DStr& d, int tp, double qv, int nmv, int cntflags) {
        if (d.n == 1) {
            int b = base_index(d.c[0]);
            if (b >= 0) {
                dense_single(pos, b, tp, (int)qv, nmv);
                return;
            }
        }
        event(0, pos, d, tp, qv, nmv, cntflags | 8, 0, qv >= cf.goodq);
    }

    __device__ __forceinline__ int tp_of(int rlen1) const { return rpe < rlen1 - rpe ? rpe + 1 : rlen1 - rpe; }

    __device__ __forceinline__ bool isNextMatched(int ci) const { return cf.performLocalRealignment && n_cigar > ci + 1 && op(ci + 1) == OP_M; }
    __device__ __forceinline__ bool isNextInsertion(int ci) const {
        return cf.performLocalRealignment && n_cigar > ci + 1 && op(ci + 1) == OP_I;
    }
    __device__ __forceinline__ bool isNextAfterNumMatched(int ci, int number) const { return n_cigar > ci + number && op(ci + 1) == OP_M; }
    __device__ __forceinline__ bool isIDWithNextMatched(int ci) const {
        return cf.performLocalRealignment && n_cigar > ci + 2 && len(ci + 1) <= cf.vext && op(ci + 1) == OP_M &&
               (op(ci + 2) == OP_I || op(ci + 2) == OP_D) && op(ci + 3) != OP_I && op(ci + 3) != OP_D;
    }

    struct Off {
        int offset;
        int from;  // read offset of the appended sequence
        long q_sum;
        int nm;
    };
    __device__ __forceinline__ Off findOffset(int refPos, int readPos, int cigLen) {
        Off o{0, readPos, 0, 0};
        int vsn = 0;
        for (int vi = 0; vsn <= cf.vext && vi < cigLen; vi++) {
            if (sq(readPos + vi) == 'N') break;
            if (q(readPos + vi) < cf.goodq) break;
            if (ref.has(refPos + vi)) {
                if (sq(readPos + vi) != ref.at(refPos + vi)) {
                    o.offset = vi + 1;
                    o.nm++;
                    vsn = 0;
                } else {
                    vsn++;
                }
            }
        }
        if (o.offset > 0) {
            for (int i = readPos; i < readPos + o.offset; i++) o.q_sum += q(i);
            for (int k = 0; k < o.offset; k++) cov(refPos + k);
        }
        return o;
    }

    __device__ __forceinline__ void appendSegments(int ci, DStr& desc, long& qs, int& qc, int mLen, int indelLen, int begin, bool isIns) {
        desc.push('#');
        sub(desc, begin, mLen);
        for (int i = begin; i < begin + mLen; i++) qs += q(i);
        desc.push('^');
        if (op(ci + 2) == OP_I)
            sub(desc, begin + mLen, indelLen);
        else
            desc.append_int(indelLen);
        if (isIns) {
            if (op(ci + 2) == OP_I) {
                for (int i = begin + mLen; i < begin + mLen + indelLen; i++) qs += q(i);
                qc += indelLen;
            } else {
                qs += q(begin + mLen);
                qc += 1;
            }
        } else if (op(ci + 2) == OP_I) {
            for (int i = begin + mLen; i < begin + mLen + indelLen; i++) qs += q(i);
            qc += indelLen;
        }
    }

    __device__ __forceinline__ void emit_clip(int side, int pos, int cel_, int nhq, int sumq, int from, int step) {
        for (int k = 0; k < nhq; k++) (void)q(from + step * k);  // range checks on every lane
        if (!leader) return;
        unsigned idx = atomicAdd(S.n_clips, 1u);
        unsigned long long off = atomicAdd(S.clip_used, (unsigned long long)(2 * nhq));
        if (idx >= S.max_clips || off + (unsigned long long)(2 * nhq) > S.clip_cap) {
            atomicAdd(&S.ctr->overflow, 1ull);
            return;
        }
        for (int k = 0; k < nhq; k++) {
            int ri = from + step * k;
            S.clip_arena[off + k] = sq(ri);
            S.clip_arena[off + nhq + k] = (char)q(ri);
        }
        ClipRec r;
        r.pos = pos;
        r.ord = ord;
        r.side = side;
        r.dir = dir ? 1 : 0;
        r.mapq = mapq;
        r.nm = nm;
        r.cel = cel_;
        r.nhq = nhq;
        r.sumq = sumq;
        r.bases_off = (uint32_t)off;
        S.clips[idx] = r;
    }

    __device__ __forceinline__ void rescue(int pos, int rp, int qv) {
        char rb = ref.at(pos);
        int b = base_index(rb);
        if (b >= 0) {
            dense_addcnt(pos, b, rp, qv);
        } else {
            DStr& d = walk_str[wave][0];
            d.clear();
            d.push(rb);
            event(2, pos, d, 0, (double)qv, nm, 0, rp, (double)qv >= (double)(float)cf.goodq);
        }
        cov(pos);
    }

    __device__ __forceinline__ void softclip(int ci, int totalLen) {
        if (ci == 0) {
            while (cel - 1 >= 0 && start - 1 > 0 && start - 1 <= chrLen && ref.hasEq(start - 1, sq(cel - 1)) && q(cel - 1) > 10) {
                rescue(start - 1, cel, q(cel - 1));
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
                if (nhq >= 1 && nhq > nlq && start >= Rs && start <= Re) emit_clip(5, start, cel, nhq, sumq, cel - 1, -1);
            }
            cel = len(ci);
        } else if (ci == n_cigar - 1) {
            while (rpi < lseq && ref.hasEq(start, sq(rpi)) && q(rpi) > 10) {
                rescue(start, totalLen - rpe, q(rpi));
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
                if (nhq >= 1 && nhq > nlq && start >= Rs && start <= Re) emit_clip(3, start, cel, nhq, sumq, rpi, 1);
            }
        }
        rpi += cel;
        offset = 0;
        start = position;
    }

    __device__ __forceinline__ int insertion(int ci, int rlen1) {
        if ((n_cigar > ci + 1 && op(ci + 1) == OP_N) || (ci > 0 && op(ci - 1) == OP_N)) {
            rpi += cel;
            return ci;
        }
        DStr& desc = walk_str[wave][0];
        desc.clear();
        sub(desc, rpi, cel);
        long qs = 0;
        int qc = 0;
        for (int i = rpi; i < rpi + cel; i++) qs += q(i);
        qc += cel;
        int ss_from = 0, ss_len = 0;
        int multoffs = 0, multoffp = 0, nmoff = 0;
        if (isIDWithNextMatched(ci)) {
            int mLen = len(ci + 1), indelLen = len(ci + 2);
            int begin = rpi + cel;
            appendSegments(ci, desc, qs, qc, mLen, indelLen, begin, true);
            multoffs += mLen + (op(ci + 2) == OP_D ? indelLen : 0);
            multoffp += mLen + (op(ci + 2) == OP_I ? indelLen : 0);
            int ci6 = n_cigar > ci + 3 ? len(ci + 3) : 0;
            if (ci6 != 0 && op(ci + 3) == OP_M) {
                Off o = findOffset(start + multoffs, rpi + cel + multoffp, ci6);
                offset = o.offset;
                ss_from = o.from;
                ss_len = o.offset;
                qs += o.q_sum;
                qc += offset;
            }
            ci += 2;
        } else if (isNextMatched(ci)) {
            int vsn = 0;
            for (int vi = 0; vsn <= cf.vext && vi < len(ci + 1); vi++) {
                if (sq(rpi + cel + vi) == 'N') break;
                if (q(rpi + cel + vi) < cf.goodq) break;
                if (ref.has(start + vi)) {
                    if (sq(rpi + cel + vi) != ref.at(start + vi)) {
                        offset = vi + 1;
                        nmoff++;
                        vsn = 0;
                    } else {
                        vsn++;
                    }
                }
            }
            if (offset != 0) {
                ss_from = rpi + cel;
                ss_len = offset;
                for (int i = rpi + cel; i < rpi + cel + offset; i++) qs += q(i);
                qc += offset;
                for (int k = 0; k < offset; k++) cov(start + k);
            }
        }
        if (offset > 0) {
            desc.push('&');
            sub(desc, ss_from, ss_len);
        }
        if (start - 1 >= Rs && start - 1 <= Re && desc.find('N') < 0) {
            int inspos = start - 1;
            bool allATGC = desc.n > 0;
            for (int i = 0; i < desc.n; i++)
                if (base_index(desc.c[i]) < 0) allATGC = false;
            if (allATGC) {
                int bi = start - 1;
                int n = 1;
                int L = desc.n;
                int guard = 0;
                while (ref.get(bi) == desc.c[L - n]) {
                    n++;
                    if (n > L) n = 1;
                    bi--;
                    if (++guard > (1 << 24)) {
                        bad = true;
                        break;
                    }
                }
                if (n > 1) {
                    // rotate right by n-1: last n-1 characters move to the front
                    DStr& t = walk_str[wave][1];
                    t.clear();
                    t.append(desc.c + (L - (n - 1)), n - 1);
                    t.append(desc.c, L - (n - 1));
                    copy_str(desc, t);
                }
                if (rpi - 1 - (start - 1 - bi) > 0) inspos = bi;
            }
            DStr& key = walk_str[wave][1];
            key.clear();
            key.push('+');
            key.append(desc);
            int tp = tp_of(rlen1);
            double tmpq = (double)qs / qc;
            event(1, inspos, key, tp, tmpq, nm - nmoff, 2 | 8, 0, tmpq >= cf.goodq);
            int index = rpi - 1 - (start - 1 - inspos);
            if (inspos > position && ref.hasEq(inspos, sq(index))) {
                char ab = sq(index);
                int b = base_index(ab);
                const int qa = q(index);
                if (b >= 0 && !leader) {
                } else if (b >= 0) {
                    unsigned ai = atomicAdd(S.n_anchors, 1u);
                    if (ai < S.max_anchors) {
                        Anchor a;
                        a.pos = inspos;
                        a.ord = ord;
                        a.base = b;
                        a.dir = dir ? 1 : 0;
                        a.tp = tp;
                        a.q = qa;
                        a.mapq = mapq;
                        a.nmv = nm - nmoff;
                        S.anchors[ai] = a;
                    } else {
                        atomicAdd(&S.ctr->overflow, 1ull);
                    }
                } else if (ab != 0) {
                    DStr& d = walk_str[wave][2];
                    d.clear();
                    d.push(ab);
                    event(3, inspos, d, tp, (double)q(index), nm - nmoff, 0, 0, (double)q(index) >= cf.goodq);
                }
            }
            if (ci == 1 && (op(0) == OP_S || op(0) == OP_H) && leader) atomicAdd(&S.ctr->ttref, 1);
        }
        rpi += cel + offset + multoffp;
        rpe += cel + offset + multoffp;
        start += offset + multoffs;
        return ci;
    }

    __device__ __forceinline__ int deletion(int ci, int rlen1) {
        if ((n_cigar > ci && op(ci + 1) == OP_N) || (ci > 1 && op(ci - 1) == OP_N)) {
            rpi += cel;
            return ci;
        }
        DStr& desc = walk_str[wave][0];
        desc.clear();
        desc.push('-');
        desc.append_int(cel);
        int app_from = 0, app_len = 0;
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
                for (int vi = 0; vsn <= cf.vext && vi < len(ci + 3); vi++) {
                    if (sq(tn + vi) == 'N') break;
                    if (q(tn + vi) < cf.goodq) break;
                    if (ref.hasEq(ts + vi, 'N')) break;
                    if (ref.has(ts + vi)) {
                        if (sq(tn + vi) != ref.at(ts + vi)) {
                            offset = vi + 1;
                            nmoff++;
                            vsn = 0;
                        } else {
                            vsn++;
                        }
                    }
                }
                if (offset != 0) {
                    app_from = tn;
                    app_len = offset;
                    for (int i = tn; i < tn + offset; i++) qs += q(i);
                    qc += offset;
                }
            }
            ci += 2;
        } else if (isNextInsertion(ci)) {
            int insLen = len(ci + 1);
            desc.push('^');
            sub(desc, rpi, insLen);
            for (int i = rpi; i < rpi + insLen; i++) qs += q(i);
            qc += insLen;
            multoffp += insLen;
            ci += 1;
        } else if (isNextMatched(ci)) {
            int mLen = len(ci + 1);
            int vsn = 0;
            for (int vi = 0; vsn <= cf.vext && vi < mLen; vi++) {
                char sc = sq(rpi + vi);
                if (sc == 'N') break;
                if (q(rpi + vi) < cf.goodq) break;
                if (ref.has(start + cel + vi)) {
                    char rc = ref.at(start + cel + vi);
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
                app_from = rpi;
                app_len = offset;
                for (int i = rpi; i < rpi + offset; i++) qs += q(i);
                qc += offset;
            }
        }
        if (offset > 0) {
            desc.push('&');
            sub(desc, app_from, app_len);
        }
        if (rpi + offset >= lseq) {
            qs += qlast;
            qc += 1;
        } else {
            int qo = q(rpi + offset);
            qs += qlast > qo ? qlast : qo;
            qc += 1;
        }
        if (start >= Rs && start <= Re) {
            int tp = tp_of(rlen1);
            double tmpq = (double)qs / qc;
            event(0, start, desc, tp, tmpq, nm - nmoff, 1 | 8, 0, tmpq >= cf.goodq);
            for (int i = 0; i < cel; i++) cov(start + i);
        }
        start += cel + offset + multoffs;
        rpi += offset + multoffp;
        rpe += offset + multoffp;
        return ci;
    }

    // Bases i, i+1, ... of the current M run that the sequential steps below would each record as
    // a single-base allele (or skip, for N), one base per lane; returns how many were handled.
    // A base stops the run when it could start a multi-base allele (a high-quality mismatch with a
    // next base in the run and the region), when it is a high-quality mismatch within the last
    // `vext` bases before an adjacent D/I, or when it is not A/C/G/T/N.
    __device__ __forceinline__ int plain_chunk(int i, int ci, int totalLen, int rlen1, int nmoff) {
        const int span = min((int)warpSize, cel - i);
        const int k = lane;
        const bool act = k < span;
        const int rp = rpi + k, st = start + k, j = i + k, re1 = rpe + k;
        // Every lane loads (indices clamped into range) and the conditions combine without branches:
        // a branch per condition costs exec-mask bookkeeping and a wait per load on every lane.
        const bool inread = rp >= 0 && rp < lseq;
        const int rpc = inread ? rp : 0;
        const char ch = base_at(rpc);
        const int qv = qual_at(rpc);
        const int b = ch == 'A' ? 0 : ch == 'C' ? 1 : ch == 'G' ? 2 : ch == 'T' ? 3 : -1;
        const bool isN = inread && ch == 'N';
        const bool hasr = ref.has(st);
        const unsigned js = (unsigned)(st - ref.slo);
        char rc = walk_ref[js < (unsigned)ref.sn ? js : 0u];
        if (hasr && js >= (unsigned)ref.sn) rc = *(volatile const char*)(ref.ref + (st - ref.lo));
        const bool hq_mis = hasr && rc != ch && (double)qv >= cf.goodq;
        // the next operation is the same for every lane
        const bool next_indel =
            cf.performLocalRealignment && ci + 1 < n_cigar && (op(ci + 1) == OP_D || op(ci + 1) == OP_I);
        // a high-quality mismatch stops the run where it could start a multi-base allele (a next base
        // in the run and the region) or lies within vext of an adjacent indel
        const bool stop_mnv = (st + 1) >= Rs && (st + 1) <= Re && j + 1 < cel && rc != 'N';
        const bool stop_ind = next_indel && cel - j <= cf.vext;
        const bool stop = act && (!inread || (!isN && (b < 0 || (hq_mis && (stop_mnv || stop_ind)))));
        const unsigned long long stops = __ballot(stop);
        const int f = stops ? (__ffsll(stops) - 1) : span;
        bool doit = act && k < f && !isN;
        if (doit && cf.trimBasesAfter != 0) doit = !(!dir ? rp > cf.trimBasesAfter : totalLen - rp > cf.trimBasesAfter);
        doit = doit && st >= Rs && st <= Re;
        const int ri = st - S.dense_lo;
        const bool inwin = ri >= 0 && ri < S.dense_n;
        if (__ballot(doit && !inwin)) bad = true;  // as row() would flag it
        const unsigned long long md = __ballot(doit && inwin);
        if (doit && inwin) {
            const unsigned t = tseq + (unsigned)__popcll(md & ((1ull << k) - 1ull));
            const int tp = re1 < rlen1 - re1 ? re1 + 1 : rlen1 - re1;
            const int ti = tile_slot(st);
            if (ti >= 0) {  // key creation, counts and coverage in one tile branch
                atomicMin(&walk_tile.first_touch[ti], ((unsigned long long)(unsigned)ord << 24) | (t & 0xffffffu));
                tile_single(ti, b, tp, qv, nm - nmoff);
                atomicAdd(&walk_tile.cov[ti], 1);
            } else {
                DenseRow* r = &S.dense[ri];
                touch_key(r, st, t);
                dense_add(r, st, b, tp, qv, nm - nmoff);
                add_cov(r, st, 1);
            }
        }
        tseq += (unsigned)__popcll(md);
        return f;
    }

    // one M (or =, X, P ...) CIGAR element; returns the possibly advanced ci
    __device__ __forceinline__ int match_run(int ci, int c_op, int rlen1, int totalLen) {
        int nmoff = 0, moffset = 0;
        int i = offset;
        while (i < cel) {
            {
                const int f = plain_chunk(i, ci, totalLen, rlen1, nmoff);
                i += f;
                start += f;
                rpi += f;
                rpe += f;
                if (i >= cel) break;
            }
            bool trim = false;
            if (cf.trimBasesAfter != 0) trim = !dir ? rpi > cf.trimBasesAfter : totalLen - rpi > cf.trimBasesAfter;
            const char ch1 = sq(rpi);
            if (ch1 == 'N') {
                start++;
                rpi++;
                rpe++;
                i++;
                continue;
            }
            DStr& s = walk_str[wave][0];
            s.clear();
            s.push(ch1);
            bool swd = false;
            double qv = q(rpi);
            int qbases = 1, qibases = 0;
            int ss_n = 0;
            DStr& ss = walk_str[wave][1];
            ss.clear();
            while ((start + 1) >= Rs && (start + 1) <= Re && (i + 1) < cel && qv >= cf.goodq && ref.hasNe(start, sq(rpi)) &&
                   ref.at(start) != 'N') {
                if (q(rpi + 1) < cf.goodq + 5) break;
                char nuc = sq(rpi + 1);
                if (nuc == 'N') break;
                if (ref.hasEq(start + 1, 'N')) break;
                if (ref.get(start + 1) != nuc) {
                    ss.push(nuc);
                    qv += q(rpi + 1);
                    qbases++;
                    rpi++;
                    rpe++;
                    i++;
                    start++;
                    nmoff++;
                } else {
                    int ssn = 0;
                    for (int ssi = 1; ssi <= cf.vext; ssi++) {
                        if (i + 1 + ssi >= cel) break;
                        if (rpi + 1 + ssi < lseq && ref.hasNe(start + 1 + ssi, sq(rpi + 1 + ssi))) {
                            ssn = ssi + 1;
                            break;
                        }
                    }
                    if (ssn == 0) break;
                    if (q(rpi + ssn) < cf.goodq + 5) break;
                    for (int ssi = 1; ssi <= ssn; ssi++) {
                        ss.push(sq(rpi + ssi));
                        qv += q(rpi + ssi);
                        qbases++;
                    }
                    rpi += ssn;
                    rpe += ssn;
                    i += ssn;
                    start += ssn;
                }
            }
            ss_n = ss.n;
            if (ss_n > 0) {
                s.push('&');
                s.append(ss);
            }
            int ddlen = 0;
            if (cf.performLocalRealignment && cel - i <= cf.vext && ci + 1 < n_cigar && op(ci + 1) == OP_D && ref.has(start) &&
                (ss_n > 0 || sq(rpi) != ref.at(start)) && q(rpi) >= cf.goodq) {
                while (i + 1 < cel) {
                    s.push(sq(rpi + 1));
                    qv += q(rpi + 1);
                    qbases++;
                    i++;
                    rpi++;
                    rpe++;
                    start++;
                }
                int amp = s.find('&');
                if (amp >= 0) s.erase_at(amp);
                DStr& t = walk_str[wave][2];
                t.clear();
                t.push('-');
                t.append_int(len(ci + 1));
                t.push('&');
                t.append(s);
                copy_str(s, t);
                swd = true;
                ddlen = len(ci + 1);
                ci += 1;
                if (n_cigar > ci + 1 && op(ci + 1) == OP_I) {
                    s.push('^');
                    sub(s, rpi + 1, len(ci + 1));
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
                    Off o = findOffset(start + ddlen + 1, rpi + 1, len(ci + 1));
                    if (o.offset != 0) {
                        moffset = o.offset;
                        nmoff += o.nm;
                        s.push('&');
                        sub(s, o.from, o.offset);
                    }
                }
            } else if (cf.performLocalRealignment && cel - i <= cf.vext && ci + 1 < n_cigar && op(ci + 1) == OP_I &&
                       ref.has(start) && (ss_n > 0 || sq(rpi) != ref.at(start)) && q(rpi) >= cf.goodq) {
                while (i + 1 < cel) {
                    s.push(sq(rpi + 1));
                    qv += q(rpi + 1);
                    qbases++;
                    i++;
                    rpi++;
                    rpe++;
                    start++;
                }
                int amp = s.find('&');
                if (amp >= 0) s.erase_at(amp);
                int nextLen = len(ci + 1);
                sub(s, rpi + 1, nextLen);
                s.insert_at(nextLen, '&');
                s.insert_at(0, '+');
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
                if (pos >= Rs && pos <= Re) {
                    int tp = tp_of(rlen1);
                    qv = qv / (qbases + qibases);
                    // MNP table entry: ^[ATGC]&[ATGC]+$
                    bool mnp = false;
                    if (s.n > 2 && s.c[1] == '&' && base_index(s.c[0]) >= 0) {
                        mnp = true;
                        for (int k = 2; k < s.n; k++)
                            if (base_index(s.c[k]) < 0) mnp = false;
                    }
                    var_contrib(pos, s, tp, qv, nm - nmoff, (swd ? 1 : 0) | (mnp ? 4 : 0));
                    for (int qi = 1; qi <= qbases; qi++) cov(start - qi + 1);
                    if (swd)
                        for (int qi = 1; qi < ddlen; qi++) cov(start + qi);
                }
            }
            if (swd) start += ddlen;
            if (c_op != OP_I) start++;
            if (c_op != OP_D) {
                rpi++;
                rpe++;
            }
            i++;
        }
        if (moffset != 0) {
            offset = moffset;
            rpi += moffset;
            start += moffset;
            rpe += moffset;
        }
        return ci;
    }

    __device__ __forceinline__ bool skip_overlap(uint16_t flag, int tid, int mtid, int mateStart, int reflen, int pos0) const {
        bool pairedSame = (flag & 1) && tid == mtid;
        if (cf.uniqueModeAlignmentEnabled && pairedSame && !dir && start >= mateStart) return true;
        if (cf.uniqueModeSecondInPairEnabled && (flag & 1) && pairedSame) {
            int endpos = pos0 + (reflen > 0 ? reflen : 1);
            if (position >= mateStart) return start >= mateStart && start <= mateStart + reflen - 1;
            return start >= mateStart && mateStart <= endpos;
        }
        return false;
    }

    __device__ __forceinline__ void walk(int rlen1, int totalLen, uint16_t flag, int tid, int mtid, int mateStart, int reflen, int pos0) {
        start = position;
        offset = 0;
        rpi = rpe = 0;
        bool need_break = true;
        for (int ci = 0; ci < n_cigar; ++ci) {
            if (bad) break;
            if (need_break && skip_overlap(flag, tid, mtid, mateStart, reflen, pos0)) break;
            need_break = false;
            int c_op = op(ci);
            cel = len(ci);
            if ((ci == 0 || ci == n_cigar - 1) && c_op == OP_I) c_op = OP_S;
            if (c_op == OP_N) {
                start += cel;
                offset = 0;
                continue;
            }
            if (c_op == OP_S) {
                softclip(ci, totalLen);
                continue;
            }
            if (c_op == OP_H) {
                offset = 0;
                continue;
            }
            if (c_op == OP_I) {
                offset = 0;
                ci = insertion(ci, rlen1);
                continue;
            }
            if (c_op == OP_D) {
                offset = 0;
                ci = deletion(ci, rlen1);
                continue;
            }
            ci = match_run(ci, c_op, rlen1, totalLen);
            if (start > Re) break;
        }
        if (bad && leader) atomicAdd(&S.ctr->walk_errors, 1ull);
    }
};

}  // namespace rv
