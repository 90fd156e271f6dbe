// Per-item bodies of the pileup kernels (one read, one dense row, one anchor).  The kernels in
// rv_pileup.hip map one thread to one item.
//
// Reference stages: src/recordPreprocessor.cpp:101-180 (filters), src/parseCigar.cpp:497-628
// (per-read setup, NM filter, modifyCigar, cleanupCigar, drop rules), :629-972 (walk),
// :1471-1490 (insertion-anchor subtraction).
#pragma once
#include "rv_walk.h"

namespace rv {

// Read batch in device memory, structure of arrays (the region iterator's records in order)
struct ReadsDev {
    const int32_t* pos0;
    const uint16_t* flag;
    const uint8_t* mapq;
    const uint8_t* has_nm;
    const int32_t* nm;
    const int32_t* mpos0;
    const int32_t* tid;
    const int32_t* mtid;
    const int64_t* seq_off;
    const int64_t* cig_off;
    const uint8_t* bases;
    const uint8_t* quals;
    const uint32_t* cigar;
    int64_t n;
    const uint8_t* drop;  // -Z: per read, 1 = dropped by the downsampling draw (nullptr without -Z)
};

// Everything the prep and walk kernels read besides the per-read records, kept in device memory
// so that uniform fields are read with scalar loads rather than copied to private memory.
struct WalkArgs {
    ReadsDev R;
    Window W;
    Config cfg;
    WalkSink S;
    uint8_t* dup;  // -t: per read, 1 = dropped as a duplicate (nullptr without -t)
};

// ---------------------------------------------------------------- duplicate removal (-t)
// recordPreprocessor.cpp:153-176.  The reference keeps a set of keys, clears it whenever a record's
// alignment start differs from the start of the key it inserted last, and drops a read whose key
// is already in the set.  Keys exist only for reads whose mate starts before base 10 (start, mate
// contig name, mate start) and for paired reads flagged unmapped (start, CIGAR text).  Over the
// records in iterator order, with the ones the flag/MAPQ/length filters dropped invisible to it,
// that is: a read is a duplicate when an earlier read with the same key lies in the same run of
// consecutive records that share its alignment start.
RV_HD bool rec_filtered(const ReadsDev& R, const Config& c, int64_t i) {
    const int lseq = (int)(R.seq_off[i + 1] - R.seq_off[i]);
    return (R.flag[i] & c.samfilter) != 0 || (R.drop && R.drop[i]) || (int)R.mapq[i] < c.mappingQuality ||
           lseq == 1;
}
// 0: no key; 1: keyed by the mate; 2: keyed by the CIGAR
RV_HD int dup_kind(const ReadsDev& R, int64_t i) {
    if (R.mpos0[i] + 1 < 10) return 1;
    const uint16_t f = R.flag[i];
    return ((f & 1) && (f & 4)) ? 2 : 0;
}
// getMateReferenceName (recordPreprocessor.cpp:80-90): "*" unpaired, "=" same contig, else the
// mate contig's name (one name per contig id)
RV_HD int mate_name(const ReadsDev& R, int64_t i) {
    if (!(R.flag[i] & 1)) return -3;
    if (R.tid[i] == R.mtid[i]) return -2;
    return R.mtid[i];
}
RV_HD bool same_dup_key(const ReadsDev& R, int64_t i, int64_t j, int kind) {
    if (kind == 1) return R.mpos0[i] == R.mpos0[j] && mate_name(R, i) == mate_name(R, j);
    const int64_t a = R.cig_off[i], b = R.cig_off[j];
    const int64_t na = R.cig_off[i + 1] - a;
    if (na != R.cig_off[j + 1] - b) return false;
    for (int64_t k = 0; k < na; k++) {
        const uint32_t x = R.cigar[a + k], y = R.cigar[b + k];
        if ((x >> 4) != (y >> 4)) return false;
        // bam_cigar_opchr prints every op code past 'B' (9) as '?'
        const uint32_t ox = (x & 15u) > 9u ? 10u : (x & 15u), oy = (y & 15u) > 9u ? 10u : (y & 15u);
        if (ox != oy) return false;
    }
    return true;
}
RV_HD bool is_duplicate(const ReadsDev& R, const Config& c, int64_t i) {
    if (rec_filtered(R, c, i)) return false;
    const int kind = dup_kind(R, i);
    if (kind == 0) return false;
    const int32_t s = R.pos0[i];
    for (int64_t j = i - 1; j >= 0; j--) {
        if (rec_filtered(R, c, j)) continue;
        if (R.pos0[j] != s) return false;
        if (dup_kind(R, j) == kind && same_dup_key(R, i, j, kind)) return true;
    }
    return false;
}

__device__ __forceinline__ uint32_t packed_seq_word(const uint8_t* seq, int lseq, int w) {
    uint32_t v = 0;
    for (int b = 0; b < 4; b++) {
        int byte = 4 * w + b;
        int i0 = 2 * byte, i1 = 2 * byte + 1;
        uint32_t hi = i0 < lseq ? (uint32_t)nt16_code((char)seq[i0]) : 0u;
        uint32_t lo = i1 < lseq ? (uint32_t)nt16_code((char)seq[i1]) : 0u;
        if (i0 < lseq) v |= ((hi << 4) | lo) << (8 * b);
    }
    return v;
}

__device__ inline void prep_one(int64_t r, const ReadsDev& R, const Window& W, const Config& cfg, PrepRec* prep,
                                uint32_t* cig, WalkCounters* ctr, int* max_span, const uint8_t* dup) {
    PrepRec p;
    p.status = RD_PARSE;
    p.flags = 0;
    p.nm_eff = 0;
    p.rlen1 = 0;
    p.total_len = 0;
    p.reflen = 0;
    p.pad = 0;
    const uint16_t flag = R.flag[r];
    const int64_t so = R.seq_off[r];
    const int lseq = (int)(R.seq_off[r + 1] - so);
    const int64_t co = R.cig_off[r];
    const int ncig = (int)(R.cig_off[r + 1] - co);
    p.lseq = lseq;
    p.position = R.pos0[r] + 1;
    p.ncig = 0;
    uint32_t* rec = cig + (size_t)r * MAXC;
    // reach of the read around its alignment: used to size the dense window
    {
        int reflen = 0;
        for (int i = 0; i < ncig; i++) {
            int o = cop(R.cigar[co + i]);
            if (o == OP_M || o == OP_D || o == OP_N || o == OP_EQ || o == OP_X) reflen += clen(R.cigar[co + i]);
        }
        atomicMax(max_span, lseq + reflen + 8);
        p.reflen = reflen;
    }
    if ((flag & cfg.samfilter) != 0 || (R.drop && R.drop[r]) || (int)R.mapq[r] < cfg.mappingQuality || lseq == 1 ||
        (dup && dup[r])) {
        p.status = RD_FILTERED;
        prep[r] = p;
        return;
    }
    if (ncig <= 0) {
        p.status = RD_NOWALK;
        prep[r] = p;
        return;
    }
    int idlen = 0;
    for (int i = 0; i < ncig; i++) {
        int o = cop(R.cigar[co + i]);
        if (o == OP_I || o == OP_D) idlen += clen(R.cigar[co + i]);
    }
    int tnm = 0;
    if (R.has_nm[r]) {
        tnm = R.nm[r] - idlen;
        if (tnm > cfg.mismatch) {
            p.status = RD_NM;
            prep[r] = p;
            return;
        }
    } else if (flag & 4) {
        p.status = RD_NM;
        prep[r] = p;
        return;
    }
    p.nm_eff = tnm;
    // emulated record: CIGAR words, then the packed SEQ bytes as little-endian words
    const uint8_t* seq = R.bases + so;
    // a CIGAR longer than the record slot is refused before anything is written: the region then
    // fails with a capacity error (the host checks the counter) instead of walking a cut record
    if (ncig > MAXC_IN) {
        p.status = RD_NOWALK;
        p.flags |= 2;
        atomicAdd(&ctr->capacity, 1ull);
        prep[r] = p;
        return;
    }
    int seq_words = ((lseq + 1) / 2 + 3) / 4;
    int nrec = ncig + seq_words;
    if (nrec > MAXC) nrec = MAXC;
    for (int i = 0; i < nrec; i++) rec[i] = i < ncig ? R.cigar[co + i] : packed_seq_word(seq, lseq, i - ncig);
    for (int i = nrec; i < MAXC; i++) rec[i] = 0;
    int n_cigar = ncig;
    int position = R.pos0[r] + 1;
    if (cfg.performLocalRealignment) {
        RefView rf{W.ref, W.ref_lo, W.ref_hi};
        ReadView rd{(const char*)seq, R.quals + so, lseq};
        int n_new = 0, fl = 0;
        if (modify_cigar_record(rf, rd, position, rec, nrec, ncig, idlen, n_new, fl)) {
            n_cigar = n_new;
            p.flags |= 1;
        }
        p.flags |= fl;
        if (fl & 2) {  // rewrite text or words beyond capacity: refused as above
            p.status = RD_NOWALK;
            atomicAdd(&ctr->capacity, 1ull);
            prep[r] = p;
            return;
        }
    }
    // cleanupCigar: leading/trailing I become S until the first M/=/X
    {
        bool none = true;
        for (int i = 0; i < n_cigar && none; i++) {
            int o = cop(rec[i]);
            if (o == OP_I)
                rec[i] = (rec[i] & ~0xfu) | OP_S;
            else if (o == OP_M || o == OP_EQ || o == OP_X)
                none = false;
        }
        none = true;
        for (int i = n_cigar - 1; i >= 0 && none; i--) {
            int o = cop(rec[i]);
            if (o == OP_I)
                rec[i] = (rec[i] & ~0xfu) | OP_S;
            else if (o == OP_M || o == OP_EQ || o == OP_X)
                none = false;
        }
    }
    p.position = position;
    p.ncig = n_cigar;
    if (n_cigar <= 0) {
        p.status = RD_NOWALK;
        prep[r] = p;
        return;
    }
    if (cop(rec[0]) == OP_S && clen(rec[0]) >= 10 && cop(rec[n_cigar - 1]) == OP_S && clen(rec[n_cigar - 1]) >= 10) {
        p.status = RD_NOWALK;
        prep[r] = p;
        return;
    }
    int rlen1 = 0, tot = 0;
    for (int i = 0; i < n_cigar; i++) {
        int o = cop(rec[i]), l = clen(rec[i]);
        if (o == OP_M || o == OP_I) rlen1 += l;
        if (o == OP_M || o == OP_I || o == OP_S) tot += l;
    }
    p.rlen1 = rlen1;
    p.total_len = tot;
    if (cfg.minmatch != 0 && rlen1 < cfg.minmatch) {
        p.status = RD_NOWALK;
        prep[r] = p;
        return;
    }
    atomicMax(&ctr->max_read_length, tot);
    if (flag & 2048) {
        p.status = RD_NOWALK;
        prep[r] = p;
        return;
    }
    prep[r] = p;
}

__device__ inline void fill_row_one(DenseRow* rows, int i) {
    DenseRow z;
    for (int b = 0; b < 4; b++) {
        z.b[b].cnt = z.b[b].fwd = z.b[b].rev = z.b[b].hi = z.b[b].lo = 0;
        z.b[b].sum_tp = z.b[b].sum_q = z.b[b].sum_mq = z.b[b].sum_nm = 0;
        z.b[b].first_tpq = 0;
        z.b[b].flags = 0;
        z.b[b].min_ord = 0xffffffffu;
        z.b[b].last_tp = 0;
        z.b[b].pad[0] = z.b[b].pad[1] = z.b[b].pad[2] = 0;
    }
    z.cov = 0;
    z.rowflags = 0;
    z.first_touch = ~0ull;
    for (int k = 0; k < 4; k++) z.pad[k] = 0;
    rows[i] = z;
}

// One read, walked by a whole wavefront (lane = this thread's lane, wave = its wavefront's index in
// the workgroup; see rv_walk.h).  Contributions to positions [tile_lo, tile_lo + TILE_POS) go to the
// workgroup's tile.
__device__ inline void walk_one(int64_t r, const ReadsDev& R, const Window& W, const Config& cfg, const PrepRec* prep,
                                const uint32_t* cig, const WalkSink& sink, int lane, int wave, int tile_lo) {
    const PrepRec p = prep[r];
    if (p.status != RD_PARSE) return;
    // every per-read field up front, so that the loads are in flight together
    const int64_t so = R.seq_off[r];
    const uint8_t mapq = R.mapq[r];
    const uint16_t flag = R.flag[r];
    const int tid = R.tid[r], mtid = R.mtid[r], mpos0 = R.mpos0[r], pos0 = R.pos0[r];
    const char* gseq = (const char*)(R.bases + so);
    const uint8_t* gqual = R.quals + so;
    const uint32_t* rec = cig + (size_t)r * MAXC;
    WaveStage* st = &walk_stage[wave];
    const bool staged = p.lseq <= STAGE_LEN;
    // the previous read's steps are done with the copy before it is overwritten, and the copy is
    // complete before any lane reads it (LDS operations of one wavefront complete in order)
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int i = lane; i < MAXC; i += warpSize) st->cig[i] = rec[i];
    if (staged) {
        for (int i0 = 0; i0 < p.lseq; i0 += 8 * warpSize) {
            char b[8];
            uint8_t q[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int i = i0 + k * warpSize + lane;
                if (i < p.lseq) {
                    b[k] = gseq[i];
                    q[k] = gqual[i];
                }
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int i = i0 + k * warpSize + lane;
                if (i < p.lseq) {
                    st->seq[i] = b[k];
                    st->qual[i] = q[k];
                }
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    Walker w;
    w.lane = lane;
    w.leader = lane == 0;
    w.use_tile = true;
    w.tile_lo = tile_lo;
    w.ref.leader = lane == 0;
    w.cf = cfg;
    w.Rs = W.region_start;
    w.Re = W.region_end;
    w.chrLen = W.chr_len;
    w.ref.ref = W.ref;
    w.ref.lo = W.ref_lo;
    w.ref.hi = W.ref_hi;
    w.ref.ctr = sink.ctr;
    w.ref.slo = tile_lo - REF_BEFORE;
    w.ref.sn = REF_STAGE;
    w.S = sink;
    w.seq = gseq;
    w.qual = gqual;
    w.wave = wave;
    w.staged = staged;
    w.lseq = p.lseq;
    {
        int nb = (p.lseq + 1) / 2;
        int last = nb - 1;
        int i0 = 2 * last, i1 = 2 * last + 1;
        int hi = i0 < p.lseq ? nt16_code(w.base_at(i0)) : 0;
        int lo = i1 < p.lseq ? nt16_code(w.base_at(i1)) : 0;
        w.qual_before0 = nb > 0 ? (uint8_t)((hi << 4) | lo) : 0;
    }
    w.rec_n = MAXC;
    w.n_cigar = p.ncig;
    w.ord = (int)r;
    w.evseq = 0;
    w.tseq = 0;
    w.mapq = mapq;
    w.nm = p.nm_eff;
    w.dir = (flag & 0x10) != 0;
    w.position = p.position;
    w.bad = false;
    w.walk(p.rlen1, p.total_len, flag, tid, mtid, mpos0 + 1, p.reflen, pos0);
}

// noPassingReads (reference src/VariationRealigner.cpp:1447-1498), one record: true when the record
// overlaps [start - 1, end) (0-based, as the region query does), carries no CIGAR deletion of
// exactly end - start bases, and its M + D span reaches more than two bases past both ends.
__device__ inline bool passes_one(const ReadsDev& R, int64_t i, int start, int end) {
    const int beg0 = R.pos0[i];
    if (beg0 >= end) return false;
    const uint32_t dq = ((uint32_t)(end - start) << 4) | (uint32_t)OP_D;
    int reflen_all = 0, al = 0;
    bool has = false;
    for (int64_t k = R.cig_off[i], k1 = R.cig_off[i + 1]; k < k1; k++) {
        const uint32_t w = R.cigar[k];
        const int o = (int)(w & 0xFu), l = (int)(w >> 4);
        if (o == OP_M || o == OP_D || o == OP_N || o == OP_EQ || o == OP_X) reflen_all += l;
        if (o == OP_M || o == OP_D) al += l;
        if (w == dq) has = true;
    }
    const int end0 = beg0 + (reflen_all > 0 ? reflen_all : 1);
    if (end0 <= start - 1 || has) return false;
    const int rs = beg0 + 1, re = rs + al;
    return re > end + 2 && rs < start - 2;
}

// Tile set-up and flush, strided over the workgroup's threads.
__device__ inline void tile_init(int tid, int nthreads, const Window& W, int tile_lo) {
    const int slo = tile_lo - REF_BEFORE;
    for (int k = tid; k < REF_STAGE; k += nthreads) {
        const int pos = slo + k;
        walk_ref[k] = (pos >= W.ref_lo && pos <= W.ref_hi) ? W.ref[pos - W.ref_lo] : '\0';
    }
    for (int k = tid; k < TF_N * 4 * TILE_POS; k += nthreads) {
        const int f = k / (4 * TILE_POS), r = k - f * 4 * TILE_POS;
        walk_tile.f[f][r / TILE_POS][r % TILE_POS] = (f == TF_MINORD || f == TF_AND) ? ~0u : 0u;
    }
    for (int k = tid; k < TILE_POS; k += nthreads) {
        walk_tile.cov[k] = 0;
        walk_tile.first_touch[k] = ~0ull;
    }
}

// Adds the tile into the dense rows.  Pairs of adjacent 32-bit row fields are added with one 64-bit
// atomic (the lower field of each pair never goes negative during the walk, so no carry crosses
// into the upper one; the upper field wraps as its own 32-bit value).  For the pstd/qstd state: the
// tile's AND/OR of its (position, quality) values covers pairs inside the workgroup, and comparing
// one tile value with the row's first value covers pairs across workgroups.  The tile value is the
// AND: exact in each field where the tile's values are all equal, and where they are not, that
// field's flag is already set, so a comparison against it can only set a flag that is set anyway.
__device__ inline void tile_flush(int tile_lo, DenseRow* rows, int dense_lo, int dense_n, int tid, int nthreads) {
    const TileAcc* t = &walk_tile;
    for (int k = tid; k < 4 * TILE_POS; k += nthreads) {
        const int b = k / TILE_POS, i = k - b * TILE_POS;
        const int ri = tile_lo + i - dense_lo;
        if (ri < 0 || ri >= dense_n) continue;
        DenseRow& r = rows[ri];
        if (b == 0) {
            if (t->cov[i]) atomicAdd(&r.cov, t->cov[i]);
            if (t->first_touch[i] != ~0ull) {
                atomicOr(&r.rowflags, 1u);
                atomicMin(&r.first_touch, t->first_touch[i]);
            }
        }
        const unsigned fl = t->f[TF_FLAGS][b][i];
        if (fl == 0) continue;  // no contribution to this base
        DenseBase& d = r.b[b];
        const unsigned cw = t->f[TF_CNT][b][i];
        const unsigned fwd = cw & 0x3ffu, rev = (cw >> 10) & 0x3ffu, hi = (cw >> 20) & 0x3ffu;
        const unsigned cnt = fwd + rev, lo = cnt - hi;
        const unsigned qmq = t->f[TF_QMQ][b][i];
        auto pair = [](unsigned lo32, unsigned hi32) { return (unsigned long long)lo32 | ((unsigned long long)hi32 << 32); };
        atomicAdd((unsigned long long*)&d.cnt, pair(cnt, fwd));
        atomicAdd((unsigned long long*)&d.rev, pair(rev, hi));
        atomicAdd((unsigned long long*)&d.lo, pair(lo, t->f[TF_TP][b][i]));
        atomicAdd((unsigned long long*)&d.sum_q, pair(qmq & 0xffffu, qmq >> 16));
        if (t->f[TF_NM][b][i]) atomicAdd((unsigned*)&d.sum_nm, t->f[TF_NM][b][i]);
        atomicMin(&d.min_ord, t->f[TF_MINORD][b][i]);
        unsigned add = fl & 9u;
        if (fl & TF_HASPQ) {
            const unsigned va = t->f[TF_AND][b][i], diff = va ^ t->f[TF_OR][b][i];
            if (diff >> 8) add |= 2u;
            if (diff & 0xffu) add |= 4u;
            const unsigned first = va + 1u;
            const unsigned old = atomicCAS(&d.first_tpq, 0u, first);
            if (old != 0u && old != first) {
                const unsigned o = old - 1u;
                if ((o >> 8) != (va >> 8)) add |= 2u;
                if ((o & 0xffu) != (va & 0xffu)) add |= 4u;
            }
        }
        atomicOr(&d.flags, add);
    }
}

// Insertion-anchor subtraction (parseCigar.cpp:1471-1490): the reference subtracts from the anchor
// base's entry only when that entry exists at that moment, i.e. when some read with an ordinal <=
// the inserting read created it (a creation by the same read always precedes its insertion).
__device__ inline void anchor_one(unsigned i, const Anchor* A, DenseRow* rows, int dense_lo, int dense_n, const Config& cfg) {
    Anchor a = A[i];
    int k = a.pos - dense_lo;
    if (k < 0 || k >= dense_n) return;
    DenseBase& d = rows[k].b[a.base];
    if (d.min_ord > (unsigned)a.ord) return;
    atomicSub(&d.cnt, 1);
    atomicSub(a.dir ? &d.rev : &d.fwd, 1);
    atomicSub(&d.sum_tp, a.tp);
    atomicSub(&d.sum_q, a.q);
    atomicSub(&d.sum_mq, a.mapq);
    atomicSub(&d.sum_nm, a.nmv);
    atomicSub((double)a.q >= cfg.goodq ? &d.hi : &d.lo, 1);
}


}  // namespace rv
