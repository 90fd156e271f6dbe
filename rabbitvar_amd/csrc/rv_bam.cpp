// BAM decoding to the structure-of-arrays records the engine takes (rv_reads).
//
// The reference reads records one at a time through htslib: sam_itr_querys on "chr:start-end" and
// sam_itr_next (src/recordPreprocessor.cpp:27, :103), then decodes bases with seq_nt16_str
// (src/parseCigar.cpp:442-459).  htslib is not vendored by the reference; this follows the public
// SAM/BAM specification (BGZF blocks: gzip members with a "BC" extra field carrying the block size;
// BAM records: block_size, refID, pos, l_read_name, mapq, bin, n_cigar_op, flag, l_seq, next_refID,
// next_pos, tlen, read_name, cigar, 4-bit seq, qual, aux) and htslib's region semantics: a query
// returns, in file order, the records of the contig that start before the region's end and whose
// end (bam_endpos: start + reference length of the CIGAR, or start + 1 when that is 0 or the read
// is unmapped) lies after its start.
//
// The file is inflated once at open, its BGZF blocks in parallel (each block's ISIZE footer gives
// its place in the output), and the records are indexed; a query then converts the selected
// records to SoA on several threads.
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "rabbitvar_amd.h"
#include "rv_bam.h"

namespace {

uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
int32_t rd32s(const uint8_t* p) { return (int32_t)rd32(p); }

struct Block {
    size_t in_off, in_len;  // deflate payload
    size_t out_off, out_len;
    uint32_t crc;
};

const char NT16[] = "=ACMGRSVTWYHKDBN";

}  // namespace

struct rv_bam {
    std::string err;
    std::vector<uint8_t> data;  // the inflated stream
    std::vector<std::string> names;
    std::vector<int32_t> lens;
    std::vector<size_t> rec;  // offset of each record's block_size field
    std::vector<int32_t> rec_tid, rec_pos, rec_end;
    bool sorted = false;     // coordinate order: records of a contig contiguous, starts non-decreasing
    int32_t max_span = 0;    // longest record (end - start)
    int threads = 1;
    double inflate_ms = 0, index_ms = 0;
    uint64_t bytes_in = 0;
    rv::BamQuery last;  // the arrays of the last rv_bam_query
};

namespace {

rv_status fail(rv_bam* b, const std::string& m) {
    b->err = m;
    return RV_ERR_ARG;
}

template <class F>
void parallel_for(size_t n, int threads, F&& f) {
    const int nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(threads, 1), n));
    if (nt == 1) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&] {
            for (;;) {
                const size_t i = next.fetch_add(64);
                if (i >= n) break;
                const size_t e = std::min(n, i + 64);
                for (size_t k = i; k < e; k++) f(k);
            }
        });
    for (auto& th : pool) th.join();
}

rv_status load(rv_bam* b, const char* path) {
    FILE* fp = fopen(path, "rb");
    if (!fp) return fail(b, std::string("cannot open ") + path);
    std::vector<uint8_t> raw;
    if (fseek(fp, 0, SEEK_END) == 0) {
        const long sz = ftell(fp);
        if (sz > 0) raw.resize((size_t)sz);
        fseek(fp, 0, SEEK_SET);
    }
    const size_t got = raw.empty() ? 0 : fread(raw.data(), 1, raw.size(), fp);
    fclose(fp);
    if (got != raw.size()) return fail(b, "short read");
    b->bytes_in = raw.size();
    // ---- BGZF blocks
    std::vector<Block> blocks;
    size_t off = 0, out = 0;
    while (off < raw.size()) {
        if (raw.size() - off < 26) return fail(b, "truncated BGZF block header");
        const uint8_t* h = raw.data() + off;
        if (h[0] != 31 || h[1] != 139 || h[2] != 8 || !(h[3] & 4)) return fail(b, "not a BGZF block");
        const uint16_t xlen = rd16(h + 10);
        size_t bsize = 0;
        for (size_t x = 0; x + 4 <= xlen;) {
            const uint8_t* sf = h + 12 + x;
            const uint16_t slen = rd16(sf + 2);
            if (sf[0] == 'B' && sf[1] == 'C' && slen == 2) bsize = (size_t)rd16(sf + 4) + 1;
            x += 4 + slen;
        }
        if (bsize == 0 || off + bsize > raw.size() || bsize < (size_t)(12 + xlen + 8))
            return fail(b, "bad BGZF block size");
        Block bl;
        bl.in_off = off + 12 + xlen;
        bl.in_len = bsize - (12 + xlen) - 8;
        bl.crc = rd32(raw.data() + off + bsize - 8);
        bl.out_len = rd32(raw.data() + off + bsize - 4);
        bl.out_off = out;
        out += bl.out_len;
        blocks.push_back(bl);
        off += bsize;
    }
    b->data.resize(out);
    std::atomic<int> bad{0};
    parallel_for(blocks.size(), b->threads, [&](size_t i) {
        const Block& bl = blocks[i];
        if (bl.out_len == 0) return;
        z_stream zs;
        memset(&zs, 0, sizeof zs);
        if (inflateInit2(&zs, -15) != Z_OK) {
            bad = 1;
            return;
        }
        zs.next_in = const_cast<Bytef*>(raw.data() + bl.in_off);
        zs.avail_in = (uInt)bl.in_len;
        zs.next_out = b->data.data() + bl.out_off;
        zs.avail_out = (uInt)bl.out_len;
        const int r = inflate(&zs, Z_FINISH);
        inflateEnd(&zs);
        if (r != Z_STREAM_END || zs.avail_out != 0 ||
            crc32(0L, b->data.data() + bl.out_off, (uInt)bl.out_len) != bl.crc)
            bad = 1;
    });
    if (bad) return fail(b, "corrupt BGZF block");
    // ---- header
    const uint8_t* d = b->data.data();
    const size_t n = b->data.size();
    if (n < 12 || memcmp(d, "BAM\1", 4) != 0) return fail(b, "not a BAM file");
    size_t p = 4;
    const uint32_t l_text = rd32(d + p);
    p += 4;
    if (p + l_text + 4 > n) return fail(b, "truncated header");
    p += l_text;
    const int32_t n_ref = rd32s(d + p);
    p += 4;
    if (n_ref < 0) return fail(b, "bad reference count");
    for (int32_t r = 0; r < n_ref; r++) {
        if (p + 4 > n) return fail(b, "truncated reference list");
        const uint32_t l_name = rd32(d + p);
        p += 4;
        if (l_name == 0 || p + l_name + 4 > n) return fail(b, "truncated reference list");
        b->names.emplace_back((const char*)d + p, strnlen((const char*)d + p, l_name));
        p += l_name;
        b->lens.push_back(rd32s(d + p));
        p += 4;
    }
    // ---- records
    while (p < n) {
        if (p + 4 > n) return fail(b, "truncated record");
        const uint32_t bs = rd32(d + p);
        if (bs < 32 || p + 4 + bs > n) return fail(b, "bad record size");
        const uint8_t* r = d + p + 4;
        const uint8_t l_read_name = r[8];
        const uint16_t n_cigar = rd16(r + 12);
        const int32_t l_seq = rd32s(r + 16);
        if (l_seq < 0 || 32 + (size_t)l_read_name + 4 * (size_t)n_cigar + (size_t)(l_seq + 1) / 2 + (size_t)l_seq > bs)
            return fail(b, "record fields exceed its size");
        b->rec.push_back(p);
        p += 4 + bs;
    }
    // start, contig and end (bam_endpos) of every record, for the region queries
    const size_t nr = b->rec.size();
    b->rec_tid.resize(nr);
    b->rec_pos.resize(nr);
    b->rec_end.resize(nr);
    parallel_for(nr, b->threads, [&](size_t i) {
        const uint8_t* r = d + b->rec[i] + 4;
        const int32_t pos = rd32s(r + 4);
        const uint16_t n_cigar = rd16(r + 12), flg = rd16(r + 14);
        const uint8_t* cg = r + 32 + r[8];
        int64_t rlen = 0;
        if (!(flg & 4))
            for (uint16_t k = 0; k < n_cigar; k++) {
                const uint32_t w = rd32(cg + 4 * k);
                const int op = (int)(w & 15);
                if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rlen += w >> 4;
            }
        if (rlen == 0) rlen = 1;
        b->rec_tid[i] = rd32s(r);
        b->rec_pos[i] = pos;
        b->rec_end[i] = (int32_t)std::min<int64_t>(INT32_MAX, pos + rlen);
    });
    b->sorted = true;
    std::vector<char> seen(b->names.size() + 1, 0);
    for (size_t i = 0; i < nr; i++) {
        b->max_span = std::max(b->max_span, b->rec_end[i] - b->rec_pos[i]);
        if (i == 0 || b->rec_tid[i] != b->rec_tid[i - 1]) {
            const int32_t t = b->rec_tid[i];
            const size_t slot = t < 0 || (size_t)t >= b->names.size() ? b->names.size() : (size_t)t;
            if (seen[slot]) b->sorted = false;
            seen[slot] = 1;
        } else if (b->rec_pos[i] < b->rec_pos[i - 1]) {
            b->sorted = false;
        }
    }
    return RV_OK;
}

// bam_aux2i of the NM tag, when present
bool aux_nm(const uint8_t* a, const uint8_t* end, int32_t& v) {
    while (a + 3 <= end) {
        const char t0 = (char)a[0], t1 = (char)a[1], ty = (char)a[2];
        const uint8_t* val = a + 3;
        size_t sz;
        switch (ty) {
            case 'A': case 'c': case 'C': sz = 1; break;
            case 's': case 'S': sz = 2; break;
            case 'i': case 'I': case 'f': sz = 4; break;
            case 'Z': case 'H': {
                const uint8_t* q = val;
                while (q < end && *q) q++;
                sz = (size_t)(q - val) + 1;
                break;
            }
            case 'B': {
                if (val + 5 > end) return false;
                const char sub = (char)val[0];
                const uint32_t cnt = rd32(val + 1);
                size_t es = (sub == 'c' || sub == 'C') ? 1 : (sub == 's' || sub == 'S') ? 2 : 4;
                sz = 5 + es * cnt;
                break;
            }
            default: return false;
        }
        if (val + sz > end) return false;
        if (t0 == 'N' && t1 == 'M') {
            switch (ty) {
                case 'c': v = (int8_t)val[0]; return true;
                case 'C': v = val[0]; return true;
                case 's': v = (int16_t)rd16(val); return true;
                case 'S': v = rd16(val); return true;
                case 'i': v = rd32s(val); return true;
                case 'I': v = (int32_t)rd32(val); return true;
                default: v = 0; return true;  // bam_aux2i gives 0 for a non-integer type
            }
        }
        a = val + sz;
    }
    return false;
}

}  // namespace

extern "C" {

rv_status rv_bam_open(const char* path, int threads, rv_bam** out) {
    if (!path || !out) return RV_ERR_ARG;
    *out = nullptr;
    rv_bam* b = new (std::nothrow) rv_bam;
    if (!b) return RV_ERR_OOM;
    b->threads = threads > 0 ? threads : 1;
    rv_status s;
    try {
        s = load(b, path);
    } catch (const std::bad_alloc&) {
        s = RV_ERR_OOM;
        b->err = "out of memory";
    }
    if (s != RV_OK) {
        *out = b;  // the caller reads rv_bam_last_error, then closes
        return s;
    }
    *out = b;
    return RV_OK;
}

const char* rv_bam_last_error(const rv_bam* b) { return b ? b->err.c_str() : ""; }

int32_t rv_bam_n_targets(const rv_bam* b) { return b ? (int32_t)b->names.size() : 0; }

const char* rv_bam_target_name(const rv_bam* b, int32_t tid) {
    if (!b || tid < 0 || tid >= (int32_t)b->names.size()) return nullptr;
    return b->names[(size_t)tid].c_str();
}

int32_t rv_bam_target_len(const rv_bam* b, int32_t tid) {
    if (!b || tid < 0 || tid >= (int32_t)b->lens.size()) return -1;
    return b->lens[(size_t)tid];
}

size_t rv_bam_n_records(const rv_bam* b) { return b ? b->rec.size() : 0; }

rv_status rv_bam_query(rv_bam* b, const char* chr, int32_t start, int32_t end, rv_reads* out) {
    if (!b || !chr || !out) return RV_ERR_ARG;
    std::string err;
    const rv_status s = rv::bam_query_into(b, chr, start, end, b->last, out, err);
    if (s != RV_OK) b->err = err;
    return s;
}

void rv_bam_close(rv_bam* b) { delete b; }

}  // extern "C"

namespace rv {

rv_status bam_query_into(const rv_bam* b, const char* chr, int32_t start, int32_t end, BamQuery& q, rv_reads* out,
                         std::string& err) {
    if (!b || !chr || !out) return RV_ERR_ARG;
    int32_t t = -1;
    for (size_t i = 0; i < b->names.size(); i++)
        if (b->names[i] == chr) t = (int32_t)i;
    if (t < 0) {
        err = std::string("contig not in the BAM header: ") + chr;
        return RV_ERR_ARG;
    }
    // 1-based inclusive region -> [beg, end) 0-based, as htslib's region parser
    const int64_t beg = (int64_t)std::max(start, 1) - 1, en = end;
    std::vector<size_t> sel;
    size_t lo = 0, hi = b->rec.size();
    if (b->sorted) {
        // the contig's block, then the starts that can reach the region: [beg - max_span, end)
        auto key = [&](size_t i) { return std::make_pair(b->rec_tid[i] < 0 ? INT32_MAX : b->rec_tid[i], b->rec_pos[i]); };
        auto first_at_least = [&](std::pair<int32_t, int64_t> k) {
            size_t a = 0, z = b->rec.size();
            while (a < z) {
                const size_t m = (a + z) / 2;
                const auto km = key(m);
                if (km.first < k.first || (km.first == k.first && km.second < k.second))
                    a = m + 1;
                else
                    z = m;
            }
            return a;
        };
        lo = first_at_least({t, beg - (int64_t)b->max_span});
        hi = first_at_least({t, en});
    }
    for (size_t i = lo; i < hi; i++) {
        if (b->rec_tid[i] != t) continue;
        if ((int64_t)b->rec_pos[i] >= en) continue;
        if ((int64_t)b->rec_end[i] <= beg) continue;
        sel.push_back(i);
    }
    const size_t n = sel.size();
    const uint8_t* d = b->data.data();
    // sizes, then offsets
    q.seq_off.assign(n + 1, 0);
    q.cigar_off.assign(n + 1, 0);
    for (size_t k = 0; k < n; k++) {
        const uint8_t* r = d + b->rec[sel[k]] + 4;
        q.seq_off[k + 1] = q.seq_off[k] + rd32s(r + 16);
        q.cigar_off[k + 1] = q.cigar_off[k] + rd16(r + 12);
    }
    q.pos0.resize(n);
    q.nm.resize(n);
    q.mpos0.resize(n);
    q.tid.resize(n);
    q.mtid.resize(n);
    q.flag.resize(n);
    q.mapq.resize(n);
    q.has_nm.resize(n);
    q.bases.resize((size_t)q.seq_off[n]);
    q.quals.resize((size_t)q.seq_off[n]);
    q.cigar.resize((size_t)q.cigar_off[n]);
    parallel_for(n, b->threads, [&](size_t k) {
        const size_t o = b->rec[sel[k]];
        const uint8_t* r = d + o + 4;
        const uint8_t* rend = r + rd32(d + o);
        q.tid[k] = rd32s(r);
        q.pos0[k] = rd32s(r + 4);
        const uint8_t l_read_name = r[8];
        q.mapq[k] = r[9];
        const uint16_t n_cigar = rd16(r + 12);
        q.flag[k] = rd16(r + 14);
        const int32_t l_seq = rd32s(r + 16);
        q.mtid[k] = rd32s(r + 20);
        q.mpos0[k] = rd32s(r + 24);
        const uint8_t* cg = r + 32 + l_read_name;
        for (uint16_t c = 0; c < n_cigar; c++) q.cigar[(size_t)q.cigar_off[k] + c] = rd32(cg + 4 * c);
        const uint8_t* sq = cg + 4 * (size_t)n_cigar;
        uint8_t* bo = q.bases.data() + q.seq_off[k];
        for (int32_t q = 0; q < l_seq; q++) bo[q] = (uint8_t)NT16[(sq[q >> 1] >> ((~q & 1) << 2)) & 15];
        const uint8_t* ql = sq + (l_seq + 1) / 2;
        memcpy(q.quals.data() + q.seq_off[k], ql, (size_t)l_seq);
        int32_t v = 0;
        const bool has = aux_nm(ql + l_seq, rend, v);
        q.has_nm[k] = has ? 1 : 0;
        q.nm[k] = has ? v : 0;
    });
    rv_reads rr;
    rr.n_reads = (int64_t)n;
    rr.pos0 = q.pos0.data();
    rr.flag = q.flag.data();
    rr.mapq = q.mapq.data();
    rr.has_nm = q.has_nm.data();
    rr.nm = q.nm.data();
    rr.mpos0 = q.mpos0.data();
    rr.tid = q.tid.data();
    rr.mtid = q.mtid.data();
    rr.seq_off = q.seq_off.data();
    rr.cigar_off = q.cigar_off.data();
    rr.bases = q.bases.data();
    rr.quals = q.quals.data();
    rr.cigar = q.cigar.data();
    rr.on_device = 0;
    *out = rr;
    return RV_OK;
}

}  // namespace rv
