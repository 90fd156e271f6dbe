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
// Two modes, chosen at open from the inflated size (the BGZF blocks' ISIZE footers, read without
// inflating):
//  * cached (inflated size within the caller's bound): the file is inflated once, its blocks in
//    parallel, and every record is indexed; a query converts the selected records to SoA on several
//    threads.
//  * streamed (larger files): one pass inflates the file chunk by chunk and keeps only a per-block
//    index (file offset, sizes, the first record starting in the block and the keys of its first
//    and last record); a query reads and inflates just the blocks whose records can reach the
//    region (coordinate-sorted files: a binary search over the block keys; unsorted files: every
//    block, chunk by chunk), so memory stays at the index plus one region.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
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

namespace {
// one BGZF block of a streamed file
struct BlockIx {
    uint64_t coff;       // file offset of the block
    uint64_t uoff;       // offset of its inflated bytes in the stream
    uint32_t csize;      // bytes of the whole block in the file
    uint32_t usize;      // inflated bytes
    uint32_t first_rec;  // offset in the block of the first record starting in it (usize: none)
    int32_t t_first, p_first, t_last, p_last;  // keys of the first and last record starting in it
    uint32_t nrec;
};
}  // namespace

struct rv_bam {
    std::string err;
    bool cached = true;
    std::vector<uint8_t> data;  // cached: the inflated stream
    std::vector<std::string> names;
    std::vector<int32_t> lens;
    std::vector<size_t> rec;  // cached: offset of each record's block_size field
    std::vector<int32_t> rec_tid, rec_pos, rec_end;
    int fd = -1;                  // streamed: the open file
    std::vector<BlockIx> blocks;  // streamed: every BGZF block
    size_t n_records = 0;
    bool sorted = false;     // coordinate order: records of a contig contiguous, starts non-decreasing
    int32_t max_span = 0;    // longest record (end - start)
    int threads = 1;
    uint64_t bytes_in = 0;
    rv::BamQuery last;  // the arrays of the last rv_bam_query
    ~rv_bam() {
        if (fd >= 0) close(fd);
    }
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

// The BGZF blocks of a file, from their headers and footers (nothing is inflated).
rv_status scan_blocks(rv_bam* b, int fd, std::vector<BlockIx>& blocks) {
    struct stat st;
    if (fstat(fd, &st) != 0) return fail(b, "cannot stat the file");
    const uint64_t size = (uint64_t)st.st_size;
    b->bytes_in = size;
    constexpr size_t CHUNK = 16u << 20;
    std::vector<uint8_t> buf(CHUNK + 70000);
    uint64_t off = 0, uoff = 0;
    uint64_t buf_off = 0;  // buf holds the file's bytes [buf_off, buf_off + have)
    size_t have = 0;
    while (off < size) {
        size_t avail = (size_t)(buf_off + have - off);
        // keep at least one whole block (at most 64 KiB) in the buffer
        if (avail < 65536 + 26 && buf_off + have < size) {
            memmove(buf.data(), buf.data() + (size_t)(off - buf_off), avail);
            buf_off = off;
            have = avail;
            const size_t want = (size_t)std::min<uint64_t>(CHUNK, size - (buf_off + have));
            if (pread(fd, buf.data() + have, want, (off_t)(buf_off + have)) != (ssize_t)want) return fail(b, "short read");
            have += want;
            avail = have;
        }
        const uint8_t* h = buf.data() + (size_t)(off - buf_off);
        if (avail < 26) return fail(b, "truncated BGZF block header");
        if (h[0] != 31 || h[1] != 139 || h[2] != 8 || !(h[3] & 4)) return fail(b, "not a BGZF block");
        const uint16_t xlen = rd16(h + 10);
        if (avail < (size_t)12 + xlen) return fail(b, "truncated BGZF block header");
        size_t bsize = 0;
        for (size_t x = 0; x + 4 <= xlen;) {
            const uint8_t* sf = h + 12 + x;
            const uint16_t slen = rd16(sf + 2);
            if (sf[0] == 'B' && sf[1] == 'C' && slen == 2) bsize = (size_t)rd16(sf + 4) + 1;
            x += 4 + slen;
        }
        if (bsize == 0 || bsize > avail || bsize < (size_t)(12 + xlen + 8)) return fail(b, "bad BGZF block size");
        BlockIx bl{};
        bl.coff = off;
        bl.csize = (uint32_t)bsize;
        bl.usize = rd32(h + bsize - 4);
        bl.uoff = uoff;
        bl.first_rec = bl.usize;
        blocks.push_back(bl);
        uoff += bl.usize;
        off += bsize;
    }
    return RV_OK;
}

// Inflates blocks [k0, k1) (whose compressed bytes start at raw, file offset raw_off) into out, each
// at its stream offset minus out_base, on `threads` threads; checks every CRC.
bool inflate_blocks(const std::vector<BlockIx>& blocks, size_t k0, size_t k1, const uint8_t* raw, uint64_t raw_off,
                    uint8_t* out, uint64_t out_base, int threads) {
    std::atomic<int> bad{0};
    parallel_for(k1 - k0, threads, [&](size_t i) {
        const BlockIx& bl = blocks[k0 + i];
        if (bl.usize == 0) return;
        const uint8_t* h = raw + (bl.coff - raw_off);
        const uint16_t xlen = rd16(h + 10);
        z_stream zs;
        memset(&zs, 0, sizeof zs);
        if (inflateInit2(&zs, -15) != Z_OK) {
            bad = 1;
            return;
        }
        uint8_t* dst = out + (bl.uoff - out_base);
        zs.next_in = const_cast<Bytef*>(h + 12 + xlen);
        zs.avail_in = (uInt)(bl.csize - 12 - xlen - 8);
        zs.next_out = dst;
        zs.avail_out = (uInt)bl.usize;
        const int r = inflate(&zs, Z_FINISH);
        inflateEnd(&zs);
        if (r != Z_STREAM_END || zs.avail_out != 0 || crc32(0L, dst, (uInt)bl.usize) != rd32(h + bl.csize - 8)) bad = 1;
    });
    return !bad;
}

// The inflated bytes of blocks [k0, k1), read from the file.
bool read_blocks(const rv_bam* b, size_t k0, size_t k1, std::vector<uint8_t>& out, size_t out_at) {
    if (k0 >= k1) return true;
    const uint64_t c0 = b->blocks[k0].coff, c1 = b->blocks[k1 - 1].coff + b->blocks[k1 - 1].csize;
    std::vector<uint8_t> raw((size_t)(c1 - c0));
    if (pread(b->fd, raw.data(), raw.size(), (off_t)c0) != (ssize_t)raw.size()) return false;
    const uint64_t u0 = b->blocks[k0].uoff, u1 = b->blocks[k1 - 1].uoff + b->blocks[k1 - 1].usize;
    out.resize(out_at + (size_t)(u1 - u0));
    return inflate_blocks(b->blocks, k0, k1, raw.data(), c0, out.data() + out_at, u0, b->threads);
}

// The header at d[0, n): names and lengths; returns the bytes it takes, 0 when more are needed.
size_t parse_header(rv_bam* b, const uint8_t* d, size_t n, rv_status& st) {
    st = RV_OK;
    if (n < 12) return 0;
    if (memcmp(d, "BAM\1", 4) != 0) {
        st = fail(b, "not a BAM file");
        return 0;
    }
    size_t p = 4;
    const uint32_t l_text = rd32(d + p);
    p += 4;
    if (p + (size_t)l_text + 4 > n) return 0;
    p += l_text;
    const int32_t n_ref = rd32s(d + p);
    p += 4;
    if (n_ref < 0) {
        st = fail(b, "bad reference count");
        return 0;
    }
    std::vector<std::string> names;
    std::vector<int32_t> lens;
    for (int32_t r = 0; r < n_ref; r++) {
        if (p + 4 > n) return 0;
        const uint32_t l_name = rd32(d + p);
        p += 4;
        if (l_name == 0) {
            st = fail(b, "truncated reference list");
            return 0;
        }
        if (p + (size_t)l_name + 4 > n) return 0;
        names.emplace_back((const char*)d + p, strnlen((const char*)d + p, l_name));
        p += l_name;
        lens.push_back(rd32s(d + p));
        p += 4;
    }
    b->names = std::move(names);
    b->lens = std::move(lens);
    return p;
}

// A record's size check; false when its fields exceed its block_size.
bool record_ok(const uint8_t* r, uint32_t bs) {
    const uint8_t l_read_name = r[8];
    const uint16_t n_cigar = rd16(r + 12);
    const int32_t l_seq = rd32s(r + 16);
    return bs >= 32 && l_seq >= 0 &&
           32 + (size_t)l_read_name + 4 * (size_t)n_cigar + (size_t)(l_seq + 1) / 2 + (size_t)l_seq <= bs;
}

// bam_endpos of the record whose fields start at r (after block_size)
int32_t record_end(const uint8_t* r) {
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
    return (int32_t)std::min<int64_t>(INT32_MAX, pos + rlen);
}

int32_t key_tid(int32_t t) { return t < 0 ? INT32_MAX : t; }

// Order bookkeeping shared by both modes: coordinate-sortedness and the longest record.
struct OrderCheck {
    std::vector<char> seen;
    int32_t prev_tid = 0, prev_pos = 0;
    bool any = false, sorted = true;
    int32_t max_span = 0;
    explicit OrderCheck(size_t n_names) : seen(n_names + 1, 0) {}
    void add(int32_t tid, int32_t pos, int32_t end) {
        max_span = std::max(max_span, end - pos);
        if (!any || tid != prev_tid) {
            const size_t slot = tid < 0 || (size_t)tid >= seen.size() - 1 ? seen.size() - 1 : (size_t)tid;
            if (seen[slot]) sorted = false;
            seen[slot] = 1;
        } else if (pos < prev_pos) {
            sorted = false;
        }
        any = true;
        prev_tid = tid;
        prev_pos = pos;
    }
};

rv_status load_cached(rv_bam* b, int fd, const std::vector<BlockIx>& blocks) {
    const uint64_t total = blocks.empty() ? 0 : blocks.back().uoff + blocks.back().usize;
    b->data.clear();
    if (!read_blocks(b, 0, blocks.size(), b->data, 0)) return fail(b, "corrupt BGZF block");
    (void)fd;
    if (b->data.size() != total) return fail(b, "inflated size differs from the blocks' footers");
    const uint8_t* d = b->data.data();
    const size_t n = b->data.size();
    rv_status st;
    size_t p = parse_header(b, d, n, st);
    if (st != RV_OK) return st;
    if (p == 0) return fail(b, "truncated header");
    while (p < n) {
        if (p + 4 > n) return fail(b, "truncated record");
        const uint32_t bs = rd32(d + p);
        if (bs < 32 || p + 4 + bs > n) return fail(b, "bad record size");
        if (!record_ok(d + p + 4, bs)) return fail(b, "record fields exceed its size");
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
        b->rec_tid[i] = rd32s(r);
        b->rec_pos[i] = rd32s(r + 4);
        b->rec_end[i] = record_end(r);
    });
    OrderCheck oc(b->names.size());
    for (size_t i = 0; i < nr; i++) oc.add(b->rec_tid[i], b->rec_pos[i], b->rec_end[i]);
    b->sorted = oc.sorted;
    b->max_span = oc.max_span;
    b->n_records = nr;
    return RV_OK;
}

// Streamed: inflates the file in chunks of blocks and keeps the per-block index only.
rv_status load_streamed(rv_bam* b) {
    std::vector<BlockIx>& blocks = b->blocks;
    const size_t nb = blocks.size();
    constexpr uint64_t CHUNK_U = 256u << 20;  // inflated bytes per chunk
    std::vector<uint8_t> buf;
    uint64_t buf_uoff = 0;  // stream offset of buf[0]
    bool header = false;
    OrderCheck oc(0);
    size_t kb = 0;  // block containing the current record start
    for (size_t k0 = 0; k0 < nb;) {
        size_t k1 = k0;
        uint64_t u = 0;
        while (k1 < nb && (k1 == k0 || u + blocks[k1].usize <= CHUNK_U)) u += blocks[k1++].usize;
        // buf keeps the unparsed tail of the previous chunk, then this chunk's bytes
        const size_t keep = buf.size();
        if (!read_blocks(b, k0, k1, buf, keep)) return fail(b, "corrupt BGZF block");
        k0 = k1;
        const uint8_t* d = buf.data();
        const size_t n = buf.size();
        size_t p = 0;
        if (!header) {
            rv_status st;
            p = parse_header(b, d, n, st);
            if (st != RV_OK) return st;
            if (p == 0) {
                if (k0 >= nb) return fail(b, "truncated header");
                continue;  // the header spans into the next chunk
            }
            header = true;
            oc = OrderCheck(b->names.size());
        }
        while (p + 4 <= n) {
            const uint32_t bs = rd32(d + p);
            if (bs < 32) return fail(b, "bad record size");
            if (p + 4 + bs > n) break;  // finishes in the next chunk
            const uint8_t* r = d + p + 4;
            if (!record_ok(r, bs)) return fail(b, "record fields exceed its size");
            const uint64_t so = buf_uoff + p;
            while (kb + 1 < nb && blocks[kb].uoff + blocks[kb].usize <= so) kb++;
            BlockIx& bl = blocks[kb];
            const int32_t tid = rd32s(r), pos = rd32s(r + 4);
            if (bl.nrec == 0) {
                bl.first_rec = (uint32_t)(so - bl.uoff);
                bl.t_first = key_tid(tid);
                bl.p_first = pos;
            }
            bl.t_last = key_tid(tid);
            bl.p_last = pos;
            bl.nrec++;
            oc.add(tid, pos, record_end(r));
            b->n_records++;
            p += 4 + bs;
        }
        if (p < n && k0 >= nb) return fail(b, "truncated record");
        buf.erase(buf.begin(), buf.begin() + (ptrdiff_t)p);
        buf_uoff += p;
    }
    if (!header) return fail(b, "truncated header");
    // blocks where no record starts take the keys of the last record before them, so that the keys
    // never decrease along the file (for the query's binary searches)
    int32_t tl = INT32_MIN, pl = INT32_MIN;
    for (BlockIx& bl : blocks) {
        if (bl.nrec == 0) {
            bl.t_first = bl.t_last = tl;
            bl.p_first = bl.p_last = pl;
        } else {
            tl = bl.t_last;
            pl = bl.p_last;
        }
    }
    b->sorted = oc.sorted;
    b->max_span = oc.max_span;
    return RV_OK;
}

rv_status load(rv_bam* b, const char* path, uint64_t max_cached) {
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(b, std::string("cannot open ") + path);
    b->fd = fd;
    std::vector<BlockIx> blocks;
    rv_status s = scan_blocks(b, fd, blocks);
    if (s != RV_OK) return s;
    const uint64_t total = blocks.empty() ? 0 : blocks.back().uoff + blocks.back().usize;
    if (total <= max_cached) {
        b->cached = true;
        b->blocks = std::move(blocks);  // read_blocks reads through the index
        s = load_cached(b, fd, b->blocks);
        b->blocks.clear();
        b->blocks.shrink_to_fit();
        close(fd);
        b->fd = -1;
        return s;
    }
    b->cached = false;
    b->blocks = std::move(blocks);
    return load_streamed(b);
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

rv_status rv_bam_open_bounded(const char* path, int threads, uint64_t max_cached_bytes, rv_bam** out) {
    if (!path || !out) return RV_ERR_ARG;
    *out = nullptr;
    rv_bam* b = new (std::nothrow) rv_bam;
    if (!b) return RV_ERR_OOM;
    b->threads = threads > 0 ? threads : 1;
    rv_status s;
    try {
        s = load(b, path, max_cached_bytes);
    } catch (const std::bad_alloc&) {
        s = RV_ERR_OOM;
        b->err = "out of memory";
    }
    *out = b;  // on failure the caller reads rv_bam_last_error, then closes
    return s;
}

rv_status rv_bam_open(const char* path, int threads, rv_bam** out) {
    return rv_bam_open_bounded(path, threads, RV_BAM_DEFAULT_CACHE_BYTES, out);
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

size_t rv_bam_n_records(const rv_bam* b) { return b ? b->n_records : 0; }

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
    auto wanted = [&](const uint8_t* r) {  // r: the record's fields, after block_size
        return rd32s(r) == t && (int64_t)rd32s(r + 4) < en && (int64_t)record_end(r) > beg;
    };
    std::vector<const uint8_t*> recp;  // the selected records' block_size fields, in file order
    std::vector<uint8_t> local;        // streamed: the inflated blocks the records lie in
    if (b->cached) {
        size_t lo = 0, hi = b->rec.size();
        if (b->sorted) {
            // the contig's block, then the starts that can reach the region: [beg - max_span, end)
            auto key = [&](size_t i) { return std::make_pair(key_tid(b->rec_tid[i]), b->rec_pos[i]); };
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
            recp.push_back(b->data.data() + b->rec[i]);
        }
    } else if (b->sorted) {
        // blocks whose records can reach the region: from the first block whose last record starts
        // at or after beg - max_span, to before the first block whose first record starts at or
        // after the region's end (keys never decrease along the file)
        const auto& bl = b->blocks;
        auto before = [](int32_t ta, int64_t pa, int32_t tb, int64_t pb) { return ta < tb || (ta == tb && pa < pb); };
        const int64_t lo_pos = beg - (int64_t)b->max_span;
        // (blocks where no record starts carry the keys of the last record before them, so both
        // predicates are monotone along the file)
        size_t k0 = (size_t)(std::partition_point(bl.begin(), bl.end(), [&](const BlockIx& x) {
                                 return before(x.t_last, x.p_last, t, lo_pos);
                             }) - bl.begin());
        while (k0 < bl.size() && bl[k0].nrec == 0) k0++;
        const size_t k1 = (size_t)(std::partition_point(bl.begin(), bl.end(), [&](const BlockIx& x) {
                                       return x.nrec == 0 ? before(x.t_last, x.p_last, t, en)
                                                          : before(x.t_first, x.p_first, t, en);
                                   }) - bl.begin());
        if (k0 < k1) {
            const uint64_t s_end = k1 < bl.size() ? bl[k1].uoff : bl.back().uoff + bl.back().usize;
            size_t ke = k1;
            if (!read_blocks(b, k0, ke, local, 0)) {
                err = "corrupt BGZF block";
                return RV_ERR_ARG;
            }
            const uint64_t base = bl[k0].uoff;
            uint64_t so = base + bl[k0].first_rec;
            std::vector<size_t> at;
            while (so < s_end) {
                // the whole record in the buffer, reading further blocks when it runs past
                auto need = [&](uint64_t upto) {
                    while (base + local.size() < upto && ke < bl.size()) {
                        if (!read_blocks(b, ke, ke + 1, local, local.size())) return false;
                        ke++;
                    }
                    return base + local.size() >= upto;
                };
                if (!need(so + 4)) break;
                const uint32_t bs = rd32(local.data() + (so - base));
                if (!need(so + 4 + bs)) {
                    err = "truncated record";
                    return RV_ERR_ARG;
                }
                if (wanted(local.data() + (so - base) + 4)) at.push_back((size_t)(so - base));
                so += 4 + bs;
            }
            for (size_t o : at) recp.push_back(local.data() + o);
        }
    } else {
        // unsorted: every block, chunk by chunk; the selected records are copied out
        const auto& bl = b->blocks;
        constexpr uint64_t CHUNK_U = 256u << 20;
        std::vector<uint8_t> buf;
        std::vector<size_t> at;
        bool header = false;
        size_t hdr = 0;
        for (size_t k0 = 0; k0 < bl.size();) {
            size_t k1 = k0;
            uint64_t u = 0;
            while (k1 < bl.size() && (k1 == k0 || u + bl[k1].usize <= CHUNK_U)) u += bl[k1++].usize;
            const size_t keep = buf.size();
            if (!read_blocks(b, k0, k1, buf, keep)) {
                err = "corrupt BGZF block";
                return RV_ERR_ARG;
            }
            k0 = k1;
            size_t p = 0;
            if (!header) {
                // the header's length (its contents were read at open)
                if (buf.size() < 8) continue;
                const size_t l_text = rd32(buf.data() + 4);
                size_t q = 8 + l_text;
                if (q + 4 > buf.size()) continue;
                const int32_t n_ref = rd32s(buf.data() + q);
                q += 4;
                bool whole = true;
                for (int32_t r = 0; r < n_ref && whole; r++) {
                    if (q + 4 > buf.size()) {
                        whole = false;
                        break;
                    }
                    q += 4 + rd32(buf.data() + q) + 4;
                    if (q > buf.size()) whole = false;
                }
                if (!whole) continue;
                hdr = q;
                p = hdr;
                header = true;
            }
            while (p + 4 <= buf.size()) {
                const uint32_t bs = rd32(buf.data() + p);
                if (p + 4 + bs > buf.size()) break;
                if (wanted(buf.data() + p + 4)) {
                    at.push_back(local.size());
                    local.insert(local.end(), buf.begin() + (ptrdiff_t)p, buf.begin() + (ptrdiff_t)(p + 4 + bs));
                }
                p += 4 + bs;
            }
            buf.erase(buf.begin(), buf.begin() + (ptrdiff_t)p);
        }
        for (size_t o : at) recp.push_back(local.data() + o);
    }
    const size_t n = recp.size();
    // sizes, then offsets
    q.seq_off.assign(n + 1, 0);
    q.cigar_off.assign(n + 1, 0);
    for (size_t k = 0; k < n; k++) {
        const uint8_t* r = recp[k] + 4;
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
        const uint8_t* r = recp[k] + 4;
        const uint8_t* rend = r + rd32(recp[k]);
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
