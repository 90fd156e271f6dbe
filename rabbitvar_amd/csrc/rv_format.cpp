// Simple-mode line formatting (src/modes/simpleMode.cpp:18-64, the printer of
// print_output_variant_simple; numbers through std::to_string, i.e. "%d" and "%f").
#include "rv_format.h"

#include <cstdio>
#include <cstring>
#include <vector>

namespace rv {

namespace {

// Writers into a caller's buffer: each returns the end of what it wrote.
char* w_u(char* p, unsigned long long v) {
    char b[24];
    int n = 0;
    do {
        b[n++] = (char)('0' + (int)(v % 10));
        v /= 10;
    } while (v);
    while (n) *p++ = b[--n];
    return p;
}
char* w_i(char* p, long long v) {
    if (v < 0) {
        *p++ = '-';
        return w_u(p, (unsigned long long)(-(v + 1)) + 1ull);
    }
    return w_u(p, (unsigned long long)v);
}
char* w_s(char* p, const char* s, size_t n) {
    std::memcpy(p, s, n);
    return p + n;
}
char* w_s0(char* p, const char* s, int n) {  // empty strings print as "0"
    if (n == 0) {
        *p++ = '0';
        return p;
    }
    return w_s(p, s, (size_t)n);
}

constexpr size_t W_F_MAX = 400;  // the longest "%f" of a double (1.8e308: 309 digits + ".000000")

// "%f" (six decimals) of a double, exactly as printf prints it: the binary value times 10^6
// rounded to the nearest integer, ties to even, computed in 128-bit integers from the value's
// significand and exponent.  Values of 2^43 and above, and non-finite ones, go through snprintf.
char* w_f(char* p, double v) {
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    const bool neg = (bits >> 63) != 0;
    const int bexp = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    if (bexp == 0x7ff || (bexp != 0 && bexp - 1075 >= -10)) {  // non-finite, or |v| >= 2^43
        const int n = snprintf(p, W_F_MAX, "%f", v);
        return p + (n > 0 ? (n < (int)W_F_MAX ? n : (int)W_F_MAX - 1) : 0);
    }
    int e;
    if (bexp == 0) {
        e = -1074;  // subnormal
    } else {
        m |= 1ull << 52;
        e = bexp - 1075;
    }
    // |v| = m * 2^e with e < -10 here
    const unsigned __int128 x = (unsigned __int128)m * 1000000u;
    const int k = -e;
    unsigned long long q;
    if (k >= 100) {
        q = 0;  // x < 2^73 is below half of 2^k
    } else {
        q = (unsigned long long)(x >> k);
        const unsigned __int128 r = x & ((((unsigned __int128)1) << k) - 1);
        const unsigned __int128 half = ((unsigned __int128)1) << (k - 1);
        if (r > half || (r == half && (q & 1))) q++;
    }
    if (neg) *p++ = '-';
    p = w_u(p, q / 1000000u);
    *p++ = '.';
    unsigned long long fr = q % 1000000u;
    for (int i = 5; i >= 0; i--) {
        p[i] = (char)('0' + (int)(fr % 10));
        fr /= 10;
    }
    return p + 6;
}

}  // namespace

void put_f(std::string& o, double v) {
    char b[W_F_MAX];
    o.append(b, (size_t)(w_f(b, v) - b));
}

std::string vartype_name(const OutRec& o, const char* arena) {
    switch (o.vartype) {
        case VT_SNV: return "SNV";
        case VT_INS: return "Insertion";
        case VT_DEL: return "Deletion";
        case VT_COMPLEX: return "Complex";
        case VT_SV: {
            const char* va = arena + o.str_off + o.len[0];
            return std::string(va + 1, 3);
        }
        default: return "";
    }
}

void format_line(std::string& out, const OutRec& o, const char* arena, const LineContext& ctx) {
    const char* s = arena + o.str_off;
    const char* part[6];
    size_t strs = 0;
    for (int k = 0, off = 0; k < 6; k++) {
        part[k] = s + off;
        off += o.len[k];
        strs += o.len[k];
    }
    const size_t ls = std::strlen(ctx.sample), lg = std::strlen(ctx.gene), lc = std::strlen(ctx.chr);
    const std::string vt = vartype_name(o, arena);
    // the line's length is at most: its strings, 16 numbers of "%f", 18 integers, the separators
    const size_t bound = ls + lg + 2 * lc + strs + vt.size() + 16 * W_F_MAX + 18 * 24 + 96;
    char stackbuf[8192];
    std::vector<char> heapbuf;
    char* b = stackbuf;
    if (bound > sizeof stackbuf) {
        heapbuf.resize(bound);
        b = heapbuf.data();
    }
    char* p = b;
    p = w_s(p, ctx.sample, ls);
    *p++ = '\t';
    p = w_s(p, ctx.gene, lg);
    *p++ = '\t';
    p = w_s(p, ctx.chr, lc);
    *p++ = '\t';
    p = w_i(p, o.start);
    *p++ = '\t';
    p = w_i(p, o.end);
    *p++ = '\t';
    p = w_s(p, part[0], (size_t)o.len[0]);  // refallele
    *p++ = '\t';
    p = w_s(p, part[1], (size_t)o.len[1]);  // varallele
    *p++ = '\t';
    p = w_i(p, o.tpc);
    *p++ = '\t';
    p = w_i(p, o.poscov);
    *p++ = '\t';
    p = w_i(p, o.rfc);
    *p++ = '\t';
    p = w_i(p, o.rrc);
    *p++ = '\t';
    p = w_i(p, o.fwd);
    *p++ = '\t';
    p = w_i(p, o.rev);
    *p++ = '\t';
    p = w_s0(p, part[2], o.len[2]);  // genotype
    *p++ = '\t';
    p = w_f(p, o.freq);
    *p++ = '\t';
    p = w_s(p, part[3], (size_t)o.len[3]);  // strand bias flags
    *p++ = '\t';
    p = w_f(p, o.pmean);
    *p++ = '\t';
    *p++ = o.pstd ? '1' : '0';
    *p++ = '\t';
    p = w_f(p, o.qmean);
    *p++ = '\t';
    *p++ = o.qstd ? '1' : '0';
    *p++ = '\t';
    if (ctx.fisher) {
        p = w_f(p, o.fisher_p);
        *p++ = '\t';
        p = w_f(p, o.fisher_ratio);
        *p++ = '\t';
    }
    p = w_f(p, o.mq);
    *p++ = '\t';
    p = w_f(p, o.hlr);
    *p++ = '\t';
    p = w_f(p, o.hifreq);
    *p++ = '\t';
    p = w_f(p, o.extrafreq);
    *p++ = '\t';
    p = w_i(p, o.shift3);
    *p++ = '\t';
    p = w_f(p, o.msi);
    *p++ = '\t';
    p = w_i(p, o.msint);
    *p++ = '\t';
    if (o.nm > 0)
        p = w_f(p, o.nm);
    else
        *p++ = '0';
    *p++ = '\t';
    p = w_i(p, o.hicnt);
    *p++ = '\t';
    p = w_i(p, o.hicov);
    *p++ = '\t';
    p = w_s0(p, part[4], o.len[4]);  // leftseq
    *p++ = '\t';
    p = w_s0(p, part[5], o.len[5]);  // rightseq
    *p++ = '\t';
    p = w_s(p, ctx.chr, lc);
    *p++ = ':';
    p = w_i(p, ctx.region_start);
    *p++ = '-';
    p = w_i(p, ctx.region_end);
    *p++ = '\t';
    p = w_s(p, vt.data(), vt.size());
    *p++ = '\t';
    p = w_s(p, "0.000000", 8);  // duprate (0 unless -t)
    *p++ = '\t';
    *p++ = '0';  // structural-variant field: the path calls no SVs
    *p++ = '\n';
    out.append(b, (size_t)(p - b));
}

}  // namespace rv
