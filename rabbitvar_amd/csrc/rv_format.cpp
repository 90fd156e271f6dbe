// Simple-mode line formatting (src/modes/simpleMode.cpp:18-64, the printer of
// print_output_variant_simple; numbers through std::to_string, i.e. "%d" and "%f").
#include "rv_format.h"

#include <cstdio>
#include <cstring>

namespace rv {

namespace {

void put_u(std::string& o, unsigned long long v) {
    char b[24];
    int n = 0;
    do {
        b[n++] = (char)('0' + (int)(v % 10));
        v /= 10;
    } while (v);
    while (n) o += b[--n];
}
void put_i(std::string& o, long long v) {
    if (v < 0) {
        o += '-';
        put_u(o, (unsigned long long)(-(v + 1)) + 1ull);
    } else {
        put_u(o, (unsigned long long)v);
    }
}
}  // namespace

// "%f" (six decimals) of a finite double, exactly as printf prints it: the binary value times 10^6
// rounded to the nearest integer, ties to even, computed in 128-bit integers from the value's
// significand and exponent.  Values of 2^43 and above, and non-finite ones, go through snprintf.
void put_f(std::string& o, double v) {
    uint64_t bits;
    std::memcpy(&bits, &v, 8);
    const bool neg = (bits >> 63) != 0;
    const int bexp = (int)((bits >> 52) & 0x7ff);
    uint64_t m = bits & ((1ull << 52) - 1);
    if (bexp == 0x7ff || (bexp != 0 && bexp - 1075 >= -10)) {  // non-finite, or |v| >= 2^43
        char b[512];
        const int n = snprintf(b, sizeof b, "%f", v);
        o.append(b, (size_t)n);
        return;
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
    if (neg) o += '-';
    put_u(o, q / 1000000u);
    o += '.';
    char f[6];
    unsigned long long fr = q % 1000000u;
    for (int i = 5; i >= 0; i--) {
        f[i] = (char)('0' + (int)(fr % 10));
        fr /= 10;
    }
    o.append(f, 6);
}

namespace {
void put_s0(std::string& o, const char* p, int n) {  // empty strings print as "0"
    if (n == 0)
        o += '0';
    else
        o.append(p, (size_t)n);
}

}  // namespace

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
    for (int k = 0, off = 0; k < 6; k++) {
        part[k] = s + off;
        off += o.len[k];
    }
    out += ctx.sample;
    out += '\t';
    out += ctx.gene;
    out += '\t';
    out += ctx.chr;
    out += '\t';
    put_i(out, o.start);
    out += '\t';
    put_i(out, o.end);
    out += '\t';
    out.append(part[0], o.len[0]);  // refallele
    out += '\t';
    out.append(part[1], o.len[1]);  // varallele
    out += '\t';
    put_i(out, o.tpc);
    out += '\t';
    put_i(out, o.poscov);
    out += '\t';
    put_i(out, o.rfc);
    out += '\t';
    put_i(out, o.rrc);
    out += '\t';
    put_i(out, o.fwd);
    out += '\t';
    put_i(out, o.rev);
    out += '\t';
    put_s0(out, part[2], o.len[2]);  // genotype
    out += '\t';
    put_f(out, o.freq);
    out += '\t';
    out.append(part[3], o.len[3]);  // strand bias flags
    out += '\t';
    put_f(out, o.pmean);
    out += '\t';
    out += o.pstd ? '1' : '0';
    out += '\t';
    put_f(out, o.qmean);
    out += '\t';
    out += o.qstd ? '1' : '0';
    out += '\t';
    if (ctx.fisher) {
        put_f(out, o.fisher_p);
        out += '\t';
        put_f(out, o.fisher_ratio);
        out += '\t';
    }
    put_f(out, o.mq);
    out += '\t';
    put_f(out, o.hlr);
    out += '\t';
    put_f(out, o.hifreq);
    out += '\t';
    put_f(out, o.extrafreq);
    out += '\t';
    put_i(out, o.shift3);
    out += '\t';
    put_f(out, o.msi);
    out += '\t';
    put_i(out, o.msint);
    out += '\t';
    if (o.nm > 0)
        put_f(out, o.nm);
    else
        out += '0';
    out += '\t';
    put_i(out, o.hicnt);
    out += '\t';
    put_i(out, o.hicov);
    out += '\t';
    put_s0(out, part[4], o.len[4]);  // leftseq
    out += '\t';
    put_s0(out, part[5], o.len[5]);  // rightseq
    out += '\t';
    out += ctx.chr;
    out += ':';
    put_i(out, ctx.region_start);
    out += '-';
    put_i(out, ctx.region_end);
    out += '\t';
    out += vartype_name(o, arena);
    out += '\t';
    out += "0.000000";  // duprate (0 unless -t)
    out += '\t';
    out += '0';  // structural-variant field: the path calls no SVs
    out += '\n';
}

}  // namespace rv
