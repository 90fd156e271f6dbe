// Per-read CIGAR rewrite on the device.
//
// Restates CigarModifier::modifyCigar (reference src/cigarModifier.cpp:219-410 and its helpers
// :412-1000, rule texts in include/patterns.h:52-153).  The reference rewrites the CIGAR *text*
// with std::regex; every rule there is a sequence of `\d+`, `\d`, op letters, op classes, `\D`,
// an optional `(\d+I)?`, an optional `^(.*?)` lazy prefix and `^`/`$` anchors.  For such
// patterns the leftmost match is found by trying start offsets left to right with greedy digit
// runs (no backtracking can succeed where the greedy run fails), so a small deterministic matcher
// over the text reproduces std::regex_search / regex_replace(format_first_only) exactly,
// including texts that contain a '-' from a negative length.
#pragma once
#include "rv_common.h"

namespace rv {

typedef FStr<CIGTXT> CText;

// Capture slots: the largest rule (threeIndelsPattern's core, "nMnxnMnxnMnxnM") captures 10.
constexpr int RX_MAXCAP = 12;

struct RxMatch {
    int start, end;
    int ncap;
    int cs[RX_MAXCAP], ce[RX_MAXCAP];
};

// number of captures a pattern makes (every code letter captures)
constexpr int rx_ncap(const char* p) {
    return *p == 0 ? 0 : ((*p == 'n' || *p == 'd' || *p == 'x' || *p == '!' || *p == '?') ? 1 : 0) + rx_ncap(p + 1);
}
static_assert(rx_ncap("nMnxnMnxnMnxnM") <= RX_MAXCAP, "capture slots");

RV_HD bool is_digit(char c) { return c >= '0' && c <= '9'; }

// Pattern codes: 'n' (\d+) 'd' (\d) 'x' ([ID]) '!' (\D) '?' ((\d+I)?) -- all captured in order;
// any other character is a literal.  `beg`/`end` anchor the match.
RV_HD bool rx_match_at(const CText& s, int at, const char* pat, bool anchor_end, RxMatch& m) {
    int i = at;
    m.ncap = 0;
    for (const char* p = pat; *p; ++p) {
        char e = *p;
        if (e == 'n') {
            int b = i;
            while (i < s.n && is_digit(s.c[i])) i++;
            if (i == b) return false;
            m.cs[m.ncap] = b;
            m.ce[m.ncap++] = i;
        } else if (e == 'd') {
            if (i >= s.n || !is_digit(s.c[i])) return false;
            m.cs[m.ncap] = i;
            m.ce[m.ncap++] = i + 1;
            i++;
        } else if (e == 'x') {
            if (i >= s.n || (s.c[i] != 'I' && s.c[i] != 'D')) return false;
            m.cs[m.ncap] = i;
            m.ce[m.ncap++] = i + 1;
            i++;
        } else if (e == '!') {
            if (i >= s.n || is_digit(s.c[i])) return false;
            m.cs[m.ncap] = i;
            m.ce[m.ncap++] = i + 1;
            i++;
        } else if (e == '?') {
            int b = i, j = i;
            while (j < s.n && is_digit(s.c[j])) j++;
            if (j > b && j < s.n && s.c[j] == 'I') {
                m.cs[m.ncap] = b;
                m.ce[m.ncap++] = j + 1;
                i = j + 1;
            } else {
                m.cs[m.ncap] = b;
                m.ce[m.ncap++] = b;  // empty optional group
            }
        } else {
            if (i >= s.n || s.c[i] != e) return false;
            i++;
        }
    }
    if (anchor_end && i != s.n) return false;
    m.start = at;
    m.end = i;
    return true;
}

RV_HD bool rx_search(const CText& s, const char* pat, bool anchor_beg, bool anchor_end, RxMatch& m, int from = 0) {
    if (anchor_beg) return from == 0 && rx_match_at(s, 0, pat, anchor_end, m);
    for (int at = from; at <= s.n; at++)
        if (rx_match_at(s, at, pat, anchor_end, m)) return true;
    return false;
}

RV_HD int rx_int(const CText& s, const RxMatch& m, int g) {
    long long v = 0;
    for (int i = m.cs[g]; i < m.ce[g]; i++) v = v * 10 + (s.c[i] - '0');
    return (int)v;
}
RV_HD char rx_ch(const CText& s, const RxMatch& m, int g) { return m.ce[g] > m.cs[g] ? s.c[m.cs[g]] : 0; }
RV_HD bool rx_empty(const RxMatch& m, int g) { return m.ce[g] == m.cs[g]; }

RV_HD void rx_splice(CText& s, int a, int b, const CText& rep) {
    CText t;
    t.clear();
    t.append(s.c, a);
    t.append(rep);
    t.append(s.c + b, s.n - b);
    if (t.overflow) s.overflow = true;
    bool of = s.overflow;
    s = t;
    s.overflow = of || t.overflow;
}

// replace the leftmost match of `pat` with rep (regex_replace(..., format_first_only) and the
// anchored patterns that can only match once)
RV_HD void rx_replace_first(CText& s, const char* pat, bool ab, bool ae, const CText& rep) {
    RxMatch m;
    if (rx_search(s, pat, ab, ae, m)) rx_splice(s, m.start, m.end, rep);
}

// sum of group-1 numbers over successive matches of (\d+)[set] in s[0, lim)
RV_HD int rx_sum_prefix(const CText& s, int lim, const char* set) {
    int sum = 0, i = 0;
    while (i < lim) {
        if (is_digit(s.c[i])) {
            int b = i;
            while (i < lim && is_digit(s.c[i])) i++;
            if (i < lim) {
                bool in = false;
                for (const char* q = set; *q; ++q)
                    if (s.c[i] == *q) in = true;
                if (in) {
                    long long v = 0;
                    for (int k = b; k < i; k++) v = v * 10 + (s.c[k] - '0');
                    sum += (int)v;
                    i++;
                    continue;
                }
            }
            // a digit run not followed by a set letter: the regex retries at the next offset,
            // which lands inside the same run and fails the same way
            continue;
        }
        i++;
    }
    return sum;
}

// Reference window view used by the rewrite (count/at of the robin_hood reference map)
struct RefView {
    const char* ref;
    int lo, hi;
    RV_HD bool has(int k) const { return k >= lo && k <= hi; }
    RV_HD char at(int k) const { return ref[k - lo]; }
    RV_HD bool hasEq(int k, char ch) const { return has(k) && at(k) == ch; }
    RV_HD bool hasNe(int k, char ch) const { return has(k) && at(k) != ch; }
};

struct ReadView {
    const char* seq;
    const uint8_t* q;
    int lseq;
    RV_HD char s(int i) const { return (i >= 0 && i < lseq) ? seq[i] : '\0'; }
    RV_HD int qq(int i) const { return (i >= 0 && i < lseq) ? (int)q[i] : 0; }
};

// The end-mismatch scans of the tail rules (cigarModifier.cpp:529-568, the 5' dig): the length of
// the mismatching run at the 3' end of an M run ending at refoff / rdoff, and at the 5' end of an
// M run starting at pos (read offset 0); both stop after 3 matches past the last mismatch.
RV_HD int tail_mismatch_run(const RefView& rf, const ReadView& rd, int refoff, int rdoff, int mch) {
    int rn = 0, rrn = 0, rmch = 0;
    while (rrn < mch && rn < mch) {
        if (!rf.has(refoff - rrn - 1)) break;
        if (rrn < rdoff && rf.at(refoff - rrn - 1) != rd.s(rdoff - rrn - 1)) {
            rn = rrn + 1;
            rmch = 0;
        } else if (rrn < rdoff && rf.at(refoff - rrn - 1) == rd.s(rdoff - rrn - 1)) {
            rmch++;
        }
        rrn++;
        if (rmch >= 3) break;
    }
    return rn;
}
RV_HD int head_mismatch_run(const RefView& rf, const ReadView& rd, int pos, int mch) {
    int rn = 0, rrn = 0, rmch = 0;
    while (rrn < mch && rn < mch) {
        if (!rf.has(pos + rrn)) break;
        if (rf.hasNe(pos + rrn, rd.s(rrn))) {
            rn = rrn + 1;
            rmch = 0;
        } else if (rf.hasEq(pos + rrn, rd.s(rrn))) {
            rmch++;
        }
        rrn++;
        if (rmch >= 3) break;
    }
    return rn;
}

struct Rewriter {
    RefView rf;
    ReadView rd;
    int pos;
    CText cs;

    RV_HD void rep2(CText& r, int a, char ca, int b, char cb) {
        r.clear();
        r.append_int(a);
        r.push(ca);
        r.append_int(b);
        r.push(cb);
    }
    RV_HD void rep1(CText& r, int a, char ca) {
        r.clear();
        r.append_int(a);
        r.push(ca);
    }

    // cigarModifier.cpp:412-439
    RV_HD void begin_dig_m(int mch) {
        const int rn = head_mismatch_run(rf, rd, pos, mch);
        if (rn > 0 && rn <= 3) {
            mch -= rn;
            CText r;
            rep2(r, rn, 'S', mch, 'M');
            rx_replace_first(cs, "nM", true, false, r);
            pos += rn;
        }
    }

    // cigarModifier.cpp:445-523
    RV_HD void s_then_m(int soft, int mch) {
        int rn = 0;
        CText r;
        while (rn < soft && rf.hasEq(pos - rn - 1, rd.s(soft - rn - 1)) && rd.qq(soft - rn - 1) > 10) rn++;
        if (rn > 0) {
            mch += rn;
            soft -= rn;
            if (soft > 0)
                rep2(r, soft, 'S', mch, 'M');
            else
                rep1(r, mch, 'M');
            rx_replace_first(cs, "nSnM", true, false, r);
            pos -= rn;
            rn = 0;
        }
        if (soft > 0) {
            uint32_t seen = 0;  // set of reference letters met (RN)
            int rn_nt = 0;
            while (rn + 1 < soft && rf.hasEq(pos - rn - 2, rd.s(soft - rn - 2)) && rd.qq(soft - rn - 2) > 10) {
                rn++;
                if (rf.has(pos - rn - 2)) {
                    uint32_t bit = 1u << ((unsigned char)rf.at(pos - rn - 2) & 31);
                    if (!(seen & bit)) {
                        seen |= bit;
                        rn_nt++;
                    }
                }
            }
            if ((rn > 4 && rn_nt > 1) || rf.hasEq(pos - 1, rd.s(soft - 1))) {
                mch += rn + 1;
                soft -= rn + 1;
                if (soft > 0)
                    rep2(r, soft, 'S', mch, 'M');
                else
                    rep1(r, mch, 'M');
                rx_replace_first(cs, "nSnM", true, false, r);
                pos -= rn + 1;
            }
            if (rn == 0) {
                int rrn = 0, rmch = 0;
                while (rrn < mch && rn < mch) {
                    if (!rf.has(pos + rrn)) break;
                    if (rf.hasNe(pos + rrn, rd.s(soft + rrn))) {
                        rn = rrn + 1;
                        rmch = 0;
                    } else if (rf.hasEq(pos + rrn, rd.s(soft + rrn))) {
                        rmch++;
                    }
                    rrn++;
                    if (rmch >= 3) break;
                }
                if (rn > 0 && rn < mch) {
                    soft += rn;
                    mch -= rn;
                    rep2(r, soft, 'S', mch, 'M');
                    rx_replace_first(cs, "nSnM", true, false, r);
                    pos += rn;
                }
            }
        }
    }

    // cigarModifier.cpp:529-568 (prefix = text before the last M run)
    RV_HD void three_mismatches(int prefix_len, int mch) {
        int refoff = pos + mch, rdoff = mch;
        if (prefix_len > 0) {
            refoff += rx_sum_prefix(cs, prefix_len, "MND");
            rdoff += rx_sum_prefix(cs, prefix_len, "MIS");
        }
        int rn = tail_mismatch_run(rf, rd, refoff, rdoff, mch);
        mch -= rn;
        if (rn > 0 && rn <= 3) {
            CText r;
            rep2(r, mch, 'M', rn, 'S');
            rx_replace_first(cs, "nM", false, true, r);
        }
    }

    // cigarModifier.cpp:574-668
    RV_HD void m_then_s(int prefix_len, int mch, int soft) {
        int refoff = pos + mch, rdoff = mch;
        if (prefix_len > 0) {
            refoff += rx_sum_prefix(cs, prefix_len, "MND");
            rdoff += rx_sum_prefix(cs, prefix_len, "MIS");
        }
        int rn = 0;
        CText r;
        while (rn < soft && rf.hasEq(refoff + rn, rd.s(rdoff + rn)) && rd.qq(rdoff + rn) > 10) rn++;
        if (rn > 0) {
            mch += rn;
            soft -= rn;
            if (soft > 0)
                rep2(r, mch, 'M', soft, 'S');
            else
                rep1(r, mch, 'M');
            rx_replace_first(cs, "nMnS", false, true, r);
            rn = 0;
        }
        if (soft > 0) {
            uint32_t seen = 0;
            int rn_nt = 0;
            while (rn + 1 < soft && rf.hasEq(refoff + rn + 1, rd.s(rdoff + rn + 1)) && rd.qq(rdoff + rn + 1) > 10) {
                rn++;
                if (rf.has(refoff + rn + 1)) {
                    uint32_t bit = 1u << ((unsigned char)rf.at(refoff + rn + 1) & 31);
                    if (!(seen & bit)) {
                        seen |= bit;
                        rn_nt++;
                    }
                }
            }
            if (rn > 4 && rn_nt > 1) {
                mch += rn + 1;
                soft -= rn + 1;
                if (soft > 0)
                    rep2(r, mch, 'M', soft, 'S');
                else
                    rep1(r, mch, 'M');
                rx_replace_first(cs, "nMnS", false, true, r);
            }
            if (rn == 0) {
                int r2 = tail_mismatch_run(rf, rd, refoff, rdoff, mch);
                if (r2 > 0 && r2 < mch) {
                    soft += r2;
                    mch -= r2;
                    rep2(r, mch, 'M', soft, 'S');
                    rx_replace_first(cs, "nMnS", false, true, r);
                }
            }
        }
    }

    RV_HD int leading_matches(int rdoff, int refoff) {
        int rn = 0;
        while (rdoff + rn < rd.lseq && rf.hasEq(refoff + rn, rd.s(rdoff + rn))) rn++;
        return rn;
    }

    // cigarModifier.cpp:753-845; m groups: 1 prefix(start), 2..11 as in threeIndelsPattern
    RV_HD bool three_indels(const RxMatch& m, bool flag) {
        // core groups (without prefix): 0:M 1:n 2:[DI] 3:M 4:n 5:[DI] 6:M 7:n 8:[DI] 9:M
        int g2 = rx_int(cs, m, 0), g3 = rx_int(cs, m, 1), g5 = rx_int(cs, m, 3), g6 = rx_int(cs, m, 4);
        int g8 = rx_int(cs, m, 6), g9 = rx_int(cs, m, 7), g11 = rx_int(cs, m, 9);
        char o4 = rx_ch(cs, m, 2), o7 = rx_ch(cs, m, 5), o10 = rx_ch(cs, m, 8);
        int mid = g5 + g8;
        int tslen = mid;
        if (o4 == 'I') tslen += g3;
        if (o7 == 'I') tslen += g6;
        if (o10 == 'I') tslen += g9;
        int dlen = mid;
        if (o4 == 'D') dlen += g3;
        if (o7 == 'D') dlen += g6;
        if (o10 == 'D') dlen += g9;
        int refoff = pos + g2, rdoff = g2, RDOFF = g2, rm = g11;
        if (m.start > 0) {
            refoff += rx_sum_prefix(cs, m.start, "MND");
            rdoff += rx_sum_prefix(cs, m.start, "MIS");
        }
        int rn = leading_matches(rdoff, refoff);
        RDOFF += rn;
        dlen -= rn;
        tslen -= rn;
        CText nc;
        nc.clear();
        nc.append_int(RDOFF);
        nc.push('M');
        if (tslen <= 0) {
            dlen -= tslen;
            rm += tslen;
            if (dlen == 0) {
                RDOFF = RDOFF + rm;
                nc.clear();
                nc.append_int(RDOFF);
                nc.push('M');
            } else if (dlen < 0) {
                tslen = -dlen;
                rm += dlen;
                if (rm < 0) {
                    RDOFF = RDOFF + rm;
                    nc.clear();
                    nc.append_int(RDOFF);
                    nc.push('M');
                    nc.append_int(tslen);
                    nc.push('I');
                } else {
                    nc.append_int(tslen);
                    nc.push('I');
                    nc.append_int(rm);
                    nc.push('M');
                }
            } else {
                nc.append_int(dlen);
                nc.push('D');
                nc.append_int(rm);
                nc.push('M');
            }
        } else {
            if (dlen == 0) {
                nc.append_int(tslen);
                nc.push('I');
                nc.append_int(rm);
                nc.push('M');
            } else if (dlen < 0) {
                rm += dlen;
                nc.append_int(tslen);
                nc.push('I');
                nc.append_int(rm);
                nc.push('M');
            } else {
                nc.append_int(dlen);
                nc.push('D');
                nc.append_int(tslen);
                nc.push('I');
                nc.append_int(rm);
                nc.push('M');
            }
        }
        if (mid <= 15) {
            rx_replace_first(cs, "nMnxnMnxnMnxnM", false, false, nc);
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:854-908 (threeDeletionsPattern core groups 0..6)
    RV_HD bool three_dels(const RxMatch& m, bool flag) {
        int g2 = rx_int(cs, m, 0), g3 = rx_int(cs, m, 1), g4 = rx_int(cs, m, 2), g5 = rx_int(cs, m, 3);
        int g6 = rx_int(cs, m, 4), g7 = rx_int(cs, m, 5), g8 = rx_int(cs, m, 6);
        int tslen = g4 + g6;
        int dlen = g3 + g4 + g5 + g6 + g7;
        int mid = g4 + g6;
        int refoff = pos + g2, rdoff = g2, RDOFF = g2, rm = g8;
        if (m.start > 0) {
            refoff += rx_sum_prefix(cs, m.start, "MND");
            rdoff += rx_sum_prefix(cs, m.start, "MIS");
        }
        int rn = leading_matches(rdoff, refoff);
        RDOFF += rn;
        dlen -= rn;
        tslen -= rn;
        CText nc;
        nc.clear();
        nc.append_int(RDOFF);
        nc.push('M');
        if (tslen <= 0) {
            dlen -= tslen;
            rm += tslen;
            nc.append_int(dlen);
            nc.push('D');
            nc.append_int(rm);
            nc.push('M');
        } else {
            nc.append_int(dlen);
            nc.push('D');
            nc.append_int(tslen);
            nc.push('I');
            nc.append_int(rm);
            nc.push('M');
        }
        if (mid <= 15) {
            rx_replace_first(cs, "nMnDnMnDnMnDnM", false, false, nc);
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:918-973 (D_M_D_DD_M_D_I_D_M_D_DD core groups 0..6)
    RV_HD bool two_del_ins(const RxMatch& m, bool flag) {
        int mch = rx_int(cs, m, 0), g3 = rx_int(cs, m, 1), g4 = rx_int(cs, m, 2), g5 = rx_int(cs, m, 3);
        int g6 = rx_int(cs, m, 4), g7 = rx_int(cs, m, 5), g8 = rx_int(cs, m, 6);
        int tslen = g4 + g5 + g6;
        int dlen = g3 + g4 + g6 + g7;
        int mid = g4 + g6;
        int refoff = pos + mch, rdoff = mch, RDOFF = mch, rm = g8;
        if (m.start > 0) {
            refoff += rx_sum_prefix(cs, m.start, "MND");
            rdoff += rx_sum_prefix(cs, m.start, "MIS");
        }
        int rn = leading_matches(rdoff, refoff);
        RDOFF += rn;
        dlen -= rn;
        tslen -= rn;
        CText nc;
        nc.clear();
        nc.append_int(RDOFF);
        nc.push('M');
        if (tslen <= 0) {
            dlen -= tslen;
            rm += tslen;
            nc.append_int(dlen);
            nc.push('D');
            nc.append_int(rm);
            nc.push('M');
        } else {
            nc.append_int(dlen);
            nc.push('D');
            nc.append_int(tslen);
            nc.push('I');
            nc.append_int(rm);
            nc.push('M');
        }
        if (mid <= 15) {
            rx_replace_first(cs, "nMnDnMnInMnDnM", false, false, nc);
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:982-1000
    RV_HD bool edge_m_indel(const RxMatch& m) {
        int tmid = rx_int(cs, m, 0), g2 = rx_int(cs, m, 1);
        char o3 = rx_ch(cs, m, 2);
        int mlen = rx_int(cs, m, 3);
        int tn = 0;
        int tslen = tmid + (o3 == 'I' ? g2 : 0);
        pos += tmid + (o3 == 'D' ? g2 : 0);
        while (tn < mlen && rf.hasNe(pos + tn, rd.s(tslen + tn))) tn++;
        tslen += tn;
        mlen -= tn;
        pos += tn;
        CText r;
        rep2(r, tslen, 'S', mlen, 'M');
        rx_replace_first(cs, "dMnxnM", true, false, r);
        return true;
    }

    // modifyCigar's indel loop (cigarModifier.cpp:261-363)
    RV_HD void indel_loop(int indel) {
        bool flag = true;
        RxMatch m;
        CText r;
        int guard = 0;
        while (flag && indel > 0) {
            if (++guard > 1000 || cs.overflow) break;
            flag = false;
            if (rx_search(cs, "nSnx", true, false, m)) {
                int a = rx_int(cs, m, 0), b = rx_int(cs, m, 1);
                char o = rx_ch(cs, m, 2);
                rep1(r, a + (o == 'I' ? b : 0), 'S');
                pos += o == 'D' ? b : 0;
                rx_splice(cs, m.start, m.end, r);
                flag = true;
            }
            if (rx_search(cs, "nxnS", false, true, m)) {
                int a = rx_int(cs, m, 0), b = rx_int(cs, m, 2);
                char o = rx_ch(cs, m, 1);
                rep1(r, b + (o == 'I' ? a : 0), 'S');
                rx_splice(cs, m.start, m.end, r);
                flag = true;
            }
            if (rx_search(cs, "nSnMnx", true, false, m)) {
                int a = rx_int(cs, m, 0), tmid = rx_int(cs, m, 1), b = rx_int(cs, m, 2);
                char o = rx_ch(cs, m, 3);
                if (tmid <= 10) {
                    rep1(r, a + tmid + (o == 'I' ? b : 0), 'S');
                    pos += tmid + (o == 'D' ? b : 0);
                    rx_splice(cs, m.start, m.end, r);
                    flag = true;
                }
            }
            if (rx_search(cs, "nxnMnS", false, true, m)) {
                int a = rx_int(cs, m, 0), tmid = rx_int(cs, m, 2), b = rx_int(cs, m, 3);
                char o = rx_ch(cs, m, 1);
                if (tmid <= 10) {
                    rep1(r, b + tmid + (o == 'I' ? a : 0), 'S');
                    rx_splice(cs, m.start, m.end, r);
                    flag = true;
                }
            }
            if (rx_search(cs, "dMnxnM", true, false, m)) flag = edge_m_indel(m);
            if (rx_search(cs, "nxdM", false, true, m)) {
                int a = rx_int(cs, m, 0), tmid = rx_int(cs, m, 2);
                char o = rx_ch(cs, m, 1);
                rep1(r, tmid + (o == 'I' ? a : 0), 'S');
                rx_splice(cs, m.start, m.end, r);
                flag = true;
            }
            RxMatch mm, tim, tdm;
            bool bm = rx_search(cs, "nMnDnMnInMnDnM", false, false, mm);
            bool btim = rx_search(cs, "nMnxnMnxnMnxnM", false, false, tim);
            bool btdm = rx_search(cs, "nMnDnMnDnMnDnM", false, false, tdm);
            if (bm)
                flag = two_del_ins(mm, flag);
            else if (btdm)
                flag = three_dels(tdm, flag);
            else if (btim)
                flag = three_indels(tim, flag);
            if (rx_search(cs, "nDnMnx?", false, false, m)) {
                int g1 = rx_int(cs, m, 0), g2 = rx_int(cs, m, 1), g3 = rx_int(cs, m, 2);
                char o = rx_ch(cs, m, 3);
                if (g2 <= 15) {
                    int dlen = g1 + g2, ilen = g2;
                    if (o == 'I') {
                        ilen += g3;
                    } else if (o == 'D') {
                        dlen += g3;
                        if (!rx_empty(m, 4)) {
                            RxMatch t = m;
                            t.ce[4] = t.ce[4] - 1;
                            ilen += rx_int(cs, t, 4);
                        }
                    }
                    rep2(r, dlen, 'D', ilen, 'I');
                    rx_replace_first(cs, "nDnMnx?", false, false, r);
                    flag = true;
                }
            }
            if (rx_search(cs, "!nInMnx?", false, false, m)) {
                char pre = rx_ch(cs, m, 0);
                if (pre != 'D' && pre != 'H') {
                    int g2 = rx_int(cs, m, 1), g3 = rx_int(cs, m, 2), g4 = rx_int(cs, m, 3);
                    char o = rx_ch(cs, m, 4);
                    if (g3 <= 15) {
                        int dlen = g3, ilen = g2 + g3;
                        if (o == 'I') {
                            ilen += g4;
                        } else if (o == 'D') {
                            dlen += g4;
                            if (!rx_empty(m, 5)) {
                                RxMatch t = m;
                                t.ce[5] = t.ce[5] - 1;
                                ilen += rx_int(cs, t, 5);
                            }
                        }
                        rep2(r, dlen, 'D', ilen, 'I');
                        rx_replace_first(cs, "nInMnx?", false, false, r);
                        flag = true;
                    }
                }
            }
            if (rx_search(cs, "nDnD", false, false, m)) {
                rep1(r, rx_int(cs, m, 0) + rx_int(cs, m, 1), 'D');
                rx_splice(cs, m.start, m.end, r);
                flag = true;
            }
            if (rx_search(cs, "nInI", false, false, m)) {
                rep1(r, rx_int(cs, m, 0) + rx_int(cs, m, 1), 'I');
                rx_splice(cs, m.start, m.end, r);
                flag = true;
            }
        }
    }

    RV_HD void tails() {
        RxMatch m;
        // ^(.*?)(\d+)M(\d+)S$  else  ^(.*?)(\d+)M$
        if (rx_search(cs, "nMnS", false, true, m))
            m_then_s(m.start, rx_int(cs, m, 0), rx_int(cs, m, 1));
        else if (rx_search(cs, "nM", false, true, m))
            three_mismatches(m.start, rx_int(cs, m, 0));
        if (rx_search(cs, "nSnM", true, false, m))
            s_then_m(rx_int(cs, m, 0), rx_int(cs, m, 1));
        else if (rx_search(cs, "nM", true, false, m))
            begin_dig_m(rx_int(cs, m, 0));
    }
};

RV_HD void cigar_text(const uint32_t* w, int n, CText& out) {
    out.clear();
    for (int i = 0; i < n; i++) {
        out.append_int(clen(w[i]));
        out.push(op_char(cop(w[i])));
    }
}

RV_HD int op_code_of(char c) {
    switch (c) {
        case 'M': return OP_M;
        case 'I': return OP_I;
        case 'D': return OP_D;
        case 'N': return OP_N;
        case 'S': return OP_S;
        case 'H': return OP_H;
        case 'P': return OP_P;
        case '=': return OP_EQ;
        case 'X': return OP_X;
        case 'B': return OP_B;
        default: return OP_M;
    }
}

// Applies the rewrite to an emulated record `rec` (CIGAR words followed by packed SEQ words, as the
// reference's in-place rewrite and its reads past the CIGAR see them).  Returns true when the
// CIGAR text changed and was written back; n_new/position are updated in that case.  `flags`
// gains bit1 when the text or the words would exceed their capacity (the caller refuses the read)
// and bit2 when a negative length restored the original.
RV_HD bool modify_cigar_record(const RefView& rf, const ReadView& rd, int& position, uint32_t* rec, int rec_cap,
                               int n_orig, int indel_len, int& n_new, int& flags) {
    // A single M run of 8 or more bases (most reads): no indel rule applies, and the tail rules
    // reduce to the 3' then the 5' end-mismatch clip ("nM" -> "aMbS" -> "cSdMbS", 1-3 bases each),
    // computed here without the CIGAR text.  Lengths stay positive, so the text path's parse has
    // nothing to refuse.
    if (n_orig == 1 && cop(rec[0]) == OP_M && clen(rec[0]) >= 8) {
        int m = (int)clen(rec[0]);
        const int rn3 = tail_mismatch_run(rf, rd, position + m, m, m);
        const int s3 = (rn3 > 0 && rn3 <= 3) ? rn3 : 0;
        m -= s3;
        const int rn5 = head_mismatch_run(rf, rd, position, m);
        const int s5 = (rn5 > 0 && rn5 <= 3) ? rn5 : 0;
        m -= s5;
        if (s3 == 0 && s5 == 0) return false;  // the text is unchanged
        const int k = 1 + (s5 > 0) + (s3 > 0);
        if (k > rec_cap) {
            flags |= 2;
            return false;
        }
        int w = 0;
        if (s5) rec[w++] = ((uint32_t)s5 << 4) | (uint32_t)OP_S;
        rec[w++] = ((uint32_t)m << 4) | (uint32_t)OP_M;
        if (s3) rec[w++] = ((uint32_t)s3 << 4) | (uint32_t)OP_S;
        n_new = w;
        position += s5;
        return true;
    }
    int base = 0;
    int n = n_orig;
    CText original;
    cigar_text(rec, n_orig, original);
    int pos = position;
    if (cop(rec[0]) == OP_D) {
        pos += clen(rec[0]);
        base = 1;
    }
    if (base + n - 1 >= rec_cap) {
        flags |= 2;
        return false;
    }
    if (rec[base + n - 1] == (uint32_t)OP_D) n--;
    if (cop(rec[base]) == OP_I) rec[base] = (rec[base] & ~0xfu) | OP_D;
    if (rec[base + n - 1] == (uint32_t)OP_I) rec[base + n - 1] = (rec[base + n - 1] & ~0xfu) | OP_D;
    Rewriter rw;
    rw.rf = rf;
    rw.rd = rd;
    rw.pos = pos;
    cigar_text(rec + base, n, rw.cs);
    rw.indel_loop(indel_len);
    rw.tails();
    if (rw.cs.overflow) {
        flags |= 2;
        return false;
    }
    if (rw.cs.eq(original)) return false;
    // cigarstr_2cigar: stoi over the text; a negative length aborts and the original is restored
    int k = 0, i = 0;
    bool ok = true;
    while (i < rw.cs.n) {
        bool neg = false;
        int j = i;
        if (rw.cs.c[j] == '-' || rw.cs.c[j] == '+') {
            neg = rw.cs.c[j] == '-';
            j++;
        }
        long long v = 0;
        while (j < rw.cs.n && is_digit(rw.cs.c[j])) v = v * 10 + (rw.cs.c[j++] - '0');
        if (neg) v = -v;
        if (v < 0) {
            ok = false;
            break;
        }
        char opc = j < rw.cs.n ? rw.cs.c[j] : '\0';
        if (base + k >= rec_cap) {
            flags |= 2;
            return false;
        }
        rec[base + k] = ((uint32_t)v << 4) | (uint32_t)op_code_of(opc);
        k++;
        i = j + 1;
    }
    if (!ok) {
        flags |= 4;
        // re-encode the original text's words at rec[base..], as cigarstr_2cigar(originalCigar) does
        int kk = 0, ii = 0;
        while (ii < original.n) {
            long long v = 0;
            int j = ii;
            while (j < original.n && is_digit(original.c[j])) v = v * 10 + (original.c[j++] - '0');
            char opc = j < original.n ? original.c[j] : '\0';
            if (base + kk >= rec_cap) {
                flags |= 2;
                return false;
            }
            rec[base + kk] = ((uint32_t)v << 4) | (uint32_t)op_code_of(opc);
            kk++;
            ii = j + 1;
        }
        return false;
    }
    n_new = k;
    position = rw.pos;
    return true;
}

}  // namespace rv
