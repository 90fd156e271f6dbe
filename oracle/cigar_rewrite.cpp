// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// CIGAR rewrite, restated from src/cigarModifier.cpp:219-1000 and include/patterns.h:52-153.
// The reference works on the CIGAR *string* with std::regex search/replace; this restatement
// keeps that formulation on purpose (the product's HIP kernel implements the same rules on CIGAR
// tokens, so the two are independent formulations of one specification).  The regex texts are
// the rules' specification and are kept identical.
//
// Pinned by tests/test_oracle_cigar_kat.py: 35 reads, each answer (the walked CIGAR, its start and
// modifyCigar's return) worked out by hand from the reference text rule by rule, cover the indel
// loop (clip/indel folds at both ends, short-M shifts, close-indel merges into complex forms, the
// three-indel forms, the negative-length restore), the record quirks (leading or trailing I/D) and
// the tail rules (end mismatches to clips, 5' and 3' clip rescue, low-quality absorption).  The
// engine's separate token matcher runs the same reads through the CPU harness.
#include <regex>
#include <stdexcept>
#include <string>

#include "rv_oracle.h"

namespace orc {
namespace {

struct Pats {
    std::regex S_ID{"^(\\d+)S(\\d+)([ID])"};                        // patterns.h:52
    std::regex ID_S_END{"(\\d+)([ID])(\\d+)S$"};                    // :56
    std::regex S_M_ID{"^(\\d+)S(\\d+)M(\\d+)([ID])"};               // :61
    std::regex ID_M_S_END{"(\\d+)([ID])(\\d+)M(\\d+)S$"};           // :66
    std::regex DIGM_ID_M{"^(\\d)M(\\d+)([ID])(\\d+)M"};             // :70
    std::regex DIGM_ID_M_{"^\\dM\\d+[ID]\\d+M"};                    // :72
    std::regex ID_DIGM_END{"(\\d+)([ID])(\\d)M$"};                  // :76
    std::regex MDMIMD{"^(.*?)(\\d+)M(\\d+)D(\\d+)M(\\d+)I(\\d+)M(\\d+)D(\\d+)M"};   // :82
    std::regex MDMIMD_prim{"(\\d+)M(\\d+)D(\\d+)M(\\d+)I(\\d+)M(\\d+)D(\\d+)M"};     // :83
    std::regex three_del{"^(.*?)(\\d+)M(\\d+)D(\\d+)M(\\d+)D(\\d+)M(\\d+)D(\\d+)M"};  // :86
    std::regex three_indel{"^(.*?)(\\d+)M(\\d+)([DI])(\\d+)M(\\d+)([DI])(\\d+)M(\\d+)([DI])(\\d+)M"};  // :87
    std::regex three_indel_rep{"\\d+M\\d+[DI]\\d+M\\d+[DI]\\d+M\\d+[DI]\\d+M"};     // :88
    std::regex three_del_rep{"\\d+M\\d+D\\d+M\\d+D\\d+M\\d+D\\d+M"};                // :89
    std::regex D_M_DI{"(\\d+)D(\\d+)M(\\d+)([DI])(\\d+I)?"};        // :93
    std::regex I_M_DI{"(\\d+)I(\\d+)M(\\d+)([DI])(\\d+I)?"};        // :97
    std::regex nD_I_M_DI{"(\\D)(\\d+)I(\\d+)M(\\d+)([DI])(\\d+I)?"};  // :101
    std::regex D_D{"(\\d+)D(\\d+)D"};                               // :105
    std::regex I_I{"(\\d+)I(\\d+)I"};                               // :109
    std::regex ANY_M_END{"^(.*?)(\\d+)M$"};                         // :110
    std::regex M_END{"\\d+M$"};                                     // :112
    std::regex BEGIN_M{"^(\\d+)M"};                                 // :113
    std::regex S_M{"^(\\d+)S(\\d+)M"};                              // :117
    std::regex M_S_END{"\\d+M\\d+S$"};                              // :118
    std::regex ANY_M_S_END{"^(.*?)(\\d+)M(\\d+)S$"};                // :122
    std::regex LEN_MND{"(\\d+)[MND]"};                              // :140
    std::regex LEN_MIS{"(\\d+)[MIS]"};                              // :147
};

const Pats& pats() {
    static Pats p;
    return p;
}

int toi(const std::string& s) { return std::stoi(s); }
std::string S(int v) { return std::to_string(v); }

// util.h:186-199: sum of group 1 over successive matches
int sum_groups(const std::regex& re, const std::string& s) {
    int r = 0;
    std::smatch m;
    auto it = s.cbegin();
    while (std::regex_search(it, s.cend(), m, re)) {
        r += std::stoi(m.str(1));
        it = m.suffix().first;
    }
    return r;
}

std::string rep_first(const std::string& s, const std::regex& re, const std::string& by) {
    return std::regex_replace(s, re, by, std::regex_constants::format_first_only);
}
std::string rep_all(const std::string& s, const std::regex& re, const std::string& by) {
    return std::regex_replace(s, re, by);
}

bool hasEq(const RefMap& ref, int k, char c) { return ref.count(k) && ref.at(k) == c; }
bool hasNe(const RefMap& ref, int k, char c) { return ref.count(k) && ref.at(k) != c; }

struct Rewriter {
    const RefMap& ref;
    int& pos;
    std::string cs;
    const std::string& seq;
    const uint8_t* q;
    int lseq;

    // cigarModifier.cpp:412-439
    void begin_dig_m(const std::smatch& m) {
        int mch = toi(m.str(1));
        int rn = 0, rrn = 0, rmch = 0;
        while (rrn < mch && rn < mch) {
            if (!ref.count(pos + rrn)) break;
            if (hasNe(ref, pos + rrn, seq[rrn])) {
                rn = rrn + 1;
                rmch = 0;
            } else if (hasEq(ref, pos + rrn, seq[rrn])) {
                rmch++;
            }
            rrn++;
            if (rmch >= 3) break;
        }
        if (rn > 0 && rn <= 3) {
            mch -= rn;
            cs = rep_first(cs, pats().BEGIN_M, S(rn) + "S" + S(mch) + "M");
            pos += rn;
        }
    }

    // cigarModifier.cpp:445-523
    void s_then_m(const std::smatch& m) {
        int mch = toi(m.str(2));
        int soft = toi(m.str(1));
        int rn = 0;
        std::set<char> RN;
        while (rn < soft && hasEq(ref, pos - rn - 1, seq[soft - rn - 1]) && q[soft - rn - 1] > 10) rn++;
        if (rn > 0) {
            mch += rn;
            soft -= rn;
            cs = rep_first(cs, pats().S_M, soft > 0 ? S(soft) + "S" + S(mch) + "M" : S(mch) + "M");
            pos -= rn;
            rn = 0;
        }
        if (soft > 0) {
            while (rn + 1 < soft && hasEq(ref, pos - rn - 2, seq[soft - rn - 2]) && q[soft - rn - 2] > 10) {
                rn++;
                if (ref.count(pos - rn - 2)) RN.insert(ref.at(pos - rn - 2));
            }
            int rn_nt = (int)RN.size();
            if ((rn > 4 && rn_nt > 1) || hasEq(ref, pos - 1, seq[soft - 1])) {
                mch += rn + 1;
                soft -= rn + 1;
                cs = rep_first(cs, pats().S_M, soft > 0 ? S(soft) + "S" + S(mch) + "M" : S(mch) + "M");
                pos -= rn + 1;
            }
            if (rn == 0) {
                int rrn = 0, rmch = 0;
                while (rrn < mch && rn < mch) {
                    if (!ref.count(pos + rrn)) break;
                    if (hasNe(ref, pos + rrn, seq[soft + rrn])) {
                        rn = rrn + 1;
                        rmch = 0;
                    } else if (hasEq(ref, pos + rrn, seq[soft + rrn])) {
                        rmch++;
                    }
                    rrn++;
                    if (rmch >= 3) break;
                }
                if (rn > 0 && rn < mch) {
                    soft += rn;
                    mch -= rn;
                    cs = rep_first(cs, pats().S_M, S(soft) + "S" + S(mch) + "M");
                    pos += rn;
                }
            }
        }
    }

    // prefix offsets shared by the tail rules
    void prefix_offsets(const std::string& ov5, int& refoff, int& rdoff) {
        if (!ov5.empty()) {
            refoff += sum_groups(pats().LEN_MND, ov5);
            rdoff += sum_groups(pats().LEN_MIS, ov5);
        }
    }

    // walk back from the 3' end: last mismatch before 3 consecutive matches
    int tail_mismatch_run(int refoff, int rdoff, int mch) {
        int rn = 0, rrn = 0, rmch = 0;
        while (rrn < mch && rn < mch) {
            if (!ref.count(refoff - rrn - 1)) break;
            if (rrn < rdoff && ref.at(refoff - rrn - 1) != seq[rdoff - rrn - 1]) {
                rn = rrn + 1;
                rmch = 0;
            } else if (rrn < rdoff && ref.at(refoff - rrn - 1) == seq[rdoff - rrn - 1]) {
                rmch++;
            }
            rrn++;
            if (rmch >= 3) break;
        }
        return rn;
    }

    // cigarModifier.cpp:529-568
    void three_mismatches(const std::smatch& m) {
        std::string ov5 = m.str(1);
        int mch = toi(m.str(2));
        int refoff = pos + mch, rdoff = mch;
        prefix_offsets(ov5, refoff, rdoff);
        int rn = tail_mismatch_run(refoff, rdoff, mch);
        mch -= rn;
        if (rn > 0 && rn <= 3) cs = rep_first(cs, pats().M_END, S(mch) + "M" + S(rn) + "S");
    }

    // cigarModifier.cpp:574-668
    void m_then_s(const std::smatch& m) {
        std::string ov5 = m.str(1);
        int mch = toi(m.str(2));
        int soft = toi(m.str(3));
        int refoff = pos + mch, rdoff = mch;
        prefix_offsets(ov5, refoff, rdoff);
        int rn = 0;
        std::set<char> RN;
        while (rn < soft && hasEq(ref, refoff + rn, seq[rdoff + rn]) && q[rdoff + rn] > 10) rn++;
        if (rn > 0) {
            mch += rn;
            soft -= rn;
            cs = rep_first(cs, pats().M_S_END, soft > 0 ? S(mch) + "M" + S(soft) + "S" : S(mch) + "M");
            rn = 0;
        }
        if (soft > 0) {
            while (rn + 1 < soft && hasEq(ref, refoff + rn + 1, seq[rdoff + rn + 1]) && q[rdoff + rn + 1] > 10) {
                rn++;
                if (ref.count(refoff + rn + 1)) RN.insert(ref.at(refoff + rn + 1));
            }
            int rn_nt = (int)RN.size();
            if (rn > 4 && rn_nt > 1) {
                mch += rn + 1;
                soft -= rn + 1;
                cs = rep_first(cs, pats().M_S_END, soft > 0 ? S(mch) + "M" + S(soft) + "S" : S(mch) + "M");
            }
            if (rn == 0) {
                int r2 = tail_mismatch_run(refoff, rdoff, mch);
                if (r2 > 0 && r2 < mch) {
                    soft += r2;
                    mch -= r2;
                    cs = rep_first(cs, pats().M_S_END, S(mch) + "M" + S(soft) + "S");
                }
            }
        }
    }

    // cigarModifier.cpp:675-707
    bool close_to_one(const std::smatch& m, bool flag) {
        std::string op = m.str(5);
        int g2 = toi(m.str(2)), g3 = toi(m.str(3)), g4 = toi(m.str(4));
        if (g3 <= 15) {
            int dlen = g3, ilen = g2 + g3;
            if (op == "I") {
                ilen += g4;
            } else if (op == "D") {
                dlen += g4;
                std::string istr = m.str(6);
                if (!istr.empty()) ilen += toi(istr.substr(0, istr.size() - 1));
            }
            cs = rep_first(cs, pats().I_M_DI, S(dlen) + "D" + S(ilen) + "I");
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:717-744
    bool close_to_correct(const std::smatch& m, bool flag) {
        int g1 = toi(m.str(1)), g2 = toi(m.str(2)), g3 = toi(m.str(3));
        if (g2 <= 15) {
            std::string op = m.str(4);
            int dlen = g1 + g2, ilen = g2;
            if (op == "I") {
                ilen += g3;
            } else if (op == "D") {
                dlen += g3;
                std::string istr = m.str(5);
                if (!istr.empty()) ilen += toi(istr.substr(0, istr.size() - 1));
            }
            cs = rep_first(cs, pats().D_M_DI, S(dlen) + "D" + S(ilen) + "I");
            flag = true;
        }
        return flag;
    }

    int leading_matches(int rdoff, int refoff) {
        int rn = 0;
        while (rdoff + rn < lseq && hasEq(ref, refoff + rn, seq[rdoff + rn])) rn++;
        return rn;
    }

    // cigarModifier.cpp:753-845
    bool three_indels(const std::smatch& m, bool flag) {
        int mid = toi(m.str(5)) + toi(m.str(8));
        int tslen = mid;
        if (m.str(4) == "I") tslen += toi(m.str(3));
        if (m.str(7) == "I") tslen += toi(m.str(6));
        if (m.str(10) == "I") tslen += toi(m.str(9));
        int dlen = mid;
        if (m.str(4) == "D") dlen += toi(m.str(3));
        if (m.str(7) == "D") dlen += toi(m.str(6));
        if (m.str(10) == "D") dlen += toi(m.str(9));
        std::string ov5 = m.str(1);
        int refoff = pos + toi(m.str(2));
        int rdoff = toi(m.str(2));
        int RDOFF = toi(m.str(2));
        int rm = toi(m.str(11));
        prefix_offsets(ov5, refoff, rdoff);
        int rn = leading_matches(rdoff, refoff);
        RDOFF += rn;
        dlen -= rn;
        tslen -= rn;
        std::string nc = S(RDOFF) + "M";
        if (tslen <= 0) {
            dlen -= tslen;
            rm += tslen;
            if (dlen == 0) {
                RDOFF = RDOFF + rm;
                nc = S(RDOFF) + "M";
            } else if (dlen < 0) {
                tslen = -dlen;
                rm += dlen;
                if (rm < 0) {
                    RDOFF = RDOFF + rm;
                    nc = S(RDOFF) + "M" + S(tslen) + "I";
                } else {
                    nc += S(tslen) + "I" + S(rm) + "M";
                }
            } else {
                nc += S(dlen) + "D" + S(rm) + "M";
            }
        } else {
            if (dlen == 0) {
                nc += S(tslen) + "I" + S(rm) + "M";
            } else if (dlen < 0) {
                rm += dlen;
                nc += S(tslen) + "I" + S(rm) + "M";
            } else {
                nc += S(dlen) + "D" + S(tslen) + "I" + S(rm) + "M";
            }
        }
        if (mid <= 15) {
            cs = rep_first(cs, pats().three_indel_rep, nc);
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:854-908
    bool three_dels(const std::smatch& m, bool flag) {
        int tslen = toi(m.str(4)) + toi(m.str(6));
        int dlen = toi(m.str(3)) + toi(m.str(4)) + toi(m.str(5)) + toi(m.str(6)) + toi(m.str(7));
        int mid = toi(m.str(4)) + toi(m.str(6));
        std::string ov5 = m.str(1);
        int refoff = pos + toi(m.str(2));
        int rdoff = toi(m.str(2));
        int RDOFF = toi(m.str(2));
        int rm = toi(m.str(8));
        prefix_offsets(ov5, refoff, rdoff);
        int rn = leading_matches(rdoff, refoff);
        RDOFF += rn;
        dlen -= rn;
        tslen -= rn;
        std::string nc = S(RDOFF) + "M";
        if (tslen <= 0) {
            dlen -= tslen;
            rm += tslen;
            nc += S(dlen) + "D" + S(rm) + "M";
        } else {
            nc += S(dlen) + "D" + S(tslen) + "I" + S(rm) + "M";
        }
        if (mid <= 15) {
            cs = rep_first(cs, pats().three_del_rep, nc);
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:918-973
    bool two_del_ins(const std::smatch& m, bool flag) {
        int tslen = toi(m.str(4)) + toi(m.str(5)) + toi(m.str(6));
        int dlen = toi(m.str(3)) + toi(m.str(4)) + toi(m.str(6)) + toi(m.str(7));
        int mid = toi(m.str(4)) + toi(m.str(6));
        std::string ov5 = m.str(1);
        int mch = toi(m.str(2));
        int refoff = pos + mch, rdoff = mch, RDOFF = mch;
        int rm = toi(m.str(8));
        prefix_offsets(ov5, refoff, rdoff);
        int rn = leading_matches(rdoff, refoff);
        RDOFF += rn;
        dlen -= rn;
        tslen -= rn;
        std::string nc = S(RDOFF) + "M";
        if (tslen <= 0) {
            dlen -= tslen;
            rm += tslen;
            nc += S(dlen) + "D" + S(rm) + "M";
        } else {
            nc += S(dlen) + "D" + S(tslen) + "I" + S(rm) + "M";
        }
        if (mid <= 15) {
            cs = rep_first(cs, pats().MDMIMD_prim, nc);
            flag = true;
        }
        return flag;
    }

    // cigarModifier.cpp:982-1000
    bool edge_m_indel(const std::smatch& m) {
        int tmid = toi(m.str(1));
        int mlen = toi(m.str(4));
        int tn = 0;
        int tslen = tmid + (m.str(3) == "I" ? toi(m.str(2)) : 0);
        pos += tmid + (m.str(3) == "D" ? toi(m.str(2)) : 0);
        while (tn < mlen && hasNe(ref, pos + tn, seq[tslen + tn])) tn++;
        tslen += tn;
        mlen -= tn;
        pos += tn;
        cs = rep_first(cs, pats().DIGM_ID_M_, S(tslen) + "S" + S(mlen) + "M");
        return true;
    }

    // the indel loop of modifyCigar (cigarModifier.cpp:261-363)
    void indel_loop(int indel) {
        const Pats& P = pats();
        std::smatch m;
        bool flag = true;
        while (flag && indel > 0) {
            flag = false;
            if (std::regex_search(cs, m, P.S_ID)) {
                std::string t = S(toi(m.str(1)) + (m.str(3) == "I" ? toi(m.str(2)) : 0)) + "S";
                pos += m.str(3) == "D" ? toi(m.str(2)) : 0;
                cs = rep_all(cs, P.S_ID, t);
                flag = true;
            }
            if (std::regex_search(cs, m, P.ID_S_END)) {
                std::string t = S(toi(m.str(3)) + (m.str(2) == "I" ? toi(m.str(1)) : 0)) + "S";
                cs = rep_all(cs, P.ID_S_END, t);
                flag = true;
            }
            if (std::regex_search(cs, m, P.S_M_ID)) {
                int tmid = toi(m.str(2));
                if (tmid <= 10) {
                    std::string t = S(toi(m.str(1)) + tmid + (m.str(4) == "I" ? toi(m.str(3)) : 0)) + "S";
                    pos += tmid + (m.str(4) == "D" ? toi(m.str(3)) : 0);
                    cs = rep_all(cs, P.S_M_ID, t);
                    flag = true;
                }
            }
            if (std::regex_search(cs, m, P.ID_M_S_END)) {
                int tmid = toi(m.str(3));
                if (tmid <= 10) {
                    std::string t = S(toi(m.str(4)) + tmid + (m.str(2) == "I" ? toi(m.str(1)) : 0)) + "S";
                    cs = rep_all(cs, P.ID_M_S_END, t);
                    flag = true;
                }
            }
            if (std::regex_search(cs, m, P.DIGM_ID_M)) flag = edge_m_indel(m);
            if (std::regex_search(cs, m, P.ID_DIGM_END)) {
                int tmid = toi(m.str(3));
                std::string t = S(tmid + (m.str(2) == "I" ? toi(m.str(1)) : 0)) + "S";
                cs = rep_all(cs, P.ID_DIGM_END, t);
                flag = true;
            }
            std::smatch sm_tim, sm_tdm;
            bool mm = std::regex_search(cs, m, P.MDMIMD);
            bool tim = std::regex_search(cs, sm_tim, P.three_indel);
            bool tdm = std::regex_search(cs, sm_tdm, P.three_del);
            if (mm)
                flag = two_del_ins(m, flag);
            else if (tdm)
                flag = three_dels(sm_tdm, flag);
            else if (tim)
                flag = three_indels(sm_tim, flag);
            if (std::regex_search(cs, m, P.D_M_DI)) flag = close_to_correct(m, flag);
            if (std::regex_search(cs, m, P.nD_I_M_DI) && m.str(1) != "D" && m.str(1) != "H")
                flag = close_to_one(m, flag);
            if (std::regex_search(cs, m, P.D_D)) {
                cs = rep_first(cs, P.D_D, S(toi(m.str(1)) + toi(m.str(2))) + "D");
                flag = true;
            }
            if (std::regex_search(cs, m, P.I_I)) {
                cs = rep_first(cs, P.I_I, S(toi(m.str(1)) + toi(m.str(2))) + "I");
                flag = true;
            }
        }
    }

    void tails() {
        const Pats& P = pats();
        std::smatch m;
        if (std::regex_search(cs, m, P.ANY_M_S_END))
            m_then_s(m);
        else if (std::regex_search(cs, m, P.ANY_M_END))
            three_mismatches(m);
        if (std::regex_search(cs, m, P.S_M))
            s_then_m(m);
        else if (std::regex_search(cs, m, P.BEGIN_M))
            begin_dig_m(m);
    }
};

std::string cigar_text(const uint32_t* w, int n) {
    std::string s;
    static const char* ops = "MIDNSHP=XB";
    for (int i = 0; i < n; i++) {
        s += std::to_string(clen(w[i]));
        int o = cop(w[i]);
        s += (o < 10) ? ops[o] : '?';
    }
    return s;
}

// cigarstr_2cigar (include/VariationUtils.h:608-629): false on a negative length
bool parse_cigar_text(std::string s, uint32_t* out, int& n) {
    n = 0;
    while (!s.empty()) {
        size_t p = 0;
        int num = std::stoi(s, &p);
        if (num < 0) return false;
        char c = s[p];
        int op = 0;
        switch (c) {
            case 'M': op = OP_M; break;
            case 'I': op = OP_I; break;
            case 'D': op = OP_D; break;
            case 'N': op = OP_N; break;
            case 'S': op = OP_S; break;
            case 'H': op = OP_H; break;
            case 'P': op = OP_P; break;
            case '=': op = OP_EQ; break;
            case 'X': op = OP_X; break;
            case 'B': op = 9; break;
            default: op = 0;
        }
        out[n++] = ((uint32_t)num << 4) | (uint32_t)op;
        s = s.substr(p + 1);
    }
    return true;
}

}  // namespace

std::string cigar_to_text(const uint32_t* w, int n) { return cigar_text(w, n); }

// Applies modifyCigar to an emulated BAM record.  `rec` holds the record's CIGAR words followed
// by its packed SEQ words (so reads one word past the CIGAR see SEQ bytes, as in the
// reference).  `base` becomes 1 when a leading D is cut (the reference advances its CIGAR
// pointer, cigarModifier.cpp:247-250).  Returns true when the text changed and the new CIGAR
// was written at rec[base..base+n_new); n_new and position are updated then.
bool modify_cigar_record(const RefMap& ref, int& position, std::vector<uint32_t>& rec, int n_orig,
                         int& base, int& n_new, const std::string& seq, const uint8_t* qual, int lseq,
                         int indelLen) {
    base = 0;
    int n = n_orig;
    const std::string original = cigar_text(rec.data(), n_orig);
    int pos = position;
    if (cop(rec[0]) == OP_D) {
        pos += clen(rec[0]);
        base = 1;
    }
    if (rec[base + n - 1] == (uint32_t)OP_D) n--;
    if (cop(rec[base]) == OP_I) rec[base] = (rec[base] & ~0xfu) | OP_D;
    if (rec[base + n - 1] == (uint32_t)OP_I) rec[base + n - 1] = (rec[base + n - 1] & ~0xfu) | OP_D;
    Rewriter rw{ref, pos, cigar_text(rec.data() + base, n), seq, qual, lseq};
    rw.indel_loop(indelLen);
    rw.tails();
    if (rw.cs == original) return false;
    std::vector<uint32_t> tmp(rw.cs.size() + 4);
    int nn = 0;
    std::string copy = rw.cs;
    // cigarstr_2cigar writes word by word into the record; emulate partial writes on failure
    {
        std::string s = copy;
        int k = 0;
        bool ok = true;
        while (!s.empty()) {
            size_t p = 0;
            int num = std::stoi(s, &p);
            if (num < 0) {
                ok = false;
                break;
            }
            uint32_t w;
            int one = 0;
            std::string t = s.substr(0, p + 1);
            parse_cigar_text(t, &w, one);
            if ((size_t)(base + k) >= rec.size()) throw std::runtime_error("cigar overflow");
            rec[base + k] = w;
            k++;
            s = s.substr(p + 1);
        }
        if (!ok) {
            int k2 = 0;
            std::vector<uint32_t> o(original.size() + 4);
            parse_cigar_text(original, o.data(), k2);
            for (int i = 0; i < k2; i++) rec[base + i] = o[i];
            return false;
        }
        nn = k;
    }
    (void)tmp;
    n_new = nn;
    position = pos;
    return true;
}

}  // namespace orc
