// Matchers for the descriptor patterns the realigner applies to variant descriptors
// (src/VariationRealigner.cpp:519-543, :737-770, :805-846, :1080-1112; rule texts in
// include/patterns.h).  The reference uses std::regex (ECMAScript, leftmost match, greedy
// quantifiers); each function below gives the same match and captures for every input, by the
// argument in its comment.  The two patterns with an SV part ("<...>") keep std::regex, and only
// run when the descriptor contains '<' (this path never makes such descriptors).
#pragma once
#include <cstdlib>
#include <regex>
#include <string>

namespace rv {
namespace desc {

inline bool dig(char c) { return c >= '0' && c <= '9'; }
inline bool atgc(char c) { return c == 'A' || c == 'T' || c == 'G' || c == 'C'; }
inline bool atgnc(char c) { return atgc(c) || c == 'N'; }
inline size_t run(const std::string& s, size_t i, bool (*cls)(char)) {
    while (i < s.size() && cls(s[i])) i++;
    return i;
}

// "^-(\d+).*": the greedy digit run is maximal, and ".*" takes the rest
inline bool begin_minus_number(const std::string& s, std::string& num) {
    if (s.size() < 2 || s[0] != '-' || !dig(s[1])) return false;
    num = s.substr(1, run(s, 1, dig) - 1);
    return true;
}
// "^-\d+(.*)": the same digit run; the capture is what follows it
inline bool begin_minus_number_any(const std::string& s, std::string& rest) {
    if (s.size() < 2 || s[0] != '-' || !dig(s[1])) return false;
    rest = s.substr(run(s, 1, dig));
    return true;
}
// "\^(\d+)$": only the last '^' can be followed by digits up to the end
inline bool up_number_end(const std::string& s, std::string& num) {
    const size_t c = s.rfind('^');
    if (c == std::string::npos || c + 1 >= s.size() || run(s, c + 1, dig) != s.size()) return false;
    num = s.substr(c + 1);
    return true;
}
// "\^([ATGNC]+)": the leftmost '^' followed by one of ATGNC; the capture is the whole run
inline bool caret_atgnc(const std::string& s, std::string& seq) {
    for (size_t i = 0; i + 1 < s.size(); i++)
        if (s[i] == '^' && atgnc(s[i + 1])) {
            seq = s.substr(i + 1, run(s, i + 1, atgnc) - i - 1);
            return true;
        }
    return false;
}
// regex_replace(s, "\^|&|#", ""): every '^', '&' and '#' removed
inline std::string strip_marks3(const std::string& s) {
    std::string o;
    o.reserve(s.size());
    for (char c : s)
        if (c != '^' && c != '&' && c != '#') o += c;
    return o;
}
// "(-\d+)&[ATGC]+$": the leftmost '-' whose maximal digit run is followed by '&' and then only
// ATGC (at least one) to the end; a shorter digit run would leave a digit where '&' must be
inline bool minus_number_amp_atgcs_end(const std::string& s, std::string& head) {
    for (size_t i = 0; i + 1 < s.size(); i++) {
        if (s[i] != '-' || !dig(s[i + 1])) continue;
        const size_t e = run(s, i + 1, dig);
        if (e < s.size() && s[e] == '&' && e + 1 < s.size() && run(s, e + 1, atgc) == s.size()) {
            head = s.substr(i, e - i);
            return true;
        }
    }
    return false;
}
// "^\+([ATGC]+).*"
inline bool begin_plus_atgc(const std::string& s, std::string& seq) {
    if (s.size() < 2 || s[0] != '+' || !atgc(s[1])) return false;
    seq = s.substr(1, run(s, 1, atgc) - 1);
    return true;
}
// ".*&([ATGC]+).*": the greedy ".*" backs off from the end, so the match uses the last '&' that is
// followed by an ATGC base; the capture is the whole ATGC run after it
inline bool amp_atgc(const std::string& s, std::string& seq) {
    for (size_t i = s.size(); i-- > 0;)
        if (s[i] == '&' && i + 1 < s.size() && atgc(s[i + 1])) {
            seq = s.substr(i + 1, run(s, i + 1, atgc) - i - 1);
            return true;
        }
    return false;
}
// "#([ATGC]+).*": the leftmost '#' followed by an ATGC base, and its whole run
inline bool hash_atgc(const std::string& s, std::string& seq) {
    for (size_t i = 0; i + 1 < s.size(); i++)
        if (s[i] == '#' && atgc(s[i + 1])) {
            seq = s.substr(i + 1, run(s, i + 1, atgc) - i - 1);
            return true;
        }
    return false;
}
// "\^([ATGC]+)$": only the last '^' can be followed by ATGC up to the end
inline bool caret_atgc_end(const std::string& s, std::string& seq) {
    const size_t c = s.rfind('^');
    if (c == std::string::npos || c + 1 >= s.size() || run(s, c + 1, atgc) != s.size()) return false;
    seq = s.substr(c + 1);
    return true;
}
// "(\+[ATGC]+)&[ATGC]+$": the leftmost '+' whose maximal ATGC run is followed by '&' and ATGC to
// the end
inline bool plus_atgcs_amp_atgcs_end(const std::string& s, std::string& head) {
    for (size_t i = 0; i + 1 < s.size(); i++) {
        if (s[i] != '+' || !atgc(s[i + 1])) continue;
        const size_t e = run(s, i + 1, atgc);
        if (e < s.size() && s[e] == '&' && e + 1 < s.size() && run(s, e + 1, atgc) == s.size()) {
            head = s.substr(i, e - i);
            return true;
        }
    }
    return false;
}
// regex_replace(first only) with "^\+", "\^\d+$" and "\^" and an empty replacement
inline void drop_lead_plus(std::string& s) {
    if (!s.empty() && s[0] == '+') s.erase(0, 1);
}
inline void drop_caret_digits_end(std::string& s) {
    const size_t c = s.rfind('^');
    if (c != std::string::npos && c + 1 < s.size() && run(s, c + 1, dig) == s.size()) s.erase(c);
}
inline void drop_first_caret(std::string& s) {
    const size_t c = s.find('^');
    if (c != std::string::npos) s.erase(c, 1);
}

// the SV forms: std::regex, only when the descriptor has an SV part
inline bool minus_sv(const std::string& s, std::string& inv5, std::string& inv3) {
    if (s.find('<') == std::string::npos) return false;
    static const std::regex re{"^-\\d+\\^([ATGNC]+)<...\\d+>([ATGNC]+)$"};
    std::smatch m;
    if (!std::regex_search(s, m, re)) return false;
    inv5 = m.str(1);
    inv3 = m.str(2);
    return true;
}
inline bool dup_num_atgc(const std::string& s, std::string& num, std::string& seq) {
    if (s.find('<') == std::string::npos) return false;
    static const std::regex re{"<dup(\\d+)>([ATGC]+)$"};
    std::smatch m;
    if (!std::regex_search(s, m, re)) return false;
    num = m.str(1);
    seq = m.str(2);
    return true;
}

}  // namespace desc
}  // namespace rv
