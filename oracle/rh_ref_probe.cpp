// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// rh_ref_probe: the same operations as rhmap_probe, applied to the reference's own robin_hood
// hash map (robin_hood 3.4.3, the reference's include/robin_hood.h, compiled unmodified from
// where it lies by `make ref`; the binary goes to oracle/_ref/).  It is the ground truth for the
// iteration orders the oracle's clone (rh_map.h) and the engine's slot-order model replay.
//   stdin:  one op per line: "i KEY" (operator[] on robin_hood::unordered_map<int,int>),
//           "e KEY" (erase), "c KEY" (operator[] on robin_hood::unordered_map<char,int>)
//   stdout: "keys" then the int map's keys in iteration order, one per line; then a line
//           "chars" followed by the char map's keys in iteration order.
#include <cstdio>
#include <iostream>
#include <limits>
#include <string>

#include "robin_hood.h"

int main() {
    robin_hood::unordered_map<int, int> m;
    robin_hood::unordered_map<char, int> cm;
    std::ios::sync_with_stdio(false);
    std::string op;
    long long key;
    while (std::cin >> op >> key) {
        if (op == "i")
            m[(int)key] += 1;
        else if (op == "e")
            m.erase((int)key);
        else if (op == "c")
            cm[(char)key] += 1;
    }
    std::string out = "keys\n";
    for (const auto& kv : m) out += std::to_string(kv.first) + "\n";
    out += "chars";
    for (const auto& kv : cm) out += " " + std::to_string((int)(unsigned char)kv.first);
    out += "\n";
    fwrite(out.data(), 1, out.size(), stdout);
    return 0;
}
