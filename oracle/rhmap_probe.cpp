// ORACLE / TEST INFRASTRUCTURE ONLY.
//
// rhmap_probe: applies a sequence of operations to the oracle's robin_hood clone and prints the
// resulting table, for tests/test_oracle_rhmap.py.
//   stdin:  one op per line: "i KEY" (operator[] insert), "e KEY" (erase), "c KEY" (char-keyed
//           map: insert the byte KEY into a separate RHMap<char,int>)
//   stdout: "mask M inc I shift S" then one line per occupied slot in iteration order:
//           "slot key info" (int map), then "chars" followed by the char map's keys in order.
#include <cstdio>
#include <iostream>
#include <string>

#include "rh_map.h"

int main() {
    orc::RHMap<int, int> m;
    orc::RHMap<char, int> cm;
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
    printf("mask %zu inc %u shift %u\n", m.mask(), m.info_inc(), m.info_hash_shift());
    for (int k : m.keys_in_order()) printf("%zu %d %u\n", m.slot_of(k), k, m.info_at(m.slot_of(k)));
    printf("chars");
    for (char c : cm.keys_in_order()) printf(" %d", (int)(unsigned char)c);
    printf("\n");
    return 0;
}
