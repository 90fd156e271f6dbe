// ORACLE / TEST INFRASTRUCTURE ONLY -- never linked into the product library.
//
// RHMap<K,V>: reproduces the slot placement, and therefore the iteration order, of
// robin_hood::unordered_map<K,V> (robin_hood 3.4.3, MIT, vendored by the reference as
// include/robin_hood.h) for integral keys on x86-64.  The reference's output line order and the
// tie-breaking of its soft-clip consensus both depend on that order, so the oracle has to
// replay it exactly.
//
// Behaviour restated from the published robin_hood 3.4.3 algorithm (reference copy:
// include/robin_hood.h):
//   hash_int (umul128 mix)                      robin_hood.h:699-705
//   keyToIdx (identity mixer, info byte)        robin_hood.h:1177-1189
//   doCreateByKey / doInsert                    robin_hood.h:1862-1960
//   insert_move (rehash path)                   robin_hood.h:1277-1319
//   shiftUp / shiftDown                         robin_hood.h:1206-1243
//   erase(key)                                  robin_hood.h:1695-1713
//   increase_size / try_increase_info           robin_hood.h:1972-2019
//   init_data / calcNumElementsWithBuffer       robin_hood.h:1765-1800
//   iteration = ascending slot index            robin_hood.h:1123-1162
// MaxLoadFactor100 = 80 (robin_hood's default), InitialNumElements = 8, 5 initial info bits.
//
// Parity: pinned by tests/test_oracle_rhmap.py (structural invariants: every key findable after
// arbitrary insert/erase mixes, iteration visits each key once, slot of a lone key equals
// hash_int(key) & 7, and the displacement bookkeeping matches a brute-force linear-probe model).
#pragma once
#include <cstdint>
#include <cstddef>
#include <stdexcept>
#include <utility>
#include <vector>

namespace orc {

inline uint64_t rh_hash_int(uint64_t obj) {
    const unsigned __int128 p = (unsigned __int128)obj * (unsigned __int128)0xde5fb9d2630458e9ULL;
    return (uint64_t)(p >> 64) + (uint64_t)p;
}

template <class K, class V>
class RHMap {
  public:
    static constexpr size_t npos = (size_t)-1;

    RHMap() = default;

    size_t size() const { return num_; }
    bool empty() const { return num_ == 0; }

    // find: slot index or npos
    size_t find_slot(K key) const {
        if (mask_ == 0 && info_.empty()) return npos;
        size_t idx;
        uint32_t info;
        key_to_idx(key, idx, info);
        do {
            if (info == info_[idx] && keys_[idx] == key) return idx;
            idx++;
            info += inc_;
            if (info == info_[idx] && keys_[idx] == key) return idx;
            idx++;
            info += inc_;
        } while (info <= info_[idx]);
        return npos;
    }
    bool count(K key) const { return find_slot(key) != npos; }
    V* find(K key) {
        size_t s = find_slot(key);
        return s == npos ? nullptr : &vals_[s];
    }
    V& at(K key) {
        size_t s = find_slot(key);
        if (s == npos) throw std::out_of_range("RHMap::at");
        return vals_[s];
    }
    const V& at(K key) const {
        size_t s = find_slot(key);
        if (s == npos) throw std::out_of_range("RHMap::at");
        return vals_[s];
    }
    // operator[] / doCreateByKey: default-constructs the value when the key is new
    V& operator[](K key) { return vals_[create(key, V(), false)]; }
    // insert(value_type): leaves an existing value untouched
    bool insert(K key, const V& v) {
        size_t before = num_;
        create(key, v, true);
        return num_ != before;
    }
    size_t erase(K key) {
        if (mask_ == 0 && info_.empty()) return 0;
        size_t idx;
        uint32_t info;
        key_to_idx(key, idx, info);
        do {
            if (info == info_[idx] && keys_[idx] == key) {
                shift_down(idx);
                --num_;
                return 1;
            }
            idx++;
            info += inc_;
        } while (info <= info_[idx]);
        return 0;
    }

    // Iteration in robin_hood order (ascending slot index).
    template <class F>
    void for_each(F&& f) {
        if (num_ == 0) return;
        size_t n = with_buffer(mask_ + 1);
        for (size_t i = 0; i < n; ++i)
            if (info_[i] != 0) f(keys_[i], vals_[i]);
    }
    std::vector<K> keys_in_order() const {
        std::vector<K> out;
        if (num_ == 0) return out;
        size_t n = with_buffer(mask_ + 1);
        for (size_t i = 0; i < n; ++i)
            if (info_[i] != 0) out.push_back(keys_[i]);
        return out;
    }
    // slot of a key and table state (tests)
    size_t slot_of(K key) const { return find_slot(key); }
    unsigned info_at(size_t slot) const { return info_[slot]; }
    size_t mask() const { return mask_; }
    unsigned info_inc() const { return inc_; }
    unsigned info_hash_shift() const { return hash_shift_; }

  private:
    std::vector<K> keys_;
    std::vector<V> vals_;
    std::vector<uint8_t> info_;
    size_t mask_ = 0;
    size_t num_ = 0;
    size_t max_allowed_ = 0;
    uint32_t inc_ = 32;
    uint32_t hash_shift_ = 59;

    static size_t calc_max(size_t n) { return n * 80 / 100; }
    static size_t with_buffer(size_t n) {
        size_t m = calc_max(n);
        return n + (m < 0xFF ? m : 0xFF);
    }

    void key_to_idx(K key, size_t& idx, uint32_t& info) const {
        uint64_t h = rh_hash_int((uint64_t)(int64_t)key);
        info = inc_ + (uint32_t)(h >> hash_shift_);
        idx = (size_t)(h & mask_);
    }

    void init_data(size_t n) {
        num_ = 0;
        mask_ = n - 1;
        max_allowed_ = calc_max(n);
        size_t wb = with_buffer(n);
        keys_.assign(wb, K());
        vals_.assign(wb, V());
        info_.assign(wb + 8, 0);
        info_[wb] = 1;  // sentinel
        inc_ = 32;
        hash_shift_ = 59;
    }

    void shift_up(size_t idx, size_t ins) {
        while (idx != ins) {
            size_t prev = idx - 1;
            keys_[idx] = keys_[prev];
            vals_[idx] = vals_[prev];
            info_[idx] = (uint8_t)(info_[prev] + inc_);
            if ((uint32_t)info_[idx] + inc_ > 0xFF) max_allowed_ = 0;
            idx = prev;
        }
    }

    void shift_down(size_t idx) {
        size_t next = idx + 1;
        while (info_[next] >= 2 * inc_) {
            info_[idx] = (uint8_t)(info_[next] - inc_);
            keys_[idx] = keys_[next];
            vals_[idx] = vals_[next];
            idx = next;
            next = idx + 1;
        }
        info_[idx] = 0;
        keys_[idx] = K();
        vals_[idx] = V();
    }

    bool try_increase_info() {
        if (inc_ <= 2) return false;
        inc_ >>= 1;
        ++hash_shift_;
        size_t words = with_buffer(mask_ + 1) / 8;
        for (size_t w = 0; w < words; ++w)
            for (size_t b = 0; b < 8; ++b) {
                uint8_t& x = info_[w * 8 + b];
                x = (uint8_t)((x >> 1) & 0x7f);
            }
        max_allowed_ = calc_max(mask_ + 1);
        return true;
    }

    void insert_move(K key, const V& val) {
        if (max_allowed_ == 0 && !try_increase_info()) throw std::overflow_error("rh overflow");
        size_t idx;
        uint32_t info;
        key_to_idx(key, idx, info);
        while (info <= info_[idx]) {
            idx++;
            info += inc_;
        }
        size_t ins = idx;
        uint8_t ins_info = (uint8_t)info;
        if ((uint32_t)ins_info + inc_ > 0xFF) max_allowed_ = 0;
        while (info_[idx] != 0) {
            idx++;
            info += inc_;
        }
        if (idx != ins) shift_up(idx, ins);
        keys_[ins] = key;
        vals_[ins] = val;
        info_[ins] = ins_info;
        ++num_;
    }

    void rehash(size_t nb) {
        std::vector<K> ok = std::move(keys_);
        std::vector<V> ov = std::move(vals_);
        std::vector<uint8_t> oi = std::move(info_);
        size_t owb = with_buffer(mask_ + 1);
        init_data(nb);
        for (size_t i = 0; i < owb; ++i)
            if (oi[i] != 0) insert_move(ok[i], ov[i]);
    }

    void increase_size() {
        if (mask_ == 0) {
            init_data(8);
            return;
        }
        size_t mx = calc_max(mask_ + 1);
        if (num_ < mx && try_increase_info()) return;
        if (num_ * 2 < calc_max(mask_ + 1)) throw std::overflow_error("rh bad hash");
        rehash((mask_ + 1) * 2);
    }

    size_t create(K key, const V& v, bool /*is_insert*/) {
        while (true) {
            if (info_.empty()) {  // unallocated map: robin_hood's dummy info byte is 0
                increase_size();
                continue;
            }
            size_t idx;
            uint32_t info;
            key_to_idx(key, idx, info);
            while (info < info_[idx]) {
                idx++;
                info += inc_;
            }
            while (info == info_[idx]) {
                if (keys_[idx] == key) return idx;
                idx++;
                info += inc_;
            }
            if (num_ >= max_allowed_) {
                increase_size();
                continue;
            }
            size_t ins = idx;
            uint32_t ins_info = info;
            if (ins_info + inc_ > 0xFF) max_allowed_ = 0;
            while (info_[idx] != 0) {
                idx++;
                info += inc_;
            }
            if (idx != ins) shift_up(idx, ins);
            keys_[ins] = key;
            vals_[ins] = v;
            info_[ins] = (uint8_t)ins_info;
            ++num_;
            return ins;
        }
    }
};

}  // namespace orc
