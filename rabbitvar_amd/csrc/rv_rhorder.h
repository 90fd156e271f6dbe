// Slot-order model of robin_hood::unordered_map<int, V> (robin_hood 3.4.3 as vendored by the
// reference in include/robin_hood.h) for replaying the reference's map iteration orders: the
// simple-mode printer walks alignedVariants in slot order (src/modes/simpleMode.cpp:146), and
// ToVarsBuilder fills alignedVariants walking nonInsertionVariants in slot order
// (src/ToVarsBuilder.cpp:97-205).  Only keys and info bytes are kept; values do not affect slots.
//
// Algorithm (robin_hood.h): hash_int umul128 mix :699-705; keyToIdx :1177-1189 (index = low hash
// bits & mask, info = inc + hash >> shift); insertion walks past larger and equal info bytes and
// shifts the run up :1862-1960; erase shifts the run down :1695-1713, :1225-1243; growth doubles
// the table at 80% load and re-inserts in slot order :1277-1319, :1997-2019; an info byte that
// would overflow first halves the info increment (try_increase_info :1972-1995).
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace rv {

class RHOrder {
  public:
    size_t size() const { return num_; }
    uint32_t info_increment() const { return inc_; }  // 32 until an overflow halves it

    // operator[] / insert: adds the key when absent
    void insert(int32_t key) {
        for (;;) {
            if (info_.empty()) {
                grow();
                continue;
            }
            size_t idx;
            uint32_t info;
            place(key, idx, info);
            while (info < info_[idx]) {
                ++idx;
                info += inc_;
            }
            while (info == info_[idx]) {
                if (keys_[idx] == key) return;
                ++idx;
                info += inc_;
            }
            if (num_ >= max_allowed_) {
                grow();
                continue;
            }
            put(key, idx, info);
            return;
        }
    }

    void erase(int32_t key) {
        if (info_.empty()) return;
        size_t idx;
        uint32_t info;
        place(key, idx, info);
        do {
            if (info == info_[idx] && keys_[idx] == key) {
                size_t next = idx + 1;
                while (info_[next] >= 2 * inc_) {
                    info_[idx] = (uint8_t)(info_[next] - inc_);
                    keys_[idx] = keys_[next];
                    idx = next++;
                }
                info_[idx] = 0;
                keys_[idx] = 0;
                --num_;
                return;
            }
            ++idx;
            info += inc_;
        } while (info <= info_[idx]);
    }

    // hint: the slots `key` will probe (inserting a known key sequence, prefetching ahead hides
    // the table's cache misses)
    void prefetch(int32_t key) const {
        if (info_.empty()) return;
        const size_t idx = (size_t)(mix((uint64_t)(int64_t)key) & mask_);
        __builtin_prefetch(&info_[idx]);
        __builtin_prefetch(&keys_[idx]);
    }

    // slot of `key` (iteration order is slot order), -1 when absent (robin_hood findIdx)
    int64_t find_slot(int32_t key) const {
        if (info_.empty()) return -1;
        size_t idx;
        uint32_t info;
        place(key, idx, info);
        do {
            if (info == info_[idx] && keys_[idx] == key) return (int64_t)idx;
            ++idx;
            info += inc_;
        } while (info <= info_[idx]);
        return -1;
    }

    // keys in iteration (slot) order
    template <class F>
    void for_each(F&& f) const {
        if (num_ == 0) return;
        const size_t n = slots(mask_ + 1);
        for (size_t i = 0; i < n; ++i)
            if (info_[i]) f(keys_[i]);
    }

  private:
    std::vector<int32_t> keys_;
    std::vector<uint8_t> info_;
    size_t mask_ = 0, num_ = 0, max_allowed_ = 0;
    uint32_t inc_ = 32, shift_ = 59;

    static uint64_t mix(uint64_t k) {
        const unsigned __int128 p = (unsigned __int128)k * (unsigned __int128)0xde5fb9d2630458e9ULL;
        return (uint64_t)(p >> 64) + (uint64_t)p;
    }
    static size_t max_load(size_t n) { return n * 80 / 100; }
    static size_t slots(size_t n) {
        size_t m = max_load(n);
        return n + (m < 0xFF ? m : 0xFF);
    }
    void place(int32_t key, size_t& idx, uint32_t& info) const {
        const uint64_t h = mix((uint64_t)(int64_t)key);
        info = inc_ + (uint32_t)(h >> shift_);
        idx = (size_t)(h & mask_);
    }
    // insert at idx (the first slot whose resident has a smaller info byte), shifting the run up
    void put(int32_t key, size_t idx, uint32_t info) {
        const size_t at = idx;
        const uint32_t at_info = info;
        if (at_info + inc_ > 0xFF) max_allowed_ = 0;
        while (info_[idx] != 0) {
            ++idx;
            info += inc_;
        }
        while (idx != at) {
            keys_[idx] = keys_[idx - 1];
            info_[idx] = (uint8_t)(info_[idx - 1] + inc_);
            if ((uint32_t)info_[idx] + inc_ > 0xFF) max_allowed_ = 0;
            --idx;
        }
        keys_[at] = key;
        info_[at] = (uint8_t)at_info;
        ++num_;
    }
    void reset(size_t n) {
        num_ = 0;
        mask_ = n - 1;
        max_allowed_ = max_load(n);
        const size_t s = slots(n);
        keys_.assign(s, 0);
        info_.assign(s + 8, 0);
        info_[s] = 1;  // sentinel
        inc_ = 32;
        shift_ = 59;
    }
    bool halve_info() {
        if (inc_ <= 2) return false;
        inc_ >>= 1;
        ++shift_;
        const size_t words = slots(mask_ + 1) / 8;
        for (size_t i = 0; i < words * 8; ++i) info_[i] = (uint8_t)((info_[i] >> 1) & 0x7f);
        max_allowed_ = max_load(mask_ + 1);
        return true;
    }
    void grow() {
        if (mask_ == 0 && info_.empty()) {
            reset(8);
            return;
        }
        if (num_ < max_load(mask_ + 1) && halve_info()) return;
        if (num_ * 2 < max_load(mask_ + 1)) throw std::overflow_error("robin_hood order model: hash overflow");
        std::vector<int32_t> ok;
        ok.swap(keys_);
        std::vector<uint8_t> oi;
        oi.swap(info_);
        const size_t old = slots(mask_ + 1);
        reset((mask_ + 1) * 2);
        for (size_t i = 0; i < old; ++i)
            if (oi[i]) {
                if (max_allowed_ == 0 && !halve_info()) throw std::overflow_error("robin_hood order model: overflow");
                size_t idx;
                uint32_t info;
                place(ok[i], idx, info);
                while (info <= info_[idx]) {
                    ++idx;
                    info += inc_;
                }
                put(ok[i], idx, info);
            }
    }
};

}  // namespace rv
