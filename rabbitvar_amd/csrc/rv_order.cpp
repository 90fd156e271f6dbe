// Slot-order replay of nonInsertionVariants (see rv_order.h).
#include "rv_order.h"

namespace rv {

void replay_created(RHOrder& m, const std::vector<int32_t>& created) {
    for (int32_t p : created) m.insert(p);
}

std::vector<int32_t> noninsertion_order(const std::vector<int32_t>& created, const std::vector<KeyOp>& log) {
    RHOrder m;
    replay_created(m, created);
    return replay_finish(m, log);
}

std::vector<int32_t> replay_finish(RHOrder& m, const std::vector<KeyOp>& log) {
    for (const KeyOp& op : log) {
        if (op.insert)
            m.insert(op.pos);
        else
            m.erase(op.pos);
    }
    std::vector<int32_t> out;
    out.reserve(m.size());
    m.for_each([&](int32_t k) { out.push_back(k); });
    return out;
}

}  // namespace rv
