// emp_glutton_wide.h -- host interface of glutton_wide's breadth step.
#ifndef EMP_GLUTTON_WIDE_H
#define EMP_GLUTTON_WIDE_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "emp_enumerate.h"
#include "emp_model.h"

namespace emp {

struct WideOutcome {
    std::vector<uint64_t> parent_masks;   // [n_parents][P][nw], 0 at placed positions
    std::vector<uint32_t> parent_status;  // 0 valid, 1 a free position without option, 2 unplayable move
    std::vector<uint32_t> child_parent;   // per child, in parent then bit order
    std::vector<uint32_t> child_move;     // pos | piece << 8 | o << 16
    std::vector<uint64_t> child_masks;    // [n_children][P][nw]
    std::vector<uint32_t> child_status;   // 0 valid, 1 a free position without option
};

// Replays each parent's moves (len per parent) from the level-0 capability
// and expands it at cells[p] over every bit of that position's mask.
void glutton_wide_on_device(hipStream_t stream, Workspace &ws, const Puzzle &puzzle,
                            const std::vector<uint32_t> &parents, uint32_t len, uint32_t n_parents,
                            const std::vector<uint32_t> &cells, WideOutcome &out);

}  // namespace emp

#endif
