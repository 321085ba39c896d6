// emp_backtrack.h -- host interface of the reference-order backtracker.
#ifndef EMP_BACKTRACK_H
#define EMP_BACKTRACK_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "emp_enumerate.h"
#include "emp_model.h"

namespace emp {

struct BacktrackOutcome {
    uint64_t trace_capacity = 0;     // in: selections to record
    uint64_t max_selections = 0;     // in: stop after that many selections, 0 = none
    bool found = false;
    uint64_t trace_length = 0;       // selections made
    std::vector<uint32_t> trace;     // (lane, variable, step) per recorded selection
    std::vector<uint16_t> board;     // first solution (row-major), empty cells 0xFFFF
    double elapsed_ms = 0;
};

// Level-0 capability (P positions in strategy order x 16 u64, bit 256 * o +
// piece): the variables of create_variables (include/emp_variable_generator.h
// :336-531) set by prepare_initial (include/system_equation_for_CUDA.h
// :159-173).  Its popcount is the reference's variable count.
std::vector<uint64_t> backtrack_capability(const Puzzle &puzzle, int order);

void backtrack_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &puzzle, int order,
                         BacktrackOutcome &out);

}  // namespace emp

#endif
