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

void backtrack_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &puzzle, int order,
                         BacktrackOutcome &out);

}  // namespace emp

#endif
