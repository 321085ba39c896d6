// emp_profile.h -- host interface of the capability profiles.
#ifndef EMP_PROFILE_H
#define EMP_PROFILE_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "emp_enumerate.h"
#include "emp_model.h"

namespace emp {

// Number of glutton profiles: total, max, min, vector_op, rvector_op, each
// searched for the min then for the max (src/feature_situation_profile.cpp
// :70-76, include/feature_situation_profile.h:310-330).
constexpr uint32_t kProfileRuns = 10;

struct ProfileOutcome {
    uint32_t level = 0;                 // pieces of the situation
    uint32_t n_variables = 0;           // variables of the system equation
    std::vector<uint32_t> glutton;      // [kProfileRuns][P + 1]
    std::vector<uint16_t> surface;      // [level + 1][2P] sorted profiles
    std::vector<uint32_t> row_index;    // [level + 1] strategy index of each row (P last)
    uint64_t launches = 0;
    double kernel_ms = 0;               // profile_kernel time (HIP events)
};

// Capability profiles of a situation (row-major board, ((id - 1) << 2) | o,
// 0xFFFF empty), positions in the basic order: the reference's
// situation_profile feature (include/feature_situation_profile.h:150-483).
// Throws std::invalid_argument when a placement is not a variable.
void situation_profile_on_device(hipStream_t stream, Workspace &ws, const Puzzle &puzzle, const uint16_t *board,
                                 ProfileOutcome &out);

struct SearchOutcome {
    std::vector<uint32_t> stats;        // [6][3][P + 1]: low min/max, high min/max, total min/max; series
                                        // all / leading to a solution / others; by level
    std::vector<uint16_t> solutions;    // [n_solutions][P] row-major boards, in board order
    uint64_t n_solutions = 0;
    std::vector<uint64_t> level_sizes;  // [P + 1] situations of each level (first pass, all starts)
    std::vector<std::vector<uint16_t>> starts;   // the searches' starting situations
    uint64_t launches = 0;
    double kernel_ms = 0;               // search_children_kernel time (HIP events)
};

// The profile feature's search (include/feature_profile.h:582-868) on the
// device.  Throws Unsupported when a level holds more than max_situations.
void profile_search_on_device(hipStream_t stream, Workspace &ws, const Puzzle &puzzle, uint64_t max_situations,
                              SearchOutcome &out);

}  // namespace emp

#endif
