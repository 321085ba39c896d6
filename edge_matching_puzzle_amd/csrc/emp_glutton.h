// emp_glutton.h -- host interface of the glutton_max path.
#ifndef EMP_GLUTTON_H
#define EMP_GLUTTON_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "emp_enumerate.h"
#include "emp_model.h"

namespace emp {

// Device tables of one puzzle: colour masks, piece colours and the initial
// capability of every position, all piece-major (bit 4 * piece + o).
struct GluttonTables {
    uint32_t P = 0, W = 0, nw = 0, max_colour = 0;
    std::vector<uint64_t> cm;          // [((c - 1) * 4 + side) * nw + w]
    std::vector<uint32_t> piece_col;   // N | E << 8 | S << 16 | W << 24
    std::vector<uint64_t> init_masks;  // [P][nw]
};
GluttonTables build_glutton_tables(const Puzzle &puzzle);

struct GluttonConfig {
    uint64_t push_budget = 0;    // pushes per search, 0 = unlimited
    uint32_t waves_per_cu = 0;   // 0 = as many as LDS allows (at most 8)
};

struct GluttonOutcome {
    uint32_t max_depth = 0;
    uint64_t pushes = 0;
    uint32_t status = 0;         // 0 solved, 1 exhausted, 2 push budget, 3 bad prefix, 4 internal
    std::vector<uint16_t> board; // full solution, or the deepest situation
};

struct GluttonStats {
    double total_ms = 0, kernel_ms = 0;
    uint64_t pushes = 0, evaluations = 0, score_evaluations = 0, rounds = 0;
    uint64_t free_slot_sum = 0;   // free slots of the level, summed over the evaluations
    uint32_t waves = 0, lds_bytes = 0;
};

// Runs one glutton_max search per prefix (prefix_len moves each, packed
// pos | piece << 8 | o << 16, applied before the search; the search never
// pops below them).
void glutton_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &puzzle,
                       const std::vector<uint32_t> &prefixes, uint32_t prefix_len, uint32_t n_items,
                       const GluttonConfig &cfg, std::vector<GluttonOutcome> &out, GluttonStats &stats);

}  // namespace emp

#endif
