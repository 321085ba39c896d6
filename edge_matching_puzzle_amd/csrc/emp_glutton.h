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

// A search's stack, as the reference's CUDA_glutton_stack_XML_converter
// saves it (include/CUDA_glutton_stack_XML_converter.h:188-202), plus the
// level the search never pops below.  State words: [0] size S (= P), [1]
// pieces P, [2] level, [3] max, [4] base, then info_to_position [S],
// position_to_info [P], played_info [S], max_played_info [S] (each move
// pos | piece << 8 | o << 16) and the available pieces [8] (bit p of word
// p / 32).  Masks: S(S+1)/2 slot masks of nw u64 (bit 4 * piece + o), level l
// at offset (S(S+1) - (S-l)(S-l+1)) / 2 * nw holding S - l slots; levels
// above `level` are zero.
uint32_t glutton_state_words(uint32_t P);
uint64_t glutton_mask_words(uint32_t P);
// Throws std::invalid_argument unless the stack is one a search of this
// puzzle can be in (sizes, indices in range, free positions and slots in
// correspondence, placed moves on placed positions, no bit past 4P).
void validate_glutton_stack(const Puzzle &puzzle, const uint32_t *state, const uint64_t *masks);

// Host buffers of stack checkpoints, n_items stacks each (may be null).
struct GluttonCheckpoints {
    const uint32_t *in_state = nullptr;   // resume from these stacks instead of playing prefixes
    const uint64_t *in_masks = nullptr;
    uint32_t *out_state = nullptr;        // each search's stack when it ends
    uint64_t *out_masks = nullptr;
};

// Runs one glutton_max search per prefix (prefix_len moves each, packed
// pos | piece << 8 | o << 16, applied before the search; the search never
// pops below them), or per saved stack (ck.in_*, prefix_len 0).
void glutton_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &puzzle,
                       const std::vector<uint32_t> &prefixes, uint32_t prefix_len, uint32_t n_items,
                       const GluttonConfig &cfg, std::vector<GluttonOutcome> &out, GluttonStats &stats,
                       const GluttonCheckpoints &ck = GluttonCheckpoints{});

}  // namespace emp

#endif
