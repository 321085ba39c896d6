// emp_enum_plan.h -- per-puzzle table consumed by the enumeration kernels.
//
// A table is one contiguous blob of u64 words (header first) that the DFS
// kernel copies into LDS (single puzzle) or reads through the cache (batches
// of puzzles).  For each position kind K the candidate space is the set of
// transitions t = (kind_index << 2) | o of pieces of that kind, a bitfield of
// words[K] u64 words; this is the reference's per-kind constraint bitfield
// (include/emp_piece_db.h:1266-1312, include/emp_position_strategy.h:172-187).
//
// Instead of the reference's constraint database indexed by a packed code of
// four colours, the candidates of a position are the AND of one colour mask
// per constrained side: cm[K][side][colour] = transitions whose side `side`
// shows `colour` (colour 0 = off-board frame).  The result is identical:
// exactly the transitions matching every constrained side.
#ifndef EMP_ENUM_PLAN_H
#define EMP_ENUM_PLAN_H

#include <cstdint>
#include <vector>

#include "emp_model.h"

namespace emp {

// Position record (one u64 per strategy index):
//   bits 0-1            kind of the position
//   bits 2+11s .. +8    side s: strategy index of the neighbour when it is
//                       placed earlier, kOffBoard, or kUnconstrained
//   bits 11+11s .. +1   kind of that neighbour
constexpr uint32_t kOffBoard = 0x1FE;
constexpr uint32_t kUnconstrained = 0x1FF;

// Neighbour record (kNbrWords u64 per strategy index), the form the kernels
// use.  Word 0: number n of neighbours placed earlier (0..4) in bits 0-2, the
// position's kind in bits 3-4, the offset of the kind's words in a lane's
// available-transition region in bits 8-23, and the u64 offset of the kind's
// class-union row in bits 32-63.  Words 1..4, the i-th such neighbour (i < n):
//   bits 0-15   byte offset of its level in a wave's LDS transition stack
//               (index * 64 lanes * bytes per transition)
//   bits 16-31  byte offset in the colour bytes of the row (neighbour kind,
//               side facing this position)
//   bits 32-63  byte offset of the colour-mask block of (this kind, the side
//               facing the neighbour): 8 * (off_cm[K] + side * ncol * row)
// (the neighbour's strategy index is the stack offset / (64 * stack_bytes))
// Frame sides are folded into the position's base mask instead.
constexpr uint32_t kNbrWords = 5;

struct EnumTableHeader {
    uint32_t P, W, H, ncol;      // ncol = max colour + 1 (colour 0 = frame)
    uint32_t words[3];           // u64 words holding a kind's transitions
    uint32_t ndup[3];            // appearance classes with >= 2 transitions
    uint32_t off_pos;            // u64 offsets into the table blob
    uint32_t off_cm[3];          // [4 sides][ncol][row]
    uint32_t off_valid[3];       // [row] orientations kept per piece
    uint32_t off_dup[3];         // [ndup[K]][row], then the union of the classes [row]
    uint32_t off_rot[3];         // u32 per transition: colours N|E<<8|S<<16|W<<24
    uint32_t root;               // transition placed at index 0
    uint32_t total_words;        // size of the blob in u64
    uint32_t avail_words;        // words[0] + words[1] + words[2]
    uint32_t row;                // words of every mask row (>= each words[K]; padding is zero)
    uint32_t off_base;           // [P][row]: valid transitions of the position's kind that match its frame sides
    uint32_t off_nbr;            // [P][kNbrWords] neighbour records
    uint32_t off_ones;           // 16 all-ones words, then 16 all-zeros words
    uint32_t off_cls[3];         // u16 per transition: its appearance class, 0xFFFF for none
    uint32_t off_col;            // colour bytes: per kind, per side, one byte per transition
    // Identical pieces.  A piece slot is abase(K) * 16 + kind_index, abase(K)
    // being the offset of kind K's words in a lane's available-transition
    // region (the slot's transitions are the 4 bits at slot * 4 of the region).
    // Pieces of one kind that share their appearance classes form a chain,
    // lowest kind_index first.
    uint32_t off_succ;           // u16 per slot: 0, or the next piece of the chain's slot | its valid orientations << 12
    uint32_t off_pred;           // u32 per chained piece after the first, highest slot first: slot | previous piece's slot << 16
    uint32_t npred;              // entries at off_pred
    uint32_t stack_bytes;        // bytes per transition in the stacks the neighbour records address
};
static_assert(sizeof(EnumTableHeader) % 8 == 0, "header must be u64 aligned");
constexpr uint32_t kHeaderWords = sizeof(EnumTableHeader) / 8;

struct EnumTable {
    std::vector<uint64_t> blob;
    std::vector<std::pair<uint32_t, uint32_t>> order;
    std::vector<uint16_t> piece_of[3];   // kind_index -> piece index (id - 1)
    EnumTableHeader header() const;
};

// u64 words needed by the largest kind's transitions (4 per piece).
uint32_t natural_row_words(const Puzzle &puzzle);

// Builds the table with mask rows of max(row_words, natural_row_words) words;
// throws std::invalid_argument when the first position has no candidate.
// stack_bytes: bytes per level in the byte plane of the kernels' LDS stacks
// (1, the default; rows wider than four words keep the transitions' two high
// bits in a second plane).
EnumTable build_enum_table(const Puzzle &puzzle, int order, uint32_t row_words = 0, uint32_t stack_bytes = 0);

// Converts a stack of transitions (one per strategy index) to a row-major board.
void stack_to_board(const EnumTable &table, const uint16_t *stack, uint16_t *board);

// Transitions of the moves (pos | piece << 8 | o << 16) placed at strategy
// indices 0..len-1, each checked to be a candidate of the enumeration at its
// index after the earlier ones; throws std::invalid_argument otherwise.
void moves_to_stack(const EnumTable &table, const Puzzle &puzzle, const uint32_t *moves, uint32_t len,
                    std::vector<uint16_t> &stack);

// Transition of the move (pos | piece << 8 | o << 16) at strategy index
// `index`: the move must be at that index's position with a piece of its
// kind (std::invalid_argument otherwise); it need not be a candidate.
uint16_t move_to_transition(const EnumTable &table, const Puzzle &puzzle, uint32_t index, uint32_t move);

// Move (pos | piece << 8 | o << 16) of transition `tr` at strategy index `index`.
uint32_t transition_to_move(const EnumTable &table, uint32_t index, uint32_t tr);

// Random spiral prefixes: up to `count` distinct rows, each the fixed moves
// followed by a placement of strategy indices 0..length-1 in spiral order,
// drawn by seeded random dives over matching pieces (pieces used by the fixed
// moves are excluded).  Moves are packed pos | piece << 8 | o << 16.
void spiral_prefixes(const Puzzle &puzzle, uint32_t length, uint32_t count, uint64_t seed,
                     const std::vector<uint32_t> &fixed, std::vector<uint32_t> &rows, uint32_t &produced);

}  // namespace emp

#endif
