// emp_model.h -- puzzle model of the host side: pieces, kinds, kind
// indices, auto-similarity, position orders, situation strings and the seeded
// synthetic generator.
//
// Semantics follow the reference's piece database and strategy generators:
//   kinds                 include/emp_piece.h:126-131
//   orientation algebra   include/emp_piece.h:173-178 (side s of a piece laid
//                         with orientation o shows colours[(s + o) % 4])
//   auto-similarity       include/emp_piece.h:144-157, orientations kept by the
//                         constraint database include/emp_piece_db.h:1269-1274
//   kind indices          include/emp_piece_db.h:555-616 (file order per kind)
//   validation            include/emp_piece_db.h:898-924
//   spiral / basic order  include/emp_spiral_strategy_generator.h:52-98,
//                         include/emp_basic_strategy_generator.h:44-53
//   situation string      include/situation_string_formatter.h:60-81
#ifndef EMP_MODEL_H
#define EMP_MODEL_H

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace emp {

enum Kind : int { kCenter = 0, kBorder = 1, kCorner = 2 };
enum Order : int { kSpiral = 0, kBasic = 1 };

struct Piece {
    uint32_t id = 0;          // 1-based, file order
    uint32_t colour[4] = {};  // N, E, S, W; 0 is the frame
    int kind = 0;             // number of frame sides
    int similarity = 0;       // 0 none, 1 half similar (N=S, E=W), 2 similar (all equal)
    uint32_t kind_index = 0;  // rank among pieces of the same kind
};

inline uint32_t shown(const Piece &p, unsigned side, unsigned orientation) {
    return p.colour[(side + orientation) & 3u];
}

class Puzzle {
  public:
    // Builds and validates; throws std::invalid_argument with the reason.
    Puzzle(uint32_t width, uint32_t height, const uint32_t *colours);

    uint32_t width() const { return width_; }
    uint32_t height() const { return height_; }
    uint32_t size() const { return width_ * height_; }
    uint32_t max_colour() const { return max_colour_; }
    const Piece &piece(uint32_t index) const { return pieces_[index]; }
    const std::vector<uint32_t> &pieces_of(int kind) const { return by_kind_[kind]; }
    // Position kind of cell (x, y).
    int position_kind(uint32_t x, uint32_t y) const;
    // Orientation o of a piece is kept by the constraint database.
    static bool orientation_kept(const Piece &p, unsigned o) { return (o & ((1u << p.similarity) - 1u)) == 0; }

  private:
    uint32_t width_, height_, max_colour_ = 0;
    std::vector<Piece> pieces_;
    std::vector<uint32_t> by_kind_[3];
};

// Strategy order: element i is the (x, y) of the i-th visited cell.
std::vector<std::pair<uint32_t, uint32_t>> make_order(uint32_t width, uint32_t height, int order);

// True when the order visits every cell exactly once.  The reference's spiral
// walk does not when height >= width + 2 (it revisits cells).
bool is_permutation(const std::vector<std::pair<uint32_t, uint32_t>> &order, uint32_t width, uint32_t height);

// Row-major board cell value: ((id - 1) << 2) | orientation, 0xFFFF empty.
std::string situation_string(uint32_t nb_cells, const uint16_t *board);

// Seeded synthetic puzzle (random-edge model).  Frame edges colour 0; edges
// between two frame cells drawn from border colours 1..B, all other inner
// edges from interior colours B+1..B+I; interior pieces turned by a random
// quarter turn, frame pieces stored as laid in the planted solution; ids
// shuffled.  planted (optional) receives the planted solution as a board.
std::vector<uint32_t> generate(uint64_t seed, uint32_t width, uint32_t height,
                               uint32_t border_colours, uint32_t interior_colours,
                               std::vector<uint16_t> *planted);

// splitmix64 step (Steele, Lea, Flood 2014).
inline uint64_t splitmix64(uint64_t &state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace emp

#endif
