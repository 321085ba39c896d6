// emp_results.h -- results.bin v1 and situation strings read back (host).
//
//   results.bin v1   include/emp_situation_binary_dumper.h:72-126 (writer),
//                    include/emp_situation_binary_reader.h (reader)
//   situation string include/situation_string_formatter.h:84-134 (parser)
#ifndef EMP_RESULTS_H
#define EMP_RESULTS_H

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace emp {

// Bits of one record field: the piece number (piece index for solution
// dumps, id with 0 = empty otherwise) and two orientation bits.
uint32_t results_field_bits(uint32_t nb_cells, bool solution_dump);

// results.bin v1: header (version 1, width, height, solution_dump, then the
// (x, y) of each strategy position), one record per board -- a bitfield of
// P fields in strategy order, LSB first in little-endian u64 words, and the
// board's node index -- then the total node count.  Boards are row-major
// ((id - 1) << 2 | o, 0xFFFF empty); a solution dump stores (id - 1) << 2 | o,
// any other dump id << 2 | o with 0 for an empty cell.
std::vector<uint8_t> write_results_bin(uint32_t width, uint32_t height,
                                       const std::vector<std::pair<uint32_t, uint32_t>> &order, bool solution_dump,
                                       const uint16_t *boards, const uint64_t *node_index, uint64_t n_records,
                                       uint64_t total_nodes);

struct ResultsBin {
    uint32_t width = 0, height = 0;
    bool solution_dump = false;
    std::vector<std::pair<uint32_t, uint32_t>> order;
    std::vector<uint16_t> boards;      // [n_records][P] row-major
    std::vector<uint64_t> node_index;  // [n_records]
    uint64_t total_nodes = 0;
    bool complete = false;             // false: truncated, no total
};

// Reads a results.bin v1 image; throws std::invalid_argument when it is not one.
ResultsBin read_results_bin(const uint8_t *data, uint64_t size);

// Row-major board of a situation string (fixed-width cells of digits(P) + 1
// characters: the id right-aligned, then N/E/S/W; '-' cells are empty).
// Throws std::invalid_argument with the reference's messages on a bad size
// or orientation letter, or for an id outside 1..P.
std::vector<uint16_t> parse_situation(uint32_t width, uint32_t height, const std::string &text);

}  // namespace emp

#endif
