// emp_results.cpp -- results.bin v1 writer/reader and the situation string
// parser (see emp_results.h for the reference files each follows).
#include "emp_results.h"

#include <cstring>
#include <stdexcept>

namespace emp {

namespace {

uint32_t bits_for(uint32_t v) {
    uint32_t b = 0;
    while (v) ++b, v >>= 1;
    return b;
}

void put_u32(std::vector<uint8_t> &out, uint32_t v) {
    for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(v >> (8 * i)));
}

void put_u64(std::vector<uint8_t> &out, uint64_t v) {
    for (int i = 0; i < 8; ++i) out.push_back((uint8_t)(v >> (8 * i)));
}

uint32_t get_u32(const uint8_t *p) {
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)p[i] << (8 * i);
    return v;
}

uint64_t get_u64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

}  // namespace

uint32_t results_field_bits(uint32_t nb_cells, bool solution_dump) {
    // piece indices 0..P-1 for solutions, ids 0..P (0 = empty) otherwise
    return 2 + bits_for(solution_dump ? nb_cells - 1 : nb_cells);
}

std::vector<uint8_t> write_results_bin(uint32_t width, uint32_t height,
                                       const std::vector<std::pair<uint32_t, uint32_t>> &order, bool solution_dump,
                                       const uint16_t *boards, const uint64_t *node_index, uint64_t n_records,
                                       uint64_t total_nodes) {
    const uint32_t P = width * height;
    if (!P || order.size() != P) throw std::invalid_argument("results.bin: the order must list every cell");
    const uint32_t f = results_field_bits(P, solution_dump);
    const uint32_t words = (P * f + 63) / 64;
    std::vector<uint8_t> out;
    out.reserve(16 + 8 * (size_t)P + (size_t)n_records * (8 * words + 8) + 8);
    put_u32(out, 1);
    put_u32(out, width);
    put_u32(out, height);
    put_u32(out, solution_dump ? 1u : 0u);
    for (const auto &xy : order) {
        put_u32(out, xy.first);
        put_u32(out, xy.second);
    }
    std::vector<uint64_t> field(words);
    for (uint64_t r = 0; r < n_records; ++r) {
        std::fill(field.begin(), field.end(), 0);
        const uint16_t *b = boards + r * P;
        for (uint32_t i = 0; i < P; ++i) {
            const uint32_t cell = order[i].second * width + order[i].first;
            const uint16_t v = b[cell];
            uint64_t value;
            if (v == 0xFFFF) {
                if (solution_dump) throw std::invalid_argument("results.bin: a solution dump holds full boards");
                value = 0;
            } else {
                if ((uint32_t)(v >> 2) >= P) throw std::invalid_argument("results.bin: piece index out of range");
                value = solution_dump ? v : (uint64_t)((((v >> 2) + 1) << 2) | (v & 3));
            }
            const uint32_t bit = i * f, w = bit / 64, o = bit % 64;
            field[w] |= value << o;
            if (o + f > 64) field[w + 1] |= value >> (64 - o);
        }
        for (uint64_t x : field) put_u64(out, x);
        put_u64(out, node_index ? node_index[r] : 0);
    }
    put_u64(out, total_nodes);
    return out;
}

ResultsBin read_results_bin(const uint8_t *data, uint64_t size) {
    if (size < 16) throw std::invalid_argument("results.bin: file shorter than its header");
    ResultsBin r;
    const uint32_t version = get_u32(data);
    if (version != 1) throw std::invalid_argument("results.bin: version " + std::to_string(version) + " is not 1");
    r.width = get_u32(data + 4);
    r.height = get_u32(data + 8);
    r.solution_dump = get_u32(data + 12) != 0;
    const uint64_t P = (uint64_t)r.width * r.height;
    if (!P || P > 4096) throw std::invalid_argument("results.bin: bad board size");
    const uint64_t start = 16 + 8 * P;
    if (size < start + 8) throw std::invalid_argument("results.bin: file shorter than its header");
    std::vector<char> seen(P, 0);
    for (uint64_t i = 0; i < P; ++i) {
        const uint32_t x = get_u32(data + 16 + 8 * i), y = get_u32(data + 20 + 8 * i);
        if (x >= r.width || y >= r.height || seen[y * r.width + x])
            throw std::invalid_argument("results.bin: the position list is not a permutation of the cells");
        seen[y * r.width + x] = 1;
        r.order.emplace_back(x, y);
    }
    const uint32_t f = results_field_bits((uint32_t)P, r.solution_dump);
    const uint64_t words = (P * f + 63) / 64, record = 8 * words + 8;
    // as the reference reader: the last 8 bytes are the total unless the
    // records do not fill the space before them
    const uint64_t body = size - 8 - start;
    const uint64_t n = body / record;
    r.complete = body % record == 0;
    if (r.complete) r.total_nodes = get_u64(data + size - 8);
    r.boards.assign(n * P, 0xFFFF);
    r.node_index.resize(n);
    const uint64_t mask = (1ull << f) - 1;
    for (uint64_t k = 0; k < n; ++k) {
        const uint8_t *rec = data + start + k * record;
        for (uint64_t i = 0; i < P; ++i) {
            const uint64_t bit = i * f, w = bit / 64, o = bit % 64;
            uint64_t value = get_u64(rec + 8 * w) >> o;
            if (o + f > 64) value |= get_u64(rec + 8 * (w + 1)) << (64 - o);
            value &= mask;
            const uint32_t cell = r.order[i].second * r.width + r.order[i].first;
            if (r.solution_dump) {
                if ((value >> 2) >= P) throw std::invalid_argument("results.bin: piece index out of range");
                r.boards[k * P + cell] = (uint16_t)value;
            } else if (value >> 2) {
                if ((value >> 2) > P) throw std::invalid_argument("results.bin: piece id out of range");
                r.boards[k * P + cell] = (uint16_t)((((value >> 2) - 1) << 2) | (value & 3));
            }
        }
        r.node_index[k] = get_u64(rec + 8 * words);
    }
    return r;
}

std::vector<uint16_t> parse_situation(uint32_t width, uint32_t height, const std::string &text) {
    const uint32_t P = width * height;
    const uint32_t w = (uint32_t)std::to_string(P).size() + 1;
    if (text.size() != (size_t)P * w)
        throw std::invalid_argument("Real size (" + std::to_string(text.size()) + ") doesn`t match theoric size(" +
                                    std::to_string((size_t)P * w) + ")");
    std::vector<uint16_t> board(P, 0xFFFF);
    static const char kSide[4] = {'N', 'E', 'S', 'W'};
    for (uint32_t cell = 0; cell < P; ++cell) {
        const std::string id = text.substr((size_t)cell * w, w - 1);
        if (id == std::string(w - 1, '-')) continue;
        const char letter = text[(size_t)cell * w + w - 1];
        int o = -1;
        for (int s = 0; s < 4; ++s)
            if (letter == kSide[s]) o = s;
        if (o < 0) throw std::invalid_argument(std::string("Unkown short string orientation : \"") + letter + "\"");
        char *end = nullptr;
        const long v = std::strtol(id.c_str(), &end, 10);
        if (v < 1 || v > (long)P) throw std::invalid_argument("piece id \"" + id + "\" out of range 1.." + std::to_string(P));
        board[cell] = (uint16_t)(((v - 1) << 2) | o);
    }
    return board;
}

}  // namespace emp
