// emp_model.cpp -- host puzzle model (see emp_model.h for the reference
// citations of each rule).
#include "emp_model.h"

#include <algorithm>
#include <stdexcept>

namespace emp {

Puzzle::Puzzle(uint32_t width, uint32_t height, const uint32_t *colours)
    : width_(width), height_(height) {
    if (!colours) throw std::invalid_argument("null colour array");
    if (width < 2 || height < 2) throw std::invalid_argument("puzzle sides must be at least 2");
    if (width * height > 1024) throw std::invalid_argument("puzzle larger than 1024 pieces");
    const uint32_t n = width * height;
    pieces_.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        Piece &p = pieces_[i];
        p.id = i + 1;
        int zeros = 0;
        for (unsigned s = 0; s < 4; ++s) {
            p.colour[s] = colours[4 * i + s];
            zeros += p.colour[s] == 0;
            max_colour_ = std::max(max_colour_, p.colour[s]);
        }
        if (zeros > 2)
            throw std::invalid_argument("Unsupported kind of piece " + std::to_string(p.id));
        if (zeros == 2) {
            bool adjacent = false;
            for (unsigned s = 0; s < 4; ++s) adjacent |= p.colour[s] == 0 && p.colour[(s + 1) & 3] == 0;
            if (!adjacent)
                throw std::invalid_argument("Unable to find a corner orientation for piece " + std::to_string(p.id));
        }
        p.kind = zeros;
        if (p.colour[0] == p.colour[2] && p.colour[1] == p.colour[3])
            p.similarity = p.colour[0] == p.colour[1] ? 2 : 1;
        p.kind_index = (uint32_t)by_kind_[p.kind].size();
        by_kind_[p.kind].push_back(i);
    }
    if (by_kind_[kCorner].size() != 4 || by_kind_[kBorder].size() != 2 * (width - 2 + height - 2) ||
        by_kind_[kCenter].size() != (width - 2) * (height - 2))
        throw std::invalid_argument("piece kinds do not match a " + std::to_string(width) + "x" +
                                    std::to_string(height) + " board");
    std::vector<uint32_t> per_colour(max_colour_ + 1, 0);
    for (const Piece &p : pieces_)
        for (unsigned s = 0; s < 4; ++s)
            if (p.colour[s]) ++per_colour[p.colour[s]];
    uint64_t total = 0;
    for (uint32_t c = 1; c <= max_colour_; ++c) {
        if (per_colour[c] & 1)
            throw std::invalid_argument("Number of pieces edge with color should be multiple of 2 or the puzzle cannot be solved");
        total += per_colour[c];
    }
    uint64_t edges = (uint64_t)(width - 1) * height + (uint64_t)(height - 1) * width;
    if (total != 2 * edges)
        throw std::invalid_argument("Incoherency between number of edges (" + std::to_string(edges) +
                                    ") and color on edges (" + std::to_string(total) + ")");
}

int Puzzle::position_kind(uint32_t x, uint32_t y) const {
    bool xb = x == 0 || x == width_ - 1, yb = y == 0 || y == height_ - 1;
    return (xb && yb) ? kCorner : ((xb || yb) ? kBorder : kCenter);
}

std::vector<std::pair<uint32_t, uint32_t>> make_order(uint32_t width, uint32_t height, int order) {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    out.reserve(width * height);
    if (order == kBasic) {
        for (uint32_t y = 0; y < height; ++y)
            for (uint32_t x = 0; x < width; ++x) out.emplace_back(x, y);
        return out;
    }
    // Clockwise walk: along the top row, down the right column, back along
    // the bottom row, up the left column, then inward; the next cell is
    // chosen by comparing x and y exactly as the reference's spiral does.
    std::vector<uint8_t> taken(width * height, 0);
    uint32_t x = 0, y = 0;
    auto free_cell = [&](uint32_t cx, uint32_t cy) { return !taken[cy * width + cx]; };
    for (uint32_t i = 0; i < width * height; ++i) {
        taken[y * width + x] = 1;
        out.emplace_back(x, y);
        if (i + 1 == width * height) break;
        if (x >= y) {
            if (x + 1 < width && free_cell(x + 1, y)) ++x;
            else if (y + 1 < height && free_cell(x, y + 1)) ++y;
            else --x;
        } else {
            if (x > 0 && free_cell(x - 1, y)) --x;
            else if (y > 0 && free_cell(x, y - 1)) --y;
            else ++x;
        }
    }
    return out;
}

bool is_permutation(const std::vector<std::pair<uint32_t, uint32_t>> &order, uint32_t width, uint32_t height) {
    if (order.size() != (size_t)width * height) return false;
    std::vector<uint8_t> seen(width * height, 0);
    for (const auto &xy : order) {
        if (xy.first >= width || xy.second >= height) return false;
        uint8_t &s = seen[xy.second * width + xy.first];
        if (s) return false;
        s = 1;
    }
    return true;
}

std::string situation_string(uint32_t nb_cells, const uint16_t *board) {
    const unsigned w = (unsigned)std::to_string(nb_cells).size() + 1;
    static const char kSide[4] = {'N', 'E', 'S', 'W'};
    std::string out;
    out.reserve(nb_cells * w);
    for (uint32_t i = 0; i < nb_cells; ++i) {
        if (board[i] == 0xFFFF) {
            out.append(w, '-');
            continue;
        }
        std::string id = std::to_string((board[i] >> 2) + 1);
        out.append(w - 1 - id.size(), ' ');
        out += id;
        out += kSide[board[i] & 3];
    }
    return out;
}

std::vector<uint32_t> generate(uint64_t seed, uint32_t width, uint32_t height,
                               uint32_t border_colours, uint32_t interior_colours,
                               std::vector<uint16_t> *planted) {
    if (width < 2 || height < 2 || width * height > 1024)
        throw std::invalid_argument("generator: unsupported board size");
    if (border_colours < 1 || interior_colours < 1 || border_colours + interior_colours > 255)
        throw std::invalid_argument("generator: bad colour counts");
    uint64_t state = seed;
    // Multiply-shift reduction of the high 32 bits onto [0, n).
    auto uniform = [&](uint32_t n) { return (uint32_t)(((splitmix64(state) >> 32) * (uint64_t)n) >> 32); };
    const uint32_t n = width * height;
    // Horizontal edge (x,y)-(x+1,y) and vertical edge (x,y)-(x,y+1).
    std::vector<uint32_t> hedge((width - 1) * height), vedge(width * (height - 1));
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x + 1 < width; ++x) {
            bool frame = y == 0 || y == height - 1;
            hedge[y * (width - 1) + x] = frame ? 1 + uniform(border_colours)
                                               : 1 + border_colours + uniform(interior_colours);
        }
    for (uint32_t y = 0; y + 1 < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            bool frame = x == 0 || x == width - 1;
            vedge[y * width + x] = frame ? 1 + uniform(border_colours)
                                         : 1 + border_colours + uniform(interior_colours);
        }
    std::vector<uint32_t> cell_colours(4 * n);
    std::vector<uint32_t> cell_orientation(n, 0);
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            uint32_t c[4];
            c[0] = y > 0 ? vedge[(y - 1) * width + x] : 0;
            c[1] = x + 1 < width ? hedge[y * (width - 1) + x] : 0;
            c[2] = y + 1 < height ? vedge[y * width + x] : 0;
            c[3] = x > 0 ? hedge[y * (width - 1) + x - 1] : 0;
            uint32_t cell = y * width + x;
            bool interior = x > 0 && y > 0 && x + 1 < width && y + 1 < height;
            uint32_t r = interior ? uniform(4) : 0;
            // The piece is stored turned by r quarter turns; laid with
            // orientation (4 - r) % 4 it shows c again.
            for (unsigned s = 0; s < 4; ++s) cell_colours[4 * cell + s] = c[(s + r) & 3];
            cell_orientation[cell] = (4 - r) & 3;
        }
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    for (uint32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[uniform(i + 1)]);
    std::vector<uint32_t> colours(4 * n);
    if (planted) planted->assign(n, 0xFFFF);
    for (uint32_t id0 = 0; id0 < n; ++id0) {
        uint32_t cell = perm[id0];
        for (unsigned s = 0; s < 4; ++s) colours[4 * id0 + s] = cell_colours[4 * cell + s];
        if (planted) (*planted)[cell] = (uint16_t)((id0 << 2) | cell_orientation[cell]);
    }
    return colours;
}

}  // namespace emp
