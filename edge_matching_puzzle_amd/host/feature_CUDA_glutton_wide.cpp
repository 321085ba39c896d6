#include "feature_CUDA_glutton_wide.h"

#include <iomanip>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

namespace edge_matching_puzzle {

namespace {

// One position's capability as the reference prints it: each non-zero 32-bit
// word as a header line, the eight piece indices it covers and, for each,
// the letters of the orientations it holds.
void print_info(std::ostream &out, const uint64_t *mask, uint32_t nw) {
    static const char kSide[4] = {'N', 'E', 'S', 'W'};
    for (uint32_t w32 = 0; w32 < 2 * nw && w32 < 32; ++w32) {
        uint32_t word = (uint32_t)(mask[w32 / 2] >> (32 * (w32 % 2)));
        if (!word) continue;
        out << "/" << std::string(8, '-') << " Word[" << std::setw(2) << w32 << "] : 0x" << std::setfill('0')
            << std::setw(8) << std::hex << word << std::dec << std::setfill(' ') << " " << std::string(8, '-') << "\\"
            << std::endl;
        for (unsigned i = 0; i < 8; ++i) out << "|" << std::setw(4) << 8 * w32 + i;
        out << "|" << std::endl;
        for (unsigned i = 0; i < 8; ++i) {
            out << "|";
            for (unsigned o = 0; o < 4; ++o) out << (((word & 0xF) >> o) & 1 ? kSide[o] : ' ');
            word >>= 4;
        }
        out << "|" << std::endl;
    }
}

// A situation of `level` played moves: positions map to info indices in
// position order, the played positions dropped.
void print_situation(std::ostream &out, uint32_t level, uint32_t index, uint32_t P, uint32_t nw,
                     const uint64_t *masks, const uint64_t *theoric, const std::vector<bool> &placed) {
    out << "Situation Level " << level << ": " << std::endl;
    out << "Situation Index " << index << ": " << std::endl;
    out << "===== Position index <-> Info index =====" << std::endl;
    std::vector<uint32_t> info_pos;
    for (uint32_t p = 0; p < P; ++p) {
        if (placed[p]) {
            out << "Position[" << p << "] -> Index " << std::numeric_limits<uint32_t>::max() << std::endl;
        } else {
            out << "Position[" << p << "] -> Index " << info_pos.size() << std::endl;
            info_pos.push_back(p);
        }
    }
    for (uint32_t i = 0; i < info_pos.size(); ++i) out << "Index[" << i << "] -> Position " << info_pos[i] << std::endl;
    for (uint32_t i = 0; i < info_pos.size(); ++i) {
        out << "Info[" << i << "]:" << std::endl;
        print_info(out, masks + (size_t)info_pos[i] * nw, nw);
        out << std::endl;
    }
    for (uint32_t i = 0; i < info_pos.size(); ++i) {
        out << "Theoric info[" << i << "]:" << std::endl;
        if (theoric) print_info(out, theoric + (size_t)info_pos[i] * nw, nw);
        out << std::endl;
    }
}

}  // namespace

void feature_CUDA_glutton_wide::run() {
    if (m_depth) {
        run_levels();
        return;
    }
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    const uint32_t P = m_puzzle.width * m_puzzle.height, nw = (4 * P + 63) / 64;
    const uint32_t corner = 0;   // position of (0, 0)
    std::vector<uint64_t> level0((size_t)P * nw);
    uint32_t status0 = 0;
    const uint64_t capacity = 4 * P;
    std::vector<uint32_t> parent(capacity), move(capacity), status(capacity);
    std::vector<uint64_t> masks(capacity * P * nw);
    uint64_t n_children = 0;
    EMP_HIP_CALL(emp_hip_glutton_wide(session.get(), &puzzle, nullptr, 0, 1, &corner, level0.data(), &status0,
                                      capacity, parent.data(), move.data(), masks.data(), status.data(),
                                      &n_children));
    std::vector<bool> none(P, false);
    // level 0: capability and theoric capability are both the initial one
    print_situation(std::cout, 0, 0, P, nw, level0.data(), level0.data(), none);
    std::cout << std::endl;
    std::cout << "Corner info : " << std::endl;
    print_info(std::cout, level0.data() + (size_t)corner * nw, nw);
    std::cout << std::endl;
    for (uint64_t c = 0; c < n_children && c < capacity; ++c) {
        const uint32_t bit = 4 * ((move[c] >> 8) & 0xFF) + ((move[c] >> 16) & 3);
        std::cout << bit + 1 << std::endl;
        if (status[c] == EMP_HIP_WIDE_EMPTIED) std::cout << "A position has no option left" << std::endl;
        std::vector<bool> placed(P, false);
        placed[move[c] & 0xFF] = true;
        print_situation(std::cout, 1, (uint32_t)c, P, nw, masks.data() + c * P * nw, nullptr, placed);
        std::cout << std::endl;
    }
}

void feature_CUDA_glutton_wide::run_levels() {
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    const uint32_t n = m_puzzle.width * m_puzzle.height;
    std::cout << "Launch kernels" << std::endl;
    for (unsigned int depth = 1; depth <= m_depth; ++depth) {
        uint64_t count = 0, nodes = 0, solutions = 0;
        const uint64_t capacity = depth == m_depth ? 1024 : 0;
        std::vector<uint16_t> boards(capacity * n);
        EMP_HIP_CALL(emp_hip_expand(session.get(), &puzzle, EMP_HIP_ORDER_SPIRAL, depth, capacity,
                                    capacity ? boards.data() : nullptr, &count, &nodes, &solutions));
        std::cout << "Level " << depth << " : " << count << " situations" << std::endl;
        if (depth == m_depth)
            for (uint64_t i = 0; i < count && i < capacity; ++i)
                std::cout << situation_string(m_puzzle, boards.data() + i * n) << std::endl;
    }
}

}  // namespace edge_matching_puzzle
