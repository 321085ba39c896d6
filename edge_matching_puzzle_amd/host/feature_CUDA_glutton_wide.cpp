#include "feature_CUDA_glutton_wide.h"

#include <iostream>
#include <vector>

namespace edge_matching_puzzle {

void feature_CUDA_glutton_wide::run() {
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
