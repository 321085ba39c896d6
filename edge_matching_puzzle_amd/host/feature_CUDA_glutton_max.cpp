#include "feature_CUDA_glutton_max.h"

#include <iostream>
#include <vector>

namespace edge_matching_puzzle {

void feature_CUDA_glutton_max::run() {
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    const uint32_t n = m_puzzle.width * m_puzzle.height;
    std::vector<uint16_t> board(n);
    emp_hip_glutton_options options{0, 0, 0};
    emp_hip_glutton_result result;
    std::cout << "Launch kernels" << std::endl;
    EMP_HIP_CALL(emp_hip_glutton_max(session.get(), &puzzle, nullptr, 0, 1, &options, &result, board.data(), nullptr));
    if (result.status == EMP_HIP_GLUTTON_SOLVED) {
        std::cout << "Situation with stack played info:" << std::endl;
        std::cout << situation_string(m_puzzle, board.data()) << std::endl;
    } else {
        // The reference prints "Empty stack" when the search is exhausted.
        std::cout << "Empty stack" << std::endl;
        std::cout << "Max depth " << result.max_depth << ":" << std::endl;
        std::cout << situation_string(m_puzzle, board.data()) << std::endl;
    }
}

}  // namespace edge_matching_puzzle
