#include "feature_CUDA_backtracker.h"

#include <iostream>
#include <vector>

namespace edge_matching_puzzle {

feature_CUDA_backtracker::feature_CUDA_backtracker(const puzzle_description &p_puzzle, const std::string &p_strategy)
    : m_puzzle(p_puzzle) {
    if (p_strategy == "basic") m_order = EMP_HIP_ORDER_BASIC;
    else if (p_strategy == "spiral") m_order = EMP_HIP_ORDER_SPIRAL;
    else throw logic_exception("Unsupported strategy \"" + p_strategy + "\"", __LINE__, __FILE__);
}

void feature_CUDA_backtracker::run() {
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    const uint32_t n = m_puzzle.width * m_puzzle.height;
    std::vector<uint16_t> board(n);
    const uint64_t capacity = 1 << 20;
    std::vector<uint32_t> trace(3 * capacity);
    uint32_t found = 0;
    uint64_t length = 0;
    uint64_t variables = 0;
    EMP_HIP_CALL(emp_hip_backtrack_variables(&puzzle, m_order, &variables));
    // as prepare_transitions reports them (include/system_equation_for_CUDA.h:81-82)
    std::cout << "Max variables nb: " << n * 256 * 4 << std::endl;
    std::cout << "Nb variables: " << variables << std::endl;
    std::cout << "Launch kernels" << std::endl;
    EMP_HIP_CALL(emp_hip_backtrack_trace(session.get(), &puzzle, m_order, 0, board.data(), &found, trace.data(),
                                         capacity, &length));
    for (uint64_t i = 0; i < length && i < capacity; ++i)
        std::cout << "Thread " << trace[3 * i] << " : selected variable " << trace[3 * i + 1] << " at step "
                  << trace[3 * i + 2] << std::endl;
    if (found) std::cout << situation_string(m_puzzle, board.data()) << std::endl;
    else std::cout << "No solution" << std::endl;
}

void feature_CUDA_enumerate::run() {
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    emp_hip_enum_options options;
    emp_hip_enum_default_options(&options);
    options.order = m_order;
    options.max_solutions = 4096;
    const uint32_t n = m_puzzle.width * m_puzzle.height;
    std::vector<uint16_t> boards(options.max_solutions * n);
    uint16_t *board_ptr = boards.data();
    emp_hip_enum_result result;
    emp_hip_enum_stats stats;
    std::cout << "Launch kernels" << std::endl;
    EMP_HIP_CALL(emp_hip_enumerate(session.get(), &puzzle, 1, &options, &result, &board_ptr, &stats));
    for (uint64_t i = 0; i < result.n_stored; ++i)
        std::cout << situation_string(m_puzzle, boards.data() + i * n) << std::endl;
    std::cout << "End of algorithm" << std::endl;
    std::cout << "Total situations explored : " << result.n_nodes << std::endl;
    std::cout << "Nb solutions : " << result.n_solutions << std::endl;
    std::cout << "Total elapsed time : " << stats.total_ms / 1000.0 << "s" << std::endl;
}

}  // namespace edge_matching_puzzle
