#include "feature_new_strategy.h"

#include <fstream>
#include <iostream>
#include <vector>

namespace edge_matching_puzzle {

feature_new_strategy::feature_new_strategy(const puzzle_description &p_puzzle, const std::string &p_strategy,
                                           const std::string &p_dump_file, const std::string &p_initial_situation)
    : m_puzzle(p_puzzle), m_dump_file(p_dump_file), m_initial_situation(p_initial_situation) {
    if (p_strategy == "basic") m_order = EMP_HIP_ORDER_BASIC;
    else if (p_strategy == "spiral") m_order = EMP_HIP_ORDER_SPIRAL;
    else throw logic_exception("Unsupported strategy \"" + p_strategy + "\"", __LINE__, __FILE__);
}

void feature_new_strategy::run() {
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    const uint32_t n = m_puzzle.width * m_puzzle.height;
    emp_hip_item item{};
    std::vector<uint32_t> moves(n);
    uint32_t n_items = 0;
    if (!m_initial_situation.empty()) {
        std::cout << "Setting initial state @ \"" << m_initial_situation << "\"" << std::endl;
        std::vector<uint16_t> board(n);
        EMP_HIP_CALL(emp_hip_parse_situation(m_puzzle.width, m_puzzle.height, m_initial_situation.c_str(),
                                             board.data()));
        EMP_HIP_CALL(emp_hip_resume_item(&puzzle, m_order, board.data(), &item, moves.data(), &n_items));
        // the situation the search resumes from: its leading positions
        std::vector<uint32_t> xy(2 * n);
        EMP_HIP_CALL(emp_hip_order(m_puzzle.width, m_puzzle.height, m_order, xy.data()));
        std::vector<uint16_t> shown(n, 0xFFFF);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t cell = xy[2 * i + 1] * m_puzzle.width + xy[2 * i];
            if (board[cell] == 0xFFFF) break;
            shown[cell] = board[cell];
        }
        std::cout << situation_string(m_puzzle, shown.data()) << std::endl;
    }
    emp_hip_enum_options options;
    emp_hip_enum_default_options(&options);
    options.order = m_order;
    emp_hip_enum_result result;
    emp_hip_enum_stats stats;
    std::vector<uint16_t> boards;
    std::vector<uint64_t> node_index;
    // every solution goes to the dump: a second pass when the first store
    // was too small
    for (uint64_t capacity = 1 << 12;;) {
        options.max_solutions = capacity;
        boards.assign(capacity * n, 0);
        node_index.assign(capacity, 0);
        EMP_HIP_CALL(emp_hip_enumerate_items(session.get(), &puzzle, &options, n_items ? &item : nullptr,
                                             n_items ? moves.data() : nullptr, n_items, &result, boards.data(),
                                             node_index.data(), nullptr, nullptr, 0, nullptr, &stats));
        if (result.n_stored == result.n_solutions) break;
        capacity = result.n_solutions;
    }
    uint64_t size = 0;
    EMP_HIP_CALL(emp_hip_results_bin_write(m_puzzle.width, m_puzzle.height, m_order, 1, boards.data(),
                                           node_index.data(), result.n_stored, result.n_nodes, nullptr, 0, &size));
    std::vector<uint8_t> bytes(size);
    EMP_HIP_CALL(emp_hip_results_bin_write(m_puzzle.width, m_puzzle.height, m_order, 1, boards.data(),
                                           node_index.data(), result.n_stored, result.n_nodes, bytes.data(), size,
                                           &size));
    std::ofstream file(m_dump_file, std::ofstream::binary);
    if (!file) throw logic_exception("Unable to create file \"" + m_dump_file + "\"", __LINE__, __FILE__);
    file.write(reinterpret_cast<const char *>(bytes.data()), (std::streamsize)bytes.size());
    std::cout << "End of algorithm" << std::endl;
    std::cout << "Total situations explored : " << result.n_nodes << std::endl;
    std::cout << "Nb solutions : " << result.n_solutions << std::endl;
}

}  // namespace edge_matching_puzzle
