#include "feature_situation_profile.h"

#include <cctype>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

namespace edge_matching_puzzle {

namespace {

std::string capitalised(const std::string &s) { return std::string(1, (char)std::toupper(s[0])) + s.substr(1); }

}  // namespace

feature_situation_profile::feature_situation_profile(const puzzle_description &p_puzzle,
                                                     const std::string &p_initial_situation)
    : m_puzzle(p_puzzle), m_initial_situation(p_initial_situation) {
    // the reference draws 10,000 random situations from the clock when no
    // situation is given (include/feature_situation_profile.h:185-288); only
    // the profile of a given situation is provided
    if (m_initial_situation.empty())
        throw logic_exception("situation_profile needs an --initial_situation", __LINE__, __FILE__);
}

void feature_situation_profile::run() {
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    const uint32_t W = m_puzzle.width, H = m_puzzle.height, P = W * H;
    std::vector<uint16_t> board(P);
    EMP_HIP_CALL(emp_hip_parse_situation(W, H, m_initial_situation.c_str(), board.data()));
    uint32_t level = 0, n_variables = 0;
    std::vector<uint32_t> glutton((size_t)10 * (P + 1)), surface((size_t)(P + 1) * 2 * P), row_index(P + 1);
    EMP_HIP_CALL(emp_hip_situation_profile(session.get(), &puzzle, board.data(), &level, glutton.data(),
                                           surface.data(), row_index.data(), &n_variables));
    // prepare_transitions' report (include/system_equation_for_CUDA.h:81-82)
    std::cout << "Max variables nb: " << P * 1024 << std::endl;
    std::cout << "Nb variables: " << n_variables << std::endl;
    std::cout << level << std::endl;
    const std::string root = std::to_string(W) + "_" + std::to_string(H) + "_";

    static const char *kCriteria[5] = {"total", "max", "min", "vector_op", "rvector_op"};
    std::vector<std::string> names;
    std::vector<std::vector<uint32_t>> all;
    for (uint32_t c = 0; c < 5; ++c)
        for (uint32_t m = 0; m < 2; ++m) {
            const std::string crit = kCriteria[c], mm = m ? "max" : "min";
            const std::vector<uint32_t> s(glutton.begin() + (size_t)(2 * c + m) * (P + 1),
                                          glutton.begin() + (size_t)(2 * c + m + 1) * (P + 1));
            names.push_back(capitalised(crit) + "_" + mm);
            all.push_back(s);
            const std::string title = "Level_" + crit + "_" + mm + "_glutton";
            write_vtk_table(root + "level_" + crit + "_" + mm + "_glutton.txt", "line_plot",
                        title + " Step " + capitalised(crit) + " " + std::to_string(P + 1) + " 1",
                        {capitalised(mm)}, {s});
        }
    write_vtk_table(root + "level_glutton.txt", "line_plot", "Level_glutton Step Total " + std::to_string(P + 1) + " 10",
                names, all);

    // treat_situation_capability: each profile on stdout (3 wide), in the
    // surface file, and plus P - strategy index in the line plot
    std::vector<std::vector<uint32_t>> rows, lines;
    std::vector<std::string> level_names;
    for (uint32_t k = 0; k <= level; ++k) {
        std::vector<uint32_t> r(surface.begin() + (size_t)k * 2 * P, surface.begin() + (size_t)(k + 1) * 2 * P);
        for (uint32_t v : r) std::cout << std::setw(3) << v << " ";
        std::cout << std::endl;
        std::vector<uint32_t> l(r);
        for (uint32_t &v : l) v += P - row_index[k];
        rows.push_back(r);
        lines.push_back(l);
        level_names.push_back("Level_" + std::to_string(k));
    }
    {
        std::ofstream f(root + "surface_vtk.txt");
        if (!f) throw logic_exception("Unable to create VTK surface file", __LINE__, __FILE__);
        f << "surface\n" << 2 * P << " " << level + 1 << "\n";
        for (const auto &r : rows) {
            for (uint32_t v : r) f << v << ' ';
            f << '\n';
        }
    }
    write_vtk_table(root + "line_plot_vtk.txt", "line_plot",
                "Profile_evolution Step Possibilities " + std::to_string(2 * P) + " " + std::to_string(level + 1),
                level_names, lines);
}

}  // namespace edge_matching_puzzle
