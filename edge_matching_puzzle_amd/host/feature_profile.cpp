#include "feature_profile.h"

#include <fstream>
#include <iostream>
#include <set>
#include <vector>

namespace edge_matching_puzzle {

feature_profile::feature_profile(const puzzle_description &p_puzzle) : m_puzzle(p_puzzle) {}

void feature_profile::run() {
    const uint32_t W = m_puzzle.width, H = m_puzzle.height, P = W * H;
    // the sizes feature_profile::run instantiates (include/feature_profile.h:379-403)
    if (P != 9 && P != 16 && P != 25 && P != 36 && P != 72 && P != 256)
        throw logic_exception("Unsupported size " + std::to_string(W) + "x" + std::to_string(H), __LINE__, __FILE__);
    device_session session;
    emp_hip_puzzle puzzle = m_puzzle.view();
    std::vector<uint32_t> stats((size_t)6 * 3 * (P + 1));
    const uint64_t capacity = 1u << 20;
    std::vector<uint16_t> solutions((size_t)capacity * P), starts((size_t)6 * P);
    std::vector<uint64_t> level_sizes(P + 1);
    uint64_t n_solutions = 0;
    uint32_t n_starts = 0;
    EMP_HIP_CALL(emp_hip_profile(session.get(), &puzzle, 1ull << 26, stats.data(), solutions.data(), capacity,
                                 &n_solutions, level_sizes.data(), starts.data(), &n_starts));
    if (P == 25)
        for (uint32_t i = 0; i < n_starts; ++i) std::cout << situation_string(m_puzzle, starts.data() + (size_t)i * P) << std::endl;
    const uint32_t first = P == 25 ? 4 : 0;
    for (uint32_t l = first; l <= P; ++l) std::cout << "Level " << l << " : " << level_sizes[l] << " situations" << std::endl;

    // dump_solutions (include/feature_profile.h:975-987): sorted strings, on
    // stdout and in solutions.txt
    std::set<std::string> strings;
    for (uint64_t i = 0; i < std::min(n_solutions, capacity); ++i)
        strings.insert(situation_string(m_puzzle, solutions.data() + (size_t)i * P));
    std::ofstream dump("solutions.txt");
    for (const std::string &s : strings) {
        std::cout << s << std::endl;
        dump << s << std::endl;
    }

    // the six histograms (include/feature_profile.h:801-868)
    static const char *kFiles[6] = {"profile_low_min_vtk.txt",  "profile_low_max_vtk.txt",
                                    "profile_high_min_vtk.txt", "profile_high_max_vtk.txt",
                                    "profile_total_min_vtk.txt", "profile_total_max_vtk.txt"};
    static const char *kTitles[6] = {"Profile_Low_min",  "Profile_Low_max",   "Profile_High_min",
                                     "Profile_High_max", "Profile_Total_Min", "Profile_Total_Max"};
    for (uint32_t k = 0; k < 6; ++k) {
        std::vector<std::vector<uint32_t>> series;
        for (uint32_t s = 0; s < 3; ++s)
            series.emplace_back(stats.begin() + ((size_t)k * 3 + s) * (P + 1),
                                stats.begin() + ((size_t)k * 3 + s + 1) * (P + 1));
        write_vtk_table(kFiles[k], "histogram",
                        std::string(kTitles[k]) + " Level Transition_number " + std::to_string(P + 1) + " 3",
                        {"General", "Solution", "NoSolution"}, series);
    }
}

}  // namespace edge_matching_puzzle
