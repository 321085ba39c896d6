// emp_hip_session.cpp -- puzzle file parsing and situation strings.
#include "emp_hip_session.h"

#include <fstream>
#include <sstream>

namespace edge_matching_puzzle {

puzzle_description parse_puzzle_file(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw logic_exception("Unable to open file \"" + path + "\"", __LINE__, __FILE__);
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(in, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
        if (!line.empty()) lines.push_back(line);
    }
    if (lines.size() < 3) throw logic_exception("Bad puzzle file \"" + path + "\"", __LINE__, __FILE__);
    puzzle_description p;
    char comma;
    std::istringstream dims(lines[1]);
    if (!(dims >> p.width >> comma >> p.height))
        throw logic_exception("Bad dimensions line \"" + lines[1] + "\"", __LINE__, __FILE__);
    const uint32_t n = p.width * p.height;
    p.colours.assign(4 * n, 0);
    std::vector<char> seen(n, 0);
    for (size_t i = 3; i < lines.size(); ++i) {
        std::istringstream row(lines[i]);
        uint32_t id, c[4];
        if (!(row >> id >> comma >> c[0] >> comma >> c[1] >> comma >> c[2] >> comma >> c[3]) || id < 1 || id > n)
            throw logic_exception("Bad piece line \"" + lines[i] + "\"", __LINE__, __FILE__);
        if (seen[id - 1]) throw logic_exception("Piece " + std::to_string(id) + " defined twice", __LINE__, __FILE__);
        seen[id - 1] = 1;
        for (int s = 0; s < 4; ++s) p.colours[4 * (id - 1) + s] = c[s];
    }
    // src/main.cpp:82-91: the number of pieces must match the board
    for (uint32_t i = 0; i < n; ++i)
        if (!seen[i])
            throw logic_exception("Number of pieces does not match " + std::to_string(p.width) + "x" +
                                      std::to_string(p.height), __LINE__, __FILE__);
    return p;
}

std::string situation_string(const puzzle_description &p, const uint16_t *board) {
    std::vector<char> text(5 * p.width * p.height + 1);
    EMP_HIP_CALL(emp_hip_situation_string(p.width, p.height, board, text.data(), text.size()));
    return std::string(text.data());
}

void write_vtk_table(const std::string &file_name, const std::string &type, const std::string &header,
                     const std::vector<std::string> &names, const std::vector<std::vector<uint32_t>> &series) {
    std::ofstream f(file_name);
    if (!f) throw logic_exception("Unable to create VTK " + type + " file \"" + file_name + "\"", __LINE__, __FILE__);
    f << type << '\n' << header << '\n';
    for (size_t i = 0; i < names.size(); ++i) f << names[i] << (i + 1 < names.size() ? " " : "\n");
    for (const auto &ser : series) {
        for (uint32_t v : ser) f << v << ' ';
        f << '\n';
    }
}

}  // namespace edge_matching_puzzle
