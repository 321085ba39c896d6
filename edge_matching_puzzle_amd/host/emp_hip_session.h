// emp_hip_session.h -- puzzle description and device context shared by the
// GPU features; every call goes through the C ABI of libemp_hip.so.
#ifndef EMP_HIP_SESSION_H
#define EMP_HIP_SESSION_H

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/emp_hip.h"

namespace edge_matching_puzzle {

// The reference reports failures as quicky_exception::quicky_logic_exception
// carrying the file and line; main prints "ERROR : <what> at <file>:<line>"
// and returns -1 (src/main.cpp:137-146).
class logic_exception : public std::logic_error {
  public:
    logic_exception(const std::string &what, unsigned int line, const std::string &file)
        : std::logic_error(what), m_line(line), m_file(file) {}
    unsigned int line() const { return m_line; }
    const std::string &file() const { return m_file; }

  private:
    unsigned int m_line;
    std::string m_file;
};

#define EMP_HIP_CALL(expr)                                                                    \
    do {                                                                                      \
        if (int status_ = (expr))                                                             \
            throw logic_exception(std::string(emp_hip_last_error()) + " (status " +           \
                                      std::to_string(status_) + ")", __LINE__, __FILE__);     \
    } while (0)

// Piece colours as read from the puzzle file (include/emp_pieces_parser.h:53-104).
struct puzzle_description {
    uint32_t width = 0, height = 0;
    std::vector<uint32_t> colours;   // [P][4] N, E, S, W per piece id 1..P
    emp_hip_puzzle view() const { return emp_hip_puzzle{width, height, colours.data()}; }
};

puzzle_description parse_puzzle_file(const std::string &path);

class device_session {
  public:
    device_session() { EMP_HIP_CALL(emp_hip_create(0, &m_context)); }
    ~device_session() { emp_hip_destroy(m_context); }
    device_session(const device_session &) = delete;
    device_session &operator=(const device_session &) = delete;
    emp_hip_context *get() const { return m_context; }

  private:
    emp_hip_context *m_context = nullptr;
};

std::string situation_string(const puzzle_description &p, const uint16_t *board);

// VTK table files (include/VTK_table_based_dumper.h): type, then title, axis
// names, series length and count, then the series names, then one line per
// series, every value followed by a space.
void write_vtk_table(const std::string &file_name, const std::string &type, const std::string &header,
                     const std::vector<std::string> &names, const std::vector<std::vector<uint32_t>> &series);

}  // namespace edge_matching_puzzle

#endif
