// feature_CUDA_glutton_wide.h -- breadth expansion of situations on the GPU.
// The reference's GPU launcher only prints the initial situations
// (src/CUDA_glutton_wide.cu:27-37) and its CPU form expands the (0,0)
// corner's candidates into a level-1 array (src/CUDA_glutton_wide.cpp:29-67).
// This feature expands level by level up to a depth (default 1) and prints
// the number of situations per level and the situations of the last one.
#ifndef EMP_FEATURE_CUDA_GLUTTON_WIDE_H
#define EMP_FEATURE_CUDA_GLUTTON_WIDE_H

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class feature_CUDA_glutton_wide : public feature_if {
  public:
    feature_CUDA_glutton_wide(const puzzle_description &p_puzzle, unsigned int p_depth)
        : m_puzzle(p_puzzle), m_depth(p_depth) {}
    void run() override;

  private:
    const puzzle_description &m_puzzle;
    unsigned int m_depth;
};

}  // namespace edge_matching_puzzle

#endif
