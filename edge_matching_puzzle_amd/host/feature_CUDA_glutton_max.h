// feature_CUDA_glutton_max.h -- glutton_max from the empty board on the GPU
// (include/feature_CUDA_glutton_max.h:61-75 runs launch_CUDA_glutton_max).
// Prints "Launch kernels" and "Situation with stack played info:" followed by
// the situation, as the reference does (src/CUDA_glutton_max.cu:58,
// include/CUDA_glutton_max.h:180-212).
#ifndef EMP_FEATURE_CUDA_GLUTTON_MAX_H
#define EMP_FEATURE_CUDA_GLUTTON_MAX_H

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class feature_CUDA_glutton_max : public feature_if {
  public:
    explicit feature_CUDA_glutton_max(const puzzle_description &p_puzzle) : m_puzzle(p_puzzle) {}
    void run() override;

  private:
    const puzzle_description &m_puzzle;
};

}  // namespace edge_matching_puzzle

#endif
