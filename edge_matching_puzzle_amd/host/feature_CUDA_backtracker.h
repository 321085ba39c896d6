// feature_CUDA_backtracker.h -- exhaustive enumeration on the GPU.
// Same name and construction as the reference's feature
// (include/feature_CUDA_backtracker.h:68-72, src/feature_CUDA_backtracker.cpp
// :31-51); run() prints every solution, then the node and solution totals
// with the lines of the reference's CPU strategy (include/emp_strategy.h
// :412-414).
#ifndef EMP_FEATURE_CUDA_BACKTRACKER_H
#define EMP_FEATURE_CUDA_BACKTRACKER_H

#include <string>

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class feature_CUDA_backtracker : public feature_if {
  public:
    feature_CUDA_backtracker(const puzzle_description &p_puzzle, const std::string &p_strategy);
    void run() override;

  private:
    const puzzle_description &m_puzzle;
    uint32_t m_order;
};

}  // namespace edge_matching_puzzle

#endif
