// feature_CUDA_backtracker.h -- the reference's CUDA backtracker search on
// the GPU (include/feature_CUDA_backtracker.h:68-72,
// src/feature_CUDA_backtracker.cpp:31-51): prints the kernel's trace
// "Thread <lane> : selected variable <variable> at step <step>" and the
// first solution (src/CUDA_backtracker.cu:119, :216-226).
// feature_CUDA_enumerate lists every solution instead, with the node and
// solution totals of the reference's CPU strategy (include/emp_strategy.h
// :412-414); it has no counterpart feature in the reference.
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

  protected:
    const puzzle_description &m_puzzle;
    uint32_t m_order;
};

class feature_CUDA_enumerate : public feature_CUDA_backtracker {
  public:
    using feature_CUDA_backtracker::feature_CUDA_backtracker;
    void run() override;
};

}  // namespace edge_matching_puzzle

#endif
