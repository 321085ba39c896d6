// feature_profile.h -- the reference's profile feature
// (include/feature_profile.h): the search over every situation the system
// equation's capability allows, computed on the GPU (emp_hip_profile), and
// its outputs written as the reference writes them:
//   profile_{low,high,total}_{min,max}_vtk.txt  per-level extrema of the
//                                               capability profiles, three
//                                               series (General, Solution,
//                                               NoSolution)
//   solutions.txt                               the solutions' situation
//                                               strings, sorted
// and on stdout the starting situations (25 pieces), the situations of each
// level and the solutions.
#ifndef EMP_FEATURE_PROFILE_H
#define EMP_FEATURE_PROFILE_H

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class feature_profile : public feature_if {
  public:
    explicit feature_profile(const puzzle_description &p_puzzle);
    void run() override;

  private:
    const puzzle_description &m_puzzle;
};

}  // namespace edge_matching_puzzle

#endif
