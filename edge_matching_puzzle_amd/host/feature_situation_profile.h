// feature_situation_profile.h -- the reference's situation_profile feature
// (include/feature_situation_profile.h, src/feature_situation_profile.cpp):
// capability profiles of the --initial_situation computed on the GPU
// (emp_hip_situation_profile), written as the reference writes them:
//   <W>_<H>_level_glutton.txt           the ten glutton profiles (line plot)
//   <W>_<H>_level_<criterion>_<min|max>_glutton.txt   one each
//   <W>_<H>_surface_vtk.txt             the profile before each placement
//   <W>_<H>_line_plot_vtk.txt           the same plus P - strategy index
// and on stdout the variable counts, the situation's level and the profiles.
#ifndef EMP_FEATURE_SITUATION_PROFILE_H
#define EMP_FEATURE_SITUATION_PROFILE_H

#include <string>

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class feature_situation_profile : public feature_if {
  public:
    feature_situation_profile(const puzzle_description &p_puzzle, const std::string &p_initial_situation);
    void run() override;

  private:
    const puzzle_description &m_puzzle;
    std::string m_initial_situation;
};

}  // namespace edge_matching_puzzle

#endif
