// factory_feature.h -- creates a feature by name, as the reference's
// factory_feature::create_feature (include/factory_feature.h:49-167) does for
// CUDA_backtracker, CUDA_glutton_max, CUDA_glutton_wide and new_strategy,
// with the same command-line inputs (initial situation, dump file, strategy,
// hint).  Other names are refused with the reference's message.
#ifndef EMP_FACTORY_FEATURE_H
#define EMP_FACTORY_FEATURE_H

#include <memory>
#include <string>

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class factory_feature {
  public:
    static std::unique_ptr<feature_if> create_feature(const std::string &p_name, const puzzle_description &p_puzzle,
                                                      const std::string &p_initial_situation,
                                                      const std::string &p_dump_file_name,
                                                      const std::string &p_strategy, const std::string &p_hint);
};

}  // namespace edge_matching_puzzle

#endif
