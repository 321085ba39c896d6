// feature_new_strategy.h -- the reference's exhaustive search feature
// (emp_strategy, include/emp_strategy.h:159-467, created for "new_strategy"
// by include/factory_feature.h:127-135) on the GPU: every solution with its
// node index written to the dump file as results.bin v1
// (include/emp_situation_binary_dumper.h:72-126), an optional initial
// situation replayed as the reference does (include/emp_strategy.h:418-467),
// and the reference's closing lines ("End of algorithm", "Total situations
// explored : <n>", "Nb solutions : <n>", include/emp_strategy.h:412-414).
#ifndef EMP_FEATURE_NEW_STRATEGY_H
#define EMP_FEATURE_NEW_STRATEGY_H

#include <string>

#include "emp_hip_session.h"
#include "feature_if.h"

namespace edge_matching_puzzle {

class feature_new_strategy : public feature_if {
  public:
    feature_new_strategy(const puzzle_description &p_puzzle, const std::string &p_strategy,
                         const std::string &p_dump_file, const std::string &p_initial_situation);
    void run() override;

  private:
    const puzzle_description &m_puzzle;
    uint32_t m_order;
    std::string m_dump_file;
    std::string m_initial_situation;
};

}  // namespace edge_matching_puzzle

#endif
