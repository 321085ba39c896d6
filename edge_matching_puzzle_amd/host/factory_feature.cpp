#include "factory_feature.h"

#include "feature_CUDA_backtracker.h"
#include "feature_CUDA_glutton_max.h"
#include "feature_CUDA_glutton_wide.h"
#include "feature_new_strategy.h"
#include "feature_profile.h"
#include "feature_situation_profile.h"

namespace edge_matching_puzzle {

std::unique_ptr<feature_if> factory_feature::create_feature(const std::string &p_name,
                                                            const puzzle_description &p_puzzle,
                                                            const std::string &p_initial_situation,
                                                            const std::string &p_dump_file_name,
                                                            const std::string &p_strategy,
                                                            const std::string &p_hint) {
    if (p_name == "new_strategy")
        return std::unique_ptr<feature_if>(
            new feature_new_strategy(p_puzzle, p_strategy, p_dump_file_name, p_initial_situation));
    if (p_name == "situation_profile") {
        if (p_strategy != "basic")
            throw logic_exception("situation_profile follows the basic order only", __LINE__, __FILE__);
        return std::unique_ptr<feature_if>(new feature_situation_profile(p_puzzle, p_initial_situation));
    }
    // the GPU search features start from an empty board and write no dump
    // (include/factory_feature.h:142-153 passes them neither)
    if (!p_initial_situation.empty())
        throw logic_exception("--initial_situation is only used by new_strategy and situation_profile", __LINE__,
                              __FILE__);
    if (p_name == "profile") return std::unique_ptr<feature_if>(new feature_profile(p_puzzle));
    if (p_name == "CUDA_backtracker")
        return std::unique_ptr<feature_if>(new feature_CUDA_backtracker(p_puzzle, p_strategy));
    if (p_name == "CUDA_enumerate")
        return std::unique_ptr<feature_if>(new feature_CUDA_enumerate(p_puzzle, p_strategy));
    if (p_name == "CUDA_glutton_max") return std::unique_ptr<feature_if>(new feature_CUDA_glutton_max(p_puzzle));
    if (p_name == "CUDA_glutton_wide") {
        unsigned int depth = 0;
        if (!p_hint.empty()) depth = (unsigned int)std::stoul(p_hint);
        return std::unique_ptr<feature_if>(new feature_CUDA_glutton_wide(p_puzzle, depth));
    }
    throw logic_exception("Unsupported feature \"" + p_name + "\"", __LINE__, __FILE__);
}

}  // namespace edge_matching_puzzle
