// feature_CUDA_glutton_wide.h -- glutton_wide on the GPU.  As the reference's
// CPU form (src/CUDA_glutton_wide.cpp:29-67): prints the level-0 situation
// (every position's capability), the "Corner info" of the (0,0) position,
// then for each of its candidate bits in order the bit number and the level-1
// situation play_from makes (include/CUDA_glutton_situations.h:713-891), in
// the reference's print format (include/CUDA_glutton_situation.h:127-158,
// include/CUDA_piece_position_info2.h:315-352).  The reference stops the
// program when a position loses every option; such a child is printed with a
// line saying so.  With --hint=<depth> (not a reference option for this
// feature) it prints instead the number of situations of the enumeration's
// breadth expansion at each level up to that depth.
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
    void run_levels();
    const puzzle_description &m_puzzle;
    unsigned int m_depth;   // 0: the reference's expansion of the corner
};

}  // namespace edge_matching_puzzle

#endif
