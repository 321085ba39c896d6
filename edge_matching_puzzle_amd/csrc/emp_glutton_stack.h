// emp_glutton_stack.h -- a glutton search's stack as XML (host).
//
// The element layout is the reference's CUDA_glutton_stack_XML_converter
// (include/CUDA_glutton_stack_XML_converter.h:165-516): a CUDA_glutton_stack
// root with size / level / nb_pieces / max attributes, the arrays
// info_to_position, position_to_info, played_info, max_played_info and
// available_pieces (items "0x%08x" with an index attribute), and
// position_infos holding one level element per stack level, each with its
// size - level position_info elements of 32 words.  A base attribute (the
// level the search never pops below, 0 when absent) is added on the root.
#ifndef EMP_GLUTTON_STACK_H
#define EMP_GLUTTON_STACK_H

#include <cstdint>
#include <string>

#include "emp_model.h"

namespace emp {

// state / masks in the layout of glutton_state_words / glutton_mask_words.
std::string glutton_stack_to_xml(const Puzzle &puzzle, const uint32_t *state, const uint64_t *masks);

// Reads an XML stack into state / masks (levels above the stack's level are
// not read and left zero); throws std::invalid_argument naming what is
// missing or inconsistent, with the reference loader's messages where it has
// one, and when the stack is not one a search of this puzzle can be in.
void glutton_stack_from_xml(const Puzzle &puzzle, const std::string &xml, uint32_t *state, uint64_t *masks);

}  // namespace emp

#endif
