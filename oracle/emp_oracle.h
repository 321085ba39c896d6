/*
 * emp_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference's CPU search algorithms (quicky2000/edge_matching_puzzle).  Only
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark
 * may load it, and only as the checker / the baseline.  The product library
 * (libemp_hip.so) never links, loads or calls anything in this directory.
 *
 * Puzzles are passed as W, H and colours[P][4] (N, E, S, W) for piece ids
 * 1..P in file order, colour 0 being the frame, exactly as in the reference's
 * text format (include/emp_pieces_parser.h:53-104).
 *
 * Boards are returned row-major, one uint16 per cell:
 *   ((piece_id - 1) << 2) | orientation, or 0xFFFF for an empty cell.
 */
#ifndef EMP_ORACLE_H
#define EMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared by every entry point. */
#define ORACLE_OK            0
#define ORACLE_BAD_ARGUMENT  1
#define ORACLE_BAD_PUZZLE    2
#define ORACLE_LIMIT_REACHED 3

/* Position orders (include/emp_strategy_generator_factory.h) */
#define ORACLE_ORDER_SPIRAL 0
#define ORACLE_ORDER_BASIC  1

/*
 * Exhaustive depth-first enumeration, restating emp_strategy::specific_run
 * (include/emp_strategy.h:332-415) with HANDLE_IDENTICAL_PIECES semantics
 * for appearance-identical pieces (include/emp_position_strategy.h:189-202).
 *
 *   node_limit      stop after that many placements (0 = unlimited); status
 *                   ORACLE_LIMIT_REACHED then means the counts are partial.
 *   boards          [max_solutions][P] row-major boards of the first
 *                   max_solutions solutions found (may be NULL).
 *   solution_nodes  node index of each stored solution (may be NULL).
 *   depth_nodes     [P] placements per strategy index (may be NULL).
 *   depth_dead      [P] placements at index k after which index k+1 had no
 *                   transition at all (may be NULL).
 *   trace / trace_capacity  "New max : k" + situation lines as printed by the
 *                   reference's MAX_DISPLAY build, '\n' separated (may be NULL).
 */
int oracle_enumerate(uint32_t width, uint32_t height, const uint32_t *colours,
                     int order, uint64_t node_limit,
                     uint64_t max_solutions, uint16_t *boards,
                     uint64_t *solution_nodes,
                     uint64_t *n_solutions, uint64_t *n_nodes,
                     uint64_t *depth_nodes, uint64_t *depth_dead,
                     char *trace, uint64_t trace_capacity);

/*
 * The same DFS cut at a depth: every placement at strategy index depth-1 is
 * recorded as a prefix (depth transitions, the oracle's kind_index << 2 | o
 * numbering) and not descended.  n_nodes / n_solutions count the placements
 * and solutions above the cut (solutions only when depth reaches P).
 */
int oracle_enumerate_prefixes(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                              uint32_t depth, uint16_t *prefixes, uint64_t capacity,
                              uint64_t *n_prefixes, uint64_t *n_nodes, uint64_t *n_solutions);

/*
 * The DFS below one prefix (transitions at indices 0..prefix_len-1, as
 * returned by oracle_enumerate_prefixes); ends when index prefix_len has no
 * transition left.  n_nodes excludes the prefix's own placements.  The union
 * over all prefixes of one depth is the whole enumeration.
 */
int oracle_enumerate_subtree(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                             uint32_t prefix_len, const uint16_t *prefix, uint64_t max_solutions,
                             uint16_t *boards, uint64_t *n_solutions, uint64_t *n_nodes);

/*
 * The reference's search resumed from a situation (emp_strategy::
 * set_initial_state then specific_run, include/emp_strategy.h:332-467):
 * situation is a row-major board whose leading strategy positions hold the
 * pieces to replay; the search restarts at the last of them and ends when
 * index 1 has no transition left.  Node counts start at 0 (placements of the
 * resumed run only); solution_nodes gives each stored solution's count.
 */
int oracle_enumerate_resume(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                            const uint16_t *situation, uint64_t max_solutions, uint16_t *boards,
                            uint64_t *solution_nodes, uint64_t *n_solutions, uint64_t *n_nodes);

/*
 * Same DFS below one prefix, reporting its placements and the most pieces
 * placed in one situation (prefix_len when nothing fits below it).
 */
int oracle_subtree_depth(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                         uint32_t prefix_len, const uint16_t *prefix, uint64_t *n_nodes, uint32_t *max_depth);

/*
 * results.bin v1 image of an enumeration (include/emp_situation_binary_dumper.h
 * :72-126): header, strategy positions, one record per solution, total nodes.
 * Writes at most capacity bytes into out and returns the full size in *size.
 */
int oracle_results_bin(uint32_t width, uint32_t height, const uint32_t *colours,
                       int order, uint8_t *out, uint64_t capacity,
                       uint64_t *size);

/* Strategy order: positions[i] = (x, y) of the i-th visited position. */
int oracle_order(uint32_t width, uint32_t height, int order, uint32_t *xy);

/*
 * Position masks (piece-major, [P][ceil(4P/64)] u64, 0 at placed positions)
 * after the moves (pos, piece, o triples) are played from the level-0
 * capability as glutton_wide's play_from does; status 0 valid, 1 a free
 * position lost every option, 2 a move is not playable.
 */
int oracle_glutton_masks(uint32_t width, uint32_t height, const uint32_t *colours, uint32_t n_moves,
                         const uint32_t *moves, uint64_t *masks, uint32_t *status);

/*
 * glutton_max restatement (include/CUDA_glutton_max.h:742-1637,
 * include/CUDA_glutton_max_stack.h:38-630).
 *
 *   prefix_moves    n_prefix moves (position, piece_index, orientation),
 *                   played through the same move application as the search
 *                   itself before the search starts; the search never pops
 *                   below them.  An unplayable move is ORACLE_BAD_ARGUMENT; a
 *                   move that leaves a free cell without option ends the
 *                   search there (not solved, no push).
 *   push_limit      stop after that many pushes (0 = unlimited).
 *   board           [P] final board: the full solution when *solved, else the
 *                   deepest situation seen (the stack's max played info).
 *   max_depth       deepest level reached (pieces on board, prefix included).
 *   n_pushes        number of pushes (nodes) performed by the search.
 *   solved          1 when the stack became full.
 * Returns ORACLE_LIMIT_REACHED when the push limit stopped the search.
 */
int oracle_glutton_max(uint32_t width, uint32_t height, const uint32_t *colours,
                       uint32_t n_prefix, const uint32_t *prefix_moves,
                       uint64_t push_limit, uint16_t *board,
                       uint32_t *max_depth, uint64_t *n_pushes,
                       uint32_t *solved);

/*
 * CUDA_backtracker restatement (src/CUDA_backtracker.cu:68-178 with the
 * variables and transitions of include/emp_variable_generator.h:336-621 and
 * include/system_equation_for_CUDA.h:70-173): forward-checking search for
 * the first solution, candidates in orientation-major order.  trace receives
 * (lane, variable, step) triples as the kernel prints them
 * ("Thread <lane> : selected variable <variable> at step <step>").
 */
int oracle_backtracker(uint32_t width, uint32_t height, const uint32_t *colours, int order, uint16_t *board,
                       uint32_t *found, uint32_t *trace, uint64_t trace_capacity, uint64_t *trace_length);

/*
 * situation_profile restatement (include/feature_situation_profile.h
 * :150-483, include/situation_profile.h, include/situation_capability.h) over
 * the variables and transitions of the system equation
 * (include/system_equation_for_CUDA.h:70-173, include/emp_variable_generator.h
 * :561-621), positions in the basic order.  situation: row-major board of the
 * placed pieces (0xFFFF empty); level: pieces placed.
 *   glutton  [10][P + 1]  the glutton profiles (total of the capability after
 *            each placement) for total, max, min, vector_op, rvector_op, each
 *            searched for the min then the max
 *   surface  [level + 1][2P]  the sorted capability profile before each
 *            placement of the situation in strategy order, then at the end
 *   row_index  [level + 1]  strategy index of each surface row's placement
 *            (P for the last row)
 *   n_variables  variables of the system equation
 * Returns ORACLE_BAD_ARGUMENT when a placement is not a variable.
 */
int oracle_situation_profile(uint32_t width, uint32_t height, const uint32_t *colours, const uint16_t *situation,
                             uint32_t *level, uint32_t *glutton, uint32_t *surface, uint32_t *row_index,
                             uint32_t *n_variables);

/*
 * The profile feature (include/feature_profile.h:582-868): a level-by-level
 * search over every situation the capability allows (placements in any
 * order, valid profiles only, each situation once), with per-level extrema
 * of the profiles' min ("low"), max ("high") and total.
 *   stats      [6][3][P + 1]: low min, low max, high min, high max, total
 *              min, total max; series all situations, situations leading to
 *              a solution, the others; by level
 *   solutions  [capacity][P] the solution set (row-major boards), ordered
 * Returns ORACLE_LIMIT_REACHED when a level holds more than max_situations.
 */
int oracle_profile(uint32_t width, uint32_t height, const uint32_t *colours, uint64_t max_situations,
                   uint32_t *stats, uint16_t *solutions, uint32_t capacity, uint32_t *n_solutions);

#ifdef __cplusplus
}
#endif

#endif
