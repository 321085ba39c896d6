/*
 * emp_hip.h -- C ABI of libemp_hip.so, the MI355X (gfx950) search path of
 * edge_matching_puzzle.
 *
 * The reference binds its GPU search through three free launchers called by
 * its feature classes (include/feature_CUDA_backtracker.h:68-72,
 * include/feature_CUDA_glutton_max.h:55-57, include/feature_CUDA_glutton_wide.h
 * :57-59).  Those launchers take the piece database and board description and
 * print their result; this ABI takes the same information as plain arrays and
 * returns results in caller-owned buffers, so any host language can bind it.
 *
 * Conventions
 *   - Every entry point returns EMP_HIP_OK (0) or an error code; the message
 *     of the last error on the calling thread is emp_hip_last_error().
 *   - A puzzle is (width, height, colours[P][4]) with colours N, E, S, W of
 *     piece ids 1..P in file order and 0 for the frame, i.e. the content of
 *     the reference's puzzle text file (include/emp_pieces_parser.h:53-104).
 *   - A board is row-major, one uint16 per cell: ((id - 1) << 2) | orientation
 *     (orientation 0..3 = N, E, S, W as in include/emp_types.h), 0xFFFF empty.
 *   - A context owns one HIP stream on one device; calls on one context are
 *     stream-ordered and synchronous on return.  Contexts are
 *     thread-compatible: use one per host thread.
 */
#ifndef EMP_HIP_H
#define EMP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMP_HIP_OK              0
#define EMP_HIP_BAD_ARGUMENT    1
#define EMP_HIP_BAD_PUZZLE      2
#define EMP_HIP_DEVICE_ERROR    3
#define EMP_HIP_OUT_OF_MEMORY   4
#define EMP_HIP_UNSUPPORTED     5

#define EMP_HIP_ORDER_SPIRAL 0  /* include/emp_spiral_strategy_generator.h */
#define EMP_HIP_ORDER_BASIC  1  /* include/emp_basic_strategy_generator.h  */

typedef struct emp_hip_context emp_hip_context;

typedef struct {
    uint32_t width;
    uint32_t height;
    const uint32_t *colours;  /* [width * height][4] */
} emp_hip_puzzle;

/* Message of the last failing call on this thread ("" if none). */
const char *emp_hip_last_error(void);

/* Version string of the library. */
const char *emp_hip_version(void);

/* Creates a context on HIP device `device` (its own stream). */
int emp_hip_create(int device, emp_hip_context **context);
int emp_hip_destroy(emp_hip_context *context);

/* Device facts used by the autotuner: compute units, LDS bytes per block,
 * clock in kHz. */
int emp_hip_device_info(emp_hip_context *context, uint32_t *compute_units,
                        uint32_t *lds_per_block, uint32_t *clock_khz);

/*
 * Seeded synthetic puzzle (random-edge model): frame edges 0, edges between
 * two frame cells from border colours 1..B, other edges from interior colours
 * B+1..B+I; interior pieces given a random quarter turn; ids shuffled.
 * colours receives [W*H][4]; planted (may be NULL) the planted solution.
 * No counterpart in the reference (its puzzles come from text files).
 */
int emp_hip_generate(uint64_t seed, uint32_t width, uint32_t height,
                     uint32_t border_colours, uint32_t interior_colours,
                     uint32_t *colours, uint16_t *planted);

/* Validates a puzzle as the reference's piece database does (kinds match the
 * board, every colour on an even number of edges, colour count equals twice
 * the inner edges: include/emp_piece_db.h:466-1010).  No device work. */
int emp_hip_validate_puzzle(const emp_hip_puzzle *puzzle);

/* Strategy order (include/emp_spiral_strategy_generator.h:52-98,
 * include/emp_basic_strategy_generator.h:44-53): xy[2i], xy[2i+1]. */
int emp_hip_order(uint32_t width, uint32_t height, uint32_t order, uint32_t *xy);

/* Situation string of a board (include/situation_string_formatter.h:60-81).
 * Writes at most capacity bytes including the terminating 0. */
int emp_hip_situation_string(uint32_t width, uint32_t height, const uint16_t *board,
                             char *out, uint64_t capacity);

/* ------------------------------------------------------------------------ *
 * Exhaustive enumeration (the reference's CPU strategy emp_strategy,
 * include/emp_strategy.h:332-415, run on the GPU): index 0 takes its first
 * candidate only, candidates are visited in ascending (kind_index << 2 | o),
 * appearance-identical pieces are tried once per position.
 * Replaces: launch_CUDA_backtracker (src/CUDA_backtracker.cu:235-269) for the
 * enumeration use of the backtracker feature.
 * ------------------------------------------------------------------------ */
typedef struct {
    uint32_t order;            /* EMP_HIP_ORDER_*                               */
    uint32_t waves_per_cu;     /* 0 = tuned default                             */
    uint32_t waves_per_block;  /* 0 = sized to fit the most resident waves;
                                  at most 16 (4 for boards with more than 64
                                  pieces of one kind)                          */
    uint32_t steps_per_round;  /* steps between work refills, 0 = default (128) */
    uint64_t step_budget;      /* passes only: steps per work item per pass
                                  before it is requeued (and split in two
                                  while the queue has room), 0 = default      */
    uint64_t target_items;     /* prefix-split until this many items, 0 = auto  */
    uint64_t node_limit;       /* per puzzle: stop taking its work items once
                                  this many nodes were counted, 0 = none      */
    uint64_t max_solutions;    /* capacity of each boards[i] in solutions       */
    uint32_t use_graph;        /* capture each pass in a hipGraph               */
    uint32_t queue_items;      /* work queue capacity in items, 0 = sized from
                                  the grid (at least the batch size; a full
                                  queue only stops items splitting in two)    */
    double   time_limit_ms;    /* device time after which every search stops
                                  (results then incomplete), 0 = none         */
    uint32_t count_work;       /* 1: count the algorithmic work of every
                                  placement into the stats (an instrumented
                                  run, for the roofline; plain order only)    */
    uint32_t schedule;         /* EMP_HIP_SCHEDULE_*                            */
} emp_hip_enum_options;

/* How the DFS work is shared out.
 * PERSISTENT: one launch; work items sit in a device-side ring, idle lanes
 *   take the next one, and a lane hands over the rest of the shallowest level
 *   of its search as a new item every step_budget steps (512 while lanes wait
 *   for work), so the search spreads without host round trips.
 * PASSES: launches of bounded length; items over their step budget split
 *   into the next launch's queue, which batches regroup by puzzle before
 *   each launch (the form emp_hip_enumerate_noalloc always uses).
 * AUTO (0, the default): PERSISTENT for one puzzle (or several roots of
 *   one), PASSES for batches of several puzzles, whose tables are read
 *   through the caches and gain from the regrouping.  Ordered results
 *   (emp_hip_enumerate_items with node_index) always use PASSES. */
#define EMP_HIP_SCHEDULE_AUTO 0u
#define EMP_HIP_SCHEDULE_PASSES 1u
#define EMP_HIP_SCHEDULE_PERSISTENT 2u

typedef struct {
    uint64_t n_nodes;          /* placements (the reference's node count)      */
    uint64_t n_solutions;      /* exact when complete                          */
    uint64_t n_stored;         /* boards written (min(n_solutions, capacity))  */
    uint32_t complete;         /* 1: whole tree enumerated                     */
    uint32_t status;           /* EMP_HIP_OK or an error for this puzzle       */
    double   solve_ms;         /* first work item start to last item end       */
    uint32_t max_depth;        /* most pieces placed in one situation visited  */
    uint32_t reserved;
} emp_hip_enum_result;

typedef struct {
    double   total_ms;         /* wall time of the device part                 */
    double   dfs_kernel_ms;    /* sum of DFS pass kernel times (HIP events)    */
    double   expand_kernel_ms; /* sum of prefix-expansion kernel times         */
    uint64_t dfs_launches;
    uint64_t expand_launches;
    uint64_t items_processed;
    uint64_t splits;
    uint32_t max_words;        /* u64 words of the widest candidate mask       */
    uint32_t lds_tables;       /* 1: tables held in LDS                        */
    uint32_t blocks;           /* grid of the DFS kernel                       */
    uint32_t threads;          /* block size of the DFS kernel                 */
    uint64_t candidate_sets;   /* DFS iterations, one candidate set each       */
    double   group_ms;         /* device time grouping the work items by puzzle
                                  before the passes (batches without a node
                                  limit), outside dfs_kernel_ms               */
    uint64_t work_bytes;       /* count_work: algorithmic LDS bytes, 8 M (n + 3)
                                  per placement (M = ceil(4P/64) u64 words, n
                                  the placed neighbours of the next index)    */
    uint64_t work_ops;         /* count_work: integer ops, 2 M (n + 1) + 2 M  */
    uint64_t tickets;          /* persistent schedule: items searched (the
                                  starting items and the hand-overs)         */
} emp_hip_enum_stats;

void emp_hip_enum_default_options(emp_hip_enum_options *options);

/* boards[i] (may be NULL, or the array may be NULL) receives up to
 * max_solutions boards of puzzle i; results[i] its counts. stats may be NULL. */
int emp_hip_enumerate(emp_hip_context *context,
                      const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                      const emp_hip_enum_options *options,
                      emp_hip_enum_result *results,
                      uint16_t *const *boards,
                      emp_hip_enum_stats *stats);

/*
 * emp_hip_enumerate without any device allocation: the call's device memory
 * (tables, work queues, counters, the solution store) is carved from a fixed
 * buffer of EMP_HIP_SCRATCH_BYTES that each device holds from library load
 * until the process ends, and no hipMalloc / hipFree or other allocating
 * routine is called.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid it).
 *   - Results equal emp_hip_enumerate's; the work queues hold what fits in
 *     the buffer (fewer items in flight, never a different result).
 *   - A call needing more (many searches, a large max_solutions) fails with
 *     EMP_HIP_OUT_OF_MEMORY before any search runs; use_graph is refused.
 *   - Every call on one device shares that device's buffer and overwrites it:
 *     calls that overlap on one device (two contexts, two threads) race.
 */
#define EMP_HIP_SCRATCH_BYTES (128u << 20)
int emp_hip_enumerate_noalloc(emp_hip_context *context,
                              const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                              const emp_hip_enum_options *options,
                              emp_hip_enum_result *results,
                              uint16_t *const *boards,
                              emp_hip_enum_stats *stats);

/* ------------------------------------------------------------------------ *
 * Work items: one puzzle's enumeration resumed from given situations, with
 * the solutions in the order of the sequential search and a resumable cut.
 * ------------------------------------------------------------------------ */
#define EMP_HIP_NO_MOVE 0xFFFFFFFFu

/* The state of the reference's DFS loop (include/emp_strategy.h:332-415):
 * moves placed at strategy indices 0..level-1, the search continuing at index
 * `level` above the move `last` (EMP_HIP_NO_MOVE: from its first candidate),
 * backtracking, and ending when index `floor` has no candidate left.  The
 * moves of an item are a row of P packed moves (pos | piece << 8 | o << 16),
 * the first `level` of them used. */
typedef struct {
    uint32_t floor;     /* 1 <= floor <= level (index 0 is never revisited)  */
    uint32_t level;     /* level < P                                         */
    uint32_t last;      /* a move at index `level`, or EMP_HIP_NO_MOVE       */
    uint32_t reserved;
} emp_hip_item;

/*
 * Enumerates the union of the items (n_items = 0: the whole enumeration from
 * its root, as emp_hip_enumerate); the items' own moves are not counted.
 *   boards      [max_solutions][P] stored solutions
 *   node_index  (may be NULL) when given, the stored solutions come in the
 *               order of the sequential search, each with its node index
 *               there (the placements up to and including it, counted from
 *               the items; exact when the result is complete) -- the record
 *               of results.bin (include/emp_strategy.h:367-375)
 *   rest        (may be NULL, needs options->time_limit_ms > 0 or
 *               options->node_limit > 0) receives the work left when the time
 *               or node limit cut the run (*n_rest items, moves in rest_moves
 *               [rest_capacity][P]; the node limit is checked as items are
 *               taken, so the run stops somewhat past it); enumerating those
 *               items later completes the enumeration: node counts add up
 *               and the solution sets are disjoint.  EMP_HIP_OUT_OF_MEMORY
 *               when more than rest_capacity items are left (*n_rest holds
 *               the count, the first rest_capacity are written).
 * Replaces: the reference's CPU search resumed from a situation
 * (include/emp_strategy.h:418-467) and its save of the search state
 * (include/emp_advanced_feature_base.h:270-322), on the GPU.
 */
int emp_hip_enumerate_items(emp_hip_context *context, const emp_hip_puzzle *puzzle,
                            const emp_hip_enum_options *options,
                            const emp_hip_item *items, const uint32_t *item_moves, uint32_t n_items,
                            emp_hip_enum_result *result, uint16_t *boards, uint64_t *node_index,
                            emp_hip_item *rest, uint32_t *rest_moves, uint64_t rest_capacity,
                            uint64_t *n_rest, emp_hip_enum_stats *stats);

/* The item the reference resumes from a situation (emp_strategy::
 * set_initial_state, include/emp_strategy.h:418-467): with k leading strategy
 * positions placed, the search restarts at index k - 1 from its first
 * candidate with the moves of indices 0..k-2 kept (floor 1).  board is
 * row-major; each of the k moves must be a candidate after the earlier ones
 * (else EMP_HIP_BAD_ARGUMENT, as is k = 0).  *n_items = 1 with item / moves
 * [P] written, or 0 when k = 1 (the whole enumeration). */
int emp_hip_resume_item(const emp_hip_puzzle *puzzle, uint32_t order, const uint16_t *board,
                        emp_hip_item *item, uint32_t *moves, uint32_t *n_items);

/* Board of a situation string (include/situation_string_formatter.h:84-134):
 * fixed-width cells, the id right-aligned then N/E/S/W, '-' cells empty. */
int emp_hip_parse_situation(uint32_t width, uint32_t height, const char *text, uint16_t *board);

/* results.bin v1 (include/emp_situation_binary_dumper.h:72-126): header
 * (version, width, height, solution_dump, the (x, y) of each position of the
 * order), one record per board (its fields in strategy order, then its node
 * index), the total node count.  boards [n_records][P] row-major.  Writes at
 * most capacity bytes to out (may be NULL); *size is the file size. */
int emp_hip_results_bin_write(uint32_t width, uint32_t height, uint32_t order, uint32_t solution_dump,
                              const uint16_t *boards, const uint64_t *node_index, uint64_t n_records,
                              uint64_t total_nodes, uint8_t *out, uint64_t capacity, uint64_t *size);

/* Reads a results.bin v1 image (include/emp_situation_binary_reader.h).
 * xy [2P], boards [capacity][P] and node_index [capacity] may be NULL; P is
 * width * height of the header.  *complete = 0 when the records do not fill
 * the file (truncated, *total_nodes then 0). */
int emp_hip_results_bin_read(const uint8_t *data, uint64_t size, uint32_t *width, uint32_t *height,
                             uint32_t *solution_dump, uint32_t *xy, uint16_t *boards, uint64_t *node_index,
                             uint64_t capacity, uint64_t *n_records, uint64_t *total_nodes,
                             uint32_t *complete);

/*
 * The reference's CUDA backtracker search (kernel_and_info,
 * src/CUDA_backtracker.cu:68-178): forward checking to the first solution,
 * candidates of a position in orientation-major order (lowest orientation,
 * then lowest piece), with the selection trace it prints ("Thread <lane> :
 * selected variable <variable> at step <step>") as (lane, variable, step)
 * triples in trace[trace_capacity][3].  The reference is only consistent in
 * the basic order (its transition table is keyed by row-major position);
 * both orders are accepted here.  max_selections = 0: unlimited.
 * Replaces: launch_CUDA_backtracker (src/CUDA_backtracker.cu:235-269).
 */
int emp_hip_backtrack_trace(emp_hip_context *context, const emp_hip_puzzle *puzzle, uint32_t order,
                            uint64_t max_selections, uint16_t *board, uint32_t *found,
                            uint32_t *trace, uint64_t trace_capacity, uint64_t *trace_length);

/* Number of variables of the backtracker search: (position, piece,
 * orientation) triples of emp_variable_generator::create_variables
 * (include/emp_variable_generator.h:336-531), i.e. the set bits of the
 * level-0 capability the trace kernel starts from; the reference prints it as
 * "Nb variables: <n>" (include/system_equation_for_CUDA.h:82).  No device work. */
int emp_hip_backtrack_variables(const emp_hip_puzzle *puzzle, uint32_t order, uint64_t *n_variables);

/*
 * Capability profiles of a situation (the reference's situation_profile
 * feature).  The system equation's capability -- per position and per piece
 * the live (piece | position, orientation) variables -- is followed through
 * the situation's placements, positions in the basic order; a profile is the
 * 2P bit counts, sorted (include/situation_capability.h:145-155).
 * situation: row-major board, ((id - 1) << 2) | o, 0xFFFF empty.
 *   *level       pieces the situation places
 *   glutton      [10][P + 1]: compute_glutton (include/feature_situation_profile.h
 *                :388-483) for total, max, min, vector_op, rvector_op, each as
 *                a min then a max search -- the total after each placement,
 *                0 for an invalid one, padded with 0
 *   surface      [P + 1][2P], rows 0..level: the profile before each placement
 *                in strategy order, then at the end (treat_situation_capability,
 *                :290-305)
 *   row_index    [P + 1]: the strategy index of each row's placement (P for the
 *                last row); the reference's line plot adds P - row_index
 *   n_variables  variables of the system equation
 * EMP_HIP_BAD_ARGUMENT when a placement is not a variable of its position;
 * EMP_HIP_UNSUPPORTED above 256 pieces.
 * Replaces: the computation of feature_situation_profile::template_run
 * (include/feature_situation_profile.h:150-183); the host feature writes its
 * files.
 */
int emp_hip_situation_profile(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint16_t *situation,
                              uint32_t *level, uint32_t *glutton, uint32_t *surface, uint32_t *row_index,
                              uint32_t *n_variables);

/*
 * The profile feature's search: level by level, every situation the system
 * equation's capability allows (pieces placed at any free position, in any
 * order; a situation joins its level when every free position and every
 * unplaced piece keeps a variable, and only once), with per-level extrema of
 * the capability profiles (the 2P bit counts: "low" = the smallest count of a
 * free position or piece, "high" = the largest, "total" = their sum).
 *   stats        [6][3][P + 1] u32: low min, low max, high min, high max,
 *                total min, total max; series: every situation, those that
 *                precede a solution (and full boards), the others; by level.
 *                A minimum never updated is 0, as the reference reports it.
 *   solutions    [solution_capacity][P] the solutions (row-major boards, in
 *                board order); *n_solutions is their full count
 *   level_sizes  [P + 1] situations of each level (may be NULL)
 *   starts       [6][P] the searches' starting situations (may be NULL;
 *                *n_starts of them: the empty situation, or at 25 pieces the
 *                six corner placements)
 * EMP_HIP_UNSUPPORTED when a level holds more than max_situations (at most
 * 2^31) or above 256 pieces.
 * Replaces: the computation of feature_profile::template_run and wide
 * (include/feature_profile.h:582-868); the host feature writes the six VTK
 * histograms and solutions.txt.
 */
int emp_hip_profile(emp_hip_context *context, const emp_hip_puzzle *puzzle, uint64_t max_situations, uint32_t *stats,
                    uint16_t *solutions, uint64_t solution_capacity, uint64_t *n_solutions, uint64_t *level_sizes,
                    uint16_t *starts, uint32_t *n_starts);

/*
 * Breadth expansion: every situation the enumeration visits with `depth`
 * pieces placed (indices 0..depth-1 of the order), level by level on the GPU.
 * boards: [capacity][P] row-major (cells beyond the depth empty); *n_situations
 * is the full count (EMP_HIP_OUT_OF_MEMORY when it exceeds 2^25);
 * *n_nodes / *n_solutions count the placements above the depth.
 * Replaces: launch_CUDA_glutton_wide (src/CUDA_glutton_wide.cu:27-37), whose
 * CPU form expands the (0,0) corner's candidates (src/CUDA_glutton_wide.cpp
 * :29-67) with play_from (include/CUDA_glutton_situations.h:713-757).
 */
int emp_hip_expand(emp_hip_context *context, const emp_hip_puzzle *puzzle, uint32_t order,
                   uint32_t depth, uint64_t capacity, uint16_t *boards, uint64_t *n_situations,
                   uint64_t *n_nodes, uint64_t *n_solutions);

/* ------------------------------------------------------------------------ *
 * glutton_max (include/CUDA_glutton_max.h:1522-1637): best-first search
 * that removes every move leaving a free position or an available piece
 * without option, plays the best-scored move, and backtracks when a level
 * runs dry.  One search per prefix; each prefix (prefix_len moves packed as
 * position | piece_index << 8 | orientation << 16, position row-major, piece
 * index = id - 1) is played before the search, which never undoes it.
 * Replaces: launch_CUDA_glutton_max (src/CUDA_glutton_max.cu:38-108), which
 * runs one search from the empty board (prefix_len = 0, n_searches = 1).
 * At most 256 pieces (include/CUDA_common_struct_glutton.h:358).
 * ------------------------------------------------------------------------ */
#define EMP_HIP_GLUTTON_SOLVED     0  /* stack full: boards[i] is a solution  */
#define EMP_HIP_GLUTTON_EXHAUSTED  1  /* no move left above the prefix        */
#define EMP_HIP_GLUTTON_BUDGET     2  /* stopped after push_budget pushes     */
#define EMP_HIP_GLUTTON_BAD_PREFIX 3  /* a prefix move is not playable        */

typedef struct {
    uint64_t push_budget;      /* pushes per search, 0 = unlimited          */
    uint32_t waves_per_cu;     /* 0 = as many as LDS allows                 */
    uint32_t reserved;
} emp_hip_glutton_options;

typedef struct {
    uint32_t max_depth;        /* deepest level: pieces on the board        */
    uint32_t status;           /* EMP_HIP_GLUTTON_*                         */
    uint64_t pushes;           /* pushes of the search (prefix excluded)    */
} emp_hip_glutton_result;

typedef struct {
    double   total_ms;
    double   kernel_ms;        /* HIP events around the glutton kernel      */
    uint64_t pushes;
    uint64_t evaluations;      /* candidate evaluations                      */
    uint32_t waves;
    uint32_t lds_bytes;        /* LDS per wave                               */
    uint64_t score_evaluations; /* evaluations that scored a candidate       */
    uint64_t rounds;           /* rounds of invalid-candidate removal        */
    uint64_t free_slot_sum;    /* free positions of the level, summed over the
                                  evaluations: the reference evaluates a
                                  candidate with one mask AND per other free
                                  position and up to four colour-mask ANDs
                                  (include/CUDA_glutton_max.h:455-714)      */
} emp_hip_glutton_stats;

/* boards: [n_searches][P], the solution or the deepest situation reached
 * (the stack's max played info).  stats may be NULL. */
int emp_hip_glutton_max(emp_hip_context *context, const emp_hip_puzzle *puzzle,
                        const uint32_t *prefixes, uint32_t prefix_len, uint32_t n_searches,
                        const emp_hip_glutton_options *options,
                        emp_hip_glutton_result *results, uint16_t *boards,
                        emp_hip_glutton_stats *stats);

/*
 * glutton_wide's breadth step (src/CUDA_glutton_wide.cpp:29-67, with
 * CUDA_glutton_situations::play_from, include/CUDA_glutton_situations.h
 * :713-891).  Each parent is a row of parent_len moves (pos | piece << 8 |
 * o << 16) played from the level-0 capability (include/CUDA_common_glutton.h
 * :148-211): a move removes its piece from every position's mask and ANDs
 * the colour constraint of each side into the free neighbour that side
 * faces.  Each valid parent is then expanded at cells[p], one child per bit
 * of that position's mask in bit order (bit 4 * piece + o).
 *   masks        piece-major, nw = ceil(4 P / 64) u64 words per position,
 *                position-indexed [P][nw], 0 at placed positions
 *   parent_masks [n_parents][P][nw] (may be NULL); parent_status per parent
 *   child_*      up to `capacity` children: parent index, move, masks
 *                [capacity][P][nw] (may be NULL), status; *n_children is the
 *                full count
 *   status       EMP_HIP_WIDE_VALID, EMP_HIP_WIDE_EMPTIED (a free position
 *                lost every option -- the reference stops the program
 *                there), EMP_HIP_WIDE_BAD_MOVE (a parent move on a placed
 *                position or outside its mask)
 * Replaces: launch_CUDA_glutton_wide (src/CUDA_glutton_wide.cpp:70-75).
 */
#define EMP_HIP_WIDE_VALID    0
#define EMP_HIP_WIDE_EMPTIED  1
#define EMP_HIP_WIDE_BAD_MOVE 2
int emp_hip_glutton_wide(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *parents,
                         uint32_t parent_len, uint32_t n_parents, const uint32_t *cells, uint64_t *parent_masks,
                         uint32_t *parent_status, uint64_t capacity, uint32_t *child_parent, uint32_t *child_move,
                         uint64_t *child_masks, uint32_t *child_status, uint64_t *n_children);

/* ------------------------------------------------------------------------ *
 * glutton stack checkpoints.  The reference dumps and reloads the whole
 * stack of a glutton run (CUDA_glutton_stack_XML_converter,
 * include/CUDA_glutton_stack_XML_converter.h:165-516, written every 1024
 * steps by include/CUDA_glutton_max.h:1556-1564); here a long search is cut
 * at its push budget and resumed from its saved stack.  A stack is state
 * words (u32) and masks (u64) in the layout documented in
 * edge_matching_puzzle_amd/csrc/emp_glutton.h (glutton_state_words).
 * ------------------------------------------------------------------------ */

/* Words of one saved stack for n_pieces pieces. */
int emp_hip_glutton_stack_words(uint32_t n_pieces, uint32_t *state_words, uint64_t *mask_words);

/* emp_hip_glutton_max with stacks.  in_state / in_masks (n_searches stacks;
 * prefixes null, prefix_len 0): each search resumes from its stack instead
 * of playing a prefix.  out_state / out_masks: each search's stack is saved
 * when it ends (push budget, solved or exhausted).  A search resumed from
 * the stack its push budget cut saved continues exactly as the uncut search:
 * the same moves, max depth, status and board, pushes counted from the
 * resume.  EMP_HIP_BAD_ARGUMENT for a stack no search of this puzzle can be
 * in.  Replaces: the reload of CUDA_glutton_stack_XML_converter::load
 * (:158-186) into a stack the kernel continues. */
int emp_hip_glutton_max_stacks(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *prefixes,
                               uint32_t prefix_len, uint32_t n_searches, const emp_hip_glutton_options *options,
                               const uint32_t *in_state, const uint64_t *in_masks, uint32_t *out_state,
                               uint64_t *out_masks, emp_hip_glutton_result *results, uint16_t *boards,
                               emp_hip_glutton_stats *stats);

/* XML form of one stack: the reference converter's element layout
 * (CUDA_glutton_stack root with size / level / nb_pieces / max, the five
 * arrays, position_infos with one level per stack level of 32-word
 * position_info elements) plus a base attribute.  to_xml writes at most
 * capacity bytes to out (null: size only) and the full size to *size;
 * from_xml reads size bytes, EMP_HIP_BAD_ARGUMENT with the loader's message
 * when an element is missing or the stack is inconsistent.  No device work.
 * Replace: CUDA_glutton_stack_XML_converter::dump (:188-202) / load (:158-186). */
int emp_hip_glutton_stack_to_xml(const emp_hip_puzzle *puzzle, const uint32_t *state, const uint64_t *masks,
                                 char *out, uint64_t capacity, uint64_t *size);
int emp_hip_glutton_stack_from_xml(const emp_hip_puzzle *puzzle, const char *xml, uint64_t size, uint32_t *state,
                                   uint64_t *masks);

/* Seeded prefixes for glutton searches: each row is the n_fixed given moves
 * followed by one placement of each of the first `length` positions of the
 * spiral order, drawn by random dives that only place matching pieces
 * (distinct rows).  out: [count][n_fixed + length] packed moves; *n_out rows
 * were produced.  No counterpart in the reference (its glutton_wide level
 * expansion, include/CUDA_glutton_situations.h:713-757, enumerates instead). */
int emp_hip_spiral_prefixes(const emp_hip_puzzle *puzzle, uint32_t length, uint32_t count,
                            uint64_t seed, const uint32_t *fixed_moves, uint32_t n_fixed,
                            uint32_t *out, uint32_t *n_out);

/* emp_hip_enumerate with a visit order per puzzle: ranks[i] holds P rows
 * (strategy index k) of T_i = 4 x (pieces of puzzle i's largest kind) u16;
 * at index k the candidates are tried by increasing ranks[i][k * T_i + t]
 * (t = 4 * kind_index + orientation), each row a permutation of 0..T_i-1.
 * Index 0 still takes its first candidate in index order.  The situations
 * visited, the node count and the solution set are those of the plain
 * order; only the order of the visits changes.  No counterpart in the
 * reference: it serves the learned candidate ordering. */
int emp_hip_enumerate_ranked(emp_hip_context *context, const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                             const emp_hip_enum_options *options, const uint16_t *const *ranks,
                             emp_hip_enum_result *results, uint16_t *const *boards, emp_hip_enum_stats *stats);

/* Depth probes: one search per row from the situation the row describes --
 * moves (pos | piece << 8 | o << 16) for strategy indices 0..row_len-1, each
 * a candidate of the enumeration after the earlier ones (else
 * EMP_HIP_BAD_ARGUMENT) -- with options->node_limit placements each.
 * results[r].max_depth is the most pieces placed in one situation below the
 * row, n_nodes the placements below it.  No counterpart in the reference: it
 * labels the traces the candidate scorer is trained on. */
int emp_hip_probe(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *moves, uint32_t row_len,
                  uint32_t n_rows, const emp_hip_enum_options *options, emp_hip_enum_result *results,
                  emp_hip_enum_stats *stats);

/* Copies `bytes` (a multiple of 16) between two device buffers of the
 * context's workspace with a plain 16-byte load / store kernel, twice, and
 * returns the second copy's kernel time.  A measurement utility with no
 * counterpart in the reference: HBM byte counters of a profiler read on this
 * kernel give the scale of the same counters on the search kernels. */
int emp_hip_stream_copy(emp_hip_context *context, uint64_t bytes, double *ms);

/* Measured peaks of the two resources the search kernels are bound by: LDS
 * read bandwidth (GB/s, a conflict-free ds_read_b128 stream on every CU) and
 * 32-bit integer VALU throughput (G lane-ops/s, independent v_and_b32 /
 * v_bcnt_u32_b32 chains), each over a second timed launch.  A measurement
 * utility with no counterpart in the reference, for the roofline fractions. */
int emp_hip_peaks(emp_hip_context *context, double *lds_gbs, double *valu_gops);

/* ------------------------------------------------------------------------
 * Candidate scorer (no counterpart in the reference: it only orders the
 * candidates of the search and never removes one).  A 126-256-256-1
 * perceptron, ReLU hidden layers, BF16 dense-layer weights with FP32
 * accumulation on the matrix cores, FP32 master parameters:
 *   W1[128][256] (rows 126 and 127 stay 0), b1[256], W2[256][256], b2[256],
 *   w3[256], b3  -- EMP_HIP_MLP_PARAMS floats in that order.
 * Features are rows of EMP_HIP_MLP_INPUTS floats; targets one float per row.
 * ---------------------------------------------------------------------- */
#define EMP_HIP_MLP_INPUTS 126
#define EMP_HIP_MLP_PARAMS 99073

typedef struct {
    uint32_t epochs;           /* passes over the rows                          */
    uint32_t batch;            /* rows per Adam step, multiple of 64, <= 65536  */
    float    lr;               /* Adam step size                                */
    uint32_t reserved;
    uint64_t seed;             /* batch order                                   */
} emp_hip_mlp_options;

/* Trains params (in/out) on the mean squared error of score vs target;
 * loss_per_epoch (may be NULL) receives each epoch's mean batch loss. */
int emp_hip_mlp_train(emp_hip_context *context, const float *features, const float *targets, uint64_t n_rows,
                      const emp_hip_mlp_options *options, float *params, float *loss_per_epoch);

/* scores[i] = network output for row i. */
int emp_hip_mlp_score(emp_hip_context *context, const float *params, const float *features, uint64_t n_rows,
                      float *scores);

/* Loss and gradient of every parameter over n_rows <= 65536 rows (one batch),
 * for checking the backward kernels. */
int emp_hip_mlp_gradients(emp_hip_context *context, const float *params, const float *features,
                          const float *targets, uint64_t n_rows, float *grads, float *loss);

#ifdef __cplusplus
}
#endif

#endif
