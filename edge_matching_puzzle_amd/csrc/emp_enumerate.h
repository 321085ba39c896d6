// emp_enumerate.h -- host interface of the enumeration path.
#ifndef EMP_ENUMERATE_H
#define EMP_ENUMERATE_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "emp_model.h"

namespace emp {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct OutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Grow-only cache of named device allocations, owned by a context so that
// repeated searches allocate nothing once warm.  In arena form the blocks are
// carved in order from one caller-provided device buffer (256-byte aligned),
// nothing is allocated or freed, and a request beyond its end throws.
class Workspace {
  public:
    Workspace() = default;
    Workspace(void *arena, size_t bytes) : arena_(static_cast<char *>(arena)), arena_bytes_(bytes) {}
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    ~Workspace();
    template <typename T>
    T *get(const std::string &name, size_t count) {
        return static_cast<T *>(raw(name, (count ? count : 1) * sizeof(T)));
    }
    void release();
    bool is_arena() const { return arena_ != nullptr; }
    // Bytes still free in the arena (SIZE_MAX when blocks are allocated).
    size_t remaining() const { return arena_ ? arena_bytes_ - arena_used_ : SIZE_MAX; }
    static constexpr size_t kAlign = 256;

  private:
    void *raw(const std::string &name, size_t bytes);
    struct Block {
        void *ptr = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Block> blocks_;
    char *arena_ = nullptr;
    size_t arena_bytes_ = 0, arena_used_ = 0;
};

struct DeviceInfo {
    int device = 0;
    uint32_t compute_units = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
    uint32_t clock_khz = 0;
};

struct EnumConfig {
    int order = kSpiral;
    uint32_t waves_per_cu = 0;
    uint32_t waves_per_block = 0;
    uint32_t steps_per_round = 0;
    uint64_t step_budget = 0;
    uint64_t drain_budget = 0;
    uint64_t target_items = 0;
    uint64_t node_limit = 0;
    uint64_t solution_capacity = 1 << 16;
    uint32_t max_expansions = 0;
    uint32_t queue_items = 0;       // capacity of each work queue, 0 = sized from the grid
    double time_limit_ms = 0;       // stop every search after this device time (0 = none)
    bool use_graph = false;         // replay the DFS launch(es) from captured hipGraphs
    bool count_work = false;        // accumulate the algorithmic bytes / ops of every placement
    // How the DFS work is shared out: one persistent launch with hand-overs,
    // DFS passes (items split into the next pass's queue), or the first for
    // one puzzle and the second for batches of several.
    enum Schedule { kAuto, kPasses, kPersistent };
    Schedule schedule = kAuto;
};

struct EnumOutcome {
    uint64_t nodes = 0;
    uint64_t solutions = 0;
    uint32_t max_depth = 0;          // most pieces placed in one situation
    bool complete = true;
    double solve_ms = 0;
    std::vector<uint16_t> boards;   // stored solutions, P cells each
    // With ordered results: the stored solutions in the order of the
    // sequential DFS, each with its node index there (the placements up to
    // and including it; exact when the enumeration is complete).
    std::vector<uint64_t> node_index;
};

// A work item of a search: the transitions placed at strategy indices
// 0..level-1 (stack), resumed at index `level` above transition `last` (-1:
// every candidate), ending when index `floor` has no candidate left -- the
// state of the reference's DFS loop (include/emp_strategy.h:332-415) with
// the loop stopping when it backtracks below `floor`.
struct EnumItem {
    uint32_t floor = 1, level = 1;
    int last = -1;
    std::vector<uint16_t> stack;
};

// One puzzle's enumeration given as work items, with ordered results and
// resumable cuts.
struct EnumWork {
    std::vector<EnumItem> items;      // empty: the root item (index 0's first candidate)
    bool record_order = false;        // node index of each stored solution in DFS order
    bool keep_rest = false;           // at the deadline or the node limit, return the unfinished work in `rest`
    uint64_t max_executions = 1u << 22;   // item executions an ordered run may record
    std::vector<EnumItem> rest;       // out: what remains when a limit cut the run
};

struct EnumStats {
    double total_ms = 0, dfs_kernel_ms = 0, expand_kernel_ms = 0;
    uint64_t dfs_launches = 0, expand_launches = 0, items_processed = 0, splits = 0;
    uint32_t max_words = 0, lds_tables = 0, blocks = 0, threads = 0;
    uint64_t candidate_sets = 0;   // DFS iterations (one candidate set each)
    double group_ms = 0;           // grouping items by puzzle before the passes
    uint64_t work_bytes = 0;       // count_work: algorithmic LDS bytes of the placements
    uint64_t work_ops = 0;         // count_work: algorithmic integer ops of the placements
    uint64_t tickets = 0;          // persistent launch: items searched (starting items and hand-overs)
};

// A search that starts from a given situation of puzzle `table`: the
// transitions placed at strategy indices 0..stack.size()-1.
struct EnumRoot {
    uint32_t table = 0;
    std::vector<uint16_t> stack;
};

// Enumerates every puzzle from its root (index 0 holds its first candidate),
// or, with `roots`, each given situation (outcomes per root; their own
// placements are not counted).  With `ranks` (one table per puzzle, P rows of
// 4 x (pieces of the largest kind) u16), candidates are visited in rank order
// instead of index order; the set of situations visited is the same.  With
// `work` (one puzzle, no roots or ranks), the search is the union of its
// items (one outcome), optionally with ordered results and a resumable cut.
void enumerate_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws,
                         const std::vector<const Puzzle *> &puzzles, const EnumConfig &cfg,
                         std::vector<EnumOutcome> &out, EnumStats &stats,
                         const std::vector<EnumRoot> *roots = nullptr,
                         const std::vector<std::vector<uint16_t>> *ranks = nullptr, EnumWork *work = nullptr);

// Breadth expansion (glutton_wide's play_from applied level by level,
// include/CUDA_glutton_situations.h:713-757): every situation the DFS visits
// with `depth` pieces placed, as row-major boards (up to `capacity`).  nodes
// and solutions count the placements of indices 0..depth-1.
void expand_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &puzzle, int order,
                      uint32_t depth, uint64_t capacity, std::vector<uint16_t> &boards, uint64_t &n_items,
                      uint64_t &nodes, uint64_t &solutions);

}  // namespace emp

#endif
