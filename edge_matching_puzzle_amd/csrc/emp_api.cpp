// emp_api.cpp -- extern "C" entry points of libemp_hip.so (declared in
// include/emp_hip.h).  Exceptions never cross the ABI: each entry point maps
// them to a status code and keeps the message for emp_hip_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/emp_hip.h"
#include "emp_backtrack.h"
#include "emp_enum_plan.h"
#include "emp_enumerate.h"
#include "emp_glutton.h"
#include "emp_glutton_stack.h"
#include "emp_glutton_wide.h"
#include "emp_host_threads.h"
#include "emp_mlp.h"
#include "emp_model.h"
#include "emp_profile.h"
#include "emp_results.h"

#include <unordered_set>

struct emp_hip_context {
    emp::DeviceInfo dev;
    hipStream_t stream = nullptr;
    emp::Workspace workspace;
    void *scratch = nullptr;   // device address of this device's copy of `scratch`
};

// Fixed device buffer of the allocation-free enumeration
// (emp_hip_enumerate_noalloc).  Each device holds its own copy from module
// load until the process ends.  In host code the name is only a stand-in for
// the symbol: its device address is taken once per device with
// hipGetSymbolAddress (scratch_address below), and only that address is ever
// passed to kernels or copies.
static __device__ __align__(256) uint8_t scratch[EMP_HIP_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mutex;
std::map<int, void *> g_scratch_address;   // device -> address of its `scratch`

// Address of `scratch` on the current device (which must be `device`),
// looked up on the first call for that device.
void *scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_address.find(device);
    if (it != g_scratch_address.end()) return it->second;
    void *ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_scratch_address[device] = ptr;
    return ptr;
}

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

emp::EnumConfig config_from(const emp_hip_enum_options &opt) {
    emp::EnumConfig cfg;
    cfg.order = (int)opt.order;
    cfg.waves_per_cu = opt.waves_per_cu;
    cfg.waves_per_block = opt.waves_per_block;
    cfg.steps_per_round = opt.steps_per_round;
    cfg.step_budget = opt.step_budget;
    cfg.target_items = opt.target_items;
    cfg.node_limit = opt.node_limit;
    cfg.queue_items = opt.queue_items;
    cfg.time_limit_ms = opt.time_limit_ms;
    cfg.use_graph = opt.use_graph != 0;
    cfg.count_work = opt.count_work != 0;
    cfg.schedule = opt.schedule == EMP_HIP_SCHEDULE_PASSES       ? emp::EnumConfig::kPasses
                   : opt.schedule == EMP_HIP_SCHEDULE_PERSISTENT ? emp::EnumConfig::kPersistent
                                                                 : emp::EnumConfig::kAuto;
    // Device-side store of up to max_solutions boards per search.
    cfg.solution_capacity = opt.max_solutions;
    return cfg;
}

// Internal form of a public work item (moves checked to be candidates).
emp::EnumItem to_item(const emp::EnumTable &t, const emp::Puzzle &pz, const emp_hip_item &in, const uint32_t *moves) {
    if (in.level >= pz.size() || in.floor < 1 || in.floor > in.level)
        throw std::invalid_argument("a work item needs 1 <= floor <= level < P");
    emp::EnumItem it;
    it.floor = in.floor;
    it.level = in.level;
    emp::moves_to_stack(t, pz, moves, in.level, it.stack);
    it.last = in.last == EMP_HIP_NO_MOVE ? -1 : (int)emp::move_to_transition(t, pz, in.level, in.last);
    return it;
}

void copy_stats(const emp::EnumStats &st, emp_hip_enum_stats *stats) {
    if (!stats) return;
    stats->total_ms = st.total_ms;
    stats->dfs_kernel_ms = st.dfs_kernel_ms;
    stats->expand_kernel_ms = st.expand_kernel_ms;
    stats->dfs_launches = st.dfs_launches;
    stats->expand_launches = st.expand_launches;
    stats->items_processed = st.items_processed;
    stats->splits = st.splits;
    stats->tickets = st.tickets;
    stats->candidate_sets = st.candidate_sets;
    stats->group_ms = st.group_ms;
    stats->work_bytes = st.work_bytes;
    stats->work_ops = st.work_ops;
    stats->max_words = st.max_words;
    stats->lds_tables = st.lds_tables;
    stats->blocks = st.blocks;
    stats->threads = st.threads;
}

template <typename F>
int guarded(F &&f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const std::invalid_argument &e) {
        return fail(EMP_HIP_BAD_PUZZLE, e.what());
    } catch (const std::domain_error &e) {
        return fail(EMP_HIP_UNSUPPORTED, e.what());
    } catch (const emp::OutOfMemory &e) {
        return fail(EMP_HIP_OUT_OF_MEMORY, e.what());
    } catch (const emp::Unsupported &e) {
        return fail(EMP_HIP_UNSUPPORTED, e.what());
    } catch (const emp::DeviceError &e) {
        return fail(EMP_HIP_DEVICE_ERROR, e.what());
    } catch (const std::bad_alloc &) {
        return fail(EMP_HIP_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(EMP_HIP_DEVICE_ERROR, e.what());
    }
}

}  // namespace

extern "C" {

const char *emp_hip_last_error(void) { return g_last_error.c_str(); }

const char *emp_hip_version(void) { return "emp_hip 0.1 gfx950"; }

int emp_hip_create(int device, emp_hip_context **context) {
    return guarded([&] {
        if (!context) return fail(EMP_HIP_BAD_ARGUMENT, "null context pointer");
        *context = nullptr;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
            return fail(EMP_HIP_DEVICE_ERROR, "no HIP device available");
        if (device < 0 || device >= count) return fail(EMP_HIP_BAD_ARGUMENT, "device index out of range");
        if (hipSetDevice(device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess)
            return fail(EMP_HIP_DEVICE_ERROR, "hipGetDeviceProperties failed");
        std::unique_ptr<emp_hip_context> ctx(new emp_hip_context);
        ctx->dev.device = device;
        ctx->dev.compute_units = (uint32_t)prop.multiProcessorCount;
        ctx->dev.lds_per_block = prop.sharedMemPerBlock;
        ctx->dev.lds_per_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                    : prop.sharedMemPerBlock;
        ctx->dev.clock_khz = (uint32_t)prop.clockRate;
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
            return fail(EMP_HIP_DEVICE_ERROR, "hipStreamCreate failed");
        ctx->scratch = scratch_address(device);
        if (!ctx->scratch) {
            (void)hipStreamDestroy(ctx->stream);
            return fail(EMP_HIP_DEVICE_ERROR, "hipGetSymbolAddress of the scratch buffer failed");
        }
        *context = ctx.release();
        return EMP_HIP_OK;
    });
}

int emp_hip_destroy(emp_hip_context *context) {
    return guarded([&] {
        if (!context) return EMP_HIP_OK;
        (void)hipSetDevice(context->dev.device);
        if (context->stream) (void)hipStreamSynchronize(context->stream);
        context->workspace.release();
        if (context->stream) (void)hipStreamDestroy(context->stream);
        delete context;
        return EMP_HIP_OK;
    });
}

int emp_hip_device_info(emp_hip_context *context, uint32_t *compute_units,
                        uint32_t *lds_per_block, uint32_t *clock_khz) {
    return guarded([&] {
        if (!context) return fail(EMP_HIP_BAD_ARGUMENT, "null context");
        if (compute_units) *compute_units = context->dev.compute_units;
        if (lds_per_block) *lds_per_block = (uint32_t)context->dev.lds_per_block;
        if (clock_khz) *clock_khz = context->dev.clock_khz;
        return EMP_HIP_OK;
    });
}

int emp_hip_generate(uint64_t seed, uint32_t width, uint32_t height,
                     uint32_t border_colours, uint32_t interior_colours,
                     uint32_t *colours, uint16_t *planted) {
    return guarded([&] {
        if (!colours) return fail(EMP_HIP_BAD_ARGUMENT, "null colour buffer");
        std::vector<uint16_t> pl;
        std::vector<uint32_t> c = emp::generate(seed, width, height, border_colours, interior_colours,
                                                planted ? &pl : nullptr);
        std::copy(c.begin(), c.end(), colours);
        if (planted) std::copy(pl.begin(), pl.end(), planted);
        return EMP_HIP_OK;
    });
}

int emp_hip_validate_puzzle(const emp_hip_puzzle *puzzle) {
    return guarded([&] {
        if (!puzzle) return fail(EMP_HIP_BAD_ARGUMENT, "null puzzle");
        emp::Puzzle p(puzzle->width, puzzle->height, puzzle->colours);
        (void)p;
        return EMP_HIP_OK;
    });
}

int emp_hip_order(uint32_t width, uint32_t height, uint32_t order, uint32_t *xy) {
    return guarded([&] {
        if (!xy || width < 1 || height < 1) return fail(EMP_HIP_BAD_ARGUMENT, "bad order arguments");
        if (order != EMP_HIP_ORDER_SPIRAL && order != EMP_HIP_ORDER_BASIC)
            return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        auto o = emp::make_order(width, height, (int)order);
        if (!emp::is_permutation(o, width, height))
            return fail(EMP_HIP_UNSUPPORTED, "the spiral order is not defined for a " + std::to_string(width) + "x" +
                                                 std::to_string(height) + " board (the walk revisits cells)");
        for (size_t i = 0; i < o.size(); ++i) {
            xy[2 * i] = o[i].first;
            xy[2 * i + 1] = o[i].second;
        }
        return EMP_HIP_OK;
    });
}

int emp_hip_situation_string(uint32_t width, uint32_t height, const uint16_t *board,
                             char *out, uint64_t capacity) {
    return guarded([&] {
        if (!board || !out || !capacity) return fail(EMP_HIP_BAD_ARGUMENT, "bad buffers");
        std::string s = emp::situation_string(width * height, board);
        if (s.size() + 1 > capacity) return fail(EMP_HIP_BAD_ARGUMENT, "output buffer too small");
        std::memcpy(out, s.c_str(), s.size() + 1);
        return EMP_HIP_OK;
    });
}

void emp_hip_enum_default_options(emp_hip_enum_options *options) {
    if (!options) return;
    std::memset(options, 0, sizeof(*options));
    options->order = EMP_HIP_ORDER_SPIRAL;
    options->max_solutions = 1024;
}

namespace {

// Shared body of emp_hip_enumerate, emp_hip_enumerate_ranked and
// emp_hip_probe.  With probe rows, the single puzzle is searched from each
// row's situation and results are per row.
int run_enumeration(emp_hip_context *context, const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                    const emp_hip_enum_options *options, const uint16_t *const *ranks, const uint32_t *probe_moves,
                    uint32_t probe_len, uint32_t probe_rows, emp_hip_enum_result *results, uint16_t *const *boards,
                    emp_hip_enum_stats *stats, emp::Workspace *arena = nullptr) {
    if (!context || !puzzles || !results || !n_puzzles) return fail(EMP_HIP_BAD_ARGUMENT, "null argument or empty batch");
    emp_hip_enum_options opt;
    emp_hip_enum_default_options(&opt);
    if (options) opt = *options;
    if (opt.order != EMP_HIP_ORDER_SPIRAL && opt.order != EMP_HIP_ORDER_BASIC)
        return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
    if (opt.waves_per_block > 16) return fail(EMP_HIP_BAD_ARGUMENT, "waves_per_block must be at most 16");
    if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
    // Checked on the host threads; the first puzzle (in batch order) that
    // fails is reported.
    std::vector<std::unique_ptr<emp::Puzzle>> owned(n_puzzles);
    std::vector<std::string> bad(n_puzzles);
    emp::for_each_item(n_puzzles, [&](uint32_t i) {
        try {
            owned[i].reset(new emp::Puzzle(puzzles[i].width, puzzles[i].height, puzzles[i].colours));
        } catch (const std::invalid_argument &e) {
            bad[i] = e.what();
        }
    });
    std::vector<const emp::Puzzle *> list;
    for (uint32_t i = 0; i < n_puzzles; ++i) {
        if (!owned[i]) return fail(EMP_HIP_BAD_PUZZLE, "puzzle " + std::to_string(i) + ": " + bad[i]);
        list.push_back(owned[i].get());
    }
    std::vector<std::vector<uint16_t>> rank_tables;
    if (ranks) {
        for (uint32_t i = 0; i < n_puzzles; ++i) {
            if (!ranks[i]) return fail(EMP_HIP_BAD_ARGUMENT, "null rank table");
            uint32_t tmax = 0;
            for (int k = 0; k < 3; ++k) tmax = std::max<uint32_t>(tmax, 4 * (uint32_t)list[i]->pieces_of(k).size());
            rank_tables.emplace_back(ranks[i], ranks[i] + (size_t)list[i]->size() * tmax);
            // Each row must be a permutation of 0..T-1: two equal ranks would
            // make "above the last rank" skip a transition.
            std::vector<uint8_t> seen(tmax);
            for (uint32_t k = 0; k < list[i]->size(); ++k) {
                std::fill(seen.begin(), seen.end(), 0);
                const uint16_t *row = rank_tables.back().data() + (size_t)k * tmax;
                for (uint32_t t = 0; t < tmax; ++t) {
                    if (row[t] >= tmax || seen[row[t]])
                        return fail(EMP_HIP_BAD_ARGUMENT, "rank table " + std::to_string(i) + " row " +
                                                              std::to_string(k) + " is not a permutation of 0.." +
                                                              std::to_string(tmax - 1));
                    seen[row[t]] = 1;
                }
            }
        }
    }
    std::vector<emp::EnumRoot> roots;
    if (probe_moves) {
        const emp::EnumTable t = emp::build_enum_table(*list[0], (int)opt.order);
        roots.resize(probe_rows);
        for (uint32_t r = 0; r < probe_rows; ++r) {
            try {
                emp::moves_to_stack(t, *list[0], probe_moves + (size_t)r * probe_len, probe_len, roots[r].stack);
            } catch (const std::invalid_argument &e) {
                return fail(EMP_HIP_BAD_ARGUMENT, "row " + std::to_string(r) + ": " + e.what());
            }
        }
    }
    if (!(opt.time_limit_ms >= 0)) return fail(EMP_HIP_BAD_ARGUMENT, "time_limit_ms must be >= 0");
    const emp::EnumConfig cfg = config_from(opt);
    std::vector<emp::EnumOutcome> out;
    emp::EnumStats st;
    emp::enumerate_on_device(context->dev, context->stream, arena ? *arena : context->workspace, list, cfg, out, st,
                             probe_moves ? &roots : nullptr, ranks ? &rank_tables : nullptr);
    for (uint32_t i = 0; i < (uint32_t)out.size(); ++i) {
        emp_hip_enum_result &r = results[i];
        const uint32_t P = list[probe_moves ? 0 : i]->size();
        const uint64_t have = out[i].boards.size() / P;
        r.n_nodes = out[i].nodes;
        r.n_solutions = out[i].solutions;
        r.complete = out[i].complete ? 1 : 0;
        r.status = EMP_HIP_OK;
        r.solve_ms = out[i].solve_ms;
        r.max_depth = out[i].max_depth;
        r.reserved = 0;
        const uint64_t keep = std::min<uint64_t>(have, opt.max_solutions);
        r.n_stored = (boards && boards[i]) ? keep : 0;
        if (boards && boards[i] && keep) std::memcpy(boards[i], out[i].boards.data(), keep * P * sizeof(uint16_t));
    }
    copy_stats(st, stats);
    return EMP_HIP_OK;
}

}  // namespace

int emp_hip_enumerate(emp_hip_context *context, const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                      const emp_hip_enum_options *options, emp_hip_enum_result *results,
                      uint16_t *const *boards, emp_hip_enum_stats *stats) {
    return guarded([&] {
        return run_enumeration(context, puzzles, n_puzzles, options, nullptr, nullptr, 0, 0, results, boards, stats);
    });
}

int emp_hip_enumerate_noalloc(emp_hip_context *context, const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                              const emp_hip_enum_options *options, emp_hip_enum_result *results,
                              uint16_t *const *boards, emp_hip_enum_stats *stats) {
    return guarded([&] {
        if (!context) return fail(EMP_HIP_BAD_ARGUMENT, "null context");
        if (!context->scratch) return fail(EMP_HIP_DEVICE_ERROR, "no scratch buffer address for this device");
        if (options && options->use_graph)
            return fail(EMP_HIP_BAD_ARGUMENT, "use_graph allocates graph objects: not available without allocation");
        // Blocks are carved from this device's scratch buffer for this call
        // only; nothing is allocated or freed.
        emp::Workspace arena(context->scratch, EMP_HIP_SCRATCH_BYTES);
        return run_enumeration(context, puzzles, n_puzzles, options, nullptr, nullptr, 0, 0, results, boards, stats,
                               &arena);
    });
}

int emp_hip_enumerate_items(emp_hip_context *context, const emp_hip_puzzle *puzzle,
                            const emp_hip_enum_options *options,
                            const emp_hip_item *items, const uint32_t *item_moves, uint32_t n_items,
                            emp_hip_enum_result *result, uint16_t *boards, uint64_t *node_index,
                            emp_hip_item *rest, uint32_t *rest_moves, uint64_t rest_capacity,
                            uint64_t *n_rest, emp_hip_enum_stats *stats) {
    return guarded([&] {
        if (!context || !puzzle || !result) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (n_items && (!items || !item_moves)) return fail(EMP_HIP_BAD_ARGUMENT, "null work items");
        if (rest && (!rest_moves || !n_rest)) return fail(EMP_HIP_BAD_ARGUMENT, "null rest buffers");
        emp_hip_enum_options opt;
        emp_hip_enum_default_options(&opt);
        if (options) opt = *options;
        if (opt.order != EMP_HIP_ORDER_SPIRAL && opt.order != EMP_HIP_ORDER_BASIC)
            return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        if (!(opt.time_limit_ms >= 0)) return fail(EMP_HIP_BAD_ARGUMENT, "time_limit_ms must be >= 0");
        if (rest && !(opt.time_limit_ms > 0) && !opt.node_limit)
            return fail(EMP_HIP_BAD_ARGUMENT, "the rest of a search needs a time or node limit");
        if (node_index && opt.use_graph) return fail(EMP_HIP_BAD_ARGUMENT, "ordered results are not recorded from graph replays");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        std::unique_ptr<emp::Puzzle> pz;
        try {
            pz.reset(new emp::Puzzle(puzzle->width, puzzle->height, puzzle->colours));
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_PUZZLE, e.what());
        }
        const uint32_t P = pz->size();
        const emp::EnumTable t = emp::build_enum_table(*pz, (int)opt.order);
        emp::EnumWork work;
        work.record_order = node_index != nullptr;
        work.keep_rest = rest != nullptr;
        for (uint32_t i = 0; i < n_items; ++i) {
            try {
                work.items.push_back(to_item(t, *pz, items[i], item_moves + (size_t)i * P));
            } catch (const std::invalid_argument &e) {
                return fail(EMP_HIP_BAD_ARGUMENT, "item " + std::to_string(i) + ": " + e.what());
            }
        }
        std::vector<const emp::Puzzle *> list{pz.get()};
        std::vector<emp::EnumOutcome> out;
        emp::EnumStats st;
        emp::enumerate_on_device(context->dev, context->stream, context->workspace, list, config_from(opt), out, st,
                                 nullptr, nullptr, &work);
        const emp::EnumOutcome &o = out[0];
        const uint64_t keep = std::min<uint64_t>(o.boards.size() / P, opt.max_solutions);
        result->n_nodes = o.nodes;
        result->n_solutions = o.solutions;
        result->complete = o.complete ? 1 : 0;
        result->status = EMP_HIP_OK;
        result->solve_ms = o.solve_ms;
        result->max_depth = o.max_depth;
        result->reserved = 0;
        result->n_stored = boards ? keep : 0;
        if (boards && keep) std::memcpy(boards, o.boards.data(), keep * P * sizeof(uint16_t));
        if (node_index && boards && keep) std::memcpy(node_index, o.node_index.data(), keep * sizeof(uint64_t));
        copy_stats(st, stats);
        if (rest) {
            *n_rest = work.rest.size();
            const uint64_t m = std::min<uint64_t>(work.rest.size(), rest_capacity);
            for (uint64_t i = 0; i < m; ++i) {
                const emp::EnumItem &it = work.rest[i];
                rest[i].floor = it.floor;
                rest[i].level = it.level;
                rest[i].last = it.last < 0 ? EMP_HIP_NO_MOVE : emp::transition_to_move(t, it.level, (uint32_t)it.last);
                rest[i].reserved = 0;
                uint32_t *row = rest_moves + (size_t)i * P;
                for (uint32_t j = 0; j < P; ++j)
                    row[j] = j < it.level ? emp::transition_to_move(t, j, it.stack[j]) : EMP_HIP_NO_MOVE;
            }
            if (work.rest.size() > rest_capacity)
                return fail(EMP_HIP_OUT_OF_MEMORY, std::to_string(work.rest.size()) + " work items left, room for " +
                                                       std::to_string(rest_capacity));
        }
        return EMP_HIP_OK;
    });
}

int emp_hip_resume_item(const emp_hip_puzzle *puzzle, uint32_t order, const uint16_t *board, emp_hip_item *item,
                        uint32_t *moves, uint32_t *n_items) {
    return guarded([&] {
        if (!puzzle || !board || !item || !moves || !n_items) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (order != EMP_HIP_ORDER_SPIRAL && order != EMP_HIP_ORDER_BASIC) return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        std::unique_ptr<emp::Puzzle> pz;
        try {
            pz.reset(new emp::Puzzle(puzzle->width, puzzle->height, puzzle->colours));
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_PUZZLE, e.what());
        }
        const uint32_t P = pz->size();
        const emp::EnumTable t = emp::build_enum_table(*pz, (int)order);
        // leading strategy positions holding a piece (set_initial_state stops
        // at the first empty one)
        std::vector<uint32_t> mv;
        for (uint32_t i = 0; i < P; ++i) {
            const uint32_t cell = t.order[i].second * puzzle->width + t.order[i].first;
            if (board[cell] == 0xFFFF) break;
            mv.push_back(cell | ((uint32_t)(board[cell] >> 2) << 8) | ((uint32_t)(board[cell] & 3) << 16));
        }
        if (mv.empty()) return fail(EMP_HIP_BAD_ARGUMENT, "the situation has no piece at the first strategy position");
        std::vector<uint16_t> stack;
        try {
            emp::moves_to_stack(t, *pz, mv.data(), (uint32_t)mv.size(), stack);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, std::string("the situation is not one the search visits: ") + e.what());
        }
        if (mv.size() == 1) {
            *n_items = 0;
            return EMP_HIP_OK;
        }
        const uint32_t k = (uint32_t)mv.size();
        item->floor = 1;
        item->level = k - 1;
        item->last = EMP_HIP_NO_MOVE;
        item->reserved = 0;
        for (uint32_t j = 0; j < P; ++j) moves[j] = j < k - 1 ? mv[j] : EMP_HIP_NO_MOVE;
        *n_items = 1;
        return EMP_HIP_OK;
    });
}

int emp_hip_parse_situation(uint32_t width, uint32_t height, const char *text, uint16_t *board) {
    return guarded([&] {
        if (!text || !board || !width || !height) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        std::vector<uint16_t> b;
        try {
            b = emp::parse_situation(width, height, text);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        std::copy(b.begin(), b.end(), board);
        return EMP_HIP_OK;
    });
}

int emp_hip_results_bin_write(uint32_t width, uint32_t height, uint32_t order, uint32_t solution_dump,
                              const uint16_t *boards, const uint64_t *node_index, uint64_t n_records,
                              uint64_t total_nodes, uint8_t *out, uint64_t capacity, uint64_t *size) {
    return guarded([&] {
        if (!size || (n_records && !boards)) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (order != EMP_HIP_ORDER_SPIRAL && order != EMP_HIP_ORDER_BASIC) return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        const auto xy = emp::make_order(width, height, (int)order);
        if (!emp::is_permutation(xy, width, height))
            return fail(EMP_HIP_UNSUPPORTED, "the order does not visit every cell once");
        std::vector<uint8_t> bytes;
        try {
            bytes = emp::write_results_bin(width, height, xy, solution_dump != 0, boards, node_index, n_records,
                                           total_nodes);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        *size = bytes.size();
        if (out) std::memcpy(out, bytes.data(), std::min<uint64_t>(capacity, bytes.size()));
        return EMP_HIP_OK;
    });
}

int emp_hip_results_bin_read(const uint8_t *data, uint64_t size, uint32_t *width, uint32_t *height,
                             uint32_t *solution_dump, uint32_t *xy, uint16_t *boards, uint64_t *node_index,
                             uint64_t capacity, uint64_t *n_records, uint64_t *total_nodes, uint32_t *complete) {
    return guarded([&] {
        if (!data || !width || !height || !n_records) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        emp::ResultsBin r;
        try {
            r = emp::read_results_bin(data, size);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        const uint64_t P = (uint64_t)r.width * r.height, n = r.node_index.size(), m = std::min<uint64_t>(n, capacity);
        *width = r.width;
        *height = r.height;
        *n_records = n;
        if (solution_dump) *solution_dump = r.solution_dump ? 1 : 0;
        if (total_nodes) *total_nodes = r.total_nodes;
        if (complete) *complete = r.complete ? 1 : 0;
        if (xy)
            for (uint64_t i = 0; i < P; ++i) xy[2 * i] = r.order[i].first, xy[2 * i + 1] = r.order[i].second;
        if (boards && m) std::memcpy(boards, r.boards.data(), m * P * sizeof(uint16_t));
        if (node_index && m) std::memcpy(node_index, r.node_index.data(), m * sizeof(uint64_t));
        return EMP_HIP_OK;
    });
}

int emp_hip_enumerate_ranked(emp_hip_context *context, const emp_hip_puzzle *puzzles, uint32_t n_puzzles,
                             const emp_hip_enum_options *options, const uint16_t *const *ranks,
                             emp_hip_enum_result *results, uint16_t *const *boards, emp_hip_enum_stats *stats) {
    return guarded([&] {
        if (!ranks) return fail(EMP_HIP_BAD_ARGUMENT, "null rank tables");
        return run_enumeration(context, puzzles, n_puzzles, options, ranks, nullptr, 0, 0, results, boards, stats);
    });
}

int emp_hip_probe(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *moves, uint32_t row_len,
                  uint32_t n_rows, const emp_hip_enum_options *options, emp_hip_enum_result *results,
                  emp_hip_enum_stats *stats) {
    return guarded([&] {
        if (!moves || !n_rows || !row_len) return fail(EMP_HIP_BAD_ARGUMENT, "no rows");
        if (n_rows > (1u << 20)) return fail(EMP_HIP_BAD_ARGUMENT, "at most 2^20 rows");
        if (puzzle && row_len >= puzzle->width * puzzle->height)
            return fail(EMP_HIP_BAD_ARGUMENT, "rows place between 1 and P - 1 pieces");
        return run_enumeration(context, puzzle, 1, options, nullptr, moves, row_len, n_rows, results, nullptr, stats);
    });
}

int emp_hip_backtrack_trace(emp_hip_context *context, const emp_hip_puzzle *puzzle, uint32_t order,
                            uint64_t max_selections, uint16_t *board, uint32_t *found, uint32_t *trace,
                            uint64_t trace_capacity, uint64_t *trace_length) {
    return guarded([&] {
        if (!context || !puzzle || !board || !found || (trace_capacity && !trace))
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (order != EMP_HIP_ORDER_SPIRAL && order != EMP_HIP_ORDER_BASIC) return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        emp::Puzzle pz(puzzle->width, puzzle->height, puzzle->colours);
        emp::BacktrackOutcome out;
        out.trace_capacity = trace_capacity;
        out.max_selections = max_selections;
        emp::backtrack_on_device(context->dev, context->stream, context->workspace, pz, (int)order, out);
        std::copy(out.board.begin(), out.board.end(), board);
        *found = out.found ? 1 : 0;
        if (trace_capacity) std::copy(out.trace.begin(), out.trace.end(), trace);
        if (trace_length) *trace_length = out.trace_length;
        return EMP_HIP_OK;
    });
}

int emp_hip_expand(emp_hip_context *context, const emp_hip_puzzle *puzzle, uint32_t order, uint32_t depth,
                   uint64_t capacity, uint16_t *boards, uint64_t *n_situations, uint64_t *n_nodes,
                   uint64_t *n_solutions) {
    return guarded([&] {
        if (!context || !puzzle || !n_situations || (capacity && !boards))
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (order != EMP_HIP_ORDER_SPIRAL && order != EMP_HIP_ORDER_BASIC) return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        emp::Puzzle pz(puzzle->width, puzzle->height, puzzle->colours);
        std::vector<uint16_t> out;
        uint64_t items = 0, nodes = 0, sols = 0;
        emp::expand_on_device(context->dev, context->stream, context->workspace, pz, (int)order, depth, capacity, out,
                              items, nodes, sols);
        if (capacity) std::copy(out.begin(), out.end(), boards);
        *n_situations = items;
        if (n_nodes) *n_nodes = nodes;
        if (n_solutions) *n_solutions = sols;
        return EMP_HIP_OK;
    });
}

int emp_hip_glutton_max(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *prefixes,
                        uint32_t prefix_len, uint32_t n_searches, const emp_hip_glutton_options *options,
                        emp_hip_glutton_result *results, uint16_t *boards, emp_hip_glutton_stats *stats) {
    return emp_hip_glutton_max_stacks(context, puzzle, prefixes, prefix_len, n_searches, options, nullptr, nullptr,
                                      nullptr, nullptr, results, boards, stats);
}

int emp_hip_glutton_stack_words(uint32_t n_pieces, uint32_t *state_words, uint64_t *mask_words) {
    return guarded([&] {
        if (!state_words || !mask_words) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (!n_pieces || n_pieces > 256) return fail(EMP_HIP_UNSUPPORTED, "glutton_max handles 1 to 256 pieces");
        *state_words = emp::glutton_state_words(n_pieces);
        *mask_words = emp::glutton_mask_words(n_pieces);
        return EMP_HIP_OK;
    });
}

int emp_hip_glutton_stack_to_xml(const emp_hip_puzzle *puzzle, const uint32_t *state, const uint64_t *masks,
                                 char *out, uint64_t capacity, uint64_t *size) {
    return guarded([&] {
        if (!puzzle || !state || !masks || !size) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        emp::Puzzle pz(puzzle->width, puzzle->height, puzzle->colours);
        if (pz.size() > 256) return fail(EMP_HIP_UNSUPPORTED, "glutton_max handles at most 256 pieces");
        std::string xml;
        try {
            xml = emp::glutton_stack_to_xml(pz, state, masks);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        *size = xml.size();
        if (out) std::memcpy(out, xml.data(), std::min<uint64_t>(capacity, xml.size()));
        return EMP_HIP_OK;
    });
}

int emp_hip_glutton_stack_from_xml(const emp_hip_puzzle *puzzle, const char *xml, uint64_t size, uint32_t *state,
                                   uint64_t *masks) {
    return guarded([&] {
        if (!puzzle || !xml || !state || !masks) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        emp::Puzzle pz(puzzle->width, puzzle->height, puzzle->colours);
        if (pz.size() > 256) return fail(EMP_HIP_UNSUPPORTED, "glutton_max handles at most 256 pieces");
        try {
            emp::glutton_stack_from_xml(pz, std::string(xml, xml + size), state, masks);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        return EMP_HIP_OK;
    });
}

int emp_hip_glutton_max_stacks(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *prefixes,
                               uint32_t prefix_len, uint32_t n_searches, const emp_hip_glutton_options *options,
                               const uint32_t *in_state, const uint64_t *in_masks, uint32_t *out_state,
                               uint64_t *out_masks, emp_hip_glutton_result *results, uint16_t *boards,
                               emp_hip_glutton_stats *stats) {
    return guarded([&] {
        if (!context || !puzzle || !results || !boards || !n_searches)
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument or no search");
        if (prefix_len && !prefixes) return fail(EMP_HIP_BAD_ARGUMENT, "null prefixes");
        if ((in_state == nullptr) != (in_masks == nullptr) || (out_state == nullptr) != (out_masks == nullptr))
            return fail(EMP_HIP_BAD_ARGUMENT, "a stack needs both its state and its masks");
        if (in_state && prefix_len) return fail(EMP_HIP_BAD_ARGUMENT, "resumed searches play no prefix");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        emp::Puzzle pz(puzzle->width, puzzle->height, puzzle->colours);
        if (pz.size() > 256) return fail(EMP_HIP_UNSUPPORTED, "glutton_max handles at most 256 pieces");
        if (prefix_len >= pz.size()) return fail(EMP_HIP_BAD_ARGUMENT, "prefix as long as the board");
        emp::GluttonConfig cfg;
        if (options) {
            cfg.push_budget = options->push_budget;
            cfg.waves_per_cu = options->waves_per_cu;
        }
        std::vector<uint32_t> pre(prefixes, prefixes + (size_t)prefix_len * n_searches);
        std::vector<emp::GluttonOutcome> out;
        emp::GluttonStats st;
        emp::GluttonCheckpoints ck;
        ck.in_state = in_state;
        ck.in_masks = in_masks;
        ck.out_state = out_state;
        ck.out_masks = out_masks;
        if (in_state) {
            const uint32_t sw = emp::glutton_state_words(pz.size());
            const uint64_t mw = emp::glutton_mask_words(pz.size());
            for (uint32_t i = 0; i < n_searches; ++i) {
                try {
                    emp::validate_glutton_stack(pz, in_state + (size_t)i * sw, in_masks + (size_t)i * mw);
                } catch (const std::invalid_argument &e) {
                    return fail(EMP_HIP_BAD_ARGUMENT, "stack " + std::to_string(i) + ": " + e.what());
                }
            }
        }
        emp::glutton_on_device(context->dev, context->stream, context->workspace, pz, pre, prefix_len, n_searches,
                               cfg, out, st, ck);
        for (uint32_t i = 0; i < n_searches; ++i) {
            results[i].max_depth = out[i].max_depth;
            results[i].status = out[i].status;
            results[i].pushes = out[i].pushes;
            if (out[i].status > EMP_HIP_GLUTTON_BAD_PREFIX)
                return fail(EMP_HIP_DEVICE_ERROR, "glutton search " + std::to_string(i) + " stopped on an unplayable best move");
            std::copy(out[i].board.begin(), out[i].board.end(), boards + (size_t)i * pz.size());
        }
        if (stats) {
            stats->total_ms = st.total_ms;
            stats->kernel_ms = st.kernel_ms;
            stats->pushes = st.pushes;
            stats->evaluations = st.evaluations;
            stats->score_evaluations = st.score_evaluations;
            stats->rounds = st.rounds;
            stats->free_slot_sum = st.free_slot_sum;
            stats->waves = st.waves;
            stats->lds_bytes = st.lds_bytes;
        }
        return EMP_HIP_OK;
    });
}

int emp_hip_backtrack_variables(const emp_hip_puzzle *puzzle, uint32_t order, uint64_t *n_variables) {
    return guarded([&] {
        if (!puzzle || !n_variables) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (order != EMP_HIP_ORDER_SPIRAL && order != EMP_HIP_ORDER_BASIC) return fail(EMP_HIP_BAD_ARGUMENT, "unknown order");
        std::unique_ptr<emp::Puzzle> pz;
        try {
            pz.reset(new emp::Puzzle(puzzle->width, puzzle->height, puzzle->colours));
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_PUZZLE, e.what());
        }
        uint64_t n = 0;
        for (uint64_t w : emp::backtrack_capability(*pz, (int)order)) n += (uint64_t)__builtin_popcountll(w);
        *n_variables = n;
        return EMP_HIP_OK;
    });
}

int emp_hip_situation_profile(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint16_t *situation,
                              uint32_t *level, uint32_t *glutton, uint32_t *surface, uint32_t *row_index,
                              uint32_t *n_variables) {
    return guarded([&] {
        if (!context || !puzzle || !situation || !level || !glutton || !surface || !row_index)
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        std::unique_ptr<emp::Puzzle> pz;
        try {
            pz.reset(new emp::Puzzle(puzzle->width, puzzle->height, puzzle->colours));
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_PUZZLE, e.what());
        }
        emp::ProfileOutcome out;
        try {
            emp::situation_profile_on_device(context->stream, context->workspace, *pz, situation, out);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        *level = out.level;
        if (n_variables) *n_variables = out.n_variables;
        std::copy(out.glutton.begin(), out.glutton.end(), glutton);
        std::copy(out.surface.begin(), out.surface.end(), surface);
        std::copy(out.row_index.begin(), out.row_index.end(), row_index);
        return EMP_HIP_OK;
    });
}

int emp_hip_profile(emp_hip_context *context, const emp_hip_puzzle *puzzle, uint64_t max_situations, uint32_t *stats,
                    uint16_t *solutions, uint64_t solution_capacity, uint64_t *n_solutions, uint64_t *level_sizes,
                    uint16_t *starts, uint32_t *n_starts) {
    return guarded([&] {
        if (!context || !puzzle || !stats || !n_solutions || (solution_capacity && !solutions))
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        std::unique_ptr<emp::Puzzle> pz;
        try {
            pz.reset(new emp::Puzzle(puzzle->width, puzzle->height, puzzle->colours));
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_PUZZLE, e.what());
        }
        emp::SearchOutcome out;
        try {
            emp::profile_search_on_device(context->stream, context->workspace, *pz, max_situations, out);
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_ARGUMENT, e.what());
        }
        const uint32_t P = pz->size();
        std::copy(out.stats.begin(), out.stats.end(), stats);
        *n_solutions = out.n_solutions;
        const uint64_t kept = std::min<uint64_t>(out.n_solutions, solution_capacity);
        std::copy(out.solutions.begin(), out.solutions.begin() + kept * P, solutions);
        if (level_sizes) std::copy(out.level_sizes.begin(), out.level_sizes.end(), level_sizes);
        if (n_starts) *n_starts = (uint32_t)out.starts.size();
        if (starts)
            for (size_t i = 0; i < out.starts.size(); ++i) std::copy(out.starts[i].begin(), out.starts[i].end(), starts + i * P);
        return EMP_HIP_OK;
    });
}

int emp_hip_glutton_wide(emp_hip_context *context, const emp_hip_puzzle *puzzle, const uint32_t *parents,
                         uint32_t parent_len, uint32_t n_parents, const uint32_t *cells, uint64_t *parent_masks,
                         uint32_t *parent_status, uint64_t capacity, uint32_t *child_parent, uint32_t *child_move,
                         uint64_t *child_masks, uint32_t *child_status, uint64_t *n_children) {
    return guarded([&] {
        if (!context || !puzzle || !cells || !n_children || !n_parents || (parent_len && !parents))
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument or no parent");
        if (n_parents > (1u << 20)) return fail(EMP_HIP_BAD_ARGUMENT, "at most 2^20 parents");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        std::unique_ptr<emp::Puzzle> pz;
        try {
            pz.reset(new emp::Puzzle(puzzle->width, puzzle->height, puzzle->colours));
        } catch (const std::invalid_argument &e) {
            return fail(EMP_HIP_BAD_PUZZLE, e.what());
        }
        const uint32_t P = pz->size();
        for (uint32_t p = 0; p < n_parents; ++p)
            if (cells[p] >= P) return fail(EMP_HIP_BAD_ARGUMENT, "cell of parent " + std::to_string(p) + " out of range");
        std::vector<uint32_t> rows(parents, parents + (size_t)n_parents * parent_len);
        std::vector<uint32_t> cl(cells, cells + n_parents);
        emp::WideOutcome out;
        emp::glutton_wide_on_device(context->stream, context->workspace, *pz, rows, parent_len, n_parents, cl, out);
        const size_t per = (size_t)P * ((4 * P + 63) / 64);
        if (parent_masks) std::memcpy(parent_masks, out.parent_masks.data(), out.parent_masks.size() * 8);
        if (parent_status) std::memcpy(parent_status, out.parent_status.data(), n_parents * 4);
        const uint64_t n = out.child_parent.size(), m = std::min<uint64_t>(n, capacity);
        *n_children = n;
        if (child_parent && m) std::memcpy(child_parent, out.child_parent.data(), m * 4);
        if (child_move && m) std::memcpy(child_move, out.child_move.data(), m * 4);
        if (child_masks && m) std::memcpy(child_masks, out.child_masks.data(), m * per * 8);
        if (child_status && m) std::memcpy(child_status, out.child_status.data(), m * 4);
        return EMP_HIP_OK;
    });
}

int emp_hip_spiral_prefixes(const emp_hip_puzzle *puzzle, uint32_t length, uint32_t count, uint64_t seed,
                            const uint32_t *fixed_moves, uint32_t n_fixed, uint32_t *out, uint32_t *n_out) {
    return guarded([&] {
        if (!puzzle || !out || !n_out || (n_fixed && !fixed_moves))
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        emp::Puzzle pz(puzzle->width, puzzle->height, puzzle->colours);
        if (length == 0 || length >= pz.size()) return fail(EMP_HIP_BAD_ARGUMENT, "bad prefix length");
        std::vector<uint32_t> rows;
        uint32_t produced = 0;
        emp::spiral_prefixes(pz, length, count, seed,
                             std::vector<uint32_t>(fixed_moves, fixed_moves + n_fixed), rows, produced);
        std::copy(rows.begin(), rows.end(), out);
        *n_out = produced;
        return EMP_HIP_OK;
    });
}

}  // extern "C"

static_assert(EMP_HIP_MLP_PARAMS == emp::kMlpParams, "parameter count of the C ABI");
static_assert(EMP_HIP_MLP_INPUTS == emp::kMlpIn, "feature count of the C ABI");

int emp_hip_mlp_train(emp_hip_context *context, const float *features, const float *targets, uint64_t n_rows,
                      const emp_hip_mlp_options *options, float *params, float *loss_per_epoch) {
    return guarded([&]() -> int {
        if (!context || !features || !targets || !options || !params || !n_rows)
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument or no rows");
        if (options->batch == 0 || options->batch % 64 || options->batch > 65536)
            return fail(EMP_HIP_BAD_ARGUMENT, "batch must be a multiple of 64, at most 65536");
        if (n_rows > (1ull << 31)) return fail(EMP_HIP_BAD_ARGUMENT, "too many rows");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        emp::MlpTrainConfig cfg;
        cfg.epochs = options->epochs;
        cfg.batch = options->batch;
        cfg.lr = options->lr;
        cfg.seed = options->seed;
        std::vector<float> losses;
        emp::mlp_train_on_device(context->stream, context->workspace, features, targets, n_rows, cfg, params, losses);
        if (loss_per_epoch) std::copy(losses.begin(), losses.end(), loss_per_epoch);
        return EMP_HIP_OK;
    });
}

int emp_hip_mlp_score(emp_hip_context *context, const float *params, const float *features, uint64_t n_rows,
                      float *scores) {
    return guarded([&]() -> int {
        if (!context || !params || (n_rows && (!features || !scores))) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        emp::mlp_score_on_device(context->stream, context->workspace, params, features, n_rows, scores);
        return EMP_HIP_OK;
    });
}

int emp_hip_mlp_gradients(emp_hip_context *context, const float *params, const float *features,
                          const float *targets, uint64_t n_rows, float *grads, float *loss) {
    return guarded([&]() -> int {
        if (!context || !params || !features || !targets || !grads || !loss)
            return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (n_rows == 0 || n_rows > 65536) return fail(EMP_HIP_BAD_ARGUMENT, "rows must be in 1..65536");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        emp::mlp_gradients_on_device(context->stream, context->workspace, params, features, targets, n_rows, grads, *loss);
        return EMP_HIP_OK;
    });
}

namespace {

__global__ void stream_copy_kernel(const uint4 *src, uint4 *dst, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

namespace {

// LDS read stream: each lane reads 16 B per ds_read_b128, a wave a
// contiguous 1 KiB of a 32 KiB block buffer (no bank conflicts, constant
// offsets from a base that moves every iteration); the XOR of everything read
// decides one store, so no read can be dropped.
constexpr uint32_t kLdsWords = 2048;   // uint4 entries
__global__ void __launch_bounds__(256) lds_stream_kernel(uint32_t iters, uint32_t *sink) {
    __shared__ uint4 buf[kLdsWords];
    for (uint32_t i = threadIdx.x; i < kLdsWords; i += blockDim.x) buf[i] = make_uint4(i, 3 * i, 5 * i, 7 * i);
    __syncthreads();
    uint32_t a0 = 0, a1 = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        // index <= 1023 + 7 * 128 = 1919 < kLdsWords; the base moves every
        // iteration, so no read is loop-invariant
        const uint4 *base = buf + ((threadIdx.x + it * 64) & (kLdsWords / 2 - 1));
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r) {
            const uint4 v = base[r * 128];
            a0 ^= v.x ^ v.y;
            a1 ^= v.z ^ v.w;
        }
    }
    if ((a0 ^ a1) == 0x9E3779B9u) sink[0] = threadIdx.x;
}

// Integer VALU stream: 8 independent chains of x = popcount(x & m) + x
// (v_and_b32 and v_bcnt_u32_b32: 2 lane-ops per step).
__global__ void __launch_bounds__(256) valu_stream_kernel(uint32_t iters, uint32_t seed, uint32_t *sink) {
    const uint32_t m = seed * 0x9E3779B9u | 1u;
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = seed ^ (threadIdx.x * 0x85EBCA6Bu + (uint32_t)j);
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = (uint32_t)__popc(x[j] & m) + x[j];
    }
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) v ^= x[j];
    if (v == 0x9E3779B9u) sink[0] = threadIdx.x;
}

}  // namespace

int emp_hip_peaks(emp_hip_context *context, double *lds_gbs, double *valu_gops) {
    return guarded([&]() -> int {
        if (!context || !lds_gbs || !valu_gops) return fail(EMP_HIP_BAD_ARGUMENT, "null argument");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        uint32_t *sink = context->workspace.get<uint32_t>("peak_sink", 1);
        const uint32_t blocks = context->dev.compute_units * 8, threads = 256;
        const uint32_t lds_iters = 4096, valu_iters = 4096;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            return fail(EMP_HIP_DEVICE_ERROR, "hipEventCreate failed");
        float t_lds = 0, t_valu = 0;
        hipLaunchKernelGGL(lds_stream_kernel, dim3(blocks), dim3(threads), 0, context->stream, lds_iters, sink);
        (void)hipEventRecord(e0, context->stream);
        hipLaunchKernelGGL(lds_stream_kernel, dim3(blocks), dim3(threads), 0, context->stream, lds_iters, sink);
        (void)hipEventRecord(e1, context->stream);
        hipError_t err = hipStreamSynchronize(context->stream);
        (void)hipEventElapsedTime(&t_lds, e0, e1);
        if (err == hipSuccess) {
            hipLaunchKernelGGL(valu_stream_kernel, dim3(blocks), dim3(threads), 0, context->stream, valu_iters, 7u, sink);
            (void)hipEventRecord(e0, context->stream);
            hipLaunchKernelGGL(valu_stream_kernel, dim3(blocks), dim3(threads), 0, context->stream, valu_iters, 7u, sink);
            (void)hipEventRecord(e1, context->stream);
            err = hipStreamSynchronize(context->stream);
            (void)hipEventElapsedTime(&t_valu, e0, e1);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (err != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, hipGetErrorString(err));
        const double n = (double)blocks * threads;
        *lds_gbs = n * lds_iters * 8.0 * 16.0 / (t_lds * 1e6);
        *valu_gops = n * valu_iters * 4.0 * 8.0 * 2.0 / (t_valu * 1e6);
        return EMP_HIP_OK;
    });
}

int emp_hip_stream_copy(emp_hip_context *context, uint64_t bytes, double *ms) {
    return guarded([&]() -> int {
        if (!context || !bytes || bytes % 16) return fail(EMP_HIP_BAD_ARGUMENT, "bytes must be a positive multiple of 16");
        if (hipSetDevice(context->dev.device) != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, "hipSetDevice failed");
        uint4 *src = context->workspace.get<uint4>("copy_src", bytes / 16);
        uint4 *dst = context->workspace.get<uint4>("copy_dst", bytes / 16);
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            return fail(EMP_HIP_DEVICE_ERROR, "hipEventCreate failed");
        const uint32_t blocks = context->dev.compute_units * 8;
        hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, context->stream, src, dst, bytes / 16);
        (void)hipEventRecord(e0, context->stream);
        hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, context->stream, src, dst, bytes / 16);
        (void)hipEventRecord(e1, context->stream);
        const hipError_t err = hipStreamSynchronize(context->stream);
        float t = 0;
        (void)hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (err != hipSuccess) return fail(EMP_HIP_DEVICE_ERROR, hipGetErrorString(err));
        if (ms) *ms = t;
        return EMP_HIP_OK;
    });
}
