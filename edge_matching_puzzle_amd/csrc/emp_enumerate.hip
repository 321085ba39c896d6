// emp_enumerate.hip -- exhaustive enumeration on gfx950.
//
// Work decomposition.  A work item is a subtree of the reference's DFS
// (include/emp_strategy.h:332-415): the transitions placed at strategy indices
// 0..kmin-1 and, at index kmin, the candidates above `last` (last = -1: all).
// The root item is the first candidate of index 0 (the reference never
// revisits index 0, include/emp_strategy.h:389).
//
//   expand_kernel  one thread per item: emits one child item per candidate at
//                  index kmin (prefix split, the breadth step of the
//                  reference's glutton_wide play_from,
//                  include/CUDA_glutton_situations.h:713-757).
//   dfs_kernel     persistent: each lane of a wave64 owns one item and runs
//                  the DFS below it.  Idle lanes are refilled from the item
//                  queue with one atomic per wave (ballot / popcount / rank).
//                  A lane that exceeds the step budget splits what remains of
//                  its subtree into two items written to the next pass's
//                  queue, so heavy subtrees are shared out.
//
// Per-lane state lives in LDS (lane-interleaved so a wave's accesses fall in
// distinct banks): the stack of placed transitions (one byte per index for
// boards up to 10x10, two above) and the available transitions of each kind.
// The candidates of the current level are kept in registers; a level's
// candidates are recomputed from the colour masks of its placed neighbours
// only when the search returns to it, which keeps the stack at one or two
// bytes per level instead of one bitfield per level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "emp_enum_plan.h"
#include "emp_enumerate.h"
#include "emp_host_threads.h"

namespace emp {

namespace {

constexpr int kWave = 64;

struct DfsArgs {
    const uint64_t *tables;        // concatenated table blobs
    const uint32_t *table_off;     // [n_puzzles] u64 offset of each blob
    uint32_t lds_table_words;      // > 0: puzzle 0's blob is copied into LDS
    uint32_t avail_words;          // max avail words over puzzles (per lane)
    uint32_t stride;               // u16 per item stack (max P, even)
    uint32_t steps_per_round;
    uint64_t step_budget;          // steps before a lane splits its item
    uint64_t drain_budget;         // same, once the input queue is exhausted
    // input queue
    const uint64_t *in_hdr;        // puzzle | kmin << 32 | (last + 1) << 48
    const uint16_t *in_stack;      // [n_in][stride]
    uint32_t n_in;
    unsigned int *in_head;
    // Batches (items grouped by puzzle): the input is cut into kRegions
    // contiguous regions, region r taken first by the blocks with blockIdx
    // mod kRegions == r -- the blocks the dispatcher places on XCD r -- so
    // each XCD's L2 holds the tables of its own run of puzzles; a wave moves
    // on to the next region once its own is taken.  Null: one head (in_head).
    unsigned int *region_head;     // [kRegions] items taken from each region
    // output queue (residual items)
    uint64_t *out_hdr;
    uint16_t *out_stack;
    uint32_t out_cap;
    unsigned int *out_count;
    unsigned int *flags;           // [0] output overflow, [1] splits, [2] items, [3] drain,
                                   // [4] second halves granted (dfs splits), [5] identical
                                   // pieces out of order (never cleared per pass)
    // per puzzle
    unsigned long long *nodes;
    unsigned long long *solutions;
    unsigned long long *first_tick;
    unsigned long long *last_tick;
    const unsigned long long *node_limit;   // 0 = none
    unsigned int *truncated;
    // solution store: the s-th solution of puzzle p (in discovery order)
    // goes to slot p * sol_per_puzzle + s while s < sol_per_puzzle
    uint16_t *sol_stack;           // [n_puzzles * sol_per_puzzle][stride]
    uint32_t sol_per_puzzle;
    unsigned long long *candidate_sets;   // candidate sets computed (DFS iterations)
    unsigned int *max_depth;       // per puzzle: most pieces placed in one situation
    // visit order (ranked kernels): rank of transition t at index k of puzzle p
    // is ranks[rank_off[p] + k * rank_stride[p] + t]; lower ranks first
    const uint16_t *ranks;
    const uint64_t *rank_off;
    const uint32_t *rank_stride;
    uint64_t deadline;             // wall_clock64 tick after which searches stop (0 = none)
    unsigned int *timed_out;       // set when the deadline cut the search
    const unsigned int *n_in_dev;  // when set, the input item count is read here (graph replays)
    uint32_t keep_rest;            // past the deadline or the node limit, requeue the
                                   // unfinished work instead of dropping it (timed_out
                                   // then marks the cut)
    // Ordered results.  exec_nodes[exec_base + i]: the nodes counted while
    // input item i of this pass was searched; per stored solution (slot as in
    // sol_stack), the execution that found it and the nodes that execution
    // had counted when it placed the solution (the solution included).
    // Expansions then leave the placements they make to the executions that
    // follow them in the DFS (header field hdr_pending), so every node
    // belongs to exactly one execution.
    unsigned long long *exec_nodes;
    uint32_t exec_base;
    uint32_t *sol_exec;
    unsigned long long *sol_local;
    // Work accounting (instrumented runs): [0] algorithmic LDS bytes, [1]
    // integer ops of the placements, 8 M (n + 3) and 2 M (n + 1) + 2 M per
    // placement with M = ceil(4 P / 64) u64 words and n the placed neighbours
    // of the next strategy index (0 after the last one).
    unsigned long long *work;
    // One persistent launch (persist != 0) instead of passes.  Work items
    // live in a ring of 2^q_shift slots (in_hdr / in_stack, written through
    // q_hdr / q_stack), ticket t in slot t mod 2^q_shift: tickets below the
    // initial q_tail hold the starting items, later ones are filled by lanes
    // that hand over the rest of their floor level -- after step_budget steps
    // since their last hand-over, drain_budget while lanes wait for work
    // (q_head > q_tail), not while q_backlog items wait for lanes.  An idle
    // lane takes the next ticket (q_head).  Slot states (q_ready): 2g + 1 =
    // holds the item of generation g (t >> q_shift), 2g + 2 = that item was
    // read; a slot is filled for generation g only in state 2g (0 for g = 0):
    // a lane whose slot is still in use keeps the item owed and writes it in
    // a later round (its floor levels do not change meanwhile), so no lane
    // waits.  q_pending counts the items reserved, queued or being searched;
    // the launch ends when it reaches zero.
    uint32_t persist;              // kernel choice: the persistent form
    uint32_t limits;               // kernel choice: node limits checked while items run
    uint32_t q_shift;
    uint32_t q_backlog;
    unsigned int *q_head;
    unsigned int *q_tail;
    unsigned int *q_pending;
    unsigned int *q_ready;
    uint64_t *q_hdr;
    uint16_t *q_stack;
    unsigned int *q_stalled;       // set when a wave waited kStallTicks with nothing to do
};

// Waits until this wave's vector memory operations issued so far are
// complete (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their maxima).
// The work items of a persistent launch are written and read with
// agent-scope relaxed atomics, which go to the coherence point past the XCD's
// L2, so ordering them needs only this wait -- an agent-scope release or
// acquire would write back or invalidate the whole L2 of the XCD, and a
// workgroup-scope fence emits no wait for global memory.
__device__ __forceinline__ void wave_settle() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// A wave of a persistent launch that has waited this long (device clock,
// 100 MHz) for work with items still pending gives up and reports it: the
// search ends with an error instead of waiting without bound.
constexpr uint64_t kStallTicks = 2000000000ull;   // 20 s

// A persistent launch's lanes read the queue counters (and may hand over)
// every kQueueRounds rounds of steps_per_round steps.
constexpr uint32_t kQueueRounds = 8;

// Input regions of a batched pass: one per XCD (MI355X has eight, and the
// dispatcher deals workgroups to them round-robin).
constexpr uint32_t kRegions = 8;
// Items a wave of a batched pass takes from its region at a time.
#ifndef EMP_RUN_ITEMS
#define EMP_RUN_ITEMS 64
#endif
constexpr uint32_t kRunItems = EMP_RUN_ITEMS;

// Steps a work item runs before it splits once every input item of a pass is
// taken (and, in a persistent launch, between hand-overs while lanes wait).
#ifndef EMP_DRAIN_BUDGET
#define EMP_DRAIN_BUDGET 512
#endif
constexpr uint64_t kDrainBudget = EMP_DRAIN_BUDGET;

// Plain order: transitions above `last` in index order.  Ranked order: the
// transitions whose rank at this index is above the rank of `last`.
template <int W>
__device__ __forceinline__ void keep_above(uint64_t (&c)[W], int last) {
    const int from = last + 1;
    const int fw = from >> 6, fb = from & 63;
#pragma unroll
    for (int w = 0; w < W; ++w) c[w] &= w < fw ? 0ull : (w == fw ? ~0ull << fb : ~0ull);
}

template <int W>
__device__ __forceinline__ void keep_above_ranked(uint64_t (&c)[W], int last, const uint16_t *row) {
    if (last < 0) return;
    const uint32_t rl = row[last];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        uint64_t v = c[w];
        for (uint64_t m = v; m; m &= m - 1) {
            const uint32_t b = (uint32_t)__builtin_ctzll(m);
            if (row[w * 64 + b] <= rl) v &= ~(1ull << b);
        }
        c[w] = v;
    }
}

// Lowest-ranked transition of c, or -1.
template <int W>
__device__ __forceinline__ int first_ranked(const uint64_t (&c)[W], const uint16_t *row) {
    int best = -1;
    uint32_t br = 0x10000u;
#pragma unroll
    for (int w = 0; w < W; ++w)
        for (uint64_t m = c[w]; m; m &= m - 1) {
            const uint32_t t = w * 64 + (uint32_t)__builtin_ctzll(m);
            const uint32_t r = row[t];
            if (r < br) {
                br = r;
                best = (int)t;
            }
        }
    return best;
}

// Item header: puzzle (bits 0-19) | floor (20-30) | level (31-41) |
// last + 1 (42-53).  The item's stack row holds the transitions of indices
// 0..level-1; the search resumes at index `level` above transition `last`
// and ends when index `floor` has no candidate left.  A fresh subtree has
// level == floor.
__device__ __forceinline__ uint64_t make_hdr(uint32_t puzzle, uint32_t floor, uint32_t level, int last) {
    return (uint64_t)puzzle | ((uint64_t)floor << 20) | ((uint64_t)level << 31) |
           ((uint64_t)(uint32_t)(last + 1) << 42);
}
__device__ __forceinline__ uint32_t hdr_puzzle(uint64_t h) { return (uint32_t)h & 0xFFFFF; }
__device__ __forceinline__ uint32_t hdr_floor(uint64_t h) { return (uint32_t)(h >> 20) & 0x7FF; }
__device__ __forceinline__ uint32_t hdr_level(uint64_t h) { return (uint32_t)(h >> 31) & 0x7FF; }
__device__ __forceinline__ int hdr_last(uint64_t h) { return (int)((uint32_t)(h >> 42) & 0xFFF) - 1; }
constexpr uint32_t kDeadItem = 0x7FF;   // floor of a slot reserved but never written
// Bits 54-63 (ordered results only): placements not counted yet that come
// right before the item in the DFS -- its own placement (stack[level - 1])
// and, for the first child of an expansion, the parent's pending ones; the
// item's execution counts them.
__device__ __forceinline__ uint32_t hdr_pending(uint64_t h) { return (uint32_t)(h >> 54); }
constexpr uint64_t kPendingMask = 0x3FFull << 54;

// Read-only view of one table blob.  Per-kind fields are kept as three
// scalars and picked with selects, so a runtime kind never indexes a private
// array (which would spill it to scratch memory).  Every mask row of the blob
// has W words (the launch's row width; words past a kind's transitions are 0).
template <typename T>
__device__ __forceinline__ T sel3(int k, T a0, T a1, T a2) { return k == 0 ? a0 : (k == 1 ? a1 : a2); }

struct TableView {
    const uint64_t *b;
    uint32_t P, ncol;
    uint32_t n0, n1, n2, d0, d1, d2;
    uint32_t cm0, cm1, cm2, va0, va1, va2, du0, du1, du2, ro0, ro1, ro2;
    uint32_t off_base, off_nbr, off_ones, off_zero, cl0, cl1, cl2, off_col;
    uint32_t off_succ, off_pred, npred, stk_shift;

    __device__ __forceinline__ void load(const uint64_t *blob) {
        b = blob;
        const uint32_t *h = reinterpret_cast<const uint32_t *>(blob);
        // field order of EnumTableHeader
        P = h[0];
        ncol = h[3];
        n0 = h[4]; n1 = h[5]; n2 = h[6];
        d0 = h[7]; d1 = h[8]; d2 = h[9];
        cm0 = h[11]; cm1 = h[12]; cm2 = h[13];
        va0 = h[14]; va1 = h[15]; va2 = h[16];
        du0 = h[17]; du1 = h[18]; du2 = h[19];
        ro0 = h[20]; ro1 = h[21]; ro2 = h[22];
        off_base = h[27];
        off_nbr = h[28];
        off_ones = h[29];
        off_zero = h[29] + 16;
        cl0 = h[30]; cl1 = h[31]; cl2 = h[32];
        off_col = h[33];
        off_succ = h[34];
        off_pred = h[35];
        npred = h[36];
        stk_shift = h[37] == 2 ? 7u : 6u;
    }
    // One table for the whole grid: keep its fields in scalar registers.
    __device__ __forceinline__ void make_uniform() {
        uint32_t *f[] = {&P, &ncol, &n0, &n1, &n2, &d0, &d1, &d2, &cm0, &cm1, &cm2, &va0, &va1, &va2,
                         &du0, &du1, &du2, &ro0, &ro1, &ro2, &off_base, &off_nbr, &off_ones, &off_zero,
                         &cl0, &cl1, &cl2, &off_col, &off_succ, &off_pred, &npred, &stk_shift};
#pragma unroll
        for (uint32_t *x : f) *x = __builtin_amdgcn_readfirstlane(*x);
    }
    // natural words of kind K (per-lane available-transition layout)
    __device__ __forceinline__ uint32_t nat(int k) const { return sel3(k, n0, n1, n2); }
    __device__ __forceinline__ uint32_t abase(int k) const { return sel3(k, n2 + n1, n2, 0u); }
    __device__ __forceinline__ uint32_t ndup(int k) const { return sel3(k, d0, d1, d2); }
    __device__ __forceinline__ uint32_t off_cm(int k) const { return sel3(k, cm0, cm1, cm2); }
    __device__ __forceinline__ uint32_t off_valid(int k) const { return sel3(k, va0, va1, va2); }
    __device__ __forceinline__ uint32_t off_dup(int k) const { return sel3(k, du0, du1, du2); }
    // word 0 of position k's neighbour record: count | kind << 3
    __device__ __forceinline__ uint64_t nbr(uint32_t k) const { return b[off_nbr + k * kNbrWords]; }
    __device__ __forceinline__ const uint64_t *nbr_rec(uint32_t k) const { return b + off_nbr + k * kNbrWords; }
    __device__ __forceinline__ uint32_t colour(uint32_t byte_off) const {
        return reinterpret_cast<const uint8_t *>(b + off_col)[byte_off];
    }
    __device__ __forceinline__ uint32_t cls(int kind, uint32_t t) const {
        return reinterpret_cast<const uint16_t *>(b + sel3(kind, cl0, cl1, cl2))[t];
    }
    __device__ __forceinline__ uint32_t rot(int kind, uint32_t t) const {
        return reinterpret_cast<const uint32_t *>(b + sel3(kind, ro0, ro1, ro2))[t];
    }
    // identical-piece chain entry of a piece slot (abase * 16 + kind_index)
    __device__ __forceinline__ uint32_t succ(uint32_t slot) const {
        return reinterpret_cast<const uint16_t *>(b + off_succ)[slot];
    }
    __device__ __forceinline__ uint32_t pred_entry(uint32_t i) const {
        return reinterpret_cast<const uint32_t *>(b + off_pred)[i];
    }
};

static_assert(sizeof(EnumTableHeader) == 38 * 4, "TableView::load reads EnumTableHeader by field index");

__device__ __forceinline__ int desc_kind(uint64_t d) { return (int)(d >> 3) & 3; }
// offset of the position's kind in a lane's available-transition region
__device__ __forceinline__ uint32_t desc_abase(uint64_t d) { return (uint32_t)(d >> 8) & 0xFFFF; }
// u64 offset of the kind's class-union row in the table
__device__ __forceinline__ uint32_t desc_union(uint64_t d) { return (uint32_t)(d >> 32); }

// Lowest set bit of c, or -1: the first non-zero word is selected, then one
// count of trailing zeros.
template <int W>
__device__ __forceinline__ int first_bit(const uint64_t (&c)[W]) {
    uint64_t x = c[0];
    int base = 0;
#pragma unroll
    for (int w = 1; w < W; ++w) {
        const bool z = x == 0;
        x = z ? c[w] : x;
        base = z ? w * 64 : base;
    }
    return x ? base + (int)__builtin_ctzll(x) : -1;
}

template <int W>
__device__ __forceinline__ bool any_bit(const uint64_t (&c)[W]) {
    uint64_t v = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) v |= c[w];
    return v != 0;
}

// Candidates of strategy index k (record word 0 = d): its base mask (valid
// transitions matching the frame) AND the available transitions of its kind
// AND one colour mask per neighbour placed earlier; in each appearance class
// only the lowest piece present keeps its transitions; then only transitions
// above `last`.  The class reduction is applied to the full candidate set, as
// the reference's identical-piece masks accumulate from the first candidate
// on.  get_choice(stack_byte_offset): transition placed at the index whose
// level starts at that byte offset of a wave's stack (index * 64 lanes *
// stack_bytes);
// get_avail(d, w): available transition word w of the kind of the position
// whose record word is d (only words below the kind's natural count matter:
// the base mask is zero above).
//
// Without CLASSES the caller's available bits already hide every piece whose
// chain holds a lower available piece (see the chain entries of the table):
// identical pieces that are both available are candidates together or not at
// all, as their transitions show the same looks, so hiding the higher ones
// from the available bits removes exactly what the class reduction removes.
template <int W, bool CLASSES = true, typename GetChoice, typename GetAvail>
__device__ __forceinline__ void candidates(const TableView &tv, uint32_t k, uint64_t d, int last,
                                           GetChoice get_choice, GetAvail get_avail, uint64_t (&c)[W],
                                           const uint16_t *rank_row = nullptr) {
    const int K = desc_kind(d);
    const uint32_t n = (uint32_t)d & 7;
    const uint64_t *base = tv.b + tv.off_base + (size_t)k * W;
#pragma unroll
    for (int w = 0; w < W; ++w) c[w] = base[w] & get_avail(d, w);
    // Neighbour masks.  The loop runs to the wave's largest count; a lane
    // past its own count ANDs the all-ones row, so the body has no branch.
    const uint64_t *rec = tv.nbr_rec(k) + 1;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const bool on = i < n;
        if (__ballot(on) == 0) break;
        const uint64_t e = on ? rec[i] : 0ull;
        const uint32_t lo = (uint32_t)e, hi = (uint32_t)(e >> 32);
        const uint32_t col = tv.colour((lo >> 16) + get_choice(lo & 0xFFFF));
        // hi: byte offset of the colour-mask block in the table
        const uint64_t *row = on ? reinterpret_cast<const uint64_t *>(reinterpret_cast<const uint8_t *>(tv.b) + hi) + col * W
                                 : tv.b + tv.off_ones;
#pragma unroll
        for (int w = 0; w < W; ++w) c[w] &= row[w];
    }
    // Appearance classes present among the candidates, lowest first; most
    // candidate sets meet none, so the loop usually does not run.
    if (CLASSES) {
        const uint64_t *uni = tv.b + desc_union(d);
        uint64_t rem[W];
#pragma unroll
        for (int w = 0; w < W; ++w) rem[w] = c[w] & uni[w];
        while (__ballot(any_bit<W>(rem)) != 0) {
            const int b = first_bit<W>(rem);
            const uint64_t *dm = b >= 0 ? tv.b + tv.off_dup(K) + tv.cls(K, (uint32_t)b) * W : tv.b + tv.off_zero;
            bool seen = false;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint64_t m = c[w] & dm[w];
                // keep the transitions of the lowest piece of the class
                const uint64_t low = m ? (0xFull << (__builtin_ctzll(m) & 60)) : 0ull;
                c[w] &= ~(m & ~(seen ? 0ull : low));
                seen |= m != 0;
                rem[w] &= ~dm[w];
            }
        }
    }
    // Only transitions above `last` (last = -1 keeps all).
    if (rank_row) keep_above_ranked<W>(c, last, rank_row);
    else keep_above<W>(c, last);
}

// ---------------------------------------------------------------------------
// Prefix expansion: one thread per item.
// ---------------------------------------------------------------------------
template <int W, bool RANKED>
__global__ void expand_kernel(DfsArgs a) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= a.n_in) return;
    const uint64_t hdr = a.in_hdr[idx];
    const uint32_t kmin = hdr_floor(hdr);
    if (kmin == kDeadItem) return;
    const uint32_t puzzle = hdr_puzzle(hdr);
    const int last = hdr_last(hdr);
    TableView tv;
    tv.load(a.tables + a.table_off[puzzle]);
    const bool record = a.exec_nodes != nullptr;
    auto pass_through = [&]() {
        if (__hip_atomic_load(&a.flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
        const uint32_t o = atomicAdd(a.out_count, 1u);
        if (o >= a.out_cap) {
            atomicOr(&a.flags[0], 1u);
            return;
        }
        const uint32_t lv = hdr_level(hdr);
        for (uint32_t j = 0; j < lv; ++j) a.out_stack[(size_t)o * a.stride + j] = a.in_stack[(size_t)idx * a.stride + j];
        a.out_hdr[o] = hdr;
    };
    if (hdr_level(hdr) != kmin || (record && kmin + 1 >= tv.P)) {
        // A resumed search is passed through unchanged; so is, for ordered
        // results, an item whose children would be solutions (their
        // executions record them).
        pass_through();
        return;
    }
    const uint16_t *stack = a.in_stack + (size_t)idx * a.stride;
    uint64_t avail[3][W];
#pragma unroll
    for (int K = 0; K < 3; ++K)
#pragma unroll
        for (int w = 0; w < W; ++w) avail[K][w] = tv.b[tv.off_valid(K) + w];
    for (uint32_t j = 0; j < kmin; ++j) {
        const int K = desc_kind(tv.nbr(j));
        const uint32_t bb = (uint32_t)stack[j] & ~3u;
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int w = 0; w < W; ++w)
                if (q == K && w == (int)(bb >> 6)) avail[q][w] &= ~(0xFull << (bb & 63));
    }
    uint64_t c[W];
    const uint16_t *rank_row = RANKED ? a.ranks + a.rank_off[puzzle] + (size_t)kmin * a.rank_stride[puzzle] : nullptr;
    candidates<W>(tv, kmin, tv.nbr(kmin), last, [&](uint32_t off) { return (uint32_t)stack[off >> tv.stk_shift]; },
                  [&](uint64_t d, int w) {
                      const int kk = desc_kind(d);
                      uint64_t v = 0;
#pragma unroll
                      for (int q = 0; q < 3; ++q)
#pragma unroll
                          for (int x = 0; x < W; ++x)
                              if (q == kk && x == w) v = avail[q][x];
                      return v;
                  },
                  c, rank_row);
    uint32_t n = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) n += (uint32_t)__popcll(c[w]);
    if (!n) {
        // nothing below: an item still holding uncounted placements is
        // passed on, so that an execution counts them in their place
        if (record && hdr_pending(hdr)) pass_through();
        return;
    }
    if (!record) atomicAdd(&a.nodes[puzzle], (unsigned long long)n);
    if (a.work && !record) {
        const uint64_t M = (4ull * tv.P + 63) / 64;
        const uint64_t nn = kmin + 1 < tv.P ? (tv.nbr(kmin + 1) & 7) : 0;
        atomicAdd(&a.work[0], (unsigned long long)(n * 8 * M * (nn + 3)));
        atomicAdd(&a.work[1], (unsigned long long)(n * 2 * M * (nn + 2)));
    }
    atomicMax(&a.max_depth[puzzle], kmin + 1);
    if (kmin == tv.P - 1) {
        uint64_t s = atomicAdd(&a.solutions[puzzle], (unsigned long long)n);
        for (int w = 0; w < W; ++w) {
            uint64_t v = c[w];
            while (v) {
                const uint32_t t = w * 64 + __builtin_ctzll(v);
                v &= v - 1;
                if (s < a.sol_per_puzzle) {
                    uint16_t *dst = a.sol_stack + ((size_t)puzzle * a.sol_per_puzzle + s) * a.stride;
                    for (uint32_t j = 0; j < kmin; ++j) dst[j] = stack[j];
                    dst[kmin] = (uint16_t)t;
                }
                ++s;
            }
        }
        return;
    }
    if (__hip_atomic_load(&a.flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const uint32_t base = atomicAdd(a.out_count, n);
    if ((uint64_t)base + n > a.out_cap) {
        atomicOr(&a.flags[0], 1u);
        return;
    }
    uint32_t o = base;
    // ordered results: each child holds its own placement, and the first
    // (lowest transition, first in the DFS) the parent's pending ones too
    uint64_t pending = record ? (uint64_t)(hdr_pending(hdr) + 1) << 54 : 0ull;
    for (int w = 0; w < W; ++w) {
        uint64_t v = c[w];
        while (v) {
            const uint32_t t = w * 64 + __builtin_ctzll(v);
            v &= v - 1;
            uint16_t *dst = a.out_stack + (size_t)o * a.stride;
            for (uint32_t j = 0; j < kmin; ++j) dst[j] = stack[j];
            dst[kmin] = (uint16_t)t;
            a.out_hdr[o] = make_hdr(puzzle, kmin + 1, kmin + 1, -1) | pending;
            pending = record ? 1ull << 54 : 0ull;
            ++o;
        }
    }
}

// ---------------------------------------------------------------------------
// Persistent DFS.
//
// A lane's state is the level k it is at and, in registers, the candidates of
// level k not tried yet (c, all above `last`); levels below k are placed (the
// transition stack and the available transitions in LDS).  Each iteration
// computes one candidate set:
//   * c empty: level k is exhausted; pop to k - 1 and recompute its
//     candidates above the transition it held;
//   * else place the first candidate t of c at level k and compute the
//     candidates of level k + 1 with t's piece taken; descend only when that
//     set is not empty -- otherwise the lane stays at level k with c intact,
//     so a dead-end placement costs one candidate set, not two.
// ---------------------------------------------------------------------------
// Blocks of up to 16 waves for narrow rows; wide rows need more registers,
// so their blocks stay at 4 waves.
constexpr uint32_t dfs_max_threads(int w) { return w <= 4 ? 1024u : 256u; }

// LIMITS: the launch has node limits, checked while items run as well as
// when they are taken (compiled only where needed: the check's live state
// costs registers in the step loop).
// A wave's stacks of placed transitions in LDS, lane-interleaved (level j of
// lane l at j * 64 + l): one byte per level while every kind has at most 256
// transitions (rows of up to four words); wider rows (up to 1,024
// transitions) keep the low byte there and the two high bits of four levels
// per byte in a second plane -- 1.25 bytes per level instead of 2, which is
// what fits five waves per CU at 16x16 instead of three.
template <bool SPLIT>
struct LaneStack {
    uint8_t *lo;     // [levels][64]
    uint8_t *hi;     // [levels / 4][64] (SPLIT)
    uint32_t lane;
    __device__ __forceinline__ uint32_t get(uint32_t j) const {
        uint32_t v = lo[j * kWave + lane];
        if (SPLIT) v |= ((uint32_t)(hi[(j >> 2) * kWave + lane] >> ((j & 3) * 2)) & 3u) << 8;
        return v;
    }
    __device__ __forceinline__ void set(uint32_t j, uint32_t v) const {
        lo[j * kWave + lane] = (uint8_t)v;
        if (SPLIT) {
            uint8_t &h = hi[(j >> 2) * kWave + lane];
            const uint32_t sh = (j & 3) * 2;
            h = (uint8_t)((h & ~(3u << sh)) | (((v >> 8) & 3u) << sh));
        }
    }
    // the level a neighbour record addresses by its byte-plane offset (level * 64)
    __device__ __forceinline__ uint32_t at(uint32_t off) const {
        uint32_t v = lo[off + lane];
        if (SPLIT) v |= ((uint32_t)(hi[((off >> 8) << 6) + lane] >> ((off >> 5) & 6)) & 3u) << 8;
        return v;
    }
};

// S(j) of the kernel: reads and assignments of level j of this lane's stack.
template <bool SPLIT>
struct LevelRef {
    const LaneStack<SPLIT> &s;
    uint32_t j;
    __device__ __forceinline__ operator uint32_t() const { return s.get(j); }
    __device__ __forceinline__ LevelRef &operator=(uint32_t v) {
        s.set(j, v);
        return *this;
    }
};

template <int W, bool LDS_TABLE, bool RANKED, bool PERSIST, bool LIMITS>
__global__ void __launch_bounds__(dfs_max_threads(W)) dfs_kernel(DfsArgs a) {
    constexpr bool SPLIT = W > 4;
    using STK = typename std::conditional<(W <= 4), uint8_t, uint16_t>::type;
    extern __shared__ uint64_t smem[];
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wave = (int)(threadIdx.x / kWave);

    if (LDS_TABLE) {
        const uint64_t *src = a.tables + a.table_off[0];
        for (uint32_t i = threadIdx.x; i < a.lds_table_words; i += blockDim.x) smem[i] = src[i];
        __syncthreads();
    }
    // Per wave: available-transition words then the stack of transitions,
    // u8 when every kind has at most 256 transitions (boards up to 10x10).
    const uint32_t per_wave = a.avail_words * kWave + (a.stride * kWave + (SPLIT ? a.stride / 4 * kWave : 0u)) / 8;
    uint64_t *avail = smem + a.lds_table_words + (size_t)wave * per_wave;
    uint8_t *stk_lo = reinterpret_cast<uint8_t *>(avail + a.avail_words * kWave);
    const LaneStack<SPLIT> stack{stk_lo, stk_lo + a.stride * kWave, (uint32_t)lane};

    TableView tv;
    if (LDS_TABLE) {
        tv.load(smem);
        tv.make_uniform();
    }
    const uint32_t n_in = a.n_in_dev ? *a.n_in_dev : a.n_in;

    bool active = false;
    bool queue_empty = false;
    uint32_t region = blockIdx.x % kRegions;   // batched passes: the input region taken from
    uint32_t regions_left = kRegions;          // regions not yet seen taken
    uint32_t run_next = 0, run_end = 0;        // the wave's run of items: [run_next, run_end) not taken
    bool waiting = false;                    // persistent launch: holds a ticket not started yet
    uint32_t ticket = 0;
    bool owed = false;                       // persistent launch: a reserved ticket whose slot was in use
    uint32_t owed_ticket = 0, owed_floor = 0;
    uint64_t give_mark = 0;                  // steps of the item at its last hand-over
    uint32_t rounds = 0;                     // rounds of this lane (queue checks every kQueueRounds)
    uint64_t idle_since = 0;                 // persistent launch: the wave's wait for work began
    bool fresh = false;       // c not computed yet for the item's level
    uint32_t puzzle = 0, k = 0, kmin = 0, P = 0, depth = 0;
    uint32_t item = 0;   // input index of the item (ordered results)
    int last = -1;
    uint32_t cur_ab = 0;   // available-region offset of level k's kind
    const uint16_t *rank_base = nullptr;   // ranked: this puzzle's rank rows
    uint32_t rank_stride = 0;
    uint64_t steps = 0, nodes = 0;
    // Under a node limit the lane adds its progress to the puzzle's count
    // every `pub_step` nodes (0: the puzzle has no limit), so the many items
    // of one puzzle that run side by side (grouped batches, a persistent
    // launch) stop close to the limit instead of each running its whole
    // budget past it.  pub_at: the low 32 bits of `nodes` when last added
    // (0: nothing added yet; fewer than 2^32 nodes pass between two additions).
    uint32_t pub_at = 0, pub_step = 0;
    uint64_t work_bytes = 0, work_ops = 0;   // instrumented runs only
    uint32_t words_alg = 0;                  // M = ceil(4 P / 64) of the item's puzzle
    uint64_t c[W];
#pragma unroll
    for (int w = 0; w < W; ++w) c[w] = 0;

    // Available word w of kind K.  Words past the kind's own read the next
    // kind's words or, past the interior kind's (the last), the rest of the
    // region (sized for W words from the interior kind's offset); the base
    // mask zeroes them.
    auto A = [&](uint32_t ab, uint32_t w) -> uint64_t & { return avail[(ab + w) * kWave + lane]; };
    auto S = [&](uint32_t j) { return LevelRef<SPLIT>{stack, j}; };
    auto finish = [&]() {
        if (a.exec_nodes) a.exec_nodes[a.exec_base + item] = nodes;
        if (a.work) {
            atomicAdd(&a.work[0], (unsigned long long)work_bytes);
            atomicAdd(&a.work[1], (unsigned long long)work_ops);
        }
        // the nodes not added yet (under LIMITS fewer than 2^32 pass between
        // two additions)
        atomicAdd(&a.nodes[puzzle], LIMITS ? (unsigned long long)((uint32_t)nodes - pub_at) : (unsigned long long)nodes);
        atomicAdd(a.candidate_sets, (unsigned long long)steps);
        atomicMax(&a.max_depth[puzzle], depth);
        atomicMax(&a.last_tick[puzzle], (unsigned long long)wall_clock64());
        // the item is done: after its hand-overs, so the count never drops
        // to zero while work can still appear
        if (PERSIST) atomicSub(a.q_pending, 1u);
        active = false;
    };

    // Transition j of input item idx.  In a persistent launch the item may
    // have been written by another CU during the launch, so it is read from
    // L2 (rows start on 8-byte boundaries: stride is a multiple of 4).
    auto item_move = [&](uint32_t idx, uint32_t j) -> uint32_t {
        const uint16_t *src = a.in_stack + (size_t)idx * a.stride;
        if (!PERSIST) return src[j];
        const uint32_t w = __hip_atomic_load(const_cast<uint32_t *>(reinterpret_cast<const uint32_t *>(src)) + (j >> 1),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (w >> (16 * (j & 1))) & 0xFFFFu;
    };
    // Writes the rest of floor level f of this lane's search (the candidates
    // at index f above S(f), with S(0..f-1) placed) as the item of ticket tk,
    // if its slot is free for tk's generation; false when the slot is still
    // in use.  The item goes through the coherence point (stores that skip
    // the XCD's L2) and is settled before its state says so.
    auto write_handover = [&](uint32_t tk, uint32_t f) -> bool {
        const uint32_t slot = tk & ((1u << a.q_shift) - 1), g = tk >> a.q_shift;
        if (__hip_atomic_load(&a.q_ready[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != (g ? 2 * g : 0u))
            return false;
        uint32_t *db = reinterpret_cast<uint32_t *>(a.q_stack + (size_t)slot * a.stride);
        for (uint32_t q = 0; q < f; q += 2)
            __hip_atomic_store(db + (q >> 1), (uint32_t)S(q) | (q + 1 < f ? (uint32_t)S(q + 1) << 16 : 0u),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.q_hdr + slot, make_hdr(puzzle, f, f, (int)S(f)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        wave_settle();
        __hip_atomic_store(&a.q_ready[slot], 2 * g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    };
    // Starts input item idx on this lane; false when it is not searched here
    // (a dead slot, its puzzle over the node limit, past the deadline).
    auto take_item = [&](uint32_t idx) -> bool {
                    const uint64_t hdr = PERSIST ? __hip_atomic_load(const_cast<uint64_t *>(a.in_hdr) + idx,
                                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                                   : a.in_hdr[idx];
                    kmin = hdr_floor(hdr);
                    puzzle = hdr_puzzle(hdr);
                    bool take = kmin != kDeadItem;
                    const bool over_limit =
                        take && a.node_limit[puzzle] && a.nodes[puzzle] >= a.node_limit[puzzle];
                    if (over_limit && !a.keep_rest) {
                        atomicOr(&a.truncated[puzzle], 1u);
                        take = false;
                    }
                    if (take && (over_limit || (a.deadline && wall_clock64() > a.deadline))) {
                        atomicOr(&a.truncated[puzzle], 1u);
                        atomicOr(a.timed_out, 1u);
                        if (a.keep_rest) {
                            // the item is kept as it is for a later call (an
                            // uncounted placements are counted now)
                            if (hdr_pending(hdr)) atomicAdd(&a.nodes[puzzle], (unsigned long long)hdr_pending(hdr));
                            const uint32_t o = atomicAdd(a.out_count, 1u);
                            if (o < a.out_cap) {
                                const uint32_t lv = hdr_level(hdr);
                                for (uint32_t j = 0; j < lv; ++j)
                                    a.out_stack[(size_t)o * a.stride + j] = (uint16_t)item_move(idx, j);
                                a.out_hdr[o] = hdr & ~kPendingMask;
                            } else {
                                atomicOr(&a.flags[0], 1u);
                            }
                        }
                        take = false;
                    }
                    if (take) {
                        if (!LDS_TABLE) tv.load(a.tables + a.table_off[puzzle]);
                        P = tv.P;
                        last = hdr_last(hdr);
                        k = hdr_level(hdr);
                        steps = 0;
                        give_mark = 0;
                        nodes = hdr_pending(hdr);
                        {
                            const unsigned long long lim = a.node_limit[puzzle];
                            pub_step = lim ? (uint32_t)min(4096ull, max(64ull, lim >> 9)) : 0u;
                            pub_at = 0;
                        }
                        item = idx;
                        work_bytes = work_ops = 0;
                        words_alg = (4 * P + 63) / 64;
                        depth = k;
                        fresh = true;
                        if (RANKED) {
                            rank_base = a.ranks + a.rank_off[puzzle];
                            rank_stride = a.rank_stride[puzzle];
                        }
                        for (int K = 0; K < 3; ++K)
                            for (uint32_t w = 0, nw = tv.nat(K); w < nw; ++w)
                                avail[(tv.abase(K) + w) * kWave + lane] = tv.b[tv.off_valid(K) + w];
                        for (uint32_t j = 0; j < k; ++j) {
                            const uint32_t t = item_move(idx, j);
                            S(j) = (STK)t;
                            const uint32_t bb = t & ~3u;
                            A(desc_abase(tv.nbr(j)), bb >> 6) &= ~(0xFull << (bb & 63));
                        }
                        // Chained pieces behind an available lower piece are
                        // hidden (highest slot first, so each test reads its
                        // lower piece before that piece is itself hidden).  A
                        // placed piece whose lower piece is still available
                        // was never a candidate: the item is reported.
                        for (uint32_t i = 0; i < tv.npred; ++i) {
                            const uint32_t e = tv.pred_entry(i);
                            const uint32_t r = e & 0xFFFF, q = e >> 16;
                            uint64_t &rv = avail[(r >> 4) * kWave + lane];
                            const uint64_t rm = 0xFull << ((r & 15) * 4);
                            const bool lower = ((avail[(q >> 4) * kWave + lane] >> ((q & 15) * 4)) & 0xF) != 0;
                            if (lower && !(rv & rm)) atomicOr(&a.flags[5], 1u);
                            rv &= lower ? ~rm : ~0ull;
                        }
                        atomicMin(&a.first_tick[puzzle], (unsigned long long)wall_clock64());
                        atomicAdd(&a.flags[2], 1u);
                        active = true;
                    }
                    return take;
    };

    while (true) {
        if (PERSIST) {
            // Idle lanes take tickets, one atomic per wave; a lane holding a
            // ticket starts once its item is written.
            if (owed && !active && write_handover(owed_ticket, owed_floor)) owed = false;
            const uint64_t idle = __ballot(!active && !waiting && !owed);
            if (idle) {
                const uint32_t n_idle = (uint32_t)__popcll(idle);
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(a.q_head, n_idle);
                base = __shfl(base, leader);
                if (!active && !waiting && !owed) {
                    ticket = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1));
                    waiting = true;
                }
            }
            if (waiting) {
                const uint32_t slot = ticket & ((1u << a.q_shift) - 1), gen = ticket >> a.q_shift;
                if (__hip_atomic_load(&a.q_ready[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 2 * gen + 1) {
                    __atomic_signal_fence(__ATOMIC_ACQUIRE);   // the item's reads stay after the state's
                    waiting = false;
                    const bool took = take_item(slot);
                    // the slot is free for the next generation once read (every
                    // value read from it is in use by now, so the reads are done)
                    __hip_atomic_store(&a.q_ready[slot], 2 * gen + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (!took) atomicSub(a.q_pending, 1u);
                }
            }
            if (!__ballot(active)) {
                // Nothing to search in this wave: the launch is over once no
                // item is queued or being searched anywhere (every later ticket
                // would stay empty), else wait for a hand-over.
                uint32_t pend = 0;
                if (lane == 0) pend = __hip_atomic_load(a.q_pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__shfl(pend, 0) == 0) break;
                const uint64_t now = wall_clock64();
                if (!idle_since) idle_since = now;
                if (now - idle_since > kStallTicks) {
                    if (lane == 0) atomicOr(a.q_stalled, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
                continue;
            }
            idle_since = 0;
            if (!active) continue;
        } else {
        // Refill idle lanes with one queue atomic per wave.
        const uint64_t idle = __ballot(!active);
        if (idle && !queue_empty) {
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            const int leader = __ffsll((unsigned long long)idle) - 1;
            uint32_t base = 0, end = n_in;
            bool taken_all = false;
            if (a.region_head) {
                // The wave takes a run of kRunItems consecutive items of its
                // region and refills its lanes from that run only, so its
                // lanes hold items of one or two puzzles and a CU's waves
                // read few tables.
                while (run_next >= run_end && regions_left) {
                    const uint32_t lo = (uint32_t)((uint64_t)n_in * region / kRegions);
                    const uint32_t hi = (uint32_t)((uint64_t)n_in * (region + 1) / kRegions);
                    uint32_t b = 0;
                    if (lane == leader) b = lo + atomicAdd(&a.region_head[region], kRunItems);
                    b = __shfl(b, leader);
                    if (b < hi) {
                        run_next = b;
                        run_end = min(b + kRunItems, hi);
                    } else {
                        region = region + 1 == kRegions ? 0u : region + 1;
                        --regions_left;
                    }
                }
                base = run_next;
                end = run_end;
                run_next = min(run_next + n_idle, run_end);
                taken_all = run_next >= run_end && !regions_left;
            } else {
                if (lane == leader) base = atomicAdd(a.in_head, n_idle);
                base = __shfl(base, leader);
                taken_all = base + n_idle >= n_in;
            }
            if (taken_all) {
                queue_empty = true;
                // Every item is taken: from now on busy lanes split early so
                // the pass ends soon and the rest is shared out next pass.
                if (lane == leader) atomicOr(&a.flags[3], 1u);
            }
            if (!active) {
                const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1));
                const uint32_t idx = base + rank;
                if (idx < end) (void)take_item(idx);
            }
        }
        if (!__ballot(active)) {
            if (queue_empty) break;
            continue;
        }
        if (!active) continue;
        }

        bool over = false;
        if (LIMITS && pub_step && (uint32_t)nodes - pub_at >= pub_step) {
            const uint32_t add = (uint32_t)nodes - pub_at;
            const unsigned long long before = atomicAdd(&a.nodes[puzzle], (unsigned long long)add);
            over = before + add >= a.node_limit[puzzle];
            pub_at = (uint32_t)nodes;
        }
        if (over || (a.deadline && wall_clock64() > a.deadline)) {
            // out of time, or the puzzle's node limit reached: the rest of
            // this item is not searched here; with keep_rest it is requeued
            // as one item (level k above `last`, which is what the candidates
            // held in c stand for)
            atomicOr(&a.truncated[puzzle], 1u);
            if (!over || a.keep_rest) atomicOr(a.timed_out, 1u);
            if (a.keep_rest) {
                const uint32_t o = atomicAdd(a.out_count, 1u);
                if (o < a.out_cap) {
                    uint16_t *db = a.out_stack + (size_t)o * a.stride;
                    for (uint32_t q = 0; q < k; ++q) db[q] = S(q);
                    a.out_hdr[o] = make_hdr(puzzle, kmin, k, last);
                } else {
                    atomicOr(&a.flags[0], 1u);
                }
            }
            finish();
            continue;
        }
        if (PERSIST) {
            // Hand over the rest of the floor level (its candidates above the
            // transition this lane holds there) as a new item and go on one
            // level deeper, after step_budget steps since the last hand-over
            // -- drain_budget while lanes wait for work -- unless the queue
            // already holds q_backlog items nobody has taken.  The queue's
            // counters are read every kQueueRounds rounds: each read is a
            // round trip to the coherence point that the whole wave waits
            // for (C1: 106.2 G nodes/s reading them every round, 111.2 every
            // 4th, 112.8 every 8th).
            if (owed && write_handover(owed_ticket, owed_floor)) owed = false;
            const int first = __builtin_ctzll(__ballot(true));
            // the first active lane's count decides, so the wave agrees
            const bool check = ((uint32_t)__shfl((int)++rounds, first) & (kQueueRounds - 1)) == 0;
            uint32_t hq = 0, qq = 0;
            if (check && lane == first) {
                hq = __hip_atomic_load(a.q_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                qq = __hip_atomic_load(a.q_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            hq = __shfl(hq, first);
            qq = __shfl(qq, first);
            const bool hungry = hq > qq;
            const bool room = check && (hungry || qq - hq < a.q_backlog);
            const bool give = room && !owed && k > kmin &&
                              steps - give_mark >= (hungry ? a.drain_budget : a.step_budget);
            const uint64_t gm = __ballot(give);
            if (gm) {
                const uint32_t ng = (uint32_t)__popcll(gm);
                const int leader = __ffsll((unsigned long long)gm) - 1;
                uint32_t base = 0;
                if (lane == leader) {
                    // counted (and the count settled) before any of them can
                    // be taken and finished
                    atomicAdd(a.q_pending, ng);
                    wave_settle();
                    base = atomicAdd(a.q_tail, ng);
                    atomicAdd(&a.flags[1], ng);
                }
                base = __shfl(base, leader);
                if (give) {
                    const uint32_t tk = base + (uint32_t)__popcll(gm & ((1ull << lane) - 1));
                    if (!write_handover(tk, kmin)) {
                        owed = true;
                        owed_ticket = tk;
                        owed_floor = kmin;
                    }
                    kmin += 1;
                    give_mark = steps;
                }
            }
        }
        const bool drain =
            !PERSIST && __hip_atomic_load(&a.flags[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        const uint64_t budget = PERSIST ? ~0ull : (drain ? a.drain_budget : a.step_budget);
        for (uint32_t r = 0; r < a.steps_per_round; ++r) {
            // One candidate set per iteration, of level tk above tl:
            //   fresh item   tk = k,     tl = last
            //   pop          tk = k - 1, tl = transition held there
            //   place t      tk = k + 1, tl = -1, t's piece taken
            //   solution t   tk = k,     tl = t (the rest of level k)
            // Updates are written unconditionally where the value is inert
            // for the lanes that do not need them, so the body has no
            // divergent branch except the lane that finishes its item.
            const int t = fresh ? -1
                          : (RANKED ? first_ranked<W>(c, rank_base + (size_t)k * rank_stride) : first_bit<W>(c));
            const bool pop = !fresh && t < 0;
            if (pop && k == kmin) {
                finish();
                break;
            }
            const bool plc = t >= 0;
            const bool sol = plc && k == P - 1;
            const bool down = plc && !sol;
            const uint32_t km1 = k - (k > 0);
            const uint32_t held = S(km1);
            const uint64_t dprev = tv.nbr(km1);
            // S(k) is dead for every lane that does not place.
            S(k) = (STK)t;
            // The piece that changes hands, by piece slot: a pop gives level
            // k - 1's piece back and hides the next piece of its chain again;
            // a placement takes t's piece (before the next level's candidates
            // read the available bits) and shows the next piece of its chain.
            // Other lanes apply empty masks to slot 0.
            const bool move = pop || down;
            const uint32_t slot = pop ? desc_abase(dprev) * 16 + (held >> 2)
                                      : (down ? cur_ab * 16 + ((uint32_t)t >> 2) : 0u);
            const uint64_t pm = move ? (0xFull << ((slot & 15) * 4)) : 0ull;
            const uint32_t se = move ? tv.succ(slot) : 0u;
            const uint32_t ss = se & 0xFFF;
            const uint64_t sm = (uint64_t)(se >> 12) << ((ss & 15) * 4);
            {
                // masked merges (bit-field inserts): the piece's bits become
                // `pop`, the next piece's bits become `!pop`
                const uint64_t to_piece = pop ? ~0ull : 0ull, to_next = pop ? 0ull : ~0ull;
                uint64_t &pv = avail[(slot >> 4) * kWave + lane];
                pv ^= (pv ^ to_piece) & pm;
                uint64_t &sv = avail[(ss >> 4) * kWave + lane];
                sv ^= (sv ^ to_next) & sm;
            }
            if (sol) {
                const uint64_t sidx = atomicAdd(&a.solutions[puzzle], 1ull);
                if (sidx < a.sol_per_puzzle) {
                    const size_t slot_s = (size_t)puzzle * a.sol_per_puzzle + sidx;
                    uint16_t *dst = a.sol_stack + slot_s * a.stride;
                    for (uint32_t j = 0; j < P; ++j) dst[j] = S(j);
                    if (a.sol_exec) {
                        a.sol_exec[slot_s] = a.exec_base + item;
                        a.sol_local[slot_s] = nodes + 1;
                    }
                }
            }
            nodes += plc;
            depth = plc && k + 1 > depth ? k + 1 : depth;
            ++steps;
            const uint32_t tk = pop ? km1 : (down ? k + 1 : k);
            const int tl = pop ? (int)held : (down ? -1 : (sol ? t : last));
            fresh = false;
            const uint64_t d = tv.nbr(tk);
            if (a.work) {
                // a placement at k reads the masks of index k + 1 (none after the last)
                const uint32_t nn = sol ? 0u : (uint32_t)d & 7;
                work_bytes += plc ? 8ull * words_alg * (nn + 3) : 0ull;
                work_ops += plc ? 2ull * words_alg * (nn + 2) : 0ull;
            }
            uint64_t cc[W];
            candidates<W, false>(tv, tk, d, tl,
                          [&](uint32_t off) { return stack.at(off); },
                          [&](uint64_t dd, int w) { return A(desc_abase(dd), (uint32_t)w); },
                          cc, RANKED ? rank_base + (size_t)tk * rank_stride : nullptr);
            // A placement whose next level has no candidate is a dead end:
            // the lane stays at level k with the rest of c.
            const bool stay = down && !any_bit<W>(cc);
            const uint64_t tbit = plc ? (1ull << (t & 63)) : 0ull;
#pragma unroll
            for (int w = 0; w < W; ++w)
                c[w] = stay ? (c[w] & ~((w == (t >> 6)) ? tbit : 0ull)) : cc[w];
            // a dead end gives the piece back and hides its chain's next
            // piece again (no-ops otherwise)
            {
                uint64_t &pv = avail[(slot >> 4) * kWave + lane];
                pv |= stay ? pm : 0ull;
                uint64_t &sv = avail[(ss >> 4) * kWave + lane];
                sv &= stay ? ~sm : ~0ull;
            }
            k = stay ? k : tk;
            last = stay ? t : tl;
            cur_ab = stay ? cur_ab : desc_abase(d);
            if (steps >= budget) {
                // Requeue what remains: (b) the search in progress at level k
                // above `last`, and, when k is above the floor and room is
                // granted, split off (a) the candidates above the current
                // choice at the floor index -- usually the bulk of what
                // remains -- with (b)'s floor one index deeper.  Every input
                // item requeues at most once per pass, so the output always
                // has room for one item per input item; only the second
                // halves are counted against the rest (out_cap - n_in).
                uint32_t n = 1;
                if (k > kmin && atomicAdd(&a.flags[4], 1u) < a.out_cap - n_in) n = 2;
                const uint32_t base = atomicAdd(a.out_count, n);
                if ((uint64_t)base + n > a.out_cap) {
                    // cannot happen while n_in <= out_cap: reported as an error
                    atomicOr(&a.flags[0], 1u);
                    finish();
                    break;
                }
                uint16_t *db = a.out_stack + (size_t)base * a.stride;
                for (uint32_t q = 0; q < k; ++q) db[q] = S(q);
                a.out_hdr[base] = make_hdr(puzzle, n == 2 ? kmin + 1 : kmin, k, last);
                if (n == 2) {
                    uint16_t *da = a.out_stack + (size_t)(base + 1) * a.stride;
                    for (uint32_t q = 0; q < kmin; ++q) da[q] = S(q);
                    a.out_hdr[base + 1] = make_hdr(puzzle, kmin, kmin, (int)S(kmin));
                }
                atomicAdd(&a.flags[1], 1u);
                finish();
                break;
            }
        }
    }
}

#define EMP_CHECK(call)                                                                   \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(err_));       \
    } while (0)

__global__ void read_clock(unsigned long long *out) { *out = wall_clock64(); }

// Fill kernel used instead of hipMemsetAsync so passes can be graph-captured.
__global__ void fill_u32(unsigned int *p, uint32_t n, unsigned int v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Work items grouped by puzzle (a counting sort on the header's puzzle
// field), so that the waves of a CU work on few puzzles and their tables
// stay in the caches.  Order within a puzzle is arbitrary; the enumeration
// does not depend on the order of its items.
// Within a wave, consecutive items of one puzzle form a run (the items a pass
// emits mostly come in such runs); each run counts or reserves with one
// atomic.  run_of: the run's first lane and its length.
__device__ __forceinline__ void run_of(uint32_t p, uint32_t lane, uint32_t &first, uint32_t &len) {
    const uint32_t prev = __shfl_up(p, 1);
    const uint64_t heads = __ballot(lane == 0 || p != prev);
    first = 63u - (uint32_t)__builtin_clzll(heads & ((2ull << lane) - 1));   // lane 63: every lane
    const uint64_t later = lane == 63 ? 0ull : heads >> (lane + 1);
    const uint32_t next = later ? lane + 1 + (uint32_t)__builtin_ctzll(later) : 64u;
    len = next - lane;   // meaningful on the run's first lane
}

__global__ void items_per_puzzle(const uint64_t *hdr, uint32_t n_items, unsigned int *count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63;
    const uint32_t p = i < n_items ? hdr_puzzle(hdr[i]) : 0xFFFFFFFFu;
    uint32_t first, len;
    run_of(p, lane, first, len);
    if (first == lane && p != 0xFFFFFFFFu) atomicAdd(&count[p], len);
}

// In-place exclusive prefix sum of n counts, one block of 1024 threads.
__global__ void __launch_bounds__(1024) exclusive_scan_u32(unsigned int *v, uint32_t n) {
    __shared__ unsigned int part[1024];
    unsigned int carry = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const unsigned int x = i < n ? v[i] : 0u;
        part[threadIdx.x] = x;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            const unsigned int y = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < n) v[i] = carry + part[threadIdx.x] - x;
        const unsigned int total = part[1023];
        __syncthreads();
        carry += total;
    }
}

// Each item takes its slot (one atomic per run of one puzzle's items in the
// wave), then the wave copies the placed part of its 64 items' stack rows
// (the transitions of indices below each item's level), kScatterLanes lanes
// per row in u64 words (rows start on 8-byte boundaries: stride is a
// multiple of 4), eight rows at a time.
constexpr uint32_t kScatterLanes = 8;
__global__ void scatter_items(const uint64_t *hdr, const uint16_t *stack, uint32_t n_items, uint32_t stride,
                              unsigned int *offset, uint64_t *out_hdr, uint16_t *out_stack) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, lane = threadIdx.x & 63;
    const bool on = i < n_items;
    const uint64_t h = on ? hdr[i] : 0ull;
    const uint32_t p = on ? hdr_puzzle(h) : 0xFFFFFFFFu;
    uint32_t first, len, base = 0;
    run_of(p, lane, first, len);
    if (first == lane && on) base = atomicAdd(&offset[p], len);
    const uint32_t o = __shfl(base, (int)first) + (lane - first);
    if (on) out_hdr[o] = h;
    const uint32_t words = on ? min((hdr_level(h) + 3) / 4, stride / 4) : 0u;
    const uint32_t wave_first = i - lane;
#pragma unroll
    for (uint32_t g = 0; g < 64 / kScatterLanes; ++g) {
        const uint32_t r = g * kScatterLanes + lane / kScatterLanes, sub = lane % kScatterLanes;
        const uint32_t orow = __shfl(o, (int)r), nw = __shfl(words, (int)r);
        const uint64_t *src = reinterpret_cast<const uint64_t *>(stack + (size_t)(wave_first + r) * stride);
        uint64_t *dst = reinterpret_cast<uint64_t *>(out_stack + (size_t)orow * stride);
        for (uint32_t w = sub; w < nw; w += kScatterLanes) dst[w] = src[w];
    }
}

template <int W>
void launch_expand(const DfsArgs &a, hipStream_t s) {
    const uint32_t threads = 256, blocks = (a.n_in + threads - 1) / threads;
    if (!blocks) return;
    if (a.ranks) hipLaunchKernelGGL((expand_kernel<W, true>), dim3(blocks), dim3(threads), 0, s, a);
    else hipLaunchKernelGGL((expand_kernel<W, false>), dim3(blocks), dim3(threads), 0, s, a);
}

template <int W, bool PERSIST, bool LIMITS>
const void *dfs_fn_p(bool lds, bool ranked) {
    if (ranked)
        return lds ? reinterpret_cast<const void *>(&dfs_kernel<W, true, true, PERSIST, LIMITS>)
                   : reinterpret_cast<const void *>(&dfs_kernel<W, false, true, PERSIST, LIMITS>);
    return lds ? reinterpret_cast<const void *>(&dfs_kernel<W, true, false, PERSIST, LIMITS>)
               : reinterpret_cast<const void *>(&dfs_kernel<W, false, false, PERSIST, LIMITS>);
}

template <int W>
const void *dfs_fn(bool lds, bool ranked, bool persist, bool limits) {
    if (persist) return limits ? dfs_fn_p<W, true, true>(lds, ranked) : dfs_fn_p<W, true, false>(lds, ranked);
    return limits ? dfs_fn_p<W, false, true>(lds, ranked) : dfs_fn_p<W, false, false>(lds, ranked);
}

// VGPRs of the DFS kernel for this row width, table placement, schedule and
// node-limit form.
uint32_t dfs_vgprs(int w, bool lds, bool ranked, bool persist, bool limits) {
    const void *fn = nullptr;
    switch (w) {
        case 1: fn = dfs_fn<1>(lds, ranked, persist, limits); break;
        case 2: fn = dfs_fn<2>(lds, ranked, persist, limits); break;
        case 3: fn = dfs_fn<3>(lds, ranked, persist, limits); break;
        case 4: fn = dfs_fn<4>(lds, ranked, persist, limits); break;
        case 6: fn = dfs_fn<6>(lds, ranked, persist, limits); break;
        case 8: fn = dfs_fn<8>(lds, ranked, persist, limits); break;
        default: fn = dfs_fn<16>(lds, ranked, persist, limits); break;
    }
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, fn) != hipSuccess || attr.numRegs <= 0) {
        (void)hipGetLastError();
        return 128;
    }
    return (uint32_t)attr.numRegs;
}

template <int W, bool PERSIST, bool LIMITS>
void launch_dfs_p(const DfsArgs &a, bool lds, uint32_t blocks, uint32_t threads, size_t lds_bytes, hipStream_t s) {
    if (a.ranks) {
        if (lds) hipLaunchKernelGGL((dfs_kernel<W, true, true, PERSIST, LIMITS>), dim3(blocks), dim3(threads), lds_bytes, s, a);
        else hipLaunchKernelGGL((dfs_kernel<W, false, true, PERSIST, LIMITS>), dim3(blocks), dim3(threads), lds_bytes, s, a);
    } else {
        if (lds) hipLaunchKernelGGL((dfs_kernel<W, true, false, PERSIST, LIMITS>), dim3(blocks), dim3(threads), lds_bytes, s, a);
        else hipLaunchKernelGGL((dfs_kernel<W, false, false, PERSIST, LIMITS>), dim3(blocks), dim3(threads), lds_bytes, s, a);
    }
}

template <int W>
void launch_dfs(const DfsArgs &a, bool lds, uint32_t blocks, uint32_t threads, size_t lds_bytes, hipStream_t s) {
    if (a.persist) {
        if (a.limits) launch_dfs_p<W, true, true>(a, lds, blocks, threads, lds_bytes, s);
        else launch_dfs_p<W, true, false>(a, lds, blocks, threads, lds_bytes, s);
    } else {
        if (a.limits) launch_dfs_p<W, false, true>(a, lds, blocks, threads, lds_bytes, s);
        else launch_dfs_p<W, false, false>(a, lds, blocks, threads, lds_bytes, s);
    }
}

void dispatch_expand(int w, const DfsArgs &a, hipStream_t s) {
    switch (w) {
        case 1: launch_expand<1>(a, s); break;
        case 2: launch_expand<2>(a, s); break;
        case 3: launch_expand<3>(a, s); break;
        case 4: launch_expand<4>(a, s); break;
        case 6: launch_expand<6>(a, s); break;
        case 8: launch_expand<8>(a, s); break;
        default: launch_expand<16>(a, s); break;
    }
}

void dispatch_dfs(int w, const DfsArgs &a, bool lds, uint32_t blocks, uint32_t threads, size_t lds_bytes,
                  hipStream_t s) {
    switch (w) {
        case 1: launch_dfs<1>(a, lds, blocks, threads, lds_bytes, s); break;
        case 2: launch_dfs<2>(a, lds, blocks, threads, lds_bytes, s); break;
        case 3: launch_dfs<3>(a, lds, blocks, threads, lds_bytes, s); break;
        case 4: launch_dfs<4>(a, lds, blocks, threads, lds_bytes, s); break;
        case 6: launch_dfs<6>(a, lds, blocks, threads, lds_bytes, s); break;
        case 8: launch_dfs<8>(a, lds, blocks, threads, lds_bytes, s); break;
        default: launch_dfs<16>(a, lds, blocks, threads, lds_bytes, s); break;
    }
}

// Row widths with a compiled kernel.
uint32_t round_words(uint32_t w) {
    if (w <= 4) return w < 1 ? 1 : w;
    if (w <= 6) return 6;
    if (w <= 8) return 8;
    return 16;
}

}  // namespace

Workspace::~Workspace() { release(); }

void Workspace::release() {
    if (!arena_)
        for (auto &kv : blocks_)
            if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    blocks_.clear();
    arena_used_ = 0;
}

void *Workspace::raw(const std::string &name, size_t bytes) {
    if (arena_) {
        Block &b = blocks_[name];
        if (b.ptr && b.bytes >= bytes) return b.ptr;
        const size_t need = (bytes + kAlign - 1) / kAlign * kAlign;
        if (need > arena_bytes_ - arena_used_)
            throw OutOfMemory("the search needs more than the " + std::to_string(arena_bytes_) +
                              " bytes of its fixed scratch buffer (block '" + name + "' of " +
                              std::to_string(bytes) + " bytes)");
        b.ptr = arena_ + arena_used_;
        b.bytes = need;
        arena_used_ += need;
        return b.ptr;
    }
    Block &b = blocks_[name];
    if (b.ptr && b.bytes >= bytes) return b.ptr;
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    const size_t want = bytes + bytes / 4;
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
        b.ptr = nullptr;
        (void)hipGetLastError();
        throw OutOfMemory("hipMalloc of " + std::to_string(want) + " bytes failed");
    }
    b.bytes = want;
    return b.ptr;
}

void enumerate_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws,
                         const std::vector<const Puzzle *> &puzzles, const EnumConfig &cfg,
                         std::vector<EnumOutcome> &out, EnumStats &stats, const std::vector<EnumRoot> *roots,
                         const std::vector<std::vector<uint16_t>> *ranks, EnumWork *work) {
    const uint32_t n_tab = (uint32_t)puzzles.size();
    if (work && (n_tab != 1 || roots || ranks))
        throw std::invalid_argument("work items describe one puzzle's plain enumeration");
    const bool record = work && work->record_order;
    const bool keep_rest = work && work->keep_rest;
    if (record && cfg.use_graph) throw std::invalid_argument("ordered results are not recorded from graph replays");
    if (keep_rest && !cfg.time_limit_ms && !cfg.node_limit)
        throw std::invalid_argument("keeping the rest of a search needs a time or node limit");
    const bool has_items = work && !work->items.empty();
    // One search per root; by default each puzzle's root item (index 0 holds
    // its first candidate).
    const uint32_t n = roots ? (uint32_t)roots->size() : n_tab;
    if (n == 0 || n > (1u << 20)) throw std::invalid_argument("between 1 and 2^20 searches per launch");
    if (ranks && ranks->size() != n_tab) throw std::invalid_argument("one rank table per puzzle");
    out.assign(n, EnumOutcome{});
    stats = EnumStats{};
    auto wall0 = std::chrono::steady_clock::now();

    // Tables, every mask row padded to the launch's row width.
    uint32_t natural = 1;
    for (uint32_t i = 0; i < n_tab; ++i) natural = std::max(natural, natural_row_words(*puzzles[i]));
    if (natural > 16) throw Unsupported("candidate bitfields wider than 1024 bits");
    const uint32_t row = round_words(natural);
    // Large batches build their tables on several host threads; an error is
    // reported for the first puzzle (in batch order) that has one.
    std::vector<EnumTable> tables(n_tab);
    for_each_item(n_tab, [&](uint32_t i) { tables[i] = build_enum_table(*puzzles[i], cfg.order, row); });
    std::vector<uint32_t> table_off(n_tab);
    // max_reach: one past the last word a read of W words at a kind's offset
    // touches (the interior kind's words come last in the region)
    uint32_t total_words = 0, max_avail = 1, max_P = 0, max_table = 0, max_reach = 0;
    for (uint32_t i = 0; i < n_tab; ++i) {
        EnumTableHeader h = tables[i].header();
        table_off[i] = total_words;
        total_words += h.total_words;
        max_avail = std::max(max_avail, h.avail_words);
        max_reach = std::max(max_reach, h.words[2] + h.words[1] + row);
        max_P = std::max(max_P, h.P);
        max_table = std::max(max_table, h.total_words);
    }
    const int maxw = (int)row;
    const uint32_t stride = (max_P + 3) & ~3u;   // u16 per item, multiple of 4
    stats.max_words = natural;

    // The tables side by side, copied on the host threads into memory first
    // touched there (every word is written: the offsets tile the blob).
    std::unique_ptr<uint64_t[]> blob(new uint64_t[total_words]);
    for_each_item(n_tab, [&](uint32_t i) {
        std::copy(tables[i].blob.begin(), tables[i].blob.end(), blob.get() + table_off[i]);
    });
    const uint32_t n_init = has_items ? (uint32_t)work->items.size() : n;
    if (has_items) {
        const uint32_t P = tables[0].header().P;
        const uint32_t T = tables[0].header().row * 64;
        if (n_init > (1u << 25)) throw std::invalid_argument("at most 2^25 work items");
        for (const EnumItem &it : work->items) {
            if (it.floor < 1 || it.floor > it.level || it.level >= P || it.stack.size() != it.level)
                throw std::invalid_argument("a work item needs 1 <= floor <= level < P and level placed transitions");
            if (it.last < -1 || it.last >= (int)T) throw std::invalid_argument("a work item's last transition is out of range");
            for (uint16_t t : it.stack)
                if (t >= T) throw std::invalid_argument("a work item's transition is out of range");
        }
    }
    std::vector<uint32_t> root_table(n), root_off(n);
    for (uint32_t r = 0; r < n; ++r) {
        root_table[r] = roots ? (*roots)[r].table : r;
        if (root_table[r] >= n_tab) throw std::invalid_argument("root names a puzzle out of range");
        root_off[r] = table_off[root_table[r]];
        if (roots) {
            const size_t len = (*roots)[r].stack.size();
            if (len < 1 || len >= tables[root_table[r]].header().P)
                throw std::invalid_argument("a root places between 1 and P - 1 pieces");
        }
    }

    // One persistent launch (not with the fixed scratch buffer, whose room
    // goes to the pass queues, nor for ordered results, which record the
    // items of each pass), or DFS passes.
    const bool persist = !ws.is_arena() && !record &&
                         (cfg.schedule == EnumConfig::kPersistent || (cfg.schedule == EnumConfig::kAuto && n_tab == 1));

    // Launch geometry: as many resident waves as registers (512 VGPRs per
    // SIMD lane, 4 SIMDs per CU) and LDS (the table once per block plus each
    // wave's stacks) allow; the block size that fits the most waves wins.
    const bool narrow = row <= 4;   // every kind has at most 256 transitions: u8 stack
    // available words of the kinds, and room for reads of W words at the
    // interior kind's offset when it has fewer words than W
    const uint32_t avail_region = std::max(max_avail, max_reach);
    // stacks: a byte plane of levels, and for wide rows a plane of the
    // transitions' high bits, four levels per byte
    const size_t per_wave_bytes = (size_t)avail_region * kWave * 8 + (size_t)stride * kWave + (narrow ? 0 : (size_t)stride / 4 * kWave);
    // One table goes to LDS when it leaves room for four waves per block.
    const bool lds_table = n_tab == 1 && (size_t)max_table * 8 + 4 * per_wave_bytes <= dev.lds_per_block;
    const size_t table_bytes = lds_table ? (size_t)max_table * 8 : 0;
    const size_t lds_cu = dev.lds_per_cu ? dev.lds_per_cu : dev.lds_per_block;
    const uint32_t vgprs = dfs_vgprs(maxw, lds_table, ranks != nullptr, persist, cfg.node_limit != 0);
    const uint32_t reg_waves = 4 * std::min<uint32_t>(8, 512 / ((vgprs + 7) & ~7u));
    auto resident = [&](uint32_t w) -> uint32_t {
        const size_t bytes = table_bytes + w * per_wave_bytes;
        if (bytes > dev.lds_per_block) return 0;
        return std::min<uint32_t>((uint32_t)(lds_cu / bytes) * w, reg_waves / w * w);
    };
    const uint32_t max_wpb = dfs_max_threads(maxw) / kWave;
    uint32_t wpb = cfg.waves_per_block;
    if (!wpb) {
        wpb = 4;
        // multiples of 4 waves spread evenly over the CU's four SIMDs; smaller
        // blocks only when per-wave LDS allows no more
        for (uint32_t w : {4u, 8u, 12u, 16u, 3u, 2u, 1u})
            if (w <= max_wpb && resident(w) > resident(wpb)) wpb = w;
    }
    if (wpb > max_wpb)
        throw std::invalid_argument("at most " + std::to_string(max_wpb) + " waves per block for this puzzle size");
    const uint32_t threads = wpb * kWave;
    const size_t lds_bytes = table_bytes + wpb * per_wave_bytes;
    if (lds_bytes > dev.lds_per_block)
        throw Unsupported("per-block LDS of " + std::to_string(lds_bytes) + " bytes exceeds the device limit");
    const uint32_t wpc = cfg.waves_per_cu ? cfg.waves_per_cu : std::max(wpb, resident(wpb));
    const uint32_t blocks = std::max(1u, dev.compute_units * wpc / wpb);
    const uint64_t lanes = (uint64_t)blocks * threads;
    stats.lds_tables = lds_table;
    stats.blocks = blocks;
    stats.threads = threads;

    // Device buffers.
    uint64_t *d_tables = ws.get<uint64_t>("tables", total_words);
    uint32_t *d_table_off = ws.get<uint32_t>("table_off", n);
    EMP_CHECK(hipMemcpyAsync(d_tables, blob.get(), (size_t)total_words * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_table_off, root_off.data(), n * 4, hipMemcpyHostToDevice, stream));
    uint16_t *d_ranks = nullptr;
    uint64_t *d_rank_off = nullptr;
    uint32_t *d_rank_stride = nullptr;
    if (ranks) {
        // rank rows: P indices of 4 * (largest kind) transitions each
        std::vector<uint64_t> off(n_tab), r_off(n);
        std::vector<uint32_t> stride_t(n_tab), r_stride(n);
        uint64_t total = 0;
        for (uint32_t i = 0; i < n_tab; ++i) {
            const Puzzle &pz = *puzzles[i];
            uint32_t tmax = 0;
            for (int k = 0; k < 3; ++k) tmax = std::max<uint32_t>(tmax, 4 * (uint32_t)pz.pieces_of(k).size());
            if ((*ranks)[i].size() != (size_t)pz.size() * tmax)
                throw std::invalid_argument("a rank table holds P x (4 x pieces of the largest kind) entries");
            off[i] = total;
            stride_t[i] = tmax;
            total += (*ranks)[i].size();
        }
        std::vector<uint16_t> all(total);
        for (uint32_t i = 0; i < n_tab; ++i) std::copy((*ranks)[i].begin(), (*ranks)[i].end(), all.begin() + off[i]);
        for (uint32_t r = 0; r < n; ++r) {
            r_off[r] = off[root_table[r]];
            r_stride[r] = stride_t[root_table[r]];
        }
        d_ranks = ws.get<uint16_t>("ranks", total);
        d_rank_off = ws.get<uint64_t>("rank_off", n);
        d_rank_stride = ws.get<uint32_t>("rank_stride", n);
        EMP_CHECK(hipMemcpyAsync(d_ranks, all.data(), total * 2, hipMemcpyHostToDevice, stream));
        EMP_CHECK(hipMemcpyAsync(d_rank_off, r_off.data(), n * 8, hipMemcpyHostToDevice, stream));
        EMP_CHECK(hipMemcpyAsync(d_rank_stride, r_stride.data(), n * 4, hipMemcpyHostToDevice, stream));
        EMP_CHECK(hipStreamSynchronize(stream));
    }

    const uint64_t target = cfg.target_items ? cfg.target_items : lanes * 4;
    // Per-search buffers first, so that in arena form the work queues take
    // whatever room is left.
    unsigned long long *d_nodes = ws.get<unsigned long long>("nodes", n);
    unsigned long long *d_sols = ws.get<unsigned long long>("sols", n);
    unsigned long long *d_first = ws.get<unsigned long long>("first", n);
    unsigned long long *d_last = ws.get<unsigned long long>("last", n);
    unsigned long long *d_limit = ws.get<unsigned long long>("limit", n);
    unsigned int *d_trunc = ws.get<unsigned int>("trunc", n);
    // [0] in_head [1] out_count [2..7] flags
    unsigned int *d_counters = ws.get<unsigned int>("counters", 8);
    unsigned long long *d_csets = ws.get<unsigned long long>("candidate_sets", 1);
    unsigned int *d_depth = ws.get<unsigned int>("max_depth", n);
    unsigned int *d_timed_out = ws.get<unsigned int>("timed_out", 1);
    unsigned long long *d_work = ws.get<unsigned long long>("work", 2);
    unsigned long long *d_now = ws.get<unsigned long long>("clock_now", 1);
    // Counters saved before each expansion so that an expansion whose
    // children do not fit can be undone and its items searched instead.
    unsigned long long *d_nodes_save = ws.get<unsigned long long>("nodes_save", n);
    unsigned long long *d_sols_save = ws.get<unsigned long long>("sols_save", n);
    unsigned int *d_counters_save = ws.get<unsigned int>("counters_save", 8);
    unsigned int *d_n_in = cfg.use_graph ? ws.get<unsigned int>("n_in", 1) : nullptr;
    const uint32_t sol_per_puzzle = (uint32_t)std::min<uint64_t>(cfg.solution_capacity, (1ull << 26) / n);
    uint16_t *d_sol_stack = ws.get<uint16_t>("sol_stack", (size_t)n * std::max(sol_per_puzzle, 1u) * stride);
    const uint64_t exec_cap = record ? std::max<uint64_t>(work->max_executions, 1) : 0;
    unsigned long long *d_exec_nodes = record ? ws.get<unsigned long long>("exec_nodes", exec_cap) : nullptr;
    uint32_t *d_sol_exec =
        record ? ws.get<uint32_t>("sol_exec", (size_t)n * std::max(sol_per_puzzle, 1u)) : nullptr;
    unsigned long long *d_sol_local =
        record ? ws.get<unsigned long long>("sol_local", (size_t)n * std::max(sol_per_puzzle, 1u)) : nullptr;
    // Batches of several tables read them through the caches: group the
    // items by puzzle before each pass (roots of one table gain nothing).
    // Under a node limit the items publish their progress to the puzzle's
    // count as they run, so the grouped items of one puzzle, taken together,
    // still stop near the limit.
    const bool group_items = !lds_table && n_tab > 1;
    unsigned int *d_group = group_items ? ws.get<unsigned int>("group_count", n) : nullptr;
    unsigned int *d_region_head = group_items ? ws.get<unsigned int>("region_head", kRegions) : nullptr;

    // Work queues.  A pass emits at most two items per input item and always
    // has room for one; a full queue only stops items splitting in two.
    const uint64_t cap64 = std::max<uint64_t>({target * 4, lanes * 16, (uint64_t)n_init * 4, 1u << 16});
    uint32_t cap = cfg.queue_items ? std::max<uint32_t>(cfg.queue_items, n_init)
                                   : (uint32_t)std::min<uint64_t>(cap64, 1u << 25);
    if (ws.is_arena()) {
        // queue rows of header + stack, two queues (three when grouping)
        const size_t row = 8 + (size_t)stride * 2, queues = group_items ? 3 : 2;
        const size_t room = ws.remaining() > 8 * Workspace::kAlign ? ws.remaining() - 8 * Workspace::kAlign : 0;
        const uint64_t fit = room / (row * queues);
        if (fit < n_init)
            throw OutOfMemory("the fixed scratch buffer leaves room for " + std::to_string(fit) +
                              " work items per queue, fewer than the " + std::to_string(n_init) + " to start with");
        cap = (uint32_t)std::min<uint64_t>(cap, fit);
    }
    // prefix-split target within a quarter of the queue
    const uint64_t expand_to = std::min<uint64_t>(target, std::max<uint64_t>(cap / 4, 1));
    uint64_t *d_hdr[2] = {ws.get<uint64_t>("hdr0", cap), ws.get<uint64_t>("hdr1", cap)};
    uint16_t *d_stack[2] = {ws.get<uint16_t>("stack0", (size_t)cap * stride),
                            ws.get<uint16_t>("stack1", (size_t)cap * stride)};
    uint64_t *d_group_hdr = group_items ? ws.get<uint64_t>("group_hdr", cap) : nullptr;
    uint16_t *d_group_stack = group_items ? ws.get<uint16_t>("group_stack", (size_t)cap * stride) : nullptr;
    // The persistent launch's ring of 2^q_shift work-item slots: the queue at
    // the launch first, then one slot per hand-over.
    uint32_t q_cap = 0, q_shift = 0;
    uint64_t *d_qhdr = nullptr;
    uint16_t *d_qstack = nullptr;
    unsigned int *d_qready = nullptr, *d_qcount = nullptr;
    if (persist) {
        // A power of two holding the starting queue and eight items per lane.
        // At most a quarter of it waits untaken (plus one item per lane from
        // waves that give at the same moment), so a slot's previous item has
        // been claimed -- by a lane that waits for it, not one that owes a
        // write -- before its slot is reserved again.
        const uint64_t want_cap = std::max<uint64_t>(cap, lanes * 8);
        while ((1ull << q_shift) < want_cap) ++q_shift;
        q_cap = 1u << q_shift;
        d_qhdr = ws.get<uint64_t>("ticket_hdr", q_cap);
        d_qstack = ws.get<uint16_t>("ticket_stack", (size_t)q_cap * stride);
        d_qready = ws.get<unsigned int>("ticket_ready", q_cap);
        d_qcount = ws.get<unsigned int>("ticket_counts", 4);
    }

    // Root items: index 0 holds its first candidate (one node each), or the
    // given situations (their placements are not counted).
    std::vector<uint64_t> h_hdr(n_init);
    std::vector<uint16_t> h_stack((size_t)n_init * stride, 0);
    std::vector<unsigned long long> h_nodes(n, (roots || has_items) ? 0 : 1), h_zero(n, 0), h_first(n, ~0ull),
        h_limit(n, cfg.node_limit);
    std::vector<unsigned int> h_depth(n, 1);
    if (has_items) {
        // Given items of search 0; their placements are not counted.
        h_depth[0] = 0;
        for (uint32_t i = 0; i < n_init; ++i) {
            const EnumItem &it = work->items[i];
            h_hdr[i] = ((uint64_t)it.floor << 20) | ((uint64_t)it.level << 31) | ((uint64_t)(uint32_t)(it.last + 1) << 42);
            std::copy(it.stack.begin(), it.stack.end(), h_stack.begin() + (size_t)i * stride);
            h_depth[0] = std::max(h_depth[0], it.level);
        }
    } else {
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t len = roots ? (uint32_t)(*roots)[i].stack.size() : 1;
            h_hdr[i] = (uint64_t)i | ((uint64_t)len << 20) | ((uint64_t)len << 31);   // floor = level = len, last = -1
            if (roots) std::copy((*roots)[i].stack.begin(), (*roots)[i].stack.end(), h_stack.begin() + (size_t)i * stride);
            else h_stack[(size_t)i * stride] = (uint16_t)tables[i].header().root;
            h_depth[i] = len;
        }
    }
    EMP_CHECK(hipMemcpyAsync(d_depth, h_depth.data(), n * 4, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_hdr[0], h_hdr.data(), n_init * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_stack[0], h_stack.data(), h_stack.size() * 2, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_nodes, h_nodes.data(), n * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_sols, h_zero.data(), n * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_first, h_first.data(), n * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_last, h_zero.data(), n * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_limit, h_limit.data(), n * 8, hipMemcpyHostToDevice, stream));
    std::vector<unsigned int> h_zero32(std::max<uint32_t>(n, 8), 0);
    EMP_CHECK(hipMemcpyAsync(d_trunc, h_zero32.data(), n * 4, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_counters, h_zero32.data(), 8 * 4, hipMemcpyHostToDevice, stream));

    DfsArgs a{};
    a.tables = d_tables;
    a.table_off = d_table_off;
    a.lds_table_words = lds_table ? max_table : 0;
    a.avail_words = avail_region;
    a.stride = stride;
    a.steps_per_round = cfg.steps_per_round ? cfg.steps_per_round : 128;
    a.step_budget = cfg.step_budget ? cfg.step_budget : (1ull << 16);
    a.drain_budget = std::min<uint64_t>(a.step_budget, cfg.drain_budget ? cfg.drain_budget : kDrainBudget);
    a.nodes = d_nodes;
    a.solutions = d_sols;
    a.first_tick = d_first;
    a.last_tick = d_last;
    a.node_limit = d_limit;
    a.truncated = d_trunc;
    a.in_head = d_counters + 0;
    a.region_head = d_region_head;
    a.out_count = d_counters + 1;
    a.flags = d_counters + 2;
    a.sol_stack = d_sol_stack;
    a.sol_per_puzzle = sol_per_puzzle;
    a.candidate_sets = d_csets;
    a.max_depth = d_depth;
    a.ranks = d_ranks;
    a.rank_off = d_rank_off;
    a.rank_stride = d_rank_stride;
    a.timed_out = d_timed_out;
    a.keep_rest = keep_rest ? 1u : 0u;
    a.limits = cfg.node_limit != 0 ? 1u : 0u;
    a.work = cfg.count_work ? d_work : nullptr;
    if (cfg.count_work) {
        const unsigned long long zero2[2] = {0, 0};
        EMP_CHECK(hipMemcpyAsync(d_work, zero2, 16, hipMemcpyHostToDevice, stream));
    }
    a.exec_nodes = d_exec_nodes;
    a.sol_exec = d_sol_exec;
    a.sol_local = d_sol_local;
    {
        const unsigned int zero = 0;
        EMP_CHECK(hipMemcpyAsync(d_timed_out, &zero, 4, hipMemcpyHostToDevice, stream));
    }
    if (cfg.time_limit_ms) {
        // the device clock counts 100 MHz ticks
        hipLaunchKernelGGL(read_clock, dim3(1), dim3(1), 0, stream, d_now);
        unsigned long long now = 0;
        EMP_CHECK(hipMemcpyAsync(&now, d_now, 8, hipMemcpyDeviceToHost, stream));
        EMP_CHECK(hipStreamSynchronize(stream));
        a.deadline = now + (uint64_t)(cfg.time_limit_ms * 1e5);
    }
    {
        const unsigned long long zero = 0;
        EMP_CHECK(hipMemcpyAsync(d_csets, &zero, 8, hipMemcpyHostToDevice, stream));
    }

    hipEvent_t ev0, ev1, eg0, eg1;
    EMP_CHECK(hipEventCreate(&ev0));
    EMP_CHECK(hipEventCreate(&ev1));
    EMP_CHECK(hipEventCreate(&eg0));
    EMP_CHECK(hipEventCreate(&eg1));
    struct EventGuard {
        hipEvent_t e[4];
        ~EventGuard() {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } guard{{ev0, ev1, eg0, eg1}};

    const char *env = std::getenv("EMP_HIP_VERBOSE");
    const bool verbose = env && *env && *env != '0';

    // Graph replay of the DFS passes: one graph per queue direction holding
    // the counter clear (a fill kernel), the DFS kernel and the copy of the
    // counters to pinned host memory; the pass's input count is read from
    // device memory, which the host sets before each replay.
    struct Graphs {
        hipGraph_t g[2] = {nullptr, nullptr};
        hipGraphExec_t x[2] = {nullptr, nullptr};
        unsigned int *pinned = nullptr;
        ~Graphs() {
            for (int i = 0; i < 2; ++i) {
                if (x[i]) (void)hipGraphExecDestroy(x[i]);
                if (g[i]) (void)hipGraphDestroy(g[i]);
            }
            if (pinned) (void)hipHostFree(pinned);
        }
    } graphs;
    if (cfg.use_graph && !persist) {
        if (ws.is_arena()) throw std::invalid_argument("graph replay allocates: not available with a fixed scratch buffer");
        EMP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&graphs.pinned), 8 * sizeof(unsigned int)));
        for (int dir = 0; dir < 2; ++dir) {
            DfsArgs ga = a;
            ga.in_hdr = d_hdr[dir];
            ga.in_stack = d_stack[dir];
            ga.out_hdr = d_hdr[1 - dir];
            ga.out_stack = d_stack[1 - dir];
            ga.out_cap = cap;
            ga.n_in = 0;
            ga.n_in_dev = d_n_in;
            EMP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, stream, d_counters, 7u, 0u);
            if (d_region_head) hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, stream, d_region_head, kRegions, 0u);
            dispatch_dfs(maxw, ga, lds_table, blocks, threads, lds_bytes, stream);
            (void)hipMemcpyAsync(graphs.pinned, d_counters, 8 * sizeof(unsigned int), hipMemcpyDeviceToHost, stream);
            EMP_CHECK(hipStreamEndCapture(stream, &graphs.g[dir]));
            EMP_CHECK(hipGraphInstantiate(&graphs.x[dir], graphs.g[dir], nullptr, nullptr, 0));
        }
    }
    int cur = 0;
    bool expansion_grows = true;
    uint32_t flat_expansions = 0;   // expansions in a row that did not grow the queue
    uint64_t n_items = n_init;
    // ordered results: each DFS pass's input items (their keys) and the
    // executions recorded so far
    std::vector<uint64_t> exec_hdr;
    std::vector<uint16_t> exec_stack;
    uint64_t exec_base = 0;
    uint32_t expansions = 0;
    const uint32_t max_expansions = cfg.max_expansions ? cfg.max_expansions : 1024;
    while (n_items) {
        a.in_hdr = d_hdr[cur];
        a.in_stack = d_stack[cur];
        a.n_in = (uint32_t)n_items;
        a.out_hdr = d_hdr[1 - cur];
        a.out_stack = d_stack[1 - cur];
        a.out_cap = cap;
        // Expand while the queue is short of the target and expansions still
        // multiply it (resumed searches pass through an expansion unchanged).
        const bool expand = n_items < expand_to && expansions < max_expansions && expansion_grows;
        const bool grouped = !expand && group_items && n_items > 1;
        const bool one_launch = persist && !expand;
        if (grouped) {
            const uint32_t ni = (uint32_t)n_items, g = (ni + 255) / 256;
            EMP_CHECK(hipEventRecord(eg0, stream));
            hipLaunchKernelGGL(fill_u32, dim3((n + 255) / 256), dim3(256), 0, stream, d_group, n, 0u);
            hipLaunchKernelGGL(items_per_puzzle, dim3(g), dim3(256), 0, stream, d_hdr[cur], ni, d_group);
            hipLaunchKernelGGL(exclusive_scan_u32, dim3(1), dim3(1024), 0, stream, d_group, n);
            hipLaunchKernelGGL(scatter_items, dim3(g), dim3(256), 0, stream, d_hdr[cur], d_stack[cur], ni, stride,
                               d_group, d_group_hdr, d_group_stack);
            EMP_CHECK(hipGetLastError());
            EMP_CHECK(hipMemcpyAsync(d_hdr[cur], d_group_hdr, (size_t)ni * 8, hipMemcpyDeviceToDevice, stream));
            EMP_CHECK(hipMemcpyAsync(d_stack[cur], d_group_stack, (size_t)ni * stride * 2, hipMemcpyDeviceToDevice,
                                     stream));
            EMP_CHECK(hipEventRecord(eg1, stream));
        }
        if (one_launch) {
            // The queue becomes tickets 0..n-1 of the ticket array; one launch
            // searches them and everything handed over from them.
            const uint32_t ni = (uint32_t)n_items;
            if (ni > q_cap)
                throw Unsupported("the persistent search starts from " + std::to_string(ni) + " items, more than its " +
                                  std::to_string(q_cap) + " slots");
            EMP_CHECK(hipMemcpyAsync(d_qhdr, d_hdr[cur], (size_t)ni * 8, hipMemcpyDeviceToDevice, stream));
            EMP_CHECK(hipMemcpyAsync(d_qstack, d_stack[cur], (size_t)ni * stride * 2, hipMemcpyDeviceToDevice, stream));
            hipLaunchKernelGGL(fill_u32, dim3((ni + 255) / 256), dim3(256), 0, stream, d_qready, ni, 1u);
            if (q_cap > ni)
                hipLaunchKernelGGL(fill_u32, dim3((q_cap - ni + 255) / 256), dim3(256), 0, stream, d_qready + ni,
                                   q_cap - ni, 0u);
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, stream, d_counters, 7u, 0u);
            const unsigned int qc[4] = {0u, ni, ni, 0u};   // head, tail, pending, stalled
            EMP_CHECK(hipMemcpyAsync(d_qcount, qc, sizeof(qc), hipMemcpyHostToDevice, stream));
            DfsArgs pa = a;
            pa.persist = 1;
            pa.q_shift = q_shift;
            pa.q_backlog = std::max<uint32_t>(q_cap / 4, 1u);
            pa.q_head = d_qcount + 0;
            pa.q_tail = d_qcount + 1;
            pa.q_pending = d_qcount + 2;
            pa.q_ready = d_qready;
            pa.q_hdr = d_qhdr;
            pa.q_stack = d_qstack;
            pa.q_stalled = d_qcount + 3;
            pa.in_hdr = d_qhdr;
            pa.in_stack = d_qstack;
            pa.n_in = ni;
            pa.n_in_dev = nullptr;
            pa.out_hdr = d_hdr[1 - cur];   // the rest of a cut run
            pa.out_stack = d_stack[1 - cur];
            pa.out_cap = cap;
            pa.exec_base = 0;
            EMP_CHECK(hipEventRecord(ev0, stream));
            dispatch_dfs(maxw, pa, lds_table, blocks, threads, lds_bytes, stream);
            EMP_CHECK(hipGetLastError());
            EMP_CHECK(hipEventRecord(ev1, stream));
            unsigned int counters[8], qend[4];
            EMP_CHECK(hipMemcpyAsync(counters, d_counters, sizeof(counters), hipMemcpyDeviceToHost, stream));
            EMP_CHECK(hipMemcpyAsync(qend, d_qcount, sizeof(qend), hipMemcpyDeviceToHost, stream));
            EMP_CHECK(hipStreamSynchronize(stream));
            float ms = 0;
            EMP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
            if (grouped) {
                float gms = 0;
                EMP_CHECK(hipEventElapsedTime(&gms, eg0, eg1));
                stats.group_ms += gms;
            }
            stats.dfs_kernel_ms += ms;
            ++stats.dfs_launches;
            stats.splits += counters[3];
            stats.items_processed += counters[4];
            stats.tickets = qend[1];
            if (qend[3]) throw DeviceError("a wave of the persistent search waited 20 s for work and gave up");
            if (qend[2] != 0) throw DeviceError("the persistent search ended with work pending");
            if (counters[2]) throw DeviceError("rest queue overflow in the persistent search");
            if (counters[7])
                throw std::invalid_argument("a situation places an identical piece while a lower-numbered twin is "
                                            "still available (such a piece is never a candidate)");
            if (keep_rest) {
                unsigned int cut = 0;
                EMP_CHECK(hipMemcpyAsync(&cut, d_timed_out, 4, hipMemcpyDeviceToHost, stream));
                EMP_CHECK(hipStreamSynchronize(stream));
                if (cut) {
                    const uint32_t m = std::min<uint32_t>(counters[1], cap);
                    std::vector<uint64_t> rh(m);
                    std::vector<uint16_t> rs((size_t)m * stride);
                    EMP_CHECK(hipMemcpyAsync(rh.data(), d_hdr[1 - cur], (size_t)m * 8, hipMemcpyDeviceToHost, stream));
                    EMP_CHECK(hipMemcpyAsync(rs.data(), d_stack[1 - cur], (size_t)m * stride * 2,
                                             hipMemcpyDeviceToHost, stream));
                    EMP_CHECK(hipStreamSynchronize(stream));
                    work->rest.clear();
                    for (uint32_t i = 0; i < m; ++i) {
                        const uint64_t h = rh[i];
                        EnumItem it;
                        it.floor = (uint32_t)(h >> 20) & 0x7FF;
                        it.level = (uint32_t)(h >> 31) & 0x7FF;
                        it.last = (int)((uint32_t)(h >> 42) & 0xFFF) - 1;
                        it.stack.assign(rs.begin() + (size_t)i * stride, rs.begin() + (size_t)i * stride + it.level);
                        work->rest.push_back(std::move(it));
                    }
                }
            }
            if (verbose)
                std::fprintf(stderr, "emp_hip: persistent in=%u tickets=%u handovers=%u %.3f ms\n", ni, qend[1],
                             counters[3], ms);
            break;
        }
        if (!cfg.use_graph || expand) {
            hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, stream, d_counters, 7u, 0u);
            if (d_region_head) hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, stream, d_region_head, kRegions, 0u);
        }
        if (expand) {
            EMP_CHECK(hipMemcpyAsync(d_nodes_save, d_nodes, n * 8, hipMemcpyDeviceToDevice, stream));
            EMP_CHECK(hipMemcpyAsync(d_sols_save, d_sols, n * 8, hipMemcpyDeviceToDevice, stream));
            EMP_CHECK(hipMemcpyAsync(d_counters_save, d_counters, 8 * 4, hipMemcpyDeviceToDevice, stream));
        }
        if (record && !expand) {
            // this pass's input items are executions exec_base..+n_items:
            // keep their keys (header + stack) and clear their counts
            if (exec_base + n_items > exec_cap)
                throw Unsupported("an ordered enumeration recorded more than " + std::to_string(exec_cap) +
                                  " work items; raise the step budget");
            const uint32_t words = (uint32_t)(2 * n_items);
            hipLaunchKernelGGL(fill_u32, dim3((words + 255) / 256), dim3(256), 0, stream,
                               reinterpret_cast<unsigned int *>(d_exec_nodes + exec_base), words, 0u);
            const size_t h0 = exec_hdr.size(), s0 = exec_stack.size();
            exec_hdr.resize(h0 + n_items);
            exec_stack.resize(s0 + (size_t)n_items * stride);
            EMP_CHECK(hipMemcpyAsync(exec_hdr.data() + h0, d_hdr[cur], n_items * 8, hipMemcpyDeviceToHost, stream));
            EMP_CHECK(hipMemcpyAsync(exec_stack.data() + s0, d_stack[cur], (size_t)n_items * stride * 2,
                                     hipMemcpyDeviceToHost, stream));
            a.exec_base = (uint32_t)exec_base;
        }
        unsigned int counters[8];
        if (cfg.use_graph && !expand) {
            const unsigned int n_in32 = (uint32_t)n_items;
            EMP_CHECK(hipMemcpyAsync(d_n_in, &n_in32, 4, hipMemcpyHostToDevice, stream));
            EMP_CHECK(hipEventRecord(ev0, stream));
            EMP_CHECK(hipGraphLaunch(graphs.x[cur], stream));
            EMP_CHECK(hipEventRecord(ev1, stream));
            EMP_CHECK(hipStreamSynchronize(stream));
            std::memcpy(counters, graphs.pinned, sizeof(counters));
        } else {
            EMP_CHECK(hipEventRecord(ev0, stream));
            if (expand) dispatch_expand(maxw, a, stream);
            else dispatch_dfs(maxw, a, lds_table, blocks, threads, lds_bytes, stream);
            EMP_CHECK(hipGetLastError());
            EMP_CHECK(hipEventRecord(ev1, stream));
            EMP_CHECK(hipMemcpyAsync(counters, d_counters, sizeof(counters), hipMemcpyDeviceToHost, stream));
            EMP_CHECK(hipStreamSynchronize(stream));
        }
        float ms = 0;
        EMP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        if (grouped) {
            float gms = 0;
            EMP_CHECK(hipEventElapsedTime(&gms, eg0, eg1));
            stats.group_ms += gms;
        }
        if (expand) {
            ++expansions;
            stats.expand_kernel_ms += ms;
            ++stats.expand_launches;
            if (counters[2]) {
                // Children did not fit: undo the expansion's counts (the
                // input items are intact) and search the items instead.
                EMP_CHECK(hipMemcpyAsync(d_nodes, d_nodes_save, n * 8, hipMemcpyDeviceToDevice, stream));
                EMP_CHECK(hipMemcpyAsync(d_sols, d_sols_save, n * 8, hipMemcpyDeviceToDevice, stream));
                EMP_CHECK(hipMemcpyAsync(d_counters, d_counters_save, 8 * 4, hipMemcpyDeviceToDevice, stream));
                expansions = max_expansions;
                continue;
            }
            // DFS passes take over once an expansion stops multiplying the
            // queue (they split items themselves).  A persistent launch
            // without a node limit starts from what the expansions made, so
            // they go on through a few levels that do not grow (corners,
            // frame turns); under a node limit breadth would spend the budget
            // on shallow levels, so they stop as for passes.
            flat_expansions = counters[1] > n_items ? 0 : flat_expansions + 1;
            expansion_grows = persist && !cfg.node_limit ? counters[1] > 0 && flat_expansions < 4
                                                         : counters[1] >= n_items + n_items / 4;
        } else {
            expansion_grows = true;
            stats.dfs_kernel_ms += ms;
            ++stats.dfs_launches;
            stats.splits += counters[3];
            stats.items_processed += counters[4];
            if (counters[2]) throw DeviceError("work queue overflow in a search pass");
            if (counters[7])
                throw std::invalid_argument("a situation places an identical piece while a lower-numbered twin is "
                                            "still available (such a piece is never a candidate)");
            if (record) exec_base += n_items;
            if (keep_rest) {
                unsigned int cut = 0;
                EMP_CHECK(hipMemcpyAsync(&cut, d_timed_out, 4, hipMemcpyDeviceToHost, stream));
                EMP_CHECK(hipStreamSynchronize(stream));
                if (cut) {
                    // The output queue holds everything not searched: the
                    // items requeued at the deadline and the split halves.
                    const uint32_t m = std::min<uint32_t>(counters[1], cap);
                    std::vector<uint64_t> rh(m);
                    std::vector<uint16_t> rs((size_t)m * stride);
                    EMP_CHECK(hipMemcpyAsync(rh.data(), d_hdr[1 - cur], (size_t)m * 8, hipMemcpyDeviceToHost, stream));
                    EMP_CHECK(hipMemcpyAsync(rs.data(), d_stack[1 - cur], (size_t)m * stride * 2,
                                             hipMemcpyDeviceToHost, stream));
                    EMP_CHECK(hipStreamSynchronize(stream));
                    work->rest.clear();
                    for (uint32_t i = 0; i < m; ++i) {
                        const uint64_t h = rh[i];
                        const uint32_t fl = (uint32_t)(h >> 20) & 0x7FF;
                        if (fl == kDeadItem) continue;
                        EnumItem it;
                        it.floor = fl;
                        it.level = (uint32_t)(h >> 31) & 0x7FF;
                        it.last = (int)((uint32_t)(h >> 42) & 0xFFF) - 1;
                        it.stack.assign(rs.begin() + (size_t)i * stride, rs.begin() + (size_t)i * stride + it.level);
                        work->rest.push_back(std::move(it));
                    }
                    break;
                }
            }
        }
        n_items = std::min<uint64_t>(counters[1], cap);
        cur = 1 - cur;
        if (verbose)
            std::fprintf(stderr, "emp_hip: %s in=%u out=%u splits=%u overflow=%u %.3f ms\n",
                         expand ? "expand" : "dfs", a.n_in, counters[1], counters[3], counters[2], ms);
    }

    // Results.
    std::vector<unsigned long long> r_nodes(n), r_sols(n), r_first(n), r_last(n);
    std::vector<unsigned int> r_trunc(n);
    EMP_CHECK(hipMemcpyAsync(r_nodes.data(), d_nodes, n * 8, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipMemcpyAsync(r_sols.data(), d_sols, n * 8, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipMemcpyAsync(r_first.data(), d_first, n * 8, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipMemcpyAsync(r_last.data(), d_last, n * 8, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipMemcpyAsync(r_trunc.data(), d_trunc, n * 4, hipMemcpyDeviceToHost, stream));
    unsigned long long r_csets = 0;
    EMP_CHECK(hipMemcpyAsync(&r_csets, d_csets, 8, hipMemcpyDeviceToHost, stream));
    std::vector<unsigned int> r_depth(n);
    EMP_CHECK(hipMemcpyAsync(r_depth.data(), d_depth, n * 4, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipStreamSynchronize(stream));
    stats.candidate_sets = r_csets;
    if (cfg.count_work) {
        unsigned long long w[2] = {0, 0};
        EMP_CHECK(hipMemcpyAsync(w, d_work, 16, hipMemcpyDeviceToHost, stream));
        EMP_CHECK(hipStreamSynchronize(stream));
        stats.work_bytes = w[0];
        stats.work_ops = w[1];
        if (!roots && !has_items)
            // the root placements (index 0, counted on the host)
            for (uint32_t i = 0; i < n; ++i) {
                const EnumTableHeader h = tables[i].header();
                const uint64_t M = (4ull * h.P + 63) / 64;
                const uint64_t nn = h.P > 1 ? (tables[i].blob[h.off_nbr + kNbrWords] & 7) : 0;
                stats.work_bytes += 8 * M * (nn + 3);
                stats.work_ops += 2 * M * (nn + 2);
            }
    }
    std::vector<uint16_t> r_stack((size_t)n * sol_per_puzzle * stride);
    if (!r_stack.empty()) {
        EMP_CHECK(hipMemcpyAsync(r_stack.data(), d_sol_stack, r_stack.size() * 2, hipMemcpyDeviceToHost, stream));
        EMP_CHECK(hipStreamSynchronize(stream));
    }
    const double tick_ms = 1e-5;   // wall_clock64 runs at 100 MHz on gfx9
    for (uint32_t i = 0; i < n; ++i) {
        EnumOutcome &o = out[i];
        o.nodes = r_nodes[i];
        o.solutions = r_sols[i];
        o.complete = !r_trunc[i];
        o.max_depth = r_depth[i];
        o.solve_ms = (r_first[i] != ~0ull && r_last[i] >= r_first[i]) ? (double)(r_last[i] - r_first[i]) * tick_ms : 0.0;
    }
    // Ordered results: each execution's part of the tree is an interval of
    // the sequential DFS starting right after its key -- its stack, then its
    // `last` transition (-1 before every transition), compared transition by
    // transition; when one key extends the other, the longer one lies inside
    // the subtree the shorter one has finished, so it comes first.  The
    // executions sorted that way give each one the nodes before it.
    std::vector<uint64_t> exec_offset;
    std::vector<uint32_t> r_sol_exec;
    std::vector<unsigned long long> r_sol_local;
    if (record) {
        std::vector<unsigned long long> en(exec_base);
        r_sol_exec.resize((size_t)n * sol_per_puzzle);
        r_sol_local.resize((size_t)n * sol_per_puzzle);
        if (exec_base) EMP_CHECK(hipMemcpyAsync(en.data(), d_exec_nodes, exec_base * 8, hipMemcpyDeviceToHost, stream));
        if (!r_sol_exec.empty()) {
            EMP_CHECK(hipMemcpyAsync(r_sol_exec.data(), d_sol_exec, r_sol_exec.size() * 4, hipMemcpyDeviceToHost, stream));
            EMP_CHECK(hipMemcpyAsync(r_sol_local.data(), d_sol_local, r_sol_local.size() * 8, hipMemcpyDeviceToHost,
                                     stream));
        }
        EMP_CHECK(hipStreamSynchronize(stream));
        auto level_of = [](uint64_t h) { return (uint32_t)(h >> 31) & 0x7FF; };
        auto last_of = [](uint64_t h) { return (int)((uint32_t)(h >> 42) & 0xFFF) - 1; };
        auto key_less = [&](uint64_t x, uint64_t y) {
            const uint64_t hx = exec_hdr[x], hy = exec_hdr[y];
            const uint32_t px = (uint32_t)hx & 0xFFFFF, py = (uint32_t)hy & 0xFFFFF;
            if (px != py) return px < py;
            const uint32_t lx = level_of(hx), ly = level_of(hy);
            const uint16_t *sx = &exec_stack[x * stride], *sy = &exec_stack[y * stride];
            for (uint32_t j = 0, m = std::min(lx, ly); j <= m; ++j) {
                const int vx = j < lx ? (int)sx[j] : last_of(hx);
                const int vy = j < ly ? (int)sy[j] : last_of(hy);
                if (vx != vy) return vx < vy;
            }
            return lx > ly;
        };
        std::vector<uint64_t> by_start(exec_base);
        for (uint64_t e = 0; e < exec_base; ++e) by_start[e] = e;
        std::stable_sort(by_start.begin(), by_start.end(), key_less);
        exec_offset.assign(exec_base, 0);
        uint64_t run = 0;
        uint32_t search = ~0u;
        for (uint64_t e : by_start) {
            const uint32_t ps = (uint32_t)exec_hdr[e] & 0xFFFFF;
            if (ps != search) run = 0, search = ps;
            exec_offset[e] = run;
            run += en[e];
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        EnumOutcome &o = out[i];
        const EnumTable &tb = tables[root_table[i]];
        const uint32_t P = tb.header().P;
        const uint64_t kept = std::min<uint64_t>(o.solutions, sol_per_puzzle);
        o.boards.resize(kept * P);
        std::vector<uint64_t> sol_order(kept);
        for (uint64_t s = 0; s < kept; ++s) sol_order[s] = s;
        if (record) {
            // node index = the root placement (default root), the nodes of
            // the executions before this one, and the nodes this execution
            // had counted when it placed the solution
            const uint64_t base = (!roots && !has_items) ? 1 : 0;
            std::vector<uint64_t> idx(kept);
            for (uint64_t s = 0; s < kept; ++s) {
                const size_t slot = (size_t)i * sol_per_puzzle + s;
                idx[s] = base + exec_offset.at(r_sol_exec[slot]) + r_sol_local[slot];
            }
            std::stable_sort(sol_order.begin(), sol_order.end(), [&](uint64_t x, uint64_t y) { return idx[x] < idx[y]; });
            o.node_index.resize(kept);
            for (uint64_t s = 0; s < kept; ++s) o.node_index[s] = idx[sol_order[s]];
        }
        for (uint64_t s = 0; s < kept; ++s)
            stack_to_board(tb, &r_stack[((size_t)i * sol_per_puzzle + sol_order[s]) * stride], &o.boards[s * P]);
    }
    stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

void expand_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &pz, int order,
                      uint32_t depth, uint64_t capacity, std::vector<uint16_t> &boards, uint64_t &n_items,
                      uint64_t &nodes, uint64_t &solutions) {
    (void)dev;
    const uint32_t natural = natural_row_words(pz);
    if (natural > 16) throw Unsupported("candidate bitfields wider than 1024 bits");
    const int maxw = (int)round_words(natural);
    const EnumTable t = build_enum_table(pz, order, (uint32_t)maxw);
    const EnumTableHeader h = t.header();
    if (depth < 1 || depth > h.P) throw std::invalid_argument("expansion depth must be in 1..P");
    const uint32_t stride = (h.P + 3) & ~3u;
    // Device queues hold whole levels; `capacity` only bounds the boards returned.
    const uint32_t cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(capacity, 1u << 22), 1u << 25);
    uint64_t *d_tables = ws.get<uint64_t>("x_tables", h.total_words);
    uint32_t *d_off = ws.get<uint32_t>("x_off", 1);
    uint64_t *d_hdr[2] = {ws.get<uint64_t>("x_hdr0", cap), ws.get<uint64_t>("x_hdr1", cap)};
    uint16_t *d_stack[2] = {ws.get<uint16_t>("x_stack0", (size_t)cap * stride),
                            ws.get<uint16_t>("x_stack1", (size_t)cap * stride)};
    unsigned long long *d_nodes = ws.get<unsigned long long>("x_nodes", 1);
    unsigned long long *d_sols = ws.get<unsigned long long>("x_sols", 1);
    unsigned int *d_counters = ws.get<unsigned int>("x_counters", 8);
    uint16_t *d_sol_stack = ws.get<uint16_t>("x_sol_stack", stride);
    const uint32_t zero_off = 0;
    const unsigned long long one = 1, zero = 0;
    const uint64_t root_hdr = (1ull << 20) | (1ull << 31);
    std::vector<uint16_t> root(stride, 0);
    root[0] = (uint16_t)h.root;
    EMP_CHECK(hipMemcpyAsync(d_tables, t.blob.data(), t.blob.size() * 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_off, &zero_off, 4, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_hdr[0], &root_hdr, 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_stack[0], root.data(), stride * 2, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_nodes, &one, 8, hipMemcpyHostToDevice, stream));
    EMP_CHECK(hipMemcpyAsync(d_sols, &zero, 8, hipMemcpyHostToDevice, stream));
    DfsArgs a{};
    a.tables = d_tables;
    a.table_off = d_off;
    a.stride = stride;
    a.nodes = d_nodes;
    a.solutions = d_sols;
    a.out_count = d_counters + 1;
    a.flags = d_counters + 2;
    a.sol_stack = d_sol_stack;
    a.sol_per_puzzle = 0;
    a.max_depth = ws.get<unsigned int>("x_depth", 1);
    int cur = 0;
    uint64_t items = 1;
    for (uint32_t level = 1; level < depth && items; ++level) {
        a.in_hdr = d_hdr[cur];
        a.in_stack = d_stack[cur];
        a.n_in = (uint32_t)items;
        a.out_hdr = d_hdr[1 - cur];
        a.out_stack = d_stack[1 - cur];
        a.out_cap = cap;
        hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, stream, d_counters, 7u, 0u);
        dispatch_expand(maxw, a, stream);
        EMP_CHECK(hipGetLastError());
        unsigned int counters[8];
        EMP_CHECK(hipMemcpyAsync(counters, d_counters, sizeof(counters), hipMemcpyDeviceToHost, stream));
        EMP_CHECK(hipStreamSynchronize(stream));
        if (counters[2]) throw OutOfMemory("expansion exceeds the capacity of " + std::to_string(cap) + " situations");
        items = counters[1];
        cur = 1 - cur;
    }
    unsigned long long r_nodes = 0, r_sols = 0;
    EMP_CHECK(hipMemcpyAsync(&r_nodes, d_nodes, 8, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipMemcpyAsync(&r_sols, d_sols, 8, hipMemcpyDeviceToHost, stream));
    std::vector<uint16_t> stacks((size_t)items * stride);
    if (items) EMP_CHECK(hipMemcpyAsync(stacks.data(), d_stack[cur], stacks.size() * 2, hipMemcpyDeviceToHost, stream));
    EMP_CHECK(hipStreamSynchronize(stream));
    nodes = r_nodes;
    solutions = r_sols;
    n_items = items;
    const uint64_t keep = std::min<uint64_t>(items, capacity);
    boards.assign((size_t)keep * h.P, 0xFFFF);
    for (uint64_t i = 0; i < keep; ++i) {
        const uint16_t *st = &stacks[(size_t)i * stride];
        uint16_t *bd = &boards[(size_t)i * h.P];
        for (uint32_t k = 0; k < depth && k < h.P; ++k) {
            const int kind = (int)(t.blob[h.off_pos + k] & 3);
            bd[t.order[k].second * h.W + t.order[k].first] =
                (uint16_t)(((uint32_t)t.piece_of[kind][st[k] >> 2] << 2) | (st[k] & 3));
        }
    }
}

}  // namespace emp
