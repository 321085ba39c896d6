// emp_glutton_wide.hip -- glutton_wide's breadth step on gfx950.
//
// The reference's glutton_wide (src/CUDA_glutton_wide.cpp:29-67) takes the
// level-0 situation (every position's initial capability,
// include/CUDA_common_glutton.h:148-211), walks the candidate bits of the
// (0,0) corner in bit order and builds one child per bit with
// CUDA_glutton_situations::play_from (include/CUDA_glutton_situations.h
// :713-891): the child's masks are the parent's with the played piece removed
// everywhere (its four bits) and, at each free neighbour the played piece
// shows a colour to, the colour constraint of that side ANDed in; the played
// position's mask leaves the situation.  Here every parent is a list of moves
// played from level 0 that way, and each is expanded at one given cell.
//
//   parents_kernel   one block per parent: replays its moves on the initial
//                    capability (masks in device memory, P x nw u64) and
//                    reports whether a free position lost every option.
//   children_kernel  one block per child: the parent's masks with the move
//                    applied, and the same report (the reference stops the
//                    program there; a child is flagged instead).
// Masks are piece-major (include/CUDA_piece_position_info2.h): bit
// 4 * piece + orientation, nw = ceil(4P / 64) u64 words per position.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "emp_glutton.h"
#include "emp_glutton_wide.h"

namespace emp {

namespace {

struct WideArgs {
    uint32_t P, W, H, nw;
    const uint64_t *init;       // [P][nw]
    const uint64_t *cm;         // [((c - 1) * 4 + side) * nw]
    const uint32_t *piece_col;  // N | E << 8 | S << 16 | W << 24
};

__device__ __forceinline__ uint32_t colour_of(const WideArgs &a, uint32_t piece, uint32_t side, uint32_t o) {
    return (a.piece_col[piece] >> (8 * ((side + o) & 3))) & 0xFF;
}

// Plays move (pos, piece, o) on masks m [P][nw] of a situation whose placed
// positions are marked in placed[] (block-shared); all threads of the block
// take part.  Returns (block-uniform) whether a free position has no option.
__device__ bool play(const WideArgs &a, uint64_t *m, uint8_t *placed, uint32_t mv) {
    const uint32_t pos = mv & 0xFF, piece = (mv >> 8) & 0xFF, o = (mv >> 16) & 3;
    const uint32_t pw = piece / 16;
    const uint64_t keep = ~(0xFull << (4 * (piece % 16)));
    const uint32_t x = pos % a.W, y = pos / a.W;
    // neighbour of each side that receives a colour constraint, or P
    uint32_t nb[4], col[4];
    for (uint32_t s = 0; s < 4; ++s) {
        const int nx = (int)x + (s == 1) - (s == 3), ny = (int)y + (s == 2) - (s == 0);
        const bool inside = nx >= 0 && ny >= 0 && nx < (int)a.W && ny < (int)a.H;
        const uint32_t q = inside ? (uint32_t)ny * a.W + (uint32_t)nx : a.P;
        col[s] = colour_of(a, piece, s, o);
        nb[s] = (q < a.P && col[s] && !placed[q]) ? q : a.P;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.P * a.nw; i += blockDim.x) {
        const uint32_t q = i / a.nw, w = i % a.nw;
        uint64_t v = m[i];
        if (w == pw) v &= keep;
        for (uint32_t s = 0; s < 4; ++s)
            if (q == nb[s]) v &= a.cm[((size_t)(col[s] - 1) * 4 + s) * a.nw + w];
        m[i] = q == pos ? 0ull : v;
    }
    __syncthreads();
    if (threadIdx.x == 0) placed[pos] = 1;
    __syncthreads();
    // a free position without option
    __shared__ int dead;
    if (threadIdx.x == 0) dead = 0;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < a.P; q += blockDim.x) {
        if (placed[q]) continue;
        uint64_t any = 0;
        for (uint32_t w = 0; w < a.nw; ++w) any |= m[(size_t)q * a.nw + w];
        if (!any) atomicOr(&dead, 1);
    }
    __syncthreads();
    return dead != 0;
}

__device__ void load_initial(const WideArgs &a, uint64_t *m, uint8_t *placed) {
    for (uint32_t i = threadIdx.x; i < a.P * a.nw; i += blockDim.x) m[i] = a.init[i];
    for (uint32_t q = threadIdx.x; q < a.P; q += blockDim.x) placed[q] = 0;
    __syncthreads();
}

// status: 0 valid, 1 a free position without option, 2 a move on a placed
// position or with a used piece or outside the position's mask.
__global__ void __launch_bounds__(256) parents_kernel(WideArgs a, const uint32_t *moves, uint32_t len,
                                                      uint64_t *masks, uint32_t *status) {
    extern __shared__ uint8_t placed[];
    const uint32_t p = blockIdx.x;
    uint64_t *m = masks + (size_t)p * a.P * a.nw;
    load_initial(a, m, placed);
    uint32_t st = 0;
    for (uint32_t i = 0; i < len && st == 0; ++i) {
        const uint32_t mv = moves[(size_t)p * len + i];
        const uint32_t pos = mv & 0xFF, piece = (mv >> 8) & 0xFF, o = (mv >> 16) & 3;
        const uint32_t b = 4 * piece + o;
        if (pos >= a.P || piece >= a.P || placed[pos] || !((m[(size_t)pos * a.nw + b / 64] >> (b % 64)) & 1)) {
            st = 2;
            break;
        }
        if (play(a, m, placed, mv)) st = 1;
    }
    if (threadIdx.x == 0) status[p] = st;
}

__global__ void __launch_bounds__(256) children_kernel(WideArgs a, const uint32_t *parent_moves, uint32_t len,
                                                       const uint64_t *parent_masks, const uint32_t *child_parent,
                                                       const uint32_t *child_move, uint64_t *masks,
                                                       uint32_t *status) {
    extern __shared__ uint8_t placed[];
    const uint32_t c = blockIdx.x, p = child_parent[c];
    uint64_t *m = masks + (size_t)c * a.P * a.nw;
    const uint64_t *pm = parent_masks + (size_t)p * a.P * a.nw;
    for (uint32_t i = threadIdx.x; i < a.P * a.nw; i += blockDim.x) m[i] = pm[i];
    for (uint32_t q = threadIdx.x; q < a.P; q += blockDim.x) placed[q] = 0;
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t i = 0; i < len; ++i) placed[parent_moves[(size_t)p * len + i] & 0xFF] = 1;
    __syncthreads();
    const bool dead = play(a, m, placed, child_move[c]);
    if (threadIdx.x == 0) status[c] = dead ? 1u : 0u;
}

#define WIDE_CHECK(call)                                                                  \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(err_));       \
    } while (0)

}  // namespace

void glutton_wide_on_device(hipStream_t stream, Workspace &ws, const Puzzle &pz, const std::vector<uint32_t> &parents,
                            uint32_t len, uint32_t n_parents, const std::vector<uint32_t> &cells, WideOutcome &out) {
    const GluttonTables t = build_glutton_tables(pz);
    if (t.P > 256) throw std::domain_error("moves pack positions and pieces in 8 bits (at most 256 pieces)");
    if (parents.size() != (size_t)n_parents * len || cells.size() != n_parents)
        throw std::invalid_argument("one row of moves and one cell per parent");
    const size_t per = (size_t)t.P * t.nw;
    WideArgs a{t.P, t.W, pz.height(), t.nw, nullptr, nullptr, nullptr};
    uint64_t *d_init = ws.get<uint64_t>("w_init", per);
    uint64_t *d_cm = ws.get<uint64_t>("w_cm", t.cm.size());
    uint32_t *d_col = ws.get<uint32_t>("w_col", t.P);
    WIDE_CHECK(hipMemcpyAsync(d_init, t.init_masks.data(), per * 8, hipMemcpyHostToDevice, stream));
    WIDE_CHECK(hipMemcpyAsync(d_cm, t.cm.data(), t.cm.size() * 8, hipMemcpyHostToDevice, stream));
    WIDE_CHECK(hipMemcpyAsync(d_col, t.piece_col.data(), t.P * 4, hipMemcpyHostToDevice, stream));
    a.init = d_init;
    a.cm = d_cm;
    a.piece_col = d_col;
    const uint32_t threads = 256;
    uint32_t *d_pmoves = ws.get<uint32_t>("w_pmoves", parents.size());
    uint64_t *d_pmasks = ws.get<uint64_t>("w_pmasks", per * n_parents);
    uint32_t *d_pstatus = ws.get<uint32_t>("w_pstatus", n_parents);
    if (!parents.empty())
        WIDE_CHECK(hipMemcpyAsync(d_pmoves, parents.data(), parents.size() * 4, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(parents_kernel, dim3(n_parents), dim3(threads), t.P, stream, a, d_pmoves, len, d_pmasks,
                       d_pstatus);
    WIDE_CHECK(hipGetLastError());
    out.parent_masks.resize(per * n_parents);
    out.parent_status.resize(n_parents);
    WIDE_CHECK(hipMemcpyAsync(out.parent_masks.data(), d_pmasks, per * n_parents * 8, hipMemcpyDeviceToHost, stream));
    WIDE_CHECK(hipMemcpyAsync(out.parent_status.data(), d_pstatus, n_parents * 4, hipMemcpyDeviceToHost, stream));
    WIDE_CHECK(hipStreamSynchronize(stream));
    // Children: each parent's candidate bits at its cell, in bit order (the
    // ffs walk of the reference), from valid parents and free cells only.
    std::vector<uint32_t> child_parent, child_move;
    for (uint32_t p = 0; p < n_parents; ++p) {
        const uint32_t cell = cells[p];
        if (cell >= t.P) throw std::invalid_argument("cell out of range");
        if (out.parent_status[p] != 0) continue;
        const uint64_t *cm = &out.parent_masks[(size_t)p * per + (size_t)cell * t.nw];
        for (uint32_t w = 0; w < t.nw; ++w)
            for (uint64_t v = cm[w]; v; v &= v - 1) {
                const uint32_t b = w * 64 + (uint32_t)__builtin_ctzll(v);
                child_parent.push_back(p);
                child_move.push_back(cell | (b / 4) << 8 | (b % 4) << 16);
            }
    }
    const uint32_t n_children = (uint32_t)child_parent.size();
    out.child_parent = child_parent;
    out.child_move = child_move;
    out.child_masks.assign(per * n_children, 0);
    out.child_status.assign(n_children, 0);
    if (!n_children) return;
    uint32_t *d_cparent = ws.get<uint32_t>("w_cparent", n_children);
    uint32_t *d_cmove = ws.get<uint32_t>("w_cmove", n_children);
    uint64_t *d_cmasks = ws.get<uint64_t>("w_cmasks", per * n_children);
    uint32_t *d_cstatus = ws.get<uint32_t>("w_cstatus", n_children);
    WIDE_CHECK(hipMemcpyAsync(d_cparent, child_parent.data(), n_children * 4, hipMemcpyHostToDevice, stream));
    WIDE_CHECK(hipMemcpyAsync(d_cmove, child_move.data(), n_children * 4, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(children_kernel, dim3(n_children), dim3(threads), t.P, stream, a, d_pmoves, len, d_pmasks,
                       d_cparent, d_cmove, d_cmasks, d_cstatus);
    WIDE_CHECK(hipGetLastError());
    WIDE_CHECK(hipMemcpyAsync(out.child_masks.data(), d_cmasks, per * n_children * 8, hipMemcpyDeviceToHost, stream));
    WIDE_CHECK(hipMemcpyAsync(out.child_status.data(), d_cstatus, n_children * 4, hipMemcpyDeviceToHost, stream));
    WIDE_CHECK(hipStreamSynchronize(stream));
}

}  // namespace emp
