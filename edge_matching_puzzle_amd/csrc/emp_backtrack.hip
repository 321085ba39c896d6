// emp_backtrack.hip -- the reference's CUDA backtracker search on gfx950.
//
// kernel_and_info (src/CUDA_backtracker.cu:68-178) is a forward-checking
// search for the first solution: at step s it takes the first available
// variable of position s -- capabilities are orientation-major
// (include/CUDA_piece_position_info.h:67-88: bit 256 * o + piece), so the
// lowest orientation first, then the lowest piece -- clears it, and ANDs the
// variable's transition into every later position; an emptied position
// rejects the variable, an exhausted position backtracks one step.
//
// The reference precomputes one transition per variable
// (include/system_equation_for_CUDA.h:70-134: 2P capabilities of 1024 bits
// per variable, 10 GB at 16x16).  Here a transition is applied on the fly:
// clear the variable's piece everywhere, and at the (up to 4) neighbours keep
// only the variables whose facing side shows the same colour -- one
// orientation-major colour mask per (colour, side).  One wave runs one
// search: the 16 u64 words of a position's capability sit in a 16-lane group,
// so a wave checks four later positions per instruction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "emp_backtrack.h"
#include "emp_enumerate.h"

namespace emp {

namespace {

constexpr int kWave = 64;
constexpr uint32_t kWords = 16;   // 1024-bit capability

struct BtArgs {
    uint32_t P, max_colour;
    const uint64_t *init;        // [P][16] level-0 capability per strategy index
    const uint64_t *cm;          // [(colour * 4 + side) * 16]: variables whose side `side` shows colour
    const uint32_t *piece_col;   // [P] N | E << 8 | S << 16 | W << 24
    const int32_t *neighbour;    // [P][4] strategy index of the neighbour on each side, -1 off board
    uint64_t *caps;              // [(P + 1) levels][P][16]
    uint32_t *chosen;            // [P] variable bit chosen at each step
    uint32_t *trace;             // [trace_capacity][3]
    uint64_t trace_capacity;
    unsigned long long *trace_length;
    uint32_t *found;
    uint64_t max_selections;     // 0 = unlimited
};

__global__ void __launch_bounds__(64) backtrack_kernel(BtArgs a) {
    const uint32_t lane = threadIdx.x;
    const uint32_t P = a.P;
    auto cap = [&](uint32_t level, uint32_t pos) { return a.caps + ((size_t)level * P + pos) * kWords; };
    for (uint32_t k = lane; k < P * kWords; k += kWave) a.caps[k] = a.init[k];
    __syncthreads();
    uint32_t step = 0;
    unsigned long long traced = 0;
    bool found = false;
    while (step < P && !(a.max_selections && traced >= a.max_selections)) {
        // First available variable of position `step`.
        uint64_t w = lane < kWords ? cap(step, step)[lane] : 0;
        const uint64_t nz = __ballot(w != 0);
        if (!nz) {
            if (step == 0) break;
            --step;
            continue;
        }
        const uint32_t wl = (uint32_t)__builtin_ctzll(nz);
        const uint64_t word = __shfl(w, wl);
        const uint32_t bit = wl * 64 + (uint32_t)__builtin_ctzll(word);
        const uint32_t o = bit >> 8, piece = bit & 255;
        if (lane == 0) {
            if (traced < a.trace_capacity) {
                a.trace[traced * 3 + 0] = bit >> 5;                 // the reference's lane
                a.trace[traced * 3 + 1] = (step << 10) | bit;       // its variable id
                a.trace[traced * 3 + 2] = step;
            }
            cap(step, step)[wl] = word & ~(1ull << (bit & 63));
            a.chosen[step] = bit;
        }
        ++traced;
        __syncthreads();
        if (step == P - 1) {
            found = true;
            break;
        }
        // Forward check of positions step+1 .. P-1, four per iteration.
        const uint32_t pcol = a.piece_col[piece];
        const int32_t *nb = a.neighbour + (size_t)step * 4;
        bool emptied = false;
        const uint32_t first = step + 1;
        for (uint32_t base = first; base < P; base += kWave / kWords) {
            const uint32_t j = base + lane / kWords, wi = lane % kWords;
            uint64_t v = 0;
            if (j < P) {
                v = cap(step, j)[wi];
                // clear the piece in every orientation
                if ((wi & 3) == (piece >> 6)) v &= ~(1ull << (piece & 63));
                // at a neighbour keep the variables matching the shared edge
                for (uint32_t s = 0; s < 4; ++s)
                    if (nb[s] == (int32_t)j) {
                        const uint32_t col = (pcol >> (8 * ((s + o) & 3))) & 0xFF;
                        v &= a.cm[((size_t)col * 4 + ((s + 2) & 3)) * kWords + wi];
                    }
                cap(step + 1, j)[wi] = v;
            }
            const uint64_t group = __ballot(v != 0);
            for (uint32_t g = 0; g < kWave / kWords; ++g)
                if (base + g < P && ((group >> (g * kWords)) & 0xFFFFull) == 0) emptied = true;
            if (emptied) break;
        }
        __syncthreads();
        if (!emptied) ++step;
    }
    if (lane == 0) {
        *a.trace_length = traced;
        *a.found = found ? 1 : 0;
    }
}

#define BT_CHECK(call)                                                                    \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(err_));       \
    } while (0)

}  // namespace

std::vector<uint64_t> backtrack_capability(const Puzzle &pz, int order) {
    const uint32_t P = pz.size();
    if (P > 256) throw Unsupported("the backtracker's capabilities hold at most 256 pieces");
    const auto ord = make_order(pz.width(), pz.height(), order);
    if (!is_permutation(ord, pz.width(), pz.height()))
        throw std::domain_error("the spiral order is not defined for this board (the walk revisits cells)");
    // Level-0 capability: every kept (piece, o) of the position's kind that
    // matches the frame (create_variables + prepare_initial).
    std::vector<uint64_t> init((size_t)P * kWords, 0);
    for (uint32_t i = 0; i < P; ++i) {
        const uint32_t x = ord[i].first, y = ord[i].second;
        const bool off[4] = {y == 0, x + 1 == pz.width(), y + 1 == pz.height(), x == 0};
        for (uint32_t idx : pz.pieces_of(pz.position_kind(x, y))) {
            const Piece &p = pz.piece(idx);
            for (unsigned o = 0; o < 4; ++o) {
                if (!Puzzle::orientation_kept(p, o)) continue;
                bool ok = true;
                for (unsigned s = 0; s < 4 && ok; ++s) ok = (shown(p, s, o) == 0) == off[s];
                if (!ok) continue;
                const uint32_t b = 256 * o + idx;
                init[(size_t)i * kWords + (b >> 6)] |= 1ull << (b & 63);
            }
        }
    }
    return init;
}

void backtrack_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &pz, int order,
                         BacktrackOutcome &out) {
    (void)dev;
    const uint32_t P = pz.size();
    if (P > 256) throw Unsupported("the backtracker's capabilities hold at most 256 pieces");
    const auto ord = make_order(pz.width(), pz.height(), order);
    if (!is_permutation(ord, pz.width(), pz.height()))
        throw std::domain_error("the spiral order is not defined for this board (the walk revisits cells)");
    std::vector<uint32_t> index_of(P);
    for (uint32_t i = 0; i < P; ++i) index_of[ord[i].second * pz.width() + ord[i].first] = i;
    const std::vector<uint64_t> init = backtrack_capability(pz, order);
    std::vector<uint64_t> cm((size_t)(pz.max_colour() + 1) * 4 * kWords, 0);
    std::vector<uint32_t> pcol(P);
    std::vector<int32_t> nb((size_t)P * 4, -1);
    for (uint32_t i = 0; i < P; ++i) {
        const uint32_t x = ord[i].first, y = ord[i].second;
        if (y > 0) nb[i * 4 + 0] = (int32_t)index_of[(y - 1) * pz.width() + x];
        if (x + 1 < pz.width()) nb[i * 4 + 1] = (int32_t)index_of[y * pz.width() + x + 1];
        if (y + 1 < pz.height()) nb[i * 4 + 2] = (int32_t)index_of[(y + 1) * pz.width() + x];
        if (x > 0) nb[i * 4 + 3] = (int32_t)index_of[y * pz.width() + x - 1];
    }
    for (uint32_t q = 0; q < P; ++q) {
        const Piece &p = pz.piece(q);
        pcol[q] = p.colour[0] | p.colour[1] << 8 | p.colour[2] << 16 | p.colour[3] << 24;
        for (unsigned o = 0; o < 4; ++o)
            for (unsigned s = 0; s < 4; ++s) {
                const uint32_t c = shown(p, s, o), b = 256 * o + q;
                cm[((size_t)c * 4 + s) * kWords + (b >> 6)] |= 1ull << (b & 63);
            }
    }
    const uint64_t cap_words = (uint64_t)(P + 1) * P * kWords;
    const uint64_t trace_cap = out.trace_capacity;
    uint64_t *d_init = ws.get<uint64_t>("bt_init", init.size());
    uint64_t *d_cm = ws.get<uint64_t>("bt_cm", cm.size());
    uint32_t *d_pcol = ws.get<uint32_t>("bt_pcol", P);
    int32_t *d_nb = ws.get<int32_t>("bt_nb", nb.size());
    uint64_t *d_caps = ws.get<uint64_t>("bt_caps", cap_words);
    uint32_t *d_chosen = ws.get<uint32_t>("bt_chosen", P);
    uint32_t *d_trace = ws.get<uint32_t>("bt_trace", std::max<uint64_t>(trace_cap, 1) * 3);
    unsigned long long *d_len = ws.get<unsigned long long>("bt_len", 1);
    uint32_t *d_found = ws.get<uint32_t>("bt_found", 1);
    BT_CHECK(hipMemcpyAsync(d_init, init.data(), init.size() * 8, hipMemcpyHostToDevice, stream));
    BT_CHECK(hipMemcpyAsync(d_cm, cm.data(), cm.size() * 8, hipMemcpyHostToDevice, stream));
    BT_CHECK(hipMemcpyAsync(d_pcol, pcol.data(), P * 4, hipMemcpyHostToDevice, stream));
    BT_CHECK(hipMemcpyAsync(d_nb, nb.data(), nb.size() * 4, hipMemcpyHostToDevice, stream));
    BtArgs a{};
    a.P = P;
    a.max_colour = pz.max_colour();
    a.init = d_init;
    a.cm = d_cm;
    a.piece_col = d_pcol;
    a.neighbour = d_nb;
    a.caps = d_caps;
    a.chosen = d_chosen;
    a.trace = d_trace;
    a.trace_capacity = trace_cap;
    a.trace_length = d_len;
    a.found = d_found;
    a.max_selections = out.max_selections;
    auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(backtrack_kernel, dim3(1), dim3(kWave), 0, stream, a);
    BT_CHECK(hipGetLastError());
    unsigned long long len = 0;
    uint32_t found = 0;
    std::vector<uint32_t> chosen(P);
    BT_CHECK(hipMemcpyAsync(&len, d_len, 8, hipMemcpyDeviceToHost, stream));
    BT_CHECK(hipMemcpyAsync(&found, d_found, 4, hipMemcpyDeviceToHost, stream));
    BT_CHECK(hipMemcpyAsync(chosen.data(), d_chosen, P * 4, hipMemcpyDeviceToHost, stream));
    BT_CHECK(hipStreamSynchronize(stream));
    out.trace.assign((size_t)std::min<uint64_t>(len, trace_cap) * 3, 0);
    if (!out.trace.empty())
        BT_CHECK(hipMemcpyAsync(out.trace.data(), d_trace, out.trace.size() * 4, hipMemcpyDeviceToHost, stream));
    BT_CHECK(hipStreamSynchronize(stream));
    out.elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out.trace_length = len;
    out.found = found != 0;
    out.board.assign(P, 0xFFFF);
    if (out.found)
        for (uint32_t i = 0; i < P; ++i)
            out.board[ord[i].second * pz.width() + ord[i].first] = (uint16_t)(((chosen[i] & 255) << 2) | (chosen[i] >> 8));
}

}  // namespace emp
