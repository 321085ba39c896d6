// emp_profile.hip -- capability profiles on gfx950.
//
// The reference's situation_profile feature (include/feature_situation_profile.h
// :150-483) follows the system equation's capability (the backtracker's: one
// info per position and one per piece, each a set of (piece | position,
// orientation) variables, include/situation_capability.h) through the
// placements of a given situation.  A profile is the 2P bit counts of the
// infos, sorted (compute_profile, include/situation_capability.h:145-155).
// It prints the profile before each placement in strategy order, and for
// five criteria (total, max, min, vector_op, rvector_op) a glutton run that
// places the situation's pieces one at a time, each time the one whose
// placed capability is the min (or the max) of the criterion, recording the
// total after each placement (compute_glutton, :388-483).
//
// A capability is stored as its position infos only, 16 u64 per position
// (bit 256 * o + piece, the backtracker's layout); a piece's info is the same
// bits read by piece.  Placing variable v = (position, piece, o) applies v's
// transition (prepare_transitions, include/system_equation_for_CUDA.h:70-134):
// v's position loses every variable, v's piece is cleared everywhere, and at
// each neighbour only the variables whose facing side shows v's colour stay.
//
//   profile_kernel  one block of 256 threads per candidate placement (or
//                   none): the placed capability is built in LDS (32 KB for
//                   P = 256) and written out, its 2P counts are taken (by
//                   position: popcounts; by piece: one thread per piece
//                   reading its bit in every position, the wave's pieces
//                   sharing each word) and sorted (bitonic, in LDS).
// The glutton choice between the candidates' profiles is made on the host
// from the sorted counts; the chosen capability is already in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "emp_backtrack.h"
#include "emp_profile.h"

namespace emp {

namespace {

constexpr uint32_t kCapWords = 16;         // 1024 bits: 256 * o + piece
constexpr uint32_t kProfThreads = 256;
constexpr uint32_t kNoPlacement = 0xFFFFFFFFu;

struct ProfArgs {
    uint32_t P;
    uint32_t sort_n;            // power of two >= 2P, at most 512
    const uint64_t *cap;        // [P][16] capability the candidates apply to
    const uint64_t *cm;         // [(colour * 4 + side) * 16] variables showing colour on side
    const uint32_t *piece_col;  // N | E << 8 | S << 16 | W << 24
    const int32_t *nb;          // [P][4] neighbour strategy index per side, -1 off the board
    const uint32_t *cand;       // [n] position | piece << 8 | o << 16, or kNoPlacement
    uint64_t *cap_out;          // [n][P][16] placed capabilities
    uint16_t *out;              // [n][2P] sorted profiles
};

__global__ void __launch_bounds__(kProfThreads) profile_kernel(ProfArgs a) {
    __shared__ uint64_t cap[256 * kCapWords];
    __shared__ uint16_t val[512];
    const uint32_t tid = threadIdx.x;
    const uint32_t P = a.P, n_words = P * kCapWords;
    const uint32_t cv = a.cand[blockIdx.x];
    const bool place = cv != kNoPlacement;
    const uint32_t pos = cv & 0xFF, piece = (cv >> 8) & 0xFF, o = (cv >> 16) & 3;
    uint64_t *dst = a.cap_out + (size_t)blockIdx.x * n_words;
    for (uint32_t t = tid; t < n_words; t += kProfThreads) {
        const uint32_t j = t / kCapWords, w = t % kCapWords;
        uint64_t v = a.cap[t];
        if (place) {
            if (j == pos) {
                v = 0;
            } else {
                // the piece's four variables (word o' * 4 + piece / 64)
                if ((w & 3) == (piece >> 6)) v &= ~(1ull << (piece & 63));
#pragma unroll
                for (uint32_t s = 0; s < 4; ++s)
                    if (a.nb[pos * 4 + s] == (int32_t)j) {
                        const uint32_t col = (a.piece_col[piece] >> (8 * ((s + o) & 3))) & 0xFF;
                        v &= a.cm[((size_t)col * 4 + ((s + 2) & 3)) * kCapWords + w];
                    }
            }
        }
        cap[t] = v;
        dst[t] = v;
    }
    __syncthreads();
    for (uint32_t i = tid; i < a.sort_n; i += kProfThreads) {
        uint32_t c = 0xFFFF;   // padding sorts last
        if (i < P) {
            c = 0;
            for (uint32_t w = 0; w < kCapWords; ++w) c += (uint32_t)__popcll(cap[i * kCapWords + w]);
        } else if (i < 2 * P) {
            const uint32_t q = i - P, qw = q >> 6, qb = q & 63;
            c = 0;
            for (uint32_t j = 0; j < P; ++j)
#pragma unroll
                for (uint32_t oo = 0; oo < 4; ++oo)
                    c += (uint32_t)(cap[j * kCapWords + oo * 4 + qw] >> qb) & 1u;
        }
        val[i] = (uint16_t)c;
    }
    __syncthreads();
    // bitonic sort of sort_n values, ascending
    for (uint32_t k = 2; k <= a.sort_n; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = tid; i < a.sort_n; i += kProfThreads) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint16_t x = val[i], y = val[ixj];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        val[i] = y;
                        val[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = tid; i < 2 * P; i += kProfThreads) a.out[(size_t)blockIdx.x * 2 * P + i] = val[i];
}

#define PF_CHECK(call)                                                                    \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(err_));       \
    } while (0)

// A sorted profile and the level it is compared at (situation_profile,
// include/situation_profile.h).  Totals are taken in 32-bit arithmetic as the
// reference's are: its start profiles of the extremum search hold 0 or
// 0xFFFFFFFF in every value, whose total wraps.
struct Profile {
    uint32_t level = 0;
    std::vector<uint32_t> v;
    uint32_t total() const {
        uint32_t t = 0;
        for (uint32_t x : v) t += x;
        return t;
    }
    uint32_t max() const { return v.back(); }
    uint32_t min() const { return 2 * level < v.size() ? v[2 * level] : 0u; }
    // the 2 * level infos of the placed positions and pieces are empty; the
    // next one must not be
    bool valid() const { return 2 * level < v.size() ? v[2 * level] != 0 : true; }
};

// less_than_total / _max / _min / _vector (include/situation_profile.h
// :190-233); the reference's rvector_op criterion compares with
// less_than_vector too (src/feature_situation_profile.cpp:75).
bool profile_less(uint32_t criterion, const Profile &a, const Profile &b) {
    switch (criterion) {
        case 0: return a.total() < b.total();
        case 1: return std::make_pair(a.max(), a.total()) < std::make_pair(b.max(), b.total());
        case 2: return std::make_pair(a.min(), a.total()) < std::make_pair(b.min(), b.total());
        default: return a.v < b.v;
    }
}

struct Device {
    hipStream_t stream;
    ProfArgs args;
    uint32_t *d_cand;
    uint64_t *d_caps[2];
    uint16_t *d_out;
    hipEvent_t e0, e1;
    uint64_t launches = 0;
    double kernel_ms = 0;

    // Profiles of the placements `cand` applied to capability `cap`; the
    // placed capabilities go to buffer `slot`.
    std::vector<uint16_t> run(const uint64_t *cap, const std::vector<uint32_t> &cand, int slot) {
        const uint32_t P = args.P;
        PF_CHECK(hipMemcpyAsync(d_cand, cand.data(), cand.size() * 4, hipMemcpyHostToDevice, stream));
        ProfArgs a = args;
        a.cap = cap;
        a.cand = d_cand;
        a.cap_out = d_caps[slot];
        a.out = d_out;
        PF_CHECK(hipEventRecord(e0, stream));
        hipLaunchKernelGGL(profile_kernel, dim3((uint32_t)cand.size()), dim3(kProfThreads), 0, stream, a);
        PF_CHECK(hipGetLastError());
        PF_CHECK(hipEventRecord(e1, stream));
        std::vector<uint16_t> out(cand.size() * 2 * P);
        PF_CHECK(hipMemcpyAsync(out.data(), d_out, out.size() * 2, hipMemcpyDeviceToHost, stream));
        PF_CHECK(hipStreamSynchronize(stream));
        float ms = 0;
        PF_CHECK(hipEventElapsedTime(&ms, e0, e1));
        kernel_ms += ms;
        ++launches;
        return out;
    }
};

}  // namespace

void situation_profile_on_device(hipStream_t stream, Workspace &ws, const Puzzle &pz, const uint16_t *board,
                                 ProfileOutcome &out) {
    const uint32_t P = pz.size(), W = pz.width(), H = pz.height();
    if (P > 256) throw Unsupported("capability profiles hold at most 256 pieces");
    const auto ord = make_order(W, H, kBasic);
    std::vector<uint32_t> index_of(P);
    for (uint32_t i = 0; i < P; ++i) index_of[ord[i].second * W + ord[i].first] = i;
    const std::vector<uint64_t> init = backtrack_capability(pz, kBasic);
    out.n_variables = 0;
    for (uint64_t w : init) out.n_variables += (uint32_t)__builtin_popcountll(w);
    // the situation's placements in strategy order, each a variable
    std::vector<uint32_t> placed, placed_index;
    for (uint32_t i = 0; i < P; ++i) {
        const uint32_t cell = ord[i].second * W + ord[i].first;
        const uint16_t b = board[cell];
        if (b == 0xFFFF) continue;
        const uint32_t piece = b >> 2, o = b & 3;
        const uint32_t bit = 256 * o + piece;
        if (piece >= P || !((init[(size_t)i * kCapWords + bit / 64] >> (bit % 64)) & 1))
            throw std::invalid_argument("the situation places piece " + std::to_string(piece + 1) + " at (" +
                                        std::to_string(ord[i].first) + "," + std::to_string(ord[i].second) +
                                        ") in an orientation that is not a variable there");
        placed.push_back(i | piece << 8 | o << 16);
        placed_index.push_back(i);
    }
    const uint32_t L = (uint32_t)placed.size();
    out.level = L;

    std::vector<uint64_t> cm((size_t)(pz.max_colour() + 1) * 4 * kCapWords, 0);
    std::vector<uint32_t> pcol(P);
    std::vector<int32_t> nb((size_t)P * 4, -1);
    for (uint32_t i = 0; i < P; ++i) {
        const uint32_t x = ord[i].first, y = ord[i].second;
        if (y > 0) nb[i * 4 + 0] = (int32_t)index_of[(y - 1) * W + x];
        if (x + 1 < W) nb[i * 4 + 1] = (int32_t)index_of[y * W + x + 1];
        if (y + 1 < H) nb[i * 4 + 2] = (int32_t)index_of[(y + 1) * W + x];
        if (x > 0) nb[i * 4 + 3] = (int32_t)index_of[y * W + x - 1];
    }
    for (uint32_t q = 0; q < P; ++q) {
        const Piece &p = pz.piece(q);
        pcol[q] = p.colour[0] | p.colour[1] << 8 | p.colour[2] << 16 | p.colour[3] << 24;
        for (unsigned o = 0; o < 4; ++o)
            for (unsigned s = 0; s < 4; ++s) {
                const uint32_t b = 256 * o + q;
                cm[((size_t)shown(p, s, o) * 4 + s) * kCapWords + b / 64] |= 1ull << (b % 64);
            }
    }
    const size_t cap_words = (size_t)P * kCapWords;
    Device dv{};
    dv.stream = stream;
    uint64_t *d_init = ws.get<uint64_t>("pf_init", cap_words);
    uint64_t *d_cm = ws.get<uint64_t>("pf_cm", cm.size());
    uint32_t *d_pcol = ws.get<uint32_t>("pf_pcol", P);
    int32_t *d_nb = ws.get<int32_t>("pf_nb", nb.size());
    dv.d_cand = ws.get<uint32_t>("pf_cand", P);
    dv.d_caps[0] = ws.get<uint64_t>("pf_caps0", (size_t)std::max(P, 1u) * cap_words);
    dv.d_caps[1] = ws.get<uint64_t>("pf_caps1", (size_t)std::max(P, 1u) * cap_words);
    dv.d_out = ws.get<uint16_t>("pf_out", (size_t)P * 2 * P);
    PF_CHECK(hipMemcpyAsync(d_init, init.data(), cap_words * 8, hipMemcpyHostToDevice, stream));
    PF_CHECK(hipMemcpyAsync(d_cm, cm.data(), cm.size() * 8, hipMemcpyHostToDevice, stream));
    PF_CHECK(hipMemcpyAsync(d_pcol, pcol.data(), P * 4, hipMemcpyHostToDevice, stream));
    PF_CHECK(hipMemcpyAsync(d_nb, nb.data(), nb.size() * 4, hipMemcpyHostToDevice, stream));
    dv.args.P = P;
    dv.args.sort_n = 1;
    while (dv.args.sort_n < 2 * P) dv.args.sort_n <<= 1;
    dv.args.cm = d_cm;
    dv.args.piece_col = d_pcol;
    dv.args.nb = d_nb;
    PF_CHECK(hipEventCreate(&dv.e0));
    PF_CHECK(hipEventCreate(&dv.e1));
    struct Events {
        Device &d;
        ~Events() {
            (void)hipEventDestroy(d.e0);
            (void)hipEventDestroy(d.e1);
        }
    } events{dv};

    auto to_profile = [&](const std::vector<uint16_t> &rows, size_t r, uint32_t level) {
        Profile p;
        p.level = level;
        p.v.assign(rows.begin() + r * 2 * P, rows.begin() + (r + 1) * 2 * P);
        return p;
    };

    // glutton runs (compute_glutton): start from the level-0 capability; at
    // each level every piece of the situation not placed yet is tried, in
    // strategy order, and the first invalid one, else the extremum, is
    // placed (a min keeps the earliest of equal profiles, a max the latest)
    out.glutton.assign((size_t)kProfileRuns * (P + 1), 0);
    for (uint32_t run = 0; run < kProfileRuns; ++run) {
        const uint32_t criterion = run / 2;
        const bool want_max = run % 2 == 1;
        uint32_t *series = out.glutton.data() + (size_t)run * (P + 1);
        series[0] = 2 * out.n_variables;
        const uint64_t *cur = d_init;
        int slot = 0;
        std::vector<char> used(L, 0);
        for (uint32_t level = 0; level < L; ++level) {
            std::vector<uint32_t> cand, which;
            for (uint32_t c = 0; c < L; ++c)
                if (!used[c]) {
                    cand.push_back(placed[c]);
                    which.push_back(c);
                }
            const std::vector<uint16_t> rows = dv.run(cur, cand, slot);
            Profile ext;
            ext.level = 0;
            ext.v.assign(2 * P, want_max ? 0u : 0xFFFFFFFFu);
            uint32_t ext_score = 0;
            size_t chosen = cand.size();
            for (size_t r = 0; r < cand.size(); ++r) {
                const Profile p = to_profile(rows, r, level + 1);
                const bool ok = p.valid();
                if (!ok || want_max != profile_less(criterion, p, ext)) {
                    ext = p;
                    ext_score = ok ? p.total() : 0;
                    chosen = r;
                    if (!ok) break;
                }
            }
            if (chosen == cand.size()) throw std::logic_error("no selection at level " + std::to_string(level));
            used[which[chosen]] = 1;
            series[level + 1] = ext_score;
            cur = dv.d_caps[slot] + chosen * cap_words;
            slot ^= 1;
        }
    }

    // the profile before each placement of the situation, then at the end
    // (treat_situation_capability, include/feature_situation_profile.h:290-305)
    out.surface.assign((size_t)(L + 1) * 2 * P, 0);
    out.row_index.assign(L + 1, P);
    {
        const std::vector<uint16_t> row = dv.run(d_init, {kNoPlacement}, 0);
        std::copy(row.begin(), row.end(), out.surface.begin());
        const uint64_t *cur = d_init;
        int slot = 1;
        for (uint32_t k = 0; k < L; ++k) {
            const std::vector<uint16_t> r = dv.run(cur, {placed[k]}, slot);
            std::copy(r.begin(), r.end(), out.surface.begin() + (size_t)(k + 1) * 2 * P);
            out.row_index[k] = placed_index[k];
            cur = dv.d_caps[slot];
            slot ^= 1;
        }
    }
    out.launches = dv.launches;
    out.kernel_ms = dv.kernel_ms;
}

}  // namespace emp
