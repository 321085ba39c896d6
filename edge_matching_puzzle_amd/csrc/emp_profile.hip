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

// ---------------------------------------------------------------------------
// Profile search (feature_profile, include/feature_profile.h:582-868): level
// by level, every live variable of every situation of a level is placed; a
// child whose profile is valid joins the next level (each situation once)
// and its profile's min, max and total update that level's extrema.
//
//   search_children_kernel  one block per parent: the parent's capability is
//                           built in LDS from the level-0 one (one transition
//                           per placed cell) with its 2P counts; its live
//                           variables are dealt evenly to the threads (a
//                           running count per capability word, searched by
//                           bisection), and each child's counts are the
//                           parent's minus what the placement removes (only
//                           the placed piece's bits, the placed cell and its
//                           neighbours are read); extrema are reduced in LDS,
//                           then one global atomic per block and value; valid
//                           children's boards are appended.
//   search_dedup_kernel     one thread per appended board: an open-addressing
//                           table of board indices (CAS on empty slots, full
//                           board compare on a hit, linear probing) keeps the
//                           first copy of each situation for the next level.
// ---------------------------------------------------------------------------
constexpr uint32_t kSearchThreads = 256;
constexpr uint16_t kEmptyCell = 0xFFFF;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kStatValues = 6;        // low min, low max, high min, high max, total min, total max
constexpr uint32_t kMaxPrecedes = 1024;    // solutions a parent precedes, listed in LDS

struct SearchArgs {
    uint32_t P;
    const uint64_t *cap0;       // [P][16] level-0 capability
    const uint64_t *cm;         // [(colour * 4 + side) * 16]
    const uint32_t *piece_col;  // N | E << 8 | S << 16 | W << 24
    const int32_t *nb;          // [P][4] neighbour cell per side, -1 off the board
    const uint16_t *parents;    // [n][P] row-major boards, ((piece << 2) | o), 0xFFFF empty
    uint32_t level;             // pieces placed in every parent
    uint16_t *cand;             // [cand_cap][P] valid children
    uint64_t cand_cap;
    unsigned long long *n_cand;
    uint32_t *stats;            // [3 series][P + 1 levels][kStatValues]
    const uint16_t *sols;       // second pass: the solutions [n_sol][P]; null in the first
    uint32_t n_sol;
};

// Word w of position j after placing (pos, piece, o) on a capability whose
// word it was: the transition of the system equation's variable
// (prepare_transitions, include/system_equation_for_CUDA.h:70-134): the
// placed position loses every variable, the piece is cleared everywhere, and
// at each neighbour only the variables showing the placed piece's colour on
// the shared edge stay.  nbm[s] is the neighbour on side s (-1: none), nrow[s]
// that side's colour-mask row.
__device__ __forceinline__ uint64_t transition_word(uint64_t v, uint32_t j, uint32_t w, uint32_t pos, uint32_t piece,
                                                   const int32_t (&nbm)[4], const uint64_t *const (&nrow)[4]) {
    if (j == pos) return 0;
    if ((w & 3) == (piece >> 6)) v &= ~(1ull << (piece & 63));
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s)
        if (nbm[s] == (int32_t)j) v &= nrow[s][w];
    return v;
}

__device__ __forceinline__ void placement_rows(const SearchArgs &a, uint32_t pos, uint32_t piece, uint32_t o,
                                               int32_t (&nbm)[4], const uint64_t *(&nrow)[4]) {
#pragma unroll
    for (uint32_t s = 0; s < 4; ++s) {
        nbm[s] = a.nb[pos * 4 + s];
        const uint32_t col = (a.piece_col[piece] >> (8 * ((s + o) & 3))) & 0xFF;
        nrow[s] = a.cm + ((size_t)col * 4 + ((s + 2) & 3)) * kCapWords;
    }
}

__global__ void __launch_bounds__(kSearchThreads) search_children_kernel(SearchArgs a) {
    __shared__ uint64_t cap[256 * kCapWords];
    __shared__ uint16_t board[256];
    __shared__ uint32_t cnt[512];                  // the parent's counts: by position, then by piece
    __shared__ uint32_t wsum[256 * kCapWords + 1]; // live variables before each capability word
    __shared__ uint32_t st[3 * kStatValues];
    __shared__ uint32_t precedes[kMaxPrecedes];
    __shared__ uint32_t n_precedes;
    const uint32_t P = a.P, tid = threadIdx.x, n_words = P * kCapWords, L1 = a.level + 1;
    const uint16_t *pb = a.parents + (size_t)blockIdx.x * P;
    for (uint32_t t = tid; t < P; t += kSearchThreads) board[t] = pb[t];
    for (uint32_t t = tid; t < n_words; t += kSearchThreads) cap[t] = a.cap0[t];
    if (tid < 3 * kStatValues) st[tid] = (tid % 2 == 0) ? 0xFFFFFFFFu : 0u;
    if (tid == 0) n_precedes = 0;
    __syncthreads();
    // the parent's capability: one transition per placed cell
    for (uint32_t cell = 0; cell < P; ++cell) {
        const uint32_t bv = board[cell];
        if (bv == kEmptyCell) continue;
        int32_t nbm[4];
        const uint64_t *nrow[4];
        placement_rows(a, cell, bv >> 2, bv & 3, nbm, nrow);
        for (uint32_t t = tid; t < n_words; t += kSearchThreads)
            cap[t] = transition_word(cap[t], t / kCapWords, t % kCapWords, cell, bv >> 2, nbm, nrow);
        __syncthreads();
    }
    // the parent's 2P counts, and a running count of live variables per word
    for (uint32_t i = tid; i < 2 * P; i += kSearchThreads) {
        uint32_t c = 0;
        if (i < P) {
            for (uint32_t w = 0; w < kCapWords; ++w) c += (uint32_t)__popcll(cap[i * kCapWords + w]);
        } else {
            const uint32_t q = i - P, qw = q >> 6, qb = q & 63;
            for (uint32_t j = 0; j < P; ++j)
#pragma unroll
                for (uint32_t oo = 0; oo < 4; ++oo) c += (uint32_t)(cap[j * kCapWords + oo * 4 + qw] >> qb) & 1u;
        }
        cnt[i] = c;
    }
    if (tid == 0) {
        // exclusive sum over the words (P <= 256: at most 4,096 words)
        uint32_t run = 0;
        for (uint32_t t = 0; t < n_words; ++t) {
            wsum[t] = run;
            run += (uint32_t)__popcll(cap[t]);
        }
        wsum[n_words] = run;
    }
    // second pass: the solutions the parent precedes (its placed cells equal theirs)
    bool scan_all = false;
    if (a.sols) {
        for (uint32_t s2 = tid; s2 < a.n_sol; s2 += kSearchThreads) {
            const uint16_t *sb = a.sols + (size_t)s2 * P;
            bool pre = true;
            for (uint32_t c = 0; c < P && pre; ++c) pre = board[c] == kEmptyCell || board[c] == sb[c];
            if (pre) {
                const uint32_t i = atomicAdd(&n_precedes, 1u);
                if (i < kMaxPrecedes) precedes[i] = s2;
            }
        }
    }
    __syncthreads();
    scan_all = a.sols && n_precedes > kMaxPrecedes;
    // children dealt evenly: variable v of the parent (in word order) to thread v mod 256
    const uint32_t n_vars = wsum[n_words];
    for (uint32_t v = tid; v < n_vars; v += kSearchThreads) {
        // the word holding variable v, then its bit
        uint32_t lo = 0, hi = n_words;   // wsum[lo] <= v < wsum[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (wsum[mid] <= v) lo = mid;
            else hi = mid;
        }
        uint64_t bits = cap[lo];
        for (uint32_t r = v - wsum[lo]; r; --r) bits &= bits - 1;
        const uint32_t pos = lo / kCapWords, w = lo % kCapWords;
        const uint32_t piece = (w & 3) * 64 + (uint32_t)__builtin_ctzll(bits), o = w >> 2;
        const uint32_t pw = piece >> 6, pbit = piece & 63;
        int32_t nbm[4];
        const uint64_t *nrow[4];
        placement_rows(a, pos, piece, o, nbm, nrow);
        // The child's counts from the parent's: its position and piece lose
        // every variable; another position loses the piece's variables, and a
        // neighbour also those the colour constraint removes; another piece
        // loses its variables at the position and those removed at the
        // neighbours.
        uint32_t zeros = 0, mn = 0xFFFFFFFFu, mx = 0, total = 0;
        auto count = [&](uint32_t c) {
            zeros += c == 0;
            mn = c ? min(mn, c) : mn;
            mx = max(mx, c);
            total += c;
        };
        for (uint32_t j = 0; j < P; ++j) {
            uint32_t c = 0;
            bool neighbour = false;
#pragma unroll
            for (uint32_t sd = 0; sd < 4; ++sd) neighbour |= nbm[sd] == (int32_t)j;
            if (j == pos) {
                c = 0;
            } else if (neighbour) {
                for (uint32_t ww = 0; ww < kCapWords; ++ww)
                    c += (uint32_t)__popcll(transition_word(cap[j * kCapWords + ww], j, ww, pos, piece, nbm, nrow));
            } else {
                c = cnt[j];
#pragma unroll
                for (uint32_t oo = 0; oo < 4; ++oo) c -= (uint32_t)(cap[j * kCapWords + oo * 4 + pw] >> pbit) & 1u;
            }
            count(c);
        }
        for (uint32_t q = 0; q < P; ++q) {
            uint32_t c = 0;
            if (q != piece) {
                const uint32_t qw = q >> 6, qb = q & 63;
                c = cnt[P + q];
#pragma unroll
                for (uint32_t oo = 0; oo < 4; ++oo) c -= (uint32_t)(cap[pos * kCapWords + oo * 4 + qw] >> qb) & 1u;
#pragma unroll
                for (uint32_t sd = 0; sd < 4; ++sd) {
                    if (nbm[sd] < 0) continue;
                    const uint32_t n = (uint32_t)nbm[sd];
#pragma unroll
                    for (uint32_t oo = 0; oo < 4; ++oo) {
                        const uint32_t ww = oo * 4 + qw;
                        c -= (uint32_t)((cap[n * kCapWords + ww] & ~nrow[sd][ww]) >> qb) & 1u;
                    }
                }
            }
            count(c);
        }
        // every placed position and piece has no variable left; a valid
        // profile leaves no other empty (include/situation_profile.h:170-178)
        if (L1 < P && zeros != 2 * L1) continue;
        const uint32_t pmin = L1 < P ? mn : 0u;
        const uint32_t cv = piece << 2 | o;
        uint32_t series = 0;
        if (a.sols) {
            bool leads = L1 == P;
            if (scan_all) {
                for (uint32_t s2 = 0; s2 < a.n_sol && !leads; ++s2) {
                    const uint16_t *sb = a.sols + (size_t)s2 * P;
                    bool pre = sb[pos] == cv;
                    for (uint32_t c = 0; c < P && pre; ++c) pre = board[c] == kEmptyCell || board[c] == sb[c];
                    leads = pre;
                }
            } else {
                for (uint32_t i = 0; i < n_precedes && !leads; ++i) leads = a.sols[(size_t)precedes[i] * P + pos] == cv;
            }
            series = leads ? 1u : 2u;
        }
        uint32_t *sv = st + series * kStatValues;
        atomicMin(&sv[0], pmin);
        atomicMax(&sv[1], pmin);
        atomicMin(&sv[2], mx);
        atomicMax(&sv[3], mx);
        atomicMin(&sv[4], total);
        atomicMax(&sv[5], total);
        const unsigned long long slot = atomicAdd(a.n_cand, 1ull);
        if (slot < a.cand_cap) {
            uint16_t *dst = a.cand + slot * P;
            for (uint32_t c = 0; c < P; ++c) dst[c] = c == pos ? (uint16_t)cv : board[c];
        }
    }
    __syncthreads();
    if (tid < 3 * kStatValues) {
        const uint32_t series = tid / kStatValues, k = tid % kStatValues;
        const uint32_t v = st[tid];
        uint32_t *g = a.stats + ((size_t)series * (P + 1) + L1) * kStatValues + k;
        if (k % 2 == 0) {
            if (v != 0xFFFFFFFFu) atomicMin(g, v);
        } else if (st[series * kStatValues + k - 1] != 0xFFFFFFFFu) {
            atomicMax(g, v);
        }
    }
}

__device__ __forceinline__ uint64_t board_hash(const uint16_t *b, uint32_t P) {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    for (uint32_t c = 0; c < P; ++c) {
        h ^= b[c];
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
    }
    return h;
}

__global__ void search_dedup_kernel(const uint16_t *cand, uint64_t n, uint32_t P, uint32_t *table, uint32_t mask,
                                    uint16_t *next, uint64_t next_cap, unsigned long long *n_next) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t *b = cand + i * P;
    uint32_t h = (uint32_t)board_hash(b, P) & mask;
    // at most half the table is used, so an empty slot ends every probe
    while (true) {
        const uint32_t old = atomicCAS(&table[h], kNoSlot, (uint32_t)i);
        if (old == kNoSlot) {
            const unsigned long long o = atomicAdd(n_next, 1ull);
            if (o < next_cap)
                for (uint32_t c = 0; c < P; ++c) next[o * P + c] = b[c];
            return;
        }
        const uint16_t *ob = cand + (uint64_t)old * P;
        bool same = true;
        for (uint32_t c = 0; c < P && same; ++c) same = ob[c] == b[c];
        if (same) return;
        h = (h + 1) & mask;
    }
}

__global__ void fill_words(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
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

// Profile search (feature_profile::template_run, include/feature_profile.h
// :686-868): see the kernels above.  The starts: the empty situation, or at
// 25 pieces the four corners placed -- corner 0 at (0, 0), the other three at
// (W-1, 0), (0, H-1), (W-1, H-1) in the six orders of next_permutation, each
// in its first orientation that is a variable there (the corners of a 5x5
// board share no edge, so a corner's variables there are those of the
// level-0 capability) -- from which the search runs the remaining 21 levels.
// The first pass records the extrema of all situations and takes the last
// level's set of the first start whose search leaves the solutions unset as
// the solutions; the second runs each start that precedes a solution again,
// splitting the extrema between situations that lead to a solution and the
// others.  Every series also holds the level-0 profile.
void profile_search_on_device(hipStream_t stream, Workspace &ws, const Puzzle &pz, uint64_t max_situations,
                              SearchOutcome &out) {
    const uint32_t P = pz.size(), W = pz.width(), H = pz.height();
    if (P > 256) throw Unsupported("capability profiles hold at most 256 pieces");
    if (max_situations == 0 || max_situations > (1ull << 31))
        throw std::invalid_argument("max_situations must be between 1 and 2^31");
    const std::vector<uint64_t> init = backtrack_capability(pz, kBasic);   // basic order: position = cell
    std::vector<uint64_t> cm((size_t)(pz.max_colour() + 1) * 4 * kCapWords, 0);
    std::vector<uint32_t> pcol(P);
    std::vector<int32_t> nb((size_t)P * 4, -1);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            const uint32_t i = y * W + x;
            if (y > 0) nb[i * 4 + 0] = (int32_t)(i - W);
            if (x + 1 < W) nb[i * 4 + 1] = (int32_t)(i + 1);
            if (y + 1 < H) nb[i * 4 + 2] = (int32_t)(i + W);
            if (x > 0) nb[i * 4 + 3] = (int32_t)(i - 1);
        }
    for (uint32_t q = 0; q < P; ++q) {
        const Piece &p = pz.piece(q);
        pcol[q] = p.colour[0] | p.colour[1] << 8 | p.colour[2] << 16 | p.colour[3] << 24;
        for (unsigned o = 0; o < 4; ++o)
            for (unsigned sd = 0; sd < 4; ++sd) {
                const uint32_t b = 256 * o + q;
                cm[((size_t)shown(p, sd, o) * 4 + sd) * kCapWords + b / 64] |= 1ull << (b % 64);
            }
    }
    const size_t cap_words = (size_t)P * kCapWords;
    uint64_t *d_init = ws.get<uint64_t>("ps_init", cap_words);
    uint64_t *d_cm = ws.get<uint64_t>("ps_cm", cm.size());
    uint32_t *d_pcol = ws.get<uint32_t>("ps_pcol", P);
    int32_t *d_nb = ws.get<int32_t>("ps_nb", nb.size());
    const size_t stat_words = (size_t)3 * (P + 1) * kStatValues;
    uint32_t *d_stats = ws.get<uint32_t>("ps_stats", stat_words);
    unsigned long long *d_counts = ws.get<unsigned long long>("ps_counts", 2);
    PF_CHECK(hipMemcpyAsync(d_init, init.data(), cap_words * 8, hipMemcpyHostToDevice, stream));
    PF_CHECK(hipMemcpyAsync(d_cm, cm.data(), cm.size() * 8, hipMemcpyHostToDevice, stream));
    PF_CHECK(hipMemcpyAsync(d_pcol, pcol.data(), P * 4, hipMemcpyHostToDevice, stream));
    PF_CHECK(hipMemcpyAsync(d_nb, nb.data(), nb.size() * 4, hipMemcpyHostToDevice, stream));
    {
        std::vector<uint32_t> st0(stat_words);
        for (size_t i = 0; i < stat_words; ++i) st0[i] = (i % kStatValues) % 2 == 0 ? 0xFFFFFFFFu : 0u;
        PF_CHECK(hipMemcpyAsync(d_stats, st0.data(), stat_words * 4, hipMemcpyHostToDevice, stream));
    }
    hipEvent_t e0, e1;
    PF_CHECK(hipEventCreate(&e0));
    PF_CHECK(hipEventCreate(&e1));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } events{e0, e1};

    // level-0 profile: the counts of the level-0 capability, by position and by piece
    std::vector<uint32_t> prof0(2 * P, 0);
    for (uint32_t j = 0; j < P; ++j)
        for (uint32_t w = 0; w < kCapWords; ++w) {
            const uint64_t v = init[(size_t)j * kCapWords + w];
            prof0[j] += (uint32_t)__builtin_popcountll(v);
            for (uint64_t m = v; m; m &= m - 1) ++prof0[P + (w & 3) * 64 + (uint32_t)__builtin_ctzll(m)];
        }
    uint32_t p0_min = 0xFFFFFFFFu, p0_max = 0, p0_total = 0;
    for (uint32_t c : prof0) {
        p0_min = std::min(p0_min, c);
        p0_max = std::max(p0_max, c);
        p0_total += c;
    }

    // starts
    std::vector<std::vector<uint16_t>> starts;
    uint32_t levels = P;
    const std::vector<uint16_t> empty(P, kEmptyCell);
    auto variable = [&](uint32_t cell, uint32_t piece, uint32_t o) {
        const uint32_t b = 256 * o + piece;
        return (init[(size_t)cell * kCapWords + b / 64] >> (b % 64)) & 1;
    };
    if (P != 25) {
        starts.push_back(empty);
    } else {
        levels = P - 4;
        const std::vector<uint32_t> &corners = pz.pieces_of(kCorner);
        auto set_corner = [&](std::vector<uint16_t> &b, uint32_t piece, uint32_t x, uint32_t y) {
            for (uint32_t o = 0; o < 4; ++o)
                if (variable(y * W + x, piece, o)) {
                    b[y * W + x] = (uint16_t)(piece << 2 | o);
                    break;
                }
        };
        std::vector<uint16_t> c0 = empty;
        set_corner(c0, corners[0], 0, 0);
        const uint32_t cx[3] = {W - 1, 0, W - 1}, cy[3] = {0, H - 1, H - 1};
        std::vector<uint32_t> others{1, 2, 3};
        do {
            std::vector<uint16_t> b = c0;
            for (uint32_t i = 0; i < 3; ++i) set_corner(b, corners[others[i]], cx[i], cy[i]);
            starts.push_back(b);
        } while (std::next_permutation(others.begin(), others.end()));
    }
    out.starts = starts;

    uint64_t cand_cap = 1u << 16, next_cap = std::min<uint64_t>(max_situations, 1u << 16);
    out.level_sizes.assign(P + 1, 0);
    out.launches = 0;
    out.kernel_ms = 0;
    // one breadth search from `start`; returns the last level's set
    auto search = [&](const std::vector<uint16_t> &start, const uint16_t *d_sols, uint32_t n_sol, bool record) {
        uint64_t n_cur = 1;
        uint32_t level = 0;
        for (uint32_t c : start) level += c != kEmptyCell;
        uint16_t *d_parents = ws.get<uint16_t>("ps_parents", (size_t)P);
        PF_CHECK(hipMemcpyAsync(d_parents, start.data(), P * 2, hipMemcpyHostToDevice, stream));
        for (uint32_t l = 0; l < levels && n_cur; ++l, ++level) {
            if (n_cur > 0xFFFFFFFFull) throw Unsupported("a level of the profile search holds more than 2^32 situations");
            SearchArgs a{};
            a.P = P;
            a.cap0 = d_init;
            a.cm = d_cm;
            a.piece_col = d_pcol;
            a.nb = d_nb;
            a.level = level;
            a.stats = d_stats;
            a.sols = d_sols;
            a.n_sol = n_sol;
            a.n_cand = d_counts;
            uint64_t n_cand = 0;
            while (true) {
                a.parents = d_parents;
                a.cand = ws.get<uint16_t>("ps_cand", (size_t)cand_cap * P);
                a.cand_cap = cand_cap;
                PF_CHECK(hipMemsetAsync(d_counts, 0, 16, stream));
                PF_CHECK(hipEventRecord(e0, stream));
                hipLaunchKernelGGL(search_children_kernel, dim3((uint32_t)n_cur), dim3(kSearchThreads), 0, stream, a);
                PF_CHECK(hipGetLastError());
                PF_CHECK(hipEventRecord(e1, stream));
                PF_CHECK(hipMemcpyAsync(&n_cand, d_counts, 8, hipMemcpyDeviceToHost, stream));
                PF_CHECK(hipStreamSynchronize(stream));
                float ms = 0;
                PF_CHECK(hipEventElapsedTime(&ms, e0, e1));
                out.kernel_ms += ms;
                ++out.launches;
                if (n_cand <= cand_cap) break;
                // the children did not fit: the level runs again with room for
                // all of them (its extrema are minima and maxima, which a second
                // run leaves as they are; the parents keep their own buffer)
                if (n_cand > (1ull << 31)) throw Unsupported("the profile search makes more than 2^31 children in a level");
                cand_cap = n_cand;
            }
            // distinct children
            uint32_t tbits = 1;
            while ((1ull << tbits) < 2 * std::max<uint64_t>(n_cand, 1)) ++tbits;
            uint32_t *d_table = ws.get<uint32_t>("ps_table", (size_t)1 << tbits);
            hipLaunchKernelGGL(fill_words, dim3(1024), dim3(256), 0, stream, d_table, (uint64_t)1 << tbits, kNoSlot);
            uint16_t *d_next = ws.get<uint16_t>("ps_next", (size_t)next_cap * P);
            uint64_t n_next = 0;
            while (true) {
                PF_CHECK(hipMemsetAsync(d_counts + 1, 0, 8, stream));
                if (n_cand)
                    hipLaunchKernelGGL(search_dedup_kernel, dim3((uint32_t)((n_cand + 255) / 256)), dim3(256), 0, stream,
                                       a.cand, n_cand, P, d_table, (uint32_t)((1ull << tbits) - 1), d_next, next_cap,
                                       d_counts + 1);
                PF_CHECK(hipGetLastError());
                PF_CHECK(hipMemcpyAsync(&n_next, d_counts + 1, 8, hipMemcpyDeviceToHost, stream));
                PF_CHECK(hipStreamSynchronize(stream));
                ++out.launches;
                if (n_next <= next_cap) break;
                if (n_next > max_situations)
                    throw Unsupported("a level of the profile search holds " + std::to_string(n_next) +
                                      " situations, more than max_situations (" + std::to_string(max_situations) + ")");
                next_cap = n_next;
                d_next = ws.get<uint16_t>("ps_next", (size_t)next_cap * P);
                hipLaunchKernelGGL(fill_words, dim3(1024), dim3(256), 0, stream, d_table, (uint64_t)1 << tbits, kNoSlot);
            }
            if (record) out.level_sizes[level + 1] += n_next;
            // the next level's parents
            d_parents = ws.get<uint16_t>("ps_parents", (size_t)std::max<uint64_t>(n_next, 1) * P);
            PF_CHECK(hipMemcpyAsync(d_parents, d_next, (size_t)n_next * P * 2, hipMemcpyDeviceToDevice, stream));
            n_cur = n_next;
        }
        std::vector<uint16_t> last((size_t)n_cur * P);
        if (n_cur) PF_CHECK(hipMemcpyAsync(last.data(), d_parents, last.size() * 2, hipMemcpyDeviceToHost, stream));
        PF_CHECK(hipStreamSynchronize(stream));
        return std::make_pair(last, n_cur);
    };

    // first pass: every situation; the solutions
    bool have_solutions = false;
    std::vector<uint16_t> sols;
    uint64_t n_sol = 0;
    for (const std::vector<uint16_t> &st : starts) {
        uint32_t lv = 0;
        for (uint16_t c : st) lv += c != kEmptyCell;
        out.level_sizes[lv] += 1;
        auto r = search(st, nullptr, 0, true);
        if (!have_solutions && r.second) {
            sols = std::move(r.first);
            n_sol = r.second;
            have_solutions = true;
        }
    }
    // the first pass wrote series 0 only; the second writes series 1 and 2
    if (n_sol > 0xFFFFFFFFull) throw Unsupported("more than 2^32 solutions");
    if (n_sol) {
        uint16_t *d_sols = ws.get<uint16_t>("ps_sols", (size_t)n_sol * P);
        PF_CHECK(hipMemcpyAsync(d_sols, sols.data(), sols.size() * 2, hipMemcpyHostToDevice, stream));
        for (const std::vector<uint16_t> &st : starts) {
            bool precedes = false;
            uint32_t lv = 0;
            for (uint16_t c : st) lv += c != kEmptyCell;
            for (uint64_t k = 0; k < n_sol && !precedes && lv < P; ++k) {
                bool pre = true;
                for (uint32_t c = 0; c < P && pre; ++c) pre = st[c] == kEmptyCell || st[c] == sols[k * P + c];
                precedes = pre;
            }
            if (precedes) (void)search(st, d_sols, (uint32_t)n_sol, false);
        }
    }
    std::vector<uint32_t> dst(stat_words);
    PF_CHECK(hipMemcpyAsync(dst.data(), d_stats, stat_words * 4, hipMemcpyDeviceToHost, stream));
    PF_CHECK(hipStreamSynchronize(stream));
    // [series][level][value] -> [value][series][level], level 0 from the
    // level-0 profile, minima never updated -> 0 (simple_stat::get_min)
    out.stats.assign((size_t)kStatValues * 3 * (P + 1), 0);
    for (uint32_t series = 0; series < 3; ++series)
        for (uint32_t l = 0; l <= P; ++l)
            for (uint32_t k = 0; k < kStatValues; ++k) {
                uint32_t v = dst[((size_t)series * (P + 1) + l) * kStatValues + k];
                if (l == 0) {
                    const uint32_t p = k < 2 ? p0_min : (k < 4 ? p0_max : p0_total);
                    v = k % 2 == 0 ? std::min(v, p) : std::max(v, p);
                }
                if (k % 2 == 0 && v == 0xFFFFFFFFu) v = 0;
                out.stats[((size_t)k * 3 + series) * (P + 1) + l] = v;
            }
    // solutions, in board order (the reference keeps them in a set)
    std::vector<size_t> order((size_t)n_sol);
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return std::lexicographical_compare(sols.begin() + x * P, sols.begin() + (x + 1) * P, sols.begin() + y * P,
                                            sols.begin() + (y + 1) * P);
    });
    out.solutions.resize((size_t)n_sol * P);
    for (size_t i = 0; i < order.size(); ++i)
        std::copy(sols.begin() + order[i] * P, sols.begin() + (order[i] + 1) * P, out.solutions.begin() + i * P);
    out.n_solutions = n_sol;
}

}  // namespace emp
