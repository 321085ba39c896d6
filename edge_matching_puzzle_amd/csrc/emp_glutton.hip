// emp_glutton.hip -- glutton_max on gfx950: one wave64 per search.
//
// The search is the reference's glutton_max kernel loop
// (include/CUDA_glutton_max.h:1522-1637): at each level remove every
// candidate that would leave a free position or an available piece without
// option (remove_invalid_transitions, :742-954), play the best-scored
// candidate (get_best_candidate, :1007-1301; apply_best_candidate,
// :1303-1476), and pop when a level runs out of candidates.  Masks are
// piece-major (bit 4 * piece + orientation, include/CUDA_piece_position_info2.h)
// with slot bookkeeping identical to include/CUDA_glutton_max_stack.h:434-522,
// so the tie-breaks of the reference's known answers are kept.
//
// Parallel form, identical results:
//   * remove_invalid_transitions reaches the greatest set of candidates that
//     are all valid with respect to each other; validity only shrinks when
//     masks shrink, so that set does not depend on the order of removals.
//     Each lane evaluates candidates against the live masks and clears
//     invalid ones at once (LDS atomic AND); rounds repeat until a round
//     removes nothing.
//   * the score of get_best_candidate for candidate (slot i, piece p, o) with
//     colour-constrained free neighbours j_l and results r_l is
//        total = 2 (PC - T(p) - pc_i + cnt_i(p) - sum_l (pc_jl - cnt_jl(p)))
//              + 2 sum_l |r_l|
//     where pc_j is the popcount of slot j, cnt_j(q) the orientations of piece
//     q in slot j, T(q) = sum_j cnt_j(q) and PC = sum_j pc_j (every mask bit
//     belongs to an available piece).  min / max run over the per-slot
//     popcounts after the move and the per-piece option counts
//     T(q) - cnt_i(q) - sum_l (cnt_jl(q) - cnt_{r_l}(q)).
//     The best candidate is the highest (total, (max << 16) + min), first in
//     (slot, bit) order on ties; the last candidate at the last level.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "emp_enumerate.h"
#include "emp_glutton.h"

namespace emp {

namespace {

constexpr int kWave = 64;

struct GlArgs {
    uint32_t P, W, S, nw, max_colour;
    uint32_t max_mult;           // most sides of one piece showing one (non-frame) colour
    const uint64_t *cm;          // [((c - 1) * 4 + side) * nw + w]
    const uint32_t *piece_col;   // [P] N | E << 8 | S << 16 | W << 24
    const uint64_t *init_masks;  // [P][nw] initial capability per position
    const uint32_t *prefixes;    // [n_items][prefix_len] pos | piece << 8 | o << 16
    uint32_t n_items, prefix_len;
    uint64_t push_budget;
    uint64_t *stacks;            // per wave: S(S+1)/2 * nw words
    uint32_t *max_played;        // per wave: S words
    uint32_t *played;            // per wave: S words, the moves of the current stack
    unsigned int *next_item;
    uint32_t *out_max_depth;     // [n_items]
    uint64_t *out_pushes;        // [n_items]
    uint32_t *out_status;        // [n_items]
    uint16_t *out_board;         // [n_items][P]
    unsigned long long *evals;   // [0] candidates evaluated, [1] of them scored, [2] removal rounds,
                                 // [3] free slots summed over the evaluations
    // Stack checkpoints (layout: glutton_state_words / glutton_mask_words):
    // with ck_in_* each search resumes from its saved stack instead of
    // playing a prefix; with ck_out_* each search saves its stack at the end.
    const uint32_t *ck_in_state;
    const uint64_t *ck_in_masks;
    uint32_t *ck_out_state;
    uint64_t *ck_out_masks;
    uint32_t ck_state_words;
    uint64_t ck_mask_words;
};

// Per-wave LDS layout (offsets in bytes, all 8-byte aligned).
struct GlLayout {
    uint32_t masks, pos_of_slot, slot_of_pos, pc, sp, T, single, crit, risk, pcol, avail, scratch, sets, cnt, snb, lo, hi,
        total;
    __host__ __device__ static GlLayout make(uint32_t S, uint32_t P, uint32_t nw, uint32_t max_colour) {
        GlLayout l;
        uint32_t o = 0;
        auto take = [&](uint32_t bytes) { uint32_t r = o; o += (bytes + 7) & ~7u; return r; };
        l.masks = take(S * nw * 8);
        l.pos_of_slot = take(S * 2);
        l.slot_of_pos = take(P * 2);
        l.pc = take(S * 2);
        l.sp = take(S * 2);
        l.T = take(256 * 2);
        l.single = take(256 * 4);
        l.crit = take(256 * 2);
        l.risk = take(256 * 2);
        l.pcol = take(P * 4);
        l.avail = take(4 * 8);
        l.scratch = take(16);
        l.sets = take(6 * 32 + 8 * 8);
        l.cnt = take(4 * (max_colour + 1) * 2);
        l.snb = take(4 * 4);
        l.lo = take(256 * 4);
        l.hi = take(256 * 4);
        l.total = o;
        return l;
    }
};

__device__ __forceinline__ uint32_t nib(uint64_t v, uint32_t q) {
    return (uint32_t)__popcll((v >> (4 * (q & 15))) & 0xFull);
}

// Orientations o in the 4-bit set `ori` at which a piece with colours qcol
// shows colour col on side (l + 2) & 3 -- the side facing a neighbour on side l.
__device__ __forceinline__ uint32_t facing(uint32_t ori, uint32_t qcol, uint32_t l, uint32_t col) {
    uint32_t n = 0;
#pragma unroll
    for (uint32_t o = 0; o < 4; ++o)
        n += ((ori >> o) & 1) && ((qcol >> (8 * ((l + 2 + o) & 3))) & 0xFF) == col;
    return n;
}

__device__ __forceinline__ uint32_t nibble(uint64_t v, uint32_t q) { return (uint32_t)(v >> (4 * (q & 15))) & 0xF; }

// The 16 pieces of a mask word that have an option in it, as 16 bits.
__device__ __forceinline__ uint64_t word_pieces(uint64_t v) {
    uint64_t x = (v | (v >> 1) | (v >> 2) | (v >> 3)) & 0x1111111111111111ull;
    x = (x | (x >> 3)) & 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    x = (x | (x >> 24)) & 0xFFFFull;
    return x;
}

// Largest facing(ori, qcol, l, col) over colours col: the most options in
// `ori` that one colour toward the slot on side l can leave.
__device__ __forceinline__ uint32_t most_facing(uint32_t ori, uint32_t qcol, uint32_t l) {
    uint32_t m = 0;
#pragma unroll
    for (uint32_t o = 0; o < 4; ++o) {
        const uint32_t f = ((ori >> o) & 1) ? facing(ori, qcol, l, (qcol >> (8 * ((l + 2 + o) & 3))) & 0xFF) : 0;
        m = f > m ? f : m;
    }
    return m;
}

// Colours that piece colours qcol shows toward the slot on side l (its side
// (l + 2 + o) & 3 at orientation o) in byte o, for the orientations in the
// 4-bit set `ori`; 0 for the others (0 is never a candidate's colour there).
__device__ __forceinline__ uint32_t facing_colours(uint32_t ori, uint32_t qcol, uint32_t l) {
    const uint32_t r = 8 * ((l + 2) & 3);
    const uint32_t rot = r ? (qcol >> r) | (qcol << (32 - r)) : qcol;
    return rot & (((ori * 0x00204081u) & 0x01010101u) * 0xFFu);
}

// Bytes of x equal to zero.
__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
    const uint32_t y = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return (uint32_t)__popc(~(y | x | 0x7F7F7F7Fu));
}

// The 64 smallest of the 256 values v[0..3] of the 64 lanes, in ascending
// order: lane r returns the r-th smallest.  Bitonic sort of each column
// across the lanes, then two rounds of merges (the smallest 64 of two sorted
// lists are min(x[r], y[63 - r]), a bitonic sequence, cleaned in 6 steps).
__device__ __forceinline__ uint32_t smallest64(uint32_t (&v)[4], uint32_t lane) {
#pragma unroll
    for (uint32_t k = 2; k <= 64; k <<= 1)
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r) {
                const uint32_t pv = (uint32_t)__shfl_xor((int)v[r], (int)j);
                v[r] = keep_min ? (pv < v[r] ? pv : v[r]) : (pv > v[r] ? pv : v[r]);
            }
        }
    auto merge = [lane](uint32_t x, uint32_t y) {
        const uint32_t yr = (uint32_t)__shfl((int)y, (int)(63 - lane));
        uint32_t c = yr < x ? yr : x;
#pragma unroll
        for (uint32_t j = 32; j > 0; j >>= 1) {
            const uint32_t pv = (uint32_t)__shfl_xor((int)c, (int)j);
            c = (lane & j) == 0 ? (pv < c ? pv : c) : (pv > c ? pv : c);
        }
        return c;
    };
    return merge(merge(v[0], v[1]), merge(v[2], v[3]));
}

struct Wave {
    const GlArgs &a;
    uint64_t *M;
    uint16_t *pos_of_slot, *slot_of_pos, *pc, *sp, *T, *crit;
    uint32_t *single;   // per piece: free slots holding only that piece
    uint32_t *played;   // HBM
    uint32_t *pcol;     // per piece N | E << 8 | S << 16 | W << 24
    uint64_t *avail;
    uint32_t *scratch;
    // 256-bit sets: slots changed in the previous / current removal round,
    // pieces that lost options in the previous / current round, single
    // pieces of the slots changed in the previous round, critical pieces
    uint64_t *ch_prev, *ch_cur, *d_prev, *d_cur, *e_single, *crit_set;
    bool zero_piece;    // an available piece without option (stats())
    // the last move played by apply() (position, piece); push_known when the
    // masks were a fixpoint right before it (a search move, not a prefix move
    // or a pop), so the next removal rounds start from its changes only
    uint32_t push_pos, push_piece;
    bool push_known;
    // [0..3] pieces q with single[q] > 0 (stats()); [4..7] removal slot
    // context, colours below 64: per side, the colours whose neighbour count
    // is at most max_mult (only those can leave the neighbour without
    // option: a candidate holds at most max_mult of them, one per side of its
    // piece showing the colour)
    uint64_t *single_set, *thin;
    uint32_t *lo, *hi;  // per piece q: min / max over free slots j of pc_j - cnt_j(q) (stats())
    uint16_t *cnt;      // slot context: [side l][colour c] options of the side-l neighbour matching c
    uint16_t *risk;     // slot context: critical pieces whose options all lie in the slot and its neighbours
    uint32_t *snb;      // slot context: free neighbour slot on each side (0xFFFFFFFF: none)
    uint64_t *stack;
    uint32_t *max_played;
    uint32_t lane;
    uint32_t level, max, base;
    // slot context for scoring, one key per lane, ascending across lanes:
    // the 64 available pieces q with the smallest lower bound lb(q) on the
    // options q keeps after a move into the slot (key lb << 8 | q), and the
    // 64 with the largest upper bound ub(q) (key (0xFFFF - ub) << 8 | q).
    uint32_t key_lo, key_hi;
    // for the piece of each listed key: the colours it shows toward the slot
    // from each free neighbour (facing_colours), and lb for the upper list
    uint32_t lo_f0, lo_f1, lo_f2, lo_f3, hi_lb, hi_f0, hi_f1, hi_f2, hi_f3;
    // slot context: free neighbour slot of side l in bits 16l..16l+15
    // (0xFFFF: none), that neighbour's option count (pc) and single piece
    // (sp) likewise; the slot's own pc
    uint64_t snbp, pcnp, spnp;
    uint32_t pc_i, sp_i;
    // the current mask word of each free neighbour (0: none)
    uint64_t nbw0, nbw1, nbw2, nbw3;
    // slot context: the first 64 at-risk pieces, one per lane, with the
    // colours they show toward the slot (facing_colours per side)
    uint32_t rk_q, rk_f0, rk_f1, rk_f2, rk_f3;

    __device__ Wave(const GlArgs &args, unsigned char *lds, uint32_t wave_id) : a(args) {
        GlLayout L = GlLayout::make(a.S, a.P, a.nw, a.max_colour);
        M = reinterpret_cast<uint64_t *>(lds + L.masks);
        pos_of_slot = reinterpret_cast<uint16_t *>(lds + L.pos_of_slot);
        slot_of_pos = reinterpret_cast<uint16_t *>(lds + L.slot_of_pos);
        pc = reinterpret_cast<uint16_t *>(lds + L.pc);
        sp = reinterpret_cast<uint16_t *>(lds + L.sp);
        T = reinterpret_cast<uint16_t *>(lds + L.T);
        single = reinterpret_cast<uint32_t *>(lds + L.single);
        crit = reinterpret_cast<uint16_t *>(lds + L.crit);
        pcol = reinterpret_cast<uint32_t *>(lds + L.pcol);
        played = a.played + (size_t)wave_id * a.S;
        avail = reinterpret_cast<uint64_t *>(lds + L.avail);
        scratch = reinterpret_cast<uint32_t *>(lds + L.scratch);
        ch_prev = reinterpret_cast<uint64_t *>(lds + L.sets);
        ch_cur = ch_prev + 4;
        d_prev = ch_prev + 8;
        d_cur = ch_prev + 12;
        e_single = ch_prev + 16;
        crit_set = ch_prev + 20;
        single_set = ch_prev + 24;
        thin = ch_prev + 28;
        lo = reinterpret_cast<uint32_t *>(lds + L.lo);
        cnt = reinterpret_cast<uint16_t *>(lds + L.cnt);
        risk = reinterpret_cast<uint16_t *>(lds + L.risk);
        snb = reinterpret_cast<uint32_t *>(lds + L.snb);
        hi = reinterpret_cast<uint32_t *>(lds + L.hi);
        stack = a.stacks + (size_t)wave_id * ((size_t)a.S * (a.S + 1) / 2) * a.nw;
        max_played = a.max_played + (size_t)wave_id * a.S;
        lane = threadIdx.x;
    }

    // position offset of the neighbour on side l (north, east, south, west)
    __device__ int off(uint32_t l) const {
        return l == 0 ? -(int)a.W : (l == 1 ? 1 : (l == 2 ? (int)a.W : -1));
    }
    __device__ uint32_t n() const { return a.S - level; }
    __device__ uint64_t nbw(uint32_t l) const { return l == 0 ? nbw0 : (l == 1 ? nbw1 : (l == 2 ? nbw2 : nbw3)); }
    __device__ uint32_t field(uint64_t packed, uint32_t l) const { return (uint32_t)(packed >> (16 * l)) & 0xFFFFu; }
    // Lane w holds mask word w of slot i (nw <= 64); the candidate packing
    // reads the words with readlane.  The neighbours' words (nbw0..3) are
    // read per candidate from the masks.
    uint64_t my_mi;
    __device__ void slot_words(uint32_t i) {
        const uint32_t wl = lane < a.nw ? lane : 0;
        my_mi = M[i * a.nw + wl];
    }
    static __device__ uint64_t lane_word(uint64_t v, uint32_t w) {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)w);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)w);
        return (uint64_t)hi << 32 | lo;
    }
    // a.evals: [0] candidate evaluations, [1] of them for scoring, [2]
    // removal rounds, [3] free slots (this level's) summed over evaluations
    __device__ void count(uint64_t evaluated, bool scoring) {
        for (int o = 32; o > 0; o >>= 1) evaluated += __shfl_xor(evaluated, o);
        if (lane == 0 && evaluated) {
            atomicAdd(&a.evals[0], (unsigned long long)evaluated);
            if (scoring) atomicAdd(&a.evals[1], (unsigned long long)evaluated);
            atomicAdd(&a.evals[3], (unsigned long long)(evaluated * n()));
        }
    }
    __device__ bool is_free(uint32_t pos) const { return slot_of_pos[pos] < n(); }
    __device__ bool piece_avail(uint32_t q) const { return (avail[q >> 6] >> (q & 63)) & 1; }
    __device__ size_t level_offset(uint32_t l) const {
        return ((size_t)a.S * (a.S + 1) - (size_t)(a.S - l) * (a.S - l + 1)) / 2 * a.nw;
    }

    // This lane's colour-mask rows for the colour-count entries lane and
    // lane + 64 (entry e: side e / nc, colour e % nc); zero past nw words
    // and for colour 0.
    uint64_t cmr0[16], cmr1[16];
    __device__ void load_colour_rows() {
        const uint32_t nc = a.max_colour + 1;
#pragma unroll
        for (uint32_t r = 0; r < 2; ++r) {
            const uint32_t e = r * kWave + lane, l = e / nc, c = e % nc;
            const bool use = e < 4 * nc && c != 0;
            const uint64_t *cm = a.cm + (use ? ((size_t)(c - 1) * 4 + l) * a.nw : 0);
#pragma unroll
            for (uint32_t w = 0; w < 16; ++w) {
                const uint64_t v = use && w < a.nw ? cm[w] : 0ull;
                if (r == 0) cmr0[w] = v;
                else cmr1[w] = v;
            }
        }
    }

    __device__ void init() {
        for (uint32_t k = lane; k < a.S * a.nw; k += kWave) M[k] = a.init_masks[k];
        for (uint32_t j = lane; j < a.S; j += kWave) {
            pos_of_slot[j] = (uint16_t)j;
            played[j] = 0xFFFFFFFFu;
        }
        for (uint32_t q = lane; q < a.P; q += kWave) {
            slot_of_pos[q] = (uint16_t)q;
            pcol[q] = a.piece_col[q];
        }
        if (lane < 4) {
            uint64_t v = 0;
            for (uint32_t b = 0; b < 64; ++b)
                if (lane * 64 + b < a.P) v |= 1ull << b;
            avail[lane] = v;
        }
        level = max = base = 0;
        push_known = false;
        __syncthreads();
    }

    // Per-slot popcount, single-piece slot marker and per-piece totals;
    // returns false when a slot is empty.  Sets PC.  With `extremes`, also
    // the per-piece min / max over free slots of pc_j - cnt_j(q) and the slot
    // holding it (scoring reads these instead of scanning every slot).
    __device__ bool stats(uint32_t &PC, bool extremes = false) {
        const uint32_t nn = n();
        uint32_t local = 0;
        bool empty = false;
        for (uint32_t q = lane; q < 256; q += kWave) single[q] = 0;
        __syncthreads();
        for (uint32_t j = lane; j < nn; j += kWave) {
            uint32_t cnt = 0, pieces = 0, first = 0xFFFF;
            for (uint32_t w = 0; w < a.nw; ++w) {
                const uint64_t v = M[j * a.nw + w];
                cnt += (uint32_t)__popcll(v);
                const uint64_t nz = (v | (v >> 1) | (v >> 2) | (v >> 3)) & 0x1111111111111111ull;
                if (nz) {
                    if (first == 0xFFFF) first = w * 16 + (uint32_t)__builtin_ctzll(nz) / 4;
                    pieces += (uint32_t)__popcll(nz);
                }
            }
            pc[j] = (uint16_t)cnt;
            sp[j] = pieces == 1 ? (uint16_t)first : (uint16_t)0xFFFF;
            if (pieces == 1) atomicAdd(&single[first], 1u);
            local += cnt;
            empty |= cnt == 0;
        }
        // Per-piece totals, and the list of available pieces with at most 20
        // options: a move removes at most 4 options from each of 5 cells, so
        // only those pieces can be left without option.
        // T: lane (word w, slot group g) sums the per-piece option counts
        // of word w over the slots j = g mod 4, as nibble popcounts in 8-bit
        // fields (even and odd pieces; at most 63 slots before the fields
        // go into 16-bit sums), then the four groups are added across lanes.
        {
            const uint32_t w = lane & 15, g = lane >> 4;
            uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;   // 16-bit sums of pieces 16w + 4h + r
            if (w < a.nw) {
                for (uint32_t j0 = g; j0 < nn; j0 += 4 * 63) {
                    uint64_t ev = 0, od = 0;
                    const uint32_t j1 = j0 + 4 * 63 < nn ? j0 + 4 * 63 : nn;
                    for (uint32_t j = j0; j < j1; j += 4) {
                        const uint64_t v = M[j * a.nw + w];
                        uint64_t c = v - ((v >> 1) & 0x5555555555555555ull);
                        c = (c & 0x3333333333333333ull) + ((c >> 2) & 0x3333333333333333ull);
                        ev += c & 0x0F0F0F0F0F0F0F0Full;
                        od += (c >> 4) & 0x0F0F0F0F0F0F0F0Full;
                    }
                    s0 += ev & 0x00FF00FF00FF00FFull;
                    s2 += (ev >> 8) & 0x00FF00FF00FF00FFull;
                    s1 += od & 0x00FF00FF00FF00FFull;
                    s3 += (od >> 8) & 0x00FF00FF00FF00FFull;
                }
            }
            auto sum_groups = [&](uint64_t v) {
                for (int o = 16; o <= 32; o <<= 1) {
                    const uint32_t lo32 = (uint32_t)__shfl_xor((int)(uint32_t)v, o);
                    const uint32_t hi32 = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), o);
                    v += (uint64_t)hi32 << 32 | lo32;
                }
                return v;
            };
            s0 = sum_groups(s0);
            s1 = sum_groups(s1);
            s2 = sum_groups(s2);
            s3 = sum_groups(s3);
            // lane (w, g) writes pieces 16w + 4g .. 16w + 4g + 3
            if (w < a.nw) {
                const uint32_t q = 16 * w + 4 * g, sh = 16 * g;
                T[q] = (uint16_t)(s0 >> sh);
                T[q + 1] = (uint16_t)(s1 >> sh);
                T[q + 2] = (uint16_t)(s2 >> sh);
                T[q + 3] = (uint16_t)(s3 >> sh);
            }
        }
        __syncthreads();
        uint32_t n_crit = 0;
        bool zero = false;
        for (uint32_t base = 0; base < a.P; base += kWave) {
            const uint32_t q = base + lane;
            const uint32_t t = q < a.P ? (uint32_t)T[q] : 0u;
            const bool c = q < a.P && t <= 20 && piece_avail(q);
            const uint64_t m = __ballot(c);
            if (c) crit[n_crit + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = (uint16_t)q;
            n_crit += (uint32_t)__popcll(m);
            if (lane == 0) crit_set[base / kWave] = m;
            zero |= __ballot(c && t == 0) != 0;
        }
        for (uint32_t base = a.P & ~(uint32_t)(kWave - 1); base < 256; base += kWave)
            if (lane == 0 && base >= a.P) crit_set[base / kWave] = 0;
        zero_piece = zero;
        if (lane == 0) scratch[0] = n_crit;
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            const uint64_t m = __ballot(r * kWave + lane < a.P && single[r * kWave + lane] != 0);
            if (lane == 0) single_set[r] = m;
        }
        for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
        PC = local;
        const bool any_empty = __ballot(empty) != 0;
        __syncthreads();
        if (extremes) {
            // Per piece q: the min of pc_j - cnt_j(q) over free slots j, the
            // slot s holding it, and the min over the free slots farther
            // than 2 steps (Manhattan) from s; likewise for the max.  Packed
            // value | s << 11 | far value << 19 (far min 0x7FF: none).
            for (uint32_t q = lane; q < a.P; q += kWave) {
                uint32_t mnv = 0x7FFu, mns = 0, mxv = 0, mxs = 0;
                const uint32_t w = q >> 4;
                for (uint32_t j = 0; j < nn; ++j) {
                    const uint32_t v = pc[j] - nib(M[j * a.nw + w], q);
                    if (v < mnv) { mnv = v; mns = j; }
                    if (v > mxv || j == 0) { mxv = v; mxs = j; }
                }
                const uint32_t pl = pos_of_slot[mns], ph = pos_of_slot[mxs];
                const int xl = (int)(pl % a.W), yl = (int)(pl / a.W), xh = (int)(ph % a.W), yh = (int)(ph / a.W);
                uint32_t mn2 = 0x7FFu, mx2 = 0;
                for (uint32_t j = 0; j < nn; ++j) {
                    const uint32_t pos = pos_of_slot[j];
                    const int x = (int)(pos % a.W), y = (int)(pos / a.W);
                    const uint32_t v = pc[j] - nib(M[j * a.nw + w], q);
                    if (abs(x - xl) + abs(y - yl) > 2) mn2 = v < mn2 ? v : mn2;
                    if (abs(x - xh) + abs(y - yh) > 2) mx2 = v > mx2 ? v : mx2;
                }
                lo[q] = mnv | mns << 11 | mn2 << 19;
                hi[q] = mxv | mxs << 11 | mx2 << 19;
            }
            __syncthreads();
        }
        return !any_empty;
    }

    // The neighbour part of candidate (slot i, piece p, orientation o).
    struct Cand {
        // constrained neighbours: slot of side l in bits 16l..16l+15 (0xFFFF:
        // none); colour the candidate shows toward it in bits 8l..8l+7, and
        // that colour repeated in the 4 bytes of b[l] (0: none)
        uint64_t nbp;
        uint32_t csp, b0, b1, b2, b3;
        uint32_t rsum, nb_pc_minus;   // sum of r_l; sum of pc_jl - cnt_jl(p)
        uint32_t mn, mx;              // extremes of r_l
        // every free neighbour constrained (the upper bounds of the slot
        // context assume it; a frame colour only ever faces the frame)
        bool bounded;
    };

    // Colour-constrained free neighbours (apply_color_constraints :528-612):
    // the options of the neighbour matching the candidate's side, from the
    // slot context, less those of piece p itself; then the other slots holding
    // only piece p.  False when either leaves a slot without option.
    // Unrolled and branch-free: the four cnt reads go out together.
    __device__ bool neighbours(uint32_t i, uint32_t p, uint32_t o, Cand &cd) {
        const uint32_t pc_p = pcol[p];
        const uint32_t ncol = a.max_colour + 1;
        cd.nbp = ~0ull;
        cd.csp = 0;
        cd.rsum = cd.nb_pc_minus = 0;
        cd.mn = 0xFFFFFFFFu;
        cd.mx = 0;
        cd.bounded = true;
        bool invalid = false;
#pragma unroll
        for (uint32_t l = 0; l < 4; ++l) {
            const uint32_t col = (pc_p >> (8 * ((l + o) & 3))) & 0xFF;
            const uint32_t j = field(snbp, l);
            const bool con = col != 0 && j != 0xFFFFu;
            cd.bounded &= col != 0 || j == 0xFFFFu;
            const uint32_t ori = nibble(nbw(l), p);
            const uint32_t rp = cnt[l * ncol + col] - facing(ori, pc_p, l, col);
            if (con) {
                invalid |= rp == 0;
                cd.nbp &= ~(0xFFFFull << (16 * l)) | ((uint64_t)j << (16 * l));
                cd.csp |= col << (8 * l);
                cd.rsum += rp;
                cd.nb_pc_minus += field(pcnp, l) - (uint32_t)__popc(ori);
                cd.mn = rp < cd.mn ? rp : cd.mn;
                cd.mx = rp > cd.mx ? rp : cd.mx;
            }
        }
        cd.b0 = (cd.csp & 0xFF) * 0x01010101u;
        cd.b1 = ((cd.csp >> 8) & 0xFF) * 0x01010101u;
        cd.b2 = ((cd.csp >> 16) & 0xFF) * 0x01010101u;
        cd.b3 = (cd.csp >> 24) * 0x01010101u;
        if (invalid) return false;
        // Other free slots: invalid when piece p is all one of them holds.
        uint32_t own_single = sp_i == p;
#pragma unroll
        for (uint32_t l = 0; l < 4; ++l) own_single += field(cd.nbp, l) != 0xFFFFu && field(spnp, l) == p;
        return single[p] <= own_single;
    }

    // Options piece q loses by the move: its orientations in slot i and
    // those the colour constraints remove at the neighbours.
    __device__ uint32_t lost(uint32_t i, const Cand &cd, uint32_t q) const {
        const uint32_t w = q >> 4;
        uint32_t d = nib(M[i * a.nw + w], q);
        const uint32_t qc = pcol[q];
        for (uint32_t l = 0; l < 4; ++l) {
            const uint32_t j = field(cd.nbp, l);
            if (j == 0xFFFFu) continue;
            const uint32_t ori = nibble(M[j * a.nw + w], q);
            d += (uint32_t)__popc(ori) - facing(ori, qc, l, (cd.csp >> (8 * l)) & 0xFF);
        }
        return d;
    }

    // With every free neighbour constrained, T(q) - lost(q) = lb(q) + the
    // options of q at the neighbours that show the candidate's colour toward
    // slot i (byte compares against the listed facing colours).
    __device__ static uint32_t matches(const Cand &cd, uint32_t f0, uint32_t f1, uint32_t f2, uint32_t f3) {
        return (cd.b0 ? zero_bytes(f0 ^ cd.b0) : 0u) + (cd.b1 ? zero_bytes(f1 ^ cd.b1) : 0u) +
               (cd.b2 ? zero_bytes(f2 ^ cd.b2) : 0u) + (cd.b3 ? zero_bytes(f3 ^ cd.b3) : 0u);
    }

    // Validity of candidate (slot i, piece p, orientation o) against the live
    // masks (remove_invalid_transitions), for the lanes with `active` set.
    // Called by the whole wave: the walk over the at-risk list reads other
    // lanes' entries (readlane), which is only defined while every lane
    // executes it, so that walk runs with all lanes and the per-lane work
    // around it is predicated.
    __device__ bool valid(uint32_t i, uint32_t p, uint32_t o, bool active) {
        const uint32_t nc = scratch[1];
        bool ok = active, decided = !active;
        if (active && a.max_colour < 64) {
            // Fast form: only the thin colours of a side need the count; a
            // piece without a single-piece slot passes the single check; with
            // no piece at risk and every free neighbour constrained nothing
            // else can fail.
            const uint32_t pc_p = pcol[p];
            const uint32_t ncol = a.max_colour + 1;
            bool invalid = false, bounded = true;
            uint32_t own_single = sp_i == p;
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l) {
                const uint32_t col = (pc_p >> (8 * ((l + o) & 3))) & 0xFF;
                const uint32_t j = field(snbp, l);
                const bool con = col != 0 && j != 0xFFFFu;
                bounded &= col != 0 || j == 0xFFFFu;
                own_single += con && field(spnp, l) == p;
                if (con && ((thin[l] >> col) & 1))
                    invalid |= cnt[l * ncol + col] == facing(nibble(nbw(l), p), pc_p, l, col);
            }
            if (invalid || (((single_set[p >> 6] >> (p & 63)) & 1) && single[p] > own_single)) {
                ok = false;
                decided = true;
            } else if (nc == 0 && bounded) {
                decided = true;
            }
        }
        Cand cd;
        if (!decided && !neighbours(i, p, o, cd)) {
            ok = false;
            decided = true;
        }
        // Only the critical pieces at risk in this slot (lb = 0) can be left
        // without option (slot context); lane k holds the k-th of the first 64.
        bool walking = !decided && cd.bounded;
        const uint32_t listed = nc < kWave ? nc : kWave;
        for (uint32_t k = 0; k < listed; ++k) {
            if (!__ballot(walking)) break;
            const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)rk_q, (int)k);
            const uint32_t f0 = __builtin_amdgcn_readlane((int)rk_f0, (int)k);
            const uint32_t f1 = __builtin_amdgcn_readlane((int)rk_f1, (int)k);
            const uint32_t f2 = __builtin_amdgcn_readlane((int)rk_f2, (int)k);
            const uint32_t f3 = __builtin_amdgcn_readlane((int)rk_f3, (int)k);
            if (walking && q != p && !matches(cd, f0, f1, f2, f3)) {
                ok = false;
                decided = true;
                walking = false;
            }
        }
        if (!decided)
            for (uint32_t k = cd.bounded ? listed : 0; k < nc; ++k) {
                const uint32_t q = risk[k];
                if (q == p) continue;
                if ((uint32_t)T[q] == lost(i, cd, q)) {
                    ok = false;
                    break;
                }
            }
        return ok;
    }

    // Score total of a candidate whose neighbour part is cd (the algebraic
    // form of get_best_candidate's sum, see the file header).
    __device__ uint32_t score_total(uint32_t i, uint32_t p, uint32_t PC, const Cand &cd) const {
        return 2 * (PC - T[p] - pc_i + nib(M[i * a.nw + (p >> 4)], p) - cd.nb_pc_minus) + 2 * cd.rsum;
    }

    // Min / max of the per-slot and per-piece option counts after the move,
    // (max << 16) + min, for the lanes with `active` set; false when the move
    // leaves a piece without option (and for inactive lanes).  Needs the
    // slot's scoring lists (score_lists()).  Called by the whole wave: the
    // walks over the bound lists read other lanes' entries (readlane), so they
    // run with all lanes and the per-lane work in them is predicated.
    __device__ bool score_minmax(uint32_t i, uint32_t p, const Cand &cd, uint32_t &minmax, bool active) {
        const uint32_t nn = n();
        const uint32_t pw = p >> 4;
        uint32_t mn = cd.mn, mx = cd.mx;
        // Free slots other than i and the constrained neighbours: the
        // per-piece extremes, unless they sit in one of those slots.
        auto excluded = [&](uint32_t j) {
            return j == i || j == field(cd.nbp, 0) || j == field(cd.nbp, 1) || j == field(cd.nbp, 2) ||
                   j == field(cd.nbp, 3);
        };
        // When the extreme's slot s is excluded, every excluded slot lies
        // within 2 steps of s: the far value covers the slots beyond, the 13
        // cells around s are visited.
        auto near = [&](uint32_t s, auto fold) {
            const uint32_t ps = pos_of_slot[s];
            const int xs = (int)(ps % a.W), ys = (int)(ps / a.W), H = (int)(a.P / a.W);
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2 + abs(dy); dx <= 2 - abs(dy); ++dx) {
                    const int x = xs + dx, y = ys + dy;
                    if (x < 0 || x >= (int)a.W || y < 0 || y >= H) continue;
                    const uint32_t j = slot_of_pos[y * (int)a.W + x];
                    if (j >= nn || excluded(j)) continue;
                    fold(pc[j] - nib(M[j * a.nw + pw], p));
                }
        };
        if (active && nn > 1) {
            const uint32_t l = lo[p], h = hi[p];
            if (!excluded((l >> 11) & 0xFF)) {
                mn = (l & 0x7FF) < mn ? (l & 0x7FF) : mn;
            } else {
                if ((l >> 19) != 0x7FFu) mn = (l >> 19) < mn ? (l >> 19) : mn;
                near((l >> 11) & 0xFF, [&](uint32_t v) { mn = v < mn ? v : mn; });
            }
            if (!excluded((h >> 11) & 0xFF)) {
                mx = (h & 0x7FF) > mx ? (h & 0x7FF) : mx;
            } else {
                mx = (h >> 19) > mx ? (h >> 19) : mx;
                near((h >> 11) & 0xFF, [&](uint32_t v) { mx = v > mx ? v : mx; });
            }
        }
        // Option counts of the other available pieces after the move: T(q)
        // less what the move takes; their min (0: invalid move) and max.
        // Pieces are visited in the order of the slot context's bounds, until
        // the next bound cannot change the extreme; past the 64 listed pieces
        // (or when a free neighbour is left unconstrained, which the bounds and
        // the listed colours do not cover), every available piece.
        uint32_t pmn = 0xFFFFFFFFu, pmx = 0;
        bool open = active, walking = active && cd.bounded;
        for (uint32_t r = 0; r < kWave; ++r) {
            if (!__ballot(walking)) break;
            const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key_lo, (int)r);
            const uint32_t f0 = __builtin_amdgcn_readlane((int)lo_f0, (int)r);
            const uint32_t f1 = __builtin_amdgcn_readlane((int)lo_f1, (int)r);
            const uint32_t f2 = __builtin_amdgcn_readlane((int)lo_f2, (int)r);
            const uint32_t f3 = __builtin_amdgcn_readlane((int)lo_f3, (int)r);
            if (!walking) continue;
            if (k == 0xFFFFFFFFu || (k >> 8) >= pmn) {
                open = false;
                walking = false;
                continue;
            }
            if ((k & 0xFF) == p) continue;
            const uint32_t c = (k >> 8) + matches(cd, f0, f1, f2, f3);
            pmn = c < pmn ? c : pmn;
        }
        if (open)
            for (uint32_t q = 0; q < a.P; ++q)
                if (q != p && piece_avail(q)) {
                    const uint32_t c = (uint32_t)T[q] - lost(i, cd, q);
                    pmn = c < pmn ? c : pmn;
                }
        const bool ok = active && pmn != 0;
        // The max in steps of 4 listed pieces (independent, so their reads
        // and compares overlap); a piece past the stopping bound cannot raise
        // the max, so taking whole steps changes nothing.
        open = ok && !cd.bounded;
        walking = ok && cd.bounded;
        for (uint32_t r = 0; r < kWave; r += 4) {
            if (!__ballot(walking)) break;
            uint32_t k[4], lb[4], f0[4], f1[4], f2[4], f3[4];
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                k[e] = (uint32_t)__builtin_amdgcn_readlane((int)key_hi, (int)(r + e));
                lb[e] = (uint32_t)__builtin_amdgcn_readlane((int)hi_lb, (int)(r + e));
                f0[e] = __builtin_amdgcn_readlane((int)hi_f0, (int)(r + e));
                f1[e] = __builtin_amdgcn_readlane((int)hi_f1, (int)(r + e));
                f2[e] = __builtin_amdgcn_readlane((int)hi_f2, (int)(r + e));
                f3[e] = __builtin_amdgcn_readlane((int)hi_f3, (int)(r + e));
            }
            if (!walking) continue;
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                if (k[e] == 0xFFFFFFFFu || (k[e] & 0xFF) == p) continue;
                const uint32_t c = lb[e] + matches(cd, f0[e], f1[e], f2[e], f3[e]);
                pmx = c > pmx ? c : pmx;
            }
            if (k[3] == 0xFFFFFFFFu || 0xFFFFu - (k[3] >> 8) <= pmx) walking = false;
            else if (r + 4 == kWave) open = true;
        }
        if (open)
            for (uint32_t q = 0; q < a.P; ++q)
                if (q != p && piece_avail(q)) {
                    const uint32_t c = (uint32_t)T[q] - lost(i, cd, q);
                    pmx = c > pmx ? c : pmx;
                }
        if (!ok) return false;
        mn = pmn < mn ? pmn : mn;
        mx = pmx > mx ? pmx : mx;
        minmax = (mx << 16) + mn;
        return true;
    }

    // For the piece q = key & 0xFF (none: key 0xFFFFFFFF): lb = T(q) less its
    // options in slot i and the free neighbours, and per side the colours q
    // shows toward slot i from that neighbour (facing_colours; 0: none).
    __device__ void payload(uint32_t i, uint32_t key, uint32_t &lb, uint32_t &f0, uint32_t &f1, uint32_t &f2,
                            uint32_t &f3) {
        lb = f0 = f1 = f2 = f3 = 0;
        if (key == 0xFFFFFFFFu) return;
        const uint32_t q = key & 0xFF, w = q >> 4, qc = pcol[q];
        uint32_t touched = nib(M[i * a.nw + w], q);
        auto side = [&](uint32_t l) {
            const uint32_t j = (uint32_t)(snbp >> (16 * l)) & 0xFFFFu;
            if (j == 0xFFFFu) return 0u;
            const uint32_t ori = nibble(M[j * a.nw + w], q);
            touched += (uint32_t)__popc(ori);
            return facing_colours(ori, qc, l);
        };
        f0 = side(0);
        f1 = side(1);
        f2 = side(2);
        f3 = side(3);
        lb = (uint32_t)T[q] - touched;
    }

    // Slot context shared by every candidate of slot i: its free neighbour
    // on each side and, for each side and colour, how many options of that
    // neighbour show the colour on the facing side.
    __device__ void slot_context(uint32_t i, bool scoring = false) {
        const uint32_t nc = a.max_colour + 1;
        if (lane < 4) {
            const uint32_t pos = pos_of_slot[i], x = pos % a.W;
            const bool inside = lane == 0   ? pos >= a.W
                                : lane == 1 ? x + 1 < a.W
                                : lane == 2 ? pos + a.W < a.P
                                            : x > 0;
            const uint32_t q = (uint32_t)((int)pos + off(lane));
            snb[lane] = (inside && is_free(q)) ? (uint32_t)slot_of_pos[q] : 0xFFFFFFFFu;
        }
        __syncthreads();
        snbp = pcnp = spnp = 0;
#pragma unroll
        for (uint32_t l = 0; l < 4; ++l) {
            const uint32_t j = snb[l];
            snbp |= (uint64_t)(j & 0xFFFFu) << (16 * l);
            pcnp |= (uint64_t)(j != 0xFFFFFFFFu ? pc[j] : 0u) << (16 * l);
            spnp |= (uint64_t)(j != 0xFFFFFFFFu ? sp[j] : 0xFFFFu) << (16 * l);
        }
        pc_i = pc[i];
        sp_i = sp[i];
        // cnt[l][c] = popcount(M[j_l] & cm[c][l]).  Entries lane and lane + 64
        // take their colour-mask rows from registers (loaded once per
        // wave); further entries (more than 31 colours) read them, 16 words
        // without a guard (index clamped, extra words masked) so all the
        // global reads of a lane are in flight together.
#pragma unroll
        for (uint32_t r = 0; r < 2; ++r) {
            const uint32_t e = r * kWave + lane;
            if (e < 4 * nc) {
                const uint32_t l = e / nc, c = e % nc, j = snb[l];
                uint32_t v = 0;
                if (c && j != 0xFFFFFFFFu) {
                    const uint64_t *mj = M + j * a.nw;
#pragma unroll
                    for (uint32_t w = 0; w < 16; ++w) {
                        const uint32_t ww = w < a.nw ? w : 0;
                        v += (uint32_t)__popcll((r == 0 ? cmr0[w] : cmr1[w]) & mj[ww]);
                    }
                }
                cnt[e] = (uint16_t)v;
            }
        }
        for (uint32_t e = 2 * kWave + lane; e < 4 * nc; e += kWave) {
            const uint32_t l = e / nc, c = e % nc, j = snb[l];
            uint32_t v = 0;
            if (c && j != 0xFFFFFFFFu) {
                const uint64_t *cm = a.cm + ((size_t)(c - 1) * 4 + l) * a.nw;
                const uint64_t *mj = M + j * a.nw;
                uint64_t x[16];
#pragma unroll
                for (uint32_t w = 0; w < 16; ++w) {
                    const uint32_t ww = w < a.nw ? w : 0;
                    x[w] = cm[ww] & mj[ww];
                }
#pragma unroll
                for (uint32_t w = 0; w < 16; ++w) v += w < a.nw ? (uint32_t)__popcll(x[w]) : 0u;
            }
            cnt[e] = (uint16_t)v;
        }
        if (scoring) {
            __syncthreads();
            return;
        }
        __syncthreads();
        if (a.max_colour < 64) {
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l) {
                const uint64_t m = __ballot(lane < nc && cnt[l * nc + lane] <= a.max_mult);
                if (lane == 0) thin[l] = m;
            }
        }
        // A candidate of slot i can leave a piece without option only if
        // every option of that piece lies in slot i or its free neighbours.
        const uint32_t n_crit = scratch[0];
        uint32_t n_risk = 0;
        for (uint32_t base = 0; base < n_crit; base += kWave) {
            const uint32_t k = base + lane;
            bool at_risk = false;
            uint32_t q = 0;
            if (k < n_crit) {
                q = crit[k];
                const uint32_t w = q >> 4;
                uint32_t here = nib(M[i * a.nw + w], q);
#pragma unroll
                for (uint32_t l = 0; l < 4; ++l)
                    if (snb[l] != 0xFFFFFFFFu) here += nib(M[snb[l] * a.nw + w], q);
                at_risk = here == (uint32_t)T[q];
            }
            const uint64_t m = __ballot(at_risk);
            if (at_risk) risk[n_risk + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = (uint16_t)q;
            n_risk += (uint32_t)__popcll(m);
        }
        if (lane == 0) scratch[1] = n_risk;
        __syncthreads();
        uint32_t lb_unused;
        rk_q = lane < n_risk ? (uint32_t)risk[lane] : 0xFFFFFFFFu;
        payload(i, rk_q, lb_unused, rk_f0, rk_f1, rk_f2, rk_f3);
    }

    // Scoring lists of slot i (after its slot context), for score_minmax().
    // Per available piece q: lb = T(q) less its options in slot i and in the
    // free neighbours (what q keeps after the worst move into slot i); ub =
    // lb plus, per free neighbour, the most options there that one colour
    // toward slot i can leave (what q keeps after the best move).  A move's
    // effect on q lies in [lb, ub].  Wave-wide: every lane calls it.
    __device__ void score_lists(uint32_t i) {
        uint32_t kl[4], kh[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t q = r * 64 + lane;
            kl[r] = kh[r] = 0xFFFFFFFFu;
            if (q < a.P && piece_avail(q)) {
                const uint32_t w = q >> 4, qc = pcol[q];
                uint32_t touched = nib(M[i * a.nw + w], q), keep = 0;
#pragma unroll
                for (uint32_t l = 0; l < 4; ++l) {
                    const uint32_t j = snb[l];
                    if (j == 0xFFFFFFFFu) continue;
                    const uint32_t ori = nibble(M[j * a.nw + w], q);
                    touched += (uint32_t)__popc(ori);
                    keep += most_facing(ori, qc, l);
                }
                const uint32_t lb = (uint32_t)T[q] - touched;
                kl[r] = lb << 8 | q;
                kh[r] = (0xFFFFu - (lb + keep)) << 8 | q;
            }
        }
        key_lo = smallest64(kl, lane);
        key_hi = smallest64(kh, lane);
        // payload of each listed piece
        uint32_t lb_unused;
        payload(i, key_lo, lb_unused, lo_f0, lo_f1, lo_f2, lo_f3);
        payload(i, key_hi, hi_lb, hi_f0, hi_f1, hi_f2, hi_f3);
    }

    // Whether slot i may hold a candidate made invalid by the changes of
    // the previous round (and of this one so far), given that every
    // candidate of slot i was valid when it was last evaluated.  A
    // candidate's validity reads the masks of slot i and its free
    // neighbours (colour counts, single pieces, at-risk pieces), single[]
    // and the totals T of critical pieces; so slot i is affected only when
    // it or a free neighbour lost options, when a critical piece that lost
    // options elsewhere has options there (it may now be at risk), or when
    // a slot changed into holding only a piece that slot i holds.
    __device__ bool affected(uint32_t i) {
        const uint32_t pos = pos_of_slot[i], x = pos % a.W;
        auto changed = [&](uint32_t j) { return ((ch_prev[j >> 6] | ch_cur[j >> 6]) >> (j & 63)) & 1; };
        if (changed(i)) return true;
        uint32_t nb[4];
#pragma unroll
        for (uint32_t l = 0; l < 4; ++l) {
            const bool inside = l == 0 ? pos >= a.W : (l == 1 ? x + 1 < a.W : (l == 2 ? pos + a.W < a.P : x > 0));
            const uint32_t q = (uint32_t)((int)pos + off(l));
            nb[l] = (inside && is_free(q)) ? (uint32_t)slot_of_pos[q] : 0xFFFFFFFFu;
            if (nb[l] != 0xFFFFFFFFu && changed(nb[l])) return true;
        }
        // pieces: lane w takes mask word w (16 pieces); a piece bit spread
        // to its 4 orientation bits
        auto spread = [](uint64_t v16) {
            uint64_t v = v16 & 0xFFFF;
            v = (v | (v << 24)) & 0x000000FF000000FFull;
            v = (v | (v << 12)) & 0x000F000F000F000Full;
            v = (v | (v << 6)) & 0x0303030303030303ull;
            v = (v | (v << 3)) & 0x1111111111111111ull;
            return v * 0xF;
        };
        bool hit = false;
        for (uint32_t w = lane; w < a.nw; w += kWave) {
            const uint32_t sh = 16 * (w & 3);
            const uint64_t dm = spread(((d_prev[w >> 2] & crit_set[w >> 2]) >> sh));
            const uint64_t em = spread(e_single[w >> 2] >> sh);
            const uint64_t own = M[i * a.nw + w];
            uint64_t near = own;
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l)
                if (nb[l] != 0xFFFFFFFFu) near |= M[nb[l] * a.nw + w];
            hit |= (near & dm) != 0 || (own & em) != 0;
        }
        return __ballot(hit) != 0;
    }

    // affected() for slots i0 .. i0 + 3 at once (bit k: slot i0 + k; lanes
    // 16k .. 16k + 15 take the mask words of slot i0 + k; nw <= 16).
    __device__ uint32_t affected4(uint32_t i0) {
        const uint32_t k = lane >> 4, w = lane & 15, i = i0 + k;
        const bool in = i < n();
        auto changed = [&](uint32_t j) { return ((ch_prev[j >> 6] | ch_cur[j >> 6]) >> (j & 63)) & 1; };
        bool hit = false;
        uint32_t nb[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        if (in) {
            const uint32_t pos = pos_of_slot[i], x = pos % a.W;
            hit = changed(i);
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l) {
                const bool inside = l == 0 ? pos >= a.W : (l == 1 ? x + 1 < a.W : (l == 2 ? pos + a.W < a.P : x > 0));
                const uint32_t q = (uint32_t)((int)pos + off(l));
                nb[l] = (inside && is_free(q)) ? (uint32_t)slot_of_pos[q] : 0xFFFFFFFFu;
                hit |= nb[l] != 0xFFFFFFFFu && changed(nb[l]);
            }
        }
        if (in && !hit && w < a.nw) {
            auto spread = [](uint64_t v16) {
                uint64_t v = v16 & 0xFFFF;
                v = (v | (v << 24)) & 0x000000FF000000FFull;
                v = (v | (v << 12)) & 0x000F000F000F000Full;
                v = (v | (v << 6)) & 0x0303030303030303ull;
                v = (v | (v << 3)) & 0x1111111111111111ull;
                return v * 0xF;
            };
            const uint32_t sh = 16 * (w & 3);
            const uint64_t dm = spread(((d_prev[w >> 2] & crit_set[w >> 2]) >> sh));
            const uint64_t em = spread(e_single[w >> 2] >> sh);
            const uint64_t own = M[i * a.nw + w];
            uint64_t near = own;
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l)
                if (nb[l] != 0xFFFFFFFFu) near |= M[nb[l] * a.nw + w];
            hit = (near & dm) != 0 || (own & em) != 0;
        }
        const uint64_t m = __ballot(hit);
        return ((m & 0xFFFFull) != 0) | ((m >> 16 & 0xFFFFull) != 0) << 1 | ((m >> 32 & 0xFFFFull) != 0) << 2 |
               ((m >> 48) != 0) << 3;
    }

    // First removal round after a search move (piece p at position s) from
    // masks that were a fixpoint: a candidate of slot i can only have lost
    // its validity when slot i lies within 2 steps of s (the neighbours of s
    // took its colours, s left the free slots), when a critical piece whose
    // total dropped (it had options at s or at a neighbour of s) has options
    // in slot i or its neighbours, when slot i holds a piece that is the
    // single piece of a slot, or when a free neighbour that lost options of p
    // now counts at most max_mult options showing one of p's colours toward
    // slot i (the only counts p's removal lowered; a candidate holds at most
    // max_mult of them).  The slots that lost options of p are in ch_prev.
    __device__ bool push_affected(uint32_t i) {
        const uint32_t pos = pos_of_slot[i], x = pos % a.W;
        {
            const int dx = (int)x - (int)(push_pos % a.W), dy = (int)(pos / a.W) - (int)(push_pos / a.W);
            if (abs(dx) + abs(dy) <= 2) return true;
        }
        uint64_t nbk = ~0ull;   // free neighbour slot of side l in bits 16l..16l+15
#pragma unroll
        for (uint32_t l = 0; l < 4; ++l) {
            const bool inside = l == 0 ? pos >= a.W : (l == 1 ? x + 1 < a.W : (l == 2 ? pos + a.W < a.P : x > 0));
            const uint32_t q = (uint32_t)((int)pos + off(l));
            if (inside && is_free(q)) nbk &= ~(0xFFFFull << (16 * l)) | ((uint64_t)slot_of_pos[q] << (16 * l));
        }
        auto spread = [](uint64_t v16) {
            uint64_t v = v16 & 0xFFFF;
            v = (v | (v << 24)) & 0x000000FF000000FFull;
            v = (v | (v << 12)) & 0x000F000F000F000Full;
            v = (v | (v << 6)) & 0x0303030303030303ull;
            v = (v | (v << 3)) & 0x1111111111111111ull;
            return v * 0xF;
        };
        bool hit = false;
        for (uint32_t w = lane; w < a.nw; w += kWave) {
            const uint32_t sh = 16 * (w & 3);
            const uint64_t dm = spread((d_prev[w >> 2] & crit_set[w >> 2]) >> sh);
            const uint64_t em = spread(e_single[w >> 2] >> sh);
            const uint64_t own = M[i * a.nw + w];
            uint64_t near = own;
#pragma unroll
            for (uint32_t l = 0; l < 4; ++l) {
                const uint32_t j = field(nbk, l);
                if (j != 0xFFFFu) near |= M[j * a.nw + w];
            }
            hit |= (near & dm) != 0 || (own & em) != 0;
        }
        // lanes: side l, colour index ci of p, word group wg (4 words)
        const uint32_t l = lane >> 4, ci = (lane >> 2) & 3, wg = lane & 3;
        const uint32_t col = (pcol[push_piece] >> (8 * ci)) & 0xFF;
        uint32_t j = field(nbk, l);
        if (j != 0xFFFFu && !((ch_prev[j >> 6] >> (j & 63)) & 1)) j = 0xFFFFu;
        uint32_t v = 0;
        if (j != 0xFFFFu && col)
            for (uint32_t w = 4 * wg; w < 4 * wg + 4 && w < a.nw; ++w)
                v += (uint32_t)__popcll(M[j * a.nw + w] & a.cm[((size_t)(col - 1) * 4 + l) * a.nw + w]);
        v += (uint32_t)__shfl_xor((int)v, 1);
        v += (uint32_t)__shfl_xor((int)v, 2);
        hit |= wg == 0 && j != 0xFFFFu && col && v <= a.max_mult;
        return __ballot(hit) != 0;
    }

    // remove_invalid_transitions: false when a slot runs empty.  Slot by
    // slot; lane b takes bit b of each word of the slot's mask.  The first
    // round evaluates every slot; later rounds only the slots the changes
    // since their last evaluation can affect (affected()), so the round
    // that ends the loop still finds every candidate valid on the final
    // masks.
    __device__ bool fixpoint() {
        uint64_t evaluated = 0;
        bool first = true;
        while (true) {
            uint32_t PC;
            if (!stats(PC)) return false;
            const uint32_t nn = n();
            // round start: the current round's changes become the previous
            // round's; single pieces of the slots they touched
            if (lane < 4) {
                ch_prev[lane] = ch_cur[lane];
                ch_cur[lane] = 0;
                d_prev[lane] = d_cur[lane];
                d_cur[lane] = 0;
                e_single[lane] = 0;
            }
            __syncthreads();
            const bool pushed = first && push_known;
            if (!first || pushed)
                for (uint32_t j = lane; j < nn; j += kWave)
                    if ((pushed || ((ch_prev[j >> 6] >> (j & 63)) & 1)) && sp[j] != 0xFFFF)
                        atomicOr(reinterpret_cast<unsigned long long *>(&e_single[sp[j] >> 6]), 1ull << (sp[j] & 63));
            __syncthreads();
            // a piece already without option makes every candidate invalid
            const bool all = (first && !push_known) || zero_piece;
            bool removed = false;
            // later rounds: affected() four slots at a time, again after a
            // slot that removed candidates (its changes may affect the rest)
            const bool packed = !all && !pushed && a.nw <= 16;
            uint32_t aff = 0, aff_base = 0, aff_n = 0;
            for (uint32_t i = 0; i < nn; ++i) {
                if (packed) {
                    if (i >= aff_base + aff_n) {
                        aff = affected4(i);
                        aff_base = i;
                        aff_n = 4;
                    }
                    if (!((aff >> (i - aff_base)) & 1)) continue;
                } else if (!all && !(pushed ? push_affected(i) : affected(i))) {
                    continue;
                }
                slot_context(i);
                slot_words(i);
                // The slot's candidates in batches of up to 64, whole mask
                // words at a time: bit b of word w goes to list entry fill +
                // (its rank among the word's bits), lane j evaluates entry j.
                // Word by word is the order that counts: a word's candidates
                // see the removals of the words before it (validity compares
                // round-start totals with live masks, so the order changes
                // the result).  A batch therefore keeps its results up to the
                // first word with a removal and starts again after it.  The
                // list lives in the per-piece extremes area, which best()
                // rewrites before reading it.
                for (uint32_t w0 = 0; w0 < a.nw;) {
                    uint32_t fill = 0, w = w0;
                    for (; w < a.nw; ++w) {
                        const uint64_t v = lane_word(my_mi, w);
                        const uint32_t c = (uint32_t)__popcll(v);
                        if (fill + c > kWave) break;
                        if ((v >> lane) & 1) {
                            const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(v >> 32),
                                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)v, 0u));
                            lo[fill + r] = w * 64 + lane;
                        }
                        fill += c;
                    }
                    if (fill == 0) break;   // the remaining words are empty
                    __syncthreads();
                    uint32_t bit = 0;
                    if (lane < fill) {
                        bit = lo[lane];
                        // neighbour words of the candidate's word (the
                        // neighbours' masks do not change while slot i is
                        // evaluated)
                        const uint32_t bw = bit >> 6;
                        uint32_t j = field(snbp, 0);
                        nbw0 = j != 0xFFFFu ? M[j * a.nw + bw] : 0ull;
                        j = field(snbp, 1);
                        nbw1 = j != 0xFFFFu ? M[j * a.nw + bw] : 0ull;
                        j = field(snbp, 2);
                        nbw2 = j != 0xFFFFu ? M[j * a.nw + bw] : 0ull;
                        j = field(snbp, 3);
                        nbw3 = j != 0xFFFFu ? M[j * a.nw + bw] : 0ull;
                    }
                    const bool passes = valid(i, bit >> 2, bit & 3, lane < fill);   // whole wave
                    const bool bad = lane < fill && !passes;
                    __syncthreads();
                    const uint64_t bads = __ballot(bad);
                    // words up to the first with a removal are final
                    const uint32_t last_w = bads ? (uint32_t)__builtin_amdgcn_readlane(
                                                       (int)(bit >> 6), (int)__builtin_ctzll(bads))
                                                 : w - 1;
                    if (lane < fill && (bit >> 6) <= last_w) {
                        ++evaluated;
                        if (bad) {
                            atomicAnd(reinterpret_cast<unsigned long long *>(&M[i * a.nw + (bit >> 6)]),
                                      ~(1ull << (bit & 63)));
                            atomicOr(reinterpret_cast<unsigned long long *>(&ch_cur[i >> 6]), 1ull << (i & 63));
                            const uint32_t pp = bit >> 2;
                            atomicOr(reinterpret_cast<unsigned long long *>(&d_cur[pp >> 6]), 1ull << (pp & 63));
                            removed = true;
                        }
                    }
                    w0 = last_w + 1;
                    __syncthreads();
                }
                __syncthreads();
                if (packed && ((ch_cur[i >> 6] >> (i & 63)) & 1)) aff_n = 0;
            }
            first = false;
            push_known = false;
            if (lane == 0) atomicAdd(&a.evals[2], 1ull);
            if (!__ballot(removed)) break;
        }
        count(evaluated, false);
        uint32_t PC;
        return stats(PC);
    }

    // get_best_candidate: returns (slot << 16) | bit, or 0xFFFFFFFF.  The
    // order is (total, (max << 16) + min), so the min / max (walks and slot
    // extremes, and the slot's scoring lists) are only computed for
    // candidates whose total reaches the best total found so far: a lower
    // total cannot win whatever its min / max.  At the last level, where the
    // last valid candidate wins, every candidate is evaluated in full.
    // A candidate of slot j scores at most 2 (PC - pc_j) (the options of the
    // neighbours and of the piece elsewhere only lower it), so slots whose
    // bound is below the best total so far are skipped, and the 64 slots
    // with the fewest options -- the highest bounds -- go first.  The
    // comparison carries the (slot, bit) tie-break, so the order of the
    // slots does not change the choice.
    __device__ uint32_t best() {
        uint32_t PC;
        stats(PC, true);
        const bool last_level = level == a.S - 1;
        uint64_t best_key = 0;
        uint32_t best_idx = 0xFFFFFFFFu;
        bool have = false;
        uint32_t wave_total = 0;   // best total of the wave after the previous slots
        bool have_any = false;     // some lane holds a valid candidate
        uint64_t evaluated = 0;
        const uint32_t nn = n();
        uint32_t ks[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t j = r * kWave + lane;
            ks[r] = j < nn ? ((uint32_t)pc[j] << 8 | j) : 0xFFFFFFFFu;
        }
        const uint32_t skey = smallest64(ks, lane);
        const uint32_t klast = (uint32_t)__builtin_amdgcn_readlane((int)skey, kWave - 1);
        // steps 0..63: the listed slots; then every other slot in index order
        for (uint32_t step = 0; step < kWave + nn; ++step) {
            uint32_t j;
            if (step < kWave) {
                const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)skey, (int)step);
                if (k == 0xFFFFFFFFu) continue;
                j = k & 0xFF;
            } else {
                j = step - kWave;
                if (klast == 0xFFFFFFFFu || ((uint32_t)pc[j] << 8 | j) <= klast) continue;
            }
            if (!last_level && have_any && 2 * (PC - pc[j]) < wave_total) continue;
            slot_context(j, true);
            slot_words(j);
            bool lists = false;
            // The slot's candidates packed 64 per pass, as in the removal
            // rounds (scoring changes nothing, and each lane keeps its best
            // under the (key, slot, bit) order, so the packing does not change
            // the choice).  The list lives in the at-risk list, which scoring
            // does not use and the next slot context rewrites.
            for (uint32_t w0 = 0; w0 < a.nw;) {
                uint32_t fill = 0, w = w0;
                for (; w < a.nw; ++w) {
                    const uint64_t v = lane_word(my_mi, w);
                    const uint32_t c = (uint32_t)__popcll(v);
                    if (fill + c > kWave) break;
                    if ((v >> lane) & 1) {
                        const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(v >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)v, 0u));
                        risk[fill + r] = (uint16_t)(w * 64 + lane);
                    }
                    fill += c;
                }
                if (fill == 0) break;   // the remaining words are empty
                w0 = w;
                __syncthreads();
                uint32_t bit = 0, p = 0;
                Cand cd;
                uint32_t total = 0;
                bool need = false;
                if (lane < fill) {
                    bit = risk[lane];
                    p = bit >> 2;
                    const uint32_t bw = bit >> 6;
                    uint32_t jn = field(snbp, 0);
                    nbw0 = jn != 0xFFFFu ? M[jn * a.nw + bw] : 0ull;
                    jn = field(snbp, 1);
                    nbw1 = jn != 0xFFFFu ? M[jn * a.nw + bw] : 0ull;
                    jn = field(snbp, 2);
                    nbw2 = jn != 0xFFFFu ? M[jn * a.nw + bw] : 0ull;
                    jn = field(snbp, 3);
                    nbw3 = jn != 0xFFFFu ? M[jn * a.nw + bw] : 0ull;
                    ++evaluated;
                    if (neighbours(j, p, bit & 3, cd)) {
                        total = score_total(j, p, PC, cd);
                        const uint32_t thr = have && ((uint32_t)(best_key >> 32)) > wave_total
                                                 ? (uint32_t)(best_key >> 32) : wave_total;
                        need = last_level || total >= thr;
                    }
                }
                __syncthreads();
                if (!__ballot(need)) continue;
                if (!lists) {
                    score_lists(j);
                    lists = true;
                }
                uint32_t mm = 0;
                const bool scored = score_minmax(j, p, cd, mm, need);
                if (!scored) continue;
                const uint64_t key = last_level ? 0 : (((uint64_t)total << 32) | mm);
                const uint32_t idx = (j << 16) | bit;
                // higher key wins; on equal keys the earlier candidate, except
                // at the last level where the later one wins
                const bool better = !have || key > best_key ||
                                    (key == best_key && (last_level ? idx > best_idx : idx < best_idx));
                if (better) {
                    best_key = key;
                    best_idx = idx;
                    have = true;
                }
            }
            __syncthreads();
            uint32_t t = have ? (uint32_t)(best_key >> 32) : 0u;
            for (int o = 32; o > 0; o >>= 1) {
                const uint32_t ot = (uint32_t)__shfl_xor((int)t, o);
                t = ot > t ? ot : t;
            }
            wave_total = t;
            have_any = __ballot(have) != 0;
        }
        count(evaluated, true);
        for (int o = 32; o > 0; o >>= 1) {
            const uint64_t ok = __shfl_xor(best_key, o);
            const uint32_t oi = __shfl_xor(best_idx, o);
            const bool oh = __shfl_xor((int)have, o) != 0;
            const bool better = oh && (!have || ok > best_key ||
                                       (ok == best_key && (last_level ? oi > best_idx : oi < best_idx)));
            if (better) {
                best_key = ok;
                best_idx = oi;
                have = true;
            }
        }
        return have ? best_idx : 0xFFFFFFFFu;
    }

    // CUDA_glutton_max_stack::push (:434-495), bookkeeping by lane 0.
    __device__ void push(uint32_t slot, uint32_t pos, uint32_t piece, uint32_t o) {
        if (lane == 0) {
            played[level] = (o << 16) | (piece << 8) | pos;
            avail[piece >> 6] &= ~(1ull << (piece & 63));
        }
        __syncthreads();
        const uint32_t nl = level + 1;
        if (nl > max) {
            for (uint32_t k = lane; k < nl; k += kWave) max_played[k] = played[k];
            max = nl;
        }
        if (lane == 0) {
            const uint32_t last = a.S - nl;
            const uint32_t last_pos = pos_of_slot[last];
            pos_of_slot[slot] = (uint16_t)slot;
            const uint16_t t = pos_of_slot[last];
            pos_of_slot[last] = pos_of_slot[slot];
            pos_of_slot[slot] = t;
            const uint16_t u = slot_of_pos[last_pos];
            slot_of_pos[last_pos] = slot_of_pos[pos];
            slot_of_pos[pos] = u;
        }
        level = nl;
        __syncthreads();
    }

    // The stack as CUDA_glutton_stack_XML_converter::dump saves it
    // (include/CUDA_glutton_stack_XML_converter.h:188-202), plus the level the
    // search never pops below: state words [0] size S, [1] pieces P, [2]
    // level, [3] max, [4] base, then info_to_position (pos_of_slot, S),
    // position_to_info (slot_of_pos, P), played (S), max_played (S) and the
    // available pieces (8 words of 32 bits); masks: every level's slot masks
    // at level_offset(l), this level's the live ones, deeper levels zero.
    __device__ void save(uint32_t *st, uint64_t *mk) {
        const uint32_t S = a.S, P = a.P;
        if (lane == 0) {
            st[0] = S;
            st[1] = P;
            st[2] = level;
            st[3] = max;
            st[4] = base;
        }
        // moves past the level (and past max) are not part of the stack
        for (uint32_t j = lane; j < S; j += kWave) {
            st[5 + j] = pos_of_slot[j];
            st[5 + S + P + j] = j < level ? played[j] : 0xFFFFFFFFu;
            st[5 + 2 * S + P + j] = j < max ? max_played[j] : 0xFFFFFFFFu;
        }
        for (uint32_t q = lane; q < P; q += kWave) st[5 + S + q] = slot_of_pos[q];
        if (lane < 8) st[5 + 3 * S + P + lane] = (uint32_t)(avail[lane >> 1] >> (32 * (lane & 1)));
        const size_t here = level_offset(level), live = (size_t)n() * a.nw, total = level_offset(S);
        for (size_t k = lane; k < here; k += kWave) mk[k] = stack[k];
        for (size_t k = lane; k < live; k += kWave) mk[here + k] = M[k];
        for (size_t k = here + live + lane; k < total; k += kWave) mk[k] = 0;
        __syncthreads();
    }

    // Reverse of save() (after init()): the stack is as it was when saved.
    // The masks are not known to be a fixpoint.
    __device__ void restore(const uint32_t *st, const uint64_t *mk) {
        const uint32_t S = a.S, P = a.P;
        level = st[2];
        max = st[3];
        base = st[4];
        for (uint32_t j = lane; j < S; j += kWave) {
            pos_of_slot[j] = (uint16_t)st[5 + j];
            played[j] = st[5 + S + P + j];
            max_played[j] = st[5 + 2 * S + P + j];
        }
        for (uint32_t q = lane; q < P; q += kWave) slot_of_pos[q] = (uint16_t)st[5 + S + q];
        if (lane < 4) avail[lane] = (uint64_t)st[5 + 3 * S + P + 2 * lane] | (uint64_t)st[5 + 3 * S + P + 2 * lane + 1] << 32;
        const size_t here = level_offset(level), live = (size_t)n() * a.nw;
        for (size_t k = lane; k < here; k += kWave) stack[k] = mk[k];
        for (size_t k = lane; k < live; k += kWave) M[k] = mk[here + k];
        push_known = false;
        __syncthreads();
    }

    // CUDA_glutton_max_stack::pop (:497-522), then reload the level's masks.
    __device__ uint32_t pop() {
        uint32_t slot = 0;
        if (lane == 0) {
            const uint32_t last = a.S - level;
            slot = pos_of_slot[last];
            const uint32_t pl = played[level - 1];
            const uint32_t pos = pl & 0xFF, piece = (pl >> 8) & 0xFF;
            avail[piece >> 6] |= 1ull << (piece & 63);
            pos_of_slot[last] = (uint16_t)pos;
            const uint32_t other = pos_of_slot[slot];
            const uint16_t t = pos_of_slot[last];
            pos_of_slot[last] = pos_of_slot[slot];
            pos_of_slot[slot] = t;
            const uint16_t u = slot_of_pos[other];
            slot_of_pos[other] = slot_of_pos[pos];
            slot_of_pos[pos] = u;
        }
        slot = __shfl(slot, 0);
        level -= 1;
        push_known = false;
        __syncthreads();
        const uint64_t *src = stack + level_offset(level);
        for (uint32_t k = lane; k < n() * a.nw; k += kWave) M[k] = src[k];
        __syncthreads();
        return slot;
    }

    // apply_best_candidate (:1303-1476).  False where the reference stops
    // (a position emptied by the move).
    __device__ bool apply(uint32_t slot, uint32_t bit) {
        const uint32_t piece = bit >> 2, o = bit & 3;
        const uint32_t nn = n();
        if (lane == 0) M[slot * a.nw + (bit >> 6)] &= ~(1ull << (bit & 63));
        __syncthreads();
        // Save this level (with the played bit cleared) for a later pop.
        uint64_t *dst = stack + level_offset(level);
        for (uint32_t k = lane; k < nn * a.nw; k += kWave) dst[k] = M[k];
        const uint32_t pos = pos_of_slot[slot];
        // pieces whose totals drop: those with options in the played slot
        for (uint32_t w = lane; w < a.nw; w += kWave) {
            const uint64_t pm = word_pieces(M[slot * a.nw + w]);
            if (pm) atomicOr(reinterpret_cast<unsigned long long *>(&d_cur[w >> 2]), pm << (16 * (w & 3)));
        }
        push_pos = pos;
        push_piece = piece;
        push_known = true;
        const uint32_t pw = piece >> 4;
        const uint64_t keep_p = ~(0xFull << (4 * (piece & 15)));
        bool empty = false;
        if (level + 1 < a.S) {
            // Next level in place: slot `slot` takes the last slot's masks.
            for (uint32_t k = lane; k < (nn - 1) * a.nw; k += kWave) {
                const uint32_t j = k / a.nw, w = k % a.nw;
                uint64_t v = (j == slot) ? M[(nn - 1) * a.nw + w] : M[k];
                if (w == pw) {
                    // slots that lose options of the played piece
                    if (v & ~keep_p) atomicOr(reinterpret_cast<unsigned long long *>(&ch_cur[j >> 6]), 1ull << (j & 63));
                    v &= keep_p;
                }
                M[k] = v;
            }
            __syncthreads();
            for (uint32_t j = lane; j + 1 < nn; j += kWave) {
                uint64_t any = 0;
                for (uint32_t w = 0; w < a.nw; ++w) any |= M[j * a.nw + w];
                empty |= any == 0;
            }
        }
        const bool bad = __ballot(empty) != 0;
        __syncthreads();
        if (bad) return false;
        push(slot, pos, piece, o);
        const uint32_t pcol = a.piece_col[piece];
        bool emptied = false;
#pragma unroll
        for (uint32_t l = 0; l < 4; ++l) {
            const uint32_t col = (pcol >> (8 * ((l + o) & 3))) & 0xFF;
            if (!col) continue;
            const uint32_t q = (uint32_t)((int)pos + off(l));
            if (q >= a.P || !is_free(q)) continue;
            const uint32_t j = slot_of_pos[q];
            const uint64_t *c = a.cm + ((size_t)(col - 1) * 4 + l) * a.nw;
            uint64_t any = 0;
            for (uint32_t w = lane; w < a.nw; w += kWave) {
                const uint64_t old = M[j * a.nw + w];
                const uint64_t v = old & c[w];
                M[j * a.nw + w] = v;
                any |= v;
                // pieces that lose options here
                const uint64_t pm = word_pieces(old & ~v);
                if (pm) atomicOr(reinterpret_cast<unsigned long long *>(&d_cur[w >> 2]), pm << (16 * (w & 3)));
            }
            emptied |= __ballot(any != 0) == 0;
            __syncthreads();
        }
        return !emptied;
    }

    // A choice of best() that can be played: its slot free, its bit one of
    // the slot's options.  Anything else is an internal error, stopped here
    // before apply() indexes with it.
    __device__ bool playable(uint32_t b) const {
        const uint32_t slot = b >> 16, bit = b & 0xFFFF;
        return b != 0xFFFFFFFFu && slot < n() && bit < 4 * a.P && ((M[slot * a.nw + (bit >> 6)] >> (bit & 63)) & 1);
    }

    __device__ bool slot_empty(uint32_t slot) {
        uint64_t any = 0;
        for (uint32_t w = lane; w < a.nw; w += kWave) any |= M[slot * a.nw + w];
        return __ballot(any != 0) == 0;
    }
};

// Status codes of one search.
constexpr uint32_t kSolved = 0, kExhausted = 1, kLimit = 2, kBadPrefix = 3, kInternal = 4;

// One wave per SIMD (the LDS holds four searches per CU), so the whole
// register file is available.
__global__ void __launch_bounds__(64) __attribute__((flatten)) glutton_kernel(GlArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    Wave wv(a, lds, blockIdx.x);
    wv.load_colour_rows();
    while (true) {
        uint32_t item = 0;
        if (wv.lane == 0) item = atomicAdd(a.next_item, 1u);
        item = __shfl(item, 0);
        if (item >= a.n_items) break;
        wv.init();
        uint32_t status = kExhausted;
        bool ok = true, dead = false;
        const bool resumed = a.ck_in_state != nullptr;
        if (resumed)
            wv.restore(a.ck_in_state + (size_t)item * a.ck_state_words, a.ck_in_masks + (size_t)item * a.ck_mask_words);
        for (uint32_t m = 0; !resumed && m < a.prefix_len && ok && !dead; ++m) {
            const uint32_t mv = a.prefixes[(size_t)item * a.prefix_len + m];
            const uint32_t pos = mv & 0xFF, piece = (mv >> 8) & 0xFF, o = (mv >> 16) & 3;
            ok = pos < a.P && piece < a.P && wv.is_free(pos) && wv.piece_avail(piece);
            if (ok) {
                const uint32_t slot = wv.slot_of_pos[pos];
                const uint32_t bit = 4 * piece + o;
                ok = (wv.M[slot * a.nw + (bit >> 6)] >> (bit & 63)) & 1;
                // A playable move that leaves a free cell without option ends
                // the search at the depth reached.
                if (ok) dead = !wv.apply(slot, bit);
            }
        }
        // the masks before the last prefix move were no fixpoint
        wv.push_known = false;
        uint64_t pushes = 0;
        if (!ok) status = kBadPrefix;
        else if (dead) status = kExhausted;
        else {
            if (!resumed) wv.base = wv.level;
            bool new_level = true;
            uint32_t pop_slot = 0;
            status = kSolved;
            while (wv.level < a.S) {
                if (!new_level && wv.slot_empty(pop_slot)) {
                    if (wv.level == wv.base) { status = kExhausted; break; }
                    pop_slot = wv.pop();
                    new_level = false;
                    continue;
                }
                if (!wv.fixpoint()) {
                    if (wv.level == wv.base) { status = kExhausted; break; }
                    pop_slot = wv.pop();
                    new_level = false;
                    continue;
                }
                const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)wv.best());
                if (!wv.playable(b) || !wv.apply(b >> 16, b & 0xFFFF)) { status = kInternal; break; }
                ++pushes;
                new_level = true;
                if (a.push_budget && pushes >= a.push_budget && wv.level < a.S) { status = kLimit; break; }
            }
        }
        // Result: the full board when solved, else the deepest situation.
        const bool full = status == kSolved && wv.level == a.S;
        uint16_t *board = a.out_board + (size_t)item * a.P;
        for (uint32_t q = wv.lane; q < a.P; q += kWave) board[q] = 0xFFFF;
        __syncthreads();
        const uint32_t cnt = full ? a.S : (wv.max < a.S ? wv.max : a.S);
        for (uint32_t k = wv.lane; k < cnt; k += kWave) {
            const uint32_t pl = full ? wv.played[k] : wv.max_played[k];
            if ((pl & 0xFF) < a.P) board[pl & 0xFF] = (uint16_t)((((pl >> 8) & 0xFF) << 2) | ((pl >> 16) & 3));
        }
        if (wv.lane == 0) {
            a.out_max_depth[item] = wv.max;
            a.out_pushes[item] = pushes;
            a.out_status[item] = status;
        }
        __syncthreads();
        if (a.ck_out_state)
            wv.save(a.ck_out_state + (size_t)item * a.ck_state_words, a.ck_out_masks + (size_t)item * a.ck_mask_words);
    }
}

#define GL_CHECK(call)                                                                    \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(err_));       \
    } while (0)

}  // namespace

GluttonTables build_glutton_tables(const Puzzle &pz) {
    GluttonTables t;
    t.P = pz.size();
    t.W = pz.width();
    t.nw = (4 * t.P + 63) / 64;
    t.max_colour = pz.max_colour();
    t.cm.assign((size_t)t.max_colour * 4 * t.nw, 0);
    t.piece_col.assign(t.P, 0);
    t.init_masks.assign((size_t)t.P * t.nw, 0);
    auto set_bit = [&](uint64_t *m, uint32_t piece, uint32_t o) {
        const uint32_t b = 4 * piece + o;
        m[b >> 6] |= 1ull << (b & 63);
    };
    for (uint32_t p = 0; p < t.P; ++p) {
        const Piece &pc = pz.piece(p);
        t.piece_col[p] = pc.colour[0] | pc.colour[1] << 8 | pc.colour[2] << 16 | pc.colour[3] << 24;
        // Colour masks (include/CUDA_common_glutton.h:109-146): entry (c, side)
        // holds (piece, o) whose side opposite to `side` shows c.
        for (uint32_t o = 0; o < 4; ++o)
            for (uint32_t side = 0; side < 4; ++side) {
                const uint32_t c = shown(pc, (side + 2) & 3, o);
                if (c) set_bit(&t.cm[((size_t)(c - 1) * 4 + side) * t.nw], p, o);
            }
    }
    // Initial capability (include/CUDA_common_glutton.h:148-211).
    const uint32_t W = pz.width(), H = pz.height();
    for (uint32_t pos = 0; pos < t.P; ++pos) {
        const uint32_t x = pos % W, y = pos / W;
        uint64_t *m = &t.init_masks[(size_t)pos * t.nw];
        const int kind = pz.position_kind(x, y);
        if (kind == kCorner) {
            // corners (0,0),(W-1,0),(W-1,H-1),(0,H-1): frame sides (N,W),(E,N),(S,E),(W,S)
            const unsigned ci = (x == 0 && y == 0) ? 0 : (y == 0 ? 1 : (x != 0 ? 2 : 3));
            const unsigned b1 = ci, b2 = (ci + 3) & 3;
            for (uint32_t idx : pz.pieces_of(kCorner))
                for (unsigned o = 0; o < 4; ++o)
                    if (shown(pz.piece(idx), b1, o) == 0 && shown(pz.piece(idx), b2, o) == 0) {
                        set_bit(m, idx, o);
                        break;
                    }
        } else if (kind == kBorder) {
            const unsigned side = y == 0 ? 0 : (x == W - 1 ? 1 : (y == H - 1 ? 2 : 3));
            for (uint32_t idx : pz.pieces_of(kBorder)) {
                unsigned zero = 0;
                while (pz.piece(idx).colour[zero]) ++zero;
                set_bit(m, idx, (zero + 4 - side) & 3);
            }
        } else {
            for (uint32_t idx : pz.pieces_of(kCenter))
                for (unsigned o = 0; o < 4; ++o) set_bit(m, idx, o);
        }
    }
    return t;
}

uint32_t glutton_state_words(uint32_t P) { return 13 + 4 * P; }

uint64_t glutton_mask_words(uint32_t P) { return (uint64_t)P * (P + 1) / 2 * ((4 * P + 63) / 64); }

void validate_glutton_stack(const Puzzle &pz, const uint32_t *st, const uint64_t *mk) {
    const uint32_t P = pz.size(), S = P, nw = (4 * P + 63) / 64;
    auto bad = [](const std::string &what) { throw std::invalid_argument("glutton stack: " + what); };
    if (st[0] != S) bad("size " + std::to_string(st[0]) + " is not the puzzle's " + std::to_string(S));
    if (st[1] != P) bad("nb_pieces " + std::to_string(st[1]) + " is not the puzzle's " + std::to_string(P));
    const uint32_t level = st[2], max = st[3], base = st[4];
    if (!(base <= level && level <= max && max <= S)) bad("base <= level <= max <= size does not hold");
    const uint32_t *i2p = st + 5, *p2i = st + 5 + S, *played = st + 5 + S + P, *maxp = st + 5 + 2 * S + P;
    const uint32_t *av = st + 5 + 3 * S + P;
    const uint32_t n = S - level;
    for (uint32_t j = 0; j < S; ++j)
        if (i2p[j] >= P) bad("info_to_position[" + std::to_string(j) + "] out of range");
    for (uint32_t q = 0; q < P; ++q)
        if (p2i[q] >= S) bad("position_to_info[" + std::to_string(q) + "] out of range");
    for (uint32_t j = 0; j < n; ++j)
        if (p2i[i2p[j]] != j) bad("free info " + std::to_string(j) + " and its position disagree");
    uint32_t free_positions = 0, n_avail = 0;
    for (uint32_t q = 0; q < P; ++q) free_positions += p2i[q] < n;
    if (free_positions != n) bad("free positions and level disagree");
    for (uint32_t w = 0; w < 8; ++w)
        for (uint32_t b = 0; b < 32; ++b)
            if ((av[w] >> b) & 1) {
                if (32 * w + b >= P) bad("available piece past the last piece");
                ++n_avail;
            }
    if (n_avail != n) bad("available pieces and level disagree");
    auto check_move = [&](uint32_t mv, const std::string &what, bool placed) {
        const uint32_t pos = mv & 0xFF, piece = (mv >> 8) & 0xFF;
        if ((mv >> 18) || pos >= P || piece >= P) bad(what + " is not a move");
        if (placed && (p2i[pos] < n || ((av[piece / 32] >> (piece % 32)) & 1)))
            bad(what + " is not on the board");
    };
    for (uint32_t l = 0; l < level; ++l) check_move(played[l], "played_info[" + std::to_string(l) + "]", true);
    for (uint32_t l = 0; l < max; ++l) check_move(maxp[l], "max_played_info[" + std::to_string(l) + "]", false);
    const uint64_t last = (4 * P) % 64 ? ~0ull << ((4 * P) % 64) : 0ull;
    for (uint32_t l = 0; l <= level && l < S; ++l) {
        const uint64_t off = ((uint64_t)S * (S + 1) - (uint64_t)(S - l) * (S - l + 1)) / 2 * nw;
        for (uint32_t j = 0; j < S - l; ++j)
            if (mk[off + (uint64_t)j * nw + nw - 1] & last) bad("mask bit past the last piece");
    }
}

void glutton_on_device(const DeviceInfo &dev, hipStream_t stream, Workspace &ws, const Puzzle &pz,
                       const std::vector<uint32_t> &prefixes, uint32_t prefix_len, uint32_t n_items,
                       const GluttonConfig &cfg, std::vector<GluttonOutcome> &out, GluttonStats &stats,
                       const GluttonCheckpoints &ck) {
    if (pz.size() > 256) throw Unsupported("glutton_max handles at most 256 pieces");
    if ((uint64_t)prefixes.size() != (uint64_t)prefix_len * n_items)
        throw std::invalid_argument("prefix array size does not match n_items * prefix_len");
    if ((ck.in_state == nullptr) != (ck.in_masks == nullptr) || (ck.out_state == nullptr) != (ck.out_masks == nullptr))
        throw std::invalid_argument("a checkpoint needs both its state and its masks");
    if (ck.in_state && prefix_len) throw std::invalid_argument("resumed searches play no prefix");
    const uint32_t ck_state = glutton_state_words(pz.size());
    const uint64_t ck_masks = glutton_mask_words(pz.size());
    if (ck.in_state)
        for (uint32_t i = 0; i < n_items; ++i)
            validate_glutton_stack(pz, ck.in_state + (size_t)i * ck_state, ck.in_masks + (size_t)i * ck_masks);
    auto wall0 = std::chrono::steady_clock::now();
    GluttonTables t = build_glutton_tables(pz);
    const uint32_t S = t.P;
    const GlLayout L = GlLayout::make(S, t.P, t.nw, t.max_colour);
    if (L.total > dev.lds_per_block) throw Unsupported("glutton LDS layout exceeds the device limit");
    const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu
                                                   : std::max<uint32_t>(1, std::min<uint32_t>(8, (uint32_t)(160 * 1024 / L.total)));
    const uint32_t waves = std::max<uint32_t>(1, std::min<uint32_t>(dev.compute_units * waves_per_cu, std::max<uint32_t>(n_items, 1)));
    const size_t tri = (size_t)S * (S + 1) / 2 * t.nw;

    uint64_t *d_cm = ws.get<uint64_t>("gl_cm", t.cm.size());
    uint32_t *d_pcol = ws.get<uint32_t>("gl_pcol", t.piece_col.size());
    uint64_t *d_init = ws.get<uint64_t>("gl_init", t.init_masks.size());
    uint32_t *d_pre = ws.get<uint32_t>("gl_prefixes", std::max<size_t>(prefixes.size(), 1));
    uint64_t *d_stacks = ws.get<uint64_t>("gl_stacks", tri * waves);
    uint32_t *d_maxp = ws.get<uint32_t>("gl_max_played", (size_t)S * waves);
    uint32_t *d_played = ws.get<uint32_t>("gl_played", (size_t)S * waves);
    unsigned int *d_next = ws.get<unsigned int>("gl_next", 2);
    uint32_t *d_depth = ws.get<uint32_t>("gl_depth", n_items);
    uint64_t *d_pushes = ws.get<uint64_t>("gl_pushes", n_items);
    uint32_t *d_status = ws.get<uint32_t>("gl_status", n_items);
    uint16_t *d_board = ws.get<uint16_t>("gl_board", (size_t)n_items * t.P);
    unsigned long long *d_evals = ws.get<unsigned long long>("gl_evals", 4);
    const bool any_ck = ck.in_state || ck.out_state;
    uint32_t *d_ck_state = any_ck ? ws.get<uint32_t>("gl_ck_state", (size_t)n_items * ck_state) : nullptr;
    uint64_t *d_ck_masks = any_ck ? ws.get<uint64_t>("gl_ck_masks", (size_t)n_items * ck_masks) : nullptr;
    if (ck.in_state) {
        GL_CHECK(hipMemcpyAsync(d_ck_state, ck.in_state, (size_t)n_items * ck_state * 4, hipMemcpyHostToDevice, stream));
        GL_CHECK(hipMemcpyAsync(d_ck_masks, ck.in_masks, (size_t)n_items * ck_masks * 8, hipMemcpyHostToDevice, stream));
    }
    GL_CHECK(hipMemcpyAsync(d_cm, t.cm.data(), t.cm.size() * 8, hipMemcpyHostToDevice, stream));
    GL_CHECK(hipMemcpyAsync(d_pcol, t.piece_col.data(), t.piece_col.size() * 4, hipMemcpyHostToDevice, stream));
    GL_CHECK(hipMemcpyAsync(d_init, t.init_masks.data(), t.init_masks.size() * 8, hipMemcpyHostToDevice, stream));
    if (!prefixes.empty())
        GL_CHECK(hipMemcpyAsync(d_pre, prefixes.data(), prefixes.size() * 4, hipMemcpyHostToDevice, stream));
    const unsigned int zero2[2] = {0, 0};
    const unsigned long long zero64[4] = {0, 0, 0, 0};
    GL_CHECK(hipMemcpyAsync(d_next, zero2, sizeof(zero2), hipMemcpyHostToDevice, stream));
    GL_CHECK(hipMemcpyAsync(d_evals, zero64, sizeof(zero64), hipMemcpyHostToDevice, stream));

    GlArgs a{};
    a.P = t.P;
    a.W = t.W;
    a.S = S;
    a.nw = t.nw;
    a.max_colour = t.max_colour;
    a.max_mult = 1;
    for (uint32_t q = 0; q < t.P; ++q)
        for (uint32_t s1 = 0; s1 < 4; ++s1) {
            const uint32_t c1 = (t.piece_col[q] >> (8 * s1)) & 0xFF;
            uint32_t m = 0;
            for (uint32_t s2 = 0; s2 < 4; ++s2) m += c1 && ((t.piece_col[q] >> (8 * s2)) & 0xFF) == c1;
            a.max_mult = std::max(a.max_mult, m);
        }
    a.cm = d_cm;
    a.piece_col = d_pcol;
    a.init_masks = d_init;
    a.prefixes = d_pre;
    a.n_items = n_items;
    a.prefix_len = prefix_len;
    a.push_budget = cfg.push_budget;
    a.stacks = d_stacks;
    a.max_played = d_maxp;
    a.played = d_played;
    a.next_item = d_next;
    a.out_max_depth = d_depth;
    a.out_pushes = d_pushes;
    a.out_status = d_status;
    a.out_board = d_board;
    a.evals = d_evals;
    a.ck_in_state = ck.in_state ? d_ck_state : nullptr;
    a.ck_in_masks = ck.in_state ? d_ck_masks : nullptr;
    a.ck_out_state = ck.out_state ? d_ck_state : nullptr;
    a.ck_out_masks = ck.out_state ? d_ck_masks : nullptr;
    a.ck_state_words = ck_state;
    a.ck_mask_words = ck_masks;

    hipEvent_t ev0, ev1;
    GL_CHECK(hipEventCreate(&ev0));
    GL_CHECK(hipEventCreate(&ev1));
    GL_CHECK(hipEventRecord(ev0, stream));
    hipLaunchKernelGGL(glutton_kernel, dim3(waves), dim3(kWave), L.total, stream, a);
    GL_CHECK(hipGetLastError());
    GL_CHECK(hipEventRecord(ev1, stream));
    GL_CHECK(hipStreamSynchronize(stream));
    float ms = 0;
    GL_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);

    std::vector<uint32_t> depth(n_items), status(n_items);
    std::vector<uint64_t> pushes(n_items);
    std::vector<uint16_t> boards((size_t)n_items * t.P);
    unsigned long long evals[4] = {0, 0, 0, 0};
    GL_CHECK(hipMemcpyAsync(depth.data(), d_depth, n_items * 4, hipMemcpyDeviceToHost, stream));
    GL_CHECK(hipMemcpyAsync(status.data(), d_status, n_items * 4, hipMemcpyDeviceToHost, stream));
    GL_CHECK(hipMemcpyAsync(pushes.data(), d_pushes, n_items * 8, hipMemcpyDeviceToHost, stream));
    GL_CHECK(hipMemcpyAsync(boards.data(), d_board, boards.size() * 2, hipMemcpyDeviceToHost, stream));
    GL_CHECK(hipMemcpyAsync(evals, d_evals, sizeof(evals), hipMemcpyDeviceToHost, stream));
    if (ck.out_state) {
        GL_CHECK(hipMemcpyAsync(ck.out_state, d_ck_state, (size_t)n_items * ck_state * 4, hipMemcpyDeviceToHost, stream));
        GL_CHECK(hipMemcpyAsync(ck.out_masks, d_ck_masks, (size_t)n_items * ck_masks * 8, hipMemcpyDeviceToHost, stream));
    }
    GL_CHECK(hipStreamSynchronize(stream));
    out.assign(n_items, GluttonOutcome{});
    stats = GluttonStats{};
    for (uint32_t i = 0; i < n_items; ++i) {
        out[i].max_depth = depth[i];
        out[i].pushes = pushes[i];
        out[i].status = status[i];
        out[i].board.assign(boards.begin() + (size_t)i * t.P, boards.begin() + (size_t)(i + 1) * t.P);
        stats.pushes += pushes[i];
    }
    stats.kernel_ms = ms;
    stats.evaluations = evals[0];
    stats.score_evaluations = evals[1];
    stats.rounds = evals[2];
    stats.free_slot_sum = evals[3];
    stats.waves = waves;
    stats.lds_bytes = L.total;
    stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

}  // namespace emp
