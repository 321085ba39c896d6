/*
 * emp_oracle.cpp -- CPU oracle: plain C++ restatement of the reference's
 * search algorithms, used as the checker of the HIP product path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the CPU
 * baseline leg of bench.py; never by the product library.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository quicky2000/edge_matching_puzzle).
 *
 * Pinning (what checks each part, beyond retyping it):
 *   - enumeration: node totals 37 / 3284 / 37938 and solution sets of
 *     tests/new_strategy_{3_3,4_4,5_5}; byte-identical results.bin images of
 *     the same tests; per-depth histograms of tests/compute_stats_N_N; the
 *     "New max" trace lines of tests/new_strategy_N_N/test.info; brute force on
 *     small synthetic puzzles (including identical / self-similar pieces).
 *   - glutton_max: first solutions of tests/CUDA_glutton_max_{3_3,4_4,5_5}.
 */
#include "emp_oracle.h"

#include <algorithm>
#include <cstring>
#include <set>
#include <string>
#include <vector>

namespace {

enum Kind { CENTER = 0, BORDER = 1, CORNER = 2 };

struct Piece {
    uint32_t id;         // 1-based id in file order
    uint32_t c[4];       // colours N, E, S, W
    int kind;            // number of frame (0) sides, include/emp_piece.h:126-131
    int sim;             // 0 none, 1 half similar, 2 similar, include/emp_piece.h:144-157
    uint32_t kind_index; // rank among pieces of the same kind, include/emp_piece_db.h:555-616
};

// Colour shown on side `side` by a piece laid with orientation o
// (include/emp_piece.h:173-178).
inline uint32_t shown(const Piece &p, unsigned side, unsigned o) { return p.c[(side + o) & 3u]; }

unsigned nb_bits(uint32_t v) {
    unsigned n = 0;
    while (v) { ++n; v >>= 1; }
    return n;
}

struct Model {
    uint32_t W = 0, H = 0, P = 0;
    std::vector<Piece> pieces;
    std::vector<uint32_t> by_kind[3];
    uint32_t max_colour = 0, border_id = 0, cs = 0;
};

// Puzzle model and validation: include/emp_piece_db.h:466-1010 (kinds,
// kind indices, border_color_id = max + 1 at :651, colour parity checks at
// :898-924) and the corner orientation requirement of
// include/emp_piece_corner.h:224-239.
bool build_model(uint32_t W, uint32_t H, const uint32_t *colours, Model &m) {
    if (W < 2 || H < 2 || W * H > 1024 || !colours) return false;
    m.W = W; m.H = H; m.P = W * H;
    m.pieces.resize(m.P);
    std::vector<uint32_t> count_by_colour;
    for (uint32_t i = 0; i < m.P; ++i) {
        Piece &p = m.pieces[i];
        p.id = i + 1;
        int zeros = 0;
        for (int s = 0; s < 4; ++s) {
            p.c[s] = colours[4 * i + s];
            zeros += p.c[s] == 0;
            m.max_colour = std::max(m.max_colour, p.c[s]);
        }
        if (zeros > 2) return false;
        if (zeros == 2) {
            bool adjacent = false;
            for (int s = 0; s < 4; ++s) adjacent |= p.c[s] == 0 && p.c[(s + 1) & 3] == 0;
            if (!adjacent) return false;
        }
        p.kind = zeros;
        if (p.c[0] == p.c[2] && p.c[1] == p.c[3]) p.sim = p.c[0] == p.c[1] ? 2 : 1;
        else p.sim = 0;
        p.kind_index = (uint32_t)m.by_kind[p.kind].size();
        m.by_kind[p.kind].push_back(i);
    }
    if (m.by_kind[CORNER].size() != 4) return false;
    if (m.by_kind[BORDER].size() != 2 * (W - 2 + H - 2)) return false;
    if (m.by_kind[CENTER].size() != (W - 2) * (H - 2)) return false;
    count_by_colour.assign(m.max_colour + 1, 0);
    for (auto &p : m.pieces)
        for (int s = 0; s < 4; ++s)
            if (p.c[s]) ++count_by_colour[p.c[s]];
    uint64_t total = 0;
    for (uint32_t c = 1; c <= m.max_colour; ++c) {
        if (count_by_colour[c] & 1) return false;
        total += count_by_colour[c];
    }
    uint64_t edges = (uint64_t)(W - 1) * H + (uint64_t)(H - 1) * W;
    if (total != 2 * edges) return false;
    m.border_id = m.max_colour + 1;
    m.cs = nb_bits(m.border_id);
    return true;
}

// Position orders: include/emp_spiral_strategy_generator.h:52-98 and
// include/emp_basic_strategy_generator.h:44-53.
std::vector<std::pair<uint32_t, uint32_t>> make_order(uint32_t W, uint32_t H, int order) {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    if (order == ORACLE_ORDER_BASIC) {
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) out.emplace_back(x, y);
        return out;
    }
    std::vector<char> used(W * H, 0);
    uint32_t x = 0, y = 0;
    for (uint32_t i = 0; i < W * H; ++i) {
        used[y * W + x] = 1;
        out.emplace_back(x, y);
        if (i + 1 == W * H) break;
        if (x >= y) {
            if (x + 1 < W && !used[y * W + x + 1]) ++x;
            else if (y + 1 < H && !used[(y + 1) * W + x]) ++y;
            else --x;
        } else {
            if (x && !used[y * W + x - 1]) --x;
            else if (y && !used[(y - 1) * W + x]) --y;
            else ++x;
        }
    }
    return out;
}

// Situation string (include/situation_string_formatter.h:60-81): row-major
// cells of width digits(P)+1, id right-aligned, then N/E/S/W; '-' if empty.
std::string situation_string(const Model &m, const std::vector<uint16_t> &board) {
    unsigned w = (unsigned)std::to_string(m.P).size() + 1;
    std::string out(m.P * w, '-');
    static const char kOrient[4] = {'N', 'E', 'S', 'W'};
    for (uint32_t i = 0; i < m.P; ++i) {
        if (board[i] == 0xFFFF) continue;
        std::string id = std::to_string((board[i] >> 2) + 1);
        std::string cell = std::string(w - 1 - id.size(), ' ') + id + kOrient[board[i] & 3];
        out.replace(i * w, w, cell);
    }
    return out;
}

// ---------------------------------------------------------------------------
// Exhaustive DFS: emp_strategy (include/emp_strategy.h:159-415) and
// emp_position_strategy (include/emp_position_strategy.h:125-208).
// ---------------------------------------------------------------------------
struct Enumerator {
    const Model &m;
    std::vector<std::pair<uint32_t, uint32_t>> order;
    uint32_t P;
    unsigned words[3];          // u64 words of a bitfield of each kind: 4 * n_kind bits
    unsigned nw;                // max words
    // Constraint database (include/emp_piece_db.h:1266-1312): code -> index
    // into db_store (nw words per entry), -1 when no piece matches.
    std::vector<int32_t> db_index;
    std::vector<uint64_t> db_store;
    // Identical-piece masks (include/emp_piece_db.h:848-884): for each kind
    // and transition, the bits cleared from the position's available
    // transitions once that transition has been tried.
    std::vector<uint64_t> ident[3];
    // Per strategy index.
    std::vector<int> kind;                 // position kind
    std::vector<uint32_t> neighbour;       // [P][4] strategy index or P (off board)
    std::vector<int> prev_same_kind;       // previous index of same kind or -1
    std::vector<uint64_t> avail_pieces;    // [P+1][nw]
    std::vector<uint64_t> avail_trans;     // [P+1][nw]
    std::vector<int32_t> placed;           // transition placed at index (-1 none)
    std::vector<uint64_t> initial[3];      // all pieces of a kind available

    explicit Enumerator(const Model &model, int ord) : m(model), P(model.P) {
        order = make_order(m.W, m.H, ord);
        for (int k = 0; k < 3; ++k) words[k] = (unsigned)((4 * m.by_kind[k].size() + 63) / 64);
        nw = std::max({words[0], words[1], words[2], 1u});
        build_db();
        build_identical();
        build_positions();
    }

    uint64_t *entry(int32_t idx) { return &db_store[(size_t)idx * nw]; }

    void build_db() {
        size_t codes = (size_t)1 << (4 * m.cs);
        db_index.assign(codes, -1);
        for (const Piece &p : m.pieces) {
            unsigned step = 1u << p.sim;
            for (unsigned o = 0; o < 4; o += step) {
                uint32_t ext = (p.kind_index << 2) + o;
                for (unsigned skel = 0; skel < 16; ++skel) {
                    uint32_t code = 0;
                    int nb_border = 0;
                    for (unsigned s = 0; s < 4; ++s) {
                        if (!(skel & (1u << s))) continue;
                        uint32_t col = shown(p, s, o);
                        if (!col) { col = m.border_id; ++nb_border; }
                        // The constraint slot of a side is the opposite side.
                        code |= col << (m.cs * ((s + 2) & 3));
                    }
                    if (nb_border != p.kind) continue;
                    int32_t &idx = db_index[code];
                    if (idx < 0) {
                        idx = (int32_t)(db_store.size() / nw);
                        db_store.resize(db_store.size() + nw, 0);
                    }
                    entry(idx)[ext >> 6] |= 1ull << (ext & 63);
                }
            }
        }
    }

    // Appearance identity between two pieces of the same kind: piece b turned
    // by c shows the same colours as piece a (include/emp_piece.h:181-211).
    static bool same_appearance(const Piece &a, const Piece &b, unsigned c) {
        for (unsigned s = 0; s < 4; ++s)
            if (a.c[s] != b.c[(s + c) & 3]) return false;
        return true;
    }

    void build_identical() {
        for (int k = 0; k < 3; ++k) {
            size_t n = m.by_kind[k].size();
            ident[k].assign(4 * n * nw, ~0ull);
            for (size_t ki = 0; ki < n; ++ki) {
                const Piece &a = m.pieces[m.by_kind[k][ki]];
                for (unsigned o = 0; o < 4; ++o) {
                    uint64_t *mask = &ident[k][(4 * ki + o) * nw];
                    uint32_t self = (uint32_t)(4 * ki + o);
                    mask[self >> 6] &= ~(1ull << (self & 63));
                    for (size_t kj = 0; kj < n; ++kj) {
                        if (kj == ki) continue;
                        const Piece &b = m.pieces[m.by_kind[k][kj]];
                        for (unsigned c = 0; c < 4; ++c) {
                            if (!same_appearance(a, b, c)) continue;
                            uint32_t other = (uint32_t)(4 * kj + ((c + o) & 3));
                            mask[other >> 6] &= ~(1ull << (other & 63));
                        }
                    }
                }
            }
            initial[k].assign(nw, 0);
            for (size_t b = 0; b < 4 * n; ++b) initial[k][b >> 6] |= 1ull << (b & 63);
        }
    }

    void build_positions() {
        std::vector<uint32_t> index_of(P);
        for (uint32_t i = 0; i < P; ++i) index_of[order[i].second * m.W + order[i].first] = i;
        kind.resize(P);
        neighbour.resize(4 * P);
        prev_same_kind.resize(P);
        int last[3] = {-1, -1, -1};
        for (uint32_t i = 0; i < P; ++i) {
            uint32_t x = order[i].first, y = order[i].second;
            bool xb = x == 0 || x == m.W - 1, yb = y == 0 || y == m.H - 1;
            kind[i] = (xb && yb) ? CORNER : ((xb || yb) ? BORDER : CENTER);
            prev_same_kind[i] = last[kind[i]];
            last[kind[i]] = (int)i;
            neighbour[4 * i + 0] = y > 0 ? index_of[(y - 1) * m.W + x] : P;
            neighbour[4 * i + 1] = x + 1 < m.W ? index_of[y * m.W + x + 1] : P;
            neighbour[4 * i + 2] = y + 1 < m.H ? index_of[(y + 1) * m.W + x] : P;
            neighbour[4 * i + 3] = x > 0 ? index_of[y * m.W + x - 1] : P;
        }
        avail_pieces.assign((size_t)(P + 1) * nw, 0);
        avail_trans.assign((size_t)(P + 1) * nw, 0);
        placed.assign(P + 1, -1);
    }

    const uint64_t *previous_available(uint32_t i) const {
        int prev = prev_same_kind[i];
        return prev < 0 ? initial[kind[i]].data() : &avail_pieces[(size_t)prev * nw];
    }

    const Piece &piece_of(int k, int32_t transition) const {
        return m.pieces[m.by_kind[k][(uint32_t)transition >> 2]];
    }

    // emp_position_strategy::compute_constraint (:172-181): colour of the
    // placed neighbour's facing side, border id off the board, 0 when the
    // neighbour is still empty.
    uint32_t constraint(uint32_t i) const {
        uint32_t code = 0;
        for (unsigned s = 0; s < 4; ++s) {
            uint32_t nb = neighbour[4 * i + s];
            uint32_t col;
            if (nb == P) col = m.border_id;
            else if (placed[nb] < 0) col = 0;
            else {
                col = shown(piece_of(kind[nb], placed[nb]), (s + 2) & 3, (unsigned)placed[nb] & 3);
                if (!col) col = m.border_id;
            }
            code |= col << (m.cs * ((s + 2) & 3));
        }
        return code;
    }

    // emp_strategy::compute_available_transitions (include/emp_strategy.h
    // :282-287) and emp_position_strategy::compute_available_transitions (:184-187).
    void compute_available(uint32_t i) {
        uint64_t *t = &avail_trans[(size_t)i * nw];
        if (i == P) { std::fill(t, t + nw, 0); return; }
        int32_t idx = db_index[constraint(i)];
        const uint64_t *prev = previous_available(i);
        if (idx < 0) { std::fill(t, t + nw, 0); return; }
        const uint64_t *db = &db_store[(size_t)idx * nw];
        for (unsigned w = 0; w < nw; ++w) t[w] = db[w] & prev[w];
    }

    // ffs over the bitfield, 0 when empty (emp_position_strategy.h:205-208).
    uint32_t next_transition(uint32_t i) const {
        const uint64_t *t = &avail_trans[(size_t)i * nw];
        for (unsigned w = 0; w < nw; ++w)
            if (t[w]) return w * 64 + (uint32_t)__builtin_ctzll(t[w]) + 1;
        return 0;
    }

    // emp_position_strategy::select_piece (:189-202), HANDLE_IDENTICAL_PIECES.
    void select(uint32_t i, uint32_t transition) {
        const uint64_t *prev = previous_available(i);
        uint64_t *ap = &avail_pieces[(size_t)i * nw];
        std::copy(prev, prev + nw, ap);
        uint32_t base = transition & ~3u;
        ap[base >> 6] &= ~(0xFull << (base & 63));
        uint64_t *t = &avail_trans[(size_t)i * nw];
        const uint64_t *id = &ident[kind[i]][(size_t)transition * nw];
        for (unsigned w = 0; w < nw; ++w) t[w] &= id[w];
    }

    std::vector<uint16_t> board(uint32_t upto) const {
        std::vector<uint16_t> b(P, 0xFFFF);
        for (uint32_t i = 0; i <= upto && i < P; ++i) {
            if (placed[i] < 0) continue;
            const Piece &p = piece_of(kind[i], placed[i]);
            b[order[i].second * m.W + order[i].first] = (uint16_t)(((p.id - 1) << 2) | ((uint32_t)placed[i] & 3));
        }
        return b;
    }

    struct Result {
        uint64_t nodes = 0, solutions = 0;
        bool limited = false;
        std::vector<std::vector<uint16_t>> boards;
        std::vector<uint64_t> solution_nodes;
        std::vector<uint64_t> depth_nodes, depth_dead;
        std::string trace;
    };

    // Places the given transitions at indices 0..len-1 as the DFS would
    // (emp_strategy::set_initial_state, include/emp_strategy.h:418-467);
    // false when one of them is not a candidate.
    bool preset(const uint16_t *prefix, uint32_t len) {
        for (uint32_t i = 0; i < len; ++i) {
            compute_available(i);
            uint32_t t = prefix[i];
            if (t >= 64 * nw || !(avail_trans[(size_t)i * nw + t / 64] >> (t % 64) & 1)) return false;
            placed[i] = (int32_t)t;
            select(i, t);
        }
        return true;
    }

    // emp_strategy::set_initial_state (include/emp_strategy.h:418-467): for
    // each of the k leading positions of the situation, the transitions of
    // the index are taken in order (each selection masking it and its
    // identical twins) until the situation's own; the search then restarts at
    // index k - 1 (specific_run recomputes that index's transitions).
    // Returns the start index, or -1 when a piece is not reached that way.
    int resume_from(const std::vector<int32_t> &situation) {
        uint32_t k = 0;
        while (k < P && situation[k] >= 0) ++k;
        if (k == 0) return -1;
        for (uint32_t i = 0; i < k; ++i) {
            compute_available(i);
            const uint32_t searched = (uint32_t)situation[i];
            uint32_t next = 0;
            do {
                next = next_transition(i);
                if (!next) return -1;
                --next;
                select(i, next);
            } while (next != searched);
            placed[i] = (int32_t)searched;
        }
        return (int)k - 1;
    }

    // emp_strategy::specific_run (include/emp_strategy.h:332-415).  With
    // start = 0 this is the reference's loop: index 0 takes its first
    // candidate and the search ends when index 1 has no transition left.
    // With start = L >= 1 (indices below L preset) it enumerates the subtree
    // of that prefix and ends when index L has no transition left.  With
    // stop_depth = D > 0 placements at index D-1 are recorded as prefixes and
    // not descended.
    // floor_index > 0 overrides where the search ends (resumed runs end, as
    // the reference's loop, when index 1 has no transition left).
    void run(Result &r, uint64_t node_limit, uint64_t keep, bool want_trace,
             uint32_t start = 0, uint32_t stop_depth = 0,
             std::vector<uint16_t> *prefixes = nullptr, uint32_t floor_index = 0) {
        r.depth_nodes.assign(P, 0);
        r.depth_dead.assign(P, 0);
        uint32_t index = start, max = 0;
        const uint32_t floor = floor_index ? floor_index : (start ? start : 1);
        compute_available(index);
        bool go_on = true;
        while (go_on) {
            uint32_t next = next_transition(index);
            if (next) {
                --next;
                placed[index] = (int32_t)next;
                select(index, next);
                ++r.nodes;
                ++r.depth_nodes[index];
                if (want_trace && index > max) {
                    r.trace += "New max : " + std::to_string(index) + "\n";
                    r.trace += situation_string(m, board(index)) + "\n";
                    max = index;
                }
                if (index == P - 1) {
                    ++r.solutions;
                    if (r.boards.size() < keep) {
                        r.boards.push_back(board(index));
                        r.solution_nodes.push_back(r.nodes);
                    }
                } else if (stop_depth && index + 1 == stop_depth) {
                    if (prefixes)
                        for (uint32_t j = 0; j <= index; ++j) prefixes->push_back((uint16_t)placed[j]);
                    if (node_limit && r.nodes >= node_limit) { r.limited = true; return; }
                    continue;
                }
                ++index;
                compute_available(index);
                if (index < P && !next_transition(index)) ++r.depth_dead[index - 1];
                if (node_limit && r.nodes >= node_limit) { r.limited = true; return; }
            } else {
                if (index < P) placed[index] = -1;
                go_on = index > floor;
                --index;
            }
        }
    }
};

// The reference's spiral walk revisits cells and never reaches the last rows
// when height >= width + 2; such an order is refused.
bool is_permutation(const Model &m, int order) {
    auto o = make_order(m.W, m.H, order);
    std::vector<char> seen(m.P, 0);
    for (auto &xy : o) {
        uint32_t cell = xy.second * m.W + xy.first;
        if (seen[cell]) return false;
        seen[cell] = 1;
    }
    return true;
}

int copy_order(const Model &m, int order, uint32_t *xy) {
    auto o = make_order(m.W, m.H, order);
    for (size_t i = 0; i < o.size(); ++i) { xy[2 * i] = o[i].first; xy[2 * i + 1] = o[i].second; }
    return ORACLE_OK;
}

// ---------------------------------------------------------------------------
// glutton_max: include/CUDA_glutton_max.h and include/CUDA_glutton_max_stack.h
// Masks are piece-major (include/CUDA_piece_position_info2.h:46-361): bit
// 4 * piece_index + orientation.
// ---------------------------------------------------------------------------
struct Glutton {
    const Model &m;
    uint32_t P, S, nw;
    int offset[4];
    std::vector<uint64_t> colour_mask;     // [(c-1)*4 + side][nw]
    std::vector<uint64_t> masks;           // triangular, S(S+1)/2 entries of nw words
    std::vector<uint32_t> pos_of_slot, slot_of_pos, played, max_played;
    std::vector<char> avail;               // per piece index (256 entries)
    uint32_t level = 0, max = 0;
    uint32_t base = 0;                     // levels below are prefix moves
    std::vector<uint16_t> counters;        // per piece option counters (u16)

    explicit Glutton(const Model &model) : m(model), P(model.P), S(model.P) {
        nw = (4 * P + 63) / 64;
        offset[0] = -(int)m.W; offset[1] = 1; offset[2] = (int)m.W; offset[3] = -1;
        build_colour_masks();
        masks.assign((size_t)S * (S + 1) / 2 * nw, 0);
        pos_of_slot.assign(S, 0xFFFFFFFFu);
        slot_of_pos.assign(P, 0xFFFFFFFFu);
        played.assign(S, 0xFFFFFFFFu);
        max_played.assign(S, 0xFFFFFFFFu);
        avail.assign(256, 0);
        counters.assign(256, 0);
        prepare_stack();
    }

    static void set_bit(uint64_t *mk, uint32_t piece, uint32_t o) {
        uint32_t b = 4 * piece + o;
        mk[b >> 6] |= 1ull << (b & 63);
    }

    // prepare_color_constraints (include/CUDA_common_glutton.h:109-146):
    // info(c-1, side) = (piece, o) whose side opposite(side) shows c.
    void build_colour_masks() {
        colour_mask.assign((size_t)m.max_colour * 4 * nw, 0);
        for (uint32_t c = 1; c <= m.max_colour; ++c)
            for (unsigned side = 0; side < 4; ++side)
                for (uint32_t p = 0; p < P; ++p)
                    for (unsigned o = 0; o < 4; ++o)
                        if (shown(m.pieces[p], (side + 2) & 3, o) == c)
                            set_bit(&colour_mask[((size_t)(c - 1) * 4 + side) * nw], p, o);
    }

    const uint64_t *cmask(uint32_t c, unsigned side) const {
        return &colour_mask[((size_t)(c - 1) * 4 + side) * nw];
    }

    uint32_t level_offset(uint32_t l) const { return (S * (S + 1) - (S - l) * (S - l + 1)) / 2; }
    uint64_t *info(uint32_t l, uint32_t slot) { return &masks[((size_t)level_offset(l) + slot) * nw]; }
    uint64_t *cur(uint32_t slot) { return info(level, slot); }
    uint32_t nb_info() const { return S - level; }
    bool is_free(uint32_t pos) const { return slot_of_pos[pos] < S - level; }
    bool valid(const uint64_t *mk) const {
        for (unsigned w = 0; w < nw; ++w) if (mk[w]) return true;
        return false;
    }

    // prepare_initial_capability (include/CUDA_common_glutton.h:148-211)
    // and prepare_stack for an empty start (include/CUDA_glutton_max.h:120-156).
    void prepare_stack() {
        for (uint32_t pos = 0; pos < P; ++pos) {
            uint64_t *mk = info(0, pos);
            uint32_t x = pos % m.W, y = pos / m.W;
            bool xb = x == 0 || x == m.W - 1, yb = y == 0 || y == m.H - 1;
            if (xb && yb) {
                // corner positions (0,0),(W-1,0),(W-1,H-1),(0,H-1) have frame
                // sides (N,W),(E,N),(S,E),(W,S) (include/emp_FSM_info.h:284-303)
                unsigned ci = (x == 0 && y == 0) ? 0 : (y == 0 ? 1 : (x != 0 ? 2 : 3));
                unsigned b1 = ci, b2 = (ci + 3) & 3;
                for (uint32_t idx : m.by_kind[CORNER]) {
                    const Piece &p = m.pieces[idx];
                    for (unsigned o = 0; o < 4; ++o)
                        if (shown(p, b1, o) == 0 && shown(p, b2, o) == 0) { set_bit(mk, idx, o); break; }
                }
            } else if (xb || yb) {
                unsigned side = y == 0 ? 0 : (x == m.W - 1 ? 1 : (y == m.H - 1 ? 2 : 3));
                for (uint32_t idx : m.by_kind[BORDER]) {
                    const Piece &p = m.pieces[idx];
                    unsigned zero = 0;
                    while (p.c[zero]) ++zero;
                    set_bit(mk, idx, (zero + 4 - side) & 3);
                }
            } else {
                for (uint32_t idx : m.by_kind[CENTER])
                    for (unsigned o = 0; o < 4; ++o) set_bit(mk, idx, o);
            }
            pos_of_slot[pos] = pos;
            slot_of_pos[pos] = pos;
        }
        for (uint32_t p = 0; p < P; ++p) avail[p] = 1;
    }

    // CUDA_glutton_max_stack::clear_piece_info (include/CUDA_glutton_max_stack.h:596-615)
    void clear_piece_info() {
        for (uint32_t q = 0; q < 256; ++q) counters[q] = avail[q] ? 0 : 0xFFFF;
    }

    // analyze_info (include/CUDA_glutton_max.h:258-288): add per-piece
    // orientation counts of a result mask to the u16 counters.
    void analyze(const uint64_t *r) {
        for (unsigned w = 0; w < nw; ++w) {
            uint64_t v = r[w];
            while (v) {
                unsigned b = (unsigned)__builtin_ctzll(v);
                uint32_t piece = w * 16 + b / 4;
                uint64_t nib = (v >> (b & ~3u)) & 0xF;
                counters[piece] = (uint16_t)(counters[piece] + __builtin_popcountll(nib));
                v &= ~(0xFull << (b & ~3u));
            }
        }
    }

    static unsigned popcount(const uint64_t *r, unsigned nw) {
        unsigned n = 0;
        for (unsigned w = 0; w < nw; ++w) n += (unsigned)__builtin_popcountll(r[w]);
        return n;
    }

    struct Stats {
        uint32_t min = 0xFFFFFFFFu, max = 0, total = 0;
        void update(uint32_t v) { max = v > max ? v : max; min = v < min ? v : min; total += v; }
    };

    // Evaluation of one candidate: the body shared by
    // remove_invalid_transitions (:779-933) and get_best_candidate (:1045-1291).
    // Returns false when the candidate leaves a free position or an
    // available piece without option; fills st when requested.
    bool evaluate(uint32_t slot, uint32_t piece, uint32_t o, Stats *st) {
        uint32_t pos = pos_of_slot[slot];
        avail[piece] = 0;                       // unvailability_lock_gard
        clear_piece_info();
        std::vector<uint64_t> r(nw);
        uint32_t related[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        uint32_t pw = piece / 16;
        uint64_t apply = ~(0xFull << (4 * (piece % 16)));
        bool ok = true;
        // apply_color_constraints (:528-612)
        for (unsigned l = 0; l < 4 && ok; ++l) {
            uint32_t c = shown(m.pieces[piece], l, o);
            if (!c) continue;
            uint32_t q = (uint32_t)((int)pos + offset[l]);
            if (!is_free(q)) continue;
            uint32_t j = slot_of_pos[q];
            const uint64_t *mj = cur(j), *cm = cmask(c, l);
            for (unsigned w = 0; w < nw; ++w) r[w] = mj[w] & cm[w] & (w == pw ? apply : ~0ull);
            if (!valid(r.data())) { ok = false; break; }
            related[l] = j;
            analyze(r.data());
            if (st) st->update(popcount(r.data(), nw));
        }
        // apply_simple_mask over [0, slot) then (slot, n) (:455-526)
        for (uint32_t j = 0; j < nb_info() && ok; ++j) {
            if (j == slot) continue;
            if (j == related[0] || j == related[1] || j == related[2] || j == related[3]) continue;
            const uint64_t *mj = cur(j);
            for (unsigned w = 0; w < nw; ++w) r[w] = mj[w] & (w == pw ? apply : ~0ull);
            if (!valid(r.data())) { ok = false; break; }
            analyze(r.data());
            if (st) st->update(popcount(r.data(), nw));
        }
        if (ok) {
            for (uint32_t q = 0; q < 256; ++q)
                if (!counters[q]) { ok = false; break; }
        }
        if (ok && st) {
            Stats ps;
            for (uint32_t q = 0; q < 256; ++q)
                if (avail[q]) ps.update(counters[q]);
            st->total += ps.total;
            st->min = ps.min < st->min ? ps.min : st->min;
            st->max = ps.max > st->max ? ps.max : st->max;
        }
        avail[piece] = 1;
        return ok;
    }

    // remove_invalid_transitions (include/CUDA_glutton_max.h:742-954)
    bool remove_invalid() {
        bool changed;
        do {
            changed = false;
            for (uint32_t slot = 0; slot < nb_info(); ++slot) {
                std::vector<uint64_t> snap(cur(slot), cur(slot) + nw);
                for (unsigned w = 0; w < nw; ++w) {
                    uint64_t v = snap[w];
                    while (v) {
                        unsigned b = (unsigned)__builtin_ctzll(v);
                        v &= v - 1;
                        uint32_t bit = w * 64 + b;
                        if (!evaluate(slot, bit / 4, bit % 4, nullptr)) {
                            changed = true;
                            cur(slot)[w] &= ~(1ull << b);
                        }
                    }
                }
                if (!valid(cur(slot))) return false;
            }
        } while (changed);
        return true;
    }

    // get_best_candidate (include/CUDA_glutton_max.h:1007-1301): first best
    // by (total, (max << 16) + min); at the last level the last candidate.
    bool best(uint32_t &b_slot, uint32_t &b_bit) {
        uint32_t best_total = 0, best_minmax = 0;
        bool found = false;
        for (uint32_t slot = 0; slot < nb_info(); ++slot) {
            std::vector<uint64_t> snap(cur(slot), cur(slot) + nw);
            for (unsigned w = 0; w < nw; ++w) {
                uint64_t v = snap[w];
                while (v) {
                    unsigned b = (unsigned)__builtin_ctzll(v);
                    v &= v - 1;
                    uint32_t bit = w * 64 + b;
                    Stats st;
                    if (!evaluate(slot, bit / 4, bit % 4, &st)) continue;  // not reached after remove_invalid
                    uint32_t minmax = (st.max << 16) + st.min;
                    if (st.total > best_total || (st.total == best_total && minmax > best_minmax) ||
                        level == S - 1) {
                        best_total = st.total;
                        best_minmax = minmax;
                        b_slot = slot;
                        b_bit = bit;
                        found = true;
                    }
                }
            }
        }
        return found;
    }

    // CUDA_glutton_max_stack::push (:434-495)
    void push(uint32_t slot, uint32_t pos, uint32_t piece, uint32_t o) {
        played[level] = (o << 16) | (piece << 8) | pos;
        avail[piece] = 0;
        ++level;
        if (level > max) {
            for (uint32_t k = 0; k < level; ++k) max_played[k] = played[k];
            max = level;
        }
        uint32_t last = S - level;
        uint32_t last_pos = pos_of_slot[last];
        pos_of_slot[slot] = slot;
        std::swap(pos_of_slot[last], pos_of_slot[slot]);
        std::swap(slot_of_pos[last_pos], slot_of_pos[pos]);
    }

    // CUDA_glutton_max_stack::pop (:497-522)
    uint32_t pop() {
        uint32_t last = S - level;
        uint32_t slot = pos_of_slot[last];
        uint32_t pl = played[level - 1];
        uint32_t pos = pl & 0xFF, piece = (pl >> 8) & 0xFF;
        --level;
        avail[piece] = 1;
        pos_of_slot[last] = pos;
        uint32_t other = pos_of_slot[slot];
        std::swap(pos_of_slot[last], pos_of_slot[slot]);
        std::swap(slot_of_pos[other], slot_of_pos[pos]);
        return slot;
    }

    // apply_best_candidate (include/CUDA_glutton_max.h:1303-1476). Returns
    // false where the reference would stop (its "should not be reached"
    // branches): an emptied position.
    bool apply(uint32_t slot, uint32_t bit) {
        uint32_t piece = bit / 4, o = bit % 4;
        cur(slot)[bit / 64] &= ~(1ull << (bit % 64));
        uint32_t pos = pos_of_slot[slot];
        uint32_t pw = piece / 16;
        uint64_t apply_mask = ~(0xFull << (4 * (piece % 16)));
        uint32_t n = nb_info();
        bool next_level = level + 1 < S;
        if (next_level) {
            for (uint32_t j = 0; j + 1 < n; ++j) {
                if (j == slot) continue;
                const uint64_t *src = cur(j);
                uint64_t *dst = info(level + 1, j);
                for (unsigned w = 0; w < nw; ++w) dst[w] = src[w] & (w == pw ? apply_mask : ~0ull);
                if (!valid(dst)) return false;
            }
            if (slot < n - 1) {
                const uint64_t *src = cur(n - 1);
                uint64_t *dst = info(level + 1, slot);
                for (unsigned w = 0; w < nw; ++w) dst[w] = src[w] & (w == pw ? apply_mask : ~0ull);
                if (!valid(dst)) return false;
            }
        }
        push(slot, pos, piece, o);
        for (unsigned l = 0; l < 4; ++l) {
            uint32_t c = shown(m.pieces[piece], l, o);
            if (!c) continue;
            uint32_t q = (uint32_t)((int)pos + offset[l]);
            if (!is_free(q)) continue;
            uint64_t *mj = cur(slot_of_pos[q]);
            const uint64_t *cm = cmask(c, l);
            for (unsigned w = 0; w < nw; ++w) mj[w] &= cm[w];
            if (!valid(mj)) return false;
        }
        return true;
    }

    // Prefix seeding: each move is applied like a chosen candidate (the
    // reference's prepare_stack skips the colour constraints of pieces
    // already on the board, include/CUDA_glutton_max.h:136-152).
    // Returns 0 when played, 1 when the move is not playable (occupied cell,
    // used piece, or not a candidate of the cell), 2 when it was played but
    // left a free cell without option (the search then ends at once).
    int seed(uint32_t pos, uint32_t piece, uint32_t o) {
        if (pos >= P || piece >= P || o > 3 || !is_free(pos) || !avail[piece]) return 1;
        uint32_t slot = slot_of_pos[pos];
        uint32_t bit = 4 * piece + o;
        if (!(cur(slot)[bit / 64] >> (bit % 64) & 1)) return 1;
        if (!apply(slot, bit)) return 2;
        base = level;
        return 0;
    }

    // kernel (include/CUDA_glutton_max.h:1522-1637)
    int run(uint64_t push_limit, uint64_t &pushes) {
        bool new_level = true;
        uint32_t pop_slot = 0;
        pushes = 0;
        while (level < S) {
            if (!new_level && !valid(cur(pop_slot))) {
                if (level == base) return ORACLE_OK;
                pop_slot = pop();
                new_level = false;
                continue;
            }
            if (!remove_invalid()) {
                if (level == base) return ORACLE_OK;
                pop_slot = pop();
                new_level = false;
                continue;
            }
            uint32_t bs = 0, bb = 0;
            if (!best(bs, bb)) return ORACLE_BAD_ARGUMENT;
            if (!apply(bs, bb)) return ORACLE_BAD_ARGUMENT;
            ++pushes;
            new_level = true;
            if (push_limit && pushes >= push_limit && level < S) return ORACLE_LIMIT_REACHED;
        }
        return ORACLE_OK;
    }
};

// ---------------------------------------------------------------------------
// CUDA_backtracker (src/CUDA_backtracker.cu:68-178) with its inputs:
// variables of emp_variable_generator::create_variables
// (include/emp_variable_generator.h:336-531: per position in strategy order,
// every kept (piece, orientation) of the position's kind matching the frame),
// level-0 capability of system_equation_for_CUDA::prepare_initial
// (include/system_equation_for_CUDA.h:159-173) and the transitions of
// prepare_transitions (:70-134): a variable clears every variable of its
// position and of its piece, and the colour-incompatible variables of its
// four neighbours.  Capabilities are orientation-major
// (include/CUDA_piece_position_info.h:67-88): bit 256 * o + piece, so the
// first set bit is the lowest orientation, then the lowest piece.  The
// position index of a capability is the strategy index; the kernel looks a
// transition up by step * 1024 + 256 * o + piece while the table is keyed by
// the row-major position, which agree for the basic order only.
struct Backtracker {
    const Model &m;
    std::vector<std::pair<uint32_t, uint32_t>> order;
    uint32_t P;
    static constexpr unsigned kWords = 16;   // 1024 bits
    std::vector<uint64_t> caps;              // [(P + 1) stack levels][P positions][kWords]
    std::vector<uint32_t> var_of_step;

    Backtracker(const Model &model, int ord) : m(model), P(model.P) {
        order = make_order(m.W, m.H, ord);
    }

    static void set(uint64_t *c, uint32_t piece, uint32_t o) {
        uint32_t b = 256 * o + piece;
        c[b >> 6] |= 1ull << (b & 63);
    }
    uint64_t *cap(uint32_t level, uint32_t pos) { return &caps[((size_t)level * P + pos) * kWords]; }

    // Does (piece, o) at cell (x, y) match the frame, and is it kept?
    bool frame_ok(const Piece &p, unsigned o, uint32_t x, uint32_t y) const {
        if (!Puzzle_orientation_kept(p, o)) return false;
        const bool off[4] = {y == 0, x + 1 == m.W, y + 1 == m.H, x == 0};
        for (unsigned s = 0; s < 4; ++s)
            if ((shown(p, s, o) == 0) != off[s]) return false;
        return true;
    }
    static bool Puzzle_orientation_kept(const Piece &p, unsigned o) { return (o & ((1u << p.sim) - 1u)) == 0; }

    int kind_at(uint32_t x, uint32_t y) const {
        bool xb = x == 0 || x == m.W - 1, yb = y == 0 || y == m.H - 1;
        return (xb && yb) ? CORNER : ((xb || yb) ? BORDER : CENTER);
    }

    // Transition of variable (piece, o) at strategy index s applied to the
    // capability of a later index j: same piece cleared, and at a neighbour
    // the variables whose facing colour differs.
    void apply_transition(uint32_t s, uint32_t piece, uint32_t o, uint32_t j, uint64_t *c) const {
        for (unsigned oo = 0; oo < 4; ++oo) {
            uint32_t b = 256 * oo + piece;
            c[b >> 6] &= ~(1ull << (b & 63));
        }
        const uint32_t xs = order[s].first, ys = order[s].second, xj = order[j].first, yj = order[j].second;
        int side = -1;   // side of the variable facing cell j
        if (xj == xs && yj + 1 == ys) side = 0;
        else if (xj == xs + 1 && yj == ys) side = 1;
        else if (xj == xs && yj == ys + 1) side = 2;
        else if (xj + 1 == xs && yj == ys) side = 3;
        if (side < 0) return;
        const uint32_t col = shown(m.pieces[piece], (unsigned)side, o);
        for (uint32_t q = 0; q < P; ++q)
            for (unsigned oo = 0; oo < 4; ++oo) {
                uint32_t b = 256 * oo + q;
                if (!(c[b >> 6] >> (b & 63) & 1)) continue;
                if (q == piece) continue;
                if (shown(m.pieces[q], ((unsigned)side + 2) & 3, oo) != col) c[b >> 6] &= ~(1ull << (b & 63));
            }
    }

    // Returns true when a solution was found; trace gets (lane, variable, step).
    bool run(std::vector<uint32_t> &trace) {
        caps.assign((size_t)(P + 1) * P * kWords, 0);
        for (uint32_t i = 0; i < P; ++i) {
            const uint32_t x = order[i].first, y = order[i].second;
            const int k = kind_at(x, y);
            for (uint32_t idx : m.by_kind[k])
                for (unsigned o = 0; o < 4; ++o)
                    if (frame_ok(m.pieces[idx], o, x, y)) set(cap(0, i), idx, o);
        }
        var_of_step.assign(P, 0);
        uint32_t step = 0;
        while (step < P) {
            uint64_t *avail = cap(step, step);
            int bit = -1;
            for (unsigned w = 0; w < kWords && bit < 0; ++w)
                if (avail[w]) bit = (int)(w * 64 + __builtin_ctzll(avail[w]));
            if (bit < 0) {
                if (step == 0) return false;
                --step;
                continue;
            }
            const uint32_t o = (uint32_t)bit / 256, piece = (uint32_t)bit % 256;
            const uint32_t lane = (uint32_t)bit / 32, variable = (step << 10) | (uint32_t)bit;
            trace.push_back(lane);
            trace.push_back(variable);
            trace.push_back(step);
            avail[bit >> 6] &= ~(1ull << (bit & 63));
            var_of_step[step] = (uint32_t)bit;
            if (step == P - 1) return true;
            bool valid = true;
            for (uint32_t j = step + 1; j < P && valid; ++j) {
                uint64_t *dst = cap(step + 1, j);
                std::copy(cap(step, j), cap(step, j) + kWords, dst);
                apply_transition(step, piece, o, j, dst);
                uint64_t any = 0;
                for (unsigned w = 0; w < kWords; ++w) any |= dst[w];
                valid = any != 0;
            }
            if (valid) ++step;
        }
        return true;
    }
};

}  // namespace

extern "C" {

int oracle_order(uint32_t width, uint32_t height, int order, uint32_t *xy) {
    if (!xy || width < 1 || height < 1) return ORACLE_BAD_ARGUMENT;
    Model m;
    m.W = width; m.H = height; m.P = width * height;
    return copy_order(m, order, xy);
}

int oracle_enumerate(uint32_t width, uint32_t height, const uint32_t *colours,
                     int order, uint64_t node_limit,
                     uint64_t max_solutions, uint16_t *boards,
                     uint64_t *solution_nodes,
                     uint64_t *n_solutions, uint64_t *n_nodes,
                     uint64_t *depth_nodes, uint64_t *depth_dead,
                     char *trace, uint64_t trace_capacity) {
    Model m;
    if (order != ORACLE_ORDER_SPIRAL && order != ORACLE_ORDER_BASIC) return ORACLE_BAD_ARGUMENT;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!is_permutation(m, order)) return ORACLE_BAD_ARGUMENT;
    Enumerator e(m, order);
    Enumerator::Result r;
    e.run(r, node_limit, boards ? max_solutions : 0, trace != nullptr);
    for (size_t s = 0; s < r.boards.size(); ++s) {
        std::copy(r.boards[s].begin(), r.boards[s].end(), boards + s * m.P);
        if (solution_nodes) solution_nodes[s] = r.solution_nodes[s];
    }
    if (n_solutions) *n_solutions = r.solutions;
    if (n_nodes) *n_nodes = r.nodes;
    if (depth_nodes) std::copy(r.depth_nodes.begin(), r.depth_nodes.end(), depth_nodes);
    if (depth_dead) std::copy(r.depth_dead.begin(), r.depth_dead.end(), depth_dead);
    if (trace && trace_capacity) {
        size_t n = std::min<size_t>(r.trace.size(), trace_capacity - 1);
        std::memcpy(trace, r.trace.data(), n);
        trace[n] = 0;
    }
    return r.limited ? ORACLE_LIMIT_REACHED : ORACLE_OK;
}

int oracle_enumerate_prefixes(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                              uint32_t depth, uint16_t *prefixes, uint64_t capacity,
                              uint64_t *n_prefixes, uint64_t *n_nodes, uint64_t *n_solutions) {
    Model m;
    if (order != ORACLE_ORDER_SPIRAL && order != ORACLE_ORDER_BASIC) return ORACLE_BAD_ARGUMENT;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!is_permutation(m, order) || depth < 2 || depth >= m.P) return ORACLE_BAD_ARGUMENT;
    Enumerator e(m, order);
    Enumerator::Result r;
    std::vector<uint16_t> pre;
    e.run(r, 0, 0, false, 0, depth, &pre);
    uint64_t n = pre.size() / depth;
    if (n_prefixes) *n_prefixes = n;
    if (n_nodes) *n_nodes = r.nodes;
    if (n_solutions) *n_solutions = r.solutions;
    if (prefixes) std::copy(pre.begin(), pre.begin() + std::min<uint64_t>(n, capacity) * depth, prefixes);
    return n > capacity ? ORACLE_LIMIT_REACHED : ORACLE_OK;
}

int oracle_enumerate_subtree(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                             uint32_t prefix_len, const uint16_t *prefix, uint64_t max_solutions,
                             uint16_t *boards, uint64_t *n_solutions, uint64_t *n_nodes) {
    Model m;
    if (order != ORACLE_ORDER_SPIRAL && order != ORACLE_ORDER_BASIC) return ORACLE_BAD_ARGUMENT;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!is_permutation(m, order) || prefix_len < 1 || prefix_len >= m.P || !prefix) return ORACLE_BAD_ARGUMENT;
    Enumerator e(m, order);
    if (!e.preset(prefix, prefix_len)) return ORACLE_BAD_ARGUMENT;
    Enumerator::Result r;
    e.run(r, 0, boards ? max_solutions : 0, false, prefix_len);
    for (size_t s = 0; s < r.boards.size(); ++s) std::copy(r.boards[s].begin(), r.boards[s].end(), boards + s * m.P);
    if (n_solutions) *n_solutions = r.solutions;
    if (n_nodes) *n_nodes = r.nodes;
    return ORACLE_OK;
}

int oracle_enumerate_resume(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                            const uint16_t *situation, uint64_t max_solutions, uint16_t *boards,
                            uint64_t *solution_nodes, uint64_t *n_solutions, uint64_t *n_nodes) {
    Model m;
    if (order != ORACLE_ORDER_SPIRAL && order != ORACLE_ORDER_BASIC) return ORACLE_BAD_ARGUMENT;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!is_permutation(m, order) || !situation) return ORACLE_BAD_ARGUMENT;
    Enumerator e(m, order);
    // situation cells (row-major) -> transitions by strategy index
    std::vector<int32_t> tr(m.P, -1);
    for (uint32_t i = 0; i < m.P; ++i) {
        const uint16_t v = situation[e.order[i].second * m.W + e.order[i].first];
        if (v == 0xFFFF) continue;
        const uint32_t piece = v >> 2;
        if (piece >= m.P) return ORACLE_BAD_ARGUMENT;
        const auto &list = m.by_kind[e.kind[i]];
        int32_t ki = -1;
        for (size_t j = 0; j < list.size(); ++j)
            if (list[j] == piece) ki = (int32_t)j;
        if (ki < 0) return ORACLE_BAD_ARGUMENT;
        tr[i] = 4 * ki + (v & 3);
    }
    const int start = e.resume_from(tr);
    if (start < 0) return ORACLE_BAD_ARGUMENT;
    Enumerator::Result r;
    e.run(r, 0, boards ? max_solutions : 0, false, (uint32_t)start, 0, nullptr, 1);
    for (size_t s = 0; s < r.boards.size(); ++s) {
        std::copy(r.boards[s].begin(), r.boards[s].end(), boards + s * m.P);
        if (solution_nodes) solution_nodes[s] = r.solution_nodes[s];
    }
    if (n_solutions) *n_solutions = r.solutions;
    if (n_nodes) *n_nodes = r.nodes;
    return ORACLE_OK;
}

int oracle_subtree_depth(uint32_t width, uint32_t height, const uint32_t *colours, int order,
                         uint32_t prefix_len, const uint16_t *prefix, uint64_t *n_nodes, uint32_t *max_depth) {
    Model m;
    if (order != ORACLE_ORDER_SPIRAL && order != ORACLE_ORDER_BASIC) return ORACLE_BAD_ARGUMENT;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!is_permutation(m, order) || prefix_len < 1 || prefix_len >= m.P || !prefix) return ORACLE_BAD_ARGUMENT;
    Enumerator e(m, order);
    if (!e.preset(prefix, prefix_len)) return ORACLE_BAD_ARGUMENT;
    Enumerator::Result r;
    e.run(r, 0, 0, false, prefix_len);
    uint32_t depth = prefix_len;
    for (uint32_t i = 0; i < m.P; ++i)
        if (r.depth_nodes[i]) depth = std::max(depth, i + 1);
    if (n_nodes) *n_nodes = r.nodes;
    if (max_depth) *max_depth = depth;
    return ORACLE_OK;
}

// results.bin v1 (include/emp_situation_binary_dumper.h:72-126, records
// written by include/emp_strategy.h:258-266 / :367-375 and :410).
int oracle_results_bin(uint32_t width, uint32_t height, const uint32_t *colours,
                       int order, uint8_t *out, uint64_t capacity,
                       uint64_t *size) {
    Model m;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!is_permutation(m, order)) return ORACLE_BAD_ARGUMENT;
    Enumerator e(m, order);
    Enumerator::Result r;
    e.run(r, 0, ~0ull, false);
    std::vector<uint8_t> img;
    auto put32 = [&](uint32_t v) { for (int i = 0; i < 4; ++i) img.push_back((uint8_t)(v >> (8 * i))); };
    auto put64 = [&](uint64_t v) { for (int i = 0; i < 8; ++i) img.push_back((uint8_t)(v >> (8 * i))); };
    put32(1); put32(width); put32(height); put32(1);
    for (auto &xy : e.order) { put32(xy.first); put32(xy.second); }
    unsigned f = nb_bits(m.P - 1) + 2;
    size_t nwords = ((size_t)m.P * f + 63) / 64;
    for (size_t s = 0; s < r.boards.size(); ++s) {
        std::vector<uint64_t> bits(nwords, 0);
        for (uint32_t i = 0; i < m.P; ++i) {
            uint64_t v = r.boards[s][e.order[i].second * width + e.order[i].first];
            size_t off = (size_t)i * f;
            bits[off / 64] |= v << (off % 64);
            if (off % 64 + f > 64) bits[off / 64 + 1] |= v >> (64 - off % 64);
        }
        for (auto w : bits) put64(w);
        put64(r.solution_nodes[s]);
    }
    put64(r.nodes);
    *size = img.size();
    if (out) std::memcpy(out, img.data(), std::min<uint64_t>(capacity, img.size()));
    return ORACLE_OK;
}

int oracle_backtracker(uint32_t width, uint32_t height, const uint32_t *colours, int order, uint16_t *board,
                       uint32_t *found, uint32_t *trace, uint64_t trace_capacity, uint64_t *trace_length) {
    Model m;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (m.P > 256 || !is_permutation(m, order)) return ORACLE_BAD_ARGUMENT;
    Backtracker b(m, order);
    std::vector<uint32_t> tr;
    bool ok = b.run(tr);
    if (found) *found = ok ? 1 : 0;
    if (board) {
        std::fill(board, board + m.P, 0xFFFF);
        if (ok)
            for (uint32_t i = 0; i < m.P; ++i) {
                const uint32_t bit = b.var_of_step[i];
                board[b.order[i].second * width + b.order[i].first] = (uint16_t)(((bit % 256) << 2) | (bit / 256));
            }
    }
    if (trace_length) *trace_length = tr.size() / 3;
    if (trace) std::copy(tr.begin(), tr.begin() + std::min<uint64_t>(tr.size(), 3 * trace_capacity), trace);
    return ORACLE_OK;
}

int oracle_glutton_masks(uint32_t width, uint32_t height, const uint32_t *colours, uint32_t n_moves,
                         const uint32_t *moves, uint64_t *masks, uint32_t *status) {
    // glutton_wide's play_from (include/CUDA_glutton_situations.h:713-891)
    // makes the masks apply_best_candidate makes (include/CUDA_glutton_max.h
    // :1303-1476): each move is replayed as a chosen candidate.
    Model m;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (m.P > 256 || !masks || !status) return ORACLE_BAD_ARGUMENT;
    Glutton g(m);
    *status = 0;
    for (uint32_t i = 0; i < n_moves && !*status; ++i) {
        const int r = g.seed(moves[3 * i], moves[3 * i + 1], moves[3 * i + 2]);
        *status = r == 1 ? 2u : (r == 2 ? 1u : 0u);
    }
    std::fill(masks, masks + (size_t)m.P * g.nw, 0ull);
    for (uint32_t q = 0; q < m.P; ++q)
        if (g.is_free(q)) std::copy(g.cur(g.slot_of_pos[q]), g.cur(g.slot_of_pos[q]) + g.nw, masks + (size_t)q * g.nw);
    return ORACLE_OK;
}

int oracle_glutton_max(uint32_t width, uint32_t height, const uint32_t *colours,
                       uint32_t n_prefix, const uint32_t *prefix_moves,
                       uint64_t push_limit, uint16_t *board,
                       uint32_t *max_depth, uint64_t *n_pushes,
                       uint32_t *solved) {
    Model m;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (m.P > 256) return ORACLE_BAD_ARGUMENT;   // include/CUDA_common_struct_glutton.h:358
    Glutton g(m);
    bool dead = false;
    for (uint32_t i = 0; i < n_prefix && !dead; ++i) {
        int r = g.seed(prefix_moves[3 * i], prefix_moves[3 * i + 1], prefix_moves[3 * i + 2]);
        if (r == 1) return ORACLE_BAD_ARGUMENT;
        dead = r == 2;
    }
    uint64_t pushes = 0;
    int status = dead ? ORACLE_OK : g.run(push_limit, pushes);
    bool full = g.level == g.S;
    if (board) {
        std::fill(board, board + m.P, 0xFFFF);
        const std::vector<uint32_t> &src = full ? g.played : g.max_played;
        uint32_t n = full ? g.S : g.max;
        for (uint32_t k = 0; k < n; ++k) {
            uint32_t pl = src[k];
            board[pl & 0xFF] = (uint16_t)((((pl >> 8) & 0xFF) << 2) | ((pl >> 16) & 3));
        }
    }
    if (max_depth) *max_depth = g.max;
    if (n_pushes) *n_pushes = pushes;
    if (solved) *solved = full ? 1 : 0;
    return status;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// situation_profile: feature_situation_profile::template_run and
// compute_glutton (include/feature_situation_profile.h:150-483).  A
// capability is the set of live variables: each variable sits in the info of
// its position and in the info of its piece, so a profile -- the 2P bit
// counts, sorted (include/situation_capability.h:145-155) -- counts the live
// variables per position and per piece.
// ---------------------------------------------------------------------------
struct Profiler {
    struct Var {
        uint32_t pos, x, y, piece, o;
    };
    const Model &m;
    uint32_t P;
    std::vector<std::pair<uint32_t, uint32_t>> order;
    std::vector<Var> vars;
    std::vector<int32_t> var_at;   // [P cells][P pieces][4] -> variable index or -1

    Profiler(const Model &model) : m(model), P(model.P) {
        order = make_order(m.W, m.H, ORACLE_ORDER_BASIC);
        var_at.assign((size_t)P * P * 4, -1);
        // create_variables: every kept orientation of a piece of the cell's
        // kind that shows the frame exactly on the frame sides
        for (uint32_t i = 0; i < P; ++i) {
            const uint32_t x = order[i].first, y = order[i].second;
            const bool off[4] = {y == 0, x + 1 == m.W, y + 1 == m.H, x == 0};
            for (uint32_t q = 0; q < P; ++q) {
                const Piece &p = m.pieces[q];
                for (unsigned o = 0; o < 4; ++o) {
                    if (o & ((1u << p.sim) - 1u)) continue;
                    bool ok = true;
                    for (unsigned s = 0; s < 4 && ok; ++s) ok = (shown(p, s, o) == 0) == off[s];
                    if (!ok) continue;
                    var_at[((size_t)(y * m.W + x) * P + q) * 4 + o] = (int32_t)vars.size();
                    vars.push_back({i, x, y, q, o});
                }
            }
        }
    }

    // Does the transition of v (prepare_transitions) mask u?  Same position,
    // same piece, or an adjacent cell whose shared edge shows another colour
    // (treat_piece_relation_equation, pieces differing).
    bool masks(const Var &v, const Var &u) const {
        if (u.pos == v.pos || u.piece == v.piece) return true;
        int side = -1;   // side of v facing u
        if (u.x == v.x && u.y + 1 == v.y) side = 0;
        else if (u.x == v.x + 1 && u.y == v.y) side = 1;
        else if (u.x == v.x && u.y == v.y + 1) side = 2;
        else if (u.x + 1 == v.x && u.y == v.y) side = 3;
        if (side < 0) return false;
        return shown(m.pieces[v.piece], (unsigned)side, v.o) != shown(m.pieces[u.piece], ((unsigned)side + 2) & 3, u.o);
    }

    void apply(std::vector<char> &live, uint32_t v) const {
        for (size_t u = 0; u < vars.size(); ++u)
            if (live[u] && masks(vars[v], vars[u])) live[u] = 0;
    }

    std::vector<uint32_t> profile(const std::vector<char> &live) const {
        std::vector<uint32_t> c(2 * P, 0);
        for (size_t u = 0; u < vars.size(); ++u)
            if (live[u]) {
                ++c[vars[u].pos];
                ++c[P + vars[u].piece];
            }
        std::sort(c.begin(), c.end());
        return c;
    }

    // situation_profile's comparisons; a profile carries its level.  The
    // total is accumulated in 32-bit arithmetic, as the reference's is
    // (the reference's min / max start profiles fill every value with 0 or
    // 0xFFFFFFFF, whose total wraps).
    struct Prof {
        uint32_t level;
        std::vector<uint32_t> v;
        uint32_t total() const {
            uint32_t t = 0;
            for (uint32_t x : v) t += x;
            return t;
        }
        uint32_t max() const { return v.back(); }
        uint32_t min() const { return 2 * level < v.size() ? v[2 * level] : 0; }
        bool valid() const { return level < v.size() / 2 ? v[2 * level] != 0 : true; }
    };
    static bool less(int criterion, const Prof &a, const Prof &b) {
        switch (criterion) {
            case 0: return a.total() < b.total();
            case 1: return std::make_pair(a.max(), a.total()) < std::make_pair(b.max(), b.total());
            case 2: return std::make_pair(a.min(), a.total()) < std::make_pair(b.min(), b.total());
            default: return a.v < b.v;   // vector_op, and rvector_op, which the reference compares the same way
        }
    }

    // compute_glutton for one criterion: from the level-0 capability, place
    // the situation's pieces one at a time, each time the one whose
    // capability is the extremum (the first invalid one at once); the total
    // after each placement (0 for an invalid one), padded to P + 1 values.
    std::vector<uint32_t> glutton(const std::vector<char> &live0, const std::vector<uint32_t> &placed_vars,
                                  int criterion, bool want_max) const {
        std::vector<uint32_t> out;
        std::vector<char> live = live0;
        out.push_back(Prof{0, profile(live)}.total());
        std::vector<char> used(placed_vars.size(), 0);
        for (size_t level = 0; level < placed_vars.size(); ++level) {
            Prof ext{0, std::vector<uint32_t>(2 * P, want_max ? 0u : 0xFFFFFFFFu)};
            uint32_t ext_score = 0;
            int chosen = -1;
            std::vector<char> chosen_live;
            for (size_t c = 0; c < placed_vars.size(); ++c) {
                if (used[c]) continue;
                std::vector<char> next = live;
                apply(next, placed_vars[c]);
                Prof cur{(uint32_t)level + 1, profile(next)};
                const bool ok = cur.valid();
                const uint32_t score = ok ? cur.total() : 0;
                if (!ok || want_max != less(criterion, cur, ext)) {
                    ext = cur;
                    ext_score = score;
                    chosen = (int)c;
                    chosen_live = next;
                    if (!ok) break;
                }
            }
            used[chosen] = 1;
            live = chosen_live;
            out.push_back(ext_score);
        }
        while (out.size() < P + 1) out.push_back(0);
        return out;
    }
};

int oracle_situation_profile(uint32_t width, uint32_t height, const uint32_t *colours, const uint16_t *situation,
                             uint32_t *level, uint32_t *glutton, uint32_t *surface, uint32_t *row_index,
                             uint32_t *n_variables) {
    Model m;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!situation || !level || !glutton || !surface || !row_index) return ORACLE_BAD_ARGUMENT;
    Profiler pr(m);
    const uint32_t P = m.P;
    if (n_variables) *n_variables = (uint32_t)pr.vars.size();
    // the situation's placements in strategy order
    std::vector<uint32_t> placed;
    std::vector<uint32_t> placed_index;
    for (uint32_t i = 0; i < P; ++i) {
        const uint32_t cell = pr.order[i].second * width + pr.order[i].first;
        const uint16_t b = situation[cell];
        if (b == 0xFFFF) continue;
        if ((uint32_t)(b >> 2) >= P) return ORACLE_BAD_ARGUMENT;
        const int32_t v = pr.var_at[((size_t)cell * P + (b >> 2)) * 4 + (b & 3)];
        if (v < 0) return ORACLE_BAD_ARGUMENT;
        placed.push_back((uint32_t)v);
        placed_index.push_back(i);
    }
    *level = (uint32_t)placed.size();
    std::vector<char> live0(pr.vars.size(), 1);
    for (int criterion = 0; criterion < 5; ++criterion)
        for (int mx = 0; mx < 2; ++mx) {
            const std::vector<uint32_t> g = pr.glutton(live0, placed, criterion, mx != 0);
            std::copy(g.begin(), g.end(), glutton + (size_t)(2 * criterion + mx) * (P + 1));
        }
    // the profile before each placement in strategy order, then at the end
    std::vector<char> live = live0;
    for (size_t k = 0; k < placed.size(); ++k) {
        const std::vector<uint32_t> pf = pr.profile(live);
        std::copy(pf.begin(), pf.end(), surface + k * 2 * P);
        row_index[k] = placed_index[k];
        pr.apply(live, placed[k]);
    }
    const std::vector<uint32_t> pf = pr.profile(live);
    std::copy(pf.begin(), pf.end(), surface + placed.size() * 2 * P);
    row_index[placed.size()] = P;
    return ORACLE_OK;
}

// ---------------------------------------------------------------------------
// profile: feature_profile::template_run and wide (include/feature_profile.h
// :582-868).  A level-by-level search over every situation the capability
// allows, placements in any order: from each situation of a level, every
// live variable (a piece, orientation at a free position) is placed; a child
// whose profile is valid joins the next level's set (include/situation_
// profile.h:170-178: no free position or piece left without a variable) and
// updates the per-level extrema of its profile's min, max and total.  The
// first pass records them over all situations and takes the last level's
// set as the solutions; the second splits them between situations that lead
// to a solution (predecessors, include/emp_situation.h:567-584, and full
// boards) and the others.  At 25 pieces the four corners are placed first:
// corner 0 at (0, 0), the other three at the other corners in the six orders
// of next_permutation, each in its first orientation that is a variable
// there (:686-764), and the search runs the remaining 21 levels from each;
// the solutions are then the last level's set of the first start that
// leaves the recorded set empty before it (:766-776).
// ---------------------------------------------------------------------------
namespace {

struct ExtremumStat {   // simple_stat (include/feature_profile.h:201-237)
    bool updated = false;
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    void update(uint32_t v) {
        updated = true;
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    uint32_t get_min() const { return updated ? mn : 0u; }
    uint32_t get_max() const { return mx; }
};

struct LevelStats {     // profile_level_stats: low = profile min, high = profile max
    ExtremumStat low, high, total;
};

using Board = std::vector<uint16_t>;

struct ProfileSearch {
    const Model &m;
    const Profiler &pr;
    uint64_t max_situations;
    bool too_large = false;
    // live variables as bitsets: keep[v] = the variables v's transition
    // leaves (Profiler::masks), at[i] = the variables of position i (i < P)
    // or of piece i - P
    size_t nw = 0;
    std::vector<uint64_t> keep, at;

    ProfileSearch(const Model &model, const Profiler &profiler, uint64_t limit)
        : m(model), pr(profiler), max_situations(limit) {
        const size_t nv = pr.vars.size();
        nw = (nv + 63) / 64;
        keep.assign(nv * nw, 0);
        at.assign((size_t)2 * m.P * nw, 0);
        for (size_t v = 0; v < nv; ++v) {
            for (size_t u = 0; u < nv; ++u)
                if (!pr.masks(pr.vars[v], pr.vars[u])) keep[v * nw + u / 64] |= 1ull << (u % 64);
            at[(size_t)pr.vars[v].pos * nw + v / 64] |= 1ull << (v % 64);
            at[(size_t)(m.P + pr.vars[v].piece) * nw + v / 64] |= 1ull << (v % 64);
        }
    }
    std::vector<uint64_t> all_live() const {
        std::vector<uint64_t> live(nw, 0);
        for (size_t v = 0; v < pr.vars.size(); ++v) live[v / 64] |= 1ull << (v % 64);
        return live;
    }
    void apply(std::vector<uint64_t> &live, size_t v) const {
        for (size_t w = 0; w < nw; ++w) live[w] &= keep[v * nw + w];
    }
    std::vector<uint32_t> profile(const std::vector<uint64_t> &live) const {
        std::vector<uint32_t> c(2 * m.P, 0);
        for (size_t i = 0; i < c.size(); ++i)
            for (size_t w = 0; w < nw; ++w) c[i] += (uint32_t)__builtin_popcountll(live[w] & at[i * nw + w]);
        std::sort(c.begin(), c.end());
        return c;
    }
    std::vector<uint64_t> rebuild(const Board &b) const {
        std::vector<uint64_t> live = all_live();
        for (uint32_t cell = 0; cell < m.P; ++cell)
            if (b[cell] != 0xFFFF) apply(live, (size_t)pr.var_at[((size_t)cell * m.P + (b[cell] >> 2)) * 4 + (b[cell] & 3)]);
        return live;
    }
    static uint32_t level_of(const Board &b) {
        uint32_t n = 0;
        for (uint16_t c : b) n += c != 0xFFFF;
        return n;
    }
    static bool predecessor(const Board &a, const Board &b) {
        if (level_of(a) >= level_of(b)) return false;
        for (size_t i = 0; i < a.size(); ++i)
            if (a[i] != 0xFFFF && a[i] != b[i]) return false;
        return true;
    }
    static void update(std::vector<LevelStats> &st, const Profiler::Prof &p) {
        st[p.level].low.update(p.min());
        st[p.level].high.update(p.max());
        st[p.level].total.update(p.total());
    }

    void wide(std::set<Board> &solutions, std::vector<LevelStats> &lead, std::vector<LevelStats> &other,
              const Board &start, uint32_t levels) {
        std::set<Board> cur{start}, next;
        for (uint32_t l = 0; l < levels && !too_large; ++l) {
            next.clear();
            for (const Board &s : cur) {
                const std::vector<uint64_t> live = rebuild(s);
                for (size_t w = 0; w < nw; ++w)
                    for (uint64_t bits = live[w]; bits; bits &= bits - 1) {
                        const size_t v = w * 64 + (size_t)__builtin_ctzll(bits);
                        std::vector<uint64_t> nl = live;
                        apply(nl, v);
                        Board c = s;
                        const Profiler::Var &var = pr.vars[v];
                        c[var.y * m.W + var.x] = (uint16_t)(var.piece << 2 | var.o);
                        const Profiler::Prof p{level_of(c), profile(nl)};
                        if (!p.valid() || next.count(c)) continue;
                        next.insert(c);
                        if (next.size() > max_situations) too_large = true;
                        bool leads = solutions.empty() || p.level == m.P;
                        for (auto it = solutions.begin(); !leads && it != solutions.end(); ++it)
                            leads = predecessor(c, *it);
                        update(leads ? lead : other, p);
                    }
            }
            std::swap(cur, next);
        }
        if (solutions.empty()) solutions = cur;
    }
};

}  // namespace

int oracle_profile(uint32_t width, uint32_t height, const uint32_t *colours, uint64_t max_situations,
                   uint32_t *stats, uint16_t *solutions, uint32_t capacity, uint32_t *n_solutions) {
    Model m;
    if (!build_model(width, height, colours, m)) return ORACLE_BAD_PUZZLE;
    if (!stats || !n_solutions) return ORACLE_BAD_ARGUMENT;
    const Profiler pr(m);
    const uint32_t P = m.P;
    ProfileSearch ps(m, pr, max_situations);
    const std::vector<char> live0(pr.vars.size(), 1);
    const Profiler::Prof p0{0, pr.profile(live0)};

    std::vector<Board> starts;
    uint32_t levels = P;
    Board empty(P, 0xFFFF);
    if (P != 25) {
        starts.push_back(empty);
    } else {
        levels = P - 4;
        auto set_corner = [&](Board &b, std::vector<char> &live, uint32_t piece, uint32_t x, uint32_t y) {
            for (uint32_t o = 0; o < 4; ++o) {
                const int32_t v = pr.var_at[((size_t)(y * width + x) * P + piece) * 4 + o];
                if (v >= 0 && live[(size_t)v]) {
                    b[y * width + x] = (uint16_t)(piece << 2 | o);
                    pr.apply(live, (uint32_t)v);
                    break;
                }
            }
        };
        const std::vector<uint32_t> &corners = m.by_kind[2];
        Board c0 = empty;
        std::vector<char> live_c0 = live0;
        set_corner(c0, live_c0, corners[0], 0, 0);
        const uint32_t cx[3] = {width - 1, 0, width - 1}, cy[3] = {0, height - 1, height - 1};
        std::vector<uint32_t> others{1, 2, 3};
        do {
            Board b = c0;
            std::vector<char> live = live_c0;
            for (uint32_t i = 0; i < 3; ++i) set_corner(b, live, corners[others[i]], cx[i], cy[i]);
            starts.push_back(b);
        } while (std::next_permutation(others.begin(), others.end()));
    }

    std::vector<LevelStats> all(P + 1), lead(P + 1), other(P + 1);
    ProfileSearch::update(all, p0);
    std::set<Board> sols;
    for (const Board &s : starts) ps.wide(sols, all, all, s, levels);
    ProfileSearch::update(lead, p0);
    ProfileSearch::update(other, p0);
    for (const Board &s : starts) {
        bool has = false;
        for (auto it = sols.begin(); !has && it != sols.end(); ++it) has = ProfileSearch::predecessor(s, *it);
        if (has) ps.wide(sols, lead, other, s, levels);
    }
    if (ps.too_large) return ORACLE_LIMIT_REACHED;

    const std::vector<LevelStats> *series[3] = {&all, &lead, &other};
    for (uint32_t k = 0; k < 3; ++k)
        for (uint32_t l = 0; l <= P; ++l) {
            const LevelStats &st = (*series[k])[l];
            const uint32_t v[6] = {st.low.get_min(), st.low.get_max(), st.high.get_min(),
                                   st.high.get_max(), st.total.get_min(), st.total.get_max()};
            for (uint32_t s = 0; s < 6; ++s) stats[((size_t)s * 3 + k) * (P + 1) + l] = v[s];
        }
    *n_solutions = (uint32_t)sols.size();
    uint32_t i = 0;
    for (auto it = sols.begin(); it != sols.end() && i < capacity; ++it, ++i)
        if (solutions) std::copy(it->begin(), it->end(), solutions + (size_t)i * P);
    return ORACLE_OK;
}
