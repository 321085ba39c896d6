// emp_enum_plan.cpp -- builds the enumeration table of a puzzle.
#include "emp_enum_plan.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <unordered_set>

namespace emp {

EnumTableHeader EnumTable::header() const {
    EnumTableHeader h;
    std::memcpy(&h, blob.data(), sizeof(h));
    return h;
}

uint32_t natural_row_words(const Puzzle &pz) {
    uint32_t w = 1;
    for (int k = 0; k < 3; ++k) w = std::max<uint32_t>(w, (4 * (uint32_t)pz.pieces_of(k).size() + 63) / 64);
    return w;
}

EnumTable build_enum_table(const Puzzle &pz, int order, uint32_t row_words, uint32_t stack_bytes) {
    EnumTable t;
    EnumTableHeader h{};
    const uint32_t P = pz.size();
    if (P >= kOffBoard) throw std::domain_error("enumeration tables index at most 509 positions");
    if (pz.max_colour() > 255) throw std::domain_error("enumeration tables hold colours up to 255");
    h.P = P;
    h.W = pz.width();
    h.H = pz.height();
    h.ncol = pz.max_colour() + 1;
    t.order = make_order(pz.width(), pz.height(), order);
    if (!is_permutation(t.order, pz.width(), pz.height()))
        throw std::domain_error("the spiral order is not defined for a " + std::to_string(pz.width()) + "x" +
                                std::to_string(pz.height()) + " board (the walk revisits cells when height >= width + 2)");
    uint32_t n_kind[3];
    for (int k = 0; k < 3; ++k) {
        n_kind[k] = (uint32_t)pz.pieces_of(k).size();
        h.words[k] = (4 * n_kind[k] + 63) / 64;
        if (h.words[k] == 0) h.words[k] = 1;
        for (uint32_t idx : pz.pieces_of(k)) t.piece_of[k].push_back((uint16_t)idx);
    }
    h.avail_words = h.words[0] + h.words[1] + h.words[2];
    h.row = std::max(row_words, natural_row_words(pz));
    const uint32_t R = h.row;

    // Appearance classes: kept transitions of the same kind showing the same
    // four colours, recorded when they span at least two pieces.  Trying
    // (A, o) clears the same-looking transitions of the other pieces from the
    // position (HANDLE_IDENTICAL_PIECES select_piece,
    // include/emp_position_strategy.h:189-202, masks of include/emp_piece_db.h
    // :848-884), so among a position's candidates only the lowest piece of a
    // class keeps its transitions.  A half-similar piece is kept at
    // orientations 0 and 2 (the database steps by 2^similarity,
    // include/emp_piece_db.h:1266-1272); both look alike and both stay.
    std::vector<std::vector<uint64_t>> dup[3];
    for (int k = 0; k < 3; ++k) {
        std::vector<uint32_t> look(4 * n_kind[k]);
        std::vector<char> valid(4 * n_kind[k], 0);
        for (uint32_t ki = 0; ki < n_kind[k]; ++ki) {
            const Piece &p = pz.piece(t.piece_of[k][ki]);
            for (unsigned o = 0; o < 4; ++o) {
                look[4 * ki + o] = shown(p, 0, o) | shown(p, 1, o) << 8 | shown(p, 2, o) << 16 | shown(p, 3, o) << 24;
                valid[4 * ki + o] = Puzzle::orientation_kept(p, o);
            }
        }
        // Valid transitions sorted by (look, index): each run of one look is a
        // class, its first entry the lowest transition; classes are listed
        // by their lowest transition.
        std::vector<std::pair<uint32_t, uint32_t>> by_look;
        by_look.reserve(4 * n_kind[k]);
        for (uint32_t a = 0; a < 4 * n_kind[k]; ++a)
            if (valid[a]) by_look.emplace_back(look[a], a);
        std::sort(by_look.begin(), by_look.end());
        std::vector<std::pair<uint32_t, size_t>> first_of;   // (lowest transition, run start)
        for (size_t i = 0; i < by_look.size(); ++i)
            if (i == 0 || by_look[i].first != by_look[i - 1].first) first_of.emplace_back(by_look[i].second, i);
        std::sort(first_of.begin(), first_of.end());
        for (const auto &run : first_of) {
            const uint32_t a = run.first;
            std::vector<uint64_t> mask(R, 0);
            bool several_pieces = false;
            for (size_t i = run.second; i < by_look.size() && by_look[i].first == look[a]; ++i) {
                const uint32_t b = by_look[i].second;
                mask[b >> 6] |= 1ull << (b & 63);
                several_pieces |= (b >> 2) != (a >> 2);
            }
            if (several_pieces) dup[k].push_back(mask);
        }
        h.ndup[k] = (uint32_t)dup[k].size();
    }

    // Layout.
    uint32_t off = kHeaderWords;
    h.off_pos = off;
    off += P;
    h.off_nbr = off;
    off += P * kNbrWords;
    h.off_base = off;
    off += P * R;
    h.off_ones = off;
    off += 32;
    for (int k = 0; k < 3; ++k) {
        h.off_cm[k] = off;
        off += 4 * h.ncol * R;
        h.off_valid[k] = off;
        off += R;
        h.off_dup[k] = off;
        off += h.ndup[k] * R + R;
        h.off_cls[k] = off;
        off += (4 * n_kind[k] + 3) / 4;
        h.off_rot[k] = off;
        off += (4 * n_kind[k] + 1) / 2;
    }
    // Identical-piece chains: the pieces an appearance class spans, lowest
    // first.  Pieces showing one look show the same kept looks (they are
    // rotations of each other with the same symmetry), so every class of a
    // piece names the same chain.
    std::vector<int> succ[3], pred[3];
    for (int k = 0; k < 3; ++k) {
        succ[k].assign(n_kind[k], -1);
        pred[k].assign(n_kind[k], -1);
        for (const std::vector<uint64_t> &mask : dup[k]) {
            std::vector<int> pieces;
            for (uint32_t w = 0; w < R; ++w)
                for (uint64_t v = mask[w]; v; v &= v - 1) {
                    const int p = (int)((w * 64 + __builtin_ctzll(v)) >> 2);
                    if (pieces.empty() || pieces.back() != p) pieces.push_back(p);
                }
            for (size_t i = 0; i + 1 < pieces.size(); ++i) {
                const int p = pieces[i], q = pieces[i + 1];
                if ((succ[k][p] >= 0 && succ[k][p] != q) || (pred[k][q] >= 0 && pred[k][q] != p))
                    throw std::logic_error("appearance classes of one piece span different pieces");
                succ[k][p] = q;
                pred[k][q] = p;
            }
        }
    }
    // Offsets of the kinds' words in a lane's available-transition region:
    // corners first, interior last, so that reads of `row` words at any
    // kind's offset stay within words[2] + words[1] + row words.
    const uint32_t abase_of[3] = {h.words[2] + h.words[1], h.words[2], 0};
    if (h.avail_words * 16 > 4096) throw std::domain_error("identical-piece chains index at most 4096 piece slots");
    std::vector<uint32_t> pred_list;
    for (int k = 0; k < 3; ++k)   // kinds by decreasing offset: highest slot first
        for (int ki = (int)n_kind[k] - 1; ki >= 0; --ki)
            if (pred[k][ki] >= 0)
                pred_list.push_back((abase_of[k] * 16 + (uint32_t)ki) | (abase_of[k] * 16 + (uint32_t)pred[k][ki]) << 16);
    h.off_succ = off;
    off += (h.avail_words * 16 * 2 + 7) / 8;
    h.off_pred = off;
    h.npred = (uint32_t)pred_list.size();
    off += (h.npred * 4 + 7) / 8;

    uint32_t col_bytes[3], col_total = 0;
    for (int k = 0; k < 3; ++k) {
        col_bytes[k] = col_total;
        col_total += 4 * 4 * n_kind[k];
    }
    if (col_total > 0xFFFF) throw std::domain_error("colour rows exceed 64 KiB");
    h.off_col = off;
    off += (col_total + 7) / 8;
    h.total_words = off;
    t.blob.assign(off, 0);
    uint64_t *b = t.blob.data();

    // Positions.
    std::vector<uint32_t> index_of(P);
    for (uint32_t i = 0; i < P; ++i) index_of[t.order[i].second * h.W + t.order[i].first] = i;
    std::vector<int> kind_at(P);
    for (uint32_t i = 0; i < P; ++i) kind_at[i] = pz.position_kind(t.order[i].first, t.order[i].second);
    for (uint32_t i = 0; i < P; ++i) {
        uint32_t x = t.order[i].first, y = t.order[i].second;
        uint64_t rec = (uint64_t)kind_at[i];
        for (unsigned s = 0; s < 4; ++s) {
            int nx = (int)x + (s == 1) - (s == 3), ny = (int)y + (s == 2) - (s == 0);
            uint32_t code, nk = 0;
            if (nx < 0 || ny < 0 || nx >= (int)h.W || ny >= (int)h.H) code = kOffBoard;
            else {
                uint32_t j = index_of[ny * h.W + nx];
                if (j < i) { code = j; nk = (uint32_t)kind_at[j]; }
                else code = kUnconstrained;
            }
            rec |= (uint64_t)code << (2 + 11 * s);
            rec |= (uint64_t)nk << (11 + 11 * s);
        }
        b[h.off_pos + i] = rec;
    }

    // Colour masks, valid-orientation masks, classes, rotated colours.
    for (int k = 0; k < 3; ++k) {
        uint64_t *cm = b + h.off_cm[k];
        uint64_t *valid = b + h.off_valid[k];
        uint32_t *rot = reinterpret_cast<uint32_t *>(b + h.off_rot[k]);
        for (uint32_t ki = 0; ki < n_kind[k]; ++ki) {
            const Piece &p = pz.piece(t.piece_of[k][ki]);
            for (unsigned o = 0; o < 4; ++o) {
                uint32_t tr = 4 * ki + o;
                rot[tr] = shown(p, 0, o) | shown(p, 1, o) << 8 | shown(p, 2, o) << 16 | shown(p, 3, o) << 24;
                if (!Puzzle::orientation_kept(p, o)) continue;
                valid[tr >> 6] |= 1ull << (tr & 63);
                for (unsigned s = 0; s < 4; ++s) {
                    uint32_t c = shown(p, s, o);
                    cm[((size_t)s * h.ncol + c) * R + (tr >> 6)] |= 1ull << (tr & 63);
                }
            }
        }
        for (uint32_t d = 0; d < h.ndup[k]; ++d)
            for (uint32_t w = 0; w < R; ++w) b[h.off_dup[k] + d * R + w] = dup[k][d][w];
        uint16_t *cls = reinterpret_cast<uint16_t *>(b + h.off_cls[k]);
        for (uint32_t tr = 0; tr < 4 * n_kind[k]; ++tr) cls[tr] = 0xFFFF;
        for (uint32_t d = 0; d < h.ndup[k]; ++d)
            for (uint32_t w = 0; w < R; ++w) {
                b[h.off_dup[k] + h.ndup[k] * R + w] |= dup[k][d][w];
                for (uint64_t v = dup[k][d][w]; v; v &= v - 1) cls[w * 64 + __builtin_ctzll(v)] = (uint16_t)d;
            }
    }

    {
        uint16_t *sc = reinterpret_cast<uint16_t *>(b + h.off_succ);
        for (int k = 0; k < 3; ++k)
            for (uint32_t ki = 0; ki < n_kind[k]; ++ki) {
                if (succ[k][ki] < 0) continue;
                const uint32_t q = (uint32_t)succ[k][ki];
                const uint32_t vn = (uint32_t)(b[h.off_valid[k] + (4 * q >> 6)] >> (4 * q & 63)) & 0xF;
                sc[abase_of[k] * 16 + ki] = (uint16_t)((abase_of[k] * 16 + q) | vn << 12);
            }
        uint32_t *pl = reinterpret_cast<uint32_t *>(b + h.off_pred);
        for (uint32_t i = 0; i < h.npred; ++i) pl[i] = pred_list[i];
    }

    for (uint32_t w = 0; w < 16; ++w) b[h.off_ones + w] = ~0ull;
    const uint32_t sb = stack_bytes ? stack_bytes : 1;
    h.stack_bytes = sb;
    {
        uint8_t *col = reinterpret_cast<uint8_t *>(b + h.off_col);
        for (int k = 0; k < 3; ++k)
            for (uint32_t ki = 0; ki < n_kind[k]; ++ki) {
                const Piece &p = pz.piece(t.piece_of[k][ki]);
                for (unsigned o = 0; o < 4; ++o)
                    for (unsigned q = 0; q < 4; ++q)
                        col[col_bytes[k] + q * 4 * n_kind[k] + 4 * ki + o] = (uint8_t)shown(p, q, o);
            }
    }

    // Base masks (valid transitions matching the frame sides) and neighbour
    // descriptors.
    for (uint32_t i = 0; i < P; ++i) {
        const uint64_t rec = b[h.off_pos + i];
        const int k = (int)(rec & 3);
        uint64_t *base = b + h.off_base + (size_t)i * R;
        for (uint32_t w = 0; w < R; ++w) base[w] = b[h.off_valid[k] + w];
        uint64_t *rec_out = b + h.off_nbr + (size_t)i * kNbrWords;
        uint32_t n = 0;
        for (unsigned s = 0; s < 4; ++s) {
            const uint32_t code = (uint32_t)(rec >> (2 + 11 * s)) & 0x1FF;
            if (code == kUnconstrained) continue;
            if (code == kOffBoard) {
                const uint64_t *m = b + h.off_cm[k] + ((size_t)s * h.ncol + 0) * R;
                for (uint32_t w = 0; w < R; ++w) base[w] &= m[w];
                continue;
            }
            const uint32_t nk = (uint32_t)(rec >> (11 + 11 * s)) & 3;
            const uint32_t facing = (s + 2) & 3;
            const uint64_t stk_off = (uint64_t)code * 64 * sb;
            const uint64_t col_off = col_bytes[nk] + facing * 4 * n_kind[nk];
            const uint64_t cm_block = h.off_cm[k] + (uint64_t)s * h.ncol * R;
            if (stk_off > 0xFFFF || cm_block >= (1u << 29)) throw std::domain_error("table offsets exceed their fields");
            rec_out[1 + n] = stk_off | col_off << 16 | (cm_block * 8) << 32;
            ++n;
        }
        const uint64_t abase = abase_of[k];
        const uint64_t uni = h.off_dup[k] + (uint64_t)h.ndup[k] * R;
        rec_out[0] = uni << 32 | abase << 8 | (uint64_t)k << 3 | n;
    }

    // Root: first candidate of index 0, where no neighbour is placed yet.
    {
        uint64_t rec = b[h.off_pos];
        int k = (int)(rec & 3);
        std::vector<uint64_t> c(b + h.off_valid[k], b + h.off_valid[k] + R);
        for (unsigned s = 0; s < 4; ++s) {
            uint32_t code = (uint32_t)(rec >> (2 + 11 * s)) & 0x1FF;
            if (code != kOffBoard) continue;
            const uint64_t *m = b + h.off_cm[k] + ((size_t)s * h.ncol + 0) * R;
            for (uint32_t w = 0; w < R; ++w) c[w] &= m[w];
        }
        int root = -1;
        for (uint32_t w = 0; w < R && root < 0; ++w)
            if (c[w]) root = (int)(w * 64 + __builtin_ctzll(c[w]));
        if (root < 0) throw std::invalid_argument("no piece fits the first position");
        h.root = (uint32_t)root;
    }
    std::memcpy(b, &h, sizeof(h));
    return t;
}

void stack_to_board(const EnumTable &t, const uint16_t *stack, uint16_t *board) {
    EnumTableHeader h = t.header();
    for (uint32_t i = 0; i < h.P; ++i) board[i] = 0xFFFF;
    for (uint32_t i = 0; i < h.P; ++i) {
        int k = (int)(t.blob[h.off_pos + i] & 3);
        uint32_t tr = stack[i];
        uint32_t piece = t.piece_of[k][tr >> 2];
        board[t.order[i].second * h.W + t.order[i].first] = (uint16_t)((piece << 2) | (tr & 3));
    }
}

uint16_t move_to_transition(const EnumTable &t, const Puzzle &pz, uint32_t index, uint32_t mv) {
    const EnumTableHeader h = t.header();
    if (index >= h.P) throw std::invalid_argument("strategy index out of range");
    const uint32_t pos = mv & 0xFF, piece = (mv >> 8) & 0xFF, o = (mv >> 16) & 3;
    if (pos != t.order[index].second * h.W + t.order[index].first)
        throw std::invalid_argument("move is not at strategy index " + std::to_string(index));
    const int K = (int)(t.blob[h.off_pos + index] & 3);
    for (uint32_t ki = 0; ki < t.piece_of[K].size(); ++ki)
        if (t.piece_of[K][ki] == piece) return (uint16_t)(4 * ki + o);
    (void)pz;
    throw std::invalid_argument("move at strategy index " + std::to_string(index) + ": piece of another kind");
}

uint32_t transition_to_move(const EnumTable &t, uint32_t index, uint32_t tr) {
    const EnumTableHeader h = t.header();
    const int K = (int)(t.blob[h.off_pos + index] & 3);
    const uint32_t piece = t.piece_of[K].at(tr >> 2);
    const uint32_t pos = t.order[index].second * h.W + t.order[index].first;
    return pos | (piece << 8) | ((tr & 3) << 16);
}

void moves_to_stack(const EnumTable &t, const Puzzle &pz, const uint32_t *moves, uint32_t len,
                    std::vector<uint16_t> &stack) {
    const EnumTableHeader h = t.header();
    const uint64_t *b = t.blob.data();
    const uint32_t R = h.row;
    if (len > h.P) throw std::invalid_argument("more moves than positions");
    std::vector<int> kind_of(pz.size(), -1), index_of(pz.size(), -1);
    for (int k = 0; k < 3; ++k)
        for (uint32_t ki = 0; ki < t.piece_of[k].size(); ++ki) {
            kind_of[t.piece_of[k][ki]] = k;
            index_of[t.piece_of[k][ki]] = (int)ki;
        }
    std::vector<uint64_t> avail[3];
    for (int k = 0; k < 3; ++k) avail[k].assign(b + h.off_valid[k], b + h.off_valid[k] + R);
    const uint8_t *colour = reinterpret_cast<const uint8_t *>(b + h.off_col);
    stack.assign(len, 0);
    for (uint32_t i = 0; i < len; ++i) {
        const uint32_t pos = moves[i] & 0xFF, piece = (moves[i] >> 8) & 0xFF, o = (moves[i] >> 16) & 3;
        const uint32_t x = t.order[i].first, y = t.order[i].second;
        if (pos != y * h.W + x) throw std::invalid_argument("move " + std::to_string(i) + " is not at strategy index " + std::to_string(i));
        const uint64_t *rec = b + h.off_nbr + (size_t)i * kNbrWords;
        const int K = (int)(rec[0] >> 3) & 3;
        if (piece >= pz.size() || kind_of[piece] != K) throw std::invalid_argument("move " + std::to_string(i) + ": piece of another kind");
        const uint32_t tr = 4 * (uint32_t)index_of[piece] + o;
        std::vector<uint64_t> c(R);
        for (uint32_t w = 0; w < R; ++w) c[w] = b[h.off_base + (size_t)i * R + w] & avail[K][w];
        for (uint32_t q = 0, n = (uint32_t)(rec[0] & 7); q < n; ++q) {
            const uint64_t e = rec[1 + q];
            const uint32_t j = ((uint32_t)e & 0xFFFF) / (64 * h.stack_bytes);
            const uint32_t col = colour[((uint32_t)e >> 16) + stack[j]];
            const uint64_t *row = b + (e >> 35) + (size_t)col * R;
            for (uint32_t w = 0; w < R; ++w) c[w] &= row[w];
        }
        for (uint32_t d = 0; d < h.ndup[K]; ++d) {
            const uint64_t *dm = b + h.off_dup[K] + (size_t)d * R;
            bool seen = false;
            for (uint32_t w = 0; w < R; ++w) {
                const uint64_t m = c[w] & dm[w];
                const uint64_t low = m ? (0xFull << (__builtin_ctzll(m) & 60)) : 0ull;
                c[w] &= ~(m & ~(seen ? 0ull : low));
                seen |= m != 0;
            }
        }
        if (!((c[tr >> 6] >> (tr & 63)) & 1))
            throw std::invalid_argument("move " + std::to_string(i) + " is not a candidate at its index");
        stack[i] = (uint16_t)tr;
        avail[K][tr >> 6] &= ~(0xFull << (tr & 60));
    }
}

void spiral_prefixes(const Puzzle &pz, uint32_t length, uint32_t count, uint64_t seed,
                     const std::vector<uint32_t> &fixed, std::vector<uint32_t> &rows, uint32_t &produced) {
    if (pz.size() > 256) throw std::domain_error("prefix moves pack positions and pieces in 8 bits");
    const EnumTable t = build_enum_table(pz, kSpiral);
    const EnumTableHeader h = t.header();
    const uint64_t *b = t.blob.data();
    std::vector<char> fixed_piece(pz.size(), 0), fixed_cell(pz.size(), 0);
    for (uint32_t mv : fixed) {
        if ((mv & 0xFF) >= pz.size() || ((mv >> 8) & 0xFF) >= pz.size()) throw std::invalid_argument("bad fixed move");
        fixed_cell[mv & 0xFF] = 1;
        fixed_piece[(mv >> 8) & 0xFF] = 1;
    }
    uint64_t state = seed;
    std::unordered_set<std::string> seen;
    const uint32_t width = (uint32_t)fixed.size() + length;
    rows.clear();
    produced = 0;
    std::vector<uint16_t> stack(length);
    const uint64_t attempts = (uint64_t)count * 8 + 64;
    for (uint64_t attempt = 0; attempt < attempts && produced < count; ++attempt) {
        // Available transitions per kind, minus the fixed pieces.
        std::vector<uint64_t> avail[3];
        for (int k = 0; k < 3; ++k) {
            avail[k].assign(b + h.off_valid[k], b + h.off_valid[k] + h.row);
            for (uint32_t ki = 0; ki < t.piece_of[k].size(); ++ki)
                if (fixed_piece[t.piece_of[k][ki]]) avail[k][(4 * ki) >> 6] &= ~(0xFull << ((4 * ki) & 63));
        }
        bool ok = true;
        for (uint32_t i = 0; i < length && ok; ++i) {
            const uint64_t rec = b[h.off_pos + i];
            const int k = (int)(rec & 3);
            const uint32_t x = t.order[i].first, y = t.order[i].second;
            if (fixed_cell[y * h.W + x]) { ok = false; break; }
            std::vector<uint64_t> c(avail[k]);
            for (unsigned s = 0; s < 4; ++s) {
                const uint32_t code = (uint32_t)(rec >> (2 + 11 * s)) & 0x1FF;
                if (code == kUnconstrained) continue;
                uint32_t col = 0;
                if (code != kOffBoard) {
                    const int nk = (int)(rec >> (11 + 11 * s)) & 3;
                    col = (reinterpret_cast<const uint32_t *>(b + h.off_rot[nk])[stack[code]] >> (8 * ((s + 2) & 3))) & 0xFF;
                }
                const uint64_t *m = b + h.off_cm[k] + ((size_t)s * h.ncol + col) * h.row;
                for (uint32_t w = 0; w < h.row; ++w) c[w] &= m[w];
            }
            uint32_t n = 0;
            for (uint64_t v : c) n += (uint32_t)__builtin_popcountll(v);
            if (!n) { ok = false; break; }
            uint32_t pick = (uint32_t)(((splitmix64(state) >> 32) * (uint64_t)n) >> 32);
            uint32_t tr = 0;
            for (uint32_t w = 0; w < h.row; ++w) {
                uint64_t v = c[w];
                const uint32_t pc = (uint32_t)__builtin_popcountll(v);
                if (pick >= pc) { pick -= pc; continue; }
                while (pick--) v &= v - 1;
                tr = w * 64 + (uint32_t)__builtin_ctzll(v);
                break;
            }
            stack[i] = (uint16_t)tr;
            avail[k][(tr & ~3u) >> 6] &= ~(0xFull << ((tr & ~3u) & 63));
        }
        if (!ok) continue;
        std::string key(reinterpret_cast<const char *>(stack.data()), length * sizeof(uint16_t));
        if (!seen.insert(key).second) continue;
        for (uint32_t mv : fixed) rows.push_back(mv);
        for (uint32_t i = 0; i < length; ++i) {
            const int k = (int)(b[h.off_pos + i] & 3);
            const uint32_t piece = t.piece_of[k][stack[i] >> 2];
            const uint32_t pos = t.order[i].second * h.W + t.order[i].first;
            rows.push_back(pos | (piece << 8) | ((uint32_t)(stack[i] & 3) << 16));
        }
        ++produced;
    }
    (void)width;
}

}  // namespace emp
