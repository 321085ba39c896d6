// emp_glutton_stack.cpp -- XML form of a glutton search's stack (see
// emp_glutton_stack.h for the layout and the reference file it follows).
#include "emp_glutton_stack.h"

#include <cctype>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <utility>
#include <vector>

#include "emp_glutton.h"

namespace emp {

namespace {

// ---------------------------------------------------------------------------
// A small XML reader: elements, attributes, text; declarations, comments and
// the five predefined entities.  Enough for the converter's files.
// ---------------------------------------------------------------------------
struct Node {
    std::string name, text;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<Node> children;

    const std::string *attr(const char *n) const {
        for (const auto &a : attrs)
            if (a.first == n) return &a.second;
        return nullptr;
    }
    const Node *child(const char *n) const {
        for (const Node &c : children)
            if (c.name == n) return &c;
        return nullptr;
    }
};

class Reader {
  public:
    explicit Reader(const std::string &s) : s_(s) {}

    Node document() {
        skip_misc();
        if (!peek('<')) fail("no root element");
        Node root = element();
        skip_misc();
        if (i_ != s_.size()) fail("text after the root element");
        return root;
    }

  private:
    [[noreturn]] void fail(const std::string &what) const {
        throw std::invalid_argument("XML: " + what + " at byte " + std::to_string(i_));
    }
    bool peek(char c) const { return i_ < s_.size() && s_[i_] == c; }
    bool starts(const char *p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
    void space() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    void skip_until(const char *end) {
        const size_t e = s_.find(end, i_);
        if (e == std::string::npos) fail(std::string("unterminated markup, no \"") + end + "\"");
        i_ = e + std::strlen(end);
    }
    // declarations, comments, DOCTYPE and white space between elements
    void skip_misc() {
        for (;;) {
            space();
            if (starts("<?")) skip_until("?>");
            else if (starts("<!--")) skip_until("-->");
            else if (starts("<!")) skip_until(">");
            else return;
        }
    }
    std::string name() {
        const size_t b = i_;
        while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == '-' ||
                                  s_[i_] == ':' || s_[i_] == '.'))
            ++i_;
        if (b == i_) fail("a name was expected");
        return s_.substr(b, i_ - b);
    }
    std::string decode(const std::string &raw) const {
        std::string out;
        for (size_t k = 0; k < raw.size(); ++k) {
            if (raw[k] != '&') {
                out += raw[k];
                continue;
            }
            static const std::pair<const char *, char> kEnt[5] = {
                {"&lt;", '<'}, {"&gt;", '>'}, {"&amp;", '&'}, {"&quot;", '"'}, {"&apos;", '\''}};
            bool done = false;
            for (const auto &e : kEnt)
                if (raw.compare(k, std::strlen(e.first), e.first) == 0) {
                    out += e.second;
                    k += std::strlen(e.first) - 1;
                    done = true;
                    break;
                }
            if (!done) throw std::invalid_argument("XML: unknown entity in \"" + raw + "\"");
        }
        return out;
    }
    Node element() {
        Node n;
        ++i_;   // '<'
        n.name = name();
        for (;;) {
            space();
            if (starts("/>")) {
                i_ += 2;
                return n;
            }
            if (peek('>')) {
                ++i_;
                break;
            }
            std::string an = name();
            space();
            if (!peek('=')) fail("'=' expected after attribute " + an);
            ++i_;
            space();
            if (!peek('"') && !peek('\'')) fail("quoted value expected for attribute " + an);
            const char q = s_[i_++];
            const size_t e = s_.find(q, i_);
            if (e == std::string::npos) fail("unterminated value of attribute " + an);
            n.attrs.emplace_back(an, decode(s_.substr(i_, e - i_)));
            i_ = e + 1;
        }
        for (;;) {
            const size_t lt = s_.find('<', i_);
            if (lt == std::string::npos) fail("element " + n.name + " is not closed");
            n.text += decode(s_.substr(i_, lt - i_));
            i_ = lt;
            if (starts("</")) {
                i_ += 2;
                if (name() != n.name) fail("closing tag does not match " + n.name);
                space();
                if (!peek('>')) fail("'>' expected");
                ++i_;
                return n;
            }
            if (starts("<!--")) skip_until("-->");
            else if (starts("<?")) skip_until("?>");
            else n.children.push_back(element());
        }
    }

    const std::string &s_;
    size_t i_ = 0;
};

std::string trim(const std::string &t) {
    const size_t b = t.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return "";
    return t.substr(b, t.find_last_not_of(" \t\r\n") - b + 1);
}

// ---------------------------------------------------------------------------
// The converter's elements (messages of its loader, the file named "the
// stack").
// ---------------------------------------------------------------------------
uint32_t attribute(const Node &n, const char *a) {
    const std::string *v = n.attr(a);
    if (!v) throw std::invalid_argument("Missing " + std::string(a) + " attribute in " + n.name + " node of the stack");
    try {
        size_t used = 0;
        const unsigned long x = std::stoul(*v, &used, 0);
        if (used != v->size() || x > 0xFFFFFFFFul) throw std::invalid_argument("range");
        return (uint32_t)x;
    } catch (const std::exception &) {
        throw std::invalid_argument("attribute " + std::string(a) + " of " + n.name + " is not a number: \"" + *v + "\"");
    }
}

void check(const Node &n, const char *a, uint32_t expected, const std::string &owner) {
    const uint32_t v = attribute(n, a);
    if (v != expected)
        throw std::invalid_argument(std::string(a) + " from XML (" + std::to_string(v) + ") is different from " + a +
                                    " of " + owner + " (" + std::to_string(expected) + ")");
}

uint32_t word(const Node &n, const std::string &where) {
    const std::string t = trim(n.text);
    if (t.empty()) throw std::invalid_argument("Empty text in node " + where + " of the stack");
    try {
        size_t used = 0;
        const unsigned long x = std::stoul(t, &used, 0);
        if (used != t.size() || x > 0xFFFFFFFFul) throw std::invalid_argument("range");
        return (uint32_t)x;
    } catch (const std::exception &) {
        throw std::invalid_argument("node " + where + " of the stack does not hold a 32-bit word: \"" + t + "\"");
    }
}

// children [0, size) of a named array, each an item with index = its rank
void load_array(const Node &root, const char *name, uint32_t size, uint32_t *dst) {
    const Node *arr = root.child(name);
    if (!arr) throw std::invalid_argument("No node \"" + std::string(name) + "\" in node \"" + root.name + "\" of the stack");
    check(*arr, "size", size, std::string("node ") + name);
    if (arr->children.size() < size)
        throw std::invalid_argument("No node for item " + std::to_string(arr->children.size()) + " in node \"" + name +
                                    "\" of the stack");
    for (uint32_t k = 0; k < size; ++k) {
        const Node &it = arr->children[k];
        check(it, "index", k, std::string(name) + " item");
        dst[k] = word(it, std::string(name) + "[" + std::to_string(k) + "]");
    }
}

std::string hex(uint32_t v) {
    char b[16];
    std::snprintf(b, sizeof b, "0x%08x", v);
    return b;
}

void put_array(std::string &o, const char *name, uint32_t size, const uint32_t *v) {
    o += "  <";
    o += name;
    o += " size=\"" + std::to_string(size) + "\">\n";
    for (uint32_t k = 0; k < size; ++k) o += "    <item index=\"" + std::to_string(k) + "\">" + hex(v[k]) + "</item>\n";
    o += "  </";
    o += name;
    o += ">\n";
}

}  // namespace

std::string glutton_stack_to_xml(const Puzzle &pz, const uint32_t *st, const uint64_t *mk) {
    validate_glutton_stack(pz, st, mk);
    const uint32_t P = pz.size(), S = st[0], level = st[2], max = st[3], base = st[4];
    const uint32_t nw = (4 * P + 63) / 64;
    std::string o;
    o.reserve((size_t)S * (S + 1) / 2 * 32 * 40 + 4096);
    o += "<CUDA_glutton_stack size=\"" + std::to_string(S) + "\" level=\"" + std::to_string(level) + "\" nb_pieces=\"" +
         std::to_string(P) + "\" max=\"" + std::to_string(max) + "\" base=\"" + std::to_string(base) + "\">\n";
    put_array(o, "info_to_position", S, st + 5);
    put_array(o, "position_to_info", P, st + 5 + S);
    put_array(o, "played_info", S, st + 5 + S + P);
    put_array(o, "max_played_info", max, st + 5 + 2 * S + P);
    put_array(o, "available_pieces", 8, st + 5 + 3 * S + P);
    o += "  <position_infos>\n";
    for (uint32_t l = 0; l < S; ++l) {
        o += "    <level index=\"" + std::to_string(l) + "\">\n";
        const uint64_t off = ((uint64_t)S * (S + 1) - (uint64_t)(S - l) * (S - l + 1)) / 2 * nw;
        for (uint32_t j = 0; j < S - l; ++j) {
            o += "      <position_info index=\"" + std::to_string(j) + "\">\n";
            for (uint32_t w = 0; w < 32; ++w) {
                const uint64_t v = w / 2 < nw ? mk[off + (uint64_t)j * nw + w / 2] : 0ull;
                o += "        <word index=\"" + std::to_string(w) + "\">" + hex((uint32_t)(v >> (32 * (w & 1)))) +
                     "</word>\n";
            }
            o += "      </position_info>\n";
        }
        o += "    </level>\n";
    }
    o += "  </position_infos>\n</CUDA_glutton_stack>\n";
    return o;
}

void glutton_stack_from_xml(const Puzzle &pz, const std::string &xml, uint32_t *st, uint64_t *mk) {
    const Node root = Reader(xml).document();
    if (root.name != "CUDA_glutton_stack")
        throw std::invalid_argument("the root element is " + root.name + ", not CUDA_glutton_stack");
    const uint32_t P = pz.size(), S = P, nw = (4 * P + 63) / 64;
    std::memset(st, 0, (size_t)glutton_state_words(P) * 4);
    std::memset(mk, 0, (size_t)glutton_mask_words(P) * 8);
    check(root, "size", S, "stack");
    check(root, "nb_pieces", P, "stack");
    st[0] = S;
    st[1] = P;
    const uint32_t level = attribute(root, "level"), max = attribute(root, "max");
    if (level > S || max > S) throw std::invalid_argument("level or max past the stack size");
    st[2] = level;
    st[3] = max;
    st[4] = root.attr("base") ? attribute(root, "base") : 0u;
    load_array(root, "info_to_position", S, st + 5);
    load_array(root, "position_to_info", P, st + 5 + S);
    load_array(root, "played_info", S, st + 5 + S + P);
    load_array(root, "max_played_info", max, st + 5 + 2 * S + P);
    for (uint32_t l = max; l < S; ++l) st[5 + 2 * S + P + l] = 0xFFFFFFFFu;   // never reached
    load_array(root, "available_pieces", 8, st + 5 + 3 * S + P);
    const Node *infos = root.child("position_infos");
    if (!infos || infos->children.empty()) throw std::invalid_argument("Node position infos of the stack is empty");
    // levels 0..level; deeper ones hold nothing a resumed search reads
    for (uint32_t l = 0; l <= level && l < S; ++l) {
        if (infos->children.size() <= l)
            throw std::invalid_argument("Node level[" + std::to_string(l) + "] of the stack is empty");
        const Node &ln = infos->children[l];
        check(ln, "index", l, "level");
        if (ln.children.size() < S - l)
            throw std::invalid_argument("Node position_info[" + std::to_string(ln.children.size()) + "] of level " +
                                        std::to_string(l) + " of the stack is empty");
        const uint64_t off = ((uint64_t)S * (S + 1) - (uint64_t)(S - l) * (S - l + 1)) / 2 * nw;
        for (uint32_t j = 0; j < S - l; ++j) {
            const Node &pn = ln.children[j];
            check(pn, "index", j, "position_info");
            if (pn.children.size() < 32)
                throw std::invalid_argument("Node word[" + std::to_string(pn.children.size()) + "] in position_info[" +
                                            std::to_string(j) + "] of the stack is empty");
            for (uint32_t w = 0; w < 32; ++w) {
                const Node &wn = pn.children[w];
                check(wn, "index", w, "word");
                const uint32_t v = word(wn, "word[" + std::to_string(w) + "]");
                if (w / 2 >= nw) {
                    if (v) throw std::invalid_argument("position_info bit past the last piece");
                    continue;
                }
                mk[off + (uint64_t)j * nw + w / 2] |= (uint64_t)v << (32 * (w & 1));
            }
        }
    }
    validate_glutton_stack(pz, st, mk);
}

}  // namespace emp
