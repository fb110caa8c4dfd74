// frep_compile.hip -- expression DAG -> tape compiler behind frep_compile()
// (host code only).
//
// Produces the tape csg::Tape::Tape(csg::Expr) builds (vulkan/src/csg/
// tape.cpp:14-21) for the same DAG:
//   1. ConstantFold (vulkan/src/csg/simplifier.cpp:31-104): bottom-up, each
//      rebuilt node gets one fold step -- all-constant operations evaluated,
//      identities (x+0, x*1, x*0, 0-x, x/1, x/c -> x*(1/c), --x, -(a-b) ->
//      b-a, ...) rewritten;
//   2. MergeDuplicates (simplifier.cpp:236-259), after NormalForm
//      (simplifier.cpp:183-206): operands of ADD, MUL, MIN and MAX ordered
//      by ThreeWayCompare (simplifier.cpp:133-180), then structurally equal
//      nodes merged bottom-up;
//   3. depth-first post-order of the result, inputs in order
//      (vulkan/src/csg/expression.cpp:298-314), constant pool by first
//      appearance (tape.cpp:44-55), liveness = last reader (tape.cpp:66-79),
//      slots 0..3 holding x, y, z, t, inputs released before the output
//      takes the lowest free slot (tape.cpp:151-228).
//
// The DAG arrives as a node list; node identity is the list index, so the
// sharing of the caller's DAG is kept exactly (two CONST nodes of equal
// value are distinct until merged, as two csg::Expr(float) are).
//
// Points where the result follows the reference's code rather than its
// intent, kept so the tapes are the reference's:
//   - ADD of two NEG operands becomes NEG(ADD(NEG a, NEG b))
//     (simplifier.cpp:51-53), which evaluates to a + b, not -(a + b);
//   - two-operand sorting swaps equal operands: libstdc++'s std::sort on two
//     elements is an insertion sort that moves the second element first when
//     comp(second, first) holds, and the comparator is ThreeWayCompare <= 0.
// Where the reference asserts, this compiler returns a valid tape instead:
//   - a root that is a bare y, z or t axis gets a COPY into slot 0 (a bare x
//     root is the empty tape);
//   - more than 256 constants produce a wide-format tape (CONST index in the
//     A and B bytes);
//   - constant division by zero folds to the IEEE result.
// NaN constants given in the node list are rejected.  A NaN that constant
// folding produces is kept as the reference keeps it -- never merged with
// another node, one pool entry per occurrence -- but its CONST reads its own
// entry, where the reference's constant map cannot find a NaN key and reads
// entry 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/frep.h"
#include "frep_device.hpp"

namespace {

// csg::Operator order (vulkan/include/csg/expression.h:14-22), which is
// also ThreeWayCompare's OpOrder (simplifier.cpp:109-130).
enum EOp : uint32_t {
    E_X = 0, E_Y, E_Z, E_T, E_CONST,
    E_SIN, E_COS, E_EXP, E_NEG, E_SQRT,
    E_ADD, E_SUB, E_MUL, E_DIV, E_MIN, E_MAX, E_COUNT
};

inline uint32_t arity(uint32_t op) { return op <= E_CONST ? 0u : (op <= E_SQRT ? 1u : 2u); }
inline bool is_input_op(uint32_t op) { return op > E_CONST && op < E_COUNT; }

struct Node {
    uint32_t op;
    uint32_t in[2];
    float c;
};

constexpr uint32_t kNone = 0xFFFFFFFFu;

struct Dag {
    std::vector<Node> n;
    uint32_t make(uint32_t op, uint32_t a = kNone, uint32_t b = kNone) {
        n.push_back(Node{op, {a, b}, 0.0f});
        return (uint32_t)n.size() - 1;
    }
    uint32_t constant(float v) {
        n.push_back(Node{E_CONST, {kNone, kNone}, v});
        return (uint32_t)n.size() - 1;
    }
    bool is_const(uint32_t e) const { return n[e].op == E_CONST; }
    bool is_const(uint32_t e, float v) const { return n[e].op == E_CONST && n[e].c == v; }
    bool is_op(uint32_t e, uint32_t op) const { return n[e].op == op; }
    uint32_t in(uint32_t e, int i) const { return n[e].in[i]; }
};

// Depth-first post-order from root, inputs in order, each node once
// (expression.cpp:298-314), without recursion.
std::vector<uint32_t> topo(const Dag& d, uint32_t root) {
    std::vector<uint32_t> order;
    std::vector<uint8_t> seen(d.n.size(), 0);
    std::vector<std::pair<uint32_t, uint32_t>> st;   // (node, next input)
    st.push_back({root, 0});
    seen[root] = 1;
    while (!st.empty()) {
        auto& top = st.back();
        const uint32_t e = top.first;
        const uint32_t k = arity(d.n[e].op);
        if (top.second < k) {
            const uint32_t c = d.n[e].in[top.second++];
            if (!seen[c]) {
                seen[c] = 1;
                st.push_back({c, 0});
            }
            continue;
        }
        order.push_back(e);
        st.pop_back();
    }
    return order;
}

// csg::ApplyOperator (expression.cpp:113-146) with this library's
// transcendentals and GLM's min / max.
float apply(uint32_t op, float a, float b) {
    switch (op) {
    case E_SIN: return frep::dev_sinf(a);
    case E_COS: return frep::dev_cosf(a);
    case E_EXP: return frep::dev_expf(a);
    case E_NEG: return -a;
    case E_SQRT: return std::sqrt(a);
    case E_ADD: return a + b;
    case E_SUB: return a - b;
    case E_MUL: return a * b;
    case E_DIV: return a / b;
    case E_MIN: return b < a ? b : a;
    case E_MAX: return a < b ? b : a;
    default: return 0.0f;
    }
}

// One fold step on a freshly built node (simplifier.cpp:31-94).
uint32_t fold_step(Dag& d, uint32_t e) {
    const uint32_t op = d.n[e].op;
    if (!is_input_op(op)) return e;
    const uint32_t k = arity(op);
    const uint32_t a = d.in(e, 0), b = k > 1 ? d.in(e, 1) : kNone;
    if (d.is_const(a) && (k == 1 || d.is_const(b)))
        return d.constant(apply(op, d.n[a].c, k > 1 ? d.n[b].c : 0.0f));
    auto neg = [&](uint32_t x) { return d.make(E_NEG, x); };
    switch (op) {
    case E_ADD:
        if (d.is_const(a, 0.0f)) return b;
        if (d.is_const(b, 0.0f)) return a;
        if (d.is_op(a, E_NEG) && d.is_op(b, E_NEG)) return neg(d.make(E_ADD, a, b));
        if (d.is_op(a, E_NEG)) return d.make(E_SUB, b, d.in(a, 0));
        if (d.is_op(b, E_NEG)) return d.make(E_SUB, a, d.in(b, 0));
        break;
    case E_MUL:
        if (d.is_const(a, 0.0f)) return d.constant(0.0f);
        if (d.is_const(b, 0.0f)) return d.constant(0.0f);
        if (d.is_const(a, 1.0f)) return b;
        if (d.is_const(b, 1.0f)) return a;
        if (d.is_const(a, -1.0f)) return fold_step(d, neg(b));
        if (d.is_const(b, -1.0f)) return fold_step(d, neg(a));
        if (d.is_op(a, E_NEG) || d.is_op(b, E_NEG)) {
            const uint32_t na = fold_step(d, neg(a));
            const uint32_t nb = fold_step(d, neg(b));
            return d.make(E_MUL, na, nb);
        }
        break;
    case E_SUB:
        if (d.is_const(a, 0.0f)) return fold_step(d, neg(b));
        if (d.is_const(b, 0.0f)) return a;
        if (d.is_op(b, E_NEG)) return fold_step(d, d.make(E_ADD, a, d.in(b, 0)));
        break;
    case E_DIV:
        if (d.is_const(a, 0.0f)) return d.constant(0.0f);
        if (d.is_const(b, 1.0f)) return a;
        if (d.is_const(b, -1.0f)) return fold_step(d, neg(a));
        if (d.is_const(b)) {
            const uint32_t r = d.constant(1.0f / d.n[b].c);
            return fold_step(d, d.make(E_MUL, a, r));
        }
        break;
    case E_NEG:
        if (d.is_op(a, E_NEG)) return d.in(a, 0);
        if (d.is_op(a, E_SUB)) return fold_step(d, d.make(E_SUB, d.in(a, 1), d.in(a, 0)));
        break;
    default:
        break;
    }
    return e;
}

// Bottom-up rebuild: axes and constants are kept, every other node is
// rebuilt from its mapped inputs and then rewritten by `step`.
template <class Step>
uint32_t topo_map(Dag& d, uint32_t root, Step step) {
    const std::vector<uint32_t> order = topo(d, root);
    std::unordered_map<uint32_t, uint32_t> m;
    m.reserve(order.size() * 2);
    for (uint32_t e : order) {
        const uint32_t op = d.n[e].op;
        if (!is_input_op(op)) {
            m[e] = e;
            continue;
        }
        const uint32_t a = m[d.n[e].in[0]];
        const uint32_t b = arity(op) > 1 ? m[d.n[e].in[1]] : kNone;
        m[e] = step(d.make(op, a, b));
    }
    return m[root];
}

int op_order(uint32_t op) { return (int)op; }

// ThreeWayCompare (simplifier.cpp:133-180).  A node compared with itself is
// equal, which the reference reaches by walking both sides.
int compare3(const Dag& d, uint32_t a, uint32_t b) {
    if (a == b) return 0;
    const Node& A = d.n[a];
    const Node& B = d.n[b];
    if (op_order(A.op) < op_order(B.op)) return -1;
    if (op_order(A.op) > op_order(B.op)) return 1;
    if (A.op == E_CONST) {
        if (A.c < B.c) return -1;
        if (A.c > B.c) return 1;
        return 0;
    }
    const uint32_t k = arity(A.op);
    for (uint32_t i = 0; i < k; ++i) {
        const int c = compare3(d, A.in[i], B.in[i]);
        if (c != 0) return c;
    }
    return 0;
}

// NormalFormStep (simplifier.cpp:183-206): the two operands of a
// commutative op in std::sort order with comparator compare3 <= 0.
uint32_t normal_step(Dag& d, uint32_t e) {
    const uint32_t op = d.n[e].op;
    if (op == E_ADD || op == E_MUL || op == E_MIN || op == E_MAX) {
        if (compare3(d, d.n[e].in[1], d.n[e].in[0]) <= 0) std::swap(d.n[e].in[0], d.n[e].in[1]);
    }
    return e;
}

struct DupKey {
    uint32_t op, a, b, bits;
    bool operator==(const DupKey& o) const {
        return op == o.op && a == o.a && b == o.b && bits == o.bits;
    }
};
struct DupHash {
    size_t operator()(const DupKey& k) const {
        uint64_t h = k.op * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)k.a + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.b + 0x85EBCA77C2B2AE63ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)k.bits + 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

// MergeDuplicates (simplifier.cpp:213-259).  The reference scans its list of
// unique nodes for the first AreDuplicate match; duplicates form classes of
// (op, constant value with -0 == 0, input identities), each with one member
// in the list, so a map keyed by the class finds the same node.
uint32_t merge_duplicates(Dag& d, uint32_t root) {
    root = topo_map(d, root, [&](uint32_t e) { return normal_step(d, e); });
    const std::vector<uint32_t> order = topo(d, root);
    std::unordered_map<uint32_t, uint32_t> m;
    std::unordered_map<DupKey, uint32_t, DupHash> uniq;
    m.reserve(order.size() * 2);
    uniq.reserve(order.size() * 2);
    for (uint32_t e : order) {
        const Node nd = d.n[e];
        uint32_t ne;
        DupKey key{nd.op, kNone, kNone, 0};
        if (nd.op == E_CONST) {
            ne = d.constant(nd.c);
            const float v = nd.c == 0.0f ? 0.0f : nd.c;
            std::memcpy(&key.bits, &v, 4);
            if (std::isnan(nd.c)) {   // NaN == NaN is false: never a duplicate
                key.a = ne;
                key.b = ne;
            }
        } else {
            const uint32_t k = arity(nd.op);
            const uint32_t a = k > 0 ? m[nd.in[0]] : kNone;
            const uint32_t b = k > 1 ? m[nd.in[1]] : kNone;
            ne = d.make(nd.op, a, b);
            key.a = a;
            key.b = b;
        }
        auto it = uniq.find(key);
        if (it != uniq.end()) {
            m[e] = it->second;
        } else {
            uniq.emplace(key, ne);
            m[e] = ne;
        }
    }
    return m[root];
}

// Tape op codes (vulkan/include/csg/tape.h:15-30) of the expression ops.
uint32_t tape_op(uint32_t op) {
    switch (op) {
    case E_CONST: return 0;
    case E_SIN: return 1;
    case E_COS: return 2;
    case E_ADD: return 3;
    case E_SUB: return 4;
    case E_MUL: return 5;
    case E_DIV: return 6;
    case E_MIN: return 7;
    case E_MAX: return 8;
    case E_EXP: return 9;
    case E_NEG: return 10;
    case E_SQRT: return 11;
    default: return 0xFF;
    }
}

struct Compiled {
    std::vector<uint32_t> words;
    std::vector<float> consts;
    uint32_t slots = 4;
    uint32_t simplified = 0;
};

bool build_tape(const Dag& d, uint32_t root, Compiled* out, std::string* err) {
    const std::vector<uint32_t> order = topo(d, root);
    out->simplified = (uint32_t)order.size();
    std::unordered_map<uint32_t, uint32_t> pos;
    pos.reserve(order.size() * 2);
    for (uint32_t i = 0; i < order.size(); ++i) pos[order[i]] = i;
    // constant pool in order of first appearance, keyed by value (-0 == 0);
    // every NaN occurrence takes an entry of its own
    std::unordered_map<uint32_t, uint32_t> cidx;
    std::unordered_map<uint32_t, uint32_t> nan_idx;   // node -> entry
    for (uint32_t e : order) {
        if (d.n[e].op != E_CONST) continue;
        if (std::isnan(d.n[e].c)) {
            nan_idx[e] = (uint32_t)out->consts.size();
            out->consts.push_back(d.n[e].c);
            continue;
        }
        const float v = d.n[e].c == 0.0f ? 0.0f : d.n[e].c;
        uint32_t bits;
        std::memcpy(&bits, &v, 4);
        if (cidx.find(bits) == cidx.end()) {
            cidx[bits] = (uint32_t)out->consts.size();
            out->consts.push_back(d.n[e].c);
        }
    }
    if (out->consts.size() > 65536) {
        *err = "more than 65536 distinct constants";
        return false;
    }
    // liveness: index of the last reader, -1 if none
    std::vector<int64_t> live(order.size(), -1);
    for (uint32_t i = 0; i < order.size(); ++i) {
        const Node& nd = d.n[order[i]];
        for (uint32_t j = 0; j < arity(nd.op); ++j) {
            const uint32_t p = pos[nd.in[j]];
            if ((int64_t)i > live[p]) live[p] = i;
        }
    }
    // slots 0..3 = x, y, z, t (kNone: axis unused, slot free)
    std::vector<uint32_t> slots(4, kNone);
    for (uint32_t e : order)
        if (d.n[e].op <= E_T) slots[d.n[e].op] = e;
    auto current = [&](uint32_t e) -> uint32_t {
        for (uint32_t s = 0; s < slots.size(); ++s)
            if (slots[s] == e) return s;
        return kNone;
    };
    auto release = [&](uint32_t e) {
        for (auto& s : slots)
            if (s == e) {
                s = kNone;
                return;
            }
    };
    std::vector<uint32_t> instrs;
    for (uint32_t i = 0; i < order.size(); ++i) {
        const uint32_t e = order[i];
        const Node& nd = d.n[e];
        if (nd.op <= E_T) continue;
        uint32_t a = 0, b = 0;
        if (nd.op == E_CONST) {
            const float v = nd.c == 0.0f ? 0.0f : nd.c;
            uint32_t bits;
            std::memcpy(&bits, &v, 4);
            const uint32_t ci = std::isnan(nd.c) ? nan_idx[e] : cidx[bits];
            a = ci & 0xFFu;
            b = ci >> 8;
        } else {
            const uint32_t k = arity(nd.op);
            a = current(nd.in[0]);
            b = k > 1 ? current(nd.in[1]) : 0u;
            if (live[pos[nd.in[0]]] == (int64_t)i) release(nd.in[0]);
            if (k > 1 && live[pos[nd.in[1]]] == (int64_t)i) release(nd.in[1]);
        }
        uint32_t o = kNone;
        for (uint32_t s = 0; s < slots.size(); ++s)
            if (slots[s] == kNone) {
                o = s;
                break;
            }
        if (o == kNone) {
            if (slots.size() >= 256) {
                *err = "the tape needs more than 256 slots";
                return false;
            }
            slots.push_back(kNone);
            o = (uint32_t)slots.size() - 1;
        }
        slots[o] = e;
        instrs.push_back(tape_op(nd.op) | (o << 8) | (a << 16) | (b << 24));
    }
    const uint32_t rop = d.n[root].op;
    if (rop >= E_Y && rop <= E_T) instrs.push_back(12u | (0u << 8) | (rop << 16));   // COPY
    if (!instrs.empty() && ((instrs.back() >> 8) & 0xFFu) != 0) {
        *err = "internal error: the result is not in slot 0";
        return false;
    }
    out->words.clear();
    out->words.push_back((uint32_t)instrs.size());
    out->words.insert(out->words.end(), instrs.begin(), instrs.end());
    out->slots = (uint32_t)slots.size();
    return true;
}

thread_local std::string g_compile_err;

frep_status compile_fail(const std::string& msg) {
    g_compile_err = msg;
    return FREP_INVALID_ARG;
}

}  // namespace

extern "C" {

const char* frep_compile_last_error(void) { return g_compile_err.c_str(); }

frep_status frep_compile(const frep_expr_node* nodes, uint32_t n_nodes, uint32_t root,
                         uint32_t* words, uint32_t words_cap, float* consts,
                         uint32_t consts_cap, frep_compile_info* info) {
    g_compile_err.clear();
    if (!nodes || n_nodes == 0 || !info) return compile_fail("null or empty argument");
    if (root >= n_nodes) return compile_fail("root index out of range");
    std::memset(info, 0, sizeof(*info));
    Dag d;
    d.n.reserve((size_t)n_nodes * 4);
    for (uint32_t i = 0; i < n_nodes; ++i) {
        const frep_expr_node& x = nodes[i];
        if (x.op >= E_COUNT) return compile_fail("node " + std::to_string(i) + ": unknown op");
        const uint32_t k = arity(x.op);
        if ((k > 0 && x.a >= i) || (k > 1 && x.b >= i))
            return compile_fail("node " + std::to_string(i) +
                                ": inputs must be earlier nodes of the list");
        if (x.op == E_CONST && std::isnan(x.value))
            return compile_fail("node " + std::to_string(i) + ": NaN constant");
        d.n.push_back(Node{x.op, {k > 0 ? x.a : kNone, k > 1 ? x.b : kNone},
                           x.op == E_CONST ? x.value : 0.0f});
    }
    // csg::Tape::Simplify (tape.cpp:23-29): ConstantFold, then MergeDuplicates
    uint32_t r = topo_map(d, root, [&](uint32_t e) { return fold_step(d, e); });
    r = merge_duplicates(d, r);
    Compiled c;
    std::string err;
    if (!build_tape(d, r, &c, &err)) return compile_fail(err);
    info->n_words = (uint32_t)c.words.size();
    info->n_consts = (uint32_t)c.consts.size();
    info->slot_count = c.slots;
    info->format = c.consts.size() > 256 ? FREP_TAPE_FORMAT_WIDE : FREP_TAPE_FORMAT_REF;
    info->dag_nodes = n_nodes;
    info->tape_nodes = c.simplified;
    if (words) {
        if (words_cap < c.words.size()) return compile_fail("words buffer too small");
        std::memcpy(words, c.words.data(), c.words.size() * 4);
    }
    if (consts) {
        if (consts_cap < c.consts.size()) return compile_fail("constant buffer too small");
        if (!c.consts.empty()) std::memcpy(consts, c.consts.data(), c.consts.size() * 4);
    }
    return FREP_OK;
}

}  // extern "C"
