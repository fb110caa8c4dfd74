/*
 * frep_csg.hpp -- C++ CSG expressions, shape library and demo shapes over
 * frep_compile (frep.h), with the names and semantics of the reference's
 * csg namespace (MathisBD/f-rep-voxel-rendering):
 *
 *   Expr, X..T, Constant, Sin .. Max, operators, e(x, y, z, t)
 *                   vulkan/include/csg/expression.h:14-115
 *   Vector3, Square .. RotateZ     vulkan/include/csg/lib.h:7-67
 *   Tape(Expr)      vulkan/include/csg/tape.h:44-56 (compiled by frep_compile)
 *   Screw .. OrganicBalls          vulkan/src/shapes.cpp:5-117
 *
 * so that the reference's
 *     m_voxels.Init(&m_device, m_params.shape);    (tape = csg::Tape(shape))
 * becomes
 *     storage.Init(&device, frep_csg::Tape(shape));
 * frep_csg::Tape is a frep_engine::Tape (words, constant pool, slot count)
 * and keeps the reference's instruction list and GetSlotCount().  Nodes are
 * shared_ptr as in the reference; Compile() flattens the DAG to the node
 * list frep_compile takes, keeping the sharing.  Errors throw
 * frep_engine::Error.
 *
 * Header-only; link with libfrep.so.
 */
#ifndef FREP_CSG_HPP
#define FREP_CSG_HPP

#include <cstdint>
#include <memory>
#include <unordered_map>
#include <utility>
#include <vector>

#include "frep.h"
#include "frep_engine.hpp"

namespace frep_csg {

enum class Operator : uint32_t {
    X = FREP_EXPR_X, Y, Z, T, CONST,
    SIN, COS, EXP, NEG, SQRT,
    ADD, SUB, MUL, DIV, MIN, MAX
};

class Expr;

struct Node {
    Operator op;
    std::vector<Expr> inputs;
    float constant = 0.0f;
};

class Expr {
  public:
    std::shared_ptr<Node> node;

    Expr() = default;
    explicit Expr(std::shared_ptr<Node> n) : node(std::move(n)) {}
    Expr(float c) : node(std::make_shared<Node>(Node{Operator::CONST, {}, c})) {}   // NOLINT
    Expr(double c) : Expr((float)c) {}                                              // NOLINT
    Expr(int c) : Expr((float)c) {}                                                 // NOLINT

    static Expr Make(Operator op, std::vector<Expr> in) {
        return Expr(std::make_shared<Node>(Node{op, std::move(in), 0.0f}));
    }
    bool IsOp(Operator op) const { return node->op == op; }

    // The DAG with x, y, z, t replaced (expression.cpp:281-296).
    Expr operator()(Expr nx, Expr ny, Expr nz, Expr nt) const {
        std::unordered_map<const Node*, Expr> m;
        std::vector<std::pair<Expr, size_t>> st{{*this, 0}};
        m.reserve(64);
        std::unordered_map<const Node*, bool> seen{{node.get(), true}};
        while (!st.empty()) {
            auto& top = st.back();
            const Expr e = top.first;
            if (top.second < e.node->inputs.size()) {
                const Expr c = e.node->inputs[top.second++];
                if (!seen[c.node.get()]) {
                    seen[c.node.get()] = true;
                    st.push_back({c, 0});
                }
                continue;
            }
            st.pop_back();
            switch (e.node->op) {
            case Operator::X: m[e.node.get()] = nx; break;
            case Operator::Y: m[e.node.get()] = ny; break;
            case Operator::Z: m[e.node.get()] = nz; break;
            case Operator::T: m[e.node.get()] = nt; break;
            case Operator::CONST: m[e.node.get()] = e; break;
            default: {
                std::vector<Expr> in;
                for (const Expr& i : e.node->inputs) in.push_back(m[i.node.get()]);
                m[e.node.get()] = Make(e.node->op, std::move(in));
            }
            }
        }
        return m[node.get()];
    }
};

inline Expr X() { return Expr::Make(Operator::X, {}); }
inline Expr Y() { return Expr::Make(Operator::Y, {}); }
inline Expr Z() { return Expr::Make(Operator::Z, {}); }
inline Expr T() { return Expr::Make(Operator::T, {}); }
inline Expr Constant(float c) { return Expr(c); }
inline Expr Sin(Expr e) { return Expr::Make(Operator::SIN, {e}); }
inline Expr Cos(Expr e) { return Expr::Make(Operator::COS, {e}); }
inline Expr Exp(Expr e) { return Expr::Make(Operator::EXP, {e}); }
inline Expr Sqrt(Expr e) { return Expr::Make(Operator::SQRT, {e}); }
inline Expr Min(Expr a, Expr b) { return Expr::Make(Operator::MIN, {a, b}); }
inline Expr Max(Expr a, Expr b) { return Expr::Make(Operator::MAX, {a, b}); }
inline Expr operator-(Expr e) { return Expr::Make(Operator::NEG, {e}); }
inline Expr operator+(Expr a, Expr b) { return Expr::Make(Operator::ADD, {a, b}); }
inline Expr operator-(Expr a, Expr b) { return Expr::Make(Operator::SUB, {a, b}); }
inline Expr operator*(Expr a, Expr b) { return Expr::Make(Operator::MUL, {a, b}); }
inline Expr operator/(Expr a, Expr b) { return Expr::Make(Operator::DIV, {a, b}); }

// ---- lib.cpp ----
struct Vector3 {
    Expr x, y, z;
    Vector3(Expr x_, Expr y_, Expr z_) : x(x_), y(y_), z(z_) {}
};
inline Vector3 operator+(Vector3 a, Vector3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vector3 operator-(Vector3 a, Vector3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vector3 operator*(Vector3 a, Vector3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Vector3 operator/(Vector3 a, Vector3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline Vector3 operator-(Vector3 v) { return {-v.x, -v.y, -v.z}; }
inline Vector3 operator*(Vector3 v, Expr e) { return {v.x * e, v.y * e, v.z * e}; }
inline Vector3 operator*(Expr e, Vector3 v) { return {e * v.x, e * v.y, e * v.z}; }
inline Vector3 operator/(Vector3 v, Expr e) { return {v.x / e, v.y / e, v.z / e}; }

inline Expr Empty() { return Expr(1.0f); }
inline Vector3 Axes() { return {X(), Y(), Z()}; }
inline Expr Pi() { return Expr(3.14159265358979323846f); }
inline Expr Square(Expr x) { return x * x; }
inline Expr Abs(Expr x) { return Sqrt(Square(x)); }
inline Expr Pow(Expr x, uint32_t p) {
    if (p == 0) return Constant(1.0f);
    if (p == 1) return x;
    if (p % 2 == 0) return Square(Pow(x, p / 2));
    return Square(Pow(x, p / 2)) * x;
}
inline Expr Dot(Vector3 a, Vector3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Expr Norm(Vector3 a) { return Sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline Expr Norm2(Vector3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline Expr Dist(Vector3 a, Vector3 b) { return Norm(b - a); }
inline Expr Dist2(Vector3 a, Vector3 b) { return Norm2(b - a); }
inline Expr Union(Expr a, Expr b) { return Min(a, b); }
inline Expr Intersect(Expr a, Expr b) { return Max(a, b); }
inline Expr Complement(Expr a) { return -a; }
inline Expr Diff(Expr a, Expr b) { return Intersect(a, Complement(b)); }
inline Expr Sphere(Vector3 center, Expr radius) { return Dist2(Axes(), center) - Square(radius); }
inline Expr Box(Vector3 low, Vector3 size) {
    auto a = Square(X() - (low.x + size.x / 2)) - Square(size.x / 2);
    auto b = Square(Y() - (low.y + size.y / 2)) - Square(size.y / 2);
    auto c = Square(Z() - (low.z + size.z / 2)) - Square(size.z / 2);
    return Intersect(a, Intersect(b, c));
}
inline Expr TranslateX(Expr a, Expr d) { return a(X() - d, Y(), Z(), T()); }
inline Expr TranslateY(Expr a, Expr d) { return a(X(), Y() - d, Z(), T()); }
inline Expr TranslateZ(Expr a, Expr d) { return a(X(), Y(), Z() - d, T()); }
inline Expr TranslateXYZ(Expr a, Vector3 d) { return a(X() - d.x, Y() - d.y, Z() - d.z, T()); }
inline Expr ScaleX(Expr a, Expr s) { return a(X() / s, Y(), Z(), T()); }
inline Expr ScaleY(Expr a, Expr s) { return a(X(), Y() / s, Z(), T()); }
inline Expr ScaleZ(Expr a, Expr s) { return a(X(), Y(), Z() / s, T()); }
inline Expr ScaleXYZ(Expr a, Expr s) { return a(X() / s, Y() / s, Z() / s, T()); }
inline Expr ScaleXYZ(Expr a, Vector3 s) { return a(X() / s.x, Y() / s.y, Z() / s.z, T()); }
inline Expr RotateX(Expr a, Expr angle) {
    auto c = Cos(-angle);
    auto s = Sin(-angle);
    return a(X(), c * Y() - s * Z(), s * Y() + c * Z(), T());
}
inline Expr RotateY(Expr a, Expr angle) {
    auto c = Cos(-angle);
    auto s = Sin(-angle);
    return a(c * X() + s * Z(), Y(), -s * X() + c * Z(), T());
}
inline Expr RotateZ(Expr a, Expr angle) {
    auto c = Cos(-angle);
    auto s = Sin(-angle);
    return a(c * X() - s * Y(), s * X() + c * Y(), Z(), T());
}

// ---- compiling: csg::Tape (tape.h:15-56) through frep_compile ----
struct Instr {
    uint8_t op, outSlot, inSlotA, inSlotB;
};

// The DAG as frep_compile's node list (inputs before users, sharing kept).
inline std::vector<frep_expr_node> NodeList(const Expr& root, uint32_t* root_index) {
    std::vector<frep_expr_node> out;
    std::unordered_map<const Node*, uint32_t> idx;
    std::vector<std::pair<Expr, size_t>> st{{root, 0}};
    idx.reserve(256);
    std::unordered_map<const Node*, bool> seen{{root.node.get(), true}};
    while (!st.empty()) {
        auto& top = st.back();
        const Expr e = top.first;
        if (top.second < e.node->inputs.size()) {
            const Expr c = e.node->inputs[top.second++];
            if (!seen[c.node.get()]) {
                seen[c.node.get()] = true;
                st.push_back({c, 0});
            }
            continue;
        }
        st.pop_back();
        frep_expr_node n{(uint32_t)e.node->op, 0, 0, e.node->constant};
        if (!e.node->inputs.empty()) n.a = idx[e.node->inputs[0].node.get()];
        if (e.node->inputs.size() > 1) n.b = idx[e.node->inputs[1].node.get()];
        idx[e.node.get()] = (uint32_t)out.size();
        out.push_back(n);
    }
    *root_index = idx[root.node.get()];
    return out;
}

class Tape : public frep_engine::Tape {
  public:
    std::vector<Instr> instructions;
    uint32_t format = FREP_TAPE_FORMAT_REF;

    Tape() = default;
    explicit Tape(const Expr& e) {
        uint32_t root = 0;
        const std::vector<frep_expr_node> nodes = NodeList(e, &root);
        frep_compile_info info{};
        Compile(nodes, root, &info);
        words.resize(info.n_words);
        constantPool.resize(info.n_consts);
        Compile(nodes, root, &info);
        slotCount = info.slot_count;
        format = info.format;
        for (size_t i = 1; i < words.size(); ++i) {
            const uint32_t w = words[i];
            instructions.push_back(Instr{(uint8_t)(w & 0xFF), (uint8_t)(w >> 8), (uint8_t)(w >> 16),
                                         (uint8_t)(w >> 24)});
        }
    }
    uint32_t GetSlotCount() const { return slotCount; }

  private:
    void Compile(const std::vector<frep_expr_node>& nodes, uint32_t root, frep_compile_info* info) {
        const frep_status st =
            frep_compile(nodes.data(), (uint32_t)nodes.size(), root,
                         words.empty() ? nullptr : words.data(), (uint32_t)words.size(),
                         constantPool.empty() ? nullptr : constantPool.data(),
                         (uint32_t)constantPool.size(), info);
        if (st != FREP_OK)
            throw frep_engine::Error(st, std::string("frep_compile: ") + frep_compile_last_error());
    }
};

// ---- shapes.cpp ----
inline Expr Screw() {
    auto shape = Box({-15, -2, -2}, {30, 4, 4});
    shape = RotateX(shape, X());
    auto scale = 2 * Exp(X() / 40);
    shape = ScaleXYZ(shape, Vector3{1, scale, scale});
    shape = RotateZ(shape, T() / 2);
    shape = RotateY(shape, T() / 5);
    return shape;
}

inline Expr ElasticCube() {
    auto shape = Box({-10, -10, -10}, {20, 20, 20});
    auto scale = 1 / (1 + Pow(Y() / 5, 4));
    shape = ScaleX(shape, scale);
    shape = ScaleX(shape, 1 + 0.2f * Sin(Z() + 10 * T()));
    auto radius = (1 + 0.1f * Sin(T())) * 10;
    shape = Diff(shape, Sphere({0, 0, 0}, radius));
    shape = RotateY(shape, T() / 5);
    return shape;
}

inline Expr TangleCube() {
    auto shape = Norm2(Axes() * Axes()) - 8 * Norm2(Axes()) + 25;
    return ScaleXYZ(shape, 4);
}

inline Expr Morph() {
    auto shape1 = Norm2(Axes() * Axes()) - 8 * Norm2(Axes()) + 25;
    shape1 = ScaleXYZ(shape1, 4);
    auto shape2 = Sphere({0, 0, 0}, 10);
    auto blend = (1 + Sin(T())) / 2;
    auto shape = blend * shape1 + (1 - blend) * shape2;
    return RotateY(shape, T() / 5);
}

inline Expr Menger(int level) {
    if (level <= 0) return Box({0, 0, 0}, {1, 1, 1});
    auto block = ScaleXYZ(Menger(level - 1), Vector3{1 / 3.0f, 1 / 3.0f, 1});
    auto menger = Empty();
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y)
            if (!(x == 1 && y == 1)) {
                auto tb = TranslateXYZ(block, Vector3{x / 3.0f, y / 3.0f, 0});
                menger = Union(menger, tb);
            }
    return menger;
}

inline Expr MengerSponge(int level) {
    auto shape = Menger(level);
    shape = TranslateXYZ(shape, {-0.5f, -0.5f, -0.5f});
    shape = ScaleXYZ(shape, Vector3{30, 30, 10});
    return Diff(shape, Sphere({10, 10, 0}, 10));
}

inline Expr TwistedTower(int level, float angle, float scale) {
    auto box = Box({-0.5f, 0, -0.5f}, {1, 1, 1});
    if (level <= 0) return box;
    auto shape = TwistedTower(level - 1, angle, scale);
    shape = ScaleXYZ(shape, scale);
    shape = RotateY(shape, angle);
    shape = TranslateY(shape, scale);
    return Union(shape, box);
}

inline Expr OrganicBalls() {
    auto s1 = Sphere({10, 0, 0}, 8);
    s1 = TranslateX(s1, 5 * Sin(T()));
    auto s2 = Sphere({-10, 0, 0}, 10);
    auto f = Min(Min(X(), Y()), Sqrt(X() * Y()) - 40);
    return f(s1, s2, Z(), T());
}

}  // namespace frep_csg

#endif
