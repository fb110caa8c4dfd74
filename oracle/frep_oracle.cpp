/*
 * frep_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY; see frep_oracle.h).
 *
 * Built with g++ -O2 -ffp-contract=off (no FMA contraction, no fast-math),
 * SSE2 scalar float arithmetic: every +, -, *, /, sqrt is IEEE-754 binary32
 * round-to-nearest, the same as the GPU path.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_sinf/cosf/expf   vs libm in double precision (ulp bound), identities
 *   orc_eval_point       vs direct evaluation of the expression DAG that the
 *                        tape was compiled from; closed forms (sphere SDF)
 *   orc_eval_interval    Moore interval rules per op on hand cases; dense
 *                        point sampling stays inside [lo, hi]
 *   orc_shorten          hand KATs (min of disjoint intervals -> COPY);
 *                        shortened == full tape bit-for-bit at every sampled
 *                        point of the box; kept set == backward closure
 *   orc_eval_gradient    closed forms (sphere normal, plane normal) and
 *                        finite differences
 *   orc_voxelize         flattened occupancy == dense point sampling
 *                        (orc_dense_occupancy) on power-of-two grids; node
 *                        counts vs brute force
 *   orc_render           DDA: every ray whose first occupied cell is
 *                        unambiguous (an exact float64 voxel walk over the
 *                        tree's leaf cells) hits exactly that cell, at the
 *                        level whose leaf bit covers it -- axis-aligned rays
 *                        through cell centres and perspective rays on 8^3..
 *                        64^3 grids; shading: half spaces n.p - d (constant
 *                        gradient) show the closed-form colour ambient +
 *                        sum max(dot(-L, n), 0) colour, rounded to 8 bits,
 *                        on every hit pixel, and k-of-spp averages; misses
 *                        show the background.  tests/test_oracle_mutants.py
 *                        plants light-sign, light-colour, ambient, rounding,
 *                        cell-index, step-sign and descent mistakes in
 *                        copies of this file and checks each one fails.
 *                        The reference's "t < 0 -> white" branch
 *                        (raytracer/main.comp:148-151) is kept but cannot be
 *                        reached: a hit's t is at least EPS.
 *   orc_dense_occupancy  vs orc_eval_point (same definition, different loop)
 *
 * Canonical tree blob (u32 words), shared with the GPU library's export:
 *   [0] 0x54505246 ('FRPT')  [1] 1 (version)  [2] n_levels
 *   [3 .. 3+n)        dims
 *   [3+n .. 3+2n)     node counts per level
 *   then for each level L, for each node in canonical order (sorted by its
 *   coords x, then y, then z):
 *       x, y, z, k (tape length), k instruction words,
 *       leaf mask (dim^3/32 words),
 *       and for L < n-1: interior mask (dim^3/32 words), then dim^3 child
 *       entries holding the canonical index of the child in level L+1, or
 *       0xFFFFFFFF where the cell has no interior child.
 */
#include "frep_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

/* Op codes: vulkan/include/csg/tape.h:15-30. */
enum Op : uint32_t {
    OP_CONST = 0, OP_SIN = 1, OP_COS = 2, OP_ADD = 3, OP_SUB = 4, OP_MUL = 5,
    OP_DIV = 6, OP_MIN = 7, OP_MAX = 8, OP_EXP = 9, OP_NEG = 10, OP_SQRT = 11,
    OP_COPY = 12, OP_COUNT = 13
};

/* Min/max choices: vulkan/shaders/tape/tape.comp:10-13. */
enum : uint8_t { MM_NONE = 0, MM_FIRST = 1, MM_SECOND = 2, MM_BOTH = 3 };

struct Instr { uint32_t op, out, a, b; };

/* Decode: vulkan/shaders/tape/tape.comp:45-56.  The CONST pool index is the
 * 16-bit field formed by the A and B bytes; reference tapes always carry
 * B = 0 for CONST (tape.cpp:172 zero-initialises the instruction), so this
 * is the reference's decode for every reference tape. */
inline Instr decode(uint32_t w) {
    return Instr{w & 0xFFu, (w >> 8) & 0xFFu, (w >> 16) & 0xFFu, w >> 24};
}
inline uint32_t const_index(const Instr& i) { return i.a | (i.b << 8); }
inline uint32_t encode(const Instr& i) {
    return (i.op & 0xFFu) | ((i.out & 0xFFu) << 8) | ((i.a & 0xFFu) << 16) |
           ((i.b & 0xFFu) << 24);
}

/* GLM / GLSL scalar min and max for non-NaN inputs, and their exact
 * behaviour on NaN and signed zero: min(x,y) = y < x ? y : x,
 * max(x,y) = x < y ? y : x (used by tape.cpp:291-292). */
inline float fmin_glm(float x, float y) { return y < x ? y : x; }
inline float fmax_glm(float x, float y) { return x < y ? y : x; }

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

const float kPi = 3.14159265358979323846f;   /* utils.comp:2 as a float */
const float kInf = std::numeric_limits<float>::infinity();

}  // namespace

/* ======================================================================
 * Transcendentals.  The reference calls GLM / GLSL built-ins whose results
 * are not pinned by anything in the reference, so this build defines its
 * own, identical on CPU and GPU:
 *   exp: Cody & Waite (1980) reduction x = k ln2 + r with a two-part ln2,
 *        Taylor polynomial of e^r to degree 7 on |r| <= ln2/2, scaled by
 *        2^k in two exact steps.
 *   sin/cos: Cody & Waite reduction by pi/2 with a three-part pi/2 (exact
 *        products for |k| < 4096, i.e. |x| < ~6400), Taylor polynomials of
 *        degree 9 (sin) and 10 (cos) on |r| <= pi/4, quadrant selection.
 * Only +, -, *, floorf and integer bit operations are used, all exact or
 * correctly rounded, so the GPU kernels reproduce these bit for bit.
 * ====================================================================== */

extern "C" float orc_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return kInf;
    if (x < -103.972084045410156f) return 0.0f;
    const float kf = std::floor(x * 1.44269502162933349609375f + 0.5f);
    const float r = (x - kf * 0.693145751953125f) - kf * 1.428606765330187e-06f;
    float p = 1.98412701138295233250e-04f;            /* 1/7! */
    p = 1.38888892251998186111e-03f + r * p;          /* 1/6! */
    p = 8.33333376795053482056e-03f + r * p;          /* 1/5! */
    p = 4.16666679084300994873e-02f + r * p;          /* 1/4! */
    p = 1.66666671633720397949e-01f + r * p;          /* 1/3! */
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    const int k = (int)kf;
    const int k1 = k / 2;        /* truncation toward zero */
    const int k2 = k - k1;
    const float s1 = u2f((uint32_t)(k1 + 127) << 23);
    const float s2 = u2f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

namespace {
/* Polynomials on |r| <= pi/4 (Taylor coefficients rounded to float). */
inline float sin_poly(float r) {
    const float z = r * r;
    float p = 2.75573188446287530661e-06f;            /* 1/9! */
    p = -1.98412701138295233250e-04f + z * p;         /* -1/7! */
    p = 8.33333376795053482056e-03f + z * p;          /* 1/5! */
    p = -1.66666671633720397949e-01f + z * p;         /* -1/3! */
    return r + (r * z) * p;
}
inline float cos_poly(float r) {
    const float z = r * r;
    float p = -2.75573199814971303567e-07f;           /* -1/10! */
    p = 2.48015876422869041562e-05f + z * p;          /* 1/8! */
    p = -1.38888892251998186111e-03f + z * p;         /* -1/6! */
    p = 4.16666679084300994873e-02f + z * p;          /* 1/4! */
    p = -0.5f + z * p;
    return 1.0f + z * p;
}
/* Returns the reduced argument and the quadrant (k mod 4). */
inline float reduce_half_pi(float x, int* quadrant) {
    const float kf = std::floor(x * 0.636619746685028076171875f + 0.5f);
    const float r = ((x - kf * 1.5703125f) - kf * 4.837512969970703125e-04f) -
                    kf * 7.549790126404332e-08f;
    /* k mod 4, exact in float for every finite k */
    *quadrant = (int)(kf - 4.0f * std::floor(kf * 0.25f));
    return r;
}
}  // namespace

extern "C" float orc_sinf(float x) {
    if (!(std::fabs(x) <= 3.0e38f)) return x - x;     /* NaN for NaN / inf */
    int q;
    const float r = reduce_half_pi(x, &q);
    switch (q) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}

extern "C" float orc_cosf(float x) {
    if (!(std::fabs(x) <= 3.0e38f)) return x - x;
    int q;
    const float r = reduce_half_pi(x, &q);
    switch (q) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}

/* ====================================================================== */

extern "C" int orc_tape_validate(const uint32_t* tape, uint32_t n_words,
                                 uint32_t n_consts, uint32_t slot_count) {
    if (n_words < 1) return 1;
    const uint32_t n = tape[0];
    if ((uint64_t)n + 1 > n_words) return 2;
    if (slot_count < 4 || slot_count > 256) return 3;
    for (uint32_t k = 0; k < n; ++k) {
        const Instr i = decode(tape[1 + k]);
        if (i.op >= OP_COUNT) return 4;
        if (i.out >= slot_count) return 5;
        if (i.op == OP_CONST) {
            if (const_index(i) >= n_consts) return 6;
            continue;
        }
        if (i.a >= slot_count) return 7;
        const bool binary = i.op == OP_ADD || i.op == OP_SUB || i.op == OP_MUL ||
                            i.op == OP_DIV || i.op == OP_MIN || i.op == OP_MAX;
        if (binary && i.b >= slot_count) return 8;
    }
    return 0;
}

/* Point rules: vulkan/shaders/tape/point.comp:2-63 and the CPU twin
 * vulkan/src/csg/tape.cpp:274-302 (slot 0..3 = x, y, z, t; result slot 0). */
extern "C" float orc_eval_point(const uint32_t* tape, const float* consts,
                                float x, float y, float z, float t) {
    float s[256];
    s[0] = x; s[1] = y; s[2] = z; s[3] = t;
    const uint32_t n = tape[0];
    for (uint32_t k = 0; k < n; ++k) {
        const Instr i = decode(tape[1 + k]);
        const float a = s[i.a], b = s[i.b];
        float r;
        switch (i.op) {
            case OP_CONST: r = consts[const_index(i)]; break;
            case OP_SIN:   r = orc_sinf(a); break;
            case OP_COS:   r = orc_cosf(a); break;
            case OP_ADD:   r = a + b; break;
            case OP_SUB:   r = a - b; break;
            case OP_MUL:   r = a * b; break;
            case OP_DIV:   r = a / b; break;
            case OP_MIN:   r = fmin_glm(a, b); break;
            case OP_MAX:   r = fmax_glm(a, b); break;
            case OP_EXP:   r = orc_expf(a); break;
            case OP_NEG:   r = -a; break;
            case OP_SQRT:  r = std::sqrt(a); break;
            default:       r = a; break;                 /* COPY */
        }
        s[i.out] = r;
    }
    return s[0];
}

namespace {
struct Itv { float lo, hi; };

/* interval_multiply: interval.comp:8-16, MIN_4/MAX_4 of utils.comp:16-17. */
inline Itv imul(Itv a, Itv b) {
    const float p0 = a.lo * b.lo, p1 = a.lo * b.hi, p2 = a.hi * b.lo, p3 = a.hi * b.hi;
    return Itv{fmin_glm(fmin_glm(p0, p1), fmin_glm(p2, p3)),
               fmax_glm(fmax_glm(p0, p1), fmax_glm(p2, p3))};
}

/* contains_multiple: interval.comp:20-25. */
inline bool contains_multiple(float lo, float hi, float x) {
    const float ql = lo / x, qh = hi / x;
    const float fl = std::floor(ql), fh = std::floor(qh);
    return fl < fh || fl == ql || fh == qh;
}
}  // namespace

/* Interval rules: vulkan/shaders/tape/interval.comp:36-111; choice
 * recording: interval.comp:140-155 (strict comparisons on the inputs). */
extern "C" uint32_t orc_eval_interval(const uint32_t* tape, const float* consts,
                                      const float* lo, const float* hi, float t,
                                      float* out, uint8_t* choices) {
    Itv s[256];
    s[0] = Itv{lo[0], hi[0]};
    s[1] = Itv{lo[1], hi[1]};
    s[2] = Itv{lo[2], hi[2]};
    s[3] = Itv{t, t};
    const float two_pi = 2 * kPi, half_pi = kPi / 2;
    uint32_t mm = 0;
    const uint32_t n = tape[0];
    for (uint32_t k = 0; k < n; ++k) {
        const Instr i = decode(tape[1 + k]);
        const Itv a = s[i.a], b = s[i.b];
        Itv r;
        switch (i.op) {
            case OP_CONST: { const float c = consts[const_index(i)]; r = Itv{c, c}; } break;
            case OP_COS:
                r.lo = contains_multiple(a.lo - kPi, a.hi - kPi, two_pi)
                           ? -1.0f : fmin_glm(orc_cosf(a.lo), orc_cosf(a.hi));
                r.hi = contains_multiple(a.lo, a.hi, two_pi)
                           ? 1.0f : fmax_glm(orc_cosf(a.lo), orc_cosf(a.hi));
                break;
            case OP_SIN:
                r.lo = contains_multiple(a.lo + half_pi, a.hi + half_pi, two_pi)
                           ? -1.0f : fmin_glm(orc_sinf(a.lo), orc_sinf(a.hi));
                r.hi = contains_multiple(a.lo - half_pi, a.hi - half_pi, two_pi)
                           ? 1.0f : fmax_glm(orc_sinf(a.lo), orc_sinf(a.hi));
                break;
            case OP_ADD: r = Itv{a.lo + b.lo, a.hi + b.hi}; break;
            case OP_SUB: r = Itv{a.lo - b.hi, a.hi - b.lo}; break;
            case OP_MUL: r = imul(a, b); break;
            case OP_DIV:
                if ((b.lo == 0 && b.hi == 0) || (b.lo < 0 && b.hi > 0)) r = Itv{-kInf, kInf};
                else if (b.lo == 0) r = imul(a, Itv{1.0f / b.hi, kInf});
                else if (b.hi == 0) r = imul(a, Itv{-kInf, 1.0f / b.lo});
                else r = imul(a, Itv{1.0f / b.hi, 1.0f / b.lo});
                break;
            case OP_MIN: r = Itv{fmin_glm(a.lo, b.lo), fmin_glm(a.hi, b.hi)}; break;
            case OP_MAX: r = Itv{fmax_glm(a.lo, b.lo), fmax_glm(a.hi, b.hi)}; break;
            case OP_EXP: r = Itv{orc_expf(a.lo), orc_expf(a.hi)}; break;
            case OP_NEG: r = Itv{-a.hi, -a.lo}; break;
            case OP_SQRT: r = Itv{std::sqrt(fmax_glm(0.0f, a.lo)), std::sqrt(fmax_glm(0.0f, a.hi))}; break;
            default: r = a; break;                        /* COPY */
        }
        if (i.op == OP_MIN) {
            const uint8_t c = a.hi < b.lo ? MM_FIRST : (b.hi < a.lo ? MM_SECOND : MM_BOTH);
            if (choices) choices[mm] = c;
            ++mm;
        } else if (i.op == OP_MAX) {
            const uint8_t c = a.hi < b.lo ? MM_SECOND : (b.hi < a.lo ? MM_FIRST : MM_BOTH);
            if (choices) choices[mm] = c;
            ++mm;
        }
        s[i.out] = r;
    }
    out[0] = s[0].lo;
    out[1] = s[0].hi;
    return mm;
}

/* Tape shortening: backward liveness pass (shorten.comp:164-186, get_choice
 * at shorten.comp:27-51, output slot cleared before the inputs are set),
 * "not shorter -> keep the parent tape" (shorten.comp:188-191), forward copy
 * with MIN/MAX rewritten to COPY (shorten.comp:61-88).  Slots are not
 * renumbered. */
extern "C" uint32_t orc_shorten(const uint32_t* tape, const uint8_t* choices,
                                uint32_t* out, int* reused) {
    const uint32_t n = tape[0];
    uint32_t n_mm = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t op = tape[1 + k] & 0xFFu;
        n_mm += (op == OP_MIN || op == OP_MAX);
    }
    std::vector<uint8_t> keep(n, 0);
    bool active[256] = {false};
    active[0] = true;
    uint32_t kept = 0;
    int64_t mm = (int64_t)n_mm - 1;
    for (int64_t k = (int64_t)n - 1; k >= 0; --k) {
        const Instr i = decode(tape[1 + k]);
        const bool is_mm = (i.op == OP_MIN || i.op == OP_MAX);
        if (active[i.out]) {
            keep[k] = 1;
            ++kept;
            uint8_t c;
            switch (i.op) {
                case OP_CONST: c = MM_NONE; break;
                case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: c = MM_BOTH; break;
                case OP_MIN: case OP_MAX: c = choices[mm]; break;
                default: c = MM_FIRST; break;             /* unary, COPY */
            }
            active[i.out] = false;
            if (c & MM_FIRST) active[i.a] = true;
            if (c & MM_SECOND) active[i.b] = true;
        }
        if (is_mm) --mm;
    }
    if (kept >= n) {
        std::memcpy(out, tape, sizeof(uint32_t) * (n + 1));
        *reused = 1;
        return n;
    }
    *reused = 0;
    out[0] = kept;
    uint32_t pos = 1;
    mm = 0;
    for (uint32_t k = 0; k < n; ++k) {
        Instr i = decode(tape[1 + k]);
        const bool is_mm = (i.op == OP_MIN || i.op == OP_MAX);
        if (keep[k]) {
            if (is_mm) {
                const uint8_t c = choices[mm];
                if (c == MM_FIRST) { i.op = OP_COPY; }
                else if (c == MM_SECOND) { i.op = OP_COPY; i.a = i.b; }
            }
            out[pos++] = encode(i);
        }
        if (is_mm) ++mm;
    }
    return kept;
}

/* Gradient rules: vulkan/shaders/tape/gradient.comp:4-88.  vec4 = (dx, dy,
 * dz, value); componentwise operations evaluated in GLSL's order. */
extern "C" void orc_eval_gradient(const uint32_t* tape, const float* consts,
                                  float x, float y, float z, float t, float* out) {
    struct V4 { float x, y, z, w; };
    V4 s[256];
    s[0] = V4{1, 0, 0, x};
    s[1] = V4{0, 1, 0, y};
    s[2] = V4{0, 0, 1, z};
    s[3] = V4{0, 0, 0, t};
    const uint32_t n = tape[0];
    for (uint32_t k = 0; k < n; ++k) {
        const Instr i = decode(tape[1 + k]);
        const V4 a = s[i.a], b = s[i.b];
        V4 r;
        switch (i.op) {
            case OP_CONST: r = V4{0, 0, 0, consts[const_index(i)]}; break;
            case OP_SIN: { const float c = orc_cosf(a.w);
                r = V4{c * a.x, c * a.y, c * a.z, orc_sinf(a.w)}; } break;
            case OP_COS: { const float ms = -orc_sinf(a.w);
                r = V4{ms * a.x, ms * a.y, ms * a.z, orc_cosf(a.w)}; } break;
            case OP_ADD: r = V4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; break;
            case OP_SUB: r = V4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; break;
            case OP_MUL:
                r = V4{a.x * b.w + a.w * b.x, a.y * b.w + a.w * b.y,
                       a.z * b.w + a.w * b.z, a.w * b.w};
                break;
            case OP_DIV: { const float d = b.w * b.w;
                r = V4{(a.x * b.w - a.w * b.x) / d, (a.y * b.w - a.w * b.y) / d,
                       (a.z * b.w - a.w * b.z) / d, a.w / b.w}; } break;
            case OP_MIN: r = (a.w < b.w) ? a : b; break;
            case OP_MAX: r = (a.w > b.w) ? a : b; break;
            case OP_EXP: { const float e = orc_expf(a.w);
                r = V4{e * a.x, e * a.y, e * a.z, e * 1.0f}; } break;
            case OP_NEG: r = V4{-a.x, -a.y, -a.z, -a.w}; break;
            case OP_SQRT: { const float sq = std::sqrt(a.w), d = 2 * sq;
                r = V4{a.x / d, a.y / d, a.z / d, sq}; } break;
            default: r = a; break;                        /* COPY */
        }
        s[i.out] = r;
    }
    out[0] = s[0].x; out[1] = s[0].y; out[2] = s[0].z; out[3] = s[0].w;
}

/* ======================================================================
 * Tree build: restatement of the GPU voxelizer.  Per level L (main.comp:
 * 154-170): every cell of every node; cell position = grid low +
 * float(node.coords * dim + cell) * cell_size[L] (main.comp:113-116);
 * cell_size[L] = world / d0 / ... / dL by sequential float division
 * (voxelizer.cpp:210-216).  Non-last level: interval classification
 * (main.comp:88-108) + shortening of ambiguous cells; last level: point
 * evaluation at the cell's low corner, leaf iff density < 0 (main.comp:74-83).
 * Mask bit index INDEX3D = z + dim*y + dim*dim*x (utils.comp:6).
 * Node ids here are assigned in (parent id, bit index) order; the GPU assigns
 * them in a different order, so the two are compared in canonical form.
 * ====================================================================== */

struct orc_node {
    uint32_t coords[3];
    std::vector<uint32_t> tape;        /* [k, instrs] */
    std::vector<uint32_t> leaf;        /* dim^3/32 words */
    std::vector<uint32_t> interior;    /* dim^3/32 words (non-last levels) */
    std::vector<uint32_t> child;       /* dim^3 entries, UINT32_MAX = none */
};

struct orc_tree {
    orc_grid grid;
    std::vector<float> consts;
    std::vector<float> cell_size;
    std::vector<std::vector<orc_node>> levels;
};

namespace {
std::vector<float> cell_sizes(const orc_grid& g) {
    std::vector<float> cs(g.n_levels);
    float c = g.world_size;
    for (uint32_t L = 0; L < g.n_levels; ++L) {
        c /= (float)g.dims[L];
        cs[L] = c;
    }
    return cs;
}
inline uint32_t tape_n_mm(const uint32_t* tape) {
    uint32_t n = 0;
    for (uint32_t k = 0; k < tape[0]; ++k) {
        const uint32_t op = tape[1 + k] & 0xFFu;
        n += (op == OP_MIN || op == OP_MAX);
    }
    return n;
}
}  // namespace

namespace {
/* One node of level L: its leaf mask, and for inner levels its interior mask
 * and, in cell order, the children of its ambiguous cells (coords + the
 * shortened tape).  child[] is left unset: the caller numbers children. */
void eval_node(orc_node& nd, uint32_t L, const orc_grid* grid, const float* consts,
               const std::vector<float>& cell_size, std::vector<orc_node>* kids,
               std::vector<uint32_t>* kid_cells) {
    const uint32_t dim = grid->dims[L];
    const uint32_t cells = dim * dim * dim;
    const float cs = cell_size[L];
    const bool last = (L + 1 == grid->n_levels);
    nd.leaf.assign(cells / 32, 0);
    if (!last) {
        nd.interior.assign(cells / 32, 0);
        nd.child.assign(cells, 0xFFFFFFFFu);
    }
    const uint32_t n_mm = tape_n_mm(nd.tape.data());
    std::vector<uint8_t> ch(n_mm + 1);
    std::vector<uint32_t> st(nd.tape.size());
    for (uint32_t idx = 0; idx < cells; ++idx) {
        const uint32_t cx = idx / (dim * dim), cy = (idx / dim) % dim, cz = idx % dim;
        const uint32_t gc[3] = {nd.coords[0] * dim + cx, nd.coords[1] * dim + cy,
                                nd.coords[2] * dim + cz};
        const float pos[3] = {grid->low[0] + (float)gc[0] * cs,
                              grid->low[1] + (float)gc[1] * cs,
                              grid->low[2] + (float)gc[2] * cs};
        const uint32_t q = idx >> 5, bit = 1u << (idx & 31u);
        if (last) {
            const float d = orc_eval_point(nd.tape.data(), consts, pos[0], pos[1],
                                           pos[2], grid->tape_time);
            if (d < 0.0f) nd.leaf[q] |= bit;
            continue;
        }
        const float hi[3] = {pos[0] + cs, pos[1] + cs, pos[2] + cs};
        float itv[2];
        orc_eval_interval(nd.tape.data(), consts, pos, hi, grid->tape_time, itv, ch.data());
        if (itv[0] > 0) continue;                   /* empty */
        if (itv[1] < 0) { nd.leaf[q] |= bit; continue; }   /* full */
        nd.interior[q] |= bit;
        orc_node c;
        c.coords[0] = gc[0]; c.coords[1] = gc[1]; c.coords[2] = gc[2];
        int reused = 0;
        const uint32_t k = orc_shorten(nd.tape.data(), ch.data(), st.data(), &reused);
        c.tape.assign(st.begin(), st.begin() + k + 1);
        kids->push_back(std::move(c));
        kid_cells->push_back(idx);
    }
}
}  // namespace

extern "C" orc_tree* orc_voxelize(const uint32_t* tape, const float* consts,
                                  const orc_grid* grid) {
    orc_tree* tr = new orc_tree();
    tr->grid = *grid;
    uint32_t max_const = 0;
    for (uint32_t k = 0; k < tape[0]; ++k) {
        const Instr i = decode(tape[1 + k]);
        if (i.op == OP_CONST) max_const = std::max(max_const, const_index(i) + 1);
    }
    tr->consts.assign(consts, consts + max_const);
    tr->cell_size = cell_sizes(*grid);
    const uint32_t nl = grid->n_levels;
    tr->levels.resize(nl);
    orc_node root;
    root.coords[0] = root.coords[1] = root.coords[2] = 0;
    root.tape.assign(tape, tape + tape[0] + 1);
    tr->levels[0].push_back(std::move(root));
    for (uint32_t L = 0; L < nl; ++L) {
        std::vector<orc_node>& nodes = tr->levels[L];
        std::vector<orc_node> next;
        std::vector<uint32_t> cells_of;
        for (size_t j = 0; j < nodes.size(); ++j) {
            const size_t first = next.size();
            cells_of.clear();
            eval_node(nodes[j], L, grid, consts, tr->cell_size, &next, &cells_of);
            for (size_t c = 0; c < cells_of.size(); ++c)
                nodes[j].child[cells_of[c]] = (uint32_t)(first + c);
        }
        if (L + 1 < nl) tr->levels[L + 1] = std::move(next);
    }
    return tr;
}

extern "C" void orc_tree_free(orc_tree* tree) { delete tree; }

extern "C" uint32_t orc_tree_node_count(const orc_tree* tree, uint32_t level) {
    if (level >= tree->levels.size()) return 0;
    return (uint32_t)tree->levels[level].size();
}

extern "C" uint64_t orc_tree_canonical(const orc_tree* tree, uint32_t* buf, uint64_t cap) {
    const uint32_t nl = tree->grid.n_levels;
    /* canonical order per level: sort by coords */
    std::vector<std::vector<uint32_t>> order(nl), rank(nl);
    for (uint32_t L = 0; L < nl; ++L) {
        const auto& nodes = tree->levels[L];
        order[L].resize(nodes.size());
        for (uint32_t j = 0; j < nodes.size(); ++j) order[L][j] = j;
        std::sort(order[L].begin(), order[L].end(), [&](uint32_t a, uint32_t b) {
            const uint32_t* ca = nodes[a].coords;
            const uint32_t* cb = nodes[b].coords;
            if (ca[0] != cb[0]) return ca[0] < cb[0];
            if (ca[1] != cb[1]) return ca[1] < cb[1];
            return ca[2] < cb[2];
        });
        rank[L].resize(nodes.size());
        for (uint32_t r = 0; r < order[L].size(); ++r) rank[L][order[L][r]] = r;
    }
    std::vector<uint32_t> v;
    v.push_back(0x54505246u);
    v.push_back(1u);
    v.push_back(nl);
    for (uint32_t L = 0; L < nl; ++L) v.push_back(tree->grid.dims[L]);
    for (uint32_t L = 0; L < nl; ++L) v.push_back((uint32_t)tree->levels[L].size());
    for (uint32_t L = 0; L < nl; ++L) {
        const bool last = (L + 1 == nl);
        for (uint32_t r = 0; r < order[L].size(); ++r) {
            const orc_node& nd = tree->levels[L][order[L][r]];
            v.insert(v.end(), nd.coords, nd.coords + 3);
            v.insert(v.end(), nd.tape.begin(), nd.tape.end());
            v.insert(v.end(), nd.leaf.begin(), nd.leaf.end());
            if (!last) {
                v.insert(v.end(), nd.interior.begin(), nd.interior.end());
                for (uint32_t c : nd.child)
                    v.push_back(c == 0xFFFFFFFFu ? c : rank[L + 1][c]);
            }
        }
    }
    if (buf && cap >= v.size()) std::memcpy(buf, v.data(), v.size() * 4);
    return v.size();
}

/* ---- trees too large to build whole on the CPU -----------------------
 *
 * A 4096^3 tree holds millions of last-level nodes; the oracle evaluates its
 * first levels in full and a sample of the rest.  Blob records are located by
 * one walk over the blob (records are variable-length: k tape words). */

namespace {
struct BlobIndex {
    uint32_t nl = 0;
    std::vector<uint32_t> dims, counts;
    std::vector<std::vector<uint64_t>> rec;    /* word offset of each record */
};

/* Returns false if the blob is not a well-formed canonical tree blob. */
bool index_blob(const uint32_t* b, uint64_t n, BlobIndex* ix) {
    if (n < 3 || b[0] != 0x54505246u || b[1] != 1u) return false;
    const uint32_t nl = b[2];
    if (nl < 1 || nl > 8 || n < 3 + 2ull * nl) return false;
    ix->nl = nl;
    ix->dims.assign(b + 3, b + 3 + nl);
    ix->counts.assign(b + 3 + nl, b + 3 + 2 * nl);
    ix->rec.assign(nl, {});
    uint64_t pos = 3 + 2ull * nl;
    for (uint32_t L = 0; L < nl; ++L) {
        const uint64_t cells = (uint64_t)ix->dims[L] * ix->dims[L] * ix->dims[L];
        const bool last = (L + 1 == nl);
        ix->rec[L].resize(ix->counts[L]);
        for (uint32_t j = 0; j < ix->counts[L]; ++j) {
            if (pos + 4 > n) return false;
            ix->rec[L][j] = pos;
            pos += 4 + (uint64_t)b[pos + 3];
            pos += cells / 32 + (last ? 0 : cells / 32 + cells);
            if (pos > n) return false;
        }
    }
    return pos == n;
}

bool coords_less(const orc_node& a, const orc_node& b) {
    if (a.coords[0] != b.coords[0]) return a.coords[0] < b.coords[0];
    if (a.coords[1] != b.coords[1]) return a.coords[1] < b.coords[1];
    return a.coords[2] < b.coords[2];
}

template <class F>
void parallel_for(size_t n, uint32_t threads, F fn) {
    const size_t T = std::max<size_t>(1, std::min<size_t>(threads ? threads : 1, n));
    if (T == 1) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t t = 0; t < T; ++t)
        pool.emplace_back([&, t] {
            for (size_t i = t; i < n; i += T) fn(i);
        });
    for (auto& th : pool) th.join();
}
}  // namespace

extern "C" int64_t orc_check_blob(const uint32_t* tape, const float* consts,
                                  const orc_grid* grid, const uint32_t* blob,
                                  uint64_t n_words, uint32_t n_full, uint32_t sample_every,
                                  uint32_t dense_every, uint32_t threads, uint64_t* report) {
    for (int i = 0; i < 24; ++i) report[i] = 0;
    report[4] = report[5] = ~0ull;
    uint64_t bad = 0;
    auto fail = [&](uint32_t L, uint64_t r, uint64_t kind) {
        if (bad == 0) { report[4] = L; report[5] = r; report[6] = kind; }
        ++bad;
    };
    BlobIndex ix;
    if (!index_blob(blob, n_words, &ix)) { report[6] = 1; return -1; }
    const uint32_t nl = grid->n_levels;
    if (ix.nl != nl) { report[6] = 1; return -1; }
    for (uint32_t L = 0; L < nl; ++L)
        if (ix.dims[L] != grid->dims[L]) { report[6] = 1; return -1; }
    const std::vector<float> cs = cell_sizes(*grid);
    if (sample_every == 0) sample_every = 1;
    if (dense_every == 0) dense_every = 1;
    auto coords_at = [&](uint32_t L, uint64_t r) { return blob + ix.rec[L][r]; };
    /* record of level L with the given coords (records are in coords order) */
    auto find = [&](uint32_t L, const uint32_t* c) -> uint64_t {
        uint64_t lo = 0, hi = ix.rec[L].size();
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            const uint32_t* m = coords_at(L, mid);
            const bool less = m[0] != c[0] ? m[0] < c[0] : m[1] != c[1] ? m[1] < c[1] : m[2] < c[2];
            if (less) lo = mid + 1; else hi = mid;
        }
        if (lo < ix.rec[L].size()) {
            const uint32_t* m = coords_at(L, lo);
            if (m[0] == c[0] && m[1] == c[1] && m[2] == c[2]) return lo;
        }
        return ~0ull;
    };

    /* `level` holds oracle nodes of level L (coords + tape): every node of
     * the level when `complete` (then in canonical order), else the
     * children of the nodes evaluated at level L-1 */
    std::vector<orc_node> level(1);
    level[0].coords[0] = level[0].coords[1] = level[0].coords[2] = 0;
    level[0].tape.assign(tape, tape + tape[0] + 1);
    bool complete = true;
    uint64_t dense_seen = 0;
    for (uint32_t L = 0; L < nl && !level.empty(); ++L) {
        const bool last = (L + 1 == nl);
        const bool full = (L < n_full);
        const uint32_t dim = grid->dims[L];
        const uint64_t cells = (uint64_t)dim * dim * dim, W = cells / 32;
        std::vector<uint64_t> rec(level.size());
        if (complete) {
            report[8 + L] = level.size();
            if (ix.counts[L] != level.size()) { fail(L, ~0ull, 2); break; }
            for (size_t j = 0; j < level.size(); ++j) rec[j] = j;
        } else {
            for (size_t j = 0; j < level.size(); ++j) rec[j] = find(L, level[j].coords);
        }
        /* coords and tape of every node */
        std::vector<char> ok(level.size(), 1);
        for (size_t j = 0; j < level.size(); ++j) {
            const orc_node& nd = level[j];
            if (rec[j] == ~0ull) { fail(L, j, 3); ok[j] = 0; continue; }
            const uint32_t* p = coords_at(L, rec[j]);
            if (p[0] != nd.coords[0] || p[1] != nd.coords[1] || p[2] != nd.coords[2]) {
                fail(L, rec[j], 3);
                ok[j] = 0;
                continue;
            }
            const uint32_t k = p[3];
            if (k + 1 != nd.tape.size() || std::memcmp(p + 3, nd.tape.data(), 4ull * (k + 1)) != 0) {
                fail(L, rec[j], 4);
                ok[j] = 0;
            }
        }
        /* which nodes are evaluated */
        std::vector<size_t> todo;
        for (size_t j = 0; j < level.size(); ++j)
            if (ok[j] && (full || j % sample_every == 0)) todo.push_back(j);
        std::vector<std::vector<orc_node>> kids(todo.size());
        std::vector<std::vector<uint32_t>> kid_cells(todo.size());
        std::vector<uint64_t> dense_bad(todo.size(), 0);
        std::vector<char> dense(todo.size(), 0);
        if (last && !full)
            for (size_t t = 0; t < todo.size(); ++t) dense[t] = (dense_seen++ % dense_every) == 0;
        parallel_for(todo.size(), threads, [&](size_t t) {
            orc_node& nd = level[todo[t]];
            eval_node(nd, L, grid, consts, cs, &kids[t], &kid_cells[t]);
            if (!dense[t]) return;
            /* the leaf bits against point evaluation of the full tape */
            for (uint32_t idx = 0; idx < cells; ++idx) {
                const uint32_t cx = idx / (dim * dim), cy = (idx / dim) % dim, cz = idx % dim;
                const float px = grid->low[0] + (float)(nd.coords[0] * dim + cx) * cs[L];
                const float py = grid->low[1] + (float)(nd.coords[1] * dim + cy) * cs[L];
                const float pz = grid->low[2] + (float)(nd.coords[2] * dim + cz) * cs[L];
                const bool in = orc_eval_point(tape, consts, px, py, pz, grid->tape_time) < 0.0f;
                dense_bad[t] += in != (((nd.leaf[idx >> 5] >> (idx & 31u)) & 1u) != 0);
            }
        });
        for (size_t t = 0; t < todo.size(); ++t) {
            report[3] += dense_bad[t];
            if (dense[t]) report[2] += cells;
        }
        report[full ? 0 : 1] += todo.size();
        /* the next level: every child (in canonical order) below a full
         * level, the children of the evaluated nodes below a sampled one */
        std::vector<orc_node> next;
        if (!last) {
            std::vector<std::pair<size_t, uint32_t>> src;   /* (todo index, cell) */
            for (size_t t = 0; t < todo.size(); ++t)
                for (size_t c = 0; c < kids[t].size(); ++c) {
                    next.push_back(std::move(kids[t][c]));
                    src.emplace_back(t, kid_cells[t][c]);
                }
            std::vector<uint32_t> order(next.size());
            for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
            std::sort(order.begin(), order.end(),
                      [&](uint32_t a, uint32_t b) { return coords_less(next[a], next[b]); });
            std::vector<orc_node> sorted(next.size());
            for (uint32_t r = 0; r < order.size(); ++r) {
                sorted[r] = std::move(next[order[r]]);
                const auto& sc = src[order[r]];
                /* a full level knows every child's rank; below a sampled
                 * level the rank is read from the blob and its record's
                 * coords are checked at the next level */
                if (complete && full) level[todo[sc.first]].child[sc.second] = r;
            }
            next = std::move(sorted);
        }
        /* masks and child entries of the evaluated nodes */
        for (size_t t = 0; t < todo.size(); ++t) {
            const size_t j = todo[t];
            const orc_node& nd = level[j];
            const uint32_t* p = blob + ix.rec[L][rec[j]] + 4 + blob[ix.rec[L][rec[j]] + 3];
            if (std::memcmp(p, nd.leaf.data(), 4 * W) != 0) { fail(L, rec[j], 5); continue; }
            if (last) continue;
            if (std::memcmp(p + W, nd.interior.data(), 4 * W) != 0) { fail(L, rec[j], 6); continue; }
            const uint32_t* ch = p + 2 * W;
            if (complete && full) {
                if (std::memcmp(ch, nd.child.data(), 4 * cells) != 0) fail(L, rec[j], 7);
                continue;
            }
            /* entries: none where the interior bit is clear; where it is set,
             * a record of level L+1 with the child's coords */
            bool good = true;
            for (uint32_t idx = 0; idx < cells && good; ++idx) {
                const bool in = (nd.interior[idx >> 5] >> (idx & 31u)) & 1u;
                if (!in) { good = ch[idx] == 0xFFFFFFFFu; continue; }
                const uint32_t cx = idx / (dim * dim), cy = (idx / dim) % dim, cz = idx % dim;
                const uint32_t want[3] = {nd.coords[0] * dim + cx, nd.coords[1] * dim + cy,
                                          nd.coords[2] * dim + cz};
                good = ch[idx] < ix.counts[L + 1] &&
                       std::memcmp(coords_at(L + 1, ch[idx]), want, 12) == 0;
            }
            if (!good) fail(L, rec[j], 7);
        }
        if (!full) complete = false;
        level = std::move(next);
    }
    report[7] = bad;
    return (int64_t)(bad + report[3]);
}

extern "C" orc_tree* orc_tree_from_blob(const uint32_t* blob, uint64_t n_words,
                                        const float* consts, uint32_t n_consts,
                                        const orc_grid* grid) {
    BlobIndex ix;
    if (!index_blob(blob, n_words, &ix) || ix.nl != grid->n_levels) return nullptr;
    orc_tree* tr = new orc_tree();
    tr->grid = *grid;
    tr->consts.assign(consts, consts + n_consts);
    tr->cell_size = cell_sizes(*grid);
    tr->levels.resize(ix.nl);
    for (uint32_t L = 0; L < ix.nl; ++L) {
        const uint64_t cells = (uint64_t)ix.dims[L] * ix.dims[L] * ix.dims[L], W = cells / 32;
        const bool last = (L + 1 == ix.nl);
        auto& nodes = tr->levels[L];
        nodes.resize(ix.counts[L]);
        for (uint32_t j = 0; j < ix.counts[L]; ++j) {
            const uint32_t* p = blob + ix.rec[L][j];
            orc_node& nd = nodes[j];
            nd.coords[0] = p[0]; nd.coords[1] = p[1]; nd.coords[2] = p[2];
            nd.tape.assign(p + 3, p + 4 + p[3]);
            p += 4 + p[3];
            nd.leaf.assign(p, p + W);
            if (last) continue;
            nd.interior.assign(p + W, p + 2 * W);
            nd.child.assign(p + 2 * W, p + 2 * W + cells);
            for (uint32_t c : nd.child)
                if (c != 0xFFFFFFFFu && c >= ix.counts[L + 1]) {
                    delete tr;
                    return nullptr;
                }
        }
    }
    return tr;
}

namespace {
void flatten_rec(const orc_tree* tr, uint32_t L, uint32_t node, uint8_t* bits, uint64_t F) {
    const orc_node& nd = tr->levels[L][node];
    const uint32_t dim = tr->grid.dims[L];
    const bool last = (L + 1 == tr->grid.n_levels);
    uint64_t span = 1;   /* fine voxels per cell edge at this level */
    for (uint32_t k = L + 1; k < tr->grid.n_levels; ++k) span *= tr->grid.dims[k];
    for (uint32_t idx = 0; idx < dim * dim * dim; ++idx) {
        const uint32_t q = idx >> 5, bit = 1u << (idx & 31u);
        const uint64_t cx = (uint64_t)nd.coords[0] * dim + idx / (dim * dim);
        const uint64_t cy = (uint64_t)nd.coords[1] * dim + (idx / dim) % dim;
        const uint64_t cz = (uint64_t)nd.coords[2] * dim + idx % dim;
        if (nd.leaf[q] & bit) {
            for (uint64_t x = cx * span; x < (cx + 1) * span; ++x)
                for (uint64_t y = cy * span; y < (cy + 1) * span; ++y)
                    std::memset(bits + (x * F + y) * F + cz * span, 1, span);
        } else if (!last && (nd.interior[q] & bit)) {
            flatten_rec(tr, L + 1, nd.child[idx], bits, F);
        }
    }
}
/* Flop weights per executed tape op (point, interval). */
inline void op_weights(uint32_t op, double* pw, double* iw) {
    switch (op) {
        case OP_CONST: case OP_COPY: *pw = 0; *iw = 0; break;
        case OP_ADD: case OP_SUB: case OP_NEG: case OP_EXP: *pw = 1; *iw = 2; break;
        case OP_MUL: *pw = 1; *iw = 10; break;
        case OP_DIV: *pw = 1; *iw = 12; break;
        case OP_MIN: case OP_MAX: case OP_SQRT: *pw = 1; *iw = 4; break;
        default: *pw = 1; *iw = 12; break;                 /* SIN, COS */
    }
}
}  // namespace

extern "C" void orc_tree_flatten(const orc_tree* tree, uint8_t* bits) {
    uint64_t F = 1;
    for (uint32_t L = 0; L < tree->grid.n_levels; ++L) F *= tree->grid.dims[L];
    std::memset(bits, 0, F * F * F);
    flatten_rec(tree, 0, 0, bits, F);
}

extern "C" void orc_tree_stats(const orc_tree* tree, double* out) {
    for (uint32_t L = 0; L < tree->grid.n_levels; ++L) {
        double nodes = 0, len = 0, pf = 0, itf = 0;
        for (const orc_node& nd : tree->levels[L]) {
            nodes += 1;
            len += nd.tape[0];
            for (uint32_t k = 0; k < nd.tape[0]; ++k) {
                double pw, iw;
                op_weights(nd.tape[1 + k] & 0xFFu, &pw, &iw);
                pf += pw;
                itf += iw;
            }
        }
        out[L * 4 + 0] = nodes;
        out[L * 4 + 1] = len;
        out[L * 4 + 2] = pf;
        out[L * 4 + 3] = itf;
    }
}

/* ======================================================================
 * Plain dense CPU evaluator: the reference's SceneBuilder::BuildNode
 * (scene_builder.cpp:91-156) samples csg::Tape::Eval (tape.cpp:274-302) at
 * every fine voxel position VoxelStorage::WorldPosition = low +
 * vec3(c) * world / fineDim (voxel_storage.h:58-61), leaf iff < 0.  The
 * "faithful" mode allocates the slot vector per call as tape.cpp:276 does.
 * ====================================================================== */
namespace {
float eval_faithful(const uint32_t* tape, const float* consts, uint32_t slot_count,
                    float x, float y, float z, float t) {
    std::vector<float> s(slot_count);
    s[0] = x; s[1] = y; s[2] = z; s[3] = t;
    const uint32_t n = tape[0];
    for (uint32_t k = 0; k < n; ++k) {
        const Instr i = decode(tape[1 + k]);
        switch (i.op) {
            case OP_CONST: s[i.out] = consts[const_index(i)]; break;
            case OP_SIN:   s[i.out] = orc_sinf(s[i.a]); break;
            case OP_COS:   s[i.out] = orc_cosf(s[i.a]); break;
            case OP_ADD:   s[i.out] = s[i.a] + s[i.b]; break;
            case OP_SUB:   s[i.out] = s[i.a] - s[i.b]; break;
            case OP_MUL:   s[i.out] = s[i.a] * s[i.b]; break;
            case OP_DIV:   s[i.out] = s[i.a] / s[i.b]; break;
            case OP_MIN:   s[i.out] = fmin_glm(s[i.a], s[i.b]); break;
            case OP_MAX:   s[i.out] = fmax_glm(s[i.a], s[i.b]); break;
            case OP_EXP:   s[i.out] = orc_expf(s[i.a]); break;
            case OP_NEG:   s[i.out] = -s[i.a]; break;
            case OP_SQRT:  s[i.out] = std::sqrt(s[i.a]); break;
            default:       s[i.out] = s[i.a]; break;
        }
    }
    return s[0];
}
}  // namespace

extern "C" void orc_dense_occupancy(const uint32_t* tape, const float* consts,
                                    uint32_t slot_count, const orc_grid* grid,
                                    uint32_t x_begin, uint32_t x_end, int faithful,
                                    int threads, uint8_t* out) {
    uint64_t F = 1;
    for (uint32_t L = 0; L < grid->n_levels; ++L) F *= grid->dims[L];
    const float fine = (float)F;
    auto work = [&](uint32_t xa, uint32_t xb) {
        for (uint32_t x = xa; x < xb; ++x)
            for (uint32_t y = 0; y < F; ++y)
                for (uint32_t z = 0; z < F; ++z) {
                    const float px = grid->low[0] + ((float)x * grid->world_size) / fine;
                    const float py = grid->low[1] + ((float)y * grid->world_size) / fine;
                    const float pz = grid->low[2] + ((float)z * grid->world_size) / fine;
                    const float d = faithful
                        ? eval_faithful(tape, consts, slot_count, px, py, pz, 0.0f)
                        : orc_eval_point(tape, consts, px, py, pz, 0.0f);
                    out[((uint64_t)(x - x_begin) * F + y) * F + z] = d < 0.0f;
                }
    };
    if (threads <= 1) { work(x_begin, x_end); return; }
    std::vector<std::thread> pool;
    const uint32_t span = x_end - x_begin;
    for (int th = 0; th < threads; ++th) {
        const uint32_t a = x_begin + (uint32_t)((uint64_t)span * th / threads);
        const uint32_t b = x_begin + (uint32_t)((uint64_t)span * (th + 1) / threads);
        if (a < b) pool.emplace_back(work, a, b);
    }
    for (auto& th : pool) th.join();
}

/* ======================================================================
 * Render: raygrid_intersect (raytracer/main.comp:24-43), kd-restart dda
 * (main.comp:46-111), shade (main.comp:115-139), ray generation
 * (main.comp:161-184).  Deviations, identical on the GPU path:
 *   - cells computed by floor() are clamped to [0, dim-1] before use (the
 *     reference would read outside the node on a rounding slip);
 *   - the dda stops after kMaxSteps inner iterations and reports a miss;
 *   - the sub-pixel jitter is a seeded integer hash instead of the
 *     time-seeded sin-hash of random.comp:3-15;
 *   - an orthographic camera (ray origin offset on the image plane, fixed
 *     direction) is provided beside the reference's perspective one;
 *   - normalize(v) is v / sqrt(dot(v, v)).
 * Hit id = level << 60 | gx << 40 | gy << 20 | gz with (gx, gy, gz) the
 * global cell coordinates of the hit cell at its level.
 * ====================================================================== */
namespace {
const float kEps = 0.0001f;                 /* utils.comp:19 */
const uint32_t kMaxSteps = 1u << 14;

struct F3 { float x, y, z; };
inline F3 add(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 muls(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 normalize(F3 a) { const float l = std::sqrt(dot(a, a)); return F3{a.x / l, a.y / l, a.z / l}; }

inline uint32_t hash32(uint32_t v) {           /* PCG output permutation */
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
inline float jitter01(uint32_t seed, uint32_t x, uint32_t y, uint32_t sample, uint32_t axis) {
    const uint32_t h = hash32(seed ^ hash32(x ^ hash32(y ^ hash32(sample * 2u + axis))));
    return (float)(h >> 8) * 5.9604644775390625e-08f;   /* 2^-24 */
}

struct Hit { float t; uint32_t node, level, cell_index; uint32_t gc[3]; };

bool dda(const orc_tree* tr, F3 orig, F3 dir, Hit* hit) {
    const orc_grid& g = tr->grid;
    const F3 inv = F3{1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
    /* raygrid_intersect */
    float t0[3], t1[3];
    const float lowv[3] = {g.low[0], g.low[1], g.low[2]};
    const float o[3] = {orig.x, orig.y, orig.z};
    const float iv[3] = {inv.x, inv.y, inv.z};
    for (int c = 0; c < 3; ++c) {
        t0[c] = ((lowv[c] + kEps) - o[c]) * iv[c];
        t1[c] = (((lowv[c] + g.world_size) - kEps) - o[c]) * iv[c];
        if (iv[c] < 0.0f) std::swap(t0[c], t1[c]);
    }
    const float t_enter = fmax_glm(t0[0], fmax_glm(t0[1], t0[2]));
    const float t_leave = fmin_glm(t1[0], fmin_glm(t1[1], t1[2]));
    if (!(t_enter <= t_leave && t_leave >= 0.0f)) return false;

    const float sgn[3] = {iv[0] >= 0.0f ? 1.0f : 0.0f, iv[1] >= 0.0f ? 1.0f : 0.0f,
                          iv[2] >= 0.0f ? 1.0f : 0.0f};
    float t_curr = fmax_glm(t_enter, 0.0f) + kEps;
    float t_max = t_curr;
    uint32_t steps = 0;
    while (t_max < t_leave - kEps) {
        uint32_t lvl = 0, node = 0;
        F3 node_pos = F3{g.low[0], g.low[1], g.low[2]};
        t_curr = t_max + kEps;
        t_max = t_leave;
        while (t_curr < t_max - kEps) {
            if (++steps > kMaxSteps) return false;
            const float cs = tr->cell_size[lvl];
            const uint32_t dim = g.dims[lvl];
            const float s = t_curr + kEps;
            const F3 p = sub(add(orig, muls(dir, s)), node_pos);
            const float norm[3] = {p.x / cs, p.y / cs, p.z / cs};
            int cell[3];
            float tn[3];
            for (int c = 0; c < 3; ++c) {
                int ci = (int)std::floor(norm[c]);
                ci = ci < 0 ? 0 : (ci > (int)dim - 1 ? (int)dim - 1 : ci);
                cell[c] = ci;
                tn[c] = s + ((((float)ci + sgn[c]) - norm[c]) * cs) * iv[c];
            }
            const uint32_t idx = (uint32_t)cell[2] + dim * (uint32_t)cell[1] +
                                 dim * dim * (uint32_t)cell[0];
            const orc_node& nd = tr->levels[lvl][node];
            const uint32_t q = idx >> 5, bit = 1u << (idx & 31u);
            const bool last = (lvl + 1 == g.n_levels);
            if (nd.leaf[q] & bit) {
                hit->t = t_curr;
                hit->node = node;
                hit->level = lvl;
                hit->cell_index = idx;
                for (int c = 0; c < 3; ++c) hit->gc[c] = nd.coords[c] * dim + (uint32_t)cell[c];
                return true;
            } else if (!last && (nd.interior[q] & bit)) {
                node = nd.child[idx];
                node_pos = add(node_pos, F3{(float)cell[0] * cs, (float)cell[1] * cs,
                                            (float)cell[2] * cs});
                t_max = fmin_glm(tn[0], fmin_glm(tn[1], tn[2]));
                ++lvl;
            } else {
                t_curr = fmin_glm(tn[0], fmin_glm(tn[1], tn[2]));
            }
        }
    }
    return false;
}
}  // namespace

extern "C" void orc_camera_basis(orc_camera* cam, const float* initial_up, float fov_deg) {
    const F3 f = F3{cam->forward[0], cam->forward[1], cam->forward[2]};
    const F3 u0 = F3{initial_up[0], initial_up[1], initial_up[2]};
    const F3 cr = F3{f.y * u0.z - f.z * u0.y, f.z * u0.x - f.x * u0.z, f.x * u0.y - f.y * u0.x};
    const F3 r = normalize(cr);
    const F3 up = F3{r.y * f.z - r.z * f.y, r.z * f.x - r.x * f.z, r.x * f.y - r.y * f.x};
    cam->right[0] = r.x; cam->right[1] = r.y; cam->right[2] = r.z;
    cam->up[0] = up.x; cam->up[1] = up.y; cam->up[2] = up.z;
    if (!cam->ortho) {
        const float rad = fov_deg * (3.14159265358979323846f / 180.0f);
        cam->screen_world[0] = 2.0f * std::tan(rad / 2.0f);
        cam->screen_world[1] = cam->screen_world[0] * ((float)cam->height / (float)cam->width);
    }
}

extern "C" void orc_render(const orc_tree* tree, const orc_camera* cam,
                           uint64_t* hit_ids, float* normals, uint8_t* rgba) {
    orc_render_rows(tree, cam, 0, 1, hit_ids, normals, rgba);
}

extern "C" void orc_render_rows(const orc_tree* tree, const orc_camera* cam,
                                uint32_t row_begin, uint32_t row_step,
                                uint64_t* hit_ids, float* normals, uint8_t* rgba) {
    const F3 pos = F3{cam->pos[0], cam->pos[1], cam->pos[2]};
    const F3 fwd = F3{cam->forward[0], cam->forward[1], cam->forward[2]};
    const F3 right = F3{cam->right[0], cam->right[1], cam->right[2]};
    const F3 up = F3{cam->up[0], cam->up[1], cam->up[2]};
    const uint32_t W = cam->width, H = cam->height, spp = cam->spp ? cam->spp : 1;
    if (row_step == 0) row_step = 1;
    for (uint32_t y = row_begin; y < H; y += row_step)
        for (uint32_t x = 0; x < W; ++x) {
            float acc[4] = {0, 0, 0, 0};
            for (uint32_t s = 0; s < spp; ++s) {
                float ox = 0.0f, oy = 0.0f;
                if (cam->jitter) {
                    ox = 2.0f * (jitter01(cam->seed, x, y, s, 0) - 0.5f);
                    oy = 2.0f * (jitter01(cam->seed, x, y, s, 1) - 0.5f);
                }
                const float dx = 2.0f * (((float)x + ox) / (float)W) - 1.0f;
                const float dy = 2.0f * (((float)y + oy) / (float)H) - 1.0f;
                F3 orig, dir;
                if (cam->ortho) {
                    const float hx = (dx * cam->screen_world[0]) * 0.5f;
                    const float hy = (dy * cam->screen_world[1]) * 0.5f;
                    orig = add(add(pos, muls(right, hx)), muls(up, hy));
                    dir = fwd;
                } else {
                    orig = pos;
                    dir = normalize(add(add(fwd, muls(right, dx * cam->screen_world[0])),
                                        muls(up, dy * cam->screen_world[1])));
                }
                Hit h;
                const uint64_t sidx = ((uint64_t)y * W + x) * spp + s;
                float col[4];
                if (dda(tree, orig, dir, &h)) {
                    if (hit_ids)
                        hit_ids[sidx] = ((uint64_t)h.level << 60) | ((uint64_t)h.gc[0] << 40) |
                                        ((uint64_t)h.gc[1] << 20) | (uint64_t)h.gc[2];
                    const orc_node& nd = tree->levels[h.level][h.node];
                    const F3 p = add(orig, muls(dir, h.t));
                    float g[4];
                    orc_eval_gradient(nd.tape.data(), tree->consts.data(), p.x, p.y, p.z,
                                      tree->grid.tape_time, g);
                    const F3 n = normalize(F3{g[0], g[1], g[2]});
                    if (normals) {
                        normals[sidx * 3 + 0] = n.x;
                        normals[sidx * 3 + 1] = n.y;
                        normals[sidx * 3 + 2] = n.z;
                    }
                    if (h.t < 0.0f) {
                        col[0] = col[1] = col[2] = col[3] = 1.0f;
                    } else {
                        float diffuse[4] = {0, 0, 0, 0};
                        for (uint32_t l = 0; l < cam->light_count; ++l) {
                            const F3 ld = F3{-cam->light_dir[l][0], -cam->light_dir[l][1],
                                             -cam->light_dir[l][2]};
                            const float it = fmax_glm(dot(ld, n), 0.0f);
                            for (int c = 0; c < 4; ++c)
                                diffuse[c] = diffuse[c] + it * cam->light_color[l][c];
                        }
                        for (int c = 0; c < 4; ++c) col[c] = cam->ambient[c] + diffuse[c];
                    }
                } else {
                    if (hit_ids) hit_ids[sidx] = ~(uint64_t)0;
                    if (normals) {
                        const float nan = std::numeric_limits<float>::quiet_NaN();
                        normals[sidx * 3 + 0] = normals[sidx * 3 + 1] = normals[sidx * 3 + 2] = nan;
                    }
                    for (int c = 0; c < 4; ++c) col[c] = cam->background[c];
                }
                for (int c = 0; c < 4; ++c) acc[c] = acc[c] + col[c];
            }
            if (rgba) {
                for (int c = 0; c < 4; ++c) {
                    float v = acc[c] / (float)spp;
                    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                    rgba[((uint64_t)y * W + x) * 4 + c] = (uint8_t)(v * 255.0f + 0.5f);
                }
            }
        }
}
