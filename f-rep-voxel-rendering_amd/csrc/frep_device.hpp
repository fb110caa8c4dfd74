// frep_device.hpp -- device-side arithmetic and tape interpreters (gfx950).
//
// Everything here is compiled with -ffp-contract=off: each +, -, *, / and
// sqrt rounds once (IEEE binary32, round to nearest), exactly like the
// reference's CPU evaluator (vulkan/src/csg/tape.cpp:274-302).
//
// Tape word layout (vulkan/include/csg/tape.h:33-39): op | out<<8 | A<<16 |
// B<<24; CONST reads pool[A | B<<8].  Tapes are read through the constant
// address space so wave-uniform instruction fetches become scalar loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frep_internal.hpp"

namespace frep {

typedef const __attribute__((address_space(4))) uint32_t* ctape_t;
typedef const __attribute__((address_space(4))) float* cconst_t;

__device__ __forceinline__ ctape_t as_const(const uint32_t* p) { return (ctape_t)p; }
__device__ __forceinline__ cconst_t as_const(const float* p) { return (cconst_t)p; }

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// GLM scalar min / max (y < x ? y : x, x < y ? y : x): exact behaviour on
// NaN and signed zero, as used by vulkan/src/csg/tape.cpp:291-292.
__device__ __forceinline__ float fmin_glm(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float fmax_glm(float x, float y) { return x < y ? y : x; }

__host__ __device__ __forceinline__ float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// ---------------------------------------------------------------------------
// Transcendentals.  Cody & Waite (1980) argument reduction and Taylor
// polynomials, using only correctly rounded operations and floor(), so the
// CPU and GPU produce identical bits (the tape compiler's constant folding
// calls the same functions on the host).
//   exp:  x = k ln2 + r, |r| <= ln2/2, two-part ln2, degree-7 polynomial,
//         result scaled by 2^(k/2) and 2^(k - k/2) (two exact products).
//   sin / cos: x = k pi/2 + r with a three-part pi/2 whose products with k
//         are exact for |k| < 4096, degree-9 / degree-10 polynomials.
// ---------------------------------------------------------------------------
//
// Each function computes its main path on an in-range stand-in argument and
// selects the special-case result at the end, with no data-dependent
// branch: inside a tape interpreter's loop a divergent branch would make the
// compiler restructure the whole uniform dispatch around it.
__host__ __device__ __forceinline__ float dev_expf(float x) {
    const bool big = x > 88.72283935546875f, tiny = x < -103.972084045410156f;
    const float xs = (x != x || big || tiny) ? 0.0f : x;
    const float kf = __builtin_floorf(xs * 1.44269502162933349609375f + 0.5f);
    const float r = (xs - kf * 0.693145751953125f) - kf * 1.428606765330187e-06f;
    float p = 1.98412701138295233250e-04f;
    p = 1.38888892251998186111e-03f + r * p;
    p = 8.33333376795053482056e-03f + r * p;
    p = 4.16666679084300994873e-02f + r * p;
    p = 1.66666671633720397949e-01f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    const int k = (int)kf;
    const int k1 = k / 2;
    const int k2 = k - k1;
    const float s1 = u2f((uint32_t)(k1 + 127) << 23);
    const float s2 = u2f((uint32_t)(k2 + 127) << 23);
    const float e = (p * s1) * s2;
    return x != x ? x : (big ? __builtin_inff() : (tiny ? 0.0f : e));
}

__host__ __device__ __forceinline__ float sin_poly(float r) {
    const float z = r * r;
    float p = 2.75573188446287530661e-06f;
    p = -1.98412701138295233250e-04f + z * p;
    p = 8.33333376795053482056e-03f + z * p;
    p = -1.66666671633720397949e-01f + z * p;
    return r + (r * z) * p;
}

__host__ __device__ __forceinline__ float cos_poly(float r) {
    const float z = r * r;
    float p = -2.75573199814971303567e-07f;
    p = 2.48015876422869041562e-05f + z * p;
    p = -1.38888892251998186111e-03f + z * p;
    p = 4.16666679084300994873e-02f + z * p;
    p = -0.5f + z * p;
    return 1.0f + z * p;
}

__host__ __device__ __forceinline__ float reduce_half_pi(float x, int* q) {
    const float kf = __builtin_floorf(x * 0.636619746685028076171875f + 0.5f);
    const float r = ((x - kf * 1.5703125f) - kf * 4.837512969970703125e-04f) -
                    kf * 7.549790126404332e-08f;
    // k mod 4 computed exactly in float for every finite k.
    *q = (int)(kf - 4.0f * __builtin_floorf(kf * 0.25f));
    return r;
}

__host__ __device__ __forceinline__ float dev_sinf(float x) {
    const bool fin = __builtin_fabsf(x) <= 3.0e38f;
    int q;
    const float r = reduce_half_pi(fin ? x : 0.0f, &q);
    const float s = sin_poly(r), c = cos_poly(r);
    // q = 0..3: s, c, -s, -c (bit selects, which stay branch-free)
    const float u = (q & 1) ? c : s;
    const float v = (q & 2) ? -u : u;
    return fin ? v : x - x;
}

__host__ __device__ __forceinline__ float dev_cosf(float x) {
    const bool fin = __builtin_fabsf(x) <= 3.0e38f;
    int q;
    const float r = reduce_half_pi(fin ? x : 0.0f, &q);
    const float s = sin_poly(r), c = cos_poly(r);
    // q = 0..3: c, -s, -c, s
    const float u = (q & 1) ? s : c;
    const float v = ((q + 1) & 2) ? -u : u;
    return fin ? v : x - x;
}

// ---------------------------------------------------------------------------
// Point rules (vulkan/shaders/tape/point.comp:19-59), K points per lane.
// ---------------------------------------------------------------------------
// Reads the next four tape words at once (one 16-byte scalar load); words
// past the end of a tape lie in the padding every tape buffer carries.
__device__ __forceinline__ void fetch4(ctape_t words, uint32_t i, uint32_t (&w)[4]) {
    w[0] = words[i]; w[1] = words[i + 1]; w[2] = words[i + 2]; w[3] = words[i + 3];
}

// Lane l of a VGPR as a wave-uniform value (v_readlane_b32; ignores EXEC).
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float rdlanef(float v, uint32_t l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}

// Tape window: 64 consecutive instruction words staged across the lanes of
// one VGPR (lane l holds word s + l), so an instruction is fetched with one
// v_readlane instead of a scalar load whose latency the interpreter would
// wait on.  Words outside [0, n) read as 0xFFFFFFFF (op 255: no constant).
// Must be called with every lane of the wavefront active.
__device__ __forceinline__ uint32_t window_load(const uint32_t* words, int32_t s, uint32_t n,
                                                uint32_t lane) {
    const int32_t i = s + (int32_t)lane;
    return (i >= 0 && i < (int32_t)n) ? words[i] : 0xFFFFFFFFu;
}
// Constant operand of each CONST word of a window (gathered once per window).
__device__ __forceinline__ float window_consts(uint32_t w, const float* consts) {
    return (w & 0xFFu) == 0u ? consts[w >> 16] : 0.0f;
}

#define FREP_BANKS8(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)

// Scheduling fence: keeps each operand's register reads together so the
// compiler opens one s_set_gpr_idx window per operand instead of one per
// register move.
#define FREP_FENCE() __builtin_amdgcn_sched_barrier(0)

// Slot index kept inside a register file of MAXS slots (tapes are validated
// on upload, so this only guards the register-file access): a mask for
// power-of-two files, a clamp otherwise.
template <int MAXS>
__device__ __forceinline__ uint32_t slot_in(uint32_t x) {
    if constexpr ((MAXS & (MAXS - 1)) == 0) return x & (MAXS - 1);
    else return x < (uint32_t)MAXS ? x : (uint32_t)(MAXS - 1);
}

// Slot field of a tape word for the interval levels' register files:
// masked to the power of two at or above MAXS, which folds into the field
// extract (one s_bfe, where a clamp costs an extra s_min per operand).
// Every slot of an uploaded tape is below its slot count (validated on
// upload, and the shortened tapes the levels write keep their parents' slot
// fields), and a tape runs on a file of at least that many slots, so the
// mask never changes a field; it only bounds the index.
template <int MAXS>
__device__ __forceinline__ uint32_t slot_field(uint32_t x) {
    constexpr uint32_t span = MAXS <= 1 ? 1u : (1u << (32 - __builtin_clz((uint32_t)MAXS - 1u)));
    return x & (span - 1u);
}

// Evaluates a tape at K <= 8 points per lane.
//
// Register file: each point has its own slot array.  The compiler keeps every
// array of <= 32 registers in VGPRs and indexes it with the wave-uniform slot
// number (s_set_gpr_idx / M0), whereas a single MAXS*K array would spill to
// scratch.  Per tape instruction the K reads of operand A share one index
// window, the K reads of B another, and the K writes of the result a third,
// so an ADD costs 3K register moves + K adds + ~6 scalar instructions.
//
// Decode: instruction words arrive four at a time with one scalar load, and
// the next group is requested before the current one is interpreted.
//
// Dispatch: a balanced test on the op-code bits (4 uniform branches) instead
// of a switch, which the compiler lowers to a long compare chain.
//
// init(k) -> float3 gives point k; out(k, v) receives its density.
template <int MAXS, int K, class Init, class Out>
__device__ __forceinline__ void run_point_tape(ctape_t tape, cconst_t consts, float t, Init init,
                                               Out out) {
    static_assert(K >= 1 && K <= 8, "1..8 points per lane");
    float s0[MAXS], s1[MAXS], s2[MAXS], s3[MAXS], s4[MAXS], s5[MAXS], s6[MAXS], s7[MAXS];
#define FREP_INIT(k)                                                         \
    if constexpr (k < K) {                                                   \
        const float3 p = init(k);                                            \
        s##k[0] = p.x; s##k[1] = p.y; s##k[2] = p.z; s##k[3] = t;            \
    }
    FREP_BANKS8(FREP_INIT)
#undef FREP_INIT
    const uint32_t n = tape[0];
    uint32_t nq[4];
    fetch4(tape, 1, nq);
    for (uint32_t i = 0; i < n; i += 4) {
        uint32_t wq[4] = {nq[0], nq[1], nq[2], nq[3]};
        if (i + 4 < n) fetch4(tape, 5 + i, nq);                 // prefetch next group
        const uint32_t m = n - i < 4 ? n - i : 4;
        // the four words rotate through wq[0] so the body is emitted once
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t w = wq[0];
            wq[0] = wq[1]; wq[1] = wq[2]; wq[2] = wq[3];
            // slot fields masked to the register file; CONST uses the raw
            // 16-bit pool index
            const uint32_t op = w & 0xFFu, o = slot_in<MAXS>((w >> 8) & 0xFFu);
            const uint32_t a = slot_in<MAXS>((w >> 16) & 0xFFu), b = slot_in<MAXS>(w >> 24);
#define FREP_RA(k) float A##k = 0.0f; if constexpr (k < K) { A##k = s##k[a]; }
#define FREP_RB(k) float B##k = 0.0f; if constexpr (k < K) { B##k = s##k[b]; }
#define FREP_B(k, EXPR) if constexpr (k < K) { r##k = EXPR; }
#define FREP_ADD(k) FREP_B(k, A##k + B##k)
#define FREP_SUB(k) FREP_B(k, A##k - B##k)
#define FREP_MUL(k) FREP_B(k, A##k * B##k)
#define FREP_DIV(k) FREP_B(k, A##k / B##k)
#define FREP_MIN(k) FREP_B(k, fmin_glm(A##k, B##k))
#define FREP_MAX(k) FREP_B(k, fmax_glm(A##k, B##k))
#define FREP_SQRT(k) FREP_B(k, __builtin_sqrtf(A##k))
#define FREP_NEG(k) FREP_B(k, -A##k)
#define FREP_COPY(k) FREP_B(k, A##k)
#define FREP_CST(k) FREP_B(k, c)
#define FREP_SIN(k) FREP_B(k, dev_sinf(A##k))
#define FREP_COS(k) FREP_B(k, dev_cosf(A##k))
#define FREP_EXP(k) FREP_B(k, dev_expf(A##k))
            FREP_BANKS8(FREP_RA)
            FREP_FENCE();
            float r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0, r5 = 0, r6 = 0, r7 = 0;
            // op codes: 0 CONST 1 SIN 2 COS 3 ADD 4 SUB 5 MUL 6 DIV 7 MIN
            //           8 MAX 9 EXP 10 NEG 11 SQRT 12 COPY
            if (op & 8u) {
                if (op & 4u) {
                    FREP_BANKS8(FREP_COPY)
                } else if (op & 2u) {
                    if (op & 1u) { FREP_BANKS8(FREP_SQRT) } else { FREP_BANKS8(FREP_NEG) }
                } else if (op & 1u) {
                    FREP_BANKS8(FREP_EXP)
                } else {
                    FREP_BANKS8(FREP_RB)
                    FREP_FENCE();
                    FREP_BANKS8(FREP_MAX)
                }
            } else if (op & 4u) {
                FREP_BANKS8(FREP_RB)
                FREP_FENCE();
                if (op & 2u) {
                    if (op & 1u) { FREP_BANKS8(FREP_MIN) } else { FREP_BANKS8(FREP_DIV) }
                } else {
                    if (op & 1u) { FREP_BANKS8(FREP_MUL) } else { FREP_BANKS8(FREP_SUB) }
                }
            } else if (op & 2u) {
                if (op & 1u) {
                    FREP_BANKS8(FREP_RB)
                    FREP_FENCE();
                    FREP_BANKS8(FREP_ADD)
                } else {
                    FREP_BANKS8(FREP_COS)
                }
            } else if (op & 1u) {
                FREP_BANKS8(FREP_SIN)
            } else {
                const float c = consts[w >> 16];
                FREP_BANKS8(FREP_CST)
            }
            FREP_FENCE();
#define FREP_WR(k) if constexpr (k < K) { s##k[o] = r##k; }
            FREP_BANKS8(FREP_WR)
#undef FREP_WR
#undef FREP_RA
#undef FREP_RB
#undef FREP_B
#undef FREP_ADD
#undef FREP_SUB
#undef FREP_MUL
#undef FREP_DIV
#undef FREP_MIN
#undef FREP_MAX
#undef FREP_SQRT
#undef FREP_NEG
#undef FREP_COPY
#undef FREP_CST
#undef FREP_SIN
#undef FREP_COS
#undef FREP_EXP
        }
    }
#define FREP_OUT(k) if constexpr (k < K) { out(k, s##k[0]); }
    FREP_BANKS8(FREP_OUT)
#undef FREP_OUT
}

// ---------------------------------------------------------------------------
// Interval rules (vulkan/shaders/tape/interval.comp:8-111) and min/max choice
// recording (interval.comp:140-155).
// ---------------------------------------------------------------------------
struct Itv { float lo, hi; };

__device__ __forceinline__ Itv imul(Itv a, Itv b) {
    const float p0 = a.lo * b.lo, p1 = a.lo * b.hi, p2 = a.hi * b.lo, p3 = a.hi * b.hi;
    return Itv{fmin_glm(fmin_glm(p0, p1), fmin_glm(p2, p3)),
               fmax_glm(fmax_glm(p0, p1), fmax_glm(p2, p3))};
}

__device__ __forceinline__ bool contains_multiple(float lo, float hi, float x) {
    const float ql = lo / x, qh = hi / x;
    const float fl = __builtin_floorf(ql), fh = __builtin_floorf(qh);
    return fl < fh || fl == ql || fh == qh;
}

__device__ __forceinline__ Itv interval_op(uint32_t op, Itv a, Itv b, float c) {
    const float kPi = 3.14159265358979323846f;
    const float two_pi = 2 * kPi, half_pi = kPi / 2;
    const float inf = __builtin_inff();
    Itv r;
    switch (op) {
        case OP_CONST: r = Itv{c, c}; break;
        // values first, then selects: no data-dependent branch (see dev_expf)
        case OP_COS: {
            const float cl = dev_cosf(a.lo), ch = dev_cosf(a.hi);
            const bool m_lo = contains_multiple(a.lo - kPi, a.hi - kPi, two_pi);
            const bool m_hi = contains_multiple(a.lo, a.hi, two_pi);
            r.lo = m_lo ? -1.0f : fmin_glm(cl, ch);
            r.hi = m_hi ? 1.0f : fmax_glm(cl, ch);
            break;
        }
        case OP_SIN: {
            const float sl = dev_sinf(a.lo), sh = dev_sinf(a.hi);
            const bool m_lo = contains_multiple(a.lo + half_pi, a.hi + half_pi, two_pi);
            const bool m_hi = contains_multiple(a.lo - half_pi, a.hi - half_pi, two_pi);
            r.lo = m_lo ? -1.0f : fmin_glm(sl, sh);
            r.hi = m_hi ? 1.0f : fmax_glm(sl, sh);
            break;
        }
        case OP_ADD: r = Itv{a.lo + b.lo, a.hi + b.hi}; break;
        case OP_SUB: r = Itv{a.lo - b.hi, a.hi - b.lo}; break;
        case OP_MUL: r = imul(a, b); break;
        case OP_DIV: {
            const bool whole = (b.lo == 0 && b.hi == 0) || (b.lo < 0 && b.hi > 0);
            const float rh = 1.0f / b.hi, rl = 1.0f / b.lo;
            const Itv q = b.lo == 0 ? Itv{rh, inf} : (b.hi == 0 ? Itv{-inf, rl} : Itv{rh, rl});
            const Itv m = imul(a, q);
            r = whole ? Itv{-inf, inf} : m;
            break;
        }
        case OP_MIN: r = Itv{fmin_glm(a.lo, b.lo), fmin_glm(a.hi, b.hi)}; break;
        case OP_MAX: r = Itv{fmax_glm(a.lo, b.lo), fmax_glm(a.hi, b.hi)}; break;
        case OP_EXP: r = Itv{dev_expf(a.lo), dev_expf(a.hi)}; break;
        case OP_NEG: r = Itv{-a.hi, -a.lo}; break;
        case OP_SQRT:
            r = Itv{__builtin_sqrtf(fmax_glm(0.0f, a.lo)), __builtin_sqrtf(fmax_glm(0.0f, a.hi))};
            break;
        default: r = a; break;
    }
    return r;
}

__device__ __forceinline__ uint32_t mm_choice(uint32_t op, Itv a, Itv b) {
    if (op == OP_MIN) return a.hi < b.lo ? MM_FIRST : (b.hi < a.lo ? MM_SECOND : MM_BOTH);
    return a.hi < b.lo ? MM_SECOND : (b.hi < a.lo ? MM_FIRST : MM_BOTH);
}

// ---------------------------------------------------------------------------
// Gradient rules (vulkan/shaders/tape/gradient.comp:24-84).  Slot = (dx, dy,
// dz, value).
// ---------------------------------------------------------------------------
struct V4 { float x, y, z, w; };

__device__ __forceinline__ V4 gradient_op(uint32_t op, V4 a, V4 b, float c) {
    V4 r;
    switch (op) {
        case OP_CONST: r = V4{0, 0, 0, c}; break;
        case OP_SIN: { const float cc = dev_cosf(a.w);
            r = V4{cc * a.x, cc * a.y, cc * a.z, dev_sinf(a.w)}; } break;
        case OP_COS: { const float ms = -dev_sinf(a.w);
            r = V4{ms * a.x, ms * a.y, ms * a.z, dev_cosf(a.w)}; } break;
        case OP_ADD: r = V4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; break;
        case OP_SUB: r = V4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; break;
        case OP_MUL:
            r = V4{a.x * b.w + a.w * b.x, a.y * b.w + a.w * b.y, a.z * b.w + a.w * b.z,
                   a.w * b.w};
            break;
        case OP_DIV: { const float d = b.w * b.w;
            r = V4{(a.x * b.w - a.w * b.x) / d, (a.y * b.w - a.w * b.y) / d,
                   (a.z * b.w - a.w * b.z) / d, a.w / b.w}; } break;
        case OP_MIN: r = (a.w < b.w) ? a : b; break;
        case OP_MAX: r = (a.w > b.w) ? a : b; break;
        case OP_EXP: { const float e = dev_expf(a.w);
            r = V4{e * a.x, e * a.y, e * a.z, e * 1.0f}; } break;
        case OP_NEG: r = V4{-a.x, -a.y, -a.z, -a.w}; break;
        case OP_SQRT: { const float sq = __builtin_sqrtf(a.w), d = 2 * sq;
            r = V4{a.x / d, a.y / d, a.z / d, sq}; } break;
        default: r = a; break;
    }
    return r;
}

}  // namespace frep
