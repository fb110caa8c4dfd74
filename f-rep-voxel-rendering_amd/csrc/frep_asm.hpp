// frep_asm.hpp -- hand-scheduled CDNA point interpreter (gfx950).
//
// M0 (the index register behind s_set_gpr_idx) is reserved by the compiler,
// which sets it before each of its own uses, so it is not listed as clobbered.
//
// Evaluates one tape at 8 points per lane with the whole instruction loop in
// a single inline-assembly block, for tapes whose ops are all in
// {CONST, ADD, SUB, MUL, MIN, MAX, NEG, SQRT, COPY} (the signed-distance
// op set; tapes using DIV/SIN/COS/EXP take the C++ interpreter).
//
// Register file: slot s of point k lives in v[16 + 8*s + k] (slot-major, the
// 8 points of a slot adjacent), so one s_set_gpr_idx window (M0 = 8*s)
// covers an operand for all 8 points.  Points arrive in slots 0..2 (x, y, z,
// pinned operands) and the density leaves in slot 0.
//
// Per tape instruction: 9 SALU to decode (slot fields masked to MAXS-1),
// one window to copy operand A into v[RA..RA+7], a balanced 4-test branch on
// the op-code bits, and one DST window writing the 8 results back.  ADD reads
// B straight through the index window; SUB/MUL/MIN/MAX copy B first.  The
// next instruction word is requested before the current one executes.
//
// Arithmetic is bit-identical to the C++ paths: IEEE add/sub/mul, GLM
// min/max as compare + select (v_cmp_lt + v_cndmask), NEG as a sign flip,
// and SQRT as the correctly rounded sequence LLVM emits for fsqrt.f32
// (v_sqrt_f32 estimate, one-ulp candidates checked with fused residuals,
// scaling for tiny inputs, +-0/+inf passed through).  Hazards the hardware
// does not interlock (VALU-written VCC/SGPR lane masks read by v_cndmask,
// v_sqrt results read by VALU) are covered with s_nop as LLVM does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frep {

typedef float v8f __attribute__((ext_vector_type(8)));

// Ops the assembly interpreter implements.
__host__ __device__ constexpr bool asm_op_ok(uint32_t op) {
    return op == 0 || op == 3 || op == 4 || op == 5 || op == 7 || op == 8 || op == 10 ||
           op == 11 || op == 12;
}

#define FREP_ASM_CLOBBER_40_63                                                                \
    "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47", "v48", "v49", "v50", "v51", "v52", \
        "v53", "v54", "v55", "v56", "v57", "v58", "v59", "v60", "v61", "v62", "v63"
#define FREP_ASM_CLOBBER_64_99                                                                \
    "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", "v76", \
        "v77", "v78", "v79", "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", "v88",    \
        "v89", "v90", "v91", "v92", "v93", "v94", "v95", "v96", "v97", "v98", "v99"
#define FREP_ASM_CLOBBER_100_163                                                               \
    "v100", "v101", "v102", "v103", "v104", "v105", "v106", "v107", "v108", "v109", "v110",    \
        "v111", "v112", "v113", "v114", "v115", "v116", "v117", "v118", "v119", "v120", "v121", \
        "v122", "v123", "v124", "v125", "v126", "v127", "v128", "v129", "v130", "v131", "v132", \
        "v133", "v134", "v135", "v136", "v137", "v138", "v139", "v140", "v141", "v142", "v143", \
        "v144", "v145", "v146", "v147", "v148", "v149", "v150", "v151", "v152", "v153", "v154", \
        "v155", "v156", "v157", "v158", "v159", "v160", "v161", "v162", "v163"

// The assembly text.  Symbols: FS first slot register, MS slot-file size,
// RA / RR / RT operand, result and sqrt temporaries.
#define FREP_ASM_POINT_BODY                                                                   \
    ".set FS, 16\n"                                                                           \
    ".set MS, %[maxs]\n"                                                                      \
    ".set RA, FS + 8*MS\n"                                                                    \
    ".set RR, RA + 8\n"                                                                       \
    ".set RT, RR + 8\n"                                                                       \
    ".irp k, 0,1,2,3,4,5,6,7\n"                                                               \
    "  v_mov_b32 v[FS + 24 + \\k], %[t]\n"                                                    \
    ".endr\n"                                                                                 \
    "s_mov_b32 %[cls], 0x260\n"                                                               \
    "s_mov_b32 %[tiny], 0xf800000\n"                                                          \
    "s_load_dword %[n], %[tape], 0x0\n"                                                       \
    "s_mov_b32 %[i], 0\n"                                                                     \
    "s_waitcnt lgkmcnt(0)\n"                                                                  \
    "s_cmp_eq_u32 %[n], 0\n"                                                                  \
    "s_cbranch_scc1 .Lfa_done%=\n"                                                            \
    "s_load_dword %[w], %[tape], 0x4\n"                                                       \
    "s_waitcnt lgkmcnt(0)\n"                                                                  \
    ".Lfa_loop%=:\n"                                                                          \
    "  s_lshl_b32 %[off], %[i], 2\n"                                                          \
    "  s_add_u32 %[off], %[off], 8\n"                                                         \
    "  s_load_dword %[wn], %[tape], %[off]\n"                                                 \
    "  s_lshr_b32 %[a], %[w], 16\n"                                                           \
    "  s_and_b32 %[a], %[a], MS - 1\n"                                                        \
    "  s_lshl_b32 %[a], %[a], 3\n"                                                            \
    "  s_lshr_b32 %[b], %[w], 24\n"                                                           \
    "  s_and_b32 %[b], %[b], MS - 1\n"                                                        \
    "  s_lshl_b32 %[b], %[b], 3\n"                                                            \
    "  s_lshr_b32 %[o], %[w], 8\n"                                                            \
    "  s_and_b32 %[o], %[o], MS - 1\n"                                                        \
    "  s_lshl_b32 %[o], %[o], 3\n"                                                            \
    "  s_set_gpr_idx_on %[a], gpr_idx(SRC0)\n"                                                \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mov_b32 v[RA + \\k], v[FS + \\k]\n"                                                \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    "  s_bitcmp1_b32 %[w], 3\n"                                                               \
    "  s_cbranch_scc1 .Lfa_hi%=\n"                                                            \
    "  s_bitcmp1_b32 %[w], 2\n"                                                               \
    "  s_cbranch_scc1 .Lfa_mid%=\n"                                                           \
    "  s_bitcmp1_b32 %[w], 1\n"                                                               \
    "  s_cbranch_scc1 .Lfa_add%=\n"                                                           \
    /* op 0: CONST, pool index = bits 16..31 */                                               \
    "  s_lshr_b32 %[c], %[w], 16\n"                                                           \
    "  s_lshl_b32 %[c], %[c], 2\n"                                                            \
    "  s_load_dword %[c], %[cst], %[c]\n"                                                     \
    "  s_waitcnt lgkmcnt(0)\n"                                                                \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mov_b32 v[RR + \\k], %[c]\n"                                                       \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    /* op 3: ADD, B read through the index window */                                          \
    ".Lfa_add%=:\n"                                                                           \
    "  s_set_gpr_idx_on %[b], gpr_idx(SRC0)\n"                                                \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_add_f32 v[RR + \\k], v[FS + \\k], v[RA + \\k]\n"                                   \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    "  s_branch .Lfa_write%=\n"                                                               \
    /* ops 4 SUB, 5 MUL, 7 MIN: B copied to RR first */                                       \
    ".Lfa_mid%=:\n"                                                                           \
    "  s_set_gpr_idx_on %[b], gpr_idx(SRC0)\n"                                                \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mov_b32 v[RR + \\k], v[FS + \\k]\n"                                                \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    "  s_bitcmp1_b32 %[w], 1\n"                                                               \
    "  s_cbranch_scc1 .Lfa_min%=\n"                                                           \
    "  s_bitcmp1_b32 %[w], 0\n"                                                               \
    "  s_cbranch_scc1 .Lfa_mul%=\n"                                                           \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_sub_f32 v[RR + \\k], v[RA + \\k], v[RR + \\k]\n"                                   \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    ".Lfa_mul%=:\n"                                                                           \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mul_f32 v[RR + \\k], v[RA + \\k], v[RR + \\k]\n"                                   \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    /* min(A,B) = B < A ? B : A; four compares into four SGPR pairs, then selects */          \
    ".Lfa_min%=:\n"                                                                           \
    "  .irp q, 0,4\n"                                                                         \
    "    v_cmp_lt_f32_e64 %[p0], v[RR + \\q + 0], v[RA + \\q + 0]\n"                          \
    "    v_cmp_lt_f32_e64 %[p1], v[RR + \\q + 1], v[RA + \\q + 1]\n"                          \
    "    v_cmp_lt_f32_e64 %[p2], v[RR + \\q + 2], v[RA + \\q + 2]\n"                          \
    "    v_cmp_lt_f32_e64 %[p3], v[RR + \\q + 3], v[RA + \\q + 3]\n"                          \
    "    v_cndmask_b32_e64 v[RR + \\q + 0], v[RA + \\q + 0], v[RR + \\q + 0], %[p0]\n"        \
    "    v_cndmask_b32_e64 v[RR + \\q + 1], v[RA + \\q + 1], v[RR + \\q + 1], %[p1]\n"        \
    "    v_cndmask_b32_e64 v[RR + \\q + 2], v[RA + \\q + 2], v[RR + \\q + 2], %[p2]\n"        \
    "    v_cndmask_b32_e64 v[RR + \\q + 3], v[RA + \\q + 3], v[RR + \\q + 3], %[p3]\n"        \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    /* ops 8 MAX, 10 NEG, 11 SQRT, 12 COPY */                                                 \
    ".Lfa_hi%=:\n"                                                                            \
    "  s_bitcmp1_b32 %[w], 2\n"                                                               \
    "  s_cbranch_scc1 .Lfa_copy%=\n"                                                          \
    "  s_bitcmp1_b32 %[w], 1\n"                                                               \
    "  s_cbranch_scc1 .Lfa_negsqrt%=\n"                                                       \
    /* max(A,B) = A < B ? B : A */                                                            \
    "  s_set_gpr_idx_on %[b], gpr_idx(SRC0)\n"                                                \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mov_b32 v[RR + \\k], v[FS + \\k]\n"                                                \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    "  .irp q, 0,4\n"                                                                         \
    "    v_cmp_lt_f32_e64 %[p0], v[RA + \\q + 0], v[RR + \\q + 0]\n"                          \
    "    v_cmp_lt_f32_e64 %[p1], v[RA + \\q + 1], v[RR + \\q + 1]\n"                          \
    "    v_cmp_lt_f32_e64 %[p2], v[RA + \\q + 2], v[RR + \\q + 2]\n"                          \
    "    v_cmp_lt_f32_e64 %[p3], v[RA + \\q + 3], v[RR + \\q + 3]\n"                          \
    "    v_cndmask_b32_e64 v[RR + \\q + 0], v[RA + \\q + 0], v[RR + \\q + 0], %[p0]\n"        \
    "    v_cndmask_b32_e64 v[RR + \\q + 1], v[RA + \\q + 1], v[RR + \\q + 1], %[p1]\n"        \
    "    v_cndmask_b32_e64 v[RR + \\q + 2], v[RA + \\q + 2], v[RR + \\q + 2], %[p2]\n"        \
    "    v_cndmask_b32_e64 v[RR + \\q + 3], v[RA + \\q + 3], v[RR + \\q + 3], %[p3]\n"        \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    ".Lfa_negsqrt%=:\n"                                                                       \
    "  s_bitcmp1_b32 %[w], 0\n"                                                               \
    "  s_cbranch_scc1 .Lfa_sqrt%=\n"                                                          \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_xor_b32 v[RR + \\k], 0x80000000, v[RA + \\k]\n"                                    \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    /* correctly rounded sqrt, one point at a time; RT..RT+3 temporaries */                   \
    ".Lfa_sqrt%=:\n"                                                                          \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mul_f32_e32 v[RT + 0], 0x4f800000, v[RA + \\k]\n"                                  \
    "    v_cmp_gt_f32_e32 vcc, %[tiny], v[RA + \\k]\n"                                        \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e32 v[RT + 0], v[RA + \\k], v[RT + 0], vcc\n"                          \
    "    v_sqrt_f32_e32 v[RT + 1], v[RT + 0]\n"                                               \
    "    s_nop 0\n"                                                                           \
    "    v_add_u32_e32 v[RT + 2], -1, v[RT + 1]\n"                                            \
    "    v_fma_f32 v[RR + \\k], -v[RT + 2], v[RT + 1], v[RT + 0]\n"                           \
    "    v_add_u32_e32 v[RT + 3], 1, v[RT + 1]\n"                                             \
    "    v_cmp_ge_f32_e64 %[p0], 0, v[RR + \\k]\n"                                            \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e64 v[RT + 2], v[RT + 1], v[RT + 2], %[p0]\n"                          \
    "    v_fma_f32 v[RT + 1], -v[RT + 3], v[RT + 1], v[RT + 0]\n"                             \
    "    v_cmp_lt_f32_e64 %[p0], 0, v[RT + 1]\n"                                              \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e64 v[RT + 2], v[RT + 2], v[RT + 3], %[p0]\n"                          \
    "    v_mul_f32_e32 v[RT + 3], 0x37800000, v[RT + 2]\n"                                    \
    "    v_cndmask_b32_e32 v[RT + 2], v[RT + 2], v[RT + 3], vcc\n"                            \
    "    v_cmp_class_f32_e64 vcc, v[RT + 0], %[cls]\n"                                        \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e32 v[RR + \\k], v[RT + 2], v[RT + 0], vcc\n"                          \
    "  .endr\n"                                                                               \
    "  s_branch .Lfa_write%=\n"                                                               \
    ".Lfa_copy%=:\n"                                                                          \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mov_b32 v[RR + \\k], v[RA + \\k]\n"                                                \
    "  .endr\n"                                                                               \
    ".Lfa_write%=:\n"                                                                         \
    "  s_set_gpr_idx_on %[o], gpr_idx(DST)\n"                                                 \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mov_b32 v[FS + \\k], v[RR + \\k]\n"                                                \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    "  s_add_u32 %[i], %[i], 1\n"                                                             \
    "  s_waitcnt lgkmcnt(0)\n"                                                                \
    "  s_mov_b32 %[w], %[wn]\n"                                                               \
    "  s_cmp_lt_u32 %[i], %[n]\n"                                                             \
    "  s_cbranch_scc1 .Lfa_loop%=\n"                                                          \
    ".Lfa_done%=:\n"

// x, y, z: the 8 points of this lane (one register each); returns their
// densities.  tape: [n, instrs] with 4 readable words past the end; consts:
// the shape's pool.  MAXS must be 8 or 16 (slot file <= 128 registers).
template <int MAXS>
__device__ __forceinline__ v8f run_point_tape_asm8(const uint32_t* tape, const float* consts,
                                                   float t, v8f x, v8f y, v8f z) {
    static_assert(MAXS == 8 || MAXS == 16, "assembly path covers 8 or 16 slots");
    uint32_t n, w, wn, a, b, o, i, off, c, cls, tiny;
    uint64_t p0, p1, p2, p3;
    const uint64_t tp = (uint64_t)tape, cp = (uint64_t)consts;
    if constexpr (MAXS == 16) {
        asm volatile(FREP_ASM_POINT_BODY
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), [n] "=&s"(n),
                       [w] "=&s"(w), [wn] "=&s"(wn), [a] "=&s"(a), [b] "=&s"(b), [o] "=&s"(o),
                       [i] "=&s"(i), [off] "=&s"(off), [c] "=&s"(c), [cls] "=&s"(cls),
                       [tiny] "=&s"(tiny), [p0] "=&s"(p0), [p1] "=&s"(p1), [p2] "=&s"(p2),
                       [p3] "=&s"(p3)
                     : [tape] "s"(tp), [cst] "s"(cp), [t] "v"(t), [maxs] "n"(MAXS)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_99, FREP_ASM_CLOBBER_100_163,
                       "vcc", "memory");
    } else {
        asm volatile(FREP_ASM_POINT_BODY
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), [n] "=&s"(n),
                       [w] "=&s"(w), [wn] "=&s"(wn), [a] "=&s"(a), [b] "=&s"(b), [o] "=&s"(o),
                       [i] "=&s"(i), [off] "=&s"(off), [c] "=&s"(c), [cls] "=&s"(cls),
                       [tiny] "=&s"(tiny), [p0] "=&s"(p0), [p1] "=&s"(p1), [p2] "=&s"(p2),
                       [p3] "=&s"(p3)
                     : [tape] "s"(tp), [cst] "s"(cp), [t] "v"(t), [maxs] "n"(MAXS)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_99, "vcc", "memory");
    }
    return x;
}

}  // namespace frep
