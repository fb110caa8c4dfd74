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
// pinned operands) and the density leaves in slot 0.  A second body
// (FREP_ASM_POINT_BODY1, run_point_tape_asm1) evaluates one point per lane
// the same way, slot s in v[16 + s], for nodes of one wavefront chunk.
//
// Tape words: staged in VGPR windows of 64 words (lane l of W0 holds word
// s + l); the constant of every CONST word of a window is gathered into C0
// once per window, so the interpreter never waits on a scalar load.  Two
// windows are kept ahead.
//
// Dispatch is threaded: once per window every lane decodes its own word
// (slot fields masked to MAXS-1 and scaled by 8, and the address of the
// word's handler in a table of fixed-size slots, one per op and in-place
// variant), and each handler reads the next word's decoded fields and
// handler with two v_readlane and ends with an s_setpc_b64 to it: one jump
// per instruction, no decode or compare chain.
// Handlers (every operand copy and result move being 4 v_mov_b64):
//   CONST          4 64-bit moves of {c, c} through the DST window;
//   ADD/SUB/MUL    in place (SRC0+DST window on the output slot) when the
//                  output slot is one of the operand slots, else operand A
//                  copied, B read through the window, result moved back;
//   NEG            in place when the output slot is the operand slot;
//   COPY           nothing when the output slot is the operand slot;
//   MIN/MAX/SQRT   operands copied, result moved back.
//
// Arithmetic is bit-identical to the C++ paths: IEEE add/sub/mul (a - b is
// computed as v_sub or v_subrev of the same operands), GLM min/max as
// compare + select (v_cmp_lt + v_cndmask), NEG as a sign flip, and SQRT as
// the correctly rounded sequence LLVM emits for fsqrt.f32 (v_sqrt_f32
// estimate, one-ulp candidates checked with fused residuals, scaling for
// tiny inputs, +-0/+inf passed through).  Hazards the hardware does not
// interlock (VALU-written SGPRs read by VALU, lane masks read by
// v_cndmask, v_sqrt results) are covered with s_nop as LLVM does.
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
#define FREP_ASM_CLOBBER_64_100                                                               \
    "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", "v76", \
        "v77", "v78", "v79", "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", "v88",    \
        "v89", "v90", "v91", "v92", "v93", "v94", "v95", "v96", "v97", "v98", "v99", "v100"
// Clobbers of the slot files minus v[WB + 0] and v[WB + 3] (window 0 and
// its constants, passed in): WB = 96 (8 slots), 112 (10), 160 (16); the
// W lists are the other window registers and NS / NT.
#define FREP_ASM_CLOBBER_64_95                                                                \
    "v64", "v65", "v66", "v67", "v68", "v69", "v70", "v71", "v72", "v73", "v74", "v75", "v76", \
        "v77", "v78", "v79", "v80", "v81", "v82", "v83", "v84", "v85", "v86", "v87", "v88",    \
        "v89", "v90", "v91", "v92", "v93", "v94", "v95"
#define FREP_ASM_CLOBBER_W8 "v97", "v98", "v100", "v101", "v102"
#define FREP_ASM_CLOBBER_101_111                                                              \
    "v101", "v102", "v103", "v104", "v105", "v106", "v107", "v108", "v109", "v110", "v111"
#define FREP_ASM_CLOBBER_W10 "v113", "v114", "v116", "v117", "v118"
#define FREP_ASM_CLOBBER_101_159                                                              \
    "v101", "v102", "v103", "v104", "v105", "v106", "v107", "v108", "v109", "v110",            \
        "v111", "v112", "v113", "v114", "v115", "v116", "v117", "v118", "v119", "v120", "v121", \
        "v122", "v123", "v124", "v125", "v126", "v127", "v128", "v129", "v130", "v131", "v132", \
        "v133", "v134", "v135", "v136", "v137", "v138", "v139", "v140", "v141", "v142", "v143", \
        "v144", "v145", "v146", "v147", "v148", "v149", "v150", "v151", "v152", "v153", "v154", \
        "v155", "v156", "v157", "v158", "v159"
#define FREP_ASM_CLOBBER_W16 "v161", "v162", "v164", "v165", "v166"
#define FREP_ASM_CLOBBER_101_116                                                               \
    "v101", "v102", "v103", "v104", "v105", "v106", "v107", "v108", "v109", "v110",            \
        "v111", "v112", "v113", "v114", "v115", "v116"
#define FREP_ASM_CLOBBER_101_164                                                               \
    "v101", "v102", "v103", "v104", "v105", "v106", "v107", "v108", "v109", "v110",            \
        "v111", "v112", "v113", "v114", "v115", "v116", "v117", "v118", "v119", "v120", "v121", \
        "v122", "v123", "v124", "v125", "v126", "v127", "v128", "v129", "v130", "v131", "v132", \
        "v133", "v134", "v135", "v136", "v137", "v138", "v139", "v140", "v141", "v142", "v143", \
        "v144", "v145", "v146", "v147", "v148", "v149", "v150", "v151", "v152", "v153", "v154", \
        "v155", "v156", "v157", "v158", "v159", "v160", "v161", "v162", "v163", "v164"

// Building blocks of the assembly text.  Symbols: FS first slot register,
// MS slot-file size, LG log2(MS), RA / RB operand and result registers,
// WB tape windows (W0 W1 W2 = WB+0..2, C0 C1 = WB+3..4), NS / NT the
// decoded fields and handler offsets of the window's words shifted down one
// lane (lane l: word l + 1).
// Operand and result moves go two registers at a time (v_mov_b64 issues at
// the rate of v_mov_b32, and the index window offsets 64-bit operands like
// 32-bit ones); the slot file, RA and RB start at even registers.
//
// A word's decoded fields sit in one SGPR: a, b, o scaled by the slot size
// (8 registers, or 1 in the one-voxel body; shift KSH) in bits 0..7, 8..15
// and 16..23.  s_set_gpr_idx_on takes its index from bits 0..7 of the
// register, so operand A indexes from it directly and B / O after a shift.
#define FREP_ASM_IDX_B "  s_lshr_b32 %[tmp], %[sel], 8\n"
#define FREP_ASM_IDX_O "  s_lshr_b32 %[tmp], %[sel], 16\n"
#define FREP_ASM_READ(R, S)                                                                   \
    "  s_set_gpr_idx_on " S ", gpr_idx(SRC0)\n"                                               \
    "  .irp k, 0,2,4,6\n"                                                                     \
    "    v_mov_b64 v[" R " + \\k:" R " + \\k + 1], v[FS + \\k:FS + \\k + 1]\n"                 \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM_READ_A(R) FREP_ASM_READ(R, "%[sel]")
#define FREP_ASM_READ_B(R) FREP_ASM_IDX_B FREP_ASM_READ(R, "%[tmp]")
#define FREP_ASM_WRITE(R)                                                                     \
    FREP_ASM_IDX_O                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  .irp k, 0,2,4,6\n"                                                                     \
    "    v_mov_b64 v[FS + \\k:FS + \\k + 1], v[" R " + \\k:" R " + \\k + 1]\n"                 \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"
// Window of tape words starting at word %[sidx] (+ lane), index clamped to
// the last word, into v[W]; uses RA+0..1.
#define FREP_ASM_LOADW(W)                                                                     \
    "  v_mbcnt_lo_u32_b32 v[RA + 1], -1, 0\n"                                                 \
    "  v_mbcnt_hi_u32_b32 v[RA + 1], -1, v[RA + 1]\n"                                         \
    "  v_add_u32_e32 v[RA + 0], %[sidx], v[RA + 1]\n"                                         \
    "  v_min_u32_e32 v[RA + 0], %[nm1], v[RA + 0]\n"                                          \
    "  v_lshlrev_b32_e32 v[RA + 0], 2, v[RA + 0]\n"                                           \
    "  global_load_dword v[" W "], v[RA + 0], %[tw]\n"
// Constant of each CONST word of window v[W] (pool index 0 otherwise) into
// v[C]; v[W] must have arrived.  Uses RA+2..3.
#define FREP_ASM_GATHER(C, W)                                                                 \
    "  v_and_b32_e32 v[RA + 2], 0xff, v[" W "]\n"                                             \
    "  v_cmp_eq_u32_e32 vcc, 0, v[RA + 2]\n"                                                  \
    "  v_lshrrev_b32_e32 v[RA + 3], 16, v[" W "]\n"                                           \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e32 v[RA + 3], 0, v[RA + 3], vcc\n"                                      \
    "  v_lshlrev_b32_e32 v[RA + 3], 2, v[RA + 3]\n"                                           \
    "  global_load_dword v[" C "], v[RA + 3], %[cst]\n"
// CONST words fused with the word after them: a lane whose word is a CONST
// (slot k) followed, in this window, by an ADD/SUB/MUL that reads slot k
// once and works in place -- on its other operand (FV_A: a, FV_B: b, slot k
// still written) or on slot k itself (FC_A: k is operand a, FC_B: k is b) --
// takes the fused handler of region 13 + (op - 3), slot FV_A 0, FV_B 1,
// FC_A 2, FC_B 3, which does both words.  NS (the next word's fields) is set
// here.  Temporaries RB+0..7, masks p0..p3.
#define FREP_ASM_FUSE                                                                         \
    "  s_nop 1\n"                                                                             \
    "  v_mov_b32_dpp v[NS], v[RA + 5] wave_shl:1 bound_ctrl:0\n"                              \
    "  v_and_b32_e32 v[RB + 0], 0xff, v[WB + 0]\n"                /* op */                    \
    "  s_nop 1\n"                                                                             \
    "  v_mov_b32_dpp v[RB + 1], v[RB + 0] wave_shl:1 bound_ctrl:0\n" /* next op */            \
    "  v_bfe_u32 v[RB + 2], v[NS], 0, 8\n"                       /* next 8a */               \
    "  v_bfe_u32 v[RB + 3], v[NS], 8, 8\n"                       /* next 8b */               \
    "  v_bfe_u32 v[RB + 4], v[NS], 16, 8\n"                      /* next 8o */               \
    "  v_bfe_u32 v[RB + 5], v[RA + 5], 16, 8\n"                  /* 8k: this word's 8o */    \
    "  v_subrev_u32_e32 v[RB + 6], 3, v[RB + 1]\n"               /* next op - 3 */           \
    "  v_add_u32_e32 v[RB + 7], %[s], v[RA + 1]\n"               /* word index */            \
    /* p0: a CONST followed by ADD/SUB/MUL within this window and the tape */                 \
    "  v_cmp_eq_u32_e64 %[p0], 0, v[RB + 0]\n"                                                \
    "  v_cmp_gt_u32_e64 %[p1], 3, v[RB + 6]\n"                                                \
    "  s_and_b64 %[p0], %[p0], %[p1]\n"                                                       \
    "  v_cmp_gt_u32_e64 %[p1], %[nm1], v[RB + 7]\n"                                           \
    "  s_and_b64 %[p0], %[p0], %[p1]\n"                                                       \
    "  v_cmp_gt_u32_e64 %[p1], 63, v[RA + 1]\n"                                               \
    "  s_and_b64 %[p0], %[p0], %[p1]\n"                                                       \
    /* p3: ... whose operand a alone is slot k, p2: whose operand b alone is */               \
    "  v_cmp_eq_u32_e64 %[p1], v[RB + 2], v[RB + 5]\n"                                        \
    "  v_cmp_eq_u32_e64 %[p2], v[RB + 3], v[RB + 5]\n"                                        \
    "  s_andn2_b64 %[p3], %[p1], %[p2]\n"                                                     \
    "  s_andn2_b64 %[p2], %[p2], %[p1]\n"                                                     \
    "  s_and_b64 %[p3], %[p3], %[p0]\n"                                                       \
    "  s_and_b64 %[p2], %[p2], %[p0]\n"                                                       \
    /* slot in v[RB + 7], fused lanes in p0: FC_A 2 and FC_B 3 (next o == k), */              \
    /* FV_B 1 (a is k, next o == b), FV_A 0 (b is k, next o == a) */                          \
    "  v_cmp_eq_u32_e64 %[p1], v[RB + 4], v[RB + 5]\n"                                        \
    "  s_and_b64 vcc, %[p3], %[p1]\n"                                                         \
    "  v_cndmask_b32_e64 v[RB + 7], 0, 2, vcc\n"                                              \
    "  s_mov_b64 %[p0], vcc\n"                                                                \
    "  s_and_b64 vcc, %[p2], %[p1]\n"                                                         \
    "  v_cndmask_b32_e64 v[RB + 7], v[RB + 7], 3, vcc\n"                                      \
    "  s_or_b64 %[p0], %[p0], vcc\n"                                                          \
    "  v_cmp_eq_u32_e64 %[p1], v[RB + 4], v[RB + 3]\n"                                        \
    "  s_and_b64 vcc, %[p3], %[p1]\n"                                                         \
    "  v_cndmask_b32_e64 v[RB + 7], v[RB + 7], 1, vcc\n"                                      \
    "  s_or_b64 %[p0], %[p0], vcc\n"                                                          \
    "  v_cmp_eq_u32_e64 %[p1], v[RB + 4], v[RB + 2]\n"                                        \
    "  s_and_b64 vcc, %[p2], %[p1]\n"                                                         \
    "  s_or_b64 %[p0], %[p0], vcc\n"                                                          \
    /* fused lanes: region 13 + (next op - 3), slot v[RB + 7] */                              \
    "  v_add_u32_e32 v[RB + 6], 13, v[RB + 6]\n"                                              \
    "  v_lshlrev_b32_e32 v[RB + 6], ROPSH, v[RB + 6]\n"                                       \
    "  v_lshl_or_b32 v[RB + 6], v[RB + 7], RVSH, v[RB + 6]\n"                                 \
    "  v_cndmask_b32_e64 v[RA + 7], v[RA + 7], v[RB + 6], %[p0]\n"

// Decode window W0 (words %[s] + lane), all lanes at once: RA+5 the fields,
// RA+7 the handler offset -- (op & 15) << ROPSH plus variant << RVSH, variant 1
// when the output slot is operand A's, else 2 when it is B's, else 0, and 3
// when A, B and the output are one slot; kept only for ops with in-place
// forms (the mask of allowed variants is 2 bits per op in the word
// 0x1100FC0: ADD, SUB, MUL all, NEG and COPY the first, which 3 becomes)
// -- or the end for words past the tape.  NS / NT get them shifted down a
// lane, NT's last lane the next window's decode (ADV) or the end; the first
// word's fields and offset go to %[sel] and %[ntg], %[j] = 0.
#define FREP_ASM_DECODE                                                                       \
    "  v_mbcnt_lo_u32_b32 v[RA + 1], -1, 0\n"                                                 \
    "  v_mbcnt_hi_u32_b32 v[RA + 1], -1, v[RA + 1]\n"                                         \
    "  v_bfe_u32 v[RA + 2], v[WB + 0], 16, LG\n"                                              \
    "  v_bfe_u32 v[RA + 3], v[WB + 0], 24, LG\n"                                              \
    "  v_bfe_u32 v[RA + 4], v[WB + 0], 8, LG\n"                                               \
    "  v_lshlrev_b32_e32 v[RA + 5], KSH, v[RA + 2]\n"                                         \
    "  v_lshl_or_b32 v[RA + 5], v[RA + 3], 8 + KSH, v[RA + 5]\n"                              \
    "  v_lshl_or_b32 v[RA + 5], v[RA + 4], 16 + KSH, v[RA + 5]\n"                             \
    "  v_cmp_eq_u32_e32 vcc, v[RA + 4], v[RA + 3]\n"                                          \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e64 v[RA + 6], 0, 2, vcc\n"                                              \
    "  v_cmp_eq_u32_e32 vcc, v[RA + 4], v[RA + 2]\n"                                          \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e64 v[RA + 6], v[RA + 6], 1, vcc\n"                                      \
    "  v_cmp_eq_u32_e32 vcc, v[RA + 2], v[RA + 3]\n"                                          \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e64 v[RA + 3], 0, 2, vcc\n"                                              \
    "  v_lshlrev_b32_e32 v[RA + 4], 1, v[RA + 6]\n"                                           \
    "  v_and_or_b32 v[RA + 6], v[RA + 3], v[RA + 4], v[RA + 6]\n"                             \
    "  v_and_b32_e32 v[RA + 7], 15, v[WB + 0]\n"                                              \
    "  v_lshlrev_b32_e32 v[RA + 2], 1, v[RA + 7]\n"                                           \
    "  s_mov_b32 %[tmp], 0x1100fc0\n"                                                         \
    "  v_lshrrev_b32_e64 v[RA + 2], v[RA + 2], %[tmp]\n"                                      \
    "  v_and_b32_e32 v[RA + 6], v[RA + 6], v[RA + 2]\n"                                       \
    "  v_lshlrev_b32_e32 v[RA + 7], ROPSH, v[RA + 7]\n"                                       \
    "  v_lshl_or_b32 v[RA + 7], v[RA + 6], RVSH, v[RA + 7]\n"                                 \
    FREP_ASM_FUSE                                                                             \
    "  v_add_u32_e32 v[RA + 2], %[s], v[RA + 1]\n"                                            \
    "  v_cmp_gt_u32_e32 vcc, %[n], v[RA + 2]\n"                                               \
    "  v_mov_b32_e32 v[RA + 3], .Lfa_end%= - .Lfa_base%=\n"                                   \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e32 v[RA + 7], v[RA + 3], v[RA + 7], vcc\n"                              \
    "  s_nop 1\n"                                                                             \
    "  v_mov_b32_dpp v[NT], v[RA + 7] wave_shl:1 bound_ctrl:0\n"                              \
    "  s_add_u32 %[tmp], %[s], 64\n"                                                          \
    "  s_cmp_lt_u32 %[tmp], %[n]\n"                                                           \
    "  s_mov_b32 %[tmp], .Lfa_end%= - .Lfa_base%=\n"                                          \
    "  s_cselect_b32 %[tmp], .Lfa_adv%= - .Lfa_base%=, %[tmp]\n"                              \
    "  v_writelane_b32 v[NT], %[tmp], 63\n"                                                   \
    "  v_readlane_b32 %[sel], v[RA + 5], 0\n"                                                 \
    "  v_readlane_b32 %[ntg], v[RA + 7], 0\n"                                                 \
    "  s_mov_b32 %[j], 0\n"                                                                   \
    "  s_add_u32 vcc_lo, %[blo], %[ntg]\n"                                                    \
    "  s_addc_u32 vcc_hi, %[bhi], 0\n"                                                        \
    "  s_setpc_b64 vcc\n"
// Every handler starts by reading the next word's handler offset (lane %[j]
// of NT) and ends, once it is done with %[sel], by reading the next word's
// fields into it (lane %[j] of NS) and jumping to the handler.
#define FREP_ASM_NEXT "  v_readlane_b32 %[ntg], v[NT], %[j]\n"
#define FREP_ASM_GO                                                                           \
    "  v_readlane_b32 %[sel], v[NS], %[j]\n"                                                  \
    "  s_add_u32 vcc_lo, %[blo], %[ntg]\n"                                                    \
    "  s_addc_u32 vcc_hi, %[bhi], 0\n"                                                        \
    "  s_add_u32 %[j], %[j], 1\n"                                                             \
    "  s_setpc_b64 vcc\n"
// Handler slot of op OP, variant V (1 << ROPSH bytes per op, 1 << RVSH per variant;
// .org fails the build if a handler outgrows its slot).
#define FREP_ASM_SLOT(OP, V) "  .org .Lfa_base%= + (" #OP " << ROPSH) + (" #V " << RVSH)\n"
#define FREP_ASM_NOP FREP_ASM_NEXT FREP_ASM_GO
// In-place binary op: v[S] = v[S] OP v[R].
#define FREP_ASM_INPLACE(OPC, S, R)                                                           \
    "  s_set_gpr_idx_on " S ", gpr_idx(SRC0,DST)\n"                                           \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    " OPC " v[FS + \\k], v[FS + \\k], v[" R " + \\k]\n"                                 \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"
// Binary op with variants: 0 RB = v[b] OP RA (RA = a), moved to the output
// slot; 1 in place on a (B copied to RB); 2 in place on b (A copied to RA);
// 3 a, b and the output one slot (x OP x in place, no copy).
// OPC_GEN / OPC_A / OPC_B are the opcodes of the three forms.
#define FREP_ASM_BINARY(OP, OPC_GEN, OPC_A, OPC_B)                                            \
    FREP_ASM_SLOT(OP, 0)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM_READ_A("RA")                                                                     \
    FREP_ASM_IDX_B                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(SRC0)\n"                                              \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    " OPC_GEN " v[RB + \\k], v[FS + \\k], v[RA + \\k]\n"                                \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_WRITE("RB")                                                                      \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(OP, 1)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM_READ_B("RB")                                                                     \
    FREP_ASM_INPLACE(OPC_A, "%[sel]", "RB")                                                   \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(OP, 2)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM_READ_A("RA")                                                                     \
    FREP_ASM_IDX_B                                                                            \
    FREP_ASM_INPLACE(OPC_B, "%[tmp]", "RA")                                                   \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(OP, 3)                                                                      \
    FREP_ASM_NEXT                                                                             \
    "  s_set_gpr_idx_on %[sel], gpr_idx(SRC0,SRC1,DST)\n"                                     \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    " OPC_A " v[FS + \\k], v[FS + \\k], v[FS + \\k]\n"                                 \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO
// min(A,B) = B < A ? B : A (LT_X < LT_Y selects RB); four compares into four
// SGPR pairs, then selects, the result in RB moved to the output slot.
#define FREP_ASM_MINMAX(LT_X, LT_Y)                                                           \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM_READ_A("RA")                                                                     \
    FREP_ASM_READ_B("RB")                                                                     \
    "  .irp q, 0,4\n"                                                                         \
    "    v_cmp_lt_f32_e64 %[p0], v[" LT_X " + \\q + 0], v[" LT_Y " + \\q + 0]\n"              \
    "    v_cmp_lt_f32_e64 %[p1], v[" LT_X " + \\q + 1], v[" LT_Y " + \\q + 1]\n"              \
    "    v_cmp_lt_f32_e64 %[p2], v[" LT_X " + \\q + 2], v[" LT_Y " + \\q + 2]\n"              \
    "    v_cmp_lt_f32_e64 %[p3], v[" LT_X " + \\q + 3], v[" LT_Y " + \\q + 3]\n"              \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e64 v[RB + \\q + 0], v[RA + \\q + 0], v[RB + \\q + 0], %[p0]\n"        \
    "    v_cndmask_b32_e64 v[RB + \\q + 1], v[RA + \\q + 1], v[RB + \\q + 1], %[p1]\n"        \
    "    v_cndmask_b32_e64 v[RB + \\q + 2], v[RA + \\q + 2], v[RB + \\q + 2], %[p2]\n"        \
    "    v_cndmask_b32_e64 v[RB + \\q + 3], v[RA + \\q + 3], v[RB + \\q + 3], %[p3]\n"        \
    "  .endr\n"                                                                               \
    FREP_ASM_WRITE("RB")                                                                      \
    FREP_ASM_GO

// Fused CONST + binary handlers (FREP_ASM_FUSE): c in vcc_lo, the binary
// word's fields in %[nx]; the binary word is skipped (%[j] advanced past it
// before the next word's offset is read).  FV_A / FV_B write {c, c} to slot
// k and apply the op in place on the other operand with c as a scalar
// operand; FC_A / FC_B copy the other operand and write the result to slot
// k.  OPC_* take (c, operand) in that order.
#define FREP_ASM_FUSED_HEAD                                                                   \
    "  v_readlane_b32 vcc_lo, v[WB + 3], %[j]\n"                                              \
    "  v_readlane_b32 %[nx], v[NS], %[j]\n"                                                   \
    "  s_add_u32 %[j], %[j], 1\n" FREP_ASM_NEXT
#define FREP_ASM_CONST_WRITE                                                                  \
    FREP_ASM_IDX_O                                                                            \
    "  s_mov_b32 vcc_hi, vcc_lo\n"                                                            \
    "  s_nop 1\n"                                                                             \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  .irp k, 0,2,4,6\n"                                                                     \
    "    v_mov_b64 v[FS + \\k:FS + \\k + 1], vcc\n"                                             \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM_SCALAR_OP(OPC, S, R)                                                         \
    "  s_set_gpr_idx_on " S ", gpr_idx(SRC1,DST)\n"                                           \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    " OPC " v[FS + \\k], vcc_lo, v[" R " + \\k]\n"                                        \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM_FUSED(R, OPC_VA, OPC_VB, OPC_CA, OPC_CB)                                     \
    FREP_ASM_SLOT(R, 0)                                                                       \
    FREP_ASM_FUSED_HEAD FREP_ASM_CONST_WRITE                                                  \
    FREP_ASM_SCALAR_OP(OPC_VA, "%[nx]", "FS")                                                 \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(R, 1)                                                                       \
    FREP_ASM_FUSED_HEAD FREP_ASM_CONST_WRITE                                                  \
    "  s_lshr_b32 %[tmp], %[nx], 8\n"                                                         \
    FREP_ASM_SCALAR_OP(OPC_VB, "%[tmp]", "FS")                                                \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(R, 2)                                                                       \
    FREP_ASM_FUSED_HEAD                                                                       \
    "  s_lshr_b32 %[tmp], %[nx], 8\n"                                                         \
    FREP_ASM_READ("RB", "%[tmp]")                                                             \
    FREP_ASM_IDX_O                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    " OPC_CA " v[FS + \\k], vcc_lo, v[RB + \\k]\n"                                         \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(R, 3)                                                                       \
    FREP_ASM_FUSED_HEAD                                                                       \
    FREP_ASM_READ("RB", "%[nx]")                                                              \
    FREP_ASM_IDX_O                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    " OPC_CB " v[FS + \\k], vcc_lo, v[RB + \\k]\n"                                         \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO

#define FREP_ASM_POINT_BODY                                                                   \
    ".set FS, 16\n"                                                                           \
    ".set MS, %[maxs]\n"                                                                      \
    ".set LG, %[lg]\n"                                                                        \
    /* 8 voxels per lane: a slot is 8 registers, fields scaled by 8, 1 KiB                    \
       of handlers per op and 256 bytes per variant */                                        \
    ".set KSH, 3\n"                                                                           \
    ".set ROPSH, 10\n"                                                                        \
    ".set RVSH, 8\n"                                                                          \
    ".set RA, FS + 8*MS\n"                                                                    \
    ".set RB, RA + 8\n"                                                                       \
    ".set WB, RB + 8\n"                                                                       \
    ".set NS, WB + 5\n"                                                                       \
    ".set NT, WB + 6\n"                                                                       \
    /* slot 3 is the input t when the tape reads it before writing it */                      \
    "s_cmp_eq_u32 %[ut], 0\n"                                                                 \
    "s_cbranch_scc1 .Lfa_nt%=\n"                                                              \
    ".irp k, 0,1,2,3,4,5,6,7\n"                                                               \
    "  v_mov_b32 v[FS + 24 + \\k], %[t]\n"                                                    \
    ".endr\n"                                                                                 \
    ".Lfa_nt%=:\n"                                                                            \
    "s_mov_b32 %[cls], 0x260\n"                                                               \
    "s_mov_b32 %[tiny], 0xf800000\n"                                                          \
    "s_mov_b32 %[sgn], 0x80000000\n"                                                          \
    /* handler table base: the address after s_getpc plus the distance */                    \
    "s_getpc_b64 vcc\n"                                                                       \
    ".Lfa_pc%=:\n"                                                                            \
    "s_add_u32 %[blo], vcc_lo, .Lfa_base%= - .Lfa_pc%=\n"                                     \
    "s_addc_u32 %[bhi], vcc_hi, 0\n"                                                          \
    "s_cmp_eq_u32 %[n], 0\n"                                                                  \
    "s_cbranch_scc1 .Lfa_end%=\n"                                                             \
    "s_sub_u32 %[nm1], %[n], 1\n"                                                             \
    "s_mov_b32 %[s], 0\n"                                                                     \
    /* window 0 and its constants arrive in v[WB + 0] and v[WB + 3]; the                      \
       next two windows are loaded only for tapes longer than 64 words */                     \
    "s_cmp_le_u32 %[n], 64\n"                                                                 \
    "s_cbranch_scc1 .Lfa_dec%=\n"                                                             \
    "s_mov_b32 %[sidx], 64\n"                                                                 \
    FREP_ASM_LOADW("WB + 1")                                                                  \
    "s_mov_b32 %[sidx], 128\n"                                                                \
    FREP_ASM_LOADW("WB + 2")                                                                  \
    "s_waitcnt vmcnt(0)\n"                                                                    \
    FREP_ASM_GATHER("WB + 4", "WB + 1")                                                       \
    "s_waitcnt vmcnt(0)\n"                                                                    \
    ".Lfa_dec%=:\n"                                                                           \
    FREP_ASM_DECODE                                                                           \
    ".Lfa_base%=:\n"                                                                          \
    /* op 0: CONST -- the window's constant of this word as the pair {c, c} in                \
       vcc, written to the slot two registers at a time */                                    \
    FREP_ASM_NEXT                                                                             \
    "  v_readlane_b32 vcc_lo, v[WB + 3], %[j]\n"                                              \
    FREP_ASM_IDX_O                                                                            \
    "  s_mov_b32 vcc_hi, vcc_lo\n"                                                            \
    "  s_nop 1\n"                                                                             \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  .irp k, 0,2,4,6\n"                                                                     \
    "    v_mov_b64 v[FS + \\k:FS + \\k + 1], vcc\n"                                             \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO                                                                               \
    /* ops 1 SIN and 2 COS never reach this interpreter */                                    \
    FREP_ASM_SLOT(1, 0) FREP_ASM_NOP                                                          \
    FREP_ASM_SLOT(2, 0) FREP_ASM_NOP                                                          \
    /* op 3 ADD, 4 SUB (a - b as v_sub with A in place, v_subrev otherwise), 5 MUL */         \
    FREP_ASM_BINARY(3, "v_add_f32_e32", "v_add_f32_e32", "v_add_f32_e32")                     \
    FREP_ASM_BINARY(4, "v_subrev_f32_e32", "v_sub_f32_e32", "v_subrev_f32_e32")               \
    FREP_ASM_BINARY(5, "v_mul_f32_e32", "v_mul_f32_e32", "v_mul_f32_e32")                     \
    FREP_ASM_SLOT(6, 0) FREP_ASM_NOP                                                          \
    /* op 7 MIN: B < A ? B : A; op 8 MAX: A < B ? B : A */                                    \
    FREP_ASM_SLOT(7, 0) FREP_ASM_MINMAX("RB", "RA")                                           \
    FREP_ASM_SLOT(8, 0) FREP_ASM_MINMAX("RA", "RB")                                           \
    FREP_ASM_SLOT(9, 0) FREP_ASM_NOP                                                          \
    /* op 10 NEG: sign flip, in place when the output slot is the operand slot */             \
    FREP_ASM_SLOT(10, 0)                                                                      \
    FREP_ASM_NEXT                                                                             \
    "  s_set_gpr_idx_on %[sel], gpr_idx(SRC0)\n"                                              \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_xor_b32_e64 v[RB + \\k], v[FS + \\k], %[sgn]\n"                                    \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_WRITE("RB")                                                                      \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(10, 1)                                                                      \
    FREP_ASM_NEXT                                                                             \
    "  s_set_gpr_idx_on %[sel], gpr_idx(SRC0,DST)\n"                                          \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_xor_b32_e64 v[FS + \\k], v[FS + \\k], %[sgn]\n"                                    \
    "  .endr\n"                                                                               \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO                                                                               \
    /* op 11 SQRT: its body follows the table */                                              \
    FREP_ASM_SLOT(11, 0)                                                                      \
    "  s_branch .Lfa_sqrt%=\n"                                                                \
    /* op 12 COPY: moved through RA, nothing to do in place */                                \
    FREP_ASM_SLOT(12, 0)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM_READ_A("RA")                                                                     \
    FREP_ASM_WRITE("RA")                                                                      \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(12, 1) FREP_ASM_NOP                                                         \
    /* regions 13-15 (op codes no tape uses): a CONST fused with the ADD, SUB or              \
       MUL after it; SUB is a - b as v_sub (c - x) or v_subrev (x - c) */                     \
    FREP_ASM_FUSED(13, "v_add_f32_e32", "v_add_f32_e32", "v_add_f32_e32", "v_add_f32_e32")     \
    FREP_ASM_FUSED(14, "v_subrev_f32_e32", "v_sub_f32_e32", "v_sub_f32_e32", "v_subrev_f32_e32") \
    FREP_ASM_FUSED(15, "v_mul_f32_e32", "v_mul_f32_e32", "v_mul_f32_e32", "v_mul_f32_e32")     \
    FREP_ASM_SLOT(16, 0)                                                                      \
    /* next window: the loads issued one window ago have had a window to                      \
       arrive */                                                                              \
    ".Lfa_adv%=:\n"                                                                           \
    "  s_add_u32 %[s], %[s], 64\n"                                                            \
    "  s_waitcnt vmcnt(0)\n"                                                                  \
    "  v_mov_b32 v[WB + 0], v[WB + 1]\n"                                                      \
    "  v_mov_b32 v[WB + 3], v[WB + 4]\n"                                                      \
    "  v_mov_b32 v[WB + 1], v[WB + 2]\n"                                                      \
    FREP_ASM_GATHER("WB + 4", "WB + 1")                                                       \
    "  s_add_u32 %[sidx], %[s], 128\n"                                                        \
    FREP_ASM_LOADW("WB + 2")                                                                  \
    FREP_ASM_DECODE                                                                           \
    /* correctly rounded sqrt, one point at a time: RB+0..4 temporaries, the                  \
       result replaces the operand copy in RA */                                              \
    ".Lfa_sqrt%=:\n"                                                                          \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM_READ_A("RA")                                                                     \
    "  .irp k, 0,1,2,3,4,5,6,7\n"                                                             \
    "    v_mul_f32_e32 v[RB + 0], 0x4f800000, v[RA + \\k]\n"                                  \
    "    v_cmp_gt_f32_e32 vcc, %[tiny], v[RA + \\k]\n"                                        \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e32 v[RB + 0], v[RA + \\k], v[RB + 0], vcc\n"                          \
    "    v_sqrt_f32_e32 v[RB + 1], v[RB + 0]\n"                                               \
    "    s_nop 0\n"                                                                           \
    "    v_add_u32_e32 v[RB + 2], -1, v[RB + 1]\n"                                            \
    "    v_fma_f32 v[RB + 4], -v[RB + 2], v[RB + 1], v[RB + 0]\n"                             \
    "    v_add_u32_e32 v[RB + 3], 1, v[RB + 1]\n"                                             \
    "    v_cmp_ge_f32_e64 %[p0], 0, v[RB + 4]\n"                                              \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e64 v[RB + 2], v[RB + 1], v[RB + 2], %[p0]\n"                          \
    "    v_fma_f32 v[RB + 1], -v[RB + 3], v[RB + 1], v[RB + 0]\n"                             \
    "    v_cmp_lt_f32_e64 %[p0], 0, v[RB + 1]\n"                                              \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e64 v[RB + 2], v[RB + 2], v[RB + 3], %[p0]\n"                          \
    "    v_mul_f32_e32 v[RB + 3], 0x37800000, v[RB + 2]\n"                                    \
    "    v_cndmask_b32_e32 v[RB + 2], v[RB + 2], v[RB + 3], vcc\n"                            \
    "    v_cmp_class_f32_e64 vcc, v[RB + 0], %[cls]\n"                                        \
    "    s_nop 1\n"                                                                           \
    "    v_cndmask_b32_e32 v[RA + \\k], v[RB + 2], v[RB + 0], vcc\n"                          \
    "  .endr\n"                                                                               \
    FREP_ASM_WRITE("RA")                                                                      \
    FREP_ASM_GO                                                                               \
    /* no load may still be writing these registers when the block ends */                    \
    ".Lfa_end%=:\n"                                                                           \
    "s_waitcnt vmcnt(0)\n"

// One voxel per lane (FREP_ASM_POINT_BODY1, for last levels whose nodes
// are one wavefront chunk of 64 voxels): the same window decode, CONST
// fusion and threaded dispatch with a slot of one register (slot s is
// register FS + s, fields unscaled) and handlers of one instruction where
// the 8-voxel ones take eight, in slots of 512 bytes per op and 128 per
// variant.
#define FREP_ASM1_READ(R, S)                                                                  \
    "  s_set_gpr_idx_on " S ", gpr_idx(SRC0)\n"                                               \
    "  v_mov_b32 v[" R "], v[FS]\n"                                                           \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM1_READ_A(R) FREP_ASM1_READ(R, "%[sel]")
#define FREP_ASM1_READ_B(R) FREP_ASM_IDX_B FREP_ASM1_READ(R, "%[tmp]")
#define FREP_ASM1_WRITE(R)                                                                    \
    FREP_ASM_IDX_O                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  v_mov_b32 v[FS], v[" R "]\n"                                                           \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM1_INPLACE(OPC, S, R)                                                          \
    "  s_set_gpr_idx_on " S ", gpr_idx(SRC0,DST)\n"                                           \
    "  " OPC " v[FS], v[FS], v[" R "]\n"                                                      \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM1_BINARY(OP, OPC_GEN, OPC_A, OPC_B)                                           \
    FREP_ASM_SLOT(OP, 0)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM1_READ_A("RA")                                                                    \
    FREP_ASM_IDX_B                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(SRC0)\n"                                              \
    "  " OPC_GEN " v[RB], v[FS], v[RA]\n"                                                     \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM1_WRITE("RB")                                                                     \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(OP, 1)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM1_READ_B("RB")                                                                    \
    FREP_ASM1_INPLACE(OPC_A, "%[sel]", "RB")                                                  \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(OP, 2)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM1_READ_A("RA")                                                                    \
    FREP_ASM_IDX_B                                                                            \
    FREP_ASM1_INPLACE(OPC_B, "%[tmp]", "RA")                                                  \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(OP, 3)                                                                      \
    FREP_ASM_NEXT                                                                             \
    "  s_set_gpr_idx_on %[sel], gpr_idx(SRC0,SRC1,DST)\n"                                     \
    "  " OPC_A " v[FS], v[FS], v[FS]\n"                                                       \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO
#define FREP_ASM1_MINMAX(LT_X, LT_Y)                                                          \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM1_READ_A("RA")                                                                    \
    FREP_ASM1_READ_B("RB")                                                                    \
    "  v_cmp_lt_f32_e64 %[p0], v[" LT_X "], v[" LT_Y "]\n"                                    \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e64 v[RB], v[RA], v[RB], %[p0]\n"                                        \
    FREP_ASM1_WRITE("RB")                                                                     \
    FREP_ASM_GO
#define FREP_ASM1_CONST_WRITE                                                                 \
    FREP_ASM_IDX_O                                                                            \
    "  s_nop 1\n"                                                                             \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  v_mov_b32 v[FS], vcc_lo\n"                                                             \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM1_SCALAR_OP(OPC, S)                                                           \
    "  s_set_gpr_idx_on " S ", gpr_idx(SRC1,DST)\n"                                           \
    "  " OPC " v[FS], vcc_lo, v[FS]\n"                                                        \
    "  s_set_gpr_idx_off\n"
#define FREP_ASM1_FUSED(R, OPC_VA, OPC_VB, OPC_CA, OPC_CB)                                    \
    FREP_ASM_SLOT(R, 0)                                                                       \
    FREP_ASM_FUSED_HEAD FREP_ASM1_CONST_WRITE                                                 \
    FREP_ASM1_SCALAR_OP(OPC_VA, "%[nx]")                                                      \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(R, 1)                                                                       \
    FREP_ASM_FUSED_HEAD FREP_ASM1_CONST_WRITE                                                 \
    "  s_lshr_b32 %[tmp], %[nx], 8\n"                                                         \
    FREP_ASM1_SCALAR_OP(OPC_VB, "%[tmp]")                                                     \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(R, 2)                                                                       \
    FREP_ASM_FUSED_HEAD                                                                       \
    "  s_lshr_b32 %[tmp], %[nx], 8\n"                                                         \
    FREP_ASM1_READ("RB", "%[tmp]")                                                            \
    FREP_ASM_IDX_O                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  " OPC_CA " v[FS], vcc_lo, v[RB]\n"                                                     \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(R, 3)                                                                       \
    FREP_ASM_FUSED_HEAD                                                                       \
    FREP_ASM1_READ("RB", "%[nx]")                                                             \
    FREP_ASM_IDX_O                                                                            \
    "  s_set_gpr_idx_on %[tmp], gpr_idx(DST)\n"                                               \
    "  " OPC_CB " v[FS], vcc_lo, v[RB]\n"                                                     \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO
#define FREP_ASM_POINT_BODY1                                                                  \
    ".set FS, 16\n"                                                                           \
    ".set MS, %[maxs]\n"                                                                      \
    ".set LG, %[lg]\n"                                                                        \
    ".set KSH, 0\n"                                                                           \
    ".set ROPSH, 9\n"                                                                         \
    ".set RVSH, 7\n"                                                                          \
    ".set RA, FS + MS\n"                                                                      \
    ".set RB, RA + 8\n"                                                                       \
    ".set WB, RB + 8\n"                                                                       \
    ".set NS, WB + 5\n"                                                                       \
    ".set NT, WB + 6\n"                                                                       \
    "s_cmp_eq_u32 %[ut], 0\n"                                                                 \
    "s_cbranch_scc1 .Lfa_nt%=\n"                                                              \
    "v_mov_b32 v[FS + 3], %[t]\n"                                                             \
    ".Lfa_nt%=:\n"                                                                            \
    "s_mov_b32 %[cls], 0x260\n"                                                               \
    "s_mov_b32 %[tiny], 0xf800000\n"                                                          \
    "s_mov_b32 %[sgn], 0x80000000\n"                                                          \
    "s_getpc_b64 vcc\n"                                                                       \
    ".Lfa_pc%=:\n"                                                                            \
    "s_add_u32 %[blo], vcc_lo, .Lfa_base%= - .Lfa_pc%=\n"                                     \
    "s_addc_u32 %[bhi], vcc_hi, 0\n"                                                          \
    "s_cmp_eq_u32 %[n], 0\n"                                                                  \
    "s_cbranch_scc1 .Lfa_end%=\n"                                                             \
    "s_sub_u32 %[nm1], %[n], 1\n"                                                             \
    "s_mov_b32 %[s], 0\n"                                                                     \
    "s_cmp_le_u32 %[n], 64\n"                                                                 \
    "s_cbranch_scc1 .Lfa_dec%=\n"                                                             \
    "s_mov_b32 %[sidx], 64\n"                                                                 \
    FREP_ASM_LOADW("WB + 1")                                                                  \
    "s_mov_b32 %[sidx], 128\n"                                                                \
    FREP_ASM_LOADW("WB + 2")                                                                  \
    "s_waitcnt vmcnt(0)\n"                                                                    \
    FREP_ASM_GATHER("WB + 4", "WB + 1")                                                       \
    "s_waitcnt vmcnt(0)\n"                                                                    \
    ".Lfa_dec%=:\n"                                                                           \
    FREP_ASM_DECODE                                                                           \
    ".Lfa_base%=:\n"                                                                          \
    /* op 0: CONST */                                                                         \
    FREP_ASM_NEXT                                                                             \
    "  v_readlane_b32 vcc_lo, v[WB + 3], %[j]\n"                                              \
    FREP_ASM1_CONST_WRITE                                                                     \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(1, 0) FREP_ASM_NOP                                                          \
    FREP_ASM_SLOT(2, 0) FREP_ASM_NOP                                                          \
    FREP_ASM1_BINARY(3, "v_add_f32_e32", "v_add_f32_e32", "v_add_f32_e32")                    \
    FREP_ASM1_BINARY(4, "v_subrev_f32_e32", "v_sub_f32_e32", "v_subrev_f32_e32")              \
    FREP_ASM1_BINARY(5, "v_mul_f32_e32", "v_mul_f32_e32", "v_mul_f32_e32")                    \
    FREP_ASM_SLOT(6, 0) FREP_ASM_NOP                                                          \
    FREP_ASM_SLOT(7, 0) FREP_ASM1_MINMAX("RB", "RA")                                          \
    FREP_ASM_SLOT(8, 0) FREP_ASM1_MINMAX("RA", "RB")                                          \
    FREP_ASM_SLOT(9, 0) FREP_ASM_NOP                                                          \
    FREP_ASM_SLOT(10, 0)                                                                      \
    FREP_ASM_NEXT                                                                             \
    "  s_set_gpr_idx_on %[sel], gpr_idx(SRC0)\n"                                              \
    "  v_xor_b32_e64 v[RB], v[FS], %[sgn]\n"                                                  \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM1_WRITE("RB")                                                                     \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(10, 1)                                                                      \
    FREP_ASM_NEXT                                                                             \
    "  s_set_gpr_idx_on %[sel], gpr_idx(SRC0,DST)\n"                                          \
    "  v_xor_b32_e64 v[FS], v[FS], %[sgn]\n"                                                  \
    "  s_set_gpr_idx_off\n"                                                                   \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(11, 0)                                                                      \
    "  s_branch .Lfa_sqrt%=\n"                                                                \
    FREP_ASM_SLOT(12, 0)                                                                      \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM1_READ_A("RA")                                                                    \
    FREP_ASM1_WRITE("RA")                                                                     \
    FREP_ASM_GO                                                                               \
    FREP_ASM_SLOT(12, 1) FREP_ASM_NOP                                                         \
    FREP_ASM1_FUSED(13, "v_add_f32_e32", "v_add_f32_e32", "v_add_f32_e32", "v_add_f32_e32")    \
    FREP_ASM1_FUSED(14, "v_subrev_f32_e32", "v_sub_f32_e32", "v_sub_f32_e32", "v_subrev_f32_e32") \
    FREP_ASM1_FUSED(15, "v_mul_f32_e32", "v_mul_f32_e32", "v_mul_f32_e32", "v_mul_f32_e32")    \
    FREP_ASM_SLOT(16, 0)                                                                      \
    ".Lfa_adv%=:\n"                                                                           \
    "  s_add_u32 %[s], %[s], 64\n"                                                            \
    "  s_waitcnt vmcnt(0)\n"                                                                  \
    "  v_mov_b32 v[WB + 0], v[WB + 1]\n"                                                      \
    "  v_mov_b32 v[WB + 3], v[WB + 4]\n"                                                      \
    "  v_mov_b32 v[WB + 1], v[WB + 2]\n"                                                      \
    FREP_ASM_GATHER("WB + 4", "WB + 1")                                                       \
    "  s_add_u32 %[sidx], %[s], 128\n"                                                        \
    FREP_ASM_LOADW("WB + 2")                                                                  \
    FREP_ASM_DECODE                                                                           \
    /* correctly rounded sqrt (the sequence of the 8-voxel body, one voxel) */                \
    ".Lfa_sqrt%=:\n"                                                                          \
    FREP_ASM_NEXT                                                                             \
    FREP_ASM1_READ_A("RA")                                                                    \
    "  v_mul_f32_e32 v[RB + 0], 0x4f800000, v[RA]\n"                                          \
    "  v_cmp_gt_f32_e32 vcc, %[tiny], v[RA]\n"                                                \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e32 v[RB + 0], v[RA], v[RB + 0], vcc\n"                                  \
    "  v_sqrt_f32_e32 v[RB + 1], v[RB + 0]\n"                                                 \
    "  s_nop 0\n"                                                                             \
    "  v_add_u32_e32 v[RB + 2], -1, v[RB + 1]\n"                                              \
    "  v_fma_f32 v[RB + 4], -v[RB + 2], v[RB + 1], v[RB + 0]\n"                               \
    "  v_add_u32_e32 v[RB + 3], 1, v[RB + 1]\n"                                               \
    "  v_cmp_ge_f32_e64 %[p0], 0, v[RB + 4]\n"                                                \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e64 v[RB + 2], v[RB + 1], v[RB + 2], %[p0]\n"                            \
    "  v_fma_f32 v[RB + 1], -v[RB + 3], v[RB + 1], v[RB + 0]\n"                               \
    "  v_cmp_lt_f32_e64 %[p0], 0, v[RB + 1]\n"                                                \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e64 v[RB + 2], v[RB + 2], v[RB + 3], %[p0]\n"                            \
    "  v_mul_f32_e32 v[RB + 3], 0x37800000, v[RB + 2]\n"                                      \
    "  v_cndmask_b32_e32 v[RB + 2], v[RB + 2], v[RB + 3], vcc\n"                              \
    "  v_cmp_class_f32_e64 vcc, v[RB + 0], %[cls]\n"                                          \
    "  s_nop 1\n"                                                                             \
    "  v_cndmask_b32_e32 v[RA], v[RB + 2], v[RB + 0], vcc\n"                                  \
    FREP_ASM1_WRITE("RA")                                                                     \
    FREP_ASM_GO                                                                               \
    ".Lfa_end%=:\n"                                                                           \
    "s_waitcnt vmcnt(0)\n"

// Scalar registers of the point interpreter's assembly block.
#define FREP_ASM_POINT_DECLS                                                                  \
    uint32_t sel, ntg, nx, j, tmp, blo, bhi, s, sidx, nm1, cls, tiny, sgn;                    \
    uint64_t p0, p1, p2, p3
#define FREP_ASM_POINT_SCALARS                                                                \
    [sel] "=&s"(sel), [ntg] "=&s"(ntg), [nx] "=&s"(nx), [j] "=&s"(j), [tmp] "=&s"(tmp),       \
        [blo] "=&s"(blo), [bhi] "=&s"(bhi), [s] "=&s"(s), [sidx] "=&s"(sidx),                 \
        [nm1] "=&s"(nm1), [cls] "=&s"(cls), [tiny] "=&s"(tiny), [sgn] "=&s"(sgn),             \
        [p0] "=&s"(p0), [p1] "=&s"(p1), [p2] "=&s"(p2), [p3] "=&s"(p3)

// Prefetch around the interpreter (run_point_tape_asm8_pf2): issued before
// the body, waited after it, so nothing is in flight when the block ends.
//   pfw  lane l's word of 64 at pfa (the next tape's window 0)
//   pfn  the word at pfb (the next tape's length)
//   pfh  four words at pfc (the header of the node after that)
//   pfk  the constants of the next node, from its window pw1 (lane l's word
//        of a tape of pn1 words): the constant a CONST word names for lanes
//        below pn1, the pool's first constant otherwise
#define FREP_ASM_PF_WAIT "s_waitcnt vmcnt(0) lgkmcnt(0)\n"
#define FREP_ASM_PF2_ISSUE                                                                    \
    "v_mbcnt_lo_u32_b32 %[pft], -1, 0\n"                                                      \
    "v_mbcnt_hi_u32_b32 %[pft], -1, %[pft]\n"                                                 \
    "v_cmp_gt_u32_e64 %[pfs], %[pn1], %[pft]\n"                                               \
    "v_lshlrev_b32_e32 %[pft], 2, %[pft]\n"                                                   \
    "global_load_dword %[pfw], %[pft], %[pfa]\n"                                              \
    "s_load_dword %[pfn], %[pfb], 0x0\n"                                                      \
    "s_load_dwordx4 %[pfh], %[pfc], 0x0\n"                                                    \
    "v_and_b32_e32 %[pft], 0xff, %[pw1]\n"                                                    \
    "v_cmp_eq_u32_e32 vcc, 0, %[pft]\n"                                                       \
    "s_and_b64 vcc, vcc, %[pfs]\n"                                                            \
    "v_lshrrev_b32_e32 %[pft], 16, %[pw1]\n"                                                  \
    "s_nop 1\n"                                                                               \
    "v_cndmask_b32_e32 %[pft], 0, %[pft], vcc\n"                                              \
    "v_lshlrev_b32_e32 %[pft], 2, %[pft]\n"                                                   \
    "global_load_dword %[pfk], %[pft], %[cst]\n"

// Window 0 of a tape for the assembly interpreter: lane l's instruction
// word min(l, n - 1) and the constant it names if it is a CONST (the pool's
// first constant otherwise, and for an empty tape).
__device__ __forceinline__ void point_window0(const uint32_t* tape, const float* consts,
                                              uint32_t lane, uint32_t* w0, float* c0,
                                              uint32_t* len = nullptr) {
    typedef const __attribute__((address_space(4))) uint32_t* cword_t;
    const uint32_t n = ((cword_t)tape)[0];
    if (len) *len = n;
    const uint32_t nm1 = n ? n - 1u : 0u;
    const uint32_t w = tape[1 + (lane < nm1 ? lane : nm1)];
    *w0 = w;
    *c0 = consts[(n != 0 && (w & 0xFFu) == 0u) ? (w >> 16) : 0u];
}

// x, y, z: the 8 points of this lane (one register each); returns their
// densities.  tape: [n, instrs]; consts: the shape's pool (index 0 readable
// even when empty).  w0, c0: the first window (lane l: instruction word
// min(l, n - 1)) and its constants (point_window0).  use_t: slot 3 is
// filled with t (0 when the tape writes slot 3 before any instruction reads
// it).  MAXS must be 8, 10 or 16.  Every lane of the wavefront must be
// active.
template <int MAXS>
__device__ __forceinline__ v8f run_point_tape_asm8(const uint32_t* tape, const float* consts,
                                                   float t, v8f x, v8f y, v8f z, uint32_t w0,
                                                   float c0, uint32_t use_t = 1) {
    static_assert(MAXS == 8 || MAXS == 10 || MAXS == 16, "assembly path covers 8, 10 or 16 slots");
    typedef const __attribute__((address_space(4))) uint32_t* cword_t;
    const uint32_t n = ((cword_t)tape)[0];
    const uint64_t twp = (uint64_t)(tape + 1), cp = (uint64_t)consts;
    FREP_ASM_POINT_DECLS;
    if constexpr (MAXS == 16) {
        asm volatile(FREP_ASM_POINT_BODY
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), "+{v160}"(w0), "+{v163}"(c0),
                       FREP_ASM_POINT_SCALARS
                     : [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(4)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_100, FREP_ASM_CLOBBER_101_159,
                       FREP_ASM_CLOBBER_W16, "vcc", "memory");
    } else if constexpr (MAXS == 10) {
        asm volatile(FREP_ASM_POINT_BODY
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), "+{v112}"(w0), "+{v115}"(c0),
                       FREP_ASM_POINT_SCALARS
                     : [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(4)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_100,
                       FREP_ASM_CLOBBER_101_111, FREP_ASM_CLOBBER_W10, "vcc", "memory");
    } else {
        asm volatile(FREP_ASM_POINT_BODY
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), "+{v96}"(w0), "+{v99}"(c0),
                       FREP_ASM_POINT_SCALARS
                     : [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(3)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_95, FREP_ASM_CLOBBER_W8,
                       "vcc", "memory");
    }
    return x;
}

// One voxel per lane (FREP_ASM_POINT_BODY1): x, y, z are the lane's point,
// the rest as run_point_tape_asm8.  MAXS must be 8 or 16.
template <int MAXS>
__device__ __forceinline__ float run_point_tape_asm1(const uint32_t* tape, const float* consts,
                                                     float t, float x, float y, float z,
                                                     uint32_t w0, float c0, uint32_t use_t = 1) {
    static_assert(MAXS == 8 || MAXS == 16, "one-voxel assembly path covers 8 or 16 slots");
    typedef const __attribute__((address_space(4))) uint32_t* cword_t;
    const uint32_t n = ((cword_t)tape)[0];
    const uint64_t twp = (uint64_t)(tape + 1), cp = (uint64_t)consts;
    FREP_ASM_POINT_DECLS;
    if constexpr (MAXS == 16) {
        asm volatile(FREP_ASM_POINT_BODY1
                     : "+{v16}"(x), "+{v17}"(y), "+{v18}"(z), "+{v48}"(w0), "+{v51}"(c0),
                       FREP_ASM_POINT_SCALARS
                     : [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(4)
                     : "v19", "v20", "v21", "v22", "v23", "v24", "v25", "v26", "v27", "v28",
                       "v29", "v30", "v31", "v32", "v33", "v34", "v35", "v36", "v37", "v38",
                       "v39", "v40", "v41", "v42", "v43", "v44", "v45", "v46", "v47",
                       "v49", "v50", "v52", "v53", "v54", "vcc", "memory");
    } else {
        asm volatile(FREP_ASM_POINT_BODY1
                     : "+{v16}"(x), "+{v17}"(y), "+{v18}"(z), "+{v40}"(w0), "+{v43}"(c0),
                       FREP_ASM_POINT_SCALARS
                     : [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(3)
                     : "v19", "v20", "v21", "v22", "v23", "v24", "v25", "v26", "v27", "v28",
                       "v29", "v30", "v31", "v32", "v33", "v34", "v35", "v36", "v37", "v38",
                       "v39",
                       "v41", "v42", "v44", "v45", "v46", "vcc", "memory");
    }
    return x;
}

// The same with the tape's length n given, loading while the tape runs:
// pf_win[lane] into *w1, pf_len[0] into *n1 and pf_hdr[0..3] into h2 (all
// readable addresses), and the constants of the next node (its window nx_w,
// lane l's word of a tape of nx_n words) into *nx_c, as point_window0
// gathers them.
template <int MAXS>
__device__ __forceinline__ v8f run_point_tape_asm8_pf2(const uint32_t* tape, const float* consts,
                                                   float t, v8f x, v8f y, v8f z, uint32_t w0,
                                                   float c0, uint32_t n, const uint32_t* pf_win,
                                                   const uint32_t* pf_len, const uint32_t* pf_hdr,
                                                   uint32_t* w1, uint32_t* n1, uint32_t h2[4],
                                                   uint32_t nx_w, uint32_t nx_n, float* nx_c,
                                                   uint32_t use_t = 1) {
    static_assert(MAXS == 8 || MAXS == 10 || MAXS == 16, "assembly path covers 8, 10 or 16 slots");
    const uint64_t twp = (uint64_t)(tape + 1), cp = (uint64_t)consts;
    FREP_ASM_POINT_DECLS;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    uint32_t pfw, pft, pfn;
    float pfk;
    uint64_t pfs;
    u32x4 pfh;
    const uint64_t pfa = (uint64_t)pf_win, pfb = (uint64_t)pf_len, pfc = (uint64_t)pf_hdr;
    if constexpr (MAXS == 16) {
        asm volatile(FREP_ASM_PF2_ISSUE FREP_ASM_POINT_BODY FREP_ASM_PF_WAIT
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), "+{v160}"(w0), "+{v163}"(c0),
                       FREP_ASM_POINT_SCALARS,
                       [pfw] "=&v"(pfw), [pft] "=&v"(pft), [pfn] "=&s"(pfn), [pfh] "=&s"(pfh),
                       [pfk] "=&v"(pfk), [pfs] "=&s"(pfs)
                     : [pw1] "v"(nx_w), [pn1] "s"(nx_n), [pfa] "s"(pfa), [pfb] "s"(pfb), [pfc] "s"(pfc), [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(4)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_100, FREP_ASM_CLOBBER_101_159,
                       FREP_ASM_CLOBBER_W16, "vcc", "memory");
    } else if constexpr (MAXS == 10) {
        asm volatile(FREP_ASM_PF2_ISSUE FREP_ASM_POINT_BODY FREP_ASM_PF_WAIT
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), "+{v112}"(w0), "+{v115}"(c0),
                       FREP_ASM_POINT_SCALARS,
                       [pfw] "=&v"(pfw), [pft] "=&v"(pft), [pfn] "=&s"(pfn), [pfh] "=&s"(pfh),
                       [pfk] "=&v"(pfk), [pfs] "=&s"(pfs)
                     : [pw1] "v"(nx_w), [pn1] "s"(nx_n), [pfa] "s"(pfa), [pfb] "s"(pfb), [pfc] "s"(pfc), [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(4)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_100,
                       FREP_ASM_CLOBBER_101_111, FREP_ASM_CLOBBER_W10, "vcc", "memory");
    } else {
        asm volatile(FREP_ASM_PF2_ISSUE FREP_ASM_POINT_BODY FREP_ASM_PF_WAIT
                     : "+{v[16:23]}"(x), "+{v[24:31]}"(y), "+{v[32:39]}"(z), "+{v96}"(w0), "+{v99}"(c0),
                       FREP_ASM_POINT_SCALARS,
                       [pfw] "=&v"(pfw), [pft] "=&v"(pft), [pfn] "=&s"(pfn), [pfh] "=&s"(pfh),
                       [pfk] "=&v"(pfk), [pfs] "=&s"(pfs)
                     : [pw1] "v"(nx_w), [pn1] "s"(nx_n), [pfa] "s"(pfa), [pfb] "s"(pfb), [pfc] "s"(pfc), [tw] "s"(twp), [cst] "s"(cp), [n] "s"(n), [t] "v"(t), [ut] "s"(use_t), [maxs] "n"(MAXS),
                       [lg] "n"(3)
                     : FREP_ASM_CLOBBER_40_63, FREP_ASM_CLOBBER_64_95, FREP_ASM_CLOBBER_W8,
                       "vcc", "memory");
    }
    *w1 = pfw;
    *n1 = pfn;
    *nx_c = pfk;
    h2[0] = pfh.x; h2[1] = pfh.y; h2[2] = pfh.z; h2[3] = pfh.w;
    return x;
}

}  // namespace frep
