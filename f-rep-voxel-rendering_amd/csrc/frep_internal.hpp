// frep_internal.hpp -- structures shared by the kernels and the host driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frep {

constexpr uint32_t kMaxLevels = 8;
constexpr uint32_t kNoChild = 0xFFFFFFFFu;

// Tape op codes (vulkan/include/csg/tape.h:15-30) and min/max choices
// (vulkan/shaders/tape/tape.comp:10-13).
enum : uint32_t {
    OP_CONST = 0, OP_SIN = 1, OP_COS = 2, OP_ADD = 3, OP_SUB = 4, OP_MUL = 5,
    OP_DIV = 6, OP_MIN = 7, OP_MAX = 8, OP_EXP = 9, OP_NEG = 10, OP_SQRT = 11,
    OP_COPY = 12
};
enum : uint32_t { MM_NONE = 0, MM_FIRST = 1, MM_SECOND = 2, MM_BOTH = 3 };

// Device counter block of one tree (u32 words).
//   [2L, 2L+1]  u64 allocation counter of level L (L = 1..7):
//               (nodes << kAllocShift) | tape-pool words used by level L.
//               One packed atomic per wavefront reserves both child ids and
//               tape space.  Level 0 always has exactly one node.
//   [16]        pool offset where level-1 tapes start
//   [17]        overflow flags, two bits per level L (bits 2L, 2L+1): the
//               launch of level L ran out of node capacity (OVF_NODES) or
//               tape pool (OVF_POOL) for level L+1
//   [18]        blocks of the last level's launch that have finished
//               (finish_tree)
// Level L's tapes occupy [base_L, base_L + words_L) with base_1 = [16] and
// base_{L+1} = base_L + words_L.
enum : uint32_t { CNT_ALLOC = 0, CNT_BASE1 = 16, CNT_OVERFLOW = 17, CNT_DONE = 18, CNT_WORDS = 32 };
enum : uint32_t { OVF_NODES = 1u, OVF_POOL = 2u };
// Overflow bits of the levels before `level`: a launch of `level` has nothing
// to do for a tree whose earlier level overflowed (the host rebuilds it).
__host__ __device__ inline uint32_t ovf_before(uint32_t level) { return (1u << (2 * level)) - 1u; }
constexpr uint32_t kAllocShift = 38;
constexpr uint64_t kAllocWordMask = (1ull << kAllocShift) - 1;
constexpr uint32_t kMaxLevelNodes = (1u << (64 - kAllocShift)) - 1;   // 2^26 - 1

__host__ __device__ inline uint32_t alloc_nodes(uint64_t v) { return (uint32_t)(v >> kAllocShift); }
__host__ __device__ inline uint64_t alloc_words(uint64_t v) { return v & kAllocWordMask; }

// Node record (u32 words), per level, in the reference's order
// (vulkan/shaders/node.comp:4-15, vulkan/include/engine/voxel_storage.h:63-88):
//   [0] tape offset in the tree's tape pool (words)
//   [1..3] node coords (x, y, z) in units of this level's nodes
//   [4 .. 4+W)            leaf mask, W = dim^3/32
//   [4+W .. 4+2W)         interior mask       (non-last levels)
//   [4+2W .. 4+2W+dim^3)  child ids, 0xFFFFFFFF = none  (non-last levels)
__host__ __device__ inline uint32_t node_words(uint32_t dim, bool last) {
    const uint32_t cells = dim * dim * dim;
    return last ? 4u + cells / 32u : 4u + cells / 16u + cells;
}

struct LevelArgs {
    uint32_t* nodes;          // this level's records
    uint32_t* child_nodes;    // next level's records (non-last levels)
    uint32_t* pool;           // tape pool
    const float* consts;      // constant pool of the shape
    uint32_t* counters;       // CNT_* block
    uint64_t pool_cap;        // words
    uint32_t level;
    uint32_t dim;
    uint32_t dim_shift;       // log2(dim) when dim is a power of two, else 0
    uint32_t node_cap;        // capacity of this level
    uint32_t child_cap;       // capacity of the next level
    uint32_t ns;              // record words, this level
    uint32_t child_ns;        // record words, next level
    uint32_t mm_stride;       // LDS choice rows per wavefront (>= MIN/MAX count)
    float cs;                 // cell size at this level
    float low[3];
    float tape_time;
    uint32_t pk;              // point level: voxels per lane (8, 4 or 1)
    uint32_t use_asm;         // point level: assembly interpreter (pk == 8, slot file <= 16)
    uint32_t finish;          // last level of a plain build: finish_tree (counters copied
                              // to host_counters, then reset for the next build)
    uint32_t* host_counters;  // pinned host mirror of the counters (finish)
    uint32_t reads_t;         // the tape reads the input t: the assembly interpreter fills
                              // slot 3 with tape_time (otherwise the tape writes it first)
    uint32_t balance;         // point level, plain kernel: a wavefront lowers its issue
                              // priority (3 -> 0) as it works through its share
    uint32_t log_cap;         // interval level: shortening-log entries per wavefront (set at
                              // launch from the LDS the launch gives it)
};
// Word CNT_DONE of a host mirror written by finish_tree.
constexpr uint32_t kCountersWritten = 0xC0DE0001u;

// Root level evaluated in passes over the tape's instructions
// (k_root_interval).  sched holds, in u32 words:
//   [c0 .. c0 + n_const)   constants used: instruction i | pool index << 16
//   [m0 .. m0 + n_pass)    op types of each pass (bit op set when present)
//   [e0 ..)                n_pass passes (a multiple of 4) of lanes_per_cell
//                          entries, two words each: op | out slot << 4 | i << 16,
//                          operand A | operand B << 16; op kRootNop pads a
//                          short pass
// An operand is a value slot (0..3 hold x, y, z, t), or kRootConst | k for
// the k-th constant.  Slots are reused once a value's last reader has run.
struct RootArgs {
    LevelArgs a;              // level 0 of the tree
    const uint32_t* sched;
    const uint32_t* tape;     // [n, instrs] of the root tape
    uint32_t n;               // instructions
    uint32_t n_pass;
    uint32_t n_const;
    uint32_t c0;
    uint32_t m0;
    uint32_t e0;              // even (entries are read as 8-byte pairs)
    uint32_t n_slot;
    uint32_t res;             // operand holding the result (slot 0 at the end of the tape)
    uint32_t table_bytes;     // LDS of the constant table: 8 n_const (values) + 2 n_const (instrs), 16-aligned
    uint32_t cell_bytes;      // LDS per cell: 8 n_slot (values) + n rounded to 4 (flags) + n_slot (live), 16-aligned
    uint32_t cells_per_block; // 16 or 8 (root_cells_per_block)
    uint32_t lanes_per_cell;  // 64 or 16: the schedule's largest pass
};
constexpr uint32_t kRootConst = 0x8000u;
constexpr uint32_t kRootNop = 0xFu;
constexpr uint32_t kRootMaxInstr = 65535;      // 16-bit instruction and constant ids
constexpr uint32_t kRootMaxSlots = 4095;       // 12-bit slot ids
constexpr size_t kRootStaticLds = 256;          // bound on the kernel's static LDS

// One tree of a batch: what the init kernel needs.
struct InitArgs {
    uint32_t* counters;
    uint32_t* nodes0;
    uint32_t* pool;
    const uint32_t* root_tape;
    uint32_t root_off;
    uint32_t pad;
};

// A batched level launch: one persistent grid over the chunks of every
// shape in the list (at most kBatchMax shapes, one per lane of the plan).
struct BatchArgs {
    const LevelArgs* shapes;  // device array
    uint32_t n_shapes;
    uint32_t mm_stride;       // LDS choice rows per wavefront (max over shapes)
    uint32_t log_cap;         // shortening-log entries per wavefront (set at launch)
    uint32_t kc;              // interval launch: cells per lane, 1 or 2 (2: slot files of at
                              // most 16; shapes whose nodes are not a multiple of 128 cells
                              // still take one)
};
constexpr uint32_t kBatchMax = 64;

struct RenderLevel {
    const uint32_t* nodes;
    uint32_t ns;
    uint32_t dim;
    float cs;
    uint32_t cap;             // records the level's buffer holds
};

struct RenderArgs {
    RenderLevel lv[kMaxLevels];
    uint32_t n_levels;
    const uint32_t* pool;
    uint64_t pool_words;      // readable words of the tape pool
    const float* consts;
    const uint32_t* counters; // the tree's counter block (overflow flags)
    float low[3];
    float world_size;
    float tape_time;
    // camera
    float pos[3], fwd[3], right[3], up[3];
    float sw[2];
    uint32_t width, height, spp, ortho, jitter, seed;
    float background[4];
    uint32_t light_count;
    float light_dir[4][4];
    float light_color[4][4];
    float ambient[4];
    // outputs (any may be null)
    uint8_t* rgba;
    unsigned long long* hit_ids;
    float* normals;
};

// Launchers (frep_kernels.hip).  Each returns hipSuccess or the launch error.
hipError_t launch_init_tree(hipStream_t st, uint32_t* counters, uint32_t* nodes0, uint32_t* pool,
                            const uint32_t* root_tape, uint32_t root_off, uint32_t n_levels);
// Level kernels run a persistent grid of min(max_grid, cu_count x resident
// blocks per CU) blocks; max_grid bounds the per-wave scratch rows.
// slot_count: the tape's slot count (at most maxs), which picks the slot file.
// nodes_hint: the level's node count in an earlier build (0 when unknown);
// a level of few nodes keeps one cell per lane, for more wavefronts.
hipError_t launch_level_interval(hipStream_t st, const LevelArgs& a, uint32_t maxs,
                                 uint32_t slot_count, uint32_t cu_count, uint32_t max_grid,
                                 uint32_t nodes_hint = 0);
// asm_ok: every op of the tape is in the assembly interpreter's set.
// slots: the tape's slot count (at most maxs), which picks the slot file.
hipError_t launch_level_point(hipStream_t st, const LevelArgs& a, uint32_t maxs,
                              uint32_t slots, uint32_t cu_count, uint32_t max_grid, bool asm_ok);
// Cells per block of k_root_interval for its LDS sizes and the device's
// LDS per block: 16 or 8, or 0 when not even 8 cells fit.
uint32_t root_cells_per_block(uint32_t table_bytes, uint32_t cell_bytes, size_t lds_per_block);
hipError_t launch_root_interval(hipStream_t st, const RootArgs& r);
// Roots of n <= kBatchMax trees: d_args[j] (cells_per_block and
// lanes_per_cell the same for all), tree j owning blocks d_starts[j] ..
// d_starts[j+1]-1 of `blocks`; lds = the largest table_bytes +
// cells_per_block * cell_bytes.
hipError_t launch_root_interval_batch(hipStream_t st, const RootArgs* d_args,
                                      const uint32_t* d_starts, uint32_t n, uint32_t blocks,
                                      uint32_t cells_per_block, uint32_t lanes_per_cell,
                                      size_t lds);
// Batched forms: init of n trees, one level over many shapes, and a copy of
// the n counter blocks into out (n * CNT_WORDS words) for one readback.
// max_grid caps the persistent grid (the autotuner varies it).
hipError_t launch_init_batch(hipStream_t st, const InitArgs* d_args, uint32_t n);
hipError_t launch_level_interval_batch(hipStream_t st, const BatchArgs& b, uint32_t maxs,
                                       uint32_t slot_count, uint32_t cu_count, uint32_t max_grid);
// asm_slots > 0: every shape takes the assembly interpreter (8 voxels per
// lane) and has at most asm_slots (<= 16) slots.
hipError_t launch_level_point_batch(hipStream_t st, const BatchArgs& b, uint32_t maxs,
                                    uint32_t cu_count, uint32_t max_grid, uint32_t asm_slots,
                                    bool small);
hipError_t launch_gather_counters(hipStream_t st, const InitArgs* d_args, uint32_t n,
                                  uint32_t* out);
uint32_t point_k_used(uint32_t maxs, uint32_t dim);
// Adds {sum of tape lengths, max length, weighted op count} of one level's
// nodes to out[0..2] (interval weights if `interval`, else point weights).
hipError_t launch_tree_stats(hipStream_t st, const uint32_t* nodes, uint32_t ns, uint32_t n_nodes,
                             const uint32_t* pool, uint64_t pool_words, bool interval,
                             unsigned long long* out);
hipError_t launch_render(hipStream_t st, const RenderArgs& a, uint32_t maxs, uint32_t slots);
// Mean of n_layers rgba8 layers of `pixels` pixels each into out (rgba8).
hipError_t launch_resolve(hipStream_t st, const void* layers, uint32_t pixels, uint32_t n_layers,
                          void* out);
hipError_t launch_eval_points(hipStream_t st, const uint32_t* tape, const float* consts,
                              const float* xyz, uint32_t n, float t, float* out, uint32_t maxs,
                              bool asm_ok);
hipError_t launch_eval_intervals(hipStream_t st, const uint32_t* tape, const float* consts,
                                 const float* lo, const float* hi, uint32_t n, float t,
                                 float* out, uint8_t* choices, uint32_t n_mm, uint32_t maxs);
hipError_t launch_eval_gradients(hipStream_t st, const uint32_t* tape, const float* consts,
                                 const float* xyz, uint32_t n, float t, float* out, uint32_t maxs);
hipError_t launch_math(hipStream_t st, uint32_t fn, const float* x, uint32_t n, float* out);

// Threads per block of the persistent level kernels (4 wavefronts).
constexpr uint32_t kLevelBlock = 256;

// Min/max choice rows of the interval kernel live in LDS: 16 bytes per
// MIN/MAX instruction per wavefront, 4 wavefronts per block.  A block may
// use up to the device's LDS per block (160 KiB on gfx950: 2544 MIN/MAX
// instructions); rows past 64 KiB per block cost occupancy.
// 1 KiB of the block's LDS is left for the kernels' static LDS (the
// per-block reservation words).
constexpr size_t kMaxChoiceLds = 159 * 1024;
inline uint32_t max_minmax(size_t lds_per_block) {
    const size_t b = lds_per_block - 1024 < kMaxChoiceLds ? lds_per_block - 1024 : kMaxChoiceLds;
    return (uint32_t)(b / 16 / (kLevelBlock / 64));
}

}  // namespace frep
