// frep_host.hpp -- host-side objects of the library: uploaded tapes, trees,
// the context, and the plans of batched voxelizations.  Shared by the C ABI
// (frep_api.hip) and the request server (frep_server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/frep.h"
#include "frep_internal.hpp"

namespace frep {

constexpr uint32_t kRootOff = 16;      // words of slack before the root tape
constexpr uint32_t kPoolPad = 64;      // words of slack after the pool end
constexpr uint64_t kMaxPoolWords = 0xFFFFFF00ull;

// One level-0 schedule of a tape (k_root_interval; RootArgs describes the
// layout).  Offsets are words from p0, the schedule's start in d_sched.
struct RootSched {
    bool ok = false;
    uint32_t lpc = 0;                  // lanes per cell: the largest pass
    uint32_t p0 = 0;
    uint32_t pass = 0, nconst = 0, c0 = 0, m0 = 0, e0 = 0, res = 0, slots = 0;
    uint32_t table_bytes = 0, cell_bytes = 0, cpb = 0;
};
constexpr uint32_t kRootVariants = 2;
constexpr uint32_t kRootLanes[kRootVariants] = {64, 16};

struct TapeObj {
    std::vector<uint32_t> words;       // [n, instrs]
    std::vector<float> consts;
    uint32_t slot_count = 0, maxs = 0, n_mm = 0;
    bool asm_ok = false;               // all ops in the assembly interpreter's set
    bool reads_t = true;               // some instruction reads slot 3 (t) before any writes it
    uint32_t* d_words = nullptr;
    float* d_consts = nullptr;
    uint64_t uid = 0;                  // unique per upload in its context
    // level-0 schedules of k_root_interval with 64 and 16 lanes per cell
    // (d_sched null when neither qualifies)
    uint32_t* d_sched = nullptr;
    RootSched root[kRootVariants];
};

// Layered render target (the reference's RenderTarget image array,
// vulkan/include/engine/render_target.h:13-19): `layers` rgba8 images of
// width x height, layer-major, in device memory.
struct TargetObj {
    uint32_t width = 0, height = 0, layers = 0;
    void* d = nullptr;
};

struct TreeObj {
    frep_grid grid{};
    bool built = false;
    uint32_t* d_nodes[kMaxLevels] = {};
    uint32_t node_cap[kMaxLevels] = {};
    uint32_t ns[kMaxLevels] = {};
    float cs[kMaxLevels] = {};
    uint32_t* d_pool = nullptr;
    uint64_t pool_cap = 0;             // usable words (allocation has kPoolPad more)
    float* d_consts = nullptr;
    uint32_t consts_cap = 0;
    uint32_t n_consts = 0;             // constants of the tape d_consts holds
    uint32_t* d_counters = nullptr;
    uint32_t* h_counters = nullptr;    // pinned mirror
    uint32_t slot_count = 0, maxs = 0;
    bool asm_ok = false;
    bool reads_t = true;               // the tape reads the input t (slot 3)
    uint32_t retries = 0;
    uint64_t consts_uid = 0;           // upload whose constants d_consts holds (0: none)
    // the plain voxelization's launches (init, levels, counter readback) as
    // a graph, replayed while its key (every launch argument) is unchanged
    hipGraphExec_t vexec = nullptr;
    uint64_t vkey = 0;
    uint64_t vkey_seen = 0;            // key of the last direct launch: captured when it repeats
    // Device-state generation: bumped whenever a node or pool buffer is
    // (re)allocated or a batched build writes the tree.  A plain build that
    // completes without overflow leaves the buffers as k_init_tree would for
    // its tape (finish_tree) and records (generation, tape upload); a plain
    // build that finds both unchanged skips the init launch.
    uint64_t dev_gen = 1;
    uint64_t ready_gen = 0;
    uint64_t ready_tape = 0;
};

// Caching sub-allocator (frep_pool.hip): device or pinned host blocks carved
// from large slabs and recycled by size class.  Thread-safe.  A block may be
// freed only after the work using it has completed.
class BlockPool {
  public:
    explicit BlockPool(bool pinned);
    ~BlockPool();
    BlockPool(const BlockPool&) = delete;
    BlockPool& operator=(const BlockPool&) = delete;
    hipError_t alloc(size_t bytes, void** out);
    void free(void* p);
    void release_all();
    size_t reserved_bytes();   // slabs + blocks outside the slabs

  private:
    static constexpr uint32_t kDirectClass = 0xFFFFFFFFu;
    struct Slab {
        char* base;
        size_t size, used;
    };
    struct Live {
        uint32_t cls;
        size_t size;
    };
    bool pinned_;
    std::mutex mu_;
    std::vector<Slab> slabs_;
    std::unordered_map<void*, Live> live_;
    std::vector<std::vector<void*>> free_;
    size_t bytes_ = 0;
};

// Autotuner state of one batch signature.
struct TuneRecord {
    uint32_t trial = 0;                // candidates timed so far
    std::vector<float> best_ms;        // per launch
    std::vector<uint32_t> grid;        // per launch: best persistent grid so far
    std::vector<uint32_t> kc;          // per launch: cells per lane of the best (interval launches)
    hipGraphExec_t exec = nullptr;     // graph of the tuned batch
};

// Argument tables (pinned + device) and gathered counter blocks of one
// batched voxelization in flight.
struct BatchStage {
    void* h_stage = nullptr;
    void* d_stage = nullptr;
    size_t stage_bytes = 0;
    uint32_t* h_cnt = nullptr;
    uint32_t* d_cnt = nullptr;
    size_t cnt_words = 0;
};

// Capacities to start a tree with (node counts per level, tape-pool words),
// e.g. learned from earlier trees of the same grid.
struct CapHint {
    uint32_t nodes[kMaxLevels] = {};
    uint64_t pool_words = 0;
};

// One voxelization of a batch (or the only one of a plain call): capacities
// requested for the next attempt.
struct Job {
    TapeObj* tp = nullptr;
    TreeObj* t = nullptr;
    uint32_t want[kMaxLevels] = {};
    uint64_t pool_want = 0;
    uint32_t mm_stride = 1;
};

// Launch list of a batched voxelization: groups of <= kBatchMax trees, and
// per group one interval and/or one point launch per level.
struct BatchLaunch {
    uint32_t group, level;
    bool point;
    uint32_t count, maxs, mm;
    size_t off;                        // byte offset of its LevelArgs (RootArgs) table
    // root launch (level 0 of the trees whose tape has a schedule)
    bool root = false;
    uint32_t cpb = 0, blocks = 0;      // cells per block (all trees), total blocks
    uint32_t variant = 0;              // schedule variant (lanes per cell kRootLanes[variant])
    uint32_t slots = 0;                // largest slot count of its shapes
    uint32_t asm_slots = 0;            // point launch: every shape on the assembly interpreter with
                                       // at most this many slots (0: not all)
    size_t lds = 0;                    // dynamic LDS per block
    size_t starts_off = 0;             // byte offset of the block starts (count + 1 words)
    bool kc2_ok = false;               // interval launch: two cells per lane possible (for the
                                       // shapes with nodes of a multiple of 128 cells; slot
                                       // files of at most 16)
    uint32_t pkset = 0;                // point launch: the shapes whose last level runs 8 voxels
                                       // per lane (1) or the others (2), each in its own launch
};
struct BatchGroup {
    uint32_t first, count;
    size_t init_off;                   // byte offset of its InitArgs table
    uint32_t lane = 0;                 // stream lane (0: the enqueueing stream)
};
// Streams and events that let the lanes of a plan run side by side: lane
// k > 0 runs on st[k - 1], forked from and joined back into the enqueueing
// stream.
constexpr uint32_t kBatchLanes = 3;
struct LaneSet {
    hipStream_t st[kBatchLanes - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kBatchLanes - 1] = {};
};
struct BatchPlan {
    uint32_t m = 0;
    bool root_ops = false;             // level 0 by k_root_interval where the tape has a schedule
    std::vector<BatchGroup> groups;
    std::vector<BatchLaunch> launches;
    size_t bytes = 0;
};

}  // namespace frep

struct frep_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t cu_count = 0;
    uint32_t level_grid = 0;
    std::string err;
    std::mutex mu;
    std::vector<std::unique_ptr<frep::TapeObj>> tapes;
    std::vector<std::unique_ptr<frep::TreeObj>> trees;
    bool timing = false;
    bool no_asm = false;               // FREP_NO_ASM=1: C++ interpreter only
    bool graphs = false;               // FREP_GRAPH=1: repeated plain builds replayed as a graph
    bool root_ops = true;              // FREP_ROOT_OPS=0: level 0 by the per-cell kernel
    bool root_ops_required = false;    // FREP_ROOT_OPS=2: a tape without a level-0 schedule is an error
    uint32_t point_balance = 4;        // FREP_POINT_BALANCE=n: last-level priority steps every 1/n of a share (0: off)
    bool root_balance = true;          // FREP_ROOT_BALANCE=0: no level-0 issue priority by progress
    uint32_t batch_lanes = frep::kBatchLanes;  // FREP_BATCH_LANES=1..3: lanes of a synchronous batch
    size_t lds_per_block = 65536;      // device limit (hipDeviceAttributeMaxSharedMemoryPerBlock)
    uint64_t next_uid = 1;
    hipEvent_t ev[frep::kMaxLevels + 4] = {};
    frep_timing last{};
    // render output staging
    void* d_out = nullptr;
    size_t out_bytes = 0;
    // device address of the fixed staging array of frep_render_noalloc
    void* noalloc_staging = nullptr;
    // staging of the synchronous batched voxelization
    frep::BatchStage stage;
    // layered render targets
    std::vector<std::unique_ptr<frep::TargetObj>> targets;
    // tree buffers (device) and small pinned mirrors / request outputs
    frep::BlockPool dpool{false};
    frep::BlockPool hpool{true};
    // autotuner records and graphs of batches, keyed by their argument tables
    bool autotune = true;              // FREP_NO_AUTOTUNE=1 turns it off
    std::unordered_map<uint64_t, frep::TuneRecord> tune;
    std::vector<hipEvent_t> bev;       // per-launch events of tuning trials
    frep::LaneSet lanes;               // side streams of the synchronous batched path
};

// Returns from the enclosing function (which has a frep_ctx* ctx) with the
// status of a failed HIP call and its message recorded in the context.
#define HIP_TRY(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::frep::fail(ctx, e_ == hipErrorOutOfMemory ? FREP_OUT_OF_MEMORY      \
                                                               : FREP_DEVICE,          \
                                std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

namespace frep {

uint32_t maxs_for(uint32_t slot_count);
// Level-0 schedule of k_root_interval for tape words w ([n, instrs]) with
// passes of at most lpc instructions: fills out and rs (offsets from the
// start of out), or returns false when the tape reads a slot above 3 before
// writing it, has more than kRootMaxInstr instructions, would average fewer
// than two instructions per pass, or needs more LDS than 8 cells per block
// allow.
bool build_root_schedule(const uint32_t* w, uint32_t lpc, size_t lds_per_block,
                         std::vector<uint32_t>* out, RootSched* rs);
// Arguments of k_root_interval for level 0 of tree t with schedule rs of tp.
RootArgs root_args(const TapeObj* tp, const TreeObj* t, const RootSched& rs, uint32_t mm_stride,
                   uint32_t cells_per_block);
// The schedule of tp for a root of `cells` cells with the lower modelled
// cost (passes x rounds of wavefronts over the GPU's SIMDs), or -1.
int pick_root_schedule(const TapeObj* tp, uint32_t cells, uint32_t cu_count);
frep_status fail(frep_ctx* c, frep_status s, const std::string& msg);
TapeObj* get_tape(frep_ctx* ctx, frep_tape_t h);
TreeObj* get_tree(frep_ctx* ctx, frep_tree_t h);
frep_status check_grid(frep_ctx* ctx, const frep_grid* g);
void free_tree_buffers(frep_ctx* ctx, TreeObj* t);
uint64_t host_alloc(const TreeObj* t, uint32_t L);
uint32_t host_nodes(const TreeObj* t, uint32_t L);
uint64_t host_pool_top(const TreeObj* t);
// Device bytes held by a tree's buffers (records, tape pool, constants).
uint64_t tree_device_bytes(const TreeObj* t);

// New tree object (no level buffers yet); caller holds ctx->mu.
frep_status new_tree(frep_ctx* ctx, frep_tree_t* out);

frep_status prepare_job(frep_ctx* ctx, TapeObj* tp, TreeObj* t, const frep_grid* grid, Job* j,
                        hipStream_t st, const CapHint* hint);
// Device bytes prepare_job + ensure_job would give a fresh tree.
uint64_t estimate_job_bytes(const TapeObj* tp, const frep_grid* grid, const CapHint* hint);
frep_status ensure_job(frep_ctx* ctx, Job* j);
frep_status regrow_job(frep_ctx* ctx, Job* j, uint32_t attempt);
LevelArgs level_args(const TreeObj* t, uint32_t L, uint32_t mm_stride);

// lanes > 1: idx must list the jobs by decreasing tape length; the first
// eighth of them, the next quarter and the rest form separate groups on
// lanes 0, 1, 2 (a job count below 8 keeps one lane).
BatchPlan plan_batch(const std::vector<Job>& jobs, const std::vector<uint32_t>& idx,
                     uint32_t lanes = 1, bool root_ops = false, uint32_t cu_count = 256);
double root_cost(const RootSched& rs, uint32_t cells, uint32_t cu_count);
void fill_batch(const BatchPlan& P, const std::vector<Job>& jobs, const std::vector<uint32_t>& idx,
                uint8_t* hs);
frep_status ensure_stage(frep_ctx* ctx, BatchStage* S, const BatchPlan& P);
void free_stage(BatchStage* S);
// Enqueues a filled plan on st: argument copy, launches (persistent grids
// capped at grid[i]), counter gather and readback into S->h_cnt.
// With lanes, the groups of lane k > 0 run on lanes->st[k - 1] beside those
// of lane 0 (ev must then be null).
frep_status enqueue_batch(frep_ctx* ctx, BatchStage* S, hipStream_t st, const BatchPlan& P,
                          const std::vector<uint32_t>& grid, hipEvent_t* ev,
                          const LaneSet* lanes = nullptr,
                          const std::vector<uint32_t>* kcs = nullptr);

// Kernel arguments of a render of tree t (outputs are device pointers).
RenderArgs render_args(const TreeObj* t, const frep_camera* cam, uint8_t* rgba,
                       unsigned long long* hit_ids, float* normals);
frep_status check_camera(frep_ctx* ctx, const frep_camera* cam);

}  // namespace frep
