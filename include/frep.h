/*
 * frep.h -- C ABI of the MI355X F-rep voxelizer / raytracer (libfrep.so).
 *
 * Drop-in boundary for the reference's hot path (MathisBD/f-rep-voxel-
 * rendering, vulkan/): tape upload -> voxelize -> render.  Plain C types
 * only; the library owns all device memory and hands out opaque handles.
 * Every entry point returns a frep_status and never aborts; the message of
 * the last failure is available from frep_last_error().
 *
 * Reference interfaces each entry point replaces:
 *   frep_tape_upload  VoxelStorage::Init + Voxelizer::UploadTape
 *                     (vulkan/include/engine/voxel_storage.h:35-49,
 *                      vulkan/src/engine/voxelizer.cpp:385-427); the constant
 *                     pool that the reference copies into every dispatch's
 *                     UBO (voxelizer.cpp:222-223) travels with the tape.
 *   frep_voxelize     Voxelizer::Voxelize (vulkan/include/engine/voxelizer.h:21,
 *                     vulkan/src/engine/voxelizer.cpp:429-440); the grid is
 *                     VoxelStorage::gridDims / lowVertex / worldSize
 *                     (vulkan/include/engine/voxel_storage.h:15-17).
 *   frep_render       Raytracer::Trace (vulkan/include/engine/raytracer.h:24,
 *                     vulkan/src/engine/raytracer.cpp:243-272) with the
 *                     camera basis of camera.cpp:16-27 and the UBO fields of
 *                     raytracer.cpp:157-212.
 *   frep_tree_export  canonical dump of VoxelStorage's node/tape buffers
 *                     (vulkan/include/engine/voxel_storage.h:63-88).
 *   frep_eval_*       batch forms of the tape evaluators
 *                     (vulkan/shaders/tape/point.comp:2-63,
 *                      vulkan/shaders/tape/interval.comp:121-160,
 *                      vulkan/shaders/tape/gradient.comp:4-88).
 */
#ifndef FREP_H
#define FREP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum frep_status {
    FREP_OK = 0,
    FREP_INVALID_ARG = 1,   /* null pointer, bad size, unknown handle */
    FREP_BAD_TAPE = 2,      /* malformed tape (op, slot or constant out of range) */
    FREP_OUT_OF_MEMORY = 3, /* device allocation failed or budget exceeded */
    FREP_DEVICE = 4,        /* HIP runtime error */
    FREP_UNSUPPORTED = 5    /* valid input outside what this build supports */
} frep_status;

typedef struct frep_ctx frep_ctx;
typedef uint32_t frep_tape_t;   /* 0 is never a valid handle */
typedef uint32_t frep_tree_t;

#define FREP_TAPE_FORMAT_REF 1u   /* reference words: op|out<<8|A<<16|B<<24, <=256 consts */
#define FREP_TAPE_FORMAT_WIDE 2u  /* same words; CONST index = A | B<<8 (<=65536 consts) */

#define FREP_MAX_LEVELS 8u

/* Grid of the tree (VoxelStorage::gridDims/lowVertex/worldSize + tape time).
 * dims[L]^3 must be a multiple of 64 (the reference asserts the same,
 * voxelizer.cpp:270), i.e. dims[L] in {4, 8, 12, 16, ...}, dims[L] <= 32. */
typedef struct frep_grid {
    uint32_t n_levels;
    uint32_t dims[FREP_MAX_LEVELS];
    float low[3];
    float world_size;
    float tape_time;
} frep_grid;

/* Camera and shading parameters (the reference's raytracer UBO,
 * raytracer/inputs.comp:46-73).  right/up/screen_world are normally filled
 * by frep_camera_look(). */
typedef struct frep_camera {
    float pos[3];
    float forward[3];
    float right[3];
    float up[3];
    float screen_world[2];  /* perspective: 2tan(fov/2), *H/W; ortho: view extent */
    uint32_t width, height;
    uint32_t spp;           /* samples per pixel, >= 1 */
    uint32_t ortho;         /* 0 perspective (reference), 1 orthographic */
    uint32_t jitter;        /* 0: no sub-pixel offset; 1: seeded jitter */
    uint32_t seed;
    float background[4];
    uint32_t light_count;   /* <= 4 */
    float light_dir[4][4];
    float light_color[4][4];
    float ambient[4];
} frep_camera;

typedef struct frep_tree_info {
    uint32_t n_levels;
    uint32_t dims[FREP_MAX_LEVELS];
    uint32_t node_count[FREP_MAX_LEVELS];
    uint32_t tape_len_max[FREP_MAX_LEVELS];
    uint32_t slot_count;
    uint32_t retries;                        /* pool regrowths of the last voxelize */
    uint32_t pad;
    /* sum of node tape lengths per level (the reference's ShaderStats) */
    uint64_t tape_len_sum[FREP_MAX_LEVELS];
    /* work of the level: sum over its nodes of tape length x dim^3 cells,
     * and the same weighted per op as FP32 operations -- above the last
     * level interval rules (ADD/SUB/NEG 2, MUL 10, DIV 12, MIN/MAX 4,
     * SQRT 4, EXP 2, SIN/COS 12), at the last level point rules (1 per op);
     * CONST and COPY count 0 */
    uint64_t op_evals[FREP_MAX_LEVELS];
    uint64_t work_flops[FREP_MAX_LEVELS];
    uint64_t tape_pool_words;                /* words of the tape pool in use */
    uint64_t device_bytes;                   /* device memory held by the tree */
} frep_tree_info;

/* Per-launch GPU durations of the last frep_voxelize / frep_render, measured
 * with HIP events on the library's stream when timing is enabled. */
typedef struct frep_timing {
    uint32_t n_levels;
    float level_ms[FREP_MAX_LEVELS];  /* one kernel per tree level */
    float init_ms;
    float render_ms;
    float batch_ms;   /* frep_voxelize_batch: all launches + counter readback */
} frep_timing;

/* ---- context ---- */
frep_status frep_init(int device, frep_ctx** out);
void        frep_shutdown(frep_ctx* ctx);
const char* frep_last_error(const frep_ctx* ctx);
frep_status frep_set_timing(frep_ctx* ctx, int enable);
frep_status frep_get_timing(frep_ctx* ctx, frep_timing* out);
/* Number of compute units and the persistent grid size used for level kernels. */
frep_status frep_device_info(frep_ctx* ctx, uint32_t* cu_count, uint32_t* level_grid);

/* ---- tape upload ---- */
/* words = [n, instr_0 .. instr_{n-1}], n_words >= n + 1.  Validates every
 * operand (slot < slot_count, const index < n_consts, op < 13) and copies
 * the tape and its constant pool to the device. */
frep_status frep_tape_upload(frep_ctx* ctx, const uint32_t* words, uint32_t n_words,
                             const float* consts, uint32_t n_consts, uint32_t slot_count,
                             uint32_t format, frep_tape_t* out);
frep_status frep_tape_free(frep_ctx* ctx, frep_tape_t tape);

/* ---- voxelize ---- */
frep_status frep_tree_create(frep_ctx* ctx, frep_tree_t* out);
/* Builds the sparse tree of `tape` over `grid` into `tree` (blocking, like
 * Voxelizer::Voxelize).  Buffers of the tree are reused and grown as needed;
 * a repeated build of the same tape into the same buffers starts from the
 * state the previous build left (no init launch).  FREP_DEVICE if the
 * device's copy of the build counters did not reach the host. */
frep_status frep_voxelize(frep_ctx* ctx, frep_tape_t tape, const frep_grid* grid,
                          frep_tree_t tree);
/* Batched form of frep_voxelize (the reference voxelizes one shape per
 * Voxelizer::Voxelize call and has no batched counterpart): voxelizes n
 * (tape, grid, tree) triples -- shapes with their
 * own tape lengths, dims and resolutions -- with one persistent launch per
 * tree level (and kind) per group of up to 64 trees, and one counter
 * readback for the whole batch.  Each tree ends up exactly as frep_voxelize
 * would leave it.  The trees must be distinct. */
frep_status frep_voxelize_batch(frep_ctx* ctx, uint32_t n, const frep_tape_t* tapes,
                                const frep_grid* grids, const frep_tree_t* trees);
frep_status frep_tree_info_get(frep_ctx* ctx, frep_tree_t tree, frep_tree_info* out);
/* Canonical blob of the tree (layout in the INTEGRATION notes); returns the
 * word count in *n_words and writes the blob when buf has room (cap words). */
frep_status frep_tree_export(frep_ctx* ctx, frep_tree_t tree, uint32_t* buf, uint64_t cap,
                             uint64_t* n_words);
frep_status frep_tree_free(frep_ctx* ctx, frep_tree_t tree);
/* Tree files (a resident tree written out and read back, e.g. to warm-start
 * a server or keep fixtures).  The reference keeps trees only in device
 * buffers whose node layout is its interface (vulkan/README.md:22-43,
 * vulkan/include/engine/voxel_storage.h:63-88); it has no file form.
 * Layout (u32 words): 'FRPF', 1, n_levels, dims[8], low[3] (f32), world_size
 * (f32), tape_time (f32), slot_count, n_consts, n_consts f32 constants, then
 * the canonical blob of frep_tree_export.  frep_tree_save writes it (size in
 * *n_words; written when buf has room); frep_tree_load checks every field
 * (grid, tapes, masks, child links, coordinates, canonical order) and
 * creates a new built tree that renders, exports and re-voxelizes like the
 * tree it came from. */
frep_status frep_tree_save(frep_ctx* ctx, frep_tree_t tree, uint32_t* buf, uint64_t cap,
                           uint64_t* n_words);
frep_status frep_tree_load(frep_ctx* ctx, const uint32_t* buf, uint64_t n_words,
                           frep_tree_t* out);

/* ---- render ---- */
/* Fills right/up/screen_world from forward, initial up and a horizontal fov
 * (camera.cpp:16-27, raytracer.cpp:177-181); ortho keeps screen_world. */
frep_status frep_camera_look(frep_camera* cam, const float initial_up[3], float fov_deg);
/* rgba: W*H*4 bytes; hit_ids: W*H*spp u64 (level<<60 | x<<40 | y<<20 | z of
 * the hit cell at its level, UINT64_MAX = miss); normals: W*H*spp*3 floats.
 * Host pointers; any may be NULL. */
frep_status frep_render(frep_ctx* ctx, frep_tree_t tree, const frep_camera* cam,
                        uint8_t* rgba, uint64_t* hit_ids, float* normals);

/* ---- layered render target and its presentation ----
 * The reference renders one jittered sample per frame into layer
 * (frame % temporalSampleCount) of its RenderTarget, an array of rgba8
 * images (vulkan/include/engine/render_target.h:13-19,
 * vulkan/src/engine/raytracer.cpp:159-162), and presents the mean of the
 * layers (vulkan/shaders/renderer/triangle.frag:14-23).  A target is that
 * image array on the device, zeroed at creation (RenderTarget::ZeroOutImage).
 * frep_render_layer renders cam (spp 1, the target's size) into one layer;
 * frep_target_resolve writes the fragment shader's result as rgba8: per
 * pixel the mean over the layers of rgb/255, alpha 1, rounded to nearest
 * (before any swapchain colour conversion, which is outside this library). */
typedef uint32_t frep_target_t;
frep_status frep_target_create(frep_ctx* ctx, uint32_t width, uint32_t height, uint32_t layers,
                               frep_target_t* out);
frep_status frep_target_free(frep_ctx* ctx, frep_target_t target);
frep_status frep_render_layer(frep_ctx* ctx, frep_tree_t tree, const frep_camera* cam,
                              frep_target_t target, uint32_t layer);
frep_status frep_target_resolve(frep_ctx* ctx, frep_target_t target, uint8_t* rgba);

/* Allocation-free form of frep_render.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and outputs as frep_render, but the device staging of
 * the outputs is a fixed array of FREP_NOALLOC_STAGING_BYTES held by the
 * library on each GPU it runs on until the process ends, so the call makes
 * no hipMalloc/hipFree.  Requests whose outputs need more staging than that
 * (rgba 4 B per pixel, hit ids 8 B and normals 12 B per sample, each region
 * rounded up to 256 B, plus 256 B) fail with FREP_INVALID_ARG.  All calls on
 * one GPU share its array: calls that overlap on one GPU (from several
 * threads or contexts) race. */
#define FREP_NOALLOC_STAGING_BYTES (64u << 20)
frep_status frep_render_noalloc(frep_ctx* ctx, frep_tree_t tree, const frep_camera* cam,
                                uint8_t* rgba, uint64_t* hit_ids, float* normals);

/* ---- batch tape evaluation (host pointers) ---- */
frep_status frep_eval_points(frep_ctx* ctx, frep_tape_t tape, const float* xyz, uint32_t n,
                             float t, float* out);
/* out: n * {low, high}; choices (optional): n * n_mm bytes (1 FIRST, 2 SECOND,
 * 3 BOTH) with n_mm = number of MIN/MAX instructions of the tape. */
frep_status frep_eval_intervals(frep_ctx* ctx, frep_tape_t tape, const float* lo,
                                const float* hi, uint32_t n, float t, float* out,
                                uint8_t* choices);
frep_status frep_eval_gradients(frep_ctx* ctx, frep_tape_t tape, const float* xyz, uint32_t n,
                                float t, float* out4);
/* fn: 0 sin, 1 cos, 2 exp -- the device implementations used by the tapes. */
frep_status frep_math(frep_ctx* ctx, uint32_t fn, const float* x, uint32_t n, float* out);

/* ---- tape compiler (host only, no device or context needed) ----
 * Replaces csg::Tape::Tape(csg::Expr) (vulkan/src/csg/tape.cpp:14-21), which
 * VoxelStorage::Init runs on the shape (vulkan/include/engine/voxel_storage.h:
 * 35-49): ConstantFold, NormalForm + MergeDuplicates
 * (vulkan/src/csg/simplifier.cpp:31-259), then topological order, constant
 * pool, liveness and slot allocation (tape.cpp:23-228).  The expression DAG
 * is a node list; node i's inputs a (and b) must be earlier nodes, and node
 * identity is the list index (two equal constants at different indices are
 * two nodes, as two csg::Expr(float) are).  Op codes follow csg::Operator
 * (vulkan/include/csg/expression.h:14-22). */
#define FREP_EXPR_X 0u
#define FREP_EXPR_Y 1u
#define FREP_EXPR_Z 2u
#define FREP_EXPR_T 3u
#define FREP_EXPR_CONST 4u  /* value */
#define FREP_EXPR_SIN 5u    /* a */
#define FREP_EXPR_COS 6u
#define FREP_EXPR_EXP 7u
#define FREP_EXPR_NEG 8u
#define FREP_EXPR_SQRT 9u
#define FREP_EXPR_ADD 10u   /* a, b */
#define FREP_EXPR_SUB 11u
#define FREP_EXPR_MUL 12u
#define FREP_EXPR_DIV 13u
#define FREP_EXPR_MIN 14u
#define FREP_EXPR_MAX 15u

typedef struct frep_expr_node {
    uint32_t op;
    uint32_t a, b;        /* input node indices (unused inputs ignored) */
    float value;          /* CONST only */
} frep_expr_node;

typedef struct frep_compile_info {
    uint32_t n_words;     /* tape words [n, instrs]: n + 1 */
    uint32_t n_consts;    /* constant pool size */
    uint32_t slot_count;  /* slots the tape uses (>= 4) */
    uint32_t format;      /* FREP_TAPE_FORMAT_REF, or _WIDE past 256 constants */
    uint32_t dag_nodes;   /* nodes given */
    uint32_t tape_nodes;  /* nodes of the simplified DAG */
} frep_compile_info;

/* Compiles the DAG rooted at node `root`.  info is always filled; words
 * (n_words) and consts (n_consts) are written when non-NULL, and
 * FREP_INVALID_ARG is returned if their capacities are too small.  Inputs
 * that are not a valid DAG, NaN constants and tapes past 256 slots are
 * FREP_INVALID_ARG with the reason in frep_compile_last_error() (per
 * thread).  The result uploads with frep_tape_upload(words, n_words,
 * consts, n_consts, slot_count, format). */
frep_status frep_compile(const frep_expr_node* nodes, uint32_t n_nodes, uint32_t root,
                         uint32_t* words, uint32_t words_cap, float* consts,
                         uint32_t consts_cap, frep_compile_info* info);
const char* frep_compile_last_error(void);

/* ---- level-0 schedule (host only, no device needed) ----
 * No counterpart in the reference: the reference evaluates the root tape one
 * instruction at a time (vulkan/shaders/tape/interval.comp:121-160).  This
 * returns the schedule frep_tape_upload builds for the pass-parallel root
 * level, for inspection and tests.  words: [n, instrs] (n_words >= n + 1);
 * lanes_per_cell: 64 or 16; lds_per_block: the device's LDS per block in
 * bytes.  info (12 words): scheduled (0/1), passes, constants, constant
 * offset, pass-mask offset, entry offset, result operand, value slots,
 * constant-table bytes, bytes per cell, cells per block, schedule words.
 * The schedule is copied to out when out_cap >= info[11]. */
frep_status frep_root_schedule(const uint32_t* words, uint32_t n_words, uint32_t lanes_per_cell,
                               uint32_t lds_per_block, uint32_t* info, uint32_t* out,
                               uint32_t out_cap);

/* ---- serving: request loop with continuous batching ----
 * The reference has no serving layer (its application voxelizes and renders
 * one scene per frame, vulkan/src/application.cpp:60-93); these entry
 * points have no counterpart there.  A server owns a host thread and its own
 * streams on the context's device.  A scene is a (tape, grid) pair whose
 * tree the server builds on first use and keeps resident; every request
 * renders one scene with its own camera.  Requests that arrive while others
 * run are batched (their trees built by the batched level kernels) and
 * batches overlap on the device.  Scenes are built only while the resident
 * trees fit the device budget; idle trees are evicted least recently used
 * first and rebuilt on demand.  Each request's outputs are what
 * frep_voxelize + frep_render give for its scene (at its tape time) and
 * camera.  A request at a tape time other than the one the scene's resident
 * tree was built at rebuilds the tree once no request in flight reads it;
 * the requests of one scene are admitted in submission order.  The tapes of
 * registered scenes must outlive the server. */
typedef struct frep_server frep_server;
typedef uint32_t frep_scene_t;

typedef struct frep_server_config {
    uint64_t device_budget_bytes;   /* trees + tape pools; 0 = 64 GiB */
    uint32_t max_batch;             /* requests per batch, 1..64; 0 = 64 */
    uint32_t max_inflight;          /* batches on the device at once, 1..4; 0 = 2 */
} frep_server_config;

/* Request flag: render the scene's tree built at the request's tape_time
 * (the reference re-voxelizes every frame at the current time when
 * voxelizeRealTime is set, vulkan/src/application.cpp:68-82).  Without it
 * a request uses the tape time the scene was registered with. */
#define FREP_REQUEST_AT_TIME 1u

typedef struct frep_request {
    frep_scene_t scene;
    uint32_t flags;                 /* FREP_REQUEST_AT_TIME or 0 */
    frep_camera camera;
    float tape_time;                /* used with FREP_REQUEST_AT_TIME */
    uint8_t* rgba;                  /* host outputs as in frep_render (any may be */
    uint64_t* hit_ids;              /* NULL); written before the request completes */
    float* normals;
} frep_request;

typedef struct frep_result {
    frep_status status;
    uint32_t voxelized;             /* 1: the request's batch (re)built the scene's tree */
    uint32_t batch_size;            /* requests in its batch */
    float latency_ms;               /* submit -> outputs written */
} frep_result;

typedef struct frep_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t batches, voxelizations, evictions;
    uint64_t resident_bytes, budget_bytes;
    uint32_t resident_scenes, queued;
    float p50_ms, p99_ms, mean_ms, max_ms;   /* latency of completed requests */
    float mean_batch;                        /* requests per batch */
    float elapsed_s;                         /* first submit -> last completion */
} frep_server_stats;

frep_status frep_server_create(frep_ctx* ctx, const frep_server_config* cfg, frep_server** out);
frep_status frep_server_destroy(frep_server* srv);
const char* frep_server_last_error(const frep_server* srv);
frep_status frep_server_scene(frep_server* srv, frep_tape_t tape, const frep_grid* grid,
                              frep_scene_t* out);
frep_status frep_server_submit(frep_server* srv, const frep_request* req, uint64_t* ticket);
/* Blocks until the request completes; the ticket is released. */
frep_status frep_server_wait(frep_server* srv, uint64_t ticket, frep_result* out);
/* *done = 1 (and the ticket released, *out filled) if the request completed. */
frep_status frep_server_poll(frep_server* srv, uint64_t ticket, frep_result* out, int* done);
frep_status frep_server_stats_get(frep_server* srv, frep_server_stats* out);
frep_status frep_server_reset_stats(frep_server* srv);

#ifdef __cplusplus
}
#endif
#endif
