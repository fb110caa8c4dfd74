/*
 * frep_oracle.h -- CPU oracle for the F-rep voxelize / render path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product library
 * (libfrep.so) never links, loads or calls it.
 *
 * It restates, in plain C++17 compiled with -ffp-contract=off, the algorithm
 * of the reference (MathisBD/f-rep-voxel-rendering, vulkan/):
 *   - tape format and decode        vulkan/include/csg/tape.h:15-42,
 *                                   vulkan/shaders/tape/tape.comp:40-56
 *   - point evaluation              vulkan/src/csg/tape.cpp:274-302,
 *                                   vulkan/shaders/tape/point.comp:2-63
 *   - interval evaluation + choices vulkan/shaders/tape/interval.comp:1-160
 *   - tape shortening               vulkan/shaders/tape/shorten.comp:20-51,148-223
 *   - gradient (forward AD)         vulkan/shaders/tape/gradient.comp:4-88
 *   - level-by-level tree build     vulkan/shaders/voxelizer/main.comp:19-170,
 *                                   vulkan/src/engine/voxelizer.cpp:200-273,429-440
 *   - DDA raytrace + shading        vulkan/shaders/raytracer/main.comp:24-184
 *   - plain dense CPU evaluator     vulkan/src/engine/scene_builder.cpp:91-156
 *
 * Every function is exported with C linkage so tests can call it via ctypes.
 * Parity status of each function is listed in the header comment of
 * frep_oracle.cpp.
 */
#ifndef FREP_ORACLE_H
#define FREP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- transcendentals shared by point / interval / gradient rules ---- */
float orc_sinf(float x);
float orc_cosf(float x);
float orc_expf(float x);

/* ---- tape primitives.  A tape is [n, instr_0 .. instr_{n-1}] (u32). ---- */

/* Returns 0 if the tape is well formed for the given pool / slot counts,
 * otherwise a positive code naming the first problem. */
int orc_tape_validate(const uint32_t* tape, uint32_t n_words,
                      uint32_t n_consts, uint32_t slot_count);

float orc_eval_point(const uint32_t* tape, const float* consts,
                     float x, float y, float z, float t);

/* Interval evaluation over the box [lo, hi].  out[0..1] = {low, high} of
 * slot 0.  choices (optional) receives one byte per MIN/MAX instruction in
 * tape order: 1 = FIRST, 2 = SECOND, 3 = BOTH.  Returns the number of
 * MIN/MAX instructions. */
uint32_t orc_eval_interval(const uint32_t* tape, const float* consts,
                           const float* lo, const float* hi, float t,
                           float* out, uint8_t* choices);

/* Tape shortening.  Writes [k, instrs] into out (capacity n+1 words).
 * Returns k.  *reused = 1 when nothing could be removed (the reference then
 * keeps the parent tape unchanged and out is a verbatim copy of it). */
uint32_t orc_shorten(const uint32_t* tape, const uint8_t* choices,
                     uint32_t* out, int* reused);

/* Forward-mode gradient.  out = {d/dx, d/dy, d/dz, value}. */
void orc_eval_gradient(const uint32_t* tape, const float* consts,
                       float x, float y, float z, float t, float* out);

/* ---- level-by-level tree build (restatement of the GPU algorithm) ---- */

typedef struct orc_grid {
    uint32_t n_levels;       /* 1..8 */
    uint32_t dims[8];        /* branching per level, dims[L]^3 % 32 == 0 */
    float    low[3];         /* world position of the grid's low corner */
    float    world_size;     /* edge length of the (cubic) grid */
    float    tape_time;      /* value of slot 3 (t) */
} orc_grid;

typedef struct orc_tree orc_tree;

orc_tree* orc_voxelize(const uint32_t* tape, const float* consts,
                       const orc_grid* grid);
void      orc_tree_free(orc_tree* tree);
uint32_t  orc_tree_node_count(const orc_tree* tree, uint32_t level);
/* Canonical blob (layout documented in frep_oracle.cpp).  Returns the
 * number of u32 words; writes them when buf != NULL and cap is enough. */
uint64_t  orc_tree_canonical(const orc_tree* tree, uint32_t* buf, uint64_t cap);
/* Flattened occupancy of the finest grid: bit (x*F + y)*F + z, F = fine dim. */
void      orc_tree_flatten(const orc_tree* tree, uint8_t* bits);
/* Per-level sums: [L*4+0] nodes, [L*4+1] sum of tape lengths,
 * [L*4+2] point-flops per cell summed over nodes, [L*4+3] interval-flops. */
void      orc_tree_stats(const orc_tree* tree, double* out);

/* ---- plain dense CPU evaluator (the reference's SceneBuilder path) ---- */

/* Occupancy of fine voxels with x in [x_begin, x_end): one byte per voxel,
 * index ((x - x_begin)*F + y)*F + z.  faithful != 0 keeps the reference's
 * per-call slot-vector allocation.  threads <= 1 runs single threaded. */
void orc_dense_occupancy(const uint32_t* tape, const float* consts,
                         uint32_t slot_count, const orc_grid* grid,
                         uint32_t x_begin, uint32_t x_end, int faithful,
                         int threads, uint8_t* out);

/* ---- DDA render ---- */

typedef struct orc_camera {
    float pos[3];
    float forward[3];
    float right[3];
    float up[3];
    float screen_world[2];   /* perspective: 2tan(fov/2) and *H/W;
                                orthographic: view extent in world units */
    uint32_t width, height;
    uint32_t spp;
    uint32_t ortho;          /* 0 perspective, 1 orthographic */
    uint32_t jitter;         /* 0: no sub-pixel offset, 1: seeded jitter */
    uint32_t seed;
    float background[4];
    uint32_t light_count;    /* <= 4 */
    float light_dir[4][4];
    float light_color[4][4];
    float ambient[4];
} orc_camera;

/* hit_ids: W*H*spp u64 (UINT64_MAX = miss), normals: W*H*spp*3 floats
 * (NaN for misses), rgba: W*H*4 bytes.  Any output may be NULL. */
void orc_render(const orc_tree* tree, const orc_camera* cam,
                uint64_t* hit_ids, float* normals, uint8_t* rgba);

/* orc_render over rows row_begin, row_begin + row_step, ... only (outputs
 * are full-frame buffers; other rows are left untouched). */
void orc_render_rows(const orc_tree* tree, const orc_camera* cam,
                     uint32_t row_begin, uint32_t row_step,
                     uint64_t* hit_ids, float* normals, uint8_t* rgba);

/* ---- trees too large to build whole on the CPU ---- */

/* Checks a canonical tree blob against the oracle without building the
 * whole tree.  Levels 0..n_full-1 are evaluated in full (threads workers):
 * node counts, and every record's coords, tape, leaf and interior masks and
 * child ranks compared.  Below them the check descends by sampling: at level
 * n_full every sample_every-th record (canonical order) is evaluated, and at
 * each deeper level every sample_every-th child of the nodes evaluated one
 * level up.  For every node the oracle reaches, the blob must hold a record
 * with its coords and tape (found by coords); for every node evaluated, its
 * leaf and interior masks must match and each child entry must name a record
 * with the child's coords.  At the last level, every dense_every-th evaluated
 * node's leaf bits are also checked against point evaluation of the FULL
 * tape at each voxel's low corner (dense sampling; independent of
 * shortening).  report (24 u64): [0] records evaluated at full levels,
 * [1] records evaluated by sampling, [2] voxels dense-checked, [3] dense
 * mismatches, [4] level and [5] record of the first mismatch, [6] its kind
 * (1 malformed, 2 node count, 3 coords, 4 tape, 5 leaf, 6 interior, 7 child
 * entries), [7] mismatching records, [8 + L] the oracle's node count of
 * level L for L <= n_full (0 where unknown).  Returns the number of
 * mismatches (records + dense voxels), or -1 for a malformed blob. */
int64_t orc_check_blob(const uint32_t* tape, const float* consts, const orc_grid* grid,
                       const uint32_t* blob, uint64_t n_words, uint32_t n_full,
                       uint32_t sample_every, uint32_t dense_every, uint32_t threads,
                       uint64_t* report);

/* A tree read from a canonical blob (e.g. one the GPU library exported), so
 * that the oracle's raytracer can render it.  NULL if malformed. */
orc_tree* orc_tree_from_blob(const uint32_t* blob, uint64_t n_words,
                             const float* consts, uint32_t n_consts,
                             const orc_grid* grid);

/* Camera basis as the reference computes it (camera.cpp:16-27,
 * raytracer.cpp:177-181): fills right/up/screen_world from forward,
 * initial up and a horizontal fov in degrees. */
void orc_camera_basis(orc_camera* cam, const float* initial_up, float fov_deg);

#ifdef __cplusplus
}
#endif
#endif
