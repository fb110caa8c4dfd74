/*
 * frep_engine.hpp -- C++ host classes over the C ABI of frep.h, with the
 * names, fields and call order of the reference's engine
 * (MathisBD/f-rep-voxel-rendering):
 *
 *   VoxelStorage  vulkan/include/engine/voxel_storage.h:9-88
 *   Voxelizer     vulkan/include/engine/voxelizer.h:12-24
 *   Raytracer     vulkan/include/engine/raytracer.h:14-28
 *   Camera        vulkan/include/engine/camera.h:6-28, vulkan/src/engine/camera.cpp:16-27
 *   RenderTarget  vulkan/include/engine/render_target.h:9-30 (+ the presentation
 *                 pass's layer mean, vulkan/shaders/renderer/triangle.frag:14-23)
 *
 * so that the reference's driver (vulkan/src/application.cpp:60-93)
 *
 *     voxelizer.Voxelize(waitSem, time);  raytracer.Trace(waitSem, &camera, time);
 *
 * becomes  voxelizer.Voxelize(time);  raytracer.Trace(&camera, time);
 *
 * Differences, all from the boundary: a Device wraps the frep context
 * instead of a Vulkan device; Voxelize/Trace block until done, so the
 * semaphore arguments go away; VoxelStorage::Init takes an already compiled
 * tape (words, constant pool, slot count -- csg::Tape's contents) instead
 * of a csg::Expr; Trace writes an rgba8 image (and optionally hit ids and
 * normals) that Image() returns instead of a render-target layer.  Where
 * the reference asserts or exits on an error, these classes throw
 * frep_engine::Error carrying the frep_status and message.
 *
 * Header-only; link with libfrep.so.
 */
#ifndef FREP_ENGINE_HPP
#define FREP_ENGINE_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "frep.h"

namespace frep_engine {

class Error : public std::runtime_error {
  public:
    Error(frep_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    frep_status status;
};

inline void check(frep_ctx* ctx, frep_status st, const char* what) {
    if (st != FREP_OK)
        throw Error(st, std::string(what) + ": " + (ctx ? frep_last_error(ctx) : "no context"));
}

using vec3 = std::array<float, 3>;

// Stands in for vkw::Device: one frep context on one GPU.
class Device {
  public:
    frep_ctx* ctx = nullptr;
    void Init(int ordinal = 0) { check(nullptr, frep_init(ordinal, &ctx), "frep_init"); }
    void Cleanup() {
        frep_shutdown(ctx);
        ctx = nullptr;
    }
};

// A compiled tape: csg::Tape's instruction words ([n, instr_0 ..]), constant
// pool and slot count (vulkan/include/csg/tape.h:15-42).
struct Tape {
    std::vector<uint32_t> words;
    std::vector<float> constantPool;
    uint32_t slotCount = 0;
};

class VoxelStorage {
  public:
    // Filled before Init() (voxel_storage.h:12-17).
    std::vector<uint32_t> gridDims;
    vec3 lowVertex{};
    float worldSize = 0.0f;

    uint32_t gridLevels = 0;
    uint32_t fineGridDim = 0;
    // Node count of each level after a voxelization (the reference's
    // interiorNodeCount, filled by Voxelizer::Voxelize).
    std::vector<uint32_t> interiorNodeCount;
    Tape tape;

    // Device objects (the reference's nodeBuffer / tapeBuffer).
    frep_tape_t tapeHandle = 0;
    frep_tree_t treeHandle = 0;

    void Init(Device* device, const Tape& t) {
        m_device = device;
        gridLevels = (uint32_t)gridDims.size();
        fineGridDim = 1;
        for (uint32_t d : gridDims) fineGridDim *= d;
        interiorNodeCount.assign(gridLevels, 0);
        tape = t;
        const uint32_t fmt = tape.constantPool.size() <= 256 ? FREP_TAPE_FORMAT_REF
                                                              : FREP_TAPE_FORMAT_WIDE;
        check(ctx(), frep_tape_upload(ctx(), tape.words.data(), (uint32_t)tape.words.size(),
                                      tape.constantPool.data(), (uint32_t)tape.constantPool.size(),
                                      tape.slotCount, fmt, &tapeHandle),
              "frep_tape_upload");
        check(ctx(), frep_tree_create(ctx(), &treeHandle), "frep_tree_create");
    }

    void Cleanup() {
        if (treeHandle) frep_tree_free(ctx(), treeHandle);
        if (tapeHandle) frep_tape_free(ctx(), tapeHandle);
        treeHandle = tapeHandle = 0;
    }

    // The coordinates are given in the finest grid dimensions
    // (voxel_storage.h:58-61).
    vec3 WorldPosition(const std::array<uint32_t, 3>& pos) const {
        vec3 w;
        for (int c = 0; c < 3; ++c)
            w[c] = lowVertex[c] + ((float)pos[c] * worldSize) / (float)fineGridDim;
        return w;
    }

    // Record size in bytes (voxel_storage.h:63-70).
    uint32_t NodeSize(uint32_t level) const {
        const uint32_t cells = gridDims[level] * gridDims[level] * gridDims[level];
        return level == gridLevels - 1 ? 16 + cells / 4 : 16 + 17 * (cells / 4);
    }

    frep_grid Grid(float tapeTime) const {
        frep_grid g{};
        g.n_levels = gridLevels;
        for (uint32_t L = 0; L < gridLevels && L < FREP_MAX_LEVELS; ++L) g.dims[L] = gridDims[L];
        for (int c = 0; c < 3; ++c) g.low[c] = lowVertex[c];
        g.world_size = worldSize;
        g.tape_time = tapeTime;
        return g;
    }

    frep_ctx* ctx() const { return m_device ? m_device->ctx : nullptr; }

  private:
    Device* m_device = nullptr;
};

class Voxelizer {
  public:
    void Init(Device* device, VoxelStorage* voxels) {
        m_device = device;
        m_voxels = voxels;
    }
    void Cleanup() { m_voxels = nullptr; }

    // Builds the tree of the storage's tape at time tapeTime (blocking;
    // voxelizer.h:21, voxelizer.cpp:429-440).
    void Voxelize(float tapeTime) {
        frep_ctx* c = m_device->ctx;
        const frep_grid g = m_voxels->Grid(tapeTime);
        check(c, frep_voxelize(c, m_voxels->tapeHandle, &g, m_voxels->treeHandle), "frep_voxelize");
        frep_tree_info info{};
        check(c, frep_tree_info_get(c, m_voxels->treeHandle, &info), "frep_tree_info_get");
        for (uint32_t L = 0; L < m_voxels->gridLevels; ++L)
            m_voxels->interiorNodeCount[L] = info.node_count[L];
        m_last = info;
    }

    // Tape length statistics per level (the reference's ShaderStats).
    void PrintStats() const {
        for (uint32_t L = 0; L < m_last.n_levels; ++L)
            std::printf("level %u: %u nodes, tape length avg %.1f max %u\n", L,
                        m_last.node_count[L],
                        m_last.node_count[L] ? (double)m_last.tape_len_sum[L] / m_last.node_count[L]
                                             : 0.0,
                        m_last.tape_len_max[L]);
    }

  private:
    Device* m_device = nullptr;
    VoxelStorage* m_voxels = nullptr;
    frep_tree_info m_last{};
};

class Camera {
  public:
    vec3 position{};
    vec3 forward{0.0f, 0.0f, -1.0f};
    vec3 initialUp{0.0f, 1.0f, 0.0f};
    float fovDeg = 45.0f;       // horizontal field of view in degrees
    float aspectRatio = 1.0f;   // screen width / screen height

    void Init() {}

    // Orthonormal basis (camera.cpp:16-27): right = normalize(cross(forward,
    // initialUp)), up = cross(right, forward).
    vec3 Right() const {
        vec3 r{forward[1] * initialUp[2] - forward[2] * initialUp[1],
               forward[2] * initialUp[0] - forward[0] * initialUp[2],
               forward[0] * initialUp[1] - forward[1] * initialUp[0]};
        const float l = std::sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
        for (float& v : r) v = v / l;
        return r;
    }
    vec3 Up() const {
        const vec3 r = Right();
        return {r[1] * forward[2] - r[2] * forward[1], r[2] * forward[0] - r[0] * forward[2],
                r[0] * forward[1] - r[1] * forward[0]};
    }
};

// Layered render target (render_target.h:9-30): temporalSampleCount rgba8
// images; Present() is the fragment shader's mean of the layers
// (vulkan/shaders/renderer/triangle.frag:14-23), as rgba8.
class RenderTarget {
  public:
    uint32_t width = 0, height = 0;
    uint32_t temporalSampleCount = 1;
    frep_target_t handle = 0;

    void Init(Device* device, uint32_t w, uint32_t h, uint32_t samples) {
        m_device = device;
        width = w;
        height = h;
        temporalSampleCount = samples;
        check(device->ctx, frep_target_create(device->ctx, w, h, samples, &handle),
              "frep_target_create");
    }
    void Cleanup() {
        if (handle) frep_target_free(m_device->ctx, handle);
        handle = 0;
    }
    std::vector<uint8_t> Present() const {
        std::vector<uint8_t> img((size_t)width * height * 4);
        check(m_device->ctx, frep_target_resolve(m_device->ctx, handle, img.data()),
              "frep_target_resolve");
        return img;
    }

  private:
    Device* m_device = nullptr;
};

class Raytracer {
  public:
    uint32_t samplesPerPixel = 1;
    bool jitter = false;
    uint32_t seed = 0;
    bool keepHitIds = false;
    bool keepNormals = false;

    void Init(Device* device, uint32_t width, uint32_t height, VoxelStorage* voxels) {
        m_device = device;
        m_voxels = voxels;
        m_width = width;
        m_height = height;
    }
    // The reference's form (raytracer.h:18-21): every Trace writes one
    // jittered sample per pixel into the target's next layer
    // (raytracer.cpp:159-162), the jitter seeded by the frame count.
    void Init(Device* device, RenderTarget* target, VoxelStorage* voxels) {
        Init(device, target->width, target->height, voxels);
        m_target = target;
        m_targetImgLayer = 0;
    }
    void Cleanup() { m_voxels = nullptr; }

    void SetBackgroundColor(const vec3& color) { m_background = {color[0], color[1], color[2], 1.0f}; }

    // Renders the storage's tree from `camera` (raytracer.h:24,
    // raytracer.cpp:243-272).  Shading uses the time of the last
    // voxelization, which the reference's driver passes to both calls.
    void Trace(const Camera* camera, float tapeTime) {
        (void)tapeTime;
        frep_ctx* c = m_device->ctx;
        frep_camera cam{};
        for (int k = 0; k < 3; ++k) {
            cam.pos[k] = camera->position[k];
            cam.forward[k] = camera->forward[k];
        }
        cam.width = m_width;
        cam.height = m_height;
        cam.spp = m_target ? 1u : samplesPerPixel;
        cam.jitter = (jitter || m_target) ? 1u : 0u;
        cam.seed = m_target ? seed + m_frame : seed;
        for (int k = 0; k < 4; ++k) cam.background[k] = m_background[k];
        // the reference's lights (raytracer.cpp:187-192) and ambient term
        const float l0[3] = {-1.0f, -1.0f, -0.2f}, l1[3] = {1.0f, -1.0f, -0.2f};
        cam.light_count = 2;
        normalize_into(l0, cam.light_dir[0]);
        normalize_into(l1, cam.light_dir[1]);
        cam.light_color[0][0] = 1.0f;
        cam.light_color[1][2] = 2.0f;
        for (int k = 0; k < 4; ++k) cam.ambient[k] = k < 3 ? 0.2f : 1.0f;
        const float up0[3] = {camera->initialUp[0], camera->initialUp[1], camera->initialUp[2]};
        check(c, frep_camera_look(&cam, up0, camera->fovDeg), "frep_camera_look");
        if (m_target) {
            check(c, frep_render_layer(c, m_voxels->treeHandle, &cam, m_target->handle,
                                       m_targetImgLayer),
                  "frep_render_layer");
            m_targetImgLayer = (m_targetImgLayer + 1) % m_target->temporalSampleCount;
            ++m_frame;
            return;
        }
        const size_t px = (size_t)m_width * m_height;
        m_image.resize(px * 4);
        m_hits.resize(keepHitIds ? px * cam.spp : 0);
        m_normals.resize(keepNormals ? px * cam.spp * 3 : 0);
        check(c, frep_render(c, m_voxels->treeHandle, &cam, m_image.data(),
                             keepHitIds ? m_hits.data() : nullptr,
                             keepNormals ? m_normals.data() : nullptr),
              "frep_render");
    }

    const std::vector<uint8_t>& Image() const { return m_image; }
    const std::vector<uint64_t>& HitIds() const { return m_hits; }
    const std::vector<float>& Normals() const { return m_normals; }

  private:
    static void normalize_into(const float v[3], float out[4]) {
        const float l = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        for (int k = 0; k < 3; ++k) out[k] = v[k] / l;
        out[3] = 0.0f;
    }
    Device* m_device = nullptr;
    VoxelStorage* m_voxels = nullptr;
    RenderTarget* m_target = nullptr;
    uint32_t m_targetImgLayer = 0, m_frame = 0;
    uint32_t m_width = 0, m_height = 0;
    std::array<float, 4> m_background{0.0f, 0.0f, 0.0f, 1.0f};
    std::vector<uint8_t> m_image;
    std::vector<uint64_t> m_hits;
    std::vector<float> m_normals;
};

}  // namespace frep_engine

#endif
