// engine_demo.cpp -- the reference's per-frame driver
// (vulkan/src/application.cpp:60-93: Voxelize, then Trace) written against
// the C++ engine classes of include/frep_engine.hpp.  Voxelizes one tape,
// traces one frame and writes the canonical tree blob, the rgba8 image and
// the per-pixel hit ids; then traces 4 frames into a 4-layer render target
// and writes the presented image.
//
// usage: engine_demo TAPE.bin d0,d1,... WIDTH HEIGHT OUT_PREFIX
//   TAPE.bin: u32 n_words, n_words u32 words ([n, instrs]), u32 n_consts,
//             n_consts f32 constants, u32 slot_count (little endian)
//   writes OUT_PREFIX.tree (u32 blob), OUT_PREFIX.rgba, OUT_PREFIX.hits,
//   OUT_PREFIX.present (rgba8)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/frep_engine.hpp"

namespace {

template <class T>
bool read_vec(std::ifstream& in, std::vector<T>& v) {
    uint32_t n = 0;
    if (!in.read(reinterpret_cast<char*>(&n), 4)) return false;
    v.resize(n);
    return n == 0 || (bool)in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)n * sizeof(T));
}

template <class T>
bool write_file(const std::string& path, const std::vector<T>& v) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
    return (bool)out;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s TAPE.bin d0,d1,... WIDTH HEIGHT OUT_PREFIX\n", argv[0]);
        return 2;
    }
    frep_engine::Tape tape;
    {
        std::ifstream in(argv[1], std::ios::binary);
        if (!read_vec(in, tape.words) || !read_vec(in, tape.constantPool) ||
            !in.read(reinterpret_cast<char*>(&tape.slotCount), 4)) {
            std::fprintf(stderr, "cannot read tape file %s\n", argv[1]);
            return 2;
        }
    }
    std::vector<uint32_t> dims;
    {
        std::stringstream ss(argv[2]);
        std::string tok;
        while (std::getline(ss, tok, ',')) dims.push_back((uint32_t)std::strtoul(tok.c_str(), nullptr, 10));
    }
    const uint32_t width = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    const uint32_t height = (uint32_t)std::strtoul(argv[4], nullptr, 10);
    const std::string prefix = argv[5];

    try {
        frep_engine::Device device;
        device.Init(0);

        frep_engine::VoxelStorage voxels;
        voxels.gridDims = dims;
        voxels.lowVertex = {-1.0f, -1.0f, -1.0f};
        voxels.worldSize = 2.0f;
        voxels.Init(&device, tape);

        frep_engine::Voxelizer voxelizer;
        voxelizer.Init(&device, &voxels);
        frep_engine::Raytracer raytracer;
        raytracer.Init(&device, width, height, &voxels);
        raytracer.SetBackgroundColor({0.0f, 0.0f, 0.0f});
        raytracer.keepHitIds = true;

        frep_engine::Camera camera;
        camera.position = {0.0f, 0.0f, 3.0f};
        camera.forward = {0.0f, 0.0f, -1.0f};
        camera.initialUp = {0.0f, 1.0f, 0.0f};
        camera.fovDeg = 45.0f;
        camera.aspectRatio = (float)width / (float)height;
        camera.Init();

        const float time = 0.0f;
        voxelizer.Voxelize(time);
        raytracer.Trace(&camera, time);
        voxelizer.PrintStats();

        // the reference's frame loop with a layered target: one jittered
        // sample per frame into the next layer, then the presented mean
        frep_engine::RenderTarget target;
        target.Init(&device, width, height, 4);
        frep_engine::Raytracer layered;
        layered.Init(&device, &target, &voxels);
        for (int frame = 0; frame < 4; ++frame) layered.Trace(&camera, time);
        const std::vector<uint8_t> presented = target.Present();
        target.Cleanup();

        uint64_t n = 0;
        frep_engine::check(device.ctx, frep_tree_export(device.ctx, voxels.treeHandle, nullptr, 0, &n),
                           "frep_tree_export");
        std::vector<uint32_t> blob(n);
        frep_engine::check(device.ctx,
                           frep_tree_export(device.ctx, voxels.treeHandle, blob.data(), n, &n),
                           "frep_tree_export");
        if (!write_file(prefix + ".tree", blob) || !write_file(prefix + ".rgba", raytracer.Image()) ||
            !write_file(prefix + ".hits", raytracer.HitIds()) ||
            !write_file(prefix + ".present", presented)) {
            std::fprintf(stderr, "cannot write outputs under %s\n", prefix.c_str());
            return 1;
        }
        voxels.Cleanup();
        device.Cleanup();
    } catch (const frep_engine::Error& e) {
        std::fprintf(stderr, "frep error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
