// csg_demo.cpp -- the reference's demo shapes (vulkan/src/shapes.cpp:5-117)
// built with include/frep_csg.hpp and compiled by frep_compile, the way the
// reference's main (vulkan/src/main.cpp:16-31) builds csg::Tape(shape) and
// prints it.  One line per shape:
//   name n_instr slot_count format words(hex,...) consts(hex bits,...)
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "../../include/frep_csg.hpp"

int main() {
    using namespace frep_csg;
    const std::vector<std::pair<std::string, std::function<Expr()>>> shapes = {
        {"screw", [] { return Screw(); }},
        {"elastic_cube", [] { return ElasticCube(); }},
        {"tangle_cube", [] { return TangleCube(); }},
        {"morph", [] { return Morph(); }},
        {"menger_sponge3", [] { return MengerSponge(3); }},
        {"twisted_tower", [] { return TwistedTower(6, 0.35f, 0.8f); }},
        {"organic_balls", [] { return OrganicBalls(); }},
    };
    try {
        for (const auto& s : shapes) {
            const Tape tape(s.second());
            uint32_t minmax = 0;
            for (const Instr& i : tape.instructions) minmax += (i.op == 7 || i.op == 8);
            std::printf("%s %zu %u %u %u ", s.first.c_str(), tape.instructions.size(),
                        tape.GetSlotCount(), tape.format, minmax);
            for (size_t i = 0; i < tape.words.size(); ++i)
                std::printf("%s%x", i ? "," : "", tape.words[i]);
            std::printf(" ");
            for (size_t i = 0; i < tape.constantPool.size(); ++i) {
                uint32_t b;
                std::memcpy(&b, &tape.constantPool[i], 4);
                std::printf("%s%x", i ? "," : "", b);
            }
            std::printf("\n");
        }
    } catch (const frep_engine::Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
