// Runs the blobs of an .ffea script on the GPU for num_steps + 1 updates (the
// reference's World::run loop length, World.cpp:1747) and prints progress every
// `check` steps: the mass-less blob path of `ffea <script>` (ffea.cpp:53-308).
//   run_script <script.ffea> [max_updates] [fp64|mixed]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "ffea_gpu.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <script.ffea> [max_updates] [fp64|mixed]\n", argv[0]);
        return 2;
    }
    try {
        ffea_gpu::Script script(argv[1]);
        const ffea_script_params p = script.params();
        long long updates = p.num_steps + 1;
        if (argc > 2) updates = std::min(updates, std::atoll(argv[2]));
        const int precision = (argc > 3 && std::strcmp(argv[3], "mixed") == 0) ? FFEA_GPU_MIXED : FFEA_GPU_FP64;
        std::vector<std::unique_ptr<ffea_gpu::Blob>> blobs;
        for (int b = 0; b < p.num_blobs; ++b) blobs.emplace_back(new ffea_gpu::Blob(script, b, precision));
        const auto t0 = std::chrono::steady_clock::now();
        for (long long step = 0; step < updates; ++step) {
            for (auto &b : blobs) b->update_internal_forces();
            for (auto &b : blobs) b->update_positions();
            if (p.check > 0 && step % p.check == 0)
                std::printf("step %lld: cg iterations %d\n", step, blobs[0]->cg_iterations());
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<double> pos = blobs[0]->positions();
        std::printf("done: %lld updates in %.3f s; blob 0 node 0 at (%.10e, %.10e, %.10e)\n", updates, s, pos[0],
                    pos[1], pos[2]);
    } catch (const ffea_gpu::Exception &e) {
        std::fprintf(stderr, "FFEA error (status %d): %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
