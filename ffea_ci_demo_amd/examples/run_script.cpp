// `ffea <script>` for the mass-less blob path (ffea.cpp:53-308): every blob of an .ffea
// script on the GPU through the World loop, writing the script's trajectory, measurement
// and checkpoint files (restart = 1 resumes from them).
//   run_script <script.ffea> [max_updates] [fp64|mixed]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ffea_gpu.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <script.ffea> [max_updates] [fp64|mixed]\n", argv[0]);
        return 2;
    }
    try {
        const long long max_updates = argc > 2 ? std::atoll(argv[2]) : -1;
        const int precision = (argc > 3 && std::strcmp(argv[3], "mixed") == 0) ? FFEA_GPU_MIXED : FFEA_GPU_FP64;
        ffea_gpu::World world(argv[1], precision);
        const long long first = world.step();
        const auto t0 = std::chrono::steady_clock::now();
        world.run(max_updates);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ffea_gpu_blob_info info{};
        ffea_gpu_blob *b0 = world.blob(0);
        ffea_gpu_blob_get_info(b0, &info);
        std::vector<double> pos((size_t)3 * info.num_nodes);
        ffea_gpu_blob_get_positions(b0, pos.data());
        int iters = 0;
        ffea_gpu_blob_get_cg_iterations(b0, &iters, nullptr);
        std::printf("done: %lld updates in %.3f s (last cg iterations %d); blob 0 node 0 at (%.10e, %.10e, %.10e)\n",
                    world.step() - first, s, iters, pos[0], pos[1], pos[2]);
    } catch (const ffea_gpu::Exception &e) {
        std::fprintf(stderr, "FFEA error (status %d): %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
