// `ffea <script>` for the mass-less blob path (ffea.cpp:53-308): every blob of an .ffea
// script on the GPU through the World loop, writing the script's trajectory, measurement,
// detailed measurement and checkpoint files (restart = 1 resumes from them); -d writes no
// detailed measurement file, as ffea -d; -l N drops N final frames on restart, as ffea -l.
//   run_script [-d] [-l N] <script.ffea> [max_updates] [fp64|mixed]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ffea_gpu.hpp"

int main(int argc, char **argv) {
    bool detailed = true;
    int delete_frames = 0;
    std::vector<char *> args;  // the positional arguments
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-d") == 0 || std::strcmp(argv[i], "--no-detail") == 0) detailed = false;
        else if ((std::strcmp(argv[i], "-l") == 0 || std::strcmp(argv[i], "--delete-frames") == 0) && i + 1 < argc)
            delete_frames = std::atoi(argv[++i]);
        else args.push_back(argv[i]);
    }
    if (args.empty()) {
        std::fprintf(stderr, "usage: %s [-d] [-l N] <script.ffea> [max_updates] [fp64|mixed]\n", argv[0]);
        return 2;
    }
    try {
        const long long max_updates = args.size() > 1 ? std::atoll(args[1]) : -1;
        const int precision = (args.size() > 2 && std::strcmp(args[2], "mixed") == 0) ? FFEA_GPU_MIXED : FFEA_GPU_FP64;
        ffea_gpu::World world(args[0], precision, 0, detailed, delete_frames);
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
