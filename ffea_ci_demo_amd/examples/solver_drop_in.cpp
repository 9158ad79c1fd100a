// The solver-level drop-in: the reference's Blob keeps its element loop and hands each
// step's right-hand side to NoMassCGSolver::solve (Blob.cpp:3243; Solver.h:42-51). Here a
// GPU blob stands in for that element loop: each step, at the positions its element pass
// used, ffea_gpu::NoMassSolver solves the blob's right-hand side in place, and the
// solution is compared with the velocities the blob's own step produced.
//   solver_drop_in <script.ffea> [steps]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ffea_gpu.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <script.ffea> [steps]\n", argv[0]);
        return 2;
    }
    try {
        const int steps = argc > 2 ? std::atoi(argv[2]) : 5;
        ffea_gpu::Script script(argv[1]);
        const ffea_script_params sp = script.params();
        int nt = 0, ne = 0, npin = 0, nctf = 0;
        ffea_script_blob_sizes(script.get(), 0, &nt, &ne, &npin, &nctf, nullptr);
        std::vector<double> pos((size_t)3 * nt), A(ne), B(ne), G(ne), E(ne), rho(ne), ctf_force((size_t)3 * nctf),
            radius(nt);
        std::vector<int> top((size_t)10 * ne), pinned(npin), ctf_node(nctf);
        if (int rc = ffea_script_blob_arrays(script.get(), 0, pos.data(), top.data(), A.data(), B.data(), G.data(),
                                             E.data(), rho.data(), pinned.data(), ctf_node.data(), ctf_force.data(),
                                             radius.data()))
            throw ffea_gpu::Exception(rc, ffea_gpu_last_error());
        std::vector<double> drag(nt);  // Blob::init (Blob.cpp:264-268)
        for (int i = 0; i < nt; ++i) drag[i] = 6.0 * M_PI * sp.stokes_visc * radius[i];
        ffea_gpu_params p{sp.dt, sp.kT, sp.epsilon2, sp.max_iterations_cg, sp.calc_noise, sp.calc_stokes,
                          FFEA_GPU_FP64, 0};
        ffea_gpu::Blob blob(script, 0);
        ffea_gpu::NoMassSolver solver(p, nt, ne, pos.data(), top.data(), A.data(), B.data(), npin, pinned.data(),
                                      drag.data());
        double worst = 0.0;
        for (int s = 0; s < steps; ++s) {
            const std::vector<double> x0 = blob.positions();
            blob.update_internal_forces();
            blob.update_positions();
            std::vector<double> x = blob.rhs();
            solver.solve(x0.data(), x.data());
            const std::vector<double> v = blob.velocities();
            double dmax = 0.0, vmax = 0.0;
            for (size_t i = 0; i < v.size(); ++i) {
                dmax = std::fmax(dmax, std::fabs(x[i] - v[i]));
                vmax = std::fmax(vmax, std::fabs(v[i]));
            }
            const double rel = vmax > 0.0 ? dmax / vmax : dmax;
            worst = std::fmax(worst, rel);
            std::printf("step %d: solver iterations %d, blob iterations %d, max rel diff %.3e\n", s,
                        solver.cg_iterations(), blob.cg_iterations(), rel);
        }
        std::printf("done: %d steps, max rel diff %.3e\n", steps, worst);
    } catch (const ffea_gpu::Exception &e) {
        std::fprintf(stderr, "FFEA error (status %d): %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
