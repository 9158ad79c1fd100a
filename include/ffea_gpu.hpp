// ffea_gpu.hpp — header-only C++ wrapper of the C ABI (ffea_gpu.h, ffea_script.h) for
// host code written like the reference's: a blob object with the per-step pair
// update_internal_forces() / update_positions() (Blob.h:130, 141) and failures thrown
// as exceptions carrying the reference's messages (FFEA_return_codes.h:52-69).
#ifndef FFEA_GPU_HPP
#define FFEA_GPU_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ffea_gpu.h"
#include "ffea_script.h"
#include "ffea_world.h"

namespace ffea_gpu {

class Exception : public std::runtime_error {
public:
    Exception(int status, const std::string &msg) : std::runtime_error(msg), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

class Script {
public:
    explicit Script(const std::string &path) {
        if (int rc = ffea_script_load(path.c_str(), &s_)) throw Exception(rc, ffea_gpu_last_error());
    }
    ~Script() { ffea_script_free(s_); }
    Script(const Script &) = delete;
    Script &operator=(const Script &) = delete;
    ffea_script_params params() const {
        ffea_script_params p{};
        ffea_script_get_params(s_, &p);
        return p;
    }
    const ffea_script *get() const { return s_; }

private:
    ffea_script *s_ = nullptr;
};

// One blob on the GPU: the reference's Blob for the mass-less dynamic path.
class Blob {
public:
    Blob(const Script &script, int blob_index, int precision = FFEA_GPU_FP64, int device = 0) {
        if (int rc = ffea_gpu_blob_create_from_script(script.get(), blob_index, precision, device, &b_))
            throw Exception(rc, ffea_gpu_last_error());
        int nn = 0;
        ffea_script_blob_sizes(script.get(), blob_index, &nn, nullptr, nullptr, nullptr, nullptr);
        num_nodes_ = nn;
    }
    Blob(const ffea_gpu_params &p, int num_nodes, int num_elements, const double *pos, const int *top,
         const double *A, const double *B, const double *G, const double *E, int num_pinned, const int *pinned,
         int num_ctf, const int *ctf_node, const double *ctf_force, const uint32_t rng_state[6])
        : num_nodes_(num_nodes) {
        if (int rc = ffea_gpu_blob_create(&p, num_nodes, num_elements, pos, top, A, B, G, E, num_pinned, pinned,
                                          num_ctf, ctf_node, ctf_force, nullptr, rng_state, &b_))
            throw Exception(rc, ffea_gpu_last_error());
    }
    ~Blob() { ffea_gpu_blob_destroy(b_); }
    Blob(const Blob &) = delete;
    Blob &operator=(const Blob &) = delete;

    void update_internal_forces() { check(ffea_gpu_blob_update_internal_forces(b_)); }
    void update_positions() { check(ffea_gpu_blob_update_positions(b_)); }
    void step(int n = 1) { check(ffea_gpu_blob_step(b_, n)); }

    std::vector<double> positions() { return get(ffea_gpu_blob_get_positions); }
    std::vector<double> velocities() { return get(ffea_gpu_blob_get_velocities); }
    std::vector<double> rhs() { return get(ffea_gpu_blob_get_rhs); }
    void set_positions(const std::vector<double> &p) { check(ffea_gpu_blob_set_positions(b_, p.data())); }
    int cg_iterations() {
        int last = 0;
        ffea_gpu_blob_get_cg_iterations(b_, &last, nullptr);
        return last;
    }
    std::vector<uint32_t> rng_state() {
        std::vector<uint32_t> s(6);
        check(ffea_gpu_blob_get_rng_state(b_, s.data()));
        return s;
    }
    int num_nodes() const { return num_nodes_; }
    ffea_gpu_blob *handle() { return b_; }

    // gradients (noise off, some nodes pinned): per-element A, B, G, E; a checkpointed run;
    // its reverse from dL/dpos at the end (.node order) to dL/d(A, B, G, E) and dL/dpos0
    void set_material(const std::vector<double> &A, const std::vector<double> &B, const std::vector<double> &G,
                      const std::vector<double> &E) {
        check(ffea_gpu_blob_set_material(b_, A.data(), B.data(), G.data(), E.data()));
    }
    void run_checkpointed(int nsteps, int interval = 20) { check(ffea_gpu_blob_run_checkpointed(b_, nsteps, interval)); }
    void adjoint(const std::vector<double> &dL_dpos, std::vector<double> &dA, std::vector<double> &dB,
                 std::vector<double> &dG, std::vector<double> &dE, std::vector<double> *dL_dpos0 = nullptr) {
        ffea_gpu_blob_info info{};
        check(ffea_gpu_blob_get_info(b_, &info));
        for (auto *v : {&dA, &dB, &dG, &dE}) v->assign(info.num_elements, 0.0);
        if (dL_dpos0) dL_dpos0->assign((size_t)3 * num_nodes_, 0.0);
        check(ffea_gpu_blob_adjoint(b_, dL_dpos.data(), dA.data(), dB.data(), dG.data(), dE.data(),
                                    dL_dpos0 ? dL_dpos0->data() : nullptr));
    }
    void use_graph(bool on, int iterations = 0) { check(ffea_gpu_blob_use_graph(b_, on ? 1 : 0, iterations)); }
    // FFEA_GPU_PRECOND_JACOBI (the reference's) or FFEA_GPU_PRECOND_MULTIGRID
    ffea_gpu_multigrid_info set_preconditioner(int kind) {
        check(ffea_gpu_blob_set_preconditioner(b_, kind));
        ffea_gpu_multigrid_info mi{};
        check(ffea_gpu_blob_get_multigrid_info(b_, &mi));
        return mi;
    }
    int autotune(int iterations = 20) {
        int chosen = 0;
        check(ffea_gpu_blob_autotune(b_, iterations, &chosen));
        return chosen;
    }

private:
    void check(int rc) {
        if (rc) throw Exception(rc, ffea_gpu_blob_last_error(b_));
    }
    std::vector<double> get(int (*f)(ffea_gpu_blob *, double *)) {
        std::vector<double> v((size_t)3 * num_nodes_);
        check(f(b_, v.data()));
        return v;
    }
    ffea_gpu_blob *b_ = nullptr;
    int num_nodes_ = 0;
};

// The reference's NoMassCGSolver on the GPU behind the shape of its Solver interface
// (Solver.h:31-52), for a Blob that keeps its CPU element loop: constructed where
// NoMassCGSolver::init runs (Blob.cpp:247) and asked to solve in place where Blob calls
// solver->solve(force) (Blob.cpp:3243). pos are the node positions the element loop used
// for the step ([num_nodes][3]); x holds the right-hand side on entry, the velocities on
// return.
class NoMassSolver {
public:
    NoMassSolver(const ffea_gpu_params &p, int num_nodes, int num_elements, const double *pos, const int *top,
                 const double *A, const double *B, int num_pinned, const int *pinned,
                 const double *stokes_drag = nullptr)
        : num_nodes_(num_nodes) {
        if (int rc = ffea_gpu_solver_create(&p, num_nodes, num_elements, pos, top, A, B, num_pinned, pinned,
                                            stokes_drag, &b_))
            throw Exception(rc, ffea_gpu_last_error());
    }
    ~NoMassSolver() { ffea_gpu_blob_destroy(b_); }
    NoMassSolver(const NoMassSolver &) = delete;
    NoMassSolver &operator=(const NoMassSolver &) = delete;

    void solve(const double *pos, double *x) { check(ffea_gpu_solver_solve(b_, pos, x)); }
    void apply_matrix(const double *pos, const double *in, double *out) {
        check(ffea_gpu_solver_apply_matrix(b_, pos, in, out));
    }
    // FFEA_GPU_PRECOND_JACOBI (the reference's) or FFEA_GPU_PRECOND_MULTIGRID
    void set_preconditioner(int kind) { check(ffea_gpu_blob_set_preconditioner(b_, kind)); }
    int cg_iterations() {
        int last = 0;
        check(ffea_gpu_blob_get_cg_iterations(b_, &last, nullptr));
        return last;
    }
    int num_nodes() const { return num_nodes_; }
    ffea_gpu_blob *handle() { return b_; }

private:
    void check(int rc) {
        if (rc) throw Exception(rc, ffea_gpu_blob_last_error(b_));
    }
    ffea_gpu_blob *b_ = nullptr;
    int num_nodes_ = 0;
};

// Every blob of a script with the reference's output files: World::init / World::run
// (World.cpp:155-930, 1747-2060), restart included.
class World {
public:
    // detailed = false: no detailed measurement file (.fdm), as the reference's ffea -d;
    // delete_frames: on restart, complete frames dropped first, as ffea -l
    explicit World(const std::string &script_path, int precision = FFEA_GPU_FP64, int device = 0,
                   bool detailed = true, int delete_frames = 0) {
        if (int rc = ffea_world_create_ex(script_path.c_str(), precision, device, detailed ? 0 : FFEA_WORLD_NO_DETAIL,
                                          delete_frames, &w_))
            throw Exception(rc, ffea_gpu_last_error());
    }
    ~World() { ffea_world_destroy(w_); }
    World(const World &) = delete;
    World &operator=(const World &) = delete;
    // to num_steps (inclusive), or at most max_steps loop iterations
    void run(long long max_steps = -1) {
        if (int rc = ffea_world_run(w_, max_steps)) throw Exception(rc, ffea_world_last_error(w_));
    }
    long long step() const { return ffea_world_current_step(w_); }
    ffea_gpu_blob *blob(int index) {
        ffea_gpu_blob *b = nullptr;
        if (int rc = ffea_world_blob(w_, index, &b)) throw Exception(rc, "no such blob");
        return b;
    }

private:
    ffea_world *w_ = nullptr;
};

}  // namespace ffea_gpu

#endif
