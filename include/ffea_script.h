/*
 * ffea_script.h — host-side reader of FFEA input (no GPU needed), and creation of
 * GPU blobs from it. Part of libffea_gpu.so.
 *
 * Reads what the reference's World::init reads for the blob path:
 *   .ffea script       FFEA_input_reader.cpp:38-196 (comments, <param>, <system>,
 *                      <blob>, <conformation>, <interactions><ctforces>);
 *                      keys and defaults of SimulationParams.cpp:26-117, 239-776
 *                      (dt, kT, epsilon -> epsilon^2, stokes_visc, scale in meso units)
 *   .node .top .mat    Blob.cpp:1911-2085, 2241-2293
 *   .stokes .pin       Blob.cpp:2297-2358, 3024-3076
 *   .surf              Blob.cpp:2090-2170 (faces must be midside sub-triangles,
 *                      tetra_element_linear.cpp:629-671)
 *   .ctforces          Blob.cpp:2478-2806: linear node forces, rotational forces (axes
 *                      by point and direction; 'n' axes are refused) and linear surface
 *                      forces
 *   .springs           World.cpp:2291-2305, 3543-3643 (springs between nodes of any blobs)
 *   rng_seed           World.cpp:259-295 (glibc srand/rand -> 6 seeds; stream 0)
 * Positions are placed as World::init leaves them: blob centroid moved to the
 * `centroid` tag if given (World.cpp:2748-2768), then the system centroid moved to
 * the centre of the (empty) box when move_into_box is 1 (World.cpp:442-474).
 */
#ifndef FFEA_SCRIPT_H
#define FFEA_SCRIPT_H

#include <stdint.h>

#include "ffea_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ffea_script ffea_script;

typedef struct {
    int num_blobs;
    long long num_steps;      /* the run loop does num_steps + 1 updates (World.cpp:1747) */
    int check;
    int rng_seed;
    int restart;
    int calc_noise, calc_stokes, calc_ctforces, calc_ssint, calc_steric, calc_es, calc_kinetics;
    int calc_springs, calc_preComp, move_into_box;
    int max_iterations_cg;
    double dt, kT, epsilon2, stokes_visc;  /* meso units */
} ffea_script_params;

/* parses the script and every blob's files; 0 on success, else an ffea_gpu_status with
 * the reference's message in ffea_gpu_last_error() */
int ffea_script_load(const char *path, ffea_script **out);
void ffea_script_free(ffea_script *s);
int ffea_script_get_params(const ffea_script *s, ffea_script_params *out);

/* blob b (first conformation): sizes, then arrays (any pointer may be NULL) */
int ffea_script_blob_sizes(const ffea_script *s, int blob, int *num_nodes, int *num_elements, int *num_pinned,
                           int *num_ctf, int *num_surface_faces);
int ffea_script_blob_arrays(const ffea_script *s, int blob, double *pos, int *top, double *A, double *B,
                            double *G, double *E, double *rho, int *pinned, int *ctf_node, double *ctf_force,
                            double *stokes_radius);

/* blob b's linear surface constant forces (.ctforces "linear surface forces", meso units),
 * as ffea_gpu_surface_forces lays them out: counts first, then the arrays (any pointer
 * may be NULL) */
int ffea_script_blob_surface_forces(const ffea_script *s, int blob, int *num_faces, int *num_sets, int *num_entries,
                                    int *face_nodes, double *set_force, int *set_ptr, int *set_faces);

/* blob b's rotational constant forces (.ctforces "rotational forces", meso units), as
 * ffea_gpu_rotational_forces lays them out: the count first, then the arrays (any pointer
 * may be NULL; type holds 2 chars per entry) */
int ffea_script_blob_rotational_forces(const ffea_script *s, int blob, int *num, int *node, double *force,
                                       double *axis, char *type);

/* blob b's solver (World.cpp:2600-2625): FFEA_GPU_SOLVER_NOMASS_CG or _MASSLUMPED */
int ffea_script_blob_solver(const ffea_script *s, int blob, int *solver);

/* the springs of the script (calc_springs = 1), file order: k and l in meso units
 * (k * area / Energy, l / length), blob / conf / node [num][2] for the two ends, active [num]
 * (World::activate_springs: both ends in the first conformation). Any pointer may be NULL. */
int ffea_script_springs(const ffea_script *s, int *num, double *k, double *l, int *blob, int *conf, int *node,
                        int *active);

/* the noise stream state of the single-thread run (rng[0] after SetPackageSeed) */
int ffea_script_rng_state(const ffea_script *s, uint32_t state[6]);

/* blob b of the script on the GPU (ffea_gpu_blob_create with the script's params). The
 * nodes active springs pull are the handle's external force nodes; the spring forces
 * themselves are the World's to apply (ffea_world_run does, World::apply_springs). */
int ffea_gpu_blob_create_from_script(const ffea_script *s, int blob, int precision, int device,
                                     ffea_gpu_blob **out);

#ifdef __cplusplus
}
#endif
#endif
