/*
 * ffea_world.h — the World loop of an FFEA script on the GPU blob path (libffea_gpu.so).
 *
 * A world is every blob of one .ffea script, each a GPU blob handle, driven the way the
 * reference's World drives them, with the reference's output files:
 *
 *   World::init (World.cpp:155-930)      ffea_world_create: builds the blobs, opens the
 *                                        trajectory, measurement and detailed measurement
 *                                        files (headers, World.cpp:519-646) or, with
 *                                        restart = 1, reads
 *                                        the checkpoint's noise stream state
 *                                        (World.cpp:302-426) and the last complete
 *                                        trajectory frame, truncates the trajectory and
 *                                        measurement files there and appends
 *                                        (World.cpp:680-930); existing output files are
 *                                        first moved aside as __name__bckp.N
 *                                        (SimulationParams.cpp:779-802, 996-1009)
 *   World::run (World.cpp:1747-2060)     ffea_world_run: steps step_initial..num_steps;
 *                                        every `check` steps the checkpoint (noise stream
 *                                        state before the forces, World.cpp:4502-4546),
 *                                        then after update_internal_forces the trajectory
 *                                        frame (Blob.cpp:855-889, World.cpp:4461-4488)
 *                                        and the measurement line: time, strain energy,
 *                                        centroid, RMSD (Blob.cpp:1066-1202,
 *                                        World.cpp:4598-4692), and the detailed line: the
 *                                        same per blob, spring energy per blob pair
 *                                        (Blob.cpp:1204-1213, World.cpp:4504-4537,
 *                                        4694-4721); then update_positions
 *
 * The supported subset is that of ffea_gpu_blob_create_from_script (the mass-less, lumped-
 * mass and consistent-mass solvers, DYNAMIC blobs, one conformation, constant forces,
 * springs; no surface interactions, electrostatics, PreComp, kinetics or rods). Single-threaded noise semantics: one thermal stream, shared by the blobs
 * in order, as the reference's ffea with one thread.
 */
#ifndef FFEA_WORLD_H
#define FFEA_WORLD_H

#include "ffea_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ffea_world ffea_world;

/* precision: ffea_gpu_precision; device: HIP device ordinal. Writes the detailed measurement
 * file as the reference's ffea does by default: det_measurement_out_fname, else the
 * measurement file's name with the extension .fdm (SimulationParams.cpp:690-709); a script
 * that names no measurement file gets none. */
int ffea_world_create(const char *ffea_script_path, int precision, int device, ffea_world **out);

/* ffea_world_create with the reference runner's options (ffea.cpp:86-88, 166):
 * flags FFEA_WORLD_NO_DETAIL writes no detailed measurement file (-d / --no-detail);
 * frames_to_delete >= 0: with restart = 1, the last complete frames dropped from the
 * trajectory and measurement files before the run resumes (-l / --delete-frames,
 * World.cpp:697-912); the noise stream still comes from the checkpoint */
#define FFEA_WORLD_NO_DETAIL 1
int ffea_world_create_ex(const char *ffea_script_path, int precision, int device, int flags, int frames_to_delete,
                         ffea_world **out);

/* runs the loop from the current step; max_steps < 0 runs to num_steps (inclusive, as
 * the reference's `step < num_steps + 1`), otherwise at most max_steps loop iterations.
 * A failing step (inversion, CG non-convergence) stops the run with its status; the
 * files hold everything written before it. */
int ffea_world_run(ffea_world *world, long long max_steps);

/* the next step the loop will run (step_initial after a restart) */
long long ffea_world_current_step(const ffea_world *world);

/* the GPU blob handle of blob `index` (owned by the world) */
int ffea_world_blob(ffea_world *world, int index, ffea_gpu_blob **blob);

/* closes the files and releases every blob; NULL is accepted */
void ffea_world_destroy(ffea_world *world);

const char *ffea_world_last_error(const ffea_world *world);

#ifdef __cplusplus
}
#endif
#endif
