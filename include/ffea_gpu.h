/*
 * ffea_gpu.h — C ABI of the MI355X blob time-step library (libffea_gpu.so).
 *
 * A handle is one FFEA blob on one GPU (ffea_gpu_blob_create), or an ensemble of
 * independent blobs stepped together (ffea_gpu_batch_create): meshes, materials,
 * rest states, the mass-less viscosity solver and the MRG32k3a noise streams live
 * in device memory. The entry points replace, for each blob, the reference's
 *
 *   Blob::init + NoMassCGSolver::init   Blob.cpp:161-247, NoMassCGSolver.cpp:56-153
 *   Blob::update_internal_forces        Blob.cpp:400-491 (Blob.h:130)
 *   Blob::update_positions              Blob.cpp:493-508 (Blob.h:141), with
 *     Blob::aggregate_forces_and_solve  Blob.cpp:3147-3244
 *     NoMassCGSolver::solve             NoMassCGSolver.cpp:156-188 (Solver.h:42-51)
 *     Blob::euler_integrate             Blob.cpp:3249-3263
 *   node positions / velocities         mesh_node::pos / vel (mesh_node.h)
 *   RngStream::GetState / SetSeed       RngStream.cpp:439-443, 251-259
 *
 * Conventions: plain pointers and sizes; arrays are host memory owned by the
 * caller and copied in or out; node vectors are [num_nodes][3] doubles in .node
 * order (all nodes, midside included); element arrays follow .top order.
 * Every call returns an ffea_gpu_status; a message for the last failure is kept
 * per handle (ffea_gpu_blob_last_error) or per thread when no handle exists
 * (ffea_gpu_last_error). A handle is not re-entrant; distinct handles may be
 * used from different host threads. Each handle owns one HIP stream.
 */
#ifndef FFEA_GPU_H
#define FFEA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FFEA_GPU_OK = 0,
    FFEA_GPU_ERR_INVERSION = 1,      /* "N elements have inverted since the last step" (Blob.cpp:484-490) */
    FFEA_GPU_ERR_CG_NO_CONVERGE = 2, /* "Could not converge after N iterations" (NoMassCGSolver.cpp:187) */
    FFEA_GPU_ERR_INVALID = 3,        /* bad argument or mesh */
    FFEA_GPU_ERR_HIP = 4,            /* HIP runtime failure (message holds the HIP error) */
    FFEA_GPU_ERR_NO_DEVICE = 5,
    FFEA_GPU_ERR_IO = 6              /* an input or output file of a world run (FFEAFileException) */
} ffea_gpu_status;

typedef enum {
    FFEA_GPU_FP64 = 0,   /* all arithmetic in FP64, the reference's algorithm and orders */
    FFEA_GPU_MIXED = 1   /* element algebra and the viscosity matrix in FP32, matrix products
                            gathering an FP32 copy of their input vector and accumulating in
                            FP64; CG vectors, inner products, residual norms and positions in
                            FP64 */
} ffea_gpu_precision;

/* Run parameters in FFEA's internal units (SimulationParams after assign/validate). */
typedef struct {
    double dt;               /* SimulationParams::dt */
    double kT;               /* SimulationParams::kT */
    double epsilon2;         /* SimulationParams::epsilon2 (epsilon squared) */
    int max_iterations_cg;   /* SimulationParams::max_iterations_cg */
    int calc_noise;          /* SimulationParams::calc_noise */
    int calc_stokes;         /* SimulationParams::calc_stokes (1: per-node Stokes drag and noise) */
    int precision;           /* ffea_gpu_precision */
    int device;              /* HIP device ordinal */
} ffea_gpu_params;

typedef struct ffea_gpu_blob ffea_gpu_blob;

/* sizes of the device-side problem */
typedef struct {
    int num_nodes, num_elements, num_linear_nodes;
    long long num_blocks;        /* 3x3 blocks of the viscosity matrix (pattern, no padding) */
    long long num_block_slots;   /* stored block slots (with slice padding) */
    int precision;
    long long device_bytes;      /* device memory held by the handle */
    int num_blobs;
    int num_chunks;              /* reduction chunks (up to 1024 node rows of one blob each) */
    int row_blocks;              /* workgroups of the row kernels (gather, CG, update) */
    int graph_iterations;        /* CG iterations in the captured step graph (0: none) */
    int single_launch_cg;        /* 1: every blob is small enough (up to 1024 corner nodes) that
                                    its whole CG solve runs in one launch, one workgroup per blob */
} ffea_gpu_blob_info;

/* Blob::init + NoMassCGSolver::init (Blob.cpp:161-247, NoMassCGSolver.cpp:56-153).
 *   pos        [num_nodes][3] node positions (meso units, .node order, scale applied)
 *   top        [num_elements][10] node indices (.top order; 4 corners then the
 *              midside nodes of edges 01 02 03 12 13 23)
 *   A, B, G, E [num_elements] shear viscosity, second viscosity coefficient, shear and
 *              bulk modulus in meso units (Blob::load_material_params, Blob.cpp:2282-2287)
 *   pinned     [num_pinned] node indices (Blob::load_pinned_nodes, Blob.cpp:3024-3076)
 *   ctf_node, ctf_force  [num_ctf], [num_ctf][3] constant linear node forces in meso
 *              units, applied in list order (Blob::apply_ctforces, Blob.cpp:1679-1683)
 *   stokes_drag [num_nodes] 6 pi stokes_visc stokes_radius per node (meso units; Blob.cpp:264-268),
 *              used when params->calc_stokes = 1: drag on the viscosity matrix diagonal of
 *              unpinned nodes (NoMassCGSolver.cpp:119-127) and, with calc_noise, 3 draws per
 *              node of noise force after the element draws (Blob.cpp:3200-3222)
 *   rng_state  6 x uint32: the state of the blob's noise stream (rng[0] of World,
 *              World.cpp:259-295)
 * On success *out holds the handle. */
int ffea_gpu_blob_create(const ffea_gpu_params *params, int num_nodes, int num_elements,
                         const double *pos, const int *top, const double *A, const double *B,
                         const double *G, const double *E, int num_pinned, const int *pinned,
                         int num_ctf, const int *ctf_node, const double *ctf_force,
                         const double *stokes_drag, const uint32_t rng_state[6],
                         ffea_gpu_blob **out);

/* Linear surface constant forces of a blob (Blob::load_ctforces "linear surface forces",
 * Blob.cpp:2730-2806; applied by Blob::apply_ctforces, Blob.cpp:1721-1740): set s spreads
 * set_force[s] over its faces in proportion to their areas at the step's positions, a
 * third of each face's share on each of its three nodes, aggregated after the element
 * forces in face order (Blob.cpp:3160-3170); midside shares are folded by linearise_force.
 *   face_nodes  [num_faces][3] node indices of the .surf faces, in file order (Face::n[0..2])
 *   set_force   [num_sets][3] F * unit direction / force unit (meso units)
 *   set_ptr     [num_sets+1], set_faces [set_ptr[num_sets]]: each set's faces, in order */
typedef struct {
    int num_faces;
    const int *face_nodes;
    int num_sets;
    const double *set_force;
    const int *set_ptr;
    const int *set_faces;
} ffea_gpu_surface_forces;

/* Rotational constant forces of a blob (Blob::load_ctforces "rotational forces",
 * Blob.cpp:2650-2726: the members ctf_r_nodes, ctf_r_forces, ctf_r_axis, ctf_r_type;
 * applied by Blob::apply_ctforces CTF 2, Blob.cpp:1684-1719, after the linear constant
 * forces): entry i acts on node[i] at the step's positions. With r the vector from the
 * node to its foot on the axis through axis[6i..6i+2] along axis[6i+3..6i+5], the node
 * receives force[i] (r x d) for type 'f' or (force[i] / |r|) (r x d) for type 't'; midside
 * shares are folded by linearise_force.
 *   node   [num] node indices ("all" expanded to every node, in order)
 *   force  [num] F / force unit
 *   axis   [num][6] point / length unit, then the unit direction
 *   type   [num][2] 'p' (axis by point and direction), then 'f' or 't'. The reference
 *          loader also accepts 'n' (axis by two nodes) but its force loop never forms r
 *          for it (Blob.cpp:1697-1706); such entries are refused here. */
typedef struct {
    int num;
    const int *node;
    const double *force;
    const double *axis;
    const char *type;
} ffea_gpu_rotational_forces;

/* ffea_gpu_blob_create with surface and rotational constant forces (either may be NULL:
 * none) and the nodes that may receive per-step external forces through
 * ffea_gpu_blob_add_node_forces (external_nodes [num_external_nodes], may be 0 / NULL; the
 * nodes World::apply_springs pulls, World.cpp:4063-4108). Position-dependent constant
 * forces make a handle refuse ffea_gpu_blob_run_checkpointed / ffea_gpu_blob_adjoint. */
int ffea_gpu_blob_create_ex(const ffea_gpu_params *params, int num_nodes, int num_elements,
                            const double *pos, const int *top, const double *A, const double *B,
                            const double *G, const double *E, int num_pinned, const int *pinned,
                            int num_ctf, const int *ctf_node, const double *ctf_force,
                            const double *stokes_drag, const uint32_t rng_state[6],
                            const ffea_gpu_surface_forces *surface_forces,
                            const ffea_gpu_rotational_forces *rotational_forces, int num_external_nodes,
                            const int *external_nodes, ffea_gpu_blob **out);

/* An ensemble of independent blobs in one handle (no counterpart in the reference, which
 * steps the blobs of a World one after another, World.cpp:2031-2059). Blob b is given as
 * for ffea_gpu_blob_create by element b of each array: num_nodes[b], pos[b], top[b], ...;
 * num_pinned / pinned and num_ctf / ctf_node / ctf_force may be NULL (none); rng_states
 * is [num_blobs][6], one noise stream per blob, each moving 7 draws per element of its
 * blob per step (ffea_gpu_rng_stream_state gives independent RngStream streams). Every
 * kernel launch covers all blobs; each blob has its own CG scalars and stops its solve
 * on its own convergence test; stokes_drag (may be NULL without calc_stokes) is per blob as for
 * ffea_gpu_blob_create. Single-blob calls act on blob 0; the ffea_gpu_batch_*
 * calls take the blob index. A step fails if any blob fails (the message names it). */
int ffea_gpu_batch_create(const ffea_gpu_params *params, int num_blobs, const int *num_nodes,
                          const int *num_elements, const double *const *pos, const int *const *top,
                          const double *const *A, const double *const *B, const double *const *G,
                          const double *const *E, const int *num_pinned, const int *const *pinned,
                          const int *num_ctf, const int *const *ctf_node,
                          const double *const *ctf_force, const double *const *stokes_drag,
                          const uint32_t *rng_states, ffea_gpu_blob **out);

/* Builds the device tables of ffea_gpu_batch_create's arguments on the host, without a
 * device, and checks that every index the kernels form from them stays inside the
 * buffer it addresses (the kernels do not check). Fills *info (may be NULL; fields that
 * need a device are 0). ffea_gpu_batch_create runs the same check before any upload.
 * No counterpart in the reference. */
int ffea_gpu_check_layout(const ffea_gpu_params *params, int num_blobs, const int *num_nodes,
                          const int *num_elements, const double *const *pos, const int *const *top,
                          const double *const *A, const double *const *B, const double *const *G,
                          const double *const *E, const int *num_pinned, const int *const *pinned,
                          const int *num_ctf, const int *const *ctf_node,
                          const double *const *ctf_force, const double *const *stokes_drag,
                          const uint32_t *rng_states, ffea_gpu_blob_info *info);

/* The mass-less viscosity solver on its own, for a caller whose Blob keeps its CPU element
 * loop and hands the solve to the GPU through the reference's Solver interface
 * (Solver.h:31-52; Blob creates its solver by enum, Blob.cpp:228-247, and calls
 * solver->solve(force) in Blob::aggregate_forces_and_solve, Blob.cpp:3243).
 *   ffea_gpu_solver_create   NoMassCGSolver::init (NoMassCGSolver.cpp:56-153): the mesh
 *                            (pos, top as for ffea_gpu_blob_create), the per-element A, B,
 *                            the pinned nodes and, with params->calc_stokes, the per-node
 *                            drag. The moduli and the noise belong to the caller's element
 *                            loop and are not taken. The handle is an ffea_gpu_blob for
 *                            ffea_gpu_blob_destroy, _last_error, _get_info,
 *                            _get_cg_iterations, _set_preconditioner, _autotune and
 *                            _set_timing; it refuses the step, force, solver and adjoint calls.
 *   ffea_gpu_solver_solve    NoMassCGSolver::solve (NoMassCGSolver.cpp:156-188), in place: x
 *                            [num_nodes][3] holds the right-hand side (the force vector after
 *                            linearise_force and pin zeroing) on entry and the solution on
 *                            return. pos [num_nodes][3] are the node positions the caller's
 *                            element loop used for this step: the viscosity matrix is built at
 *                            them (the element matrices, tetra_element_linear.cpp:222-327, and
 *                            SparseMatrixFixedPattern::build, SparseMatrixFixedPattern.cpp:86-94).
 *                            Jacobi-PCG from x = 0 with the reference's stopping test (or the
 *                            multigrid preconditioner when set); FFEA_GPU_ERR_CG_NO_CONVERGE
 *                            carries the reference's message. Pinned rows are identity rows.
 *                            Midside rows are identity rows that linearise_force leaves zero:
 *                            a nonzero midside entry is refused, and they return zero.
 *   ffea_gpu_solver_apply_matrix  out = V in [num_nodes][3] with the matrix solve builds at
 *                            pos, midside and pinned rows as the identity rows of the
 *                            reference's matrix (with the drag of an unpinned node under
 *                            calc_stokes). The reference's NoMassCGSolver::apply_matrix is an
 *                            empty stub (NoMassCGSolver.h:103). */
int ffea_gpu_solver_create(const ffea_gpu_params *params, int num_nodes, int num_elements, const double *pos,
                           const int *top, const double *A, const double *B, int num_pinned, const int *pinned,
                           const double *stokes_drag, ffea_gpu_blob **out);
int ffea_gpu_solver_solve(ffea_gpu_blob *solver, const double *pos, double *x);
int ffea_gpu_solver_apply_matrix(ffea_gpu_blob *solver, const double *pos, const double *in, double *out);

/* releases device memory and the stream; NULL is accepted */
void ffea_gpu_blob_destroy(ffea_gpu_blob *blob);

/* Blob::update_internal_forces (Blob.cpp:400-491): ctforces + per-element elastic and
 * fluctuating stress -> element nodal forces; counts inverted elements. */
int ffea_gpu_blob_update_internal_forces(ffea_gpu_blob *blob);

/* Blob::update_positions (Blob.cpp:493-508): gather, linearise_force, pins, viscosity
 * matrix build, Jacobi-PCG solve, Euler update, midside re-linearisation. */
int ffea_gpu_blob_update_positions(ffea_gpu_blob *blob);

/* nsteps x { zero forces; update_internal_forces; update_positions } — the blob's part
 * of World::run (World.cpp:1786-2059). Stops at the first failing step. */
int ffea_gpu_blob_step(ffea_gpu_blob *blob, int nsteps);

/* The reference's solvers behind its Solver interface (Blob.cpp:228-243,
 * FFEA_return_codes.h:35-38) that this library runs */
#define FFEA_GPU_SOLVER_ITERATIVE 1
#define FFEA_GPU_SOLVER_MASSLUMPED 2
#define FFEA_GPU_SOLVER_NOMASS_CG 3

/* Chooses the blob's solver (single-blob handles). NOMASS_CG (the default) is the
 * mass-less viscous solve. ITERATIVE is ConjugateGradientSolver
 * (ConjugateGradientSolver.cpp:51-335): the linear tets' consistent mass matrix (.1 rho
 * vol_0 on a corner's own entry, .05 rho vol_0 between two corners of an element, element
 * order; pinned rows identity), built once, and a Jacobi-PCG solve of M a = f each step
 * with the mass-less CG's stopping rule, then the same inertial update as MASSLUMPED.
 * MASSLUMPED is MassLumpedSolver (MassLumpedSolver.cpp:27-73)
 * with the inertial time step it implies: each corner node's mass is the sum over its
 * elements, in element order, of rho vol_0 / 4 (rho [num_elements] in meso units, .mat
 * order); midside and pinned nodes have mass 1. A step then forms the element forces with
 * the viscous term V v_e from the node velocities (Blob.cpp:460-466), subtracts drag v from
 * every node under calc_stokes (Blob.cpp:3172-3196), and updates a = f / m, v += a dt,
 * x += v dt (Blob.cpp:3249-3280): no linear solve. Switching sets the velocities to zero;
 * rho may be NULL for NOMASS_CG. The multigrid preconditioner, graphs and the adjoint are
 * for NOMASS_CG only. A refused call (unknown solver, a batch handle, missing or
 * non-positive densities, multigrid on) returns an error and leaves the handle as it was. */
int ffea_gpu_blob_set_solver(ffea_gpu_blob *blob, int solver, const double *rho);

/* Blob::make_measurements' kinetic energy (Blob.cpp:1076-1118): half the sum over elements
 * of v_e . (M_e v_e) with the element mass matrix (tetra_element_linear.cpp:524-542); 0
 * for the mass-less solver */
int ffea_gpu_blob_kinetic_energy(ffea_gpu_blob *blob, double *energy);

/* Blob::add_force_to_node (Blob.cpp:1770-1774) for the next step only: force[i] (meso
 * units) is added to node[i], in list order and across calls in call order, before the
 * constant forces of the next step's force formation (World.cpp:1862-1910: forces are
 * zeroed, the springs are applied, then update_internal_forces adds the constant forces).
 * The nodes must be among the external_nodes the handle was created with. */
int ffea_gpu_blob_add_node_forces(ffea_gpu_blob *blob, int num, const int *node, const double *force);

/* Blob::get_node (Blob.cpp:1762-1764): positions [num][3] of the given nodes (midside
 * nodes at their edge midpoints) */
int ffea_gpu_blob_get_node_positions(ffea_gpu_blob *blob, int num, const int *node, double *pos);

/* node state, [num_nodes][3] (midside nodes at their edge midpoints) */
int ffea_gpu_blob_get_positions(ffea_gpu_blob *blob, double *pos);
int ffea_gpu_blob_set_positions(ffea_gpu_blob *blob, const double *pos);
int ffea_gpu_blob_get_velocities(ffea_gpu_blob *blob, double *vel);
/* node velocities [num_nodes][3] (corner nodes; the inertial solver's state, read back by a
 * restart, Blob.cpp:1002-1010) */
int ffea_gpu_blob_set_velocities(ffea_gpu_blob *blob, const double *vel);

/* solver right-hand side of the last update_positions ([num_nodes][3], after gather,
 * linearise_force and pin zeroing — the `force` vector Blob.cpp:3243 hands to solve) */
int ffea_gpu_blob_get_rhs(ffea_gpu_blob *blob, double *rhs);

/* element nodal forces of the last update_internal_forces, [num_elements][4][3]
 * (tetra_element_linear::node_force[0..3]) */
int ffea_gpu_blob_get_element_forces(ffea_gpu_blob *blob, double *force);

/* CG iterations of the last solve and the total since creation */
int ffea_gpu_blob_get_cg_iterations(ffea_gpu_blob *blob, int *last, long long *total);

/* noise stream state (RngStream::GetState / SetSeed) */
int ffea_gpu_blob_get_rng_state(ffea_gpu_blob *blob, uint32_t state[6]);
int ffea_gpu_blob_set_rng_state(ffea_gpu_blob *blob, const uint32_t state[6]);

/* Noise stream shared with other blobs: `state` is this blob's stream state for its next
 * element pass and every pass then moves the stream `draws_per_step` draws on (7 per
 * element of every blob that draws from the same stream in a step; the default is 7
 * per element of this blob). World::run draws blob after blob from rng[0] in one
 * thread (World.cpp:2031-2038), so blob b starts 7 * (elements of blobs 0..b-1) on. */
int ffea_gpu_blob_set_rng_stream(ffea_gpu_blob *blob, const uint32_t state[6], long long draws_per_step);

/* With calc_stokes and calc_noise a step draws 7 per element, then 3 per node. World::run
 * takes the element draws of all blobs first and the node draws after them
 * (World.cpp:2034-2058), so in a multi-blob world blob b's node draws start `offset` draws
 * after its own element draws (default: 7 x its elements). */
int ffea_gpu_batch_set_node_draw_offset(ffea_gpu_blob *blob, int index, long long offset);

/* strain energy of every blob of the handle at the current positions, energy[num_blobs]
 * (Blob::calculate_strain_energy, Blob.cpp:1052-1064; the measurement file's StrainEnergy) */
int ffea_gpu_batch_strain_energy(ffea_gpu_blob *blob, double *energy);

/* ffea_gpu_batch_strain_energy that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Its per-element energies and per-blob sums live in
 * a fixed 64 MiB scratch array per GPU, held until the process ends, so the handle's
 * elements plus blobs may not exceed 8,388,608 (FFEA_GPU_ERR_INVALID otherwise; the
 * 6.0M-tet blob fits). Results equal ffea_gpu_batch_strain_energy's bit for bit. Every call on
 * one GPU uses that GPU's one array: calls that overlap on one GPU (handles of the same device
 * used from different threads at once) race and must not be made. */
int ffea_gpu_batch_strain_energy_noalloc(ffea_gpu_blob *blob, double *energy);

/* state after `ndraws` draws from `state` (exact jump-ahead; host only) */
int ffea_gpu_advance_rng_state(const uint32_t state[6], long long ndraws, uint32_t out[6]);

/* initial state of the stream-th RngStream created after RngStream::SetPackageSeed(
 * package_seed): streams start 2^127 draws apart (RngStream.cpp:327-345, 380-390);
 * World gives thread t stream t (World.cpp:278-295). Host only. */
int ffea_gpu_rng_stream_state(const uint32_t package_seed[6], long long stream, uint32_t out[6]);

/* per-blob forms of the node-state, force and stream calls above (blob = 0..num_blobs-1;
 * arrays sized for that blob). ffea_gpu_batch_set_rng_stream with draws_per_step < 0
 * keeps the blob's stride. ffea_gpu_batch_get_cg_iterations fills [num_blobs] arrays. */
int ffea_gpu_batch_get_positions(ffea_gpu_blob *blob, int index, double *pos);
int ffea_gpu_batch_set_positions(ffea_gpu_blob *blob, int index, const double *pos);
int ffea_gpu_batch_get_velocities(ffea_gpu_blob *blob, int index, double *vel);
int ffea_gpu_batch_get_rhs(ffea_gpu_blob *blob, int index, double *rhs);
int ffea_gpu_batch_get_element_forces(ffea_gpu_blob *blob, int index, double *force);
int ffea_gpu_batch_get_cg_iterations(ffea_gpu_blob *blob, int *last, long long *total);
int ffea_gpu_batch_get_rng_state(ffea_gpu_blob *blob, int index, uint32_t state[6]);
int ffea_gpu_batch_set_rng_stream(ffea_gpu_blob *blob, int index, const uint32_t state[6],
                                  long long draws_per_step);

int ffea_gpu_blob_get_info(ffea_gpu_blob *blob, ffea_gpu_blob_info *info);

/* Differentiation of a blob trajectory (single-blob handles, calc_noise = 0, some nodes
 * pinned, no surface, rotational or external forces). No counterpart in the reference.
 *   ffea_gpu_blob_set_material   replaces the per-element A, B, G, E (meso units, .top order);
 *                                a run kept by ffea_gpu_blob_run_checkpointed before it is
 *                                dropped (it was made with the earlier material), as it is
 *                                by ffea_gpu_blob_set_solver with an inertial solver
 *   ffea_gpu_blob_run_checkpointed  nsteps steps of ffea_gpu_blob_step from the current
 *                                state, keeping the positions of every interval-th step on
 *                                the device
 *   ffea_gpu_blob_adjoint        reverse of that run (discrete adjoint: each step's
 *                                x' = x + dt K(x)^-1 P f(x) reversed exactly, one CG solve
 *                                with K per step, segments recomputed from the checkpoints):
 *                                from dL/dpos at the end of the run ([num_nodes][3], .node
 *                                order) it returns dL/dA, dL/dB, dL/dG, dL/dE ([num_elements]
 *                                each, any may be NULL) and dL/dpos at the start (may be
 *                                NULL; zero on midside nodes, which are derived from the
 *                                corners). The blob is left in the state at the end of the run. */
int ffea_gpu_blob_set_material(ffea_gpu_blob *blob, const double *A, const double *B, const double *G,
                               const double *E);
int ffea_gpu_blob_run_checkpointed(ffea_gpu_blob *blob, int nsteps, int interval);
int ffea_gpu_blob_adjoint(ffea_gpu_blob *blob, const double *dL_dpos, double *dL_dA, double *dL_dB,
                          double *dL_dG, double *dL_dE, double *dL_dpos0);

/* FP64 refinement of the FP32-mixed form's solves (no counterpart in the reference, whose
 * solve is FP64). The FP32 viscosity matrix's rounding is amplified by the solve in the
 * smooth, slowly relaxing part of a solution (a static load on a pinned blob; the adjoint's
 * solves). With tolerance > 0, after each converged mass-less solve V v = b (the step's,
 * ffea_gpu_solver_solve's and the adjoint's) the true residual b - V v is formed in FP64
 * (matrix-free: shape derivatives and volumes recomputed in FP64 from the positions, FP64
 * viscosities), an FP32 correction solve with the CG's own kernels and preconditioner is
 * added, and this repeats until |b - V v| <= tolerance |b| (at most 30 correction solves;
 * FFEA_GPU_ERR_INVALID otherwise). The correction solves' iterations count in the step's
 * CG iterations; refined steps run eagerly (no graph). 0 turns it off (the default).
 * FP32-mixed single-blob mass-less handles with a nonsingular matrix (pinned nodes or
 * Stokes drag), without surface, rotational or external forces.
 * ffea_gpu_blob_get_refinement: correction solves of the last solve and in all. */
int ffea_gpu_blob_set_refinement(ffea_gpu_blob *blob, double tolerance);
int ffea_gpu_blob_get_refinement(ffea_gpu_blob *blob, int *last_sweeps, long long *total_sweeps);

/* Launch shape. Totals are summed per reduction chunk in a fixed order and every row and
 * block entry in a fixed order, so the shape never changes results. The knobs: the number
 * of workgroups of the row kernels (rounded to whole chunks per workgroup and a multiple
 * of 8, one share per XCD; ffea_gpu_blob_set_row_blocks), the SpMV's row unroll (rows of
 * up to 16 or up to 8 blocks issue all their block loads together) and the assembly's
 * mapping: 1, 4 or 16 lanes per row reading each block source's shape record from memory,
 * or -1: one wavefront per group of 8 rows that first copies the records of its rows'
 * elements to LDS. ffea_gpu_blob_autotune times `iterations` CG iterations of
 * the current matrix at each candidate workgroup count, then both unrolls at the fastest
 * count, then three assemblies with each mapping, keeps the fastest of each and returns
 * the workgroup count in *chosen (may be NULL); it leaves positions, velocities and noise
 * streams untouched. ffea_gpu_blob_get_launch_shape reads the three knobs;
 * ffea_gpu_blob_set_launch_shape sets them (0: unchanged). No counterpart in the
 * reference. */
int ffea_gpu_blob_set_row_blocks(ffea_gpu_blob *blob, int blocks);
int ffea_gpu_blob_autotune(ffea_gpu_blob *blob, int iterations, int *chosen);
int ffea_gpu_blob_get_launch_shape(ffea_gpu_blob *blob, int *row_blocks, int *spmv_unroll, int *assembly_lanes);
int ffea_gpu_blob_set_launch_shape(ffea_gpu_blob *blob, int row_blocks, int spmv_unroll, int assembly_lanes);

/* enable != 0: ffea_gpu_blob_step runs each step as one HIP graph launch (element pass,
 * stream advance, assembly, gather, CG iterations, update), continuing eagerly only if a
 * blob needs more CG iterations than the graph holds. The graph holds `iterations` CG
 * iterations, or (iterations <= 0) a little over the last step's count, recaptured when
 * that count moves. Results are identical to eager steps. Per-phase timing does not see
 * graph steps. No counterpart in the reference. */
int ffea_gpu_blob_use_graph(ffea_gpu_blob *blob, int enable, int iterations);

/* Preconditioner of the viscosity-matrix CG. JACOBI is the reference's
 * (NoMassCGSolver.cpp:268-281, parallel_apply_preconditioner) and the default.
 * MULTIGRID has no counterpart in the reference: a smoothed-aggregation multigrid V-cycle
 * whose coarse unknowns are the 3 translations and 3 rotations of node aggregates. The
 * same smoothing runs before and after every coarse correction: a damped block-Jacobi
 * step (the 3x3 diagonal blocks) on the finest level; on level 1 a damped-Jacobi step, or
 * a degree-2 Chebyshev polynomial for a single blob of more than 10^4 level-1 aggregates;
 * four damped-Jacobi steps on any level between level 1 and the coarsest, which is solved
 * with its dense inverse. Its
 * hierarchy is built when the preconditioner is set, from the viscosity matrix at the
 * current positions (an element pass and an assembly that leave positions, velocities
 * and noise streams untouched); every later step applies its own matrix on the finest
 * level. The CG keeps the reference's start (x = 0) and stopping test, so each solve
 * meets the same tolerance in far fewer iterations; the solution differs from the
 * Jacobi-preconditioned one within that tolerance. A free blob's solution is returned in
 * the rigid-mode gauge of Jacobi-PCG (Z^T D x = 0, Z the rigid modes, D the diagonal).
 * Single-blob handles with more than 1024 corner nodes; the adjoint's solves
 * (ffea_gpu_blob_adjoint) use it too. */
typedef enum {
    FFEA_GPU_PRECOND_JACOBI = 0,
    FFEA_GPU_PRECOND_MULTIGRID = 1
} ffea_gpu_preconditioner;

#define FFEA_GPU_MG_MAX_LEVELS 8
typedef struct {
    int num_levels;                                 /* 0: Jacobi; else levels including the coarsest */
    int rows[FFEA_GPU_MG_MAX_LEVELS];               /* block rows (level 0: corner nodes; then aggregates) */
    int block[FFEA_GPU_MG_MAX_LEVELS];              /* unknowns per block row: 3 on level 0, 6 above */
    long long blocks[FFEA_GPU_MG_MAX_LEVELS];       /* blocks of each level's operator */
    long long transfer_blocks[FFEA_GPU_MG_MAX_LEVELS]; /* blocks of the prolongator from level l to l+1 */
    double omega[FFEA_GPU_MG_MAX_LEVELS];           /* smoothing weights */
    double lmax[FFEA_GPU_MG_MAX_LEVELS];            /* estimates of the largest eigenvalue of D^-1 A */
    int coarsest_unknowns;                          /* the densely inverted level */
    double setup_seconds;
} ffea_gpu_multigrid_info;

int ffea_gpu_blob_set_preconditioner(ffea_gpu_blob *blob, int kind);
int ffea_gpu_blob_get_multigrid_info(ffea_gpu_blob *blob, ffea_gpu_multigrid_info *info);

/* Host-only (no device): the multigrid hierarchy of a corner-node matrix given in 3x3
 * block CSR (num_rows block rows, rowptr [num_rows+1], col [nnzb], val [nnzb][9]
 * row-major blocks; pinned rows hold an identity block only), positions [num_rows][3],
 * pinned [num_rows] flags (may be NULL), free_blob = 1 for a singular (unpinned, no
 * drag) matrix. Fills *info; where the pointers are not NULL also copies the first
 * prolongator P0 (rows num_rows, 3x6 blocks: p_rowptr [num_rows+1], p_col
 * [transfer_blocks[0]], p_val [transfer_blocks[0]][18]) and the first coarse operator
 * A1 = P0^T A P0 (6x6 blocks: a_rowptr [rows[1]+1], a_col [blocks[1]], a_val
 * [blocks[1]][36]). */
int ffea_gpu_multigrid_inspect(int num_rows, const int *rowptr, const int *col, const double *val,
                               const double *pos, const unsigned char *pinned, int free_blob,
                               ffea_gpu_multigrid_info *info, int *p_rowptr, int *p_col, double *p_val,
                               int *a_rowptr, int *a_col, double *a_val);

/* Kernel timing with HIP events recorded on the handle's stream around each launch of
 * the phases selected by phase_mask (bit i = phase i; 0 turns timing off).
 * Phases: 0 element forces, 1 matrix assembly, 2 gather + CG start, 3 CG SpMV step,
 * 4 CG vector step, 5 position update. Times in ms, summed since the last set_timing;
 * launches of phases 3 and 4 count CG iterations done (launches queued after
 * convergence return at once and are not counted). */
#define FFEA_GPU_NUM_PHASES 6
int ffea_gpu_blob_set_timing(ffea_gpu_blob *blob, unsigned phase_mask);
int ffea_gpu_blob_get_timing(ffea_gpu_blob *blob, double ms[FFEA_GPU_NUM_PHASES],
                             long long launches[FFEA_GPU_NUM_PHASES]);

/* Achievable-bandwidth probes (no counterpart in the reference): mean time over `iters`
 * launches of a streaming kernel that copies `bytes` device to device (`bytes` in and
 * `bytes` out; also the calibration kernel of the hardware byte counters), or that reads
 * `bytes` once (`ffea_gpu_read_probe`). */
int ffea_gpu_copy_probe(int device, long long bytes, int iters, double *ms_per_copy);
int ffea_gpu_read_probe(int device, long long bytes, int iters, double *ms_per_read);

/* waits for all work queued on the handle's stream */
int ffea_gpu_blob_synchronize(ffea_gpu_blob *blob);

const char *ffea_gpu_blob_last_error(const ffea_gpu_blob *blob);
const char *ffea_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
