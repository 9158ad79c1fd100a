/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, single-threaded, FP64 restatement of FFEA's blob time step (the
 * reference's `Blob::update_internal_forces` + `Blob::update_positions` with the
 * `NoMassCGSolver`), written from the reference's algorithm and used as the
 * checker for the HIP path. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product library never links or calls it.
 *
 * Every function names the reference file:line whose arithmetic it follows.
 * Pinning of each part (known answers, closed forms, the reference's own
 * translation units built under oracle/_ref) is listed in the header of
 * ffea_oracle.c.
 */
#ifndef FFEA_ORACLE_H
#define FFEA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- MRG32k3a streams (RngStream.cpp) ---------------- */
typedef struct {
    double Cg[6], Bg[6], Ig[6];
    int anti, inc_prec;
} orc_rng;

/* package seed holder: the "next stream" state (RngStream.cpp nextSeed) */
typedef struct {
    double next[6];
} orc_rng_package;

void   orc_rng_package_default(orc_rng_package *pk);
int    orc_rng_package_set_seed(orc_rng_package *pk, const uint32_t seed[6]);
void   orc_rng_new_stream(orc_rng *g, orc_rng_package *pk);
int    orc_rng_set_seed(orc_rng *g, const uint32_t seed[6]);
void   orc_rng_get_state(const orc_rng *g, uint32_t seed[6]);
double orc_rng_u01(orc_rng *g);
double orc_rng_rand_u01(orc_rng *g);
int    orc_rng_rand_int(orc_rng *g, int low, int high);
void   orc_rng_fill_u01(orc_rng *g, int n, double *out);
void   orc_rng_advance_state(orc_rng *g, int32_t e, int32_t c);
void   orc_rng_reset_start_stream(orc_rng *g);
void   orc_rng_reset_start_substream(orc_rng *g);
void   orc_rng_reset_next_substream(orc_rng *g);
void   orc_rng_set_antithetic(orc_rng *g, int a);
void   orc_rng_increased_precis(orc_rng *g, int incp);
/* World.cpp:272-276 seed derivation from an integer rng_seed (glibc rand) */
void   orc_sixseed_from_rng_seed(int rng_seed, uint32_t sixseed[6]);

/* ---------------- blob time step ---------------- */
typedef struct {
    double dt;          /* meso time units */
    double kT;          /* meso energy units */
    double epsilon2;    /* squared CG tolerance (SimulationParams.cpp:259-265) */
    int max_iterations_cg;
    int calc_noise;
    int calc_stokes;
} orc_params;

typedef struct orc_blob orc_blob;

/* Builds a blob from the loaded node/topology/material arrays, in the order of
 * Blob::init (Blob.cpp:161-247): linearise, rest state, connectivity, solver
 * init. `top` is [ne][10] (.top order), A/B/G/E are per element in meso units,
 * `stokes_drag` is per node (may be NULL when calc_stokes == 0), ctforces are
 * per-entry (node, 3 meso forces) applied in list order. rng_state is stream 0
 * (the single-thread stream) as 6 uint32. Returns NULL on bad input. */
orc_blob *orc_blob_create(int nt, int ne, const double *pos, const int *top,
                          const double *A, const double *B, const double *G, const double *E,
                          int npin, const int *pinned,
                          int nctf, const int *ctf_node, const double *ctf_force,
                          const double *stokes_drag,
                          const orc_params *params, const uint32_t rng_state[6]);
void orc_blob_destroy(orc_blob *b);

/* linear surface constant forces (Blob::load_ctforces, Blob.cpp:2730-2806): faces as
 * [nfaces][3] node indices (.surf order, Face::n[0..2]); set s applies set_force[s]
 * (meso units) over faces set_faces[set_ptr[s] .. set_ptr[s+1]). Returns 0 or 1 (bad input). */
int orc_blob_set_surface_forces(orc_blob *b, int nfaces, const int *face_nodes, int nsets, const double *set_force,
                                const int *set_ptr, const int *set_faces);

/* rotational constant forces (Blob::load_ctforces, Blob.cpp:2650-2726), axes given by a
 * point and a direction ('p' forces; Blob.cpp:1697-1706 leaves node-pair axes unwritten):
 * entry i pulls node[i] round the axis through axis[6i..6i+2] (meso units) along the unit
 * direction axis[6i+3..6i+5] with force[i] (torque[i] = 0, 'f': |f| = force |r|) or
 * constant torque (torque[i] = 1, 't': |f| = force). Returns 0 or 1 (bad input). */
int orc_blob_set_rotational_forces(orc_blob *b, int num, const int *node, const double *force, const double *axis,
                                   const int *torque);
/* the force of one such entry on a node at x (Blob.cpp:1693-1714) */
void orc_rotational_force(const double *x, const double *axis, double F, int torque, double *f);

/* MassLumpedSolver (MassLumpedSolver.cpp:27-73) in place of the mass-less CG: rho [ne]
 * densities (meso units); velocities are set to zero. Returns 0 or 1 (bad density). */
int orc_blob_set_mass_lumped(orc_blob *b, const double *rho);
/* ConjugateGradientSolver (ConjugateGradientSolver.cpp:51-335) in place of the mass-less
 * CG: the consistent mass matrix of the linear tets, solved by Jacobi-PCG each step;
 * velocities are set to zero. Returns 0 or 1. */
int orc_blob_set_mass_consistent(orc_blob *b, const double *rho);
/* the kinetic energy of Blob::make_measurements (Blob.cpp:1076-1118), 0 without mass */
double orc_blob_kinetic_energy(const orc_blob *b);

/* Blob::add_force_to_node (Blob.cpp:1770-1774) for the next orc_blob_step only, applied
 * after its forces are zeroed and before the constant forces. Returns 0 or 1 (bad node). */
int orc_blob_add_node_forces(orc_blob *b, int num, const int *node, const double *force);
/* one spring of World::apply_springs (World.cpp:4063-4108): forces on its two nodes at x0
 * and x1 for constant k and length l (meso units); returns 1 for coincident nodes (the
 * reference's normalize throws there), else 0 */
int orc_spring_forces(const double *x0, const double *x1, double k, double l, double *f0, double *f1);
/* its energy 0.5 k (|x0 - x1| - l)^2 (World.cpp:4645-4648) */
double orc_spring_energy(const double *x0, const double *x1, double k, double l);

/* Blob.cpp:400-491 ; returns the number of inverted elements */
int  orc_blob_update_internal_forces(orc_blob *b);
/* Blob.cpp:493-508 ; returns 0, or 1 when CG does not converge */
int  orc_blob_update_positions(orc_blob *b);
/* one World::run step for one blob: zero forces, internal forces, positions.
 * Returns 0, 1 (inversion) or 2 (CG non-convergence). */
int  orc_blob_step(orc_blob *b);

int  orc_blob_num_nodes(const orc_blob *b);
int  orc_blob_num_elements(const orc_blob *b);
void orc_blob_get_positions(const orc_blob *b, double *out);      /* [nt][3] */
void orc_blob_set_positions(orc_blob *b, const double *in);       /* [nt][3] */
void orc_blob_get_velocities(const orc_blob *b, double *out);     /* [nt][3] */
void orc_blob_set_velocities(orc_blob *b, const double *in);
/* the solver right-hand side of the last update_positions (after gather,
 * stokes noise, linearise_force and pin zeroing) */
void orc_blob_get_rhs(const orc_blob *b, double *out);            /* [nt][3] */
/* per-element nodal forces of the last update_internal_forces [ne][4][3] */
void orc_blob_get_element_forces(const orc_blob *b, double *out);
int  orc_blob_last_cg_iterations(const orc_blob *b);
void orc_blob_get_rng_state(const orc_blob *b, uint32_t s[6]);
void orc_blob_set_rng_state(orc_blob *b, const uint32_t s[6]);
/* Blob.cpp:1052-1064 (uses the current positions) */
double orc_blob_strain_energy(orc_blob *b);
/* per-element rest state: Jinv0 [ne][9], vol0 [ne] */
void orc_blob_get_rest_state(const orc_blob *b, double *jinv0, double *vol0);

/* Fixed-pattern viscosity matrix over 3*nt scalar rows (NoMassCGSolver.cpp:56-153).
 * orc_blob_build_matrix recomputes every element's viscosity matrix at the current
 * positions and sums the sources (SparseMatrixFixedPattern.cpp:86-94, 305-312). */
void orc_blob_build_matrix(orc_blob *b);
int  orc_blob_matrix_nnz(const orc_blob *b);
void orc_blob_get_matrix(const orc_blob *b, int *rowptr, int *col, double *val);
void orc_blob_get_inverse_diagonal(const orc_blob *b, double *inv_diag);
/* SparseMatrixFixedPattern.cpp:148-187 on [nt][3] vectors */
void orc_blob_apply_matrix(const orc_blob *b, const double *in, double *out);
/* NoMassCGSolver.cpp:156-188 on an explicit right-hand side (in/out [nt][3]);
 * uses the matrix from the last build. Returns iterations, or -1 when not converged. */
int  orc_blob_solve(orc_blob *b, double *x);

/* element-level helpers for closed-form checks (tetra_element_linear.cpp) */
/* 4 node positions [4][3], rest Jinv0 [9], vol0, G, E -> 4x3 node forces (no noise) */
int orc_element_elastic_force(const double *x4, const double *jinv0, double vol0,
                              double G, double E, double *force12);
/* 4 node positions -> dpsi[12], vol; returns det */
double orc_element_shape(const double *x4, double *dpsi, double *vol);
/* dpsi, vol, A, B -> 12x12 viscosity matrix (row-major, reference index 4*i+a) */
void orc_element_viscosity_matrix(const double *dpsi, double vol, double A, double B, double *V144);
/* 7 uniforms in draw order -> symmetric fluctuating stress (3x3 row-major) added to zero */
void orc_element_noise_stress(double vol, double A, double B, double kT, double dt,
                              const double *u7, double *stress9);

#ifdef __cplusplus
}
#endif
#endif
