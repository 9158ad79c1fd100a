/*
 * ORACLE — test infrastructure only (see ffea_oracle.h).
 *
 * FP64, single-threaded restatement of FFEA's blob step, in the reference's
 * summation orders (the reference run with OMP_NUM_THREADS=1). Citations are
 * into the reference tree.
 *
 * What pins each part (tests/test_oracle_*.py):
 *   RNG (MRG32k3a streams, substreams, AdvanceState, antithetic,
 *     increased precision)       known answer of the reference's
 *                                tests/consistency/rngstream/testRngStream.cpp:129
 *                                (39.697547445251) replayed on this code, and
 *                                draw-by-draw equality with the reference's own
 *                                RngStream.cpp built under oracle/_ref
 *   seed derivation              fixture of sixseed values (glibc rand)
 *   jacobian / dpsi / volume     closed forms: dpsi rows = rows of J^-1,
 *                                sum_a grad psi_a = 0, vol = |det|/6, regular
 *                                cube corner tets; J^-1_0 equal to the
 *                                reference's mat3_invert (oracle/_ref)
 *   elastic stress and forces    F = I gives zero force; uniform dilation closed
 *                                form; rigid rotation gives zero force; zero net
 *                                force and torque; central differences of the
 *                                strain energy (Blob.cpp:1052-1064) = -force
 *   viscosity matrix             symmetric, positive semi-definite, kills rigid
 *                                modes, equals vol * sigma(v) . grad psi with
 *                                sigma = A (grad v + grad v^T) + B div(v) I
 *   noise stress                 closed-form variances; mean zero
 *   fixed-pattern matrix,
 *     source sums, SpMV, Jacobi  bit-equality with the reference's own
 *                                SparsityPattern / SparseMatrixFixedPattern /
 *                                SparseMatrixTypes built under oracle/_ref, fed the
 *                                same element matrices
 *   PCG                          dense direct solve (numpy) on small meshes
 *   gather, linearise_force,
 *     pins, ctforces, Euler      conservation and closed-form checks
 */
#include "ffea_oracle.h"

/* OMP(...) parallelises the step's loops in the timing build only (libffea_oracle_omp.so,
 * -fopenmp with the reference's Release flags, as the reference's OpenMP loops,
 * Blob.cpp:419-481, NoMassCGSolver.cpp:199-281): that build times the CPU baseline and
 * nothing else. The default build has no OpenMP and runs every loop in order. */
#ifdef _OPENMP
#include <omp.h>
#define OMP(x) _Pragma(#x)
#else
#define OMP(x)
#endif

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ======================================================================
 * MRG32k3a (RngStream.cpp; L'Ecuyer, Simard, Chen, Kelton, Oper. Res. 50(6), 2002)
 * The state is kept as integer-valued doubles as in the reference; all modular
 * arithmetic here is done exactly in 128-bit integers.
 * ====================================================================== */
#define ORC_M1 4294967087LL
#define ORC_M2 4294944443LL
#define ORC_A12 1403580LL
#define ORC_A13N 810728LL
#define ORC_A21 527612LL
#define ORC_A23N 1370589LL

typedef long long i64;
typedef __int128 i128;

static i64 mod_canon(i128 v, i64 m) {
    i128 r = v % m;
    if (r < 0) r += m;
    return (i64)r;
}

/* C = A*B mod m (3x3) */
static void mat_mul_mod(const i64 A[3][3], const i64 B[3][3], i64 C[3][3], i64 m) {
    i64 T[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            i128 s = 0;
            for (int k = 0; k < 3; ++k) s += (i128)A[i][k] * B[k][j];
            T[i][j] = mod_canon(s, m);
        }
    memcpy(C, T, sizeof(T));
}

static void mat_vec_mod(const i64 A[3][3], const double *s, double *out, i64 m) {
    i64 v[3];
    for (int i = 0; i < 3; ++i) {
        i128 acc = 0;
        for (int k = 0; k < 3; ++k) acc += (i128)A[i][k] * (i64)s[k];
        v[i] = mod_canon(acc, m);
    }
    for (int i = 0; i < 3; ++i) out[i] = (double)v[i];
}

/* one-step transition matrices of the two components (RngStream.cpp:76-86) */
static void comp_matrix(int comp, i64 A[3][3]) {
    memset(A, 0, 9 * sizeof(i64));
    A[0][1] = 1;
    A[1][2] = 1;
    if (comp == 0) {
        A[2][0] = -ORC_A13N;
        A[2][1] = ORC_A12;
    } else {
        A[2][0] = -ORC_A23N;
        A[2][2] = ORC_A21;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = mod_canon(A[i][j], comp == 0 ? ORC_M1 : ORC_M2);
}

static i64 inv_mod(i64 a, i64 m) {
    /* extended Euclid */
    i64 t = 0, nt = 1, r = m, nr = mod_canon(a, m);
    while (nr != 0) {
        i64 q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return mod_canon(t, m);
}

/* inverse of the one-step matrix: the recurrence run backwards */
static void comp_matrix_inv(int comp, i64 Ai[3][3]) {
    const i64 m = comp == 0 ? ORC_M1 : ORC_M2;
    memset(Ai, 0, 9 * sizeof(i64));
    /* new state (s1, s2, x) with x = c1*s1 + c2*s2 - c0*s0  =>  s0 = (c1*s1' + c2*s2'... ) */
    if (comp == 0) {
        /* (s0,s1,s2) -> (s1, s2, a12*s1 - a13n*s0); s0 = (a12*t0 - t2) / a13n */
        const i64 inv = inv_mod(ORC_A13N, m);
        Ai[0][0] = mod_canon((i128)ORC_A12 * inv, m);
        Ai[0][2] = mod_canon(-(i128)inv, m);
    } else {
        /* (s0,s1,s2) -> (s1, s2, a21*s2 - a23n*s0); s0 = (a21*t1 - t2) / a23n */
        const i64 inv = inv_mod(ORC_A23N, m);
        Ai[0][1] = mod_canon((i128)ORC_A21 * inv, m);
        Ai[0][2] = mod_canon(-(i128)inv, m);
    }
    Ai[1][0] = 1;
    Ai[2][1] = 1;
}

static void mat_two_pow(const i64 A[3][3], i64 B[3][3], i64 m, int e) {
    memcpy(B, A, 9 * sizeof(i64));
    for (int i = 0; i < e; ++i) mat_mul_mod(B, B, B, m);
}

static void mat_pow(const i64 A[3][3], i64 B[3][3], i64 m, i64 n) {
    i64 W[3][3];
    memcpy(W, A, sizeof(W));
    memset(B, 0, 9 * sizeof(i64));
    B[0][0] = B[1][1] = B[2][2] = 1;
    while (n > 0) {
        if (n & 1) mat_mul_mod(W, B, B, m);
        mat_mul_mod(W, W, W, m);
        n >>= 1;
    }
}

static int check_seed(const uint32_t s[6]) {
    for (int i = 0; i < 3; ++i)
        if ((i64)s[i] >= ORC_M1) return -1;
    for (int i = 3; i < 6; ++i)
        if ((i64)s[i] >= ORC_M2) return -1;
    if (s[0] == 0 && s[1] == 0 && s[2] == 0) return -1;
    if (s[3] == 0 && s[4] == 0 && s[5] == 0) return -1;
    return 0;
}

void orc_rng_package_default(orc_rng_package *pk) {
    for (int i = 0; i < 6; ++i) pk->next[i] = 12345.0; /* RngStream.cpp:328-331 */
}

int orc_rng_package_set_seed(orc_rng_package *pk, const uint32_t seed[6]) {
    if (check_seed(seed)) return 0;
    for (int i = 0; i < 6; ++i) pk->next[i] = seed[i];
    return 1;
}

/* RngStream::RngStream (RngStream.cpp:336-352): take the package state, then
 * move the package state 2^127 steps ahead */
void orc_rng_new_stream(orc_rng *g, orc_rng_package *pk) {
    g->anti = 0;
    g->inc_prec = 0;
    for (int i = 0; i < 6; ++i) g->Bg[i] = g->Cg[i] = g->Ig[i] = pk->next[i];
    i64 A[3][3], P[3][3];
    comp_matrix(0, A);
    mat_two_pow(A, P, ORC_M1, 127);
    mat_vec_mod(P, pk->next, pk->next, ORC_M1);
    comp_matrix(1, A);
    mat_two_pow(A, P, ORC_M2, 127);
    mat_vec_mod(P, pk->next + 3, pk->next + 3, ORC_M2);
}

int orc_rng_set_seed(orc_rng *g, const uint32_t seed[6]) {
    if (check_seed(seed)) return 0;
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i] = g->Ig[i] = seed[i];
    return 1;
}

void orc_rng_get_state(const orc_rng *g, uint32_t seed[6]) {
    for (int i = 0; i < 6; ++i) seed[i] = (uint32_t)g->Cg[i];
}

/* RngStream::U01 (RngStream.cpp:266-289) */
double orc_rng_u01(orc_rng *g) {
    const double norm = 1.0 / (4294967087.0 + 1.0);
    i64 p1 = mod_canon((i128)ORC_A12 * (i64)g->Cg[1] - (i128)ORC_A13N * (i64)g->Cg[0], ORC_M1);
    g->Cg[0] = g->Cg[1];
    g->Cg[1] = g->Cg[2];
    g->Cg[2] = (double)p1;
    i64 p2 = mod_canon((i128)ORC_A21 * (i64)g->Cg[5] - (i128)ORC_A23N * (i64)g->Cg[3], ORC_M2);
    g->Cg[3] = g->Cg[4];
    g->Cg[4] = g->Cg[5];
    g->Cg[5] = (double)p2;
    const double d1 = (double)p1, d2 = (double)p2;
    const double u = (d1 > d2) ? (d1 - d2) * norm : (d1 - d2 + 4294967087.0) * norm;
    return g->anti ? (1 - u) : u;
}

/* RngStream::U01d (RngStream.cpp:295-308) */
static double orc_rng_u01d(orc_rng *g) {
    const double fact = 5.9604644775390625e-8; /* 2^-24 */
    double u = orc_rng_u01(g);
    if (g->anti) {
        u += (orc_rng_u01(g) - 1.0) * fact;
        return (u < 0.0) ? u + 1.0 : u;
    }
    u += orc_rng_u01(g) * fact;
    return (u < 1.0) ? u : (u - 1.0);
}

double orc_rng_rand_u01(orc_rng *g) { return g->inc_prec ? orc_rng_u01d(g) : orc_rng_u01(g); }

int orc_rng_rand_int(orc_rng *g, int low, int high) {
    return low + (int)((high - low + 1.0) * orc_rng_rand_u01(g)); /* RngStream.cpp:521-524 */
}

void orc_rng_fill_u01(orc_rng *g, int n, double *out) {
    for (int i = 0; i < n; ++i) out[i] = orc_rng_rand_u01(g);
}

/* RngStream::AdvanceState (RngStream.cpp:408-435) */
void orc_rng_advance_state(orc_rng *g, int32_t e, int32_t c) {
    for (int comp = 0; comp < 2; ++comp) {
        const i64 m = comp == 0 ? ORC_M1 : ORC_M2;
        i64 A[3][3], Ai[3][3], Bm[3][3], Cm[3][3];
        comp_matrix(comp, A);
        comp_matrix_inv(comp, Ai);
        if (e > 0) mat_two_pow(A, Bm, m, e);
        else if (e < 0) mat_two_pow(Ai, Bm, m, -e);
        if (c >= 0) mat_pow(A, Cm, m, c);
        else mat_pow(Ai, Cm, m, -(i64)c);
        if (e) mat_mul_mod(Bm, Cm, Cm, m);
        mat_vec_mod(Cm, g->Cg + 3 * comp, g->Cg + 3 * comp, m);
    }
}

void orc_rng_reset_start_stream(orc_rng *g) {
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i] = g->Ig[i];
}

void orc_rng_reset_start_substream(orc_rng *g) {
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i];
}

/* substreams are 2^76 steps apart (RngStream.cpp:371-377) */
void orc_rng_reset_next_substream(orc_rng *g) {
    for (int comp = 0; comp < 2; ++comp) {
        const i64 m = comp == 0 ? ORC_M1 : ORC_M2;
        i64 A[3][3], P[3][3];
        comp_matrix(comp, A);
        mat_two_pow(A, P, m, 76);
        mat_vec_mod(P, g->Bg + 3 * comp, g->Bg + 3 * comp, m);
    }
    for (int i = 0; i < 6; ++i) g->Cg[i] = g->Bg[i];
}

void orc_rng_set_antithetic(orc_rng *g, int a) { g->anti = a ? 1 : 0; }
void orc_rng_increased_precis(orc_rng *g, int incp) { g->inc_prec = incp ? 1 : 0; }

/* World.cpp:269-276: srand(rng_seed); sixseed[i] = (rand() + rand()) % {m1,m2}[i/3],
 * the sum taken as unsigned 32-bit */
void orc_sixseed_from_rng_seed(int rng_seed, uint32_t sixseed[6]) {
    const uint32_t two_max[2] = {4294967087u, 4294944443u};
    srand((unsigned)rng_seed);
    for (int i = 0; i < 6; ++i) {
        uint32_t a = (uint32_t)rand();
        uint32_t b = (uint32_t)rand();
        sixseed[i] = (a + b) % two_max[i / 3];
    }
}

/* ======================================================================
 * element algebra (tetra_element_linear.cpp, mat_vec_fns.cpp)
 * ====================================================================== */

/* tetra_element_linear.cpp:191-203 ; J rows = x1-x0, x2-x0, x3-x0 */
static void elem_jacobian(const double *x0, const double *x1, const double *x2, const double *x3,
                          double J[3][3]) {
    for (int c = 0; c < 3; ++c) {
        J[0][c] = x1[c] - x0[c];
        J[1][c] = x2[c] - x0[c];
        J[2][c] = x3[c] - x0[c];
    }
}

/* tetra_element_linear.cpp:222-292 ; returns 1 on inversion (sign change of det
 * against *last_det), leaving dpsi unscaled and vol untouched as the reference does */
static int elem_shape(const double J[3][3], double dpsi[12], double *vol, double *last_det) {
    dpsi[1] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
    dpsi[2] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    dpsi[3] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
    dpsi[5] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
    dpsi[6] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    dpsi[7] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
    dpsi[9] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
    dpsi[10] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    dpsi[11] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
    double det = J[0][0] * dpsi[1] + J[1][0] * dpsi[2] + J[2][0] * dpsi[3];
    if (*last_det * det < 0) return 1;
    *last_det = det;
    *vol = (1.0 / 6.0) * fabs(det);
    det = 1.0 / det;
    dpsi[1] *= det; dpsi[2] *= det; dpsi[3] *= det;
    dpsi[5] *= det; dpsi[6] *= det; dpsi[7] *= det;
    dpsi[9] *= det; dpsi[10] *= det; dpsi[11] *= det;
    dpsi[0] = -(dpsi[1] + dpsi[2] + dpsi[3]);
    dpsi[4] = -(dpsi[5] + dpsi[6] + dpsi[7]);
    dpsi[8] = -(dpsi[9] + dpsi[10] + dpsi[11]);
    return 0;
}

/* mat_vec_fns.cpp:125-154 */
static void mat3_invert(const double m[3][3], double mi[3][3], double *det_m) {
    mi[0][0] = m[2][2] * m[1][1] - m[2][1] * m[1][2];
    mi[0][1] = m[2][1] * m[0][2] - m[2][2] * m[0][1];
    mi[0][2] = m[1][2] * m[0][1] - m[1][1] * m[0][2];
    mi[1][0] = m[2][0] * m[1][2] - m[2][2] * m[1][0];
    mi[1][1] = m[2][2] * m[0][0] - m[2][0] * m[0][2];
    mi[1][2] = m[1][0] * m[0][2] - m[1][2] * m[0][0];
    mi[2][0] = m[2][1] * m[1][0] - m[2][0] * m[1][1];
    mi[2][1] = m[2][0] * m[0][1] - m[2][1] * m[0][0];
    mi[2][2] = m[1][1] * m[0][0] - m[1][0] * m[0][1];
    double det = m[0][0] * mi[0][0] + m[1][0] * mi[0][1] + m[2][0] * mi[0][2];
    *det_m = det;
    det = 1.0 / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) mi[i][j] *= det;
}

/* tetra_element_linear.cpp:298-327 with the macros of tetra_element_linear.h:56-117
 * and calc_del2_matrix/add_diffusion_matrix (:585-622). Entry (4i+a, 4j+b). */
static void elem_viscosity(const double d[12], double vol, double A, double B, double V[12][12]) {
    /* diagonal 4x4 blocks, upper triangle */
    for (int I = 0; I < 12; I += 4)
        for (int a = 0; a < 4; ++a)
            for (int b = a; b < 4; ++b) V[I + a][I + b] = d[I + b] * d[I + a] * (A + B);
    /* off-diagonal blocks I<J: V[I+a][J+b] = B*K[a][b] + A*K[b][a], K[a][b] = d[I+a]*d[J+b];
     * the a==b entries are (B + A) * K[a][a] */
    static const int IJ[3][2] = {{0, 4}, {0, 8}, {4, 8}};
    for (int t = 0; t < 3; ++t) {
        const int I = IJ[t][0], J = IJ[t][1];
        double K[4][4];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) K[a][b] = d[I + a] * d[J + b];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                V[I + a][J + b] = (a == b) ? (B + A) * K[a][a] : B * K[a][b] + A * K[b][a];
    }
    /* del2 (upper triangle) and the diffusion blocks */
    double del2[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = a; b < 4; ++b) del2[a][b] = (d[a] * d[b] + d[4 + a] * d[4 + b] + d[8 + a] * d[8 + b]);
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 4; ++a)
            for (int b = a; b < 4; ++b) V[4 * i + a][4 * i + b] += del2[a][b] * A;
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j <= i; ++j) V[j][i] *= vol;
    for (int i = 1; i < 12; ++i)
        for (int j = 0; j < i; ++j) V[i][j] = V[j][i];
}

/* tetra_element_linear.cpp:375-392 (+ mat_vec_fns.cpp:70-103) ; F kept in Fout */
static void elem_shear_stress(const double J[3][3], const double Ji0[3][3], double G, double vol0,
                              double vol, double s[3][3], double F[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += J[k][i] * Ji0[j][k];
            F[i][j] = acc;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = s[i][j];
            for (int k = 0; k < 3; ++k) acc += F[i][k] * F[j][k];
            s[i][j] = acc;
        }
    const double sc = G * vol0 / vol;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s[i][j] *= sc;
    s[0][0] -= G;
    s[1][1] -= G;
    s[2][2] -= G;
}

/* tetra_element_linear.cpp:397-403 */
static void elem_bulk_stress(double G, double E, double vol0, double vol, double s[3][3]) {
    const double c2 = E - G * 2.0 / 3.0;
    const double c = G * (1.0 - (vol0 / vol)) + 0.5 * c2 * ((vol / vol0) - (vol0 / vol));
    s[0][0] += c;
    s[1][1] += c;
    s[2][2] += c;
}

/* tetra_element_linear.cpp:412-431 ; u[7] in draw order, RAND(-.5,.5) = -.5 + 1.0*u */
static void elem_noise_stress(double vol, double A, double B, double kT, double dt, const double u[7],
                              double s[3][3]) {
    const double c = sqrt((24 * kT) / (vol * dt));
#define ORC_RAND(x) ((-.5) + ((.5) - (-.5)) * (x))
    const double bf = sqrt(B) * ORC_RAND(u[0]);
    s[0][0] += c * (sqrt(2 * A) * ORC_RAND(u[1]) + bf);
    s[1][1] += c * (sqrt(2 * A) * ORC_RAND(u[2]) + bf);
    s[2][2] += c * (sqrt(2 * A) * ORC_RAND(u[3]) + bf);
    s[0][1] += c * sqrt(A) * ORC_RAND(u[4]);
    s[0][2] += c * sqrt(A) * ORC_RAND(u[5]);
    s[1][2] += c * sqrt(A) * ORC_RAND(u[6]);
#undef ORC_RAND
    s[1][0] = s[0][1];
    s[2][0] = s[0][2];
    s[2][1] = s[1][2];
}

/* tetra_element_linear.cpp:436-443 then add_element_force_vector :468-483
 * (du starts at zero for the mass-less solver, Blob.cpp:462-467) */
static double vol_term(const double d[12], double vol, const double s[3][3], int i, int a) {
    return vol * (d[a] * s[i][0] + d[4 + a] * s[i][1] + d[8 + a] * s[i][2]);
}
static void elem_apply_stress(const double d[12], double vol, const double s[3][3], double du[12]) {
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 4; ++a) du[4 * i + a] = 0 + vol_term(d, vol, s, i, a);
}

/* ======================================================================
 * blob
 * ====================================================================== */
#define SRC_ONE 0xFFFFFFFFu
#define SRC_DRAG 0xFFFFFFFEu

struct orc_blob {
    int nt, ne;
    double *pos, *vel, *force, *rhs;
    int *top;
    double *A, *B, *G, *E;
    double *jinv0, *vol0, *last_det, *dpsi, *vol, *V, *node_force;
    int *is_pinned;
    int npin;
    int *pinned;
    int nctf;
    int *ctf_node;
    double *ctf_force;
    double *stokes_drag;
    orc_params p;
    orc_rng rng;
    int *contrib_ptr, *contrib; /* node -> e*10+slot, element order (Blob.cpp:3284-3314) */
    int nrows, nnz;
    int *rowptr, *col, *diag;
    double *val;
    int64_t *src_ptr;
    uint32_t *src;
    double *inv_diag;
    double *r, *pv, *z, *q, *f;
    int last_iters;
    int all_pinned;
    /* linear surface constant forces (Blob::load_ctforces, Blob.cpp:2730-2806) */
    int nsf_faces, nsf_sets;
    int *sf_face_nodes;        /* [nsf_faces][3] Face::n[0..2] */
    double *sf_set_force;      /* [nsf_sets][3] */
    int *sf_set_ptr, *sf_set_faces;
    double *sf_face_force;     /* [nsf_faces][3] Face::force[0..2] of this step */
    /* rotational constant forces (Blob::load_ctforces, Blob.cpp:2650-2726) */
    int nrf;
    int *rf_node;
    double *rf_force;          /* [nrf] F / force unit */
    double *rf_axis;           /* [nrf][6] point / length unit, unit direction */
    int *rf_torque;            /* [nrf] 1: 't' (constant torque), 0: 'f' */
    /* MassLumpedSolver (Blob.cpp:232-234): inertial steps */
    int mass;                  /* 1: MassLumpedSolver, 2: ConjugateGradientSolver */
    double *inv_m;             /* [nt] */
    double *rho;               /* [ne] */
    int *m_ptr, *m_col;        /* ConjugateGradientSolver's mass matrix, scalar CSR over nt rows */
    double *m_val, *m_pre;
    /* external forces for the next step (Blob::add_force_to_node) */
    int next, next_cap;
    int *ext_node;
    double *ext_force;
};

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* tetra_element_linear.cpp:553-577 */
static void linearise_element(orc_blob *b, int e) {
    static const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    const int *n = b->top + 10 * e;
    for (int k = 0; k < 6; ++k) {
        double *m = b->pos + 3 * n[4 + k];
        const double *p0 = b->pos + 3 * n[ends[k][0]], *p1 = b->pos + 3 * n[ends[k][1]];
        for (int c = 0; c < 3; ++c) m[c] = .5 * (p0[c] + p1[c]);
    }
}

static void elem_geometry(orc_blob *b, int e, double J[3][3]) {
    const int *n = b->top + 10 * e;
    elem_jacobian(b->pos + 3 * n[0], b->pos + 3 * n[1], b->pos + 3 * n[2], b->pos + 3 * n[3], J);
}

typedef struct {
    int m;      /* neighbour node */
    int e;      /* element */
    int a, bb;  /* local slots of row node and neighbour */
} nb_entry;

static int nb_cmp(const void *x, const void *y) {
    const nb_entry *p = (const nb_entry *)x, *q = (const nb_entry *)y;
    if (p->m != q->m) return p->m < q->m ? -1 : 1;
    if (p->e != q->e) return p->e < q->e ? -1 : 1;
    return 0;
}

/* NoMassCGSolver::init registration (NoMassCGSolver.cpp:81-127) restated per row:
 * a scalar entry's sources are its element contributions in element order (then
 * the stokes drag), columns ascending (SparsityPattern.cpp:35-56, 73-110). */
static int build_pattern(orc_blob *b) {
    const int nt = b->nt;
    b->nrows = 3 * nt;
    /* node -> incident (e, slot) lists, element order */
    b->contrib_ptr = xcalloc(nt + 1, sizeof(int));
    for (int e = 0; e < b->ne; ++e)
        for (int s = 0; s < 10; ++s) b->contrib_ptr[b->top[10 * e + s] + 1]++;
    for (int i = 0; i < nt; ++i) b->contrib_ptr[i + 1] += b->contrib_ptr[i];
    b->contrib = xcalloc((size_t)b->contrib_ptr[nt], sizeof(int));
    int *fill = xcalloc(nt, sizeof(int));
    for (int e = 0; e < b->ne; ++e)
        for (int s = 0; s < 10; ++s) {
            const int n = b->top[10 * e + s];
            b->contrib[b->contrib_ptr[n] + fill[n]++] = 10 * e + s;
        }
    free(fill);

    /* pass 1: count entries and sources per node row */
    int64_t nnz = 0, nsrc = 0;
    nb_entry *tmp = NULL;
    size_t tmpcap = 0;
    b->rowptr = xcalloc((size_t)b->nrows + 1, sizeof(int));
    for (int pass = 0; pass < 2; ++pass) {
        nnz = 0;
        nsrc = 0;
        for (int n = 0; n < nt; ++n) {
            const int c0 = b->contrib_ptr[n], c1 = b->contrib_ptr[n + 1];
            size_t cnt = 0;
            int has_lin = 0;
            if (!b->is_pinned[n]) {
                for (int c = c0; c < c1; ++c) {
                    const int s = b->contrib[c] % 10;
                    if (s < 4) { has_lin = 1; cnt += 4; }
                }
            }
            if (cnt > tmpcap) {
                tmpcap = cnt * 2;
                tmp = realloc(tmp, tmpcap * sizeof(nb_entry));
            }
            size_t k = 0;
            if (has_lin) {
                for (int c = c0; c < c1; ++c) {
                    const int e = b->contrib[c] / 10, s = b->contrib[c] % 10;
                    if (s >= 4) continue;
                    for (int bb = 0; bb < 4; ++bb) {
                        const int m = b->top[10 * e + bb];
                        if (b->is_pinned[m]) continue;
                        tmp[k].m = m; tmp[k].e = e; tmp[k].a = s; tmp[k].bb = bb;
                        ++k;
                    }
                }
                qsort(tmp, k, sizeof(nb_entry), nb_cmp);
            }
            /* does the node carry element sources on its diagonal? (a linear slot in
             * some element, unpinned). Otherwise its 3 diagonal entries get `one`. */
            int ndistinct = 0;
            for (size_t t = 0; t < k; ++t)
                if (t == 0 || tmp[t].m != tmp[t - 1].m) ++ndistinct;
            int diag_from_elem = 0;
            for (size_t t = 0; t < k; ++t)
                if (tmp[t].m == n) { diag_from_elem = 1; break; }
            const int drag = b->p.calc_stokes && !b->is_pinned[n];
            for (int i = 0; i < 3; ++i) {
                const int row = 3 * n + i;
                if (pass == 0) {
                    int row_nnz = ndistinct * 3 + (diag_from_elem ? 0 : 1);
                    b->rowptr[row + 1] = row_nnz;
                } else {
                    int pos = b->rowptr[row];
                    int64_t sp = b->src_ptr[pos];
                    int placed_diag = diag_from_elem;
                    /* entries in ascending column: neighbours ascending, j ascending;
                     * a `one` diagonal is placed in column order as well */
                    size_t t = 0;
                    while (t < k || !placed_diag) {
                        const int mnext = (t < k) ? tmp[t].m : 0x7fffffff;
                        if (!placed_diag && n <= mnext) {
                            /* identity diagonal (midside or pinned row) */
                            b->col[pos] = 3 * n + i;
                            b->diag[row] = pos;
                            b->src[sp++] = SRC_ONE;
                            if (drag) b->src[sp++] = SRC_DRAG;
                            b->src_ptr[++pos] = sp;
                            placed_diag = 1;
                            continue;
                        }
                        size_t t1 = t;
                        while (t1 < k && tmp[t1].m == mnext) ++t1;
                        for (int j = 0; j < 3; ++j) {
                            b->col[pos] = 3 * mnext + j;
                            if (mnext == n && j == i) b->diag[row] = pos;
                            for (size_t u = t; u < t1; ++u) {
                                const int e = tmp[u].e, a = tmp[u].a, bb = tmp[u].bb;
                                b->src[sp++] = (uint32_t)((int64_t)e * 144 + (a + 4 * i) * 12 + (bb + 4 * j));
                            }
                            if (mnext == n && j == i && drag) b->src[sp++] = SRC_DRAG;
                            b->src_ptr[++pos] = sp;
                        }
                        t = t1;
                    }
                }
            }
            if (pass == 0) {
                nsrc += (int64_t)3 * (3 * (int64_t)k + (diag_from_elem ? 0 : 1) + (drag ? 1 : 0));
            }
        }
        if (pass == 0) {
            for (int r = 0; r < b->nrows; ++r) b->rowptr[r + 1] += b->rowptr[r];
            b->nnz = b->rowptr[b->nrows];
            b->col = xcalloc((size_t)b->nnz, sizeof(int));
            b->val = xcalloc((size_t)b->nnz, sizeof(double));
            b->src_ptr = xcalloc((size_t)b->nnz + 1, sizeof(int64_t));
            b->src = xcalloc((size_t)nsrc, sizeof(uint32_t));
            b->diag = xcalloc((size_t)b->nrows, sizeof(int));
            /* src_ptr at each row start is filled as rows are written in order */
        }
    }
    free(tmp);
    (void)nnz;
    return 0;
}

orc_blob *orc_blob_create(int nt, int ne, const double *pos, const int *top, const double *A,
                          const double *B, const double *G, const double *E, int npin, const int *pinned,
                          int nctf, const int *ctf_node, const double *ctf_force, const double *stokes_drag,
                          const orc_params *params, const uint32_t rng_state[6]) {
    if (nt <= 0 || ne <= 0 || (int64_t)ne * 144 >= 0xFFFFFFF0LL) return NULL;
    for (int i = 0; i < 10 * ne; ++i)
        if (top[i] < 0 || top[i] >= nt) return NULL;
    orc_blob *b = xcalloc(1, sizeof(orc_blob));
    b->nt = nt;
    b->ne = ne;
    b->p = *params;
    b->pos = xcalloc((size_t)3 * nt, sizeof(double));
    memcpy(b->pos, pos, sizeof(double) * 3 * nt);
    b->vel = xcalloc((size_t)3 * nt, sizeof(double));
    b->force = xcalloc((size_t)3 * nt, sizeof(double));
    b->rhs = xcalloc((size_t)3 * nt, sizeof(double));
    b->top = xcalloc((size_t)10 * ne, sizeof(int));
    memcpy(b->top, top, sizeof(int) * 10 * ne);
    b->A = xcalloc(ne, sizeof(double));
    b->B = xcalloc(ne, sizeof(double));
    b->G = xcalloc(ne, sizeof(double));
    b->E = xcalloc(ne, sizeof(double));
    memcpy(b->A, A, sizeof(double) * ne);
    memcpy(b->B, B, sizeof(double) * ne);
    memcpy(b->G, G, sizeof(double) * ne);
    memcpy(b->E, E, sizeof(double) * ne);
    b->jinv0 = xcalloc((size_t)9 * ne, sizeof(double));
    b->vol0 = xcalloc(ne, sizeof(double));
    b->last_det = xcalloc(ne, sizeof(double));
    b->dpsi = xcalloc((size_t)12 * ne, sizeof(double));
    b->vol = xcalloc(ne, sizeof(double));
    b->V = xcalloc((size_t)144 * ne, sizeof(double));
    b->node_force = xcalloc((size_t)30 * ne, sizeof(double));
    b->is_pinned = xcalloc(nt, sizeof(int));
    b->npin = npin;
    b->pinned = xcalloc(npin, sizeof(int));
    for (int i = 0; i < npin; ++i) {
        if (pinned[i] < 0 || pinned[i] >= nt) { orc_blob_destroy(b); return NULL; }
        b->pinned[i] = pinned[i];
        b->is_pinned[pinned[i]] = 1;
    }
    b->nctf = nctf;
    b->ctf_node = xcalloc(nctf, sizeof(int));
    b->ctf_force = xcalloc((size_t)3 * nctf, sizeof(double));
    for (int i = 0; i < nctf; ++i) {
        if (ctf_node[i] < 0 || ctf_node[i] >= nt) { orc_blob_destroy(b); return NULL; }
        b->ctf_node[i] = ctf_node[i];
        for (int c = 0; c < 3; ++c) b->ctf_force[3 * i + c] = ctf_force[3 * i + c];
    }
    b->stokes_drag = xcalloc(nt, sizeof(double));
    if (stokes_drag) memcpy(b->stokes_drag, stokes_drag, sizeof(double) * nt);
    orc_rng_set_seed(&b->rng, rng_state);

    /* Blob::init order (Blob.cpp:203-247): linearise, rest state, connectivity, solver */
    for (int e = 0; e < ne; ++e) linearise_element(b, e);
    for (int e = 0; e < ne; ++e) { /* calc_rest_state_info, Blob.cpp:3080-3143 */
        double J[3][3], Ji[3][3], tmp;
        elem_geometry(b, e, J);
        mat3_invert(J, Ji, &tmp);
        memcpy(b->jinv0 + 9 * e, Ji, sizeof(Ji));
        elem_shape(J, b->dpsi + 12 * e, &b->vol[e], &b->last_det[e]);
        b->vol0[e] = b->vol[e];
    }
    b->all_pinned = 0;
    {
        int cnt = 0;
        for (int n = 0; n < nt; ++n) cnt += b->is_pinned[n];
        /* Blob.cpp:405 compares the pin list length with the node count */
        b->all_pinned = (npin == nt) && cnt > 0;
    }
    build_pattern(b);
    /* NoMassCGSolver::init recomputes shape derivatives for every element (:83) */
    for (int e = 0; e < ne; ++e) {
        double J[3][3];
        elem_geometry(b, e, J);
        elem_shape(J, b->dpsi + 12 * e, &b->vol[e], &b->last_det[e]);
    }
    b->inv_diag = xcalloc((size_t)3 * nt, sizeof(double));
    b->r = xcalloc((size_t)3 * nt, sizeof(double));
    b->pv = xcalloc((size_t)3 * nt, sizeof(double));
    b->z = xcalloc((size_t)3 * nt, sizeof(double));
    b->q = xcalloc((size_t)3 * nt, sizeof(double));
    b->f = xcalloc((size_t)3 * nt, sizeof(double));
    return b;
}

void orc_blob_destroy(orc_blob *b) {
    if (!b) return;
    double *d[] = {b->pos, b->vel, b->force, b->rhs, b->A, b->B, b->G, b->E, b->jinv0, b->vol0,
                   b->last_det, b->dpsi, b->vol, b->V, b->node_force, b->ctf_force, b->stokes_drag,
                   b->val, b->inv_diag, b->r, b->pv, b->z, b->q, b->f};
    for (size_t i = 0; i < sizeof(d) / sizeof(d[0]); ++i) free(d[i]);
    free(b->top); free(b->is_pinned); free(b->pinned); free(b->ctf_node);
    free(b->sf_face_nodes); free(b->sf_set_force); free(b->sf_set_ptr); free(b->sf_set_faces); free(b->sf_face_force);
    free(b->rf_node); free(b->rf_force); free(b->rf_axis); free(b->rf_torque);
    free(b->ext_node); free(b->ext_force); free(b->inv_m); free(b->rho);
    free(b->m_ptr); free(b->m_col); free(b->m_val); free(b->m_pre);
    free(b->contrib_ptr); free(b->contrib); free(b->rowptr); free(b->col); free(b->diag);
    free(b->src_ptr); free(b->src);
    free(b);
}

static void surface_ctforces(orc_blob *b);
static void rotational_ctforces(orc_blob *b);
static int mass_cg_solve(orc_blob *b, double *x);

/* Blob::update_internal_forces (Blob.cpp:400-491), single thread (rng[0]) */
int orc_blob_update_internal_forces(orc_blob *b) {
    if (b->all_pinned) return 0;
    /* apply_ctforces linear part (Blob.cpp:1679-1683), rotational part (:1684-1719),
     * surface part (:1721-1740) */
    for (int i = 0; i < b->nctf; ++i)
        for (int c = 0; c < 3; ++c) b->force[3 * b->ctf_node[i] + c] += b->ctf_force[3 * i + c];
    if (b->nrf) rotational_ctforces(b);
    if (b->nsf_sets) surface_ctforces(b);
    int inversions = 0;
#ifdef _OPENMP
    /* each thread draws from its own stream, the reference's rng[omp_get_thread_num()]
     * (Blob.cpp:424-428): substreams of the blob's stream here */
    const int nthr = omp_get_max_threads();
    orc_rng *trng = (orc_rng *)malloc(sizeof(orc_rng) * (size_t)nthr);
    for (int t = 0; t < nthr; ++t) {
        trng[t] = b->rng;
        for (int k = 0; k < t; ++k) orc_rng_reset_next_substream(&trng[t]);
    }
#endif
    OMP(omp parallel for schedule(guided) reduction(+ : inversions))
    for (int e = 0; e < b->ne; ++e) {
        double J[3][3], s[3][3] = {{0}}, F[3][3], Ji0[3][3];
        double *d = b->dpsi + 12 * e;
        elem_geometry(b, e, J);
        if (elem_shape(J, d, &b->vol[e], &b->last_det[e])) inversions++;
        elem_viscosity(d, b->vol[e], b->A[e], b->B[e], (double(*)[12])(b->V + 144 * e));
        memcpy(Ji0, b->jinv0 + 9 * e, sizeof(Ji0));
        elem_shear_stress(J, Ji0, b->G[e], b->vol0[e], b->vol[e], s, F);
        elem_bulk_stress(b->G[e], b->E[e], b->vol0[e], b->vol[e], s);
        if (b->p.calc_noise == 1) {
            double u[7];
#ifdef _OPENMP
            for (int k = 0; k < 7; ++k) u[k] = orc_rng_rand_u01(&trng[omp_get_thread_num()]);
#else
            for (int k = 0; k < 7; ++k) u[k] = orc_rng_rand_u01(&b->rng);
#endif
            elem_noise_stress(b->vol[e], b->A[e], b->B[e], b->p.kT, b->p.dt, u, s);
        }
        double du[12];
        elem_apply_stress(d, b->vol[e], s, du);
        if (b->mass) {
            /* Blob.cpp:460-466 with mass: du = V v_e (get_element_velocity_vector,
             * tetra_element_linear.cpp:448-463; mat12_apply, mat_vec_fns.cpp:28-36), then the
             * stress terms are added to it */
            const int *nn = b->top + 10 * e;
            const double(*V)[12] = (const double(*)[12])(b->V + 144 * e);
            double v[12], tv[12];
            for (int a = 0; a < 4; ++a)
                for (int c = 0; c < 3; ++c) v[4 * c + a] = b->vel[3 * nn[a] + c];
            for (int i = 0; i < 12; ++i) {
                tv[i] = 0;
                for (int j = 0; j < 12; ++j) tv[i] += V[i][j] * v[j];
            }
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < 4; ++a)
                    du[4 * i + a] = tv[4 * i + a] + vol_term(d, b->vol[e], s, i, a);
        }
        double *out = b->node_force + 30 * e;
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 3; ++c) out[3 * a + c] -= du[4 * c + a];
    }
#ifdef _OPENMP
    b->rng = trng[0];
    free(trng);
#endif
    return inversions;
}

/* SparseMatrixFixedPattern::build + calc_inverse_diagonal (SparseMatrixFixedPattern.cpp:86-94,
 * 305-312 ; SparseMatrixTypes.cpp:40-46) */
static void matrix_build(orc_blob *b) {
    OMP(omp parallel for schedule(static))
    for (int k = 0; k < b->nnz; ++k) {
        double sum = 0.0;
        for (int64_t t = b->src_ptr[k]; t < b->src_ptr[k + 1]; ++t) {
            const uint32_t s = b->src[t];
            if (s == SRC_ONE) sum += 1.0;
            else if (s == SRC_DRAG) sum += b->stokes_drag[b->col[k] / 3];
            else sum += b->V[s];
        }
        b->val[k] = sum;
    }
    OMP(omp parallel for schedule(static))
    for (int r = 0; r < b->nrows; ++r) b->inv_diag[r] = 1.0 / b->val[b->diag[r]];
}

void orc_blob_build_matrix(orc_blob *b) {
    for (int e = 0; e < b->ne; ++e) {
        double J[3][3];
        elem_geometry(b, e, J);
        elem_shape(J, b->dpsi + 12 * e, &b->vol[e], &b->last_det[e]);
        elem_viscosity(b->dpsi + 12 * e, b->vol[e], b->A[e], b->B[e], (double(*)[12])(b->V + 144 * e));
    }
    matrix_build(b);
}

/* SparseMatrixFixedPattern::apply(arr3) (:148-187) */
void orc_blob_apply_matrix(const orc_blob *b, const double *in, double *out) {
    OMP(omp parallel for schedule(static))
    for (int r = 0; r < b->nrows; ++r) {
        double acc = 0.0;
        for (int k = b->rowptr[r]; k < b->rowptr[r + 1]; ++k) acc += b->val[k] * in[b->col[k]];
        out[r] = acc;
    }
}

/* NoMassCGSolver::solve (NoMassCGSolver.cpp:156-188, 196-281; mat_vec_fns.cpp:187-212) */
static int cg_solve(orc_blob *b, double *x) {
    const int nn = b->nt;
    double *r = b->r, *p = b->pv, *z = b->z, *q = b->q, *f = b->f, *D = b->inv_diag;
    double delta_new = 0;
    OMP(omp parallel for schedule(static) reduction(+ : delta_new))
    for (int i = 0; i < nn; ++i) {
        for (int c = 0; c < 3; ++c) {
            r[3 * i + c] = x[3 * i + c];
            f[3 * i + c] = x[3 * i + c];
            x[3 * i + c] = 0;
            z[3 * i + c] = D[3 * i + c] * r[3 * i + c];
            p[3 * i + c] = z[3 * i + c];
        }
        delta_new += r[3 * i] * z[3 * i] + r[3 * i + 1] * z[3 * i + 1] + r[3 * i + 2] * z[3 * i + 2];
    }
    for (int it = 0; it < b->p.max_iterations_cg; ++it) {
        orc_blob_apply_matrix(b, p, q);
        double pTq = 0;
        OMP(omp parallel for schedule(static) reduction(+ : pTq))
        for (int i = 0; i < nn; ++i)
            pTq += p[3 * i] * q[3 * i] + p[3 * i + 1] * q[3 * i + 1] + p[3 * i + 2] * q[3 * i + 2];
        const double alpha = delta_new / pTq;
        OMP(omp parallel for schedule(static))
        for (int k = 0; k < 3 * nn; ++k) {
            x[k] += alpha * p[k];
            r[k] += -alpha * q[k];
        }
        double r2 = 0, f2 = 0;
        OMP(omp parallel for schedule(static) reduction(+ : r2, f2))
        for (int i = 0; i < nn; ++i) {
            r2 += r[3 * i] * r[3 * i] + r[3 * i + 1] * r[3 * i + 1] + r[3 * i + 2] * r[3 * i + 2];
            f2 += f[3 * i] * f[3 * i] + f[3 * i + 1] * f[3 * i + 1] + f[3 * i + 2] * f[3 * i + 2];
        }
        const double res = (f2 == 0.0) ? 0.0 : r2 / f2;
        if (res < b->p.epsilon2) {
            b->last_iters = it + 1;
            return it + 1;
        }
        const double delta_old = delta_new;
        delta_new = 0;
        OMP(omp parallel for schedule(static) reduction(+ : delta_new))
        for (int i = 0; i < nn; ++i) {
            for (int c = 0; c < 3; ++c) z[3 * i + c] = D[3 * i + c] * r[3 * i + c];
            delta_new += r[3 * i] * z[3 * i] + r[3 * i + 1] * z[3 * i + 1] + r[3 * i + 2] * z[3 * i + 2];
        }
        const double beta = delta_new / delta_old;
        OMP(omp parallel for schedule(static))
        for (int k = 0; k < 3 * nn; ++k) p[k] = beta * p[k] + z[k];
    }
    b->last_iters = b->p.max_iterations_cg;
    return -1;
}

int orc_blob_solve(orc_blob *b, double *x) { return cg_solve(b, x); }

/* Blob::linearise_force (Blob.cpp:781-799) */
static void linearise_force(orc_blob *b) {
    for (int e = 0; e < b->ne; ++e) {
        const int *n = b->top + 10 * e;
        double *F = b->force;
        for (int j = 0; j < 3; ++j) {
            F[3 * n[0] + j] += 0.5 * (F[3 * n[4] + j] + F[3 * n[5] + j] + F[3 * n[6] + j]);
            F[3 * n[1] + j] += 0.5 * (F[3 * n[4] + j] + F[3 * n[7] + j] + F[3 * n[8] + j]);
            F[3 * n[2] + j] += 0.5 * (F[3 * n[5] + j] + F[3 * n[7] + j] + F[3 * n[9] + j]);
            F[3 * n[3] + j] += 0.5 * (F[3 * n[6] + j] + F[3 * n[8] + j] + F[3 * n[9] + j]);
        }
        for (int s = 4; s < 10; ++s)
            for (int c = 0; c < 3; ++c) F[3 * n[s] + c] = 0.;
    }
}

int orc_blob_set_surface_forces(orc_blob *b, int nfaces, const int *face_nodes, int nsets, const double *set_force,
                                const int *set_ptr, const int *set_faces) {
    if (nfaces < 0 || nsets < 0 || (nsets && (!set_ptr || set_ptr[0] != 0))) return 1;
    for (int i = 0; i < 3 * nfaces; ++i)
        if (face_nodes[i] < 0 || face_nodes[i] >= b->nt) return 1;
    for (int s = 0; s < nsets; ++s)
        if (set_ptr[s + 1] < set_ptr[s]) return 1;
    for (int k = 0; nsets && k < set_ptr[nsets]; ++k)
        if (set_faces[k] < 0 || set_faces[k] >= nfaces) return 1;
    const int ne = nsets ? set_ptr[nsets] : 0;
    free(b->sf_face_nodes); free(b->sf_set_force); free(b->sf_set_ptr); free(b->sf_set_faces); free(b->sf_face_force);
    b->nsf_faces = nfaces;
    b->nsf_sets = nsets;
    b->sf_face_nodes = xcalloc((size_t)3 * nfaces, sizeof(int));
    memcpy(b->sf_face_nodes, face_nodes, sizeof(int) * 3 * (size_t)nfaces);
    b->sf_set_force = xcalloc((size_t)3 * nsets, sizeof(double));
    memcpy(b->sf_set_force, set_force, sizeof(double) * 3 * (size_t)nsets);
    b->sf_set_ptr = xcalloc((size_t)nsets + 1, sizeof(int));
    if (nsets) memcpy(b->sf_set_ptr, set_ptr, sizeof(int) * ((size_t)nsets + 1));
    b->sf_set_faces = xcalloc((size_t)ne, sizeof(int));
    memcpy(b->sf_set_faces, set_faces, sizeof(int) * (size_t)ne);
    b->sf_face_force = xcalloc((size_t)3 * nfaces, sizeof(double));
    return 0;
}

/* Face::get_area (Face.cpp:169-187) */
static double face_area(const orc_blob *b, int f) {
    const int *n = b->sf_face_nodes + 3 * f;
    double a[3], c[3], t[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = b->pos[3 * n[1] + i] - b->pos[3 * n[0] + i];
        c[i] = b->pos[3 * n[2] + i] - b->pos[3 * n[0] + i];
    }
    t[0] = a[1] * c[2] - a[2] * c[1];
    t[1] = a[2] * c[0] - a[0] * c[2];
    t[2] = a[0] * c[1] - a[1] * c[0];
    double m = 0.0;
    for (int i = 0; i < 3; ++i) m += t[i] * t[i];
    return .5 * sqrt(m);
}

/* Blob::apply_ctforces CTF 3 (Blob.cpp:1721-1740): each set's force is spread over its
 * faces by area, a third of each face's share on each of the face's three nodes
 * (Face::add_force_to_node, Face.cpp:238-241), at the positions of the step */
static void surface_ctforces(orc_blob *b) {
    memset(b->sf_face_force, 0, sizeof(double) * 3 * (size_t)b->nsf_faces);
    for (int s = 0; s < b->nsf_sets; ++s) {
        const int k0 = b->sf_set_ptr[s], k1 = b->sf_set_ptr[s + 1];
        double total = 0;
        double *areas = xcalloc((size_t)(k1 - k0), sizeof(double));
        for (int k = k0; k < k1; ++k) {
            areas[k - k0] = face_area(b, b->sf_set_faces[k]);
            total += areas[k - k0];
        }
        for (int k = k0; k < k1; ++k) {
            const double sc = areas[k - k0] / (3 * total);
            double *ff = b->sf_face_force + 3 * b->sf_set_faces[k];
            for (int c = 0; c < 3; ++c) ff[c] += sc * b->sf_set_force[3 * s + c];
        }
        free(areas);
    }
}

/* u x v as arr3arr3VectorProduct (mat_vec_fns_II.h:369-373) */
static void cross3(const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - v[1] * u[2];
    w[1] = -u[0] * v[2] + v[0] * u[2];
    w[2] = u[0] * v[1] - v[0] * u[1];
}

/* Blob::apply_ctforces CTF 2 (Blob.cpp:1684-1719), axes given by point and direction: r
 * runs from the node to its foot on the axis (intersectingPointToLine,
 * mat_vec_fns_II.h:755-777: foot = p + c (d x ((p - a) x d)), c from detByRows,
 * mat_vec_fns_II.h:379-386; a node on the axis, |(p - a) x d| < 3 eps, gets r = 0), and
 * the node receives F (r x d) ('f') or (F / |r|) (r x d) ('t'), in list order */
void orc_rotational_force(const double *x, const double *axis, double F, int torque, double *f) {
    const double three_eps = 3.0 * DBL_EPSILON;
    const double *a = axis, *d = axis + 3;
    double pa[3], v2[3], cp[3], v1v2[3], foot[3], r[3], w[3];
    for (int c = 0; c < 3; ++c) pa[c] = x[c] - a[c];
    cross3(pa, d, v2);
    double m = 0.0;
    for (int c = 0; c < 3; ++c) m += v2[c] * v2[c];
    if (sqrt(m) < three_eps) {
        for (int c = 0; c < 3; ++c) foot[c] = x[c];
    } else {
        cross3(d, v2, cp);
        cross3(d, cp, v1v2);
        double det = pa[0] * (d[1] * v1v2[2] - d[2] * v1v2[1]);
        det += pa[1] * (d[2] * v1v2[0] - d[0] * v1v2[2]);
        det += pa[2] * (d[0] * v1v2[1] - d[1] * v1v2[0]);
        const double cc = det / (v1v2[0] * v1v2[0] + v1v2[1] * v1v2[1] + v1v2[2] * v1v2[2]);
        for (int c = 0; c < 3; ++c) foot[c] = x[c] + cc * cp[c];
    }
    for (int c = 0; c < 3; ++c) r[c] = foot[c] - x[c];
    cross3(r, d, w);
    double s = F;
    if (torque) {
        double rr = 0.0;
        for (int c = 0; c < 3; ++c) rr += r[c] * r[c];
        s = F / sqrt(rr);
    }
    for (int c = 0; c < 3; ++c) f[c] = s * w[c];
}

static void rotational_ctforces(orc_blob *b) {
    for (int i = 0; i < b->nrf; ++i) {
        double f[3];
        orc_rotational_force(b->pos + 3 * b->rf_node[i], b->rf_axis + 6 * i, b->rf_force[i], b->rf_torque[i], f);
        for (int c = 0; c < 3; ++c) b->force[3 * b->rf_node[i] + c] += f[c];
    }
}

int orc_blob_set_rotational_forces(orc_blob *b, int num, const int *node, const double *force, const double *axis,
                                   const int *torque) {
    if (num < 0) return 1;
    for (int i = 0; i < num; ++i)
        if (node[i] < 0 || node[i] >= b->nt || (torque[i] != 0 && torque[i] != 1)) return 1;
    free(b->rf_node); free(b->rf_force); free(b->rf_axis); free(b->rf_torque);
    b->nrf = num;
    b->rf_node = xcalloc((size_t)num, sizeof(int));
    b->rf_force = xcalloc((size_t)num, sizeof(double));
    b->rf_axis = xcalloc((size_t)6 * num, sizeof(double));
    b->rf_torque = xcalloc((size_t)num, sizeof(int));
    if (num) {
        memcpy(b->rf_node, node, sizeof(int) * (size_t)num);
        memcpy(b->rf_force, force, sizeof(double) * (size_t)num);
        memcpy(b->rf_axis, axis, sizeof(double) * 6 * (size_t)num);
        memcpy(b->rf_torque, torque, sizeof(int) * (size_t)num);
    }
    return 0;
}

/* Blob::update_positions -> aggregate_forces_and_solve (Blob.cpp:3147-3244),
 * euler_integrate (:3249-3263), linearise_element loop (:505-507) */
int orc_blob_update_positions(orc_blob *b) {
    const int nt = b->nt;
    double *F = b->force;
    OMP(omp parallel for schedule(static))
    for (int n = 0; n < nt; ++n)
        for (int m = b->contrib_ptr[n]; m < b->contrib_ptr[n + 1]; ++m) {
            const double *src = b->node_force + 3 * b->contrib[m];
            F[3 * n] += src[0];
            F[3 * n + 1] += src[1];
            F[3 * n + 2] += src[2];
        }
    /* surface forces onto nodes, face order (Blob.cpp:3160-3170; the opposite node, i = 3,
     * receives no constant force) */
    for (int f = 0; f < b->nsf_faces; ++f)
        for (int i = 0; i < 3; ++i) {
            const int n = b->sf_face_nodes[3 * f + i];
            for (int c = 0; c < 3; ++c) F[3 * n + c] += b->sf_face_force[3 * f + c];
        }
    if (b->p.calc_stokes == 1 && b->mass) {  /* Blob.cpp:3172-3196 */
        for (int i = 0; i < nt; ++i) {
            for (int c = 0; c < 3; ++c) F[3 * i + c] -= b->vel[3 * i + c] * b->stokes_drag[i];
            if (b->p.calc_noise == 1) {
                const double sc = sqrt((24 * b->p.kT * b->stokes_drag[i]) / (b->p.dt));
                for (int c = 0; c < 3; ++c)
                    F[3 * i + c] -= ((-.5) + ((.5) - (-.5)) * orc_rng_rand_u01(&b->rng)) * sc;
            }
        }
    } else if (b->p.calc_stokes == 1 && b->p.calc_noise == 1) {
        for (int i = 0; i < nt; ++i) {
            const double sc = sqrt((24 * b->p.kT * b->stokes_drag[i]) / (b->p.dt));
            for (int c = 0; c < 3; ++c) F[3 * i + c] -= ((-.5) + ((.5) - (-.5)) * orc_rng_rand_u01(&b->rng)) * sc;
        }
    }
    linearise_force(b);
    for (int k = 0; k < b->npin; ++k)
        for (int c = 0; c < 3; ++c) F[3 * b->pinned[k] + c] = 0;
    memcpy(b->rhs, F, sizeof(double) * 3 * nt);
    if (b->mass) {
        /* MassLumpedSolver::solve (MassLumpedSolver.cpp:66-73) or ConjugateGradientSolver's,
         * then the inertial euler_integrate (Blob.cpp:3266-3280) */
        if (b->mass == 2) {
            if (mass_cg_solve(b, F) < 0) return 1;
        } else {
            for (int i = 0; i < nt; ++i)
                for (int c = 0; c < 3; ++c) F[3 * i + c] = F[3 * i + c] * b->inv_m[i];
        }
        for (int i = 0; i < nt; ++i)
            for (int c = 0; c < 3; ++c) {
                b->vel[3 * i + c] += F[3 * i + c] * b->p.dt;
                b->pos[3 * i + c] += b->vel[3 * i + c] * b->p.dt;
            }
        for (int e = 0; e < b->ne; ++e) linearise_element(b, e);
        return 0;
    }
    /* solver->solve(force): build V from the element matrices of this step */
    matrix_build(b);
    int it = cg_solve(b, F);
    if (it < 0) return 1;
    OMP(omp parallel for schedule(static))
    for (int i = 0; i < nt; ++i)
        for (int c = 0; c < 3; ++c) {
            b->vel[3 * i + c] = F[3 * i + c];
            b->pos[3 * i + c] += F[3 * i + c] * b->p.dt;
        }
    OMP(omp parallel for schedule(static))
    for (int e = 0; e < b->ne; ++e) linearise_element(b, e);
    return 0;
}

/* one World::run iteration for one blob (World.cpp:1786-1866, 2031-2059) */
/* MassLumpedSolver::init (MassLumpedSolver.cpp:27-63) */
int orc_blob_set_mass_lumped(orc_blob *b, const double *rho) {
    for (int e = 0; e < b->ne; ++e)
        if (!(rho[e] > 0)) return 1;
    free(b->inv_m);
    b->inv_m = xcalloc((size_t)b->nt, sizeof(double));
    for (int i = 0; i < b->nt; ++i) b->inv_m[i] = 0;
    for (int e = 0; e < b->ne; ++e) {
        const int *n = b->top + 10 * e;
        for (int i = 0; i < 10; ++i) {
            if (i < 4) b->inv_m[n[i]] += .25 * rho[e] * b->vol0[e];
            else b->inv_m[n[i]] = 1;
        }
    }
    for (int k = 0; k < b->npin; ++k) b->inv_m[b->pinned[k]] = 1.0;
    for (int i = 0; i < b->nt; ++i) b->inv_m[i] = 1.0 / b->inv_m[i];
    memset(b->vel, 0, sizeof(double) * 3 * (size_t)b->nt);
    free(b->rho);
    b->rho = xcalloc((size_t)b->ne, sizeof(double));
    memcpy(b->rho, rho, sizeof(double) * (size_t)b->ne);
    b->mass = 1;
    return 0;
}

/* ConjugateGradientSolver::init (ConjugateGradientSolver.cpp:51-190): the linear tets'
 * consistent mass matrix over the corner nodes, .1 rho vol_0 on a node's own entry and .05
 * rho vol_0 between two corners of an element, added in element order; a pair with a pinned
 * node gets nothing except a 1 on a pinned node's diagonal; midside nodes are identity
 * rows. Entries that stay exactly 0 are dropped; columns ascend. Jacobi preconditioner. */
static int cmp_int(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }
static int find_entry(const orc_blob *b, int i, int j) {
    int lo = b->m_ptr[i], hi = b->m_ptr[i + 1] - 1;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (b->m_col[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

int orc_blob_set_mass_consistent(orc_blob *b, const double *rho) {
    if (orc_blob_set_mass_lumped(b, rho)) return 1;  /* densities, zero velocities */
    const int nt = b->nt, ne = b->ne;
    /* pattern: each node's neighbours through elements (corners), or itself */
    int *cnt = xcalloc((size_t)nt + 1, sizeof(int));
    int *nb = xcalloc((size_t)nt * 64, sizeof(int));  /* <= 64 distinct neighbours per node */
    for (int e = 0; e < ne; ++e) {
        const int *n = b->top + 10 * e;
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) {
                if (!((i < 4 && j < 4) || i == j)) continue;
                const int ni = n[i], nj = n[j];
                int found = 0;
                for (int k = 0; k < cnt[ni]; ++k)
                    if (nb[64 * ni + k] == nj) { found = 1; break; }
                if (!found) {
                    if (cnt[ni] >= 64) { free(cnt); free(nb); return 1; }
                    nb[64 * ni + cnt[ni]++] = nj;
                }
            }
    }
    free(b->m_ptr); free(b->m_col); free(b->m_val); free(b->m_pre);
    b->m_ptr = xcalloc((size_t)nt + 1, sizeof(int));
    for (int i = 0; i < nt; ++i) {
        qsort(nb + 64 * i, (size_t)cnt[i], sizeof(int), cmp_int);
        b->m_ptr[i + 1] = b->m_ptr[i] + cnt[i];
    }
    b->m_col = xcalloc((size_t)b->m_ptr[nt], sizeof(int));
    b->m_val = xcalloc((size_t)b->m_ptr[nt], sizeof(double));
    for (int i = 0; i < nt; ++i)
        for (int k = 0; k < cnt[i]; ++k) b->m_col[b->m_ptr[i] + k] = nb[64 * i + k];
    free(cnt);
    free(nb);
    /* the entry (ni, nj) in the CSR */
    #define MENTRY(ni, nj) find_entry(b, ni, nj)
    for (int e = 0; e < ne; ++e) {
        const int *n = b->top + 10 * e;
        for (int i = 0; i < 10; ++i)
            for (int j = 0; j < 10; ++j) {
                if (i < 4 && j < 4) {
                    const int ni = n[i], nj = n[j];
                    double *v = b->m_val + MENTRY(ni, nj);
                    if (!b->is_pinned[ni] && !b->is_pinned[nj]) {
                        if (i == j) *v += .1 * rho[e] * b->vol0[e];
                        else *v += .05 * rho[e] * b->vol0[e];
                    } else if (i == j) {
                        *v = 1;
                    }
                } else if (i == j) {
                    b->m_val[MENTRY(n[i], n[j])] = 1;
                }
            }
    }
    #undef MENTRY
    b->m_pre = xcalloc((size_t)nt, sizeof(double));
    for (int i = 0; i < nt; ++i) b->m_pre[i] = 1.0 / b->m_val[find_entry(b, i, i)];
    b->mass = 2;
    return 0;
}

/* ConjugateGradientSolver::solve (ConjugateGradientSolver.cpp:192-215, 226-335): Jacobi-PCG
 * on M a = f, x_0 = 0, stop when sum r^2 / sum f^2 < eps^2 at the start of an iteration.
 * Returns the iterations, or -1 without convergence. Dropped zero entries change no sum. */
static int mass_cg_solve(orc_blob *b, double *x) {
    const int nn = b->nt;
    double *r = b->r, *d = b->pv, *s = b->z, *q = b->q, *f = b->f;
    const double *P = b->m_pre;
    double delta_new = 0;
    OMP(omp parallel for schedule(static) reduction(+ : delta_new))
    for (int i = 0; i < nn; ++i) {
        for (int c = 0; c < 3; ++c) {
            r[3 * i + c] = x[3 * i + c];
            f[3 * i + c] = x[3 * i + c];
            x[3 * i + c] = 0;
            d[3 * i + c] = P[i] * r[3 * i + c];
        }
        delta_new += r[3 * i] * d[3 * i] + r[3 * i + 1] * d[3 * i + 1] + r[3 * i + 2] * d[3 * i + 2];
    }
    for (int it = 0; it < b->p.max_iterations_cg; ++it) {
        double r2 = 0, f2 = 0;
        for (int i = 0; i < nn; ++i) {
            r2 += r[3 * i] * r[3 * i] + r[3 * i + 1] * r[3 * i + 1] + r[3 * i + 2] * r[3 * i + 2];
            f2 += f[3 * i] * f[3 * i] + f[3 * i + 1] * f[3 * i + 1] + f[3 * i + 2] * f[3 * i + 2];
        }
        if ((f2 == 0.0 ? 0.0 : r2 / f2) < b->p.epsilon2) {
            b->last_iters = it;
            return it;
        }
        double dTq = 0;
        for (int i = 0; i < nn; ++i) {
            q[3 * i] = q[3 * i + 1] = q[3 * i + 2] = 0;
            for (int k = b->m_ptr[i]; k < b->m_ptr[i + 1]; ++k) {
                const int j = b->m_col[k];
                for (int c = 0; c < 3; ++c) q[3 * i + c] += b->m_val[k] * d[3 * j + c];
            }
            dTq += d[3 * i] * q[3 * i] + d[3 * i + 1] * q[3 * i + 1] + d[3 * i + 2] * q[3 * i + 2];
        }
        const double alpha = delta_new / dTq;
        for (int i = 0; i < 3 * nn; ++i) x[i] += alpha * d[i];
        for (int i = 0; i < 3 * nn; ++i) r[i] += -alpha * q[i];
        const double delta_old = delta_new;
        delta_new = 0;
        for (int i = 0; i < nn; ++i) {
            for (int c = 0; c < 3; ++c) s[3 * i + c] = P[i] * r[3 * i + c];
            delta_new += r[3 * i] * s[3 * i] + r[3 * i + 1] * s[3 * i + 1] + r[3 * i + 2] * s[3 * i + 2];
        }
        for (int i = 0; i < 3 * nn; ++i) d[i] = s[i] + (delta_new / delta_old) * d[i];
    }
    return -1;
}

/* Blob::make_measurements, kinetic energy (Blob.cpp:1076-1118; apply_element_mass_matrix,
 * tetra_element_linear.cpp:524-542); 0 without mass */
double orc_blob_kinetic_energy(const orc_blob *b) {
    static const double M[4][4] = {{.1, .05, .05, .05}, {.05, .1, .05, .05}, {.05, .05, .1, .05}, {.05, .05, .05, .1}};
    if (!b->mass) return 0.0;
    double kenergy = 0.0;
    for (int e = 0; e < b->ne; ++e) {
        const int *n = b->top + 10 * e;
        double vec[12], tmp[12] = {0};
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 4; ++a) vec[4 * c + a] = b->vel[3 * n[a] + c];
        for (int k = 0; k < 12; k += 4)
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) tmp[i + k] += M[i][j] * vec[j + k];
        double du[12];
        for (int i = 0; i < 12; ++i) du[i] = tmp[i] * (b->vol0[e] * b->rho[e]);
        double t = 0.0;
        for (int i = 0; i < 12; ++i) t = i == 0 ? vec[0] * du[0] : t + vec[i] * du[i];
        kenergy += t;
    }
    return kenergy * 0.5;
}

int orc_blob_add_node_forces(orc_blob *b, int num, const int *node, const double *force) {
    for (int i = 0; i < num; ++i)
        if (node[i] < 0 || node[i] >= b->nt) return 1;
    if (b->next + num > b->next_cap) {
        const int cap = 2 * (b->next + num);
        b->ext_node = realloc(b->ext_node, sizeof(int) * (size_t)cap);
        b->ext_force = realloc(b->ext_force, sizeof(double) * 3 * (size_t)cap);
        b->next_cap = cap;
    }
    for (int i = 0; i < num; ++i) {
        b->ext_node[b->next + i] = node[i];
        for (int c = 0; c < 3; ++c) b->ext_force[3 * (b->next + i) + c] = force[3 * i + c];
    }
    b->next += num;
    return 0;
}

/* World::apply_springs for one spring (World.cpp:4070-4106) */
int orc_spring_forces(const double *x0, const double *x1, double k, double l, double *f0, double *f1) {
    double sep[3], sep_norm[3], sum = 0.0, mag = 0.0;
    for (int c = 0; c < 3; ++c) sep[c] = x1[c] - x0[c];
    for (int c = 0; c < 3; ++c) sum += sep[c] * sep[c];  /* normalize, mat_vec_fns_II.h:188-198 */
    if (sum == 0) return 1;
    sum = sqrt(sum);
    for (int c = 0; c < 3; ++c) sep_norm[c] = sep[c] / sum;
    for (int c = 0; c < 3; ++c) mag += sep[c] * sep[c];  /* magnitude, mat_vec_fns_II.h:326-332 */
    const double force_mag = k * (sqrt(mag) - l);
    for (int c = 0; c < 3; ++c) {
        f0[c] = force_mag * sep_norm[c];
        f1[c] = -1 * force_mag * sep_norm[c];
    }
    return 0;
}

/* World::make_measurements, one spring's energy (World.cpp:4645-4648) */
double orc_spring_energy(const double *x0, const double *x1, double k, double l) {
    double c[3], m = 0.0;
    for (int i = 0; i < 3; ++i) c[i] = x0[i] - x1[i];
    for (int i = 0; i < 3; ++i) m += c[i] * c[i];
    m = sqrt(m);
    return 0.5 * k * (m - l) * (m - l);
}

int orc_blob_step(orc_blob *b) {
    memset(b->node_force, 0, sizeof(double) * 30 * (size_t)b->ne);
    memset(b->force, 0, sizeof(double) * 3 * (size_t)b->nt);
    /* forces added before the step (World::apply_springs, World.cpp:1910), in call order */
    for (int i = 0; i < b->next; ++i)
        for (int c = 0; c < 3; ++c) b->force[3 * b->ext_node[i] + c] += b->ext_force[3 * i + c];
    b->next = 0;
    if (orc_blob_update_internal_forces(b) != 0) return 1;
    if (b->all_pinned) return 0;
    return orc_blob_update_positions(b) ? 2 : 0;
}

int orc_blob_num_nodes(const orc_blob *b) { return b->nt; }
int orc_blob_num_elements(const orc_blob *b) { return b->ne; }
void orc_blob_get_positions(const orc_blob *b, double *out) { memcpy(out, b->pos, sizeof(double) * 3 * b->nt); }
void orc_blob_set_positions(orc_blob *b, const double *in) { memcpy(b->pos, in, sizeof(double) * 3 * b->nt); }
void orc_blob_get_velocities(const orc_blob *b, double *out) { memcpy(out, b->vel, sizeof(double) * 3 * b->nt); }
void orc_blob_set_velocities(orc_blob *b, const double *in) { memcpy(b->vel, in, sizeof(double) * 3 * b->nt); }
void orc_blob_get_rhs(const orc_blob *b, double *out) { memcpy(out, b->rhs, sizeof(double) * 3 * b->nt); }
void orc_blob_get_element_forces(const orc_blob *b, double *out) {
    for (int e = 0; e < b->ne; ++e) memcpy(out + 12 * e, b->node_force + 30 * e, sizeof(double) * 12);
}
int orc_blob_last_cg_iterations(const orc_blob *b) { return b->last_iters; }
void orc_blob_get_rng_state(const orc_blob *b, uint32_t s[6]) { orc_rng_get_state(&b->rng, s); }
void orc_blob_set_rng_state(orc_blob *b, const uint32_t s[6]) { orc_rng_set_seed(&b->rng, s); }
void orc_blob_get_rest_state(const orc_blob *b, double *jinv0, double *vol0) {
    memcpy(jinv0, b->jinv0, sizeof(double) * 9 * b->ne);
    memcpy(vol0, b->vol0, sizeof(double) * b->ne);
}
int orc_blob_matrix_nnz(const orc_blob *b) { return b->nnz; }
void orc_blob_get_matrix(const orc_blob *b, int *rowptr, int *col, double *val) {
    memcpy(rowptr, b->rowptr, sizeof(int) * (b->nrows + 1));
    memcpy(col, b->col, sizeof(int) * b->nnz);
    memcpy(val, b->val, sizeof(double) * b->nnz);
}
void orc_blob_get_inverse_diagonal(const orc_blob *b, double *inv_diag) {
    memcpy(inv_diag, b->inv_diag, sizeof(double) * b->nrows);
}

/* Blob::calculate_strain_energy (Blob.cpp:1052-1064) without touching last_det */
double orc_blob_strain_energy(orc_blob *b) {
    double se = 0.0;
    for (int e = 0; e < b->ne; ++e) {
        double J[3][3], d[12], vol = 0, ld = 0, F[3][3], Ji0[3][3];
        elem_geometry(b, e, J);
        elem_shape(J, d, &vol, &ld);
        memcpy(Ji0, b->jinv0 + 9 * e, sizeof(Ji0));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double acc = 0;
                for (int k = 0; k < 3; ++k) acc += J[k][i] * Ji0[j][k];
                F[i][j] = acc;
            }
        const double dc = F[0][0] * F[0][0] + F[1][1] * F[1][1] + F[2][2] * F[2][2] + F[0][1] * F[0][1] +
                          F[1][0] * F[1][0] + F[0][2] * F[0][2] + F[2][0] * F[2][0] + F[1][2] * F[1][2] +
                          F[2][1] * F[2][1];
        const double C = b->E[e] - b->G[e] * 2.0 / 3.0;
        const double detF = vol / b->vol0[e];
        se += b->vol0[e] * (b->G[e] * (dc - 3) + 0.5 * C * (detF * detF - 1) - ((2 * b->G[e]) + C) * log(detF));
    }
    return 0.5 * se;
}

/* -------- element helpers for closed-form tests -------- */
double orc_element_shape(const double *x4, double *dpsi, double *vol) {
    double J[3][3], ld = 0;
    elem_jacobian(x4, x4 + 3, x4 + 6, x4 + 9, J);
    elem_shape(J, dpsi, vol, &ld);
    return ld;
}

int orc_element_elastic_force(const double *x4, const double *jinv0, double vol0, double G, double E,
                              double *force12) {
    double J[3][3], d[12], vol = 0, ld = 0, s[3][3] = {{0}}, F[3][3], Ji0[3][3], du[12];
    elem_jacobian(x4, x4 + 3, x4 + 6, x4 + 9, J);
    elem_shape(J, d, &vol, &ld);
    memcpy(Ji0, jinv0, sizeof(Ji0));
    elem_shear_stress(J, Ji0, G, vol0, vol, s, F);
    elem_bulk_stress(G, E, vol0, vol, s);
    elem_apply_stress(d, vol, s, du);
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 3; ++c) force12[3 * a + c] = 0.0 - du[4 * c + a];
    return 0;
}

void orc_element_viscosity_matrix(const double *dpsi, double vol, double A, double B, double *V144) {
    elem_viscosity(dpsi, vol, A, B, (double(*)[12])V144);
}

void orc_element_noise_stress(double vol, double A, double B, double kT, double dt, const double *u7,
                              double *stress9) {
    double s[3][3] = {{0}};
    elem_noise_stress(vol, A, B, kT, dt, u7, s);
    memcpy(stress9, s, sizeof(s));
}

/* ---- timing sample for the CPU baseline: one step with the CG capped at
 * `cg_iters` iterations; wall times of the element pass (update_internal_forces),
 * of gather + matrix build, and per CG iteration. Positions are not updated. */
static double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_blob_time_step(orc_blob *b, int cg_iters, double *t_elem, double *t_build, double *t_iter) {
    memset(b->node_force, 0, sizeof(double) * 30 * (size_t)b->ne);
    memset(b->force, 0, sizeof(double) * 3 * (size_t)b->nt);
    double t0 = orc_now();
    int inv = orc_blob_update_internal_forces(b);
    double t1 = orc_now();
    const int nt = b->nt;
    double *F = b->force;
    for (int n = 0; n < nt; ++n)
        for (int m = b->contrib_ptr[n]; m < b->contrib_ptr[n + 1]; ++m) {
            const double *src = b->node_force + 3 * b->contrib[m];
            F[3 * n] += src[0];
            F[3 * n + 1] += src[1];
            F[3 * n + 2] += src[2];
        }
    linearise_force(b);
    for (int k = 0; k < b->npin; ++k)
        for (int c = 0; c < 3; ++c) F[3 * b->pinned[k] + c] = 0;
    matrix_build(b);
    double t2 = orc_now();
    const int keep = b->p.max_iterations_cg;
    b->p.max_iterations_cg = cg_iters;
    double *x = malloc(sizeof(double) * 3 * (size_t)nt);
    memcpy(x, F, sizeof(double) * 3 * (size_t)nt);
    double t3 = orc_now();
    int it = cg_solve(b, x);
    double t4 = orc_now();
    b->p.max_iterations_cg = keep;
    free(x);
    *t_elem = t1 - t0;
    *t_build = t2 - t1;
    *t_iter = (t4 - t3) / (it > 0 ? it : cg_iters);
    return inv;
}

/* element viscosity matrices of the last build, [ne][12][12] */
void orc_blob_get_element_matrices(const orc_blob *b, double *out) {
    memcpy(out, b->V, sizeof(double) * 144 * (size_t)b->ne);
}

/* pinned flags per node */
void orc_blob_get_pinned_flags(const orc_blob *b, int *out) { memcpy(out, b->is_pinned, sizeof(int) * b->nt); }
