// ORACLE — test infrastructure only.
//
// C entry points over the reference's own translation units (RngStream.cpp,
// SparsityPattern.cpp, SparseMatrixFixedPattern.cpp, SparseMatrixTypes.cpp,
// mat_vec_fns.cpp, and the templates of mat_vec_fns_II.h), compiled from the reference tree into oracle/_ref by
// oracle/Makefile. Tests use it to check the C restatement in ffea_oracle.c
// draw-for-draw and bit-for-bit. Nothing here ships or runs on the GPU path.
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "RngStream.h"
#include "SparsityPattern.h"
#include "SparseMatrixFixedPattern.h"
#include "mat_vec_fns.h"
#include "mat_vec_fns_II.h"

extern "C" {

// Stream `stream_index` after setting the package seed (World.cpp:281 then the
// per-thread construction of World.cpp:288), advanced `substreams` substreams,
// then `n` RandU01 draws.
int ref_rng_stream_draws(const uint32_t package_seed[6], int stream_index, int substreams, int n,
                         double *out, uint32_t state_after[6]) {
    if (!RngStream::SetPackageSeed(package_seed)) return -1;
    std::vector<RngStream> streams(stream_index + 1);
    RngStream &g = streams[stream_index];
    for (int i = 0; i < substreams; ++i) g.ResetNextSubstream();
    for (int i = 0; i < n; ++i) out[i] = g.RandU01();
    g.GetState(state_after);
    return 0;
}

// SetSeed, AdvanceState(e, c), then n draws.
int ref_rng_advance_draws(const uint32_t seed[6], int e, int c, int n, double *out, uint32_t state_after[6]) {
    RngStream g;
    if (!g.SetSeed(seed)) return -1;
    g.AdvanceState(e, c);
    for (int i = 0; i < n; ++i) out[i] = g.RandU01();
    g.GetState(state_after);
    return 0;
}

void ref_mat3_invert(const double *m9, double *mi9, double *det) {
    matrix3 m, mi;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = m9[3 * i + j];
    mat3_invert(m, mi, det);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) mi9[3 * i + j] = mi[i][j];
}

// The mass-less solver's fixed-pattern viscosity matrix for given element
// matrices V[ne][12][12]: sources registered with the loop of
// NoMassCGSolver.cpp:81-127 (written out here), then the reference's
// create_sparse_matrix / build / calc_inverse_diagonal / apply.
// Call with rowptr == nullptr to get nnz only.
int ref_nomass_matrix(int nt, int ne, const int *top, const int *is_pinned, const double *V,
                      const double *stokes_drag, int *rowptr, int *col, double *val, double *inv_diag,
                      const double *x_in, double *y_out) {
    SparsityPattern sp;
    const int num_rows = 3 * nt;
    sp.init(num_rows);
    double one = 1;
    std::vector<double> Vcopy(V, V + (size_t)144 * ne);
    std::vector<double> drag(nt, 0.0);
    if (stokes_drag) std::memcpy(drag.data(), stokes_drag, sizeof(double) * nt);
    for (int n = 0; n < ne; ++n) {
        for (int ni = 0; ni < 10; ++ni)
            for (int nj = 0; nj < 10; ++nj) {
                const int a = top[10 * n + ni], b = top[10 * n + nj];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        const int r = 3 * a + i, c = 3 * b + j;
                        if (!is_pinned[a] && !is_pinned[b]) {
                            if (ni < 4 && nj < 4) {
                                sp.register_contribution(r, c, &Vcopy[(size_t)144 * n + (ni + 4 * i) * 12 + nj + 4 * j]);
                            } else if (ni == nj && i == j && !sp.check_for_contribution(r, c)) {
                                sp.register_contribution(r, c, &one);
                            }
                        } else if (ni == nj && i == j && !sp.check_for_contribution(r, c)) {
                            sp.register_contribution(r, c, &one);
                        }
                    }
            }
    }
    if (stokes_drag)
        for (int n = 0; n < nt; ++n)
            if (!is_pinned[n])
                for (int j = 0; j < 3; ++j) sp.register_contribution(3 * n + j, 3 * n + j, &drag[n]);
    std::shared_ptr<SparseMatrixFixedPattern> M = sp.create_sparse_matrix();
    const int nnz = M->get_num_nonzero_elements();
    if (!rowptr) return nnz;
    M->build();
    std::vector<scalar> inv(num_rows);
    M->calc_inverse_diagonal(inv);
    std::vector<sparse_entry> &ent = M->get_entries();
    std::vector<int> &key = M->get_key();
    for (int r = 0; r <= num_rows; ++r) rowptr[r] = key[r];
    for (int k = 0; k < nnz; ++k) {
        col[k] = ent[k].column_index;
        val[k] = ent[k].val;
    }
    for (int r = 0; r < num_rows; ++r) inv_diag[r] = inv[r];
    if (x_in && y_out) {
        std::vector<arr3> xin(nt), yout(nt);
        for (int n = 0; n < nt; ++n)
            for (int c = 0; c < 3; ++c) xin[n][c] = x_in[3 * n + c];
        M->apply(xin, yout);
        for (int n = 0; n < nt; ++n)
            for (int c = 0; c < 3; ++c) y_out[3 * n + c] = yout[n][c];
    }
    return nnz;
}

// One rotational constant force as Blob::apply_ctforces CTF 2 forms it
// (Blob.cpp:1693-1714, axis given by point and direction): the reference's
// intersectingPointToLine, sub, arr3arr3VectorProduct, magnitude and resize,
// called in that order. torque: 0 'f', 1 't'.
void ref_rotational_force(const double *pos3, const double *axis6, double F, int torque, double *f3) {
    arr3 p{pos3[0], pos3[1], pos3[2]}, a{axis6[0], axis6[1], axis6[2]}, d{axis6[3], axis6[4], axis6[5]};
    arr3 r, f;
    intersectingPointToLine(p, a, d, r);
    sub(r, p, r);
    arr3arr3VectorProduct(r, d, f);
    if (torque) {
        scalar rsize = magnitude(r);
        resize(F / rsize, f);
    } else {
        resize(F, f);
    }
    for (int c = 0; c < 3; ++c) f3[c] = f[c];
}

// One spring as World::apply_springs forms it (World.cpp:4070-4106) with the reference's
// normalize (throws for a zero vector: returns 1 here) and magnitude.
int ref_spring_forces(const double *x0, const double *x1, double k, double l, double *f0, double *f1) {
    arr3 n0{x0[0], x0[1], x0[2]}, n1{x1[0], x1[1], x1[2]}, sep, sep_norm;
    sep[0] = n1[0] - n0[0];
    sep[1] = n1[1] - n0[1];
    sep[2] = n1[2] - n0[2];
    try {
        normalize(sep, sep_norm);
    } catch (const FFEAException &) {
        return 1;
    }
    const scalar force_mag = k * (magnitude(sep) - l);
    for (int c = 0; c < 3; ++c) {
        f0[c] = force_mag * sep_norm[c];
        f1[c] = -1 * force_mag * sep_norm[c];
    }
    return 0;
}

// World::make_measurements' term for one spring (World.cpp:4645-4648)
double ref_spring_energy(const double *x0, const double *x1, double k, double l) {
    arr3 a{x0[0], x0[1], x0[2]}, b{x1[0], x1[1], x1[2]}, c;
    sub(a, b, c);
    return 0.5 * k * (magnitude(c) - l) * (magnitude(c) - l);
}

}  // extern "C"
