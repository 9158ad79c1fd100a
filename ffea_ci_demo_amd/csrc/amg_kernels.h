// Device code of the multigrid preconditioner (MI355X, gfx950). Included by ffea_blob.hip
// after kernels.h. A single blob's kernels return at once when its solve has stopped. An
// ensemble's levels are the blobs' own levels side by side (block-diagonal operators);
// its coarse-level kernels run for every blob (a stopped blob's CG no longer reads z),
// its level-0 kernels skip the chunks of stopped blobs, and the smoother weights are the
// blobs' own: per row on level 0, folded into the stored diagonal above.
//
// One V-cycle z = M r, with level 0 the corner-node viscosity matrix K (3x3 blocks, the
// CG's own sliced-ELL storage) and levels 1.. the Galerkin operators over aggregates
// (6x6 blocks). The same smoothing runs before and after each coarse correction, which
// keeps M symmetric: damped steps x += w B^-1 (b - A x), B the 3x3 diagonal blocks on
// level 0 (k_mg_fine) and the diagonal on coarser levels, where a step may also carry a
// polynomial smoother's alpha x + beta xold terms (k_mg_csr). Level-0 vectors are the
// CG's 64-row tiled FP64 vectors; coarse vectors are FP64 [rows][6]. The preconditioner's
// matrices - a copy of K written by the assembly, the transfers and the coarse operators -
// are FP16 scaled by a power of two per matrix (largest entry near 2^10; a sum is
// multiplied back by the exact inverse), with FP64 accumulation: M only has to be a
// fixed symmetric positive operator, and half the bytes make the V-cycle cheaper.
// Prolongations (few blocks per row) are in a sliced ELL: 64 block rows per slice, slot
// width W = the slice's longest row, block column at base + j*64 + l (-1: padding); a
// slot's values are lane-major vectors (ell_val_index), so a lane reads its block's
// BR*BC FP16 values in 16-byte loads.
#pragma once
#include "kernels.h"

namespace ffk {

constexpr int kMgBlock = 256;  // threads per workgroup of the coarse-level kernels

// a coarse-level kernel has nothing to do once the solve it serves has stopped: blob 0's
// (S) for a single blob, every blob's (num_done of active) for an ensemble
struct MgStop {
    const CgScalars *S;
    const int *num_done;
    int active;
    __device__ __forceinline__ bool stopped() const {
        return S ? S->done != 0 : (num_done && *num_done >= active);
    }
};

// vector layouts: level 0 tiled by 64 rows (vi), coarse levels plain [rows][6]
template <int B>
struct VecLayout;
template <>
struct VecLayout<3> {
    static __device__ __forceinline__ size_t at(int n, int k) { return vi(n, k); }
};
template <>
struct VecLayout<6> {
    static __device__ __forceinline__ size_t at(int n, int k) { return (size_t)n * 6 + k; }
};

enum MgMode { MG_RESID = 0, MG_RESTRICT = 1, MG_PROLONG = 2, MG_POST = 3 };

// Value k of the BB-value block of lane l in the ELL slot starting at `slot` (= base +
// j*64): the first BB/8*8 values as vectors of 8 (vector t of lanes 0..63 contiguous),
// the rest (BB % 8) as one vector per lane after them. Each vector load of a wavefront
// is one contiguous span.
__host__ __device__ __forceinline__ size_t ell_val_index(long long slot, int bb, int lane, int k) {
    const size_t b = (size_t)slot * bb;
    const int full = bb / 8 * 8, rem = bb - full;
    return k < full ? b + (size_t)(k / 8) * 8 * 64 + (size_t)lane * 8 + (k % 8)
                    : b + (size_t)full * 64 + (size_t)lane * rem + (k - full);
}

template <int BB>
__device__ __forceinline__ void load_ell_block(const _Float16 *val, long long slot, int lane, float (&w)[BB]) {
    typedef _Float16 vh8 __attribute__((ext_vector_type(8)));
    constexpr int full = BB / 8 * 8, rem = BB - full;
    static_assert(rem == 0 || rem == 2 || rem == 4, "ELL blocks: a remainder of 2 or 4 values");
    const _Float16 *b = val + (size_t)slot * BB;
#pragma unroll
    for (int t = 0; t < BB / 8; ++t) {
        const vh8 u = __builtin_nontemporal_load((const vh8 *)(b + (size_t)t * 8 * 64) + lane);
#pragma unroll
        for (int i = 0; i < 8; ++i) w[8 * t + i] = (float)u[i];
    }
    if constexpr (rem == 4) {
        typedef _Float16 vh4 __attribute__((ext_vector_type(4)));
        const vh4 u = __builtin_nontemporal_load((const vh4 *)(b + (size_t)full * 64) + lane);
#pragma unroll
        for (int i = 0; i < 4; ++i) w[full + i] = (float)u[i];
    } else if constexpr (rem == 2) {
        typedef _Float16 vh2 __attribute__((ext_vector_type(2)));
        const vh2 u = __builtin_nontemporal_load((const vh2 *)(b + (size_t)full * 64) + lane);
        w[full] = (float)u.x;
        w[full + 1] = (float)u.y;
    }
}

struct MgArgs {
    MgStop stop;
    int n;                  // block rows
    const int *sw;          // [slices] slot widths
    const long long *sb;    // [slices] slot bases
    const int *col;         // [slots]
    const _Float16 *val;    // [slots * BR*BC], scaled
    double mscale;          // inverse of the values' scale
    const double *x;        // input gathered by block column
    const double *b;        // RESID, POST: right-hand side (row layout)
    const double *base;     // PROLONG: y = bscale base + A x (row layout)
    const double *dinv;     // RESTRICT: the next level's D^-1; POST: this level's (row layout)
    double omega, bscale;
    double *y, *y2;         // RESTRICT: y = A x, y2 = omega D^-1 y (if not null)
    const double *bscale_row = nullptr;  // PROLONG on level 0 of an ensemble: bscale per row
    float *y32 = nullptr;   // PROLONG on level 0 of the FP32-mixed form: y's packed FP32 copy
};

// y = A x and the mode's row update, one thread per block row:
//   RESID     y = b - A x
//   RESTRICT  y = A x (A = P^T), y2 = omega D^-1 y (the pre-smoothing step on the next level)
//   PROLONG   y = bscale base + A x (A = P)
//   POST      y = x + omega D^-1 (b - A x) (the post-smoothing step; A square)
template <int BR, int BC, int MODE>
__global__ __launch_bounds__(kMgBlock) void k_mg_apply(MgArgs a) {
    if (a.stop.stopped()) return;
    const int row = blockIdx.x * kMgBlock + threadIdx.x;
    if (row >= a.n) return;
    const int sl = row >> 6, lane = row & 63;
    const int W = __builtin_amdgcn_readfirstlane(a.sw[sl]);  // wave-uniform: a slice is one wavefront's rows
    const long long base = a.sb[sl];
    double acc[BR];
#pragma unroll
    for (int p = 0; p < BR; ++p) acc[p] = 0.0;
    // groups of G slots issue all their loads before any product; padding slots (column -1)
    // hold zero blocks and gather row 0, slots past W repeat slot W-1 with a zero x: both add
    // +-0 and leave the sums unchanged
    constexpr int G = BR * BC <= 18 ? 4 : 2;
    for (int g = 0; g < W; g += G) {
        int c[G];
#pragma unroll
        for (int h = 0; h < G; ++h) {
            const int j = g + h < W ? g + h : W - 1;
            const int cj = a.col[base + (long long)j * 64 + lane];
            c[h] = cj < 0 ? 0 : cj;
        }
        double xv[G][BC];
        float w[G][BR * BC];
#pragma unroll
        for (int h = 0; h < G; ++h) {
            const int j = g + h < W ? g + h : W - 1;
            if constexpr (BC == 6) {  // a coarse row [n][6]: three 16-byte loads
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const vd2 u = ((const vd2 *)(a.x + (size_t)c[h] * 6))[t];
                    xv[h][2 * t] = u.x;
                    xv[h][2 * t + 1] = u.y;
                }
            } else {
#pragma unroll
                for (int q = 0; q < BC; ++q) xv[h][q] = a.x[VecLayout<BC>::at(c[h], q)];
            }
            load_ell_block<BR * BC>(a.val, base + (long long)j * 64, lane, w[h]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int h = 0; h < G; ++h) {
            const bool live = g + h < W;
#pragma unroll
            for (int p = 0; p < BR; ++p)
#pragma unroll
                for (int q = 0; q < BC; ++q) acc[p] += (double)w[h][p * BC + q] * (live ? xv[h][q] : 0.0);
        }
    }
#pragma unroll
    for (int p = 0; p < BR; ++p) {
        acc[p] *= a.mscale;
        const size_t o = VecLayout<BR>::at(row, p);
        if constexpr (MODE == MG_RESID) {
            a.y[o] = a.b[o] - acc[p];
        } else if constexpr (MODE == MG_RESTRICT) {
            a.y[o] = acc[p];
            if (a.y2) a.y2[o] = a.omega * a.dinv[o] * acc[p];
        } else if constexpr (MODE == MG_PROLONG) {
            const double yv = (a.bscale_row ? a.bscale_row[row] : a.bscale) * a.base[o] + acc[p];
            a.y[o] = yv;
            acc[p] = yv;  // for the FP32 copy
        } else {
            a.y[o] = a.x[o] + a.omega * a.dinv[o] * (a.b[o] - acc[p]);
        }
    }
    if constexpr (MODE == MG_PROLONG && BR == 3)
        if (a.y32) store_f32x3(a.y32, row, acc);
}

// The same operations for matrices with few, long rows (the coarse operators, 6x6 blocks
// and ~45 per row, and the restrictions, ~100 blocks per row): block CSR and one
// wavefront per block row. Lane l takes the row's blocks l, l+64, ...; the lanes'
// partial sums are combined by a fixed butterfly, so results do not depend on timing.
// A row's W blocks are stored vector-major: with V-value vectors (V = 4 for 6x6 blocks,
// 2 for 6x3), vector t of block j is at (rp[row] * BR*BC) + (t W + j) V, so one vector
// load of a wavefront reads consecutive blocks' vectors (contiguous bytes).
struct MgCsrArgs {
    MgStop stop;
    int n;
    const int *rp, *ci;
    const _Float16 *val;    // [nnzb][BR*BC], scaled
    double mscale;          // inverse of the values' scale
    const double *x, *b, *dinv;
    double omega;
    double *y, *y2;
    const float *x32 = nullptr;  // level-0 restriction of the FP32-mixed form: x's packed FP32 copy, gathered
    // POST as a step of a polynomial smoother: y = alpha x + beta xold + omega D^-1 (b - A x)
    // (xold read at the row itself only, so y may overwrite it); alpha 1 and no xold is the
    // damped-Jacobi sweep
    double alpha = 1.0, beta = 0.0;
    const double *xold = nullptr;
};

template <int BR, int BC, int MODE>
__global__ __launch_bounds__(kMgBlock) void k_mg_csr(MgCsrArgs a) {
    if (a.stop.stopped()) return;
    const int row = blockIdx.x * (kMgBlock / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= a.n) return;  // whole wavefronts
    double acc[BR];
#pragma unroll
    for (int p = 0; p < BR; ++p) acc[p] = 0.0;
    const int r0 = a.rp[row], W = a.rp[row + 1] - r0;
    const _Float16 *rowv = a.val + (size_t)r0 * (BR * BC);
    for (int j = lane; j < W; j += 64) {
        const int c = a.ci[r0 + j];
        double xv[BC];
        if constexpr (BC == 3) {
            if (a.x32) gather3(a.x32, c, xv);
            else gather3(a.x, c, xv);
        } else {
#pragma unroll
            for (int q = 0; q < BC; ++q) xv[q] = a.x[VecLayout<BC>::at(c, q)];
        }
        float w[BR * BC];
        if constexpr ((BR * BC) % 4 == 0) {
            typedef _Float16 vh4 __attribute__((ext_vector_type(4)));
#pragma unroll
            for (int t = 0; t < BR * BC / 4; ++t) {
                const vh4 u = __builtin_nontemporal_load((const vh4 *)rowv + (size_t)t * W + j);
                w[4 * t] = u.x; w[4 * t + 1] = u.y; w[4 * t + 2] = u.z; w[4 * t + 3] = u.w;
            }
        } else {
            typedef _Float16 vh2 __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int t = 0; t < BR * BC / 2; ++t) {
                const vh2 u = __builtin_nontemporal_load((const vh2 *)rowv + (size_t)t * W + j);
                w[2 * t] = u.x; w[2 * t + 1] = u.y;
            }
        }
#pragma unroll
        for (int p = 0; p < BR; ++p)
#pragma unroll
            for (int q = 0; q < BC; ++q) acc[p] += (double)w[p * BC + q] * xv[q];
    }
#pragma unroll
    for (int p = 0; p < BR; ++p) acc[p] = wave_sum(acc[p]) * a.mscale;
    double mine = acc[0];
#pragma unroll
    for (int p = 1; p < BR; ++p)
        if (lane == p) mine = acc[p];
    if (lane >= BR) return;
    const size_t o = VecLayout<BR>::at(row, lane);
    if constexpr (MODE == MG_RESID) {
        a.y[o] = a.b[o] - mine;
    } else if constexpr (MODE == MG_RESTRICT) {
        a.y[o] = mine;
        if (a.y2) a.y2[o] = a.omega * a.dinv[o] * mine;
    } else {
        static_assert(MODE == MG_POST, "k_mg_csr: residual, restriction or post-smoothing");
        if (a.xold == nullptr && a.alpha == 1.0) {
            a.y[o] = a.x[o] + a.omega * a.dinv[o] * (a.b[o] - mine);
        } else {
            const double xo = a.xold ? a.beta * a.xold[o] : 0.0;
            a.y[o] = a.alpha * a.x[o] + xo + a.omega * a.dinv[o] * (a.b[o] - mine);
        }
    }
}

// coarsest level: e = C b with the dense inverse C (m x m, FP32, symmetric: row `row`
// is read as contiguous bytes, rows padded with zeros to dense_ld(m), a multiple of 4
// floats, so a lane reads 16-byte vectors), one wavefront per row. An ensemble has one
// inverse per blob (row_blob non-null): row `row` belongs to blob q = row_blob[row],
// whose rows are [off[q], off[q] + m[q]) and whose inverse starts at C + coff[q].
__host__ __device__ __forceinline__ int dense_ld(int m) { return (m + 3) / 4 * 4; }
struct MgDenseArgs {
    MgStop stop;
    int n;
    const float *C;
    const double *b;
    double *e;
    const int *row_blob, *off, *m;
    const long long *coff;
};
__global__ __launch_bounds__(kMgBlock) void k_mg_dense(MgDenseArgs a) {
    if (a.stop.stopped()) return;
    const int row = blockIdx.x * (kMgBlock / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= a.n) return;
    int o = 0, m = a.n;
    const float *C = a.C;
    if (a.row_blob) {
        const int q = a.row_blob[row];
        o = a.off[q];
        m = a.m[q];
        C += a.coff[q];
    }
    const int ld = dense_ld(m);
    const vf4 *Cr = (const vf4 *)(C + (size_t)(row - o) * ld);
    const double *b = a.b + o;
    double acc = 0.0;
#pragma unroll 4
    for (int k4 = lane; k4 < ld / 4; k4 += 64) {
        const vf4 c = __builtin_nontemporal_load(Cr + k4);
        const int k = 4 * k4;
        // the padding's zeros multiply nothing: past m no b is read
        acc += (double)c.x * b[k];
        if (k + 1 < m) acc += (double)c.y * b[k + 1];
        if (k + 2 < m) acc += (double)c.z * b[k + 2];
        if (k + 3 < m) acc += (double)c.w * b[k + 3];
    }
    acc = wave_sum(acc);
    if (lane == 0) a.e[row] = acc;
}

// Level-0 steps on the CG's own matrix storage, one chunk of rows per workgroup
// (kRowBlock threads, one row each):
//   FINE_RESID  y = r - omega K z          (z = D^-1 r from the vector step: the
//                                           pre-smoothed iterate is omega z)
//   FINE_POST   z = xb + omega D^-1 (r - K xb), and the chunk partials of r.z
constexpr int MG_FINE_RESID = 0, MG_FINE_POST = 1;

template <typename T>
struct MgFineArgs {
    const CgScalars *S;    // [nblob]: a chunk of a stopped blob is skipped
    const Chunk *chunks;
    int nchunks;
    const int *slice_w;
    const long long *slice_base;
    const int *col;
    const T *val;          // the assembly's scaled FP16 copy of K (the CG's slots)
    double mscale;         // inverse of its scale
    const double *x;       // z (RESID) or xb (POST): gathered (FP64 form), own row
    const double *r, *dinv;
    double omega;
    double *y;             // RESID: residual; POST: z
    double *P;             // POST: partials row (P[c])
    const double *omega_row = nullptr;  // an ensemble's smoother weight per row (else omega)
    const float *x32 = nullptr;         // FP32-mixed form: x's packed FP32 copy, gathered
    float *y32 = nullptr;               // FP32-mixed form: y's packed FP32 copy (the residual, z)
    const float *binv = nullptr;        // POST: the rows' diagonal-block inverses (B^-1 in place of D^-1)
};

// blocks whose loads one row issues together (slice_row_times)
#ifndef FFEA_MG_FINE_GROUP
#define FFEA_MG_FINE_GROUP 2
#endif

// BS threads per chunk: kRowBlock for chunks of up to 1024 rows, one row a thread; an
// ensemble's mostly 256-row chunks take kMgBlock, looping over longer chunks
// V: the gathered vector's type, double (FP64 form: x) or float (FP32-mixed form: x32)
template <typename T, int MODE, int BS = kRowBlock, typename V = double>
__global__ __launch_bounds__(BS) void k_mg_fine(MgFineArgs<T> a) {
    const int c = blockIdx.x;
    const Chunk ch = a.chunks[c];
    if (a.S[ch.blob].done) return;
    double v[1] = {0.0};
    for (int n = ch.begin + threadIdx.x; n < ch.end; n += BS) {
        const double omega = a.omega_row ? a.omega_row[n] : a.omega;
        const int sl = n / kSlice, lane = n % kSlice;
        const int W = __builtin_amdgcn_readfirstlane(a.slice_w[sl]);  // wave-uniform
        const long long base = a.slice_base[sl];
        double acc[3] = {0.0, 0.0, 0.0};
        if constexpr (std::is_same<V, float>::value)
            slice_row_times<T, 0, FFEA_MG_FINE_GROUP>(a.col + base + lane, a.val + (size_t)base * 9, lane, W, a.x32, acc);
        else
            slice_row_times<T, 0, FFEA_MG_FINE_GROUP>(a.col + base + lane, a.val + (size_t)base * 9, lane, W, a.x, acc);
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] *= a.mscale;
        double z3[3];
        if constexpr (MODE == MG_FINE_RESID) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const size_t o = vi(n, k);
                z3[k] = a.r[o] - omega * acc[k];
                a.y[o] = z3[k];
            }
        } else {
            double r3[3], d3[3], s3[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                r3[k] = a.r[vi(n, k)];
                d3[k] = r3[k] - acc[k];
            }
            jacobi_times(a.dinv, a.binv, n, d3, s3);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const size_t o = vi(n, k);
                const double zk = a.x[o] + omega * s3[k];
                a.y[o] = zk;
                z3[k] = zk;
                v[0] += r3[k] * zk;
            }
        }
        if (a.y32) store_f32x3(a.y32, n, z3);
    }
    if constexpr (MODE == MG_FINE_POST) store_partials<1, BS>(v, a.P, a.nchunks, c);
}

// D-weighted rigid-mode gauge of a free blob's solution. Jacobi-PCG from x = 0 returns
// the solution with Z^T D x = 0 (its iterates lie in D^-1 range(K)); the multigrid
// preconditioner does not keep that, so after the solve x <- x - Z (Z^T D Z)^-1 Z^T D x
// with D the matrix diagonal. Partials per chunk: Z^T D x (6) and the upper triangle
// of Z^T D Z (21).
struct MgRigidArgs {
    const CgScalars *S;     // [nblob]
    const Chunk *chunks;
    int nchunks;
    const double *x, *dinv, *pos;
    const double3 *centre;  // [nblob] rigid-mode centres
    double *P;      // [27][nchunks]
    double *coef;   // [nblob][6]
    double *xout;   // x (corrected in place)
    const int2 *seg;        // [nblob] chunk range of each blob
    const int *free_blob;   // [nblob] 1: a free blob, whose solution is put in the gauge
};

__device__ __forceinline__ void rigid_basis(const double *d, double (&Z)[3][6]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 6; ++k) Z[i][k] = 0.0;
    Z[0][0] = Z[1][1] = Z[2][2] = 1.0;
    // column 3+k: e_k cross d
    Z[1][3] = -d[2]; Z[2][3] = d[1];
    Z[0][4] = d[2];  Z[2][4] = -d[0];
    Z[0][5] = -d[1]; Z[1][5] = d[0];
}

// 256 threads per chunk (4 rows each): the 27 sums stay in registers
__global__ __launch_bounds__(kMgBlock) void k_mg_rigid_sums(MgRigidArgs a) {
    const int c = blockIdx.x;
    const Chunk ch = a.chunks[c];
    const CgScalars &S = a.S[ch.blob];
    if (!S.converged || S.updated || !a.free_blob[ch.blob]) return;
    const double3 cen = a.centre[ch.blob];
    double v[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) v[k] = 0.0;
    for (int n = ch.begin + threadIdx.x; n < ch.end; n += kMgBlock) {
        const double d[3] = {a.pos[3 * n] - cen.x, a.pos[3 * n + 1] - cen.y, a.pos[3 * n + 2] - cen.z};
        double Z[3][6];
        rigid_basis(d, Z);
        double w[3], xd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double di = a.dinv[vi(n, i)];
            w[i] = di > 0.0 ? 1.0 / di : 0.0;
            xd[i] = w[i] * a.x[vi(n, i)];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) v[k] += Z[0][k] * xd[0] + Z[1][k] * xd[1] + Z[2][k] * xd[2];
        int t = 6;
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int l = k; l < 6; ++l) v[t++] += w[0] * Z[0][k] * Z[0][l] + w[1] * Z[1][k] * Z[1][l] + w[2] * Z[2][k] * Z[2][l];
    }
    store_partials<27, kMgBlock>(v, a.P, a.nchunks, c);
}

// one workgroup per blob
__global__ __launch_bounds__(kMgBlock) void k_mg_rigid_coef(MgRigidArgs a) {
    const int q = blockIdx.x;
    const CgScalars &S = a.S[q];
    if (!S.converged || S.updated || !a.free_blob[q]) return;
    double t[27];
    sum_partials<27, kMgBlock>(a.P, a.nchunks, a.seg[q].x, a.seg[q].y, t);
    if (threadIdx.x == 0) {
        double M[36], Mi[36];
        int u = 6;
        for (int k = 0; k < 6; ++k)
            for (int l = k; l < 6; ++l) {
                M[6 * k + l] = M[6 * l + k] = t[u];
                ++u;
            }
        invert6(M, Mi);
        for (int k = 0; k < 6; ++k) {
            double acc = 0.0;
            for (int l = 0; l < 6; ++l) acc += Mi[6 * k + l] * t[l];
            a.coef[6 * q + k] = acc;
        }
    }
}

__global__ __launch_bounds__(kRowBlock) void k_mg_rigid_fix(MgRigidArgs a) {
    const Chunk ch = a.chunks[blockIdx.x];
    const CgScalars &S = a.S[ch.blob];
    if (!S.converged || S.updated || !a.free_blob[ch.blob]) return;
    const int n = ch.begin + threadIdx.x;
    if (n >= ch.end) return;
    const double3 cen = a.centre[ch.blob];
    const double d[3] = {a.pos[3 * n] - cen.x, a.pos[3 * n + 1] - cen.y, a.pos[3 * n + 2] - cen.z};
    double cf[6], zc[3];
#pragma unroll
    for (int k = 0; k < 6; ++k) cf[k] = a.coef[6 * ch.blob + k];
    z_times(d, cf, zc);
#pragma unroll
    for (int i = 0; i < 3; ++i) a.xout[vi(n, i)] -= zc[i];
}

}  // namespace ffk
