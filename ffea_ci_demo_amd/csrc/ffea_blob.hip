// MI355X (gfx950) blob time step: element forces, node gather, viscosity matrix
// assembly, Jacobi-PCG, Euler update — behind the C ABI of include/ffea_gpu.h.
//
// Data layout in HBM (one blob):
//   corner ("linear") nodes are renumbered 0..Nl-1 in ascending .node order;
//   positions, velocities and all CG vectors are FP64 [Nl][3];
//   element data are structure-of-arrays over elements (conn [4][Ne], J^-1_0 [9][Ne],
//   vol_0, G, E, A, B, last_det) so one-element-per-lane loads coalesce;
//   element nodal forces are [Ne][12] (4 nodes x 3), the shape derivatives kept for
//   the matrix are [10][Ne] (the 9 independent dpsi entries and the volume);
//   the viscosity matrix is 3x3 blocks over corner nodes in a sliced-ELL layout:
//   64 rows per slice, per slice width W = its longest row, values
//   [slot_base*9 + (j*9 + k)*64 + lane] and block columns [slot_base + j*64 + lane],
//   so lane-per-row loads of one (j, k) are 512 contiguous bytes.
//
// The node-force scatter is a gather over a node->element CSR in element order
// (no atomics): results are bit-reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <type_traits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ffea_gpu.h"
#include "mrg32k3a.h"

namespace {

constexpr int kBlock = 256;   // threads per workgroup: 4 wavefronts of 64
constexpr int kSlice = 64;    // rows per ELL slice (one wavefront)
constexpr uint32_t kIdentity = 0xFFFFFFFFu;
constexpr int kMaxUnrolledRow = 16;  // rows up to this many blocks take the unrolled SpMV path

// CG node vectors are tiled by 64 nodes: x[64] y[64] z[64] per tile, so row-wise
// loads and stores of one component are 512 contiguous bytes and stencil gathers of
// neighbouring rows stay coalesced
__host__ __device__ __forceinline__ size_t vi(int n, int k) {
    return ((size_t)(n >> 6) * 3 + (size_t)k) * 64 + (size_t)(n & 63);
}

thread_local std::string g_last_error;

#define FFEA_HIP_TRY(expr)                                                                 \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) throw HipError(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct HipError {
    std::string msg;
    explicit HipError(std::string m) : msg(std::move(m)) {}
};
struct Invalid {
    std::string msg;
    explicit Invalid(std::string m) : msg(std::move(m)) {}
};

std::string fmt(const char *f, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

// scalars shared by the CG kernels (device memory). Every reduction is finished by
// the kernel that consumes it: each block of that kernel adds the producer's
// per-block partials in the same fixed order, so all blocks hold bit-identical
// totals and no atomics or cross-block fences are needed.
struct CgScalars {
    double delta[2];   // delta[(i-1)&1]: r.z entering iteration i (written by the SpMV step)
    double f2;         // |f|^2 of the right-hand side
    double G[36];      // (Z^T Z)^-1 for the rigid-mode projection of the mixed form
    int iter;          // iterations done when the solve stopped
    int done;          // 1: stop (converged, failed or skipped)
    int converged;     // 1: converged this solve
    int inverted;      // inverted elements counted by the element pass (relaxed atomics)
    int inverted_step; // inverted elements of the step being solved
    int pad;
};

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block sum of NV values; result valid in every thread
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[kBlock / 64][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh[wid][k] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s += sh[w][k];
        v[k] = s;
    }
    __syncthreads();
}

// per-block partials, stored [k][nb]
template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *P, int nb, int L) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) P[(size_t)k * nb + L] = v[k];
}

// totals of partials written by the previous kernel: identical in every block
template <int NV>
__device__ __forceinline__ void sum_partials(const double *P, int nb, double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += kBlock) acc += P[(size_t)k * nb + b];
        out[k] = acc;
    }
    block_sum<NV>(out);
}

// Row range of a block in the row kernels: a fixed grid of nb blocks, each owning a
// contiguous range of rows; blocks are mapped so that each of the 8 XCDs (which
// receive workgroups round-robin) owns one contiguous eighth of the rows.
struct RowRange {
    int L, begin, end;
};
__device__ __forceinline__ RowRange row_range(int nrows, int rows_per_block) {
    const int nb = gridDim.x, b = blockIdx.x;
    const int L = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int begin = L * rows_per_block;
    const int end = min(nrows, begin + rows_per_block);
    return {L, begin, end};
}

// rigid-mode basis of a corner node at x (relative to a fixed centre): columns are
// the 3 translations and the 3 infinitesimal rotations e_k x x. Z^T v = (v, x cross v).
__device__ __forceinline__ void zt_times(const double *x, const double *v, double *out6) {
    out6[0] = v[0];
    out6[1] = v[1];
    out6[2] = v[2];
    out6[3] = x[1] * v[2] - x[2] * v[1];
    out6[4] = x[2] * v[0] - x[0] * v[2];
    out6[5] = x[0] * v[1] - x[1] * v[0];
}
// Z c = c_t + c_r cross x
__device__ __forceinline__ void z_times(const double *x, const double *c, double *out3) {
    out3[0] = c[0] + (c[4] * x[2] - c[5] * x[1]);
    out3[1] = c[1] + (c[5] * x[0] - c[3] * x[2]);
    out3[2] = c[2] + (c[3] * x[1] - c[4] * x[0]);
}

// 6x6 inverse by Gauss-Jordan with partial pivoting (FP64)
__device__ void invert6(const double *M, double *Minv) {
    double a[6][12];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 12; ++j) a[i][j] = j < 6 ? M[6 * i + j] : (j - 6 == i ? 1.0 : 0.0);
    for (int c = 0; c < 6; ++c) {
        int piv = c;
        for (int r = c + 1; r < 6; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (piv != c)
            for (int j = 0; j < 12; ++j) {
                const double t = a[c][j];
                a[c][j] = a[piv][j];
                a[piv][j] = t;
            }
        const double d = 1.0 / a[c][c];
        for (int j = 0; j < 12; ++j) a[c][j] *= d;
        for (int r = 0; r < 6; ++r)
            if (r != c) {
                const double f = a[r][c];
                for (int j = 0; j < 12; ++j) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) Minv[6 * i + j] = a[i][j + 6];
}

// Z^T Z from the moments N, sum x, sum x_i x_j (m[0], m[1..3], m[4..9]: xx yy zz xy xz yz)
__device__ void ztz_from_moments(const double *m, double *M) {
    for (int i = 0; i < 36; ++i) M[i] = 0.0;
    const double N = m[0], sx = m[1], sy = m[2], sz = m[3];
    const double xx = m[4], yy = m[5], zz = m[6], xy = m[7], xz = m[8], yz = m[9];
    M[0] = M[7] = M[14] = N;
    // translation-rotation block: sum_n (e_k x x_n)_i
    const double Tr[3][3] = {{0, sz, -sy}, {-sz, 0, sx}, {sy, -sx, 0}};  // [i][k]
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            M[6 * i + 3 + k] = Tr[i][k];
            M[6 * (3 + k) + i] = Tr[i][k];
        }
    const double r2 = xx + yy + zz;
    const double R[3][3] = {{r2 - xx, -xy, -xz}, {-xy, r2 - yy, -yz}, {-xz, -yz, r2 - zz}};
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) M[6 * (3 + k) + 3 + l] = R[k][l];
}

// ---------------------------------------------------------------- element pass
template <typename T>
struct ElemArgs {
    int ne;
    const int *conn;          // [4][ne] corner-node indices
    const double *pos;        // [nl][3]
    const T *jinv0;           // [9][ne]
    const T *vol0, *G, *E, *A, *B;
    double *last_det;         // [ne]
    T *force;                 // [ne][12]
    T *geo;                   // [10][ne]: dpsi 1,2,3,5,6,7,9,10,11 then vol
    int calc_noise;
    double kT, dt;
    const mrg::Jump *jump_group;  // [ceil(ne/64)]: A^(7*64*g)
    const mrg::Jump *jump_lane;   // [64]: A^(7*l)
    uint32_t rng[6];              // stream state at the start of this pass
    CgScalars *S;
};

// tetra_element_linear.cpp:191-443 and :468-483 for one element (mass-less solver:
// du starts at zero). Arithmetic in the reference's order, no contraction.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_element(ElemArgs<T> a) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.ne) return;
    const int ne = a.ne;
    uint32_t st[6];
    if (a.calc_noise) {
        // this element's 7 draws start 7*e draws into the pass's segment of the stream
        for (int i = 0; i < 6; ++i) st[i] = a.rng[i];
        mrg::d_apply(a.jump_group[e >> 6], st);
        mrg::d_apply(a.jump_lane[e & 63], st);
    }
    int n[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) n[k] = a.conn[k * ne + e];
    double x0[3], J64[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) x0[c] = a.pos[3 * n[0] + c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J64[r][c] = a.pos[3 * n[r + 1] + c] - x0[c];
    T J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = (T)J64[r][c];

    // shape function derivatives and volume (tetra_element_linear.cpp:222-292)
    T d[12];
    d[1] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
    d[2] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    d[3] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
    d[5] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
    d[6] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    d[7] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
    d[9] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
    d[10] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    d[11] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
    T det = J[0][0] * d[1] + J[1][0] * d[2] + J[2][0] * d[3];
    const double ld = a.last_det[e];
    T vol;
    if (ld * (double)det < 0) {
        __hip_atomic_fetch_add(&a.S->inverted, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vol = a.geo[9 * ne + e];  // the reference keeps the previous volume
    } else {
        a.last_det[e] = (double)det;
        vol = (T)(1.0 / 6.0) * (T)fabs(det);
        const T rdet = (T)1.0 / det;
        d[1] *= rdet; d[2] *= rdet; d[3] *= rdet;
        d[5] *= rdet; d[6] *= rdet; d[7] *= rdet;
        d[9] *= rdet; d[10] *= rdet; d[11] *= rdet;
    }
    d[0] = -(d[1] + d[2] + d[3]);
    d[4] = -(d[5] + d[6] + d[7]);
    d[8] = -(d[9] + d[10] + d[11]);

    // shear elastic stress (tetra_element_linear.cpp:375-392)
    T Ji[3][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) Ji[k / 3][k % 3] = a.jinv0[k * ne + e];
    T F[3][3], s[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T acc = (T)0;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc += J[k][i] * Ji[j][k];
            F[i][j] = acc;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T acc = (T)0;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc += F[i][k] * F[j][k];
            s[i][j] = acc;
        }
    const T G = a.G[e], E = a.E[e], v0 = a.vol0[e];
    const T sc = G * v0 / vol;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s[i][j] *= sc;
    s[0][0] -= G;
    s[1][1] -= G;
    s[2][2] -= G;
    // bulk elastic stress (:397-403)
    {
        const T c2 = E - G * (T)2.0 / (T)3.0;
        const T c = G * ((T)1.0 - (v0 / vol)) + (T)0.5 * c2 * ((vol / v0) - (v0 / vol));
        s[0][0] += c;
        s[1][1] += c;
        s[2][2] += c;
    }
    // fluctuating stress (:412-431), draws bf, s00, s11, s22, s01, s02, s12
    if (a.calc_noise) {
        const T A = a.A[e], B = a.B[e];
        double u[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) u[k] = mrg::d_u01(st);
        T r[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) r[k] = (T)((-.5) + ((.5) - (-.5)) * u[k]);
        const T c = sqrt(((T)24 * (T)a.kT) / (vol * (T)a.dt));
        const T bf = sqrt(B) * r[0];
        s[0][0] += c * (sqrt((T)2 * A) * r[1] + bf);
        s[1][1] += c * (sqrt((T)2 * A) * r[2] + bf);
        s[2][2] += c * (sqrt((T)2 * A) * r[3] + bf);
        s[0][1] += c * sqrt(A) * r[4];
        s[0][2] += c * sqrt(A) * r[5];
        s[1][2] += c * sqrt(A) * r[6];
        s[1][0] = s[0][1];
        s[2][0] = s[0][2];
        s[2][1] = s[1][2];
    }
    // apply_stress_tensor (:436-443) then node_force -= du (:468-483)
    T out[12];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T du = (T)0 + vol * (d[q] * s[i][0] + d[4 + q] * s[i][1] + d[8 + q] * s[i][2]);
            out[3 * q + i] = (T)0 - du;
        }
    T *fo = a.force + (size_t)12 * e;
#pragma unroll
    for (int k = 0; k < 12; ++k) fo[k] = out[k];
    const int gi[9] = {1, 2, 3, 5, 6, 7, 9, 10, 11};
#pragma unroll
    for (int k = 0; k < 9; ++k) a.geo[(size_t)k * ne + e] = d[gi[k]];
    a.geo[(size_t)9 * ne + e] = vol;
}

// ---------------------------------------------------------------- matrix assembly
template <typename T>
struct AsmArgs {
    int nl, ne;
    const int *slice_w;          // [nslices]
    const long long *slice_base; // [nslices] first slot of the slice
    const uint32_t *cptr;        // [nslots+1]
    const uint32_t *contrib;     // (e<<4)|(a<<2)|b, element order; kIdentity = identity block
    const unsigned char *diag_j; // [nl] slot index j of the diagonal block
    const T *geo;                // [10][ne]
    const T *A, *B;
    T *val;                      // [nslots*9]
    double *dinv;                // [nl][3]
};

// Entry (4i+a, 4j+b) of an element's 12x12 viscosity matrix, evaluated as the
// reference's create_viscosity_matrix does (tetra_element_linear.cpp:298-327, macros
// tetra_element_linear.h:56-117, del2 :585-622): upper triangle computed, the lower
// one mirrored, everything times vol at the end.
template <typename T>
__device__ __forceinline__ T visc_entry(const T *d, T A, T B, T vol, int i, int a, int j, int b) {
#pragma clang fp contract(off)
    int I = 4 * i, Jc = 4 * j, p = a, q = b;
    if (I > Jc || (I == Jc && p > q)) {  // mirror to the upper triangle
        int t = I; I = Jc; Jc = t;
        t = p; p = q; q = t;
    }
    T v;
    if (I == Jc) {
        const T del2 = (d[p] * d[q] + d[4 + p] * d[4 + q] + d[8 + p] * d[8 + q]);
        v = d[I + q] * d[I + p] * (A + B);
        v += del2 * A;
    } else {
        if (p == q) v = (B + A) * (d[I + p] * d[Jc + p]);
        else v = B * (d[I + p] * d[Jc + q]) + A * (d[I + q] * d[Jc + p]);
    }
    return v * vol;
}

// SparseMatrixFixedPattern::build + calc_inverse_diagonal (SparseMatrixFixedPattern.cpp:86-94,
// 305-312): every block entry is the sum of its element sources in element order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs<T> a) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= a.nl) return;
    const int sl = n / kSlice, lane = n % kSlice;
    const int W = a.slice_w[sl];
    const long long base = a.slice_base[sl];
    const int ne = a.ne;
    for (int j = 0; j < W; ++j) {
        const long long slot = base + (long long)j * kSlice + lane;
        const uint32_t c0 = a.cptr[slot], c1 = a.cptr[slot + 1];
        double acc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = 0.0;
        for (uint32_t c = c0; c < c1; ++c) {
            const uint32_t code = a.contrib[c];
            if (code == kIdentity) {
                acc[0] += 1.0; acc[4] += 1.0; acc[8] += 1.0;
                continue;
            }
            const int e = (int)(code >> 4), pa = (code >> 2) & 3, pb = code & 3;
            T d[12];
            d[1] = a.geo[(size_t)0 * ne + e]; d[2] = a.geo[(size_t)1 * ne + e]; d[3] = a.geo[(size_t)2 * ne + e];
            d[5] = a.geo[(size_t)3 * ne + e]; d[6] = a.geo[(size_t)4 * ne + e]; d[7] = a.geo[(size_t)5 * ne + e];
            d[9] = a.geo[(size_t)6 * ne + e]; d[10] = a.geo[(size_t)7 * ne + e]; d[11] = a.geo[(size_t)8 * ne + e];
            const T vol = a.geo[(size_t)9 * ne + e];
            d[0] = -(d[1] + d[2] + d[3]);
            d[4] = -(d[5] + d[6] + d[7]);
            d[8] = -(d[9] + d[10] + d[11]);
            const T A = a.A[e], B = a.B[e];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) acc[3 * i + jj] += (double)visc_entry<T>(d, A, B, vol, i, pa, jj, pb);
        }
        T *v = a.val + (size_t)base * 9 + (size_t)j * 9 * kSlice + lane;
#pragma unroll
        for (int k = 0; k < 9; ++k) v[(size_t)k * kSlice] = (T)acc[k];
        if (j == a.diag_j[n]) {
#pragma unroll
            for (int i = 0; i < 3; ++i) a.dinv[vi(n, i)] = 1.0 / (double)(T)acc[4 * i];
        }
    }
}

// ---------------------------------------------------------------- gather + CG start
template <typename T>
struct GatherArgs {
    int nl, rows_per_block;
    const uint32_t *inc_ptr;    // [nl+1]
    const uint32_t *inc;        // e*4 + slot, element order
    const T *force;             // [ne][12]
    const double *ctf;          // [nl][3] or null: constant forces summed in list order
    const uint32_t *fold_ptr;   // [nl+1] or null
    const double *fold;         // [..][3]: linearise_force increments in element order
    const unsigned char *pinned;
    const double *dinv;
    const double *pos;          // [nl][3] (rigid modes)
    double3 centre;
    double *rhs, *r, *z, *x;
    double *P;                  // partials
    CgScalars *S;
};

// Blob::aggregate_forces_and_solve (Blob.cpp:3152-3238) for corner nodes, then the
// start of NoMassCGSolver::solve (NoMassCGSolver.cpp:196-220). With PROJ (free blob,
// mixed form) the CG start is left to k_cg_start and the partials are |f|^2, Z^T f
// and the moments of Z^T Z.
template <typename T, bool PROJ>
__global__ __launch_bounds__(kBlock) void k_gather(GatherArgs<T> a) {
#pragma clang fp contract(off)
    const RowRange rr = row_range(a.nl, a.rows_per_block);
    constexpr int NV = PROJ ? 17 : 2;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int n = rr.begin + threadIdx.x; n < rr.end; n += kBlock) {
        double f[3] = {0.0, 0.0, 0.0};
        if (a.ctf) {
            f[0] = a.ctf[3 * n]; f[1] = a.ctf[3 * n + 1]; f[2] = a.ctf[3 * n + 2];
        }
        for (uint32_t k = a.inc_ptr[n]; k < a.inc_ptr[n + 1]; ++k) {
            const uint32_t es = a.inc[k];
            const T *src = a.force + (size_t)12 * (es >> 2) + 3 * (es & 3);
            f[0] += (double)src[0];
            f[1] += (double)src[1];
            f[2] += (double)src[2];
        }
        if (a.fold_ptr)
            for (uint32_t k = a.fold_ptr[n]; k < a.fold_ptr[n + 1]; ++k) {
                f[0] += a.fold[3 * k];
                f[1] += a.fold[3 * k + 1];
                f[2] += a.fold[3 * k + 2];
            }
        if (a.pinned[n]) f[0] = f[1] = f[2] = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) a.rhs[vi(n, c)] = f[c];
        if constexpr (!PROJ) {
            double zz[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a.r[vi(n, c)] = f[c];
                a.x[vi(n, c)] = 0.0;
                zz[c] = a.dinv[vi(n, c)] * f[c];
                a.z[vi(n, c)] = zz[c];
            }
            v[0] += f[0] * zz[0] + f[1] * zz[1] + f[2] * zz[2];
            v[1] += f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        } else {
            const double x[3] = {a.pos[3 * n] - a.centre.x, a.pos[3 * n + 1] - a.centre.y,
                                 a.pos[3 * n + 2] - a.centre.z};
            double zf[6];
            zt_times(x, f, zf);
            v[0] += f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
#pragma unroll
            for (int k = 0; k < 6; ++k) v[1 + k] += zf[k];
            v[7] += 1.0;
            v[8] += x[0]; v[9] += x[1]; v[10] += x[2];
            v[11] += x[0] * x[0]; v[12] += x[1] * x[1]; v[13] += x[2] * x[2];
            v[14] += x[0] * x[1]; v[15] += x[0] * x[2]; v[16] += x[1] * x[2];
        }
    }
    store_partials<NV>(v, a.P, gridDim.x, rr.L);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.S->inverted_step = a.S->inverted;
        a.S->done = 0;
        a.S->converged = 0;
        a.S->iter = 0;
    }
}

// Rigid-mode projection of the right-hand side for a free blob in the mixed form:
// r = f - Z (Z^T Z)^-1 Z^T f, then z = D^-1 r, x = 0, partials r.z.
__global__ __launch_bounds__(kBlock) void k_cg_start(int nl, int rows_per_block, const double *Pg, const double *rhs,
                                                    const double *pos, double3 centre, const double *dinv,
                                                    double *r, double *z, double *x, double *P, CgScalars *S) {
    const RowRange rr = row_range(nl, rows_per_block);
    double t[17];
    sum_partials<17>(Pg, gridDim.x, t);
    double M[36], Gi[36], cf[6];
    const double mom[10] = {t[7], t[8], t[9], t[10], t[11], t[12], t[13], t[14], t[15], t[16]};
    ztz_from_moments(mom, M);
    invert6(M, Gi);
    for (int i = 0; i < 6; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += Gi[6 * i + j] * t[1 + j];
        cf[i] = acc;
    }
    double v[1] = {0.0};
    for (int n = rr.begin + threadIdx.x; n < rr.end; n += kBlock) {
        const double xx[3] = {pos[3 * n] - centre.x, pos[3 * n + 1] - centre.y, pos[3 * n + 2] - centre.z};
        double zc[3], rl[3], zl[3];
        z_times(xx, cf, zc);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rl[c] = rhs[vi(n, c)] - zc[c];
            r[vi(n, c)] = rl[c];
            x[vi(n, c)] = 0.0;
            zl[c] = dinv[vi(n, c)] * rl[c];
            z[vi(n, c)] = zl[c];
        }
        v[0] += rl[0] * zl[0] + rl[1] * zl[1] + rl[2] * zl[2];
    }
    store_partials<1>(v, P, gridDim.x, rr.L);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->f2 = t[0];
        for (int i = 0; i < 36; ++i) S->G[i] = Gi[i];
    }
}

// ---------------------------------------------------------------- CG iteration
template <typename T>
struct SpmvArgs {
    int nl, rows_per_block, iter, max_iter;
    double epsilon2;
    const int *slice_w;
    const long long *slice_base;
    const int *col;        // [nslots]
    const T *val;          // [nslots*9]
    const double *z;
    const double *pos;
    double3 centre;
    double *p, *q;         // updated in place
    const double *Pin;     // partials of the previous kernel (start: r.z [and |f|^2]; later: r.r, r.z)
    double *P1;            // partials of this kernel
    CgScalars *S;
};

// Iteration `iter`: first the convergence test of the previous iteration
// (|r|^2/|f|^2 < epsilon^2, NoMassCGSolver.cpp:223-237), then
// p = beta p + z (mat_vec_fns.cpp:202-212) and q = V p, formed as q = V z + beta q
// so the row gathers only z (SparseMatrixFixedPattern.cpp:148-187), and the p.q
// partials (NoMassCGSolver.cpp:251-265).
template <typename T, bool PROJ>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(SpmvArgs<T> a) {
    CgScalars *S = a.S;
    if (S->done) return;
    const int i = a.iter;
    const bool first = (i == 0);
    double beta = 0.0;
    if (first) {
        if (S->inverted_step > 0) {  // an inverted mesh is not solved
            if (blockIdx.x == 0 && threadIdx.x == 0) S->done = 1;
            return;
        }
        if constexpr (PROJ) {
            double t[1];
            sum_partials<1>(a.Pin, gridDim.x, t);
            if (blockIdx.x == 0 && threadIdx.x == 0) S->delta[1] = t[0];
        } else {
            double t[2];
            sum_partials<2>(a.Pin, gridDim.x, t);
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                S->delta[1] = t[0];
                S->f2 = t[1];
            }
        }
    } else {
        double t[2];
        sum_partials<2>(a.Pin, gridDim.x, t);  // r.r, r.z of iteration i-1
        const double f2 = S->f2;
        const double res = (f2 == 0.0) ? 0.0 : t[0] / f2;
        if (res < a.epsilon2 || i >= a.max_iter) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                S->converged = res < a.epsilon2 ? 1 : 0;
                S->iter = i;
                S->done = 1;
            }
            return;
        }
        beta = t[1] / S->delta[(i - 2) & 1];
        if (blockIdx.x == 0 && threadIdx.x == 0) S->delta[(i - 1) & 1] = t[1];
    }
    const RowRange rr = row_range(a.nl, a.rows_per_block);
    constexpr int NV = PROJ ? 13 : 1;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    for (int n = rr.begin + threadIdx.x; n < rr.end; n += kBlock) {
        const int sl = n / kSlice, lane = n % kSlice;
        const int W = a.slice_w[sl];
        const long long base = a.slice_base[sl];
        const int *colp = a.col + base + lane;
        const T *vp = a.val + (size_t)base * 9 + lane;
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        auto block = [&](int j, int c) {
            const T *vv = vp + (size_t)j * 9 * kSlice;
            const double z0 = a.z[vi(c, 0)], z1 = a.z[vi(c, 1)], z2 = a.z[vi(c, 2)];
            acc0 += (double)vv[0 * kSlice] * z0; acc0 += (double)vv[1 * kSlice] * z1; acc0 += (double)vv[2 * kSlice] * z2;
            acc1 += (double)vv[3 * kSlice] * z0; acc1 += (double)vv[4 * kSlice] * z1; acc1 += (double)vv[5 * kSlice] * z2;
            acc2 += (double)vv[6 * kSlice] * z0; acc2 += (double)vv[7 * kSlice] * z1; acc2 += (double)vv[8 * kSlice] * z2;
        };
        if (W <= kMaxUnrolledRow) {
            // all block columns of the row first, so the gathers and value loads of every
            // block can be in flight together
            int cc[kMaxUnrolledRow];
#pragma unroll
            for (int j = 0; j < kMaxUnrolledRow; ++j)
                if (j < W) cc[j] = colp[(size_t)j * kSlice];
#pragma unroll
            for (int j = 0; j < kMaxUnrolledRow; ++j)
                if (j < W) block(j, cc[j]);
        } else {
#pragma unroll 2
            for (int j = 0; j < W; ++j) block(j, colp[(size_t)j * kSlice]);
        }
        double pn[3], qn[3] = {acc0, acc1, acc2};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double zk = a.z[vi(n, k)];
            if (first) {
                pn[k] = zk;
            } else {
                pn[k] = beta * a.p[vi(n, k)] + zk;
                qn[k] = beta * a.q[vi(n, k)] + qn[k];
            }
            a.p[vi(n, k)] = pn[k];
            a.q[vi(n, k)] = qn[k];
        }
        v[0] += pn[0] * qn[0] + pn[1] * qn[1] + pn[2] * qn[2];
        if constexpr (PROJ) {
            const double x[3] = {a.pos[3 * n] - a.centre.x, a.pos[3 * n + 1] - a.centre.y,
                                 a.pos[3 * n + 2] - a.centre.z};
            double zq[6], zp[6];
            zt_times(x, qn, zq);
            zt_times(x, pn, zp);
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                v[1 + k] += zq[k];
                v[7 + k] += zp[k];
            }
        }
    }
    store_partials<NV>(v, a.P1, gridDim.x, rr.L);
}

struct VecArgs {
    int nl, rows_per_block, iter;
    const double *p, *q, *dinv;
    const double *pos;
    double3 centre;
    double *x, *r, *z;
    const double *P1;  // partials of the SpMV step
    double *P2;        // r.r, r.z partials of this step
    CgScalars *S;
};

// alpha = delta / p.q, x += alpha p, r -= alpha q (mat_vec_fns.cpp:187-197),
// z = D^-1 r (NoMassCGSolver.cpp:268-281), partials of r.r and r.z.
// With PROJ, q is first projected onto the range of the operator:
// q <- q - Z c, c = (Z^T Z)^-1 Z^T q, and p.q is corrected to match.
template <bool PROJ>
__global__ __launch_bounds__(kBlock) void k_cg_vec(VecArgs a) {
    CgScalars *S = a.S;
    if (S->done) return;
    constexpr int NV = PROJ ? 13 : 1;
    double t[NV];
    sum_partials<NV>(a.P1, gridDim.x, t);
    double pq = t[0];
    double cq[6] = {0, 0, 0, 0, 0, 0};
    if constexpr (PROJ) {
        for (int k = 0; k < 6; ++k) {
            double acc = 0.0;
            for (int j = 0; j < 6; ++j) acc += S->G[6 * k + j] * t[1 + j];
            cq[k] = acc;
        }
        for (int k = 0; k < 6; ++k) pq -= t[7 + k] * cq[k];
    }
    const double alpha = S->delta[(a.iter - 1) & 1] / pq;
    const RowRange rr = row_range(a.nl, a.rows_per_block);
    double v[2] = {0.0, 0.0};
    for (int n = rr.begin + threadIdx.x; n < rr.end; n += kBlock) {
        double qn[3] = {a.q[vi(n, 0)], a.q[vi(n, 1)], a.q[vi(n, 2)]};
        if constexpr (PROJ) {
            const double x[3] = {a.pos[3 * n] - a.centre.x, a.pos[3 * n + 1] - a.centre.y,
                                 a.pos[3 * n + 2] - a.centre.z};
            double zc[3];
            z_times(x, cq, zc);
            qn[0] -= zc[0];
            qn[1] -= zc[1];
            qn[2] -= zc[2];
        }
        double r3[3], z3[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.x[vi(n, k)] += alpha * a.p[vi(n, k)];
            r3[k] = a.r[vi(n, k)] + (-alpha) * qn[k];
            a.r[vi(n, k)] = r3[k];
            z3[k] = a.dinv[vi(n, k)] * r3[k];
            a.z[vi(n, k)] = z3[k];
        }
        v[0] += r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2];
        v[1] += r3[0] * z3[0] + r3[1] * z3[1] + r3[2] * z3[2];
    }
    store_partials<2>(v, a.P2, gridDim.x, rr.L);
}

// Blob::euler_integrate (Blob.cpp:3249-3263), mass-less branch: vel = x, pos += x dt
__global__ __launch_bounds__(kBlock) void k_update(int nl, const double *x, double *vel, double *pos, double dt,
                                                   CgScalars *S) {
    const int converged = S->converged;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) S->inverted = 0;
    if (!converged) return;
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nl) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double v = x[vi(n, k)];
        vel[3 * n + k] = v;
        pos[3 * n + k] += v * dt;
    }
}

// ---------------------------------------------------------------- copy probe
// y = x over n doubles (16-byte vectors), grid-stride: reads n*8 and writes n*8 bytes
__global__ __launch_bounds__(kBlock) void k_copy(const double2 *__restrict__ x, double2 *__restrict__ y, size_t n2) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n2; i += stride) y[i] = x[i];
}

// ================================================================ host side
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count, size_t *total) {
        n = count;
        FFEA_HIP_TRY(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)));
        *total += std::max<size_t>(1, count) * sizeof(T);
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        FFEA_HIP_TRY(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
};

struct Phase {
    std::vector<hipEvent_t> ev;  // pairs
    std::vector<int> iter;       // CG phases: iteration of each timed pair
    double ms = 0;
    long long launches = 0;
};

struct BlobBase {
    virtual ~BlobBase() = default;
};

}  // namespace

// ---------------------------------------------------------------- the handle
struct ffea_gpu_blob {
    ffea_gpu_params prm{};
    int nt = 0, ne = 0, nl = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool mixed = false;
    size_t dev_bytes = 0;

    // host plan
    std::vector<int> lin_of, orig_of_lin;
    std::vector<int> mid_a, mid_b;  // for every node: edge ends (orig) for midside nodes, -1 for corners
    int nslices = 0;
    long long nslots = 0, nblocks = 0;
    bool all_pinned = false;

    // device arrays (element data in T via void pointers sized by precision)
    DevBuf<int> conn, slice_w, col;
    DevBuf<long long> slice_base;
    DevBuf<uint32_t> cptr, contrib, inc_ptr, inc, fold_ptr;
    DevBuf<unsigned char> diag_j, pinned;
    DevBuf<double> pos, vel, last_det, dinv, rhs, r, z, x, q, p, ctf, fold;
    DevBuf<double> Pg, Pi, P1, P2;  // per-block partials: gather, CG start, SpMV step, vector step
    DevBuf<double> d_jinv0, d_vol0, d_G, d_E, d_A, d_B, d_force, d_geo, d_val;  // FP64 form
    DevBuf<float> f_jinv0, f_vol0, f_G, f_E, f_A, f_B, f_force, f_geo, f_val;   // mixed form
    DevBuf<mrg::Jump> jump_group, jump_lane;
    DevBuf<CgScalars> S;
    mrg::Jump step_jump{};
    uint32_t rng[6] = {0, 0, 0, 0, 0, 0};  // noise stream state (host copy)
    bool have_ctf = false, have_fold = false;
    bool project = false;     // rigid-mode projection in the CG (mixed form, free blob)
    double3 centre{0, 0, 0};  // rest centroid of the corner nodes
    int grid_nodes = 0, grid_elems = 0;
    int nb = 0, rows_per_block = 0;  // row kernels: fixed grid and rows per block

    // host-visible status
    CgScalars *h_status = nullptr;  // pinned [3]
    hipEvent_t chunk_ev[2] = {nullptr, nullptr};
    int last_iters = 0;
    long long total_iters = 0;
    int expect_iters = 0;

    // timing
    unsigned timing = 0;  // bit i: time phase i
    int cg_sample = 8;    // CG phases: time one iteration in cg_sample
    Phase phase[FFEA_GPU_NUM_PHASES];
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;

    ~ffea_gpu_blob() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : chunk_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        if (h_status) (void)hipHostFree(h_status);
        if (stream) (void)hipStreamDestroy(stream);
    }

    hipEvent_t take_event() {
        if (ev_used == ev_pool.size()) {
            hipEvent_t e;
            FFEA_HIP_TRY(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        return ev_pool[ev_used++];
    }
    void t_begin(int ph, hipEvent_t *out) {
        *out = nullptr;
        if (!(timing >> ph & 1u) && !(ph >= 3 && ph <= 4)) return;
        *out = take_event();
        FFEA_HIP_TRY(hipEventRecord(*out, stream));
    }
    void t_end(int ph, hipEvent_t b, int it = -1) {
        if (!b) return;
        hipEvent_t e = take_event();
        FFEA_HIP_TRY(hipEventRecord(e, stream));
        phase[ph].ev.push_back(b);
        phase[ph].ev.push_back(e);
        phase[ph].iter.push_back(it);
    }
    // CG phases are timed on every cg_sample-th iteration only (events between
    // kernels delay the next launch); pairs of iterations past the converged one
    // (launches that return at once) are dropped when collected
    bool t_cg(int ph, int it) const { return (timing >> ph & 1u) && it % cg_sample == 0; }
    void collect_timing(int last_iter = -1) {
        if (!timing) return;
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
        for (auto &p : phase) {
            for (size_t i = 0; i + 1 < p.ev.size(); i += 2) {
                const int it = p.iter[i / 2];
                if (it >= 0 && last_iter >= 0 && it >= last_iter) continue;
                float ms = 0;
                FFEA_HIP_TRY(hipEventElapsedTime(&ms, p.ev[i], p.ev[i + 1]));
                p.ms += ms;
                p.launches++;
            }
            p.ev.clear();
            p.iter.clear();
        }
        ev_used = 0;
    }

    void build(const double *pos_h, const int *top, const double *A, const double *B, const double *G,
               const double *E, int npin, const int *pin, int nctf, const int *ctf_node, const double *ctf_force,
               const uint32_t *rng_state);
    template <typename T>
    void upload_elements(const std::vector<double> &jinv0, const std::vector<double> &vol0, const double *A,
                         const double *B, const double *G, const double *E);
    void launch_element();
    template <typename T, bool PROJ>
    void launch_solve(const T *val, const T *force);
    int check_status_after_forces();
    int solve_and_update();
};

namespace {

// closed-form rest state (Blob::calc_rest_state_info, Blob.cpp:3080-3143;
// mat3_invert, mat_vec_fns.cpp:125-154): J^-1_0 and vol_0 = |det J|/6 per element
void rest_state(const double *p0, const double *p1, const double *p2, const double *p3, double *jinv, double *vol0,
                double *det_out) {
    double m[3][3];
    for (int c = 0; c < 3; ++c) {
        m[0][c] = p1[c] - p0[c];
        m[1][c] = p2[c] - p0[c];
        m[2][c] = p3[c] - p0[c];
    }
    double mi[3][3];
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
    *det_out = det;
    const double rdet = 1.0 / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) jinv[3 * i + j] = mi[i][j] * rdet;
    *vol0 = (1.0 / 6.0) * fabs(det);
}

}  // namespace

template <typename T>
void ffea_gpu_blob::upload_elements(const std::vector<double> &jinv0, const std::vector<double> &vol0,
                                    const double *A, const double *B, const double *G, const double *E) {
    DevBuf<T> *bj, *bv, *bG, *bE, *bA, *bB, *bf, *bg;
    if constexpr (std::is_same<T, double>::value) {
        bj = &d_jinv0; bv = &d_vol0; bG = &d_G; bE = &d_E; bA = &d_A; bB = &d_B; bf = &d_force; bg = &d_geo;
    } else {
        bj = &f_jinv0; bv = &f_vol0; bG = &f_G; bE = &f_E; bA = &f_A; bB = &f_B; bf = &f_force; bg = &f_geo;
    }
    std::vector<T> tmp((size_t)9 * ne);
    for (int k = 0; k < 9; ++k)
        for (int e = 0; e < ne; ++e) tmp[(size_t)k * ne + e] = (T)jinv0[(size_t)9 * e + k];
    bj->alloc(tmp.size(), &dev_bytes);
    bj->upload(tmp, stream);
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    auto up = [&](DevBuf<T> *b, const double *src) {
        std::vector<T> t(ne);
        for (int e = 0; e < ne; ++e) t[e] = (T)src[e];
        b->alloc(ne, &dev_bytes);
        b->upload(t, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
    };
    up(bv, vol0.data());
    up(bA, A);
    up(bB, B);
    up(bG, G);
    up(bE, E);
    bf->alloc((size_t)12 * ne, &dev_bytes);
    bg->alloc((size_t)10 * ne, &dev_bytes);
    // the geometry array is read for the previous volume of an inverted element
    std::vector<T> g0((size_t)10 * ne, (T)0);
    for (int e = 0; e < ne; ++e) g0[(size_t)9 * ne + e] = (T)vol0[e];
    bg->upload(g0, stream);
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
}

void ffea_gpu_blob::build(const double *pos_h, const int *top, const double *A, const double *B, const double *G,
                          const double *E, int npin, const int *pin, int nctf, const int *ctf_node,
                          const double *ctf_force, const uint32_t *rng_state) {
    // ---- corner / midside classification (Blob::load_topology, Blob.cpp:2071-2077)
    std::vector<signed char> kind(nt, 0);  // 1 corner, 2 midside
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 10; ++s) {
            const int v = top[10 * e + s];
            if (v < 0 || v >= nt) throw Invalid(fmt("Error: Element %d references an out of bounds node index.", e));
            const signed char k = s < 4 ? 1 : 2;
            if (kind[v] && kind[v] != k)
                throw Invalid(fmt("node %d is a corner of one element and a midside node of another", v));
            kind[v] = k;
        }
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 4; ++s)
            for (int t = s + 1; t < 4; ++t)
                if (top[10 * e + s] == top[10 * e + t]) throw Invalid(fmt("element %d repeats a corner node", e));
    lin_of.assign(nt, -1);
    orig_of_lin.clear();
    for (int v = 0; v < nt; ++v)
        if (kind[v] == 1) {
            lin_of[v] = (int)orig_of_lin.size();
            orig_of_lin.push_back(v);
        }
    nl = (int)orig_of_lin.size();
    mid_a.assign(nt, -1);
    mid_b.assign(nt, -1);
    static const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 6; ++k) {
            const int m = top[10 * e + 4 + k];
            mid_a[m] = top[10 * e + ends[k][0]];
            mid_b[m] = top[10 * e + ends[k][1]];
        }

    // ---- pins
    std::vector<unsigned char> pin_node(nt, 0), pin_lin(nl, 0);
    for (int i = 0; i < npin; ++i) {
        if (pin[i] < 0 || pin[i] >= nt) throw Invalid(fmt("Error: Pinned node %d references an out of bounds node index.", i));
        pin_node[pin[i]] = 1;
    }
    for (int n = 0; n < nl; ++n) pin_lin[n] = pin_node[orig_of_lin[n]];
    all_pinned = (npin == nt);  // Blob.cpp:405 compares the list length with the node count

    // ---- linearised rest positions and rest state (Blob.cpp:203-208)
    std::vector<double> p0((size_t)3 * nl);
    for (int n = 0; n < nl; ++n)
        for (int c = 0; c < 3; ++c) p0[3 * n + c] = pos_h[3 * orig_of_lin[n] + c];
    std::vector<int> conn_h((size_t)4 * ne);
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 4; ++s) conn_h[(size_t)s * ne + e] = lin_of[top[10 * e + s]];
    std::vector<double> jinv0((size_t)9 * ne), vol0(ne), det0(ne);
    for (int e = 0; e < ne; ++e)
        rest_state(&p0[3 * conn_h[e]], &p0[3 * conn_h[(size_t)ne + e]], &p0[3 * conn_h[(size_t)2 * ne + e]],
                   &p0[3 * conn_h[(size_t)3 * ne + e]], &jinv0[(size_t)9 * e], &vol0[e], &det0[e]);

    // ---- node -> element incidence in element order (Blob.cpp:3284-3314)
    std::vector<uint32_t> ip(nl + 1, 0), ic((size_t)4 * ne);
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 4; ++s) ip[conn_h[(size_t)s * ne + e] + 1]++;
    for (int n = 0; n < nl; ++n) ip[n + 1] += ip[n];
    {
        std::vector<uint32_t> fill(ip.begin(), ip.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int s = 0; s < 4; ++s) ic[fill[conn_h[(size_t)s * ne + e]]++] = (uint32_t)e * 4u + (uint32_t)s;
    }

    // ---- block pattern (NoMassCGSolver.cpp:81-127 restricted to corner nodes):
    // row n holds a block for each unpinned corner neighbour m (ascending), whose
    // sources are the elements containing both, in element order; pinned rows and
    // rows without element sources hold an identity diagonal block only.
    struct Nb { int m; uint32_t code; };
    std::vector<int> rowlen(nl);
    nslices = (nl + kSlice - 1) / kSlice;
    std::vector<int> sw(nslices, 0);
    // first pass: row lengths
    std::vector<Nb> tmp;
    auto row_entries = [&](int n, std::vector<Nb> &out) {
        out.clear();
        if (pin_lin[n]) return;
        for (uint32_t k = ip[n]; k < ip[n + 1]; ++k) {
            const int e = (int)(ic[k] >> 2), a = (int)(ic[k] & 3);
            for (int b = 0; b < 4; ++b) {
                const int m = conn_h[(size_t)b * ne + e];
                if (pin_lin[m]) continue;
                out.push_back({m, ((uint32_t)e << 4) | ((uint32_t)a << 2) | (uint32_t)b});
            }
        }
        std::stable_sort(out.begin(), out.end(), [](const Nb &x, const Nb &y) { return x.m < y.m; });
    };
    long long total_contrib = 0;
    for (int n = 0; n < nl; ++n) {
        row_entries(n, tmp);
        int distinct = 0;
        for (size_t t = 0; t < tmp.size(); ++t)
            if (t == 0 || tmp[t].m != tmp[t - 1].m) ++distinct;
        if (distinct == 0) distinct = 1;  // identity diagonal
        rowlen[n] = distinct;
        total_contrib += tmp.empty() ? 1 : (long long)tmp.size();
        sw[n / kSlice] = std::max(sw[n / kSlice], distinct);
    }
    if ((long long)ne * 16 >= (1ll << 32) - 1 || (long long)ne >= (1ll << 28))
        throw Invalid("mesh too large for 32-bit contributor lists");
    std::vector<long long> sb(nslices + 1, 0);
    for (int s = 0; s < nslices; ++s) sb[s + 1] = sb[s] + (long long)sw[s] * kSlice;
    nslots = sb[nslices];
    if (nslots >= (1ll << 31)) throw Invalid("mesh too large for 32-bit slot indices");
    std::vector<int> col_h(nslots);
    std::vector<uint32_t> cnt(nslots + 1, 0);
    std::vector<unsigned char> dj(nl, 0);
    // second pass: counts per slot, columns
    std::vector<uint32_t> contrib_h;
    contrib_h.reserve((size_t)total_contrib);
    nblocks = 0;
    // count pass
    for (int n = 0; n < nl; ++n) {
        const int s = n / kSlice, lane = n % kSlice;
        row_entries(n, tmp);
        const long long base = sb[s];
        int j = 0;
        if (tmp.empty()) {
            const long long slot = base + lane;
            col_h[slot] = n;
            cnt[slot + 1] = 1;
            dj[n] = 0;
            j = 1;
            nblocks += 1;
        } else {
            size_t t = 0;
            while (t < tmp.size()) {
                size_t t1 = t;
                while (t1 < tmp.size() && tmp[t1].m == tmp[t].m) ++t1;
                const long long slot = base + (long long)j * kSlice + lane;
                col_h[slot] = tmp[t].m;
                cnt[slot + 1] = (uint32_t)(t1 - t);
                if (tmp[t].m == n) dj[n] = (unsigned char)j;
                ++j;
                t = t1;
            }
            nblocks += j;
        }
        for (; j < sw[s]; ++j) col_h[base + (long long)j * kSlice + lane] = n;  // padding: zero block
    }
    // rows past nl in the last slice: pad columns with row 0
    for (int n = nl; n < nslices * kSlice; ++n) {
        const int s = n / kSlice, lane = n % kSlice;
        for (int j = 0; j < sw[s]; ++j) col_h[sb[s] + (long long)j * kSlice + lane] = 0;
    }
    for (long long k = 0; k < nslots; ++k) cnt[k + 1] += cnt[k];
    contrib_h.assign(cnt[nslots], 0);
    for (int n = 0; n < nl; ++n) {
        const int s = n / kSlice, lane = n % kSlice;
        row_entries(n, tmp);
        const long long base = sb[s];
        if (tmp.empty()) {
            contrib_h[cnt[base + lane]] = kIdentity;
            continue;
        }
        size_t t = 0;
        int j = 0;
        while (t < tmp.size()) {
            const long long slot = base + (long long)j * kSlice + lane;
            uint32_t w = cnt[slot];
            while (t < tmp.size() && tmp[t].m == col_h[slot]) contrib_h[w++] = tmp[t++].code;
            ++j;
        }
    }

    // ---- constant forces and the linearise_force increments of midside ones
    std::vector<double> ctf_node_total((size_t)3 * nt, 0.0);
    std::vector<unsigned char> has_ctf(nt, 0);
    for (int i = 0; i < nctf; ++i) {
        const int v = ctf_node[i];
        if (v < 0 || v >= nt) throw Invalid(fmt("constant force %d references an out of bounds node index", i));
        for (int c = 0; c < 3; ++c) ctf_node_total[3 * v + c] += ctf_force[3 * i + c];
        has_ctf[v] = 1;
    }
    std::vector<double> ctf_h((size_t)3 * nl, 0.0);
    for (int n = 0; n < nl; ++n)
        for (int c = 0; c < 3; ++c) ctf_h[3 * n + c] = ctf_node_total[3 * orig_of_lin[n] + c];
    have_ctf = nctf > 0;
    // Blob::linearise_force (Blob.cpp:781-799): a midside force is folded half onto each
    // edge end by the first element (in element order) that holds it, then zeroed.
    std::vector<std::vector<std::array<double, 3>>> folds(nl);
    bool any_mid_ctf = false;
    for (int v = 0; v < nt; ++v)
        if (has_ctf[v] && kind[v] == 2) any_mid_ctf = true;
    if (any_mid_ctf) {
        std::vector<double> fm(ctf_node_total);  // only midside entries are read
        static const int adj[4][3] = {{4, 5, 6}, {4, 7, 8}, {5, 7, 9}, {6, 8, 9}};
        for (int e = 0; e < ne; ++e) {
            const int *nn = top + 10 * e;
            for (int a = 0; a < 4; ++a) {
                std::array<double, 3> inc3;
                bool nz = false;
                for (int c = 0; c < 3; ++c) {
                    inc3[c] = 0.5 * (fm[3 * nn[adj[a][0]] + c] + fm[3 * nn[adj[a][1]] + c] + fm[3 * nn[adj[a][2]] + c]);
                    if (inc3[c] != 0.0) nz = true;
                }
                if (nz) folds[lin_of[nn[a]]].push_back(inc3);
            }
            for (int k = 4; k < 10; ++k)
                for (int c = 0; c < 3; ++c) fm[3 * nn[k] + c] = 0.0;
        }
    }
    std::vector<uint32_t> fp(nl + 1, 0);
    std::vector<double> fv;
    for (int n = 0; n < nl; ++n) {
        for (auto &f : folds[n]) fv.insert(fv.end(), f.begin(), f.end());
        fp[n + 1] = fp[n] + (uint32_t)folds[n].size();
    }
    have_fold = any_mid_ctf;

    // ---- noise stream jump tables
    const int ngroups = (ne + 63) / 64;
    std::vector<mrg::Jump> jg(ngroups), jl(64);
    const mrg::Jump a7 = mrg::h_pow(7), a448 = mrg::h_pow(7ull * 64ull);
    jl[0] = mrg::h_identity();
    for (int l = 1; l < 64; ++l) jl[l] = mrg::h_mul(a7, jl[l - 1]);
    jg[0] = mrg::h_identity();
    for (int g = 1; g < ngroups; ++g) jg[g] = mrg::h_mul(a448, jg[g - 1]);
    step_jump = mrg::h_pow(7ull * (unsigned long long)ne);

    // ---- device upload
    grid_nodes = (nl + kBlock - 1) / kBlock;
    grid_elems = (ne + kBlock - 1) / kBlock;
    conn.alloc(conn_h.size(), &dev_bytes); conn.upload(conn_h, stream);
    pos.alloc((size_t)3 * nl, &dev_bytes); pos.upload(p0, stream);
    vel.alloc((size_t)3 * nl, &dev_bytes);
    FFEA_HIP_TRY(hipMemsetAsync(vel.p, 0, sizeof(double) * 3 * nl, stream));
    last_det.alloc(ne, &dev_bytes); last_det.upload(det0, stream);
    inc_ptr.alloc(ip.size(), &dev_bytes); inc_ptr.upload(ip, stream);
    inc.alloc(ic.size(), &dev_bytes); inc.upload(ic, stream);
    std::vector<unsigned char> pin_u(pin_lin);
    pinned.alloc(nl, &dev_bytes); pinned.upload(pin_u, stream);
    slice_w.alloc(nslices, &dev_bytes); slice_w.upload(sw, stream);
    std::vector<long long> sbase(sb.begin(), sb.end() - 1);
    slice_base.alloc(nslices, &dev_bytes); slice_base.upload(sbase, stream);
    col.alloc(nslots, &dev_bytes); col.upload(col_h, stream);
    cptr.alloc(cnt.size(), &dev_bytes); cptr.upload(cnt, stream);
    contrib.alloc(contrib_h.size(), &dev_bytes); contrib.upload(contrib_h, stream);
    diag_j.alloc(nl, &dev_bytes); diag_j.upload(dj, stream);
    const size_t vlen = (size_t)nslices * 3 * kSlice;  // tiled node vectors (vi)
    for (DevBuf<double> *b : {&dinv, &rhs, &r, &z, &x, &q, &p}) {
        b->alloc(vlen, &dev_bytes);
        FFEA_HIP_TRY(hipMemsetAsync(b->p, 0, sizeof(double) * vlen, stream));
    }
    {
        // row kernels: a fixed grid of at most max_row_blocks blocks (tuning knob)
        int max_row_blocks = 1024;
        if (const char *e = getenv("FFEA_GPU_MAX_ROW_BLOCKS")) max_row_blocks = std::max(8, atoi(e));
        nb = std::min((nl + kBlock - 1) / kBlock, max_row_blocks);
    }
    if (nb >= 8) nb -= nb % 8;
    rows_per_block = ((nl + nb - 1) / nb + kSlice - 1) / kSlice * kSlice;
    nb = (nl + rows_per_block - 1) / rows_per_block;
    if (nb >= 8 && nb % 8) {  // keep the XCD mapping: grow to a multiple of 8 blocks
        nb += 8 - nb % 8;
    }
    Pg.alloc((size_t)17 * nb, &dev_bytes);
    Pi.alloc((size_t)2 * nb, &dev_bytes);
    P1.alloc((size_t)13 * nb, &dev_bytes);
    P2.alloc((size_t)2 * nb, &dev_bytes);
    if (have_ctf) { ctf.alloc(ctf_h.size(), &dev_bytes); ctf.upload(ctf_h, stream); }
    if (have_fold) {
        fold_ptr.alloc(fp.size(), &dev_bytes); fold_ptr.upload(fp, stream);
        fold.alloc(fv.size(), &dev_bytes); fold.upload(fv, stream);
    }
    jump_group.alloc(ngroups, &dev_bytes); jump_group.upload(jg, stream);
    jump_lane.alloc(64, &dev_bytes); jump_lane.upload(jl, stream);
    S.alloc(1, &dev_bytes);
    CgScalars s0{};
    S.upload(&s0, 1, stream);
    for (int i = 0; i < 6; ++i) rng[i] = rng_state[i];
    // free blob (no pins, no stokes drag): the viscosity matrix is singular with the 6
    // rigid modes as null space. The mixed form projects them out of every SpMV result.
    {
        double c[3] = {0, 0, 0};
        for (int n = 0; n < nl; ++n)
            for (int k = 0; k < 3; ++k) c[k] += p0[3 * n + k];
        centre = make_double3(c[0] / nl, c[1] / nl, c[2] / nl);
        bool connected = true;
        {  // union-find over element corners
            std::vector<int> par(nl);
            std::iota(par.begin(), par.end(), 0);
            std::function<int(int)> find = [&](int v) {
                while (par[v] != v) v = par[v] = par[par[v]];
                return v;
            };
            for (int e = 0; e < ne; ++e)
                for (int k = 1; k < 4; ++k) {
                    int ra = find(conn_h[e]), rb = find(conn_h[(size_t)k * ne + e]);
                    if (ra != rb) par[ra] = rb;
                }
            const int r0 = find(0);
            for (int n = 1; n < nl && connected; ++n) connected = find(n) == r0;
        }
        project = mixed && npin == 0 && connected;
    }
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    if (mixed) {
        upload_elements<float>(jinv0, vol0, A, B, G, E);
        f_val.alloc((size_t)nslots * 9, &dev_bytes);
        FFEA_HIP_TRY(hipMemsetAsync(f_val.p, 0, sizeof(float) * 9 * nslots, stream));
    } else {
        upload_elements<double>(jinv0, vol0, A, B, G, E);
        d_val.alloc((size_t)nslots * 9, &dev_bytes);
        FFEA_HIP_TRY(hipMemsetAsync(d_val.p, 0, sizeof(double) * 9 * nslots, stream));
    }
    FFEA_HIP_TRY(hipHostMalloc(&h_status, sizeof(CgScalars) * 3, hipHostMallocDefault));
    for (auto &e : chunk_ev) FFEA_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
}

void ffea_gpu_blob::launch_element() {
    hipEvent_t t0;
    t_begin(0, &t0);
    if (mixed) {
        ElemArgs<float> a{ne, conn.p, pos.p, f_jinv0.p, f_vol0.p, f_G.p, f_E.p, f_A.p, f_B.p, last_det.p,
                          f_force.p, f_geo.p, prm.calc_noise, prm.kT, prm.dt, jump_group.p, jump_lane.p,
                          {rng[0], rng[1], rng[2], rng[3], rng[4], rng[5]}, S.p};
        hipLaunchKernelGGL(k_element<float>, dim3(grid_elems), dim3(kBlock), 0, stream, a);
    } else {
        ElemArgs<double> a{ne, conn.p, pos.p, d_jinv0.p, d_vol0.p, d_G.p, d_E.p, d_A.p, d_B.p, last_det.p,
                           d_force.p, d_geo.p, prm.calc_noise, prm.kT, prm.dt, jump_group.p, jump_lane.p,
                           {rng[0], rng[1], rng[2], rng[3], rng[4], rng[5]}, S.p};
        hipLaunchKernelGGL(k_element<double>, dim3(grid_elems), dim3(kBlock), 0, stream, a);
    }
    FFEA_HIP_TRY(hipGetLastError());
    t_end(0, t0);
    // the pass consumed 7 draws per element (tetra_element_linear.cpp:416-426)
    if (prm.calc_noise) mrg::h_apply(step_jump, rng, rng);
}

int ffea_gpu_blob::check_status_after_forces() {
    FFEA_HIP_TRY(hipMemcpyAsync(h_status, S.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    const int inv = h_status->inverted;
    if (inv > 0) {
        err = inv == 1 ? "1 element has inverted since the last step. Aborting simulation.\n"
                       : fmt("%d elements have inverted since the last step. Aborting simulation.\n", inv);
        return FFEA_GPU_ERR_INVERSION;
    }
    return FFEA_GPU_OK;
}

template <typename T, bool PROJ>
void ffea_gpu_blob::launch_solve(const T *val, const T *force) {
    hipEvent_t t0;
    t_begin(2, &t0);
    GatherArgs<T> g{nl, rows_per_block, inc_ptr.p, inc.p, force, have_ctf ? ctf.p : nullptr,
                    have_fold ? fold_ptr.p : nullptr, have_fold ? fold.p : nullptr, pinned.p, dinv.p, pos.p,
                    centre, rhs.p, r.p, z.p, x.p, PROJ ? Pg.p : Pi.p, S.p};
    hipLaunchKernelGGL((k_gather<T, PROJ>), dim3(nb), dim3(kBlock), 0, stream, g);
    if (PROJ)
        hipLaunchKernelGGL(k_cg_start, dim3(nb), dim3(kBlock), 0, stream, nl, rows_per_block, (const double *)Pg.p,
                           (const double *)rhs.p, (const double *)pos.p, centre, (const double *)dinv.p, r.p, z.p,
                           x.p, Pi.p, S.p);
    FFEA_HIP_TRY(hipGetLastError());
    t_end(2, t0);

    // CG iterations, queued in chunks. Iteration i is an SpMV step (which first tests
    // the convergence of iteration i-1) and a vector step. Kernels queued after the
    // solve stopped return at once; the host stops queuing once a completed chunk
    // reports done.
    const int imax = prm.max_iterations_cg;
    int launched = 0, chunk_idx = 0;
    int chunk = std::max(4, expect_iters - 2);
    bool stop = false;
    while (launched <= imax && !stop) {
        const int nit = std::min(chunk, imax + 1 - launched);
        for (int k = 0; k < nit; ++k) {
            const int it = launched + k;
            hipEvent_t ts = nullptr;
            if (t_cg(3, it)) t_begin(3, &ts);
            SpmvArgs<T> sa{nl, rows_per_block, it, imax, prm.epsilon2, slice_w.p, slice_base.p, col.p, val, z.p,
                           pos.p, centre, p.p, q.p, it == 0 ? Pi.p : P2.p, P1.p, S.p};
            hipLaunchKernelGGL((k_cg_spmv<T, PROJ>), dim3(nb), dim3(kBlock), 0, stream, sa);
            t_end(3, ts, it);
            if (it == imax) break;  // that SpMV step only tests the last iteration
            hipEvent_t tv = nullptr;
            if (t_cg(4, it)) t_begin(4, &tv);
            VecArgs va{nl, rows_per_block, it, p.p, q.p, dinv.p, pos.p, centre, x.p, r.p, z.p, P1.p, P2.p, S.p};
            hipLaunchKernelGGL((k_cg_vec<PROJ>), dim3(nb), dim3(kBlock), 0, stream, va);
            t_end(4, tv, it);
        }
        FFEA_HIP_TRY(hipGetLastError());
        launched += nit;
        const int slot = chunk_idx & 1;
        FFEA_HIP_TRY(hipMemcpyAsync(&h_status[slot], S.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipEventRecord(chunk_ev[slot], stream));
        if (chunk_idx > 0) {
            FFEA_HIP_TRY(hipEventSynchronize(chunk_ev[slot ^ 1]));
            if (h_status[slot ^ 1].done) stop = true;
        }
        ++chunk_idx;
        chunk = 4;
    }
}

int ffea_gpu_blob::solve_and_update() {
    // matrix assembly
    hipEvent_t t0;
    t_begin(1, &t0);
    if (mixed) {
        AsmArgs<float> a{nl, ne, slice_w.p, slice_base.p, cptr.p, contrib.p, diag_j.p, f_geo.p, f_A.p, f_B.p,
                         f_val.p, dinv.p};
        hipLaunchKernelGGL(k_assemble<float>, dim3(grid_nodes), dim3(kBlock), 0, stream, a);
    } else {
        AsmArgs<double> a{nl, ne, slice_w.p, slice_base.p, cptr.p, contrib.p, diag_j.p, d_geo.p, d_A.p, d_B.p,
                          d_val.p, dinv.p};
        hipLaunchKernelGGL(k_assemble<double>, dim3(grid_nodes), dim3(kBlock), 0, stream, a);
    }
    FFEA_HIP_TRY(hipGetLastError());
    t_end(1, t0);
    if (mixed) {
        if (project) launch_solve<float, true>(f_val.p, f_force.p);
        else launch_solve<float, false>(f_val.p, f_force.p);
    } else {
        launch_solve<double, false>(d_val.p, d_force.p);
    }
    // Euler update (skipped unless converged)
    t_begin(5, &t0);
    hipLaunchKernelGGL(k_update, dim3(grid_nodes), dim3(kBlock), 0, stream, nl, (const double *)x.p, vel.p, pos.p,
                       prm.dt, S.p);
    FFEA_HIP_TRY(hipGetLastError());
    t_end(5, t0);
    FFEA_HIP_TRY(hipMemcpyAsync(&h_status[2], S.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    const CgScalars &st = h_status[2];
    collect_timing(st.iter);
    last_iters = st.iter;
    total_iters += st.iter;
    if (st.iter > 0) expect_iters = st.iter;
    if (st.inverted_step > 0) {
        const int inv = st.inverted_step;
        err = inv == 1 ? "1 element has inverted since the last step. Aborting simulation.\n"
                       : fmt("%d elements have inverted since the last step. Aborting simulation.\n", inv);
        return FFEA_GPU_ERR_INVERSION;
    }
    if (!st.converged) {
        err = fmt("Conjugate gradient solver: Could not converge after %d iterations.\n\tEither epsilon or "
                  "max_iterations_cg are set too low, or something went wrong with the simulation.\n",
                  prm.max_iterations_cg);
        return FFEA_GPU_ERR_CG_NO_CONVERGE;
    }
    return FFEA_GPU_OK;
}

// ================================================================ C ABI
void ffea_internal_set_error(const std::string &msg) { g_last_error = msg; }

namespace {
template <typename F>
int guarded(ffea_gpu_blob *b, F &&f) {
    try {
        return f();
    } catch (const HipError &e) {
        if (b) b->err = e.msg;
        g_last_error = e.msg;
        return FFEA_GPU_ERR_HIP;
    } catch (const Invalid &e) {
        if (b) b->err = e.msg;
        g_last_error = e.msg;
        return FFEA_GPU_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        if (b) b->err = "host allocation failed";
        g_last_error = "host allocation failed";
        return FFEA_GPU_ERR_INVALID;
    }
}
}  // namespace

extern "C" {

int ffea_gpu_blob_create(const ffea_gpu_params *params, int num_nodes, int num_elements, const double *pos,
                         const int *top, const double *A, const double *B, const double *G, const double *E,
                         int num_pinned, const int *pinned, int num_ctf, const int *ctf_node,
                         const double *ctf_force, const double *stokes_drag, const uint32_t rng_state[6],
                         ffea_gpu_blob **out) {
    if (out) *out = nullptr;
    if (!params || !out || !pos || !top || !A || !B || !G || !E || !rng_state || num_nodes <= 0 ||
        num_elements <= 0 || num_pinned < 0 || num_ctf < 0 || (num_pinned && !pinned) ||
        (num_ctf && (!ctf_node || !ctf_force))) {
        g_last_error = "ffea_gpu_blob_create: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    if (params->calc_stokes || stokes_drag) {
        g_last_error = "ffea_gpu_blob_create: calc_stokes is not supported by this version";
        return FFEA_GPU_ERR_INVALID;
    }
    if (params->precision != FFEA_GPU_FP64 && params->precision != FFEA_GPU_MIXED) {
        g_last_error = "ffea_gpu_blob_create: unknown precision";
        return FFEA_GPU_ERR_INVALID;
    }
    if (!(params->dt > 0) || params->max_iterations_cg <= 0 || !(params->epsilon2 >= 0)) {
        g_last_error = "ffea_gpu_blob_create: dt, epsilon2 and max_iterations_cg must be positive";
        return FFEA_GPU_ERR_INVALID;
    }
    if (!mrg::h_valid_state(rng_state)) {
        g_last_error = "ffea_gpu_blob_create: invalid MRG32k3a state";
        return FFEA_GPU_ERR_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_last_error = "ffea_gpu_blob_create: no HIP device";
        return FFEA_GPU_ERR_NO_DEVICE;
    }
    if (params->device < 0 || params->device >= ndev) {
        g_last_error = "ffea_gpu_blob_create: device ordinal out of range";
        return FFEA_GPU_ERR_INVALID;
    }
    std::unique_ptr<ffea_gpu_blob> b(new ffea_gpu_blob());
    b->prm = *params;
    b->nt = num_nodes;
    b->ne = num_elements;
    b->device = params->device;
    b->mixed = params->precision == FFEA_GPU_MIXED;
    int rc = guarded(b.get(), [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        FFEA_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->build(pos, top, A, B, G, E, num_pinned, pinned, num_ctf, ctf_node, ctf_force, rng_state);
        return (int)FFEA_GPU_OK;
    });
    if (rc != FFEA_GPU_OK) {
        g_last_error = b->err;
        return rc;
    }
    *out = b.release();
    return FFEA_GPU_OK;
}

void ffea_gpu_blob_destroy(ffea_gpu_blob *blob) {
    if (!blob) return;
    (void)hipSetDevice(blob->device);
    delete blob;
}

int ffea_gpu_blob_update_internal_forces(ffea_gpu_blob *b) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (b->all_pinned) return (int)FFEA_GPU_OK;
        b->launch_element();
        return b->check_status_after_forces();
    });
}

int ffea_gpu_blob_update_positions(ffea_gpu_blob *b) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        int rc = b->solve_and_update();
        b->collect_timing();
        return rc;
    });
}

int ffea_gpu_blob_step(ffea_gpu_blob *b, int nsteps) {
    if (!b || nsteps < 0) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        for (int s = 0; s < nsteps; ++s) {
            if (b->all_pinned) continue;
            b->launch_element();
            int rc = b->solve_and_update();
            if (rc != FFEA_GPU_OK) {
                b->collect_timing();
                return rc;
            }
        }
        b->collect_timing();
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_positions(ffea_gpu_blob *b, double *out) {
    if (!b || !out) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        std::vector<double> p((size_t)3 * b->nl);
        FFEA_HIP_TRY(hipMemcpyAsync(p.data(), b->pos.p, sizeof(double) * p.size(), hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        for (int n = 0; n < b->nl; ++n)
            for (int c = 0; c < 3; ++c) out[3 * b->orig_of_lin[n] + c] = p[3 * n + c];
        // tetra_element_linear::linearise_element (tetra_element_linear.cpp:553-577)
        for (int v = 0; v < b->nt; ++v)
            if (b->lin_of[v] < 0) {
                if (b->mid_a[v] < 0) {
                    out[3 * v] = out[3 * v + 1] = out[3 * v + 2] = 0.0;
                    continue;
                }
                const double *pa = &p[3 * b->lin_of[b->mid_a[v]]], *pb = &p[3 * b->lin_of[b->mid_b[v]]];
                for (int c = 0; c < 3; ++c) out[3 * v + c] = .5 * (pa[c] + pb[c]);
            }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_set_positions(ffea_gpu_blob *b, const double *in) {
    if (!b || !in) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        std::vector<double> p((size_t)3 * b->nl);
        for (int n = 0; n < b->nl; ++n)
            for (int c = 0; c < 3; ++c) p[3 * n + c] = in[3 * b->orig_of_lin[n] + c];
        FFEA_HIP_TRY(hipMemcpyAsync(b->pos.p, p.data(), sizeof(double) * p.size(), hipMemcpyHostToDevice, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

static int get_node_vec(ffea_gpu_blob *b, const double *dev, double *out, bool tiled) {
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        std::vector<double> p(tiled ? (size_t)b->nslices * 3 * kSlice : (size_t)3 * b->nl);
        FFEA_HIP_TRY(hipMemcpyAsync(p.data(), dev, sizeof(double) * p.size(), hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        std::fill(out, out + (size_t)3 * b->nt, 0.0);
        for (int n = 0; n < b->nl; ++n)
            for (int c = 0; c < 3; ++c) out[3 * b->orig_of_lin[n] + c] = tiled ? p[vi(n, c)] : p[3 * n + c];
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_velocities(ffea_gpu_blob *b, double *out) {
    if (!b || !out) return FFEA_GPU_ERR_INVALID;
    return get_node_vec(b, b->vel.p, out, false);
}

int ffea_gpu_blob_get_rhs(ffea_gpu_blob *b, double *out) {
    if (!b || !out) return FFEA_GPU_ERR_INVALID;
    return get_node_vec(b, b->rhs.p, out, true);
}

int ffea_gpu_blob_get_element_forces(ffea_gpu_blob *b, double *out) {
    if (!b || !out) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const size_t n = (size_t)12 * b->ne;
        if (b->mixed) {
            std::vector<float> f(n);
            FFEA_HIP_TRY(hipMemcpyAsync(f.data(), b->f_force.p, sizeof(float) * n, hipMemcpyDeviceToHost, b->stream));
            FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
            for (size_t i = 0; i < n; ++i) out[i] = f[i];
        } else {
            FFEA_HIP_TRY(hipMemcpyAsync(out, b->d_force.p, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
            FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_cg_iterations(ffea_gpu_blob *b, int *last, long long *total) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (last) *last = b->last_iters;
    if (total) *total = b->total_iters;
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_get_rng_state(ffea_gpu_blob *b, uint32_t state[6]) {
    if (!b || !state) return FFEA_GPU_ERR_INVALID;
    for (int i = 0; i < 6; ++i) state[i] = b->rng[i];
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_set_rng_state(ffea_gpu_blob *b, const uint32_t state[6]) {
    if (!b || !state) return FFEA_GPU_ERR_INVALID;
    if (!mrg::h_valid_state(state)) {
        b->err = "invalid MRG32k3a state";
        return FFEA_GPU_ERR_INVALID;
    }
    for (int i = 0; i < 6; ++i) b->rng[i] = state[i];
    return FFEA_GPU_OK;
}

int ffea_gpu_advance_rng_state(const uint32_t state[6], long long ndraws, uint32_t out[6]) {
    if (!state || !out || ndraws < 0 || !mrg::h_valid_state(state)) return FFEA_GPU_ERR_INVALID;
    uint32_t t[6];
    mrg::h_apply(mrg::h_pow((unsigned long long)ndraws), state, t);
    for (int i = 0; i < 6; ++i) out[i] = t[i];
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_set_rng_stream(ffea_gpu_blob *b, const uint32_t state[6], long long draws_per_step) {
    if (!b || !state || draws_per_step < 0) return FFEA_GPU_ERR_INVALID;
    if (!mrg::h_valid_state(state)) {
        b->err = "invalid MRG32k3a state";
        return FFEA_GPU_ERR_INVALID;
    }
    for (int i = 0; i < 6; ++i) b->rng[i] = state[i];
    b->step_jump = mrg::h_pow((unsigned long long)draws_per_step);
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_get_info(ffea_gpu_blob *b, ffea_gpu_blob_info *info) {
    if (!b || !info) return FFEA_GPU_ERR_INVALID;
    info->num_nodes = b->nt;
    info->num_elements = b->ne;
    info->num_linear_nodes = b->nl;
    info->num_blocks = b->nblocks;
    info->num_block_slots = b->nslots;
    info->precision = b->prm.precision;
    info->device_bytes = (long long)b->dev_bytes;
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_set_timing(ffea_gpu_blob *b, unsigned phase_mask) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        b->collect_timing();
        b->timing = phase_mask & ((1u << FFEA_GPU_NUM_PHASES) - 1u);
        for (auto &p : b->phase) {
            p.ms = 0;
            p.launches = 0;
            p.ev.clear();
        }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_timing(ffea_gpu_blob *b, double ms[FFEA_GPU_NUM_PHASES], long long launches[FFEA_GPU_NUM_PHASES]) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        b->collect_timing();
        for (int i = 0; i < FFEA_GPU_NUM_PHASES; ++i) {
            if (ms) ms[i] = b->phase[i].ms;
            if (launches) launches[i] = b->phase[i].launches;
        }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_synchronize(ffea_gpu_blob *b) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_copy_probe(int device, long long bytes, int iters, double *ms_per_copy) {
    if (bytes <= 0 || iters <= 0 || !ms_per_copy) return FFEA_GPU_ERR_INVALID;
    return guarded(nullptr, [&]() {
        FFEA_HIP_TRY(hipSetDevice(device));
        const size_t n2 = (size_t)bytes / 16;
        double2 *x = nullptr, *y = nullptr;
        FFEA_HIP_TRY(hipMalloc(&x, n2 * 16));
        FFEA_HIP_TRY(hipMalloc(&y, n2 * 16));
        FFEA_HIP_TRY(hipMemset(x, 0, n2 * 16));
        hipEvent_t e0, e1;
        FFEA_HIP_TRY(hipEventCreate(&e0));
        FFEA_HIP_TRY(hipEventCreate(&e1));
        const int grid = 256 * 32;
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, 0, x, y, n2);  // warm-up
        FFEA_HIP_TRY(hipEventRecord(e0, 0));
        for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, 0, x, y, n2);
        FFEA_HIP_TRY(hipEventRecord(e1, 0));
        FFEA_HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        FFEA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(x);
        (void)hipFree(y);
        return (int)FFEA_GPU_OK;
    });
}

const char *ffea_gpu_blob_last_error(const ffea_gpu_blob *b) { return b ? b->err.c_str() : g_last_error.c_str(); }
const char *ffea_gpu_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
