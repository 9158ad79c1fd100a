// Device code of the blob time step (MI355X, gfx950). Included by ffea_blob.hip.
//
// A handle holds one or more independent blobs. Their corner nodes, elements and
// matrix slices are concatenated; every blob's rows start on a 64-row slice boundary.
// Rows are cut into chunks of up to kChunkRows rows inside each blob; every chunk
// contributes one partial sum per reduced quantity, and the kernel that needs a
// blob's total adds that blob's chunk partials in chunk order. Totals are therefore
// independent of the grid size and of which block processed which chunk, and the
// grid of the row kernels is a free tuning parameter. A row-kernel workgroup has one
// thread per row of a chunk, so a chunk's rows are all in flight at once.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <type_traits>

#include "mrg32k3a.h"

namespace ffk {

constexpr int kBlock = 256;          // threads per workgroup of the element kernels: 4 wavefronts of 64
constexpr int kRowBlock = 1024;      // threads per workgroup of the row kernels: one row of a chunk each
constexpr int kMixedBlock = 256;     // ... of the FP32-mixed form's: its projected kernels keep 13-17 FP64
                                     // sums per row, more than the 128 VGPRs of a 1024-thread workgroup;
                                     // its unprojected ones match them, so a blob sums alike alone or
                                     // batched with projected blobs
template <typename T>
constexpr int row_block = std::is_same<T, double>::value ? kRowBlock : kMixedBlock;
constexpr int kSlice = 64;           // rows per ELL slice (one wavefront)
// T per element shape record: dpsi 1,2,3,5,6,7,9,10,11, vol, A, B and 4 padding, so that a
// record is 64 (FP32) or 128 (FP64) aligned bytes that the assembly reads in one piece
constexpr int kGeo = 16;
constexpr int kChunkRows = 1024;     // rows per reduction chunk (one per thread of a row workgroup)
constexpr uint32_t kIdentity = 0xFFFFFFFFu;
constexpr int kFusedMaxRows = 1024;  // blobs up to this many rows solve in one launch (k_cg_fused)
constexpr int kMaxUnrolledRow = 16;  // rows with up to this many blocks take the unrolled SpMV path

// CG node vectors are tiled by 64 nodes: x[64] y[64] z[64] per tile, so row-wise
// loads and stores of one component are 512 contiguous bytes and stencil gathers of
// neighbouring rows stay coalesced
__host__ __device__ __forceinline__ size_t vi(int n, int k) {
    return ((size_t)(n >> 6) * 3 + (size_t)k) * 64 + (size_t)(n & 63);
}

// the FP32-mixed form's packed FP32 copy of a gathered vector, row n: three contiguous
// floats, so a gather of one row is one 12-byte load
__device__ __forceinline__ void store_f32x3(float *y, int n, const double (&v)[3]) {
    float *o = y + 3 * (size_t)n;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = (float)v[k];
}

// Block values of one slot (64 rows x 9 entries) are ordered so that a lane reads its 9
// entries as 16-byte vectors: FP64 as pairs (0,1) (2,3) (4,5) (6,7) then 8, FP32 as
// quads (0..3) (4..7) then 8. Each wave load is one contiguous 1 KiB span.
template <typename T>
__host__ __device__ __forceinline__ size_t val_index(long long slot_base, int j, int lane, int k) {
    const size_t b = (size_t)slot_base * 9 + (size_t)j * 9 * kSlice;
    constexpr int per = 16 / sizeof(T);  // entries per 16-byte vector
    return k < 8 ? b + (size_t)(k / per) * per * kSlice + (size_t)lane * per + (k % per) : b + 8 * kSlice + lane;
}

// per-blob CG state (device memory)
struct CgScalars {
    double delta[2];   // delta[(i-1)&1]: r.z entering iteration i (written by the SpMV step)
    double f2;         // |f|^2 of the right-hand side
    int iter;          // iterations done when the solve stopped
    int done;          // 1: stop (converged, failed or skipped)
    int converged;     // 1: converged this solve
    int inverted;      // inverted elements counted by the element pass (relaxed atomics)
    int inverted_step; // inverted elements of the step being solved
    int updated;       // positions of this solve already applied
};

struct Chunk {
    int blob, begin, end, first;  // rows [begin, end) of blob; first = first chunk of the blob
};

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block sum of NV values; result valid in every thread
template <int NV, int BS>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double sh[BS / 64][NV];
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
        for (int w = 0; w < BS / 64; ++w) s += sh[w][k];
        v[k] = s;
    }
    __syncthreads();
}

// chunk partials, stored [k][nchunks]
template <int NV, int BS>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *P, int nchunks, int c) {
    block_sum<NV, BS>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) P[(size_t)k * nchunks + c] = v[k];
}

// totals of a blob's chunk partials [c0, c1): identical in every block
// The NV loads of one chunk are issued together (one memory latency per chunk a thread
// visits, not one per sum); each sum still adds its chunks in ascending order.
template <int NV, int BS>
__device__ __forceinline__ void sum_partials(const double *P, int nchunks, int c0, int c1, double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = 0.0;
    for (int c = c0 + threadIdx.x; c < c1; c += BS) {
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = P[(size_t)k * nchunks + c];
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] += v[k];
    }
    block_sum<NV, BS>(out);
}

// The chunks of this block: a contiguous range; blocks are mapped so that each of the
// 8 XCDs (which receive workgroups round-robin) owns one contiguous eighth of them.
struct ChunkRange {
    int c0, c1;
};
__device__ __forceinline__ ChunkRange my_chunks(int nchunks, int cpb) {
    const int nb = gridDim.x, b = blockIdx.x;
    const int L = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int c0 = min(nchunks, L * cpb);
    return {c0, min(nchunks, c0 + cpb)};
}

// rigid-mode basis of a corner node at x (relative to a fixed centre): columns are
// the 3 translations and the 3 infinitesimal rotations e_k cross x. Z^T v = (v, x cross v).
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
__device__ inline void invert6(const double *M, double *Minv) {
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
__device__ inline void ztz_from_moments(const double *m, double *M) {
    for (int i = 0; i < 36; ++i) M[i] = 0.0;
    const double N = m[0], sx = m[1], sy = m[2], sz = m[3];
    const double xx = m[4], yy = m[5], zz = m[6], xy = m[7], xz = m[8], yz = m[9];
    M[0] = M[7] = M[14] = N;
    const double Tr[3][3] = {{0, sz, -sy}, {-sz, 0, sx}, {sy, -sx, 0}};  // sum_n (e_k x x_n)_i as [i][k]
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

// Entry (4i+a, 4j+b) of an element's 12x12 viscosity matrix, evaluated as the
// reference's create_viscosity_matrix does (tetra_element_linear.cpp:298-327, macros
// tetra_element_linear.h:56-117, del2 :585-622): upper triangle computed, the lower
// one mirrored, everything times vol at the end.
template <typename T>
__host__ __device__ __forceinline__ T visc_entry(const T *d, T A, T B, T vol, int i, int a, int j, int b) {
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

// The 3x3 block (a, b) of the same matrix for corners known only at run time, entry (i, j)
// equal to visc_entry(d, A, B, vol, i, a, j, b) bit for bit (the same products and sums:
// IEEE products and sums of two are commutative). The two corners' shape gradients
// ga = d[4c + a], gb = d[4c + b] are picked once; indexing d with a run-time corner in every
// entry instead costs a chain of selects per access.
template <typename T>
__host__ __device__ __forceinline__ void visc_block_g(const T (&ga)[3], const T (&gb)[3], bool same, T A, T B, T vol,
                                                      T (&out)[9]) {
#pragma clang fp contract(off)
    const T del2 = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T v;
            if (i == j) {
                v = ga[i] * gb[i] * (A + B);
                v += del2 * A;
            } else {
                const T vs = (B + A) * (ga[i] * ga[j]);
                const T vd = B * (ga[i] * gb[j]) + A * (gb[i] * ga[j]);
                v = same ? vs : vd;
            }
            out[3 * i + j] = v * vol;
        }
}
template <typename T>
__host__ __device__ __forceinline__ void visc_block(const T (&d)[12], T A, T B, T vol, int a, int b, T (&out)[9]) {
    T ga[3], gb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ga[c] = a == 0 ? d[4 * c] : a == 1 ? d[4 * c + 1] : a == 2 ? d[4 * c + 2] : d[4 * c + 3];
        gb[c] = b == 0 ? d[4 * c] : b == 1 ? d[4 * c + 1] : b == 2 ? d[4 * c + 2] : d[4 * c + 3];
    }
    visc_block_g<T>(ga, gb, a == b, A, B, vol, out);
}

// ---------------------------------------------------------------- element pass
template <typename T>
struct ElemArgs {
    int ne;                       // elements of all blobs
    const int *conn;              // [4][ne] corner-node indices
    const double *pos;            // [nl][3]
    const T *jinv0;               // [9][ne]
    const T *vol0, *G, *E, *A, *B;
    double *last_det;             // [ne]
    T *force;                     // [4 ne][3]: each corner's force at its place in the incidence lists
    T *geo;                       // [ne][kGeo]: dpsi 1,2,3,5,6,7,9,10,11, vol (A, B left as set)
    int calc_noise;
    double kT, dt;
    const int *elem_blob;         // [ne] blob of each element (null: one blob)
    const int *elem_off;          // [nblob] first element of each blob
    const mrg::Jump *jump_group;  // [ceil(max blob ne / 64)]: A^(7*64*g)
    const mrg::Jump *jump_lane;   // [64]: A^(7*l)
    const uint32_t *rng;          // [nblob][6]: stream state at the start of this pass
    CgScalars *S;                 // [nblob]
    const double *rest;           // [nl][3] rest positions (mixed form: cancellation-free stress)
    const double *vel = nullptr;  // [nl][3] node velocities: inertial solvers' viscous term, or null
    const uint32_t *inc_pos = nullptr;  // [4][ne] where corner q of element e goes in force
};

// Elastic stress without cancellation, for FP32: with H = F - I formed from displacements
// (F - I = (J - J_0)^T J_0^-T, the edge displacements taken in FP64 before rounding),
// F F^T - I = H + H^T + H H^T and det F - 1 = tr H + ((tr H)^2 - tr H^2)/2 + det H, the
// reference's stress (tetra_element_linear.cpp:375-403)
//   s = (G v0/vol) F F^T - G I + [G (1 - v0/vol) + (E - 2G/3)(vol/v0 - v0/vol)/2] I
// is, exactly, s = (G/j)(F F^T - I) + (E - 2G/3) e (2 + e) / (2 j) I with j = 1 + e = vol/v0.
template <typename S, typename T>
__device__ __forceinline__ void strain_terms(const S (&H)[3][3], S (&FFm)[3][3], S *e) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            FFm[i][j] = H[i][j] + H[j][i] + (H[i][0] * H[j][0] + H[i][1] * H[j][1] + H[i][2] * H[j][2]);
    const S tr = H[0][0] + H[1][1] + H[2][2];
    S tr2 = H[0][0] * H[0][0];
    tr2 = tr2 + H[1][1] * H[1][1] + H[2][2] * H[2][2] + (H[0][1] * H[1][0] + H[0][2] * H[2][0] + H[1][2] * H[2][1]) * (T)2;
    const S det = H[0][0] * (H[1][1] * H[2][2] - H[1][2] * H[2][1]) - H[0][1] * (H[1][0] * H[2][2] - H[1][2] * H[2][0]) +
                  H[0][2] * (H[1][0] * H[2][1] - H[1][1] * H[2][0]);
    *e = tr + (tr * tr - tr2) * (T)0.5 + det;
}

// tetra_element_linear.cpp:191-443 and :468-483 for one element (mass-less solver:
// du starts at zero). Arithmetic in the reference's order, no contraction.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_element(ElemArgs<T> a) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.ne) return;
    const int ne = a.ne;
    const int blob = a.elem_blob ? a.elem_blob[e] : 0;
    uint32_t st[6];
    if (a.calc_noise) {
        // the element's 7 draws start 7*e_local draws into the pass's segment of its blob's stream
        const int el = e - a.elem_off[blob];
        for (int i = 0; i < 6; ++i) st[i] = a.rng[6 * blob + i];
        mrg::d_apply(a.jump_group[el >> 6], st);
        mrg::d_apply(a.jump_lane[el & 63], st);
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
        __hip_atomic_fetch_add(&a.S[blob].inverted, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vol = a.geo[(size_t)kGeo * e + 9];  // the reference keeps the previous volume
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
    for (int k = 0; k < 9; ++k) Ji[k / 3][k % 3] = a.jinv0[(size_t)k * ne + e];
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
    if (std::is_same<T, float>::value && a.rest && ld * (double)det >= 0) {
        // FP32: the same stress from the displacement gradient (see strain_terms)
        double u0[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) u0[c] = x0[c] - a.rest[3 * n[0] + c];
        T DJ[3][3], H[3][3], FFm[3][3], ee;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                DJ[r][c] = (T)((a.pos[3 * n[r + 1] + c] - a.rest[3 * n[r + 1] + c]) - u0[c]);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) H[i][j] = DJ[0][i] * Ji[j][0] + DJ[1][i] * Ji[j][1] + DJ[2][i] * Ji[j][2];
        strain_terms<T, T>(H, FFm, &ee);
        const T jr = (T)1 / ((T)1 + ee);
        const T bulk = (E - G * (T)2.0 / (T)3.0) * ee * ((T)2 + ee) * (T)0.5 * jr;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s[i][j] = G * jr * FFm[i][j] + (i == j ? bulk : (T)0);
    } else {
        const T sc = G * v0 / vol;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) s[i][j] *= sc;
        s[0][0] -= G;
        s[1][1] -= G;
        s[2][2] -= G;
        // bulk elastic stress (:397-403)
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
    // inertial solvers: du = V v_e, the element viscosity matrix times its corner
    // velocities (get_element_velocity_vector, :448-463; mat12_apply, mat_vec_fns.cpp:28-36)
    T visc[12];
#pragma unroll
    for (int r = 0; r < 12; ++r) visc[r] = (T)0;
    if (a.vel) {
        const T A = a.A[e], B = a.B[e];
        T v[12];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[4 * i + q] = (T)a.vel[3 * n[q] + i];
        for (int r = 0; r < 12; ++r) {
            T acc = (T)0;
            for (int j = 0; j < 12; ++j) acc += visc_entry<T>(d, A, B, vol, r >> 2, r & 3, j >> 2, j & 3) * v[j];
            visc[r] = acc;
        }
    }
    // apply_stress_tensor (:436-443) then node_force -= du (:468-483)
    T out[12];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T du = visc[4 * i + q] + vol * (d[q] * s[i][0] + d[4 + q] * s[i][1] + d[8 + q] * s[i][2]);
            out[3 * q + i] = (T)0 - du;
        }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        T *fo = a.force + 3 * (size_t)a.inc_pos[(size_t)q * ne + e];
#pragma unroll
        for (int i = 0; i < 3; ++i) fo[i] = out[3 * q + i];
    }
    const int gi[9] = {1, 2, 3, 5, 6, 7, 9, 10, 11};
    T *go = a.geo + (size_t)kGeo * e;
#pragma unroll
    for (int k = 0; k < 9; ++k) go[k] = d[gi[k]];
    go[9] = vol;
}

// the pass consumed 7 draws per element of each blob (tetra_element_linear.cpp:416-426)
__global__ void k_rng_advance(int nblob, uint32_t *rng, const mrg::Jump *step_jump) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblob) return;
    uint32_t s[6];
    for (int i = 0; i < 6; ++i) s[i] = rng[6 * b + i];
    mrg::d_apply(step_jump[b], s);
    for (int i = 0; i < 6; ++i) rng[6 * b + i] = s[i];
}

// Stokes noise (Blob.cpp:3200-3222): after the element pass's draws, 3 draws per node for
// every node of the blob's .node file (midside and pinned ones included), in node order;
// node i's force loses RAND(-.5,.5) * sqrt(24 kT drag_i / dt) per component. The
// subtrahend is stored for the gather, which applies and folds it.
struct NoiseArgs {
    int nnodes;                    // nodes of all blobs
    const int *node_blob;          // [nnodes] or null (one blob)
    const int *node_off;           // [nblob] first node of each blob
    const double *sc;              // [nnodes] sqrt(24 kT drag / dt)
    const uint32_t *rng;           // [nblob][6] stream state at the start of the step's draws
    const mrg::Jump *node_jump;    // [nblob] A^(7 * elements of the blob)
    const mrg::Jump *jump_group;   // [ceil(max nodes / 64)] A^(3*64*g)
    const mrg::Jump *jump_lane;    // [64] A^(3*l)
    double *noise;                 // [nnodes][3]
};
__global__ __launch_bounds__(kBlock) void k_node_noise(NoiseArgs a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.nnodes) return;
    const int blob = a.node_blob ? a.node_blob[i] : 0;
    const int il = i - a.node_off[blob];
    uint32_t st[6];
    for (int k = 0; k < 6; ++k) st[k] = a.rng[6 * blob + k];
    mrg::d_apply(a.node_jump[blob], st);
    mrg::d_apply(a.jump_group[il >> 6], st);
    mrg::d_apply(a.jump_lane[il & 63], st);
    const double sc = a.sc[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double u = mrg::d_u01(st);
        a.noise[3 * i + c] = ((-.5) + ((.5) - (-.5)) * u) * sc;
    }
}

// ---------------------------------------------------------------- matrix assembly
// The multigrid smoother's inverse of a row's 3x3 diagonal block B: the symmetric inverse
// (adjugate over determinant) of the symmetrised block, stored as 6 FP32 values (xx, xy,
// xz, yy, yz, zz) per row; where B is not positive definite, the inverses of its diagonal.
__device__ __forceinline__ void store_block_inverse(float *binv, int n, const double (&m)[9]) {
    const double a = m[0], d = m[4], f = m[8];
    const double b = 0.5 * (m[1] + m[3]), c = 0.5 * (m[2] + m[6]), e = 0.5 * (m[5] + m[7]);
    const double A = d * f - e * e, B = c * e - b * f, C = b * e - c * d;
    const double det = a * A + b * B + c * C;
    float o[6];
    if (det > 0.0 && a > 0.0 && a * d - b * b > 0.0 && det < 1e300) {
        const double s = 1.0 / det;
        o[0] = (float)(A * s);
        o[1] = (float)(B * s);
        o[2] = (float)(C * s);
        o[3] = (float)((a * f - c * c) * s);
        o[4] = (float)((b * c - a * e) * s);
        o[5] = (float)((a * d - b * b) * s);
    } else {
        o[0] = (float)(1.0 / a);
        o[1] = o[2] = o[4] = 0.0f;
        o[3] = (float)(1.0 / d);
        o[5] = (float)(1.0 / f);
    }
    float2 *dst = (float2 *)(binv + 6 * (size_t)n);
    dst[0] = make_float2(o[0], o[1]);
    dst[1] = make_float2(o[2], o[3]);
    dst[2] = make_float2(o[4], o[5]);
}

// z = D^-1 r (Jacobi), or z = B^-1 r where the multigrid preconditioner's block smoother
// is on (binv: the rows' stored block inverses, else null)
__device__ __forceinline__ void jacobi_times(const double *dinv, const float *binv, int n, const double (&r)[3],
                                             double (&z)[3]) {
    if (binv) {
        const float2 *src = (const float2 *)(binv + 6 * (size_t)n);
        const float2 u = src[0], v = src[1], w = src[2];
        const double xx = u.x, xy = u.y, xz = v.x, yy = v.y, yz = w.x, zz = w.y;
        z[0] = xx * r[0] + xy * r[1] + xz * r[2];
        z[1] = xy * r[0] + yy * r[1] + yz * r[2];
        z[2] = xz * r[0] + yz * r[1] + zz * r[2];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) z[k] = dinv[vi(n, k)] * r[k];
    }
}

template <typename T>
struct AsmArgs {
    int nrows, ne;
    const int *slice_w;          // [nslices]
    const long long *slice_base; // [nslices] first slot of the slice
    const uint32_t *cptr;        // [nslots+1]
    const uint32_t *contrib;     // (e<<4)|(a<<2)|b, element order; kIdentity = identity block
    const unsigned char *diag_j; // [nrows] slot index j of the diagonal block
    const T *geo;                // [ne][kGeo]
    T *val;                      // [nslots*9]
    double *dinv;                // tiled [nrows][3]
    const double *drag;          // [nrows] Stokes drag of unpinned rows, or null
    _Float16 *hval;              // [nslots*9] the multigrid preconditioner's scaled copy, or null
    double hscale;               // its scale
    float *binv = nullptr;       // [nrows][6] the multigrid smoother's diagonal-block inverses, or null
};

// the first 12 T of a shape record in 16-byte loads
template <typename T>
__device__ __forceinline__ void load_geo(const T *gr, T (&g)[12]) {
    if constexpr (std::is_same<T, float>::value) {
        const float4 *v = reinterpret_cast<const float4 *>(gr);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float4 x = v[i];
            g[4 * i] = x.x; g[4 * i + 1] = x.y; g[4 * i + 2] = x.z; g[4 * i + 3] = x.w;
        }
    } else {
        const double2 *v = reinterpret_cast<const double2 *>(gr);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const double2 x = v[i];
            g[2 * i] = x.x; g[2 * i + 1] = x.y;
        }
    }
}

// A and B of every element into its shape record (set up, and after new materials)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_geo_material(int ne, const T *A, const T *B, T *geo) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    geo[(size_t)kGeo * e + 10] = A[e];
    geo[(size_t)kGeo * e + 11] = B[e];
}


// the 12 used values of a shape record as the 3x4 shape-derivative array d (column 0 from
// the other three), its volume and A, B
template <typename T>
__device__ __forceinline__ void unpack_geo(const T (&gr)[12], T (&d)[12], T &vol, T &A, T &B) {
#pragma clang fp contract(off)
    d[1] = gr[0]; d[2] = gr[1]; d[3] = gr[2];
    d[5] = gr[3]; d[6] = gr[4]; d[7] = gr[5];
    d[9] = gr[6]; d[10] = gr[7]; d[11] = gr[8];
    vol = gr[9];
    d[0] = -(d[1] + d[2] + d[3]);
    d[4] = -(d[5] + d[6] + d[7]);
    d[8] = -(d[9] + d[10] + d[11]);
    A = gr[10];
    B = gr[11];
}

// a summed block (+ the row's drag on its diagonal) into the matrix, its FP16 copy and D^-1
template <typename T>
__device__ __forceinline__ void store_block(const AsmArgs<T> &a, int n, int j, int lane, long long base,
                                            double (&acc)[9]) {
#pragma clang fp contract(off)
    if (a.drag && j == a.diag_j[n]) {  // NoMassCGSolver.cpp:119-127: after the element sources
        const double dr = a.drag[n];
        acc[0] += dr;
        acc[4] += dr;
        acc[8] += dr;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) a.val[val_index<T>(base, j, lane, k)] = (T)acc[k];
    if (a.hval)
#pragma unroll
        for (int k = 0; k < 9; ++k) a.hval[val_index<_Float16>(base, j, lane, k)] = (_Float16)(acc[k] * a.hscale);
    if (j == a.diag_j[n]) {
#pragma unroll
        for (int i = 0; i < 3; ++i) a.dinv[vi(n, i)] = 1.0 / (double)(T)acc[4 * i];
        if (a.binv) {
            double m[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) m[k] = (double)(T)acc[k];
            store_block_inverse(a.binv, n, m);
        }
    }
}

// The slot a row takes at step p: its diagonal block first (about 24 sources in a tet mesh,
// against 1-6 for an off-diagonal one), then the others in slot order. Rows of a wavefront
// have their diagonals in different slots; taken in slot order, every step would wait for
// the lane at its diagonal. The order of the blocks does not change any block's sum.
__device__ __forceinline__ int diag_first(int p, int dj) {
    if (dj == 0xFF) return p;
    return p == 0 ? dj : (p <= dj ? p - 1 : p);
}

// one row's slots, steps s, s + L, ... of diag_first: each block the sum of its element
// sources in element order (+ drag on the diagonal), written to the matrix, its FP16 copy
// and D^-1. source(c, gr) returns the record of source c in gr, or false for an identity
// source.
template <typename T, int L, typename Source>
__device__ __forceinline__ void assemble_row(const AsmArgs<T> &a, int n, int s, Source source) {
#pragma clang fp contract(off)
    const int sl = n / kSlice, lane = n % kSlice;
    const int W = a.slice_w[sl];
    const long long base = a.slice_base[sl];
    const int dj = a.diag_j[n];
    for (int p = s; p < W; p += L) {
        const int j = diag_first(p, dj);
        const long long slot = base + (long long)j * kSlice + lane;
        const uint32_t c0 = a.cptr[slot], c1 = a.cptr[slot + 1];
        double acc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = 0.0;
        for (uint32_t c = c0; c < c1; ++c) {
            T gr[12];
            int pa, pb;
            if (!source(c, gr, pa, pb)) {
                acc[0] += 1.0; acc[4] += 1.0; acc[8] += 1.0;
                continue;
            }
            T d[12], vol, A, B, blk[9];
            unpack_geo<T>(gr, d, vol, A, B);
            visc_block<T>(d, A, B, vol, pa, pb, blk);
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[k] += (double)blk[k];
        }
        store_block<T>(a, n, j, lane, base, acc);
    }
}

// the global source code (e<<4)|(a<<2)|b: the record from the element's shape record
template <typename T>
__device__ __forceinline__ bool global_source(const AsmArgs<T> &a, uint32_t c, T (&gr)[12], int &pa, int &pb) {
    const uint32_t code = a.contrib[c];
    if (code == kIdentity) return false;
    const int e = (int)(code >> 4);
    pa = (code >> 2) & 3;
    pb = code & 3;
    load_geo(a.geo + (size_t)kGeo * e, gr);
    return true;
}

// SparseMatrixFixedPattern::build + calc_inverse_diagonal (SparseMatrixFixedPattern.cpp:86-94,
// 305-312): every block entry is the sum of its element sources in element order.
// L lanes share a row, lane s taking the row's slots s, s + L, ...: L = 1 for large
// meshes (a workgroup reads the shape records of its rows' elements while they are
// cached; measured at 100^3 against L = 4 and 16, which read them together across lanes
// but idle on the short off-diagonal chains), L = 16 for small ones (more lanes in
// flight). Each block is the same sum whatever the mapping.
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs<T> a) {
    // XCD-aware order: workgroups reach the 8 XCDs round-robin, so workgroup b takes the
    // (b % 8)-th eighth of the rows; the rows an XCD works on at once are then neighbours
    // in the mesh, whose shape records its L2 shares
    const int nbk = gridDim.x, bk = blockIdx.x;
    const int lb = (nbk % 8 == 0) ? (bk % 8) * (nbk / 8) + bk / 8 : bk;
    const long long t = (long long)lb * kBlock + threadIdx.x;
    const int n = (int)(t / L), s = (int)(t % L);
    if (n >= a.nrows) return;
    assemble_row<T, L>(a, n, s, [&](uint32_t c, T(&gr)[12], int &pa, int &pb) {
        return global_source<T>(a, c, gr, pa, pb);
    });
}

// The same blocks with the shape records staged in LDS. One wavefront per group of
// kStageRows<T> consecutive rows, kStageLanes<T> lanes per row. The group's elements (every
// element with a corner among its rows, ascending) are listed in gel[gptr[g] .. gptr[g+1]);
// the wavefront copies their records to LDS once, as the four corners' shape gradients
// (d[4c + a] at 3a + c, column 0 formed here as unpack_geo forms it) then vol, A, B, so a
// source reads its two corners' gradients at run-time offsets with no selects. Each source
// c names its record by its place in that list, cloc[c] = (place<<4)|(a<<2)|b (0xFFFF:
// identity). A 6M-tet blob's record is thus fetched about 3 times (once per group it
// touches) instead of 16 times (once per block source). A group with more than
// kStageCap<T> elements reads the records from memory instead. Same sums in the same order
// as k_assemble.
#ifndef FFEA_STAGE_ROWS_F32
#define FFEA_STAGE_ROWS_F32 8
#endif
#ifndef FFEA_STAGE_ROWS_F64
#define FFEA_STAGE_ROWS_F64 8
#endif
template <typename T> constexpr int kStageRows = std::is_same<T, double>::value ? FFEA_STAGE_ROWS_F64 : FFEA_STAGE_ROWS_F32;
template <typename T> constexpr int kStageLanes = kSlice / kStageRows<T>;
// a group of a regular tet mesh has about 18.4 elements per row; at 8 rows its records take
// 9.6 KB of LDS in FP32 and 19.2 KB in FP64
template <typename T> constexpr int kStageCap = 20 * kStageRows<T>;
constexpr int kStageRec = 15;  // values per staged record
constexpr uint16_t kLocalIdentity = 0xFFFF;
constexpr int kAsmStaged = -1;  // the assembly launch-shape value that selects k_assemble_staged

template <typename T>
__global__ __launch_bounds__(kSlice) void k_assemble_staged(AsmArgs<T> a, const uint32_t *gptr, const uint32_t *gel,
                                                            const uint16_t *cloc) {
#pragma clang fp contract(off)
    constexpr int G = kStageRows<T>, L = kStageLanes<T>, CAP = kStageCap<T>;
    constexpr int RPT = (CAP + kSlice - 1) / kSlice;  // records per thread
    constexpr int kBatch = 8;                          // sources whose codes load together
    __shared__ T rec[CAP * kStageRec];
    const int nbk = gridDim.x, bk = blockIdx.x;
    const int g = (nbk % 8 == 0) ? (bk % 8) * (nbk / 8) + bk / 8 : bk;
    const int r0 = g * G;
    if (r0 >= a.nrows) return;  // the same for the whole workgroup
    const uint32_t e0 = gptr[g], cnt = gptr[g + 1] - e0;
    const bool staged = cnt <= (uint32_t)CAP;
    if (staged && cnt > 0) {
        // every record first, then the LDS copies, so that all loads are in flight together
        // (addresses past the list repeat its last entry; those copies are not kept)
        T gr[RPT][12];
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            load_geo(a.geo + (size_t)kGeo * gel[e0 + min((uint32_t)(threadIdx.x + k * kSlice), cnt - 1)], gr[k]);
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = threadIdx.x + k * kSlice;
            if (r < (int)cnt) {
                T d[12], vol, A, B;
                unpack_geo<T>(gr[k], d, vol, A, B);
                T *o = rec + r * kStageRec;
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int c = 0; c < 3; ++c) o[3 * q + c] = d[4 * c + q];
                o[12] = vol;
                o[13] = A;
                o[14] = B;
            }
        }
    }
    __syncthreads();
    const int n = r0 + threadIdx.x / L, s = threadIdx.x % L;
    if (!staged) {
        assemble_row<T, L>(a, n, s, [&](uint32_t c, T(&gr)[12], int &pa, int &pb) {
            return global_source<T>(a, c, gr, pa, pb);
        });
        return;
    }
    const int sl = n / kSlice, lane = n % kSlice;
    const int W = a.slice_w[sl];
    const long long base = a.slice_base[sl];
    const int dj = a.diag_j[n];
    for (int p = s; p < W; p += L) {
        const int j = diag_first(p, dj);
        const long long slot = base + (long long)j * kSlice + lane;
        const uint32_t c0 = a.cptr[slot], c1 = a.cptr[slot + 1];
        double acc[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) acc[k] = 0.0;
        for (uint32_t cb = c0; cb < c1; cb += kBatch) {
            uint16_t code[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) code[k] = cloc[min(cb + k, c1 - 1)];
#pragma unroll
            for (int k = 0; k < kBatch; ++k) {
                if (cb + k >= c1) break;
                if (code[k] == kLocalIdentity) {
                    acc[0] += 1.0; acc[4] += 1.0; acc[8] += 1.0;
                    continue;
                }
                const int pa = (code[k] >> 2) & 3, pb = code[k] & 3;
                const T *rp = rec + (code[k] >> 4) * kStageRec;
                const T ga[3] = {rp[3 * pa], rp[3 * pa + 1], rp[3 * pa + 2]};
                const T gb[3] = {rp[3 * pb], rp[3 * pb + 1], rp[3 * pb + 2]};
                T blk[9];
                visc_block_g<T>(ga, gb, pa == pb, rp[13], rp[14], rp[12], blk);
#pragma unroll
                for (int q = 0; q < 9; ++q) acc[q] += (double)blk[q];
            }
        }
        store_block<T>(a, n, j, lane, base, acc);
    }
}

// ConjugateGradientSolver::init (ConjugateGradientSolver.cpp:51-190): the linear tets'
// consistent mass matrix in the viscosity matrix's slots, m I per 3x3 block with m the sum
// over the block's element sources, in element order, of .1 rho vol_0 (a corner with
// itself) or .05 rho vol_0 (two corners); identity sources (pinned rows) give 1. D^-1 from
// the diagonal. One lane per row.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_assemble_mass(AsmArgs<T> a, const double *rho, const T *vol0) {
#pragma clang fp contract(off)
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= a.nrows) return;
    const int sl = n / kSlice, lane = n % kSlice;
    const int W = a.slice_w[sl];
    const long long base = a.slice_base[sl];
    for (int j = 0; j < W; ++j) {
        const long long slot = base + (long long)j * kSlice + lane;
        double m = 0.0;
        for (uint32_t c = a.cptr[slot]; c < a.cptr[slot + 1]; ++c) {
            const uint32_t code = a.contrib[c];
            if (code == kIdentity) {
                m = 1.0;
                continue;
            }
            const int e = (int)(code >> 4), pa = (code >> 2) & 3, pb = code & 3;
            m += (pa == pb ? .1 : .05) * rho[e] * (double)vol0[e];
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) a.val[val_index<T>(base, j, lane, k)] = (T)(k % 4 == 0 ? m : 0.0);
        if (j == a.diag_j[n]) {
#pragma unroll
            for (int i = 0; i < 3; ++i) a.dinv[vi(n, i)] = 1.0 / (double)(T)m;
        }
    }
}

// ---------------------------------------------------------------- shared row-kernel arguments
struct RowArgs {
    int nchunks, cpb;
    const Chunk *chunks;
    const int2 *seg;       // [nblob]: chunk range of each blob
    CgScalars *S;          // [nblob]
    int *num_done;         // blobs whose solve stopped
};

// ---------------------------------------------------------------- surface constant forces
// Blob::apply_ctforces CTF 3 (Blob.cpp:1721-1740) for a single blob, one workgroup: each
// set's face areas at the step's positions (Face::get_area, Face.cpp:169-187; a midside
// node sits at .5 (a + b) of its edge ends, tetra_element_linear.cpp:553-577), their sum in
// face order, and per face the sum over its set entries (set order) of
// (area / (3 total)) F: the force Face::add_force_to_node puts on each of the face's nodes.
struct SurfArgs {
    int nfaces, nsets, nent;
    const int *rows;        // [faces][3][2] corner rows of each face node
    const double *pos;      // [rows][3]
    const double *force;    // [sets][3]
    const int *set_ptr;     // [sets+1]
    const int *entry_face, *entry_set;
    const int *fe_ptr, *fe; // per face: its entries
    double *area, *total;   // scratch [entries], [sets]
    double *face_force;     // [faces][3]
};
__global__ __launch_bounds__(kRowBlock) void k_surface_forces(SurfArgs a) {
#pragma clang fp contract(off)
    for (int e = threadIdx.x; e < a.nent; e += kRowBlock) {
        const int *r = a.rows + 6 * a.entry_face[e];
        double p[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c) p[i][c] = .5 * (a.pos[3 * r[2 * i] + c] + a.pos[3 * r[2 * i + 1] + c]);
        double u[3], v[3], t[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u[c] = p[1][c] - p[0][c];
            v[c] = p[2][c] - p[0][c];
        }
        t[0] = u[1] * v[2] - u[2] * v[1];
        t[1] = u[2] * v[0] - u[0] * v[2];
        t[2] = u[0] * v[1] - u[1] * v[0];
        double m = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) m += t[c] * t[c];
        a.area[e] = .5 * sqrt(m);
    }
    __syncthreads();
    for (int st = threadIdx.x; st < a.nsets; st += kRowBlock) {  // serial sum per set, as the reference
        double tot = 0.0;
        for (int k = a.set_ptr[st]; k < a.set_ptr[st + 1]; ++k) tot += a.area[k];
        a.total[st] = tot;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < a.nfaces; f += kRowBlock) {
        double ff[3] = {0.0, 0.0, 0.0};
        for (int k = a.fe_ptr[f]; k < a.fe_ptr[f + 1]; ++k) {
            const int e = a.fe[k], st = a.entry_set[e];
            const double sc = a.area[e] / (3 * a.total[st]);
#pragma unroll
            for (int c = 0; c < 3; ++c) ff[c] += sc * a.force[3 * st + c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) a.face_force[3 * f + c] = ff[c];
    }
}

// ---------------------------------------------------------------- per-step node forces
// The force of each node whose force is formed each step, one lane per node: its external
// force (Blob::add_force_to_node calls before the step, Blob.cpp:1770-1774, summed in call
// order on the host), then its linear constant forces one by one and then each rotational
// one's pull round its axis at the node's position this step (Blob::apply_ctforces CTF 1
// and 2, Blob.cpp:1679-1719), in list order. A midside node sits at .5 (a + b) of its edge
// ends (tetra_element_linear.cpp:553-577). The total goes to the node's entry of the
// per-node table read by the midside folds (when the handle forms folds each step) and,
// for a corner node, to its row's constant force.
struct DynArgs {
    int n;
    const int *node;        // [n] node
    const int *row;         // [n] its corner row, or -1
    const int *ends;        // [n][2] corner rows whose midpoint is its position
    const int *ext;         // [n] its external force slot, or -1
    const double *ext_force;// [slots][3]
    const int *lptr;        // [n+1] its linear constant forces
    const double *lval;     // [..][3]
    const int *rptr;        // [n+1] its rotational forces
    const double *axis;     // [entries][6] point, unit direction
    const double *F;        // [entries]
    const int *torque;      // [entries] 1: constant torque ('t'), 0: 'f'
    const double *pos;      // [rows][3]
    double *node_ctf;       // [nodes][3], or null
    double *ctf;            // [rows][3]
};

__device__ inline void cross3(const double *u, const double *v, double *w) {
#pragma clang fp contract(off)
    w[0] = u[1] * v[2] - v[1] * u[2];
    w[1] = -u[0] * v[2] + v[0] * u[2];
    w[2] = u[0] * v[1] - v[0] * u[1];
}

// one entry (Blob.cpp:1693-1714): r runs from the node x to its foot on the axis, formed as
// intersectingPointToLine does (mat_vec_fns_II.h:755-777: v2 = (x - a) x d; a node with
// |v2| < 3 eps lies on the axis and gets r = 0; else foot = x + c (d x v2) with
// c = det[(x - a); d; d x (d x v2)] / |d x (d x v2)|^2, mat_vec_fns_II.h:379-386); the
// force is F (r x d), or (F / |r|) (r x d) for a constant torque
__device__ inline void rotational_force(const double *x, const double *ax, double F, int torque, double *f) {
#pragma clang fp contract(off)
    const double *a = ax, *d = ax + 3;
    double pa[3], v2[3], cp[3], w[3], r[3], g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) pa[c] = x[c] - a[c];
    cross3(pa, d, v2);
    double m = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) m += v2[c] * v2[c];
    if (sqrt(m) < 3.0 * __DBL_EPSILON__) {
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = x[c] - x[c];
    } else {
        cross3(d, v2, cp);
        cross3(d, cp, w);
        double det = pa[0] * (d[1] * w[2] - d[2] * w[1]);
        det += pa[1] * (d[2] * w[0] - d[0] * w[2]);
        det += pa[2] * (d[0] * w[1] - d[1] * w[0]);
        const double cc = det / (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = (x[c] + cc * cp[c]) - x[c];
    }
    cross3(r, d, g);
    double s = F;
    if (torque) {
        double rr = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) rr += r[c] * r[c];
        s = F / sqrt(rr);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = s * g[c];
}

__global__ __launch_bounds__(kBlock) void k_node_forces(DynArgs a) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= a.n) return;
    const int ra = a.ends[2 * k], rb = a.ends[2 * k + 1], x_slot = a.ext[k];
    double x[3], f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        x[c] = ra == rb ? a.pos[3 * ra + c] : .5 * (a.pos[3 * ra + c] + a.pos[3 * rb + c]);
        f[c] = x_slot >= 0 ? a.ext_force[3 * x_slot + c] : 0.0;
    }
    for (int e = a.lptr[k]; e < a.lptr[k + 1]; ++e)
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] += a.lval[3 * e + c];
    for (int e = a.rptr[k]; e < a.rptr[k + 1]; ++e) {
        double g[3];
        rotational_force(x, a.axis + 6 * e, a.F[e], a.torque[e], g);
#pragma unroll
        for (int c = 0; c < 3; ++c) f[c] += g[c];
    }
    const int v = a.node[k], row = a.row[k];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (a.node_ctf) a.node_ctf[3 * v + c] = f[c];
        if (row >= 0) a.ctf[3 * row + c] = f[c];
    }
}

// positions of selected rows (-1: none, zeros)
__global__ __launch_bounds__(kBlock) void k_take_rows(int n, const int *rows, const double *pos, double *out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int r = rows[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) out[3 * i + c] = r >= 0 ? pos[3 * r + c] : 0.0;
}

// ---------------------------------------------------------------- gather + CG start
template <typename T>
struct GatherArgs {
    RowArgs ra;
    const uint32_t *inc_ptr;    // [nrows+1]
    const T *force;             // [4 ne][3]: the corner forces of each row's elements, element order
    const double *ctf;          // [nrows][3] or null: constant forces summed in list order
    const uint32_t *fold_ptr;   // [nrows+1] or null
    const double *fold;         // [..][3]: linearise_force increments in element order (static)
    const double *noise;        // [nodes][3] Stokes noise subtrahends, or null
    const int *row_node;        // [nrows] node (noise index) of each corner row (with noise)
    const int3 *fold_mid;       // [..]: midside nodes folded onto the row per record (per-step folds), or null
    const double *node_ctf;     // [nodes][3] constant forces per node (per-step folds), or null
    const int *sf_nf_ptr;       // [nodes+1] surface constant forces: faces holding each node, or null
    const int *sf_nf;           // [..] used faces, face order
    const double *sf_face_force;// [faces][3] force on each of a face's nodes this step
    const unsigned char *pinned;
    const double *dinv;
    const double *pos;          // [nrows][3] (rigid modes)
    const double3 *centre;      // [nblob]
    double *rhs, *r, *z, *x;
    double *P;                  // chunk partials
    const double *mass_vel = nullptr;   // inertial solvers under calc_stokes: [rows][3] velocities
    const double *mass_drag = nullptr;  // and [rows] drags, for the drag force -drag v
    const double *rhs_in = nullptr;     // tiled [rows][3]: a right-hand side handed over whole
                                        // (the solver's own entry point), or null: formed here
    float *z32 = nullptr;               // FP32-mixed form: z's packed FP32 copy
    const float *binv = nullptr;        // multigrid block smoother: z = B^-1 r (jacobi_times)
};

// Blob::aggregate_forces_and_solve (Blob.cpp:3152-3238) for corner nodes, then the
// start of NoMassCGSolver::solve (NoMassCGSolver.cpp:196-220). With PROJ (free blob,
// mixed form) the CG start is left to k_cg_start and the partials are |f|^2, Z^T f
// and the moments of Z^T Z.
template <typename T, bool PROJ>
__global__ __launch_bounds__(row_block<T>) void k_gather(GatherArgs<T> a) {
#pragma clang fp contract(off)
    constexpr int BS = row_block<T>;
    const ChunkRange cr = my_chunks(a.ra.nchunks, a.ra.cpb);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.ra.num_done = 0;
    constexpr int NV = PROJ ? 17 : 2;
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = a.ra.chunks[c];
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = 0.0;
        const double3 cen = a.centre[ch.blob];
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += BS) {
            double f[3] = {0.0, 0.0, 0.0};
            if (a.rhs_in) {  // NoMassCGSolver::solve's x on entry: pinned rows are identity rows
#pragma unroll
                for (int k = 0; k < 3; ++k) f[k] = a.rhs_in[vi(n, k)];
            } else {
                if (a.ctf) {
                    f[0] = a.ctf[3 * n]; f[1] = a.ctf[3 * n + 1]; f[2] = a.ctf[3 * n + 2];
                }
                for (uint32_t k = a.inc_ptr[n]; k < a.inc_ptr[n + 1]; ++k) {
                    const T *src = a.force + 3 * (size_t)k;
                    f[0] += (double)src[0];
                    f[1] += (double)src[1];
                    f[2] += (double)src[2];
                }
                if (a.sf_nf_ptr) {  // surface forces after the element forces, face order (Blob.cpp:3160-3170)
                    const int g = a.row_node[n];
                    for (int k = a.sf_nf_ptr[g]; k < a.sf_nf_ptr[g + 1]; ++k) {
                        const double *ff = a.sf_face_force + 3 * a.sf_nf[k];
                        f[0] += ff[0];
                        f[1] += ff[1];
                        f[2] += ff[2];
                    }
                }
                if (a.mass_vel) {  // inertial solvers: -drag v before the noise (Blob.cpp:3185-3190)
#pragma unroll
                    for (int k = 0; k < 3; ++k) f[k] -= a.mass_vel[3 * n + k] * a.mass_drag[n];
                }
                if (a.fold_mid) {
                    // the node's own noise, then Blob::linearise_force (Blob.cpp:781-799): each
                    // record is one (element, corner) in element order; its midside nodes not
                    // folded by an earlier element contribute half of (constant force + surface
                    // forces - noise)
                    if (a.noise) {
                        const int g = a.row_node[n];
#pragma unroll
                        for (int k = 0; k < 3; ++k) f[k] -= a.noise[3 * g + k];
                    }
                    for (uint32_t k = a.fold_ptr[n]; k < a.fold_ptr[n + 1]; ++k) {
                        const int3 mm = a.fold_mid[k];
                        const int m3[3] = {mm.x, mm.y, mm.z};
                        double s3[3] = {0.0, 0.0, 0.0};
#pragma unroll
                        for (int t = 0; t < 3; ++t) {
                            double val[3] = {0.0, 0.0, 0.0};
                            if (m3[t] >= 0) {
#pragma unroll
                                for (int c = 0; c < 3; ++c) val[c] = a.node_ctf ? a.node_ctf[3 * m3[t] + c] : 0.0;
                                if (a.sf_nf_ptr)
                                    for (int q = a.sf_nf_ptr[m3[t]]; q < a.sf_nf_ptr[m3[t] + 1]; ++q) {
                                        const double *ff = a.sf_face_force + 3 * a.sf_nf[q];
#pragma unroll
                                        for (int c = 0; c < 3; ++c) val[c] += ff[c];
                                    }
                                if (a.noise)
#pragma unroll
                                    for (int c = 0; c < 3; ++c) val[c] -= a.noise[3 * m3[t] + c];
                            }
#pragma unroll
                            for (int c = 0; c < 3; ++c) s3[c] = t == 0 ? val[c] : s3[c] + val[c];
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) f[c] += 0.5 * s3[c];
                    }
                } else if (a.fold_ptr) {
                    for (uint32_t k = a.fold_ptr[n]; k < a.fold_ptr[n + 1]; ++k) {
                        f[0] += a.fold[3 * k];
                        f[1] += a.fold[3 * k + 1];
                        f[2] += a.fold[3 * k + 2];
                    }
                }
                if (a.pinned[n]) f[0] = f[1] = f[2] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) a.rhs[vi(n, k)] = f[k];
            if constexpr (!PROJ) {
                double zz[3];
                const double f3[3] = {f[0], f[1], f[2]};
                jacobi_times(a.dinv, a.binv, n, f3, zz);
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    a.r[vi(n, k)] = f[k];
                    a.x[vi(n, k)] = 0.0;
                    a.z[vi(n, k)] = zz[k];
                }
                if (a.z32) store_f32x3(a.z32, n, zz);
                v[0] += f[0] * zz[0] + f[1] * zz[1] + f[2] * zz[2];
                v[1] += f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
            } else {
                const double x[3] = {a.pos[3 * n] - cen.x, a.pos[3 * n + 1] - cen.y, a.pos[3 * n + 2] - cen.z};
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
        store_partials<NV, BS>(v, a.P, a.ra.nchunks, c);
        if (c == ch.first && threadIdx.x == 0) {
            CgScalars *S = a.ra.S + ch.blob;
            S->inverted_step = S->inverted;
            S->done = 0;
            S->converged = 0;
            S->iter = 0;
            S->updated = 0;
        }
    }
}

// Rigid-mode projection of the right-hand side for a free blob in the mixed form:
// r = f - Z (Z^T Z)^-1 Z^T f, then z = D^-1 r, x = 0, partials r.z.
struct StartArgs {
    RowArgs ra;
    const double *Pg, *rhs, *pos, *dinv;
    const double3 *centre;
    const int *project;         // [nblob]: 1 = free blob, project; 0 = leave r = f
    double *r, *z, *x, *P, *G;  // G: [nblob][36] ((Z^T Z)^-1, zero for unprojected blobs)
    float *z32 = nullptr;       // FP32-mixed form: z's packed FP32 copy
    const float *binv = nullptr;  // multigrid block smoother: z = B^-1 r (jacobi_times)
};
__global__ __launch_bounds__(kMixedBlock) void k_cg_start(StartArgs a) {
    const ChunkRange cr = my_chunks(a.ra.nchunks, a.ra.cpb);
    // (Z^T Z)^-1 and its product with Z^T f: formed once per blob by thread 0, shared
    __shared__ double sGi[36], scf[6];
    int cur = -1;
    double cf[6], f2 = 0.0;
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = a.ra.chunks[c];
        if (ch.blob != cur) {
            cur = ch.blob;
            const int2 sg = a.ra.seg[cur];
            double t[17];
            sum_partials<17, kMixedBlock>(a.Pg, a.ra.nchunks, sg.x, sg.y, t);
            if (threadIdx.x == 0) {
                double M[36], Gi[36];
                const double mom[10] = {t[7], t[8], t[9], t[10], t[11], t[12], t[13], t[14], t[15], t[16]};
                if (a.project[cur]) {
                    ztz_from_moments(mom, M);
                    invert6(M, Gi);
                } else {
                    for (int i = 0; i < 36; ++i) Gi[i] = 0.0;
                }
                for (int i = 0; i < 36; ++i) sGi[i] = Gi[i];
                for (int i = 0; i < 6; ++i) {
                    double acc = 0.0;
                    for (int j = 0; j < 6; ++j) acc += Gi[6 * i + j] * t[1 + j];
                    scf[i] = acc;
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 6; ++i) cf[i] = scf[i];
            f2 = t[0];
        }
        const double3 cen = a.centre[cur];
        double v[1] = {0.0};
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kMixedBlock) {
            const double xx[3] = {a.pos[3 * n] - cen.x, a.pos[3 * n + 1] - cen.y, a.pos[3 * n + 2] - cen.z};
            double zc[3], rl[3], zl[3];
            z_times(xx, cf, zc);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                rl[k] = a.rhs[vi(n, k)] - zc[k];
                a.r[vi(n, k)] = rl[k];
                a.x[vi(n, k)] = 0.0;
            }
            jacobi_times(a.dinv, a.binv, n, rl, zl);
#pragma unroll
            for (int k = 0; k < 3; ++k) a.z[vi(n, k)] = zl[k];
            if (a.z32) store_f32x3(a.z32, n, zl);
            v[0] += rl[0] * zl[0] + rl[1] * zl[1] + rl[2] * zl[2];
        }
        store_partials<1, kMixedBlock>(v, a.P, a.ra.nchunks, c);  // its block sum also orders the sGi reads before reuse
        if (c == ch.first && threadIdx.x == 0) {
            a.ra.S[cur].f2 = f2;
            for (int i = 0; i < 36; ++i) a.G[36 * cur + i] = sGi[i];
        }
    }
}

// ---------------------------------------------------------------- CG iteration
template <typename T>
struct SpmvArgs {
    RowArgs ra;
    int iter, max_iter;
    double epsilon2;
    int zero_rhs_solved;   // 1: a zero right-hand side has the solution x = 0 (adjoint solves);
                           // 0: the reference's arithmetic (0/0 in the first step)
    const int *slice_w;
    const long long *slice_base;
    const int *col;        // [nslots]
    const T *val;          // [nslots*9]
    const double *z;
    const float *z32;      // FP32-mixed form: z's packed FP32 copy [rows][3], which the row products gather
    const double *pos;
    const double3 *centre;
    double *p, *q;         // updated in place
    const double *Pin;     // chunk partials of the previous kernel (start: r.z [and |f|^2]; later: r.r, r.z)
    double *P1;            // chunk partials of this kernel
};

// Iteration `iter` of every blob still running: first the convergence test of the
// blob's previous iteration (|r|^2/|f|^2 < epsilon^2, NoMassCGSolver.cpp:223-237), then
// p = beta p + z (mat_vec_fns.cpp:202-212) and q = V p, formed as q = V z + beta q so
// the row gathers only z (SparseMatrixFixedPattern.cpp:148-187), and the p.q partials
// (NoMassCGSolver.cpp:251-265).
typedef double vd2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));

// one lane's 9 block entries of one slot (val_index order), non-temporal 16-byte loads
template <typename T>
__device__ __forceinline__ void load_block(const T *slot, int lane, T (&w)[9]);
template <>
__device__ __forceinline__ void load_block<double>(const double *slot, int lane, double (&w)[9]) {
#pragma unroll
    for (int h = 0; h < 4; ++h) {
        const vd2 v = __builtin_nontemporal_load((const vd2 *)(slot + h * 2 * kSlice) + lane);
        w[2 * h] = v.x;
        w[2 * h + 1] = v.y;
    }
    w[8] = __builtin_nontemporal_load(slot + 8 * kSlice + lane);
}
template <>
__device__ __forceinline__ void load_block<_Float16>(const _Float16 *slot, int lane, _Float16 (&w)[9]) {
    typedef _Float16 vh8 __attribute__((ext_vector_type(8)));
    const vh8 v = __builtin_nontemporal_load((const vh8 *)slot + lane);
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = v[k];
    w[8] = __builtin_nontemporal_load(slot + 8 * kSlice + lane);
}
template <>
__device__ __forceinline__ void load_block<float>(const float *slot, int lane, float (&w)[9]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const vf4 v = __builtin_nontemporal_load((const vf4 *)(slot + h * 4 * kSlice) + lane);
        w[4 * h] = v.x;
        w[4 * h + 1] = v.y;
        w[4 * h + 2] = v.z;
        w[4 * h + 3] = v.w;
    }
    w[8] = __builtin_nontemporal_load(slot + 8 * kSlice + lane);
}

// acc += (one row of V) z, the row being lane `lane` of a slice whose W slots (W
// wave-uniform: padding slots hold zero blocks) start at colp / vp; blocks are summed in
// slot order (SparseMatrixFixedPattern.cpp:148-187 on the block matrix).
// Loads are issued in groups of G blocks, each group's value loads and z gathers before
// any of them is used (a scheduling barrier keeps them together), so G blocks are in
// flight at once; one branch per slot instead made every block wait for the one before.
// Rows of up to U slots (U > 0) load all their block columns first. Slots past W in a row's last
// group repeat slot W-1 with a zero z, which adds +-0 and leaves the sums unchanged.
// The matrix streams through once per product: non-temporal loads keep it from evicting
// the gathered z and the row vectors that the next kernel reads.
// blocks whose loads one row issues together (slice_row_times), FP64 and FP32 values
#ifndef FFEA_SPMV_GROUP_F64
#define FFEA_SPMV_GROUP_F64 1
#endif
#ifndef FFEA_SPMV_GROUP_F32
#define FFEA_SPMV_GROUP_F32 4
#endif
// The gathered vector: FP64 tiled by 64 rows (vi), or the FP32-mixed form's packed FP32
// copy [rows][3], whose three values a lane reads in one 12-byte load
__device__ __forceinline__ void gather3(const double *z, int c, double (&o)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = z[vi(c, k)];
}
__device__ __forceinline__ void gather3(const float *z, int c, double (&o)[3]) {
    const float *g = z + 3 * (size_t)c;
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k] = (double)g[k];
}

template <typename T, int U, int G = sizeof(T) == 8 ? FFEA_SPMV_GROUP_F64 : FFEA_SPMV_GROUP_F32, typename V = double>
__device__ __forceinline__ void slice_row_times(const int *colp, const T *vp, int lane, int W, const V *z,
                                                double (&acc)[3]) {
    auto group = [&](int g, const int (&cc)[G]) {
        T w[G][9];
        double zz[G][3];
#pragma unroll
        for (int h = 0; h < G; ++h) {
            const int j = g + h < W ? g + h : W - 1;
            load_block<T>(vp + (size_t)j * 9 * kSlice, lane, w[h]);
            gather3(z, cc[h], zz[h]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int h = 0; h < G; ++h) {
            const bool live = g + h < W;
            const double z0 = live ? zz[h][0] : 0.0, z1 = live ? zz[h][1] : 0.0, z2 = live ? zz[h][2] : 0.0;
            acc[0] += (double)w[h][0] * z0; acc[0] += (double)w[h][1] * z1; acc[0] += (double)w[h][2] * z2;
            acc[1] += (double)w[h][3] * z0; acc[1] += (double)w[h][4] * z1; acc[1] += (double)w[h][5] * z2;
            acc[2] += (double)w[h][6] * z0; acc[2] += (double)w[h][7] * z1; acc[2] += (double)w[h][8] * z2;
        }
    };
    if (U > 0 && W <= U) {
        int ca[U > 0 ? U : 1];
#pragma unroll
        for (int j = 0; j < U; ++j) ca[j] = __builtin_nontemporal_load(colp + (size_t)(j < W ? j : W - 1) * kSlice);
#pragma unroll
        for (int g = 0; g < U; g += G) {
            if (g < W) {
                int cc[G];
#pragma unroll
                for (int h = 0; h < G; ++h) cc[h] = ca[g + h];
                group(g, cc);
            }
        }
    } else {
        for (int g = 0; g < W; g += G) {
            int cc[G];
#pragma unroll
            for (int h = 0; h < G; ++h)
                cc[h] = __builtin_nontemporal_load(colp + (size_t)(g + h < W ? g + h : W - 1) * kSlice);
            group(g, cc);
        }
    }
}

// U: rows with up to U blocks load all their block columns first (slice_row_times)
template <typename T, bool PROJ, int BS = row_block<T>, int U = kMaxUnrolledRow>
__device__ __forceinline__ void cg_spmv(const SpmvArgs<T> &a, const ChunkRange cr) {
    const int i = a.iter;
    const bool first = (i == 0);
    int cur = -1;
    bool skip = false;
    double beta = 0.0;
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = a.ra.chunks[c];
        if (ch.blob != cur) {
            cur = ch.blob;
            CgScalars *S = a.ra.S + cur;
            const int2 sg = a.ra.seg[cur];
            const bool lead = (c == ch.first) && threadIdx.x == 0;
            skip = S->done != 0;
            if (!skip && first) {
                if (S->inverted_step > 0) {  // an inverted mesh is not solved
                    if (lead) {
                        S->done = 1;
                        __hip_atomic_fetch_add(a.ra.num_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    skip = true;
                } else if constexpr (PROJ) {
                    double t[1];
                    sum_partials<1, BS>(a.Pin, a.ra.nchunks, sg.x, sg.y, t);
                    if (lead) S->delta[1] = t[0];
                } else {
                    double t[2];
                    sum_partials<2, BS>(a.Pin, a.ra.nchunks, sg.x, sg.y, t);
                    if (a.zero_rhs_solved && t[1] == 0.0) {
                        if (lead) {
                            S->converged = 1;
                            S->iter = 0;
                            S->done = 1;
                            __hip_atomic_fetch_add(a.ra.num_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        skip = true;
                    } else if (lead) {
                        S->delta[1] = t[0];
                        S->f2 = t[1];
                    }
                }
            } else if (!skip) {
                double t[2];
                sum_partials<2, BS>(a.Pin, a.ra.nchunks, sg.x, sg.y, t);  // r.r, r.z of iteration i-1
                const double f2 = S->f2;
                const double res = (f2 == 0.0) ? 0.0 : t[0] / f2;
                if (res < a.epsilon2 || i >= a.max_iter) {
                    if (lead) {
                        S->converged = res < a.epsilon2 ? 1 : 0;
                        S->iter = i;
                        S->done = 1;
                        __hip_atomic_fetch_add(a.ra.num_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    skip = true;
                } else {
                    beta = t[1] / S->delta[(i - 2) & 1];
                    if (lead) S->delta[(i - 1) & 1] = t[1];
                }
            }
        }
        if (skip) continue;
        const double3 cen = a.centre[cur];
        constexpr int NV = PROJ ? 13 : 1;
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = 0.0;
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += BS) {
            const int sl = n / kSlice, lane = n % kSlice;
            // wave-uniform: a slice is one wavefront's 64 rows
            const int W = __builtin_amdgcn_readfirstlane(a.slice_w[sl]);
            const long long base = a.slice_base[sl];
            double acc[3] = {0.0, 0.0, 0.0};
            if constexpr (std::is_same<T, float>::value)
                slice_row_times<T, U>(a.col + base + lane, a.val + (size_t)base * 9, lane, W, a.z32, acc);
            else
                slice_row_times<T, U>(a.col + base + lane, a.val + (size_t)base * 9, lane, W, a.z, acc);
            double pn[3], qn[3] = {acc[0], acc[1], acc[2]};
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
                // the rigid-mode sums need FP64: their FP32 rounding is as large as the
                // null-space error the projection removes
                const double x[3] = {a.pos[3 * n] - cen.x, a.pos[3 * n + 1] - cen.y, a.pos[3 * n + 2] - cen.z};
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
        store_partials<NV, BS>(v, a.P1, a.ra.nchunks, c);
    }
}

// BS: threads per workgroup; the FP64 form takes kRowBlock, or kMixedBlock when no chunk
// has more than 256 rows (either way each thread owns at most one row of a chunk, so the
// sums are the same)
template <typename T, bool PROJ, int BS = row_block<T>, int U = kMaxUnrolledRow>
__global__ __launch_bounds__(BS) void k_cg_spmv(SpmvArgs<T> a) {
    cg_spmv<T, PROJ, BS, U>(a, my_chunks(a.ra.nchunks, a.ra.cpb));
}

// the projected FP32 step is held to 128 VGPRs (4 waves per SIMD): one 1024-row chunk
// per workgroup then fits the 6M-tet blob's 1007 chunks on the 256 CUs in one round
#ifndef FFEA_SPMV_PROJ_WAVES
#define FFEA_SPMV_PROJ_WAVES 4
#endif
template <>
__global__ __launch_bounds__(kMixedBlock) __attribute__((amdgpu_waves_per_eu(FFEA_SPMV_PROJ_WAVES))) void
k_cg_spmv<float, true, kMixedBlock, kMaxUnrolledRow>(SpmvArgs<float> a) {
    cg_spmv<float, true>(a, my_chunks(a.ra.nchunks, a.ra.cpb));
}
template <>
__global__ __launch_bounds__(kMixedBlock) __attribute__((amdgpu_waves_per_eu(FFEA_SPMV_PROJ_WAVES))) void
k_cg_spmv<float, true, kMixedBlock, kMaxUnrolledRow / 2>(SpmvArgs<float> a) {
    cg_spmv<float, true, kMixedBlock, kMaxUnrolledRow / 2>(a, my_chunks(a.ra.nchunks, a.ra.cpb));
}

// out = V in over all rows (tiled vectors), one lane per row, blocks in slot order:
// SparseMatrixFixedPattern::apply (SparseMatrixFixedPattern.cpp:148-187) on the block
// matrix; padding slots hold zero blocks
template <typename T>
__global__ __launch_bounds__(kBlock) void k_apply_matrix(int nrows, const int *slice_w, const long long *slice_base,
                                                         const int *col, const T *val, const double *in,
                                                         double *out) {
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= nrows) return;
    const int sl = n / kSlice, lane = n % kSlice;
    const int W = slice_w[sl];
    const long long base = slice_base[sl];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < W; ++j) {
        const int cc = col[base + (long long)j * kSlice + lane];
        T w[9];
        load_block<T>(val + (size_t)base * 9 + (size_t)j * 9 * kSlice, lane, w);
        const double v0 = in[vi(cc, 0)], v1 = in[vi(cc, 1)], v2 = in[vi(cc, 2)];
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[r] += (double)w[3 * r] * v0 + (double)w[3 * r + 1] * v1 + (double)w[3 * r + 2] * v2;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) out[vi(n, r)] = acc[r];
}

struct VecArgs {
    RowArgs ra;
    int iter;
    const double *p, *q, *dinv;
    const double *pos;
    const double3 *centre;
    const double *G;   // [nblob][36]
    double *x, *r, *z;
    const double *P1;  // chunk partials of the SpMV step
    double *P2;        // r.r, r.z partials of this step
    float *z32 = nullptr;  // FP32-mixed form: z's packed FP32 copy
    const float *binv = nullptr;  // multigrid block smoother: z = B^-1 r (jacobi_times)
};

// alpha = delta / p.q, x += alpha p, r -= alpha q (mat_vec_fns.cpp:187-197),
// z = D^-1 r (NoMassCGSolver.cpp:268-281), partials of r.r and r.z.
// With PROJ, q is first projected onto the range of the operator:
// q <- q - Z c, c = (Z^T Z)^-1 Z^T q, and p.q is corrected to match.
// the projected step's coefficients from the SpMV step's totals t (p.q, Z^T q, Z^T p):
// cq = (Z^T Z)^-1 Z^T q (Gi = (Z^T Z)^-1) and alpha = r.z / (p.q - (Z^T p).cq), the
// projected q being q - Z cq
__device__ __forceinline__ void proj_coef(const double (&t)[13], const double *Gi, double rz, double &alpha,
                                          double (&cq)[6]) {
    double pq = t[0];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) acc += Gi[6 * k + j] * t[1 + j];
        cq[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) pq -= t[7 + k] * cq[k];
    alpha = rz / pq;
}

template <bool PROJ, int BS>
__device__ __forceinline__ void cg_vec(const VecArgs &a, const ChunkRange cr) {
    int cur = -1;
    bool skip = false;
    double alpha = 0.0;
    double cq[6] = {0, 0, 0, 0, 0, 0};
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = a.ra.chunks[c];
        if (ch.blob != cur) {
            cur = ch.blob;
            const CgScalars *S = a.ra.S + cur;
            skip = S->done != 0;
            if (!skip) {
                if constexpr (PROJ) {
                    // every workgroup forms the blob's coefficients from the SpMV step's
                    // totals itself (identical in every block): no kernel between the steps
                    const int2 sg = a.ra.seg[cur];
                    double t[13];
                    sum_partials<13, BS>(a.P1, a.ra.nchunks, sg.x, sg.y, t);
                    proj_coef(t, a.G + 36 * cur, S->delta[(a.iter - 1) & 1], alpha, cq);
                } else {
                    const int2 sg = a.ra.seg[cur];
                    double t[1];
                    sum_partials<1, BS>(a.P1, a.ra.nchunks, sg.x, sg.y, t);
                    alpha = S->delta[(a.iter - 1) & 1] / t[0];
                }
            }
        }
        if (skip) continue;
        const double3 cen = a.centre[cur];
        double v[2] = {0.0, 0.0};
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += BS) {
            double qn[3] = {a.q[vi(n, 0)], a.q[vi(n, 1)], a.q[vi(n, 2)]};
            if constexpr (PROJ) {
                const double x[3] = {a.pos[3 * n] - cen.x, a.pos[3 * n + 1] - cen.y, a.pos[3 * n + 2] - cen.z};
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
            }
            jacobi_times(a.dinv, a.binv, n, r3, z3);
#pragma unroll
            for (int k = 0; k < 3; ++k) a.z[vi(n, k)] = z3[k];
            if (a.z32) store_f32x3(a.z32, n, z3);
            v[0] += r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2];
            v[1] += r3[0] * z3[0] + r3[1] * z3[1] + r3[2] * z3[2];
        }
        store_partials<2, BS>(v, a.P2, a.ra.nchunks, c);
    }
}
template <bool PROJ, int BS>
__global__ __launch_bounds__(BS) void k_cg_vec(VecArgs a) {
    cg_vec<PROJ, BS>(a, my_chunks(a.ra.nchunks, a.ra.cpb));
}

// Small blobs: the whole solve in one launch, one workgroup per blob running the SpMV
// step and the vector step of each iteration back to back over the blob's chunks. Partials, scalars and every sum are those of
// the separate kernels, so the results are identical; a barrier with a workgroup fence
// between the steps stands in for the kernel boundary.
template <typename T, bool PROJ>
__global__ __launch_bounds__(row_block<T>) void k_cg_fused(SpmvArgs<T> sa, VecArgs va, const double *Pstart) {
    const int b = blockIdx.x;
    const int2 sg = sa.ra.seg[b];
    const ChunkRange cr{sg.x, sg.y};
    const CgScalars *S = sa.ra.S + b;
    double *const P2 = va.P2;
    for (int it = 0; it <= sa.max_iter; ++it) {
        sa.iter = it;
        sa.Pin = it == 0 ? Pstart : P2;
        cg_spmv<T, PROJ>(sa, cr);
        __syncthreads();
        if (S->done || it == sa.max_iter) break;  // the step of max_iter only tests max_iter - 1
        va.iter = it;
        cg_vec<PROJ, row_block<T>>(va, cr);
        __syncthreads();
    }
}

// Blob::euler_integrate (Blob.cpp:3249-3263), mass-less branch: vel = x, pos += x dt,
// for every blob whose solve converged and is not yet applied
__global__ __launch_bounds__(kRowBlock) void k_update(RowArgs ra, const double *x, double *vel, double *pos,
                                                   double dt) {
    const ChunkRange cr = my_chunks(ra.nchunks, ra.cpb);
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = ra.chunks[c];
        const CgScalars &S = ra.S[ch.blob];
        if (!S.converged || S.updated) continue;
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double v = x[vi(n, k)];
                vel[3 * n + k] = v;
                pos[3 * n + k] += v * dt;
            }
    }
}

// MassLumpedSolver::solve and the inertial Blob::euler_integrate (MassLumpedSolver.cpp:66-73,
// Blob.cpp:3266-3280): a = f / m, v += a dt, x += v dt, for blobs without inverted elements
__global__ __launch_bounds__(kRowBlock) void k_mass_update(RowArgs ra, const double *rhs, const double *inv_m,
                                                        double *vel, double *pos, double dt) {
#pragma clang fp contract(off)
    const ChunkRange cr = my_chunks(ra.nchunks, ra.cpb);
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = ra.chunks[c];
        if (ra.S[ch.blob].inverted_step > 0) continue;
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double acc = rhs[vi(n, k)] * inv_m[n];
                const double v = vel[3 * n + k] + acc * dt;
                vel[3 * n + k] = v;
                pos[3 * n + k] += v * dt;
            }
    }
}

// the inertial Blob::euler_integrate after ConjugateGradientSolver::solve (Blob.cpp:3266-3280):
// v += a dt, x += v dt with a the CG solution, for blobs whose solve converged
__global__ __launch_bounds__(kRowBlock) void k_mass_update_cg(RowArgs ra, const double *acc, double *vel, double *pos,
                                                           double dt) {
#pragma clang fp contract(off)
    const ChunkRange cr = my_chunks(ra.nchunks, ra.cpb);
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = ra.chunks[c];
        const CgScalars &S = ra.S[ch.blob];
        if (!S.converged || S.updated) continue;
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double v = vel[3 * n + k] + acc[vi(n, k)] * dt;
                vel[3 * n + k] = v;
                pos[3 * n + k] += v * dt;
            }
    }
}

// MassLumpedSolver::init (MassLumpedSolver.cpp:27-63): 1 / (sum over the row's elements, in
// element order, of .25 rho vol_0); pinned rows 1
template <typename T>
__global__ __launch_bounds__(kRowBlock) void k_lumped_mass(RowArgs ra, const uint32_t *inc_ptr, const uint32_t *inc,
                                                        const double *rho, const T *vol0, const unsigned char *pinned,
                                                        double *inv_m) {
#pragma clang fp contract(off)
    const ChunkRange cr = my_chunks(ra.nchunks, ra.cpb);
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = ra.chunks[c];
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock) {
            double m = 0.0;
            for (uint32_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
                const uint32_t e = inc[k] >> 2;
                m += .25 * rho[e] * (double)vol0[e];
            }
            if (pinned[n]) m = 1.0;
            inv_m[n] = 1.0 / m;
        }
    }
}

// after k_update: mark applied solves and clear the inversion counters of the pass
__global__ void k_mark_updated(int nblob, CgScalars *S) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblob) return;
    if (S[b].converged) S[b].updated = 1;
    S[b].inverted = 0;
}

// ---------------------------------------------------------------- FP64 refinement of a solve
// The FP32-mixed form solves with an FP32 viscosity matrix, whose rounding (~1e-7 of each
// entry) the solve amplifies in the smooth, low-eigenvalue part of a solution. Refinement
// forms the true residual b - V x with an FP64 matrix-free product (shape derivatives and
// volumes recomputed in FP64 from the positions, FP64 viscosities) and solves the FP32
// system again for the correction, until that residual meets its tolerance.

// Each element's product with v (tiled node vector): (V_e v)_a = vol sigma(v) grad psi_a with
// sigma = A (L + L^T) + B tr(L) I and L = sum_b v_b (x) grad psi_b, the action of the 12x12
// matrix of tetra_element_linear.cpp:298-327; written where the row gather reads it.
struct ViscApplyArgs {
    int ne;
    const int *conn;           // [4][ne]
    const double *pos;         // [rows][3]
    const double *A, *B;       // [ne] FP64
    const double *v;           // tiled [rows][3]
    const uint32_t *inc_pos;   // [4][ne]
    double *w;                 // [4 ne][3]
};
__global__ __launch_bounds__(kBlock) void k_visc_apply64(ViscApplyArgs a) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.ne) return;
    const int ne = a.ne;
    int n[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) n[k] = a.conn[k * ne + e];
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) J[r][c] = a.pos[3 * n[r + 1] + c] - a.pos[3 * n[0] + c];
    double d[12];
    d[1] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
    d[2] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    d[3] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
    d[5] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
    d[6] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    d[7] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
    d[9] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
    d[10] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    d[11] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
    const double det = J[0][0] * d[1] + J[1][0] * d[2] + J[2][0] * d[3];
    const double rdet = 1.0 / det, vol = fabs(det) / 6.0;
#pragma unroll
    for (int k = 1; k < 12; ++k)
        if (k & 3) d[k] *= rdet;
    d[0] = -(d[1] + d[2] + d[3]);
    d[4] = -(d[5] + d[6] + d[7]);
    d[8] = -(d[9] + d[10] + d[11]);
    double L[3][3];  // L[i][j] = sum_b v_b[i] d_j psi_b, d_j psi_b = d[4 j + b]
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) L[i][j] = 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        double vb[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) vb[i] = a.v[vi(n[b], i)];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) L[i][j] += vb[i] * d[4 * j + b];
    }
    const double A = a.A[e], B = a.B[e], tr = L[0][0] + L[1][1] + L[2][2];
    double sg[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) sg[i][j] = A * (L[i][j] + L[j][i]) + (i == j ? B * tr : 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double *wo = a.w + 3 * (size_t)a.inc_pos[(size_t)q * ne + e];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            wo[i] = vol * (sg[i][0] * d[q] + sg[i][1] * d[4 + q] + sg[i][2] * d[8 + q]);
    }
}

// res = b - V x per corner row (V x from the element products in element order, the drag
// on unpinned rows; pinned rows are identity rows), with chunk partials |res|^2 and |b|^2
struct ResidArgs {
    RowArgs ra;
    const uint32_t *inc_ptr;
    const double *w;             // [4 ne][3] from k_visc_apply64
    const double *drag;          // [rows] or null
    const unsigned char *pinned;
    const double *b, *x;         // tiled
    double *res;                 // tiled
    double *P;                   // [2][nchunks]
};
__global__ __launch_bounds__(kRowBlock) void k_resid64(ResidArgs a) {
    const ChunkRange cr = my_chunks(a.ra.nchunks, a.ra.cpb);
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = a.ra.chunks[c];
        double v[2] = {0.0, 0.0};
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock) {
            double kx[3] = {0.0, 0.0, 0.0};
            if (a.pinned[n]) {
#pragma unroll
                for (int i = 0; i < 3; ++i) kx[i] = a.x[vi(n, i)];
            } else {
                for (uint32_t k = a.inc_ptr[n]; k < a.inc_ptr[n + 1]; ++k)
#pragma unroll
                    for (int i = 0; i < 3; ++i) kx[i] += a.w[3 * (size_t)k + i];
                if (a.drag)
#pragma unroll
                    for (int i = 0; i < 3; ++i) kx[i] += a.drag[n] * a.x[vi(n, i)];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double bi = a.b[vi(n, i)], ri = bi - kx[i];
                a.res[vi(n, i)] = ri;
                v[0] += ri * ri;
                v[1] += bi * bi;
            }
        }
        store_partials<2, kRowBlock>(v, a.P, a.ra.nchunks, c);
    }
}

// y += x over a whole tiled vector
__global__ __launch_bounds__(kBlock) void k_add_vec(long long n, const double *x, double *y) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// ---------------------------------------------------------------- adjoint (reverse of a step)
// One mass-less step is x' = x + dt v with K(x) v = P f(x) (P zeroes pinned rows). For
// lambda' = dL/dx' the reverse is: solve K mu = P lambda'; then
//   lambda = lambda' + dt [ d(mu . f)/dx - d(mu^T K(x) v)/dx ]
//   dL/dtheta += dt [ d(mu . f)/dtheta - d(mu^T K v)/dtheta ]   (theta = G, E, A, B per element)
// K comes from tetra_element_linear.cpp:298-327, which for an element is
//   (K_e v)_a = vol sigma(v) grad psi_a,  sigma(v) = A (L + L^T) + B tr(L) I,  L = sum_b v_b (x) grad psi_b,
// so mu^T K_e v = vol sigma(v) : L_mu. The elastic force is f_a = -vol s grad psi_a with s the
// shear + bulk stress of tetra_element_linear.cpp:375-403. Both scalars are functions of the
// element Jacobian J only; their derivatives are taken in forward mode over the 9 entries of J.

template <typename T>
struct Dual {
    T v, d;  // value and derivative along one direction
};
template <typename T>
__device__ __forceinline__ Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator*(T b, Dual<T> a) { return {a.v * b, a.d * b}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator/(Dual<T> a, Dual<T> b) {
    const T r = (T)1 / b.v;
    return {a.v * r, (a.d - a.v * r * b.d) * r};
}
template <typename T>
__device__ __forceinline__ Dual<T> operator/(T a, Dual<T> b) {
    const T r = (T)1 / b.v;
    return {a * r, -a * r * r * b.d};
}
template <typename T>
__device__ __forceinline__ Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator-(T b, Dual<T> a) { return {b - a.v, -a.d}; }
template <typename T>
__device__ __forceinline__ Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <typename T>
__device__ __forceinline__ T value_of(Dual<T> a) { return a.v; }
template <typename T>
__device__ __forceinline__ T value_of(T a) { return a; }

// mu . f_el(J) - mu^T K(J) v for one element (noise off, element not inverted) and, for
// the value pass, its partial derivatives in G, E, A, B
template <typename S, typename T>
__device__ __forceinline__ S adjoint_scalar(const S (&J)[3][3], const S (&DJ)[3][3], const T (&Ji0)[3][3], T v0, T G,
                                            T E, T A, T B, const T (&mu)[4][3], const T (&vel)[4][3], T *part) {
    S d[12];
    d[1] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
    d[2] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    d[3] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
    d[5] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
    d[6] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    d[7] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
    d[9] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
    d[10] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    d[11] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
    const S det = J[0][0] * d[1] + J[1][0] * d[2] + J[2][0] * d[3];
    const S vol = (value_of(det) < (T)0 ? -det : det) * (T)(1.0 / 6.0);
    const S rdet = (T)1 / det;
#pragma unroll
    for (int k = 1; k < 12; ++k)
        if (k & 3) d[k] = d[k] * rdet;
    d[0] = -(d[1] + d[2] + d[3]);
    d[4] = -(d[5] + d[6] + d[7]);
    d[8] = -(d[9] + d[10] + d[11]);
    // H = F - I from the edge displacements DJ = J - J_0; F F^T - I and e = det F - 1 (strain_terms)
    S H[3][3], FF[3][3], e;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) H[i][j] = DJ[0][i] * Ji0[j][0] + DJ[1][i] * Ji0[j][1] + DJ[2][i] * Ji0[j][2];
    strain_terms<S, T>(H, FF, &e);
    // velocity gradients L[i][j] = sum_q w_q[i] grad psi_q[j], grad psi_q[j] = d[4j+q]
    S Lm[3][3], Lv[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Lm[i][j] = mu[0][i] * d[4 * j] + mu[1][i] * d[4 * j + 1] + mu[2][i] * d[4 * j + 2] + mu[3][i] * d[4 * j + 3];
            Lv[i][j] = vel[0][i] * d[4 * j] + vel[1][i] * d[4 * j + 1] + vel[2][i] * d[4 * j + 2] +
                       vel[3][i] * d[4 * j + 3];
        }
    const S trm = Lm[0][0] + Lm[1][1] + Lm[2][2];
    const S trv = Lv[0][0] + Lv[1][1] + Lv[2][2];
    // stress = G sG + E sE with j = 1 + e = vol/v0:
    //   sG = (F F^T - I)/j - e (2 + e)/(3 j) I,   sE = e (2 + e)/(2 j) I
    S ffm = FF[0][0] * Lm[0][0];
    ffm = ffm + FF[0][1] * Lm[0][1] + FF[0][2] * Lm[0][2] + FF[1][0] * Lm[1][0] + FF[1][1] * Lm[1][1] +
          FF[1][2] * Lm[1][2] + FF[2][0] * Lm[2][0] + FF[2][1] * Lm[2][1] + FF[2][2] * Lm[2][2];
    const S jr = (T)1 / (e + (T)1);
    const S vb = e * (e + (T)2) * jr;                            // vol/v0 - v0/vol
    const S sG_m = jr * ffm - vb * trm * (T)(1.0 / 3.0);         // sG : L_mu
    const S sE_m = vb * trm * (T)0.5;                            // sE : L_mu
    S sym = Lv[0][0] * Lm[0][0] * (T)2;
    sym = sym + (Lv[0][1] + Lv[1][0]) * (Lm[0][1] + Lm[1][0]) + (Lv[0][2] + Lv[2][0]) * (Lm[0][2] + Lm[2][0]) +
          (Lv[1][2] + Lv[2][1]) * (Lm[1][2] + Lm[2][1]) + Lv[1][1] * Lm[1][1] * (T)2 + Lv[2][2] * Lm[2][2] * (T)2;
    const S tt = trv * trm;
    const S phi = -(vol * (G * sG_m + E * sE_m));   // mu . f_el
    const S psi = vol * (A * sym + B * tt);         // mu^T K v
    if (part) {
        part[0] = -value_of(vol * sym);   // d/dA of (phi - psi)
        part[1] = -value_of(vol * tt);    // d/dB
        part[2] = -value_of(vol * sG_m);  // d/dG
        part[3] = -value_of(vol * sE_m);  // d/dE
    }
    return phi - psi;
}

template <typename T>
struct AdjArgs {
    int ne;
    const int *conn;      // [4][ne]
    const double *pos;    // [rows][3] positions at the step
    const double *rest;   // [rows][3] rest positions
    const double *mu;     // tiled: the adjoint solve
    const double *vel;    // [rows][3]: the step's velocity solution
    const T *jinv0, *vol0, *G, *E, *A, *B;
    double dt;
    T *w;                 // [ne][12]: d(phi - psi)/dx per corner
    double *grad;         // [4][ne]: dL/dA, dL/dB, dL/dG, dL/dE accumulated
};

// reverse of the element pass and of the viscosity operator for one step, one element
// per lane; every element writes only its own outputs (no atomics)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_element_adjoint(AdjArgs<T> a) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.ne) return;
    const int ne = a.ne;
    int n[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) n[k] = a.conn[k * ne + e];
    double x0[3], u0[3];
    T Jv[3][3], DJv[3][3], mu[4][3], vel[4][3], Ji0[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        x0[c] = a.pos[3 * n[0] + c];
        u0[c] = x0[c] - a.rest[3 * n[0] + c];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Jv[r][c] = (T)(a.pos[3 * n[r + 1] + c] - x0[c]);
            DJv[r][c] = (T)((a.pos[3 * n[r + 1] + c] - a.rest[3 * n[r + 1] + c]) - u0[c]);
        }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            mu[q][c] = (T)a.mu[vi(n[q], c)];
            vel[q][c] = (T)a.vel[3 * n[q] + c];
        }
#pragma unroll
    for (int k = 0; k < 9; ++k) Ji0[k / 3][k % 3] = a.jinv0[(size_t)k * ne + e];
    const T v0 = a.vol0[e], G = a.G[e], E = a.E[e], A = a.A[e], B = a.B[e];
    T part[4];
    (void)adjoint_scalar<T, T>(Jv, DJv, Ji0, v0, G, E, A, B, mu, vel, part);
    T dJ[3][3];
    for (int k = 0; k < 9; ++k) {
        Dual<T> J[3][3], DJ[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T t = (T)((3 * r + c) == k ? 1 : 0);
                J[r][c] = Dual<T>{Jv[r][c], t};
                DJ[r][c] = Dual<T>{DJv[r][c], t};
            }
        dJ[k / 3][k % 3] = adjoint_scalar<Dual<T>, T>(J, DJ, Ji0, v0, G, E, A, B, mu, vel, nullptr).d;
    }
    // J rows are x_{r+1} - x_0
    T *wo = a.w + (size_t)12 * e;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        wo[c] = -(dJ[0][c] + dJ[1][c] + dJ[2][c]);
        wo[3 + c] = dJ[0][c];
        wo[6 + c] = dJ[1][c];
        wo[9 + c] = dJ[2][c];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) a.grad[(size_t)k * ne + e] += a.dt * (double)part[k];
}

// adjoint right-hand side and CG start: rhs = P lambda (tiled), r = rhs, x = 0, z = D^-1 r,
// chunk partials r.z and |rhs|^2 as k_gather leaves them
struct AdjStartArgs {
    RowArgs ra;
    const double *lam;            // tiled
    const unsigned char *pinned;
    const double *dinv;
    double *rhs, *r, *z, *x, *P;
    float *z32 = nullptr;         // FP32-mixed form: z's packed FP32 copy
    const float *binv = nullptr;  // multigrid block smoother: z = B^-1 r (jacobi_times)
};
__global__ __launch_bounds__(kRowBlock) void k_adjoint_start(AdjStartArgs a) {
    const ChunkRange cr = my_chunks(a.ra.nchunks, a.ra.cpb);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.ra.num_done = 0;
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = a.ra.chunks[c];
        double v[2] = {0.0, 0.0};
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock) {
            const bool pin = a.pinned[n] != 0;
            double f3[3], z3[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) f3[k] = pin ? 0.0 : a.lam[vi(n, k)];
            jacobi_times(a.dinv, a.binv, n, f3, z3);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double f = f3[k], zz = z3[k];
                a.rhs[vi(n, k)] = f;
                a.r[vi(n, k)] = f;
                a.x[vi(n, k)] = 0.0;
                a.z[vi(n, k)] = zz;
                v[0] += f * zz;
                v[1] += f * f;
            }
            if (a.z32) store_f32x3(a.z32, n, z3);
        }
        store_partials<2, kRowBlock>(v, a.P, a.ra.nchunks, c);
        if (c == ch.first && threadIdx.x == 0) {
            CgScalars *S = a.ra.S + ch.blob;
            S->inverted_step = 0;
            S->done = 0;
            S->converged = 0;
            S->iter = 0;
            S->updated = 1;  // an adjoint solve never moves the blob
        }
    }
}

// lambda += dt * sum over the node's element corners (element order) of w
template <typename T>
__global__ __launch_bounds__(kRowBlock) void k_adjoint_gather(RowArgs ra, const uint32_t *inc_ptr, const uint32_t *inc,
                                                           const T *w, double dt, double *lam) {
    const ChunkRange cr = my_chunks(ra.nchunks, ra.cpb);
    for (int c = cr.c0; c < cr.c1; ++c) {
        const Chunk ch = ra.chunks[c];
        for (int n = ch.begin + threadIdx.x; n < ch.end; n += kRowBlock) {
            double s[3] = {0.0, 0.0, 0.0};
            for (uint32_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
                const uint32_t es = inc[k];
                const T *src = w + (size_t)12 * (es >> 2) + 3 * (es & 3);
                s[0] += (double)src[0];
                s[1] += (double)src[1];
                s[2] += (double)src[2];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) lam[vi(n, k)] += dt * s[k];
        }
    }
}

// ---------------------------------------------------------------- measurements
// Blob::calculate_strain_energy (Blob.cpp:1052-1064) per element, in FP64 from the current
// positions: W = vol0 [G (tr F^T F - 3) + C (det F^2 - 1)/2 - (2G + C) ln det F], C = E - 2G/3,
// with F = I + H formed from displacements (tr F^T F - 3 = 2 tr H + |H|^2,
// det F = 1 + e as in strain_terms) so that small strains keep their digits
template <typename T>
__global__ __launch_bounds__(kBlock) void k_strain_energy(int ne, const int *conn, const double *pos, const double *rest,
                                                          const T *jinv0, const T *vol0, const T *G, const T *E,
                                                          double *w) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    int n[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) n[k] = conn[k * ne + e];
    double u0[3], DJ[3][3], Ji[3][3], H[3][3], FFm[3][3], ee;
#pragma unroll
    for (int c = 0; c < 3; ++c) u0[c] = pos[3 * n[0] + c] - rest[3 * n[0] + c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) DJ[r][c] = (pos[3 * n[r + 1] + c] - rest[3 * n[r + 1] + c]) - u0[c];
#pragma unroll
    for (int k = 0; k < 9; ++k) Ji[k / 3][k % 3] = (double)jinv0[(size_t)k * ne + e];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) H[i][j] = DJ[0][i] * Ji[j][0] + DJ[1][i] * Ji[j][1] + DJ[2][i] * Ji[j][2];
    strain_terms<double, double>(H, FFm, &ee);
    double hh = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) hh += H[i][j] * H[i][j];
    const double g = (double)G[e], C = (double)E[e] - g * 2.0 / 3.0;
    const double trc = 2.0 * (H[0][0] + H[1][1] + H[2][2]) + hh;
    // vol / vol_0 = |det F| (the element volume is unsigned), so an inverted element takes log|det F|
    const double lj = ee > -1.0 ? log1p(ee) : log(-1.0 - ee);
    w[e] = (double)vol0[e] * (g * trc + 0.5 * C * ee * (2.0 + ee) - (2.0 * g + C) * lj);
}

// per blob (one workgroup each): half the sum of its elements' W
// Blob::make_measurements, one element's kinetic energy term (Blob.cpp:1079-1102):
// v_e . (M v_e) with the element mass matrix of apply_element_mass_matrix
// (tetra_element_linear.cpp:524-542: .1 on the diagonal, .05 off it, per component, times
// vol_0 rho)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_kinetic_energy(int ne, const int *conn, const double *vel,
                                                        const double *rho, const T *vol0, double *w) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= ne) return;
    double v[12], t[12];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int n = conn[a * ne + e];
#pragma unroll
        for (int c = 0; c < 3; ++c) v[4 * c + a] = vel[3 * n + c];
    }
#pragma unroll
    for (int k = 0; k < 12; k += 4)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j) acc += (i == j ? .1 : .05) * v[j + k];
            t[i + k] = acc;
        }
    const double s = (double)vol0[e] * rho[e];
    double ke = v[0] * (t[0] * s);
#pragma unroll
    for (int i = 1; i < 12; ++i) ke += v[i] * (t[i] * s);
    w[e] = ke;
}

__global__ __launch_bounds__(kRowBlock) void k_blob_energy(const int *elem_off, const int *elem_cnt, const double *w,
                                                        double *out) {
    const int b = blockIdx.x;
    double v[1] = {0.0};
    for (int e = elem_off[b] + threadIdx.x; e < elem_off[b] + elem_cnt[b]; e += kRowBlock) v[0] += w[e];
    block_sum<1, kRowBlock>(v);
    if (threadIdx.x == 0) out[b] = 0.5 * v[0];
}

// per-blob sums of per-element energies whose blob b holds elements [elem_off[b],
// elem_off[b+1]) (the last blob ends at ne): the allocation-free strain-energy call, which
// has no per-blob count table
__global__ __launch_bounds__(kRowBlock) void k_blob_energy_range(int nblob, int ne, const int *elem_off, const double *w,
                                                              double *out) {
    const int b = blockIdx.x;
    const int e0 = elem_off[b], e1 = b + 1 < nblob ? elem_off[b + 1] : ne;
    double v[1] = {0.0};
    for (int e = e0 + threadIdx.x; e < e1; e += kRowBlock) v[0] += w[e];
    block_sum<1, kRowBlock>(v);
    if (threadIdx.x == 0) out[b] = 0.5 * v[0];
}

// ---------------------------------------------------------------- copy probe
// y = x over n2 16-byte vectors, grid-stride: reads n2*16 and writes n2*16 bytes
// Streaming probes of the achievable HBM rate. A thread moves kProbeVecs 16-byte vectors
// a grid stride apart (each instruction of a wavefront covers 1 KiB of contiguous bytes),
// all loads issued before any use, non-temporal. k_copy: y = x (the calibration kernel of
// the hardware byte counters); k_read: every vector read once, one sum per workgroup.
constexpr int kProbeVecs = 4;
__global__ __launch_bounds__(kBlock) void k_copy(const vd2 *__restrict__ x, vd2 *__restrict__ y, size_t n2) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x; i0 < n2; i0 += kProbeVecs * stride) {
        vd2 v[kProbeVecs];
#pragma unroll
        for (int u = 0; u < kProbeVecs; ++u)
            if (i0 + u * stride < n2) v[u] = __builtin_nontemporal_load(x + i0 + u * stride);
#pragma unroll
        for (int u = 0; u < kProbeVecs; ++u)
            if (i0 + u * stride < n2) __builtin_nontemporal_store(v[u], y + i0 + u * stride);
    }
}
__global__ __launch_bounds__(kBlock) void k_read(const vd2 *__restrict__ x, size_t n2, double *__restrict__ out) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    double s[1] = {0.0};
    for (size_t i0 = (size_t)blockIdx.x * kBlock + threadIdx.x; i0 < n2; i0 += kProbeVecs * stride) {
        vd2 v[kProbeVecs];
#pragma unroll
        for (int u = 0; u < kProbeVecs; ++u)
            v[u] = i0 + u * stride < n2 ? __builtin_nontemporal_load(x + i0 + u * stride) : vd2{0.0, 0.0};
#pragma unroll
        for (int u = 0; u < kProbeVecs; ++u) s[0] += v[u].x + v[u].y;
    }
    block_sum<1, kBlock>(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

}  // namespace ffk
