// MI355X (gfx950) blob time step behind the C ABI of include/ffea_gpu.h: host side.
// The kernels are in kernels.h.
//
// A handle holds one blob (ffea_gpu_blob_create) or many independent blobs
// (ffea_gpu_batch_create) stepped together: each kernel launch covers every blob, and
// each blob keeps its own CG scalars, noise stream and convergence state.
//
// Data layout in HBM:
//   the corner ("linear") nodes of a blob are renumbered 0..Nl-1 in ascending .node
//   order; blobs are concatenated, each starting on a 64-row boundary;
//   positions and velocities are FP64 [rows][3]; CG vectors are FP64, tiled by 64 rows;
//   element data are structure-of-arrays over the elements of all blobs;
//   element nodal forces are [Ne][12]; shape records [Ne][16] (dpsi, vol, A, B, padding);
//   the viscosity matrix is 3x3 blocks in a sliced-ELL layout: 64 rows per slice, slice
//   width W = its longest row, values [slot_base*9 + (j*9 + k)*64 + lane] and block
//   columns [slot_base + j*64 + lane], so lane-per-row loads of one (j, k) are 512
//   contiguous bytes.
// The node-force scatter is a gather over a node->element CSR in element order and
// every reduction is summed in a fixed order: results are bit-reproducible and do not
// depend on the launch grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/ffea_gpu.h"
#include "amg_kernels.h"
#include "amg_setup.h"
#include "kernels.h"
#include "mrg32k3a.h"

using namespace ffk;

// The multigrid V-cycle's damped-Jacobi sweeps before and after the correction from the
// level below: one on levels 0 and 1, except two on a single blob's level 1 of more than
// kMgBigLevel1 aggregates, and kMgCoarserSweeps on levels 2 .. L-2
constexpr int kMgBigLevel1 = 10000;
constexpr int kMgCoarserSweeps = 4;
// a level 1 with two sweeps smooths with a degree-2 Chebyshev polynomial of D^-1 A over
// [l/kMgChebRatio, l] for l = 1.1 times its largest-eigenvalue estimate, at the cost of
// two damped-Jacobi sweeps (false: the sweeps)
constexpr bool kMgChebyshev1 = true;
constexpr double kMgChebRatio = 10.0;
// unknowns at which a single blob's coarsening stops: its coarsest level, up to this size,
// is solved with a dense inverse (an ensemble's blobs keep amg::Options' default)
constexpr int kMgMaxCoarseSingle = 4800;

// Set-up options of one blob's hierarchy in a handle of nblob blobs, the blob having nl
// corner nodes. Every blob coarsens at least once, so that it has a coarse level (a
// single blob of up to kMgMaxCoarseSingle unknowns would otherwise stop at level 0); an
// ensemble's blobs keep amg::Options' coarse size and never coarsen to fewer than 4
// aggregates (a one-aggregate level of a free blob is pure null space: with the FP32
// matrix its rounding is not positive).
static amg::Options mg_options(int nblob, int nl, int threads) {
    amg::Options opt;
    opt.threads = threads;
    if (nblob == 1) opt.max_coarse = kMgMaxCoarseSingle;
    opt.max_coarse = std::min(opt.max_coarse, std::max(6, 3 * nl / 8));
    if (nblob > 1) opt.min_aggregates = 4;
    return opt;
}

// Scratch of the allocation-free entry point (ffea_gpu_batch_strain_energy_noalloc): 64 MiB
// per GPU, held by every GPU the library runs on until the process ends. Host code never
// uses this name as an address: each device's copy is found once with hipGetSymbolAddress
// (scratch_on below) and that pointer is what kernels and copies receive.
constexpr size_t kScratchDoubles = ((size_t)64 << 20) / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

thread_local std::string g_last_error;

#define FFEA_HIP_TRY(expr)                                                                        \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
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

// the device address of `scratch` on `device` (which must be the current device), looked
// up on the first call for that device and kept
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxScratchDevices] = {};
double *scratch_on(int device) {
    if (device < 0 || device >= kMaxScratchDevices) throw Invalid("device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (double *)ptr;
    }
    return g_scratch_ptr[device];
}

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
        n = std::max<size_t>(1, count);
        FFEA_HIP_TRY(hipMalloc(&p, n * sizeof(T)));
        *total += n * sizeof(T);
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count) FFEA_HIP_TRY(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
    // not for code that may be captured into a graph
    void zero(hipStream_t s) { FFEA_HIP_TRY(hipMemsetAsync(p, 0, n * sizeof(T), s)); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Phase {
    std::vector<hipEvent_t> ev;  // pairs
    std::vector<int> iter;       // CG phases: iteration of each timed pair
    double ms = 0;
    long long launches = 0;
};

// one blob as the caller hands it over (host pointers)
struct BlobInput {
    int nt = 0, ne = 0;
    const double *pos = nullptr;
    const int *top = nullptr;
    const double *A = nullptr, *B = nullptr, *G = nullptr, *E = nullptr;
    int npin = 0;
    const int *pin = nullptr;
    int nctf = 0;
    const int *ctf_node = nullptr;
    const double *ctf_force = nullptr;
    const uint32_t *rng = nullptr;
    const double *drag = nullptr;  // [nt] Stokes drag per node (calc_stokes), else null
    const ffea_gpu_surface_forces *sf = nullptr;  // surface constant forces, or null
    const ffea_gpu_rotational_forces *rf = nullptr;  // rotational constant forces, or null
    int next = 0;
    const int *ext = nullptr;  // nodes that may receive per-step external forces
};

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

// host plan of one blob, in local indices
struct BlobPlan {
    int nt = 0, ne = 0, nl = 0, nslices = 0, npin = 0;
    bool all_pinned = false, connected = true, have_ctf = false, have_fold = false;
    std::vector<int> lin_of, orig_of_lin, mid_a, mid_b;
    std::vector<unsigned char> pin_lin;
    std::vector<double> p0;         // [nl][3] linearised rest positions
    std::vector<int> conn;          // [4][ne]
    std::vector<double> jinv0, vol0, det0;
    std::vector<uint32_t> ip, ic;   // node -> element incidence (e*4 + corner)
    std::vector<int> sw;            // slice widths
    std::vector<long long> sb;      // slice bases (nslices+1)
    std::vector<int> col;           // [nslots] block columns
    std::vector<uint32_t> cnt;      // [nslots+1] source prefix
    std::vector<uint32_t> contrib;  // (e<<4)|(a<<2)|b in element order, or kIdentity
    std::vector<unsigned char> dj;  // [nl] slot index of the diagonal block
    std::vector<double> ctf;        // [nl][3]
    std::vector<uint32_t> fp;       // [nl+1]
    std::vector<double> fv;         // linearise_force increments [..][3] (constant forces only)
    // Stokes noise: per (element, corner) in element order, the midside nodes first folded
    // there (local node indices, -1 if already folded), and per node its constant force
    std::vector<std::array<int, 3>> fm;
    std::vector<double> node_ctf;   // [nt][3]
    std::vector<double> drag_lin;   // [nl] drag of unpinned corner rows
    // surface constant forces: the faces some set uses, in face order ("used faces")
    bool has_sf = false;
    std::vector<int> sf_rows;       // [nfu][3][2] corner rows whose midpoint is each face node's position
    std::vector<double> sf_force;   // [ns][3]
    std::vector<int> sf_set_ptr;    // [ns+1] entries of each set
    std::vector<int> sf_entry_face; // [entries] used face of each entry (set order)
    std::vector<int> sf_entry_set;  // [entries]
    std::vector<int> sf_fe_ptr, sf_fe;  // per used face: its entries in set order
    std::vector<int> sf_nf_ptr, sf_nf;  // per node of the .node file: used faces holding it, face order
    // nodes whose force is formed each step (rotational constant forces, external forces),
    // ascending node index
    bool has_dn = false, has_rf = false;
    bool dn_mid = false;            // some are midside nodes: per-step folds
    std::vector<int> ext_nodes;     // [next] external force nodes, in the caller's order (first occurrence)
    std::vector<int> dn_node;       // [ndn] node
    std::vector<int> dn_ext;        // [ndn] its external force slot, or -1
    std::vector<int> dn_row;        // [ndn] its corner row, or -1 (midside)
    std::vector<int> dn_ends;       // [ndn][2] corner rows whose midpoint is its position
    std::vector<int> dn_lptr;       // [ndn+1] its linear constant forces, list order
    std::vector<double> dn_lval;    // [..][3]
    std::vector<int> dn_rptr;       // [ndn+1] its rotational forces, list order
    std::vector<double> rf_axis;    // [entries][6]
    std::vector<double> rf_F;       // [entries]
    std::vector<int> rf_torque;     // [entries] 1: 't'
    double centre[3] = {0, 0, 0};
    long long nblocks = 0;
    long long nslots() const { return sb.empty() ? 0 : sb.back(); }
};

BlobPlan make_plan(const BlobInput &in, int blob_index, bool stokes_noise, int threads) {
    BlobPlan P;
    const int nt = in.nt, ne = in.ne;
    const int *top = in.top;
    P.nt = nt;
    P.ne = ne;
    if ((long long)ne >= (1ll << 28) - 1) throw Invalid("mesh too large for 32-bit source lists");
    const std::string who = blob_index >= 0 ? fmt("blob %d: ", blob_index) : std::string();
    // ---- corner / midside classification (Blob::load_topology, Blob.cpp:2071-2077)
    std::vector<signed char> kind(nt, 0);  // 1 corner, 2 midside
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 10; ++s) {
            const int v = top[10 * e + s];
            if (v < 0 || v >= nt)
                throw Invalid(who + fmt("Error: Element %d references an out of bounds node index.", e));
            const signed char k = s < 4 ? 1 : 2;
            if (kind[v] && kind[v] != k)
                throw Invalid(who + fmt("node %d is a corner of one element and a midside node of another", v));
            kind[v] = k;
        }
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 4; ++s)
            for (int t = s + 1; t < 4; ++t)
                if (top[10 * e + s] == top[10 * e + t]) throw Invalid(who + fmt("element %d repeats a corner node", e));
    P.lin_of.assign(nt, -1);
    for (int v = 0; v < nt; ++v)
        if (kind[v] == 1) {
            P.lin_of[v] = (int)P.orig_of_lin.size();
            P.orig_of_lin.push_back(v);
        }
    const int nl = P.nl = (int)P.orig_of_lin.size();
    P.mid_a.assign(nt, -1);
    P.mid_b.assign(nt, -1);
    static const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 6; ++k) {
            const int m = top[10 * e + 4 + k];
            P.mid_a[m] = top[10 * e + ends[k][0]];
            P.mid_b[m] = top[10 * e + ends[k][1]];
        }
    // ---- pins
    std::vector<unsigned char> pin_node(nt, 0);
    P.pin_lin.assign(nl, 0);
    for (int i = 0; i < in.npin; ++i) {
        if (in.pin[i] < 0 || in.pin[i] >= nt)
            throw Invalid(who + fmt("Error: Pinned node %d references an out of bounds node index.", i));
        pin_node[in.pin[i]] = 1;
    }
    for (int n = 0; n < nl; ++n) P.pin_lin[n] = pin_node[P.orig_of_lin[n]];
    P.npin = in.npin;
    P.all_pinned = (in.npin == nt);  // Blob.cpp:405 compares the list length with the node count
    // ---- linearised rest positions and rest state (Blob.cpp:203-208)
    P.p0.resize((size_t)3 * nl);
    for (int n = 0; n < nl; ++n)
        for (int c = 0; c < 3; ++c) P.p0[3 * n + c] = in.pos[3 * P.orig_of_lin[n] + c];
    P.conn.resize((size_t)4 * ne);
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 4; ++s) P.conn[(size_t)s * ne + e] = P.lin_of[top[10 * e + s]];
    P.jinv0.resize((size_t)9 * ne);
    P.vol0.resize(ne);
    P.det0.resize(ne);
    for (int e = 0; e < ne; ++e)
        rest_state(&P.p0[3 * P.conn[e]], &P.p0[3 * P.conn[(size_t)ne + e]], &P.p0[3 * P.conn[(size_t)2 * ne + e]],
                   &P.p0[3 * P.conn[(size_t)3 * ne + e]], &P.jinv0[(size_t)9 * e], &P.vol0[e], &P.det0[e]);
    for (int n = 0; n < nl; ++n)
        for (int k = 0; k < 3; ++k) P.centre[k] += P.p0[3 * n + k];
    for (int k = 0; k < 3; ++k) P.centre[k] /= nl;
    // ---- node -> element incidence in element order (Blob.cpp:3284-3314)
    P.ip.assign(nl + 1, 0);
    P.ic.resize((size_t)4 * ne);
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < 4; ++s) P.ip[P.conn[(size_t)s * ne + e] + 1]++;
    for (int n = 0; n < nl; ++n) P.ip[n + 1] += P.ip[n];
    {
        std::vector<uint32_t> fill(P.ip.begin(), P.ip.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int s = 0; s < 4; ++s) P.ic[fill[P.conn[(size_t)s * ne + e]]++] = (uint32_t)e * 4u + (uint32_t)s;
    }
    {  // are the corner nodes one connected mesh? (union-find over elements)
        std::vector<int> par(nl);
        std::iota(par.begin(), par.end(), 0);
        std::function<int(int)> find = [&](int v) {
            while (par[v] != v) v = par[v] = par[par[v]];
            return v;
        };
        for (int e = 0; e < ne; ++e)
            for (int k = 1; k < 4; ++k) {
                const int ra = find(P.conn[e]), rb = find(P.conn[(size_t)k * ne + e]);
                if (ra != rb) par[ra] = rb;
            }
        const int r0 = find(0);
        for (int n = 1; n < nl && P.connected; ++n) P.connected = find(n) == r0;
    }
    // ---- block pattern (NoMassCGSolver.cpp:81-127 restricted to corner nodes):
    // row n holds a block for each unpinned corner neighbour m (ascending), whose
    // sources are the elements containing both, in element order; pinned rows and
    // rows without element sources hold an identity diagonal block only.
    struct Nb {
        int m;
        uint32_t code;
    };
    P.nslices = (nl + kSlice - 1) / kSlice;
    P.sw.assign(P.nslices, 0);
    auto row_entries = [&](int n, std::vector<Nb> &out) {
        out.clear();
        if (P.pin_lin[n]) return;
        for (uint32_t k = P.ip[n]; k < P.ip[n + 1]; ++k) {
            const int e = (int)(P.ic[k] >> 2), a = (int)(P.ic[k] & 3);
            for (int b = 0; b < 4; ++b) {
                const int m = P.conn[(size_t)b * ne + e];
                if (P.pin_lin[m]) continue;
                out.push_back({m, ((uint32_t)e << 4) | ((uint32_t)a << 2) | (uint32_t)b});
            }
        }
        std::stable_sort(out.begin(), out.end(), [](const Nb &x, const Nb &y) { return x.m < y.m; });
    };
    // three passes over the rows, each over whole slices in parallel ranges (every row
    // writes only its own lane of its slice's slots)
    std::vector<int> sbounds(threads + 1);
    for (int t = 0; t <= threads; ++t) sbounds[t] = (int)((long long)P.nslices * t / threads);
    auto over_slices = [&](const std::function<void(int, int, std::vector<Nb> &)> &row) {
        amg::parallel_ranges(sbounds, [&](int s0, int s1, int) {
            std::vector<Nb> tmp;
            for (int n = s0 * kSlice; n < std::min(nl, s1 * kSlice); ++n) row(n, s1, tmp);
        });
    };
    std::atomic<int> crowded{-1};  // the lowest row with more than 253 neighbours
    over_slices([&](int n, int, std::vector<Nb> &tmp) {
        row_entries(n, tmp);
        int distinct = 0;
        for (size_t t = 0; t < tmp.size(); ++t)
            if (t == 0 || tmp[t].m != tmp[t - 1].m) ++distinct;
        if (distinct == 0) distinct = 1;  // identity diagonal
        if (distinct > 254) {  // the lowest such row is reported
            int cur = crowded.load();
            while ((cur < 0 || n < cur) && !crowded.compare_exchange_weak(cur, n)) {
            }
        }
        P.sw[n / kSlice] = std::max(P.sw[n / kSlice], distinct);
    });
    if (crowded >= 0) throw Invalid(who + fmt("node %d has more than 253 neighbours", P.orig_of_lin[crowded]));
    P.sb.assign(P.nslices + 1, 0);
    for (int s = 0; s < P.nslices; ++s) P.sb[s + 1] = P.sb[s] + (long long)P.sw[s] * kSlice;
    const long long nslots = P.nslots();
    P.col.assign(nslots, 0);
    P.cnt.assign(nslots + 1, 0);
    P.dj.assign(nl, 0);
    std::vector<long long> row_blocks(nl, 0);
    over_slices([&](int n, int, std::vector<Nb> &tmp) {
        const int s = n / kSlice, lane = n % kSlice;
        row_entries(n, tmp);
        const long long base = P.sb[s];
        int j = 0;
        if (tmp.empty()) {
            P.col[base + lane] = n;
            P.cnt[base + lane + 1] = 1;
            P.dj[n] = 0;
            j = 1;
        } else {
            size_t t = 0;
            while (t < tmp.size()) {
                size_t t1 = t;
                while (t1 < tmp.size() && tmp[t1].m == tmp[t].m) ++t1;
                const long long slot = base + (long long)j * kSlice + lane;
                P.col[slot] = tmp[t].m;
                P.cnt[slot + 1] = (uint32_t)(t1 - t);
                if (tmp[t].m == n) P.dj[n] = (unsigned char)j;
                ++j;
                t = t1;
            }
        }
        row_blocks[n] = j;
        for (; j < P.sw[s]; ++j) P.col[base + (long long)j * kSlice + lane] = n;  // padding: zero block
    });
    for (int n = 0; n < nl; ++n) P.nblocks += row_blocks[n];
    for (long long k = 0; k < nslots; ++k) P.cnt[k + 1] += P.cnt[k];
    P.contrib.assign(P.cnt[nslots], 0);
    over_slices([&](int n, int, std::vector<Nb> &tmp) {
        const int s = n / kSlice, lane = n % kSlice;
        row_entries(n, tmp);
        const long long base = P.sb[s];
        if (tmp.empty()) {
            P.contrib[P.cnt[base + lane]] = kIdentity;
            return;
        }
        size_t t = 0;
        int j = 0;
        while (t < tmp.size()) {
            const long long slot = base + (long long)j * kSlice + lane;
            uint32_t w = P.cnt[slot];
            while (t < tmp.size() && tmp[t].m == P.col[slot]) P.contrib[w++] = tmp[t++].code;
            ++j;
        }
    });
    // ---- constant forces and the linearise_force increments of midside ones
    std::vector<double> ctf_total((size_t)3 * nt, 0.0);
    std::vector<unsigned char> has_ctf(nt, 0);
    for (int i = 0; i < in.nctf; ++i) {
        const int v = in.ctf_node[i];
        if (v < 0 || v >= nt) throw Invalid(who + fmt("constant force %d references an out of bounds node index", i));
        for (int c = 0; c < 3; ++c) ctf_total[3 * v + c] += in.ctf_force[3 * i + c];
        has_ctf[v] = 1;
    }
    P.have_ctf = in.nctf > 0;
    P.ctf.assign((size_t)3 * nl, 0.0);
    for (int n = 0; n < nl; ++n)
        for (int c = 0; c < 3; ++c) P.ctf[3 * n + c] = ctf_total[3 * P.orig_of_lin[n] + c];
    // Blob::linearise_force (Blob.cpp:781-799): a midside force is folded half onto each
    // edge end by the first element (in element order) that holds it, then zeroed.
    std::vector<std::vector<std::array<double, 3>>> folds(nl);
    for (int v = 0; v < nt; ++v)
        if (has_ctf[v] && kind[v] == 2) P.have_fold = true;
    if (P.have_fold) {
        std::vector<double> fm(ctf_total);  // only midside entries are read
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
                if (nz) folds[P.lin_of[nn[a]]].push_back(inc3);
            }
            for (int k = 4; k < 10; ++k)
                for (int c = 0; c < 3; ++c) fm[3 * nn[k] + c] = 0.0;
        }
    }
    // ---- surface constant forces (Blob::load_ctforces, Blob.cpp:2730-2806)
    if (in.sf && in.sf->num_sets > 0) {
        const ffea_gpu_surface_forces &F = *in.sf;
        if (F.num_faces < 0 || (F.num_faces && !F.face_nodes) || !F.set_force || !F.set_ptr || !F.set_faces ||
            F.set_ptr[0] != 0)
            throw Invalid(who + "surface constant forces: invalid argument");
        for (int k = 0; k < 3 * F.num_faces; ++k)
            if (F.face_nodes[k] < 0 || F.face_nodes[k] >= nt)
                throw Invalid(who + fmt("surface face %d references an out of bounds node index", k / 3));
        for (int st = 0; st < F.num_sets; ++st)
            if (F.set_ptr[st + 1] < F.set_ptr[st]) throw Invalid(who + "surface constant forces: set_ptr decreases");
        const int nent = F.set_ptr[F.num_sets];
        std::vector<int> compact(F.num_faces, -1);
        for (int k = 0; k < nent; ++k) {
            const int f = F.set_faces[k];
            if (f < 0 || f >= F.num_faces)
                throw Invalid(who + fmt("surface constant force on an out of bounds face index %d", f));
            compact[f] = 0;
        }
        int nfu = 0;
        std::vector<int> used;
        for (int f = 0; f < F.num_faces; ++f)
            if (compact[f] == 0) {
                compact[f] = nfu++;
                used.push_back(f);
            }
        P.has_sf = nent > 0;
        P.sf_rows.resize((size_t)6 * nfu);
        for (int u = 0; u < nfu; ++u)
            for (int i = 0; i < 3; ++i) {
                const int v = F.face_nodes[3 * used[u] + i];
                int ra, rb;
                if (P.lin_of[v] >= 0) {
                    ra = rb = P.lin_of[v];
                } else if (P.mid_a[v] >= 0) {
                    ra = P.lin_of[P.mid_a[v]];
                    rb = P.lin_of[P.mid_b[v]];
                } else {
                    throw Invalid(who + fmt("surface face node %d belongs to no element", v));
                }
                P.sf_rows[6 * u + 2 * i] = ra;
                P.sf_rows[6 * u + 2 * i + 1] = rb;
            }
        P.sf_force.assign(F.set_force, F.set_force + 3 * (size_t)F.num_sets);
        P.sf_set_ptr.assign(F.set_ptr, F.set_ptr + F.num_sets + 1);
        P.sf_entry_face.resize(nent);
        P.sf_entry_set.resize(nent);
        for (int st = 0; st < F.num_sets; ++st)
            for (int k = F.set_ptr[st]; k < F.set_ptr[st + 1]; ++k) {
                P.sf_entry_face[k] = compact[F.set_faces[k]];
                P.sf_entry_set[k] = st;
            }
        P.sf_fe_ptr.assign(nfu + 1, 0);
        for (int k = 0; k < nent; ++k) P.sf_fe_ptr[P.sf_entry_face[k] + 1]++;
        for (int u = 0; u < nfu; ++u) P.sf_fe_ptr[u + 1] += P.sf_fe_ptr[u];
        P.sf_fe.resize(nent);
        {
            std::vector<int> fill(P.sf_fe_ptr.begin(), P.sf_fe_ptr.end() - 1);
            for (int k = 0; k < nent; ++k) P.sf_fe[fill[P.sf_entry_face[k]]++] = k;  // entries ascend: set order
        }
        P.sf_nf_ptr.assign(nt + 1, 0);
        for (int u = 0; u < nfu; ++u)
            for (int i = 0; i < 3; ++i) P.sf_nf_ptr[F.face_nodes[3 * used[u] + i] + 1]++;
        for (int v = 0; v < nt; ++v) P.sf_nf_ptr[v + 1] += P.sf_nf_ptr[v];
        P.sf_nf.resize(P.sf_nf_ptr[nt]);
        {
            std::vector<int> fill(P.sf_nf_ptr.begin(), P.sf_nf_ptr.end() - 1);
            for (int u = 0; u < nfu; ++u)  // face order
                for (int i = 0; i < 3; ++i) P.sf_nf[fill[F.face_nodes[3 * used[u] + i]]++] = u;
        }
    }
    // ---- nodes whose force is formed each step: rotational constant forces
    // (Blob::load_ctforces, Blob.cpp:2650-2726) and external forces (the nodes
    // World::apply_springs pulls). Per node: its external force, then its linear constant
    // forces one by one, then its rotational ones, each in list order, as
    // Blob::add_force_to_node and Blob::apply_ctforces add them (Blob.cpp:1676-1719).
    const int nrot = (in.rf && in.rf->num > 0) ? in.rf->num : 0;
    if (nrot > 0 || in.next > 0) {
        std::vector<int> rcount(nt + 1, 0), lcount(nt + 1, 0);
        std::vector<int> ext_of(nt, -1);
        auto on_element = [&](int v) { return P.lin_of[v] >= 0 || P.mid_a[v] >= 0; };
        if (nrot > 0) {
            const ffea_gpu_rotational_forces &R = *in.rf;
            if (!R.node || !R.force || !R.axis || !R.type)
                throw Invalid(who + "rotational constant forces: invalid argument");
            for (int i = 0; i < R.num; ++i) {
                const int v = R.node[i];
                if (v < 0 || v >= nt)
                    throw Invalid(who + fmt("rotational force %d references an out of bounds node index", i));
                if (R.type[2 * i] != 'p')
                    throw Invalid(who + fmt("rotational force %d: only axes given by point and direction ('p') are "
                                            "applied (the reference leaves 'n' axes unformed)", i));
                if (R.type[2 * i + 1] != 'f' && R.type[2 * i + 1] != 't')
                    throw Invalid(who + fmt("rotational force %d: type must be 'f' or 't'", i));
                if (!on_element(v))
                    throw Invalid(who + fmt("rotational force %d acts on node %d, which belongs to no element", i, v));
                rcount[v + 1]++;
            }
        }
        if (in.next > 0) {
            if (!in.ext) throw Invalid(who + "external force nodes: invalid argument");
            for (int i = 0; i < in.next; ++i) {
                const int v = in.ext[i];
                if (v < 0 || v >= nt)
                    throw Invalid(who + fmt("external force node %d is an out of bounds node index", i));
                if (!on_element(v))
                    throw Invalid(who + fmt("external force node %d belongs to no element", v));
                if (ext_of[v] < 0) {
                    ext_of[v] = (int)P.ext_nodes.size();
                    P.ext_nodes.push_back(v);
                }
            }
        }
        for (int i = 0; i < in.nctf; ++i) lcount[in.ctf_node[i] + 1]++;
        for (int v = 0; v < nt; ++v) {
            rcount[v + 1] += rcount[v];
            lcount[v + 1] += lcount[v];
        }
        std::vector<int> rorder(nrot), lorder(in.nctf);
        {
            std::vector<int> fill(rcount.begin(), rcount.end() - 1);
            for (int i = 0; i < nrot; ++i) rorder[fill[in.rf->node[i]]++] = i;  // list order within a node
            std::vector<int> lfill(lcount.begin(), lcount.end() - 1);
            for (int i = 0; i < in.nctf; ++i) lorder[lfill[in.ctf_node[i]]++] = i;
        }
        P.has_rf = nrot > 0;
        P.has_dn = true;
        P.dn_lptr.push_back(0);
        P.dn_rptr.push_back(0);
        for (int v = 0; v < nt; ++v) {
            if (rcount[v + 1] == rcount[v] && ext_of[v] < 0) continue;
            P.dn_node.push_back(v);
            P.dn_ext.push_back(ext_of[v]);
            if (P.lin_of[v] >= 0) {
                P.dn_row.push_back(P.lin_of[v]);
                P.dn_ends.push_back(P.lin_of[v]);
                P.dn_ends.push_back(P.lin_of[v]);
            } else {
                P.dn_mid = true;
                P.dn_row.push_back(-1);
                P.dn_ends.push_back(P.lin_of[P.mid_a[v]]);
                P.dn_ends.push_back(P.lin_of[P.mid_b[v]]);
            }
            for (int k = lcount[v]; k < lcount[v + 1]; ++k)
                P.dn_lval.insert(P.dn_lval.end(), in.ctf_force + 3 * (size_t)lorder[k],
                                 in.ctf_force + 3 * (size_t)lorder[k] + 3);
            P.dn_lptr.push_back((int)(P.dn_lval.size() / 3));
            for (int k = rcount[v]; k < rcount[v + 1]; ++k) {
                const int i = rorder[k];
                P.rf_axis.insert(P.rf_axis.end(), in.rf->axis + 6 * (size_t)i, in.rf->axis + 6 * (size_t)i + 6);
                P.rf_F.push_back(in.rf->force[i]);
                P.rf_torque.push_back(in.rf->type[2 * i + 1] == 't' ? 1 : 0);
            }
            P.dn_rptr.push_back((int)P.rf_F.size());
        }
        P.have_ctf = true;  // the per-step totals live in the constant-force tables
    }
    P.fp.assign(nl + 1, 0);
    if (stokes_noise || P.has_sf || P.dn_mid) {
        // every node carries noise: the folds are formed each step from the records
        std::vector<std::vector<std::array<int, 3>>> rec(nl);
        std::vector<unsigned char> folded(nt, 0);
        static const int adj[4][3] = {{4, 5, 6}, {4, 7, 8}, {5, 7, 9}, {6, 8, 9}};
        for (int e = 0; e < ne; ++e) {
            const int *nn = top + 10 * e;
            for (int a = 0; a < 4; ++a) {
                std::array<int, 3> r3{-1, -1, -1};
                bool any = false;
                for (int t = 0; t < 3; ++t)
                    if (!folded[nn[adj[a][t]]]) {
                        r3[t] = nn[adj[a][t]];
                        any = true;
                    }
                if (any) rec[P.lin_of[nn[a]]].push_back(r3);
            }
            for (int k = 4; k < 10; ++k) folded[nn[k]] = 1;
        }
        for (int n = 0; n < nl; ++n) {
            P.fm.insert(P.fm.end(), rec[n].begin(), rec[n].end());
            P.fp[n + 1] = P.fp[n] + (uint32_t)rec[n].size();
        }
        P.have_fold = true;
        if (P.have_ctf) P.node_ctf = ctf_total;
    } else {
        for (int n = 0; n < nl; ++n) {
            for (auto &f : folds[n]) P.fv.insert(P.fv.end(), f.begin(), f.end());
            P.fp[n + 1] = P.fp[n] + (uint32_t)folds[n].size();
        }
    }
    if (in.drag) {
        P.drag_lin.assign(nl, 0.0);
        for (int n = 0; n < nl; ++n)
            if (!P.pin_lin[n]) P.drag_lin[n] = in.drag[P.orig_of_lin[n]];
    }
    return P;
}

// per-blob bookkeeping kept on the host
struct BlobMeta {
    int nt = 0, ne = 0, nl = 0;
    int row_off = 0, elem_off = 0;
    long long nblocks = 0;
    bool all_pinned = false, project = false;
    bool free_blob = false;  // no pins, one connected mesh, no drag: singular viscosity matrix
    std::vector<int> lin_of, orig_of_lin, mid_a, mid_b;
    int last_iters = 0;
    long long total_iters = 0;
};

// Every table the kernels read, for all blobs of a handle, built on the host.
struct Tables {
    int nblob = 0, nrows = 0, ne = 0, nslices = 0, nchunks = 0, max_blob_ne = 0, active_blobs = 0;
    long long nslots = 0;
    bool have_ctf = false, have_fold = false, any_project = false;
    std::vector<BlobMeta> meta;
    std::vector<int> conn, sw, col, eblob, eoff, proj;
    std::vector<long long> sb;
    std::vector<uint32_t> cnt, contrib, ip, ic, fp, rng;
    std::vector<uint32_t> inc_pos;       // [4][ne]: position of (element, corner) in the incidence lists
    std::vector<unsigned char> dj, pin;
    std::vector<double> pos, det, ctf, fv;
    std::vector<double> jinv0, vol0, A, B, G, E;  // [9][ne] and [ne], element order
    std::vector<Chunk> chunks;
    std::vector<int2> seg;
    std::vector<double3> centre;
    std::vector<mrg::Jump> step_jump, jump_group, jump_lane;
    // Stokes branch
    bool stokes = false, stokes_noise = false;
    int nnodes = 0;                      // nodes of all blobs (.node files), noise index
    std::vector<double> drag;            // [nrows]
    std::vector<int> node_blob, node_off, row_node;
    std::vector<double> sc, node_ctf;    // [nnodes], [nnodes][3]
    std::vector<int3> fold_mid;
    std::vector<mrg::Jump> node_jump, njump_group, njump_lane;
    // per-step midside folds (Stokes noise or surface constant forces): node-indexed tables
    bool dyn_fold = false;
    // surface constant forces (single-blob handles)
    bool have_sf = false;
    int sf_nfaces = 0, sf_nsets = 0;
    std::vector<int> sf_rows, sf_set_ptr, sf_entry_face, sf_entry_set, sf_fe_ptr, sf_fe, sf_nf_ptr, sf_nf;
    std::vector<double> sf_force;
    // nodes whose force is formed each step (single-blob handles)
    bool have_dn = false, have_rf = false;
    int next = 0;
    std::vector<int> ext_nodes, dn_node, dn_ext, dn_row, dn_ends, dn_lptr, dn_rptr, rf_torque;
    std::vector<double> dn_lval, rf_axis, rf_F;
};

// Rows per reduction chunk of a blob, a function of its own size only (so a blob sums alike
// alone and in an ensemble): blobs of a million rows take kChunkRows, one chunk per
// 1024-thread workgroup, and so do blobs that solve in one workgroup (k_cg_fused, one chunk);
// those between take 256, so that a blob of some 10^5 rows still spreads over the 256 CUs
// (69k rows: 270 chunks rather than 68)
int chunk_rows(int nl) { return (nl <= kFusedMaxRows || nl >= (1 << 18)) ? kChunkRows : kChunkRows / 4; }

Tables make_tables(const std::vector<BlobInput> &in, bool mixed, const ffea_gpu_params &prm) {
    Tables t;
    const int nblob = t.nblob = (int)in.size();
    t.stokes = prm.calc_stokes != 0;
    t.stokes_noise = t.stokes && prm.calc_noise != 0;
    // per-blob plans are independent: build them on up to 16 host threads
    std::vector<BlobPlan> plans(nblob);
    {
        const int nthreads = std::max(1, std::min({nblob, 16, (int)std::thread::hardware_concurrency()}));
        std::atomic<int> next{0};
        std::vector<std::string> errors(nblob);  // the first failing blob is reported
        auto work = [&]() {
            for (int b = next++; b < nblob; b = next++) {
                try {
                    // a single blob's plan takes the host threads itself
                    plans[b] = make_plan(in[b], nblob > 1 ? b : -1, t.stokes_noise,
                                         nblob > 1 ? 1 : std::max(1, std::min(16, (int)std::thread::hardware_concurrency())));
                } catch (const Invalid &e) {
                    errors[b] = e.msg;
                } catch (const std::bad_alloc &) {
                    errors[b] = "host allocation failed";
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
        for (const auto &e : errors)
            if (!e.empty()) throw Invalid(e);
    }
    for (int b = 0; b < nblob; ++b) t.have_sf |= plans[b].has_sf;
    if (t.have_sf && nblob > 1) throw Invalid("surface constant forces need a single-blob handle");
    bool dn_mid = false;
    for (int b = 0; b < nblob; ++b) {
        t.have_dn |= plans[b].has_dn;
        t.have_rf |= plans[b].has_rf;
        dn_mid |= plans[b].dn_mid;
    }
    if (t.have_dn && nblob > 1)
        throw Invalid("rotational constant forces and external forces need a single-blob handle");
    t.dyn_fold = t.stokes_noise || t.have_sf || dn_mid;
    // ---- concatenation offsets: every blob starts on a 64-row slice boundary
    t.meta.resize(nblob);
    long long rows = 0, elems = 0, slots = 0, slices = 0, contribs = 0, nodes = 0;
    for (int b = 0; b < nblob; ++b) {
        BlobMeta &m = t.meta[b];
        const BlobPlan &P = plans[b];
        m.nt = P.nt;
        m.ne = P.ne;
        m.nl = P.nl;
        m.row_off = (int)std::min<long long>(rows, 1ll << 30);
        m.elem_off = (int)std::min<long long>(elems, 1ll << 30);
        m.nblocks = P.nblocks;
        m.all_pinned = P.all_pinned;
        // free blob (no pins): the viscosity matrix is singular with the 6 rigid modes as
        // null space; the mixed form projects them out of every SpMV result
        m.project = mixed && P.npin == 0 && P.connected && !P.all_pinned && !t.stokes;
        m.free_blob = P.npin == 0 && P.connected && !P.all_pinned && !t.stokes;
        m.lin_of = P.lin_of;
        m.orig_of_lin = P.orig_of_lin;
        m.mid_a = P.mid_a;
        m.mid_b = P.mid_b;
        rows += (long long)P.nslices * kSlice;
        elems += P.ne;
        slots += P.nslots();
        slices += P.nslices;
        contribs += (long long)P.contrib.size();
        nodes += P.nt;
        t.max_blob_ne = std::max(t.max_blob_ne, P.ne);
        if (!m.all_pinned) ++t.active_blobs;
        t.have_ctf |= P.have_ctf;
        t.have_fold |= P.have_fold;
        t.any_project |= m.project;
    }
    if (rows >= (1ll << 30) || elems >= (1ll << 28) - 1 || slots >= (1ll << 31) || contribs >= (1ll << 32) - 1 ||
        nodes >= (1ll << 31) / 3)
        throw Invalid("problem too large for 32-bit indices");
    const int nrows = t.nrows = (int)rows;
    const int ne = t.ne = (int)elems;
    const long long nslots = t.nslots = slots;
    t.nslices = (int)slices;
    // ---- concatenated arrays
    t.conn.assign((size_t)4 * ne, 0);
    t.sw.assign(t.nslices, 0);
    t.sb.assign(t.nslices, 0);
    t.col.assign(nslots, 0);
    t.cnt.assign(nslots + 1, 0);
    t.contrib.reserve((size_t)contribs);
    t.ip.assign(nrows + 1, 0);
    t.ic.assign((size_t)4 * ne, 0);
    t.fp.assign(nrows + 1, 0);
    t.dj.assign(nrows, 0xFF);
    t.pin.assign(nrows, 0);
    t.pos.assign((size_t)3 * nrows, 0.0);
    t.det.assign(ne, 0.0);
    if (t.have_ctf) t.ctf.assign((size_t)3 * nrows, 0.0);
    if (nblob > 1) t.eblob.assign(ne, 0);
    t.eoff.assign(nblob, 0);
    t.proj.assign(nblob, 0);
    t.seg.resize(nblob);
    t.centre.resize(nblob);
    t.step_jump.resize(nblob);
    t.rng.assign((size_t)6 * nblob, 0);
    t.nnodes = (int)nodes;
    if (t.stokes) t.drag.assign(nrows, 0.0);
    if (t.dyn_fold) {
        t.node_off.assign(nblob, 0);
        t.row_node.assign(nrows, -1);
        if (t.have_ctf) t.node_ctf.assign((size_t)3 * t.nnodes, 0.0);
    }
    if (t.stokes_noise) {
        if (nblob > 1) t.node_blob.assign(t.nnodes, 0);
        t.sc.assign(t.nnodes, 0.0);
        t.node_jump.resize(nblob);
    }
    t.jinv0.assign((size_t)9 * ne, 0.0);
    t.vol0.assign(ne, 0.0);
    t.A.assign(ne, 0.0);
    t.B.assign(ne, 0.0);
    t.G.assign(ne, 0.0);
    t.E.assign(ne, 0.0);
    long long slot_off = 0;
    int slice_off = 0, node_off = 0, max_nt = 0;
    for (int b = 0; b < nblob; ++b) {
        BlobPlan &P = plans[b];
        if (t.stokes)
            for (int n = 0; n < P.nl; ++n) t.drag[t.meta[b].row_off + n] = P.drag_lin[n];
        if (t.dyn_fold) {
            t.node_off[b] = node_off;
            max_nt = std::max(max_nt, P.nt);
            for (int i = 0; i < P.nt; ++i) {
                if (t.stokes_noise) {
                    if (nblob > 1) t.node_blob[node_off + i] = b;
                    // Blob.cpp:3214-3216, evaluated on the host in IEEE double
                    t.sc[node_off + i] = std::sqrt((24 * prm.kT * in[b].drag[i]) / (prm.dt));
                }
                if (t.have_ctf && !P.node_ctf.empty())
                    for (int c = 0; c < 3; ++c) t.node_ctf[3 * (size_t)(node_off + i) + c] = P.node_ctf[3 * (size_t)i + c];
            }
            for (int n = 0; n < P.nl; ++n) t.row_node[t.meta[b].row_off + n] = node_off + P.orig_of_lin[n];
            for (const auto &r3 : P.fm)
                t.fold_mid.push_back(make_int3(r3[0] < 0 ? -1 : node_off + r3[0], r3[1] < 0 ? -1 : node_off + r3[1],
                                               r3[2] < 0 ? -1 : node_off + r3[2]));
            if (t.stokes_noise) t.node_jump[b] = mrg::h_pow(7ull * (unsigned long long)P.ne);
        }
        if (P.has_sf) {  // single blob: rows and nodes need no offsets
            t.sf_nfaces = (int)P.sf_fe_ptr.size() - 1;
            t.sf_nsets = (int)P.sf_set_ptr.size() - 1;
            t.sf_rows = std::move(P.sf_rows);
            t.sf_force = std::move(P.sf_force);
            t.sf_set_ptr = std::move(P.sf_set_ptr);
            t.sf_entry_face = std::move(P.sf_entry_face);
            t.sf_entry_set = std::move(P.sf_entry_set);
            t.sf_fe_ptr = std::move(P.sf_fe_ptr);
            t.sf_fe = std::move(P.sf_fe);
            t.sf_nf_ptr = std::move(P.sf_nf_ptr);
            t.sf_nf = std::move(P.sf_nf);
        }
        if (P.has_dn) {  // single blob: rows and nodes need no offsets
            t.next = (int)P.ext_nodes.size();
            t.ext_nodes = std::move(P.ext_nodes);
            t.dn_node = std::move(P.dn_node);
            t.dn_ext = std::move(P.dn_ext);
            t.dn_row = std::move(P.dn_row);
            t.dn_ends = std::move(P.dn_ends);
            t.dn_lptr = std::move(P.dn_lptr);
            t.dn_lval = std::move(P.dn_lval);
            t.dn_rptr = std::move(P.dn_rptr);
            t.rf_axis = std::move(P.rf_axis);
            t.rf_F = std::move(P.rf_F);
            t.rf_torque = std::move(P.rf_torque);
        }
        const BlobMeta &m = t.meta[b];
        const int ro = m.row_off, eo = m.elem_off;
        t.eoff[b] = eo;
        t.proj[b] = m.project ? 1 : 0;
        for (int s = 0; s < 4; ++s)
            for (int e = 0; e < P.ne; ++e) t.conn[(size_t)s * ne + eo + e] = P.conn[(size_t)s * P.ne + e] + ro;
        for (int e = 0; e < P.ne; ++e) {
            if (nblob > 1) t.eblob[eo + e] = b;
            t.det[eo + e] = P.det0[e];
            for (int k = 0; k < 9; ++k) t.jinv0[(size_t)k * ne + eo + e] = P.jinv0[(size_t)9 * e + k];
            t.vol0[eo + e] = P.vol0[e];
            t.A[eo + e] = in[b].A[e];
            t.B[eo + e] = in[b].B[e];
            t.G[eo + e] = in[b].G[e];
            t.E[eo + e] = in[b].E[e];
        }
        for (int s = 0; s < P.nslices; ++s) {
            t.sw[slice_off + s] = P.sw[s];
            t.sb[slice_off + s] = slot_off + P.sb[s];
        }
        const uint32_t cbase = (uint32_t)t.contrib.size();
        for (long long k = 0; k < P.nslots(); ++k) {
            t.col[slot_off + k] = P.col[k] + ro;
            t.cnt[slot_off + k + 1] = cbase + P.cnt[k + 1];
        }
        for (uint32_t code : P.contrib)
            t.contrib.push_back(code == kIdentity ? kIdentity : code + ((uint32_t)eo << 4));
        for (int n = 0; n < P.nl; ++n) {
            t.dj[ro + n] = P.dj[n];
            t.pin[ro + n] = P.pin_lin[n];
            for (int k = 0; k < 3; ++k) t.pos[3 * (size_t)(ro + n) + k] = P.p0[3 * n + k];
            if (t.have_ctf)
                for (int k = 0; k < 3; ++k) t.ctf[3 * (size_t)(ro + n) + k] = P.ctf[3 * n + k];
            // per-row counts, prefix-summed over all rows below
            t.ip[ro + n + 1] = P.ip[n + 1] - P.ip[n];
            t.fp[ro + n + 1] = P.fp[n + 1] - P.fp[n];
        }
        for (size_t k = 0; k < P.ic.size(); ++k) t.ic[(size_t)4 * eo + k] = P.ic[k] + 4u * (uint32_t)eo;
        t.fv.insert(t.fv.end(), P.fv.begin(), P.fv.end());
        // reduction chunks: up to chunk_rows(nl) rows of one blob each
        t.seg[b].x = (int)t.chunks.size();
        if (!m.all_pinned) {
            const int first = (int)t.chunks.size(), cr = chunk_rows(P.nl);
            for (int r0 = 0; r0 < P.nl; r0 += cr)
                t.chunks.push_back(Chunk{b, ro + r0, ro + std::min(P.nl, r0 + cr), first});
        }
        t.seg[b].y = (int)t.chunks.size();
        t.centre[b] = make_double3(P.centre[0], P.centre[1], P.centre[2]);
        // noise stream: 7 draws per element per pass; none for an all-pinned blob, whose
        // update_internal_forces returns at once (Blob.cpp:405-407)
        // with Stokes noise the step also draws 3 per node (Blob.cpp:3214-3216)
        t.step_jump[b] = m.all_pinned ? mrg::h_identity()
                                      : mrg::h_pow(7ull * (unsigned long long)P.ne +
                                                   (t.stokes_noise ? 3ull * (unsigned long long)P.nt : 0ull));
        for (int i = 0; i < 6; ++i) t.rng[6 * b + i] = in[b].rng[i];
        slot_off += P.nslots();
        slice_off += P.nslices;
        node_off += P.nt;
        P = BlobPlan();  // release the blob's host plan
    }
    for (int n = 0; n < nrows; ++n) {
        t.ip[n + 1] += t.ip[n];
        t.fp[n + 1] += t.fp[n];
    }
    // the element pass writes each corner's force where the gather of its row reads it
    t.inc_pos.assign((size_t)4 * ne, 0xFFFFFFFFu);
    for (size_t k = 0; k < t.ic.size(); ++k) t.inc_pos[(size_t)(t.ic[k] & 3) * ne + (t.ic[k] >> 2)] = (uint32_t)k;
    t.nchunks = (int)t.chunks.size();
    // ---- noise jump tables: A^(7*64*g) per group of 64 elements, A^(7*l) per lane
    const int ngroups = std::max(1, (t.max_blob_ne + 63) / 64);
    t.jump_group.resize(ngroups);
    t.jump_lane.resize(64);
    const mrg::Jump a7 = mrg::h_pow(7), a448 = mrg::h_pow(7ull * 64ull);
    t.jump_lane[0] = mrg::h_identity();
    for (int l = 1; l < 64; ++l) t.jump_lane[l] = mrg::h_mul(a7, t.jump_lane[l - 1]);
    t.jump_group[0] = mrg::h_identity();
    for (int g = 1; g < ngroups; ++g) t.jump_group[g] = mrg::h_mul(a448, t.jump_group[g - 1]);
    if (t.stokes_noise) {  // node draws: A^(3*64*g) per group of 64 nodes, A^(3*l) per lane
        const int nng = std::max(1, (max_nt + 63) / 64);
        t.njump_group.resize(nng);
        t.njump_lane.resize(64);
        const mrg::Jump a3 = mrg::h_pow(3), a192 = mrg::h_pow(3ull * 64ull);
        t.njump_lane[0] = mrg::h_identity();
        for (int l = 1; l < 64; ++l) t.njump_lane[l] = mrg::h_mul(a3, t.njump_lane[l - 1]);
        t.njump_group[0] = mrg::h_identity();
        for (int g = 1; g < nng; ++g) t.njump_group[g] = mrg::h_mul(a192, t.njump_group[g - 1]);
    }
    return t;
}

// Tables of the LDS-staged assembly (k_assemble_staged): rows in groups of G, each group's
// elements (every element with a corner among its rows, from the rows' incidence lists)
// ascending in gel[gptr[g] .. gptr[g+1]), and for each block source its element's place in
// the list of its row's group, cloc = (place<<4)|(a<<2)|b (kLocalIdentity for an identity
// source). Groups with more than cap elements keep cloc unset; the kernel reads their
// records from memory. Built on up to 16 host threads.
struct StageTables {
    std::vector<uint32_t> gptr, gel;
    std::vector<uint16_t> cloc;
    int overflow = 0;
};

StageTables stage_tables(const Tables &t, int G, int cap) {
    StageTables st;
    if (t.nrows % G != 0) throw Invalid("internal layout error: rows are not whole assembly groups");
    const int ngroups = t.nrows / G;
    std::vector<std::vector<uint32_t>> lists(ngroups);
    st.cloc.assign(t.contrib.size(), kLocalIdentity);
    std::atomic<int> next{0}, bad{0}, over{0};
    auto work = [&]() {
        constexpr int kBatch = 256;
        for (int g0 = (next += kBatch) - kBatch; g0 < ngroups; g0 = (next += kBatch) - kBatch) {
            for (int g = g0; g < std::min(ngroups, g0 + kBatch); ++g) {
                std::vector<uint32_t> &L = lists[g];
                for (int n = g * G; n < (g + 1) * G; ++n)
                    for (uint32_t k = t.ip[n]; k < t.ip[n + 1]; ++k) L.push_back(t.ic[k] >> 2);
                std::sort(L.begin(), L.end());
                L.erase(std::unique(L.begin(), L.end()), L.end());
                if ((int)L.size() > cap) {
                    ++over;
                    continue;
                }
                for (int n = g * G; n < (g + 1) * G; ++n) {
                    const int sl = n / kSlice, lane = n % kSlice;
                    for (int j = 0; j < t.sw[sl]; ++j) {
                        const long long slot = t.sb[sl] + (long long)j * kSlice + lane;
                        for (uint32_t c = t.cnt[slot]; c < t.cnt[slot + 1]; ++c) {
                            const uint32_t code = t.contrib[c];
                            if (code == kIdentity) continue;
                            const auto it = std::lower_bound(L.begin(), L.end(), code >> 4);
                            if (it == L.end() || *it != (code >> 4)) {
                                ++bad;
                                continue;
                            }
                            st.cloc[c] = (uint16_t)(((uint32_t)(it - L.begin()) << 4) | (code & 15u));
                        }
                    }
                }
            }
        }
    };
    const int nthreads = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int k = 1; k < nthreads; ++k) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    if (bad) throw Invalid("internal layout error: a block source outside its row's element list");
    st.overflow = over;
    st.gptr.assign(ngroups + 1, 0);
    for (int g = 0; g < ngroups; ++g) st.gptr[g + 1] = st.gptr[g] + (uint32_t)lists[g].size();
    st.gel.reserve(st.gptr[ngroups]);
    for (auto &L : lists) {
        st.gel.insert(st.gel.end(), L.begin(), L.end());
        std::vector<uint32_t>().swap(L);
    }
    for (uint32_t e : st.gel)
        if ((long long)e >= t.ne) throw Invalid("internal layout error: assembly group element");
    return st;
}

// Every index a kernel forms from the tables lies inside the buffer it addresses
// (the kernels themselves do not check). Throws Invalid naming the first violation.
void validate(const Tables &t) {
    auto fail = [](const char *what, long long i) { throw Invalid(fmt("internal layout error: %s at %lld", what, i)); };
    const long long nrows = t.nrows, ne = t.ne, nslots = t.nslots, nc = t.nchunks;
    if (nrows != (long long)t.nslices * kSlice) fail("row count is not whole slices", nrows);
    if ((long long)t.conn.size() != 4 * ne) fail("conn size", (long long)t.conn.size());
    for (size_t i = 0; i < t.conn.size(); ++i)
        if (t.conn[i] < 0 || t.conn[i] >= nrows) fail("conn", (long long)i);
    long long expect_base = 0;
    for (int s = 0; s < t.nslices; ++s) {
        if (t.sb[s] != expect_base || t.sw[s] < 1 || t.sw[s] > 254) fail("slice base/width", s);
        expect_base += (long long)t.sw[s] * kSlice;
    }
    if (expect_base != nslots || (long long)t.col.size() != nslots) fail("slot count", expect_base);
    for (long long k = 0; k < nslots; ++k)
        if (t.col[k] < 0 || t.col[k] >= nrows) fail("block column", k);
    if ((long long)t.cnt.size() != nslots + 1 || t.cnt[0] != 0 || t.cnt[nslots] != t.contrib.size())
        fail("source prefix ends", nslots);
    for (long long k = 0; k < nslots; ++k)
        if (t.cnt[k + 1] < t.cnt[k]) fail("source prefix order", k);
    for (size_t i = 0; i < t.contrib.size(); ++i)
        if (t.contrib[i] != kIdentity && (long long)(t.contrib[i] >> 4) >= ne) fail("source element", (long long)i);
    if ((long long)t.dj.size() != nrows || (long long)t.pin.size() != nrows) fail("row table size", nrows);
    for (long long n = 0; n < nrows; ++n)
        if (t.dj[n] != 0xFF && t.dj[n] >= t.sw[n / kSlice]) fail("diagonal slot", n);
    if ((long long)t.ip.size() != nrows + 1 || t.ip[0] != 0 || (long long)t.ip[nrows] != 4 * ne ||
        (long long)t.ic.size() != 4 * ne)
        fail("incidence ends", nrows);
    for (long long n = 0; n < nrows; ++n)
        if (t.ip[n + 1] < t.ip[n]) fail("incidence order", n);
    for (size_t i = 0; i < t.ic.size(); ++i)
        if ((long long)(t.ic[i] >> 2) >= ne) fail("incidence element", (long long)i);
    if ((long long)t.inc_pos.size() != 4 * ne) fail("incidence positions", (long long)t.inc_pos.size());
    for (size_t i = 0; i < t.inc_pos.size(); ++i)  // each (element, corner) once in the lists
        if ((long long)t.inc_pos[i] >= 4 * ne || t.ic[t.inc_pos[i]] != (uint32_t)(4 * (i % ne) + i / ne))
            fail("incidence position", (long long)i);
    if (t.have_fold && !t.dyn_fold) {
        if ((long long)t.fp.size() != nrows + 1 || t.fp[0] != 0 || 3 * (size_t)t.fp[nrows] != t.fv.size())
            fail("fold ends", nrows);
        for (long long n = 0; n < nrows; ++n)
            if (t.fp[n + 1] < t.fp[n]) fail("fold order", n);
    }
    if (t.have_ctf && (long long)t.ctf.size() != 3 * nrows) fail("ctf size", (long long)t.ctf.size());
    if ((long long)t.pos.size() != 3 * nrows || (long long)t.det.size() != ne) fail("node/element sizes", nrows);
    for (const std::vector<double> *v : {&t.vol0, &t.A, &t.B, &t.G, &t.E})
        if ((long long)v->size() != ne) fail("element array size", (long long)v->size());
    if ((long long)t.jinv0.size() != 9 * ne) fail("rest-state size", (long long)t.jinv0.size());
    if ((int)t.meta.size() != t.nblob || (int)t.seg.size() != t.nblob || (int)t.eoff.size() != t.nblob ||
        (int)t.proj.size() != t.nblob || (int)t.centre.size() != t.nblob || (int)t.step_jump.size() != t.nblob ||
        (long long)t.rng.size() != 6ll * t.nblob)
        fail("per-blob table size", t.nblob);
    if (t.nblob > 1 && (long long)t.eblob.size() != ne) fail("element blob size", (long long)t.eblob.size());
    const long long ngroups = (long long)t.jump_group.size();
    for (int b = 0; b < t.nblob; ++b) {
        const BlobMeta &m = t.meta[b];
        if (m.elem_off != t.eoff[b] || m.elem_off + (long long)m.ne > ne || m.row_off % kSlice ||
            m.row_off + (long long)m.nl > nrows)
            fail("blob offsets", b);
        if (((long long)m.ne + 63) / 64 > ngroups) fail("jump group table", b);
        for (int e = 0; e < m.ne && t.nblob > 1; ++e)
            if (t.eblob[m.elem_off + e] != b) fail("element blob", (long long)m.elem_off + e);
        const int2 sg = t.seg[b];
        if (sg.x < 0 || sg.x > sg.y || sg.y > nc) fail("chunk segment", b);
        for (int c = sg.x; c < sg.y; ++c) {
            const Chunk &ch = t.chunks[c];
            if (ch.blob != b || ch.first != sg.x || ch.begin < m.row_off || ch.end > m.row_off + m.nl ||
                ch.begin >= ch.end)
                fail("chunk", c);
        }
    }
    if ((long long)t.chunks.size() != nc || t.jump_lane.size() != 64) fail("chunk/jump table size", nc);
    if (t.stokes && (long long)t.drag.size() != nrows) fail("drag size", (long long)t.drag.size());
    if (t.dyn_fold) {
        const long long nn = t.nnodes;
        if ((long long)t.row_node.size() != nrows || (int)t.node_off.size() != t.nblob ||
            (t.have_ctf && (long long)t.node_ctf.size() != 3 * nn))
            fail("node table size", nn);
        if (t.stokes_noise && ((long long)t.sc.size() != nn || (int)t.node_jump.size() != t.nblob ||
                               t.njump_lane.size() != 64 || (t.nblob > 1 && (long long)t.node_blob.size() != nn)))
            fail("noise table size", nn);
        if ((long long)t.fp.size() != nrows + 1 || (size_t)t.fp[nrows] != t.fold_mid.size()) fail("fold record ends", nrows);
        for (long long n = 0; n < nrows; ++n)
            if (t.fp[n + 1] < t.fp[n]) fail("fold record order", n);
        for (int b = 0; b < t.nblob; ++b) {
            const BlobMeta &m = t.meta[b];
            if (t.node_off[b] + (long long)m.nt > nn) fail("node range", b);
            if (t.stokes_noise && ((long long)m.nt + 63) / 64 > (long long)t.njump_group.size())
                fail("noise node range", b);
            for (int n = 0; n < m.nl; ++n) {
                const int g = t.row_node[m.row_off + n];
                if (g < t.node_off[b] || g >= t.node_off[b] + m.nt) fail("row node", (long long)m.row_off + n);
            }
            for (int i = 0; t.stokes_noise && i < m.nt && t.nblob > 1; ++i)
                if (t.node_blob[t.node_off[b] + i] != b) fail("node blob", (long long)t.node_off[b] + i);
        }
        for (size_t k = 0; k < t.fold_mid.size(); ++k) {
            const int m3[3] = {t.fold_mid[k].x, t.fold_mid[k].y, t.fold_mid[k].z};
            for (int q : m3)
                if (q < -1 || q >= nn) fail("fold record node", (long long)k);
        }
    }
    if (t.have_sf) {
        const long long nf = t.sf_nfaces, ns = t.sf_nsets, nn = t.nnodes;
        const long long nent = t.sf_entry_face.size();
        if (t.nblob != 1 || nf <= 0 || ns <= 0 || (long long)t.sf_rows.size() != 6 * nf ||
            (long long)t.sf_force.size() != 3 * ns || (long long)t.sf_set_ptr.size() != ns + 1 || t.sf_set_ptr[0] != 0 ||
            t.sf_set_ptr[ns] != nent || (long long)t.sf_entry_set.size() != nent ||
            (long long)t.sf_fe_ptr.size() != nf + 1 || t.sf_fe_ptr[0] != 0 || t.sf_fe_ptr[nf] != nent ||
            (long long)t.sf_fe.size() != nent || (long long)t.sf_nf_ptr.size() != nn + 1 || t.sf_nf_ptr[0] != 0 ||
            (long long)t.sf_nf.size() != t.sf_nf_ptr[nn] || !t.dyn_fold)
            fail("surface force table size", nf);
        for (long long k = 0; k < 6 * nf; ++k)
            if (t.sf_rows[k] < 0 || t.sf_rows[k] >= nrows) fail("surface face row", k);
        for (long long i = 0; i < ns; ++i)
            if (t.sf_set_ptr[i + 1] < t.sf_set_ptr[i]) fail("surface set order", i);
        for (long long k = 0; k < nent; ++k)
            if (t.sf_entry_face[k] < 0 || t.sf_entry_face[k] >= nf || t.sf_entry_set[k] < 0 || t.sf_entry_set[k] >= ns ||
                t.sf_fe[k] < 0 || t.sf_fe[k] >= nent)
                fail("surface entry", k);
        for (long long f = 0; f < nf; ++f)
            if (t.sf_fe_ptr[f + 1] < t.sf_fe_ptr[f]) fail("surface face entries", f);
        for (long long v = 0; v < nn; ++v)
            if (t.sf_nf_ptr[v + 1] < t.sf_nf_ptr[v]) fail("surface node faces", v);
        for (size_t k = 0; k < t.sf_nf.size(); ++k)
            if (t.sf_nf[k] < 0 || t.sf_nf[k] >= nf) fail("surface node face", (long long)k);
    }
    if (t.have_dn) {
        const long long n = t.dn_node.size(), nn = t.nnodes;
        if (t.nblob != 1 || n <= 0 || (long long)t.dn_ext.size() != n || (long long)t.dn_row.size() != n ||
            (long long)t.dn_ends.size() != 2 * n || (long long)t.dn_lptr.size() != n + 1 || t.dn_lptr[0] != 0 ||
            (long long)t.dn_rptr.size() != n + 1 || t.dn_rptr[0] != 0 || !t.have_ctf ||
            (long long)t.ctf.size() != 3 * nrows || (long long)t.ext_nodes.size() != t.next ||
            (t.dyn_fold && (long long)t.node_ctf.size() != 3 * nn))
            fail("dynamic node table size", n);
        const long long nl = t.dn_lptr[n], nr = t.dn_rptr[n];
        if ((long long)t.dn_lval.size() != 3 * nl || (long long)t.rf_axis.size() != 6 * nr ||
            (long long)t.rf_F.size() != nr || (long long)t.rf_torque.size() != nr)
            fail("dynamic node entries", nr);
        for (long long k = 0; k < n; ++k) {
            if (t.dn_node[k] < 0 || t.dn_node[k] >= nn || t.dn_row[k] < -1 || t.dn_row[k] >= nrows ||
                t.dn_ends[2 * k] < 0 || t.dn_ends[2 * k] >= nrows || t.dn_ends[2 * k + 1] < 0 ||
                t.dn_ends[2 * k + 1] >= nrows || t.dn_lptr[k + 1] < t.dn_lptr[k] || t.dn_rptr[k + 1] < t.dn_rptr[k] ||
                t.dn_ext[k] < -1 || t.dn_ext[k] >= t.next)
                fail("dynamic node", k);
            if (t.dn_row[k] < 0 && !t.dyn_fold) fail("dynamic midside node without per-step folds", k);
        }
    }
}

}  // namespace

// ---------------------------------------------------------------- the handle
struct ffea_gpu_blob {
    ffea_gpu_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool mixed = false, any_project = false;
    size_t dev_bytes = 0;

    int nblob = 0, nrows = 0, ne = 0, nslices = 0, nchunks = 0, max_blob_ne = 0;
    long long nslots = 0;
    std::vector<BlobMeta> meta;
    int active_blobs = 0;  // blobs that are solved (not all pinned)

    // device arrays
    DevBuf<int> conn, slice_w, col, elem_blob, elem_off, project;
    DevBuf<long long> slice_base;
    DevBuf<uint32_t> cptr, contrib, inc_ptr, inc, inc_pos, fold_ptr, rng;
    DevBuf<uint32_t> st_ptr, st_el;     // LDS-staged assembly: group element lists
    DevBuf<uint16_t> st_loc;            // ... and each source's place in its group's list
    DevBuf<unsigned char> diag_j, pinned;
    DevBuf<double> pos, rest, vel, last_det, dinv, rhs, r, z, x, q, p, ctf, fold, Gmat;
    DevBuf<double> Pg, Pi, P1, P2;  // chunk partials: gather, CG start, SpMV step, vector step
    DevBuf<float> z32;              // FP32-mixed form: packed FP32 copy of z, gathered by the row products
    DevBuf<double> d_jinv0, d_vol0, d_G, d_E, d_A, d_B, d_force, d_geo, d_val;  // FP64 form
    DevBuf<float> f_jinv0, f_vol0, f_G, f_E, f_A, f_B, f_force, f_geo, f_val;   // mixed form
    DevBuf<mrg::Jump> jump_group, jump_lane, step_jump;
    DevBuf<CgScalars> S;
    DevBuf<Chunk> chunks;
    DevBuf<int2> seg;
    DevBuf<double3> centre;
    DevBuf<int> num_done;
    bool have_ctf = false, have_fold = false;
    // Stokes branch
    bool stokes = false, stokes_noise = false;
    int nnodes = 0, grid_nodes = 0;
    DevBuf<double> drag, node_sc, node_ctf, noise;
    DevBuf<int> node_blob, node_off, row_node;
    DevBuf<int3> fold_mid;
    DevBuf<mrg::Jump> node_jump, njump_group, njump_lane;
    bool dyn_fold = false;  // midside folds formed per step (Stokes noise, surface or rotational forces)
    // surface constant forces (single blob)
    bool have_sf = false;
    int sf_nfaces = 0, sf_nsets = 0, sf_nent = 0;
    DevBuf<int> sf_rows, sf_set_ptr, sf_entry_face, sf_entry_set, sf_fe_ptr, sf_fe, sf_nf_ptr, sf_nf;
    DevBuf<double> sf_force, sf_area, sf_total, sf_face_force;
    // nodes whose force is formed each step (single blob): rotational constant forces,
    // external forces (staged on the host by ffea_gpu_blob_add_node_forces)
    bool have_dn = false, have_rf = false;
    int dn_n = 0, next = 0;
    DevBuf<int> dn_node, dn_ext, dn_row, dn_ends, dn_lptr, dn_rptr, rf_torque;
    DevBuf<double> dn_lval, rf_axis, rf_F, ext_force;
    std::vector<int> ext_slot;        // [num_nodes] external force slot of each node, or -1
    std::vector<double> ext_staged;   // [next][3] forces added since the last force formation
    bool ext_dirty = false;           // ext_staged holds forces not yet uploaded
    bool ext_on_device = false;       // ext_force holds forces of the coming step (to clear after it)
    DevBuf<int> take_rows;            // ffea_gpu_blob_get_node_positions scratch
    DevBuf<double> take_out;
    // inertial solvers (ffea_gpu_blob_set_solver): MassLumpedSolver's explicit steps, or
    // ConjugateGradientSolver's (mass_cg: the consistent mass matrix in the CG's slots)
    bool mass = false, mass_cg = false;
    bool project_saved = false;        // any_project of the mass-less solver, off for the mass CG
    DevBuf<double> inv_m, rho_dev;     // [rows] 1 / lumped mass, [ne] densities
    int grid_rows = 0, grid_elems = 0, grid_blobs = 0;
    // solver-only handle (ffea_gpu_solver_create): the mass-less viscosity solve of a caller
    // whose Blob keeps its own element loop; no element forces, steps or adjoint
    bool solver_only = false;
    const double *rhs_in = nullptr;    // set while a handed-over right-hand side is gathered
    std::vector<double> mid_diag;      // [num_nodes] diagonal of each midside node's identity row
                                       // (1, plus the drag of an unpinned node under calc_stokes)
    std::vector<double> h_vec;         // tiled host staging of one node vector
    // FP64 refinement of the solves (ffea_gpu_blob_set_refinement): the true residual by an
    // FP64 matrix-free product, FP32 correction solves, until |b - V x| <= refine_tol |b|
    double refine_tol = 0.0;           // 0: off
    double eps2_override = 0.0;        // > 0: the CG's stopping threshold for a correction solve
    int refine_sweeps_last = 0;
    long long refine_sweeps_total = 0;
    DevBuf<double> w64, xacc, bsave, res64, Pr, A64, B64;
    std::vector<double> h_pr;
    double cg_eps2() const { return eps2_override > 0.0 ? eps2_override : prm.epsilon2; }
    // every blob has at most kFusedMaxRows rows: each CG solve is one k_cg_fused launch
    bool fused_cg = false;
    bool small_chunks = false;  // no chunk has more than kMixedBlock rows
    int nb = 0, cpb = 0;  // row-kernel grid and reduction chunks per block

    // host-visible status (pinned)
    CgScalars *h_S = nullptr;  // [nblob]
    int *h_done = nullptr;     // [2]
    hipEvent_t chunk_ev[2] = {nullptr, nullptr};
    int expect_iters = 0;

    // one whole step captured as a graph
    bool use_graph = false, capturing = false;
    hipGraphExec_t graph_exec = nullptr;
    int graph_iters = 0;
    int graph_fixed_iters = 0;  // > 0: CG iterations captured; 0: from the last step's count

    // checkpointed run and its reverse (single blob, noise off)
    int zero_rhs_solved = 0;  // CG: a zero right-hand side is solved by x = 0 (adjoint solves)
    int run_steps = 0, run_interval = 0;
    DevBuf<double> ckpt, seg_x, seg_v, x_end, v_end, lam, grad;
    DevBuf<double> energy_w, energy_out;  // strain energy per element / per blob
    DevBuf<int> elem_cnt;
    DevBuf<double> d_w;
    DevBuf<float> f_w;

    // multigrid preconditioner (single blob): levels 0..L-1, level L-1 inverted densely
    struct MgEll {
        int n = 0;
        double mscale = 1.0;   // inverse of the values' power-of-two scale
        DevBuf<int> sw, col;
        DevBuf<long long> sb;
        DevBuf<_Float16> val;
    };
    struct MgCsr {
        int n = 0;
        double mscale = 1.0;
        DevBuf<int> rp, ci;
        DevBuf<_Float16> val;
    };
    struct MgLevel {
        int n = 0;             // block rows
        double omega = 0.0;    // smoother weight
        double lmax = 0.0, fold = 0.0;  // levels >= 1 of a single blob: the power estimate of lambda_max(D^-1 A)
                                        // and the weight folded into dinv
        MgCsr A, R;            // A: levels >= 1 (level 0 is the CG's matrix); R: to level + 1
        MgEll P;               // prolongation from level + 1 (few blocks per row: a thread per row)
        DevBuf<double> dinv, b, xa, res, xb, e;  // levels >= 1, [n][6]
    };
    int precond = FFEA_GPU_PRECOND_JACOBI;
    std::vector<std::unique_ptr<MgLevel>> mg;
    DevBuf<float> mg_cinv;
    int mg_nc = 0;
    DevBuf<_Float16> mg_hval;  // scaled FP16 copy of the viscosity matrix, written by the assembly
    DevBuf<float> mg_binv;     // [nrows][6] level-0 smoother: inverses of the diagonal blocks, written by the assembly
    double mg_hscale = 1.0;
    DevBuf<double> mg_res0, mg_xb0, mg_rp, mg_coef;
    DevBuf<float> mg_xb32, mg_res32;  // FP32-mixed form: packed FP32 copies of mg_xb0, mg_res0
    DevBuf<double> mg_omega0;                           // ensemble: level-0 smoother weight per row
    DevBuf<int> mg_crow_blob, mg_coff_row, mg_cm, mg_free;  // ensemble coarsest tables; free blobs
    DevBuf<long long> mg_ccoff;
    bool mg_any_free = false;
    ffea_gpu_multigrid_info mg_info{};
    bool mg_on() const { return precond == FFEA_GPU_PRECOND_MULTIGRID; }
    // the CG's z = B^-1 r under the multigrid preconditioner (its pre-smoothed level-0
    // iterate, up to omega); null: Jacobi-PCG's z = D^-1 r
    const float *mg_smoother() const { return mg_on() ? (const float *)mg_binv.p : nullptr; }

    // timing
    unsigned timing = 0;  // bit i: time phase i
    int cg_sample = 8;    // CG phases: time one iteration in cg_sample
    Phase phase[FFEA_GPU_NUM_PHASES];
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;

    ~ffea_gpu_blob() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (auto &e : chunk_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        if (h_S) (void)hipHostFree(h_S);
        if (h_done) (void)hipHostFree(h_done);
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
    bool timing_on(int ph) const { return (timing >> ph & 1u) && !capturing; }
    void t_begin(int ph, hipEvent_t *out) {
        *out = nullptr;
        if (!timing_on(ph)) return;
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
    // CG phases are timed on every cg_sample-th iteration only (events between kernels
    // delay the next launch); pairs of iterations at or past the last one of the step
    // (launches that return at once) are dropped when collected
    bool t_cg(int ph, int it) const { return timing_on(ph) && it % cg_sample == 0; }
    void collect_timing(int last_iter = -1) {
        if (!timing) return;
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
        for (auto &ph : phase) {
            for (size_t i = 0; i + 1 < ph.ev.size(); i += 2) {
                const int it = ph.iter[i / 2];
                if (it >= 0 && last_iter >= 0 && it >= last_iter) continue;
                float ms = 0;
                FFEA_HIP_TRY(hipEventElapsedTime(&ms, ph.ev[i], ph.ev[i + 1]));
                ph.ms += ms;
                ph.launches++;
            }
            ph.ev.clear();
            ph.iter.clear();
        }
        ev_used = 0;
    }

    void build(const std::vector<BlobInput> &in);
    template <typename T>
    void upload_elements(const Tables &t);
    void set_row_blocks(int blocks);
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        graph_exec = nullptr;
    }
    RowArgs row_args() { return RowArgs{nchunks, cpb, chunks.p, seg.p, S.p, num_done.p}; }
    // assembly grid: a workgroup per slot and row block for small meshes, per row block
    // (its lanes walking their rows' slots) from 1024 row blocks on
    // lanes per row of the assembly: 1 once the rows alone fill 1024 workgroups, else 16
    int asm_lanes() const { return asm_lanes_set ? asm_lanes_set : (nrows > kFusedMaxRows ? kAsmStaged : 16); }
    int st_groups = 0;                  // groups of the LDS-staged assembly
    // launch-shape knobs the autotuner sets (results do not depend on them)
    int asm_lanes_set = 0;              // lanes per row of the assembly (1, 4 or 16), kAsmStaged; 0: by size
    int spmv_unroll = kMaxUnrolledRow;  // blocks per row the SpMV issues together (16 or 8)
    dim3 st_grid() const { return dim3((unsigned)((st_groups + 7) / 8 * 8)); }
    dim3 asm_grid() const {  // a multiple of 8 workgroups: each XCD takes one contiguous eighth of the rows
        long long g = ((long long)nrows * asm_lanes() + kBlock - 1) / kBlock;
        if (g >= 8 && g % 8) g += 8 - g % 8;
        return dim3((unsigned)g);
    }
    void launch_element(bool advance = true);
    void launch_assemble();
    void launch_assemble_gather();
    void prepare_ext();
    int step_mass();
    void set_solver(int kind, const double *rho);
    void mg_setup();
    void launch_mg_vcycle(double *Pout);
    void launch_mg_gauge();
    template <typename T>
    void set_material_t(const double *A, const double *B, const double *G, const double *E);
    int forward_checkpointed(int nsteps, int interval);
    int adjoint(const double *dLdx, double *dA, double *dB, double *dG, double *dE, double *dLdx0);
    template <typename T>
    void launch_adjoint_step(const double *pos_k, const double *vel_k, T *w);
    template <typename T, bool PROJ>
    void launch_gather(const T *force);
    void launch_cg_iteration(int it, bool spmv_only);
    template <typename T, bool PROJ>
    void launch_cg_fused_t(const T *val);
    void launch_cg_fused();
    void launch_cg_vec(int it);
    template <typename T, bool PROJ>
    void launch_cg_iteration_t(int it, bool spmv_only, const T *val);
    void launch_update();
    int run_cg_eager(int start_iter);
    int finish_step();
    int step_once();
    void capture_graph(int iters);
    void set_refinement(double tol);
    void refine_solution();
    void solver_matrix_at(const double *pos_nodes);
    int solver_solve(const double *pos_nodes, double *x);
    void solver_apply(const double *pos_nodes, const double *in, double *out);
    int check_status_after_forces();
};

template <typename T>
static void upload_converted(DevBuf<T> &buf, const std::vector<double> &h, hipStream_t s, size_t *total) {
    std::vector<T> t(h.size());
    for (size_t i = 0; i < h.size(); ++i) t[i] = (T)h[i];
    buf.alloc(t.size(), total);
    buf.upload(t, s);
    FFEA_HIP_TRY(hipStreamSynchronize(s));  // t goes out of scope
}

template <typename T>
void ffea_gpu_blob::upload_elements(const Tables &t) {
    DevBuf<T> *bj, *bv, *bG, *bE, *bA, *bB, *bf, *bg, *bval;
    if constexpr (std::is_same<T, double>::value) {
        bj = &d_jinv0; bv = &d_vol0; bG = &d_G; bE = &d_E; bA = &d_A; bB = &d_B; bf = &d_force; bg = &d_geo;
        bval = &d_val;
    } else {
        bj = &f_jinv0; bv = &f_vol0; bG = &f_G; bE = &f_E; bA = &f_A; bB = &f_B; bf = &f_force; bg = &f_geo;
        bval = &f_val;
    }
    if constexpr (!std::is_same<T, double>::value) {  // FP64 viscosities for the refinement's product
        A64.alloc(ne, &dev_bytes);
        A64.upload(t.A, stream);
        B64.alloc(ne, &dev_bytes);
        B64.upload(t.B, stream);
    }
    upload_converted(*bj, t.jinv0, stream, &dev_bytes);
    upload_converted(*bv, t.vol0, stream, &dev_bytes);
    upload_converted(*bA, t.A, stream, &dev_bytes);
    upload_converted(*bB, t.B, stream, &dev_bytes);
    upload_converted(*bG, t.G, stream, &dev_bytes);
    upload_converted(*bE, t.E, stream, &dev_bytes);
    bf->alloc((size_t)12 * ne, &dev_bytes);
    bf->zero(stream);
    // shape records: the volume is read for the previous volume of an inverted element;
    // A and B ride in each record for the assembly
    std::vector<double> g0((size_t)kGeo * ne, 0.0);
    for (int e = 0; e < ne; ++e) {
        g0[(size_t)kGeo * e + 9] = t.vol0[e];
        g0[(size_t)kGeo * e + 10] = t.A[e];
        g0[(size_t)kGeo * e + 11] = t.B[e];
    }
    upload_converted(*bg, g0, stream, &dev_bytes);
    bval->alloc((size_t)nslots * 9, &dev_bytes);
    bval->zero(stream);
}

void ffea_gpu_blob::build(const std::vector<BlobInput> &in) {
    Tables t = make_tables(in, mixed, prm);
    validate(t);
    nblob = t.nblob;
    nrows = t.nrows;
    ne = t.ne;
    nslices = t.nslices;
    nslots = t.nslots;
    nchunks = t.nchunks;
    max_blob_ne = t.max_blob_ne;
    active_blobs = t.active_blobs;
    have_ctf = t.have_ctf;
    have_fold = t.have_fold;
    any_project = t.any_project;
    meta = std::move(t.meta);
    grid_rows = (nrows + kBlock - 1) / kBlock;
    fused_cg = nblob > 0;
    for (const BlobMeta &m : meta) fused_cg = fused_cg && m.nl <= kFusedMaxRows;
    grid_elems = (ne + kBlock - 1) / kBlock;
    grid_blobs = (nblob + kBlock - 1) / kBlock;
    small_chunks = !t.chunks.empty();
    for (const Chunk &c : t.chunks) small_chunks = small_chunks && c.end - c.begin <= kMixedBlock;
    conn.alloc(t.conn.size(), &dev_bytes); conn.upload(t.conn, stream);
    pos.alloc(t.pos.size(), &dev_bytes); pos.upload(t.pos, stream);
    rest.alloc(t.pos.size(), &dev_bytes); rest.upload(t.pos, stream);
    vel.alloc(t.pos.size(), &dev_bytes); vel.zero(stream);
    last_det.alloc(t.det.size(), &dev_bytes); last_det.upload(t.det, stream);
    inc_ptr.alloc(t.ip.size(), &dev_bytes); inc_ptr.upload(t.ip, stream);
    inc.alloc(t.ic.size(), &dev_bytes); inc.upload(t.ic, stream);
    inc_pos.alloc(t.inc_pos.size(), &dev_bytes); inc_pos.upload(t.inc_pos, stream);
    pinned.alloc(t.pin.size(), &dev_bytes); pinned.upload(t.pin, stream);
    slice_w.alloc(t.sw.size(), &dev_bytes); slice_w.upload(t.sw, stream);
    slice_base.alloc(t.sb.size(), &dev_bytes); slice_base.upload(t.sb, stream);
    col.alloc(t.col.size(), &dev_bytes); col.upload(t.col, stream);
    cptr.alloc(t.cnt.size(), &dev_bytes); cptr.upload(t.cnt, stream);
    contrib.alloc(t.contrib.size(), &dev_bytes); contrib.upload(t.contrib, stream);
    {
        const StageTables st = mixed ? stage_tables(t, kStageRows<float>, kStageCap<float>)
                                     : stage_tables(t, kStageRows<double>, kStageCap<double>);
        st_groups = (int)st.gptr.size() - 1;
        st_ptr.alloc(st.gptr.size(), &dev_bytes); st_ptr.upload(st.gptr, stream);
        st_el.alloc(st.gel.size(), &dev_bytes); st_el.upload(st.gel, stream);
        st_loc.alloc(st.cloc.size(), &dev_bytes); st_loc.upload(st.cloc, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));  // the host tables go out of scope
    }
    diag_j.alloc(t.dj.size(), &dev_bytes); diag_j.upload(t.dj, stream);
    if (nblob > 1) {
        elem_blob.alloc(t.eblob.size(), &dev_bytes);
        elem_blob.upload(t.eblob, stream);
    }
    elem_off.alloc(t.eoff.size(), &dev_bytes); elem_off.upload(t.eoff, stream);
    project.alloc(t.proj.size(), &dev_bytes); project.upload(t.proj, stream);
    const size_t vlen = (size_t)nslices * 3 * kSlice;  // tiled node vectors (vi)
    if (mixed) {
        z32.alloc((size_t)nslices * 3 * kSlice, &dev_bytes);
        z32.zero(stream);
    }
    for (DevBuf<double> *buf : {&dinv, &rhs, &r, &z, &x, &q, &p}) {
        buf->alloc(vlen, &dev_bytes);
        buf->zero(stream);
    }
    const size_t nc = (size_t)std::max(1, nchunks);
    Pg.alloc(17 * nc, &dev_bytes); Pg.zero(stream);
    Pi.alloc(2 * nc, &dev_bytes); Pi.zero(stream);
    P1.alloc(13 * nc, &dev_bytes); P1.zero(stream);
    P2.alloc(2 * nc, &dev_bytes); P2.zero(stream);
    if (have_ctf) {
        ctf.alloc(t.ctf.size(), &dev_bytes);
        ctf.upload(t.ctf, stream);
    }
    if (have_fold) {
        fold_ptr.alloc(t.fp.size(), &dev_bytes);
        fold_ptr.upload(t.fp, stream);
        fold.alloc(t.fv.size(), &dev_bytes);
        fold.upload(t.fv, stream);
    }
    stokes = t.stokes;
    stokes_noise = t.stokes_noise;
    nnodes = t.nnodes;
    if (stokes) {
        drag.alloc(t.drag.size(), &dev_bytes);
        drag.upload(t.drag, stream);
    }
    dyn_fold = t.dyn_fold;
    if (dyn_fold) {
        node_off.alloc(t.node_off.size(), &dev_bytes); node_off.upload(t.node_off, stream);
        row_node.alloc(t.row_node.size(), &dev_bytes); row_node.upload(t.row_node, stream);
        fold_mid.alloc(t.fold_mid.size(), &dev_bytes); fold_mid.upload(t.fold_mid, stream);
        if (have_ctf) { node_ctf.alloc(t.node_ctf.size(), &dev_bytes); node_ctf.upload(t.node_ctf, stream); }
    }
    if (stokes_noise) {
        grid_nodes = (nnodes + kBlock - 1) / kBlock;
        node_sc.alloc(t.sc.size(), &dev_bytes); node_sc.upload(t.sc, stream);
        node_jump.alloc(t.node_jump.size(), &dev_bytes); node_jump.upload(t.node_jump, stream);
        njump_group.alloc(t.njump_group.size(), &dev_bytes); njump_group.upload(t.njump_group, stream);
        njump_lane.alloc(t.njump_lane.size(), &dev_bytes); njump_lane.upload(t.njump_lane, stream);
        noise.alloc((size_t)3 * nnodes, &dev_bytes); noise.zero(stream);
        if (nblob > 1) { node_blob.alloc(t.node_blob.size(), &dev_bytes); node_blob.upload(t.node_blob, stream); }
    }
    have_sf = t.have_sf;
    if (have_sf) {
        sf_nfaces = t.sf_nfaces;
        sf_nsets = t.sf_nsets;
        sf_nent = (int)t.sf_entry_face.size();
        sf_rows.alloc(t.sf_rows.size(), &dev_bytes); sf_rows.upload(t.sf_rows, stream);
        sf_force.alloc(t.sf_force.size(), &dev_bytes); sf_force.upload(t.sf_force, stream);
        sf_set_ptr.alloc(t.sf_set_ptr.size(), &dev_bytes); sf_set_ptr.upload(t.sf_set_ptr, stream);
        sf_entry_face.alloc(t.sf_entry_face.size(), &dev_bytes); sf_entry_face.upload(t.sf_entry_face, stream);
        sf_entry_set.alloc(t.sf_entry_set.size(), &dev_bytes); sf_entry_set.upload(t.sf_entry_set, stream);
        sf_fe_ptr.alloc(t.sf_fe_ptr.size(), &dev_bytes); sf_fe_ptr.upload(t.sf_fe_ptr, stream);
        sf_fe.alloc(t.sf_fe.size(), &dev_bytes); sf_fe.upload(t.sf_fe, stream);
        sf_nf_ptr.alloc(t.sf_nf_ptr.size(), &dev_bytes); sf_nf_ptr.upload(t.sf_nf_ptr, stream);
        sf_nf.alloc(t.sf_nf.size(), &dev_bytes); sf_nf.upload(t.sf_nf, stream);
        sf_area.alloc(sf_nent, &dev_bytes);
        sf_total.alloc(sf_nsets, &dev_bytes);
        sf_face_force.alloc((size_t)3 * sf_nfaces, &dev_bytes); sf_face_force.zero(stream);
    }
    have_dn = t.have_dn;
    have_rf = t.have_rf;
    if (have_dn) {
        dn_n = (int)t.dn_node.size();
        next = t.next;
        dn_node.alloc(t.dn_node.size(), &dev_bytes); dn_node.upload(t.dn_node, stream);
        dn_ext.alloc(t.dn_ext.size(), &dev_bytes); dn_ext.upload(t.dn_ext, stream);
        dn_row.alloc(t.dn_row.size(), &dev_bytes); dn_row.upload(t.dn_row, stream);
        dn_ends.alloc(t.dn_ends.size(), &dev_bytes); dn_ends.upload(t.dn_ends, stream);
        dn_lptr.alloc(t.dn_lptr.size(), &dev_bytes); dn_lptr.upload(t.dn_lptr, stream);
        dn_rptr.alloc(t.dn_rptr.size(), &dev_bytes); dn_rptr.upload(t.dn_rptr, stream);
        dn_lval.alloc(std::max<size_t>(1, t.dn_lval.size()), &dev_bytes);
        if (!t.dn_lval.empty()) dn_lval.upload(t.dn_lval, stream);
        rf_torque.alloc(std::max<size_t>(1, t.rf_torque.size()), &dev_bytes);
        if (!t.rf_torque.empty()) rf_torque.upload(t.rf_torque, stream);
        rf_axis.alloc(std::max<size_t>(1, t.rf_axis.size()), &dev_bytes);
        if (!t.rf_axis.empty()) rf_axis.upload(t.rf_axis, stream);
        rf_F.alloc(std::max<size_t>(1, t.rf_F.size()), &dev_bytes);
        if (!t.rf_F.empty()) rf_F.upload(t.rf_F, stream);
        ext_force.alloc((size_t)3 * std::max(1, next), &dev_bytes);
        ext_force.zero(stream);
        ext_slot.assign(meta[0].nt, -1);
        for (int i = 0; i < next; ++i) ext_slot[t.ext_nodes[i]] = i;
        ext_staged.assign((size_t)3 * next, 0.0);
    }
    jump_group.alloc(t.jump_group.size(), &dev_bytes); jump_group.upload(t.jump_group, stream);
    jump_lane.alloc(t.jump_lane.size(), &dev_bytes); jump_lane.upload(t.jump_lane, stream);
    step_jump.alloc(t.step_jump.size(), &dev_bytes); step_jump.upload(t.step_jump, stream);
    rng.alloc(t.rng.size(), &dev_bytes); rng.upload(t.rng, stream);
    std::vector<CgScalars> s0(nblob);
    for (int b = 0; b < nblob; ++b) {
        s0[b] = CgScalars{};
        s0[b].done = meta[b].all_pinned ? 1 : 0;
    }
    S.alloc(nblob, &dev_bytes); S.upload(s0, stream);
    chunks.alloc(t.chunks.size(), &dev_bytes); chunks.upload(t.chunks, stream);
    seg.alloc(t.seg.size(), &dev_bytes); seg.upload(t.seg, stream);
    centre.alloc(t.centre.size(), &dev_bytes); centre.upload(t.centre, stream);
    Gmat.alloc((size_t)36 * nblob, &dev_bytes); Gmat.zero(stream);
    num_done.alloc(1, &dev_bytes); num_done.zero(stream);
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    if (mixed) upload_elements<float>(t);
    else upload_elements<double>(t);
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    FFEA_HIP_TRY(hipHostMalloc(&h_S, sizeof(CgScalars) * nblob, hipHostMallocDefault));
    FFEA_HIP_TRY(hipHostMalloc(&h_done, sizeof(int) * 2, hipHostMallocDefault));
    for (auto &e : chunk_ev) FFEA_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // every buffer a kernel reads must exist
    const void *need[] = {conn.p, pos.p, rest.p, vel.p, last_det.p, inc_ptr.p, inc.p, pinned.p, slice_w.p, slice_base.p,
                          col.p, cptr.p, contrib.p, diag_j.p, elem_off.p, project.p, dinv.p, rhs.p, r.p, z.p, x.p,
                          q.p, p.p, Pg.p, Pi.p, P1.p, P2.p, jump_group.p, jump_lane.p, step_jump.p, rng.p, S.p,
                          chunks.p, seg.p, centre.p, Gmat.p, num_done.p,
                          mixed ? (const void *)f_jinv0.p : (const void *)d_jinv0.p,
                          mixed ? (const void *)f_vol0.p : (const void *)d_vol0.p,
                          mixed ? (const void *)f_A.p : (const void *)d_A.p,
                          mixed ? (const void *)f_B.p : (const void *)d_B.p,
                          mixed ? (const void *)f_G.p : (const void *)d_G.p,
                          mixed ? (const void *)f_E.p : (const void *)d_E.p,
                          mixed ? (const void *)f_force.p : (const void *)d_force.p,
                          mixed ? (const void *)f_geo.p : (const void *)d_geo.p,
                          mixed ? (const void *)f_val.p : (const void *)d_val.p};
    for (const void *ptr : need)
        if (!ptr) throw Invalid("internal error: a device buffer was not allocated");
    if ((nblob > 1 && !elem_blob.p) || (have_ctf && !ctf.p) || (have_fold && (!fold_ptr.p || !fold.p)) ||
        (stokes && !drag.p) || (dyn_fold && (!node_off.p || !row_node.p || !fold_mid.p || (have_ctf && !node_ctf.p))) ||
        (stokes_noise && (!node_sc.p || !node_jump.p || !njump_group.p || !njump_lane.p || !noise.p ||
                          (nblob > 1 && !node_blob.p))) ||
        (have_sf && (!sf_rows.p || !sf_force.p || !sf_set_ptr.p || !sf_entry_face.p || !sf_entry_set.p ||
                     !sf_fe_ptr.p || !sf_fe.p || !sf_nf_ptr.p || !sf_nf.p || !sf_area.p || !sf_total.p ||
                     !sf_face_force.p)) ||
        (have_dn && (!dn_node.p || !dn_ext.p || !dn_row.p || !dn_ends.p || !dn_lptr.p || !dn_rptr.p || !dn_lval.p ||
                     !rf_torque.p || !rf_axis.p || !rf_F.p || !ext_force.p || !ctf.p || (dyn_fold && !node_ctf.p))))
        throw Invalid("internal error: a device buffer was not allocated");
    // row-kernel grid: one block per chunk up to 1024 blocks (tunable, result-neutral)
    set_row_blocks(1024);
}

void ffea_gpu_blob::set_row_blocks(int blocks) {
    const int nc = std::max(1, nchunks);
    blocks = std::max(1, std::min(blocks, nc));
    cpb = (nc + blocks - 1) / blocks;
    nb = (nc + cpb - 1) / cpb;
    if (nb >= 8 && nb % 8) nb += 8 - nb % 8;  // the XCD mapping wants a multiple of 8 blocks
    drop_graph();                             // the graph holds the old grid
}

void ffea_gpu_blob::launch_element(bool advance) {
    hipEvent_t t0;
    t_begin(0, &t0);
    const int *eb = nblob > 1 ? elem_blob.p : nullptr;
    if (mixed) {
        ElemArgs<float> a{ne, conn.p, pos.p, f_jinv0.p, f_vol0.p, f_G.p, f_E.p, f_A.p, f_B.p, last_det.p, f_force.p,
                          f_geo.p, prm.calc_noise, prm.kT, prm.dt, eb, elem_off.p, jump_group.p, jump_lane.p,
                          rng.p, S.p, rest.p, mass ? vel.p : nullptr, inc_pos.p};
        hipLaunchKernelGGL(k_element<float>, dim3(grid_elems), dim3(kBlock), 0, stream, a);
    } else {
        ElemArgs<double> a{ne, conn.p, pos.p, d_jinv0.p, d_vol0.p, d_G.p, d_E.p, d_A.p, d_B.p, last_det.p, d_force.p,
                           d_geo.p, prm.calc_noise, prm.kT, prm.dt, eb, elem_off.p, jump_group.p, jump_lane.p,
                           rng.p, S.p, nullptr, mass ? vel.p : nullptr, inc_pos.p};  // FP64: the reference's arithmetic
        hipLaunchKernelGGL(k_element<double>, dim3(grid_elems), dim3(kBlock), 0, stream, a);
    }
    FFEA_HIP_TRY(hipGetLastError());
    t_end(0, t0);
    if (!advance) return;  // forces and shape data only (multigrid set-up)
    if (stokes_noise) {  // the node draws follow the element draws of the step
        NoiseArgs na{nnodes, nblob > 1 ? node_blob.p : nullptr, node_off.p, node_sc.p, rng.p, node_jump.p,
                     njump_group.p, njump_lane.p, noise.p};
        hipLaunchKernelGGL(k_node_noise, dim3(grid_nodes), dim3(kBlock), 0, stream, na);
    }
    // the pass consumed 7 draws per element (tetra_element_linear.cpp:416-426)
    if (prm.calc_noise) {
        hipLaunchKernelGGL(k_rng_advance, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, rng.p,
                           (const mrg::Jump *)step_jump.p);
        FFEA_HIP_TRY(hipGetLastError());
    }
}

static std::string inversion_message(int inv) {
    return inv == 1 ? std::string("1 element has inverted since the last step. Aborting simulation.\n")
                    : fmt("%d elements have inverted since the last step. Aborting simulation.\n", inv);
}

int ffea_gpu_blob::check_status_after_forces() {
    FFEA_HIP_TRY(hipMemcpyAsync(h_S, S.p, sizeof(CgScalars) * nblob, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    for (int b = 0; b < nblob; ++b) {
        if (meta[b].all_pinned || h_S[b].inverted <= 0) continue;
        err = (nblob > 1 ? fmt("blob %d: ", b) : std::string()) + inversion_message(h_S[b].inverted);
        return FFEA_GPU_ERR_INVERSION;
    }
    return FFEA_GPU_OK;
}

template <typename T, bool PROJ>
void ffea_gpu_blob::launch_gather(const T *force) {
    if (have_dn) {  // external forces, then Blob::apply_ctforces CTF 1 + 2 at this step's positions
        DynArgs da{dn_n, dn_node.p, dn_row.p, dn_ends.p, dn_ext.p, ext_force.p, dn_lptr.p, dn_lval.p, dn_rptr.p,
                   rf_axis.p, rf_F.p, rf_torque.p, pos.p, dyn_fold ? node_ctf.p : nullptr, ctf.p};
        hipLaunchKernelGGL(k_node_forces, dim3((dn_n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, da);
    }
    if (have_sf) {  // Blob::apply_ctforces CTF 3 at this step's positions
        SurfArgs sa{sf_nfaces, sf_nsets, sf_nent, sf_rows.p, pos.p, sf_force.p, sf_set_ptr.p, sf_entry_face.p,
                    sf_entry_set.p, sf_fe_ptr.p, sf_fe.p, sf_area.p, sf_total.p, sf_face_force.p};
        hipLaunchKernelGGL(k_surface_forces, dim3(1), dim3(kRowBlock), 0, stream, sa);
    }
    GatherArgs<T> g{row_args(), inc_ptr.p, force, have_ctf ? ctf.p : nullptr,
                    have_fold ? fold_ptr.p : nullptr, have_fold ? fold.p : nullptr,
                    stokes_noise ? noise.p : nullptr, dyn_fold ? row_node.p : nullptr,
                    dyn_fold ? fold_mid.p : nullptr, dyn_fold && have_ctf ? node_ctf.p : nullptr,
                    have_sf ? sf_nf_ptr.p : nullptr, have_sf ? sf_nf.p : nullptr, have_sf ? sf_face_force.p : nullptr,
                    pinned.p, dinv.p, pos.p, centre.p, rhs.p, r.p, z.p, x.p, PROJ ? Pg.p : Pi.p,
                    mass && stokes ? vel.p : nullptr, mass && stokes ? drag.p : nullptr, rhs_in};
    g.z32 = z32.p;
    g.binv = mg_smoother();
    hipLaunchKernelGGL((k_gather<T, PROJ>), dim3(nb), dim3(row_block<T>), 0, stream, g);
    if (PROJ) {
        StartArgs sa{row_args(), Pg.p, rhs.p, pos.p, dinv.p, centre.p, project.p, r.p, z.p, x.p, Pi.p, Gmat.p};
        sa.z32 = z32.p;
        sa.binv = mg_smoother();
        hipLaunchKernelGGL(k_cg_start, dim3(nb), dim3(kMixedBlock), 0, stream, sa);
    }
}

void ffea_gpu_blob::launch_assemble_gather() {
    launch_assemble();
    hipEvent_t t0;
    t_begin(2, &t0);
    if (mixed) {
        if (any_project) launch_gather<float, true>(f_force.p);
        else launch_gather<float, false>(f_force.p);
    } else {
        launch_gather<double, false>(d_force.p);
    }
    FFEA_HIP_TRY(hipGetLastError());
    if (mg_on()) launch_mg_vcycle(Pi.p);  // z = M r of the start; its r.z replaces row 0
    t_end(2, t0);
}

void ffea_gpu_blob::launch_assemble() {
    hipEvent_t t0;
    t_begin(1, &t0);
    if (mixed) {
        AsmArgs<float> a{nrows, ne, slice_w.p, slice_base.p, cptr.p, contrib.p, diag_j.p, f_geo.p,
                         f_val.p, dinv.p, stokes ? drag.p : nullptr, mg_hval.p, mg_hscale, mg_binv.p};
        if (asm_lanes() == kAsmStaged)
            hipLaunchKernelGGL(k_assemble_staged<float>, st_grid(), dim3(kSlice), 0, stream, a, st_ptr.p, st_el.p, st_loc.p);
        else if (asm_lanes() == 1) hipLaunchKernelGGL((k_assemble<float, 1>), asm_grid(), dim3(kBlock), 0, stream, a);
        else if (asm_lanes() == 4) hipLaunchKernelGGL((k_assemble<float, 4>), asm_grid(), dim3(kBlock), 0, stream, a);
        else hipLaunchKernelGGL((k_assemble<float, 16>), asm_grid(), dim3(kBlock), 0, stream, a);
    } else {
        AsmArgs<double> a{nrows, ne, slice_w.p, slice_base.p, cptr.p, contrib.p, diag_j.p, d_geo.p,
                          d_val.p, dinv.p, stokes ? drag.p : nullptr, mg_hval.p, mg_hscale, mg_binv.p};
        if (asm_lanes() == kAsmStaged)
            hipLaunchKernelGGL(k_assemble_staged<double>, st_grid(), dim3(kSlice), 0, stream, a, st_ptr.p, st_el.p,
                               st_loc.p);
        else if (asm_lanes() == 1) hipLaunchKernelGGL((k_assemble<double, 1>), asm_grid(), dim3(kBlock), 0, stream, a);
        else if (asm_lanes() == 4) hipLaunchKernelGGL((k_assemble<double, 4>), asm_grid(), dim3(kBlock), 0, stream, a);
        else hipLaunchKernelGGL((k_assemble<double, 16>), asm_grid(), dim3(kBlock), 0, stream, a);
    }
    FFEA_HIP_TRY(hipGetLastError());
    t_end(1, t0);
}

template <typename T, bool PROJ>
void ffea_gpu_blob::launch_cg_iteration_t(int it, bool spmv_only, const T *val) {
    hipEvent_t ts = nullptr;
    if (t_cg(3, it)) t_begin(3, &ts);
    SpmvArgs<T> sa{row_args(), it, prm.max_iterations_cg, cg_eps2(), zero_rhs_solved, slice_w.p, slice_base.p, col.p, val, z.p,
                   z32.p, pos.p, centre.p, p.p, q.p, it == 0 ? Pi.p : P2.p, P1.p};
    constexpr int U2 = kMaxUnrolledRow / 2;
    const bool half = spmv_unroll != kMaxUnrolledRow;
    if (std::is_same<T, double>::value && small_chunks) {
        if (half) hipLaunchKernelGGL((k_cg_spmv<T, PROJ, kMixedBlock, U2>), dim3(nb), dim3(kMixedBlock), 0, stream, sa);
        else hipLaunchKernelGGL((k_cg_spmv<T, PROJ, kMixedBlock>), dim3(nb), dim3(kMixedBlock), 0, stream, sa);
    } else {
        if (half) hipLaunchKernelGGL((k_cg_spmv<T, PROJ, row_block<T>, U2>), dim3(nb), dim3(row_block<T>), 0, stream, sa);
        else hipLaunchKernelGGL((k_cg_spmv<T, PROJ>), dim3(nb), dim3(row_block<T>), 0, stream, sa);
    }
    t_end(3, ts, it);
    if (!spmv_only) launch_cg_vec(it);
}

void ffea_gpu_blob::launch_cg_vec(int it) {
    hipEvent_t tv = nullptr;
    if (t_cg(4, it)) t_begin(4, &tv);
    VecArgs va{row_args(), it, p.p, q.p, dinv.p, pos.p, centre.p, Gmat.p, x.p, r.p, z.p, P1.p, P2.p, z32.p};
    va.binv = mg_smoother();
    if (mixed && any_project) {
        hipLaunchKernelGGL((k_cg_vec<true, kMixedBlock>), dim3(nb), dim3(kMixedBlock), 0, stream, va);
    } else {
        if (mixed || small_chunks)
            hipLaunchKernelGGL((k_cg_vec<false, kMixedBlock>), dim3(nb), dim3(kMixedBlock), 0, stream, va);
        else
            hipLaunchKernelGGL((k_cg_vec<false, kRowBlock>), dim3(nb), dim3(kRowBlock), 0, stream, va);
    }
    if (mg_on()) launch_mg_vcycle(P2.p + nchunks);  // z = M r; its r.z replaces row 1
    t_end(4, tv, it);
}

void ffea_gpu_blob::launch_cg_iteration(int it, bool spmv_only) {
    if (mixed) {
        if (any_project) launch_cg_iteration_t<float, true>(it, spmv_only, f_val.p);
        else launch_cg_iteration_t<float, false>(it, spmv_only, f_val.p);
    } else {
        launch_cg_iteration_t<double, false>(it, spmv_only, d_val.p);
    }
}

template <typename T, bool PROJ>
void ffea_gpu_blob::launch_cg_fused_t(const T *val) {
    SpmvArgs<T> sa{row_args(), 0, prm.max_iterations_cg, cg_eps2(), zero_rhs_solved, slice_w.p, slice_base.p, col.p, val, z.p,
                   z32.p, pos.p, centre.p, p.p, q.p, Pi.p, P1.p};
    VecArgs va{row_args(), 0, p.p, q.p, dinv.p, pos.p, centre.p, Gmat.p, x.p, r.p, z.p, P1.p, P2.p, z32.p};
    hipLaunchKernelGGL((k_cg_fused<T, PROJ>), dim3(nblob), dim3(row_block<T>), 0, stream, sa, va, (const double *)Pi.p);
    FFEA_HIP_TRY(hipGetLastError());
}

void ffea_gpu_blob::launch_cg_fused() {
    hipEvent_t ts = nullptr;
    if (t_cg(3, 0)) t_begin(3, &ts);
    if (mixed) {
        if (any_project) launch_cg_fused_t<float, true>(f_val.p);
        else launch_cg_fused_t<float, false>(f_val.p);
    } else {
        launch_cg_fused_t<double, false>(d_val.p);
    }
    t_end(3, ts, 0);
}

void ffea_gpu_blob::launch_update() {
    hipEvent_t t0;
    t_begin(5, &t0);
    if (mg_on() && mg_any_free) launch_mg_gauge();
    hipLaunchKernelGGL(k_update, dim3(nb), dim3(kRowBlock), 0, stream, row_args(), (const double *)x.p, vel.p, pos.p,
                       prm.dt);
    hipLaunchKernelGGL(k_mark_updated, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, S.p);
    FFEA_HIP_TRY(hipGetLastError());
    t_end(5, t0);
}

// ---------------------------------------------------------------- multigrid preconditioner
namespace {
// a block CSR level matrix in the sliced ELL of amg_kernels.h (FP32 values, -1 padding)
struct EllHost {
    std::vector<int> sw, col;
    std::vector<long long> sb;
    std::vector<_Float16> val;
};
// power of two s with max|v| s in [512, 1024): FP16 keeps 64x headroom above it
double fp16_scale(double maxabs) {
    if (!(maxabs > 0.0) || !std::isfinite(maxabs)) return 1.0;
    int e = 0;
    (void)std::frexp(maxabs, &e);
    return std::ldexp(1.0, 10 - e);
}
double max_abs(const std::vector<double> &v, int threads) {
    const int nb = (int)std::min<size_t>(v.size() / 9 + 1, (size_t)1 << 30);  // ranges of whole 9-value groups
    const std::vector<int> bounds = amg::split_rows(nb, threads);
    std::vector<double> part(bounds.size() - 1, 0.0);
    amg::parallel_ranges(bounds, [&](int b0, int b1, int t) {
        double m = 0.0;
        const size_t e = std::min(v.size(), (size_t)b1 * 9);
        for (size_t i = (size_t)b0 * 9; i < e; ++i) m = std::max(m, std::fabs(v[i]));
        part[t] = m;
    });
    double m = 0.0;
    for (double x : part) m = std::max(m, x);
    return m;
}
EllHost to_ell(const amg::BCsr &A, double scale, int threads) {
    EllHost E;
    const int ns = (A.n + 63) / 64, bb = A.br * A.bc;
    E.sw.assign(std::max(1, ns), 1);
    E.sb.assign(std::max(1, ns), 0);
    long long slots = 0;
    for (int s = 0; s < ns; ++s) {
        int w = 1;
        for (int i = s * 64; i < std::min(A.n, s * 64 + 64); ++i) w = std::max(w, A.rp[i + 1] - A.rp[i]);
        E.sw[s] = w;
        E.sb[s] = slots;
        slots += (long long)w * 64;
    }
    for (long long k = 0; k < A.nnzb(); ++k)
        if (A.ci[k] < 0 || A.ci[k] >= A.ncols) throw Invalid("internal error: multigrid block column out of range");
    E.col.assign(std::max<long long>(1, slots), -1);
    E.val.assign((size_t)std::max<long long>(1, slots) * bb, (_Float16)0.0f);
    amg::parallel_rows(ns, threads, [&](int s0, int s1, int) {  // whole slices per range
        for (int i = s0 * 64; i < std::min(A.n, s1 * 64); ++i) {
            const int s = i / 64, lane = i % 64;
            for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const long long slot = E.sb[s] + (long long)(k - A.rp[i]) * 64;
                E.col[slot + lane] = A.ci[k];
                for (int e = 0; e < bb; ++e)
                    E.val[ell_val_index(slot, bb, lane, e)] = (_Float16)(A.v[(size_t)k * bb + e] * scale);
            }
        }
    });
    return E;
}
}  // namespace

// An identity transfer between two copies of a coarse level (6x6 blocks): a blob whose own
// hierarchy is shallower than the ensemble's repeats its coarsest level, whose exact solve
// the padded levels pass straight through
static amg::BCsr identity6(int n) {
    amg::BCsr I;
    I.n = I.ncols = n;
    I.br = I.bc = 6;
    I.rp.resize(n + 1);
    I.ci.resize(n);
    I.v.assign((size_t)n * 36, 0.0);
    for (int i = 0; i < n; ++i) {
        I.rp[i] = i;
        I.ci[i] = i;
        for (int k = 0; k < 6; ++k) I.v[(size_t)i * 36 + 7 * k] = 1.0;
    }
    I.rp[n] = n;
    return I;
}

// blobs' matrices side by side: part q's rows at row_off[q], its columns shifted by
// col_off[q]; rows no part covers are empty
static amg::BCsr side_by_side(const std::vector<const amg::BCsr *> &part, const std::vector<int> &row_off,
                              const std::vector<int> &col_off, int nrows, int ncols) {
    amg::BCsr C;
    C.n = nrows;
    C.ncols = ncols;
    C.br = part[0]->br;
    C.bc = part[0]->bc;
    const int bb = C.br * C.bc;
    std::vector<int> cnt(nrows, 0);
    for (size_t q = 0; q < part.size(); ++q)
        for (int i = 0; i < part[q]->n; ++i) cnt[row_off[q] + i] = part[q]->rp[i + 1] - part[q]->rp[i];
    C.rp.assign(nrows + 1, 0);
    for (int i = 0; i < nrows; ++i) C.rp[i + 1] = C.rp[i] + cnt[i];
    C.ci.resize(C.rp[nrows]);
    C.v.resize((size_t)C.rp[nrows] * bb);
    for (size_t q = 0; q < part.size(); ++q) {
        const amg::BCsr &P = *part[q];
        for (int i = 0; i < P.n; ++i) {
            const int r = row_off[q] + i;
            for (int k = P.rp[i]; k < P.rp[i + 1]; ++k) {
                const int d = C.rp[r] + (k - P.rp[i]);
                C.ci[d] = P.ci[k] + col_off[q];
                std::copy(P.v.begin() + (size_t)k * bb, P.v.begin() + (size_t)(k + 1) * bb, C.v.begin() + (size_t)d * bb);
            }
        }
    }
    return C;
}

// Builds the multigrid hierarchy from the viscosity matrix at the current positions: an
// element pass (no stream advance) and an assembly on the device, the matrix read back,
// the set-up on the host (amg::build; an ensemble's blobs each their own, on up to 16
// threads), the levels uploaded. An ensemble's levels hold its blobs' levels side by side:
// a blob with fewer levels repeats its coarsest one behind identity transfers; each
// blob's smoother weights are folded into its rows of the stored diagonal (per row on
// level 0); each blob's coarsest level has its own dense inverse. The hierarchy is kept
// for the rest of the run; each step's CG still applies that step's own matrix and
// diagonal on level 0.
void ffea_gpu_blob::mg_setup() {
    for (const BlobMeta &m : meta)
        if (m.all_pinned) throw Invalid("multigrid preconditioner: a blob has every node pinned, nothing is solved");
    if (fused_cg)
        throw Invalid("multigrid preconditioner: a blob of up to 1024 corner nodes solves in one launch; "
                      "the preconditioner is for larger blobs");
    const auto t0 = std::chrono::steady_clock::now();
    launch_element(false);
    launch_assemble();
    hipLaunchKernelGGL(k_mark_updated, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, S.p);  // pass's inversion count
    FFEA_HIP_TRY(hipGetLastError());
    std::vector<int> h_sw(nslices), h_col(nslots);
    std::vector<long long> h_sb(nslices);
    std::vector<unsigned char> h_pin(nrows);
    std::vector<double> h_pos((size_t)3 * nrows);
    FFEA_HIP_TRY(hipMemcpyAsync(h_sw.data(), slice_w.p, sizeof(int) * nslices, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(h_sb.data(), slice_base.p, sizeof(long long) * nslices, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(h_col.data(), col.p, sizeof(int) * nslots, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(h_pin.data(), pinned.p, nrows, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(h_pos.data(), pos.p, sizeof(double) * 3 * nrows, hipMemcpyDeviceToHost, stream));
    std::vector<double> h_val((size_t)nslots * 9);
    if (mixed) {
        std::vector<float> f((size_t)nslots * 9);
        FFEA_HIP_TRY(hipMemcpyAsync(f.data(), f_val.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
        for (size_t i = 0; i < f.size(); ++i) h_val[i] = f[i];
    } else {
        FFEA_HIP_TRY(hipMemcpyAsync(h_val.data(), d_val.p, sizeof(double) * h_val.size(), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
    }
    // blob b's corner-node matrix in block CSR over its own rows (slice padding left out),
    // rows in parallel ranges on `threads` threads: a count pass, then a fill pass
    auto fine_of = [&](int b, int threads) {
        const BlobMeta &m = meta[b];
        const int nl = m.nl, ro = m.row_off;
        amg::BCsr fine;
        fine.n = fine.ncols = nl;
        fine.br = fine.bc = 3;
        fine.rp.assign(nl + 1, 0);
        auto value = [&](int s, int j, int lane, int k) {
            return h_val[mixed ? val_index<float>(h_sb[s], j, lane, k) : val_index<double>(h_sb[s], j, lane, k)];
        };
        auto live = [&](int s, int j, int lane) {
            bool nz = false;
            for (int k = 0; k < 9; ++k) nz |= value(s, j, lane, k) != 0.0;
            return nz;
        };
        const std::vector<int> bounds = amg::split_rows(nl, threads);
        amg::parallel_ranges(bounds, [&](int n0, int n1, int) {
            for (int n = n0; n < n1; ++n) {
                const int g = ro + n, s = g / kSlice, lane = g % kSlice;
                int cnt = 0;
                for (int j = 0; j < h_sw[s]; ++j) cnt += live(s, j, lane) ? 1 : 0;
                fine.rp[n + 1] = cnt;
            }
        });
        for (int n = 0; n < nl; ++n) fine.rp[n + 1] += fine.rp[n];
        fine.ci.resize(fine.rp[nl]);
        fine.v.resize((size_t)fine.rp[nl] * 9);
        std::atomic<int> bad{0};  // 1: a column outside the blob, 2: a repeated column
        amg::parallel_ranges(bounds, [&](int n0, int n1, int) {
            std::vector<std::pair<int, int>> ent;  // (local column, slot j)
            for (int n = n0; n < n1; ++n) {
                const int g = ro + n, s = g / kSlice, lane = g % kSlice;
                ent.clear();
                for (int j = 0; j < h_sw[s]; ++j) {
                    if (!live(s, j, lane)) continue;
                    const int c = h_col[h_sb[s] + (long long)j * kSlice + lane];
                    if (c < ro || c >= ro + nl) {
                        bad = 1;
                        return;
                    }
                    ent.push_back({c - ro, j});
                }
                std::sort(ent.begin(), ent.end());
                for (size_t t = 1; t < ent.size(); ++t)
                    if (ent[t].first == ent[t - 1].first) {
                        bad = 2;
                        return;
                    }
                int d = fine.rp[n];
                for (auto &e : ent) {
                    fine.ci[d] = e.first;
                    for (int k = 0; k < 9; ++k) fine.v[(size_t)d * 9 + k] = value(s, e.second, lane, k);
                    ++d;
                }
            }
        });
        if (bad == 1) throw Invalid("internal error: matrix column outside the blob");
        if (bad == 2) throw Invalid("internal error: repeated matrix column");
        return fine;
    };
    const int hw = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<amg::Hierarchy> H(nblob);
    double hscale_max = 0.0;  // the FP16 copy of the whole level-0 matrix shares one scale
    {
        std::vector<std::string> errors(nblob);
        std::vector<double> vmax(nblob, 0.0);
        std::atomic<int> next{0};
        auto work = [&]() {
            for (int b = next++; b < nblob; b = next++) {
                try {
                    const int threads = nblob == 1 ? hw : 1;
                    amg::BCsr fine = fine_of(b, threads);
                    vmax[b] = max_abs(fine.v, threads);
                    const amg::Options opt = mg_options(nblob, meta[b].nl, threads);
                    const int ro = meta[b].row_off;
                    H[b] = amg::build(std::move(fine), h_pos.data() + 3 * (size_t)ro, h_pin.data() + ro,
                                      meta[b].free_blob, opt);
                } catch (const std::runtime_error &e) {
                    errors[b] = (nblob > 1 ? fmt("blob %d: ", b) : std::string()) + e.what();
                } catch (const Invalid &e) {
                    errors[b] = e.msg;
                } catch (const std::bad_alloc &) {
                    errors[b] = "host allocation failed";
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < std::min(hw, nblob); ++t) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
        for (const auto &e : errors)
            if (!e.empty()) throw Invalid(e);
        for (double v : vmax) hscale_max = std::max(hscale_max, v);
    }
    h_val.clear();
    h_val.shrink_to_fit();
    int L = 0;
    for (const auto &h : H) L = std::max(L, (int)h.lv.size());
    for (int b = 0; b < nblob; ++b)
        if ((int)H[b].lv.size() < 2 && L > 1)
            throw Invalid(fmt("multigrid preconditioner: blob %d does not coarsen", b));
    if (L < 2 || L > 8) throw Invalid(fmt("multigrid preconditioner: %d levels (needs 2 to 8)", L));
    for (auto &h : H)  // pad shallower hierarchies: the coarsest level repeated behind identity transfers
        while ((int)h.lv.size() < L) {
            amg::Level &last = h.lv.back();
            const int n = last.A.n;
            amg::Level copy;
            copy.A = last.A;
            copy.dinv = last.dinv;
            copy.lmax = last.lmax;
            copy.omega = last.omega;
            last.P = identity6(n);
            last.R = identity6(n);
            h.lv.push_back(std::move(copy));
        }
    // row offsets of each blob on each level: level 0 is the handle's own rows
    std::vector<std::vector<int>> off(L, std::vector<int>(nblob, 0));
    std::vector<int> nlev(L, 0);
    nlev[0] = nrows;
    for (int b = 0; b < nblob; ++b) off[0][b] = meta[b].row_off;
    for (int l = 1; l < L; ++l)
        for (int b = 0; b < nblob; ++b) {
            off[l][b] = nlev[l];
            nlev[l] += H[b].lv[l].A.n;
        }
    // upload
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    mg.clear();
    mg_cinv.reset();
    mg_hval.reset();
    mg_binv.reset();
    for (DevBuf<double> *bf : {&mg_res0, &mg_xb0, &mg_rp, &mg_coef, &mg_omega0}) bf->reset();
    mg_xb32.reset();
    mg_res32.reset();
    for (DevBuf<int> *bf : {&mg_crow_blob, &mg_coff_row, &mg_cm, &mg_free}) bf->reset();
    mg_ccoff.reset();
    // one matrix of the handle's level: a single blob's own, or the blobs' side by side
    auto level_matrix = [&](const std::vector<const amg::BCsr *> &part, const std::vector<int> &row_off,
                            const std::vector<int> &col_off, int nr, int nc, amg::BCsr &keep) -> const amg::BCsr & {
        if (part.size() == 1 && row_off[0] == 0 && col_off[0] == 0 && part[0]->n == nr && part[0]->ncols == nc)
            return *part[0];
        keep = side_by_side(part, row_off, col_off, nr, nc);
        return keep;
    };
    auto up_ell = [&](MgEll &D, const amg::BCsr &A) {
        const double sc = fp16_scale(max_abs(A.v, hw));
        const EllHost E = to_ell(A, sc, hw);
        D.n = A.n;
        D.mscale = 1.0 / sc;
        D.sw.alloc(E.sw.size(), &dev_bytes); D.sw.upload(E.sw, stream);
        D.sb.alloc(E.sb.size(), &dev_bytes); D.sb.upload(E.sb, stream);
        D.col.alloc(E.col.size(), &dev_bytes); D.col.upload(E.col, stream);
        D.val.alloc(E.val.size(), &dev_bytes); D.val.upload(E.val, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));  // E goes out of scope
    };
    auto up_csr = [&](MgCsr &D, const amg::BCsr &A) {
        for (long long k = 0; k < A.nnzb(); ++k)
            if (A.ci[k] < 0 || A.ci[k] >= A.ncols) throw Invalid("internal error: multigrid block column out of range");
        // vector-major rows (k_mg_csr): vector t of block j of row i at (rp[i] BB) + (t W + j) V
        const int BB = A.br * A.bc, V = BB % 4 == 0 ? 4 : 2;
        if (BB % V) throw Invalid("internal error: multigrid block size");
        const double sc = fp16_scale(max_abs(A.v, hw));
        std::vector<_Float16> v(A.v.size());
        amg::parallel_rows(A.n, hw, [&](int i0, int i1, int) {
            for (int i = i0; i < i1; ++i) {
                const int r0 = A.rp[i], W = A.rp[i + 1] - r0;
                for (int j = 0; j < W; ++j)
                    for (int t = 0; t < BB / V; ++t)
                        for (int u = 0; u < V; ++u)
                            v[(size_t)r0 * BB + ((size_t)t * W + j) * V + u] =
                                (_Float16)(A.v[((size_t)r0 + j) * BB + t * V + u] * sc);
            }
        });
        D.n = A.n;
        D.mscale = 1.0 / sc;
        D.rp.alloc(A.rp.size(), &dev_bytes); D.rp.upload(A.rp, stream);
        D.ci.alloc(A.ci.size(), &dev_bytes); D.ci.upload(A.ci, stream);
        D.val.alloc(v.size(), &dev_bytes); D.val.upload(v, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));  // v goes out of scope
    };
    // the blob whose smoother weights the info reports: the largest
    int big = 0, total_nl = 0;
    for (const BlobMeta &m : meta) total_nl += m.nl;
    for (int b = 1; b < nblob; ++b)
        if (meta[b].nl > meta[big].nl) big = b;
    mg_info = ffea_gpu_multigrid_info{};
    mg_info.num_levels = L;
    for (int l = 0; l < L; ++l) {
        auto lev = std::make_unique<MgLevel>();
        lev->n = nlev[l];
        lev->omega = l == 0 ? H[0].lv[0].omega : 1.0;  // levels above 0: folded into dinv
        lev->lmax = H[0].lv[l].lmax;
        lev->fold = H[0].lv[l].omega;
        long long blocks = 0;
        for (int b = 0; b < nblob; ++b) blocks += H[b].lv[l].A.nnzb();
        if (l >= 1) {
            for (int b = 0; b < nblob; ++b)
                if (H[b].lv[l].A.br != 6 || H[b].lv[l].A.bc != 6)
                    throw Invalid("internal error: coarse blocks are not 6x6");
            std::vector<const amg::BCsr *> parts;
            for (int b = 0; b < nblob; ++b) parts.push_back(&H[b].lv[l].A);
            amg::BCsr keep;
            if (l < L - 1) up_csr(lev->A, level_matrix(parts, off[l], off[l], nlev[l], nlev[l], keep));
            const size_t nv = (size_t)nlev[l] * 6;
            std::vector<double> dv(nv, 0.0);
            for (int b = 0; b < nblob; ++b) {
                const amg::Level &hl = H[b].lv[l];
                for (size_t i = 0; i < hl.dinv.size(); ++i) dv[(size_t)off[l][b] * 6 + i] = hl.omega * hl.dinv[i];
            }
            lev->dinv.alloc(nv, &dev_bytes);
            lev->dinv.upload(dv, stream);
            FFEA_HIP_TRY(hipStreamSynchronize(stream));
            for (DevBuf<double> *bf : {&lev->b, &lev->xa, &lev->res, &lev->xb, &lev->e}) {
                bf->alloc(nv, &dev_bytes);
                bf->zero(stream);
            }
        }
        if (l < L - 1) {
            std::vector<const amg::BCsr *> pp, rr;
            long long tb = 0;
            for (int b = 0; b < nblob; ++b) {
                const amg::Level &hl = H[b].lv[l];
                if (hl.P.n != hl.A.n || hl.P.ncols != H[b].lv[l + 1].A.n || hl.R.n != hl.P.ncols ||
                    hl.R.ncols != hl.P.n)
                    throw Invalid("internal error: multigrid transfer shapes");
                pp.push_back(&hl.P);
                rr.push_back(&hl.R);
                tb += hl.P.nnzb();
            }
            amg::BCsr keep;
            up_ell(lev->P, level_matrix(pp, off[l], off[l + 1], nlev[l], nlev[l + 1], keep));
            up_csr(lev->R, level_matrix(rr, off[l + 1], off[l], nlev[l + 1], nlev[l], keep));
            mg_info.transfer_blocks[l] = tb;
        }
        mg_info.rows[l] = l == 0 ? total_nl : nlev[l];
        mg_info.block[l] = l == 0 ? 3 : 6;
        mg_info.blocks[l] = blocks;
        mg_info.omega[l] = H[big].lv[l].omega;
        mg_info.lmax[l] = H[big].lv[l].lmax;
        mg.push_back(std::move(lev));
    }
    // the coarsest level: each blob's dense inverse
    mg_nc = nlev[L - 1] * 6;
    {
        std::vector<float> ci;
        std::vector<int> row_blob, coff_row(nblob), cm(nblob);
        std::vector<long long> ccoff(nblob);
        for (int b = 0; b < nblob; ++b) {
            const int m = H[b].nc;
            if (m != H[b].lv[L - 1].A.n * 6 || (long long)H[b].cinv.size() != (long long)m * m)
                throw Invalid("internal error: coarsest inverse size");
            coff_row[b] = off[L - 1][b] * 6;
            cm[b] = m;
            ccoff[b] = (long long)ci.size();  // a multiple of 4 floats: every row starts 16-byte aligned
            const int ld = dense_ld(m);
            for (int i = 0; i < m; ++i) {
                ci.insert(ci.end(), H[b].cinv.begin() + (size_t)i * m, H[b].cinv.begin() + (size_t)(i + 1) * m);
                ci.insert(ci.end(), (size_t)(ld - m), 0.0f);
            }
            row_blob.insert(row_blob.end(), m, b);
        }
        mg_cinv.alloc(ci.size(), &dev_bytes);
        mg_cinv.upload(ci, stream);
        if (nblob > 1) {
            mg_crow_blob.alloc(row_blob.size(), &dev_bytes); mg_crow_blob.upload(row_blob, stream);
            mg_coff_row.alloc(nblob, &dev_bytes); mg_coff_row.upload(coff_row, stream);
            mg_cm.alloc(nblob, &dev_bytes); mg_cm.upload(cm, stream);
            mg_ccoff.alloc(nblob, &dev_bytes); mg_ccoff.upload(ccoff, stream);
        }
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
    }
    mg_info.coarsest_unknowns = mg_nc;
    // level 0: the FP16 copy of K (the assembly writes it), per-row smoother weights of an
    // ensemble, the rigid-mode gauge of free blobs
    mg_hscale = fp16_scale(hscale_max);
    mg_hval.alloc((size_t)nslots * 9, &dev_bytes);
    mg_binv.alloc((size_t)nrows * 6, &dev_bytes);
    mg_binv.zero(stream);
    launch_assemble();  // fills the FP16 copy and the block inverses at the current positions
    if (nblob > 1) {
        std::vector<double> om(nrows, 0.0);
        for (int b = 0; b < nblob; ++b)
            for (int n = 0; n < meta[b].nl; ++n) om[meta[b].row_off + n] = H[b].lv[0].omega;
        mg_omega0.alloc(nrows, &dev_bytes);
        mg_omega0.upload(om, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
    }
    mg_any_free = false;
    {
        std::vector<int> fr(nblob, 0);
        for (int b = 0; b < nblob; ++b) {
            fr[b] = meta[b].free_blob ? 1 : 0;
            mg_any_free |= meta[b].free_blob;
        }
        mg_free.alloc(nblob, &dev_bytes);
        mg_free.upload(fr, stream);
    }
    const size_t vlen = (size_t)nslices * 3 * kSlice;
    mg_res0.alloc(vlen, &dev_bytes); mg_res0.zero(stream);
    mg_xb0.alloc(vlen, &dev_bytes); mg_xb0.zero(stream);
    if (mixed) {
        mg_xb32.alloc(vlen, &dev_bytes);
        mg_xb32.zero(stream);
        mg_res32.alloc(vlen, &dev_bytes);
        mg_res32.zero(stream);
    }
    mg_rp.alloc((size_t)27 * std::max(1, nchunks), &dev_bytes); mg_rp.zero(stream);
    mg_coef.alloc((size_t)6 * nblob, &dev_bytes); mg_coef.zero(stream);
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    mg_info.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// z = M r (one V-cycle), r and z = D^-1 r given by the CG start or vector step; the
// chunk partials of r.z go to Pout[c]
void ffea_gpu_blob::launch_mg_vcycle(double *Pout) {
    const int L = (int)mg.size();
    auto grid = [](int n) { return dim3((unsigned)std::max(1, (n + kMgBlock - 1) / kMgBlock)); };
    MgLevel &f = *mg[0];
    const double hms = 1.0 / mg_hscale;
    // coarse kernels stop with blob 0's solve, an ensemble's with every blob's
    const MgStop S0{nblob == 1 ? (const CgScalars *)S.p : nullptr, num_done.p, active_blobs};
    const bool ens = nblob > 1;  // an ensemble's mostly 256-row chunks: 256-thread level-0 workgroups
    const double *om0 = nblob == 1 ? nullptr : (const double *)mg_omega0.p;
    // the FP32-mixed form's level-0 products gather the packed FP32 copies (z32, mg_xb32),
    // its first restriction mg_res32
    MgFineArgs<_Float16> fa{S.p, chunks.p, nchunks, slice_w.p, slice_base.p, col.p, mg_hval.p, hms, z.p, r.p,
                            dinv.p, f.omega, mg_res0.p, nullptr, om0, z32.p, mg_res32.p};
    if (mixed) {
        if (ens) hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_RESID, kMgBlock, float>), dim3(nchunks), dim3(kMgBlock), 0, stream, fa);
        else hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_RESID, kRowBlock, float>), dim3(nchunks), dim3(kRowBlock), 0, stream, fa);
    } else {
        if (ens) hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_RESID, kMgBlock>), dim3(nchunks), dim3(kMgBlock), 0, stream, fa);
        else hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_RESID>), dim3(nchunks), dim3(kRowBlock), 0, stream, fa);
    }
    auto wgrid = [](int n) { return dim3((unsigned)std::max(1, (n + kMgBlock / 64 - 1) / (kMgBlock / 64))); };
    // levels 1 .. L-2 take sweeps(l) damped-Jacobi sweeps before and after their coarse
    // correction (the first pre-sweep from zero is the restriction's y2); the sweeps
    // alternate between two of the level's vectors, and pre[l] / out[l] record where the
    // pre-smoothed iterate and the level's correction end up
    const int sweeps1 = nblob == 1 && L > 1 && mg[1]->n > kMgBigLevel1 ? 2 : 1;
    auto sweeps = [&](int l) { return l == 1 ? sweeps1 : kMgCoarserSweeps; };
    std::vector<double *> pre(L, nullptr), out(L, nullptr);
    // y = alpha x + beta xold + w (omega D^-1, folded into dinv) (b - A x)
    auto post_sweep = [&](MgLevel &lv, const double *x, double *y, double w = 1.0, double alpha = 1.0,
                          double beta = 0.0, const double *xold = nullptr) {
        MgCsrArgs po{S0, lv.n, lv.A.rp.p, lv.A.ci.p, lv.A.val.p, lv.A.mscale, x, lv.b.p, lv.dinv.p, lv.omega * w, y,
                     nullptr};
        po.alpha = alpha;
        po.beta = beta;
        po.xold = xold;
        hipLaunchKernelGGL((k_mg_csr<6, 6, MG_POST>), wgrid(lv.n), dim3(kMgBlock), 0, stream, po);
    };
    // level 1's degree-2 Chebyshev smoother (kMgChebyshev1): x1 = x0 + D^-1 (b - A x0) / theta,
    // x2 = x1 + rho1 rho0 (x1 - x0) + (2 rho1 / delta) D^-1 (b - A x1), theta and delta the
    // centre and half-width of the interval, rho0 = delta / theta, rho1 = 1 / (2 theta / delta
    // - rho0); before the correction x0 = 0, after it x0 is the prolonged iterate. The same
    // polynomial both times keeps M symmetric. Weights relative to the folded one:
    const bool cheb1 = kMgChebyshev1 && sweeps1 == 2;
    double cw1 = 1.0, ca2 = 1.0, cb2 = 0.0, cw2 = 1.0;
    if (cheb1) {
        const double hi = 1.1 * mg[1]->lmax, lo = hi / kMgChebRatio;
        const double th = 0.5 * (hi + lo), de = 0.5 * (hi - lo), sg = th / de;
        const double rho0 = 1.0 / sg, rho1 = 1.0 / (2.0 * sg - rho0);
        cw1 = 1.0 / (th * mg[1]->fold);
        ca2 = 1.0 + rho1 * rho0;
        cb2 = -rho1 * rho0;
        cw2 = 2.0 * rho1 / de / mg[1]->fold;
    }
    for (int l = 0; l + 1 < L; ++l) {  // down: restrict and pre-smooth, residual
        MgLevel &cur = *mg[l], &nx = *mg[l + 1];
        const bool coarsest = l + 1 == L - 1;
        MgCsrArgs ra{S0, nx.n, cur.R.rp.p, cur.R.ci.p, cur.R.val.p, cur.R.mscale, l == 0 ? mg_res0.p : cur.res.p,
                     nullptr, coarsest ? nullptr : nx.dinv.p, nx.omega * (l == 0 && cheb1 ? cw1 : 1.0), nx.b.p,
                     coarsest ? nullptr : nx.xa.p, l == 0 ? mg_res32.p : nullptr};
        if (l == 0) hipLaunchKernelGGL((k_mg_csr<6, 3, MG_RESTRICT>), wgrid(nx.n), dim3(kMgBlock), 0, stream, ra);
        else hipLaunchKernelGGL((k_mg_csr<6, 6, MG_RESTRICT>), wgrid(nx.n), dim3(kMgBlock), 0, stream, ra);
        if (!coarsest) {
            double *xp = nx.xa.p;
            for (int k = 1; k < sweeps(l + 1); ++k) {
                double *to = xp == nx.xa.p ? nx.e.p : nx.xa.p;  // e is free until the way up
                if (l == 0 && cheb1) post_sweep(nx, xp, to, cw2, ca2);  // x0 = 0: no xold term
                else post_sweep(nx, xp, to);
                xp = to;
            }
            pre[l + 1] = xp;
            MgCsrArgs rs{S0, nx.n, nx.A.rp.p, nx.A.ci.p, nx.A.val.p, nx.A.mscale, xp, nx.b.p, nullptr, 0.0,
                         nx.res.p, nullptr};
            hipLaunchKernelGGL((k_mg_csr<6, 6, MG_RESID>), wgrid(nx.n), dim3(kMgBlock), 0, stream, rs);
        }
    }
    MgLevel &c = *mg[L - 1];
    MgDenseArgs da{S0, mg_nc, mg_cinv.p, c.b.p, c.e.p, nblob == 1 ? nullptr : (const int *)mg_crow_blob.p,
                   nblob == 1 ? nullptr : (const int *)mg_coff_row.p, nblob == 1 ? nullptr : (const int *)mg_cm.p,
                   nblob == 1 ? nullptr : (const long long *)mg_ccoff.p};
    hipLaunchKernelGGL(k_mg_dense, wgrid(mg_nc), dim3(kMgBlock), 0, stream, da);
    out[L - 1] = c.e.p;
    for (int l = L - 2; l >= 1; --l) {  // up: prolong, post-smooth
        MgLevel &cur = *mg[l];
        MgArgs pa{S0, cur.n, cur.P.sw.p, cur.P.sb.p, cur.P.col.p, cur.P.val.p, cur.P.mscale, out[l + 1], nullptr,
                  pre[l], nullptr, 0.0, 1.0, cur.xb.p, nullptr};
        hipLaunchKernelGGL((k_mg_apply<6, 6, MG_PROLONG>), grid(cur.n), dim3(kMgBlock), 0, stream, pa);
        double *xp = cur.xb.p;
        for (int k = 0; k < sweeps(l); ++k) {
            double *to = xp == cur.xb.p ? cur.e.p : cur.xb.p;
            if (l == 1 && cheb1) {
                // the second step reads x0 (the prolonged iterate, in xb) at its own row and
                // overwrites it
                if (k == 0) post_sweep(cur, xp, to, cw1);
                else post_sweep(cur, xp, to, cw2, ca2, cb2, to);
            } else {
                post_sweep(cur, xp, to);
            }
            xp = to;
        }
        out[l] = xp;
    }
    MgArgs p0{S0, f.n, f.P.sw.p, f.P.sb.p, f.P.col.p, f.P.val.p, f.P.mscale, out[1], nullptr, z.p, nullptr,
              0.0, f.omega, mg_xb0.p, nullptr, om0, mg_xb32.p};
    hipLaunchKernelGGL((k_mg_apply<3, 6, MG_PROLONG>), grid(f.n), dim3(kMgBlock), 0, stream, p0);
    MgFineArgs<_Float16> fp{S.p, chunks.p, nchunks, slice_w.p, slice_base.p, col.p, mg_hval.p, hms, mg_xb0.p, r.p,
                            dinv.p, f.omega, z.p, Pout, om0, mg_xb32.p, z32.p, mg_binv.p};
    if (mixed) {
        if (ens) hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_POST, kMgBlock, float>), dim3(nchunks), dim3(kMgBlock), 0, stream, fp);
        else hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_POST, kRowBlock, float>), dim3(nchunks), dim3(kRowBlock), 0, stream, fp);
    } else {
        if (ens) hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_POST, kMgBlock>), dim3(nchunks), dim3(kMgBlock), 0, stream, fp);
        else hipLaunchKernelGGL((k_mg_fine<_Float16, MG_FINE_POST>), dim3(nchunks), dim3(kRowBlock), 0, stream, fp);
    }
    FFEA_HIP_TRY(hipGetLastError());
}

// free blobs' solutions in the Jacobi-PCG gauge, Z^T D x = 0, before the update
void ffea_gpu_blob::launch_mg_gauge() {
    MgRigidArgs ra{S.p, chunks.p, nchunks, x.p, dinv.p, pos.p, centre.p, mg_rp.p, mg_coef.p, x.p, seg.p, mg_free.p};
    hipLaunchKernelGGL(k_mg_rigid_sums, dim3(nchunks), dim3(kMgBlock), 0, stream, ra);
    hipLaunchKernelGGL(k_mg_rigid_coef, dim3(nblob), dim3(kMgBlock), 0, stream, ra);
    hipLaunchKernelGGL(k_mg_rigid_fix, dim3(nchunks), dim3(kRowBlock), 0, stream, ra);
    FFEA_HIP_TRY(hipGetLastError());
}

// CG iterations from start_iter on, queued in chunks. Iteration i is an SpMV step (which
// first tests each blob's iteration i-1) and a vector step. Kernels of blobs whose solve
// stopped return at once; the host stops queuing once a completed chunk reports every
// solved blob done. Returns the first iteration not queued.
int ffea_gpu_blob::run_cg_eager(int start_iter) {
    const int imax = prm.max_iterations_cg;
    if (fused_cg && start_iter == 0) {  // the whole solve in one launch
        launch_cg_fused();
        return imax + 1;
    }
    int launched = start_iter, chunk_idx = 0;
    int chunk = std::max(4, expect_iters - 2 - start_iter);
    bool stop = false;
    while (launched <= imax && !stop) {
        const int nit = std::min(chunk, imax + 1 - launched);
        // the SpMV step of iteration imax only tests iteration imax-1
        for (int k = 0; k < nit; ++k) launch_cg_iteration(launched + k, launched + k == imax);
        FFEA_HIP_TRY(hipGetLastError());
        launched += nit;
        const int slot = chunk_idx & 1;
        FFEA_HIP_TRY(hipMemcpyAsync(&h_done[slot], num_done.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipEventRecord(chunk_ev[slot], stream));
        if (chunk_idx > 0) {
            FFEA_HIP_TRY(hipEventSynchronize(chunk_ev[slot ^ 1]));
            if (h_done[slot ^ 1] >= active_blobs) stop = true;
        }
        ++chunk_idx;
        chunk = 4;
    }
    return launched;
}

// status of the step just queued (one synchronisation)
int ffea_gpu_blob::finish_step() {
    FFEA_HIP_TRY(hipMemcpyAsync(h_S, S.p, sizeof(CgScalars) * nblob, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    int max_it = 0;
    for (int b = 0; b < nblob; ++b) {
        if (meta[b].all_pinned) continue;
        meta[b].last_iters = h_S[b].iter;
        meta[b].total_iters += h_S[b].iter;
        max_it = std::max(max_it, h_S[b].iter);
    }
    collect_timing(max_it);
    if (max_it > 0) expect_iters = max_it;
    for (int b = 0; b < nblob; ++b) {
        if (meta[b].all_pinned) continue;
        const std::string who = nblob > 1 ? fmt("blob %d: ", b) : std::string();
        if (h_S[b].inverted_step > 0) {
            err = who + inversion_message(h_S[b].inverted_step);
            return FFEA_GPU_ERR_INVERSION;
        }
        if (!h_S[b].converged) {
            err = who + fmt("Conjugate gradient solver: Could not converge after %d iterations.\n\tEither epsilon or "
                            "max_iterations_cg are set too low, or something went wrong with the simulation.\n",
                            prm.max_iterations_cg);
            return FFEA_GPU_ERR_CG_NO_CONVERGE;
        }
    }
    return FFEA_GPU_OK;
}

// One step captured as a graph: element pass, stream advance, assembly, gather, `iters`
// CG iterations, the SpMV step that tests the last of them, update. Kernels of blobs
// whose solve stopped return at once; blobs still running after the graph continue
// eagerly. No memset or copy is captured: every node is a kernel.
void ffea_gpu_blob::capture_graph(int iters) {
    drop_graph();
    iters = std::min(iters, prm.max_iterations_cg);
    hipGraph_t g = nullptr;
    FFEA_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    try {
        launch_element();
        launch_assemble_gather();
        if (fused_cg) {
            launch_cg_fused();
            iters = prm.max_iterations_cg;
        } else {
            for (int it = 0; it <= iters; ++it) launch_cg_iteration(it, it == iters);
        }
        launch_update();
    } catch (...) {
        capturing = false;
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    capturing = false;
    FFEA_HIP_TRY(hipStreamEndCapture(stream, &g));
    const hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    FFEA_HIP_TRY(e);
    graph_iters = iters;
}

// External forces staged since the last force formation go to the device for the coming
// step; those of an earlier step are cleared (Blob::set_forces_to_zero each step,
// World.cpp:1862). Runs outside any graph.
void ffea_gpu_blob::prepare_ext() {
    if (!have_dn || next == 0) return;
    if (ext_dirty) {
        ext_force.upload(ext_staged, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));  // the copy has read ext_staged
        std::fill(ext_staged.begin(), ext_staged.end(), 0.0);
        ext_dirty = false;
        ext_on_device = true;
    } else if (ext_on_device) {
        ext_force.zero(stream);
        ext_on_device = false;
    }
}

// One inertial step of the lumped-mass solver: element pass (with the viscous term),
// gather (drag under calc_stokes, folds, pins), then a = f / m, v += a dt, x += v dt.
int ffea_gpu_blob::step_mass() {
    prepare_ext();
    if (mixed) launch_gather<float, false>(f_force.p);
    else launch_gather<double, false>(d_force.p);
    if (mass_cg) {  // ConjugateGradientSolver::solve with the constant mass matrix, then the update
        // its convergence test comes before each update: a zero right-hand side stops it at once
        // with a = 0 (ConjugateGradientSolver.cpp:194-200, 319-334)
        zero_rhs_solved = 1;
        run_cg_eager(0);
        zero_rhs_solved = 0;
        hipLaunchKernelGGL(k_mass_update_cg, dim3(nb), dim3(kRowBlock), 0, stream, row_args(), (const double *)x.p,
                           vel.p, pos.p, prm.dt);
        hipLaunchKernelGGL(k_mark_updated, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, S.p);
        FFEA_HIP_TRY(hipGetLastError());
        return finish_step();
    }
    hipLaunchKernelGGL(k_mass_update, dim3(nb), dim3(kRowBlock), 0, stream, row_args(), (const double *)rhs.p,
                       (const double *)inv_m.p, vel.p, pos.p, prm.dt);
    hipLaunchKernelGGL(k_mark_updated, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, S.p);
    FFEA_HIP_TRY(hipGetLastError());
    FFEA_HIP_TRY(hipMemcpyAsync(h_S, S.p, sizeof(CgScalars) * nblob, hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    for (int b = 0; b < nblob; ++b) {
        if (meta[b].all_pinned || h_S[b].inverted_step <= 0) continue;
        err = (nblob > 1 ? fmt("blob %d: ", b) : std::string()) + inversion_message(h_S[b].inverted_step);
        return FFEA_GPU_ERR_INVERSION;
    }
    return FFEA_GPU_OK;
}

void ffea_gpu_blob::set_solver(int kind, const double *rho_h) {
    if (kind != FFEA_GPU_SOLVER_NOMASS_CG) {  // every check before any state changes
        if (kind != FFEA_GPU_SOLVER_MASSLUMPED && kind != FFEA_GPU_SOLVER_ITERATIVE)
            throw Invalid(fmt("ffea_gpu_blob_set_solver: solver %d is not one this library runs (1 = ITERATIVE, 2 = "
                              "MASSLUMPED, 3 = NOMASS_CG)", kind));
        if (nblob != 1) throw Invalid("ffea_gpu_blob_set_solver: the inertial solvers need a single-blob handle");
        if (!rho_h) throw Invalid("ffea_gpu_blob_set_solver: the inertial solvers need the element densities");
        if (mg_on()) throw Invalid("ffea_gpu_blob_set_solver: switch the multigrid preconditioner off first");
        for (int e = 0; e < ne; ++e)
            if (!(rho_h[e] > 0.0) || !std::isfinite(rho_h[e]))
                throw Invalid(fmt("ffea_gpu_blob_set_solver: element %d has density %g", e, rho_h[e]));
    }
    if (mass_cg) any_project = project_saved;  // leaving the mass CG: the mass-less solver's projection again
    if (kind == FFEA_GPU_SOLVER_NOMASS_CG) {
        mass = mass_cg = false;
        inv_m.reset();
        rho_dev.reset();
        return;
    }
    drop_graph();
    run_steps = 0;  // the adjoint reverses mass-less runs only
    rho_dev.reset();
    rho_dev.alloc(ne, &dev_bytes);
    rho_dev.upload(rho_h, ne, stream);
    if (!inv_m.p) inv_m.alloc(nrows, &dev_bytes);
    if (mixed)
        hipLaunchKernelGGL(k_lumped_mass<float>, dim3(nb), dim3(kRowBlock), 0, stream, row_args(), inc_ptr.p, inc.p,
                           (const double *)rho_dev.p, (const float *)f_vol0.p, pinned.p, inv_m.p);
    else
        hipLaunchKernelGGL(k_lumped_mass<double>, dim3(nb), dim3(kRowBlock), 0, stream, row_args(), inc_ptr.p, inc.p,
                           (const double *)rho_dev.p, (const double *)d_vol0.p, pinned.p, inv_m.p);
    FFEA_HIP_TRY(hipGetLastError());
    mass_cg = kind == FFEA_GPU_SOLVER_ITERATIVE;
    if (mass_cg) {  // the consistent mass matrix, once: it depends on the rest state only
        if (mixed) {
            AsmArgs<float> am{nrows, ne, slice_w.p, slice_base.p, cptr.p, contrib.p, diag_j.p, f_geo.p, f_val.p, dinv.p,
                              nullptr, nullptr, 1.0};
            hipLaunchKernelGGL(k_assemble_mass<float>, dim3(grid_rows), dim3(kBlock), 0, stream, am,
                               (const double *)rho_dev.p, (const float *)f_vol0.p);
        } else {
            AsmArgs<double> am{nrows, ne, slice_w.p, slice_base.p, cptr.p, contrib.p, diag_j.p, d_geo.p, d_val.p,
                               dinv.p, nullptr, nullptr, 1.0};
            hipLaunchKernelGGL(k_assemble_mass<double>, dim3(grid_rows), dim3(kBlock), 0, stream, am,
                               (const double *)rho_dev.p, (const double *)d_vol0.p);
        }
        FFEA_HIP_TRY(hipGetLastError());
        project_saved = any_project;
        any_project = false;  // the mass matrix is nonsingular: no rigid-mode projection
    }
    FFEA_HIP_TRY(hipMemsetAsync(vel.p, 0, sizeof(double) * vel.n, stream));  // nodes start at rest
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    mass = true;
}

int ffea_gpu_blob::step_once() {
    if (active_blobs == 0) return FFEA_GPU_OK;
    if (mass) {
        launch_element();
        return step_mass();
    }
    prepare_ext();
    if (use_graph && expect_iters > 0 && refine_tol == 0.0) {
        // enough iterations to cover the last step's count with a margin, in steps of 16
        int want = std::min(prm.max_iterations_cg, ((expect_iters + expect_iters / 8 + 8 + 15) / 16) * 16);
        if (graph_fixed_iters > 0) want = std::min(prm.max_iterations_cg, graph_fixed_iters);
        if (!graph_exec || (!fused_cg && (want > graph_iters || want < graph_iters / 2))) capture_graph(want);
        FFEA_HIP_TRY(hipGraphLaunch(graph_exec, stream));
        FFEA_HIP_TRY(hipMemcpyAsync(&h_done[0], num_done.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
        if (h_done[0] < active_blobs && graph_iters < prm.max_iterations_cg) {
            // iteration graph_iters has had its SpMV step in the graph: finish it, go on
            launch_cg_vec(graph_iters);
            FFEA_HIP_TRY(hipGetLastError());
            run_cg_eager(graph_iters + 1);
            launch_update();
        }
        return finish_step();
    }
    launch_element();
    launch_assemble_gather();
    run_cg_eager(0);
    if (refine_tol > 0.0) refine_solution();
    launch_update();
    return finish_step();
}

// ---------------------------------------------------------------- FP64 refinement
// tol > 0 turns refinement on (FP32-mixed single-blob handles whose viscosity matrix is
// nonsingular: some nodes pinned or Stokes drag)
void ffea_gpu_blob::set_refinement(double tol) {
    if (!(tol >= 0.0) || !std::isfinite(tol)) throw Invalid("ffea_gpu_blob_set_refinement: tolerance must be >= 0");
    if (tol == 0.0) {
        refine_tol = 0.0;
        return;
    }
    if (!mixed) throw Invalid("ffea_gpu_blob_set_refinement: the FP64 form needs no refinement");
    if (nblob != 1) throw Invalid("ffea_gpu_blob_set_refinement: single-blob handles only");
    if (mass) throw Invalid("ffea_gpu_blob_set_refinement: the mass-less solver only");
    if (meta[0].free_blob || any_project)
        throw Invalid("ffea_gpu_blob_set_refinement: a free blob's viscosity matrix is singular; refinement needs "
                      "pinned nodes or Stokes drag");
    if (have_sf || have_dn)
        throw Invalid("ffea_gpu_blob_set_refinement: not with surface, rotational or external forces");
    const size_t vlen = (size_t)nslices * 3 * kSlice;
    if (!w64.p) w64.alloc((size_t)12 * ne, &dev_bytes);
    for (DevBuf<double> *bf : {&xacc, &bsave, &res64})
        if (!bf->p) bf->alloc(vlen, &dev_bytes);
    if (!Pr.p) Pr.alloc((size_t)2 * std::max(1, nchunks), &dev_bytes);
    drop_graph();  // refined steps run eagerly
    refine_tol = tol;
}

// After a converged solve of V x = b (x in x, b in rhs): sweeps of r = b - V64 x in FP64,
// an FP32 correction solve V d = r (the CG's own kernels, stopping at a fraction of r that
// still meets the tolerance overall), x += d, until |r| <= refine_tol |b|. Leaves x and rhs
// as a plain solve would, the step's CG iteration count including the correction solves.
void ffea_gpu_blob::refine_solution() {
    refine_sweeps_last = 0;
    FFEA_HIP_TRY(hipMemcpyAsync(h_S, S.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    if (!h_S[0].converged || h_S[0].inverted_step > 0) return;  // reported by the caller
    const CgScalars main = h_S[0];
    const size_t vlen = (size_t)nslices * 3 * kSlice;
    FFEA_HIP_TRY(hipMemcpyAsync(xacc.p, x.p, sizeof(double) * vlen, hipMemcpyDeviceToDevice, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(bsave.p, rhs.p, sizeof(double) * vlen, hipMemcpyDeviceToDevice, stream));
    const int nc = std::max(1, nchunks);
    h_pr.resize((size_t)2 * nc);
    int extra_iters = 0;
    constexpr int kMaxSweeps = 30;
    constexpr double kInner2 = 1e-6;  // a correction solve reduces its residual by 1e3 at least
    for (int sweep = 0;; ++sweep) {
        ViscApplyArgs va{ne, conn.p, pos.p, A64.p, B64.p, xacc.p, inc_pos.p, w64.p};
        hipLaunchKernelGGL(k_visc_apply64, dim3(grid_elems), dim3(kBlock), 0, stream, va);
        ResidArgs ra{row_args(), inc_ptr.p, w64.p, stokes ? drag.p : nullptr, pinned.p, bsave.p, xacc.p, res64.p, Pr.p};
        hipLaunchKernelGGL(k_resid64, dim3(nb), dim3(kRowBlock), 0, stream, ra);
        FFEA_HIP_TRY(hipGetLastError());
        FFEA_HIP_TRY(hipMemcpyAsync(h_pr.data(), Pr.p, sizeof(double) * h_pr.size(), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
        double r2 = 0.0, b2 = 0.0;
        for (int c = 0; c < nchunks; ++c) {
            r2 += h_pr[c];
            b2 += h_pr[(size_t)nchunks + c];
        }
        if (!std::isfinite(r2)) throw Invalid("FP64 refinement: the residual is not finite");
        if (r2 <= refine_tol * refine_tol * b2) break;
        if (sweep == kMaxSweeps)
            throw Invalid(fmt("FP64 refinement: |b - V x| / |b| = %.3e after %d correction solves (tolerance %.3e)",
                              std::sqrt(r2 / b2), kMaxSweeps, refine_tol));
        eps2_override = std::max(0.25 * refine_tol * refine_tol * b2 / r2, kInner2);
        rhs_in = res64.p;
        try {
            launch_gather<float, false>(f_force.p);  // the correction's start on r
            FFEA_HIP_TRY(hipGetLastError());
            if (mg_on()) launch_mg_vcycle(Pi.p);
            run_cg_eager(0);
        } catch (...) {
            rhs_in = nullptr;
            eps2_override = 0.0;
            throw;
        }
        rhs_in = nullptr;
        eps2_override = 0.0;
        FFEA_HIP_TRY(hipMemcpyAsync(h_S, S.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream));
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
        if (!h_S[0].converged)
            throw Invalid(fmt("FP64 refinement: a correction solve could not converge after %d iterations",
                              prm.max_iterations_cg));
        extra_iters += h_S[0].iter;
        const long long nv = (long long)vlen;
        hipLaunchKernelGGL(k_add_vec, dim3((unsigned)((nv + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, nv,
                           (const double *)x.p, xacc.p);
        ++refine_sweeps_last;
    }
    refine_sweeps_total += refine_sweeps_last;
    FFEA_HIP_TRY(hipMemcpyAsync(x.p, xacc.p, sizeof(double) * vlen, hipMemcpyDeviceToDevice, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(rhs.p, bsave.p, sizeof(double) * vlen, hipMemcpyDeviceToDevice, stream));
    h_S[0] = main;
    h_S[0].iter = main.iter + extra_iters;
    FFEA_HIP_TRY(hipMemcpyAsync(S.p, h_S, sizeof(CgScalars), hipMemcpyHostToDevice, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------- solver-only handles
// The viscosity matrix at the caller's node positions: NoMassCGSolver::solve builds it from
// the element matrices that Blob::update_internal_forces has just formed from those
// positions (NoMassCGSolver.cpp:158-159, tetra_element_linear.cpp:222-327), so the handle
// forms shape data and volumes from the same positions and assembles.
void ffea_gpu_blob::solver_matrix_at(const double *pos_nodes) {
    const BlobMeta &m = meta[0];
    std::vector<double> p((size_t)3 * m.nl);
    for (int n = 0; n < m.nl; ++n)
        for (int c = 0; c < 3; ++c) p[3 * n + c] = pos_nodes[3 * m.orig_of_lin[n] + c];
    FFEA_HIP_TRY(hipMemcpyAsync(pos.p, p.data(), sizeof(double) * p.size(), hipMemcpyHostToDevice, stream));
    launch_element(false);
    launch_assemble();
    FFEA_HIP_TRY(hipStreamSynchronize(stream));  // p goes out of scope
}

// NoMassCGSolver::solve (NoMassCGSolver.cpp:156-188): x holds the right-hand side on entry
// and the solution on return, [num_nodes][3]. Corner rows go through the handle's CG (its
// preconditioner, the reference's start and stopping test); pinned rows are identity rows
// there. Midside rows are identity rows that Blob::linearise_force hands over zero
// (Blob.cpp:781-799): a nonzero one is refused, and they return zero.
int ffea_gpu_blob::solver_solve(const double *pos_nodes, double *xio) {
    const BlobMeta &m = meta[0];
    for (int v = 0; v < m.nt; ++v)
        if (m.lin_of[v] < 0)
            for (int c = 0; c < 3; ++c)
                if (xio[3 * v + c] != 0.0)
                    throw Invalid(fmt("ffea_gpu_solver_solve: node %d is a midside node with a nonzero right-hand "
                                      "side (linearise_force leaves midside entries zero)", v));
    h_vec.assign((size_t)nslices * 3 * kSlice, 0.0);
    for (int n = 0; n < m.nl; ++n)
        for (int c = 0; c < 3; ++c) h_vec[vi(n, c)] = xio[3 * m.orig_of_lin[n] + c];
    FFEA_HIP_TRY(hipMemcpyAsync(r.p, h_vec.data(), sizeof(double) * h_vec.size(), hipMemcpyHostToDevice, stream));
    solver_matrix_at(pos_nodes);
    // r holds the right-hand side; the gather copies it into rhs, then starts the CG on it
    rhs_in = r.p;
    try {
        hipEvent_t t0;
        t_begin(2, &t0);
        if (mixed) {
            if (any_project) launch_gather<float, true>(f_force.p);
            else launch_gather<float, false>(f_force.p);
        } else {
            launch_gather<double, false>(d_force.p);
        }
        FFEA_HIP_TRY(hipGetLastError());
        if (mg_on()) launch_mg_vcycle(Pi.p);
        t_end(2, t0);
    } catch (...) {
        rhs_in = nullptr;
        throw;
    }
    rhs_in = nullptr;
    run_cg_eager(0);
    if (refine_tol > 0.0) refine_solution();
    if (mg_on() && mg_any_free) launch_mg_gauge();  // Jacobi-PCG's rigid-mode gauge
    hipLaunchKernelGGL(k_mark_updated, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, S.p);
    FFEA_HIP_TRY(hipGetLastError());
    FFEA_HIP_TRY(hipMemcpyAsync(h_vec.data(), x.p, sizeof(double) * h_vec.size(), hipMemcpyDeviceToHost, stream));
    const int rc = finish_step();
    if (rc != FFEA_GPU_OK) return rc;  // finish_step synchronised the stream
    std::fill(xio, xio + (size_t)3 * m.nt, 0.0);
    for (int n = 0; n < m.nl; ++n)
        for (int c = 0; c < 3; ++c) xio[3 * m.orig_of_lin[n] + c] = h_vec[vi(n, c)];
    return FFEA_GPU_OK;
}

// out = V in with the viscosity matrix solver_solve would use at these positions, every
// row of the reference's matrix included: corner rows from the block matrix, midside rows
// identity rows (with the drag of an unpinned node under calc_stokes,
// NoMassCGSolver.cpp:97-127)
void ffea_gpu_blob::solver_apply(const double *pos_nodes, const double *in, double *out) {
    const BlobMeta &m = meta[0];
    h_vec.assign((size_t)nslices * 3 * kSlice, 0.0);
    for (int n = 0; n < m.nl; ++n)
        for (int c = 0; c < 3; ++c) h_vec[vi(n, c)] = in[3 * m.orig_of_lin[n] + c];
    FFEA_HIP_TRY(hipMemcpyAsync(z.p, h_vec.data(), sizeof(double) * h_vec.size(), hipMemcpyHostToDevice, stream));
    solver_matrix_at(pos_nodes);
    if (mixed)
        hipLaunchKernelGGL(k_apply_matrix<float>, dim3(grid_rows), dim3(kBlock), 0, stream, nrows,
                           (const int *)slice_w.p, (const long long *)slice_base.p, (const int *)col.p,
                           (const float *)f_val.p, (const double *)z.p, q.p);
    else
        hipLaunchKernelGGL(k_apply_matrix<double>, dim3(grid_rows), dim3(kBlock), 0, stream, nrows,
                           (const int *)slice_w.p, (const long long *)slice_base.p, (const int *)col.p,
                           (const double *)d_val.p, (const double *)z.p, q.p);
    hipLaunchKernelGGL(k_mark_updated, dim3(grid_blobs), dim3(kBlock), 0, stream, nblob, S.p);
    FFEA_HIP_TRY(hipGetLastError());
    FFEA_HIP_TRY(hipMemcpyAsync(h_vec.data(), q.p, sizeof(double) * h_vec.size(), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    for (int v = 0; v < m.nt; ++v)
        for (int c = 0; c < 3; ++c) out[3 * v + c] = m.lin_of[v] >= 0 ? h_vec[vi(m.lin_of[v], c)] : mid_diag[v] * in[3 * v + c];
}

// ---------------------------------------------------------------- checkpointed run and reverse
template <typename T>
void ffea_gpu_blob::set_material_t(const double *A, const double *B, const double *G, const double *E) {
    DevBuf<T> *bA, *bB, *bG, *bE;
    if constexpr (std::is_same<T, double>::value) {
        bA = &d_A; bB = &d_B; bG = &d_G; bE = &d_E;
    } else {
        bA = &f_A; bB = &f_B; bG = &f_G; bE = &f_E;
    }
    const int ne0 = meta[0].ne;
    std::vector<T> t(ne0);
    const std::pair<DevBuf<T> *, const double *> items[4] = {{bA, A}, {bB, B}, {bG, G}, {bE, E}};
    for (const auto &it : items) {
        for (int e = 0; e < ne0; ++e) t[e] = (T)it.second[e];
        it.first->upload(t, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
    }
    if constexpr (!std::is_same<T, double>::value) {
        A64.upload(A, ne0, stream);
        B64.upload(B, ne0, stream);
        FFEA_HIP_TRY(hipStreamSynchronize(stream));
    }
    T *geo = std::is_same<T, double>::value ? (T *)d_geo.p : (T *)f_geo.p;
    hipLaunchKernelGGL(k_geo_material<T>, dim3((ne0 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, ne0,
                       (const T *)bA->p, (const T *)bB->p, geo);
    FFEA_HIP_TRY(hipGetLastError());
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
}

static void ensure(DevBuf<double> &b, size_t n, size_t *total) {
    if (b.p && b.n >= n) return;
    if (b.p) {
        (void)hipFree(b.p);
        *total -= b.n * sizeof(double);
        b.p = nullptr;
    }
    b.alloc(n, total);
}

// nsteps forward steps from the current state; the positions at every interval-th step
// are kept on the device for ffea_gpu_blob_adjoint
int ffea_gpu_blob::forward_checkpointed(int nsteps, int interval) {
    if (nblob != 1 || prm.calc_noise || meta[0].all_pinned)
        throw Invalid("a checkpointed run needs a single blob, calc_noise = 0 and unpinned nodes");
    if (have_sf || have_dn)
        throw Invalid("a checkpointed run needs forces that do not depend on positions: the adjoint does not "
                      "differentiate surface, rotational or external forces");
    if (mass) throw Invalid("a checkpointed run needs the mass-less solver");
    const size_t vec = (size_t)3 * nrows;
    const int nck = (nsteps + interval - 1) / interval;
    ensure(ckpt, vec * std::max(1, nck), &dev_bytes);
    ensure(x_end, vec, &dev_bytes);
    ensure(v_end, vec, &dev_bytes);
    run_steps = 0;
    for (int k = 0; k < nsteps; ++k) {
        if (k % interval == 0)
            FFEA_HIP_TRY(hipMemcpyAsync(ckpt.p + vec * (k / interval), pos.p, vec * sizeof(double),
                                        hipMemcpyDeviceToDevice, stream));
        const int rc = step_once();
        if (rc != FFEA_GPU_OK) return rc;
    }
    FFEA_HIP_TRY(hipMemcpyAsync(x_end.p, pos.p, vec * sizeof(double), hipMemcpyDeviceToDevice, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(v_end.p, vel.p, vec * sizeof(double), hipMemcpyDeviceToDevice, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    run_steps = nsteps;
    run_interval = interval;
    return FFEA_GPU_OK;
}

// one reverse step at positions pos_k with the step's velocity vel_k: K(x_k) mu = P lambda,
// then the element reverse and lambda += dt * gather(w)
template <typename T>
void ffea_gpu_blob::launch_adjoint_step(const double *pos_k, const double *vel_k, T *w) {
    const size_t vec = (size_t)3 * nrows;
    FFEA_HIP_TRY(hipMemcpyAsync(pos.p, pos_k, vec * sizeof(double), hipMemcpyDeviceToDevice, stream));
    launch_element();  // shape data of x_k for the operator
    launch_assemble();
    AdjStartArgs st{row_args(), lam.p, pinned.p, dinv.p, rhs.p, r.p, z.p, x.p, Pi.p, z32.p};
    st.binv = mg_smoother();
    hipLaunchKernelGGL(k_adjoint_start, dim3(nb), dim3(kRowBlock), 0, stream, st);
    FFEA_HIP_TRY(hipGetLastError());
    if (mg_on()) launch_mg_vcycle(Pi.p);  // z = M r of the start; its r.z replaces row 0
    zero_rhs_solved = 1;
    try {
        run_cg_eager(0);
    } catch (...) {
        zero_rhs_solved = 0;
        throw;
    }
    zero_rhs_solved = 0;
    FFEA_HIP_TRY(hipMemcpyAsync(h_S, S.p, sizeof(CgScalars), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    if (!h_S[0].converged)
        throw Invalid(fmt("adjoint solve: conjugate gradient could not converge after %d iterations",
                          prm.max_iterations_cg));
    if (refine_tol > 0.0) refine_solution();
    AdjArgs<T> a{ne, conn.p, pos.p, rest.p, x.p, vel_k, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, prm.dt, w,
                 grad.p};
    if constexpr (std::is_same<T, double>::value) {
        a.jinv0 = d_jinv0.p; a.vol0 = d_vol0.p; a.G = d_G.p; a.E = d_E.p; a.A = d_A.p; a.B = d_B.p;
    } else {
        a.jinv0 = f_jinv0.p; a.vol0 = f_vol0.p; a.G = f_G.p; a.E = f_E.p; a.A = f_A.p; a.B = f_B.p;
    }
    hipLaunchKernelGGL(k_element_adjoint<T>, dim3(grid_elems), dim3(kBlock), 0, stream, a);
    hipLaunchKernelGGL(k_adjoint_gather<T>, dim3(nb), dim3(kRowBlock), 0, stream, row_args(), (const uint32_t *)inc_ptr.p,
                       (const uint32_t *)inc.p, (const T *)w, prm.dt, lam.p);
    FFEA_HIP_TRY(hipGetLastError());
}

// reverse of the last checkpointed run: dL/dx at its end in, dL/d(A, B, G, E) per element
// and dL/dx at its start out; the blob is left in the state at the end of the run
int ffea_gpu_blob::adjoint(const double *dLdx, double *dA, double *dB, double *dG, double *dE, double *dLdx0) {
    if (run_steps <= 0) throw Invalid("no checkpointed run to reverse");
    const BlobMeta &m = meta[0];
    const size_t vec = (size_t)3 * nrows, vlen = (size_t)nslices * 3 * kSlice;
    ensure(lam, vlen, &dev_bytes);
    ensure(grad, (size_t)4 * ne, &dev_bytes);
    ensure(seg_x, vec * run_interval, &dev_bytes);
    ensure(seg_v, vec * run_interval, &dev_bytes);
    if (mixed && !f_w.p) f_w.alloc((size_t)12 * ne, &dev_bytes);
    if (!mixed && !d_w.p) d_w.alloc((size_t)12 * ne, &dev_bytes);
    // dL/dx on the corner rows: a midside node sits at its edge midpoint, so half of its
    // derivative goes to each edge end
    std::vector<double> l(vlen, 0.0);
    for (int v = 0; v < m.nt; ++v) {
        if (m.lin_of[v] >= 0) {
            for (int c = 0; c < 3; ++c) l[vi(m.lin_of[v], c)] += dLdx[3 * v + c];
        } else if (m.mid_a[v] >= 0) {
            for (int c = 0; c < 3; ++c) {
                l[vi(m.lin_of[m.mid_a[v]], c)] += 0.5 * dLdx[3 * v + c];
                l[vi(m.lin_of[m.mid_b[v]], c)] += 0.5 * dLdx[3 * v + c];
            }
        }
    }
    lam.upload(l, stream);
    grad.zero(stream);
    const bool keep_graph = use_graph;
    use_graph = false;  // recomputed steps run eagerly (identical results)
    try {
        const int nseg = (run_steps + run_interval - 1) / run_interval;
        for (int sg = nseg - 1; sg >= 0; --sg) {
            const int k0 = sg * run_interval, k1 = std::min(run_steps, k0 + run_interval);
            FFEA_HIP_TRY(hipMemcpyAsync(pos.p, ckpt.p + vec * sg, vec * sizeof(double), hipMemcpyDeviceToDevice, stream));
            for (int k = k0; k < k1; ++k) {
                FFEA_HIP_TRY(hipMemcpyAsync(seg_x.p + vec * (k - k0), pos.p, vec * sizeof(double),
                                            hipMemcpyDeviceToDevice, stream));
                const int rc = step_once();
                if (rc != FFEA_GPU_OK) throw Invalid("recomputed step failed: " + err);
                FFEA_HIP_TRY(hipMemcpyAsync(seg_v.p + vec * (k - k0), vel.p, vec * sizeof(double),
                                            hipMemcpyDeviceToDevice, stream));
            }
            for (int k = k1 - 1; k >= k0; --k) {
                if (mixed) launch_adjoint_step<float>(seg_x.p + vec * (k - k0), seg_v.p + vec * (k - k0), f_w.p);
                else launch_adjoint_step<double>(seg_x.p + vec * (k - k0), seg_v.p + vec * (k - k0), d_w.p);
            }
        }
    } catch (...) {
        use_graph = keep_graph;
        (void)hipMemcpyAsync(pos.p, x_end.p, vec * sizeof(double), hipMemcpyDeviceToDevice, stream);
        (void)hipMemcpyAsync(vel.p, v_end.p, vec * sizeof(double), hipMemcpyDeviceToDevice, stream);
        (void)hipStreamSynchronize(stream);
        throw;
    }
    use_graph = keep_graph;
    FFEA_HIP_TRY(hipMemcpyAsync(pos.p, x_end.p, vec * sizeof(double), hipMemcpyDeviceToDevice, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(vel.p, v_end.p, vec * sizeof(double), hipMemcpyDeviceToDevice, stream));
    std::vector<double> g((size_t)4 * ne);
    FFEA_HIP_TRY(hipMemcpyAsync(g.data(), grad.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipMemcpyAsync(l.data(), lam.p, l.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    FFEA_HIP_TRY(hipStreamSynchronize(stream));
    double *outs[4] = {dA, dB, dG, dE};
    for (int k = 0; k < 4; ++k)
        if (outs[k]) std::copy(g.begin() + (size_t)k * ne, g.begin() + (size_t)(k + 1) * ne, outs[k]);
    if (dLdx0) {
        // midside positions are re-derived from the corners at set-up: they carry no derivative
        std::fill(dLdx0, dLdx0 + (size_t)3 * m.nt, 0.0);
        for (int n = 0; n < m.nl; ++n)
            for (int c = 0; c < 3; ++c) dLdx0[3 * m.orig_of_lin[n] + c] = l[vi(n, c)];
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

int check_params(const char *who, const ffea_gpu_params *params) {
    const std::string w(who);
    if (params->precision != FFEA_GPU_FP64 && params->precision != FFEA_GPU_MIXED) {
        g_last_error = w + ": unknown precision";
        return FFEA_GPU_ERR_INVALID;
    }
    if (!(params->dt > 0) || params->max_iterations_cg <= 0 || !(params->epsilon2 >= 0)) {
        g_last_error = w + ": dt, epsilon2 and max_iterations_cg must be positive";
        return FFEA_GPU_ERR_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_last_error = w + ": no HIP device";
        return FFEA_GPU_ERR_NO_DEVICE;
    }
    if (params->device < 0 || params->device >= ndev) {
        g_last_error = w + ": device ordinal out of range";
        return FFEA_GPU_ERR_INVALID;
    }
    return FFEA_GPU_OK;
}

bool valid_blob_input(const BlobInput &b) {
    return b.pos && b.top && b.A && b.B && b.G && b.E && b.rng && b.nt > 0 && b.ne > 0 && b.npin >= 0 && b.nctf >= 0 &&
           (b.npin == 0 || b.pin) && (b.nctf == 0 || (b.ctf_node && b.ctf_force));
}

int create_handle(const char *who, const ffea_gpu_params *params, const std::vector<BlobInput> &in,
                  ffea_gpu_blob **out) {
    if (int rc = check_params(who, params)) return rc;
    std::unique_ptr<ffea_gpu_blob> b(new ffea_gpu_blob());
    b->prm = *params;
    b->device = params->device;
    b->mixed = params->precision == FFEA_GPU_MIXED;
    const int rc = guarded(b.get(), [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        FFEA_HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        (void)scratch_on(b->device);  // set-up: this device's scratch address, once per device
        b->build(in);
        return (int)FFEA_GPU_OK;
    });
    if (rc != FFEA_GPU_OK) {
        g_last_error = b->err;
        return rc;
    }
    *out = b.release();
    return FFEA_GPU_OK;
}

bool bad_blob(const ffea_gpu_blob *b, int blob) { return !b || blob < 0 || blob >= b->nblob; }

// per-blob inputs from the arrays of ffea_gpu_batch_create / ffea_gpu_check_layout
int batch_inputs(const char *who, int num_blobs, const int *num_nodes, const int *num_elements,
                 const double *const *pos, const int *const *top, const double *const *A, const double *const *B,
                 const double *const *G, const double *const *E, const int *num_pinned, const int *const *pinned,
                 const int *num_ctf, const int *const *ctf_node, const double *const *ctf_force,
                 const double *const *stokes_drag, bool calc_stokes, const uint32_t *rng_states,
                 std::vector<BlobInput> *out) {
    if (num_blobs <= 0 || !num_nodes || !num_elements || !pos || !top || !A || !B || !G || !E || !rng_states ||
        (calc_stokes && !stokes_drag)) {
        g_last_error = std::string(who) + ": invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    out->resize(num_blobs);
    for (int b = 0; b < num_blobs; ++b) {
        (*out)[b] = BlobInput{num_nodes[b], num_elements[b], pos[b], top[b], A[b], B[b], G[b], E[b],
                              num_pinned ? num_pinned[b] : 0, pinned ? pinned[b] : nullptr,
                              num_ctf ? num_ctf[b] : 0, ctf_node ? ctf_node[b] : nullptr,
                              ctf_force ? ctf_force[b] : nullptr, rng_states + 6 * b,
                              calc_stokes ? stokes_drag[b] : nullptr};
        if (!valid_blob_input((*out)[b]) || (calc_stokes && !(*out)[b].drag)) {
            g_last_error = fmt("%s: invalid argument for blob %d", who, b);
            return FFEA_GPU_ERR_INVALID;
        }
        if (!mrg::h_valid_state((*out)[b].rng)) {
            g_last_error = fmt("%s: invalid MRG32k3a state for blob %d", who, b);
            return FFEA_GPU_ERR_INVALID;
        }
    }
    return FFEA_GPU_OK;
}
}  // namespace

extern "C" {

int ffea_gpu_blob_create(const ffea_gpu_params *params, int num_nodes, int num_elements, const double *pos,
                         const int *top, const double *A, const double *B, const double *G, const double *E,
                         int num_pinned, const int *pinned, int num_ctf, const int *ctf_node,
                         const double *ctf_force, const double *stokes_drag, const uint32_t rng_state[6],
                         ffea_gpu_blob **out) {
    return ffea_gpu_blob_create_ex(params, num_nodes, num_elements, pos, top, A, B, G, E, num_pinned, pinned, num_ctf,
                                   ctf_node, ctf_force, stokes_drag, rng_state, nullptr, nullptr, 0, nullptr, out);
}

int ffea_gpu_blob_create_ex(const ffea_gpu_params *params, int num_nodes, int num_elements, const double *pos,
                            const int *top, const double *A, const double *B, const double *G, const double *E,
                            int num_pinned, const int *pinned, int num_ctf, const int *ctf_node,
                            const double *ctf_force, const double *stokes_drag, const uint32_t rng_state[6],
                            const ffea_gpu_surface_forces *surface_forces,
                            const ffea_gpu_rotational_forces *rotational_forces, int num_external_nodes,
                            const int *external_nodes, ffea_gpu_blob **out) {
    if (out) *out = nullptr;
    if (num_external_nodes < 0 || (num_external_nodes > 0 && !external_nodes)) {
        g_last_error = "ffea_gpu_blob_create: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    if (!params || !out) {
        g_last_error = "ffea_gpu_blob_create: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    // the drag is used only with calc_stokes = 1 (NoMassCGSolver.cpp:119-127, Blob.cpp:3172)
    const BlobInput bi{num_nodes,      num_elements, pos,       top, A, B, G, E, num_pinned, pinned, num_ctf, ctf_node,
                       ctf_force,      rng_state,    params->calc_stokes ? stokes_drag : nullptr, surface_forces,
                       rotational_forces, num_external_nodes, external_nodes};
    if (!valid_blob_input(bi)) {
        g_last_error = "ffea_gpu_blob_create: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    if (params->calc_stokes && !stokes_drag) {
        g_last_error = "ffea_gpu_blob_create: calc_stokes = 1 needs the per-node Stokes drag";
        return FFEA_GPU_ERR_INVALID;
    }
    if (!mrg::h_valid_state(rng_state)) {
        g_last_error = "ffea_gpu_blob_create: invalid MRG32k3a state";
        return FFEA_GPU_ERR_INVALID;
    }
    return create_handle("ffea_gpu_blob_create", params, {bi}, out);
}

int ffea_gpu_solver_create(const ffea_gpu_params *params, int num_nodes, int num_elements, const double *pos,
                           const int *top, const double *A, const double *B, int num_pinned, const int *pinned,
                           const double *stokes_drag, ffea_gpu_blob **out) {
    if (out) *out = nullptr;
    if (!params || !out || num_nodes <= 0 || num_elements <= 0 || !pos || !top || !A || !B || num_pinned < 0 ||
        (num_pinned > 0 && !pinned)) {
        g_last_error = "ffea_gpu_solver_create: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    if (params->calc_stokes && !stokes_drag) {
        g_last_error = "ffea_gpu_solver_create: calc_stokes = 1 needs the per-node Stokes drag";
        return FFEA_GPU_ERR_INVALID;
    }
    for (int i = 0; i < num_pinned; ++i)
        if (pinned[i] < 0 || pinned[i] >= num_nodes) {
            g_last_error = fmt("Error: Pinned node %d references an out of bounds node index.", i);
            return FFEA_GPU_ERR_INVALID;
        }
    // the solver needs the mesh, the viscosities, the pins and the drag only: no moduli, no
    // noise (the caller's element loop forms the forces)
    ffea_gpu_params prm = *params;
    prm.calc_noise = 0;
    const std::vector<double> zero((size_t)num_elements, 0.0);
    static const uint32_t kAnyState[6] = {12345u, 12345u, 12345u, 12345u, 12345u, 12345u};
    const BlobInput bi{num_nodes, num_elements, pos,     top,       A,       B,       zero.data(),
                       zero.data(), num_pinned, pinned,  0,         nullptr, nullptr, kAnyState,
                       prm.calc_stokes ? stokes_drag : nullptr};
    const int rc = create_handle("ffea_gpu_solver_create", &prm, {bi}, out);
    if (rc != FFEA_GPU_OK) return rc;
    ffea_gpu_blob *b = *out;
    b->solver_only = true;
    std::vector<unsigned char> pin(num_nodes, 0);
    for (int i = 0; i < num_pinned; ++i) pin[pinned[i]] = 1;
    b->mid_diag.assign(num_nodes, 1.0);
    for (int v = 0; v < num_nodes; ++v)
        if (prm.calc_stokes && !pin[v]) b->mid_diag[v] = 1.0 + stokes_drag[v];  // &one, then &stokes_drag
    return FFEA_GPU_OK;
}

int ffea_gpu_solver_solve(ffea_gpu_blob *b, const double *pos, double *x) {
    if (!b || !pos || !x) return FFEA_GPU_ERR_INVALID;
    if (!b->solver_only) {
        b->err = "ffea_gpu_solver_solve: not a handle from ffea_gpu_solver_create";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (b->active_blobs == 0) {  // every node pinned: identity rows only
            const BlobMeta &m = b->meta[0];
            for (int v = 0; v < m.nt; ++v)
                if (m.lin_of[v] < 0)
                    for (int c = 0; c < 3; ++c) x[3 * v + c] = 0.0;
            return (int)FFEA_GPU_OK;
        }
        return b->solver_solve(pos, x);
    });
}

int ffea_gpu_solver_apply_matrix(ffea_gpu_blob *b, const double *pos, const double *in, double *out) {
    if (!b || !pos || !in || !out) return FFEA_GPU_ERR_INVALID;
    if (!b->solver_only) {
        b->err = "ffea_gpu_solver_apply_matrix: not a handle from ffea_gpu_solver_create";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        b->solver_apply(pos, in, out);
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_create(const ffea_gpu_params *params, int num_blobs, const int *num_nodes,
                          const int *num_elements, const double *const *pos, const int *const *top,
                          const double *const *A, const double *const *B, const double *const *G,
                          const double *const *E, const int *num_pinned, const int *const *pinned,
                          const int *num_ctf, const int *const *ctf_node, const double *const *ctf_force,
                          const double *const *stokes_drag, const uint32_t *rng_states, ffea_gpu_blob **out) {
    if (out) *out = nullptr;
    if (!params || !out) {
        g_last_error = "ffea_gpu_batch_create: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    std::vector<BlobInput> in;
    if (int rc = batch_inputs("ffea_gpu_batch_create", num_blobs, num_nodes, num_elements, pos, top, A, B, G, E,
                              num_pinned, pinned, num_ctf, ctf_node, ctf_force, stokes_drag, params->calc_stokes != 0,
                              rng_states, &in))
        return rc;
    return create_handle("ffea_gpu_batch_create", params, in, out);
}

int ffea_gpu_check_layout(const ffea_gpu_params *params, int num_blobs, const int *num_nodes,
                          const int *num_elements, const double *const *pos, const int *const *top,
                          const double *const *A, const double *const *B, const double *const *G,
                          const double *const *E, const int *num_pinned, const int *const *pinned,
                          const int *num_ctf, const int *const *ctf_node, const double *const *ctf_force,
                          const double *const *stokes_drag, const uint32_t *rng_states, ffea_gpu_blob_info *info) {
    if (!params) {
        g_last_error = "ffea_gpu_check_layout: invalid argument";
        return FFEA_GPU_ERR_INVALID;
    }
    const int precision = params->precision;
    std::vector<BlobInput> in;
    if (int rc = batch_inputs("ffea_gpu_check_layout", num_blobs, num_nodes, num_elements, pos, top, A, B, G, E,
                              num_pinned, pinned, num_ctf, ctf_node, ctf_force, stokes_drag,
                              params->calc_stokes != 0, rng_states, &in))
        return rc;
    return guarded(nullptr, [&]() {
        const Tables t = make_tables(in, precision == FFEA_GPU_MIXED, *params);
        validate(t);
        if (precision == FFEA_GPU_MIXED) (void)stage_tables(t, kStageRows<float>, kStageCap<float>);
        else (void)stage_tables(t, kStageRows<double>, kStageCap<double>);
        if (info) {
            *info = ffea_gpu_blob_info{};
            for (const auto &m : t.meta) {
                info->num_nodes += m.nt;
                info->num_linear_nodes += m.nl;
                info->num_blocks += m.nblocks;
            }
            info->num_elements = t.ne;
            info->num_block_slots = t.nslots;
            info->precision = precision;
            info->num_blobs = t.nblob;
            info->num_chunks = t.nchunks;
        }
        return (int)FFEA_GPU_OK;
    });
}

void ffea_gpu_blob_destroy(ffea_gpu_blob *blob) {
    if (!blob) return;
    (void)hipSetDevice(blob->device);
    delete blob;
}

namespace {
// a solver-only handle holds the viscosity solve only
int solver_refuses(ffea_gpu_blob *b, const char *who) {
    b->err = std::string(who) + ": a handle from ffea_gpu_solver_create only solves (ffea_gpu_solver_solve)";
    return FFEA_GPU_ERR_INVALID;
}
}  // namespace

int ffea_gpu_blob_update_internal_forces(ffea_gpu_blob *b) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (b->solver_only) return solver_refuses(b, "ffea_gpu_blob_update_internal_forces");
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (b->active_blobs == 0) return (int)FFEA_GPU_OK;
        b->launch_element();
        return b->check_status_after_forces();
    });
}

int ffea_gpu_blob_update_positions(ffea_gpu_blob *b) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (b->solver_only) return solver_refuses(b, "ffea_gpu_blob_update_positions");
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (b->active_blobs == 0) return (int)FFEA_GPU_OK;
        if (b->mass) return b->step_mass();
        b->prepare_ext();
        b->launch_assemble_gather();
        b->run_cg_eager(0);
        if (b->refine_tol > 0.0) b->refine_solution();
        b->launch_update();
        return b->finish_step();
    });
}

int ffea_gpu_blob_step(ffea_gpu_blob *b, int nsteps) {
    if (!b || nsteps < 0) return FFEA_GPU_ERR_INVALID;
    if (b->solver_only) return solver_refuses(b, "ffea_gpu_blob_step");
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        for (int s = 0; s < nsteps; ++s) {
            const int rc = b->step_once();
            if (rc != FFEA_GPU_OK) return rc;
        }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_set_solver(ffea_gpu_blob *b, int solver, const double *rho) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (b->solver_only) return solver_refuses(b, "ffea_gpu_blob_set_solver");
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        b->set_solver(solver, rho);
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_add_node_forces(ffea_gpu_blob *b, int num, const int *node, const double *force) {
    if (!b || num < 0 || (num > 0 && (!node || !force))) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        if (num == 0) return (int)FFEA_GPU_OK;
        if (!b->have_dn || b->next == 0)
            throw Invalid("ffea_gpu_blob_add_node_forces: the handle was created without external force nodes");
        for (int i = 0; i < num; ++i)
            if (node[i] < 0 || node[i] >= (int)b->ext_slot.size() || b->ext_slot[node[i]] < 0)
                throw Invalid(fmt("ffea_gpu_blob_add_node_forces: node %d is not an external force node of this handle",
                                  node[i]));
        for (int i = 0; i < num; ++i) {  // Blob::add_force_to_node, in call order
            double *f = &b->ext_staged[3 * (size_t)b->ext_slot[node[i]]];
            for (int c = 0; c < 3; ++c) f[c] += force[3 * i + c];
        }
        b->ext_dirty = true;
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_node_positions(ffea_gpu_blob *b, int num, const int *node, double *out) {
    if (!b || num < 0 || (num > 0 && (!node || !out))) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        if (num == 0) return (int)FFEA_GPU_OK;
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const BlobMeta &m = b->meta[0];
        std::vector<int> rows((size_t)2 * num);
        for (int i = 0; i < num; ++i) {
            const int v = node[i];
            if (v < 0 || v >= m.nt) throw Invalid(fmt("ffea_gpu_blob_get_node_positions: node %d is out of bounds", v));
            int ra = -1, rb = -1;
            if (m.lin_of[v] >= 0) {
                ra = rb = m.row_off + m.lin_of[v];
            } else if (m.mid_a[v] >= 0) {
                ra = m.row_off + m.lin_of[m.mid_a[v]];
                rb = m.row_off + m.lin_of[m.mid_b[v]];
            }
            rows[2 * i] = ra;
            rows[2 * i + 1] = rb;
        }
        const size_t nr = rows.size();
        if (!b->take_rows.p || b->take_rows.n < nr) {
            b->take_rows.reset();
            b->take_rows.alloc(nr, &b->dev_bytes);
            b->take_out.reset();
            b->take_out.alloc(3 * nr, &b->dev_bytes);
        }
        b->take_rows.upload(rows, b->stream);
        hipLaunchKernelGGL(k_take_rows, dim3((unsigned)((nr + kBlock - 1) / kBlock)), dim3(kBlock), 0, b->stream,
                           (int)nr, b->take_rows.p, b->pos.p, b->take_out.p);
        FFEA_HIP_TRY(hipGetLastError());
        std::vector<double> g(3 * nr);
        FFEA_HIP_TRY(hipMemcpyAsync(g.data(), b->take_out.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost,
                                    b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        for (int i = 0; i < num; ++i) {
            const double *pa = &g[6 * (size_t)i], *pb = &g[6 * (size_t)i + 3];
            for (int c = 0; c < 3; ++c)  // tetra_element_linear::linearise_element for midside nodes
                out[3 * i + c] = rows[2 * i] == rows[2 * i + 1] ? pa[c] : .5 * (pa[c] + pb[c]);
        }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_get_positions(ffea_gpu_blob *b, int blob, double *out) {
    if (bad_blob(b, blob) || !out) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const BlobMeta &m = b->meta[blob];
        std::vector<double> p((size_t)3 * m.nl);
        FFEA_HIP_TRY(hipMemcpyAsync(p.data(), b->pos.p + 3 * (size_t)m.row_off, sizeof(double) * p.size(),
                                    hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        for (int n = 0; n < m.nl; ++n)
            for (int c = 0; c < 3; ++c) out[3 * m.orig_of_lin[n] + c] = p[3 * n + c];
        // tetra_element_linear::linearise_element (tetra_element_linear.cpp:553-577)
        for (int v = 0; v < m.nt; ++v)
            if (m.lin_of[v] < 0) {
                if (m.mid_a[v] < 0) {
                    out[3 * v] = out[3 * v + 1] = out[3 * v + 2] = 0.0;
                    continue;
                }
                const double *pa = &p[3 * m.lin_of[m.mid_a[v]]], *pb = &p[3 * m.lin_of[m.mid_b[v]]];
                for (int c = 0; c < 3; ++c) out[3 * v + c] = .5 * (pa[c] + pb[c]);
            }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_set_positions(ffea_gpu_blob *b, int blob, const double *in) {
    if (bad_blob(b, blob) || !in) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const BlobMeta &m = b->meta[blob];
        std::vector<double> p((size_t)3 * m.nl);
        for (int n = 0; n < m.nl; ++n)
            for (int c = 0; c < 3; ++c) p[3 * n + c] = in[3 * m.orig_of_lin[n] + c];
        FFEA_HIP_TRY(hipMemcpyAsync(b->pos.p + 3 * (size_t)m.row_off, p.data(), sizeof(double) * p.size(),
                                    hipMemcpyHostToDevice, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_set_velocities(ffea_gpu_blob *b, const double *in) {
    if (!b || !in) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const BlobMeta &m = b->meta[0];
        std::vector<double> v((size_t)3 * m.nl);
        for (int n = 0; n < m.nl; ++n)
            for (int c = 0; c < 3; ++c) v[3 * n + c] = in[3 * m.orig_of_lin[n] + c];
        FFEA_HIP_TRY(hipMemcpyAsync(b->vel.p + 3 * (size_t)m.row_off, v.data(), sizeof(double) * v.size(),
                                    hipMemcpyHostToDevice, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

static int get_node_vec(ffea_gpu_blob *b, int blob, const double *dev, double *out, bool tiled) {
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const BlobMeta &m = b->meta[blob];
        const size_t ns = (size_t)(m.nl + kSlice - 1) / kSlice;
        std::vector<double> p(tiled ? ns * 3 * kSlice : (size_t)3 * m.nl);
        const double *src = tiled ? dev + vi(m.row_off, 0) : dev + 3 * (size_t)m.row_off;
        FFEA_HIP_TRY(hipMemcpyAsync(p.data(), src, sizeof(double) * p.size(), hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        std::fill(out, out + (size_t)3 * m.nt, 0.0);
        for (int n = 0; n < m.nl; ++n)
            for (int c = 0; c < 3; ++c) out[3 * m.orig_of_lin[n] + c] = tiled ? p[vi(n, c)] : p[3 * n + c];
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_get_velocities(ffea_gpu_blob *b, int blob, double *out) {
    if (bad_blob(b, blob) || !out) return FFEA_GPU_ERR_INVALID;
    return get_node_vec(b, blob, b->vel.p, out, false);
}

int ffea_gpu_batch_get_rhs(ffea_gpu_blob *b, int blob, double *out) {
    if (bad_blob(b, blob) || !out) return FFEA_GPU_ERR_INVALID;
    return get_node_vec(b, blob, b->rhs.p, out, true);
}

int ffea_gpu_batch_get_element_forces(ffea_gpu_blob *b, int blob, double *out) {
    if (bad_blob(b, blob) || !out) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        // the blob's corner forces fill incidence positions [4 elem_off, 4 (elem_off + ne)) in
        // row order; inc_pos says where each (element, corner) sits
        const BlobMeta &m = b->meta[blob];
        const size_t n = (size_t)12 * m.ne, off = (size_t)12 * m.elem_off;
        std::vector<double> f(n);
        if (b->mixed) {
            std::vector<float> ff(n);
            FFEA_HIP_TRY(hipMemcpyAsync(ff.data(), b->f_force.p + off, sizeof(float) * n, hipMemcpyDeviceToHost,
                                        b->stream));
            FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
            for (size_t i = 0; i < n; ++i) f[i] = ff[i];
        } else {
            FFEA_HIP_TRY(hipMemcpyAsync(f.data(), b->d_force.p + off, sizeof(double) * n, hipMemcpyDeviceToHost,
                                        b->stream));
        }
        std::vector<uint32_t> ip((size_t)4 * m.ne);
        for (int q = 0; q < 4; ++q)
            FFEA_HIP_TRY(hipMemcpyAsync(ip.data() + (size_t)q * m.ne, b->inc_pos.p + (size_t)q * b->ne + m.elem_off,
                                        sizeof(uint32_t) * m.ne, hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        const size_t base = (size_t)4 * m.elem_off;
        for (int e = 0; e < m.ne; ++e)
            for (int q = 0; q < 4; ++q) {
                const size_t k = ip[(size_t)q * m.ne + e] - base;
                for (int i = 0; i < 3; ++i) out[12 * (size_t)e + 3 * q + i] = f[3 * k + i];
            }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_get_cg_iterations(ffea_gpu_blob *b, int *last, long long *total) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    for (int i = 0; i < b->nblob; ++i) {
        if (last) last[i] = b->meta[i].last_iters;
        if (total) total[i] = b->meta[i].total_iters;
    }
    return FFEA_GPU_OK;
}

int ffea_gpu_batch_get_rng_state(ffea_gpu_blob *b, int blob, uint32_t state[6]) {
    if (bad_blob(b, blob) || !state) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        FFEA_HIP_TRY(hipMemcpyAsync(state, b->rng.p + 6 * (size_t)blob, sizeof(uint32_t) * 6, hipMemcpyDeviceToHost,
                                    b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_set_rng_stream(ffea_gpu_blob *b, int blob, const uint32_t state[6], long long draws_per_step) {
    if (bad_blob(b, blob) || !state) return FFEA_GPU_ERR_INVALID;
    if (!mrg::h_valid_state(state)) {
        b->err = "invalid MRG32k3a state";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        FFEA_HIP_TRY(hipMemcpyAsync(b->rng.p + 6 * (size_t)blob, state, sizeof(uint32_t) * 6, hipMemcpyHostToDevice,
                                    b->stream));
        mrg::Jump j{};
        if (draws_per_step >= 0) {  // a graph reads the jump from device memory: it stays valid
            j = b->meta[blob].all_pinned ? mrg::h_identity() : mrg::h_pow((unsigned long long)draws_per_step);
            FFEA_HIP_TRY(hipMemcpyAsync(b->step_jump.p + blob, &j, sizeof(j), hipMemcpyHostToDevice, b->stream));
        }
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_set_node_draw_offset(ffea_gpu_blob *b, int blob, long long offset) {
    if (bad_blob(b, blob) || offset < 0) return FFEA_GPU_ERR_INVALID;
    if (!b->stokes_noise) {
        b->err = "node draws exist only with calc_stokes = 1 and calc_noise = 1";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const mrg::Jump j = mrg::h_pow((unsigned long long)offset);
        FFEA_HIP_TRY(hipMemcpyAsync(b->node_jump.p + blob, &j, sizeof(j), hipMemcpyHostToDevice, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_kinetic_energy(ffea_gpu_blob *b, double *energy) {
    if (!b || !energy) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        *energy = 0.0;
        if (!b->mass) return (int)FFEA_GPU_OK;  // the mass-less solver has none (Blob.cpp:1079)
        if (!b->energy_w.p) {
            b->energy_w.alloc(b->ne, &b->dev_bytes);
            b->energy_out.alloc(b->nblob, &b->dev_bytes);
            std::vector<int> cnt(b->nblob);
            for (int i = 0; i < b->nblob; ++i) cnt[i] = b->meta[i].ne;
            b->elem_cnt.alloc(b->nblob, &b->dev_bytes);
            b->elem_cnt.upload(cnt, b->stream);
        }
        if (b->mixed)
            hipLaunchKernelGGL(k_kinetic_energy<float>, dim3(b->grid_elems), dim3(kBlock), 0, b->stream, b->ne,
                               (const int *)b->conn.p, (const double *)b->vel.p, (const double *)b->rho_dev.p,
                               (const float *)b->f_vol0.p, b->energy_w.p);
        else
            hipLaunchKernelGGL(k_kinetic_energy<double>, dim3(b->grid_elems), dim3(kBlock), 0, b->stream, b->ne,
                               (const int *)b->conn.p, (const double *)b->vel.p, (const double *)b->rho_dev.p,
                               (const double *)b->d_vol0.p, b->energy_w.p);
        hipLaunchKernelGGL(k_blob_energy, dim3(b->nblob), dim3(kRowBlock), 0, b->stream, (const int *)b->elem_off.p,
                           (const int *)b->elem_cnt.p, (const double *)b->energy_w.p, b->energy_out.p);
        FFEA_HIP_TRY(hipGetLastError());
        FFEA_HIP_TRY(hipMemcpyAsync(energy, b->energy_out.p, sizeof(double), hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_batch_strain_energy(ffea_gpu_blob *b, double *energy) {
    if (!b || !energy) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (!b->energy_w.p) {
            b->energy_w.alloc(b->ne, &b->dev_bytes);
            b->energy_out.alloc(b->nblob, &b->dev_bytes);
            std::vector<int> cnt(b->nblob);
            for (int i = 0; i < b->nblob; ++i) cnt[i] = b->meta[i].ne;
            b->elem_cnt.alloc(b->nblob, &b->dev_bytes);
            b->elem_cnt.upload(cnt, b->stream);
        }
        if (b->mixed)
            hipLaunchKernelGGL(k_strain_energy<float>, dim3(b->grid_elems), dim3(kBlock), 0, b->stream, b->ne,
                               (const int *)b->conn.p, (const double *)b->pos.p, (const double *)b->rest.p,
                               (const float *)b->f_jinv0.p, (const float *)b->f_vol0.p, (const float *)b->f_G.p,
                               (const float *)b->f_E.p, b->energy_w.p);
        else
            hipLaunchKernelGGL(k_strain_energy<double>, dim3(b->grid_elems), dim3(kBlock), 0, b->stream, b->ne,
                               (const int *)b->conn.p, (const double *)b->pos.p, (const double *)b->rest.p,
                               (const double *)b->d_jinv0.p, (const double *)b->d_vol0.p, (const double *)b->d_G.p,
                               (const double *)b->d_E.p, b->energy_w.p);
        hipLaunchKernelGGL(k_blob_energy, dim3(b->nblob), dim3(kRowBlock), 0, b->stream, (const int *)b->elem_off.p,
                           (const int *)b->elem_cnt.p, (const double *)b->energy_w.p, b->energy_out.p);
        FFEA_HIP_TRY(hipGetLastError());
        FFEA_HIP_TRY(hipMemcpyAsync(energy, b->energy_out.p, sizeof(double) * b->nblob, hipMemcpyDeviceToHost,
                                    b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

// ffea_gpu_batch_strain_energy without its lazily allocated buffers: the per-element
// energies and the per-blob sums live in this device's `scratch` (through its looked-up
// address), every entry of which the kernels write before they read it
int ffea_gpu_batch_strain_energy_noalloc(ffea_gpu_blob *b, double *energy) {
    if (!b || !energy) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        const size_t need = (size_t)b->ne + (size_t)b->nblob;  // energies, then the per-blob sums
        if (need > kScratchDoubles)
            throw Invalid(fmt("ffea_gpu_batch_strain_energy_noalloc: %d elements and %d blobs need %zu doubles of "
                              "scratch, more than its %zu (64 MiB); use ffea_gpu_batch_strain_energy",
                              b->ne, b->nblob, need, kScratchDoubles));
        double *w = scratch_on(b->device);
        double *sums = w + b->ne;
        if (b->mixed)
            hipLaunchKernelGGL(k_strain_energy<float>, dim3(b->grid_elems), dim3(kBlock), 0, b->stream, b->ne,
                               (const int *)b->conn.p, (const double *)b->pos.p, (const double *)b->rest.p,
                               (const float *)b->f_jinv0.p, (const float *)b->f_vol0.p, (const float *)b->f_G.p,
                               (const float *)b->f_E.p, w);
        else
            hipLaunchKernelGGL(k_strain_energy<double>, dim3(b->grid_elems), dim3(kBlock), 0, b->stream, b->ne,
                               (const int *)b->conn.p, (const double *)b->pos.p, (const double *)b->rest.p,
                               (const double *)b->d_jinv0.p, (const double *)b->d_vol0.p, (const double *)b->d_G.p,
                               (const double *)b->d_E.p, w);
        hipLaunchKernelGGL(k_blob_energy_range, dim3(b->nblob), dim3(kRowBlock), 0, b->stream, b->nblob, b->ne,
                           (const int *)b->elem_off.p, (const double *)w, sums);
        FFEA_HIP_TRY(hipGetLastError());
        FFEA_HIP_TRY(hipMemcpyAsync(energy, sums, sizeof(double) * b->nblob, hipMemcpyDeviceToHost, b->stream));
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_positions(ffea_gpu_blob *b, double *out) { return ffea_gpu_batch_get_positions(b, 0, out); }
int ffea_gpu_blob_set_positions(ffea_gpu_blob *b, const double *in) { return ffea_gpu_batch_set_positions(b, 0, in); }
int ffea_gpu_blob_get_velocities(ffea_gpu_blob *b, double *out) { return ffea_gpu_batch_get_velocities(b, 0, out); }
int ffea_gpu_blob_get_rhs(ffea_gpu_blob *b, double *out) { return ffea_gpu_batch_get_rhs(b, 0, out); }
int ffea_gpu_blob_get_element_forces(ffea_gpu_blob *b, double *out) {
    return ffea_gpu_batch_get_element_forces(b, 0, out);
}

int ffea_gpu_blob_get_cg_iterations(ffea_gpu_blob *b, int *last, long long *total) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (last) *last = b->meta[0].last_iters;
    if (total) *total = b->meta[0].total_iters;
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_get_rng_state(ffea_gpu_blob *b, uint32_t state[6]) { return ffea_gpu_batch_get_rng_state(b, 0, state); }

int ffea_gpu_blob_set_rng_state(ffea_gpu_blob *b, const uint32_t state[6]) {
    return ffea_gpu_batch_set_rng_stream(b, 0, state, -1);
}

int ffea_gpu_blob_set_rng_stream(ffea_gpu_blob *b, const uint32_t state[6], long long draws_per_step) {
    if (draws_per_step < 0) return FFEA_GPU_ERR_INVALID;
    return ffea_gpu_batch_set_rng_stream(b, 0, state, draws_per_step);
}

int ffea_gpu_advance_rng_state(const uint32_t state[6], long long ndraws, uint32_t out[6]) {
    if (!state || !out || ndraws < 0 || !mrg::h_valid_state(state)) return FFEA_GPU_ERR_INVALID;
    uint32_t t[6];
    mrg::h_apply(mrg::h_pow((unsigned long long)ndraws), state, t);
    for (int i = 0; i < 6; ++i) out[i] = t[i];
    return FFEA_GPU_OK;
}

int ffea_gpu_rng_stream_state(const uint32_t package_seed[6], long long stream, uint32_t out[6]) {
    if (!package_seed || !out || stream < 0 || !mrg::h_valid_state(package_seed)) return FFEA_GPU_ERR_INVALID;
    uint32_t t[6];
    mrg::h_apply(mrg::h_stream_jump((unsigned long long)stream), package_seed, t);
    for (int i = 0; i < 6; ++i) out[i] = t[i];
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_get_info(ffea_gpu_blob *b, ffea_gpu_blob_info *info) {
    if (!b || !info) return FFEA_GPU_ERR_INVALID;
    *info = ffea_gpu_blob_info{};
    for (const auto &m : b->meta) {
        info->num_nodes += m.nt;
        info->num_linear_nodes += m.nl;
        info->num_blocks += m.nblocks;
    }
    info->num_elements = b->ne;
    info->num_block_slots = b->nslots;
    info->precision = b->prm.precision;
    info->device_bytes = (long long)b->dev_bytes;
    info->num_blobs = b->nblob;
    info->num_chunks = b->nchunks;
    info->row_blocks = b->nb;
    info->graph_iterations = b->graph_exec ? b->graph_iters : 0;
    info->single_launch_cg = b->fused_cg ? 1 : 0;
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_set_row_blocks(ffea_gpu_blob *b, int blocks) {
    if (!b || blocks <= 0) return FFEA_GPU_ERR_INVALID;
    b->set_row_blocks(blocks);
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_autotune(ffea_gpu_blob *b, int iterations, int *chosen_blocks) {
    if (!b || iterations <= 0) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (b->nchunks == 0) {
            if (chosen_blocks) *chosen_blocks = b->nb;
            return (int)FFEA_GPU_OK;
        }
        std::vector<int> cand;
        for (int c : {b->nchunks, 1024, 768, 512, 384, 256, 128, 64, 32})
            if (c <= b->nchunks && std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
        // time CG iterations with every blob running (the convergence test never passes)
        std::vector<CgScalars> run(b->nblob);
        for (int i = 0; i < b->nblob; ++i) {
            run[i] = CgScalars{};
            run[i].delta[0] = run[i].delta[1] = 1.0;
            run[i].f2 = 1.0;
            run[i].done = b->meta[i].all_pinned ? 1 : 0;
        }
        const ffea_gpu_params keep = b->prm;
        const unsigned keep_timing = b->timing;
        b->prm.epsilon2 = -1.0;
        b->prm.max_iterations_cg = 1 << 30;
        b->timing = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        double best = 1e300;
        int best_blocks = b->nb;
        try {
            FFEA_HIP_TRY(hipEventCreate(&e0));
            FFEA_HIP_TRY(hipEventCreate(&e1));
            for (int c : cand) {
                b->set_row_blocks(c);
                b->S.upload(run, b->stream);
                for (int it = 1; it <= 2; ++it) b->launch_cg_iteration(it, false);  // warm-up
                FFEA_HIP_TRY(hipEventRecord(e0, b->stream));
                for (int it = 3; it < 3 + iterations; ++it) b->launch_cg_iteration(it, false);
                FFEA_HIP_TRY(hipEventRecord(e1, b->stream));
                FFEA_HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                FFEA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    best_blocks = c;
                }
            }
        } catch (...) {
            b->prm = keep;
            b->timing = keep_timing;
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            throw;
        }
        b->set_row_blocks(best_blocks);
        // then the SpMV's row unroll and the assembly's lanes per row at that grid
        try {
            double best_u = 1e300;
            int unroll = b->spmv_unroll;
            for (int u : {kMaxUnrolledRow, kMaxUnrolledRow / 2}) {  // the SpMV step alone, twice as often
                b->spmv_unroll = u;
                b->S.upload(run, b->stream);
                for (int it = 1; it <= 2; ++it) b->launch_cg_iteration(it, true);
                FFEA_HIP_TRY(hipEventRecord(e0, b->stream));
                for (int it = 3; it < 3 + 2 * iterations; ++it) b->launch_cg_iteration(it, true);
                FFEA_HIP_TRY(hipEventRecord(e1, b->stream));
                FFEA_HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                FFEA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best_u) {
                    best_u = ms;
                    unroll = u;
                }
            }
            b->spmv_unroll = unroll;
            // the assembly rewrites the matrix of the last element pass, the same values each time
            double best_a = 1e300;
            int lanes = b->asm_lanes();
            for (int l : {1, 4, 16, kAsmStaged}) {
                b->asm_lanes_set = l;
                b->launch_assemble();
                FFEA_HIP_TRY(hipEventRecord(e0, b->stream));
                for (int r = 0; r < 3; ++r) b->launch_assemble();
                FFEA_HIP_TRY(hipEventRecord(e1, b->stream));
                FFEA_HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                FFEA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best_a) {
                    best_a = ms;
                    lanes = l;
                }
            }
            b->asm_lanes_set = lanes;
        } catch (...) {
            b->prm = keep;
            b->timing = keep_timing;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->prm = keep;
        b->timing = keep_timing;
        b->drop_graph();  // a captured step holds the old launch shape
        // the CG vectors are rebuilt by the next gather; restore the solver scalars
        std::vector<CgScalars> s0(b->nblob);
        for (int i = 0; i < b->nblob; ++i) {
            s0[i] = CgScalars{};
            s0[i].done = b->meta[i].all_pinned ? 1 : 0;
        }
        b->S.upload(s0, b->stream);
        FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
        if (chosen_blocks) *chosen_blocks = b->nb;
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_set_material(ffea_gpu_blob *b, const double *A, const double *B, const double *G,
                                const double *E) {
    if (!b || !A || !B || !G || !E) return FFEA_GPU_ERR_INVALID;
    if (b->nblob != 1) {
        b->err = "ffea_gpu_blob_set_material: single-blob handles only";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        if (b->mixed) b->set_material_t<float>(A, B, G, E);
        else b->set_material_t<double>(A, B, G, E);
        b->run_steps = 0;  // a stored run was made with the earlier material: nothing left to reverse
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_run_checkpointed(ffea_gpu_blob *b, int nsteps, int interval) {
    if (!b || nsteps <= 0 || interval <= 0) return FFEA_GPU_ERR_INVALID;
    if (b->solver_only) return solver_refuses(b, "ffea_gpu_blob_run_checkpointed");
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        return b->forward_checkpointed(nsteps, interval);
    });
}

int ffea_gpu_blob_adjoint(ffea_gpu_blob *b, const double *dL_dpos, double *dL_dA, double *dL_dB, double *dL_dG,
                          double *dL_dE, double *dL_dpos0) {
    if (!b || !dL_dpos) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        return b->adjoint(dL_dpos, dL_dA, dL_dB, dL_dG, dL_dE, dL_dpos0);
    });
}

int ffea_gpu_blob_set_refinement(ffea_gpu_blob *b, double tolerance) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        b->set_refinement(tolerance);
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_refinement(ffea_gpu_blob *b, int *last_sweeps, long long *total_sweeps) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (last_sweeps) *last_sweeps = b->refine_sweeps_last;
    if (total_sweeps) *total_sweeps = b->refine_sweeps_total;
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_get_launch_shape(ffea_gpu_blob *b, int *row_blocks, int *spmv_unroll, int *assembly_lanes) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (row_blocks) *row_blocks = b->nb;
    if (spmv_unroll) *spmv_unroll = b->spmv_unroll;
    if (assembly_lanes) *assembly_lanes = b->asm_lanes();
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_set_launch_shape(ffea_gpu_blob *b, int row_blocks, int spmv_unroll, int assembly_lanes) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if ((spmv_unroll && spmv_unroll != kMaxUnrolledRow && spmv_unroll != kMaxUnrolledRow / 2) ||
        (assembly_lanes && assembly_lanes != 1 && assembly_lanes != 4 && assembly_lanes != 16 &&
         assembly_lanes != kAsmStaged) ||
        row_blocks < 0) {
        b->err = "ffea_gpu_blob_set_launch_shape: row_blocks >= 0, spmv_unroll 16 or 8, assembly_lanes 1, 4, 16 or -1 "
                 "(records staged in LDS) (0: unchanged)";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        if (row_blocks) b->set_row_blocks(row_blocks);
        if (spmv_unroll) b->spmv_unroll = spmv_unroll;
        if (assembly_lanes) b->asm_lanes_set = assembly_lanes;
        b->drop_graph();
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_use_graph(ffea_gpu_blob *b, int enable, int iterations) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    b->use_graph = enable != 0;
    if (iterations != b->graph_fixed_iters) b->drop_graph();
    b->graph_fixed_iters = std::max(0, iterations);
    if (!b->use_graph) b->drop_graph();
    return FFEA_GPU_OK;
}

int ffea_gpu_blob_set_preconditioner(ffea_gpu_blob *b, int kind) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    if (kind != FFEA_GPU_PRECOND_JACOBI && kind != FFEA_GPU_PRECOND_MULTIGRID) {
        b->err = "unknown preconditioner";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(b, [&]() {
        FFEA_HIP_TRY(hipSetDevice(b->device));
        b->drop_graph();  // a captured step holds the other preconditioner's launches
        if (kind == FFEA_GPU_PRECOND_JACOBI) {
            FFEA_HIP_TRY(hipStreamSynchronize(b->stream));
            b->precond = FFEA_GPU_PRECOND_JACOBI;
            b->mg.clear();
            b->mg_cinv.reset();
            b->mg_hval.reset();
            b->mg_binv.reset();
            for (auto *bf : {&b->mg_res0, &b->mg_xb0, &b->mg_rp, &b->mg_coef, &b->mg_omega0}) bf->reset();
            b->mg_xb32.reset();
            b->mg_res32.reset();
            for (auto *bf : {&b->mg_crow_blob, &b->mg_coff_row, &b->mg_cm, &b->mg_free}) bf->reset();
            b->mg_ccoff.reset();
            b->mg_any_free = false;
            b->mg_info = ffea_gpu_multigrid_info{};
            return (int)FFEA_GPU_OK;
        }
        if (b->mass) throw Invalid("the multigrid preconditioner is for the mass-less solver");
        b->precond = FFEA_GPU_PRECOND_JACOBI;  // until the hierarchy is complete
        b->mg_setup();
        b->precond = FFEA_GPU_PRECOND_MULTIGRID;
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_get_multigrid_info(ffea_gpu_blob *b, ffea_gpu_multigrid_info *info) {
    if (!b || !info) return FFEA_GPU_ERR_INVALID;
    *info = b->precond == FFEA_GPU_PRECOND_MULTIGRID ? b->mg_info : ffea_gpu_multigrid_info{};
    return FFEA_GPU_OK;
}

int ffea_gpu_multigrid_inspect(int num_rows, const int *rowptr, const int *col, const double *val, const double *pos,
                               const unsigned char *pinned, int free_blob, ffea_gpu_multigrid_info *info,
                               int *p_rowptr, int *p_col, double *p_val, int *a_rowptr, int *a_col, double *a_val) {
    if (num_rows <= 0 || !rowptr || !col || !val || !pos || !info) {
        g_last_error = "ffea_gpu_multigrid_inspect: missing argument";
        return FFEA_GPU_ERR_INVALID;
    }
    return guarded(nullptr, [&]() {
        if (rowptr[0] != 0) throw Invalid("ffea_gpu_multigrid_inspect: rowptr[0] must be 0");
        for (int i = 0; i < num_rows; ++i)
            if (rowptr[i + 1] < rowptr[i]) throw Invalid("ffea_gpu_multigrid_inspect: rowptr must not decrease");
        amg::BCsr A;
        A.n = A.ncols = num_rows;
        A.br = A.bc = 3;
        A.rp.assign(rowptr, rowptr + num_rows + 1);
        A.ci.assign(col, col + rowptr[num_rows]);
        A.v.assign(val, val + (size_t)rowptr[num_rows] * 9);
        // the hierarchy a single-blob handle builds for this matrix
        const amg::Options opt =
            mg_options(1, num_rows, std::max(1, std::min(16, (int)std::thread::hardware_concurrency())));
        const auto t0 = std::chrono::steady_clock::now();
        amg::Hierarchy H;
        try {
            H = amg::build(A, pos, pinned, free_blob != 0, opt);
        } catch (const std::runtime_error &e) {
            throw Invalid(e.what());
        }
        const int L = (int)H.lv.size();
        if (L > FFEA_GPU_MG_MAX_LEVELS) throw Invalid("ffea_gpu_multigrid_inspect: too many levels");
        *info = ffea_gpu_multigrid_info{};
        info->num_levels = L;
        for (int l = 0; l < L; ++l) {
            info->rows[l] = H.lv[l].A.n;
            info->block[l] = H.lv[l].A.br;
            info->blocks[l] = H.lv[l].A.nnzb();
            info->omega[l] = H.lv[l].omega;
            info->lmax[l] = H.lv[l].lmax;
            if (l + 1 < L) info->transfer_blocks[l] = H.lv[l].P.nnzb();
        }
        info->coarsest_unknowns = H.nc;
        info->setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (L >= 2 && p_rowptr && p_col && p_val) {
            const amg::BCsr &P = H.lv[0].P;
            std::copy(P.rp.begin(), P.rp.end(), p_rowptr);
            std::copy(P.ci.begin(), P.ci.end(), p_col);
            std::copy(P.v.begin(), P.v.end(), p_val);
        }
        if (L >= 2 && a_rowptr && a_col && a_val) {
            const amg::BCsr &C = H.lv[1].A;
            std::copy(C.rp.begin(), C.rp.end(), a_rowptr);
            std::copy(C.ci.begin(), C.ci.end(), a_col);
            std::copy(C.v.begin(), C.v.end(), a_val);
        }
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_blob_set_timing(ffea_gpu_blob *b, unsigned phase_mask) {
    if (!b) return FFEA_GPU_ERR_INVALID;
    return guarded(b, [&]() {
        b->collect_timing();
        b->timing = phase_mask & ((1u << FFEA_GPU_NUM_PHASES) - 1u);
        for (auto &ph : b->phase) {
            ph.ms = 0;
            ph.launches = 0;
            ph.ev.clear();
            ph.iter.clear();
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

// mean time of `iters` launches of a streaming probe kernel over `bytes` (read = 0: a copy
// of `bytes`; read = 1: `bytes` read and summed)
static int stream_probe(int device, long long bytes, int iters, int read, double *ms_per_launch) {
    if (bytes < 16 || iters <= 0 || !ms_per_launch) return FFEA_GPU_ERR_INVALID;
    return guarded(nullptr, [&]() {
        FFEA_HIP_TRY(hipSetDevice(device));
        const size_t n2 = (size_t)bytes / 16;
        const unsigned grid = (unsigned)((n2 + (size_t)kProbeVecs * kBlock - 1) / ((size_t)kProbeVecs * kBlock));
        DevBuf<vd2> xx, yy;
        DevBuf<double> part;
        size_t total = 0;
        xx.alloc(n2, &total);
        if (read) part.alloc(grid, &total);
        else yy.alloc(n2, &total);
        FFEA_HIP_TRY(hipMemset(xx.p, 0, n2 * 16));
        hipEvent_t e0, e1;
        FFEA_HIP_TRY(hipEventCreate(&e0));
        FFEA_HIP_TRY(hipEventCreate(&e1));
        auto launch = [&]() {
            if (read) hipLaunchKernelGGL(k_read, dim3(grid), dim3(kBlock), 0, 0, (const vd2 *)xx.p, n2, part.p);
            else hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, 0, (const vd2 *)xx.p, yy.p, n2);
        };
        launch();  // warm-up
        FFEA_HIP_TRY(hipEventRecord(e0, 0));
        for (int i = 0; i < iters; ++i) launch();
        FFEA_HIP_TRY(hipEventRecord(e1, 0));
        FFEA_HIP_TRY(hipEventSynchronize(e1));
        FFEA_HIP_TRY(hipGetLastError());
        float ms = 0;
        FFEA_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_launch = ms / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return (int)FFEA_GPU_OK;
    });
}

int ffea_gpu_copy_probe(int device, long long bytes, int iters, double *ms_per_copy) {
    return stream_probe(device, bytes, iters, 0, ms_per_copy);
}

int ffea_gpu_read_probe(int device, long long bytes, int iters, double *ms_per_read) {
    return stream_probe(device, bytes, iters, 1, ms_per_read);
}

const char *ffea_gpu_blob_last_error(const ffea_gpu_blob *b) { return b ? b->err.c_str() : g_last_error.c_str(); }
const char *ffea_gpu_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
