// Smoothed-aggregation multigrid set-up on the host (see amg_setup.h).
#include "amg_setup.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <thread>

namespace amg {
namespace {

// fn(begin, end) over contiguous row ranges on up to `threads` threads; results that
// each range writes to its own rows are independent of the thread count
void parallel_rows(int n, int threads, const std::function<void(int, int, int)> &fn) {
    const int nt = std::max(1, std::min(threads, (n + 4095) / 4096));
    if (nt == 1) {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) {
        const int b = (int)((long long)n * t / nt), e = (int)((long long)n * (t + 1) / nt);
        pool.emplace_back(fn, b, e, t);
    }
    for (auto &th : pool) th.join();
}

// fn(i) for every i in [0, count) on up to `threads` threads, which take items in turn;
// each item is computed whole by one thread in a fixed order, so results do not depend
// on the thread count
void parallel_items(int count, int threads, const std::function<void(int)> &fn) {
    const int nt = std::max(1, std::min(threads, count));
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < count; i = next++) fn(i);
    };
    if (nt == 1) {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(work);
    for (auto &th : pool) th.join();
}

// dot product of n contiguous values in four interleaved partial sums (a fixed order)
inline double dot4(const double *a, const double *b, int n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = 0;
    for (; k + 4 <= n; k += 4) {
        s0 += a[k] * b[k];
        s1 += a[k + 1] * b[k + 1];
        s2 += a[k + 2] * b[k + 2];
        s3 += a[k + 3] * b[k + 3];
    }
    for (; k < n; ++k) s0 += a[k] * b[k];
    return (s0 + s1) + (s2 + s3);
}

// Cholesky M = L L^T in place (lower triangle of the row-major n x n M), right-looking
// in blocks of kb columns: the diagonal block on one thread, then the panel rows below
// it and the trailing rows, one row per parallel item
void cholesky_blocked(std::vector<double> &M, int n, int threads) {
    constexpr int kb = 64;
    for (int k0 = 0; k0 < n; k0 += kb) {
        const int k1 = std::min(n, k0 + kb);
        for (int j = k0; j < k1; ++j) {
            double *rj = &M[(size_t)j * n];
            const double d = rj[j] - dot4(rj + k0, rj + k0, j - k0);
            if (!(d > 0.0)) throw std::runtime_error("multigrid set-up: coarsest operator is not positive definite");
            const double l = std::sqrt(d);
            rj[j] = l;
            for (int i = j + 1; i < k1; ++i) {
                double *ri = &M[(size_t)i * n];
                ri[j] = (ri[j] - dot4(ri + k0, rj + k0, j - k0)) / l;
            }
        }
        parallel_items(n - k1, threads, [&](int t) {  // panel: L[i, k0:k1] for i >= k1
            double *ri = &M[(size_t)(k1 + t) * n];
            for (int j = k0; j < k1; ++j) {
                const double *rj = &M[(size_t)j * n];
                ri[j] = (ri[j] - dot4(ri + k0, rj + k0, j - k0)) / rj[j];
            }
        });
        parallel_items(n - k1, threads, [&](int t) {  // trailing update of row i >= k1, columns k1..i
            const int i = k1 + t;
            double *ri = &M[(size_t)i * n];
            for (int j = k1; j <= i; ++j) ri[j] -= dot4(ri + k0, &M[(size_t)j * n] + k0, k1 - k0);
        });
    }
}

// y = D^-1 A x (scalar diagonal)
void dinv_apply(const BCsr &A, const std::vector<double> &dinv, const std::vector<double> &x, std::vector<double> &y,
                int threads) {
    const int br = A.br, bc = A.bc;
    parallel_rows(A.n, threads, [&](int b, int e, int) {
        for (int i = b; i < e; ++i) {
            for (int p = 0; p < br; ++p) {
                double acc = 0.0;
                for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                    const double *blk = &A.v[(size_t)k * br * bc + (size_t)p * bc];
                    const double *xx = &x[(size_t)A.ci[k] * bc];
                    for (int q = 0; q < bc; ++q) acc += blk[q] * xx[q];
                }
                y[(size_t)i * br + p] = dinv[(size_t)i * br + p] * acc;
            }
        }
    });
}

// largest eigenvalue of D^-1 A by power iteration from a fixed start vector
double power_lmax(const BCsr &A, const std::vector<double> &dinv, int iters, int threads) {
    const size_t n = (size_t)A.n * A.br;
    std::vector<double> v(n), w(n);
    for (size_t i = 0; i < n; ++i) v[i] = dinv[i] != 0.0 ? 1.0 + 0.5 * std::sin(0.7 * (double)i) : 0.0;
    double lam = 0.0;
    for (int it = 0; it < iters; ++it) {
        dinv_apply(A, dinv, v, w, threads);
        double nv = 0.0, nw = 0.0;
        for (size_t i = 0; i < n; ++i) {
            nv += v[i] * v[i];
            nw += w[i] * w[i];
        }
        if (nv == 0.0 || nw == 0.0) return 1.0;
        lam = std::sqrt(nw / nv);
        const double s = 1.0 / std::sqrt(nw);
        for (size_t i = 0; i < n; ++i) v[i] = w[i] * s;
    }
    return lam;
}

std::vector<double> inverse_diagonal(const BCsr &A) {
    const int br = A.br, bc = A.bc;
    std::vector<double> d((size_t)A.n * br, 0.0);
    for (int i = 0; i < A.n; ++i)
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (A.ci[k] == i)
                for (int p = 0; p < br; ++p) {
                    const double a = A.v[(size_t)k * br * bc + (size_t)p * bc + p];
                    d[(size_t)i * br + p] = a > 0.0 ? 1.0 / a : 0.0;  // dead unknown: left alone
                }
    return d;
}

// Greedy aggregation over the block graph (node order): 1) a node whose neighbours are
// all free starts an aggregate with them; 2) a left-over node joins the aggregate most
// of its neighbours belong to (lowest index on ties); 3) what remains groups with its
// free neighbours. Rows with active[i] == 0 get no aggregate (-1).
int aggregate(const BCsr &A, const std::vector<unsigned char> &active, std::vector<int> &agg) {
    const int n = A.n;
    agg.assign(n, -1);
    int na = 0;
    auto nbrs_free = [&](int i) {
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int j = A.ci[k];
            if (j != i && active[j] && agg[j] >= 0) return false;
        }
        return true;
    };
    for (int i = 0; i < n; ++i) {
        if (!active[i] || agg[i] >= 0 || !nbrs_free(i)) continue;
        agg[i] = na;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
            if (active[A.ci[k]]) agg[A.ci[k]] = na;
        ++na;
    }
    std::vector<int> cnt_idx, cnt_val;
    for (int i = 0; i < n; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        cnt_idx.clear();
        cnt_val.clear();
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int j = A.ci[k];
            if (j == i || !active[j] || agg[j] < 0) continue;
            const int a = agg[j];
            size_t t = 0;
            while (t < cnt_idx.size() && cnt_idx[t] != a) ++t;
            if (t == cnt_idx.size()) {
                cnt_idx.push_back(a);
                cnt_val.push_back(0);
            }
            cnt_val[t]++;
        }
        int best = -1, bv = 0;
        for (size_t t = 0; t < cnt_idx.size(); ++t)
            if (cnt_val[t] > bv || (cnt_val[t] == bv && cnt_idx[t] < best)) {
                best = cnt_idx[t];
                bv = cnt_val[t];
            }
        agg[i] = best;  // -1 when no neighbour is aggregated: phase 3
    }
    for (int i = 0; i < n; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        agg[i] = na;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int j = A.ci[k];
            if (active[j] && agg[j] < 0) agg[j] = na;
        }
        ++na;
    }
    return na;
}

// Tentative prolongator: per aggregate, an orthonormal basis Q of the modes Bloc restricted
// to its rows (modified Gram-Schmidt, twice; a mode that is not independent there gives a
// zero column), and the coarse modes Bc = Q^T Bglob. With Bloc = Bglob this is a QR.
// B arrays are [rows][6], rows = n * br.
void tentative(const std::vector<int> &agg, int na, const std::vector<double> &Bloc, const std::vector<double> &Bglob,
               int br, BCsr &Pt, std::vector<double> &Bc) {
    const int n = (int)agg.size();
    std::vector<int> start(na + 1, 0), order(n);
    for (int i = 0; i < n; ++i)
        if (agg[i] >= 0) start[agg[i] + 1]++;
    for (int a = 0; a < na; ++a) start[a + 1] += start[a];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; ++i)
            if (agg[i] >= 0) order[fill[agg[i]]++] = i;
    }
    Pt = BCsr{};
    Pt.n = n;
    Pt.ncols = na;
    Pt.br = br;
    Pt.bc = 6;
    Pt.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) Pt.rp[i + 1] = Pt.rp[i] + (agg[i] >= 0 ? 1 : 0);
    Pt.ci.assign(Pt.rp[n], 0);
    Pt.v.assign((size_t)Pt.rp[n] * br * 6, 0.0);
    Bc.assign((size_t)na * 36, 0.0);
    std::vector<double> Q;
    for (int a = 0; a < na; ++a) {
        const int m = start[a + 1] - start[a], rows = m * br;
        Q.assign((size_t)rows * 6, 0.0);
        for (int t = 0; t < m; ++t) {
            const int i = order[start[a] + t];
            for (int p = 0; p < br; ++p)
                for (int k = 0; k < 6; ++k) Q[(size_t)(t * br + p) * 6 + k] = Bloc[((size_t)i * br + p) * 6 + k];
        }
        for (int k = 0; k < 6; ++k) {
            double n0 = 0.0;
            for (int r = 0; r < rows; ++r) n0 += Q[(size_t)r * 6 + k] * Q[(size_t)r * 6 + k];
            n0 = std::sqrt(n0);
            for (int pass = 0; pass < 2; ++pass)
                for (int j = 0; j < k; ++j) {
                    double d = 0.0;
                    for (int r = 0; r < rows; ++r) d += Q[(size_t)r * 6 + j] * Q[(size_t)r * 6 + k];
                    for (int r = 0; r < rows; ++r) Q[(size_t)r * 6 + k] -= d * Q[(size_t)r * 6 + j];
                }
            double nk = 0.0;
            for (int r = 0; r < rows; ++r) nk += Q[(size_t)r * 6 + k] * Q[(size_t)r * 6 + k];
            nk = std::sqrt(nk);
            const double s = (n0 > 0.0 && nk > 1e-8 * n0) ? 1.0 / nk : 0.0;
            for (int r = 0; r < rows; ++r) Q[(size_t)r * 6 + k] *= s;
        }
        for (int t = 0; t < m; ++t) {
            const int i = order[start[a] + t];
            const int slot = Pt.rp[i];
            Pt.ci[slot] = a;
            for (int p = 0; p < br; ++p)
                for (int k = 0; k < 6; ++k)
                    Pt.v[(size_t)slot * br * 6 + (size_t)p * 6 + k] = Q[(size_t)(t * br + p) * 6 + k];
            // Bc += Q_t^T Bglob_t
            for (int k = 0; k < 6; ++k)
                for (int c = 0; c < 6; ++c) {
                    double acc = 0.0;
                    for (int p = 0; p < br; ++p)
                        acc += Q[(size_t)(t * br + p) * 6 + k] * Bglob[((size_t)i * br + p) * 6 + c];
                    Bc[(size_t)a * 36 + (size_t)k * 6 + c] += acc;
                }
        }
    }
}

// C = A B in block CSR (A: br x bk blocks, B: bk x bc), rows in parallel, columns sorted;
// scale: optional per-scalar-row factor applied to A's rows (C = diag(s) A B)
BCsr spgemm(const BCsr &A, const BCsr &B, int threads) {
    if (A.bc != B.br || A.ncols != B.n) throw std::runtime_error("multigrid set-up: product shape mismatch");
    const int br = A.br, bk = A.bc, bc = B.bc, n = A.n, m = B.ncols;
    const int nt = std::max(1, std::min(threads, (n + 4095) / 4096));
    std::vector<std::vector<int>> rcnt(nt), rci(nt);
    std::vector<std::vector<double>> rv(nt);
    std::vector<int> bounds(nt + 1);
    for (int t = 0; t <= nt; ++t) bounds[t] = (int)((long long)n * t / nt);
    auto work = [&](int t) {
        std::vector<int> mark(m, -1), cols;
        std::vector<double> acc;
        for (int i = bounds[t]; i < bounds[t + 1]; ++i) {
            cols.clear();
            for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int j = A.ci[k];
                for (int l = B.rp[j]; l < B.rp[j + 1]; ++l) {
                    const int c = B.ci[l];
                    if (mark[c] < 0) {
                        mark[c] = (int)cols.size();
                        cols.push_back(c);
                    }
                }
            }
            std::sort(cols.begin(), cols.end());
            for (size_t s = 0; s < cols.size(); ++s) mark[cols[s]] = (int)s;
            acc.assign(cols.size() * br * bc, 0.0);
            for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int j = A.ci[k];
                const double *a = &A.v[(size_t)k * br * bk];
                for (int l = B.rp[j]; l < B.rp[j + 1]; ++l) {
                    const double *b = &B.v[(size_t)l * bk * bc];
                    double *c = &acc[(size_t)mark[B.ci[l]] * br * bc];
                    for (int p = 0; p < br; ++p)
                        for (int q = 0; q < bk; ++q) {
                            const double apq = a[p * bk + q];
                            if (apq == 0.0) continue;
                            for (int r = 0; r < bc; ++r) c[p * bc + r] += apq * b[q * bc + r];
                        }
                }
            }
            rcnt[t].push_back((int)cols.size());
            rci[t].insert(rci[t].end(), cols.begin(), cols.end());
            rv[t].insert(rv[t].end(), acc.begin(), acc.end());
            for (int c : cols) mark[c] = -1;
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t) pool.emplace_back(work, t);
        for (auto &th : pool) th.join();
    }
    BCsr C;
    C.n = n;
    C.ncols = m;
    C.br = br;
    C.bc = bc;
    C.rp.assign(n + 1, 0);
    int row = 0;
    for (int t = 0; t < nt; ++t)
        for (int c : rcnt[t]) {
            C.rp[row + 1] = C.rp[row] + c;
            ++row;
        }
    C.ci.reserve(C.rp[n]);
    C.v.reserve((size_t)C.rp[n] * br * bc);
    for (int t = 0; t < nt; ++t) {
        C.ci.insert(C.ci.end(), rci[t].begin(), rci[t].end());
        C.v.insert(C.v.end(), rv[t].begin(), rv[t].end());
    }
    return C;
}

BCsr transpose(const BCsr &P) {
    BCsr R;
    R.n = P.ncols;
    R.ncols = P.n;
    R.br = P.bc;
    R.bc = P.br;
    R.rp.assign(R.n + 1, 0);
    for (long long k = 0; k < P.nnzb(); ++k) R.rp[P.ci[k] + 1]++;
    for (int i = 0; i < R.n; ++i) R.rp[i + 1] += R.rp[i];
    R.ci.assign(P.nnzb(), 0);
    R.v.assign(P.v.size(), 0.0);
    std::vector<int> fill(R.rp.begin(), R.rp.end() - 1);
    for (int i = 0; i < P.n; ++i)  // ascending fine rows: each row of R is sorted
        for (int k = P.rp[i]; k < P.rp[i + 1]; ++k) {
            const int s = fill[P.ci[k]]++;
            R.ci[s] = i;
            for (int p = 0; p < P.br; ++p)
                for (int q = 0; q < P.bc; ++q)
                    R.v[(size_t)s * R.br * R.bc + (size_t)q * R.bc + p] = P.v[(size_t)k * P.br * P.bc + (size_t)p * P.bc + q];
        }
    return R;
}

// P = (I - w D^-1 A) Pt
BCsr smooth_prolongator(const BCsr &A, const std::vector<double> &dinv, const BCsr &Pt, double w, int threads) {
    BCsr AP = spgemm(A, Pt, threads);
    const int br = Pt.br, bc = Pt.bc;
    for (int i = 0; i < AP.n; ++i)
        for (int k = AP.rp[i]; k < AP.rp[i + 1]; ++k)
            for (int p = 0; p < br; ++p) {
                const double s = -w * dinv[(size_t)i * br + p];
                for (int q = 0; q < bc; ++q) AP.v[(size_t)k * br * bc + (size_t)p * bc + q] *= s;
            }
    // add Pt (at most one block per row, its column is in AP's row when A has a diagonal)
    BCsr P;
    P.n = AP.n;
    P.ncols = AP.ncols;
    P.br = br;
    P.bc = bc;
    P.rp.assign(P.n + 1, 0);
    for (int i = 0; i < P.n; ++i) {
        int extra = 0;
        for (int k = Pt.rp[i]; k < Pt.rp[i + 1]; ++k) {
            bool found = false;
            for (int l = AP.rp[i]; l < AP.rp[i + 1]; ++l) found |= AP.ci[l] == Pt.ci[k];
            extra += found ? 0 : 1;
        }
        P.rp[i + 1] = P.rp[i] + (AP.rp[i + 1] - AP.rp[i]) + extra;
    }
    P.ci.resize(P.rp[P.n]);
    P.v.assign((size_t)P.rp[P.n] * br * bc, 0.0);
    for (int i = 0; i < P.n; ++i) {
        int o = P.rp[i];
        std::vector<std::pair<int, const double *>> ent;
        for (int l = AP.rp[i]; l < AP.rp[i + 1]; ++l) ent.push_back({AP.ci[l], &AP.v[(size_t)l * br * bc]});
        for (int k = Pt.rp[i]; k < Pt.rp[i + 1]; ++k) {
            bool found = false;
            for (auto &e : ent) found |= e.first == Pt.ci[k];
            if (!found) ent.push_back({Pt.ci[k], nullptr});
        }
        std::sort(ent.begin(), ent.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (auto &e : ent) {
            P.ci[o] = e.first;
            double *dst = &P.v[(size_t)o * br * bc];
            if (e.second)
                for (int t = 0; t < br * bc; ++t) dst[t] = e.second[t];
            for (int k = Pt.rp[i]; k < Pt.rp[i + 1]; ++k)
                if (Pt.ci[k] == e.first)
                    for (int t = 0; t < br * bc; ++t) dst[t] += Pt.v[(size_t)k * br * bc + t];
            ++o;
        }
    }
    return P;
}

// dense inverse of the coarsest operator (column-major), regularised on the modes B
// (n x 6, orthonormalised here) for a singular (free) operator; dead unknowns get a 1
// on the diagonal
std::vector<double> coarsest_inverse(const BCsr &A, const std::vector<double> &B, bool regularise, int threads) {
    const int n = A.n * A.br, br = A.br, bc = A.bc;
    std::vector<double> M((size_t)n * n, 0.0);  // row-major == column-major (symmetric)
    for (int i = 0; i < A.n; ++i)
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k)
            for (int p = 0; p < br; ++p)
                for (int q = 0; q < bc; ++q)
                    M[(size_t)(i * br + p) * n + (size_t)A.ci[k] * bc + q] += A.v[(size_t)k * br * bc + p * bc + q];
    for (int i = 0; i < n; ++i)  // symmetrise the rounding of the Galerkin product
        for (int j = i + 1; j < n; ++j) {
            const double s = 0.5 * (M[(size_t)i * n + j] + M[(size_t)j * n + i]);
            M[(size_t)i * n + j] = M[(size_t)j * n + i] = s;
        }
    double tr = 0.0;
    int live = 0;
    for (int i = 0; i < n; ++i)
        if (M[(size_t)i * n + i] > 0.0) {
            tr += M[(size_t)i * n + i];
            ++live;
        }
    const double scale = live ? tr / live : 1.0;
    if (regularise) {
        std::vector<double> Q(B);
        for (int k = 0; k < 6; ++k) {
            for (int pass = 0; pass < 2; ++pass)
                for (int j = 0; j < k; ++j) {
                    double d = 0.0;
                    for (int r = 0; r < n; ++r) d += Q[(size_t)r * 6 + j] * Q[(size_t)r * 6 + k];
                    for (int r = 0; r < n; ++r) Q[(size_t)r * 6 + k] -= d * Q[(size_t)r * 6 + j];
                }
            double nk = 0.0;
            for (int r = 0; r < n; ++r) nk += Q[(size_t)r * 6 + k] * Q[(size_t)r * 6 + k];
            const double s = nk > 0.0 ? 1.0 / std::sqrt(nk) : 0.0;
            for (int r = 0; r < n; ++r) Q[(size_t)r * 6 + k] *= s;
        }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 6; ++k) acc += Q[(size_t)i * 6 + k] * Q[(size_t)j * 6 + k];
                M[(size_t)i * n + j] += scale * acc;
            }
    }
    for (int i = 0; i < n; ++i) {
        bool zero = true;
        for (int j = 0; j < n && zero; ++j) zero = M[(size_t)i * n + j] == 0.0;
        if (zero) M[(size_t)i * n + i] = 1.0;
    }
    cholesky_blocked(M, n, threads);
    // W = L^-1 by columns (column c solves L w = e_c, w zero above row c), then
    // inv = W^T W: entry (i, j), j <= i, is the dot product of columns i and j from row i
    std::vector<double> W((size_t)n * n, 0.0);
    parallel_items(n, threads, [&](int c) {
        double *w = &W[(size_t)c * n];
        w[c] = 1.0 / M[(size_t)c * n + c];
        for (int i = c + 1; i < n; ++i) {
            const double *ri = &M[(size_t)i * n];
            w[i] = -dot4(ri + c, w + c, i - c) / ri[i];
        }
    });
    std::vector<double> inv((size_t)n * n, 0.0);
    parallel_items(n, threads, [&](int i) {
        const double *wi = &W[(size_t)i * n];
        for (int j = 0; j <= i; ++j) {
            const double v = dot4(wi + i, &W[(size_t)j * n] + i, n - i);
            inv[(size_t)i * n + j] = v;
            inv[(size_t)j * n + i] = v;
        }
    });
    return inv;
}

void rigid_modes(int n, const double *pos, const double c[3], std::vector<double> &B) {
    B.assign((size_t)n * 3 * 6, 0.0);
    for (int i = 0; i < n; ++i) {
        const double d[3] = {pos[3 * i] - c[0], pos[3 * i + 1] - c[1], pos[3 * i + 2] - c[2]};
        double *b = &B[(size_t)i * 18];  // rows 3i..3i+2, 6 columns
        for (int k = 0; k < 3; ++k) b[k * 6 + k] = 1.0;
        // column 3+k: e_k cross d
        b[0 * 6 + 3] = 0.0;   b[1 * 6 + 3] = -d[2]; b[2 * 6 + 3] = d[1];
        b[0 * 6 + 4] = d[2];  b[1 * 6 + 4] = 0.0;   b[2 * 6 + 4] = -d[0];
        b[0 * 6 + 5] = -d[1]; b[1 * 6 + 5] = d[0];  b[2 * 6 + 5] = 0.0;
    }
}

}  // namespace

Hierarchy build(const BCsr &fine, const double *pos, const unsigned char *pinned, bool free_blob, const Options &opt) {
    if (fine.br != 3 || fine.bc != 3 || fine.ncols != fine.n || (int)fine.rp.size() != fine.n + 1)
        throw std::runtime_error("multigrid set-up: the fine operator must be square with 3x3 blocks");
    for (long long k = 0; k < fine.nnzb(); ++k)
        if (fine.ci[k] < 0 || fine.ci[k] >= fine.n) throw std::runtime_error("multigrid set-up: block column out of range");
    Hierarchy H;
    H.lv.emplace_back();
    H.lv[0].A = fine;
    std::vector<unsigned char> active(fine.n, 1);
    double cen[3] = {0, 0, 0};
    int nfree = 0;
    for (int i = 0; i < fine.n; ++i) {
        if (pinned && pinned[i]) {
            active[i] = 0;
            continue;
        }
        for (int k = 0; k < 3; ++k) cen[k] += pos[3 * i + k];
        ++nfree;
    }
    for (int k = 0; k < 3; ++k) cen[k] /= std::max(1, nfree);
    std::vector<double> Bglob;
    rigid_modes(fine.n, pos, cen, Bglob);
    std::vector<double> B = Bglob;
    for (int l = 0;; ++l) {
        Level &L = H.lv[l];
        L.dinv = inverse_diagonal(L.A);
        const double est = power_lmax(L.A, L.dinv, 30, opt.threads);
        L.lmax = est;
        const double lhat = est * (l == 0 ? opt.smoother_safety : 1.1);
        L.omega = 2.0 / (lhat * (1.0 + 1.0 / 30.0));
        const int unknowns = L.A.n * L.A.br;
        if (unknowns <= opt.max_coarse || l + 1 >= opt.max_levels) break;
        std::vector<int> agg;
        const int na = aggregate(L.A, active, agg);
        if (na == 0 || 6 * na >= (int)(0.8 * unknowns)) break;  // no coarsening progress
        if (na < opt.min_aggregates) break;
        BCsr Pt;
        std::vector<double> Bc;
        if (l == 0) {
            // local centring per aggregate for the orthogonalisation; Bc from the global modes
            std::vector<double> Bloc((size_t)fine.n * 18, 0.0), sum((size_t)na * 3, 0.0);
            std::vector<int> cnt(na, 0);
            for (int i = 0; i < fine.n; ++i)
                if (agg[i] >= 0) {
                    for (int k = 0; k < 3; ++k) sum[(size_t)agg[i] * 3 + k] += pos[3 * i + k];
                    cnt[agg[i]]++;
                }
            for (int i = 0; i < fine.n; ++i) {
                if (agg[i] < 0) continue;
                const int a = agg[i];
                const double c[3] = {sum[3 * a] / cnt[a], sum[3 * a + 1] / cnt[a], sum[3 * a + 2] / cnt[a]};
                std::vector<double> one;
                rigid_modes(1, pos + 3 * i, c, one);
                std::copy(one.begin(), one.end(), Bloc.begin() + (size_t)i * 18);
            }
            tentative(agg, na, Bloc, Bglob, 3, Pt, Bc);
        } else {
            tentative(agg, na, B, B, L.A.br, Pt, Bc);
        }
        const double w = 4.0 / (3.0 * est * 1.05);
        BCsr P = smooth_prolongator(L.A, L.dinv, Pt, w, opt.threads);
        BCsr R = transpose(P);
        BCsr AP = spgemm(L.A, P, opt.threads);
        BCsr Ac = spgemm(R, AP, opt.threads);
        L.P = std::move(P);
        L.R = std::move(R);
        Level next;
        next.A = std::move(Ac);
        H.lv.push_back(std::move(next));
        B = std::move(Bc);
        active.assign(H.lv[l + 1].A.n, 1);
    }
    const Level &C = H.lv.back();
    H.nc = C.A.n * C.A.br;
    H.cinv = coarsest_inverse(C.A, B, free_blob, opt.threads);
    return H;
}

}  // namespace amg
