// Smoothed-aggregation multigrid set-up on the host (see amg_setup.h).
#include "amg_setup.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <functional>
#include <stdexcept>
#include <thread>

namespace amg {

// contiguous ranges of n rows for up to `threads` threads (at least 4096 rows each)
std::vector<int> split_rows(int n, int threads) {
    const int nt = std::max(1, std::min(threads, (n + 4095) / 4096));
    std::vector<int> bounds(nt + 1);
    for (int t = 0; t <= nt; ++t) bounds[t] = (int)((long long)n * t / nt);
    return bounds;
}

// fn(begin, end, t) over the ranges [bounds[t], bounds[t+1]), one thread each; results
// that each range writes to its own rows are independent of the thread count. An
// exception in a range is rethrown here, on the caller's thread, once all ranges ended.
void parallel_ranges(const std::vector<int> &bounds, const std::function<void(int, int, int)> &fn) {
    const int nt = (int)bounds.size() - 1;
    if (nt == 1) {
        fn(bounds[0], bounds[1], 0);
        return;
    }
    std::vector<std::exception_ptr> err(nt);
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            try {
                fn(bounds[t], bounds[t + 1], t);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    for (auto &th : pool) th.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

// fn(begin, end, t) over contiguous row ranges on up to `threads` threads
void parallel_rows(int n, int threads, const std::function<void(int, int, int)> &fn) {
    parallel_ranges(split_rows(n, threads), fn);
}

namespace {

// fn(i) for every i in [0, count) on up to `threads` threads, which take items in turn;
// each item is computed whole by one thread in a fixed order, so results do not depend
// on the thread count
void parallel_items(int count, int threads, const std::function<void(int)> &fn) {
    const int nt = std::max(1, std::min(threads, count));
    std::atomic<int> next{0};
    if (nt == 1) {
        for (int i = 0; i < count; ++i) fn(i);
        return;
    }
    std::vector<std::exception_ptr> err(nt);  // rethrown on the caller's thread
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t]() {
            try {
                for (int i = next++; i < count; i = next++) fn(i);
            } catch (...) {
                err[t] = std::current_exception();
                next = count;  // the other threads stop taking items
            }
        });
    for (auto &th : pool) th.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

typedef double v2d __attribute__((vector_size(16)));
struct int2_t {
    int a, b;
};

// out[a][b] = sum over r < rows of X[r ldx + a] Y[r ldy + b], a, b < 4: a 4x4 tile of
// X^T Y, rows summed in order, the accumulators in registers
inline void tile4x4(const double *X, size_t ldx, const double *Y, size_t ldy, int rows, double (&out)[4][4]) {
    v2d acc[4][2] = {};
    for (int r = 0; r < rows; ++r) {
        const double *x = X + (size_t)r * ldx, *y = Y + (size_t)r * ldy;
        v2d y0, y1;
        __builtin_memcpy(&y0, y, sizeof(v2d));
        __builtin_memcpy(&y1, y + 2, sizeof(v2d));
        for (int a = 0; a < 4; ++a) {
            const v2d xa = {x[a], x[a]};
            acc[a][0] += xa * y0;
            acc[a][1] += xa * y1;
        }
    }
    for (int a = 0; a < 4; ++a) {
        out[a][0] = acc[a][0][0];
        out[a][1] = acc[a][0][1];
        out[a][2] = acc[a][1][0];
        out[a][3] = acc[a][1][1];
    }
}

// Cholesky M = L L^T in place (lower triangle of the row-major n x n M), right-looking
// in panels of kb columns: the panel's diagonal block on one thread, its rows below in
// parallel, then the trailing rows updated in 4x4 tiles from the transposed panel
void cholesky_blocked(std::vector<double> &M, int n, int threads) {
    constexpr int kb = 64;
    const size_t np = (size_t)n + 4;   // panel rows padded so that a 4-wide tile never reads past them
    std::vector<double> PT(kb * np);   // PT[k][i] = L[i][k0 + k]
    for (int k0 = 0; k0 < n; k0 += kb) {
        const int k1 = std::min(n, k0 + kb), w = k1 - k0;
        for (int j = k0; j < k1; ++j) {
            double *rj = &M[(size_t)j * n];
            double d = rj[j];
            for (int k = k0; k < j; ++k) d -= rj[k] * rj[k];
            if (!(d > 0.0)) throw std::runtime_error("multigrid set-up: coarsest operator is not positive definite");
            const double l = std::sqrt(d);
            rj[j] = l;
            for (int i = j + 1; i < k1; ++i) {
                double *ri = &M[(size_t)i * n];
                double t = ri[j];
                for (int k = k0; k < j; ++k) t -= ri[k] * rj[k];
                ri[j] = t / l;
            }
        }
        parallel_items(n - k1, threads, [&](int t) {  // panel: L[i, k0:k1] for i >= k1
            double *ri = &M[(size_t)(k1 + t) * n];
            for (int j = k0; j < k1; ++j) {
                const double *rj = &M[(size_t)j * n];
                double u = ri[j];
                for (int k = k0; k < j; ++k) u -= ri[k] * rj[k];
                ri[j] = u / rj[j];
            }
        });
        std::fill(PT.begin(), PT.end(), 0.0);
        for (int i = k1; i < n; ++i)
            for (int k = 0; k < w; ++k) PT[(size_t)k * np + i] = M[(size_t)i * n + k0 + k];
        const int ntiles = (n - k1 + 3) / 4;
        parallel_items(ntiles, threads, [&](int ti) {  // trailing rows i0..i0+3, columns k1..i
            const int i0 = k1 + 4 * ti;
            for (int j0 = k1; j0 <= i0; j0 += 4) {
                double o[4][4];
                tile4x4(&PT[i0], np, &PT[j0], np, w, o);
                for (int a = 0; a < 4 && i0 + a < n; ++a)
                    for (int b = 0; b < 4 && j0 + b <= i0 + a; ++b) M[(size_t)(i0 + a) * n + j0 + b] -= o[a][b];
            }
        });
    }
}

// largest eigenvalue of M^-1 A by power iteration from a fixed start vector, M the scalar
// diagonal (dinv) or, with binv (3x3 blocks only: [n][9]), the block diagonal; the norms
// are sums of partial sums over fixed chunks of rows, so the estimate does not depend on
// the thread count
double power_lmax(const BCsr &A, const std::vector<double> &dinv, int iters, int threads,
                  const std::vector<double> *binv = nullptr) {
    const int br = A.br, bc = A.bc;
    const size_t n = (size_t)A.n * br;
    constexpr int kRows = 4096;
    const int nchunk = (A.n + kRows - 1) / kRows;
    std::vector<double> v(n), w(n), pv(nchunk), pw(nchunk);
    for (size_t i = 0; i < n; ++i) v[i] = dinv[i] != 0.0 ? 1.0 + 0.5 * std::sin(0.7 * (double)i) : 0.0;
    double lam = 0.0;
    for (int it = 0; it < iters; ++it) {
        parallel_items(nchunk, threads, [&](int c) {  // w = M^-1 A v and the chunk's |v|^2, |w|^2
            double sv = 0.0, sw = 0.0, t[8];
            for (int i = c * kRows; i < std::min(A.n, (c + 1) * kRows); ++i) {
                for (int p = 0; p < br; ++p) {
                    double acc = 0.0;
                    for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                        const double *blk = &A.v[(size_t)k * br * bc + (size_t)p * bc];
                        const double *xx = &v[(size_t)A.ci[k] * bc];
                        for (int q = 0; q < bc; ++q) acc += blk[q] * xx[q];
                    }
                    t[p] = acc;
                }
                for (int p = 0; p < br; ++p) {
                    const size_t r = (size_t)i * br + p;
                    if (binv) {
                        const double *bi = &(*binv)[(size_t)i * 9 + (size_t)p * 3];
                        w[r] = bi[0] * t[0] + bi[1] * t[1] + bi[2] * t[2];
                    } else {
                        w[r] = dinv[r] * t[p];
                    }
                    sv += v[r] * v[r];
                    sw += w[r] * w[r];
                }
            }
            pv[c] = sv;
            pw[c] = sw;
        });
        double nv = 0.0, nw = 0.0;
        for (int c = 0; c < nchunk; ++c) {
            nv += pv[c];
            nw += pw[c];
        }
        if (nv == 0.0 || nw == 0.0) return 1.0;
        lam = std::sqrt(nw / nv);
        const double s = 1.0 / std::sqrt(nw);
        parallel_rows(A.n, threads, [&](int i0, int i1, int) {
            for (size_t r = (size_t)i0 * br; r < (size_t)i1 * br; ++r) v[r] = w[r] * s;
        });
    }
    return lam;
}

// inverses of the 3x3 diagonal blocks, symmetrised ([n][9]; the inverse diagonal where a
// block is not positive definite): the level-0 smoother's block diagonal
std::vector<double> block_inverse3(const BCsr &A) {
    std::vector<double> bi((size_t)A.n * 9, 0.0);
    for (int i = 0; i < A.n; ++i)
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            if (A.ci[k] != i) continue;
            const double *m = &A.v[(size_t)k * 9];
            const double a = m[0], d = m[4], f = m[8];
            const double b = 0.5 * (m[1] + m[3]), c = 0.5 * (m[2] + m[6]), e = 0.5 * (m[5] + m[7]);
            const double A0 = d * f - e * e, B0 = c * e - b * f, C0 = b * e - c * d;
            const double det = a * A0 + b * B0 + c * C0;
            double *o = &bi[(size_t)i * 9];
            if (det > 0.0 && a > 0.0 && a * d - b * b > 0.0) {
                const double s = 1.0 / det;
                const double u[6] = {A0 * s, B0 * s, C0 * s, (a * f - c * c) * s, (b * c - a * e) * s, (a * d - b * b) * s};
                o[0] = u[0]; o[1] = u[1]; o[2] = u[2];
                o[3] = u[1]; o[4] = u[3]; o[5] = u[4];
                o[6] = u[2]; o[7] = u[4]; o[8] = u[5];
            } else {
                o[0] = a > 0.0 ? 1.0 / a : 0.0;
                o[4] = d > 0.0 ? 1.0 / d : 0.0;
                o[8] = f > 0.0 ? 1.0 / f : 0.0;
            }
        }
    return bi;
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
               int br, BCsr &Pt, std::vector<double> &Bc, int threads) {
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
    // aggregates in parallel ranges: each writes its own rows of Pt and Bc
    parallel_rows(na, threads, [&](int a0, int a1, int) {
    std::vector<double> Q;
    for (int a = a0; a < a1; ++a) {
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
    });
}

// C = diag(s) (A B) + E in block CSR (A: br x bk blocks, B: bk x bc; s: a factor per
// scalar row of C, or null; E: br x bc blocks, or null), rows in parallel, columns sorted.
// Two passes over the same row ranges: the first counts each row's distinct columns, the
// second sums each block in a fixed order (A's row, then B's row) straight into the
// exactly sized output, scales the blocks that A B reaches and adds E's. The result does
// not depend on the thread count.
BCsr spgemm(const BCsr &A, const BCsr &B, int threads, const double *s = nullptr, const BCsr *E = nullptr) {
    if (A.bc != B.br || A.ncols != B.n) throw std::runtime_error("multigrid set-up: product shape mismatch");
    const int br = A.br, bk = A.bc, bc = B.bc, n = A.n, m = B.ncols, bb = br * bc;
    if (E && (E->n != n || E->ncols != m || E->br != br || E->bc != bc))
        throw std::runtime_error("multigrid set-up: sum shape mismatch");
    const std::vector<int> bounds = split_rows(n, threads);
    BCsr C;
    C.n = n;
    C.ncols = m;
    C.br = br;
    C.bc = bc;
    C.rp.assign(n + 1, 0);
    // row i's columns into cols (unsorted); mark[c] = 1 for A B's columns, 2 for E's only
    auto columns = [&](int i, std::vector<int> &mark, std::vector<int> &cols) {
        cols.clear();
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
            const int j = A.ci[k];
            for (int l = B.rp[j]; l < B.rp[j + 1]; ++l) {
                const int c = B.ci[l];
                if (mark[c] < 0) {
                    mark[c] = 1;
                    cols.push_back(c);
                }
            }
        }
        if (E)
            for (int k = E->rp[i]; k < E->rp[i + 1]; ++k) {
                const int c = E->ci[k];
                if (mark[c] < 0) {
                    mark[c] = 2;
                    cols.push_back(c);
                }
            }
    };
    parallel_ranges(bounds, [&](int b0, int b1, int) {
        std::vector<int> mark(m, -1), cols;
        for (int i = b0; i < b1; ++i) {
            columns(i, mark, cols);
            C.rp[i + 1] = (int)cols.size();
            for (int c : cols) mark[c] = -1;
        }
    });
    for (int i = 0; i < n; ++i) C.rp[i + 1] += C.rp[i];
    C.ci.resize(C.rp[n]);
    C.v.assign((size_t)C.rp[n] * bb, 0.0);
    parallel_ranges(bounds, [&](int b0, int b1, int) {
        std::vector<int> mark(m, -1), cols;
        std::vector<unsigned char> only_e;
        for (int i = b0; i < b1; ++i) {
            columns(i, mark, cols);
            std::sort(cols.begin(), cols.end());
            const int r0 = C.rp[i];
            only_e.resize(cols.size());
            for (size_t t = 0; t < cols.size(); ++t) {
                C.ci[r0 + t] = cols[t];
                only_e[t] = mark[cols[t]] == 2;
                mark[cols[t]] = (int)t;
            }
            double *crow = &C.v[(size_t)r0 * bb];
            for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int j = A.ci[k];
                const double *a = &A.v[(size_t)k * br * bk];
                for (int l = B.rp[j]; l < B.rp[j + 1]; ++l) {
                    const double *bv = &B.v[(size_t)l * bk * bc];
                    double *c = crow + (size_t)mark[B.ci[l]] * bb;
                    for (int p = 0; p < br; ++p)
                        for (int q = 0; q < bk; ++q) {
                            const double apq = a[p * bk + q];
                            if (apq == 0.0) continue;
                            for (int r = 0; r < bc; ++r) c[p * bc + r] += apq * bv[q * bc + r];
                        }
                }
            }
            if (s)
                for (size_t t = 0; t < cols.size(); ++t) {
                    if (only_e[t]) continue;
                    for (int p = 0; p < br; ++p) {
                        const double sp = s[(size_t)i * br + p];
                        for (int q = 0; q < bc; ++q) crow[t * bb + (size_t)p * bc + q] *= sp;
                    }
                }
            if (E)
                for (int k = E->rp[i]; k < E->rp[i + 1]; ++k) {
                    double *c = crow + (size_t)mark[E->ci[k]] * bb;
                    for (int t = 0; t < bb; ++t) c[t] += E->v[(size_t)k * bb + t];
                }
            for (int c : cols) mark[c] = -1;
        }
    });
    return C;
}

// R = P^T; each row of R lists P's rows in ascending order. Rows of P in parallel
// ranges: range t's entries of a column follow those of the ranges before it, so the
// result does not depend on the thread count.
BCsr transpose(const BCsr &P, int threads) {
    BCsr R;
    R.n = P.ncols;
    R.ncols = P.n;
    R.br = P.bc;
    R.bc = P.br;
    const int bb = P.br * P.bc;
    const std::vector<int> bounds = split_rows(P.n, threads);
    const int nt = (int)bounds.size() - 1;
    std::vector<std::vector<int>> fill(nt, std::vector<int>(R.n, 0));
    parallel_ranges(bounds, [&](int b0, int b1, int t) {
        for (int k = P.rp[b0]; k < P.rp[b1]; ++k) fill[t][P.ci[k]]++;
    });
    R.rp.assign(R.n + 1, 0);
    for (int c = 0; c < R.n; ++c) {
        int at = R.rp[c];
        for (int t = 0; t < nt; ++t) {
            const int cnt = fill[t][c];
            fill[t][c] = at;
            at += cnt;
        }
        R.rp[c + 1] = at;
    }
    R.ci.resize(P.nnzb());
    R.v.resize(P.v.size());
    parallel_ranges(bounds, [&](int b0, int b1, int t) {
        std::vector<int> &f = fill[t];
        for (int i = b0; i < b1; ++i)
            for (int k = P.rp[i]; k < P.rp[i + 1]; ++k) {
                const int d = f[P.ci[k]]++;
                R.ci[d] = i;
                for (int p = 0; p < P.br; ++p)
                    for (int q = 0; q < P.bc; ++q)
                        R.v[(size_t)d * bb + (size_t)q * R.bc + p] = P.v[(size_t)k * bb + (size_t)p * P.bc + q];
            }
    });
    return R;
}

// P = (I - w D^-1 A) Pt = diag(-w D^-1) (A Pt) + Pt
BCsr smooth_prolongator(const BCsr &A, const std::vector<double> &dinv, const BCsr &Pt, double w, int threads) {
    std::vector<double> s(dinv.size());
    for (size_t i = 0; i < s.size(); ++i) s[i] = -w * dinv[i];
    return spgemm(A, Pt, threads, s.data(), &Pt);
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
    parallel_items(n, threads, [&](int i) {  // symmetrise the rounding of the Galerkin product
        for (int j = i + 1; j < n; ++j) {        // (M[j][i], j > i, is row i's alone)
            const double s = 0.5 * (M[(size_t)i * n + j] + M[(size_t)j * n + i]);
            M[(size_t)i * n + j] = M[(size_t)j * n + i] = s;
        }
    });
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
        parallel_rows(n, threads, [&](int i0, int i1, int) {
            for (int i = i0; i < i1; ++i)
                for (int j = 0; j < n; ++j) {
                    double acc = 0.0;
                    for (int k = 0; k < 6; ++k) acc += Q[(size_t)i * 6 + k] * Q[(size_t)j * 6 + k];
                    M[(size_t)i * n + j] += scale * acc;
                }
        });
    }
    for (int i = 0; i < n; ++i) {
        bool zero = true;
        for (int j = 0; j < n && zero; ++j) zero = M[(size_t)i * n + j] == 0.0;
        if (zero) M[(size_t)i * n + i] = 1.0;
    }
    cholesky_blocked(M, n, threads);
    // W = L^-1 in groups of kg columns: group g holds rows c0 .. n-1 of columns c0 .. c0+kg-1
    // (row-major, kg wide; W is zero above its diagonal), each row a forward substitution
    // that reads L's row once for the whole group
    constexpr int kg = 16;
    const int ng = (n + kg - 1) / kg;
    std::vector<size_t> goff(ng + 1, 0);
    for (int g = 0; g < ng; ++g) goff[g + 1] = goff[g] + (size_t)(n - g * kg) * kg + 4 * kg;  // + 4 rows of zero pad
    std::vector<double> W(goff[ng], 0.0);
    parallel_items(ng, threads, [&](int g) {
        const int c0 = g * kg;
        double *wg = &W[goff[g]];
        for (int i = c0; i < n; ++i) {
            const double *ri = &M[(size_t)i * n];
            v2d acc[kg / 2] = {};
            for (int k = c0; k < i; ++k) {
                const v2d a = {ri[k], ri[k]};
                const double *wk = wg + (size_t)(k - c0) * kg;
                for (int h = 0; h < kg / 2; ++h) {
                    v2d x;
                    __builtin_memcpy(&x, wk + 2 * h, sizeof(v2d));
                    acc[h] += a * x;
                }
            }
            double *wi = wg + (size_t)(i - c0) * kg;
            for (int c = 0; c < kg; ++c) wi[c] = ((i == c0 + c ? 1.0 : 0.0) - acc[c / 2][c % 2]) / ri[i];
        }
    });
    // inv = W^T W: block (g1, g2), g2 <= g1, sums rows r >= c0(g1) (W's columns of g1 are zero
    // above), in 4x4 tiles; written to both triangles
    std::vector<double> inv((size_t)n * n, 0.0);
    std::vector<int2_t> pairs;
    for (int g1 = 0; g1 < ng; ++g1)
        for (int g2 = 0; g2 <= g1; ++g2) pairs.push_back({g1, g2});
    parallel_items((int)pairs.size(), threads, [&](int q) {
        const int g1 = pairs[q].a, g2 = pairs[q].b, c1 = g1 * kg, c2 = g2 * kg, rows = n - c1;
        const double *x = &W[goff[g1]], *y = &W[goff[g2]] + (size_t)(c1 - c2) * kg;
        for (int a = 0; a < kg; a += 4)
            for (int b = 0; b < kg; b += 4) {
                double o[4][4];
                tile4x4(x + a, kg, y + b, kg, rows, o);
                for (int u = 0; u < 4; ++u)
                    for (int v = 0; v < 4; ++v) {
                        const int i = c1 + a + u, j = c2 + b + v;
                        if (i >= n || j >= n || j > i) continue;
                        inv[(size_t)i * n + j] = o[u][v];
                        inv[(size_t)j * n + i] = o[u][v];
                    }
            }
    });
    return inv;
}

// the 6 rigid modes at one node (3 rows of 6): translations e_k and rotations e_k x d,
// d = p - c
void rigid_row(const double *p, const double c[3], double *b) {
    const double d[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    for (int k = 0; k < 18; ++k) b[k] = 0.0;
    for (int k = 0; k < 3; ++k) b[k * 6 + k] = 1.0;
    b[0 * 6 + 3] = 0.0;   b[1 * 6 + 3] = -d[2]; b[2 * 6 + 3] = d[1];
    b[0 * 6 + 4] = d[2];  b[1 * 6 + 4] = 0.0;   b[2 * 6 + 4] = -d[0];
    b[0 * 6 + 5] = -d[1]; b[1 * 6 + 5] = d[0];  b[2 * 6 + 5] = 0.0;
}

void rigid_modes(int n, const double *pos, const double c[3], std::vector<double> &B) {
    B.assign((size_t)n * 3 * 6, 0.0);
    for (int i = 0; i < n; ++i) rigid_row(pos + 3 * i, c, &B[(size_t)i * 18]);  // rows 3i..3i+2, 6 columns
}

}  // namespace

Hierarchy build(BCsr fine_in, const double *pos, const unsigned char *pinned, bool free_blob, const Options &opt) {
    if (fine_in.br != 3 || fine_in.bc != 3 || fine_in.ncols != fine_in.n || (int)fine_in.rp.size() != fine_in.n + 1)
        throw std::runtime_error("multigrid set-up: the fine operator must be square with 3x3 blocks");
    for (long long k = 0; k < fine_in.nnzb(); ++k)
        if (fine_in.ci[k] < 0 || fine_in.ci[k] >= fine_in.n) throw std::runtime_error("multigrid set-up: block column out of range");
    Hierarchy H;
    H.lv.reserve(std::max(1, opt.max_levels));  // no reallocation: `fine` stays valid
    H.lv.emplace_back();
    H.lv[0].A = std::move(fine_in);
    const BCsr &fine = H.lv[0].A;
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
        const double est = power_lmax(L.A, L.dinv, 30, opt.threads);  // D^-1 A: the prolongator's smoothing
        // level 0 smooths with the inverses of its 3x3 diagonal blocks B, its weight from B^-1 A
        double est_s = est;
        if (l == 0) {
            const std::vector<double> bi = block_inverse3(L.A);
            est_s = power_lmax(L.A, L.dinv, 30, opt.threads, &bi);
        }
        L.lmax = est_s;
        const double lhat = est_s * (l == 0 ? opt.smoother_safety : 1.1);
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
            parallel_rows(fine.n, opt.threads, [&](int i0, int i1, int) {
                for (int i = i0; i < i1; ++i) {
                    if (agg[i] < 0) continue;
                    const int a = agg[i];
                    const double c[3] = {sum[3 * a] / cnt[a], sum[3 * a + 1] / cnt[a], sum[3 * a + 2] / cnt[a]};
                    rigid_row(pos + 3 * i, c, &Bloc[(size_t)i * 18]);
                }
            });
            tentative(agg, na, Bloc, Bglob, 3, Pt, Bc, opt.threads);
        } else {
            tentative(agg, na, B, B, L.A.br, Pt, Bc, opt.threads);
        }
        const double w = 4.0 / (3.0 * est * 1.05);
        BCsr P = smooth_prolongator(L.A, L.dinv, Pt, w, opt.threads);
        BCsr R = transpose(P, opt.threads);
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
