// Host construction of the smoothed-aggregation multigrid hierarchy that preconditions
// the blob's viscosity-matrix CG (an alternative to the reference's Jacobi
// preconditioner, NoMassCGSolver.cpp:268-281). Plain C++, no device code.
//
// Level 0 is the corner-node viscosity matrix in 3x3 blocks. Each coarser level holds
// 6 unknowns per aggregate: the 3 translations and 3 infinitesimal rotations of the
// aggregate's nodes (the operator's null space for a free blob, its near-null space
// otherwise). Transfers are the tentative prolongator of those modes smoothed by one
// damped-Jacobi step; coarse operators are Galerkin products; the coarsest level is
// inverted densely (regularised on the rigid modes for a free blob).
#pragma once
#include <functional>
#include <string>
#include <vector>

namespace amg {

// block compressed sparse rows: n block rows of br x bc blocks, row-major inside a block
struct BCsr {
    int n = 0, ncols = 0, br = 0, bc = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    long long nnzb() const { return rp.empty() ? 0 : rp.back(); }
};

struct Level {
    BCsr A;                    // operator (level 0: a copy of the fine matrix)
    std::vector<double> dinv;  // inverse scalar diagonal (0 for dead unknowns)
    double lmax = 0.0;         // power-iteration estimate of the largest eigenvalue of M^-1 A, M the
                               // smoother's diagonal: the 3x3 diagonal blocks on level 0, D above
    double omega = 0.0;        // smoother weight: x += omega M^-1 (b - A x)
    BCsr P, R;                 // transfer to the next level and its transpose (not on the coarsest)
};

struct Hierarchy {
    std::vector<Level> lv;     // lv.back() is the coarsest level
    int nc = 0;                // unknowns of the coarsest level
    std::vector<double> cinv;  // its dense inverse, column-major nc x nc
};

struct Options {
    int max_coarse = 1200;     // unknowns at which coarsening stops
    int max_levels = 8;
    double smoother_safety = 1.3;  // the fine operator drifts from the set-up one over a run
    int min_aggregates = 1;    // a coarsening to fewer aggregates stops instead: a level of one
                               // aggregate of a free blob is its null space and nothing else
    int threads = 16;
};

// fine: nl x nl matrix of 3x3 blocks over corner nodes (pinned rows: identity blocks,
// no couplings); pos: [nl][3] positions; pinned: [nl] flags (may be null); free_blob:
// the operator is singular with the 6 rigid modes as null space.
// Throws std::runtime_error on inconsistent input.
Hierarchy build(BCsr fine, const double *pos, const unsigned char *pinned, bool free_blob, const Options &opt);

// contiguous ranges of n rows for up to `threads` threads (bounds[t] .. bounds[t+1])
std::vector<int> split_rows(int n, int threads);
// fn(begin, end, t) over the ranges, one thread each; the first exception a range throws
// is rethrown on the caller's thread after every range has ended
void parallel_ranges(const std::vector<int> &bounds, const std::function<void(int, int, int)> &fn);
void parallel_rows(int n, int threads, const std::function<void(int, int, int)> &fn);

}  // namespace amg
