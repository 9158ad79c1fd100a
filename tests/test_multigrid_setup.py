"""Host-side multigrid set-up (ffea_gpu_multigrid_inspect, no device), on viscosity
matrices the oracle assembles in the reference's summation order.

What pins it: the coarse operator is the Galerkin product P0^T K P0 (recomputed here
with scipy); the prolongator reproduces the 6 rigid modes exactly (smoothing keeps any
vector K annihilates), so a free blob's coarse operator annihilates their coarse
representation; pinned rows (identity blocks) get no coarse unknowns; level 0's smoother
weight follows the largest eigenvalue of Bd^-1 K, Bd its 3x3 diagonal blocks (scipy's
eigensolver).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from ffea_ci_demo_amd import _native, workloads


def corner_bsr(oracle_mod, spec):
    ob = oracle_mod.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    ob.update_internal_forces()
    ob.build_matrix()
    rp, col, val = ob.matrix()
    K = sp.csr_matrix((val, col, rp), shape=(3 * ob.nt, 3 * ob.nt))
    corner = np.unique(spec.top[:, :4])
    dof = (3 * corner[:, None] + np.arange(3)).reshape(-1)
    Kc = K[dof][:, dof].tocsr()
    B = sp.bsr_matrix(Kc, blocksize=(3, 3))
    B.sort_indices()
    pin = np.zeros(len(corner), np.uint8)
    lin = {v: i for i, v in enumerate(corner)}
    for p in spec.pinned:
        if int(p) in lin:  # midside pins have no corner row
            pin[lin[int(p)]] = 1
    return B, ob.positions()[corner], pin, Kc


def rigid(x):
    c = x - x.mean(0)
    Z = np.zeros((3 * len(x), 6))
    for k in range(3):
        Z[k::3, k] = 1
    Z[1::3, 3], Z[2::3, 3] = -c[:, 2], c[:, 1]
    Z[0::3, 4], Z[2::3, 4] = c[:, 2], -c[:, 0]
    Z[0::3, 5], Z[1::3, 5] = -c[:, 1], c[:, 0]
    return Z


def inspect(B, x, pin, free):
    d, (prp, pcol, pval), (arp, acol, aval) = _native.multigrid_inspect(B.indptr, B.indices, B.data, x, pin, free,
                                                                        True)
    P = sp.bsr_matrix((pval, pcol, prp), shape=(3 * d["rows"][0], 6 * d["rows"][1])).tocsr()
    A1 = sp.bsr_matrix((aval, acol, arp), shape=(6 * d["rows"][1], 6 * d["rows"][1])).toarray()
    return d, P, A1


@pytest.mark.parametrize("n", [8, 11])
def test_free_blob_galerkin_and_rigid_modes(oracle_mod, n):
    spec = workloads.cube_blob(n, rng_seed=3)
    B, x, pin, K = corner_bsr(oracle_mod, spec)
    d, P, A1 = inspect(B, x, pin, True)
    assert d["num_levels"] >= 2 and d["rows"][0] == len(x)
    assert all(d["rows"][l + 1] < d["rows"][l] for l in range(d["num_levels"] - 1))
    G = (P.T @ K @ P).toarray()
    assert np.abs(A1 - G).max() <= 1e-12 * np.abs(G).max()
    Z = rigid(x)
    Bc = np.linalg.lstsq(P.toarray(), Z, rcond=None)[0]
    assert np.abs(P @ Bc - Z).max() <= 1e-10 * np.abs(Z).max()
    assert np.abs(A1 @ Bc).max() <= 1e-10 * np.abs(A1).max() * np.abs(Bc).max()
    # level 0 smooths with its 3x3 diagonal blocks Bd: the weight follows the largest
    # eigenvalue of Bd^-1 K (K x = lam Bd x, scipy's eigensolver), below that of D^-1 K
    Bd = sp.block_diag([B.data[k] for i in range(B.shape[0] // 3) for k in range(B.indptr[i], B.indptr[i + 1])
                        if B.indices[k] == i]).tocsc()
    Bd = 0.5 * (Bd + Bd.T)
    lam = sla.eigsh(K, M=Bd, k=1, which="LA", return_eigenvectors=False)[0]
    assert 0.95 * lam <= d["lmax"][0] <= 1.0001 * lam
    Dm = sp.diags(1.0 / np.sqrt(K.diagonal()))
    lam_d = sla.eigsh(Dm @ K @ Dm, k=1, which="LA", return_eigenvectors=False)[0]
    assert lam < lam_d
    assert d["omega"][0] * 1.3 * d["lmax"][0] * (1 + 1 / 30) == pytest.approx(2.0)


def test_pinned_rows_have_no_coarse_unknowns(oracle_mod):
    spec = workloads.config3(n=8)
    B, x, pin, K = corner_bsr(oracle_mod, spec)
    assert pin.sum() > 0
    d, P, A1 = inspect(B, x, pin, False)
    rows = np.flatnonzero(np.repeat(pin, 3))
    assert P[rows].nnz == 0
    G = (P.T @ K @ P).toarray()
    assert np.abs(A1 - G).max() <= 1e-12 * np.abs(G).max()
    # a pinned blob's coarse operator is positive definite on its live unknowns
    live = np.flatnonzero(np.diag(A1) > 0)
    assert np.linalg.eigvalsh(A1[np.ix_(live, live)]).min() > 0


def test_bad_input_is_refused():
    with pytest.raises(RuntimeError):
        _native.multigrid_inspect(np.array([0, 1], np.int32), np.array([5], np.int32), np.eye(3).reshape(1, 9),
                                  np.zeros((1, 3)))


@pytest.mark.parametrize("n", [10, 19])
def test_single_blob_hierarchy_has_a_coarse_level(oracle_mod, n):
    """the inspected hierarchy is the one a single-blob handle builds: a blob small enough
    for a dense coarsest level (1331 corner nodes, 3993 unknowns) still coarsens once,
    and a larger one (8000 nodes) coarsens until its coarsest level is at most 4800
    unknowns"""
    spec = workloads.config4(n=n, rng_seed=23)
    B, x, pin, K = corner_bsr(oracle_mod, spec)
    d, P, A1 = inspect(B, x, pin, True)
    assert d["num_levels"] >= 2
    assert d["coarsest_unknowns"] <= min(4800, 3 * len(x) // 8)
    assert d["coarsest_unknowns"] == d["rows"][d["num_levels"] - 1] * d["block"][d["num_levels"] - 1]
