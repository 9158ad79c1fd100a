"""Geometry of the generated workloads (CPU): the squashed cubes of the ill-conditioned
blob and a noisy oracle run on it without inversion."""
import numpy as np

from ffea_ci_demo_amd import workloads


def _vol(pos, top):
    p = pos[top[:, :4]]
    return np.linalg.det(p[:, 1:] - p[:, :1]) / 6.0


def test_config4_squashed_cubes():
    for n in (8, 40):
        spec = workloads.config4(n=n)
        v = np.abs(_vol(spec.pos, spec.top))
        med = np.median(v)
        thin = v < 0.2 * med
        frac = thin.mean()
        # about 2% of the tets squashed (the window's cubes), plus wedges at its rim
        assert 0.015 <= frac <= 0.045, (n, frac)
        # none close to degenerate: a squashed tet keeps ~1/20 of a regular one's volume
        assert v.min() >= 0.02 * med, (n, v.min() / med)
        # all tets keep the orientation of the unsquashed mesh
        assert np.all(_vol(spec.pos, spec.top) * _vol(workloads.cube_blob(n).pos, spec.top) > 0)


def test_config4_noisy_steps_without_inversion(oracle_mod):
    """300 noisy FP64 oracle steps on the 8^3 ill-conditioned blob: no element inverts and
    the squashed tets' volumes stay well away from zero"""
    spec = workloads.config4(n=8, rng_seed=5)
    v0 = _vol(spec.pos, spec.top)
    o = oracle_mod.Blob(spec.pos, spec.top, spec.A, spec.B, spec.G, spec.E, **spec.oracle_kwargs())
    lo = np.inf
    for _ in range(300):
        assert o.step() == 0
        lo = min(lo, (_vol(o.positions(), spec.top) / v0).min())
    assert lo > 0.3
