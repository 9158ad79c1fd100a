"""The reference's equipartition check of the mass-less path with noise
(tests/physics/sphere_63_120_nomass/testAverages.py:51-71): the sphere_63_120 blob (63
linear nodes, Stokes drag, thermal noise) run for 1e5 steps with a measurement every 1000;
the mean StrainEnergy over measurements 40 to the last but one must be within 3% of
kT (3 * 63 - 6) / 2, kT = 4.11e-21 J. The script is the reference's
sphere_63_120_nomass.ffea (the same <param> values, rng_seed 44) over the reference's
mesh files; this is the end-to-end check of the noise amplitude, the unit conversions and
the element draws feeding the stress entries."""
import numpy as np
import pytest

from ffea_ci_demo_amd.world import World
from test_gpu_world import sphere_script

pytestmark = pytest.mark.gpu

KBT = 4.11e-21
E_TH = KBT * (3 * 63 - 6) / 2


def read_strain_energy(path):
    """the StrainEnergy column of a measurement file, RESTART lines skipped
    (testAverages.py:27-48)"""
    with open(path) as fh:
        lines = fh.read().splitlines()
    k = lines.index("Measurements:")
    head = lines[k + 1].split()
    col = head.index("StrainEnergy")
    return np.array([float(l.split()[col]) for l in lines[k + 2:] if l.strip() and "RESTART" not in l])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_equipartition_sphere_63_120(tmp_path, precision):
    path = sphere_script(tmp_path, "sphere_63_120_nomass", num_steps="1e5", check="1e3")
    print(f"[case] equipartition sphere_63_120, 1e5 steps, {precision}", flush=True)
    w = World(str(path), precision=precision)
    w.run()
    w.close()
    se = read_strain_energy(tmp_path / "sphere_63_120_nomass_measurement.out")
    assert se.size >= 100, se.size
    sample = se[40:-1]  # ini = 40, end = -1 (testAverages.py:56-57)
    dev = abs(sample.mean() / E_TH - 1)
    print(f"[equipartition] {precision}: mean StrainEnergy {sample.mean():.4e} J over {sample.size} measurements, "
          f"theory {E_TH:.4e} J, relative deviation {dev:.4f}", flush=True)
    assert dev < 0.03
