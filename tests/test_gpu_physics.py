"""End-to-end statistical parity with the reference's own published runs.

examples/Piel2017/Figure1/{consistent,oml}/launchPiel2017Figure1*.py run
OpenDust on OpenMM (Ar+, T_e = 29011 K, T_i = 290.11 K, n = 1e14 m^-3, M = 1,
R = 3, H = 6 r_D_e, 2^16 superions, 3000 steps of 3.5148240854e-9 s, one dust
of radius 58.8 um at z = -r_D_e, charge from 0) and print the mean z-forces on
the dust over steps i > 1000 (outputPiel2017Fig1Consistent:40-41,
outputPiel2017Fig1OML:40-41).  The same scripts run here through the drop-in
API (FP32c, other random numbers).  Both runs are single stochastic
realisations, so the check is statistical: with sigma the standard error of
our mean from 10 block averages of 200 steps, the reference's mean must lie
within 5 sqrt(2) sigma of ours (two independent runs of equal variance; 5
standard deviations leaves room for the 10-block estimate of sigma).
"""

import contextlib
import io

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

# mean z-forces (N) over i > 1000 printed by the reference runs
PIEL_FIG1 = {
    "consistent": (4.495695908591751e-11, 8.97145418766947e-12),  # outputPiel2017Fig1Consistent:40-41
    "oml": (4.48957607439666e-11, 9.203742146333308e-12),         # outputPiel2017Fig1OML:40-41
}


def _block_mean(a, n_blocks=10):
    a = np.asarray(a, dtype=np.float64)
    k = len(a) // n_blocks
    blocks = a[len(a) - k * n_blocks:].reshape(n_blocks, k).mean(axis=1)
    return float(a.mean()), float(blocks.std(ddof=1) / np.sqrt(n_blocks))


@pytest.mark.parametrize("method", ["consistent", "oml"])
def test_piel2017_fig1_mean_forces(method):
    from opendust_amd import workloads

    say("Piel2017 Fig. 1", method, "N 65536, 3000 steps, fp32c")
    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.piel_fig1(N=2 ** 16, n=3000, charge=method, q_e=0.0, precision="fp32c", seed=2017,
                                 trapSeed=42)
        np.random.seed(17)
        od.simulate("0")
    d = od.dustParticles[0]
    orbit, s_orbit = _block_mean(d.forceIonsOrbit[1001:, 2])
    collect, s_collect = _block_mean(d.forceIonsCollect[1001:, 2])
    ref_orbit, ref_collect = PIEL_FIG1[method]
    z_orbit = (orbit - ref_orbit) / (np.sqrt(2.0) * s_orbit)
    z_collect = (collect - ref_collect) / (np.sqrt(2.0) * s_collect)
    say("  orbit %.4e +- %.2e (reference %.4e, z = %.2f)" % (orbit, s_orbit, ref_orbit, z_orbit))
    say("  collect %.4e +- %.2e (reference %.4e, z = %.2f)" % (collect, s_collect, ref_collect, z_collect))
    say("  charge at the end %.4e C" % d.q[-1])
    assert np.all(np.isfinite(d.q)) and d.q[-1] < 0
    assert abs(z_orbit) <= 5.0
    assert abs(z_collect) <= 5.0
