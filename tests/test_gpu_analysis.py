"""Axial ion-density profile on the device (tools/concentration.cpp:32-101)
against the oracle's restatement of the tool: bit-exact, since every node's
sum is the tool's sequential sum in ion order."""

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _edge_positions(L, n, rng, N=200_000):
    dz = L / (n - 1)
    z = rng.uniform(-0.55 * L, 0.55 * L, N)
    nodes = -L / 2 + np.arange(n) * L / (n - 1)
    special = np.concatenate([nodes, [-L / 2, L / 2, -L / 2 + 0.5 * dz, L / 2 - 0.5 * dz, -L, L,
                                      np.nextafter(-L / 2, -1.0), np.nextafter(L / 2, 1.0)]])
    z[: special.size] = special
    return z * 1e9


@pytest.mark.parametrize("n", [100, 7])
def test_density_profile_host_positions_bit_exact(oracle_lib, n):
    from opendust_amd import engine

    R, L, N_s = 1.4940882723609578e-3, 5.976353089443831e-3, 63.95267339533143
    z = _edge_positions(L, n, np.random.default_rng(n))
    say("density profile, host positions, nodes", n, "ions", z.size)
    zg, c = engine.density_profile_z(z, R, L, N_s, n)
    zo, co = oracle_lib.density_profile(z, R, L, N_s, n)
    assert np.array_equal(zg, zo)
    assert np.array_equal(c, co)


def test_density_profile_of_context_and_xyz_roundtrip(oracle_lib, tmp_path):
    """The context's profile equals the oracle's on the same ions, and equals
    the profile of the XYZ frame the simulation writes (positions printed at
    full precision read back exactly)."""
    import contextlib
    import io

    from opendust_amd import engine, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=30000, n=10, precision="fp32c", seed=3)
    alpha = workloads.synthetic_alpha(od)
    cfg, x0, v0 = workloads.prepare(od, alpha=alpha, np_seed=5)
    sp = od.simulationParametersInSIUnits
    say("density profile of a context, N", cfg.n_ions)
    e = engine.Engine(cfg)
    e.set_ions(x0, v0)
    e.run(3)
    x, _ = e.get_ions()
    zg, c = e.density_profile(sp.R, sp.H, sp.N_s, 100)
    e.close()
    zo, co = oracle_lib.density_profile(x[:, 2], sp.R, sp.H, sp.N_s, 100)
    assert np.array_equal(zg, zo) and np.array_equal(c, co)
    # XYZ frame written by the output path, parsed back
    path = tmp_path / "frame.xyz"
    od.outputParameters.xyzOutputFileName = str(path)
    od._write_frame(0, x)
    lines = path.read_text().splitlines()
    assert int(lines[0]) == od.N and lines[1] == "0"
    z_file = np.array([float(l.split("\t")[3]) for l in lines[2:]])
    assert np.array_equal(z_file, x[:, 2])
    zf, cf = engine.density_profile_z(z_file, sp.R, sp.H, sp.N_s, 100)
    assert np.array_equal(cf, c)
