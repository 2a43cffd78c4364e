"""The allocation-free density profile (od_density_profile_noalloc) against
od_density_profile, through the C ABI: bit-identical node positions and
concentrations on two calls in a row, the second on a smaller context with
fewer nodes (the shared per-device scratch keeps the first call's contents),
then once more on the first context; the node maximum is enforced."""

import contextlib
import io

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _context(N, seed):
    from opendust_amd import engine, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=N, n=2, precision="fp64", seed=seed)
        cfg, x0, v0 = workloads.prepare(od, alpha=workloads.synthetic_alpha(od), np_seed=seed, precision="fp64")
    e = engine.Engine(cfg)
    e.set_ions(x0, v0)
    return od, e


def test_noalloc_density_profile_matches_allocating_entry_point():
    from opendust_amd import engine

    engine.load_library()
    od_a, a = _context(6000, 3)
    od_b, b = _context(1500, 4)
    try:
        cases = [(a, od_a, 200), (b, od_b, 50), (a, od_a, 200)]
        args = []
        for e, od, n in cases:
            sp = od.simulationParametersInSIUnits
            args.append((e, (sp.R, sp.H, sp.N_s, n)))
        refs = [e.density_profile(*p) for e, p in args]
        # the allocation-free calls back to back: larger, smaller, larger again
        outs = []
        for e, p in args:
            say("noalloc density profile: ions", e.n, "nodes", p[3])
            outs.append(e.density_profile_noalloc(*p))
        for (zg_ref, c_ref), (zg, c) in zip(refs, outs):
            assert np.array_equal(zg, zg_ref)
            assert np.array_equal(c, c_ref)
            assert np.all(np.isfinite(c)) and np.any(c != 0.0)
        R, L, N_s = args[1][1][:3]
        with pytest.raises(RuntimeError, match="65536"):
            a.density_profile_noalloc(R, L, N_s, 65537)
        # the largest node count the scratch holds still works
        zg_ref, c_ref = b.density_profile(R, L, N_s, 65536)
        zg, c = b.density_profile_noalloc(R, L, N_s, 65536)
        assert np.array_equal(zg, zg_ref) and np.array_equal(c, c_ref)
    finally:
        a.close()
        b.close()
