"""Resident plasma solves for dust-dynamics loops (od_update_config,
opendust_amd.dynamics.ResidentPlasma) against the FP64 CPU oracle: after the
dusts move inside a live context, the next steps equal the oracle run with
the new configuration from the ions' current state.  Tolerances as in the
FP64 multi-step tests: absorbed counts exact, charges rel <= 1e-12, orbit and
collection forces rel <= 1e-9, positions <= 1e-10 R (the engine continues with
its incrementally updated ion-ion force, the oracle recomputes it, which
differs at FP64 rounding)."""

import contextlib
import io

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu


def _compare(cfg, rec, rec_o, x_g, x_o):
    assert np.array_equal(rec["absorbed"], rec_o["absorbed"])
    assert np.max(np.abs(rec["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12
    for key in ("force_orbit", "force_collect"):
        sc = np.max(np.abs(rec_o[key]))
        assert np.max(np.abs(rec[key] - rec_o[key])) <= 1e-9 * sc, key
    assert np.max(np.abs(x_g - x_o)) <= 1e-10 * cfg.R


def test_update_config_moves_dusts_in_a_live_context(oracle_lib):
    from opendust_amd import engine, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.dust_chain(N=2048, n=10, precision="fp64", seed=5)
    alpha = workloads.synthetic_alpha(od)
    cfg, x0, v0 = workloads.prepare(od, alpha=alpha, np_seed=2)
    say("update_config, dust chain fp64 N", cfg.n_ions)
    e = engine.Engine(cfg)
    e.set_ions(x0, v0)
    e.run(3, engine.RUN_GRAPH)
    x1, v1 = e.get_ions()
    cfg2 = engine.Config.from_buffer_copy(cfg)
    lam = cfg.r_D_e
    for p in range(cfg.n_dust):
        cfg2.dust_z[p] = cfg.dust_z[p] + (0.2 + 0.1 * p) * lam
        cfg2.dust_x[p] = cfg.dust_x[p] + 0.05 * lam
        cfg2.si_dust_z[p] = cfg2.dust_z[p] * 1e-9
        cfg2.si_dust_x[p] = cfg2.dust_x[p] * 1e-9
    cfg2.seed = 12345
    q2 = np.array([-1.5e4 * 1.602e-19 * (1 + 0.1 * p) for p in range(cfg.n_dust)])
    e.update_config(cfg2)
    e.set_dust_charges(q2)
    e.set_step(0)
    rec = e.run(5, engine.RUN_GRAPH)
    x2, _ = e.get_ions()
    e.close()
    x_o, v_o, q_o, rec_o = oracle_lib.run(cfg2, x1, v1, q2, 0, 5)
    _compare(cfg2, rec, rec_o, x2, x_o)
    # geometry that buffers depend on cannot change
    e = engine.Engine(cfg)
    bad = engine.Config.from_buffer_copy(cfg)
    bad.n_ions = cfg.n_ions + 1
    with pytest.raises(RuntimeError, match="must stay the same"):
        e.update_config(bad)
    e.close()


def test_resident_plasma_solves_vs_oracle(oracle_lib):
    from opendust_amd import dynamics, engine, workloads
    from opendust_amd.opendust import DustParticle

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=2048, n=6, precision="fp64", seed=9)
    alpha = workloads.synthetic_alpha(od)
    np.random.seed(4)
    rp = dynamics.ResidentPlasma(od, "0", alphaTrap=alpha)
    lam = od.plasmaParametersInSIUnits.r_D_e
    e = od.plasmaParametersInSIUnits.e
    say("resident plasma, two solves, N", od.N)
    r1 = rp.solve([DustParticle(0.0, 0.0, 0.0, 0.05 * lam, "consistent", -2e4 * abs(e))])
    assert r1.dustParticles[0].q[0] == -2e4 * abs(e)
    x1, v1 = rp.engine.get_ions()
    dust2 = [DustParticle(0.0, 0.0, 0.4 * lam, 0.06 * lam, "consistent", float(r1.dustParticles[0].q[-1]))]
    r2 = rp.solve(dust2, n=4)
    rp.close()
    cfg2 = r2.make_config(rp.tables, rp.alphaTrap, seed=r2.seed)
    x_o, v_o, q_o, rec_o = oracle_lib.run(cfg2, x1, v1, np.array([dust2[0].q]), 0, 4)
    d = r2.dustParticles[0]
    assert d.q.shape == (4,) and d.forceIonsOrbit.shape == (4, 3)
    assert np.max(np.abs(d.q - rec_o["q"][:, 0]) / np.abs(rec_o["q"][:, 0])) <= 1e-12
    sc = np.max(np.abs(rec_o["force_orbit"][:, 0]))
    assert np.max(np.abs(d.forceIonsOrbit - rec_o["force_orbit"][:, 0])) <= 1e-9 * sc
    assert r2.seed != r1.seed
