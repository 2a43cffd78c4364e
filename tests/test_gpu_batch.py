"""Batched sweeps (od_batch_*, opendust_amd.sweep) against the FP64 CPU oracle.

Each simulation of a batch is checked against the oracle run alone on the same
configuration (its sim_id included, which selects its Philox counter word) and
the same initial state.  Tolerances are those of the single-simulation tests:
FP64 records rel <= 1e-9 (charges 1e-12), positions <= 1e-10 R; FP32c after one
step per-ion |dx_i| within the drift of the pair-force error bound, |dv| <=
1e-4 max|v|, records as conftest.check_records states; flags, faces and
absorbed counts bit-exact.  The batch mixes ion counts that are not multiples
of the 1024-ion padding, 1-4 dusts and different Mach numbers.
"""

import numpy as np
import pytest

from conftest import check_records, fp32c_position_error_ok, say

pytestmark = pytest.mark.gpu


def _batch_setup(precision, n_sims=3, n=20, seed=15):
    import contextlib
    import io

    from opendust_amd import sweep, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        ods = workloads.sweep(n_sims=n_sims, n=n, precision=precision, seed=seed, log2_n=(10.4, 12.1))
    alphas = [workloads.synthetic_alpha(od) for od in ods]
    np.random.seed(3)
    cfgs, states = sweep.prepare_batch(ods, device=0, alphas=alphas)
    # craft ions into the dusts and beyond the side wall so that the first step
    # absorbs and reinjects
    rng = np.random.RandomState(11)
    for cfg, (x0, v0) in zip(cfgs, states):
        for i in range(8):
            p = i % cfg.n_dust
            d = rng.normal(size=3)
            d *= 0.5 * cfg.dust_r[p] / np.linalg.norm(d)
            x0[i] = [cfg.dust_x[p] + d[0], cfg.dust_y[p] + d[1], cfg.dust_z[p] + d[2]]
        for i in range(8, 14):
            a = rng.uniform(0, 2 * np.pi)
            x0[i] = [1.02 * cfg.R * np.cos(a), 1.02 * cfg.R * np.sin(a), rng.uniform(-0.4, 0.4) * cfg.H]
    return ods, cfgs, states


def _q0(cfg):
    return np.array([cfg.dust_q0[p] for p in range(cfg.n_dust)])


def test_batch_setup_is_uneven():
    ods, cfgs, _ = _batch_setup("fp64")
    ns = [c.n_ions for c in cfgs]
    say("batch ion counts", ns, "dusts", [c.n_dust for c in cfgs])
    assert len(set(ns)) == len(ns) and all(n % 1024 for n in ns)
    assert len({c.n_dust for c in cfgs}) > 1
    assert [c.sim_id for c in cfgs] == list(range(len(cfgs)))


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_batch_one_step_vs_oracle(oracle_lib, precision):
    from opendust_amd import engine

    ods, cfgs, states = _batch_setup(precision)
    say("batch one step", precision, [c.n_ions for c in cfgs])
    b = engine.Batch(cfgs)
    for s, (x0, v0) in enumerate(states):
        b.set_ions(s, x0, v0)
    rec = b.run(1)
    for s, (cfg, (x0, v0)) in enumerate(zip(cfgs, states)):
        x_g, v_g = b.get_ions(s)
        fl_g, face_g = b.step_info(s)
        x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, _q0(cfg), 0)
        nd = cfg.n_dust
        assert np.array_equal(fl_g, fl_o), s
        assert np.array_equal(face_g, face_o), s
        assert int((fl_o != 0).sum()) >= 10
        assert np.array_equal(rec["absorbed"][0, s, :nd], rec_o["absorbed"]), s
        dx = np.max(np.abs(x_g - x_o)) / cfg.R
        dv = np.max(np.abs(v_g - v_o)) / np.max(np.abs(v_o))
        say("  sim", s, "N", cfg.n_ions, "nd", nd, "max |dx|/R", dx, "max |dv|/max|v|", dv)
        bound = None
        if precision == "fp64":
            assert dx <= 1e-11
        else:
            ok, worst, bound = fp32c_position_error_ok(oracle_lib, cfg, x0, x_g, x_o, fl_o)
            say("  worst |dx_i| / bound_i", worst)
            assert ok, s
        assert dv <= (1e-10 if precision == "fp64" else 1e-4)
        ratios = check_records(cfg, rec[0, s, :nd], rec_o, precision, x_o, bound)
        say("  records, error / tolerance:", ratios)
        for key, r in ratios.items():
            assert r <= 1.0, (s, key)
    b.close()


def test_batch_multi_step_fp64_vs_oracle(oracle_lib):
    from opendust_amd import engine

    ods, cfgs, states = _batch_setup("fp64")
    say("batch multi step fp64", [c.n_ions for c in cfgs])
    b = engine.Batch(cfgs)
    for s, (x0, v0) in enumerate(states):
        b.set_ions(s, x0, v0)
    rec = b.run(6)
    for s, (cfg, (x0, v0)) in enumerate(zip(cfgs, states)):
        nd = cfg.n_dust
        x_g, _ = b.get_ions(s)
        x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, _q0(cfg), 0, 6)
        r = rec[:, s, :nd]
        assert np.array_equal(r["absorbed"], rec_o["absorbed"]), s
        assert np.max(np.abs(r["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12, s
        for key in ("force_orbit", "force_collect"):
            sc = np.max(np.abs(rec_o[key]))
            assert np.max(np.abs(r[key] - rec_o[key])) <= 1e-9 * sc, (s, key)
        assert np.max(np.abs(x_g - x_o)) <= 1e-10 * cfg.R, s
    b.close()


def test_batch_determinism_and_graph():
    from opendust_amd import engine

    ods, cfgs, states = _batch_setup("fp32c")
    say("batch determinism fp32c", [c.n_ions for c in cfgs])
    outs = []
    for flags in (0, engine.RUN_GRAPH, engine.RUN_GRAPH):
        b = engine.Batch(cfgs)
        for s, (x0, v0) in enumerate(states):
            b.set_ions(s, x0, v0)
        rec = b.run(37, flags)
        xs = [b.get_ions(s) for s in range(len(cfgs))]
        b.close()
        outs.append((rec, xs))
    for rec, xs in outs[1:]:
        for s, cfg in enumerate(cfgs):
            assert np.array_equal(rec[:, s, :cfg.n_dust], outs[0][0][:, s, :cfg.n_dust]), s
        for (x, v), (x0, v0) in zip(xs, outs[0][1]):
            assert np.array_equal(x, x0) and np.array_equal(v, v0)
        # slots past a simulation's dusts are zero
        assert np.array_equal(rec, outs[0][0])


def test_simulate_batch_api_vs_oracle(oracle_lib):
    """simulateBatch fills every OpenDust object's result arrays as simulate()
    does; checked against the oracle run per simulation."""
    import contextlib
    import io

    from opendust_amd import sweep, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        ods = workloads.sweep(n_sims=2, n=7, precision="fp64", seed=9, log2_n=(10.0, 11.0))
    alphas = [workloads.synthetic_alpha(od) for od in ods]
    np.random.seed(4)
    sweep.simulateBatch(ods, "0", alphas=alphas)
    np.random.seed(4)
    cfgs, states = sweep.prepare_batch(ods, device=0, alphas=alphas)
    for s, (od, cfg, (x0, v0)) in enumerate(zip(ods, cfgs, states)):
        x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, _q0(cfg), 0, od.n)
        say("  simulateBatch sim", s, "N", cfg.n_ions, "dusts", cfg.n_dust)
        for p, d in enumerate(od.dustParticles):
            assert np.max(np.abs(d.q - rec_o["q"][:, p]) / np.abs(rec_o["q"][:, p])) <= 1e-12
            sc = np.max(np.abs(rec_o["force_orbit"][:, p]))
            assert np.max(np.abs(d.forceIonsOrbit - rec_o["force_orbit"][:, p])) <= 1e-9 * sc
            assert np.array_equal(d.forceDust.shape, (od.n, 3))


def test_batch_tune_is_reproducible():
    """The autotuner's choice fixes the schedule: a tuned batch and a batch
    created with that sym_chunk and sym_tile give bitwise identical steps."""
    from opendust_amd import engine

    ods, cfgs, states = _batch_setup("fp32c")
    say("batch tune", [c.n_ions for c in cfgs])
    outs = []
    chosen = None
    for tune in (True, False):
        if not tune:
            cfgs[0].sym_chunk, cfgs[0].sym_tile = chosen
        b = engine.Batch(cfgs)
        for s, (x0, v0) in enumerate(states):
            b.set_ions(s, x0, v0)
        if tune:
            chosen = b.tune(2)
            say("  chosen sym_chunk, sym_tile", chosen)
            assert chosen[0] in (1, 2, 4, 8, 16, 32)
            assert 1 <= chosen[1] <= len(engine.sym_variants())
            # row groups need one simulation: the batch tuner never picks them
            assert engine.sym_variants()[chosen[1] - 1][3] == 1
        rec = b.run(12, engine.RUN_GRAPH)
        outs.append((rec, [b.get_ions(s) for s in range(len(cfgs))]))
        b.close()
    cfgs[0].sym_chunk = cfgs[0].sym_tile = 0
    assert np.array_equal(outs[0][0], outs[1][0])
    grouped = [i + 1 for i, v in enumerate(engine.sym_variants()) if v[3] > 1]
    assert grouped
    cfgs[0].sym_tile = grouped[0]
    with pytest.raises(RuntimeError, match="groups target blocks"):
        engine.Batch(cfgs)
    cfgs[0].sym_tile = 0
    for (x, v), (x1, v1) in zip(outs[0][1], outs[1][1]):
        assert np.array_equal(x, x1) and np.array_equal(v, v1)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_batch_mixed_geometry_and_distributions_vs_oracle(oracle_lib, precision):
    """Simulations that differ in everything a configuration holds: cylinder
    size, Maxwellian vs field-driven flow with collisions, dust count, ion
    count; each checked against the oracle alone."""
    import contextlib
    import io

    from opendust_amd import engine, sweep, workloads

    with contextlib.redirect_stdout(io.StringIO()):
        ods = [workloads.single_dust(N=2500, n=10, precision=precision, seed=31),
               workloads.dust_chain(N=3100, n=10, precision=precision, seed=32),
               workloads.piel_fig1(N=1200, n=10, precision=precision, seed=33, q_e=1e5)]
    alphas = [workloads.synthetic_alpha(od) for od in ods]
    np.random.seed(8)
    cfgs, states = sweep.prepare_batch(ods, device=0, alphas=alphas)
    assert len({(c.R, c.H) for c in cfgs}) == 3 and len({c.distribution for c in cfgs}) == 2
    say("batch mixed", precision, [(c.n_ions, c.n_dust, c.distribution) for c in cfgs])
    b = engine.Batch(cfgs)
    for s, (x0, v0) in enumerate(states):
        b.set_ions(s, x0, v0)
    nsteps = 4 if precision == "fp64" else 1
    rec = b.run(nsteps)
    for s, (cfg, (x0, v0)) in enumerate(zip(cfgs, states)):
        nd = cfg.n_dust
        x_g, v_g = b.get_ions(s)
        x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, _q0(cfg), 0, nsteps)
        r = rec[:, s, :nd]
        assert np.array_equal(r["absorbed"], rec_o["absorbed"]), s
        if precision == "fp64":
            assert np.max(np.abs(x_g - x_o)) <= 1e-10 * cfg.R, s
            assert np.max(np.abs(r["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12, s
            sc = np.max(np.abs(rec_o["force_orbit"]))
            assert np.max(np.abs(r["force_orbit"] - rec_o["force_orbit"])) <= 1e-9 * sc, s
        else:
            assert np.max(np.abs(v_g - v_o)) <= 1e-4 * np.max(np.abs(v_o)), s
            assert np.max(np.abs(r["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12, s
    b.close()
