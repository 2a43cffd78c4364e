"""Engine (HIP, through the C ABI) against the FP64 CPU oracle on identical
seeded states.

Tolerances (stated here, checked below):
  * FP64 engine vs oracle: pair forces per ion rel <= 1e-11; one step x, v
    rel <= 1e-11 (positions relative to the domain size); dust records rel
    <= 1e-9.
  * FP32-compensated engine vs FP64 oracle.  Error model: positions enter the
    pair kernel rounded to FP32 in units of r_D_e/log2(e) (|u| <= ~10), so a
    pair at distance r carries a relative error ~ ulp(|u|)/r_u ~ 1e-6/r_u on
    top of a few FP32 ulps of the arithmetic; the compensated sums add nothing
    visible.  Close pairs dominate the force, hence: pair forces norm-wise rel
    <= 1e-5, per ion |dF_i| <= 1e-4 |F_i| + 1e-5 rms|F|; after one step
    per ion |dx_i| <= dt (dt/2) / m_s (1e-4 |F_i| + 1e-5 rms|F|) (the drift of
    that force error; reinjected ions to FP64 rounding), |dv| <= 1e-4 max|v|,
    dust charges, collection and dust-dust forces rel <= 1e-5 of their scale,
    orbit sums and forces within the propagation of the per-ion position
    bounds through the orbit term (conftest.check_records).  These force
    bounds hold at the test sizes; at 2^20 ions the position quantization
    grows with density, and test_full_size_sym_forces_vs_fp64_gpu states its
    own split of arithmetic and quantization errors.
  * Integer/index outputs (flags, reinjection faces, absorbed counts) are
    bit-exact.
"""

import contextlib
import io

import numpy as np
import pytest

from conftest import check_records, fp32c_position_error_ok, say

pytestmark = pytest.mark.gpu


def _engine_mod():
    from opendust_amd import engine

    engine.load_library()
    return engine


MODES = {"rows": 1, "sym": 2}


def _setup(N=2048, precision="fp64", seed=7, np_seed=3, workload="single_dust", craft=True, mode="rows"):
    from opendust_amd import workloads

    od = getattr(workloads, workload)(N=N, n=100, precision=precision, seed=seed)
    alpha = workloads.synthetic_alpha(od)
    cfg, x0, v0 = workloads.prepare(od, alpha=alpha, np_seed=np_seed, precision=precision)
    cfg.pair_mode = MODES[mode]
    if craft:
        # ions 0..5 inside the dust(s), 6..11 outside the side wall, 12..17
        # beyond the end caps: they are absorbed / flagged this step
        rng = np.random.RandomState(11)
        nd = cfg.n_dust
        for i in range(6):
            p = i % max(nd, 1)
            if nd:
                d = rng.normal(size=3)
                d *= 0.5 * cfg.dust_r[p] / np.linalg.norm(d)
                x0[i] = [cfg.dust_x[p] + d[0], cfg.dust_y[p] + d[1], cfg.dust_z[p] + d[2]]
        for i in range(6, 12):
            a = rng.uniform(0, 2 * np.pi)
            x0[i] = [1.02 * cfg.R * np.cos(a), 1.02 * cfg.R * np.sin(a), rng.uniform(-0.4, 0.4) * cfg.H]
        for i in range(12, 18):
            x0[i, 2] = (0.5 + 0.01) * cfg.H * (1 if i % 2 else -1)
    return od, cfg, x0, v0


def test_library_loads_on_gpu():
    eng = _engine_mod()
    assert eng.check_abi()


@pytest.mark.parametrize("precision,mode", [("fp64", "rows"), ("fp32c", "rows"), ("fp32c", "sym")])
def test_pair_forces_vs_oracle(oracle_lib, precision, mode):
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=3000, precision=precision, craft=False, mode=mode)
    say("pair forces", precision, mode, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    assert e.pair_info()[0] == MODES[mode]
    f_gpu = e.pair_forces(x0)
    e.close()
    f_orc = oracle_lib.pair_forces(cfg, x0)
    fn = np.linalg.norm(f_orc, axis=1)
    err = np.linalg.norm(f_gpu - f_orc, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    norm_rel = np.linalg.norm(f_gpu - f_orc) / np.linalg.norm(f_orc)
    say("  norm-wise rel", norm_rel, "max per-ion rel", np.max(err / fn))
    if precision == "fp64":
        assert np.all(err <= 1e-11 * fn + 1e-14 * rms)
    else:
        assert norm_rel <= 1e-5
        assert np.all(err <= 1e-4 * fn + 1e-5 * rms)


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_sym_variants_pair_forces_vs_oracle(oracle_lib, tile):
    """Every variant of the symmetric kernel the autotuner may pick (tile
    shape, waves per SIMD) meets the FP32c force bound against the oracle, for
    one and for several source blocks per work-group."""
    eng = _engine_mod()
    variants = eng.sym_variants()
    if tile > len(variants):
        pytest.skip("library has %d variants" % len(variants))
    od, cfg, x0, v0 = _setup(N=5000, precision="fp32c", craft=False, mode="sym")
    f_orc = oracle_lib.pair_forces(cfg, x0)
    fn = np.linalg.norm(f_orc, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    for ch in (1, 2):
        cfg.sym_tile, cfg.sym_chunk = tile, ch
        say("sym variant", tile, variants[tile - 1], "chunk", ch, "N", cfg.n_ions)
        e = eng.Engine(cfg)
        assert e.pair_info()[0] == MODES["sym"]
        f_gpu = e.pair_forces(x0)
        e.close()
        err = np.linalg.norm(f_gpu - f_orc, axis=1)
        say("  norm-wise rel", np.linalg.norm(f_gpu - f_orc) / np.linalg.norm(f_orc))
        assert np.linalg.norm(f_gpu - f_orc) <= 1e-5 * np.linalg.norm(f_orc)
        assert np.all(err <= 1e-4 * fn + 1e-5 * rms)
    bad = eng.Config.from_buffer_copy(cfg)
    bad.sym_tile = len(variants) + 1
    with pytest.raises(RuntimeError, match="sym_tile"):
        eng.Engine(bad)


@pytest.mark.parametrize("tile", [0, 10, 12])
def test_sym_schedule_in_bands(oracle_lib, tile):
    """The symmetric kernel's triangular schedule run in bands of rows (each
    launched and reduced in order into an FP64 accumulator; the default band
    is 48 GB of partials, here 1 MiB, so every band holds one or two rows),
    for the default variant and for row groups of 4 target blocks (variants
    10 and 12, register and LDS sources: one source partial per work-group
    for all 4 blocks' waves):
    forces within the FP32c bound of the oracle and within 1e-6 of the
    one-band schedule (only the summation order differs; the engine reports
    one band, then several), and a run is bitwise reproducible, eager and
    graph-replayed.

    At 20000 ions in the cylinder close pairs are common, so the rounding of
    the positions to FP32 dominates the error against exact positions; the
    kernel's arithmetic is checked against the oracle on the rounded
    positions (per ion 1e-5 |F_i| + 1e-6 rms|F|, compensated sums) and the
    total against exact positions only up to what that rounding alone moves
    (FP64 on rounded positions vs exact, + 5e-6)."""
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=20000, precision="fp32c", craft=False, mode="sym")
    cfg.sym_tile = tile
    f_orc = oracle_lib.pair_forces(cfg, x0)
    f_rnd = oracle_lib.pair_forces(cfg, oracle_lib.fp32c_rounded_positions(cfg, x0))
    fn = np.linalg.norm(f_rnd, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    rel_q = np.linalg.norm(f_rnd - f_orc) / np.linalg.norm(f_orc)
    forces = {}
    for mib in (0, 1):
        cfg.sym_band_mib = mib
        say("sym schedule, band MiB", mib, "N", cfg.n_ions)
        e = eng.Engine(cfg)
        assert e.pair_info()[0] == MODES["sym"]
        bands = e.sym_schedule()[2]
        say("  bands of rows", bands)
        assert (bands == 1) if mib == 0 else (bands > 2)
        forces[mib] = e.pair_forces(x0)
        e.close()
        err = np.linalg.norm(forces[mib] - f_rnd, axis=1)
        rel_a = np.linalg.norm(forces[mib] - f_rnd) / np.linalg.norm(f_rnd)
        rel_b = np.linalg.norm(forces[mib] - f_orc) / np.linalg.norm(f_orc)
        say("  arithmetic: norm-wise", rel_a, "worst per-ion err / bound", np.max(err / (1e-5 * fn + 1e-6 * rms)),
            "| vs exact positions", rel_b, "(rounding alone", rel_q, ")")
        assert rel_a <= 2e-6
        assert np.all(err <= 1e-5 * fn + 1e-6 * rms)
        assert rel_b <= rel_q + 5e-6
    # the FP64 sums of a few FP32-precision partials per ion are usually exact,
    # so the two orders mostly agree bit for bit
    d = np.linalg.norm(forces[1] - forces[0]) / np.linalg.norm(forces[0])
    say("  banded vs one band: norm-wise", d)
    assert d <= 1e-6
    # steps with ions crafted into the dust and beyond the walls
    od, cfg, x0, v0 = _setup(N=20000, precision="fp32c", craft=True, mode="sym")
    cfg.sym_band_mib = 1
    cfg.sym_tile = tile
    outs = []
    for flags in (0, 0, eng.RUN_GRAPH):
        e = eng.Engine(cfg)
        e.set_ions(x0, v0)
        rec = e.run(8, flags)
        outs.append((rec, e.get_ions()[0]))
        e.close()
    for rec, x in outs[1:]:
        assert np.array_equal(rec, outs[0][0]) and np.array_equal(x, outs[0][1])


def test_full_size_sym_forces_vs_fp64_gpu():
    """At the benchmark size, 2^20 ions, the FP32c symmetric kernel against the
    plain FP64 row kernel on the GPU (independent code: ordered pairs, FP64
    law, no partials), for the default variant and schedule (sources read
    from LDS, row groups of 4 target blocks, 8 source blocks per work-group),
    for row groups of 2 with LDS sources (variant 13) and of 4 with register
    sources (variant 10), for 4 targets per lane with 16 source blocks per
    work-group, and for one row per work-group (variant 6) with 4.

    Two error sources are separated.  (a) Kernel arithmetic: against FP64 on
    the same positions rounded to FP32 in the kernel's units (what the FP32
    kernel sees), every pair term carries a few FP32 roundings (2^-24 each)
    and the sums are compensated, so per ion |dF_i| <= 1e-5 |F_i| + 1e-6
    rms|F| and norm-wise <= 2e-6: every block pair of the triangular schedule
    is present and every partial lands on its ions.  (b) Position
    quantization: against FP64 on the exact positions the rounding of the
    coordinates (ulp 4.8e-7 in units of r_D_e/log2 e at |u| ~ 7) moves a pair
    at distance r by ~ulp/r relatively; at this density (nearest neighbours
    ~0.05 r_D_e, some far closer) that dominates: norm-wise <= 1e-4 (measured
    ~4e-5), per ion not bounded by the small-N model.  (The CPU oracle would
    need ~10 minutes at this size; it pins both kernels at small sizes.)"""
    eng = _engine_mod()
    from opendust_amd import workloads

    od = workloads.single_dust(N=2 ** 20, n=1, precision="fp64", seed=3)
    alpha = workloads.synthetic_alpha(od)
    cfg, x0, v0 = workloads.prepare(od, alpha=alpha, np_seed=5, precision="fp64")
    scale = 1.4426950408889634 / cfg.r_D_e
    x32 = (x0 * scale).astype(np.float32).astype(np.float64) / scale
    say("full size: fp64 rows N", cfg.n_ions)
    e = eng.Engine(cfg)
    f64 = e.pair_forces(x0)
    f64r = e.pair_forces(x32)
    e.close()
    cfg32, _, _ = workloads.prepare(od, alpha=alpha, np_seed=5, precision="fp32c")
    for tile, ch in ((0, 0), (13, 8), (3, 16), (6, 4), (10, 8)):
        cfg32.sym_tile, cfg32.sym_chunk = tile, ch
        e = eng.Engine(cfg32)
        assert e.pair_info()[0] == MODES["sym"]
        f32 = e.pair_forces(x0)
        e.close()
        fn = np.linalg.norm(f64r, axis=1)
        rms = np.sqrt(np.mean(fn ** 2))
        err = np.linalg.norm(f32 - f64r, axis=1)
        rel_a = np.linalg.norm(f32 - f64r) / np.linalg.norm(f64r)
        rel_b = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
        rel_q = np.linalg.norm(f64r - f64) / np.linalg.norm(f64)
        say("  fp32c sym tile", tile, "chunk", ch, "| arithmetic: norm-wise", rel_a, "worst per-ion err / bound",
            np.max(err / (1e-5 * fn + 1e-6 * rms)), "| vs exact positions: norm-wise", rel_b,
            "(fp64 on rounded positions alone:", rel_q, ")")
        assert rel_a <= 2e-6
        assert np.all(err <= 1e-5 * fn + 1e-6 * rms)
        assert rel_b <= 1e-4


@pytest.mark.parametrize("precision,mode,workload", [
    ("fp64", "rows", "single_dust"), ("fp32c", "rows", "single_dust"), ("fp32c", "sym", "single_dust"),
    ("fp64", "rows", "dust_chain"), ("fp32c", "sym", "dust_chain")])
def test_one_step_vs_oracle(oracle_lib, precision, mode, workload):
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=2048, precision=precision, mode=mode, workload=workload)
    q0 = np.array([cfg.dust_q0[p] for p in range(cfg.n_dust)])
    say("one step", precision, mode, workload, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    e.set_step(5)
    rec = e.run(1)
    x_g, v_g = e.get_ions()
    fl_g, face_g, nre = e.step_info()
    e.close()
    x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, q0, 5)
    assert np.array_equal(fl_g, fl_o)
    assert np.array_equal(face_g, face_o)
    # the crafted ions (in the dusts, beyond the wall and the caps) are flagged;
    # a few in-dust ones may cross a small dust within the step
    assert nre == int((fl_o != 0).sum()) and nre >= 12
    assert np.array_equal(rec["absorbed"][0], rec_o["absorbed"])
    dx = np.max(np.abs(x_g - x_o)) / cfg.R
    dv = np.max(np.abs(v_g - v_o)) / np.max(np.abs(v_o))
    say("  max |dx|/R", dx, "max |dv|/max|v|", dv)
    bound = None
    if precision == "fp64":
        assert dx <= 1e-11
    else:
        ok, worst, bound = fp32c_position_error_ok(oracle_lib, cfg, x0, x_g, x_o, fl_o)
        say("  worst |dx_i| / bound_i", worst)
        assert ok
    assert dv <= (1e-10 if precision == "fp64" else 1e-4)
    ratios = check_records(cfg, rec[0], rec_o, precision, x_o, bound)
    say("  records, error / tolerance:", ratios)
    for key, r in ratios.items():
        assert r <= 1.0, key


@pytest.mark.parametrize("workload", ["single_dust", "dust_chain"])
def test_multi_step_fp64_records(oracle_lib, workload):
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=1024, precision="fp64", craft=True, workload=workload)
    q0 = np.array([cfg.dust_q0[p] for p in range(cfg.n_dust)])
    say("multi step fp64", workload, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    rec = e.run(6)
    x_g, v_g = e.get_ions()
    e.close()
    x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, q0, 0, 6)
    assert np.array_equal(rec["absorbed"], rec_o["absorbed"])
    assert np.max(np.abs(rec["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12
    for key in ("force_orbit", "force_collect"):
        sc = np.max(np.abs(rec_o[key]))
        assert np.max(np.abs(rec[key] - rec_o[key])) <= 1e-9 * sc, key
    assert np.max(np.abs(x_g - x_o)) <= 1e-10 * cfg.R


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_four_dusts_mixed_charging_vs_oracle(oracle_lib, precision):
    """The most dusts a configuration holds (4), off the axis, one per charge
    method plus a second consistent one, ions crafted inside every dust: six
    steps against the oracle (FP64: the multi-step tolerances; FP32c: counts
    and given charges exact, forces and positions within the FP32c bounds of
    one step, then counts over all six steps)."""
    eng = _engine_mod()
    from opendust_amd.opendust import (DustParticle, OpenDust, OutputParameters, PlasmaParametersInSIUnitsMaxwell,
                                       SimulationParametersInSIUnits)
    from opendust_amd import workloads

    with contextlib.redirect_stdout(io.StringIO()):
        pp = PlasmaParametersInSIUnitsMaxwell(290.11, 29011, 3.57e14, 1.3, 6.63e-26)
        lam = pp.r_D_e
        sp = SimulationParametersInSIUnits(4 * lam, 8 * lam, 3000, 6, 0.01 / pp.w_p_i, pp)
        e_ = abs(pp.e)
        dusts = [DustParticle(0.3 * lam, 0.0, -1.5 * lam, 0.08 * lam, "consistent", -2e4 * e_),
                 DustParticle(-0.2 * lam, 0.4 * lam, -0.4 * lam, 0.06 * lam, "oml", -1.5e4 * e_),
                 DustParticle(0.0, -0.5 * lam, 0.7 * lam, 0.07 * lam, "given", -3e4 * e_),
                 DustParticle(0.5 * lam, 0.5 * lam, 1.8 * lam, 0.05 * lam, "consistent", -1e4 * e_)]
        od = OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), dusts, "Maxwellian", precision=precision, seed=29)
    cfg, x0, v0 = workloads.prepare(od, alpha=workloads.synthetic_alpha(od), np_seed=8, precision=precision)
    assert cfg.n_dust == 4
    rng = np.random.RandomState(3)
    for i in range(8):
        p = i % 4
        d = rng.normal(size=3)
        d *= 0.3 * cfg.dust_r[p] / np.linalg.norm(d)
        x0[i] = [cfg.dust_x[p] + d[0], cfg.dust_y[p] + d[1], cfg.dust_z[p] + d[2]]
    q0 = np.array([cfg.dust_q0[p] for p in range(4)])
    say("four dusts", precision, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    rec1 = e.run(1)
    x1, _ = e.get_ions()
    rec = e.run(5)
    e.close()
    x_o1, _, _, rec_o1, fl_o1, _ = oracle_lib.step(cfg, x0, v0, q0, 0)
    assert np.all(rec1["absorbed"][0] >= 2), rec1["absorbed"]
    assert np.array_equal(rec1["absorbed"][0], rec_o1["absorbed"])
    assert rec1["q"][0, 2] == q0[2]
    if precision == "fp64":
        x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, q0, 0, 6)
        rec_all = np.concatenate([rec1, rec])
        assert np.array_equal(rec_all["absorbed"], rec_o["absorbed"])
        assert np.max(np.abs(rec_all["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12
        for key in ("force_orbit", "force_collect", "force_dust"):
            sc = np.max(np.abs(rec_o[key]))
            assert np.max(np.abs(rec_all[key] - rec_o[key])) <= 1e-9 * sc, key
    else:
        ok, worst, bound = fp32c_position_error_ok(oracle_lib, cfg, x0, x1, x_o1, fl_o1)
        say("  worst |dx_i| / bound_i", worst)
        assert ok
        ratios = check_records(cfg, rec1[0], rec_o1, precision, x_o1, bound)
        say("  records, error / tolerance:", ratios)
        for key, r in ratios.items():
            assert r <= 1.0, key


@pytest.mark.parametrize("mode", ["rows", "sym"])
def test_determinism_and_graph(oracle_lib, mode):
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=4096, precision="fp32c", craft=True, mode=mode)
    say("determinism fp32c", mode, "N", cfg.n_ions)
    outs = []
    for flags in (0, 0, eng.RUN_GRAPH):
        e = eng.Engine(cfg)
        e.set_ions(x0, v0)
        rec = e.run(40, flags)
        x, v = e.get_ions()
        e.close()
        outs.append((rec, x, v))
    for rec, x, v in outs[1:]:
        assert np.array_equal(rec, outs[0][0])
        assert np.array_equal(x, outs[0][1])
        assert np.array_equal(v, outs[0][2])


def test_tune_is_reproducible():
    """od_tune picks the symmetric kernel's source blocks per work-group and
    kernel variant; a context created with that choice reproduces the tuned
    run bitwise."""
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=5000, precision="fp32c", craft=True, mode="sym")
    say("tune fp32c sym N", cfg.n_ions)
    outs = []
    chosen = None
    for tune in (True, False):
        if not tune:
            cfg.sym_chunk, cfg.sym_tile = chosen
        e = eng.Engine(cfg)
        e.set_ions(x0, v0)
        if tune:
            chosen = e.tune(2)
            say("  chosen sym_chunk, sym_tile", chosen)
            assert chosen[0] in (1, 2, 4, 8, 16, 32)
            assert 1 <= chosen[1] <= len(eng.sym_variants())
        assert e.sym_choice() == chosen
        rec = e.run(12, eng.RUN_GRAPH)
        outs.append((rec, e.get_ions()))
        e.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1][0], outs[1][1][0]) and np.array_equal(outs[0][1][1], outs[1][1][1])


CUT_CASES = [("fp64", 1.5, None), ("fp32c", 1.5, None), ("fp32c", 3.2, 3.0), ("fp64", 3.2, 3.0)]


@pytest.mark.parametrize("precision,rc,rs", CUT_CASES)
def test_cutoff_pair_forces_and_cells(oracle_lib, precision, rc, rs):
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=3000, precision=precision, craft=False)
    cfg.cutoff = rc * cfg.r_D_e
    cfg.switch_dist = 0.0 if rs is None else rs * cfg.r_D_e
    say("cutoff pair forces", precision, rc, rs, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    assert e.pair_info()[0] == eng.PAIR_CELL
    f_gpu = e.pair_forces(x0)
    cell, order, dims = e.cells()
    e.close()
    f_orc = oracle_lib.pair_forces(cfg, x0)
    fn = np.linalg.norm(f_orc, axis=1)
    err = np.linalg.norm(f_gpu - f_orc, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    say("  norm-wise rel", np.linalg.norm(f_gpu - f_orc) / np.linalg.norm(f_orc))
    if precision == "fp64":
        assert np.all(err <= 1e-11 * fn + 1e-13 * rms)
    else:
        assert np.linalg.norm(f_gpu - f_orc) <= 1e-5 * np.linalg.norm(f_orc)
        assert np.all(err <= 1e-4 * fn + 1e-5 * rms)
    # binning is bit-exact, the order is (cell, ion index)
    ref = np.array([oracle_lib.cell_of(cfg, xi)[0] for xi in x0])
    assert dims == oracle_lib.cell_of(cfg, x0[0])[1]
    assert np.array_equal(cell, ref)
    assert np.array_equal(order, np.lexsort((np.arange(len(ref)), ref)))


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_cutoff_one_step_and_determinism(oracle_lib, precision):
    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=2048, precision=precision, craft=True)
    cfg.cutoff = 1.5 * cfg.r_D_e
    q0 = np.array([cfg.dust_q0[p] for p in range(cfg.n_dust)])
    say("cutoff one step", precision, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    e.set_step(3)
    rec = e.run(1)
    x_g, v_g = e.get_ions()
    fl_g, face_g, nre = e.step_info()
    cell, order, dims = e.cells()
    e.close()
    x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, q0, 3)
    assert np.array_equal(fl_g, fl_o) and np.array_equal(face_g, face_o) and nre >= 18
    assert np.array_equal(rec["absorbed"][0], rec_o["absorbed"])
    # the cell list of the step holds the kept ions at their drifted positions
    kept = fl_o == 0
    assert len(order) == int(kept.sum())
    bound = None
    if precision == "fp64":
        assert np.max(np.abs(x_g - x_o)) <= 1e-11 * cfg.R
    else:
        ok, worst, bound = fp32c_position_error_ok(oracle_lib, cfg, x0, x_g, x_o, fl_o)
        say("  worst |dx_i| / bound_i", worst)
        assert ok
    ratios = check_records(cfg, rec[0], rec_o, precision, x_o, bound)
    say("  records, error / tolerance:", ratios)
    for key, r in ratios.items():
        assert r <= 1.0, key
    # determinism: eager twice and graph replay, bitwise
    outs = []
    for flags in (0, 0, eng.RUN_GRAPH):
        e = eng.Engine(cfg)
        e.set_ions(x0, v0)
        r = e.run(12, flags)
        outs.append((r, e.get_ions()[0]))
        e.close()
    for r, x in outs[1:]:
        assert np.array_equal(r, outs[0][0]) and np.array_equal(x, outs[0][1])


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
@pytest.mark.parametrize("thresholds", ["reference", "bottom_heavy"])
def test_bottom_face_reinjection_vs_oracle(oracle_lib, precision, thresholds):
    """Bottom-face reinjection (opendust.py:939-951), live whenever p_b != 0:
    Maxwellian flow at M = 0.1 (p_b = 0.030), 400 ions crafted beyond the end
    caps and the side wall.  "reference": the reference's face probabilities;
    "bottom_heavy": thresholds moved to top 0.3, side 0.3, bottom 0.4, so that
    the bottom branch runs for many ions.  Faces and flags bit-exact; bottom
    ions exactly at z = +H/2 with v_z <= 0; positions and velocities of every
    reinjected ion to FP64 rounding (1e-12 of R and of max|v|), other ions as
    the one-step tests state them."""
    from opendust_amd import host, workloads

    eng = _engine_mod()
    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=2048, n=3, precision=precision, seed=31, mach=0.1)
        cfg, x0, v0 = workloads.prepare(od, alpha=workloads.synthetic_alpha(od), np_seed=4, precision=precision)
    assert cfg.use_bottom == 1 and cfg.bottom.n_seg > 0
    if thresholds == "bottom_heavy":
        cfg.thr_top = host.u32_threshold(0.3)
        cfg.thr_side = host.u32_threshold(0.6)
    rng = np.random.RandomState(17)
    for i in range(200):
        x0[i, 2] = (0.5 + rng.uniform(0.002, 0.02)) * cfg.H * (1 if i % 2 else -1)
    for i in range(200, 400):
        a = rng.uniform(0, 2 * np.pi)
        x0[i] = [1.02 * cfg.R * np.cos(a), 1.02 * cfg.R * np.sin(a), rng.uniform(-0.45, 0.45) * cfg.H]
    q0 = np.array([cfg.dust_q0[0]])
    say("bottom face", precision, thresholds, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    e.set_step(4)
    rec = e.run(1)
    x_g, v_g = e.get_ions()
    fl_g, face_g, nre = e.step_info()
    e.close()
    x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, q0, 4)
    counts = np.bincount(face_o, minlength=4)
    say("  faces [none, top, side, bottom]", counts.tolist())
    assert np.array_equal(fl_g, fl_o)
    assert np.array_equal(face_g, face_o)
    assert counts[3] >= (5 if thresholds == "reference" else 120)
    b = face_g == 3
    assert np.all(x_g[b, 2] == cfg.H / 2) and np.all(v_g[b, 2] <= 0.0)
    re = face_o != 0
    dx_re = np.max(np.abs(x_g[re] - x_o[re])) / cfg.R
    dv_re = np.max(np.abs(v_g[re] - v_o[re])) / np.max(np.abs(v_o))
    say("  reinjected ions: max |dx|/R", dx_re, "max |dv|/max|v|", dv_re)
    assert dx_re <= 1e-12 and dv_re <= 1e-12
    assert np.array_equal(rec["absorbed"][0], rec_o["absorbed"])
    bound = None
    if precision == "fp64":
        assert np.max(np.abs(x_g - x_o)) <= 1e-11 * cfg.R
    else:
        ok, worst, bound = fp32c_position_error_ok(oracle_lib, cfg, x0, x_g, x_o, fl_o)
        say("  worst |dx_i| / bound_i", worst)
        assert ok
    ratios = check_records(cfg, rec[0], rec_o, precision, x_o, bound)
    say("  records, error / tolerance:", ratios)
    for key, r in ratios.items():
        assert r <= 1.0, key


@pytest.mark.parametrize("cut", [False, True])
def test_stage_timing_covers_the_step(cut):
    """od_stage_timing: every stage of a profiled step is timed (non-negative,
    the cell-list stage only with a cutoff), the stage times add up to the
    step's, and profiling changes no result (bitwise against an unprofiled run)."""
    import time

    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=8192, precision="fp32c", craft=True)
    if cut:
        cfg.cutoff = 1.5 * cfg.r_D_e
    say("stage timing, cutoff", cut, "N", cfg.n_ions)
    outs = []
    for flags in (eng.RUN_PROFILE_STAGES, 0):
        e = eng.Engine(cfg)
        e.set_ions(x0, v0)
        t0 = time.perf_counter()
        rec = e.run(4, flags)
        wall = (time.perf_counter() - t0) / 4 * 1e3
        if flags:
            st, n = e.stage_timing(reset=True)
            assert n == 4 and e.stage_timing()[1] == 0
        outs.append((rec, e.get_ions()[0]))
        e.close()
    say("  ms per step:", {k: round(v, 4) for k, v in st.items()}, "wall", round(wall, 3))
    assert all(v >= 0.0 for v in st.values())
    assert (st["cells"] > 0.0) == cut and (st["corr"] > 0.0) == cut
    for k in ("kick_drift", "list", "pair", "kick2", "inflow", "dust"):
        assert st[k] > 0.0, k
    assert sum(st.values()) <= 4.0 * wall + 1.0
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_trap_grid_vs_oracle(oracle_lib):
    eng = _engine_mod()
    R, H = 5.0, 10.0
    rho = np.linspace(0, R, 6)
    h = np.linspace(0, H / 2, 5)
    say("trap grid")
    g = eng.trap_potential_grid(R, H, rho, h, 200000, 99)
    o = oracle_lib.trap_grid(R, H, rho, h, 200000, 99)
    assert np.max(np.abs(g - o) / np.abs(o)) <= 1e-12


def test_simulate_api_small(tmp_path):
    from opendust_amd import workloads

    od = workloads.single_dust(N=4096, n=30, trapSamples=200000, trapSeed=5)
    od.outputParameters.nFileOutput = 10
    od.outputParameters.xyzOutputFileName = str(tmp_path / "traj.xyz")
    od.outputParameters.restartFileName = str(tmp_path / "RESTART")
    np.random.seed(0)
    say("simulate small")
    od.simulate()
    d = od.dustParticles[0]
    assert d.q.shape == (30,) and d.forceIonsOrbit.shape == (30, 3)
    assert np.all(np.isfinite(d.forceIonsOrbit)) and np.all(np.isfinite(d.q))
    assert d.q[0] == d.q[0] and d.q[1] != d.q[0]
    lines = open(tmp_path / "traj.xyz").read().splitlines()
    assert lines[0] == "4096" and len(lines) == 3 * (4096 + 2)
    rs = open(tmp_path / "RESTART").read().splitlines()
    assert len(rs) == 4096 and len(rs[0].split("\t")) == 6


@pytest.mark.parametrize("N,n_dust", [(1, 1), (7, 0), (1500, 0)])
def test_edge_sizes_and_no_dust_vs_oracle(oracle_lib, N, n_dust):
    """A single ion, a handful of ions, and a cylinder without dust: the same
    kernels and padding, checked against the oracle for three steps (FP64)."""
    import ctypes

    eng = _engine_mod()
    od, cfg, x0, v0 = _setup(N=max(N, 64), precision="fp64", craft=False)
    cfg = eng.Config.from_buffer_copy(cfg)
    cfg.n_ions = N
    x0, v0 = x0[:N].copy(), v0[:N].copy()
    if n_dust == 0:
        cfg.n_dust = 0
    say("edge case N", N, "dusts", cfg.n_dust)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    rec = e.run(3)
    x_g, v_g = e.get_ions()
    e.close()
    q0 = np.array([cfg.dust_q0[p] for p in range(cfg.n_dust)])
    x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, q0, 0, 3)
    assert rec.shape == (3, cfg.n_dust)
    assert np.max(np.abs(x_g - x_o)) <= 1e-10 * cfg.R
    assert np.max(np.abs(v_g - v_o)) <= 1e-9 * max(np.max(np.abs(v_o)), 1e-30)
    if cfg.n_dust:
        assert np.array_equal(rec["absorbed"], rec_o["absorbed"])
        assert np.max(np.abs(rec["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12
