"""Parity at the workloads' own sizes and densities (engine through the C ABI
against the FP64 CPU oracle, and FP32c against the engine's FP64 kernels).

  * Cutoff mode at the density of the 2^22-ion cutoff workload (667 ions per
    r_D_e^3): columns of r_c / 4 hold several hundred ions, so the symmetric
    cell kernel runs several target groups per column (its accumulating
    reaction branch, reaches that overlap and leave gaps), and the FP64 cell
    kernel several 64-target chunks per column.  Small enough (65536 ions,
    R = 2.5, H = 5 r_D_e) for the oracle's full rows; then the full 2^22-ion
    workload itself against FP64 on the GPU and oracle rows of a sample.
  * The 16384-ion single-dust workload at full size, several steps.
  * A batched step of the sweep generator with its largest simulation and a
    2^17-ion one.

Tolerances.  FP64 engine vs oracle: per ion |dF_i| <= 1e-11 |F_i| + 1e-13
rms|F|.  FP32c: the position rounding and the arithmetic are separated, as
the FP32c kernels see positions rounded to FP32 in units of r_D_e / log2 e:
against FP64 forces on those rounded positions, per ion |dF_i| <= 1e-5 |F_i|
+ 1e-6 rms|F| and norm-wise <= 2e-6 (compensated sums; FP32 pair terms);
against exact positions norm-wise <= 1e-4 (the rounding alone, dominated by
close pairs; <= 3e-4 in the 2^22-ion cylinder, where |u| reaches 14.4).  With a hard cutoff a pair whose distance lies within the
rounding of r of the cutoff can be counted by one side and not the other:
each such pair may add one pair force at the cutoff, |f(r_c)|, to an ion's
error.  The oracle counts those pairs per row (relative window 1e-6 for
FP32 distances, 1e-12 for FP64), and the bound of that row grows by that
count times |f(r_c)|.  Integer outputs (cells, flags, faces, counts) are
bit-exact.
"""

import contextlib
import io

import numpy as np
import pytest

from conftest import check_records, say

pytestmark = pytest.mark.gpu


def _eng():
    from opendust_amd import engine

    engine.load_library()
    return engine


def _quiet(fn, *a, **kw):
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **kw)


def _dense_setup(precision, rc, rs, N=65536, R=2.5, H=5.0, seed=9, np_seed=6):
    from opendust_amd import workloads

    od = _quiet(workloads.single_dust, N=N, n=2, precision=precision, seed=seed, R=R, H=H)
    cfg, x0, v0 = _quiet(workloads.prepare, od, alpha=workloads.synthetic_alpha(od), np_seed=np_seed,
                         precision=precision)
    cfg.cutoff = rc * cfg.r_D_e
    cfg.switch_dist = 0.0 if rs is None else rs * cfg.r_D_e
    return od, cfg, x0, v0


def _check_rows(f_gpu, f_ref, edge, f_cut, rel_i, rel_rms, what):
    """Per-row bound rel_i |F_i| + rel_rms rms|F| + edge_i |f(r_c)|; returns
    the worst error/bound ratio."""
    fn = np.linalg.norm(f_ref, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    err = np.linalg.norm(f_gpu - f_ref, axis=1)
    bound = rel_i * fn + rel_rms * rms + edge * f_cut * 1.001
    worst = float(np.max(err / bound))
    say("  %s: norm-wise rel %.3g, worst per-ion error/bound %.3g, rows with pairs at the cutoff %d"
        % (what, np.linalg.norm(f_gpu - f_ref) / np.linalg.norm(f_ref), worst, int(np.sum(edge > 0))))
    assert np.all(err <= bound), what
    return worst


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,rc,rs", [("fp64", 2.0, None), ("fp32c", 2.0, None), ("fp64", 3.2, 3.0),
                                             ("fp32c", 3.2, 3.0)])
def test_cutoff_dense_cells_vs_oracle(oracle_lib, precision, rc, rs):
    eng = _eng()
    od, cfg, x0, v0 = _dense_setup(precision, rc, rs)
    say("dense cutoff forces", precision, "rc", rc, "rs", rs, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    assert e.pair_info()[0] == eng.PAIR_CELL
    f_gpu = e.pair_forces(x0)
    cell, order, dims = e.cells()
    in_range, cand = e.cutoff_pair_count()
    e.close()
    n_cells = dims[0] * dims[1] * dims[2]
    col = np.bincount(cell[cell < n_cells] // dims[2], minlength=dims[0] * dims[1])
    say("  cells", dims, "max ions per column", col.max(), "in-range / candidate pair terms", in_range, cand)
    # several target groups (up to 256 ions each) and 64-ion chunks per column
    assert col.max() > 512
    ref = np.array([oracle_lib.cell_of(cfg, xi)[0] for xi in x0])
    assert np.array_equal(cell, ref)
    assert np.array_equal(order, np.lexsort((np.arange(len(ref)), ref)))
    f_cut = oracle_lib.cut_force_at_cutoff(cfg)
    if precision == "fp64":
        f_orc = oracle_lib.pair_forces(cfg, x0)
        edge = oracle_lib.cut_edge_count(cfg, x0, 1e-12)
        _check_rows(f_gpu, f_orc, edge, f_cut, 1e-11, 1e-13, "fp64 vs oracle")
    else:
        from oracle.oracle import fp32c_rounded_positions

        xr = fp32c_rounded_positions(cfg, x0)
        f_r = oracle_lib.pair_forces(cfg, xr)
        edge = oracle_lib.cut_edge_count(cfg, xr, 1e-6)
        _check_rows(f_gpu, f_r, edge, f_cut, 1e-5, 1e-6, "fp32c vs oracle on rounded positions")
        flip = np.sqrt(np.sum((edge * f_cut) ** 2))
        assert np.linalg.norm(f_gpu - f_r) <= 2e-6 * np.linalg.norm(f_r) + flip
        f_orc = oracle_lib.pair_forces(cfg, x0)
        rel_q = np.linalg.norm(f_gpu - f_orc) / np.linalg.norm(f_orc)
        say("  fp32c vs oracle on exact positions: norm-wise rel", rel_q)
        assert rel_q <= 1e-4


@pytest.mark.timeout(600)
@pytest.mark.parametrize("rc,rs", [(2.0, None), (3.2, 3.0)])
def test_cutoff_cell_variants_vs_oracle(oracle_lib, rc, rs):
    """Every variant of the symmetric cell kernel the cutoff autotuner may
    pick (1, 2 or 4 target sub-blocks per wave, sources rotated through
    registers or read from LDS), hard and switched law, at
    the dense workload's density: FP32c bound against the oracle on the
    rounded positions, each variant bitwise reproducible."""
    eng = _eng()
    from oracle.oracle import fp32c_rounded_positions

    od, cfg, x0, v0 = _dense_setup("fp32c", rc, rs)
    xr = fp32c_rounded_positions(cfg, x0)
    f_r = oracle_lib.pair_forces(cfg, xr)
    edge = oracle_lib.cut_edge_count(cfg, xr, 1e-6)
    f_cut = oracle_lib.cut_force_at_cutoff(cfg)
    flip = np.sqrt(np.sum((edge * f_cut) ** 2))
    tiles = []
    for tile in range(1, 7):
        cfg.sym_tile = tile
        say("dense cutoff, cell kernel variant", tile, "rc", rc, "rs", rs)
        e = eng.Engine(cfg)
        f1 = e.pair_forces(x0)
        f2 = e.pair_forces(x0)
        tiles.append(e.sym_choice()[1])
        e.close()
        assert np.array_equal(f1, f2)
        _check_rows(f1, f_r, edge, f_cut, 1e-5, 1e-6, "variant %d vs oracle on rounded positions" % tile)
        assert np.linalg.norm(f1 - f_r) <= 2e-6 * np.linalg.norm(f_r) + flip
    assert tiles == [1, 2, 3, 4, 5, 6]
    bad = eng.Config.from_buffer_copy(cfg)
    bad.sym_tile = 7
    with pytest.raises(RuntimeError, match="sym_tile"):
        eng.Engine(bad)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_cutoff_many_cells_vs_oracle(oracle_lib, precision):
    """A grid of 40 x 40 x 80 = 128000 cells (R = 10, H = 40 r_D_e, cutoff 1
    r_D_e), more than 16-bit cell ids hold (three radix passes): binning and
    order bit-exact, forces as in the dense case."""
    eng = _eng()
    od, cfg, x0, v0 = _dense_setup(precision, 1.0, None, N=20000, R=10.0, H=40.0, seed=2, np_seed=3)
    say("many cells", precision, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    f_gpu = e.pair_forces(x0)
    cell, order, dims = e.cells()
    e.close()
    assert dims[0] * dims[1] * dims[2] > 65535
    ref = np.array([oracle_lib.cell_of(cfg, xi)[0] for xi in x0])
    assert ref.max() > 65535
    assert np.array_equal(cell, ref)
    assert np.array_equal(order, np.lexsort((np.arange(len(ref)), ref)))
    f_cut = oracle_lib.cut_force_at_cutoff(cfg)
    if precision == "fp64":
        _check_rows(f_gpu, oracle_lib.pair_forces(cfg, x0), oracle_lib.cut_edge_count(cfg, x0, 1e-12), f_cut,
                    1e-11, 1e-13, "fp64 vs oracle")
    else:
        from oracle.oracle import fp32c_rounded_positions

        xr = fp32c_rounded_positions(cfg, x0)
        _check_rows(f_gpu, oracle_lib.pair_forces(cfg, xr), oracle_lib.cut_edge_count(cfg, xr, 1e-6), f_cut,
                    1e-5, 1e-6, "fp32c vs oracle on rounded positions")


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_cutoff_dense_one_step_vs_oracle(oracle_lib, precision):
    """One full step at the cutoff workload's density with ions crafted into
    the dust and beyond the walls: flags, faces, counts bit-exact; positions
    to 1e-11 R (FP64) or within the measured FP32c bound (position rounding
    measured, arithmetic bounded) for every kept ion; records as the
    single-step tests state them."""
    from oracle.oracle import fp32c_position_bounds_measured

    eng = _eng()
    od, cfg, x0, v0 = _dense_setup(precision, 2.0, None, N=32768, R=2.0, H=4.0, seed=4, np_seed=2)
    rng = np.random.RandomState(5)
    for i in range(6):
        d = rng.normal(size=3)
        d *= 0.5 * cfg.dust_r[0] / np.linalg.norm(d)
        x0[i] = [cfg.dust_x[0] + d[0], cfg.dust_y[0] + d[1], cfg.dust_z[0] + d[2]]
    for i in range(6, 40):
        a = rng.uniform(0, 2 * np.pi)
        x0[i] = [1.01 * cfg.R * np.cos(a), 1.01 * cfg.R * np.sin(a), rng.uniform(-0.45, 0.45) * cfg.H]
    for i in range(40, 70):
        x0[i, 2] = 0.505 * cfg.H * (1 if i % 2 else -1)
    q0 = np.array([cfg.dust_q0[0]])
    say("dense cutoff one step", precision, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    e.set_step(2)
    rec = e.run(1)
    x_g, v_g = e.get_ions()
    fl_g, face_g, nre = e.step_info()
    e.close()
    x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, q0, 2)
    assert np.array_equal(fl_g, fl_o) and np.array_equal(face_g, face_o) and nre >= 60
    assert np.array_equal(rec["absorbed"][0], rec_o["absorbed"])
    bound = None
    if precision == "fp64":
        dx = np.max(np.abs(x_g - x_o)) / cfg.R
        say("  max |dx|/R", dx)
        assert dx <= 1e-11
    else:
        bound = fp32c_position_bounds_measured(oracle_lib, cfg, x0, fl_o)
        ratio = np.linalg.norm(x_g - x_o, axis=1) / bound
        say("  worst |dx_i| / bound_i", ratio.max())
        assert np.all(ratio <= 1.0)
    ratios = check_records(cfg, rec[0], rec_o, precision, x_o, bound)
    say("  records, error / tolerance:", ratios)
    for key, r in ratios.items():
        assert r <= 1.0, key


@pytest.mark.timeout(900)
def test_cutoff_full_size_fp32c_vs_fp64_and_oracle_rows(oracle_lib):
    """The 2^22-ion cutoff workload (R = 10, H = 20 r_D_e, hard cutoff 2
    r_D_e) at full size.  (a) FP64 cell kernel against the oracle's rows of
    2000 ions (exact positions).  (b) FP32c symmetric cell kernel against the
    FP64 cell kernel on the FP32-rounded positions: the 2000 sampled rows to
    the per-row bound with the oracle's counts of pairs at the cutoff; every
    ion either within the flip-free bound or, for at most 3 % of them
    (distance within the FP32 rounding of the cutoff), within it plus three
    pair forces at the cutoff.  (c) Against exact positions: the difference is
    the FP64 kernel's own difference between rounded and exact positions (to
    5e-6), which is <= 3e-4 norm-wise (measured 1.67e-4).
    (At this density 6 % of the ions have a pair within 1e-6 of the cutoff,
    about a quarter of that within the FP32 rounding of r; a flipped pair
    adds |f(r_c)| = 4.4e-5 rms|F| to its ion's error.)"""
    from opendust_amd import workloads
    from oracle.oracle import fp32c_rounded_positions

    eng = _eng()
    od = _quiet(workloads.cutoff_large, N=2 ** 22, n=2, precision="fp64", seed=3)
    alpha = workloads.synthetic_alpha(od)
    cfg, x0, v0 = _quiet(workloads.prepare, od, alpha=alpha, np_seed=5, precision="fp64")
    xr = fp32c_rounded_positions(cfg, x0)
    say("full-size cutoff: fp64 cell kernel N", cfg.n_ions)
    e = eng.Engine(cfg)
    f64 = e.pair_forces(x0)
    f64r = e.pair_forces(xr)
    cell, _, dims = e.cells()
    e.close()
    n_cells = dims[0] * dims[1] * dims[2]
    col = np.bincount(cell[cell < n_cells] // dims[2], minlength=dims[0] * dims[1])
    say("  cells", dims, "mean ions per occupied column %.0f, max %d" % (col[col > 0].mean(), col.max()))
    assert col.max() > 512
    cfg32, x32_0, _ = _quiet(workloads.prepare, _quiet(workloads.cutoff_large, N=2 ** 22, n=2, precision="fp32c",
                                                       seed=3), alpha=alpha, np_seed=5, precision="fp32c")
    assert np.array_equal(x32_0, x0)
    say("full-size cutoff: fp32c symmetric cell kernel")
    e = eng.Engine(cfg32)
    f32 = e.pair_forces(x0)
    e.close()
    f_cut = oracle_lib.cut_force_at_cutoff(cfg)
    i0 = cfg.n_ions // 2
    rows = (i0, i0 + 2000)
    say("  oracle rows", rows)
    fo = oracle_lib.pair_forces(cfg, x0, rows=rows)
    edge = oracle_lib.cut_edge_count(cfg, x0, 1e-12, rows=rows)
    _check_rows(f64[i0:i0 + 2000], fo, edge, f_cut, 1e-11, 1e-13, "fp64 cell kernel vs oracle rows")
    fo_r = oracle_lib.pair_forces(cfg, xr, rows=rows)
    edge_r = oracle_lib.cut_edge_count(cfg, xr, 1e-6, rows=rows)
    _check_rows(f32[i0:i0 + 2000], fo_r, edge_r, f_cut, 1e-5, 1e-6, "fp32c vs oracle rows on rounded positions")
    fn = np.linalg.norm(f64r, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    err = np.linalg.norm(f32 - f64r, axis=1)
    b0 = 1e-5 * fn + 1e-6 * rms
    over = err > b0
    say("  fp32c vs fp64 (rounded positions), all ions: norm-wise %.3g, beyond the flip-free bound %d (%.3g %%), "
        "worst excess in pair forces at the cutoff %.3g"
        % (np.linalg.norm(f32 - f64r) / np.linalg.norm(f64r), int(over.sum()), 100 * over.mean(),
           float(np.max((err - b0) / f_cut))))
    assert over.mean() <= 0.03
    # at most three flipped pairs on one ion (two seen at this seed)
    assert np.all(err <= b0 + 3.0 * f_cut)
    rel_q = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    rel_round = np.linalg.norm(f64r - f64) / np.linalg.norm(f64)
    say("  fp32c vs fp64 on exact positions: norm-wise", rel_q, "(fp64 on rounded positions alone:", rel_round, ")")
    # all of it is the rounding of the positions: |u| <= 14.4 here (ulp 9.5e-7,
    # twice that of the 2^20 all-pairs workload) and the truncated sum leaves
    # the close pairs a larger share of each force
    assert rel_q <= rel_round + 5e-6
    assert rel_round <= 3e-4


@pytest.mark.timeout(900)
def test_all_pairs_2e22_symmetric_in_bands_vs_fp64():
    """All-pairs at 2^22 ions (the cutoff workload's cylinder, no cutoff): the
    symmetric FP32c kernel's schedule holds 51.5 GB of partials at its
    default (row groups of 4 target blocks, 8 source blocks per work-group),
    so it runs in two bands of at most 48 GB, each launched and reduced in
    order.  Against the FP64 row kernel: on
    the FP32-rounded positions norm-wise <= 2e-6 and per ion <= 2e-5 |F_i| +
    2e-6 rms|F|, the 2^20 bound scaled by sqrt(N / 2^20) for the random walk
    of the per-term roundings over N sources (every block pair of every band
    lands); on the exact positions the FP64 kernel's own rounded-minus-exact
    difference to 5e-6."""
    from opendust_amd import workloads
    from oracle.oracle import fp32c_rounded_positions

    eng = _eng()
    od = _quiet(workloads.cutoff_large, N=2 ** 22, n=2, precision="fp64", seed=8, cutoff=0.0)
    od.set_cutoff(False)
    alpha = workloads.synthetic_alpha(od)
    cfg, x0, v0 = _quiet(workloads.prepare, od, alpha=alpha, np_seed=9, precision="fp64")
    assert cfg.cutoff == 0.0
    xr = fp32c_rounded_positions(cfg, x0)
    say("2^22 all-pairs: fp64 rows N", cfg.n_ions)
    e = eng.Engine(cfg)
    f64 = e.pair_forces(x0)
    f64r = e.pair_forces(xr)
    e.close()
    cfg32 = eng.Config.from_buffer_copy(cfg)
    cfg32.precision = 0
    cfg32.pair_mode = 2
    say("2^22 all-pairs: fp32c symmetric, in bands")
    e = eng.Engine(cfg32)
    sched = e.sym_schedule()
    say("  schedule (source blocks per work-group, variant, bands)", sched)
    assert e.pair_info()[0] == 2 and sched[:2] == (8, 12) and sched[2] >= 2
    f32 = e.pair_forces(x0)
    e.close()
    fn = np.linalg.norm(f64r, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    err = np.linalg.norm(f32 - f64r, axis=1)
    rel_a = np.linalg.norm(f32 - f64r) / np.linalg.norm(f64r)
    rel_q = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    rel_round = np.linalg.norm(f64r - f64) / np.linalg.norm(f64)
    say("  arithmetic: norm-wise", rel_a, "worst per-ion err / bound", np.max(err / (2e-5 * fn + 2e-6 * rms)),
        "| vs exact positions", rel_q, "(rounding alone", rel_round, ")")
    assert rel_a <= 2e-6
    assert np.all(err <= 2e-5 * fn + 2e-6 * rms)
    assert rel_q <= rel_round + 5e-6


@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,mode", [("fp64", "rows"), ("fp32c", "rows"), ("fp32c", "sym")])
def test_small_workload_full_size_vs_oracle(oracle_lib, precision, mode):
    """The 16384-ion single-dust workload (R = 5, H = 10 r_D_e, M = 1) at its
    own size, ions crafted into the dust and beyond the walls: FP64 three
    steps (positions 1e-10 R, charges 1e-12, dust forces 1e-9, counts exact);
    FP32c one step (measured FP32c position bound, records)."""
    from opendust_amd import workloads
    from oracle.oracle import fp32c_position_bounds_measured

    eng = _eng()
    od = _quiet(workloads.single_dust, N=16384, n=3, precision=precision, seed=12)
    cfg, x0, v0 = _quiet(workloads.prepare, od, alpha=workloads.synthetic_alpha(od), np_seed=7, precision=precision)
    cfg.pair_mode = {"rows": 1, "sym": 2}[mode]
    rng = np.random.RandomState(2)
    for i in range(8):
        d = rng.normal(size=3)
        d *= 0.5 * cfg.dust_r[0] / np.linalg.norm(d)
        x0[i] = [cfg.dust_x[0] + d[0], cfg.dust_y[0] + d[1], cfg.dust_z[0] + d[2]]
    for i in range(8, 30):
        a = rng.uniform(0, 2 * np.pi)
        x0[i] = [1.02 * cfg.R * np.cos(a), 1.02 * cfg.R * np.sin(a), rng.uniform(-0.4, 0.4) * cfg.H]
    q0 = np.array([cfg.dust_q0[0]])
    say("small workload full size", precision, mode, "N", cfg.n_ions)
    e = eng.Engine(cfg)
    e.set_ions(x0, v0)
    if precision == "fp64":
        rec = e.run(3)
        x_g, v_g = e.get_ions()
        e.close()
        x_o, v_o, q_o, rec_o = oracle_lib.run(cfg, x0, v0, q0, 0, 3)
        assert np.array_equal(rec["absorbed"], rec_o["absorbed"])
        assert rec_o["absorbed"].sum() >= 8
        assert np.max(np.abs(rec["q"] - rec_o["q"]) / np.abs(rec_o["q"])) <= 1e-12
        for key in ("force_orbit", "force_collect", "force_dust"):
            sc = max(np.max(np.abs(rec_o[key])), 1e-300)
            assert np.max(np.abs(rec[key] - rec_o[key])) <= 1e-9 * sc, key
        dx = np.max(np.abs(x_g - x_o)) / cfg.R
        say("  max |dx|/R after 3 steps", dx)
        assert dx <= 1e-10
        return
    rec = e.run(1)
    x_g, v_g = e.get_ions()
    fl_g, face_g, nre = e.step_info()
    e.close()
    x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, q0, 0)
    assert np.array_equal(fl_g, fl_o) and np.array_equal(face_g, face_o) and nre >= 28
    assert np.array_equal(rec["absorbed"][0], rec_o["absorbed"])
    bound = fp32c_position_bounds_measured(oracle_lib, cfg, x0, fl_o)
    ratio = np.linalg.norm(x_g - x_o, axis=1) / bound
    say("  worst |dx_i| / bound_i", ratio.max())
    assert np.all(ratio <= 1.0)
    assert np.max(np.abs(v_g - v_o)) <= 1e-4 * np.max(np.abs(v_o))
    ratios = check_records(cfg, rec[0], rec_o, precision, x_o, bound)
    say("  records, error / tolerance:", ratios)
    for key, r in ratios.items():
        assert r <= 1.0, key


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_batch_sweep_generator_at_scale_vs_oracle(oracle_lib, precision):
    """One batched step of the sweep generator's own simulations (seed 7): its
    largest (108158 ions, 2^16.7), its smallest (4199) and three more, plus a
    2^17-ion simulation from the same generator, each against the oracle run
    alone on its configuration.  Flags, faces and counts bit-exact; FP64
    positions 1e-11 R; FP32c within the measured bound; records as in the
    single-step tests."""
    from opendust_amd import sweep, workloads
    from oracle.oracle import fp32c_position_bounds_measured

    eng = _eng()
    gen = _quiet(workloads.sweep, n_sims=64, n=2, precision=precision, seed=7)
    ns = np.array([od.N for od in gen])
    pick = [int(ns.argmax()), int(ns.argmin()), 0, 1, 2]
    ods = [gen[i] for i in pick] + _quiet(workloads.sweep, n_sims=1, n=2, precision=precision, seed=7,
                                          log2_n=(17.0, 17.0))
    assert ods[-1].N == 2 ** 17 and ods[0].N == ns.max()
    alphas = [workloads.synthetic_alpha(od) for od in ods]
    np.random.seed(21)
    cfgs, states = _quiet(sweep.prepare_batch, ods, device=0, alphas=alphas)
    rng = np.random.RandomState(8)
    for cfg, (x0, v0) in zip(cfgs, states):
        for i in range(8):
            p = i % cfg.n_dust
            d = rng.normal(size=3)
            d *= 0.5 * cfg.dust_r[p] / np.linalg.norm(d)
            x0[i] = [cfg.dust_x[p] + d[0], cfg.dust_y[p] + d[1], cfg.dust_z[p] + d[2]]
        for i in range(8, 16):
            a = rng.uniform(0, 2 * np.pi)
            x0[i] = [1.02 * cfg.R * np.cos(a), 1.02 * cfg.R * np.sin(a), rng.uniform(-0.4, 0.4) * cfg.H]
    say("batch at scale", precision, [c.n_ions for c in cfgs])
    b = eng.Batch(cfgs)
    for s, (x0, v0) in enumerate(states):
        b.set_ions(s, x0, v0)
    rec = b.run(1)
    got = [(b.get_ions(s), b.step_info(s)) for s in range(len(cfgs))]
    b.close()
    for s, (cfg, (x0, v0)) in enumerate(zip(cfgs, states)):
        (x_g, v_g), (fl_g, face_g) = got[s]
        nd = cfg.n_dust
        q0 = np.array([cfg.dust_q0[p] for p in range(nd)])
        x_o, v_o, q_o, rec_o, fl_o, face_o = oracle_lib.step(cfg, x0, v0, q0, 0)
        assert np.array_equal(fl_g, fl_o), s
        assert np.array_equal(face_g, face_o), s
        assert int((fl_o != 0).sum()) >= 12
        assert np.array_equal(rec["absorbed"][0, s, :nd], rec_o["absorbed"]), s
        bound = None
        if precision == "fp64":
            dx = np.max(np.abs(x_g - x_o)) / cfg.R
            say("  sim", s, "N", cfg.n_ions, "max |dx|/R", dx)
            assert dx <= 1e-11, s
        else:
            bound = fp32c_position_bounds_measured(oracle_lib, cfg, x0, fl_o)
            ratio = np.linalg.norm(x_g - x_o, axis=1) / bound
            say("  sim", s, "N", cfg.n_ions, "worst |dx_i| / bound_i", ratio.max())
            assert np.all(ratio <= 1.0), s
        ratios = check_records(cfg, rec[0, s, :nd], rec_o, precision, x_o, bound)
        say("  records, error / tolerance:", ratios)
        for key, r in ratios.items():
            assert r <= 1.0, (s, key)
