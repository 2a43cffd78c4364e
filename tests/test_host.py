"""Host-side logic of the drop-in API (no GPU): parameter classes against the
reference's own printed logs, reinjection tables against closed forms, the
initial-state draw order against a literal transcription of the reference
loop, the trap fit, restart I/O and API quirks."""

import contextlib
import io
import json
import os

import numpy as np
import pytest
from scipy import special

from opendust_amd import host
from opendust_amd import opendust as od

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_logs.json")


def _gold():
    with open(GOLD) as fh:
        return json.load(fh)


def _capture(fn):
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        fn()
    return buf.getvalue().splitlines()


def test_maxwell_parameters_match_reference_log():
    g = _gold()["piel2017_fig1_consistent"]
    inp = g["inputs"]
    pp = od.PlasmaParametersInSIUnitsMaxwell(inp["T_i"], inp["T_e"], inp["n_inf"], inp["M"], inp["m_i"])
    sp = od.SimulationParametersInSIUnits(inp["R_over_lambda"] * pp.r_D_e, inp["H_over_lambda"] * pp.r_D_e,
                                          inp["N"], inp["n"], inp["d_t"], pp)
    lines = _capture(pp.printParameters) + _capture(sp.printParameters)
    gold = g["printout"]
    assert lines[:len(gold)] == gold and all(ln == "" for ln in lines[len(gold):])


def test_field_driven_parameters_match_reference_log():
    g = _gold()["patacchini2008_fig3_point1"]
    inp = g["inputs"]
    E = np.asarray([inp["E"]])[0]
    w_c = np.asarray([inp["w_c"]])[0]
    pp = od.PlasmaParametersInSIUnitsFieldDriven(inp["T_n"], inp["T_e"], inp["n_inf"], E, w_c, inp["m_i"])
    sp = od.SimulationParametersInSIUnits(inp["R_over_lambda"] * pp.r_D_e, inp["H_over_lambda"] * pp.r_D_e,
                                          inp["N"], inp["n"], inp["d_t"], pp)
    lines = _capture(pp.printParameters) + _capture(sp.printParameters)
    gold = g["printout"]
    assert lines[:len(gold)] == gold and all(ln == "" for ln in lines[len(gold):])


def test_alias_and_reexports():
    import opendust_amd

    assert opendust_amd.SimulatioParametersInSIUnits is opendust_amd.SimulationParametersInSIUnits
    opendust_amd.install_alias()
    from opendust.opendust import OpenDust  # noqa: F401  (resolves to this package)
    from opendust import DustParticle

    assert DustParticle is od.DustParticle


def _single(N=256, n=7, dist="Maxwellian", charge="consistent"):
    pp = od.PlasmaParametersInSIUnitsMaxwell(290.11, 29011, 1e14, 1, 6.6335209e-26)
    lam = pp.r_D_e
    sp = od.SimulationParametersInSIUnits(3 * lam, 6 * lam, N, n, 3.5148240854e-09, pp)
    d = od.DustParticle(0, 0, -lam, 58.8e-6, charge, -1e5 * abs(pp.e))
    return od.OpenDust(pp, sp, od.OutputParameters(), [d], dist)


def test_opendust_attributes_follow_reference():
    o = _single()
    # t = linspace(0, n*dt, n): the reference's spacing is n/(n-1)*dt (opendust.py:315)
    assert np.array_equal(o.t, np.linspace(0, 7 * 3.5148240854e-09, 7))
    d = o.dustParticles[0]
    assert d.q.shape == (7,) and np.all(d.q == d.q[0])
    assert d.forceIonsOrbit.shape == (7, 3) and not hasattr(d, "forceExternalField")
    assert abs(d.z / o.r_D_e + 1.0) < 1e-15
    assert host.charge_method_code("consistent") == host.CHARGE_CONSISTENT
    assert host.charge_method_code("oml") == host.CHARGE_OML
    assert host.charge_method_code("anything") == host.CHARGE_GIVEN


def test_unknown_distribution_raises_name_error():
    with pytest.raises(NameError):
        host.initial_velocities("uniform", 1.0, 0.1, 4)


def _flux_cdf_closed(v_fl, v_T, a, b):
    """int_a^b v N(v; v_fl, v_T) dv in closed form."""
    s2 = np.sqrt(2.0) * v_T
    return (v_T / np.sqrt(2 * np.pi) * (np.exp(-((a - v_fl) / s2) ** 2) - np.exp(-((b - v_fl) / s2) ** 2))
            + v_fl / 2 * (special.erf((b - v_fl) / s2) - special.erf((a - v_fl) / s2)))


@pytest.mark.parametrize("M", [0.3, 1.0, 2.0])
def test_maxwell_face_tables_against_closed_form(M):
    v_T = 0.2457
    v_fl = M * 10 * v_T
    F_top, top = host.face_table(v_fl, v_T, "top", "Maxwellian")
    v_min, v_max = max(v_fl - 4 * v_T, 0), v_fl + 4 * v_T
    assert abs(F_top - _flux_cdf_closed(v_fl, v_T, v_min, v_max)) <= 1e-10 * F_top
    # knots: the CDF at every grid speed equals the closed form
    for v, xi in zip(top.grid_v, top.grid_xi):
        assert abs(xi - _flux_cdf_closed(v_fl, v_T, v_min, v) / F_top) <= 1e-9
    # the expression-form evaluation equals linear interpolation of the knots
    # up to the %.8f rounding of its constants (5e-9 each; the rounding of
    # xi0 and the centre is amplified by the segment's slope)
    u = np.linspace(0, 1, 2001)[:-1] + 0.00025
    ref = np.interp(u, top.grid_xi, top.grid_v)
    k = np.clip(np.searchsorted(top.grid_xi, u) - 1, 0, top.n_seg - 1)
    bound = 1.5e-8 * (1.0 + np.abs(np.asarray(top.slope)[k])) + 1e-12
    assert np.all(np.abs(top(u) - ref) <= bound)
    F_side, side = host.face_table(0, v_T, "side", "Maxwellian")
    assert abs(F_side - _flux_cdf_closed(0, v_T, 0, 4 * v_T)) <= 1e-10 * F_side
    F_bot, bot = host.face_table(-v_fl, v_T, "bottom", "Maxwellian")
    if -v_fl + 4 * v_T < 0.4 * v_T:
        assert F_bot == 0 and bot is None
    else:
        assert np.all(bot.grid_v <= 0)


def test_face_probabilities_single_dust_workload():
    from opendust_amd import workloads

    o = workloads.single_dust(N=1024, n=2)
    o.reinjection_tables()
    assert abs(o.p_t - 0.8624) < 1e-4 and abs(o.p_s - 0.1376) < 1e-4 and o.p_b == 0.0
    top = host.face_table(o.v_fl, o.v_T, "top", "Maxwellian")[1]
    assert top.n_seg == 38  # int(8 v_T / (0.2 v_T)) = 39 knots after float truncation


def test_field_driven_tables_monotone_and_normalised():
    v_T, v_fl = 0.2457, 8 * 0.2457
    for side, vf in (("top", v_fl), ("bottom", -v_fl), ("side", 0)):
        F, t = host.face_table(vf, v_T, side, "fieldDriven")
        assert F > 0
        assert t.grid_xi[0] == 0.0 and abs(t.grid_xi[-1] - 1.0) < 1e-12
        assert np.all(np.diff(t.grid_xi) > 0)
    vz = host.vz_table(v_fl, v_T)
    assert np.all(np.diff(vz.grid_xi) > 0) and abs(vz.grid_xi[-1] - 1) < 1e-12


def test_u32_threshold():
    assert host.u32_threshold(0.0) == 0
    assert host.u32_threshold(1.0) == 2 ** 32
    assert host.u32_threshold(0.5) == 2 ** 31
    assert host.u32_threshold(1e-3) == int(np.floor(1e-3 * 2 ** 32))


def _reference_loop_positions(R, H, dusts, N):
    """Literal transcription of the reference's per-ion rejection loop
    (opendust.py:376-399) for comparison with the vectorised version."""
    from numpy import random

    position = []
    for _ in range(N):
        flag = True
        while flag:
            r = np.sqrt(random.uniform(0, 1))
            theta = random.uniform(0, 1)
            x = R * r * np.cos(2 * np.pi * theta)
            y = R * r * np.sin(2 * np.pi * theta)
            z = random.uniform(-H / 2, H / 2)
            inside = False
            for (dx, dy, dz, dr) in dusts:
                if np.sqrt((x - dx) ** 2 + (y - dy) ** 2 + (z - dz) ** 2) <= dr:
                    inside = True
                    break
            if not inside:
                position.append((x, y, z))
                flag = False
    return np.array(position)


def test_initial_positions_follow_reference_stream():
    R, H = 3.0, 6.0
    dusts = [(0.0, 0.0, -1.0, 1.2), (0.5, 0.0, 1.0, 0.8)]  # big dusts: many rejections
    np.random.seed(123)
    ref = _reference_loop_positions(R, H, dusts, 400)
    nxt_ref = np.random.random_sample()
    np.random.seed(123)
    got = host.initial_positions(R, H, dusts, 400)
    nxt_got = np.random.random_sample()
    assert np.max(np.abs(got - ref)) <= 4 * np.finfo(float).eps * R
    assert nxt_got == nxt_ref  # same stream position afterwards


def test_initial_velocities_same_calls():
    np.random.seed(9)
    v = host.initial_velocities("Maxwellian", 2.0, 0.25, 100)
    np.random.seed(9)
    ref = np.stack([np.random.normal(0, 0.25, 100), np.random.normal(0, 0.25, 100),
                    np.random.normal(2.0, 0.25, 100)], axis=1)
    assert np.array_equal(v, ref)


def test_trap_fit_recovers_polynomial():
    Rl, Hl = 5.0, 10.0
    n_r, n_h, n_mc = host.trap_grid_shape(Rl, Hl)
    assert (n_r, n_h) == (50, 50) and n_mc == int(np.pi * 25 * 10 / 0.05 ** 3)
    rng = np.random.RandomState(0)
    c = rng.normal(size=15) * 10.0 ** -np.arange(15)
    rho = np.linspace(0, Rl, n_r)
    hh = np.linspace(0, Hl / 2, n_h)
    R2, H2 = np.meshgrid(rho, hh, indexing="ij")
    pot = sum(c[m] * R2 ** a * H2 ** b for m, (a, b) in enumerate(host.TRAP_MONOMIALS))
    alpha = host.trap_fit(pot, Rl, Hl, 1000)
    assert np.allclose(alpha, -c * 1000, rtol=1e-6, atol=1e-9 * np.abs(c).max() * 1000)


def test_restart_roundtrip_and_errors(tmp_path):
    o = _single(N=5)
    pos = np.arange(15, dtype=float).reshape(5, 3) * 1.5
    vel = -np.arange(15, dtype=float).reshape(5, 3) * 0.25
    o.outputParameters.restartFileName = str(tmp_path / "RESTART")
    o._write_restart(pos, vel)
    x, v = host.read_restart(str(tmp_path / "RESTART"), 5)
    assert np.array_equal(x, pos) and np.array_equal(v, vel)
    bad = tmp_path / "BAD"
    bad.write_text("1\t2\t3\n")
    with pytest.raises(ValueError):
        host.read_restart(str(bad), 1)


def test_config_layout_and_constants():
    from opendust_amd import workloads

    o = workloads.single_dust(N=512, n=3)
    alpha = np.zeros(15)
    cfg, x0, v0 = workloads.prepare(o, alpha=alpha, np_seed=1)
    assert cfg.n_ions == 512 and cfg.n_dust == 1
    # v_B dt = 0.01 r_D_e at dt = 0.01/w_pi, M = 1 (v_B / w_pi = r_D_e)
    assert abs(cfg.v_fl * cfg.dt / cfg.r_D_e - 0.01) < 1e-9
    assert abs(cfg.v_fl / cfg.v_T - 10.0) < 1e-9
    assert cfg.top.n_seg == 38 and cfg.side.n_seg > 0 and cfg.bottom.n_seg == 0
    assert cfg.use_bottom == 0
    # positions inside the cylinder, outside the dust
    rho = np.hypot(x0[:, 0], x0[:, 1])
    assert np.all(rho <= cfg.R) and np.all(np.abs(x0[:, 2]) <= cfg.H / 2)
    assert np.all(np.linalg.norm(x0, axis=1) > cfg.dust_r[0])


def test_dust_langevin_step_drag_gravity_and_noise():
    """Without noise the update is the exact constant-acceleration step with
    Epstein drag; the thermal kick has the stated variance."""
    from opendust_amd.dynamics import dust_langevin_step

    r_p, rho, P, v_T, dt = 4.445e-6, 1500.0, 18.13, 395.0, 5e-5
    m_p = 4.0 / 3.0 * np.pi * r_p ** 3 * rho * 1.05
    gamma = 8.0 / 3.0 * np.sqrt(2.0 * np.pi) * r_p ** 2 * P / v_T / m_p
    x = np.array([1e-4, -2e-4, 3e-4])
    v = np.array([1e-3, 0.0, -2e-3])
    f = np.array([1e-13, 0.0, -5e-13])
    xn, vn = dust_langevin_step(x, v, f, m_p, r_p, P, v_T, 0.0, dt)
    a = (f - m_p * gamma * v) / m_p + np.array([0, 0, 9.8])
    assert np.allclose(vn, v + a * dt, rtol=1e-14, atol=0)
    assert np.allclose(xn, x + v * dt + 0.5 * a * dt ** 2, rtol=1e-14, atol=0)
    rng = np.random.RandomState(0)
    T = 300.0
    dv = np.array([dust_langevin_step(x, np.zeros(3), np.zeros(3), m_p, r_p, P, v_T, T, dt, g=0.0, rng=rng)[1]
                   for _ in range(4000)])
    var_expect = 2 * 1.380649e-23 * T * m_p * gamma / dt * (dt / m_p) ** 2
    assert abs(dv.var() / var_expect - 1) < 0.06


def test_validation_fixtures_match_the_scan_tables():
    """The reference means the GPU validation scripts compare with parse, and
    their parameter columns are the scans the scripts run (Mach numbers, d_x,
    collision frequencies as the reference printed them)."""
    import importlib.util

    here = os.path.dirname(__file__)

    def script(name):
        spec = importlib.util.spec_from_file_location(name, os.path.join(here, "validation", name + ".py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        return mod

    golden = os.path.join(here, "golden")
    f5 = np.loadtxt(os.path.join(golden, "piel2017_fig5_reference.txt"))
    assert f5.shape == (12, 3) and np.allclose(f5[:, 0], script("piel2017_fig5").MACH, rtol=0, atol=1e-5)
    # the orbit force rises to a maximum near M = 0.8 and falls again, as the reference printed it
    assert np.argmax(f5[:, 1]) == 6 and np.all(f5[:, 1] > 0)
    f6 = np.loadtxt(os.path.join(golden, "piel2017_fig6_reference.txt"))
    assert f6.shape == (7, 5) and np.allclose(f6[:, 0], script("piel2017_fig6").MACH, rtol=0, atol=1e-5)
    f7 = np.loadtxt(os.path.join(golden, "piel2017_fig7_reference.txt"))
    assert f7.shape == (11, 5) and np.allclose(f7[:, 0], script("piel2017_fig7").D_X, rtol=0, atol=1e-5)
    pt = np.loadtxt(os.path.join(golden, "patacchini2008_fig3_reference.txt"))
    s = script("patacchini2008_fig3")
    assert pt.shape == (13, 4) and np.allclose(pt[:, 0], s.W_C, rtol=1e-8, atol=0)
    # the field gives Mach 1 at every collision frequency (v_fl = e E / (m_i w_c))
    assert np.allclose(np.array(s.E_FIELD) / np.array(s.W_C), s.E_FIELD[0] / s.W_C[0], rtol=1e-6)
    assert np.all(pt[:, 1] < 0) and np.all(np.abs(pt[:, 3]) > 0)


def test_bench_traffic_lookup_labels_its_source():
    """The bench's HBM traffic of k_pair32_sym at 2^20 ions: the PMC reading
    of a profiled (variant, block count); for an unprofiled pair, that of
    variant 1 with the same block count, labelled, unless the variant spills
    registers; otherwise none."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert bench.sym_traffic_2e20(6, 4) == (bench.SYM_TRAFFIC_2E20[(6, 4)], "pmc")
    assert bench.sym_traffic_2e20(2, 16) == (bench.SYM_TRAFFIC_2E20[(2, 16)], "pmc")
    assert (8, 8) not in bench.SYM_TRAFFIC_2E20
    t, src = bench.sym_traffic_2e20(8, 8)
    assert t == bench.SYM_TRAFFIC_2E20[(1, 8)] and "variant 1" in src
    assert bench.sym_traffic_2e20(4, 64) == (None, None)
    assert (10, 1) not in bench.SYM_TRAFFIC_2E20 and (1, 1) in bench.SYM_TRAFFIC_2E20
    assert bench.sym_traffic_2e20(10, 1, rows=4) == (None, None)  # grouped: no fallback to variant 1
    assert all(v not in bench.SYM_SPILLING_VARIANTS for v in (1, 3, 5, 6, 7, 8))
    # the LDS-source variants (the defaults, and what the autotuner keeps at
    # 2^20) have their own readings at every block count it times
    for tile, rows in ((11, 1), (12, 4), (13, 2)):
        for chunk in (1, 2, 4, 8, 16, 32):
            assert bench.sym_traffic_2e20(tile, chunk, rows=rows)[1] == "pmc", (tile, chunk)


def test_bench_pair_evaluation_count():
    """The bench's count of pair-term evaluations per step: symmetric main
    launch (one per unordered pair between blocks) + in-block ordered pairs +
    the flagged ions' corrections, against a direct count for a small case."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    n, blk, nf, na = 3000, 1024, 7, 2
    sizes = [min(blk, n - b) for b in range(0, n, blk)]
    between = n * (n - 1) - sum(s * (s - 1) for s in sizes)  # ordered pairs in different blocks
    ev = bench.pair_evaluations(n, True, between, nf, na)
    assert ev["main_launch"] == between // 2
    assert ev["in_block"] == sum(s * (s - 1) for s in sizes)
    assert ev["flagged_corrections"] == 2 * (n - nf) * nf + n * (nf + na)
    assert ev["total"] == ev["main_launch"] + ev["in_block"] + ev["flagged_corrections"]
    rows = bench.pair_evaluations(n, False, n * (n - 1), nf, na)
    assert rows["main_launch"] == n * (n - 1) and rows["in_block"] == 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before
    touching a device."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr


def test_field_driven_vz_density_is_exponentially_modified_gaussian():
    """The field-driven v_z density (opendust.py:738-754) is the
    exponentially modified Gaussian of a drift v_fl with thermal spread v_T
    (an exponential of mean v_fl convolved with N(0, v_T^2)): SciPy's
    exponnorm with K = v_fl / v_T and scale v_T.  The table's CDF knots are
    that CDF renormalised on [v_min, v_max]."""
    from scipy import stats

    for v_T, v_fl in ((0.2457, 8 * 0.2457), (1.0, 0.8), (0.3, 2.5)):
        K = v_fl / v_T
        f = host._density_field_driven(v_fl, v_T)
        lo, hi = -4 * v_T, 6 * abs(v_fl) + 5 * v_T
        v = np.linspace(lo, hi, 301)
        ref = stats.exponnorm.pdf(v, K, scale=v_T)
        assert np.allclose(f(v), ref, rtol=1e-9, atol=1e-12 * ref.max())
        t = host.vz_table(v_fl, v_T)
        c = lambda x: stats.exponnorm.cdf(x, K, scale=v_T)
        xi = (c(t.grid_v) - c(lo)) / (c(hi) - c(lo))
        assert np.max(np.abs(t.grid_xi - xi)) <= 1e-8
        # the top face's flux weights the same density by v (opendust.py:673-689)
        flux = host._flux_field_driven(v_fl, v_T)
        assert np.allclose(flux(v), v * ref, rtol=1e-9, atol=1e-12 * np.max(np.abs(v * ref)))


def test_blocking_error_of_the_validation_scripts():
    """The blocking estimate the validation scripts quote: white noise gives
    sigma / sqrt(n); an AR(1) series gives the analytic standard error of its
    mean, sigma / sqrt(n) * sqrt((1 + phi) / (1 - phi)), where 10 fixed blocks of
    a short window would understate it."""
    import importlib.util

    here = os.path.dirname(__file__)
    spec = importlib.util.spec_from_file_location("fig6", os.path.join(here, "validation", "piel2017_fig6.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.RandomState(5)
    n = 2 ** 16
    w = rng.normal(size=n)
    assert abs(mod.blocking_error(w) / (1 / np.sqrt(n)) - 1) < 0.3
    phi = 0.98
    x = np.empty(n)
    x[0] = rng.normal() / np.sqrt(1 - phi ** 2)
    e = rng.normal(size=n)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + e[i]
    sigma = 1 / np.sqrt(1 - phi ** 2)
    exact = sigma / np.sqrt(n) * np.sqrt((1 + phi) / (1 - phi))
    assert abs(mod.blocking_error(x) / exact - 1) < 0.35
