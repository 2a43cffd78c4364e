"""Pins the FP64 CPU oracle (oracle/od_oracle.c) without importing the
reference: known-answer vectors, finite differences of the reference's
potential-energy expressions, closed forms, physical limits and distribution
tests.  Each check is chosen so that a plausible mistake (a dropped term, a
wrong sign or index, a swapped operand) fails it."""

import json
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy import integrate, special, stats

from opendust_amd import host, workloads

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def orc():
    from oracle import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="module")
def base():
    """single-dust workload, small N, a random trap polynomial."""
    od = workloads.single_dust(N=64, n=3, seed=17)
    alpha = np.random.RandomState(5).normal(size=15) * 10.0 ** -np.arange(15) * 0.3
    cfg, x0, v0 = workloads.prepare(od, alpha=alpha, np_seed=2)
    return od, cfg, x0, v0


# ------------------------------------------------------------------ Philox
def test_philox_known_answers(orc):
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as fh:
        kat = json.load(fh)
    for v in kat["vectors"]:
        ctr = [int(c, 16) for c in v["ctr"]]
        key = int(v["key"][0], 16) + (int(v["key"][1], 16) << 32)
        assert orc.philox(ctr, key) == [int(o, 16) for o in v["out"]]


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc needed to build the rocRAND cross-check")
def test_philox_matches_rocrand(orc, tmp_path):
    exe = tmp_path / "philox_check"
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    subprocess.run([hipcc, "-O1", "-o", str(exe), os.path.join(HERE, "golden", "philox_rocrand_check.cpp")],
                   check=True, capture_output=True)
    rng = np.random.RandomState(3)
    rows = rng.randint(0, 2 ** 32, size=(200, 6), dtype=np.uint64)
    inp = "\n".join(" ".join("%x" % int(v) for v in r) for r in rows) + "\n"
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for r, line in zip(rows, out):
        got = orc.philox([int(v) for v in r[:4]], int(r[4]) + (int(r[5]) << 32))
        assert got == [int(w, 16) for w in line.split()]


# ------------------------------------------------------------------ forces
def _pair_energy(cfg, x):
    """Total ion-ion energy, the reference's expression kqq*exp(-r/r_D_e)/(r+r_0)."""
    e = 0.0
    n = len(x)
    for i in range(n):
        d = x[i + 1:] - x[i]
        r = np.sqrt((d ** 2).sum(1))
        e += np.sum(cfg.kqq * np.exp(-r / cfg.r_D_e) / (r + cfg.r_0))
    return e


def test_pair_force_is_minus_gradient_of_reference_energy(orc, base):
    _, cfg, _, _ = base
    rng = np.random.RandomState(1)
    lam = cfg.r_D_e
    x = rng.uniform(-0.8, 0.8, size=(14, 3)) * lam
    x[3] = x[2] + np.array([0.01, -0.004, 0.02]) * lam  # one close pair
    f = orc.pair_forces(cfg, x, nthreads=2)
    h = 1e-6 * lam
    for i in (0, 2, 3, 9):
        for k in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            fd = -(_pair_energy(cfg, xp) - _pair_energy(cfg, xm)) / (2 * h)
            assert abs(fd - f[i, k]) <= 1e-6 * np.abs(f[i]).max()
    # Newton's third law: the ion-ion forces sum to zero
    assert np.linalg.norm(f.sum(0)) <= 1e-12 * np.abs(f).sum()


def _switched_energy(cfg, x, rc, rs):
    e = 0.0
    for i in range(len(x)):
        d = x[i + 1:] - x[i]
        r = np.sqrt((d ** 2).sum(1))
        u = cfg.kqq * np.exp(-r / cfg.r_D_e) / (r + cfg.r_0)
        t = np.clip((r - rs) / (rc - rs), 0.0, 1.0)
        S = 1 - 10 * t ** 3 + 15 * t ** 4 - 6 * t ** 5
        e += np.sum(np.where(r < rc, u * S, 0.0))
    return e


@pytest.mark.parametrize("switch", [True, False])
def test_truncated_force_is_minus_gradient_of_switched_energy(orc, base, switch):
    """OpenMM's switched cutoff (opendust.py:481-485): U S(x), S = 1 - 10x^3 +
    15x^4 - 6x^5 between the switching distance and the cutoff."""
    _, cfg, _, _ = base
    lam = cfg.r_D_e
    rng = np.random.RandomState(4)
    x = rng.uniform(-2.0, 2.0, size=(24, 3)) * lam
    x[1] = x[0] + np.array([3.1, 0.0, 0.0]) * lam  # inside the switching band
    rc, rs = 3.2 * lam, (3.0 * lam if switch else 3.2 * lam)
    cfg.cutoff, cfg.switch_dist = rc, (rs if switch else 0.0)
    try:
        f = orc.pair_forces(cfg, x, nthreads=2)
        h = 1e-6 * lam
        for i in (0, 1, 5):
            for k in range(3):
                xp = x.copy()
                xm = x.copy()
                xp[i, k] += h
                xm[i, k] -= h
                fd = -(_switched_energy(cfg, xp, rc, rs) - _switched_energy(cfg, xm, rc, rs)) / (2 * h)
                assert abs(fd - f[i, k]) <= 1e-6 * np.abs(f[i]).max() + 1e-12 * np.abs(f).max()
        # pairs beyond the cutoff do not interact
        far = np.array([[0.0, 0.0, 0.0], [3.3 * lam, 0.0, 0.0]])
        assert np.all(orc.pair_forces(cfg, far, nthreads=1) == 0.0)
    finally:
        cfg.cutoff = cfg.switch_dist = 0.0


def test_binning_spec(orc, base):
    """Columns of cutoff/4 in x and y, slabs of cutoff/16 in z, z fastest
    (a column's cells are consecutive), clamped to the grid."""
    _, cfg, _, _ = base
    cfg.cutoff = 1.5 * cfg.r_D_e
    try:
        w, h = cfg.cutoff / 4, cfg.cutoff / 16
        cid, dims = orc.cell_of(cfg, np.array([-cfg.R, -cfg.R, -cfg.H / 2]))
        assert cid == 0 and dims == (int(np.ceil(2 * cfg.R / w)),) * 2 + (int(np.ceil(cfg.H / h)),)
        nx, ny, nz = dims
        cid, _ = orc.cell_of(cfg, np.array([cfg.R * 2, cfg.R * 2, cfg.H]))  # clamped to the last cell
        assert cid == nx * ny * nz - 1
        p = np.array([-cfg.R + 2.5 * w, -cfg.R + 0.5 * w, -cfg.H / 2 + 1.5 * h])
        assert orc.cell_of(cfg, p)[0] == 1 + nz * (2 + nx * 0)
        # one slab up is the next cell, one column over is nz cells on
        assert orc.cell_of(cfg, p + [0, 0, h])[0] == orc.cell_of(cfg, p)[0] + 1
        assert orc.cell_of(cfg, p + [w, 0, 0])[0] == orc.cell_of(cfg, p)[0] + nz
        assert orc.cell_of(cfg, p + [0, w, 0])[0] == orc.cell_of(cfg, p)[0] + nz * nx
    finally:
        cfg.cutoff = 0.0


def test_cell_offsets_cover_the_cutoff_sphere():
    """The neighbour columns and slab reaches the cell kernels visit (77
    column offsets, slab reach 1 + max{j : j^2 < 256 - 16 (mx^2 + my^2)})
    contain every cell that can hold an ion within the cutoff of an ion in
    the target cell, and drop only cells that cannot: brute force over random
    pairs of points in cells of the grid."""
    w, h, rc = 0.25, 1.0 / 16.0, 1.0

    def m(d):
        return max(abs(d) - 1, 0)

    def nbz(dx, dy):
        v = 256 - 16 * (m(dx) ** 2 + m(dy) ** 2)
        j = 0
        while (j + 1) ** 2 < v:
            j += 1
        return j + 1

    offs = {(dx, dy) for dx in range(-4, 5) for dy in range(-4, 5) if 16 * (m(dx) ** 2 + m(dy) ** 2) < 256}
    assert len(offs) == 77 and (4, 4) not in offs and (4, 3) in offs
    fwd = [(dx, 0) for dx in range(1, 5)] + [(dx, dy) for dy in range(1, 4) for dx in range(-4, 5)] + \
        [(dx, 4) for dx in range(-3, 4)]
    assert len(fwd) == 38 and set(fwd) | {(0, 0)} | {(-a, -b) for a, b in fwd} == offs
    rng = np.random.default_rng(5)
    for dx in range(-5, 6):
        for dy in range(-5, 6):
            for dz in range(-18, 19):
                # closest approach of two points in cells (0, 0, 0) and (dx, dy, dz)
                gap = np.array([m(dx) * w, m(dy) * w, m(dz) * h])
                reachable = np.linalg.norm(gap) < rc
                listed = (dx, dy) in offs and abs(dz) <= nbz(dx, dy)
                assert reachable == listed, (dx, dy, dz)
                if reachable:
                    # the bound is attained: random points come arbitrarily close to the gap
                    a = rng.uniform(0, 1, size=(2000, 3)) * [w, w, h]
                    b = (rng.uniform(0, 1, size=(2000, 3)) + [dx, dy, dz]) * [w, w, h]
                    assert np.min(np.linalg.norm(a - b, axis=1)) >= np.linalg.norm(gap) - 1e-12


def test_binning_spec_beyond_16_bit_cell_ids(orc, base):
    """A grid of more than 65535 cells (R = 10, H = 40 r_D_e at a 1 r_D_e
    cutoff: 80 x 80 x 640 cells), where cell ids need a third 8-bit radix
    digit: the same formula, no wrap-around."""
    _, cfg, _, _ = base
    save = (cfg.R, cfg.H, cfg.cutoff)
    lam = cfg.r_D_e
    try:
        cfg.R, cfg.H, cfg.cutoff = 10 * lam, 40 * lam, 1.0 * lam
        w, h = cfg.cutoff / 4, cfg.cutoff / 16
        cid, dims = orc.cell_of(cfg, np.array([cfg.R, cfg.R, cfg.H / 2]))
        assert dims == (80, 80, 640) and cid == 80 * 80 * 640 - 1 > 65535
        p = np.array([-cfg.R + 17.5 * w, -cfg.R + 33.5 * w, -cfg.H / 2 + 471.5 * h])
        assert orc.cell_of(cfg, p)[0] == 471 + 640 * (17 + 80 * 33)
    finally:
        cfg.R, cfg.H, cfg.cutoff = save


def _ext_energy(cfg, q, x):
    e = 0.0
    for p in range(cfg.n_dust):
        r = np.linalg.norm(x - np.array([cfg.dust_x[p], cfg.dust_y[p], cfg.dust_z[p]]))
        e += q[p] * cfg.k_dust / r
    rho = np.hypot(x[0], x[1]) / cfg.r_D_e
    h = x[2] / cfg.r_D_e
    e += cfg.trap_C * host.trap_potential_poly(cfg.trap_alpha, rho, h)
    if cfg.distribution == host.DIST_FIELD_DRIVEN:
        e += -cfg.field_force * x[2]
    return e


@pytest.mark.parametrize("field", [False, True])
def test_external_force_is_minus_gradient(orc, base, field):
    _, cfg, _, _ = base
    if field:
        cfg.distribution = host.DIST_FIELD_DRIVEN
        cfg.field_force = 3.7e-3
    q = np.array([cfg.dust_q0[0]])
    lam = cfg.r_D_e
    try:
        for x in (np.array([0.31, -0.52, 1.7]) * lam, np.array([2.2, 1.4, -3.1]) * lam,
                  np.array([0.0, 0.0, 0.4]) * lam):
            f = orc.ext_force(cfg, q, x)
            h = 1e-6 * lam
            for k in range(3):
                xp = x.copy()
                xm = x.copy()
                xp[k] += h
                xm[k] -= h
                fd = -(_ext_energy(cfg, q, xp) - _ext_energy(cfg, q, xm)) / (2 * h)
                assert abs(fd - f[k]) <= 1e-6 * np.abs(f).max() + 1e-300
    finally:
        cfg.distribution = host.DIST_MAXWELLIAN
        cfg.field_force = 0.0


def test_orbit_term_is_yukawa_force_on_dust(orc, base):
    """q k q_s O / lambda^2 (SI) = -grad_dust of k q q_s exp(-r/lambda)/r."""
    _, cfg, _, _ = base
    lam_nm, lam = cfg.r_D_e, cfg.si_r_D_e
    xi = np.array([0.4, -0.3, 0.9]) * lam_nm
    o = orc.orbit_term(cfg, 0, xi)
    q = cfg.dust_q0[0]
    xp = np.array([cfg.dust_x[0], cfg.dust_y[0], cfg.dust_z[0]]) * 1e-9
    xi_m = xi * 1e-9

    def U(d):
        r = np.linalg.norm(d - xi_m)
        return cfg.si_k * q * cfg.si_q_s * np.exp(-r / lam) / r

    f = o * q * cfg.si_k * cfg.si_q_s / lam ** 2
    h = 1e-7 * lam
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd = -(U(xp + e) - U(xp - e)) / (2 * h)
        assert abs(fd - f[k]) <= 1e-6 * np.abs(f).max()


def test_out_domain_force_is_minus_gradient(orc, base):
    _, cfg, _, _ = base
    lam = cfg.si_r_D_e
    qp = -7.7e-15
    save = (cfg.si_dust_x[0], cfg.si_dust_y[0], cfg.si_dust_z[0])
    try:
        pos = np.array([0.7, -0.2, -1.1]) * lam

        def U(p):
            return (cfg.si_k * qp * cfg.si_q_s / lam
                    * host.trap_potential_poly(cfg.trap_alpha, np.hypot(p[0], p[1]) / lam, p[2] / lam))

        cfg.si_dust_x[0], cfg.si_dust_y[0], cfg.si_dust_z[0] = pos
        f = orc.out_domain_force(cfg, 0, qp)
        h = 1e-6 * lam
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fd = -(U(pos + e) - U(pos - e)) / (2 * h)
            assert abs(fd - f[k]) <= 1e-6 * np.abs(f).max()
    finally:
        cfg.si_dust_x[0], cfg.si_dust_y[0], cfg.si_dust_z[0] = save


# ------------------------------------------------------------ integrator
def _no_dust(cfg):
    cfg.n_dust = 0
    return cfg


def test_kdk_constant_force_is_exact(orc):
    od = workloads.dust_chain(N=1, n=2, seed=3)
    cfg, _, _ = workloads.prepare(od, alpha=np.zeros(15), np_seed=1)
    cfg.n_dust = 0
    cfg.thr_coll = 0
    x0 = np.array([[0.1, 0.2, 0.3]]) * cfg.r_D_e
    v0 = np.array([[0.01, -0.02, 0.05]])
    x, v, _, _, fl, _ = orc.step(cfg, x0, v0, [], 4)
    a = np.array([0.0, 0.0, cfg.field_force]) / cfg.m_s
    v1 = v0 + 0.5 * cfg.dt * a
    assert np.allclose(x, x0 + cfg.dt * v1, rtol=1e-15, atol=0)
    assert np.allclose(v, v1 + 0.5 * cfg.dt * a, rtol=1e-15, atol=0)
    assert fl[0] == 0


def test_step_is_kick_drift_kick_with_forces_at_drifted_positions(orc, base):
    _, cfg, x0, v0 = base
    cfg = _no_dust(cfg)
    try:
        x0 = x0[:24] * 0.5
        v0 = v0[:24]
        q = np.zeros(1)

        def F(x):
            return orc.pair_forces(cfg, x, nthreads=2) + np.array([orc.ext_force(cfg, q, xi) for xi in x])

        v1 = v0 + 0.5 * cfg.dt * F(x0) / cfg.m_s
        x1 = x0 + cfg.dt * v1
        v2 = v1 + 0.5 * cfg.dt * F(x1) / cfg.m_s
        x, v, _, _, fl, _ = orc.step(cfg, x0, v0, q, 1)
        assert np.all(fl == 0)
        assert np.allclose(x, x1, rtol=1e-14, atol=0)
        assert np.allclose(v, v2, rtol=1e-12, atol=1e-14 * np.abs(v2).max())
    finally:
        cfg.n_dust = 1


def test_absorbed_momentum_uses_post_kick_velocity_and_orbit_uses_final_positions(orc, base):
    _, cfg, x0, v0 = base
    x0 = x0[:16].copy()
    v0 = v0[:16].copy()
    x0[0] = [cfg.dust_x[0] + 0.4 * cfg.dust_r[0], cfg.dust_y[0], cfg.dust_z[0]]
    q = np.array([cfg.dust_q0[0]])

    def F(x):
        return orc.pair_forces(cfg, x, nthreads=2) + np.array([orc.ext_force(cfg, q, xi) for xi in x])

    v1 = v0 + 0.5 * cfg.dt * F(x0) / cfg.m_s
    x1 = x0 + cfg.dt * v1
    v2 = v1 + 0.5 * cfg.dt * F(x1) / cfg.m_s
    x, v, qn, rec, fl, face = orc.step(cfg, x0, v0, q, 2)
    assert fl[0] & 1 and face[0] in (1, 2)
    assert rec["absorbed"][0] == 1
    assert np.allclose(rec["momentum_sum"][0], v2[0], rtol=1e-12)
    assert np.allclose(rec["force_collect"][0], v2[0] * 1e3 * 6.63e-26 * workloads.single_dust(N=64).N_s
                       / cfg.si_dt, rtol=1e-12)
    orb = sum(orc.orbit_term(cfg, 0, xi) for xi in x)
    assert np.allclose(rec["orbit_sum"][0], orb, rtol=1e-12)
    # reinjected ion sits on a face of the cylinder
    rho = np.hypot(x[0, 0], x[0, 1])
    assert (face[0] == 1 and x[0, 2] == -cfg.H / 2) or (face[0] == 2 and abs(rho - cfg.R) < 1e-9 * cfg.R)


def test_flags_boundaries_use_greater_equal(orc, base):
    _, cfg, _, _ = base
    save = (cfg.dust_x[0], cfg.dust_y[0], cfg.dust_z[0], cfg.dust_r[0], cfg.R, cfg.H)
    try:
        cfg.dust_x[0] = cfg.dust_y[0] = cfg.dust_z[0] = 0.0
        cfg.dust_r[0] = 1024.0
        cfg.R = 2.0 ** 21
        cfg.H = 2.0 ** 23
        assert orc.flags(cfg, [0.0, 0.0, 1024.0]) == 1
        assert orc.flags(cfg, [0.0, 0.0, 1024.0 * (1 + 1e-15)]) == 0
        assert orc.flags(cfg, [2.0 ** 21, 0.0, 5000.0]) == 1 << 16
        assert orc.flags(cfg, [2.0 ** 21 * (1 - 1e-15), 0.0, 5000.0]) == 0
        assert orc.flags(cfg, [3000.0, 0.0, -2.0 ** 22]) == 1 << 17
        assert orc.flags(cfg, [3000.0, 0.0, 2.0 ** 22 * (1 - 1e-15)]) == 0
    finally:
        (cfg.dust_x[0], cfg.dust_y[0], cfg.dust_z[0], cfg.dust_r[0], cfg.R, cfg.H) = save


# ------------------------------------------------------------ reinjection
def _reinject_many(orc, cfg, n):
    faces = np.empty(n, dtype=int)
    xs = np.empty((n, 3))
    vs = np.empty((n, 3))
    for i in range(n):
        faces[i], xs[i], vs[i] = orc.reinject(cfg, i, 7)
    return faces, xs, vs


def test_reinjection_distributions(orc, base):
    od, cfg, _, _ = base
    n = 60000
    faces, xs, vs = _reinject_many(orc, cfg, n)
    p_t = cfg.thr_top / 2 ** 32
    frac_top = np.mean(faces == 1)
    assert abs(frac_top - p_t) <= 5 * np.sqrt(p_t * (1 - p_t) / n)
    assert np.all(faces != 3)  # p_b = 0 at M = 1
    t = faces == 1
    assert np.all(xs[t, 2] == -cfg.H / 2)
    assert stats.kstest((xs[t, 0] ** 2 + xs[t, 1] ** 2) / cfg.R ** 2, "uniform").pvalue > 1e-4
    assert stats.kstest(vs[t, 0] / cfg.v_T, "norm").pvalue > 1e-4
    # top v_z against the exact flux-weighted CDF on [v_fl - 4 v_T, v_fl + 4 v_T]
    a, b = max(cfg.v_fl - 4 * cfg.v_T, 0), cfg.v_fl + 4 * cfg.v_T
    s2 = np.sqrt(2) * cfg.v_T

    def G(v, vf):
        return cfg.v_T / np.sqrt(2 * np.pi) * (-np.exp(-((v - vf) / s2) ** 2)) + vf / 2 * special.erf((v - vf) / s2)

    cdf = lambda v: (G(np.clip(v, a, b), cfg.v_fl) - G(a, cfg.v_fl)) / (G(b, cfg.v_fl) - G(a, cfg.v_fl))
    D = stats.kstest(vs[t, 2], cdf).statistic
    assert D <= 1.63 / np.sqrt(t.sum()) + 3e-3  # + piecewise-linear interpolation error
    s = faces == 2
    rho = np.hypot(xs[s, 0], xs[s, 1])
    assert np.allclose(rho, cfg.R, rtol=1e-12)
    assert stats.kstest((xs[s, 2] + cfg.H / 2) / cfg.H, "uniform").pvalue > 1e-4
    nr = np.stack([xs[s, 0], xs[s, 1]], 1) / rho[:, None]
    v_in = -(vs[s, 0] * nr[:, 0] + vs[s, 1] * nr[:, 1])
    v_tan = vs[s, 0] * nr[:, 1] - vs[s, 1] * nr[:, 0]
    cdf_s = lambda v: (G(np.clip(v, 0, 4 * cfg.v_T), 0) - G(0, 0)) / (G(4 * cfg.v_T, 0) - G(0, 0))
    assert stats.kstest(v_in, cdf_s).statistic <= 1.63 / np.sqrt(s.sum()) + 3e-3
    assert stats.kstest(v_tan / cfg.v_T, "norm").pvalue > 1e-4
    assert stats.kstest((vs[s, 2] - cfg.v_fl) / cfg.v_T, "norm").pvalue > 1e-4


def _flux(u, v_T, a, b):
    """Flux of a Maxwellian with drift u (toward the face) through the face,
    per unit density, over normal speeds [a, b]: int_a^b v N(v; u, v_T) dv,
    from the antiderivative -v_T^2 N(v; u, v_T) + u Phi((v - u) / v_T)."""
    s2 = np.sqrt(2) * v_T
    G = lambda v: -v_T / np.sqrt(2 * np.pi) * np.exp(-((v - u) / s2) ** 2) + u / 2 * special.erf((v - u) / s2)
    return G(b) - G(a)


@pytest.mark.parametrize("mach", [0.1, 0.3, 1.0])
def test_face_probabilities_are_closed_form_fluxes(mach):
    """p_t, p_s, p_b (opendust.py:806-837) are the faces' shares of the
    inflowing flux: top pi R^2 G(v_fl), bottom pi R^2 G(-v_fl), side 2 pi R H
    G(0), with G the closed-form flux of the drifting Maxwellian over the
    speed ranges the reference integrates (opendust.py:608-632): top
    [max(v_fl - 4 v_T, 0), v_fl + 4 v_T], bottom [0, 4 v_T - v_fl] (empty,
    p_b = 0, once that is below 0.4 v_T, i.e. M >= 0.36 here), side [0, 4 v_T]."""
    import contextlib
    import io

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=64, n=3, mach=mach)
        workloads.prepare(od, alpha=np.zeros(15), np_seed=1)
    v_T, v_fl = od.v_T, od.v_fl
    G_t = _flux(v_fl, v_T, max(v_fl - 4 * v_T, 0), v_fl + 4 * v_T)
    G_s = _flux(0.0, v_T, 0.0, 4 * v_T)
    hi_b = max(-v_fl + 4 * v_T, 0)
    G_b = _flux(-v_fl, v_T, 0.0, hi_b) if hi_b >= 0.4 * v_T else 0.0
    area_t, area_s = np.pi * od.R ** 2, 2 * np.pi * od.R * od.H
    tot = area_t * (G_t + G_b) + area_s * G_s
    assert od.p_t == pytest.approx(area_t * G_t / tot, rel=1e-7)
    assert od.p_s == pytest.approx(area_s * G_s / tot, rel=1e-7)
    if mach >= 0.36:
        assert od.p_b == 0.0
    else:
        assert od.p_b > 0.0
        assert od.p_b == pytest.approx(area_t * G_b / tot, rel=1e-6)


def test_bottom_face_reinjection_distributions(orc):
    """Bottom face (opendust.py:939-951, used when p_b != 0): at M = 0.1 the
    reference's p_b = 0.030.  With the real thresholds the bottom share
    matches p_b; with thresholds moved so that 60 % of the draws choose the
    bottom, the bottom ions sit at z = +H/2 exactly, uniform over the disc,
    v_x and v_y ~ N(0, v_T), and -v_z follows the flux-weighted distribution
    of the drift -v_fl on [0, max(4 v_T - v_fl, 0)] (the table built with
    -v_fl, opendust.py:625-632)."""
    import contextlib
    import io

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=64, n=3, mach=0.1, seed=23)
        cfg, _, _ = workloads.prepare(od, alpha=np.zeros(15), np_seed=1)
    assert cfg.use_bottom == 1 and cfg.bottom.n_seg > 0
    n = 40000
    faces, _, _ = _reinject_many(orc, cfg, n)
    p_b = 1.0 - cfg.thr_side / 2 ** 32
    assert p_b == pytest.approx(od.p_b, abs=1e-9)
    assert abs(np.mean(faces == 3) - p_b) <= 5 * np.sqrt(p_b * (1 - p_b) / n)
    cfg.thr_top = host.u32_threshold(0.2)
    cfg.thr_side = host.u32_threshold(0.4)
    faces, xs, vs = _reinject_many(orc, cfg, n)
    b = faces == 3
    assert abs(np.mean(b) - 0.6) <= 5 * np.sqrt(0.24 / n)
    assert abs(np.mean(faces == 1) - 0.2) <= 5 * np.sqrt(0.16 / n)
    assert np.all(xs[b, 2] == cfg.H / 2)
    assert np.all(xs[faces == 1, 2] == -cfg.H / 2)
    assert stats.kstest((xs[b, 0] ** 2 + xs[b, 1] ** 2) / cfg.R ** 2, "uniform").pvalue > 1e-4
    assert stats.kstest(vs[b, 0] / cfg.v_T, "norm").pvalue > 1e-4
    assert stats.kstest(vs[b, 1] / cfg.v_T, "norm").pvalue > 1e-4
    assert np.all(vs[b, 2] <= 0.0)
    s2 = np.sqrt(2) * cfg.v_T
    vf = -cfg.v_fl
    hi = max(vf + 4 * cfg.v_T, 0.0)

    def G(v):
        return cfg.v_T / np.sqrt(2 * np.pi) * (-np.exp(-((v - vf) / s2) ** 2)) + vf / 2 * special.erf((v - vf) / s2)

    cdf = lambda w: (G(np.clip(w, 0, hi)) - G(0.0)) / (G(hi) - G(0.0))
    # the table's knots are the closed-form CDF of that flux; between knots the
    # sampled speed is linear in u, so its CDF is the linear interpolant of
    # the knots, which the sample must follow to the KS bound
    _, tab = host.face_table(vf, cfg.v_T, "bottom", "Maxwellian")
    knots = -tab.grid_v
    assert np.all(np.diff(knots) > 0) and knots[0] == 0.0 and knots[-1] == pytest.approx(hi, rel=1e-12)
    assert np.max(np.abs(tab.grid_xi - cdf(knots))) <= 1e-7
    D = stats.kstest(-vs[b, 2], lambda w: np.interp(w, knots, cdf(knots))).statistic
    assert D <= 1.63 / np.sqrt(b.sum())
    # a sign slip (the top table, +v_fl) is far outside that bound
    vf_wrong = cfg.v_fl
    G_w = lambda v: (cfg.v_T / np.sqrt(2 * np.pi) * (-np.exp(-((v - vf_wrong) / s2) ** 2))
                     + vf_wrong / 2 * special.erf((v - vf_wrong) / s2))
    hi_w = vf_wrong + 4 * cfg.v_T
    cdf_w = lambda w: (G_w(np.clip(w, 0, hi_w)) - G_w(0.0)) / (G_w(hi_w) - G_w(0.0))
    assert stats.kstest(-vs[b, 2], cdf_w).statistic > 10 * (1.63 / np.sqrt(b.sum()))


def test_cut_edge_count_matches_brute_force(orc):
    """The test helper counting pairs near the cutoff, against NumPy."""
    import contextlib
    import io

    with contextlib.redirect_stdout(io.StringIO()):
        od = workloads.single_dust(N=400, n=3, R=1.0, H=2.0)
        cfg, x0, _ = workloads.prepare(od, alpha=np.zeros(15), np_seed=4)
    cfg.cutoff = 0.7 * cfg.r_D_e
    d = np.linalg.norm(x0[:, None, :] - x0[None, :, :], axis=2)
    np.fill_diagonal(d, -1.0)
    for rel in (1e-2, 1e-3):
        ref = np.sum(np.abs(d - cfg.cutoff) <= rel * cfg.cutoff, axis=1)
        assert np.array_equal(orc.cut_edge_count(cfg, x0, rel, nthreads=2), ref)
    assert orc.cut_edge_count(cfg, x0, 1e-2).sum() > 0


def test_collisions_rate_and_distribution(orc):
    od = workloads.dust_chain(N=8, n=2, seed=3)
    cfg, _, _ = workloads.prepare(od, alpha=np.zeros(15), np_seed=1)
    cfg.thr_coll = host.u32_threshold(0.3)
    n = 30000
    hit = []
    new = []
    for i in range(n):
        v = orc.collide(cfg, i, 11, [1e3, 1e3, 1e3])
        hit.append(v != 1e3)
        new.extend(v[v != 1e3])
    rate = np.mean(hit)
    assert abs(rate - 0.3) <= 5 * np.sqrt(0.3 * 0.7 / (3 * n))
    assert stats.kstest(np.array(new) / cfg.v_T, "norm").pvalue > 1e-4


def test_table_eval_agrees_with_host(orc, base):
    _, cfg, _, _ = base
    od = workloads.single_dust(N=64)
    top = host.face_table(od.v_fl, od.v_T, "top", "Maxwellian")[1]
    for u in np.linspace(0, 1, 97)[:-1] + 1e-4:
        assert orc.table_eval(cfg.top, u) == top(np.array([u]))[0]


# ------------------------------------------------------------------ charge
def test_charge_given_and_step_zero(orc, base):
    _, cfg, _, _ = base
    save = cfg.dust_method[0]
    try:
        cfg.dust_method[0] = host.CHARGE_GIVEN
        assert orc.charge(cfg, 0, 5, -3e-15, 4) == -3e-15
        cfg.dust_method[0] = host.CHARGE_CONSISTENT
        assert orc.charge(cfg, 0, 0, -3e-15, 4) == -3e-15  # step 0 is never credited
    finally:
        cfg.dust_method[0] = save


def test_consistent_charge_is_textbook_electron_flux_plus_ions(orc, base):
    _, cfg, _, _ = base
    e = cfg.si_e
    r = cfg.si_dust_r[0]
    q = -4.1e-15
    phi = q / (4 * np.pi * 8.85418781762039e-12 * r)
    m_e = 9.0938356e-31  # the reference's constant (opendust.py:67)
    n_inf = 3.57e14
    # Boltzmann-repelled electron flux n sqrt(kT/(2 pi m)) exp(|e| phi / kT) onto 4 pi r^2
    flux = n_inf * np.sqrt(cfg.si_kB * cfg.si_Te / (2 * np.pi * m_e)) * np.exp(abs(e) * phi / (cfg.si_kB * cfg.si_Te))
    expected = q + 3 * cfg.si_q_s + flux * cfg.si_dt * 4 * np.pi * r ** 2 * e
    assert abs(orc.charge(cfg, 0, 9, q, 3) - expected) <= 1e-12 * abs(expected)


def test_oml_ion_current_fast_beam_limit(orc):
    """For u = v_fl/(sqrt(2) v_T) >> 1 the reference's OML ion flux tends to
    the beam flux n v_fl / 4 (1 + 2 v_T^2 chi / v_fl^2), up to its 1/(4u) term."""
    pp = workloads.PlasmaParametersInSIUnitsMaxwell(290.11, 29011, 1e14, 7.0, 6.6335209e-26)
    lam = pp.r_D_e
    sp = workloads.SimulationParametersInSIUnits(3 * lam, 6 * lam, 64, 3, 1e-9, pp)
    d = workloads.DustParticle(0, 0, 0, 1e-5, "oml", -1e-15)
    odo = workloads.OpenDust(pp, sp, workloads.OutputParameters(), [d], "Maxwellian")
    cfg, _, _ = workloads.prepare(odo, alpha=np.zeros(15), np_seed=1)
    u = pp.v_fl / pp.v_T / np.sqrt(2)
    assert u > 45
    q = -1e-15
    r = cfg.si_dust_r[0]
    S = 4 * np.pi * r ** 2
    cfg.dust_method[0] = host.CHARGE_CONSISTENT
    dq_e = orc.charge(cfg, 0, 5, q, 0) - q  # electron part only
    cfg.dust_method[0] = host.CHARGE_OML
    dq = orc.charge(cfg, 0, 5, q, 0) - q
    gamma_i = (dq - dq_e) / (cfg.si_dt * S * abs(cfg.si_e))
    phi = cfg.si_k * q / r
    chi = cfg.si_e * phi / (cfg.si_kB * cfg.si_Tn)
    beam = pp.n_inf * pp.v_fl / 4 * (1 + 2 * pp.v_T ** 2 * chi / pp.v_fl ** 2)
    assert abs(gamma_i / beam - 1) <= 1.0 / (4 * u) + 1e-3


def test_dust_dust_force_is_coulomb_gradient(orc):
    od = workloads.dust_chain(N=16, n=2, seed=3)
    cfg, x0, v0 = workloads.prepare(od, alpha=np.zeros(15), np_seed=1)
    q = np.array([-3e-15, -2.5e-15, -4e-15])
    for p in range(3):
        cfg.dust_method[p] = host.CHARGE_GIVEN
    cfg.thr_coll = 0
    _, _, _, rec, _, _ = orc.step(cfg, x0, v0, q, 1)
    pos = np.array([[cfg.si_dust_x[p], cfg.si_dust_y[p], cfg.si_dust_z[p]] for p in range(3)])

    def U(P):
        s = 0.0
        for a in range(3):
            for b in range(a + 1, 3):
                s += cfg.si_k * q[a] * q[b] / np.linalg.norm(P[a] - P[b])
        return s

    h = 1e-6 * cfg.si_r_D_e
    for p in range(3):
        for k in range(3):
            Pp = pos.copy()
            Pm = pos.copy()
            Pp[p, k] += h
            Pm[p, k] -= h
            fd = -(U(Pp) - U(Pm)) / (2 * h)
            assert abs(fd - rec["force_dust"][p][k]) <= 1e-6 * np.abs(rec["force_dust"][p]).max() + 1e-30
    # external field force on the dusts: q E along z
    assert np.allclose(rec["force_field"][:, 2], q * cfg.si_E, rtol=1e-15)


# ------------------------------------------------------------- trap MC
def test_trap_sampler_uniform_in_cylinder(orc):
    R, H = 5.0, 10.0
    s = np.array([orc.trap_sample(R, H, 77, k) for k in range(20000)])
    assert stats.kstest((s[:, 0] ** 2 + s[:, 1] ** 2) / R ** 2, "uniform").pvalue > 1e-4
    assert stats.kstest((s[:, 2] + H / 2) / H, "uniform").pvalue > 1e-4
    assert stats.kstest((np.arctan2(s[:, 1], s[:, 0]) + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-4


def test_trap_grid_center_against_quadrature(orc):
    """Mean of exp(-d)/(d+0.001) over the cylinder seen from its centre, by
    ray quadrature: (2 pi / V) int sin(t) G(L(t)) dt, G(L) = int_0^L s^2 e^-s/(s+0.001) ds."""
    R, H = 3.0, 6.0
    G = lambda L: integrate.quad(lambda s: s * s * np.exp(-s) / (s + 0.001), 0, L, limit=200)[0]
    Lr = lambda t: min(R / max(np.sin(t), 1e-300), (H / 2) / max(abs(np.cos(t)), 1e-300))
    I = 2 * np.pi * integrate.quad(lambda t: np.sin(t) * G(Lr(t)), 0, np.pi, limit=200, points=[np.arctan(2 * R / H),
                                                                                               np.pi - np.arctan(2 * R / H)])[0]
    exact = I / (np.pi * R ** 2 * H)
    n_mc = 400000
    got = orc.trap_grid(R, H, np.array([0.0]), np.array([0.0]), n_mc, 123)[0, 0]
    # MC standard error from the integrand's spread over the same samples
    s = np.array([orc.trap_sample(R, H, 123, k) for k in range(20000)])
    d = np.linalg.norm(s, axis=1)
    sig = np.std(np.exp(-d) / (d + 0.001)) / np.sqrt(n_mc)
    assert abs(got - exact) <= 5 * sig


# ---------------------------------------------------------------------------
# axial density profile (tools/concentration.cpp:32-101)
# ---------------------------------------------------------------------------
def _profile_raw(oracle_lib, z_nm, R, L, N_s, n):
    """Undo the interior normalisation to get the raw node deposits."""
    zg, c = oracle_lib.density_profile(z_nm, R, L, N_s, n)
    dz = L / (n - 1)
    return zg, c, c * (dz * np.pi * R * R / N_s)


def test_density_profile_single_ion_is_linear_interpolation(oracle_lib):
    """An interior ion splits unit weight linearly between its two nodes;
    nothing else is touched; nodes are -L/2 + i L/(n-1)."""
    R, L, N_s, n = 1.5e-3, 6e-3, 64.0, 100
    dz = L / (n - 1)
    for idx, frac in [(1, 0.25), (37, 0.5), (50, 0.9), (97, 0.1)]:
        z_m = -L / 2 + (idx + frac) * dz
        zg, c, raw = _profile_raw(oracle_lib, [z_m * 1e9], R, L, N_s, n)
        assert np.allclose(zg, -L / 2 + np.arange(n) * dz, rtol=0, atol=1e-18)
        assert abs(raw[idx] - (1 - frac)) < 1e-9 and abs(raw[idx + 1] - frac) < 1e-9
        others = np.delete(raw[1:-1], [idx - 1, idx])
        assert np.all(others == 0.0)
        # end nodes are the linear extrapolation of their neighbours
        assert c[0] == 2 * c[1] - c[2] and c[-1] == 2 * c[-2] - c[-3]


def test_density_profile_end_intervals_and_outside(oracle_lib):
    """First interval: only node 1 gets its linear share; last interval: only
    node n-2; ions below -L/2 or beyond the last node deposit nothing."""
    R, L, N_s, n = 1.0e-3, 4e-3, 10.0, 50
    dz = L / (n - 1)
    zg, c, raw = _profile_raw(oracle_lib, [(-L / 2 + 0.3 * dz) * 1e9], R, L, N_s, n)
    assert abs(raw[1] - 0.3) < 1e-9 and np.count_nonzero(raw[1:-1]) == 1
    zg, c, raw = _profile_raw(oracle_lib, [(L / 2 - 0.3 * dz) * 1e9], R, L, N_s, n)
    assert abs(raw[n - 2] - 0.3) < 1e-9 and np.count_nonzero(raw[1:-1]) == 1
    zg, c, raw = _profile_raw(oracle_lib, [(-L / 2 - 2 * dz) * 1e9, (L / 2 + 0.5 * dz) * 1e9], R, L, N_s, n)
    assert np.all(raw == 0.0) and np.all(c == 0.0)


def test_density_profile_uniform_column(oracle_lib):
    """Ions spread uniformly along z give the column density N N_s/(pi R^2 L)
    at every interior node (up to the two end intervals' dropped shares) and,
    extrapolated, at the ends."""
    R, L, N_s, n = 2e-3, 8e-3, 50.0, 40
    N = 400000
    z = (-L / 2 + (np.arange(N) + 0.5) * L / N) * 1e9
    zg, c = oracle_lib.density_profile(z, R, L, N_s, n)
    target = N * N_s / (np.pi * R * R * L)
    assert np.max(np.abs(c[2:-2] / target - 1)) < 1e-4
    assert abs(c[0] / target - 1) < 1e-3 and abs(c[-1] / target - 1) < 1e-3
