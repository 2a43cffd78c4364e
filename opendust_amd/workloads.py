"""Named ion-flow workloads (the benchmark and parity configurations) built
through the public API, plus a helper that prepares an engine configuration
and initial state without running ``simulate()``.

  single_dust     Maxwellian Ar+ flow, M = 1, T_e = 2.5 eV, T_i = 0.025 eV,
                  n = 3.57e14 m^-3, cylinder R = 5, H = 10 r_D_e, one dust
                  (r = 0.05 r_D_e, q0 = -2e4 e, consistent charge) at the
                  origin, dt = 0.01 / w_pi.  N = 16384 (small) or 2^20 (large).
  dust_chain      field-driven flow with charge-exchange collisions,
                  w_c = 0.1 w_pi, E such that M = 0.8, three dusts on the axis
                  at z = -1, 0, +1 r_D_e, R = 6, H = 14 r_D_e, N = 2^20.
  cutoff_large    single dust, N = 2^22, R = 10, H = 20 r_D_e, ion-ion force
                  truncated at 2 r_D_e through the cell list.
  sweep           64 independent single-flow simulations: N log-uniform in
                  [2^12, 2^17], M uniform in [0.5, 2], 1-4 dusts on the axis
                  with spacing uniform in [0.5, 2] r_D_e (centred), q0 uniform
                  in [-3e4, -1e4] e, the rest as single_dust; 5000 steps.
  piel_fig1       the tutorial / Piel 2017 Fig. 1 setting (n = 1e14 m^-3,
                  R = 3, H = 6 r_D_e, r = 58.8 um at z = -r_D_e, N = 2^16).
"""

import numpy as np

from . import host
from .opendust import (DustParticle, OpenDust, OutputParameters, PlasmaParametersInSIUnitsFieldDriven,
                       PlasmaParametersInSIUnitsMaxwell, SimulationParametersInSIUnits)

M_AR = 6.63e-26
T_E = 29011.3
T_I = 290.113
N_INF = 3.57e14


def single_dust(N=16384, n=2000, precision="fp32c", seed=1, charge="consistent", mach=1.0, R=5.0, H=10.0, **kw):
    """The single-dust flow; ``mach``, ``R`` and ``H`` (in r_D_e) vary the
    drift and the cylinder for tests (at M < 0.36 the bottom face reinjects)."""
    pp = PlasmaParametersInSIUnitsMaxwell(T_I, T_E, N_INF, mach, M_AR)
    lam = pp.r_D_e
    sp = SimulationParametersInSIUnits(R * lam, H * lam, int(N), int(n), 0.01 / pp.w_p_i, pp)
    dust = [DustParticle(0.0, 0.0, 0.0, 0.05 * lam, charge, -2e4 * abs(pp.e))]
    return OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), dust, "Maxwellian", precision=precision, seed=seed,
                    **kw)


def cutoff_large(N=2 ** 22, n=10000, precision="fp32c", seed=1, cutoff=2.0, switch=None, **kw):
    """Single dust as single_dust, N = 2^22 in a cylinder R = 10, H = 20 r_D_e,
    ion-ion force truncated at `cutoff` r_D_e (hard, or switched from `switch`)."""
    pp = PlasmaParametersInSIUnitsMaxwell(T_I, T_E, N_INF, 1.0, M_AR)
    lam = pp.r_D_e
    sp = SimulationParametersInSIUnits(10 * lam, 20 * lam, int(N), int(n), 0.01 / pp.w_p_i, pp)
    dust = [DustParticle(0.0, 0.0, 0.0, 0.05 * lam, "consistent", -2e4 * abs(pp.e))]
    od = OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), dust, "Maxwellian", precision=precision, seed=seed,
                  cutoffDistance=cutoff * lam, switchingDistance=None if switch is None else switch * lam, **kw)
    od.set_cutoff(True)
    return od


def dust_chain(N=2 ** 20, n=20000, precision="fp32c", seed=1, **kw):
    m_i = M_AR
    tmp = PlasmaParametersInSIUnitsMaxwell(T_I, T_E, N_INF, 1.0, m_i)
    w_c = 0.1 * tmp.w_p_i
    E = 0.8 * tmp.v_B * m_i * w_c / abs(tmp.e)
    pp = PlasmaParametersInSIUnitsFieldDriven(T_I, T_E, N_INF, E, w_c, m_i)
    lam = pp.r_D_e
    sp = SimulationParametersInSIUnits(6 * lam, 14 * lam, int(N), int(n), 0.01 / pp.w_p_i, pp)
    dust = [DustParticle(0.0, 0.0, z * lam, 0.05 * lam, "consistent", -2e4 * abs(pp.e)) for z in (-1.0, 0.0, 1.0)]
    return OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), dust, "fieldDriven", precision=precision, seed=seed,
                    **kw)


def sweep(n_sims=64, n=5000, precision="fp32c", seed=7, log2_n=(12.0, 17.0), **kw):
    """OpenDust objects of the parameter sweep (see the module docstring); the
    parameters are drawn from NumPy RandomState(seed), simulation s gets
    Philox key seed + s."""
    rs = np.random.RandomState(seed)
    ods = []
    for s in range(n_sims):
        N = int(round(2.0 ** rs.uniform(*log2_n)))
        mach = rs.uniform(0.5, 2.0)
        nd = int(rs.randint(1, 5))
        spacing = rs.uniform(0.5, 2.0)
        q0 = rs.uniform(-3e4, -1e4)
        pp = PlasmaParametersInSIUnitsMaxwell(T_I, T_E, N_INF, mach, M_AR)
        lam = pp.r_D_e
        sp = SimulationParametersInSIUnits(5 * lam, 10 * lam, N, int(n), 0.01 / pp.w_p_i, pp)
        zs = [(k - (nd - 1) / 2.0) * spacing * lam for k in range(nd)]
        dust = [DustParticle(0.0, 0.0, z, 0.05 * lam, "consistent", q0 * abs(pp.e)) for z in zs]
        ods.append(OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), dust, "Maxwellian", precision=precision,
                            seed=seed + s, **kw))
    return ods


def piel_fig1(N=2 ** 16, n=3000, charge="consistent", q_e=0.0, precision="fp32c", seed=1, **kw):
    pp = PlasmaParametersInSIUnitsMaxwell(290.11, 29011, 1e14, 1, 6.6335209e-26)
    lam = pp.r_D_e
    sp = SimulationParametersInSIUnits(3 * lam, 6 * lam, int(N), int(n), 3.5148240854e-09, pp)
    dust = [DustParticle(0, 0, -1 * lam, 58.8e-6, charge, q_e * pp.e)]
    return OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), dust, "Maxwellian", precision=precision, seed=seed,
                    **kw)


def synthetic_alpha(od):
    """A smooth, physically shaped trap fit computed on the CPU without the
    GPU Monte Carlo: the exact background potential of a uniform cylinder is
    replaced by a coarse deterministic quadrature, then fitted as the
    reference does.  Used where a test needs a trap without a GPU."""
    Rl = od.R / od.r_D_e
    Hl = od.H / od.r_D_e
    n_r, n_h = max(int(Rl * 2), 4), max(int(Hl), 4)
    rho = np.linspace(0, Rl, n_r)
    hh = np.linspace(0, Hl / 2.0, n_h)
    # midpoint quadrature over the cylinder (cell volume-weighted)
    nr, nt, nz = 24, 24, 48
    rr = (np.arange(nr) + 0.5) / nr * Rl
    tt = (np.arange(nt) + 0.5) / nt * 2 * np.pi
    zz = -Hl / 2 + (np.arange(nz) + 0.5) / nz * Hl
    R3, T3, Z3 = np.meshgrid(rr, tt, zz, indexing="ij")
    w = R3 * (Rl / nr) * (2 * np.pi / nt) * (Hl / nz)
    X, Y = R3 * np.cos(T3), R3 * np.sin(T3)
    vol = np.pi * Rl ** 2 * Hl
    pot = np.empty((n_r, n_h))
    for a, gr in enumerate(rho):
        for b, gh in enumerate(hh):
            d = np.sqrt((gr - X) ** 2 + Y ** 2 + (gh - Z3) ** 2)
            pot[a, b] = np.sum(w * np.exp(-d) / (d + 0.001)) / vol
    return host.trap_fit(pot, Rl, Hl, od.N)


def prepare(od, alpha=None, device=0, np_seed=None, precision=None):
    """Initial state, tables and engine configuration of ``od`` without
    running it.  ``alpha``: trap coefficients (None: GPU Monte-Carlo fit).
    ``np_seed`` seeds NumPy's global stream before the initial state, as a
    user of the reference would with np.random.seed."""
    if np_seed is not None:
        np.random.seed(np_seed)
    x0, v0 = od.initial_state()
    tables = od.reinjection_tables()
    if alpha is None:
        alpha = od.fit_trap(device)
    od.alphaTrap = np.asarray(alpha, dtype=np.float64)
    cfg = od.make_config(tables, od.alphaTrap, device=device, precision=precision,
                         seed=od.seed if od.seed is not None else 1)
    return cfg, x0, v0
