"""Resident plasma solver for dust-dynamics loops.

The reference's dust-dynamics drivers advance a few dust grains with a
Langevin integrator and, at every dust step, solve the plasma around them with
a full ``OpenDust(...).simulate()`` call
(examples/DustDynamics/TwoParticles/launch.py:113-220): a new OpenMM context,
ions re-read from the RESTART file, reinjection tables rebuilt, the trap
refitted by Monte Carlo.  ``ResidentPlasma`` keeps one engine context for the
whole loop: each ``solve()`` moves the dusts inside it (``od_update_config``)
and continues the ion state on the device, so a dust step costs only its
plasma steps.

Randomness: solve k uses Philox key ``seed + k * 0x9E3779B97F4A7C15`` (mod
2^64) and restarts the step counter at 0, so every solve draws fresh numbers
and applies the "i > 0" charge rule (opendust.py:1121) as a new simulate()
would.  The ions continue from their current positions and velocities (the
reference's RESTART hand-over writes the positions of the last file-output
frame, opendust.py:1103-1108; a resident context has no such lag).

``dust_langevin_step`` restates the example's dust update (section 7 of
launch.py:186-211) for callers that want the whole loop here.
"""

import secrets

import numpy as np

from .engine import RUN_GRAPH, Engine
from .opendust import OpenDust

_GOLDEN = 0x9E3779B97F4A7C15


class ResidentPlasma:
    """One engine context serving many plasma solves with moving dusts.

    openDust: an OpenDust object that defines the plasma, the cylinder, the
    superion count and the initial dusts (its number of dusts is fixed for the
    lifetime of the solver)."""

    def __init__(self, openDust, deviceIndex="0", toRestartFileName="", alphaTrap=None):
        self.template = openDust
        dev_str = str(deviceIndex).strip()
        self.device = int(dev_str.split(",")[0]) if dev_str else 0
        x0, v0 = openDust.initial_state(toRestartFileName)
        self.tables = openDust.reinjection_tables()
        self.alphaTrap = (np.asarray(alphaTrap, dtype=np.float64) if alphaTrap is not None
                          else openDust.fit_trap(self.device))
        openDust.alphaTrap = self.alphaTrap
        if openDust.seed is None:
            openDust.seed = secrets.randbits(64)
        self.seed = int(openDust.seed)
        self.n_solves = 0
        cfg = openDust.make_config(self.tables, self.alphaTrap, device=self.device, seed=self.seed)
        self.engine = Engine(cfg)
        self.engine.set_ions(x0, v0)

    def close(self):
        self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def solve(self, dustParticles, n=None):
        """Run n plasma steps (default: the template's n) around `dustParticles`
        (DustParticle objects, SI, as many as the template has).  Returns an
        OpenDust object whose dustParticles[p].{q, forceIonsOrbit,
        forceIonsCollect, forceDust[, forceExternalField]} hold the per-step
        results, as after simulate()."""
        t = self.template
        if len(dustParticles) != t.numberOfDustParticles:
            raise ValueError("the number of dust particles is fixed at %d" % t.numberOfDustParticles)
        sp = t.simulationParametersInSIUnits
        if n is not None and int(n) != sp.n:
            sp = type(sp)(sp.R, sp.H, sp.N, int(n), sp.d_t, t.plasmaParametersInSIUnits)
        od = OpenDust(t.plasmaParametersInSIUnits, sp, t.outputParameters, dustParticles, t.distributionType,
                      precision=t.precision, useGraph=t.useGraph)
        od.p_t, od.p_s, od.p_b = t.p_t, t.p_s, t.p_b
        od.alphaTrap = self.alphaTrap
        od.seed = (self.seed + self.n_solves * _GOLDEN) & (2 ** 64 - 1)
        cfg = od.make_config(self.tables, self.alphaTrap, device=self.device, seed=od.seed)
        self.engine.update_config(cfg)
        self.engine.set_dust_charges([d.q[0] for d in od.dustParticles])
        self.engine.set_step(0)
        flags = RUN_GRAPH if t.useGraph else 0
        done = 0
        while done < od.n:
            k = min(4096, od.n - done)
            od._record(done, self.engine.run(k, flags))
            done += k
        self.n_solves += 1
        return od


def dust_langevin_step(x, v, f, m_p, r_p, P, v_T, T_p, d_t_p, g=9.8, rng=np.random):
    """One dust step of the reference's driver (launch.py:186-211): Epstein
    drag m_p gamma_p v with gamma_p = 8/3 sqrt(2 pi) r_p^2 P / v_T / m_p, a
    thermal kick of variance 2 k_B T_p m_p gamma_p / d_t_p per component
    (NumPy normal draws x, y, z in that order), gravity g along +z, and
    x' = x + v d_t_p + a d_t_p^2 / 2, v' = v + a d_t_p.
    x, v, f: (3,) arrays (m, m/s, N); returns (x', v')."""
    k_B = 1.380649e-23
    gamma_p = 8.0 / 3.0 * np.sqrt(2.0 * np.pi) * r_p ** 2 * P / v_T / m_p
    s_p = np.sqrt(2 * k_B * T_p * m_p * gamma_p / d_t_p)
    f_therm = np.array([rng.normal(0, s_p, 1)[0] - m_p * gamma_p * v[c] for c in range(3)])
    a = (f_therm + np.asarray(f, dtype=np.float64)) / m_p
    a[2] += g
    x_new = np.asarray(x, dtype=np.float64) + np.asarray(v, dtype=np.float64) * d_t_p + 0.5 * a * d_t_p ** 2
    v_new = np.asarray(v, dtype=np.float64) + a * d_t_p
    return x_new, v_new
