"""OpenDust's public Python API over the MI355X engine.

Same class names, constructor signatures, attribute names, printed parameter
reports and result arrays as the reference module (opendust.py:19-362), so a
script written for ``opendust.opendust`` runs unchanged after
``import opendust_amd; opendust_amd.install_alias()``.

``OpenDust.simulate`` keeps the reference's setup (initial state, trap fit,
reinjection tables; opendust.py:370-837) on the host and replaces the OpenMM
context and the per-step Python loop (opendust.py:1039-1384) with the HIP
engine (opendust_amd/csrc/od_engine.hip) driven through the C ABI.
"""

import ctypes
import secrets

import numpy as np
from scipy import special

from . import host
from .engine import (ABI_VERSION, RUN_GRAPH, Config, Engine, trap_potential_grid, write_csv, write_restart,
                     write_xyz)


class DustParticle:
    """A dust grain (opendust.py:25-40); SI units."""

    def __init__(self, x, y, z, r, chargeCalculationMethod, q=0.0):
        self.x = x
        self.y = y
        self.z = z
        self.chargeCalculationMethod = chargeCalculationMethod
        self.q = q
        self.r = r


class _PlasmaConstants:
    k_B = 1.380649e-23
    e = -1.60217662e-19
    eps_0 = 8.85418781762039e-12
    # the reference's value (the CODATA mass is 9.10938356e-31), kept for parity
    m_e = 9.0938356e-31

    def _derived(self):
        self.r_D_e = np.sqrt((self.eps_0 * self.k_B * self.T_e) / (self.e ** 2 * self.n_inf))
        self.v_B = np.sqrt(self.k_B * self.T_e / self.m_i)
        self.w_p_e = np.sqrt((self.n_inf * self.e ** 2) / (self.m_e * self.eps_0))
        self.w_p_i = np.sqrt((self.n_inf * self.e ** 2) / (self.m_i * self.eps_0))
        self.v_T = np.sqrt(self.k_B * self.T_n / self.m_i)


def _banner(title):
    print("#########################################")
    print(title)
    print("#########################################")
    print("\n")


class PlasmaParametersInSIUnitsMaxwell(_PlasmaConstants):
    """Plasma with a shifted-Maxwellian ion flow (opendust.py:49-103)."""

    def __init__(self, T_i, T_e, n_inf, M, m_i):
        self.T_n = T_i
        self.T_e = T_e
        self.n_inf = n_inf
        self.M = M
        self.m_i = m_i
        self.k_B = _PlasmaConstants.k_B
        self.e = _PlasmaConstants.e
        self.eps_0 = _PlasmaConstants.eps_0
        self.m_e = _PlasmaConstants.m_e
        self._derived()
        self.v_fl = self.M * self.v_B

    def printParameters(self):
        _banner("##### Plasma parameters in SI units #####")
        print("Ion temperature T_i = {} K".format(self.T_n))
        print("Electron temperature T_e = {} K".format(self.T_e))
        print("Unperturbed density of the plasma in the quasineutral plasma region = {} 1/m**3".format(self.n_inf))
        print("Mach number of the ion flow M = {}".format(self.M))
        print("Ion mass m_i = {} kg".format(self.m_i))
        print("Electron Debye radius r_D_e = {} m".format(self.r_D_e))
        print("Bohm velocity v_B = {} m/s".format(self.v_B))
        print("Electron plasma frequency w_p_e = {} radian/s".format(self.w_p_e))
        print("Ion plasma frequency w_p_i = {} radian/s".format(self.w_p_i))
        print("Thermal ion velocity v_T = {} m/s".format(self.v_T))
        print("Ion flow velocity v_fl = {} m/s".format(self.v_fl))
        print("\n")


class PlasmaParametersInSIUnitsFieldDriven(_PlasmaConstants):
    """Plasma with a field-driven, collisional ion flow (opendust.py:112-171)."""

    def __init__(self, T_n, T_e, n_inf, E, w_c, m_i):
        self.T_n = T_n
        self.T_e = T_e
        self.n_inf = n_inf
        self.E = E
        self.m_i = m_i
        self.k_B = _PlasmaConstants.k_B
        self.e = _PlasmaConstants.e
        self.eps_0 = _PlasmaConstants.eps_0
        self.m_e = _PlasmaConstants.m_e
        self._derived()
        self.w_c = w_c
        self.v_fl = self.E * abs(self.e) / self.m_i / self.w_c
        self.M = self.v_fl / self.v_B

    def printParameters(self):
        _banner("##### Plasma parameters in SI units #####")
        print("Neutral gas temperature T_n = {} K".format(self.T_n))
        print("Electron temperature T_e = {} K".format(self.T_e))
        print("Unperturbed density of the plasma in the quasineutral plasma region n_inf = {} 1/m**3".format(
            self.n_inf))
        print("Mach number of the ion flow M = {}".format(self.M))
        print("Ion mass m_i = {} kg".format(self.m_i))
        print("Electron Debye radius r_D_e = {} m".format(self.r_D_e))
        print("Bohm velocity v_B = {} m/s".format(self.v_B))
        print("Electron plasma frequency w_p_e = {} radian/s".format(self.w_p_e))
        print("Ion plasma frequency w_p_i = {} radian/s".format(self.w_p_i))
        print("Thermal neutral gas velocity v_T = {} m/s".format(self.v_T))
        print("Ion flow velocity v_fl = {} m/s".format(self.v_fl))
        print("Ion-neutral collision frequency w_c = {} 1/s".format(self.w_c))
        print("External electric field E = {} V/m".format(self.E))
        print("\n")


class SimulationParametersInSIUnits:
    """Domain, superion count and time step (opendust.py:181-215)."""

    def __init__(self, R, H, N, n, d_t, plasmaParametersInSIUnits):
        n_inf = plasmaParametersInSIUnits.n_inf
        r_D_e = plasmaParametersInSIUnits.r_D_e
        self.R = R
        self.H = H
        self.N = N
        self.n = n
        self.N_s = np.pi * R ** 2 * H * n_inf / self.N
        self.d_t = d_t
        self.r_0 = r_D_e * 1.0e-3

    def printParameters(self):
        _banner("### Simulation parameters in SI units ###")
        print("Radius of the cylindrical computational domain R = {} m".format(self.R))
        print("Height of the cylindrical computational domain H = {} m".format(self.H))
        print("Number of superions N = {}".format(self.N))
        print("Number of ions in one superion N_s = {}".format(self.N_s))
        print("Integration time step d_t = {} s".format(self.d_t))
        print("Number of time steps n = {}".format(self.n))
        print("\n")


# the benchmark scripts of the reference import this spelling
# (performance/NVIDIA-V100/benchmark.py:3)
SimulatioParametersInSIUnits = SimulationParametersInSIUnits


class OutputParameters:
    """Console/file output cadence and file names (opendust.py:225-245)."""

    def __init__(self, nOutput=10, nFileOutput=10, csvOutputFileName="", xyzOutputFileName="",
                 restartFileName=""):
        self.nOutput = nOutput
        self.nFileOutput = nFileOutput
        self.csvOutputFileName = csvOutputFileName
        self.xyzOutputFileName = xyzOutputFileName
        self.restartFileName = restartFileName


class _DustState:
    """Per-dust state and per-step result arrays (opendust.py:319-338); x, y,
    z, r in nm, charges in C, forces in N."""

    def __init__(self, dustParticle, n, distributionType):
        self.x = float(dustParticle.x * host.M_TO_NM)
        self.y = float(dustParticle.y * host.M_TO_NM)
        self.z = float(dustParticle.z * host.M_TO_NM)
        self.r = float(dustParticle.r * host.M_TO_NM)
        self.q = np.full(n, float(dustParticle.q))
        self.chargeCalculationMethod = dustParticle.chargeCalculationMethod
        self.forceIonsCollect = np.zeros((n, 3))
        self.forceIonsOrbit = np.zeros((n, 3))
        self.forceDust = np.zeros((n, 3))
        if distributionType == "fieldDriven":
            self.forceExternalField = np.zeros((n, 3))


class OpenDust:
    """Main class (opendust.py:254-1384).

    Extension keywords (all optional, default behaviour = reference):
      precision  "fp32c" (FP32 pair arithmetic, compensated sums) or "fp64"
      seed       Philox key of the in-step randomness (None: fresh random key)
      useGraph   replay steps as captured HIP graphs
      trapSeed, trapSamples   Monte-Carlo seed / sample count of the trap fit
      cutoffDistance, switchingDistance   (m) truncation used when
                 simulate(cutOff=True); default 3.2 and 3 r_D_e as the reference
                 (opendust.py:481-485); switchingDistance=None with a given
                 cutoffDistance means a hard cutoff
    """

    def __init__(self, plasmaParametersInSIUnits, simulationParametersInSIUnits, outputParameters, dustParticles,
                 distributionType, precision="fp32c", seed=None, useGraph=True, trapSeed=None, trapSamples=None,
                 cutoffDistance=None, switchingDistance=None):
        self.plasmaParametersInSIUnits = plasmaParametersInSIUnits
        self.simulationParametersInSIUnits = simulationParametersInSIUnits
        self.outputParameters = outputParameters

        sp, pp = simulationParametersInSIUnits, plasmaParametersInSIUnits
        self.R = float(sp.R * host.M_TO_NM)
        self.H = float(sp.H * host.M_TO_NM)
        self.d_t = float(sp.d_t * host.S_TO_PS)
        self.r_0 = float(sp.r_0 * host.M_TO_NM)
        self.N = sp.N
        self.N_s = sp.N_s
        self.n = sp.n

        self.k = 1.0 / (pp.eps_0 * 4.0 * np.pi)
        self.eps_0 = pp.eps_0
        self.e = pp.e

        self.n_inf = float(pp.n_inf * 1e-27)
        self.r_D_e = float(pp.r_D_e * host.M_TO_NM)
        self.v_fl = float(pp.v_fl * 1e-3)
        self.v_T = float(pp.v_T * 1e-3)
        self.q_s = abs(pp.e) * self.N_s
        self.m_s = float(pp.m_i * self.N_s / host.AMU_KG)

        self.distributionType = distributionType
        if self.distributionType == "fieldDriven":
            self.E = pp.E
            self.w_c = float(pp.w_c / host.S_TO_PS)

        self.interactionConversionCoefficient = host.J_TO_KJ_PER_MOL * host.M_TO_NM
        self.externalFieldConversionCoefficient = host.J_TO_KJ_PER_MOL * host.NM_TO_M

        self.t = np.linspace(0, self.n * sp.d_t, self.n)

        self.numberOfDustParticles = len(dustParticles)
        if self.numberOfDustParticles > host.MAX_DUST:
            raise ValueError("at most %d dust particles are supported" % host.MAX_DUST)
        self.dustParticles = [_DustState(d, self.n, self.distributionType) for d in dustParticles]

        self.precision = precision
        self.seed = seed
        self.useGraph = useGraph
        self.trapSeed = trapSeed
        self.trapSamples = trapSamples
        self.cutoffDistance = cutoffDistance
        self.switchingDistance = switchingDistance
        self._cutoff_nm = 0.0
        self._switch_nm = 0.0
        self.alphaTrap = None
        self.p_t = self.p_s = self.p_b = None

    # ------------------------------------------------------------------ setup
    def initial_state(self, toRestartFileName=""):
        """Initial ion positions/velocities in nm, nm/ps (opendust.py:370-466)."""
        if toRestartFileName == "":
            dusts = [(d.x, d.y, d.z, d.r) for d in self.dustParticles]
            x = host.initial_positions(self.R, self.H, dusts, self.N)
            v = host.initial_velocities(self.distributionType, self.v_fl, self.v_T, self.N)
        else:
            x, v = host.read_restart(toRestartFileName, self.N)
        return x, v

    def reinjection_tables(self):
        """Face tables and probabilities (opendust.py:606-837)."""
        F_top, top = host.face_table(self.v_fl, self.v_T, "top", self.distributionType)
        F_bot, bottom = host.face_table(-self.v_fl, self.v_T, "bottom", self.distributionType)
        F_side, side = host.face_table(0, self.v_T, "side", self.distributionType)
        vz = host.vz_table(self.v_fl, self.v_T) if self.distributionType == "fieldDriven" else None
        self.p_t, self.p_s, self.p_b = host.face_probabilities(self.R, self.H, F_top, F_bot, F_side)
        return top, bottom, side, vz

    def fit_trap(self, device=0):
        """Background-potential fit (opendust.py:499-550) with the GPU Monte Carlo."""
        Rl = self.R / self.r_D_e
        Hl = self.H / self.r_D_e
        n_r, n_h, n_mc = host.trap_grid_shape(Rl, Hl)
        if self.trapSamples is not None:
            n_mc = int(self.trapSamples)
        seed = self.trapSeed if self.trapSeed is not None else secrets.randbits(64)
        rho = np.linspace(0, Rl, n_r)
        h = np.linspace(0, Hl / 2.0, n_h)
        pot = trap_potential_grid(Rl, Hl, rho, h, n_mc, seed, device)
        return host.trap_fit(pot, Rl, Hl, self.N)

    def set_cutoff(self, cutOff):
        """Truncation of the ion-ion force (nm) for a run; cutOff=False: all pairs."""
        if not cutOff:
            self._cutoff_nm = self._switch_nm = 0.0
        elif self.cutoffDistance is None:
            self._cutoff_nm = 3.2 * self.r_D_e
            self._switch_nm = 3.0 * self.r_D_e
        else:
            self._cutoff_nm = float(self.cutoffDistance * host.M_TO_NM)
            self._switch_nm = (float(self.switchingDistance * host.M_TO_NM)
                               if self.switchingDistance is not None else 0.0)

    def make_config(self, tables, alphaTrap, device=0, precision=None, seed=None):
        """Fill the engine configuration (include/opendust_amd.h od_config)."""
        pp, sp = self.plasmaParametersInSIUnits, self.simulationParametersInSIUnits
        top, bottom, side, vz = tables
        prec = precision or self.precision
        c = Config()
        c.abi_version = ABI_VERSION
        c.struct_size = ctypes.sizeof(Config)
        c.n_ions = int(self.N)
        c.n_dust = self.numberOfDustParticles
        c.distribution = host.DIST_FIELD_DRIVEN if self.distributionType == "fieldDriven" else host.DIST_MAXWELLIAN
        if prec == "fp32c":
            c.precision = host.PREC_FP32C
        elif prec == "fp64":
            c.precision = host.PREC_FP64
        else:
            raise ValueError("precision must be 'fp32c' or 'fp64'")
        c.device = int(device)
        c.use_bottom = 1 if self.p_b != 0 else 0
        c.j_split = 0
        c.pair_mode = 0
        c.sym_chunk = 0
        c.seed = int(seed) & (2 ** 64 - 1)
        c.R, c.H, c.dt = self.R, self.H, self.d_t
        c.r_D_e, c.r_0 = self.r_D_e, self.r_0
        c.m_s = self.m_s
        c.v_T, c.v_fl = self.v_T, self.v_fl
        kconv = self.k * self.interactionConversionCoefficient
        c.kqq = kconv * self.q_s * self.q_s
        c.k_dust = kconv * self.q_s
        c.trap_C = self.q_s * self.q_s / pp.r_D_e * self.k * host.J_TO_KJ_PER_MOL
        for m in range(15):
            c.trap_alpha[m] = float(alphaTrap[m])
        c.cutoff = self._cutoff_nm
        c.switch_dist = self._switch_nm
        if self.distributionType == "fieldDriven":
            c.field_force = self.E * self.externalFieldConversionCoefficient * self.q_s
            c.thr_coll = host.u32_threshold(self.d_t * self.w_c)
        c.thr_top = host.u32_threshold(self.p_t)
        c.thr_side = host.u32_threshold(self.p_t + self.p_s)
        for p, d in enumerate(self.dustParticles):
            c.dust_x[p], c.dust_y[p], c.dust_z[p], c.dust_r[p] = d.x, d.y, d.z, d.r
            c.dust_q0[p] = float(d.q[0])
            c.dust_method[p] = host.charge_method_code(d.chargeCalculationMethod)
            c.si_dust_x[p] = d.x * host.NM_TO_M
            c.si_dust_y[p] = d.y * host.NM_TO_M
            c.si_dust_z[p] = d.z * host.NM_TO_M
            c.si_dust_r[p] = d.r * host.NM_TO_M
        c.si_k = self.k
        c.si_e = self.e
        c.si_q_s = self.q_s
        c.si_r_D_e = pp.r_D_e
        c.si_dt = sp.d_t
        c.si_kB = pp.k_B
        c.si_Te = pp.T_e
        c.si_Tn = pp.T_n
        c.si_E = getattr(pp, "E", 0.0) if self.distributionType == "fieldDriven" else 0.0
        c.gamma_e_pref = pp.n_inf * (pp.T_e * pp.k_B / np.pi / pp.m_e / 2.0) ** (1.0 / 2.0)
        u = pp.v_fl / pp.v_T / np.sqrt(2.0)
        v_ti = pp.v_T * np.sqrt(2.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            c.oml_A = pp.n_inf * v_ti * u / 4.0
            c.oml_B = 1.0 + 1.0 / (2.0 * u * 2) if u != 0 else float("inf")
            c.oml_u2 = u ** 2
            c.oml_erf_u = float(special.erf(u))
            c.oml_tail = 1.0 / u / np.sqrt(np.pi) * np.exp(-(u ** 2)) if u != 0 else float("inf")
        c.collect_scale = host.NM_PER_PS_TO_M_PER_S * pp.m_i * self.N_s / sp.d_t
        c.top.load(top)
        c.bottom.load(bottom)
        c.side.load(side)
        c.vz.load(vz)
        return c

    # ------------------------------------------------------------ simulation
    def simulate(self, deviceIndex="0", cutOff=False, toRestartFileName=""):
        """Run the simulation (opendust.py:362-1384)."""
        self.set_cutoff(cutOff)
        device = 0
        dev_str = str(deviceIndex).strip()
        if dev_str:
            device = int(dev_str.split(",")[0])

        x0, v0 = self.initial_state(toRestartFileName)
        tables = self.reinjection_tables()
        self.alphaTrap = self.fit_trap(device)
        if self.seed is None:
            self.seed = secrets.randbits(64)
        cfg = self.make_config(tables, self.alphaTrap, device=device, seed=self.seed)
        engine = Engine(cfg)
        try:
            engine.set_ions(x0, v0)
            self._run_loop(engine)
        finally:
            engine.close()

    def _record(self, i0, rec):
        """Copy engine records of steps i0.. into the result arrays."""
        k = rec.shape[0]
        for p, d in enumerate(self.dustParticles):
            r = rec[:, p]
            d.q[i0:i0 + k] = r["q"]
            d.forceIonsOrbit[i0:i0 + k] = r["force_orbit"]
            d.forceIonsCollect[i0:i0 + k] = r["force_collect"]
            d.forceDust[i0:i0 + k] = r["force_dust"]
            if self.distributionType == "fieldDriven":
                d.forceExternalField[i0:i0 + k] = r["force_field"]

    def _run_loop(self, engine):
        out = self.outputParameters
        nOutput, nFileOutput = out.nOutput, out.nFileOutput
        if out.xyzOutputFileName != "":
            open(out.xyzOutputFileName, "w").close()
        flags = RUN_GRAPH if self.useGraph else 0
        positions = None
        i = 0
        while i < self.n:
            # run up to and including the next step that writes files
            j = i
            while j < self.n - 1 and j % nFileOutput != 0:
                j += 1
            rec = engine.run(j - i + 1, flags)
            self._record(i, rec)
            for s in range(i, j + 1):
                if s % nOutput == 0:
                    if s == 0:
                        _banner("###        Start simulation           ###")
                    print("Step: {} out of {}".format(s, self.n))
            if j % nFileOutput == 0:
                positions, _ = engine.get_ions()
                self._write_frame(j, positions)
            if j == self.n - 1 and out.restartFileName != "":
                _, vel = engine.get_ions()
                self._write_restart(positions, vel)
            i = j + 1

    def _write_frame(self, i, pos):
        """XYZ and CSV frames (opendust.py:1080-1099) through the native writer."""
        out = self.outputParameters
        idx = int(i / out.nFileOutput)
        if out.xyzOutputFileName != "":
            write_xyz(out.xyzOutputFileName, pos, idx)
        if out.csvOutputFileName != "":
            write_csv(out.csvOutputFileName + "{}.csv".format(idx), pos,
                      self.plasmaParametersInSIUnits.n_inf * 1e-27, self.N_s)

    def _write_restart(self, pos, vel):
        """RESTART at the last step.  As in the reference (opendust.py:1100-1115)
        the positions are the ones of the last file-output frame, not of the
        last step; the velocities are current."""
        write_restart(self.outputParameters.restartFileName, pos, vel)
