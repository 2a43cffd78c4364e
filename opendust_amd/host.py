"""Host-side setup of the ion-flow step: units, reinjection tables, initial ion
state, trap fit and the engine configuration.

Everything here runs once per ``simulate()`` call on the host, as in the
reference (opendust.py:370-837).  The per-step work runs in the HIP engine.

Unit handling follows the reference's conversion of SI inputs to OpenMM units
(nm, ps, amu, kJ/mol; opendust.py:274-312) with explicit constants instead of
unyt.  The constants are CODATA 2018 (N_A exact); unyt's own constant tables
differ by at most ~1e-9 relative.
"""

import math

import numpy as np
from scipy import integrate, special
from scipy.interpolate import interp1d

# --- unit conversion (replaces unyt at opendust.py:4-11, :275-312) -------------
N_A = 6.02214076e23
AMU_KG = 1.66053906660e-27
M_TO_NM = 1e9
NM_TO_M = 1e-9
S_TO_PS = 1e12
J_TO_KJ_PER_MOL = N_A / 1000.0
NM_PER_PS_TO_M_PER_S = 1e-9 / 1e-12

# engine enumerations (mirror include/opendust_amd.h)
DIST_MAXWELLIAN = 0
DIST_FIELD_DRIVEN = 1
PREC_FP32C = 0
PREC_FP64 = 1
CHARGE_GIVEN = 0
CHARGE_CONSISTENT = 1
CHARGE_OML = 2
MAX_DUST = 8
MAX_TABLE = 512


def charge_method_code(name):
    """opendust.py:1120, :1153 -- "consistent", "oml", anything else keeps q."""
    if name == "consistent":
        return CHARGE_CONSISTENT
    if name == "oml":
        return CHARGE_OML
    return CHARGE_GIVEN


# --- reinjection tables (opendust.py:606-837) ----------------------------------

def _flux_maxwell(v_fl, v_T):
    """Flux-weighted shifted Maxwellian v*N(v; v_fl, v_T) (opendust.py:654-659)."""
    norm = np.sqrt(2.0 * np.pi * v_T ** 2)

    def f(v):
        return v / norm * np.exp(-((v - v_fl) ** 2) / (2.0 * v_T ** 2))

    return f


def _flux_field_driven(v_fl, v_T):
    """Field-driven (collisional) flux distribution (opendust.py:673-689, :738-754)."""
    a = v_fl / v_T

    def f(v):
        return (
            v / 2.0 / abs(v_fl)
            * np.exp(-(2.0 * (v / v_T) * a - 1) / (2.0 * a ** 2))
            * (special.erf(((v / v_T) * a - 1.0) / (np.sqrt(2.0) * abs(a))) + 1.0)
        )

    return f


def _density_field_driven(v_fl, v_T):
    """Field-driven velocity density used for v_z (opendust.py:415-431, :738-754)."""
    a = v_fl / v_T

    def f(v):
        return (
            1 / 2.0 / abs(v_fl)
            * np.exp(-(2.0 * (v / v_T) * a - 1) / (2.0 * a ** 2))
            * (special.erf(((v / v_T) * a - 1.0) / (np.sqrt(2.0) * abs(a))) + 1.0)
        )

    return f


class InverseCDFTable:
    """Piecewise-linear inverse CDF as the reference prints it into the OpenMM
    expression: every constant is the ``%.8f``-rounded value
    (opendust.py:702-723).  ``grid_v``/``grid_xi`` keep the unrounded knots."""

    def __init__(self, v, xi, F_max):
        self.grid_v = np.asarray(v, dtype=np.float64)
        self.grid_xi = np.asarray(xi, dtype=np.float64)
        self.F_max = float(F_max)
        n = len(v)
        rnd = lambda a: float("{:.8f}".format(a))
        self.v0, self.xi0, self.slope, self.center, self.half = [], [], [], [], []
        for i in range(1, n):
            dxi = xi[i] - xi[i - 1]
            if dxi == 0.0:
                raise ValueError("degenerate reinjection table: equal CDF knots")
            self.v0.append(rnd(v[i - 1]))
            self.xi0.append(rnd(xi[i - 1]))
            self.slope.append(rnd((v[i] - v[i - 1]) / dxi))
            self.center.append(rnd(0.5 * (xi[i] + xi[i - 1])))
            self.half.append(rnd(0.5 * (xi[i] - xi[i - 1])))
        if len(self.v0) > MAX_TABLE:
            raise ValueError("reinjection table longer than %d segments" % MAX_TABLE)

    @property
    def n_seg(self):
        return len(self.v0)

    def __call__(self, u):
        """Evaluate V(u) the way the expression does (sum of masked segments)."""
        u = np.asarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        for k in range(self.n_seg):
            sel = (self.half[k] - np.abs(u - self.center[k])) >= 0.0
            out = out + (self.v0[k] + (u - self.xi0[k]) * self.slope[k]) * sel
        return out


def _cdf_table(f, v_min, v_max, v):
    """xi[i] = F(v[i]) / F(v_max), F(v) = quad(f, v_min, v) (opendust.py:691-697)."""
    F = lambda b: integrate.quad(f, v_min, b)[0]
    F_max = F(v_max)
    xi = np.zeros(len(v))
    for i in range(len(v)):
        xi[i] = F(v[i]) / F_max
    return F_max, xi


def face_table(v_fl, v_T, side, distribution):
    """One face's flux table, following opendust.py:608-724.  Returns
    (F_max, InverseCDFTable or None).  ``v_fl`` is the signed flow speed the
    reference passes for that face (top: v_fl, bottom: -v_fl, side: 0)."""
    if distribution == "Maxwellian":
        if side == "top":
            v_max = v_fl + 4 * v_T
            v_min = max(v_fl - 4 * v_T, 0)
        elif side == "bottom":
            v_max = max(v_fl + 4 * v_T, 0)
            v_min = 0
        else:
            v_max = 4 * v_T
            v_min = 0
    elif distribution == "fieldDriven":
        if side == "top":
            v_max = abs(v_fl) * 6 + 5 * v_T
            v_min = 0
        else:
            v_max = 4 * v_T
            v_min = 0
    else:
        raise ValueError("unknown distributionType %r" % (distribution,))

    if v_max < 0.4 * v_T:
        return 0, None

    if distribution == "fieldDriven" and side == "top":
        n1 = int(4 * v_T / (0.2 * v_T))
        n2 = int((v_max - 5 * v_T) / v_T)
        v = np.concatenate((np.linspace(0, 4 * v_T, n1), np.linspace(5 * v_T, v_max, n2)))
    else:
        n_v = int((v_max - v_min) / (0.2 * v_T))
        v = np.linspace(v_min, v_max, n_v)

    if distribution == "Maxwellian" or side == "side":
        f = _flux_maxwell(v_fl, v_T)
    else:
        f = _flux_field_driven(v_fl, v_T)

    F_max, xi = _cdf_table(f, v_min, v_max, v)
    if side == "bottom":
        v = -v
    return F_max, InverseCDFTable(v, xi, F_max)


def vz_table(v_fl, v_T):
    """v_z table of side reinjection in the fieldDriven case (opendust.py:726-787)."""
    v_max = abs(v_fl) * 6 + 5 * v_T
    v_min = -4 * v_T
    n1 = int(8 * v_T / (0.2 * v_T))
    n2 = int((v_max - 5 * v_T) / v_T)
    v = np.concatenate((np.linspace(v_min, 4 * v_T, n1), np.linspace(5 * v_T, v_max, n2)))
    F_max, xi = _cdf_table(_density_field_driven(v_fl, v_T), v_min, v_max, v)
    return InverseCDFTable(v, xi, F_max)


def face_probabilities(R, H, F_top, F_bottom, F_side):
    """p_t, p_s, p_b (opendust.py:806-837), R and H in nm."""
    denom = np.pi * R ** 2 * F_top + np.pi * R ** 2 * F_bottom + 2 * np.pi * R * H * F_side
    p_t = np.pi * R ** 2 * F_top / denom
    p_s = 2 * np.pi * R * H * F_side / denom
    p_b = np.pi * R ** 2 * F_bottom / denom
    return float(p_t), float(p_s), float(p_b)


def u32_threshold(p):
    """Decision threshold on a uniform u32 draw: event iff u32 < thr, so that
    P(event) = thr / 2**32 = p rounded down to the 2**-32 grid."""
    p = min(max(float(p), 0.0), 1.0)
    return min(int(math.floor(p * 4294967296.0)), 4294967296)


# --- initial ion state (opendust.py:370-466) -----------------------------------

def initial_positions(R, H, dusts, N, rng=np.random):
    """Uniform positions in the cylinder outside all dusts, drawing from the
    same legacy NumPy stream in the same order as the reference's per-ion
    rejection loop (opendust.py:376-399): per attempt u_r, u_theta, u_z.
    ``dusts`` is a list of (x, y, z, r) in nm.  Vectorised in rounds of exactly
    the number of attempts the loop is still guaranteed to make, so the stream
    position afterwards equals the loop's."""
    out = np.empty((N, 3), dtype=np.float64)
    filled = 0
    while filled < N:
        need = N - filled
        u = rng.random_sample(3 * need).reshape(need, 3)
        r = np.sqrt(u[:, 0])
        theta = u[:, 1]
        x = R * r * np.cos(2 * np.pi * theta)
        y = R * r * np.sin(2 * np.pi * theta)
        z = -H / 2 + (H / 2 - (-H / 2)) * u[:, 2]
        ok = np.ones(need, dtype=bool)
        for (dx, dy, dz, dr) in dusts:
            d = np.sqrt((x - dx) ** 2 + (y - dy) ** 2 + (z - dz) ** 2)
            ok &= ~(d <= dr)
        k = int(ok.sum())
        out[filled:filled + k, 0] = x[ok]
        out[filled:filled + k, 1] = y[ok]
        out[filled:filled + k, 2] = z[ok]
        filled += k
    return out


def field_driven_vz(v_fl, v_T, N, rng=np.random):
    """v_z of the fieldDriven initial state by inverse CDF (opendust.py:407-444)."""
    v_max = abs(v_fl) * 8 + 5 * v_T
    v_min = -4 * v_T
    n_v = int((v_max - v_min) / (0.2 * v_T))
    v = np.linspace(v_min, v_max, n_v)
    _, xi = _cdf_table(_density_field_driven(v_fl, v_T), v_min, v_max, v)
    xi_vals = rng.uniform(0, 1, N)
    return interp1d(xi, v)(xi_vals)


def initial_velocities(distribution, v_fl, v_T, N, rng=np.random):
    """opendust.py:401-448 (same NumPy calls, same order)."""
    if distribution == "Maxwellian":
        vx = rng.normal(0, v_T, N)
        vy = rng.normal(0, v_T, N)
        vz = rng.normal(v_fl, v_T, N)
    elif distribution == "fieldDriven":
        vx = rng.normal(0, v_T, N)
        vy = rng.normal(0, v_T, N)
        vz = field_driven_vz(v_fl, v_T, N, rng)
    else:
        # the reference leaves v_x undefined here and fails with NameError
        raise NameError("name 'v_x' is not defined (distributionType %r)" % (distribution,))
    return np.stack([vx, vy, vz], axis=1)


def read_restart(path, N):
    """RESTART reader: N lines 'x\\ty\\tz\\tvx\\tvy\\tvz' (opendust.py:452-466)."""
    x = np.empty((N, 3))
    v = np.empty((N, 3))
    with open(path, "r") as fh:
        for i in range(N):
            line = fh.readline()
            parts = line.split("\t")
            if len(parts) < 6:
                raise ValueError("RESTART file %s: line %d is malformed" % (path, i + 1))
            vals = [float(p) for p in parts[:6]]
            x[i] = vals[0:3]
            v[i] = vals[3:6]
    return x, v


# --- trap: background potential fit (opendust.py:499-548) ----------------------

def trap_grid_shape(R_over_lambda, H_over_lambda):
    """Grid and Monte-Carlo sample counts of opendust.py:501-509."""
    n_r = int(R_over_lambda * 10)
    n_h = int(H_over_lambda * 5)
    n_mc = int(np.pi * R_over_lambda ** 2.0 * H_over_lambda / (0.05) ** 3.0)
    return n_r, n_h, n_mc


def trap_fit(potential, R_over_lambda, H_over_lambda, N):
    """Least-squares fit of the 15 even monomials (opendust.py:521-548).
    ``potential``: [n_r][n_h] mean of exp(-d)/(d+0.001); N: superion count
    (the reference passes self.N at opendust.py:550)."""
    n_r, n_h = potential.shape
    grid_R = np.linspace(0, R_over_lambda, n_r)
    grid_H = np.linspace(0, H_over_lambda / 2.0, n_h)
    gH, gR = np.meshgrid(grid_H, grid_R, copy=False)
    gR = gR.flatten()
    gH = gH.flatten()
    A = np.vstack([
        gR * 0 + 1, gR ** 2, gH ** 2, gH ** 2 * gR ** 2, gR ** 4, gH ** 4,
        gR ** 4 * gH ** 2, gR ** 2 * gH ** 4, gR ** 6, gH ** 6, gR ** 4 * gH ** 4,
        gR ** 6 * gH ** 2, gR ** 2 * gH ** 6, gR ** 8, gH ** 8,
    ]).T
    B = np.asarray(potential, dtype=np.float64).flatten()
    return -np.linalg.lstsq(A, B, rcond=None)[0] * N


# exponents (a, b) of rho^a h^b for alpha[0..14] (opendust.py:527-545)
TRAP_MONOMIALS = [(0, 0), (2, 0), (0, 2), (2, 2), (4, 0), (0, 4), (4, 2), (2, 4),
                  (6, 0), (0, 6), (4, 4), (6, 2), (2, 6), (8, 0), (0, 8)]


def trap_potential_poly(alpha, rho, h):
    """sum_{m>=1} alpha_m rho^a h^b (the trap expression, opendust.py:552-576)."""
    s = 0.0
    for m in range(1, 15):
        a, b = TRAP_MONOMIALS[m]
        s = s + alpha[m] * rho ** a * h ** b
    return s
