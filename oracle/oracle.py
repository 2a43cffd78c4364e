"""ctypes wrapper of the FP64 CPU oracle (oracle/od_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the engine.  Build with
``make -C oracle`` (or ``oracle.build()``).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


def load(path=None):
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        path = LIB_PATH
        # make rebuilds only when od_oracle.c or include/opendust_amd.h changed;
        # a stale library would read od_config with the wrong layout
        if os.path.exists(os.path.join(HERE, "od_oracle.c")) and subprocess.run(
                ["make", "-s", "-C", HERE], capture_output=True).returncode != 0 and not os.path.exists(LIB_PATH):
            raise RuntimeError("building the oracle failed")
    from opendust_amd.engine import Config, DustRecord, Table  # structure layouts only

    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    lib.orc_philox.argtypes = [P(ctypes.c_uint32), ctypes.c_uint64, P(ctypes.c_uint32)]
    lib.orc_table_eval.argtypes = [P(Table), ctypes.c_double]
    lib.orc_table_eval.restype = ctypes.c_double
    lib.orc_pair_forces.argtypes = [P(Config), ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_cut_edge_count.argtypes = [P(Config), ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       P(ctypes.c_int), ctypes.c_int]
    lib.orc_ext_force.argtypes = [P(Config), dp, dp, dp]
    lib.orc_reinject.argtypes = [P(Config), ctypes.c_uint32, ctypes.c_uint32, dp, dp]
    lib.orc_reinject.restype = ctypes.c_int
    lib.orc_collide.argtypes = [P(Config), ctypes.c_uint32, ctypes.c_uint32, dp]
    lib.orc_flags.argtypes = [P(Config), dp]
    lib.orc_flags.restype = ctypes.c_uint32
    lib.orc_orbit_term.argtypes = [P(Config), ctypes.c_int, dp, dp]
    lib.orc_charge.argtypes = [P(Config), ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_int64]
    lib.orc_charge.restype = ctypes.c_double
    lib.orc_out_domain_force.argtypes = [P(Config), ctypes.c_int, ctypes.c_double, dp]
    lib.orc_step.argtypes = [P(Config), ctypes.c_int, dp, dp, dp, ctypes.c_int64, P(DustRecord),
                             P(ctypes.c_uint32), P(ctypes.c_uint8), ctypes.c_int]
    lib.orc_run.argtypes = [P(Config), ctypes.c_int, dp, dp, dp, ctypes.c_int64, ctypes.c_int64, P(DustRecord),
                            ctypes.c_int]
    lib.orc_trap_sample.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int64, dp]
    lib.orc_trap_grid.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, ctypes.c_int, dp,
                                  ctypes.c_int64, ctypes.c_uint64, dp, ctypes.c_int]
    lib.orc_max_threads.restype = ctypes.c_int
    lib.orc_config_size.restype = ctypes.c_int64
    lib.orc_cell_of.argtypes = [P(Config), dp, P(ctypes.c_int)]
    lib.orc_cell_of.restype = ctypes.c_int
    if lib.orc_config_size() != ctypes.sizeof(Config):
        raise RuntimeError("oracle library was built against a different od_config layout; run make -C oracle")
    _lib = lib
    return lib


def use_native_build():
    """Rebuild the oracle for the host it runs on (-O3 -march=native, OpenMP)
    in a temporary directory and make this module use that build from now on
    (the CPU-baseline leg of the benchmark; the committed Makefile targets
    x86-64-v3 so that one build runs on any host).  Returns the compile
    command."""
    global _lib
    import tempfile

    load()  # layout check of the regular build first
    out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
           "-I", os.path.join(HERE, "..", "include"), "-o", out, os.path.join(HERE, "od_oracle.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    saved, _lib = _lib, None
    try:
        load(out)
    except Exception:
        _lib = saved
        raise
    return " ".join(cmd[:6])


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo), or the platform's."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or platform.machine()


def _d(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def threads():
    return load().orc_max_threads()


def philox(ctr, key):
    lib = load()
    c = (ctypes.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    o = (ctypes.c_uint32 * 4)()
    lib.orc_philox(c, int(key) & (2 ** 64 - 1), o)
    return [int(v) for v in o]


def table_eval(table_struct, u):
    return load().orc_table_eval(ctypes.byref(table_struct), float(u))


def pair_forces(cfg, x, rows=None, nthreads=0):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    i0, i1 = (0, n) if rows is None else rows
    f = np.empty((i1 - i0, 3))
    lib.orc_pair_forces(ctypes.byref(cfg), n, _d(x), _d(f), i0, i1, nthreads or threads())
    return f


def cut_edge_count(cfg, x, rel, rows=None, nthreads=0):
    """Per row: other ions at a distance within rel * cutoff of the cutoff
    (pairs whose truncation a relative rounding of r by rel may flip)."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    i0, i1 = (0, n) if rows is None else rows
    out = np.empty(i1 - i0, dtype=np.int32)
    lib.orc_cut_edge_count(ctypes.byref(cfg), n, _d(x), i0, i1, float(rel),
                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), nthreads or threads())
    return out


def cut_force_at_cutoff(cfg):
    """Magnitude of one pair's force just inside the cutoff (kJ/mol/nm): the
    most a single flipped truncation decision can change a row, with the
    switch: zero at the cutoff."""
    if cfg.switch_dist > 0.0 and cfg.switch_dist < cfg.cutoff:
        return 0.0
    r, lam = cfg.cutoff, cfg.r_D_e
    rr = r + cfg.r_0
    return cfg.kqq * np.exp(-r / lam) * (1.0 / (lam * rr) + 1.0 / (rr * rr))


def ext_force(cfg, q, xi):
    lib = load()
    q = np.ascontiguousarray(np.resize(np.asarray(q, dtype=np.float64), max(cfg.n_dust, 1)))
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    f = np.empty(3)
    lib.orc_ext_force(ctypes.byref(cfg), _d(q), _d(xi), _d(f))
    return f


def reinject(cfg, ion, step):
    lib = load()
    x = np.zeros(3)
    v = np.zeros(3)
    face = lib.orc_reinject(ctypes.byref(cfg), ion, step, _d(x), _d(v))
    return face, x, v


def collide(cfg, ion, step, v):
    v = np.array(v, dtype=np.float64)
    load().orc_collide(ctypes.byref(cfg), ion, step, _d(v))
    return v


def flags(cfg, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return int(load().orc_flags(ctypes.byref(cfg), _d(x)))


def orbit_term(cfg, p, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    o = np.empty(3)
    load().orc_orbit_term(ctypes.byref(cfg), p, _d(x), _d(o))
    return o


def charge(cfg, p, step, q_prev, absorbed):
    return load().orc_charge(ctypes.byref(cfg), p, int(step), float(q_prev), int(absorbed))


def out_domain_force(cfg, p, qp):
    f = np.empty(3)
    load().orc_out_domain_force(ctypes.byref(cfg), p, float(qp), _d(f))
    return f


def step(cfg, x, v, q, step_index, nthreads=0):
    """One oracle step in place on copies; returns (x, v, q, records, flags, faces)."""
    from opendust_amd.engine import DustRecord, RECORD_DTYPE

    lib = load()
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    v = np.array(v, dtype=np.float64).reshape(-1, 3)
    q = np.array(np.resize(np.asarray(q, dtype=np.float64), max(cfg.n_dust, 1)))
    n = x.shape[0]
    rec = np.zeros(max(cfg.n_dust, 1), dtype=RECORD_DTYPE)
    fl = np.zeros(n, dtype=np.uint32)
    face = np.zeros(n, dtype=np.uint8)
    rc = lib.orc_step(ctypes.byref(cfg), n, _d(x), _d(v), _d(q), int(step_index),
                      rec.ctypes.data_as(ctypes.POINTER(DustRecord)),
                      fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                      face.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), nthreads or threads())
    if rc:
        raise RuntimeError("orc_step failed")
    return x, v, q[: cfg.n_dust], rec[: cfg.n_dust], fl, face


def run(cfg, x, v, q, step0, nsteps, nthreads=0):
    """nsteps oracle steps; returns (x, v, q, records[nsteps][n_dust])."""
    from opendust_amd.engine import DustRecord, RECORD_DTYPE

    lib = load()
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    v = np.array(v, dtype=np.float64).reshape(-1, 3)
    q = np.array(np.resize(np.asarray(q, dtype=np.float64), max(cfg.n_dust, 1)))
    rec = np.zeros((nsteps, max(cfg.n_dust, 1)), dtype=RECORD_DTYPE)
    rc = lib.orc_run(ctypes.byref(cfg), x.shape[0], _d(x), _d(v), _d(q), int(step0), int(nsteps),
                     rec.ctypes.data_as(ctypes.POINTER(DustRecord)), nthreads or threads())
    if rc:
        raise RuntimeError("orc_run failed")
    return x, v, q[: cfg.n_dust], rec[:, : cfg.n_dust]


def trap_sample(R, H, seed, k):
    o = np.empty(3)
    load().orc_trap_sample(float(R), float(H), int(seed) & (2 ** 64 - 1), int(k), _d(o))
    return o


def trap_grid(R, H, rho, h, n_mc, seed, nthreads=0):
    lib = load()
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    out = np.empty((len(rho), len(h)))
    rc = lib.orc_trap_grid(float(R), float(H), len(rho), _d(rho), len(h), _d(h), int(n_mc),
                           int(seed) & (2 ** 64 - 1), _d(out), nthreads or threads())
    if rc:
        raise RuntimeError("orc_trap_grid failed")
    return out


def cell_of(cfg, x):
    """(cell id, (nx, ny, nz)) of a position under the engine's binning."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dims = (ctypes.c_int * 3)()
    cid = load().orc_cell_of(ctypes.byref(cfg), _d(x), dims)
    return cid, tuple(dims)


def density_profile(z_nm, R, L, N_s, n=100):
    """(z_grid[n] m, concentration[n] 1/m^3) of ions at axial positions z_nm (nm)."""
    z = np.ascontiguousarray(z_nm, dtype=np.float64)
    zg = np.empty(n)
    c = np.empty(n)
    lib = load()
    lib.orc_density_profile.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
    lib.orc_density_profile.restype = None
    lib.orc_density_profile(_d(z), len(z), int(n), float(R), float(L), float(N_s), _d(zg), _d(c))
    return zg, c


# ---------------------------------------------------------------------------
# Error model of one FP32c step against this oracle (used by the tests and the
# smoke check): per-ion position bounds from the pair-force error bound, their
# propagation through the orbit sums, and the record comparison.
# ---------------------------------------------------------------------------
def fp32c_position_bounds(oracle_mod, cfg, x0, flags_orc):
    """Per-ion bound of |x_gpu - x_orc| after one FP32c step from x0.  The
    drift uses the ion-ion force evaluated at x0, so for a kept ion
    |dx_i| = dt * (dt/2) / m_s * |dF_i(x0)|, and the pair-force error model
    |dF_i| <= 1e-4 |F_i| + 1e-5 rms|F| bounds it; reinjected ions take
    positions computed in FP64 from the same random numbers (FP64 rounding).
    The force model holds at the test sizes (a few thousand ions): at high
    density (2^20 ions in the benchmark cylinder) the rounding of positions to
    FP32 alone exceeds it for some ions, because ever closer pairs dominate."""
    f = oracle_mod.pair_forces(cfg, x0)
    fn = np.linalg.norm(f, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    kept = np.asarray(flags_orc) == 0
    return np.where(kept, cfg.dt * 0.5 * cfg.dt / cfg.m_s * (1e-4 * fn + 1e-5 * rms), 0.0) + 1e-12 * cfg.R


def fp32c_rounded_positions(cfg, x):
    """The positions the FP32 pair kernels see: x scaled to units of
    r_D_e / log2(e), rounded to FP32 (nearest even), scaled back in FP64."""
    s = 1.4426950408889634 / cfg.r_D_e
    return (np.asarray(x, dtype=np.float64) * s).astype(np.float32).astype(np.float64) / s


def fp32c_position_bounds_measured(oracle_mod, cfg, x0, flags_orc):
    """Per-ion bound of |x_gpu - x_orc| after one FP32c step from x0 at any
    density.  The FP32c force error splits into the rounding of the positions
    (measured here exactly: FP64 forces on the rounded positions minus FP64
    forces on x0) and the kernel's arithmetic on the rounded positions
    (|dF_i| <= 1e-5 |F_i| + 1e-6 rms|F|, compensated sums); a kept ion drifts
    by dt (dt/2) / m_s times that force error."""
    f = oracle_mod.pair_forces(cfg, x0)
    xr = fp32c_rounded_positions(cfg, x0)
    q = np.linalg.norm(oracle_mod.pair_forces(cfg, xr) - f, axis=1)
    fn = np.linalg.norm(f, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    df = q + 1e-5 * fn + 1e-6 * rms
    if cfg.cutoff > 0.0:
        # a pair within the FP32 rounding of r of a hard cutoff may be counted
        # on one side only: one pair force at the cutoff each
        df = df + cut_edge_count(cfg, xr, 1e-6) * cut_force_at_cutoff(cfg) * 1.001
    kept = np.asarray(flags_orc) == 0
    return np.where(kept, cfg.dt * 0.5 * cfg.dt / cfg.m_s * df, 0.0) + 1e-12 * cfg.R


def fp32c_position_error_ok(oracle_mod, cfg, x0, x_gpu, x_orc, flags_orc):
    """(all ions within fp32c_position_bounds, worst error/bound ratio, bounds)"""
    bound = fp32c_position_bounds(oracle_mod, cfg, x0, flags_orc)
    err = np.linalg.norm(np.asarray(x_gpu) - np.asarray(x_orc), axis=1)
    ratio = err / bound
    return bool(np.all(ratio <= 1.0)), float(np.max(ratio)), bound


def orbit_sum_bound(cfg, p, x, pos_bound):
    """Bound of the error of dust p's orbit sum (FP64, over the ions at x) when
    every ion i may be displaced by pos_bound[i]: the term
    T(u) = -g(r) u, u = x_i - x_p, g(r) = e^{-r/l} (l^2/r^3 + l/r^2), has
    |dT/du| <= |g| + r |g'| with g' = -g/l - e^{-r/l} (3 l^2/r^4 + 2 l/r^3)."""
    lam = cfg.r_D_e
    u = np.asarray(x) - np.array([cfg.dust_x[p], cfg.dust_y[p], cfg.dust_z[p]])
    r = np.linalg.norm(u, axis=1)
    e = np.exp(-r / lam)
    g = e * (lam ** 2 / r ** 3 + lam / r ** 2)
    gp = -g / lam - e * (3 * lam ** 2 / r ** 4 + 2 * lam / r ** 3)
    return float(np.sum((np.abs(g) + r * np.abs(gp)) * pos_bound) + 1e-12 * np.sum(np.abs(g) * r))


def check_records(cfg, rec_gpu, rec_orc, precision, x_orc=None, pos_bound=None):
    """Per-dust records of one step, engine against oracle.  Charges,
    dust-dust and collection forces: rel <= 1e-9 (FP64) / 1e-5 (FP32c) of
    their scale.  Orbit sums and orbit forces: rel <= 1e-9 (FP64); for FP32c
    within orbit_sum_bound of the per-ion position bounds (the sums are FP64
    over positions that carry the FP32c drift error), times q k q_s / l^2 for
    the orbit force.  Returns {key: error / tolerance}."""
    out = {}
    tol_r = 1e-9 if precision == "fp64" else 1e-5
    keys = ("q", "force_collect", "force_dust") + (("force_orbit", "orbit_sum") if precision == "fp64" else ())
    for key in keys:
        a = np.asarray(rec_gpu[key], dtype=np.float64)
        b = np.asarray(rec_orc[key], dtype=np.float64)
        sc = max(np.max(np.abs(b)), 1e-300)
        out[key] = float(np.max(np.abs(a - b)) / (tol_r * sc))
    if precision != "fp64":
        for p in range(cfg.n_dust):
            ob = orbit_sum_bound(cfg, p, x_orc, pos_bound)
            d_sum = np.linalg.norm(np.asarray(rec_gpu["orbit_sum"][p]) - np.asarray(rec_orc["orbit_sum"][p]))
            q = float(rec_orc["q"][p])
            fs = abs(q * cfg.si_k * cfg.si_q_s) / cfg.si_r_D_e ** 2
            f_tol = ob * fs + 1e-12 * np.linalg.norm(rec_orc["force_orbit"][p])
            d_f = np.linalg.norm(np.asarray(rec_gpu["force_orbit"][p]) - np.asarray(rec_orc["force_orbit"][p]))
            out["orbit_sum[%d]" % p] = float(d_sum / ob)
            out["force_orbit[%d]" % p] = float(d_f / f_tol)
    return out

