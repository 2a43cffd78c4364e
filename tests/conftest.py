import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the engine through the C ABI")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    oracle.load()
    return oracle


def say(*a):
    """Progress line that names the case before a GPU step (flushed)."""
    print(*a, flush=True)


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def fp32c_position_bounds(oracle, cfg, x0, flags_orc):
    """Per-ion bound of |x_gpu - x_orc| after one FP32c step from x0.  The
    drift uses the ion-ion force evaluated at x0, so for a kept ion
    |dx_i| = dt * (dt/2) / m_s * |dF_i(x0)|, and the pair-force error model
    |dF_i| <= 1e-4 |F_i| + 1e-5 rms|F| bounds it; reinjected ions take
    positions computed in FP64 from the same random numbers (FP64 rounding)."""
    f = oracle.pair_forces(cfg, x0)
    fn = np.linalg.norm(f, axis=1)
    rms = np.sqrt(np.mean(fn ** 2))
    kept = np.asarray(flags_orc) == 0
    return np.where(kept, cfg.dt * 0.5 * cfg.dt / cfg.m_s * (1e-4 * fn + 1e-5 * rms), 0.0) + 1e-12 * cfg.R


def fp32c_position_error_ok(oracle, cfg, x0, x_gpu, x_orc, flags_orc):
    """(all ions within fp32c_position_bounds, worst error/bound ratio, bounds)"""
    bound = fp32c_position_bounds(oracle, cfg, x0, flags_orc)
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
