"""Piel 2017, Fig. 7 (examples/Piel2017/Figure7/launchPiel2017Figure7.py): the
reference's 11-point scan of the wake force between two dust grains, run here
as one batched sweep (opendust_amd.sweep.simulateBatch) and compared point by
point with the means the reference printed (tests/golden/piel2017_fig7_reference.txt).

Each point: Ar+, T_e = 29011 K, T_n = 290.11 K, n = 1e14 m^-3, M = 1, R = 3,
H = 6 r_D_e, 2^17 superions, 5000 steps of 0.5148240854e-9 s; two grains of
radius 25 um with given charge 170000 e at (0, 0, -0.5 r_D_e) and
(d_x r_D_e, 0, 0.5 r_D_e).  Reported per point: the means over steps i > 2500
of the x and z forces (orbit + dust-dust) on both grains, with the standard
error from 10 block averages, and z = (ours - reference) / (sqrt(2) sigma).

    python tests/validation/piel2017_fig7.py [--out results.json] [--points 0,1,...]

A point listed more than once (--points 5,5,5) runs independent realisations,
which measure the run-to-run spread directly.

Runs on the GPU (about 5 minutes for all 11 points on one MI355X).
"""

import argparse
import contextlib
import io
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

D_X = [0.00176, 0.10369, 0.20211, 0.30053, 0.40246, 0.50088, 0.60105, 0.70123, 0.80141, 0.89982, 1.00000]


def build(points, seed=17):
    """One OpenDust per entry of `points`; a point listed again gets an
    independent realisation (Philox key, trap seed)."""
    from opendust_amd.opendust import (DustParticle, OpenDust, OutputParameters, PlasmaParametersInSIUnitsMaxwell,
                                       SimulationParametersInSIUnits)

    ods = []
    for k, p in enumerate(points):
        rep = points[:k].count(p)
        pp = PlasmaParametersInSIUnitsMaxwell(290.11, 29011, 1e14, 1, 6.6335209e-26)
        lam = pp.r_D_e
        sp = SimulationParametersInSIUnits(3 * lam, 6 * lam, int(2 ** 17), 5000, 0.5148240854e-09, pp)
        q = 170000.0 * pp.e
        d1 = DustParticle(0, 0, -0.5 * lam, 25.0e-6, "given", q)
        d2 = DustParticle(D_X[p] * lam, 0, 0.5 * lam, 25.0e-6, "given", q)
        ods.append(OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), [d1, d2], "Maxwellian", precision="fp32c",
                            seed=seed + p + 1000 * rep, trapSeed=5 + rep))
    return ods


def block_mean(a, n_blocks=10):
    a = np.asarray(a, dtype=np.float64)
    k = len(a) // n_blocks
    blocks = a[len(a) - k * n_blocks:].reshape(n_blocks, k).mean(axis=1)
    return float(a.mean()), float(blocks.std(ddof=1) / np.sqrt(n_blocks))


def blocking_error(a, min_blocks=8):
    """Standard error of the mean of a correlated series by repeated pairwise
    blocking (Flyvbjerg and Petersen, J. Chem. Phys. 91, 461 (1989)): the
    largest estimate over the blocking levels that keep >= min_blocks blocks."""
    x = np.asarray(a, dtype=np.float64)
    best = 0.0
    while len(x) >= min_blocks:
        best = max(best, float(x.std(ddof=1) / np.sqrt(len(x))))
        n = len(x) // 2
        x = 0.5 * (x[0:2 * n:2] + x[1:2 * n:2])
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--points", default=",".join(str(i) for i in range(11)))
    args = ap.parse_args()
    points = [int(t) for t in args.points.split(",")]
    ref = np.loadtxt(os.path.join(os.path.dirname(HERE), "golden", "piel2017_fig7_reference.txt"))
    from opendust_amd import sweep

    with contextlib.redirect_stdout(io.StringIO()):
        ods = build(points)
    np.random.seed(2017)
    t0 = time.time()
    sweep.simulateBatch(ods, "0", tune=True,
                        progress=lambda i, n: print("  step %d of %d, %.0f s" % (i, n, time.time() - t0), flush=True))
    wall = time.time() - t0
    rows = []
    for od, p in zip(ods, points):
        d1, d2 = od.dustParticles
        f = [d1.forceIonsOrbit[:, 0] + d1.forceDust[:, 0], d1.forceIonsOrbit[:, 2] + d1.forceDust[:, 2],
             d2.forceIonsOrbit[:, 0] + d2.forceDust[:, 0], d2.forceIonsOrbit[:, 2] + d2.forceDust[:, 2]]
        row = {"d_x": D_X[p], "realisation": points[:len(rows)].count(p)}
        for name, series, r in zip(("F1x", "F1z", "F2x", "F2z"), f, ref[p, 1:]):
            m, s = block_mean(series[2501:])
            sb = blocking_error(series[2501:])
            # [mean, 10-block error, reference, z(10-block), blocking error, z(blocking)]
            row[name] = [m, s, float(r), float((m - r) / (np.sqrt(2.0) * s)), sb, float((m - r) / (np.sqrt(2.0) * sb))]
        rows.append(row)
        print("d_x %.5f  " % D_X[p] + "  ".join("%s %.3e+-%.1e (ref %.3e, z %+.2f; blocking +-%.1e, z %+.2f)"
                                                % (k, *row[k]) for k in ("F1x", "F1z", "F2x", "F2z")), flush=True)
    zs = np.array([row[k][3] for row in rows for k in ("F1x", "F1z", "F2x", "F2z")])
    summary = {"points": len(rows), "wall_s": wall, "max_abs_z": float(np.max(np.abs(zs))),
               "rms_z": float(np.sqrt(np.mean(zs ** 2))), "rows": rows}
    print("max |z| %.2f, rms z %.2f over %d values; batched run %.1f s" % (summary["max_abs_z"], summary["rms_z"],
                                                                        zs.size, wall), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()
