"""Piel 2017, Fig. 5 (examples/Piel2017/Figure5/launchPiel2017Figure5.py): the
reference's 12-point Mach scan of the ion forces on one grain with OML
charging, run here as one batched sweep (opendust_amd.sweep.simulateBatch)
and compared point by point with the means the reference printed
(tests/golden/piel2017_fig5_reference.txt, from outputPiel2017Fig5).

Each point: Ar+, T_e = 29011 K, T_n = 290.11 K, n = 1e14 m^-3, Mach number M
from the scan, R = 3, H = 6 r_D_e, 2^16 superions, 5000 steps of
3.5148240854e-9 s; one grain of radius 58.8 um at the origin, OML charging
from 392500 e.  Reported per point: the means over steps i > 3000 of the
orbit and collection z-forces, as the reference averages them
(launchPiel2017Figure5.py:133-141).

The averaging window (2000 steps) is shorter than the time ions need to
cross the domain (about 4000 steps at M = 0.2, 800 at M = 2), so the
within-run blocking error understates the uncertainty of one run.  Each
point therefore runs --reps independent realisations (Philox key, NumPy seed
and trap seed varied), whose spread s measures the run-to-run scatter
directly.  The reference is one such run: the difference between our mean
over the realisations and the reference has the standard deviation
s sqrt(1 + 1/reps), and z = (ours - reference) / (s sqrt(1 + 1/reps)).

    python tests/validation/piel2017_fig5.py [--out results.json] [--points 0,1,...] [--reps 3]

Runs on the GPU (12 points x 3 realisations in one batch: a few minutes on
one MI355X).
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

MACH = [0.19679, 0.30120, 0.40161, 0.49799, 0.60241, 0.69880, 0.80321, 0.89960, 0.99598, 1.19679, 1.49799,
        2.00000]
STEPS = 5000
AVERAGE_FROM = 3001  # the reference averages i > 3000


def build(points, reps, seed=55):
    """One OpenDust per (point, realisation), realisation-major."""
    from opendust_amd.opendust import (DustParticle, OpenDust, OutputParameters, PlasmaParametersInSIUnitsMaxwell,
                                       SimulationParametersInSIUnits)

    ods, keys = [], []
    for rep in range(reps):
        for p in points:
            pp = PlasmaParametersInSIUnitsMaxwell(290.11, 29011, 1e14, MACH[p], 6.6335209e-26)
            lam = pp.r_D_e
            sp = SimulationParametersInSIUnits(3 * lam, 6 * lam, int(2 ** 16), STEPS, 3.5148240854e-09, pp)
            dust = DustParticle(0, 0, 0 * lam, 58.8e-6, "oml", 392500.0 * pp.e)
            ods.append(OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), [dust], "Maxwellian",
                                precision="fp32c", seed=seed + p + 1000 * rep, trapSeed=5 + 7 * rep))
            keys.append((p, rep))
    return ods, keys


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
    ap.add_argument("--points", default=",".join(str(i) for i in range(len(MACH))))
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    points = [int(t) for t in args.points.split(",")]
    if args.reps < 2:
        raise SystemExit("--reps must be at least 2 (the run-to-run spread needs two realisations)")
    ref = np.loadtxt(os.path.join(os.path.dirname(HERE), "golden", "piel2017_fig5_reference.txt"))
    from opendust_amd import sweep

    with contextlib.redirect_stdout(io.StringIO()):
        ods, keys = build(points, args.reps)
    np.random.seed(2017)
    t0 = time.time()
    print("points %s x %d realisations, %d steps" % (points, args.reps, STEPS), flush=True)
    sweep.simulateBatch(ods, "0", tune=True,
                        progress=lambda i, n: print("  step %d of %d, %.0f s" % (i, n, time.time() - t0), flush=True))
    wall = time.time() - t0
    rows, zs = [], []
    for p in points:
        row = {"M": MACH[p], "realisations": {}}
        for name, col in (("F_orbit_z", 1), ("F_collect_z", 2)):
            runs = []
            for od, (pk, rep) in zip(ods, keys):
                if pk != p:
                    continue
                d = od.dustParticles[0]
                series = (d.forceIonsOrbit if name == "F_orbit_z" else d.forceIonsCollect)[AVERAGE_FROM:, 2]
                runs.append((float(np.mean(series)), blocking_error(series)))
            means = np.array([m for m, _ in runs])
            mean, spread = float(means.mean()), float(means.std(ddof=1))
            r = float(ref[p, col])
            sigma = spread * np.sqrt(1.0 + 1.0 / len(means))
            z = (mean - r) / sigma if sigma > 0 else float("inf")
            zs.append(z)
            row[name] = {"mean": mean, "run_to_run_sd": spread, "reference": r, "z": z,
                         "per_realisation": [{"mean": m, "blocking_error": s} for m, s in runs]}
        q = [od.dustParticles[0].q[AVERAGE_FROM:].mean() for od, (pk, _) in zip(ods, keys) if pk == p]
        row["q_mean_over_e"] = float(np.mean(q) / ods[0].plasmaParametersInSIUnits.e)
        rows.append(row)
        print("M %.5f  orbit %.3e (sd %.1e; ref %.3e, z %+.2f)  collect %.3e (sd %.1e; ref %.3e, z %+.2f)  q %.0f e"
              % (MACH[p], row["F_orbit_z"]["mean"], row["F_orbit_z"]["run_to_run_sd"], row["F_orbit_z"]["reference"],
                 row["F_orbit_z"]["z"], row["F_collect_z"]["mean"], row["F_collect_z"]["run_to_run_sd"],
                 row["F_collect_z"]["reference"], row["F_collect_z"]["z"], row["q_mean_over_e"]), flush=True)
    zs = np.array(zs)
    summary = {"points": points, "reps": args.reps, "steps": STEPS, "average_from_step": AVERAGE_FROM, "wall_s": wall,
               "max_abs_z": float(np.max(np.abs(zs))), "rms_z": float(np.sqrt(np.mean(zs ** 2))), "rows": rows}
    print("max |z| %.2f, rms z %.2f over %d values; batched run %.1f s"
          % (summary["max_abs_z"], summary["rms_z"], zs.size, wall), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(summary, fh, indent=1)


if __name__ == "__main__":
    main()
