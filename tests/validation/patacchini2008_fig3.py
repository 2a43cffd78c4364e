"""Patacchini 2008, Fig. 3 (examples/Patacchini2008/Figure3/launchPatacchini2008Figure3.py):
the reference's 13-point scan of ion-neutral collision frequency w_c for one
freely charging grain in a field-driven H+ flow, compared point by point with
the time-averaged charge and z-forces the reference printed
(tests/golden/patacchini2008_fig3_reference.txt).  This is the only reference
data for the dust charge, and it exercises collisions and the field-driven
reinjection.

Each point: H+ (m_i = 1.673557e-27 kg), T_e = 30000 K, T_n = 300 K,
n = 3.57167962497e15 m^-3, (w_c, E) from the scan, R = 3, H = 6 r_D_e, 2^17
superions, time step 1e-11 s; one grain of radius 10 um at z = -r_D_e,
consistent charging from q = 0.  The reference runs 200000 steps and averages
over i > 150000.  Here the run length is --steps (default 200000) and the
average covers the last min(50000, steps / 2) steps: the reference's window
at 200000 steps, i > 50000 at 100000 (the charge relaxes in far fewer
steps, and ions cross the domain in ~8000).  Reported: mean charge and mean
orbit and collection z-forces, each with its blocking standard error
(Flyvbjerg and Petersen, J. Chem. Phys. 91, 461 (1989)), the difference
between the averaging window's two halves (a drift check), and
z = (ours - reference) / (sqrt(2) sigma).

    python tests/validation/patacchini2008_fig3.py --points 0 [--steps 100000] [--realisation 1] [--out results.json]

--realisation R > 0 runs an independent realisation (Philox key, trap seed and
NumPy seed offset by 1000 R); R = 0 is the default run.

Runs on the GPU (one point of 100000 steps: about 9 minutes on one MI355X).
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

W_C = [1.57319303e04, 1.57319303e05, 4.87689840e05, 1.66758461e06, 4.93982612e06, 1.64084033e07, 5.00275384e07,
       1.65562835e08, 4.89121446e08, 1.65767350e09, 3.22479401e09, 8.05204244e09, 1.61113688e10]
E_FIELD = [2.58520000e00, 2.58520000e01, 8.01412000e01, 2.74031200e02, 8.11752800e02, 2.69636360e03, 8.22093600e03,
           2.72066448e04, 8.03764532e04, 2.72402524e05, 5.29924637e05, 1.32317775e06, 2.64755244e06]


def build(p, steps, seed):
    from opendust_amd.opendust import (DustParticle, OpenDust, OutputParameters, PlasmaParametersInSIUnitsFieldDriven,
                                       SimulationParametersInSIUnits)

    pp = PlasmaParametersInSIUnitsFieldDriven(300, 30000, 3.57167962497e15, E_FIELD[p], W_C[p], 1.673557e-27)
    lam = pp.r_D_e
    sp = SimulationParametersInSIUnits(3 * lam, 6 * lam, int(2 ** 17), steps, 1e-11, pp)
    dust = DustParticle(0, 0, -1 * lam, 1e-05, "consistent", 0 * pp.e)
    return OpenDust(pp, sp, OutputParameters(10 ** 9, 10 ** 9), [dust], "fieldDriven", precision="fp32c", seed=seed,
                    trapSeed=seed + 1)


def blocking_error(a, min_blocks=8):
    """Standard error of the mean of a correlated series by repeated pairwise
    blocking: the largest estimate over the levels that keep >= min_blocks blocks."""
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
    ap.add_argument("--points", default="0")
    ap.add_argument("--steps", type=int, default=200000)
    ap.add_argument("--seed", type=int, default=8)
    ap.add_argument("--realisation", type=int, default=0)
    args = ap.parse_args()
    ref = np.loadtxt(os.path.join(os.path.dirname(HERE), "golden", "patacchini2008_fig3_reference.txt"))
    results = []
    for p in (int(t) for t in args.points.split(",")):
        with contextlib.redirect_stdout(io.StringIO()):
            od = build(p, args.steps, args.seed + p + 1000 * args.realisation)
        od.outputParameters.nOutput = 10000
        od.outputParameters.nFileOutput = 10000
        np.random.seed(300 + p + 1000 * args.realisation)
        t0 = time.time()
        print("w_c %.4e, E %.4e V/m: %d steps" % (W_C[p], E_FIELD[p], args.steps), flush=True)
        od.simulate("0")
        wall = time.time() - t0
        d = od.dustParticles[0]
        start = args.steps - min(50000, args.steps // 2) + 1
        row = {"w_c": W_C[p], "steps": args.steps, "realisation": args.realisation, "average_from_step": start,
               "wall_s": wall}
        for name, series, r in (("q", d.q, ref[p, 1]), ("F_orbit_z", d.forceIonsOrbit[:, 2], ref[p, 2]),
                                ("F_collect_z", d.forceIonsCollect[:, 2], ref[p, 3])):
            a = np.asarray(series[start:], dtype=np.float64)
            m, s = float(a.mean()), blocking_error(a)
            half = len(a) // 2
            drift = float(a[half:].mean() - a[:half].mean())
            row[name] = {"mean": m, "sigma": s, "reference": float(r), "z": float((m - r) / (np.sqrt(2.0) * s)),
                         "half_difference": drift,
                         "block_means_1000": a[:len(a) // 1000 * 1000].reshape(-1, 1000).mean(axis=1).tolist()}
            print("  %-12s %.4e +- %.1e (reference %.4e, %+.1f %%, z %+.2f; second minus first half %+.1e)"
                  % (name, m, s, r, 100.0 * (m - r) / abs(r), row[name]["z"], drift), flush=True)
        print("  %.0f s" % wall, flush=True)
        results.append(row)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(results, fh, indent=1)


if __name__ == "__main__":
    main()
