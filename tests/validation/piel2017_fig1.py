"""Piel 2017, Fig. 1 (examples/Piel2017/Figure1/{consistent,oml,given}/): the
reference's three single-grain runs through the drop-in API, compared with the
mean z-forces the reference printed (outputPiel2017Fig1Consistent:40-41,
outputPiel2017Fig1OML:40-41, outputPiel2017Fig1Given:337-338).

Ar+, T_e = 29011 K, T_i = 290.11 K, n = 1e14 m^-3, M = 1, R = 3, H = 6 r_D_e,
3000 steps of 3.5148240854e-9 s, one grain of radius 58.8 um at z = -r_D_e;
consistent and OML charging from q = 0 with 2^16 superions (the row kernel),
given charge 392500 e with 2^19 superions (the symmetric kernel).  Means over
steps i > 1000, standard error from 10 block averages (an underestimate: ions
need about 800 steps to cross the domain), z = (ours - reference)/(sqrt(2) sigma).

    python tests/validation/piel2017_fig1.py [--out results.json] [--cases consistent,oml,given]
"""

import argparse
import contextlib
import io
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

# mean z-forces (N) over i > 1000 printed by the reference runs: (orbit, collection)
REFERENCE = {
    "consistent": (2 ** 16, 0.0, 4.495695908591751e-11, 8.97145418766947e-12),
    "oml": (2 ** 16, 0.0, 4.48957607439666e-11, 9.203742146333308e-12),
    "given": (2 ** 19, 392500.0, 4.8917982383641007e-11, 9.188077929110581e-12),
}


def block_mean(a, n_blocks=10):
    a = np.asarray(a, dtype=np.float64)
    k = len(a) // n_blocks
    blocks = a[len(a) - k * n_blocks:].reshape(n_blocks, k).mean(axis=1)
    return float(a.mean()), float(blocks.std(ddof=1) / np.sqrt(n_blocks))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="")
    ap.add_argument("--cases", default="consistent,oml,given")
    ap.add_argument("--seed", type=int, default=0, help="offset of the Philox key and the NumPy seed")
    args = ap.parse_args()
    from opendust_amd import workloads

    results = {}
    for k, case in enumerate(args.cases.split(",")):
        N, q_e, ref_orbit, ref_collect = REFERENCE[case]
        t0 = time.time()
        print("%s: N %d, 3000 steps" % (case, N), flush=True)
        with contextlib.redirect_stdout(io.StringIO()):
            od = workloads.piel_fig1(N=N, n=3000, charge=case, q_e=q_e, precision="fp32c",
                                     seed=101 + k + 1000 * args.seed, trapSeed=42 + args.seed)
        # progress lines every 500 steps (no files are written)
        od.outputParameters.nOutput = 500
        od.outputParameters.nFileOutput = 500
        np.random.seed(7 + k + 1000 * args.seed)
        od.simulate("0")
        d = od.dustParticles[0]
        orbit = block_mean(d.forceIonsOrbit[1001:, 2])
        collect = block_mean(d.forceIonsCollect[1001:, 2])
        z_o = (orbit[0] - ref_orbit) / (np.sqrt(2.0) * orbit[1])
        z_c = (collect[0] - ref_collect) / (np.sqrt(2.0) * collect[1])
        results[case] = {"N": N, "orbit": [orbit[0], orbit[1], ref_orbit, z_o],
                         "collect": [collect[0], collect[1], ref_collect, z_c], "q_end": float(d.q[-1]),
                         "wall_s": time.time() - t0}
        print("  orbit %.4e +- %.2e (reference %.4e, z %+.2f); collection %.4e +- %.2e (reference %.4e, z %+.2f); "
              "%.0f s" % (orbit[0], orbit[1], ref_orbit, z_o, collect[0], collect[1], ref_collect, z_c,
                          time.time() - t0), flush=True)
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(results, fh, indent=1)


if __name__ == "__main__":
    main()
