"""Batched parameter sweeps: many OpenDust simulations run together on one GPU.

The reference runs a sweep as a sequential Python loop of independent
``OpenDust(...).simulate()`` calls (e.g.
examples/Piel2017/Figure5/launchPiel2017Figure5.py:30, ``for p in range(12)``).
``simulateBatch`` takes the same OpenDust objects and advances all of them in
one engine batch (od_batch_*): each step is one launch per step kernel for all
simulations, the ion-ion forces of all simulations are one launch, and steps
replay as captured HIP graphs.  The results land in each object's
``dustParticles[p].{q, forceIonsOrbit, forceIonsCollect, forceDust,
forceExternalField}`` exactly as ``simulate()`` leaves them.

Randomness: simulation s of the batch uses its own Philox key (``od.seed``) and
counter word ``sim_id = s``; the same object run alone through
``Engine``/``simulate`` with ``sim_id = s`` draws the same numbers.
File output (xyz frames, restart files, progress lines) stays with
``simulate()``; a batch returns dust records only.
"""

import secrets

import numpy as np

from . import host
from .engine import RUN_GRAPH, Batch, trap_potential_grid

_REC_CHUNK = 1024  # steps per od_batch_run call (bounds the host record buffer)


def fit_traps(openDusts, device=0):
    """Trap coefficients of every simulation (opendust.py:499-550).  The Monte
    Carlo background potential depends only on the geometry in units of r_D_e
    (and the sample count and seed), so simulations sharing those share one GPU
    Monte Carlo grid; each is then fitted with its own superion count N."""
    grids = {}
    alphas = []
    for od in openDusts:
        Rl, Hl = od.R / od.r_D_e, od.H / od.r_D_e
        n_r, n_h, n_mc = host.trap_grid_shape(Rl, Hl)
        if od.trapSamples is not None:
            n_mc = int(od.trapSamples)
        key = (round(Rl, 12), round(Hl, 12), n_mc, od.trapSeed)
        if key not in grids:
            seed = od.trapSeed if od.trapSeed is not None else secrets.randbits(64)
            rho = np.linspace(0, Rl, n_r)
            h = np.linspace(0, Hl / 2.0, n_h)
            grids[key] = trap_potential_grid(Rl, Hl, rho, h, n_mc, seed, device)
        alphas.append(host.trap_fit(grids[key], Rl, Hl, od.N))
    return alphas


def prepare_batch(openDusts, device=0, alphas=None, precision=None):
    """Initial states and engine configurations of a batch (no GPU work except
    the trap Monte Carlo when `alphas` is None)."""
    if len({od.precision if precision is None else precision for od in openDusts}) != 1:
        raise ValueError("all simulations of a batch must share one precision")
    if any(od._cutoff_nm > 0 for od in openDusts):
        raise ValueError("batched sweeps run all-pairs forces (no cutoff)")
    if alphas is None:
        alphas = fit_traps(openDusts, device)
    cfgs, states = [], []
    for s, od in enumerate(openDusts):
        x0, v0 = od.initial_state()
        tables = od.reinjection_tables()
        od.alphaTrap = np.asarray(alphas[s], dtype=np.float64)
        if od.seed is None:
            od.seed = secrets.randbits(64)
        cfg = od.make_config(tables, od.alphaTrap, device=device, precision=precision, seed=od.seed)
        cfg.sim_id = s
        cfgs.append(cfg)
        states.append((x0, v0))
    return cfgs, states


def simulateBatch(openDusts, deviceIndex="0", useGraph=True, alphas=None, precision=None, tune=False,
                  sym_chunk=None, sym_tile=None, progress=None):
    """Run every OpenDust object of `openDusts` for its own n steps, together.
    Simulations may differ in ion count, dust count and positions, plasma
    parameters and geometry; they share precision and device.  The batch steps
    until the longest simulation ends; the extra steps of shorter ones are
    computed and discarded.  tune=True lets the autotuner pick the pair
    kernel's schedule and tile shape (od_batch_tune); the choice
    (sym_chunk, sym_tile) is returned, and results are reproducible for it
    (pass it back as sym_chunk and sym_tile).  progress(done, total) is called
    after every chunk of steps."""
    if len(openDusts) == 0:
        return
    dev_str = str(deviceIndex).strip()
    device = int(dev_str.split(",")[0]) if dev_str else 0
    cfgs, states = prepare_batch(openDusts, device, alphas, precision)
    if sym_chunk is not None:
        cfgs[0].sym_chunk = int(sym_chunk)
    if sym_tile is not None:
        cfgs[0].sym_tile = int(sym_tile)
    batch = Batch(cfgs)
    try:
        for s, (x0, v0) in enumerate(states):
            batch.set_ions(s, x0, v0)
        chosen = batch.tune() if tune else None
        n_max = max(od.n for od in openDusts)
        flags = RUN_GRAPH if useGraph else 0
        i = 0
        while i < n_max:
            k = min(_REC_CHUNK, n_max - i)
            rec = batch.run(k, flags)
            for s, od in enumerate(openDusts):
                if i >= od.n:
                    continue
                kk = min(k, od.n - i)
                od._record(i, rec[:kk, s, :od.numberOfDustParticles])
            i += k
            if progress is not None:
                progress(i, n_max)
    finally:
        batch.close()
    return chosen
