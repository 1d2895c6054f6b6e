#!/usr/bin/env python3
"""Writes tests/golden/stat_equilibrium.json: equilibrium means and standard
errors of e, m, mz, chi and radial-order bins for the state points in
tests/stat_observables.py, from long CPU runs of the oracle's FP64
restatement of MC_moves (random site with replacement, MT19937 stream,
random sphere proposals, Metropolis; src/starrynight-montecarlo-core.c:143-191).

Each state point runs CHAINS independent chains (one process each) of
BURN sweeps of burn-in, then SWEEPS sweeps with the scalars sampled every
SCALAR_EVERY sweeps and the radial bins every RDF_EVERY sweeps.

    python tests/golden/make_stat_fixtures.py [--chains 8] [--sweeps 20000]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests.stat_observables import (RDF_BINS, STATES, beta_of, block_stats,  # noqa: E402
                                    initial_lattice, sample_scalars, species_lengths)

SCALAR_EVERY = 2
RDF_EVERY = 20
BURN = 2000
NBLOCKS = 5   # blocks of 3600 sweeps per chain


def run_chain(args):
    name, chain, sweeps = args
    from oracle import oracle as O
    st = STATES[name]
    X, Y, Z = st["X"], st["Y"], st["Z"]
    N = X * Y * Z
    beta = beta_of(st["T"])
    o = O.Oracle(X, Y, Z, st["rc"], J=st["J"], E=st["E"], K=0.0, constrain_to_x=False, beta=beta)
    L = species_lengths(N, st["lengths"], st["prevalence"], seed=7)
    lat = initial_lattice(st, N, L, seed=1000 + chain)
    mt = O.MT19937(0xC0FFEE + 7919 * chain)
    o.mc_moves_f64(lat, mt, BURN * N)
    sc, rd = [], []
    for s in range(1, sweeps + 1):
        o.mc_moves_f64(lat, mt, N)
        if s % SCALAR_EVERY == 0:
            e, _ = o.energy(lat)
            sc.append(sample_scalars(e, O.moment_sum(lat), N))
        if s % RDF_EVERY == 0:
            fe, afe, _ = o.radial_order(lat)
            rd.append(np.concatenate([fe[RDF_BINS], afe[RDF_BINS]]))
    return name, np.asarray(sc), np.asarray(rd)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=8)
    ap.add_argument("--sweeps", type=int, default=20000)
    ap.add_argument("--states", nargs="*", default=list(STATES))
    ap.add_argument("--out", default=os.path.join(HERE, "stat_equilibrium.json"))
    a = ap.parse_args()
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    jobs = [(n, c, a.sweeps) for n in a.states for c in range(a.chains)]
    t0 = time.time()
    res = {n: ([], []) for n in a.states}
    with ProcessPoolExecutor(min(len(jobs), os.cpu_count() or 1)) as ex:
        for name, sc, rd in ex.map(run_chain, jobs):
            res[name][0].append(sc)
            res[name][1].append(rd)
            print(f"  {name}: chain done ({time.time() - t0:.0f} s)", flush=True)
    out = {"method": "oracle or_mc_moves_f64 (random-site MT19937 Metropolis, FP64)",
           "burn_sweeps": BURN, "scalar_every": SCALAR_EVERY, "rdf_every": RDF_EVERY,
           "blocks_per_chain": NBLOCKS, "rdf_bins": RDF_BINS, "states": {}}
    if os.path.exists(a.out):
        out["states"] = json.load(open(a.out)).get("states", {})
    for name, (scs, rds) in res.items():
        st = STATES[name]
        N = st["X"] * st["Y"] * st["Z"]
        stats = block_stats(scs, rds, NBLOCKS, beta_of(st["T"]), N)
        stats.update(chains=len(scs), sweeps_per_chain=a.sweeps)
        out["states"][name] = stats
        print(name, {k: stats[k] for k in ("e", "m", "mz", "chi")}, flush=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
