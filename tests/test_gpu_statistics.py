"""Statistical oracle: equilibrium averages of the GPU colour sweeps against
long CPU runs of the oracle's FP64 restatement of MC_moves (random site with
replacement, MT19937; tests/golden/stat_equilibrium.json, written by
tests/golden/make_stat_fixtures.py).

Both chains sample the Boltzmann distribution of the same Hamiltonian, so the
averages must agree: e = E/N, m = |sum p|/N, mz and chi within 2 sigma
(sigma = the two standard errors combined), the radial-order bins within
3 sigma.  The GPU side runs many replicas of the state point in one context.
"""
import json
import os

import numpy as np
import pytest

from tests.helpers import print_case
from tests.stat_observables import (RDF_BINS, STATES, beta_of, block_stats, random_lattice,
                                    sample_scalars, species_lengths)

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "stat_equilibrium.json")
REPLICAS = 64
BURN = 2000
SWEEPS = 2000
SCALAR_EVERY = 2
RDF_EVERY = 20
NBLOCKS = 25


def _fixture():
    if not os.path.exists(FIX):
        pytest.skip("stat_equilibrium.json not generated")
    return json.load(open(FIX))


@pytest.mark.parametrize("name", list(STATES))
def test_equilibrium_averages_match_cpu_chains(sn, name):
    fx = _fixture()
    if name not in fx["states"]:
        pytest.skip(f"{name} not in the fixture")
    want = fx["states"][name]
    st = STATES[name]
    X, Y, Z = st["X"], st["Y"], st["Z"]
    N = X * Y * Z
    beta = beta_of(st["T"])
    print_case("statistics", name=name, replicas=REPLICAS, sweeps=SWEEPS)
    L = species_lengths(N, st["lengths"], st["prevalence"], seed=7)
    reps = [sn.ReplicaParams(X, Y, Z, DipoleCutOff=st["rc"], CageStrain=st["J"], Efield=st["E"],
                             T=st["T"], seed=0xABC000 + r) for r in range(REPLICAS)]
    sc = [[] for _ in range(REPLICAS)]
    rd = [[] for _ in range(REPLICAS)]
    with sn.StarryNight(reps) as ctx:
        for r in range(REPLICAS):
            ctx.set_lattice(random_lattice(N, L, seed=5000 + r), r)
        ctx.MC_sweeps(BURN)
        for s in range(SCALAR_EVERY, SWEEPS + 1, SCALAR_EVERY):
            ctx.MC_sweeps(SCALAR_EVERY)
            for r in range(REPLICAS):
                e, _ = ctx.energy(r)
                sc[r].append(sample_scalars(e, ctx.moment(r), N))
                if s % RDF_EVERY == 0:
                    fe, afe, _ = ctx.radial_order(r)
                    rd[r].append(np.concatenate([fe[RDF_BINS], afe[RDF_BINS]]))
        cnt = ctx.counters(0)
    got = block_stats([np.asarray(x) for x in sc], [np.asarray(x) for x in rd], NBLOCKS, beta, N)
    print(f"  acceptance (replica 0) {cnt['accept'] / max(1, cnt['accept'] + cnt['reject']):.3f}")
    bad = []
    for k in ("e", "m", "mz", "chi"):
        g, gs = got[k]
        w, ws = want[k]
        sig = np.hypot(gs, ws)
        z = (g - w) / sig
        print(f"  {k:4s} gpu {g:+.5f} +- {gs:.5f}  cpu {w:+.5f} +- {ws:.5f}  z {z:+.2f}", flush=True)
        if abs(z) > 2.0:
            bad.append(k)
    for i, b in enumerate(RDF_BINS * 2):
        g, gs = got["rdf_mean"][i], got["rdf_se"][i]
        w, ws = want["rdf_mean"][i], want["rdf_se"][i]
        z = (g - w) / np.hypot(gs, ws)
        tag = "fe" if i < len(RDF_BINS) else "afe"
        print(f"  {tag}[r2={b}] gpu {g:+.5f} +- {gs:.5f}  cpu {w:+.5f} +- {ws:.5f}  z {z:+.2f}")
        if abs(z) > 3.0:
            bad.append(f"{tag}{b}")
    assert not bad, f"outside the bar: {bad}"
