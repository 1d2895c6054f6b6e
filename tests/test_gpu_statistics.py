"""Statistical oracle: equilibrium averages of the GPU colour sweeps against
long CPU runs of the oracle's FP64 restatement of MC_moves (random site with
replacement, MT19937; tests/golden/stat_equilibrium.json, written by
tests/golden/make_stat_fixtures.py).

Both chains sample the Boltzmann distribution of the same Hamiltonian, so the
averages must agree.  Per state point the family of z-tests (z = difference /
the two standard errors combined) covers e = E/N, m = |sum p|/N, mz, chi and
the fe / afe radial-order bins.  Criterion, fixed before any result was seen:
the family passes when its smallest two-sided p-value exceeds alpha / m, with
family-wise alpha = 0.01 and m tests (the first step of Holm's step-down
procedure, which decides pass / fail of the whole family; it holds under any
correlation between the observables).  Under the null hypothesis a state
point fails with probability at most 1 %; with m = 22 the bar is |z| < 3.51.
The GPU side runs many replicas of the state point in one context.
"""
import json
import os
from math import erfc, sqrt

import numpy as np
import pytest

from tests.helpers import print_case
from tests.stat_observables import (RDF_BINS, STATES, beta_of, block_stats, initial_lattice,
                                    sample_scalars, species_lengths)

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "stat_equilibrium.json")
REPLICAS = 64
BURN = 2000
SWEEPS = 2000
SCALAR_EVERY = 2
RDF_EVERY = 20
NBLOCKS = 1   # each replica is one block (independent chains)
ALPHA = 0.01   # family-wise error rate of one state point's z-tests


def family_passes(zs: dict, alpha: float):
    """First step of Holm's step-down test of a family of two-sided z-tests:
    the family passes iff the smallest p-value exceeds alpha / m.  Returns
    (passes, worst test, its p-value, alpha / m)."""
    worst = max(zs, key=lambda k: abs(zs[k]))
    p_min = erfc(abs(zs[worst]) / sqrt(2.0))
    bar = alpha / len(zs)
    return p_min > bar, worst, p_min, bar


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
            ctx.set_lattice(initial_lattice(st, N, L, seed=5000 + r), r)
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
    zs = {}
    for k in ("e", "m", "mz", "chi"):
        g, gs = got[k]
        w, ws = want[k]
        zs[k] = (g - w) / np.hypot(gs, ws)
        print(f"  {k:4s} gpu {g:+.5f} +- {gs:.5f}  cpu {w:+.5f} +- {ws:.5f}  z {zs[k]:+.2f}", flush=True)
    for i, b in enumerate(RDF_BINS * 2):
        g, gs = got["rdf_mean"][i], got["rdf_se"][i]
        w, ws = want["rdf_mean"][i], want["rdf_se"][i]
        tag = f"{'fe' if i < len(RDF_BINS) else 'afe'}[r2={b}]"
        zs[tag] = (g - w) / np.hypot(gs, ws)
        print(f"  {tag} gpu {g:+.5f} +- {gs:.5f}  cpu {w:+.5f} +- {ws:.5f}  z {zs[tag]:+.2f}")
    ok, worst, p_min, bar = family_passes(zs, ALPHA)
    print(f"  family of {len(zs)} tests: worst {worst} z {zs[worst]:+.2f}, p {p_min:.2e} vs alpha/m {bar:.2e}",
          flush=True)
    assert ok, f"{worst}: z = {zs[worst]:+.2f} (p = {p_min:.1e} <= alpha/m = {bar:.1e})"
