"""The statistical criterion of test_gpu_statistics.py, exercised on the CPU
with the oracle's own chains: short FP64 random-site chains at the fixture's
state point pass it, and the same chains with a planted mistake (the
temperature 1 % off, the cage strain dropped) fail it.  This shows the
criterion can tell a wrong chain from a right one at the GPU test's sample
size, before it is trusted on the GPU."""
import json
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np
import pytest

from tests.stat_observables import (RDF_BINS, STATES, beta_of, block_stats, random_lattice,
                                    sample_scalars, species_lengths)
from tests.test_gpu_statistics import ALPHA, FIX, family_passes

STATE = "cube12_T600"
CHAINS, BURN, SWEEPS, EVERY, RDF_EVERY, NBLOCKS = 8, 600, 1500, 2, 20, 25


def chain(args):
    k, T, J = args
    from oracle import oracle as O
    st = STATES[STATE]
    X, Y, Z = st["X"], st["Y"], st["Z"]
    N = X * Y * Z
    o = O.Oracle(X, Y, Z, st["rc"], J=J, E=st["E"], beta=beta_of(T))
    lat = random_lattice(N, species_lengths(N, st["lengths"], st["prevalence"], seed=7), seed=9000 + k)
    mt = O.MT19937(0xBADC0DE + 31 * k)
    o.mc_moves_f64(lat, mt, BURN * N)
    sc, rd = [], []
    for s in range(1, SWEEPS + 1):
        o.mc_moves_f64(lat, mt, N)
        if s % EVERY == 0:
            sc.append(sample_scalars(o.energy(lat)[0], O.moment_sum(lat), N))
        if s % RDF_EVERY == 0:
            fe, afe, _ = o.radial_order(lat)
            rd.append(np.concatenate([fe[RDF_BINS], afe[RDF_BINS]]))
    return np.asarray(sc), np.asarray(rd)


def zscores(T, J):
    st = STATES[STATE]
    want = json.load(open(FIX))["states"][STATE]
    with ProcessPoolExecutor(min(CHAINS, os.cpu_count() or 1)) as ex:
        res = list(ex.map(chain, [(k, T, J) for k in range(CHAINS)]))
    got = block_stats([r[0] for r in res], [r[1] for r in res], NBLOCKS, beta_of(st["T"]),
                      st["X"] * st["Y"] * st["Z"])
    zs = {k: (got[k][0] - want[k][0]) / np.hypot(got[k][1], want[k][1]) for k in ("e", "m", "mz", "chi")}
    for i in range(2 * len(RDF_BINS)):
        zs[f"rdf{i}"] = (got["rdf_mean"][i] - want["rdf_mean"][i]) / np.hypot(got["rdf_se"][i], want["rdf_se"][i])
    return zs


@pytest.mark.parametrize("T,J,should_pass", [(600.0, 1.0, True), (606.0, 1.0, False), (600.0, 0.0, False)],
                         ids=["right_chain", "temperature_1pc_off", "cage_strain_dropped"])
def test_criterion_separates_right_and_planted_chains(oracle, T, J, should_pass):
    if not os.path.exists(FIX):
        pytest.skip("stat_equilibrium.json not generated")
    zs = zscores(T, J)
    ok, worst, p, bar = family_passes(zs, ALPHA)
    print(f"T={T} J={J}: worst {worst} z {zs[worst]:+.2f} p {p:.1e} bar {bar:.1e}")
    assert ok == should_pass
