"""Autotuner (sn_autotune): the schedule it picks runs, keeps parity with the
oracle, and is served from the cache file on the next call."""
import os
import time

import numpy as np
import pytest

from tests.helpers import print_case, rand_lattice
from tests.test_gpu_sweep import run_pair

pytestmark = pytest.mark.gpu


def test_autotune_pick_keeps_parity_and_caches(sn, oracle, tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SN_TUNE_CACHE", str(cache))
    rp = sn.ReplicaParams(24, 24, 24, T=300, seed=77)
    print_case("autotune", X=24, budget_ms=1500)
    t0 = time.perf_counter()
    tun, rate = sn.autotune(rp, budget_ms=1500)
    first = time.perf_counter() - t0
    print(f"  picked {tun} at {rate:.3e} site-updates/s in {first:.2f} s", flush=True)
    assert rate > 0 and tun.kernel in (1, 2) and tun.fine_colouring in (1, 2, 3)
    assert cache.exists() and "24x24x24*1" in cache.read_text()
    t1 = time.perf_counter()
    tun2, rate2 = sn.autotune(rp, budget_ms=1500)
    assert tun2 == tun and rate2 == pytest.approx(rate, rel=1e-5)
    assert time.perf_counter() - t1 < first
    lat = rand_lattice(rp.N, np.random.default_rng(77), mixed=True)
    run_pair(sn, oracle, [rp], [lat], tun.__dict__, nsweeps=2)


def test_autotune_uneven_batch(sn, tmp_path, monkeypatch):
    monkeypatch.setenv("SN_TUNE_CACHE", str(tmp_path / "tune.txt"))
    reps = [sn.ReplicaParams(L, L, L, T=300, seed=L) for L in (16, 20, 24, 33)]
    print_case("autotune_batch", sizes=(16, 20, 24, 33))
    tun, rate = sn.autotune(reps, budget_ms=2000, use_cache=False)
    assert rate > 0
    with sn.StarryNight(reps, tun) as ctx:
        for r in range(len(reps)):
            ctx.init_lattice_random(r)
        ctx.MC_sweeps(2)
        cnt = ctx.counters(3)
    assert cnt["attempted"] == 2 * 33 ** 3
    assert not (tmp_path / "tune.txt").exists()


def test_autotune_species_table_keys_the_cache(sn, tmp_path, monkeypatch):
    """Trial lattices follow the species table: a mixed table (Dipoles [1, 0],
    Prevalence [.5, .5]) is tuned and cached apart from unit lengths, and its
    pick runs a mixed lattice."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("SN_TUNE_CACHE", str(cache))
    rp = sn.ReplicaParams(24, 24, 24, DipoleCutOff=4, T=300, seed=4)
    print_case("autotune_species", X=24, rc=4)
    tun_m, rate_m = sn.autotune(rp, budget_ms=1500, Dipoles=(1.0, 0.0), Prevalence=(0.5, 0.5))
    tun_u, rate_u = sn.autotune(rp, budget_ms=1500)
    text = cache.read_text()
    assert "|mixed|" in text and "|uniform|" in text and rate_m > 0 and rate_u > 0
    with pytest.raises(ValueError):
        sn.autotune(rp, Dipoles=(1.0, 0.0), Prevalence=(1.0,))
    with sn.StarryNight(rp, tun_m) as ctx:
        ctx.init_lattice_random(0, Dipoles=(1.0, 0.0), Prevalence=(0.5, 0.5))
        ctx.MC_sweeps(2)
        assert ctx.counters()["attempted"] == 2 * rp.N
