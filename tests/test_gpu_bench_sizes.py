"""Sweep parity at the sizes and schedules the benchmark reports.

The benchmark times these kernels on full-size lattices; here each one runs
one or two sweeps at that size and is compared with the oracle's ordered FP64
sweep (or_sweep_ordered_f64, the restatement of MC_move / MC_moves,
src/starrynight-montecarlo-core.c:143-191, in the library's reported visit
order with the same counter draws).  Bar: accept / reject counts exact,
lattice within 1e-12, lengths unchanged.  The visit order each context
reports is also checked conflict-free site by site on the host.

At 128^3 with 8^3 tiles a coarse-colour launch holds 512 tiles, so the
gather kernel's XCD tile remap (tile index = (block % 8) * (tiles / 8) +
block / 8) is a non-identity permutation; the smaller tests run at most 8
tiles per launch, where it reduces to the identity.
"""
import numpy as np
import pytest

from tests.helpers import oracle_model, print_case, rand_lattice
from tests.test_schedule import check_order

pytestmark = pytest.mark.gpu

# the schedule the benchmark's autotuner picks at 128^3 rc = 3 (8^3 tiles,
# BCC fine colouring, per-phase gather kernel, 256 threads: the fixed-stride
# 14^3-region instance)
HEADLINE_TUNING = dict(tile=(8, 8, 8), fine_colouring=3, kernel=2, block_threads=256)


def sweep_and_compare(sn, oracle, reps, lats, tuning, nsweeps, check=None, params_between=None):
    """Run nsweeps sweeps of every replica on the GPU, replay the replicas in
    `check` (default all) through the oracle, compare.  params_between[s]
    (optional) maps replica -> (T, E, J) applied before sweep s."""
    check = range(len(reps)) if check is None else check
    tun = sn.TuningParams(**tuning) if isinstance(tuning, dict) else tuning
    with sn.StarryNight(reps, tun) as ctx:
        cfg = ctx.sweep_config()
        print(f"  sweep config {cfg}", flush=True)
        for r, lat in enumerate(lats):
            ctx.set_lattice(lat, r)
        for s in range(nsweeps):
            for r, (T, E, J) in (params_between or {}).get(s, {}).items():
                ctx.set_params(r, T, E, J)
            ctx.MC_sweeps(1)
        got = {r: ctx.get_lattice(r) for r in check}
        cnt = {r: ctx.counters(r) for r in check}
        sched = {r: ctx.schedule(r) for r in check}
    for r in check:
        rp = reps[r]
        order, starts = sched[r]
        check_order(rp, order, starts)
        want = lats[r].copy()
        acc = rej = 0
        for s in range(nsweeps):
            T, E, J = (params_between or {}).get(s, {}).get(r, (rp.T, rp.Efield, rp.CageStrain))
            rp = sn.ReplicaParams(rp.X, rp.Y, rp.Z, DipoleCutOff=rp.DipoleCutOff, CageStrain=J, Efield=E,
                                  K=rp.K, ConstrainToX=rp.ConstrainToX, T=T, seed=rp.seed)
            a, b = oracle_model(oracle, rp).sweep_ordered(want, order, rp.seed, r, s)
            acc += a
            rej += b
        err = np.abs(got[r] - want).max()
        print(f"  replica {r} ({rp.X}x{rp.Y}x{rp.Z} T={rp.T}): max |dlat| {err:.3e}  "
              f"acc {cnt[r]['accept']}/{acc} rej {cnt[r]['reject']}/{rej}", flush=True)
        assert cnt[r]["accept"] == acc and cnt[r]["reject"] == rej
        assert cnt[r]["attempted"] == nsweeps * rp.N
        assert err <= 1e-12
        assert np.array_equal(got[r][:, 3], lats[r][:, 3])
    return cfg


def test_headline_128cube_rc3_benchmark_schedule(sn, oracle):
    """128^3, DipoleCutOff 3, T = 300 K, unit lengths: the benchmark's
    workload and schedule, two sweeps (sweep counter 0 and 1)."""
    rp = sn.ReplicaParams(128, 128, 128, DipoleCutOff=3, CageStrain=1.0, T=300, seed=0x5EED)
    print_case("headline_128cube_rc3", **HEADLINE_TUNING)
    lat = rand_lattice(rp.N, np.random.default_rng(128), mixed=False)
    cfg = sweep_and_compare(sn, oracle, [rp], [lat], HEADLINE_TUNING, nsweeps=2)
    assert cfg == {"tiles": 4096, "launches_per_sweep": 8, "lanes_per_site": 16, "fine_colours": 32}


def test_headline_128cube_rc3_default_tuning(sn, oracle):
    """The same lattice through sn_create with no tuning (the drop-in binding
    passes NULL), one sweep, field and E != 0."""
    rp = sn.ReplicaParams(128, 128, 128, DipoleCutOff=3, T=300, Efield=(0.0, 0.02, 0.0), seed=0x5EED + 1)
    print_case("headline_128cube_rc3_default")
    lat = rand_lattice(rp.N, np.random.default_rng(129), mixed=False)
    sweep_and_compare(sn, oracle, [rp], [lat], None, nsweeps=1)


def test_64cube_rc3_temperature_sweep_batched(sn, oracle):
    """64^3 rc = 3 at T = 0, 25 and 500 K as three replicas of one context
    (192 tiles per launch, non-identity XCD remap), two sweeps."""
    reps = [sn.ReplicaParams(64, 64, 64, DipoleCutOff=3, T=T, seed=640 + k) for k, T in enumerate((0, 25, 500))]
    print_case("64cube_rc3_T0_25_500", **HEADLINE_TUNING)
    rng = np.random.default_rng(64)
    lats = [rand_lattice(rp.N, rng, mixed=False) for rp in reps]
    cfg = sweep_and_compare(sn, oracle, reps, lats, HEADLINE_TUNING, nsweeps=2)
    assert cfg["tiles"] == 3 * 512


def batched_workload(sn, rc=3):
    """96 replicas of 32^3 (32 temperatures 50-500 K x E_z in {0, 0.01, 0.05})
    plus 32 replicas of 16^3..48^3 at 300 K: the batched benchmark workload."""
    reps, k = [], 0
    for i in range(32):
        T = 50.0 + 450.0 * i / 31.0
        for Ez in (0.0, 0.01, 0.05):
            reps.append(sn.ReplicaParams(32, 32, 32, DipoleCutOff=rc, T=T, Efield=(0, 0, Ez), seed=1000 + k))
            k += 1
    for i in range(32):
        Lk = 16 + round(32 * i / 31)
        reps.append(sn.ReplicaParams(Lk, Lk, Lk, DipoleCutOff=rc, T=300, seed=1000 + k))
        k += 1
    return reps


@pytest.mark.parametrize("how", ["autotuned", "default"])
def test_batched_128_replicas(sn, oracle, how):
    """The full 128-replica batch in one context, with the autotuner's pick
    for it and with the built-in schedule: every uneven replica and every
    fourth 32^3 replica are replayed through the oracle."""
    reps = batched_workload(sn)
    tun = sn.autotune(reps, budget_ms=3000, use_cache=False)[0] if how == "autotuned" else None
    print_case("batched_128", how=how, tuning=tun)
    rng = np.random.default_rng(4096)
    lats = [rand_lattice(rp.N, rng, mixed=False) for rp in reps]
    check = list(range(0, 96, 4)) + list(range(96, 128))
    sweep_and_compare(sn, oracle, reps, lats, tun, nsweeps=1, check=check)


def test_128cube_rc4_solid_solution_ramp_schedule(sn, oracle):
    """128^3, DipoleCutOff 4, Dipoles [1, 0] / Prevalence [.5, .5] (half the
    sites vacant), on the ramp's schedule (8x8x10 tiles, minimum-index
    colouring, gather kernel): E_z = -0.1 for the first sweep, -0.095 for the
    second (sn_set_params between sweeps)."""
    rp = sn.ReplicaParams(128, 128, 128, DipoleCutOff=4, T=300, Efield=(0, 0, -0.1), seed=4000)
    tuning = dict(tile=(8, 8, 10), fine_colouring=2, kernel=2, block_threads=256)
    print_case("128cube_rc4_solid_solution", **tuning)
    rng = np.random.default_rng(4000)
    lat = rand_lattice(rp.N, rng, mixed=False)
    lat[:, 3] = np.where(rng.random(rp.N) < 0.5, 1.0, 0.0)
    ramp = {0: {0: (300, (0.0, 0.0, -0.1), 1.0)}, 1: {0: (300, (0.0, 0.0, -0.095), 1.0)}}
    sweep_and_compare(sn, oracle, [rp], [lat], tuning, nsweeps=2, params_between=ramp)
