"""Colour-sweep kernel (sn_mc_sweeps) vs the oracle's ordered counter-RNG sweep.

The library reports its visit order (sn_schedule_order); the oracle replays
the same order with the same Philox draws in FP64.  Bar: identical accept /
reject counts and lattices equal to 1e-12 (the proposals are bit-reproducible;
only the order of the dE sum differs).
"""
import numpy as np
import pytest

from tests.helpers import oracle_model, print_case, rand_lattice

pytestmark = pytest.mark.gpu

CASES = {
    # the library's own choice (kernel 0): gather tiles at 16^3, whole-lattice at 12^3 and 25x25
    "16cube_auto": (dict(X=16, Y=16, Z=16, T=300, Efield=(0.0, 0.02, 0.0)), True, {}),
    "12cube_auto": (dict(X=12, Y=12, Z=12, T=300), False, {}),
    "25x25x1_auto": (dict(X=25, Y=25, Z=1, T=300, K=0.1), True, {}),
    "16cube_rc3_uniform": (dict(X=16, Y=16, Z=16, T=300), False, dict(kernel=1)),
    "16cube_rc3_mixed_field_K": (dict(X=16, Y=16, Z=16, T=250, Efield=(0.05, 0.0, 0.1), K=0.3), True, dict(kernel=1)),
    "25x25x1_rc3": (dict(X=25, Y=25, Z=1, T=300), False, dict(kernel=1)),
    "20cube_rc4_mixed": (dict(X=20, Y=20, Z=20, DipoleCutOff=4, T=300), True, dict(kernel=1)),
    "12x16x20_rc3": (dict(X=12, Y=16, Z=20, T=400, CageStrain=2.0), False, dict(kernel=1)),
    "16cube_constrain_x": (dict(X=16, Y=16, Z=16, T=300, ConstrainToX=True), False, dict(kernel=1)),
    "16cube_T0": (dict(X=16, Y=16, Z=16, T=0), False, dict(kernel=1)),
    "16cube_lanes8_minidx": (dict(X=16, Y=16, Z=16, T=300), True, dict(kernel=1, lanes_per_site=8, fine_colouring=2)),
    "16cube_lanes32_cubic": (dict(X=16, Y=16, Z=16, T=300), False, dict(kernel=1, lanes_per_site=32, fine_colouring=1)),
    "18cube_direct_launch": (dict(X=18, Y=18, Z=18, T=300), False, dict(kernel=1, use_graph=-1)),
    "16cube_gather_kernel": (dict(X=16, Y=16, Z=16, T=300, Efield=(0, 0.05, 0)), True, dict(kernel=2)),
    "25x25x1_gather_kernel": (dict(X=25, Y=25, Z=1, T=300), False, dict(kernel=2)),
    "20cube_rc4_gather_kernel": (dict(X=20, Y=20, Z=20, DipoleCutOff=4, T=300), True, dict(kernel=2)),
    "odd_tiles_17x19x21": (dict(X=17, Y=19, Z=21, T=350), True, dict(kernel=1)),
    "2d_mixed_30x26": (dict(X=30, Y=26, Z=1, T=200, K=0.2), True, dict(kernel=1)),
    "rc2_12cube": (dict(X=12, Y=12, Z=12, DipoleCutOff=2, T=300), True, dict(kernel=1)),
    "rc5_24cube": (dict(X=24, Y=24, Z=24, DipoleCutOff=5, T=300), False, dict(kernel=1)),
    "16cube_gather_512": (dict(X=16, Y=16, Z=16, T=300, K=0.2), True, dict(kernel=2, block_threads=512)),
    "20cube_rc4_gather_512": (dict(X=20, Y=20, Z=20, DipoleCutOff=4, T=300), True,
                              dict(kernel=2, block_threads=512)),
    "25x25x1_gather_512": (dict(X=25, Y=25, Z=1, T=300), False, dict(kernel=2, block_threads=512)),
    "18x20x22_gather_512_cubic": (dict(X=18, Y=20, Z=22, T=350), False,
                                  dict(kernel=2, block_threads=512, fine_colouring=1)),
    # staged region of exactly 14^3 sites: the compile-time-stride gather instances
    "16cube_gather_region14_uniform": (dict(X=16, Y=16, Z=16, T=300, K=0.1, Efield=(0.0, 0.0, 0.05)), False,
                                       dict(kernel=2, tile=(8, 8, 8))),
    "16cube_gather_region14_lanes8": (dict(X=16, Y=16, Z=16, T=300), True,
                                      dict(kernel=2, tile=(8, 8, 8), lanes_per_site=8)),
    "20cube_rc2_gather_region14": (dict(X=20, Y=20, Z=20, DipoleCutOff=2, T=300), True,
                                   dict(kernel=2, tile=(10, 10, 10), lanes_per_site=16)),
    # the gather kernel (the library's choice above 4096 sites) on the field kernel's cases
    "16cube_gather_constrain_x": (dict(X=16, Y=16, Z=16, T=300, ConstrainToX=True), True, dict(kernel=2)),
    "16cube_gather_T0": (dict(X=16, Y=16, Z=16, T=0), True, dict(kernel=2)),
    "odd_tiles_17x19x21_gather": (dict(X=17, Y=19, Z=21, T=350, Efield=(0.03, 0.0, 0.0)), True, dict(kernel=2)),
    "12cube_rc1_gather": (dict(X=12, Y=12, Z=12, DipoleCutOff=1, T=300), True, dict(kernel=2)),
    "20cube_rc2_gather_512": (dict(X=20, Y=20, Z=20, DipoleCutOff=2, T=300, K=0.1), True,
                              dict(kernel=2, block_threads=512)),
    "24cube_rc5_gather": (dict(X=24, Y=24, Z=24, DipoleCutOff=5, T=300), True, dict(kernel=2)),
    "24cube_rc5_gather_512": (dict(X=24, Y=24, Z=24, DipoleCutOff=5, T=300), False,
                              dict(kernel=2, block_threads=512)),
    "12cube_rc1_gather_512": (dict(X=12, Y=12, Z=12, DipoleCutOff=1, T=300, Efield=(0, 0, 0.05)), False,
                              dict(kernel=2, block_threads=512)),
    # exact pair counts per lane: 4 lanes x 1 pair (3 pairs), 64 lanes x 5 pairs (257 pairs)
    "12cube_rc1_gather_256": (dict(X=12, Y=12, Z=12, DipoleCutOff=1, T=300, K=0.1), True,
                              dict(kernel=2, block_threads=256)),
    "24cube_rc5_gather_256_lanes64": (dict(X=24, Y=24, Z=24, DipoleCutOff=5, T=300, Efield=(0.0, 0.02, 0.0)), True,
                                      dict(kernel=2, block_threads=256, lanes_per_site=64)),
    "2d_mixed_30x26_gather": (dict(X=30, Y=26, Z=1, T=200, K=0.2, Efield=(0.02, -0.01, 0.0)), True,
                              dict(kernel=2)),
    "12x16x20_gather_J2": (dict(X=12, Y=16, Z=20, T=400, CageStrain=2.0), True, dict(kernel=2)),
    # 8-site tiles at DipoleCutOff 4: 16-site staged z rows, padded to 17 in LDS
    "32cube_rc4_gather_padded_rows": (dict(X=32, Y=32, Z=32, DipoleCutOff=4, T=300, Efield=(0.0, 0.0, 0.03)), True,
                                      dict(kernel=2, tile=(8, 8, 8))),
    "32cube_rc4_gather_padded_rows_512": (dict(X=32, Y=32, Z=32, DipoleCutOff=4, T=250, K=0.1), True,
                                          dict(kernel=2, tile=(8, 8, 8), block_threads=512)),
    # whole-lattice sweeps (kernel 3): one workgroup per replica, periodic colouring
    "25x25x1_torus": (dict(X=25, Y=25, Z=1, T=300), False, dict(kernel=3)),
    "16cube_torus_mixed_field_K": (dict(X=16, Y=16, Z=16, T=250, Efield=(0.05, 0.0, 0.1), K=0.3), True,
                                   dict(kernel=3)),
    "12cube_rc2_torus": (dict(X=12, Y=12, Z=12, DipoleCutOff=2, T=300), True, dict(kernel=3)),
    "16cube_rc4_torus": (dict(X=16, Y=16, Z=16, DipoleCutOff=4, T=300), True, dict(kernel=3)),
    "32x24x1_torus_mixed_K": (dict(X=32, Y=24, Z=1, T=200, K=0.2), True, dict(kernel=3)),
    "16cube_torus_constrain_x": (dict(X=16, Y=16, Z=16, T=300, ConstrainToX=True), False, dict(kernel=3)),
    "25x25x1_torus_lanes16": (dict(X=25, Y=25, Z=1, T=300, Efield=(0.0, 0.02, 0.0)), True,
                              dict(kernel=3, lanes_per_site=16)),
    "16cube_torus_lanes32": (dict(X=16, Y=16, Z=16, T=300), True, dict(kernel=3, lanes_per_site=32)),
    # greedy periodic colourings (fewer colours than any sublattice of these sizes)
    "30x26x1_torus_greedy_mixed_K": (dict(X=30, Y=26, Z=1, T=250, K=0.2, Efield=(0.02, 0.0, 0.0)), True,
                                     dict(kernel=3)),
    "11x13x1_torus_greedy": (dict(X=11, Y=13, Z=1, T=300), False, dict(kernel=3)),
    "25x25x1_torus_lanes8": (dict(X=25, Y=25, Z=1, T=300), True, dict(kernel=3, lanes_per_site=8)),
    "12cube_torus_T0_field": (dict(X=12, Y=12, Z=12, T=0, Efield=(0.0, 0.0, -0.05)), True, dict(kernel=3)),
    "10cube_rc1_torus": (dict(X=10, Y=10, Z=10, DipoleCutOff=1, T=300, K=0.1), True, dict(kernel=3)),
    "24x20x1_rc5_torus": (dict(X=24, Y=20, Z=1, DipoleCutOff=5, T=300), True, dict(kernel=3)),
    "12cube_rc1_auto": (dict(X=12, Y=12, Z=12, DipoleCutOff=1, T=300), True, {}),
    # requested lanes whose one-round thread count has no instance run several rounds
    "12cube_rc2_torus_lanes4": (dict(X=12, Y=12, Z=12, DipoleCutOff=2, T=300), True,
                                dict(kernel=3, lanes_per_site=4)),
    # the library's choice above 4096 sites: DipoleCutOff 4 runs 512 threads
    # (one 256-thread workgroup's region per CU), 5 runs 64 lanes x 5 pairs
    "20cube_rc4_auto": (dict(X=20, Y=20, Z=20, DipoleCutOff=4, T=300, Efield=(0.0, 0.0, 0.04)), True, {}),
    # 2D gather tiles at every cutoff: 4 x 1 (rc 1), 16 x 1 (rc 3), 16 x 2 (rc 4), ...
    "2d_72x64_rc1_gather": (dict(X=72, Y=64, Z=1, DipoleCutOff=1, T=300, K=0.1), True, dict(kernel=2)),
    "2d_72x64_rc2_gather": (dict(X=72, Y=64, Z=1, DipoleCutOff=2, T=300), False, dict(kernel=2)),
    "2d_72x64_rc4_gather": (dict(X=72, Y=64, Z=1, DipoleCutOff=4, T=250, Efield=(0.01, 0.0, 0.0)), True,
                            dict(kernel=2)),
    "2d_80x80_rc5_gather": (dict(X=80, Y=80, Z=1, DipoleCutOff=5, T=300), True, dict(kernel=2)),
    "2d_96x80_rc3_auto": (dict(X=96, Y=80, Z=1, T=300), True, {}),
    "20cube_rc5_auto": (dict(X=20, Y=20, Z=20, DipoleCutOff=5, T=300, K=0.1), True, {}),
    "18x18x18_rc3_auto": (dict(X=18, Y=18, Z=18, T=300), True, {}),
}


def test_untileable_lattice_rejects_sweeps_but_gathers(sn, oracle):
    """16^3 at DipoleCutOff=5 has no tiling with an even tile count per axis
    and tile >= cut-off: sweeps must raise, site energies must still work."""
    rp = sn.ReplicaParams(16, 16, 16, DipoleCutOff=5, T=300, seed=3)
    print_case("untileable", X=16, rc=5)
    lat = rand_lattice(rp.N, np.random.default_rng(3), mixed=False)
    with sn.StarryNight([rp]) as ctx:
        ctx.set_lattice(lat, 0)
        with pytest.raises(sn.SnError):
            ctx.MC_sweeps(1)
        sites = np.arange(0, rp.N, 97, dtype=np.int32)
        props = rand_lattice(len(sites), np.random.default_rng(4), mixed=False)[:, :3].copy()
        dE = ctx.site_energy_batch(sites, props, 0)
    want, _ = oracle_model(oracle, rp).site_energy_batch(lat, sites, props)
    assert np.allclose(dE, want, rtol=0, atol=1e-10 * max(1.0, np.abs(want).max()))


def test_cutoff_beyond_tiles_creates_and_gathers(sn, oracle):
    """DipoleCutOff = 6 in 3D: no tile + halo fits the LDS region limit, so
    the context is created without sweeps; site energies still work."""
    rp = sn.ReplicaParams(20, 20, 20, DipoleCutOff=6, T=300, seed=6)
    print_case("rc6_no_sweeps", X=20, rc=6)
    lat = rand_lattice(rp.N, np.random.default_rng(6), mixed=True)
    with sn.StarryNight([rp]) as ctx:
        ctx.set_lattice(lat, 0)
        with pytest.raises(sn.SnError) as e:
            ctx.MC_sweeps(1)
        assert e.value.status == 5
        sites = np.arange(0, rp.N, 61, dtype=np.int32)
        props = rand_lattice(len(sites), np.random.default_rng(7), mixed=False)[:, :3].copy()
        dE = ctx.site_energy_batch(sites, props, 0)
    want, _ = oracle_model(oracle, rp).site_energy_batch(lat, sites, props)
    assert np.allclose(dE, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()) * 100)


def run_pair(sn, oracle, reps, lats, tuning, nsweeps):
    tun = sn.TuningParams(**tuning) if tuning else None
    with sn.StarryNight(reps, tun) as ctx:
        for r, lat in enumerate(lats):
            ctx.set_lattice(lat, r)
        ctx.MC_sweeps(nsweeps)
        got = [ctx.get_lattice(r) for r in range(len(reps))]
        cnt = [ctx.counters(r) for r in range(len(reps))]
        orders = [ctx.schedule(r)[0] for r in range(len(reps))]
    for r, rp in enumerate(reps):
        o = oracle_model(oracle, rp)
        want = lats[r].copy()
        acc = rej = 0
        for s in range(nsweeps):
            a, b = o.sweep_ordered(want, orders[r], rp.seed, r, s)
            acc += a; rej += b
        err = np.abs(got[r] - want).max()
        print(f"  replica {r}: max |dlat| {err:.3e}  acc {cnt[r]['accept']}/{acc} rej {cnt[r]['reject']}/{rej}",
              flush=True)
        assert cnt[r]["accept"] == acc and cnt[r]["reject"] == rej
        assert cnt[r]["attempted"] == nsweeps * rp.N
        assert err <= 1e-12
        assert np.array_equal(got[r][:, 3], lats[r][:, 3])


@pytest.mark.parametrize("name", list(CASES))
def test_sweep_parity(sn, oracle, name):
    kw, mixed, tuning = CASES[name]
    rp = sn.ReplicaParams(seed=0x5EED0000 + len(name), **kw)
    print_case("sweep", name=name, **kw)
    lat = rand_lattice(rp.N, np.random.default_rng(len(name)), mixed=mixed)
    run_pair(sn, oracle, [rp], [lat], tuning, nsweeps=3)


def test_batched_uneven_replicas_parity(sn, oracle):
    """Replicas of different sizes, temperatures and fields in one context:
    each follows exactly the trajectory the oracle gives it alone."""
    reps = [sn.ReplicaParams(16, 16, 16, T=100, seed=11),
            sn.ReplicaParams(18, 18, 18, T=300, Efield=(0, 0, 0.05), seed=12),
            sn.ReplicaParams(24, 20, 17, T=500, seed=13),
            sn.ReplicaParams(16, 16, 16, T=0, seed=14)]
    print_case("batched_uneven", n=len(reps))
    rng = np.random.default_rng(7)
    lats = [rand_lattice(r.N, rng, mixed=(i == 1)) for i, r in enumerate(reps)]
    run_pair(sn, oracle, reps, lats, {}, nsweeps=2)


def test_batched_equals_single(sn):
    """With tiles, a replica's trajectory does not depend on what else is
    batched with it.  (Whole-lattice sweeps colour a batch with one
    sublattice that holds every box, so their order may change with the
    batch; the library's own choice for a lone 16^3 replica is such sweeps,
    hence the pinned kernel.)"""
    a = sn.ReplicaParams(16, 16, 16, T=300, seed=99)
    b = sn.ReplicaParams(20, 20, 20, T=150, seed=98)
    lat = rand_lattice(a.N, np.random.default_rng(3))
    tiles = sn.TuningParams(kernel=2)
    with sn.StarryNight([a], tiles) as c1:
        c1.set_lattice(lat); c1.MC_sweeps(5); one = c1.get_lattice()
    with sn.StarryNight([b, a], tiles) as c2:
        c2.set_lattice(rand_lattice(b.N, np.random.default_rng(4)), 0)
        c2.set_lattice(lat, 1); c2.MC_sweeps(5); two = c2.get_lattice(1)
    # same schedule for `a`, replica index differs -> different RNG stream unless
    # the replica slot matches; compare with the replica in slot 0 instead
    with sn.StarryNight([a, b], tiles) as c3:
        c3.set_lattice(lat, 0); c3.set_lattice(rand_lattice(b.N, np.random.default_rng(4)), 1)
        c3.MC_sweeps(5); three = c3.get_lattice(0)
    assert np.array_equal(one, three)
    assert not np.array_equal(one, two)   # slot 1 draws a different stream


def test_graph_replay_matches_direct_launch(sn):
    rp = sn.ReplicaParams(16, 16, 16, T=300, seed=5)
    lat = rand_lattice(rp.N, np.random.default_rng(5), mixed=True)
    out = []
    for g in (0, -1):
        with sn.StarryNight(rp, sn.TuningParams(use_graph=g)) as ctx:
            ctx.set_lattice(lat)
            for _ in range(3):   # several replays with host work in between
                ctx.MC_sweeps(2)
                _ = ctx.counters()
            out.append(ctx.get_lattice())
    assert np.array_equal(out[0], out[1])


def test_set_params_between_sweeps(sn, oracle):
    """Temperature / field ramps through sn_set_params reach the captured graph."""
    rp = sn.ReplicaParams(16, 16, 16, T=300, seed=21)
    lat = rand_lattice(rp.N, np.random.default_rng(21))
    ramp = [(300, (0, 0, 0.0)), (200, (0, 0, 0.05)), (100, (0, 0, 0.1))]
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        order = ctx.schedule()[0]
        for T, E in ramp:
            ctx.set_params(0, T, E, 1.0)
            ctx.MC_sweeps(1)
        got = ctx.get_lattice()
    want = lat.copy()
    for s, (T, E) in enumerate(ramp):
        o = oracle_model(oracle, sn.ReplicaParams(16, 16, 16, T=T, Efield=E))
        o.sweep_ordered(want, order, rp.seed, 0, s)
    assert np.abs(got - want).max() <= 1e-12


def test_device_random_init(sn):
    rp = sn.ReplicaParams(32, 32, 32, seed=77)
    with sn.StarryNight(rp) as ctx:
        ctx.init_lattice_random(0, Dipoles=(1.0, 0.0), Prevalence=(0.5, 0.5))
        lat = ctx.get_lattice()
    assert np.abs(np.linalg.norm(lat[:, :3], axis=1) - 1).max() < 1e-15
    frac = (lat[:, 3] == 1.0).mean()
    assert abs(frac - 0.5) < 0.02 and set(np.unique(lat[:, 3])) == {0.0, 1.0}
    assert np.abs(lat[:, :3].mean(0)).max() < 0.02


def test_untileable_lattice_refuses_sweeps_but_gathers_work(sn, oracle):
    rp = sn.ReplicaParams(9, 9, 9, T=300)
    lat = rand_lattice(rp.N, np.random.default_rng(9))
    with sn.StarryNight(rp, sn.TuningParams(kernel=1)) as ctx:
        ctx.set_lattice(lat)
        with pytest.raises(sn.SnError) as e:
            ctx.MC_sweeps(1)
        assert e.value.status == 5
        dE = ctx.site_energy_batch(np.arange(10, dtype=np.int32), lat[10:20, :3])
    want, _ = oracle_model(oracle, rp).site_energy_batch(lat, np.arange(10), lat[10:20, :3])
    assert np.abs(dE - want).max() < 1e-12 * 100


def test_default_kernel_choice(sn):
    """Kernel 0 (and no tuning at all): whole-lattice sweeps for replicas of
    up to 4096 sites that fit in LDS, tiles with the gather kernel above that
    (at 512 threads for DipoleCutOff 1).  9^3 has no tiling at rc = 3 (tile
    kernels refuse it) but a periodic colouring; 16^3 at DipoleCutOff 2 has
    colours too large for the proposal slots of a whole-lattice sweep."""
    cases = [([sn.ReplicaParams(12, 12, 12, T=300)], True),
             ([sn.ReplicaParams(9, 9, 9, T=300)], True),
             ([sn.ReplicaParams(25, 25, 1, T=300)], True),
             ([sn.ReplicaParams(16, 16, 16, T=300)], True),
             ([sn.ReplicaParams(16, 16, 16, DipoleCutOff=2, T=300)], False),
             ([sn.ReplicaParams(20, 20, 20, T=300)], False),
             ([sn.ReplicaParams(12, 12, 12, T=300), sn.ReplicaParams(20, 20, 20, T=300)], False)]
    for reps, whole in cases:
        for tun in (None, sn.TuningParams()):
            with sn.StarryNight(reps, tun) as ctx:
                cfg = ctx.sweep_config()
                assert (cfg["tiles"] == 0) == whole, (reps[0].X, cfg)
                if not whole and reps[-1].DipoleCutOff == 3:
                    assert cfg["lanes_per_site"] == 16
    # the whole choice (kernel, threads, lanes x pairs) for NULL tuning
    R = sn.ReplicaParams
    for reps, want in (([R(128, 128, 128, T=300)], (2, 256, 16, 4)),
                       ([R(128, 128, 128, DipoleCutOff=4, T=300)], (2, 512, 32, 4)),
                       ([R(32, 32, 32, DipoleCutOff=5, T=300)], (2, 256, 64, 5)),
                       ([R(32, 32, 32, DipoleCutOff=1, T=300)], (2, 512, 4, 1)),
                       ([R(64, 64, 64, DipoleCutOff=2, T=300)], (2, 256, 4, 4)),
                       ([R(256, 256, 1, T=300)], (2, 256, 8, 2)),
                       ([R(12, 12, 12, T=300)], (3, 256, 16, 4)),
                       ([R(25, 25, 1, T=300)], (3, 512, 8, 2))):
        with sn.StarryNight(reps) as ctx:
            kc = ctx.kernel_choice()
            got = (kc["kernel"], kc["block_threads"], kc["lanes_per_site"], kc["pairs_per_lane"])
            assert got == want, (reps[0], got)
    with sn.StarryNight(R(32, 32, 32, T=300), sn.TuningParams(kernel=1)) as ctx:
        assert ctx.kernel_choice() == {"kernel": 1, "block_threads": 256, "lanes_per_site": 16, "pairs_per_lane": 0}
    # DipoleCutOff 1: 4 lanes a site (3 pairs); 5: 64 lanes a site (257
    # pairs, 5 a lane); kernel 1 reports the field kernel's 16 lanes
    for rc, tun, lanes in ((1, None, 4), (1, sn.TuningParams(kernel=2, block_threads=256), 4),
                           (5, None, 64), (5, sn.TuningParams(kernel=2), 64),
                           (5, sn.TuningParams(kernel=2, block_threads=256), 64),
                           (5, sn.TuningParams(kernel=1), 16)):
        with sn.StarryNight(sn.ReplicaParams(32, 32, 32, DipoleCutOff=rc, T=300), tun) as ctx:
            cfg = ctx.sweep_config()
            assert cfg["tiles"] > 0 and cfg["lanes_per_site"] == lanes, (rc, tun, cfg)


def test_field_and_gather_kernels_agree(sn):
    """The field-update and per-phase gather kernels visit the same sites with
    the same draws: after several sweeps the lattices match."""
    rp = sn.ReplicaParams(24, 24, 24, T=300, Efield=(0.02, 0, 0), seed=31)
    lat = rand_lattice(rp.N, np.random.default_rng(31), mixed=True)
    out = []
    for k in (1, 2):
        with sn.StarryNight(rp, sn.TuningParams(kernel=k)) as ctx:
            ctx.set_lattice(lat)
            ctx.MC_sweeps(6)
            out.append((ctx.get_lattice(), ctx.counters()))
    assert np.abs(out[0][0] - out[1][0]).max() <= 1e-12
    assert out[0][1] == out[1][1]


def test_multi_sweep_graph_matches_single_sweeps(sn):
    """MC_sweeps(n) replays a 16-sweep graph n // 16 times and a one-sweep graph
    for the rest: 37 sweeps at once equal 37 single-sweep calls, bit for bit,
    counters included."""
    rp = sn.ReplicaParams(16, 16, 16, T=350, Efield=(0.0, 0.01, 0.0), seed=37)
    print_case("multi_sweep_graph", X=16, sweeps=37)
    lat = rand_lattice(rp.N, np.random.default_rng(37), mixed=True)
    out = []
    for chunks in ([37], [1] * 37):
        with sn.StarryNight(rp) as ctx:
            ctx.set_lattice(lat)
            for n in chunks:
                ctx.MC_sweeps(n)
            out.append((ctx.get_lattice(), ctx.counters()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and out[0][1]["attempted"] == 37 * rp.N


def test_torus_batch_and_limits(sn, oracle):
    """Whole-lattice sweeps of a batch of small replicas (one workgroup each,
    one periodic colouring for all), against the oracle; sizes with no periodic
    colouring, or too large for LDS, are refused at creation."""
    reps = [sn.ReplicaParams(L, L, L, T=250 + 20 * i, seed=70 + i) for i, L in enumerate((16, 12, 8))]
    print_case("torus_batch", sizes=(16, 12, 8))
    rng = np.random.default_rng(12)
    lats = [rand_lattice(rp.N, rng, mixed=True) for rp in reps]
    run_pair(sn, oracle, reps, lats, dict(kernel=3), nsweeps=3)
    for bad in (dict(X=17, Y=19, Z=21), dict(X=24, Y=24, Z=24)):
        with pytest.raises(sn.SnError) as e:
            sn.StarryNight(sn.ReplicaParams(T=300, **bad), sn.TuningParams(kernel=3))
        assert e.value.status == 5


def test_torus_greedy_colouring(sn):
    """25 x 25 at DipoleCutOff 3: every sublattice colouring holding the box has
    25 colours (its index divides 625); the greedy colouring needs at most 16,
    and no two sites of a colour are within the cutoff through the periodic
    boundary."""
    rp = sn.ReplicaParams(25, 25, 1, T=300)
    print_case("torus_greedy_colouring", X=25, Y=25, Z=1)
    with sn.StarryNight(rp, sn.TuningParams(kernel=3)) as ctx:
        order, starts = ctx.schedule(0)
        cfg = ctx.sweep_config()
    order, starts = np.asarray(order), np.asarray(starts)
    assert cfg["fine_colours"] <= 16 and len(starts) - 1 == cfg["fine_colours"]
    assert np.array_equal(np.sort(order), np.arange(625))
    x, y = order // 25, order % 25
    for k in range(len(starts) - 1):
        cx, cy = x[starts[k]:starts[k + 1]], y[starts[k]:starts[k + 1]]
        dx = np.abs(cx[:, None] - cx[None, :]); dx = np.minimum(dx, 25 - dx)
        dy = np.abs(cy[:, None] - cy[None, :]); dy = np.minimum(dy, 25 - dy)
        d2 = dx * dx + dy * dy
        np.fill_diagonal(d2, 100)
        assert d2.min() > 9, f"colour {k}: sites within the cutoff"


def test_torus_greedy_batch(sn, oracle):
    """Whole-lattice sweeps of a batch of uneven 2D replicas with greedy
    colourings (two replicas share a box and its colouring; shorter colourings
    are padded with empty colours), against the oracle."""
    sizes = ((25, 25), (16, 20), (25, 25), (9, 7))
    reps = [sn.ReplicaParams(X, Y, 1, T=200 + 50 * i, seed=90 + i) for i, (X, Y) in enumerate(sizes)]
    print_case("torus_greedy_batch", sizes=sizes)
    rng = np.random.default_rng(13)
    lats = [rand_lattice(rp.N, rng, mixed=True) for rp in reps]
    run_pair(sn, oracle, reps, lats, dict(kernel=3), nsweeps=4)
    with sn.StarryNight(reps, sn.TuningParams(kernel=3)) as ctx:
        o0, s0 = ctx.schedule(0)
        o2, s2 = ctx.schedule(2)
        assert np.array_equal(o0, o2) and np.array_equal(s0, s2)


def test_resume_from_saved_lattice_and_sweep_index(sn):
    """Exact restart: 3 + 3 sweeps across two contexts (the lattice and the
    sweep index carried over) equal 6 sweeps in one, bit for bit, for every
    replica of a batch."""
    reps = [sn.ReplicaParams(16, 16, 16, T=300, seed=51), sn.ReplicaParams(20, 18, 16, T=250, seed=52)]
    print_case("resume", n=2)
    rng = np.random.default_rng(51)
    lats = [rand_lattice(r.N, rng, mixed=True) for r in reps]
    with sn.StarryNight(reps) as ctx:
        for r, lat in enumerate(lats):
            ctx.set_lattice(lat, r)
        ctx.MC_sweeps(6)
        straight = [ctx.get_lattice(r) for r in range(2)]
        assert [ctx.sweep_index(r) for r in range(2)] == [6, 6]
    with sn.StarryNight(reps) as ctx:
        for r, lat in enumerate(lats):
            ctx.set_lattice(lat, r)
        ctx.MC_sweeps(3)
        saved = [(ctx.get_lattice(r), ctx.sweep_index(r)) for r in range(2)]
    with sn.StarryNight(reps) as ctx:
        for r, (lat, idx) in enumerate(saved):
            ctx.set_lattice(lat, r)
            ctx.set_sweep_index(r, idx)
        ctx.MC_sweeps(3)
        resumed = [ctx.get_lattice(r) for r in range(2)]
        assert [ctx.sweep_index(r) for r in range(2)] == [6, 6]
    for a, b in zip(straight, resumed):
        assert np.array_equal(a, b)
    # without the index the restart draws the first sweeps' stream again
    with sn.StarryNight(reps) as ctx:
        ctx.set_lattice(saved[0][0], 0)
        ctx.MC_sweeps(3)
        assert not np.array_equal(ctx.get_lattice(0), straight[0])
