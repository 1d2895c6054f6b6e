"""Edge inputs of the entry points: empty batches, zero sweeps, indices out of
range, negative counts and lattices with every site vacant.  Each is either a
no-op or a clean status code; nothing reaches a kernel with a bad index."""
import numpy as np
import pytest

from tests.helpers import print_case, rand_lattice

pytestmark = pytest.mark.gpu


def test_empty_and_invalid_inputs(sn):
    rp = sn.ReplicaParams(16, 16, 16, T=300, seed=5)
    print_case("edges", X=16)
    lat = rand_lattice(rp.N, np.random.default_rng(5), mixed=True)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        # empty batches and zero sweeps change nothing
        assert ctx.site_energy_batch(np.zeros(0, np.int32), np.zeros((0, 3))).shape == (0,)
        dec, mm, nt = ctx.replay(np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros(0), np.zeros(0, np.uint8),
                                 np.zeros(0, np.uint8))
        assert len(dec) == 0 and mm == 0 and nt == 0
        ctx.MC_sweeps(0)
        assert np.array_equal(ctx.get_lattice(), lat) and ctx.counters()["attempted"] == 0
        # out of range and negative requests are refused with a status, not run
        with pytest.raises(sn.SnError) as e:
            ctx.site_energy_batch(np.array([rp.N], np.int32), np.zeros((1, 3)))
        assert e.value.status == 4
        with pytest.raises(sn.SnError) as e:
            ctx.site_energy_batch(np.array([-1], np.int32), np.zeros((1, 3)))
        assert e.value.status == 4
        with pytest.raises(sn.SnError) as e:
            ctx.MC_sweeps(-1)
        assert e.value.status == 1
        with pytest.raises(sn.SnError) as e:
            ctx.energy(replica=1)
        assert e.value.status == 4
        bad = lat.copy()
        bad[3, 0] = np.inf
        with pytest.raises(sn.SnError):
            ctx.set_lattice(bad)
        assert np.array_equal(ctx.get_lattice(), lat)


def test_all_vacant_lattice(sn):
    """Every site vacant (length 0): no move is ever made (the reference skips
    vacancies), the energy and gradient are zero, sweeps still count attempts."""
    rp = sn.ReplicaParams(16, 16, 16, T=300, CageStrain=1.0, Efield=(0.0, 0.0, 0.1), seed=6)
    print_case("edges_vacant", X=16)
    lat = rand_lattice(rp.N, np.random.default_rng(6))
    lat[:, 3] = 0.0
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        ctx.MC_sweeps(3)
        assert np.array_equal(ctx.get_lattice(), lat)
        c = ctx.counters()
        assert c["accept"] == 0 and c["attempted"] == 3 * rp.N
        # site_energy's cage and field terms carry no length factor, so only the
        # dipolar part (and its strength derivative) vanishes
        g, dp, e = ctx.gradient()
        _, parts = ctx.energy()
        assert parts[0] == 0.0 and dp[4] == 0.0
        assert np.all(np.isfinite(g))
