"""Energy and adjoint (sn_energy / sn_gradient) vs the oracle and vs finite
differences.  Bars: total energy and parameter gradients to 1e-12 relative of
the oracle, per-site gradients to 1e-12 of their scale, central differences
(h = 1e-6) of the GPU total energy to 1e-7 relative."""
import numpy as np
import pytest

from tests.helpers import oracle_model, print_case, rand_lattice

pytestmark = pytest.mark.gpu

CASES = [
    dict(X=12, Y=12, Z=12, DipoleCutOff=3, CageStrain=0.8, Efield=(0.1, -0.2, 0.3), K=0.0),
    dict(X=16, Y=16, Z=16, DipoleCutOff=3, CageStrain=1.0, Efield=(0.0, 0.0, 0.05), K=0.4),
    dict(X=25, Y=25, Z=1, DipoleCutOff=3, CageStrain=1.0, Efield=(0.0, 0.0, 0.0), K=0.0),
    dict(X=14, Y=14, Z=14, DipoleCutOff=4, CageStrain=0.5, Efield=(0.0, 0.02, 0.0), K=0.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['X']}x{c['Z']}rc{c['DipoleCutOff']}K{c['K']}")
def test_energy_gradient_parity(sn, oracle, case):
    rp = sn.ReplicaParams(**case)
    print_case("energy_gradient", **case)
    lat = rand_lattice(rp.N, np.random.default_rng(rp.N), mixed=True)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        e, parts = ctx.energy()
        g, dp, e2 = ctx.gradient()
    o = oracle_model(oracle, rp)
    ew, pw = o.energy(lat)
    gw, dpw = o.gradient(lat)
    scale = np.abs(pw).sum() + 1.0
    assert abs(e - ew) <= 1e-12 * scale and abs(e2 - ew) <= 1e-12 * scale
    assert np.abs(parts - pw).max() <= 1e-12 * scale
    gs = np.abs(gw).max() + 1.0
    assert np.abs(g - gw).max() <= 1e-12 * gs * 10
    assert np.abs(dp - dpw).max() <= 1e-12 * (np.abs(dpw).max() + 1.0) * 10


def test_adjoint_vs_finite_differences(sn):
    rp = sn.ReplicaParams(16, 16, 16, CageStrain=0.9, Efield=(0.05, 0.0, -0.02))
    rng = np.random.default_rng(42)
    lat = rand_lattice(rp.N, rng, mixed=True)
    h = 1e-6
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        g, dp, e0 = ctx.gradient()
        # per-site field: dE of a single-site change is (p'-p).dE/dp_i exactly (K = 0)
        sites = rng.integers(0, rp.N, 512).astype(np.int32)
        props = rand_lattice(512, rng)[:, :3]
        dE = ctx.site_energy_batch(sites, props)
        lin = np.einsum("ij,ij->i", props - lat[sites, :3], g[sites])
        assert np.abs(dE - lin).max() <= 1e-12 * (np.abs(g).max() + 1) * 10

        def E(J=0.9, Ef=(0.05, 0.0, -0.02)):
            ctx.set_params(0, rp.T, Ef, J)
            return ctx.energy()[0]
        fd = (E(J=0.9 + h) - E(J=0.9 - h)) / (2 * h)
        assert abs(fd - dp[0]) <= 1e-7 * max(1.0, abs(dp[0]))
        for k in range(3):
            ep = [0.05, 0.0, -0.02]; em = list(ep); ep[k] += h; em[k] -= h
            fd = (E(Ef=ep) - E(Ef=em)) / (2 * h)
            assert abs(fd - dp[1 + k]) <= 1e-7 * max(1.0, abs(dp[1 + k]))
        E()
        l2 = lat.copy(); l2[:, 3] *= 1 + h; ctx.set_lattice(l2); ep_ = ctx.energy()[0]
        l2 = lat.copy(); l2[:, 3] *= 1 - h; ctx.set_lattice(l2); em_ = ctx.energy()[0]
        assert abs((ep_ - em_) / (2 * h) - dp[4]) <= 1e-7 * max(1.0, abs(dp[4]))


def test_moment_and_counters(sn):
    rp = sn.ReplicaParams(16, 16, 16)
    lat = rand_lattice(rp.N, np.random.default_rng(0))
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        M = ctx.moment()
    assert np.abs(M - lat[:, :3].sum(0)).max() < 1e-11


@pytest.mark.parametrize("dims,mixed", [((20, 20, 20), True), ((25, 25, 1), False), ((12, 10, 9), True),
                                        ((7, 7, 7), False), ((48, 48, 48), False)])
def test_radial_order_parity(sn, oracle, dims, mixed):
    """sn_radial_order vs the oracle's radial_order_parameter restatement:
    counts exact, means to 1e-12 (only the summation order differs)."""
    rp = sn.ReplicaParams(*dims)
    print_case("radial_order", dims=dims, mixed=mixed)
    lat = rand_lattice(rp.N, np.random.default_rng(sum(dims)), mixed=mixed)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        fe, afe, cnt = ctx.radial_order()
    o = oracle_model(oracle, rp)
    few, afew, cntw = o.radial_order(lat)
    assert np.array_equal(cnt, cntw)
    assert np.abs(fe - few).max() <= 1e-12 and np.abs(afe - afew).max() <= 1e-12


@pytest.mark.parametrize("dims,mixed", [((20, 20, 20), True), ((25, 25, 1), False), ((13, 11, 9), True)])
def test_dipole_potential_and_field_parity(sn, oracle, dims, mixed):
    """sn_dipole_potential / sn_dipole_efield vs the oracle's restatements of
    dipole_potential and dipole_electricfield (cutoff 4, as the reference's
    lattice_Efield_XYZ uses): FP64, only the summation order differs."""
    rp = sn.ReplicaParams(*dims)
    print_case("potential_efield", dims=dims, mixed=mixed)
    lat = rand_lattice(rp.N, np.random.default_rng(sum(dims) + 1), mixed=mixed)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        pot = ctx.dipole_potential()
        E4 = ctx.dipole_efield(4)
        E2 = ctx.dipole_efield(2)
    o = oracle_model(oracle, rp)
    assert np.abs(pot - o.dipole_potential(lat)).max() <= 1e-12
    assert np.abs(E4 - o.dipole_efield(lat, 4)).max() <= 1e-12
    assert np.abs(E2 - o.dipole_efield(lat, 2)).max() <= 1e-12


@pytest.mark.parametrize("dims", [(20, 20, 20), (25, 25, 1)])
def test_recombination_parity(sn, oracle, dims):
    rp = sn.ReplicaParams(*dims)
    print_case("recombination", dims=dims)
    lat = rand_lattice(rp.N, np.random.default_rng(5), mixed=True)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        got = ctx.recombination()
    want = oracle_model(oracle, rp).recombination(lat)
    assert np.allclose(got, want, rtol=1e-11, atol=0)
