"""Energy and adjoint (sn_energy / sn_gradient) vs the oracle and vs finite
differences.  Bars: total energy and parameter gradients to 1e-12 relative of
the oracle, per-site gradients to 1e-12 of their scale, central differences
(h = 1e-6) of the GPU total energy to 1e-7 relative."""
import numpy as np
import pytest

from tests.helpers import oracle_model, print_case, rand_lattice

pytestmark = pytest.mark.gpu

# Lattices that have a sweep schedule take the tiled (LDS-staged) energy
# kernel for DipoleCutOff 1..5; 14^3 at rc = 4 has none and 20^3 at rc = 6
# is beyond the tiled kernel's cutoffs: both take the per-site gather kernel.
# `L` is the uniform dipole length (None: mixed lengths with vacancies).
CASES = [
    dict(X=12, Y=12, Z=12, DipoleCutOff=3, CageStrain=0.8, Efield=(0.1, -0.2, 0.3), K=0.0, L=None),
    dict(X=16, Y=16, Z=16, DipoleCutOff=3, CageStrain=1.0, Efield=(0.0, 0.0, 0.05), K=0.4, L=None),
    dict(X=25, Y=25, Z=1, DipoleCutOff=3, CageStrain=1.0, Efield=(0.0, 0.0, 0.0), K=0.0, L=None),
    dict(X=14, Y=14, Z=14, DipoleCutOff=4, CageStrain=0.5, Efield=(0.0, 0.02, 0.0), K=0.0, L=None),
    dict(X=16, Y=16, Z=16, DipoleCutOff=3, CageStrain=1.2, Efield=(0.03, 0.0, -0.01), K=0.0, L=1.0),
    dict(X=16, Y=16, Z=16, DipoleCutOff=3, CageStrain=0.6, Efield=(0.0, 0.0, 0.0), K=0.3, L=0.8),
    dict(X=12, Y=16, Z=20, DipoleCutOff=1, CageStrain=1.0, Efield=(0.0, 0.1, 0.0), K=0.0, L=None),
    dict(X=16, Y=16, Z=16, DipoleCutOff=2, CageStrain=0.7, Efield=(0.0, 0.0, 0.02), K=0.0, L=1.3),
    dict(X=24, Y=24, Z=24, DipoleCutOff=4, CageStrain=1.0, Efield=(0.0, 0.0, 0.0), K=0.0, L=1.0),
    dict(X=20, Y=20, Z=22, DipoleCutOff=5, CageStrain=0.9, Efield=(0.01, 0.0, 0.0), K=0.0, L=None),
    dict(X=20, Y=20, Z=20, DipoleCutOff=6, CageStrain=0.9, Efield=(0.0, 0.0, 0.0), K=0.0, L=None),
    dict(X=32, Y=24, Z=1, DipoleCutOff=4, CageStrain=1.0, Efield=(0.0, 0.05, 0.0), K=0.2, L=0.9),
]


def _case_lattice(rp, L):
    lat = rand_lattice(rp.N, np.random.default_rng(rp.N), mixed=L is None)
    if L is not None:
        lat[:, 3] = L
    return lat


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['X']}x{c['Y']}x{c['Z']}rc{c['DipoleCutOff']}"
                                                      f"K{c['K']}L{c['L']}")
def test_energy_gradient_parity(sn, oracle, case):
    case = dict(case)
    L = case.pop("L")
    rp = sn.ReplicaParams(**case)
    print_case("energy_gradient", L=L, **case)
    lat = _case_lattice(rp, L)
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


def test_device_lattice_and_gradient(sn, oracle):
    """sn_set_lattice_device / sn_gradient_device give the host entry points'
    results (uniform lengths detected from device memory), and a non-finite
    value is refused."""
    import torch
    dev = torch.device("cuda:0")
    for L in (0.8, None):
        rp = sn.ReplicaParams(16, 16, 16, DipoleCutOff=3, CageStrain=0.9, Efield=(0.0, 0.01, 0.02))
        print_case("device_gradient", L=L)
        lat = _case_lattice(rp, L)
        with sn.StarryNight(rp) as ctx:
            ctx.set_lattice(lat)
            g, dp, e = ctx.gradient()
            ctx.set_lattice(np.zeros_like(lat))
            ctx.set_lattice_device(torch.tensor(lat, device=dev))
            gd = torch.empty((rp.N, 3), dtype=torch.float64, device=dev)
            dpd, ed = ctx.gradient_device(gd)
            assert np.array_equal(ctx.get_lattice(), lat)
            assert np.array_equal(gd.cpu().numpy(), g) and np.array_equal(dpd, dp) and ed == e
            dpn, en = ctx.gradient_device(None)
            assert np.array_equal(dpn, dp) and en == e
            bad = torch.tensor(lat, device=dev)
            bad[77, 1] = float("nan")
            with pytest.raises(sn.SnError):
                ctx.set_lattice_device(bad)
            assert np.array_equal(ctx.get_lattice(), lat)


@pytest.mark.parametrize("kind", ["tiled", "gather"])
def test_gradient_all_replicas_in_one_pass(sn, oracle, kind):
    """sn_gradient_all_device: every replica's energy, parameter derivatives and
    per-site gradient from one pass equal the oracle's (uniform, mixed and K > 0
    replicas of different sizes; the tiled kernel, or the per-site kernel when
    the lattices have no tiling)."""
    import torch
    if kind == "tiled":
        reps = [sn.ReplicaParams(16, 16, 16, DipoleCutOff=3, CageStrain=0.9, Efield=(0.0, 0.0, 0.02)),
                sn.ReplicaParams(24, 24, 24, DipoleCutOff=3, CageStrain=1.1, Efield=(0.01, 0.0, 0.0)),
                sn.ReplicaParams(16, 16, 16, DipoleCutOff=3, CageStrain=0.7, K=0.3)]
        lens = [0.8, None, 1.0]
    else:
        reps = [sn.ReplicaParams(14, 14, 14, DipoleCutOff=4, CageStrain=0.5, Efield=(0.0, 0.02, 0.0)),
                sn.ReplicaParams(14, 14, 14, DipoleCutOff=4, CageStrain=1.0)]
        lens = [None, 1.0]
    print_case("gradient_all", kind=kind, replicas=len(reps))
    lats = [_case_lattice(rp, L) for rp, L in zip(reps, lens)]
    ntot = sum(rp.N for rp in reps)
    with sn.StarryNight(reps) as ctx:
        for r, lat in enumerate(lats):
            ctx.set_lattice(lat, r)
        g = torch.full((ntot, 3), float("nan"), dtype=torch.float64, device="cuda:0")
        dp, e = ctx.gradient_all_device(g)
        dpn, en = ctx.gradient_all_device(None)
        g = g.cpu().numpy()
    assert np.array_equal(dp, dpn) and np.array_equal(e, en)
    row = 0
    for r, (rp, lat) in enumerate(zip(reps, lats)):
        o = oracle_model(oracle, rp)
        ew, pw = o.energy(lat)
        gw, dpw = o.gradient(lat)
        scale = np.abs(pw).sum() + 1.0
        assert abs(e[r] - ew) <= 1e-12 * scale
        assert np.abs(dp[r] - dpw).max() <= 1e-11 * (np.abs(dpw).max() + 1.0)
        assert np.abs(g[row:row + rp.N] - gw).max() <= 1e-11 * (np.abs(gw).max() + 1.0)
        row += rp.N
