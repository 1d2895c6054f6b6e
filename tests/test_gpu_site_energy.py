"""site_energy drop-in (sn_site_energy_batch) vs the FP64 oracle restatement.

Bar: |dE_gpu - dE_oracle| <= 1e-12 * S, with S the sum of |term| of the
reference's per-neighbour expression (the rounding scale of both sums).
"""
import numpy as np
import pytest

from tests.helpers import oracle_model, print_case, rand_lattice

pytestmark = pytest.mark.gpu

CASES = [
    dict(X=16, Y=16, Z=16, DipoleCutOff=3, CageStrain=1.0, Efield=(0.1, -0.05, 0.2), K=0.0, mixed=True),
    dict(X=25, Y=25, Z=1, DipoleCutOff=3, CageStrain=1.0, Efield=(0.0, 0.0, 0.0), K=0.0, mixed=False),
    dict(X=20, Y=20, Z=20, DipoleCutOff=4, CageStrain=0.5, Efield=(0.0, 0.0, 0.1), K=0.3, mixed=True),
    dict(X=12, Y=14, Z=16, DipoleCutOff=3, CageStrain=2.0, Efield=(0.02, 0.0, 0.0), K=0.0, mixed=False),
    dict(X=9, Y=9, Z=9, DipoleCutOff=3, CageStrain=1.0, Efield=(0.0, 0.0, 0.0), K=0.5, mixed=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['X']}x{c['Y']}x{c['Z']}rc{c['DipoleCutOff']}")
def test_site_energy_batch_parity(sn, oracle, case):
    case = dict(case)
    mixed = case.pop("mixed")
    rp = sn.ReplicaParams(**case)
    print_case("site_energy_batch", **case)
    rng = np.random.default_rng(rp.N)
    lat = rand_lattice(rp.N, rng, mixed=mixed)
    n = min(rp.N, 4096)
    sites = rng.permutation(rp.N)[:n].astype(np.int32)
    props = rand_lattice(n, rng)[:, :3]
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        got = ctx.site_energy_batch(sites, props)
    o = oracle_model(oracle, rp)
    want, scale = o.site_energy_batch(lat, sites, props)
    err = np.abs(got - want) / np.maximum(scale, 1e-300)
    print(f"  max scaled err {err.max():.3e}", flush=True)
    assert err.max() <= 1e-12


def test_ferroelectric_flip_gpu(sn):
    for J, Ex in ((1.0, 0.0), (0.7, 0.3)):
        rp = sn.ReplicaParams(16, 16, 16, CageStrain=J, Efield=(Ex, 0, 0))
        lat = np.zeros((rp.N, 4)); lat[:, 0] = 1; lat[:, 3] = 1
        with sn.StarryNight(rp) as ctx:
            ctx.set_lattice(lat)
            dE = ctx.site_energy(3, 4, 5, (-1.0, 0.0, 0.0))
        assert abs(dE - (12 * J - 2 * Ex)) < 1e-12


def test_linear_in_proposal(sn):
    """dE(p -> p') is linear in p' when K = 0: dE(a) + dE(b) - dE(0-proposal)."""
    rp = sn.ReplicaParams(16, 16, 16, CageStrain=0.9, Efield=(0.1, 0.2, 0.3))
    rng = np.random.default_rng(1)
    lat = rand_lattice(rp.N, rng, mixed=True)
    with sn.StarryNight(rp) as ctx:
        ctx.set_lattice(lat)
        s = np.full(3, 1234, np.int32)
        a, b = rng.normal(size=3), rng.normal(size=3)
        d = ctx.site_energy_batch(s, np.stack([a, b, a + b]))
        z = ctx.site_energy_batch(s[:1], np.zeros((1, 3)))
    assert abs((d[0] - z[0]) + (d[1] - z[0]) - (d[2] - z[0])) < 1e-12 * 100
