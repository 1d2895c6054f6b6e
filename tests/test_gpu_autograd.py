"""Adjoint through torch.autograd (starrynight_amd.autograd.lattice_energy).

Bars (all FP64):
  * autograd gradients equal the adjoint kernel's and the oracle's gradient to
    1e-12 of their scale;
  * dE/dp_i matches central differences of the GPU single-site energy
    difference (h = 1e-6) to 1e-7 relative;
  * dE/d(J, Efield, s) match central differences of the GPU total energy
    (h = 1e-6) to 1e-7 relative;
  * the reduced differentiation workload (32 replicas of 48^3, rc = 3) gives
    every replica's adjoint equal to the oracle's;
  * the field-derivative dielectric susceptibility from the adjoint's
    dE/dEfield fluctuations equals the finite-field derivative of <dE/dEfield>
    (the fluctuation-response identity) within 3 standard errors.
"""
import numpy as np
import pytest
import torch

from tests.helpers import oracle_model, print_case, rand_lattice

pytestmark = pytest.mark.gpu

H = 1e-6


def _inputs(N, seed, mixed=True):
    lat = rand_lattice(N, np.random.default_rng(seed), mixed=mixed)
    return lat


def test_autograd_matches_adjoint_and_oracle(sn, oracle):
    from starrynight_amd.autograd import EnergyModel, lattice_energy
    rp = sn.ReplicaParams(12, 12, 12, DipoleCutOff=3, CageStrain=0.7, Efield=(0.02, -0.01, 0.05))
    print_case("autograd", X=12, rc=3)
    lat = _inputs(rp.N, 1)
    dev = torch.device("cuda:0")
    p = torch.tensor(lat[:, :3], device=dev, requires_grad=True)
    L = torch.tensor(lat[:, 3], device=dev)
    J = torch.tensor(0.7, dtype=torch.float64, device=dev, requires_grad=True)
    E = torch.tensor([0.02, -0.01, 0.05], dtype=torch.float64, device=dev, requires_grad=True)
    s = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
    with EnergyModel(rp) as m:
        e = lattice_energy(p, L, J, E, s, m)
        (2.5 * e).backward()
    o = oracle_model(oracle, rp)
    ew, _ = o.energy(lat)
    gw, dpw = o.gradient(lat)
    sc = np.abs(gw).max() + 1.0
    assert abs(e.item() - ew) <= 1e-12 * (abs(ew) + rp.N)
    assert np.abs(p.grad.cpu().numpy() - 2.5 * gw).max() <= 1e-11 * sc
    assert abs(J.grad.item() - 2.5 * dpw[0]) <= 1e-11 * (abs(dpw[0]) + 1)
    assert np.abs(E.grad.cpu().numpy() - 2.5 * dpw[1:4]).max() <= 1e-11 * (np.abs(dpw[1:4]).max() + 1)
    assert abs(s.grad.item() - 2.5 * dpw[4]) <= 1e-11 * (abs(dpw[4]) + 1)


@pytest.mark.parametrize("rc,dims,s0", [(3, (16, 16, 16), 1.0), (3, (16, 16, 16), 0.8),
                                        (4, (14, 14, 14), 1.3), (3, (25, 25, 1), 1.0)])
def test_autograd_vs_central_differences(sn, rc, dims, s0):
    from starrynight_amd.autograd import EnergyModel
    rp = sn.ReplicaParams(*dims, DipoleCutOff=rc, CageStrain=0.9, Efield=(0.05, 0.0, -0.02))
    print_case("autograd_fd", dims=dims, rc=rc, s=s0)
    lat = _inputs(rp.N, rc + dims[2])
    p, L = lat[:, :3].copy(), lat[:, 3].copy()
    J0, E0 = 0.9, np.array([0.05, 0.0, -0.02])
    with EnergyModel(rp) as m:
        e0, g, dp = m.evaluate(p, L, J0, E0, s0)

        def energy(J=J0, E=E0, s=s0):
            return m.evaluate(p, L, J, E, s)[0]

        fd = np.array([
            (energy(J=J0 + H) - energy(J=J0 - H)) / (2 * H),
            *[(energy(E=E0 + H * np.eye(3)[k]) - energy(E=E0 - H * np.eye(3)[k])) / (2 * H) for k in range(3)],
            (energy(s=s0 + H) - energy(s=s0 - H)) / (2 * H),
        ])
        for k in range(5):
            assert abs(fd[k] - dp[k]) <= 1e-7 * max(1.0, abs(dp[k])), (k, fd[k], dp[k])

        # per-site: central differences of the exact single-site energy difference
        m.evaluate(p, L, J0, E0, s0)
        rng = np.random.default_rng(5)
        sites = rng.choice(rp.N, 256, replace=False).astype(np.int32)
        for k in range(3):
            plus = p[sites].copy(); plus[:, k] += H
            minus = p[sites].copy(); minus[:, k] -= H
            dplus = m.ctx.site_energy_batch(sites, plus)
            dminus = m.ctx.site_energy_batch(sites, minus)
            fdk = (dplus - dminus) / (2 * H)
            err = np.abs(fdk - g[sites, k]) / np.maximum(1.0, np.abs(g[sites, k]))
            assert err.max() <= 1e-7, (k, err.max())


def test_differentiation_workload_32x48cube(sn, oracle):
    """32 replicas of 48^3 at rc=3 in one context (the differentiation
    workload): every replica's energy and adjoint equal the oracle's, and the
    field derivatives give the polarisation (dE/dEfield = sum_i p_i)."""
    rng = np.random.default_rng(48)
    reps = [sn.ReplicaParams(48, 48, 48, DipoleCutOff=3, CageStrain=0.5 + 0.02 * i,
                             Efield=(0.0, 0.0, 0.002 * i), T=300, seed=100 + i) for i in range(32)]
    print_case("diff_workload", replicas=32, L=48)
    lats = [rand_lattice(r.N, rng, mixed=(i % 4 == 0)) for i, r in enumerate(reps)]
    with sn.StarryNight(reps) as ctx:
        for i, l in enumerate(lats):
            ctx.set_lattice(l, i)
        res = [ctx.gradient(i) for i in range(len(reps))]
        # all 32 in one pass (the batched form the workload is timed with)
        g_all = torch.empty((sum(r.N for r in reps), 3), dtype=torch.float64, device="cuda:0")
        dp_all, e_all = ctx.gradient_all_device(g_all)
        g_all = g_all.cpu().numpy()
    for i, (g, dp, e) in enumerate(res):
        gi = g_all[i * reps[0].N:(i + 1) * reps[0].N]
        assert abs(e_all[i] - e) <= 1e-12 * (abs(e) + reps[i].N)
        assert np.abs(dp_all[i] - dp).max() <= 1e-11 * (np.abs(dp).max() + 1)
        assert np.abs(gi - g).max() <= 1e-11 * (np.abs(g).max() + 1)
    for i in (0, 1, 13, 31):
        g, dp, e = res[i]
        o = oracle_model(oracle, reps[i])
        ew, pw = o.energy(lats[i])
        gw, dpw = o.gradient(lats[i])
        assert abs(e - ew) <= 1e-12 * (np.abs(pw).sum() + 1)
        assert np.abs(g - gw).max() <= 1e-11 * (np.abs(gw).max() + 1)
        assert np.abs(dp - dpw).max() <= 1e-11 * (np.abs(dpw).max() + 1)
    for i, (g, dp, e) in enumerate(res):
        M = lats[i][:, :3].sum(0)
        assert np.abs(dp[1:4] - M).max() <= 1e-9 * reps[i].N


def test_field_susceptibility_fluctuation_response(sn):
    """chi_ab = beta/N Cov(dE/dE_a, dE/dE_b) at E = 0 (adjoint field
    derivatives sampled along the chain) against -(1/N) d<M_z>/dE_z by central
    differences between chains at E_z = +-delta (the same sweeps).  12^3,
    rc = 3, T = 900 K (beta = 1/3, paraelectric); beta delta = 0.1 keeps the
    response linear to 0.1%."""
    from starrynight_amd.autograd import field_susceptibility
    delta, T = 0.3, 900
    reps = [sn.ReplicaParams(12, 12, 12, DipoleCutOff=3, T=T, CageStrain=1.0, Efield=(0.0, 0.0, ez),
                             seed=900 + k) for k, ez in enumerate((0.0, delta, -delta))]
    print_case("field_susceptibility", X=12, T=T, delta=delta)
    with sn.StarryNight(reps) as ctx:
        for r in range(3):
            ctx.init_lattice_random(r)
        res = field_susceptibility(ctx, [0, 1, 2], samples=3000, sweeps_between=2, equilibration=300)
    N = reps[0].N
    chi, err = res[0]["chi"], res[0]["chi_err"]
    mp, mm = res[1]["dE_dEfield"][2], res[2]["dE_dEfield"][2]
    fd = -(mp - mm) / (2 * delta * N)
    fd_err = np.hypot(res[1]["dE_dEfield_err"][2], res[2]["dE_dEfield_err"][2]) / (2 * delta * N)
    print(f"chi diag {np.diag(chi)} +- {np.diag(err)}; finite field {fd:.5f} +- {fd_err:.5f}", flush=True)
    assert abs(chi[2, 2] - fd) <= 3 * np.hypot(err[2, 2], fd_err)
    # cubic symmetry: equal diagonal, vanishing off-diagonal
    for a in range(2):
        assert abs(chi[a, a] - chi[2, 2]) <= 3 * np.hypot(err[a, a], err[2, 2])
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(chi[a, b]) <= 3 * err[a, b] + 1e-3 * chi[2, 2]
    # the field tilts the mean polarisation against E (field term +p.E)
    assert mp < 0 < mm and fd > 0
