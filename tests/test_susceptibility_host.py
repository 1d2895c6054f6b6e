"""Host logic of starrynight_amd.autograd.field_susceptibility, without a GPU:
a stand-in context returns field derivatives with a known covariance, and the
estimator must give beta/N times that covariance, with a jackknife error that
covers the deviation."""
import numpy as np
import pytest

from starrynight_amd.autograd import _beta, field_susceptibility


class FakeContext:
    """MC_sweeps / gradient_device / replicas, with dE/dEfield ~ N(mu, C)."""

    def __init__(self, N, T, mu, C, seed):
        self.replicas = [type("R", (), {"N": N, "T": T})()]
        self.rng = np.random.default_rng(seed)
        self.mu, self.L = np.asarray(mu), np.linalg.cholesky(C)
        self.sweeps = 0

    def MC_sweeps(self, n, sync=True):
        self.sweeps += n

    def gradient_device(self, grad, replica=0):
        assert grad is None and replica == 0
        m = self.mu + self.L @ self.rng.normal(size=3)
        return np.array([0.0, *m, 0.0, 0.0]), 0.0


def test_estimator_recovers_known_covariance():
    C = np.array([[40.0, 5.0, 0.0], [5.0, 30.0, -3.0], [0.0, -3.0, 50.0]])
    ctx = FakeContext(N=1000, T=600, mu=(1.0, -2.0, 0.5), C=C, seed=7)
    r = field_susceptibility(ctx, 0, samples=20000, sweeps_between=3, equilibration=10, blocks=20)
    assert ctx.sweeps == 10 + 3 * 20000
    want = _beta(600) / 1000 * C
    assert r["beta"] == 0.5 and r["samples"] == 20000
    assert np.all(np.abs(r["chi"] - want) <= 4 * r["chi_err"])
    # the jackknife error is the right size: ~ beta/N * sqrt((C_aa C_bb + C_ab^2)/n)
    exp_err = _beta(600) / 1000 * np.sqrt((np.outer(np.diag(C), np.diag(C)) + C ** 2) / 20000)
    assert np.all(r["chi_err"] < 2 * exp_err) and np.all(r["chi_err"] > 0.5 * exp_err)
    assert np.all(np.abs(r["dE_dEfield"] - ctx.mu) <= 4 * r["dE_dEfield_err"])


def test_estimator_arguments():
    ctx = FakeContext(N=10, T=300, mu=(0, 0, 0), C=np.eye(3), seed=1)
    with pytest.raises(ValueError):
        field_susceptibility(ctx, 0, samples=10, blocks=20)
    ctx.replicas[0].T = 0
    with pytest.raises(ValueError):
        field_susceptibility(ctx, 0, samples=100, blocks=10)
    assert _beta(300) == 1.0 and _beta(450.5) == 300.0 / 450.5
