"""torch.autograd face of the adjoint kernel (sn_gradient).

    E = lattice_energy(p, lengths, J, Efield, s, model)

p (N,3) are the dipole orientations, lengths (N,) the per-site lengths (0 =
vacancy, not differentiated), J the cage strain, Efield (3,) the field and s a
dipole-strength factor multiplying every length.  The energy is the
Hamiltonian whose single-site differences are the reference's site_energy
(src/starrynight-montecarlo-core.c:76-141):

    E = 1/2 sum_i sum_r sL_i sL_j [p_i.p_j - 3(n.p_i)(n.p_j)]/d^3
        - J/2 sum_<ij> p_i.p_j + sum_i p_i.Efield - [K>0] K sum_i (|p_ix| + |p_iy|)

Forward and backward both run on the GPU in one pass of the adjoint kernel
(energy, dE/dp_i on every site, dE/d(J, Efield, s)); backward only scales the
stored gradients by the incoming grad.  K and the geometry are fixed by the
model's ReplicaParams.  GPU tensors stay in device memory (the lattice is
assembled and the gradient written there); host tensors are copied in and
out.  Both run the same GPU kernel: there is no CPU path.
"""
from __future__ import annotations

import numpy as np
import torch

from .lattice import ReplicaParams, StarryNight


class EnergyModel:
    """A one-replica context the energy function evaluates on (geometry,
    DipoleCutOff and K come from `rp`; J and Efield are inputs)."""

    def __init__(self, rp: ReplicaParams):
        self.rp = rp
        self.ctx = StarryNight(rp)

    def close(self):
        self.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def evaluate(self, p: np.ndarray, lengths: np.ndarray, J: float, E: np.ndarray, s: float):
        """Energy, dE/dp (N,3) and dE/d(J, Ex, Ey, Ez, s) at the given inputs."""
        N = self.rp.N
        lat = np.empty((N, 4), np.float64)
        lat[:, :3] = p
        lat[:, 3] = lengths * s
        self.ctx.set_params(0, self.rp.T, tuple(float(v) for v in E), float(J))
        self.ctx.set_lattice(lat)
        g, dp, e = self.ctx.gradient()
        # dE/ds: the kernel returns d/dt E(t * lengths * s) at t = 1 = s * dE/ds
        if s != 0.0:
            ds = dp[4] / s
        else:   # E_dd is quadratic in s: dE/ds = 2 s E_dd(lengths) = 0 at s = 0
            ds = 0.0
        return e, g, np.array([dp[0], dp[1], dp[2], dp[3], ds])

    def evaluate_device(self, p: torch.Tensor, lengths: torch.Tensor, J: float, E, s: float):
        """evaluate() on GPU tensors: the lattice is assembled and the gradient
        written in device memory (no host copy of the N-site arrays)."""
        N = self.rp.N
        lat = torch.empty((N, 4), dtype=torch.float64, device=p.device)
        lat[:, :3] = p
        lat[:, 3] = lengths * s
        self.ctx.set_params(0, self.rp.T, tuple(float(v) for v in E), float(J))
        self.ctx.set_lattice_device(lat)
        g = torch.empty((N, 3), dtype=torch.float64, device=p.device)
        dp, e = self.ctx.gradient_device(g)
        ds = dp[4] / s if s != 0.0 else 0.0
        return e, g, np.array([dp[0], dp[1], dp[2], dp[3], ds])


class _LatticeEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, lengths, J, E, s, model):
        En = E.detach().to("cpu", torch.float64).numpy()
        if p.is_cuda:
            e, g, dparams = model.evaluate_device(p.detach().contiguous(),
                                                  lengths.detach().to(p.device, torch.float64),
                                                  float(J), En, float(s))
        else:
            pn = p.detach().to("cpu", torch.float64).numpy()
            ln = lengths.detach().to("cpu", torch.float64).numpy()
            e, g, dparams = model.evaluate(pn, ln, float(J), En, float(s))
            g = torch.from_numpy(g)
        ctx.save_for_backward(g.to(p.device), torch.from_numpy(dparams).to(p.device))
        return torch.tensor(e, dtype=torch.float64, device=p.device)

    @staticmethod
    def backward(ctx, gout):
        g, dparams = ctx.saved_tensors
        gp = gout * g
        gJ = (gout * dparams[0]).reshape(())
        gE = gout * dparams[1:4]
        gs = (gout * dparams[4]).reshape(())
        return gp, None, gJ, gE, gs, None


def _beta(T: float) -> float:
    """beta as the library computes it (src/starrynight-main.c:215):
    1/((float)T/300) for integral T, 300/T otherwise."""
    T = float(T)
    if T <= 0.0:
        raise ValueError("the susceptibility needs T > 0")
    if T == int(T) and T < 16777216.0:
        return 1.0 / (float(np.float32(T)) / 300.0)
    return 300.0 / T


def field_susceptibility(ctx: StarryNight, replica=0, samples: int = 2000,
                         sweeps_between: int = 2, equilibration: int = 200, blocks: int = 20):
    """Field-derivative dielectric susceptibility of one replica from the
    adjoint's field derivatives.

    For every configuration dE/dEfield = sum_i p_i = M exactly (the energy's
    field term is +sum_i p_i . Efield, as in site_energy,
    src/starrynight-montecarlo-core.c:120-121), and in equilibrium

        d<M_a>/dE_b = -beta (<M_a M_b> - <M_a><M_b>),

    so chi_ab = -(1/N) d<M_a>/dE_b = beta/N Cov(dE/dE_a, dE/dE_b).  The chain
    runs `equilibration` sweeps (every replica of the context sweeps), then
    `samples` samples `sweeps_between` sweeps apart; errors are
    leave-one-block-out jackknife over `blocks` equal blocks.  `replica` is an
    index (one result) or a sequence of indices (a list, one per replica,
    sampled from the same sweeps).  A result is {"chi": (3,3), "chi_err":
    (3,3), "dE_dEfield": (3,) mean, "dE_dEfield_err": (3,), "beta", "samples"}."""
    if samples < 2 * blocks or blocks < 2:
        raise ValueError("need blocks >= 2 and at least two samples per block")
    reps = [int(replica)] if np.ndim(replica) == 0 else [int(r) for r in replica]
    betas = [_beta(ctx.replicas[r].T) for r in reps]
    ctx.MC_sweeps(equilibration)
    d = np.empty((len(reps), samples, 3))
    for k in range(samples):
        ctx.MC_sweeps(sweeps_between)
        for j, r in enumerate(reps):
            dp, _ = ctx.gradient_device(None, r)
            d[j, k] = dp[1:4]
    nb = samples // blocks
    out = []
    for j, r in enumerate(reps):
        beta, N = betas[j], ctx.replicas[r].N
        x = d[j, :nb * blocks]

        def chi_of(y):
            return beta / N * np.cov(y, rowvar=False, bias=True)

        jk = np.array([chi_of(np.delete(x, np.s_[b * nb:(b + 1) * nb], axis=0)) for b in range(blocks)])
        bm = x.reshape(blocks, nb, 3).mean(1)
        out.append({"chi": chi_of(x),
                    "chi_err": np.sqrt((blocks - 1) / blocks * ((jk - jk.mean(0)) ** 2).sum(0)),
                    "dE_dEfield": x.mean(0), "dE_dEfield_err": bm.std(0, ddof=1) / np.sqrt(blocks),
                    "beta": beta, "samples": len(x)})
    return out[0] if np.ndim(replica) == 0 else out


def lattice_energy(p: torch.Tensor, lengths: torch.Tensor, J: torch.Tensor, E: torch.Tensor,
                   s: torch.Tensor, model: EnergyModel) -> torch.Tensor:
    """Total lattice energy (FP64) with gradients to p, J, E and s."""
    if p.dtype != torch.float64:
        raise TypeError("p must be float64 (the path runs in FP64 throughout)")
    J, E, s = (torch.as_tensor(v, dtype=torch.float64, device=p.device) for v in (J, E, s))
    N = model.rp.N
    if p.shape != (N, 3) or lengths.shape != (N,) or E.shape != (3,):
        raise ValueError(f"expected p ({N},3), lengths ({N},), E (3,)")
    return _LatticeEnergy.apply(p, lengths, J, E, s, model)
