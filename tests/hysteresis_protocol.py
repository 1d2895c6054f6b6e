"""The field-ramp (hysteresis) protocol shared by the CPU fixture generator
(tests/golden/make_hysteresis_fixture.py) and the GPU test
(tests/test_gpu_hysteresis.py).

The solid-solution workload ramps E_z over +-0.1 in 41 field values per branch
(steps of 0.005) at 300 K with DipoleCutOff 4 and half the sites vacant
(Dipoles [1, 0], Prevalence [.5, .5]); the reference reads the field from
Efield (src/starrynight-config.c:132-134) and plots the loop P(E)
(figures/hysterisis/hysterisis.gpt:1).  Here the loop runs up from -0.1 to
+0.1, then down to -0.1: 82 steps.  At each step a chain runs SWEEPS sweeps
and records the mean P_z = sum_i p_iz / N over the second half of them.

Loop figures (per chain, then averaged over chains with the spread over
chains as the error):
  P_up[k], P_down[k]  the two branches at E_k = -0.1 + 0.005 k
  remanence           P at E = 0 on the descending and on the ascending branch
  coercive field      E where each branch crosses P = 0 (linear interpolation)
  area                sum over k of (P_down - P_up) * dE, the loop's area
"""
from __future__ import annotations

import numpy as np

X = Y = Z = 16
RC = 4
T = 300.0
J = 1.0
LENGTHS = (1.0, 0.0)
PREVALENCE = (0.5, 0.5)
E_VALUES = [-0.1 + 0.005 * k for k in range(41)]
SWEEPS = 500
SAMPLE_EVERY = 10


def fields():
    """(branch, k, E_z) for the 82 steps: up, then down."""
    out = [("up", k, E_VALUES[k]) for k in range(41)]
    out += [("down", k, E_VALUES[k]) for k in range(40, -1, -1)]
    return out


def initial_lattice(chain: int) -> np.ndarray:
    """Random unit dipoles with half the sites vacant, from numpy seed `chain`."""
    rng = np.random.default_rng(70000 + chain)
    N = X * Y * Z
    v = rng.normal(size=(N, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    L = np.where(rng.random(N) < PREVALENCE[0], LENGTHS[0], LENGTHS[1])
    return np.concatenate([v, L[:, None]], 1)


def zero_crossing(E, P):
    """First E at which P changes sign along the branch (linear interpolation);
    NaN if it never does."""
    for k in range(len(P) - 1):
        if P[k] == 0.0:
            return E[k]
        if (P[k] < 0) != (P[k + 1] < 0):
            return E[k] + (E[k + 1] - E[k]) * (0 - P[k]) / (P[k + 1] - P[k])
    return float("nan")


def loop_figures(p_up: np.ndarray, p_down: np.ndarray) -> dict:
    """Figures of one chain's loop; p_up / p_down indexed by k (ascending E)."""
    E = np.asarray(E_VALUES)
    k0 = 20   # E = 0
    return {
        "remanence_down": float(p_down[k0]),
        "remanence_up": float(p_up[k0]),
        "coercive_up": zero_crossing(E, p_up),
        "coercive_down": zero_crossing(E[::-1], p_down[::-1]),
        "area": float(np.sum(p_down - p_up) * 0.005),
    }


def summarise(chains_up, chains_down) -> dict:
    """Mean and standard error over chains of the branches and figures."""
    up, down = np.asarray(chains_up), np.asarray(chains_down)
    n = len(up)
    out = {"chains": n, "E": E_VALUES,
           "P_up": up.mean(0).tolist(), "P_up_se": (up.std(0, ddof=1) / np.sqrt(n)).tolist(),
           "P_down": down.mean(0).tolist(), "P_down_se": (down.std(0, ddof=1) / np.sqrt(n)).tolist()}
    figs = [loop_figures(u, d) for u, d in zip(up, down)]
    for key in figs[0]:
        v = np.array([f[key] for f in figs])
        v = v[np.isfinite(v)]
        out[key] = [float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v)))] if len(v) > 1 else [float("nan")] * 2
    return out
