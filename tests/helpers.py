"""Shared helpers for the GPU parity tests (the oracle is the checker)."""
import numpy as np


def rand_lattice(N, rng, mixed=False, vac=0.1, half=0.3):
    v = rng.normal(size=(N, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    L = np.ones((N, 1))
    if mixed:
        L[rng.random(N) < half] = 0.5
        L[rng.random(N) < vac] = 0.0
    return np.concatenate([v, L], 1)


def oracle_model(O, rp):
    """Oracle model with the same parameters as a ReplicaParams."""
    return O.Oracle(rp.X, rp.Y, rp.Z, rp.DipoleCutOff, J=rp.CageStrain, E=tuple(rp.Efield), K=rp.K,
                    constrain_to_x=rp.ConstrainToX, beta=beta_of_T(rp.T))


def beta_of_T(T):
    """beta as the library computes it: 1/((float)T/300) for integral T."""
    T = float(T)
    if T == 0.0:
        return float("inf")
    if T == int(T):
        return 1.0 / (float(np.float32(T)) / 300.0)
    return 300.0 / T


def print_case(label, **kw):
    print(f"[case] {label} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)
