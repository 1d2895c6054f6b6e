"""Equilibrium observables shared by the fixture generator (long CPU runs of
the oracle's FP64 MC_moves restatement) and the GPU statistical test.

Observables per sample of a lattice with N sites:
  e      = E / N            total energy per site (Hamiltonian whose single-
                            site differences are site_energy)
  m      = |sum_i p_i| / N  sqrt(landau_order) / N, src/starrynight-analysis.c:506-526
  mz     = sum_i p_iz / N
  m2     = m^2
  fe[k], afe[k]             radial_order_parameter bins r^2 = RDF_BINS[k],
                            src/starrynight-analysis.c:528-598
Derived: chi = beta N (<m2> - <m>^2) over all samples (finite-lattice
susceptibility of the order-parameter magnitude).  Errors come from block
means: every chain is cut into equal blocks much longer than the
autocorrelation time; the standard error of a mean is the spread of the block
means over sqrt(#blocks), that of chi a leave-one-block-out jackknife.  The
CPU chains run 5 blocks of 3600 sweeps each; every GPU replica is one block
of 2000 sweeps (after 2000 of burn-in), so its error is the spread over
independent chains.  The slowest mode is the tilt of the whole polarisation
in the ordered ferroelectric state (~10^2 sweeps); blocks of 80 sweeps
underestimated the error of mz there by about a factor of two.
"""
from __future__ import annotations

import numpy as np

RDF_BINS = [1, 2, 3, 4, 5, 6, 8, 9, 12]

# state points: geometry, couplings, the two species and how the chains run
STATES = {
    "cube12_T600": dict(X=12, Y=12, Z=12, rc=3, T=600.0, J=1.0, E=(0.0, 0.0, 0.0),
                        lengths=(1.0,), prevalence=(1.0,)),
    "cube12_T600_field_mixed": dict(X=12, Y=12, Z=12, rc=3, T=600.0, J=1.0, E=(0.0, 0.0, 0.3),
                                    lengths=(1.0, 0.5), prevalence=(0.5, 0.5)),
    "square25_T450": dict(X=25, Y=25, Z=1, rc=3, T=450.0, J=1.0, E=(0.0, 0.0, 0.0),
                          lengths=(1.0,), prevalence=(1.0,)),
    # ordered phases: ferroelectric (CageStrain 4, <|P|> ~ 0.79) and the
    # antiferroelectric order of CageStrain 1 at low temperature (E/N ~ -3.65,
    # fe[r2=12] ~ -0.73).  At 200 K random starts order within ~2000 sweeps;
    # at 150 K (acceptance 0.06) chains keep distinct long-lived defect states
    # for thousands of sweeps, so a handful of chains does not sample the
    # distribution.
    # The ferroelectric chains start polarised along +x in a field along -x
    # (the field term of site_energy is +E.p, src/starrynight-montecarlo-core.c:120-121,
    # so dipoles line up against E).  From random starts some chains stay in
    # multi-domain states for >10^4 sweeps at 12^3; without the field the
    # direction of the polarisation wanders over the sphere more slowly than
    # the chains run, so averages of direction-dependent observables (mz, afe)
    # over a handful of chains depend on where each chain happened to drift.
    # The field (0.2: about 130 theta^2 of energy for a tilt theta of the
    # whole lattice, against kT = 2.67) holds the direction near +x.
    "cube12_T800_J4_ferro_Ex": dict(X=12, Y=12, Z=12, rc=3, T=800.0, J=4.0, E=(-0.2, 0.0, 0.0),
                                    lengths=(1.0,), prevalence=(1.0,), initial="ferro_x"),
    "cube12_T200_antiferro": dict(X=12, Y=12, Z=12, rc=3, T=200.0, J=1.0, E=(0.0, 0.0, 0.0),
                                  lengths=(1.0,), prevalence=(1.0,)),
}


def beta_of(T: float) -> float:
    return 300.0 / T


def species_lengths(N: int, lengths, prevalence, seed: int) -> np.ndarray:
    """A fixed assignment of species lengths (the same on CPU and GPU)."""
    rng = np.random.default_rng(seed)
    cum = np.cumsum(prevalence)
    u = rng.random(N)
    idx = np.searchsorted(cum, u, side="right")
    return np.asarray(lengths, np.float64)[np.minimum(idx, len(lengths) - 1)]


def random_lattice(N: int, L: np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(N, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    return np.concatenate([v, L[:, None]], 1)


def initial_lattice(st: dict, N: int, L: np.ndarray, seed: int) -> np.ndarray:
    """The chains' starting lattice: random unit dipoles, or all along +x
    for a state point marked initial="ferro_x"."""
    if st.get("initial") == "ferro_x":
        lat = np.zeros((N, 4))
        lat[:, 0] = 1.0
        lat[:, 3] = L
        return lat
    return random_lattice(N, L, seed)


def sample_scalars(energy: float, M: np.ndarray, N: int) -> np.ndarray:
    """(e, m, mz, m2) of one sample."""
    m = float(np.linalg.norm(M)) / N
    return np.array([energy / N, m, M[2] / N, m * m])


def block_stats(chains_scalars, chains_rdf, nblocks: int, beta: float, N: int) -> dict:
    """Means and standard errors from equal blocks of every chain.

    chains_scalars: list of (n_samples, 4) arrays; chains_rdf: list of
    (n_rdf_samples, 2 * len(RDF_BINS)) arrays."""
    blocks = []
    blk_rdf = []
    for sc, rd in zip(chains_scalars, chains_rdf):
        n = (len(sc) // nblocks) * nblocks
        blocks += list(sc[:n].reshape(nblocks, -1, 4).mean(1))
        nr = (len(rd) // nblocks) * nblocks
        blk_rdf += list(rd[:nr].reshape(nblocks, -1, rd.shape[1]).mean(1))
    b = np.asarray(blocks)
    nb = len(b)
    out = {}
    for k, col in (("e", 0), ("m", 1), ("mz", 2)):
        out[k] = [float(b[:, col].mean()), float(b[:, col].std(ddof=1) / np.sqrt(nb))]
    # chi over all samples, its error by leave-one-block-out jackknife
    tot = b.sum(0)
    chi = beta * N * (tot[3] / nb - (tot[1] / nb) ** 2)
    jk = np.array([beta * N * ((tot[3] - r[3]) / (nb - 1) - ((tot[1] - r[1]) / (nb - 1)) ** 2) for r in b])
    out["chi"] = [float(chi), float(np.sqrt((nb - 1) / nb * ((jk - jk.mean()) ** 2).sum()))]
    r = np.asarray(blk_rdf)
    out["rdf_mean"] = [float(x) for x in r.mean(0)]
    out["rdf_se"] = [float(x) for x in r.std(0, ddof=1) / np.sqrt(len(r))]
    out["blocks"] = nb
    return out
