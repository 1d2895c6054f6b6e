#!/usr/bin/env python3
"""Writes tests/golden/hysteresis_16cube_rc4.json: the field-ramp loop of
tests/hysteresis_protocol.py from CPU chains of the oracle's FP64
restatement of MC_moves (random site with replacement, MT19937 stream,
Metropolis; src/starrynight-montecarlo-core.c:143-191), one chain per process.

    python tests/golden/make_hysteresis_fixture.py [--chains 8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tests import hysteresis_protocol as H  # noqa: E402


def run_chain(chain: int):
    from oracle import oracle as O
    N = H.X * H.Y * H.Z
    o = O.Oracle(H.X, H.Y, H.Z, H.RC, J=H.J, E=(0.0, 0.0, H.E_VALUES[0]), beta=300.0 / H.T)
    lat = H.initial_lattice(chain)
    mt = O.MT19937(0x4E57 + 104729 * chain)
    up, down = np.zeros(41), np.zeros(41)
    for branch, k, Ez in H.fields():
        o.m.E[2] = Ez
        acc = []
        for s in range(1, H.SWEEPS + 1):
            o.mc_moves_f64(lat, mt, N)
            if s > H.SWEEPS // 2 and s % H.SAMPLE_EVERY == 0:
                acc.append(O.moment_sum(lat)[2] / N)
        (up if branch == "up" else down)[k] = np.mean(acc)
    return up, down


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "hysteresis_16cube_rc4.json"))
    a = ap.parse_args()
    t0 = time.time()
    ups, downs = [], []
    with ProcessPoolExecutor(min(a.chains, os.cpu_count() or 1)) as ex:
        for up, down in ex.map(run_chain, range(a.chains)):
            ups.append(up)
            downs.append(down)
            print(f"  chain done ({time.time() - t0:.0f} s)", flush=True)
    out = {"method": "oracle or_mc_moves_f64 (random-site MT19937 Metropolis, FP64)",
           "lattice": [H.X, H.Y, H.Z], "DipoleCutOff": H.RC, "T": H.T, "CageStrain": H.J,
           "Dipoles": H.LENGTHS, "Prevalence": H.PREVALENCE, "sweeps_per_step": H.SWEEPS,
           "loop": H.summarise(ups, downs)}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print({k: out["loop"][k] for k in ("remanence_down", "remanence_up", "coercive_up", "coercive_down", "area")})


if __name__ == "__main__":
    main()
