"""Periodic colourings of whole-lattice sweeps (starrynight_amd/csrc/sn_schedule.cpp):
the greedy colouring of the periodic conflict graph (sites within DipoleCutOff
through the periodic boundary, the neighbour offsets of gen_neighbour,
src/starrynight-montecarlo-core.c:38-72) and its checker.

A small C++ harness is compiled with g++ against the schedule code (plain C++,
no HIP); the colourings it prints are checked here by brute force, and the
checker is shown to reject planted mistakes."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "starrynight_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include <utility>
#include "sn_schedule.h"
/* usage: harness X Y Z rc max_class  ->  line 1: colours valid; line 2: start; line 3: sites;
   line 4: checker on a copy with two sites of different colours swapped; line 5: with a duplicate */
int main(int argc, char **argv) {
    sn::ReplicaGeom g{atoi(argv[1]), atoi(argv[2]), atoi(argv[3])};
    const int rc = atoi(argv[4]), cap = atoi(argv[5]);
    std::vector<int32_t> s, st;
    const int n = sn::greedy_torus_classes(g, rc, cap, s, st);
    printf("%d %d\n", n, (int)sn::valid_torus_classes(g, rc, s, st));
    for (size_t k = 0; k < st.size(); k++) printf("%d ", st[k]);
    printf("\n");
    for (size_t k = 0; k < s.size(); k++) printf("%d ", s[k]);
    printf("\n");
    std::vector<int32_t> a = s;
    std::swap(a[0], a[st[1]]);
    printf("%d\n", (int)sn::valid_torus_classes(g, rc, a, st));
    std::vector<int32_t> b = s;
    b[1] = b[0];
    printf("%d\n", (int)sn::valid_torus_classes(g, rc, b, st));
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("colour")
    src = d / "harness.cpp"
    src.write_text(HARNESS)
    exe = d / "harness"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", str(exe), str(src),
                           os.path.join(CSRC, "sn_schedule.cpp")])
    return str(exe)


def run(harness, X, Y, Z, rc, cap):
    out = subprocess.check_output([harness, str(X), str(Y), str(Z), str(rc), str(cap)], text=True).splitlines()
    n, valid = map(int, out[0].split())
    starts = np.array(out[1].split(), dtype=np.int64)
    sites = np.array(out[2].split(), dtype=np.int64)
    return n, valid, starts, sites, int(out[3]), int(out[4])


def conflict_free(X, Y, Z, rc, starts, sites):
    """Brute force: no two sites of a colour within rc by the minimum-image distance."""
    z = sites % Z
    y = (sites // Z) % Y
    x = sites // (Y * Z)
    for k in range(len(starts) - 1):
        sl = slice(starts[k], starts[k + 1])
        d2 = 0
        for c, L in ((x[sl], X), (y[sl], Y), (z[sl], Z)):
            d = np.abs(c[:, None] - c[None, :])
            d = np.minimum(d, L - d)
            d2 = d2 + d * d
        np.fill_diagonal(d2, rc * rc + 1)
        if d2.min() <= rc * rc:
            return False
    return True


@pytest.mark.parametrize("X,Y,Z,rc,cap,most", [
    (25, 25, 1, 3, 64, 16),    # every sublattice colouring of this box has 25 colours
    (30, 26, 1, 3, 64, 18),
    (11, 13, 1, 3, 64, 18),
    (9, 7, 1, 3, 64, 63),
    (10, 10, 10, 2, 64, 40),
    (8, 8, 8, 1, 64, 10),
])
def test_greedy_colouring_is_a_proper_periodic_colouring(harness, X, Y, Z, rc, cap, most):
    n, valid, starts, sites, _, _ = run(harness, X, Y, Z, rc, cap)
    assert valid == 1
    assert len(starts) == n + 1 and starts[0] == 0 and starts[-1] == X * Y * Z
    assert np.array_equal(np.sort(sites), np.arange(X * Y * Z))
    assert np.diff(starts).max() <= cap
    assert n <= most
    assert conflict_free(X, Y, Z, rc, starts, sites)


def test_checker_rejects_planted_mistakes(harness):
    # swapping the first sites of colours 0 and 1 breaks the colouring unless the two
    # happen to be compatible with each other's classes; a duplicated site always does
    n, valid, starts, sites, swapped, dup = run(harness, 25, 25, 1, 3, 64)
    assert valid == 1 and dup == 0
    a = sites.copy()
    a[0], a[starts[1]] = a[starts[1]], a[0]
    assert swapped == int(conflict_free(25, 25, 1, 3, starts, a))
