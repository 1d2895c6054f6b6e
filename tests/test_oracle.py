"""CPU tests pinning the oracle (oracle/sn_oracle.c) before it is trusted.

Each check is chosen so that a plausible slip (a dropped term, a sign, a
transposed index) fails: library routines (numpy's MT19937, the C++ standard's
known answer, the reference's own mt19937ar-cok.c built from source, Random123
known-answer vectors, rocRAND's Philox), closed forms and brute force.
"""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- MT19937
def test_mt_cxx_standard_known_answer(oracle):
    # [rand.predef]: the 10000th consecutive output of a default-constructed
    # std::mt19937 (seed 5489) is 4123659995.
    assert oracle.MT19937(5489).u32(10000)[-1] == 4123659995


@pytest.mark.parametrize("seed", [5489, 0xDEADBEEF + 300, 0xDEADBEEF, 1])
def test_mt_matches_numpy_init_genrand(oracle, seed):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)   # init_genrand(seed) seeding
    want = bg.random_raw(3000).astype(np.uint32)
    assert np.array_equal(oracle.MT19937(seed).u32(3000), want)


def test_mt_matches_reference_source(oracle):
    """oracle/_ref/libref_mt.so is src/mt19937ar-cok.c compiled unmodified."""
    ref = oracle.ref_mt_stream(0xDEADBEEF + 300, 3000, 3)
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources not mounted)")
    mt = oracle.MT19937(0xDEADBEEF + 300)
    mine = []
    for i in range(3000):
        k = i % 3
        mine.append(float(mt.u32(1)[0]) if k == 0 else (mt.real1() if k == 1 else mt.real2()))
    assert np.array_equal(np.array(mine), ref)


def test_mt_real_closed_forms(oracle):
    a, b = oracle.MT19937(7), oracle.MT19937(7)
    y = a.u32(4)
    assert b.real1() == float(y[0]) * (1.0 / 4294967295.0)
    assert b.real2() == float(y[1]) * (1.0 / 4294967296.0)


# ---------------------------------------------------------------- Philox
RANDOM123_KAT = [  # philox4x32_10: counter, key -> output (Random123 kat_vectors)
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert tuple(int(v) for v in oracle.philox4x32_10(ctr, key)) == out


def test_philox_matches_rocrand(oracle):
    path = os.path.join(GOLDEN, "philox_rocrand.json")
    vecs = json.load(open(path))
    for v in vecs["vectors"]:
        got = [int(x) for x in oracle.philox4x32_10(v["counter"], v["key"])]
        assert got == v["output"], v


# ---------------------------------------------------------------- neighbours
def _ball_count(rc, two_d):
    n = 0
    zr = range(1) if two_d else range(-rc, rc + 1)
    for dx in range(-rc, rc + 1):
        for dy in range(-rc, rc + 1):
            for dz in zr:
                if 0 < dx * dx + dy * dy + dz * dz <= rc * rc:
                    n += 1
    return n


@pytest.mark.parametrize("rc,Z,count", [(3, 10, 122), (3, 1, 28), (4, 10, 256), (1, 10, 6), (2, 1, 12)])
def test_neighbour_table_closed_form(oracle, rc, Z, count):
    nb, n = oracle.neighbours(rc, Z)
    assert n == count == _ball_count(rc, Z == 1)
    offs = [(nb[i].dx, nb[i].dy, nb[i].dz) for i in range(n)]
    assert offs == sorted(offs)                        # dx, dy, dz ascending
    assert all(o[2] == 0 for o in offs) == (Z == 1)    # dz == 0 iff 2D
    nn = [o for o in offs if sum(c * c for c in o) == 1]
    assert len(nn) == (4 if Z == 1 else 6)
    for i in range(n):
        assert nb[i].d == np.float32(math.sqrt(offs[i][0] ** 2 + offs[i][1] ** 2 + offs[i][2] ** 2))


def test_product_neighbour_table_matches_oracle(oracle, sn):
    for rc, Z in ((3, 10), (3, 1), (4, 10)):
        nb, n = oracle.neighbours(rc, Z)
        assert sn.gen_neighbour(rc, Z) == [(nb[i].dx, nb[i].dy, nb[i].dz) for i in range(n)]


# ---------------------------------------------------------------- site_energy
def _rand_lattice(N, rng, mixed=False):
    v = rng.normal(size=(N, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    L = np.ones((N, 1))
    if mixed:
        L[rng.random(N) < 0.3] = 0.5
        L[rng.random(N) < 0.1] = 0.0
    return np.concatenate([v, L], 1)


def test_ferroelectric_flip_closed_form_3d(oracle):
    # all +x, flip one to -x: the dipole sum over a cubic shell vanishes, the
    # six cage bonds give 12 J, the field term -2 Ex.
    for J, Ex in ((1.0, 0.0), (0.7, 0.3), (2.0, -0.1)):
        o = oracle.Oracle(9, 9, 9, 3, J=J, E=(Ex, 0, 0))
        lat = np.zeros((729, 4)); lat[:, 0] = 1; lat[:, 3] = 1
        dE, _ = o.site_energy_batch(lat, [100], [[-1, 0, 0]])
        assert abs(dE[0] - (12 * J - 2 * Ex)) < 1e-12


def test_ferroelectric_flip_2d_brute_force(oracle):
    # 2D: 4 cage bonds and a dipole sum that does not vanish:
    # dE = 8 J - 2 Ex + 2 sum_r (3 n_x^2 - 1)/d^3  (in-plane shell)
    J, Ex = 1.0, 0.05
    o = oracle.Oracle(25, 25, 1, 3, J=J, E=(Ex, 0, 0))
    lat = np.zeros((625, 4)); lat[:, 0] = 1; lat[:, 3] = 1
    dE, _ = o.site_energy_batch(lat, [312], [[-1, 0, 0]])
    s = 0.0
    for dx in range(-3, 4):
        for dy in range(-3, 4):
            r2 = dx * dx + dy * dy
            if 0 < r2 <= 9:
                d = math.sqrt(r2)
                s += (3 * (dx / d) ** 2 - 1) / d ** 3
    assert abs(dE[0] - (8 * J - 2 * Ex + 2 * s)) < 1e-12


def test_two_dipole_textbook(oracle):
    # lone neighbour at r = (2,1,0): E = [p1.p2 - 3 (n.p1)(n.p2)]/d^3
    o = oracle.Oracle(12, 12, 12, 3, J=0.0)
    lat = np.zeros((12 ** 3, 4))
    lat[:, 2] = 1.0     # orientation irrelevant where length is 0
    i, j = (5 * 12 + 5) * 12 + 5, (7 * 12 + 6) * 12 + 5
    p1 = np.array([0.6, 0.0, 0.8]); p2 = np.array([0.0, 1.0, 0.0]); pn = np.array([1.0, 0, 0])
    lat[i, :3] = p1; lat[j, :3] = p2; lat[i, 3] = lat[j, 3] = 1.0
    n = np.array([2.0, 1.0, 0.0]) / math.sqrt(5)
    e = lambda a: (a @ p2 - 3 * (n @ a) * (n @ p2)) / 5 ** 1.5
    dE, _ = o.site_energy_batch(lat, [i], [pn])
    assert abs(dE[0] - (e(pn) - e(p1))) < 1e-14


def test_site_energy_invariants(oracle):
    rng = np.random.default_rng(3)
    o = oracle.Oracle(10, 10, 10, 3, J=0.8, E=(0.1, 0.2, -0.3), K=0.4)
    lat = _rand_lattice(1000, rng, mixed=True)
    sites = rng.integers(0, 1000, 200)
    props = _rand_lattice(200, rng)[:, :3]
    # dE(p -> p) = 0
    same, _ = o.site_energy_batch(lat, sites, lat[sites, :3])
    assert np.abs(same).max() == 0.0
    # antisymmetry: dE(p -> q) on lattice A == -dE(q -> p) on lattice with q placed
    for s, q in zip(sites[:50], props[:50]):
        a, _ = o.site_energy_batch(lat, [s], [q])
        l2 = lat.copy(); old = l2[s, :3].copy(); l2[s, :3] = q
        b, _ = o.site_energy_batch(l2, [s], [old])
        assert abs(a[0] + b[0]) < 1e-12


def test_site_energy_translation_and_cubic_rotation(oracle):
    rng = np.random.default_rng(5)
    X = 10
    o = oracle.Oracle(X, X, X, 3, J=1.1, E=(0, 0, 0))
    lat = _rand_lattice(X ** 3, rng, mixed=True).reshape(X, X, X, 4)
    q = np.array([0.0, 0.6, 0.8])
    ref, _ = o.site_energy_batch(lat.reshape(-1, 4), [(2 * X + 3) * X + 4], [q])
    # translate by (3, 5, 7)
    t = np.roll(lat, (3, 5, 7), axis=(0, 1, 2))
    v, _ = o.site_energy_batch(t.reshape(-1, 4), [(5 * X + 8) * X + 1], [q])
    assert abs(v[0] - ref[0]) < 1e-12
    # 90-degree rotation about z: (x, y) -> (-y, x) for positions and vectors
    rot = np.rot90(lat, k=1, axes=(0, 1)).copy()
    R = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    rot[..., :3] = rot[..., :3] @ R.T
    # site (2,3,4) moves to (X-1-3, 2, 4) under np.rot90 with axes (0,1)
    s2 = ((X - 1 - 3) * X + 2) * X + 4
    v, _ = o.site_energy_batch(rot.reshape(-1, 4), [s2], [R @ q])
    assert abs(v[0] - ref[0]) < 1e-12


@pytest.mark.parametrize("dims,rc", [((9, 9, 9), 3), ((12, 12, 1), 3), ((13, 13, 13), 4)])
def test_site_energy_brute_force_total_energy(oracle, dims, rc):
    rng = np.random.default_rng(11)
    X, Y, Z = dims
    N = X * Y * Z
    o = oracle.Oracle(X, Y, Z, rc, J=0.9, E=(0.05, -0.1, 0.2), K=0.3)
    lat = _rand_lattice(N, rng, mixed=True)
    e0, _ = o.energy(lat)
    for s in rng.integers(0, N, 8):
        q = _rand_lattice(1, rng)[0, :3]
        dE, sc = o.site_energy_batch(lat, [s], [q])
        l2 = lat.copy(); l2[s, :3] = q
        e1, _ = o.energy(l2)
        assert abs(dE[0] - (e1 - e0)) < 1e-10 * max(1.0, abs(e0))


def test_float_faithful_matches_fp64(oracle):
    """The float restatement agrees with the FP64 one to float accuracy."""
    lat32, _ = oracle.random_lattice_f32(12, 12, 12, 0xDEADBEEF + 300)
    o = oracle.Oracle(12, 12, 12, 3, J=1.0, E=(0.01, 0.0, 0.05))
    lat64 = lat32.astype(np.float64)
    rng = np.random.default_rng(2)
    worst = 0.0
    for s in rng.integers(0, 1728, 100):
        z, y, x = s % 12, (s // 12) % 12, s // 144
        q = _rand_lattice(1, rng)[0].astype(np.float32)
        q[3] = lat32[s, 3]
        f = o.site_energy_f32(lat32, x, y, z, q, lat32[s])
        d, sc = o.site_energy_batch(lat64, [s], [q[:3].astype(np.float64)])
        worst = max(worst, abs(f - d[0]) / max(sc[0], 1e-30))
    assert worst < 4e-6


def test_mc_moves_float_trace_consistency(oracle):
    """Trace fields satisfy the reference's acceptance rule, and the FP64 dE of
    each recorded move agrees with its float dE to the float bound."""
    lat32, mt = oracle.random_lattice_f32(8, 8, 8, 0xDEADBEEF + 300, (1.0, 0.0), (0.5, 0.5))
    o = oracle.Oracle(8, 8, 8, 3, J=1.0, beta=oracle.beta_of_T(300))
    start = lat32.copy()
    res = o.mc_moves_f32(lat32, mt, 3000, trace=True)
    tr = res["trace"]
    live = tr[tr["vacancy"] == 0]
    assert len(live) == res["accept"] + res["reject"]
    assert ((live["dE"] < 0) == (live["u_drawn"] == 0)).all()
    drawn = live[live["u_drawn"] == 1]
    assert np.array_equal(drawn["accepted"] == 1, np.exp(-drawn["dE"].astype(np.float64)) > drawn["u"])
    # vacancies: sites with length 0 are never moved
    vac = start[:, 3] == 0
    assert np.array_equal(lat32[vac], start[vac])


# ---------------------------------------------------------------- counter draw
def test_counter_draw_distribution(oracle):
    n = 20000
    P = np.empty((n, 3)); U = np.empty(n)
    for i in range(n):
        P[i], U[i] = oracle.counter_draw(0x1234567890ABCDEF, 2, 7, i)
    assert np.abs(np.linalg.norm(P, axis=1) - 1).max() < 4e-16
    assert np.abs(P.mean(0)).max() < 4 * math.sqrt(1 / 3 / n)
    C_ = P.T @ P / n
    assert np.abs(C_ - np.eye(3) / 3).max() < 0.02
    assert 0 <= U.min() and U.max() < 1 and abs(U.mean() - 0.5) < 4 * math.sqrt(1 / 12 / n)


# ---------------------------------------------------------------- energy / gradient
def test_gradient_matches_finite_differences(oracle):
    rng = np.random.default_rng(9)
    X = 9
    o = oracle.Oracle(X, X, X, 3, J=0.7, E=(0.1, -0.2, 0.3), K=0.0)
    lat = _rand_lattice(X ** 3, rng, mixed=True)
    g, dp = o.gradient(lat)
    # per-site field: local energy is linear in p_i (K = 0) -> exact
    for s in rng.integers(0, X ** 3, 10):
        z, y, x = s % X, (s // X) % X, s // X // X
        for k in range(3):
            e = np.zeros(3); e[k] = 1.0
            fd = o.local_energy(lat, x, y, z, lat[s, :3] + e) - o.local_energy(lat, x, y, z, lat[s, :3])
            assert abs(fd - g[s, k]) < 1e-12 * max(1, abs(g[s, k]))
    # dE = (p' - p) . dE/dp_i for a single-site change (K = 0)
    for s in rng.integers(0, X ** 3, 10):
        q = _rand_lattice(1, rng)[0, :3]
        dE, _ = o.site_energy_batch(lat, [s], [q])
        assert abs(dE[0] - (q - lat[s, :3]) @ g[s]) < 1e-12 * 50
    # parameter gradients by central differences of the compensated total
    h = 1e-6
    def E(J=0.7, Ef=(0.1, -0.2, 0.3), s=1.0):
        oo = oracle.Oracle(X, X, X, 3, J=J, E=Ef)
        l2 = lat.copy(); l2[:, 3] *= s
        return oo.energy(l2)[0]
    assert abs((E(J=0.7 + h) - E(J=0.7 - h)) / (2 * h) - dp[0]) < 1e-7 * max(1, abs(dp[0]))
    for k in range(3):
        ep = [0.1, -0.2, 0.3]; em = list(ep); ep[k] += h; em[k] -= h
        assert abs((E(Ef=ep) - E(Ef=em)) / (2 * h) - dp[1 + k]) < 1e-7 * max(1, abs(dp[1 + k]))
    assert abs((E(s=1 + h) - E(s=1 - h)) / (2 * h) - dp[4]) < 1e-7 * max(1, abs(dp[4]))


def test_energy_parts_closed_form_ferro(oracle):
    # all +x, length 1: E_cage = -J * N * 3 (three bonds per site), E_field = N Ex,
    # E_dd = 0 by cubic symmetry of each shell
    X = 9
    o = oracle.Oracle(X, X, X, 3, J=1.5, E=(0.2, 0, 0))
    lat = np.zeros((X ** 3, 4)); lat[:, 0] = 1; lat[:, 3] = 1
    e, parts = o.energy(lat)
    assert abs(parts[0]) < 1e-9
    assert abs(parts[1] - (-1.5 * 3 * X ** 3)) < 1e-9
    assert abs(parts[2] - 0.2 * X ** 3) < 1e-9


# ---------------------------------------------------------------- observables
def _r3(n):
    c = 0
    for x in range(-9, 10):
        for y in range(-9, 10):
            for z in range(-9, 10):
                c += x * x + y * y + z * z == n
    return c


def test_radial_order_count_invariant(oracle):
    X = 20
    o = oracle.Oracle(X, X, X, 3)
    lat = _rand_lattice(X ** 3, np.random.default_rng(0))
    fe, afe, cnt = o.radial_order(lat)
    N = X ** 3
    for r2 in (1, 2, 3, 4, 5, 9, 27, 80):
        assert cnt[r2] == N * _r3(r2)
    # the reference's RDF files hold the same count column (N = 8000)
    ref = "/root/reference/data/2015-01_Radial_Order_Work/radial_order_parameter_10000MC_moves_20x20x20_Strain1.dat"
    if os.path.exists(ref):
        rows = [l.split() for l in open(ref) if l.strip() and not l.startswith("#")]
        for r in rows[:20]:
            r2, count = int(r[0]), int(r[4])
            assert count == cnt[r2]
    # r = 0 bin: FE correlation is |p|^2 = 1
    assert abs(fe[0] - 1.0) < 1e-12 and cnt[0] == N


def _shell(b):
    return [(x, y, z) for x in range(-9, 10) for y in range(-9, 10) for z in range(-9, 10)
            if x * x + y * y + z * z == b]


@pytest.mark.parametrize("dims", [(10, 10, 10), (12, 11, 1)])
def test_radial_order_ferroelectric_closed_form(oracle, dims):
    """All dipoles along x: FE = 1 in every bin; AFE = 1 - 3<n_x^2> = 0 for
    r^2 > 0 (each shell is symmetric under coordinate permutations, so
    <n_x^2> = 1/3) and AFE = 1 at r = 0 (n = 0)."""
    X, Y, Z = dims
    lat = np.zeros((X * Y * Z, 4)); lat[:, 0] = 1.0; lat[:, 3] = 1.0
    fe, afe, cnt = oracle.Oracle(X, Y, Z, 3).radial_order(lat)
    occ = cnt > 0
    assert np.abs(fe[occ] - 1.0).max() < 1e-12
    assert abs(afe[0] - 1.0) < 1e-12 and np.abs(afe[1:][occ[1:]]).max() < 1e-12


@pytest.mark.parametrize("dims", [(10, 10, 10), (10, 9, 1), (8, 9, 9)])
def test_radial_order_layered_closed_form(oracle, dims):
    """p = (-1)^x x-hat (X even): a pair at offset r correlates as (-1)^dx,
    so FE[r^2] = <(-1)^dx> and AFE[r^2] = <(-1)^dx (1 - 3 dx^2/r^2)> over the
    shell; with Z == 1 the dz range still runs (onto the same plane), and an
    axis shorter than 9 wraps periodically.  Also checks the pair counts."""
    X, Y, Z = dims
    N = X * Y * Z
    x = np.arange(N) // (Y * Z)
    lat = np.zeros((N, 4)); lat[:, 0] = np.where(x % 2 == 0, 1.0, -1.0); lat[:, 3] = 1.0
    fe, afe, cnt = oracle.Oracle(X, Y, Z, 3).radial_order(lat)
    for b in range(81):
        sh = _shell(b)
        assert cnt[b] == N * len(sh)
        if not sh:
            continue
        sgn = np.array([(-1) ** (dx % 2) for dx, _, _ in sh], float)
        nx2 = np.array([dx * dx / b if b else 0.0 for dx, _, _ in sh])
        assert abs(fe[b] - sgn.mean()) < 1e-12
        assert abs(afe[b] - (sgn * (1 - 3 * nx2)).mean()) < 1e-12


def test_library_exports_every_declared_symbol(sn):
    """libstarrynight_amd.so loads (no device work) and exports every function
    include/starrynight_amd.h declares; the Python binding lists the same set."""
    import ctypes
    import re
    hdr = open(os.path.join(os.path.dirname(GOLDEN), "..", "include", "starrynight_amd.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(sn_\w+)\s*\(", hdr, re.M))
    assert len(declared) >= 25
    lib = ctypes.CDLL(sn.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == set(sn.EXPORTS)


@pytest.mark.parametrize("dims", [(14, 14, 14), (15, 13, 1)])
def test_dipole_potential_and_field_single_dipole(oracle, dims):
    """One dipole p at s (length L), every other site empty: the potential at
    i is the textbook L p.(s-i)/|s-i|^3 within |s-i| <= 6, the field at i is
    (3 n (n.p) - p)/d^3 with n = (s-i)/d within the cutoff (lengths not
    applied), and -p/3 at s itself; zero beyond the cutoffs."""
    X, Y, Z = dims
    N = X * Y * Z
    lat = np.zeros((N, 4))
    s = np.array([5, 6, 0 if Z == 1 else 7])
    p = np.array([0.3, -0.5, 0.81]); p /= np.linalg.norm(p)
    k = (s[0] * Y + s[1]) * Z + s[2]
    lat[k, :3] = p; lat[k, 3] = 0.7
    o = oracle.Oracle(X, Y, Z, 3)
    pot = o.dipole_potential(lat)
    E = o.dipole_efield(lat, 4)
    idx = np.arange(N)
    xyz = np.stack([idx // (Y * Z), (idx // Z) % Y, idx % Z], 1)
    d = s[None, :] - xyz
    d = (d + np.array([X, Y, Z]) // 2) % np.array([X, Y, Z]) - np.array([X, Y, Z]) // 2   # minimum image
    if Z == 1:
        # the dz range runs over -6..6 onto the same plane: each in-plane
        # offset counts once per dz with |r| <= cutoff
        want_pot = np.zeros(N); want_E = np.zeros((N, 3))
        for dz in range(-6, 7):
            r = np.concatenate([d[:, :2], np.full((N, 1), dz)], 1).astype(float)
            r2 = (r * r).sum(1)
            ok = (r2 > 0) & (r2 <= 36)
            want_pot[ok] += 0.7 * (r[ok] @ p) / r2[ok] ** 1.5
            ok4 = (r2 > 0) & (r2 <= 16)
            n = r[ok4] / np.sqrt(r2[ok4])[:, None]
            want_E[ok4] += (3 * n * (n @ p)[:, None] - p) / r2[ok4, None] ** 1.5
    else:
        r = d.astype(float)
        r2 = (r * r).sum(1)
        ok = (r2 > 0) & (r2 <= 36)
        want_pot = np.where(ok, 0.7 * (r @ p) / np.maximum(r2, 1) ** 1.5, 0.0)
        ok4 = (r2 > 0) & (r2 <= 16)
        n = r / np.sqrt(np.maximum(r2, 1))[:, None]
        want_E = np.where(ok4[:, None], (3 * n * (n @ p)[:, None] - p) / np.maximum(r2, 1)[:, None] ** 1.5, 0.0)
    want_E[k] -= p / 3
    assert np.abs(pot - want_pot).max() < 1e-13
    assert np.abs(E - want_E).max() < 1e-13


def test_dipole_potential_and_field_ferroelectric(oracle):
    """Uniform lattice: the potential sums p.r/|r|^3 over a symmetric shell (0),
    the field sums to zero by cubic symmetry, leaving -p/3 (|E| = 1/3)."""
    X = 12
    lat = np.zeros((X ** 3, 4)); lat[:, 0] = 1.0; lat[:, 3] = 1.0
    o = oracle.Oracle(X, X, X, 3)
    assert np.abs(o.dipole_potential(lat)).max() < 1e-13
    E = o.dipole_efield(lat, 3)
    assert np.abs(E - np.array([-1 / 3, 0, 0])).max() < 1e-13


def test_recombination_closed_forms(oracle):
    """Uniform lattice: the potential vanishes, so every Boltzmann factor is 1
    and every Fermi-Dirac occupation 1/2: ZB = N, ZFD = N/2, R_Boltz = 1,
    R_FD = 1 and both FD totals 1.  A single dipole: the sums follow from the
    (separately pinned) potential with numpy."""
    X = 10
    N = X ** 3
    lat = np.zeros((N, 4)); lat[:, 2] = 1.0; lat[:, 3] = 1.0
    o = oracle.Oracle(X, X, X, 3)
    r = o.recombination(lat)
    assert np.allclose(r, [N, N, N / 2, N / 2, 1.0, 1.0, 1.0, 1.0], rtol=1e-13, atol=0)
    lat = np.zeros((N, 4)); lat[123, :3] = (0.0, 0.6, 0.8); lat[123, 3] = 1.0
    pot = 0.165 / 5 * o.dipole_potential(lat)
    b = 40.0
    zbe, zbh = np.exp(-pot * b).sum(), np.exp(pot * b).sum()
    fe, fh = 1 / (np.exp(pot * b) + 1), 1 / (np.exp(-pot * b) + 1)
    want = [zbe, zbh, fe.sum(), fh.sum(), N * N / (zbe * zbh), N * ((fe / fe.sum()) * (fh / fh.sum())).sum(), 1, 1]
    assert np.allclose(o.recombination(lat), want, rtol=1e-12, atol=0)
