"""The reference's initial state and float-faithful proposal draws, on the CPU.

Pins for the oracle's float restatements of
  random_sphere_point  (src/starrynight-config.c:203-227),
  random_X_point       (src/starrynight-config.c:230-263),
  solid_solution       (src/starrynight-lattice.c:139-171),
and for the product's host initialiser sn_reference_initial_lattice (the
InitialLattice patterns, src/starrynight-lattice.c:25-137, chosen as in
src/starrynight-main.c:184-205), plus the C host's argv overrides
(src/starrynight-main.c:142-151, seed 0xDEADBEEF + T at :172-176).

The MT19937 stream under all of them is pinned separately against the
reference's own src/mt19937ar-cok.c (test_oracle.py); where that build is
present, the genrand_real1 values fed to the closed forms below come from it.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "starrynight_amd", "sn_host")


def real1_stream(oracle, seed, n):
    """genrand_real1 values of init_genrand(seed): the reference's own MT
    (oracle/_ref) when built, else the oracle's (pinned against it)."""
    ref = oracle.ref_mt_stream(seed, n, 1)
    if ref is not None:
        return ref
    return oracle.MT19937(seed).u32(n).astype(np.float64) * (1.0 / 4294967295.0)


def draws_consumed(oracle, seed, mt_after, horizon):
    """How many u32 draws a generator seeded `seed` has made, from where the
    next three outputs of `mt_after` sit in the fresh stream."""
    nxt = mt_after.u32(3)
    s = oracle.MT19937(seed).u32(horizon + 3)
    hits = np.flatnonzero((s[:-2] == nxt[0]) & (s[1:-1] == nxt[1]) & (s[2:] == nxt[2]))
    assert len(hits) == 1
    return int(hits[0])


def marsaglia_from_real1(r):
    """random_sphere_point on a stream of genrand_real1 values r, in the
    reference's types: x1, x2 = 2 r - 1 computed in double and stored as
    float; rejected while the float x1*x1 + x2*x2 exceeds 1; then
    p = (2 x1 sqrt(1-s), 2 x2 sqrt(1-s), 1 - 2 s) with float products and sums,
    sqrt of the widened float, each component stored as float.  Returns the
    points and the number of real1 values consumed."""
    f32 = np.float32
    pts, k, used = [], 0, 0
    while k + 1 < len(r):
        x1 = f32(2.0 * r[k] - 1.0)
        x2 = f32(2.0 * r[k + 1] - 1.0)
        k += 2
        s = x1 * x1 + x2 * x2                      # float * float + float * float
        if float(s) > 1.0:
            continue
        root = np.sqrt(np.float64((f32(1) - x1 * x1) - x2 * x2))
        px = f32(np.float64(f32(2) * x1) * root)
        py = f32(np.float64(f32(2) * x2) * root)
        pz = f32(1.0 - 2.0 * np.float64(s))
        pts.append((px, py, pz))
        used = k
    return np.array(pts, np.float32), used


# ---------------------------------------------------------------- sphere point
def check_sphere_closed_form(oracle):
    """Marsaglia's mapping on the reference's genrand_real1 stream, bit for
    bit, and the same number of draws consumed."""
    seed = 0xDEADBEEF + 300
    r = real1_stream(oracle, seed, 40000)
    want, used = marsaglia_from_real1(r)
    mt = oracle.MT19937(seed)
    got = oracle.sphere_points_f32(mt, len(want))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert draws_consumed(oracle, seed, mt, 50000) == used


def check_sphere_invariants(oracle):
    """|p| = 1 to float rounding; E[p] = 0 and E[p p^T] = I/3 (uniform on the
    sphere); p_z uniform on [-1, 1] (Archimedes); 8/pi draws per point on
    average (two per try, acceptance pi/4)."""
    seed, n = 12345, 120000
    mt = oracle.MT19937(seed)
    p = oracle.sphere_points_f32(mt, n).astype(np.float64)
    assert np.abs(np.linalg.norm(p, axis=1) - 1.0).max() < 4e-7
    assert np.abs(p.mean(0)).max() < 5 * np.sqrt(1.0 / 3.0 / n)
    second = p.T @ p / n
    # Var(p_a p_b) = 1/15 off the diagonal, 4/45 on it
    assert np.abs(second - np.eye(3) / 3).max() < 5 * np.sqrt(4.0 / 45.0 / n)
    hist, _ = np.histogram(p[:, 2], bins=20, range=(-1, 1))
    chi2 = ((hist - n / 20) ** 2 / (n / 20)).sum()
    assert chi2 < 50          # 19 dof: P(chi2 > 50) ~ 1e-4
    used = draws_consumed(oracle, seed, mt, int(3.5 * n))
    # tries per point ~ geometric(pi/4): draws 2 * tries, variance 4 (1 - q) / q^2
    q = np.pi / 4
    assert abs(used / n - 8 / np.pi) < 5 * np.sqrt(4 * (1 - q) / q ** 2 / n)


def test_sphere_point_closed_form_given_x1_x2(oracle):
    check_sphere_closed_form(oracle)


def test_sphere_point_distribution_invariants(oracle):
    check_sphere_invariants(oracle)


# ---------------------------------------------------------------- X point
AXES = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)], np.float32)


def check_x_table(oracle):
    """random_X_point: rand_int(6) = genrand_int32() % 6 picks +x, -x, +y,
    -y, +z, -z in the order of the reference's switch (config.c:238-257);
    one draw per call."""
    seed, n = 99, 3000
    y = oracle.MT19937(seed).u32(n)
    mt = oracle.MT19937(seed)
    got = oracle.x_points_f32(mt, n)
    assert np.array_equal(got, AXES[y % 6])
    assert draws_consumed(oracle, seed, mt, n + 10) == n


def check_x_frequencies(oracle):
    """Every call is a signed unit axis, each of the six at 1/6."""
    n = 60000
    got = oracle.x_points_f32(oracle.MT19937(2024), n)
    assert np.all(np.abs(got).sum(1) == 1) and np.all(np.abs(got).max(1) == 1)
    idx = [int(np.flatnonzero((AXES == g).all(1))[0]) for g in got[:6000]]
    counts = np.bincount(idx, minlength=6)
    assert ((counts - 1000) ** 2 / 1000).sum() < 25   # 5 dof
    assert np.abs(got.mean(0)).max() < 5 * np.sqrt(1 / 3 / n)


def test_x_point_one_draw_per_call_and_axis_table(oracle):
    check_x_table(oracle)


def test_x_point_frequencies(oracle):
    check_x_frequencies(oracle)


# ---------------------------------------------------------------- solid solution
WALK_CASES = [
    ((0.25, 0.25, 0.5), 0.0, 0),
    ((0.25, 0.25, 0.5), 0.25, 0),                                   # equality stays: sample > prev is false
    ((0.25, 0.25, 0.5), float(np.nextafter(np.float32(0.25), np.float32(1))), 1),
    ((0.25, 0.25, 0.5), 0.5, 1),
    ((0.25, 0.25, 0.5), float(np.nextafter(np.float32(0.5), np.float32(1))), 2),
    ((0.25, 0.25, 0.5), 1.0, 2),
    ((0.5, 0.0, 0.5), float(np.nextafter(np.float32(0.5), np.float32(1))), 2),   # a zero-prevalence entry is skipped
    ((0.0, 1.0), 0.0, 0),                                          # ... except by sample 0 exactly
    ((0.0, 1.0), float(np.nextafter(np.float32(0), np.float32(1))), 1),
    ((0.25, 0.25), 0.75, -1),                                       # table summing below the sample: past the end
    ((0.5,) * 2, 1.0, 1),
    ((0.1,) * 10, 0.0, 0),
]


def check_walk_boundaries(oracle):
    """solid_solution's walk (lattice.c:154-159) on planted float samples:
    `sample > prevalence` steps on, so a sample equal to a cumulative
    boundary stays on the lower species."""
    for prev, sample, want in WALK_CASES:
        assert oracle.species_walk_f32(sample, prev) == want, (prev, sample, want)


SOLID_TABLES = [
    ((1.0, 0.0), (0.5, 0.5)),
    ((1.0, 0.5, 0.8), (0.2, 0.3, 0.5)),
    (tuple(0.1 * (k + 1) for k in range(10)), (0.1,) * 10),
]


def check_solid_solution(oracle):
    """One genrand_real1 per site in lattice order, narrowed to float and
    walked down the table; species counts binomial about N * prevalence."""
    import ctypes as C
    X, Y, Z = 24, 20, 16
    N = X * Y * Z
    seed = 777
    for lengths, prev in SOLID_TABLES:
        mt = oracle.MT19937(seed)
        lat = np.zeros((N, 4), np.float32)
        ln, pv = np.asarray(lengths, np.float32), np.asarray(prev, np.float32)
        assert oracle.lib().or_solid_solution_f32(C.byref(mt.s), X, Y, Z, len(ln), oracle._p(ln, C.c_float),
                                                  oracle._p(pv, C.c_float), oracle._p(lat, C.c_float)) == 0
        assert draws_consumed(oracle, seed, mt, N + 10) == N
        r = real1_stream(oracle, seed, N)
        idx = np.array([oracle.species_walk_f32(float(np.float32(v)), prev) for v in r])
        assert np.array_equal(lat[:, 3], ln[idx])
        counts = np.bincount(idx, minlength=len(prev))
        for c, q in zip(counts, prev):
            assert abs(c - N * q) < 5 * np.sqrt(N * q * (1 - q)) + 1, (prev, counts)


def test_species_walk_at_the_prevalence_boundaries(oracle):
    check_walk_boundaries(oracle)


def test_solid_solution_one_draw_per_site_and_frequencies(oracle):
    check_solid_solution(oracle)


# ---------------------------------------------------------------- planted mistakes
# Plausible slips, planted by text substitution into a copy of sn_oracle.c
# compiled on the side; the pins above must reject every one.
PLANTS = {
    "sphere_real2_for_real1": ([("float x1, x2;\n    do {\n        x1 = 2.0 * or_mt_real1(g) - 1.0;",
                                 "float x1, x2;\n    do {\n        x1 = 2.0 * or_mt_real2(g) - 1.0;")],
                               [check_sphere_closed_form, check_sphere_invariants]),
    "sphere_dropped_factor_2_in_z": ([("p[2] = 1.0 - 2.0 * (x1 * x1 + x2 * x2);",
                                       "p[2] = 1.0 - (x1 * x1 + x2 * x2);")],
                                     [check_sphere_closed_form, check_sphere_invariants]),
    "sphere_swapped_x_y": ([("p[0] = 2 * x1 * sqrt", "p[0] = 2 * xT * sqrt"),
                            ("p[1] = 2 * x2 * sqrt", "p[1] = 2 * x1 * sqrt"),
                            ("p[0] = 2 * xT * sqrt", "p[0] = 2 * x2 * sqrt")],
                           [check_sphere_closed_form, check_sphere_invariants]),
    "x_point_axis_sign_order": ([("static const int ax[6][3] = {{1, 0, 0}, {-1, 0, 0},",
                                  "static const int ax[6][3] = {{-1, 0, 0}, {1, 0, 0},")],
                                [check_x_table, check_x_frequencies]),
    "x_point_span_5": ([("int i = rand_int(g, 6);\n    p[0]", "int i = rand_int(g, 5);\n    p[0]")],
                       [check_x_table, check_x_frequencies]),
    "walk_ge_for_gt": ([("sample > prev[i]; sample -= prev[i], i++)", "sample >= prev[i]; sample -= prev[i], i++)")],
                       [check_walk_boundaries, check_solid_solution]),
    "walk_subtracts_next_entry": ([("sample > prev[i]; sample -= prev[i], i++)",
                                    "sample > prev[i]; i++, sample -= prev[i])")],
                                  [check_walk_boundaries, check_solid_solution]),
}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_pins_reject_planted_mistakes(oracle, tmp_path, monkeypatch, plant):
    subs, checks = PLANTS[plant]
    src = open(os.path.join(ROOT, "oracle", "sn_oracle.c")).read()
    for old, new in subs:
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    (tmp_path / "sn_oracle.c").write_text(src)
    so = tmp_path / "libplant.so"
    subprocess.check_call(["gcc", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=gnu11", "-w",
                           "-I", os.path.join(ROOT, "oracle"), "-shared", "-o", str(so),
                           str(tmp_path / "sn_oracle.c"), "-lm"])
    monkeypatch.setattr(oracle, "_lib", oracle.load(str(so)))
    caught = []
    for chk in checks:
        try:
            chk(oracle)
        except AssertionError:
            caught.append(chk.__name__)
    assert caught, f"planted mistake {plant} passed every pin"


# ---------------------------------------------------------------- initial lattices
@pytest.mark.parametrize("dims,seed,lengths,prev", [
    ((16, 16, 16), 0xDEADBEEF + 300, (1.0,), (1.0,)),
    ((25, 25, 1), 0xDEADBEEF + 300, (1.0,), (1.0,)),
    ((12, 10, 14), 0xDEADBEEF + 250, (1.0, 0.5), (0.7, 0.3)),
    ((20, 20, 20), 0xDEADBEEF, (1.0, 0.0), (0.5, 0.5)),
])
def test_reference_initial_lattice_random_bit_exact(oracle, sn, dims, seed, lengths, prev):
    """sn_reference_initial_lattice("random") equals the oracle's
    initialise_lattice_random + solid_solution bit for bit (the float values,
    widened to double)."""
    got = sn.reference_initial_lattice(*dims, "random", seed=seed, Dipoles=lengths, Prevalence=prev)
    want, _ = oracle.random_lattice_f32(*dims, seed, lengths=lengths, prevalence=prev)
    assert got.dtype == np.float64 and np.array_equal(got, want.astype(np.float64))
    # any other name is the reference's default, the random lattice
    other = sn.reference_initial_lattice(*dims, "no_such_pattern", seed=seed, Dipoles=lengths, Prevalence=prev)
    assert np.array_equal(other, got)


def pattern_closed_form(name, X, Y, Z):
    """The deterministic InitialLattice patterns (lattice.c:39-137) written
    as formulas of (x, y, z)."""
    x, y, z = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    zero = np.zeros(x.shape)
    if name == "ferroelectric":
        return np.stack([np.ones(x.shape), zero, zero], 1)
    if name == "buckled":
        return np.stack([x % 2, y % 2, z % 2], 1).astype(float)
    if name == "antiferro_wall":
        a = (y < Y // 2) ^ (x > X // 2)
        return np.stack([np.where(a, 2 * ((z + y) % 2) - 1, 0), np.where(a, 0, 2 * ((x + z) % 2) - 1), zero], 1)
    if name == "ferro_wall":
        return np.stack([zero, np.where(x < X // 2, -1.0, 1.0), zero], 1)
    if name == "antiferro_slip":
        return np.stack([np.where(x < X // 2, 2 * ((z + y) % 2) - 1, 2 * ((z + y + 1) % 2) - 1), zero, zero], 1)
    if name == "spectrum":
        ang = 2 * np.pi * (x * X + y) / (X * Y)
        return np.stack([np.sin(ang), np.cos(ang), zero], 1)
    if name == "slab_delete":
        return np.stack([zero, zero, zero], 1)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["ferroelectric", "buckled", "antiferro_wall", "ferro_wall",
                                  "antiferro_slip", "spectrum", "slab_delete"])
@pytest.mark.parametrize("dims", [(8, 6, 4), (9, 7, 5), (10, 10, 1)])
def test_deterministic_patterns(oracle, sn, name, dims):
    """Each deterministic pattern against its formula, and the lengths of
    the solid solution that follows: the patterns draw nothing, so the
    solid solution starts at the seed's first genrand_real1."""
    X, Y, Z = dims
    seed = 0xDEADBEEF + 200
    got = sn.reference_initial_lattice(X, Y, Z, name, seed=seed, Dipoles=(1.0, 0.5), Prevalence=(0.6, 0.4))
    want = pattern_closed_form(name, X, Y, Z)
    # spectrum stores its angle (< 2 pi) as float, half an ulp <= 2.4e-7, then
    # rounds sin / cos to float: within 4e-7 of the exact formula
    tol = 4e-7 if name == "spectrum" else 0.0
    assert np.abs(got[:, :3] - want).max() <= tol
    idx = np.array([oracle.species_walk_f32(float(np.float32(v)), (0.6, 0.4))
                    for v in real1_stream(oracle, seed, X * Y * Z)])
    assert np.array_equal(got[:, 3], np.array([1.0, 0.5])[idx])


def test_pattern_invariants(sn):
    """Independent of the formulas: a ferroelectric lattice has moment N x,
    antiferro_slip and antiferro_wall cancel on even boxes, ferro_wall's
    halves cancel, spectrum is unit length."""
    X, Y, Z = 8, 8, 8
    N = X * Y * Z
    M = lambda name: sn.reference_initial_lattice(X, Y, Z, name)[:, :3].sum(0)
    assert np.array_equal(M("ferroelectric"), [N, 0, 0])
    assert np.array_equal(M("antiferro_slip"), [0, 0, 0])
    assert np.array_equal(M("ferro_wall"), [0, 0, 0])
    assert np.array_equal(M("antiferro_wall"), [0, 0, 0])
    s = sn.reference_initial_lattice(X, Y, Z, "spectrum")[:, :3]
    assert np.abs(np.linalg.norm(s, axis=1) - 1).max() < 1e-6


def test_prevalence_below_the_sample_is_refused(sn):
    """A Prevalence table summing below the drawn sample would send the
    reference past its Dipoles[] array; the library refuses it."""
    with pytest.raises(sn.SnError):
        sn.reference_initial_lattice(16, 16, 16, "random", Dipoles=(1.0, 0.5), Prevalence=(0.2, 0.2))


# ---------------------------------------------------------------- C host argv
CFG = """
T: 250
X = 12
Y = 10
Z = 8
InitialLattice = "random"
DipoleCutOff: 3
Dipoles    = [ 1.0, 0.5 ];
Prevalence = [ 0.7, 0.3 ];
CageStrain = 2.0
MCEqmSteps: 1
MCMegaSteps: 1
MCMoves: 1.0
"""


def dry_run(tmp_path, *args):
    (tmp_path / "starrynight.cfg").write_text(CFG)
    out = tmp_path / "init.bin"
    p = subprocess.run([HOST, "--dry-run", str(out), *args], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 0, p.stderr
    kv = dict(l.split("=", 1) for l in p.stdout.splitlines() if "=" in l and not l.startswith("replica"))
    reps = [dict(f.split("=", 1) for f in l.split()) for l in p.stdout.splitlines() if l.startswith("replica=")]
    lat = np.fromfile(out, np.float64).reshape(-1, 4)
    return kv, reps, lat, p.stderr


@pytest.mark.parametrize("args,T,J", [((), 250, 2.0), (("300",), 300, 2.0), (("125", "0.5"), 125, 0.5)])
def test_host_argv_overrides(oracle, tmp_path, args, T, J):
    """sn_host [T [CageStrain]]: argv[1] replaces T and argv[2] CageStrain
    (main.c:142-151); the MT19937 seed and the RNG key follow the new T
    (main.c:172), so the initial lattice is the oracle's for 0xDEADBEEF + T."""
    if not os.path.exists(HOST):
        pytest.skip("sn_host not built")
    kv, reps, lat, err = dry_run(tmp_path, *args)
    assert int(kv["T"]) == T and float(kv["CageStrain"]) == J
    assert len(reps) == 1 and float(reps[0]["temperature"]) == T
    assert int(reps[0]["mt_seed"]) == (0xDEADBEEF + T) & 0xFFFFFFFF == int(reps[0]["seed"])
    if args:
        assert f"Command line temperature: T = {T}" in err
    want, _ = oracle.random_lattice_f32(12, 10, 8, 0xDEADBEEF + T, lengths=(1.0, 0.5), prevalence=(0.7, 0.3))
    assert np.array_equal(lat, want.astype(np.float64))
