"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker / timed CPU baseline, never
by the product package.  See sn_oracle.h for what each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_MT_PATH = os.path.join(HERE, "_ref", "libref_mt.so")
CPU_BENCH = os.path.join(HERE, "sn_cpu_bench")


def build(ref: bool = False) -> None:
    """Compile the oracle (and, when the reference is mounted, oracle/_ref)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if ref and os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


class Neighbour(C.Structure):
    _fields_ = [("dx", C.c_int32), ("dy", C.c_int32), ("dz", C.c_int32), ("d", C.c_float)]


class Model(C.Structure):
    _fields_ = [("X", C.c_int32), ("Y", C.c_int32), ("Z", C.c_int32), ("rc", C.c_int32),
                ("J", C.c_double), ("E", C.c_double * 3), ("K", C.c_double),
                ("constrain_to_x", C.c_int32), ("beta", C.c_double)]


class MoveRec(C.Structure):
    _fields_ = [("site", C.c_int32), ("vacancy", C.c_int32), ("prop", C.c_float * 3),
                ("dE", C.c_float), ("u", C.c_double), ("u_drawn", C.c_int32),
                ("accepted", C.c_int32)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def load(path: str):
    """A liboracle.so built from sn_oracle.c (the default one, or a test's
    variant with a planted mistake), with its argument types declared."""
    L = C.CDLL(path)
    P = C.POINTER
    dp, fp, ip, u64p = P(C.c_double), P(C.c_float), P(C.c_int32), P(C.c_uint64)
    L.or_mt_seed.argtypes = [P(MT), C.c_uint32]
    L.or_mt_u32.argtypes = [P(MT)]; L.or_mt_u32.restype = C.c_uint32
    L.or_mt_real1.argtypes = [P(MT)]; L.or_mt_real1.restype = C.c_double
    L.or_mt_real2.argtypes = [P(MT)]; L.or_mt_real2.restype = C.c_double
    L.or_mt_fill_u32.argtypes = [P(MT), C.c_int64, P(C.c_uint32)]
    L.or_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
    L.or_gen_neighbour.argtypes = [C.c_int, C.c_int, P(Neighbour), C.c_int]
    L.or_gen_neighbour.restype = C.c_int
    L.or_site_energy_f32.argtypes = [P(Model), P(Neighbour), C.c_int, fp, C.c_int, C.c_int,
                                     C.c_int, fp, fp]
    L.or_site_energy_f32.restype = C.c_double
    L.or_random_sphere_point_f32.argtypes = [P(MT), fp]
    L.or_random_X_point_f32.argtypes = [P(MT), fp]
    L.or_init_lattice_random_f32.argtypes = [P(MT), C.c_int, C.c_int, C.c_int, fp]
    L.or_solid_solution_f32.argtypes = [P(MT), C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp]
    L.or_solid_solution_f32.restype = C.c_int
    L.or_species_walk_f32.argtypes = [C.c_float, C.c_int, fp]
    L.or_species_walk_f32.restype = C.c_int
    L.or_mc_moves_f32.argtypes = [P(Model), P(Neighbour), C.c_int, fp, P(MT), C.c_int64,
                                  u64p, u64p, P(MoveRec)]
    L.or_mc_moves_f32.restype = C.c_int64
    L.or_site_energy_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, C.c_int, C.c_int,
                                     C.c_int, dp, dp]
    L.or_site_energy_f64.restype = C.c_double
    L.or_site_energy_batch_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, C.c_int64, ip,
                                           dp, dp, dp]
    L.or_mc_moves_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, P(MT), C.c_int64,
                                  u64p, u64p]
    L.or_mc_moves_f64.restype = C.c_int64
    L.or_counter_draw.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                  dp, dp]
    L.or_sweep_ordered_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, ip, C.c_int64,
                                       C.c_uint64, C.c_uint32, C.c_uint32, u64p, u64p]
    L.or_energy_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, dp]
    L.or_energy_f64.restype = C.c_double
    L.or_gradient_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, dp, dp]
    L.or_local_energy_f64.argtypes = [P(Model), P(Neighbour), C.c_int, dp, C.c_int, C.c_int,
                                      C.c_int, dp]
    L.or_local_energy_f64.restype = C.c_double
    L.or_moment_sum.argtypes = [dp, C.c_int64, dp]
    L.or_radial_order_f64.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, P(C.c_int64)]
    L.or_dipole_potential_f64.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    L.or_dipole_efield_f64.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
    L.or_recombination_f64.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    return L


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# ---------------------------------------------------------------- RNGs
class MT19937:
    """MT19937 seeded like init_genrand (src/mt19937ar-cok.c:63)."""

    def __init__(self, seed: int):
        self.s = MT()
        lib().or_mt_seed(C.byref(self.s), C.c_uint32(seed & 0xFFFFFFFF))

    def u32(self, n: int) -> np.ndarray:
        out = np.empty(n, np.uint32)
        lib().or_mt_fill_u32(C.byref(self.s), n, _p(out, C.c_uint32))
        return out

    def real1(self) -> float:
        return lib().or_mt_real1(C.byref(self.s))

    def real2(self) -> float:
        return lib().or_mt_real2(C.byref(self.s))


def philox4x32_10(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, np.uint32).copy()
    k = np.asarray(key, np.uint32).copy()
    out = np.empty(4, np.uint32)
    lib().or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def ref_mt_stream(seed: int, n: int, kind: int) -> np.ndarray | None:
    """The reference's own MT19937 (oracle/_ref), or None when not built."""
    if not os.path.exists(REF_MT_PATH):
        return None
    L = C.CDLL(REF_MT_PATH)
    L.ref_mt_stream.argtypes = [C.c_uint32, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    out = np.empty(n, np.float64)
    L.ref_mt_stream(seed & 0xFFFFFFFF, n, kind, _p(out, C.c_double))
    return out


# ---------------------------------------------------------------- model
def neighbours(rc: int, Z: int):
    arr = (Neighbour * 10000)()
    n = lib().or_gen_neighbour(rc, Z, arr, 10000)
    if n < 0:
        raise ValueError("neighbour table overflow")
    return arr, n


def beta_of_T(T) -> float:
    """beta = 1/((float)T/300.0) (src/starrynight-main.c:215); T=0 gives +inf."""
    t = float(np.float32(T)) / 300.0
    return float("inf") if t == 0.0 else 1.0 / t


class Oracle:
    """One lattice model: dimensions, cutoff, couplings, beta."""

    def __init__(self, X, Y, Z, rc=3, J=1.0, E=(0.0, 0.0, 0.0), K=0.0, constrain_to_x=False,
                 beta=1.0):
        self.X, self.Y, self.Z, self.rc = X, Y, Z, rc
        self.N = X * Y * Z
        self.m = Model(X, Y, Z, rc, J, (C.c_double * 3)(*E), K, int(constrain_to_x), beta)
        self.nb, self.nnb = neighbours(rc, Z)

    # -- float-faithful
    def site_energy_f32(self, lat32, x, y, z, newp4, oldp4) -> float:
        newp = np.ascontiguousarray(newp4, np.float32)
        oldp = np.ascontiguousarray(oldp4, np.float32)
        return lib().or_site_energy_f32(C.byref(self.m), self.nb, self.nnb, _p(lat32, C.c_float),
                                        x, y, z, _p(newp, C.c_float), _p(oldp, C.c_float))

    def mc_moves_f32(self, lat32, mt: MT19937, moves: int, trace: bool = False):
        acc, rej = C.c_uint64(0), C.c_uint64(0)
        recs = (MoveRec * moves)() if trace else None
        lib().or_mc_moves_f32(C.byref(self.m), self.nb, self.nnb, _p(lat32, C.c_float),
                              C.byref(mt.s), moves, C.byref(acc), C.byref(rej), recs)
        out = {"accept": acc.value, "reject": rej.value}
        if trace:
            a = np.ctypeslib.as_array(recs)
            out["trace"] = a.copy()
        return out

    # -- FP64
    def site_energy_batch(self, lat64, sites, props):
        sites = np.ascontiguousarray(sites, np.int32)
        props = np.ascontiguousarray(props, np.float64)
        n = len(sites)
        dE = np.empty(n, np.float64)
        sc = np.empty(n, np.float64)
        lib().or_site_energy_batch_f64(C.byref(self.m), self.nb, self.nnb, _p(lat64, C.c_double),
                                       n, _p(sites, C.c_int32), _p(props, C.c_double),
                                       _p(dE, C.c_double), _p(sc, C.c_double))
        return dE, sc

    def mc_moves_f64(self, lat64, mt: MT19937, moves: int):
        acc, rej = C.c_uint64(0), C.c_uint64(0)
        lib().or_mc_moves_f64(C.byref(self.m), self.nb, self.nnb, _p(lat64, C.c_double),
                              C.byref(mt.s), moves, C.byref(acc), C.byref(rej))
        return acc.value, rej.value

    def sweep_ordered(self, lat64, order, seed, replica, sweep):
        order = np.ascontiguousarray(order, np.int32)
        acc, rej = C.c_uint64(0), C.c_uint64(0)
        lib().or_sweep_ordered_f64(C.byref(self.m), self.nb, self.nnb, _p(lat64, C.c_double),
                                   _p(order, C.c_int32), len(order), C.c_uint64(seed), replica,
                                   sweep, C.byref(acc), C.byref(rej))
        return acc.value, rej.value

    def energy(self, lat64):
        parts = np.empty(4, np.float64)
        e = lib().or_energy_f64(C.byref(self.m), self.nb, self.nnb, _p(lat64, C.c_double),
                                _p(parts, C.c_double))
        return e, parts

    def gradient(self, lat64):
        g = np.empty((self.N, 3), np.float64)
        dp = np.empty(6, np.float64)
        lib().or_gradient_f64(C.byref(self.m), self.nb, self.nnb, _p(lat64, C.c_double),
                              _p(g, C.c_double), _p(dp, C.c_double))
        return g, dp

    def local_energy(self, lat64, x, y, z, p):
        p = np.ascontiguousarray(p, np.float64)
        return lib().or_local_energy_f64(C.byref(self.m), self.nb, self.nnb, _p(lat64, C.c_double),
                                         x, y, z, _p(p, C.c_double))

    def dipole_potential(self, lat64):
        pot = np.empty(self.N, np.float64)
        lib().or_dipole_potential_f64(self.X, self.Y, self.Z, _p(lat64, C.c_double), _p(pot, C.c_double))
        return pot

    def dipole_efield(self, lat64, cutoff):
        E = np.empty((self.N, 3), np.float64)
        lib().or_dipole_efield_f64(self.X, self.Y, self.Z, int(cutoff), _p(lat64, C.c_double), _p(E, C.c_double))
        return E

    def recombination(self, lat64):
        out = np.empty(8, np.float64)
        lib().or_recombination_f64(self.X, self.Y, self.Z, _p(lat64, C.c_double), _p(out, C.c_double))
        return out

    def radial_order(self, lat64):
        fe = np.empty(81); afe = np.empty(81); cnt = np.empty(81, np.int64)
        lib().or_radial_order_f64(self.X, self.Y, self.Z, _p(lat64, C.c_double), _p(fe, C.c_double),
                                  _p(afe, C.c_double), _p(cnt, C.c_int64))
        return fe, afe, cnt


def counter_draw(seed, replica, sweep, site, constrain_to_x=False):
    p = np.empty(3, np.float64)
    u = C.c_double(0)
    lib().or_counter_draw(C.c_uint64(seed), replica, sweep, site, int(constrain_to_x),
                          _p(p, C.c_double), C.byref(u))
    return p, u.value


def moment_sum(lat64) -> np.ndarray:
    M = np.empty(3)
    lib().or_moment_sum(_p(lat64, C.c_double), lat64.shape[0], _p(M, C.c_double))
    return M


def sphere_points_f32(mt: MT19937, n: int) -> np.ndarray:
    """n successive random_sphere_point calls (float32 [n, 3]) on `mt`."""
    out = np.zeros((n, 4), np.float32)
    for i in range(n):
        lib().or_random_sphere_point_f32(C.byref(mt.s), _p(out[i], C.c_float))
    return out[:, :3].copy()


def x_points_f32(mt: MT19937, n: int) -> np.ndarray:
    """n successive random_X_point calls (float32 [n, 3]) on `mt`."""
    out = np.zeros((n, 4), np.float32)
    for i in range(n):
        lib().or_random_X_point_f32(C.byref(mt.s), _p(out[i], C.c_float))
    return out[:, :3].copy()


def species_walk_f32(sample: float, prevalence) -> int:
    """solid_solution's walk of one float sample down the prevalence list:
    the species index, or -1 past the end of the table."""
    pv = np.ascontiguousarray(prevalence, np.float32)
    return lib().or_species_walk_f32(C.c_float(sample), len(pv), _p(pv, C.c_float))


def random_lattice_f32(X, Y, Z, seed, lengths=(1.0,), prevalence=(1.0,)):
    """initialise_lattice_random + solid_solution from init_genrand(seed);
    returns (float32 lattice [N,4], the generator positioned after init)."""
    mt = MT19937(seed)
    lat = np.zeros((X * Y * Z, 4), np.float32)
    lib().or_init_lattice_random_f32(C.byref(mt.s), X, Y, Z, _p(lat, C.c_float))
    ln = np.asarray(lengths, np.float32)
    pv = np.asarray(prevalence, np.float32)
    if lib().or_solid_solution_f32(C.byref(mt.s), X, Y, Z, len(ln), _p(ln, C.c_float),
                                   _p(pv, C.c_float), _p(lat, C.c_float)) != 0:
        raise ValueError("Prevalence list sums below 1: the reference would read past it")
    return lat, mt
