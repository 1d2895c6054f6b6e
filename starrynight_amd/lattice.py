"""Python face of the C ABI, named after the reference's hot-path functions.

    gen_neighbour()            src/starrynight-montecarlo-core.c:38-72
    site_energy(x,y,z,new)     src/starrynight-montecarlo-core.c:76-141
    MC_moves(moves)            src/starrynight-montecarlo-core.c:143-149 (whole sweeps)
    MC_sweeps(n)               n sweeps of N attempts each
    replay(trace)              MC_move decisions of a recorded reference run

Errors surface as SnError (the library returns status codes; nothing exits).
Every call runs on the GPU through libstarrynight_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from ._lib import Params, SnError, Tuning, _ptr, lib


def _check_device_tensor(t, shape):
    if not getattr(t, "is_cuda", False):
        raise TypeError("expected a GPU tensor")
    if str(t.dtype) != "torch.float64" or tuple(t.shape) != shape or not t.is_contiguous():
        raise ValueError(f"expected a contiguous float64 tensor of shape {list(shape)}")


def _sync_tensor_stream(t):
    import torch
    torch.cuda.current_stream(t.device).synchronize()


@dataclass
class ReplicaParams:
    """One replica's model (the reference's globals, src/starrynight-config.c:12-93)."""
    X: int
    Y: int
    Z: int
    DipoleCutOff: int = 3
    CageStrain: float = 1.0
    Efield: Sequence[float] = (0.0, 0.0, 0.0)
    K: float = 0.0
    ConstrainToX: bool = False
    T: float = 300.0
    seed: int = 0xDEADBEEF + 300

    def to_c(self) -> Params:
        p = Params()
        p.X, p.Y, p.Z = self.X, self.Y, self.Z
        p.dipole_cutoff = self.DipoleCutOff
        p.cage_strain = self.CageStrain
        p.efield = (C.c_double * 3)(*[float(v) for v in self.Efield])
        p.K = self.K
        p.constrain_to_x = int(bool(self.ConstrainToX))
        p.temperature = float(self.T)
        p.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        return p

    @property
    def N(self) -> int:
        return self.X * self.Y * self.Z


@dataclass
class TuningParams:
    tile: Sequence[int] = (0, 0, 0)
    lanes_per_site: int = 0
    fine_colouring: int = 0
    use_graph: int = 0
    kernel: int = 0   # 0 auto, 1 field-update, 2 per-phase gather
    block_threads: int = 0   # gather kernel: 0 / 256, or 512 threads per tile

    def to_c(self) -> Tuning:
        t = Tuning()
        t.tile = (C.c_int32 * 3)(*self.tile)
        t.lanes_per_site = self.lanes_per_site
        t.fine_colouring = self.fine_colouring
        t.use_graph = self.use_graph
        t.kernel = self.kernel
        t.block_threads = self.block_threads
        return t


class StarryNight:
    """A context of one or more replicas on the current GPU."""

    def __init__(self, replicas: ReplicaParams | Sequence[ReplicaParams],
                 tuning: TuningParams | None = None):
        if isinstance(replicas, ReplicaParams):
            replicas = [replicas]
        self.replicas = list(replicas)
        arr = (Params * len(self.replicas))(*[r.to_c() for r in self.replicas])
        h = C.c_void_p()
        tun = tuning.to_c() if tuning else None
        st = lib().sn_create(arr, len(self.replicas), C.byref(tun) if tun else None, C.byref(h))
        self._h = h
        if st != 0:
            msg = lib().sn_last_error(h).decode() if h else lib().sn_status_string(st).decode()
            if h:
                lib().sn_destroy(h)
                self._h = C.c_void_p()
            raise SnError(st, msg)

    # -- plumbing
    def _chk(self, st: int):
        if st != 0:
            raise SnError(st, lib().sn_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().sn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- lattice
    def set_lattice(self, lat: np.ndarray, replica: int = 0):
        a = np.ascontiguousarray(lat, np.float64)
        if a.shape != (self.replicas[replica].N, 4):
            raise ValueError(f"lattice must be [{self.replicas[replica].N}, 4]")
        self._chk(lib().sn_set_lattice(self._h, replica, _ptr(a, C.c_double)))

    def set_lattice_device(self, lat, replica: int = 0):
        """Load replica's lattice from a contiguous float64 [N, 4] torch tensor
        on the GPU (synchronises the tensor's stream first)."""
        _check_device_tensor(lat, (self.replicas[replica].N, 4))
        _sync_tensor_stream(lat)
        self._chk(lib().sn_set_lattice_device(self._h, replica, C.c_void_p(lat.data_ptr())))

    def get_lattice(self, replica: int = 0) -> np.ndarray:
        out = np.empty((self.replicas[replica].N, 4), np.float64)
        self._chk(lib().sn_get_lattice(self._h, replica, _ptr(out, C.c_double)))
        return out

    def init_lattice_random(self, replica: int = 0, Dipoles=(1.0,), Prevalence=(1.0,)):
        ln = np.asarray(Dipoles, np.float64)
        pv = np.asarray(Prevalence, np.float64)
        self._chk(lib().sn_init_lattice_random(self._h, replica, len(ln), _ptr(ln, C.c_double),
                                               _ptr(pv, C.c_double)))

    def set_params(self, replica: int, T: float, Efield=(0.0, 0.0, 0.0), CageStrain: float = 1.0):
        e = np.asarray(Efield, np.float64)
        self._chk(lib().sn_set_params(self._h, replica, float(T), _ptr(e, C.c_double), float(CageStrain)))
        r = self.replicas[replica]
        r.T, r.Efield, r.CageStrain = T, tuple(e), CageStrain

    # -- hot path
    def site_energy(self, x: int, y: int, z: int, new, replica: int = 0) -> float:
        r = self.replicas[replica]
        site = np.array([(x * r.Y + y) * r.Z + z], np.int32)
        return float(self.site_energy_batch(site, np.asarray(new, np.float64).reshape(1, 3), replica)[0])

    def site_energy_batch(self, sites, props, replica: int = 0) -> np.ndarray:
        s = np.ascontiguousarray(sites, np.int32)
        p = np.ascontiguousarray(props, np.float64).reshape(-1, 3)
        if len(s) != len(p):
            raise ValueError("sites and proposals differ in length")
        out = np.empty(len(s), np.float64)
        self._chk(lib().sn_site_energy_batch(self._h, replica, len(s), _ptr(s, C.c_int32),
                                             _ptr(p, C.c_double), _ptr(out, C.c_double)))
        return out

    def MC_sweeps(self, nsweeps: int, sync: bool = True):
        self._chk(lib().sn_mc_sweeps(self._h, int(nsweeps)))
        if sync:
            self.sync()

    def MC_moves(self, moves: int):
        """MC_moves(moves) in whole sweeps: moves must be a multiple of N for
        every replica (one sweep = N attempted updates)."""
        for r in self.replicas:
            if moves % r.N:
                raise ValueError("moves must be a whole number of sweeps (multiple of X*Y*Z)")
        n = {moves // r.N for r in self.replicas}
        if len(n) != 1:
            raise ValueError("replicas of different sizes: use MC_sweeps")
        self.MC_sweeps(n.pop())

    def sync(self):
        self._chk(lib().sn_sync(self._h))

    def replay(self, sites, props, u, u_drawn, ref_accept, replica: int = 0):
        s = np.ascontiguousarray(sites, np.int32)
        p = np.ascontiguousarray(props, np.float64).reshape(-1, 3)
        uu = np.ascontiguousarray(u, np.float64)
        ud = np.ascontiguousarray(u_drawn, np.uint8)
        ra = np.ascontiguousarray(ref_accept, np.uint8)
        dec = np.empty(len(s), np.uint8)
        mm, nt = C.c_uint64(0), C.c_uint64(0)
        self._chk(lib().sn_mc_replay(self._h, replica, len(s), _ptr(s, C.c_int32), _ptr(p, C.c_double),
                                     _ptr(uu, C.c_double), _ptr(ud, C.c_uint8), _ptr(ra, C.c_uint8),
                                     _ptr(dec, C.c_uint8), C.byref(mm), C.byref(nt)))
        return dec, mm.value, nt.value

    def counters(self, replica: int = 0):
        a, r, t = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        self._chk(lib().sn_counters(self._h, replica, C.byref(a), C.byref(r), C.byref(t)))
        return {"accept": a.value, "reject": r.value, "attempted": t.value}

    def reset_counters(self):
        self._chk(lib().sn_reset_counters(self._h))

    def sweep_index(self, replica: int = 0) -> int:
        """The sweep counter keying replica `replica`'s counter-RNG draws."""
        v = C.c_uint64(0)
        self._chk(lib().sn_get_sweep_index(self._h, replica, C.byref(v)))
        return v.value

    def set_sweep_index(self, replica: int, index: int):
        """Resume replica `replica`'s draw stream at sweep `index` (with its
        saved lattice, a restarted run continues exactly)."""
        self._chk(lib().sn_set_sweep_index(self._h, replica, int(index)))

    # -- energy / adjoint / observables
    def energy(self, replica: int = 0):
        e = C.c_double(0)
        parts = np.empty(4, np.float64)
        self._chk(lib().sn_energy(self._h, replica, C.byref(e), _ptr(parts, C.c_double)))
        return e.value, parts

    def gradient(self, replica: int = 0, want_field: bool = True):
        g = np.empty((self.replicas[replica].N, 3), np.float64) if want_field else None
        dp = np.empty(6, np.float64)
        e = C.c_double(0)
        self._chk(lib().sn_gradient(self._h, replica, _ptr(g, C.c_double) if want_field else None,
                                    _ptr(dp, C.c_double), C.byref(e)))
        return g, dp, e.value

    def gradient_device(self, grad, replica: int = 0):
        """gradient() with dE/dp written into `grad` (contiguous float64 [N, 3]
        torch tensor on the GPU, or None); returns (dparams, energy)."""
        ptr = None
        if grad is not None:
            _check_device_tensor(grad, (self.replicas[replica].N, 3))
            _sync_tensor_stream(grad)
            ptr = C.c_void_p(grad.data_ptr())
        dp = np.empty(6, np.float64)
        e = C.c_double(0)
        self._chk(lib().sn_gradient_device(self._h, replica, ptr, _ptr(dp, C.c_double), C.byref(e)))
        return dp, e.value

    def gradient_all_device(self, grad=None):
        """Every replica's adjoint in one pass.  `grad` (contiguous float64
        [sum N, 3] torch tensor on the GPU, or None) receives dE/dp of all
        sites, replica after replica.  Returns (dparams [R, 6], energies [R])."""
        ptr = None
        if grad is not None:
            _check_device_tensor(grad, (sum(r.N for r in self.replicas), 3))
            _sync_tensor_stream(grad)
            ptr = C.c_void_p(grad.data_ptr())
        R = len(self.replicas)
        dp = np.empty((R, 6), np.float64)
        e = np.empty(R, np.float64)
        self._chk(lib().sn_gradient_all_device(self._h, ptr, _ptr(dp, C.c_double), _ptr(e, C.c_double)))
        return dp, e

    def moment(self, replica: int = 0) -> np.ndarray:
        M = np.empty(3, np.float64)
        self._chk(lib().sn_moment(self._h, replica, _ptr(M, C.c_double)))
        return M

    def radial_order(self, replica: int = 0):
        """radial_order_parameter (src/starrynight-analysis.c:528-598):
        (fe[81], afe[81], count[81]) indexed by r^2."""
        fe, afe = np.empty(81, np.float64), np.empty(81, np.float64)
        cnt = np.empty(81, np.int64)
        self._chk(lib().sn_radial_order(self._h, replica, _ptr(fe, C.c_double), _ptr(afe, C.c_double),
                                        _ptr(cnt, C.c_int64)))
        return fe, afe, cnt

    def dipole_potential(self, replica: int = 0) -> np.ndarray:
        """dipole_potential (src/starrynight-analysis.c:65-94) at every site."""
        pot = np.empty(self.replicas[replica].N, np.float64)
        self._chk(lib().sn_dipole_potential(self._h, replica, _ptr(pot, C.c_double)))
        return pot

    def dipole_efield(self, cutoff: int, replica: int = 0) -> np.ndarray:
        """dipole_electricfield (src/starrynight-analysis.c:393-457) at every
        site, as (N, 3) vectors (the reference reports their norm)."""
        E = np.empty((self.replicas[replica].N, 3), np.float64)
        self._chk(lib().sn_dipole_efield(self._h, replica, int(cutoff), _ptr(E, C.c_double)))
        return E

    def recombination(self, replica: int = 0) -> np.ndarray:
        """recombination_calculator (src/starrynight-analysis.c:97-164):
        {ZBe, ZBh, ZFDe, ZFDh, R_Boltz, R_FD, FD-Total-electron, FD-Total-hole}."""
        out = np.empty(8, np.float64)
        self._chk(lib().sn_recombination(self._h, replica, _ptr(out, C.c_double)))
        return out

    # -- schedule / timing
    def schedule(self, replica: int = 0):
        n, ph = C.c_int64(0), C.c_int32(0)
        self._chk(lib().sn_schedule_info(self._h, replica, C.byref(n), C.byref(ph)))
        order = np.empty(n.value, np.int32)
        starts = np.empty(ph.value + 1, np.int32)
        self._chk(lib().sn_schedule_order(self._h, replica, _ptr(order, C.c_int32),
                                          _ptr(starts, C.c_int32)))
        return order, starts

    def sweep_config(self):
        t, l, g, f = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        self._chk(lib().sn_sweep_config(self._h, C.byref(t), C.byref(l), C.byref(g), C.byref(f)))
        return {"tiles": t.value, "launches_per_sweep": l.value, "lanes_per_site": g.value,
                "fine_colours": f.value}

    def kernel_choice(self):
        """The sweep kernel this context runs (1 field update, 2 gather, 3
        whole lattice), its threads per workgroup, lanes per site and pairs
        per lane."""
        k, b, g, n = C.c_int32(0), C.c_int32(0), C.c_int32(0), C.c_int32(0)
        self._chk(lib().sn_kernel_choice(self._h, C.byref(k), C.byref(b), C.byref(g), C.byref(n)))
        return {"kernel": k.value, "block_threads": b.value, "lanes_per_site": g.value,
                "pairs_per_lane": n.value}

    def time_sweep_kernel(self, nsweeps: int):
        ms, n = C.c_double(0), C.c_int32(0)
        self._chk(lib().sn_time_sweep_kernel(self._h, int(nsweeps), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def stream_ptr(self) -> int:
        s = C.c_void_p()
        self._chk(lib().sn_stream(self._h, C.byref(s)))
        return s.value or 0


def gen_neighbour(rc: int, Z: int = 2):
    """Neighbour offsets in gen_neighbour order (src/starrynight-montecarlo-core.c:38-72):
    every integer r != 0 with |r| <= rc, dz == 0 when Z == 1."""
    zc = 0 if Z == 1 else rc
    out = []
    for dx in range(-rc, rc + 1):
        for dy in range(-rc, rc + 1):
            for dz in range(-zc, zc + 1):
                r2 = dx * dx + dy * dy + dz * dz
                if r2 and r2 <= rc * rc:
                    out.append((dx, dy, dz))
    return out


def plan_schedule(replicas: ReplicaParams | Sequence[ReplicaParams], replica: int = 0,
                  tuning: TuningParams | None = None):
    """Sweep visit order of one replica, planned on the host (no device):
    returns (order[N], phase_start[n_phases + 1])."""
    if isinstance(replicas, ReplicaParams):
        replicas = [replicas]
    arr = (Params * len(replicas))(*[r.to_c() for r in replicas])
    tun = tuning.to_c() if tuning else None
    n, ph = C.c_int64(0), C.c_int32(0)
    st = lib().sn_plan_schedule(arr, len(replicas), C.byref(tun) if tun else None, replica,
                                C.byref(n), C.byref(ph), None, None)
    if st != 0:
        raise SnError(st, lib().sn_status_string(st).decode())
    order = np.empty(n.value, np.int32)
    starts = np.empty(ph.value + 1, np.int32)
    st = lib().sn_plan_schedule(arr, len(replicas), C.byref(tun) if tun else None, replica, None, None,
                                _ptr(order, C.c_int32), _ptr(starts, C.c_int32))
    if st != 0:
        raise SnError(st, lib().sn_status_string(st).decode())
    return order, starts


def reference_initial_lattice(X: int, Y: int, Z: int, InitialLattice: str = "random", seed: int = None,
                              T: int = 300, Dipoles=(1.0,), Prevalence=(1.0,)) -> np.ndarray:
    """The reference's initial state (init_genrand(0xDEADBEEF + T), the
    InitialLattice pattern, solid_solution), computed on the host; [N, 4]."""
    if seed is None:
        seed = (0xDEADBEEF + int(T)) & 0xFFFFFFFF
    ln = np.ascontiguousarray(Dipoles, np.float64)
    pv = np.ascontiguousarray(Prevalence, np.float64)
    out = np.empty((X * Y * Z, 4), np.float64)
    st = lib().sn_reference_initial_lattice(X, Y, Z, InitialLattice.encode(), seed & 0xFFFFFFFF, len(ln),
                                            _ptr(ln, C.c_double), _ptr(pv, C.c_double), _ptr(out, C.c_double))
    if st != 0:
        raise SnError(st, "invalid initial-lattice request (Prevalence must cover the drawn samples)")
    return out


def autotune(replicas, budget_ms: int = 3000, use_cache: bool = True, Dipoles=None, Prevalence=None):
    """sn_autotune: the fastest measured schedule for these replicas on the
    current device, as (TuningParams, site-updates/s).  Dipoles / Prevalence
    (the species table the lattices will hold; default unit lengths) are used
    for the trial lattices: mixed lengths change the kernels' LDS use."""
    if isinstance(replicas, ReplicaParams):
        replicas = [replicas]
    arr = (Params * len(replicas))(*[r.to_c() for r in replicas])
    t = Tuning()
    rate = C.c_double(0)
    ln = np.ascontiguousarray(Dipoles if Dipoles is not None else [], np.float64)
    pv = np.ascontiguousarray(Prevalence if Prevalence is not None else [], np.float64)
    if len(ln) != len(pv):
        raise ValueError("Dipoles and Prevalence differ in length")
    st = lib().sn_autotune(arr, len(replicas), len(ln), _ptr(ln, C.c_double) if len(ln) else None,
                           _ptr(pv, C.c_double) if len(pv) else None, int(budget_ms), int(bool(use_cache)),
                           C.byref(t), C.byref(rate))
    if st != 0:
        raise SnError(st, "autotune found no runnable schedule")
    return TuningParams(tile=tuple(t.tile), lanes_per_site=t.lanes_per_site, fine_colouring=t.fine_colouring,
                        use_graph=t.use_graph, kernel=t.kernel, block_threads=t.block_threads), rate.value
