"""sn_gradient_all_device_noalloc: the all-replica adjoint with its partial
sums in the library's fixed per-GPU scratch array instead of an allocated
buffer.  Checked against sn_gradient_all_device (same kernels, same reduction
order: results equal bit for bit) on two calls in a row, the second with a
smaller input, so a stale tail of the first call's partials cannot leak in."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import print_case, rand_lattice

pytestmark = pytest.mark.gpu


def noalloc(sn, ctx, grad):
    """The entry point through the C ABI (ctypes), as a caller binds it."""
    R = len(ctx.replicas)
    dp = np.empty((R, 6), np.float64)
    e = np.empty(R, np.float64)
    ptr = C.c_void_p(grad.data_ptr()) if grad is not None else None
    st = sn.lib().sn_gradient_all_device_noalloc(ctx._h, ptr, dp.ctypes.data_as(C.POINTER(C.c_double)),
                                                 e.ctypes.data_as(C.POINTER(C.c_double)))
    return st, dp, e


def check_against_allocating(sn, reps, seed, mixed):
    import torch
    rng = np.random.default_rng(seed)
    with sn.StarryNight(reps) as ctx:
        for r, rp in enumerate(reps):
            ctx.set_lattice(rand_lattice(rp.N, rng, mixed=mixed), r)
        total = sum(rp.N for rp in reps)
        g_ref = torch.full((total, 3), float("nan"), dtype=torch.float64, device="cuda:0")
        g_new = torch.full((total, 3), float("nan"), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        dp_ref, e_ref = ctx.gradient_all_device(g_ref)
        st, dp_new, e_new = noalloc(sn, ctx, g_new)
        assert st == 0, sn.lib().sn_last_error(ctx._h)
        torch.cuda.synchronize()
        assert np.array_equal(dp_new, dp_ref) and np.array_equal(e_new, e_ref)
        assert torch.equal(g_new, g_ref)
        # energies only (no gradient buffer)
        st, dp2, e2 = noalloc(sn, ctx, None)
        assert st == 0 and np.array_equal(dp2, dp_ref) and np.array_equal(e2, e_ref)
        print(f"  {len(reps)} replicas: E = {e_new[:2]} ...", flush=True)


def test_noalloc_adjoint_matches_allocating_entry_point(sn):
    """A large call (6 replicas of 24^3 and 32^3, mixed lengths, 384 tiles of
    partials), then a smaller one (2 replicas of 16^3) on another context of
    the same GPU."""
    big = [sn.ReplicaParams(24, 24, 24, T=300, Efield=(0, 0, 0.02), seed=1 + k) for k in range(4)] + \
          [sn.ReplicaParams(32, 32, 32, T=300, K=0.1, seed=9 + k) for k in range(2)]
    print_case("noalloc_big", replicas=len(big))
    check_against_allocating(sn, big, 101, mixed=True)
    small = [sn.ReplicaParams(16, 16, 16, T=300, seed=20 + k) for k in range(2)]
    print_case("noalloc_small", replicas=len(small))
    check_against_allocating(sn, small, 102, mixed=False)


def test_noalloc_refuses_partials_beyond_the_scratch(sn):
    """128^3 in 4^3 tiles needs 32768 rows of 10 partial sums, beyond the
    131072-double scratch: refused before any launch, with a message."""
    rp = sn.ReplicaParams(128, 128, 128, T=300, seed=3)
    print_case("noalloc_too_large", X=128, tile=4)
    with sn.StarryNight(rp, sn.TuningParams(tile=(4, 4, 4))) as ctx:
        ctx.init_lattice_random(0)
        st, _, _ = noalloc(sn, ctx, None)
        assert st == 4 and "fixed scratch" in sn.lib().sn_last_error(ctx._h).decode()
        # the allocating entry point still serves it
        dp, e = ctx.gradient_all_device(None)
        assert np.isfinite(e).all()
