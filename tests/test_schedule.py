"""The colour schedule (pure host logic, sn_plan_schedule) and the C ABI's
argument checks, on the CPU.

A sweep must visit every site exactly once, and no two sites of one phase may
lie within DipoleCutOff of each other through the periodic boundary: then
updating a phase's sites concurrently equals any sequential order, which is
what lets the kernels run a phase in parallel."""
import itertools

import numpy as np
import pytest


def offsets(rc, two_d):
    out = []
    for dx, dy, dz in itertools.product(range(-rc, rc + 1), repeat=3):
        if two_d and dz:
            continue
        if 0 < dx * dx + dy * dy + dz * dz <= rc * rc:
            out.append((dx, dy, dz))
    return out


def check_schedule(sn, rp, tuning=None, coarse=None):
    order, starts = sn.plan_schedule(rp, 0, tuning)
    return check_order(rp, order, starts, coarse)


def check_order(rp, order, starts, coarse=None):
    """A visit order (a permutation grouped into phases) is conflict-free:
    no two sites of one phase within DipoleCutOff through the periodic
    boundary.  Returns the number of phases."""
    order, starts = np.asarray(order), np.asarray(starts)
    N = rp.N
    assert len(order) == N and np.array_equal(np.sort(order), np.arange(N))
    assert starts[0] == 0 and starts[-1] == N and np.all(np.diff(starts) >= 0)
    X, Y, Z = rp.X, rp.Y, rp.Z
    z, y, x = order % Z, (order // Z) % Y, order // (Y * Z)
    in_phase = np.full(N, -1, np.int64)
    for p in range(len(starts) - 1):
        in_phase[order[starts[p]:starts[p + 1]]] = p
    phase_of_order = in_phase[order]
    for dx, dy, dz in offsets(rp.DipoleCutOff, Z == 1):
        t = (((x + dx) % X) * Y + (y + dy) % Y) * Z + (z + dz) % Z
        clash = in_phase[t] == phase_of_order
        assert not clash.any(), (dx, dy, dz, int(np.argmax(clash)))
    if coarse is not None:
        assert (len(starts) - 1) % coarse == 0
    return len(starts) - 1


CASES = [
    ((16, 16, 16), 3, {}), ((17, 19, 21), 3, {}), ((24, 20, 18), 3, dict(fine_colouring=2)),
    ((12, 16, 20), 3, dict(fine_colouring=1)), ((20, 20, 20), 4, {}), ((12, 12, 12), 2, {}),
    ((24, 24, 24), 5, {}), ((25, 25, 1), 3, {}), ((30, 26, 1), 3, dict(fine_colouring=1)),
    ((33, 33, 33), 3, dict(tile=(6, 8, 10))), ((128, 16, 16), 3, dict(tile=(4, 8, 8))),
]


@pytest.mark.parametrize("dims,rc,tun", CASES, ids=[f"{d[0]}x{d[1]}x{d[2]}rc{r}{'-' + '-'.join(t) if t else ''}"
                                                   for d, r, t in CASES])
def test_sweep_is_a_conflict_free_permutation(sn, dims, rc, tun):
    rp = sn.ReplicaParams(*dims, DipoleCutOff=rc)
    nph = check_schedule(sn, rp, sn.TuningParams(**tun) if tun else None, coarse=4 if dims[2] == 1 else 8)
    assert nph >= 8


def test_phase_counts_of_the_colourings(sn):
    """BCC colouring at rc = 3: 32 fine colours; exact minimum-index
    sublattices: 26 at rc = 3, 54 at rc = 4; cubic (rc + 1)^3 = 64."""
    rp = sn.ReplicaParams(16, 16, 16, DipoleCutOff=3)
    assert check_schedule(sn, rp, sn.TuningParams(fine_colouring=3)) == 8 * 32
    assert check_schedule(sn, rp, sn.TuningParams(fine_colouring=2)) == 8 * 26
    assert check_schedule(sn, rp, sn.TuningParams(fine_colouring=1)) == 8 * 64
    rp4 = sn.ReplicaParams(24, 24, 24, DipoleCutOff=4)
    assert check_schedule(sn, rp4, sn.TuningParams(fine_colouring=2)) == 8 * 54
    # 20^3 at rc = 4 splits into tiles of 5^3 = 125 sites: a colour with no
    # site in any tile gives no phase
    assert check_schedule(sn, sn.ReplicaParams(20, 20, 20, DipoleCutOff=4), sn.TuningParams(fine_colouring=2)) <= 8 * 54


def test_replicas_are_scheduled_independently(sn):
    reps = [sn.ReplicaParams(16, 16, 16), sn.ReplicaParams(24, 20, 17), sn.ReplicaParams(18, 18, 18)]
    for r in range(3):
        order, starts = sn.plan_schedule(reps, r)
        assert len(order) == reps[r].N and np.array_equal(np.sort(order), np.arange(reps[r].N))


@pytest.mark.parametrize("reps,status,text", [
    ([dict(X=16, Y=16, Z=16, DipoleCutOff=3), dict(X=16, Y=16, Z=16, DipoleCutOff=4)], 1, "share DipoleCutOff"),
    ([dict(X=0, Y=16, Z=16)], 1, "positive"),
    ([dict(X=16, Y=16, Z=16, DipoleCutOff=9)], 5, "DipoleCutOff"),
    ([dict(X=3, Y=16, Z=16, DipoleCutOff=3)], 5, "exceed DipoleCutOff"),
    ([dict(X=16, Y=16, Z=16), dict(X=16, Y=16, Z=1)], 1, "2D"),
    ([dict(X=16, Y=16, Z=16, T=-1.0)], 1, "temperature"),
])
def test_create_rejects_bad_parameters(sn, reps, status, text):
    """sn_create checks its arguments before touching the device and reports
    a status and a message (the reference exits instead)."""
    with pytest.raises(sn.SnError) as e:
        sn.StarryNight([sn.ReplicaParams(**r) for r in reps])
    assert e.value.status == status and text in str(e.value)


def test_plan_schedule_rejects_bad_parameters(sn):
    with pytest.raises(sn.SnError):
        sn.plan_schedule([sn.ReplicaParams(16, 16, 16)], replica=1)
    with pytest.raises(sn.SnError):
        sn.plan_schedule(sn.ReplicaParams(16, 16, 16, DipoleCutOff=0))
    with pytest.raises(sn.SnError):
        sn.plan_schedule(sn.ReplicaParams(9, 9, 9, DipoleCutOff=3))   # untileable at rc = 3


@pytest.mark.parametrize("dims,rc,tun", [
    ((128, 128, 128), 3, dict(tile=(8, 8, 8), fine_colouring=3, kernel=2)),   # the headline sweep
    ((64, 64, 64), 3, dict(tile=(8, 8, 8), fine_colouring=3, kernel=2)),
    ((128, 128, 128), 4, dict(tile=(8, 8, 10), fine_colouring=2, kernel=2)),  # solid-solution ramp
], ids=["128cube_rc3_bcc_8", "64cube_rc3_bcc_8", "128cube_rc4_minidx_8x8x10"])
def test_benchmark_schedules_are_conflict_free(sn, dims, rc, tun):
    """The schedules the benchmark times at full size, checked site by site
    through the periodic boundary."""
    rp = sn.ReplicaParams(*dims, DipoleCutOff=rc)
    nph = check_schedule(sn, rp, sn.TuningParams(**tun), coarse=8)
    assert nph == 8 * (32 if rc == 3 else 54)
