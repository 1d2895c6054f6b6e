#!/usr/bin/env python3
"""Benchmark: Metropolis site-updates/s of the tile-sweep kernel on one MI355X.

Workload (one "step" = one sweep = N attempted single-site updates):
  128^3 periodic lattice, DipoleCutOff = 3 (122 neighbours), CageStrain = 1,
  Efield = 0, K = 0, T = 300 K, FP64 unit dipoles uniform on the sphere
  (synthetic, drawn on the device from a fixed seed).
Also reported: the batched form (96 replicas of 32^3 over 32 temperatures x 3
fields plus 32 replicas of uneven sizes 16^3..48^3, one launch per coarse
colour for all of them) against a single 32^3 replica, and the CPU baseline
(the float-faithful C restatement of the reference hot path, one process per
core, as the reference is run).

Prints ONE JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector (= dense FP64 matrix) peak
PEAK_HBM_GBS = 8000.0


def flops_per_update(n_nb: int, n_nn: int) -> int:
    """Algorithmic FP64 flops of one attempted update: 19 per neighbour,
    3 per nearest neighbour, 40 per site (proposal, dE, acceptance)."""
    return 19 * n_nb + 3 * n_nn + 40


BYTES_PER_UPDATE = 56        # read x,y,z,L (32 B) + write x,y,z (24 B) per attempted update
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(kernel: str, L: int, rc: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (already scaled by the copy-kernel
    calibration factors stated in that file), or None if there is none."""
    path = os.path.join(PROFILES, f"{kernel}_{L}cube_rc{rc}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    return vals.get("traffic_bytes_per_launch")


def copy_kernel(nbytes: int = 64 << 20, reps: int = 3):
    """A plain device copy of known size (the FETCH_SIZE / WRITE_SIZE
    calibration: each launch reads and writes nbytes)."""
    import torch
    a = torch.randn(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": reps}), flush=True)


def pcie_inclusive(ctx, N, sweeps=10):
    """The same sweeps with the lattice handed over in host memory: upload,
    `sweeps` sweeps, download (pageable host buffers, the C ABI's exchange)."""
    import numpy as np
    host = ctx.get_lattice()
    t0 = time.perf_counter()
    ctx.set_lattice(host)
    ctx.MC_sweeps(sweeps, sync=False)
    out = ctx.get_lattice()
    dt = time.perf_counter() - t0
    assert out.shape == host.shape and np.isfinite(out).all()
    return {"site_updates_per_s": sweeps * N / dt, "sweeps": sweeps,
            "note": "upload + sweeps + download of the 64 MiB FP64 lattice from pageable host memory"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu_info():
    """The host CPU as lscpu reports it, and how many CPUs this process may
    use (scheduler affinity, cgroup CPU quota)."""
    info = {"model": None, "sockets": None, "cores_per_socket": None, "threads_per_core": None}
    try:
        for line in subprocess.check_output(["lscpu"], text=True, env={**os.environ, "LC_ALL": "C"}).splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    if info["sockets"] and info["cores_per_socket"]:
        info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    return info


# CPUs a benchmark may keep busy beside one GPU on the shared GPU hosts
CPU_SHARE_PER_GPU = 16


def cpu_baseline(seconds_target: float = 10.0, procs: int | None = None):
    """Time oracle/sn_cpu_bench on the 128^3 rc=3 workload: the float-faithful
    port of MC_moves (gcc -O3, one thread per process) on one core, then one
    process per usable physical core at once (distinct T, as the reference's
    GNU-parallel runs), and the FP64 restatement on one core.  Sample: a
    bounded number of moves per process (about `seconds_target` s each)."""
    exe = os.path.join(ROOT, "oracle", "sn_cpu_bench")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    host = host_cpu_info()

    def run(*args):
        return json.loads(subprocess.check_output([exe, *map(str, args)]))

    # calibrate: 0.5 M moves
    rate1 = run(128, 128, 128, 3, 300, 500000)["rate"]
    moves = int(max(1e6, rate1 * seconds_target))
    single = run(128, 128, 128, 3, 300, moves)
    if procs is None:
        usable = [host.get("physical_cores") or host["affinity_cpus"], host["affinity_cpus"], CPU_SHARE_PER_GPU]
        if host["cgroup_cpu_quota"]:
            usable.append(int(host["cgroup_cpu_quota"]))
        procs = max(1, min(usable))
    ps = [subprocess.Popen([exe, "128", "128", "128", "3", str(300 + 10 * i), str(moves)],
                           stdout=subprocess.PIPE) for i in range(procs)]
    rates = [json.loads(p.communicate()[0])["rate"] for p in ps]
    f64_moves = int(max(5e5, rate1 * seconds_target / 2))
    f64 = run(128, 128, 128, 3, 300, f64_moves, "f64")
    # the 2D 25x25x1 rc=3 case (the reference's CPU-runnable configuration), one core, 3000 sweeps
    sq = run(25, 25, 1, 3, 300, 3000 * 625)
    phys = host.get("physical_cores")
    return {
        "value": sum(rates), "unit": "site-updates/s", "cores": procs, "kind": "port",
        "single_core": single["rate"],
        # not measured: the host's other cores serve other GPUs' jobs, so this
        # only scales the measured per-process rate to every physical core
        "projected_all_physical_cores": (sum(rates) / procs * phys) if phys else None,
        "physical_cores": phys,
        "single_core_fp64_oracle": f64["rate"],
        "square_25x25_rc3_single_core": sq["rate"],
        "host": host,
        "sample": f"128^3 rc=3 T=300 random lattice, {moves} moves per process "
                  f"(~{moves / single['rate']:.0f} s), float-faithful restatement of "
                  f"MC_moves/site_energy with MT19937, gcc -O3, {procs} single-threaded processes at once "
                  f"(one per usable physical core: min of lscpu's physical cores, the affinity mask, the cgroup "
                  f"quota and the {CPU_SHARE_PER_GPU} CPUs a GPU's job may use on this host); the FP64 "
                  f"restatement (or_mc_moves_f64) on one core, {f64_moves} moves",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--rc", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the built-in schedule instead of the autotuner's pick")
    ap.add_argument("--schedule", default=None,
                    help="fixed schedule 'tx,ty,tz,colouring,kernel[,threads]' instead of autotuning "
                         "(profiler runs: keeps trial launches out of the trace); with --workload it "
                         "applies to that workload's sweeps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the headline sweeps (for profiler runs)")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the other configurations (2D, temperature sweep, field ramp, adjoint)")
    ap.add_argument("--workload", default=None,
                    help="run only this one of the other configurations and print its figures "
                         "(for profiler runs): batched, 2d_25x25_rc3, 64cube_rc3_temperature_sweep, "
                         "128cube_rc4_solid_solution_ramp, 32x48cube_adjoint")
    ap.add_argument("--copy-kernel", action="store_true",
                    help="run only a 64 MiB device copy (profiler counter calibration)")
    args = ap.parse_args()
    if args.copy_kernel:
        copy_kernel()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (replicas, and slabs of "
              "one large lattice, would shard one per GPU).", file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401  (loaded before the library: one HIP runtime for both)

    import starrynight_amd as sn

    if args.workload == "batched":
        print(json.dumps(run_batched(sn, args.rc, log, autotune=not args.no_autotune)))
        return
    if args.workload:
        print(json.dumps(other_workloads(sn, log, autotune=not args.no_autotune, only=args.workload,
                                         schedule=parse_schedule(sn, args.schedule))))
        return

    L, rc = args.L, args.rc
    rp = sn.ReplicaParams(L, L, L, DipoleCutOff=rc, CageStrain=1.0, T=300, seed=0x5EED)
    N = rp.N
    n_nb = len(sn.gen_neighbour(rc, L))
    F = flops_per_update(n_nb, 6)
    if args.schedule:
        tun, tuned_rate = parse_schedule(sn, args.schedule), None
    elif args.no_autotune:
        tun, tuned_rate = None, None
    else:
        tun, tuned_rate = sn.autotune(rp, budget_ms=4000, use_cache=False)
    if tuned_rate is not None:
        log(f"[bench] autotuner pick: {tun} ({tuned_rate:.3e} site-updates/s in its trial)")
    ctx = sn.StarryNight(rp, tun)
    ctx.init_lattice_random(0)
    cfg = ctx.sweep_config()
    kc = ctx.kernel_choice()
    log(f"[bench] {L}^3 rc={rc} N={N} neighbours={n_nb} F={F} sweep config {cfg}, kernel {kc}")
    ctx.MC_sweeps(args.warmup)
    ctx.reset_counters()
    # one GPU, one process: the barrier is trivial; the library stream and
    # torch's device are both synchronised on each side of the timed steps
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.MC_sweeps(args.steps, sync=False)
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = t1 - t0
    rate = args.steps * N / dt
    cnt = ctx.counters()
    log(f"[bench] {args.steps} sweeps in {dt * 1e3:.2f} ms -> {rate:.4e} site-updates/s, "
        f"acceptance {cnt['accept'] / max(1, cnt['accept'] + cnt['reject']):.3f}")

    # dominant kernel: per-launch HIP events on the library stream
    ksweeps = max(2, min(args.steps, 20))
    mean_ms, nl = ctx.time_sweep_kernel(ksweeps)
    updates_per_launch = N * ksweeps / nl
    achieved = updates_per_launch * F / (mean_ms * 1e-3) / 1e12
    field = tun is not None and tun.kernel == 1
    kname = "field_sweep_kernel" if field else "tile_sweep_kernel"
    traffic = measured_traffic(kname, L, rc)
    hbm_alg = updates_per_launch * BYTES_PER_UPDATE / (mean_ms * 1e-3) / 1e9
    roof = {"bound": "fp64", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
            "traffic_source": (f"profiles/{kname}_{L}cube_rc{rc}_traffic.txt: rocprofv3 --pmc FETCH_SIZE and "
                               "WRITE_SIZE of this kernel on this workload, scaled by a copy kernel's readings")
            if traffic is not None else None,
            "kernel": kname, "mean_launch_ms": round(mean_ms, 5), "launches_timed": nl,
            "updates_per_launch": updates_per_launch, "flop_per_update": F,
            "algorithmic_bytes_per_update": BYTES_PER_UPDATE,
            "algorithmic_hbm_GBps": round(hbm_alg, 1), "hbm_peak_GBps": PEAK_HBM_GBS,
            "bound_note": "compute-bound: FP64 vector ALU (VALU) peak, 78.6 TFLOP/s on MI355X; "
                          "no MFMA (a stencil sum with a data-dependent update per phase, not a dense "
                          "contraction); 42 flop per algorithmic byte, far right of the HBM ridge"}
    log(f"[bench] {kname} mean {mean_ms * 1e3:.1f} us/launch over {nl} launches -> "
        f"{achieved:.2f} TFLOP/s ({achieved / PEAK_FP64_TFLOPS:.1%} of FP64 peak), "
        f"{hbm_alg:.0f} GB/s algorithmic")
    pcie = None if args.kernel_only else pcie_inclusive(ctx, N)
    ctx.close()
    if args.kernel_only:
        print(json.dumps({"metric": "site-updates/s", "value": rate, "roofline": roof}))
        return

    batched = None
    if not args.no_batched:
        batched = run_batched(sn, rc, log, autotune=not args.no_autotune)

    workloads = None
    if not args.no_workloads:
        workloads = other_workloads(sn, log, autotune=not args.no_autotune)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline()
        log(f"[bench] CPU: 1 core {cpu['single_core']:.3e}/s, {cpu['cores']} cores {cpu['value']:.3e}/s")

    line = {
        "metric": "site-updates/s and FP64-roofline fraction at 128^3 rc=3; batched vs single replica",
        "value": rate, "unit": "site-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{L}^3 periodic lattice, DipoleCutOff={rc}, single replica, one sweep per step",
                   "X": L, "Y": L, "Z": L, "DipoleCutOff": rc, "CageStrain": 1.0, "Efield": [0, 0, 0],
                   "T": 300, "neighbours": n_nb, "tiles": cfg["tiles"],
                   "launches_per_sweep": cfg["launches_per_sweep"], "lanes_per_site": cfg["lanes_per_site"],
                   "block_threads": kc["block_threads"], "pairs_per_lane": kc["pairs_per_lane"],
                   "fine_colours": cfg["fine_colours"],
                   "schedule": ("autotuned " + str(tun.__dict__)) if tun is not None else "built-in default"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_cpu_all_cores": (rate / cpu["value"]) if cpu else None,
        "speedup_vs_cpu_one_core": (rate / cpu["single_core"]) if cpu else None,
        "batched": batched,
        "workloads": workloads,
        "pcie_inclusive": pcie,
    }
    print(json.dumps(line))


def parse_schedule(sn, text):
    """TuningParams from 'tx,ty,tz,colouring,kernel[,threads]', or None."""
    if not text:
        return None
    v = [int(x) for x in text.split(",")]
    return sn.TuningParams(tile=tuple(v[:3]), fine_colouring=v[3], kernel=v[4],
                           block_threads=v[5] if len(v) > 5 else 0)


def _timed_sweeps(ctx, sweeps, sites):
    ctx.MC_sweeps(2)
    t0 = time.perf_counter()
    ctx.MC_sweeps(sweeps, sync=False)
    ctx.sync()
    return sweeps * sites / (time.perf_counter() - t0)


def other_workloads(sn, log, autotune=True, only=None, schedule=None):
    """The remaining configurations, each timed briefly (site-updates/s of the
    sweeps; the adjoint in gradient evaluations/s).  `only` names one of them
    (a profiler run of a single workload); `schedule` (TuningParams) replaces
    the autotuner's pick."""
    import torch

    def tuned(reps, ms, **species):
        if schedule is not None:
            return schedule
        return sn.autotune(reps, budget_ms=ms, use_cache=False, **species)[0] if autotune else None

    def square_2d():
        # 2D 25x25x1, rc = 3, T = 300 (the reference's CPU-runnable case): 1000 sweeps.
        # One 625-site lattice fills one workgroup on one CU; the batched form
        # (512 replicas, 50-500 K, the library's own kernel choice) fills the device.
        rp = sn.ReplicaParams(25, 25, 1, DipoleCutOff=3, T=300, seed=0xDEADBEEF + 300)
        with sn.StarryNight(rp, tuned(rp, 1500)) as ctx:
            ctx.init_lattice_random(0)
            single = _timed_sweeps(ctx, 1000, rp.N)
        reps = [sn.ReplicaParams(25, 25, 1, DipoleCutOff=3, T=50.0 + 450.0 * k / 511, seed=2500 + k)
                for k in range(512)]
        with sn.StarryNight(reps) as ctx:
            for r in range(len(reps)):
                ctx.init_lattice_random(r)
            agg = _timed_sweeps(ctx, 200, sum(r.N for r in reps))
        return {"site_updates_per_s": single, "sweeps": 1000,
                "batched_512_replicas_site_updates_per_s": agg, "batched_sweeps": 200}

    def temperature_sweep():
        # 64^3 rc = 3: one replica, and the 0-500 K sweep in 25 K steps batched as 21 replicas
        rp = sn.ReplicaParams(64, 64, 64, DipoleCutOff=3, T=300, seed=64)
        with sn.StarryNight(rp, tuned(rp, 2000)) as ctx:
            ctx.init_lattice_random(0)
            single = _timed_sweeps(ctx, 20, rp.N)
        reps = [sn.ReplicaParams(64, 64, 64, DipoleCutOff=3, T=25.0 * k, seed=640 + k) for k in range(21)]
        with sn.StarryNight(reps, tuned(reps, 3000)) as ctx:
            for r in range(len(reps)):
                ctx.init_lattice_random(r)
            agg = _timed_sweeps(ctx, 10, sum(r.N for r in reps))
        return {"single_replica": single, "21_temperatures_batched": agg, "temperatures": "0..500 K step 25"}

    def field_ramp():
        # 128^3 rc = 4 solid solution (Dipoles [1, 0], Prevalence [0.5, 0.5]), 16 replicas
        # on an E_z ramp: 3 ramp steps of 3 sweeps each (the full ramp is 41 x 500)
        reps = [sn.ReplicaParams(128, 128, 128, DipoleCutOff=4, T=300, Efield=(0, 0, -0.1), seed=4000 + k)
                for k in range(16)]
        tun = tuned(reps, 6000, Dipoles=(1.0, 0.0), Prevalence=(0.5, 0.5))
        with sn.StarryNight(reps, tun) as ctx:
            for r in range(len(reps)):
                ctx.init_lattice_random(r, Dipoles=(1.0, 0.0), Prevalence=(0.5, 0.5))
            ctx.MC_sweeps(1)
            t0 = time.perf_counter()
            steps, per = 3, 3
            for k in range(steps):
                for r in range(len(reps)):
                    ctx.set_params(r, 300, (0.0, 0.0, -0.1 + 0.005 * k), 1.0)
                ctx.MC_sweeps(per, sync=False)
            ctx.sync()
            dt = time.perf_counter() - t0
            n = sum(r.N for r in reps)
            return {"site_updates_per_s": steps * per * n / dt, "replicas": len(reps), "sweeps": steps * per,
                    "schedule": str(tun.__dict__) if tun else "built-in"}

    def adjoint():
        # 32 replicas of 48^3, rc = 3: energy + adjoint (local field and parameter derivatives)
        reps = [sn.ReplicaParams(48, 48, 48, DipoleCutOff=3, T=300, seed=480 + k) for k in range(32)]
        with sn.StarryNight(reps) as ctx:
            for r in range(len(reps)):
                ctx.init_lattice_random(r)
            g_all = torch.empty((sum(r.N for r in reps), 3), dtype=torch.float64, device="cuda:0")
            grads = [torch.empty((r.N, 3), dtype=torch.float64, device="cuda:0") for r in reps]
            ctx.gradient_all_device(g_all)
            ctx.gradient_device(grads[0], 0)
            ctx.gradient(0)
            passes = 10
            t0 = time.perf_counter()
            for _ in range(passes):
                ctx.gradient_all_device(g_all)
            dt_all = time.perf_counter() - t0
            t0 = time.perf_counter()
            for _ in range(2):
                for r in range(len(reps)):
                    ctx.gradient_device(grads[r], r)
            dt_one = time.perf_counter() - t0
            t0 = time.perf_counter()
            for r in range(len(reps)):
                ctx.gradient(r)
            dt_host = time.perf_counter() - t0
            return {"gradients_per_s": passes * len(reps) / dt_all, "replicas": len(reps),
                    "one_call_per_replica_gradients_per_s": 2 * len(reps) / dt_one,
                    "host_copy_gradients_per_s": len(reps) / dt_host,
                    "note": "each gradient: dE/dp on 110592 sites written in HBM + dE/d(J, E, s, K); "
                            "gradients_per_s: all 32 replicas per call (sn_gradient_all_device); "
                            "one_call_per_replica: sn_gradient_device per replica; host_copy: sn_gradient, "
                            "the N x 3 field copied to host memory"}

    def library_defaults():
        # the library's own kernel choice (NULL tuning) on other cutoffs and
        # sizes: 128^3 at DipoleCutOff 1, 2, 4, 5 (gather tiles), 64 replicas of
        # 12^3 and 14^3 (whole-lattice sweeps), 2D 1024^2 at DipoleCutOff 3
        res = {}
        cases = [(f"128cube_rc{rc}", [sn.ReplicaParams(128, 128, 128, DipoleCutOff=rc, T=300, seed=rc)], 6)
                 for rc in (1, 2, 4, 5)]
        cases += [(f"64x{L}cube_rc3", [sn.ReplicaParams(L, L, L, T=300, seed=L + k) for k in range(64)], 40)
                  for L in (12, 14)]
        cases.append(("2d_1024sq_rc3", [sn.ReplicaParams(1024, 1024, 1, T=300, seed=1024)], 20))
        for name, reps, sweeps in cases:
            with sn.StarryNight(reps) as ctx:
                for r in range(len(reps)):
                    ctx.init_lattice_random(r)
                res[name] = {"site_updates_per_s": _timed_sweeps(ctx, sweeps, sum(r.N for r in reps)),
                             "sweeps": sweeps, **ctx.kernel_choice()}
        return res

    table = {"2d_25x25_rc3": square_2d, "64cube_rc3_temperature_sweep": temperature_sweep,
             "128cube_rc4_solid_solution_ramp": field_ramp, "32x48cube_adjoint": adjoint,
             "library_defaults": library_defaults}
    if only is not None and only not in table:
        raise SystemExit(f"unknown workload {only!r}; one of {sorted(table)}")
    out = {}
    for name, fn in table.items():
        if only is None or name == only:
            out[name] = fn()
            log(f"[bench] {name}: {out[name]}")
    return out


def batched_replicas(sn, rc):
    reps = []
    k = 0
    for i in range(32):
        T = 50.0 + 450.0 * i / 31.0
        for Ez in (0.0, 0.01, 0.05):
            reps.append(sn.ReplicaParams(32, 32, 32, DipoleCutOff=rc, T=T, Efield=(0, 0, Ez), seed=1000 + k))
            k += 1
    for i in range(32):
        Lk = 16 + round(32 * i / 31)
        reps.append(sn.ReplicaParams(Lk, Lk, Lk, DipoleCutOff=rc, T=300, seed=1000 + k))
        k += 1
    return reps


def run_batched(sn, rc, log, sweeps=20, autotune=True):
    reps = batched_replicas(sn, rc)
    total = sum(r.N for r in reps)
    tun = sn.autotune(reps, budget_ms=6000, use_cache=False)[0] if autotune else None
    ctx = sn.StarryNight(reps, tun)
    for r in range(len(reps)):
        ctx.init_lattice_random(r)
    ctx.MC_sweeps(3)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.MC_sweeps(sweeps, sync=False)
    ctx.sync()
    dt = time.perf_counter() - t0
    agg = sweeps * total / dt
    cfg = ctx.sweep_config()
    ctx.close()
    one_rp = sn.ReplicaParams(32, 32, 32, DipoleCutOff=rc, T=300, seed=7)
    one = sn.StarryNight(one_rp, sn.autotune(one_rp, budget_ms=3000, use_cache=False)[0] if autotune else None)
    one.init_lattice_random(0)
    one.MC_sweeps(3)
    t0 = time.perf_counter()
    one.MC_sweeps(sweeps * 4, sync=False)
    one.sync()
    dt1 = time.perf_counter() - t0
    single = sweeps * 4 * 32 ** 3 / dt1
    one.close()
    log(f"[bench] batched {len(reps)} replicas ({total} sites): {agg:.4e}/s; single 32^3: {single:.4e}/s; "
        f"ratio {agg / single:.1f}")
    return {"replicas": len(reps), "sites_per_sweep": total, "aggregate": agg, "single_32cube": single,
            "ratio_vs_single_32cube": agg / single, "sweeps": sweeps, "tiles": cfg["tiles"],
            "schedule": ("autotuned " + str(tun.__dict__)) if tun is not None else "built-in default"}


if __name__ == "__main__":
    main()
