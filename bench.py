"""Benchmark: particle-steps/s of the 2D electrostatic PIC step on one MI355X.

Workload: collisionless Landau damping, 512x512 periodic grid, k lambda_D = 0.5, 64
particles/cell (16,777,216), Maxwellian vth = 1, FP32 particles, FP64 FFT Poisson,
dt = 0.05.  One "step" = field solve + fused gather/push/deposit (+ the amortised cell
sort).  Inputs are resident in HBM when the timed region starts.

Prints ONE JSON line.  Extra keys: roofline (dominant kernel, HIP-event timed), cpu_baseline
(the CPU oracle on a bounded sample, one thread) and cpu_baseline_all_cores (the same on
every host thread the process may use), stages (ms per step per stage), batched_vs_sequential
(48-run ragged ensemble), adjoint_over_forward (128^2 x 32 ppc, 100 steps) and mcc_256
(the collision workload: 256^2, 8.4M electrons, null-collision MCC), dsmc_256 (the
DSMC blob: 256^2, 2.1M particles) and two_stream_64 (config 1: 64^2, 409,600 FP64
particles), each with the oracle's CPU rate on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "msc-thesis_amd"))

METRIC = "particle-steps/s at 512²×64ppc; batched vs plain; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0
PMC_FILE = os.path.join(ROOT, "profiles", "push_deposit_landau512_traffic.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the ensemble comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing")
    ap.add_argument("--no-mcc", action="store_true", help="skip the collision workloads")
    ap.add_argument("--no-two-stream", action="store_true", help="skip the FP64 two-stream leg")
    ap.add_argument("--copy-only", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--stream", action="store_true",
                    help="run only the streaming reference: an in-place read-modify-write of "
                         "four 64 MiB FP32 arrays (the plain push's particle state) with the "
                         "caches flushed before each pass")
    ap.add_argument("--only", choices=["batch", "adjoint", "mcc", "dsmc", "two_stream"],
                    default=None,
                    help="run only the ensemble comparison or only the adjoint timing "
                         "(for profiling those legs)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tuning", default=None,
                    help="fixed tile,halo,splits,sort_every instead of the autotuner")
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=64)
    return ap.parse_args()


def read_traffic():
    """HBM bytes per push_deposit launch measured with rocprofv3 --pmc (see the file), by
    the tuning "tile,halo,splits,sort_every" they were measured at."""
    table, traffic, tuning = {}, None, None
    try:
        with open(PMC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    traffic = float(line.split("=")[1])
                elif line.startswith("traffic_tuning"):
                    tuning = line.split("=")[1].strip()
                elif line.startswith("traffic_at "):
                    key, val = line[len("traffic_at "):].split("=")
                    table[key.strip()] = float(val)
    except OSError:
        pass
    if traffic is not None and tuning:
        table[tuning] = traffic
    return table


def cpu_threads():
    """Host threads for the all-cores baseline: OMP_NUM_THREADS if set, else the CPUs this
    process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(parts, cfg, nsteps, nthreads=1):
    """The oracle's FP32-mirror step on the same particles, bounded sample: one thread (the
    reference's serial loop), or nthreads OpenMP threads (or_run_f32_threads, bit-identical
    particles)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import load
    o = load()
    n = parts[0].size
    arrs = [a.astype(np.float32) for a in parts]
    qp = -(cfg.lx * cfg.ly) / n
    t0 = time.perf_counter()
    if nthreads == 1:
        o.run_f32(*arrs, cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.dt, cfg.qm, qp, nsteps)
        fn = "or_run_f32, 1 thread"
    else:
        o.run_f32_threads(*arrs, cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.dt, cfg.qm, qp, nsteps,
                          nthreads)
        fn = f"or_run_f32_threads, {nthreads} OpenMP threads"
    dt = time.perf_counter() - t0
    return {"value": n * nsteps / dt, "unit": "particle-steps/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{nsteps} steps of the same {cfg.nx}x{cfg.ny} x {n} particle Landau "
                      f"workload through the FP32-mirror oracle ({fn}; the time includes "
                      f"its initial deposit and final field solve)",
            "seconds": dt}


def stream_probe(reps=20, n=1 << 24):
    """What a plain streaming pass over the push's particle state costs on this device: four
    in-place x += 1 passes over 16,777,216 FP32 values each (the x, y, vx, vy arrays of the
    benched workload: 536.9 MB read + written), timed with HIP events, after a 1 GiB fill
    that evicts them from the L2 and the last-level cache (cold), and repeated without the
    fill (warm: partly served from the last-level cache); then the same passes out of place,
    each array read and x + 1 written to a second array."""
    import torch
    arrs = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(4)]
    dsts = [torch.zeros(n, dtype=torch.float32, device="cuda") for _ in range(4)]
    flush = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    # in place (the plain push's pattern) and out of place into a second set of arrays (the
    # sort step's pattern: read one buffer, write the other), same bytes
    for mode in ("cold", "warm", "cold_out_of_place", "warm_out_of_place"):
        ms = []
        for r in range(reps):
            if mode.startswith("cold"):
                flush.fill_(float(r))
            e0.record()
            if mode.endswith("out_of_place"):
                for a, d in zip(arrs, dsts):
                    torch.add(a, 1.0, out=d)
            else:
                for a in arrs:
                    a.add_(1.0)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        ms.sort()
        med = ms[len(ms) // 2]
        out[mode] = {"median_us": med * 1e3, "min_us": ms[0] * 1e3,
                     "GB_per_s": 8.0 * n * 4 / (med * 1e-3) / 1e9}
    out["bytes"] = 8 * n * 4
    return out


def mcc_rate(steps=100, warmup=10):
    """Config 3: electron-neutral null-collision MCC on 256 x 256 with 8,388,608 electrons
    (Te = 2 eV, neutrals 1e21 m^-3), autotuned (tile, halo, splits, sort cadence; the
    state is restored afterwards); particle-steps/s of the whole step (field, push,
    collisions, births, sort) after warm-up steps."""
    import torch
    from pic_amd import Simulation, presets
    cfg, parts, _ = presets.mcc_256()
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.autotune(8)
        tune = sim.tuning()
        sim.step(warmup)
        sim.sync()
        torch.cuda.synchronize()
        n0 = sim.num_particles()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        counts = sim.collision_counts()
        n1 = sim.num_particles()
    return {"particle_steps_per_s": 0.5 * (n0 + n1) * steps / dt, "ms_per_step": dt * 1e3 / steps,
            "particles": n1, "steps": steps, "collision_counts": counts,
            "grid": [cfg.nx, cfg.ny], "tuning": tune}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _oracle import load
    return load()


def two_stream_rate(warmup=10):
    """Config 1: cold two-stream, 64 x 64, 409,600 FP64 particles, FP64 everything, dt = 0.1,
    autotuned, the configured 200 steps timed after warm-up steps; the growth rate of |rho_hat(k)| fitted
    over the linear phase t in [8, 16] of the timed run (cold beams, k v0 = 0.5:
    gamma = 0.3406).  CPU baseline: the oracle's FP64 step (or_run_f64, one thread) over the
    whole 200-step run on the same particles."""
    import numpy as np
    import torch
    from pic_amd import Simulation, presets
    cfg, parts, nsteps = presets.two_stream_64()
    n = parts[0].size
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.autotune(8)  # tiling / splits / sort cadence (the state is restored afterwards)
        tune = sim.tuning()
        sim.step(warmup)
        sim.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(nsteps)
        sim.sync()
        dt = time.perf_counter() - t0
    with Simulation(cfg) as sim:  # the growth fit needs the run from t = 0
        sim.set_particles(*parts)
        sim.step(nsteps)
        diag = sim.diagnostics(0, nsteps)
    t = np.arange(nsteps) * cfg.dt
    m = (t >= 8) & (t <= 16)
    gamma = float(np.polyfit(t[m], np.log(diag[m, 4]), 1)[0])
    out = {"particle_steps_per_s": n * nsteps / dt, "ms_per_step": dt * 1e3 / nsteps,
           "particles": n, "steps": nsteps, "grid": [cfg.nx, cfg.ny], "dtype": "f64",
           "growth_rate": gamma, "growth_rate_theory": 0.3406, "tuning": tune}
    return out


def two_stream_cpu(nsteps=200):
    from pic_amd import presets
    cfg, parts, _ = presets.two_stream_64()
    o = _oracle()
    arrs = [a.copy() for a in parts]
    n = arrs[0].size
    t0 = time.perf_counter()
    o.run_f64(*arrs, cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.dt, cfg.qm, -(cfg.lx * cfg.ly) / n,
              nsteps)
    dt = time.perf_counter() - t0
    return {"value": n * nsteps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"the whole {nsteps}-step 64x64 x {n} particle two-stream run through "
                      "the FP64 oracle (or_run_f64, 1 thread)", "seconds": dt}


def _collide_cpu(cfg, parts, nsteps):
    """The oracle composition of the engine's step with collisions (deposit, FP64 Poisson,
    E, FP32 push, then MCC per particle or DSMC per cell on the cell-sorted particles), one
    thread, on the workload's own particles for nsteps steps."""
    import ctypes
    import numpy as np
    from pic_amd import _lib
    o = _oracle()
    n = parts[0].size
    cap = max(int(cfg.capacity), n)
    x, y, vx, vy = (np.zeros(cap, np.float32) for _ in range(4))
    for a, b in zip((x, y, vx, vy), parts):
        a[:n] = b
    ids = np.zeros(cap, np.int32)
    ids[:n] = np.arange(n)
    qp = cfg.qp if cfg.qp != 0 else -(cfg.lx * cfg.ly) / n
    G = cfg.nx * cfg.ny
    gmax = np.full(G, cfg.dsmc_gmax0, np.float32)
    counts = np.zeros(4, np.int64)
    if cfg.collide == _lib.PIC_COLLIDE_MCC:
        kE, kMe = 1.602176634e-19, 9.1093837015e-31
        numax = _lib.lib.pic_mcc_nu_max(ctypes.byref(cfg))
        mp = (0.5 * kMe * cfg.mcc_v_unit ** 2 / kE,
              cfg.mcc_ng * cfg.mcc_v_unit * cfg.mcc_dt_phys * cfg.dt,
              1.0 - math.exp(-numax), numax, 2.0 * kMe / cfg.mcc_neutral_mass)
    m = n
    t0 = time.perf_counter()
    for s in range(nsteps):
        xs, ys, vxs, vys = x[:m], y[:m], vx[:m], vy[:m]
        acc = o.deposit_fixed(xs, ys, cfg.nx, cfg.ny, cfg.lx, cfg.ly)
        rho = o.fixed_to_rho(acc, cfg.nx, cfg.ny, cfg.lx, cfg.ly, qp)
        ex, ey = o.efield(o.poisson(rho, cfg.lx, cfg.ly), cfg.lx, cfg.ly)
        o.push_f32(xs, ys, vxs, vys, ex, ey, cfg.lx, cfg.ly, cfg.dt, cfg.qm,
                   0.5 if s == 0 else 1.0)
        if cfg.collide == _lib.PIC_COLLIDE_DSMC:
            keys = o.cell_keys(xs, ys, cfg.nx, cfg.ny, cfg.lx, cfg.ly)
            perm, _, off = o.stable_sort(keys, G)
            svx, svy = np.ascontiguousarray(vxs[perm]), np.ascontiguousarray(vys[perm])
            counts[0] += o.or_dsmc_step_f32(G, o.p(off), o.p(svx), o.p(svy), o.p(gmax),
                                            cfg.dsmc_coeff, cfg.seed, s)
            vx[perm], vy[perm] = svx, svy
        else:
            m = o.or_mcc_step_f32(m, cap, o.p(x), o.p(y), o.p(vx), o.p(vy), o.p(ids), m,
                                  *mp, cfg.seed, s, o.p(counts))
    dt = time.perf_counter() - t0
    return n * nsteps / dt, dt


def dsmc_rate(steps=100, warmup=10):
    """The DSMC workload: 256 x 256, a Gaussian blob of 2,097,152 particles (183 per cell at
    the centre), NTC hard spheres per cell on the cell-sorted state every step (nu dt ~ 0.1
    at the centre); particle-steps/s of the whole step after autotuning and warm-up steps."""
    import torch
    from pic_amd import Simulation, presets
    cfg, parts, _ = presets.dsmc_blob()
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        n = sim.num_particles()
        sim.autotune(4)  # tile / halo / splits (the cell sort runs every step)
        tune = sim.tuning()
        sim.step(warmup)
        sim.sync()
        torch.cuda.synchronize()
        c0 = sim.collision_counts()[0]
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        dt = time.perf_counter() - t0
        c1 = sim.collision_counts()[0]
    return {"particle_steps_per_s": n * steps / dt, "ms_per_step": dt * 1e3 / steps,
            "particles": n, "steps": steps, "grid": [cfg.nx, cfg.ny],
            "collisions_per_step": (c1 - c0) / steps, "tuning": tune}


def collide_cpu(which, nsteps):
    from pic_amd import presets
    cfg, parts, _ = presets.mcc_256() if which == "mcc" else presets.dsmc_blob()
    rate, dt = _collide_cpu(cfg, parts, nsteps)
    what = ("null-collision MCC on the 8,388,608-electron 256x256 workload" if which == "mcc"
            else "NTC DSMC on the 2,097,152-particle 256x256 blob")
    return {"value": rate, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{nsteps} steps of the {what} through the oracle's stages (FP32 mirror "
                      "push, FP64 Poisson, the collision step), 1 thread", "seconds": dt}


def batch_ratio(steps=300):
    """48 ragged runs: one batched context vs the same runs one at a time.  Both sides are
    autotuned first (the batch over tile/halo/sort cadence for all runs at once, each
    sequential run on its own), so the ratio compares the two launch structures."""
    import torch
    from pic_amd import Batch, Simulation, presets
    runs, _ = presets.batch_48()
    total = sum(p[0].size for _, p in runs)
    b = Batch([c for c, _ in runs])
    for r, (_, p) in enumerate(runs):
        b.set_particles(*p, run=r)
    b.autotune(16)
    btune = b.tuning()
    b.step(10)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    tb = time.perf_counter() - t0
    b.close()
    sims = []
    for c, p in runs:
        s = Simulation(c)
        s.set_particles(*p)
        s.autotune(8)
        s.step(10)
        s.sync()
        sims.append(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in sims:
        s.step(steps)
        s.sync()
    ts = time.perf_counter() - t0
    for s in sims:
        s.close()
    return {"runs": len(runs), "particles": total, "steps": steps,
            "batched_particle_steps_per_s": total * steps / tb,
            "sequential_particle_steps_per_s": total * steps / ts, "ratio": ts / tb,
            "batch_tuning": btune}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU only (got --gpus {args.gpus}); "
              "sharding across GPUs is not built.", file=sys.stderr)
        sys.exit(2)
    import torch
    from pic_amd import Simulation, presets

    torch.cuda.init()
    if args.copy_only:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 3}))
        return
    if args.stream:
        print(json.dumps({"stream_rmw": stream_probe()}))
        return
    if args.only == "batch":
        print(json.dumps({"batched_vs_sequential": batch_ratio()}))
        return
    if args.only == "mcc":
        print(json.dumps({"mcc_256": mcc_rate()}))
        return
    if args.only == "dsmc":
        print(json.dumps({"dsmc_256": dsmc_rate()}))
        return
    if args.only == "two_stream":
        print(json.dumps({"two_stream_64": two_stream_rate()}))
        return
    if args.only == "adjoint":
        from pic_amd import _lib
        from pic_amd.adjoint import time_adjoint
        print(json.dumps({"adjoint_over_forward": {"f32": time_adjoint(precision=_lib.PIC_FP32),
                                                   "f64": time_adjoint(precision=_lib.PIC_FP64)}}))
        return
    cfg, parts, _ = presets.landau_512(ppc=args.ppc, n=args.grid)
    n = parts[0].size
    sim = Simulation(cfg)
    sim.set_particles(*parts)
    if args.tuning:
        sim.set_tuning(*(int(v) for v in args.tuning.split(",")))
    elif not args.no_autotune:
        sim.autotune(trial_steps=8)
    tune = sim.tuning()
    fused = sim.sort_fused()
    sim.step(args.warmup)
    sim.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = n * args.steps / el
    ms_per_step = el * 1e3 / args.steps
    # per-stage kernel durations, HIP events on the engine stream (eager launches)
    stages = sim.stage_times(50)
    sim.check_finite()
    sim.close()

    G = cfg.nx * cfg.ny
    # push_deposit algorithmic bytes per launch: x,y,vx,vy read + written (32 B/particle),
    # padded E read (8 B/node) and fixed-point rho written (8 B/node); with the sort fused
    # into the push, one launch in sort_every also reads and writes the particle ids
    # (8 B/particle) — the launch average below spans all three kinds of launch
    alg_bytes = 32.0 * n + 16.0 * G
    if fused and tune["sort_every"] > 0:
        alg_bytes += 8.0 * n / tune["sort_every"]
    push_ms = stages["push_deposit"]
    achieved = alg_bytes / (push_ms * 1e-3) / 1e9
    # the measured traffic belongs to a tuning: reported only for one that was measured
    traffic_tuning = "%d,%d,%d,%d" % (tune["tile"], tune["halo"], tune["splits"],
                                      tune["sort_every"])
    traffic = read_traffic().get(traffic_tuning)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Landau quiet-start loader)",
        "config": {"workload": f"landau_{cfg.nx}x{cfg.ny}_{args.ppc}ppc", "grid": [cfg.nx, cfg.ny],
                   "particles": n, "dt": cfg.dt, "field": "f64 FFT Poisson",
                   "charge": "int64 fixed-point", "tuning": tune, "sort_fused": fused},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_push (fused gather+push+deposit)",
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": push_ms,
                     "traffic_tuning": traffic_tuning},
        "stages_ms": stages,
        "algorithmic_roofline_frac": value * (40.0 + 72.0 * G / n) / (HBM_PEAK_GBS * 1e9),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(parts, cfg, 40)
        nt = cpu_threads()
        if nt > 1:
            line["cpu_baseline_all_cores"] = cpu_baseline(parts, cfg, 80, nt)
    if not args.no_batch:
        try:
            line["batched_vs_sequential"] = batch_ratio()
        except Exception as e:  # reported, not hidden
            line["batched_vs_sequential"] = {"error": repr(e)}
    if not args.no_adjoint:
        try:
            from pic_amd import _lib
            from pic_amd.adjoint import time_adjoint
            line["adjoint_over_forward"] = {"f32": time_adjoint(precision=_lib.PIC_FP32),
                                            "f64": time_adjoint(precision=_lib.PIC_FP64)}
        except Exception as e:
            line["adjoint_over_forward"] = {"error": repr(e)}
    if not args.no_mcc:
        try:
            line["mcc_256"] = mcc_rate()
            if not args.no_cpu:
                line["mcc_256"]["cpu_baseline"] = collide_cpu("mcc", 3)
        except Exception as e:  # reported, not hidden
            line["mcc_256"] = {"error": repr(e)}
        try:
            line["dsmc_256"] = dsmc_rate()
            if not args.no_cpu:
                line["dsmc_256"]["cpu_baseline"] = collide_cpu("dsmc", 3)
        except Exception as e:
            line["dsmc_256"] = {"error": repr(e)}
    if args.no_two_stream:
        print(json.dumps(line))
        return
    try:
        line["two_stream_64"] = two_stream_rate()
        if not args.no_cpu:
            line["two_stream_64"]["cpu_baseline"] = two_stream_cpu()
    except Exception as e:
        line["two_stream_64"] = {"error": repr(e)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
