"""Benchmark: ms per timestep of the FP64 512^3 NS+ACDI MHIT36 step on one MI355X.

Workload (the 512^3 NS+ACDI box): random-phase divergence-free velocity,
E(k) ~ k^4 exp(-2 (k/4)^2), u_rms = 1 (seed 1234, made on the device); ABC forcing A=B=C=1 at
k0=2; nu=0.001, rho=1; 512 non-overlapping drops of radius 0.1 (seed 4321),
eps = dx, gamma = 1.2 umax, sigma = 0.1; dt = 0.1 dx / umax0 (fixed; the
explicit phase-field update is unstable above CFL ~0.14).  Inputs
are resident in HBM before the timed region.  One step = every kernel of
multi/main.f90:329-946 (phase field, momentum + CSF + AB2, divergence, FFT
Poisson solve, projection, mean removal, umax/CFL).

`value` is the FP64 step (the drop-in); the same inputs are then run in mixed
precision (MHIT_MIXED) and reported under "mixed".  "ensemble" is the serving
stream of independent boxes (ensemble_leg), one at a time against the
Ensemble loop.

Prints ONE JSON line.  --gpus above 1 is refused: this build targets one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def alg_bytes_per_step(n, phiflag):
    """Algorithmic HBM bytes of one step: 39 F (NS+ACDI) or 33 F (NS), F = 8 N^3."""
    F = 8.0 * n ** 3
    return (39.0 if phiflag else 33.0) * F


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_512cubed_ns_acdi_fp64_hbm_traffic.json")


def measured_traffic(n, phiflag, kernel):
    """HBM bytes per launch of `kernel` as rocprofv3 FETCH_SIZE + WRITE_SIZE
    measured them for this workload (scaled by a 1 GiB copy; the committed
    summary in profiles/ holds the commands), or None when not measured."""
    if n != 512 or not phiflag or not os.path.exists(TRAFFIC_FILE):
        return None
    k = json.load(open(TRAFFIC_FILE))["kernels"].get(kernel)
    return None if k is None else round(k["read_bytes"] + k["write_bytes"])


def make_inputs(n, phiflag, seed_u=1234, seed_phi=4321, ndrops=512, radius=0.1):
    """Host copies of the workload's initial fields: the random-phase velocity
    made on the device (mhit_init_random_velocity) and the drops (ic.drops)."""
    from mhit36_amd import Solver, ic
    with Solver(nx=n, dt=1e-3, mu=0.001, phiflag=0, inflow=1) as s:
        s.init_random_velocity(seed=seed_u, k_peak=4.0, urms=1.0)
        u, v, w = (s.get_field(f) for f in (1, 2, 3))
    _, dx = ic.grid(n)
    phi = None
    if phiflag:
        eps = dx
        centres = ic.drop_centres(ndrops, radius, eps, seed=seed_phi)
        phi = ic.drops(n, centres, radius, eps)
    umax0 = max(float(u.max()), float(v.max()), float(w.max()))
    # CFL 0.5 for NS only; 0.1 with ACDI, whose explicit-Euler phase-field
    # diffusion gamma*eps*dt/dx^2 <= 1/6 with gamma = 1.2 umax, eps = dx needs CFL < 0.14
    cfl = 0.1 if phiflag else 0.5
    dt = cfl * dx / umax0
    return u, v, w, phi, dt


def params_for(n, phiflag, dt):
    return dict(nx=n, dt=dt, mu=0.001, rho=1.0, f1=1.0, f2=1.0, f3=1.0, k0=2.0, radius=0.1,
                sigma=0.1, epsr=1.0, phiflag=phiflag, inflow=1, inphi=1 if phiflag else 0,
                gamma_factor=1.2)


BOX_CPU_SHARE = 16      # host threads one GPU's process may use on the GPU box


def host_cores():
    """(sockets, physical cores) of this host from lscpu's parseable listing
    (unique (socket, core) pairs), or (None, os.cpu_count()) without lscpu."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=SOCKET,CORE"], capture_output=True, text=True, check=True).stdout
        pairs = {tuple(line.split(",")) for line in out.splitlines() if line and not line.startswith("#")}
        return len({p[0] for p in pairs}), len(pairs)
    except (OSError, subprocess.CalledProcessError):
        return None, os.cpu_count() or 1


def _oracle_step_times(n, phiflag, fields, dt, threads, untimed, timed):
    """Wall time (ms) of each of `timed` oracle steps after `untimed` ones
    (the first step is the Euler start; the timed ones are AB2 steps)."""
    from oracle import orc
    orc.set_threads(threads)
    o = orc.Oracle(**params_for(n, phiflag, dt))
    for f, a in fields.items():
        o.set_field(f, a)
    o.step(untimed)
    ts = []
    for _ in range(timed):
        t0 = time.perf_counter()
        o.step(1)
        ts.append((time.perf_counter() - t0) * 1e3)
    o.close()
    return ts


def cpu_baseline(n, phiflag, fields, dt, untimed=2, timed=5):
    """Oracle (the CPU restatement, C + OpenMP, gcc -O3) on the same workload:
    median wall time of `timed` steps after `untimed` ones, on the host's
    physical cores (lscpu) up to one GPU's share of the box."""
    import statistics
    sockets, phys = host_cores()
    threads = max(1, min(phys, BOX_CPU_SHARE))
    ts = _oracle_step_times(n, phiflag, fields, dt, threads, untimed, timed)
    return {"value": round(statistics.median(ts), 3), "unit": "ms/timestep", "cores": threads, "kind": "port",
            "host": {"sockets": sockets, "physical_cores": phys, "threads_used": threads},
            "sample": f"oracle/liboracle.so, {n}^3 {'NS+ACDI' if phiflag else 'NS'} workload above, "
                      f"median of {timed} steps after {untimed} untimed, {threads} OpenMP threads"}


def cpu_baseline_1thread(n_big=512):
    """The oracle on one host thread.  One 512^3 step takes tens of seconds on
    one thread, so the figure for the bench box is extrapolated from measured
    128^3 and 256^3 steps (same generators, 8-64 drops): t(512) = t(256) *
    (t(256) / t(128)), the growth per doubling measured one size down."""
    import statistics
    meas = {}
    for n, nd, r in ((128, 8, 0.3), (256, 64, 0.2)):
        u, v, w, phi, dt = make_inputs(n, 1, ndrops=nd, radius=r)
        ts = _oracle_step_times(n, 1, {1: u, 2: v, 3: w, 5: phi}, dt, 1, 1, 3)
        meas[n] = statistics.median(ts)
    est = meas[256] * (meas[256] / meas[128])
    return {"value": round(est, 1), "unit": "ms/timestep", "cores": 1, "kind": "port",
            "measured_ms": {f"{k}^3": round(v, 3) for k, v in meas.items()},
            "sample": f"oracle/liboracle.so on 1 thread, NS+ACDI: median of 3 steps after 1 untimed at 128^3 "
                      f"(8 drops R=0.3) and 256^3 (64 drops R=0.2); {n_big}^3 extrapolated as "
                      "t(256) * t(256) / t(128)"}


def stream_copy_gbs(torch, nbytes=1 << 30, reps=10):
    """Measured stream-copy ceiling: a plain 1 GiB device-to-device copy
    (read + write bytes over its event-timed duration)."""
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    torch.cuda.empty_cache()
    return round(2.0 * nbytes / (ms * 1e-3) / 1e9, 1)


def request_latency(res):
    """Request-level latency: submit -> finish (ensemble clock) per request,
    p50 and p99 over the requests (nearest rank)."""
    import numpy as np
    lat = np.sort(np.array([r["finish_ms"] - r["submit_ms"] for r in res], dtype=float))
    rank = lambda q: float(lat[min(len(lat) - 1, max(0, int(np.ceil(q * len(lat))) - 1))])
    return {"request_latency_p50_ms": round(rank(0.50), 3), "request_latency_p99_ms": round(rank(0.99), 3),
            "requests": int(len(lat))}


def ensemble_leg(count=16, steps=(200, 800), inflight=2, single_box_gcell_steps_s=None):
    """The serving stream: `count` seeded Poisson-arrival requests (40/s) of
    independent boxes, N in {64, 96, 128, 192, 256}, nu in [0.002, 0.02],
    half with ACDI drops, mixed precision.  Run one box at a time (create,
    step, copy out, destroy) and through Ensemble (submitted at their arrival
    times, continuous admission), compare results, report throughput in cell
    updates per second and the step latency percentiles."""
    import numpy as np
    from mhit36_amd import serve
    reqs = serve.serving_stream(count=count, steps=steps)
    cells = serve.cell_steps(reqs)
    serve.pretune(reqs)                      # launch shapes cached before either timed run
    t_one, out_one = serve.run_one_at_a_time(reqs)
    wall, out_ens, lat, res = serve.run_ensemble(reqs, follow_arrivals=True, inflight=inflight)
    same = all(np.array_equal(out_one[q], out_ens[q]) for q in out_one)
    small = small_box_leg(steps=steps)
    return {"workload": f"{count} requests, N in {list(serve.SIZES)}, {steps[0]}-{steps[1]} steps, "
                        "half ACDI, mixed precision, Poisson arrivals 40/s",
            "cell_steps": cells, "one_at_a_time_s": round(t_one, 4), "ensemble_s": round(wall, 4),
            "throughput_gcell_steps_s": round(cells / wall / 1e9, 3),
            "one_at_a_time_gcell_steps_s": round(cells / t_one / 1e9, 3),
            "speedup": round(t_one / wall, 3), "step_latency_p50_ms": round(lat["p50_ms"], 4),
            "step_latency_p99_ms": round(lat["p99_ms"], 4), "steps_timed": lat["steps"],
            **request_latency(res),
            "single_box_ceiling_gcell_steps_s": single_box_gcell_steps_s,
            "fraction_of_single_box_ceiling": (round(cells / wall / 1e9 / single_box_gcell_steps_s, 3)
                                               if single_box_gcell_steps_s else None),
            "same_alone_and_in_mix": bool(same), "status": sorted(set(r["status"] for r in res)),
            "inflight": inflight, "small_boxes": small}


def small_box_leg(count=32, steps=(200, 800)):
    """A stream of small boxes only (N = 64, otherwise as the serving stream),
    all submitted at once: one box at a time, the ensemble with each box on its
    own stream (gang_size 1), and the ensemble stepping boxes of one class by
    shared launches (gangs of 8)."""
    import numpy as np
    from mhit36_amd import serve
    reqs = serve.serving_stream(count=count, steps=steps, sizes=(64,))
    cells = serve.cell_steps(reqs)
    serve.pretune(reqs)
    t_one, out_one = serve.run_one_at_a_time(reqs)
    out = {"workload": f"{count} requests, N = 64, {steps[0]}-{steps[1]} steps, half ACDI, mixed precision, "
                       "submitted at once", "cell_steps": cells, "one_at_a_time_s": round(t_one, 4)}
    for gang, key in ((1, "streams"), (8, "gangs")):
        wall, o, lat, res = serve.run_ensemble(reqs, follow_arrivals=False, gang_size=gang)
        out[key] = {"ensemble_s": round(wall, 4), "speedup": round(t_one / wall, 3),
                    "throughput_gcell_steps_s": round(cells / wall / 1e9, 3),
                    "step_latency_p50_ms": round(lat["p50_ms"], 4), "step_latency_p99_ms": round(lat["p99_ms"], 4),
                    **request_latency(res),
                    "same_alone_and_in_mix": bool(all(np.array_equal(out_one[q], o[q]) for q in out_one))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--phiflag", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble-serving leg")
    ap.add_argument("--no-autotune", action="store_true", help="heuristic launch shapes only")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (3 times), the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from mhit36_amd import Solver

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 8}), flush=True)
        return
    copy_gbs = stream_copy_gbs(torch)
    n, phiflag = args.n, args.phiflag
    t_ic = time.perf_counter()
    u, v, w, phi, dt = make_inputs(n, phiflag)
    fields = {1: u, 2: v, 3: w}
    if phiflag:
        fields[5] = phi
    print(f"[bench] inputs {n}^3 ready in {time.perf_counter() - t_ic:.1f}s, dt={dt:.4e}",
          file=sys.stderr, flush=True)

    tunings = {}

    def timed_run(precision, kernel_timing):
        s = Solver(**params_for(n, phiflag, dt), precision=precision)
        if not args.no_autotune:
            tunings["fp64" if precision == 0 else "mixed"] = s.autotune()
        for f, a in fields.items():
            s.set_field(f, a)
        s.step(args.warmup)
        if kernel_timing:
            s.kernel_timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d = s.step(args.steps)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        stats = s.kernel_stats() if kernel_timing else []
        s.kernel_timing(False)
        dev_bytes = s.device_bytes()
        s.close()
        return elapsed * 1e3 / args.steps, d, stats, dev_bytes

    ms, d, stats, dev_bytes = timed_run(0, not args.no_kernel_timing)
    balg = alg_bytes_per_step(n, phiflag)
    print(f"[bench] FP64 {args.steps} steps: {ms:.3f} ms/step, gumax={d.gumax:.4f} cou={d.cou:.3f}",
          file=sys.stderr, flush=True)
    mixed = None
    if not args.no_mixed:
        ms_m, d_m, stats_m, bytes_m = timed_run(1, not args.no_kernel_timing)
        print(f"[bench] mixed {args.steps} steps: {ms_m:.3f} ms/step", file=sys.stderr, flush=True)
        mixed = {"ms_per_step": round(ms_m, 4), "speedup_vs_fp64": round(ms / ms_m, 3),
                 "device_bytes": bytes_m, "gumax": d_m.gumax,
                 "kernels": [{"name": k["name"], "avg_ms": round(k["avg_ms"], 4)} for k in stats_m],
                 "storage": "u,v,w,u*,rhs_o,phi FP32; p, spectrum, reductions FP64"}

    roof = None
    kernels = []
    if stats:
        for k in stats:
            avg = k["avg_ms"]
            gbs = k["bytes"] / (avg * 1e-3) / 1e9 if avg > 0 else 0.0
            kernels.append({"name": k["name"], "avg_ms": round(avg, 4), "launches": k["launches"],
                            "alg_bytes": k["bytes"], "gbs": round(gbs, 1)})
        dom = max(stats, key=lambda k: k["total_ms"])
        avg = dom["avg_ms"]
        ach = dom["bytes"] / (avg * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n, phiflag, dom["name"]),
                "kernel": dom["name"], "avg_ms": round(avg, 4), "alg_bytes_per_launch": dom["bytes"]}

    ens = None
    if not args.no_ensemble:
        # the rate one large box reaches on this GPU: 512^3 cells per mixed step
        ceiling = round(n ** 3 / (mixed["ms_per_step"] * 1e-3) / 1e9, 3) if mixed else None
        ens = ensemble_leg(single_box_gcell_steps_s=ceiling)
        print(f"[bench] ensemble {ens['ensemble_s']:.3f}s vs one at a time {ens['one_at_a_time_s']:.3f}s, "
              f"p99 {ens['step_latency_p99_ms']:.3f} ms", file=sys.stderr, flush=True)

    cpu = cpu1 = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, phiflag, fields, dt)
        cpu1 = cpu_baseline_1thread()

    out = {
        "metric": "ms/timestep (FP64 vs mixed) at 512^3 NS+ACDI; HBM GB/s; p99 ensemble latency",
        "value": round(ms, 4),
        "unit": "ms/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 {'NS+ACDI' if phiflag else 'NS-only'} FP64 forced HIT"
                               + (", 512 drops R=0.1" if phiflag else ""),
                   "nx": n, "phiflag": phiflag, "precision": "fp64", "dt": dt,
                   "kernel_timing_events": not args.no_kernel_timing},
        "hbm_gbs_step": round(balg / (ms * 1e-3) / 1e9, 1),
        "step_roofline_frac": round(balg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "alg_bytes_per_step": balg,
        "device_bytes": dev_bytes,
        "stream_copy_gbs": copy_gbs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "mixed": mixed,
        "tuning": {k: v["value"] for k, v in tunings.items()},
        "ensemble": ens,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
