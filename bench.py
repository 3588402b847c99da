#!/usr/bin/env python3
"""Benchmark of the FP64 SU(3) leapfrog (Dynamics::SUNc::Update) on one MI355X.

Workload: the large single-event configuration, 256 x 256 x 1024 sites (67 108 864), a = 1, dt = a/4,
open z boundaries.  A "step" is one leapfrog step of the whole lattice (one launch of the fused step
kernel).  Inputs are resident in HBM before timing starts.  The fields are an MV collision initial state
built on the device (--init mv, the default) or synthetic random SU(3) links and random E / J
(--init random); the per-site work of the step does not depend on the field values.

Prints ONE JSON line:  metric, value (site-updates/s), roofline (HBM, 1760 algorithmic bytes per
site-update / measured kernel time), cpu_baseline (the oracle timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_SITE_UPDATE = 1760  # U 432 + E 192 + J+,J-,J+stat,J-stat 4*64, each read once and written once
HBM_PEAK_GBS = 8000.0         # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "leapfrog_step_256x256x1024_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per step-kernel launch from the committed rocprofv3 --pmc summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def host_cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sample_n, steps):
    """The oracle (CPU restatement, OpenMP) on a bounded sample of the same workload."""
    from oracle import oracle as O
    from soerenpragya3dglasma_amd import Glasma
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.lib().ref_set_threads(threads)
    with Glasma(sample_n, (1.0, 1.0, 1.0), 0.25) as g:
        g.fill_random(seed=7, link_scale=0.5, e_scale=0.1, j_scale=0.05)
        d = g.get_fields()
    f = O.Fields(sample_n)
    f.U, f.E, f.Jp, f.Jm, f.JpS, f.JmS = d["U"], d["E"], d["Jp"], d["Jm"], d["JpS"], d["JmS"]
    O.update(f, (1.0, 1.0, 1.0), 0.25, 1)  # warm-up step
    t0 = time.perf_counter()
    O.update(f, (1.0, 1.0, 1.0), 0.25, steps)
    dt = time.perf_counter() - t0
    sites = sample_n[0] * sample_n[1] * sample_n[2]
    return {
        "value": sites * steps / dt,
        "unit": "site-updates/s",
        "cores": threads,
        "cpu_model": host_cpu_model(),
        "logical_cpus": os.cpu_count(),
        "kind": "port",
        "sample": f"oracle (C++/OpenMP restatement of Dynamics::SUNc::Update) on a {sample_n[0]}x{sample_n[1]}x"
                  f"{sample_n[2]} slab, {steps} timed steps after 1 warm-up ({dt:.1f} s)",
    }


def reference_case(threads: int) -> dict:
    """The reference's own CPU-sized case in full: a 32 x 32 x 64 left-moving MV nucleus (g^2 mu 0.5, width 4,
    m 0.1, charge seed 1, dt = a/4), 100 steps, on the GPU and with the oracle on the host; the energy after
    100 steps is compared as a parity check."""
    from oracle import oracle as O
    from soerenpragya3dglasma_amd import Glasma
    N, a, dt, steps = (32, 32, 64), (1.0, 1.0, 1.0), 0.25, 100
    xl, xr = Glasma.draw_charges(1, N[0], N[1], 0.5, 0.5)
    kw = dict(lam=1.65, meff=0.1, sigma_l=4.0, sigma_r=4.0, center_l=N[2] // 2, center_r=N[2] // 2)
    with Glasma(N, a, dt) as g:
        g.mv_initial_state(1, Glasma.mv_params(**kw), xl, xr)
        g.Update(1)
        g.synchronize()
        t0 = time.perf_counter()
        g.update_async(steps)
        g.synchronize()
        gpu_s = time.perf_counter() - t0
        g.mv_initial_state(1, Glasma.mv_params(**kw), xl, xr)
        g.Update(steps)
        t00_gpu = g.bulk()["T00"]
    f = O.mv_initial_state(1, N, a, dt, xl, xr, **kw)
    t0 = time.perf_counter()
    O.update(f, a, dt, steps)
    cpu_s = time.perf_counter() - t0
    t00_cpu = O.bulk_T(O.bulk(f, a), a)["T00"]
    return {"lattice": list(N), "steps": steps, "gpu_ms_per_step": 1e3 * gpu_s / steps,
            "cpu_ms_per_step": 1e3 * cpu_s / steps, "cpu_cores": threads, "gpu_over_cpu": cpu_s / gpu_s,
            "t00_rel_diff": abs(t00_gpu - t00_cpu) / abs(t00_cpu)}


def batched_ensemble(nevents: int, steps: int, lattice=None) -> dict:
    """The ensemble workload: events drawn with master seed 12345 (numpy RandomState), N_perp uniform in
    {32, 48, 64, 96}, Nz uniform in {128, 192, 256, 384} (or all on `lattice`), g^2 mu uniform in [0.3, 0.8],
    charge seeds 2k+1 (left) and 2k+2 (right), MV collision initial states built on the device; `steps`
    leapfrog steps of every event, batched (one launch per step over all events, replayed from a HIP graph)
    against serial per-event launches.  Both runs start from the same initial states."""
    from soerenpragya3dglasma_amd import Glasma, GlasmaBatch
    rng = np.random.RandomState(12345)
    specs, qs = [], []
    for _ in range(nevents):
        nperp = int(rng.choice([32, 48, 64, 96]))
        nz = int(rng.choice([128, 192, 256, 384]))
        qs.append(float(rng.uniform(0.3, 0.8)))
        specs.append((tuple(lattice) if lattice else (nperp, nperp, nz), (1.0, 1.0, 1.0), 0.25))
    sites = sum(sp[0][0] * sp[0][1] * sp[0][2] for sp in specs)
    log(f"[bench] ensemble: {nevents} events, {sites} sites in total, {steps} steps")
    out = {"events": nevents, "steps": steps, "sites_total": sites}
    with GlasmaBatch(specs) as b:
        t0 = time.perf_counter()
        for k, (ev, sp, q) in enumerate(zip(b.events, specs, qs)):
            N = sp[0]
            xl, _ = Glasma.draw_charges(2 * k + 1, N[0], N[1], q, q)
            _, xr = Glasma.draw_charges(2 * k + 2, N[0], N[1], q, q)
            mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=4.0, sigma_r=4.0,
                                  center_l=int(0.85 * N[2] / 2 - 1), center_r=int(1.15 * N[2] / 2 - 1))
            ev.mv_initial_state(0, mv, xl, xr)
        log(f"[bench] ensemble initial states in {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        b.tune(reps=2)  # one tile order per distinct lattice, used by the batched and the serial runs
        log(f"[bench] ensemble tile orders tuned in {time.perf_counter() - t0:.1f} s")
        b.update(1)  # graph capture, warm-up
        b.update_serial(1)
        b.update(steps)
        t_batch = b.last_ms() * 1e-3
        b.update_serial(steps)
        t_serial = b.last_ms() * 1e-3
    out.update({"batched_s": t_batch, "serial_s": t_serial, "events_per_s_batched": nevents / t_batch,
                "events_per_s_serial": nevents / t_serial, "batched_over_serial": t_serial / t_batch,
                "site_updates_per_s_batched": sites * steps / t_batch})
    log(f"[bench] ensemble: batched {t_batch:.2f} s, serial {t_serial:.2f} s, ratio {t_serial / t_batch:.2f}")
    return out


def gauge_fixing_collision(iters: int) -> dict:
    """Coulomb gauge fixing of the 64 x 64 x 256 MV collision (charge seeds 1 / 2, width 6, centres 64 apart)
    after 50 steps, the state the reference's collision workload fixes every 50 steps: ms per iteration."""
    from soerenpragya3dglasma_amd import Glasma
    N = (64, 64, 256)
    xl, _ = Glasma.draw_charges(1, N[0], N[1], 0.5, 0.5)
    _, xr = Glasma.draw_charges(2, N[0], N[1], 0.5, 0.5)
    mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=6.0, sigma_r=6.0, center_l=128 - 32, center_r=128 + 32)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        g.mv_initial_state(0, mv, xl, xr)
        g.Update(50)
        g.coulomb_gauge(tol=0.0, maxit=1, alpha=2.0)
        t0 = time.perf_counter()
        st = g.coulomb_gauge(tol=0.0, maxit=iters, alpha=2.0)
        t1 = time.perf_counter()
        g.gauge_reset()
        conv = converging_gauge(g, 1e-10, 500)
    n = max(1, int(st[0]))
    return {"iterations": int(st[0]), "ms_per_iteration": 1e3 * (t1 - t0) / n,
            "iterations_per_s": n / (t1 - t0), "max_deviation": float(st[1]), "consistent_to_1e-10": conv}


def converging_gauge(g, tol: float, maxit: int) -> dict:
    """The boundary-consistent Fourier update (GLASMA_GAUGE_FOURIER_CONSISTENT) from G = 1 to max|Delta| <= tol."""
    t0 = time.perf_counter()
    st = g.coulomb_gauge_algorithm("consistent", tol, maxit, 2.0)
    dt = time.perf_counter() - t0
    n = max(1, int(st[0]))
    return {"iterations": int(st[0]), "max_deviation": float(st[1]), "s_total": dt, "ms_per_iteration": 1e3 * dt / n,
            "converged": bool(st[1] <= tol)}


def adjoint_cost(N=(48, 48, 128), steps: int = 200, ckpt: int = 20) -> dict:
    """Gradient of the slab energy (|z - Nz/2| < 8) after `steps` steps of a left-moving MV sheet with
    respect to its charge density rho(x), against the forward evaluation of the same energy."""
    from soerenpragya3dglasma_amd import Glasma
    xl, _ = Glasma.draw_charges(1, N[0], N[1], 0.5, 0.5)
    z = np.arange(N[2])
    center = int(0.85 * N[2] / 2 - 1)
    prof = np.exp(-0.5 * ((z - center) / 4.0) ** 2) / np.sqrt(2 * np.pi * 16.0)
    rho = prof[:, None, None, None] * xl[None]
    mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=4.0, sigma_r=4.0, center_l=center, center_r=center)
    with Glasma(N, (1.0, 1.0, 1.0), 0.25) as g:
        # warm-up calls of both at the timed sizes: the context keeps their work arrays, as a caller that repeats
        # them (a training loop) sees
        g.slab_energy_from_charge(mv, rho, 2, N[2] // 2, 8)
        g.slab_energy_gradient(mv, rho, steps, ckpt, N[2] // 2, 8)
        t0 = time.perf_counter()
        e = g.slab_energy_from_charge(mv, rho, steps, N[2] // 2, 8)
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        e2, grad = g.slab_energy_gradient(mv, rho, steps, ckpt, N[2] // 2, 8)
        t_adj = time.perf_counter() - t0
    log(f"[bench] adjoint: forward {t_fwd:.3f} s, forward+gradient {t_adj:.3f} s, ratio {t_adj / t_fwd:.1f}")
    return {"lattice": list(N), "steps": steps, "checkpoint_every": ckpt, "energy": e, "forward_s": t_fwd,
            "gradient_s": t_adj, "gradient_over_forward": t_adj / t_fwd, "grad_norm": float(np.linalg.norm(grad)),
            "energies_agree": abs(e - e2) <= 1e-12 * abs(e)}


def calibrate_copy(gib: float) -> None:
    """A plain device-to-device copy kernel of known size, for scaling the HBM byte counters of rocprofv3.

    The copy reads and writes exactly ``gib`` GiB per launch; the FETCH_SIZE / WRITE_SIZE that the
    profiler reports for it, divided by that size, is the counter's scale factor on this device.
    """
    import torch
    n = int(gib * 2 ** 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    # the copy's own rate: the bandwidth a plain read-and-write stream reaches on this device
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                      "launches": 13, "ms_per_copy": ms, "read_plus_write_GBps": 2 * n * 8 / ms / 1e6}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-copy", type=float, default=None, metavar="GIB",
                    help="run only a plain copy kernel of GIB GiB (profiler counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, nargs=3, default=[256, 256, 1024])
    ap.add_argument("--init", choices=["mv", "random"], default="mv")
    ap.add_argument("--cpu-sample", type=int, nargs=3, default=[256, 256, 128])
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-events", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile order instead of tuning it")
    ap.add_argument("--no-gauge", action="store_true", help="skip the Coulomb gauge-fixing timing")
    ap.add_argument("--gauge-iters", type=int, default=5)
    args = ap.parse_args()

    if args.calibrate_copy is not None:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU (z-slab) decomposition is a design, "
              "not an implementation. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from soerenpragya3dglasma_amd import Glasma
    N = tuple(args.n)
    S = N[0] * N[1] * N[2]
    a = (1.0, 1.0, 1.0)
    dt = 0.25
    log(f"[bench] lattice {N} ({S} sites), init={args.init}")
    g = Glasma(N, a, dt)
    init = args.init
    ic_ms = None
    if init == "mv":
        # left nucleus from seed 1, right nucleus from seed 2, each drawn as the reference draws it
        xl, _ = Glasma.draw_charges(1, N[0], N[1], 0.5, 0.5)
        _, xr = Glasma.draw_charges(2, N[0], N[1], 0.5, 0.5)
        mv = Glasma.mv_params(lam=1.65, meff=0.1, sigma_l=8.0, sigma_r=8.0,
                              center_l=int(0.85 * N[2] / 2 - 1), center_r=int(1.15 * N[2] / 2 - 1))
        t0 = time.perf_counter()
        g.mv_initial_state(0, mv, xl, xr)
        ic_ms = 1e3 * (time.perf_counter() - t0)
        log(f"[bench] MV collision initial state built on the device in {ic_ms / 1e3:.2f} s")
    else:
        g.fill_random(seed=1, link_scale=0.5, e_scale=0.1, j_scale=0.05)
    tune = None
    if not args.no_tune:
        t0 = time.perf_counter()
        tune = g.tune_step(reps=3)
        log(f"[bench] tile order tuned in {time.perf_counter() - t0:.2f} s: {tune['chosen']} "
            f"({', '.join(f'{k} {v:.2f}' for k, v in tune['candidates_ms'].items())} ms)")
    log(f"[bench] warmup {args.warmup} steps")
    g.update_async(args.warmup)
    g.synchronize()
    log(f"[bench] timing {args.steps} steps")
    g.synchronize()
    t0 = time.perf_counter()
    g.update_async(args.steps)
    g.synchronize()
    t1 = time.perf_counter()
    kern_ms = g.last_step_kernel_ms()
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = S * args.steps / wall
    achieved = BYTES_PER_SITE_UPDATE * S / (kern_ms * 1e-3) / 1e9
    # the device's ceiling for exactly this traffic: a copy of the step's 880 + 880 B per site with its layout
    ceil_ms = g.copy_ceiling(5)
    ceiling = {"copy_ms": ceil_ms, "GBps": BYTES_PER_SITE_UPDATE * S / (ceil_ms * 1e-3) / 1e9,
               "step_over_copy": kern_ms / ceil_ms}
    log(f"[bench] copy of the step's compulsory traffic: {ceil_ms:.3f} ms = {ceiling['GBps']:.0f} GB/s")
    umax, _ = g.unitarity()
    g_order = g.tile_order()[0]
    gauge = None
    if not args.no_gauge and args.gauge_iters > 0:
        # Coulomb gauge fixing of the evolved state (Fourier-accelerated, two 3D FFTs of 8 components per
        # iteration); the first call allocates the work arrays and is timed apart
        t0 = time.perf_counter()
        g.coulomb_gauge(tol=0.0, maxit=1, alpha=2.0)
        t1 = time.perf_counter()
        st = g.coulomb_gauge(tol=0.0, maxit=args.gauge_iters, alpha=2.0)
        t2 = time.perf_counter()
        n1 = max(1, int(st[0]))
        gauge = {"iterations": int(st[0]), "ms_per_iteration": 1e3 * (t2 - t1) / n1,
                 "iterations_per_s": n1 / (t2 - t1),
                 "first_call_ms": 1e3 * (t1 - t0), "max_deviation": float(st[1])}
        # a call of four times as many iterations (continuing from the same G): the cost of one more iteration,
        # and the fixed cost of a call (U^G and E^G formed once from the final G), which the short call's
        # ms_per_iteration spreads over its few iterations
        st2 = g.coulomb_gauge(tol=0.0, maxit=4 * args.gauge_iters, alpha=2.0)
        t3 = time.perf_counter()
        n2 = int(st2[0])
        if n2 > n1:
            per_it = ((t3 - t2) - (t2 - t1)) / (n2 - n1)
            gauge["ms_per_added_iteration"] = 1e3 * per_it
            gauge["ms_fixed_per_call"] = 1e3 * ((t2 - t1) - n1 * per_it)
        log(f"[bench] Coulomb gauge fixing: {gauge['ms_per_iteration']:.1f} ms per iteration over {n1} iterations"
            + (f", {gauge['ms_per_added_iteration']:.1f} ms per added iteration and "
               f"{gauge['ms_fixed_per_call']:.1f} ms fixed per call" if n2 > n1 else ""))
        g.gauge_reset()
        gauge["consistent_to_1e-10"] = converging_gauge(g, 1e-10, 300)
        log(f"[bench] consistent gauge fixing to 1e-10: {gauge['consistent_to_1e-10']}")
        gauge["collision_64x64x256"] = gauge_fixing_collision(args.gauge_iters)
    log(f"[bench] {ms_per_step:.3f} ms/step wall, {kern_ms:.3f} ms/kernel, {achieved:.0f} GB/s algorithmic, "
        f"unitarity max {umax:.2e}")
    g.close()

    batched = None
    if not args.no_batch:
        batched = batched_ensemble(args.batch_events, args.batch_steps)
        # where batching pays: events too small to fill the device alone, whose serial launches are bound by
        # launch latency (the reference's own CPU-sized lattice, and one of an eighth of its sites)
        batched["small_events"] = [
            {"lattice": list(lat), **{k: v for k, v in batched_ensemble(ne, 200, lat).items()
                                      if k in ("events", "steps", "batched_s", "serial_s", "batched_over_serial")}}
            for lat, ne in (((32, 32, 64), 512), ((16, 16, 32), 256))]

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_cost()

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle on the host)")
        try:
            cpu = cpu_baseline(tuple(args.cpu_sample), args.cpu_steps)
            cpu["reference_case"] = reference_case(cpu["cores"])
        except Exception as e:
            log(f"[bench] cpu baseline failed: {e}")

    out = {
        "metric": "FP64 site-updates/s @256^2x1024 (SU(3) leapfrog)",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (" + ("MV collision initial state built on the device, g2mu 0.5, sigma 8, m 0.1, "
                                 "Lambda 1.65, charge seeds 1 / 2" if init == "mv" else "random SU(3) fields") + ")",
        "config": {"workload": "large single-event FP64 collision, 256x256x1024, a=1, dt=a/4, open z",
                   "lattice": list(N), "sites": S, "init": init, "parallelism": "single GPU",
                   "tile_order": g_order, "tile_orders_ms": tune["candidates_ms"] if tune else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_step4", "kernel_ms": kern_ms, "bytes_per_site_update": BYTES_PER_SITE_UPDATE,
                     "compulsory_copy_ceiling": ceiling},
        "cpu_baseline": cpu,
        "ic_ms": ic_ms,
        "gauge_fixing": gauge,
        "batched": batched,
        "adjoint": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
