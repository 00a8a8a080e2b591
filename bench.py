"""Benchmark of the D3Q19 ComputeUnit main loop on one MI355X.

Workload (the reference's GPU configuration): a Rushton-turbine tank on a 512^3
grid (T = 512, impeller D = T/3, 6 blades, clearance T/3, tip speed 0.1,
Re = 1e5 with nu from calcNuAndRe_m), FP32 storage and arithmetic,
Eggels-Somers collision with Smagorinsky LES (cs0 = 0.1), fluid initialised at
rest (rho = 1).  One step is one iteration of the reference main loop
(tdlbcpp/src/main.cpp:336-455): blade rotation, collision, bounce-back,
streaming, moments and IB forcing, all on the device.

Every streaming x layout variant is run, and Esotwist/LIJK also with 16-bit
population storage (fp16, bf16, e3m12); the headline value is FP32 Esotwist/LIJK.
value = interior cells x steps / seconds / 1e6 (MLUPS), timed with HIP events
on the ComputeUnit's stream around exactly --steps steps after --warmup steps.
The roofline entry uses the stream-collide kernel's own event-timed average
duration and its algorithmic bytes (2 x 18 x 4 B populations + 2 x 4 B Nu per
cell).  cpu_baseline times the CPU oracle (a restatement of the reference's
single-threaded CPU path) on one host core on a bounded sample, the same tank
at 96^3 for ~15 s, and reports beside it, labelled, the oracle with collision
and moments on OpenMP threads, and the host CPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & %HBM roofline, FP32 vs FP16-store; p50/p99 latency of batched tanks"
HBM_PEAK_GBS = 8000.0


class Hip:
    """HIP events on a given stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.L.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.L.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.L.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.L.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.L.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def device_sync(self):
        assert self.L.hipDeviceSynchronize() == 0


def bytes_per_cell(pop_bytes=4, les=True, nu_bytes=4):
    """Algorithmic bytes of one fused stream-collide cell update: 18 populations
    read and written, plus Nu read and written with LES."""
    return 2 * 18 * pop_bytes + (2 * nu_bytes if les else 0)


POP_BYTES = {0: 4, 1: 8, 2: 2, 3: 2, 4: 2}


def run_variant(td, hip, grid, streaming, layout, steps, warmup, precision=0, re=1.0e5, uav=0.1):
    flow = td.flow_for_tank(grid, re=re, uav=uav, cs0=0.1, use_les=1)
    cu = td.ComputeUnit(grid, layout=layout, streaming=streaming, precision=precision, **flow)
    rp = td.rushton_defaults(grid, tip_speed=uav, startup_steps=30)
    t0 = time.time()
    cu.initialise(1.0)
    cu.bind_rushton(rp, step=1, angle=0.0)
    cu.synchronize()
    setup_s = time.time() - t0
    cu.run(warmup)
    cu.synchronize()
    s = cu.stream()
    e0, e1 = hip.event(), hip.event()
    cu.profile(True)
    hip.record(e0, s)
    cu.run(steps)
    hip.record(e1, s)
    total_ms = hip.elapsed_ms(e0, e1)
    coll_ms, ncoll, force_ms, nforce = cu.profile_read()
    cu.profile(False)
    n_static, n_rot, n_levels = cu.forcing_info()
    cells = grid ** 3
    res = dict(mlups=cells * steps / (total_ms / 1e3) / 1e6, ms_per_step=total_ms / steps,
               collide_ms=coll_ms / max(ncoll, 1), forcing_ms=force_ms / max(nforce, 1),
               ib_points_static=n_static, ib_points_rotating=n_rot, forcing_levels=n_levels,
               setup_s=round(setup_s, 2))
    res["bytes_per_cell"] = bytes_per_cell(POP_BYTES[precision], nu_bytes=POP_BYTES[precision] if precision >= 2
                                           else 8 if precision == 1 else 4)
    res["collide_gbs"] = res["bytes_per_cell"] * cells / (res["collide_ms"] / 1e3) / 1e9
    cu.close()
    return res


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu": model, "nproc": os.cpu_count(), "cpus_available": avail}


def cpu_baseline(sample_grid=96, max_seconds=20.0, omp_seconds=8.0, omp_threads=16):
    """The oracle (CPU restatement of the reference path) on the same tank at
    sample_grid^3, Simple streaming, IJKL, FP32, LES: the configuration main.cpp
    runs on the CPU.  The baseline is one host core, as the reference is
    single-threaded; a second leg runs the oracle's collision and moments on
    OpenMP threads (push, bounce-back and forcing stay serial), labelled apart."""
    import tdlbcpp_amd as td
    from oracle import oracle as O
    from tests.helpers import ReferenceLoop, flow_for_tank
    oflow, _ = flow_for_tank(sample_grid, re=7300.0)
    rp = td.rushton_defaults(sample_grid)

    def timed(threads, seconds):
        O.set_threads(threads)
        o = O.OracleCU(sample_grid, dtype=O.F32, layout=O.IJKL, streaming=O.SIMPLE, use_les=True, **oflow)
        o.initialise(1.0)
        loop = ReferenceLoop(o, rp)
        loop.run(1)  # warm caches
        steps, t0 = 0, time.time()
        while True:
            loop.run(1)
            steps += 1
            el = time.time() - t0
            if el > seconds or steps >= 500:
                break
        return sample_grid ** 3 * steps / el / 1e6, steps, el

    host = host_cpu()
    try:
        mlups, steps, el = timed(1, max_seconds)
        out = dict(value=round(mlups, 3), unit="MLUPS", cores=1, kind="port",
                   sample=f"oracle/liboracle.so: Rushton tank {sample_grid}^3, Simple/IJKL, FP32, LES, "
                          f"{steps} main-loop steps in {el:.1f} s on one host core", host=host)
        threads = min(omp_threads, host["cpus_available"] or 1)
        if threads > 1 and omp_seconds > 0:
            mlups, steps, el = timed(threads, omp_seconds)
            out["openmp"] = dict(value=round(mlups, 3), unit="MLUPS", cores=threads,
                                 sample=f"same tank, collision and moments on {threads} OpenMP threads, push, "
                                        f"bounce-back and forcing serial: {steps} steps in {el:.1f} s")
    finally:
        O.set_threads(1)
    return out


def batch_tanks(n_tanks=64, steps=2000, seed=1):
    """Independent Rushton tanks: grid edge drawn from {48, 64, 96, 128, 160},
    tip speed U[0.05, 0.12], Re U[2e3, 2e4] (numpy PCG64 seeded with `seed`)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    edges = rng.choice([48, 64, 96, 128, 160], size=n_tanks)
    tips = rng.uniform(0.05, 0.12, size=n_tanks)
    res = rng.uniform(2e3, 2e4, size=n_tanks)
    return [dict(edge=int(e), tip=float(t), re=float(r), steps=steps) for e, t, r in zip(edges, tips, res)]


def make_tank(td, spec, precision, streaming, layout):
    flow = td.flow_for_tank(spec["edge"], re=spec["re"], uav=spec["tip"], cs0=0.1, use_les=1)
    cu = td.ComputeUnit(spec["edge"], layout=layout, streaming=streaming, precision=precision, **flow)
    cu.initialise(1.0)
    cu.bind_rushton(td.rushton_defaults(spec["edge"], tip_speed=spec["tip"], startup_steps=100))
    return cu


def run_batch_workload(td, hip, n_tanks, steps, precision, streaming, layout, solo_sample):
    """Batched (one graph-replayed launch sequence for all tanks) against
    one-at-a-time runs of the same tanks."""
    specs = batch_tanks(n_tanks, steps)
    cells = sum(s["edge"] ** 3 for s in specs)
    t0 = time.time()
    tanks = [make_tank(td, s, precision, streaming, layout) for s in specs]
    setup = time.time() - t0
    b = td.Batch(precision, layout, streaming)
    t0 = time.time()
    for cu, s in zip(tanks, specs):
        b.add(cu, s["steps"])
    tune = time.time() - t0
    b.run(2)  # settles each member and captures the graph
    hip.device_sync()
    t0 = time.time()
    b.run(steps)
    hip.device_sync()
    el = time.time() - t0
    info = b.info()
    b.close()
    done = steps - 2
    batched = cells * done / el / 1e6
    # one at a time: the same tanks, each alone for its steps (timed on a sample of the tanks)
    sample = specs[:solo_sample] if solo_sample else specs
    scells = sum(s["edge"] ** 3 for s in sample)
    for cu in tanks:
        cu.close()
    solo_s = 0.0
    for s in sample:
        cu = make_tank(td, s, precision, streaming, layout)
        cu.run(2)
        hip.device_sync()
        t0 = time.time()
        cu.run(done)
        hip.device_sync()
        solo_s += time.time() - t0
        cu.close()
    solo = scells * done / solo_s / 1e6
    return dict(tanks=n_tanks, steps=done, cells=cells, batched_mlups=batched, batched_s=el, solo_mlups=solo,
                solo_sample_tanks=len(sample), speedup=batched / solo, setup_s=round(setup, 1),
                autotune_s=round(tune, 2), launches_per_step=info["launches_per_step"],
                tiles={int(e): t for e, t in zip([s["edge"] for s in specs], info["tiles"])})


def serve_requests(duration, rate=4.0, seed=1):
    """Poisson arrivals at `rate` per second over `duration` seconds: grid edge
    log-uniform in [64, 320], tip speed U[0.05, 0.12], steps U[500, 5000],
    Re U[2e3, 2e4] (numpy PCG64 seeded with `seed`)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        edge = int(round(float(np.exp(rng.uniform(np.log(64), np.log(320))))))
        out.append(dict(t=t, edge=edge, tip=float(rng.uniform(0.05, 0.12)), re=float(rng.uniform(2e3, 2e4)),
                        steps=int(rng.integers(500, 5001))))


def run_serve_workload(td, duration, rate, tick_steps, max_resident, seed=1):
    """Feed the arrivals to a tdlb server in real time for `duration` seconds;
    latency = finish - submit of the requests completed inside the window."""
    import numpy as np
    reqs = serve_requests(duration, rate, seed)
    s = td.Server(td.PREC_F16, td.LAYOUT_LIJK, td.STREAM_ESOTWIST, tick_steps=tick_steps, max_resident=max_resident)
    t0 = time.time()
    i, results, ticks, last = 0, [], 0, 0.0
    while True:
        now = time.time() - t0
        while i < len(reqs) and reqs[i]["t"] <= now:
            r = reqs[i]
            s.submit(r["edge"], r["tip"], r["re"], r["steps"])
            i += 1
        if now >= duration:
            break
        resident, queued, n = s.tick()
        ticks += 1
        results += s.poll()
        if resident == 0 and queued == 0 and i < len(reqs):
            time.sleep(max(0.0, min(0.05, reqs[i]["t"] - (time.time() - t0))))
        if now - last >= 30:
            print(f"[serve] t={now:.0f}s submitted {i} done {len(results)} resident {resident} queued {queued}",
                  file=sys.stderr, flush=True)
            last = now
    elapsed = time.time() - t0
    stats = s.stats()
    s.close()
    ok = [r for r in results if r["error"] == 0]
    lat = np.array([r["finish_s"] - r["submit_s"] for r in ok]) if ok else np.array([0.0])
    updates = sum(r["edge"] ** 3 * r["steps"] for r in ok)
    return dict(duration_s=round(elapsed, 1), rate_per_s=rate, submitted=i, completed=len(ok),
                failed=len(results) - len(ok), backlog=i - len(results), ticks=ticks, tick_steps=tick_steps,
                max_resident=max_resident, p50_latency_s=round(float(np.percentile(lat, 50)), 2),
                p99_latency_s=round(float(np.percentile(lat, 99)), 2), aggregate_mlups=updates / elapsed / 1e6,
                budget_gb=round(stats["budget_bytes"] / 1e9, 1))


def copy_calibration(gib):
    """A plain device copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel
    give the factor between those counters and bytes moved."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.time() - t0
    print(json.dumps({"copy_bytes_each_way": n * 4, "copies": 5, "gbs": round(2 * 5 * n * 4 / el / 1e9, 1)}),
          flush=True)
    return 0


def measured_copy_gbs(gib=2.0, reps=20):
    """Bytes read + written per second of torch's device copy of `gib` GiB,
    timed with events on torch's stream: the bandwidth a plain streaming
    kernel reaches on this GPU, beside the 8 TB/s peak."""
    import torch
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    gbs = 2 * reps * n * 4 / (e0.elapsed_time(e1) / 1e3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--variants", default="all", help="all, precision (fp32 + fp16/bf16/e3m12 Esotwist/LIJK) or headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-omp-seconds", type=float, default=8.0,
                    help="seconds of the OpenMP leg of the CPU baseline (0 skips it)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="seconds of oracle main-loop steps timed for cpu_baseline")
    ap.add_argument("--workload", default="tank", choices=["tank", "batch", "serve", "accuracy"],
                    help="tank: one 512^3 tank (headline); batch: independent tanks batched vs one at a time; "
                         "serve: a Poisson stream of tank requests; accuracy: fp32/fp16/bf16/e3m12 against the fp64 "
                         "build (use --grid 96 --steps 5000)")
    ap.add_argument("--accuracy-re", type=float, nargs="+", default=[1.0e5, 7300.0])
    ap.add_argument("--accuracy-streaming", nargs="+", default=["Esotwist"], choices=["Esotwist", "Simple"],
                    help="accuracy: streamings (the Simple tank is unstable at these Reynolds numbers, as the "
                         "reference's Simple wall treatment is)")
    ap.add_argument("--accuracy-precisions", nargs="+", default=["fp32", "fp16", "bf16", "e3m12"],
                    choices=["fp32", "fp16", "bf16", "e3m12"])
    ap.add_argument("--duration", type=float, default=600.0, help="serve: seconds of arrivals")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: arrivals per second")
    ap.add_argument("--tick-steps", type=int, default=50, help="serve: steps per scheduling tick")
    ap.add_argument("--max-resident", type=int, default=0, help="serve: tanks at once (0: memory-bound; 1: one at a time)")
    ap.add_argument("--tanks", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--solo-sample", type=int, default=0,
                    help="time one-at-a-time runs on the first N tanks only (0: all)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only copy GIB GiB device-to-device 5 times (reference kernel for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs one ComputeUnit on one GPU; --gpus {args.gpus} is not supported "
              "(one GPU runs one tank or a batch of tanks; a slab decomposition over GPUs is not built)",
              file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)

    import tdlbcpp_amd as td
    hip = Hip()
    if args.workload == "batch":
        r = run_batch_workload(td, hip, args.tanks, args.batch_steps, td.PREC_F16, td.STREAM_ESOTWIST,
                               td.LAYOUT_LIJK, args.solo_sample)
        print(f"[bench] batch: {r['batched_mlups']:.0f} MLUPS batched vs {r['solo_mlups']:.0f} one at a time "
              f"({r['speedup']:.2f}x)", file=sys.stderr, flush=True)
        out = {"metric": METRIC, "value": round(r["batched_mlups"], 1), "unit": "MLUPS", "n_gpus": 1,
               "steps": r["steps"], "warmup": 2, "ms_per_step": round(1e3 * r["batched_s"] / r["steps"], 4),
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16-store/f32",
               "data": "synthetic: Rushton tanks at rest, geometry from this library's generator",
               "config": {"workload": f"batched_rushton_tanks_{args.tanks}_fp16_esotwist_lijk", **r}}
        print(json.dumps(out), flush=True)
        return 0
    if args.workload == "accuracy":
        from tdlbcpp_amd import precision as P
        reports = {}
        names = {"fp32": td.PREC_F32, "fp16": td.PREC_F16, "bf16": td.PREC_BF16, "e3m12": td.PREC_E3M12}
        precs = tuple(names[p] for p in args.accuracy_precisions)
        for name in args.accuracy_streaming:
            st = td.STREAM_ESOTWIST if name == "Esotwist" else td.STREAM_SIMPLE
            for re_ in args.accuracy_re:
                r = P.precision_report(T=args.grid, steps=args.steps, precisions=precs, streaming=st, re=re_,
                                       samples=20)
                reports[f"{name}/LIJK/Re{re_:g}"] = r
                print(f"[bench] accuracy {name} Re={re_:g}: " + ", ".join(
                    f"{k} L2 {v['velocity_l2']:.2e} (mean field {v['velocity_l2_mean']:.2e}) Po err "
                    f"{v['po_rel_err']:.2e} mass drift {v['mass_drift']:.1e}" for k, v in r["variants"].items()),
                    file=sys.stderr, flush=True)
        print(json.dumps({"workload": f"accuracy_{args.grid}cubed_{args.steps}steps_vs_fp64", "reports": reports}),
              flush=True)
        return 0
    if args.workload == "serve":
        r = run_serve_workload(td, args.duration, args.rate, args.tick_steps, args.max_resident)
        print(f"[bench] serve: p50 {r['p50_latency_s']} s, p99 {r['p99_latency_s']} s, "
              f"{r['aggregate_mlups']:.0f} MLUPS, {r['completed']}/{r['submitted']} done",
              file=sys.stderr, flush=True)
        out = {"metric": METRIC, "value": round(r["aggregate_mlups"], 1), "unit": "MLUPS", "n_gpus": 1,
               "steps": r["ticks"], "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f16-store/f32",
               "data": "synthetic: Poisson stream of Rushton tank requests, geometry from this library's generator",
               "config": {"workload": "served_rushton_tanks_fp16_esotwist_lijk", **r}}
        print(json.dumps(out), flush=True)
        return 0
    F32, F16, BF16, E3M12 = td.PREC_F32, td.PREC_F16, td.PREC_BF16, td.PREC_E3M12
    variants = [("Esotwist/LIJK", td.STREAM_ESOTWIST, td.LAYOUT_LIJK, F32)]
    if args.variants in ("all", "precision"):
        variants += [("Esotwist/LIJK/fp16", td.STREAM_ESOTWIST, td.LAYOUT_LIJK, F16),
                     ("Esotwist/LIJK/bf16", td.STREAM_ESOTWIST, td.LAYOUT_LIJK, BF16),
                     ("Esotwist/LIJK/e3m12", td.STREAM_ESOTWIST, td.LAYOUT_LIJK, E3M12)]
    if args.variants == "all":
        variants += [("Esotwist/IJKL", td.STREAM_ESOTWIST, td.LAYOUT_IJKL, F32),
                     ("Simple/LIJK", td.STREAM_SIMPLE, td.LAYOUT_LIJK, F32),
                     ("Simple/IJKL", td.STREAM_SIMPLE, td.LAYOUT_IJKL, F32),
                     ("Esotwist/IJKL_SOA", td.STREAM_ESOTWIST, td.LAYOUT_IJKL_SOA, F32),
                     ("Simple/IJKL_SOA", td.STREAM_SIMPLE, td.LAYOUT_IJKL_SOA, F32),
                     ("Simple/LIJK/fp16", td.STREAM_SIMPLE, td.LAYOUT_LIJK, F16)]
    results = {}
    for name, st, ly, prec in variants:
        r = run_variant(td, hip, args.grid, st, ly, args.steps, args.warmup, precision=prec)
        results[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        print(f"[bench] {name}: {r['mlups']:.0f} MLUPS, {r['ms_per_step']:.3f} ms/step, "
              f"collide {r['collide_ms']:.3f} ms ({r['collide_gbs']:.0f} GB/s), forcing {r['forcing_ms']:.3f} ms",
              file=sys.stderr, flush=True)
    head = results["Esotwist/LIJK"]
    achieved = head["collide_gbs"]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_esotwist_lijk_512.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC, "value": round(head["mlups"], 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: Rushton tank at rest, geometry from this library's generator",
        "config": {"workload": f"rushton_tank_{args.grid}cubed_esotwist_lijk_fp32_les", "grid": args.grid,
                   "streaming": "Esotwist", "layout": "LIJK", "precision": "fp32", "les_cs0": 0.1,
                   "reynolds": 1.0e5, "tip_speed": 0.1, "variants": results},
        "fp16_store_speedup": (round(results["Esotwist/LIJK/fp16"]["mlups"] / head["mlups"], 3)
                               if "Esotwist/LIJK/fp16" in results else None),
        "e3m12_store_speedup": (round(results["Esotwist/LIJK/e3m12"]["mlups"] / head["mlups"], 3)
                                if "Esotwist/LIJK/e3m12" in results else None),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_esotwist<float,LIJK,FUSED>",
                     "bytes_per_cell": bytes_per_cell(), "cells_per_launch": args.grid ** 3},
    }
    copy = measured_copy_gbs()
    out["roofline"]["copy_gbs"] = round(copy, 1)
    out["roofline"]["frac_of_copy"] = round(achieved / copy, 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(max_seconds=args.cpu_seconds, omp_seconds=args.cpu_omp_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
