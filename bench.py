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
cell); roofline.copy_gbs is the speed of light of that access pattern, the
faster of two schedules of a probe that reads and writes back 19 planes of
grid^3 values in place (tdlb_probe_planes).  accuracy: the 16-bit storage
formats against the fp64 build of the same kernels on the 96^3 tank over 5000
steps (velocity L2 at the last step and of the mean field).  serving: a short
Poisson stream of tank requests below capacity, drained, p50/p99 over every
request.  cpu_baseline times the CPU oracle (a restatement of the reference's
single-threaded CPU path) on one host core: the headline tank for 3 steps
(extrapolated), the 60^3 tank for its full 20 steps, and, labelled apart, the
60^3 run with collision and moments on OpenMP threads; and the host CPU.
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


def cpu_baseline(grid=512, steps=3, omp_threads=16):
    """The oracle (CPU restatement of the reference path, one host core, as the
    reference's CPU build is single-threaded) on the reference's CPU
    configuration: Simple streaming, IJKL, FP32, LES cs0 = 0.1, Rushton tank at
    rest.  Legs: the headline tank at grid^3 for `steps` main-loop steps (the
    value: a sample of the 1000-step run, extrapolated), the 60^3 tank at
    Re 7300 for its full 20 steps, and that 60^3 run again with the oracle's
    collision and moments on OpenMP threads (push, bounce-back and forcing
    serial), labelled apart."""
    import tdlbcpp_amd as td
    from oracle import oracle as O
    from tests.helpers import ReferenceLoop, flow_for_tank

    def timed(G, n, threads, re):
        O.set_threads(threads)
        oflow, _ = flow_for_tank(G, re=re)
        o = O.OracleCU(G, dtype=O.F32, layout=O.IJKL, streaming=O.SIMPLE, use_les=True, **oflow)
        o.initialise(1.0)
        loop = ReferenceLoop(o, td.rushton_defaults(G))  # binds the geometry: its first forcing pass
        t0 = time.time()
        loop.run(n)
        el = time.time() - t0
        del loop, o
        return G ** 3 * n / el / 1e6, el

    host = host_cpu()
    try:
        mlups, el = timed(grid, steps, 1, 1.0e5)
        out = dict(value=round(mlups, 3), unit="MLUPS", cores=1, kind="port",
                   sample=f"oracle/liboracle.so: Rushton tank {grid}^3 at Re 1e5, Simple/IJKL, FP32, LES, {steps} "
                          f"main-loop steps in {el:.1f} s on one host core (a sample of the 1000-step run: "
                          f"1000 steps extrapolate to {1000 * el / steps / 3600:.1f} h)", host=host)
        c1, el1 = timed(60, 20, 1, 7300.0)
        out["config1"] = dict(value=round(c1, 3), unit="MLUPS", cores=1,
                              sample=f"60^3 tank at Re 7300, Simple/IJKL, FP32, LES: the full 20 steps in {el1:.2f} s")
        threads = min(omp_threads, host["cpus_available"] or 1)
        if threads > 1:
            c1t, el1t = timed(60, 20, threads, 7300.0)
            out["openmp"] = dict(value=round(c1t, 3), unit="MLUPS", cores=threads,
                                 sample=f"the 60^3 run with collision and moments on {threads} OpenMP threads, "
                                        f"push, bounce-back and forcing serial: 20 steps in {el1t:.2f} s")
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
    solo_s, graph_s = 0.0, 0.0
    for s in sample:
        # plain: tdlb_cu_run, the launches of each step issued from the host
        cu = make_tank(td, s, precision, streaming, layout)
        cu.run(2)
        hip.device_sync()
        t0 = time.time()
        cu.run(done)
        hip.device_sync()
        solo_s += time.time() - t0
        cu.close()
        # graph: the same tank alone in a batch (its own tuned tile, steps replayed from a HIP graph)
        cu = make_tank(td, s, precision, streaming, layout)
        b1 = td.Batch(precision, layout, streaming)
        b1.add(cu, s["steps"])
        b1.run(2)
        hip.device_sync()
        t0 = time.time()
        b1.run(done)
        hip.device_sync()
        graph_s += time.time() - t0
        b1.close()
        cu.close()
    solo = scells * done / solo_s / 1e6
    solo_graph = scells * done / graph_s / 1e6
    return dict(tanks=n_tanks, steps=done, cells=cells, batched_mlups=batched, batched_s=el, solo_mlups=solo,
                solo_graph_mlups=solo_graph, solo_sample_tanks=len(sample), speedup=batched / solo,
                speedup_vs_graph=batched / solo_graph, setup_s=round(setup, 1), autotune_s=round(tune, 2),
                launches_per_step=info["launches_per_step"], block_size=info.get("block_size"),
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


def run_serve_workload(td, duration, rate, tick_steps, max_resident, seed=1, drain_s=600.0):
    """Feed the arrivals to a tdlb server in real time for `duration` seconds,
    then keep ticking, with no new arrivals, until every request has finished or
    `drain_s` more seconds have passed.  Latency = finish - submit; a request
    still unfinished at the end counts with end - submit, a lower bound
    (censored), so p50/p99 cover every request submitted in the window."""
    import numpy as np
    reqs = serve_requests(duration, rate, seed)
    s = td.Server(td.PREC_F16, td.LAYOUT_LIJK, td.STREAM_ESOTWIST, tick_steps=tick_steps, max_resident=max_resident)
    t0 = time.time()
    i, results, ticks, last = 0, [], 0, 0.0
    submitted = {}  # ticket -> submit time (bench clock)
    while True:
        now = time.time() - t0
        while i < len(reqs) and reqs[i]["t"] <= now:
            r = reqs[i]
            submitted[s.submit(r["edge"], r["tip"], r["re"], r["steps"])] = time.time() - t0
            i += 1
        if now >= duration and (len(results) >= len(submitted) or now >= duration + drain_s):
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
    end = time.time() - t0
    stats = s.stats()
    s.close()
    ok = [r for r in results if r["error"] == 0]
    finished = {r["ticket"] for r in results}
    lat = [r["finish_s"] - r["submit_s"] for r in ok]
    censored = [end - t for k, t in submitted.items() if k not in finished]
    allv = np.array(lat + censored) if (lat or censored) else np.array([0.0])
    done_lat = np.array(lat) if lat else np.array([0.0])
    updates = sum(r["edge"] ** 3 * r["steps"] for r in ok)
    offered = sum(r["edge"] ** 3 * r["steps"] for r in reqs) / max(duration, 1e-9) / 1e6
    return dict(arrivals_s=duration, duration_s=round(end, 1), rate_per_s=rate, submitted=len(submitted),
                completed=len(ok), failed=len(results) - len(ok), unfinished=len(censored), ticks=ticks,
                tick_steps=tick_steps, max_resident=max_resident,
                p50_latency_s=round(float(np.percentile(allv, 50)), 2),
                p99_latency_s=round(float(np.percentile(allv, 99)), 2),
                p50_latency_completed_s=round(float(np.percentile(done_lat, 50)), 2),
                p99_latency_completed_s=round(float(np.percentile(done_lat, 99)), 2),
                offered_mlups=round(offered, 1), aggregate_mlups=updates / end / 1e6,
                budget_gb=round(stats["budget_bytes"] / 1e9, 1))


def sol_probe(grid, elem_bytes=4, planes=19, passes=10):
    """Speed of light of the fused kernel's access pattern (tdlb_probe_planes):
    `planes` planes of grid^3 values, as far apart as the planes of the LIJK
    lattice, each value read and written back in place per pass, with and
    without non-temporal hints."""
    import tdlbcpp_amd as td
    cells = grid ** 3
    stride = td.lijk_plane_bytes(grid, elem_bytes)
    out = {"planes": planes, "elem_bytes": elem_bytes, "cells": cells, "plane_stride_bytes": stride}
    for nt in (True, False):
        ms = td.probe_planes(elem_bytes, cells, planes, nontemporal=nt, passes=passes, stride_bytes=stride)
        out["nontemporal" if nt else "plain"] = {"ms": round(ms, 4),
                                                 "gbs": round(2 * planes * cells * elem_bytes / (ms / 1e3) / 1e9, 1)}
    out["gbs"] = max(out["nontemporal"]["gbs"], out["plain"]["gbs"])
    return out


def probe_calibration(gib):
    """One plane of `gib` GiB of 4-byte values read and written back in place,
    2 + 5 passes of k_probe_planes: the known-size kernel against which the
    FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import tdlbcpp_amd as td
    cells = int(gib * (1 << 30)) // 4
    ms = td.probe_planes(4, cells, 1, nontemporal=False, passes=5)
    print(json.dumps({"kernel": "k_probe_planes<uint32_t,false>", "bytes_read_per_dispatch": cells * 4,
                      "bytes_written_per_dispatch": cells * 4, "dispatches": 7, "ms_per_pass": round(ms, 4),
                      "gbs": round(2 * cells * 4 / (ms / 1e3) / 1e9, 1)}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--variants", default="all", help="all, precision (fp32 + fp16/bf16/e3m12 Esotwist/LIJK) or headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="main-loop steps of the CPU oracle on the headline tank (cpu_baseline)")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="tank: skip the accuracy object (16-bit storage against the fp64 build, 96^3 x 5000 steps)")
    ap.add_argument("--no-serving", action="store_true",
                    help="tank: skip the serving object (a short Poisson stream below capacity, drained)")
    ap.add_argument("--serving-duration", type=float, default=30.0,
                    help="tank: seconds of arrivals of the serving object")
    ap.add_argument("--serving-rate", type=float, default=1.5,
                    help="tank: arrivals per second of the serving object")
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
                    help="only run the access-pattern probe over one plane of GIB GiB (read and written back in "
                         "place, 7 passes): the known-size kernel for the HBM byte counters")
    ap.add_argument("--drain", type=float, default=600.0,
                    help="serve: seconds to keep ticking after the arrivals end (unfinished requests are censored)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs one ComputeUnit on one GPU; --gpus {args.gpus} is not supported "
              "(one GPU runs one tank or a batch of tanks; a slab decomposition over GPUs is not built)",
              file=sys.stderr)
        return 2
    if args.copy_calibration > 0:
        return probe_calibration(args.copy_calibration)

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
        r = run_serve_workload(td, args.duration, args.rate, args.tick_steps, args.max_resident, drain_s=args.drain)
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
    probe = sol_probe(args.grid)
    out["roofline"]["copy_gbs"] = probe["gbs"]
    out["roofline"]["frac_of_copy"] = round(achieved / probe["gbs"], 4)
    out["roofline"]["sol_probe"] = probe
    if not args.no_accuracy:
        from tdlbcpp_amd import precision as P
        names = {"fp16": "Esotwist/LIJK/fp16", "bf16": "Esotwist/LIJK/bf16", "e3m12": "Esotwist/LIJK/e3m12"}
        r = P.precision_report(T=96, steps=5000, precisions=(td.PREC_F32, td.PREC_F16, td.PREC_BF16, td.PREC_E3M12),
                               re=1.0e5, samples=20)
        acc = {"grid": 96, "steps": 5000, "reynolds": 1.0e5, "against": "fp64 build of the same kernels",
               "bar_velocity_l2": 1e-3, "formats": {}}
        for k, v in r["variants"].items():
            e = {"velocity_l2": v["velocity_l2"], "velocity_l2_mean_field": v["velocity_l2_mean"],
                 "po_rel_err": v["po_rel_err"], "mass_drift": v["mass_drift"],
                 "meets_bar": v["velocity_l2"] <= 1e-3, "meets_bar_mean_field": v["velocity_l2_mean"] <= 1e-3}
            if k in names and names[k] in results:
                e["store_speedup"] = round(results[names[k]]["mlups"] / head["mlups"], 3)
            acc["formats"][k] = e
        out["accuracy"] = acc
        print("[bench] accuracy 96^3 x 5000: " + ", ".join(
            f"{k} L2 {v['velocity_l2']:.2e} (mean field {v['velocity_l2_mean_field']:.2e})"
            for k, v in acc["formats"].items()), file=sys.stderr, flush=True)
    if not args.no_serving:
        r = run_serve_workload(td, args.serving_duration, args.serving_rate, args.tick_steps, 0, drain_s=120.0)
        out["serving"] = r
        print(f"[bench] serving: p50 {r['p50_latency_s']} s, p99 {r['p99_latency_s']} s over {r['submitted']} "
              f"requests ({r['unfinished']} unfinished), {r['aggregate_mlups']:.0f} MLUPS", file=sys.stderr,
              flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(grid=args.grid, steps=args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
