"""Accuracy of the 16-bit population storage against the fp64 build of the same
kernels: one Rushton tank run per precision, then the relative velocity L2 over
the fluid cells and the power number from the FP64 device torque sum.

Power number: P = torque * dtheta (work per step of the impeller force on the
fluid, torque from tdlb_cu_reductions), N = dtheta / (2 pi) revolutions per
step, Po = P / (rho0 * N^3 * D^5) with D = 2 * blade_outer_radius; averaged
over the samples taken in the last quarter of the run.
"""
import math
import time

import numpy as np

from . import (LAYOUT_LIJK, PREC_F64, STREAM_ESOTWIST, ComputeUnit, flow_for_tank, rushton_defaults,
               rushton_step_increment)


def tank_run(T, precision, steps, streaming=STREAM_ESOTWIST, layout=LAYOUT_LIJK, re=1e5, tip_speed=0.1,
             startup_steps=200, samples=20, flow=None):
    """Run one tank; returns {"mac": (x,y,z,4) float64 at the last step, "mac_mean": its mean over the
    sampling window (the last quarter of the run), "po": mean power number, "mass": [...], ...}."""
    if flow is None:
        flow = flow_for_tank(T, re=re, uav=tip_speed, precision=PREC_F64)
    rp = rushton_defaults(T, tip_speed=tip_speed, startup_steps=startup_steps)
    cu = ComputeUnit(T, layout=layout, streaming=streaming, precision=precision, **flow)
    try:
        cu.initialise(1.0)
        cu.bind_rushton(rp)
        m0 = cu.reductions()["mass"]
        sample_from = steps - steps // 4
        every = max(1, (steps - sample_from) // max(1, samples))
        pos, masses = [], []
        mean, nmean = None, 0  # macroscopic fields summed over the sampling window
        D = 2.0 * rp.blade_outer_radius
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = every if done >= sample_from else min(sample_from - done, steps - done)
            n = min(n, steps - done)
            cu.run(n)
            done += n
            if done > sample_from or done == steps:
                red = cu.reductions()
                dth = rushton_step_increment(rp, cu.step)
                N = dth / (2.0 * math.pi)
                pos.append(red["torque"] * dth / (1.0 * N ** 3 * D ** 5))
                masses.append(red["mass"])
                m = cu.macroscopic().astype(np.float64)
                mean = m if mean is None else mean + m
                nmean += 1
        cu.synchronize()
        secs = time.perf_counter() - t0
        mac = cu.macroscopic().astype(np.float64)
        solid = cu.solid_mask().reshape(cu.xg, cu.yg, cu.zg)[1:-1, 1:-1, 1:-1] != 0
        return {"mac": mac, "mac_mean": mean / nmean, "fluid": ~solid, "po": float(np.mean(pos)),
                "po_samples": pos, "mass0": m0, "mass": masses, "seconds": secs}
    finally:
        cu.close()


def velocity_l2(a, b, fluid):
    """||u_a - u_b||_2 / ||u_b||_2 over the fluid cells."""
    ua, ub = a[..., 1:4][fluid], b[..., 1:4][fluid]
    return float(np.linalg.norm(ua - ub) / np.linalg.norm(ub))


def precision_report(T=96, steps=5000, precisions=(0, 2, 3), **kw):
    """Each precision against fp64: velocity L2 at the last step and of the velocity averaged over the
    last quarter of the run, power-number error, mass drift."""
    ref = tank_run(T, PREC_F64, steps, **kw)
    names = {0: "fp32", 1: "fp64", 2: "fp16", 3: "bf16", 4: "e3m12"}
    out = {"grid": T, "steps": steps, "po_fp64": ref["po"], "mass_drift_fp64": ref["mass"][-1] / ref["mass0"] - 1.0,
           "variants": {}}
    for p in precisions:
        r = tank_run(T, p, steps, **kw)
        out["variants"][names[p]] = {
            "velocity_l2": velocity_l2(r["mac"], ref["mac"], ref["fluid"]),
            "velocity_l2_mean": velocity_l2(r["mac_mean"], ref["mac_mean"], ref["fluid"]),
            "po": r["po"],
            "po_rel_err": abs(r["po"] - ref["po"]) / abs(ref["po"]) if ref["po"] else None,
            "mass_drift": r["mass"][-1] / r["mass0"] - 1.0,
            "seconds": r["seconds"],
        }
    return out
