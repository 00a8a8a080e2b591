"""Shared helpers of the parity tests: the reference main-loop sequence on the
oracle, seeded states and the tolerance of the parity bar."""
import numpy as np

import tdlbcpp_amd as td
from oracle import oracle as O


def f32(v):
    return float(np.float32(v))


def flow_for_tank(T, re=7300.0, uav=0.1, les=True, cs0=0.1, dtype=O.F32):
    """FlowParams after calcNuAndRe_m with the integer blade radius int(T/6) (main.cpp:274-276)."""
    nu, rem = O.calc_nu(re, uav, int(T / 6), dtype)
    return dict(initialRho=1.0, reMNonDimensional=re, uav=uav, cs0=cs0, g3=0.8, nu=nu, Re_m=rem,
                alpha=0.97, beta=1.9), dict(initial_rho=1.0, re_m_nondimensional=re, uav=uav, cs0=cs0, g3=0.8,
                                            nu=nu, re_m=rem, alpha=0.97, beta=1.9, use_les=1 if les else 0)


class ReferenceLoop:
    """main.cpp:270-455 on an oracle ComputeUnit with this library's host generator lists."""

    def __init__(self, ocu, rp, step=1, angle=0.0, alfa=0.97, beta=1.9, precision=O.F32):
        self.cu, self.rp, self.step, self.angle = ocu, rp, step, angle
        self.alfa, self.beta = alfa, beta
        rnd = f32 if precision == O.F32 else float
        self.rnd = rnd
        delta = rnd(td.rushton_step_increment(rp, step))
        self.fixed = td.rushton_generate(rp, td.GEOM_FIXED)
        self.nonupd = td.rushton_generate(rp, td.GEOM_ROTATING_NON_UPDATING, 0.0, delta)
        rot = td.rushton_generate(rp, td.GEOM_ROTATING, angle, delta)
        self.cu.forcing(np.concatenate([self.fixed, self.nonupd, rot]), alfa, beta)

    def run(self, n):
        for _ in range(n):
            self.step += 1
            delta = self.rnd(td.rushton_step_increment(self.rp, self.step))
            self.angle += delta
            rot = td.rushton_generate(self.rp, td.GEOM_ROTATING, self.angle, delta)
            self.cu.collision()
            self.cu.bounce_back()
            self.cu.streaming_push()
            self.cu.moments()
            self.rot = rot  # the rotating list of the last step (output exclusion)
            self.cu.forcing(np.concatenate([self.fixed, self.nonupd, rot]), self.alfa, self.beta)


def assert_close_moments(got, ref, rho0=1.0, rtol=1e-5, what="Q"):
    """Parity bar: |a - b| <= rtol * max(|b|, rho0/24) elementwise (near-zero higher
    moments make a pure relative error meaningless)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.maximum(np.abs(ref), rho0 / 24.0)
    err = np.abs(got - ref) / scale
    worst = np.unravel_index(np.argmax(err), err.shape) if err.size else ()
    assert err.max(initial=0.0) <= rtol, f"{what}: max scaled error {err.max():.3e} at {worst}"
    return float(err.max(initial=0.0))


def random_moments(shape, rng, rho0=1.0, amp=0.01):
    m = np.zeros(shape + (18,))
    m[..., 0] = rho0 * (1.0 + amp * rng.standard_normal(shape))
    m[..., 1:4] = amp * 0.1 * rng.standard_normal(shape + (3,))
    m[..., 4:10] = amp * 0.01 * rng.standard_normal(shape + (6,))
    m[..., 10:16] = amp * 0.001 * rng.standard_normal(shape + (6,))
    m[..., 16:] = amp * 0.001 * rng.standard_normal(shape + (2,))
    return m
