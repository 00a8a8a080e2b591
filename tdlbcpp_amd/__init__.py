"""Python binding of the MI355X ComputeUnit C ABI (include/tdlb.h, libtdlb_hip.so).

The host-side product code is C++ (the C ABI and the tdlb_run driver); this
module is the thin ctypes layer the tests and bench.py drive it through.
Method names follow the reference's ComputeUnitBase (tdlbcpp/src/ComputeUnit.h:97-201).
The library must be built (``make -C tdlbcpp_amd``); loading fails loudly if it
is missing, there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtdlb_hip.so")

LAYOUT_IJKL, LAYOUT_LIJK, LAYOUT_IJKL_SOA = 0, 1, 2  # IJKL_SOA: IJKL at the boundary, LIJK in HBM
STREAM_SIMPLE, STREAM_ESOTWIST = 0, 1
PREC_F32, PREC_F64, PREC_F16, PREC_BF16, PREC_E3M12 = 0, 1, 2, 3, 4
FIELD_Q, FIELD_F, FIELD_NU, FIELD_O = 0, 1, 2, 3
GEOM_FIXED, GEOM_ROTATING_NON_UPDATING, GEOM_ROTATING = 0, 1, 2

POINT_DTYPE = np.dtype([("i", "<i8"), ("j", "<i8"), ("k", "<i8"), ("i_frac", "<f8"), ("k_frac", "<f8"),
                        ("u_delta", "<f8"), ("v_delta", "<f8"), ("w_delta", "<f8")])


# savePlaneXZ records (tDiskGrid<float, 4> and <float, 3>)
PLANE_Q_DTYPE = np.dtype({"names": ["i", "j", "k", "q"], "formats": ["<u2", "<u2", "<u2", ("<f4", 4)],
                          "offsets": [0, 2, 4, 8], "itemsize": 24})
PLANE_F_DTYPE = np.dtype({"names": ["i", "j", "k", "q"], "formats": ["<u2", "<u2", "<u2", ("<f4", 3)],
                          "offsets": [0, 2, 4, 8], "itemsize": 20})


class CuParams(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int64), ("y", ctypes.c_int64), ("z", ctypes.c_int64), ("ghost", ctypes.c_int64),
                ("i0", ctypes.c_int64), ("j0", ctypes.c_int64), ("k0", ctypes.c_int64),
                ("resolution", ctypes.c_int64),
                ("idi", ctypes.c_int32), ("idj", ctypes.c_int32), ("idk", ctypes.c_int32),
                ("node_id", ctypes.c_int32), ("device_id", ctypes.c_int32),
                ("layout", ctypes.c_int32), ("streaming", ctypes.c_int32), ("precision", ctypes.c_int32)]


class Flow(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("initial_rho", "re_m_nondimensional", "uav", "cs0", "g3", "nu",
                                               "fx0", "fy0", "fz0", "re_m", "re_f", "uf", "alpha", "beta")] + \
               [("use_les", ctypes.c_int32), ("collision", ctypes.c_int32)]


class RushtonParams(ctypes.Structure):
    _fields_ = [("tank_diameter", ctypes.c_int64), ("height", ctypes.c_int64),
                ("center_x", ctypes.c_double), ("center_z", ctypes.c_double), ("tank_radius", ctypes.c_double),
                ("n_baffles", ctypes.c_int32), ("baffle_width", ctypes.c_double),
                ("baffle_angle0", ctypes.c_double), ("disk_plane", ctypes.c_int64),
                ("disk_radius", ctypes.c_double), ("hub_radius", ctypes.c_double),
                ("shaft_radius", ctypes.c_double), ("n_blades", ctypes.c_int32),
                ("blade_outer_radius", ctypes.c_double), ("blade_inner_radius", ctypes.c_double),
                ("blade_plane_lo", ctypes.c_int64), ("blade_planes", ctypes.c_int64),
                ("tip_speed", ctypes.c_double), ("startup_steps", ctypes.c_int64)]


# FlowParams<T> defaults (tdlbcpp/src/Params/FlowParams.hpp:27-61)
FLOW_DEFAULTS = dict(initial_rho=8.0, re_m_nondimensional=7300.0, uav=0.1, cs0=0.12, g3=0.8, nu=0.0,
                     fx0=0.0, fy0=0.0, fz0=0.0, re_m=0.0, re_f=0.0, uf=0.0, alpha=0.97, beta=1.9,
                     use_les=0, collision=0)


class TdlbError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libtdlb_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TdlbError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                            "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, u64, d, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_int
        P = ctypes.POINTER
        sig = {
            "tdlb_last_error": ([], ctypes.c_char_p),
            "tdlb_version": ([], i32),
            "tdlb_device_count": ([P(i32)], i32),
            "tdlb_flow_calc_nu_re_m": ([P(Flow), ctypes.c_int32, ctypes.c_int32], i32),
            "tdlb_cu_create": ([P(CuParams), P(Flow), P(vp)], i32),
            "tdlb_cu_destroy": ([vp], i32),
            "tdlb_cu_set_flow": ([vp, P(Flow)], i32),
            "tdlb_cu_initialise": ([vp, d], i32),
            "tdlb_cu_collision": ([vp], i32),
            "tdlb_cu_moments": ([vp], i32),
            "tdlb_cu_stream_push": ([vp], i32),
            "tdlb_cu_stream_pull": ([vp], i32),
            "tdlb_cu_bounce_back": ([vp], i32),
            "tdlb_cu_forcing": ([vp, vp, i64, d, d], i32),
            "tdlb_cu_field_bytes": ([vp, i32, P(ctypes.c_size_t)], i32),
            "tdlb_cu_download": ([vp, i32, vp, ctypes.c_size_t], i32),
            "tdlb_cu_upload": ([vp, i32, vp, ctypes.c_size_t], i32),
            "tdlb_cu_get_even_step": ([vp, P(i32)], i32),
            "tdlb_cu_set_even_step": ([vp, i32], i32),
            "tdlb_cu_dims": ([vp, P(i64), P(i64), P(i64), P(i64)], i32),
            "tdlb_cu_synchronize": ([vp], i32),
            "tdlb_cu_stream": ([vp, P(vp)], i32),
            "tdlb_rushton_defaults": ([i64, i64, d, i64, P(RushtonParams)], i32),
            "tdlb_rushton_generate": ([P(RushtonParams), i32, d, d, vp, i64, P(i64)], i32),
            "tdlb_rushton_step_increment": ([P(RushtonParams), u64, P(d)], i32),
            "tdlb_rushton_solid_mask": ([P(RushtonParams), i64, i64, i64, i64, vp], i32),
            "tdlb_cu_bind_rushton": ([vp, P(RushtonParams), u64, d, d, d], i32),
            "tdlb_cu_run": ([vp, i64, P(u64), P(d)], i32),
            "tdlb_cu_rotating_points": ([vp, vp, i64, P(i64)], i32),
            "tdlb_cu_solid_mask": ([vp, vp, ctypes.c_size_t], i32),
            "tdlb_cu_forcing_info": ([vp, P(i64), P(i64), P(i64)], i32),
            "tdlb_cu_one_launch": ([vp, P(i32)], i32),
            "tdlb_cu_one_launch_phases": ([vp, P(d), i32], i32),
            "tdlb_cu_profile": ([vp, i32], i32),
            "tdlb_cu_profile_read": ([vp, P(d), P(i64), P(d), P(i64)], i32),
            "tdlb_cu_reductions": ([vp, P(d)], i32),
            "tdlb_cu_reductions_noalloc": ([vp, vp], i32),
            "tdlb_probe_planes": ([ctypes.c_int32, ctypes.c_int32, i64, i64, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int32, P(d)], i32),
            "tdlb_cu_macroscopic": ([vp, vp, ctypes.c_size_t], i32),
            "tdlb_cu_double_resolution": ([vp], i32),
            "tdlb_cu_output_exclude": ([vp, vp, ctypes.c_size_t], i32),
            "tdlb_cu_plane_xz": ([vp, i64, vp, vp, P(i64), vp, P(i64)], i32),
            "tdlb_cu_vorticity": ([vp, ctypes.c_int32, i64, vp, vp, ctypes.c_size_t], i32),
            "tdlb_batch_create": ([i32, i32, i32, i32, P(vp)], i32),
            "tdlb_batch_destroy": ([vp], i32),
            "tdlb_batch_set_options": ([vp, i32, i32], i32),
            "tdlb_batch_add": ([vp, vp, u64, d, i64], i32),
            "tdlb_batch_remove": ([vp, vp], i32),
            "tdlb_batch_run": ([vp, i64, P(i64)], i32),
            "tdlb_batch_member": ([vp, vp, P(u64), P(d), P(i64)], i32),
            "tdlb_batch_info": ([vp, P(i64), P(i64), P(ctypes.c_int32), i64], i32),
            "tdlb_batch_tune": ([vp, vp], i32),
            "tdlb_batch_shape": ([vp, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32)], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols():
    """Names declared in include/tdlb.h that the library must export."""
    hdr = os.path.join(os.path.dirname(HERE), "include", "tdlb.h")
    import re
    names = re.findall(r"^(?:int|const char\*)\s+(tdlb_\w+)\s*\(", open(hdr).read(), flags=re.M)
    return names


def check(rc):
    if rc != 0:
        raise TdlbError(f"tdlb error {rc}: {lib().tdlb_last_error().decode()}")


def device_count():
    n = ctypes.c_int()
    check(lib().tdlb_device_count(ctypes.byref(n)))
    return n.value


def probe_planes(elem_bytes, cells, planes, nontemporal=True, passes=10, device=0, stride_bytes=0):
    """Milliseconds per pass of the access-pattern probe (tdlb_probe_planes): every value of `planes`
    planes of `cells` values, stride_bytes apart (0: rounded up to 1 MiB), read and written back in place;
    bytes per pass 2 * planes * cells * elem_bytes."""
    ms = ctypes.c_double()
    check(lib().tdlb_probe_planes(device, elem_bytes, cells, stride_bytes, planes, 1 if nontemporal else 0, passes,
                                  ctypes.byref(ms)))
    return ms.value


def lijk_plane_bytes(grid, elem_bytes):
    """Bytes between the population planes of a grid^3 LIJK lattice in HBM: the interior k-rows of every
    (i, j) row including the ghost rows, rounded to 128 B lines, then the two k-ghost cells of every row,
    the whole rounded up to 1 MiB when it reaches 16 MiB (else to a 128 B line)."""
    xg = yg = grid + 2
    line = 128 // elem_bytes
    side = (xg * yg * grid + line - 1) // line * line
    cells = side + 2 * xg * yg
    align = (1 << 20) // elem_bytes if cells * elem_bytes >= (16 << 20) else line
    return (cells + align - 1) // align * align * elem_bytes


def make_flow(**kw):
    vals = dict(FLOW_DEFAULTS)
    vals.update(kw)
    f = Flow()
    for k, v in vals.items():
        setattr(f, k, v)
    return f


def flow_for_tank(tank_diameter, re=7300.0, uav=0.1, cs0=0.1, use_les=1, precision=PREC_F32, **kw):
    """FlowParams as main.cpp:274-276 leaves them: nu and Re_m from calcNuAndRe_m
    with the integer impeller blade radius int(T/6)."""
    vals = dict(initial_rho=1.0, re_m_nondimensional=re, uav=uav, cs0=cs0, use_les=use_les)
    vals.update(kw)
    f = make_flow(**vals)
    check(lib().tdlb_flow_calc_nu_re_m(ctypes.byref(f), int(tank_diameter / 6), precision))
    return {n: getattr(f, n) for n, _ in Flow._fields_}


def rushton_defaults(tank_diameter, height=None, tip_speed=0.1, startup_steps=30):
    p = RushtonParams()
    check(lib().tdlb_rushton_defaults(tank_diameter, height or tank_diameter, tip_speed, startup_steps,
                                      ctypes.byref(p)))
    return p


def rushton_generate(p, group, angle=0.0, delta_angle=0.0):
    n = ctypes.c_int64()
    check(lib().tdlb_rushton_generate(ctypes.byref(p), group, angle, delta_angle, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, dtype=POINT_DTYPE)
    check(lib().tdlb_rushton_generate(ctypes.byref(p), group, angle, delta_angle,
                                      out.ctypes.data if n.value else None, n.value, ctypes.byref(n)))
    return out


def rushton_step_increment(p, step):
    d = ctypes.c_double()
    check(lib().tdlb_rushton_step_increment(ctypes.byref(p), step, ctypes.byref(d)))
    return d.value


def rushton_solid_mask(p, xg, yg, zg, ghost=1):
    out = np.empty(xg * yg * zg, dtype=np.uint8)
    check(lib().tdlb_rushton_solid_mask(ctypes.byref(p), xg, yg, zg, ghost, out.ctypes.data))
    return out


class ComputeUnit:
    """One ComputeUnit in HBM: the reference's ComputeUnitBase<T,18,Layout> surface."""

    def __init__(self, x, y=None, z=None, ghost=1, layout=LAYOUT_LIJK, streaming=STREAM_ESOTWIST,
                 precision=PREC_F32, device_id=0, **flow):
        p = CuParams()
        p.x, p.y, p.z = x, (x if y is None else y), (x if z is None else z)
        p.ghost, p.layout, p.streaming, p.precision, p.device_id = ghost, layout, streaming, precision, device_id
        p.resolution = 1
        self.params = p
        self.flow = make_flow(**flow)
        self._h = ctypes.c_void_p()
        check(lib().tdlb_cu_create(ctypes.byref(p), ctypes.byref(self.flow), ctypes.byref(self._h)))
        xg, yg, zg, size = (ctypes.c_int64() for _ in range(4))
        check(lib().tdlb_cu_dims(self._h, ctypes.byref(xg), ctypes.byref(yg), ctypes.byref(zg), ctypes.byref(size)))
        self.xg, self.yg, self.zg, self.size = xg.value, yg.value, zg.value, size.value
        self.layout = layout
        self.precision = precision
        # every field is float except with fp64 (the 16-bit modes store only the lattice in 16 bits)
        self.np_dtype = np.float64 if precision == PREC_F64 else np.float32
        self.step = 0
        self.angle = 0.0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(lib().tdlb_cu_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- reference operator surface
    def set_flow(self, **kw):
        vals = {n: getattr(self.flow, n) for n, _ in Flow._fields_}
        vals.update(kw)
        self.flow = make_flow(**vals)
        check(lib().tdlb_cu_set_flow(self._h, ctypes.byref(self.flow)))

    def initialise(self, rho):
        check(lib().tdlb_cu_initialise(self._h, rho))

    def collision(self):
        check(lib().tdlb_cu_collision(self._h))

    def moments(self):
        check(lib().tdlb_cu_moments(self._h))

    def streaming_push(self):
        check(lib().tdlb_cu_stream_push(self._h))

    def streaming_pull(self):
        check(lib().tdlb_cu_stream_pull(self._h))

    def bounce_back(self):
        check(lib().tdlb_cu_bounce_back(self._h))

    def forcing(self, points, alfa, beta):
        pts = np.ascontiguousarray(points, dtype=POINT_DTYPE)
        check(lib().tdlb_cu_forcing(self._h, pts.ctypes.data if len(pts) else None, len(pts), alfa, beta))

    @property
    def even_step(self):
        e = ctypes.c_int()
        check(lib().tdlb_cu_get_even_step(self._h, ctypes.byref(e)))
        return bool(e.value)

    @even_step.setter
    def even_step(self, v):
        check(lib().tdlb_cu_set_even_step(self._h, 1 if v else 0))

    def synchronize(self):
        check(lib().tdlb_cu_synchronize(self._h))

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().tdlb_cu_stream(self._h, ctypes.byref(s)))
        return s.value

    # ---- fields in the reference's in-memory layout
    def field_bytes(self, field):
        b = ctypes.c_size_t()
        check(lib().tdlb_cu_field_bytes(self._h, field, ctypes.byref(b)))
        return b.value

    def get(self, field, dtype=None):
        dt = np.uint8 if field == FIELD_O else (dtype or self.np_dtype)
        nb = self.field_bytes(field)
        out = np.empty(nb // np.dtype(dt).itemsize, dtype=dt)
        check(lib().tdlb_cu_download(self._h, field, out.ctypes.data, nb))
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(arr)
        if field == FIELD_O:
            a = np.ascontiguousarray(a, dtype=np.uint8)
        nb = self.field_bytes(field)
        if a.nbytes != nb:
            raise TdlbError(f"upload needs {nb} bytes, got {a.nbytes}")
        check(lib().tdlb_cu_upload(self._h, field, a.ctypes.data, nb))

    def q_ijkl(self, dtype=None):
        q = self.get(FIELD_Q, dtype)
        if self.layout in (LAYOUT_IJKL, LAYOUT_IJKL_SOA):
            return q.reshape(self.xg, self.yg, self.zg, 18)
        return np.moveaxis(q.reshape(18, self.xg, self.yg, self.zg), 0, -1)

    def set_q_ijkl(self, q4):
        q4 = np.asarray(q4)
        if self.layout in (LAYOUT_IJKL, LAYOUT_IJKL_SOA):
            self.set(FIELD_Q, np.ascontiguousarray(q4).reshape(-1))
        else:
            self.set(FIELD_Q, np.ascontiguousarray(np.moveaxis(q4, -1, 0)).reshape(-1))

    # ---- fused main loop with the bound Rushton tank
    def bind_rushton(self, rp, step=1, angle=0.0, alfa=None, beta=None):
        self.step, self.angle = step, angle
        check(lib().tdlb_cu_bind_rushton(self._h, ctypes.byref(rp), step, angle,
                                         self.flow.alpha if alfa is None else alfa,
                                         self.flow.beta if beta is None else beta))

    def run(self, nsteps):
        s = ctypes.c_uint64(self.step)
        a = ctypes.c_double(self.angle)
        check(lib().tdlb_cu_run(self._h, nsteps, ctypes.byref(s), ctypes.byref(a)))
        self.step, self.angle = s.value, a.value

    def rotating_points(self):
        n = ctypes.c_int64()
        check(lib().tdlb_cu_rotating_points(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=POINT_DTYPE)
        check(lib().tdlb_cu_rotating_points(self._h, out.ctypes.data, n.value, ctypes.byref(n)))
        return out

    def solid_mask(self):
        out = np.empty(self.size, dtype=np.uint8)
        check(lib().tdlb_cu_solid_mask(self._h, out.ctypes.data, out.nbytes))
        return out

    def profile(self, enable=True):
        check(lib().tdlb_cu_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        """(collide_ms, collide_launches, forcing_ms, forcing_passes) since profile(True)."""
        a, c = ctypes.c_double(), ctypes.c_double()
        b, d = ctypes.c_int64(), ctypes.c_int64()
        check(lib().tdlb_cu_profile_read(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                         ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    @property
    def one_launch(self):
        """True when run() advances this bound tank with the one-launch step."""
        v = ctypes.c_int32()
        check(lib().tdlb_cu_one_launch(self._h, ctypes.byref(v)))
        return bool(v.value)

    def one_launch_phases(self):
        """Per XCD (rows written, forcing gathers done, forcing levels done), in
        microseconds from the first work item of the last one-launch step."""
        out = (ctypes.c_double * 24)()
        check(lib().tdlb_cu_one_launch_phases(self._h, out, 24))
        return [tuple(out[3 * x:3 * x + 3]) for x in range(8)]

    def forcing_info(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().tdlb_cu_forcing_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def double_resolution(self):
        """doubleResolutionFullCU: the grid doubles; bind the tank of the new size before run()."""
        check(lib().tdlb_cu_double_resolution(self._h))
        xg, yg, zg, size = (ctypes.c_int64() for _ in range(4))
        check(lib().tdlb_cu_dims(self._h, ctypes.byref(xg), ctypes.byref(yg), ctypes.byref(zg), ctypes.byref(size)))
        self.xg, self.yg, self.zg, self.size = xg.value, yg.value, zg.value, size.value

    def output_exclude(self):
        out = np.empty(self.size, dtype=np.uint8)
        check(lib().tdlb_cu_output_exclude(self._h, out.ctypes.data, out.nbytes))
        return out

    def plane_xz(self, cut, exclude=None):
        """savePlaneXZ records: (QVec records, F3 records) as structured arrays."""
        n = (self.xg - 2) * (self.zg - 2)
        q = np.zeros(n, dtype=PLANE_Q_DTYPE)
        f = np.zeros(n, dtype=PLANE_F_DTYPE)
        nq, nf = ctypes.c_int64(), ctypes.c_int64()
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint8)
        check(lib().tdlb_cu_plane_xz(self._h, cut, None if ex is None else ex.ctypes.data, q.ctypes.data,
                                     ctypes.byref(nq), f.ctypes.data, ctypes.byref(nf)))
        return q[:nq.value], f[:nf.value]

    def vorticity(self, axis, cut, exclude=None):
        """log |curl u|^2 on plane j = cut (axis 0, (x, z)) or k = cut (axis 1, (x, y))."""
        out = np.empty((self.xg - 2, (self.zg if axis == 0 else self.yg) - 2), dtype=self.np_dtype)
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint8)
        check(lib().tdlb_cu_vorticity(self._h, axis, cut, None if ex is None else ex.ctypes.data, out.ctypes.data,
                                      out.nbytes))
        return out

    def reductions(self):
        """FP64 device sums: {"mass", "momentum" (3), "force" (3), "torque"}."""
        out = (ctypes.c_double * 8)()
        check(lib().tdlb_cu_reductions(self._h, out))
        v = list(out)
        return {"mass": v[0], "momentum": tuple(v[1:4]), "force": tuple(v[4:7]), "torque": v[7]}

    def macroscopic(self):
        """(x, y, z, 4) array of {rho, ux, uy, uz} over the interior cells."""
        nx, ny, nz = self.xg - 2, self.yg - 2, self.zg - 2
        out = np.empty((nx, ny, nz, 4), dtype=self.np_dtype)
        check(lib().tdlb_cu_macroscopic(self._h, out.ctypes.data, out.nbytes))
        return out


class Batch:
    """Many bound ComputeUnits advanced together (tdlb_batch_*): one launch per
    kernel of the main-loop step for all members, the step replayed from a HIP graph."""

    def __init__(self, precision=PREC_F16, layout=LAYOUT_LIJK, streaming=STREAM_ESOTWIST, device_id=0,
                 use_graph=True, autotune=True):
        self._b = ctypes.c_void_p()
        check(lib().tdlb_batch_create(precision, layout, streaming, device_id, ctypes.byref(self._b)))
        check(lib().tdlb_batch_set_options(self._b, 1 if use_graph else 0, 1 if autotune else 0))
        self.members = []

    def add(self, cu, nsteps):
        check(lib().tdlb_batch_add(self._b, cu._h, cu.step, cu.angle, nsteps))
        self.members.append(cu)

    def remove(self, cu):
        check(lib().tdlb_batch_remove(self._b, cu._h))
        self.members.remove(cu)

    def member(self, cu):
        """(step, angle, remaining) of a member."""
        s, a, r = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_int64()
        check(lib().tdlb_batch_member(self._b, cu._h, ctypes.byref(s), ctypes.byref(a), ctypes.byref(r)))
        return s.value, a.value, r.value

    def run(self, nsteps):
        n = ctypes.c_int64()
        check(lib().tdlb_batch_run(self._b, nsteps, ctypes.byref(n)))
        for cu in self.members:
            cu.step, cu.angle, _ = self.member(cu)
        return n.value

    def tune(self, cu):
        """Autotune cu's grid shape on cu's own arrays (before cu.initialise)."""
        check(lib().tdlb_batch_tune(self._b, cu._h))

    def info(self):
        n, l = ctypes.c_int64(), ctypes.c_int64()
        tiles = (ctypes.c_int32 * max(1, len(self.members)))()
        check(lib().tdlb_batch_info(self._b, ctypes.byref(n), ctypes.byref(l), tiles, len(self.members)))
        bs, gs, fb = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().tdlb_batch_shape(self._b, ctypes.byref(bs), ctypes.byref(gs), ctypes.byref(fb)))
        return {"members": n.value, "launches_per_step": l.value, "tiles": list(tiles)[:len(self.members)],
                "block_size": bs.value, "graph_steps": gs.value, "forcing_block": fb.value}

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            check(lib().tdlb_batch_destroy(self._b))
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TankRequest(ctypes.Structure):
    _fields_ = [("edge", ctypes.c_int64), ("tip_speed", ctypes.c_double), ("reynolds", ctypes.c_double),
                ("steps", ctypes.c_int64), ("startup_steps", ctypes.c_int64)]


class TankResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_int64), ("edge", ctypes.c_int64), ("steps", ctypes.c_int64),
                ("error", ctypes.c_int32)] + \
               [(n, ctypes.c_double) for n in ("submit_s", "admit_s", "start_s", "finish_s", "mass", "mass_drift",
                                               "torque", "power_number")]


def _server_sigs(L):
    vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "tdlb_server_create": ([i32, i32, i32, i32, d, i64, i64, i32, P(vp)], ctypes.c_int),
        "tdlb_server_destroy": ([vp], ctypes.c_int),
        "tdlb_server_submit": ([vp, P(TankRequest), P(i64)], ctypes.c_int),
        "tdlb_server_tick": ([vp, P(i64), P(i64), P(i64)], ctypes.c_int),
        "tdlb_server_poll": ([vp, P(TankResult), i64, P(i64)], ctypes.c_int),
        "tdlb_server_take": ([vp, i64, P(vp)], ctypes.c_int),
        "tdlb_server_stats": ([vp, P(d), P(d), P(d)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


class TankHandle(ComputeUnit):
    """A ComputeUnit handle the server created (tdlb_server_take); same methods."""

    def __init__(self, h, edge, precision):
        self._h = h
        xg, yg, zg, size = (ctypes.c_int64() for _ in range(4))
        check(lib().tdlb_cu_dims(self._h, ctypes.byref(xg), ctypes.byref(yg), ctypes.byref(zg), ctypes.byref(size)))
        self.xg, self.yg, self.zg, self.size = xg.value, yg.value, zg.value, size.value
        self.precision = precision
        self.np_dtype = np.float64 if precision == PREC_F64 else np.float32
        self.step, self.angle = 0, 0.0


class Server:
    """Serving loop over tdlb_server_*: submit requests, tick, poll results."""

    def __init__(self, precision=PREC_F16, layout=LAYOUT_LIJK, streaming=STREAM_ESOTWIST, device_id=0,
                 hbm_budget=0.0, tick_steps=50, max_resident=0, keep_handles=False):
        _server_sigs(lib())
        self._s = ctypes.c_void_p()
        self.precision = precision
        check(lib().tdlb_server_create(precision, layout, streaming, device_id, float(hbm_budget), tick_steps,
                                       max_resident, 1 if keep_handles else 0, ctypes.byref(self._s)))

    def submit(self, edge, tip_speed, reynolds, steps, startup_steps=100):
        r = TankRequest(edge, tip_speed, reynolds, steps, startup_steps)
        t = ctypes.c_int64()
        check(lib().tdlb_server_submit(self._s, ctypes.byref(r), ctypes.byref(t)))
        return t.value

    def tick(self):
        """(resident, queued, steps_run) after one tick."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().tdlb_server_tick(self._s, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def poll(self, cap=256):
        buf = (TankResult * cap)()
        n = ctypes.c_int64()
        check(lib().tdlb_server_poll(self._s, buf, cap, ctypes.byref(n)))
        return [{f: getattr(buf[a], f) for f, _ in TankResult._fields_} for a in range(n.value)]

    def take(self, ticket, edge):
        h = ctypes.c_void_p()
        check(lib().tdlb_server_take(self._s, ticket, ctypes.byref(h)))
        return TankHandle(h, edge, self.precision)

    def stats(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(lib().tdlb_server_stats(self._s, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"now_s": a.value, "reserved_bytes": b.value, "budget_bytes": c.value}

    def close(self):
        if getattr(self, "_s", None) and self._s.value:
            check(lib().tdlb_server_destroy(self._s))
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
