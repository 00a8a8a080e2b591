"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TDLB_ORACLE_LIB names another build of the same source (the mutation test of
# tests/test_oracle_mutants.py loads oracles with planted mistakes this way)
LIB_PATH = os.environ.get("TDLB_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

F32, F64, TAGS = 0, 1, 2
IJKL, LIJK = 0, 1
SIMPLE, ESOTWIST = 0, 1
FIELD_Q, FIELD_F, FIELD_NU, FIELD_O = 0, 1, 2, 3

# Same layout as tdlb_ib_point in include/tdlb.h
POINT_DTYPE = np.dtype([("i", "<i8"), ("j", "<i8"), ("k", "<i8"), ("i_frac", "<f8"), ("k_frac", "<f8"),
                        ("u_delta", "<f8"), ("v_delta", "<f8"), ("w_delta", "<f8")])

FLOW_KEYS = ("initialRho", "reMNonDimensional", "uav", "cs0", "g3", "nu", "fx0", "fy0", "fz0",
             "Re_m", "Re_f", "uf", "alpha", "beta")
FLOW_DEFAULTS = dict(initialRho=8.0, reMNonDimensional=7300.0, uav=0.1, cs0=0.12, g3=0.8, nu=0.0,
                     fx0=0.0, fy0=0.0, fz0=0.0, Re_m=0.0, Re_f=0.0, uf=0.0, alpha=0.97, beta=1.9)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def set_threads(n):
    """OpenMP threads for collision and moments (1, the default, is the reference's
    single-threaded path; results are the same for any count)."""
    lib().orc_set_threads(int(n))
    return lib().orc_get_threads()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i64, i64, i64, i64,
                                 ctypes.POINTER(d), ctypes.c_int]
        for name in ("orc_destroy", "orc_collision", "orc_moments", "orc_stream_push", "orc_stream_pull",
                     "orc_bounce_back"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.orc_initialise.argtypes = [vp, d]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_set_threads(1)  # the reference's path is single-threaded
        L.orc_forcing.argtypes = [vp, vp, i64, d, d]
        L.orc_field_bytes.argtypes = [vp, ctypes.c_int]
        L.orc_field_bytes.restype = i64
        L.orc_get.argtypes = [vp, ctypes.c_int, vp]
        L.orc_set.argtypes = [vp, ctypes.c_int, vp]
        L.orc_even_step.argtypes = [vp]
        L.orc_set_even_step.argtypes = [vp, ctypes.c_int]
        L.orc_dim.argtypes = [vp, ctypes.c_int]
        L.orc_dim.restype = i64
        L.orc_cells.argtypes = [vp, ctypes.c_int, vp]
        L.orc_calc_nu_f32.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
        L.orc_calc_nu_f32.restype = ctypes.c_float
        L.orc_calc_nu_f64.argtypes = [d, d, ctypes.c_int, ctypes.POINTER(d)]
        L.orc_calc_nu_f64.restype = d
        L.orc_weight_f32.argtypes = [ctypes.c_float]
        L.orc_weight_f32.restype = ctypes.c_float
        L.orc_weight_f64.argtypes = [d]
        L.orc_weight_f64.restype = d
        L.orc_moments_of_f64.argtypes = [vp, vp]
        L.orc_moments_of_f32.argtypes = [vp, vp]
        L.orc_double_resolution.argtypes = [vp]
        L.orc_set_flow.argtypes = [vp, ctypes.POINTER(d), ctypes.c_int]
        L.orc_plane_xz.argtypes = [vp, i64, vp, i64, i64, vp, ctypes.POINTER(i64), vp, ctypes.POINTER(i64)]
        L.orc_vorticity.argtypes = [vp, ctypes.c_int, i64, vp, vp]
        _lib = L
    return _lib


# savePlaneXZ records (Output.hpp:22-25: tDiskGrid<float, 4> and <float, 3>)
PLANE_Q_DTYPE = np.dtype({"names": ["i", "j", "k", "q"], "formats": ["<u2", "<u2", "<u2", ("<f4", 4)],
                          "offsets": [0, 2, 4, 8], "itemsize": 24})
PLANE_F_DTYPE = np.dtype({"names": ["i", "j", "k", "q"], "formats": ["<u2", "<u2", "<u2", ("<f4", 3)],
                          "offsets": [0, 2, 4, 8], "itemsize": 20})


class OracleCU:
    """One reference-semantics ComputeUnit on the CPU (float, double or uint64 tags)."""

    def __init__(self, x, y=None, z=None, ghost=1, dtype=F32, layout=LIJK, streaming=SIMPLE,
                 use_les=False, **flow):
        y = x if y is None else y
        z = x if z is None else z
        vals = dict(FLOW_DEFAULTS)
        vals.update(flow)
        arr = (ctypes.c_double * 14)(*[float(vals[k]) for k in FLOW_KEYS])
        self.dtype = dtype
        self.np_dtype = {F32: np.float32, F64: np.float64, TAGS: np.uint64}[dtype]
        self.layout = layout
        self._h = lib().orc_create(dtype, layout, streaming, x, y, z, ghost, arr, 1 if use_les else 0)
        self.xg, self.yg, self.zg, self.size = (lib().orc_dim(self._h, w) for w in range(4))
        self.ghost = ghost
        self.flow = vals
        self.use_les = use_les

    def set_flow(self, **flow):
        """Replace flow parameters (the ComputeUnit's `flow` member)."""
        self.flow.update(flow)
        arr = (ctypes.c_double * 14)(*[float(self.flow[k]) for k in FLOW_KEYS])
        lib().orc_set_flow(self._h, arr, 1 if self.use_les else 0)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    # reference surface
    def initialise(self, rho):
        lib().orc_initialise(self._h, rho)

    def collision(self):
        lib().orc_collision(self._h)

    def moments(self):
        lib().orc_moments(self._h)

    def streaming_push(self):
        lib().orc_stream_push(self._h)

    def streaming_pull(self):
        lib().orc_stream_pull(self._h)

    def bounce_back(self):
        lib().orc_bounce_back(self._h)

    def forcing(self, points, alfa, beta):
        pts = np.ascontiguousarray(points, dtype=POINT_DTYPE)
        lib().orc_forcing(self._h, pts.ctypes.data if len(pts) else None, len(pts), alfa, beta)

    def double_resolution(self):
        """doubleResolutionFullCU: the grid doubles in x, y, z."""
        lib().orc_double_resolution(self._h)
        self.xg, self.yg, self.zg, self.size = (lib().orc_dim(self._h, w) for w in range(4))

    def plane_xz(self, cut, exclude=None, i0=0, k0=0):
        """savePlaneXZ records: (QVec records, F3 records) as structured arrays."""
        n = (self.xg - 2) * (self.zg - 2)
        q = np.zeros(n, dtype=PLANE_Q_DTYPE)
        f = np.zeros(n, dtype=PLANE_F_DTYPE)
        nq, nf = ctypes.c_int64(), ctypes.c_int64()
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint8)
        lib().orc_plane_xz(self._h, cut, None if ex is None else ex.ctypes.data, i0, k0, q.ctypes.data,
                           ctypes.byref(nq), f.ctypes.data, ctypes.byref(nf))
        return q[:nq.value], f[:nf.value]

    def vorticity(self, axis, cut, exclude=None):
        """log |curl u|^2 on plane j = cut (axis 0, shape (x, z)) or k = cut
        (axis 1, shape (x, y)); NaN where excluded."""
        out = np.zeros((self.xg - 2, (self.zg if axis == 0 else self.yg) - 2), dtype=self.np_dtype)
        ex = None if exclude is None else np.ascontiguousarray(exclude, dtype=np.uint8)
        lib().orc_vorticity(self._h, axis, cut, None if ex is None else ex.ctypes.data, out.ctypes.data)
        return out

    def interior_shape(self):
        g = 2 * self.ghost
        return (self.xg - g, self.yg - g, self.zg - g, 18)

    def read_cells(self):
        """AccessField::read of every interior cell -> [x, y, z, 18]."""
        out = np.empty(self.interior_shape(), dtype=self.np_dtype)
        lib().orc_cells(self._h, 0, out.ctypes.data)
        return out

    def write_cells(self, vals, as_moments=True):
        """AccessField::writeMoments (or ::write) of every interior cell from [x, y, z, 18]."""
        a = np.ascontiguousarray(vals, dtype=self.np_dtype).reshape(self.interior_shape())
        lib().orc_cells(self._h, 1 if as_moments else 2, a.ctypes.data)

    @property
    def even_step(self):
        return bool(lib().orc_even_step(self._h))

    @even_step.setter
    def even_step(self, v):
        lib().orc_set_even_step(self._h, 1 if v else 0)

    # raw fields in the CU's in-memory layout
    def get(self, field):
        nbytes = lib().orc_field_bytes(self._h, field)
        dt = np.uint8 if field == FIELD_O else self.np_dtype
        out = np.empty(nbytes // np.dtype(dt).itemsize, dtype=dt)
        lib().orc_get(self._h, field, out.ctypes.data)
        return out

    def set(self, field, arr):
        dt = np.uint8 if field == FIELD_O else self.np_dtype
        a = np.ascontiguousarray(arr, dtype=dt).ravel()
        assert a.nbytes == lib().orc_field_bytes(self._h, field)
        lib().orc_set(self._h, field, a.ctypes.data)

    def q_ijkl(self):
        """Q as [xg, yg, zg, 18] regardless of memory layout."""
        q = self.get(FIELD_Q)
        if self.layout == IJKL:
            return q.reshape(self.xg, self.yg, self.zg, 18)
        return np.moveaxis(q.reshape(18, self.xg, self.yg, self.zg), 0, -1)

    def set_q_ijkl(self, q4):
        q4 = np.asarray(q4, dtype=self.np_dtype)
        if self.layout == IJKL:
            self.set(FIELD_Q, q4.reshape(-1))
        else:
            self.set(FIELD_Q, np.moveaxis(q4, -1, 0).reshape(-1))


def calc_nu(re_m_nondim, uav, outer_radius, dtype=F32):
    if dtype == F32:
        rm = ctypes.c_float()
        nu = lib().orc_calc_nu_f32(re_m_nondim, uav, int(outer_radius), ctypes.byref(rm))
        return float(nu), float(rm.value)
    rm = ctypes.c_double()
    nu = lib().orc_calc_nu_f64(re_m_nondim, uav, int(outer_radius), ctypes.byref(rm))
    return float(nu), float(rm.value)


def weight(x, dtype=F32):
    return float(lib().orc_weight_f32(x) if dtype == F32 else lib().orc_weight_f64(x))


def moments_of(q, dtype=F64):
    npdt = np.float64 if dtype == F64 else np.float32
    q = np.ascontiguousarray(q, dtype=npdt)
    m = np.empty(18, dtype=npdt)
    (lib().orc_moments_of_f64 if dtype == F64 else lib().orc_moments_of_f32)(q.ctypes.data, m.ctypes.data)
    return m


def fill_for_test(cu, offset=0):
    """Integer-tag fill of ParamsCommon::fillForTest (ParamsCommon.hpp:167-206): Q, F, Nu, O."""
    i, j, k = np.meshgrid(np.arange(cu.xg), np.arange(cu.yg), np.arange(cu.zg), indexing="ij")
    base = (((offset * 100 + i) * 100 + j) * 100 + k).astype(np.uint64) * np.uint64(100)
    q = base[..., None] + np.arange(18, dtype=np.uint64)
    f = base[..., None] + np.arange(3, dtype=np.uint64)
    return q, f, base + np.uint64(1)
