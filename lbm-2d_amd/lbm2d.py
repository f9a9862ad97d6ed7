"""ctypes binding of liblbm2d.so (include/lbm_gpu.h) for tests and benchmarks.

The library is the product: every call here goes straight to the C ABI, whose
kernels run on the GPU.  There is no CPU fallback -- loading fails loudly when
the library has not been built (``make -C lbm-2d_amd``).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LBM2D_LIB selects another build of the same library (tuning experiments).
LIB_PATH = os.environ.get("LBM2D_LIB") or os.path.join(HERE, "liblbm2d.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "lbm_gpu.h")

STORAGE = {"f64": 0, "f32": 1, "f16s": 2, "bf16s": 3}
BYTES_PER_POP = {"f64": 8, "f32": 4, "f16s": 2, "bf16s": 2}
MRT, BGK = 0, 1
EDGE_PERIODIC, EDGE_OPEN, EDGE_WALL, EDGE_FREE_SLIP = 0, 1, 2, 3
INLET_NONE, INLET_ZOU_HE, INLET_PARABOLIC, INLET_EQUILIBRIUM = 0, 1, 2, 3
BB_SIMPLE, BB_MEI, BB_BOUZIDI = 0, 1, 2

# CaseType (src/lbm_types.hpp:20-23)
CYLINDER, CAVITY, STEP, RIBS, URBAN_CANYON, DOWNWASH, SQUARE_CYLINDER, FLAT_PLATE, \
    SPORTS_BALL, PERIODIC_HILLS, CYLINDER_NEAR_WALL, SIDE_BY_SIDE, ROTATING_CYLINDER, \
    ORIFICE_PLATE = range(14)


class LbmDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("collision", C.c_int32), ("les", C.c_int32), ("cs", C.c_double),
        ("x_edge", C.c_int32), ("y_edge", C.c_int32),
        ("inlet", C.c_int32), ("inlet_y0", C.c_int32), ("inlet_h", C.c_int32),
        ("outlet", C.c_int32), ("lid", C.c_int32), ("forces", C.c_int32),
        ("body_force_x", C.c_double), ("tau", C.c_double), ("u_in", C.c_double),
    ]


class LbmLink(C.Structure):
    _fields_ = [
        ("cell", C.c_int32), ("dir", C.c_int32), ("scheme", C.c_int32), ("slot", C.c_int32),
        ("q", C.c_double), ("uwx", C.c_double), ("uwy", C.c_double),
    ]


class LbmInfo(C.Structure):
    _fields_ = [
        ("cells", C.c_int64), ("fluid_cells", C.c_int64), ("link_cells", C.c_int64),
        ("les_cells", C.c_int64), ("bytes_per_pop", C.c_int32), ("n_lattices", C.c_int32),
        ("device_bytes", C.c_int64), ("block_x", C.c_int32), ("rows", C.c_int32),
        ("nodes_per_thread", C.c_int32), ("launch_options", C.c_int32),
        ("bulk_tiles", C.c_int32), ("other_tiles", C.c_int32), ("edge_tiles", C.c_int32),
        ("empty_tiles", C.c_int32),
    ]


class LbmThermalRule(C.Structure):
    _fields_ = [("t", C.c_double), ("kind", C.c_int32), ("pad", C.c_int32)]


THERMAL_EQ_FLOW, THERMAL_EQ_REST = 0, 1
THERMAL_MIRROR = 16


class LbmError(RuntimeError):
    pass


_lib = None
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_i32p = C.POINTER(C.c_int32)
_h = C.c_void_p


def declared_functions() -> list:
    """Names of the functions declared in include/lbm_gpu.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(lbm_\w+)\s*\(", text, re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LbmError(f"{LIB_PATH} is missing: build it with `make -C lbm-2d_amd` "
                       "(there is no fallback path)")
    L = C.CDLL(LIB_PATH)
    i32, i64, dbl = C.c_int32, C.c_int64, C.c_double
    sig = {
        "lbm_last_error": (C.c_char_p, []),
        "lbm_desc_for_case": (C.c_int, [C.c_int, i32, i32, C.POINTER(LbmDesc)]),
        "lbm_create": (C.c_int, [C.POINTER(LbmDesc), i32, i32, i32, C.c_void_p, C.POINTER(_h)]),
        "lbm_destroy": (C.c_int, [_h]),
        "lbm_population_layout": (C.c_int, [C.POINTER(LbmDesc), i32, C.POINTER(i64), C.POINTER(i64),
                                            C.POINTER(i64)]),
        "lbm_set_obstacle": (C.c_int, [_h, i32, _u8p]),
        "lbm_set_bb_circle": (C.c_int, [_h, i32, dbl, dbl, dbl, i32]),
        "lbm_set_bb_polygon": (C.c_int, [_h, i32, _dp, i32, i32]),
        "lbm_set_moving_wall": (C.c_int, [_h, i32, dbl, dbl, dbl]),
        "lbm_set_force_radius": (C.c_int, [_h, i32, dbl, dbl, dbl]),
        "lbm_set_links": (C.c_int, [_h, i32, C.POINTER(LbmLink), i32]),
        "lbm_set_force_slots": (C.c_int, [_h, i32, i32]),
        "lbm_set_params": (C.c_int, [_h, i32, dbl, dbl]),
        "lbm_upload": (C.c_int, [_h, i32, _dp]),
        "lbm_download": (C.c_int, [_h, i32, _dp]),
        "lbm_init_equilibrium": (C.c_int, [_h, i32, dbl, dbl, dbl]),
        "lbm_set_storage_shift": (C.c_int, [_h, i32, dbl, dbl]),
        "lbm_set_stochastic_rounding": (C.c_int, [_h, i32, C.c_uint32]),
        "lbm_macros": (C.c_int, [_h, i32, i32, i32, _dp, _dp, _dp]),
        "lbm_step": (C.c_int, [_h, i32]),
        "lbm_step_lattice": (C.c_int, [_h, i32, i32]),
        "lbm_debug_poison": (C.c_int, [_h, i32]),
        "lbm_step_noalloc": (C.c_int, [_h, i32]),
        "lbm_set_launch": (C.c_int, [_h, i32, i32, i32, i32, i32]),
        "lbm_autotune": (C.c_int, [_h, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                   C.POINTER(i32), C.POINTER(C.c_float)]),
        "lbm_forces": (C.c_int, [_h, i32, _dp, _dp, i32, C.POINTER(i32)]),
        "lbm_mass": (C.c_int, [_h, i32, _dp]),
        "lbm_nan_flag": (C.c_int, [_h, i32, C.POINTER(i32)]),
        "lbm_steps_done": (C.c_int, [_h, i32, C.POINTER(i64)]),
        "lbm_get_info": (C.c_int, [_h, C.POINTER(LbmInfo)]),
        "lbm_sync": (C.c_int, [_h]),
        "lbm_wall_distance": (C.c_int, [_h, i32, _dp]),
        "lbm_perturb_velocity": (C.c_int, [_h, i32, i32, i32, _dp]),
        "lbm_place_obstacles": (C.c_int, [_h, i32, _dp, i32, _dp, _i32p, i32]),
        "lbm_get_obstacle": (C.c_int, [_h, i32, _u8p]),
        "lbm_set_thermal": (C.c_int, [_h, i32, dbl, dbl, dbl, dbl]),
        "lbm_set_thermal_rules": (C.c_int, [_h, i32, _u8p, C.POINTER(LbmThermalRule), i32]),
        "lbm_upload_thermal": (C.c_int, [_h, i32, _dp]),
        "lbm_download_thermal": (C.c_int, [_h, i32, _dp]),
        "lbm_temperature": (C.c_int, [_h, i32, i32, i32, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().lbm_last_error().decode(errors="replace")
        raise LbmError(f"status {rc}: {msg}")


def desc_for_case(case: int, nx: int, ny: int, *, tau: float = 1.0, u_in: float = 0.1,
                  les: bool = False, cs: float = 0.12, collision: int = MRT,
                  body_force_x: float = 0.0) -> LbmDesc:
    d = LbmDesc()
    _check(lib().lbm_desc_for_case(case, nx, ny, C.byref(d)))
    d.tau, d.u_in, d.les, d.cs = tau, u_in, int(les), cs
    d.collision, d.body_force_x = collision, body_force_x
    return d


def population_layout(descs: Sequence[LbmDesc]):
    """(f_off per lattice, pool values, guard values) of the population buffers a handle
    of these lattices holds; no device needed."""
    arr = (LbmDesc * len(descs))(*descs)
    off = (C.c_int64 * len(descs))()
    pool, guard = C.c_int64(), C.c_int64()
    _check(lib().lbm_population_layout(arr, len(descs), off, C.byref(pool), C.byref(guard)))
    return list(off), pool.value, guard.value


class Ensemble:
    """A handle holding one or more lattices that step together on the GPU."""

    def __init__(self, descs: Sequence[LbmDesc], storage: str = "f32", scale_exp: int = 8,
                 stream: Optional[int] = None):
        self.storage = storage
        self.descs = list(descs)
        arr = (LbmDesc * len(self.descs))(*self.descs)
        h = _h()
        _check(lib().lbm_create(arr, len(self.descs), STORAGE[storage], scale_exp,
                                C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def shape(self, l=0):
        d = self.descs[l]
        return d.ny, d.nx

    # ---- geometry
    def set_obstacle(self, l: int, mask: np.ndarray):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        assert m.shape == self.shape(l), (m.shape, self.shape(l))
        _check(lib().lbm_set_obstacle(self._h, l, m.ctypes.data_as(_u8p)))

    def place_obstacles(self, l: int, circles=(), polygons=()):
        """place_cylinder / place_polygon on the device: circles as (cx, cy, r), polygons
        as (n, 2) vertex arrays; cells only become solid."""
        c = np.ascontiguousarray(np.asarray(circles, np.float64).reshape(-1, 3))
        polys = [np.asarray(p, np.float64).reshape(-1, 2) for p in polygons]
        xy = np.ascontiguousarray(np.concatenate(polys) if polys else np.zeros((0, 2)))
        counts = np.array([len(p) for p in polys], np.int32)
        _check(lib().lbm_place_obstacles(self._h, l, c.ctypes.data_as(_dp), len(c), xy.ctypes.data_as(_dp),
                                         counts.ctypes.data_as(_i32p), len(polys)))

    def obstacle(self, l: int = 0) -> np.ndarray:
        m = np.zeros(self.shape(l), np.uint8)
        _check(lib().lbm_get_obstacle(self._h, l, m.ctypes.data_as(_u8p)))
        return m

    def set_bb_circle(self, l, cx, cy, r, scheme=BB_MEI):
        _check(lib().lbm_set_bb_circle(self._h, l, cx, cy, r, scheme))

    def set_bb_polygon(self, l, xy, scheme=BB_MEI):
        p = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
        _check(lib().lbm_set_bb_polygon(self._h, l, p.ctypes.data_as(_dp), len(p) // 2, scheme))

    def set_moving_wall(self, l, omega, rcx, rcy):
        _check(lib().lbm_set_moving_wall(self._h, l, omega, rcx, rcy))

    def set_force_radius(self, l, cx, cy, r):
        _check(lib().lbm_set_force_radius(self._h, l, cx, cy, r))

    def set_force_slots(self, l, n):
        _check(lib().lbm_set_force_slots(self._h, l, n))

    def set_links(self, l, links: Sequence[tuple]):
        """links: iterable of (cell, dir, scheme, slot, q, uwx, uwy)."""
        arr = (LbmLink * max(1, len(links)))()
        for k, t in enumerate(links):
            arr[k] = LbmLink(*t)
        _check(lib().lbm_set_links(self._h, l, arr, len(links)))

    def set_params(self, l, tau, u_in):
        _check(lib().lbm_set_params(self._h, l, tau, u_in))
        self.descs[l].tau, self.descs[l].u_in = tau, u_in

    # ---- state
    def upload(self, l: int, f: np.ndarray):
        ny, nx = self.shape(l)
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(ny, nx, 9)
        _check(lib().lbm_upload(self._h, l, a.ctypes.data_as(_dp)))

    def set_storage_shift(self, l: int, u: float, v: float = 0.0):
        """16-bit forms: store f - feq(1, u, v) rather than f - w (call before upload)."""
        _check(lib().lbm_set_storage_shift(self._h, l, u, v))

    def set_stochastic_rounding(self, on: bool = True, seed: int = 0):
        """16-bit forms: stochastic instead of round-to-nearest-even storage."""
        _check(lib().lbm_set_stochastic_rounding(self._h, int(on), seed & 0xFFFFFFFF))

    def init_equilibrium(self, l: int, rho: float = 1.0, u: float = 0.0, v: float = 0.0):
        _check(lib().lbm_init_equilibrium(self._h, l, rho, u, v))

    def perturb_velocity(self, l: int, x0: int, x1: int, dv: np.ndarray):
        """dv: (x1 - x0, ny), added to the transverse velocity of the fluid nodes of
        columns [x0, x1) (see seed42_draws for the drivers' values)."""
        ny, nx = self.shape(l)
        a = np.ascontiguousarray(dv, dtype=np.float64).reshape(x1 - x0, ny)
        _check(lib().lbm_perturb_velocity(self._h, l, x0, x1, a.ctypes.data_as(_dp)))

    def macros(self, l: int = 0, y0: int = 0, y1: Optional[int] = None):
        """(rho, u, v), each (y1-y0, nx) FP64, of rows [y0, y1) of the current state."""
        ny, nx = self.shape(l)
        y1 = ny if y1 is None else y1
        out = [np.empty((max(0, y1 - y0), nx)) for _ in range(3)]
        _check(lib().lbm_macros(self._h, l, y0, y1, *(o.ctypes.data_as(_dp) for o in out)))
        return tuple(out)

    def download(self, l: int = 0) -> np.ndarray:
        ny, nx = self.shape(l)
        out = np.empty((ny, nx, 9), dtype=np.float64)
        _check(lib().lbm_download(self._h, l, out.ctypes.data_as(_dp)))
        return out

    # ---- thermal double distribution (one-lattice handles, f64 / f32)
    def set_thermal(self, l: int, omega_k: float, beta: float = 0.0, g_buoyancy: float = 0.0,
                    t_ref: float = 1.0):
        _check(lib().lbm_set_thermal(self._h, l, omega_k, beta, g_buoyancy, t_ref))

    def set_thermal_rules(self, l: int, code: Optional[np.ndarray], rules: Sequence = ()):
        """code: (ny, nx) uint8 (low 4 bits: rule index + 1, THERMAL_MIRROR bit);
        rules: (t, kind) pairs, kind THERMAL_EQ_FLOW or THERMAL_EQ_REST."""
        arr = (LbmThermalRule * max(1, len(rules)))()
        for k, (t, kind) in enumerate(rules):
            arr[k].t, arr[k].kind = t, kind
        if code is None:
            cp = None
        else:
            c = np.ascontiguousarray(code, dtype=np.uint8)
            assert c.shape == self.shape(l), (c.shape, self.shape(l))
            cp = c.ctypes.data_as(_u8p)
        _check(lib().lbm_set_thermal_rules(self._h, l, cp, arr, len(rules)))

    def upload_thermal(self, l: int, g: np.ndarray):
        ny, nx = self.shape(l)
        a = np.ascontiguousarray(g, dtype=np.float64).reshape(ny, nx, 9)
        _check(lib().lbm_upload_thermal(self._h, l, a.ctypes.data_as(_dp)))

    def download_thermal(self, l: int = 0) -> np.ndarray:
        ny, nx = self.shape(l)
        out = np.empty((ny, nx, 9), dtype=np.float64)
        _check(lib().lbm_download_thermal(self._h, l, out.ctypes.data_as(_dp)))
        return out

    def temperature(self, l: int = 0, y0: int = 0, y1: Optional[int] = None) -> np.ndarray:
        ny, nx = self.shape(l)
        y1 = ny if y1 is None else y1
        out = np.empty((max(0, y1 - y0), nx))
        _check(lib().lbm_temperature(self._h, l, y0, y1, out.ctypes.data_as(_dp)))
        return out

    def step(self, n: int = 1):
        _check(lib().lbm_step(self._h, n))

    def step_noalloc(self, n: int = 1):
        """lbm_step without allocating: one launch per step, force partials in the
        library's fixed device scratch (call step(0) after changes first)."""
        _check(lib().lbm_step_noalloc(self._h, n))

    def debug_poison(self, l: int = 0):
        """NaN into lattice l's planes of the buffer the next launch writes."""
        _check(lib().lbm_debug_poison(self._h, l))

    def step_lattice(self, l: int, n: int = 1):
        _check(lib().lbm_step_lattice(self._h, l, n))

    def set_launch(self, block_x=0, rows=0, nodes_per_thread=0, options=-1, graph_steps=0):
        _check(lib().lbm_set_launch(self._h, block_x, rows, nodes_per_thread, options, graph_steps))

    def autotune(self):
        """Returns ((block_x, rows, nodes_per_thread, options), ms per step) of the fastest."""
        bx, rows, v, opt, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        _check(lib().lbm_autotune(self._h, C.byref(bx), C.byref(rows), C.byref(v), C.byref(opt),
                                  C.byref(ms)))
        return (bx.value, rows.value, v.value, opt.value), ms.value

    def forces(self, l: int = 0, n_slots: int = 1):
        """Per-step FP64 force totals since the last call: arrays (steps, n_slots)."""
        chunks_x, chunks_y = [], []
        cap = 8192
        while True:
            fx = np.empty(cap * n_slots)
            fy = np.empty(cap * n_slots)
            n = C.c_int32()
            _check(lib().lbm_forces(self._h, l, fx.ctypes.data_as(_dp), fy.ctypes.data_as(_dp),
                                    cap, C.byref(n)))
            chunks_x.append(fx[: n.value * n_slots])
            chunks_y.append(fy[: n.value * n_slots])
            if n.value < cap:
                break
        fx = np.concatenate(chunks_x).reshape(-1, n_slots)
        fy = np.concatenate(chunks_y).reshape(-1, n_slots)
        return fx, fy

    def mass(self, l: int = 0) -> float:
        m = C.c_double()
        _check(lib().lbm_mass(self._h, l, C.byref(m)))
        return m.value

    def nan_flag(self, l: int = 0) -> bool:
        f = C.c_int32()
        _check(lib().lbm_nan_flag(self._h, l, C.byref(f)))
        return bool(f.value)

    def steps_done(self, l: int = 0) -> int:
        s = C.c_int64()
        _check(lib().lbm_steps_done(self._h, l, C.byref(s)))
        return s.value

    def info(self) -> LbmInfo:
        i = LbmInfo()
        _check(lib().lbm_get_info(self._h, C.byref(i)))
        return i

    def sync(self):
        _check(lib().lbm_sync(self._h))

    def wall_distance(self, l: int = 0) -> np.ndarray:
        ny, nx = self.shape(l)
        out = np.empty((ny, nx))
        _check(lib().lbm_wall_distance(self._h, l, out.ctypes.data_as(_dp)))
        return out


# ---- host-side helpers mirroring the reference's geometry and initial conditions ----------

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
CXV = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CYV = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
_CX = tuple(int(c) for c in CXV)
_CY = tuple(int(c) for c in CYV)


def equilibrium(rho, u, v) -> np.ndarray:
    """compute_equilibrium (src/lbm_types.hpp:269-273), broadcast over the last axis."""
    rho, u, v = np.asarray(rho, float), np.asarray(u, float), np.asarray(v, float)
    eu = CXV * u[..., None] + CYV * v[..., None]
    uu = (u * u + v * v)[..., None]
    return W * rho[..., None] * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu)


class Mt19937:
    """std::mt19937 (Matsumoto & Nishimura 1998) with libstdc++'s
    uniform_real_distribution<double>: generate_canonical over two 32-bit draws,
    (g1 + g2 * 2^32) / 2^64, kept below 1.  The generator of the drivers' seed-42
    perturbation (src/main.cpp:68-69)."""

    def __init__(self, seed: int = 5489):
        mt = [seed & 0xFFFFFFFF]
        for i in range(1, 624):
            mt.append((1812433253 * (mt[-1] ^ (mt[-1] >> 30)) + i) & 0xFFFFFFFF)
        self._mt = mt
        self._i = 624

    def _twist(self):
        mt = self._mt
        for k in range(624):
            y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
            mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        self._i = 0

    def next(self) -> int:
        if self._i >= 624:
            self._twist()
        y = self._mt[self._i]
        self._i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF

    def uniform(self, a: float, b: float) -> float:
        lo = float(self.next())
        r = (lo + float(self.next()) * 4294967296.0) / 18446744073709551616.0
        if r >= 1.0:
            r = 1.0 - 2.0 ** -53
        return r * (b - a) + a


def seed42_draws(obstacle: np.ndarray, x0: int, x1: int, amp: float = 1e-4) -> np.ndarray:
    """The drivers' perturbation values (src/main.cpp:67-83) as a (x1 - x0, ny) array for
    Ensemble.perturb_velocity: std::mt19937(42) uniform(-amp, amp) draws, x outer, y inner,
    one per fluid node; 0 on solid nodes."""
    g = Mt19937(42)
    ny, nx = obstacle.shape
    x1 = min(nx, x1)
    dv = np.zeros((max(0, x1 - x0), ny))
    for x in range(x0, x1):
        for y in range(ny):
            if not obstacle[y, x]:
                dv[x - x0, y] = g.uniform(-amp, amp)
    return dv


def perturb_seed42(f: np.ndarray, obstacle: np.ndarray, x0: int, x1: int, amp: float = 1e-4) -> None:
    """The cylinder drivers' shedding trigger (src/main.cpp:67-83), in place on AoS
    populations f (ny, nx, 9): columns x0 <= x < x1, rows in order, fluid nodes get
    feq(rho, u, v + U(-amp, amp)) with the node's own moments."""
    g = Mt19937(42)
    ny, nx = obstacle.shape
    for x in range(x0, min(nx, x1)):
        for y in range(ny):
            if obstacle[y, x]:
                continue
            dv = g.uniform(-amp, amp)
            fr = f[y, x].tolist()
            rho = 0.0
            jx = 0.0
            jy = 0.0
            for i in range(9):
                rho += fr[i]
                jx += fr[i] * _CX[i]
                jy += fr[i] * _CY[i]
            u, v = jx / rho, jy / rho
            f[y, x] = equilibrium(np.array(rho), np.array(u), np.array(v + dv))


# ---- force analysis ----------------------------------------------------------------

def coefficients(fx, fy, length: float, u: float):
    """Cd, Cl = 2F / (L u^2) per step (src/main.cpp:112-115)."""
    k = 2.0 / (length * u * u)
    return np.asarray(fx, float) * k, np.asarray(fy, float) * k


def late_stats(cd, cl):
    """Mean Cd and Cl amplitude (max - min) / 2 over the latter half of the history,
    with Cl's extremes started at 0 (src/main.cpp:185-195)."""
    cd, cl = np.asarray(cd, float), np.asarray(cl, float)
    n_late = int(len(cd) * 0.5)
    cd_mean = float(np.cumsum(cd[n_late:])[-1] / (len(cd) - n_late))
    tail = cl[n_late:]
    return cd_mean, float((max(0.0, tail.max()) - min(0.0, tail.min())) / 2.0)


def welch_psd(x, nperseg: int, noverlap: int):
    """One-sided Welch power spectral density at fs = 1: periodic Hann window,
    per-segment mean removed, density scaling (the estimator scripts/postprocess.py:423-427
    takes from scipy.signal.welch)."""
    x = np.asarray(x, float)
    step = nperseg - noverlap
    n_seg = (len(x) - noverlap) // step
    win = 0.5 - 0.5 * np.cos(2.0 * np.pi * np.arange(nperseg) / nperseg)
    acc = np.zeros(nperseg // 2 + 1)
    for k in range(n_seg):
        seg = x[k * step:k * step + nperseg]
        acc += np.abs(np.fft.rfft((seg - seg.mean()) * win)) ** 2
    psd = acc / n_seg / (win * win).sum()
    psd[1:-1 if nperseg % 2 == 0 else None] *= 2.0
    return np.fft.rfftfreq(nperseg, 1.0), psd


def strouhal_welch(cl, length: float, u: float):
    """Strouhal number as scripts/postprocess.py:386-441 computes it: drop the first fifth,
    Welch with nperseg = min(n/4, 16384) and 50% overlap, peak in [0.01, 0.5] U/L,
    St = f L / U.  None when the history is too short or the band empty."""
    cl = np.asarray(cl, float)
    t = cl[len(cl) // 5:]
    if len(t) < 64:
        return None
    nperseg = min(len(t) // 4, 16384)
    if nperseg < 4:
        return None
    f, psd = welch_psd(t - t.mean(), nperseg, nperseg // 2)
    band = (f >= 0.01 * u / length) & (f <= 0.5 * u / length)
    if not band.any():
        return None
    return float(f[band][np.argmax(psd[band])] * length / u)


def strouhal_crossings(cl, length: float, u: float, tail: float = 0.5, hysteresis: float = 0.5):
    """Strouhal number from the mean period between upward crossings of Cl's mean over the
    last `tail` of the history.  The crossings are those of a Schmitt trigger: one counts
    when the signal, having been below -hysteresis * rms, rises above +hysteresis * rms,
    and it is placed (by linear interpolation) at the last upward zero crossing before
    that rise, so noise near the mean or near either threshold adds no crossings.
    Resolves St far below the Welch bin width 1/nperseg when the history holds few
    periods."""
    c = np.asarray(cl, float)
    c = c[int(len(c) * (1.0 - tail)):]
    c = c - c.mean()
    h = hysteresis * float(np.sqrt(np.mean(c * c)))
    up = (c[:-1] < 0.0) & (c[1:] >= 0.0)
    times, armed, last_up = [], False, -1
    for j in range(len(c)):
        if j > 0 and up[j - 1]:
            last_up = j - 1
        if c[j] < -h:
            armed = True
        elif armed and c[j] > h:
            i = last_up
            times.append(i + (-c[i]) / (c[i + 1] - c[i]))
            armed = False
    if len(times) < 2:
        return None
    return float(length / ((times[-1] - times[0]) / (len(times) - 1)) / u)


def circle_mask(nx, ny, cx, cy, r) -> np.ndarray:
    """place_cylinder rasterisation: sqrt(dx^2+dy^2) < r (src/lbm.hpp:931-944)."""
    y, x = np.mgrid[0:ny, 0:nx]
    dx = (x - cx).astype(float)
    dy = (y - cy).astype(float)
    return (np.sqrt(dx * dx + dy * dy) < r).astype(np.uint8)


def walls_mask(nx, ny) -> np.ndarray:
    """place_walls: x=0, x=nx-1 and y=0 are solid (src/lbm.hpp:79-87)."""
    m = np.zeros((ny, nx), np.uint8)
    m[0, :] = 1
    m[:, 0] = 1
    m[:, nx - 1] = 1
    return m


def plate_polygon(cx, cy, chord, aoa_deg, thickness=1.0) -> np.ndarray:
    """Corners of the flat-plate driver's rotated chord x thickness rectangle
    (src/flat_plate.cpp:90-105), shape (4, 2)."""
    rad = -aoa_deg * np.pi / 180.0
    ca, sa = np.cos(rad), np.sin(rad)
    hx, hy = chord / 2.0, thickness / 2.0
    return np.array([(px * ca - py * sa + cx, px * sa + py * ca + cy)
                     for px, py in ((-hx, -hy), (hx, -hy), (hx, hy), (-hx, hy))])


def polygon_mask(nx, ny, xy) -> np.ndarray:
    """place_polygon with the even-odd ray-crossing rule (src/geometry.hpp:120-134)."""
    xy = np.asarray(xy, float).reshape(-1, 2)
    y, x = np.mgrid[0:ny, 0:nx]
    px, py = x.astype(float), y.astype(float)
    inside = np.zeros((ny, nx), bool)
    n = len(xy)
    j = n - 1
    for i in range(n):
        xi, yi = xy[i]
        xj, yj = xy[j]
        cond = (yi > py) != (yj > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (xj - xi) * (py - yi) / (yj - yi) + xi
        inside ^= cond & (px < xint)
        j = i
    return inside.astype(np.uint8)
