"""FP64 CPU oracle for the lbm-2d lattice update -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` as the checker.  The product library never imports this package.

``Oracle`` wraps ``liblbm_oracle.so`` (the C restatement in lbm_oracle.c).
``RefLattice`` wraps ``_ref/libref_check.so`` / ``_ref/libref_fast.so``: the
reference solver itself, compiled from its own headers by ``oracle/Makefile``
where the reference sources are present.  Only one ``RefLattice`` may exist per
library at a time because the reference keeps its grid size in process globals
(src/lbm_types.hpp:15-32).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LBM_ORACLE_SO selects another build of the oracle (the sanitizer build of the tests).
ORACLE_SO = os.environ.get("LBM_ORACLE_SO") or os.path.join(HERE, "liblbm_oracle.so")
REF_CHECK_SO = os.path.join(HERE, "_ref", "libref_check.so")
REF_FAST_SO = os.path.join(HERE, "_ref", "libref_fast.so")

# CaseType (src/lbm_types.hpp:20-23)
CYLINDER, CAVITY, STEP, RIBS, URBAN_CANYON, DOWNWASH, SQUARE_CYLINDER, FLAT_PLATE, \
    SPORTS_BALL, PERIODIC_HILLS, CYLINDER_NEAR_WALL, SIDE_BY_SIDE, ROTATING_CYLINDER, \
    ORIFICE_PLATE = range(14)
MRT, BGK = 0, 1

_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref where the reference is present)."""
    import subprocess
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        L.orc_equilibrium.restype = C.c_double
        L.orc_equilibrium.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_weight.restype = C.c_double
        L.orc_weight.argtypes = [C.c_int]
        L.orc_macros.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_mrt_rates.argtypes = [C.c_double, _dp]
        L.orc_mrt_collide.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp,
                                      C.c_double, C.c_double, C.c_double]
        L.orc_q_circle.restype = C.c_double
        L.orc_q_circle.argtypes = [C.c_double] * 5 + [C.c_int]
        L.orc_q_polygon.restype = C.c_double
        L.orc_q_polygon.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_point_in_polygon.restype = C.c_int
        L.orc_point_in_polygon.argtypes = [C.c_double, C.c_double, _dp, C.c_int]
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name in ("orc_f", "orc_wall_dist", "orc_fx_body", "orc_fy_body"):
            getattr(L, name).restype = _dp
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_obstacle.restype = _u8p
        L.orc_obstacle.argtypes = [C.c_void_p]
        L.orc_set_case.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_collision.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_les.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_set_body_force.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_mei.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_moving_wall.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_set_faithful.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_spec.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_place_walls.argtypes = [C.c_void_p]
        L.orc_place_cylinder.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_place_polygon.argtypes = [C.c_void_p, _dp, C.c_int]
        L.orc_obstacle_changed.argtypes = [C.c_void_p]
        L.orc_init_equilibrium.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_perturb_seed42.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_step.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_steps.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_total_force.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_mass.restype = C.c_double
        L.orc_mass.argtypes = [C.c_void_p]
        L.orc_wall_distance_sweep.argtypes = [_u8p, C.c_int, C.c_int, _dp]
        L.orc_wall_distance_bfs.argtypes = [_u8p, C.c_int, C.c_int, _dp]
        L.orc_mt_seed.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_mt_next.argtypes = [C.c_void_p]
        L.orc_g.restype = _dp
        L.orc_g.argtypes = [C.c_void_p]
        L.orc_set_thermal.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 4
        L.orc_thermal_equilibrium.restype = C.c_double
        L.orc_thermal_equilibrium.argtypes = [C.c_int] + [C.c_double] * 3
        L.orc_thermal_isothermal_solids.argtypes = [C.c_void_p, C.c_double]
        L.orc_thermal_cavity_walls.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_uniform_real.restype = C.c_double
        L.orc_uniform_real.argtypes = [C.c_void_p, C.c_double, C.c_double]
        _lib = L
    return _lib


def _arr(ptr, n, dtype=np.float64):
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


# ---- per-node closed forms -------------------------------------------------------

def equilibrium(i, rho, u, v) -> float:
    return lib().orc_equilibrium(i, rho, u, v)


def thermal_equilibrium(i, T, u, v) -> float:
    return lib().orc_thermal_equilibrium(i, T, u, v)


def macros(f9):
    f9 = np.ascontiguousarray(f9, dtype=np.float64)
    r, u, v = C.c_double(), C.c_double(), C.c_double()
    lib().orc_macros(_dptr(f9), C.byref(r), C.byref(u), C.byref(v))
    return r.value, u.value, v.value


def mrt_rates(tau):
    out = np.zeros(3)
    lib().orc_mrt_rates(tau, _dptr(out))
    return out


def mrt_collide(f9, tau, cs_sq=0.0, y_plus=0.0, rates=None):
    f = np.array(f9, dtype=np.float64)
    rho, u, v = macros(f)
    r = mrt_rates(tau) if rates is None else np.asarray(rates, dtype=np.float64)
    lib().orc_mrt_collide(_dptr(f), rho, u, v, _dptr(r), cs_sq, tau, y_plus)
    return f


def q_circle(ccx, ccy, r, xf, yf, i) -> float:
    return lib().orc_q_circle(ccx, ccy, r, xf, yf, i)


def q_polygon(xy, xf, yf, i) -> float:
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
    return lib().orc_q_polygon(_dptr(xy), len(xy) // 2, xf, yf, i)


def point_in_polygon(px, py, xy) -> bool:
    xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
    return bool(lib().orc_point_in_polygon(px, py, _dptr(xy), len(xy) // 2))


def wall_distance(obstacle: np.ndarray, method: str = "bfs") -> np.ndarray:
    obs = np.ascontiguousarray(obstacle, dtype=np.uint8)
    ny, nx = obs.shape
    d = np.zeros(obs.size)
    fn = lib().orc_wall_distance_bfs if method == "bfs" else lib().orc_wall_distance_sweep
    fn(obs.ctypes.data_as(_u8p), nx, ny, _dptr(d))
    return d.reshape(ny, nx)


class MT19937:
    """std::mt19937 restated in C (lbm_oracle.c)."""

    def __init__(self, seed: int):
        self._state = C.create_string_buffer(624 * 4 + 8)
        lib().orc_mt_seed(self._state, seed)

    def next(self) -> int:
        return lib().orc_mt_next(self._state)

    def uniform(self, a, b) -> float:
        return lib().orc_uniform_real(self._state, a, b)


# ---- lattice ------------------------------------------------------------------------

class Oracle:
    """One lattice of the FP64 restatement; arrays are numpy views of its buffers."""

    def __init__(self, nx, ny, case=CYLINDER, collision=MRT, les=False, cs=0.12, faithful=False):
        self.nx, self.ny = nx, ny
        self._h = lib().orc_create(nx, ny)
        lib().orc_set_case(self._h, case)
        lib().orc_set_collision(self._h, collision)
        lib().orc_set_les(self._h, int(les), cs)
        lib().orc_set_faithful(self._h, int(faithful))
        self.case = case

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    @property
    def n(self):
        return self.nx * self.ny

    @property
    def f(self) -> np.ndarray:
        """AoS populations, shape (ny, nx, 9); a live view (re-read after each step)."""
        return _arr(lib().orc_f(self._h), self.n * 9).reshape(self.ny, self.nx, 9)

    @property
    def obstacle(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_obstacle(self._h), shape=(self.ny, self.nx))

    @property
    def wall_dist(self):
        return _arr(lib().orc_wall_dist(self._h), self.n).reshape(self.ny, self.nx)

    @property
    def fx_body(self):
        return _arr(lib().orc_fx_body(self._h), self.n).reshape(self.ny, self.nx)

    @property
    def fy_body(self):
        return _arr(lib().orc_fy_body(self._h), self.n).reshape(self.ny, self.nx)

    def set_body_force(self, fx):
        lib().orc_set_body_force(self._h, fx)

    def set_mei(self, on: bool):
        lib().orc_set_mei(self._h, int(on))

    def set_moving_wall(self, omega, rcx, rcy, on=True):
        lib().orc_set_moving_wall(self._h, int(on), omega, rcx, rcy)

    def set_spec(self, inlet_equilibrium: bool = True, free_slip: bool = True):
        """The configs' boundaries instead of the reference's (cylinder-type cases):
        equilibrium inlet feq(1, u_in, 0) and free-slip y edges."""
        lib().orc_set_spec(self._h, int(inlet_equilibrium), int(free_slip))

    def set_les(self, on: bool, cs: float):
        lib().orc_set_les(self._h, int(on), cs)

    def set_obstacle(self, mask: np.ndarray):
        self.obstacle[:] = np.asarray(mask, dtype=np.uint8)
        lib().orc_obstacle_changed(self._h)

    def place_walls(self):
        lib().orc_place_walls(self._h)

    def place_cylinder(self, cx, cy, r):
        lib().orc_place_cylinder(self._h, cx, cy, r)

    def place_polygon(self, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
        lib().orc_place_polygon(self._h, _dptr(xy), len(xy) // 2)

    def init_equilibrium(self, rho=1.0, u=0.0, v=0.0):
        lib().orc_init_equilibrium(self._h, rho, u, v)

    def perturb_seed42(self, x0, x1):
        # the restated loop (like the driver's) does not clamp its column range
        if not 0 <= x0 < x1 <= self.nx:
            raise ValueError(f"perturbation columns [{x0}, {x1}) outside a lattice {self.nx} wide")
        lib().orc_perturb_seed42(self._h, x0, x1)

    def set_f(self, f):
        self.f[...] = np.asarray(f, dtype=np.float64).reshape(self.ny, self.nx, 9)

    def step(self, tau, u_in, n=1):
        lib().orc_steps(self._h, n, tau, u_in)

    @property
    def g(self) -> np.ndarray:
        """AoS thermal populations, shape (ny, nx, 9); a live view."""
        return _arr(lib().orc_g(self._h), self.n * 9).reshape(self.ny, self.nx, 9)

    def set_thermal(self, omega_k, beta=0.0, g_buoyancy=0.0, t_ref=1.0, on=True):
        lib().orc_set_thermal(self._h, int(on), omega_k, beta, g_buoyancy, t_ref)

    def set_g(self, g):
        self.g[...] = np.asarray(g, dtype=np.float64).reshape(self.ny, self.nx, 9)

    def thermal_isothermal_solids(self, t_wall):
        lib().orc_thermal_isothermal_solids(self._h, t_wall)

    def thermal_cavity_walls(self, t_hot, t_cold):
        lib().orc_thermal_cavity_walls(self._h, t_hot, t_cold)

    def total_force(self):
        fx, fy = C.c_double(), C.c_double()
        lib().orc_total_force(self._h, C.byref(fx), C.byref(fy))
        return fx.value, fy.value

    def mass(self):
        return lib().orc_mass(self._h)


# ---- the reference itself (oracle/_ref) -----------------------------------------

_ref_libs: dict = {}


def ref_available(fast: bool = False) -> bool:
    return os.path.exists(REF_FAST_SO if fast else REF_CHECK_SO)


def _ref_lib(fast: bool) -> C.CDLL:
    path = REF_FAST_SO if fast else REF_CHECK_SO
    if path not in _ref_libs:
        L = C.CDLL(path)
        L.ref_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ref_place_cylinder.argtypes = [C.c_int] * 3
        L.ref_place_polygon.argtypes = [_dp, C.c_int]
        L.ref_set_mei.argtypes = [C.c_int]
        L.ref_set_moving_wall.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_set_body_force.argtypes = [C.c_double]
        L.ref_init_equilibrium.argtypes = [C.c_double] * 3
        L.ref_perturb_seed42.argtypes = [C.c_int, C.c_int]
        L.ref_uniform_draws.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.ref_step.argtypes = [C.c_int, C.c_double, C.c_double]
        L.ref_time_steps.restype = C.c_double
        L.ref_time_steps.argtypes = [C.c_int, C.c_double, C.c_double]
        for name in ("ref_f", "ref_fx_body", "ref_fy_body", "ref_wall_dist"):
            getattr(L, name).restype = _dp
        L.ref_obstacle.restype = _u8p
        L.ref_mrt_collide.argtypes = [_dp] + [C.c_double] * 6
        L.ref_set_thermal.argtypes = [C.c_int] + [C.c_double] * 4
        L.ref_g.restype = _dp
        L.ref_thermal_isothermal_solids.argtypes = [C.c_double]
        L.ref_thermal_cavity_walls.argtypes = [C.c_double, C.c_double]
        L.ref_q.restype = C.c_double
        L.ref_q.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.c_int,
                            C.c_double, C.c_double, C.c_int]
        _ref_libs[path] = L
    return _ref_libs[path]


class RefLattice:
    """The reference solver (compiled from its sources into oracle/_ref)."""

    def __init__(self, nx, ny, case=CYLINDER, collision=MRT, les=False, cs=0.12, fast=False):
        self.L = _ref_lib(fast)
        self.nx, self.ny = nx, ny
        self.L.ref_setup(nx, ny, case, collision, int(les), cs)

    @property
    def n(self):
        return self.nx * self.ny

    @property
    def f(self):
        return _arr(self.L.ref_f(), self.n * 9).reshape(self.ny, self.nx, 9)

    @property
    def obstacle(self):
        return np.ctypeslib.as_array(self.L.ref_obstacle(), shape=(self.ny, self.nx))

    @property
    def wall_dist(self):
        return _arr(self.L.ref_wall_dist(), self.n).reshape(self.ny, self.nx)

    @property
    def fx_body(self):
        return _arr(self.L.ref_fx_body(), self.n).reshape(self.ny, self.nx)

    @property
    def fy_body(self):
        return _arr(self.L.ref_fy_body(), self.n).reshape(self.ny, self.nx)

    def place_walls(self):
        self.L.ref_place_walls()

    def place_cylinder(self, cx, cy, r):
        self.L.ref_place_cylinder(cx, cy, r)

    def place_polygon(self, xy):
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
        self.L.ref_place_polygon(_dptr(xy), len(xy) // 2)

    def set_obstacle(self, mask):
        self.obstacle[:] = np.asarray(mask, dtype=np.uint8)

    def set_mei(self, on):
        self.L.ref_set_mei(int(on))

    def set_moving_wall(self, omega, rcx, rcy, on=True):
        self.L.ref_set_moving_wall(int(on), omega, rcx, rcy)

    def set_body_force(self, fx):
        self.L.ref_set_body_force(fx)

    def init_equilibrium(self, rho=1.0, u=0.0, v=0.0):
        self.L.ref_init_equilibrium(rho, u, v)

    def perturb_seed42(self, x0, x1):
        # the reference's loop does not clamp its column range
        if not 0 <= x0 < x1 <= self.nx:
            raise ValueError(f"perturbation columns [{x0}, {x1}) outside a lattice {self.nx} wide")
        self.L.ref_perturb_seed42(x0, x1)

    def set_f(self, f):
        self.f[...] = np.asarray(f, dtype=np.float64).reshape(self.ny, self.nx, 9)

    @property
    def g(self):
        return _arr(self.L.ref_g(), self.n * 9).reshape(self.ny, self.nx, 9)

    def set_thermal(self, omega_k, beta=0.0, g_buoyancy=0.0, t_ref=1.0, on=True):
        self.L.ref_set_thermal(int(on), omega_k, beta, g_buoyancy, t_ref)

    def set_g(self, g):
        self.g[...] = np.asarray(g, dtype=np.float64).reshape(self.ny, self.nx, 9)

    def thermal_isothermal_solids(self, t_wall):
        self.L.ref_thermal_isothermal_solids(t_wall)

    def thermal_cavity_walls(self, t_hot, t_cold):
        self.L.ref_thermal_cavity_walls(t_hot, t_cold)

    def step(self, tau, u_in, n=1):
        self.L.ref_step(n, tau, u_in)

    def time_steps(self, n, tau, u_in) -> float:
        return self.L.ref_time_steps(n, tau, u_in)

    def total_force(self):
        """Sum over nodes in node order, as the drivers do (src/main.cpp:106-110)."""
        return float(np.cumsum(self.fx_body.ravel())[-1]), float(np.cumsum(self.fy_body.ravel())[-1])


def ref_uniform_draws(n, a, b, fast=False):
    out = np.zeros(n)
    _ref_lib(fast).ref_uniform_draws(n, a, b, _dptr(out))
    return out


def ref_mrt_collide(f9, tau, cs_sq=0.0, y_plus=0.0, fast=False):
    f = np.array(f9, dtype=np.float64)
    rho, u, v = macros(f)
    _ref_lib(fast).ref_mrt_collide(_dptr(f), rho, u, v, tau, cs_sq, y_plus)
    return f


def ref_q(xf, yf, i, circle=None, polygon=None, fast=False):
    L = _ref_lib(fast)
    if polygon is not None:
        xy = np.ascontiguousarray(polygon, dtype=np.float64).reshape(-1)
        return L.ref_q(1, 0.0, 0.0, 0.0, _dptr(xy), len(xy) // 2, xf, yf, i)
    ccx, ccy, r = circle
    dummy = np.zeros(2)
    return L.ref_q(0, ccx, ccy, r, _dptr(dummy), 0, xf, yf, i)
