"""Small lattices exercising every branch of the step, built identically in the
FP64 oracle and in the GPU library.

Each case mirrors one of the reference's case drivers (src/*.cpp) at a size the
oracle runs in well under a second per hundred steps.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Optional

import numpy as np

import oracle as O
import lbm2d as G


@dataclasses.dataclass
class Case:
    name: str
    nx: int
    ny: int
    case: int
    tau: float
    u: float
    steps: int
    setup: Callable  # (obj) -> None, works for Oracle and RefLattice
    init: tuple = (1.0, 0.1, 0.0)
    pert: Optional[tuple] = None
    les: bool = False
    cs: float = 0.12
    collision: int = O.MRT
    body_force: float = 0.0
    # GPU-side geometry: list of ("circle", cx, cy, r, scheme) / ("polygon", xy, scheme) /
    # ("moving", omega, rcx, rcy) / ("force_radius", cx, cy, r)
    gpu_geom: tuple = ()


def _cyl(cx, cy, r):
    return lambda s: s.place_cylinder(cx, cy, r)


def _rot(cx, cy, r, omega, mei=True):
    def f(s):
        s.place_cylinder(cx, cy, r)
        s.set_moving_wall(omega, cx, cy)
        if not mei:
            s.set_mei(False)
    return f


def plate_polygon(cx, cy, chord, aoa_deg, thickness=1.0):
    """Rotated chord x thickness rectangle of the flat-plate driver (src/flat_plate.cpp:90-105)."""
    rad = -aoa_deg * np.pi / 180.0
    ca, sa = np.cos(rad), np.sin(rad)
    hx, hy = chord / 2.0, thickness / 2.0
    pts = []
    for px, py in ((-hx, -hy), (hx, -hy), (hx, hy), (-hx, hy)):
        pts.append((px * ca - py * sa + cx, px * sa + py * ca + cy))
    return np.array(pts)


def _plate(xy):
    return lambda s: s.place_polygon(xy)


def _step_geometry(nx, ny):
    m = np.zeros((ny, nx), np.uint8)
    m[0, :] = 1
    m[ny - 1, :] = 1
    m[: ny // 3, : nx // 4] = 1
    return m


def _hills(nx, ny):
    """Periodic-hills obstacle (src/periodic_hills.cpp:95-115) at a small size."""
    H = ny * 2 // 3
    h_max = max(1, int(H / 6))
    m = np.zeros((ny, nx), np.uint8)
    for x in range(nx):
        xr = x / nx * 3
        hill = H - h_max * 0.5 * (1 - np.cos(2 * np.pi * xr)) * (1 - np.cos(2 * np.pi * xr))
        hy = int(np.ceil(hill))
        m[ny - 1, x] = 1
        m[:hy, x] = 1
    return m


def _mask_setup(maskfn):
    def f(s):
        s.set_obstacle(maskfn)
    return f


def _near_wall(cx, cy, r):
    def f(s):
        s.place_cylinder(cx, cy, r)
        m = s.obstacle.copy()
        m[0, :] = 1  # bottom wall (src/cylinder_near_wall.cpp:93-96)
        s.set_obstacle(m)
    return f


PLATE = plate_polygon(24, 24, 22.0, 10.0)
HILLS = _hills(96, 36)
STEPM = _step_geometry(96, 36)

CASES = {
    "cavity": Case("cavity", 32, 32, O.CAVITY, 0.8, 0.1, 200, lambda s: s.place_walls(),
                   init=(1.0, 0.0, 0.0)),
    "cylinder": Case("cylinder", 96, 48, O.CYLINDER, 0.56, 0.1, 100, _cyl(24, 25, 6),
                     pert=(29, 84), gpu_geom=(("circle", 24, 25, 6, G.BB_MEI),)),
    "rotating": Case("rotating", 96, 48, O.ROTATING_CYLINDER, 0.6, 0.1, 60,
                     _rot(24, 24, 6, 0.1 * 0.1 / 6),
                     gpu_geom=(("circle", 24, 24, 6, G.BB_MEI), ("moving", 0.1 * 0.1 / 6, 24, 24))),
    "bouzidi": Case("bouzidi", 96, 48, O.ROTATING_CYLINDER, 0.6, 0.1, 60,
                    _rot(24, 24, 6, 0.01, mei=False),
                    gpu_geom=(("circle", 24, 24, 6, G.BB_BOUZIDI), ("moving", 0.01, 24, 24))),
    "les": Case("les", 96, 48, O.CYLINDER, 0.52, 0.1, 60, _cyl(24, 25, 6), pert=(29, 84),
                les=True, cs=0.1, gpu_geom=(("circle", 24, 25, 6, G.BB_MEI),)),
    "bgk": Case("bgk", 96, 48, O.CYLINDER, 0.7, 0.1, 60, _cyl(24, 25, 6), pert=(29, 84),
                collision=O.BGK, gpu_geom=(("circle", 24, 25, 6, G.BB_MEI),)),
    "plate": Case("plate", 96, 48, O.FLAT_PLATE, 0.6, 0.1, 60, _plate(PLATE),
                  gpu_geom=(("polygon", PLATE, G.BB_MEI),)),
    "step": Case("step", 96, 36, O.STEP, 0.7, 0.1, 80, _mask_setup(STEPM), init=(1.0, 0.0, 0.0)),
    "hills": Case("hills", 96, 36, O.PERIODIC_HILLS, 0.8, 0.1, 60, _mask_setup(HILLS),
                  init=(1.0, 0.0, 0.0), body_force=2e-5),
    "near_wall": Case("near_wall", 96, 48, O.CYLINDER_NEAR_WALL, 0.6, 0.1, 60,
                      _near_wall(24, 14, 6), pert=(29, 84),
                      gpu_geom=(("circle", 24, 14, 6, G.BB_MEI), ("force_radius", 24, 14, 6))),
}


def _cyl_and_edge_block(cx, cy, r):
    """A circle, plus a block on the top edge row (the free-slip reflection's solid
    emitters)."""
    def f(s):
        s.place_cylinder(cx, cy, r)
        m = s.obstacle.copy()
        m[-5:, 60:64] = 1
        s.set_obstacle(m)
    return f


# The configs' boundary wording (equilibrium inlet, free-slip y edges): no reference
# counterpart, so no reference fixtures; the oracle's option is pinned by invariants
# (tests/test_oracle_spec_bcs.py).
SPEC_CASES = {
    "spec_cylinder": Case("spec_cylinder", 96, 48, O.CYLINDER, 0.56, 0.1, 100, _cyl_and_edge_block(24, 25, 6),
                          pert=(29, 84), gpu_geom=(("circle", 24, 25, 6, G.BB_MEI),)),
    "spec_les": Case("spec_les", 96, 48, O.CYLINDER, 0.52, 0.1, 60, _cyl(24, 25, 6), pert=(29, 84),
                     les=True, cs=0.1, gpu_geom=(("circle", 24, 25, 6, G.BB_MEI),)),
    "spec_bouzidi": Case("spec_bouzidi", 96, 48, O.ROTATING_CYLINDER, 0.6, 0.1, 60, _rot(24, 24, 6, 0.01, mei=False),
                         gpu_geom=(("circle", 24, 24, 6, G.BB_BOUZIDI), ("moving", 0.01, 24, 24))),
}

# Widths whose last 64-node column block holds only the outlet column (nx % 64 == 1):
# an ensemble cuts their rows into right-aligned two-step tiles.  The perturbation
# columns stay inside the lattice (the restated loop does not clamp them).
NARROW_CASES = {
    "cylinder_129": dataclasses.replace(CASES["cylinder"], name="cylinder_129", nx=129),
    "les_65": dataclasses.replace(CASES["les"], name="les_65", nx=65, pert=(29, 60)),
}


def build_oracle(c: Case, cls=O.Oracle):
    o = cls(c.nx, c.ny, c.case, collision=c.collision, les=c.les, cs=c.cs)
    if c.name.startswith("spec_"):
        o.set_spec(inlet_equilibrium=True, free_slip=True)
    c.setup(o)
    if c.body_force:
        o.set_body_force(c.body_force)
    o.init_equilibrium(*c.init)
    if c.pert:
        o.perturb_seed42(*c.pert)
    return o


def gpu_desc(c: Case):
    d = G.desc_for_case(c.case, c.nx, c.ny, tau=c.tau, u_in=c.u, les=c.les, cs=c.cs,
                        collision=c.collision, body_force_x=c.body_force)
    if c.name.startswith("spec_"):
        d.inlet, d.y_edge = G.INLET_EQUILIBRIUM, G.EDGE_FREE_SLIP
    return d


def apply_gpu_geometry(ens: G.Ensemble, l: int, c: Case, obstacle: np.ndarray):
    ens.set_obstacle(l, obstacle)
    for g in c.gpu_geom:
        if g[0] == "circle":
            ens.set_bb_circle(l, g[1], g[2], g[3], g[4])
        elif g[0] == "polygon":
            ens.set_bb_polygon(l, g[1], g[2])
        elif g[0] == "moving":
            ens.set_moving_wall(l, g[1], g[2], g[3])
        elif g[0] == "force_radius":
            ens.set_force_radius(l, g[1], g[2], g[3])


def build_gpu(c: Case, storage: str, o=None, scale_exp: int = 8):
    """GPU ensemble of one lattice with the oracle's obstacle mask and initial state."""
    o = o if o is not None else build_oracle(c)
    ens = G.Ensemble([gpu_desc(c)], storage=storage, scale_exp=scale_exp)
    apply_gpu_geometry(ens, 0, c, o.obstacle.copy())
    ens.upload(0, o.f.copy())
    return ens, o


# ---- thermal double distribution ----------------------------------------------------

@dataclasses.dataclass
class ThermalCase:
    """A momentum case with the thermal populations on (src/lbm.hpp:356-397, 489-528).

    t_field(x, y) is the initial temperature of fluid nodes (g = thermal equilibrium at
    the node's initial velocity); solid nodes start as the drivers leave them before their
    loop.  walls is the driver's loop after each step: ("solids", t_wall) as
    src/heated_cylinder.cpp:155-164, ("cavity", t_hot, t_cold) as
    src/thermal_cavity.cpp:153-188, or None."""
    name: str
    base: Case
    omega_k: float
    beta: float
    g_buoyancy: float
    t_ref: float
    t_field: Callable
    walls: Optional[tuple] = None
    steps: int = 60


def _blob(x0, y0, amp, width):
    return lambda x, y: 1.0 + amp * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / width)


def thermal_eq(T, u, v):
    """Thermal equilibrium of every slot, (..., 9) (src/thermal.hpp:24-30)."""
    cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    T, u, v = (np.asarray(a, float)[..., None] for a in (T, u, v))
    eu = cx * u + cy * v
    return w * T * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * (u * u + v * v))


def _alpha(tau):
    return (tau - 0.5) / 3.0 / 0.71  # Pr = 0.71 (src/heated_cylinder.cpp:103)


def _cavity_walls(nx, ny):
    m = np.zeros((ny, nx), np.uint8)
    m[:, 0] = m[:, nx - 1] = 1
    m[0, :] = m[ny - 1, :] = 1
    return m


CAVW = _cavity_walls(32, 32)
_RA, _TC_TAU = 1e5, 0.8
_GB = _RA * ((_TC_TAU - 0.5) / 3.0) * _alpha(_TC_TAU) / (1.0 * 30.0 ** 3)

THERMAL_CASES = {
    # heated cylinder with a warm blob and strong buoyancy (src/heated_cylinder.cpp)
    "heated": ThermalCase("heated", CASES["cylinder"], 1.0 / (0.5 + 3.0 * _alpha(0.56)), 0.01, -1e-3, 1.0,
                          _blob(45, 20, 0.5, 60.0), walls=("solids", 1.5)),
    "heated_les": ThermalCase("heated_les", CASES["les"], 1.0 / (0.5 + 3.0 * _alpha(0.52)), 0.01, -1e-3, 1.0,
                              _blob(40, 30, 0.5, 40.0), walls=("solids", 1.5), steps=40),
    "heated_bgk": ThermalCase("heated_bgk", CASES["bgk"], 1.0 / (0.5 + 3.0 * _alpha(0.7)), 0.0, 0.0, 1.0,
                              _blob(40, 24, 0.3, 80.0), walls=("solids", 1.2), steps=40),
    # natural convection in a walled box on the default (cylinder) case, no inflow
    # (src/thermal_cavity.cpp: Ra = 1e5, tau = 0.8, beta = 0.001)
    "convection": ThermalCase("convection",
                              Case("convection", 32, 32, O.CYLINDER, _TC_TAU, 0.0, 200, _mask_setup(CAVW),
                                   init=(1.0, 0.0, 0.0)),
                              1.0 / (0.5 + 3.0 * _alpha(_TC_TAU)), 0.001, -_GB / 0.001, 1.0,
                              lambda x, y: np.ones_like(x, float), walls=("cavity", 1.5, 0.5), steps=200),
    # lid-driven cavity: momentum walls bounce, the thermal populations wrap in y
    "lid": ThermalCase("lid", CASES["cavity"], 1.4, 0.005, -2e-3, 1.0, _blob(16, 12, 0.4, 20.0), steps=80),
    # periodic hills: the thermal populations wrap in x too
    "hills": ThermalCase("hills", CASES["hills"], 1.2, 0.01, -1e-3, 1.0, _blob(30, 26, 0.5, 30.0)),
    # rotating cylinder with Bouzidi links and the Ladd wall term
    "rotating": ThermalCase("rotating", CASES["bouzidi"], 1.0 / (0.5 + 3.0 * _alpha(0.6)), 0.01, -1e-3, 1.0,
                            _blob(36, 30, 0.4, 50.0), walls=("solids", 1.3), steps=40),
}


def thermal_initial_g(tc: ThermalCase, obstacle: np.ndarray) -> np.ndarray:
    c = tc.base
    y, x = np.mgrid[0:c.ny, 0:c.nx].astype(float)
    rho0, u0, v0 = c.init
    g = thermal_eq(tc.t_field(x, y), np.full(x.shape, u0), np.full(x.shape, v0))
    solid = obstacle != 0
    if tc.walls and tc.walls[0] == "solids":
        g[solid] = thermal_eq(tc.walls[1], 0.0, 0.0)
    elif tc.walls and tc.walls[0] == "cavity":
        # the driver's setup before its loop (src/thermal_cavity.cpp:127-144): hot / cold
        # side walls, top and bottom walls mirrored in place in slot order
        th, tcold = tc.walls[1], tc.walls[2]
        g[:, 0] = thermal_eq(th, 0.0, 0.0)
        g[:, c.nx - 1] = thermal_eq(tcold, 0.0, 0.0)
        opp = [0, 3, 4, 1, 2, 7, 8, 5, 6]
        for yy in (0, c.ny - 1):
            for xx in range(1, c.nx - 1):
                for i in range(9):
                    g[yy, xx, i] = g[yy, xx, opp[i]]
    return g


def build_thermal_oracle(tc: ThermalCase, cls=O.Oracle):
    o = build_oracle(tc.base, cls=cls)
    o.set_thermal(tc.omega_k, tc.beta, tc.g_buoyancy, tc.t_ref)
    o.set_g(thermal_initial_g(tc, o.obstacle))
    return o


def thermal_walls(o, tc: ThermalCase):
    """The driver's loop after a step, on an Oracle or RefLattice."""
    if not tc.walls:
        return
    if tc.walls[0] == "solids":
        o.thermal_isothermal_solids(tc.walls[1])
    else:
        o.thermal_cavity_walls(tc.walls[1], tc.walls[2])


def thermal_step(o, tc: ThermalCase):
    o.step(tc.base.tau, tc.base.u)
    thermal_walls(o, tc)


def gpu_thermal_rules(tc: ThermalCase, obstacle: np.ndarray):
    """The driver's wall loop as per-node conditions of the GPU library: (code, rules)."""
    if not tc.walls:
        return None, []
    ny, nx = obstacle.shape
    code = np.zeros((ny, nx), np.uint8)
    if tc.walls[0] == "solids":
        code[obstacle != 0] = 1
        return code, [(tc.walls[1], G.THERMAL_EQ_REST)]
    code[1:ny - 1, 1] = 1
    code[1:ny - 1, nx - 2] = 2
    code[1, 1:nx - 1] |= G.THERMAL_MIRROR
    code[ny - 2, 1:nx - 1] |= G.THERMAL_MIRROR
    return code, [(tc.walls[1], G.THERMAL_EQ_FLOW), (tc.walls[2], G.THERMAL_EQ_FLOW)]


def build_gpu_thermal(tc: ThermalCase, storage: str):
    """GPU lattice with the thermal populations, state and conditions of the oracle's."""
    o = build_thermal_oracle(tc)
    ens, _ = build_gpu(tc.base, storage, o=o)
    ens.set_thermal(0, tc.omega_k, tc.beta, tc.g_buoyancy, tc.t_ref)
    ens.upload_thermal(0, o.g.copy())
    code, rules = gpu_thermal_rules(tc, o.obstacle)
    if code is not None:
        ens.set_thermal_rules(0, code, rules)
    return ens, o


# ---- classes of two-step tiles (lbm_kernels.hip tile_class_kernel), restated on the mask
#
# Node flags as the geometry pass sets them: a solid node, a fluid node with a solid
# neighbour (boundary links; neighbours wrap across periodic edges only) and every edge
# node are general-path nodes.  Inlet and outlet column nodes (CELL_COLUMN, rows 1..ny-2 of
# an open-x lattice): the inlet's from row 1 (Zou/He, equilibrium) or inlet_y0 (parabolic),
# unless solid or with links; the outlet's unless it or the node before it is.  A node
# whose general path reduces to wrapped pulls (wrap_only) lies on periodic edges only and
# is no solid, link, inlet, outlet or lid node.

_CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
_CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def node_kinds(obstacle, d):
    """(special, allowed) masks: general-path nodes, and the nodes an edge tile may hold
    (plain, column or wrap-only)."""
    solid = np.asarray(obstacle).astype(bool)
    ny, nx = solid.shape
    xp, yp = d.x_edge == G.EDGE_PERIODIC, d.y_edge == G.EDGE_PERIODIC
    links = np.zeros_like(solid)
    for cx, cy in zip(_CX[1:], _CY[1:]):
        s = np.roll(np.roll(solid, -cy, axis=0), -cx, axis=1)
        if not xp:
            s[:, -1 if cx == 1 else 0] &= cx == 0
        if not yp:
            s[-1 if cy == 1 else 0, :] &= cy == 0
        links |= s
    links &= ~solid
    edge = np.zeros_like(solid)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    special = solid | links | edge
    blocked = solid | links
    column = np.zeros_like(solid)
    y = np.arange(ny)
    if d.x_edge == G.EDGE_OPEN and nx >= 4:
        if d.inlet in (G.INLET_ZOU_HE, G.INLET_EQUILIBRIUM):
            start = 1
        elif d.inlet == G.INLET_PARABOLIC:
            start = max(1, d.inlet_y0)
        else:
            start = ny
        r = (y >= start) & (y <= ny - 2)
        column[r, 0] = ~blocked[r, 0]
        if d.outlet:
            r = (y >= 1) & (y <= ny - 2)
            column[r, nx - 1] = ~blocked[r, nx - 1] & ~blocked[r, nx - 2]
    yy, xx = np.mgrid[0:ny, 0:nx]
    wrap = ~blocked & ~column
    if not xp:
        wrap &= (xx != 0) & (xx != nx - 1)
    if not yp:
        wrap &= (yy != 0) & (yy != ny - 1)
    if d.lid:
        wrap &= yy != ny - 1
    if d.inlet != G.INLET_NONE:
        wrap &= xx != 0
    if d.outlet:
        wrap &= xx != nx - 1
    return special, ~special | column | wrap


def ring_nodes(d, nx, ny, x0, y0, rows):
    """Lattice rows and columns of a tile's ring (66 x (rows + 2) from (x0 - 1, y0 - 1)),
    wrapped around every edge, and the mask of positions on the lattice or across a
    periodic edge, at most one node off it."""
    ys, xs = y0 - 1 + np.arange(rows + 2), x0 - 1 + np.arange(66)
    iny = ((ys >= 0) & (ys < ny) | (d.y_edge == G.EDGE_PERIODIC)) & (ys >= -1) & (ys <= ny)
    inx = ((xs >= 0) & (xs < nx) | (d.x_edge == G.EDGE_PERIODIC)) & (xs >= -1) & (xs <= nx)
    return ys % ny, xs % nx, iny[:, None] & inx[None, :]


def tile_classes(obstacle, d, rows, x_start=0, written_solids=False):
    """[(x0, y0, cls)] of a lattice's 64 x rows two-step tiles: 1 bulk, 2 edge, 3 empty
    (every node solid and, on a lattice without Bouzidi links, never written), 0 other."""
    special, allowed = node_kinds(obstacle, d)
    solid = np.asarray(obstacle).astype(bool)
    ny, nx = special.shape
    out = []
    for y0 in range(0, ny, rows):
        for x0 in range(x_start, nx, 64):
            inside = x0 >= 1 and x0 + 64 <= nx - 1 and y0 >= 1 and y0 + rows <= ny - 1
            # edge tiles start on a lattice node; an outlet node needs the column before it
            whole = x0 >= 0 and not (d.outlet and x0 == nx - 1)
            cls = 0
            if not written_solids and solid[y0:y0 + rows, max(x0, 0):x0 + 64].all():
                cls = 3
            elif inside and not special[y0 - 1:y0 + rows + 1, x0 - 1:x0 + 65].any():
                cls = 1
            elif whole:
                ys, xs, inm = ring_nodes(d, nx, ny, x0, y0, rows)
                if allowed[np.ix_(ys, xs)][inm].all():
                    cls = 2
            out.append((x0, y0, cls))
    return out


def split_counts(obstacle, d, rows, x_start=0, written_solids=False):
    """(bulk, edge, other, empty) tile counts of a split two-step launch: the other tiles
    are cut into 64 x min(rows, 16) tiles."""
    ny = np.asarray(obstacle).shape[0]
    bulk = edge = other = empty = 0
    for x0, y0, cls in tile_classes(obstacle, d, rows, x_start, written_solids):
        if cls == 1:
            bulk += 1
        elif cls == 2:
            edge += 1
        elif cls == 3:
            empty += 1
        else:
            other += -(-min(rows, ny - y0) // min(rows, 16))
    return bulk, edge, other, empty
