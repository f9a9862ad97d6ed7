"""Pins the FP64 oracle without the reference: its own unit tests restated
(src/lbm_test.cpp), textbook closed forms, brute force on tiny grids, invariants.

Each check is chosen so that a plausible slip in the oracle (dropped term, sign,
index, transposed operand) fails it; none retypes the oracle's formula.
"""
import math

import numpy as np
import pytest

import oracle as O

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


# ---------------------------------------------------------------- src/lbm_test.cpp restated

def test_equilibrium_sum_equals_rho():  # src/lbm_test.cpp:12-20
    assert abs(sum(O.equilibrium(i, 1.2, 0.1, 0.05) for i in range(9)) - 1.2) < 1e-12


def test_equilibrium_momentum():  # :22-33
    mx = sum(O.equilibrium(i, 1.0, 0.15, 0.0) * CX[i] for i in range(9))
    my = sum(O.equilibrium(i, 1.0, 0.15, 0.0) * CY[i] for i in range(9))
    assert abs(mx - 0.15) < 1e-12 and abs(my) < 1e-12


def test_equilibrium_rest_state_and_positivity():  # :35-52
    for i in range(9):
        assert abs(O.equilibrium(i, 1.0, 0.0, 0.0) - W[i]) < 1e-15
        assert O.equilibrium(i, 1.0, 0.1, 0.05) > 0


def test_macros_recover_input_and_zero_velocity():  # :57-86
    f = [O.equilibrium(i, 1.5, 0.12, -0.08) for i in range(9)]
    r, u, v = O.macros(f)
    assert abs(r - 1.5) < 1e-12 and abs(u - 0.12) < 1e-12 and abs(v + 0.08) < 1e-12
    r, u, v = O.macros(W)
    assert abs(r - 1) < 1e-15 and abs(u) < 1e-15 and abs(v) < 1e-15


def test_opposites():  # :101-111
    for i in range(9):
        j = O.lib().orc_opposite(i)
        assert CX[j] == -CX[i] and CY[j] == -CY[i]
        assert O.lib().orc_opposite(j) == i
        assert O.lib().orc_cx(i) == CX[i] and O.lib().orc_cy(i) == CY[i]
        assert O.lib().orc_weight(i) == W[i]


def test_cylinder_placement_area():  # :116-141
    o = O.Oracle(800, 300)
    o.place_cylinder(200, 150, 30)
    ob = o.obstacle
    assert ob[150, 200] == 1 and ob[0, 0] == 0 and ob[299, 799] == 0
    ratio = ob.sum() / (math.pi * 30 * 30)
    assert 0.6 < ratio < 1.5
    # strict inequality sqrt(dx^2+dy^2) < r: the node at distance exactly r is fluid
    assert ob[150, 230] == 0 and ob[150, 229] == 1


def test_time_step_mass_conservation():  # :158-199 (10 steps, tau=0.56, within 1%)
    o = O.Oracle(800, 300)
    o.place_cylinder(200, 150, 30)
    o.init_equilibrium(1.0, 0.1, 0.0)
    m0 = o.mass()
    o.step(0.56, 0.1, 10)
    assert abs(o.mass() - m0) < 0.01 * m0


def test_tau_from_re():  # :204-219 (uses NX rather than D, as the reference test does)
    for re, tau in ((100.0, 2.9), (200.0, 1.7)):
        assert abs(0.5 + 3.0 * (0.1 * 800 / re) - tau) < 1e-12


# ---------------------------------------------------------------- MRT against textbook algebra

def dhumieres_matrix():
    """D2Q9 moment basis (Lallemand & Luo 2000), rows rho,e,eps,jx,qx,jy,qy,pxx,pxy,
    built from the velocity set, not typed in."""
    rows = []
    for i in range(9):
        cx, cy = CX[i], CY[i]
        c2 = cx * cx + cy * cy
        rows.append([1, -4 + 3 * c2, 4 - 21 * c2 / 2 + 9 * c2 * c2 / 2, cx, (-5 + 3 * c2) * cx, cy,
                     (-5 + 3 * c2) * cy, cx * cx - cy * cy, cx * cy])
    return np.array(rows, dtype=float).T


def textbook_mrt(f, rates, s_shear=None):
    M = dhumieres_matrix()
    m = M @ f
    rho, jx, jy = m[0], m[3], m[5]
    u, v = jx / rho, jy / rho
    meq = np.array([rho, -2 * rho + 3 * (jx * jx + jy * jy) / rho, rho - 3 * (jx * jx + jy * jy) / rho,
                    jx, -jx, jy, -jy, (jx * jx - jy * jy) / rho, jx * jy / rho])
    ss = rates[0] if s_shear is None else s_shear
    S = np.diag([0, rates[1], rates[1], 0, rates[2], 0, rates[2], ss, ss])
    return f - np.linalg.solve(M, S @ (m - meq))


def random_state(rng, rho=1.0, u=0.05, v=-0.03, noise=2e-3):
    return np.array([O.equilibrium(i, rho, u, v) for i in range(9)]) * (1 + noise * rng.standard_normal(9))


@pytest.mark.parametrize("tau", [0.51, 0.68, 1.268, 3.0])
def test_mrt_matches_matrix_form(tau):
    rng = np.random.default_rng(1)
    for _ in range(20):
        f = random_state(rng)
        got = O.mrt_collide(f, tau)
        exp = textbook_mrt(f, O.mrt_rates(tau))
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-14)


def test_mrt_conserves_mass_and_momentum_and_fixes_equilibrium():
    rng = np.random.default_rng(2)
    f = random_state(rng, noise=1e-2)
    g = O.mrt_collide(f, 0.6)
    for w in (np.ones(9), CX, CY):
        assert abs((g - f) @ w) < 1e-15
    feq = np.array([O.equilibrium(i, 1.1, 0.07, 0.02) for i in range(9)])
    np.testing.assert_allclose(O.mrt_collide(feq, 0.6), feq, rtol=0, atol=1e-15)


def test_mrt_with_equal_rates_is_bgk():
    rng = np.random.default_rng(3)
    tau = 0.8
    f = random_state(rng)
    rho, u, v = O.macros(f)
    feq = np.array([O.equilibrium(i, rho, u, v) for i in range(9)])
    got = O.mrt_collide(f, tau, rates=[1 / tau, 1 / tau, 1 / tau])
    np.testing.assert_allclose(got, f - (f - feq) / tau, rtol=0, atol=1e-15)


def test_mrt_rate_clamp():  # src/lbm_types.hpp:72-76
    assert list(O.mrt_rates(0.4)) == [1.99, 1.2, 1.0]
    assert list(O.mrt_rates(3.0)) == [0.5, 1.2, 1.0]
    assert O.mrt_rates(0.68)[0] == 1 / 0.68


def test_les_shear_rate_closed_form():
    """Smagorinsky + Van Driest (src/lbm.hpp:155-182): tau_eff = (tau + sqrt(tau^2 + 4A))/2,
    A = 9 cs_d^2 Q / (2 rho), cs_d = cs (1 - exp(-y+/26)), Q = sqrt(2 (Pxx_neq^2 + Pxy_neq^2))."""
    rng = np.random.default_rng(4)
    tau, cs, yp = 0.52, 0.1, 7.5
    for _ in range(10):
        f = random_state(rng, noise=5e-2)
        M = dhumieres_matrix()
        m = M @ f
        rho, jx, jy = m[0], m[3], m[5]
        pxx_neq = m[7] - (jx * jx - jy * jy) / rho
        pxy_neq = m[8] - jx * jy / rho
        Q = math.sqrt(2 * (pxx_neq ** 2 + pxy_neq ** 2))
        csd = cs * (1 - math.exp(-yp / 26))
        A = 9 * csd * csd * Q / (2 * rho)
        s = min(1.99, max(0.5, 2 / (tau + math.sqrt(tau * tau + 4 * A))))
        got = O.mrt_collide(f, tau, cs_sq=cs * cs, y_plus=yp)
        exp = textbook_mrt(f, O.mrt_rates(tau), s_shear=s)
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-14)
        # y+ = 0 keeps the undamped constant
        csd = cs
        A = 9 * csd * csd * Q / (2 * rho)
        s0 = min(1.99, max(0.5, 2 / (tau + math.sqrt(tau * tau + 4 * A))))
        np.testing.assert_allclose(O.mrt_collide(f, tau, cs_sq=cs * cs, y_plus=0.0),
                                   textbook_mrt(f, O.mrt_rates(tau), s_shear=s0), rtol=0, atol=1e-14)


# ---------------------------------------------------------------- geometry

def _q_circle_bisect(cx, cy, r, xf, yf, i):
    """First crossing of the circle along the link by dense sampling + bisection."""
    ex, ey = CX[i], CY[i]
    g = lambda t: (xf + t * ex - cx) ** 2 + (yf + t * ey - cy) ** 2 - r * r
    ts = np.linspace(0.0, 1.0, 4001)
    vals = g(ts)
    for k in range(1, len(ts)):
        if vals[k] == 0.0 and ts[k] > 0:
            return ts[k]
        if np.sign(vals[k]) != np.sign(vals[k - 1]) and ts[k] > 0:
            lo, hi = ts[k - 1], ts[k]
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if np.sign(g(mid)) == np.sign(g(lo)):
                    lo = mid
                else:
                    hi = mid
            return 0.5 * (lo + hi)
    return 1.0


def test_q_circle_matches_bisection_on_every_link():
    cx, cy, r = 24.3, 25.1, 6.7
    checked = 0
    for y in range(15, 36):
        for x in range(14, 35):
            if (x - cx) ** 2 + (y - cy) ** 2 < r * r:
                continue
            for i in range(1, 9):
                tx, ty = x + CX[i], y + CY[i]
                if (tx - cx) ** 2 + (ty - cy) ** 2 >= r * r:
                    continue
                q = O.q_circle(cx, cy, r, x, y, i)
                assert 0 < q <= 1
                assert abs(q - _q_circle_bisect(cx, cy, r, x, y, i)) < 1e-9
                checked += 1
    assert checked > 50
    # rays that miss give the documented 1.0 fallback (src/lbm_types.hpp:129,137)
    assert O.q_circle(0, 0, 1, 10, 10, 1) == 1.0


def test_q_polygon_matches_linear_solve():
    xy = np.array([[10.2, 10.1], [20.7, 12.4], [18.3, 19.9], [9.1, 17.2]])
    n = len(xy)
    for (x, y) in [(9, 12), (21, 15), (15, 9), (15, 21), (10, 19), (19, 11)]:
        for i in range(1, 9):
            best = 1.0
            for v in range(n):
                p1, p2 = xy[v], xy[(v + 1) % n]
                A = np.array([[CX[i], -(p2[0] - p1[0])], [CY[i], -(p2[1] - p1[1])]], float)
                if abs(np.linalg.det(A)) < 1e-12:
                    continue
                t, s = np.linalg.solve(A, p1 - np.array([x, y], float))
                if 0 < t <= best and 0 <= s <= 1:
                    best = t
            assert abs(O.q_polygon(xy, x, y, i) - best) < 1e-12


def test_point_in_polygon_rotated_rectangle_half_planes():
    rng = np.random.default_rng(5)
    ang = 0.3
    c, s = math.cos(ang), math.sin(ang)
    hx, hy = 7.0, 2.0
    xy = np.array([[-hx, -hy], [hx, -hy], [hx, hy], [-hx, hy]]) @ np.array([[c, s], [-s, c]]) + [20, 20]
    for _ in range(500):
        px, py = rng.uniform(5, 35, 2)
        lx = (px - 20) * c + (py - 20) * s
        ly = -(px - 20) * s + (py - 20) * c
        inside = abs(lx) < hx - 1e-9 and abs(ly) < hy - 1e-9
        outside = abs(lx) > hx + 1e-9 or abs(ly) > hy + 1e-9
        got = O.point_in_polygon(px, py, xy)
        if inside:
            assert got
        if outside:
            assert not got


def _bfs_python(obs):
    ny, nx = obs.shape
    d = np.full((ny, nx), 1e9)
    q = [(y, x) for y in range(ny) for x in range(nx) if obs[y, x]]
    for y, x in q:
        d[y, x] = 0
    h = 0
    while h < len(q):
        y, x = q[h]
        h += 1
        for dy, dx in ((0, -1), (0, 1), (-1, 0), (1, 0)):
            yy, xx = y + dy, x + dx
            if 0 <= yy < ny and 0 <= xx < nx and d[yy, xx] > d[y, x] + 1:
                d[yy, xx] = d[y, x] + 1
                q.append((yy, xx))
    return d


def test_wall_distance_sweep_equals_bfs_equals_bruteforce():
    rng = np.random.default_rng(6)
    obs = (rng.random((23, 31)) < 0.04).astype(np.uint8)
    obs[5:9, 10:20] = 1
    sweep = O.wall_distance(obs, "sweep")
    bfs = O.wall_distance(obs, "bfs")
    np.testing.assert_array_equal(sweep, bfs)
    np.testing.assert_array_equal(bfs, _bfs_python(obs))
    # no solid at all: every node keeps the 1e9 sentinel
    assert (O.wall_distance(np.zeros((5, 6), np.uint8)) == 1e9).all()


# ---------------------------------------------------------------- streaming, boundaries, forces

def test_resting_fluid_stays_at_rest_periodic():
    o = O.Oracle(20, 12, O.RIBS)
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.step(0.7, 0.0, 5)
    # a fixed point up to FP64 rounding of the moment transform
    np.testing.assert_allclose(o.f, np.broadcast_to(W, o.f.shape), rtol=0, atol=2e-16)


def test_periodic_mass_conservation_with_obstacles():
    rng = np.random.default_rng(7)
    o = O.Oracle(40, 24, O.RIBS)
    m = np.zeros((24, 40), np.uint8)
    m[10:14, 15:19] = 1
    o.set_obstacle(m)
    f = np.array([[[O.equilibrium(i, 1.0, 0.05, 0.0) for i in range(9)]] * 40] * 24)
    f *= 1 + 1e-3 * rng.standard_normal(f.shape)
    o.set_f(f)
    m0 = o.mass()
    o.step(0.6, 0.0, 30)
    assert abs(o.mass() - m0) < 1e-11 * m0


def numpy_push_periodic(f, obs):
    """Independent push streaming with simple bounce-back on a fully periodic grid."""
    ny, nx, _ = f.shape
    g = np.zeros_like(f)
    for i in range(9):
        for y in range(ny):
            for x in range(nx):
                if obs[y, x]:
                    continue
                tx, ty = (x + CX[i]) % nx, (y + CY[i]) % ny
                j = [0, 3, 4, 1, 2, 7, 8, 5, 6][i]
                if obs[ty, tx]:
                    g[y, x, j] = f[y, x, i]
                else:
                    g[ty, tx, i] = f[y, x, i]
    return g


def test_streaming_bruteforce_tiny_periodic_grid():
    """BGK with tau = 1e30 leaves f unchanged in FP64, so a step is pure streaming."""
    rng = np.random.default_rng(8)
    obs = np.zeros((7, 9), np.uint8)
    obs[3, 4] = obs[3, 5] = obs[5, 1] = 1
    f = rng.random((7, 9, 9))
    o = O.Oracle(9, 7, O.RIBS, collision=O.BGK)
    o.set_obstacle(obs)
    o.set_f(f)
    o.step(1e30, 0.0)
    exp = numpy_push_periodic(f, obs)
    fluid = obs == 0
    np.testing.assert_array_equal(o.f[fluid], exp[fluid])


def test_zou_he_sets_inlet_velocity_and_outflow_copies():
    o = O.Oracle(40, 20, O.CYLINDER)
    o.place_cylinder(12, 10, 3)
    o.init_equilibrium(1.0, 0.05, 0.0)
    o.step(0.7, 0.08, 3)
    f = o.f
    rho = f[:, 0].sum(axis=1)
    ux = (f[:, 0] * CX).sum(axis=1) / rho
    np.testing.assert_allclose(ux, 0.08, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(f[:, -1], f[:, -2])


def test_lid_row_is_equilibrium_at_lid_velocity():
    o = O.Oracle(16, 16, O.CAVITY)
    o.place_walls()
    o.init_equilibrium(1.0, 0.0, 0.0)
    o.step(0.8, 0.1, 4)
    top = o.f[-1]
    for x in range(1, 15):
        rho = top[x].sum()
        exp = [O.equilibrium(i, rho, 0.1, 0.0) for i in range(9)]
        np.testing.assert_allclose(top[x], exp, rtol=0, atol=1e-15)
    assert (top[0] == 0).all() and (top[15] == 0).all()  # solid corners stay zero


def test_momentum_exchange_matches_independent_link_sum():
    o = O.Oracle(50, 30, O.CYLINDER)
    o.place_cylinder(15, 15, 5)
    o.init_equilibrium(1.0, 0.1, 0.0)
    o.step(0.7, 0.1, 7)
    f, obs = o.f, o.obstacle
    fx = fy = 0.0
    for y in range(30):
        for x in range(50):
            if obs[y, x]:
                continue
            for i in range(1, 9):
                tx, ty = x + CX[i], (y + CY[i]) % 30
                if 0 <= tx < 50 and obs[ty, tx]:
                    fb = f[y, x, [0, 3, 4, 1, 2, 7, 8, 5, 6][i]]
                    fx += CX[i] * 2.0 * fb
                    fy += CY[i] * 2.0 * fb
    gx, gy = o.total_force()
    assert abs(gx - fx) < 1e-12 * abs(fx) and abs(gy - fy) < 1e-12 + 1e-12 * abs(fy)


def _wall_case(scheme_bouzidi):
    """Closed box (wall edges, as the cavity case) whose bottom wall is a polygon edge
    at y = 0.5: every link into row 0 has q = 1/2 exactly (axis and diagonal)."""
    o = O.Oracle(30, 16, O.CAVITY)
    poly = np.array([[-10.0, -10.0], [40.0, -10.0], [40.0, 0.5], [-10.0, 0.5]])
    o.place_polygon(poly)
    o.set_mei(not scheme_bouzidi)
    o.init_equilibrium(1.0, 0.05, 0.0)
    return o


def test_bouzidi_at_q_half_is_halfway_bounce_back():
    a = _wall_case(True)
    assert a.obstacle[0].all() and not a.obstacle[1:].any()
    poly = np.array([[-10.0, -10.0], [40.0, -10.0], [40.0, 0.5], [-10.0, 0.5]])
    assert [O.q_polygon(poly, 5, 1, i) for i in (4, 7, 8)] == [0.5, 0.5, 0.5]
    b = O.Oracle(30, 16, O.CAVITY)  # same mask, no interpolation geometry: simple BB
    b.set_obstacle(a.obstacle.copy())
    b.init_equilibrium(1.0, 0.05, 0.0)
    a.step(0.7, 0.05, 5)
    b.step(0.7, 0.05, 5)
    np.testing.assert_array_equal(a.f, b.f)


def test_mei_at_q_one_gives_wall_equilibrium():
    """A cylinder far from the bottom wall: links into the wall miss the circle, q = 1,
    and Mei returns feq_i(rho, 0) with rho of the post-collision node (src/lbm.hpp:255)."""
    o = O.Oracle(40, 24, O.CYLINDER_NEAR_WALL)
    o.place_cylinder(30, 15, 3)
    m = o.obstacle.copy()
    m[0, :] = 1
    o.set_obstacle(m)
    o.init_equilibrium(1.0, 0.05, 0.0)
    f0 = o.f.copy()
    o.step(0.7, 0.05)
    x, y = 10, 1
    post = O.mrt_collide(f0[y, x], 0.7)
    rho = post.sum()
    for i in (4, 7, 8):  # links from (10,1) into the wall row
        j = [0, 3, 4, 1, 2, 7, 8, 5, 6][i]
        assert abs(o.f[y, x, j] - O.equilibrium(i, rho, 0.0, 0.0)) < 1e-15


# ---------------------------------------------------------------- random numbers

def test_mt19937_known_answer():
    """[rand.predef]: the 10000th consecutive output of mt19937 default-seeded (5489) is 4123659995."""
    g = O.MT19937(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 4123659995


def test_uniform_real_two_draw_canonical():
    a = O.MT19937(42)
    b = O.MT19937(42)
    for _ in range(100):
        lo, hi = b.next(), b.next()
        r = (lo + hi * 2 ** 32) / 2 ** 64
        assert a.uniform(-1e-4, 1e-4) == r * 2e-4 + -1e-4
